/*
 * ref_harness.c — command-line driver that calls the reference
 * CPU_vti2dfd (2D-3D-vti-fd-cuda-gtk-GUI/CPU_vti2dfd_kernels.c:366-367)
 * with the argument list its GTK caller passes
 * (GTK_cpuvti2dfd.c:57).  Built by oracle/Makefile into oracle/_ref/ only;
 * used to write golden fixtures in this container.  TEST INFRASTRUCTURE.
 */
#include <stdio.h>
#include <stdlib.h>

void CPU_vti2dfd(int nx, int nz, int dx, int dz, int npd, int SV,
                 const char FNv[], const char FNe[], const char FNd[],
                 int favg, int ns, int fs, int ds, int zs,
                 const char FNshot[], const char FNsnap[], int nt, int dt,
                 int run_count);

int main(int argc, char **argv)
{
    if (argc != 19) {
        fprintf(stderr, "usage: %s nx nz dx dz npd SV vel eps del favg ns fs ds zs shot snap nt dt_us\n", argv[0]);
        return 2;
    }
    CPU_vti2dfd(atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]),
                atoi(argv[5]), atoi(argv[6]), argv[7], argv[8], argv[9],
                atoi(argv[10]), atoi(argv[11]), atoi(argv[12]), atoi(argv[13]),
                atoi(argv[14]), argv[15], argv[16], atoi(argv[17]),
                atoi(argv[18]), 0);
    return 0;
}
