/* Test infrastructure: prints the byte offset and size of the trace-header
 * keys vti_su_write fills, as the reference's own header declares them
 * (utils-binary2segy/hc/cjbsegy.h, the struct binary2segy.c reads and
 * writes), so the tests pin the product's layout to the reference's. */
#include <stddef.h>
#include <stdio.h>

#include "hc/cjbsegy.h"

#define KEY(k) printf("%s %zu %zu\n", #k, offsetof(cjbsegy, k), sizeof(((cjbsegy *)0)->k))

int main(void)
{
    printf("size %zu 0\n", sizeof(cjbsegy));
    KEY(tracl); KEY(tracr); KEY(fldr); KEY(tracf); KEY(ep); KEY(trid); KEY(offset); KEY(gelev);
    KEY(sdepth); KEY(scalel); KEY(scalco); KEY(sx); KEY(sy); KEY(gx); KEY(gy); KEY(counit); KEY(ns); KEY(dt);
    return 0;
}
