"""Build libvti.so for gfx950 with hipcc (in-tree, next to this file).

    python -m geophysics_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvti.so")
SOURCES = ["vti_kernels.hip", "vti_api.hip", "vti_tables.cpp", "vti_dropin.cpp", "vti_rtm.hip", "vti_grad.hip", "vti_batch.hip", "vti_sweep.cpp", "vti_nn.hip", "vti_ae.hip", "vti_bcodec.hip", "vti_bqcodec.hip", "vti_rtm_learned.hip", "vti_rtm_eb.hip", "vti_adcig.hip", "vti_segy.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no multiply-add fusion, so every step is the reference's
# float arithmetic term by term.  -packed-fp32-ops off: no v_pk_*_f32.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = []
    for s in srcs:
        o = os.path.join(CSRC, os.path.basename(s) + ".o")
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-x", "hip", "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(o)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
