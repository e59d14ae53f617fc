"""TEST INFRASTRUCTURE: compile libhfx's unmodified sources with g++ against the
HIP stand-in in tests/cpuemu/hip/hip_runtime.h, giving
tests/cpuemu/libhfx_cpuemu.so (one lane per workgroup) and, with --lanes,
tests/cpuemu/libhfx_lanes.so (every workgroup as its blockDim.x lanes in
64-lane waves), both with the C ABI of hfcode_amd/libhfx.so.  Used by the CPU
test suite only, to check host logic and the generated kernels against the
oracle on a machine without a GPU.

    python tests/cpuemu/build_cpuemu.py [--lanes]
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "hfcode_amd", "csrc")
GEN = os.path.join(CSRC, "gen")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libhfx_cpuemu.so")
OBJ_LANES = os.path.join(HERE, "_obj_lanes")
LIB_LANES = os.path.join(HERE, "libhfx_lanes.so")
FLAGS = ["-x", "c++", "-std=c++17", "-O1", "-fPIC", "-w", "-I" + HERE, "-I" + os.path.join(ROOT, "include"),
         "-I" + CSRC, "-I" + GEN]


def _one(src, lanes=False):
    obj = os.path.join(OBJ_LANES if lanes else OBJ, os.path.basename(src) + ".o")
    deps = [src, os.path.join(HERE, "hip", "hip_runtime.h"), os.path.join(HERE, "rocblas", "rocblas.h"),
            os.path.join(HERE, "rocsolver", "rocsolver.h"), os.path.join(CSRC, "hfx_device.h"),
            os.path.join(CSRC, "hfx_kernels.h"), os.path.join(CSRC, "hfx_internal.h"), os.path.join(GEN, "eri_dispatch.h")]
    if os.path.exists(obj) and all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps):
        return None
    r = subprocess.run(["g++"] + FLAGS + (["-DHFX_EMU_LANES"] if lanes else []) + ["-c", src, "-o", obj],
                       stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    return r.stdout if r.returncode else None


def build(jobs=8, lanes=False):
    sys.path.insert(0, ROOT)
    from hfcode_amd.gen import gen_eri, gen_onee
    gen_eri.main(GEN)
    gen_onee.main(GEN)
    obj_dir, lib = (OBJ_LANES, LIB_LANES) if lanes else (OBJ, LIB)
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [os.path.join(CSRC, "hfx_kernels.hip"), os.path.join(CSRC, "hfx_host.hip"), os.path.join(CSRC, "hfx_scf.hip"),
            os.path.join(CSRC, "hfx_batch.hip"),
            os.path.join(HERE, "emu_smem.cpp"), os.path.join(CSRC, "hfx_onee.cpp"),
            os.path.join(CSRC, "hfx_input.cpp")]
    srcs += sorted(os.path.join(GEN, f) for f in os.listdir(GEN) if f.endswith(".hip"))
    srcs.sort(key=lambda s: -os.path.getsize(s))
    with cf.ThreadPoolExecutor(jobs) as ex:
        errs = [e for e in ex.map(lambda s: _one(s, lanes), srcs) if e]
    if errs:
        raise RuntimeError("g++ failed:\n" + "\n".join(e[-3000:] for e in errs))
    objs = [os.path.join(obj_dir, os.path.basename(s) + ".o") for s in srcs]
    subprocess.run(["g++", "-shared", "-o", lib + ".tmp"] + objs, check=True)
    os.replace(lib + ".tmp", lib)
    return lib


if __name__ == "__main__":
    print(build(lanes="--lanes" in sys.argv[1:]))
