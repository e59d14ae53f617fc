"""Build libhfx.so (gfx950) in-tree.

1. regenerate the per-class kernels (hfcode_amd/gen/gen_eri.py) into
   hfcode_amd/csrc/gen/;
2. compile every .hip translation unit with hipcc for gfx950, in parallel;
3. link hfcode_amd/libhfx.so (with rocBLAS and rocSOLVER for the SCF driver).

    python -m hfcode_amd.build [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "gen")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhfx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOST_SOURCES = ("hfx_onee.cpp", "hfx_input.cpp")

HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-munsafe-fp-atomics",                                   # hardware FP64 atomic add
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + GEN,
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def sources(gen=GEN):
    srcs = [os.path.join(CSRC, f) for f in ("hfx_kernels.hip", "hfx_host.hip", "hfx_scf.hip", "hfx_batch.hip")]
    srcs += sorted(os.path.join(gen, f) for f in os.listdir(gen) if f.endswith(".hip"))
    return srcs


def compile_one(src, gen=GEN, obj_dir=OBJ):
    obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
    deps = [src, os.path.join(CSRC, "hfx_device.h"), os.path.join(CSRC, "hfx_kernels.h"),
            os.path.join(CSRC, "hfx_internal.h")]
    if os.path.basename(src).startswith("hfx_"):
        deps += [os.path.join(gen, "eri_dispatch.h"),
                 os.path.join(ROOT, "include", "hfx.h")]
    if os.path.exists(obj) and all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps if os.path.exists(d)):
        return obj, None
    flags = [f if f != "-I" + GEN else "-I" + gen for f in HIPFLAGS]
    cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        return obj, r.stdout
    return obj, None


def build(jobs=None, verbose=True, gen=GEN, obj_dir=OBJ, lib=LIB, budgets=None):
    """Generate and compile everything into `lib`.  `gen`, `obj_dir`, `lib`
    and `budgets` ({class: (fp64, fp32)} grouping budgets, or one pair for
    every class) let the grouping autotuner build variants beside the
    library."""
    sys.path.insert(0, ROOT)
    from hfcode_amd.gen import gen_eri, gen_onee
    saved = gen_eri.GROUP_BUDGETS
    if budgets is not None:
        gen_eri.GROUP_BUDGETS = (budgets if isinstance(budgets, dict)
                                 else {c: tuple(budgets) for c in gen_eri.CLASSES})
    try:
        gen_eri.main(gen)
    finally:
        gen_eri.GROUP_BUDGETS = saved
    gen_onee.main(gen)
    os.makedirs(obj_dir, exist_ok=True)
    srcs = sources(gen)
    jobs = jobs or min(16, os.cpu_count() or 4)
    errs = []
    objs = []
    # the biggest translation units first
    srcs.sort(key=lambda s: -os.path.getsize(s))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: compile_one(s, gen, obj_dir), srcs):
            objs.append(obj)
            if err:
                errs.append(err)
    for name in HOST_SOURCES:     # host-only C++: one-electron set-up, input readers
        host_src = os.path.join(CSRC, name)
        host_obj = os.path.join(obj_dir, name + ".o")
        r = subprocess.run(["g++", "-O2", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-c", host_src,
                            "-o", host_obj], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode:
            errs.append(r.stdout)
        objs.append(host_obj)
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(e[-4000:] for e in errs))
    # rocBLAS / rocSOLVER: the products and eigensolves of the SCF driver (hfx_scf.hip)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib + ".tmp"] + sorted(objs) + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(lib + ".tmp", lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build(a.j)
