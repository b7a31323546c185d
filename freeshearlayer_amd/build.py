"""Build of the native libraries (hipcc for gfx950, gcc for the ANSI-C host).

Outputs (git-ignored, built in-tree so they travel with the source tree):
  freeshearlayer_amd/lib/libfsl.so         HIP kernels + C-ABI + ANSI-C host layer
  freeshearlayer_amd/lib/libfsl_layer2.so  reference-named drop-in entry points
  freeshearlayer_amd/lib/fsl_layer2        layer2-compatible executable
  oracle/liboracle.so                      CPU restatement (test-only checker)
  oracle/_ref/*.so                         compiled reference (where it exists)

Usage: python -m freeshearlayer_amd.build [--no-oracle] [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
REFERENCE_SRC = "/root/reference/src"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

# -ffp-contract=off keeps the FP64 path in the reference's rounding order
# (except in the fast build named below);
# packed FP32 instructions are disabled (the FP32 kernels use scalar VALU).
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + INCLUDE, "-I" + CSRC,
]
C_FLAGS = ["-O2", "-std=c99", "-Wall", "-fPIC", "-ffp-contract=off", "-I" + INCLUDE]

HIP_SOURCES = ["fsl_api.hip", "fsl_kernels.hip", "fsl_sgs.hip", "fsl_fused.hip", "fsl_fused_fast.hip",
               "fsl_fused_mixed.hip", "fsl_mlp.hip", "fsl_learned.hip", "fsl_output.hip", "fsl_sgs_march.hip"]
# the fast build of the FP64 stage kernel (FSL_PREC_FP64_FAST) contracts a * b + c into FMA
CONTRACTED = {"fsl_fused_fast.hip"}
# scheduler strategies per file, from A/B timings at 384x256x256 on one GPU: max-ilp makes the
# mixed stage kernel 0.8 % and the network kernel 1.3 % faster (no spills; the FP64 stage kernel
# spills with it), iterative-ilp the FP64 dynamic model 6.5 % and the first FP32 dynamic-model
# kernel (k_sgs_a<float>) 20 % faster
SCHED = {"fsl_fused_mixed.hip": "max-ilp", "fsl_mlp.hip": "max-ilp", "fsl_sgs_march.hip": "iterative-ilp",
         "fsl_sgs.hip": "iterative-ilp"}
HOST_SOURCES = ["fsl_host.c", "fsl_driver.c", "fsl_stats.c", "fsl_files.c"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = r.stdout
    # hipcc's host pass does not know the packed-fp32 target feature; drop that noise
    out = "\n".join(l for l in out.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def _newer(target, deps):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps if os.path.exists(d))


def _headers():
    hs = [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h")]
    hs += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    return hs


def build_native(jobs=8, force=True, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    hdrs = _headers()
    tasks = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s.replace(".hip", ".o"))
        deps = [src] + hdrs + ([os.path.join(CSRC, "fsl_fused.hip")] if s in CONTRACTED else [])
        if force or not _newer(obj, deps):
            flags = [("-ffp-contract=fast" if f == "-ffp-contract=off" else f) for f in HIP_FLAGS] \
                if s in CONTRACTED else list(HIP_FLAGS)
            if s in SCHED:
                flags += ["-mllvm", "-amdgpu-sched-strategy=" + SCHED[s]]
            tasks.append([HIPCC] + flags + ["-c", src, "-o", obj])
    for s in HOST_SOURCES + ["fsl_layer2_api.c", "fsl_layer2_main.c", "fsl_layer2_p.c"]:
        src = os.path.join(HOST, s)
        obj = os.path.join(BUILD, s.replace(".c", ".o"))
        if force or not _newer(obj, [src] + hdrs):
            tasks.append([CC] + C_FLAGS + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, tasks):
            if verbose and out.strip():
                print(out)
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in HIP_SOURCES]
    objs += [os.path.join(BUILD, s.replace(".c", ".o")) for s in HOST_SOURCES]
    libfsl = os.path.join(LIB, "libfsl.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", libfsl + ".tmp"] + objs +
         ["-Wl,-Bsymbolic-functions", "-Wl,-soname,libfsl.so", "-lm"])
    os.replace(libfsl + ".tmp", libfsl)
    # reference-named drop-in layer over libfsl.so
    l2 = os.path.join(LIB, "libfsl_layer2.so")
    _run([CC, "-shared", "-fPIC", "-o", l2 + ".tmp", os.path.join(BUILD, "fsl_layer2_api.o"),
          "-L" + LIB, "-lfsl", "-Wl,-rpath,$ORIGIN", "-Wl,-Bsymbolic-functions", "-lm"])
    os.replace(l2 + ".tmp", l2)
    exe = os.path.join(LIB, "fsl_layer2")
    _run([CC, "-o", exe + ".tmp", os.path.join(BUILD, "fsl_layer2_main.o"), "-L" + LIB, "-lfsl",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(exe + ".tmp", exe)
    # the Favre-averaging post-processor (favre-average-src/layer2_p.c)
    exe = os.path.join(LIB, "fsl_layer2_p")
    _run([CC, "-o", exe + ".tmp", os.path.join(BUILD, "fsl_layer2_p.o"), "-L" + LIB, "-lfsl",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(exe + ".tmp", exe)
    return libfsl


def build_test_programs():
    """Host programs of the test suite, linked against the built libraries
    (tests/_bin/, git-ignored): layer2_dropin drives libfsl_layer2.so in the
    reference's call order."""
    tdir = os.path.join(ROOT, "tests")
    src = os.path.join(tdir, "layer2_dropin.c")
    if not os.path.exists(src):
        return
    out = os.path.join(tdir, "_bin")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, "layer2_dropin")
    _run([CC] + C_FLAGS + [src, "-o", exe + ".tmp", "-L" + LIB, "-lfsl_layer2", "-lfsl",
                           "-Wl,-rpath,$ORIGIN/../../freeshearlayer_amd/lib", "-lm"])
    os.replace(exe + ".tmp", exe)


def build_oracle(with_reference=True):
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "oracle"])
    if with_reference and os.path.isdir(REFERENCE_SRC):
        _run(["make", "-s", "-C", odir, "ref", "REF=" + REFERENCE_SRC])


def build(jobs=8, oracle=True, verbose=False):
    lib = build_native(jobs=jobs, force=True, verbose=verbose)
    build_test_programs()
    if oracle:
        build_oracle()
    return lib


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    lib = build(jobs=a.jobs, oracle=not a.no_oracle, verbose=a.verbose)
    print("built", lib)


if __name__ == "__main__":
    sys.exit(main())
