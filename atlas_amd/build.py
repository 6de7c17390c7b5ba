"""Build libatlas_mi355x.so in-tree for gfx950.

C host sources are compiled with gcc (C99, HIP runtime API headers only);
HIP kernel sources with hipcc --offload-arch=gfx950; hipcc links the
shared library.  Object files go to atlas_amd/_build/ (git-ignored).

    python -m atlas_amd.build        # or: python atlas_amd/build.py
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libatlas_mi355x.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")

C_SOURCES = ["runtime.c", "xerbla.c", "blas3.c", "lapack.c", "mixed.c", "batched.c",
             "f77.c", "cblas3c.c"]
HIP_SOURCES = ["k_gemm.hip", "k_aux.hip", "k_lu.hip", "k_chol.hip", "k_mp.hip", "k_trsv.hip", "k_tslu.hip", "k_batched.hip",
               "k_qr.hip", "noalloc.hip", "k_cplx.hip"]

CFLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wno-unused-function",
          "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
          "-I" + os.path.join(ROOT, "include")]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in C_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        jobs.append(["gcc"] + CFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        jobs.append([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    # -Bsymbolic-functions: calls between this library's own entry points (the F77
    # layer into cblas_/clapack_) bind here, whatever else the process has loaded;
    # the error handlers in interposable.list stay replaceable by the program
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-Bsymbolic-functions",
          "-Wl,--dynamic-list=" + os.path.join(CSRC, "interposable.list"),
          "-o", tmp] + objs +
         ["-lpthread", "-lm"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
