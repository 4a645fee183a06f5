"""Build the native libraries in-tree.

  lib/libsns_hip.so  hand-written HIP kernels for gfx950 + the C-ABI (include/sns_hip.h)
  lib/libsns.so      C host layer with the reference's call surface (include/*.h)
  lib/sns_run        driver with the reference's time loop (host/main.c)

Usage: python -m simplenavierstokessolver_amd.build   (or build() from Python)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")

HIP_SOURCES = ["sns_stencil.hip", "sns_solve.hip", "sns_poisson.hip", "sns_fused.hip", "sns_pred.hip",
               "sns_diag.hip", "sns_capi.hip", "sns_batch.hip", "sns_ysolve.hip"]
HOST_SOURCES = ["common.c", "parallel.c", "param.c", "fluid.c", "temperature.c", "fileio.c",
                "save.c", "logging.c", "statistics.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: products stay separately rounded as in the reference's
# gcc build; packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I", INC, "-I", CSRC]
CC_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
            "-Wno-unused-parameter", "-I", INC]

# include/parallel.h keeps the reference's MPI-typed prototypes, so the host
# layer compiles against an MPI implementation's <mpi.h> (it links no MPI
# library: one rank needs no communication)
MPI_INCLUDE_CANDIDATES = ["/opt/conda/include", "/usr/include/x86_64-linux-gnu/mpich",
                          "/usr/include/mpich", "/usr/lib/x86_64-linux-gnu/openmpi/include",
                          "/usr/include/openmpi"]


def mpi_include():
    """Directory holding mpi.h: $SNS_MPI_INCLUDE, else the first candidate
    that has one."""
    env = os.environ.get("SNS_MPI_INCLUDE")
    for d in ([env] if env else []) + MPI_INCLUDE_CANDIDATES:
        if os.path.exists(os.path.join(d, "mpi.h")):
            return d
    raise RuntimeError("build failed: no mpi.h found (set SNS_MPI_INCLUDE to the directory "
                       "of an MPI implementation's mpi.h)")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        futs = [ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(srcs, objs)]
        for f in futs:
            out = f.result()
            if verbose and out:
                print(out)
    hip_so = os.path.join(LIB, "libsns_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", hip_so + ".tmp"] + objs)
    os.replace(hip_so + ".tmp", hip_so)
    host_so = os.path.join(LIB, "libsns.so")
    cc_flags = CC_FLAGS + ["-I", mpi_include()]
    _run(["gcc"] + cc_flags + ["-shared", "-o", host_so + ".tmp"]
         + [os.path.join(HOST, s) for s in HOST_SOURCES]
         + ["-L" + LIB, "-lsns_hip", "-lm", "-Wl,-rpath,$ORIGIN"])
    os.replace(host_so + ".tmp", host_so)
    exe = os.path.join(LIB, "sns_run")
    _run(["gcc"] + cc_flags + ["-o", exe + ".tmp", os.path.join(HOST, "main.c"),
                               "-L" + LIB, "-lsns", "-lsns_hip", "-lm", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return hip_so, host_so, exe


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
