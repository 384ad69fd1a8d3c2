"""Build the native parts in-tree: libap26gpu.so (HIP, gfx950), the C host
(libap26host.so + the AP26 command) and, for tests only, the oracle.

No cmake/ninja: plain hipcc / gcc command lines, so the built files sit next to
their sources and travel with the tree.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ap26_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")

LIB_GPU = os.path.join(PKG, "libap26gpu.so")
LIB_HOST = os.path.join(PKG, "libap26host.so")
BIN_AP26 = os.path.join(PKG, "bin", "AP26")
BIN_PROBE = os.path.join(PKG, "bin", "ceiling_probe")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are switched off for every HIP object.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    # hipcc's host pass warns about the packed-fp32 feature; keep other output
    return r


def _atomic_link(tmp, dst):
    os.replace(tmp, dst)


def build_gpu(verbose=False):
    """Compile the HIP sources to objects in parallel, then link the library."""
    srcs = [os.path.join(CSRC, "ap26_gpu.hip"), os.path.join(CSRC, "ap26_batch.hip")]
    srcs += [os.path.join(CSRC, f"ap26_pvariants_{i}.hip") for i in range(1, 6)]
    srcs = [s for s in srcs if os.path.exists(s)]
    objs = [os.path.join(PKG, os.path.basename(s) + ".o") for s in srcs]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", "-o", o, s] for s, o in zip(srcs, objs)]
    # the issue/gather ceiling probe (a stand-alone program, not linked into the library)
    os.makedirs(os.path.dirname(BIN_PROBE), exist_ok=True)
    probe_src = os.path.join(CSRC, "ceiling_probe.hip")
    if os.path.exists(probe_src):
        cmds.append([HIPCC, *HIP_FLAGS, "-o", BIN_PROBE + ".tmp", probe_src])
    if verbose:
        for c in cmds:
            print(" ".join(c))
    procs = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for c in cmds]
    failed = []
    for c, pr in zip(cmds, procs):
        out, err = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(" ".join(c))
    if failed:
        raise RuntimeError("build step failed: " + "; ".join(failed))
    tmp = LIB_GPU + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    for o in objs:
        os.remove(o)
    _atomic_link(tmp, LIB_GPU)
    if os.path.exists(BIN_PROBE + ".tmp"):
        _atomic_link(BIN_PROBE + ".tmp", BIN_PROBE)
    return LIB_GPU


def build_host(verbose=False):
    os.makedirs(os.path.dirname(BIN_AP26), exist_ok=True)
    cflags = ["-O2", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
              f"-I{INCLUDE}", f"-I{CSRC}"]
    tmp = LIB_HOST + ".tmp"
    _run(["gcc", *cflags, "-fPIC", "-shared", "-o", tmp,
          os.path.join(HOST, "ap26_host.c")])
    _atomic_link(tmp, LIB_HOST)
    if os.path.exists(os.path.join(HOST, "ap26_main.c")):
        tmpb = BIN_AP26 + ".tmp"
        _run(["gcc", *cflags, "-o", tmpb,
              os.path.join(HOST, "ap26_main.c"), os.path.join(HOST, "ap26_host.c"),
              f"-L{PKG}", "-lap26gpu", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,$ORIGIN"])
        _atomic_link(tmpb, BIN_AP26)
    return LIB_HOST


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", ORACLE, "-B"])
    ref = "/root/reference"
    if os.path.isdir(ref) and shutil.which("g++"):
        _run(["make", "-s", "-f", os.path.join(ORACLE, "Makefile_ref"), f"REF={ref}",
              f"OUT={os.path.join(ORACLE, '_ref')}"], cwd=ROOT)


def build_all(verbose=False):
    build_gpu(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB_GPU, LIB_HOST, BIN_AP26)
