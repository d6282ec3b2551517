"""Build the D3Q19 shear-cell libraries in-tree.

Products (all under this package directory, git-ignored, shipped with the tree):
  lib/libd3q19.so      HIP C-ABI library for gfx950 (hipcc)
  lib/libd3q19host.so  C host parser / set-up / writers (gcc, no GPU)
  bin/d3q19_sim        C host program linking both (the reference's main, sim_main.c)

Usage: python build.py [--jobs N]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops off: packed FP32 forms are not used anywhere in this library.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC, "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
HIP_SOURCES = ["abi.hip", "lbm.hip", "ibm.hip", "particles.hip", "batch.hip", "checkpoint.hip", "probe.hip"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build_host():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = [os.path.join(HOST, f) for f in ("params.c", "output.c")]
    run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", "-I" + INC, "-o",
         os.path.join(LIB, "libd3q19host.so")] + srcs + ["-lm"])


LLVM = "/opt/rocm/llvm/bin"


def check_kernel_registration(obj):
    """Every kernel (and device variable) that the host half of `obj` registers by name must have a
    definition in its gfx950 code object; a launch of a registered kernel without one aborts in the
    HIP runtime ("Cannot find Symbol"), so such an object fails the build instead.  The host half
    holds the registered names as string literals; the code object is the gfx950 entry of the
    object's .hip_fatbin bundle, whose kernels carry <name>.kd descriptors."""
    import tempfile
    r = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "-p", ".rodata.str1.1", obj], capture_output=True, text=True)
    host = set()
    for line in r.stdout.splitlines():
        if "]" in line:
            name = line.split("]", 1)[1].strip()
            if name.startswith("_Z"):
                host.add(name)
    if not host:
        return
    with tempfile.TemporaryDirectory() as tmp:
        fb, co = os.path.join(tmp, "fatbin"), os.path.join(tmp, "gfx950.co")
        run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section", ".hip_fatbin=" + fb, obj, os.path.join(tmp, "discard.o")])
        run([os.path.join(LLVM, "clang-offload-bundler"), "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--input=" + fb,
             "--output=" + co, "--unbundle"])
        syms = run([os.path.join(LLVM, "llvm-readelf"), "--syms", co]).stdout.split()
    device = {t[:-3] if t.endswith(".kd") else t for t in syms}
    missing = sorted(n for n in host if n not in device)
    if missing:
        raise RuntimeError("%s registers kernels without gfx950 code: %s" % (os.path.basename(obj), ", ".join(missing[:5])))


def build_hip(jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)

    def one(src):
        out = os.path.join(OBJ, src.replace(".hip", ".o"))
        run([HIPCC] + HIP_FLAGS + ["-c", "-x", "hip", os.path.join(CSRC, src), "-o", out])
        return out

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    for o in objs:
        check_kernel_registration(o)
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", os.path.join(LIB, "libd3q19.so")] + objs)


def build_sim():
    run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I" + INC, "-o", os.path.join(BIN, "d3q19_sim"),
         os.path.join(HOST, "sim_main.c"), "-L" + LIB, "-ld3q19host", "-ld3q19",
         "-Wl,-rpath,$ORIGIN/../lib", "-lm"])


def build_all(jobs=4):
    build_host()
    build_hip(jobs)
    build_sim()


if __name__ == "__main__":
    j = 4
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build_all(j)
    print("built", LIB, BIN)
