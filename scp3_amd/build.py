"""Build libscp3g.so (the HIP kernels and the C ABI) in-tree for gfx950.

    python -m scp3_amd.build          # or scp3_amd.build.build()

Objects go to scp3_amd/_build/, the library to scp3_amd/libscp3g.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libscp3g.so")

SOURCES = [
    "sobol.hip",
    "map.hip",
    "pes_qtip4pf.hip",
    "pes_ttm3f.hip",
    "basis.hip",
    "syrk.hip",
    "syrk_mixed.hip",
    "syrk_fast.hip",
    "syrk64_fast.hip",
    "eigen.hip",
    "mlp.hip",
    "grad.hip",
    "scp1.hip",
    "context.cpp",
    "modes.cpp",
    "pes_cpu.cpp",
    "small_eigh.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SCP3G_ARCH", "gfx950")
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    # packed FP32 VALU instructions (v_pk_*_f32) stay disabled: some of their
    # forms have produced wrong results, with no error, on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT, src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature: harmless
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), err))
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cli(verbose)
    return LIB


HDF5_INC = os.environ.get("SCP3G_HDF5_INC", "/opt/conda/include")
HDF5_LIB = os.environ.get("SCP3G_HDF5_LIB", "/opt/conda/lib")
BIN = os.path.join(HERE, "bin")


def build_cli(verbose=False):
    """The SCP3, eigensolver and SCP1 drivers, linked against libscp3g.so (and
    the HDF5 C library for sM_*.h5 when its header is present)."""
    os.makedirs(BIN, exist_ok=True)
    hdf5 = os.path.exists(os.path.join(HDF5_INC, "hdf5.h"))
    extra = []
    if hdf5:
        # the HDF5 directory also holds an older libstdc++ than the ROCm
        # toolchain needs, so the driver's run path gets a directory with a
        # link to libhdf5 alone
        real = os.path.realpath(os.path.join(HDF5_LIB, "libhdf5.so"))
        soname = "libhdf5.so." + real.split("libhdf5.so.")[1].split(".")[0]
        ldir = os.path.join(BIN, "hdf5")
        os.makedirs(ldir, exist_ok=True)
        link = os.path.join(ldir, soname)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink(real, link)
        extra = ["-DSCP3G_HAVE_HDF5", "-I" + HDF5_INC, real, "-lz", "-Wl,-rpath,$ORIGIN/hdf5"]
    for src, exe in (("scp3_main.cpp", "SCP3"), ("eigensolver_main.cpp", "eigensolver"), ("scp1_main.cpp", "SCP1")):
        cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-pthread", os.path.join(CSRC, src), "-o", os.path.join(BIN, exe),
               "-L" + HERE, "-lscp3g", "-Wl,-rpath,$ORIGIN/.."] + extra
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("CLI build failed: %s\n%s" % (" ".join(cmd), r.stderr))
        if verbose:
            print("built", os.path.join(BIN, exe), "(HDF5)" if hdf5 else "(no HDF5)")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
