"""Build recipe of libchs.so (hipcc for gfx950, in-tree) and of the Fortran
shim module.  Called by __graft_entry__.build(); also runnable as
`python -m chiralheisenberg_amd.build`."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libchs.so")
SOURCES = [os.path.join(HERE, "csrc", "chs_api.hip"), os.path.join(HERE, "csrc", "chs_rng.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CHS_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are switched off for all device code: the compiler
# would otherwise form v_pk_* FP32 operations from float2 arithmetic.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def build_lib(verbose=False):
    """Compile libchs.so in place; returns its path.

    The host random-number code is compiled by the system C++ compiler with
    floating-point contraction off, so that its Box-Muller Gaussians round
    exactly as the reference's (and the oracle's) do; the HIP sources go
    through hipcc."""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    rng_obj = os.path.join(bdir, "chs_rng.o")
    cxx = os.environ.get("CXX", "g++")
    cmd0 = [cxx, "-O2", "-fPIC", "-ffp-contract=off", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-c",
            SOURCES[1], "-o", rng_obj]
    tmp = LIB + ".tmp"
    # the object goes first: hipcc compiles every file after a .hip source as HIP
    cmd = [HIPCC] + HIP_FLAGS + ["-I" + os.path.join(ROOT, "include"), rng_obj, SOURCES[0], "-o", tmp]
    for c in (cmd0, cmd):
        if verbose:
            print(" ".join(c), flush=True)
        res = subprocess.run(c, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("%s failed:\n%s" % (c[0], res.stderr[-4000:]))
    shutil.move(tmp, LIB)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
SHIM_SIZES = (8, 8, 8)   # ksize, ksizet, kthird of the shim library the tests drive


def shim_lib_path(sizes=SHIM_SIZES):
    return os.path.join(HERE, "fortran", "libchs_shim_l%d_lt%d_ls%d.so" % sizes)


def build_fortran_shim(sizes=SHIM_SIZES, verbose=False):
    """Compile the Fortran drop-in binding (fortran/chs_shim.F90) for the
    given compile-time sizes, with the reference's FP64 default kinds
    (-fdefault-real-8), into a shared library linked against libchs.so.
    Returns its path, or None when no flang is available."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(HERE, "fortran", "chs_shim.F90")
    bdir = os.path.join(HERE, "_build", "shim_l%d_lt%d_ls%d" % sizes)
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, "chs_shim.o")
    out = shim_lib_path(sizes)
    cmds = [[FLANG, "-cpp", "-O2", "-fPIC", "-fdefault-real-8", "-fdefault-double-8",
             "-DKSIZE=%d" % sizes[0], "-DKSIZET=%d" % sizes[1], "-DKTHIRD=%d" % sizes[2],
             "-module-dir", bdir, "-c", src, "-o", obj],
            [FLANG, "-shared", "-o", out, obj, "-L" + HERE, "-lchs", "-Wl,-rpath,$ORIGIN/.."]]
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        res = subprocess.run(c, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("flang failed:\n" + res.stderr[-4000:])
    return out


if __name__ == "__main__":
    print(build_lib(verbose=True))
