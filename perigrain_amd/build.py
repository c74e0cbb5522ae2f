"""Build step of the package, in-tree:

  perigrain_amd/libpg.so          the HIP stepper (gfx950)
  perigrain_amd/libpgio.so        config / HDF5 file I/O of the drop-in
  perigrain_amd/bin/simulate2d    drop-ins for the reference's executables
  perigrain_amd/bin/simulate3d

    python -m perigrain_amd.build

Every HIP source is compiled with -ffp-contract=off (each +,-,*,/ a single
IEEE binary64 rounding, as in the reference's FMA-free build) and with packed
FP32 instructions disabled.  The HDF5 C library comes from the image
(/opt/conda); bin/hdf5/ holds a link to it so that the run-time search path
of the built files names that one library and nothing else of /opt/conda.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC", "-Wall", "-Wno-unused-result",
]

LIBPG_SOURCES = ["pg_kernels.hip", "pg_adjoint.hip", "pg_api.hip"]
HOST = os.path.join(HERE, "host")
BIN = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
HDF5_PREFIXES = [os.environ.get("PG_HDF5_PREFIX", ""), "/opt/conda", "/usr"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32-ops target feature and
    # says so; that line is expected and dropped here.
    lines = [ln for ln in p.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if lines:
        print("\n".join(lines), flush=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}")


def build_libpg():
    out = os.path.join(HERE, "libpg.so")
    srcs = [os.path.join(CSRC, s) for s in LIBPG_SOURCES]
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-I", INCLUDE, "-I", CSRC] + srcs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def find_hdf5():
    """(include dir, path of the versioned libhdf5 shared object)."""
    for pre in HDF5_PREFIXES:
        if not pre:
            continue
        inc = os.path.join(pre, "include")
        libs = sorted(glob.glob(os.path.join(pre, "lib", "libhdf5.so.[0-9]*")),
                      key=len)
        if os.path.exists(os.path.join(inc, "hdf5.h")) and libs:
            return inc, libs[0]
    raise RuntimeError("HDF5 C library (hdf5.h, libhdf5.so.N) not found; set PG_HDF5_PREFIX")


def build_io():
    inc, lib = find_hdf5()
    hdir = os.path.join(BIN, "hdf5")
    os.makedirs(hdir, exist_ok=True)
    link = os.path.join(hdir, os.path.basename(lib))
    if os.path.islink(link) or os.path.exists(link):
        os.remove(link)
    os.symlink(lib, link)
    soname = os.path.basename(lib)
    cxx = [CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE, "-I", inc]
    out_io = os.path.join(HERE, "libpgio.so")
    _run(cxx + ["-shared", os.path.join(HOST, "pgio.cpp"), "-o", out_io + ".tmp",
                "-L", hdir, "-l:" + soname, "-Wl,-rpath,$ORIGIN/bin/hdf5"])
    os.replace(out_io + ".tmp", out_io)
    outs = [out_io]
    for dim in (2, 3):
        exe = os.path.join(BIN, f"simulate{dim}d")
        _run(cxx + ["-pthread", f"-DPG_DIM={dim}", os.path.join(HOST, "simulate.cpp"), "-o", exe + ".tmp",
                    "-L", HERE, "-lpg", "-lpgio", "-L", hdir, "-l:" + soname,
                    "-Wl,-rpath,$ORIGIN/..:$ORIGIN/hdf5"])
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


def build_io_sanitized(out_dir):
    """libpgio built with the address and undefined-behaviour sanitizers into
    out_dir (host C++ only, for the CPU test that runs the file-side tests
    against it); returns the library's path."""
    inc, lib = find_hdf5()
    hdir = os.path.join(BIN, "hdf5")
    if not os.path.exists(os.path.join(hdir, os.path.basename(lib))):
        build_io()
    out = os.path.join(out_dir, "libpgio.so")
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-g", "-O1"]
    _run([CXX] + san + ["-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE, "-I", inc, "-shared",
                        os.path.join(HOST, "pgio.cpp"), "-o", out, "-L", hdir,
                        "-l:" + os.path.basename(lib), "-Wl,-rpath," + hdir])
    return out


def build_all():
    return [build_libpg()] + build_io()


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
