"""Build step of the package: compiles libpg.so for gfx950 in-tree.

    python -m perigrain_amd.build

Every HIP source is compiled with -ffp-contract=off (each +,-,*,/ a single
IEEE binary64 rounding, as in the reference's FMA-free build) and with packed
FP32 instructions disabled.
"""
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

LIBPG_SOURCES = ["pg_kernels.hip", "pg_api.hip"]


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


def build_all():
    return [build_libpg()]


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
