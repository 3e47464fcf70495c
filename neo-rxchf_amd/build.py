"""In-tree build of libneorx_hip.so (gfx950) and its Fortran drop-in shim.

    python neo-rxchf_amd/build.py          # or __graft_entry__.build()

hipcc cross-compiles for gfx950 without a GPU present.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

HIPFLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip():
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    out = os.path.join(HERE, "lib", "libneorx_hip.so")
    tmp = out + ".tmp"
    run([HIPCC] + HIPFLAGS + ["-o", tmp, os.path.join(HERE, "csrc", "nrx_lib.hip")])
    os.replace(tmp, out)
    return out


def build_shim():
    """iso_c_binding shim exporting the reference's subroutine names."""
    src = os.path.join(HERE, "fortran", "neorx_shim.f90")
    if not os.path.exists(src):
        return None
    lib = os.path.join(HERE, "lib")
    out = os.path.join(lib, "libneorx_fshim.so")
    run([FLANG, "-O2", "-fPIC", "-shared", "-fdefault-integer-8", "-J", lib, "-o", out, src,
         "-L" + lib, "-lneorx_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_shim_driver():
    """Fortran test program calling the drop-in subroutines (tests/fortran)."""
    src = os.path.join(ROOT, "tests", "fortran", "drive_shim.f90")
    if not os.path.exists(src):
        return None
    lib = os.path.join(HERE, "lib")
    outdir = os.path.join(ROOT, "tests", "_build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "drive_shim")
    run([FLANG, "-O2", "-fdefault-integer-8", "-J", outdir, "-o", out, src, "-L" + lib, "-lneorx_fshim",
         "-lneorx_hip", "-Wl,-rpath,$ORIGIN/../../neo-rxchf_amd/lib"])
    return out


def build_all():
    build_hip()
    build_shim()
    build_shim_driver()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
