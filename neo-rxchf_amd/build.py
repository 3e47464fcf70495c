"""In-tree build of libneorx_hip.so (gfx950) and its Fortran drop-in shim.

    python neo-rxchf_amd/build.py          # or __graft_entry__.build()

hipcc cross-compiles for gfx950 without a GPU present.  Each HIP translation
unit of csrc/ is compiled to an object in parallel, then linked.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

HIPFLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip():
    lib = os.path.join(HERE, "lib")
    objdir = os.path.join(lib, "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    procs, objs = [], []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [HIPCC] + HIPFLAGS + ["-c", "-o", obj, src]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    bad = [cmd for cmd, p in procs if p.wait() != 0]
    if bad:
        raise subprocess.CalledProcessError(1, bad[0])
    out = os.path.join(lib, "libneorx_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
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
    run([FLANG, "-O2", "-fdefault-integer-8", "-J", outdir, "-I", lib, "-o", out, src, "-L" + lib, "-lneorx_fshim",
         "-lneorx_hip", "-Wl,-rpath,$ORIGIN/../../neo-rxchf_amd/lib"])
    return out


def build_all():
    build_hip()
    build_shim()
    build_shim_driver()


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "hip":
        build_hip()
    else:
        build_all()
    sys.exit(0)
