"""Builds libflint_amd.so in-tree: regenerates the mechanism kernels, compiles every HIP/C++
source for gfx950 with hipcc and links one shared library next to this file."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflint_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # keep hipcc from emitting packed FP32 instructions (v_pk_*_f32)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def sources():
    out = []
    for d in (CSRC, os.path.join(CSRC, "gen")):
        for f in sorted(os.listdir(d)):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(d, f))
    return out


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(BUILD, rel + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


FORTRAN = os.path.join(HERE, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose=False):
    """The iso_c_binding drop-in module (fortran/flint_amd.f90) and its check program, linked
    against libflint_amd.so; skipped when no flang is installed."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran module not built")
        return None
    out = os.path.join(FORTRAN, "_build")
    os.makedirs(out, exist_ok=True)
    mod = os.path.join(FORTRAN, "flint_amd.f90")
    obj = os.path.join(out, "flint_amd.o")
    exe = os.path.join(out, "drop_in_check")
    steps = [[FLANG, "-O2", "-fPIC", "-c", mod, "-module-dir", out, "-o", obj],
             [FLANG, "-O2", os.path.join(FORTRAN, "drop_in_check.f90"), obj, "-module-dir", out, "-I", out,
              "-L" + HERE, "-lflint_amd", "-Wl,-rpath,$ORIGIN/../..", "-o", exe + ".tmp"]]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=out)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", exe)
    return exe


def build(verbose=False, jobs=8):
    from .codegen import generate
    generate()
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True, jobs=int(sys.argv[1]) if len(sys.argv) > 1 else 8)
