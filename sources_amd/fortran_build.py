"""Builds the Fortran host (module sphg + driver sph_gpu) with flang,
linked against the in-tree libsphg.so."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(HERE, "fortran")
EXE = os.path.join(FDIR, "sph_gpu")
EXE_CAS = os.path.join(FDIR, "sph_cas")


def flang():
    return shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


def build(verbose=False):
    out = os.path.join(FDIR, "_build")
    os.makedirs(out, exist_ok=True)
    lib_dir = HERE
    cmds = [
        [flang(), "-O2", "-c", os.path.join(FDIR, "sphg_mod.f90"), "-module-dir", out,
         "-o", os.path.join(out, "sphg_mod.o")],
        [flang(), "-O2", "-I", out, os.path.join(FDIR, "sph_main.f90"), os.path.join(out, "sphg_mod.o"),
         "-L", lib_dir, "-lsphg", f"-Wl,-rpath,$ORIGIN/..", "-o", EXE],
        [flang(), "-O2", "-c", "-I", out, os.path.join(FDIR, "sphg_io.f90"), "-module-dir", out,
         "-o", os.path.join(out, "sphg_io.o")],
        [flang(), "-O2", "-I", out, os.path.join(FDIR, "sph_cas.f90"), os.path.join(out, "sphg_io.o"),
         os.path.join(out, "sphg_mod.o"), "-L", lib_dir, "-lsphg", f"-Wl,-rpath,$ORIGIN/..", "-o", EXE_CAS],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=FDIR)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    return EXE
