"""Build recipe of the HIP library (hipcc, gfx950) and the Fortran module.

python feicar3d_3.0_febio_amd/build.py      -> csrc/_build/libfc_poisson.so
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(CSRC, "_build")
LIB = os.path.join(OUT, "libfc_poisson.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["fc_kernels.hip", "fc_poisson.hip", "fc_serve.hip", "fc_geometry.hip", "fc_restart.hip"]
# the geometry refresh rounds every product and sum on its own, in the
# reference's order (no fused multiply-add)
EXTRA_FLAGS = {"fc_geometry.hip": ["-ffp-contract=off"]}
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_library(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        cmd = [HIPCC] + HIP_FLAGS + EXTRA_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
FORT = os.path.join(PKG, "fortran")
DRIVER = os.path.join(OUT, "fc_driver")
STEP_DRIVER = os.path.join(OUT, "fc_step_driver")
REF_DRIVER = os.path.join(OUT, "fc_ref_driver")
RESTART_CHECK = os.path.join(OUT, "fc_restart_check")


def build_fortran(verbose=False):
    """MODULE fc_poisson and the fc_driver host program, linked against the
    HIP library with an rpath to it."""
    os.makedirs(OUT, exist_ok=True)
    mod_obj = os.path.join(OUT, "fc_poisson_mod.o")
    drv_obj = os.path.join(OUT, "fc_driver.o")
    for src, obj in (("fc_poisson_mod.f90", mod_obj), ("fc_driver.f90", drv_obj)):
        cmd = [FLANG, "-O2", "-module-dir", OUT, "-c", os.path.join(FORT, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
    cmd = [FLANG, "-O2", "-o", DRIVER + ".tmp", drv_obj, mod_obj, "-L" + OUT, "-lfc_poisson",
           "-Wl,-rpath," + OUT, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(DRIVER + ".tmp", DRIVER)
    # the time-step driver (projection and advection-diffusion wrappers)
    step_obj = os.path.join(OUT, "fc_step_driver.o")
    cmd = [FLANG, "-O2", "-module-dir", OUT, "-c", os.path.join(FORT, "fc_step_driver.f90"), "-o", step_obj]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    cmd = [FLANG, "-O2", "-o", STEP_DRIVER + ".tmp", step_obj, mod_obj, "-L" + OUT, "-lfc_poisson",
           "-Wl,-rpath," + OUT, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(STEP_DRIVER + ".tmp", STEP_DRIVER)
    # MODULE fc_poisson_ref (the reference's routine names) and its driver
    ref_obj = os.path.join(OUT, "fc_poisson_ref.o")
    rdrv_obj = os.path.join(OUT, "fc_ref_driver.o")
    for src, obj in (("fc_poisson_ref.f90", ref_obj), ("fc_ref_driver.f90", rdrv_obj)):
        cmd = [FLANG, "-O2", "-module-dir", OUT, "-c", os.path.join(FORT, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
    cmd = [FLANG, "-O2", "-o", REF_DRIVER + ".tmp", rdrv_obj, ref_obj, mod_obj, "-L" + OUT, "-lfc_poisson",
           "-Wl,-rpath," + OUT, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(REF_DRIVER + ".tmp", REF_DRIVER)
    # the restart-file check: Fortran's own unformatted I/O in the restart layout
    cmd = [FLANG, "-O1", "-o", RESTART_CHECK + ".tmp", os.path.join(FORT, "fc_restart_check.f90")]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(RESTART_CHECK + ".tmp", RESTART_CHECK)
    return DRIVER


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_fortran(verbose=True))
