"""MI355X (gfx950) pressure-Poisson hot path of FEICAR3D.

The directory name is not a Python identifier; load it with
``load_package()`` from this file's parent (tests/conftest.py, bench.py and
__graft_entry__.py do so) and it registers as ``feicar3d_amd``.

Contents: capi (ctypes mirror of include/fc_poisson.h), poisson (host-side
mirror of the reference's solve_poisson / mg_solve / GCM_ghostcell_pressure
interface over the HIP library), geometry (synthetic workloads), build
(hipcc recipe of the HIP library and the Fortran module).
"""
