"""The drop-in boundary without a GPU: the HIP library loads and exports every
entry point include/*.h declares (fc_poisson.h, fc_geometry.h, fc_restart.h), the ctypes
mirrors of the geometry structs have the C layout, and the ctypes and Fortran
mirrors of the header's structs have the C layout (sizes and offsets
printed by gcc and flang from the header and the module)."""
import ctypes as C
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "fc_poisson.h")
HEADERS = sorted(os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include"))
                 if f.endswith(".h"))


def _declared():
    names = []
    for h in HEADERS:
        for line in open(h):
            m = re.match(r"^[a-z][\w\s\*]*?\b(fc_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) == len(set(names))
    for must in ("fc_pson_create", "fc_pson_solve", "fc_pson_mg_solve", "fc_pson_ghostcell_pressure",
                 "fc_pson_itsolv", "fc_pson_calc_residual", "fc_pson_rhs", "fc_tdma_batched",
                 "fc_serve_create", "fc_serve_submit", "fc_serve_step", "fc_serve_poll", "fc_serve_destroy",
                 "fc_correct_vel", "fc_correct_ghostcell_velocity", "fc_face_vel", "fc_update_pressure",
                 "fc_project_resident", "fc_geom_create", "fc_geom_set_body", "fc_geom_move_body",
                 "fc_geom_refresh", "fc_geom_get_ghosts", "fc_geom_get_dead", "fc_geom_get_masks",
                 "fc_geom_set_marker_velocity", "fc_geom_body_intercept_values", "fc_geom_get_bi_values",
                 "fc_geom_marker_stress", "fc_geom_marker_stress_device", "fc_geom_get_marker_forces",
                 "fc_geom_get_stress_timing", "fc_restart_write", "fc_restart_read", "fc_restart_name",
                 "fc_geom_marker_stress_resident", "fc_face_vel_resident", "fc_pson_rhs_resident",
                 "fc_pson_resident_copy", "fc_geom_get_dead_ip"):
        assert must in names


def test_geometry_struct_layout(tmp_path):
    """sizeof / offsetof of the geometry and restart structs, printed by gcc
    from include/fc_geometry.h and fc_restart.h, against the ctypes mirrors
    in geom.py and restart.py."""
    import feicar3d_amd.geom as geom
    import feicar3d_amd.restart as rs
    src = tmp_path / "probe.c"
    fields = {"fc_geom_params": geom.FcGeomParams, "fc_geom_stats": geom.FcGeomStats,
              "fc_geom_ghosts": geom.FcGeomGhosts, "fc_geom_dead": geom.FcGeomDead,
              "fc_geom_surface": geom.FcGeomSurface, "fc_geom_forces": geom.FcGeomForces,
              "fc_restart_flow": rs.FcRestartFlow, "fc_restart_body": rs.FcRestartBody}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "fc_geometry.h"', '#include "fc_restart.h"',
             "int main(void) {"]
    for cname, py in fields.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-I" + os.path.join(REPO, "include"), str(src), "-o", str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, py in fields.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[cname + "." + f]) == getattr(py, f).offset, (cname, f)


def test_library_exports(capi):
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("HIP library not built: " + capi.LIB_PATH)
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.fc_pson_version.restype = C.c_char_p
    assert lib.fc_pson_version()


_C_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "fc_poisson.h"
#define F(T, f) printf(#T " " #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("fc_grid size %zu\n", sizeof(fc_grid));
  printf("fc_pson_params size %zu\n", sizeof(fc_pson_params));
  printf("fc_pson_stats size %zu\n", sizeof(fc_pson_stats));
  F(fc_grid, nx); F(fc_grid, x); F(fc_grid, z);
  F(fc_pson_params, it_solver_type); F(fc_pson_params, omega_pson); F(fc_pson_params, itermax_pson);
  F(fc_pson_params, restol_pson); F(fc_pson_params, iterInter); F(fc_pson_params, mgrids_z);
  F(fc_pson_params, restol_gcp); F(fc_pson_params, itermax_gc); F(fc_pson_params, mix_GC_form);
  F(fc_pson_params, is_PBC_homogeneous); F(fc_pson_params, pbc); F(fc_pson_params, ppp);
  F(fc_pson_params, bc_vel); F(fc_pson_params, infoconv); F(fc_pson_params, precision);
  F(fc_pson_params, outer_rule); F(fc_pson_params, rel_tol); F(fc_pson_params, fp32_cycles);
  F(fc_pson_params, aa_depth); F(fc_pson_params, launch_mode); F(fc_pson_params, cycle);
  F(fc_pson_params, vc_walls); F(fc_pson_params, coarse_solve); F(fc_pson_params, reserved);
  F(fc_pson_stats, outer_iters); F(fc_pson_stats, converged); F(fc_pson_stats, maxres);
  F(fc_pson_stats, rel_resid); F(fc_pson_stats, time_ms); F(fc_pson_stats, reserved);
  printf("fc_serve_config size %zu\n", sizeof(fc_serve_config));
  printf("fc_serve_request size %zu\n", sizeof(fc_serve_request));
  printf("fc_serve_status size %zu\n", sizeof(fc_serve_status));
  F(fc_serve_config, max_active); F(fc_serve_config, admit_threads); F(fc_serve_request, params); F(fc_serve_request, iblank);
  F(fc_serve_request, nGhost); F(fc_serve_request, nDead); F(fc_serve_request, xyzDCIP);
  F(fc_serve_request, pgrad); F(fc_serve_request, div); F(fc_serve_request, pPrime);
  F(fc_serve_status, steps); F(fc_serve_status, pool_capacity);
  return 0;
}
"""


def test_struct_layout_matches_header(capi, tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(_C_PROBE)
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    types = {"fc_grid": capi.FcGrid, "fc_pson_params": capi.FcParams, "fc_pson_stats": capi.FcStats,
             "fc_serve_config": capi.FcServeConfig, "fc_serve_request": capi.FcServeRequest,
             "fc_serve_status": capi.FcServeStatus}
    n = 0
    for line in out:
        if not line:
            continue
        t, f, v = line.split()
        cls = types[t]
        got = C.sizeof(cls) if f == "size" else getattr(cls, f).offset
        assert got == int(v), (t, f, got, v)
        n += 1
    assert n > 20


_F_PROBE = """
program probe
  use iso_c_binding
  use fc_poisson
  implicit none
  type(fc_params_t) :: p
  type(fc_stats_t) :: s
  print '(i0,1x,i0)', c_sizeof(p), c_sizeof(s)
end program probe
"""


def test_fortran_types_match_header(capi, tmp_path):
    import feicar3d_amd.build as bld
    mod = os.path.join(bld.OUT, "fc_poisson.mod")
    if not os.path.exists(mod):
        pytest.fail("Fortran module not built: " + mod)
    src = tmp_path / "probe.f90"
    src.write_text(_F_PROBE)
    exe = tmp_path / "fprobe"
    subprocess.check_call([bld.FLANG, "-I", bld.OUT, "-c", str(src), "-o", str(tmp_path / "probe.o")],
                          cwd=str(tmp_path))
    subprocess.check_call([bld.FLANG, str(tmp_path / "probe.o"), os.path.join(bld.OUT, "fc_poisson_mod.o"),
                           "-L" + bld.OUT, "-lfc_poisson", "-Wl,-rpath," + bld.OUT, "-o", str(exe)])
    a, b = subprocess.check_output([str(exe)], text=True).split()
    assert int(a) == C.sizeof(capi.FcParams) and int(b) == C.sizeof(capi.FcStats)
