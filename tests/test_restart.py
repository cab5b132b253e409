"""The restart file (include/fc_restart.h; write_restart / read_restart,
MAIN.F90:983-1122) against Fortran's own unformatted sequential I/O: the
program fc_restart_check (feicar3d_3.0_febio_amd/fortran, built with flang)
writes and reads the reference's record lists with WRITE and READ.

* the file fc_restart_write makes from a pattern is byte-identical to the
  one Fortran writes from the same pattern, and Fortran's READ finds the
  pattern in it;
* fc_restart_read reads the pattern from Fortran's file;
* the subrecord framing of records longer than the marker limit (the limit
  lowered so that small files need it): markers, signs and lengths as the
  compilers lay them out, and the round trip;
* wrong dimensions and truncated files are refused, not misread.

Host arrays only: no HIP call is made (the device form is in
test_gpu_restart.py)."""
import os
import struct
import subprocess

import numpy as np
import pytest

NX, NY, NZ = 5, 4, 3


def _check_prog():
    import feicar3d_amd.build as b
    if not os.path.exists(b.RESTART_CHECK):
        pytest.skip("fc_restart_check not built (python feicar3d_3.0_febio_amd/build.py)")
    return b.RESTART_CHECK


def pattern(nx=NX, ny=NY, nz=NZ):
    """the pattern fc_restart_check.f90 fills (its header states it)"""
    import feicar3d_amd.restart as rs
    shapes = rs.flow_shapes(nx, ny, nz)
    flow = {}
    for a, n in enumerate(rs.FLOW_NAMES[:26]):
        sh = shapes[n]
        q = np.arange(int(np.prod(sh)), dtype=np.float64)
        flow[n] = np.asfortranarray((1000.0 * (a + 1) + 0.5 * q).reshape(sh, order="F"))
    q = np.arange(int(np.prod(shapes["iblank"])))
    flow["iblank"] = np.asfortranarray(((q % 3) - 1).astype(np.int8).reshape(shapes["iblank"], order="F"))
    flow["ghostCellMark"] = np.asfortranarray((q % 2).astype(np.int8).reshape(shapes["iblank"], order="F"))
    bodies = []
    for b, (npts, nelm) in enumerate(((7, 4), (5, 3)), start=1):
        d = {}
        for qq, n in enumerate(rs.BODY_ARRAYS, start=1):
            d[n] = 100.0 * b + 10.0 * qq + 0.25 * np.arange(npts, dtype=np.float64)
        e, r = np.meshgrid(np.arange(nelm), np.arange(1, 4), indexing="xy")
        d["triElemNeig"] = np.asfortranarray((3 * e + r + b).astype(np.int32))
        for s, n in enumerate(rs.BODY_SCALARS, start=1):
            d[n] = 0.5 * (b + s)
        bodies.append(d)
    return flow, bodies


def _empty_like(flow, bodies):
    f = {k: np.asfortranarray(np.full_like(v, -7)) for k, v in flow.items()}
    bs = []
    for d in bodies:
        e = {k: (np.full_like(v, -7) if isinstance(v, np.ndarray) else 0.0) for k, v in d.items()}
        e["triElemNeig"] = np.asfortranarray(e["triElemNeig"])
        bs.append(e)
    return f, bs


def _same(flow, bodies, f2, b2):
    for k in flow:
        assert np.array_equal(flow[k], f2[k]), k
    for d, e in zip(bodies, b2):
        for k in d:
            assert np.array_equal(np.asarray(d[k]), np.asarray(e[k])), k


def test_restart_name(capi):
    import feicar3d_amd.restart as rs
    assert rs.restart_name(0) == "restart_0000.dat"
    assert rs.restart_name(37) == "restart_0037.dat"
    assert rs.restart_name(9999) == "restart_9999.dat"
    from feicar3d_amd.poisson import FcError
    with pytest.raises(FcError):
        rs.restart_name(10000)


def test_writer_bytes_equal_fortran(capi, tmp_path):
    import feicar3d_amd.restart as rs
    prog = _check_prog()
    fort = tmp_path / "fortran.dat"
    subprocess.run([prog, "write", str(fort), str(NX), str(NY), str(NZ)], check=True, timeout=60)
    flow, bodies = pattern()
    ours = tmp_path / "ours.dat"
    rs.write_restart(ours, NX, NY, NZ, 1.25, 77, flow, bodies)
    a, b = fort.read_bytes(), ours.read_bytes()
    assert len(a) == len(b) and a == b
    r = subprocess.run([prog, "read", str(ours), str(NX), str(NY), str(NZ)], timeout=60,
                       stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout


def test_reader_reads_fortran_file(capi, tmp_path):
    import feicar3d_amd.restart as rs
    prog = _check_prog()
    fort = tmp_path / "restart_0000.dat"
    subprocess.run([prog, "write", str(fort), str(NX), str(NY), str(NZ)], check=True, timeout=60)
    flow, bodies = pattern()
    f2, b2 = _empty_like(flow, bodies)
    t, nt = rs.read_restart(fort, NX, NY, NZ, f2, b2)
    assert t == 1.25 and nt == 77
    _same(flow, bodies, f2, b2)


def _records(data):
    """(leading, trailing, length) of every subrecord, walking the markers"""
    out, pos = [], 0
    while pos < len(data):
        lead = struct.unpack_from("<i", data, pos)[0]
        n = abs(lead)
        trail = struct.unpack_from("<i", data, pos + 4 + n)[0]
        out.append((lead, trail, n))
        pos += 8 + n
    assert pos == len(data)
    return out


def test_subrecords(capi, tmp_path):
    import feicar3d_amd.restart as rs
    L = rs.library()
    flow, bodies = pattern()
    one = tmp_path / "one.dat"
    rs.write_restart(one, NX, NY, NZ, 1.25, 77, flow, bodies)
    whole = _records(one.read_bytes())
    assert [r[2] for r in whole][:1] == [12] and all(r[0] == r[1] == r[2] for r in whole)
    limit = 1000
    old = L.fc_restart_set_max_subrecord(limit)
    try:
        split = tmp_path / "split.dat"
        rs.write_restart(split, NX, NY, NZ, 1.25, 77, flow, bodies)
        recs = _records(split.read_bytes())
        # regroup the subrecords into logical records: leading marker
        # negative while more follow, trailing negative after the first
        logical, cur = [], []
        for lead, trail, n in recs:
            assert n <= limit
            assert (trail < 0) == (len(cur) > 0)
            cur.append(n)
            if lead > 0 or n == 0:
                logical.append(cur)
                cur = []
            else:
                assert n == limit
        assert not cur
        assert [sum(c) for c in logical] == [r[2] for r in whole]
        assert max(len(c) for c in logical) > 1
        f2, b2 = _empty_like(flow, bodies)
        t, nt = rs.read_restart(split, NX, NY, NZ, f2, b2)
        assert (t, nt) == (1.25, 77)
        _same(flow, bodies, f2, b2)
    finally:
        L.fc_restart_set_max_subrecord(old)


def test_refuses_wrong_dims_and_truncation(capi, tmp_path):
    import feicar3d_amd.restart as rs
    from feicar3d_amd.poisson import FcError
    flow, bodies = pattern()
    path = tmp_path / "r.dat"
    rs.write_restart(path, NX, NY, NZ, 1.25, 77, flow, bodies)
    f3, b3 = pattern(NX + 1, NY, NZ)
    f3, b3 = _empty_like(f3, b3)
    with pytest.raises(FcError, match=r"\(-4\)"):
        rs.read_restart(path, NX + 1, NY, NZ, f3, b3)
    data = path.read_bytes()
    cut = tmp_path / "cut.dat"
    cut.write_bytes(data[:len(data) - 10])
    f2, b2 = _empty_like(flow, bodies)
    with pytest.raises(FcError):
        rs.read_restart(cut, NX, NY, NZ, f2, b2)
    bad = bytearray(data)
    bad[-1] ^= 0x01                             # the last trailing marker
    (tmp_path / "bad.dat").write_bytes(bytes(bad))
    with pytest.raises(FcError, match=r"\(-4\)"):
        rs.read_restart(tmp_path / "bad.dat", NX, NY, NZ, f2, b2)
    with pytest.raises(FcError, match=r"\(-7\)"):
        rs.read_restart(tmp_path / "missing.dat", NX, NY, NZ, f2, b2)
