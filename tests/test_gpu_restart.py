"""The restart file written from and read into device-resident arrays
(fc_restart_flow.on_device = 1: staged through pinned host buffers in
32 MiB chunks, copies overlapping the file I/O; the device arrays are
allocated through the library's own HIP runtime): the same bytes as the
host-array form, and a read that restores every array bit for bit.  The
larger case has grid arrays of several chunks."""
import numpy as np
import pytest

from test_restart import _empty_like, _same, pattern

pytestmark = pytest.mark.gpu


def _to_device(flow):
    from _hipmem import DeviceBuffer
    out = {}
    for k, v in flow.items():
        b = DeviceBuffer(v.nbytes)
        b.upload(v)
        out[k] = b
    return out


def _from_device(dflow, like):
    return {k: v.download(like[k]) for k, v in dflow.items()}


@pytest.mark.parametrize("dims", [(5, 4, 3), (170, 168, 166)])
def test_device_restart_matches_host(dims, capi, tmp_path):
    import feicar3d_amd.restart as rs
    nx, ny, nz = dims
    print("restart dims", dims, flush=True)
    flow, bodies = pattern(nx, ny, nz)
    host = tmp_path / "host.dat"
    dev = tmp_path / "dev.dat"
    rs.write_restart(host, nx, ny, nz, 3.5, 12, flow, bodies)
    dflow = _to_device(flow)
    rs.write_restart(dev, nx, ny, nz, 3.5, 12, dflow, bodies, on_device=True)
    assert host.read_bytes() == dev.read_bytes()
    f2, b2 = _empty_like(flow, bodies)
    d2 = _to_device(f2)
    t, nt = rs.read_restart(host, nx, ny, nz, d2, b2, on_device=True)
    assert (t, nt) == (3.5, 12)
    _same(flow, bodies, _from_device(d2, flow), b2)
