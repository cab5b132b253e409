import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _pkg  # noqa: E402

_pkg.load_package()
_pkg.oracle_dir()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def capi():
    import feicar3d_amd.capi as m
    return m


@pytest.fixture(scope="session")
def geo():
    import feicar3d_amd.geometry as m
    return m


@pytest.fixture(scope="session")
def oracle():
    import fc_oracle
    fc_oracle.build()
    return fc_oracle


@pytest.fixture(scope="session")
def gpu():
    import feicar3d_amd.poisson as m
    m.library()
    return m


def interior(a):
    return a[1:-2, 1:-2, 1:-2]


def no_edges(a):
    """mask of the array without the edge ghosts (cells on two outer ghost
    planes), which no interior stencil reads"""
    sh = a.shape
    m = np.ones(sh, dtype=bool)
    gi = np.zeros(sh[0], bool); gi[[0, sh[0] - 2, sh[0] - 1]] = True
    gj = np.zeros(sh[1], bool); gj[[0, sh[1] - 2, sh[1] - 1]] = True
    gk = np.zeros(sh[2], bool); gk[[0, sh[2] - 2, sh[2] - 1]] = True
    cnt = gi[:, None, None].astype(int) + gj[None, :, None].astype(int) + gk[None, None, :].astype(int)
    m[cnt >= 2] = False
    # the index n+1 layers are never written by the reference
    m[-1, :, :] = False
    m[:, -1, :] = False
    m[:, :, -1] = False
    return m


@pytest.fixture(scope="session")
def geom():
    import feicar3d_amd.geom as m
    return m


@pytest.fixture(scope="session")
def surface():
    import feicar3d_amd.surface as m
    return m


@pytest.fixture(scope="session")
def geom_oracle():
    import fc_oracle
    fc_oracle.build()
    import fc_geom_oracle
    return fc_geom_oracle
