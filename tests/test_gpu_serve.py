"""The request server (fc_serve_*) through the C-ABI.

A request solved by the server, alone or batched with others that join and
leave at different steps, must give the same bits as fc_pson_solve on its
own handle with the same parameters (accelerated loop, both precisions):
the batched kernels give every request the block grid and reduction order
of its own launch.  The single-handle solve is itself pinned against the
oracle (test_gpu_solve.py), so this carries that parity over to serving.
Also: admission by device memory and the error for a request that cannot
fit the pool."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(20, 16, 24, 11), (24, 24, 18, 12), (16, 20, 20, 13), (28, 18, 16, 14)]


@pytest.fixture(scope="module")
def reqs(geo):
    return [geo.config_request(a, b, c, seed=s) for a, b, c, s in SHAPES]


def _params(capi, prec):
    # the server runs FP64 requests in the reference's form (V-cycles on p,
    # held masked cells), so the single-handle reference uses it too
    return capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-12, itermax_gc=50,
                               precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-9,
                               vc_walls=1 if prec == capi.PREC_MIXED else 0)


def _single(gpu, w, prm):
    s = gpu.PoissonSolver(*w.grid, prm)
    try:
        s.set_workload(w)
        p = np.zeros(w.div.shape, order="F")
        rc, st, _ = s.solve_poisson(p, w.div, np.zeros(w.div.shape, order="F"))
    finally:
        s.close()
    return rc, st, p


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_single_handle(capi, gpu, reqs, prec):
    prm = _params(capi, prec)
    srv = gpu.Server(pool_bytes=2 << 30, max_active=8)
    try:
        ids = [srv.submit(w, prm)[0] for w in reqs]
        res = srv.drain()
    finally:
        srv.close()
    for rid, w in zip(ids, reqs):
        rc, st, p = _single(gpu, w, prm)
        assert rc == 0 and st.converged == 1
        _, rc2, st2, out = res[rid]
        print("request %d: outer %d / %d, rel %.3e" % (rid, st.outer_iters, st2.outer_iters, st2.rel_resid))
        assert rc2 == 0 and st2.outer_iters == st.outer_iters
        assert np.array_equal(out, p)


def test_alone_and_in_a_mix_are_identical(capi, gpu, reqs):
    """Requests of both precisions arriving at different steps, batched with
    each other, against each request solved alone by a server."""
    prms = [_params(capi, k % 2) for k in range(len(reqs))]
    alone = []
    for w, prm in zip(reqs, prms):
        srv = gpu.Server(pool_bytes=1 << 30, max_active=1)
        try:
            rid, _ = srv.submit(w, prm)
            alone.append(srv.drain()[rid])
        finally:
            srv.close()
    srv = gpu.Server(pool_bytes=2 << 30, max_active=3)
    try:
        ids = [srv.submit(reqs[0], prms[0])[0], srv.submit(reqs[1], prms[1])[0]]
        res = {}
        for k in range(7):
            srv.step()
            while True:
                r = srv.poll()
                if r is None:
                    break
                res[r[0]] = r
        ids += [srv.submit(reqs[2], prms[2])[0], srv.submit(reqs[3], prms[3])[0]]
        res.update(srv.drain())
    finally:
        srv.close()
    for k, rid in enumerate(ids):
        a = alone[k]
        b = res[rid]
        assert a[1] == 0 and b[1] == 0
        assert a[2].outer_iters == b[2].outer_iters
        assert np.array_equal(a[3], b[3]), k


def test_admission_waits_for_memory(capi, gpu, reqs):
    """A pool that holds about two of these requests: the others wait in
    arrival order and every answer is still the one it gets alone."""
    prm = _params(capi, 1)
    # the pool's use once the largest request is admitted
    srv = gpu.Server(pool_bytes=1 << 30, max_active=1)
    try:
        srv.submit(reqs[1], prm)
        need = 0
        for _ in range(1000):
            st = srv.step()
            if st.active == 1:
                need = st.pool_used
                break
        assert need > 0
        srv.drain()
    finally:
        srv.close()
    pool = int(need * 2.6)
    srv = gpu.Server(pool_bytes=pool, max_active=8)
    try:
        ids = [srv.submit(w, prm)[0] for w in reqs]
        res = {}
        peak = 0
        for _ in range(5000):
            st = srv.step()
            peak = max(peak, st.active)
            assert st.pool_used <= st.pool_capacity
            while True:
                r = srv.poll()
                if r is None:
                    break
                res[r[0]] = r
            if st.active == 0 and st.queued == 0:
                break
    finally:
        srv.close()
    assert len(res) == len(reqs) and peak < len(reqs)
    for rid, w in zip(ids, reqs):
        rc, st1, p = _single(gpu, w, prm)
        assert res[rid][1] == 0 and np.array_equal(res[rid][3], p)


def test_request_larger_than_pool(capi, gpu, reqs):
    prm = _params(capi, 1)
    srv = gpu.Server(pool_bytes=1 << 20, max_active=2)
    try:
        rid, _ = srv.submit(reqs[0], prm)
        srv.step()
        r = srv.poll()
    finally:
        srv.close()
    assert r is not None and r[0] == rid and r[1] == -6


# requests at the bounds of the serving workload (32..192 cells per axis)
C4_SHAPES = [(32, 192, 97, 1001), (192, 40, 150, 1002), (160, 160, 32, 1003)]


@pytest.fixture(scope="module")
def c4reqs(geo):
    return [geo.config_request(a, b, c, seed=s) for a, b, c, s in C4_SHAPES]


def _flag_err(w, p):
    f = (w.iblank == 0) & (w.ghost == 0)
    m = np.zeros_like(f)
    m[1:-2, 1:-2, 1:-2] = True
    f &= m
    x = p[f] - p[f].mean()
    y = w.p_exact[f] - w.p_exact[f].mean()
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("prec", [1, 0])
def test_serving_sizes_exact_and_mix(capi, gpu, c4reqs, prec):
    """Requests of 32 to 192 cells per axis: batched together they give the
    bits each gets alone, and the pressure is within 1e-8 of the manufactured
    exact pressure with mixed precision at the relative-residual stop 1e-10,
    within 1e-10 in FP64 at the stop 1e-12."""
    prm = _params(capi, prec)
    prm.rel_tol = 1e-10 if prec == 1 else 1e-12
    alone = []
    for w in c4reqs:
        srv = gpu.Server(pool_bytes=8 << 30, max_active=1)
        try:
            rid, _ = srv.submit(w, prm)
            alone.append(srv.drain()[rid])
        finally:
            srv.close()
    srv = gpu.Server(pool_bytes=8 << 30, max_active=3)
    try:
        ids = [srv.submit(w, prm)[0] for w in c4reqs]
        res = srv.drain()
    finally:
        srv.close()
    for k, (rid, w) in enumerate(zip(ids, c4reqs)):
        a, b = alone[k], res[rid]
        err = _flag_err(w, b[3])
        print("request %s: outer %d, rel %.2e, vs exact %.2e" % (C4_SHAPES[k][:3], b[2].outer_iters, b[2].rel_resid,
                                                                err), flush=True)
        assert a[1] == 0 and b[1] == 0 and b[2].converged == 1
        assert a[2].outer_iters == b[2].outer_iters
        assert np.array_equal(a[3], b[3]), k
        assert err <= (1e-8 if prec == 1 else 1e-10)
