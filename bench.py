#!/usr/bin/env python3
"""Benchmark: 256^3 masked Poisson solves per second to 1e-8 relative
residual on one MI355X, mixed precision (FP32 V-cycles inside FP64 defect
correction) beside the plain FP64 GPU solve, with p50/p99 solve latency.

Workload: 256^3 cells, tanh-stretched grid, a sphere and a rotated ellipsoid
with GCM ghost/dead cells, manufactured right-hand side (its exact solution
is known and checked).  One step = one solve_poisson from a zero pressure to
flag-cell relative residual <= 1e-8 with the GCM relations satisfied to the
same tolerance, on buffers already resident in HBM.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tests"))
import _pkg  # noqa: E402

_pkg.load_package()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mean_free_err(a, b, mask):
    x = a[mask] - a[mask].mean()
    y = b[mask] - b[mask].mean()
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def run_mode(poisson, capi, w, prec, steps, warmup, rel_tol, aa_depth, launch_mode=0, mgrids=0, walls=1,
             host_steps=0, coarse=0):
    """steps timed solves on resident buffers (plus, with host_steps, that
    many through fc_pson_solve with host arrays: the PCIe-inclusive rate of
    the Fortran drop-in); walls: vc_walls of the correction cycles"""
    p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=500, restol_gcp=1e-12, itermax_gc=50,
                            precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=rel_tol, vc_walls=walls)
    p.aa_depth = aa_depth
    p.mgrids_x = p.mgrids_y = p.mgrids_z = mgrids
    p.coarse_solve = coarse if walls else 0
    p.launch_mode = launch_mode
    s = poisson.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    s.upload(np.zeros(w.div.shape, order="F"), w.div, np.zeros(w.div.shape, order="F"))
    for _ in range(warmup):
        s.reset()
        s.solve_resident()
    lat = []
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s.reset()
        t1 = time.perf_counter()
        rc, st = s.solve_resident()
        lat.append(time.perf_counter() - t1)
        stats.append(st.as_dict())
        if rc != 0:
            raise RuntimeError("solve did not converge: %s" % st.as_dict())
    total = time.perf_counter() - t0
    pres = s.download()
    host_s = None
    if host_steps:
        hp = np.zeros(w.div.shape, order="F")
        t1 = time.perf_counter()
        for _ in range(host_steps):
            hp[:] = 0.0
            s.solve_poisson(hp, w.div)
        host_s = (time.perf_counter() - t1) / host_steps
    # one more solve with HIP events around every finest-level sweep (on the
    # solver's stream); per-kernel events keep this solve out of the graph
    s.timing(True)
    s.reset()
    s.solve_resident()
    sweep_ms, launches = s.timing_read()
    s.timing(False)
    res = dict(total_s=total, lat=lat, stats=stats, sweep_ms=sweep_ms, launches=launches, pres=pres,
               bytes=s.device_bytes(), host_s=host_s)
    s.close()
    return res


def cpu_baseline(capi, w, n_iter_gpu, budget_s=20.0, mgrids=0):
    """The oracle's outer iteration (GCM update + one reference-ordering
    V-cycle + residuals, FP64, one host thread) timed on the 256^3 workload
    for a bounded number of iterations; solves/s = 1 / (seconds per
    iteration x outer iterations the FP64 GPU solve needed)."""
    _pkg.oracle_dir()
    import fc_oracle
    fc_oracle.build()
    p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=1, restol_gcp=0.0, itermax_gc=50)
    p.mgrids_x = p.mgrids_y = p.mgrids_z = mgrids
    t = time.perf_counter()
    o = fc_oracle.Oracle(capi, *w.grid, p, ordering=fc_oracle.ORDER_REFERENCE)
    o.set_workload(w)
    setup = time.perf_counter() - t
    pp = np.zeros(w.div.shape, order="F")
    times = []
    t0 = time.perf_counter()
    while True:
        t1 = time.perf_counter()
        o.solve_poisson(pp, w.div)
        times.append(time.perf_counter() - t1)
        if time.perf_counter() - t0 > budget_s or len(times) >= 50:
            break
    o.close()
    per_iter = float(np.median(times))
    return dict(value=1.0 / (per_iter * n_iter_gpu), unit="solves/s", cores=1, kind="port",
                sample="%d outer iterations (GCM update + 1 V-cycle, reference ordering, FP64) of the CPU "
                       "restatement on the 256^3 workload, %.2f s each, scaled by the %d outer iterations of "
                       "the FP64 GPU solve; setup %.1f s not counted" % (len(times), per_iter, n_iter_gpu, setup))


def kernel_only(capi, geo, poisson, args):
    """Finest-level FP32 half sweeps (all 3 directions x 2 colours, --steps
    rounds) on the 256^3 workload, then --steps device copies of 256 MiB
    (the calibration kernel for the HBM byte counters)."""
    import torch
    w = geo.config_c2(args.n)
    s = poisson.PoissonSolver(*w.grid, capi.default_params(precision=capi.PREC_MIXED))
    s.set_workload(w)
    rng = np.random.default_rng(3)
    sh = s.level_shape(1)
    s.lvl_write(1, 10, np.asfortranarray(rng.standard_normal(sh)))
    s.lvl_write(1, 11, np.asfortranarray(rng.standard_normal(sh)))
    for _ in range(args.steps):
        for d in range(3):
            for c in range(2):
                s.sweep(1, d, c, prec=capi.PREC_MIXED)
    s.close()
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"kernel_only": True, "half_sweeps": 6 * args.steps, "copy_bytes_each_way": a.numel() * 4,
                      "copies": args.steps}), flush=True)


def projection_bench(capi, geo, poisson, args):
    """The velocity projection after the solve (CORE_TIME_STEP.F90:211-222:
    correct_vel, correct_ghostcell_velocity, face_vel, update_pressure) on
    the n^3 workload, device resident.  One step = the four kernels on
    buffers already in HBM.  Prints its own JSON line."""
    w = geo.config_c2(args.n)
    nx, ny, nz = (a.size - 2 for a in w.grid)
    prm = capi.default_params()
    s = poisson.PoissonSolver(*w.grid, prm)
    s.set_workload(w)
    rng = np.random.default_rng(17)
    sh = w.shape
    f = lambda shape: np.asfortranarray(rng.standard_normal(shape))  # noqa: E731
    st = dict(P=f(sh), u=f(sh), v=f(sh), w=f(sh), fu=f(sh), fv=f(sh), fw=f(sh),
              bcxu=f((2, ny + 2, nz + 2)), bcyv=f((nx + 2, 2, nz + 2)), bczw=f((nx + 2, ny + 2, 2)))
    s.set_vel_bc(st["bcxu"], st["bcyv"], st["bczw"])
    s.vel_upload(st["u"], st["v"], st["w"], st["fu"], st["fv"], st["fw"])
    s.upload(pPrime=st["P"])
    dt = 1e-3
    for _ in range(args.warmup):
        s.project_resident(dt, dt)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.project_resident(dt, dt)
    total = time.perf_counter() - t0
    s.timing(True)
    ms = np.zeros(4)
    for _ in range(args.steps):
        s.project_resident(dt, dt)
        ms += np.array(s.project_timing())
    s.timing(False)
    ms /= args.steps
    nint = float((nx - 1) * (ny - 1) * (nz - 1))
    nfield = float((nx + 2) * (ny + 2) * (nz + 2))
    ng = int(w.gcm["iGhost"].size)
    algo = [57.0 * nint, 128.0 * ng, 48.0 * nint, 17.0 * nfield]
    names = ["k_correct_vel", "k_correct_ghost_vel", "k_face_vel", "k_update_pressure"]
    kern = {n: {"avg_ms": round(float(m), 5), "bytes": b, "GB_s": round(b / (m * 1e-3) / 1e9, 1) if m > 0 else None}
            for n, m, b in zip(names, ms, algo)}
    achieved = algo[0] / (ms[0] * 1e-3) / 1e9
    s.close()
    cpu = None
    if not args.no_cpu:
        _pkg.oracle_dir()
        import fc_oracle
        fc_oracle.build()
        o = fc_oracle.Oracle(capi, *w.grid, prm)
        o.set_workload(w)
        o.set_vel_bc(st["bcxu"], st["bcyv"], st["bczw"])
        u, v, ww, fu, fv, fw = (st[k].copy(order="F") for k in ("u", "v", "w", "fu", "fv", "fw"))
        P = st["P"].copy(order="F")
        p = np.zeros_like(P)
        times = []
        for _ in range(3):
            t = time.perf_counter()
            o.correct_vel(P, dt, u, v, ww)
            o.correct_ghostcell_velocity(P, dt, u, v, ww)
            o.face_vel(u, v, ww, fu, fv, fw)
            o.update_pressure(P, p)
            times.append(time.perf_counter() - t)
        o.close()
        cpu = dict(value=round(1.0 / float(np.median(times)), 4), unit="projection steps/s", cores=1, kind="port",
                   sample="3 steps of the oracle's correct_vel, correct_ghostcell_velocity, face_vel, "
                          "update_pressure on the same %d^3 state, median %.3f s" % (args.n, float(np.median(times))))
    out = {
        "metric": "%d^3 velocity projection steps/s (correct_vel, correct_ghostcell_velocity, face_vel, "
                  "update_pressure), FP64, device resident" % args.n,
        "value": round(args.steps / total, 3),
        "unit": "projection steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random velocities and pressure, seeded; the n^3 sphere + ellipsoid masks and ghost list)",
        "config": {"workload": "%d^3 cells tanh-stretched, sphere + rotated ellipsoid, %d ghost cells" % (args.n, ng)},
        "kernels": kern,
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s",
                         frac=round(achieved / 8000.0, 4),
                         traffic=row_traffic("k_correct_vel") if args.n == 256 else None, kernel="k_correct_vel",
                         bytes_per_launch=algo[0], avg_launch_ms=round(float(ms[0]), 5)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def ad_bench(capi, geo, poisson, args):
    """The advection-diffusion line solves (CORE_SOLVE_AD.F90 itsolv_ad: x-, y-
    and z-line solves of one velocity component with Crank-Nicolson
    advection coefficients) on the n^3 workload, device resident, Re = 200.
    One step = one itsolv_ad call.  Prints its own JSON line."""
    w = geo.config_c2(args.n)
    nx, ny, nz = (a.size - 2 for a in w.grid)
    reinv, dt = 1.0 / 200.0, 2e-3
    s = poisson.PoissonSolver(*w.grid, capi.default_params())
    s.set_workload(w)
    s.set_solve_ad(reinv, dt, 1, 1.0)
    rng = np.random.default_rng(23)
    sh = w.shape
    f = lambda: np.asfortranarray(rng.standard_normal(sh))  # noqa: E731
    st = dict(u=f(), v=f(), w=f(), fu=f(), fv=f(), fw=f(), nlu=f())
    s.vel_upload(st["u"], st["v"], st["w"], st["fu"], st["fv"], st["fw"])
    s.ad_upload(nlu=st["nlu"])
    for _ in range(args.warmup):
        s.itsolv_ad_resident(0, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.itsolv_ad_resident(0, 1)
    total = time.perf_counter() - t0
    s.timing(True)
    ms = []
    for _ in range(args.steps):
        s.itsolv_ad_resident(0, 1)
        ms.append(s.ad_timing())
    s.timing(False)
    resm, loc, _ = s.calc_residual_ad_resident(0)
    s.close()
    ncell = float((nx - 1) * (ny - 1) * (nz - 1))
    bytes_call = 147.0 * ncell
    avg = float(np.mean(ms))
    achieved = bytes_call / (avg * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        _pkg.oracle_dir()
        import fc_oracle
        fc_oracle.build()
        o = fc_oracle.Oracle(capi, *w.grid, capi.default_params(), ordering=fc_oracle.ORDER_REFERENCE)
        o.set_workload(w)
        o.set_solve_ad(reinv, dt, 1, 1.0)
        var = st["u"].copy(order="F")
        t = time.perf_counter()
        o.itsolv_ad(var, st["nlu"], st["fu"], st["fv"], st["fw"])
        el = time.perf_counter() - t
        o.close()
        cpu = dict(value=round(1.0 / el, 4), unit="itsolv_ad calls/s", cores=1, kind="port",
                   sample="1 itsolv_ad call of the oracle (reference line order, tdma) on the same %d^3 state, %.2f s"
                          % (args.n, el))
    out = {
        "metric": "%d^3 itsolv_ad calls/s (x-, y-, z-line solves of one velocity component, Crank-Nicolson "
                  "advection), FP64, device resident" % args.n,
        "value": round(args.steps / total, 3),
        "unit": "itsolv_ad calls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random velocities, face velocities and right-hand side, seeded; the n^3 masks)",
        "config": {"workload": "%d^3 cells tanh-stretched, sphere + rotated ellipsoid masks, Re 200, dt %g, "
                               "Crank-Nicolson" % (args.n, dt)},
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s",
                         frac=round(achieved / 8000.0, 4),
                         traffic=(round(sum(row_traffic("k_ad_sweep<%d, 5, 4>" % dd) for dd in range(3)) / 3)
                                  if args.n == 256 else None),
                         kernel="k_ad_sweep (six half sweeps per call)", bytes_per_launch=bytes_call / 6.0,
                         avg_call_ms=round(avg, 5)),
        "residual_after": resm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def rhs_ad_bench(capi, geo, poisson, args):
    """The explicit right-hand side of the advection-diffusion step
    (rhs_advec_diff, CORE_SET_SOLVERS.F90:10-61: hybrid upwind advection and
    diffusion of u, v and w) on the n^3 workload, device resident, Re = 200,
    alfa_upwind 0.5, viscTot = reinv.  One step = one rhs_advec_diff call.
    Prints its own JSON line."""
    w = geo.config_c2(args.n)
    nx, ny, nz = (a.size - 2 for a in w.grid)
    reinv, dt, alfa = 1.0 / 200.0, 2e-3, 0.5
    s = poisson.PoissonSolver(*w.grid, capi.default_params())
    s.set_workload(w)
    s.set_solve_ad(reinv, dt, 1, 1.0)
    rng = np.random.default_rng(29)
    sh = w.shape
    f = lambda: np.asfortranarray(rng.standard_normal(sh))  # noqa: E731
    st = dict(u=f(), v=f(), w=f(), fu=f(), fv=f(), fw=f())
    s.vel_upload(st["u"], st["v"], st["w"], st["fu"], st["fv"], st["fw"])
    for _ in range(args.warmup):
        s.rhs_advec_diff_resident(alfa, dt)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.rhs_advec_diff_resident(alfa, dt)
    total = time.perf_counter() - t0
    s.timing(True)
    ms = []
    for _ in range(args.steps):
        s.rhs_advec_diff_resident(alfa, dt)
        ms.append(s.ad_timing())
    s.timing(False)
    s.close()
    ncell = float((nx - 1) * (ny - 1) * (nz - 1))
    bytes_call = 72.0 * ncell
    avg = float(np.mean(ms))
    achieved = bytes_call / (avg * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        _pkg.oracle_dir()
        import fc_oracle
        fc_oracle.build()
        o = fc_oracle.Oracle(capi, *w.grid, capi.default_params())
        o.set_workload(w)
        o.set_solve_ad(reinv, dt, 1, 1.0)
        nl = [np.zeros(sh, order="F") for _ in range(3)]
        t = time.perf_counter()
        o.rhs_advec_diff(st["u"], st["v"], st["w"], st["fu"], st["fv"], st["fw"], alfa, dt, *nl)
        el = time.perf_counter() - t
        o.close()
        cpu = dict(value=round(1.0 / el, 4), unit="rhs_advec_diff calls/s", cores=1, kind="port",
                   sample="1 rhs_advec_diff call of the oracle (the reference's loops; bit-identical to the compiled "
                          "reference) on the same %d^3 state, %.2f s" % (args.n, el))
    out = {
        "metric": "%d^3 rhs_advec_diff calls/s (hybrid upwind advection + diffusion right-hand side of u, v, w), "
                  "FP64, device resident" % args.n,
        "value": round(args.steps / total, 3),
        "unit": "rhs_advec_diff calls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random velocities and face velocities, seeded; the n^3 grid)",
        "config": {"workload": "%d^3 cells tanh-stretched, Re 200, dt %g, alfa_upwind %g, viscTot = reinv"
                               % (args.n, dt, alfa)},
        "roofline": dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s",
                         frac=round(achieved / 8000.0, 4),
                         traffic=row_traffic("k_rhs_advec_diff") if args.n == 256 else None,
                         kernel="k_rhs_advec_diff", bytes_per_launch=bytes_call, avg_call_ms=round(avg, 5)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def geometry_workload(n=384, markers=101, seed=5):
    """The C5 shape of the moving-body refresh: n^3 cells tanh-stretched
    toward a 0.4 x 0.4 plate at the box centre (smallest spacing 1/1024 across
    the plate, 1/768 along it), the plate a membrane of (markers-1)^2 * 2
    triangles (101: 20000) with a smooth seeded perturbation of amplitude
    0.02 along its normal, membrane thickness = sqrt(dxMin^2 + dyMin^2 +
    dzMin^2)."""
    import feicar3d_amd.geometry as geo
    import feicar3d_amd.surface as surf
    x = geo.tanh_axis(n + 1, [0.5], 1.0 / 768.0, width=0.45)
    y = geo.tanh_axis(n + 1, [0.5], 1.0 / 1024.0, width=0.25)
    z = geo.tanh_axis(n + 1, [0.5], 1.0 / 768.0, width=0.45)
    plate = surf.plate((0.5, 0.5, 0.5), side=0.4, n=markers, amplitude=0.02, seed=seed, smooth=True)
    return x, y, z, plate, surf.membrane_thickness(x, y, z, 1.0)


def geometry_bench(capi, args):
    """The immersed-boundary geometry refresh of a moving body, one time step
    per step: the C5 plate pitching as 0.3 sin(2 pi t) (t = 0.01 k) is moved
    (host marker rotation, upload) and refreshed on the device
    (calculate_arclength_norm_ds, set_iblank_body_fast,
    GCM_set_internal_boundary).  Prints its own JSON line."""
    import feicar3d_amd.geom as geom
    import feicar3d_amd.surface as surf
    n = args.n if args.n != 256 else 384
    x, y, z, plate, tk = geometry_workload(n)
    c = (0.5, 0.5, 0.5)
    g = geom.GeometryRefresh(x, y, z, [plate], membrane_tkns=tk)
    bodies = [surf.pitch(plate, c, surf.pitch_angle(0.01 * k)) for k in range(args.warmup + 2 * args.steps)]
    for k in range(args.warmup):
        g.move_body(1, bodies[k])
        g.refresh()
    t0 = time.perf_counter()
    stats = []
    for k in range(args.warmup, args.warmup + args.steps):
        g.move_body(1, bodies[k])
        stats.append(g.refresh())
    total = time.perf_counter() - t0
    st = stats[-1]
    log("geometry: %d^3, %d markers, nGhost %d nDead %d undecided %d smoothed %d rounds %d, device %.2f ms"
        % (n, plate.npts, st.nGhost, st.nDead, st.n_undecided, st.n_smoothed, st.smooth_rounds, st.time_ms))
    # the closest-marker searches (k_classify, k_ghost_rel, k_dead_rel): 8 FP64
    # operations per cell-marker distance (3 differences, 3 squares, 2 sums),
    # counted over the markers the binned searches actually compared
    scan_ms = float(np.mean([s.scan_ms for s in stats]))
    pairs = float(np.mean([s.scan_pairs for s in stats]))
    flops = 8.0 * pairs
    achieved = flops / (scan_ms * 1e-3) / 1e12
    brute = float(st.n_undecided + st.nGhost + st.nDead) * plate.npts
    cpu = None
    if not args.no_cpu:
        _pkg.oracle_dir()
        import fc_oracle
        import fc_geom_oracle
        fc_oracle.build()
        r = g.results()
        o = fc_geom_oracle.GeomOracle(capi, geom, x, y, z, [bodies[args.warmup + args.steps - 1]], membrane_tkns=tk)
        idx = np.linspace(0, r["nGhost"] - 1, min(200, r["nGhost"])).astype(int)
        t = time.perf_counter()
        for q in idx:
            o.body_intercept(1, int(r["iGhost"][q]), int(r["jGhost"][q]), int(r["kGhost"][q]))
        per = (time.perf_counter() - t) / idx.size
        o.close()
        est = per * (st.n_undecided + st.nGhost + st.nDead)
        cpu = dict(value=round(1.0 / est, 5), unit="geometry refreshes/s", cores=1, kind="port",
                   sample="%d body intercepts (calc_bodyIntercept_Unstruc, the oracle's restatement, bit-identical to "
                          "the compiled reference) at ghost cells of the last step: %.3f ms each, times the %d "
                          "intercepts of one refresh" % (idx.size, 1e3 * per,
                                                         st.n_undecided + st.nGhost + st.nDead))
    g.close()
    out = {
        "metric": "%d^3 moving-plate geometry refreshes/s (surface normals, iblank, smoothing, ghost and dead cells, "
                  "image points, GCM coefficients), FP64" % n,
        "value": round(args.steps / total, 3),
        "unit": "geometry refreshes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed plate, prescribed pitching)",
        "config": {"workload": "%d^3 cells stretched to 1/1024 across and 1/768 along a 0.4 x 0.4 membrane plate of "
                               "%d triangles (smooth perturbation, amplitude 0.02), pitching 0.3 sin(2 pi t), t = 0.01 k; per step: marker "
                               "upload + refresh; nGhost %d, nDead %d" % (n, plate.nelem, st.nGhost, st.nDead)},
        "roofline": dict(bound="fp64", achieved=round(achieved, 3), peak=78.6, unit="TFLOP/s",
                         frac=round(achieved / 78.6, 4), traffic=None, kernel="closest-marker searches (k_classify, "
                         "k_ghost_rel, k_dead_rel)", flops_per_refresh=flops, pairs_per_refresh=pairs,
                         pairs_brute_force=brute, avg_scan_ms=round(scan_ms, 4),
                         device_ms=round(float(np.mean([s.time_ms for s in stats])), 4)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def drag_bench(capi, args):
    """The drag and lift of the moving body, one time step per step: the C5
    plate refreshed once at t = 0.05, then per step new marker velocities of
    the pitching motion (omega = 0.6 pi cos(2 pi t), t = 0.05 + 0.01 k,
    uploaded), GCM_SetBodyInterceptValues and compute_marker_stress on
    device-resident p, u, v, w, with the nine totals copied back.  Prints
    its own JSON line."""
    import torch
    import feicar3d_amd.geom as geom
    import feicar3d_amd.surface as surf
    n = args.n if args.n != 256 else 384
    x, y, z, plate, tk = geometry_workload(n)
    c = (0.5, 0.5, 0.5)
    body = surf.pitch(plate, c, surf.pitch_angle(0.05))
    g = geom.GeometryRefresh(x, y, z, [body], membrane_tkns=tk)
    st = g.refresh()
    shape = g.shape[::-1]
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(17)
    fields = [torch.rand(shape, dtype=torch.float64, device="cuda:0", generator=gen) * 2.0 - 1.0 for _ in range(4)]
    reinv = 1.0 / 200.0
    dt = 0.01

    def velocities(t):
        om = 0.3 * 2.0 * np.pi * np.cos(2.0 * np.pi * t)
        u, v = -om * (body.y - c[1]), om * (body.x - c[0])
        return u, v, np.zeros_like(u)

    def step(k):
        t = 0.05 + 0.01 * k
        g.set_marker_velocity(1, *velocities(t), *velocities(t - dt))
        g.body_intercept_values(dt, 2)
        return g.marker_stress_device(*fields, reinv)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timing = []
    for k in range(args.warmup, args.warmup + args.steps):
        tot = step(k)
        timing.append(g.stress_timing())
    total = time.perf_counter() - t0
    bi_ms, interp_ms, sum_ms = (float(np.mean([t[q] for t in timing])) for q in range(3))
    log("drag: %d^3, %d elements, nGhost %d nDead %d; device BI values %.4f ms, element stresses %.4f ms, "
        "totals + marker forces %.4f ms; totals %s" % (n, plate.nelem, st.nGhost, st.nDead, bi_ms, interp_ms, sum_ms,
                                                        np.array2string(tot[0], precision=4)))
    # k_marker_stress, algorithmic bytes per element: centroid, normal and area
    # (7 doubles), the 8 pressures and 8 iblank/dead_cell pairs of the host
    # cell, the nearest dead cell's 8 coefficients, its 6 index words, probe
    # length, 3 intercept velocities and the 3 velocities of 8 cells (24
    # doubles, the cell itself included), the 6 force words written
    per_elem = 8 * (7 + 8 + 8 + 1 + 3 + 24 + 6) + 8 * 2 + 6 * 4
    traffic = per_elem * plate.nelem
    achieved = traffic / (interp_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        _pkg.oracle_dir()
        import fc_oracle
        import fc_geom_oracle
        fc_oracle.build()
        log("drag: oracle refresh at %d^3 (not timed) ..." % n)
        o = fc_geom_oracle.GeomOracle(capi, geom, x, y, z, [body], membrane_tkns=tk)
        o.refresh()
        k = args.warmup + args.steps - 1
        t = 0.05 + 0.01 * k
        o.set_marker_velocity(1, *velocities(t), *velocities(t - dt))
        hf = [np.asfortranarray(f.cpu().numpy().transpose(2, 1, 0)) for f in fields]
        tm = time.perf_counter()
        o.body_intercept_values(dt, 2)
        to = o.marker_stress(*hf, reinv)
        per = time.perf_counter() - tm
        o.close()
        log("drag: oracle totals %s (GPU %s)" % (np.array2string(to[0], precision=4),
                                                  np.array2string(tot[0], precision=4)))
        cpu = dict(value=round(1.0 / per, 4), unit="drag/lift evaluations/s", cores=1, kind="port",
                   sample="one full evaluation (og_body_intercept_values + og_marker_stress, the oracle's "
                          "restatement of GCM_SetBodyInterceptValues and compute_marker_stress) on the last step's "
                          "inputs: %.2f s; totals equal to the GPU's: %s" % (per, bool(np.array_equal(to, tot))))
    g.close()
    out = {
        "metric": "%d^3 moving-plate drag/lift evaluations/s (body-intercept velocities and dp/dn, element pressure "
                  "and viscous forces, per-marker forces, body totals), FP64" % n,
        "value": round(args.steps / total, 3),
        "unit": "drag/lift evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded plate, prescribed pitching velocities, uniform random p, u, v, w on the device)",
        "config": {"workload": "%d^3 cells stretched to 1/1024 across and 1/768 along a 0.4 x 0.4 membrane plate of "
                               "%d triangles pitched to 0.3 sin(0.1 pi); per step: marker velocities upload, "
                               "body-intercept values of %d ghost and %d dead cells, marker stresses, totals "
                               "download" % (n, plate.nelem, st.nGhost, st.nDead)},
        "roofline": dict(bound="hbm", achieved=round(achieved, 2), peak=8000.0, unit="GB/s",
                         frac=round(achieved / 8000.0, 5), traffic=None, kernel="k_marker_stress",
                         bytes_per_launch=traffic, avg_ms=round(interp_ms, 5), bi_values_ms=round(bi_ms, 5),
                         totals_and_marker_forces_ms=round(sum_ms, 5)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def c5_bench(capi, args):
    """The full fractional step of C5: flow at Re = 200 past the plate of
    geometry_workload pitching as 0.3 sin(2 pi t), one time step per step
    (timestep.FractionalStep: geometry refresh, body-intercept values,
    ghost-cell velocities, rhs_advec_diff and solve_ad, rhs_poisson and the
    mixed-precision accelerated solve_poisson, projection, marker stresses).
    Freestream u = 1 on every face, p = 0 on the x = 1 face.  Prints its own
    JSON line."""
    import feicar3d_amd.geom as geom
    import feicar3d_amd.poisson as poisson
    import feicar3d_amd.timestep as ts
    n = args.n if args.n != 256 else 384
    x, y, z, plate, tk = geometry_workload(n)
    sh = (x.size, y.size, z.size)
    c = (0.5, 0.5, 0.5)
    dt = 1.0e-3
    pbc = [capi.PBC_NEUMANN, capi.PBC_DIRICHLET] + [capi.PBC_NEUMANN] * 4
    prm = capi.default_params(bc_vel=[1] * 6, pbc=pbc, restol_pson=0.0, iterInter=0, itermax_pson=1500,
                              restol_gcp=1e-10, itermax_gc=200, outer_rule=capi.OUTER_RELRES, rel_tol=args.rel_tol,
                              precision=capi.PREC_FP64 if args.c5_fp64 else capi.PREC_MIXED)
    # automatic level count (MG_initial's)
    prm.mgrids_x = prm.mgrids_y = prm.mgrids_z = 0
    prm.aa_depth = args.aa
    t = time.perf_counter()
    flow = poisson.PoissonSolver(x, y, z, prm)
    flow.vel_set_outer(np.array([1.0, 1.0, 0.0, 0.0, 0.0, 0.0] * 3))
    s0 = ts.pitching_plate_motion(plate, c, 0.0, dt)[0]
    g = geom.GeometryRefresh(x, y, z, [s0], membrane_tkns=tk)
    vbc = (np.ones((2, sh[1], sh[2])), np.zeros((sh[0], 2, sh[2])), np.zeros((sh[0], sh[1], 2)))
    fs = ts.FractionalStep(flow, g, sh, 1.0 / 200.0, dt, 2.0, itermax_ad=100, restol_ad=1e-9, restol_gcv=1e-10,
                           u0=(1.0, 0.0, 0.0), vel_bc=vbc, resident=not args.c5_host)
    log("c5: %d^3 set up in %.1f s" % (n, time.perf_counter() - t))
    outs = []

    def one():
        tt = fs.time + dt
        s, mv = ts.pitching_plate_motion(plate, c, tt, dt)
        r = fs.step([s], [mv])
        st = r["poisson"]
        log("c5 step %d: %.2f s; nGhost %d nDead %d; AD passes %d (res %.1e); Poisson rc %d, %d outer, rel %.1e; "
            "cx %.4f cy %.4f; phases %s" % (r["ntime"], sum(r["phases_s"].values()), r["nGhost"], r["nDead"],
                                            r["ad_iters"], r["ad_res"], r["poisson_rc"], st.outer_iters,
                                            st.rel_resid, r["totals"][0][0], r["totals"][0][1],
                                            {k: round(v, 3) for k, v in r["phases_s"].items()}))
        if r["poisson_rc"] != 0 or r["ad_rc"] != 0:
            raise RuntimeError("a C5 solve did not converge (Poisson rc %d, solve_ad rc %d)"
                               % (r["poisson_rc"], r["ad_rc"]))
        return r

    for _ in range(args.warmup):
        one()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs.append(one())
    total = time.perf_counter() - t0
    phases = {k: round(float(np.mean([o["phases_s"][k] for o in outs])), 4) for k in outs[0]["phases_s"]}
    last = outs[-1]
    flow.close()
    g.close()
    out = {
        "metric": "%d^3 moving-plate fractional steps/s (C5: geometry refresh, advection-diffusion, %s "
                  "Poisson, projection, drag/lift), FP64 state" % (n, "FP64" if args.c5_fp64 else "mixed-precision"),
        "value": round(args.steps / total, 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded plate, prescribed pitching, freestream start)",
        "config": {"workload": "%d^3 cells stretched to 1/1024 across and 1/768 along a 0.4 x 0.4 membrane plate of "
                               "%d triangles pitching 0.3 sin(2 pi t), Re 200, dt %.0e, freestream u = 1 on every "
                               "face, p = 0 at x = 1; Poisson: %s, accelerated, rel %.0e, %s levels; flow state %s"
                               % (n, plate.nelem, dt, "FP64" if args.c5_fp64 else "mixed", args.rel_tol,
                                  "automatic",
                                  "on the host between calls (PCIe included)" if args.c5_host else
                                  "resident in HBM (masks and lists through the host each step)"),
                   "phases_s_per_step": phases,
                   "poisson_outer_iters": [o["poisson"].outer_iters for o in outs],
                   "ad_passes": [o["ad_iters"] for o in outs],
                   "last_totals": [round(float(v), 6) for v in last["totals"][0]]},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def restart_bench(capi, args):
    """write_restart / read_restart of a 384^3 flow state held on the device
    (fc_restart_write / fc_restart_read with device arrays: staged through
    pinned memory, copies overlapping the file I/O), to a file under the
    temporary directory.  Prints its own JSON line."""
    import tempfile
    import torch
    import feicar3d_amd.restart as rs
    n = args.n if args.n != 256 else 384
    torch.cuda.init()
    shapes = rs.flow_shapes(n, n, n)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(3)
    flow = {}
    for k in rs.FLOW_NAMES:
        sh = shapes[k][::-1]
        if k in ("iblank", "ghostCellMark"):
            flow[k] = torch.randint(0, 2, sh, dtype=torch.int8, device="cuda:0", generator=gen)
        else:
            flow[k] = torch.rand(sh, dtype=torch.float64, device="cuda:0", generator=gen)
    m = 10201
    body = {a: np.linspace(0.0, 1.0, m) for a in rs.BODY_ARRAYS}
    body["triElemNeig"] = np.asfortranarray(np.ones((3, 20000), dtype=np.int32))
    nbytes = sum(v.numel() * v.element_size() for v in flow.values())
    path = os.path.join(tempfile.mkdtemp(), rs.restart_name(0))
    torch.cuda.synchronize()
    wt, rt = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        rs.write_restart(path, n, n, n, 1.0, 10, flow, [body], on_device=True)
        wt.append(time.perf_counter() - t0)
        back = {k: torch.empty_like(v) for k, v in flow.items()}
        t0 = time.perf_counter()
        rs.read_restart(path, n, n, n, back, [dict(body)], on_device=True)
        torch.cuda.synchronize()
        rt.append(time.perf_counter() - t0)
    same = all(torch.equal(flow[k], back[k]) for k in flow)
    fsize = os.path.getsize(path)
    os.remove(path)
    w, r = float(np.median(wt)), float(np.median(rt))
    log("restart: %d^3, %.2f GB per file; write %.3f s (%.2f GB/s), read %.3f s (%.2f GB/s); read == written: %s"
        % (n, fsize / 1e9, w, fsize / w / 1e9, r, fsize / r / 1e9, same))
    out = {
        "metric": "%d^3 restart file write (write_restart) from device-resident state, GB/s" % n,
        "value": round(fsize / w / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": round(1e3 * w, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform random fields on the device)",
        "config": {"workload": "restart_0000.dat of a %d^3 rank: 26 real*8 and 2 integer*1 arrays, one 10201-marker "
                               "body; %d bytes of device arrays; file on the temporary directory's file system" %
                               (n, nbytes),
                   "read_GBps": round(fsize / r / 1e9, 3), "file_bytes": fsize, "round_trip_equal": bool(same)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def _gen_request(args):
    import feicar3d_amd.geometry as geo
    nx, ny, nz, seed = args
    return geo.config_request(nx, ny, nz, seed)


def serve_payloads(n_distinct, seed=4, lo=32, hi=192, workers=16):
    """n_distinct seeded serving requests: cells per axis uniform in
    [lo, hi], one random sphere or ellipsoid each, manufactured right-hand
    side.  Built in forked worker processes before any GPU call."""
    import multiprocessing as mp
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(n_distinct, 3))
    jobs = [(int(a), int(b), int(c), 1000 + k) for k, (a, b, c) in enumerate(dims)]
    with mp.get_context("fork").Pool(min(workers, n_distinct)) as pool:
        return pool.map(_gen_request, jobs)


def serve_run(capi, poisson, payloads, prec, n_req, rate, seed, max_active, pool_bytes, rel_tol, mgrids=0):
    """Open-loop serving: n_req requests with exponential inter-arrival times
    (mean 1/rate, seeded), request r using payload r % len(payloads); each is
    submitted at its arrival time and stepped with whatever else is resident.
    Latency = retirement time - scheduled arrival time (host clock)."""
    prm = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=600, restol_gcp=1e-12, itermax_gc=50,
                              precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=rel_tol)
    prm.mgrids_x = prm.mgrids_y = prm.mgrids_z = mgrids
    rng = np.random.default_rng(seed)
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    srv = poisson.Server(pool_bytes=pool_bytes, max_active=max_active)
    lat, iters, rel, act = [], [], [], []
    sub = {}
    nsub = 0
    t0 = time.perf_counter()
    t_last = t0
    try:
        while len(lat) < n_req:
            now = time.perf_counter() - t0
            while nsub < n_req and arrive[nsub] <= now:
                w = payloads[nsub % len(payloads)]
                rid, _ = srv.submit(w, prm, out=np.empty(w.div.shape, order="F"))
                sub[rid] = arrive[nsub]
                nsub += 1
            st = srv.step()
            if st.active == 0 and st.queued == 0 and nsub < n_req:
                time.sleep(max(0.0, arrive[nsub] - (time.perf_counter() - t0)))
            act.append(st.active)
            while True:
                r = srv.poll()
                if r is None:
                    break
                rid, rc, stt, _ = r
                if rc < 0:
                    raise RuntimeError("serving request %d failed (%d)" % (rid, rc))
                t_last = time.perf_counter()
                lat.append((t_last - t0 - sub.pop(rid)) * 1e3)
                iters.append(stt.outer_iters)
                rel.append(stt.rel_resid)
    finally:
        srv.close()
    span = t_last - t0 - arrive[0]
    lat = np.array(lat)
    return {"requests": n_req, "throughput_req_s": round(n_req / span, 2), "span_s": round(span, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 1), "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "mean_active": round(float(np.mean(act)), 1), "outer_iters_mean": round(float(np.mean(iters)), 1),
            "outer_iters_max": int(np.max(iters)), "rel_resid_max": float(np.max(rel))}


def row_traffic(kernel):
    """HBM bytes per launch of a projection or advection-diffusion kernel at
    256^3 from the committed counter summary (profiles/projection_ad_traffic.json),
    or None."""
    f = os.path.join(REPO, "profiles", "projection_ad_traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        e = json.load(fh).get(kernel)
    return None if e is None else e["traffic_bytes"]


def sweep_traffic():
    """HBM bytes per finest-level half sweep from the committed counter
    summary (profiles/sweep_traffic.json), or None."""
    f = os.path.join(REPO, "profiles", "sweep_traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--config", choices=["c2", "c3"], default="c2",
                    help="c2: the n^3 sphere + ellipsoid workload (the headline); c3: the 512 x 384 x 384 "
                         "semi-coarsened plate workload (a parity and scale case, not the headline)")
    ap.add_argument("--rel-tol", type=float, default=1e-10,
                    help="stop rule: flag-cell relative residual and max outer step relative to max|p| (1e-10: the "
                         "pressure then lies within 1e-8 of the exact and of the FP64 solve at 256^3; 1e-8 leaves "
                         "4e-7)")
    ap.add_argument("--fp64-tol", type=float, default=1e-12,
                    help="stop rule of the plain FP64 drop-in at its own accuracy target (pressure within 1e-10)")
    ap.add_argument("--aa", type=int, default=0, help="Anderson depth (0 = library default)")
    ap.add_argument("--mgrids", type=int, default=None,
                    help="multigrid levels per axis for both precisions (mgrids_x/y/z of input.dat; 0 = automatic, "
                         "as MG_initial: 7 at 256^3).  Default 4 for c2, the fastest measured at 256^3 (the GCM "
                         "coupling, not the coarse grids, sets the outer iteration count); automatic for c3, whose "
                         "point is the semi-coarsened hierarchy")
    ap.add_argument("--coarse-solve", type=int, default=0,
                    help="1: FP64 direct solve of the coarsest level of the wall correction cycles (needs a coarsest "
                         "level of at most 1024 free cells, e.g. --mgrids 6 or 0 at 256^3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--launch-mode", type=int, default=0,
                    help="1: launch every kernel from the host instead of one HIP graph per outer iteration")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (request stream) measurement")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean arrivals per second")
    ap.add_argument("--serve-distinct", type=int, default=64, help="distinct request payloads generated")
    ap.add_argument("--serve-active", type=int, default=128, help="requests batched together")
    ap.add_argument("--serve-mgrids", type=int, default=0,
                    help="multigrid levels per axis of every request (0 = automatic, as MG_initial)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving measurement (prints its dict)")
    ap.add_argument("--ad", action="store_true",
                    help="measure the advection-diffusion line solves (itsolv_ad) instead of the solve")
    ap.add_argument("--rhs-ad", action="store_true",
                    help="measure the advection-diffusion right-hand side (rhs_advec_diff) instead of the solve")
    ap.add_argument("--geometry", action="store_true",
                    help="measure the moving-body geometry refresh (C5 plate, 384^3) instead of the solve")
    ap.add_argument("--drag", action="store_true",
                    help="measure the moving-body drag and lift (body-intercept values and marker stresses)")
    ap.add_argument("--restart", action="store_true",
                    help="measure the restart file write/read from device-resident state (384^3)")
    ap.add_argument("--c5", action="store_true",
                    help="measure the full fractional step of the pitching plate (C5, 384^3)")
    ap.add_argument("--c5-fp64", action="store_true", help="with --c5: the Poisson solve in FP64")
    ap.add_argument("--c5-host", action="store_true",
                    help="with --c5: the flow state on the host between calls (the reference's call sites)")
    ap.add_argument("--projection", action="store_true",
                    help="measure the velocity projection after the solve instead of the solve")
    ap.add_argument("--kernel-only", action="store_true",
                    help="run only finest-level FP32 half sweeps and a 256 MiB device copy (for counter passes)")
    args = ap.parse_args()
    if args.mgrids is None:
        args.mgrids = 0 if args.config == "c3" else 4
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the Poisson path is a single-domain solve on one MI355X "
              "(no halo exchange or sharding is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import feicar3d_amd.capi as capi
    import feicar3d_amd.geometry as geo
    import feicar3d_amd.poisson as poisson
    if args.kernel_only:
        kernel_only(capi, geo, poisson, args)
        return
    if args.geometry:
        geometry_bench(capi, args)
        return
    if args.drag:
        drag_bench(capi, args)
        return
    if args.c5:
        c5_bench(capi, args)
        return
    if args.restart:
        restart_bench(capi, args)
        return
    if args.projection:
        projection_bench(capi, geo, poisson, args)
        return
    if args.ad:
        ad_bench(capi, geo, poisson, args)
        return
    if args.rhs_ad:
        rhs_ad_bench(capi, geo, poisson, args)
        return

    payloads = None
    if not args.no_serve:
        t = time.perf_counter()
        payloads = serve_payloads(args.serve_distinct)
        log("serving: %d distinct request payloads built in %.1f s, mean cells %.0f" % (
            len(payloads), time.perf_counter() - t, np.mean([np.prod(p.nodes()) for p in payloads])))
    if args.serve_only:
        for name, prec in (("mixed", capi.PREC_MIXED), ("fp64", capi.PREC_FP64)):
            r = serve_run(capi, poisson, payloads, prec, args.serve_requests, args.serve_rate, 4, args.serve_active,
                          96 << 30, args.rel_tol, args.serve_mgrids)
            print(json.dumps({"serving": name, "mgrids": args.serve_mgrids, **{k: float(v) if hasattr(v, "item")
                              else v for k, v in r.items()}}), flush=True)
        return
    t = time.perf_counter()
    if args.config == "c3":
        w = geo.config_c3()
        shape = "512x384x384"
        desc = ("512x384x384 cells tanh-stretched (x one level deeper than y, z, then held), thin tilted plate "
                "0.4 wide, GCM ghost/dead cells, Neumann box, rel resid %.0e" % args.rel_tol)
    else:
        w = geo.config_c2(args.n)
        shape = "%d^3" % args.n
        hx = float(np.min(np.diff(w.grid[0][1:-1])))
        desc = ("%d^3 cells tanh-stretched (min spacing 1/%.0f), sphere r=0.15 + ellipsoid 0.2/0.1/0.05 rot 30deg, "
                "GCM ghost/dead cells, Neumann box, manufactured RHS (exact pressure known), %s multigrid levels, "
                "stop at rel resid %.0e" % (args.n, 1.0 / hx, args.mgrids or "automatic", args.rel_tol))
    log("workload %s built in %.1f s: nGhost %d nDead %d" % (w.info["name"], time.perf_counter() - t,
                                                             w.info["nGhost"], w.info["nDead"]))
    n1 = float(np.prod([a.size - 3 for a in w.grid]))
    mask = (w.iblank == 0) & (w.ghost == 0)
    mask[0] = mask[-2:] = False
    mask[:, 0] = mask[:, -2:] = False
    mask[:, :, 0] = mask[:, :, -2:] = False

    mixed = run_mode(poisson, capi, w, capi.PREC_MIXED, args.steps, args.warmup, args.rel_tol, args.aa,
                     args.launch_mode, args.mgrids, walls=1, host_steps=2, coarse=args.coarse_solve)
    log("mixed: %.3f s for %d solves, outer iterations %s" % (mixed["total_s"], args.steps,
                                                            [s["outer_iters"] for s in mixed["stats"]]))
    k64 = args.fp64_steps if args.fp64_steps is not None else max(2, args.steps // 2)
    # the plain FP64 GPU solve: the reference's form (V-cycles on p with the
    # masked cells held, GCM update, Anderson-accelerated outer loop), at the
    # mixed solve's stop rule and at the FP64 drop-in's own (1e-10 pressure)
    fp64 = run_mode(poisson, capi, w, capi.PREC_FP64, k64, 1, args.rel_tol, args.aa, args.launch_mode,
                    args.mgrids, walls=0)
    log("fp64 (reference form): %.3f s for %d solves, outer iterations %s" % (
        fp64["total_s"], k64, [s["outer_iters"] for s in fp64["stats"]]))
    # the same algorithm as the mixed solve in FP64 (wall correction cycles,
    # FP64 defect correction): what the precision alone buys
    fp64w = run_mode(poisson, capi, w, capi.PREC_FP64, k64, 1, args.rel_tol, args.aa, args.launch_mode,
                     args.mgrids, walls=1, coarse=args.coarse_solve)
    log("fp64 (wall cycles): %.3f s for %d solves, outer iterations %s" % (
        fp64w["total_s"], k64, [s["outer_iters"] for s in fp64w["stats"]]))
    fp64t = run_mode(poisson, capi, w, capi.PREC_FP64, 2, 1, args.fp64_tol, args.aa, args.launch_mode,
                     args.mgrids, walls=0)
    log("fp64 (reference form, rel %.0e): %.3f s for 2 solves" % (args.fp64_tol, fp64t["total_s"]))
    err_mixed = mean_free_err(mixed["pres"], w.p_exact, mask)
    err_fp64 = mean_free_err(fp64["pres"], w.p_exact, mask)
    err_mix_vs_64 = mean_free_err(mixed["pres"], fp64t["pres"], mask)
    value = args.steps / mixed["total_s"]
    v64 = k64 / fp64["total_s"]
    v64w = k64 / fp64w["total_s"]
    v64t = 2 / fp64t["total_s"]
    lat = np.array(mixed["lat"]) * 1e3
    lat64 = np.array(fp64["lat"]) * 1e3
    # dominant kernel: finest-level line-SOR half sweep (FP32 in the mixed path)
    s_bytes = 4
    bytes_per_launch = (3 * s_bytes + 1) * n1 / 2.0
    avg_ms = mixed["sweep_ms"] / max(mixed["launches"], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s",
                frac=round(achieved / 8000.0, 4),
                traffic=sweep_traffic() if (args.config == "c2" and args.n == 256) else None,
                kernel="k_xsweep/k_yzsweep<float,8> finest level, one colour of one direction",
                bytes_per_launch=bytes_per_launch, avg_launch_ms=round(avg_ms, 5), launches=mixed["launches"])
    cpu = None
    if not args.no_cpu:
        n_iter64 = int(np.median([s["outer_iters"] for s in fp64["stats"]]))
        cpu = cpu_baseline(capi, w, n_iter64, mgrids=args.mgrids)
        log("cpu baseline: %s" % cpu)
    serving = None
    if payloads is not None:
        serving = {"config": {"requests": args.serve_requests, "arrivals": "Poisson, mean %.0f req/s, seed 4"
                              % args.serve_rate, "cells_per_axis": "uniform 32..192", "distinct_payloads":
                              len(payloads), "max_active": args.serve_active, "rel_tol": args.rel_tol}}
        for name, prec in (("mixed", capi.PREC_MIXED), ("fp64", capi.PREC_FP64)):
            serving[name] = serve_run(capi, poisson, payloads, prec, args.serve_requests, args.serve_rate, 4,
                                      args.serve_active, 96 << 30, args.rel_tol, args.serve_mgrids)
            log("serving %s: %s" % (name, serving[name]))
    hmin = [float(np.min(np.diff(a[1:-1]))) for a in w.grid]
    out = {
        "metric": "%s masked Poisson solves/s @1e-8 rel resid (mixed vs FP64); p99 latency" % shape,
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["total_s"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 smoother/transfers + f64 residual, GCM, update",
        "data": "synthetic (manufactured RHS, seeded)",
        "config": {"workload": desc, "hmin": [round(h, 7) for h in hmin], "rhs": w.info.get("rhs", "manufactured"),
                   "nGhost": w.info["nGhost"], "nDead": w.info["nDead"], "outer": "Anderson(%s)" % (args.aa or "default"),
                   "mg_levels": args.mgrids or "auto",
                   "stop_rule": "flag-cell rel resid <= %.0e and max outer step <= %.0e max|p|" % (args.rel_tol,
                                                                                                  args.rel_tol),
                   "correction_cycles": "bodies as zero-flux walls (vc_walls=1)",
                   "coarsest_level": "FP64 direct solve" if args.coarse_solve else "one line-SOR pass"},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "fp64": {"value": round(v64, 4), "p50_ms": round(float(np.percentile(lat64, 50)), 3),
                 "p99_ms": round(float(np.percentile(lat64, 99)), 3),
                 "form": "reference (V-cycles on p, masked cells held), same stop rule",
                 "outer_iters": [s["outer_iters"] for s in fp64["stats"]]},
        "fp64_wall_cycles": {"value": round(v64w, 4), "form": "the mixed solve's algorithm in FP64",
                             "outer_iters": [s["outer_iters"] for s in fp64w["stats"]],
                             "err_vs_exact": mean_free_err(fp64w["pres"], w.p_exact, mask)},
        "fp64_drop_in": {"value": round(v64t, 4), "rel_tol": args.fp64_tol,
                         "outer_iters": [s["outer_iters"] for s in fp64t["stats"]],
                         "err_vs_exact": mean_free_err(fp64t["pres"], w.p_exact, mask)},
        "mixed_over_fp64": round(value / v64, 3),
        "mixed_over_fp64_same_cycles": round(value / v64w, 3),
        "pcie_inclusive": {"value": round(1.0 / mixed["host_s"], 4),
                           "what": "fc_pson_solve with host arrays (upload div, download p)"},
        "outer_iters": [s["outer_iters"] for s in mixed["stats"]],
        "vcycles_fp32": [s["vcycles_fp32"] for s in mixed["stats"]],
        "rel_resid": max(s["rel_resid"] for s in mixed["stats"]),
        "err_vs_exact": {"mixed": err_mixed, "fp64": err_fp64, "mixed_vs_fp64": err_mix_vs_64,
                         "mixed_vs_fp64_what": "against the FP64 drop-in solved to rel %.0e" % args.fp64_tol},
        "hbm_bytes": mixed["bytes"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
