"""One time step of flow past a moving triangulated body, in the order of
the reference's time_step_viscous for a prescribed motion without the
structural solver (CORE_TIME_STEP.F90:100-222, one FSI pass):

  markers moved (move_boundary's prescribed positions and velocities)
  set_iblank_body_fast, GCM_set_internal_boundary        geometry refresh
  GCM_SetBodyInterceptValues
  GCM_ghostcell_velocity(1, uOld, vOld, wOld), face_vel(uOld, vOld, wOld)
  set_solve_ad
  rhs_advec_diff, solve_ad, face_vel(u, v, w)
  rhs_poisson, solve_poisson
  correct_vel, correct_ghostcell_velocity(pPrime), face_vel, update_pressure
  compute_marker_stress

Every routine runs in the HIP library: the geometry on a GeometryRefresh
handle, the flow on a PoissonSolver handle.  The refresh's masks and lists
reach the flow handle through the host (fc_geom_get_* then
fc_pson_set_masks / fc_pson_set_gcm / fc_set_gcm_vel).  The flow state u, v,
w, p, pPrime and the face velocities is either held in host arrays between
the calls, which copy it in and out (the form of the reference's own call
sites, and the one the CPU oracle can stand in for), or, with
resident=True, kept in the flow handle's HBM buffers from step to step
(fc_ghostcell_velocity, fc_face_vel_resident, fc_rhs_advec_diff_resident,
fc_solve_ad_resident, fc_pson_rhs_resident, fc_pson_solve_resident,
fc_project_resident, fc_geom_marker_stress_resident); download() then
brings it to the host arrays.  One difference in
the resident form: the ghost-cell velocity update before the step acts on
the one velocity field (the reference updates the dead cells of its copy
uOld only); solve_ad recomputes every ghost and dead cell before it
converges, so the step's result is the same to the solver tolerances."""
import time

import numpy as np

from . import surface as surf


def gcm_lists(r, bi):
    """The lists fc_pson_set_gcm takes (GCM_arrays of the pressure update),
    from a refresh's handoff r and its body-intercept values bi."""
    g = {k: r[k] for k in ("iGhost", "jGhost", "kGhost", "iCellIndex", "jCellIndex", "kCellIndex", "coeffGCMN",
                           "alphaGhost", "iDead", "jDead", "kDead", "iDeadCellIndex", "jDeadCellIndex",
                           "kDeadCellIndex", "coeffGCMDeadN", "probeLengthDead")}
    g["dpdnBodyIntercept"] = bi["dpdnBodyIntercept"]
    g["dpdnInterceptDead"] = bi["dpdnInterceptDead"]
    g["xyzDCIP"] = r["xyzDCIP_dead"]
    return g


def pitching_plate_motion(plate, center, t, dt):
    """The prescribed pitching 0.3 sin(2 pi t) about z through center: the
    markers at time t and their velocities at t and t - dt (rigid rotation,
    omega = 0.6 pi cos(2 pi t))."""
    s = surf.pitch(plate, center, surf.pitch_angle(t))

    def vel(tt):
        om = 0.6 * np.pi * np.cos(2.0 * np.pi * tt)
        return -om * (s.y - center[1]), om * (s.x - center[0]), np.zeros(s.npts)

    return s, vel(t) + vel(t - dt)


class FractionalStep:
    """The flow state of one rank and the step sequence above.

    flow: a PoissonSolver (params with the outer velocity and pressure
    conditions), geom: a GeometryRefresh of the same grid.  Both are driven
    through their methods only; an object with the same methods (the CPU
    oracle's, in the tests) can stand in for either."""

    def __init__(self, flow, geom, shape, reinv, dt, probeD, alfa_upwind=1.0, crank_nicolson=1, omega_ad=1.0,
                 itermax_ad=200, restol_ad=1e-10, restol_gcv=1e-12, u0=(0.0, 0.0, 0.0), vel_bc=None,
                 resident=False):
        self.flow, self.geom = flow, geom
        self.shape = shape
        self.reinv, self.dt, self.probeD = reinv, dt, probeD
        self.alfa_upwind = alfa_upwind
        self.cn, self.omega_ad = crank_nicolson, omega_ad
        self.itermax_ad, self.restol_ad, self.restol_gcv = itermax_ad, restol_ad, restol_gcv
        z = lambda: np.zeros(shape, order="F")  # noqa: E731
        self.u, self.v, self.w = (np.asfortranarray(np.full(shape, c)) for c in u0)
        self.p, self.pPrime = z(), z()
        self.face_u, self.face_v, self.face_w = (np.asfortranarray(np.full(shape, c)) for c in u0)
        self.time = 0.0
        self.ntime = 0
        # the outer-face velocities bcxu, bcyv, bczw face_vel starts from
        # (set once the flow handle has masks; solve_ad updates them after)
        self._vel_bc = vel_bc
        self.resident = resident
        self._uploaded = False

    def step(self, bodies, marker_vel):
        """One time step: bodies = the surfaces at the new time, marker_vel =
        per body (u, v, w, uOld, vOld, wOld) of the markers.  Returns a dict
        of what the step reports (list sizes, solve statistics, the nine
        force totals per body)."""
        f, g = self.flow, self.geom
        self.ntime += 1
        self.time += self.dt
        out = {"ntime": self.ntime, "time": self.time}
        clock = [time.perf_counter()]
        phases = {}

        def lap(name):
            now = time.perf_counter()
            phases[name] = now - clock[0]
            clock[0] = now

        for b, s in enumerate(bodies):
            g.move_body(b + 1, s)
        st = g.refresh()
        lap("refresh")
        r = g.handoff()
        out["nGhost"], out["nDead"] = r["nGhost"], r["nDead"]
        for b, m in enumerate(marker_vel):
            g.set_marker_velocity(b + 1, *m)
        g.body_intercept_values(self.dt, self.ntime)
        bi = g.bi_values()
        lap("lists_to_host")
        f.set_masks(r["iblank"], r["ghostCellMark"], r["dead_cell"])
        if self._vel_bc is not None:
            f.set_vel_bc(*self._vel_bc)
            self._vel_bc = None
        f.set_gcm(gcm_lists(r, bi))
        f.set_gcm_vel(r["coeffGCMD"], r["coeffGCMDeadD"],
                      np.stack([bi["uBodyIntercept"], bi["vBodyIntercept"], bi["wBodyIntercept"]], 1),
                      np.stack([bi["uBodyInterceptDead"], bi["vBodyInterceptDead"], bi["wBodyInterceptDead"]], 1),
                      self.probeD)
        lap("lists_to_flow")
        if self.resident:
            self._flow_resident(out, lap)
        else:
            self._flow_host(out, lap)
        out["refresh"] = st
        out["phases_s"] = phases
        return out

    def _flow_host(self, out, lap):
        f, g = self.flow, self.geom
        # save_lastStep, then the dead cells of uOld and its face velocities
        uo, vo, wo = (a.copy(order="F") for a in (self.u, self.v, self.w))
        f.ghostcell_velocity_host(uo, vo, wo, 1, self.restol_gcv)
        f.face_vel(uo, vo, wo, self.face_u, self.face_v, self.face_w)
        f.set_solve_ad(self.reinv, self.dt, self.cn, self.omega_ad)
        nl = [np.zeros(self.shape, order="F") for _ in range(3)]
        f.rhs_advec_diff(uo, vo, wo, self.face_u, self.face_v, self.face_w, self.alfa_upwind, self.dt, *nl)
        rc, it, res, _ = f.solve_ad(self.u, self.v, self.w, *nl, self.face_u, self.face_v, self.face_w, self.time,
                                    self.itermax_ad, self.restol_ad, self.restol_gcv)
        out["ad_rc"], out["ad_iters"], out["ad_res"] = rc, it, res
        f.face_vel(self.u, self.v, self.w, self.face_u, self.face_v, self.face_w)
        lap("advection_diffusion")
        div, sumd, resdiv = f.rhs_poisson(self.face_u, self.face_v, self.face_w, 1.0 / self.dt)
        out["sumd"], out["resDiv"] = sumd, resdiv
        rc, pst, _ = f.solve_poisson(self.pPrime, div, self.p)
        out["poisson_rc"], out["poisson"] = rc, pst
        lap("poisson")
        f.correct_vel(self.pPrime, self.dt, self.u, self.v, self.w)
        f.correct_ghostcell_velocity(self.pPrime, self.dt, self.u, self.v, self.w)
        f.face_vel(self.u, self.v, self.w, self.face_u, self.face_v, self.face_w)
        f.update_pressure(self.pPrime, self.p)
        lap("projection")
        out["totals"] = np.array(g.marker_stress(self.p, self.u, self.v, self.w, self.reinv))
        lap("marker_stress")

    def _flow_resident(self, out, lap):
        f, g = self.flow, self.geom
        if not self._uploaded:
            f.vel_upload(self.u, self.v, self.w, self.face_u, self.face_v, self.face_w)
            f.upload(self.pPrime, None, self.p)
            self._uploaded = True
        f.ghostcell_velocity(1, self.restol_gcv)
        f.face_vel_resident()
        f.set_solve_ad(self.reinv, self.dt, self.cn, self.omega_ad)
        f.rhs_advec_diff_resident(self.alfa_upwind, self.dt)
        rc, it, res, _ = f.solve_ad_resident(self.time, self.itermax_ad, self.restol_ad, self.restol_gcv)
        out["ad_rc"], out["ad_iters"], out["ad_res"] = rc, it, res
        f.face_vel_resident()
        lap("advection_diffusion")
        out["sumd"], out["resDiv"] = f.rhs_resident(1.0 / self.dt)
        rc, pst = f.solve_resident()
        out["poisson_rc"], out["poisson"] = rc, pst
        lap("poisson")
        f.project_resident(self.dt, self.dt)
        lap("projection")
        out["totals"] = np.array(g.marker_stress_resident(f, self.reinv))
        lap("marker_stress")

    def download(self):
        """The resident state into the host arrays (resident=True)."""
        if self.resident and self._uploaded:
            f = self.flow
            self.u, self.v, self.w, self.face_u, self.face_v, self.face_w = f.vel_download()
            self.p = f.download_p()
            self.pPrime = f.download()
