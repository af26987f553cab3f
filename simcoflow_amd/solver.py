"""Time stepping of the hot path on one device context.

``TimeStepper.step`` follows ``AdamBasforthCrankNicolson``
(Solver_Mod.f90:246-319) for solids at rest in a closed box:

  CopyNewCell (time-n copies of the P / U / V interface fields)
  ComputeTimeStep (:321-344)                       -> sg_time_step_size
  if itt > 1: Coupled_LS_VOF (Clsvof_Mod.f90:730)  -> sg_clsvof_advect_resident
              Initial_ClsVofUV (:402) for U and V  -> sg_interface_to_uv_resident
  Boundary_Condition_Var (closed-box form)         -> sg_closed_box_bc
  UpdatePUV (Solver_Mod.f90:312):
      Predictor_UV                                 -> sg_predictor_resident
      PoissonEquationSolver                        -> sg_poisson_build/solve_resident
      velocity correction                          -> sg_velocity_correct_resident

With ``moving`` the solid moves as a rigid body (ComputeForceObject ->
acceleration, ObjectMovement + SolidVolumeFraction inside the transport's
sub-steps, Grid_Preprocess + NewCellFace on the device, InterNewVar,
moving-wall terms); particles are tracked with ``Projection.particles_*``
beside the step.  Every field stays on the device between steps; the host
reads dt and the solver reports.  Left out: the wave channel's
inlet/outlet profiles (a closed box instead).  Non-dimensional groups use
Uref = 1 m/s and Lref = 1 m (StateVariables.f90:97-98): Re = 1/nu_l,
Fr = 1/sqrt(g).
"""
import math

import numpy as np

from . import _lib
from . import cases as C
from ._lib import F, check
from .projection import Projection, params_for


def physics(fluids):
    nu_l = fluids.mu_l / fluids.rho_l
    return dict(nu_l=nu_l, nu_g=fluids.mu_g / fluids.rho_g, nu_ref=nu_l, Re=1.0 / nu_l, Fr=1.0 / math.sqrt(fluids.g))


class TimeStepper:
    def __init__(self, case, precision=_lib.SG_PREC_FP64, tol=None, cfl=0.25, vint=0.0, uref=1.0, max_iter=500,
                 learned=None, stags=None, learned_width=256, previous_pressure=False, diagnostics=False,
                 device_geometry=False, moving=None):
        """learned: (weights, level, out_scale, deflate) of a trained warm
        start, level None for the first multigrid level no wider than
        learned_width; stags: the (U, V) cases.StagGeometry if already
        computed; previous_pressure: start each Poisson solve from the last
        step's pressure instead of zero (the reference starts from zero,
        ProjectionP_Mod.f90:319; it solves for the full pressure, so the
        last one is a close start), with learned: the last pressure plus the
        net's correction predicted from its residual; diagnostics: record a
        row of device diagnostics after every step (read with history());
        device_geometry: build the cut-cell geometry of case.shape on the
        device (sg_geometry_rebuild, equal to the host builder bit for bit)
        instead of uploading case.geo and the host's staggered geometry;
        moving: dict(mass=M, g=gravity, start=itt) to let case.shape move
        as a rigid body under gravity and the fluid force (ComputeForceObject
        -> asy = (F - M g)/M, held at rest before step `start`,
        Solver_Mod.f90:277-290), moved inside the transport's sub-steps
        (ObjectMovement + SolidVolumeFraction), its cut cells rebuilt after
        the transport (Grid_Preprocess) and the velocities of uncovered cells
        set by InterNewVar (Solver_Mod.f90:296-307); needs device_geometry."""
        if moving and (not device_geometry or case.shape[0] == "none"):
            raise ValueError("a moving solid needs device_geometry and a case.shape")
        self.moving = dict(moving, shape=(tuple(case.shape) + (0.0,) * 6)[:6], us=0.0, vs=0.0,
                           asy=0.0) if moving else None
        self.case = case
        g = case.grid
        self.cfl, self.vint, self.uref = cfl, vint, uref
        ph = physics(case.fluids)
        self.params = params_for(case.fluids, tol=case.tol if tol is None else tol, precision=precision,
                                 max_iter=max_iter, bc=case.bc, Re=ph["Re"], Fr=ph["Fr"],
                                 warm_start=(3 if previous_pressure else 2) if learned else
                                 (1 if previous_pressure else 0))
        # params_for takes nu from the fluids already
        self.pj = pj = Projection(g, self.params)
        if device_geometry:
            pj.geometry_rebuild(*case.shape)
        else:
            pj.set_geometry(case.geo)
            U, V = stags if stags is not None else C.staggered_geometry(g, case.geo, case.solid)
            pj.set_staggered_geometry(0, U)
            pj.set_staggered_geometry(1, V)
        itf = case.itf
        pj.upload("VOF", itf.vof)
        pj.upload("PHI", itf.phi)
        pj.upload("NX", itf.nx)
        pj.upload("NY", itf.ny)
        z = np.zeros((g.nj + 2, g.ni + 2))
        for name in ("U", "V", "P", "UPRED", "VPRED", "PP", "DPU", "DPV", "FLUXN1U", "FLUXN1V"):
            pj.upload(name, z)
        check(pj.lib.sg_interface_to_uv_resident(pj.ctx))
        if learned:
            from . import learned as L
            w, level, scale, deflate = learned
            L.enable(pj, w, L.pick_level(pj, learned_width) if level is None else level, scale, deflate)
        self.itt = 0
        self.time = 0.0
        self.last = {}
        self.diagnostics = diagnostics
        self.device_geometry = device_geometry

    def close(self):
        self.pj.close()

    def step(self):
        pj, lib, ctx = self.pj, self.pj.lib, self.pj.ctx
        self.itt += 1
        for dst, src in (("VOFO", "VOF"), ("PHIO", "PHI"), ("NXO", "NX"), ("NYO", "NY"), ("UVOFO", "UVOF"),
                         ("VVOFO", "VVOF")):
            check(lib.sg_field_copy(ctx, F[dst], F[src]))
        mv = self.moving
        if mv:
            pj.geometry_snapshot()
            force = pj.object_force(mv["vs"])[0]
            mv["asy"] = (force - mv["mass"] * mv["g"]) / mv["mass"]
            if self.itt < mv.get("start", 10):
                mv["us"] = mv["vs"] = mv["asy"] = 0.0
            mv["force"] = force
        vs, asy = (mv["vs"], mv["asy"]) if mv else (0.0, 0.0)
        dt = _lib.ctypes.c_double(0.0)
        check(lib.sg_time_step_size(ctx, self.cfl, self.vint, self.uref, vs, asy, _lib.ctypes.byref(dt)))
        dt = dt.value
        if self.itt > 1:
            if mv:
                kind, cx, cy, a, b, cc = mv["shape"]
                pj.set_moving_solid(kind, cx, cy, a, b, cc, us=mv["us"], vs=mv["vs"], asx=0.0, asy=mv["asy"])
            pj.clsvof_advect_resident(dt, self.time)
            if mv:
                # Initial_ClsVofUV updates the U / V solid fractions for the
                # moved solid before it computes and clips the U / V volume
                # fractions (SolidVolumeFraction, Clsvof_Mod.f90:416,681-684):
                # rebuild first, so the clip sees the new solid
                (_, cx, cy, a, b, cc), mv["us"], mv["vs"] = pj.get_moving_solid()
                pj.set_moving_solid("none")
                mv["shape"] = (kind, cx, cy, a, b, cc)
                pj.geometry_rebuild(kind, cx, cy, a, b, cc)
            check(lib.sg_interface_to_uv_resident(ctx))
        check(lib.sg_closed_box_bc(ctx))
        vb = mv["vs"] if mv else 0.0
        if mv and self.itt > 1:
            pj.inter_new_var(vb)
        pinfo = pj.predictor_resident(dt, vb=vb, first_step=(self.itt == 1))
        pj.build(vb)
        sinfo = pj.solve()
        pj.velocity_correct(vb)
        if self.diagnostics:
            pj.diagnostics_record()
        self.time += dt
        self.last = dict(dt=dt, predictor=pinfo.as_dict(), poisson=sinfo.as_dict())
        return self.last

    def history(self):
        """Diagnostics rows recorded since the last call, as a dict of
        columns (dt, liquid and gas volume, gas centroid, max |mres|, max
        |u|, max |v|)."""
        rows, lost = self.pj.diagnostics_read()
        if lost:
            raise RuntimeError(f"{lost} diagnostics rows were overwritten before they were read")
        return {k: rows[:, c] for c, k in enumerate(_lib.DIAG_NAMES)}

    def save(self, path):
        """Checkpoint of the run (device state, step counter, time; a moving
        solid's shape, velocity, acceleration and mass)."""
        from .output import save_checkpoint
        extra = {}
        if self.moving:
            extra["moving"] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in self.moving.items()}
        save_checkpoint(path, self.pj, self.itt, self.time, extra=extra)

    def load(self, path):
        """Resumes from a checkpoint written by save on the same case; a
        moving solid is put back where it was and its cut-cell geometry is
        rebuilt on the device there (the state holds the fields, not the
        geometry)."""
        from .output import load_checkpoint
        meta = load_checkpoint(path, self.pj)
        self.itt, self.time = meta["itt"], meta["time"]
        mv = meta.get("extra", {}).get("moving")
        if mv:
            if not self.moving:
                raise ValueError(f"{path}: saved with a moving solid; make the stepper with moving=...")
            mv["shape"] = tuple(mv["shape"])
            self.moving = mv
            self.pj.geometry_rebuild(*mv["shape"])

    def write_vtr(self, directory, fmt="ascii"):
        """Print_Result_VTK_2D's file for the current step
        (<directory>/Paraview/FlowField_<itt>.vtr); returns its path."""
        from .output import vtr_name, write_vtr
        path = vtr_name(directory, self.itt)
        st = self.state()
        if self.device_geometry:
            # the geometry lives on the device (it moves with a moving solid)
            phiS, vofS = self.pj.download("PHIS"), self.pj.download("VOFS")
        else:
            phiS, vofS = self.case.geo.phiS, self.case.geo.vofS
        write_vtr(path, self.case.grid, st["u"], st["v"], st["p"], st["phi"], phiS, st["vof"], vofS, fmt=fmt)
        return path

    def state(self):
        """Host copies of the interface and the velocity / pressure fields."""
        return {k: self.pj.download(n) for k, n in (("vof", "VOF"), ("phi", "PHI"), ("nx", "NX"), ("ny", "NY"),
                                                    ("u", "U"), ("v", "V"), ("p", "P"))}
