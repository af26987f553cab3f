"""The oracle's time step (test infrastructure: the checker of the GPU
TimeStepper and the CPU-baseline leg of bench.py; nothing in the product path
imports it).

``OracleStepper.step`` runs, in the order of simcoflow_amd.solver.TimeStepper
(AdamBasforthCrankNicolson, Solver_Mod.f90:246-319, solids at rest, closed
box): copies of the time-n interface, ComputeTimeStep, Coupled_LS_VOF and
Initial_ClsVofUV (from the second step), the closed-box ghost layer,
Predictor_UV (direct sparse momentum solves), PoissonEquationSolver (the
oracle's PCG-multigrid to `tol`) and the velocity correction."""
from types import SimpleNamespace

import numpy as np

from simcoflow_amd import cases as C


def phys_of(fl, Re=None, Fr=None):
    """Non-dimensional groups with Uref = 1 m/s, Lref = 1 m
    (StateVariables.f90: Re = Uref Lref / nuref, Fr = Uref / sqrt(g Lref))."""
    nu_l = fl.mu_l / fl.rho_l
    return dict(rho_l=fl.rho_l, rho_g=fl.rho_g, rho_ref=fl.rho_ref, nu_l=nu_l, nu_g=fl.mu_g / fl.rho_g,
                nu_ref=nu_l, Re=1.0 / nu_l if Re is None else Re, Fr=1.0 / np.sqrt(fl.g) if Fr is None else Fr)


def oracle_predictor(O, d, first_step=False, vb=0.0):
    """The oracle's step: rows, direct solves, scatter and ghost copies."""
    g = d["grid"]
    r = O.predictor_assemble(g, d["U"], d["V"], d["uvo"], d["uvn"], d["vvo"], d["vvn"], d["pcell"], d["u"], d["v"],
                             d["phys"], d["dt"], vb=vb, first_step=first_step, fluxn1=d["fluxn1"], DpU=d["DpU"],
                             DpV=d["DpV"])
    au = d["U"].active.copy()
    au[:, -1] = 0
    av = d["V"].active.copy()
    av[-1, :] = 0
    xu = O.momentum_solve(r["matr_u"], r["rhs_u"], au)
    xv = O.momentum_solve(r["matr_v"], r["rhs_v"], av)
    up = d["upred"].copy()
    vp = d["vpred"].copy()
    up[1:-1, 1:-1][:, :-1] = np.where(au[:, :-1] == 1, xu[:, :-1], 0.0)
    vp[1:-1, 1:-1][:-1, :] = np.where(av[:-1, :] == 1, xv[:-1, :], 0.0)
    DU, DV, vp = O.predictor_finish(g, r["DpU"], r["DpV"], vp)
    r.update(upred=up, vpred=vp, DpU=DU, DpV=DV, xu=xu, xv=xv)
    return r


class OracleStepper:
    def __init__(self, O, case, cfl=0.25, vint=0.0, uref=1.0, tol=1e-12):
        self.O, self.case, self.g = O, case, case.grid
        self.cfl, self.vint, self.uref, self.tol = cfl, vint, uref, tol
        g, geo = case.grid, case.geo
        self.cg = O.ClsGeom(g, geo)
        self.U, self.V = C.staggered_geometry(g, geo, case.solid)
        itf = case.itf
        self.vof, self.phi = itf.vof.copy(), itf.phi.copy()
        self.nx, self.ny = itf.nx.copy(), itf.ny.copy()
        gz = np.zeros((g.nj + 2, g.ni + 2))
        self.u, self.v, self.p = gz.copy(), gz.copy(), gz.copy()
        self.upred, self.vpred, self.DpU, self.DpV = gz.copy(), gz.copy(), gz.copy(), gz.copy()
        self.fluxn1 = np.zeros((2, g.nj + 2, g.ni + 2))
        self.uvof = self._to_uv(0)
        self.vvof = self._to_uv(1)
        self.phys = phys_of(case.fluids)
        self.itt = 0
        self.time = 0.0

    # the stepper's state under the device's field names (ghosted u, v, p,
    # predictor outputs, Flux_n1; cell-sized interface fields)
    STATE = {"VOF": "vof", "PHI": "phi", "NX": "nx", "NY": "ny", "U": "u", "V": "v", "P": "p", "UPRED": "upred",
             "VPRED": "vpred", "DPU": "DpU", "DPV": "DpV", "UVOF": "uvof", "VVOF": "vvof"}

    def set_state(self, fields, itt, time):
        """Continue from another run's state (fields keyed as STATE, plus
        FLUXN1U / FLUXN1V): the next step() is that run's step itt + 1."""
        for key, attr in self.STATE.items():
            setattr(self, attr, np.array(fields[key], dtype=np.float64))
        self.fluxn1 = np.stack([np.array(fields["FLUXN1U"], dtype=np.float64),
                                np.array(fields["FLUXN1V"], dtype=np.float64)])
        self.itt, self.time = int(itt), float(time)

    def state(self):
        out = {key: getattr(self, attr) for key, attr in self.STATE.items()}
        out["FLUXN1U"], out["FLUXN1V"] = self.fluxn1[0], self.fluxn1[1]
        return out

    def _to_uv(self, which):
        st = self.U if which == 0 else self.V
        return self.O.interface_to_uv(self.cg, which, self.vof, self.phi, self.nx, self.ny, st.vofS, st.phiS)[0]

    def step(self):
        O, g, case = self.O, self.g, self.case
        geo = case.geo
        self.itt += 1
        vofo, phio, nxo, nyo = self.vof.copy(), self.phi.copy(), self.nx.copy(), self.ny.copy()
        uvofo, vvofo = self.uvof.copy(), self.vvof.copy()
        dt = O.time_step(g, self.u, self.v, self.cfl, self.phys["Fr"], self.vint, self.uref)
        if self.itt > 1:
            self.vof, self.phi, self.nx, self.ny = O.clsvof_advect(self.cg, self.vof, self.phi, self.u, self.v, dt,
                                                                   time=self.time)
            self.uvof = self._to_uv(0)
            self.vvof = self._to_uv(1)
        O.closed_box_bc(self.u, self.v, self.p)
        d = dict(grid=g, U=self.U, V=self.V, uvo=uvofo, uvn=self.uvof, vvo=vvofo, vvn=self.vvof,
                 pcell=dict(vof=vofo, phi=phio, nx=nxo, ny=nyo, vofS=geo.vofS, nxS=geo.nxS, nyS=geo.nyS,
                            phiS=geo.phiS),
                 u=self.u, v=self.v, phys=self.phys, dt=dt, fluxn1=self.fluxn1, DpU=self.DpU, DpV=self.DpV,
                 upred=self.upred, vpred=self.vpred)
        r = oracle_predictor(O, d, first_step=(self.itt == 1))
        self.fluxn1, self.DpU, self.DpV = r["fluxn1"], r["DpU"], r["DpV"]
        self.upred, self.vpred = r["upred"], r["vpred"]
        f = SimpleNamespace(vof=self.vof, phi=self.phi, vofS=geo.vofS)
        M = O.poisson_matrix(g, geo, O.poisson_coef(g, f, self.DpU, self.DpV, case.fluids), case.bc)
        b = O.poisson_rhs(g, geo, self.upred, self.vpred)
        x, it, rr = O.pcg_mg_solve(g, M, b, tol=self.tol, max_iter=3000)
        Pp = O.pressure_scatter_bc(g, geo, x, case.bc)
        self.u, self.v, self.p, mres = O.velocity_correct(g, geo, f, case.fluids, self.DpU, self.DpV, self.upred,
                                                          self.vpred, Pp, self.u, self.v, self.p)
        self.time += dt
        return dict(dt=dt, iterations=it, mres=float(np.abs(mres).max()))


def gas_centroid(g, vof, vofS):
    """Centroid of the gas (1 - vof - vofS weighted) in cell units."""
    X, Y = np.meshgrid(g.xc, g.yc)
    w = np.clip(1.0 - vof - vofS, 0.0, None) * (g.dx[None, :] * g.dy[:, None])
    s = w.sum()
    return float((w * X).sum() / s / g.dx[0]), float((w * Y).sum() / s / g.dy[0])
