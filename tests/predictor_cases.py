"""Inputs of the momentum predictor for tests: a case's P-cell state at time
n, U/V cell geometry and liquid fractions, face velocities, Flux_n1 and the
previous Dp, all seeded."""
import numpy as np

from oracle.step import oracle_predictor, phys_of  # noqa: F401  (re-exported for the tests)
from simcoflow_amd import cases as C


def predictor_inputs(case, seed=0, amp=0.4, Re=None, shift=0.3):
    g = case.grid
    rng = np.random.default_rng(seed)
    U, V = C.staggered_geometry(g, case.geo, case.solid)
    # vof at time n and after a small shift of the interface (the transport)
    ls = case.levelset

    def shifted(X, Y):
        return ls(X - shift * g.dx[0], Y + 0.5 * shift * g.dy[0])
    uvo = C.staggered_vof(g, U, 0, ls)
    vvo = C.staggered_vof(g, V, 1, ls)
    uvn = C.staggered_vof(g, U, 0, shifted)
    vvn = C.staggered_vof(g, V, 1, shifted)
    u, v = C.transport_velocity(g, "vortex", amp=amp, seed=seed + 1)
    u += 0.05 * amp * rng.standard_normal(u.shape)
    v += 0.05 * amp * rng.standard_normal(v.shape)
    itf = case.itf
    pcell = dict(vof=itf.vof, phi=itf.phi, nx=itf.nx, ny=itf.ny, vofS=case.geo.vofS, nxS=case.geo.nxS,
                 nyS=case.geo.nyS, phiS=case.geo.phiS)
    ng = (g.nj + 2, g.ni + 2)
    fluxn1 = 1e-3 * rng.standard_normal((2,) + ng)
    DpU = case.dt * (1.0 + 0.1 * rng.uniform(size=ng))
    DpV = case.dt * (1.0 + 0.1 * rng.uniform(size=ng))
    upred = rng.standard_normal(ng)
    vpred = rng.standard_normal(ng)
    return dict(grid=g, U=U, V=V, uvo=uvo, uvn=uvn, vvo=vvo, vvn=vvn, pcell=pcell, u=np.ascontiguousarray(u),
                v=np.ascontiguousarray(v), phys=phys_of(case.fluids, Re=Re), dt=case.dt, fluxn1=fluxn1,
                DpU=DpU, DpV=DpV, upred=upred, vpred=vpred, geo=case.geo)


def two_phase_stream(flow):
    """A uniform stream (along x or y) in a 32x24 tank across a water-air
    interface inclined at 30 degrees, no gravity: the input of the
    mass-momentum consistency checks (the stream must stay uniform)."""
    g = C.uniform_grid(32, 24, 1.0, 0.75)
    geo = C.build_geometry(g, None)
    U, V = C.staggered_geometry(g, geo, None)
    th = np.deg2rad(30.0)
    nx, ny = np.sin(th), np.cos(th)
    X, Y = np.meshgrid(g.xc, g.yc)
    phi = nx * X + ny * Y - 0.45
    h = g.dx[0]
    vof = C.frac(nx, ny, phi, g.dx[None, :], g.dy[:, None])
    uvof = C.frac(nx, ny, phi + nx * 0.5 * h, g.udx[None, :], g.dy[:, None])
    vvof = C.frac(nx, ny, phi + ny * 0.5 * h, g.dx[None, :], g.vdy[:, None])
    pcell = dict(vof=vof, phi=phi, nx=np.full(g.shape, nx), ny=np.full(g.shape, ny), vofS=geo.vofS, nxS=geo.nxS,
                 nyS=geo.nyS, phiS=geo.phiS)
    ng = (g.nj + 2, g.ni + 2)
    uc, vc = (0.37, 0.0) if flow == "x" else (0.0, 0.29)
    return dict(grid=g, U=U, V=V, uvo=uvof, uvn=uvof, vvo=vvof, vvn=vvof, pcell=pcell, u=np.full(ng, uc),
                v=np.full(ng, vc), phys=phys_of(C.Fluids(), Fr=1e30), dt=0.004, fluxn1=np.zeros((2,) + ng),
                DpU=np.zeros(ng), DpV=np.zeros(ng), upred=np.zeros(ng), vpred=np.zeros(ng), geo=geo, uc=uc, vc=vc)


def device_predictor(d, first_step=False, vb=0.0, momentum_method=0):
    """The same step through the C ABI; returns the device results."""
    from simcoflow_amd.projection import Projection, default_params
    g = d["grid"]
    ph = d["phys"]
    prm = default_params(rho_l=ph["rho_l"], rho_g=ph["rho_g"], rho_ref=ph["rho_ref"], nu_l=ph["nu_l"],
                         nu_g=ph["nu_g"], nu_ref=ph["nu_ref"], Re=ph["Re"], Fr=ph["Fr"],
                         momentum_method=momentum_method)
    pj = Projection(g, prm)
    try:
        pj.set_geometry(d["geo"])
        pj.set_staggered_geometry(0, d["U"])
        pj.set_staggered_geometry(1, d["V"])
        pc = d["pcell"]
        for name, key in (("VOFO", "vof"), ("PHIO", "phi"), ("NXO", "nx"), ("NYO", "ny")):
            pj.upload(name, pc[key])
        pj.upload("UVOFO", d["uvo"])
        pj.upload("UVOF", d["uvn"])
        pj.upload("VVOFO", d["vvo"])
        pj.upload("VVOF", d["vvn"])
        pj.upload("U", d["u"])
        pj.upload("V", d["v"])
        pj.upload("FLUXN1U", d["fluxn1"][0])
        pj.upload("FLUXN1V", d["fluxn1"][1])
        pj.upload("DPU", d["DpU"])
        pj.upload("DPV", d["DpV"])
        pj.upload("UPRED", d["upred"])
        pj.upload("VPRED", d["vpred"])
        info = pj.predictor_resident(d["dt"], vb=vb, first_step=first_step)
        mu, ru, mv, rv = pj.predictor_rows()
        out = dict(info=info.as_dict(), matr_u=mu, rhs_u=ru, matr_v=mv, rhs_v=rv,
                   fluxn1=np.stack([pj.download("FLUXN1U"), pj.download("FLUXN1V")]),
                   DpU=pj.download("DPU"), DpV=pj.download("DPV"), upred=pj.download("UPRED"),
                   vpred=pj.download("VPRED"))
    finally:
        pj.close()
    return out
