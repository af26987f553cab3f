"""Parity of the HIP interface transport (Coupled_LS_VOF) with the CPU oracle,
through the C ABI.

Both sides evaluate the reference's expressions in its order without FP
contraction, so vof, phi, nx and ny are compared bit for bit after whole
calls of nsub = 6 sub-steps (each sub-step: two sweeps, three
reconstructions, re-distancing).  Cases cover two-phase bubbles, cut solids
(three-phase cells, Ridders plane search), the west inflow wave, the edge
chains of the reference's reconstruction, random states whose solid line
disagrees with vofS (plane constants left unassigned), minimal and
non-uniform grids, and the full 2048 x 2048 bubble swarm.
"""
import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def _ctx(case_grid, geo):
    from simcoflow_amd.projection import Projection, default_params
    pj = Projection(case_grid, default_params())
    pj.set_geometry(geo)
    return pj


def _equal(what, a, b):
    if np.array_equal(a, b, equal_nan=True):
        return
    d = np.abs(a - b)
    d = np.where(np.isnan(d), np.inf, d)
    k = np.unravel_index(int(np.argmax(d)), a.shape)
    raise AssertionError(f"{what}: {int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b)))))} cells differ; "
                         f"max |diff| {d[k]:.3e} at (j,i)={k}: gpu {a[k]!r} oracle {b[k]!r}")


def _run_both(O, grid, geo, vof, phi, u, v, dt, steps=2, nsub=6, wave=None, t0=0.0):
    cg = O.ClsGeom(grid, geo)
    pj = _ctx(grid, geo)
    try:
        for s in range(steps):
            t = t0 + s * dt
            if wave is None:
                ref = O.clsvof_advect(cg, vof, phi, u, v, dt, time=t, nsub=nsub)
                got = pj.clsvof_advect(vof, phi, u, v, dt, time=t, nsub=nsub)
            else:
                ref = O.clsvof_advect(cg, vof, phi, u, v, dt, time=t, nsub=nsub, wave=wave)
                got = pj.clsvof_advect(vof, phi, u, v, dt, time=t, nsub=nsub, wave=wave)
            for name, a, b in zip(("vof", "phi", "nx", "ny"), got, ref):
                _equal(f"call {s} {name}", a, b)
            vof, phi = ref[0], ref[1]
    finally:
        pj.close()
    return vof, phi


CASES = [
    ("bubble-rotation", lambda: C.case_bubble(64, 128), "rotation"),
    ("bubble-vortex", lambda: C.case_bubble(64, 128), "vortex"),
    ("dambreak-vortex", lambda: C.case_dambreak(128, 64), "vortex"),
    ("dambreak-rotation", lambda: C.case_dambreak(128, 64), "rotation"),
    ("waves-vortex", lambda: C.case_waves(96, 48, seed=2), "vortex"),
    ("swarm-vortex", lambda: C.case_swarm(128, 128, seed=3), "vortex"),
    ("dambreak-odd-vortex", lambda: C.case_dambreak(131, 67), "vortex"),
]


@pytest.mark.parametrize("name,mk,kind", CASES, ids=[c[0] for c in CASES])
def test_advect_bit_exact(oracle_lib, name, mk, kind):
    case = mk()
    u, v = C.transport_velocity(case.grid, kind, amp=1.0, seed=1)
    print(f"[clsvof] {name} {case.grid.ni}x{case.grid.nj}", flush=True)
    _run_both(oracle_lib, case.grid, case.geo, case.itf.vof, case.itf.phi, u, v, case.dt, steps=3)


def test_inflow_wave_bit_exact(oracle_lib):
    """Open west/south faces with the wave condition of X_Sweep :1057-1068."""
    case = C.case_waves(96, 48, seed=2)
    u, v = C.transport_velocity(case.grid, "uniform", amp=0.4, walls=False)
    wave = (2.0, 0.3, 2 * np.pi / 2.0, 2 * np.pi / 1.3)
    print("[clsvof] inflow wave 96x48", flush=True)
    _run_both(oracle_lib, case.grid, case.geo, case.itf.vof, case.itf.phi, u, v, case.dt, steps=3,
              wave=wave, t0=0.37)


def _edge_film_case(ni=48, nj=64):
    """Liquid films along the west and east walls and a layer at the bottom:
    whole edge columns of interface cells (long stale-normal chains)."""
    g = C.uniform_grid(ni, nj, 1.0, nj / ni)
    geo = C.build_geometry(g, None)
    vof = np.zeros((nj, ni))
    vof[:, 0] = 0.45
    vof[:, -1] = 0.6
    vof[:4, :] = 1.0
    vof[4, 1:-1] = 0.3
    X, Y = np.meshgrid(g.xc, g.yc)
    h = g.dx[0]
    phi = np.minimum(np.minimum(X - 0.45 * h, (1.0 - 0.6 * h) - X), Y - 4.3 * h)
    phi = -phi  # gas (> 0) in the middle
    phi = np.where(vof >= 1.0, -np.abs(phi), phi)
    return g, geo, np.ascontiguousarray(vof), np.ascontiguousarray(phi)


def test_edge_chains_bit_exact(oracle_lib):
    g, geo, vof, phi = _edge_film_case()
    u, v = C.transport_velocity(g, "vortex", amp=0.5, seed=5)
    print("[clsvof] edge films 48x64", flush=True)
    _run_both(oracle_lib, g, geo, vof, phi, u, v, 0.25 * g.dx[0], steps=2)


def _random_state(ni, nj, seed, nonuniform=False):
    """Random cut-cell state: fractional, full and empty cells, solids whose
    line is drawn independently of vofS (so the Ridders search of
    DistanceFluidCalculate often fails to bracket and leaves the plane
    constant to the previous cell), random apertures and face flags."""
    rng = np.random.default_rng(seed)
    if nonuniform:
        dx = rng.uniform(0.5, 1.5, ni) / ni
        dy = rng.uniform(0.5, 1.5, nj) / nj
        xn = np.concatenate([[0.0], np.cumsum(dx)])
        yn = np.concatenate([[0.0], np.cumsum(dy)])
        xc, yc = 0.5 * (xn[1:] + xn[:-1]), 0.5 * (yn[1:] + yn[:-1])
        udx = np.append(xc[1:] - xc[:-1], dx[-1])
        vdy = np.append(yc[1:] - yc[:-1], dy[-1])
        g = C.Grid(ni, nj, 0.0, 0.0, dx, dy, udx, vdy, xc, yc)
    else:
        g = C.uniform_grid(ni, nj, 1.0, nj / ni)
    shp = (nj, ni)
    r = rng.uniform(size=shp)
    vofS = np.where(r < 0.7, 0.0, np.where(r < 0.9, rng.uniform(1e-3, 1 - 1e-3, shp), 1.0))
    th = rng.uniform(0, 2 * np.pi, shp)
    nxS, nyS = np.cos(th), np.sin(th)
    phiS = rng.uniform(-0.6, 0.6, shp) * g.dx[0]
    r = rng.uniform(size=shp)
    vof = np.where(r < 0.3, 0.0, np.where(r < 0.6, 1.0 - vofS, rng.uniform(0, 1, shp) * (1.0 - vofS)))
    phi = rng.uniform(-3, 3, shp) * g.dx[0]

    def ap():
        a = np.zeros((nj + 2, ni + 2))
        x = rng.uniform(size=shp)
        a[1:-1, 1:-1] = np.where(x < 0.6, 1.0, np.where(x < 0.8, rng.uniform(size=shp), 0.0))
        return a
    z = np.zeros(shp)
    geo = C.Geometry(vofS=np.ascontiguousarray(vofS), phiS=np.ascontiguousarray(phiS),
                     nxS=np.ascontiguousarray(nxS), nyS=np.ascontiguousarray(nyS), WlLh=z.copy(),
                     aE=ap(), aW=ap(), aN=ap(), aS=ap(), vofSU=z.copy(), vofSV=z.copy(),
                     moexU=np.zeros(shp, np.int32), moexV=np.zeros(shp, np.int32),
                     uflag=(rng.uniform(size=shp) < 0.9).astype(np.int32),
                     vflag=(rng.uniform(size=shp) < 0.9).astype(np.int32))
    return g, geo, np.ascontiguousarray(vof), np.ascontiguousarray(phi)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_states_bit_exact(oracle_lib, seed):
    g, geo, vof, phi = _random_state(40, 36, seed, nonuniform=(seed == 2))
    u, v = C.transport_velocity(g, "random", amp=0.5, seed=seed, walls=(seed != 1))
    print(f"[clsvof] random state seed {seed}", flush=True)
    _run_both(oracle_lib, g, geo, vof, phi, u, v, 0.2 * float(g.dx.min()), steps=2, nsub=6)


@pytest.mark.parametrize("fill", [0.0, 1.0])
def test_uniform_fill(oracle_lib, fill):
    """All gas / all liquid: nothing to reconstruct, re-distancing gives the
    far-field band value everywhere."""
    g = C.uniform_grid(32, 16, 2.0, 1.0)
    geo = C.build_geometry(g, None)
    vof = np.full(g.shape, fill)
    phi = np.full(g.shape, 0.3 if fill == 0.0 else -0.3)
    u, v = C.transport_velocity(g, "vortex", amp=1.0, seed=2)
    vof2, phi2 = _run_both(oracle_lib, g, geo, vof, phi, u, v, 0.01, steps=1)
    # the explicit second sweep is not clipped (Clsvof_Mod.f90:844-850): a
    # full domain keeps its volume to round-off
    assert np.abs(vof2 - fill).max() <= 1e-12
    band = 5.0 * np.sqrt(g.dx[0] ** 2 + g.dy[0] ** 2)
    assert np.allclose(phi2, band if fill == 0.0 else -band, rtol=0, atol=0)


def test_minimal_grid(oracle_lib):
    g, geo, vof, phi = _random_state(4, 4, 11)
    u, v = C.transport_velocity(g, "random", amp=0.3, seed=3)
    print("[clsvof] 4x4", flush=True)
    _run_both(oracle_lib, g, geo, vof, phi, u, v, 0.05, steps=2)


def test_rejects_small_grid():
    from simcoflow_amd._lib import SimcoError
    g = C.uniform_grid(3, 8, 1.0, 1.0)
    geo = C.build_geometry(g, None)
    pj = _ctx(g, geo)
    try:
        with pytest.raises(SimcoError):
            pj.clsvof_advect(np.zeros(g.shape), np.ones(g.shape), np.zeros((10, 5)), np.zeros((10, 5)), 0.1)
    finally:
        pj.close()


def test_full_size_swarm_bit_exact(oracle_lib):
    """The 2048 x 2048 bubble swarm, one call (6 sub-steps), bit for bit."""
    case = C.case_swarm(2048, 2048, seed=0)
    u, v = C.transport_velocity(case.grid, "vortex", amp=0.5, seed=4)
    print("[clsvof] swarm 2048x2048", flush=True)
    vof, phi = _run_both(oracle_lib, case.grid, case.geo, case.itf.vof, case.itf.phi, u, v, case.dt, steps=1)
    # the discretely divergence-free field keeps the liquid volume to
    # round-off plus the split scheme's O(dt^2) imbalance
    dv = abs(vof.sum() - case.itf.vof.sum()) / case.itf.vof.sum()
    assert dv < 1e-5, dv


# ---- interface on the U / V cells (Initial_ClsVofUV) -------------------------
def _uv_both(O, g, geo, stags, vof, phi, nx, ny):
    cg = O.ClsGeom(g, geo)
    pj = _ctx(g, geo)
    try:
        for w in (0, 1):
            pj.set_staggered_geometry(w, stags[w])
        for name, a in (("VOF", vof), ("PHI", phi), ("NX", nx), ("NY", ny)):
            pj.upload(name, a)
        got = pj.interface_to_uv()
    finally:
        pj.close()
    for w in (0, 1):
        ref = O.interface_to_uv(cg, w, vof, phi, nx, ny, stags[w].vofS, stags[w].phiS)
        for name, a, b in zip(("vof", "phi", "nx", "ny"), got[w], ref):
            _equal(f"{'UV'[w]} {name}", a, b)


UV_CASES = [
    ("bubble", lambda: C.case_bubble(64, 128)),
    ("dambreak", lambda: C.case_dambreak(128, 64)),
    ("waves", lambda: C.case_waves(96, 48, seed=2)),
    ("dambreak-odd", lambda: C.case_dambreak(131, 67)),
]


@pytest.mark.parametrize("name,mk", UV_CASES, ids=[c[0] for c in UV_CASES])
def test_interface_to_uv_bit_exact(oracle_lib, name, mk):
    O = oracle_lib
    case = mk()
    g = case.grid
    print(f"[to_uv] {name} {g.ni}x{g.nj}", flush=True)
    nx, ny, _ = O.reconstruct(O.ClsGeom(g, case.geo), case.itf.vof, case.itf.phi)
    stags = C.staggered_geometry(g, case.geo, case.solid)
    _uv_both(O, g, case.geo, stags, case.itf.vof, case.itf.phi, nx, ny)


@pytest.mark.parametrize("seed,ni,nj", [(0, 40, 36), (1, 37, 41), (2, 40, 36), (3, 4, 4)])
def test_interface_to_uv_random_bit_exact(oracle_lib, seed, ni, nj):
    """Random cut-cell states, including full-liquid cells that are cut by
    solids (no branch assigns the half fraction: the value of the previous
    staggered cell in loop order is kept) and a non-uniform grid."""
    import dataclasses
    g, geo, vof, phi = _random_state(ni, nj, seed, nonuniform=(seed == 2))
    rng = np.random.default_rng(100 + seed)
    th = rng.uniform(0, 2 * np.pi, g.shape)
    nx, ny = np.cos(th), np.sin(th)
    # full liquid where the solid fraction is small but above the threshold
    vof = np.where(rng.uniform(size=g.shape) < 0.1, 1.0, vof)
    stags = []
    for st in C.staggered_geometry(g, geo, None):
        r = rng.uniform(size=g.shape)
        vs = np.where(r < 0.7, 0.0, np.where(r < 0.9, rng.uniform(1e-3, 1 - 1e-3, g.shape), 1.0))
        stags.append(dataclasses.replace(st, vofS=np.ascontiguousarray(vs),
                                         phiS=np.ascontiguousarray(rng.uniform(-0.6, 0.6, g.shape) * g.dx[0])))
    print(f"[to_uv] random seed {seed} {ni}x{nj}", flush=True)
    _uv_both(oracle_lib, g, geo, stags, np.ascontiguousarray(vof), phi, nx, ny)


def test_interface_to_uv_full_size(oracle_lib):
    O = oracle_lib
    case = C.case_swarm(2048, 2048, seed=0)
    g = case.grid
    print("[to_uv] swarm 2048x2048", flush=True)
    nx, ny, _ = O.reconstruct(O.ClsGeom(g, case.geo), case.itf.vof, case.itf.phi)
    stags = C.staggered_geometry(g, case.geo, case.solid)
    _uv_both(O, g, case.geo, stags, case.itf.vof, case.itf.phi, nx, ny)
