"""The time-step orchestration on the GPU against the oracle: the CFL step
(ComputeTimeStep) and the closed-box ghost layer bit for bit, then 200 full
steps of a rising bubble (predictor, projection, correction, interface
transport) against the oracle's own 200 steps: the interface position (gas
centroid) within 0.01 cell, the liquid volume, and the mixed-precision
solver against the FP64 one on the same steps."""
import ctypes

import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def _ctx(g):
    from simcoflow_amd.projection import Projection, default_params
    return Projection(g, default_params(Fr=0.3))


@pytest.mark.parametrize("seed,vint,vs,asy", [(0, 0.0, 0.0, 0.0), (1, 2.5, 0.0, 0.0), (2, 0.0, 0.3, -1.2)])
def test_time_step_bit_exact(oracle_lib, seed, vint, vs, asy):
    rng = np.random.default_rng(seed)
    g = C.uniform_grid(96, 40, 3.0, 1.0)
    u = rng.standard_normal((g.nj + 2, g.ni + 2)) * 3.0
    v = rng.standard_normal((g.nj + 2, g.ni + 2)) * 3.0
    u[rng.uniform(size=u.shape) < 0.1] = 0.0
    v[rng.uniform(size=v.shape) < 0.1] = 0.0
    pj = _ctx(g)
    try:
        pj.upload("U", u)
        pj.upload("V", v)
        dt = ctypes.c_double()
        from simcoflow_amd._lib import check
        check(pj.lib.sg_time_step_size(pj.ctx, 0.25, vint, 1.3, vs, asy, ctypes.byref(dt)))
    finally:
        pj.close()
    ref = oracle_lib.time_step(g, u, v, 0.25, 0.3, vint, 1.3, vs, asy)
    assert dt.value == ref, (dt.value, ref)


def test_closed_box_bc_bit_exact(oracle_lib):
    rng = np.random.default_rng(3)
    g = C.uniform_grid(37, 21, 1.0, 1.0)
    u, v, p = (rng.standard_normal((g.nj + 2, g.ni + 2)) for _ in range(3))
    pj = _ctx(g)
    try:
        for n, a in (("U", u), ("V", v), ("P", p)):
            pj.upload(n, a)
        from simcoflow_amd._lib import check
        check(pj.lib.sg_closed_box_bc(pj.ctx))
        got = [pj.download(n) for n in ("U", "V", "P")]
    finally:
        pj.close()
    ref = oracle_lib.closed_box_bc(u.copy(), v.copy(), p.copy())
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def _run_gpu(case, steps, precision, tol, previous=False, iters=None):
    from simcoflow_amd.solver import TimeStepper
    st = TimeStepper(case, precision=precision, tol=tol, previous_pressure=previous)
    try:
        for k in range(steps):
            rep = st.step()
            assert rep["poisson"]["status"] == 0, (k, rep)
            if iters is not None:
                iters.append(rep["poisson"]["iterations"])
        return st.state(), st.time
    finally:
        st.close()


def test_200_steps_match_oracle(oracle_lib):
    from oracle import step as S
    case = C.case_bubble(64, 128)
    g = case.grid
    print("[step] bubble 64x128, 200 steps, GPU FP64", flush=True)
    gs, gt = _run_gpu(case, 200, 0, 1e-12)
    print("[step] oracle, 200 steps", flush=True)
    st = S.OracleStepper(oracle_lib, case, tol=1e-12)
    for _ in range(200):
        st.step()
    cg = S.gas_centroid(g, gs["vof"], case.geo.vofS)
    co = S.gas_centroid(g, st.vof, case.geo.vofS)
    c0 = S.gas_centroid(g, case.itf.vof, case.geo.vofS)
    dvol = abs(gs["vof"].sum() - st.vof.sum()) / st.vof.sum()
    dp = np.linalg.norm(gs["p"][1:-1, 1:-1] - st.p[1:-1, 1:-1]) / np.linalg.norm(st.p[1:-1, 1:-1])
    print(f"time {gt:.4f} vs {st.time:.4f}; centroid gpu {cg} oracle {co} (start {c0}); "
          f"volume rel diff {dvol:.2e}; max |dvof| {np.abs(gs['vof'] - st.vof).max():.2e}; pressure rel L2 {dp:.2e}")
    assert abs(gt - st.time) <= 1e-9 * st.time
    assert co[1] - c0[1] > 1.0          # the bubble rose more than a cell
    assert max(abs(cg[0] - co[0]), abs(cg[1] - co[1])) < 0.01
    assert dvol < 1e-12
    assert dp < 1e-6


def test_mixed_precision_interface_error():
    """200 steps with the mixed-precision solver (FP32 multigrid, FP64
    outer) against the FP64 solver, both to a relative residual of 1e-10:
    interface position within 0.01 cell, pressures and volumes reported."""
    from oracle import step as S
    case = C.case_bubble(64, 128)
    g = case.grid
    print("[step] bubble 64x128, 200 steps, FP64 and mixed", flush=True)
    a, _ = _run_gpu(case, 200, 0, 1e-10)
    b, _ = _run_gpu(case, 200, 1, 1e-10)
    ca = S.gas_centroid(g, a["vof"], case.geo.vofS)
    cb = S.gas_centroid(g, b["vof"], case.geo.vofS)
    dvol = abs(a["vof"].sum() - b["vof"].sum()) / a["vof"].sum()
    dp = np.linalg.norm(a["p"][1:-1, 1:-1] - b["p"][1:-1, 1:-1]) / np.linalg.norm(a["p"][1:-1, 1:-1])
    print(f"centroid fp64 {ca} mixed {cb}; volume rel diff {dvol:.2e}; pressure rel L2 {dp:.2e}")
    assert max(abs(ca[0] - cb[0]), abs(ca[1] - cb[1])) < 0.01
    assert dvol < 1e-12


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_previous_pressure_start(precision):
    """Starting each solve from the last step's pressure (not the
    reference's zero start) reaches the same states: 100 steps of the
    rising bubble to a relative residual of 1e-10, interface within 0.01
    cell, liquid volume 1e-12, pressure 1e-6 relative L2; the solves need
    fewer iterations on the whole."""
    from oracle import step as S
    case = C.case_bubble(64, 128)
    g = case.grid
    print(f"[step] bubble 64x128, 100 steps, precision {precision}, zero and previous-pressure starts", flush=True)
    ia, ib = [], []
    a, ta = _run_gpu(case, 100, precision, 1e-10, iters=ia)
    b, tb = _run_gpu(case, 100, precision, 1e-10, previous=True, iters=ib)
    ca = S.gas_centroid(g, a["vof"], case.geo.vofS)
    cb = S.gas_centroid(g, b["vof"], case.geo.vofS)
    dvol = abs(a["vof"].sum() - b["vof"].sum()) / a["vof"].sum()
    dp = np.linalg.norm(a["p"][1:-1, 1:-1] - b["p"][1:-1, 1:-1]) / np.linalg.norm(a["p"][1:-1, 1:-1])
    print(f"iterations zero start {sum(ia)}, previous pressure {sum(ib)}; centroid {ca} vs {cb}; "
          f"volume rel diff {dvol:.2e}; pressure rel L2 {dp:.2e}")
    assert abs(ta - tb) <= 1e-6 * ta
    assert max(abs(ca[0] - cb[0]), abs(ca[1] - cb[1])) < 0.01
    assert dvol < 1e-12
    assert dp < 1e-6
    assert sum(ib) < sum(ia)


STATE_FIELDS = ("VOF", "PHI", "NX", "NY", "U", "V", "P", "UPRED", "VPRED", "DPU", "DPV", "FLUXN1U", "FLUXN1V",
                "UVOF", "VVOF")
# transport, U/V interface, Dp and Flux_n1: bit for bit from the same state
EXACT_FIELDS = ("VOF", "PHI", "NX", "NY", "UVOF", "VVOF", "DPU", "DPV", "FLUXN1U", "FLUXN1V")


def _c1_diagnostics(g, vof, vofS, gas_centroid):
    area = g.dx[None, :] * g.dy[:, None]
    cx, cy = gas_centroid(g, vof, vofS)
    return float((vof * area).sum()), cx, cy, float(vof[0, :].sum())


def test_dambreak_500_steps_match_oracle(oracle_lib):
    """The 512 x 256 dam break with its rectangular boom, 500 FP64 steps to
    a relative residual of 1e-10 (AdamBasforthCrankNicolson,
    Solver_Mod.f90:196-242,246-319), against the oracle's own 500 steps
    (fixtures of tests/golden/make_dambreak_c1.py).

    * Step by step: at six points of the run the oracle takes one step from
      the device's state; the transport, the U/V interface, Dp and Flux_n1
      come out bit for bit, the predictor within 1e-12 and the pressure and
      velocities within 1e-9 relative (the two PCG runs round differently).
    * Trajectory: the first 200 steps agree to dt 1e-10, liquid volume
      1e-12 relative, gas centroid and bottom-row liquid length 1e-9 cell.
      The step holds discrete decisions (the sweeps' clipping, the
      density-ratio branch of the convective term) that a 1e-12 difference
      can flip: the oracle's own run splits from itself at step 213 when
      only its solve tolerance changes (1e-10 -> 1e-12, the second fixture),
      the device's splits from the oracle's at step 276.  After a split the
      runs stay within dt 1e-3 relative, volume 1e-5 relative, centroid
      1e-3 cell and bottom-row length 0.1 cell of each other."""
    import os
    from oracle import step as S
    O = oracle_lib
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ref = np.load(os.path.join(here, "dambreak_512x256_500.npz"))
    ref12 = np.load(os.path.join(here, "dambreak_512x256_500_tol1e-12.npz"))
    case = C.case_dambreak(512, 256)
    g = case.grid
    from simcoflow_amd.solver import TimeStepper
    st = TimeStepper(case, precision=0, tol=1e-10)
    ost = S.OracleStepper(O, case, tol=1e-10)
    checks = (0, 99, 199, 274, 399, 498)   # one oracle step from the state before these steps
    rows = []
    try:
        for k in range(500):
            if k in checks:
                before = {n: st.pj.download(n) for n in STATE_FIELDS}
                itt, t = st.itt, st.time
            rep = st.step()
            assert rep["poisson"]["status"] == 0, (k, rep)
            s = st.state()
            rows.append((rep["dt"],) + _c1_diagnostics(g, s["vof"], case.geo.vofS, S.gas_centroid))
            if k in checks:
                ost.set_state(before, itt, t)
                r = ost.step()
                after = {n: st.pj.download(n) for n in STATE_FIELDS}
                mine = ost.state()
                worst = {}
                for n in STATE_FIELDS:
                    d = np.abs(after[n] - mine[n]).max()
                    worst[n] = d / max(np.abs(mine[n]).max(), 1e-300)
                print(f"[step] dam break step {k + 1}: one oracle step from the device state: dt "
                      f"{abs(r['dt'] - rep['dt']) / r['dt']:.1e}, max rel diffs "
                      + ", ".join(f"{n} {v:.1e}" for n, v in worst.items() if v > 0), flush=True)
                assert r["dt"] == rep["dt"], k
                for n in EXACT_FIELDS:
                    assert np.array_equal(after[n], mine[n]), (k, n)
                assert worst["UPRED"] < 1e-12 and worst["VPRED"] < 1e-12, (k, worst)
                assert worst["U"] < 1e-9 and worst["V"] < 1e-9 and worst["P"] < 1e-9, (k, worst)
            if k % 50 == 49:
                print(f"[step] dam break 512x256, step {k + 1}/500", flush=True)
    finally:
        st.close()
    rows = np.array(rows)
    keys = ("dt", "volume", "centroid_x", "centroid_y", "bottom_length")
    gpu = {key: rows[:, c] for c, key in enumerate(keys)}

    def dev(a, b, key):
        d = np.abs(a[key] - b[key])
        return d / np.abs(b[key]) if key in ("dt", "volume") else d
    dg = {key: dev(gpu, ref, key) for key in keys}
    do = {key: dev(ref12, ref, key) for key in keys}
    def split(d):
        over = d["dt"] > 1e-9
        return int(np.argmax(over)) + 1 if over.any() else None
    print(f"[step] dam break: device vs oracle (tol 1e-10), worst over steps 1-200 / 1-500; oracle tol 1e-12 vs "
          f"1e-10 over 1-500: " + "; ".join(f"{key} {dg[key][:200].max():.1e} / {dg[key].max():.1e} / "
                                           f"{do[key].max():.1e}" for key in keys)
          + f"; first step with dt apart by 1e-9: device {split(dg)}, oracle {split(do)}", flush=True)
    assert ref["bottom_length"][-1] - ref["bottom_length"][0] > 20.0   # the water ran out along the floor
    assert dg["dt"][:200].max() < 1e-10 and dg["volume"][:200].max() < 1e-12
    for key in ("centroid_x", "centroid_y", "bottom_length"):
        assert dg[key][:200].max() < 1e-9, key
    assert dg["dt"].max() < 1e-3 and dg["volume"].max() < 1e-5
    assert dg["centroid_x"].max() < 1e-3 and dg["centroid_y"].max() < 1e-3 and dg["bottom_length"].max() < 0.1
