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
