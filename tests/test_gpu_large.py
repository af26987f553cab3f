"""The largest configuration and the device-built geometry in time stepping.

* Dam break, 10 full steps with the geometry built on the device
  (TimeStepper(device_geometry=True)) equal 10 steps with the uploaded host
  geometry bit for bit.
* The largest configuration: the wave tank at 4096 x 2048 (8 x 4 m, cells of
  1/512 m, 8.4 M cells, an ellipse at the surface), mixed precision,
  geometry built on the device: 4 full steps with 10^5 tracked particles
  (log-normal d about 50 um) and a diagnostics row per step; every solve
  converges, the liquid volume holds to 1e-6 relative (1e-7 measured), the
  mass residual stays small, no particle fails.
* C4 for 100 steps, learned start in mixed precision against plain FP64."""
import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def test_device_geometry_steps_equal_host_geometry():
    from simcoflow_amd.solver import TimeStepper
    case = C.case_dambreak(128, 64)
    out = []
    for dev in (False, True):
        print(f"[large] dam break 128x64, 10 steps, device geometry {dev}", flush=True)
        st = TimeStepper(case, precision=0, tol=1e-10, device_geometry=dev)
        try:
            for _ in range(10):
                st.step()
            out.append(st.state())
        finally:
            st.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_largest_grid_with_particles():
    from simcoflow_amd.solver import TimeStepper
    print("[large] building the 4096x2048 wave tank", flush=True)
    case = C.case_waves(4096, 2048, seed=3)
    f = case.fluids
    P = dict(rho_l=f.rho_l, rho_g=f.rho_g, nu_l=f.mu_l / f.rho_l, nu_g=f.mu_g / f.rho_g, rho_p=2500.0, gx=0.0,
             gy=-f.g, lref=1.0, uref=1.0, us=0.0, vs=0.0)
    rng = np.random.default_rng(1)
    n = 100000
    parts = (rng.uniform(0.0, 8.0, n), rng.uniform(0.0, 1.9, n), np.zeros(n) + 1e-3, np.zeros(n),
             np.exp(rng.normal(np.log(50e-6), 0.5, n)))
    st = TimeStepper(case, precision=1, tol=1e-8, max_iter=300, diagnostics=True, device_geometry=True)
    try:
        st.pj.particles_set(*parts)
        for k in range(4):
            rep = st.step()
            print(f"[large] step {k}: {rep['poisson']['iterations']} iterations, relres {rep['poisson']['relres']:.2e}",
                  flush=True)
            assert rep["poisson"]["status"] == 0 and rep["poisson"]["relres"] < 5e-8, rep["poisson"]
            assert st.pj.particles_track(rep["dt"], P) == 0
        h = st.history()
        x, y, u, v = st.pj.particles_get()
    finally:
        st.close()
    vol = h["liquid_volume"]
    print(f"liquid volume {vol}, max |mres| {h['max_abs_mres']}")
    assert vol.size == 4
    # the sweeps' clipping at the three-phase cells of the ellipse and the
    # solid-cell quirks of X_Sweep / Y_Sweep (kept bit for bit) move the
    # liquid volume by ~1e-7 relative over these steps; no mass is lost on a
    # larger scale
    assert np.abs(vol - vol[0]).max() <= 1e-6 * vol[0]
    assert np.all(h["max_abs_mres"] < 1e-6)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))


def test_largest_grid_100_steps_learned_vs_fp64():
    """C4 for 100 steps: the 4096 x 2048 wave tank with 10^5 particles,
    stepped twice -- mixed precision with the learned start (the net, trained
    here on this tank's own first Poisson systems, corrects the previous
    pressure: warm_start = 3), and plain FP64 from x0 = 0.  Every solve
    converges to 1e-8; the two runs' liquid volumes agree step by step and the
    final interfaces agree, within bounds a little above the measured
    differences (volumes 8e-13 relative apart, final vof 7.5e-10 apart on
    average and 2.5e-3 at the worst cell, particles 7e-8 m; each run's own
    volume drifts by 1.9e-7 over the 100 steps); no particle fails in either
    run."""
    from simcoflow_amd import learned as Lr
    from simcoflow_amd.solver import TimeStepper
    print("[large] C4: training pairs from the 4096x2048 tank's first steps", flush=True)

    def mk(seed):
        return C.case_waves(4096, 2048, seed=3)
    X, T, M, level = Lr.collect_step_pairs(mk, seeds=[3], steps=16, skip=2, tol=1e-9)
    w, K, hist = Lr.train(X, T, M, epochs=12, batch=8, lr=3e-3)
    case = mk(3)
    f = case.fluids
    P = dict(rho_l=f.rho_l, rho_g=f.rho_g, nu_l=f.mu_l / f.rho_l, nu_g=f.mu_g / f.rho_g, rho_p=2500.0, gx=0.0,
             gy=-f.g, lref=1.0, uref=1.0, us=0.0, vs=0.0)
    rng = np.random.default_rng(1)
    n = 100000
    parts = (rng.uniform(0.0, 8.0, n), rng.uniform(0.0, 1.9, n), np.zeros(n) + 1e-3, np.zeros(n),
             np.exp(rng.normal(np.log(50e-6), 0.5, n)))
    runs = {}
    for name, kw in (("learned_mixed", dict(precision=1, learned=(w, level, K, False), previous_pressure=True)),
                     ("fp64", dict(precision=0))):
        print(f"[large] C4: 100 steps, {name}", flush=True)
        st = TimeStepper(case, tol=1e-8, max_iter=300, diagnostics=True, device_geometry=True, **kw)
        its = []
        try:
            st.pj.particles_set(*parts)
            for k in range(100):
                rep = st.step()
                assert rep["poisson"]["status"] == 0 and rep["poisson"]["relres"] < 5e-8, (name, k, rep["poisson"])
                assert st.pj.particles_track(rep["dt"], P) == 0, (name, k)
                its.append(rep["poisson"]["iterations"])
            h = st.history()
            vof = st.state()["vof"]
            x, y, _, _ = st.pj.particles_get()
        finally:
            st.close()
        runs[name] = dict(vol=h["liquid_volume"], mres=h["max_abs_mres"], vof=vof, x=x, y=y, its=its)
        print(f"[large] C4 {name}: iterations {its[:3]} ... {its[-3:]} (mean {np.mean(its):.1f}), volume drift "
              f"{(h['liquid_volume'][-1] - h['liquid_volume'][0]) / h['liquid_volume'][0]:.2e}", flush=True)
    a, b = runs["learned_mixed"], runs["fp64"]
    dvol = np.abs(a["vol"] - b["vol"]).max() / b["vol"][0]
    dvof = np.abs(a["vof"] - b["vof"])
    dpart = max(np.abs(a["x"] - b["x"]).max(), np.abs(a["y"] - b["y"]).max())
    print(f"[large] C4 learned_mixed vs fp64 over 100 steps: max volume diff {dvol:.2e} rel, final vof max "
          f"{dvof.max():.2e} / mean {dvof.mean():.2e}, particle positions {dpart:.2e} m", flush=True)
    assert a["vol"].size == 100 and b["vol"].size == 100
    for r in (a, b):
        assert np.abs(r["vol"] - r["vol"][0]).max() <= 1e-6 * r["vol"][0]
        assert np.all(r["mres"] < 1e-6)
    assert dvol < 1e-10
    assert dvof.max() < 0.05 and dvof.mean() < 1e-8
    assert dpart < 1e-5
