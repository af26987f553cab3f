"""Checkpoint and restart on the GPU: a run of the rising bubble (FP64,
previous-pressure starts, 200 tracked particles) saved after 12 steps and
resumed in a fresh context repeats the uninterrupted run bit for bit over
the next 12 steps (interface, velocity, pressure, particles, time); a state
saved from another grid is refused; the VTR file of the GPU state holds the
device fields as Float32."""
import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def _particle_params(case):
    f = case.fluids
    return dict(rho_l=f.rho_l, rho_g=f.rho_g, nu_l=f.mu_l / f.rho_l, nu_g=f.mu_g / f.rho_g, rho_p=2500.0, gx=0.0,
                gy=-f.g, lref=1.0, uref=1.0, us=0.0, vs=0.0)


def _advance(st, steps, P):
    for _ in range(steps):
        rep = st.step()
        assert rep["poisson"]["status"] == 0
        st.pj.particles_track(rep["dt"], P)


def test_restart_repeats_the_run(tmp_path):
    from simcoflow_amd.solver import TimeStepper
    case = C.case_bubble(64, 128)
    P = _particle_params(case)
    rng = np.random.default_rng(4)
    n = 200
    parts = (rng.uniform(0.05, 0.95, n), rng.uniform(0.05, 1.9, n), rng.normal(0, 0.01, n), rng.normal(0, 0.01, n),
             np.exp(rng.normal(np.log(1e-4), 0.3, n)))
    ck = str(tmp_path / "bubble.ckpt")
    print("[restart] bubble 64x128: 24 steps straight through, checkpoint at 12", flush=True)
    a = TimeStepper(case, precision=0, tol=1e-10, previous_pressure=True)
    try:
        a.pj.particles_set(*parts)
        _advance(a, 12, P)
        a.save(ck)
        _advance(a, 12, P)
        sa, ta, ia, pa = a.state(), a.time, a.itt, a.pj.particles_get()
    finally:
        a.close()
    print("[restart] fresh context from the checkpoint, 12 steps", flush=True)
    b = TimeStepper(case, precision=0, tol=1e-10, previous_pressure=True)
    try:
        b.load(ck)
        assert b.itt == 12
        _advance(b, 12, P)
        sb, tb, ib, pb = b.state(), b.time, b.itt, b.pj.particles_get()
        vtr = b.write_vtr(str(tmp_path))
    finally:
        b.close()
    assert (ia, ta) == (ib, tb)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y)
    from simcoflow_amd.output import read_vtr
    arr, ext = read_vtr(vtr)
    assert ext == [1, 64, 1, 128, 1, 1]
    assert np.array_equal(arr["Vof"], sb["vof"].astype(np.float32).ravel())
    assert np.array_equal(arr["Pressure"], sb["p"][1:-1, 1:-1].astype(np.float32).ravel())


def test_state_from_another_grid_is_refused():
    from simcoflow_amd._lib import SimcoError
    from simcoflow_amd.projection import Projection, params_for
    a = C.case_bubble(32, 64)
    b = C.case_bubble(64, 64)
    pa = Projection(a.grid, params_for(a.fluids))
    pb = Projection(b.grid, params_for(b.fluids))
    try:
        blob = pa.state_save()
        with pytest.raises(SimcoError):
            pb.state_load(blob)
        with pytest.raises(SimcoError):
            pa.state_load(b"X" * len(blob))
        pa.state_load(blob)
    finally:
        pa.close()
        pb.close()


def test_restart_of_a_moving_solid(tmp_path):
    """A heavy cylinder sinking in a tank (the whole moving-solid loop, device
    geometry): saved after 6 of 12 steps and resumed in a fresh context, the
    run repeats the uninterrupted one bit for bit -- fields, the solid's
    centre and velocity -- and the VTR file shows the moved solid.  Its mass
    is twice what it displaces, above its added mass, so the reference's
    explicit force coupling (the force of the step before) is stable; a body
    lighter than its added mass oscillates with growing amplitude."""
    from simcoflow_amd.output import read_vtr
    from simcoflow_amd.solver import TimeStepper
    n = 64
    g = C.uniform_grid(n, n, 1.0, 1.0)
    shape = ("circle", 0.5, 0.45, 0.12, 0.0, 0.0)
    solid = lambda X, Y: C.sdf_circle(X, Y, 0.5, 0.45, 0.12)
    geo = C.build_geometry(g, solid)

    def ls(X, Y):
        return Y - 0.8, np.zeros(X.shape), np.ones(X.shape)
    itf = C.interface_from_levelset(g, geo, ls)
    case = C.Case("tank", g, C.Fluids(), geo, itf, dt=0.25 / n, tol=1e-8, solid=solid, levelset=ls, shape=shape)

    def make():
        return TimeStepper(case, precision=0, tol=1e-10, device_geometry=True,
                           moving=dict(mass=2.0 * 1000.0 * np.pi * 0.12 ** 2, g=case.fluids.g, start=3))
    # every field a step leaves, the predictor's Dp included: a U/V cell the
    # solid made inactive keeps its last Dp, which a later Poisson assembly
    # reads, so the state must carry it
    def fields(st):
        return {f: st.pj.download(f) for f in ("U", "V", "P", "VOF", "PHI", "DPU", "DPV", "UPRED", "VPRED")}
    ck = str(tmp_path / "moving.ckpt")
    print("[restart] sinking cylinder 64^2: 12 steps straight through, checkpoint at 6", flush=True)
    a = make()
    steps_a = []
    try:
        for _ in range(6):
            a.step()
        a.save(ck)
        for _ in range(6):
            a.step()
            steps_a.append(fields(a))
        sa, ma = a.state(), dict(a.moving)
    finally:
        a.close()
    print("[restart] fresh context from the checkpoint, 6 steps", flush=True)
    b = make()
    try:
        b.load(ck)
        assert b.itt == 6
        for k in range(6):
            b.step()
            fb = fields(b)
            for f in fb:
                assert np.array_equal(steps_a[k][f], fb[f]), (7 + k, f)
        sb, mb = b.state(), dict(b.moving)
        vtr = b.write_vtr(str(tmp_path))
        vofS = b.pj.download("VOFS")
    finally:
        b.close()
    assert ma["shape"] == mb["shape"] and ma["vs"] == mb["vs"] and ma["shape"][2] < 0.45 and ma["vs"] < 0.0
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    arr, _ = read_vtr(vtr)
    assert np.array_equal(arr["SolidVof"], vofS.astype(np.float32).ravel())
    assert not np.array_equal(vofS, geo.vofS)
