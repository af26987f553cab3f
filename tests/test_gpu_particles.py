"""The HIP particle tracker (TrackingParticles, Particles.f90:71-249) against
the CPU oracle through the C ABI.

Particles are independent of each other, so the device result of any subset
can be checked against the oracle run on that subset alone: small sets on
the dam break (a solid block, an interface, particles outside the grid and
past the last cell centre) over several steps, and a sample of a million
particles on a 512^2 bubble swarm.  Tolerance: the drag law's pow and the
integrator's exp are the device library's on one side and the C library's
on the other (each within an ulp or two), so positions and velocities agree
to 1e-12 of their scale rather than bit for bit; every particle's branch
(fluid, solid, outside, parked) and the failure count agree exactly."""
import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def _params(case, **kw):
    f = case.fluids
    p = dict(rho_l=f.rho_l, rho_g=f.rho_g, nu_l=f.mu_l / f.rho_l, nu_g=f.mu_g / f.rho_g, rho_p=2500.0, gx=0.0,
             gy=-f.g, lref=1.0, uref=1.0, us=0.0, vs=0.0)
    p.update(kw)
    return p


def _particles(g, n, seed, with_edges=True):
    rng = np.random.default_rng(seed)
    lx, ly = g.xc[-1] + g.dx[-1] / 2, g.yc[-1] + g.dy[-1] / 2
    x = rng.uniform(0.0, lx, n)
    y = rng.uniform(0.0, ly, n)
    if with_edges:
        k = n // 8
        x[:k] = g.xc[0] - g.dx[0] / 2 + g.dx[0] * rng.integers(0, g.ni + 1, k)  # on cell edges
        x[k:k + 4] = [-0.1, lx + 0.1, g.xc[-1] + 1e-9, g.xc[-1]]  # outside / past the last centre
        y[k + 4:k + 8] = [-0.1, ly + 0.1, g.yc[-1] + 1e-9, g.yc[-1]]
    u = rng.normal(0.0, 0.05, n)
    v = rng.normal(0.0, 0.05, n)
    d = np.exp(rng.normal(np.log(50e-6), 0.5, n))
    return x, y, u, v, d


def _close(what, a, b, scale):
    err = np.abs(a - b)
    bad = ~(err <= 1e-12 * scale)
    if bad.any():
        k = int(np.argmax(np.where(bad, err, 0)))
        raise AssertionError(f"{what}: {int(bad.sum())} particles differ; worst #{k} gpu {a[k]!r} oracle {b[k]!r}")


def _device(case, u, v, parts, params, dt, steps):
    from simcoflow_amd.projection import Projection, params_for
    g = case.grid
    pj = Projection(g, params_for(case.fluids))
    try:
        pj.set_geometry(case.geo)
        pj.upload("VOF", case.itf.vof)
        pj.upload("U", u)
        pj.upload("V", v)
        pj.particles_set(*parts)
        flagged = [pj.particles_track(dt, params) for _ in range(steps)]
        return pj.particles_get(), flagged
    finally:
        pj.close()


def test_dambreak_particles_match_oracle(oracle_lib):
    case = C.case_dambreak(128, 64)
    g = case.grid
    u, v, _, _ = C.predictor_fields(case, seed=4, amp=0.3)
    parts = _particles(g, 3000, 7)
    # some particles inside the solid block, moving with the solid
    solid = np.argwhere(case.geo.vofS > 1.0 - 1e-3)
    assert len(solid) > 0
    parts[0][-len(solid):] = g.xc[solid[:, 1]]
    parts[1][-len(solid):] = g.yc[solid[:, 0]]
    P = _params(case, us=0.01, vs=-0.02)
    dt = 2.0e-3
    print(f"[particles] dambreak {g.ni}x{g.nj}, {len(parts[0])} particles, 5 steps", flush=True)
    (x, y, pu, pv), flagged = _device(case, u, v, parts, P, dt, 5)
    ox, oy, ou, ov, _ = parts
    ref_flags = []
    for _ in range(5):
        ox, oy, ou, ov, bad = oracle_lib.particle_track(g, case.itf.vof, case.geo.vofS, u, v, ox, oy, ou, ov,
                                                        parts[4], P, dt)
        ref_flags.append(bad)
    assert flagged == ref_flags
    parked = ox == 1.0e3
    assert parked.sum() >= 2 and np.array_equal(x == 1.0e3, parked)
    assert np.array_equal(x == parts[0], ox == parts[0])  # the particles left alone
    _close("x", x, ox, 4.0)
    _close("y", y, oy, 4.0)
    _close("u", pu, ou, 1.0)
    _close("v", pv, ov, 1.0)


def test_million_particle_sample_matches_oracle(oracle_lib):
    case = C.case_swarm(512, 512, seed=5)
    g = case.grid
    u, v, _, _ = C.predictor_fields(case, seed=2, amp=0.2)
    n = 1 << 20
    parts = _particles(g, n, 11, with_edges=False)
    P = _params(case)
    dt = 1.0e-3
    print(f"[particles] swarm {g.ni}x{g.nj}, {n} particles, 2 steps", flush=True)
    (x, y, pu, pv), flagged = _device(case, u, v, parts, P, dt, 2)
    assert flagged == [0, 0]
    sel = np.random.default_rng(0).choice(n, 4000, replace=False)
    o = [a[sel] for a in parts]
    for _ in range(2):
        o[0], o[1], o[2], o[3], bad = oracle_lib.particle_track(g, case.itf.vof, case.geo.vofS, u, v, o[0], o[1],
                                                                o[2], o[3], o[4], P, dt)
        assert bad == 0
    _close("x", x[sel], o[0], 1.0)
    _close("y", y[sel], o[1], 1.0)
    _close("u", pu[sel], o[2], 1.0)
    _close("v", pv[sel], o[3], 1.0)
