"""Per-step diagnostics recorded on the device (sg_diagnostics_record)
against the same quantities computed on the host from the downloaded
fields: dt, the maxima of |mres|, |u| and |v| bit for bit (a maximum does
not depend on the order); the volumes and the gas centroid to 1e-13
relative (two-phase sums in another order).  The ring: rows come back in
recording order, partial reads resume where they stopped, overwritten rows
are counted, a NaN in a field shows as NaN."""
import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def _expected(g, vof, vofS, u, v, mres, dt):
    area = np.outer(g.dy, g.dx)
    gas = 1.0 - vof - vofS
    gv = float((gas * area).sum())
    return np.array([dt, float((vof * area).sum()), gv, float((gas * area * g.xc[None, :]).sum()) / gv,
                     float((gas * area * g.yc[:, None]).sum()) / gv, float(np.abs(mres).max()),
                     float(np.abs(u[1:-1, 1:-1]).max()), float(np.abs(v[1:-1, 1:-1]).max())])


def test_rows_match_host_reductions():
    from simcoflow_amd.solver import TimeStepper
    case = C.case_dambreak(128, 64)
    g = case.grid
    print("[diagnostics] dam break 128x64, 15 steps", flush=True)
    st = TimeStepper(case, precision=0, tol=1e-10, diagnostics=True)
    try:
        want = []
        for _ in range(15):
            rep = st.step()
            s = st.state()
            mres = st.pj.download("MRES")
            want.append(_expected(g, s["vof"], case.geo.vofS, s["u"], s["v"], mres, rep["dt"]))
        h = st.history()
        assert st.history()["dt"].size == 0
    finally:
        st.close()
    from simcoflow_amd._lib import DIAG_NAMES
    got = np.stack([h[k] for k in DIAG_NAMES], axis=1)
    want = np.array(want)
    assert got.shape == (15, 8)
    for c in (0, 5, 6, 7):
        assert np.array_equal(got[:, c], want[:, c]), DIAG_NAMES[c]
    for c in (1, 2, 3, 4):
        assert np.all(np.abs(got[:, c] - want[:, c]) <= 1e-13 * np.abs(want[:, c])), DIAG_NAMES[c]
    assert np.all(want[:, 5] > 0.0)  # the mass residual is live, not a zero field


def test_ring_order_partial_reads_overflow_and_nan():
    from simcoflow_amd._lib import DIAG_RING
    from simcoflow_amd.projection import Projection, params_for
    case = C.case_bubble(32, 64)
    g = case.grid
    pj = Projection(g, params_for(case.fluids))
    try:
        pj.set_geometry(case.geo)
        pj.upload("VOF", case.itf.vof)
        u = np.zeros((g.nj + 2, g.ni + 2))
        print("[diagnostics] ring: ordered rows, partial reads", flush=True)
        for k in range(5):
            u[3, 4] = float(k + 1)
            pj.upload("U", u)
            pj.diagnostics_record()
        a, lost = pj.diagnostics_read(2)
        b, _ = pj.diagnostics_read()
        assert lost == 0 and a.shape[0] == 2 and b.shape[0] == 3
        assert list(np.concatenate([a, b])[:, 6]) == [1.0, 2.0, 3.0, 4.0, 5.0]
        print("[diagnostics] ring: overflow", flush=True)
        for _ in range(DIAG_RING + 4):
            pj.diagnostics_record()
        rows, lost = pj.diagnostics_read()
        assert rows.shape[0] == DIAG_RING and lost == 4
        assert pj.diagnostics_read()[0].shape[0] == 0
        u[5, 6] = np.nan
        pj.upload("U", u)
        pj.diagnostics_record()
        rows, _ = pj.diagnostics_read()
        assert np.isnan(rows[0, 6]) and rows[0, 7] == 0.0
    finally:
        pj.close()
