"""The allocation-free CFL step (sg_time_step_size_noalloc) against the
context entry point sg_time_step_size on the same device fields: bit for
bit, on two calls in a row, the second on a smaller grid, so the shared
scratch array holds the first call's partial minima when the second starts.
Its argument checks run on the host and launch nothing."""
import ctypes

import numpy as np
import pytest

from simcoflow_amd import cases as C

pytestmark = pytest.mark.gpu


def _dt_both(g, seed, vint, vs, asy):
    import torch
    from simcoflow_amd._lib import check
    from simcoflow_amd.projection import Projection, default_params
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((g.nj + 2, g.ni + 2)) * 3.0
    v = rng.standard_normal((g.nj + 2, g.ni + 2)) * 3.0
    u[rng.uniform(size=u.shape) < 0.1] = 0.0
    v[rng.uniform(size=v.shape) < 0.1] = 0.0
    fr = 0.3
    pj = Projection(g, default_params(Fr=fr))
    try:
        pj.upload("U", u)
        pj.upload("V", v)
        ref = ctypes.c_double()
        check(pj.lib.sg_time_step_size(pj.ctx, 0.25, vint, 1.3, vs, asy, ctypes.byref(ref)))
        dev = torch.device("cuda:0")
        udx = torch.from_numpy(np.ascontiguousarray(g.udx)).to(dev)
        vdy = torch.from_numpy(np.ascontiguousarray(g.vdy)).to(dev)
        out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        rc = pj.lib.sg_time_step_size_noalloc(g.ni, g.nj, udx.data_ptr(), vdy.data_ptr(), pj.field_ptr("U"),
                                              pj.field_ptr("V"), 0.25, fr, vint, 1.3, vs, asy, out.data_ptr(), None)
        check(rc)
        torch.cuda.synchronize()
        got = float(out.cpu()[0])
    finally:
        pj.close()
    return got, ref.value


def test_noalloc_time_step_matches_context_entry_point():
    print("noalloc: 1200x900 then 96x40", flush=True)
    # 1.08M cells: 4219 blocks of 256 capped at the scratch's 2048 partials
    big = _dt_both(C.uniform_grid(1200, 900, 3.0, 2.0), 0, 2.5, 0.3, -1.2)
    small = _dt_both(C.uniform_grid(96, 40, 3.0, 1.0), 1, 0.0, 0.0, 0.0)
    for got, ref in (big, small):
        assert got == ref, (got, ref)
    assert big[0] != small[0]


def test_noalloc_time_step_rejects_bad_sizes():
    from simcoflow_amd._lib import SG_ERR_ARG, load
    L = load()
    one = ctypes.c_double(0.0)
    p = ctypes.cast(ctypes.pointer(one), ctypes.c_void_p)
    # host-side checks: nothing is launched
    assert L.sg_time_step_size_noalloc(0, 10, p, p, p, p, 0.25, 0.3, 0.0, 1.0, 0.0, 0.0, p, None) == SG_ERR_ARG
    assert L.sg_time_step_size_noalloc(70000, 70000, p, p, p, p, 0.25, 0.3, 0.0, 1.0, 0.0, 0.0, p, None) == SG_ERR_ARG
    assert L.sg_time_step_size_noalloc(10, 10, None, p, p, p, 0.25, 0.3, 0.0, 1.0, 0.0, 0.0, p, None) == SG_ERR_ARG
