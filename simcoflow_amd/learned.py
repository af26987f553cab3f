"""Learned warm start of the projection solve.

The network (a 4-level convolutional encoder-decoder, 16/32/64/128
channels, 3x3 kernels, BF16 weights, FP32 accumulation) runs inside the
library on a coarse level of the solve's multigrid hierarchy; this module
holds the host side of its life cycle, all of it over the C ABI:

* ``collect_pairs``: (input, pressure) pairs made by the plain FP64 solver
  itself on seeded variants of a case (the right-hand side varies through
  seeded smooth predictor velocities on the device);
* ``Trainer``: forward, gradient and Adam steps of the library's own HIP
  kernels on batches held in device memory;
* ``save`` / ``load``: the weight file (numpy ``.npz``: FP32 weights, the
  level, the output scale), loaded with ``allow_pickle=False``;
* ``enable``: switch a ``Projection`` context to the learned warm start.

Device buffers for the training data are torch tensors (plumbing only: the
library reads and writes them by pointer).  No counterpart in the reference,
which starts PCG from zero (ProjectionP_Mod.f90:319).
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check

_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)


def num_weights():
    return int(_lib.load().sg_nn_num_weights())


def init_weights(seed=0):
    w = np.zeros(num_weights(), dtype=np.float32)
    check(_lib.load().sg_nn_init_weights(w.ctypes.data_as(_FP), int(seed)))
    return w


# layer table of the network (input channels, output channels, level); the
# weights of layer l are a (cout, kpad(cin)) row-major block, column
# tap*cin + ci, tap = 3*dy + dx, zero beyond 9*cin
SPEC = [(8, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 3),
        (128, 128, 3), (128, 64, 2), (64, 64, 2), (64, 32, 1), (32, 32, 1), (32, 16, 0), (16, 16, 0),
        (16, 16, 0)]


def kpad(c):
    return (9 * c + 31) // 32 * 32


def split_weights(w):
    """The flat weight vector as per-layer (cout, cin, 3, 3) arrays (the
    cross-correlation kernels of torch.nn.functional.conv2d)."""
    out = []
    o = 0
    for cin, cout, _ in SPEC:
        kp = kpad(cin)
        blk = np.asarray(w[o:o + cout * kp], dtype=np.float64).reshape(cout, kp)[:, :9 * cin]
        out.append(blk.reshape(cout, 3, 3, cin).transpose(0, 3, 1, 2).copy())
        o += cout * kp
    return out


class Trainer:
    """A net on level grids of H x W (H % 8 == 0, W % 128 == 0) with batch
    capacity `batch`.  x: torch bfloat16 (nb, H+2, W+2, 8) with a zero
    border; t, mask: torch float32 (nb, H, W); all on the GPU."""

    def __init__(self, H, W, batch, weights=None, seed=0):
        self.lib = _lib.load()
        self.H, self.W, self.batch = H, W, batch
        h = ctypes.c_void_p()
        check(self.lib.sg_nn_create(H, W, batch, ctypes.byref(h)))
        self.h = h
        self.set_weights(init_weights(seed) if weights is None else weights)

    def close(self):
        if self.h:
            self.lib.sg_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == num_weights()
        check(self.lib.sg_nn_set_weights(self.h, w.ctypes.data_as(_FP)))

    def get_weights(self):
        w = np.zeros(num_weights(), dtype=np.float32)
        check(self.lib.sg_nn_get_weights(self.h, w.ctypes.data_as(_FP)))
        return w

    def forward(self, x, y, nb):
        check(self.lib.sg_nn_forward(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), nb))

    def gradient(self, x, t, m, nb):
        g = np.zeros(num_weights(), dtype=np.float32)
        loss = ctypes.c_float(0.0)
        check(self.lib.sg_nn_gradient(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                      ctypes.c_void_p(m.data_ptr()), nb, g.ctypes.data_as(_FP),
                                      ctypes.byref(loss)))
        return g, loss.value

    def step(self, x, t, m, nb, lr, want_loss=False):
        loss = ctypes.c_float(0.0)
        check(self.lib.sg_nn_train_step(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                         ctypes.c_void_p(m.data_ptr()), nb, float(lr),
                                         ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None


def level_dims(pj, level):
    H, W = ctypes.c_int(), ctypes.c_int()
    check(pj.lib.sg_nn_level_dims(pj.ctx, int(level), ctypes.byref(H), ctypes.byref(W)))
    return H.value, W.value


def pick_level(pj, max_width=256):
    """The first multigrid level no wider than max_width whose grid the net
    accepts (nj % 8 == 0, ni % 128 == 0)."""
    for lv in range(32):
        try:
            H, W = level_dims(pj, lv)
        except _lib.SimcoError:
            break
        if W <= max_width and H % 8 == 0 and W % 128 == 0:
            return lv
    raise ValueError("no multigrid level fits the net (needs ni % 128 == 0, nj % 8 == 0 on some level)")


def features(pj, level, x, t=None, m=None):
    """Net input (and target / mask) of the last solved system into the
    torch buffers x (H+2, W+2, 8) bf16, t, m (H, W) float32; returns the
    rms scale of the restricted right-hand side."""
    sc = ctypes.c_double(0.0)
    check(pj.lib.sg_nn_features(pj.ctx, int(level), ctypes.c_void_p(x.data_ptr()),
                                ctypes.c_void_p(t.data_ptr()) if t is not None else None,
                                ctypes.c_void_p(m.data_ptr()) if m is not None else None, ctypes.byref(sc)))
    return sc.value


def smooth_velocities(torch, shape, g_dt, gen, modes=8, amp=0.2):
    """Seeded smooth predictor velocities on the device (ghosted (nj+2,
    ni+2) float64): a sum of random separable sine/cosine modes plus a
    gravity increment -g dt s (s in [0.5, 1.5]), walls closed as the
    reference's predictor leaves them (u = 0 on the west/east faces, v = 0
    on the bottom face)."""
    nj2, ni2 = shape
    dev = "cuda"
    X = torch.arange(ni2, device=dev, dtype=torch.float64)[None, :] / (ni2 - 2)
    Y = torch.arange(nj2, device=dev, dtype=torch.float64)[:, None] / (nj2 - 2)

    def field():
        f = torch.zeros(shape, device=dev, dtype=torch.float64)
        for _ in range(modes):
            kx, ky = (torch.rand(2, generator=gen, dtype=torch.float64) * 6.0 + 0.5).tolist()
            px, py = (torch.rand(2, generator=gen, dtype=torch.float64) * 6.283185307179586).tolist()
            a = float(torch.randn(1, generator=gen, dtype=torch.float64)) * amp / modes ** 0.5
            f += a * torch.sin(6.283185307179586 * kx * X + px) * torch.cos(6.283185307179586 * ky * Y + py)
        return f
    u = field()
    v = field() - g_dt * (0.5 + float(torch.rand(1, generator=gen, dtype=torch.float64)))
    ni = ni2 - 2
    u[:, 0] = 0.0
    u[:, ni] = 0.0
    u[:, ni + 1] = 0.0
    v[0, :] = 0.0
    return u.contiguous(), v.contiguous()


def collect_pairs(make_case, seeds, per_case, level=None, tol=1e-9, progress=None, seed=1234, max_width=256):
    """(x, t, mask) training pairs from the plain FP64 solver: for each case
    seed, `per_case` right-hand sides from seeded smooth predictor
    velocities, on `level` (default: the first level no wider than
    max_width).  Returns torch tensors on the GPU and the level."""
    import torch
    from . import cases as C
    from .projection import Projection, params_for
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    X = T = M = None
    k = 0
    n = len(seeds) * per_case
    t0 = time.time()
    for s in seeds:
        case = make_case(s)
        g = case.grid
        _, _, DpU, DpV = C.predictor_fields(case, seed=s + 7)
        pj = Projection(g, params_for(case.fluids, tol=tol, max_iter=500))
        try:
            pj.set_geometry(case.geo)
            pj.set_interface(case.itf.vof, case.itf.phi)
            pj.upload("DPU", DpU)
            pj.upload("DPV", DpV)
            g_dt = case.fluids.g * case.dt
            for r in range(per_case):
                u, v = smooth_velocities(torch, (g.nj + 2, g.ni + 2), g_dt, gen)
                torch.cuda.synchronize()
                check(pj.lib.sg_upload_device(pj.ctx, _lib.F["UPRED"], ctypes.c_void_p(u.data_ptr())))
                check(pj.lib.sg_upload_device(pj.ctx, _lib.F["VPRED"], ctypes.c_void_p(v.data_ptr())))
                pj.build(0.0)
                pj.solve()
                if X is None:
                    if level is None:
                        level = pick_level(pj, max_width)
                    H, W = level_dims(pj, level)
                    X = torch.zeros((n, H + 2, W + 2, 8), device="cuda", dtype=torch.bfloat16)
                    T = torch.zeros((n, H, W), device="cuda", dtype=torch.float32)
                    M = torch.zeros((n, H, W), device="cuda", dtype=torch.float32)
                    torch.cuda.synchronize()
                features(pj, level, X[k], T[k], M[k])
                k += 1
            if progress:
                progress(f"pairs {k}/{n} ({time.time() - t0:.1f} s)")
        finally:
            pj.close()
    return X, T, M, level


def collect_step_pairs(make_case, seeds, steps, skip=2, level=None, tol=1e-9, progress=None, max_width=256):
    """(x, t, mask) training pairs from time stepping: for each case seed,
    the right-hand side and pressure of every Poisson solve of `steps` full
    time steps (TimeStepper, FP64 solver) after `skip` warm-up steps, on
    `level` (default: the first level no wider than max_width).  These are
    the systems the solver meets in a run, unlike collect_pairs' seeded
    predictor fields.  Returns torch tensors on the GPU and the level."""
    import torch
    from .solver import TimeStepper
    X = T = M = None
    k = 0
    n = len(seeds) * steps
    t0 = time.time()
    for s in seeds:
        st = TimeStepper(make_case(s), precision=_lib.SG_PREC_FP64, tol=tol, max_iter=500)
        pj = st.pj
        try:
            for _ in range(skip):
                st.step()
            for _ in range(steps):
                st.step()
                if X is None:
                    if level is None:
                        level = pick_level(pj, max_width)
                    H, W = level_dims(pj, level)
                    X = torch.zeros((n, H + 2, W + 2, 8), device="cuda", dtype=torch.bfloat16)
                    T = torch.zeros((n, H, W), device="cuda", dtype=torch.float32)
                    M = torch.zeros((n, H, W), device="cuda", dtype=torch.float32)
                    torch.cuda.synchronize()
                features(pj, level, X[k], T[k], M[k])
                k += 1
            if progress:
                progress(f"step pairs {k}/{n} ({time.time() - t0:.1f} s)")
        finally:
            st.close()
    return X, T, M, level


def train(X, T, M, epochs=10, batch=16, lr=2e-3, seed=0, weights=None, progress=None, lr_final=2e-4):
    """Adam on the normalised targets T / K (K = rms of T over the fluid
    points); cosine learning-rate decay from lr to lr_final.  Returns (weights,
    K, loss history)."""
    import torch
    n, H2, W2, _ = X.shape
    H, W = H2 - 2, W2 - 2
    K = float(torch.sqrt((T * T * M).sum() / M.sum().clamp_min(1.0)))
    Tn = (T / K).contiguous()
    tr = Trainer(H, W, batch, weights=weights, seed=seed)
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    hist = []
    steps = epochs * max(1, n // batch)
    it = 0
    try:
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen)
            tot = 0.0
            cnt = 0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch].cuda()
                xb = X.index_select(0, idx).contiguous()
                tb = Tn.index_select(0, idx).contiguous()
                mb = M.index_select(0, idx).contiguous()
                torch.cuda.synchronize()
                cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * it / max(1, steps)))
                loss = tr.step(xb, tb, mb, batch, cur, want_loss=True)
                tot += loss
                cnt += 1
                it += 1
            hist.append(tot / max(cnt, 1))
            if progress:
                progress(f"epoch {ep}: loss {hist[-1]:.4e}")
        w = tr.get_weights()
    finally:
        tr.close()
    return w, K, hist


def save(path, weights, level, out_scale, **meta):
    np.savez(path, weights=np.asarray(weights, dtype=np.float32), level=np.int64(level),
             out_scale=np.float64(out_scale), **{k: np.asarray(v) for k, v in meta.items()})


def load(path):
    with np.load(path, allow_pickle=False) as d:
        return d["weights"].astype(np.float32), int(d["level"]), float(d["out_scale"])


def set_prediction(pj, level, e, deflate=False):
    """A level-`level` prediction (array of the level's (H, W) points) in
    place of the net's output for warm_start 2 / 3; None switches it off."""
    if e is None:
        check(pj.lib.sg_set_learned_prediction(pj.ctx, int(level), None, 0))
        return
    H, W = level_dims(pj, level)
    a = np.ascontiguousarray(e, dtype=np.float64)
    if a.shape != (H, W):
        raise ValueError(f"prediction of shape {a.shape}, level {level} has {(H, W)}")
    check(pj.lib.sg_set_learned_prediction(pj.ctx, int(level), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           int(bool(deflate))))


def sample_at_level(pj, level, fine):
    """The fine cell array sampled at a level's points the way the training
    targets are: the point covering the block's first cell."""
    H, W = level_dims(pj, level)
    nj, ni = fine.shape
    sy = nj // H if H else 1
    sx = max(1, ni // W)
    return np.ascontiguousarray(fine[::sy, ::sx][:H, :W])


def enable(pj, weights, level, out_scale, deflate=True):
    """Switch a Projection to the learned warm start (solves with
    warm_start = 2 use it)."""
    w = np.ascontiguousarray(weights, dtype=np.float32)
    check(pj.lib.sg_set_learned(pj.ctx, w.ctypes.data_as(_FP), int(level), float(out_scale), int(bool(deflate))))


def disable(pj):
    check(pj.lib.sg_set_learned(pj.ctx, None, 0, 1.0, 0))
