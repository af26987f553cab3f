"""CPU pins of the learned warm-start network's reference (oracle/nn_reference)
and of the weight layout shared by the library and the reference."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _weights(rng, scale=1.0):
    from simcoflow_amd.learned import SPEC
    return [torch.from_numpy(rng.standard_normal((co, ci, 3, 3)) * scale * np.sqrt(2.0 / (9 * ci)))
            for ci, co, _ in SPEC]


def test_reference_gradient_matches_finite_differences():
    """Autograd of the FP64 reference (no BF16 rounding) against central
    differences along random weight directions on a tiny input."""
    from oracle import nn_reference as R
    rng = np.random.default_rng(0)
    Ws = _weights(rng)
    x = torch.from_numpy(rng.standard_normal((2, 8, 8, 8)))
    t = torch.from_numpy(rng.standard_normal((2, 8, 8)))
    m = torch.from_numpy((rng.uniform(size=(2, 8, 8)) < 0.8).astype(np.float64))
    loss, _, grads = R.loss_and_grads(x, t, m, Ws, bf16=False)
    for trial in range(3):
        d = [torch.from_numpy(rng.standard_normal(w.shape)) for w in Ws]
        eps = 1e-6
        lp = R.loss_and_grads(x, t, m, [w + eps * dd for w, dd in zip(Ws, d)], bf16=False)[0]
        lm = R.loss_and_grads(x, t, m, [w - eps * dd for w, dd in zip(Ws, d)], bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        an = float(sum((g * dd).sum() for g, dd in zip(grads, d)))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_reference_layers_are_wired_as_documented():
    """Zeroing the decoder's last convolution kills the output; with identity
    -like kernels on channel 0 the network passes a positive input through
    the encoder, the bottleneck and the skips (the output is a positive
    multiple of a smoothed input)."""
    from oracle import nn_reference as R
    rng = np.random.default_rng(1)
    Ws = _weights(rng)
    x = torch.from_numpy(rng.uniform(0.1, 1.0, (1, 8, 16, 16)))
    Ws0 = [w.clone() for w in Ws]
    Ws0[14] = torch.zeros_like(Ws0[14])
    assert float(R.forward(x, Ws0, bf16=False).abs().max()) == 0.0
    Wi = [torch.zeros_like(w) for w in Ws]
    for w in Wi:
        w[0, 0, 1, 1] = 1.0      # channel 0 -> channel 0, centre tap
    y = R.forward(x, Wi, bf16=False)
    # e0 = x0, d-path averages, decoder adds skips: y = x0 (skip e0) + up(...) > x0
    assert torch.all(y > x[:, 0])


def test_weight_layout_matches_torch_kernels():
    """Column tap*cin + ci of row co (tap = 3 dy + dx) is the torch kernel
    element (co, ci, dy, dx)."""
    from simcoflow_amd.learned import SPEC, kpad, num_weights, split_weights
    n = sum(co * kpad(ci) for ci, co, _ in SPEC)
    w = np.zeros(n, dtype=np.float32)
    # layer 2 (cin 16, cout 32), weights start after layers 0 and 1
    off = 16 * kpad(8) + 16 * kpad(16)
    co, ci, dy, dx = 5, 3, 2, 1
    w[off + co * kpad(16) + (3 * dy + dx) * 16 + ci] = 1.0
    Ws = split_weights(w)
    assert Ws[2][co, ci, dy, dx] == 1.0 and Ws[2].sum() == 1.0
    assert sum(a.sum() for a in Ws) == 1.0


def test_library_weight_init():
    """He-normal initial weights of the library: per-layer std ~ sqrt(2/(9
    cin)), zero padding columns, only row 0 of the output layer."""
    from simcoflow_amd.learned import SPEC, init_weights, kpad, num_weights
    assert num_weights() == sum(co * kpad(ci) for ci, co, _ in SPEC)
    w = init_weights(3)
    o = 0
    for l, (ci, co, _) in enumerate(SPEC):
        blk = w[o:o + co * kpad(ci)].reshape(co, kpad(ci))
        o += co * kpad(ci)
        assert np.all(blk[:, 9 * ci:] == 0.0)
        if l == len(SPEC) - 1:
            assert np.all(blk[1:] == 0.0)
            continue
        sd = blk[:, :9 * ci].std()
        assert abs(sd / np.sqrt(2.0 / (9 * ci)) - 1.0) < 0.15, (l, sd)
    assert not np.array_equal(w, init_weights(4))
