"""CPU FP64 reference of the learned warm-start network (torch autograd).

TEST INFRASTRUCTURE ONLY: the checker of the library's BF16 forward and
backward kernels; nothing in the product path imports it.

The network: 15 3x3 convolutions (padding 1, no bias), ReLU, 2x2 average
pooling, nearest upsampling, additive skips,

  e0a = relu(c0(x))      e0 = relu(c1(e0a))     d1 = pool(e0)
  e1a = relu(c2(d1))     e1 = relu(c3(e1a))     d2 = pool(e1)
  e2a = relu(c4(d2))     e2 = relu(c5(e2a))     d3 = pool(e2)
  e3a = relu(c6(d3))     e3 = relu(c7(e3a))
  g2a = relu(c8(up(e3)) + e2)   g2 = relu(c9(g2a))
  g1a = relu(c10(up(g2)) + e1)  g1 = relu(c11(g1a))
  g0a = relu(c12(up(g1)) + e0)  g0 = relu(c13(g0a))
  y = c14(g0)[channel 0]

loss = mean over all points of mask * (y - t)^2.

With ``bf16=True`` every tensor the GPU stores in BF16 is rounded to BF16
(round to nearest even) in the forward pass and its gradient is rounded in
the backward pass, and the gradient at y is rounded, so the reference sees
the same roundings as the GPU and differs only by FP32-vs-FP64 sums.
"""
import numpy as np
import torch
import torch.nn.functional as Fn


class _RoundBoth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


class _RoundGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


def forward(x, Ws, bf16=True):
    """x: (B, 8, H, W) float64; Ws: 15 tensors (cout, cin, 3, 3)."""
    R = _RoundBoth.apply if bf16 else (lambda t: t)

    def c(t, l):
        return Fn.conv2d(t, Ws[l], padding=1)

    def pool(t):
        return Fn.avg_pool2d(t, 2)

    def up(t):
        return t.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)

    r = torch.relu
    e0a = R(r(c(x, 0)))
    e0 = R(r(c(e0a, 1)))
    d1 = R(pool(e0))
    e1a = R(r(c(d1, 2)))
    e1 = R(r(c(e1a, 3)))
    d2 = R(pool(e1))
    e2a = R(r(c(d2, 4)))
    e2 = R(r(c(e2a, 5)))
    d3 = R(pool(e2))
    e3a = R(r(c(d3, 6)))
    e3 = R(r(c(e3a, 7)))
    u2 = R(up(e3))
    g2a = R(r(c(u2, 8) + e2))
    g2 = R(r(c(g2a, 9)))
    u1 = R(up(g2))
    g1a = R(r(c(u1, 10) + e1))
    g1 = R(r(c(g1a, 11)))
    u0 = R(up(g1))
    g0a = R(r(c(u0, 12) + e0))
    g0 = R(r(c(g0a, 13)))
    y = c(g0, 14)[:, 0]
    if bf16:
        y = _RoundGrad.apply(y)
    return y


def loss_and_grads(x, t, m, Ws, bf16=True):
    """x (B,8,H,W), t, m (B,H,W) float64; returns (loss, y, [dW_l])."""
    Ws = [w.detach().clone().requires_grad_(True) for w in Ws]
    y = forward(x, Ws, bf16=bf16)
    loss = (m * (y - t) ** 2).sum() / y.numel()
    loss.backward()
    return float(loss.detach()), y.detach(), [w.grad for w in Ws]


def bf16_round(a):
    """numpy float array rounded to BF16 (round to nearest even), as float64."""
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
