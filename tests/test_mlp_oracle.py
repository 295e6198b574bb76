"""Pins of the warm-start network oracle (oracle/mlp.py) on the CPU.

* BF16 rounding against hand-computed bit patterns (exact values, ties to even,
  carries into the exponent).
* Gradients of the unquantised float64 loss against central finite
  differences: a dropped activation derivative, a transposed weight or a wrong
  1/rows factor all fail it.
* The quantised gradient equals the unquantised gradient evaluated at the BF16
  weights (straight-through: gradients are taken w.r.t. the rounded values).
"""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-7, -2.5, 3.0e38, 1.0 + 2**-8 + 2**-20],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2**-6, 1.0 + 2**-7, -2.5, 2.9980278e38, 1.0 + 2**-7], dtype=np.float32)
    got = mlp.bf16(x)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 is a tie -> even (1+2^-6); above the tie rounds up
    assert np.array_equal(got[:5], want[:5])
    assert got[6] == want[6]
    assert (got.view(np.uint32) & 0xFFFF).max() == 0


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    widths = [5, 7, 6, 4]
    p = rng.normal(size=mlp.n_params(widths)) * 0.5
    X = rng.normal(size=(9, 5))
    Y = rng.normal(size=(9, 4))
    loss, g = mlp.loss_grad(p, widths, X, Y, quantize=False)
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp.loss_grad(p + e, widths, X, Y, quantize=False)[0]
                 - mlp.loss_grad(p - e, widths, X, Y, quantize=False)[0]) / (2 * h)
    assert np.abs(g - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
    # loss closed form: a network with zero weights predicts zero
    z = np.zeros_like(p)
    assert abs(mlp.loss_grad(z, widths, X, Y, quantize=False)[0] - 0.5 * np.sum(Y * Y) / 9) < 1e-12


def test_quantised_gradient_is_gradient_at_rounded_weights():
    rng = np.random.default_rng(4)
    widths = [3, 8, 2]
    p = rng.normal(size=mlp.n_params(widths)).astype(np.float32)
    pq = p.copy()
    for w_off, b_off, no, ni in mlp.layout(widths):
        pq[w_off:w_off + no * ni] = mlp.bf16(p[w_off:w_off + no * ni])
    X = rng.normal(size=(4, 3))
    Y = rng.normal(size=(4, 2))
    l1, g1 = mlp.loss_grad(p, widths, X, Y, quantize=True)
    l2, g2 = mlp.loss_grad(pq, widths, X, Y, quantize=False)
    assert abs(l1 - l2) < 1e-12 and np.abs(g1 - g2).max() < 1e-12
