"""Pins the CPU oracle of the learned proposal's network (oracle/mlp_oracle.py): BF16
rounding known answers, a hand-computed forward pass, and the analytic gradient against
central finite differences of the loss."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 0.0, 3.0e38], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 ties to 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 0.0, 2.9910e38], dtype=np.float32)
    got = mo.bf16_round(x)
    assert np.array_equal(got[:6], want[:6])
    assert abs(got[6] / 3.0e38 - 1) < 2 ** -8


def test_forward_hand_computed():
    # two inputs used, three hidden units; values exactly representable in BF16
    w1 = np.zeros((64, 40), np.float32)
    w1[0, 0], w1[0, 1] = 1.0, -2.0
    w1[1, 0], w1[1, 1] = 0.5, 0.5
    w1[2, 0] = -1.0
    b1 = np.zeros(64, np.float32)
    b1[1] = -1.0
    w2 = np.zeros(64, np.float32)
    w2[:3] = [1.0, 2.0, 4.0]
    f = np.zeros((2, 40), np.float32)
    f[0, :2] = [3.0, 1.0]   # h = relu(3-2, 1.5+0.5-1, -3) = (1, 1, 0) -> y = 1 + 2 + 0.25
    f[1, :2] = [-1.0, 2.0]  # h = relu(-5, 0.5-1, 1) = (0, 0, 1) -> y = 4 + 0.25
    y, h, _ = mo.forward(f, w1, b1, w2, 0.25, np.zeros(40), np.ones(40))
    assert np.allclose(h[0, :3], [1, 1, 0]) and np.allclose(h[1, :3], [0, 0, 1])
    assert np.allclose(y, [3.25, 4.25])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    B = 37
    f = rng.standard_normal((B, 40))
    t = rng.standard_normal(B)
    w1 = rng.standard_normal((64, 40)) * 0.3
    b1 = rng.standard_normal(64) * 0.1
    w2 = rng.standard_normal(64) * 0.3
    b2 = 0.1
    mean, scale = np.zeros(40), np.ones(40)
    loss, dW1, db1, dw2, db2 = mo.loss_and_grad(f, t, w1, b1, w2, b2, mean, scale, quantize=False, quantize_dh=False)

    def L(w1_, b1_, w2_, b2_):
        return mo.loss_and_grad(f, t, w1_, b1_, w2_, b2_, mean, scale, quantize=False, quantize_dh=False)[0]

    eps = 1e-6
    for (h, k) in [(0, 0), (5, 17), (63, 39), (30, 2)]:
        wp, wm = w1.copy(), w1.copy()
        wp[h, k] += eps
        wm[h, k] -= eps
        assert abs((L(wp, b1, w2, b2) - L(wm, b1, w2, b2)) / (2 * eps) - dW1[h, k]) < 1e-6 * (1 + abs(dW1[h, k]))
    for h in [0, 9, 63]:
        bp, bm = b1.copy(), b1.copy()
        bp[h] += eps
        bm[h] -= eps
        assert abs((L(w1, bp, w2, b2) - L(w1, bm, w2, b2)) / (2 * eps) - db1[h]) < 1e-6 * (1 + abs(db1[h]))
        vp, vm = w2.copy(), w2.copy()
        vp[h] += eps
        vm[h] -= eps
        assert abs((L(w1, b1, vp, b2) - L(w1, b1, vm, b2)) / (2 * eps) - dw2[h]) < 1e-6 * (1 + abs(dw2[h]))
    assert abs((L(w1, b1, w2, b2 + eps) - L(w1, b1, w2, b2 - eps)) / (2 * eps) - db2) < 1e-6 * (1 + abs(db2))
