"""CPU FP64 oracle of the learned proposal's network -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the GPU kernels in
<pkg>/csrc/proposal.hip (k_mlp_forward, k_mlp_grad).  The product never calls it.

The network (no reference counterpart): x = bf16((features - mean) * scale), 40 inputs;
h = relu(x W1^T + b1), 64 hidden units, W1 in BF16; y = h . w2 + b2 (predicted log DTW
distance).  Loss: mean squared error over the batch.  The GPU accumulates in FP32 on BF16
operands and rounds the hidden gradient dH to BF16 before the weight-gradient product; this
oracle evaluates the same expressions in FP64 on the same BF16 operands (quantize_dh=True), or
the exact gradient of the unquantised expression (quantize_dh=False), which the finite-
difference test pins.
"""
import numpy as np


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def network_input(features, mean, scale):
    f = np.asarray(features, dtype=np.float32)
    return bf16_round((f - np.asarray(mean, np.float32)) * np.asarray(scale, np.float32)).astype(np.float64)


def forward(features, w1, b1, w2, b2, mean, scale, quantize=True):
    """Returns (y, h, x) in FP64."""
    x = network_input(features, mean, scale) if quantize else (
        (np.asarray(features, np.float64) - np.asarray(mean, np.float64)) * np.asarray(scale, np.float64))
    W1 = bf16_round(w1).astype(np.float64) if quantize else np.asarray(w1, np.float64)
    h = np.maximum(x @ W1.T + np.asarray(b1, np.float64), 0.0)
    y = h @ np.asarray(w2, np.float64) + float(b2)
    return y, h, x


def loss_and_grad(features, targets, w1, b1, w2, b2, mean, scale, quantize=True, quantize_dh=True):
    """Mean squared error and its gradient (dW1 [64 x 40], db1, dw2, db2)."""
    y, h, x = forward(features, w1, b1, w2, b2, mean, scale, quantize)
    t = np.asarray(targets, np.float64)
    B = y.shape[0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / B
    dh = dy[:, None] * np.asarray(w2, np.float64)[None, :] * (h > 0)
    if quantize_dh:
        dh = bf16_round(dh.astype(np.float32)).astype(np.float64)
    dW1 = dh.T @ x
    db1 = dh.sum(axis=0)
    dw2 = h.T @ dy
    db2 = float(dy.sum())
    return loss, dW1, db1, dw2, db2
