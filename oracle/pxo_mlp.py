"""ORACLE -- test infrastructure only.

float64 numpy restatement of the learned warm start's network: forward, loss and analytic
gradients (checked by finite differences in tests/test_mlp_oracle.py), plus a forward pass that
rounds inputs, weights and hidden activations to BF16 the way the GPU's matrix-core inference does.
"""
import numpy as np


def forward(p, X):
    x = (np.asarray(X, np.float64) - p["in_shift"]) * p["in_scale"]
    h1 = np.maximum(x @ p["w1"].T + p["b1"], 0.0)
    h2 = np.maximum(h1 @ p["w2"].T + p["b2"], 0.0)
    y = h2 @ p["w3"].T + p["b3"]
    return x, h1, h2, y


def predict(p, X):
    return forward(p, X)[3] * p["out_scale"]


def loss_and_grads(p, X, Y):
    """Loss = mean over samples of sum over outputs of (y - Y / out_scale)^2."""
    x, h1, h2, y = forward(p, X)
    t = np.asarray(Y, np.float64) / p["out_scale"]
    n = len(X)
    err = y - t
    loss = (err ** 2).sum() / n
    dy = 2.0 * err / n
    g = {"w3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ p["w3"]) * (h2 > 0)
    g["w2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["w2"]) * (h1 > 0)
    g["w1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return loss, g


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def predict_bf16(p, X):
    """The GPU inference's arithmetic: BF16 inputs/weights/activations, FP32-exact products."""
    f32 = lambda a: np.asarray(a, np.float32)
    x = bf16((f32(X) - f32(p["in_shift"])) * f32(p["in_scale"])).astype(np.float64)
    h1 = bf16(np.maximum(x @ bf16(p["w1"]).astype(np.float64).T + p["b1"], 0.0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ bf16(p["w2"]).astype(np.float64).T + p["b2"], 0.0).astype(np.float32)).astype(np.float64)
    y = h2 @ bf16(p["w3"]).astype(np.float64).T + p["b3"]
    return y * p["out_scale"]
