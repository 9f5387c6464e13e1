"""The float64 MLP oracle (oracle/pxo_mlp.py) pinned by finite differences and BF16 known values."""
import numpy as np

import pxo_mlp


def rand_params(rng):
    return {"w1": rng.normal(0, 0.3, (64, 24)), "b1": rng.normal(0, 0.1, 64), "w2": rng.normal(0, 0.2, (64, 64)),
            "b2": rng.normal(0, 0.1, 64), "w3": rng.normal(0, 0.2, (3, 64)), "b3": rng.normal(0, 0.1, 3),
            "in_shift": rng.normal(0, 1, 24), "in_scale": rng.uniform(0.5, 2, 24), "out_scale": rng.uniform(0.5, 2, 3)}


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = rand_params(rng)
    X = rng.normal(0, 1, (40, 24))
    Y = rng.normal(0, 1, (40, 3))
    loss, g = pxo_mlp.loss_and_grads(p, X, Y)
    eps = 1e-6
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        flat = p[name].ravel()
        for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = pxo_mlp.loss_and_grads(p, X, Y)
            flat[idx] = old - eps
            lm, _ = pxo_mlp.loss_and_grads(p, X, Y)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name].ravel()[idx]) <= 1e-6 + 1e-5 * abs(fd), (name, idx)


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.140625, 65504.0, 1e-3], np.float32)
    r = pxo_mlp.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0            # halfway between 1 and 1.0078125: ties to even (1.0)
    assert r[2] == 1.0078125      # above halfway rounds up
    assert r[3] == -3.140625      # exactly representable
    assert abs(r[5] - 1e-3) / 1e-3 < 2 ** -8
