"""Learned warm start: dataset recording, GPU training and inference of the 24-64-64-3 MLP.

The network predicts the converged (normal, tangent1, tangent2) impulses of a contact point from
24 pre-solve features (see include/pxg.h). The library applies it to warm-start misses only; the
plain solver then runs as usual, so the residual judges the prediction.
"""
import ctypes
import os

import numpy as np

from . import _abi, scenes
from ._abi import ptr
from ._lib import lib, check_mlp

N_IN, N_H, N_OUT = 24, 64, 3


class MlpParams(ctypes.Structure):
    """pxg_mlp (include/pxg.h): FP32 weights [out][in], biases and the input/output scaling."""
    _fields_ = [("w1", ctypes.c_float * (N_H * N_IN)), ("b1", ctypes.c_float * N_H),
                ("w2", ctypes.c_float * (N_H * N_H)), ("b2", ctypes.c_float * N_H),
                ("w3", ctypes.c_float * (N_OUT * N_H)), ("b3", ctypes.c_float * N_OUT),
                ("in_shift", ctypes.c_float * N_IN), ("in_scale", ctypes.c_float * N_IN),
                ("out_scale", ctypes.c_float * N_OUT)]

    def arrays(self):
        return {f: np.ctypeslib.as_array(getattr(self, f)).copy() for f, _ in self._fields_}

    @classmethod
    def from_arrays(cls, d):
        p = cls()
        for f, _ in cls._fields_:
            np.ctypeslib.as_array(getattr(p, f))[:] = np.asarray(d[f], np.float32).ravel()
        return p

    def save(self, path):
        np.savez(path, **self.arrays())

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls.from_arrays({k: z[k] for k in z.files})


class TrainConfig(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("reserved0", ctypes.c_uint32), ("lr", ctypes.c_float), ("lr_decay", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("weight_decay", ctypes.c_float)]


def init_params(X, Y, seed=0):
    """He-initialised weights; inputs standardised by the dataset's mean / std, outputs scaled by
    their RMS."""
    rng = np.random.default_rng(seed)
    d = {}
    d["w1"] = rng.normal(0, np.sqrt(2.0 / N_IN), (N_H, N_IN))
    d["b1"] = np.zeros(N_H)
    d["w2"] = rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H))
    d["b2"] = np.zeros(N_H)
    d["w3"] = rng.normal(0, np.sqrt(1.0 / N_H), (N_OUT, N_H))
    d["b3"] = np.zeros(N_OUT)
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    d["in_shift"] = X.mean(0)
    sd = X.std(0)
    d["in_scale"] = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 0.0)
    rms = np.sqrt((Y ** 2).mean(0))
    d["out_scale"] = np.where(rms > 1e-9, rms, 1.0)
    return MlpParams.from_arrays(d)


def infer(net, X):
    X = np.ascontiguousarray(np.asarray(X, np.float32))
    out = np.zeros((len(X), N_OUT), np.float32)
    if len(X):
        check_mlp(lib().pxg_mlp_infer(ctypes.byref(net), ptr(X), len(X), ptr(out)), "pxg_mlp_infer")
    return out


def gradients(net, X, Y):
    X = np.ascontiguousarray(np.asarray(X, np.float32))
    Y = np.ascontiguousarray(np.asarray(Y, np.float32))
    g = MlpParams()
    loss = ctypes.c_float()
    check_mlp(lib().pxg_mlp_gradients(ctypes.byref(net), ptr(X), ptr(Y), len(X), ctypes.byref(g), ctypes.byref(loss)),
              "pxg_mlp_gradients")
    return g, loss.value


def train(net, X, Y, epochs=30, batch=2048, lr=2e-3, lr_decay=0.93, seed=1, weight_decay=0.0):
    X = np.ascontiguousarray(np.asarray(X, np.float32))
    Y = np.ascontiguousarray(np.asarray(Y, np.float32))
    c = TrainConfig()
    c.epochs, c.batch, c.seed = epochs, batch, seed
    c.lr, c.lr_decay, c.beta1, c.beta2, c.eps, c.weight_decay = lr, lr_decay, 0.9, 0.999, 1e-8, weight_decay
    losses = np.zeros(epochs, np.float32)
    check_mlp(lib().pxg_mlp_train(ctypes.byref(net), ptr(X), ptr(Y), len(X), ctypes.byref(c), ptr(losses)),
              "pxg_mlp_train")
    return net, losses


def record_dataset(seeds, steps=300, velocity_iterations=64, log=None, scene=None):
    """Runs config-1 piles (one per seed; or `scene(seed)`) with the plain solver at
    `velocity_iterations` and returns (features, converged impulses, missed) of every active contact
    point of every step."""
    from .world import World
    Xs, Ys, Ms = [], [], []
    for k, seed in enumerate(seeds):
        sc = scenes.config1_scene(seed=seed) if scene is None else scene(seed)
        cfg = _abi.default_config(**dict(sc.config, velocity_iterations=velocity_iterations))
        w = World(cfg, sc.bodies, sc.shapes)
        w.set_recording(True)
        for _ in range(steps):
            w.step_fixed()
            f, y, m = w.records()
            if len(f):
                Xs.append(f)
                Ys.append(y)
                Ms.append(m)
        w.close()
        if log:
            log("recorded pile %d/%d (seed %d)" % (k + 1, len(seeds), seed))
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms)


def subsample(X, Y, M, n, seed=0):
    """Keeps every warm-start miss and fills up to n records with randomly chosen hits."""
    rng = np.random.default_rng(seed)
    miss = np.flatnonzero(M != 0)
    hit = np.flatnonzero(M == 0)
    k = max(0, n - len(miss))
    pick = np.concatenate([miss, rng.choice(hit, size=min(k, len(hit)), replace=False)]) if k else miss[:n]
    rng.shuffle(pick)
    return X[pick], Y[pick], M[pick]


DEFAULT_NET = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mlp_config1.npz")


def train_config1(n_piles=32, records=200000, steps=300, epochs=30, seed0=1000, log=None):
    """The recipe of the learned warm start: `records` contact records (every warm-start miss, topped
    up with hits) from `n_piles` seeded config-1 piles solved at 64 velocity iterations, trained on
    the GPU. Returns (net, losses)."""
    X, Y, M = record_dataset(list(range(seed0, seed0 + n_piles)), steps=steps, velocity_iterations=64, log=log)
    X, Y, M = subsample(X, Y, M, records)
    net = init_params(X, Y, seed=0)
    net, losses = train(net, X, Y, epochs=epochs, batch=2048, lr=2e-3)
    return net, losses


def main():
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="train the learned warm start's network on the GPU")
    ap.add_argument("--out", default=DEFAULT_NET)
    ap.add_argument("--piles", type=int, default=32)
    ap.add_argument("--records", type=int, default=200000)
    ap.add_argument("--epochs", type=int, default=30)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    net, losses = train_config1(a.piles, a.records, epochs=a.epochs, log=log)
    net.save(a.out)
    log("loss %.4f -> %.4f, saved %s" % (losses[0], losses[-1], a.out))


if __name__ == "__main__":
    main()
