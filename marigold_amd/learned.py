"""Host side of the learned build method: training data and training run.

The split proposer is trained on the split decisions the plain builder emits
for the 65,536-triangle soup and the 1024x512-cell heightfield (every node with
more than 32 objects split by the binned sweep), by the library's own GPU
kernels (mgbvh_model_train_device).
"""
from __future__ import annotations

import time

import numpy as np

from . import _native as N
from . import scenes

# learned-method settings used by the scene builds (per-node-size dispatch,
# window half-width, guard); `autotune` may replace them
DEFAULT_LEARNED = {"learned_min_objects": 32768, "learned_window": 2, "learned_guard": 0.02}


def training_meshes():
    return [scenes.soup(65536, seed=1), scenes.heightfield(1024, 512, seed=2)]


def collect(ctx: N.Context, meshes, cap: int = 200_000):
    """Split samples of plain builds of `meshes` -> (device features, device
    labels, count)."""
    d_f, d_l = ctx.alloc(cap * N.MODEL_FEATURES * 4), ctx.alloc(cap * 4)
    tf, tl = ctx.alloc(cap * N.MODEL_FEATURES * 4), ctx.alloc(cap * 4)
    feats, labels = [], []
    for pos, tris in meshes:
        dp, dt = ctx.to_device(pos), ctx.to_device(tris)
        k = ctx.collect_samples(dp, 12, pos.shape[0], dt, tris.shape[0], N.default_settings(), 33,
                                tf, tl, cap)
        feats.append(tf.download(np.float32, k * N.MODEL_FEATURES))
        labels.append(tl.download(np.uint32, k))
        dp.free()
        dt.free()
    tf.free()
    tl.free()
    f = np.concatenate(feats)[: cap * N.MODEL_FEATURES]
    lab = np.concatenate(labels)[:cap]
    d_f.upload(f)
    d_l.upload(lab)
    return d_f, d_l, int(lab.shape[0])


def train_split_model(ctx: N.Context, seed: int = 1, epochs=(30, 15), batch: int = 512):
    """Trains a fresh model on the GPU; returns (model, info)."""
    t0 = time.perf_counter()
    d_f, d_l, n = collect(ctx, training_meshes())
    model = N.Model(ctx, seed=seed)
    model.train(d_f, d_l, n, epochs=epochs[0], batch=batch, lr=2e-3, seed=seed)
    loss, acc = model.train(d_f, d_l, n, epochs=epochs[1], batch=batch, lr=5e-4, seed=seed + 1)
    d_f.free()
    d_l.free()
    return model, {"samples": n, "loss": loss, "accuracy": acc,
                   "seconds": time.perf_counter() - t0}
