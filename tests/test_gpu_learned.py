"""Split proposer network on the GPU against its CPU oracle (oracle/mlp.py).

* split samples: the features and label the plain builder emits for the root
  equal the oracle's (root centroid bins restated in numpy) and the root's
  split record;
* gradients of the FP32 training kernels equal the oracle's analytic
  gradients (themselves pinned by finite differences);
* BF16 MFMA inference equals the oracle's BF16 restatement within tolerance;
* training on the GPU lowers the loss.
"""
import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def samples(ctx):
    pos, tris = scenes.soup(65536, seed=1)
    n = tris.shape[0]
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    cap = 8192
    d_f, d_l = ctx.alloc(cap * 102 * 4), ctx.alloc(cap * 4)
    cnt = ctx.collect_samples(d_pos, 12, pos.shape[0], d_tris, n, N.default_settings(), 33, d_f,
                              d_l, cap)
    yield dict(pos=pos, tris=tris, d_f=d_f, d_l=d_l, count=cnt,
               f=d_f.download(np.float32, cnt * 102).reshape(cnt, 102),
               l=d_l.download(np.uint32, cnt))
    for b in (d_pos, d_tris, d_f, d_l):
        b.free()


def test_samples_root_features_and_label(ctx, samples):
    print("[case] split samples of the 65,536-triangle soup", flush=True)
    pos, tris = samples["pos"], samples["tris"]
    ab = O.tri_aabbs(pos, tris)
    mn, mx = ab[:, :3], ab[:, 3:]
    c = ((mn + mx) * np.float32(0.5)).astype(np.float32)
    cmin, cmax = c.min(axis=0), c.max(axis=0)
    cext = (cmax - cmin).astype(np.float32)
    counts = np.zeros((3, 32))
    for a in range(3):
        b = np.floor(((c[:, a] - cmin[a]) / cext[a]) * np.float32(32.0)).astype(np.int64)
        counts[a] = np.bincount(np.clip(b, 0, 31), minlength=32)
    want = M.node_features(counts, tris.shape[0], mn.min(axis=0), mx.max(axis=0), cext)
    got = samples["f"][0]
    assert np.abs(got - want).max() < 1e-6, (got[:8], want[:8])
    t = tris.copy()
    _, _, _, splits, _ = ctx.build_tris(pos, t, N.default_settings(), want_splits=True)
    rec = splits[0]
    assert rec["axis_method"] >> 8 == 0
    # one sample per binned split (method 0)
    assert samples["count"] == int(((splits["axis_method"] >> 8) == 0).sum()) > 1000
    assert samples["l"][0] == (rec["axis_method"] & 0xFF) * 32 + rec["value"]
    assert np.all(samples["l"] % 32 >= 1) and np.all(samples["l"] < 96)
    assert np.all((samples["f"] >= 0) & (samples["f"] <= 1))


def test_samples_same_with_one_node_per_lane(ctx, samples):
    """The sweep's one-node-per-lane mode (crowded levels) splits undecided
    nodes by whole wavefronts too and writes the same samples."""
    print("[case] split samples with the sweep one node per lane on every level", flush=True)
    pos, tris = samples["pos"], samples["tris"]
    n = tris.shape[0]
    old = ctx.get_tuning()
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    cap = 8192
    d_f, d_l = ctx.alloc(cap * 102 * 4), ctx.alloc(cap * 4)
    try:
        ctx.set_tuning(sweep_pack_min_nodes=1)
        cnt = ctx.collect_samples(d_pos, 12, pos.shape[0], d_tris, n, N.default_settings(), 33,
                                  d_f, d_l, cap)
        f = d_f.download(np.float32, cnt * 102).reshape(cnt, 102)
        lab = d_l.download(np.uint32, cnt)
    finally:
        ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})
        for b in (d_pos, d_tris, d_f, d_l):
            b.free()
    assert cnt == samples["count"]

    def rows(f, lab):  # samples in a canonical order (their slots follow atomics)
        r = np.concatenate([lab[:, None], f.view(np.uint32)], axis=1)
        return r[np.lexsort(r.T[::-1])]

    assert np.array_equal(rows(f, lab), rows(samples["f"], samples["l"]))


def test_gradients_match_oracle(ctx, samples):
    print("[case] training gradients vs oracle", flush=True)
    m = N.Model(ctx, seed=7)
    n = 512
    d_g = ctx.alloc(N.MODEL_PARAMS * 4)
    loss = m.gradients(samples["d_f"], samples["d_l"], n, d_g)
    g = d_g.download(np.float32, N.MODEL_PARAMS)
    p = m.params()
    assert np.all(M.unpack(p)["W1"][:, M.IN:] == 0)
    ref_loss, ref = M.loss_and_grads(p.astype(np.float64), samples["f"][:n], samples["l"][:n])
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss)
    err = np.abs(g - ref)
    assert err.max() <= 2e-4 * np.abs(ref).max() + 1e-7, (err.max(), np.abs(ref).max(), err.argmax())
    m.close()
    d_g.free()


def test_infer_bf16_matches_oracle(ctx, samples):
    print("[case] BF16 MFMA inference vs oracle", flush=True)
    m = N.Model(ctx, seed=9)
    n = 1000  # not a multiple of the 16-row wave tile
    d_z = ctx.alloc(n * 96 * 4)
    m.infer(samples["d_f"], n, d_z)
    z = d_z.download(np.float32, n * 96).reshape(n, 96)
    ref = M.infer_bf16(m.params(), samples["f"][:n])
    scale = np.abs(ref).max()
    assert np.abs(z - ref).max() <= 2e-2 * scale, (np.abs(z - ref).max(), scale)
    assert (z.argmax(1) == ref.argmax(1)).mean() > 0.95
    m.close()
    d_z.free()


def test_training_lowers_loss(ctx, samples):
    print("[case] training on the GPU", flush=True)
    m = N.Model(ctx, seed=3)
    n = samples["count"]
    before = m.gradients(samples["d_f"], samples["d_l"], n)
    loss, acc = m.train(samples["d_f"], samples["d_l"], n, epochs=8, batch=256, lr=2e-3, seed=5)
    after = m.gradients(samples["d_f"], samples["d_l"], n)
    print(f"loss {before:.3f} -> {after:.3f}, last-epoch accuracy {acc:.3f}", flush=True)
    assert after < 0.7 * before
    assert acc > 0.1
    m.close()
