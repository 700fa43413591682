"""Pins for the CPU oracle of the split proposer network (oracle/mlp.py)."""
import numpy as np

from oracle import mlp as M


def _rand_params(seed=0):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(M.NPARAM) * 0.1).astype(np.float32)
    w = M.unpack(p)
    w["W1"][:, M.IN:] = 0.0
    return M.pack(w)


def test_layout_size():
    assert M.NPARAM == 45408
    p = np.arange(M.NPARAM, dtype=np.float32)
    assert np.array_equal(M.pack(M.unpack(p)), p)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = _rand_params(1).astype(np.float64)
    X = rng.random((12, M.IN))
    y = rng.integers(0, M.OUT, 12)
    loss, g = M.loss_and_grads(p, X, y, np.float64)
    assert abs(loss - M.loss_only(p, X, y)) < 1e-12
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False),
                          [0, 16384, 16511, 16512, 33023, 33024, 45311, 45407]])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = M.loss_only(q, X, y)
        q[i] -= 2 * eps
        lm = M.loss_only(q, X, y)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_softmax_ce_known_value():
    # zero weights: uniform softmax, loss = log 96 and the b3 gradient is (1/96 - onehot) / B
    p = np.zeros(M.NPARAM)
    X = np.zeros((2, M.IN))
    y = np.array([3, 95])
    loss, g = M.loss_and_grads(p, X, y)
    assert abs(loss - np.log(96.0)) < 1e-12
    gb3 = M.unpack(g)["b3"]
    want = np.full(96, 1 / 96.0)
    want[3] -= 0.5
    want[95] -= 0.5
    assert np.allclose(gb3, want, atol=1e-15)


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 0.0,
                     1.0 + 2 ** -8 + 2 ** -20, 3.0e38], np.float32)
    got = M.bf16(vals).view(np.uint32) >> 16
    # 1 + 2^-8 is a tie between 0x3F80 and 0x3F81: even 0x3F80; 1 + 3*2^-8 ties to 0x3F82
    assert list(got) == [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x0000, 0x3F81, 0x7F62]


def test_features_known_answer():
    counts = np.zeros((3, 32))
    counts[0, 0] = 40
    counts[1, 5] = 40
    counts[2, 31] = 40
    f = M.node_features(counts, 40, [0, 0, 0], [2, 1, 0.5], [1, 0.5, 0])
    assert f[0] == 1.0 and f[32 + 5] == 1.0 and f[64 + 31] == 1.0 and f[1] == 0.0
    assert np.allclose(f[96:99], [1.0, 0.5, 0.25]) and np.allclose(f[99:102], [0.5, 0.5, 0.0])


def test_infer_bf16_close_to_fp32():
    rng = np.random.default_rng(2)
    p = _rand_params(2)
    X = rng.random((64, M.IN)).astype(np.float32)
    a = M.forward(p, X)[3]
    b = M.infer_bf16(p, X)
    assert np.abs(a - b).max() < 0.05 * np.abs(a).max()
