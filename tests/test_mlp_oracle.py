"""Pins the CPU oracle of the learned initialiser (no GPU): bfloat16 rounding
on known bit patterns, the backward pass against central finite differences
of the unrounded network, the feature and target formulas on hand-made
inputs, and the C-ABI constants."""
import numpy as np
import pytest

from oracle import mlp as M


def test_bf16_rounding_known_patterns():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40, 0.0, -0.0], dtype=np.float32)
    b = M.bf16_bits(v)
    # 1+2^-8 is halfway between 0x3f80 and 0x3f81 -> even (0x3f80); 1+3*2^-8 -> 0x3f82
    assert list(b[:4]) == [0x3F80, 0x3F80, 0x3F82, 0xC020]
    assert b[4] == 0x7F62          # 3e38 rounds within range
    assert b[5] == 0x0001          # subnormal keeps its top bits (1e-40 = 0x000116c2)
    assert b[6] == 0x0000 and b[7] == 0x8000
    nan = M.bf16_bits(np.array([np.nan], dtype=np.float32))[0]
    assert (nan & 0x7F80) == 0x7F80 and (nan & 0x7F) != 0
    # round trip of every bf16 value that is finite is the identity
    allb = np.arange(0, 1 << 16, dtype=np.uint32).astype(np.uint16)
    f = M.bf16_to_f32(allb)
    fin = np.isfinite(f)
    assert np.array_equal(M.bf16_bits(f[fin]), allb[fin])


def _rand_params(rng, scale=1.0):
    d = {}
    for name, shape in M.SIZES:
        fan_in = shape[1] if len(shape) == 2 else 1
        d[name] = rng.normal(0, scale / np.sqrt(fan_in), size=shape)
    return M.pack(d)


@pytest.mark.parametrize("mapped", [False, True])
def test_backward_matches_finite_differences(mapped):
    rng = np.random.default_rng(7)
    p = _rand_params(rng)
    x = rng.uniform(0, 2, size=(6, M.IN))
    t = np.concatenate([rng.normal(0, 1, size=(6, 6)), rng.dirichlet(np.ones(5), size=6),
                        np.array([[1], [0], [1], [1], [0], [1]])], axis=1)
    omap = M.IDENTITY_MAP.astype(np.float64)
    valid = None
    if mapped:
        omap = np.concatenate([rng.normal(0, 2, 6), rng.uniform(0.5, 3, 6), rng.uniform(0.1, 5, 8)])
        valid = t[:, 11].copy()          # invalid samples keep only the confidence term
    L, g = M.gradients(p, x, t, round_bf16=False, dtype=np.float64, omap=omap, valid=valid)
    assert np.isclose(L, M.loss(p, x, t, round_bf16=False, dtype=np.float64, omap=omap, valid=valid))
    idx = np.concatenate([rng.choice(np.arange(o, o + n), size=8, replace=False)
                          for o, n in _offsets()])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss(pp, x, t, False, np.float64, omap, valid) -
              M.loss(pm, x, t, False, np.float64, omap, valid)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def _offsets():
    o = 0
    for _, shape in M.SIZES:
        n = int(np.prod(shape))
        yield o, n
        o += n


def test_rounded_gradients_close_to_exact():
    """The bf16 rounding points perturb the gradient by roughly bf16 precision."""
    rng = np.random.default_rng(3)
    p = _rand_params(rng).astype(np.float32)
    x = M.bf16(rng.uniform(0, 1, size=(64, M.IN)).astype(np.float32))
    t = np.concatenate([rng.normal(0, 1, size=(64, 6)), rng.dirichlet(np.ones(5), size=64),
                        rng.integers(0, 2, size=(64, 1))], axis=1)
    _, g32 = M.gradients(p, x, t)
    _, g64 = M.gradients(p, x, t, round_bf16=False, dtype=np.float64)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 2e-2


def test_features_formula():
    c = np.zeros((2, 300), dtype=np.uint16)
    c[0, 0], c[0, 1], c[0, 10], c[0, 255], c[0, 256] = 3, 1, 4, 8, 100   # bin 256 is outside the 128 features
    x = M.features(c)
    total = 16.0
    assert x.shape == (2, 128)
    assert x[0, 0] == M.bf16(np.float32(16.0 * 4 / total)) and x[0, 5] == np.float32(16.0 * 4 / total)
    assert x[0, 127] == np.float32(16.0 * 8 / total)
    assert np.count_nonzero(x[0]) == 3
    assert not x[1].any()          # empty histogram -> zero input


def test_targets_decode_round_trip():
    rng = np.random.default_rng(11)
    Z = 50
    lo = rng.uniform(-60, -30, Z).astype(np.float32)
    bs = rng.uniform(1.5, 3.0, Z).astype(np.float32)
    diag = np.zeros((18, Z), dtype=np.float32)
    rho = rng.dirichlet(np.ones(5), size=Z).T
    diag[0:5] = rho
    diag[5] = rng.normal(0, 2, Z)
    diag[6:10] = np.array([180, 120, 80, 50])[:, None] * rng.uniform(0.5, 1.5, Z)
    diag[10] = rng.uniform(100, 200, Z)
    diag[15] = rng.uniform(0, 1, Z)                  # confidence
    diag[16] = rng.uniform(size=Z) > 0.2             # converged
    diag[17] = np.where(rng.uniform(size=Z) > 0.9, 4, 0)   # status bits
    t, ok = M.targets(lo, bs, diag, min_conf=0.3)
    assert np.array_equal(ok, (diag[16] > 0) & (diag[17] == 0) & (diag[15] >= np.float32(0.3)))
    assert np.array_equal(t[:, 11], ok.astype(np.float32))
    omap = np.concatenate([rng.normal(0, 3, 6), rng.uniform(0.5, 4, 6), np.ones(8)]).astype(np.float32)
    y = t.copy()
    y[:, :6] = (t[:, :6] - omap[:6]) / omap[6:12]   # outputs that map back onto the targets
    y[:, 6:11] = np.log(t[:, 6:11])     # logits whose softmax is the target weights
    r, mu, v0 = M.decode(y, lo, bs, omap)
    assert np.allclose(r, rho.T, rtol=1e-5)
    assert np.allclose(mu[:, 0], diag[5], atol=1e-3)
    assert np.allclose(mu[:, 1:], diag[6:10].T, rtol=1e-5)
    assert np.allclose(v0, diag[10], rtol=1e-5)


def test_param_count_matches_library(H):
    assert M.NPARAM == H.MLP_PARAMS == 101900
