"""GPU checks of the learned initialiser's kernels against the CPU oracle
(oracle/mlp.py), through the C-ABI.

* Forward pass with integer-valued weights and inputs: every product and
  partial sum is an exactly representable integer, so the result does not
  depend on the accumulation order and must equal the oracle bit for bit.
  This pins the matrix-core operand and accumulator lane layouts and the
  bias / ReLU / bf16 epilogues.
* Features from lane histograms: bit-exact.
* Gradients on realistic data: relative Frobenius error per tensor below
  2e-2 (the two sides accumulate in different orders, so an element near a
  bfloat16 rounding boundary can land one bf16 step apart), loss within 1e-4.
* Training with the GPU's own forward/backward/Adam lowers the loss above
  its floor (the entropy of the target weights) tenfold.
"""
import numpy as np
import pytest

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _upload_bf16(H, x):
    return H.dev_from_numpy(M.bf16_bits(np.asarray(x, dtype=np.float32)))


def _int_params(rng):
    d = {}
    for name, shape in M.SIZES:
        if name.startswith("W"):
            d[name] = rng.integers(-1, 2, size=shape).astype(np.float32)
        else:
            d[name] = rng.integers(-2, 3, size=shape).astype(np.float32)
    return M.pack(d).astype(np.float32)


@pytest.mark.parametrize("rows", [64, 200, 1024])
def test_forward_exact_integer_network(H, gpu, rows):
    rng = np.random.default_rng(rows)
    init = H.Initializer(seed=5)
    p = _int_params(rng)
    init.set_params(p)
    assert np.array_equal(init.params(), p)
    x = rng.integers(0, 2, size=(rows, M.IN)).astype(np.float32)
    d_x = _upload_bf16(H, x)
    d_y = H.dev_bytes(rows * M.OUT * 4)
    init.forward(d_x, rows, d_y)
    H.sync()
    y = H.to_numpy(d_y, np.float32, (rows, M.OUT))
    ref = M.forward(p, x)["y"]
    assert np.abs(ref).max() > 100          # the products are not trivially small
    assert np.array_equal(y, ref)


def test_features_match_oracle(H, T, gpu):
    rng = np.random.default_rng(2)
    lanes = 3
    hist = np.zeros(lanes, dtype=T.LANE_HIST)
    hist["bin_count"] = rng.integers(0, 60, size=(lanes, 300, 64)).astype(np.uint16)
    hist["bin_count"][1, :, 5] = 0            # an empty ZMW
    hist["bin_count"][2, 200:, :] = 0
    d_h = H.dev_from_numpy(hist)
    d_x = H.dev_bytes(lanes * 64 * M.IN * 2)
    H.mlp_features(d_h, lanes, d_x)
    H.sync()
    got = H.to_numpy(d_x, np.uint16, (lanes * 64, M.IN))
    counts = hist["bin_count"].transpose(0, 2, 1).reshape(lanes * 64, 300)
    want = M.bf16_bits(M.features(counts))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_from_histograms_equals_features_then_forward(H, T, gpu, seed):
    # The inference kernel computes the inputs from the lane histograms itself; its
    # outputs equal features + forward bit for bit (random, not integer, weights).
    rng = np.random.default_rng(10 + seed)
    lanes = 5
    hist = np.zeros(lanes, dtype=T.LANE_HIST)
    hist["bin_count"] = rng.integers(0, 90, size=(lanes, 300, 64)).astype(np.uint16)
    hist["bin_count"][0, :, 7] = 0           # an empty ZMW
    hist["bin_count"][3, :, :] //= 7
    init = H.Initializer(seed=3 + seed)
    d_h = H.dev_from_numpy(hist)
    d_x = H.dev_bytes(lanes * 64 * M.IN * 2)
    d_y1, d_y2 = H.dev_bytes(lanes * 64 * M.OUT * 4), H.dev_bytes(lanes * 64 * M.OUT * 4)
    H.mlp_features(d_h, lanes, d_x)
    init.forward(d_x, lanes * 64, d_y1)
    init.forward_hist(d_h, lanes, d_y2)
    H.sync()
    y1 = H.to_numpy(d_y1, np.float32, (lanes * 64, M.OUT))
    y2 = H.to_numpy(d_y2, np.float32, (lanes * 64, M.OUT))
    assert np.abs(y1).max() > 0
    assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))


def _realistic_batch(rng, B):
    # histogram-like non-negative inputs and plausible targets
    x = np.zeros((B, M.IN), dtype=np.float32)
    centers = rng.uniform(5, 15, B)
    for i in range(B):
        k = np.arange(M.IN)
        x[i] = 12 * np.exp(-0.5 * ((k - centers[i]) / 2.5) ** 2) / (2.5 * np.sqrt(2 * np.pi))
        x[i] += rng.uniform(0, 0.3, M.IN) * (k < 60)
    x = M.bf16(x)
    t = np.concatenate([rng.normal(0.5, 0.3, size=(B, 6)), rng.dirichlet(np.ones(5) * 2, size=B),
                        (rng.uniform(size=(B, 1)) > 0.1)], axis=1).astype(np.float32)
    return x, t


@pytest.mark.parametrize("batch", [256, 200])
def test_gradients_match_oracle(H, gpu, batch):
    rng = np.random.default_rng(batch)
    init = H.Initializer(seed=9)
    p = init.params()
    omap = np.concatenate([rng.normal(0, 1, 6), rng.uniform(0.5, 2, 6), rng.uniform(0.2, 3, 8)]).astype(np.float32)
    init.set_output_map(omap[:6], omap[6:12], omap[12:])
    assert np.array_equal(init.output_map(), omap)
    x, t = _realistic_batch(rng, batch)
    valid = t[:, 11].astype(np.uint8)
    d_x, d_t, d_v = _upload_bf16(H, x), H.dev_from_numpy(t), H.dev_from_numpy(valid)
    d_g = H.dev_bytes(M.NPARAM * 4)
    d_l = H.dev_bytes(4)
    init.gradients(d_x, d_t, batch, d_g, d_l, d_valid=d_v)
    H.sync()
    g = H.to_numpy(d_g, np.float32, (M.NPARAM,))
    loss = H.to_numpy(d_l, np.float32, (1,))[0]
    # oracle: invalid samples contribute only the confidence term, the mean stays over the batch
    L_ref, g_ref = M.gradients(p, x, t, omap=omap, valid=valid)
    assert abs(loss - L_ref) <= 1e-4 * abs(L_ref)
    o = 0
    for name, shape in M.SIZES:
        n = int(np.prod(shape))
        a, b = g[o:o + n], g_ref[o:o + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err < 2e-2, (name, err)
        o += n


def test_training_lowers_loss(H, gpu):
    """Adam on the GPU fits a teacher network's outputs."""
    rng = np.random.default_rng(4)
    B = 512
    x, _ = _realistic_batch(rng, 4 * B)
    teacher = M.forward(H.Initializer(seed=77).params(), x)["y"]
    t = np.concatenate([teacher[:, :6] * 0.1,
                        np.exp(teacher[:, 6:11]) / np.exp(teacher[:, 6:11]).sum(1, keepdims=True),
                        (teacher[:, 11:12] > 0)], axis=1).astype(np.float32)
    init = H.Initializer(seed=3)
    d_x, d_t = _upload_bf16(H, x), H.dev_from_numpy(t)
    d_l = H.dev_bytes(4)
    losses = []
    for step in range(300):
        off = (step % 4) * B
        init.train_step(d_x.view(off * M.IN * 2, B * M.IN * 2), d_t.view(off * M.OUT * 4, B * M.OUT * 4), B, 1e-3,
                        d_loss=d_l)
        if step % 50 == 0 or step == 299:
            H.sync()
            losses.append(float(H.to_numpy(d_l, np.float32, (1,))[0]))
    assert np.all(np.isfinite(losses))
    # the cross-entropy term cannot go below the entropy of the target weights
    floor = float(np.mean(-(t[:, 6:11] * np.log(t[:, 6:11])).sum(axis=1)))
    excess = [v - floor for v in losses]
    assert excess[-1] < 0.1 * excess[0], (floor, losses)


def test_targets_match_oracle(H, T, gpu):
    """hwpa_mlp_targets against oracle/mlp.py targets on hand-made EM results:
    the raw-unit means and weights bit for bit, log(var / u^2) within 3 ulp (the
    library's log against numpy's), and the validity flag / confidence target
    exactly (converged, status OK, confidence >= min_confidence)."""
    rng = np.random.default_rng(12)
    L = 3
    Z = L * 64
    hist = np.zeros(L, dtype=T.LANE_HIST)
    hist["low_bound"] = rng.uniform(-60, -30, (L, 64)).astype(np.float32)
    hist["bin_size"] = rng.uniform(1.5, 3.0, (L, 64)).astype(np.float32)
    diag = np.zeros((L, H.EM_DIAG_ROWS, 64), dtype=np.float32)
    diag[:, 0:5] = rng.dirichlet(np.ones(5), size=(L, 64)).transpose(0, 2, 1)
    diag[:, 5] = rng.normal(0, 2, (L, 64))
    diag[:, 6:10] = np.array([180, 120, 80, 50])[None, :, None] * rng.uniform(0.5, 1.5, (L, 4, 64))
    diag[:, 10] = rng.uniform(100, 200, (L, 64))
    diag[:, 15] = rng.uniform(0, 1, (L, 64))
    diag[:, 16] = rng.uniform(size=(L, 64)) > 0.2
    diag[:, 17] = np.where(rng.uniform(size=(L, 64)) > 0.9, 4, 0)
    d_t, d_v = H.dev_bytes(Z * H.MLP_OUT * 4), H.dev_bytes(Z)
    H.mlp_targets(H.dev_from_numpy(hist), H.dev_from_numpy(diag), L, 0.3, d_t, d_v)
    g_t = H.to_numpy(d_t, np.float32, (Z, H.MLP_OUT))
    g_v = H.to_numpy(d_v, np.uint8, (Z,))
    flat = diag.transpose(1, 0, 2).reshape(H.EM_DIAG_ROWS, Z)
    o_t, o_ok = M.targets(hist["low_bound"].reshape(Z), hist["bin_size"].reshape(Z), flat, min_conf=0.3)
    assert np.array_equal(g_v.astype(bool), o_ok)
    cols = [0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]
    assert np.array_equal(g_t[:, cols], o_t[:, cols])
    spacing = np.spacing(np.abs(o_t[:, 1]))
    assert np.all(np.abs(g_t[:, 1] - o_t[:, 1]) <= 3 * spacing)
