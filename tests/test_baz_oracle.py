"""Pins of the oracle's BAZ packet encoding (no GPU) against the reference's
own tests: bazio/encoding/test/TestPulseToBaz.cu (encoded lengths 16 / 32 /
24 / 86 and decoded values of the three Kestrel layouts and InternalPulses,
group parameters), TestFieldSerializers.cpp (truncate / simple / compact
round trips and overflow byte counts), TestFieldTransforms.cpp
(FloatFixedCodec), TestBazRoundTrip.cpp; plus decode(encode(x)) == x over
random chunked pulse streams with the delta state carried across chunks."""
import math

import numpy as np
import pytest

NAN, INF = float("nan"), float("inf")
WIDTH = [4, 120, 55, 32, 182, 5, 1, 300]
GAPS = [130, 5, 18, 22, 500, 64, 33]
LABEL = [0, 1, 2, 3, 0, 1, 2, 3]


def _starts():
    st = [14]
    for i in range(7):
        st.append(st[-1] + WIDTH[i] + GAPS[i])
    return st


def _pulses():
    return dict(label=LABEL, start=_starts(), width=WIDTH)


@pytest.mark.parametrize("layout,length", [(2, 16), (3, 32), (4, 24), (0, 24)])
def test_kestrel_layout_lengths_and_round_trip(O, layout, length):
    data, last = O.baz_encode_fields(layout, _pulses())
    assert len(data) == length
    assert last == _starts()[-1]
    out, used, _ = O.baz_decode_fields(layout, data, 8)
    assert used == length
    assert out["label"].tolist() == LABEL
    if layout == 2:   # lossy truncate: width and start delta modulo 128
        assert out["width"].tolist() == [w % 128 for w in WIDTH]
        d_in = np.diff([0] + _starts())
        d_out = np.diff([0] + out["start"].astype(np.int64).tolist())
        assert (d_in % 128).tolist() == d_out.tolist()
    else:
        assert out["width"].tolist() == WIDTH and out["start"].tolist() == _starts()


def test_internal_mode_length_and_values(O):
    # TestPulseToBaz InternalMode: 86 bytes; floats within 1/scale below the 2-byte range,
    # non-finite beyond it, NaN kept (MidSignal and SignalM2 are NaN below 3 frames).
    mx = [0, 423.2, 223.645, 623.24, NAN, 2123.12, 3613.36, INF]
    mean = [423.2, 223.645, 623.24, NAN, 2123.12, 3613.36, INF, 0]
    mid = [223.645, 623.24, NAN, 2123.12, 3613.36, INF, 0, 423.2]
    m2 = [623.24, NAN, 4123.12, 6613.36, INF, 0, 423.2, 223.645]
    mid = [m if w >= 3 else NAN for m, w in zip(mid, WIDTH)]
    m2 = [m if w >= 3 else NAN for m, w in zip(m2, WIDTH)]
    f = dict(_pulses(), pkmax=mx, pkmean=mean, pkmid=mid, pkvar=m2, isbase=[1] * 8)
    data, _ = O.baz_encode_fields(1, f)
    assert len(data) == 86
    out, used, _ = O.baz_decode_fields(1, data, 8)
    assert used == 86
    assert out["start"].tolist() == _starts() and out["width"].tolist() == WIDTH
    for name, truth, mx_val in (("pkmax", mx, 3276.7), ("pkmean", mean, 3276.7), ("pkmid", mid, 3276.7),
                                ("pkvar", m2, 6553.5)):
        for t, c in zip(truth, out[name]):
            if math.isnan(t):
                assert math.isnan(c), name
            elif t < mx_val:
                assert abs(t - c) <= 0.1 + 1e-6, name
            else:
                assert not math.isfinite(c), name


def test_layout_groups(O):
    # TestPulseToBaz Params: production one 16-bit group of 3 fields; internal 5 groups,
    # the last holding Pkvar (7) and IsBase (1); nominal bytes 2 and 6 (PrelimHQFilter.cpp:67-68).
    bits, nf = O.baz_layout_groups(0)
    assert bits == [16] and nf == [3]
    bits, nf = O.baz_layout_groups(1)
    assert bits == [16, 8, 8, 8, 8] and nf == [3, 1, 1, 1, 2]
    assert sum((b + 7) // 8 for b in bits) == 6


def _s(v):
    return v & (2 ** 64 - 1)


def _signed(v):
    return v - 2 ** 64 if v >= 2 ** 63 else v


@pytest.mark.parametrize("v,bits,signed,exp", [
    (0, 1, False, 0), (1, 1, False, 1), (2, 1, False, 0), (3, 1, False, 1), (3, 3, False, 3), (5, 3, False, 5),
    (9, 3, False, 1), (1 << 33, 32, False, 0), (1 << 33, 34, False, 1 << 33),
    (2, 2, True, -2), (3, 2, True, -1), (9, 4, True, -7), (1 << 33, 33, True, 0), (1 << 33, 35, True, 1 << 33),
    (-1, 2, True, -1), (-3, 2, True, 1), (-9, 4, True, 7), (-(1 << 33), 33, True, 0), (-(1 << 33), 35, True, -(1 << 33))])
def test_truncate_round_trip(O, v, bits, signed, exp):
    r, ov = O.baz_serial_roundtrip(O.BAZ_TRUNC, bits, 0, signed, _s(v))
    assert (_signed(r) if signed else r) == exp and ov == 0


@pytest.mark.parametrize("v,bits,ob,signed", [
    (0, 1, 1, False), (2, 1, 1, False), (3, 1, 1, False), (9, 3, 1, False), (1 << 33, 32, 8, False),
    (2, 2, 1, True), (3, 2, 1, True), (9, 4, 1, True), (1 << 33, 33, 8, True), (-1, 2, 1, True), (-1, 3, 1, True),
    (-3, 2, 1, True), (-9, 4, 1, True), (-(1 << 33), 33, 8, True)])
def test_simple_overflow_round_trip(O, v, bits, ob, signed):
    r, _ = O.baz_serial_roundtrip(O.BAZ_SIMPLE, bits, ob, signed, _s(v))
    assert (_signed(r) if signed else r) == v


def test_simple_overflow_bytes_and_truncation(O):
    assert O.baz_serial_roundtrip(O.BAZ_SIMPLE, 4, 1, True, 1 << 2)[1] == 0
    assert O.baz_serial_roundtrip(O.BAZ_SIMPLE, 4, 1, True, 1 << 5)[1] == 1
    assert O.baz_serial_roundtrip(O.BAZ_SIMPLE, 4, 4, True, 1 << 5)[1] == 4
    assert O.baz_serial_roundtrip(O.BAZ_SIMPLE, 7, 4, True, 1 << 5)[1] == 0
    assert O.baz_serial_roundtrip(O.BAZ_SIMPLE, 7, 4, True, 1 << 9)[1] == 4
    # not enough overflow bytes: truncated
    assert _signed(O.baz_serial_roundtrip(O.BAZ_SIMPLE, 30, 4, True, _s(-(1 << 33)))[0]) == 0


@pytest.mark.parametrize("v,bits,signed", [
    (0, 2, False), (3, 2, False), (9, 4, False), (1 << 33, 33, False), (1 << 33, 35, False),
    (0, 3, True), (3, 3, True), (9, 5, True), (1 << 33, 34, True), (-1, 3, True), (-3, 3, True), (-9, 5, True),
    (-(1 << 33), 34, True), (-(1 << 33), 36, True), (2 ** 40 + 12345, 7, False), (2 ** 63 + 5, 7, False)])
def test_compact_overflow_round_trip(O, v, bits, signed):
    r, _ = O.baz_serial_roundtrip(O.BAZ_COMPACT, bits, 0, signed, _s(v))
    assert (_signed(r) if signed else r) == v


def test_compact_overflow_bytes(O):
    assert O.baz_serial_roundtrip(O.BAZ_COMPACT, 4, 0, False, 1 << 2)[1] == 0
    assert O.baz_serial_roundtrip(O.BAZ_COMPACT, 4, 0, True, 1 << 2)[1] == 1
    assert O.baz_serial_roundtrip(O.BAZ_COMPACT, 4, 0, False, 1 << 5)[1] == 1
    assert O.baz_serial_roundtrip(O.BAZ_COMPACT, 7, 0, False, 1 << 5)[1] == 0
    assert O.baz_serial_roundtrip(O.BAZ_COMPACT, 7, 0, True, 1 << 9)[1] == 1


def test_float_fixed_codec(O):
    # TestFieldTransforms FloatFixedCodec round trips
    for f, signed, scale, exp in ((3.125412, False, 10, 3.1), (3.125412, False, 100, 3.13), (3.125412, False, 1000, 3.125),
                                  (3.125412, False, 30, 3.13333333), (-3.125412, True, 10, -3.1),
                                  (-3.125412, True, 100, -3.13), (-3.125412, True, 30, -3.13333333)):
        _, back = O.baz_floatfixed(f, signed, scale, 2)
        assert back == pytest.approx(exp, rel=1e-6)
    for signed in (False, True):
        for nb in (1, 2, 4, 8):
            assert O.baz_floatfixed(NAN, signed, 10, nb)[0] == 0 and math.isnan(O.baz_floatfixed(NAN, signed, 10, nb)[1])
            assert O.baz_floatfixed(INF, signed, 10, nb) == (1, INF)
            assert O.baz_floatfixed(-INF, signed, 10, nb) == (2, -INF)
    # values out of the byte range become infinities, finite ones keep their offset of 3
    # (signed 2 bytes: |scaled| <= 2^15 - 1 - 3)
    assert O.baz_floatfixed(3276.4, True, 10, 2)[0] == 32767
    assert O.baz_floatfixed(3276.5, True, 10, 2)[0] == 1
    assert O.baz_floatfixed(-3276.4, True, 10, 2)[0] == 2 ** 64 - 32764
    assert O.baz_floatfixed(0.0, False, 10, 2)[0] == 3


def random_chunks(T, zmws, max_calls, chunks, seed):
    """Pulse chunks [zmws * max_calls] with counts, in time order per ZMW, with rejects,
    wide pulses, long gaps and the occasional count above the slot number."""
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 100, zmws).astype(np.int64)
    out = []
    for c in range(chunks):
        p = np.zeros((zmws, max_calls), dtype=T.PULSE)
        counts = rng.integers(0, max_calls + 4, zmws).astype(np.uint32)
        for z in range(zmws):
            for j in range(min(int(counts[z]), max_calls)):
                t[z] += int(rng.choice([0, 1, 20, 200, 70000], p=[0.3, 0.3, 0.3, 0.09, 0.01]))
                w = int(rng.choice([1, 2, 3, 8, 40, 200, 5000], p=[0.1, 0.1, 0.2, 0.4, 0.1, 0.08, 0.02]))
                q = p[z, j]
                q["start"], q["width"] = t[z], w
                q["label"] = rng.integers(0, 4)
                q["mean_signal"], q["mid_signal"] = rng.integers(0, 65535, 2)
                q["max_signal"] = rng.integers(0, 65535)
                q["signal_m2"] = rng.choice([rng.uniform(0, 7000), np.float32(np.inf), 0.0])
                q["is_reject"] = rng.random() < 0.15
                t[z] += w
        out.append((p.reshape(-1), counts))
    return out


@pytest.mark.parametrize("internal", [0, 1])
def test_batch_stream_decodes_to_the_pulses(O, T, internal):
    zmws, mc = 40, 12
    last = np.zeros(zmws, dtype=np.uint64)
    dec_last = np.zeros(zmws, dtype=np.uint64)
    for c, (p, counts) in enumerate(random_chunks(T, zmws, mc, 4, seed=3 + internal)):
        data, offsets, events = O.baz_encode_batch(internal, p, counts, mc, last)
        assert len(data) == int(offsets[-1])
        pz = p.reshape(zmws, mc)
        for z in range(zmws):
            sel = [q for q in pz[z, :min(int(counts[z]), mc)] if internal or not q["is_reject"]]
            assert events[z] == len(sel)
            seg = data[int(offsets[z]):int(offsets[z + 1])]
            out, used, dec_last[z] = O.baz_decode_fields(internal, seg, len(sel), int(dec_last[z]))
            assert used == len(seg)
            assert out["start"].tolist() == [int(q["start"]) for q in sel]
            assert out["width"].tolist() == [int(q["width"]) for q in sel]
            assert out["label"].tolist() == [int(q["label"]) for q in sel]
            if internal:
                assert out["isbase"].tolist() == [0 if q["is_reject"] else 1 for q in sel]
                for q, mxv in zip(sel, out["pkmax"]):
                    truth = np.float32(q["max_signal"]) * np.float32(0.1)
                    assert (abs(mxv - truth) <= 0.05 + 1e-3) if truth < 3276.7 else not np.isfinite(mxv)
        assert np.array_equal(dec_last, last)


def test_batch_too_small_buffer_writes_nothing(O, T):
    zmws, mc = 16, 8
    (p, counts), (p2, c2) = random_chunks(T, zmws, mc, 2, seed=9)
    last = np.zeros(zmws, dtype=np.uint64)
    data, offsets, _ = O.baz_encode_batch(0, p, counts, mc, last)
    keep = last.copy()
    small, offs2, _ = O.baz_encode_batch(0, p2, c2, mc, last, cap=3)
    assert small is None and np.array_equal(last, keep) and offs2[-1] > 3
    full, offs3, _ = O.baz_encode_batch(0, p2, c2, mc, last)
    assert np.array_equal(offs2, offs3) and len(full) == int(offs3[-1])
