"""GPU BAZ packet encoding (hwpa_baz_*) against the oracle's restatement of
ObjectToBaz: the whole stream, offsets and event counts byte for byte over
chunks with the delta state carried across them, in both layouts, on random
pulse records and on the analyzer's own pulses; a too-small buffer writes
nothing and keeps the state."""
import numpy as np
import pytest

from conftest import gen_traces
from test_baz_oracle import random_chunks

pytestmark = pytest.mark.gpu


def _run(H, enc, p, counts, mc, cap):
    zmws = len(counts)
    d_p, d_c = H.dev_from_numpy(p), H.dev_from_numpy(counts)
    d_out = H.dev_bytes(max(cap, 1))
    d_off, d_ev = H.dev_bytes((zmws + 1) * 8), H.dev_bytes(zmws * 4)
    enc.encode(d_p, d_c, mc, d_out, cap, d_off, d_ev)
    off = H.to_numpy(d_off, np.uint64, (zmws + 1,))
    ev = H.to_numpy(d_ev, np.uint32, (zmws,))
    total = int(off[-1])
    data = H.to_numpy(d_out, np.uint8, (max(cap, 1),))[:total].tobytes() if total <= cap else None
    return data, off, ev


@pytest.mark.parametrize("internal,zmws,mc", [(0, 320, 16), (1, 320, 16), (0, 128, 150), (1, 64, 100)])
def test_baz_random_chunks_bit_exact(H, O, T, gpu, internal, zmws, mc):
    enc = H.BazEncoder(zmws, internal=bool(internal))
    last = np.zeros(zmws, dtype=np.uint64)
    for c, (p, counts) in enumerate(random_chunks(T, zmws, mc, 3, seed=21 + internal)):
        o_data, o_off, o_ev = O.baz_encode_batch(internal, p, counts, mc, last)
        g_data, g_off, g_ev = _run(H, enc, p, counts, mc, len(o_data) + 64)
        assert np.array_equal(g_off, o_off), c
        assert np.array_equal(g_ev, o_ev), c
        assert g_data == o_data, c


def test_baz_small_buffer_keeps_state(H, O, T, gpu):
    zmws, mc = 192, 12
    enc = H.BazEncoder(zmws)
    last = np.zeros(zmws, dtype=np.uint64)
    chunks = random_chunks(T, zmws, mc, 2, seed=5)
    o1, _, _ = O.baz_encode_batch(0, chunks[0][0], chunks[0][1], mc, last)
    g1, _, _ = _run(H, enc, chunks[0][0], chunks[0][1], mc, len(o1))
    assert g1 == o1
    # size query: nothing written, state kept, then the real call
    none, q_off, _ = _run(H, enc, chunks[1][0], chunks[1][1], mc, 0)
    assert none is None
    o2, o_off, _ = O.baz_encode_batch(0, chunks[1][0], chunks[1][1], mc, last)
    assert np.array_equal(q_off, o_off)
    g2, _, _ = _run(H, enc, chunks[1][0], chunks[1][1], mc, int(q_off[-1]))
    assert g2 == o2


def test_baz_analyzer_pulses(H, O, T, gpu):
    """Encode the GPU analyzer's pulses chunk after chunk (production layout) and compare
    with the oracle encoding of the same pulse records."""
    L, F, C = 4, 512, 6
    cfg = H.bench_config(F)
    ga = H.BatchAnalyzer(0, L, F, cfg)
    enc = H.BazEncoder(L * 64)
    last = np.zeros(L * 64, dtype=np.uint64)
    produced = 0
    for c, tr in enumerate(gen_traces(O, T, H, L, F, C, seed=17)):
        ga(tr, c * F)
        if not ga.flags()[0]:
            continue
        p, n, _ = ga.pulses()
        mc = p.shape[1]
        o_data, o_off, o_ev = O.baz_encode_batch(0, p.reshape(-1), n, mc, last)
        g_data, g_off, g_ev = _run(H, enc, p.reshape(-1), n, mc, len(o_data))
        assert g_data == o_data and np.array_equal(g_off, o_off) and np.array_equal(g_ev, o_ev), c
        produced += int(o_ev.sum())
    assert produced > 1000
