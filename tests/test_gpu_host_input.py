"""Chunks handed over in host memory (hwpa_analyzer_process with on_device = 0).

The analyzer copies a host chunk on its own stream into one of two staging
buffers in turn, so the copy of one chunk overlaps the previous chunk's stages,
and the call returns only when its copy is done (BatchAnalyzer::operator() takes
its TraceBatch by reference and does not keep it: analyzer/BatchAnalyzer.cpp:149-170).
Checked: an analyzer fed pinned host chunks back to back, with each host buffer
overwritten as soon as its call returns, ends every chunk with the same outputs
and state as one fed the same chunks from device memory; the same for pageable
(numpy) chunks; and results read after every chunk agree chunk by chunk (what a chunk
produced: labels and models only once the labeler runs, pulse records up to each
ZMW's count)."""
import numpy as np
import pytest

from conftest import gen_traces

pytestmark = pytest.mark.gpu


def _snapshot(a):
    """What the last chunk produced: the baselined trace and pulse counts always; labels and
    models when the labeler ran (flags[2]); each ZMW's pulse records up to its count."""
    p, cnt, ovf = a.pulses()
    flags = tuple(a.flags()[:3])
    snap = {"flags": flags, "baselined": a.baselined(), "counts": cnt}
    if flags[2]:
        snap["labels"] = a.labels()
        snap["models"] = np.ascontiguousarray(a.models()).view(np.uint8)
    recs = np.ascontiguousarray(p).view(np.uint8).reshape(len(cnt), -1)
    per = recs.shape[1] // max(p.shape[1], 1)
    snap["pulses"] = np.concatenate([recs[z, :per * min(int(cnt[z]), p.shape[1])] for z in range(len(cnt))])
    return snap


def _same(x, y, what):
    assert x.keys() == y.keys(), what
    for k in x:
        if isinstance(x[k], tuple):
            assert x[k] == y[k], "%s %s" % (what, k)
        else:
            assert np.array_equal(x[k], y[k]), "%s %s" % (what, k)


def test_host_chunks_match_device_chunks(H, O, T, gpu):
    L, F, C = 6, 512, 9
    cfg = H.bench_config(F)
    chunks = gen_traces(O, T, H, L, F, C, seed=31)
    dev = [H.dev_from_numpy(c) for c in chunks]
    pinned = [H.host_pinned(c.nbytes) for c in chunks]

    ref = H.BatchAnalyzer(0, L, F, cfg)
    pin = H.BatchAnalyzer(0, L, F, cfg)
    page = H.BatchAnalyzer(0, L, F, cfg)
    step = H.BatchAnalyzer(0, L, F, cfg)
    for c in range(C):
        ref(dev[c], c * F)
        want = _snapshot(ref)
        # results read after every chunk
        step(chunks[c], c * F)
        _same(_snapshot(step), want, "per-chunk host input, chunk %d" % c)
    final = _snapshot(ref)

    # back to back, no reads between chunks; each buffer is clobbered when its call returns
    for c in range(C):
        pinned[c].array(np.int16)[:] = chunks[c].reshape(-1)
        pin(pinned[c], c * F, on_device=False)
        pinned[c].array(np.int16)[:] = -12345
        buf = chunks[c].copy()
        page(buf, c * F)
        buf.fill(-12345)
    _same(_snapshot(pin), final, "pinned host input")
    _same(_snapshot(page), final, "pageable host input")


def test_host_uint8_chunks_match_device_chunks(H, O, T, gpu):
    """uint8 chunks (half the bytes per sample of the staging buffers) from pinned and
    pageable host memory, back to back, against the same chunks from device memory."""
    L, F, C = 4, 512, 7
    cfg = H.bench_config(F)
    cfg.movie.pedestal = 64
    chunks = [np.clip(c.astype(np.int32) // 2 + 64, 0, 255).astype(np.uint8)
              for c in gen_traces(O, T, H, L, F, C, seed=37)]
    dev = [H.dev_from_numpy(c) for c in chunks]
    ref = H.BatchAnalyzer(0, L, F, cfg)
    pin = H.BatchAnalyzer(0, L, F, cfg)
    page = H.BatchAnalyzer(0, L, F, cfg)
    pinned = [H.host_pinned(c.nbytes) for c in chunks]
    for c in range(C):
        ref(dev[c], c * F, encoding=T.ENCODING_UINT8)
        pinned[c].array(np.uint8)[:] = chunks[c].reshape(-1)
        pin(pinned[c], c * F, on_device=False, encoding=T.ENCODING_UINT8)
        pinned[c].array(np.uint8)[:] = 7
        page(chunks[c].copy(), c * F)
    final = _snapshot(ref)
    _same(_snapshot(pin), final, "pinned uint8 host input")
    _same(_snapshot(page), final, "pageable uint8 host input")
    assert final["flags"][2]
