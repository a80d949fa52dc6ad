"""GPU block repacker (hwpa_repack) against the reference's test pattern and
the oracle, int16 and uint8, several packet shapes, batches partially covered
by packets (untouched samples keep their previous contents)."""
import numpy as np
import pytest

from test_repack_oracle import expected_batch, packets_chunk

pytestmark = pytest.mark.gpu


def _upload(H, pk):
    return [(H.dev_from_numpy(d), z0, f0, d.shape[0], d.shape[1], d.shape[2]) for d, z0, f0 in pk]


@pytest.mark.parametrize("dtype,block_width", [(np.int16, 32), (np.int16, 64), (np.int16, 128), (np.uint8, 64),
                                               (np.uint8, 128)])
@pytest.mark.parametrize("packet_frames", [64, 256])
def test_repack_pattern(H, T, gpu, dtype, block_width, packet_frames):
    enc = T.ENCODING_UINT8 if dtype == np.uint8 else T.ENCODING_INT16
    num_zmw, num_frames, lanes = 16384, 256, 16
    r = H.Repacker()
    es = np.dtype(dtype).itemsize
    d_batch = H.dev_bytes(lanes * num_frames * 64 * es)
    for chunk in range(3):
        pk = _upload(H, packets_chunk(dtype, chunk * num_frames, num_frames, num_zmw, packet_frames, 50, block_width))
        for b0 in (0, 5 * lanes * 64, num_zmw - lanes * 64):
            r.run(pk, enc, d_batch, b0, chunk * num_frames, lanes, num_frames)
            got = H.to_numpy(d_batch, dtype, (lanes, num_frames, 64))
            assert np.array_equal(got, expected_batch(dtype, b0, chunk * num_frames, lanes, num_frames)), (chunk, b0)


@pytest.mark.parametrize("dtype", [np.int16, np.uint8])
def test_repack_partial_overlap_random(H, O, T, gpu, dtype):
    rng = np.random.default_rng(3)
    enc = T.ENCODING_UINT8 if dtype == np.uint8 else T.ENCODING_INT16
    es = np.dtype(dtype).itemsize
    lanes, frames = 8, 128
    bw = 64 if dtype == np.uint8 else 32
    # packets covering only part of the batch, some straddling its edges in ZMW and frame
    host = []
    for z0, f0, nb, pf in ((0, 0, 3, 64), (bw * 3, 32, 4, 64), (lanes * 64 - 2 * bw, 96, 5, 64), (128, 1000, 2, 32)):
        info = np.iinfo(dtype)
        host.append((rng.integers(info.min, info.max, size=(nb, pf, bw)).astype(dtype), z0, f0))
    sentinel = np.full((lanes, frames, 64), 0x55, dtype=dtype)
    d_batch = H.dev_from_numpy(sentinel)
    r = H.Repacker()
    first_zmw, first_frame = 64, 16
    r.run(_upload(H, host), enc, d_batch, first_zmw, first_frame, lanes, frames)
    got = H.to_numpy(d_batch, dtype, (lanes, frames, 64))
    exp = O.repack(host, first_zmw, first_frame, lanes, frames, dtype)
    cover = O.repack([(np.ones_like(d), z0, f0) for d, z0, f0 in host], first_zmw, first_frame, lanes, frames, dtype) != 0
    assert cover.any() and (~cover).any()
    assert np.array_equal(got[cover], exp[cover])
    assert np.all(got[~cover] == sentinel[~cover])
