"""The oracle's block repacker against the reference's own test pattern
(appModules/test/TestBlockRepacker.cpp:63-90, 359-460): every sample holds its
(frame mod 2^(b/2)) in the high half and (zmw mod 2^(b/2)) in the low half, so
a correctly repacked batch is known in closed form (no GPU)."""
import numpy as np
import pytest


def pattern(dtype, zmws, frames):
    """Sample value of the reference's pattern at ZMW indices `zmws` and frames `frames`."""
    bits = np.dtype(dtype).itemsize * 8 // 2
    mask = (1 << bits) - 1
    f = (np.asarray(frames)[:, None] & mask) << bits
    z = np.asarray(zmws)[None, :] & mask
    return (f | z).astype(dtype)


def packets_chunk(dtype, start_frame, num_frames, num_zmw, packet_frames, packet_blocks, block_width):
    """GeneratePacketsChunk: packets of packet_blocks blocks (fewer for the last ZMWs) over
    the chip, one per packet_frames frames."""
    out = []
    for z0 in range(0, num_zmw, packet_blocks * block_width):
        nb = min(packet_blocks, (num_zmw - z0) // block_width)
        for f0 in range(start_frame, start_frame + num_frames, packet_frames):
            zm = np.arange(z0, z0 + nb * block_width)
            pat = pattern(dtype, zm, np.arange(f0, f0 + packet_frames))          # [frame][zmw]
            data = np.ascontiguousarray(pat.reshape(packet_frames, nb, block_width).transpose(1, 0, 2))
            out.append((data, z0, f0))
    return out


def expected_batch(dtype, first_zmw, first_frame, lanes, frames):
    pat = pattern(dtype, np.arange(first_zmw, first_zmw + lanes * 64), np.arange(first_frame, first_frame + frames))
    return np.ascontiguousarray(pat.reshape(frames, lanes, 64).transpose(1, 0, 2))


@pytest.mark.parametrize("dtype,block_width", [(np.int16, 32), (np.int16, 64), (np.int16, 128), (np.uint8, 64),
                                               (np.uint8, 128)])
@pytest.mark.parametrize("packet_frames", [64, 256])
def test_oracle_repack_pattern(O, dtype, block_width, packet_frames):
    num_zmw, num_frames, batch_lanes = 16384, 256, 16
    for chunk in range(3):
        pk = packets_chunk(dtype, chunk * num_frames, num_frames, num_zmw, packet_frames, 50, block_width)
        for b0 in (0, 5 * batch_lanes * 64, num_zmw - batch_lanes * 64):
            got = O.repack(pk, b0, chunk * num_frames, batch_lanes, num_frames, dtype)
            assert np.array_equal(got, expected_batch(dtype, b0, chunk * num_frames, batch_lanes, num_frames))
