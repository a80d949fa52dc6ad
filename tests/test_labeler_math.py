"""The labeler's divide-free ROI quotient and word-parallel ROI backward pass
(include/hwpa_math.h) against the arithmetic they replace.

RoiFilterDefault::SmoothedVal (RoiHost.h:91-99) divides window sums by
sqrt(2.0f) and sqrt(3.0f).  hw_div_by_const must give the IEEE quotient: the
check covers every mantissa of both signs at exponents near the bottom of its
stated range (2^-100), around 1, and at the top (where x * (1/d) could
overflow), plus zeros, infinities, NaN and the extreme finite values.  Within an
exponent range where no step under- or overflows, scaling x by 2 scales every
step exactly, so those mantissa sweeps stand for every exponent between them.
Below 2^-100 the residual stops being exact; the sweep shows that too, so the
stated bound is a real one.

Roi::BackwardRecursion (RoiHost.h:195-243) walks frames backwards with
final_k = roi_k | (mid_k & final_k+1); hw_roi_backward_word does 16 frames at
once with one addition.  It is checked against that loop on random words,
words with every 8-field pattern, and the carry in both states."""
import ctypes

import numpy as np
import pytest

S2 = np.float32(np.sqrt(np.float32(2.0)))
S3 = np.float32(np.sqrt(np.float32(3.0)))


def _div_sweep(O, d, exps):
    e = np.ascontiguousarray(exps, dtype=np.int32)
    out = (ctypes.c_uint64 * 2)()
    O.lib.orc_div_by_const_sweep(ctypes.c_float(float(d)), e.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_int32(len(e)), out)
    return out[0], out[1]


@pytest.mark.parametrize("d", [S2, S3], ids=["sqrt2", "sqrt3"])
def test_div_by_const_matches_ieee(O, d):
    assert float(S2) == float(np.float32(1.41421354)) and float(S3) == float(np.float32(1.73205078))
    # 2^-100 has biased exponent 27; 127 is [1, 2); 252-254 the top binades
    checked, wrong = _div_sweep(O, d, [27, 28, 29, 126, 127, 128, 252, 253, 254])
    assert checked == 9 * 2 * (1 << 23) + 7
    assert wrong == 0


@pytest.mark.parametrize("d", [S2, S3], ids=["sqrt2", "sqrt3"])
def test_div_by_const_fails_below_its_range(O, d):
    # the stated lower bound is where exactness really ends: far below it results differ
    _, wrong = _div_sweep(O, d, [1, 2])
    assert wrong > 0


def _backward_ref(word, carry):
    """Roi::BackwardRecursion's per-frame loop over one 16-field word."""
    pr = carry
    out = 0
    for k in range(15, -1, -1):
        r = (word >> (2 * k)) & 3
        b = (r & 1) or ((r & 2) and pr)
        pr = 1 if b else 0
        out |= pr << (2 * k)
    return out, pr


def _backward_lib(O, words, carries):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    c = np.ascontiguousarray(carries, dtype=np.uint32)
    ow = np.zeros_like(w)
    oc = np.zeros_like(c)
    O.lib.orc_roi_backward_words(w.ctypes.data_as(ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_int32(len(w)), ow.ctypes.data_as(ctypes.c_void_p),
                                 oc.ctypes.data_as(ctypes.c_void_p))
    return ow, oc


def test_roi_backward_word_matches_frame_loop(O):
    rng = np.random.default_rng(5)
    words = [int(x) for x in rng.integers(0, 1 << 32, 3000, dtype=np.uint64)]
    # every pattern of the low 8 fields, high fields random, both carries
    for pat in range(0, 1 << 16, 7):
        words.append(pat | (int(rng.integers(0, 1 << 16)) << 16))
    # long propagation runs: all mid, mid with one roi, fields zero above a valid prefix
    words += [0xAAAAAAAA, 0xAAAAAAAB, 0x2AAAAAAA, 0x0000AAAA, 0xAAAA0000, 0x00000003, 0xC0000000, 0x80000000,
              0x00000002, 0xFFFFFFFF, 0x55555555, 0]
    carries = [i & 1 for i in range(len(words))] + [1 - (i & 1) for i in range(len(words))]
    words = words + words
    got_w, got_c = _backward_lib(O, words, carries)
    for w, c, gw, gc in zip(words, carries, got_w, got_c):
        ew, ec = _backward_ref(w, c)
        assert (int(gw), int(gc)) == (ew, ec), hex(w)


def test_roi_backward_words_chain_like_the_oracle(O):
    """Words chained from the top with the carry, against the oracle's whole-sequence loop
    (the forward pass leaves fields past the last frame zero)."""
    rng = np.random.default_rng(9)
    for nroi in (1, 15, 16, 17, 33, 536):
        fields = rng.integers(0, 4, nroi)
        seq_out = []
        pr = 0
        for k in range(nroi - 1, -1, -1):
            r = int(fields[k])
            pr = 1 if ((r & 1) or ((r & 2) and pr)) else 0
            seq_out.append(pr)
        seq_out = seq_out[::-1]
        nw = (nroi + 15) // 16
        words = [0] * nw
        for k in range(nroi):
            words[k // 16] |= int(fields[k]) << (2 * (k % 16))
        carry = 0
        got = [0] * nroi
        for wi in range(nw - 1, -1, -1):
            ow, oc = _backward_lib(O, [words[wi]], [carry])
            carry = int(oc[0])
            for k in range(16):
                if wi * 16 + k < nroi:
                    got[wi * 16 + k] = (int(ow[0]) >> (2 * k)) & 3
        assert got == seq_out, nroi
