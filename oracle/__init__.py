"""CPU oracle for the trace-analysis path -- TEST INFRASTRUCTURE ONLY.

ctypes view of liboracle.so, the scalar C++ restatement of the reference's
Host* strategies (oracle/oracle.cpp).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may use it, as the checker and the CPU
baseline.  All arrays are numpy, [lane][frame][zmw] like the GPU path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("liboracle.so is not built; run `make -C oracle`")
    return ctypes.CDLL(LIB_PATH)


lib = _load()
_vp = ctypes.c_void_p
LANE = 64
LATENT = 24


def _p(a):
    return a.ctypes.data_as(_vp)


class DmeDiag(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32 * LANE), ("iter_count", ctypes.c_int32 * LANE),
                ("status", ctypes.c_int32 * LANE), ("iterations_run", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("log_like", ctypes.c_float * LANE), ("conf_factors", (ctypes.c_float * LANE) * 8),
                ("scale_factor", ctypes.c_float * LANE), ("confidence", ctypes.c_float * LANE),
                ("est_fp32", ((ctypes.c_float * LANE) * 3) * 5), ("out_fp32", ((ctypes.c_float * LANE) * 3) * 5),
                ("capture_iter", ctypes.c_int32 * LANE)]


# DME forms (oracle.h): "reference" follows DmeEmHost's arithmetic (log-sum-exp E-step,
# the reference's Cephes exp/log, libm exp2/log2/pow); "mirror" is the GPU kernels'
# restated E-step and math routines, bit for bit; "reference_libm" is "reference"
# with libm expf/logf in place of the Cephes routines (a sensitivity yardstick).
FORMS = {"reference": 0, "mirror": 1, "reference_libm": 2}


def threads(n):
    return lib.orc_threads(int(n))


def tracegen(params, lanes, frames, first_frame, state, init, lane_offset=0, truth=False):
    """Seeded synthetic traces [lanes, frames, 64]; with truth=True also the
    generating state per frame (0 = baseline, a+1 = analog a)."""
    out = np.empty((lanes, frames, LANE), dtype=np.int16)
    tr = np.empty((lanes, frames, LANE), dtype=np.int16) if truth else None
    lib.orc_tracegen(ctypes.byref(params), ctypes.c_int32(lanes), ctypes.c_int32(lane_offset),
                     ctypes.c_int32(frames), ctypes.c_uint32(first_frame), _p(state), ctypes.c_int32(1 if init else 0),
                     _p(out), _p(tr) if truth else None)
    return (out, tr) if truth else out


def transition_matrix(cfg, tr):
    lib.orc_transition_matrix(ctypes.byref(cfg), ctypes.byref(tr))
    return tr


class Baseliner:
    def __init__(self, cfg, lanes):
        lib.orc_baseliner_create.restype = _vp
        self.h = _vp(lib.orc_baseliner_create(ctypes.byref(cfg), ctypes.c_int32(lanes)))
        self.lanes = lanes

    def run(self, raw, stats_dtype):
        raw = np.ascontiguousarray(raw, dtype=np.int16)
        frames = raw.shape[1]
        out = np.empty_like(raw)
        stats = np.zeros(self.lanes, dtype=stats_dtype)
        rc = lib.orc_baseliner_process(self.h, _p(raw), ctypes.c_int32(frames), _p(out), _p(stats))
        if rc != 0:
            raise RuntimeError("oracle baseliner failed: %d" % rc)
        return out, stats

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_baseliner_destroy(self.h)
            self.h = None


class Aggregator:
    def __init__(self, lanes):
        lib.orc_agg_create.restype = _vp
        self.h = _vp(lib.orc_agg_create(ctypes.c_int32(lanes)))
        self.lanes = lanes

    def add(self, stats):
        lib.orc_agg_add(self.h, _p(np.ascontiguousarray(stats)))

    def get(self, stats_dtype):
        out = np.zeros(self.lanes, dtype=stats_dtype)
        lib.orc_agg_get(self.h, _p(out))
        return out

    def reset(self):
        lib.orc_agg_reset(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_agg_destroy(self.h)
            self.h = None


class Histogram:
    def __init__(self, cfg, lanes):
        lib.orc_hist_create.restype = _vp
        self.h = _vp(lib.orc_hist_create(ctypes.byref(cfg), ctypes.c_int32(lanes)))
        self.lanes = lanes

    def reset(self, stats):
        return lib.orc_hist_reset(self.h, _p(np.ascontiguousarray(stats)))

    def add(self, traces, models):
        traces = np.ascontiguousarray(traces, dtype=np.int16)
        lib.orc_hist_add(self.h, _p(traces), ctypes.c_int32(traces.shape[1]), _p(np.ascontiguousarray(models)))

    def get(self, hist_dtype):
        out = np.zeros(self.lanes, dtype=hist_dtype)
        lib.orc_hist_get(self.h, _p(out))
        return out

    def set(self, hist):
        lib.orc_hist_set(self.h, _p(np.ascontiguousarray(hist)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_hist_destroy(self.h)
            self.h = None


def dme_init_models(cfg, stats, model_dtype):
    out = np.zeros(len(stats), dtype=model_dtype)
    lib.orc_dme_init_models(ctypes.byref(cfg), ctypes.c_int32(len(stats)), _p(np.ascontiguousarray(stats)), _p(out))
    return out


def dme_estimate(cfg, hist, stats, models, model_fi, est_fi, want_diag=False, form="reference"):
    models = np.ascontiguousarray(models).copy()
    lanes = len(models)
    mfi = (ctypes.c_int32 * 2)(*model_fi)
    efi = (ctypes.c_int32 * 2)(*est_fi)
    diag = (DmeDiag * lanes)() if want_diag else None
    lib.orc_dme_estimate(ctypes.c_int32(FORMS[form]), ctypes.byref(cfg), ctypes.c_int32(lanes),
                         _p(np.ascontiguousarray(hist)),
                         _p(np.ascontiguousarray(stats)), _p(models), mfi, efi, diag)
    return (models, diag) if want_diag else models


def dme_estimate_learned(cfg, hist, stats, models, model_fi, est_fi, y, omap):
    """Learned-start estimation (mirror arithmetic): y float32 [lanes*64][12] network
    outputs, omap the 12 floats {offset[6], scale[6]}.  Returns (models, fallback
    flags [lanes*64], fallback count)."""
    models = np.ascontiguousarray(models).copy()
    lanes = len(models)
    y = np.ascontiguousarray(y, dtype=np.float32).reshape(lanes * LANE, -1)
    omap = np.ascontiguousarray(omap, dtype=np.float32)[:12]
    fb = np.zeros(lanes * LANE, dtype=np.uint8)
    n = lib.orc_dme_estimate_learned(ctypes.byref(cfg), ctypes.c_int32(lanes), _p(np.ascontiguousarray(hist)),
                                     _p(np.ascontiguousarray(stats)), _p(models), (ctypes.c_int32 * 2)(*model_fi),
                                     (ctypes.c_int32 * 2)(*est_fi), _p(y), _p(omap), _p(fb))
    return models, fb, int(n)


class Labeler:
    def __init__(self, cfg, lanes):
        lib.orc_labeler_create.restype = _vp
        self.h = _vp(lib.orc_labeler_create(ctypes.byref(cfg), ctypes.c_int32(lanes)))
        self.lanes = lanes

    def run(self, traces, models):
        traces = np.ascontiguousarray(traces, dtype=np.int16)
        frames = traces.shape[1]
        labels = np.empty_like(traces)
        latent = np.empty((self.lanes, LATENT, LANE), dtype=np.int16)
        score = np.empty((self.lanes, LANE), dtype=np.float32)
        lib.orc_labeler_process(self.h, _p(traces), ctypes.c_int32(frames), _p(np.ascontiguousarray(models)),
                                _p(labels), _p(latent), _p(score))
        return labels, latent, score

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_labeler_destroy(self.h)
            self.h = None


class Pulser:
    def __init__(self, cfg, lanes):
        lib.orc_pulse_create.restype = _vp
        self.h = _vp(lib.orc_pulse_create(ctypes.byref(cfg), ctypes.c_int32(lanes)))
        self.lanes = lanes
        self.max_calls = cfg.max_calls_per_zmw

    def run(self, labels, latent, traces, first_frame, models, pulse_dtype, stat_dtype):
        labels = np.ascontiguousarray(labels, dtype=np.int16)
        frames = labels.shape[1]
        n = self.lanes * LANE
        pulses = np.zeros((n, self.max_calls), dtype=pulse_dtype)
        counts = np.zeros(n, dtype=np.uint32)
        ovf = np.zeros(n, dtype=np.uint32)
        bl = np.zeros(self.lanes, dtype=stat_dtype)
        lib.orc_pulse_process(self.h, _p(labels), _p(np.ascontiguousarray(latent, dtype=np.int16)),
                              _p(np.ascontiguousarray(traces, dtype=np.int16)), ctypes.c_int32(frames),
                              ctypes.c_int32(first_frame), _p(np.ascontiguousarray(models)), _p(pulses), _p(counts),
                              _p(ovf), _p(bl))
        return pulses, counts, ovf, bl

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_pulse_destroy(self.h)
            self.h = None


class Result(ctypes.Structure):
    _fields_ = [("pulses", _vp), ("pulse_counts", _vp), ("pulse_overflows", _vp), ("pulse_bl_stats", _vp),
                ("baselined", _vp), ("bl_stats", _vp), ("labels", _vp), ("viterbi_score", _vp), ("models", _vp),
                ("metrics", _vp), ("flags", ctypes.c_int32 * 4)]


class Analyzer:
    """BatchAnalyzer (the configured mode) restated on the CPU; em selects the
    DME form ("reference" or "mirror", see FORMS)."""

    def __init__(self, cfg, pool_id, lanes, frames, em="reference"):
        lib.orc_analyzer_create.restype = _vp
        self.h = _vp(lib.orc_analyzer_create(ctypes.byref(cfg), ctypes.c_uint32(pool_id), ctypes.c_int32(lanes),
                                             ctypes.c_int32(frames)))
        lib.orc_analyzer_set_em_form(self.h, ctypes.c_int32(FORMS[em]))
        self.lanes, self.frames, self.pool_id = lanes, frames, pool_id
        self.max_calls = cfg.max_calls_per_zmw

    def __call__(self, traces, first_frame, dtypes, pool_id=None, want=("pulses", "labels", "models")):
        pulse_dt, stat_dt, bls_dt, model_dt = dtypes
        L, F, n = self.lanes, self.frames, self.lanes * LANE
        out = {}
        r = Result()
        if "pulses" in want:
            out["pulses"] = np.zeros((n, self.max_calls), dtype=pulse_dt)
            out["counts"] = np.zeros(n, dtype=np.uint32)
            out["overflows"] = np.zeros(n, dtype=np.uint32)
            out["pulse_bl_stats"] = np.zeros(L, dtype=stat_dt)
            r.pulses, r.pulse_counts = _p(out["pulses"]), _p(out["counts"])
            r.pulse_overflows, r.pulse_bl_stats = _p(out["overflows"]), _p(out["pulse_bl_stats"])
        if "baselined" in want:
            out["baselined"] = np.zeros((L, F, LANE), dtype=np.int16)
            r.baselined = _p(out["baselined"])
        if "bl_stats" in want:
            out["bl_stats"] = np.zeros(L, dtype=bls_dt)
            r.bl_stats = _p(out["bl_stats"])
        if "labels" in want:
            out["labels"] = np.zeros((L, F, LANE), dtype=np.int16)
            r.labels = _p(out["labels"])
        if "score" in want:
            out["score"] = np.zeros((L, LANE), dtype=np.float32)
            r.viterbi_score = _p(out["score"])
        if "models" in want:
            out["models"] = np.zeros(L, dtype=model_dt)
            r.models = _p(out["models"])
        if "metrics" in want:
            from hw_amd import types as T
            out["metrics"] = np.zeros(L, dtype=T.METRICS)
            r.metrics = _p(out["metrics"])
        enc = 1 if np.asarray(traces).dtype == np.uint8 else 0      # HWPA_ENCODING_UINT8 / _INT16
        traces = np.ascontiguousarray(traces, dtype=np.uint8 if enc else np.int16)
        rc = lib.orc_analyzer_process_encoded(self.h, _p(traces), ctypes.c_int32(enc),
                                              ctypes.c_uint32(self.pool_id if pool_id is None else pool_id),
                                              ctypes.c_int32(first_frame), ctypes.byref(r))
        out["rc"] = rc
        out["flags"] = list(r.flags)
        return out

    def set_activity_tree(self, tree):
        """tree: ActivityTree (hw_amd.types) or None."""
        lib.orc_analyzer_set_activity_tree(self.h, ctypes.byref(tree) if tree is not None else None)

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_analyzer_destroy(self.h)
            self.h = None


class Metrics:
    """HFMetricsFilterHost restated on the CPU: one call per chunk of pulses."""

    def __init__(self, cfg, lanes):
        lib.orc_metrics_create.restype = _vp
        self.h = _vp(lib.orc_metrics_create(ctypes.byref(cfg), ctypes.c_int32(lanes)))
        self.lanes = lanes

    def set_activity_tree(self, tree):
        lib.orc_metrics_set_tree(self.h, ctypes.byref(tree) if tree is not None else None)

    def process(self, pulses, counts, max_calls, frames, last_frame, bl_stats, score, pd_stats, models, metrics_dt):
        out = np.zeros(self.lanes, dtype=metrics_dt)
        done = lib.orc_metrics_process(self.h, _p(np.ascontiguousarray(pulses)), _p(np.ascontiguousarray(counts)),
                                       ctypes.c_uint32(max_calls), ctypes.c_int32(frames), ctypes.c_int32(last_frame),
                                       _p(np.ascontiguousarray(bl_stats)),
                                       _p(np.ascontiguousarray(score, dtype=np.float32)),
                                       _p(np.ascontiguousarray(pd_stats)), _p(np.ascontiguousarray(models)), _p(out))
        return done, out

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_metrics_destroy(self.h)
            self.h = None


def autocorr(x, pieces=1):
    """AutocorrAccumulator (lag 4) autocorrelation of a series, built in merged parts:
    `pieces` equal parts, or a list of part sizes, merged left to right."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    lib.orc_autocorr.restype = ctypes.c_float
    if isinstance(pieces, int):
        return float(lib.orc_autocorr(_p(x), ctypes.c_int32(len(x)), None, ctypes.c_int32(pieces)))
    sizes = np.ascontiguousarray(pieces, dtype=np.int32)
    return float(lib.orc_autocorr(_p(x), ctypes.c_int32(len(x)), _p(sizes), ctypes.c_int32(len(sizes))))


# ---------------------------------------------------------------------------
# BAZ packet encoding (layouts: 0 ProductionPulses, 1 InternalPulses, 2 lossy
# truncate, 3 simple overflow, 4 compact overflow)
# ---------------------------------------------------------------------------

BAZ_TRUNC, BAZ_SIMPLE, BAZ_COMPACT = 0, 1, 2
_BAZ_FIELDS = (("label", np.uint8), ("start", np.uint32), ("width", np.uint16), ("isbase", np.uint8),
               ("pkmax", np.float32), ("pkmid", np.float32), ("pkmean", np.float32), ("pkvar", np.float32))


def baz_encode_fields(layout, fields, last=0):
    """fields: dict of per-pulse arrays (label, start, width; optional isbase and the four
    float signals).  Returns (bytes, last start frame)."""
    n = len(fields["label"])
    arrs = {k: (np.ascontiguousarray(fields[k], dtype=dt) if k in fields else None) for k, dt in _BAZ_FIELDS}
    lst = ctypes.c_uint64(last)
    ptrs = [(_p(arrs[k]) if arrs[k] is not None else None) for k, _ in _BAZ_FIELDS]
    lib.orc_baz_encode_fields.restype = ctypes.c_int64
    size = lib.orc_baz_encode_fields(ctypes.c_int32(layout), ctypes.c_int32(n), *ptrs, ctypes.byref(ctypes.c_uint64(last)),
                                     None)
    out = np.zeros(max(size, 1), dtype=np.uint8)
    lib.orc_baz_encode_fields(ctypes.c_int32(layout), ctypes.c_int32(n), *ptrs, ctypes.byref(lst), _p(out))
    return out[:size].tobytes(), lst.value


def baz_decode_fields(layout, data, n, last=0):
    """Returns (fields dict, bytes consumed, last start frame)."""
    buf = np.frombuffer(bytes(data) + b"\0" * 16, dtype=np.uint8).copy()
    out = {k: np.zeros(n, dtype=dt) for k, dt in _BAZ_FIELDS}
    lst = ctypes.c_uint64(last)
    lib.orc_baz_decode_fields.restype = ctypes.c_int64
    used = lib.orc_baz_decode_fields(ctypes.c_int32(layout), ctypes.c_int32(n), _p(buf), ctypes.byref(lst),
                                     *[_p(out[k]) for k, _ in _BAZ_FIELDS])
    return out, int(used), lst.value


def baz_layout_groups(layout):
    bits = np.zeros(16, dtype=np.int32)
    nf = np.zeros(16, dtype=np.int32)
    ng = lib.orc_baz_layout_groups(ctypes.c_int32(layout), _p(bits), _p(nf), ctypes.c_int32(16))
    return list(bits[:ng]), list(nf[:ng])


def baz_serial_roundtrip(ser, bits, obytes, signed, v):
    """(decoded value as uint64, overflow bytes) of one serialiser round trip."""
    ov = ctypes.c_int32(0)
    lib.orc_baz_serial_roundtrip.restype = ctypes.c_uint64
    r = lib.orc_baz_serial_roundtrip(ctypes.c_int32(ser), ctypes.c_int32(bits), ctypes.c_int32(obytes),
                                     ctypes.c_int32(1 if signed else 0), ctypes.c_uint64(v & (2 ** 64 - 1)),
                                     ctypes.byref(ov))
    return r, ov.value


def baz_floatfixed(f, signed, scale, nbytes):
    lib.orc_baz_floatfixed_apply.restype = ctypes.c_uint64
    lib.orc_baz_floatfixed_revert.restype = ctypes.c_float
    code = lib.orc_baz_floatfixed_apply(ctypes.c_float(f), ctypes.c_int32(1 if signed else 0), ctypes.c_uint32(scale),
                                        ctypes.c_int32(nbytes))
    back = lib.orc_baz_floatfixed_revert(ctypes.c_uint64(code), ctypes.c_int32(1 if signed else 0),
                                         ctypes.c_uint32(scale))
    return code, back


def baz_encode_batch(internal, pulses, counts, max_calls, last, cap=None):
    """A chunk of pulse records [zmws * max_calls] into one stream ordered by ZMW.
    `last` (uint64 [zmws]) is the delta state, updated in place when the stream fits.
    Returns (stream bytes or None, offsets [zmws + 1], events [zmws])."""
    zmws = len(counts)
    offsets = np.zeros(zmws + 1, dtype=np.uint64)
    events = np.zeros(zmws, dtype=np.uint32)
    lib.orc_baz_encode_batch.restype = ctypes.c_int64
    pulses = np.ascontiguousarray(pulses)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    assert last.dtype == np.uint64 and last.flags.c_contiguous
    if cap is None:
        tmp = last.copy()
        cap = int(lib.orc_baz_encode_batch(ctypes.c_int32(internal), _p(pulses), _p(counts), ctypes.c_uint32(max_calls),
                                           ctypes.c_int32(zmws), _p(tmp), None, ctypes.c_uint64(0), _p(offsets),
                                           _p(events)))
    out = np.zeros(max(cap, 1), dtype=np.uint8)
    total = int(lib.orc_baz_encode_batch(ctypes.c_int32(internal), _p(pulses), _p(counts), ctypes.c_uint32(max_calls),
                                         ctypes.c_int32(zmws), _p(last), _p(out), ctypes.c_uint64(cap), _p(offsets),
                                         _p(events)))
    return (out[:total].tobytes() if total <= cap else None), offsets, events


def repack(packets, first_zmw, first_frame, lanes, frames, dtype):
    """BlockRepacker's BatchFiller::Process (appModules/BlockRepacker.cpp:255-330) restated:
    packets = [(samples [blocks, packet_frames, block_width], start_zmw, start_frame)];
    returns the batch [lanes, frames, 64] with the overlapping samples copied (zeros elsewhere)."""
    out = np.zeros((frames, lanes * LANE), dtype=dtype)
    for data, z0, f0 in packets:
        nb, pf, bw = data.shape
        rows = np.ascontiguousarray(data.transpose(1, 0, 2)).reshape(pf, nb * bw)   # [frame][zmw]
        za, zb = max(z0, first_zmw), min(z0 + nb * bw, first_zmw + lanes * LANE)
        fa, fb = max(f0, first_frame), min(f0 + pf, first_frame + frames)
        if za >= zb or fa >= fb:
            continue
        out[fa - first_frame:fb - first_frame, za - first_zmw:zb - first_zmw] = rows[fa - f0:fb - f0, za - z0:zb - z0]
    return np.ascontiguousarray(out.reshape(frames, lanes, LANE).transpose(1, 0, 2))
