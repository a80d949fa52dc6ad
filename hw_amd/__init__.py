"""hw_amd -- MI355X (gfx950) trace-analysis path of the smrt-basecaller.

Python view of libhwpa.so's C-ABI (include/hwpa.h).  The library does all
analysis work on the GPU; this module only loads it, mirrors its structures
and moves buffers through the C-ABI's own device helpers.  Importing it when the
library has not been built raises immediately: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import types as T
from .types import (BASELINER_STATS, LANE_HIST, LANE_MODEL, PULSE, STAT_ACC, TRACEGEN_STATE, Config, FilterParams,
                    TracegenParams, Transition, half_to_float)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhwpa.so")


class HwpaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hwpa error %d (%s): %s" % (code, T.STATUS_NAMES.get(code, "?"), msg))
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libhwpa.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.hwpa_last_error.restype = ctypes.c_char_p
    lib.hwpa_edge_state_bytes.restype = ctypes.c_size_t
    return lib


lib = _load()
_vp = ctypes.c_void_p


def check(rc):
    if rc != T.OK:
        raise HwpaError(rc, lib.hwpa_last_error().decode())
    return rc


def _sizes_ok():
    n = lib.hwpa_layout_sizes(None, 0)
    arr = (ctypes.c_int64 * n)()
    lib.hwpa_layout_sizes(arr, n)
    for (name, mine), theirs in zip(T.LAYOUT_ORDER, arr):
        if mine != theirs:
            raise ImportError("layout mismatch for %s: python %d vs C %d" % (name, mine, theirs))


_sizes_ok()


def default_config(**overrides):
    cfg = Config()
    check(lib.hwpa_config_default(ctypes.byref(cfg)))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    check(lib.hwpa_config_validate(ctypes.byref(cfg)))
    return cfg


def load_json(cfg, text):
    check(lib.hwpa_config_load_json(ctypes.byref(cfg), text.encode()))
    return cfg


def bench_config(frames=512, estimate_frames=None):
    """The measurement configuration: sigma EMA off (baseliner start-up 1024
    frames) and one model estimate per chunk, or per `estimate_frames`
    frames when that is longer than a chunk."""
    cfg = default_config()
    est = max(int(frames), int(estimate_frames or 0))
    cfg.sigma_ema_scale_strides = 0.0
    cfg.min_frames_for_estimate = est
    cfg.num_frames_pre_accum_stats = est
    check(lib.hwpa_config_validate(ctypes.byref(cfg)))
    return cfg


def filter_params(ftype):
    fp = FilterParams()
    check(lib.hwpa_filter_params_get(int(ftype), ctypes.byref(fp)))
    return fp


def transition_matrix(cfg):
    tr = Transition()
    check(lib.hwpa_transition_matrix(ctypes.byref(cfg), ctypes.byref(tr)))
    return tr


def tracegen_default(**overrides):
    p = TracegenParams()
    check(lib.hwpa_tracegen_default(ctypes.byref(p)))
    for k, v in overrides.items():
        if k == "amplitude":
            for i, a in enumerate(v):
                p.amplitude[i] = a
        else:
            setattr(p, k, v)
    return p


# ---- device buffers ----------------------------------------------------------

def device_count():
    n = ctypes.c_int32(0)
    check(lib.hwpa_device_count(ctypes.byref(n)))
    return n.value


class DevBuf:
    """A device allocation owned by Python (hipMalloc through the C-ABI)."""

    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        self.p = _vp()
        check(lib.hwpa_dev_alloc(ctypes.byref(self.p), ctypes.c_size_t(self.nbytes)))

    def data_ptr(self):
        return self.p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise HwpaError(1, "upload larger than the buffer")
        if a.nbytes:
            check(lib.hwpa_memcpy_htod(self.p, a.ctypes.data_as(_vp), ctypes.c_size_t(a.nbytes)))
        return self

    def download(self, dtype, shape=None, nbytes=None):
        n = self.nbytes if nbytes is None else nbytes
        if shape is not None:
            n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        out = _dtoh(self.p.value, n).view(dtype)
        return out.reshape(shape) if shape is not None else out

    def view(self, offset, nbytes):
        """A non-owning window [offset, offset + nbytes) of this buffer."""
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes:
            raise HwpaError(1, "view outside the buffer")
        return DevView(self, offset, nbytes)

    def __del__(self):
        if getattr(self, "p", None) is not None and self.p.value:
            lib.hwpa_dev_free(self.p)
            self.p = None


class _Ptr:
    """A borrowed device pointer."""

    def __init__(self, p):
        self.p = p

    def data_ptr(self):
        return self.p


class DevView(DevBuf):
    """Window into a DevBuf; keeps the parent alive, frees nothing."""

    def __init__(self, parent, offset, nbytes):
        self.parent = parent
        self.nbytes = nbytes
        self.p = _vp(parent.data_ptr() + offset)

    def __del__(self):
        pass


def dev_bytes(nbytes):
    return DevBuf(nbytes)


class HostBuf:
    """Page-locked host memory (hwpa_host_alloc); `array(dtype)` views it."""

    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        self.p = _vp()
        check(lib.hwpa_host_alloc(ctypes.byref(self.p), ctypes.c_size_t(self.nbytes)))

    def data_ptr(self):
        return self.p.value

    def array(self, dtype=np.int16):
        n = self.nbytes // np.dtype(dtype).itemsize
        return np.ctypeslib.as_array((ctypes.c_uint8 * (n * np.dtype(dtype).itemsize)).from_address(self.p.value)).view(dtype)

    def __del__(self):
        if getattr(self, "p", None) is not None and self.p.value:
            lib.hwpa_host_free(self.p)
            self.p = None


def host_pinned(nbytes):
    return HostBuf(nbytes)


def dev_from_numpy(a):
    a = np.ascontiguousarray(a)
    return DevBuf(a.nbytes).upload(a)


def to_numpy(buf, dtype, shape=None):
    sync()
    return buf.download(dtype, shape)


def fetch(dev, dtype, shape):
    """Copies `shape` elements of `dtype` from a device buffer or borrowed pointer."""
    sync()
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    return _dtoh(dev.data_ptr(), n).view(dtype).reshape(shape)


def ptr(buf):
    return _vp(buf.data_ptr())


def sync():
    check(lib.hwpa_device_sync())


def count_equal_i16(a, b, n):
    """Equal elements among the first n int16 values of two device buffers."""
    out = ctypes.c_uint64(0)
    check(lib.hwpa_count_equal_i16(ptr(a), ptr(b), ctypes.c_size_t(n), ctypes.byref(out)))
    return out.value


TUNE_HISTOGRAM, TUNE_BASELINE, TUNE_LABEL, TUNE_PULSE = 0, 1, 2, 3
_TUNE_NAMES = ("histogram", "baseline", "labels", "pulses")


def tune_variants(kernel):
    n = ctypes.c_int32(0)
    check(lib.hwpa_tune_variants(ctypes.c_int32(kernel), ctypes.byref(n)))
    return n.value


def tune_set(kernel, lanes, frames, variant):
    check(lib.hwpa_tune_set(ctypes.c_int32(kernel), ctypes.c_uint32(lanes), ctypes.c_uint32(frames),
                            ctypes.c_int32(variant)))


def tune_get(kernel, lanes, frames):
    v = ctypes.c_int32(0)
    check(lib.hwpa_tune_get(ctypes.c_int32(kernel), ctypes.c_uint32(lanes), ctypes.c_uint32(frames), ctypes.byref(v)))
    return v.value


def autotune(cfg, lanes, frames, repeats=3):
    """Times every kernel variant on this shape and keeps the fastest; returns
    {"histogram": [ms...], "baseline": [...], "labels": [...], "pulses": [...],
    "chosen": (one variant per kernel, in that order)}."""
    counts = [tune_variants(k) for k in range(len(_TUNE_NAMES))]
    ms = (ctypes.c_float * sum(counts))()
    check(lib.hwpa_autotune(ctypes.byref(cfg), ctypes.c_uint32(lanes), ctypes.c_uint32(frames),
                            ctypes.c_int32(repeats), ms, ctypes.c_int32(sum(counts))))
    out, o = {}, 0
    for name, n in zip(_TUNE_NAMES, counts):
        out[name] = list(ms)[o:o + n]
        o += n
    out["chosen"] = tuple(tune_get(k, lanes, frames) for k in range(len(_TUNE_NAMES)))
    return out


def copy_bandwidth(nbytes=1 << 30, launches=10):
    """Mean ms per launch of the library's plain copy kernel over nbytes, and
    the resulting read+write GB/s."""
    src, dst = dev_bytes(nbytes), dev_bytes(nbytes)
    ms = ctypes.c_float(0.0)
    check(lib.hwpa_copy_bandwidth(ptr(dst), ptr(src), ctypes.c_size_t(nbytes), launches, ctypes.byref(ms)))
    return ms.value, 2.0 * nbytes / (ms.value * 1e-3) / 1e9


# ---- stage wrappers ------------------------------------------------------------

class Baseliner:
    """HostMultiScaleBaseliner/DeviceMultiScaleBaseliner counterpart: one
    persistent filter state per lane, one call per chunk."""

    def __init__(self, cfg, lanes):
        self.h = _vp()
        self.lanes = lanes
        check(lib.hwpa_baseliner_create(ctypes.byref(cfg), ctypes.c_uint32(lanes), ctypes.byref(self.h)))

    def run(self, d_raw, frames, d_out, d_stats, stream=None, encoding=T.ENCODING_INT16):
        check(lib.hwpa_baseliner_run_encoded(self.h, ptr(d_raw), ctypes.c_int32(encoding), ctypes.c_uint32(frames),
                                             ptr(d_out), ptr(d_stats), _vp(stream)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.hwpa_baseliner_destroy(self.h)
            self.h = None


def stats_reset(d_agg, lanes, stream=None):
    check(lib.hwpa_stats_reset(ptr(d_agg), ctypes.c_uint32(lanes), _vp(stream)))


def stats_merge(d_agg, d_chunk, lanes, stream=None):
    check(lib.hwpa_stats_merge(ptr(d_agg), ptr(d_chunk), ctypes.c_uint32(lanes), _vp(stream)))


def edge_state_bytes():
    return lib.hwpa_edge_state_bytes()


def edge_state_init(d_edge, lanes, stream=None):
    check(lib.hwpa_edge_state_init(ptr(d_edge), ctypes.c_uint32(lanes), _vp(stream)))


def hist_reset(cfg, d_hist, d_stats, lanes, stream=None):
    check(lib.hwpa_hist_reset(ctypes.byref(cfg), ptr(d_hist), ptr(d_stats), ctypes.c_uint32(lanes), _vp(stream)))


def hist_add(d_hist, d_edge, d_traces, frames, d_models, lanes, stream=None):
    check(lib.hwpa_hist_add(ptr(d_hist), ptr(d_edge), ptr(d_traces), ctypes.c_uint32(frames), ptr(d_models),
                            ctypes.c_uint32(lanes), _vp(stream)))


def dme_init_models(cfg, d_stats, d_models, lanes, stream=None):
    check(lib.hwpa_dme_init_models(ctypes.byref(cfg), ptr(d_stats), ptr(d_models), ctypes.c_uint32(lanes),
                                   _vp(stream)))


def dme_estimate(cfg, d_hist, d_stats, d_models, model_fi, est_fi, lanes, d_iters=None, stream=None):
    mfi = (ctypes.c_int32 * 2)(*model_fi)
    efi = (ctypes.c_int32 * 2)(*est_fi)
    check(lib.hwpa_dme_estimate(ctypes.byref(cfg), ptr(d_hist), ptr(d_stats), ptr(d_models), mfi, efi,
                                ctypes.c_uint32(lanes), ptr(d_iters) if d_iters is not None else None, _vp(stream)))


def dme_estimate_diag(cfg, d_hist, d_stats, d_models, model_fi, est_fi, lanes, d_diag, d_iters=None, stream=None):
    """As dme_estimate, also writing the raw per-ZMW EM result to d_diag
    (float32 [lanes][EM_DIAG_ROWS][64])."""
    mfi = (ctypes.c_int32 * 2)(*model_fi)
    efi = (ctypes.c_int32 * 2)(*est_fi)
    check(lib.hwpa_dme_estimate_diag(ctypes.byref(cfg), ptr(d_hist), ptr(d_stats), ptr(d_models), mfi, efi,
                                     ctypes.c_uint32(lanes), ptr(d_iters) if d_iters is not None else None,
                                     ptr(d_diag), _vp(stream)))


def dme_estimate_learned(cfg, init, d_hist, d_stats, d_models, model_fi, est_fi, lanes, d_iters=None, stream=None):
    """Learned-start estimation; returns the number of ZMWs re-estimated by the plain EM."""
    mfi = (ctypes.c_int32 * 2)(*model_fi)
    efi = (ctypes.c_int32 * 2)(*est_fi)
    n = ctypes.c_uint32(0)
    check(lib.hwpa_dme_estimate_learned(ctypes.byref(cfg), init.h, ptr(d_hist), ptr(d_stats), ptr(d_models), mfi,
                                        efi, ctypes.c_uint32(lanes), ptr(d_iters) if d_iters is not None else None,
                                        ctypes.byref(n), _vp(stream)))
    return n.value


NOALLOC_MAX_LANES = 256


def dme_estimate_learned_noalloc(cfg, init, d_hist, d_stats, d_models, model_fi, est_fi, lanes, d_iters=None,
                                 stream=None):
    """As dme_estimate_learned, with the workspace in the library's fixed device
    array (at most NOALLOC_MAX_LANES lanes; no device allocation)."""
    mfi = (ctypes.c_int32 * 2)(*model_fi)
    efi = (ctypes.c_int32 * 2)(*est_fi)
    n = ctypes.c_uint32(0)
    check(lib.hwpa_dme_estimate_learned_noalloc(ctypes.byref(cfg), init.h, ptr(d_hist), ptr(d_stats), ptr(d_models),
                                                mfi, efi, ctypes.c_uint32(lanes),
                                                ptr(d_iters) if d_iters is not None else None, ctypes.byref(n),
                                                _vp(stream)))
    return n.value


MLP_IN, MLP_HIDDEN, MLP_OUT = 128, 256, 12
MLP_PARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_OUT * MLP_HIDDEN + MLP_OUT
EM_DIAG_ROWS = 18


class Initializer:
    """The learned EM initialiser (128 -> 256 -> 256 -> 11 perceptron)."""

    def __init__(self, seed=1):
        self.h = _vp()
        check(lib.hwpa_initializer_create(ctypes.c_uint64(seed), ctypes.byref(self.h)))

    def params(self):
        p = np.empty(MLP_PARAMS, dtype=np.float32)
        check(lib.hwpa_initializer_get_params(self.h, p.ctypes.data_as(_vp), ctypes.c_size_t(MLP_PARAMS)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib.hwpa_initializer_set_params(self.h, p.ctypes.data_as(_vp), ctypes.c_size_t(p.size)))

    def set_output_map(self, offset, scale, weight):
        """weight: 8 loss weights (6 regression outputs, mixture cross-entropy,
        confidence logistic loss); 7 values keep the confidence weight at 1."""
        weight = list(weight) + [1.0] * (8 - len(weight))
        off = (ctypes.c_float * 6)(*[float(v) for v in offset])
        sc = (ctypes.c_float * 6)(*[float(v) for v in scale])
        w = (ctypes.c_float * 8)(*[float(v) for v in weight])
        check(lib.hwpa_initializer_set_output_map(self.h, off, sc, w))

    def output_map(self):
        out = (ctypes.c_float * 20)()
        check(lib.hwpa_initializer_get_output_map(self.h, out))
        return np.array(out, dtype=np.float32)

    def forward(self, d_x, rows, d_y, stream=None):
        check(lib.hwpa_mlp_forward(self.h, ptr(d_x), ctypes.c_uint32(rows), ptr(d_y), _vp(stream)))

    def forward_hist(self, d_hist, lanes, d_y, stream=None):
        check(lib.hwpa_mlp_forward_hist(self.h, ptr(d_hist), ctypes.c_uint32(lanes), ptr(d_y), _vp(stream)))

    def gradients(self, d_x, d_t, batch, d_grads, d_loss, d_valid=None, stream=None):
        check(lib.hwpa_mlp_gradients(self.h, ptr(d_x), ptr(d_t), _opt(d_valid), ctypes.c_uint32(batch), ptr(d_grads),
                                     ptr(d_loss), _vp(stream)))

    def train_step(self, d_x, d_t, batch, lr, d_loss=None, d_valid=None, stream=None):
        check(lib.hwpa_mlp_train_step(self.h, ptr(d_x), ptr(d_t), _opt(d_valid), ctypes.c_uint32(batch),
                                      ctypes.c_float(lr), _opt(d_loss), _vp(stream)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.hwpa_initializer_destroy(self.h)
            self.h = _vp()


def _opt(buf):
    return ptr(buf) if buf is not None else None


def mlp_features(d_hist, lanes, d_x, stream=None):
    check(lib.hwpa_mlp_features(ptr(d_hist), ctypes.c_uint32(lanes), ptr(d_x), _vp(stream)))


def mlp_targets(d_hist, d_diag, lanes, min_conf, d_t, d_valid, stream=None):
    check(lib.hwpa_mlp_targets(ptr(d_hist), ptr(d_diag), ctypes.c_uint32(lanes), ctypes.c_float(min_conf), ptr(d_t),
                               ptr(d_valid), _vp(stream)))


class Labeler:
    """FrameLabelerHost/FrameLabelerDevice counterpart."""

    def __init__(self, cfg, lanes, max_frames):
        self.h = _vp()
        check(lib.hwpa_labeler_create(ctypes.byref(cfg), ctypes.c_uint32(lanes), ctypes.c_uint32(max_frames),
                                      ctypes.byref(self.h)))

    def run(self, d_traces, frames, d_models, d_labels, d_latent, d_score, stream=None):
        check(lib.hwpa_labeler_run(self.h, ptr(d_traces), ctypes.c_uint32(frames), ptr(d_models), ptr(d_labels),
                                   ptr(d_latent), ptr(d_score), _vp(stream)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.hwpa_labeler_destroy(self.h)
            self.h = None


class Pulser:
    """HostPulseAccumulator/DevicePulseAccumulator counterpart."""

    def __init__(self, cfg, lanes):
        self.h = _vp()
        check(lib.hwpa_pulser_create(ctypes.byref(cfg), ctypes.c_uint32(lanes), ctypes.byref(self.h)))

    def run(self, d_labels, d_latent, d_traces, frames, first_frame, d_models, d_pulses, d_counts, d_ovf,
            d_blstats, stream=None):
        check(lib.hwpa_pulser_run(self.h, ptr(d_labels), ptr(d_latent), ptr(d_traces), ctypes.c_uint32(frames),
                                  ctypes.c_int32(first_frame), ptr(d_models), ptr(d_pulses), ptr(d_counts),
                                  ptr(d_ovf), ptr(d_blstats), _vp(stream)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.hwpa_pulser_destroy(self.h)
            self.h = None


def tracegen(params, lanes, frames, first_frame, d_state, init, d_out, lane_offset=0, stream=None):
    check(lib.hwpa_tracegen_run(ctypes.byref(params), ctypes.c_uint32(lanes), ctypes.c_uint32(lane_offset),
                                ctypes.c_uint32(frames), ctypes.c_uint32(first_frame), ptr(d_state),
                                ctypes.c_int32(1 if init else 0), ptr(d_out), _vp(stream)))


def activity_tree(children_left, children_right, feature, threshold, value, hsw_curve, max_acceptable_hsw_rate):
    """A trained activity-labeler tree as the C-ABI takes it; returns
    (ActivityTree, arrays) -- keep the tuple alive while the tree is in use."""
    arrs = (np.ascontiguousarray(children_left, dtype=np.int16), np.ascontiguousarray(children_right, dtype=np.int16),
            np.ascontiguousarray(feature, dtype=np.int8), np.ascontiguousarray(threshold, dtype=np.float32),
            np.ascontiguousarray(value, dtype=np.int8))
    t = T.ActivityTree()
    t.num_nodes = len(arrs[0])
    t.max_acceptable_hsw_rate = float(max_acceptable_hsw_rate)
    t.hsw_curve[0], t.hsw_curve[1] = float(hsw_curve[0]), float(hsw_curve[1])
    t.children_left, t.children_right, t.feature, t.threshold, t.value = [a.ctypes.data for a in arrs]
    return t, arrs


class Metrics:
    """HFMetricsFilter counterpart (traceAnalysis/HFMetricsFilterHost.cpp): one
    process call per chunk of pulses; returns True when a block ended and its
    BasecallingMetrics were written to `out`."""

    def __init__(self, cfg, lanes):
        self.h = _vp()
        self.lanes = lanes
        check(lib.hwpa_metrics_create(ctypes.byref(cfg), ctypes.c_uint32(lanes), ctypes.byref(self.h)))

    def set_activity_tree(self, tree):
        check(lib.hwpa_metrics_set_activity_tree(self.h, ctypes.byref(tree[0]) if tree is not None else None))

    def process(self, d_pulses, d_counts, max_calls, frames, last_frame, d_bl_stats, d_score, d_pd_stats, d_models,
                d_out, stream=None):
        produced = ctypes.c_int32(0)
        check(lib.hwpa_metrics_process(self.h, ptr(d_pulses), ptr(d_counts), ctypes.c_uint32(max_calls),
                                       ctypes.c_uint32(frames), ctypes.c_int32(last_frame), ptr(d_bl_stats),
                                       ptr(d_score), ptr(d_pd_stats), ptr(d_models), ptr(d_out),
                                       ctypes.byref(produced), _vp(stream)))
        return produced.value != 0

    def __del__(self):
        if getattr(self, "h", None):
            lib.hwpa_metrics_destroy(self.h)
            self.h = None


class Repacker:
    """BlockRepacker counterpart: dense sensor packets (device buffers) into a lane-major batch."""

    def __init__(self, max_packets=4096):
        self.h = _vp()
        check(lib.hwpa_repacker_create(ctypes.c_uint32(max_packets), ctypes.byref(self.h)))

    def run(self, packets, encoding, d_batch, first_zmw, first_frame, lanes, frames, stream=None):
        """packets: list of (device buffer, start_zmw, start_frame, blocks, frames, block_width)."""
        n = len(packets)
        arr = (T.PacketDesc * n)()
        for i, (buf, z0, f0, nb, nf, bw) in enumerate(packets):
            arr[i] = T.PacketDesc(buf.data_ptr(), z0, f0, nb, nf, bw, 0)
        check(lib.hwpa_repack(self.h, arr, ctypes.c_uint32(n), ctypes.c_int32(encoding), ptr(d_batch),
                              ctypes.c_uint32(first_zmw), ctypes.c_uint32(first_frame), ctypes.c_uint32(lanes),
                              ctypes.c_uint32(frames), _vp(stream)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.hwpa_repacker_destroy(self.h)
            self.h = None


class BazEncoder:
    """BAZ packet encoder of a set of ZMWs (ProductionPulses, or InternalPulses when
    internal=True) with the per-ZMW StartFrame delta state kept across chunks."""

    def __init__(self, zmws, internal=False):
        self.h = _vp()
        self.zmws = zmws
        check(lib.hwpa_baz_encoder_create(ctypes.c_uint32(zmws), ctypes.c_int32(1 if internal else 0),
                                          ctypes.byref(self.h)))

    def encode(self, d_pulses, d_counts, max_calls, d_out, capacity, d_offsets, d_events, stream=None):
        check(lib.hwpa_baz_encode(self.h, ptr(d_pulses), ptr(d_counts), ctypes.c_uint32(max_calls),
                                  ptr(d_out) if d_out is not None else None, ctypes.c_uint64(capacity),
                                  ptr(d_offsets), ptr(d_events), _vp(stream)))

    def reset(self, stream=None):
        check(lib.hwpa_baz_encoder_reset(self.h, _vp(stream)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.hwpa_baz_encoder_destroy(self.h)
            self.h = None


class BatchAnalyzer:
    """Per-pool analyzer with the reference's BatchAnalyzer shape
    (analyzer/BatchAnalyzer.h:51-106): construct once per pool, call once per
    chunk in frame order; results stay on the device until fetched."""

    def __init__(self, pool_id, lanes, frames_per_chunk, cfg):
        self.h = _vp()
        self.pool_id = pool_id
        self.lanes = lanes
        self.frames = frames_per_chunk
        self.cfg = cfg
        check(lib.hwpa_analyzer_create(ctypes.c_uint32(pool_id), ctypes.c_uint32(lanes),
                                       ctypes.c_uint32(frames_per_chunk), ctypes.byref(cfg), ctypes.byref(self.h)))

    def pool_id_(self):
        return self.pool_id

    def __call__(self, traces, first_frame, pool_id=None, on_device=None, encoding=None):
        """traces: numpy int16 or uint8 [lanes, frames, 64] (host) or a device
        buffer (anything with data_ptr()) already in HBM; `encoding` (ENCODING_*)
        is taken from a numpy array's dtype, and defaults to int16 for device buffers.
        on_device=False marks a data_ptr() buffer as host memory (a pinned tensor):
        the chunk is copied in, overlapping the previous chunk's work, and the call
        returns once the copy is done."""
        pid = self.pool_id if pool_id is None else pool_id
        if isinstance(traces, np.ndarray):
            enc = T.ENCODING_UINT8 if traces.dtype == np.uint8 else T.ENCODING_INT16
            a = np.ascontiguousarray(traces, dtype=np.uint8 if enc == T.ENCODING_UINT8 else np.int16)
            if a.size != self.lanes * self.frames * T.LANE:
                raise HwpaError(1, "trace batch shape does not match the analyzer")
            rc = lib.hwpa_analyzer_process_encoded(self.h, a.ctypes.data_as(_vp), ctypes.c_int32(enc),
                                                   ctypes.c_int32(0), ctypes.c_uint32(pid),
                                                   ctypes.c_int32(first_frame))
        else:
            enc = T.ENCODING_INT16 if encoding is None else encoding
            dev = 1 if on_device is None or on_device else 0
            rc = lib.hwpa_analyzer_process_encoded(self.h, ptr(traces), ctypes.c_int32(enc), ctypes.c_int32(dev),
                                                   ctypes.c_uint32(pid), ctypes.c_int32(first_frame))
        check(rc)
        return self

    def sync(self):
        check(lib.hwpa_analyzer_sync(self.h))

    def stream(self):
        s = _vp()
        check(lib.hwpa_analyzer_stream(self.h, ctypes.byref(s)))
        return s.value

    def flags(self):
        f = (ctypes.c_int32 * 4)()
        check(lib.hwpa_result_flags(self.h, f))
        return list(f)

    def _fetch(self, getter, nbytes, dtype, shape):
        p = _vp()
        check(getter(self.h, ctypes.byref(p)))
        self.sync()
        return _dtoh(p.value, nbytes).view(dtype).reshape(shape)

    def pulses_device(self):
        """(pulses, counts, max_calls) of the last chunk as borrowed device pointers."""
        pp, pc, po = _vp(), _vp(), _vp()
        mc = ctypes.c_uint32()
        check(lib.hwpa_result_pulses(self.h, ctypes.byref(pp), ctypes.byref(pc), ctypes.byref(po), ctypes.byref(mc)))
        return _Ptr(pp.value), _Ptr(pc.value), mc.value

    def labels_device(self):
        """Device pointer of the last labels (valid until the next call)."""
        p = ctypes.c_void_p()
        check(lib.hwpa_result_labels(self.h, ctypes.byref(p)))
        return _Ptr(p.value)

    def labels(self):
        return self._fetch(lib.hwpa_result_labels, self.lanes * self.frames * T.LANE * 2, np.int16,
                           (self.lanes, self.frames, T.LANE))

    def baselined(self):
        return self._fetch(lib.hwpa_result_baselined, self.lanes * self.frames * T.LANE * 2, np.int16,
                           (self.lanes, self.frames, T.LANE))

    def models(self):
        return self._fetch(lib.hwpa_result_models, self.lanes * LANE_MODEL.itemsize, LANE_MODEL, (self.lanes,))

    def baseliner_stats(self):
        return self._fetch(lib.hwpa_result_baseliner_stats, self.lanes * BASELINER_STATS.itemsize,
                           BASELINER_STATS, (self.lanes,))

    def viterbi_score(self):
        return self._fetch(lib.hwpa_result_viterbi_score, self.lanes * T.LANE * 4, np.float32,
                           (self.lanes, T.LANE))

    def pulse_baseline_stats(self):
        return self._fetch(lib.hwpa_result_pulse_baseline_stats, self.lanes * STAT_ACC.itemsize, STAT_ACC,
                           (self.lanes,))

    def metrics(self):
        """BasecallingMetrics of the block that ended with the last chunk (flags()[3])."""
        return self._fetch(lib.hwpa_result_metrics, self.lanes * T.METRICS.itemsize, T.METRICS, (self.lanes,))

    def set_activity_tree(self, tree):
        """tree: an activity_tree(...) result or None."""
        check(lib.hwpa_analyzer_set_activity_tree(self.h, ctypes.byref(tree[0]) if tree is not None else None))

    def pulses(self):
        """Returns (pulses [lanes*64, max_calls], counts [lanes*64], overflows [lanes*64])."""
        pp, pc, po = _vp(), _vp(), _vp()
        mc = ctypes.c_uint32()
        check(lib.hwpa_result_pulses(self.h, ctypes.byref(pp), ctypes.byref(pc), ctypes.byref(po),
                                     ctypes.byref(mc)))
        n = self.lanes * T.LANE
        mc = mc.value
        self.sync()
        res = []
        for p, nbytes, dt, shape in ((pp, n * mc * PULSE.itemsize, PULSE, (n, mc)), (pc, n * 4, np.uint32, (n,)),
                                     (po, n * 4, np.uint32, (n,))):
            res.append(_dtoh(p.value, nbytes).view(dt).reshape(shape))
        return tuple(res)

    def set_profiling(self, on=True):
        check(lib.hwpa_analyzer_set_profiling(self.h, ctypes.c_int32(1 if on else 0)))

    def stage_ms(self):
        """[baseline, aggregate, histogram, dme, labels, pulses, metrics] ms of the last call."""
        ms = (ctypes.c_float * 8)()
        check(lib.hwpa_analyzer_stage_ms(self.h, ms))
        return list(ms)[:7]

    def em_iterations(self):
        it = ctypes.c_int32()
        check(lib.hwpa_analyzer_em_iterations(self.h, ctypes.byref(it)))
        return it.value

    def set_initializer(self, init):
        self._init = init
        check(lib.hwpa_analyzer_set_initializer(self.h, init.h if init is not None else None))

    def capture_training(self, d_x, d_t, d_valid, min_conf=0.1):
        """Arm (or with d_x None, disarm) capture of (input, target, valid) at each plain estimate."""
        self._cap = (d_x, d_t, d_valid)
        check(lib.hwpa_analyzer_capture_training(self.h, _opt(d_x), _opt(d_t), _opt(d_valid),
                                                 ctypes.c_float(min_conf)))

    def fallback_count(self):
        n = ctypes.c_uint32()
        check(lib.hwpa_analyzer_fallback_count(self.h, ctypes.byref(n)))
        return n.value

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_owned", True):
            lib.hwpa_analyzer_destroy(self.h)
        self.h = None


class _PoolView(BatchAnalyzer):
    """Result view of one pool of a Batch (the batch owns the analyzer)."""

    def __init__(self, batch, index, pool_id, lanes, frames, cfg):
        self.h = _vp()
        check(lib.hwpa_batch_pool(batch.h, ctypes.c_uint32(index), ctypes.byref(self.h)))
        self._owned = False
        self._batch = batch
        self.pool_id, self.lanes, self.frames, self.cfg = pool_id, lanes, frames, cfg


class Batch:
    """Several pools run one chunk step at a time, captured into one hipGraph
    per step (pools concurrent)."""

    def __init__(self, pools):
        """pools: list of (pool_id, lanes, frames_per_chunk, cfg)."""
        n = len(pools)
        ids = (ctypes.c_uint32 * n)(*[p[0] for p in pools])
        lanes = (ctypes.c_uint32 * n)(*[p[1] for p in pools])
        frames = (ctypes.c_uint32 * n)(*[p[2] for p in pools])
        cfgs = (Config * n)(*[p[3] for p in pools])
        self.h = _vp()
        check(lib.hwpa_batch_create(ctypes.c_uint32(n), ids, lanes, frames, cfgs, ctypes.byref(self.h)))
        self.specs = list(pools)
        self.views = [_PoolView(self, i, *p) for i, p in enumerate(pools)]

    def set_graph(self, on=True):
        check(lib.hwpa_batch_set_graph(self.h, ctypes.c_int32(1 if on else 0)))

    def set_initializer(self, init):
        self._init = init
        check(lib.hwpa_batch_set_initializer(self.h, init.h if init is not None else None))

    def process(self, chunks, first_frames, encodings=None):
        """chunks: one device buffer per pool; encodings: ENCODING_* per pool (int16 if None)."""
        n = len(self.specs)
        ptrs = (ctypes.c_void_p * n)(*[c.data_ptr() for c in chunks])
        ff = (ctypes.c_int32 * n)(*first_frames)
        if encodings is None:
            check(lib.hwpa_batch_process(self.h, ptrs, ff))
        else:
            enc = (ctypes.c_int32 * n)(*encodings)
            check(lib.hwpa_batch_process_encoded(self.h, ptrs, enc, ff))

    def sync(self):
        check(lib.hwpa_batch_sync(self.h))

    def autotune(self, repeats=3):
        check(lib.hwpa_batch_autotune(self.h, ctypes.c_int32(repeats)))

    def graph_stats(self):
        st = (ctypes.c_uint64 * 3)()
        check(lib.hwpa_batch_graph_stats(self.h, st))
        return {"steps": st[0], "instantiations": st[1], "updates": st[2]}

    def pool(self, i):
        return self.views[i]

    def __del__(self):
        if getattr(self, "h", None):
            for v in getattr(self, "views", []):
                v.h = None
            lib.hwpa_batch_destroy(self.h)
            self.h = None


def _dtoh(src, nbytes):
    out = np.empty(nbytes, dtype=np.uint8)
    if nbytes:
        check(lib.hwpa_memcpy_dtoh(out.ctypes.data_as(_vp), _vp(src), ctypes.c_size_t(nbytes)))
    return out
