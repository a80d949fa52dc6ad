"""ctypes / numpy mirrors of include/hwpa_types.h.

The numpy dtypes describe the per-lane structure-of-arrays records exactly
(C alignment), so device buffers copied to the host can be viewed as
structured arrays; their sizes are checked against the library's own
sizeof() at load time.
"""
import ctypes

import numpy as np

LANE = 64
NUM_ANALOGS = 4
NUM_BINS = 300
NUM_STATES = 13
LATENT_FRAMES = 24

OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "BAD_CONFIG", 3: "BAD_POOL", 4: "BAD_FRAME",
                5: "HIP", 6: "NO_DEVICE", 7: "UNSUPPORTED", 8: "OOM"}
ERR_BAD_POOL = 3
ERR_BAD_FRAME = 4
ERR_BAD_CONFIG = 2

MODEL_FIXED, MODEL_INITIAL, MODEL_DYNAMIC = 0, 1, 2
ENCODING_INT16, ENCODING_UINT8 = 0, 1

FILTER_TYPES = {"MultiScaleLarge": 0, "MultiScaleMedium": 1, "MultiScaleSmall": 2,
                "TwoScaleLarge": 3, "TwoScaleMedium": 4, "TwoScaleSmall": 5}


class Analog(ctypes.Structure):
    _fields_ = [("rel_amplitude", ctypes.c_float), ("excess_noise_cv", ctypes.c_float),
                ("inter_pulse_distance", ctypes.c_float), ("pulse_width", ctypes.c_float),
                ("pw2_slow_step_ratio", ctypes.c_float), ("ipd2_slow_step_ratio", ctypes.c_float),
                ("base_label", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 3)]


class MovieInfo(ctypes.Structure):
    _fields_ = [("frame_rate", ctypes.c_float), ("photoelectron_sensitivity", ctypes.c_float),
                ("ref_snr", ctypes.c_float), ("shot_var_coeff", ctypes.c_float),
                ("analogs", Analog * NUM_ANALOGS), ("pedestal", ctypes.c_int16), ("pad", ctypes.c_int16)]


class Config(ctypes.Structure):
    _fields_ = [
        ("filter_type", ctypes.c_int32), ("mean_bias_adjust", ctypes.c_float),
        ("sigma_bias_adjust", ctypes.c_float), ("mean_ema_scale_strides", ctypes.c_float),
        ("sigma_ema_scale_strides", ctypes.c_float), ("jump_suppression", ctypes.c_float),
        ("bin_size_coeff", ctypes.c_float), ("bin_size_low_bound_coeff", ctypes.c_float),
        ("baseline_stat_min_frame_count", ctypes.c_uint32), ("fallback_baseline_sigma", ctypes.c_float),
        ("analog_mix_fraction_threshold", ctypes.c_float * 2), ("scale_snr_conf_tol", ctypes.c_float),
        ("em_iteration_limit", ctypes.c_uint32), ("gtest_stat_factor", ctypes.c_float),
        ("iterate_to_limit", ctypes.c_int32), ("min_analog_snr_thresh0", ctypes.c_float),
        ("min_analog_snr_thresh1", ctypes.c_float), ("pulse_amp_regularization", ctypes.c_float),
        ("model_update_method", ctypes.c_uint32), ("snr_drop_thresh", ctypes.c_float),
        ("success_confidence_thresh", ctypes.c_float), ("min_frames_for_estimate", ctypes.c_uint32),
        ("num_frames_pre_accum_stats", ctypes.c_uint32), ("confidence_half_life", ctypes.c_float * 2),
        ("viterbi_alpha", ctypes.c_float), ("viterbi_beta", ctypes.c_float), ("viterbi_gamma", ctypes.c_float),
        ("roi_upper_threshold", ctypes.c_float), ("roi_lower_threshold", ctypes.c_float),
        ("max_calls_per_zmw", ctypes.c_uint32), ("xsp_amp_thresh", ctypes.c_double),
        ("xsp_width_thresh", ctypes.c_float), ("sandwich_tolerance", ctypes.c_uint32),
        ("frames_per_hf_metric_block", ctypes.c_uint32), ("realtime_activity_labels", ctypes.c_int32),
        ("model_estimation_mode", ctypes.c_int32), ("static_baseline_mean", ctypes.c_float),
        ("static_baseline_variance", ctypes.c_float),
        ("learned_init", ctypes.c_int32),
        ("learned_conf_threshold", ctypes.c_float), ("learned_em_iteration_limit", ctypes.c_uint32),
        ("learned_conv_tol", ctypes.c_float), ("learned_gate_threshold", ctypes.c_float),
        ("learned_lane_min_confident", ctypes.c_uint32), ("movie", MovieInfo)]


class FilterParams(ctypes.Structure):
    _fields_ = [("num_stages", ctypes.c_int32), ("strides", ctypes.c_int32 * 3), ("widths", ctypes.c_int32 * 3),
                ("sigma_bias", ctypes.c_float), ("mean_bias", ctypes.c_float),
                ("aggregate_stride", ctypes.c_int32), ("latent_size", ctypes.c_int32)]


class Transition(ctypes.Structure):
    _fields_ = [("logp", (ctypes.c_float * NUM_STATES) * NUM_STATES), ("first_col", ctypes.c_int32 * NUM_STATES),
                ("seg_first", ctypes.c_int32 * NUM_STATES), ("seg_last", ctypes.c_int32 * NUM_STATES)]


class TracegenParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("baseline_mean", ctypes.c_float), ("baseline_sigma", ctypes.c_float),
                ("drift_per_chunk", ctypes.c_float), ("drift_chunk", ctypes.c_float), ("step_drift", ctypes.c_float),
                ("step_span", ctypes.c_float), ("empty_fraction", ctypes.c_float),
                ("amplitude", ctypes.c_float * NUM_ANALOGS), ("amplitude_cv", ctypes.c_float),
                ("mean_pulse_width", ctypes.c_float), ("mean_ipd", ctypes.c_float), ("snr_lo", ctypes.c_float),
                ("snr_hi", ctypes.c_float)]


def _f(n):
    return (np.float32, (n,))


STAT_ACC = np.dtype([("offset", _f(LANE)[0], (LANE,)), ("m0", np.float32, (LANE,)),
                     ("m1", np.float32, (LANE,)), ("m2", np.float32, (LANE,))], align=True)
AUTOCORR = np.dtype([("basic", STAT_ACC), ("lag_m2", np.float32, (LANE,)), ("fbuf", np.float32, (4, LANE)),
                     ("bbuf", np.float32, (4, LANE)), ("bidx", np.uint8, (2, LANE))], align=True)
BASELINER_STATS = np.dtype([("full", AUTOCORR), ("trace_min", np.int16, (LANE,)), ("trace_max", np.int16, (LANE,)),
                            ("baseline", STAT_ACC), ("raw_baseline_sum", np.float32, (LANE,)),
                            ("background", STAT_ACC)], align=True)
LANE_HIST = np.dtype([("low_bound", np.float32, (LANE,)), ("bin_size", np.float32, (LANE,)),
                      ("outlier_low", np.uint16, (LANE,)), ("outlier_high", np.uint16, (LANE,)),
                      ("bin_count", np.uint16, (NUM_BINS, LANE))], align=True)
MODE_H = np.dtype([("means", np.uint16, (LANE,)), ("vars", np.uint16, (LANE,)), ("weights", np.uint16, (LANE,))],
                  align=True)
LANE_MODEL = np.dtype([("analogs", MODE_H, (NUM_ANALOGS,)), ("baseline", MODE_H), ("confidence", np.uint16, (LANE,))],
                      align=True)
PULSE = np.dtype([("start", np.uint32), ("width", np.uint16), ("mean_signal", np.uint16),
                  ("mid_signal", np.uint16), ("max_signal", np.uint16), ("signal_m2", np.float32),
                  ("label", np.uint8), ("is_reject", np.uint8), ("pad", np.uint8, (2,))], align=True)
TRACEGEN_STATE = np.dtype([("slope", np.float32, (LANE,)), ("step_frame", np.float32, (LANE,)),
                           ("in_pulse", np.uint8, (LANE,)), ("analog", np.uint8, (LANE,)),
                           ("empty", np.uint8, (LANE,)), ("pad", np.uint8, (LANE,))], align=True)

_U16 = (np.uint16, (LANE,))
_A16 = (np.uint16, (NUM_ANALOGS, LANE))
_AF = (np.float32, (NUM_ANALOGS, LANE))
METRICS = np.dtype([("activity_label", np.uint8, (LANE,)), ("num_pulse_frames",) + _U16,
                    ("num_base_frames",) + _U16, ("num_sandwiches",) + _U16, ("num_half_sandwiches",) + _U16,
                    ("num_pulse_label_stutters",) + _U16, ("num_bases",) + _U16, ("num_pulses",) + _U16,
                    ("pk_mid_signal",) + _AF, ("bp_zvar",) + _AF, ("pk_zvar",) + _AF, ("pk_max",) + _AF,
                    ("num_pk_mid_frames",) + _A16, ("num_pk_mid_bases_by_analog",) + _A16,
                    ("num_bases_by_analog",) + _A16, ("num_pulses_by_analog",) + _A16,
                    ("frame_baseline_dws", np.float32, (LANE,)), ("frame_baseline_variance_dws", np.float32, (LANE,)),
                    ("num_frames_baseline",) + _U16, ("start_frame", np.uint32, (LANE,)), ("num_frames",) + _U16,
                    ("autocorrelation", np.float32, (LANE,)), ("pulse_detection_score", np.float32, (LANE,)),
                    ("pixel_checksum", np.int16, (LANE,))], align=True)
assert METRICS.itemsize == 137 * LANE


class PacketDesc(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("start_zmw", ctypes.c_uint32), ("start_frame", ctypes.c_uint32),
                ("blocks", ctypes.c_uint32), ("frames", ctypes.c_uint32), ("block_width", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class ActivityTree(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_int32), ("max_acceptable_hsw_rate", ctypes.c_float),
                ("hsw_curve", ctypes.c_float * 2), ("children_left", ctypes.c_void_p),
                ("children_right", ctypes.c_void_p), ("feature", ctypes.c_void_p), ("threshold", ctypes.c_void_p),
                ("value", ctypes.c_void_p)]


LAYOUT_ORDER = [("hwpa_config", ctypes.sizeof(Config)), ("hwpa_movie_info", ctypes.sizeof(MovieInfo)),
                ("hwpa_stat_acc", STAT_ACC.itemsize), ("hwpa_autocorr_acc", AUTOCORR.itemsize),
                ("hwpa_baseliner_stats", BASELINER_STATS.itemsize), ("hwpa_lane_hist", LANE_HIST.itemsize),
                ("hwpa_lane_model", LANE_MODEL.itemsize), ("hwpa_pulse", PULSE.itemsize),
                ("hwpa_tracegen_params", ctypes.sizeof(TracegenParams)),
                ("hwpa_tracegen_state", TRACEGEN_STATE.itemsize), ("hwpa_transition", ctypes.sizeof(Transition)),
                ("hwpa_basecalling_metrics", METRICS.itemsize)]


def half_to_float(a):
    """IEEE binary16 bit patterns -> float32 values."""
    return np.asarray(a, dtype=np.uint16).view(np.float16).astype(np.float32)
