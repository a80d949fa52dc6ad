/*
 * hwpa_types.h -- plain-old-data layouts shared by the HIP library, its
 * C-ABI callers and the CPU oracle.
 *
 * Every per-lane structure is structure-of-arrays over the 64 ZMWs of a lane
 * so that one wave64 (one thread per ZMW) touches 64 consecutive elements per
 * access.  Layouts that cross the analyzer boundary keep the reference's
 * field order and sizes so a caller can reinterpret them without copying:
 *   hwpa_stat_acc        <-> StatAccumState          (common/StatAccumState.h:43-58)
 *   hwpa_autocorr_acc    <-> AutocorrAccumState      (common/AutocorrAccumState.h:43-57)
 *   hwpa_baseliner_stats <-> BaselinerStatAccumState (dataTypes/BaselinerStatAccumState.h:15-46)
 *   hwpa_lane_hist       <-> LaneHistogram<float,u16>(dataTypes/PoolHistogram.h:17-48)
 *   hwpa_lane_model      <-> LaneModelParameters<PBHalf,64> (dataTypes/LaneDetectionModel.h:88-143)
 *   hwpa_pulse           <-> Pulse                   (dataTypes/Pulse.h:48-237)
 */
#ifndef HWPA_TYPES_H
#define HWPA_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HWPA_LANE_SIZE 64
#define HWPA_NUM_ANALOGS 4
#define HWPA_NUM_BINS 300
#define HWPA_NUM_STATES 13
#define HWPA_AUTOCORR_LAG 4
/* Viterbi stitching: 16 look-back frames + 2 ROI look-forward + 6 ROI stitch. */
#define HWPA_VITERBI_LOOKBACK 16
#define HWPA_ROI_LOOKFORWARD 2
#define HWPA_ROI_STITCH 6
#define HWPA_LATENT_FRAMES (HWPA_VITERBI_LOOKBACK + HWPA_ROI_LOOKFORWARD + HWPA_ROI_STITCH)
#define HWPA_MAX_FILTER_STAGES 3

/* Status codes returned by every C-ABI entry point. */
enum hwpa_status {
    HWPA_OK = 0,
    HWPA_ERR_INVALID_ARG = 1,
    HWPA_ERR_BAD_CONFIG = 2,
    HWPA_ERR_BAD_POOL = 3,
    HWPA_ERR_BAD_FRAME = 4,
    HWPA_ERR_HIP = 5,
    HWPA_ERR_NO_DEVICE = 6,
    HWPA_ERR_UNSUPPORTED = 7,
    HWPA_ERR_OOM = 8
};

/* ZMW status bits of the EM estimator (traceAnalysis/CoreDMEstimator.h:59-65). */
enum hwpa_zmw_status {
    HWPA_ZS_OK = 0,
    HWPA_ZS_NO_CONVERGE = 1,
    HWPA_ZS_INSUF_DATA = 2,
    HWPA_ZS_VLOW_SIGNAL = 4
};

/* Baseline filter tables (traceAnalysis/BaselinerParams.cpp:10-72). */
enum hwpa_filter_type {
    HWPA_FILTER_MULTISCALE_LARGE = 0,
    HWPA_FILTER_MULTISCALE_MEDIUM = 1,
    HWPA_FILTER_MULTISCALE_SMALL = 2,
    HWPA_FILTER_TWOSCALE_LARGE = 3,
    HWPA_FILTER_TWOSCALE_MEDIUM = 4,
    HWPA_FILTER_TWOSCALE_SMALL = 5
};

/* BasecallerAlgorithmConfig::ModelEstimationMode (dataTypes/configs/BasecallerAlgorithmConfig.h:64-65):
 * which BatchAnalyzer the analyzer follows (analyzer/BatchAnalyzer.cpp:120-373). */
enum hwpa_model_estimation_mode {
    HWPA_MODEL_FIXED = 0,    /* FixedModelBatchAnalyzer: StaticDetModelConfig model, no DME */
    HWPA_MODEL_INITIAL = 1,  /* SingleEstimateBatchAnalyzer: first estimate kept forever */
    HWPA_MODEL_DYNAMIC = 2   /* DynamicEstimateBatchAnalyzer (default) */
};

/* Raw trace encodings of TraceBatchVariant (dataTypes/TraceBatch.h:97-128); uint8 samples are
 * zero-extended to int16 before any arithmetic (dataTypes/BatchData.h:160-182). */
enum hwpa_trace_encoding {
    HWPA_ENCODING_INT16 = 0,
    HWPA_ENCODING_UINT8 = 1
};

/* A dense SensorPacket (pa-common DataSource::SensorPacket with BLOCK_LAYOUT_DENSE, as the
 * reference's BlockRepacker takes it, appModules/BlockRepacker.h:41-62): `blocks` blocks of
 * `frames` x `block_width` samples, [block][frame][zmw], ZMWs start_zmw .. and frames
 * start_frame .. of the chip.  data is a device pointer. */
typedef struct hwpa_packet_desc {
    const void* data;
    uint32_t start_zmw;
    uint32_t start_frame;
    uint32_t blocks;
    uint32_t frames;
    uint32_t block_width;
    uint32_t pad;
} hwpa_packet_desc;

typedef struct hwpa_stat_acc {
    float offset[HWPA_LANE_SIZE];
    float m0[HWPA_LANE_SIZE];
    float m1[HWPA_LANE_SIZE];
    float m2[HWPA_LANE_SIZE];
} hwpa_stat_acc;

typedef struct hwpa_autocorr_acc {
    hwpa_stat_acc basic;
    float lag_m2[HWPA_LANE_SIZE];
    float fbuf[HWPA_AUTOCORR_LAG][HWPA_LANE_SIZE];
    float bbuf[HWPA_AUTOCORR_LAG][HWPA_LANE_SIZE];
    uint8_t bidx[2][HWPA_LANE_SIZE];
} hwpa_autocorr_acc;

typedef struct hwpa_baseliner_stats {
    hwpa_autocorr_acc full;
    int16_t trace_min[HWPA_LANE_SIZE];
    int16_t trace_max[HWPA_LANE_SIZE];
    hwpa_stat_acc baseline;
    float raw_baseline_sum[HWPA_LANE_SIZE];
    hwpa_stat_acc background;
} hwpa_baseliner_stats;

/* BasecallingMetrics of one lane (dataTypes/BasecallingMetrics.h:53-80): 137
 * bytes per ZMW, same field order and types.  activity_label holds
 * HQRFPhysicalStates (0 EMPTY, 1 SINGLE, 2 MULTI, 3 LOW). */
typedef struct hwpa_basecalling_metrics {
    uint8_t activity_label[HWPA_LANE_SIZE];
    uint16_t num_pulse_frames[HWPA_LANE_SIZE];
    uint16_t num_base_frames[HWPA_LANE_SIZE];
    uint16_t num_sandwiches[HWPA_LANE_SIZE];
    uint16_t num_half_sandwiches[HWPA_LANE_SIZE];
    uint16_t num_pulse_label_stutters[HWPA_LANE_SIZE];
    uint16_t num_bases[HWPA_LANE_SIZE];
    uint16_t num_pulses[HWPA_LANE_SIZE];
    float pk_mid_signal[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    float bp_zvar[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    float pk_zvar[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    float pk_max[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    uint16_t num_pk_mid_frames[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    uint16_t num_pk_mid_bases_by_analog[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    uint16_t num_bases_by_analog[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    uint16_t num_pulses_by_analog[HWPA_NUM_ANALOGS][HWPA_LANE_SIZE];
    float frame_baseline_dws[HWPA_LANE_SIZE];
    float frame_baseline_variance_dws[HWPA_LANE_SIZE];
    uint16_t num_frames_baseline[HWPA_LANE_SIZE];
    uint32_t start_frame[HWPA_LANE_SIZE];
    uint16_t num_frames[HWPA_LANE_SIZE];
    float autocorrelation[HWPA_LANE_SIZE];
    float pulse_detection_score[HWPA_LANE_SIZE];
    int16_t pixel_checksum[HWPA_LANE_SIZE];
} hwpa_basecalling_metrics;

/* A trained activity-labeler decision tree (the shape of
 * ActivityLabeler::TrainedCart, dataTypes/TrainedCartParams.h:45-54).  Node
 * i tests feature[i] of the 11 block features (< 0: leaf of class value[i])
 * against threshold[i] and goes to children_left[i] if <=, else to
 * children_right[i]; children must have larger indices than their parent.
 * The trained values are the caller's data: pass them with
 * hwpa_metrics_set_activity_tree / hwpa_analyzer_set_activity_tree. */
typedef struct hwpa_activity_tree {
    int32_t num_nodes;
    float max_acceptable_hsw_rate;
    float hsw_curve[2];
    const int16_t* children_left;
    const int16_t* children_right;
    const int8_t* feature;
    const float* threshold;
    const int8_t* value;
} hwpa_activity_tree;

typedef struct hwpa_lane_hist {
    float low_bound[HWPA_LANE_SIZE];
    float bin_size[HWPA_LANE_SIZE];
    uint16_t outlier_low[HWPA_LANE_SIZE];
    uint16_t outlier_high[HWPA_LANE_SIZE];
    uint16_t bin_count[HWPA_NUM_BINS][HWPA_LANE_SIZE];
} hwpa_lane_hist;

/* One mixture component, IEEE binary16 bit patterns. */
typedef struct hwpa_mode_h {
    uint16_t means[HWPA_LANE_SIZE];
    uint16_t vars[HWPA_LANE_SIZE];
    uint16_t weights[HWPA_LANE_SIZE];
} hwpa_mode_h;

typedef struct hwpa_lane_model {
    hwpa_mode_h analogs[HWPA_NUM_ANALOGS];
    hwpa_mode_h baseline;
    uint16_t confidence[HWPA_LANE_SIZE];
} hwpa_lane_model;

/* 20-byte pulse record with the reference's fixed-point fields. */
typedef struct hwpa_pulse {
    uint32_t start;
    uint16_t width;
    uint16_t mean_signal;  /* value / 0.1, rounded half up */
    uint16_t mid_signal;
    uint16_t max_signal;
    float signal_m2;
    uint8_t label;         /* 0=A 1=C 2=G 3=T */
    uint8_t is_reject;
    uint8_t pad[2];
} hwpa_pulse;

/* Analog (dye) description (pacbio AnalogMode fields read by the path). */
typedef struct hwpa_analog {
    float rel_amplitude;
    float excess_noise_cv;
    float inter_pulse_distance;  /* seconds */
    float pulse_width;           /* seconds */
    float pw2_slow_step_ratio;
    float ipd2_slow_step_ratio;
    uint8_t base_label;          /* 'A','C','G','T' */
    uint8_t pad[3];
} hwpa_analog;

typedef struct hwpa_movie_info {
    float frame_rate;
    float photoelectron_sensitivity;
    float ref_snr;
    float shot_var_coeff;        /* sum(X^2 * PSF); 1 for an empty PSF */
    hwpa_analog analogs[HWPA_NUM_ANALOGS];  /* descending relative amplitude */
    int16_t pedestal;
    int16_t pad;
} hwpa_movie_info;

/* Knobs of the analysis path with the reference defaults
 * (dataTypes/configs/Basecaller*Config.h). */
typedef struct hwpa_config {
    /* BasecallerBaselinerConfig */
    int32_t filter_type;
    float mean_bias_adjust;
    float sigma_bias_adjust;
    float mean_ema_scale_strides;
    float sigma_ema_scale_strides;
    float jump_suppression;
    /* BasecallerTraceHistogramConfig */
    float bin_size_coeff;
    float bin_size_low_bound_coeff;
    uint32_t baseline_stat_min_frame_count;
    float fallback_baseline_sigma;
    /* BasecallerDmeConfig */
    float analog_mix_fraction_threshold[2];
    float scale_snr_conf_tol;
    uint32_t em_iteration_limit;
    float gtest_stat_factor;
    int32_t iterate_to_limit;
    float min_analog_snr_thresh0;
    float min_analog_snr_thresh1;
    float pulse_amp_regularization;
    uint32_t model_update_method;
    float snr_drop_thresh;
    float success_confidence_thresh;
    uint32_t min_frames_for_estimate;
    uint32_t num_frames_pre_accum_stats;
    float confidence_half_life[2];
    /* BasecallerFrameLabelerConfig */
    float viterbi_alpha;
    float viterbi_beta;
    float viterbi_gamma;
    float roi_upper_threshold;
    float roi_lower_threshold;
    /* BasecallerPulseAccumConfig */
    uint32_t max_calls_per_zmw;
    double xsp_amp_thresh;
    float xsp_width_thresh;
    /* BasecallerMetricsConfig */
    uint32_t sandwich_tolerance;
    uint32_t frames_per_hf_metric_block;
    int32_t realtime_activity_labels;
    /* BasecallerAlgorithmConfig::modelEstimationMode and StaticDetModelConfig
     * (dataTypes/configs/StaticDetModelConfig.h:43-44) */
    int32_t model_estimation_mode;
    float static_baseline_mean;
    float static_baseline_variance;
    /* Learned initialiser (no reference counterpart): 0 = plain EM. */
    int32_t learned_init;
    float learned_conf_threshold;
    uint32_t learned_em_iteration_limit;
    float learned_conv_tol;      /* relative log-likelihood tolerance of the capped EM */
    float learned_gate_threshold; /* network confidence below which the capped EM is skipped (0: never) */
    uint32_t learned_lane_min_confident; /* a lane with fewer ZMWs above the gate sends all 64 to the plain EM */
    hwpa_movie_info movie;
} hwpa_config;

/* Derived filter table (traceAnalysis/BaselinerParams.h:14-77). */
typedef struct hwpa_filter_params {
    int32_t num_stages;
    int32_t strides[HWPA_MAX_FILTER_STAGES];
    int32_t widths[HWPA_MAX_FILTER_STAGES];
    float sigma_bias;
    float mean_bias;
    int32_t aggregate_stride;
    int32_t latent_size;
} hwpa_filter_params;

/* Sparse 13x13 log transition matrix: row = next state, each row has one
 * leading column and one contiguous column segment
 * (traceAnalysis/SubframeScorer.h:54-83, SparseMatrix.h). */
typedef struct hwpa_transition {
    float logp[HWPA_NUM_STATES][HWPA_NUM_STATES];  /* dense copy, 0 where absent */
    int32_t first_col[HWPA_NUM_STATES];
    int32_t seg_first[HWPA_NUM_STATES];
    int32_t seg_last[HWPA_NUM_STATES];              /* exclusive */
} hwpa_transition;

/* Synthetic trace generator description (seeded, counter-based). */
typedef struct hwpa_tracegen_params {
    uint64_t seed;
    float baseline_mean;
    float baseline_sigma;
    float drift_per_chunk;       /* slope ~ U[-d, +d] counts per drift_chunk frames */
    float drift_chunk;
    float step_drift;            /* +step at a uniformly drawn frame in [0, step_span) */
    float step_span;
    float empty_fraction;        /* ZMWs with baseline only */
    float amplitude[HWPA_NUM_ANALOGS];
    float amplitude_cv;          /* multiplicative per-frame noise */
    float mean_pulse_width;      /* frames, geometric */
    float mean_ipd;              /* frames, geometric */
    float snr_lo, snr_hi;        /* if snr_hi > 0: per-ZMW amplitude scale so that
                                    amplitude[0] / baseline_sigma ~ U[snr_lo, snr_hi] */
} hwpa_tracegen_params;

/* Per-ZMW generator state carried between chunks. */
typedef struct hwpa_tracegen_state {
    float slope[HWPA_LANE_SIZE];
    float step_frame[HWPA_LANE_SIZE];
    uint8_t in_pulse[HWPA_LANE_SIZE];
    uint8_t analog[HWPA_LANE_SIZE];
    uint8_t empty[HWPA_LANE_SIZE];
    uint8_t pad[HWPA_LANE_SIZE];
} hwpa_tracegen_state;

#ifdef __cplusplus
}
#endif

#endif /* HWPA_TYPES_H */
