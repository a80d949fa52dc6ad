/*
 * hwpa.h -- C-ABI of the MI355X trace-analysis library (libhwpa.so).
 *
 * The analyzer entry points replace the reference's per-pool analyzer
 * boundary; the per-stage entry points replace the Device* strategy
 * classes behind it and exist so each stage can be checked on its own.
 * All functions return an hwpa_status code and never throw; the message of
 * the last failure on the calling thread is available from hwpa_last_error().
 *
 * Pointers named d_* are device pointers (HBM) in the [lane][frame][zmw]
 * layout of include/hwpa_types.h; `stream` is a hipStream_t (NULL = the
 * default stream).  Stage functions enqueue work and return without
 * synchronising.
 */
#ifndef HWPA_H
#define HWPA_H

#include <stddef.h>
#include <stdint.h>

#include "hwpa_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors / configuration ------------------------------------------- */

/* Thread-local message of the last failed call ("" if none). */
const char* hwpa_last_error(void);

/* Fills *cfg with the reference defaults (dataTypes/configs/Basecaller*Config.h,
 * CoreDMEstimator.cpp:36-87) and a default movie description. */
int32_t hwpa_config_default(hwpa_config* cfg);

/* Applies a JSON object of overrides, keyed like the reference's
 * BasecallerAlgorithmConfig (e.g. {"baselinerConfig":{"SigmaEmaScaleStrides":0}}).
 * Replaces PBConfig JSON loading (applications/SmrtBasecaller.cpp:985-1029). */
int32_t hwpa_config_load_json(hwpa_config* cfg, const char* json);

/* Checks knob ranges (ValidateConfig<...> specialisations in dataTypes/configs). */
int32_t hwpa_config_validate(const hwpa_config* cfg);

/* sizeof() of the public structures, in this order: hwpa_config,
 * hwpa_movie_info, hwpa_stat_acc, hwpa_autocorr_acc, hwpa_baseliner_stats,
 * hwpa_lane_hist, hwpa_lane_model, hwpa_pulse, hwpa_tracegen_params,
 * hwpa_tracegen_state, hwpa_transition, hwpa_basecalling_metrics.  Returns how
 * many were written. */
int32_t hwpa_layout_sizes(int64_t* out, int32_t n);

/* Filter table lookup (traceAnalysis/BaselinerParams.cpp:10-72). */
int32_t hwpa_filter_params_get(int32_t filter_type, hwpa_filter_params* out);

/* 13-state log transition matrix (traceAnalysis/SubframeScorer.cpp:35-223). */
int32_t hwpa_transition_matrix(const hwpa_config* cfg, hwpa_transition* out);

/* ---- per-pool analyzer (analyzer/BatchAnalyzer.h:51-106) -------------- */

typedef struct hwpa_analyzer hwpa_analyzer;
typedef struct hwpa_initializer hwpa_initializer;  /* learned EM initialiser, below */
typedef struct hwpa_metrics hwpa_metrics;          /* HF metrics stage, below */

/* BatchAnalyzer ctor (analyzer/BatchAnalyzer.cpp:71-118) for the analyzer that
 * cfg->model_estimation_mode selects (appModules/Basecaller.h:103-116): Fixed (static model),
 * Initial (single estimate) or Dynamic (default). */
int32_t hwpa_analyzer_create(uint32_t pool_id, uint32_t lanes, uint32_t frames_per_chunk,
                             const hwpa_config* cfg, hwpa_analyzer** out);
int32_t hwpa_analyzer_destroy(hwpa_analyzer* a);

/* BatchAnalyzer::operator()(const TraceBatchVariant&) (analyzer/BatchAnalyzer.cpp:149-170).
 * traces: int16 [lanes][frames][64]; traces_on_device selects d_ vs host pointer.
 * Errors: HWPA_ERR_BAD_POOL (wrong pool), HWPA_ERR_BAD_FRAME (out of order). */
int32_t hwpa_analyzer_process(hwpa_analyzer* a, const int16_t* traces, int32_t traces_on_device,
                              uint32_t pool_id, int32_t first_frame);
/* Same for either TraceBatchVariant alternative (dataTypes/TraceBatch.h:97-128): traces are
 * [lanes][frames][64] samples of HWPA_ENCODING_INT16 or HWPA_ENCODING_UINT8 (zero-extended,
 * dataTypes/BatchData.h:160-182; the movie pedestal is subtracted as for int16). */
int32_t hwpa_analyzer_process_encoded(hwpa_analyzer* a, const void* traces, int32_t encoding,
                                      int32_t traces_on_device, uint32_t pool_id, int32_t first_frame);

/* Blocks until the analyzer's stream is idle. */
int32_t hwpa_analyzer_sync(hwpa_analyzer* a);

/* Result accessors for the last process call (device pointers owned by the
 * analyzer, valid until the next process call).  flags: [0] pulses emitted,
 * [1] full estimation ran, [2] frame labels ran, [3] a metrics block ended. */
int32_t hwpa_result_flags(const hwpa_analyzer* a, int32_t flags[4]);
int32_t hwpa_result_pulses(const hwpa_analyzer* a, const hwpa_pulse** d_pulses,
                           const uint32_t** d_counts, const uint32_t** d_overflows,
                           uint32_t* max_calls_per_zmw);
int32_t hwpa_result_labels(const hwpa_analyzer* a, const int16_t** d_labels);
int32_t hwpa_result_baselined(const hwpa_analyzer* a, const int16_t** d_traces);
int32_t hwpa_result_models(const hwpa_analyzer* a, const hwpa_lane_model** d_models);
int32_t hwpa_result_baseliner_stats(const hwpa_analyzer* a, const hwpa_baseliner_stats** d_stats);
int32_t hwpa_result_pulse_baseline_stats(const hwpa_analyzer* a, const hwpa_stat_acc** d_stats);
int32_t hwpa_result_viterbi_score(const hwpa_analyzer* a, const float** d_score);
/* BasecallingMetrics of the block that ended with the last chunk (flags[3]);
 * BatchResult::metrics of the reference (AN/BatchAnalyzer.cpp:366-372). */
int32_t hwpa_result_metrics(const hwpa_analyzer* a, const hwpa_basecalling_metrics** d_metrics);
/* Trained activity-labeler tree for the analyzer's metrics (copied; NULL clears it:
 * activity labels then stay EMPTY, as with realtime_activity_labels = 0). */
int32_t hwpa_analyzer_set_activity_tree(hwpa_analyzer* a, const hwpa_activity_tree* tree);

/* Per-stage device time of the last process call in milliseconds, measured
 * with HIP events on the analyzer's stream when profiling is enabled:
 * [0] baseline [1] aggregate [2] histogram [3] dme [4] labels [5] pulses
 * [6] metrics. */
int32_t hwpa_analyzer_set_profiling(hwpa_analyzer* a, int32_t enable);
int32_t hwpa_analyzer_stage_ms(const hwpa_analyzer* a, float ms[8]);
/* Number of EM iterations the last estimation executed (max over lanes). */
int32_t hwpa_analyzer_em_iterations(const hwpa_analyzer* a, int32_t* iters);
/* Stream used by the analyzer (hipStream_t). */
int32_t hwpa_analyzer_stream(const hwpa_analyzer* a, void** stream);

/* ---- batched pools (no reference counterpart: the reference runs one
 * analyzer per pool on per-thread streams, Basecaller.h:160,246-250) -------
 * n_pools analyzers (pool_ids[i], lanes[i], frames[i], cfgs[i]), each with its
 * own state and stream.  hwpa_batch_process runs one chunk of every pool:
 * d_traces[i] is pool i's device chunk, first_frames[i] its first frame.  By
 * default the whole step is captured into one hipGraph (pools concurrent) and
 * launched; executable graphs are cached per combination of the pools' stage
 * sequences and updated in place.  Per-pool results come from the pool's
 * analyzer (hwpa_batch_pool) after hwpa_batch_sync.  Stage profiling must be
 * off while graphs are on. */
typedef struct hwpa_batch hwpa_batch;
int32_t hwpa_batch_create(uint32_t n_pools, const uint32_t* pool_ids, const uint32_t* lanes, const uint32_t* frames,
                          const hwpa_config* cfgs, hwpa_batch** out);
int32_t hwpa_batch_destroy(hwpa_batch* b);
int32_t hwpa_batch_pool(hwpa_batch* b, uint32_t index, hwpa_analyzer** a);
int32_t hwpa_batch_set_graph(hwpa_batch* b, int32_t enable);
int32_t hwpa_batch_set_initializer(hwpa_batch* b, hwpa_initializer* m);
int32_t hwpa_batch_process(hwpa_batch* b, const int16_t* const* d_traces, const int32_t* first_frames);
/* Per-pool encodings (HWPA_ENCODING_*); NULL means int16 for every pool. */
int32_t hwpa_batch_process_encoded(hwpa_batch* b, const void* const* d_traces, const int32_t* encodings,
                                   const int32_t* first_frames);
int32_t hwpa_batch_sync(hwpa_batch* b);
/* [0] steps run, [1] graph instantiations, [2] in-place graph updates. */
int32_t hwpa_batch_graph_stats(const hwpa_batch* b, uint64_t stats[3]);

/* Autotunes every pool's shape (hwpa_autotune on its lanes, frames, config). */
int32_t hwpa_batch_autotune(hwpa_batch* b, int32_t repeats);

/* ---- kernel-shape autotuner (no reference counterpart) -------------------
 * Tunable kernels run in variants that give the same bits with different
 * shapes: kernel 0 = histogram (variants 0-8: one wave per lane, 8, 16 or 32
 * frames per prefetched batch x 128, 160 or 192 bins kept in LDS; 9-15: 2, 4 or
 * 8 waves per lane sharing its LDS counters; 16-21: all 300 bins in LDS, 8 or 16
 * frames per batch x 4, 8 or 2 waves per lane), kernel 1 = baseline filter (1, 2 or 4 lanes per
 * workgroup, then the same with two groups of raw frames in flight), kernel 2 = Viterbi labeler (16 or 8 frames per prefetched
 * batch, then the same with registers held to 4 waves per SIMD), kernel 3 = pulse accumulator (frames
 * per batch x batches in the load ring x record queue: 8x1, 8x2, 16x2 writing records directly; 8x1,
 * 8x2, 4x2, 4x1, 4x4, 8x3, 4x8, 4x12, 8x4, 2x16 writing them through an LDS queue).
 * The launchers look up the variant for their (lanes, frames);
 * unset shapes use variant 0.  hwpa_autotune times every variant on
 * synthetic traces of the shape (`repeats` timed runs each, HIP events) and
 * keeps the fastest; ms (optional, ms_len entries) receives the mean times
 * of the variants of kernel 0, then 1, 2 and 3. */
int32_t hwpa_tune_variants(int32_t kernel, int32_t* count);
int32_t hwpa_tune_set(int32_t kernel, uint32_t lanes, uint32_t frames, int32_t variant);
int32_t hwpa_tune_get(int32_t kernel, uint32_t lanes, uint32_t frames, int32_t* variant);
int32_t hwpa_autotune(const hwpa_config* cfg, uint32_t lanes, uint32_t frames, int32_t repeats, float* ms,
                      int32_t ms_len);

/* ---- HF metrics (HFMetricsFilterHost, traceAnalysis/HFMetricsFilterHost.cpp:37-116) ----
 * Per-ZMW basecalling metrics over blocks of frames_per_hf_metric_block
 * frames.  hwpa_metrics_process takes one chunk's pulse batch (device
 * pointers; last_frame = the batch's LastFrame, i.e. the chunk's end minus
 * the 24 latent frames), the chunk's baseliner statistics, Viterbi scores,
 * pulse-detector baseline statistics and current models; when the block ends
 * it writes d_out[lanes] and sets *produced. */
int32_t hwpa_metrics_create(const hwpa_config* cfg, uint32_t lanes, hwpa_metrics** out);
int32_t hwpa_metrics_destroy(hwpa_metrics* m);
int32_t hwpa_metrics_set_activity_tree(hwpa_metrics* m, const hwpa_activity_tree* tree);
int32_t hwpa_metrics_process(hwpa_metrics* m, const hwpa_pulse* d_pulses, const uint32_t* d_counts,
                             uint32_t max_calls, uint32_t frames, int32_t last_frame,
                             const hwpa_baseliner_stats* d_bl_stats, const float* d_viterbi_score,
                             const hwpa_stat_acc* d_pd_stats, const hwpa_lane_model* d_models,
                             hwpa_basecalling_metrics* d_out, int32_t* produced, void* stream);

/* ---- BAZ packet encoding of pulses (bazio/encoding/ObjectToBaz.h, FieldTransforms.h,
 * FieldSerializers.h; layouts basecaller/dataTypes/PulseGroups.h:41-94) --------------
 * One encoder per set of ZMWs keeps each ZMW's StartFrame delta state across chunks, as the
 * reference keeps one serializer per ZMW (appModules/PrelimHQFilter.cpp:148-160).
 * internal_mode 0 = ProductionPulses (Label, PulseWidth, StartFrame; bases only), 1 =
 * InternalPulses (plus Pkmax, Pkmid, Pkmean, Pkvar, IsBase; every pulse).
 * hwpa_baz_encode serialises one chunk (d_pulses [zmws][max_calls], d_counts [zmws], as the
 * analyzer produces them) into one stream ordered by ZMW: d_offsets [zmws + 1] receives each
 * ZMW's byte offset and the total at [zmws], d_events [zmws] the pulses written per ZMW.  If
 * the total exceeds `capacity`, nothing is written to d_out and the delta state is unchanged
 * (a size query: capacity 0). */
typedef struct hwpa_baz_encoder hwpa_baz_encoder;
/* zmws: a multiple of 64 (whole lanes); max_calls at most 2^20. */
int32_t hwpa_baz_encoder_create(uint32_t zmws, int32_t internal_mode, hwpa_baz_encoder** out);
int32_t hwpa_baz_encoder_destroy(hwpa_baz_encoder* e);
int32_t hwpa_baz_encoder_reset(hwpa_baz_encoder* e, void* stream);
int32_t hwpa_baz_encode(hwpa_baz_encoder* e, const hwpa_pulse* d_pulses, const uint32_t* d_counts,
                        uint32_t max_calls, uint8_t* d_out, uint64_t capacity, uint64_t* d_offsets,
                        uint32_t* d_events, void* stream);

/* ---- block repacker (appModules/BlockRepacker.cpp:200-330, BatchFiller::Process) -----
 * Copies the parts of n dense packets that overlap a batch [batch_lanes][batch_frames][64]
 * starting at (batch_first_zmw, batch_first_frame) into it, samples of HWPA_ENCODING_INT16
 * or HWPA_ENCODING_UINT8 kept as they are.  Packets need 16-byte aligned data and block
 * widths / start ZMWs of whole 16-byte pieces; the batch starts on a lane.  One launch per
 * call.  Descriptor slots rotate over 4 calls and each slot's reuse waits (on the new call's
 * stream) for the kernel that last read it, so calls may go to different streams. */
typedef struct hwpa_repacker hwpa_repacker;
int32_t hwpa_repacker_create(uint32_t max_packets, hwpa_repacker** out);
int32_t hwpa_repacker_destroy(hwpa_repacker* r);
int32_t hwpa_repack(hwpa_repacker* r, const hwpa_packet_desc* packets, uint32_t n, int32_t encoding, void* d_batch,
                    uint32_t batch_first_zmw, uint32_t batch_first_frame, uint32_t batch_lanes,
                    uint32_t batch_frames, void* stream);

/* ---- stage entry points (Device* strategies) ----------------------------- */

/* Baseliner: DeviceMultiScaleBaseliner / HostMultiScaleBaseliner::FilterBaseline
 * (traceAnalysis/HostMultiScaleBaseliner.cpp:57-176).  State is per lane. */
typedef struct hwpa_baseliner hwpa_baseliner;
int32_t hwpa_baseliner_create(const hwpa_config* cfg, uint32_t lanes, hwpa_baseliner** out);
int32_t hwpa_baseliner_destroy(hwpa_baseliner* b);
int32_t hwpa_baseliner_run(hwpa_baseliner* b, const int16_t* d_raw, uint32_t frames,
                           int16_t* d_out, hwpa_baseliner_stats* d_stats, void* stream);
/* The same on HWPA_ENCODING_INT16 or HWPA_ENCODING_UINT8 samples (DeviceMultiScaleBaseliner's
 * uint8 / int16 filters, traceAnalysis/DeviceMultiScaleBaseliner.cu:84-101). */
int32_t hwpa_baseliner_run_encoded(hwpa_baseliner* b, const void* d_raw, int32_t encoding, uint32_t frames,
                                   int16_t* d_out, hwpa_baseliner_stats* d_stats, void* stream);

/* Stats aggregator: BaselineStatsAggregatorDevice::AddMetricsImpl / ResetImpl
 * (traceAnalysis/BaselineStatsAggregatorHost.cpp:40-71). */
int32_t hwpa_stats_reset(hwpa_baseliner_stats* d_agg, uint32_t lanes, void* stream);
int32_t hwpa_stats_merge(hwpa_baseliner_stats* d_agg, const hwpa_baseliner_stats* d_chunk,
                         uint32_t lanes, void* stream);

/* Histogram: TraceHistogramAccumHost::{ResetImpl,AddBlock}
 * (traceAnalysis/TraceHistogramAccumHost.cpp:92-169).  The edge-frame
 * classifier state (d_edge, hwpa_edge_state_bytes() per lane) persists
 * across resets. */
size_t hwpa_edge_state_bytes(void);
int32_t hwpa_edge_state_init(void* d_edge, uint32_t lanes, void* stream);
int32_t hwpa_hist_reset(const hwpa_config* cfg, hwpa_lane_hist* d_hist,
                        const hwpa_baseliner_stats* d_stats, uint32_t lanes, void* stream);
int32_t hwpa_hist_add(hwpa_lane_hist* d_hist, void* d_edge, const int16_t* d_traces,
                      uint32_t frames, const hwpa_lane_model* d_models, uint32_t lanes, void* stream);

/* DME: DmeEmHost::InitDetectionModels / EstimateImpl
 * (traceAnalysis/DmeEmHost.cpp:131-159,855-925).  Intervals are [lo, hi). */
int32_t hwpa_dme_init_models(const hwpa_config* cfg, const hwpa_baseliner_stats* d_stats,
                             hwpa_lane_model* d_models, uint32_t lanes, void* stream);
int32_t hwpa_dme_estimate(const hwpa_config* cfg, const hwpa_lane_hist* d_hist,
                          const hwpa_baseliner_stats* d_stats, hwpa_lane_model* d_models,
                          const int32_t model_fi[2], const int32_t est_fi[2], uint32_t lanes,
                          int32_t* d_iters, void* stream);

/* As hwpa_dme_estimate, and also writes each ZMW's raw EM result (before the
 * confidence-weighted update) to d_diag, fp32 [lanes][HWPA_EM_DIAG_ROWS][64]:
 * rows 0-4 mixture weights, 5-9 means, 10-14 variances (baseline first),
 * 15 confidence, 16 converged (0/1), 17 status bits.  No reference
 * counterpart (the reference's DmeDiagnostics are not exported per ZMW);
 * these are the training targets of the learned initialiser. */
int32_t hwpa_dme_estimate_diag(const hwpa_config* cfg, const hwpa_lane_hist* d_hist,
                               const hwpa_baseliner_stats* d_stats, hwpa_lane_model* d_models,
                               const int32_t model_fi[2], const int32_t est_fi[2], uint32_t lanes,
                               int32_t* d_iters, float* d_diag, void* stream);

/* Frame labeler: FrameLabelerHost::Process (traceAnalysis/FrameLabelerHost.cpp:312-612).
 * d_latent_out receives the 24 trace frames preceding this chunk (the
 * LabelsBatch latent trace). */
typedef struct hwpa_labeler hwpa_labeler;
int32_t hwpa_labeler_create(const hwpa_config* cfg, uint32_t lanes, uint32_t max_frames,
                            hwpa_labeler** out);
int32_t hwpa_labeler_destroy(hwpa_labeler* l);
int32_t hwpa_labeler_run(hwpa_labeler* l, const int16_t* d_traces, uint32_t frames,
                         const hwpa_lane_model* d_models, int16_t* d_labels,
                         int16_t* d_latent_out, float* d_score, void* stream);

/* Pulse accumulator: HostPulseAccumulator::Process
 * (traceAnalysis/HostPulseAccumulator.cpp:48-127). */
typedef struct hwpa_pulser hwpa_pulser;
int32_t hwpa_pulser_create(const hwpa_config* cfg, uint32_t lanes, hwpa_pulser** out);
int32_t hwpa_pulser_destroy(hwpa_pulser* p);
int32_t hwpa_pulser_run(hwpa_pulser* p, const int16_t* d_labels, const int16_t* d_latent,
                        const int16_t* d_traces, uint32_t frames, int32_t first_frame,
                        const hwpa_lane_model* d_models, hwpa_pulse* d_pulses, uint32_t* d_counts,
                        uint32_t* d_overflows, hwpa_stat_acc* d_bl_stats, void* stream);

/* ---- learned initialiser of the EM (no reference counterpart) -----------
 * A 128 -> 256 -> 256 -> 12 perceptron with ReLU hidden layers maps each
 * ZMW's histogram (re-binned to 128 bins: bin i = bins 2i and 2i+1 of the
 * 300, scaled by 16 / their total) to a start model for the EM: baseline
 * mean, log variance and four analog means in units of u = 2 binSize from the
 * histogram's lower bound (through the output map below), five
 * mixture-weight logits, and a confidence logit: the network's estimate that
 * the ZMW's EM converges with status OK (trained on that outcome of the plain
 * EM, a logistic loss).  ZMWs whose confidence is below
 * cfg->learned_gate_threshold skip the capped EM and go straight to the plain EM.  Matrices are bfloat16 for the matrix cores with fp32 accumulation;
 * an fp32 master copy and Adam moments live on the device.  The start replaces
 * the EM's (DmeEmHost.cpp:236-249 onwards); the EM then runs at most
 * cfg->learned_em_iteration_limit (default 5) iterations with the convergence test
 * delta LL < cfg->learned_conv_tol * |LL| (the plain EM uses 1e-4), and ZMWs
 * that did not converge (or whose model confidence is below
 * cfg->learned_conf_threshold, default 0) are re-estimated by the plain EM. */
#define HWPA_MLP_IN 128
#define HWPA_MLP_HIDDEN 256
#define HWPA_MLP_OUT 12
/* W1[256][128] b1[256] W2[256][256] b2[256] W3[12][256] b3[12] */
#define HWPA_MLP_PARAMS (HWPA_MLP_HIDDEN * HWPA_MLP_IN + HWPA_MLP_HIDDEN + HWPA_MLP_HIDDEN * HWPA_MLP_HIDDEN + \
                         HWPA_MLP_HIDDEN + HWPA_MLP_OUT * HWPA_MLP_HIDDEN + HWPA_MLP_OUT)
#define HWPA_EM_DIAG_ROWS 18

/* He-uniform weights from a 64-bit seed, zero biases. */
int32_t hwpa_initializer_create(uint64_t seed, hwpa_initializer** out);
int32_t hwpa_initializer_destroy(hwpa_initializer* m);
/* fp32 parameters in the order above (n must equal HWPA_MLP_PARAMS);
 * setting them also resets the optimiser state. */
int32_t hwpa_initializer_get_params(const hwpa_initializer* m, float* host_params, size_t n);
int32_t hwpa_initializer_set_params(hwpa_initializer* m, const float* host_params, size_t n);
/* Output map: network output j < 6 stands for offset[j] + scale[j] * y[j] in
 * the raw target units of hwpa_mlp_targets; weight[0..5] weigh their squared
 * errors in the loss, weight[6] the cross-entropy of the mixture weights,
 * weight[7] the logistic loss of the confidence output.
 * Default: offset 0, scale 1, weights 1.  get returns {offset, scale, weight}. */
int32_t hwpa_initializer_set_output_map(hwpa_initializer* m, const float offset[6], const float scale[6],
                                        const float weight[8]);
int32_t hwpa_initializer_get_output_map(const hwpa_initializer* m, float out[20]);
/* Network input from lane histograms: bf16 [lanes*64][128] (16-byte aligned). */
int32_t hwpa_mlp_features(const hwpa_lane_hist* d_hist, uint32_t lanes, uint16_t* d_x, void* stream);
/* Training targets from hwpa_dme_estimate_diag output: fp32 [lanes*64][12] in
 * raw units, u = 2 binSize from the lower bound lo: (mu_bl - lo)/u,
 * log(var_bl/u^2), (mu_a - lo)/u for the four analogs, the five mixture
 * weights, and 1 or 0 for the confidence output; d_valid u8 [lanes*64] = that
 * last target: converged, status OK, confidence >= min_confidence. */
int32_t hwpa_mlp_targets(const hwpa_lane_hist* d_hist, const float* d_diag, uint32_t lanes,
                         float min_confidence, float* d_t, uint8_t* d_valid, void* stream);
/* Network outputs fp32 [rows][11]. */
int32_t hwpa_mlp_forward(hwpa_initializer* m, const uint16_t* d_x, uint32_t rows, float* d_y, void* stream);
/* Network outputs fp32 [lanes*64][11] straight from lane histograms: equal to
 * hwpa_mlp_features followed by hwpa_mlp_forward, without the input matrix. */
int32_t hwpa_mlp_forward_hist(hwpa_initializer* m, const hwpa_lane_hist* d_hist, uint32_t lanes, float* d_y,
                              void* stream);
/* Mean loss over a batch (squared error on outputs 0-5, softmax cross-entropy
 * on 6-10, logistic loss of the confidence output 11 against target 11) and
 * its gradient, fp32 [HWPA_MLP_PARAMS] in parameter order.  Samples with
 * d_valid[i] == 0 (d_valid may be NULL) contribute only the confidence term;
 * the mean is taken over the whole batch. */
int32_t hwpa_mlp_gradients(hwpa_initializer* m, const uint16_t* d_x, const float* d_t, const uint8_t* d_valid,
                           uint32_t batch, float* d_grads, float* d_loss, void* stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on a batch; d_loss (optional)
 * receives the batch loss before the step. */
int32_t hwpa_mlp_train_step(hwpa_initializer* m, const uint16_t* d_x, const float* d_t, const uint8_t* d_valid,
                            uint32_t batch, float lr, float* d_loss, void* stream);
/* Learned-start estimation of one pool: *fallback_count (host, optional)
 * receives the number of ZMWs re-estimated by the plain EM.  Synchronous. */
int32_t hwpa_dme_estimate_learned(const hwpa_config* cfg, hwpa_initializer* m, const hwpa_lane_hist* d_hist,
                                  const hwpa_baseliner_stats* d_stats, hwpa_lane_model* d_models,
                                  const int32_t model_fi[2], const int32_t est_fi[2], uint32_t lanes,
                                  int32_t* d_iters, uint32_t* fallback_count, void* stream);
/* As hwpa_dme_estimate_learned, but allocates nothing: its workspace (network
 * outputs, fallback flags, per-lane counts and offsets, the compacted
 * fallback lanes and their histograms, statistics and models) lives in a
 * fixed device array of the library, sized for HWPA_NOALLOC_MAX_LANES lanes
 * and held on each GPU that calls it until the process ends.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  More lanes than the maximum -> HWPA_ERR_INVALID_ARG.  Every
 * call on one GPU uses the same array, so calls that overlap on one GPU (two
 * streams or two threads at once) race.  Synchronous, like the variant above. */
#define HWPA_NOALLOC_MAX_LANES 256
int32_t hwpa_dme_estimate_learned_noalloc(const hwpa_config* cfg, hwpa_initializer* m,
                                          const hwpa_lane_hist* d_hist, const hwpa_baseliner_stats* d_stats,
                                          hwpa_lane_model* d_models, const int32_t model_fi[2],
                                          const int32_t est_fi[2], uint32_t lanes, int32_t* d_iters,
                                          uint32_t* fallback_count, void* stream);
/* Attaches an initialiser to an analyzer (NULL detaches); with
 * cfg->learned_init set, its estimations use the learned start.  The
 * initialiser must outlive the analyzer or be detached first. */
int32_t hwpa_analyzer_set_initializer(hwpa_analyzer* a, hwpa_initializer* m);
/* Fallback count of the last learned estimation (synchronises the stream). */
int32_t hwpa_analyzer_fallback_count(const hwpa_analyzer* a, uint32_t* count);
/* Arms training-pair capture: every later plain estimation also writes the
 * network input of each ZMW (d_x, bf16 [lanes*64][128], 16-byte aligned),
 * its target from the EM result (d_t, fp32 [lanes*64][11]) and d_valid
 * (u8 [lanes*64], see hwpa_mlp_targets with min_confidence), overwriting the
 * previous capture.  NULL d_x disarms. */
int32_t hwpa_analyzer_capture_training(hwpa_analyzer* a, uint16_t* d_x, float* d_t, uint8_t* d_valid,
                                       float min_confidence);

/* Device memory helpers for callers without their own allocator (tests,
 * benchmarks, language bindings).  Copies are synchronous. */
int32_t hwpa_device_count(int32_t* count);
int32_t hwpa_device_sync(void);
int32_t hwpa_dev_alloc(void** d_ptr, size_t bytes);
int32_t hwpa_dev_free(void* d_ptr);
/* Page-locked host memory (hipHostMalloc): host chunks in it reach the device
 * by DMA, so hwpa_analyzer_process with on_device = 0 overlaps their copy with
 * the previous chunk's work. */
int32_t hwpa_host_alloc(void** h_ptr, size_t bytes);
int32_t hwpa_host_free(void* h_ptr);
int32_t hwpa_memcpy_htod(void* d_dst, const void* host_src, size_t bytes);
int32_t hwpa_memcpy_dtoh(void* host_dst, const void* d_src, size_t bytes);
int32_t hwpa_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
/* Times streaming copy kernels (1, 4 and 8 16-byte elements per thread, plain
 * and non-temporal), each `launches` times on the null stream after one
 * untimed launch, and returns the fastest form's mean duration per launch (HIP
 * events).  No reference counterpart: the achievable-bandwidth yardstick of the
 * roofline figures.  bytes and both pointers must be multiples of 16. */
int32_t hwpa_copy_bandwidth(void* d_dst, const void* d_src, size_t bytes, int32_t launches,
                            float* ms_per_launch);

/* Number of equal elements of two int16 device arrays (synchronous); used to
 * measure frame-label agreement between two pipelines without a host copy. */
int32_t hwpa_count_equal_i16(const int16_t* d_a, const int16_t* d_b, size_t n, uint64_t* equal);

/* ---- synthetic input (no reference counterpart: bench/test data) ------- */
int32_t hwpa_tracegen_default(hwpa_tracegen_params* p);
int32_t hwpa_tracegen_run(const hwpa_tracegen_params* p, uint32_t lanes, uint32_t lane_offset,
                          uint32_t frames, uint32_t first_frame, hwpa_tracegen_state* d_state,
                          int32_t init, int16_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HWPA_H */
