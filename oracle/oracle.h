/*
 * oracle.h -- CPU oracle for the trace-analysis path (TEST INFRASTRUCTURE).
 *
 * A scalar C++ restatement of the reference's Host* strategies, used only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker and CPU baseline.  Nothing in the product links or calls it.
 * Each function cites the reference file:line it follows.
 */
#ifndef HWPA_ORACLE_H
#define HWPA_ORACLE_H

#include <stdint.h>
#include "../include/hwpa_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-lane DME diagnostics (traceAnalysis/DmeDiagnostics.h, MaxLikelihoodDiagnostics.h). */
typedef struct orc_dme_diag {
    int32_t converged[HWPA_LANE_SIZE];
    int32_t iter_count[HWPA_LANE_SIZE];
    int32_t status[HWPA_LANE_SIZE];
    int32_t iterations_run;                 /* lane-wide EM iterations executed */
    int32_t pad;
    float log_like[HWPA_LANE_SIZE];
    float conf_factors[8][HWPA_LANE_SIZE];
    float scale_factor[HWPA_LANE_SIZE];
    float confidence[HWPA_LANE_SIZE];       /* estimate confidence before blending */
    float est_fp32[5][3][HWPA_LANE_SIZE];   /* [mode][weight,mean,var] of the estimate */
    float out_fp32[5][3][HWPA_LANE_SIZE];   /* blended model before half export */
    int32_t capture_iter[HWPA_LANE_SIZE];   /* last iteration whose parameters became the estimate, -1 if none */
} orc_dme_diag;

/* Forms of the DME: ORC_FORM_REFERENCE follows DmeEmHost's arithmetic (log-sum-exp
 * E-step, the reference's Cephes exp/log and libm exp2/log2/pow); ORC_FORM_MIRROR
 * is the GPU kernels' restated E-step and math routines, bit for bit. */
#define ORC_FORM_REFERENCE 0
#define ORC_FORM_MIRROR 1
/* ORC_FORM_REFERENCE with libm expf/logf in place of the Cephes routines (a
 * sensitivity yardstick: the same algorithm, different last bits of exp/log). */
#define ORC_FORM_REFERENCE_LIBM 2

typedef struct orc_result {
    hwpa_pulse* pulses;              /* [L*64*max_calls] */
    uint32_t* pulse_counts;          /* [L*64] */
    uint32_t* pulse_overflows;       /* [L*64] */
    hwpa_stat_acc* pulse_bl_stats;   /* [L] */
    int16_t* baselined;              /* [L][F][64] */
    hwpa_baseliner_stats* bl_stats;  /* [L] */
    int16_t* labels;                 /* [L][F][64] */
    float* viterbi_score;            /* [L*64] */
    hwpa_lane_model* models;         /* [L] */
    hwpa_basecalling_metrics* metrics; /* [L], written when flags[3] */
    int32_t flags[4];                /* [0] pulses emitted, [1] estimation ran, [2] labels ran,
                                        [3] a metrics block ended */
} orc_result;

int orc_threads(int n);              /* set worker thread count, returns previous */

void orc_filter_params(int32_t type, hwpa_filter_params* out);
void orc_transition_matrix(const hwpa_config* cfg, hwpa_transition* out);

void* orc_baseliner_create(const hwpa_config* cfg, int32_t lanes);
void orc_baseliner_destroy(void* h);
int32_t orc_baseliner_process(void* h, const int16_t* raw, int32_t frames,
                              int16_t* out, hwpa_baseliner_stats* stats);

void* orc_agg_create(int32_t lanes);
void orc_agg_destroy(void* h);
void orc_agg_add(void* h, const hwpa_baseliner_stats* stats);
void orc_agg_get(void* h, hwpa_baseliner_stats* out);
void orc_agg_reset(void* h);

void* orc_hist_create(const hwpa_config* cfg, int32_t lanes);
void orc_hist_destroy(void* h);
int32_t orc_hist_reset(void* h, const hwpa_baseliner_stats* stats);
void orc_hist_add(void* h, const int16_t* traces, int32_t frames, const hwpa_lane_model* models);
void orc_hist_get(void* h, hwpa_lane_hist* out);
void orc_hist_set(void* h, const hwpa_lane_hist* in);
/* Every int16 sample against n (low bound, bin size) pairs: hw_shortbin, the
 * histogram kernels' divide-free bin, beside the reference's
 * floor((x - lo) / binSize) with the IEEE divide.  out[0] = samples whose bin is
 * in [0, 300), out[1] = those hw_shortbin decides, out[2] = samples it decides
 * wrongly or decides although their bin is outside [0, 300). */
void orc_shortbin_sweep(const float* lo, const float* bs, int32_t n, uint64_t out[3]);
/* Every float with a biased exponent in exps[0..n) (both signs, all 2^23
 * mantissas) and the special values: hw_div_by_const(x, d, 1/d), the labeler's
 * divide-free quotient, against the IEEE quotient x / d, zeros compared up to
 * sign.  out[0] = values checked, out[1] = values where they differ. */
void orc_div_by_const_sweep(float d, const int32_t* exps, int32_t n, uint64_t out[2]);
/* hw_roi_backward_word, the labeler's word-parallel ROI backward pass, over n
 * (word, carry) pairs; the checker is the oracle's per-frame loop (RoiZ). */
void orc_roi_backward_words(const uint32_t* words, const uint32_t* carries, int32_t n, uint32_t* out_words,
                            uint32_t* out_carries);

void orc_dme_init_models(const hwpa_config* cfg, int32_t lanes,
                         const hwpa_baseliner_stats* stats, hwpa_lane_model* out);
/* Learned-start estimation (hwpa_dme_estimate_learned's arithmetic, mirror form):
 * network outputs y [lanes*64][12] (the last the confidence logit, gated by
 * cfg->learned_gate_threshold), output map omap {offset[6], scale[6]};
 * fallback [lanes*64] optional; returns the number of ZMWs re-estimated by the plain EM. */
int32_t orc_dme_estimate_learned(const hwpa_config* cfg, int32_t lanes, const hwpa_lane_hist* hist,
                                 const hwpa_baseliner_stats* stats, hwpa_lane_model* models, const int32_t model_fi[2],
                                 const int32_t est_fi[2], const float* y, const float* omap, uint8_t* fallback);
void orc_dme_estimate(int32_t form, const hwpa_config* cfg, int32_t lanes, const hwpa_lane_hist* hist,
                      const hwpa_baseliner_stats* stats, hwpa_lane_model* models,
                      const int32_t model_fi[2], const int32_t est_fi[2], orc_dme_diag* diag);

void* orc_labeler_create(const hwpa_config* cfg, int32_t lanes);
void orc_labeler_destroy(void* h);
void orc_labeler_process(void* h, const int16_t* traces, int32_t frames,
                         const hwpa_lane_model* models, int16_t* labels,
                         int16_t* latent_trace, float* scores);

void* orc_pulse_create(const hwpa_config* cfg, int32_t lanes);
void orc_pulse_destroy(void* h);
void orc_pulse_process(void* h, const int16_t* labels, const int16_t* latent_trace,
                       const int16_t* cur_trace, int32_t frames, int32_t first_frame,
                       const hwpa_lane_model* models, hwpa_pulse* pulses, uint32_t* counts,
                       uint32_t* overflows, hwpa_stat_acc* bl_stats);

void* orc_metrics_create(const hwpa_config* cfg, int32_t lanes);
void orc_metrics_destroy(void* h);
void orc_metrics_set_tree(void* h, const hwpa_activity_tree* t);
int32_t orc_metrics_process(void* h, const hwpa_pulse* pulses, const uint32_t* counts, uint32_t max_calls,
                            int32_t frames, int32_t last_frame, const hwpa_baseliner_stats* bl_stats,
                            const float* viterbi_score, const hwpa_stat_acc* pd_stats,
                            const hwpa_lane_model* models, hwpa_basecalling_metrics* out);
float orc_autocorr(const float* x, int32_t n, const int32_t* sizes, int32_t pieces);

void* orc_analyzer_create(const hwpa_config* cfg, uint32_t pool_id, int32_t lanes, int32_t frames);
void orc_analyzer_destroy(void* h);
int32_t orc_analyzer_process_encoded(void* h, const void* traces, int32_t encoding, uint32_t pool_id,
                                     int32_t first_frame, orc_result* res);
int32_t orc_analyzer_process(void* h, const int16_t* traces, uint32_t pool_id,
                             int32_t first_frame, orc_result* res);
void orc_analyzer_set_activity_tree(void* h, const hwpa_activity_tree* t);
void orc_analyzer_set_em_form(void* h, int32_t form);   /* ORC_FORM_*; default ORC_FORM_REFERENCE */

void orc_tracegen(const hwpa_tracegen_params* p, int32_t lanes, int32_t lane_offset,
                  int32_t frames, uint32_t first_frame, hwpa_tracegen_state* state,
                  int32_t init, int16_t* out, int16_t* truth /* nullable: 0 = baseline, a+1 = analog a */);

/* BAZ packet encoding (layouts: 0 ProductionPulses, 1 InternalPulses, 2-4 reference test layouts). */
int64_t orc_baz_encode_fields(int32_t layout, int32_t n, const uint8_t* label, const uint32_t* start,
                              const uint16_t* width, const uint8_t* isbase, const float* pkmax, const float* pkmid,
                              const float* pkmean, const float* pkvar, uint64_t* last, uint8_t* out);
int64_t orc_baz_decode_fields(int32_t layout, int32_t n, const uint8_t* in, uint64_t* last, uint8_t* label,
                              uint32_t* start, uint16_t* width, uint8_t* isbase, float* pkmax, float* pkmid,
                              float* pkmean, float* pkvar);
int32_t orc_baz_layout_groups(int32_t layout, int32_t* group_bits, int32_t* group_fields, int32_t cap);
uint64_t orc_baz_serial_roundtrip(int32_t ser, int32_t bits, int32_t obytes, int32_t sgn, uint64_t v,
                                  int32_t* overflow_bytes);
uint64_t orc_baz_floatfixed_apply(float f, int32_t sgn, uint32_t scale, int32_t bytes);
float orc_baz_floatfixed_revert(uint64_t v, int32_t sgn, uint32_t scale);
int64_t orc_baz_encode_batch(int32_t internal, const hwpa_pulse* pulses, const uint32_t* counts, uint32_t max_calls,
                             int32_t zmws, uint64_t* last, uint8_t* out, uint64_t cap, uint64_t* offsets,
                             uint32_t* events);

#ifdef __cplusplus
}
#endif

#endif /* HWPA_ORACLE_H */
