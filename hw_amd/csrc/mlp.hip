// mlp.hip -- the learned initialiser of the EM: a 128 -> 256 -> 256 -> 11
// perceptron with ReLU hidden layers, inference and training on the matrix
// cores (bfloat16 operands, fp32 accumulation).  No reference counterpart.
//
// Parameter vector (fp32 master copy, PyTorch Linear convention W[out][in]):
//   W1[256][128] b1[256] W2[256][256] b2[256] W3[11][256] b3[11]
// A bf16 mirror of the whole vector feeds the matrix cores; biases are added
// in fp32 from the master copy.
//
// One generic GEMM kernel serves every layer and every gradient:
//   C[m][n] = sum_k A[m][k] * B[k][n]
// with element strides for A and B (so transposed operands need no copies),
// a 64x64 tile per 256-thread workgroup, one 32x32 sub-tile per wave and
// v_mfma_f32_32x32x16_bf16 over k in steps of 16.  Operand lanes: lane l
// supplies A[l%32][k0+8(l/32)+j] and B[k0+8(l/32)+j][l%32], j = 0..7; the 16
// accumulators of lane l hold C[8(i/4)+4(l/32)+i%4][l%32].

#include "internal.h"

#include <cmath>
#include <vector>

namespace hwpa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MI = HWPA_MLP_IN, MH = HWPA_MLP_HIDDEN, MO = HWPA_MLP_OUT;
constexpr size_t OFF_W1 = 0;
constexpr size_t OFF_B1 = OFF_W1 + (size_t)MH * MI;
constexpr size_t OFF_W2 = OFF_B1 + MH;
constexpr size_t OFF_B2 = OFF_W2 + (size_t)MH * MH;
constexpr size_t OFF_W3 = OFF_B2 + MH;
constexpr size_t OFF_B3 = OFF_W3 + (size_t)MO * MH;
constexpr size_t NPARAM = OFF_B3 + MO;
static_assert(NPARAM == HWPA_MLP_PARAMS, "parameter count");

enum { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_RELU_BF16 = 2, EPI_MASK_BF16 = 3 };

struct GemmArgs {
    const uint16_t* A;
    int64_t a_m, a_k;
    const uint16_t* B;
    int64_t b_k, b_n;
    int32_t M, N, K;
    const float* bias;
    const uint16_t* mask;
    int64_t mk_m, mk_n;
    float* Cf;
    uint16_t* Cb;
    int64_t c_m, c_n;
    // Split-K (EPI_F32 only): slice ks of kSplit covers k in [ks * kSlice,
    // (ks + 1) * kSlice) and writes its partial product at Cf + ks * sliceStride.
    int32_t kSplit, kSlice;
    int64_t sliceStride;
};

template <bool CONTIG>
__device__ __forceinline__ bf16x8 LoadK(const uint16_t* __restrict__ p, bool valid, int64_t off, int k,
                                        int64_t sk, int K)
{
    if (CONTIG && valid && k + 8 <= K) {
        const uint4 q = *reinterpret_cast<const uint4*>(p + off + k);
        return __builtin_bit_cast(bf16x8, q);
    }
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k0 = k + 2 * j, k1 = k0 + 1;
        const uint32_t lo = (valid && k0 < K) ? p[off + (int64_t)k0 * sk] : 0u;
        const uint32_t hi = (valid && k1 < K) ? p[off + (int64_t)k1 * sk] : 0u;
        w[j] = lo | (hi << 16);
    }
    const uint4 q = make_uint4(w[0], w[1], w[2], w[3]);
    return __builtin_bit_cast(bf16x8, q);
}

__device__ __forceinline__ bool PositiveBf16(uint16_t b) { return b != 0 && (b & 0x8000u) == 0; }

template <bool AK, bool BK, int EPI>
__global__ __launch_bounds__(256) void GemmKernel(GemmArgs g, uint32_t nTiles, uint32_t tiles)
{
    const uint32_t tile = blockIdx.x % tiles, ks = blockIdx.x / tiles;
    const int mt = (int)(tile / nTiles), nt = (int)(tile % nTiles);
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m0 = mt * 64 + (wave >> 1) * 32, n0 = nt * 64 + (wave & 1) * 32;
    const int r = l & 31, kq = (l >> 5) * 8;
    const int am = m0 + r, bn = n0 + r;
    const bool aValid = am < g.M, bValid = bn < g.N;
    const int64_t aOff = (int64_t)am * g.a_m, bOff = (int64_t)bn * g.b_n;
    f32x16 acc = {};
    const int kBeg = (int)ks * g.kSlice;
    const int kEnd = (kBeg + g.kSlice < g.K) ? kBeg + g.kSlice : g.K;
    for (int k0 = kBeg; k0 < kEnd; k0 += 16) {
        const bf16x8 a = LoadK<AK>(g.A, aValid, aOff, k0 + kq, g.a_k, kEnd);
        const bf16x8 b = LoadK<BK>(g.B, bValid, bOff, k0 + kq, g.b_k, kEnd);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    float* const cf = g.Cf + (int64_t)ks * g.sliceStride;
    const int n = n0 + r;
    if (n >= g.N) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = m0 + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3);
        if (m >= g.M) continue;
        const int64_t c = (int64_t)m * g.c_m + (int64_t)n * g.c_n;
        const float v = acc[i];
        if (EPI == EPI_F32) {
            cf[c] = v;
        } else if (EPI == EPI_BIAS_F32) {
            g.Cf[c] = v + g.bias[n];
        } else if (EPI == EPI_BIAS_RELU_BF16) {
            const float z = v + g.bias[n];
            g.Cb[c] = hw_f2bf_nonnan(z > 0.0f ? z : 0.0f);
        } else {
            const uint16_t mk = g.mask[(int64_t)m * g.mk_m + (int64_t)n * g.mk_n];
            g.Cb[c] = PositiveBf16(mk) ? hw_f2bf(v) : (uint16_t)0;
        }
    }
}

template <bool AK, bool BK, int EPI>
static int32_t Gemm(const GemmArgs& g, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return Fail(HWPA_ERR_INVALID_ARG, "gemm: empty shape");
    // Contiguous operands are read 8 elements (16 bytes) at a time.
    if (AK && (((uintptr_t)g.A & 15) != 0 || (g.a_m & 7) != 0 || g.a_k != 1))
        return Fail(HWPA_ERR_INVALID_ARG, "gemm: A not 16-byte aligned along k");
    if (BK && (((uintptr_t)g.B & 15) != 0 || (g.b_n & 7) != 0 || g.b_k != 1))
        return Fail(HWPA_ERR_INVALID_ARG, "gemm: B not 16-byte aligned along k");
    GemmArgs h = g;
    if (h.kSplit <= 1) {
        h.kSplit = 1;
        h.kSlice = h.K;
        h.sliceStride = 0;
    } else if (EPI != EPI_F32 || h.kSlice % 16 != 0 || (int64_t)h.kSplit * h.kSlice < h.K) {
        return Fail(HWPA_ERR_INVALID_ARG, "gemm: bad k split");
    }
    const uint32_t mt = (uint32_t)((h.M + 63) / 64), nt = (uint32_t)((h.N + 63) / 64);
    hipLaunchKernelGGL((GemmKernel<AK, BK, EPI>), dim3(mt * nt * (uint32_t)h.kSplit), dim3(256), 0, s, h, nt, mt * nt);
    return CheckLaunch("GemmKernel");
}

// out[i] = sum over the kSplit slices of part[ks * n + i], slices in order.
__global__ __launch_bounds__(256) void SliceSumKernel(const float* __restrict__ part, int32_t slices, int64_t n,
                                                      float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int32_t k = 0; k < slices; ++k) acc += part[(int64_t)k * n + i];
    out[i] = acc;
}

// Weight gradient out[M][N] (row-major) = A^T-like product over a long K
// (the batch): K is cut into up to kMaxSplit slices of at least 128 rows
// whose partial products are summed in a fixed order.
constexpr int kMaxSplit = 32;

template <bool AK, bool BK>
static int32_t GemmSplitK(GemmArgs g, float* part, float* out, hipStream_t s)
{
    int split = (g.K + 127) / 128;
    split = split < kMaxSplit ? split : kMaxSplit;
    if (split < 1) split = 1;
    const int slice = (((g.K + split - 1) / split) + 15) / 16 * 16;
    split = (g.K + slice - 1) / slice;
    g.kSplit = split;
    g.kSlice = slice;
    g.Cf = part;
    g.c_m = g.N;
    g.c_n = 1;
    g.sliceStride = (int64_t)g.M * g.N;
    int32_t rc = Gemm<AK, BK, EPI_F32>(g, s);
    if (rc != HWPA_OK) return rc;
    const int64_t n = (int64_t)g.M * g.N;
    hipLaunchKernelGGL(SliceSumKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, split, n, out);
    return CheckLaunch("SliceSumKernel");
}

static GemmArgs MakeGemm(int M, int N, int K, const uint16_t* A, int64_t a_m, int64_t a_k, const uint16_t* B,
                         int64_t b_k, int64_t b_n)
{
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.A = A;
    g.a_m = a_m;
    g.a_k = a_k;
    g.B = B;
    g.b_k = b_k;
    g.b_n = b_n;
    return g;
}

// ---- features: 300-bin lane histogram -> 128 re-binned, normalised bf16 ----
// x[i] = 16 * (cnt[2i] + cnt[2i+1]) / N,  N = sum of cnt[0..255] (x = 0 if N = 0).
__global__ __launch_bounds__(64) void FeaturesKernel(const hwpa_lane_hist* __restrict__ hists, uint16_t* __restrict__ x)
{
    const int lane = blockIdx.x, z = threadIdx.x;
    const hwpa_lane_hist& h = hists[lane];
    uint32_t total = 0;
    for (int b = 0; b < 2 * MI; ++b) total += h.bin_count[b][z];
    const float scale = total ? 16.0f / (float)total : 0.0f;
    uint4* out = reinterpret_cast<uint4*>(x + ((size_t)lane * LS + z) * MI);
    for (int i = 0; i < MI; i += 8) {
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b0 = 2 * (i + 2 * j), b1 = b0 + 2;
            const float v0 = (float)(h.bin_count[b0][z] + h.bin_count[b0 + 1][z]) * scale;
            const float v1 = (float)(h.bin_count[b1][z] + h.bin_count[b1 + 1][z]) * scale;
            w[j] = (uint32_t)hw_f2bf(v0) | ((uint32_t)hw_f2bf(v1) << 16);
        }
        out[i / 8] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// ---- training targets from the EM diagnostics ---------------------------------
// Raw units: u = 2 * binSize (one re-binned bin), origin at the histogram's
// lower bound:  r0 = (mu_bl - lo) / u,  r1 = log(var_bl / u^2),
// r2..5 = (mu_a - lo) / u,  r6..10 = mixture weights (softmax targets),
// r11 = 1 when the estimate is a valid target (converged, status OK, enough
// confidence), else 0 (the confidence output's target).
__global__ __launch_bounds__(64) void TargetsKernel(const hwpa_lane_hist* __restrict__ hists,
                                                    const float* __restrict__ diag, float minConf,
                                                    float* __restrict__ t, uint8_t* __restrict__ valid)
{
    const int lane = blockIdx.x, z = threadIdx.x;
    const float* dg = diag + (size_t)lane * HWPA_EM_DIAG_ROWS * LS + z;
    const float lo = hists[lane].low_bound[z], u = 2.0f * hists[lane].bin_size[z];
    float* o = t + ((size_t)lane * LS + z) * MO;
    const float var0 = dg[(10 + 0) * LS];
    o[0] = (dg[(5 + 0) * LS] - lo) / u;
    o[1] = hw_logf(var0 / (u * u));
    for (int a = 0; a < NA; ++a) o[2 + a] = (dg[(5 + 1 + a) * LS] - lo) / u;
    for (int k = 0; k < 5; ++k) o[6 + k] = dg[k * LS];
    const bool ok = dg[16 * LS] > 0.0f && dg[17 * LS] == 0.0f && dg[15 * LS] >= minConf &&
                    hw_isfinite(o[0]) && hw_isfinite(o[1]) && var0 > 0.0f;
    o[11] = ok ? 1.0f : 0.0f;
    valid[(size_t)lane * LS + z] = ok ? 1 : 0;
}

// ---- loss ----------------------------------------------------------------------
// Network output j < 6 stands for the raw target off[j] + scale[j] * y[j]; the
// loss is sum_j w[j] (y[j] - (r[j] - off[j]) / scale[j])^2 + w[6] * CE(softmax(y[6..10]), r[6..10])
// + w[7] * BCE(sigmoid(y[11]), r[11]), the last as softplus(y) - r y in its stable form
// max(y, 0) - r y + log(1 + exp(-|y|)).  Samples marked invalid keep only the last term.
// omap = {off[6], scale[6], w[8]}.
__global__ __launch_bounds__(256) void LossKernel(const float* __restrict__ y, const float* __restrict__ t,
                                                  const uint8_t* __restrict__ valid, uint32_t batch,
                                                  const float* __restrict__ omap, float* __restrict__ dy,
                                                  uint16_t* __restrict__ dyb, float* __restrict__ lossPer)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const float* yb = y + (size_t)b * MO;
    const float* tb = t + (size_t)b * MO;
    const float invB = 1.0f / (float)batch;
    float g[MO];
    // Confidence output: logistic loss against target 11.
    const float wg = omap[19];
    const float yc = yb[11], tc = tb[11];
    const float ez = hw_expf(-hw_fabsf(yc));
    float loss = wg * ((hw_maxf(yc, 0.0f) - tc * yc) + hw_logf(1.0f + ez));
    const float sg = (yc >= 0.0f) ? 1.0f / (1.0f + ez) : ez / (1.0f + ez);
    g[11] = (wg * (sg - tc)) * invB;
    if (valid && valid[b] == 0) {
        for (int j = 0; j < MO; ++j) {
            const float gj = (j == 11) ? g[11] : 0.0f;
            dy[(size_t)b * MO + j] = gj;
            dyb[(size_t)b * MO + j] = hw_f2bf(gj);
        }
        lossPer[b] = loss;
        return;
    }
    for (int j = 0; j < 6; ++j) {
        const float tn = (tb[j] - omap[j]) / omap[6 + j];
        const float d = yb[j] - tn;
        const float w = omap[12 + j];
        loss += w * (d * d);
        g[j] = ((2.0f * w) * d) * invB;
    }
    const float wce = omap[18];
    float mx = yb[6];
    for (int k = 1; k < 5; ++k) mx = hw_maxf(mx, yb[6 + k]);
    float e[5], se = 0.0f;
    for (int k = 0; k < 5; ++k) {
        e[k] = hw_expf(yb[6 + k] - mx);
        se += e[k];
    }
    const float lse = mx + hw_logf(se);
    for (int k = 0; k < 5; ++k) {
        const float p = e[k] / se;
        loss -= wce * (tb[6 + k] * (yb[6 + k] - lse));
        g[6 + k] = (wce * (p - tb[6 + k])) * invB;
    }
    for (int j = 0; j < MO; ++j) {
        dy[(size_t)b * MO + j] = g[j];
        dyb[(size_t)b * MO + j] = hw_f2bf(g[j]);
    }
    lossPer[b] = loss;
}

// Deterministic mean of per-sample losses (one workgroup, fixed order).
__global__ __launch_bounds__(256) void MeanKernel(const float* __restrict__ v, uint32_t n, float* __restrict__ out)
{
    __shared__ float part[256];
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += 256) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0] / (float)n;
}

// Column sums of a [rows][cols] matrix (bf16 or fp32), two fixed-order passes.
// Rows per partial sum of the bias-gradient column sums: batch / 32 workgroups, each
// thread with its 32 loads in flight; the partials are then added in slice order.
constexpr int COLSUM_ROWS = 32;
template <typename T>
__device__ __forceinline__ float LoadVal(const T* p, size_t i);
template <>
__device__ __forceinline__ float LoadVal<float>(const float* p, size_t i) { return p[i]; }
template <>
__device__ __forceinline__ float LoadVal<uint16_t>(const uint16_t* p, size_t i) { return hw_bf2f(p[i]); }

template <typename T>
__global__ __launch_bounds__(256) void ColSumPartKernel(const T* __restrict__ d, uint32_t rows, uint32_t cols,
                                                        float* __restrict__ part)
{
    const uint32_t c = threadIdx.x + blockIdx.y * 256;
    if (c >= cols) return;
    const uint32_t r0 = blockIdx.x * COLSUM_ROWS;
    const uint32_t r1 = min(rows, r0 + COLSUM_ROWS);
    // All loads of the slice are issued before the (row-ordered) adds.
    float v[COLSUM_ROWS];
#pragma unroll
    for (int i = 0; i < COLSUM_ROWS; ++i) v[i] = (r0 + i < r1) ? LoadVal<T>(d, (size_t)(r0 + i) * cols + c) : 0.0f;
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < COLSUM_ROWS; ++i) s += v[i];
    part[(size_t)blockIdx.x * cols + c] = s;
}

__global__ __launch_bounds__(256) void ColSumFinalKernel(const float* __restrict__ part, uint32_t chunks,
                                                         uint32_t cols, float* __restrict__ out)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    // Partials in slice order, 16 loads in flight at a time.
    float s = 0.0f;
    for (uint32_t k0 = 0; k0 < chunks; k0 += 16) {
        float v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = (k0 + i < chunks) ? part[(size_t)(k0 + i) * cols + c] : 0.0f;
#pragma unroll
        for (int i = 0; i < 16; ++i) s += v[i];
    }
    out[c] = s;
}

__global__ __launch_bounds__(256) void AdamKernel(float* __restrict__ p, const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  uint16_t* __restrict__ pb, uint32_t n, float lr, float b1,
                                                  float b2, float c1, float c2, float eps)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    const float pi = p[i] - lr * mh / (hw_sqrtf(vh) + eps);
    p[i] = pi;
    pb[i] = hw_f2bf(pi);
}

__global__ __launch_bounds__(256) void ToBf16Kernel(const float* __restrict__ p, uint16_t* __restrict__ pb, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = hw_f2bf(p[i]);
}

}  // namespace hwpa

using namespace hwpa;

struct hwpa_initializer {
    float* P = nullptr;     // fp32 master parameters
    uint16_t* Pb = nullptr; // bf16 mirror
    float* G = nullptr;     // gradients
    float* Mo = nullptr;    // Adam moments
    float* Ve = nullptr;
    int64_t step = 0;
    // activations, sized for `rows`
    uint32_t rows = 0;
    uint16_t* H1 = nullptr;
    uint16_t* H2 = nullptr;
    float* Y = nullptr;
    // training buffers, sized for `batchCap`
    uint32_t batchCap = 0;
    float* dY = nullptr;
    uint16_t* dYb = nullptr;
    uint16_t* dH1 = nullptr;
    uint16_t* dH2 = nullptr;
    float* lossPer = nullptr;
    float* colPart = nullptr;
    float* lossDev = nullptr;
    float* omap = nullptr;      // device {off[6], scale[6], w[8]}
    float omapHost[20] = {};
    float* kPart = nullptr;     // split-K partial weight gradients, kMaxSplit x MH x MH
};

namespace hwpa {

static void FreeActs(hwpa_initializer* m)
{
    DeviceFree(m->H1);
    DeviceFree(m->H2);
    DeviceFree(m->Y);
    m->H1 = m->H2 = nullptr;
    m->Y = nullptr;
    m->rows = 0;
}

static void FreeTrain(hwpa_initializer* m)
{
    DeviceFree(m->dY);
    DeviceFree(m->dYb);
    DeviceFree(m->dH1);
    DeviceFree(m->dH2);
    DeviceFree(m->lossPer);
    DeviceFree(m->colPart);
    m->dY = nullptr;
    m->dYb = m->dH1 = m->dH2 = nullptr;
    m->lossPer = m->colPart = nullptr;
    m->batchCap = 0;
}

static int32_t EnsureRows(hwpa_initializer* m, uint32_t rows)
{
    if (rows <= m->rows) return HWPA_OK;
    // Buffers may still be in use by queued work on any stream.
    HWPA_HIP(hipDeviceSynchronize());
    FreeActs(m);
    int32_t rc = DeviceAlloc((void**)&m->H1, (size_t)rows * MH * 2, "mlp activations");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->H2, (size_t)rows * MH * 2, "mlp activations");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->Y, (size_t)rows * MO * 4, "mlp outputs");
    if (rc != HWPA_OK) {
        FreeActs(m);
        return rc;
    }
    m->rows = rows;
    return HWPA_OK;
}

static int32_t EnsureBatch(hwpa_initializer* m, uint32_t batch)
{
    int32_t rc = EnsureRows(m, batch);
    if (rc != HWPA_OK || batch <= m->batchCap) return rc;
    HWPA_HIP(hipDeviceSynchronize());
    FreeTrain(m);
    const size_t chunks = (batch + COLSUM_ROWS - 1) / COLSUM_ROWS;
    rc = DeviceAlloc((void**)&m->dY, (size_t)batch * MO * 4, "mlp gradients");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->dYb, (size_t)batch * MO * 2, "mlp gradients");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->dH1, (size_t)batch * MH * 2, "mlp gradients");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->dH2, (size_t)batch * MH * 2, "mlp gradients");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->lossPer, (size_t)batch * 4, "mlp loss");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->colPart, chunks * MH * 4, "mlp bias sums");
    if (rc != HWPA_OK) {
        FreeTrain(m);
        return rc;
    }
    m->batchCap = batch;
    return HWPA_OK;
}

static int32_t ColSum(const void* d, bool bf16, uint32_t rows, uint32_t cols, float* part, float* out, hipStream_t s)
{
    const uint32_t chunks = (rows + COLSUM_ROWS - 1) / COLSUM_ROWS;
    const dim3 grid(chunks, (cols + 255) / 256);
    if (bf16)
        hipLaunchKernelGGL(ColSumPartKernel<uint16_t>, grid, dim3(256), 0, s, (const uint16_t*)d, rows, cols, part);
    else
        hipLaunchKernelGGL(ColSumPartKernel<float>, grid, dim3(256), 0, s, (const float*)d, rows, cols, part);
    int32_t rc = CheckLaunch("ColSumPartKernel");
    if (rc != HWPA_OK) return rc;
    hipLaunchKernelGGL(ColSumFinalKernel, dim3((cols + 255) / 256), dim3(256), 0, s, part, chunks, cols, out);
    return CheckLaunch("ColSumFinalKernel");
}

// Fused forward pass: 32 rows per wave, kFusedWaves (8) waves per workgroup.  Each layer is the
// 32x32x16 bf16 MFMA over 32-column tiles with k ascending -- the arithmetic of
// GemmKernel, so outputs equal the three-GEMM pass bit for bit -- but the hidden
// activations stay in LDS (row-major, read back as the next layer's operand) instead of
// HBM; h1 / h2 are written out only when given (training keeps them for the backward
// pass).  The weights of each 16-wide k-step of layers 1 and 2 (all 256 output columns:
// 8 KB) are staged once per workgroup in LDS, double-buffered with one barrier per step, so
// the 8 waves (256 rows) of a workgroup share one read of them instead of each wave
// fetching all 192 KB from L2.
constexpr int kFusedRows = 32, kFusedWaves = 8, kHsStride = MH + 8;
constexpr int kSlab = MH * 16;   // bf16 weights of one k-step: [256 columns][16 k]

__device__ __forceinline__ bf16x8 LoadRow8(const uint16_t* __restrict__ p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// HIST: the inputs are computed from the lane histograms as FeaturesKernel does (row =
// lane * 64 + ZMW), so inference reads the counters once and writes no input matrix.  Each
// half-wave needs half of the 128 re-binned inputs of its row (k-chunks kq = 0 / 8 of every
// 16-wide k-step): it sums exactly those bin pairs, the two halves' partial totals give the
// row total, and the scaled bf16 inputs stay in registers.
template <bool HIST>
__global__ __launch_bounds__(64 * kFusedWaves) void MlpFusedForwardKernel(const uint16_t* __restrict__ x,
                                                             const hwpa_lane_hist* __restrict__ hists, uint32_t rows,
                                                             const uint16_t* __restrict__ Wb,
                                                             const float* __restrict__ P, float* __restrict__ y,
                                                             uint16_t* __restrict__ h1out, uint16_t* __restrict__ h2out)
{
    __shared__ __align__(16) uint16_t hs[kFusedWaves][kFusedRows][kHsStride];
    __shared__ __align__(16) uint16_t ws[2][kSlab];
    static_assert(64 * kFusedWaves == 2 * MH, "one 16-byte half column of a slab per thread");
    // 132 KiB of activations + 16 KiB of weight slabs: one workgroup per CU's 160 KiB of LDS.
    static_assert(sizeof(hs) + sizeof(ws) <= 160 * 1024, "LDS budget of a CU");
    // Thread t stages half t & 1 of column t >> 1 of a k-step's weights (16 bytes).
    const int sn = threadIdx.x >> 1, sh = (threadIdx.x & 1) * 8;
    auto slabLoad = [&](int off, int kdim, int k0) {
        return *reinterpret_cast<const uint4*>(Wb + off + (size_t)sn * kdim + k0 + sh);
    };
    auto slabStore = [&](int buf, uint4 v) { *reinterpret_cast<uint4*>(&ws[buf][sn * 16 + sh]) = v; };
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r = l & 31, kq = (l >> 5) * 8;
    const int64_t m0 = ((int64_t)blockIdx.x * kFusedWaves + wave) * kFusedRows;
    const bool rowValid = m0 + r < (int64_t)rows;
    uint16_t (*h)[kHsStride] = hs[wave];
    auto hp = [&](int row, int col) { return &h[row][col]; };
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));

    // Every product is (weights) x (activations): the weights are the A operand and a wave's 32
    // rows the B operand, so accumulator element i of lane l is output feature outRow(i) of
    // the tile (f32x16 layout of the 32x32 MFMA) for row r: each lane holds four runs of four
    // consecutive features of its own row.
    auto outRow = [&](int i) { return (i >> 2) * 8 + (l >> 5) * 4 + (i & 3); };
    // Bias, ReLU, bf16 and the row's activations back into LDS (and to hout when given): one
    // 8-byte store per run of four features.
    auto hiddenOut = [&](const f32x16 (&a)[MH / 32], int boff, uint16_t* hout) {
#pragma unroll
        for (int j = 0; j < MH / 32; ++j) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int f0 = j * 32 + g * 8 + (l >> 5) * 4;
                const float4 bias = *reinterpret_cast<const float4*>(P + boff + f0);
                const float bv[4] = {bias.x, bias.y, bias.z, bias.w};
                uint32_t hv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float zv = a[j][4 * g + q] + bv[q];
                    hv[q] = hw_f2bf_nonnan(zv > 0.0f ? zv : 0.0f);
                }
                const uint2 pk = make_uint2(hv[0] | (hv[1] << 16), hv[2] | (hv[3] << 16));
                *reinterpret_cast<uint2*>(hp(r, f0)) = pk;
                if (hout && rowValid) *reinterpret_cast<uint2*>(hout + (m0 + r) * MH + f0) = pk;
            }
        }
    };
    static_assert(OFF_B1 % 4 == 0 && OFF_B2 % 4 == 0 && (kHsStride * 2) % 8 == 0, "aligned runs of four");

    // ---- inputs from the histogram (HIST)
    uint32_t xin[MI / 16][4];
    if (HIST) {
        // The wave's 32 ZMWs x 2 MI bins of counters go through its activation buffer: 16-byte
        // loads of 8 ZMWs of one bin (a wave instruction moves 16 bins), then each lane reads
        // its ZMW's bins from LDS.  Rows come in whole half-lanes, so a wave is valid or not.
        static_assert(2 * MI * kFusedRows <= kFusedRows * kHsStride, "counters fit the activation buffer");
        uint16_t* hc = &h[0][0];   // [bin][32 ZMWs]
        if (m0 < (int64_t)rows) {
            const uint16_t* lb = &hists[m0 / LS].bin_count[0][m0 % LS];
            constexpr int NIT = 2 * MI / 16;
            const int piece = (l & 3) * 8;
            uint4 v[NIT];   // every load in flight before the first store
#pragma unroll
            for (int it = 0; it < NIT; ++it)
                v[it] = *reinterpret_cast<const uint4*>(lb + (size_t)(16 * it + (l >> 2)) * LS + piece);
#pragma unroll
            for (int it = 0; it < NIT; ++it)
                *reinterpret_cast<uint4*>(hc + (16 * it + (l >> 2)) * kFusedRows + piece) = v[it];
        }
        __syncthreads();
        uint32_t pair[MI / 16][8];
        uint32_t part = 0;
#pragma unroll
        for (int s = 0; s < MI / 16; ++s)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int i = 16 * s + kq + j;
                const uint32_t v =
                    rowValid ? (uint32_t)hc[(2 * i) * kFusedRows + r] + (uint32_t)hc[(2 * i + 1) * kFusedRows + r] : 0u;
                pair[s][j] = v;
                part += v;
            }
        const uint32_t total = part + (uint32_t)__shfl_xor((int)part, 32, 64);
        const float scale = total ? 16.0f / (float)total : 0.0f;
#pragma unroll
        for (int s = 0; s < MI / 16; ++s)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                xin[s][q] = (uint32_t)hw_f2bf_nonnan((float)pair[s][2 * q] * scale) |   // count x finite scale
                            ((uint32_t)hw_f2bf_nonnan((float)pair[s][2 * q + 1] * scale) << 16);
    }

    // ---- layer 1: [32 x MI] x [MI x MH]
    f32x16 acc[MH / 32];
#pragma unroll
    for (int j = 0; j < MH / 32; ++j) acc[j] = f32x16{};
    const uint16_t* xr = HIST ? nullptr : x + (m0 + r) * MI;
    // Weight slabs: step k + 1's slab is in flight during step k and stored after its
    // products; step k + 2's load is issued at the start of step k.
    uint4 next = slabLoad(OFF_W1, MI, 16);
    slabStore(0, slabLoad(OFF_W1, MI, 0));
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < MI; k0 += 16) {
        const int buf = (k0 / 16) & 1;
        uint4 next2{};
        if (k0 + 32 < MI) next2 = slabLoad(OFF_W1, MI, k0 + 32);
        bf16x8 a;
        if (HIST)
            a = __builtin_bit_cast(bf16x8, make_uint4(xin[k0 / 16][0], xin[k0 / 16][1], xin[k0 / 16][2],
                                                      xin[k0 / 16][3]));
        else
            a = rowValid ? LoadRow8(xr + k0 + kq) : zero8;
        bf16x8 b[MH / 32];
#pragma unroll
        for (int j = 0; j < MH / 32; ++j) b[j] = LoadRow8(&ws[buf][(j * 32 + r) * 16 + kq]);
#pragma unroll
        for (int j = 0; j < MH / 32; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a, acc[j], 0, 0, 0);
        if (k0 + 16 < MI) slabStore(buf ^ 1, next);
        next = next2;
        __syncthreads();
    }
    hiddenOut(acc, OFF_B1, h1out);
    __syncthreads();

    // ---- layer 2: [32 x MH] x [MH x MH]
#pragma unroll
    for (int j = 0; j < MH / 32; ++j) acc[j] = f32x16{};
    // Layer 1's last barrier has passed: every wave is done with both slab buffers.
    next = slabLoad(OFF_W2, MH, 16);
    slabStore(0, slabLoad(OFF_W2, MH, 0));
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < MH; k0 += 16) {
        const int buf = (k0 / 16) & 1;
        uint4 next2{};
        if (k0 + 32 < MH) next2 = slabLoad(OFF_W2, MH, k0 + 32);
        const bf16x8 a = LoadRow8(hp(r, k0 + kq));
        bf16x8 b[MH / 32];
#pragma unroll
        for (int j = 0; j < MH / 32; ++j) b[j] = LoadRow8(&ws[buf][(j * 32 + r) * 16 + kq]);
#pragma unroll
        for (int j = 0; j < MH / 32; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a, acc[j], 0, 0, 0);
        if (k0 + 16 < MH) slabStore(buf ^ 1, next);
        next = next2;
        __syncthreads();
    }
    hiddenOut(acc, OFF_B2, h2out);
    __syncthreads();

    // ---- layer 3: [32 x MH] x [MH x MO], one 32-column tile of which MO columns are real
    f32x16 out = f32x16{};
    bf16x8 b3[MH / 16];   // all of this lane's layer-3 weights in flight at once
#pragma unroll
    for (int k = 0; k < MH / 16; ++k) b3[k] = r < MO ? LoadRow8(Wb + OFF_W3 + (size_t)r * MH + 16 * k + kq) : zero8;
#pragma unroll
    for (int k = 0; k < MH / 16; ++k) {
        const bf16x8 a = LoadRow8(hp(r, 16 * k + kq));
        out = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b3[k], a, out, 0, 0, 0);
    }
    if (rowValid) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int f = outRow(i);
            if (f < MO) y[(m0 + r) * MO + f] = out[i] + P[OFF_B3 + f];
        }
    }
}

// Forward pass over `rows` inputs through the activation buffers h1, h2 into y.
static int32_t Forward(const hwpa_initializer* m, const uint16_t* x, uint32_t rows, float* y, uint16_t* h1,
                       uint16_t* h2, hipStream_t s)
{
    if (rows == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp forward: no rows");
    if (((uintptr_t)x & 15) != 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp forward: input not 16-byte aligned");
    const uint32_t per = (uint32_t)(kFusedRows * kFusedWaves);
    hipLaunchKernelGGL(MlpFusedForwardKernel<false>, dim3((rows + per - 1) / per), dim3(64 * kFusedWaves), 0, s, x,
                       nullptr, rows,
                       m->Pb, m->P, y, h1, h2);
    return CheckLaunch("MlpFusedForwardKernel");
}

// Inference straight from lane histograms: FeaturesKernel + Forward in one launch.
static int32_t ForwardFromHist(const hwpa_initializer* m, const hwpa_lane_hist* hist, uint32_t lanes, float* y,
                               hipStream_t s)
{
    if (lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp forward: no lanes");
    // Whole lanes, so every wave's 32 rows are all valid or all past the end: the
    // HIST input staging reads a wave's half-lane of counters only when m0 < rows.
    static_assert(LS % kFusedRows == 0, "a wave's rows lie in one lane");
    const uint32_t rows = lanes * (uint32_t)LS;
    const uint32_t per = (uint32_t)(kFusedRows * kFusedWaves);
    hipLaunchKernelGGL(MlpFusedForwardKernel<true>, dim3((rows + per - 1) / per), dim3(64 * kFusedWaves), 0, s, nullptr,
                       hist, rows,
                       m->Pb, m->P, y, nullptr, nullptr);
    return CheckLaunch("MlpFusedForwardKernel");
}

// Loss and gradients of the mean loss over a batch; gradients into m->G.
static int32_t Backward(hwpa_initializer* m, const uint16_t* x, const float* t, const uint8_t* valid, uint32_t batch,
                        hipStream_t s)
{
    int32_t rc = Forward(m, x, batch, m->Y, m->H1, m->H2, s);
    if (rc != HWPA_OK) return rc;
    hipLaunchKernelGGL(LossKernel, dim3((batch + 255) / 256), dim3(256), 0, s, m->Y, t, valid, batch, m->omap, m->dY,
                       m->dYb, m->lossPer);
    if ((rc = CheckLaunch("LossKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(MeanKernel, dim3(1), dim3(256), 0, s, m->lossPer, batch, m->lossDev);
    if ((rc = CheckLaunch("MeanKernel")) != HWPA_OK) return rc;
    const uint16_t* Wb = m->Pb;
    const int B = (int)batch;
    // dW3[o][h] = sum_b dY[b][o] H2[b][h]
    GemmArgs g = MakeGemm(MO, MH, B, m->dYb, 1, MO, m->H2, MH, 1);
    if ((rc = GemmSplitK<false, false>(g, m->kPart, m->G + OFF_W3, s)) != HWPA_OK) return rc;
    if ((rc = ColSum(m->dY, false, batch, MO, m->colPart, m->G + OFF_B3, s)) != HWPA_OK) return rc;
    // dH2[b][h] = (sum_o dY[b][o] W3[o][h]) * [H2 > 0]
    g = MakeGemm(B, MH, MO, m->dYb, MO, 1, Wb + OFF_W3, MH, 1);
    g.mask = m->H2;
    g.mk_m = MH;
    g.mk_n = 1;
    g.Cb = m->dH2;
    g.c_m = MH;
    g.c_n = 1;
    if ((rc = Gemm<false, false, EPI_MASK_BF16>(g, s)) != HWPA_OK) return rc;
    // dW2[o][i] = sum_b dH2[b][o] H1[b][i]
    g = MakeGemm(MH, MH, B, m->dH2, 1, MH, m->H1, MH, 1);
    if ((rc = GemmSplitK<false, false>(g, m->kPart, m->G + OFF_W2, s)) != HWPA_OK) return rc;
    if ((rc = ColSum(m->dH2, true, batch, MH, m->colPart, m->G + OFF_B2, s)) != HWPA_OK) return rc;
    // dH1[b][i] = (sum_o dH2[b][o] W2[o][i]) * [H1 > 0]
    g = MakeGemm(B, MH, MH, m->dH2, MH, 1, Wb + OFF_W2, MH, 1);
    g.mask = m->H1;
    g.mk_m = MH;
    g.mk_n = 1;
    g.Cb = m->dH1;
    g.c_m = MH;
    g.c_n = 1;
    if ((rc = Gemm<true, false, EPI_MASK_BF16>(g, s)) != HWPA_OK) return rc;
    // dW1[o][i] = sum_b dH1[b][o] X[b][i]
    g = MakeGemm(MH, MI, B, m->dH1, 1, MH, x, MI, 1);
    if ((rc = GemmSplitK<false, false>(g, m->kPart, m->G + OFF_W1, s)) != HWPA_OK) return rc;
    return ColSum(m->dH1, true, batch, MH, m->colPart, m->G + OFF_B1, s);
}

// Deterministic He-uniform initialisation from a 64-bit seed (biases zero).
static void InitParams(std::vector<float>& p, uint64_t seed)
{
    p.assign(NPARAM, 0.0f);
    uint64_t state = seed ^ 0x6a09e667f3bcc909ull;
    auto next = [&]() {
        state += 0x9e3779b97f4a7c15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](size_t off, int fanOut, int fanIn) {
        const double lim = std::sqrt(6.0 / fanIn);
        for (size_t i = 0; i < (size_t)fanOut * fanIn; ++i) p[off + i] = (float)((2.0 * next() - 1.0) * lim);
    };
    fill(OFF_W1, MH, MI);
    fill(OFF_W2, MH, MH);
    fill(OFF_W3, MO, MH);
}

int32_t MlpFeatures(const hwpa_lane_hist* hist, uint32_t lanes, uint16_t* x, hipStream_t s)
{
    hipLaunchKernelGGL(FeaturesKernel, dim3(lanes), dim3(LS), 0, s, hist, x);
    return CheckLaunch("FeaturesKernel");
}

const float* MlpOutputMap(const hwpa_initializer* m) { return m->omap; }

int32_t MlpTargets(const hwpa_lane_hist* hist, const float* diag, uint32_t lanes, float minConf, float* t,
                   uint8_t* valid, hipStream_t s)
{
    hipLaunchKernelGGL(TargetsKernel, dim3(lanes), dim3(LS), 0, s, hist, diag, minConf, t, valid);
    return CheckLaunch("TargetsKernel");
}

int32_t MlpForward(const hwpa_initializer* m, const uint16_t* x, uint32_t rows, float* y, uint16_t* h1, uint16_t* h2,
                   hipStream_t s)
{
    return Forward(m, x, rows, y, h1, h2, s);
}

int32_t MlpForwardFromHist(const hwpa_initializer* m, const hwpa_lane_hist* hist, uint32_t lanes, float* y,
                           hipStream_t s)
{
    return ForwardFromHist(m, hist, lanes, y, s);
}

}  // namespace hwpa

extern "C" {

int32_t hwpa_initializer_create(uint64_t seed, hwpa_initializer** out)
{
    if (!out) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    hwpa_initializer* m = new (std::nothrow) hwpa_initializer();
    if (!m) return Fail(HWPA_ERR_OOM, "out of host memory");
    int32_t rc = DeviceAlloc((void**)&m->P, NPARAM * 4, "mlp parameters");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->Pb, NPARAM * 2 + 16, "mlp parameters");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->G, NPARAM * 4, "mlp gradients");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->Mo, NPARAM * 4, "mlp moments");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->Ve, NPARAM * 4, "mlp moments");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->lossDev, 16, "mlp loss");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->omap, 20 * sizeof(float), "mlp output map");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&m->kPart, (size_t)kMaxSplit * MH * MH * 4, "mlp split-K partials");
    if (rc != HWPA_OK) {
        hwpa_initializer_destroy(m);
        return rc;
    }
    // Identity output map, unit weights, until the trainer sets one.
    for (int j = 0; j < 6; ++j) {
        m->omapHost[j] = 0.0f;
        m->omapHost[6 + j] = 1.0f;
        m->omapHost[12 + j] = 1.0f;
    }
    m->omapHost[18] = 1.0f;
    m->omapHost[19] = 1.0f;
    if (hipMemcpy(m->omap, m->omapHost, sizeof(m->omapHost), hipMemcpyHostToDevice) != hipSuccess) {
        hwpa_initializer_destroy(m);
        return Fail(HWPA_ERR_HIP, "output map upload failed");
    }
    std::vector<float> p;
    InitParams(p, seed);
    if ((rc = hwpa_initializer_set_params(m, p.data(), p.size())) != HWPA_OK) {
        hwpa_initializer_destroy(m);
        return rc;
    }
    *out = m;
    return HWPA_OK;
}

int32_t hwpa_initializer_destroy(hwpa_initializer* m)
{
    if (!m) return HWPA_OK;
    (void)hipDeviceSynchronize();
    FreeActs(m);
    FreeTrain(m);
    DeviceFree(m->P);
    DeviceFree(m->Pb);
    DeviceFree(m->G);
    DeviceFree(m->Mo);
    DeviceFree(m->Ve);
    DeviceFree(m->lossDev);
    DeviceFree(m->omap);
    DeviceFree(m->kPart);
    delete m;
    return HWPA_OK;
}

int32_t hwpa_initializer_get_params(const hwpa_initializer* m, float* host_params, size_t n)
{
    if (!m || !host_params) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    if (n != NPARAM) return Fail(HWPA_ERR_INVALID_ARG, "parameter count must be HWPA_MLP_PARAMS");
    HWPA_HIP(hipDeviceSynchronize());
    HWPA_HIP(hipMemcpy(host_params, m->P, NPARAM * 4, hipMemcpyDeviceToHost));
    return HWPA_OK;
}

int32_t hwpa_initializer_set_params(hwpa_initializer* m, const float* host_params, size_t n)
{
    if (!m || !host_params) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    if (n != NPARAM) return Fail(HWPA_ERR_INVALID_ARG, "parameter count must be HWPA_MLP_PARAMS");
    HWPA_HIP(hipDeviceSynchronize());
    HWPA_HIP(hipMemcpy(m->P, host_params, NPARAM * 4, hipMemcpyHostToDevice));
    HWPA_HIP(hipMemset(m->Mo, 0, NPARAM * 4));
    HWPA_HIP(hipMemset(m->Ve, 0, NPARAM * 4));
    HWPA_HIP(hipMemset(m->G, 0, NPARAM * 4));
    m->step = 0;
    hipLaunchKernelGGL(ToBf16Kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, nullptr, m->P, m->Pb, (uint32_t)NPARAM);
    int32_t rc = CheckLaunch("ToBf16Kernel");
    if (rc != HWPA_OK) return rc;
    HWPA_HIP(hipDeviceSynchronize());
    return HWPA_OK;
}

int32_t hwpa_initializer_set_output_map(hwpa_initializer* m, const float offset[6], const float scale[6],
                                        const float weight[7])
{
    if (!m || !offset || !scale || !weight) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    for (int j = 0; j < 6; ++j) {
        if (!(scale[j] > 0.0f) || !hw_isfinite(offset[j]) || !(weight[j] >= 0.0f))
            return Fail(HWPA_ERR_INVALID_ARG, "output map: scales must be positive, weights non-negative");
    }
    if (!(weight[6] >= 0.0f) || !(weight[7] >= 0.0f))
        return Fail(HWPA_ERR_INVALID_ARG, "output map: weights must be non-negative");
    for (int j = 0; j < 6; ++j) {
        m->omapHost[j] = offset[j];
        m->omapHost[6 + j] = scale[j];
        m->omapHost[12 + j] = weight[j];
    }
    m->omapHost[18] = weight[6];
    m->omapHost[19] = weight[7];
    HWPA_HIP(hipDeviceSynchronize());
    HWPA_HIP(hipMemcpy(m->omap, m->omapHost, sizeof(m->omapHost), hipMemcpyHostToDevice));
    return HWPA_OK;
}

int32_t hwpa_initializer_get_output_map(const hwpa_initializer* m, float out[20])
{
    if (!m || !out) return Fail(HWPA_ERR_INVALID_ARG, "null argument");
    for (int j = 0; j < 20; ++j) out[j] = m->omapHost[j];
    return HWPA_OK;
}

int32_t hwpa_mlp_features(const hwpa_lane_hist* d_hist, uint32_t lanes, uint16_t* d_x, void* stream)
{
    if (!d_hist || !d_x || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp features: null argument or no lanes");
    if (((uintptr_t)d_x & 15) != 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp features: output must be 16-byte aligned");
    return MlpFeatures(d_hist, lanes, d_x, (hipStream_t)stream);
}

int32_t hwpa_mlp_targets(const hwpa_lane_hist* d_hist, const float* d_diag, uint32_t lanes, float min_confidence,
                         float* d_t, uint8_t* d_valid, void* stream)
{
    if (!d_hist || !d_diag || !d_t || !d_valid || lanes == 0)
        return Fail(HWPA_ERR_INVALID_ARG, "mlp targets: null argument or no lanes");
    return MlpTargets(d_hist, d_diag, lanes, min_confidence, d_t, d_valid, (hipStream_t)stream);
}

int32_t hwpa_mlp_forward(hwpa_initializer* m, const uint16_t* d_x, uint32_t rows, float* d_y, void* stream)
{
    if (!m || !d_x || !d_y || rows == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp forward: null argument or no rows");
    if (((uintptr_t)d_x & 15) != 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp forward: input must be 16-byte aligned");
    // Inference keeps the hidden activations on chip (the fused kernel writes them only for training).
    return Forward(m, d_x, rows, d_y, nullptr, nullptr, (hipStream_t)stream);
}

int32_t hwpa_mlp_forward_hist(hwpa_initializer* m, const hwpa_lane_hist* d_hist, uint32_t lanes, float* d_y,
                              void* stream)
{
    if (!m || !d_hist || !d_y || lanes == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp forward: null argument or no lanes");
    return ForwardFromHist(m, d_hist, lanes, d_y, (hipStream_t)stream);
}

int32_t hwpa_mlp_gradients(hwpa_initializer* m, const uint16_t* d_x, const float* d_t, const uint8_t* d_valid,
                           uint32_t batch, float* d_grads, float* d_loss, void* stream)
{
    if (!m || !d_x || !d_t || batch == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp gradients: null argument or empty batch");
    if (((uintptr_t)d_x & 15) != 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp gradients: input must be 16-byte aligned");
    int32_t rc = EnsureBatch(m, batch);
    if (rc != HWPA_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = Backward(m, d_x, d_t, d_valid, batch, s)) != HWPA_OK) return rc;
    if (d_grads) HWPA_HIP(hipMemcpyAsync(d_grads, m->G, NPARAM * 4, hipMemcpyDeviceToDevice, s));
    if (d_loss) HWPA_HIP(hipMemcpyAsync(d_loss, m->lossDev, 4, hipMemcpyDeviceToDevice, s));
    return HWPA_OK;
}

int32_t hwpa_mlp_train_step(hwpa_initializer* m, const uint16_t* d_x, const float* d_t, const uint8_t* d_valid,
                            uint32_t batch, float lr, float* d_loss, void* stream)
{
    if (!m || !d_x || !d_t || batch == 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp train: null argument or empty batch");
    if (((uintptr_t)d_x & 15) != 0) return Fail(HWPA_ERR_INVALID_ARG, "mlp train: input must be 16-byte aligned");
    if (!(lr > 0.0f)) return Fail(HWPA_ERR_INVALID_ARG, "mlp train: learning rate must be positive");
    int32_t rc = EnsureBatch(m, batch);
    if (rc != HWPA_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = Backward(m, d_x, d_t, d_valid, batch, s)) != HWPA_OK) return rc;
    m->step += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - std::pow(b1, (double)m->step));
    const float c2 = (float)(1.0 - std::pow(b2, (double)m->step));
    hipLaunchKernelGGL(AdamKernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, m->P, m->G, m->Mo, m->Ve, m->Pb,
                       (uint32_t)NPARAM, lr, (float)b1, (float)b2, c1, c2, 1e-8f);
    if ((rc = CheckLaunch("AdamKernel")) != HWPA_OK) return rc;
    if (d_loss) HWPA_HIP(hipMemcpyAsync(d_loss, m->lossDev, 4, hipMemcpyDeviceToDevice, s));
    return HWPA_OK;
}

}  // extern "C"
