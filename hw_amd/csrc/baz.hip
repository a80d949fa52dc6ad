// baz.hip -- BAZ packet encoding of a chunk's pulses on the GPU.
//
// The reference serialises each ZMW's pulses with a stateful ObjectToBaz
// encoder (bazio/encoding/ObjectToBaz.h) whose field transforms and
// serialisers are bazio/encoding/FieldTransforms.h and FieldSerializers.h,
// one encoder per ZMW kept across chunks (appModules/PrelimHQFilter.cpp:
// serializers_, :148-160), writing into per-ZMW packet buffers
// (bazio/writing/PacketBuffer.h:142-200).  Two layouts are used
// (basecaller/dataTypes/PulseGroups.h:41-94):
//   ProductionPulses: {Label 2b truncate, PulseWidth 7b compact, StartFrame
//                      delta 7b compact}                      -- bases only
//   InternalPulses:   the same group, then Pkmax, Pkmid, Pkmean (signed
//                      fixed-point x10 in 2 bytes, 8b simple overflow), Pkvar
//                      (unsigned, 7b simple overflow) grouped with IsBase (1b)
//                                                             -- every pulse
// A group's fields are packed LSB first into ceil(bits/8) little-endian bytes,
// followed by the overflow bytes of its fields in field order.
//
// Here the whole chunk becomes one contiguous stream ordered by ZMW: a size
// pass (one wave per lane of 64 ZMWs), an exclusive scan of the lane sizes,
// and an encode pass in which each wave writes its lane's ZMWs in order and
// keeps every ZMW's delta state (last start frame) for the next chunk.  When the
// stream would not fit the caller's buffer nothing is written and the state
// is left unchanged, so the call can be repeated with a larger buffer.

#include "internal.h"

#include <new>

namespace hwpa {

constexpr int kBazScanBlock = 1024;

struct BazPulseFields {
    uint32_t label, width, midS, maxS, meanS;
    uint32_t start;
    float m2;
    bool reject;
};

__device__ __forceinline__ BazPulseFields BazLoad(const uint32_t* r)
{
    BazPulseFields f;
    f.start = r[0];
    f.width = r[1] & 0xffffu;
    f.meanS = r[1] >> 16;
    f.midS = r[2] & 0xffffu;
    f.maxS = r[2] >> 16;
    f.m2 = __uint_as_float(r[3]);
    f.label = r[4] & 0xffu;
    f.reject = ((r[4] >> 8) & 0xffu) != 0u;
    return f;
}

// CompactOverflow::OverflowBytes, unsigned (FieldSerializers.h): 7-bit groups above the
// main portion's numBits - 1 value bits.
template <int BITS>
__device__ __forceinline__ uint32_t CompactBytes(uint64_t t)
{
    t >>= (BITS - 1);
    uint32_t n = 0;
    while (t) {
        ++n;
        t >>= 7;
    }
    return n;
}

// FloatFixedCodec::Apply (FieldTransforms.h) with scale 10 in 2 bytes; special codes
// NaN 0, +inf 1, -inf 2, finite values offset by 3 when non-negative.
template <bool SIGNED>
__device__ __forceinline__ uint64_t FloatFixed10x2(float f)
{
    float scaled = roundf(f * 10.0f);
    if (SIGNED) {
        const int64_t mx = (int64_t)((1ull << 15) - 1 - 3);
        if ((float)(-mx) <= scaled && (float)mx >= scaled) {
            if (scaled >= 0.0f) scaled += 3.0f;
            return (uint64_t)(int64_t)scaled;
        }
    } else {
        const uint64_t mx = (1ull << 16) - 1 - 3;
        if (0.0f <= scaled && (float)mx >= scaled) {
            scaled += 3.0f;
            return (uint64_t)scaled;
        }
    }
    if (f != f) return 0;
    if (scaled < 0.0f) return 2;
    return 1;
}

// SimpleOverflow::NeedsOverflow (FieldSerializers.h): the all-ones main pattern marks overflow.
template <int BITS, bool SIGNED>
__device__ __forceinline__ bool SimpleNeeds(uint64_t v)
{
    if (SIGNED) v += (1ull << (BITS - 1)) - 1;
    return v >= (1ull << BITS) - 1;
}

// Float signal fields of the Mongo pulse (dataTypes/Pulse.h:104-126): fixed point x 0.1,
// MidSignal and SignalM2 NaN below 3 frames.
struct BazSignals {
    uint64_t pkmax, pkmid, pkmean, pkvar;
};
__device__ __forceinline__ BazSignals BazSignalCodes(const BazPulseFields& f)
{
    const float nan = __uint_as_float(0x7fc00000u);
    BazSignals s;
    s.pkmax = FloatFixed10x2<true>((float)f.maxS * 0.1f);
    s.pkmid = FloatFixed10x2<true>(f.width < 3u ? nan : (float)f.midS * 0.1f);
    s.pkmean = FloatFixed10x2<true>((float)f.meanS * 0.1f);
    s.pkvar = FloatFixed10x2<false>(f.width < 3u ? nan : f.m2);
    return s;
}

template <bool INTERNAL>
__device__ __forceinline__ uint32_t BazPulseBytes(const BazPulseFields& f, uint64_t delta)
{
    uint32_t n = 2u + CompactBytes<7>(f.width) + CompactBytes<7>(delta);
    if (INTERNAL) {
        const BazSignals s = BazSignalCodes(f);
        n += 4u;
        n += SimpleNeeds<8, true>(s.pkmax) ? 2u : 0u;
        n += SimpleNeeds<8, true>(s.pkmid) ? 2u : 0u;
        n += SimpleNeeds<8, true>(s.pkmean) ? 2u : 0u;
        n += SimpleNeeds<7, false>(s.pkvar) ? 2u : 0u;
    }
    return n;
}

// CompactOverflow::ToBinary (unsigned): returns the 7-bit main portion, writes the 7-bit
// continuation bytes.
__device__ __forceinline__ uint32_t CompactPut7(uint64_t v, uint8_t*& p)
{
    uint32_t n = CompactBytes<7>(v);
    uint32_t main = (uint32_t)(v & 0x3Fu);
    v >>= 6;
    if (n) main |= 0x40u;
    while (n) {
        uint8_t b = (uint8_t)(v & 0x7Fu);
        v >>= 7;
        --n;
        if (n) b |= 0x80u;
        *p++ = b;
    }
    return main;
}

// SimpleOverflow::ToBinary with 2 overflow bytes.
template <int BITS, bool SIGNED>
__device__ __forceinline__ uint32_t SimplePut(uint64_t v, uint8_t*& p)
{
    if (SimpleNeeds<BITS, SIGNED>(v)) {
        p[0] = (uint8_t)(v & 0xffu);
        p[1] = (uint8_t)((v >> 8) & 0xffu);
        p += 2;
        return (1u << BITS) - 1u;
    }
    if (SIGNED) v += (1ull << (BITS - 1)) - 1;
    return (uint32_t)v;
}

template <bool INTERNAL>
__device__ __forceinline__ void BazPutPulse(const BazPulseFields& f, uint64_t delta, uint8_t*& p)
{
    // Group 1: Label (2) | PulseWidth (7) | StartFrame delta (7) = 2 bytes, then overflows.
    uint8_t* head = p;
    p += 2;
    const uint32_t w = CompactPut7(f.width, p);
    const uint32_t d = CompactPut7(delta, p);
    const uint32_t g = (f.label & 3u) | ((w & 0x7Fu) << 2) | ((d & 0x7Fu) << 9);
    head[0] = (uint8_t)(g & 0xffu);
    head[1] = (uint8_t)(g >> 8);
    if (INTERNAL) {
        const BazSignals s = BazSignalCodes(f);
        // Pkmax, Pkmid, Pkmean: one byte each plus their own overflow bytes.
        uint64_t v3[3] = {s.pkmax, s.pkmid, s.pkmean};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint8_t* h = p++;
            *h = (uint8_t)SimplePut<8, true>(v3[k], p);
        }
        // Pkvar (7 bits) | IsBase (1 bit) in one byte, then Pkvar's overflow.
        uint8_t* h = p++;
        const uint32_t pv = SimplePut<7, false>(s.pkvar, p);
        *h = (uint8_t)((pv & 0x7Fu) | ((f.reject ? 0u : 1u) << 7));
    }
}

// One lane (64 ZMWs) per wave.  The wave walks the lane's ZMWs in order; for each ZMW
// thread t handles pulse slot t of a window of 64 slots: the records load coalesced (20 B
// apart), the StartFrame delta takes the previous selected pulse's start from the lane that
// holds it (the carried state for the first), and a wave prefix sum of the pulse sizes gives
// every pulse its byte position, so the ZMW's bytes are written by the wave side by side.
// ENCODE = false only sums the lane's bytes (size pass).
template <bool INTERNAL, bool ENCODE>
__global__ __launch_bounds__(256) void BazLaneKernel(const hwpa_pulse* __restrict__ pulses,
                                                     const uint32_t* __restrict__ counts, uint32_t maxCalls,
                                                     uint32_t lanes, uint64_t* __restrict__ last,
                                                     uint32_t* __restrict__ laneBytes,
                                                     const uint32_t* __restrict__ local,
                                                     const uint64_t* __restrict__ blockOff, uint8_t* __restrict__ out,
                                                     uint64_t capacity, uint64_t* __restrict__ offsets,
                                                     uint32_t* __restrict__ events)
{
    const uint32_t lane = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t t = threadIdx.x & 63u;
    if (lane >= lanes) return;
    uint64_t total = 0, run = 0;
    if (ENCODE) {
        total = blockOff[(lanes + kBazScanBlock - 1) / kBazScanBlock];
        run = blockOff[lane / kBazScanBlock] + local[lane];
        if (lane == 0 && t == 0) offsets[(size_t)lanes * LS] = total;
    }
    const bool write = ENCODE && total <= capacity;
    const uint64_t below = (t == 0) ? 0ull : (~0ull >> (64u - t));
    // The lane's counts and delta states, one per thread, broadcast per ZMW below.
    const size_t zt = (size_t)lane * LS + t;
    const uint32_t myN = counts[zt] < maxCalls ? counts[zt] : maxCalls;
    const uint64_t myPrev = last[zt];
    const uint32_t* lp = reinterpret_cast<const uint32_t*>(pulses + (size_t)lane * LS * maxCalls);
    const size_t zStride = (size_t)maxCalls * 5u;  // dwords per ZMW
    auto loadWindow = [&](uint32_t zi, uint32_t base, uint32_t n, uint32_t* r) {
        const uint32_t i = base + t;
#pragma unroll
        for (int k = 0; k < 5; ++k) r[k] = 0u;
        if (i < n) {
            const uint32_t* q = lp + zi * zStride + 5u * i;
#pragma unroll
            for (int k = 0; k < 5; ++k) r[k] = q[k];
        }
    };
    uint32_t cur[5];
    uint32_t nCur = __shfl(myN, 0, 64);
    loadWindow(0, 0, nCur, cur);
    uint64_t myOff = 0, myLast = 0;
    uint32_t myEv = 0;
    for (uint32_t zi = 0; zi < (uint32_t)LS; ++zi) {
        const uint32_t n = nCur;
        const uint32_t prevLo = __shfl((uint32_t)myPrev, (int)zi, 64);
        const uint32_t prevHi = __shfl((uint32_t)(myPrev >> 32), (int)zi, 64);
        uint64_t prev = ((uint64_t)prevHi << 32) | prevLo;
        // Next ZMW's first window in flight while this one is encoded.
        uint32_t nxt[5];
        const uint32_t zn = zi + 1 < (uint32_t)LS ? zi + 1 : zi;
        const uint32_t nNext = __shfl(myN, (int)zn, 64);
        loadWindow(zn, 0, zi + 1 < (uint32_t)LS ? nNext : 0u, nxt);
        const uint64_t zStart = run;
        uint32_t ev = 0;
        for (uint32_t base = 0; base < n; base += 64u) {
            uint32_t r[5];
            if (base == 0) {
#pragma unroll
                for (int k = 0; k < 5; ++k) r[k] = cur[k];
            } else {
                loadWindow(zi, base, n, r);
            }
            const uint32_t i = base + t;
            const BazPulseFields f = BazLoad(r);
            const bool sel = i < n && (INTERNAL || !f.reject);
            const uint64_t mask = __ballot(sel);
            const uint64_t mb = mask & below;
            const int src = mb ? 63 - __clzll((long long)mb) : 0;
            const uint32_t srcStart = __shfl(f.start, src, 64);
            const uint64_t pstart = mb ? (uint64_t)srcStart : prev;
            const uint64_t delta = (uint64_t)f.start - pstart;
            const uint32_t bytes = sel ? BazPulseBytes<INTERNAL>(f, delta) : 0u;
            uint32_t incl = bytes;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(incl, d, 64);
                if (t >= (uint32_t)d) incl += y;
            }
            if (write && sel) {
                uint8_t* p = out + run + (incl - bytes);
                BazPutPulse<INTERNAL>(f, delta, p);
            }
            if (mask) prev = (uint64_t)__shfl(f.start, 63 - __clzll((long long)mask), 64);
            run += __shfl(incl, 63, 64);
            ev += (uint32_t)__popcll(mask);
        }
        if (t == zi) {
            myOff = zStart;
            myEv = ev;
            myLast = prev;
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) cur[k] = nxt[k];
        nCur = nNext;
    }
    if (!ENCODE) {
        if (t == 0) laneBytes[lane] = (uint32_t)run;
        return;
    }
    const size_t z = (size_t)lane * LS + t;
    offsets[z] = myOff;
    events[z] = myEv;
    if (write) last[z] = myLast;
}

// Block-local exclusive scan of the sizes (1024 per block) and the block totals.
__global__ __launch_bounds__(kBazScanBlock) void BazScanBlocksKernel(const uint32_t* __restrict__ sizes, uint32_t zmws,
                                                                     uint32_t* __restrict__ local,
                                                                     uint64_t* __restrict__ blockSums)
{
    __shared__ uint32_t waveSum[kBazScanBlock / 64];
    const uint32_t i = blockIdx.x * kBazScanBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t v = i < zmws ? sizes[i] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) waveSum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t s = lane < kBazScanBlock / 64 ? waveSum[lane] : 0u;
#pragma unroll
        for (int d = 1; d < kBazScanBlock / 64; d <<= 1) {
            const uint32_t y = __shfl_up(s, d, 64);
            if (lane >= (uint32_t)d) s += y;
        }
        if (lane < kBazScanBlock / 64) waveSum[lane] = s;  // inclusive over waves
    }
    __syncthreads();
    const uint32_t before = wave ? waveSum[wave - 1] : 0u;
    if (i < zmws) local[i] = before + x - v;
    if (threadIdx.x == kBazScanBlock - 1) blockSums[blockIdx.x] = waveSum[kBazScanBlock / 64 - 1];
}

// Exclusive scan of the block totals in one workgroup; total at blockSums[nb].
__global__ __launch_bounds__(kBazScanBlock) void BazScanTotalsKernel(uint64_t* __restrict__ blockSums, uint32_t nb)
{
    __shared__ uint64_t waveSum[kBazScanBlock / 64];
    __shared__ uint64_t carry;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += kBazScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t v = i < nb ? blockSums[i] : 0ull;
        uint64_t x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t y = __shfl_up(x, d, 64);
            if (lane >= (uint32_t)d) x += y;
        }
        if (lane == 63) waveSum[wave] = x;
        __syncthreads();
        if (wave == 0) {
            uint64_t s = lane < kBazScanBlock / 64 ? waveSum[lane] : 0ull;
#pragma unroll
            for (int d = 1; d < kBazScanBlock / 64; d <<= 1) {
                const uint64_t y = __shfl_up(s, d, 64);
                if (lane >= (uint32_t)d) s += y;
            }
            if (lane < kBazScanBlock / 64) waveSum[lane] = s;
        }
        __syncthreads();
        const uint64_t c = carry;
        const uint64_t before = wave ? waveSum[wave - 1] : 0ull;
        if (i < nb) blockSums[i] = c + before + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + waveSum[kBazScanBlock / 64 - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) blockSums[nb] = carry;
}

__global__ void BazZeroKernel(uint64_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace hwpa

using namespace hwpa;

struct hwpa_baz_encoder {
    uint32_t zmws = 0;
    int32_t internal = 0;
    uint64_t* last = nullptr;       // per ZMW: last start frame serialised (DeltaCompression state)
    uint32_t* sizes = nullptr;      // per lane bytes of the current chunk
    uint32_t* local = nullptr;      // block-local exclusive lane offsets
    uint64_t* blockSums = nullptr;  // block offsets, then the total
};

extern "C" {

int32_t hwpa_baz_encoder_create(uint32_t zmws, int32_t internal_mode, hwpa_baz_encoder** out)
{
    if (!out || zmws == 0 || zmws % LS != 0)
        return Fail(HWPA_ERR_INVALID_ARG, "baz_encoder_create: zmws must be a positive multiple of 64");
    *out = nullptr;
    auto* e = new (std::nothrow) hwpa_baz_encoder();
    if (!e) return Fail(HWPA_ERR_OOM, "out of host memory");
    e->zmws = zmws;
    e->internal = internal_mode != 0;
    const uint32_t lanes = zmws / LS;
    const uint32_t nb = (lanes + kBazScanBlock - 1) / kBazScanBlock;
    int32_t rc = DeviceAlloc((void**)&e->last, sizeof(uint64_t) * zmws, "baz state");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&e->sizes, sizeof(uint32_t) * lanes, "baz lane sizes");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&e->local, sizeof(uint32_t) * lanes, "baz lane offsets");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&e->blockSums, sizeof(uint64_t) * (nb + 1), "baz block sums");
    if (rc == HWPA_OK) {
        hipLaunchKernelGGL(BazZeroKernel, dim3((zmws + 255) / 256), dim3(256), 0, nullptr, e->last, zmws);
        rc = CheckLaunch("BazZeroKernel");
        if (rc == HWPA_OK && hipDeviceSynchronize() != hipSuccess) rc = Fail(HWPA_ERR_HIP, "baz init failed");
    }
    if (rc != HWPA_OK) {
        hwpa_baz_encoder_destroy(e);
        return rc;
    }
    *out = e;
    return HWPA_OK;
}

int32_t hwpa_baz_encoder_destroy(hwpa_baz_encoder* e)
{
    if (!e) return HWPA_OK;
    DeviceFree(e->last);
    DeviceFree(e->sizes);
    DeviceFree(e->local);
    DeviceFree(e->blockSums);
    delete e;
    return HWPA_OK;
}

int32_t hwpa_baz_encoder_reset(hwpa_baz_encoder* e, void* stream)
{
    if (!e) return Fail(HWPA_ERR_INVALID_ARG, "null encoder");
    hipLaunchKernelGGL(BazZeroKernel, dim3((e->zmws + 255) / 256), dim3(256), 0, AsStream(stream), e->last, e->zmws);
    return CheckLaunch("BazZeroKernel");
}

int32_t hwpa_baz_encode(hwpa_baz_encoder* e, const hwpa_pulse* d_pulses, const uint32_t* d_counts, uint32_t max_calls,
                        uint8_t* d_out, uint64_t capacity, uint64_t* d_offsets, uint32_t* d_events, void* stream)
{
    if (!e || !d_pulses || !d_counts || !d_offsets || !d_events || max_calls == 0 || (!d_out && capacity != 0))
        return Fail(HWPA_ERR_INVALID_ARG, "baz_encode: bad argument");
    if (((uintptr_t)d_pulses & 3u) != 0) return Fail(HWPA_ERR_INVALID_ARG, "baz_encode: pulses must be 4-byte aligned");
    // A lane's bytes are counted in 32 bits: at most 64 x max_calls x 30 bytes.
    if (max_calls > (1u << 20)) return Fail(HWPA_ERR_INVALID_ARG, "baz_encode: max_calls above 2^20");
    hipStream_t s = AsStream(stream);
    const uint32_t L = e->zmws / (uint32_t)LS;
    const uint32_t nb = (L + kBazScanBlock - 1) / kBazScanBlock;
    const dim3 grid((L + 3) / 4), block(256);
    if (e->internal)
        hipLaunchKernelGGL((BazLaneKernel<true, false>), grid, block, 0, s, d_pulses, d_counts, max_calls, L, e->last,
                           e->sizes, nullptr, nullptr, nullptr, 0ull, nullptr, nullptr);
    else
        hipLaunchKernelGGL((BazLaneKernel<false, false>), grid, block, 0, s, d_pulses, d_counts, max_calls, L, e->last,
                           e->sizes, nullptr, nullptr, nullptr, 0ull, nullptr, nullptr);
    int32_t rc = CheckLaunch("BazLaneKernel (sizes)");
    if (rc != HWPA_OK) return rc;
    hipLaunchKernelGGL(BazScanBlocksKernel, dim3(nb), dim3(kBazScanBlock), 0, s, e->sizes, L, e->local, e->blockSums);
    if ((rc = CheckLaunch("BazScanBlocksKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(BazScanTotalsKernel, dim3(1), dim3(kBazScanBlock), 0, s, e->blockSums, nb);
    if ((rc = CheckLaunch("BazScanTotalsKernel")) != HWPA_OK) return rc;
    if (e->internal)
        hipLaunchKernelGGL((BazLaneKernel<true, true>), grid, block, 0, s, d_pulses, d_counts, max_calls, L, e->last,
                           nullptr, e->local, e->blockSums, d_out, capacity, d_offsets, d_events);
    else
        hipLaunchKernelGGL((BazLaneKernel<false, true>), grid, block, 0, s, d_pulses, d_counts, max_calls, L, e->last,
                           nullptr, e->local, e->blockSums, d_out, capacity, d_offsets, d_events);
    return CheckLaunch("BazLaneKernel (encode)");
}

}  // extern "C"
