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
// pass (one thread per ZMW), an exclusive scan of the sizes, and an encode
// pass in which every thread writes its ZMW's bytes at its offset and keeps
// the ZMW's delta state (last start frame) for the next chunk.  When the
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

// Size pass: bytes and events per ZMW of this chunk (production: bases only).
template <bool INTERNAL>
__global__ __launch_bounds__(256) void BazSizeKernel(const hwpa_pulse* __restrict__ pulses,
                                                     const uint32_t* __restrict__ counts, uint32_t maxCalls,
                                                     uint32_t zmws, const uint64_t* __restrict__ last,
                                                     uint32_t* __restrict__ sizes, uint32_t* __restrict__ events)
{
    const uint32_t z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= zmws) return;
    const uint32_t n = counts[z] < maxCalls ? counts[z] : maxCalls;
    const uint32_t* zp = reinterpret_cast<const uint32_t*>(pulses + (size_t)z * maxCalls);
    uint64_t prev = last[z];
    uint32_t bytes = 0, ev = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const BazPulseFields f = BazLoad(zp + 5 * i);
        if (!INTERNAL && f.reject) continue;
        const uint64_t s = f.start;
        bytes += BazPulseBytes<INTERNAL>(f, s - prev);
        prev = s;
        ++ev;
    }
    sizes[z] = bytes;
    events[z] = ev;
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

// Encode pass: every ZMW's bytes at its offset; offsets[z] (and offsets[zmws] = total)
// for the caller; the delta state advances only when the whole stream fits.
template <bool INTERNAL>
__global__ __launch_bounds__(256) void BazEncodeKernel(const hwpa_pulse* __restrict__ pulses,
                                                       const uint32_t* __restrict__ counts, uint32_t maxCalls,
                                                       uint32_t zmws, uint64_t* __restrict__ last,
                                                       const uint32_t* __restrict__ local,
                                                       const uint64_t* __restrict__ blockOff, uint8_t* __restrict__ out,
                                                       uint64_t capacity, uint64_t* __restrict__ offsets)
{
    const uint32_t z = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = blockOff[(zmws + kBazScanBlock - 1) / kBazScanBlock];
    if (z == 0) offsets[zmws] = total;
    if (z >= zmws) return;
    const uint64_t off = blockOff[z / kBazScanBlock] + local[z];
    offsets[z] = off;
    if (total > capacity) return;
    const uint32_t n = counts[z] < maxCalls ? counts[z] : maxCalls;
    const uint32_t* zp = reinterpret_cast<const uint32_t*>(pulses + (size_t)z * maxCalls);
    uint8_t* p = out + off;
    uint64_t prev = last[z];
    for (uint32_t i = 0; i < n; ++i) {
        const BazPulseFields f = BazLoad(zp + 5 * i);
        if (!INTERNAL && f.reject) continue;
        const uint64_t s = f.start;
        BazPutPulse<INTERNAL>(f, s - prev, p);
        prev = s;
    }
    last[z] = prev;
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
    uint32_t* sizes = nullptr;      // per ZMW bytes of the current chunk
    uint32_t* local = nullptr;      // block-local exclusive offsets
    uint64_t* blockSums = nullptr;  // block offsets, then the total
};

extern "C" {

int32_t hwpa_baz_encoder_create(uint32_t zmws, int32_t internal_mode, hwpa_baz_encoder** out)
{
    if (!out || zmws == 0) return Fail(HWPA_ERR_INVALID_ARG, "baz_encoder_create: bad argument");
    *out = nullptr;
    auto* e = new (std::nothrow) hwpa_baz_encoder();
    if (!e) return Fail(HWPA_ERR_OOM, "out of host memory");
    e->zmws = zmws;
    e->internal = internal_mode != 0;
    const uint32_t nb = (zmws + kBazScanBlock - 1) / kBazScanBlock;
    int32_t rc = DeviceAlloc((void**)&e->last, sizeof(uint64_t) * zmws, "baz state");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&e->sizes, sizeof(uint32_t) * zmws, "baz sizes");
    if (rc == HWPA_OK) rc = DeviceAlloc((void**)&e->local, sizeof(uint32_t) * zmws, "baz offsets");
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
    hipStream_t s = AsStream(stream);
    const uint32_t Z = e->zmws;
    const uint32_t nb = (Z + kBazScanBlock - 1) / kBazScanBlock;
    const dim3 g256((Z + 255) / 256);
    if (e->internal)
        hipLaunchKernelGGL(BazSizeKernel<true>, g256, dim3(256), 0, s, d_pulses, d_counts, max_calls, Z, e->last,
                           e->sizes, d_events);
    else
        hipLaunchKernelGGL(BazSizeKernel<false>, g256, dim3(256), 0, s, d_pulses, d_counts, max_calls, Z, e->last,
                           e->sizes, d_events);
    int32_t rc = CheckLaunch("BazSizeKernel");
    if (rc != HWPA_OK) return rc;
    hipLaunchKernelGGL(BazScanBlocksKernel, dim3(nb), dim3(kBazScanBlock), 0, s, e->sizes, Z, e->local, e->blockSums);
    if ((rc = CheckLaunch("BazScanBlocksKernel")) != HWPA_OK) return rc;
    hipLaunchKernelGGL(BazScanTotalsKernel, dim3(1), dim3(kBazScanBlock), 0, s, e->blockSums, nb);
    if ((rc = CheckLaunch("BazScanTotalsKernel")) != HWPA_OK) return rc;
    if (e->internal)
        hipLaunchKernelGGL(BazEncodeKernel<true>, g256, dim3(256), 0, s, d_pulses, d_counts, max_calls, Z, e->last,
                           e->local, e->blockSums, d_out, capacity, d_offsets);
    else
        hipLaunchKernelGGL(BazEncodeKernel<false>, g256, dim3(256), 0, s, d_pulses, d_counts, max_calls, Z, e->last,
                           e->local, e->blockSums, d_out, capacity, d_offsets);
    return CheckLaunch("BazEncodeKernel");
}

}  // extern "C"
