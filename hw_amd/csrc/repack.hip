// repack.hip -- sensor packets into the analyzer's lane-major trace batches on the GPU.
//
// The reference's BlockRepacker (appModules/BlockRepacker.cpp:200-330, BatchFiller::Process)
// copies each dense SensorPacket -- [block][frame][blockWidth] samples of int16 or uint8,
// blocks of a cache-line multiple of ZMWs, starting at (startZmw, startFrame) -- into the
// overlapping part of a TraceBatch laid out [lane][frame][64].  The data type is kept.
//
// Here one launch handles every packet of a call: grid.y is the packet, each thread moves
// one 16-byte piece of a packet row (8 int16 or 16 uint8 ZMWs, never straddling a lane
// because packets and batches start on 16-ZMW boundaries) with one vector load and one
// vector store.  Pieces outside the batch are skipped, so a packet may overlap a batch
// partially, as in BatchFiller.

#include "internal.h"

#include <algorithm>
#include <new>
#include <vector>

namespace hwpa {

struct PacketDev {
    const uint8_t* data;
    uint32_t startZmw, startFrame, blocks, frames, blockWidth, pad;
};

__global__ __launch_bounds__(256) void RepackKernel(const PacketDev* __restrict__ packets, uint32_t elemBytes,
                                                    uint8_t* __restrict__ batch, uint32_t batchZmw0,
                                                    uint32_t batchFrame0, uint32_t lanes, uint32_t frames)
{
    const PacketDev pk = packets[blockIdx.y];
    const uint32_t zPerPiece = 16u / elemBytes;
    const uint32_t piecesPerRow = pk.blockWidth / zPerPiece;
    const uint64_t pieces = (uint64_t)pk.blocks * pk.frames * piecesPerRow;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= pieces) return;
    const uint32_t piece = (uint32_t)(idx % piecesPerRow);
    const uint64_t row = idx / piecesPerRow;  // block * frames + frame
    const uint32_t f = (uint32_t)(row % pk.frames);
    const uint32_t b = (uint32_t)(row / pk.frames);
    const uint32_t zmw = pk.startZmw + b * pk.blockWidth + piece * zPerPiece;
    const uint32_t frame = pk.startFrame + f;
    if (zmw < batchZmw0 || zmw - batchZmw0 >= lanes * (uint32_t)LS) return;
    if (frame < batchFrame0 || frame - batchFrame0 >= frames) return;
    const uint32_t zr = zmw - batchZmw0;
    const uint32_t lane = zr / (uint32_t)LS, z = zr % (uint32_t)LS;
    const uint4 v = *reinterpret_cast<const uint4*>(pk.data + idx * 16u);
    uint8_t* dst = batch + (((size_t)lane * frames + (frame - batchFrame0)) * LS + z) * elemBytes;
    *reinterpret_cast<uint4*>(dst) = v;
}

}  // namespace hwpa

using namespace hwpa;

// Descriptor slots used in rotation.  Each slot records an event after its kernel; a
// call that reuses the slot makes its stream wait for that event before the new
// descriptors are copied in, so calls on different streams never overwrite
// descriptors a kernel still reads.
constexpr int kRepackSlots = 4;

struct hwpa_repacker {
    uint32_t maxPackets = 0;
    PacketDev* d = nullptr;  // kRepackSlots x maxPackets
    std::vector<PacketDev> h;
    hipEvent_t done[kRepackSlots] = {};
    bool used[kRepackSlots] = {};
    int slot = 0;
};

extern "C" {

int32_t hwpa_repacker_create(uint32_t max_packets, hwpa_repacker** out)
{
    if (!out || max_packets == 0 || max_packets > 65535)
        return Fail(HWPA_ERR_INVALID_ARG, "repacker_create: max_packets must be in [1, 65535]");
    *out = nullptr;
    auto* r = new (std::nothrow) hwpa_repacker();
    if (!r) return Fail(HWPA_ERR_OOM, "out of host memory");
    r->maxPackets = max_packets;
    r->h.resize((size_t)kRepackSlots * max_packets);
    int32_t rc = DeviceAlloc((void**)&r->d, sizeof(PacketDev) * kRepackSlots * max_packets, "packet descriptors");
    for (int i = 0; i < kRepackSlots && rc == HWPA_OK; ++i)
        if (hipEventCreateWithFlags(&r->done[i], hipEventDisableTiming) != hipSuccess)
            rc = Fail(HWPA_ERR_HIP, "repacker_create: hipEventCreate failed");
    if (rc != HWPA_OK) {
        hwpa_repacker_destroy(r);
        return rc;
    }
    *out = r;
    return HWPA_OK;
}

int32_t hwpa_repacker_destroy(hwpa_repacker* r)
{
    if (!r) return HWPA_OK;
    for (int i = 0; i < kRepackSlots; ++i)
        if (r->done[i]) (void)hipEventDestroy(r->done[i]);
    DeviceFree(r->d);
    delete r;
    return HWPA_OK;
}

int32_t hwpa_repack(hwpa_repacker* r, const hwpa_packet_desc* packets, uint32_t n, int32_t encoding, void* d_batch,
                    uint32_t batch_first_zmw, uint32_t batch_first_frame, uint32_t batch_lanes, uint32_t batch_frames,
                    void* stream)
{
    if (!r || !packets || !d_batch || n == 0 || n > r->maxPackets || batch_lanes == 0 || batch_frames == 0)
        return Fail(HWPA_ERR_INVALID_ARG, "repack: bad argument");
    if (encoding != HWPA_ENCODING_INT16 && encoding != HWPA_ENCODING_UINT8)
        return Fail(HWPA_ERR_INVALID_ARG, "repack: unknown encoding");
    const uint32_t eb = encoding == HWPA_ENCODING_UINT8 ? 1u : 2u;
    const uint32_t zPerPiece = 16u / eb;
    if (batch_first_zmw % LS != 0 || ((uintptr_t)d_batch & 15u) != 0)
        return Fail(HWPA_ERR_INVALID_ARG, "repack: the batch must start on a lane and be 16-byte aligned");
    uint64_t maxPieces = 0;
    PacketDev* hs = r->h.data() + (size_t)r->slot * r->maxPackets;
    PacketDev* ds = r->d + (size_t)r->slot * r->maxPackets;
    for (uint32_t i = 0; i < n; ++i) {
        const hwpa_packet_desc& p = packets[i];
        // BlockRepacker requires blocks of whole cache lines (BlockRepacker.cpp:261-265).
        if (!p.data || ((uintptr_t)p.data & 15u) != 0 || p.block_width == 0 || p.block_width % zPerPiece != 0 ||
            p.start_zmw % zPerPiece != 0 || p.blocks == 0 || p.frames == 0)
            return Fail(HWPA_ERR_INVALID_ARG, "repack: packet layout not supported (16-byte aligned data, block width "
                                              "and start ZMW multiples of 16 bytes of samples)");
        hs[i] = PacketDev{static_cast<const uint8_t*>(p.data), p.start_zmw, p.start_frame, p.blocks, p.frames,
                            p.block_width, 0u};
        maxPieces = std::max(maxPieces, (uint64_t)p.blocks * p.frames * (p.block_width / zPerPiece));
    }
    const uint64_t gx = (maxPieces + 255) / 256;
    if (gx > 0x7fffffffull) return Fail(HWPA_ERR_INVALID_ARG, "repack: packet too large");
    hipStream_t s = AsStream(stream);
    const int slot = r->slot;
    if (r->used[slot]) HWPA_HIP(hipStreamWaitEvent(s, r->done[slot], 0));
    HWPA_HIP(hipMemcpyAsync(ds, hs, sizeof(PacketDev) * n, hipMemcpyHostToDevice, s));
    r->slot = (slot + 1) % kRepackSlots;
    hipLaunchKernelGGL(RepackKernel, dim3((unsigned)gx, n), dim3(256), 0, s, ds, eb, static_cast<uint8_t*>(d_batch),
                       batch_first_zmw, batch_first_frame, batch_lanes, batch_frames);
    const int32_t rc = CheckLaunch("RepackKernel");
    if (rc != HWPA_OK) return rc;
    HWPA_HIP(hipEventRecord(r->done[slot], s));
    r->used[slot] = true;
    return HWPA_OK;
}

}  // extern "C"
