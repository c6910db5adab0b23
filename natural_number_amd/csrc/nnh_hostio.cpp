// nnh_hostio.cpp — transfers between the caller's host buffers and the
// device for the host-pointer entry points (nnh_mul, nnh_sqr, the natural
// layer's calls).
//
// The reference's naturals live in malloc'd (pageable) memory
// (basic_natural.h:16-34), which the DMA engines cannot read directly; the
// runtime's own pageable path moves it through one staging buffer at
// 30 GB/s host-to-device and 17.6 GB/s device-to-host on MI355X (the 2^30-bit
// product's 512 MiB of operands and result take 24 of its 28 ms).  Large
// transfers here go through a ring of pinned slots instead: the host copies
// chunk c+1 between the caller's buffer and a slot (split over six
// threads) while the DMA engine moves chunk c, each slot reused once the
// event recorded after its last DMA has completed.  Pinned DMA runs at
// 57 GB/s each way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/nnh.h"
#include "nnh_copyteam.hpp"
#include "nnh_host.hpp"

namespace nnh {

namespace {

constexpr u64 CHUNK = 16ull << 20;      // bytes per slot
constexpr int NSLOT = 4;                // slots in the ring (64 MiB pinned)
constexpr u64 STAGED_MIN = 8ull << 20;  // smaller transfers use the runtime's path

struct Ring {
    void* slot[NSLOT]{};
    hipEvent_t ev[NSLOT]{};
    bool busy[NSLOT]{};  // an event was recorded after the slot's last DMA
    CopyTeam<5> team;  // five copy threads beside the caller (faster into untouched result pages than three)
};
Ring g_ring;
std::mutex g_ring_mu;

int ring_ready(Ring& R) {
    for (int i = 0; i < NSLOT; i++) {
        if (!R.slot[i]) {
            hipError_t e = hipHostMalloc(&R.slot[i], CHUNK, hipHostMallocDefault);
            if (e != hipSuccess) {
                R.slot[i] = nullptr;
                return set_error(e == hipErrorOutOfMemory ? NNH_ERR_NOMEM : NNH_ERR_HIP,
                                 std::string("pinned staging: ") + hipGetErrorString(e));
            }
        }
        if (!R.ev[i]) {
            hipError_t e = hipEventCreateWithFlags(&R.ev[i], hipEventDisableTiming);
            if (e != hipSuccess) {
                R.ev[i] = nullptr;
                return set_error(NNH_ERR_HIP, std::string("staging event: ") + hipGetErrorString(e));
            }
        }
    }
    return NNH_OK;
}

// Wait until slot i's last DMA has completed.
int slot_wait(Ring& R, int i) {
    if (!R.busy[i]) return NNH_OK;
    hipError_t e = hipEventSynchronize(R.ev[i]);
    R.busy[i] = false;
    return e == hipSuccess ? NNH_OK : set_error(NNH_ERR_HIP, std::string("staging wait: ") + hipGetErrorString(e));
}

int hip_rc(hipError_t e, const char* what) {
    return e == hipSuccess ? NNH_OK : set_error(NNH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// The runtime's own pageable copies, for small transfers and when the ring
// cannot be set up (no pinned memory left): slower, never wrong.
static int plain_upload(void* d_dst, const void* h_src, u64 bytes, hipStream_t st) {
    return hip_rc(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, st), "upload");
}
static int plain_download(void* h_dst, const void* d_src, u64 bytes, hipStream_t st) {
    if (int rc = hip_rc(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, st), "download")) return rc;
    return hip_rc(hipStreamSynchronize(st), "download");
}

int host_upload(void* d_dst, const void* h_src, u64 bytes, hipStream_t st) {
    if (!bytes) return NNH_OK;
    if (bytes < STAGED_MIN) return plain_upload(d_dst, h_src, bytes, st);
    std::lock_guard<std::mutex> g(g_ring_mu);
    Ring& R = g_ring;
    if (ring_ready(R) != NNH_OK) return plain_upload(d_dst, h_src, bytes, st);
    const u64 nchunk = (bytes + CHUNK - 1) / CHUNK;
    for (u64 c = 0; c < nchunk; c++) {
        const int i = (int)(c % NSLOT);
        const u64 off = c * CHUNK, n = std::min<u64>(CHUNK, bytes - off);
        if (int rc = slot_wait(R, i)) return rc;
        R.team.copy(R.slot[i], (const char*)h_src + off, n);
        if (int rc = hip_rc(hipMemcpyAsync((char*)d_dst + off, R.slot[i], n, hipMemcpyHostToDevice, st), "upload"))
            return rc;
        if (int rc = hip_rc(hipEventRecord(R.ev[i], st), "upload event")) return rc;
        R.busy[i] = true;
    }
    return NNH_OK;
}

int host_download(void* h_dst, const void* d_src, u64 bytes, hipStream_t st) {
    if (!bytes) return NNH_OK;
    if (bytes < STAGED_MIN) return plain_download(h_dst, d_src, bytes, st);
    std::lock_guard<std::mutex> g(g_ring_mu);
    Ring& R = g_ring;
    if (ring_ready(R) != NNH_OK) return plain_download(h_dst, d_src, bytes, st);
    const u64 nchunk = (bytes + CHUNK - 1) / CHUNK;
    auto issue = [&](u64 c) -> int {
        const int i = (int)(c % NSLOT);
        const u64 off = c * CHUNK, n = std::min<u64>(CHUNK, bytes - off);
        if (int rc = slot_wait(R, i)) return rc;
        if (int rc = hip_rc(hipMemcpyAsync(R.slot[i], (const char*)d_src + off, n, hipMemcpyDeviceToHost, st),
                            "download"))
            return rc;
        if (int rc = hip_rc(hipEventRecord(R.ev[i], st), "download event")) return rc;
        R.busy[i] = true;
        return NNH_OK;
    };
    for (u64 c = 0; c < std::min<u64>(nchunk, NSLOT); c++)
        if (int rc = issue(c)) return rc;
    for (u64 c = 0; c < nchunk; c++) {
        const int i = (int)(c % NSLOT);
        const u64 off = c * CHUNK, n = std::min<u64>(CHUNK, bytes - off);
        if (int rc = slot_wait(R, i)) return rc;
        R.team.copy((char*)h_dst + off, R.slot[i], n);
        if (c + NSLOT < nchunk)
            if (int rc = issue(c + NSLOT)) return rc;
    }
    return NNH_OK;
}

void hostio_release() {
    std::lock_guard<std::mutex> g(g_ring_mu);
    Ring& R = g_ring;
    R.team.stop();
    for (int i = 0; i < NSLOT; i++) {
        if (R.busy[i]) (void)hipEventSynchronize(R.ev[i]);
        R.busy[i] = false;
        if (R.slot[i]) (void)hipHostFree(R.slot[i]);
        if (R.ev[i]) (void)hipEventDestroy(R.ev[i]);
        R.slot[i] = nullptr;
        R.ev[i] = nullptr;
    }
}

}  // namespace nnh
