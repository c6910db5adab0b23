// nnh_host.hpp — host-side entry points of the product engine used by the
// device-resident natural layer (nnh_natural.cpp).  Pointers are device
// pointers; work is queued on st.  Callers hold host_mutex().
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "nnh_internal.hpp"

namespace nnh {

int host_init();                  // initialise device 0 if nnh_init was not called
hipStream_t host_stream();
std::mutex& host_mutex();
void host_release();              // free the natural layer's cached device buffers
void host_trim();                 // free the natural layer's idle cached buffers (caller holds host_mutex())
int set_error(int code, const std::string& msg);
// Order a call's queued work after the previous call's on another stream,
// and record the end of this call's work (the shared workspace).
int host_ws_acquire(hipStream_t st);
void host_ws_release(hipStream_t st);
// NNH_ERR_HIP (with its message set) when nnh_debug_fail_next asked for a failure.
int injected_failure();

// Caller's host buffer -> device (nnh_hostio.cpp): ordered on st; the host
// buffer has been read completely when the call returns.
int host_upload(void* d_dst, const void* h_src, u64 bytes, hipStream_t st);
// Device -> caller's host buffer, after the work queued on st; complete when
// the call returns.
int host_download(void* h_dst, const void* d_src, u64 bytes, hipStream_t st);
// Free the pinned staging slots and stop the copy threads (nnh_shutdown).
void hostio_release();

// out[0 .. an+bn) = a * b, the reference dispatch (natural.h:444-450).
int dev_mul(u64* out, const u64* a, u64 an, const u64* b, u64 bn, hipStream_t st);
// out[0 .. 2 an) = a^2 (natural.h:529-535).
int dev_sqr(u64* out, const u64* a, u64 an, hipStream_t st);
// a * b through a transform of 2^s words with the reference's save rule
// (ntt_info.h:237-248): when save_size - 1 > 2^s the product modulo
// B^(2^s) - 1 (2^s words), otherwise the exact product (save_size = an+bn
// words).  an, bn <= 2^s.  out holds max(an+bn, 2^s + 2) words.
int dev_mul_scale(u64* out, const u64* a, u64 an, const u64* b, u64 bn, int s, u64 save_size, u64* n_out,
                  hipStream_t st);

// Plan, enqueue, synchronise and release one batch.
int dev_batch(const IReq* reqs, u64 n, hipStream_t st);

// Register-resident butterfly rate of prime pr (nnh_peak.hip).
int butterfly_peak(int pr, u32 iters, double* rate);

}  // namespace nnh
