// nnh_api.cpp — host driver and C ABI of libnnh.so.
//
// Owns the per-device state (root-base constants, Montgomery-form twiddle
// tables grown on demand like the reference's static NTT_info::root/iroot
// caches, ntt_info.h:22-41 and 141-145, and a growing workspace) and
// sequences the kernels of one product:
//   top passes (limbs -> residues, transform levels h >= 4096)
//   -> fused chunk kernel (remaining levels of A and B, base multiply,
//      inverse levels h < 4096)
//   -> inverse top passes -> Garner CRT + carry scan.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nnh.h"
#include "nnh_host.hpp"
#include "nnh_internal.hpp"
#include "nnh_nat.hpp"
#include "nnh_plan.hpp"

namespace nnh {

typedef unsigned __int128 u128;

constexpr u64 P[3] = {27ull << 56 | 1, 58ull << 56 | 1, 87ull << 56 | 1};
// Primitive 2^56-th roots of unity: the last entry of each row of the
// reference's root_base table (ntt_info.h:89, 98, 107); every other entry of
// that table is a repeated square of it.
constexpr u64 ROOT_GEN[3] = {11, 2, 310};

static u64 mulmod(u64 a, u64 b, u64 p) { return (u64)((u128)a * b % p); }
static u64 powmod(u64 b, u64 e, u64 p) {
    u64 r = 1;
    for (b %= p; e; e >>= 1, b = mulmod(b, b, p))
        if (e & 1) r = mulmod(r, b, p);
    return r;
}
static u64 to_mont(u64 x, u64 p) { return (u64)(((u128)(x % p) << 64) % p); }

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int set_error(int code, const std::string& msg) { return fail(code, msg); }
#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess)                                                      \
            return fail(e_ == hipErrorOutOfMemory ? NNH_ERR_NOMEM : NNH_ERR_HIP,   \
                        std::string(#x) + ": " + hipGetErrorString(e_));           \
    } while (0)

struct Buf {
    u64* p = nullptr;
    u64 words = 0;
    bool fixed = false;  // caller-provided memory of a set size: never freed or grown
};

// Montgomery-form twiddle tables of the three primes for transforms up to
// 2^scale words.  Shared: a batch keeps the tables its launches (or its
// captured graph) read, so growing the device's tables for a larger
// product, or nnh_shutdown, leaves them alive until the batch is destroyed.
struct Twiddles {
    int scale = 0;
    u64* f[3]{};
    u64* i[3]{};
    ~Twiddles() {
        (void)hipDeviceSynchronize();  // queued work on any stream may still read them
        for (int k = 0; k < 3; k++) {
            if (f[k]) (void)hipFree(f[k]);
            if (i[k]) (void)hipFree(i[k]);
        }
    }
};

// ---------------------------------------------------------------------------
// Allocation-free products (nnh_mul_device_noalloc): products of up to
// 2^NOALLOC_SCALE words through transforms of at most that length, with the
// twiddle tables, the transform workspace, the schoolbook columns / split
// addend and the CRT tile flags carved out of one file-scope device array.
// Each GPU holds its own copy for the life of the process (84 MiB); its
// address comes from hipGetSymbolAddress at nnh_init, per device.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_SCALE = 20;
constexpr u64 NA_TW = 1ull << (NOALLOC_SCALE - 2);       // entries per twiddle table (6 tables)
constexpr u64 NA_WS = 6ull << NOALLOC_SCALE;             // A and B transforms, 3 primes each
constexpr u64 NA_TMP = (3ull << NOALLOC_SCALE) + 64;     // schoolbook columns or split addend
constexpr u64 NA_TILES = 4096;                           // CRT tile carry functions (words)
constexpr u64 NA_WORDS = 6 * NA_TW + NA_WS + NA_TMP + NA_TILES;
constexpr int NA_MAX_DEVICES = 64;
static __device__ __align__(256) u64 scratch[NA_WORDS];

struct Device {
    int id = -1;
    hipStream_t stream = nullptr;
    // The transform stages of p2 run on stream2 beside those of p0/p1 on the
    // caller's stream (the prime chains are independent until the CRT), so
    // one chain's memory-heavy top passes overlap the other's VALU-bound
    // chunk kernel and the launch tails overlap.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Batches spread their buckets (independent until the shared CRT) over
    // the caller's (or capturing) stream and nwork worker streams (work[0] is
    // stream2; NNH_BATCH_STREAMS=0..4 sets nwork): one bucket's p0/p1 chain
    // on one lane, its p2 chain on the next, forked from and joined back into
    // the caller's stream; under graph capture they become parallel branches.
    // 3 workers by default, four lanes for the default 4 hardware queues:
    // the 4096-product mix 18.55-18.59 ms against 18.67-18.81 ms with 4
    // workers and 18.58-18.68 ms with 2 (workers against none: -3 %).  (Four
    // freshly created worker streams instead of three plus stream2 slowed
    // the two-stream product by 3 %.)
    static constexpr int NWORK = 4;
    int nwork = 3;
    hipStream_t work[NWORK]{};
    hipEvent_t ev_bfork = nullptr, ev_bjoin[NWORK]{};
    // Internal batches (dev_batch) upload their job tables from this pinned
    // staging buffer on the caller's stream and return without waiting;
    // ev_tab marks the last upload, which must finish before the staging
    // buffer is rewritten.
    u8* tab_host = nullptr;
    u64 tab_bytes = 0;
    hipEvent_t ev_tab = nullptr;
    bool tab_pending = false;
    bool two_streams = true;
    RootBases rb{};
    CrtConsts crt{};
    u64 rb_plain[3][56]{}, irb_plain[3][56]{};
    int tw_scale = 0;          // tables hold 2^(tw_scale-2) entries
    const u64* twf[3]{};       // the tables the next launches read (tw's, or fixed ones)
    const u64* twi[3]{};
    std::shared_ptr<Twiddles> tw;
    bool tw_fixed = false;     // twf/twi are fixed tables of tw_scale: never grown
    // The workspace below is shared by every call on the device.  Each call
    // records ev_ws on its stream after its last launch; a call on another
    // stream waits for that event before its first launch, so queued work of
    // two calls never overlaps in the workspace.
    hipEvent_t ev_ws = nullptr;
    hipStream_t ws_stream = nullptr;  // stream ev_ws was last recorded on (null: none pending)
    Buf ws;                    // 6 L words: A and B transforms
    Buf tiles;                 // tile carry functions and carry-ins (bytes)
    Buf stage;                 // host-API staging: a, b, dst
    Buf tmp;                   // split addend / cyclic words / columns
    Buf bws, bjobs, btiles;    // workspace of internal batches (reused across calls)
    // device address of `scratch` per device id (null until nnh_init on it)
    u64* na_ptr[NA_MAX_DEVICES]{};
    // tuned level splits per scale, for single products [0] and batch
    // buckets [1] (cb = 0: the formula)
    LevelPlan tuned[2][MAX_SCALE + 1];
    std::mutex mu;
};

static Device g_dev;
static bool g_inited = false;
static std::mutex g_init_mu;

// ---------------------------------------------------------------------------
// Optional per-kernel timing: HIP events recorded on the launch stream around
// each kernel family, summed when read (nnh_kernel_times).
// ---------------------------------------------------------------------------
enum KernelKind { K_TOP_LOAD = 0, K_TOP_FWD, K_MID, K_TOP_INV, K_CRT, K_SCHOOL, K_FOLD, K_NKINDS };
static const char* KIND_NAMES[K_NKINDS] = {"top_pass_load", "top_pass_fwd", "mid_fused", "top_pass_inv",
                                           "crt_carry", "school_cols", "cyclic_fold"};
struct ProfRec {
    hipEvent_t a, b;
    int kind;
    u64 bytes;
};
static bool g_prof = false;
static std::vector<ProfRec> g_prof_recs;
static double g_prof_ms[K_NKINDS];
static u64 g_prof_n[K_NKINDS];
static double g_prof_bytes[K_NKINDS];
struct ProfScope {
    ProfRec r{};
    hipStream_t st;
    bool on;
    ProfScope(int kind, u64 bytes, hipStream_t s) : st(s), on(g_prof) {
        if (!on) return;
        r.kind = kind;
        r.bytes = bytes;
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        (void)hipEventRecord(r.a, st);
    }
    ~ProfScope() {
        if (!on) return;
        (void)hipEventRecord(r.b, st);
        g_prof_recs.push_back(r);
    }
};
static void prof_collect() {
    for (ProfRec& r : g_prof_recs) {
        float ms = 0;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        g_prof_ms[r.kind] += ms;
        g_prof_n[r.kind] += 1;
        g_prof_bytes[r.kind] += (double)r.bytes;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
}

static int ensure(Buf& b, u64 words) {
    if (b.words >= words) return NNH_OK;
    if (b.fixed)
        return fail(NNH_ERR_ARG, "workspace of " + std::to_string(words) + " words exceeds the fixed limit of " +
                                     std::to_string(b.words));
    // queued work on any stream may still use the old buffer
    if (b.p) HIPCHK(hipDeviceSynchronize());
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.words = 0;
    u64 w = std::bit_ceil(std::max<u64>(words, 1024));
    hipError_t e = hipMalloc((void**)&b.p, w * sizeof(u64));
    if (e == hipErrorOutOfMemory) {
        // idle buffers cached by the natural layer's pool may hold the memory
        (void)hipGetLastError();
        host_trim();
        e = hipMalloc((void**)&b.p, w * sizeof(u64));
    }
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(e == hipErrorOutOfMemory ? NNH_ERR_NOMEM : NNH_ERR_HIP,
                    std::string("workspace allocation: ") + hipGetErrorString(e));
    }
    b.words = w;
    return NNH_OK;
}

// alpha_s with alpha^2 = 2^64 * 2^-(s-1) (mod p): scales every loaded
// residue so that the base multiply's Montgomery factor 2^-64 and the
// unscaled inverse's 2^(s-1) cancel.  sqrt(2) = z + z^-1 for a primitive
// 8th root z = rb[2].
static u64 alpha_plain(const Device& D, int i, int s) {
    u64 p = P[i];
    int e = 65 - s;
    u64 a = powmod(2, (u64)(e / 2), p);
    if (e & 1) {
        u64 sqrt2 = (D.rb_plain[i][2] + D.irb_plain[i][2]) % p;
        a = mulmod(a, sqrt2, p);
    }
    return a;
}

static void build_constants(Device& D) {
    for (int i = 0; i < 3; i++) {
        u64 p = P[i];
        D.rb_plain[i][55] = ROOT_GEN[i];
        for (int j = 54; j >= 0; j--) D.rb_plain[i][j] = mulmod(D.rb_plain[i][j + 1], D.rb_plain[i][j + 1], p);
        for (int j = 0; j < 56; j++) {
            D.irb_plain[i][j] = powmod(D.rb_plain[i][j], p - 2, p);
            D.rb.rb_m[i][j] = to_mont(D.rb_plain[i][j], p);
            D.rb.irb_m[i][j] = to_mont(D.irb_plain[i][j], p);
        }
        D.rb.one_m[i] = to_mont(1, p);
    }
    u64 m0 = P[0], m1 = P[1], m2 = P[2];
    D.crt.c02m = to_mont(powmod(m0, m2 - 2, m2), m2);
    const u64 c1 = powmod(mulmod(m0 % m1, m2 % m1, m1), m1 - 2, m1);
    D.crt.c1m = to_mont(c1, m1);
    D.crt.km = to_mont((m1 - mulmod(m0 % m1, c1, m1)) % m1, m1);
}

static int ensure_twiddles(Device& D, int s) {
    if (s <= D.tw_scale) return NNH_OK;
    if (D.tw_fixed)
        return fail(NNH_ERR_ARG, "transform of 2^" + std::to_string(s) + " words exceeds the fixed twiddle tables (2^" +
                                     std::to_string(D.tw_scale) + ")");
    int ns = std::max(s, 12);
    u64 n = 1ull << (ns - 2);
    auto T = std::make_shared<Twiddles>();
    for (int i = 0; i < 3; i++) {
        HIPCHK(hipMalloc((void**)&T->f[i], n * sizeof(u64)));
        HIPCHK(hipMalloc((void**)&T->i[i], n * sizeof(u64)));
        launch_twiddles(T->f[i], n, D.rb, i, 0, D.stream);
        launch_twiddles(T->i[i], n, D.rb, i, 1, D.stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(D.stream));
    T->scale = ns;
    // the previous tables are freed here unless a batch still holds them
    D.tw = std::move(T);
    for (int i = 0; i < 3; i++) {
        D.twf[i] = D.tw->f[i];
        D.twi[i] = D.tw->i[i];
    }
    D.tw_scale = ns;
    return NNH_OK;
}

static void drop_twiddles(Device& D) {
    D.tw.reset();
    for (int i = 0; i < 3; i++) D.twf[i] = D.twi[i] = nullptr;
    D.tw_scale = 0;
}

// Workspace hand-over between calls (see Device::ev_ws).  Streams being
// captured into a graph are left alone: their work runs when the caller
// replays the graph, which the caller orders.
static bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}
static int ws_acquire(Device& D, hipStream_t st) {
    if (!D.ws_stream || D.ws_stream == st || capturing(st)) return NNH_OK;
    HIPCHK(hipStreamWaitEvent(st, D.ev_ws, 0));
    return NNH_OK;
}
static void ws_release(Device& D, hipStream_t st) {
    if (capturing(st)) return;
    if (hipEventRecord(D.ev_ws, st) == hipSuccess) D.ws_stream = st;
}
struct WsScope {
    Device& D;
    hipStream_t st;
    int rc;
    WsScope(Device& d, hipStream_t s) : D(d), st(s), rc(ws_acquire(d, s)) {}
    ~WsScope() { ws_release(D, st); }
};

// Test hook: the next n compute calls fail with NNH_ERR_HIP before they
// enqueue anything (nnh_debug_fail_next).
static std::atomic<int> g_fail_next{0};
int injected_failure() {
    int n = g_fail_next.load();
    while (n > 0)
        if (g_fail_next.compare_exchange_weak(n, n - 1)) return fail(NNH_ERR_HIP, "injected failure (nnh_debug_fail_next)");
    return NNH_OK;
}

static int device_init(int dev) {
    std::lock_guard<std::mutex> g(g_init_mu);
    if (g_inited && g_dev.id == dev) return NNH_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= dev || dev < 0)
        return fail(NNH_ERR_NODEV, "no HIP device " + std::to_string(dev));
    HIPCHK(hipSetDevice(dev));
    if (!g_dev.stream) HIPCHK(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
    if (!g_dev.stream2) HIPCHK(hipStreamCreateWithFlags(&g_dev.stream2, hipStreamNonBlocking));
    if (!g_dev.ev_fork) HIPCHK(hipEventCreateWithFlags(&g_dev.ev_fork, hipEventDisableTiming));
    if (!g_dev.ev_join) HIPCHK(hipEventCreateWithFlags(&g_dev.ev_join, hipEventDisableTiming));
    if (!g_dev.ev_ws) HIPCHK(hipEventCreateWithFlags(&g_dev.ev_ws, hipEventDisableTiming));
    {
        const char* e = std::getenv("NNH_BATCH_STREAMS");
        if (e) g_dev.nwork = std::max(0, std::min(Device::NWORK, std::atoi(e)));
    }
    for (int w = 0; w < g_dev.nwork; w++) {
        if (!g_dev.work[w]) {
            if (w == 0) g_dev.work[w] = g_dev.stream2;
            else HIPCHK(hipStreamCreateWithFlags(&g_dev.work[w], hipStreamNonBlocking));
        }
        if (!g_dev.ev_bjoin[w]) HIPCHK(hipEventCreateWithFlags(&g_dev.ev_bjoin[w], hipEventDisableTiming));
    }
    if (!g_dev.ev_bfork) HIPCHK(hipEventCreateWithFlags(&g_dev.ev_bfork, hipEventDisableTiming));
    if (!g_dev.ev_tab) HIPCHK(hipEventCreateWithFlags(&g_dev.ev_tab, hipEventDisableTiming));
    {
        const char* e = std::getenv("NNH_ONE_STREAM");
        g_dev.two_streams = !(e && e[0] == '1');
    }
    g_dev.id = dev;
    build_constants(g_dev);
    if (dev < NA_MAX_DEVICES && !g_dev.na_ptr[dev]) {
        // the allocation-free path's twiddle tables, built once per device
        u64* p = nullptr;
        HIPCHK(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
        for (int i = 0; i < 3; i++) {
            launch_twiddles(p + (2 * i) * NA_TW, NA_TW, g_dev.rb, i, 0, g_dev.stream);
            launch_twiddles(p + (2 * i + 1) * NA_TW, NA_TW, g_dev.rb, i, 1, g_dev.stream);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(g_dev.stream));
        g_dev.na_ptr[dev] = p;
    }
    g_inited = true;
    return NNH_OK;
}

// For the duration of one allocation-free call: the workspace buffers and
// twiddle tables the product functions read become fixed views of the
// device's scratch array (ensure() and ensure_twiddles() then check sizes
// instead of allocating); the previous ones are restored afterwards.
struct ScratchScope {
    Device& D;
    Buf ws, tmp, tiles;
    const u64* twf[3];
    const u64* twi[3];
    int tw_scale;
    bool prof;
    ScratchScope(Device& d, u64* base)
        : D(d), ws(d.ws), tmp(d.tmp), tiles(d.tiles), tw_scale(d.tw_scale), prof(g_prof) {
        g_prof = false;  // per-kernel timing creates events
        for (int i = 0; i < 3; i++) {
            twf[i] = D.twf[i];
            twi[i] = D.twi[i];
            D.twf[i] = base + (2 * i) * NA_TW;
            D.twi[i] = base + (2 * i + 1) * NA_TW;
        }
        D.tw_scale = NOALLOC_SCALE;
        D.tw_fixed = true;
        u64* w = base + 6 * NA_TW;
        D.ws = Buf{w, NA_WS, true};
        D.tmp = Buf{w + NA_WS, NA_TMP, true};
        D.tiles = Buf{w + NA_WS + NA_TMP, NA_TILES, true};
    }
    ~ScratchScope() {
        D.ws = ws;
        D.tmp = tmp;
        D.tiles = tiles;
        for (int i = 0; i < 3; i++) {
            D.twf[i] = twf[i];
            D.twi[i] = twi[i];
        }
        D.tw_scale = tw_scale;
        D.tw_fixed = false;
        g_prof = prof;
    }
};

// ---------------------------------------------------------------------------
// Transform building blocks (device pointers, queued on st)
// ---------------------------------------------------------------------------
struct Plan {
    int s;
    u64 L;
    int cb;          // fused chunk bits
    int ntop;        // top levels (s - cb)
    LevelPlan lp;    // the split (chunk bits and top passes)
    u64 alpha_m[3];
};

// The level split of a transform of 2^s words: the tuned one for single
// products or batch buckets when the autotuner installed one, else the
// formula (formula_plan: the chunk kernel is VALU-bound and the top passes
// HBM-bound with VALU to spare, so above 2^20 words the top passes take as
// many levels as the fewest passes allow; 3.5 -> 3.4 ms for the 2^30-bit
// product).
static Plan make_plan(const Device& D, int s, bool batched = false) {
    Plan pl;
    pl.s = s;
    pl.L = 1ull << s;
    const LevelPlan& t = D.tuned[batched ? 1 : 0][s];
    pl.lp = t.cb ? t : formula_plan(s);
    pl.cb = pl.lp.cb;
    pl.ntop = s - pl.cb;
    for (int i = 0; i < 3; i++) pl.alpha_m[i] = to_mont(alpha_plain(D, i, s), P[i]);
    return pl;
}

// Top passes as (lowest global level, levels), in forward order.
static std::vector<std::pair<int, int>> top_passes(const Plan& pl) {
    std::vector<std::pair<int, int>> passes;
    int g_hi = pl.s - 1;
    for (int k = 0; k < pl.lp.np; k++) {
        const int nl = pl.lp.lev[k];
        passes.push_back({g_hi - nl + 1, nl});
        g_hi -= nl;
    }
    return passes;
}

static void set3(u64* dst[3], u64* base, u64 L) {
    for (int i = 0; i < 3; i++) dst[i] = base + (u64)i * L;
}

// Stream for the p2 stages: stream2 after it has waited for everything
// queued on st so far, or st itself (one stream while per-kernel timing is
// on, so that the events bracket single kernels).
static int fork_p2(Device& D, hipStream_t st, hipStream_t* st2) {
    *st2 = st;
    if (!D.two_streams || g_prof || st == D.stream2) return NNH_OK;
    HIPCHK(hipEventRecord(D.ev_fork, st));
    HIPCHK(hipStreamWaitEvent(D.stream2, D.ev_fork, 0));
    *st2 = D.stream2;
    return NNH_OK;
}
// st waits for the p2 stages queued on st2.
static int join_p2(Device& D, hipStream_t st, hipStream_t st2) {
    if (st2 == st) return NNH_OK;
    HIPCHK(hipEventRecord(D.ev_join, st2));
    HIPCHK(hipStreamWaitEvent(st, D.ev_join, 0));
    return NNH_OK;
}

// Forward top levels (global bits s-1 .. 12) of operand x into f, and of
// operand y into g when y is set (each pass one launch for both).
static void fwd_top(const Device& D, const Plan& pl, const u64* x, u64 xn, u64* f, hipStream_t st,
                    hipStream_t st2, const u64* y = nullptr, u64 yn = 0, u64* g = nullptr) {
    bool first = true;
    for (auto [g_lo, nl] : top_passes(pl)) {
        TopPassDesc td{};
        set3(td.data, f, pl.L);
        for (int i = 0; i < 3; i++) {
            td.tw[i] = D.twf[i];
            td.alpha_m[i] = pl.alpha_m[i];
        }
        td.limbs = first ? x : nullptr;
        td.an = xn;
        td.L = pl.L;
        td.g_lo = (u32)g_lo;
        td.nlev = (u32)nl;
        td.inverse = false;
        td.two = y != nullptr;
        if (td.two) {
            set3(td.data2, g, pl.L);
            td.limbs2 = first ? y : nullptr;
            td.an2 = yn;
        }
        {
            // algorithmic bytes: limbs read once (first pass) or residues read, residues written
            u64 rd = first ? 8 * std::min<u64>(xn, pl.L) : 24 * pl.L;
            if (td.two) rd += first ? 8 * std::min<u64>(yn, pl.L) + 24 * pl.L : 48 * pl.L;
            ProfScope ps(first ? K_TOP_LOAD : K_TOP_FWD, rd + 24 * pl.L, st);
            launch_top_pass(td, st, st2);
        }
        first = false;
    }
}

// Inverse top levels (global bits 12 .. s-1), mirrored passes.
static void inv_top(const Device& D, const Plan& pl, u64* f, hipStream_t st, hipStream_t st2) {
    std::vector<std::pair<int, int>> passes = top_passes(pl);
    for (auto it = passes.rbegin(); it != passes.rend(); ++it) {
        TopPassDesc td{};
        set3(td.data, f, pl.L);
        for (int i = 0; i < 3; i++) {
            td.tw[i] = D.twi[i];
            td.alpha_m[i] = pl.alpha_m[i];
        }
        td.limbs = nullptr;
        td.L = pl.L;
        td.g_lo = (u32)it->first;
        td.nlev = (u32)it->second;
        td.inverse = true;
        {
            ProfScope ps(K_TOP_INV, 48 * pl.L, st);
            launch_top_pass(td, st, st2);
        }
    }
}

// Fused chunk kernel.  a/b are limbs when the transform fits one chunk and
// the operand is not yet transformed.
static void mid(const Device& D, const Plan& pl, u64* fa, const u64* fb, const u64* la, u64 an,
                const u64* lb, u64 bn, bool sqr, bool fwd_only, bool a_done, bool b_done, hipStream_t st,
                hipStream_t st2) {
    MidDesc md{};
    set3(md.fa, fa, pl.L);
    for (int i = 0; i < 3; i++) {
        md.fb[i] = fb + (u64)i * pl.L;
        md.twf[i] = D.twf[i];
        md.twi[i] = D.twi[i];
        md.alpha_m[i] = pl.alpha_m[i];
    }
    md.la = la;
    md.lb = lb;
    md.an = an;
    md.bn = bn;
    md.L = pl.L;
    md.cb = pl.cb;
    md.sqr = sqr;
    md.fwd_only = fwd_only;
    md.a_done = a_done;
    md.b_done = b_done;
    // bytes: operands read (limbs or residues), result written
    u64 ra = la ? 8 * std::min<u64>(an, pl.L) : 24 * pl.L;
    u64 rb = fwd_only || sqr ? 0 : (lb ? 8 * std::min<u64>(bn, pl.L) : 24 * pl.L);
    ProfScope ps(K_MID, ra + rb + 24 * pl.L, st);
    launch_mid(md, st, st2);
}

static int crt(Device& D, const Plan& pl, const u64* f, u64 n_res, u64* out, u64 n_out, const u64* addend,
               u64 n_add, hipStream_t st) {
    u64 tiles = crt_tiles_single(n_out);
    int rc = ensure(D.tiles, (2 * tiles + 7) / 8 + 16);
    if (rc) return rc;
    CrtDesc cd{};
    for (int i = 0; i < 3; i++) cd.r[i] = f + (u64)i * pl.L;
    cd.columns = 0;
    cd.n_res = n_res;
    cd.out = out;
    cd.n_out = n_out;
    cd.addend = addend;
    cd.n_add = n_add;
    cd.tile_fn = (u8*)D.tiles.p;
    cd.tile_cin = (u8*)D.tiles.p + tiles;
    cd.C = D.crt;
    {
        ProfScope ps(K_CRT, 24 * n_res + 8 * n_out + 8 * n_add, st);
        launch_crt(cd, st);
    }
    return NNH_OK;
}

// Exact product of transforms already prepared: fa holds A (limbs loaded by
// the caller's top passes, or a finished transform if a_done), similarly fb.
// Writes out[0 .. n_out) (+ addend).
static int product_core(Device& D, const Plan& pl, const u64* a, u64 an, const u64* b, u64 bn, bool sqr,
                        const u64* fb_done, u64* out, u64 n_out, const u64* addend, u64 n_add, hipStream_t st) {
    u64* fa = D.ws.p;
    u64* fb = D.ws.p + 3 * pl.L;
    const u64* fbp = fb_done ? fb_done : fb;
    bool small = pl.ntop == 0;
    hipStream_t st2;
    int rc = fork_p2(D, st, &st2);
    if (rc) return rc;
    if (!small) {
        if (!sqr && !fb_done) fwd_top(D, pl, a, an, fa, st, st2, b, bn, fb);
        else fwd_top(D, pl, a, an, fa, st, st2);
    }
    mid(D, pl, fa, fbp, small ? a : nullptr, an, (small && !sqr && !fb_done) ? b : nullptr, bn, sqr, false,
        false, fb_done != nullptr, st, st2);
    if (!small) inv_top(D, pl, fa, st, st2);
    if ((rc = join_p2(D, st, st2))) return rc;
    u64 n_res = std::min<u64>(an + (sqr ? an : bn) - 1, pl.L);
    return crt(D, pl, fa, n_res, out, n_out, addend, n_add, st);
}

// Product of operands whose transforms are complete: A's (3 L words at
// fa_src) is copied into the workspace, which the product overwrites; B's
// is read where it is (fb_src), or A's is squared.
static int product_transformed(Device& D, const Plan& pl, const u64* fa_src, u64 an, const u64* fb_src, u64 bn,
                               bool sqr, u64* out, u64 n_out, hipStream_t st) {
    u64* fa = D.ws.p;
    HIPCHK(hipMemcpyAsync(fa, fa_src, 3 * pl.L * sizeof(u64), hipMemcpyDeviceToDevice, st));
    hipStream_t st2;
    int rc = fork_p2(D, st, &st2);
    if (rc) return rc;
    mid(D, pl, fa, sqr ? fa : fb_src, nullptr, an, nullptr, sqr ? an : bn, sqr, false, true, !sqr, st, st2);
    if (pl.ntop) inv_top(D, pl, fa, st, st2);
    if ((rc = join_p2(D, st, st2))) return rc;
    const u64 n_res = std::min<u64>(an + (sqr ? an : bn) - 1, pl.L);
    return crt(D, pl, fa, n_res, out, n_out, nullptr, 0, st);
}

// Complete forward transform of x into dst (3 L words).
static int transform_full(Device& D, const Plan& pl, const u64* x, u64 xn, u64* dst, hipStream_t st) {
    bool small = pl.ntop == 0;
    hipStream_t st2;
    int rc = fork_p2(D, st, &st2);
    if (rc) return rc;
    if (!small) fwd_top(D, pl, x, xn, dst, st, st2);
    mid(D, pl, dst, dst, small ? x : nullptr, xn, nullptr, 0, false, true, false, false, st, st2);
    return join_p2(D, st, st2);
}

static int school(Device& D, const u64* a, u64 an, const u64* b, u64 bn, u64* out, hipStream_t st) {
    u64 n = an + bn - 1;
    int rc = ensure(D.tmp, 3 * n + 64);
    if (rc) return rc;
    u64* c0 = D.tmp.p;
    u64* c1 = c0 + n;
    u64* c2 = c1 + n;
    {
        ProfScope ps(K_SCHOOL, 8 * (an + bn) + 24 * n, st);
        launch_school_cols(a, an, b, bn, c0, c1, c2, st);
    }
    u64 n_out = an + bn;
    u64 tiles = crt_tiles_single(n_out);
    rc = ensure(D.tiles, (2 * tiles + 7) / 8 + 16);
    if (rc) return rc;
    CrtDesc cd{};
    cd.r[0] = c0; cd.r[1] = c1; cd.r[2] = c2;
    cd.columns = 1;
    cd.n_res = n;
    cd.out = out;
    cd.n_out = n_out;
    cd.tile_fn = (u8*)D.tiles.p;
    cd.tile_cin = (u8*)D.tiles.p + tiles;
    cd.C = D.crt;
    {
        ProfScope ps(K_CRT, 24 * n + 8 * n_out, st);
        launch_crt(cd, st);
    }
    return NNH_OK;
}

// out[0 .. an+bn) = a * b (device pointers), reference natural::_mul:
// small operands by schoolbook columns, otherwise the transform with the
// reference's one-level split of an unbalanced larger operand
// (natural.h:391-431): B is transformed once and reused for both halves.
static int mul_device(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, hipStream_t st) {
    if (an < bn) {
        std::swap(a, b);
        std::swap(an, bn);
    }
    if (bn < NTT_MUL_THRESHOLD) return school(D, a, an, b, bn, out, st);
    u64 a0n = (an + 1) >> 1;
    int s = ntt_scale(an, bn), s_opt = ntt_scale(a0n, bn);
    if (s > MAX_SCALE) return fail(NNH_ERR_ARG, "operands too large");
    int rc;
    if (s == s_opt) {
        Plan pl = make_plan(D, s);
        if ((rc = ensure_twiddles(D, s)) || (rc = ensure(D.ws, 6 * pl.L))) return rc;
        return product_core(D, pl, a, an, b, bn, false, nullptr, out, an + bn, nullptr, 0, st);
    }
    Plan pl = make_plan(D, s_opt);
    if ((rc = ensure_twiddles(D, s_opt)) || (rc = ensure(D.ws, 6 * pl.L)) || (rc = ensure(D.tmp, bn + 64)))
        return rc;
    u64* fb = D.ws.p + 3 * pl.L;
    if ((rc = transform_full(D, pl, b, bn, fb, st))) return rc;
    // low half: a0 * b -> out[0 .. a0n + bn)
    if ((rc = product_core(D, pl, a, a0n, b, bn, false, fb, out, a0n + bn, nullptr, 0, st))) return rc;
    // high half: a1 * b + (overlap) -> out[a0n .. an + bn)
    HIPCHK(hipMemcpyAsync(D.tmp.p, out + a0n, bn * sizeof(u64), hipMemcpyDeviceToDevice, st));
    return product_core(D, pl, a + a0n, an - a0n, b, bn, false, fb, out + a0n, an - a0n + bn, D.tmp.p, bn, st);
}

// ---------------------------------------------------------------------------
// Explicit algorithm arms (every arm is exact; the choice only costs time).
// ---------------------------------------------------------------------------
static int kara_device(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, hipStream_t st) {
    if (an < bn) {
        std::swap(a, b);
        std::swap(an, bn);
    }
    if (bn < 2) return school(D, a, an, b, bn, out, st);
    u64 nc = an + bn - 1;
    int rc = ensure(D.tmp, kara_ws_words(an, bn) + 3 * nc + 64);
    if (rc) return rc;
    u64* cols = D.tmp.p + kara_ws_words(an, bn);
    {
        ProfScope ps(K_SCHOOL, 8 * (an + bn) + 24 * nc, st);
        launch_kara(a, an, b, bn, D.tmp.p, cols, st);
    }
    u64 n_out = an + bn;
    u64 tiles = crt_tiles_single(n_out);
    if ((rc = ensure(D.tiles, (2 * tiles + 7) / 8 + 16))) return rc;
    CrtDesc cd{};
    cd.r[0] = cols; cd.r[1] = cols + nc; cd.r[2] = cols + 2 * nc;
    cd.columns = 1;
    cd.n_res = nc;
    cd.out = out;
    cd.n_out = n_out;
    cd.tile_fn = (u8*)D.tiles.p;
    cd.tile_cin = (u8*)D.tiles.p + tiles;
    cd.C = D.crt;
    {
        ProfScope ps(K_CRT, 24 * nc + 8 * n_out, st);
        launch_crt(cd, st);
    }
    return NNH_OK;
}

// One transform of scale max(get_NTT_scale, 8) for the whole product.
static int ntt_direct(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, hipStream_t st) {
    int s = std::max(ntt_scale(an, bn), MIN_MID_BITS);
    if (s > MAX_SCALE) return fail(NNH_ERR_ARG, "operands too large");
    Plan pl = make_plan(D, s);
    int rc;
    if ((rc = ensure_twiddles(D, s)) || (rc = ensure(D.ws, 6 * pl.L))) return rc;
    return product_core(D, pl, a, an, b, bn, false, nullptr, out, an + bn, nullptr, 0, st);
}

// The larger operand cut into pieces of P = 2^sp - bn + 1 words, each
// multiplied by b through a transform of scale sp; piece i's product
// (P + bn words) starts at word i P.
//
// Sequential form: one product per piece reusing b's transform, each piece
// added through the CRT addend (used when the pieces are so short that the
// grouped form below would need more than 4 buffers).
static int ntt_chunked_seq(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, int sp,
                           hipStream_t st) {
    Plan pl = make_plan(D, sp);
    u64 P = pl.L - bn + 1;
    int rc;
    if ((rc = ensure_twiddles(D, sp)) || (rc = ensure(D.ws, 6 * pl.L)) || (rc = ensure(D.tmp, bn + 64))) return rc;
    u64* fb = D.ws.p + 3 * pl.L;
    if ((rc = transform_full(D, pl, b, bn, fb, st))) return rc;
    for (u64 off = 0; off < an; off += P) {
        u64 pn = std::min(P, an - off);
        if (off == 0) {
            if ((rc = product_core(D, pl, a, pn, b, bn, false, fb, out, pn + bn, nullptr, 0, st))) return rc;
        } else {
            HIPCHK(hipMemcpyAsync(D.tmp.p, out + off, bn * sizeof(u64), hipMemcpyDeviceToDevice, st));
            if ((rc = product_core(D, pl, a + off, pn, b, bn, false, fb, out + off, pn + bn, D.tmp.p, bn, st)))
                return rc;
        }
    }
    return NNH_OK;
}

// Grouped form: all pieces run as one batch whose jobs share b's transform.
// Pieces G = ceil((P + bn) / P) apart do not overlap, so piece i lands in
// buffer i mod G (buffer 0 is `out`; all buffers zero outside their
// pieces) and G - 1 carry-lookahead additions sum the buffers.  The batch
// synchronises the stream.
static int ntt_chunked(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, int sp, hipStream_t st) {
    if (an < bn) {
        std::swap(a, b);
        std::swap(an, bn);
    }
    if (sp <= 0) sp = ntt_scale(bn, bn);
    sp = std::max(sp, MIN_MID_BITS);
    if (sp > MAX_SCALE || (1ull << sp) < bn) return fail(NNH_ERR_ARG, "bad piece scale");
    const u64 L = 1ull << sp, P = L - bn + 1, n = an + bn;
    int rc;
    if (an <= P) {  // one piece: a single product at scale sp
        Plan pl = make_plan(D, sp);
        if ((rc = ensure_twiddles(D, sp)) || (rc = ensure(D.ws, 6 * pl.L))) return rc;
        return product_core(D, pl, a, an, b, bn, false, nullptr, out, n, nullptr, 0, st);
    }
    const u64 G = (P + bn + P - 1) / P;
    if (G > 4) return ntt_chunked_seq(D, out, a, an, b, bn, sp, st);
    // the pieces run as an internal batch, which a graph cannot hold (refused
    // here, before anything is enqueued)
    if (capturing(st)) return fail(NNH_ERR_ARG, "an internal batch (chunked arm) cannot be captured into a graph");
    if ((rc = ensure(D.tmp, (G - 1) * n + 64)) || (rc = ensure(D.tiles, 2 * ((n + LIMB_TILE - 1) / LIMB_TILE) + 16)))
        return rc;
    auto buf = [&](u64 g) { return g == 0 ? out : D.tmp.p + (g - 1) * n; };
    for (u64 g = 0; g < G; g++) launch_fill(buf(g), n, 0, st);
    std::vector<IReq> rq;
    for (u64 i = 0, off = 0; off < an; i++, off += P) {
        const u64 pn = std::min(P, an - off);
        rq.push_back(IReq{buf(i % G) + off, a + off, pn, b, bn, pn + bn, nullptr, 0});
    }
    if ((rc = dev_batch(rq.data(), rq.size(), st))) return rc;
    for (u64 g = 1; g < G; g++)
        launch_add_one(AddJob{out, out, buf(g), n, n, n, 0, nullptr, nullptr, 0, 0}, (u8*)D.tiles.p, st);
    HIPCHK(hipGetLastError());
    return NNH_OK;
}

// The reference's dispatch without its schoolbook threshold: transform,
// split once when halving the larger operand saves a power of two.
static int ntt_refsplit(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, hipStream_t st) {
    if (an < bn) {
        std::swap(a, b);
        std::swap(an, bn);
    }
    u64 a0n = (an + 1) >> 1;
    int s = ntt_scale(an, bn), s_opt = ntt_scale(a0n, bn);
    if (s == s_opt || s_opt < MIN_MID_BITS || a0n < 1 || an - a0n < 1) return ntt_direct(D, out, a, an, b, bn, st);
    Plan pl = make_plan(D, s_opt);
    int rc;
    if ((rc = ensure_twiddles(D, s_opt)) || (rc = ensure(D.ws, 6 * pl.L)) || (rc = ensure(D.tmp, bn + 64))) return rc;
    u64* fb = D.ws.p + 3 * pl.L;
    if ((rc = transform_full(D, pl, b, bn, fb, st))) return rc;
    if ((rc = product_core(D, pl, a, a0n, b, bn, false, fb, out, a0n + bn, nullptr, 0, st))) return rc;
    HIPCHK(hipMemcpyAsync(D.tmp.p, out + a0n, bn * sizeof(u64), hipMemcpyDeviceToDevice, st));
    return product_core(D, pl, a + a0n, an - a0n, b, bn, false, fb, out + a0n, an - a0n + bn, D.tmp.p, bn, st);
}

// The installed cost-model policy (nnh_policy_set); read under the device lock.
static PolicyTable g_policy;

static int mul_algo_device(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, int algo, int param,
                           hipStream_t st);

// The library's dispatch for one product among `count` back-to-back ones:
// the arm the installed policy picks, else the reference's thresholds.
static int mul_dispatch(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, u64 count,
                        hipStream_t st) {
    const int arm = g_policy.choose(an, bn, count);
    if (arm < 0) return mul_device(D, out, a, an, b, bn, st);
    int algo, param;
    arm_call(arm, an, bn, &algo, &param);
    return mul_algo_device(D, out, a, an, b, bn, algo, param, st);
}

static int mul_algo_device(Device& D, u64* out, const u64* a, u64 an, const u64* b, u64 bn, int algo, int param,
                           hipStream_t st) {
    switch (algo) {
        case ALGO_AUTO: return mul_dispatch(D, out, a, an, b, bn, 1, st);
        case ALGO_REFERENCE: return mul_device(D, out, a, an, b, bn, st);
        case ALGO_MODEL:
            if (!g_policy.on) return fail(NNH_ERR_ARG, "no cost-model policy installed");
            return mul_dispatch(D, out, a, an, b, bn, std::max(param, 1), st);
        case ALGO_SCHOOL: return an >= bn ? school(D, a, an, b, bn, out, st) : school(D, b, bn, a, an, out, st);
        case ALGO_KARATSUBA: return kara_device(D, out, a, an, b, bn, st);
        case ALGO_NTT: return ntt_direct(D, out, a, an, b, bn, st);
        case ALGO_NTT_SPLIT: return ntt_refsplit(D, out, a, an, b, bn, st);
        case ALGO_NTT_CHUNK: return ntt_chunked(D, out, a, an, b, bn, param, st);
        default: return fail(NNH_ERR_ARG, "unknown algorithm");
    }
}

static int sqr_device(Device& D, u64* out, const u64* a, u64 an, hipStream_t st) {
    if (an < NTT_MUL_THRESHOLD) return school(D, a, an, a, an, out, st);
    int s = ntt_scale(an, an);
    if (s > MAX_SCALE) return fail(NNH_ERR_ARG, "operand too large");
    Plan pl = make_plan(D, s);
    int rc;
    if ((rc = ensure_twiddles(D, s)) || (rc = ensure(D.ws, 6 * pl.L))) return rc;
    return product_core(D, pl, a, an, a, an, true, nullptr, out, 2 * an, nullptr, 0, st);
}

static int stage_host(Device& D, u64 words) { return ensure(D.stage, words); }

// ---------------------------------------------------------------------------
// Batched products: jobs bucketed by transform scale, one device job table,
// one workspace; enqueued directly or captured once into a hipGraph.
// ---------------------------------------------------------------------------
}  // namespace nnh

// A forward transform kept in device memory (the counterpart of an NTT_info
// kept by mul_save_NTT_info / sqr_save_NTT_info, natural.h:452-501).
struct nnh_xf {
    int scale = 0;
    nnh::u64 held = 0;       // words of the transformed operand (0: none yet)
    nnh::u64* f = nullptr;   // 3 * 2^scale words, prime-major
};

struct nnh_batch {
    nnh::BatchPlan plan;
    nnh::BJob* d_jobs = nullptr;
    nnh::u32* d_sj = nullptr;
    nnh::u8* d_tiles = nullptr;
    nnh::AddJob* d_adds = nullptr;  // chunked requests: odd pieces added into dst
    nnh::u8* d_add_tiles = nullptr;
    nnh::u64* ws = nullptr;
    // twiddle tables read by the batch's launches (and baked into its graph),
    // kept alive by tw for as long as the batch exists
    std::shared_ptr<nnh::Twiddles> tw;
    const nnh::u64* twf[3]{};
    const nnh::u64* twi[3]{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    bool borrowed = false;  // buffers belong to the device's internal batch workspace
};

namespace nnh {

static void batch_free(nnh_batch* B) {
    if (!B) return;
    if (!B->borrowed || B->exec) (void)hipDeviceSynchronize();
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->cap_stream) (void)hipStreamDestroy(B->cap_stream);
    if (!B->borrowed) {
        for (void* p : {(void*)B->d_jobs, (void*)B->d_sj, (void*)B->d_tiles, (void*)B->d_adds, (void*)B->d_add_tiles,
                        (void*)B->ws})
            if (p) (void)hipFree(p);
    }
    delete B;
}

// Enqueue every launch of the batch on st.
static int batch_enqueue(Device& D, nnh_batch* B, hipStream_t st) {
    const BatchPlan& P = B->plan;
    const bool fan = D.two_streams && D.nwork > 0 && P.buckets.size() > 0;
    // the buckets' two chains take lanes st, work[0], work[1], ... in turn:
    // only the workers they reach are forked and joined
    const int used = fan ? (int)std::min<u64>((u64)D.nwork, 2 * P.buckets.size() - 1) : 0;
    if (fan) {
        HIPCHK(hipEventRecord(D.ev_bfork, st));
        for (int w = 0; w < used; w++) HIPCHK(hipStreamWaitEvent(D.work[w], D.ev_bfork, 0));
    }
    // lanes: st and the workers
    int lane = 0;
    auto next_stream = [&]() {
        const int l = lane++ % (D.nwork + 1);
        return (fan && l) ? D.work[l - 1] : st;
    };
    if (P.n_school) launch_school_batch(B->d_jobs, B->d_sj, P.n_school, P.ncols, st);
    for (const Bucket& bk : P.buckets) {
        const hipStream_t sa = next_stream(), sb = next_stream();
        Plan pl = make_plan(D, bk.s, true);
        const BJob* jobs = B->d_jobs + bk.first;
        TopBatchParams tp{};
        for (int i = 0; i < 3; i++) tp.alpha_m[i] = pl.alpha_m[i];
        tp.L = pl.L;
        // forward top passes, both operands
        std::vector<std::pair<int, int>> passes = top_passes(pl);
        for (int op = 0; op < 2; op++)
            for (size_t k = 0; k < passes.size(); k++) {
                for (int i = 0; i < 3; i++) tp.tw[i] = B->twf[i];
                tp.g_lo = (u32)passes[k].first;
                tp.nlev = (u32)passes[k].second;
                launch_top_pass_batch(jobs, bk.count, op, tp, false, k == 0, sa, sb);
            }
        MidBatchParams mp{};
        for (int i = 0; i < 3; i++) {
            mp.twf[i] = B->twf[i];
            mp.twi[i] = B->twi[i];
            mp.alpha_m[i] = pl.alpha_m[i];
        }
        mp.L = pl.L;
        if (bk.own_count)
            launch_mid_batch(B->d_jobs + bk.own_first, bk.own_count, mp, pl.cb, pl.ntop == 0, true, sa, sb);
        launch_mid_batch(jobs, bk.count, mp, pl.cb, pl.ntop == 0, false, sa, sb);
        for (auto it = passes.rbegin(); it != passes.rend(); ++it) {
            for (int i = 0; i < 3; i++) tp.tw[i] = B->twi[i];
            tp.g_lo = (u32)it->first;
            tp.nlev = (u32)it->second;
            launch_top_pass_batch(jobs, bk.count, 0, tp, true, false, sa, sb);
        }
    }
    if (fan)
        for (int w = 0; w < used; w++) {
            HIPCHK(hipEventRecord(D.ev_bjoin[w], D.work[w]));
            HIPCHK(hipStreamWaitEvent(st, D.ev_bjoin[w], 0));
        }
    launch_crt_batch(B->d_jobs, P.n_req_jobs, P.ntiles, B->d_tiles, B->d_tiles + P.ntiles, D.crt, st);
    if (!P.adds.empty()) launch_add_batch(B->d_adds, (u32)P.adds.size(), P.add_tiles, B->d_add_tiles, st);
    return NNH_OK;
}

// flags bit 0: capture into a hipGraph; bit 1: internal batch (buffers from
// the device's reusable batch workspace; no policy); bit 2: plan each
// request's arm with the installed policy; bit 4: leave the job tables to
// the caller (upload_tables, on its stream).  forced_arm >= 0: every request
// runs that arm (timing).
static int batch_create(Device& D, const IReq* reqs, u64 n, int flags, nnh_batch** out, int forced_arm = -1) {
    nnh_batch* B = new nnh_batch();
    BatchPlan& P = B->plan;
    std::string err;
    const bool use_policy = (flags & 4) && !(flags & 2) && g_policy.on;
    if (plan_batch(reqs, n, use_policy ? &g_policy : nullptr, std::min<u64>(n, 64), P, err, forced_arm)) {
        delete B;
        return fail(NNH_ERR_ARG, err);
    }
    int rc;
    if (P.max_s && (rc = ensure_twiddles(D, P.max_s))) {
        batch_free(B);
        return rc;
    }
    B->tw = D.tw;
    for (int i = 0; i < 3; i++) {
        B->twf[i] = D.twf[i];
        B->twi[i] = D.twi[i];
    }
    const u64 ws = P.ws_words, tiles = P.ntiles;
    if (flags & 2) {
        // internal batch: carve the buffers out of the device's reusable
        // workspace (job table, then the additions; CRT tile flags, then the
        // schoolbook list, then the additions' tile flags)
        B->borrowed = true;
        const u64 jw = (P.jobs.size() * sizeof(BJob) + 7) / 8, aw = (P.adds.size() * sizeof(AddJob) + 7) / 8;
        const u64 tw = (2 * tiles + 7) / 8 + 1, sw = (P.n_school * 4 + 4 + 7) / 8 + 1, atw = (2 * P.add_tiles + 7) / 8;
        if ((rc = ensure(D.bws, ws + 1)) || (rc = ensure(D.bjobs, jw + aw + 1)) || (rc = ensure(D.btiles, tw + sw + atw + 1))) {
            batch_free(B);
            return rc;
        }
        B->ws = D.bws.p;
        B->d_jobs = (BJob*)D.bjobs.p;
        B->d_adds = (AddJob*)(D.bjobs.p + jw);
        B->d_tiles = (u8*)D.btiles.p;
        B->d_sj = (u32*)(D.btiles.p + tw);
        B->d_add_tiles = (u8*)(D.btiles.p + tw + sw);
    } else {
        auto alloc = [&](void** p, u64 bytes) -> int {
            hipError_t e = hipMalloc(p, std::max<u64>(bytes, 64));
            if (e != hipSuccess) return fail(NNH_ERR_NOMEM, std::string("batch allocation: ") + hipGetErrorString(e));
            return NNH_OK;
        };
        if ((rc = alloc((void**)&B->ws, ws * 8)) || (rc = alloc((void**)&B->d_jobs, P.jobs.size() * sizeof(BJob))) ||
            (rc = alloc((void**)&B->d_tiles, 2 * tiles)) || (rc = alloc((void**)&B->d_sj, P.n_school * 4 + 4)) ||
            (!P.adds.empty() && ((rc = alloc((void**)&B->d_adds, P.adds.size() * sizeof(AddJob))) ||
                                 (rc = alloc((void**)&B->d_add_tiles, 2 * P.add_tiles))))) {
            batch_free(B);
            return rc;
        }
    }
    rebase_plan(P, B->ws);
    std::vector<u32> sj(P.n_school);
    for (u32 i = 0; i < P.n_school; i++) sj[i] = i;  // school jobs sort first
    if (!(flags & 16)) {
        HIPCHK(hipMemcpy(B->d_jobs, P.jobs.data(), P.jobs.size() * sizeof(BJob), hipMemcpyHostToDevice));
        if (P.n_school) HIPCHK(hipMemcpy(B->d_sj, sj.data(), sj.size() * 4, hipMemcpyHostToDevice));
        if (!P.adds.empty())
            HIPCHK(hipMemcpy(B->d_adds, P.adds.data(), P.adds.size() * sizeof(AddJob), hipMemcpyHostToDevice));
    }
    if (flags & 1) {
        HIPCHK(hipStreamCreateWithFlags(&B->cap_stream, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(B->cap_stream, hipStreamCaptureModeThreadLocal));
        const int erc = batch_enqueue(D, B, B->cap_stream);
        const hipError_t ce = hipStreamEndCapture(B->cap_stream, &B->graph);
        const hipError_t ie =
            (erc || ce != hipSuccess) ? hipSuccess : hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0);
        if (erc || ce != hipSuccess || ie != hipSuccess) {
            batch_free(B);
            if (erc) return erc;
            return fail(NNH_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(ce != hipSuccess ? ce : ie));
        }
    }
    *out = B;
    return NNH_OK;
}

// ---------------------------------------------------------------------------
// The cost-model policy: the network runs on the device once over every
// point of the policy table (nnh_mlp_forward), and the predictions replace
// the table the dispatch reads.
// ---------------------------------------------------------------------------
static int install_policy(nnh_mlp* m) {
    const int want = PolicyTable::N_IN * 64 + 64 + 64 * 64 + 64 + N_ARMS * 64 + N_ARMS;
    if (nnh_mlp_num_params(m) != want) return fail(NNH_ERR_ARG, "policy network must be [3 -> 64 -> 64 -> 6]");
    std::vector<float> x, y((size_t)PolicyTable::N_POINTS * N_ARMS);
    PolicyTable::features(x);
    int rc = nnh_mlp_forward(m, x.data(), PolicyTable::N_POINTS, y.data());
    if (rc) return fail(rc, "policy network forward pass failed");
    for (float v : y)
        if (!std::isfinite(v)) return fail(NNH_ERR_ARG, "policy network predicts a non-finite time");
    PolicyTable t;
    t.set(y.data(), N_ARMS);
    std::lock_guard<std::mutex> g(g_dev.mu);
    g_policy = std::move(t);
    return NNH_OK;
}

// Policy file: a line "nnh-policy 1 <n_in> <n_out> <n_params>", then the
// network's parameters (nnh_mlp layout), whitespace-separated decimal floats.
static int load_policy_file(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(NNH_ERR_ARG, std::string("cannot open policy file ") + path);
    char magic[32] = {0};
    int ver = 0, ni = 0, no = 0, np = 0;
    int got = std::fscanf(f, "%31s %d %d %d %d", magic, &ver, &ni, &no, &np);
    std::vector<float> p;
    if (got == 5 && std::string(magic) == "nnh-policy" && ver == 1 && ni == PolicyTable::N_IN && no == N_ARMS &&
        np > 0 && np < (1 << 20)) {
        p.resize(np);
        for (int k = 0; k < np; k++)
            if (std::fscanf(f, "%f", &p[k]) != 1) {
                p.clear();
                break;
            }
    }
    std::fclose(f);
    if (p.empty()) return fail(NNH_ERR_ARG, std::string("malformed policy file ") + path);
    nnh_mlp* m = nullptr;
    int rc = nnh_mlp_create(ni, no, 1, &m);
    if (rc) return fail(rc, "policy network allocation failed");
    rc = nnh_mlp_set_params(m, p.data(), p.size());
    if (!rc) rc = install_policy(m);
    else rc = fail(rc, "policy parameters do not fit the network");
    nnh_mlp_destroy(m);
    return rc;
}

// At nnh_init: NNH_POLICY=<file> installs that policy, NNH_POLICY=reference
// keeps the reference's thresholds, and otherwise the policy shipped beside
// the library (policy_<arch>.txt) is installed when it exists.
static void default_policy() {
    const char* e = std::getenv("NNH_POLICY");
    std::string path;
    if (e && *e) {
        if (std::string(e) == "reference") return;
        path = e;
    } else {
        Dl_info info{};
        if (!dladdr((void*)&default_policy, &info) || !info.dli_fname) return;
        std::string lib = info.dli_fname;
        const size_t cut = lib.find_last_of('/');
        path = (cut == std::string::npos ? std::string(".") : lib.substr(0, cut)) + "/policy_gfx950.txt";
        if (FILE* f = std::fopen(path.c_str(), "r")) std::fclose(f);
        else return;
    }
    const std::string keep = g_err;
    if (load_policy_file(path.c_str()) != NNH_OK)
        std::fprintf(stderr, "libnnh: %s; using the reference thresholds\n", g_err.c_str());
    g_err = keep;
}

// ---------------------------------------------------------------------------
// Level-split plans: the table make_plan reads, its file form and the
// autotuner that fills it.
// ---------------------------------------------------------------------------
static std::string device_key() {
    hipDeviceProp_t pr{};
    if (hipGetDeviceProperties(&pr, g_dev.id) != hipSuccess) return "unknown 0";
    std::string arch = pr.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
    return arch + " " + std::to_string(pr.multiProcessorCount);
}

// File: "nnh-plans 1 <arch> <compute units>", then one line per tuned scale,
// "single|batch <s> <cb> <np> <lev0> <lev1> <lev2> <lev3>".
static int load_plans_file(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(NNH_ERR_ARG, std::string("cannot open plan file ") + path);
    char magic[32] = {0}, arch[64] = {0};
    int ver = 0, cus = 0;
    LevelPlan t[2][MAX_SCALE + 1];
    std::string bad;
    if (std::fscanf(f, "%31s %d %63s %d", magic, &ver, arch, &cus) != 4 || std::string(magic) != "nnh-plans" || ver != 1)
        bad = "malformed plan file";
    else if (std::string(arch) + " " + std::to_string(cus) != device_key())
        bad = std::string("plan file made on another device (") + arch + " " + std::to_string(cus) + ")";
    char kind[16];
    int sc;
    LevelPlan lp;
    while (bad.empty() && std::fscanf(f, "%15s %d %d %d %d %d %d %d", kind, &sc, &lp.cb, &lp.np, &lp.lev[0], &lp.lev[1],
                                      &lp.lev[2], &lp.lev[3]) == 8) {
        const std::string k = kind;
        if ((k != "single" && k != "batch") || !plan_valid(sc, lp)) bad = "malformed plan file";
        else t[k == "batch"][sc] = lp;
    }
    std::fclose(f);
    if (!bad.empty()) return fail(NNH_ERR_ARG, bad + " " + path);
    std::lock_guard<std::mutex> g(g_dev.mu);
    std::copy(&t[0][0], &t[0][0] + 2 * (MAX_SCALE + 1), &g_dev.tuned[0][0]);
    return NNH_OK;
}

static std::string beside_library(const char* name) {
    Dl_info info{};
    if (!dladdr((void*)&beside_library, &info) || !info.dli_fname) return "";
    std::string lib = info.dli_fname;
    const size_t cut = lib.find_last_of('/');
    return (cut == std::string::npos ? std::string(".") : lib.substr(0, cut)) + "/" + name;
}

// At nnh_init: NNH_PLANS=<file> loads that table, NNH_PLANS=formula keeps
// the formula, otherwise plans_gfx950.txt beside the library is loaded when
// present and made on this kind of device.
static void default_plans() {
    const char* e = std::getenv("NNH_PLANS");
    std::string path;
    if (e && *e) {
        if (std::string(e) == "formula") return;
        path = e;
    } else {
        path = beside_library("plans_gfx950.txt");
        FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
        if (!f) return;
        std::fclose(f);
    }
    const std::string keep = g_err;
    if (load_plans_file(path.c_str()) != NNH_OK) std::fprintf(stderr, "libnnh: %s; using the formula\n", g_err.c_str());
    g_err = keep;
}

// Device set-up, then (once per process) the default policy and plans.
static std::once_flag g_policy_once;
static int lib_init(int dev) {
    int rc = device_init(dev);
    if (!rc)
        std::call_once(g_policy_once, [] {
            default_policy();
            default_plans();
        });
    return rc;
}

// ---------------------------------------------------------------------------
// Entry points for the device-resident natural layer (nnh_host.hpp).
// ---------------------------------------------------------------------------
int host_init() { return g_inited ? NNH_OK : lib_init(0); }
hipStream_t host_stream() { return g_dev.stream; }
std::mutex& host_mutex() { return g_dev.mu; }
int host_ws_acquire(hipStream_t st) { return ws_acquire(g_dev, st); }
void host_ws_release(hipStream_t st) { ws_release(g_dev, st); }

int dev_mul(u64* out, const u64* a, u64 an, const u64* b, u64 bn, hipStream_t st) {
    return mul_dispatch(g_dev, out, a, an, b, bn, 1, st);
}
int dev_sqr(u64* out, const u64* a, u64 an, hipStream_t st) { return sqr_device(g_dev, out, a, an, st); }

int dev_mul_scale(u64* out, const u64* a, u64 an, const u64* b, u64 bn, int s, u64 save_size, u64* n_out,
                  hipStream_t st) {
    Device& D = g_dev;
    u64 L = 1ull << s;
    if (s < MIN_MID_BITS || s > MAX_SCALE || an > L || bn > L) return fail(NNH_ERR_ARG, "bad transform scale");
    if (save_size - 1 <= L) {
        *n_out = an + bn;
        return mul_device(D, out, a, an, b, bn, st);
    }
    Plan pl = make_plan(D, s);
    int rc;
    if ((rc = ensure_twiddles(D, s)) || (rc = ensure(D.ws, 6 * pl.L))) return rc;
    if ((rc = product_core(D, pl, a, an, b, bn, false, nullptr, out, L + 2, nullptr, 0, st))) return rc;
    {
        ProfScope ps(K_FOLD, 16 * L, st);
        launch_cyclic_fold(out, L, st);
    }
    *n_out = L;
    return NNH_OK;
}

// Job tables of an internal batch, copied on st from the device's pinned
// staging buffer (job table, schoolbook list, additions).  The buffers they
// land in are shared by every internal batch: the copy is ordered on st after
// the previous batch's work, whose last launches run on st (the lanes join
// it), and a call on another stream first waits for the previous call's work
// (the workspace hand-over of the entry points).
static int upload_tables(Device& D, nnh_batch* B, hipStream_t st) {
    const BatchPlan& P = B->plan;
    const u64 jb = P.jobs.size() * sizeof(BJob), sb = (u64)P.n_school * 4, ab = P.adds.size() * sizeof(AddJob);
    const u64 o_sj = (jb + 255) & ~255ull, o_ad = (o_sj + sb + 255) & ~255ull, total = o_ad + ab;
    if (D.tab_pending) {
        HIPCHK(hipEventSynchronize(D.ev_tab));
        D.tab_pending = false;
    }
    if (total > D.tab_bytes) {
        if (D.tab_host) (void)hipHostFree(D.tab_host);
        D.tab_host = nullptr;
        D.tab_bytes = 0;
        const u64 cap = std::bit_ceil(std::max<u64>(total, 65536));
        HIPCHK(hipHostMalloc((void**)&D.tab_host, cap, hipHostMallocDefault));
        D.tab_bytes = cap;
    }
    std::memcpy(D.tab_host, P.jobs.data(), jb);
    u32* sj = reinterpret_cast<u32*>(D.tab_host + o_sj);
    for (u32 i = 0; i < P.n_school; i++) sj[i] = i;  // school jobs sort first
    if (ab) std::memcpy(D.tab_host + o_ad, P.adds.data(), ab);
    HIPCHK(hipMemcpyAsync(B->d_jobs, D.tab_host, jb, hipMemcpyHostToDevice, st));
    if (sb) HIPCHK(hipMemcpyAsync(B->d_sj, D.tab_host + o_sj, sb, hipMemcpyHostToDevice, st));
    if (ab) HIPCHK(hipMemcpyAsync(B->d_adds, D.tab_host + o_ad, ab, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(D.ev_tab, st));
    D.tab_pending = true;
    return NNH_OK;
}

// An internal batch, queued on st without waiting for it: planning the next
// call overlaps this one's kernels (a synchronising version left the GPU
// idle ~0.1 ms per call while the host planned and uploaded).
int dev_batch(const IReq* reqs, u64 n, hipStream_t st) {
    if (!n) return NNH_OK;
    // A captured upload would read the staging buffer at every replay,
    // after later calls have rewritten it.
    if (capturing(st)) return fail(NNH_ERR_ARG, "an internal batch (chunked arm) cannot be captured into a graph");
    nnh_batch* B = nullptr;
    int rc = batch_create(g_dev, reqs, n, 2 | 16, &B);
    if (rc) return rc;
    rc = upload_tables(g_dev, B, st);
    if (!rc) rc = batch_enqueue(g_dev, B, st);
    const hipError_t e = hipGetLastError();
    if (rc || e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        batch_free(B);
        return rc ? rc : fail(NNH_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
    }
    batch_free(B);  // the plan's host side only: the device tables stay in the shared buffers
    return NNH_OK;
}

}  // namespace nnh

// ===========================================================================
// C ABI
// ===========================================================================
using namespace nnh;

#define NEED_INIT()                                                              \
    do {                                                                         \
        if (!g_inited) {                                                         \
            int rc_ = lib_init(0);                                               \
            if (rc_) return rc_;                                                 \
        }                                                                        \
    } while (0)

// Start of a compute call on stream ST_: initialise, take the device lock,
// honour an injected failure, and order this call's queued work after the
// previous call's (the shared workspace).  Defines `st`.
#define ENTER(ST_)                                                               \
    NEED_INIT();                                                                 \
    std::lock_guard<std::mutex> lock_(g_dev.mu);                                 \
    if (int rc_ = injected_failure()) return rc_;                                \
    hipStream_t st = (ST_);                                                      \
    WsScope ws_scope_(g_dev, st);                                                \
    if (ws_scope_.rc) return ws_scope_.rc

extern "C" {

int nnh_init(int device) { return lib_init(device); }

int nnh_debug_fail_next(int n) { return g_fail_next.exchange(n > 0 ? n : 0); }

int nnh_policy_set(nnh_mlp* m) {
    NEED_INIT();
    if (!m) {
        std::lock_guard<std::mutex> g(g_dev.mu);
        g_policy = PolicyTable{};
        return NNH_OK;
    }
    return install_policy(m);
}

int nnh_policy_load(const char* path) {
    if (!path) return fail(NNH_ERR_ARG, "null path");
    NEED_INIT();
    return load_policy_file(path);
}

int nnh_policy_active(void) {
    std::lock_guard<std::mutex> g(g_dev.mu);
    return g_policy.on ? 1 : 0;
}

int nnh_policy_arm(uint64_t an, uint64_t bn, uint64_t count, int* arm, int* algo, int* param) {
    if (!an || !bn || !arm || !algo || !param) return fail(NNH_ERR_ARG, "bad policy query");
    std::lock_guard<std::mutex> g(g_dev.mu);
    *arm = g_policy.choose(an, bn, count);
    if (*arm < 0) {
        *algo = ALGO_REFERENCE;
        *param = 0;
    } else {
        arm_call(*arm, an, bn, algo, param);
    }
    return NNH_OK;
}

int nnh_arm_feasible(uint64_t an, uint64_t bn, float* feasible) {
    if (!an || !bn || !feasible) return fail(NNH_ERR_ARG, "bad arguments");
    arm_feasible(an, bn, feasible);
    return NNH_OK;
}

int nnh_arm_call(int arm, uint64_t an, uint64_t bn, int* algo, int* param) {
    if (arm < 0 || arm >= N_ARMS || !an || !bn || !algo || !param) return fail(NNH_ERR_ARG, "bad arguments");
    arm_call(arm, an, bn, algo, param);
    return NNH_OK;
}

int nnh_shutdown(void) {
    std::lock_guard<std::mutex> g(g_dev.mu);
    if (!g_inited) return NNH_OK;
    (void)hipDeviceSynchronize();
    g_dev.ws_stream = nullptr;
    drop_twiddles(g_dev);  // batches still alive keep the tables they use
    host_release();
    hostio_release();
    for (Buf* b : {&g_dev.ws, &g_dev.tiles, &g_dev.stage, &g_dev.tmp, &g_dev.bws, &g_dev.bjobs, &g_dev.btiles}) {
        if (b->p) (void)hipFree(b->p);
        *b = Buf{};
    }
    if (g_dev.tab_host) (void)hipHostFree(g_dev.tab_host);
    g_dev.tab_host = nullptr;
    g_dev.tab_bytes = 0;
    g_dev.tab_pending = false;
    return NNH_OK;
}

uint64_t nnh_threshold_words(void) { return NTT_MUL_THRESHOLD; }

int nnh_kernel_timing(int enable) {
    std::lock_guard<std::mutex> g(g_dev.mu);
    prof_collect();
    g_prof = enable != 0;
    for (int k = 0; k < K_NKINDS; k++) {
        g_prof_ms[k] = 0;
        g_prof_n[k] = 0;
        g_prof_bytes[k] = 0;
    }
    return NNH_OK;
}

int nnh_kernel_times(int kind, double* total_ms, uint64_t* launches, double* bytes, const char** name) {
    if (kind < 0 || kind >= K_NKINDS) return fail(NNH_ERR_ARG, "bad kernel kind");
    std::lock_guard<std::mutex> g(g_dev.mu);
    prof_collect();
    if (total_ms) *total_ms = g_prof_ms[kind];
    if (launches) *launches = g_prof_n[kind];
    if (bytes) *bytes = g_prof_bytes[kind];
    if (name) *name = KIND_NAMES[kind];
    return NNH_OK;
}

const char* nnh_last_error(void) { return g_err.c_str(); }

int nnh_mul_device_async(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn,
                         void* stream) {
    if (!d_dst || !d_a || !d_b || !an || !bn) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    int rc = mul_dispatch(g_dev, d_dst, d_a, an, d_b, bn, 1, st);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return NNH_OK;
}

int nnh_mul_device_noalloc(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn,
                           void* stream) {
    if (!d_dst || !d_a || !d_b || !an || !bn) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    if (an + bn > (1ull << NOALLOC_SCALE))
        return fail(NNH_ERR_ARG, "product of " + std::to_string(an + bn) + " words exceeds the allocation-free limit of 2^" +
                                     std::to_string(NOALLOC_SCALE) + " words");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    Device& D = g_dev;
    u64* base = D.id >= 0 && D.id < NA_MAX_DEVICES ? D.na_ptr[D.id] : nullptr;
    if (!base) return fail(NNH_ERR_NODEV, "no scratch array on this device");
    ScratchScope sc(D, base);
    int rc = mul_device(D, d_dst, d_a, an, d_b, bn, st);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return NNH_OK;
}

int nnh_sqr_device_async(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, void* stream) {
    if (!d_dst || !d_a || !an) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    int rc = sqr_device(g_dev, d_dst, d_a, an, st);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return NNH_OK;
}

int nnh_mul(uint64_t* dst, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn) {
    if (!dst || !a || !b || !an || !bn) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    int rc = stage_host(D, 2 * (an + bn) + 16);
    if (rc) return rc;
    u64* da = D.stage.p;
    u64* db = da + an;
    u64* dd = db + bn;
    if (int rc_ = host_upload(da, a, an * 8, st)) return rc_;
    if (int rc_ = host_upload(db, b, bn * 8, st)) return rc_;
    if ((rc = mul_dispatch(D, dd, da, an, db, bn, 1, st))) return rc;
    HIPCHK(hipGetLastError());
    if (int rc_ = host_download(dst, dd, (an + bn) * 8, st)) return rc_;
    HIPCHK(hipStreamSynchronize(st));
    return NNH_OK;
}

int nnh_mul_algo_device_async(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn,
                              int algo, int param, void* stream) {
    if (!d_dst || !d_a || !d_b || !an || !bn) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    int rc = mul_algo_device(g_dev, d_dst, d_a, an, d_b, bn, algo, param, st);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return NNH_OK;
}

int nnh_transform_plan(int scale, int* chunk_bits, int* n_passes, int* pass_levels) {
    if (scale < 2 || scale > MAX_SCALE || !chunk_bits || !n_passes || !pass_levels)
        return fail(NNH_ERR_ARG, "bad transform-plan arguments");
    LevelPlan lp = formula_plan(scale);
    if (g_inited) {
        std::lock_guard<std::mutex> g(g_dev.mu);
        if (g_dev.tuned[0][scale].cb) lp = g_dev.tuned[0][scale];
    }
    *chunk_bits = lp.cb;
    *n_passes = lp.np;
    for (int i = 0; i < lp.np; i++) pass_levels[i] = lp.lev[i];
    return NNH_OK;
}

int nnh_butterfly_peak(int prime, uint64_t iters, double* butterflies_per_s) {
    if (prime < 0 || prime > 2 || iters < 1 || iters > (1u << 20) || !butterflies_per_s)
        return fail(NNH_ERR_ARG, "bad butterfly-peak arguments");
    NEED_INIT();
    std::lock_guard<std::mutex> g(g_dev.mu);
    return butterfly_peak(prime, (u32)iters, butterflies_per_s);
}

int nnh_time_algo(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn, int algo,
                  int param, int count, int reps, float* ms_per_product) {
    if (!d_dst || !d_a || !d_b || !an || !bn || count < 1 || reps < 1 || !ms_per_product)
        return fail(NNH_ERR_ARG, "bad timing arguments");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    int rc = mul_algo_device(D, d_dst, d_a, an, d_b, bn, algo, param, st);  // warm-up, grows buffers
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps && !rc; r++)
        for (int c = 0; c < count && !rc; c++) rc = mul_algo_device(D, d_dst, d_a, an, d_b, bn, algo, param, st);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms_per_product = ms / (float)(reps * count);
    return NNH_OK;
}

int nnh_time_batch(uint64_t* d_dst, uint64_t dst_cap, const uint64_t* d_a, uint64_t a_cap, uint64_t an,
                   const uint64_t* d_b, uint64_t b_cap, uint64_t bn, int arm, int count, int reps,
                   float* ms_per_product) {
    if (!d_dst || !d_a || !d_b || !an || !bn || an > a_cap || bn > b_cap || arm < 0 || arm >= N_ARMS || count < 1 ||
        reps < 1 || !ms_per_product || (u64)count * (an + bn) > dst_cap)
        return fail(NNH_ERR_ARG, "bad batch timing arguments");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    // count requests with operands at different offsets (distinct B
    // pointers, so that no transform is shared) and disjoint outputs
    std::vector<IReq> rq(count);
    for (int i = 0; i < count; i++) {
        const u64 oa = ((u64)i * an) % (a_cap - an + 1), ob = ((u64)i * bn + i) % (b_cap - bn + 1);
        rq[i] = IReq{d_dst + (u64)i * (an + bn), d_a + oa, an, d_b + ob, bn, an + bn, nullptr, 0};
    }
    nnh_batch* B = nullptr;
    int rc = batch_create(D, rq.data(), rq.size(), 1 | 2, &B, arm);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0;
    auto run = [&]() -> int {
        HIPCHK(hipGraphLaunch(B->exec, st));  // warm-up
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) HIPCHK(hipGraphLaunch(B->exec, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        return NNH_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    batch_free(B);
    if (rc) return rc;
    *ms_per_product = ms / (float)(reps * count);
    return NNH_OK;
}

int nnh_mul_algo(uint64_t* dst, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn, int algo, int param) {
    if (!dst || !a || !b || !an || !bn) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    int rc = stage_host(D, 2 * (an + bn) + 16);
    if (rc) return rc;
    u64* da = D.stage.p;
    u64* db = da + an;
    u64* dd = db + bn;
    if (int rc_ = host_upload(da, a, an * 8, st)) return rc_;
    if (int rc_ = host_upload(db, b, bn * 8, st)) return rc_;
    if ((rc = mul_algo_device(D, dd, da, an, db, bn, algo, param, st))) return rc;
    HIPCHK(hipGetLastError());
    if (int rc_ = host_download(dst, dd, (an + bn) * 8, st)) return rc_;
    HIPCHK(hipStreamSynchronize(st));
    return NNH_OK;
}

int nnh_sqr(uint64_t* dst, const uint64_t* a, uint64_t an) {
    if (!dst || !a || !an) return fail(NNH_ERR_ARG, "null pointer or empty operand");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    int rc = stage_host(D, 3 * an + 16);
    if (rc) return rc;
    u64* da = D.stage.p;
    u64* dd = da + an;
    if (int rc_ = host_upload(da, a, an * 8, st)) return rc_;
    if ((rc = sqr_device(D, dd, da, an, st))) return rc;
    HIPCHK(hipGetLastError());
    if (int rc_ = host_download(dst, dd, 2 * an * 8, st)) return rc_;
    HIPCHK(hipStreamSynchronize(st));
    return NNH_OK;
}

// ---- device plumbing (so callers need no second HIP runtime) ----------------
int nnh_device_alloc(void** ptr, uint64_t bytes) {
    if (!ptr) return fail(NNH_ERR_ARG, "null pointer");
    NEED_INIT();
    *ptr = nullptr;
    HIPCHK(hipMalloc(ptr, bytes ? bytes : 8));
    return NNH_OK;
}
int nnh_device_free(void* ptr) {
    if (!ptr) return NNH_OK;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(ptr));
    return NNH_OK;
}
int nnh_copy_to_device(void* dst, const void* src, uint64_t bytes) {
    NEED_INIT();
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return NNH_OK;
}
int nnh_copy_to_host(void* dst, const void* src, uint64_t bytes) {
    NEED_INIT();
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return NNH_OK;
}
int nnh_copy_device(void* dst, const void* src, uint64_t bytes, void* stream) {
    NEED_INIT();
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream ? (hipStream_t)stream : g_dev.stream));
    return NNH_OK;
}
int nnh_stream_create(void** stream) {
    if (!stream) return fail(NNH_ERR_ARG, "null pointer");
    NEED_INIT();
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
    *stream = st;
    return NNH_OK;
}
int nnh_stream_destroy(void* stream) {
    if (!stream) return NNH_OK;
    {
        // a later stream may reuse the address: the workspace hand-over must not mistake it for this one
        std::lock_guard<std::mutex> g(g_dev.mu);
        if (g_dev.ws_stream == (hipStream_t)stream) {
            HIPCHK(hipStreamSynchronize((hipStream_t)stream));
            g_dev.ws_stream = nullptr;
        }
    }
    HIPCHK(hipStreamDestroy((hipStream_t)stream));
    return NNH_OK;
}
int nnh_synchronize(void* stream) {
    NEED_INIT();
    if (stream) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    else HIPCHK(hipDeviceSynchronize());
    return NNH_OK;
}
int nnh_event_create(void** ev) {
    if (!ev) return fail(NNH_ERR_ARG, "null pointer");
    NEED_INIT();
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    *ev = e;
    return NNH_OK;
}
int nnh_event_record(void* ev, void* stream) {
    HIPCHK(hipEventRecord((hipEvent_t)ev, stream ? (hipStream_t)stream : g_dev.stream));
    return NNH_OK;
}
int nnh_event_elapsed_ms(void* start, void* stop, float* ms) {
    HIPCHK(hipEventSynchronize((hipEvent_t)stop));
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return NNH_OK;
}
int nnh_event_destroy(void* ev) {
    if (ev) HIPCHK(hipEventDestroy((hipEvent_t)ev));
    return NNH_OK;
}

int nnh_batch_create(const nnh_req* reqs, uint64_t n, int flags, nnh_batch** out) {
    if (!reqs || !n || !out) return fail(NNH_ERR_ARG, "empty batch");
    NEED_INIT();
    std::lock_guard<std::mutex> g(g_dev.mu);
    *out = nullptr;
    if (int rc = injected_failure()) return rc;
    if (flags & ~(NNH_BATCH_GRAPH | NNH_BATCH_REFERENCE)) return fail(NNH_ERR_ARG, "unknown batch flags");
    std::vector<IReq> ir(n);
    for (u64 i = 0; i < n; i++)
        ir[i] = IReq{reqs[i].dst, reqs[i].a, reqs[i].an, reqs[i].b, reqs[i].bn, reqs[i].an + reqs[i].bn, nullptr, 0};
    return batch_create(g_dev, ir.data(), n, (flags & NNH_BATCH_GRAPH ? 1 : 0) | (flags & NNH_BATCH_REFERENCE ? 0 : 4),
                        out);
}

int nnh_batch_run(nnh_batch* b, void* stream) {
    if (!b) return fail(NNH_ERR_ARG, "null batch");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    if (b->exec) HIPCHK(hipGraphLaunch(b->exec, st));
    else if (int rc = batch_enqueue(g_dev, b, st)) return rc;
    HIPCHK(hipGetLastError());
    return NNH_OK;
}

int nnh_batch_destroy(nnh_batch* b) {
    std::lock_guard<std::mutex> g(g_dev.mu);
    batch_free(b);
    return NNH_OK;
}

int nnh_mul_batch(const nnh_req* reqs, uint64_t n, void* stream) {
    nnh_batch* b = nullptr;
    int rc = nnh_batch_create(reqs, n, 0, &b);
    if (rc) return rc;
    rc = nnh_batch_run(b, stream);
    if (!rc) {
        hipStream_t st = stream ? (hipStream_t)stream : g_dev.stream;
        if (hipStreamSynchronize(st) != hipSuccess) rc = fail(NNH_ERR_HIP, "batch synchronisation failed");
    }
    nnh_batch_destroy(b);
    return rc;
}

int nnh_alpha(int scale, uint64_t out[3]) {
    if (!out || scale < MIN_MID_BITS || scale > MAX_SCALE) return fail(NNH_ERR_ARG, "bad scale");
    NEED_INIT();
    for (int i = 0; i < 3; i++) out[i] = alpha_plain(g_dev, i, scale);
    return NNH_OK;
}

int nnh_transform_forward(uint64_t* out, const uint64_t* a, uint64_t an, int scale) {
    if (!out || !a || !an || scale < MIN_MID_BITS || scale > MAX_SCALE || an > (1ull << scale))
        return fail(NNH_ERR_ARG, "bad transform arguments");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, scale);
    int rc;
    if ((rc = ensure_twiddles(D, scale)) || (rc = ensure(D.ws, 6 * pl.L)) || (rc = stage_host(D, an + 16)))
        return rc;
    if (int rc_ = host_upload(D.stage.p, a, an * 8, st)) return rc_;
    if ((rc = transform_full(D, pl, D.stage.p, an, D.ws.p, st))) return rc;
    HIPCHK(hipGetLastError());
    if (int rc_ = host_download(out, D.ws.p, 3 * pl.L * 8, st)) return rc_;
    HIPCHK(hipStreamSynchronize(st));
    return NNH_OK;
}

int nnh_mul_transformed(uint64_t* dst, const uint64_t* fa, uint64_t an, const uint64_t* b, uint64_t bn,
                        int scale) {
    if (!dst || !fa || !b || !an || !bn || scale < MIN_MID_BITS || scale > MAX_SCALE ||
        an + bn - 1 > (1ull << scale))
        return fail(NNH_ERR_ARG, "bad arguments");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, scale);
    int rc;
    if ((rc = ensure_twiddles(D, scale)) || (rc = ensure(D.ws, 6 * pl.L)) ||
        (rc = stage_host(D, bn + an + bn + 16)))
        return rc;
    u64* fsaved = D.ws.p + 3 * pl.L;  // the reused transform plays operand B
    u64* db = D.stage.p;
    u64* dd = db + bn;
    if (int rc_ = host_upload(fsaved, fa, 3 * pl.L * 8, st)) return rc_;
    if (int rc_ = host_upload(db, b, bn * 8, st)) return rc_;
    if ((rc = product_core(D, pl, db, bn, nullptr, an, false, fsaved, dd, an + bn, nullptr, 0, st))) return rc;
    HIPCHK(hipGetLastError());
    if (int rc_ = host_download(dst, dd, (an + bn) * 8, st)) return rc_;
    HIPCHK(hipStreamSynchronize(st));
    return NNH_OK;
}

int nnh_mul_cyclic(uint64_t* dst, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn, int scale) {
    if (!dst || !a || !b || !an || !bn || scale < MIN_MID_BITS || scale > MAX_SCALE)
        return fail(NNH_ERR_ARG, "bad arguments");
    u64 L = 1ull << scale;
    if (an > L || bn > L) return fail(NNH_ERR_ARG, "operand longer than the transform");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, scale);
    bool cyclic = an + bn - 1 > L;
    u64 n_out = cyclic ? L + 2 : an + bn;
    int rc;
    if ((rc = ensure_twiddles(D, scale)) || (rc = ensure(D.ws, 6 * pl.L)) ||
        (rc = stage_host(D, an + bn + n_out + 16)))
        return rc;
    u64* da = D.stage.p;
    u64* db = da + an;
    u64* dd = db + bn;
    if (int rc_ = host_upload(da, a, an * 8, st)) return rc_;
    if (int rc_ = host_upload(db, b, bn * 8, st)) return rc_;
    if ((rc = product_core(D, pl, da, an, db, bn, false, nullptr, dd, n_out, nullptr, 0, st))) return rc;
    if (cyclic) {
        ProfScope ps(K_FOLD, 16 * L, st);
        launch_cyclic_fold(dd, L, st);
    }
    HIPCHK(hipGetLastError());
    if (int rc_ = host_download(dst, dd, (cyclic ? L : an + bn) * 8, st)) return rc_;
    HIPCHK(hipStreamSynchronize(st));
    return NNH_OK;
}


// ---- level-split plans and their autotuner ---------------------------------

int nnh_plan_set(int scale, int batched, int chunk_bits, int n_passes, const int* pass_levels) {
    if (scale < MIN_MID_BITS || scale > MAX_SCALE || (batched != 0 && batched != 1))
        return fail(NNH_ERR_ARG, "bad plan scale");
    LevelPlan lp;
    if (chunk_bits) {
        if (n_passes < 0 || n_passes > MAX_PASSES || (n_passes && !pass_levels))
            return fail(NNH_ERR_ARG, "bad plan");
        lp.cb = chunk_bits;
        lp.np = n_passes;
        for (int k = 0; k < n_passes; k++) lp.lev[k] = pass_levels[k];
        if (!plan_valid(scale, lp)) return fail(NNH_ERR_ARG, "plan not valid at this scale");
    }
    NEED_INIT();
    std::lock_guard<std::mutex> g(g_dev.mu);
    (void)hipDeviceSynchronize();  // queued launches were planned with the previous split
    g_dev.tuned[batched][scale] = lp;
    return NNH_OK;
}

int nnh_plan_get(int scale, int batched, int* chunk_bits, int* n_passes, int* pass_levels) {
    if (scale < MIN_MID_BITS || scale > MAX_SCALE || (batched != 0 && batched != 1) || !chunk_bits || !n_passes ||
        !pass_levels)
        return fail(NNH_ERR_ARG, "bad plan query");
    LevelPlan lp = formula_plan(scale);
    int tuned = 0;
    if (g_inited) {
        std::lock_guard<std::mutex> g(g_dev.mu);
        if (g_dev.tuned[batched][scale].cb) {
            lp = g_dev.tuned[batched][scale];
            tuned = 1;
        }
    }
    *chunk_bits = lp.cb;
    *n_passes = lp.np;
    for (int k = 0; k < MAX_PASSES; k++) pass_levels[k] = k < lp.np ? lp.lev[k] : 0;
    return tuned;
}

int nnh_plans_save(const char* path) {
    if (!path) return fail(NNH_ERR_ARG, "null path");
    NEED_INIT();
    std::lock_guard<std::mutex> g(g_dev.mu);
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(NNH_ERR_ARG, std::string("cannot write ") + path);
    std::fprintf(f, "nnh-plans 1 %s\n", device_key().c_str());
    for (int b = 0; b < 2; b++)
        for (int sc = 0; sc <= MAX_SCALE; sc++) {
            const LevelPlan& lp = g_dev.tuned[b][sc];
            if (lp.cb)
                std::fprintf(f, "%s %d %d %d %d %d %d %d\n", b ? "batch" : "single", sc, lp.cb, lp.np, lp.lev[0],
                             lp.lev[1], lp.lev[2], lp.lev[3]);
        }
    std::fclose(f);
    return NNH_OK;
}

int nnh_plans_load(const char* path) {
    if (!path) return fail(NNH_ERR_ARG, "null path");
    NEED_INIT();
    return load_plans_file(path);
}

int nnh_plan_candidates(int scale, nnh_plan_time* res, int cap, int* n) {
    if (scale < MIN_MID_BITS || scale > MAX_SCALE || !n || (cap > 0 && !res)) return fail(NNH_ERR_ARG, "bad arguments");
    const std::vector<LevelPlan> cands = plan_candidates(scale);
    *n = (int)cands.size();
    for (int c = 0; c < std::min<int>(cap, (int)cands.size()); c++) {
        res[c] = nnh_plan_time{cands[c].cb, cands[c].np, {cands[c].lev[0], cands[c].lev[1], cands[c].lev[2], cands[c].lev[3]},
                               0.f};
    }
    return NNH_OK;
}

// Single products: every candidate split of `scale` timed on a product of
// two 2^(scale-1)-word operands, each product followed by an event record
// as the entry points do.  The device first runs products for about 100 ms
// (clocks settle), then `rounds` rounds time every candidate in turn over
// `reps` products, and each candidate keeps its fastest round.  A candidate
// replaces the formula only when it is faster by more than `margin`.
// Batch buckets (batched = 1): a batch of about 2^22 transform words of
// such products per run, replayed as a graph.
int nnh_tune_plans(int scale, int batched, int reps, nnh_plan_time* res, int cap, int* n) {
    if (scale < MIN_MID_BITS || scale > MAX_SCALE || (batched != 0 && batched != 1) || reps < 1 || !n ||
        (cap > 0 && !res))
        return fail(NNH_ERR_ARG, "bad tuning arguments");
    ENTER(g_dev.stream);
    Device& D = g_dev;
    constexpr int ROUNDS = 3;
    constexpr float MARGIN = 0.01f;
    const std::vector<LevelPlan> cands = plan_candidates(scale);
    const u64 L = 1ull << scale, half = L / 2;
    const u64 count = batched ? std::max<u64>(1, (1ull << 22) >> scale) : 1;
    const u64 words = count * 2 * L;  // operands (2 half) and output (L) per product
    u64* buf = nullptr;
    HIPCHK(hipMalloc((void**)&buf, words * sizeof(u64)));
    // random operands: the power drawn (and so the clock) depends on the bits switched
    launch_fill_hash(buf, words, 0x4E4E2E19ull + (u64)scale, st);
    const LevelPlan saved = D.tuned[batched][scale];
    std::vector<float> ms(cands.size(), 0.f);
    hipEvent_t e0 = nullptr, e1 = nullptr, ec = nullptr;
    int rc = NNH_OK;
    std::vector<nnh_batch*> graphs(cands.size(), nullptr);
    auto finish = [&](int code) {
        for (nnh_batch* B : graphs) batch_free(B);
        for (hipEvent_t e : {e0, e1, ec})
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamSynchronize(st);
        (void)hipFree(buf);
        if (code) D.tuned[batched][scale] = saved;
        return code;
    };
    if ((rc = ensure_twiddles(D, scale)) || (!batched && (rc = ensure(D.ws, 6 * L)))) return finish(rc);
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipEventCreateWithFlags(&ec, hipEventDisableTiming) != hipSuccess)
        return finish(fail(NNH_ERR_HIP, "tuning events"));
    std::vector<IReq> rq(count);
    for (u64 i = 0; i < count; i++) {
        u64* base = buf + i * 2 * L;
        rq[i] = IReq{base + L, base, half, base + half, half, L, nullptr, 0};
    }
    if (batched)  // one graph (with its own buffers) per candidate, planned while its split is installed
        for (size_t c = 0; c < cands.size() && !rc; c++) {
            D.tuned[1][scale] = cands[c];
            rc = batch_create(D, rq.data(), count, 1, &graphs[c], ARM_NTT);
        }
    if (rc) return finish(rc);
    // one run of candidate c: a product or a graph replay
    auto run = [&](size_t c) -> int {
        if (batched) return hipGraphLaunch(graphs[c]->exec, st) == hipSuccess ? NNH_OK : fail(NNH_ERR_HIP, "tuning launch");
        D.tuned[0][scale] = cands[c];
        const Plan pl = make_plan(D, scale);
        int r = product_core(D, pl, rq[0].a, half, rq[0].b, half, false, nullptr, rq[0].dst, L, nullptr, 0, st);
        if (!r && hipEventRecord(ec, st) != hipSuccess) r = fail(NNH_ERR_HIP, "tuning event");
        return r;
    };
    // settle the clocks: about 100 ms of the formula's runs
    {
        float t = 0.f;
        if ((rc = run(0)) || hipEventRecord(e0, st) != hipSuccess || (rc = run(0)) || hipEventRecord(e1, st) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess)
            return finish(rc ? rc : fail(NNH_ERR_HIP, "tuning event"));
        (void)hipEventElapsedTime(&t, e0, e1);
        const int warm = (int)std::min(2000.f, 100.f / std::max(t, 1e-3f));
        for (int w = 0; w < warm && !rc; w++) rc = run(0);
        if (rc) return finish(rc);
    }
    for (int round = 0; round < ROUNDS && !rc; round++)
        for (size_t c = 0; c < cands.size() && !rc; c++) {
            float t = 0.f;
            if ((rc = run(c))) break;  // one untimed run with this split first
            if (hipEventRecord(e0, st) != hipSuccess) rc = fail(NNH_ERR_HIP, "tuning event");
            for (int r = 0; r < reps && !rc; r++) rc = run(c);
            if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                rc = fail(NNH_ERR_HIP, "tuning event");
            if (rc) break;
            (void)hipEventElapsedTime(&t, e0, e1);
            t /= (float)(reps * count);
            ms[c] = round == 0 ? t : std::min(ms[c], t);
        }
    if (rc) return finish(rc);
    size_t best = std::min_element(ms.begin(), ms.end()) - ms.begin();
    if (ms[best] > ms[0] * (1.f - MARGIN)) best = 0;
    D.tuned[batched][scale] = best == 0 ? LevelPlan{} : cands[best];
    *n = (int)cands.size();
    for (int c = 0; c < std::min<int>(cap, (int)cands.size()); c++) {
        res[c].chunk_bits = cands[c].cb;
        res[c].n_passes = cands[c].np;
        for (int k = 0; k < 4; k++) res[c].pass_levels[k] = cands[c].lev[k];
        res[c].ms = ms[c];
    }
    return finish(NNH_OK);
}

// ---- transform handles (mul_save / mul_load / sqr_save, natural.h:452-527) --

int nnh_xf_create(int scale, nnh_xf** out) {
    if (!out || scale < MIN_MID_BITS || scale > MAX_SCALE) return fail(NNH_ERR_ARG, "bad transform scale");
    NEED_INIT();
    *out = nullptr;
    nnh_xf* x = new nnh_xf();
    x->scale = scale;
    hipError_t e = hipMalloc((void**)&x->f, (3ull << scale) * sizeof(u64));
    if (e != hipSuccess) {
        delete x;
        return fail(e == hipErrorOutOfMemory ? NNH_ERR_NOMEM : NNH_ERR_HIP,
                    std::string("transform allocation: ") + hipGetErrorString(e));
    }
    *out = x;
    return NNH_OK;
}

int nnh_xf_destroy(nnh_xf* x) {
    if (!x) return NNH_OK;
    (void)hipDeviceSynchronize();  // queued work may still read the transform
    (void)hipFree(x->f);
    delete x;
    return NNH_OK;
}

int nnh_xf_forward(nnh_xf* x, const uint64_t* d_a, uint64_t an, void* stream) {
    if (!x || !d_a || !an || an > (1ull << x->scale)) return fail(NNH_ERR_ARG, "bad transform arguments");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, x->scale);
    int rc;
    if ((rc = ensure_twiddles(D, x->scale))) return rc;
    if ((rc = transform_full(D, pl, d_a, an, x->f, st))) return rc;
    HIPCHK(hipGetLastError());
    x->held = an;
    return NNH_OK;
}

// The save rule of NTT_info::save (ntt_info.h:237-248): a product of
// size = an + bn words whose size - 1 exceeds the transform comes out modulo
// B^L - 1 (L words), otherwise exactly (size words).
static int xf_save(Device& D, const Plan& pl, u64 size, u64* d_dst, uint64_t* n_out, hipStream_t st,
                   const std::function<int(u64)>& product) {
    const bool cyclic = size - 1 > pl.L;
    int rc = product(cyclic ? pl.L + 2 : size);
    if (rc) return rc;
    if (cyclic) {
        ProfScope ps(K_FOLD, 16 * pl.L, st);
        launch_cyclic_fold(d_dst, pl.L, st);
    }
    HIPCHK(hipGetLastError());
    if (n_out) *n_out = cyclic ? pl.L : size;
    return NNH_OK;
}

int nnh_xf_mul(uint64_t* d_dst, uint64_t* n_out, const nnh_xf* xa, const uint64_t* d_b, uint64_t bn, void* stream) {
    if (!d_dst || !xa || !xa->held || !d_b || !bn || bn > (1ull << xa->scale))
        return fail(NNH_ERR_ARG, "bad transform-product arguments");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, xa->scale);
    int rc;
    if ((rc = ensure_twiddles(D, xa->scale)) || (rc = ensure(D.ws, 6 * pl.L))) return rc;
    // the kept transform plays operand B of the product
    return xf_save(D, pl, xa->held + bn, d_dst, n_out, st, [&](u64 n) {
        return product_core(D, pl, d_b, bn, nullptr, xa->held, false, xa->f, d_dst, n, nullptr, 0, st);
    });
}

int nnh_xf_mul2(uint64_t* d_dst, uint64_t* n_out, const nnh_xf* xa, const nnh_xf* xb, void* stream) {
    if (!d_dst || !xa || !xb || !xa->held || !xb->held || xa->scale != xb->scale)
        return fail(NNH_ERR_ARG, "bad transform-product arguments");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, xa->scale);
    int rc;
    if ((rc = ensure_twiddles(D, xa->scale)) || (rc = ensure(D.ws, 3 * pl.L))) return rc;
    return xf_save(D, pl, xa->held + xb->held, d_dst, n_out, st, [&](u64 n) {
        return product_transformed(D, pl, xa->f, xa->held, xb->f, xb->held, false, d_dst, n, st);
    });
}

int nnh_xf_sqr(uint64_t* d_dst, uint64_t* n_out, const nnh_xf* xa, void* stream) {
    if (!d_dst || !xa || !xa->held) return fail(NNH_ERR_ARG, "bad transform-square arguments");
    ENTER(stream ? (hipStream_t)stream : g_dev.stream);
    Device& D = g_dev;
    Plan pl = make_plan(D, xa->scale);
    int rc;
    if ((rc = ensure_twiddles(D, xa->scale)) || (rc = ensure(D.ws, 3 * pl.L))) return rc;
    return xf_save(D, pl, 2 * xa->held, d_dst, n_out, st, [&](u64 n) {
        return product_transformed(D, pl, xa->f, xa->held, nullptr, xa->held, true, d_dst, n, st);
    });
}

}  // extern "C"
