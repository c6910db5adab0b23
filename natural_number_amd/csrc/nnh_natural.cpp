// nnh_natural.cpp — device-resident naturals and the reference algorithms built
// on the GPU multiply: Newton division (natural.h:566-769), pow
// (natural.h:554-564), decimal output (natural.h:904-1039) and the e
// workload (calc_e natural.h:841-857, recipe test.cpp:12-19).
//
// Limbs live in device memory from input to output.  The host walks the
// reference's control flow, which depends on operand sizes and a few
// leading words; those are read back through small pinned copies.  Every
// product, addition, shift and comparison runs as a device kernel.  Sizes
// follow the reference exactly (normalised where it normalises, structural
// where it resizes), because they steer the Newton iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nnh.h"
#include "nnh_host.hpp"
#include "nnh_nat.hpp"

namespace nnh {
namespace {

typedef unsigned __int128 u128;

constexpr u64 NTT_DIV_THRESHOLD = 256;   // ntt_info.h:7
constexpr int NTT_SCALE_THRESHOLD = 10;  // ntt_info.h:8
constexpr u64 DEC_BASE0 = 1000000000000000000ull;  // natural.h:904
constexpr int DEC_WG_MAX_WORDS = 256;    // bases split by one work-group (kernel limit)
constexpr int DEC_LEAF_LEVEL = 3;        // pieces < 10^(18*16) become digits per thread
constexpr int E_WG_MAX_WORDS = 256;      // e levels combined by one work-group
constexpr int E_COMBINE_MAX_WORDS = 512;  // e levels whose combines run one per work-group
constexpr int E_LEAF_MAX_WORDS = 14;

int ntt_scale(u64 s1, u64 s2) { return 64 - std::countl_zero(s1 + s2 - 2); }  // natural.h:384-386

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }
// NNH_TRACE=1 prints phase timings of the long computations to stderr.
const bool g_trace = std::getenv("NNH_TRACE") != nullptr;

struct Err {
    int code;
    std::string msg;
};
[[noreturn]] void raise(int code, const std::string& m) { throw Err{code, m}; }
void hipchk(hipError_t e, const char* what) {
    if (e != hipSuccess)
        raise(e == hipErrorOutOfMemory ? NNH_ERR_NOMEM : NNH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
void rcchk(int rc) {
    if (rc) throw Err{rc, nnh_last_error()};
}

// Device buffers recycled by power-of-two capacity.  One pool serves every
// call for the life of the process (calls hold host_mutex() and end with
// their stream synchronised, so a buffer is reused only after the work that
// used it); nnh_shutdown releases it.  Within a call, reuse is safe because
// every kernel runs on one stream.
class Pool {
   public:
    void release() {
        (void)hipDeviceSynchronize();
        for (u64* p : all_) (void)hipFree(p);
        all_.clear();
        free_.clear();
    }
    // Free the idle buffers (those not handed out).
    void trim() {
        if (free_.empty()) return;
        (void)hipDeviceSynchronize();
        for (auto& [cap, p] : free_) {
            (void)hipFree(p);
            all_.erase(std::find(all_.begin(), all_.end(), p));
        }
        free_.clear();
    }
    u64* get(u64 words, u64& cap) {
        cap = std::bit_ceil(std::max<u64>(words, 64));
        auto it = free_.find(cap);
        if (it != free_.end()) {
            u64* p = it->second;
            free_.erase(it);
            return p;
        }
        u64* p = nullptr;
        hipError_t e = hipMalloc((void**)&p, cap * sizeof(u64));
        if (e == hipErrorOutOfMemory && !free_.empty()) {
            // idle buffers of other size classes may hold the memory
            (void)hipGetLastError();
            trim();
            e = hipMalloc((void**)&p, cap * sizeof(u64));
        }
        hipchk(e, "device allocation");
        all_.push_back(p);
        return p;
    }
    void put(u64* p, u64 cap) { free_.emplace(cap, p); }

   private:
    std::multimap<u64, u64*> free_;
    std::vector<u64*> all_;
};

// Never destroyed: at process exit the HIP runtime may already be gone.
Pool& shared_pool() {
    static Pool* p = new Pool;
    return *p;
}

struct Block {
    Pool* pool;
    u64* p;
    u64 cap;
    Block(Pool* pl, u64* pp, u64 c) : pool(pl), p(pp), cap(c) {}
    Block(const Block&) = delete;
    Block& operator=(const Block&) = delete;
    ~Block() { pool->put(p, cap); }
};

// A natural in device memory: words [p, p + n), n >= 1.  Views share the
// owning block.
struct DN {
    std::shared_ptr<Block> blk;
    u64* p = nullptr;
    u64 n = 0;
    DN view(u64 off, u64 len) const {
        DN v = *this;
        v.p = p + off;
        v.n = len;
        return v;
    }
};

// Transform kept for reuse (NTT_info in mul_save/mul_load_NTT_info):
// its scale and the size of the operand it holds.
struct Xf {
    int scale = 0;
    u64 held_n = 0;
};

class Ctx {
   public:
    explicit Ctx(hipStream_t s) : st(s), pool(shared_pool()) {
        hipchk(hipHostMalloc((void**)&h_, 4096, hipHostMallocDefault), "pinned allocation");
        res_ = zeros(8);
        mailbox();
    }
    ~Ctx() {
        (void)hipStreamSynchronize(st);
        if (h_) (void)hipHostFree(h_);
        if (d_jobs_) (void)hipFree(d_jobs_);
        if (d_tiles_) (void)hipFree(d_tiles_);
        if (h_jobs_) (void)hipHostFree(h_jobs_);
    }
    hipStream_t st;
    Pool& pool;

    DN alloc(u64 n) {
        u64 cap;
        u64* p = pool.get(n, cap);
        DN x;
        x.blk = std::make_shared<Block>(&pool, p, cap);
        x.p = p;
        x.n = n;
        return x;
    }
    DN zeros(u64 n) {
        DN x = alloc(n);
        launch_fill(x.p, n, 0, st);
        return x;
    }
    DN word(u64 v) {
        DN x = alloc(1);
        launch_set_word(x.p, v, st);
        return x;
    }
    DN upload(const u64* h, u64 n) {
        DN x = alloc(n);
        rcchk(host_upload(x.p, h, n * sizeof(u64), st));
        sync();
        return x;
    }
    void download(u64* h, const u64* d, u64 n) { rcchk(host_download(h, d, n * sizeof(u64), st)); }
    void copy(u64* dst, const u64* src, u64 n) {
        if (n) hipchk(hipMemcpyAsync(dst, src, n * sizeof(u64), hipMemcpyDeviceToDevice, st), "copy");
    }
    void sync() {
        hipchk(hipGetLastError(), "launch");
        hipchk(hipStreamSynchronize(st), "synchronise");
    }
    // Small readbacks: a one-thread kernel writes the words and a sequence
    // number into pinned, host-coherent memory; the host polls for the
    // number (short waits), falling back to a stream synchronisation.
    struct Mailbox {
        u64* host = nullptr;
        u64* dev = nullptr;
        u64 seq = 0;
    };
    static Mailbox& mailbox() {
        static Mailbox m;  // process lifetime: device-written host memory is never freed
        if (!m.host) {
            hipchk(hipHostMalloc((void**)&m.host, 64, hipHostMallocMapped | hipHostMallocCoherent), "mailbox");
            hipchk(hipHostGetDevicePointer((void**)&m.dev, m.host, 0), "mailbox");
            std::memset(m.host, 0, 64);
        }
        return m;
    }
    const u64* wait_mail(u64 seq) {
        Mailbox& m = mailbox();
        volatile u64* flag = m.host + 3;
        auto t0 = Clock::now();
        for (u32 spin = 0; *flag != seq; spin++)
            if ((spin & 1023) == 1023 && ms_since(t0) > 0.5) {
                sync();
                if (*flag != seq) raise(NNH_ERR_HIP, "readback did not arrive");
                break;
            }
        std::atomic_thread_fence(std::memory_order_acquire);
        volatile const u64* v = m.host;
        for (int i = 0; i < 3; i++) mail_[i] = v[i];
        return mail_;
    }
    u64 read(const u64* p) {
        Mailbox& m = mailbox();
        launch_peek(p, 1, m.dev, ++m.seq, st);
        return wait_mail(m.seq)[0];
    }
    u64 top(const DN& x) { return read(x.p + x.n - 1); }
    // 1 + highest qualifying index (0: none); wa/wb receive the words there.
    u64 find(const u64* a, const u64* b, u64 n, int mode, u64* wa = nullptr, u64* wb = nullptr) {
        Mailbox& m = mailbox();
        launch_find(a, b, n, mode, res_.p, m.dev, ++m.seq, st);
        const u64* r = wait_mail(m.seq);
        if (wa) *wa = r[1];
        if (wb) *wb = r[2];
        return r[0];
    }
    // natural.h:82-93; returns the top word.
    u64 normalize(DN& x) {
        if (x.n == 0) {
            x = word(0);
            return 0;
        }
        u64 w = 0;
        u64 h = find(x.p, nullptr, x.n, 0, &w);
        x.n = h ? h : 1;
        return w;
    }
    // Normalise a product or sum of normalised operands: its value is at
    // least B^(n-2) unless an operand was zero, so the top two words decide.
    u64 normalize_top(DN& x) {
        if (x.n < 2) return read(x.p);
        Mailbox& m = mailbox();
        launch_peek(x.p + x.n - 2, 2, m.dev, ++m.seq, st);
        const u64* r = wait_mail(m.seq);
        if (r[1]) return r[1];
        if (r[0]) {
            x.n -= 1;
            return r[0];
        }
        return normalize(x);
    }
    bool is_zero(const DN& x) { return x.n == 1 && read(x.p) == 0; }
    // Compare normalised naturals (natural.h:145-160).
    int cmp(const DN& a, const DN& b) {
        if (a.n != b.n) return a.n < b.n ? -1 : 1;
        u64 wa = 0, wb = 0;
        u64 h = find(a.p, b.p, a.n, 2, &wa, &wb);
        if (!h) return 0;
        return wa > wb ? 1 : -1;
    }

    // ---- additions (base.h:7-16, asmAdd1) ----------------------------------
    void run_adds(std::vector<AddJob>& jobs) {
        if (jobs.empty()) return;
        u64 tiles = 0;
        for (AddJob& J : jobs) {
            J.tile0 = tiles;
            tiles += (J.n + LIMB_TILE - 1) / LIMB_TILE;
        }
        u64 jb = jobs.size() * sizeof(AddJob);
        if (jb > jobs_cap_) {
            if (d_jobs_) (void)hipFree(d_jobs_);
            if (h_jobs_) (void)hipHostFree(h_jobs_);
            d_jobs_ = nullptr;
            h_jobs_ = nullptr;
            jobs_cap_ = std::bit_ceil(std::max<u64>(jb, 4096));
            hipchk(hipMalloc((void**)&d_jobs_, jobs_cap_), "job table");
            hipchk(hipHostMalloc((void**)&h_jobs_, jobs_cap_, hipHostMallocDefault), "job staging");
        }
        if (2 * tiles > tiles_cap_) {
            if (d_tiles_) (void)hipFree(d_tiles_);
            tiles_cap_ = std::bit_ceil(std::max<u64>(2 * tiles, 4096));
            hipchk(hipMalloc((void**)&d_tiles_, tiles_cap_), "tile flags");
        }
        std::memcpy(h_jobs_, jobs.data(), jb);
        hipchk(hipMemcpyAsync(d_jobs_, h_jobs_, jb, hipMemcpyHostToDevice, st), "job upload");
        launch_add_batch(d_jobs_, (u32)jobs.size(), tiles, d_tiles_, st);
        sync();  // the staging buffer is reused by the next call
    }
    // One addition enqueued with its job by value (no host synchronisation).
    void add_one(const AddJob& j) {
        u64 need = 2 * ((j.n + LIMB_TILE - 1) / LIMB_TILE);
        if (need > tiles_cap_) {
            sync();
            if (d_tiles_) (void)hipFree(d_tiles_);
            tiles_cap_ = std::bit_ceil(std::max<u64>(need, 4096));
            hipchk(hipMalloc((void**)&d_tiles_, tiles_cap_), "tile flags");
        }
        launch_add_one(j, d_tiles_, st);
    }
    // dst[0..as) = a + b (as >= bs), the carry dropped or stored to *cout.
    void add_to(u64* dst, const u64* a, u64 as, const u64* b, u64 bs, u32* cout = nullptr,
                const u32* cin_ptr = nullptr) {
        if (bs > as) raise(NNH_ERR_DOMAIN, "addition with a longer second operand");
        add_one(AddJob{dst, a, b, as, as, bs, 0, cin_ptr, cout, 0, 0});
    }
    // dst[0..as) = a - b (as >= bs), the borrow dropped.
    void sub_from(u64* dst, const u64* a, u64 as, const u64* b, u64 bs) {
        if (bs > as) raise(NNH_ERR_DOMAIN, "subtraction with a longer second operand");
        add_one(AddJob{dst, a, b, as, as, bs, 0, nullptr, nullptr, 1, 0});
    }
    // natural.h:303-311
    DN add(const DN& x, const DN& y) {
        const DN& a = x.n >= y.n ? x : y;
        const DN& b = x.n >= y.n ? y : x;
        DN d = alloc(a.n + 1);
        add_one(AddJob{d.p, a.p, b.p, a.n + 1, a.n, b.n, 0, nullptr, nullptr, 0, 0});
        normalize_top(d);
        return d;
    }
    // natural.h:324-342 for x >= y
    DN sub(const DN& x, const DN& y) {
        DN d = alloc(x.n);
        add_one(AddJob{d.p, x.p, y.p, x.n, x.n, y.n, 0, nullptr, nullptr, 1, 0});
        normalize(d);
        return d;
    }

    // natural.h:324-342: the larger minus the smaller (both normalised)
    DN absdiff(const DN& x, const DN& y) { return cmp(x, y) >= 0 ? sub(x, y) : sub(y, x); }

    // ---- shifts (natural.h:264-295) ------------------------------------------
    // Word shifts keep the size structural; bit shifts normalise.  a != 0.
    DN shl(const DN& a, u64 bits) {
        u64 n = bits >> 6;
        u32 b = (u32)(bits & 63);
        if (b == 0) {
            DN d = alloc(a.n + n);
            launch_fill(d.p, n, 0, st);
            copy(d.p + n, a.p, a.n);
            return d;
        }
        DN d = alloc(a.n + n + 1);
        launch_fill(d.p, n, 0, st);
        launch_shl_bits(d.p + n, a.p, a.n, b, st);
        normalize(d);
        return d;
    }
    DN shr(const DN& a, u64 bits) {
        u64 n = bits >> 6;
        if (a.n <= n) return word(0);
        u32 b = (u32)(bits & 63);
        if (b == 0) return a.view(n, a.n - n);
        DN d = alloc(a.n - n);
        launch_shr_bits(d.p, a.p + n, a.n - n, b, st);
        normalize(d);
        return d;
    }
    DN complement(const u64* p, u64 n) {
        DN d = alloc(n);
        launch_not(d.p, p, n, st);
        return d;
    }

    // ---- products ------------------------------------------------------------
    // Products of normalised operands (every caller normalises its inputs).
    DN mul(const DN& a, const DN& b) {
        DN d = alloc(a.n + b.n);
        rcchk(dev_mul(d.p, a.p, a.n, b.p, b.n, st));
        normalize_top(d);
        return d;
    }
    DN sqr(const DN& a) {
        DN d = alloc(2 * a.n);
        rcchk(dev_sqr(d.p, a.p, a.n, st));
        normalize_top(d);
        return d;
    }
    // NTT_info load/mul/INTT/save at scale s (ntt_info.h:147-248).
    DN mul_scale(const DN& a, const DN& b, int s, u64 save_size) {
        u64 L = 1ull << s;
        DN d = alloc(std::max(a.n + b.n, L + 2));
        u64 n = 0;
        rcchk(dev_mul_scale(d.p, a.p, a.n, b.p, b.n, s, save_size, &n, st));
        d.n = n;
        if (n == a.n + b.n) normalize_top(d);  // exact product
        else normalize(d);                     // residue mod B^L - 1
        return d;
    }
    // mul_save_NTT_info (natural.h:452-467): d = a * b, x keeps a's transform.
    DN mul_save(Xf& x, const DN& a, const DN& b) {
        if (x.scale < NTT_SCALE_THRESHOLD) {
            x.scale = 0;
            return mul(a, b);
        }
        x.held_n = a.n;
        return mul_scale(a, b, x.scale, a.n + b.n);
    }
    // mul_load_NTT_info (natural.h:504-515): d = a * b with a's kept transform.
    DN mul_load(const Xf& x, const DN& a, const DN& b) {
        if (x.scale == 0) return mul(a, b);
        if (a.n != x.held_n) raise(NNH_ERR_DOMAIN, "kept transform holds a different operand");
        return mul_scale(a, b, x.scale, a.n + b.n);
    }

   private:
    u64* h_ = nullptr;
    u64 mail_[3] = {0, 0, 0};
    DN res_;
    AddJob* d_jobs_ = nullptr;
    AddJob* h_jobs_ = nullptr;
    u64 jobs_cap_ = 0;
    u8* d_tiles_ = nullptr;
    u64 tiles_cap_ = 0;
};

// ---------------------------------------------------------------------------
// Division (natural.h:566-769)
// ---------------------------------------------------------------------------
// Start estimate of both reciprocals (natural.h:579-583, 622-624).
DN initial_estimate(Ctx& c, const DN& d) {
    u64 w = c.read(d.p + d.n - 2);
    u64 v = (1ull << 63) | (w >> 1);
    if (v == ~0ull) {
        u64 h[2] = {1ull << 63, 1};
        return c.upload(h, 2);
    }
    return c.word((u64)((((u128)1 << 63) << 64) / ((u128)v + 1)));
}

// natural.h:576-617 (reciprocal without a cached divisor transform)
DN reciprocal_plain(Ctx& c, const DN& divs, u64 divd_size) {
    u64 expected = divd_size + 1 - divs.n;
    DN ans = initial_estimate(c, divs);
    for (;;) {
        u64 cur = ans.n;
        if (cur >= expected + 1) return c.shr(ans, (cur - expected) << 6);
        DN tmp = divs.n > 2 * cur + 1 ? c.shr(divs, (divs.n - 2 * cur - 1) << 6)
                                      : c.shl(divs, (2 * cur + 1 - divs.n) << 6);
        Xf xa;
        xa.scale = ntt_scale(tmp.n, 1);
        tmp = c.mul_save(xa, ans, tmp);
        if (tmp.n <= cur) raise(NNH_ERR_DOMAIN, "reciprocal: product shorter than the estimate");
        tmp = c.complement(tmp.p + cur, tmp.n - cur);
        int lz = std::countl_zero(c.normalize(tmp));
        u64 nxt = std::min(2 * cur, (2 * cur - tmp.n) * 2 + (u64)((lz + 31) >> 5));
        if (nxt == NTT_DIV_THRESHOLD) nxt = NTT_DIV_THRESHOLD - 1;
        if (nxt <= cur)  // no progress: the reference would shift by a negative count
            raise(NNH_ERR_DOMAIN, "reciprocal: accuracy did not grow");
        u64 drop = 2 * cur - nxt;
        tmp = tmp.n <= drop ? c.word(1) : c.shr(tmp, drop << 6);
        tmp = c.mul_load(xa, ans, tmp);
        ans = c.shl(ans, (nxt - cur) << 6);
        if (tmp.n > cur) c.add_to(ans.p, ans.p, ans.n, tmp.p + cur, tmp.n - cur);
    }
}

// natural.h:620-687 (reciprocal keeping the divisor transform in xdiv)
DN reciprocal_ntt(Ctx& c, const DN& divisor, u64 dividend_size, Xf& xdiv) {
    u64 expected = dividend_size + 1 - divisor.n;
    DN ans = initial_estimate(c, divisor);
    for (u64 i = 1; i < divisor.n; i *= 2) {
        u64 prev = ans.n, seg_size = 2 * prev + 1;
        if (2 * i >= divisor.n) seg_size = divisor.n;
        DN seg = divisor.view(divisor.n - seg_size, seg_size);
        DN tmp;
        Xf xans;
        if (2 * i <= NTT_DIV_THRESHOLD) {
            tmp = c.mul(seg, ans);
        } else {
            int s = ntt_scale(i, i);
            xans.scale = s;
            xans.held_n = ans.n;
            xdiv.scale = s;
            xdiv.held_n = seg_size;
            tmp = c.mul_scale(ans, seg, s, seg_size + prev);
        }
        u64 n_shl = seg_size - 1 - prev;
        u64 topw = 0, h = 0;
        if (tmp.n > n_shl + 1) {
            h = c.find(tmp.p + n_shl + 1, nullptr, tmp.n - n_shl - 1, 1, &topw);
            tmp.n = n_shl + 1 + h;
        }
        if (!h) topw = c.top(tmp);
        int lz = std::countl_zero(~topw);
        u64 cur = std::min(2 * prev, (2 * prev - tmp.n + n_shl) * 2 + (u64)((lz + 31) >> 5));
        if (cur == NTT_DIV_THRESHOLD) cur -= 1;
        if (cur < prev)  // the reference's natural.h:675-677 then writes ~2^64 words
            raise(NNH_ERR_DOMAIN, "divisor outside the reference reciprocal's domain");
        n_shl += 2 * prev - cur;
        if (tmp.n <= n_shl) {
            tmp = c.word(1);
        } else {
            tmp = c.complement(tmp.p + n_shl, tmp.n - n_shl);
            c.normalize(tmp);
        }
        tmp = c.mul_load(xans, ans, tmp);
        tmp = c.shr(tmp, prev << 6);
        u64 inc = cur - prev;
        DN na = c.alloc(cur);
        if (inc >= tmp.n) {
            c.copy(na.p, tmp.p, tmp.n);
            launch_fill(na.p + tmp.n, inc - tmp.n, 0, c.st);
            c.copy(na.p + inc, ans.p, prev);
        } else {
            c.copy(na.p, tmp.p, inc);
            c.add_to(na.p + inc, ans.p, prev, tmp.p + inc, tmp.n - inc);
        }
        ans = na;
        if (cur >= expected + 1) return c.shr(ans, (cur - expected) << 6);
    }
    return c.shr(ans, 64);
}

// natural.h:689-747
void div_iterative(Ctx& c, const DN& divd, const DN& divs, DN& q, DN& r) {
    int n_shl = (std::countl_zero(c.top(divs)) + 1) & 63;
    DN b = c.shl(divs, (u64)n_shl);
    r = c.shl(divd, (u64)n_shl);
    Xf xb;
    auto t0 = Clock::now();
    DN rec = reciprocal_ntt(c, b, r.n, xb);
    if (g_trace) {
        c.sync();
        std::fprintf(stderr, "[nnh] division reciprocal (%llu words)   %9.3f ms\n", (unsigned long long)rec.n,
                     ms_since(t0));
    }
    Xf xrec;
    xrec.scale = ntt_scale(rec.n, rec.n);
    int sb = ntt_scale(b.n, 1);
    u64 Lb = 1ull << sb;
    if (xb.scale != sb && sb >= NTT_SCALE_THRESHOLD - 1) {
        xb.scale = sb;
        xb.held_n = b.n;
    }
    q = c.zeros(r.n - b.n + 1);
    bool first = true;
    DN cf_word = c.zeros(1);  // carry between the two halves of the fold
    u32* cf = (u32*)cf_word.p;
    for (int iter = 0;; iter++) {
        if (iter > 4096) raise(NNH_ERR_DOMAIN, "division: remainder did not shrink");
        u64 idx = 0;
        if (r.n > rec.n + b.n - 1) idx = r.n - rec.n - b.n + 1;
        DN top = c.shr(r, (idx + b.n - 1) << 6);
        DN tmp = first ? c.mul_save(xrec, rec, top) : c.mul_load(xrec, rec, top);
        first = false;
        tmp = c.shr(tmp, rec.n << 6);
        c.add_to(q.p + idx, q.p + idx, q.n - idx, tmp.p, tmp.n);
        tmp = c.mul_load(xb, b, tmp);
        c.sub_from(r.p + idx, r.p + idx, r.n - idx, tmp.p, tmp.n);
        c.normalize(r);
        if (xb.scale) {
            if (r.n - idx > Lb) {
                c.add_to(r.p + idx, r.p + idx, Lb, r.p + idx + Lb, r.n - idx - Lb, cf);
                c.add_to(r.p + idx, r.p + idx, Lb, nullptr, 0, nullptr, cf);
                r.n = idx + Lb;
            }
            if (r.n == idx + Lb && (c.top(r) >> 63)) r.n = idx;
            c.normalize(r);
        }
        if (idx == 0) break;
    }
    DN one = c.word(1);
    for (int fix = 0; c.cmp(r, b) >= 0; fix++) {
        if (fix >= 64) raise(NNH_ERR_DOMAIN, "division: final correction did not converge");
        q = c.add(q, one);
        r = c.sub(r, b);
    }
    r = c.shr(r, (u64)n_shl);
}

// natural.h:749-769 (_div): q and r normalised.
void divmod(Ctx& c, const DN& a, const DN& b, DN& q, DN& r) {
    if (c.is_zero(b)) {
        q = c.word(c.is_zero(a) ? 1 : 0);
        r = c.word(0);
        return;
    }
    if (c.cmp(a, b) < 0) {
        q = c.word(0);
        r = a;
        return;
    }
    if (b.n == 1) {
        u64 d0 = c.read(b.p);
        u32 l = (u32)std::countl_zero(d0);
        u64 d = d0 << l;
        u64 v = (u64)((~(u128)0) / d - ((u128)1 << 64));
        q = c.alloc(a.n);
        DN rem = c.alloc(1);
        DN scr = c.alloc(div_word_scratch_words(a.n));
        launch_div_word(q.p, a.p, a.n, d, l, v, rem.p, scr.p, c.st);
        c.normalize(q);
        r = rem;
        return;
    }
    div_iterative(c, a, b, q, r);
    c.normalize(q);
}

// natural.h:554-564
DN pow(Ctx& c, const DN& b, u64 e) {
    if (e == 0) return c.word(1);
    if (e == 1 || (b.n == 1 && c.read(b.p) < 2)) return b;
    DN p = pow(c, b, e >> 1);
    p = c.sqr(p);
    if (e & 1) p = c.mul(p, b);
    return p;
}

// ---------------------------------------------------------------------------
// Decimal output (natural.h:904-1039).  dec_base[i] = 10^(18 2^i); a piece
// below dec_base[i+1] splits into hi = x / dec_base[i] and lo = x mod
// dec_base[i] with the reference's reciprocal estimate and correction loop.
// All pieces of one level are split together; pieces known to be zero
// from the digit count (the reference's hi == 0 pruning, natural.h:978-981)
// are skipped.
// ---------------------------------------------------------------------------
struct DecBases {
    std::vector<DN> base, rec;
    std::vector<u32> n_shl;
};

void dec_big_level(Ctx& c, const DN& cur, u64 sx, u64 j0, u64 count, const DN& base, const DN& rec, u32 n_shl,
                   DN& nxt) {
    const u64 w = base.n, xn = sx - (w - 1), rn = rec.n, pn = xn + rn, m = count - j0;
    const u64 sbits = (w + 2) * 64 - n_shl;
    DN P = c.alloc(m * pn);
    std::vector<IReq> rq(m);
    for (u64 j = 0; j < m; j++)
        rq[j] = IReq{P.p + j * pn, cur.p + (j0 + j) * sx + (w - 1), xn, rec.p, rn, pn, nullptr, 0};
    rcchk(dev_batch(rq.data(), m, c.st));
    // hi -> nxt[2j], stride 2w between consecutive hi pieces
    launch_shr_pieces(nxt.p + 2 * j0 * w, 2 * w, P.p, pn, pn, sbits, w, m, c.st);
    DN Q = c.alloc(m * 2 * w);
    for (u64 j = 0; j < m; j++)
        rq[j] = IReq{Q.p + j * 2 * w, nxt.p + 2 * (j0 + j) * w, w, base.p, w, 2 * w, nullptr, 0};
    rcchk(dev_batch(rq.data(), m, c.st));
    // lo = x - hi * base over 2w words (borrow flags must stay 1)
    DN lo = c.alloc(m * 2 * w);
    DN fl = c.zeros((m + 1) / 2 + 1);
    u32* flags = (u32*)fl.p;
    std::vector<AddJob> jobs(m);
    for (u64 j = 0; j < m; j++)
        jobs[j] = AddJob{lo.p + j * 2 * w, cur.p + (j0 + j) * sx, Q.p + j * 2 * w, 2 * w, sx, 2 * w, 0, nullptr,
                         flags + j, 1, 0};
    c.run_adds(jobs);
    std::vector<u32> hf(m);
    hipchk(hipMemcpyAsync(hf.data(), flags, m * sizeof(u32), hipMemcpyDeviceToHost, c.st), "flags");
    c.sync();
    for (u64 j = 0; j < m; j++)
        if (hf[j] != 1) raise(NNH_ERR_DOMAIN, "radix conversion: quotient estimate too large");
    // lo < 3 base < B^(w+1) when the estimate is within 2 of the quotient
    // (it is for the reference's reciprocal); words w+1 .. 2w-1 must be zero.
    DN zf = c.zeros(1);
    launch_nonzero_pieces(lo.p, 2 * w, w + 1, 2 * w, m, zf.p, c.st);
    if (c.read(zf.p)) raise(NNH_ERR_DOMAIN, "radix conversion: quotient estimate too small");
    // correction: while lo >= base: hi += 1, lo -= base (natural.h:962-965)
    DN cf = c.alloc((m + 7) / 8 + 1);
    u8* cmpf = (u8*)cf.p;
    std::vector<u8> hc(m);
    for (int round = 0;; round++) {
        launch_cmp_pieces(lo.p, 2 * w, w + 1, base.p, w, m, cmpf, c.st);
        hipchk(hipMemcpyAsync(hc.data(), cmpf, m, hipMemcpyDeviceToHost, c.st), "compare flags");
        c.sync();
        jobs.clear();
        for (u64 j = 0; j < m; j++)
            if (hc[j]) {
                jobs.push_back(AddJob{lo.p + j * 2 * w, lo.p + j * 2 * w, base.p, 2 * w, 2 * w, w, 0, nullptr, nullptr,
                                      1, 0});
                u64* h = nxt.p + 2 * (j0 + j) * w;
                jobs.push_back(AddJob{h, h, nullptr, w, w, 0, 0, nullptr, nullptr, 0, 1});
            }
        if (jobs.empty()) break;
        if (round >= 8) raise(NNH_ERR_DOMAIN, "radix conversion: correction did not converge");
        c.run_adds(jobs);
    }
    // lo -> nxt[2j+1]
    launch_shr_pieces(nxt.p + (2 * j0 + 1) * w, 2 * w, lo.p, 2 * w, 2 * w, 0, w, m, c.st);
}

// Decimal digits of x (no leading zeros) written to out with a NUL; returns
// the digit count.  The digits of pieces below dec_base[0] come from the
// leaf kernel as 18-character groups of the zero-padded 18 * 2^(I+1)-digit
// representation; only the part from the first nonzero digit is copied.
u64 to_dec_into(Ctx& c, DN x, char* out, u64 cap) {
    auto t0 = Clock::now();
    auto tick = [&](const char* what) {
        if (!g_trace) return;
        c.sync();
        std::fprintf(stderr, "[nnh] to_dec %-28s %9.3f ms\n", what, ms_since(t0));
        t0 = Clock::now();
    };
    c.normalize(x);
    DecBases B;
    B.base.push_back(c.word(DEC_BASE0));
    int I = -1;  // dec_init (natural.h:1013-1018)
    while (c.cmp(x, B.base[I + 1]) >= 0) {
        B.base.push_back(c.sqr(B.base[I + 1]));
        I++;
    }
    if (I < 0) {
        std::string s = std::to_string(c.read(x.p));
        if (s.size() + 1 > cap) raise(NNH_ERR_ARG, "output buffer too small");
        std::memcpy(out, s.c_str(), s.size() + 1);
        return s.size();
    }
    tick("bases");
    B.rec.resize(I + 1);
    B.n_shl.assign(I + 1, 0);
    // calc_dec_base (natural.h:909-923) runs the reciprocal's Newton ladder
    // (natural.h:575-617) for every base; for these divisors it returns
    // rec[i] = floor(2^(64 S_i) / D_i), D_i = base[i] << n_shl[i],
    // S_i = 2 n_i + 1.  As base[i+1] = base[i]^2, (base[i] * rec[i+1]) >> t_i
    // with t_i = 64 (S_{i+1} - S_i) - n_shl[i+1] + n_shl[i] equals
    // floor(2^(64 S_i) / D_i - e), where e < 2 base[i] / 2^t_i is far below 1
    // when rec[i+1] is its floor or one less: rec[i] or rec[i] - 1, an
    // underestimate the split's correction loop absorbs like the ladder's.
    // Only the top base runs the ladder; each lower base costs one product.
    auto S = [&](int i) { return 2 * B.base[i].n + 1; };
    for (int i = DEC_LEAF_LEVEL + 1; i <= I; i++)
        B.n_shl[i] = (u32)((std::countl_zero(c.top(B.base[i])) + 1) & 63);
    if (I > DEC_LEAF_LEVEL) B.rec[I] = reciprocal_plain(c, c.shl(B.base[I], B.n_shl[I]), S(I));
    for (int i = I - 1; i > DEC_LEAF_LEVEL; i--) {
        const u64 t = 64 * (S(i + 1) - S(i)) - B.n_shl[i + 1] + B.n_shl[i];
        B.rec[i] = c.shr(c.mul(B.base[i], B.rec[i + 1]), t);
    }
    tick("reciprocals");
    const u64 total = 18ull << (I + 1);
    // digits of x: at most floor(bits * log10 2) + 1
    u64 bits = 64 * (x.n - 1) + (64 - std::countl_zero(c.top(x)));
    u64 ndig = (u64)((double)bits * 0.30102999566398120) + 2;
    if (ndig > total) ndig = total;
    DN digits_dev = c.alloc((total + 7) / 8);
    char* dd = (char*)digits_dev.p;
    DN err = c.zeros(1);
    int lvl = I;
    u64 sx = B.base[I + 1].n;
    DN cur = c.zeros(sx);
    c.copy(cur.p, x.p, x.n);
    auto zero_pieces = [&](int level) -> u64 {  // leading pieces of 18*2^(level+1) digits that are zero
        u64 pd = 18ull << (level + 1);
        return (total - ndig) / pd;
    };
    for (; lvl > DEC_LEAF_LEVEL; lvl--) {
        const DN& base = B.base[lvl];
        const u64 w = base.n;
        const u64 count = 1ull << (I - lvl);
        const u64 j0 = std::min(zero_pieces(lvl), count - 1);
        DN nxt = c.alloc(2 * count * w);
        if (j0) launch_fill(nxt.p, 2 * j0 * w, 0, c.st);
        const DN& rec = B.rec[lvl];
        if (w <= DEC_WG_MAX_WORDS && rec.n <= w + 16 && sx <= 2 * w) {
            launch_dec_split_wg(cur.p + j0 * sx, sx, count - j0, base.p, w, rec.p, rec.n, B.n_shl[lvl],
                                nxt.p + 2 * j0 * w, err.p, c.st);
            tick("level (work-group)");
        } else {
            dec_big_level(c, cur, sx, j0, count, base, rec, B.n_shl[lvl], nxt);
            tick("level (batched products)");
        }
        cur = nxt;
        sx = w;
    }
    // leaves: pieces below dec_base[lvl+1] (lvl <= 3), one thread each
    const u64 count = 1ull << (I - lvl);
    const u64 j0 = std::min(zero_pieces(lvl), count - 1);
    const u64 pd = 18ull << (lvl + 1);
    if (sx > 16) raise(NNH_ERR_DOMAIN, "radix conversion: leaf pieces too long");
    launch_dec_leaves(cur.p + j0 * sx, sx, count - j0, lvl, dd + j0 * pd, c.st);
    if (c.read(err.p)) raise(NNH_ERR_DOMAIN, "radix conversion: quotient estimate out of range");
    tick("leaves");
    // first nonzero digit: scan the leading pieces on the host
    std::vector<char> head(pd);
    u64 first = total;
    for (u64 j = j0; j < count && first == total; j++) {
        hipchk(hipMemcpyAsync(head.data(), dd + j * pd, pd, hipMemcpyDeviceToHost, c.st), "digits");
        c.sync();
        for (u64 k = 0; k < pd; k++)
            if (head[k] != '0') {
                first = j * pd + k;
                break;
            }
    }
    if (first == total) first = total - 1;  // x == 0 is handled above; keep one digit
    const u64 len = total - first;
    if (len + 1 > cap) raise(NNH_ERR_ARG, "output buffer too small (" + std::to_string(len + 1) + " bytes needed)");
    rcchk(host_download(out, dd + first, len, c.st));
    out[len] = 0;
    tick("copy out");
    return len;
}

// ---------------------------------------------------------------------------
// e by binary splitting over [0, 2^k), level-synchronous: level l holds
// the 2^(k-l) ranges of 2^l terms at a common stride S(l) words.
// ---------------------------------------------------------------------------
u64 e_stride(int k, int l) {
    // den of the last range [2^k - 2^l, 2^k) is (2^k - 1)! / (2^k - 2^l - 1)!,
    // the largest of the level; num < 3 den.
    long double e = std::ldexp(1.0L, k), b = e - std::ldexp(1.0L, l);
    long double bits = (std::lgamma(e) - (b > 0 ? std::lgamma(b) : 0.0L)) / std::log(2.0L);
    return (u64)std::ceil((double)((bits + 2.0L) / 64.0L)) + 1;
}

void calc_e(Ctx& c, int k, DN& num, DN& den) {
    if (k < 6 || k > 26) raise(NNH_ERR_ARG, "calc_e: k must be in [6, 26]");
    u64 S = e_stride(k, 5);
    if (S > E_LEAF_MAX_WORDS) raise(NNH_ERR_ARG, "calc_e: leaf stride too large");
    u64 nb = 1ull << (k - 5);
    DN cn = c.alloc(nb * S), cd = c.alloc(nb * S);
    auto t0 = Clock::now();
    launch_e_leaves(cn.p, cd.p, nb, S, c.st);
    for (int l = 6; l <= k; l++) {
        if (g_trace) {
            c.sync();
            std::fprintf(stderr, "[nnh] calc_e level %2d (stride %8llu) %9.3f ms\n", l - 1, (unsigned long long)S,
                         ms_since(t0));
            t0 = Clock::now();
        }
        u64 S2 = e_stride(k, l), ncomb = 1ull << (k - l);
        DN nn = c.alloc(ncomb * S2), nd = c.alloc(ncomb * S2);
        if (S <= E_COMBINE_MAX_WORDS && S2 <= 4096) {
            launch_e_combine(cn.p, cd.p, S, nn.p, nd.p, S2, ncomb, c.st);
        } else {
            std::vector<IReq> rq(2 * ncomb);
            for (u64 j = 0; j < ncomb; j++) {
                const u64 L = 2 * j * S, R = (2 * j + 1) * S;
                // natural.h:852-856: num = num * den1 + num1, den = den * den1
                rq[2 * j] = IReq{nn.p + j * S2, cn.p + L, S, cd.p + R, S, S2, cn.p + R, S};
                rq[2 * j + 1] = IReq{nd.p + j * S2, cd.p + L, S, cd.p + R, S, S2, nullptr, 0};
            }
            rcchk(dev_batch(rq.data(), rq.size(), c.st));
        }
        cn = nn;
        cd = nd;
        S = S2;
    }
    cn.n = S;
    cd.n = S;
    c.normalize(cn);
    c.normalize(cd);
    num = cn;
    den = cd;
}


// ---------------------------------------------------------------------------
// factorial (natural.h:795-839): n! = 2^e2 * prod over bits i (high to low)
// of [ans = ans^2 * P_i], P_i the product of the odd primes whose exponent
// in n! has bit i set.  The reference multiplies each P_i through a
// smallest-first priority queue; the product is the same in any order, so
// here it is a level-synchronous product tree on the device.
// ---------------------------------------------------------------------------
constexpr u64 TREE_LEAF_FACTORS = 16;
constexpr u64 TREE_LEAF_MAX_WORDS = 9;

DN product_tree(Ctx& c, const std::vector<u64>& f, u32 bits) {
    if (f.empty()) return c.word(1);
    if (bits > 32) raise(NNH_ERR_ARG, "product tree: factors must be below 2^32");
    u64 nleaf = std::bit_ceil((f.size() + TREE_LEAF_FACTORS - 1) / TREE_LEAF_FACTORS);
    std::vector<u64> padded(nleaf * TREE_LEAF_FACTORS, 1);
    std::copy(f.begin(), f.end(), padded.begin());
    DN df = c.upload(padded.data(), padded.size());
    // entries of level l hold 2^l leaves of 16 factors below 2^bits
    auto stride = [&](u64 leaves) { return (leaves * TREE_LEAF_FACTORS * bits + 63) / 64 + 1; };
    u64 S = std::max<u64>(stride(1), TREE_LEAF_MAX_WORDS);
    DN cur = c.alloc(nleaf * S);
    launch_tree_leaves(df.p, nleaf, cur.p, S, c.st);
    for (u64 n = nleaf, leaves = 2; n > 1; n >>= 1, leaves <<= 1) {
        u64 S2 = stride(leaves), ncomb = n / 2;
        DN nxt = c.alloc(ncomb * S2);
        if (S <= E_WG_MAX_WORDS && S2 <= 4096) {
            launch_tree_combine(cur.p, S, nxt.p, S2, ncomb, c.st);
        } else {
            std::vector<IReq> rq(ncomb);
            for (u64 j = 0; j < ncomb; j++)
                rq[j] = IReq{nxt.p + j * S2, cur.p + 2 * j * S, S, cur.p + (2 * j + 1) * S, S, S2, nullptr, 0};
            rcchk(dev_batch(rq.data(), ncomb, c.st));
        }
        cur = nxt;
        S = S2;
    }
    cur.n = S;
    c.normalize(cur);
    return cur;
}

DN factorial(Ctx& c, u64 n) {
    if (n <= 1) return c.word(1);
    if (n == 2) return c.word(2);
    if (n >= (1ull << 32)) raise(NNH_ERR_ARG, "factorial: n must be below 2^32");
    // sieve and Legendre exponents (natural.h:800-816)
    std::vector<u8> composite(n + 1, 0);
    std::vector<u64> prime, pexp;
    for (u64 i = 2; i <= n; i++) {
        if (!composite[i]) {
            prime.push_back(i);
            u64 e = 0;
            for (u64 m = n; m; m /= i) e += m / i;
            pexp.push_back(e);
            for (u64 j = i * i; j <= n; j += i) composite[j] = 1;
        }
    }
    const u32 bits = 64 - std::countl_zero(n);
    DN ans = c.word(1);
    for (int i = 63 - std::countl_zero(pexp[1]); i >= 0; i--) {  // natural.h:818-834
        std::vector<u64> f;
        for (u64 j = 1; j < prime.size(); j++) {
            if (pexp[j] < (1ull << i)) break;
            if (pexp[j] & (1ull << i)) f.push_back(prime[j]);
        }
        DN p = product_tree(c, f, bits);
        ans = c.mul(c.sqr(ans), p);
    }
    return c.shl(ans, pexp[0]);  // natural.h:835
}

// ---------------------------------------------------------------------------
// pi (natural.h:859-902, recipe test.cpp:22-32).  calc_pi runs
// level-synchronously over [1, 2^k + 1): one thread per pair of terms
// (the pair level subtracts), then levels of num = num_l den_r +
// numc_l num_r, den = den_l den_r, numc = numc_l numc_r at strides from
// per-level bit bounds.
// ---------------------------------------------------------------------------
struct PiLevel {
    u64 S;
    DN num, den, numc;
};

void calc_pi(Ctx& c, int k, DN& num, DN& den, DN& numc) {
    if (k < 2 || k > 20) raise(NNH_ERR_ARG, "calc_pi: k must be in [2, 20]");
    auto bl = [](u64 x) { return (double)(64 - std::countl_zero(x)); };
    const u64 bmax = 1ull << k;
    double bd = 3 * bl(bmax) + 54, bc = bl(2 * bmax) + 2 * bl(6 * bmax), bn = bl(0x207e2da6ull * bmax + 0xcf6371) + bc;
    double nd = 2 * bd, nc = 2 * bc, nn = std::max(bc + bn, bn + bd);  // after the pair level
    auto words = [](double bits) { return (u64)std::ceil(bits / 64.0) + 1; };
    u64 S = std::max<u64>(words(std::max({nd, nc, nn})), 5);
    u64 npairs = bmax / 2;
    DN cn = c.alloc(npairs * S), cd = c.alloc(npairs * S), cc = c.alloc(npairs * S);
    launch_pi_pairs(1, npairs, cn.p, cd.p, cc.p, S, c.st);
    for (u64 n = npairs; n > 1; n >>= 1) {
        double nn2 = std::max(nn + nd, nc + nn) + 1;
        nd *= 2;
        nc *= 2;
        nn = nn2;
        const u64 S2 = words(std::max({nd, nc, nn})), ncomb = n / 2;
        DN xn = c.alloc(ncomb * S2), xd = c.alloc(ncomb * S2), xc = c.alloc(ncomb * S2);
        if (S <= E_WG_MAX_WORDS && S2 <= 4096) {
            launch_pi_combine(cn.p, cd.p, cc.p, S, xn.p, xd.p, xc.p, S2, ncomb, c.st);
        } else {
            // natural.h:893-901: the numc_l num_r product first, then
            // num_l den_r with it as the CRT addend.  Each batch repeats one
            // B operand per combination (numc_l, then den_r), whose transform
            // the batch computes once.
            DN tmp = c.alloc(ncomb * S2);
            std::vector<IReq> rq(2 * ncomb);
            for (u64 j = 0; j < ncomb; j++) {
                const u64 L = 2 * j * S, R = (2 * j + 1) * S;
                rq[2 * j] = IReq{tmp.p + j * S2, cn.p + R, S, cc.p + L, S, S2, nullptr, 0};
                rq[2 * j + 1] = IReq{xc.p + j * S2, cc.p + R, S, cc.p + L, S, S2, nullptr, 0};
            }
            rcchk(dev_batch(rq.data(), rq.size(), c.st));
            for (u64 j = 0; j < ncomb; j++) {
                const u64 L = 2 * j * S, R = (2 * j + 1) * S;
                rq[2 * j] = IReq{xd.p + j * S2, cd.p + L, S, cd.p + R, S, S2, nullptr, 0};
                rq[2 * j + 1] = IReq{xn.p + j * S2, cn.p + L, S, cd.p + R, S, S2, tmp.p + j * S2, S2};
            }
            rcchk(dev_batch(rq.data(), rq.size(), c.st));
        }
        cn = xn;
        cd = xd;
        cc = xc;
        S = S2;
    }
    cn.n = cd.n = cc.n = S;
    c.normalize(cn);
    c.normalize(cd);
    c.normalize(cc);
    num = cn;
    den = cd;
    numc = cc;
}

// natural.h:859-872
void sqrt_10005(Ctx& c, u64 acc, DN& num, DN& den) {
    num = c.word(4001);
    den = c.word(40);
    DN k10005 = c.word(10005);
    while (num.n + den.n - 2 < acc) {
        DN c1 = c.sqr(den);
        DN c2 = c.sqr(num);
        den = c.shl(c.mul(num, den), 1);
        num = c.add(c.mul(c1, k10005), c2);
    }
}

template <class F>
int guarded(F&& f) {
    int rc = host_init();
    if (rc) return rc;
    std::lock_guard<std::mutex> g(host_mutex());
    if ((rc = injected_failure())) return rc;
    if ((rc = host_ws_acquire(host_stream()))) return rc;
    try {
        Ctx c(host_stream());
        f(c);
        c.sync();
        host_ws_release(host_stream());
        return NNH_OK;
    } catch (const Err& e) {
        (void)hipStreamSynchronize(host_stream());
        host_ws_release(host_stream());
        return set_error(e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(NNH_ERR_NOMEM, "host allocation failed");
    }
}

u64 emit(Ctx& c, const DN& x, u64* out, u64 cap) {
    if (x.n > cap) raise(NNH_ERR_ARG, "output buffer too small (" + std::to_string(x.n) + " words needed)");
    c.download(out, x.p, x.n);
    return x.n;
}

}  // namespace

void host_release() { shared_pool().release(); }
void host_trim() { shared_pool().trim(); }

}  // namespace nnh

using namespace nnh;

extern "C" {

int nnh_divmod(uint64_t* q, uint64_t q_cap, uint64_t* qn, uint64_t* r, uint64_t r_cap, uint64_t* rn,
               const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn) {
    if (!q || !r || !qn || !rn || !a || !b || !an || !bn) return set_error(NNH_ERR_ARG, "null pointer or empty operand");
    return guarded([&](Ctx& c) {
        DN da = c.upload(a, an), db = c.upload(b, bn);
        c.normalize(da);
        c.normalize(db);
        DN dq, dr;
        divmod(c, da, db, dq, dr);
        c.normalize(dr);
        *qn = emit(c, dq, q, q_cap);
        *rn = emit(c, dr, r, r_cap);
    });
}

int nnh_pow(uint64_t* dst, uint64_t cap, uint64_t* n, const uint64_t* b, uint64_t bn, uint64_t e) {
    if (!dst || !n || !b || !bn) return set_error(NNH_ERR_ARG, "null pointer or empty operand");
    return guarded([&](Ctx& c) {
        DN db = c.upload(b, bn);
        c.normalize(db);
        DN p = pow(c, db, e);
        *n = emit(c, p, dst, cap);
    });
}

int nnh_to_dec(char* out, uint64_t cap, uint64_t* len, const uint64_t* a, uint64_t an) {
    if (!out || !len || !a || !an) return set_error(NNH_ERR_ARG, "null pointer or empty operand");
    return guarded([&](Ctx& c) {
        DN da = c.upload(a, an);
        *len = to_dec_into(c, da, out, cap);
    });
}

int nnh_calc_e(uint64_t* num, uint64_t num_cap, uint64_t* num_n, uint64_t* den, uint64_t den_cap, uint64_t* den_n,
               int k) {
    if (!num || !den || !num_n || !den_n) return set_error(NNH_ERR_ARG, "null pointer");
    return guarded([&](Ctx& c) {
        DN dn, dd;
        calc_e(c, k, dn, dd);
        *num_n = emit(c, dn, num, num_cap);
        *den_n = emit(c, dd, den, den_cap);
    });
}

int nnh_factorial(uint64_t* out, uint64_t cap, uint64_t* n_out, uint64_t n) {
    if (!out || !n_out) return set_error(NNH_ERR_ARG, "null pointer");
    return guarded([&](Ctx& c) {
        DN f = factorial(c, n);
        *n_out = emit(c, f, out, cap);
    });
}

int nnh_calc_pi(uint64_t* num, uint64_t* den, uint64_t* numc, uint64_t cap, int k, uint64_t* sizes) {
    if (!num || !den || !numc || !sizes) return set_error(NNH_ERR_ARG, "null pointer");
    return guarded([&](Ctx& c) {
        DN a, b, d;
        calc_pi(c, k, a, b, d);
        sizes[0] = emit(c, a, num, cap);
        sizes[1] = emit(c, b, den, cap);
        sizes[2] = emit(c, d, numc, cap);
    });
}

int nnh_pi_digits(char* out, uint64_t cap, uint64_t* len, int k, uint64_t digits, uint64_t acc, double* stage_ms) {
    if (!out || !len) return set_error(NNH_ERR_ARG, "null pointer");
    return guarded([&](Ctx& c) {
        double ms[NNH_PI_STAGES] = {0, 0, 0, 0, 0};
        auto t = Clock::now();
        DN num, den, numc, sn, sd;
        calc_pi(c, k, num, den, numc);  // test.cpp:24
        c.sync();
        ms[0] = ms_since(t);
        t = Clock::now();
        sqrt_10005(c, acc, sn, sd);  // test.cpp:25
        c.sync();
        ms[1] = ms_since(t);
        t = Clock::now();
        num = c.absdiff(c.mul(c.word(0xcf6371), den), num);                              // test.cpp:26
        den = c.mul(den, c.mul(c.mul(c.word(0x68380), sn), pow(c, c.word(10), digits)));  // test.cpp:27
        num = c.mul(num, sd);                                                            // test.cpp:28
        c.sync();
        ms[2] = ms_since(t);
        t = Clock::now();
        DN q, r;
        divmod(c, den, num, q, r);  // test.cpp:29
        c.sync();
        ms[3] = ms_since(t);
        t = Clock::now();
        *len = to_dec_into(c, q, out, cap);  // test.cpp:31
        ms[4] = ms_since(t);
        if (stage_ms)
            for (int i = 0; i < NNH_PI_STAGES; i++) stage_ms[i] = ms[i];
    });
}

int nnh_e_digits(char* out, uint64_t cap, uint64_t* len, int k, uint64_t digits, double* stage_ms) {
    if (!out || !len) return set_error(NNH_ERR_ARG, "null pointer");
    return guarded([&](Ctx& c) {
        double ms[NNH_E_STAGES] = {0, 0, 0, 0, 0};
        auto t = Clock::now();
        DN num, den;
        calc_e(c, k, num, den);  // test.cpp:14
        c.sync();
        ms[0] = ms_since(t);
        t = Clock::now();
        DN p = pow(c, c.word(10), digits);  // test.cpp:15
        c.sync();
        ms[1] = ms_since(t);
        t = Clock::now();
        num = c.mul(num, p);
        c.sync();
        ms[2] = ms_since(t);
        t = Clock::now();
        DN q, r;
        divmod(c, num, den, q, r);  // test.cpp:16
        c.sync();
        ms[3] = ms_since(t);
        t = Clock::now();
        *len = to_dec_into(c, q, out, cap);  // test.cpp:18-19
        ms[4] = ms_since(t);
        if (stage_ms)
            for (int i = 0; i < NNH_E_STAGES; i++) stage_ms[i] = ms[i];
    });
}

}  // extern "C"
