// nnh_plan.cpp — host-only planning: algorithm arms, the cost-model policy
// table and batch plans (see nnh_plan.hpp).
#include "nnh_plan.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <utility>

namespace nnh {

// ---- level splits -----------------------------------------------------------

static void balance(int s, int cb, int np, bool deeper_first, LevelPlan& p) {
    p.cb = cb;
    p.np = np;
    int rem = s - cb;
    for (int k = 0; k < np; k++) {
        const int nl = (rem + (np - k) - 1) / (np - k);
        p.lev[k] = nl;
        rem -= nl;
    }
    if (!deeper_first) std::reverse(p.lev, p.lev + np);
}

LevelPlan formula_plan(int s) {
    LevelPlan p;
    int cb;
    if (s <= 20) {
        cb = std::min(s, TILE_BITS);
    } else {
        const int npass = (s - TILE_BITS + 7) / 8;
        cb = std::max(MIN_CHUNK_BITS, s - 8 * npass);
    }
    balance(s, cb, (s - cb + 7) / 8, true, p);
    return p;
}

bool plan_valid(int s, const LevelPlan& p) {
    if (s < MIN_MID_BITS || s > MAX_SCALE || p.cb < MIN_MID_BITS || p.cb > TILE_BITS || p.cb > s) return false;
    // up to one tile, one chunk; beside top passes, chunks of at least one wave
    if (s <= TILE_BITS && p.cb != s) return false;
    if (s > TILE_BITS && p.cb < MIN_CHUNK_BITS) return false;
    if (p.np < 0 || p.np > MAX_PASSES || (p.np == 0) != (p.cb == s)) return false;
    int sum = 0;
    for (int k = 0; k < p.np; k++) {
        if (p.lev[k] < 1 || p.lev[k] > 8) return false;
        sum += p.lev[k];
    }
    if (sum != s - p.cb) return false;
    // the lowest pass starts at bit cb; its tile of 2^(12 - lev) columns must fit one stride block
    return p.np == 0 || p.cb + p.lev[p.np - 1] >= TILE_BITS;
}

std::vector<LevelPlan> plan_candidates(int s) {
    std::vector<LevelPlan> out;
    auto add = [&](const LevelPlan& p) {
        if (!plan_valid(s, p)) return;
        for (const LevelPlan& q : out)
            if (q.cb == p.cb && q.np == p.np && std::equal(q.lev, q.lev + q.np, p.lev)) return;
        out.push_back(p);
    };
    add(formula_plan(s));
    for (int cb = MIN_CHUNK_BITS; cb <= std::min(s, TILE_BITS); cb++) {
        const int top = s - cb;
        const int np0 = (top + 7) / 8;
        for (int np = np0; np <= std::min(np0 + 1, MAX_PASSES); np++) {
            if (np == 0 && top) continue;
            LevelPlan p;
            balance(s, cb, np, true, p);
            add(p);
            balance(s, cb, np, false, p);
            add(p);
            // deepest first: 8-level passes, the rest last
            if (np >= 2) {
                LevelPlan q;
                q.cb = cb;
                q.np = np;
                int rem = top;
                for (int k = 0; k < np; k++) {
                    q.lev[k] = std::min(8, rem - (np - 1 - k));
                    rem -= q.lev[k];
                }
                add(q);
            }
        }
    }
    return out;
}

int chunk_scale(int arm, u64 small) { return std::max(ntt_scale(small, small), MIN_MID_BITS) + (arm == ARM_NTT_CHUNK1); }

u64 chunk_pieces(u64 big, u64 small, int sp) {
    const u64 L = 1ull << sp;
    if (L < small) return 0;
    const u64 piece = L - small + 1;
    return (big + piece - 1) / piece;
}

static bool mac_ok(u64 big, u64 small) { return small == 0 || big <= SCHOOL_MAX_MAC / small; }

void arm_feasible(u64 an, u64 bn, float f[N_ARMS]) {
    const u64 big = std::max(an, bn), small = std::min(an, bn);
    f[ARM_SCHOOL] = mac_ok(big, small) ? 1.f : 0.f;
    f[ARM_KARATSUBA] = small >= 2 && mac_ok(big, small) ? 1.f : 0.f;
    const bool t_ok = std::max(ntt_scale(an, bn), MIN_MID_BITS) <= MAX_SCALE;
    f[ARM_NTT] = t_ok ? 1.f : 0.f;
    f[ARM_NTT_SPLIT] = t_ok ? 1.f : 0.f;
    for (int arm : {ARM_NTT_CHUNK, ARM_NTT_CHUNK1}) {
        const int sp = chunk_scale(arm, small);
        const u64 pieces = sp <= MAX_SCALE ? chunk_pieces(big, small, sp) : 0;
        f[arm] = pieces >= 2 && pieces <= CHUNK_MAX_PIECES ? 1.f : 0.f;
    }
}

void arm_call(int arm, u64 an, u64 bn, int* algo, int* param) {
    const u64 small = std::min(an, bn);
    *param = 0;
    switch (arm) {
        case ARM_SCHOOL: *algo = ALGO_SCHOOL; break;
        case ARM_KARATSUBA: *algo = ALGO_KARATSUBA; break;
        case ARM_NTT: *algo = ALGO_NTT; break;
        case ARM_NTT_SPLIT: *algo = ALGO_NTT_SPLIT; break;
        default:
            *algo = ALGO_NTT_CHUNK;
            *param = chunk_scale(arm, small);
            break;
    }
}

// ---- policy table ---------------------------------------------------------

void PolicyTable::request_features(u64 an, u64 bn, u64 count, float x[N_IN]) {
    const double big = (double)std::max(an, bn), small = (double)std::min(an, bn);
    x[0] = (float)(std::log2(64.0 * big) / 30.0);
    x[1] = (float)(std::log2(64.0 * small) / 30.0);
    x[2] = (float)(std::log2((double)std::max<u64>(count, 1)) / 8.0);
}

void PolicyTable::features(std::vector<float>& x) {
    x.assign((size_t)N_POINTS * N_IN, 0.f);
    size_t k = 0;
    for (int c = 0; c < NCOUNT; c++)
        for (int i = 0; i < NPT; i++)
            for (int j = 0; j <= i; j++) {
                x[k++] = (float)((6.0 + LO + (double)i / Q) / 30.0);
                x[k++] = (float)((6.0 + LO + (double)j / Q) / 30.0);
                x[k++] = (float)(3.0 * c / 8.0);
            }
}

void PolicyTable::set(const float* y, u64 n_out_per_point) {
    pred.assign((size_t)N_POINTS * N_ARMS, 0.f);
    for (size_t p = 0; p < (size_t)N_POINTS; p++)
        for (int a = 0; a < N_ARMS; a++) pred[p * N_ARMS + a] = y[p * n_out_per_point + a];
    on = true;
}

int PolicyTable::choose(u64 an, u64 bn, u64 count) const {
    if (!on || !an || !bn) return -1;
    const u64 big = std::max(an, bn), small = std::min(an, bn);
    if (small < (1ull << LO) || big > (1ull << HI)) return -1;
    auto idx = [](u64 w) {
        long q = std::lround((std::log2((double)w) - LO) * Q);
        return (int)std::clamp<long>(q, 0, NPT - 1);
    };
    const int i = idx(big), j = std::min(idx(small), i);
    const int c = (int)std::clamp<long>(std::lround(std::log2((double)std::max<u64>(count, 1)) / 3.0), 0, NCOUNT - 1);
    const float* row = &pred[((size_t)c * NTRI + (size_t)i * (i + 1) / 2 + j) * N_ARMS];
    float f[N_ARMS];
    arm_feasible(an, bn, f);
    int best = -1;
    for (int a = 0; a < N_ARMS; a++)
        if (f[a] > 0.f && (best < 0 || row[a] < row[best])) best = a;
    return best;
}

// ---- batch plans ----------------------------------------------------------

namespace {

// How one request runs.
struct Route {
    int s = 0;        // transform scale, 0 = schoolbook
    int sp = 0;       // chunked: piece scale (pieces >= 2), else 0
    const u64* big = nullptr;
    u64 big_n = 0;
    const u64* small = nullptr;
    u64 small_n = 0;
    u64 piece = 0;    // piece length
    u64 pieces = 0;
    u64 tmp_off = 0;  // workspace offset of the odd pieces' buffer
};

// One job before sorting: a whole request, or piece k of a chunked one.
struct Item {
    u64 req;
    int s;
    u64 k;      // piece index (chunked), 0 otherwise
    bool piece;
};

// Every buffer carved from the batch workspace starts on a 64-word (512-byte)
// boundary.  The transform kernels move rows of 16-32 consecutive words
// (128-256 bytes): packed after the odd-sized schoolbook and chunk buffers,
// a bucket's residues could start mid cache line, and every row segment then
// touched one line more (the 2^19-word bucket's strided passes ran ~2x
// slower per job than the same bucket planned alone).
constexpr u64 WS_ALIGN_WORDS = 64;
u64 ws_align(u64 w) { return (w + WS_ALIGN_WORDS - 1) & ~(WS_ALIGN_WORDS - 1); }

u64* wsptr(u64 off) { return reinterpret_cast<u64*>(static_cast<uintptr_t>(off)); }
u64 wsoff(const void* p) { return static_cast<u64>(reinterpret_cast<uintptr_t>(p)); }

}  // namespace

int plan_batch(const IReq* reqs, u64 n, const PolicyTable* pol, u64 count, BatchPlan& P, std::string& err,
               int forced_arm) {
    P = BatchPlan{};
    std::vector<Route> route(n);
    P.arms.assign(n, -1);
    u64 ws = 0;
    for (u64 i = 0; i < n; i++) {
        const IReq& r = reqs[i];
        if (!r.dst || !r.a || !r.b || !r.an || !r.bn || !r.n_out || r.n_out > r.an + r.bn + 1 ||
            (r.n_add && !r.add) || r.n_add > r.n_out) {
            err = "batch request " + std::to_string(i) + ": null pointer or bad size";
            return 1;
        }
        const bool sq = r.a == r.b && r.an == r.bn;
        Route& R = route[i];
        const u64 mn = std::min(r.an, r.bn);
        int arm = -1;
        const bool eligible = !sq && !r.n_add && r.n_out == r.an + r.bn;
        if (eligible && forced_arm >= 0 && forced_arm < N_ARMS) arm = forced_arm;
        else if (eligible && pol && pol->on) arm = pol->choose(r.an, r.bn, count);
        P.arms[i] = arm;
        if (arm < 0) {
            R.s = mn < NTT_MUL_THRESHOLD ? 0 : ntt_scale(r.an, r.bn);
        } else if (arm == ARM_SCHOOL || arm == ARM_KARATSUBA) {
            R.s = 0;
        } else if (arm == ARM_NTT) {
            R.s = std::max(ntt_scale(r.an, r.bn), MIN_MID_BITS);
        } else {
            // the reference split is the chunked form at the split's scale
            const u64 big = std::max(r.an, r.bn), small = mn;
            int sp = arm == ARM_NTT_SPLIT ? std::max(ntt_scale((big + 1) >> 1, small), MIN_MID_BITS)
                                          : chunk_scale(arm, small);
            const u64 pieces = sp <= MAX_SCALE ? chunk_pieces(big, small, sp) : 0;
            // pieces of one parity must not overlap: more than two pieces need
            // a piece at least as long as the shorter operand
            const bool disjoint = pieces == 2 || (1ull << sp) >= 2 * small - 1;
            if (pieces >= 2 && pieces <= CHUNK_MAX_PIECES && disjoint) {
                R.s = sp;
                R.sp = sp;
                const bool a_big = r.an >= r.bn;
                R.big = a_big ? r.a : r.b;
                R.big_n = big;
                R.small = a_big ? r.b : r.a;
                R.small_n = small;
                R.piece = (1ull << sp) - small + 1;
                R.pieces = pieces;
                R.tmp_off = ws;
                ws = ws_align(ws + r.an + r.bn - R.piece);
            } else {
                R.s = std::max(ntt_scale(r.an, r.bn), MIN_MID_BITS);
            }
        }
        if (R.s > MAX_SCALE) {
            err = "batch request too large";
            return 1;
        }
        P.max_s = std::max(P.max_s, R.s);
    }
    std::vector<Item> items;
    items.reserve(n);
    for (u64 i = 0; i < n; i++) {
        if (route[i].sp)
            for (u64 k = 0; k < route[i].pieces; k++) items.push_back({i, route[i].s, k, true});
        else
            items.push_back({i, route[i].s, 0, false});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.s < y.s; });
    u64 tiles = 0, cols = 0;
    // B operands repeated within a bucket (same pointer and size) are
    // transformed once: the first such job owns the transform, later ones read it
    std::map<std::pair<const u64*, u64>, u32> owner_of;
    for (const Item& it : items) {
        const IReq& r = reqs[it.req];
        const Route& R = route[it.req];
        BJob J{};
        bool out_ws = false;
        if (it.piece) {
            const u64 off = it.k * R.piece, nn = r.an + r.bn;
            J.a = R.big + off;
            J.an = std::min(R.piece, R.big_n - off);
            J.b = R.small;
            J.bn = R.small_n;
            // even pieces write windows of dst, odd ones of the buffer holding
            // dst[P ..): each window covers its piece's product and the gap
            // to the next piece of the same parity, and the last window of
            // each parity runs to the end of the product
            const bool last_of_parity = it.k + 2 >= R.pieces;
            J.n_out = last_of_parity ? nn - off : 2 * R.piece;
            if (it.k % 2 == 0) {
                J.out = r.dst + off;
            } else {
                J.out = wsptr(R.tmp_off + (off - R.piece));
                out_ws = true;
            }
            P.n_pieces++;
        } else {
            J.a = r.a;
            J.b = r.b;
            J.an = r.an;
            J.bn = r.bn;
            J.out = r.dst;
            J.n_out = r.n_out;
            J.add = r.add;
            J.n_add = r.n_add;
        }
        const bool sq = J.a == J.b && J.an == J.bn;
        J.tile0 = tiles;
        tiles += crt_tiles(J.n_out);
        if (it.s == 0) {
            J.school = 1;
            J.n_res = J.an + J.bn - 1;
            J.col0 = cols;
            cols += J.n_res;
            J.fa = wsptr(ws);
            ws = ws_align(ws + 3 * J.n_res);
        } else {
            const u64 L = 1ull << it.s;
            J.sqr = sq;
            J.n_res = std::min<u64>(J.an + J.bn - 1, L);
            J.fa = wsptr(ws);
            for (int p = 0; p < 3; p++) J.r[p] = wsptr(ws + (u64)p * L);
            ws += 3 * L;
            if (P.buckets.empty() || P.buckets.back().s != it.s) {
                P.buckets.push_back({it.s, (u32)P.jobs.size(), 0, 0, 0});
                owner_of.clear();
            }
            P.buckets.back().count++;
            auto key = std::make_pair(J.b, J.bn);
            auto own = sq ? owner_of.end() : owner_of.find(key);
            if (own != owner_of.end()) {
                BJob& O = P.jobs[own->second];
                O.share = 1;
                J.share = 2;
                J.fb = O.fb;
            } else {
                J.fb = wsptr(ws);
                ws += 3 * L;
                if (!sq) owner_of[key] = (u32)P.jobs.size();
            }
        }
        if (it.s == 0) P.n_school++;
        P.jobs.push_back(J);
        P.out_in_ws.push_back(out_ws);
    }
    // the batched kernels launch one work-group per (job, tile) item of a bucket
    for (const Bucket& bk : P.buckets)
        if ((u64)bk.count * ((1ull << bk.s) >> MIN_MID_BITS) >= (1ull << 31)) {
            err = "batch bucket too large";
            return 1;
        }
    // forward-only entries: transform each shared B once into its owner's fb
    P.n_req_jobs = (u32)P.jobs.size();
    for (Bucket& bk : P.buckets) {
        bk.own_first = (u32)P.jobs.size();
        for (u32 j = bk.first; j < bk.first + bk.count; j++) {
            const BJob& O = P.jobs[j];
            if (O.share != 1) continue;
            BJob F{};
            F.a = O.b;
            F.an = O.bn;
            F.b = O.b;
            F.bn = O.bn;
            F.fa = O.fb;
            F.fb = O.fb;
            F.n_res = O.n_res;
            P.jobs.push_back(F);
            P.out_in_ws.push_back(false);
        }
        bk.own_count = (u32)P.jobs.size() - bk.own_first;
    }
    // chunked requests: dst[P .. n) += the odd pieces' buffer
    for (u64 i = 0; i < n; i++) {
        const Route& R = route[i];
        if (!R.sp) continue;
        const u64 m = reqs[i].an + reqs[i].bn - R.piece;
        AddJob A{};
        A.out = reqs[i].dst + R.piece;
        A.a = reqs[i].dst + R.piece;
        A.b = wsptr(R.tmp_off);
        A.n = A.na = A.nb = m;
        A.tile0 = P.add_tiles;
        P.add_tiles += (m + LIMB_TILE - 1) / LIMB_TILE;
        P.adds.push_back(A);
    }
    P.ncols = cols;
    P.ntiles = tiles;
    P.ws_words = ws;
    return 0;
}

void rebase_plan(BatchPlan& P, u64* ws) {
    for (u32 j = 0; j < P.jobs.size(); j++) {
        BJob& J = P.jobs[j];
        J.fa = ws + wsoff(J.fa);
        if (P.out_in_ws[j]) J.out = ws + wsoff(J.out);
        if (J.school) {
            for (int i = 0; i < 3; i++) J.r[i] = J.fa + (u64)i * J.n_res;
        } else {
            J.fb = ws + wsoff(J.fb);
            if (j < P.n_req_jobs)
                for (int i = 0; i < 3; i++) J.r[i] = ws + wsoff(J.r[i]);
            if (J.sqr) J.fb = J.fa;
        }
    }
    for (AddJob& A : P.adds) A.b = ws + wsoff(A.b);
}

}  // namespace nnh
