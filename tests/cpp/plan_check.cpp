// plan_check.cpp — TEST INFRASTRUCTURE.  The host-only batch planner
// (natural_number_amd/csrc/nnh_plan.cpp) built for the CPU, under the
// address and undefined-behaviour sanitizers by tests/test_plan.py.
//
//   plan_check SIZES_FILE   plans the requests listed in the file (one
//                           "an bn" per line) three ways -- the reference's
//                           rule, a policy that prefers the chunked arms,
//                           and one that prefers the reference split -- and
//                           checks every invariant of each plan;
//   plan_check emulate      plans small random requests under those
//                           policies and executes each plan on the CPU the
//                           way the device does (every job's product into
//                           its output window, zeros past the product, then
//                           the additions), comparing every result with a
//                           schoolbook product.
// Prints "plan_check ok" on success; exits 1 on the first violation.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "nnh_plan.hpp"

using namespace nnh;

static void die(const std::string& m) {
    std::printf("FAIL %s\n", m.c_str());
    std::exit(1);
}
#define REQUIRE(c, m)       \
    do {                    \
        if (!(c)) die(m);   \
    } while (0)

static uint64_t sm_next(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Policy tables with fixed preferences (predicted log2 times per arm).
static PolicyTable make_policy(const float pref[N_ARMS]) {
    PolicyTable t;
    std::vector<float> y((size_t)PolicyTable::N_POINTS * N_ARMS);
    for (size_t p = 0; p < (size_t)PolicyTable::N_POINTS; p++)
        for (int a = 0; a < N_ARMS; a++) y[p * N_ARMS + a] = pref[a];
    t.set(y.data(), N_ARMS);
    return t;
}

struct Req {
    uint64_t an, bn;
};

// Offsets of workspace pointers (before rebase the planner stores word offsets in them).
static uint64_t off(const void* p) { return (uint64_t)(uintptr_t)p; }

static void check_plan(const std::vector<IReq>& R, const BatchPlan& P, const char* what) {
    const std::string w = what;
    const uint64_t n = R.size();
    REQUIRE(P.out_in_ws.size() == P.jobs.size(), w + ": flag per job");
    REQUIRE(P.n_req_jobs <= P.jobs.size(), w + ": request jobs");
    // schoolbook jobs first, then buckets in increasing scale covering the rest
    for (uint32_t j = 0; j < P.n_school; j++) REQUIRE(P.jobs[j].school == 1, w + ": schoolbook jobs first");
    uint32_t next = P.n_school;
    int prev_s = 0;
    for (const Bucket& b : P.buckets) {
        REQUIRE(b.first == next && b.count > 0, w + ": buckets contiguous");
        REQUIRE(b.s > prev_s && b.s <= MAX_SCALE, w + ": bucket scales increase");
        prev_s = b.s;
        next += b.count;
        const uint64_t L = 1ull << b.s;
        std::map<std::pair<const uint64_t*, uint64_t>, uint32_t> owners;
        for (uint32_t j = b.first; j < b.first + b.count; j++) {
            const BJob& J = P.jobs[j];
            REQUIRE(!J.school, w + ": no schoolbook job in a bucket");
            REQUIRE(J.an + J.bn - 1 <= L, w + ": transform holds the product");
            REQUIRE(J.n_res == J.an + J.bn - 1, w + ": residue count");
            if (J.share == 1) owners[{J.b, J.bn}] = j;
        }
        for (uint32_t j = b.first; j < b.first + b.count; j++) {
            const BJob& J = P.jobs[j];
            if (J.share != 2) continue;
            auto it = owners.find({J.b, J.bn});
            REQUIRE(it != owners.end() && it->second < j, w + ": reader has an earlier owner in its bucket");
            REQUIRE(P.jobs[it->second].fb == J.fb, w + ": reader reads its owner's transform");
        }
        // forward-only entries: one per owner, writing the owner's fb
        uint32_t n_own = 0;
        for (uint32_t j = b.first; j < b.first + b.count; j++) n_own += P.jobs[j].share == 1;
        REQUIRE(b.own_count == n_own, w + ": one forward-only entry per owner");
        for (uint32_t k = b.own_first; k < b.own_first + b.own_count; k++)
            REQUIRE(k >= P.n_req_jobs && k < P.jobs.size(), w + ": forward-only entries after the requests");
    }
    REQUIRE(next == P.n_req_jobs, w + ": buckets cover the transform jobs");
    // CRT tiles contiguous
    uint64_t t = 0;
    for (uint32_t j = 0; j < P.n_req_jobs; j++) {
        REQUIRE(P.jobs[j].tile0 == t, w + ": tiles contiguous");
        t += crt_tiles(P.jobs[j].n_out);
    }
    REQUIRE(t == P.ntiles, w + ": tile count");
    // workspace intervals pairwise disjoint and inside ws_words
    std::vector<std::pair<uint64_t, uint64_t>> iv;
    std::map<uint64_t, uint64_t> fb_seen;
    for (uint32_t j = 0; j < P.n_req_jobs; j++) {
        const BJob& J = P.jobs[j];
        if (J.school) {
            iv.push_back({off(J.fa), off(J.fa) + 3 * J.n_res});
            continue;
        }
        uint64_t Lj = 0;
        for (const Bucket& b : P.buckets)
            if (j >= b.first && j < b.first + b.count) Lj = 1ull << b.s;
        iv.push_back({off(J.fa), off(J.fa) + 3 * Lj});
        if (!J.sqr && J.share != 2 && !fb_seen.count(off(J.fb))) {
            fb_seen[off(J.fb)] = 1;
            iv.push_back({off(J.fb), off(J.fb) + 3 * Lj});
        }
    }
    // every residue and transform buffer starts on a 64-word boundary (the
    // transform kernels' row segments then never straddle an extra line),
    // and so does every chunked request's odd-piece buffer
    for (uint32_t j = 0; j < P.jobs.size(); j++) {
        const BJob& J = P.jobs[j];
        REQUIRE(off(J.fa) % 64 == 0, w + ": residues aligned");
        if (!J.school) REQUIRE(off(J.fb) % 64 == 0, w + ": transform aligned");
    }
    for (const AddJob& A : P.adds) REQUIRE(off(A.b) % 64 == 0, w + ": odd-piece buffer aligned");
    // odd-piece buffers
    for (uint32_t j = 0; j < P.n_req_jobs; j++)
        if (P.out_in_ws[j]) REQUIRE(off(P.jobs[j].out) + P.jobs[j].n_out <= P.ws_words, w + ": odd window in ws");
    for (const AddJob& A : P.adds) iv.push_back({off(A.b), off(A.b) + A.nb});
    std::sort(iv.begin(), iv.end());
    for (size_t k = 0; k < iv.size(); k++) {
        REQUIRE(iv[k].second <= P.ws_words, w + ": workspace interval inside ws");
        if (k) REQUIRE(iv[k - 1].second <= iv[k].first, w + ": workspace intervals disjoint");
    }
    // every request's output covered: whole jobs, or even windows tiling [0, n)
    // of dst and odd windows tiling [P, n) of the request's buffer
    std::map<const uint64_t*, uint64_t> req_of_dst;
    for (uint64_t i = 0; i < n; i++) req_of_dst[R[i].dst] = i;
    uint64_t n_add = 0, pieces = 0;
    for (const AddJob& A : P.adds) {
        // A.out = dst + P: find the request whose dst range holds it
        auto it = req_of_dst.upper_bound(A.out);
        REQUIRE(it != req_of_dst.begin(), w + ": addition inside a request");
        --it;
        const IReq& r = R[it->second];
        const uint64_t piece = (uint64_t)(A.out - r.dst), nn = r.an + r.bn;
        REQUIRE(A.a == A.out && A.n == nn - piece && A.na == A.n && A.nb == A.n, w + ": addition covers dst[P, n)");
        // the request's piece jobs
        std::vector<std::pair<uint64_t, uint64_t>> even, odd;  // (start, n_out) in product coordinates
        const uint64_t big = std::max(r.an, r.bn), small = std::min(r.an, r.bn);
        const uint64_t* bigp = r.an >= r.bn ? r.a : r.b;
        for (uint32_t j = 0; j < P.n_req_jobs; j++) {
            const BJob& J = P.jobs[j];
            if (P.out_in_ws[j]) {
                if (off(J.out) < off(A.b) || off(J.out) >= off(A.b) + A.nb) continue;
                const uint64_t start = off(J.out) - off(A.b) + piece;
                REQUIRE((uint64_t)(J.a - bigp) == start, w + ": odd piece reads its part of the longer operand");
                odd.push_back({start, J.n_out});
            } else {
                if (J.out < r.dst || J.out >= r.dst + nn || J.a < bigp || J.a >= bigp + big) continue;
                const uint64_t start = (uint64_t)(J.out - r.dst);
                REQUIRE((uint64_t)(J.a - bigp) == start,
                        w + ": even piece reads its part of the longer operand (job " + std::to_string(j) + ", an " +
                            std::to_string(r.an) + ", bn " + std::to_string(r.bn) + ", start " + std::to_string(start) +
                            ", a offset " + std::to_string((uint64_t)(J.a - bigp)) + ")");
                even.push_back({start, J.n_out});
            }
            REQUIRE(J.an <= piece && J.bn == small, w + ": piece sizes");
            REQUIRE(J.n_out >= J.an + J.bn, w + ": window holds the piece product");
            pieces++;
        }
        std::sort(even.begin(), even.end());
        std::sort(odd.begin(), odd.end());
        uint64_t e = 0;
        for (auto& [s0, len] : even) {
            REQUIRE(s0 == e, w + ": even windows tile dst");
            e += len;
        }
        REQUIRE(e == nn, w + ": even windows reach the end");
        uint64_t o = piece;
        for (auto& [s0, len] : odd) {
            REQUIRE(s0 == o, w + ": odd windows tile [P, n)");
            o += len;
        }
        REQUIRE(odd.empty() || o == nn, w + ": odd windows reach the end");
        REQUIRE(even.size() >= 1 && odd.size() >= 1, w + ": at least two pieces");
        // pieces two apart must not overlap: P >= the shorter operand unless there are only two
        REQUIRE(even.size() + odd.size() == 2 || piece >= small, w + ": same-parity pieces disjoint");
        n_add++;
    }
    REQUIRE(pieces == P.n_pieces, w + ": every piece job belongs to an addition");
    REQUIRE(n_add == P.adds.size(), w + ": additions");
}

static BatchPlan plan_or_die(const std::vector<IReq>& R, const PolicyTable* pol, const char* what) {
    BatchPlan P;
    std::string err;
    if (plan_batch(R.data(), R.size(), pol, std::min<uint64_t>(R.size(), 64), P, err))
        die(std::string(what) + ": plan_batch: " + err);
    return P;
}

static const float PREF_CHUNK[N_ARMS] = {9, 9, 5, 6, 1, 2};
static const float PREF_SPLIT[N_ARMS] = {9, 9, 5, 1, 6, 7};
static const float PREF_SCHOOL[N_ARMS] = {1, 2, 5, 6, 7, 8};

static int run_sizes(const char* file) {
    std::vector<Req> rs;
    FILE* f = std::fopen(file, "r");
    REQUIRE(f, "open sizes file");
    unsigned long long an, bn;
    while (std::fscanf(f, "%llu %llu", &an, &bn) == 2) rs.push_back({an, bn});
    std::fclose(f);
    REQUIRE(!rs.empty(), "sizes file empty");
    // requests 9k are squares; requests i+1..i+4 for i = 3, 100, 197, ...
    // pass request i's B operand (as calc_e and the decimal levels share one)
    const size_t nr = rs.size();
    std::vector<long> share_from(nr, -1);
    for (size_t i = 3; i + 5 < nr; i += 97)
        for (size_t k = 1; k <= 4; k++)
            if ((i + k) % 9 != 0) share_from[i + k] = (long)i;
    std::vector<uint64_t> bn_of(nr);
    for (size_t i = 0; i < nr; i++) {
        bn_of[i] = i % 9 == 0 ? rs[i].an : rs[i].bn;
        if (share_from[i] >= 0) bn_of[i] = rs[share_from[i]].bn;
    }
    uint64_t total = 0;
    for (size_t i = 0; i < nr; i++) total += rs[i].an + rs[i].bn + rs[i].an + bn_of[i];
    // operands and outputs point into one host allocation (never touched)
    uint64_t* mem = (uint64_t*)std::malloc(total * sizeof(uint64_t));
    REQUIRE(mem, "allocation");
    std::vector<IReq> R(nr);
    uint64_t o = 0;
    for (size_t i = 0; i < nr; i++) {
        IReq& q = R[i];
        q.a = mem + o;
        q.an = rs[i].an;
        q.b = mem + o + q.an;
        q.bn = rs[i].bn;
        o += rs[i].an + rs[i].bn;
        if (i % 9 == 0) {
            q.b = q.a;
            q.bn = q.an;
        }
    }
    for (size_t i = 0; i < nr; i++)
        if (share_from[i] >= 0) {
            R[i].b = R[share_from[i]].b;
            R[i].bn = R[share_from[i]].bn;
        }
    uint64_t n_dst = 0;
    for (size_t i = 0; i < nr; i++) n_dst += R[i].an + R[i].bn;
    uint64_t* dmem = (uint64_t*)std::malloc(n_dst * sizeof(uint64_t));
    REQUIRE(dmem && o <= total, "layout");
    o = 0;
    for (size_t i = 0; i < nr; i++) {
        R[i].dst = dmem + o;
        R[i].n_out = R[i].an + R[i].bn;
        o += R[i].n_out;
    }
    BatchPlan P0 = plan_or_die(R, nullptr, "reference rule");
    check_plan(R, P0, "reference rule");
    PolicyTable pc = make_policy(PREF_CHUNK), ps = make_policy(PREF_SPLIT), pk = make_policy(PREF_SCHOOL);
    BatchPlan P1 = plan_or_die(R, &pc, "chunk policy");
    check_plan(R, P1, "chunk policy");
    BatchPlan P2 = plan_or_die(R, &ps, "split policy");
    check_plan(R, P2, "split policy");
    BatchPlan P3 = plan_or_die(R, &pk, "schoolbook policy");
    check_plan(R, P3, "schoolbook policy");
    std::printf("requests %zu: reference %zu jobs; chunk policy %zu jobs (%llu pieces, %zu additions); split policy "
                "%zu jobs (%llu pieces); schoolbook policy %u schoolbook jobs\n",
                R.size(), (size_t)P0.n_req_jobs, (size_t)P1.n_req_jobs, (unsigned long long)P1.n_pieces,
                P1.adds.size(), (size_t)P2.n_req_jobs, (unsigned long long)P2.n_pieces, P3.n_school);
    REQUIRE(P1.n_pieces > 0 && P2.n_pieces > 0, "the policies produced chunked requests");
    // a bad request is refused
    std::vector<IReq> bad = {R[1]};
    bad[0].n_out = bad[0].an + bad[0].bn + 2;
    BatchPlan Pb;
    std::string err;
    REQUIRE(plan_batch(bad.data(), 1, nullptr, 1, Pb, err) != 0 && !err.empty(), "bad size refused");
    std::free(mem);
    std::free(dmem);
    std::printf("plan_check ok\n");
    return 0;
}

// ---- exact CPU execution of plans of small requests ----------------------

using u128 = unsigned __int128;

static void school_mul(uint64_t* out, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn) {
    std::memset(out, 0, (an + bn) * 8);
    for (uint64_t i = 0; i < an; i++) {
        uint64_t c = 0;
        for (uint64_t j = 0; j < bn; j++) {
            u128 t = (u128)a[i] * b[j] + out[i + j] + c;
            out[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
        out[i + bn] = c;
    }
}

static int run_emulate() {
    uint64_t seed = 0x4E4E504C414E;
    const PolicyTable pc = make_policy(PREF_CHUNK), ps = make_policy(PREF_SPLIT);
    for (int round = 0; round < 4; round++) {
        std::vector<std::vector<uint64_t>> A, B, Dst;
        std::vector<IReq> R;
        const int nreq = 60;
        for (int i = 0; i < nreq; i++) {
            uint64_t an = 1 + sm_next(seed) % 4000, bn = 16 + sm_next(seed) % 120;
            if (i % 3 == 0) std::swap(an, bn);
            A.emplace_back(an);
            B.emplace_back(bn);
            for (auto& w : A.back()) w = sm_next(seed);
            for (auto& w : B.back()) w = sm_next(seed);
            if (i % 11 == 5) std::fill(A.back().begin(), A.back().end(), ~0ull);  // all-ones: long carries
            if (i % 11 == 5) std::fill(B.back().begin(), B.back().end(), ~0ull);
            Dst.emplace_back(an + bn, 0xdeadbeefull);
        }
        for (int i = 0; i < nreq; i++) {
            IReq q{};
            q.dst = Dst[i].data();
            q.a = A[i].data();
            q.an = A[i].size();
            q.b = B[i].data();
            q.bn = B[i].size();
            q.n_out = q.an + q.bn;
            R.push_back(q);
        }
        // every few requests share the B operand of the previous one
        for (int i = 1; i < nreq; i += 4) {
            R[i].b = R[i - 1].b;
            R[i].bn = R[i - 1].bn;
            R[i].n_out = R[i].an + R[i].bn;
            Dst[i].assign(R[i].n_out, 0xdeadbeefull);
            R[i].dst = Dst[i].data();
        }
        const PolicyTable* pol = round % 2 ? &ps : &pc;
        BatchPlan P = plan_or_die(R, pol, "emulated plan");
        check_plan(R, P, "emulated plan");
        REQUIRE(P.n_pieces > 0, "emulated plan has pieces");
        std::vector<uint64_t> ws(P.ws_words + 1, 0xfeedfaceull);
        rebase_plan(P, ws.data());
        // the device's semantics: each request job writes a*b into out[0, n_out)
        for (uint32_t j = 0; j < P.n_req_jobs; j++) {
            const BJob& J = P.jobs[j];
            std::vector<uint64_t> prod(J.an + J.bn);
            school_mul(prod.data(), J.a, J.an, J.b, J.bn);
            for (uint64_t k = 0; k < J.n_out; k++) J.out[k] = k < prod.size() ? prod[k] : 0;
        }
        for (const AddJob& Ad : P.adds) {
            uint64_t c = 0;
            for (uint64_t k = 0; k < Ad.n; k++) {
                u128 t = (u128)(k < Ad.na ? Ad.a[k] : 0) + (k < Ad.nb ? Ad.b[k] : 0) + c;
                Ad.out[k] = (uint64_t)t;
                c = (uint64_t)(t >> 64);
            }
            REQUIRE(c == 0, "addition carries out of the product");
        }
        for (int i = 0; i < nreq; i++) {
            std::vector<uint64_t> want(R[i].an + R[i].bn);
            school_mul(want.data(), R[i].a, R[i].an, R[i].b, R[i].bn);
            REQUIRE(std::memcmp(want.data(), Dst[i].data(), want.size() * 8) == 0,
                    "emulated product " + std::to_string(i) + " (" + std::to_string(R[i].an) + " x " +
                        std::to_string(R[i].bn) + ") arm " + std::to_string(P.arms[i]));
        }
        std::printf("round %d: %d requests, %llu pieces, %zu additions, exact\n", round, nreq,
                    (unsigned long long)P.n_pieces, P.adds.size());
    }
    std::printf("plan_check ok\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: plan_check SIZES_FILE | emulate\n");
        return 2;
    }
    if (std::string(argv[1]) == "emulate") return run_emulate();
    return run_sizes(argv[1]);
}
