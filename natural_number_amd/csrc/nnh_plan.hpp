// nnh_plan.hpp — host-only planning of products (no HIP): the exact
// algorithm arms and when each applies, the learned cost model's policy
// table, and batch plans (buckets, job table, workspace offsets).  Kept free
// of GPU types so that the planner builds and runs on a CPU under address
// and undefined-behaviour sanitizers.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "nnh_jobs.hpp"

namespace nnh {

constexpr u64 NTT_MUL_THRESHOLD = 384;  // reference NTTmul_threshold (ntt_info.h:6)
constexpr int MIN_MID_BITS = 8;         // smallest fused-chunk transform (2^8 words)
constexpr int MAX_SCALE = 30;           // largest transform handled (2^30 words per prime)

// get_NTT_scale (natural.h:384-386): the smallest s with 2^s >= s1 + s2 - 1.
inline int ntt_scale(u64 s1, u64 s2) {
    const u64 x = s1 + s2 - 2;
    return x ? 64 - __builtin_clzll(x) : 0;
}

// ---- level split of one transform ------------------------------------------
// A transform of 2^s words runs s - 1 butterfly levels: the fused chunk
// kernel holds chunks of 2^cb words (its cb - 1 levels of each operand, the
// base multiply and the inverse levels), and np strided top passes, in
// forward order the highest levels first, run lev[k] levels each over tiles
// of 4096 words.  The inverse runs the passes in reverse.
constexpr int TILE_BITS = 12;       // top-pass tile and largest fused chunk: 4096 words
constexpr int MIN_CHUNK_BITS = 10;  // smallest chunk the formula gives beside top passes (one wave)
constexpr int MAX_PASSES = 4;
struct LevelPlan {
    int cb = 0;  // 0: no plan (the formula applies)
    int np = 0;
    int lev[MAX_PASSES] = {0, 0, 0, 0};
};
// The hand-derived split: up to 2^20 words the largest chunk (2^12) and
// balanced passes; above, as many levels in the (HBM-bound) top passes as
// the fewest passes allow, the (VALU-bound) chunk keeping at least 2^10 words.
LevelPlan formula_plan(int s);
// Whether plan p is one the library runs at scale s: cb = s up to 2^12
// words; above, 10 <= cb <= 12 (work-groups of at least one wave), at most
// four passes of 1..8 levels summing to s - cb, and every pass's tile inside
// one stride block (lowest level + depth >= 12).
bool plan_valid(int s, const LevelPlan& p);
// The autotuner's candidates at scale s (all valid, the formula's first):
// chunk sizes 2^10..2^12 with the fewest passes and one more, the passes
// balanced with the deeper or the shallower first, or the deepest first.
std::vector<LevelPlan> plan_candidates(int s);

// Algorithm selectors of nnh_mul_algo (include/nnh.h).
enum Algo {
    ALGO_AUTO = 0,       // the library's dispatch: the installed policy, else ALGO_REFERENCE
    ALGO_SCHOOL = 1,
    ALGO_KARATSUBA = 2,
    ALGO_NTT = 3,
    ALGO_NTT_SPLIT = 4,
    ALGO_NTT_CHUNK = 5,
    ALGO_REFERENCE = 6,  // the reference's fixed thresholds (natural.h:444-450)
    ALGO_MODEL = 7,      // the installed cost-model policy (an error when none is installed)
};

// Candidate arms of the cost model, in the order of its outputs.
enum Arm { ARM_SCHOOL = 0, ARM_KARATSUBA, ARM_NTT, ARM_NTT_SPLIT, ARM_NTT_CHUNK, ARM_NTT_CHUNK1, N_ARMS };

// Caps that keep hopeless candidates out of the search (every arm is exact).
constexpr u64 SCHOOL_MAX_MAC = 1ull << 28;  // schoolbook / Karatsuba word-products
constexpr u64 CHUNK_MAX_PIECES = 4096;      // chunked transform: pieces of one product

// Piece scale of the chunked arms for a shorter operand of `small` words:
// the smallest transform holding it twice (at least 2^8), and one more.
int chunk_scale(int arm, u64 small);
// Pieces of the longer operand at piece scale sp (piece length 2^sp - small + 1).
u64 chunk_pieces(u64 big, u64 small, int sp);
// f[arm] = 1 when arm is a candidate for an an x bn product, else 0.
void arm_feasible(u64 an, u64 bn, float f[N_ARMS]);
// (algo, param) of nnh_mul_algo that runs `arm`.
void arm_call(int arm, u64 an, u64 bn, int* algo, int* param);

// The cost model compiled into a table: predicted log2 milliseconds of
// every arm at grid points (larger, smaller operand words) spaced 1/Q
// octave over [2^LO, 2^HI] words (the model's training range, 2^10 .. 2^30
// bits) and at the counts it was trained on: 1 (a product on its own) and
// 8, 64 (products inside one batch of that many).  The
// network itself runs on the device when the table is built (one forward
// launch over every point); a request is then a table lookup and an
// arg-min over the arms feasible at its exact sizes.
struct PolicyTable {
    static constexpr int N_IN = 3;
    static constexpr int Q = 16;
    static constexpr int LO = 4;
    static constexpr int HI = 24;
    static constexpr int NPT = (HI - LO) * Q + 1;          // points per axis
    static constexpr int NTRI = NPT * (NPT + 1) / 2;       // (i >= j) pairs
    static constexpr int NCOUNT = 3;                       // counts 1, 8, 64
    static constexpr int N_POINTS = NCOUNT * NTRI;
    bool on = false;
    std::vector<float> pred;  // [N_POINTS][N_ARMS]

    // Model inputs of every table point, N_IN floats each, in table order.
    static void features(std::vector<float>& x);
    // Inputs of one request (log2 bits of the larger and smaller operand
    // over 30, log2 count over 8).
    static void request_features(u64 an, u64 bn, u64 count, float x[N_IN]);
    void set(const float* y, u64 n_out_per_point);
    // Arm for an an x bn product alone (count 1) or in a batch of `count`; -1 when no
    // table is installed or the sizes lie outside it (the reference's
    // thresholds then apply).
    int choose(u64 an, u64 bn, u64 count) const;
};

// A bucket of transform jobs of one scale.
struct Bucket {
    int s;
    u32 first, count;
    u32 own_first, own_count;  // forward-only entries of the bucket's shared B transforms
};

// A batch plan.  Pointers in `jobs` and `adds` that point into the batch
// workspace hold word offsets until rebase() (the *_in_ws flags say which).
struct BatchPlan {
    std::vector<BJob> jobs;  // the requests' jobs (schoolbook first, then by scale), then forward-only entries
    u32 n_req_jobs = 0;
    std::vector<Bucket> buckets;
    u32 n_school = 0;
    u64 ncols = 0;
    u64 ntiles = 0;
    u64 ws_words = 0;
    int max_s = 0;
    // requests split into pieces (chunked arm): odd pieces land in a
    // workspace buffer that one addition per request adds into dst
    std::vector<AddJob> adds;
    u64 add_tiles = 0;
    u64 n_pieces = 0;  // piece jobs made from chunked requests
    std::vector<u8> out_in_ws;  // per job: out is a workspace offset
    // arm chosen per request (-1: the reference's rule)
    std::vector<int> arms;
};

// Plan n requests.  With a policy (non-null, installed), requests without
// an addend whose result is an + bn words run the arm it picks for a batch
// of `count`: schoolbook (also for Karatsuba), one transform, or
// the chunked transform (the reference split is the chunked form at the
// split's scale).  Otherwise schoolbook below 384 words and one transform
// of get_NTT_scale words above.  Returns 0 or an nnh_status with `err`.
// forced_arm >= 0 runs every eligible request with that arm instead (timing
// of the arms in batch context).
int plan_batch(const IReq* reqs, u64 n, const PolicyTable* pol, u64 count, BatchPlan& P, std::string& err,
               int forced_arm = -1);
// Replace workspace offsets by addresses in ws.
void rebase_plan(BatchPlan& P, u64* ws);

}  // namespace nnh
