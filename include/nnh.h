/*
 * nnh.h — C ABI of the MI355X three-prime NTT multiply (libnnh.so).
 *
 * Limbs are little-endian uint64_t words, least significant first, exactly
 * the layout of the reference's basic_natural (basic_natural.h:11).  Every
 * entry point returns an nnh_status; nothing throws across the boundary.
 * Buffers are caller-owned; the library keeps per-device twiddle tables and
 * a growing workspace.  Calls on one device are serialised by an internal
 * lock; each call completes on the given stream before it returns unless the
 * name ends in _async.
 *
 * The entry points replace the reference's multiply choice points:
 *   natural::_mul / _mul_NTT     natural.h:391-450   -> nnh_mul
 *   natural::_sqr / _sqr_NTT     natural.h:433-442, 529-535 -> nnh_sqr
 *   NTT_info::load+NTT           ntt_info.h:147-198  -> nnh_xf_forward (device), nnh_transform_forward (host)
 *   mul_save/mul_load/sqr_save   natural.h:452-527   -> nnh_xf_mul, nnh_xf_mul2, nnh_xf_sqr (device),
 *                                                       nnh_mul_transformed (host)
 *   NTT_info::save (cyclic mode) ntt_info.h:237-248  -> nnh_mul_cyclic
 * and, with no counterpart in the reference (which multiplies one pair at
 * a time), nnh_mul_batch for many independent products.
 */
#ifndef NNH_H
#define NNH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nnh_status {
    NNH_OK = 0,
    NNH_ERR_ARG = 1,       /* null pointer, zero length, size out of range */
    NNH_ERR_NOMEM = 2,     /* device allocation failed */
    NNH_ERR_HIP = 3,       /* a HIP runtime call failed */
    NNH_ERR_NODEV = 4,     /* no usable device / nnh_init not called */
    NNH_ERR_DOMAIN = 5     /* input outside the reference algorithm's domain */
} nnh_status;

/* Select and initialise a device (builds the root-base constants).
 * Reference: the static tables of NTT_info (ntt_info.h:69-145). */
int nnh_init(int device);

/* Release every device buffer held by the library, and the pinned host
 * staging ring and copy threads of the host-pointer calls.  Batches that
 * are still alive keep the twiddle tables they read until
 * nnh_batch_destroy. */
int nnh_shutdown(void);

/* Test hook (no counterpart in the reference): the next n compute calls
 * (products, batches, transforms, natural-layer calls) return NNH_ERR_HIP
 * before anything is queued, so that a caller's fallback to its CPU path can
 * be exercised; n = 0 cancels.  Counts across threads.  Returns how many
 * injected failures were still pending before this call. */
int nnh_debug_fail_next(int n);

/* dst[0 .. an+bn) = a * b for host buffers.  dst is not normalised: the
 * caller trims top zero words (natural::std, natural.h:82-93).  dst may
 * overlap a or b (a *= a, natural.h:357-362): both operands are staged in
 * device memory before any result word is written back.
 * Reference: natural::_mul (natural.h:444-450). */
int nnh_mul(uint64_t* dst, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn);

/* dst[0 .. 2an) = a^2; dst may overlap a.  Reference: natural::_sqr
 * (natural.h:529-535). */
int nnh_sqr(uint64_t* dst, const uint64_t* a, uint64_t an);

/* Explicit algorithm arms (no counterpart in the reference, whose choice is
 * fixed by thresholds, natural.h:444-450 and base.h:18,37).  Every arm is
 * exact; `param` is the piece transform scale for NNH_ALGO_NTT_CHUNK
 * (0 = smallest that holds the shorter operand twice).  Used by the
 * autotuner and the cost model. */
#define NNH_ALGO_AUTO 0        /* the library's dispatch: the installed cost-model policy (nnh_policy_*),
                                  else the reference's (schoolbook below 384 words, else NTT with one split) */
#define NNH_ALGO_SCHOOL 1      /* column-wise schoolbook (asmMul1) */
#define NNH_ALGO_KARATSUBA 2   /* one Karatsuba level over schoolbook columns (toom22mul) */
#define NNH_ALGO_NTT 3         /* one transform for the whole product */
#define NNH_ALGO_NTT_SPLIT 4   /* the reference's one-level unbalanced split (_mul_NTT) */
#define NNH_ALGO_NTT_CHUNK 5   /* larger operand in pieces, all pieces one batch sharing the shorter
                                  operand's transform (planned on the host, queued without waiting) */
#define NNH_ALGO_REFERENCE 6   /* the reference's fixed thresholds, whatever policy is installed */
#define NNH_ALGO_MODEL 7       /* the installed cost-model policy (NNH_ERR_ARG if none); param = count of
                                  back-to-back products the choice assumes (0 = 1) */
int nnh_mul_algo(uint64_t* dst, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn, int algo, int param);
int nnh_mul_algo_device_async(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn,
                              int algo, int param, void* stream);

/* Autotuner timing: average device milliseconds per product of `count`
 * back-to-back products through one arm, over `reps` repetitions after one
 * warm-up, on the library stream (device pointers). */
int nnh_time_algo(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn, int algo,
                  int param, int count, int reps, float* ms_per_product);
/* Autotuner timing in batch context: device milliseconds per product of a
 * batch of `count` an x bn products (operands at different offsets of the
 * a_cap / b_cap-word buffers, outputs consecutive in the dst_cap-word
 * buffer, count * (an + bn) <= dst_cap) planned with every request on cost-
 * model arm `arm` (0..5, see nnh_policy_set) and replayed as a hipGraph
 * `reps` times after one warm-up. */
int nnh_time_batch(uint64_t* d_dst, uint64_t dst_cap, const uint64_t* d_a, uint64_t a_cap, uint64_t an,
                   const uint64_t* d_b, uint64_t b_cap, uint64_t bn, int arm, int count, int reps,
                   float* ms_per_product);

/* Learned cost model (no counterpart in the reference): an FP32 MLP
 * [n_in -> 64 -> 64 -> n_out] with ReLU, trained on the device by
 * hand-written kernels (forward, backward, Adam) on masked mean-squared
 * error; outputs are predicted log2 times of the candidate arms.  Parameter
 * layout (flat float32): W1[64][n_in], b1[64], W2[64][64], b2[64],
 * W3[n_out][64], b3[n_out]. */
typedef struct nnh_mlp nnh_mlp;
int nnh_mlp_create(int n_in, int n_out, uint64_t seed, nnh_mlp** out);
int nnh_mlp_destroy(nnh_mlp* m);
int nnh_mlp_num_params(nnh_mlp* m);
int nnh_mlp_set_params(nnh_mlp* m, const float* params, uint64_t n);
int nnh_mlp_get_params(nnh_mlp* m, float* params, uint64_t n);
int nnh_mlp_forward(nnh_mlp* m, const float* x, uint64_t n, float* y);
/* Gradient of the masked MSE (mean over unmasked targets) w.r.t. every
 * parameter, and the loss. */
int nnh_mlp_gradients(nnh_mlp* m, const float* x, const float* target, const float* mask, uint64_t n, float* grad,
                      float* loss);
/* `steps` full-batch Adam steps with learning rate lr; final loss returned. */
int nnh_mlp_train(nnh_mlp* m, const float* x, const float* target, const float* mask, uint64_t n, int steps,
                  float lr, float* final_loss);
/* Index of the smallest predicted output among those with feasible > 0. */
int nnh_mlp_choose(nnh_mlp* m, const float* x, const float* feasible, uint64_t n, int* choice);

/* The cost model as the library's dispatch policy (replaces the fixed
 * thresholds of natural::_mul, natural.h:444-450, ntt_info.h:6-8, base.h:18,37;
 * no counterpart in the reference).  Arms, in the network's output order:
 * 0 schoolbook, 1 Karatsuba, 2 one transform, 3 the reference split,
 * 4 chunked at the smallest piece scale holding the shorter operand twice,
 * 5 chunked one scale up.  nnh_policy_set runs the network [3 -> 64 -> 64 -> 6]
 * on the device over a table of operand sizes (1/16 octave, 2^10 .. 2^30 bits
 * each, 1/8/64 back-to-back products) and installs the predictions: from then
 * on nnh_mul, nnh_mul_device_async, the natural-layer calls and batches pick
 * per request the arm with the smallest predicted time among those feasible
 * at its exact sizes; sizes outside the table keep the reference's rule.
 * Every arm is exact, so results do not change, only time.  NULL removes the
 * policy.  nnh_policy_load reads a file ("nnh-policy 1 3 6 <n>" then n
 * parameters); at nnh_init the library installs $NNH_POLICY, or the
 * policy_gfx950.txt beside it when present (NNH_POLICY=reference: none). */
int nnh_policy_set(nnh_mlp* m);
int nnh_policy_load(const char* path);
int nnh_policy_active(void);
/* The arm the dispatch takes for an an x bn product among `count`
 * back-to-back ones: arm = -1 (and algo NNH_ALGO_REFERENCE) when no policy
 * applies.  algo/param are those of nnh_mul_algo. */
int nnh_policy_arm(uint64_t an, uint64_t bn, uint64_t count, int* arm, int* algo, int* param);
/* feasible[0..5] = 1 for the arms that are candidates at these sizes. */
int nnh_arm_feasible(uint64_t an, uint64_t bn, float* feasible);
/* (algo, param) of nnh_mul_algo that runs arm `arm` at these sizes. */
int nnh_arm_call(int arm, uint64_t an, uint64_t bn, int* algo, int* param);

/* Device-pointer variants: operands and result already in HBM, work is
 * queued on `stream` (a hipStream_t, NULL = the library's stream) and the
 * call returns without synchronising (when the dispatch, the installed
 * cost-model policy, takes the chunked arm, the host plans its pieces as a
 * batch before the call returns; that arm refuses a stream under graph
 * capture with NNH_ERR_ARG).  d_dst must not overlap
 * d_a or d_b (the split path writes its first half before loading the
 * second).  Calls on different streams are ordered by the library (each
 * waits for the previous call's work in the shared workspace). */
int nnh_mul_device_async(uint64_t* d_dst, const uint64_t* d_a, uint64_t an,
                         const uint64_t* d_b, uint64_t bn, void* stream);
int nnh_sqr_device_async(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, void* stream);

/* Allocation-free form of nnh_mul_device_async.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same product, same reference dispatch (schoolbook below 384
 * words, else the transform with the one-level split); it calls no
 * hipMalloc/hipFree.  Its workspace and twiddle tables are a fixed array in
 * device memory (84 MiB per GPU, held from nnh_init until the process
 * ends), which limits the product to an + bn <= 2^20 words (NNH_ERR_ARG
 * beyond).  Every call on a GPU shares that array: the library orders calls
 * it sees one after another like its other calls, but replays of captured
 * calls that overlap on one GPU race on it.  d_dst must not overlap d_a or
 * d_b. */
int nnh_mul_device_noalloc(uint64_t* d_dst, const uint64_t* d_a, uint64_t an, const uint64_t* d_b, uint64_t bn,
                           void* stream);

/* Cyclic product (a*b) mod (2^(64*2^scale) - 1) as the reference's
 * NTT_info::save computes it when the product is longer than the transform
 * (ntt_info.h:238-242): dst receives 2^scale words.  Requires
 * an, bn <= 2^scale.  Host buffers. */
int nnh_mul_cyclic(uint64_t* dst, const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn, int scale);

/* Forward transform of a at `scale` (NTT_info::load + NTT_info::NTT,
 * ntt_info.h:147-198) into out[3 * 2^scale] (prime-major).  The words are
 * the reference's transform multiplied by alpha_scale (see nnh_alpha),
 * lazily reduced: in [0, 4p) for p0 and p1, [0, 2p) for p2.  Host buffers. */
int nnh_transform_forward(uint64_t* out, const uint64_t* a, uint64_t an, int scale);

/* The constant alpha_scale mod p_i (i = 0..2) by which device transforms
 * differ from the reference's. */
int nnh_alpha(int scale, uint64_t out[3]);

/* dst = a * b where fa is a forward transform of a at `scale` produced by
 * nnh_transform_forward (mul_load_NTT_info, natural.h:504-515).  Requires
 * an + bn - 1 <= 2^scale.  Host buffers. */
int nnh_mul_transformed(uint64_t* dst, const uint64_t* fa, uint64_t an, const uint64_t* b, uint64_t bn, int scale);

/* Transform handles: a forward transform kept in device memory, the
 * counterpart of the NTT_info that the reference's mul_save_NTT_info /
 * sqr_save_NTT_info keep and mul_load_NTT_info reuses (natural.h:452-527).
 * A handle holds 3 * 2^scale words (8 <= scale <= 30); the caller keeps it
 * as long as it likes (nnh_shutdown does not free it).  Device pointers,
 * work queued on `stream` (NULL = the library's stream).
 *   nnh_xf_forward  NTT_info(a, scale) + NTT(): keep a's transform (an <= 2^scale)
 *   nnh_xf_mul      d = a * b with a's kept transform (mul_load_NTT_info; with
 *                   nnh_xf_forward first, mul_save_NTT_info), bn <= 2^scale
 *   nnh_xf_mul2     d = a * b from two kept transforms of one scale
 *                   (mul_load_NTT_info with ni_a and ni_b)
 *   nnh_xf_sqr      d = a^2 from a's kept transform (sqr_save_NTT_info)
 * Every product follows NTT_info::save (ntt_info.h:237-248): with
 * size = an + bn words, when size - 1 > 2^scale the result is the product
 * modulo B^(2^scale) - 1 in 2^scale words, otherwise the exact product in
 * size words; *n_out receives that count.  d_dst holds max(size, 2^scale + 2)
 * words and must not overlap an operand.  The result is not normalised. */
typedef struct nnh_xf nnh_xf;
int nnh_xf_create(int scale, nnh_xf** out);
int nnh_xf_destroy(nnh_xf* xf);
int nnh_xf_forward(nnh_xf* xf, const uint64_t* d_a, uint64_t an, void* stream);
int nnh_xf_mul(uint64_t* d_dst, uint64_t* n_out, const nnh_xf* xa, const uint64_t* d_b, uint64_t bn, void* stream);
int nnh_xf_mul2(uint64_t* d_dst, uint64_t* n_out, const nnh_xf* xa, const nnh_xf* xb, void* stream);
int nnh_xf_sqr(uint64_t* d_dst, uint64_t* n_out, const nnh_xf* xa, void* stream);

/* Batched products (no counterpart in the reference, which multiplies one
 * pair at a time; needed for binary-splitting workloads such as calc_e,
 * natural.h:841-857).  Requests carry DEVICE pointers; dst holds an+bn words.
 * A request with a == b and an == bn is a squaring.  Products are bucketed
 * by transform length; each transform stage is one launch per
 * bucket and one CRT/carry sequence serves the whole batch.  Requests that
 * repeat a B operand (same pointer and size, same transform length) share
 * one transform of it, as the reference's mul_save/mul_load reuse does
 * (natural.h:452-527); the products are the same. */
typedef struct nnh_req {
    uint64_t* dst;
    const uint64_t* a;
    uint64_t an;
    const uint64_t* b;
    uint64_t bn;
} nnh_req;

typedef struct nnh_batch nnh_batch;

/* Plan a batch: buckets, workspace, device job table.  Flags:
 * NNH_BATCH_GRAPH captures the launch sequence into a hipGraph replayed by
 * nnh_batch_run; with a cost-model policy installed each request runs the
 * arm it picks (schoolbook, one transform, or the chunked transform whose
 * pieces join the buckets and share the shorter operand's transform; the
 * reference split runs as the chunked form at its scale), unless
 * NNH_BATCH_REFERENCE asks for the reference's rule (schoolbook below 384
 * words, else one transform).  The plan keeps the request pointers:
 * rerunning recomputes from their current contents. */
#define NNH_BATCH_GRAPH 1
#define NNH_BATCH_REFERENCE 8
int nnh_batch_create(const nnh_req* reqs, uint64_t n, int flags, nnh_batch** out);
/* Queue one execution of the batch on stream (NULL = library stream). */
int nnh_batch_run(nnh_batch* b, void* stream);
int nnh_batch_destroy(nnh_batch* b);
/* One-shot: plan, run, synchronise, release. */
int nnh_mul_batch(const nnh_req* reqs, uint64_t n, void* stream);

/* Smallest operand size (words) at which nnh_mul uses the transform
 * (the reference's NTTmul_threshold, ntt_info.h:6). */
uint64_t nnh_threshold_words(void);

/* Kernel timing (no counterpart in the reference; for benchmarks): with
 * timing enabled the library records HIP events on the launch stream around
 * each kernel family; nnh_kernel_times reads the accumulated milliseconds,
 * launch count, algorithmic HBM bytes and name of family `kind`
 * (0 .. NNH_KERNEL_KINDS-1).  Enabling or disabling clears the totals. */
#define NNH_KERNEL_KINDS 7
int nnh_kernel_timing(int enable);
int nnh_kernel_times(int kind, double* total_ms, uint64_t* launches, double* bytes, const char** name);

/* Butterfly issue ceiling (no counterpart in the reference; for
 * benchmarks): butterflies per second of prime `prime`'s transform
 * butterfly (the lazy one for primes 0 and 1, the strict one for prime 2)
 * on register-resident data, `iters` iterations of 16 butterflies per
 * thread over 2^19 threads.  Transform kernels run the same butterfly
 * plus loads, stores, LDS exchanges and twiddle reads, so this rate
 * bounds them from above. */
int nnh_butterfly_peak(int prime, uint64_t iters, double* butterflies_per_s);

/* Active level split of a transform of 2^scale words for single products
 * (for benchmarks; the tuned one when installed, else the formula;
 * the counterpart of the reference's cache blocking of NTT/INTT into global
 * levels and 2^20-word chunks, ntt_info.h:187-196, 209-219): the fused
 * chunk kernel's chunk is 2^chunk_bits words (it runs chunk_bits - 1 levels
 * of each transform), and n_passes <= 4 strided top passes run
 * pass_levels[i] levels each, in forward order (the inverse runs them in
 * reverse).  Needs no device. */
int nnh_transform_plan(int scale, int* chunk_bits, int* n_passes, int* pass_levels);

/* Level splits and their autotuner (the counterpart of the reference's
 * fixed cache blocking of NTT/INTT into global levels and 2^20-word chunks,
 * ntt_info.h:179-221).  A transform of 2^scale words runs its low
 * chunk_bits levels in the fused chunk kernel and the rest in n_passes
 * strided top passes (pass_levels[k] levels each, highest first).  The
 * library keeps one table for single products and one for batch buckets
 * (batched = 1); scales without an entry use the built-in formula.
 * nnh_tune_plans times every candidate split of `scale` (chunks of 2^10 ..
 * 2^12 words beside the top passes, the fewest passes and one more, several
 * orders) on a product of two 2^(scale-1)-word operands, or on a batch of
 * such products of about 2^22 transform words: after about 100 ms of
 * warm-up runs, three rounds time every candidate in turn over `reps` runs
 * and each keeps its fastest round; the fastest candidate is installed when
 * it beats the formula's split by more than 1 %, else the scale keeps the
 * formula.  Up to cap timings are returned (res[0] is the formula's split)
 * and *n receives the number of candidates; nnh_plan_candidates lists the
 * candidates without timing them (ms = 0).  nnh_plans_save/load write and read
 * the table ("nnh-plans 1 <arch> <compute units>", one line per entry); at
 * nnh_init the library loads $NNH_PLANS, or plans_gfx950.txt beside it when
 * it was made on the same kind of device (NNH_PLANS=formula: none).
 * nnh_plan_get returns 1 when the split comes from the table, 0 for the
 * formula; nnh_plan_set with chunk_bits = 0 returns a scale to the formula. */
typedef struct nnh_plan_time {
    int chunk_bits;
    int n_passes;
    int pass_levels[4];
    float ms;
} nnh_plan_time;
int nnh_tune_plans(int scale, int batched, int reps, nnh_plan_time* res, int cap, int* n);
int nnh_plan_candidates(int scale, nnh_plan_time* res, int cap, int* n);
int nnh_plan_set(int scale, int batched, int chunk_bits, int n_passes, const int* pass_levels);
int nnh_plan_get(int scale, int batched, int* chunk_bits, int* n_passes, int* pass_levels);
int nnh_plans_save(const char* path);
int nnh_plans_load(const char* path);

/* Device plumbing (no counterpart in the reference): device memory, copies,
 * streams and events through the library's own HIP runtime, so that a
 * caller needs no second GPU runtime in its process.  Streams are created
 * non-blocking at default priority; NULL stream = the library's stream. */
int nnh_device_alloc(void** ptr, uint64_t bytes);
int nnh_device_free(void* ptr);
int nnh_copy_to_device(void* dst, const void* src, uint64_t bytes);
int nnh_copy_to_host(void* dst, const void* src, uint64_t bytes);   /* synchronises the device first */
int nnh_copy_device(void* dst, const void* src, uint64_t bytes, void* stream);
int nnh_stream_create(void** stream);
int nnh_stream_destroy(void* stream);
int nnh_synchronize(void* stream);                                   /* NULL: whole device */
int nnh_event_create(void** ev);
int nnh_event_record(void* ev, void* stream);
int nnh_event_elapsed_ms(void* start, void* stop, float* ms);
int nnh_event_destroy(void* ev);

/* Human-readable text of the last error on this thread. */
/* ---- Device-resident natural layer ------------------------------------------
 * Host buffers in and out; limbs stay in device memory in between and every
 * product, addition, shift and comparison runs on the GPU.  Outputs are
 * normalised (top word nonzero unless the value is 0). */

/* q = a / b, r = a % b: the reference natural::_div (natural.h:749-769) with
 * its Newton reciprocal and div_iterative (natural.h:566-747), including the
 * cyclic products of NTT_info::save (ntt_info.h:237-248).  Returns
 * NNH_ERR_DOMAIN for divisors outside the reference reciprocal's domain
 * (natural.h:675-677 would write ~2^64 words). */
int nnh_divmod(uint64_t* q, uint64_t q_cap, uint64_t* qn, uint64_t* r, uint64_t r_cap, uint64_t* rn,
               const uint64_t* a, uint64_t an, const uint64_t* b, uint64_t bn);
/* b^e by the reference's square-and-multiply recursion (pow, natural.h:554-564). */
int nnh_pow(uint64_t* dst, uint64_t cap, uint64_t* n, const uint64_t* b, uint64_t bn, uint64_t e);
/* Decimal digits of a (operator<< with std::dec, natural.h:1020-1039 via
 * dec_init/_print_dec/_dec_split, natural.h:904-992); out gets a
 * NUL-terminated string of *len digits. */
int nnh_to_dec(char* out, uint64_t cap, uint64_t* len, const uint64_t* a, uint64_t an);
/* n! (factorial, natural.h:795-839): Legendre exponents, for each exponent
 * bit the product of the odd primes carrying it (a device product tree; the
 * reference uses a priority queue, same product), ans = ans^2 * P_i, then the
 * power of two as a shift.  n < 2^32. */
int nnh_factorial(uint64_t* out, uint64_t cap, uint64_t* n_out, uint64_t n);
/* calc_pi(num, den, numc, 1, 2^k + 1) (natural.h:874-902), level-synchronous;
 * sizes[0..2] receive the words of num, den, numc.  k <= 20. */
int nnh_calc_pi(uint64_t* num, uint64_t* den, uint64_t* numc, uint64_t cap, int k, uint64_t* sizes);
/* The reference's pi recipe (test.cpp:22-32): calc_pi over [1, 2^k + 1),
 * sqrt_10005 to acc words (natural.h:859-872), num = 13591409 den - num,
 * den *= 426880 numc 10^digits, num *= denc, den /= num, decimal output.
 * stage_ms (NNH_PI_STAGES entries, may be NULL): calc_pi, sqrt, products,
 * division, decimal output. */
#define NNH_PI_STAGES 5
int nnh_pi_digits(char* out, uint64_t cap, uint64_t* len, int k, uint64_t digits, uint64_t acc, double* stage_ms);
/* calc_e(num, den, 0, 2^k) (natural.h:841-857), level-synchronous. */
int nnh_calc_e(uint64_t* num, uint64_t num_cap, uint64_t* num_n, uint64_t* den, uint64_t den_cap, uint64_t* den_n,
               int k);
/* The reference's e recipe (test.cpp:12-19): calc_e over [0, 2^k), times
 * 10^digits, divided by den, printed in decimal.  stage_ms (NNH_E_STAGES
 * entries, may be NULL) receives the wall time of calc_e, pow, the
 * product, the division and the decimal output. */
#define NNH_E_STAGES 5
int nnh_e_digits(char* out, uint64_t cap, uint64_t* len, int k, uint64_t digits, double* stage_ms);

const char* nnh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NNH_H */
