// nnh_nat.hpp — device-side limb arithmetic used by the device-resident
// natural layer (division, radix conversion, binary splitting).
#pragma once
#include "nnh_internal.hpp"

namespace nnh {

// Enqueue a batch of additions; jobs/tile scratch are device buffers of at
// least njobs AddJobs and 2*ntiles bytes (ntiles = sum ceil(n/4096)).
u64 add_tiles_needed(const AddJob* jobs, u64 njobs);
void launch_add_batch(AddJob* d_jobs, u32 njobs, u64 ntiles, u8* d_tiles, hipStream_t st);
// One addition passed by value (no job table); d_tiles holds 2*ceil(n/4096) bytes.
void launch_add_one(AddJob job, u8* d_tiles, hipStream_t st);

void launch_not(u64* out, const u64* a, u64 n, hipStream_t st);
void launch_fill(u64* out, u64 n, u64 v, hipStream_t st);
// out[k] = splitmix64(seed + (k + 1) gamma): random-looking words (timing inputs).
void launch_fill_hash(u64* out, u64 n, u64 seed, hipStream_t st);
// out[0..n] = a << b (1 <= b <= 63): n + 1 words.
void launch_shl_bits(u64* out, const u64* a, u64 n, u32 b, hipStream_t st);
// out[0..n) = a >> b (1 <= b <= 63).
void launch_shr_bits(u64* out, const u64* a, u64 n, u32 b, hipStream_t st);
// Highest index k < n with: mode 0 a[k] != 0, mode 1 a[k] != ~0,
// mode 2 a[k] != b[k].  hout (host-visible, 4 words) receives {k+1 or 0,
// a[k], b[k], seq}, seq written last; res (device word, zero on entry) is
// the search's accumulator and is cleared again for the next search.
void launch_find(const u64* a, const u64* b, u64 n, int mode, u64* res, u64* hout, u64 seq, hipStream_t st);
// hout[0..n) = p[0..n) (n <= 3), then hout[3] = seq.
void launch_peek(const u64* p, u32 n, u64* hout, u64 seq, hipStream_t st);
void launch_set_word(u64* p, u64 v, hipStream_t st);
// q[0..n) = a / d0 and *rem = a mod d0, where d = d0 << l has its top bit
// set and v = floor((2^128 - 1) / d) - 2^64; scratch holds
// div_word_scratch_words(n) words (the chunks' remainders).
u64 div_word_scratch_words(u64 n);
void launch_div_word(u64* q, const u64* a, u64 n, u64 d, u32 l, u64 v, u64* rem, u64* scratch, hipStream_t st);

// Binary splitting of sum 1/k! (reference calc_e, natural.h:841-857).
// Leaves: thread j folds the 16 pairs [32j, 32j+32) into (num, den) at
// stride words.  Requires stride >= the value's words (<= 14 for terms < 2^25).
void launch_e_leaves(u64* num, u64* den, u64 nblocks, u64 stride, hipStream_t st);
// One level: combine c folds ranges 2c and 2c+1 (stride S <= 512) into
// num2 = num_l den_r + num_r, den2 = den_l den_r (stride S2 <= 4096).
void launch_e_combine(const u64* num, const u64* den, u64 S, u64* num2, u64* den2, u64 S2, u64 ncomb,
                      hipStream_t st);

// pi by binary splitting (calc_pi, natural.h:874-902).  Pairs: entry j
// combines terms first+2j and first+2j+1 (terms < 2^21) into num, den,
// numc at stride S >= 4.  Levels: entries 2c, 2c+1 (stride S <= 256) into
// entry c (stride S2 <= 4096), num = num_l den_r + numc_l num_r.
void launch_pi_pairs(u64 first, u64 npairs, u64* num, u64* den, u64* numc, u64 S, hipStream_t st);
void launch_pi_combine(const u64* num, const u64* den, const u64* numc, u64 S, u64* num2, u64* den2, u64* numc2,
                       u64 S2, u64 ncomb, hipStream_t st);

// Product trees of one-word factors (< 2^32): leaves multiply 16 factors
// each (f holds 16 * nleaves words, padded with 1) into `stride` words
// (>= 9); a level multiplies entries 2c and 2c+1 (stride S <= 256) into
// entry c of stride S2 <= 4096.
void launch_tree_leaves(const u64* f, u64 nleaves, u64* out, u64 stride, hipStream_t st);
void launch_tree_combine(const u64* in, u64 S, u64* out, u64 S2, u64 ncomb, hipStream_t st);

// Radix conversion (reference _dec_split/_print_dec, natural.h:953-992).
// Leaves: piece j (< 10^(18*2^(lvl+1)), stride words <= 15) becomes
// 18*2^(lvl+1) zero-padded digits at out + j * that many chars.
void launch_dec_leaves(const u64* x, u64 stride, u64 count, int lvl, char* out, hipStream_t st);
// One level of splits by base = 10^(18*2^i) with w = |base| <= 256 words:
// piece j (stride sx words) -> hi at out[(2j) w], lo at out[(2j+1) w],
// using the reference's reciprocal r (rn words) and normalising shift n_shl.
// err (device word) is set nonzero when an estimate exceeds the quotient.
void launch_dec_split_wg(const u64* x, u64 sx, u64 count, const u64* base, u64 w, const u64* r, u64 rn, u32 n_shl,
                         u64* out, u64* err, hipStream_t st);
// Batched helpers for the large levels of the conversion.
// out_j[0..n_out) = (in_j >> shift bits), pieces at strides so / si with n_in words.
void launch_shr_pieces(u64* out, u64 so, const u64* in, u64 si, u64 n_in, u64 shift, u64 n_out, u64 count,
                       hipStream_t st);
// *flag = 1 when any word [lo, hi) of any piece (stride sx) is nonzero.
void launch_nonzero_pieces(const u64* x, u64 sx, u64 lo, u64 hi, u64 count, u64* flag, hipStream_t st);
// flags[j] = (x_j >= y) for pieces x_j (stride sx, nx words) and one y (ny words).
void launch_cmp_pieces(const u64* x, u64 sx, u64 nx, const u64* y, u64 ny, u64 count, u8* flags, hipStream_t st);

}  // namespace nnh
