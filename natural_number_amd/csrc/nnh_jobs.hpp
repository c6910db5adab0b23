// nnh_jobs.hpp — plain job descriptors shared by the kernels, the host
// driver and the host-only planner (no HIP types: the planner and its CPU
// tests include this without a GPU toolchain).
#pragma once
#include <cstdint>

namespace nnh {

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

// Words per CRT / carry-lookahead tile (16 rounds of 256 positions).
constexpr u64 CRT_TILE_WORDS = 4096;
inline u64 crt_tiles(u64 n_out) { return (n_out + CRT_TILE_WORDS - 1) / CRT_TILE_WORDS; }
// A single product's CRT: outputs shorter than 2^22 words use tiles of 512
// words (two rounds), so that enough work-groups run side by side (a
// 2^17-word output is 32 tiles of 4096 words, each 16 rounds in sequence);
// batches keep 4096-word tiles.
constexpr u64 CRT_SHORT_TILE_WORDS = 512;
constexpr u64 CRT_SHORT_BELOW = 1ull << 22;
inline u64 crt_tile_words_single(u64 n_out) { return n_out < CRT_SHORT_BELOW ? CRT_SHORT_TILE_WORDS : CRT_TILE_WORDS; }
inline u64 crt_tiles_single(u64 n_out) {
    const u64 w = crt_tile_words_single(n_out);
    return (n_out + w - 1) / w;
}

// One product of a batch (device-resident table).
struct BJob {
    const u64* a;
    const u64* b;
    u64 an, bn;
    u64* fa;          // transform workspace 3L (or school columns 3 n_res)
    u64* fb;          // second transform 3L
    const u64* r[3];  // CRT inputs: fa + i L, or the three column arrays
    u64 n_res;        // positions with residues / columns
    u64* out;
    u64 n_out;
    u64 tile0;        // first global CRT tile of this job
    u64 col0;         // first global school column (school jobs)
    const u64* add;   // optional addend (n_add words) summed into out
    u64 n_add;
    u32 school;       // 1: schoolbook columns
    u32 sqr;          // 1: b == a (single transform)
    u32 share;        // B transform shared within the batch: 0 no, 1 owner, 2 reader of the owner's fb
};

// One modular addition of a batch: out[0..n) = a + b (+ carry-in) or
// a - b (a + ~b + 1), with a zero-extended beyond na and b beyond nb; the
// carry out of word n-1 is dropped (stored to *cout when non-null: for a
// subtraction 1 means no borrow).  out may alias a or b.
struct AddJob {
    u64* out;
    const u64* a;
    const u64* b;
    u64 n, na, nb;
    u64 tile0;            // first 4096-word tile of this job (set before upload)
    const u32* cin_ptr;   // device carry-in (add only), overrides cin when non-null
    u32* cout;
    u32 sub;
    u32 cin;
};

constexpr u64 LIMB_TILE = 4096;

// One request of a batch: dst[0 .. n_out) = a * b + add (the value must fit
// n_out words; n_out <= an + bn + 1).
struct IReq {
    u64* dst;
    const u64* a;
    u64 an;
    const u64* b;
    u64 bn;
    u64 n_out;
    const u64* add;
    u64 n_add;
};

}  // namespace nnh
