// Shared device/host helpers for the gfx950 MSA-search kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/sseqs_amd.h"

namespace sseqs {

// Last error message for the calling thread (sseqs_last_error()).
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Checks hipGetLastError() after a launch; returns 0 or the hip error code.
int check_launch(const char* what);

// Residue order used by every table: ARNDCQEGHILKMFPSTWYV.
// BLOSUM62 (sseqs/sw.cu:7-29), host copy; kernels receive it by value.
extern const int8_t h_blosum62[400];
struct Blosum {
    int8_t v[400];
};
inline Blosum blosum62() {
    Blosum b;
    for (int i = 0; i < 400; ++i) b.v[i] = h_blosum62[i];
    return b;
}

// ASCII -> 0..19 residue index, 20 otherwise (the CUDA SW kernels score any
// non-residue as -4, sseqs/sw.cu:128-130).
__host__ __device__ inline int aa_index(uint32_t c) {
    switch (c) {
    case 'A': return 0;  case 'R': return 1;  case 'N': return 2;  case 'D': return 3;
    case 'C': return 4;  case 'Q': return 5;  case 'E': return 6;  case 'G': return 7;
    case 'H': return 8;  case 'I': return 9;  case 'L': return 10; case 'K': return 11;
    case 'M': return 12; case 'F': return 13; case 'P': return 14; case 'S': return 15;
    case 'T': return 16; case 'W': return 17; case 'Y': return 18; case 'V': return 19;
    default: return 20;
    }
}

// Prefilter database code (sseqs/fft.cu:5-15): residues 0..19, X - * -> 20,
// any other byte -> 21.
__host__ __device__ inline int corr_code(uint32_t c) {
    int a = aa_index(c);
    if (a < 20) return a;
    if (c == 'X' || c == '-' || c == '*') return 20;
    return 21;
}

// One 5-bit symbol (as ASCII) at symbol index q of an MSB-first packed
// stream; reads bytes q*5/8 and q*5/8+1 (database buffers carry 16 bytes of
// tail padding so the second byte always exists).
__device__ inline uint32_t packed_sym(const uint8_t* __restrict__ p, int64_t q) {
    int64_t bit = q * 5;
    int64_t b = bit >> 3;
    uint32_t w = ((uint32_t)p[b] << 8) | (uint32_t)p[b + 1];
    return ((w >> (11 - (uint32_t)(bit & 7))) & 31u) + 64u;
}

__device__ inline int wave_max_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Blocks for a grid-stride launch over `items` units of `per_block` each.
// An HSA dispatch counts work-items in 32 bits, so kernels that may cover a
// whole database loop instead of launching one thread per unit; 65536
// blocks fill the 256 CUs many times over.
inline unsigned grid_cap(int64_t items, int64_t per_block) {
    int64_t b = ceil_div(items > 0 ? items : 1, per_block);
    return (unsigned)(b < 65536 ? b : 65536);
}

}  // namespace sseqs
