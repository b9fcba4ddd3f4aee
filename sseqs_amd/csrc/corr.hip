// Low-rank correlation prefilter on MFMA.
//
// The reference (sseqs/filter.cu:92-159, sseqs/fft.cu:17-415) computes, per
// database segment d and rank r < k, an N-point real FFT of
//   b_r[t] = LUT_B[r][code(d_t)]   (t < min(len, maxlen, N), else 0)
// multiplies it by the FFT of the reversed, LUT_A-weighted query
//   a_r[t] = LUT_A[r][qrev[t]]     (t < min(Lq, N), else 0)
// inverts, sums the ranks and keeps max over even m of (1/N) * (H-point
// unnormalised inverse), i.e.
//   score = max_{m even} 0.5 * sum_r sum_p a_r[(m-p) mod N] * b_r[p].
//
// Here that sum is evaluated directly as a matrix product: for a tile of 32
// segments (MFMA columns) and 32 even shifts m = 2i (MFMA rows),
//   C[i][s] = sum_{r,p} T_r[i][p] * b_r^s[p],   T_r[i][p] = 0.5 a_r[(2i-p) mod N]
// with K = (r, p) in steps of 16 on v_mfma_f32_32x32x16_f16.  T is Toeplitz,
// so one LDS row per rank, qtab[r][u] = 0.5 a_r[(-u) mod N] (extended by 8
// entries to avoid wrap handling), yields every A fragment with one 16-byte
// read.  B fragments are table lookups of the staged residue codes.  Inputs
// are fp16, accumulation fp32, so there is no fp16 overflow (the reference's
// all-fp16 evaluation saturates to inf on strong hits).
#include <cstdio>
#include <type_traits>

#include "common.h"

namespace sseqs {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CORR_PAD = 39;   // code index whose table value is 0
constexpr int CORR_NB = 40;    // table entries per rank
constexpr int CORR_MT = 8;     // m-tiles (32 shifts each) accumulated per pass

// Segments of an ASCII batch: offsets/lengths as in sseqs/msa.py:266-287.
struct CorrAscii {
    const uint8_t* flat;
    const int32_t* off;
    const int32_t* len;
    __device__ int64_t offset(int64_t b) const { return off[b]; }
    __device__ int length(int64_t b) const { return len[b]; }
    // table index: prefilter code 0..21 (sseqs/fft.cu:5-15)
    __device__ uint32_t code(int64_t o, int p) const { return (uint32_t)corr_code(flat[o + p]); }
};

// Sequences of the HBM-resident packed database.
struct CorrPacked {
    const uint8_t* packed;
    const int64_t* off;
    const int32_t* len;
    __device__ int64_t offset(int64_t b) const { return off[b]; }
    __device__ int length(int64_t b) const { return len[b]; }
    // table index: the 5-bit symbol itself (the table maps it through the code)
    __device__ uint32_t code(int64_t o, int p) const { return packed_sym(packed, o + p) - 64u; }
};

// Copies the packed bytes of one tile of database sequences into the LDS
// stage.  lo / hi: smallest and largest symbol offset in the tile; ehi: end
// (offset + residues used) of the sequence reaching furthest; span: symbols
// past its offset that a lane addresses with 8-symbol windows (the ones
// beyond its length are masked).  Returns whether the tile fits the stage
// (else its symbols are read from global memory) and the stage's first
// byte.  Only bytes holding residues in use are copied, so a tile never
// reads more than 24 bytes past the last residue of the database.
struct Stage {
    int64_t b_lo;
    bool fast;
};
__device__ inline Stage stage_tile(const uint8_t* __restrict__ packed, uint8_t* stg, int CAPB,
                                   int64_t lo, int64_t hi, int64_t ehi, int span, int lane) {
    Stage st;
    st.b_lo = ((lo * 5) >> 3) & ~(int64_t)15;
    // the last window of a lane starts at byte 5*(o+span-8)/8 and is 16 bytes
    st.fast = ((((hi + span) * 5) >> 3) + 24 - st.b_lo) <= CAPB;
    const int64_t b_end = ((ehi * 5) >> 3) + 8;
    __syncthreads();  // the previous tile no longer reads the stage
    if (st.fast) {
        const uint4* src = reinterpret_cast<const uint4*>(packed + st.b_lo);
        uint4* dst = reinterpret_cast<uint4*>(stg);
        const int64_t nchunk = (b_end - st.b_lo + 15) >> 4;
        for (int64_t c = lane; c < nchunk; c += 64) dst[c] = src[c];
    }
    __syncthreads();
    return st;
}

// 8 symbols (5-bit, minus 64) of a sequence starting at bit `bit` of the
// stage: two aligned 8-byte LDS words, byte-swapped to big-endian order.
__device__ inline void stage_syms(const uint8_t* stg, int64_t bit, uint32_t* sym) {
    const int64_t byte = bit >> 3;
    const uint64_t* s64 = reinterpret_cast<const uint64_t*>(stg + (byte & ~7));
    const uint64_t B0 = __builtin_bswap64(s64[0]);
    const uint64_t B1 = __builtin_bswap64(s64[1]);
    const int sh = (int)(byte & 7) * 8 + (int)(bit & 7);
    // (B1 >> 1) >> (63 - sh) is B1 >> (64 - sh), and 0 for sh == 0
    const uint64_t top = (B0 << sh) | ((B1 >> 1) >> (63 - sh));
#pragma unroll
    for (int x = 0; x < 8; ++x) sym[x] = (uint32_t)(top >> (59 - 5 * x)) & 31u;
}

template <class OutT>
__device__ inline void store_score(OutT* out, int64_t i, float v);
template <>
__device__ inline void store_score<float>(float* out, int64_t i, float v) { out[i] = v; }
template <>
__device__ inline void store_score<_Float16>(_Float16* out, int64_t i, float v) {
    out[i] = (_Float16)v;
}

// One wave (64 lanes) per block; the block walks tiles of 32 segments.
// Dynamic LDS: qtab k*(N+8) fp16 | btab k*40 fp16 | codes 32*(KPmax+8) bytes.
template <class Src, class OutT, bool PACKED_CODES>
__global__ __launch_bounds__(64) void k_corr(Src src, int64_t s_begin, int64_t nseq,
                                             const uint8_t* __restrict__ qrev, int Lq,
                                             const float* __restrict__ lutA,
                                             const float* __restrict__ lutB, int V, int k, int N,
                                             int maxlen, int KPmax, OutT* __restrict__ out) {
    extern __shared__ __align__(16) uint8_t smem[];
    _Float16* qtab = reinterpret_cast<_Float16*>(smem);
    const int QS = N + 8;
    _Float16* btab = qtab + k * QS;
    uint8_t* codes = reinterpret_cast<uint8_t*>(btab + k * CORR_NB);
    const int KPS = KPmax + 8;  // row stride of the code tile (bank spread)

    const int lane = threadIdx.x;
    const int Lqe = min(Lq, N);
    // query table: qtab[r][u] = 0.5 * A[r][qrev[(-u) mod N]]
    for (int idx = lane; idx < k * QS; idx += 64) {
        int r = idx / QS, u = idx - r * QS;
        int uu = u & (N - 1);
        int t = (N - uu) & (N - 1);
        float v = 0.f;
        if (t < Lqe) v = 0.5f * lutA[r * V + (int)qrev[t]];
        qtab[idx] = (_Float16)v;
    }
    // database table, flat indexing as the reference (code 21 -> next row col 0)
    for (int idx = lane; idx < k * CORR_NB; idx += 64) {
        int r = idx / CORR_NB, c = idx - r * CORR_NB;
        float v = 0.f;
        if (PACKED_CODES) {
            if (c < 32) v = lutB[r * V + corr_code((uint32_t)c + 64u)];
        } else {
            if (c < 22) v = lutB[r * V + c];
        }
        btab[idx] = (_Float16)v;
    }

    const int sl = lane & 31, h = lane >> 5;
    const int NMT = N >= 64 ? N / 64 : 1;
    for (int64_t tile = blockIdx.x; tile * 32 < nseq; tile += gridDim.x) {
        const int64_t t0 = tile * 32;
        const int64_t sid = t0 + sl;
        const bool in = sid < nseq;
        int Le = 0;
        int64_t o = 0;
        if (in) {
            o = src.offset(s_begin + sid);
            Le = max(0, min(min(src.length(s_begin + sid), maxlen), N));
        }
        int kp = Le;
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) kp = max(kp, __shfl_xor(kp, off, 64));
        kp = (kp + 15) & ~15;
        __syncthreads();  // codes of the previous tile are consumed; tables are ready
        // stage codes: lane (sl, h) fills half of row sl, 4 codes per LDS store
        {
            const int hw = kp >> 1;
            const int p0 = h * hw;
            uint32_t* row = reinterpret_cast<uint32_t*>(codes + sl * KPS + p0);
            for (int p = 0; p < hw; p += 4) {
                uint32_t w = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    int pp = p0 + p + b;
                    uint32_t c = pp < Le ? src.code(o, pp) : (uint32_t)CORR_PAD;
                    w |= c << (8 * b);
                }
                row[p >> 2] = w;
            }
        }
        __syncthreads();

        float best = -INFINITY;
        for (int mp = 0; mp < NMT; mp += CORR_MT) {
            f32x16 acc[CORR_MT];
#pragma unroll
            for (int t = 0; t < CORR_MT; ++t)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[t][e] = 0.f;
            for (int r = 0; r < k; ++r) {
                const _Float16* qr = qtab + r * QS;
                const _Float16* br = btab + r * CORR_NB;
                for (int p0 = 0; p0 < kp; p0 += 16) {
                    const int pb = p0 + 8 * h;
                    uint64_t cw = *reinterpret_cast<const uint64_t*>(codes + sl * KPS + pb);
                    half8 bfrag;
#pragma unroll
                    for (int x = 0; x < 8; ++x) bfrag[x] = br[(cw >> (8 * x)) & 0xff];
#pragma unroll
                    for (int t = 0; t < CORR_MT; ++t) {
                        if (mp + t < NMT) {
                            const int i = (mp + t) * 32 + sl;
                            const int u0 = (pb - 2 * i) & (N - 1);
                            const uint32_t* q32 = reinterpret_cast<const uint32_t*>(qr + u0);
                            union {
                                uint32_t w[4];
                                half8 v;
                            } a;
                            a.w[0] = q32[0];
                            a.w[1] = q32[1];
                            a.w[2] = q32[2];
                            a.w[3] = q32[3];
                            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.v, bfrag, acc[t], 0,
                                                                            0, 0);
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < CORR_MT; ++t) {
                if (mp + t < NMT) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) best = fmaxf(best, acc[t][e]);
                }
            }
        }
        // lanes sl and sl+32 hold the two row halves of column sl
        best = fmaxf(best, __shfl_xor(best, 32, 64));
        if (h == 0 && in) store_score<OutT>(out, sid, best);
    }
}

// ---------------------------------------------------------------------------
// Database form: fused unpack + prefilter over the packed, HBM-resident
// stream.  A tile's 32 sequences are adjacent in a length-sorted database,
// so their packed bytes form one short range that the wave copies into LDS
// with coalesced 16-byte loads; B fragments are then decoded straight from
// that copy (8 symbols = one 40-bit window, read as two aligned 8-byte LDS
// words and byte-swapped to big-endian), once per 16-column chunk and
// shared by all ranks.  A tile whose range does not fit (an index that is
// not in stream order) reads its symbols from global memory instead.
// Dynamic LDS: qtab k*(N+8) fp16 | btab k*40 fp16 | stage CAPB+64 bytes.
// ---------------------------------------------------------------------------
template <int MT>  // m-tiles (32 even shifts each) accumulated per pass; NMT % MT == 0
__global__ __launch_bounds__(64) void k_corr_db(const uint8_t* __restrict__ packed,
                                                const int64_t* __restrict__ off,
                                                const int32_t* __restrict__ len, int64_t s_begin,
                                                int64_t nseq, const uint8_t* __restrict__ qrev,
                                                int Lq, const float* __restrict__ lutA,
                                                const float* __restrict__ lutB, int V, int k,
                                                int N, int maxlen, int CAPB,
                                                float* __restrict__ out) {
    extern __shared__ __align__(16) uint8_t smem[];
    _Float16* qtab = reinterpret_cast<_Float16*>(smem);
    const int QS = N + 8;
    _Float16* btab = qtab + k * QS;
    uint8_t* stg = reinterpret_cast<uint8_t*>(btab + k * CORR_NB);

    const int lane = threadIdx.x;
    const int Lqe = min(Lq, N);
    for (int idx = lane; idx < k * QS; idx += 64) {
        int r = idx / QS, u = idx - r * QS;
        int t = (N - (u & (N - 1))) & (N - 1);
        qtab[idx] = (_Float16)(t < Lqe ? 0.5f * lutA[r * V + (int)qrev[t]] : 0.f);
    }
    for (int idx = lane; idx < k * CORR_NB; idx += 64) {
        int r = idx / CORR_NB, c = idx - r * CORR_NB;
        btab[idx] = (_Float16)(c < 32 ? lutB[r * V + corr_code((uint32_t)c + 64u)] : 0.f);
    }

    const int sl = lane & 31, h = lane >> 5;
    const int NMT = N >= 64 ? N / 64 : 1;
    for (int64_t tile = blockIdx.x; tile * 32 < nseq; tile += gridDim.x) {
        const int64_t sid = tile * 32 + sl;
        const bool in = sid < nseq;
        int64_t o = 0;
        int Le = 0;
        if (in) {
            o = off[s_begin + sid];
            Le = max(0, min(min(len[s_begin + sid], maxlen), N));
        }
        int kp = Le;
        int64_t lo = in ? o : INT64_MAX, hi = in ? o : 0, ehi = in ? o + Le : 0;
#pragma unroll
        for (int d = 16; d > 0; d >>= 1) {
            kp = max(kp, __shfl_xor(kp, d, 64));
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
            ehi = max(ehi, __shfl_xor(ehi, d, 64));
        }
        kp = (kp + 15) & ~15;
        if (!in) o = lo;  // idle lanes decode (and mask) inside the stage
        const Stage st = stage_tile(packed, stg, CAPB, lo, hi, ehi, kp, lane);
        const bool fast = st.fast;
        const int64_t relbit = 5 * o - 8 * st.b_lo;  // bit of this sequence in the stage

        float best = -INFINITY;
        for (int mp = 0; mp < NMT; mp += MT) {
            f32x16 acc[MT];
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[t][e] = 0.f;
            for (int p0 = 0; p0 < kp; p0 += 16) {
                const int pb = p0 + 8 * h;
                uint32_t sym[8];
                if (fast) {
                    stage_syms(stg, relbit + 5 * (int64_t)pb, sym);
                } else {
#pragma unroll
                    for (int x = 0; x < 8; ++x)
                        sym[x] = pb + x < Le ? packed_sym(packed, o + pb + x) - 64u : 0u;
                }
#pragma unroll
                for (int x = 0; x < 8; ++x) sym[x] = pb + x < Le ? sym[x] : (uint32_t)CORR_PAD;
                // A-fragment base for m-tile 0 of this pass; m-tile t is 64*t earlier
                const int ubase = pb - 2 * (mp * 32 + sl);
                // operands of rank r are loaded while the MFMAs of rank r-1 run
                half8 bcur, bnxt;
                half8 acur[MT], anxt[MT];
                auto load = [&](int r, half8& b, half8* a) {
                    const _Float16* br = btab + r * CORR_NB;
                    const _Float16* qr = qtab + r * QS;
#pragma unroll
                    for (int x = 0; x < 8; ++x) b[x] = br[sym[x]];
#pragma unroll
                    for (int t = 0; t < MT; ++t) {
                        const uint32_t* q32 =
                            reinterpret_cast<const uint32_t*>(qr + ((ubase - 64 * t) & (N - 1)));
                        union {
                            uint32_t w[4];
                            half8 v;
                        } u;
                        u.w[0] = q32[0];
                        u.w[1] = q32[1];
                        u.w[2] = q32[2];
                        u.w[3] = q32[3];
                        a[t] = u.v;
                    }
                };
                load(0, bcur, acur);
                for (int r = 0; r < k; ++r) {
                    if (r + 1 < k) load(r + 1, bnxt, anxt);
#pragma unroll
                    for (int t = 0; t < MT; ++t)
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(acur[t], bcur, acc[t], 0,
                                                                        0, 0);
                    bcur = bnxt;
#pragma unroll
                    for (int t = 0; t < MT; ++t) acur[t] = anxt[t];
                }
            }
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int e = 0; e < 16; ++e) best = fmaxf(best, acc[t][e]);
        }
        best = fmaxf(best, __shfl_xor(best, 32, 64));
        if (h == 0 && in) out[sid] = best;
    }
}

// Whether the A block whose lane-0 column starts at U (mod N) has a nonzero
// entry.  Over the 64 lanes the block reads qtab[U-62 .. U+15], and qtab[u]
// is nonzero only where (-u) mod N < Lq, i.e. for -u in the 78-value window
// starting at (-U-15) mod N.
__device__ inline bool band_nonzero(int U, int N, int Lq) {
    if (Lq <= 0) return false;
    if (N <= 78) return true;
    const int v0 = (-U - 15) & (N - 1);
    return v0 < Lq || v0 + 77 >= N;
}

// ---------------------------------------------------------------------------
// Rank-4 database kernel with a register ring of A fragments.
//
// For m-tile t and 16-column chunk c the A fragment depends only on
// 16c - 64t (mod N): the fragment of m-tile t at chunk c+4 is the one m-tile
// t-1 used at chunk c.  Walking the chunks of one residue class mod 4 in
// order therefore needs one new fragment per rank and step (m-tile 0's)
// instead of MT; the other MT-1 rotate through a register ring whose slot
// for m-tile t at unrolled step s is (t - s) mod MT.  B values of all four
// ranks come from one 8-byte table entry per symbol.
// Dynamic LDS: qtab 4*(N+8) fp16 | btab4 32 x 4 fp16 | stage CAPB+64 bytes.
// ---------------------------------------------------------------------------
template <int MT>
__global__ __launch_bounds__(64) void k_corr_db4(const uint8_t* __restrict__ packed,
                                                 const int64_t* __restrict__ off,
                                                 const int32_t* __restrict__ len, int64_t s_begin,
                                                 int64_t nseq, const uint8_t* __restrict__ qrev,
                                                 int Lq, const float* __restrict__ lutA,
                                                 const float* __restrict__ lutB, int V, int N,
                                                 int maxlen, int CAPB, float* __restrict__ out) {
    constexpr int K = 4;
    extern __shared__ __align__(16) uint8_t smem[];
    _Float16* qtab = reinterpret_cast<_Float16*>(smem);
    const int QS = N + 8;
    uint64_t* btab4 = reinterpret_cast<uint64_t*>(qtab + K * QS);  // [40] x 4 fp16
    uint8_t* stg = reinterpret_cast<uint8_t*>(btab4 + CORR_NB);

    const int lane = threadIdx.x;
    const int Lqe = min(Lq, N);
    for (int idx = lane; idx < K * QS; idx += 64) {
        int r = idx / QS, u = idx - r * QS;
        int t = (N - (u & (N - 1))) & (N - 1);
        qtab[idx] = (_Float16)(t < Lqe ? 0.5f * lutA[r * V + (int)qrev[t]] : 0.f);
    }
    for (int c = lane; c < CORR_NB; c += 64) {
        union {
            _Float16 h[4];
            uint64_t u;
        } e;
#pragma unroll
        for (int r = 0; r < K; ++r)
            e.h[r] = (_Float16)(c < 32 ? lutB[r * V + corr_code((uint32_t)c + 64u)] : 0.f);
        btab4[c] = e.u;
    }

    const int sl = lane & 31, h = lane >> 5;
    const int NMT = N >= 64 ? N / 64 : 1;
    for (int64_t tile = blockIdx.x; tile * 32 < nseq; tile += gridDim.x) {
        const int64_t sid = tile * 32 + sl;
        const bool in = sid < nseq;
        int64_t o = 0;
        int Le = 0;
        if (in) {
            o = off[s_begin + sid];
            Le = max(0, min(min(len[s_begin + sid], maxlen), N));
        }
        int kp = Le;
        int64_t lo = in ? o : INT64_MAX, hi = in ? o : 0, ehi = in ? o + Le : 0;
#pragma unroll
        for (int d = 16; d > 0; d >>= 1) {
            kp = max(kp, __shfl_xor(kp, d, 64));
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
            ehi = max(ehi, __shfl_xor(ehi, d, 64));
        }
        kp = (kp + 15) & ~15;
        if (!in) o = lo;
        const Stage st = stage_tile(packed, stg, CAPB, lo, hi, ehi, kp, lane);
        const bool fast = st.fast;
        const int64_t relbit = 5 * o - 8 * st.b_lo;
        const int C = kp >> 4;  // chunks holding residues

        // B fragments of the 4 ranks for chunk c (this lane: 8 columns)
        auto bfrags = [&](int c, half8* b) {
            const int pb = 16 * c + 8 * h;
            uint32_t sym[8];
            if (fast) {
                stage_syms(stg, relbit + 5 * (int64_t)pb, sym);
            } else {
#pragma unroll
                for (int x = 0; x < 8; ++x)
                    sym[x] = pb + x < Le ? packed_sym(packed, o + pb + x) - 64u : 0u;
            }
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const uint32_t ix = pb + x < Le ? sym[x] : (uint32_t)CORR_PAD;
                union {
                    uint64_t u;
                    _Float16 h4[4];
                } e;
                e.u = btab4[ix];
#pragma unroll
                for (int r = 0; r < K; ++r) b[r][x] = e.h4[r];
            }
        };
        // A fragment of rank r whose first column is u (mod N) for this lane
        auto afrag = [&](int r, int u) -> half8 {
            const uint32_t* q32 =
                reinterpret_cast<const uint32_t*>(qtab + r * QS + (u & (N - 1)));
            union {
                uint32_t w[4];
                half8 v;
            } a;
            a.w[0] = q32[0];
            a.w[1] = q32[1];
            a.w[2] = q32[2];
            a.w[3] = q32[3];
            return a.v;
        };

        float best = -INFINITY;
        for (int mp = 0; mp < NMT; mp += MT) {
            f32x16 acc[MT];
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[t][e] = 0.f;
            // lane base: m-tile (mp + t), column 8h - 2*sl of the fragment
            const int lb = 8 * h - 2 * (mp * 32 + sl);
            for (int j = 0; j < 4 && j < C; ++j) {
                half8 ring[K][MT];
                uint32_t nzmask = 0;  // bit k: ring slot k holds a fragment with a nonzero
#pragma unroll
                for (int t = 0; t < MT; ++t) {
#pragma unroll
                    for (int r = 0; r < K; ++r) ring[r][t] = afrag(r, lb + 16 * j - 64 * t);
                    nzmask |= (uint32_t)band_nonzero(16 * (j - 4 * (mp + t)), N, Lqe) << t;
                }
                for (int cg = j; cg < C; cg += 4 * MT) {
#pragma unroll
                    for (int s = 0; s < MT; ++s) {
                        const int c = cg + 4 * s;
                        if (c < C) {
                            half8 b[K];
                            bfrags(c, b);
                            if (c != j) {
                                const int slot = (MT - s) % MT;
#pragma unroll
                                for (int r = 0; r < K; ++r) ring[r][slot] = afrag(r, lb + 16 * c);
                                nzmask = (nzmask & ~(1u << slot)) |
                                         ((uint32_t)band_nonzero(16 * (c - 4 * mp), N, Lqe)
                                          << slot);
                            }
                            // blocks whose A fragment is all zero add exact zeros: skip
#pragma unroll
                            for (int r = 0; r < K; ++r)
#pragma unroll
                                for (int t = 0; t < MT; ++t) {
                                    const int slot = (t - s + MT) % MT;
                                    if (nzmask & (1u << slot))
                                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                                            ring[r][slot], b[r], acc[t], 0, 0, 0);
                                }
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int e = 0; e < 16; ++e) best = fmaxf(best, acc[t][e]);
        }
        best = fmaxf(best, __shfl_xor(best, 32, 64));
        if (h == 0 && in) out[sid] = best;
    }
}

// ---------------------------------------------------------------------------
// Rank-4 band kernel on v_mfma_f32_16x16x32_f16.
//
// Rows are 16 even shifts (m-tile I: m = 2i, i in [16I, 16I+16)), columns 16
// sequences.  A chunk c covers positions [32c, 32c+32) and takes 4 MFMAs per
// m-tile, each with K = 8 positions x 4 ranks: ranks sit next to each other
// in K, so a B fragment is two 8-byte table entries (4 ranks of 2 symbols)
// and an A fragment is 16 contiguous bytes of the rank-interleaved query
// table.  The A block of (I, c) depends only on the diagonal
// d = (I - c) mod ND, ND = N/32, and is nonzero only for a circular run of
// Rn diagonals starting at d0, because the query occupies Lq of the N
// correlation lags.  The A fragments of those diagonals stay in registers
// for the whole kernel, so each chunk is decoded once and feeds Rn*4 MFMAs.
//
// Accumulators form a ring over the R m-tiles a chunk can reach.  With
// relative m-tile index rel = I - d0 (before reduction mod ND), chunk v
// (taken in order) updates rel = v .. v+R-1 and rel = v gets its last update
// at chunk v, so after that chunk it is folded into the running max and its
// slot reused (slot = rel mod R).  rel >= ND is the circular wrap of the
// correlation: shift tiles that the tail of a long sequence reaches again.
// Their first x = C - ND + Rn - 1 counterparts (rel < x) are parked in LDS
// when they close instead of being folded, and completed at the end.
// m-tiles that no chunk reaches are exactly zero.  FULL: every diagonal is
// nonzero (R == ND, d0 == 0); all m-tiles stay live and are folded at the end.
//
// Staging is asynchronous: while a tile is computed, the next tile's packed
// bytes are copied global -> LDS (global_load_lds_dwordx4) into the second
// stage buffer (DBL) and the index entries of the one after are loaded.
//
// Dynamic LDS: btab4 40 x 8 B | qtab4 (N+8) x 8 B during set-up, then
// (not FULL) save R x 64 lanes x 16 B | 1 or 2 (DBL) stages of CAPB+64
// bytes (rounded to 16).
// ---------------------------------------------------------------------------
__device__ inline float max4(f32x4 v) { return fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3])); }

// Reductions over the 16 lanes of a DPP row (lanes 16r .. 16r+15) by
// rotations row_ror:8, 4, 2, 1; afterwards every lane of the row holds the
// result.
template <int CTRL>
__device__ inline int dpp_row(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline int64_t dpp_row64(int64_t v) {
    const int lo = dpp_row<CTRL>((int)(uint32_t)v), hi = dpp_row<CTRL>((int)(v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ inline int row_max(int v) {
    v = max(v, dpp_row<0x128>(v));
    v = max(v, dpp_row<0x124>(v));
    v = max(v, dpp_row<0x122>(v));
    v = max(v, dpp_row<0x121>(v));
    return v;
}
template <bool MAX>
__device__ inline int64_t row_ext64(int64_t v) {
    int64_t o;
    o = dpp_row64<0x128>(v); v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    o = dpp_row64<0x124>(v); v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    o = dpp_row64<0x122>(v); v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    o = dpp_row64<0x121>(v); v = MAX ? (o > v ? o : v) : (o < v ? o : v);
    return v;
}

template <int R, bool FULL, bool DBL, int G>
__global__ __launch_bounds__(64, (G == 1 && R <= 8 ? 2 : 1)) void k_corr_band(
    const uint8_t* __restrict__ packed, const int64_t* __restrict__ off,
    const int32_t* __restrict__ len, int64_t s_begin, int64_t nseq,
    const uint8_t* __restrict__ qrev, int Lq, const float* __restrict__ lutA,
    const float* __restrict__ lutB, int V, int N, int maxlen, int CAPB, int d0, int Rn,
    float* __restrict__ out) {
    constexpr int K = 4;
    constexpr int TS = 16 * G;  // sequences per tile: G groups of 16 columns
    extern __shared__ __align__(16) uint8_t smem[];
    uint64_t* btab4 = reinterpret_cast<uint64_t*>(smem);  // [40] x 4 fp16
    uint8_t* region = smem + CORR_NB * 8;
    uint64_t* qtab4 = reinterpret_cast<uint64_t*>(region);  // [N+8] x 4 fp16, set-up only
    f32x4* save = reinterpret_cast<f32x4*>(region);         // [G][R][64]
    uint8_t* stg = region + (FULL ? 0 : G * R * 64 * 16);

    const int lane = threadIdx.x;
    const int Lqe = min(Lq, N);
    // query table, ranks interleaved: qtab4[u] = 0.5 * A[r][qrev[(-u) mod N]], r = 0..3
    for (int u = lane; u < N + 8; u += 64) {
        const int t = (N - (u & (N - 1))) & (N - 1);
        union {
            _Float16 h[4];
            uint64_t w;
        } e;
#pragma unroll
        for (int r = 0; r < K; ++r)
            e.h[r] = (_Float16)(t < Lqe ? 0.5f * lutA[r * V + (int)qrev[t]] : 0.f);
        qtab4[u] = e.w;
    }
    // database table by 5-bit symbol.  Symbol 0 ('@') never occurs inside a
    // sequence, so its entry is zero and stands for masked positions.
    for (int c = lane; c < CORR_NB; c += 64) {
        union {
            _Float16 h[4];
            uint64_t u;
        } e;
#pragma unroll
        for (int r = 0; r < K; ++r)
            e.h[r] = (_Float16)(c > 0 && c < 32 ? lutB[r * V + corr_code((uint32_t)c + 64u)]
                                                : 0.f);
        btab4[c] = e.u;
    }
    __syncthreads();

    // lane (col, g): A row col, B/C column col (of every group), positions
    // 8g..8g+7 of a chunk
    const int col = lane & 15, g = lane >> 4;
    // K is (position, rank) interleaved: element x of lane (col, g) in
    // sub-chunk q of chunk c is position 32c + 8g + 2q + (x >> 2), rank x & 3.
    // A fragment of diagonal d0+j, sub-chunk q: lags u, u+1 with
    // u = 8g + 2q - 2*col - 32*(I-c) (mod N), i.e. 16 bytes of qtab4.
    half8 A[R][4];
#pragma unroll
    for (int j = 0; j < R; ++j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int u = (8 * g + 2 * q - 2 * col - 32 * (d0 + j)) & (N - 1);
            union {
                uint64_t w[2];
                half8 v;
            } a;
            a.w[0] = qtab4[u];
            a.w[1] = qtab4[u + 1];
            A[j][q] = a.v;
        }
    }
    __syncthreads();  // the region now holds save + stage

    const int ND = N >> 5;
    const int64_t ntiles = (nseq + TS - 1) / TS;
    int64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    // stage buffers: the next tile's bytes land in one (asynchronous
    // global -> LDS copies) while the current tile is computed from the other
    const int SB = (CAPB + 64 + 15) & ~15;
    const int SB2 = DBL ? SB : 0;  // offset of the second stage buffer

    // raw index entries of tile t's sequences (clamped into range, so the
    // loads are unconditional and nothing waits for them before use)
    auto meta = [&](int64_t t, int64_t* o, int* L) {
#pragma unroll
        for (int gg = 0; gg < G; ++gg) {
            const int64_t sid = min(t * TS + 16 * gg + col, nseq - 1);
            o[gg] = off[s_begin + sid];
            L[gg] = len[s_begin + sid];
        }
    };
    // A tile is staged either as one byte range (its sequences lie close
    // together: a length-sorted stream) or, when that range exceeds the
    // stage, as TS slots of scb 16-byte chunks, slot 16gg+col holding the
    // sequence of column col of group gg from its 16-byte-aligned first
    // byte (sequences scattered over the stream); scb = 0 for a range.
    struct Tile {
        int64_t o[G], b_lo, nchunk;
        int Le[G], C, scb;
        bool in[G], fast;
    };
    auto info = [&](int64_t t, const int64_t* o, const int* L) -> Tile {
        Tile T;
        int kp = 0;
        int64_t lo = INT64_MAX, hi = 0, ehi = 0;
        // this lane's sequences (one per group)
#pragma unroll
        for (int gg = 0; gg < G; ++gg) {
            T.in[gg] = t * TS + 16 * gg + col < nseq;
            T.Le[gg] = T.in[gg] ? max(0, min(min(L[gg], maxlen), N)) : 0;
            if (T.in[gg]) {
                kp = max(kp, T.Le[gg]);
                lo = min(lo, o[gg]);
                hi = max(hi, o[gg]);
                ehi = max(ehi, o[gg] + T.Le[gg]);
            }
        }
        // Extent of the tile (the 16 columns are the 16 lanes of every DPP
        // row).  In a length-sorted stream offsets and lengths rise across
        // a one-group tile, so its first and last sequences bound it;
        // otherwise reduce over the row.
        bool rising = false;
        int nin = 0;
        if (G == 1) {
            nin = (int)min((int64_t)16, nseq - t * 16);
            const int64_t o_prev = dpp_row64<0x111>(o[0]);  // row_shr:1, column col-1
            const int L_prev = dpp_row<0x111>(T.Le[0]);
            rising = __ballot(!(col == 0 || col >= nin ||
                                (o[0] > o_prev && T.Le[0] >= L_prev))) == 0;
        }
        if (rising) {
            kp = __builtin_amdgcn_readlane(T.Le[0], nin - 1);
            lo = ((int64_t)__builtin_amdgcn_readfirstlane((int)(o[0] >> 32)) << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)o[0]);
            hi = ((int64_t)__builtin_amdgcn_readlane((int)(o[0] >> 32), nin - 1) << 32) |
                 (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)o[0], nin - 1);
            ehi = hi + kp;
        } else {
            kp = row_max(kp);
            lo = row_ext64<false>(lo);
            hi = row_ext64<true>(hi);
            ehi = row_ext64<true>(ehi);
        }
        T.C = __builtin_amdgcn_readfirstlane((kp + 31) >> 5);  // chunks holding residues, <= ND
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
            T.o[gg] = T.in[gg] ? o[gg] : lo;  // idle lanes decode (and mask) inside the stage
        T.b_lo = ((lo * 5) >> 3) & ~(int64_t)15;
        // the last window of a lane starts at byte 5*(o+32C-8)/8 and is 16 bytes
        T.fast = ((((hi + 32 * T.C) * 5) >> 3) + 24 - T.b_lo) <= CAPB;
        // copy only bytes holding residues in use: at most 24 bytes past the
        // last residue of the database are read
        T.nchunk = T.fast ? ((((ehi * 5) >> 3) + 8 - T.b_lo + 15) >> 4) : 0;
        T.scb = 0;
        if (!T.fast) {
            // slots: a sequence's copied bytes are at most 15 (alignment)
            // + 5*kp/8 + 1 + 8; its windows end at most 20C+26 bytes into
            // its slot, and the last slot's must end inside the stage
            const int scb = (((5 * kp) >> 3) + 24 + 15) >> 4;
            const int used = max(TS * 16 * scb, (TS - 1) * 16 * scb + 20 * T.C + 26);
            if (used <= CAPB) {
                T.scb = scb;
                T.fast = true;
                T.nchunk = TS * scb;
            }
        }
        return T;
    };
    // One global_load_lds_dwordx4 copies 64 x 16 bytes: lane l's 16 bytes go
    // to M0 + 16 l.  Issued from asm so that the compiler does not add waits
    // for it before unrelated LDS reads; the loop waits for it explicitly.
    auto issue = [&](const Tile& T, uint8_t* buf) {
        const uint32_t lds0 =
            (uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint8_t*)buf;
        if (T.scb == 0) {
            for (int64_t c0 = 0; c0 < T.nchunk; c0 += 64)
                if (c0 + lane < T.nchunk) {
                    const uint8_t* src = packed + T.b_lo + 16 * (c0 + lane);
                    asm volatile("global_load_lds_dwordx4 %1, off"
                                 :
                                 : "{m0}"(lds0 + 16u * (uint32_t)c0), "v"(src)
                                 : "memory");
                }
            return;
        }
        // slots: chunk c is chunk j of slot s; the slot's sequence comes
        // from lane s & 15 (every row holds the same column)
        for (int c0 = 0; c0 < (int)T.nchunk; c0 += 64) {
            const int c = c0 + lane;
            const int s = c / T.scb, j = c - s * T.scb;
            int64_t os = 0;
            int ls = 0;
#pragma unroll
            for (int gg = 0; gg < G; ++gg) {
                const int64_t v = __shfl(T.o[gg], s & 15);
                const int l = __shfl(T.Le[gg], s & 15);
                if ((s >> 4) == gg) {
                    os = v;
                    ls = l;
                }
            }
            const int64_t bb = ((os * 5) >> 3) & ~(int64_t)15;
            const int64_t nb = (((os + ls) * 5) >> 3) + 8 - bb;  // bytes holding its residues
            if (c < (int)T.nchunk && ls > 0 && 16 * (int64_t)j < nb) {
                const uint8_t* src = packed + bb + 16 * j;
                asm volatile("global_load_lds_dwordx4 %1, off"
                             :
                             : "{m0}"(lds0 + 16u * (uint32_t)c0), "v"(src)
                             : "memory");
            }
        }
    };

    int64_t o_n[G], o_c[G];
    int Le_n[G], Le_c[G];
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
        o_n[gg] = 0;
        Le_n[gg] = 0;
    }
    meta(tile, o_c, Le_c);
    if (tile + gridDim.x < ntiles) meta(tile + gridDim.x, o_n, Le_n);
    Tile cur = info(tile, o_c, Le_c);
    issue(cur, stg);
    int sb = 0;
    int64_t prev_sid[G];  // this lane's results of the previous tile, stored late
    float prev_best[G];
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
        prev_sid[gg] = -1;
        prev_best[gg] = 0.f;
    }
    for (;;) {
        // the current tile's bytes (and the next tile's index entries) are in
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        const int64_t tn = tile + gridDim.x;
        Tile nxt{};
        if (tn < ntiles) {
            nxt = info(tn, o_n, Le_n);
            if (DBL) issue(nxt, stg + (sb ^ 1) * SB2);
        }
        if (tn + gridDim.x < ntiles) meta(tn + gridDim.x, o_n, Le_n);
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
            if (prev_sid[gg] >= 0) out[prev_sid[gg]] = prev_best[gg];
        const uint8_t* sbuf = stg + sb * SB2;
        const int C = cur.C;
        int64_t relbit[G];  // bit of each sequence of this lane in the stage
#pragma unroll
        for (int gg = 0; gg < G; ++gg)
            relbit[gg] = cur.scb ? 128 * (int64_t)cur.scb * (16 * gg + col) + 5 * cur.o[gg] -
                                       8 * (((cur.o[gg] * 5) >> 3) & ~(int64_t)15)
                                 : 5 * cur.o[gg] - 8 * cur.b_lo;

        // B fragments (4 sub-chunks) of chunk v of group gg: this lane's 8
        // positions 32v+8g.., symbol 0 past the sequence's end
        auto decode = [&](auto fast_tag, int gg, int v, half8* b) {
            const int pb = 32 * v + 8 * g;
            const int Le = cur.Le[gg];
            uint32_t sym[8];
            // the 40-bit window of the 8 symbols, from the stage (fast) or
            // straight from the packed stream (16 bytes at an 8-aligned
            // address, inside the stream's tail padding at its end)
            const uint8_t* wbase;
            int64_t bit;
            if constexpr (decltype(fast_tag)::value) {
                wbase = sbuf;
                bit = relbit[gg] + 5 * (int64_t)pb;
            } else {
                wbase = packed;
                bit = 5 * (cur.o[gg] + (int64_t)pb);
            }
            const int64_t byte = bit >> 3;
            const uint64_t* s64 = reinterpret_cast<const uint64_t*>(wbase + (byte & ~7));
            const uint64_t B0 = __builtin_bswap64(s64[0]);
            const uint64_t B1 = __builtin_bswap64(s64[1]);
            const int sh = (int)(byte & 7) * 8 + (int)(bit & 7);
            uint64_t top = (B0 << sh) | ((B1 >> 1) >> (63 - sh));
            // keep the first min(max(Le - pb, 0), 8) symbols of the window
            const int nv = min(max(Le - pb, 0), 8);
            top &= ~(~0ull >> (5 * nv));
#pragma unroll
            for (int e = 0; e < 8; ++e) sym[e] = (uint32_t)(top >> (59 - 5 * e)) & 31u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                union {
                    uint64_t w[2];
                    half8 v;
                } u;
                u.w[0] = btab4[sym[2 * q]];
                u.w[1] = btab4[sym[2 * q + 1]];
                b[q] = u.v;
            }
        };
        // The tile's chunks; chunk v+1 is decoded in the same basic block as
        // chunk v's MFMAs, so the two interleave.  Two instances: symbols
        // from the stage (fast) or from global memory.
        auto run = [&](auto fast_tag, float* best) {
            f32x4 acc[G][R];
#pragma unroll
            for (int gg = 0; gg < G; ++gg) {
                best[gg] = -INFINITY;
#pragma unroll
                for (int t = 0; t < R; ++t) acc[gg][t] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
            const int x = FULL ? 0 : max(0, C - ND + Rn - 1);
            // B fragments of the current and the next chunk alternate between
            // bb[0] and bb[1] (R is even, so chunk v always uses bb[v & 1])
            static_assert(R % 2 == 0, "ring size must be even");
            half8 bb[2][G][4];
            if (C > 0)
#pragma unroll
                for (int gg = 0; gg < G; ++gg) decode(fast_tag, gg, 0, bb[0][gg]);
            for (int vb = 0; vb < C; vb += R) {
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const int v = vb + s;
                    if (v < C) {
#pragma unroll
                        for (int gg = 0; gg < G; ++gg)
                            decode(fast_tag, gg, min(v + 1, C - 1), bb[(s + 1) & 1][gg]);
                        // consecutive MFMAs go to different accumulators
#pragma unroll
                        for (int q = 0; q < 4; ++q)
#pragma unroll
                            for (int j = 0; j < R; ++j)
#pragma unroll
                                for (int gg = 0; gg < G; ++gg)
                                    acc[gg][(s + j) % R] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                        A[j][q], bb[s & 1][gg][q], acc[gg][(s + j) % R], 0, 0,
                                        0);
                        if (!FULL) {
                            // relative m-tile v is complete: park it (slot R-1
                            // takes the stores of the ones that are not)
                            const bool park = v < x;
#pragma unroll
                            for (int gg = 0; gg < G; ++gg) {
                                save[((gg * R) + (park ? v : R - 1)) * 64 + lane] = acc[gg][s];
                                const float m = max4(acc[gg][s]);
                                best[gg] = park ? best[gg] : fmaxf(best[gg], m);
                                acc[gg][s] = f32x4{0.f, 0.f, 0.f, 0.f};
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int gg = 0; gg < G; ++gg) {
                if (FULL) {
#pragma unroll
                    for (int t = 0; t < R; ++t) best[gg] = fmaxf(best[gg], max4(acc[gg][t]));
                } else {
                    // live: rel = C .. C+R-2 in slots rel mod R
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        const int rel = C + ((t - C) % R + R) % R;
                        if (rel == C + R - 1) continue;
                        if (rel < ND)
                            best[gg] = fmaxf(best[gg], max4(acc[gg][t]));
                        else if (rel - ND < x)
                            best[gg] = fmaxf(best[gg],
                                             max4(acc[gg][t] +
                                                  save[(gg * R + rel - ND) * 64 + lane]));
                    }
                    if (C + R - 1 < ND) best[gg] = fmaxf(best[gg], 0.f);  // unreached m-tiles
                }
            }
        };
        float best[G];
        if (cur.fast)
            run(std::true_type{}, best);
        else
            run(std::false_type{}, best);
#pragma unroll
        for (int gg = 0; gg < G; ++gg) {
            // the four lane groups hold different rows of the same columns
            float bm = fmaxf(best[gg], __shfl_xor(best[gg], 16, 64));
            bm = fmaxf(bm, __shfl_xor(bm, 32, 64));
            prev_best[gg] = bm;
            prev_sid[gg] = (g == 0 && cur.in[gg]) ? tile * TS + 16 * gg + col : -1;
        }

        tile = tn;
        if (tile >= ntiles) break;
        if (!DBL) {
            __syncthreads();  // the stage is no longer read
            issue(nxt, stg);
        } else {
            sb ^= 1;
        }
        cur = nxt;
    }
#pragma unroll
    for (int gg = 0; gg < G; ++gg)
        if (prev_sid[gg] >= 0) out[prev_sid[gg]] = prev_best[gg];
}

// Nonzero diagonals of the band kernel: diagonal d's A blocks read lags
// (32d + e) mod N, e in [-31, 30], and the query fills lags [0, Lqe).  The
// nonzero ones form one circular run; returns its start and length.
static void band_diagonals(int N, int Lqe, int* d0, int* rn) {
    const int ND = N / 32;
    bool nz[64];
    int cnt = 0;
    for (int d = 0; d < ND; ++d) {
        nz[d] = false;
        for (int e = -31; e <= 30 && !nz[d]; ++e) nz[d] = ((32 * d + e) & (N - 1)) < Lqe;
        cnt += nz[d];
    }
    *d0 = 0;
    *rn = cnt;
    if (cnt == 0 || cnt == ND) return;
    for (int d = 0; d < ND; ++d)
        if (nz[d] && !nz[(d + ND - 1) % ND]) *d0 = d;
}

template <class Src, class OutT, bool PACKED_CODES>
static int launch_corr(const Src& src, int64_t s_begin, int64_t nseq, const uint8_t* qrev, int Lq,
                       const float* lutA, const float* lutB, int V, int k, int N, int maxlen,
                       OutT* out, hipStream_t stream) {
    if (nseq <= 0) return 0;
    if (N < 16 || N > 2048 || (N & (N - 1)))
        return fail(-1, "prefilter: N must be a power of two in 16..2048");
    if (k < 1 || k > 20) return fail(-1, "prefilter: k must be 1..20");
    if (V < 1) return fail(-1, "prefilter: vocab_size must be >= 1");
    int mlen = maxlen < 0 ? 0 : maxlen;
    int KPmax = (min(mlen, N) + 15) & ~15;
    size_t lds = (size_t)k * (N + 8) * 2 + (size_t)k * CORR_NB * 2 + (size_t)32 * (KPmax + 8);
    lds = (lds + 15) & ~(size_t)15;
    if (lds > 160 * 1024) return fail(-1, "prefilter: tables exceed LDS (reduce k or N)");
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_corr<Src, OutT, PACKED_CODES>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return fail(-3, "prefilter: cannot enable large LDS");
    int64_t tiles = ceil_div(nseq, 32);
    unsigned grid = (unsigned)(tiles < 8192 ? tiles : 8192);
    hipLaunchKernelGGL((k_corr<Src, OutT, PACKED_CODES>), dim3(grid), dim3(64), lds, stream, src,
                       s_begin, nseq, qrev, Lq, lutA, lutB, V, k, N, mlen, KPmax, out);
    return check_launch("prefilter");
}

}  // namespace sseqs

using namespace sseqs;

extern "C" int launch_warp_corr(uint16_t* corr, const uint8_t* qrev, const float* lutA,
                                const float* W_db, const uint8_t* flat, const int32_t* offsets,
                                const int32_t* lengths, float* out_stats, int B, int vocab_size,
                                int k, int query_len, int maxlen, int N_full, hipStream_t stream,
                                uint16_t* query_rfft) {
    (void)out_stats;
    (void)query_rfft;
    if (B < 0) return fail(-1, "warp_corr: negative segment count");
    if (B == 0) return 0;
    CorrAscii src{flat, offsets, lengths};
    return launch_corr<CorrAscii, _Float16, false>(src, 0, B, qrev, query_len, lutA, W_db,
                                                   vocab_size, k, N_full, maxlen,
                                                   reinterpret_cast<_Float16*>(corr), stream);
}

// Which kernel and launch shape sseqs_db_prefilter uses for a length range.
struct PrefilterPlan {
    enum Kind { BAND, DB4, DB } kind;
    int R = 0;        // band: ring size (diagonals held in registers)
    int G = 1;        // band: groups of 16 sequences per tile
    bool full = false, dbl = false;
    int d0 = 0, rn = 0;
    int nmt = 1;      // dense kernels: m-tiles per pass
    int capb = 0;     // stage capacity (bytes)
    size_t lds = 0;   // dynamic LDS (bytes)
    int tile = 32;    // sequences per tile
};

static PrefilterPlan plan_prefilter(int query_len, int k, int N, int mlen, bool allow_band = true) {
    PrefilterPlan P;
    const int Lqe = query_len < N ? query_len : N;
    if (allow_band && k == 4 && N >= 128 && Lqe > 0 && !getenv("SSEQS_CORR_NO_BAND")) {
        band_diagonals(N, Lqe, &P.d0, &P.rn);
        const int ND = N / 32;
        P.full = P.rn == ND;
        if (P.full) {
            P.R = ND <= 16 ? ND : 0;
        } else {
            for (int c : {4, 6, 8, 12, 16, 18})
                if (!P.R && P.rn <= c) P.R = c;
        }
        if (P.R) {
            P.kind = PrefilterPlan::BAND;
            // one group of 16 sequences per tile: two groups per tile at one
            // wave per SIMD measured 27 % slower than one group at two waves
            P.G = 1;
            P.tile = 16 * P.G;
            // a tile of length-sorted sequences spans about that many of them
            const int span = (min(mlen, N) + 31) & ~31;
            const int capb = (int)((((int64_t)P.tile * (span + 1) * 5) / 8 + 48 + 15) & ~15);
            // a tile of scattered sequences of up to span symbols, one
            // 16-byte-aligned slot each (the kernel's slot layout)
            const int scb = (((5 * span) >> 3) + 24 + 15) >> 4;
            const int slots = (max(P.tile * 16 * scb,
                                   (P.tile - 1) * 16 * scb + 20 * (span / 32) + 26) + 15) & ~15;
            const size_t qt = (size_t)4 * (N + 8) * 2;
            const size_t save = P.full ? 0 : (size_t)P.G * P.R * 64 * 16;
            auto lds_for = [&](int cap, int nbuf) {
                const size_t run = save + nbuf * (size_t)((cap + 64 + 15) & ~15);
                return ((size_t)CORR_NB * 8 + (qt > run ? qt : run) + 15) & ~(size_t)15;
            };
            // two stage buffers while the blocks of 8 waves per CU still fit
            const size_t budget = (size_t)(P.G == 1 ? 20 : 40) * 1024;
            P.dbl = lds_for(capb, 2) <= budget;
            if (const char* e = getenv("SSEQS_CORR_DBL")) P.dbl = atoi(e) != 0;
            const int nbuf = P.dbl ? 2 : 1;
            // room for the slot layout where that budget allows it
            P.capb = slots > capb && lds_for(slots, nbuf) <= budget ? slots : capb;
            P.lds = lds_for(P.capb, nbuf);
            return P;
        }
    }
    P.kind = k == 4 ? PrefilterPlan::DB4 : PrefilterPlan::DB;
    const int KPmax = (min(mlen, N) + 15) & ~15;
    // a length-sorted tile spans about 32 sequences of at most KPmax symbols
    P.capb = (int)((((int64_t)33 * (KPmax + 17) * 5) / 8 + 64 + 15) & ~15);
    P.lds = ((size_t)k * (N + 8) * 2 + (size_t)k * CORR_NB * 2 + (size_t)P.capb + 64 + 15) &
            ~(size_t)15;
    P.nmt = N >= 64 ? N / 64 : 1;
    // m-tiles per pass: more reuse of each decoded B fragment against more
    // registers (8 -> 1 wave per SIMD, 4 -> 2).  Measured on gfx950 with
    // rank 4: N = 512 runs fastest at 4, N = 1024 and 2048 at 8.
    if (P.nmt > 4 && N <= 512) P.nmt = 4;
    if (P.nmt > 8) P.nmt = 8;
    // SSEQS_CORR_MT caps it further (tuning knob; results do not depend on it)
    if (const char* e = getenv("SSEQS_CORR_MT")) {
        int cap = atoi(e);
        while (P.nmt > 1 && P.nmt > cap) P.nmt >>= 1;
    }
    return P;
}

static std::string plan_name(const PrefilterPlan& P) {
    char b[64];
    if (P.kind == PrefilterPlan::BAND)
        snprintf(b, sizeof b, "k_corr_band<%d, %s, %s, %d>", P.R, P.full ? "true" : "false",
                 P.dbl ? "true" : "false", P.G);
    else
        snprintf(b, sizeof b, "%s<%d>", P.kind == PrefilterPlan::DB4 ? "k_corr_db4" : "k_corr_db",
                 P.nmt);
    return b;
}

extern "C" int sseqs_db_prefilter_kernel(int query_len, int k, int N, int maxlen, char* name,
                                         int name_len) {
    if (N < 16 || N > 2048 || (N & (N - 1)))
        return fail(-1, "db_prefilter_kernel: N must be a power of two in 16..2048");
    if (k < 1 || k > 20) return fail(-1, "db_prefilter_kernel: k must be 1..20");
    if (!name || name_len < 1) return fail(-1, "db_prefilter_kernel: no output buffer");
    const std::string s = plan_name(plan_prefilter(query_len, k, N, maxlen < 0 ? 0 : maxlen));
    snprintf(name, (size_t)name_len, "%s", s.c_str());
    return 0;
}

static int db_prefilter(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                        int64_t s_begin, int64_t s_end, const uint8_t* qrev, int query_len,
                        const float* lutA, const float* lutB, int vocab_size, int k, int N,
                        int maxlen, float* scores, hipStream_t stream, bool allow_band) {
    if (s_end < s_begin) return fail(-1, "db_prefilter: s_end < s_begin");
    const int64_t nseq = s_end - s_begin;
    if (nseq == 0) return 0;
    if (N < 16 || N > 2048 || (N & (N - 1)))
        return fail(-1, "db_prefilter: N must be a power of two in 16..2048");
    if (k < 1 || k > 20) return fail(-1, "db_prefilter: k must be 1..20");
    if (vocab_size < 1) return fail(-1, "db_prefilter: vocab_size must be >= 1");
    const int mlen = maxlen < 0 ? 0 : maxlen;
    const PrefilterPlan P = plan_prefilter(query_len, k, N, mlen, allow_band);
    if (P.lds > 160 * 1024) return fail(-1, "db_prefilter: tables exceed LDS (reduce k or N)");
    const int64_t tiles = ceil_div(nseq, P.tile);
    const unsigned grid = (unsigned)(tiles < 8192 ? tiles : 8192);
    auto allow_lds = [&](const void* kern) -> bool {
        return P.lds <= 64 * 1024 ||
               hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)P.lds) == hipSuccess;
    };
    if (P.kind == PrefilterPlan::BAND) {
        auto gob = [&](auto kern) -> int {
            if (!allow_lds(reinterpret_cast<const void*>(kern)))
                return fail(-3, "db_prefilter: cannot enable large LDS");
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64), P.lds, stream, packed, seq_off, seq_len,
                               s_begin, nseq, qrev, query_len, lutA, lutB, vocab_size, N, mlen,
                               P.capb, P.d0, P.rn, scores);
            return check_launch("db_prefilter");
        };
        auto pick = [&](auto dbl_tag) -> int {
            constexpr bool D = decltype(dbl_tag)::value;
            if (P.full)
                return P.R == 4   ? gob(&k_corr_band<4, true, D, 1>)
                       : P.R == 8 ? gob(&k_corr_band<8, true, D, 1>)
                                  : gob(&k_corr_band<16, true, D, 1>);
            switch (P.R) {
            case 4: return gob(&k_corr_band<4, false, D, 1>);
            case 6: return gob(&k_corr_band<6, false, D, 1>);
            case 8: return gob(&k_corr_band<8, false, D, 1>);
            case 12: return gob(&k_corr_band<12, false, D, 1>);
            case 16: return gob(&k_corr_band<16, false, D, 1>);
            default: return gob(&k_corr_band<18, false, D, 1>);
            }
        };
        return P.dbl ? pick(std::true_type{}) : pick(std::false_type{});
    }
    if (P.kind == PrefilterPlan::DB4) {
        auto go4 = [&](auto kern) -> int {
            if (!allow_lds(reinterpret_cast<const void*>(kern)))
                return fail(-3, "db_prefilter: cannot enable large LDS");
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64), P.lds, stream, packed, seq_off, seq_len,
                               s_begin, nseq, qrev, query_len, lutA, lutB, vocab_size, N, mlen,
                               P.capb, scores);
            return check_launch("db_prefilter");
        };
        switch (P.nmt) {
        case 1: return go4(&k_corr_db4<1>);
        case 2: return go4(&k_corr_db4<2>);
        case 4: return go4(&k_corr_db4<4>);
        default: return go4(&k_corr_db4<8>);
        }
    }
    auto go = [&](auto kern) -> int {
        if (!allow_lds(reinterpret_cast<const void*>(kern)))
            return fail(-3, "db_prefilter: cannot enable large LDS");
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64), P.lds, stream, packed, seq_off, seq_len,
                           s_begin, nseq, qrev, query_len, lutA, lutB, vocab_size, k, N, mlen,
                           P.capb, scores);
        return check_launch("db_prefilter");
    };
    switch (P.nmt) {
    case 1: return go(&k_corr_db<1>);
    case 2: return go(&k_corr_db<2>);
    case 4: return go(&k_corr_db<4>);
    default: return go(&k_corr_db<8>);
    }
}

extern "C" int sseqs_db_prefilter(const uint8_t* packed, const int64_t* seq_off,
                                  const int32_t* seq_len, int64_t s_begin, int64_t s_end,
                                  const uint8_t* qrev, int query_len, const float* lutA,
                                  const float* lutB, int vocab_size, int k, int N, int maxlen,
                                  float* scores, hipStream_t stream) {
    return db_prefilter(packed, seq_off, seq_len, s_begin, s_end, qrev, query_len, lutA, lutB,
                        vocab_size, k, N, maxlen, scores, stream, true);
}

// Spans that may hold '@' inside them (the shifted segments of the per-batch
// cascade): the generic kernels only, whose table gives '@' the value of
// code 21 like the ASCII operator, while the band kernel's symbol 0 masks.
extern "C" int sseqs_db_prefilter_spans(const uint8_t* packed, const int64_t* seq_off,
                                        const int32_t* seq_len, int64_t s_begin, int64_t s_end,
                                        const uint8_t* qrev, int query_len, const float* lutA,
                                        const float* lutB, int vocab_size, int k, int N,
                                        int maxlen, float* scores, hipStream_t stream) {
    return db_prefilter(packed, seq_off, seq_len, s_begin, s_end, qrev, query_len, lutA, lutB,
                        vocab_size, k, N, maxlen, scores, stream, false);
}
