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
#include "common.h"

namespace sseqs {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

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
        int64_t lo = in ? o : INT64_MAX, hi = in ? o : 0;
#pragma unroll
        for (int d = 16; d > 0; d >>= 1) {
            kp = max(kp, __shfl_xor(kp, d, 64));
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
        }
        kp = (kp + 15) & ~15;
        if (!in) o = lo;  // idle lanes decode (and mask) inside the stage
        // Symbols read: [lo, hi + kp).  The last window of a lane starts at
        // byte 5*(o+kp-8)/8 and spans 16 bytes, so 24 bytes past 5*(hi+kp)/8
        // cover every read; the start is 16-byte aligned for the copy.
        const int64_t b_lo = ((lo * 5) >> 3) & ~(int64_t)15;
        const int64_t b_hi = (((hi + kp) * 5) >> 3) + 24;
        const bool fast = (b_hi - b_lo) <= CAPB;
        __syncthreads();  // the previous tile no longer reads the stage
        if (fast) {
            const uint4* src = reinterpret_cast<const uint4*>(packed + b_lo);
            uint4* dst = reinterpret_cast<uint4*>(stg);
            const int nchunk = (int)((b_hi - b_lo + 15) >> 4);
            for (int c = lane; c < nchunk; c += 64) dst[c] = src[c];
        }
        __syncthreads();
        const int64_t relbit = 5 * o - 8 * b_lo;  // bit of this sequence in the stage

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
                    const int64_t bit = relbit + 5 * (int64_t)pb;
                    const int64_t byte = bit >> 3;
                    const uint64_t* s64 = reinterpret_cast<const uint64_t*>(stg + (byte & ~7));
                    const uint64_t B0 = __builtin_bswap64(s64[0]);
                    const uint64_t B1 = __builtin_bswap64(s64[1]);
                    const int sh = (int)(byte & 7) * 8 + (int)(bit & 7);
                    const uint64_t top = sh ? (B0 << sh) | (B1 >> (64 - sh)) : B0;
#pragma unroll
                    for (int x = 0; x < 8; ++x) sym[x] = (uint32_t)(top >> (59 - 5 * x)) & 31u;
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
        int64_t lo = in ? o : INT64_MAX, hi = in ? o : 0;
#pragma unroll
        for (int d = 16; d > 0; d >>= 1) {
            kp = max(kp, __shfl_xor(kp, d, 64));
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
        }
        kp = (kp + 15) & ~15;
        if (!in) o = lo;
        const int64_t b_lo = ((lo * 5) >> 3) & ~(int64_t)15;
        const int64_t b_hi = (((hi + kp) * 5) >> 3) + 24;
        const bool fast = (b_hi - b_lo) <= CAPB;
        __syncthreads();
        if (fast) {
            const uint4* src = reinterpret_cast<const uint4*>(packed + b_lo);
            uint4* dst = reinterpret_cast<uint4*>(stg);
            const int nchunk = (int)((b_hi - b_lo + 15) >> 4);
            for (int c = lane; c < nchunk; c += 64) dst[c] = src[c];
        }
        __syncthreads();
        const int64_t relbit = 5 * o - 8 * b_lo;
        const int C = kp >> 4;  // chunks holding residues

        // B fragments of the 4 ranks for chunk c (this lane: 8 columns)
        auto bfrags = [&](int c, half8* b) {
            const int pb = 16 * c + 8 * h;
            uint32_t sym[8];
            if (fast) {
                const int64_t bit = relbit + 5 * (int64_t)pb;
                const int64_t byte = bit >> 3;
                const uint64_t* s64 = reinterpret_cast<const uint64_t*>(stg + (byte & ~7));
                const uint64_t B0 = __builtin_bswap64(s64[0]);
                const uint64_t B1 = __builtin_bswap64(s64[1]);
                const int sh = (int)(byte & 7) * 8 + (int)(bit & 7);
                const uint64_t top = sh ? (B0 << sh) | (B1 >> (64 - sh)) : B0;
#pragma unroll
                for (int x = 0; x < 8; ++x) sym[x] = (uint32_t)(top >> (59 - 5 * x)) & 31u;
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

extern "C" int sseqs_db_prefilter(const uint8_t* packed, const int64_t* seq_off,
                                  const int32_t* seq_len, int64_t s_begin, int64_t s_end,
                                  const uint8_t* qrev, int query_len, const float* lutA,
                                  const float* lutB, int vocab_size, int k, int N, int maxlen,
                                  float* scores, hipStream_t stream) {
    if (s_end < s_begin) return fail(-1, "db_prefilter: s_end < s_begin");
    const int64_t nseq = s_end - s_begin;
    if (nseq == 0) return 0;
    if (N < 16 || N > 2048 || (N & (N - 1)))
        return fail(-1, "db_prefilter: N must be a power of two in 16..2048");
    if (k < 1 || k > 20) return fail(-1, "db_prefilter: k must be 1..20");
    if (vocab_size < 1) return fail(-1, "db_prefilter: vocab_size must be >= 1");
    const int mlen = maxlen < 0 ? 0 : maxlen;
    const int KPmax = (min(mlen, N) + 15) & ~15;
    // a length-sorted tile spans about 32 sequences of at most KPmax symbols
    int CAPB = (int)(((int64_t)33 * (KPmax + 17) * 5) / 8 + 64);
    CAPB = (CAPB + 15) & ~15;
    size_t lds = (size_t)k * (N + 8) * 2 + (size_t)k * CORR_NB * 2 + (size_t)CAPB + 64;
    lds = (lds + 15) & ~(size_t)15;
    if (lds > 160 * 1024) return fail(-1, "db_prefilter: tables exceed LDS (reduce k or N)");
    const int64_t tiles = ceil_div(nseq, 32);
    const unsigned grid = (unsigned)(tiles < 8192 ? tiles : 8192);
    int nmt = N >= 64 ? N / 64 : 1;
    // m-tiles per pass: more reuse of each decoded B fragment against more
    // registers (8 -> 1 wave per SIMD, 4 -> 2).  Measured on gfx950 with
    // rank 4: N = 512 runs fastest at 4, N = 1024 and 2048 at 8.
    if (nmt > 4 && N <= 512) nmt = 4;
    // SSEQS_CORR_MT caps it further (tuning knob; results do not depend on it)
    if (const char* e = getenv("SSEQS_CORR_MT")) {
        int cap = atoi(e);
        while (nmt > 1 && nmt > cap) nmt >>= 1;
    }
    auto go = [&](auto kern) -> int {
        if (lds > 64 * 1024 &&
            hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return fail(-3, "db_prefilter: cannot enable large LDS");
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, stream, packed, seq_off, seq_len,
                           s_begin, nseq, qrev, query_len, lutA, lutB, vocab_size, k, N, mlen,
                           CAPB, scores);
        return check_launch("db_prefilter");
    };
    if (k == 4) {
        auto go4 = [&](auto kern) -> int {
            if (lds > 64 * 1024 &&
                hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
                return fail(-3, "db_prefilter: cannot enable large LDS");
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, stream, packed, seq_off, seq_len,
                               s_begin, nseq, qrev, query_len, lutA, lutB, vocab_size, N, mlen,
                               CAPB, scores);
            return check_launch("db_prefilter");
        };
        switch (nmt) {
        case 1: return go4(&k_corr_db4<1>);
        case 2: return go4(&k_corr_db4<2>);
        case 4: return go4(&k_corr_db4<4>);
        default: return go4(&k_corr_db4<8>);
        }
    }
    switch (nmt) {
    case 1: return go(&k_corr_db<1>);
    case 2: return go(&k_corr_db<2>);
    case 4: return go(&k_corr_db<4>);
    default: return go(&k_corr_db<8>);
    }
}
