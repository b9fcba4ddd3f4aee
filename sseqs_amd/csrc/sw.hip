// Smith-Waterman scoring with the reference's 32-stripe rule.
//
// Semantics (sseqs/sw.cu:55-325): rows are target residues, columns query
// residues; the query is cut into 32 stripes of TPT = ceil(Lq/32) columns and
// the horizontal (left) operand is 0 at the first column of every stripe,
// while the diagonal operand crosses stripes.  Linear:
//   H = max(0, Hdiag+s, Hup+g, Hleft+g)
// Affine (gaps open from M only):
//   M = max(0, s + max(M,X,Y)diag), X = max(0, Mleft-o, Xleft-e),
//   Y = max(0, Mup-o, Yup-e), score = max over cells of max(M,X,Y).
//
// Layout on gfx950: one wave64 scores two targets, lanes 0-31 and 32-63 each
// own the 32 stripes of one target and sweep its rows; per row a lane reads
// its stripe of the query profile (int8, [residue][lane][TPT4]) from LDS and
// takes the diagonal operand of its first column from the left lane with one
// cross-lane shift.  Both halves run max(T0, T1) rows; the shorter target is
// padded with a non-residue row (score -4 everywhere), which can never raise
// the running maximum, so scores are unchanged.  Columns past the query end
// are padded the same way.
#include <type_traits>

#include "common.h"

namespace sseqs {

constexpr int SW_CH = 256;  // target rows staged per LDS refill

// Targets given as the reference passes them: good_idx into a delimiter
// position array P over an ASCII batch (sseqs/sw.cu:67-77).
struct SrcAscii {
    const uint8_t* ascii;
    const int32_t* P;
    const int32_t* good_idx;
    __device__ bool target(int64_t g, int64_t& start, int& len) const {
        int gi = good_idx[g];
        int64_t prevp = gi > 0 ? (int64_t)P[gi - 1] : -1;  // P[-1] taken as -1
        int64_t st = prevp + 1, stop = (int64_t)P[gi] + 1;
        int64_t length = stop - st - 1;
        const bool ok = !(length == 0 || stop < st);
        // a skipped target reports its delimiter, which lies inside the batch
        start = ok ? st : (int64_t)P[gi];
        len = ok && length > 0 ? (int)length : 0;
        return ok;
    }
    __device__ uint32_t sym(int64_t start, int p) const { return ascii[start + p]; }
};

// Targets given as candidate ids of the HBM-resident packed database.
struct SrcPacked {
    const uint8_t* packed;
    const int64_t* off;
    const int32_t* len;
    const int64_t* cand;
    __device__ bool target(int64_t g, int64_t& start, int& l) const {
        int64_t s = cand[g];
        start = off[s];
        l = len[s];
        return l > 0;
    }
    __device__ uint32_t sym(int64_t start, int p) const { return packed_sym(packed, start + p); }
};

template <int TPT, bool AFFINE, class Src>
__global__ __launch_bounds__(64) void k_sw(Src src, const uint8_t* __restrict__ query, int Lq,
                                           int64_t n, int p1, int p2, int32_t* __restrict__ out,
                                           Blosum bl) {
    constexpr int TPT4 = (TPT + 3) & ~3;
    constexpr int NW = TPT4 / 4;
    __shared__ int8_t s_bl[400];
    __shared__ uint32_t s_prof[21 * 32 * NW];
    __shared__ uint8_t s_res[2][SW_CH];
    __shared__ uint8_t s_code[256];  // byte -> residue index (aa_index)

    const int tid = threadIdx.x;
    const int half = tid >> 5, lane = tid & 31;

    for (int i = tid; i < 400; i += 64) s_bl[i] = bl.v[i];
    for (int i = tid; i < 256; i += 64) s_code[i] = (uint8_t)aa_index((uint32_t)i);
    __syncthreads();
    // profile word w of (residue c, lane ln) packs columns ln*TPT + 4w .. +3
    for (int idx = tid; idx < 21 * 32 * NW; idx += 64) {
        int c = idx / (32 * NW);
        int rem = idx - c * 32 * NW;
        int ln = rem / NW, w = rem - ln * NW;
        uint32_t word = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int k = 4 * w + b;
            int j = ln * TPT + k;
            int v = -4;
            if (k < TPT && j < Lq && c < 20) {
                int qc = query[j];
                v = qc < 20 ? (int)s_bl[c * 20 + qc] : -4;
            }
            word |= (uint32_t)(uint8_t)(int8_t)v << (8 * b);
        }
        s_prof[idx] = word;
    }
    __syncthreads();

    const int go = p1, ge = p2;  // affine: open, extend; linear: p1 = gap (<= 0)
    for (int64_t pair = blockIdx.x; pair * 2 < n; pair += gridDim.x) {
        const int64_t g = pair * 2 + half;
        int64_t start = 0;
        int len = 0;
        bool valid = false;
        if (g < n) valid = src.target(g, start, len);
        const int T = valid ? len : 0;
        const int Tmax = max(T, __shfl_xor(T, 32, 64));

        int Hp[TPT], Mp[TPT], Yp[TPT];
#pragma unroll
        for (int k = 0; k < TPT; ++k) { Hp[k] = 0; Mp[k] = 0; Yp[k] = 0; }
        int best = 0;

        for (int c0 = 0; c0 < Tmax; c0 += SW_CH) {
            const int rows = min(SW_CH, Tmax - c0);
            __syncthreads();
            // four rows per lane per step, loads first; a row past the
            // target's end re-reads its last residue (or, for an empty or
            // skipped target, the symbol at `start`, which is in bounds) and
            // stores 20
            for (int r0 = lane; r0 < rows; r0 += 128) {
                uint32_t sy[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    sy[u] = src.sym(start, max(min(c0 + r0 + 32 * u, T - 1), 0));
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = r0 + 32 * u;
                    if (r < rows) s_res[half][r] = c0 + r < T ? s_code[sy[u] & 255u] : (uint8_t)20;
                }
            }
            __syncthreads();
            for (int r = 0; r < rows; ++r) {
                const int c = s_res[half][r];
                const uint32_t* pr = &s_prof[(c * 32 + lane) * NW];
                uint32_t pw[NW];
#pragma unroll
                for (int w = 0; w < NW; ++w) pw[w] = pr[w];
                int din = __shfl_up(Hp[TPT - 1], 1, 32);
                if (lane == 0) din = 0;
                int dg = din;
                if constexpr (AFFINE) {
                    int lM = 0, lX = 0;
#pragma unroll
                    for (int k = 0; k < TPT; ++k) {
                        const int s = (int)(int8_t)(pw[k >> 2] >> (8 * (k & 3)));
                        const int m = max(dg + s, 0);
                        const int x = max(max(lM - go, lX - ge), 0);
                        const int y = max(max(Mp[k] - go, Yp[k] - ge), 0);
                        const int h = max(m, max(x, y));
                        dg = Hp[k];
                        Hp[k] = h;
                        Mp[k] = m;
                        Yp[k] = y;
                        lM = m;
                        lX = x;
                        best = max(best, h);
                    }
                } else {
                    int left = 0;
#pragma unroll
                    for (int k = 0; k < TPT; ++k) {
                        const int s = (int)(int8_t)(pw[k >> 2] >> (8 * (k & 3)));
                        const int h = max(max(dg + s, Hp[k] + go), max(left + go, 0));
                        dg = Hp[k];
                        Hp[k] = h;
                        left = h;
                        best = max(best, h);
                    }
                }
            }
        }
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, 32));
        if (lane == 0 && g < n) out[g] = valid ? best : -1;
    }
}

// Packed 16-bit variant for queries of up to 1024 residues (TPT <= 32).
// A target is held by the 16 lanes of one DPP row; lane l owns stripes l
// (low halves) and l + 16 (high halves) of the 32-stripe rule, so one
// packed instruction advances two cells, and a wave scores four targets.
// The diagonal operand of a stripe's first column is the previous row's
// last column of the stripe before it: row rotation by one lane gives
// stripe l-1 (low) and l+15 (high) for lanes 1..15; lane 0 takes 0 for
// stripe 0 and stripe 15 (lane 15, low half) for stripe 16.
// Exactness: every H, M, X and Y is in [0, 11 * Lq] (11 = largest BLOSUM62
// entry), and the operands subtracted from them are the gap parameters, so
// 16-bit arithmetic is exact while 11 * 1024 + gap parameters < 32768 (the
// unclamped M and X of the affine update stay above -4 - 2 * 16000); the
// host sends other parameters to the 32-bit kernel.
// The query profile ([residue][word][lane] x 4 packed pairs) is built once
// per block of SW16_WAVES waves; each wave stages its own targets' residues.
typedef short i16x2 __attribute__((ext_vector_type(2)));
constexpr int SW16_WAVES = 4;

__device__ inline i16x2 as_i16x2(uint32_t u) { return __builtin_bit_cast(i16x2, u); }
__device__ inline uint32_t as_u32(i16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ inline i16x2 vmax(i16x2 a, i16x2 b) { return __builtin_elementwise_max(a, b); }
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
// max(a - b, 0) for a, b >= 0 in one instruction (unsigned saturating subtract)
__device__ inline i16x2 subsat(i16x2 a, u16x2 b) {
    return __builtin_bit_cast(
        i16x2, __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, a), b));
}

template <int CTRL>
__device__ inline int sw_dpp(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}

// Query profile of the packed kernel: s_prof[(c * NW + w) * 16 + ln] packs
// the scores of residue c against columns (ln + 16h) * TPT + 4w .. +3 of
// the query (h = 0 low halves, h = 1 high halves), -4 past the query end.
template <int TPT>
__device__ void sw16_profile(const uint8_t* __restrict__ query, int Lq,
                             const int8_t* __restrict__ s_bl, uint4* __restrict__ s_prof,
                             int tid, int nthr) {
    constexpr int NW = (TPT + 3) / 4;
    for (int idx = tid; idx < 21 * NW * 16; idx += nthr) {
        const int c = idx / (NW * 16);
        const int rem = idx - c * NW * 16;
        const int w = rem >> 4, ln = rem & 15;
        uint32_t v[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = 4 * w + b;
            int sv[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int j = (ln + 16 * h) * TPT + k;
                int sc = -4;
                if (k < TPT && j < Lq && c < 20) {
                    const int qc = query[j];
                    sc = qc < 20 ? (int)s_bl[c * 20 + qc] : -4;
                }
                sv[h] = sc;
            }
            v[b] = (uint32_t)(uint16_t)(int16_t)sv[0] | ((uint32_t)(uint16_t)(int16_t)sv[1] << 16);
        }
        s_prof[idx] = make_uint4(v[0], v[1], v[2], v[3]);
    }
}

// Scores of the four targets of one quad (quad * 4 + grp, grp = lane / 16)
// by one wave: out[g] for g < n.  `res` is this row group's residue buffer
// of SW_CH bytes in LDS.
template <int TPT, bool AFFINE, class Src>
__device__ __forceinline__ void sw16_quad(const Src& src, int64_t quad, int64_t n,
                                          int32_t* __restrict__ out,
                                          const uint4* __restrict__ s_prof,
                                          const uint8_t* __restrict__ s_code,
                                          uint8_t* __restrict__ res, int lane, int p1, int p2) {
    constexpr int NW = (TPT + 3) / 4;
    const int grp = lane >> 4, l = lane & 15;
    const i16x2 zero = {0, 0};
    // byte selector of the diagonal operand: the rotated word as it is, or
    // for lane 0 (stripes 0 and 16) 0 below and the word's low half above
    const uint32_t dg_sel = l == 0 ? 0x01000c0cu : 0x03020100u;
    // affine: open, extend; linear: p1 = gap (<= 0)
    const i16x2 c1 = {(short)p1, (short)p1}, c2 = {(short)p2, (short)p2};
    // the gap costs as unsigned amounts: extend (affine), -gap (linear)
    const u16x2 u2 = {(unsigned short)p2, (unsigned short)p2};
    const u16x2 ug = {(unsigned short)-p1, (unsigned short)-p1};
    const int64_t g = quad * 4 + grp;
    int64_t start = 0;
    int len = 0;
    bool valid = false;
    if (g < n) valid = src.target(g, start, len);
    const int T = valid ? len : 0;
    int Tmax = max(T, __shfl_xor(T, 16, 64));
    Tmax = max(Tmax, __shfl_xor(Tmax, 32, 64));
    Tmax = __builtin_amdgcn_readfirstlane(Tmax);  // uniform: loops run on scalars

    // Mo holds M - open of the previous row: the one value both the next
    // row (Y) and the next column (X) take from M
    i16x2 Hp[TPT], Mo[TPT], Yp[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        Hp[k] = zero;
        Mo[k] = zero - c1;
        Yp[k] = zero;
    }
    i16x2 best = zero;

    for (int c0 = 0; c0 < Tmax; c0 += SW_CH) {
        const int rows = min(SW_CH, Tmax - c0);
        // this wave's LDS reads of the previous rows are done before
        // the rows are overwritten, and the writes before the reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // four rows per lane per step, loads first; a row past the
        // target's end re-reads its last residue (or, for an empty or
        // skipped target, the symbol at `start`, which is in bounds) and
        // stores 20
        for (int r0 = l; r0 < rows; r0 += 64) {
            uint32_t sy[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                sy[u] = src.sym(start, max(min(c0 + r0 + 16 * u, T - 1), 0));
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = r0 + 16 * u;
                if (r < rows) res[r] = c0 + r < T ? s_code[sy[u] & 255u] : (uint8_t)20;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // one row of the DP; several per step so that the register
        // rotation of Hp between rows needs no copies
        auto row = [&](const int r) {
            const int c = res[r];
            const uint4* pr = &s_prof[c * NW * 16 + l];
            // row rotation by one lane (every lane of the row reads, so
            // no old value is needed), then lane 0's halves rearranged
            const uint32_t rot = (uint32_t)__builtin_amdgcn_mov_dpp(
                (int)as_u32(Hp[TPT - 1]), 0x121, 0xf, 0xf, false);
            i16x2 dg = as_i16x2(__builtin_amdgcn_perm(rot, rot, dg_sel));
            if constexpr (AFFINE) {
                i16x2 lMo = zero - c1, lX = zero;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const uint4 pw = pr[w * 16];
                    const uint32_t pv[4] = {pw.x, pw.y, pw.z, pw.w};
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int k = 4 * w + b;
                        if (k < TPT) {
                            // m and x without their clamp at 0: m >= -4, and
                            // x >= m_left - open, so neither drifts; only
                            // max(m, x, y) with y >= 0 is used, which equals
                            // the clamped recurrence's H
                            const i16x2 sc = as_i16x2(pv[b]);
                            const i16x2 m = dg + sc;
                            const i16x2 x = vmax(lMo, lX - c2);
                            // max(M_up - open, Yp - extend, 0): Yp >= 0, so
                            // the saturated second term carries the 0
                            const i16x2 y = vmax(Mo[k], subsat(Yp[k], u2));
                            const i16x2 h = vmax(m, vmax(x, y));
                            const i16x2 mo = m - c1;
                            dg = Hp[k];
                            Hp[k] = h;
                            Mo[k] = mo;
                            Yp[k] = y;
                            lMo = mo;
                            lX = x;
                            best = vmax(best, h);
                        }
                    }
                }
            } else {
                i16x2 left = zero;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const uint4 pw = pr[w * 16];
                    const uint32_t pv[4] = {pw.x, pw.y, pw.z, pw.w};
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int k = 4 * w + b;
                        if (k < TPT) {
                            const i16x2 sc = as_i16x2(pv[b]);
                            // max(dg + sc, Hp + gap, left + gap, 0): Hp and
                            // left are >= 0, so max(Hp, left) + gap
                            // saturated at 0 carries the 0
                            const i16x2 h = vmax(dg + sc, subsat(vmax(Hp[k], left), ug));
                            dg = Hp[k];
                            Hp[k] = h;
                            left = h;
                            best = vmax(best, h);
                        }
                    }
                }
            }
        };
        // short stripes: four rows per step (the per-row overhead is a
        // larger share of a short row); long stripes keep their registers
        constexpr int RS = TPT <= 8 ? 4 : 2;
        int r = 0;
        for (; r + RS - 1 < rows; r += RS) {
#pragma unroll
            for (int u = 0; u < RS; ++u) row(r + u);
        }
        for (; r < rows; ++r) row(r);
    }
    int bv = max((int)best.x, (int)best.y);
    bv = max(bv, sw_dpp<0x128>(bv));
    bv = max(bv, sw_dpp<0x124>(bv));
    bv = max(bv, sw_dpp<0x122>(bv));
    bv = max(bv, sw_dpp<0x121>(bv));
    if (l == 0 && g < n) out[g] = valid ? bv : -1;
}

template <int TPT, bool AFFINE, class Src>
__global__ __launch_bounds__(64 * SW16_WAVES) void k_sw16(Src src,
                                                         const uint8_t* __restrict__ query,
                                                         int Lq, int64_t n, int p1, int p2,
                                                         int32_t* __restrict__ out, Blosum bl) {
    constexpr int NW = (TPT + 3) / 4;
    __shared__ int8_t s_bl[400];
    __shared__ uint4 s_prof[21 * NW * 16];
    __shared__ uint8_t s_res[SW16_WAVES][4][SW_CH];
    __shared__ uint8_t s_code[256];  // byte -> residue index (aa_index)

    const int tid = threadIdx.x;
    const int wid = tid >> 6, lane = tid & 63;
    const int grp = lane >> 4;

    for (int i = tid; i < 400; i += 64 * SW16_WAVES) s_bl[i] = bl.v[i];
    for (int i = tid; i < 256; i += 64 * SW16_WAVES) s_code[i] = (uint8_t)aa_index((uint32_t)i);
    __syncthreads();
    sw16_profile<TPT>(query, Lq, s_bl, s_prof, tid, 64 * SW16_WAVES);
    __syncthreads();
    uint8_t* res = s_res[wid][grp];
    for (int64_t quad = (int64_t)blockIdx.x * SW16_WAVES + wid; quad * 4 < n;
         quad += (int64_t)gridDim.x * SW16_WAVES)
        sw16_quad<TPT, AFFINE>(src, quad, n, out, s_prof, s_code, res, lane, p1, p2);
}

// ---------------------------------------------------------------------------
// Batched sweep: affine SW of many queries' candidate lists, one persistent
// launch per stripe width (length bucket) present.
//
// Work item i is chunk item_j[i] (of `item` candidates) of query item_q[i]'s
// candidate list cand[c_off[q] .. c_off[q+1]); the score of candidate c
// goes to out[c].  The host sorts the items of a class by stripe width
// and query; each stripe width (ceil(Lq / 32)) has its own queue.  A block takes the next item from `counter` (one global
// atomic by one lane per item) and builds the query profile in LDS only
// when the item's query differs from the previous one; its waves then score
// the item's candidates four at a time with the same code as k_sw16, so a
// candidate's score is the one a single-query launch gives.  One kernel
// instance per stripe width keeps each bucket at k_sw16's register budget
// (one instance covering several widths is sized for the widest and its
// occupancy drops to 1-2 waves per SIMD).  Every block leaves the loop when
// the queue is empty; the counter is reset by k_sw_reset before the launch.
struct SwBatch {
    const uint8_t* packed;
    const int64_t* off;
    const int32_t* len;
    const uint8_t* qcodes;  // query q: qcodes[q_off[q] .. q_off[q+1]), codes 0..19
    const int64_t* q_off;
    const int64_t* cand;  // candidates of query q: cand[c_off[q] .. c_off[q+1])
    const int64_t* c_off;
    const int32_t* item_q;
    const int32_t* item_j;
    int64_t n_items;
    int item;
    uint32_t* counter;
    int p1, p2;
    int32_t* out;
};

__global__ void k_sw_reset(uint32_t* counter) { *counter = 0u; }

template <int TPT, bool AFFINE>
__global__ __launch_bounds__(64 * SW16_WAVES) void k_sw16_batch(SwBatch b, Blosum bl) {
    constexpr int NW = (TPT + 3) / 4;
    __shared__ int8_t s_bl[400];
    __shared__ uint4 s_prof[21 * NW * 16];
    __shared__ uint8_t s_res[SW16_WAVES][4][SW_CH];
    __shared__ uint8_t s_code[256];
    __shared__ int s_item;

    const int tid = threadIdx.x;
    const int wid = tid >> 6, lane = tid & 63;
    const int grp = lane >> 4;
    for (int i = tid; i < 400; i += 64 * SW16_WAVES) s_bl[i] = bl.v[i];
    for (int i = tid; i < 256; i += 64 * SW16_WAVES) s_code[i] = (uint8_t)aa_index((uint32_t)i);
    uint8_t* res = s_res[wid][grp];
    int cur_q = -1;
    for (;;) {
        if (tid == 0) s_item = (int)min(atomicAdd(b.counter, 1u), 0x7fffffffu);
        __syncthreads();  // also: every wave is done with the previous item
        // uniform across the block: item fields are loaded as scalars
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= b.n_items) break;
        const int q = __builtin_amdgcn_readfirstlane(b.item_q[it]);
        const int64_t qb = b.q_off[q];
        const int Lq = __builtin_amdgcn_readfirstlane((int)(b.q_off[q + 1] - qb));
        if (q != cur_q) {
            sw16_profile<TPT>(b.qcodes + qb, Lq, s_bl, s_prof, tid, 64 * SW16_WAVES);
            cur_q = q;
        }
        __syncthreads();
        const int64_t cb = b.c_off[q] + (int64_t)__builtin_amdgcn_readfirstlane(b.item_j[it]) *
                                            b.item;
        const int64_t n = min(cb + (int64_t)b.item, b.c_off[q + 1]) - cb;
        const SrcPacked src{b.packed, b.off, b.len, b.cand + cb};
        for (int64_t quad = wid; quad * 4 < n; quad += SW16_WAVES)
            sw16_quad<TPT, AFFINE>(src, quad, n, b.out + cb, s_prof, s_code, res, lane, b.p1,
                                   b.p2);
        __syncthreads();  // s_item is rewritten at the top
    }
}

// The packed kernel is exact for these parameters (see k_sw16).
static bool sw16_ok(int Lq, bool affine, int p1, int p2) {
    if (Lq < 1 || Lq > 1024 || getenv("SSEQS_SW_INT32")) return false;
    if (affine) return p1 >= 0 && p2 >= 0 && p1 <= 16000 && p2 <= 16000;
    return p1 <= 0 && p1 >= -16000;
}

template <bool AFFINE, class Src>
static int launch_sw(const Src& src, const uint8_t* q, int Lq, int64_t n, int p1, int p2,
                     int32_t* out, hipStream_t stream) {
    const Blosum bl = blosum62();
    if constexpr (std::is_same<Src, SrcPacked>::value) {
        if (sw16_ok(Lq, AFFINE, p1, p2)) {
            const int64_t quads = ceil_div(n, 4 * SW16_WAVES);
            const unsigned grid16 = (unsigned)(quads < 8192 ? quads : 8192);
            switch ((Lq + 31) >> 5) {
#define SW16_CASE(T)                                                                            \
    case T:                                                                                     \
        hipLaunchKernelGGL((k_sw16<T, AFFINE, Src>), dim3(grid16), dim3(64 * SW16_WAVES), 0,   \
                           stream, src, q, Lq, n, p1, p2, out, bl);                             \
        break;
                SW16_CASE(1) SW16_CASE(2) SW16_CASE(3) SW16_CASE(4) SW16_CASE(5) SW16_CASE(6)
                SW16_CASE(7) SW16_CASE(8) SW16_CASE(9) SW16_CASE(10) SW16_CASE(11) SW16_CASE(12)
                SW16_CASE(13) SW16_CASE(14) SW16_CASE(15) SW16_CASE(16) SW16_CASE(17)
                SW16_CASE(18) SW16_CASE(19) SW16_CASE(20) SW16_CASE(21) SW16_CASE(22)
                SW16_CASE(23) SW16_CASE(24) SW16_CASE(25) SW16_CASE(26) SW16_CASE(27)
                SW16_CASE(28) SW16_CASE(29) SW16_CASE(30) SW16_CASE(31) SW16_CASE(32)
#undef SW16_CASE
            }
            return check_launch(AFFINE ? "sw affine" : "sw linear");
        }
    }
    int64_t pairs = ceil_div(n, 2);
    unsigned grid = (unsigned)(pairs < 16384 ? pairs : 16384);
    const int tpt = (Lq + 31) >> 5;
    switch (tpt) {
#define SW_CASE(T)                                                                              \
    case T:                                                                                     \
        hipLaunchKernelGGL((k_sw<T, AFFINE, Src>), dim3(grid), dim3(64), 0, stream, src, q, Lq, \
                           n, p1, p2, out, bl);                                                 \
        break;
        SW_CASE(1) SW_CASE(2) SW_CASE(3) SW_CASE(4) SW_CASE(5) SW_CASE(6) SW_CASE(7) SW_CASE(8)
        SW_CASE(9) SW_CASE(10) SW_CASE(11) SW_CASE(12) SW_CASE(13) SW_CASE(14) SW_CASE(15)
        SW_CASE(16) SW_CASE(17) SW_CASE(18) SW_CASE(19) SW_CASE(20) SW_CASE(21) SW_CASE(22)
        SW_CASE(23) SW_CASE(24) SW_CASE(25) SW_CASE(26) SW_CASE(27) SW_CASE(28) SW_CASE(29)
        SW_CASE(30) SW_CASE(31) SW_CASE(32) SW_CASE(33) SW_CASE(34) SW_CASE(35) SW_CASE(36)
        SW_CASE(37) SW_CASE(38) SW_CASE(39) SW_CASE(40) SW_CASE(41) SW_CASE(42) SW_CASE(43)
        SW_CASE(44) SW_CASE(45) SW_CASE(46) SW_CASE(47) SW_CASE(48) SW_CASE(49) SW_CASE(50)
        SW_CASE(51) SW_CASE(52) SW_CASE(53) SW_CASE(54) SW_CASE(55) SW_CASE(56) SW_CASE(57)
        SW_CASE(58) SW_CASE(59) SW_CASE(60) SW_CASE(61) SW_CASE(62) SW_CASE(63) SW_CASE(64)
#undef SW_CASE
    default:
        return fail(-1, "sw: query length must be 1..2048");
    }
    return check_launch(AFFINE ? "sw affine" : "sw linear");
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static int fill_minus_one(int32_t* out, int64_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, out, n,
                       -1);
    return check_launch("sw fill");
}

}  // namespace sseqs

using namespace sseqs;

extern "C" int launch_sw_cuda_blosum62(const uint8_t* q, int Lq, const int32_t* good_idx,
                                       const uint8_t* ascii, const int32_t* P, int n, int32_t* out,
                                       int gap_val, hipStream_t stream) {
    if (n < 0) return fail(-1, "sw_cuda_blosum62: negative count");
    if (n == 0) return 0;
    if (Lq < 0 || Lq > 2048) return fail(-1, "sw_cuda_blosum62: query length must be 0..2048");
    if (gap_val > 0) return fail(-1, "sw_cuda_blosum62: gap_penalty must be <= 0");
    if (Lq == 0) return fill_minus_one(out, n, stream);
    SrcAscii src{ascii, P, good_idx};
    return launch_sw<false>(src, q, Lq, n, gap_val, 0, out, stream);
}

extern "C" int launch_sw_cuda_affine(const uint8_t* q, int Lq, const int32_t* good_idx,
                                     const uint8_t* ascii, const int32_t* P, int n, int32_t* out,
                                     int gap_open, int gap_extend, hipStream_t stream) {
    if (n < 0) return fail(-1, "sw_cuda_affine: negative count");
    if (n == 0) return 0;
    if (Lq < 0 || Lq > 2048) return fail(-1, "sw_cuda_affine: query length must be 0..2048");
    if (gap_open < 0 || gap_extend < 0)
        return fail(-1, "sw_cuda_affine: gap_open and gap_extend must be >= 0");
    if (Lq == 0) return fill_minus_one(out, n, stream);
    SrcAscii src{ascii, P, good_idx};
    return launch_sw<true>(src, q, Lq, n, gap_open, gap_extend, out, stream);
}

extern "C" int sseqs_db_sw_affine(const uint8_t* packed, const int64_t* seq_off,
                                  const int32_t* seq_len, const int64_t* cand, int64_t n,
                                  const uint8_t* query, int Lq, int gap_open, int gap_extend,
                                  int32_t* out, hipStream_t stream) {
    if (n < 0) return fail(-1, "db_sw_affine: negative count");
    if (n == 0) return 0;
    if (Lq < 1 || Lq > 2048) return fail(-1, "db_sw_affine: query length must be 1..2048");
    if (gap_open < 0 || gap_extend < 0)
        return fail(-1, "db_sw_affine: gap_open and gap_extend must be >= 0");
    SrcPacked src{packed, seq_off, seq_len, cand};
    return launch_sw<true>(src, query, Lq, n, gap_open, gap_extend, out, stream);
}

extern "C" int sseqs_db_sw_linear(const uint8_t* packed, const int64_t* seq_off,
                                  const int32_t* seq_len, const int64_t* cand, int64_t n,
                                  const uint8_t* query, int Lq, int gap_val, int32_t* out,
                                  hipStream_t stream) {
    if (n < 0) return fail(-1, "db_sw_linear: negative count");
    if (n == 0) return 0;
    if (Lq < 1 || Lq > 2048) return fail(-1, "db_sw_linear: query length must be 1..2048");
    if (gap_val > 0) return fail(-1, "db_sw_linear: gap_val must be <= 0");
    SrcPacked src{packed, seq_off, seq_len, cand};
    return launch_sw<false>(src, query, Lq, n, gap_val, 0, out, stream);
}

extern "C" int sseqs_db_sw_batch(const uint8_t* packed, const int64_t* seq_off,
                                 const int32_t* seq_len, const uint8_t* qcodes,
                                 const int64_t* q_off, const int64_t* cand, const int64_t* c_off,
                                 const int32_t* item_q, const int32_t* item_j, int64_t n_items,
                                 int item, int tpt, int grid, int gap_open,
                                 int gap_extend, int32_t* out, uint32_t* counter,
                                 hipStream_t stream) {
    if (n_items < 0 || item < 1 || grid < 0) return fail(-1, "db_sw_batch: bad sizes");
    if (gap_open < 0 || gap_extend < 0 || gap_open > 16000 || gap_extend > 16000)
        return fail(-1, "db_sw_batch: gap_open and gap_extend must be in 0..16000");
    if (n_items == 0) return 0;
    if (!packed || !seq_off || !seq_len || !qcodes || !q_off || !cand || !c_off || !item_q ||
        !item_j || !out || !counter)
        return fail(-1, "db_sw_batch: null pointer");
    if (n_items > 0x7fffffffLL) return fail(-1, "db_sw_batch: too many work items");
    const SwBatch b{packed, seq_off, seq_len, qcodes, q_off, cand, c_off, item_q, item_j,
                    n_items, item, counter, gap_open, gap_extend, out};
    const Blosum bl = blosum62();
    const unsigned g = (unsigned)(grid > 0 ? grid : 1024);
    hipLaunchKernelGGL(k_sw_reset, dim3(1), dim3(1), 0, stream, counter);
    if (int r = check_launch("db_sw_batch reset")) return r;
    switch (tpt) {
#define SWB_CASE(T)                                                                            \
    case T:                                                                                    \
        hipLaunchKernelGGL((k_sw16_batch<T, true>), dim3(g), dim3(64 * SW16_WAVES), 0, stream, \
                           b, bl);                                                             \
        break;
        SWB_CASE(1) SWB_CASE(2) SWB_CASE(3) SWB_CASE(4) SWB_CASE(5) SWB_CASE(6) SWB_CASE(7)
        SWB_CASE(8) SWB_CASE(9) SWB_CASE(10) SWB_CASE(11) SWB_CASE(12) SWB_CASE(13)
        SWB_CASE(14) SWB_CASE(15) SWB_CASE(16) SWB_CASE(17) SWB_CASE(18) SWB_CASE(19)
        SWB_CASE(20) SWB_CASE(21) SWB_CASE(22) SWB_CASE(23) SWB_CASE(24) SWB_CASE(25)
        SWB_CASE(26) SWB_CASE(27) SWB_CASE(28) SWB_CASE(29) SWB_CASE(30) SWB_CASE(31)
        SWB_CASE(32)
#undef SWB_CASE
    default:
        return fail(-1, "db_sw_batch: stripe width must be 1..32 (queries of 1..1024 residues)");
    }
    return check_launch("db_sw_batch");
}
