// Affine DP + traceback for a3m rows (replaces the Triton kernels
// sw_diag_batch_affine / backtrack_kernel_batched_affine, sseqs/sw.py:408-768).
//
// DP (i = query 1..Lq, j = target 1..T, zero borders, -5 for non-residues):
//   M  = s + H(i-1,j-1)
//   Ix = max(H(i-1,j)-o, (i-1==0 ? -16384 : Ix(i-1,j)) - e)
//   Iy = max(H(i,j-1)-o, (j-1==0 ? -16384 : Iy(i,j-1)) - e)
//   H  = max(0, M, Ix, Iy)
// The reference stores the three int16 matrices and re-derives the move at
// every traceback cell (H == M -> diagonal, else H == Ix -> up, else
// H == Iy -> left, stop at H == 0).  Those comparisons depend only on the
// cell, so two bits per cell (the move) are stored instead of six bytes.
//
// One wave per target.  The query is processed in strips of 64 rows; lane l
// owns row 64k+l+1 and handles column j at step j+l-1, receiving H, Ix and the
// target residue of the cell above from lane l-1 through one DPP wave shift
// per step (a systolic anti-diagonal sweep).  The strip's last row is kept in
// LDS for the next strip.  The 64 moves of one step are two 64-bit ballots,
// stored as one 16-byte record: record (k, s) holds the moves of cells
// (64k+l+1, s-l+1), l = 0..63.  End cell: the first cell in row-major order
// that holds the maximum (sseqs/sw.py:620-637); the wave then follows the
// moves through windows of 128 records staged in LDS.
#include "common.h"

namespace sseqs {

constexpr int BT_NEG = -16384;  // VERY_NEGATIVE_SCORE of sseqs/sw.py:517

// Same DP and traceback for targets too long for k_backtrack's LDS row
// buffer: one move byte per cell and the strip boundary row in global memory.
__global__ __launch_bounds__(64) void k_backtrack_global(
    const uint8_t* __restrict__ query, int Lq, const uint8_t* __restrict__ targets,
    const int64_t* __restrict__ t_off, const int32_t* __restrict__ t_len,
    const int64_t* __restrict__ cell_off, int B, int go, int ge, int A, uint8_t* __restrict__ out_q,
    uint8_t* __restrict__ out_t, int32_t* __restrict__ info, uint8_t* __restrict__ dirs_all,
    int2* __restrict__ rowbuf_all, Blosum bl) {
    __shared__ int8_t s_bl[400];
    const int lane = threadIdx.x;
    for (int i = lane; i < 400; i += 64) s_bl[i] = bl.v[i];
    __syncthreads();
    const int b = blockIdx.x;
    if (b >= B) return;
    const int T = t_len[b];
    const uint8_t* tq = targets + t_off[b];
    uint8_t* dirs = dirs_all + cell_off[b];
    volatile int* rb = reinterpret_cast<int*>(rowbuf_all + (Lq > 0 ? cell_off[b] / Lq : 0));

    int best = 0;
    int64_t bestpos = INT64_MAX;
    if (Lq > 0 && T > 0) {
        for (int i0 = 0; i0 < Lq; i0 += 64) {
            const int i = i0 + lane + 1;  // 1-based row of this lane
            const bool row_ok = i <= Lq;
            const int qa = row_ok ? aa_index(query[i - 1]) : 20;
            int Hleft = 0, Yleft = BT_NEG;   // H(i,0) = 0; Iy(i,0) treated as -16384
            int diagH = 0;                   // H(i-1, j-1), starts at H(i-1, 0) = 0
            int sendH = 0, sendX = BT_NEG;
            const int steps = T + 63;
            for (int s = 0; s < steps; ++s) {
                const int j = s - lane + 1;
                int upH = __shfl_up(sendH, 1, 64);
                int upX = __shfl_up(sendX, 1, 64);
                if (lane == 0) {
                    if (j >= 1 && j <= T && i0 > 0) {
                        upH = rb[2 * (j - 1)];
                        upX = rb[2 * (j - 1) + 1];
                    } else {
                        upH = 0;
                        upX = BT_NEG;
                    }
                }
                if (j >= 1 && j <= T && row_ok) {
                    const int tb = aa_index(tq[j - 1]);
                    const int sc = (qa < 20 && tb < 20) ? (int)s_bl[qa * 20 + tb] : -5;
                    const int m = sc + diagH;
                    const int x = max(upH - go, (i == 1 ? BT_NEG : upX) - ge);
                    const int y = max(Hleft - go, (j == 1 ? BT_NEG : Yleft) - ge);
                    const int h = max(0, max(m, max(x, y)));
                    const uint8_t d = h == 0 ? 0 : (h == m ? 1 : (h == x ? 2 : 3));
                    const int64_t pos = (int64_t)(i - 1) * T + (j - 1);
                    dirs[pos] = d;
                    if (h > best || (h == best && pos < bestpos)) {
                        best = h;
                        bestpos = pos;
                    }
                    Hleft = h;
                    Yleft = y;
                    sendH = h;
                    sendX = x;
                    if (lane == 63) {
                        rb[2 * (j - 1)] = h;
                        rb[2 * (j - 1) + 1] = x;
                    }
                } else if (j < 1) {
                    sendH = 0;
                    sendX = BT_NEG;
                }
                diagH = upH;
            }
            __threadfence();
        }
        // best value, first row-major position among equals
        for (int off = 32; off > 0; off >>= 1) {
            int ob = __shfl_xor(best, off, 64);
            int64_t op = __shfl_xor(bestpos, off, 64);
            if (ob > best || (ob == best && op < bestpos)) {
                best = ob;
                bestpos = op;
            }
        }
    }
    if (lane != 0) return;
    int32_t* inf = info + (int64_t)b * 6;
    int n = 0, ci = -1, cj = -1, ei = -1, ej = -1;
    if (best > 0 && Lq > 0 && T > 0) {
        ei = (int)(bestpos / T) + 1;
        ej = (int)(bestpos % T) + 1;
        int i = ei, j = ej;
        int w = A - 1;
        const volatile uint8_t* vd = dirs;
        while (i > 0 && j > 0 && w >= 0) {
            const uint8_t d = vd[(int64_t)(i - 1) * T + (j - 1)];
            if (d == 0) break;
            uint8_t qc, tc;
            if (d == 1) {
                qc = query[i - 1];
                tc = tq[j - 1];
                --i;
                --j;
            } else if (d == 2) {
                qc = query[i - 1];
                tc = '-';
                --i;
            } else {
                qc = '-';
                tc = tq[j - 1];
                --j;
            }
            out_q[(int64_t)b * A + w] = qc;
            out_t[(int64_t)b * A + w] = tc;
            --w;
            ++n;
        }
        ci = i;
        cj = j;
    }
    if (n > 0) {
        inf[0] = n;
        inf[1] = ci;
        inf[2] = ei - 1;
        inf[3] = cj;
        inf[4] = ej - 1;
    } else {
        inf[0] = 0;
        inf[1] = -1;
        inf[2] = -1;
        inf[3] = -1;
        inf[4] = -1;
    }
    inf[5] = best;
}

__device__ inline int dpp_shr1(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}
// wave_shr:1 with bound_ctrl: lane 0 reads 0 (callers replace lane 0's value)
__device__ inline uint32_t dpp_shr1u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}

// Two 16-bit lanes per register: low half the strip's upper row of a lane,
// high half its lower row.
typedef short bt_i16x2 __attribute__((ext_vector_type(2)));
__device__ inline bt_i16x2 bt_v(uint32_t u) { return __builtin_bit_cast(bt_i16x2, u); }
__device__ inline uint32_t bt_u(bt_i16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ inline bt_i16x2 bt_max(bt_i16x2 a, bt_i16x2 b) { return __builtin_elementwise_max(a, b); }

constexpr int BT_WIN = 128;  // steps per traceback window (two 16-byte records each)
// LDS bytes shared by the row buffer and the traceback window
__host__ __device__ inline int bt_region_bytes(int tcap) {
    const int rbb = 8 * (tcap > 0 ? tcap : 0), wb = BT_WIN * 32;
    return ((rbb > wb ? rbb : wb) + 15) & ~15;
}
constexpr uint32_t BT_NEG2 = 0xC000C000u;  // BT_NEG in both halves

// Move records of target b start at uint4 2 * NS * (cell_off[b] / Lq + 128 b),
// NS = ceil(Lq / 128) strips of T + 128 steps, two records per step.
//
// One wave per target.  The query is processed in strips of 128 rows; lane l
// owns rows 128k + 2l + 1 (the "a" row, low halves) and 128k + 2l + 2 (the
// "b" row, high halves) and at step s handles column s - 2l + 1 of its a row
// and one column less of its b row, every cell update in packed 16-bit
// arithmetic for both rows at once (DP values stay within +-22528 + the gap
// parameters, so 16 bits are exact).  The a row takes H and Ix of the cell
// above from lane l-1's b row through one DPP wave shift per step; the b row
// takes them from the lane's own a row one step earlier.  The strip's last
// row (lane 63's b row) is kept in LDS for the next strip.  A step's moves
// are four 64-bit ballots (two per row), stored as two 16-byte records:
// record 2 (k (T + 128) + s) + h holds the moves of the cells of row
// 128k + 2l + 1 + h at column s - 2l + 1 - h, bit l.  End cell: the first
// cell in row-major order that holds the maximum (sseqs/sw.py:620-637); the
// wave then follows the moves through windows of 128 steps staged in LDS.
//
// Dynamic LDS: one region for the row buffer (tcap x int2, the DP) and then
// the traceback window (BT_WIN x 32 B) | BLOSUM 400 B | target residues tcap
// B.  Sharing the region keeps a wave's LDS near 5 KB for 500-residue
// targets, so that eight waves fit a SIMD to hide each step's latency.
// Targets longer than tcap are rejected (info all -2).
__global__ __launch_bounds__(64) void k_backtrack(
    const uint8_t* __restrict__ query, int Lq, const uint8_t* __restrict__ targets,
    const int64_t* __restrict__ t_off, const int32_t* __restrict__ t_len,
    const int64_t* __restrict__ cell_off, int go, int ge, int A, uint8_t* __restrict__ out_q,
    uint8_t* __restrict__ out_t, int32_t* __restrict__ info, uint4* __restrict__ recs_all,
    int tcap, Blosum bl) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint4* win = reinterpret_cast<uint4*>(smem);
    int2* rb = reinterpret_cast<int2*>(smem);
    const int region = bt_region_bytes(tcap);
    int8_t* s_bl = reinterpret_cast<int8_t*>(smem + region);
    uint8_t* tq = reinterpret_cast<uint8_t*>(s_bl + 400);
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int T = t_len[b];
    int32_t* inf = info + (int64_t)b * 6;
    if (T > tcap) {
        if (lane < 6) inf[lane] = -2;
        return;
    }
    for (int i = lane; i < 400; i += 64) s_bl[i] = bl.v[i];
    const uint8_t* tg = targets + t_off[b];
    for (int j = lane; j < T; j += 64) tq[j] = (uint8_t)aa_index(tg[j]);
    __syncthreads();

    const int NS = (Lq + 127) >> 7;
    const int RS = T + 128;  // steps per strip
    uint4* recs = recs_all + (Lq > 0 ? 2LL * NS * (cell_off[b] / Lq + 128LL * b) : 0);
    const bt_i16x2 GO = bt_v((uint32_t)(uint16_t)go * 0x10001u);
    const bt_i16x2 GE = bt_v((uint32_t)(uint16_t)ge * 0x10001u);
    const bt_i16x2 ZERO = bt_v(0u);
    int best = 0, bi = 0, bj = 0;  // this lane's first maximum, in (row, column) order
    if (Lq > 0 && T > 0) {
        for (int k = 0; k < NS; ++k) {
            const int ia = 128 * k + 2 * lane + 1, ib = ia + 1;
            const bool oka = ia <= Lq, okb = ib <= Lq;
            const int qa = oka ? aa_index(query[ia - 1]) : 20;
            const int qb = okb ? aa_index(query[ib - 1]) : 20;
            const int8_t* browa = s_bl + (qa < 20 ? qa : 0) * 20;
            const int8_t* browb = s_bl + (qb < 20 ? qb : 0) * 20;
            const bool resa = qa < 20, resb = qb < 20;
            // state of both rows: H (the cell to the left, and what the next
            // lane sees above), Ix, Iy; the cell above-left (diagonal)
            uint32_t H = 0, X = BT_NEG2, Y = BT_NEG2, diag = 0;
            int bestA = 0, bjA = 0, bestB = 0, bjB = 0;
            // Software pipeline: each step reads the next step's substitution
            // scores and lane 0's inputs (the row above from rb, the target
            // residue two steps ahead), so no LDS read sits in a step's chain.
            // The a row's residue at step s+1 is lane l-1's b-row residue at
            // step s; the b row's is the lane's own a-row residue at step s.
            int2 r_next = k > 0 && T > 0 ? rb[0] : make_int2(0, BT_NEG);
            int tA = lane == 0 && T > 0 ? (int)tq[0] : 20, tB = 20;
            int t_next = T > 1 ? (int)tq[1] : 20;
            int sa = (resa && tA < 20) ? (int)browa[tA] : -5;
            int sb = (resb && tB < 20) ? (int)browb[tB] : -5;
            const int steps = T + 127;
            for (int s = 0; s < steps; ++s) {
                const int ja = s - 2 * lane + 1, jb = ja - 1;
                const int2 r = r_next;
                const uint32_t sc = ((uint32_t)sa & 0xffffu) | ((uint32_t)sb << 16);
                int tA_next = (int)dpp_shr1u((uint32_t)tB);
                tA_next = lane == 0 ? t_next : tA_next;
                const int tB_next = tA;
                // unconditional in-range reads, then the select
                const int ga = (int)browa[tA_next < 20 ? tA_next : 0];
                const int gb = (int)browb[tB_next < 20 ? tB_next : 0];
                const int sa_next = (resa && tA_next < 20) ? ga : -5;
                const int sb_next = (resb && tB_next < 20) ? gb : -5;
                const int sn = s + 1;  // lane 0's a-row column at step sn is sn + 1
                const int2 rr = rb[sn < T ? sn : T - 1];
                r_next = k > 0 && sn < T ? rr : make_int2(0, BT_NEG);
                const int tt = (int)tq[s + 2 < T ? s + 2 : T - 1];
                const int t_nn = s + 2 < T ? tt : 20;
                // the cells above: the a row's from lane l-1's b row (lane 0:
                // the previous strip's last row), the b row's from its a row
                uint32_t hin = dpp_shr1u(H), xin = dpp_shr1u(X);
                hin = lane == 0 ? (uint32_t)r.x << 16 : hin;
                xin = lane == 0 ? (uint32_t)r.y << 16 : xin;
                const uint32_t up = __builtin_amdgcn_alignbit(H, hin, 16);
                const uint32_t upX = __builtin_amdgcn_alignbit(X, xin, 16);
                const bool va = ja >= 1 && ja <= T && oka;
                const bool vb = jb >= 1 && jb <= T && okb;
                const bt_i16x2 m = bt_v(sc) + bt_v(diag);
                const bt_i16x2 x = bt_max(bt_v(up) - GO, bt_v(upX) - GE);
                const bt_i16x2 y = bt_max(bt_v(H) - GO, bt_v(Y) - GE);
                const bt_i16x2 h = bt_max(bt_max(m, x), bt_max(y, ZERO));
                // move: 0 stop (h == 0), 1 diagonal (h == m), 2 up (h == x), 3 left
                const bool nza = h.x != 0, hma = h.x == m.x, hxa = h.x == x.x;
                const bool nzb = h.y != 0, hmb = h.y == m.y, hxb = h.y == x.y;
                const uint64_t p0a = __ballot(va && nza && (hma || !hxa));
                const uint64_t p1a = __ballot(va && nza && !hma);
                const uint64_t p0b = __ballot(vb && nzb && (hmb || !hxb));
                const uint64_t p1b = __ballot(vb && nzb && !hmb);
                if (lane == 0) {
                    uint4* rec = recs + 2 * ((int64_t)k * RS + s);
                    rec[0] = make_uint4((uint32_t)p0a, (uint32_t)(p0a >> 32), (uint32_t)p1a,
                                        (uint32_t)(p1a >> 32));
                    rec[1] = make_uint4((uint32_t)p0b, (uint32_t)(p0b >> 32), (uint32_t)p1b,
                                        (uint32_t)(p1b >> 32));
                }
                const bool betterA = va && h.x > bestA, betterB = vb && h.y > bestB;
                bestA = betterA ? h.x : bestA;
                bjA = betterA ? ja : bjA;
                bestB = betterB ? h.y : bestB;
                bjB = betterB ? jb : bjB;
                // rows update only on their valid cells; before a row starts its
                // state stays (H 0, Ix and Iy BT_NEG), the left boundary
                const uint32_t msk = (va ? 0xffffu : 0u) | (vb ? 0xffff0000u : 0u);
                H = (bt_u(h) & msk) | (H & ~msk);
                X = (bt_u(x) & msk) | (X & ~msk);
                Y = (bt_u(y) & msk) | (Y & ~msk);
                if (lane == 63 && vb) rb[jb - 1] = make_int2(h.y, x.y);
                diag = up;
                tA = tA_next;
                tB = tB_next;
                sa = sa_next;
                sb = sb_next;
                t_next = t_nn;
            }
            // the lane's rows in order: a strictly larger value replaces, so
            // ties keep the earlier row (and the earlier column within a row)
            if (bestA > best) {
                best = bestA;
                bi = ia;
                bj = bjA;
            }
            if (bestB > best) {
                best = bestB;
                bi = ib;
                bj = bjB;
            }
            __syncthreads();  // the strip's last row is in rb
        }
        // best value, then the first cell in row-major order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            const int oj = __shfl_xor(bj, off, 64);
            if (ob > best || (ob == best && (oi < bi || (oi == bi && oj < bj)))) {
                best = ob;
                bi = oi;
                bj = oj;
            }
        }
        __threadfence();  // the move records are visible to the loads below
    }
    // traceback: every lane follows the same path; lane 0 writes the rows
    int n = 0, ci = -1, cj = -1, ei = -1, ej = -1;
    if (best > 0) {
        ei = bi;
        ej = bj;
        int i = ei, j = ej;
        int w = A - 1;
        int wk = -1, wlo = 0;  // strip and first step held in the window
        while (i > 0 && j > 0 && w >= 0) {
            const int k = (i - 1) >> 7, rr = (i - 1) & 127, l = rr >> 1, hh = rr & 1;
            const int st = j - 1 + 2 * l + hh;
            if (k != wk || st < wlo || st >= wlo + BT_WIN) {
                wk = k;
                wlo = max(0, st - (BT_WIN - 1));
                __syncthreads();
                for (int t = lane; t < 2 * BT_WIN; t += 64)
                    if (wlo + (t >> 1) < RS) win[t] = recs[2 * ((int64_t)k * RS + wlo) + t];
                __syncthreads();
            }
            const uint4 r = win[2 * (st - wlo) + hh];
            const uint32_t w0 = l < 32 ? r.x : r.y, w1 = l < 32 ? r.z : r.w;
            const int d = (int)((w0 >> (l & 31)) & 1u) | (int)(((w1 >> (l & 31)) & 1u) << 1);
            if (d == 0) break;
            uint8_t qc, tc;
            if (d == 1) {
                qc = query[i - 1];
                tc = tg[j - 1];
                --i;
                --j;
            } else if (d == 2) {
                qc = query[i - 1];
                tc = '-';
                --i;
            } else {
                qc = '-';
                tc = tg[j - 1];
                --j;
            }
            if (lane == 0) {
                out_q[(int64_t)b * A + w] = qc;
                out_t[(int64_t)b * A + w] = tc;
            }
            --w;
            ++n;
        }
        ci = i;
        cj = j;
    }
    if (lane != 0) return;
    if (n > 0) {
        inf[0] = n;
        inf[1] = ci;
        inf[2] = ei - 1;
        inf[3] = cj;
        inf[4] = ej - 1;
    } else {
        inf[0] = 0;
        inf[1] = -1;
        inf[2] = -1;
        inf[3] = -1;
        inf[4] = -1;
    }
    inf[5] = best;
}

}  // namespace sseqs

using namespace sseqs;

// Workspace: the move records of k_backtrack or, for targets too long for
// its LDS row buffer, the direction bytes + row buffer of k_backtrack_global.
static size_t bt_records_bytes(int query_len, int64_t total_cells, int B) {
    if (query_len <= 0) return 0;
    const int64_t ns = (query_len + 127) / 128;
    return (size_t)(ns * 32 * (total_cells / query_len + 128LL * B));
}
static size_t bt_global_bytes(int query_len, int64_t total_cells) {
    const int64_t rows = query_len > 0 ? total_cells / query_len : 0;
    return (size_t)((total_cells + 255) & ~255LL) + (size_t)rows * 8;
}

extern "C" size_t sseqs_backtrack_ws_bytes(int query_len, int64_t total_cells, int B) {
    const size_t a = bt_records_bytes(query_len, total_cells, B);
    const size_t g = bt_global_bytes(query_len, total_cells);
    return (a > g ? a : g) + 256;
}

extern "C" int sseqs_backtrack(const uint8_t* query, int query_len, const uint8_t* targets,
                               const int64_t* t_off, const int32_t* t_len, const int64_t* cell_off,
                               int64_t total_cells, int B, int gap_open, int gap_extend,
                               int max_align, uint8_t* out_q, uint8_t* out_t, int32_t* info,
                               void* ws, size_t ws_bytes, hipStream_t stream) {
    if (B < 0 || query_len < 0 || total_cells < 0) return fail(-1, "backtrack: negative size");
    if (B == 0) return 0;
    if (max_align < 1) return fail(-1, "backtrack: max_align must be >= 1");
    if (ws_bytes < sseqs_backtrack_ws_bytes(query_len, total_cells, B))
        return fail(-2, "backtrack: workspace too small");
    // targets are at most max_align - query_len - 1 long (the documented bound)
    const int tcap = max_align - query_len - 1;
    const size_t lds = ((size_t)bt_region_bytes(tcap) + 400 + (size_t)(tcap > 0 ? tcap : 0) + 15) &
                       ~(size_t)15;
    if (tcap >= 1 && lds <= 48 * 1024) {
        hipLaunchKernelGGL(k_backtrack, dim3((unsigned)B), dim3(64), lds, stream, query,
                           query_len, targets, t_off, t_len, cell_off, gap_open, gap_extend,
                           max_align, out_q, out_t, info, static_cast<uint4*>(ws), tcap,
                           blosum62());
        return check_launch("backtrack");
    }
    uint8_t* dirs = static_cast<uint8_t*>(ws);
    int2* rowbuf = reinterpret_cast<int2*>(dirs + ((total_cells + 255) & ~255LL));
    hipLaunchKernelGGL(k_backtrack_global, dim3((unsigned)B), dim3(64), 0, stream, query,
                       query_len, targets, t_off, t_len, cell_off, B, gap_open, gap_extend,
                       max_align, out_q, out_t, info, dirs, rowbuf, blosum62());
    return check_launch("backtrack");
}
