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

// Move records of target b start at record NS * (cell_off[b] / Lq + 64 b),
// NS = ceil(Lq / 64) strips of T + 64 records each.
__device__ inline int dpp_shr1(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}

constexpr int BT_WIN = 128;  // records per traceback window

// Dynamic LDS: window BT_WIN x 16 B | row buffer tcap x int2 | BLOSUM 400 B |
// target residues tcap B.  Targets longer than tcap are rejected (info all -2).
__global__ __launch_bounds__(64) void k_backtrack(
    const uint8_t* __restrict__ query, int Lq, const uint8_t* __restrict__ targets,
    const int64_t* __restrict__ t_off, const int32_t* __restrict__ t_len,
    const int64_t* __restrict__ cell_off, int go, int ge, int A, uint8_t* __restrict__ out_q,
    uint8_t* __restrict__ out_t, int32_t* __restrict__ info, uint4* __restrict__ recs_all,
    int tcap, Blosum bl) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint4* win = reinterpret_cast<uint4*>(smem);
    int2* rb = reinterpret_cast<int2*>(smem + BT_WIN * 16);
    int8_t* s_bl = reinterpret_cast<int8_t*>(rb + tcap);
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

    const int NS = (Lq + 63) >> 6;
    const int RS = T + 64;  // records per strip
    uint4* recs = recs_all + (Lq > 0 ? (int64_t)NS * (cell_off[b] / Lq + 64LL * b) : 0);
    int best = 0, bi = 0, bj = 0;  // this lane's first maximum, in (row, column) order
    if (Lq > 0 && T > 0) {
        for (int k = 0; k < NS; ++k) {
            const int i = 64 * k + lane + 1;
            const bool row_ok = i <= Lq;
            const int qa = row_ok ? aa_index(query[i - 1]) : 20;
            const int8_t* brow = s_bl + (qa < 20 ? qa : 0) * 20;
            int Hleft = 0, Yleft = BT_NEG, diagH = 0;
            int sendH = 0, sendX = BT_NEG, sendT = 20;
            const int steps = T + 63;
            for (int s = 0; s < steps; ++s) {
                const int j = s - lane + 1;
                int upH = dpp_shr1(sendH), upX = dpp_shr1(sendX), tb = dpp_shr1(sendT);
                if (lane == 0) {
                    const bool in = s < T;  // column s + 1
                    const int2 r = k > 0 && in ? rb[s] : make_int2(0, BT_NEG);
                    upH = r.x;
                    upX = r.y;
                    tb = in ? tq[s] : 20;
                }
                const bool valid = j >= 1 && j <= T && row_ok;
                const int sc = (qa < 20 && tb < 20) ? (int)brow[tb] : -5;
                const int m = sc + diagH;
                const int x = max(upH - go, (i == 1 ? BT_NEG : upX) - ge);
                const int y = max(Hleft - go, (j == 1 ? BT_NEG : Yleft) - ge);
                const int h = max(0, max(m, max(x, y)));
                const int d = h == 0 ? 0 : (h == m ? 1 : (h == x ? 2 : 3));
                const uint64_t p0 = __ballot(valid && (d & 1));
                const uint64_t p1 = __ballot(valid && (d & 2));
                if (lane == 0)
                    recs[(int64_t)k * RS + s] = make_uint4((uint32_t)p0, (uint32_t)(p0 >> 32),
                                                           (uint32_t)p1, (uint32_t)(p1 >> 32));
                if (valid) {
                    if (h > best) {
                        best = h;
                        bi = i;
                        bj = j;
                    }
                    Hleft = h;
                    Yleft = y;
                    sendH = h;
                    sendX = x;
                    if (lane == 63) rb[j - 1] = make_int2(h, x);
                } else if (j < 1) {
                    sendH = 0;
                    sendX = BT_NEG;
                }
                sendT = tb;
                diagH = upH;
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
        int wk = -1, wlo = 0;  // strip and first record held in the window
        while (i > 0 && j > 0 && w >= 0) {
            const int k = (i - 1) >> 6, l = (i - 1) & 63, st = j - 1 + l;
            if (k != wk || st < wlo || st >= wlo + BT_WIN) {
                wk = k;
                wlo = max(0, st - (BT_WIN - 1));
                __syncthreads();
                for (int t = lane; t < BT_WIN; t += 64)
                    if (wlo + t < RS) win[t] = recs[(int64_t)k * RS + wlo + t];
                __syncthreads();
            }
            const uint4 r = win[st - wlo];
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
    const int64_t ns = (query_len + 63) / 64;
    return (size_t)(ns * 16 * (total_cells / query_len + 64LL * B));
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
    const size_t lds = ((size_t)BT_WIN * 16 + (size_t)(tcap > 0 ? tcap : 0) * 9 + 400 + 15) &
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
