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
// cell, so one direction byte per cell is stored instead of six score bytes.
//
// One wave per target.  The query is processed in strips of 64 rows; lane l
// owns row i0+l+1 and handles column j at step j+l-1, receiving H and Ix of
// the cell above from lane l-1 through one lane shift per step (a systolic
// anti-diagonal sweep).  The strip's last row is kept for the next strip in
// a per-target row buffer.  End cell: the first cell in row-major order that
// holds the maximum (sseqs/sw.py:620-637); lane 0 then follows the bytes.
#include "common.h"

namespace sseqs {

constexpr int BT_NEG = -16384;  // VERY_NEGATIVE_SCORE of sseqs/sw.py:517

__global__ __launch_bounds__(64) void k_backtrack(
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

}  // namespace sseqs

using namespace sseqs;

extern "C" size_t sseqs_backtrack_ws_bytes(int query_len, int64_t total_cells, int B) {
    (void)B;
    int64_t rows = query_len > 0 ? total_cells / query_len : 0;
    return (size_t)((total_cells + 255) & ~255LL) + (size_t)rows * 8 + 256;
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
    uint8_t* dirs = static_cast<uint8_t*>(ws);
    int2* rowbuf = reinterpret_cast<int2*>(dirs + ((total_cells + 255) & ~255LL));
    hipLaunchKernelGGL(k_backtrack, dim3((unsigned)B), dim3(64), 0, stream, query, query_len,
                       targets, t_off, t_len, cell_off, B, gap_open, gap_extend, max_align, out_q,
                       out_t, info, dirs, rowbuf, blosum62());
    return check_launch("backtrack");
}
