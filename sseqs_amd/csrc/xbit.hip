// 5-bit database codec and database index (delimiter scan).
//
// Format (sseqs/xbit.cu:8-62): symbol = byte - 64 (5 bits); 8 symbols form a
// 40-bit big-endian group of 5 bytes, symbol j of the group in bits
// 39-5j..35-5j, i.e. the stream is MSB-first and symbol q starts at bit 5q.
#include "common.h"

namespace sseqs {

// One 8-symbol group per thread and step: 5 byte loads (the group is not
// aligned), one 8-byte store when the group is whole.  Grid-stride: an HSA
// dispatch counts work-items in 32 bits, so a database-sized stream
// (> 2^32 groups) must not be one thread per group.
__global__ __launch_bounds__(256) void k_unpack(const uint8_t* __restrict__ in,
                                                uint8_t* __restrict__ out, int64_t n) {
    const int64_t groups = (n + 7) / 8;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q0 = g * 8;
        const uint8_t* p = in + g * 5;
        uint64_t w = ((uint64_t)p[0] << 32) | ((uint64_t)p[1] << 24) | ((uint64_t)p[2] << 16) |
                     ((uint64_t)p[3] << 8) | (uint64_t)p[4];
        uint64_t o = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) o |= (((w >> (35 - 5 * k)) & 31u) + 64u) << (8 * k);
        if (q0 + 8 <= n && ((reinterpret_cast<uintptr_t>(out + q0) & 7) == 0)) {
            *reinterpret_cast<uint64_t*>(out + q0) = o;
        } else {
            for (int k = 0; k < 8 && q0 + k < n; ++k) out[q0 + k] = (uint8_t)(o >> (8 * k));
        }
    }
}

// One group per thread and step; a short final group is padded with '@'
// (0) and only bytes below ceil(5n/8) are written.
__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ in,
                                              uint8_t* __restrict__ out, int64_t n) {
    const int64_t groups = (n + 7) / 8;
    const int64_t nout = (n * 5 + 7) / 8;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q0 = g * 8;
        uint64_t w = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint64_t s = (q0 + k < n) ? (uint64_t)((uint8_t)(in[q0 + k] - 64) & 31u) : 0u;
            w |= s << (35 - 5 * k);
        }
#pragma unroll
        for (int b = 0; b < 5; ++b) {
            int64_t o = g * 5 + b;
            if (o < nout) out[o] = (uint8_t)(w >> (32 - 8 * b));
        }
    }
}

// Wide forms for aligned streams: 32 symbols (four groups, 20 packed bytes)
// per chunk, moved as two 16-byte and five 4-byte accesses.  Each thread
// loads WIDE_U chunks (at a grid stride, clamped into range) before it
// converts and stores any, so that enough bytes are in flight per wave.
// They cover the first n/32 whole chunks; the kernels above take the tail.
constexpr int WIDE_U = 4;

__global__ __launch_bounds__(256) void k_unpack32(const uint32_t* __restrict__ in,
                                                  uint4* __restrict__ out, int64_t chunks) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c0 < chunks;
         c0 += WIDE_U * stride) {
        uint32_t d[WIDE_U][5];
#pragma unroll
        for (int u = 0; u < WIDE_U; ++u) {
            const int64_t c = min(c0 + u * stride, chunks - 1);
#pragma unroll
            for (int i = 0; i < 5; ++i) d[u][i] = __builtin_bswap32(in[5 * c + i]);
        }
#pragma unroll
        for (int u = 0; u < WIDE_U; ++u) {
            const int64_t c = c0 + u * stride;
            uint32_t o[8];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                // group g: bits 40g .. 40g+39 of the 160-bit big-endian chunk
                const int b = 40 * g, i = b >> 5, sh = b & 31;
                const uint64_t hi = ((uint64_t)d[u][i] << 32) | d[u][i + 1];
                const uint64_t w = (hi << sh) >> 24;  // the group in bits 39..0
                uint64_t v = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) v |= (((w >> (35 - 5 * k)) & 31u) + 64u) << (8 * k);
                o[2 * g] = (uint32_t)v;
                o[2 * g + 1] = (uint32_t)(v >> 32);
            }
            if (c < chunks) {
                out[2 * c] = make_uint4(o[0], o[1], o[2], o[3]);
                out[2 * c + 1] = make_uint4(o[4], o[5], o[6], o[7]);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_pack32(const uint4* __restrict__ in,
                                                uint32_t* __restrict__ out, int64_t chunks) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c0 < chunks;
         c0 += WIDE_U * stride) {
        uint4 a[WIDE_U][2];
#pragma unroll
        for (int u = 0; u < WIDE_U; ++u) {
            const int64_t c = min(c0 + u * stride, chunks - 1);
            a[u][0] = in[2 * c];
            a[u][1] = in[2 * c + 1];
        }
#pragma unroll
        for (int u = 0; u < WIDE_U; ++u) {
            const int64_t c = c0 + u * stride;
            const uint32_t x[8] = {a[u][0].x, a[u][0].y, a[u][0].z, a[u][0].w,
                                   a[u][1].x, a[u][1].y, a[u][1].z, a[u][1].w};
            uint32_t d[5] = {0u, 0u, 0u, 0u, 0u};  // big-endian chunk
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                uint64_t w = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t byte = (x[2 * g + (k >> 2)] >> (8 * (k & 3))) & 255u;
                    w |= (uint64_t)((byte - 64u) & 31u) << (35 - 5 * k);
                }
                const int bit = 40 * g, i = bit >> 5, sh = bit & 31;
                // place the 40 bits at chunk bits bit..bit+39 (from the top);
                // sh <= 24, so they stay inside words i and i + 1
                const uint64_t top = w << 24;  // group in bits 63..24
                d[i] |= (uint32_t)((top >> 32) >> sh);
                d[i + 1] |= (uint32_t)((top << (32 - sh)) >> 32);
            }
            if (c < chunks) {
#pragma unroll
                for (int i = 0; i < 5; ++i) out[5 * c + i] = __builtin_bswap32(d[i]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Database index: sequences are the maximal non-'@' runs of the stream.
// The stream is cut into tiles of IDX_TILE symbols, 64 per thread.
//   count: per tile, the number of runs that start in it and the position of
//          its first '@' (-1 for none);
//   scan:  exclusive scan of the tile counts (three launches);
//   write: (offset, length) of every run that starts in the tile.  A run's
//          end is found from the same bit masks when it lies in the tile;
//          the one run that may leave the tile looks its end up in the
//          first-'@' table of the tiles that follow.
// A run starts at q when sym(q) != '@' and (q == 0 or sym(q-1) == '@').
// ---------------------------------------------------------------------------
constexpr int IDX_TILE = 16384;  // symbols per tile: 256 threads x 64
constexpr int SCAN_PER = 8192;   // tile counts per block of the scan: 1024 x 8

// Bit j of the result: symbol q0 + j (q0 a multiple of 64) is a residue,
// i.e. not '@' and below n.  The 64 symbols are the 40 bytes at 5*q0/8, read
// as ten aligned dwords (within the stream's tail pad).
__device__ inline uint64_t residue_mask64(const uint8_t* __restrict__ p, int64_t n, int64_t q0) {
    if (q0 >= n) return 0ull;
    const uint32_t* d = reinterpret_cast<const uint32_t*>(p + (q0 >> 3) * 5);
    uint32_t D[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) D[i] = __builtin_bswap32(d[i]);
    uint64_t m = 0;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        // group h: bits 40h .. 40h+39 of the 320-bit big-endian window
        const int b = 40 * h, i = b >> 5, sh = b & 31;
        const uint64_t w = ((((uint64_t)D[i] << 32) | D[i + 1]) << sh) >> 24;
        // OR of each 5-bit field lands on the field's lowest bit
        const uint64_t f = w | (w >> 1) | (w >> 2) | (w >> 3) | (w >> 4);
#pragma unroll
        for (int k = 0; k < 8; ++k) m |= ((f >> (35 - 5 * k)) & 1ull) << (8 * h + k);
    }
    if (q0 + 64 > n) m &= (1ull << (uint32_t)(n - q0)) - 1ull;
    return m;
}

__device__ inline bool residue_at(const uint8_t* p, int64_t n, int64_t q) {
    return q >= 0 && q < n && packed_sym(p, q) != 64u;
}

__device__ inline int wave_sum_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline int wave_min_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ __launch_bounds__(256) void k_idx_count(const uint8_t* __restrict__ p, int64_t n,
                                                   int64_t nt, int64_t* __restrict__ tile_cnt,
                                                   int64_t* __restrict__ first_at) {
    __shared__ uint64_t mk[256];
    __shared__ int red[2][4];
    __shared__ int prev_res;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int64_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const int64_t base = tile * IDX_TILE;
        const uint64_t m = residue_mask64(p, n, base + 64 * tid);
        mk[tid] = m;
        if (tid == 0) prev_res = residue_at(p, n, base - 1);
        __syncthreads();
        const uint64_t prev = tid ? mk[tid - 1] >> 63 : (uint64_t)prev_res;
        const uint64_t starts = m & ~((m << 1) | prev);
        // first '@' below n: a zero bit of m at a position < n
        const int64_t lim = n - (base + 64 * tid);
        const uint64_t valid = lim >= 64 ? ~0ull : lim <= 0 ? 0ull : (1ull << (uint32_t)lim) - 1ull;
        const uint64_t at = ~m & valid;
        const int c = wave_sum_i32(__popcll(starts));
        const int z = wave_min_i32(at ? 64 * tid + __builtin_ctzll(at) : IDX_TILE);
        if (lane == 0) {
            red[0][wid] = c;
            red[1][wid] = z;
        }
        __syncthreads();
        if (tid == 0) {
            tile_cnt[tile] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
            const int zz = min(min(red[1][0], red[1][1]), min(red[1][2], red[1][3]));
            first_at[tile] = zz < IDX_TILE ? base + zz : -1;
        }
        __syncthreads();
    }
}

// Scan of the tile counts: per-block sums, a one-block exclusive scan of the
// sums (k_scan_i64), then each block's exclusive scan with its offset.
// Exclusive scan of nt values in one block of 1024 threads; total -> *total.
__global__ __launch_bounds__(1024) void k_scan_i64(int64_t* __restrict__ v, int64_t nt,
                                                   int64_t* __restrict__ total) {
    __shared__ int64_t part[1024];
    int64_t per = (nt + 1023) / 1024;
    int64_t lo = threadIdx.x * per, hi = min(nt, lo + per);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t t = (threadIdx.x >= off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t run = part[threadIdx.x] - s;
    for (int64_t i = lo; i < hi; ++i) {
        int64_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (threadIdx.x == 1023) *total = part[1023];
}

__global__ __launch_bounds__(1024) void k_scan_up(const int64_t* __restrict__ v, int64_t nt,
                                                  int64_t* __restrict__ part) {
    __shared__ int64_t red[16];
    const int tid = threadIdx.x;
    for (int64_t b = blockIdx.x; b * SCAN_PER < nt; b += gridDim.x) {
        const int64_t lo = b * SCAN_PER + 8 * (int64_t)tid;
        int64_t s = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) s += lo + k < nt ? v[lo + k] : 0;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            int64_t t = 0;
            for (int w = 0; w < 16; ++w) t += red[w];
            part[b] = t;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_scan_down(int64_t* __restrict__ v, int64_t nt,
                                                    const int64_t* __restrict__ part) {
    __shared__ int64_t wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int64_t b = blockIdx.x; b * SCAN_PER < nt; b += gridDim.x) {
        const int64_t lo = b * SCAN_PER + 8 * (int64_t)tid;
        int64_t x[8], s = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            x[k] = lo + k < nt ? v[lo + k] : 0;
            s += x[k];
        }
        int64_t inc = s;  // inclusive scan over the wave
        for (int off = 1; off < 64; off <<= 1) {
            int64_t t = __shfl_up(inc, off, 64);
            if (lane >= off) inc += t;
        }
        if (lane == 63) wsum[wid] = inc;
        __syncthreads();
        int64_t run = part[b] + inc - s;
        for (int w = 0; w < wid; ++w) run += wsum[w];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (lo + k < nt) v[lo + k] = run;
            run += x[k];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_idx_write(const uint8_t* __restrict__ p, int64_t n,
                                                   int64_t nt, const int64_t* __restrict__ tile_base,
                                                   const int64_t* __restrict__ first_at,
                                                   int64_t* __restrict__ off,
                                                   int32_t* __restrict__ len, int64_t cap) {
    __shared__ uint64_t mk[256];
    __shared__ uint16_t st[IDX_TILE / 2];  // local start of the tile's j-th run
    __shared__ int wsum[4];
    __shared__ int edge[2];                // residue just before / just after the tile
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int64_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const int64_t base = tile * IDX_TILE;
        const int64_t out = tile_base[tile];
        const uint64_t m = residue_mask64(p, n, base + 64 * tid);
        mk[tid] = m;
        if (tid == 0) {
            edge[0] = residue_at(p, n, base - 1);
            edge[1] = residue_at(p, n, base + IDX_TILE);
        }
        __syncthreads();
        const uint64_t prev = tid ? mk[tid - 1] >> 63 : (uint64_t)edge[0];
        const uint64_t next = tid < 255 ? mk[tid + 1] & 1ull : (uint64_t)edge[1];
        const uint64_t starts = m & ~((m << 1) | prev);
        const uint64_t ends = m & ~((m >> 1) | (next << 63));
        const int c = __popcll(starts);
        int inc = c;  // inclusive scan of start counts over the wave
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[wid] = inc;
        __syncthreads();
        int ex = inc - c;
        for (int w = 0; w < wid; ++w) ex += wsum[w];
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        {
            uint64_t s = starts;
            int j = ex;
            while (s) {
                const int b = __builtin_ctzll(s);
                s &= s - 1;
                st[j] = (uint16_t)(64 * tid + b);
                if (out + j < cap) off[out + j] = base + 64 * tid + b;
                ++j;
            }
        }
        __syncthreads();
        {
            // an end at local e closes the latest run started at or before e;
            // none started in the tile means the run began in an earlier tile,
            // whose block writes its length
            uint64_t e = ends;
            while (e) {
                const int b = __builtin_ctzll(e);
                e &= e - 1;
                // starts at or below bit b (2 << 63 wraps to 0: all bits)
                const int j = ex + __popcll(starts & ((2ull << b) - 1ull)) - 1;
                if (j >= 0 && out + j < cap) len[out + j] = 64 * tid + b + 1 - st[j];
            }
        }
        if (tid == 255 && total > 0 && (m >> 63) && next) {
            // the tile's last run continues past it: it ends at the first '@'
            // of a later tile, or at the end of the stream
            const int j = total - 1;
            if (out + j < cap) {
                int64_t t = tile + 1;
                while (t < nt && first_at[t] < 0) ++t;
                const int64_t e = t < nt ? first_at[t] : n;
                len[out + j] = (int32_t)(e - (base + st[j]));
            }
        }
        __syncthreads();
    }
}

// Hit gather (sseqs/msa.py:349-370): row i of `out` receives the first
// min(len, W) residues of database sequence ids[i] as ASCII.
__global__ __launch_bounds__(64) void k_gather(const uint8_t* __restrict__ p,
                                               const int64_t* __restrict__ off,
                                               const int32_t* __restrict__ len,
                                               const int64_t* __restrict__ ids, int64_t n, int W,
                                               uint8_t* __restrict__ out,
                                               int32_t* __restrict__ out_len) {
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        int64_t s = ids[i];
        int L = min(len[s], W);
        int64_t o = off[s];
        for (int q = threadIdx.x; q < L; q += 64) out[i * W + q] = (uint8_t)packed_sym(p, o + q);
        if (threadIdx.x == 0) out_len[i] = L;
    }
}

}  // namespace sseqs

using namespace sseqs;

extern "C" int sseqs_db_gather(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                               const int64_t* ids, int64_t n, int width, uint8_t* out,
                               int32_t* out_len, hipStream_t stream) {
    if (n < 0 || width < 1) return fail(-1, "db_gather: bad size");
    if (n == 0) return 0;
    unsigned grid = (unsigned)(n < 65536 ? n : 65536);
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(64), 0, stream, packed, seq_off, seq_len, ids, n,
                       width, out, out_len);
    return check_launch("db_gather");
}

extern "C" int decompress_launcher(const uint8_t* d_in, uint8_t* d_out, int64_t n_syms,
                                   int64_t lo, hipStream_t stream) {
    if (n_syms < 0 || lo < 0) return fail(-1, "decompress: negative size or offset");
    if (n_syms == 0) return 0;
    if (!d_in || !d_out) return fail(-1, "decompress: null pointer");
    const uint8_t* src = d_in + lo;
    int64_t chunks = 0;
    if ((reinterpret_cast<uintptr_t>(src) & 3) == 0 && (reinterpret_cast<uintptr_t>(d_out) & 15) == 0)
        chunks = n_syms / 32;
    if (chunks) {
        hipLaunchKernelGGL(k_unpack32, dim3(grid_cap(chunks, 256 * WIDE_U)), dim3(256), 0, stream,
                           reinterpret_cast<const uint32_t*>(src), reinterpret_cast<uint4*>(d_out),
                           chunks);
        if (int r = check_launch("decompress")) return r;
    }
    const int64_t rest = n_syms - 32 * chunks;
    if (rest == 0) return 0;
    hipLaunchKernelGGL(k_unpack, dim3(grid_cap(ceil_div(rest, 8), 256)), dim3(256), 0, stream,
                       src + 20 * chunks, d_out + 32 * chunks, rest);
    return check_launch("decompress");
}

extern "C" int compress_launcher(const uint8_t* d_in, uint8_t* d_out, int64_t n_syms,
                                 hipStream_t stream) {
    if (n_syms < 0) return fail(-1, "compress: negative size");
    if (n_syms == 0) return 0;
    if (!d_in || !d_out) return fail(-1, "compress: null pointer");
    int64_t chunks = 0;
    if ((reinterpret_cast<uintptr_t>(d_in) & 15) == 0 && (reinterpret_cast<uintptr_t>(d_out) & 3) == 0)
        chunks = n_syms / 32;
    if (chunks) {
        hipLaunchKernelGGL(k_pack32, dim3(grid_cap(chunks, 256 * WIDE_U)), dim3(256), 0, stream,
                           reinterpret_cast<const uint4*>(d_in), reinterpret_cast<uint32_t*>(d_out),
                           chunks);
        if (int r = check_launch("compress")) return r;
    }
    const int64_t rest = n_syms - 32 * chunks;
    if (rest == 0) return 0;
    hipLaunchKernelGGL(k_pack, dim3(grid_cap(ceil_div(rest, 8), 256)), dim3(256), 0, stream,
                       d_in + 32 * chunks, d_out + 20 * chunks, rest);
    return check_launch("compress");
}

extern "C" size_t sseqs_db_index_ws_bytes(int64_t n_syms) {
    const int64_t nt = ceil_div(n_syms > 0 ? n_syms : 1, IDX_TILE);
    return (size_t)(2 * nt + ceil_div(nt, SCAN_PER) + 1) * sizeof(int64_t);
}

extern "C" int sseqs_db_index(const uint8_t* packed, int64_t n_syms, int64_t* seq_off,
                              int32_t* seq_len, int64_t cap, int64_t* d_nseq, void* ws,
                              size_t ws_bytes, hipStream_t stream) {
    if (n_syms <= 0) return fail(-1, "db_index: empty database");
    if (!packed || !d_nseq || !ws || (cap > 0 && (!seq_off || !seq_len)))
        return fail(-1, "db_index: null pointer");
    if (reinterpret_cast<uintptr_t>(packed) & 3) return fail(-1, "db_index: packed must be 4-byte aligned");
    if (ws_bytes < sseqs_db_index_ws_bytes(n_syms)) return fail(-2, "db_index: workspace too small");
    const int64_t nt = ceil_div(n_syms, IDX_TILE), nb = ceil_div(nt, SCAN_PER);
    int64_t* tiles = static_cast<int64_t*>(ws);
    int64_t* first_at = tiles + nt;
    int64_t* part = first_at + nt;
    hipLaunchKernelGGL(k_idx_count, dim3(grid_cap(nt, 1)), dim3(256), 0, stream, packed, n_syms, nt,
                       tiles, first_at);
    if (int r = check_launch("db_index count")) return r;
    hipLaunchKernelGGL(k_scan_up, dim3(grid_cap(nb, 1)), dim3(1024), 0, stream, tiles, nt, part);
    if (int r = check_launch("db_index scan")) return r;
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, stream, part, nb, d_nseq);
    if (int r = check_launch("db_index scan")) return r;
    hipLaunchKernelGGL(k_scan_down, dim3(grid_cap(nb, 1)), dim3(1024), 0, stream, tiles, nt, part);
    if (int r = check_launch("db_index scan")) return r;
    if (cap <= 0) return 0;
    hipLaunchKernelGGL(k_idx_write, dim3(grid_cap(nt, 1)), dim3(256), 0, stream, packed, n_syms, nt,
                       tiles, first_at, seq_off, seq_len, cap);
    return check_launch("db_index write");
}
