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

// ---------------------------------------------------------------------------
// Database index: sequences are the maximal non-'@' runs of the stream.
// Pass 1 counts run starts per tile, pass 2 scans tile counts (one block),
// pass 3 writes (offset, length) for every run that starts in the tile.
// A run starts at q when sym(q) != '@' and (q == 0 or sym(q-1) == '@').
// ---------------------------------------------------------------------------
constexpr int IDX_TILE = 4096;  // symbols per block in passes 1 and 3

__device__ inline bool run_start(const uint8_t* p, int64_t q) {
    if (packed_sym(p, q) == 64u) return false;
    return q == 0 || packed_sym(p, q - 1) == 64u;
}

__global__ __launch_bounds__(256) void k_idx_count(const uint8_t* __restrict__ p, int64_t n,
                                                   int64_t nt, int64_t* __restrict__ tile_cnt) {
    __shared__ int red[256];
    for (int64_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const int64_t base = tile * IDX_TILE;
        int c = 0;
        for (int i = threadIdx.x; i < IDX_TILE; i += 256) {
            int64_t q = base + i;
            if (q < n && run_start(p, q)) ++c;
        }
        red[threadIdx.x] = c;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) tile_cnt[tile] = red[0];
        __syncthreads();
    }
}

// Exclusive scan of ntile counts in one block of 1024 threads; total -> *total.
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

__global__ __launch_bounds__(256) void k_idx_write(const uint8_t* __restrict__ p, int64_t n,
                                                   int64_t nt, const int64_t* __restrict__ tile_base,
                                                   int64_t* __restrict__ off,
                                                   int32_t* __restrict__ len, int64_t cap) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const int64_t base = tile * IDX_TILE;
    int64_t out = tile_base[tile];
    for (int i0 = 0; i0 < IDX_TILE; i0 += 256) {
        int64_t q = base + i0 + threadIdx.x;
        bool st = q < n && run_start(p, q);
        uint64_t m = __ballot(st);
        int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wid] = __popcll(m);
        __syncthreads();
        int wb = 0;
        for (int w = 0; w < wid; ++w) wb += wsum[w];
        int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (st) {
            int64_t slot = out + wb + before;
            if (slot < cap) {
                int64_t e = q;
                while (e < n && packed_sym(p, e) != 64u) ++e;
                off[slot] = q;
                len[slot] = (int32_t)(e - q);
            }
        }
        out += tot;
        __syncthreads();
    }
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
    hipLaunchKernelGGL(k_unpack, dim3(grid_cap(ceil_div(n_syms, 8), 256)), dim3(256), 0, stream,
                       d_in + lo, d_out, n_syms);
    return check_launch("decompress");
}

extern "C" int compress_launcher(const uint8_t* d_in, uint8_t* d_out, int64_t n_syms,
                                 hipStream_t stream) {
    if (n_syms < 0) return fail(-1, "compress: negative size");
    if (n_syms == 0) return 0;
    if (!d_in || !d_out) return fail(-1, "compress: null pointer");
    hipLaunchKernelGGL(k_pack, dim3(grid_cap(ceil_div(n_syms, 8), 256)), dim3(256), 0, stream, d_in,
                       d_out, n_syms);
    return check_launch("compress");
}

extern "C" size_t sseqs_db_index_ws_bytes(int64_t n_syms) {
    return (size_t)(ceil_div(n_syms, IDX_TILE) + 1) * sizeof(int64_t);
}

extern "C" int sseqs_db_index(const uint8_t* packed, int64_t n_syms, int64_t* seq_off,
                              int32_t* seq_len, int64_t cap, int64_t* d_nseq, void* ws,
                              size_t ws_bytes, hipStream_t stream) {
    if (n_syms <= 0) return fail(-1, "db_index: empty database");
    if (ws_bytes < sseqs_db_index_ws_bytes(n_syms)) return fail(-2, "db_index: workspace too small");
    int64_t nt = ceil_div(n_syms, IDX_TILE);
    int64_t* tiles = static_cast<int64_t*>(ws);
    hipLaunchKernelGGL(k_idx_count, dim3(grid_cap(nt, 1)), dim3(256), 0, stream, packed, n_syms, nt,
                       tiles);
    if (int r = check_launch("db_index count")) return r;
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, stream, tiles, nt, d_nseq);
    if (int r = check_launch("db_index scan")) return r;
    hipLaunchKernelGGL(k_idx_write, dim3(grid_cap(nt, 1)), dim3(256), 0, stream, packed, n_syms, nt,
                       tiles, seq_off, seq_len, cap);
    return check_launch("db_index write");
}
