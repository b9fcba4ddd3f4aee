// Synthetic length-sorted protein database, generated in HBM.
//
// Not a reference function: the reference's .xbit database is a download.
// The generator follows the workload definition (i.i.d. residues from the
// BLOSUM62 background, log-normal lengths, planted homologs with
// BLOSUM62-weighted substitutions and geometric indels) with a counter-based
// integer RNG, so every residue is a pure function of (seed, sequence,
// position) and the CPU restatement produces the same bytes.
//
// Layout: sequence i has length cls_len[c] for the length class c holding
// it (classes sorted by length, cls_start = cumulative counts); its first
// residue sits at symbol cls_sym[c] + (i - cls_start[c]) * (len + 1) and a
// '@' follows every sequence.  One sequence per block of hom_period is a
// homolog of query (block * 2654435761) % nq.
#include "common.h"

namespace sseqs {

constexpr int HOM_STRIDE = 1024;  // bytes reserved per homolog in the scratch

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ inline uint32_t rnd24(uint64_t seed, uint64_t stream, uint64_t ctr) {
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL ^ mix64(stream * 0xD1B54A32D192ED03ULL + ctr));
    return (uint32_t)(h >> 40);
}
__device__ inline int draw(const int32_t* cdf, int n, uint32_t r) {
    int a = 0;
    while (a < n - 1 && (int32_t)r >= cdf[a]) ++a;
    return a;
}

__constant__ char c_aa[21] = "ARNDCQEGHILKMFPSTWYV";

struct GenCtx {
    sseqs_gen_params gp;
    const uint8_t* qcodes;
    const int64_t* q_off;
    int nq;
    const int64_t* cls_start;
    const int32_t* cls_len;
    const int64_t* cls_sym;
    int K;
    int64_t n;
    int64_t n_syms;
};

__device__ inline int class_of_seq(const GenCtx& c, int64_t i) {
    int lo = 0, hi = c.K - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c.cls_start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ inline int class_of_sym(const GenCtx& c, int64_t q) {
    int lo = 0, hi = c.K - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c.cls_sym[mid] <= q) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ inline bool homolog_of(const GenCtx& c, int64_t i, int64_t& blk) {
    const int P = c.gp.hom_period;
    blk = i / P;
    int64_t slot = (int64_t)(mix64(c.gp.seed ^ (0xA5A5A5A5ULL + (uint64_t)blk)) % (uint64_t)P);
    return c.nq > 0 && (i % P) == slot;
}
__device__ inline uint8_t bg_residue(const GenCtx& c, int64_t i, int p) {
    return (uint8_t)c_aa[draw(c.gp.bg_cdf, 20, rnd24(c.gp.seed, (uint64_t)i, (uint64_t)p))];
}

// Mutated copy of query q for homolog h.  Residue number t of the copy is
// written to out[t + shift] when 0 <= t + shift < L (out may be null to
// measure).  Returns the copy's length (capped at 4096).
__device__ int mutate(const GenCtx& c, int64_t h, const uint8_t* q, int Lq, uint8_t* out,
                      int shift, int L) {
    const sseqs_gen_params& gp = c.gp;
    const uint64_t S = (1ULL << 61) | (uint64_t)h;
    uint64_t k = 0;
    uint32_t idr = rnd24(gp.seed, S, k++);
    int32_t id = gp.id_lo + (int32_t)(((uint64_t)idr * (uint64_t)gp.id_span) >> 24);
    int pos = 0, n = 0;
    const int cap = 4096;
    while (pos < Lq && n < cap) {
        if ((int32_t)rnd24(gp.seed, S, k++) < gp.p_indel) {
            int len = 1;
            while (len < 64 && (int32_t)rnd24(gp.seed, S, k++) < gp.p_ext) ++len;
            if ((int32_t)rnd24(gp.seed, S, k++) < gp.p_ins) {
                for (int x = 0; x < len && n < cap; ++x) {
                    int a = draw(gp.bg_cdf, 20, rnd24(gp.seed, S, k++));
                    int p = n + shift;
                    if (out && p >= 0 && p < L) out[p] = (uint8_t)c_aa[a];
                    ++n;
                }
            } else {
                pos += len;
            }
            continue;
        }
        int a = q[pos];
        if ((int32_t)rnd24(gp.seed, S, k++) >= id)
            a = draw(gp.sub_cdf + 20 * a, 20, rnd24(gp.seed, S, k++));
        int p = n + shift;
        if (out && p >= 0 && p < L) out[p] = (uint8_t)c_aa[a];
        ++n;
        ++pos;
    }
    return n;
}

// Homolog of block h: background residues, then the mutated query copy
// (whole copy at a random shift when it fits, else a random window of it).
__device__ void gen_homolog(const GenCtx& c, uint8_t* __restrict__ hs, int64_t h) {
    const int P = c.gp.hom_period;
    int64_t slot = (int64_t)(mix64(c.gp.seed ^ (0xA5A5A5A5ULL + (uint64_t)h)) % (uint64_t)P);
    int64_t i = h * P + slot;
    if (i >= c.n) return;
    int cls = class_of_seq(c, i);
    int L = c.cls_len[cls];
    uint8_t* o = hs + h * HOM_STRIDE;
    for (int p = 0; p < L; ++p) o[p] = bg_residue(c, i, p);
    int qi = (int)((uint64_t)(h * 2654435761ULL) % (uint64_t)c.nq);
    const uint8_t* q = c.qcodes + c.q_off[qi];
    int Lq = (int)(c.q_off[qi + 1] - c.q_off[qi]);
    int m = mutate(c, h, q, Lq, nullptr, 0, 0);
    const uint64_t S2 = (1ULL << 62) | (uint64_t)h;
    uint32_t r = rnd24(c.gp.seed, S2, 0);
    if (m <= L) {
        int shift = (int)(((uint64_t)r * (uint64_t)(L - m + 1)) >> 24);
        mutate(c, h, q, Lq, o, shift, L);
    } else {
        int shift = (int)(((uint64_t)r * (uint64_t)(m - L + 1)) >> 24);
        mutate(c, h, q, Lq, o, -shift, L);
    }
}

// Grid-stride over homolog blocks (an HSA dispatch counts work-items in 32
// bits, so database-sized launches loop instead of one thread per unit).
__global__ __launch_bounds__(256) void k_gen_homologs(GenCtx c, uint8_t* __restrict__ hs) {
    if (c.nq <= 0) return;
    const int P = c.gp.hom_period;
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h * P < c.n;
         h += (int64_t)gridDim.x * blockDim.x)
        gen_homolog(c, hs, h);
}

// One 8-symbol group of the packed output.
__device__ void gen_group(const GenCtx& c, const uint8_t* __restrict__ hs,
                          uint8_t* __restrict__ packed, int64_t g) {
    int64_t q0 = g * 8;
    int cls = class_of_sym(c, q0);
    int L = c.cls_len[cls];
    int64_t rel = q0 - c.cls_sym[cls];
    int64_t i = c.cls_start[cls] + rel / (L + 1);
    int p = (int)(rel % (L + 1));
    int64_t blk;
    bool hom = homolog_of(c, i, blk);
    uint64_t w = 0;
    for (int k = 0; k < 8; ++k) {
        uint32_t sym = 0;  // '@' (also pads the final group)
        if (q0 + k < c.n_syms) {
            if (p < L) {
                uint8_t ch = hom ? hs[blk * HOM_STRIDE + p] : bg_residue(c, i, p);
                sym = (uint32_t)(ch - 64) & 31u;
            }
            // advance to the next symbol
            if (++p > L) {
                p = 0;
                ++i;
                if (i < c.n) {
                    if (cls + 1 < c.K && i >= c.cls_start[cls + 1]) {
                        ++cls;
                        L = c.cls_len[cls];
                    }
                    hom = homolog_of(c, i, blk);
                }
            }
        }
        w |= (uint64_t)sym << (35 - 5 * k);
    }
    int64_t nout = (c.n_syms * 5 + 7) / 8;
    for (int b = 0; b < 5; ++b) {
        int64_t o = g * 5 + b;
        if (o < nout) packed[o] = (uint8_t)(w >> (32 - 8 * b));
    }
}

__global__ __launch_bounds__(256) void k_gen_pack(GenCtx c, const uint8_t* __restrict__ hs,
                                                  uint8_t* __restrict__ packed) {
    const int64_t groups = (c.n_syms + 7) / 8;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
         g += (int64_t)gridDim.x * blockDim.x)
        gen_group(c, hs, packed, g);
}

__device__ void gen_index_one(const GenCtx& c, int64_t i, int64_t* __restrict__ off,
                              int32_t* __restrict__ len, int32_t* __restrict__ homq) {
    int cls = class_of_seq(c, i);
    int L = c.cls_len[cls];
    off[i] = c.cls_sym[cls] + (i - c.cls_start[cls]) * (int64_t)(L + 1);
    len[i] = L;
    int64_t blk;
    bool hom = homolog_of(c, i, blk);
    if (homq) homq[i] = hom ? (int32_t)((uint64_t)(blk * 2654435761ULL) % (uint64_t)c.nq) : -1;
}

__global__ __launch_bounds__(256) void k_gen_index(GenCtx c, int64_t* __restrict__ off,
                                                   int32_t* __restrict__ len,
                                                   int32_t* __restrict__ homq) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n;
         i += (int64_t)gridDim.x * blockDim.x)
        gen_index_one(c, i, off, len, homq);
}

}  // namespace sseqs

using namespace sseqs;

extern "C" size_t sseqs_gen_ws_bytes(int64_t n, int hom_period) {
    if (hom_period < 1) hom_period = 1;
    return (size_t)(ceil_div(n, hom_period) + 1) * HOM_STRIDE;
}

extern "C" int sseqs_gen_db(const sseqs_gen_params* h_params, const uint8_t* qcodes,
                            const int64_t* q_off, int nq, const int64_t* cls_start,
                            const int32_t* cls_len, const int64_t* cls_sym, int K, int64_t n,
                            int64_t n_syms, uint8_t* packed, int64_t* seq_off, int32_t* seq_len,
                            int32_t* hom_query, void* ws, size_t ws_bytes, hipStream_t stream) {
    if (!h_params) return fail(-1, "gen_db: null parameters");
    if (n <= 0 || K <= 0 || n_syms <= 0) return fail(-1, "gen_db: empty database");
    if (h_params->hom_period < 1) return fail(-1, "gen_db: hom_period must be >= 1");
    if (ws_bytes < sseqs_gen_ws_bytes(n, h_params->hom_period))
        return fail(-2, "gen_db: workspace too small");
    GenCtx c;
    c.gp = *h_params;
    c.qcodes = qcodes;
    c.q_off = q_off;
    c.nq = nq;
    c.cls_start = cls_start;
    c.cls_len = cls_len;
    c.cls_sym = cls_sym;
    c.K = K;
    c.n = n;
    c.n_syms = n_syms;
    uint8_t* hs = static_cast<uint8_t*>(ws);
    int64_t nh = ceil_div(n, h_params->hom_period);
    if (nq > 0) {
        hipLaunchKernelGGL(k_gen_homologs, dim3(grid_cap(nh, 256)), dim3(256), 0, stream, c, hs);
        if (int r = check_launch("gen homologs")) return r;
    }
    hipLaunchKernelGGL(k_gen_pack, dim3(grid_cap(ceil_div(n_syms, 8), 256)), dim3(256), 0, stream,
                       c, hs, packed);
    if (int r = check_launch("gen pack")) return r;
    hipLaunchKernelGGL(k_gen_index, dim3(grid_cap(n, 256)), dim3(256), 0, stream, c,
                       seq_off, seq_len, hom_query);
    return check_launch("gen index");
}
