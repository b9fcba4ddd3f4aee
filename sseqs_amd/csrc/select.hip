// Deterministic device top-k.
//
// Used wherever the reference calls torch.topk / torch.sort on the search
// path (sseqs/msa.py:297, 307, 324, 335).  Scores become order-preserving
// 32-bit keys; three histogram passes (11 + 11 + 10 bits, most significant
// first; two, 11 + 8, for fp16 scores, whose keys end in 13 cleared bits) find
// the k-th largest key T and how many keys equal to T are taken;
// a compaction pass then writes every index with key > T plus the first
// ones with key == T, in ascending index order.  The optional sort orders
// them by (key desc, index asc) with a bitonic sort in LDS.  The tie rule
// (smaller index first) is the build's definition; torch.topk leaves it open.
#include <mutex>
#include <string>

#include "common.h"

namespace sseqs {

constexpr int SEL_BINS = 2048;
constexpr int SEL_TILE = 16384;     // elements per block in count/write passes
constexpr int SEL_SORT_MAX = 16384;

struct SelState {
    uint32_t prefix, mask, k_rem, pad;
};

__device__ inline uint32_t key_f32(float x) {
    if (x != x) return 0u;  // NaN ranks lowest
    uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

template <int DT>
__device__ inline uint32_t load_key(const void* s, int64_t i) {
    if constexpr (DT == SSEQS_F32) return key_f32(static_cast<const float*>(s)[i]);
    // an fp16 value as fp32 has 13 constant low bits (zeros, ones once
    // complemented for a negative value): clearing them keeps the order
    if constexpr (DT == SSEQS_F16)
        return key_f32((float)static_cast<const _Float16*>(s)[i]) & ~0x1FFFu;
    return (uint32_t) static_cast<const int32_t*>(s)[i] ^ 0x80000000u;
}

__global__ void k_sel_init(uint32_t* hist, SelState* st, uint32_t k) {
    for (int i = threadIdx.x; i < SEL_BINS; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x == 0) {
        st->prefix = 0;
        st->mask = 0;
        st->k_rem = k;
        st->pad = 0;
    }
}

template <int DT>
__global__ __launch_bounds__(256) void k_sel_hist(const void* s, int64_t n,
                                                  const SelState* st, uint32_t* hist, int shift,
                                                  int bits) {
    __shared__ uint32_t lh[SEL_BINS];
    for (int i = threadIdx.x; i < SEL_BINS; i += 256) lh[i] = 0;
    __syncthreads();
    const uint32_t prefix = st->prefix, mask = st->mask, bm = (1u << bits) - 1u;
    // four unconditional (index-clamped) loads in flight per thread
    for (int64_t i0 = (int64_t)blockIdx.x * 1024; i0 < n; i0 += (int64_t)gridDim.x * 1024) {
        uint32_t key[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) key[e] = load_key<DT>(s, min(i0 + e * 256 + threadIdx.x, n - 1));
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (i0 + e * 256 + threadIdx.x < n && (key[e] & mask) == prefix)
                atomicAdd(&lh[(key[e] >> shift) & bm], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SEL_BINS; i += 256)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// One block: fix the next digit of T, clear hist.  The digit is the bin b
// (from the top) where the keys above it fall short of k_rem and the keys
// above plus those in it do not; the counts above every bin come from one
// block-wide scan of the bins in descending order.
__global__ __launch_bounds__(1024) void k_sel_pick(SelState* st, uint32_t* hist, int shift,
                                                   int bits) {
    __shared__ uint32_t wsum[16];
    __shared__ int pick_b;
    __shared__ uint32_t pick_above, cnt0;
    const int nb = 1 << bits;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    // thread t holds bins hi = nb-1-2t and lo = hi-1, so ascending t walks
    // the bins from the top
    const int hi = nb - 1 - 2 * t, lo = hi - 1;
    const uint32_t c_hi = hi >= 0 ? hist[hi] : 0u, c_lo = lo >= 0 ? hist[lo] : 0u;
    if (t == 0) pick_b = -1;
    __syncthreads();
    if (hi >= 0) hist[hi] = 0u;
    if (lo >= 0) hist[lo] = 0u;
    if (lo == 0) cnt0 = c_lo;
    const uint32_t c = c_hi + c_lo;
    uint32_t inc = c;  // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t above = inc - c;  // keys in the bins above hi
    for (int w = 0; w < wid; ++w) above += wsum[w];
    const uint32_t need = st->k_rem;
    if (hi >= 0 && above < need && need <= above + c_hi) {
        pick_b = hi;
        pick_above = above;
    } else if (lo >= 0 && above + c_hi < need && need <= above + c) {
        pick_b = lo;
        pick_above = above + c_hi;
    }
    __syncthreads();
    if (t == 0) {
        int b = pick_b;
        uint32_t ab = pick_above;
        if (b < 0) {  // fewer keys than k_rem: bin 0, with every other bin above it
            b = 0;
            ab = 0;
            for (int w = 0; w < 16; ++w) ab += wsum[w];
            ab -= cnt0;
        }
        st->prefix |= (uint32_t)b << shift;
        st->mask |= ((1u << bits) - 1u) << shift;
        st->k_rem = need - ab;
    }
}

template <int DT>
__global__ __launch_bounds__(256) void k_sel_count(const void* s, int64_t n,
                                                   int64_t nb, const SelState* st, int64_t* cgt,
                                                   int64_t* ceq) {
    const uint32_t T = st->prefix;
    __shared__ int rg[256], re[256];
    for (int64_t tile = blockIdx.x; tile < nb; tile += gridDim.x) {
    const int64_t base = tile * SEL_TILE;
    int gt = 0, eq = 0;
#pragma unroll 4
    for (int j = threadIdx.x; j < SEL_TILE; j += 256) {
        const int64_t i = base + j;
        const uint32_t key = load_key<DT>(s, min(i, n - 1));
        if (i < n) {
            gt += key > T;
            eq += key == T;
        }
    }
    rg[threadIdx.x] = gt;
    re[threadIdx.x] = eq;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rg[threadIdx.x] += rg[threadIdx.x + o];
            re[threadIdx.x] += re[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cgt[tile] = rg[0];
        ceq[tile] = re[0];
    }
    __syncthreads();
    }
}

// Exclusive scan of two arrays of nb counts, one block.
__global__ __launch_bounds__(1024) void k_sel_scan(int64_t* a, int64_t* b, int64_t nb) {
    __shared__ int64_t pa[1024], pb[1024];
    int64_t per = (nb + 1023) / 1024;
    int64_t lo = threadIdx.x * per, hi = min(nb, lo + per);
    int64_t sa = 0, sb = 0;
    for (int64_t i = lo; i < hi; ++i) {
        sa += a[i];
        sb += b[i];
    }
    pa[threadIdx.x] = sa;
    pb[threadIdx.x] = sb;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t ta = threadIdx.x >= off ? pa[threadIdx.x - off] : 0;
        int64_t tb = threadIdx.x >= off ? pb[threadIdx.x - off] : 0;
        __syncthreads();
        pa[threadIdx.x] += ta;
        pb[threadIdx.x] += tb;
        __syncthreads();
    }
    int64_t ra = pa[threadIdx.x] - sa, rb = pb[threadIdx.x] - sb;
    for (int64_t i = lo; i < hi; ++i) {
        int64_t xa = a[i], xb = b[i];
        a[i] = ra;
        b[i] = rb;
        ra += xa;
        rb += xb;
    }
}

template <int DT>
__global__ __launch_bounds__(256) void k_sel_write(const void* s, int64_t n,
                                                   int64_t nb, const SelState* st,
                                                   const int64_t* bgt, const int64_t* beq,
                                                   int64_t k, int64_t* out) {
    const uint32_t T = st->prefix;
    const int64_t keq = st->k_rem;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    __shared__ int wg[4][4], we[4][4];  // [step][wave]
    for (int64_t tile = blockIdx.x; tile < nb; tile += gridDim.x) {
        int64_t gt_run = bgt[tile], eq_run = beq[tile];
        const int64_t base = tile * SEL_TILE;
        // four steps of 256 elements per barrier pair; loads first, in flight together
        for (int j0 = 0; j0 < SEL_TILE; j0 += 1024) {
            uint32_t key[4];
            uint64_t mg[4], me[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                key[t] = load_key<DT>(s, min(base + j0 + t * 256 + threadIdx.x, n - 1));
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bool in = base + j0 + t * 256 + threadIdx.x < n;
                mg[t] = __ballot(in && key[t] > T);
                me[t] = __ballot(in && key[t] == T);
                if (lane == 0) {
                    wg[t][wid] = __popcll(mg[t]);
                    we[t][wid] = __popcll(me[t]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                int og = 0, oe = 0, tg = 0, te = 0;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    if (w < wid) {
                        og += wg[t][w];
                        oe += we[t][w];
                    }
                    tg += wg[t][w];
                    te += we[t][w];
                }
                const bool g = (mg[t] >> lane) & 1ull, e = (me[t] >> lane) & 1ull;
                const int64_t gb = gt_run + og + __popcll(mg[t] & below);  // keys > T before i
                const int64_t eb = eq_run + oe + __popcll(me[t] & below);  // keys == T before i
                const int64_t slot = gb + min(eb, keq);
                if ((g || (e && eb < keq)) && slot < k) out[slot] = base + j0 + t * 256 + threadIdx.x;
                gt_run += tg;
                eq_run += te;
            }
            __syncthreads();
        }
    }
}

// Sort the k selected indices by (key desc, index asc); k <= SEL_SORT_MAX.
// ---- fp16 scores, eight per thread --------------------------------------
//
// The passes above give each thread one 2-byte score per load.  For fp16
// scores at a 16-byte aligned address these variants load eight at once
// (one 16-byte load per thread), and the write pass places a step of 2048
// scores with one block barrier (a wave scan of the threads' packed counts)
// instead of two barriers per 1024.  Same keys, same T, same output.

// Keys of scores i .. i+7 (entries at or past n are flagged invalid).
__device__ inline uint32_t keys8_f16(const _Float16* s, int64_t i, int64_t n, uint32_t key[8]) {
    if (i + 8 <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(s + i);
        const _Float16* e = reinterpret_cast<const _Float16*>(&v);
#pragma unroll
        for (int x = 0; x < 8; ++x) key[x] = key_f32((float)e[x]) & ~0x1FFFu;
        return 0xFFu;
    }
    uint32_t valid = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const bool in = i + x < n;
        key[x] = in ? key_f32((float)s[i + x]) & ~0x1FFFu : 0u;
        valid |= (uint32_t)in << x;
    }
    return valid;
}

__global__ __launch_bounds__(256) void k_sel_hist_h8(const _Float16* __restrict__ s, int64_t n,
                                                     const SelState* st, uint32_t* hist, int shift,
                                                     int bits) {
    __shared__ uint32_t lh[SEL_BINS];
    for (int i = threadIdx.x; i < SEL_BINS; i += 256) lh[i] = 0;
    __syncthreads();
    const uint32_t prefix = st->prefix, mask = st->mask, bm = (1u << bits) - 1u;
    for (int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 8; i < n;
         i += (int64_t)gridDim.x * 2048) {
        uint32_t key[8];
        const uint32_t valid = keys8_f16(s, i, n, key);
        // neighbouring scores often share a bin: one LDS atomic per run
        uint32_t run_bin = 0, run = 0;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            if (!((valid >> x) & 1u) || (key[x] & mask) != prefix) continue;
            const uint32_t b = (key[x] >> shift) & bm;
            if (run && b == run_bin) {
                ++run;
            } else {
                if (run) atomicAdd(&lh[run_bin], run);
                run_bin = b;
                run = 1;
            }
        }
        if (run) atomicAdd(&lh[run_bin], run);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SEL_BINS; i += 256)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

__global__ __launch_bounds__(256) void k_sel_count_h8(const _Float16* __restrict__ s, int64_t n,
                                                      int64_t nb, const SelState* st,
                                                      int64_t* cgt, int64_t* ceq) {
    const uint32_t T = st->prefix;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __shared__ int rg[4], re[4];
    for (int64_t tile = blockIdx.x; tile < nb; tile += gridDim.x) {
        const int64_t base = tile * SEL_TILE;
        int gt = 0, eq = 0;
#pragma unroll 2
        for (int j = 8 * threadIdx.x; j < SEL_TILE; j += 2048) {
            uint32_t key[8];
            const uint32_t valid = keys8_f16(s, base + j, n, key);
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const bool in = (valid >> x) & 1u;
                gt += in && key[x] > T;
                eq += in && key[x] == T;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            gt += __shfl_xor(gt, o, 64);
            eq += __shfl_xor(eq, o, 64);
        }
        if (lane == 0) {
            rg[wid] = gt;
            re[wid] = eq;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            cgt[tile] = rg[0] + rg[1] + rg[2] + rg[3];
            ceq[tile] = re[0] + re[1] + re[2] + re[3];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_sel_write_h8(const _Float16* __restrict__ s, int64_t n,
                                                      int64_t nb, const SelState* st,
                                                      const int64_t* bgt, const int64_t* beq,
                                                      int64_t k, int64_t* out) {
    const uint32_t T = st->prefix;
    const int64_t keq = st->k_rem;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // wave totals of a step, (count > T) << 16 | (count == T), by step parity:
    // a wave writes a step's slot only after every wave has passed the
    // previous step's barrier, i.e. has read the slot two steps back
    __shared__ uint32_t wt[2][4];
    int par = 0;
    for (int64_t tile = blockIdx.x; tile < nb; tile += gridDim.x) {
        int64_t gt_run = bgt[tile], eq_run = beq[tile];
        const int64_t base = tile * SEL_TILE;
        for (int j0 = 0; j0 < SEL_TILE; j0 += 2048) {
            const int64_t i0 = base + j0 + 8 * threadIdx.x;
            uint32_t key[8];
            const uint32_t valid = keys8_f16(s, i0, n, key);
            uint32_t gm = 0, em = 0;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const bool in = (valid >> x) & 1u;
                gm |= (uint32_t)(in && key[x] > T) << x;
                em |= (uint32_t)(in && key[x] == T) << x;
            }
            const uint32_t c = ((uint32_t)__popc(gm) << 16) | (uint32_t)__popc(em);
            uint32_t inc = c;  // inclusive scan over the wave (each field < 2^16)
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(inc, o, 64);
                if (lane >= o) inc += v;
            }
            if (lane == 63) wt[par][wid] = inc;
            __syncthreads();
            uint32_t before = inc - c, total = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t t = wt[par][w];
                if (w < wid) before += t;
                total += t;
            }
            par ^= 1;
            if (gm | em) {
                int64_t gb = gt_run + (before >> 16);     // keys > T before score i0
                int64_t eb = eq_run + (before & 0xFFFFu);  // keys == T before score i0
#pragma unroll
                for (int x = 0; x < 8; ++x) {
                    const bool g = (gm >> x) & 1u, e = (em >> x) & 1u;
                    const int64_t slot = gb + min(eb, keq);
                    if ((g || (e && eb < keq)) && slot < k) out[slot] = i0 + x;
                    gb += g;
                    eb += e;
                }
            }
            gt_run += total >> 16;
            eq_run += total & 0xFFFFu;
        }
    }
}

template <int DT>
__global__ __launch_bounds__(1024) void k_sel_sort(const void* s, const int64_t* sel,
                                                   int64_t k, int P, int64_t* out) {
    extern __shared__ uint64_t comp[];
    for (int i = threadIdx.x; i < P; i += 1024) {
        if (i < k) {
            int64_t idx = sel[i];
            uint32_t key = load_key<DT>(s, idx);
            comp[i] = ((uint64_t)(~key) << 32) | (uint64_t)(uint32_t)idx;
        } else {
            comp[i] = ~0ull;
        }
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < P; i += 1024) {
                int j = i ^ stride;
                if (j > i) {
                    bool up = (i & size) == 0;
                    uint64_t a = comp[i], b = comp[j];
                    if ((a > b) == up) {
                        comp[i] = b;
                        comp[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < k; i += 1024) out[i] = (int64_t)(uint32_t)comp[i];
}

// ------------------------------------------------------------------------
// Segmented top-k: the per-batch selects of the reference's cascade
// (sseqs/msa.py:297-310, 324, 335) for every batch of a sweep in one launch.
// One block of SEG_THREADS threads per segment: histogram passes over the
// segment's keys in LDS fix the k-th largest key T, then one pass in index
// order writes the selection (keys > T, then the first keys == T), or the
// selection is staged in LDS and sorted by (key desc, index asc).

constexpr int SEG_THREADS = 1024;
constexpr int SEG_WAVES = SEG_THREADS / 64;
constexpr int SEG_SORT_MAX = 4096;
constexpr int SEG_E = 8;  // elements per thread per pass step

// 16-bit order key of an fp16 score (NaN lowest, as key_f32)
__device__ inline uint32_t key_h16(uint16_t h) {
    if ((h & 0x7c00u) == 0x7c00u && (h & 0x03ffu)) return 0u;
    return (h & 0x8000u) ? (~h & 0xffffu) : (h | 0x8000u);
}

template <int DT>
__device__ inline uint32_t seg_key(const void* s, int64_t i) {
    if constexpr (DT == SSEQS_F16)
        return key_h16(static_cast<const uint16_t*>(s)[i]);
    else
        return load_key<DT>(s, i);
}

template <int DT>
__global__ __launch_bounds__(SEG_THREADS) void k_seg_topk(
    const void* __restrict__ s, const int64_t* __restrict__ seg_begin,
    const int32_t* __restrict__ seg_n, const int32_t* __restrict__ seg_k,
    const int64_t* __restrict__ out_base, int64_t nseg, int sorted, int drop_front,
    int drop_back, int add_begin, int64_t* __restrict__ out) {
    __shared__ uint32_t hist[SEL_BINS];
    __shared__ uint64_t comp[SEG_SORT_MAX];
    __shared__ uint32_t wsum[SEG_WAVES];
    __shared__ uint32_t s_pick[2];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int npass = DT == SSEQS_F16 ? 2 : 3;
    const int shifts16[2] = {5, 0}, bits16[2] = {11, 5};
    const int shifts32[3] = {21, 10, 0}, bits32[3] = {11, 11, 10};
    for (int64_t j = blockIdx.x; j < nseg; j += gridDim.x) {
        const int64_t beg = seg_begin[j];
        const int n = seg_n[j];
        int k = min(seg_k[j], n);
        if (sorted) k = min(k, SEG_SORT_MAX);
        if (k <= 0) continue;  // uniform over the block
        const void* base = DT == SSEQS_F16
                               ? static_cast<const void*>(static_cast<const uint16_t*>(s) + beg)
                           : DT == SSEQS_F32
                               ? static_cast<const void*>(static_cast<const float*>(s) + beg)
                               : static_cast<const void*>(static_cast<const int32_t*>(s) + beg);
        uint32_t prefix = 0, mask = 0, krem = (uint32_t)k;
        for (int p = 0; p < npass; ++p) {
            const int shift = DT == SSEQS_F16 ? shifts16[p] : shifts32[p];
            const int bits = DT == SSEQS_F16 ? bits16[p] : bits32[p];
            const uint32_t bm = (1u << bits) - 1u;
            for (int i = tid; i < SEL_BINS; i += SEG_THREADS) hist[i] = 0;
            __syncthreads();
            // SEG_E independent loads in flight per thread
            for (int i0 = 0; i0 < n; i0 += SEG_THREADS * SEG_E) {
                uint32_t key[SEG_E];
#pragma unroll
                for (int e = 0; e < SEG_E; ++e) {
                    // clamped index: the loads are unconditional and all in flight
                    const int i = min(i0 + e * SEG_THREADS + tid, n - 1);
                    key[e] = seg_key<DT>(base, i);
                }
#pragma unroll
                for (int e = 0; e < SEG_E; ++e) {
                    const int i = i0 + e * SEG_THREADS + tid;
                    if (i < n && (key[e] & mask) == prefix)
                        atomicAdd(&hist[(key[e] >> shift) & bm], 1u);
                }
            }
            __syncthreads();
            if (wid == 0) {
                // lane L holds bins 32L..32L+31; the lane whose bins contain
                // the krem-th largest key walks them from the top
                uint32_t sum = 0;
                for (int t = 0; t < 32; ++t) sum += hist[lane * 32 + t];
                uint32_t incl = sum;  // keys in bins of lanes >= this one
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t t = __shfl_down(incl, off, 64);
                    if (lane + off < 64) incl += t;
                }
                uint32_t above = incl - sum;
                if (above < krem && krem <= incl) {
                    int b = lane * 32 + 31;
                    for (; b > lane * 32; --b) {
                        if (above + hist[b] >= krem) break;
                        above += hist[b];
                    }
                    s_pick[0] = (uint32_t)b;
                    s_pick[1] = krem - above;
                }
            }
            __syncthreads();
            prefix |= s_pick[0] << shift;
            mask |= bm << shift;
            krem = s_pick[1];
            __syncthreads();
        }
        // keys > T are all taken, keys == T while fewer than krem came
        // before.  Each thread takes SEG_E consecutive elements of a chunk;
        // a block scan of the per-thread counts (> T, == T), packed as
        // (gt << 16 | eq), gives every element its place in index order.
        const uint32_t T = prefix;
        const int keq = (int)krem;
        const int hi = k - drop_back;
        const int64_t ob = out_base[j] - drop_front;
        const int64_t add = add_begin ? beg : 0;
        int gt_run = 0, eq_run = 0;
        for (int i0 = 0; i0 < n; i0 += SEG_THREADS * SEG_E) {
            const int ib = i0 + tid * SEG_E;
            uint32_t key[SEG_E];
#pragma unroll
            for (int e = 0; e < SEG_E; ++e) key[e] = seg_key<DT>(base, min(ib + e, n - 1));
            uint32_t cnt = 0;
#pragma unroll
            for (int e = 0; e < SEG_E; ++e) {
                const bool in = ib + e < n;
                cnt += (in && key[e] > T) ? 0x10000u : 0u;
                cnt += (in && key[e] == T) ? 1u : 0u;
            }
            // inclusive scan over the wave, then over the waves
            uint32_t inc = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            if (lane == 63) wsum[wid] = inc;
            __syncthreads();
            uint32_t before = inc - cnt, total = 0;
#pragma unroll
            for (int w = 0; w < SEG_WAVES; ++w) {
                const uint32_t t = wsum[w];
                before += w < wid ? t : 0u;
                total += t;
            }
            int gb = gt_run + (int)(before >> 16), eb = eq_run + (int)(before & 0xffffu);
#pragma unroll
            for (int e = 0; e < SEG_E; ++e) {
                const bool in = ib + e < n;
                const bool g = in && key[e] > T;
                const bool q = in && key[e] == T;
                const int slot = gb + min(eb, keq);
                if ((g || (q && eb < keq)) && slot < k) {
                    if (sorted)
                        comp[slot] = ((uint64_t)(~key[e]) << 32) | (uint64_t)(uint32_t)(ib + e);
                    else if (slot >= drop_front && slot < hi)
                        out[ob + slot] = add + ib + e;
                }
                gb += g;
                eb += q;
            }
            gt_run += (int)(total >> 16);
            eq_run += (int)(total & 0xffffu);
            __syncthreads();
        }
        if (sorted) {
            int P = 1;
            while (P < k) P <<= 1;
            for (int i = k + tid; i < P; i += SEG_THREADS) comp[i] = ~0ull;
            __syncthreads();
            for (int size = 2; size <= P; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int i = tid; i < P; i += SEG_THREADS) {
                        const int m = i ^ stride;
                        if (m > i) {
                            const bool up = (i & size) == 0;
                            const uint64_t a = comp[i], b = comp[m];
                            if ((a > b) == up) {
                                comp[i] = b;
                                comp[m] = a;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = drop_front + tid; i < hi; i += SEG_THREADS)
                out[ob + i] = add + (int64_t)(uint32_t)comp[i];
            __syncthreads();
        }
    }
}

}  // namespace sseqs

using namespace sseqs;

// Launch the instance of a select kernel for the score type `dtype`.
#define SEL_DISPATCH(KERNEL, ...)                                                   \
    do {                                                                          \
        if (dtype == SSEQS_F32)                                                   \
            hipLaunchKernelGGL(KERNEL<SSEQS_F32>, __VA_ARGS__);                   \
        else if (dtype == SSEQS_F16)                                              \
            hipLaunchKernelGGL(KERNEL<SSEQS_F16>, __VA_ARGS__);                   \
        else                                                                      \
            hipLaunchKernelGGL(KERNEL<SSEQS_I32>, __VA_ARGS__);                   \
    } while (0)

extern "C" size_t sseqs_topk_ws_bytes(int64_t n) {
    int64_t nb = ceil_div(n > 0 ? n : 1, SEL_TILE);
    return (size_t)SEL_BINS * 4 + 64 + (size_t)nb * 16 + (size_t)SEL_SORT_MAX * 8 + 256;
}

extern "C" int sseqs_topk(const void* scores, int dtype, int64_t n, int64_t k, int sorted,
                          int64_t* out_idx, void* ws, size_t ws_bytes, hipStream_t stream) {
    if (n < 0 || k < 0) return fail(-1, "topk: negative size");
    if (k > n) return fail(-1, "topk: k larger than the number of scores");
    if (dtype != SSEQS_F32 && dtype != SSEQS_F16 && dtype != SSEQS_I32)
        return fail(-1, "topk: unsupported dtype");
    if (k == 0) return 0;
    if (sorted && k > SEL_SORT_MAX) return fail(-1, "topk: sorted output limited to k <= 16384");
    if (sorted && n > 0xffffffffLL) return fail(-1, "topk: sorted output needs n < 2^32");
    if (ws_bytes < sseqs_topk_ws_bytes(n)) return fail(-2, "topk: workspace too small");
    uint8_t* w = static_cast<uint8_t*>(ws);
    uint32_t* hist = reinterpret_cast<uint32_t*>(w);
    SelState* st = reinterpret_cast<SelState*>(w + SEL_BINS * 4);
    const int64_t nb = ceil_div(n, SEL_TILE);
    int64_t* cgt = reinterpret_cast<int64_t*>(w + SEL_BINS * 4 + 64);
    int64_t* ceq = cgt + nb;
    int64_t* tmp = ceq + nb;  // SEL_SORT_MAX entries for the sorted path

    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, stream, hist, st, (uint32_t)k);
    if (int r = check_launch("topk init")) return r;
    // keys of fp16 scores have their low 13 bits zero: two passes fix the rest
    const int shifts32[3] = {21, 10, 0}, bits32[3] = {11, 11, 10};
    const int shifts16[2] = {21, 13}, bits16[2] = {11, 8};
    const bool half = dtype == SSEQS_F16;
    const int* shifts = half ? shifts16 : shifts32;
    const int* bits = half ? bits16 : bits32;
    int64_t hb = ceil_div(n, 256 * 16);
    unsigned hgrid = (unsigned)(hb < 4096 ? (hb > 0 ? hb : 1) : 4096);
    // fp16 scores at a 16-byte aligned address: eight per thread and load
    const bool h8 = half && (reinterpret_cast<uintptr_t>(scores) & 15) == 0;
    const _Float16* sh = static_cast<const _Float16*>(scores);
    for (int p = 0; p < (half ? 2 : 3); ++p) {
        if (h8)
            hipLaunchKernelGGL(k_sel_hist_h8, dim3(hgrid), dim3(256), 0, stream, sh, n, st, hist,
                               shifts[p], bits[p]);
        else
            SEL_DISPATCH(k_sel_hist, dim3(hgrid), dim3(256), 0, stream, scores, n, st, hist,
                         shifts[p], bits[p]);
        if (int r = check_launch("topk hist")) return r;
        hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(1024), 0, stream, st, hist, shifts[p], bits[p]);
        if (int r = check_launch("topk pick")) return r;
    }
    if (h8)
        hipLaunchKernelGGL(k_sel_count_h8, dim3(grid_cap(nb, 1)), dim3(256), 0, stream, sh, n, nb,
                           st, cgt, ceq);
    else
        SEL_DISPATCH(k_sel_count, dim3(grid_cap(nb, 1)), dim3(256), 0, stream, scores, n,
                     nb, st, cgt, ceq);
    if (int r = check_launch("topk count")) return r;
    hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, stream, cgt, ceq, nb);
    if (int r = check_launch("topk scan")) return r;
    int64_t* dst = sorted ? tmp : out_idx;
    if (h8)
        hipLaunchKernelGGL(k_sel_write_h8, dim3(grid_cap(nb, 1)), dim3(256), 0, stream, sh, n, nb,
                           st, cgt, ceq, k, dst);
    else
        SEL_DISPATCH(k_sel_write, dim3(grid_cap(nb, 1)), dim3(256), 0, stream, scores, n,
                     nb, st, cgt, ceq, k, dst);
    if (int r = check_launch("topk write")) return r;
    if (sorted) {
        int P = 1;
        while (P < k) P <<= 1;
        size_t lds = (size_t)P * 8;
        const void* sort_fn = dtype == SSEQS_F32   ? reinterpret_cast<const void*>(&k_sel_sort<SSEQS_F32>)
                              : dtype == SSEQS_F16 ? reinterpret_cast<const void*>(&k_sel_sort<SSEQS_F16>)
                                                   : reinterpret_cast<const void*>(&k_sel_sort<SSEQS_I32>);
        if (lds > 64 * 1024 &&
            hipFuncSetAttribute(sort_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
                hipSuccess)
            return fail(-3, "topk: cannot enable large LDS for the sort");
        SEL_DISPATCH(k_sel_sort, dim3(1), dim3(1024), lds, stream, scores, tmp, k, P,
                           out_idx);
        if (int r = check_launch("topk sort")) return r;
    }
    return 0;
}

// ------------------------------------------------ sseqs_topk_noalloc --
//
// sseqs_topk on a library-owned scratch array instead of a caller
// workspace, for callers that cannot allocate.  `scratch` lives in device
// memory of every GPU the library runs on (one copy per device, held until
// the process ends).  Host code never passes the array's name to a kernel:
// the name is a host-side stand-in, so each device's address is looked up
// once with hipGetSymbolAddress (with that device current) and kept in
// g_scratch_ptr.  Every call on one device shares its copy; k_sel_init
// clears the histogram and state at the start of each call (a kernel, not
// hipMemsetAsync), and the count pass writes every per-tile counter, so
// nothing depends on the previous call's contents.  Calls that overlap on
// one device (two streams, or two threads) race on the array.

constexpr size_t NOALLOC_SCRATCH_BYTES = 512 * 1024;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Largest n whose workspace fits the scratch array.
int64_t noalloc_max_n() {
    const size_t fixed = (size_t)SEL_BINS * 4 + 64 + (size_t)SEL_SORT_MAX * 8 + 256;
    return (int64_t)((NOALLOC_SCRATCH_BYTES - fixed) / 16) * SEL_TILE;
}

// Device address of this device's scratch copy (looked up on first use).
int scratch_ptr(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(-4, "topk_noalloc: hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return fail(-4, "topk_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail((int)(e != hipSuccess ? e : hipErrorInvalidSymbol),
                        std::string("topk_noalloc: hipGetSymbolAddress failed: ") +
                            hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" int64_t sseqs_topk_noalloc_max_n(void) { return noalloc_max_n(); }

extern "C" int sseqs_topk_noalloc(const void* scores, int dtype, int64_t n, int64_t k, int sorted,
                                  int64_t* out_idx, hipStream_t stream) {
    if (n > noalloc_max_n())
        return fail(-1, "topk_noalloc: n = " + std::to_string(n) + " exceeds the scratch limit of " +
                            std::to_string(noalloc_max_n()) + " scores");
    void* ws = nullptr;
    if (int r = scratch_ptr(&ws)) return r;
    return sseqs_topk(scores, dtype, n, k, sorted, out_idx, ws, NOALLOC_SCRATCH_BYTES, stream);
}

extern "C" int sseqs_topk_segments(const void* scores, int dtype, const int64_t* seg_begin,
                                   const int32_t* seg_n, const int32_t* seg_k,
                                   const int64_t* out_base, int64_t nseg, int max_k, int sorted,
                                   int drop_front, int drop_back, int add_begin,
                                   int64_t* out_idx, hipStream_t stream) {
    if (nseg < 0 || max_k < 0 || drop_front < 0 || drop_back < 0)
        return fail(-1, "topk_segments: negative size");
    if (dtype != SSEQS_F32 && dtype != SSEQS_F16 && dtype != SSEQS_I32)
        return fail(-1, "topk_segments: unsupported dtype");
    if (sorted && max_k > SEG_SORT_MAX)
        return fail(-1, "topk_segments: sorted output limited to k <= 4096 per segment");
    if (nseg == 0 || max_k == 0) return 0;
    if (!scores || !seg_begin || !seg_n || !seg_k || !out_base || !out_idx)
        return fail(-1, "topk_segments: null pointer");
    const unsigned grid = (unsigned)(nseg < 8192 ? nseg : 8192);
    SEL_DISPATCH(k_seg_topk, dim3(grid), dim3(SEG_THREADS), 0, stream, scores,
                       seg_begin, seg_n, seg_k, out_base, nseg, sorted, drop_front, drop_back,
                       add_begin, out_idx);
    return check_launch("topk_segments");
}
