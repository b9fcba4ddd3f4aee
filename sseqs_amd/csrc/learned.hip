// Learned prefilter: the sequence encoder and the candidate proposal.
//
// Encoder of one sequence (codes 0..19, 20 for anything else):
//   x0[p] = E[code[p]]                              E: 22 x 64, rows 20, 21 zero
//   h1    = GELU(conv9(x0; W1: 64 -> 128) + b1)
//   h2    = GELU(conv9(h1; W2: 128 -> 128) + b2)
//   h3    = GELU(conv9(h2; W3: 128 -> 256) + b3)
//   z     = mean over positions of h3,  e = z / |z|      (256-d)
// conv9(x; W)[p][o] = sum_{t<9} sum_i x[p + t - 4][i] * W[t][i][o], zero
// outside the sequence.  The similarity of two sequences is e_q . e_t.
//
// Layout: a batch of sequences is one flat array of rows (positions) with
// four zero "pad" rows before every sequence and after the last one, and
// the row count rounded up to a multiple of 128 with pad rows; mask[p] = 1
// on residue rows, 0 on pad rows.  A layer's input row p + t - 4 of a
// residue row never leaves its sequence and its pads, so a conv layer is
// one GEMM over all rows: C[p][o] = sum_k A[p][k] * Bt[o][k] with
// A[p][k] = X[(p - 4) * Cin + k], k < 9 * Cin -- row p of A is the
// contiguous window of rows p-4 .. p+4 of X -- and Bt the weights as
// [o][t * Cin + i].  Every layer's output is zeroed on pad rows, so pads
// stay zero for the next layer.  The backward pass over the input is the
// same GEMM on dU (gradient before the GELU) with the weights flipped and
// transposed, Bt'[i][t * Cout + o] = W[8 - t][i][o]; the weight gradient
// dW[t * Cin + i][o] = sum_p A[p][t * Cin + i] * dU[p][o] reduces over rows.
//
// bf16 operands, fp32 accumulation on v_mfma_f32_32x32x16_bf16.  The
// proposal scores e_q . e_t of many queries against an FP8 (e4m3) table of
// every database sequence on v_mfma_f32_16x16x32_fp8_fp8.
#include "common.h"

namespace sseqs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(__bf16 v) { return (float)v; }
__device__ inline __bf16 f2bf(float v) { return (__bf16)v; }

// GELU, tanh form, and its derivative.
__device__ inline float gelu(float x) {
    const float c = 0.7978845608028654f;  // sqrt(2 / pi)
    const float u = c * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.f + tanhf(u));
}
__device__ inline float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

// ------------------------------------------------------------- conv GEMM --
//
// Block of 256 threads (4 waves) computes a 128 x BN tile of C; wave
// (wm, wn) the 64 x BN/2 sub-tile, as 2 x BN/64 MFMA tiles of 32 x 32.
// K advances 32 per stage: A (128 x 32) and Bt (BN x 32) tiles go through
// LDS (rows padded to 40 elements), the next stage's global loads are
// issued before the current stage's MFMAs.
//
// MFMA 32x32x16 operands: lane (r = lane % 32, h = lane / 32) supplies
// row r, elements 8h .. 8h+7 of K of both A and Bt; result element e of
// lane (c = lane % 32, h) is row 8 * (e / 4) + 4 * h + (e % 4), column c.
constexpr int GM = 128, GK = 32, GPAD = 40;

enum ConvEpi {
    EPI_FWD = 0,      // u = acc + bias; U = u (optional); Y = gelu(u) * mask
    EPI_BWD_GELU = 1, // dUprev = acc * gelu'(Uprev) * mask
    EPI_BWD_F32 = 2,  // dX (fp32) = acc * mask (input of layer 1)
};

struct ConvArgs {
    const __bf16* A;     // A row p at A + (p - 4) * lda (rows -4 .. M + 3 readable)
    int lda;             // elements between consecutive rows of A (Cin or Cout)
    const __bf16* Bt;    // N rows of K
    int M, N, K;
    const float* bias;   // EPI_FWD
    const uint8_t* mask; // [M]
    __bf16* U;           // EPI_FWD: pre-activation out (may be null)
    __bf16* Y;           // EPI_FWD: activation out; EPI_BWD_GELU: gradient out
    const __bf16* Uprev; // EPI_BWD_GELU
    float* Yf;           // EPI_BWD_F32
};

template <int BN, int EPI>
__global__ __launch_bounds__(256) void k_conv_gemm(ConvArgs a) {
    constexpr int NT = BN / 64;  // MFMA tiles per wave along N
    __shared__ __align__(16) __bf16 sA[2][GM * GPAD];
    __shared__ __align__(16) __bf16 sB[2][BN * GPAD];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int r = lane & 31, h = lane >> 5;
    const int64_t m0 = (int64_t)blockIdx.x * GM;
    const int n0 = blockIdx.y * BN;

    // global -> register staging: A tile 128 x 32 = 512 chunks of 8, B tile
    // BN x 32 = BN * 4 chunks; thread i takes chunks i and i + 256
    const int crow = tid >> 2, ccol = (tid & 3) * 8;
    const __bf16* gA0 = a.A + (m0 + crow - 4) * a.lda + ccol;
    const __bf16* gA1 = gA0 + (int64_t)64 * a.lda;
    const __bf16* gB0 = a.Bt + (int64_t)(n0 + crow) * a.K + ccol;
    const __bf16* gB1 = gB0 + (int64_t)64 * a.K;
    uint4 ra0, ra1, rb0, rb1 = make_uint4(0, 0, 0, 0);
#define CONV_GLOAD(k0)                                                     \
    do {                                                                   \
        ra0 = *reinterpret_cast<const uint4*>(gA0 + (k0));                 \
        ra1 = *reinterpret_cast<const uint4*>(gA1 + (k0));                 \
        rb0 = *reinterpret_cast<const uint4*>(gB0 + (k0));                 \
        if (BN == 128) rb1 = *reinterpret_cast<const uint4*>(gB1 + (k0));  \
    } while (0)
#define CONV_SSTORE(b)                                                                   \
    do {                                                                                 \
        *reinterpret_cast<uint4*>(&sA[b][crow * GPAD + ccol]) = ra0;                     \
        *reinterpret_cast<uint4*>(&sA[b][(crow + 64) * GPAD + ccol]) = ra1;              \
        *reinterpret_cast<uint4*>(&sB[b][crow * GPAD + ccol]) = rb0;                     \
        if (BN == 128) *reinterpret_cast<uint4*>(&sB[b][(crow + 64) * GPAD + ccol]) = rb1; \
    } while (0)

    f32x16 acc[2][NT];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

    const int nk = a.K / GK;
    CONV_GLOAD(0);
    CONV_SSTORE(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) CONV_GLOAD((kt + 1) * GK);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            bf16x8 fa[2], fb[NT];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const bf16x8*>(
                    &sA[buf][(64 * wm + 32 * i + r) * GPAD + 16 * s + 8 * h]);
#pragma unroll
            for (int j = 0; j < NT; ++j)
                fb[j] = *reinterpret_cast<const bf16x8*>(
                    &sB[buf][((BN / 2) * wn + 32 * j + r) * GPAD + 16 * s + 8 * h]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j],
                                                                        0, 0, 0);
        }
        if (kt + 1 < nk) CONV_SSTORE(buf ^ 1);
        __syncthreads();
    }
#undef CONV_GLOAD
#undef CONV_SSTORE

    // epilogue
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int n = n0 + (BN / 2) * wn + 32 * j + r;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t m = m0 + 64 * wm + 32 * i + 8 * (e >> 2) + 4 * h + (e & 3);
                const float v = acc[i][j][e];
                const float mk = a.mask[m] ? 1.f : 0.f;
                const int64_t o = m * a.N + n;
                if constexpr (EPI == EPI_FWD) {
                    const float u = v + a.bias[n];
                    if (a.U) a.U[o] = f2bf(u);
                    a.Y[o] = f2bf(gelu(u) * mk);
                } else if constexpr (EPI == EPI_BWD_GELU) {
                    a.Y[o] = f2bf(v * gelu_grad(bf2f(a.Uprev[o])) * mk);
                } else {
                    a.Yf[o] = v * mk;
                }
            }
        }
    }
}

// ---------------------------------------------------------- weight grad --
//
// dW[k][o] += sum_p A[p][k] * dU[p][o] over the rows p of this block's
// range, k = t * Cin + i.  Block: 64 rows of k x 64 columns of o, rows p in
// steps of 32 through LDS (X rows p-4 .. p+27+4 as the window source, dU
// rows p .. p+31), MFMA with K = rows: lane (r, h) gathers 8 consecutive
// rows of one column from LDS.  Partial sums of a block's row range are
// added to dW (fp32) atomically.
constexpr int WP = 32;   // rows per step
constexpr int WROWS = 2048;  // rows per block

__global__ __launch_bounds__(256) void k_conv_wgrad(const __bf16* __restrict__ X, int Cin,
                                                    const __bf16* __restrict__ dU, int Cout,
                                                    int64_t M, float* __restrict__ dW) {
    // tile: k in [k0, k0 + 64) (k = t * Cin + i), o in [o0, o0 + 64); the
    // k-columns of row p are X[(p - 4) * Cin + k0 .. + 64), contiguous
    const int k0 = blockIdx.x * 64, o0 = blockIdx.y * 64;
    const int64_t p_lo = (int64_t)blockIdx.z * WROWS;
    const int64_t p_hi = min(p_lo + WROWS, M);
    __shared__ __align__(16) __bf16 sA[WP * 64];
    __shared__ __align__(16) __bf16 sD[WP * 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r = lane & 31, h = lane >> 5;
    const int wm = w >> 1, wn = w & 1;  // wave sub-tile: 32 k x 32 o
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.f;
    for (int64_t p0 = p_lo; p0 < p_hi; p0 += WP) {
        __syncthreads();
        {   // 32 rows x 64 columns of A and of dU, 8 elements per thread
            const int row = tid >> 3, col = (tid & 7) * 8;
            const int64_t p = p0 + row;
            uint4 va = make_uint4(0, 0, 0, 0), vd = make_uint4(0, 0, 0, 0);
            if (p < p_hi) {
                va = *reinterpret_cast<const uint4*>(X + (p - 4) * Cin + k0 + col);
                vd = *reinterpret_cast<const uint4*>(dU + p * Cout + o0 + col);
            }
            *reinterpret_cast<uint4*>(&sA[row * 64 + col]) = va;
            *reinterpret_cast<uint4*>(&sD[row * 64 + col]) = vd;
        }
        __syncthreads();
        // K of the MFMA runs over rows: lane (r, h) gathers 8 rows of one column
#pragma unroll
        for (int s = 0; s < WP / 16; ++s) {
            bf16x8 fa, fb;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const int pp = 16 * s + 8 * h + x;
                fa[x] = sA[pp * 64 + 32 * wm + r];
                fb[x] = sD[pp * 64 + 32 * wn + r];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int k = k0 + 32 * wm + 8 * (e >> 2) + 4 * h + (e & 3);
        const int o = o0 + 32 * wn + r;
        atomicAdd(&dW[(int64_t)k * Cout + o], acc[e]);
    }
}

// db[o] += sum over rows of dU[p][o]: a block sums 1024 rows per column
// (one thread per column of 64, four row phases), then one atomic each.
__global__ __launch_bounds__(256) void k_colsum(const __bf16* __restrict__ dU, int C, int64_t M,
                                                float* __restrict__ db) {
    __shared__ float s[4][64];
    const int c = blockIdx.y * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
    const int64_t lo = (int64_t)blockIdx.x * 1024, hi = min(lo + 1024, M);
    float acc = 0.f;
    for (int64_t p = lo + ph; p < hi; p += 4) acc += bf2f(dU[p * C + c]);
    s[ph][threadIdx.x & 63] = acc;
    __syncthreads();
    if (ph == 0) atomicAdd(&db[c], s[0][threadIdx.x] + s[1][threadIdx.x] + s[2][threadIdx.x] +
                                       s[3][threadIdx.x]);
}

// ------------------------------------------------------------ small ops --

// x0[p] = E[code[p]] (bf16 rows of 64)
__global__ void k_embed(const uint8_t* __restrict__ code, int64_t M, const float* __restrict__ E,
                        __bf16* __restrict__ x0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one thread: 8 channels
    if (i >= M * 8) return;
    const int64_t p = i >> 3;
    const int c0 = (int)(i & 7) * 8;
    const int cd = code[p];
    bf16x8 v;
#pragma unroll
    for (int x = 0; x < 8; ++x) v[x] = f2bf(E[cd * 64 + c0 + x]);
    *reinterpret_cast<bf16x8*>(x0 + p * 64 + c0) = v;
}

// dE[code][c] += dX0[p][c]: per-block sums in LDS, then one atomic per entry
__global__ __launch_bounds__(256) void k_embed_grad(const uint8_t* __restrict__ code, int64_t M,
                                                    const float* __restrict__ dx0,
                                                    float* __restrict__ dE) {
    __shared__ float s[22 * 64];
    for (int i = threadIdx.x; i < 22 * 64; i += 256) s[i] = 0.f;
    __syncthreads();
    const int c = threadIdx.x & 63, sub = threadIdx.x >> 6;
    for (int64_t p = (int64_t)blockIdx.x * 4 + sub; p < M; p += (int64_t)gridDim.x * 4)
        atomicAdd(&s[code[p] * 64 + c], dx0[p * 64 + c]);
    __syncthreads();
    for (int i = threadIdx.x; i < 20 * 64; i += 256)
        if (s[i] != 0.f) atomicAdd(&dE[i], s[i]);
}

// Sequence s occupies rows [row0[s], row0[s] + len[s]).  z = mean of h3
// rows, e = z / |z| (|z| = 0 leaves e = 0); one block of 256 threads per
// sequence.  Thread t reads channels 4 (t % 64) .. +3 of rows p = t / 64
// (mod 4), 8 bytes per load and 16 rows of the block in flight per step;
// the four row classes are summed in LDS, in a fixed order.
__global__ __launch_bounds__(256) void k_pool(const __bf16* __restrict__ h3,
                                              const int64_t* __restrict__ row0,
                                              const int32_t* __restrict__ len, int64_t nseq,
                                              float* __restrict__ e, float* __restrict__ znorm) {
    __shared__ float part[4][256];
    __shared__ float red[4];
    const int t = threadIdx.x, c4 = 4 * (t & 63), ph = t >> 6;
    for (int64_t s = blockIdx.x; s < nseq; s += gridDim.x) {
        const int64_t b = row0[s];
        const int L = len[s];
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        int p = ph;
        for (; p + 12 < L; p += 16) {
            uint2 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                v[u] = *reinterpret_cast<const uint2*>(h3 + (b + p + 4 * u) * 256 + c4);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a0 += __uint_as_float(v[u].x << 16);
                a1 += __uint_as_float(v[u].x & 0xffff0000u);
                a2 += __uint_as_float(v[u].y << 16);
                a3 += __uint_as_float(v[u].y & 0xffff0000u);
            }
        }
        for (; p < L; p += 4) {
            const uint2 v = *reinterpret_cast<const uint2*>(h3 + (b + p) * 256 + c4);
            a0 += __uint_as_float(v.x << 16);
            a1 += __uint_as_float(v.x & 0xffff0000u);
            a2 += __uint_as_float(v.y << 16);
            a3 += __uint_as_float(v.y & 0xffff0000u);
        }
        part[ph][c4] = a0;
        part[ph][c4 + 1] = a1;
        part[ph][c4 + 2] = a2;
        part[ph][c4 + 3] = a3;
        __syncthreads();
        const float acc = (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
        const float z = L > 0 ? acc / (float)L : 0.f;
        float sq = z * z;
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
        if ((t & 63) == 0) red[t >> 6] = sq;
        __syncthreads();
        const float nrm = sqrtf(red[0] + red[1] + red[2] + red[3]);
        e[s * 256 + t] = nrm > 0.f ? z / nrm : 0.f;
        if (t == 0 && znorm) znorm[s] = nrm;
        __syncthreads();
    }
}

// Backward of the pool and of layer 3's GELU: dz = (de - e (e . de)) / |z|,
// dU3[p][c] = dz[c] / L * gelu'(U3[p][c]) on the sequence's rows.
__global__ __launch_bounds__(256) void k_pool_grad(const float* __restrict__ de,
                                                   const float* __restrict__ e,
                                                   const float* __restrict__ znorm,
                                                   const __bf16* __restrict__ U3,
                                                   const int64_t* __restrict__ row0,
                                                   const int32_t* __restrict__ len, int64_t nseq,
                                                   __bf16* __restrict__ dU3) {
    __shared__ float red[4];
    const int c = threadIdx.x;
    for (int64_t s = blockIdx.x; s < nseq; s += gridDim.x) {
        const float ev = e[s * 256 + c], dv = de[s * 256 + c];
        float dot = ev * dv;
        for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
        if ((c & 63) == 0) red[c >> 6] = dot;
        __syncthreads();
        const float d = red[0] + red[1] + red[2] + red[3];
        const float nrm = znorm[s];
        const int L = len[s];
        const float dz = nrm > 0.f && L > 0 ? (dv - ev * d) / nrm / (float)L : 0.f;
        const int64_t b = row0[s];
        for (int p = 0; p < L; ++p) {
            const int64_t o = (b + p) * 256 + c;
            dU3[o] = f2bf(dz * gelu_grad(bf2f(U3[o])));
        }
        __syncthreads();
    }
}

// Pair loss: for pair j, cos = e[qi] . e[ti], loss += (cos - y)^2 / npair;
// de[qi] += 2 (cos - y) / npair * e[ti] and symmetrically.  One wave per
// pair (4 channels per lane).
__global__ __launch_bounds__(256) void k_pair_loss(const float* __restrict__ e, int64_t nseq,
                                                   const int32_t* __restrict__ qi,
                                                   const int32_t* __restrict__ ti,
                                                   const float* __restrict__ y,
                                                   const float* __restrict__ wt, int64_t npair,
                                                   float* __restrict__ de, float* __restrict__ cos_out,
                                                   float* __restrict__ loss) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= npair) return;
    // pairs naming a sequence outside e are skipped (and counted nowhere)
    if (qi[j] < 0 || qi[j] >= nseq || ti[j] < 0 || ti[j] >= nseq) return;
    const float* a = e + (int64_t)qi[j] * 256;
    const float* b = e + (int64_t)ti[j] * 256;
    float d = 0.f;
#pragma unroll
    for (int x = 0; x < 4; ++x) d += a[lane * 4 + x] * b[lane * 4 + x];
    for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
    const float wj = wt ? wt[j] : 1.f;
    const float g = 2.f * wj * (d - y[j]) / (float)npair;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        atomicAdd(&de[(int64_t)qi[j] * 256 + lane * 4 + x], g * b[lane * 4 + x]);
        atomicAdd(&de[(int64_t)ti[j] * 256 + lane * 4 + x], g * a[lane * 4 + x]);
    }
    if (lane == 0) {
        if (cos_out) cos_out[j] = d;
        if (loss) atomicAdd(loss, wj * (d - y[j]) * (d - y[j]) / (float)npair);
    }
}

// Soft-target InfoNCE over a step's sequences: rows 0..nq-1 of e are the
// queries, rows nq..nq+nt-1 the targets.  For query q the logits are
// z_t = e_q . e_t / tau over every target and, with the target distribution
// P[q][.] (a row summing to 1, or all zero to leave the query out), the
// loss is -scale * sum_t P_qt log softmax(z)_t.  With
// G_qt = scale * (softmax(z)_t - P_qt), de_q += sum_t G_qt e_t / tau here and
// de_t += sum_q G_qt e_q / tau in k_infonce_targets.  One block per query;
// the logits live in LDS (nt <= INFONCE_MAX_T).
constexpr int INFONCE_MAX_T = 8192;

__device__ inline float block_reduce_256(float v, float* red, bool is_max) {
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(v, off, 64);
        v = is_max ? fmaxf(v, o) : v + o;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return is_max ? fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))
                  : red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_infonce(const float* __restrict__ e, int nq, int nt,
                                                 const float* __restrict__ P, float inv_tau,
                                                 float scale, float* __restrict__ G,
                                                 float* __restrict__ de, float* __restrict__ loss) {
    __shared__ float s_q[256];
    __shared__ float s_z[INFONCE_MAX_T];
    __shared__ float red[4];
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    const float* Pq = P + (int64_t)q * nt;
    float* Gq = G + (int64_t)q * nt;
    s_q[tid] = e[(int64_t)q * 256 + tid];
    __syncthreads();
    float mx = -INFINITY, psum = 0.f;
    for (int t = tid; t < nt; t += 256) {
        const float4* r = reinterpret_cast<const float4*>(e + (int64_t)(nq + t) * 256);
        float d = 0.f;
#pragma unroll 8
        for (int x = 0; x < 64; ++x) {
            const float4 v = r[x];
            d += v.x * s_q[4 * x] + v.y * s_q[4 * x + 1] + v.z * s_q[4 * x + 2] +
                 v.w * s_q[4 * x + 3];
        }
        const float z = d * inv_tau;
        s_z[t] = z;
        mx = fmaxf(mx, z);
        psum += Pq[t];
    }
    mx = block_reduce_256(mx, red, true);
    psum = block_reduce_256(psum, red, false);
    if (!(psum > 0.f)) {  // the query takes no part: zero gradient row
        for (int t = tid; t < nt; t += 256) Gq[t] = 0.f;
        return;
    }
    float se = 0.f;
    for (int t = tid; t < nt; t += 256) se += __expf(s_z[t] - mx);
    se = block_reduce_256(se, red, false);
    const float lse = mx + __logf(se);
    float lsum = 0.f;
    for (int t = tid; t < nt; t += 256) {
        const float pt = Pq[t];
        const float z = s_z[t];
        lsum -= pt * (z - lse);
        const float g = scale * (__expf(z - lse) - pt);
        Gq[t] = g;
        s_z[t] = g;
    }
    lsum = block_reduce_256(lsum, red, false);
    if (tid == 0) atomicAdd(loss, scale * lsum);
    // de_q: thread tid owns channel tid (reads of row t are coalesced)
    float acc = 0.f;
    for (int t = 0; t < nt; ++t) acc += s_z[t] * e[(int64_t)(nq + t) * 256 + tid];
    de[(int64_t)q * 256 + tid] += acc * inv_tau;
}

// de_t += sum_q G_qt e_q / tau: one block per target, one thread per channel.
__global__ __launch_bounds__(256) void k_infonce_targets(const float* __restrict__ e, int nq,
                                                         int nt, const float* __restrict__ G,
                                                         float inv_tau, float* __restrict__ de) {
    const int t = blockIdx.x;
    const int c = threadIdx.x;
    float acc = 0.f;
    for (int q = 0; q < nq; ++q) acc += G[(int64_t)q * nt + t] * e[(int64_t)q * 256 + c];
    de[(int64_t)(nq + t) * 256 + c] += acc * inv_tau;
}

// AdamW step on fp32 parameters.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float wd, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
}

// Weight layouts for the GEMMs from W[t][i][o] (fp32): forward
// Bt[o][t * Cin + i], and backward-data Bt'[i][t * Cout + o] = W[8 - t][i][o].
__global__ void k_pack_weights(const float* __restrict__ W, int Cin, int Cout,
                               __bf16* __restrict__ fwd, __bf16* __restrict__ bwd) {
    const int64_t n = (int64_t)9 * Cin * Cout;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int o = (int)(idx % Cout);
    const int i = (int)((idx / Cout) % Cin);
    const int t = (int)(idx / ((int64_t)Cout * Cin));
    const __bf16 v = f2bf(W[idx]);
    fwd[(int64_t)o * 9 * Cin + t * Cin + i] = v;
    if (bwd) bwd[(int64_t)i * 9 * Cout + (8 - t) * Cout + o] = v;
}

// ------------------------------------------------------------- table --

// Flat rows of a range of database sequences: codes (0..19, 20 other,
// 21 pad) and mask; sequence s of the range starts at row row0[s].
__global__ void k_db_rows(const uint8_t* __restrict__ packed, const int64_t* __restrict__ off,
                          const int32_t* __restrict__ len, int64_t s_begin, int64_t nseq,
                          const int64_t* __restrict__ row0, int max_len,
                          uint8_t* __restrict__ code, uint8_t* __restrict__ mask) {
    for (int64_t s = blockIdx.x; s < nseq; s += gridDim.x) {
        const int64_t o = off[s_begin + s];
        const int L = min(len[s_begin + s], max_len);
        const int64_t b = row0[s];
        for (int p = threadIdx.x; p < L; p += blockDim.x) {
            const int a = aa_index(packed_sym(packed, o + p));
            code[b + p] = (uint8_t)a;
            mask[b + p] = 1;
        }
    }
}

// Table rows: fp8 (e4m3) of scale * e, 256 bytes per sequence.
__global__ void k_to_fp8(const float* __restrict__ e, int64_t n, float scale,
                         uint8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // 4 values per thread
    if (i * 4 >= n) return;
    const float4 v = *reinterpret_cast<const float4*>(e + i * 4);
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(scale * v.x, scale * v.y, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(scale * v.z, scale * v.w, w, true);
    *reinterpret_cast<int*>(out + i * 4) = w;
}

// ---------------------------------------------------------- proposal --
//
// S[q][s] = sum_c T[s][c] * Q[q][c] over the fp8 table T (n x 256) and G
// query rows Q (G x 256, G a multiple of 16, at most 128), written as fp16
// to S[q * ldS + s].  A block takes PROP_ROWS consecutive table rows per
// step (4 waves x 32 PROP_SUB rows, kept in registers for every query group)
// against all G queries held in LDS.  MFMA 16x16x32 fp8: lane (r = lane %
// 16, g = lane / 16) supplies 8 bytes of row r per 32-byte K slice; result
// element i of lane (c, g) is row 4g + i, column c.  A dot product does not
// depend on the order of its terms, so the K slices are taken in the order
// that lets each lane read 16 bytes at once: slice s of lane (r, g) is bytes
// 64 (s / 2) + 16 g + 8 (s % 2) of its row, for the table and the query
// alike (16-byte loads read the table at 6.2 TB/s where 8-byte loads of the
// natural order reach 4.8).  Each group of 16 queries' 16 x PROP_ROWS scores
// goes through LDS so that the block writes every query's 2 PROP_ROWS
// contiguous bytes with 16-byte stores.
constexpr int PROP_SUB = 2;                // 32-row tiles per wave and step
constexpr int PROP_ROWS = 128 * PROP_SUB;  // table rows per block and step
constexpr int PROP_MAX_SLOTS = 32;         // output rows of a MAXQ launch
static_assert(PROP_ROWS == 256, "the MAXQ fold gives each of 256 threads one table row");

// Orders a wave's LDS accesses across its lanes (no barrier between waves):
// wavefront-scope release/acquire fences keep the compiler's memory
// operations on either side, and the hardware runs one wave's LDS
// instructions in order.
__device__ __forceinline__ void lds_wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Output rows of a MAXQ launch, as runs of query rows: query rows lo .. hi-1
// are folded, bit q of starts marks the first row of a run, and output row o
// is the maximum over the rows of run o.
struct PropSlots {
    uint64_t starts;
    int lo, hi;
};

// MAXQ: the kernel writes, per output row, the maximum of its query rows'
// scores (taken in ascending row order, as sseqs_rows_max takes them)
// instead of every query row's scores.  Thread t keeps the running maximum
// of table row t of the step in a register while it walks the query rows in
// order, and stores an output row's maximum to LDS once, when its run of
// query rows ends; no value is read back, so the fold costs one LDS read per
// query row and score.
template <bool MAXQ>
__global__ __launch_bounds__(256) void k_propose(const uint8_t* __restrict__ T, int64_t n,
                                                 const uint8_t* __restrict__ Q, int G,
                                                 PropSlots slots, int nslot,
                                                 _Float16* __restrict__ S, int64_t ldS) {
    constexpr int QROWS = MAXQ ? 64 : 128;
    __shared__ __align__(16) uint8_t sQ[QROWS * 256];
    __shared__ __align__(16) _Float16 sS[16 * PROP_ROWS];
    __shared__ __align__(16) _Float16 sM[MAXQ ? PROP_MAX_SLOTS * PROP_ROWS : 8];
    for (int i = threadIdx.x; i < G * 16; i += 256)
        reinterpret_cast<uint4*>(sQ)[i] = reinterpret_cast<const uint4*>(Q)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    // the cooperative store: thread -> (row oq of 16, scores 8 oc .. 8 oc + 7
    // of each 128-score part)
    const int oq = threadIdx.x >> 4, oc = threadIdx.x & 15;
    const bool vec = (ldS & 7) == 0 && (reinterpret_cast<uintptr_t>(S) & 15) == 0;
    for (int64_t b0 = (int64_t)blockIdx.x * PROP_ROWS; b0 < n;
         b0 += (int64_t)gridDim.x * PROP_ROWS) {
        const int64_t t0 = b0 + 32 * PROP_SUB * w;
        // this wave's 32 PROP_SUB table rows (two MFMA row tiles per 32), all
        // 8 K slices, kept for every query group
        long fa[PROP_SUB][2][8];
#pragma unroll
        for (int u = 0; u < PROP_SUB; ++u)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int64_t row = min(t0 + 32 * u + 16 * i + r, n - 1);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint4 v =
                        *reinterpret_cast<const uint4*>(T + row * 256 + 64 * j + 16 * g);
                    fa[u][i][2 * j] = (long)(((uint64_t)v.y << 32) | v.x);
                    fa[u][i][2 * j + 1] = (long)(((uint64_t)v.w << 32) | v.z);
                }
            }
        int cur_sl = -1;   // MAXQ: the output row of the run being folded
        _Float16 cur = 0;  // and its running maximum (table row threadIdx.x)
        for (int q0 = 0; q0 < G; q0 += 16) {
            f32x4 acc[PROP_SUB][2];
#pragma unroll
            for (int u = 0; u < PROP_SUB; ++u)
#pragma unroll
                for (int i = 0; i < 2; ++i) acc[u][i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const long fb = *reinterpret_cast<const long*>(
                    &sQ[(q0 + r) * 256 + 64 * (s >> 1) + 16 * g + 8 * (s & 1)]);
#pragma unroll
                for (int u = 0; u < PROP_SUB; ++u)
#pragma unroll
                    for (int i = 0; i < 2; ++i)
                        acc[u][i] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(fa[u][i][s], fb,
                                                                               acc[u][i], 0, 0, 0);
            }
            // rows 4g .. 4g+3 of each row tile are consecutive: one 8-byte LDS store
#pragma unroll
            for (int u = 0; u < PROP_SUB; ++u)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    union {
                        _Float16 h[4];
                        uint2 u2;
                    } v;
#pragma unroll
                    for (int x = 0; x < 4; ++x) v.h[x] = (_Float16)acc[u][i][x];
                    *reinterpret_cast<uint2*>(
                        &sS[r * PROP_ROWS + 32 * PROP_SUB * w + 32 * u + 16 * i + 4 * g]) = v.u2;
                }
            // MAXQ: thread t folds table row t, which its own wave wrote (the
            // wave's rows are 32 PROP_SUB w .. + 32 PROP_SUB), and a wave's LDS
            // accesses complete in order: the waves need not wait for each other
            if constexpr (MAXQ)
                lds_wave_fence();
            else
                __syncthreads();
            if constexpr (MAXQ) {
                const int c = threadIdx.x;
                _Float16 x[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) x[q] = sS[q * PROP_ROWS + c];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int row = q0 + q;  // the run structure is uniform
                    if (row < slots.lo || row >= slots.hi) continue;
                    if ((slots.starts >> row) & 1) {  // a run starts: the last one is done
                        if (cur_sl >= 0 && cur_sl < PROP_MAX_SLOTS) sM[cur_sl * PROP_ROWS + c] = cur;
                        ++cur_sl;
                        cur = x[q];
                    } else {
                        cur = x[q] > cur ? x[q] : cur;
                    }
                }
            } else {
#pragma unroll
                for (int part = 0; part < PROP_SUB; ++part) {
                    const int64_t row = b0 + 128 * part + 8 * oc;
                    _Float16* dst = S + (int64_t)(q0 + oq) * ldS + row;
                    const _Float16* src = &sS[oq * PROP_ROWS + 128 * part + 8 * oc];
                    if (vec && row + 8 <= n) {
                        *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
                    } else {
                        for (int x = 0; x < 8; ++x)
                            if (row + x < n) dst[x] = src[x];
                    }
                }
            }
            if constexpr (MAXQ)
                lds_wave_fence();  // the fold's reads before the next tile's writes
            else
                __syncthreads();
        }
        if constexpr (MAXQ) {
            if (cur_sl >= 0 && cur_sl < PROP_MAX_SLOTS) sM[cur_sl * PROP_ROWS + threadIdx.x] = cur;
            __syncthreads();
            // every output row's PROP_ROWS scores, 16 bytes per thread and pass
            for (int e = threadIdx.x; e < nslot * (PROP_ROWS / 8); e += 256) {
                const int sl = e / (PROP_ROWS / 8), c8 = e % (PROP_ROWS / 8);
                const int64_t row = b0 + 8 * c8;
                _Float16* dst = S + (int64_t)sl * ldS + row;
                const _Float16* src = &sM[sl * PROP_ROWS + 8 * c8];
                if (vec && row + 8 <= n) {
                    *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
                } else {
                    for (int x = 0; x < 8; ++x)
                        if (row + x < n) dst[x] = src[x];
                }
            }
            __syncthreads();
        }
    }
}

// out[j] = max over r < nrows of S[r * ld + j], j < n (fp16; ld a multiple
// of 8 and S, out 16-byte aligned: 8 scores per thread and step).
__global__ __launch_bounds__(256) void k_rows_max(const _Float16* __restrict__ S, int64_t ld,
                                                  int nrows, int64_t n,
                                                  _Float16* __restrict__ out) {
    const int64_t n8 = n / 8;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c * 8 < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (c < n8) {
            uint4 v = *reinterpret_cast<const uint4*>(S + c * 8);
            _Float16* m = reinterpret_cast<_Float16*>(&v);
            for (int r = 1; r < nrows; ++r) {
                uint4 w = *reinterpret_cast<const uint4*>(S + r * ld + c * 8);
                const _Float16* x = reinterpret_cast<const _Float16*>(&w);
#pragma unroll
                for (int e = 0; e < 8; ++e) m[e] = x[e] > m[e] ? x[e] : m[e];
            }
            *reinterpret_cast<uint4*>(out + c * 8) = v;
        } else {
            for (int64_t j = c * 8; j < n; ++j) {
                _Float16 m = S[j];
                for (int r = 1; r < nrows; ++r) m = S[r * ld + j] > m ? S[r * ld + j] : m;
                out[j] = m;
            }
        }
    }
}

}  // namespace sseqs

using namespace sseqs;

// ---------------------------------------------------------------- C ABI --

extern "C" int sseqs_enc_conv(const void* A, int lda, const void* Bt, int64_t M, int N, int K,
                              const float* bias, const uint8_t* mask, void* U, void* Y,
                              const void* Uprev, float* Yf, int epi, hipStream_t stream) {
    if (M <= 0 || M % GM) return fail(-1, "enc_conv: M must be a positive multiple of 128");
    if (K <= 0 || K % GK) return fail(-1, "enc_conv: K must be a positive multiple of 32");
    if (N % 64) return fail(-1, "enc_conv: N must be a multiple of 64");
    if (M / GM > 0x7fffffffLL) return fail(-1, "enc_conv: too many rows");
    ConvArgs a{static_cast<const __bf16*>(A), lda, static_cast<const __bf16*>(Bt), (int)0, N, K,
               bias, mask, static_cast<__bf16*>(U), static_cast<__bf16*>(Y),
               static_cast<const __bf16*>(Uprev), Yf};
    a.M = (int)(M < 0x7fffffffLL ? M : 0x7fffffffLL);
    const bool wide = N % 128 == 0;
    const dim3 grid((unsigned)(M / GM), (unsigned)(N / (wide ? 128 : 64)));
    switch (epi) {
    case EPI_FWD:
        if (!Y || !bias) return fail(-1, "enc_conv: forward needs Y and bias");
        if (wide) hipLaunchKernelGGL((k_conv_gemm<128, EPI_FWD>), grid, dim3(256), 0, stream, a);
        else hipLaunchKernelGGL((k_conv_gemm<64, EPI_FWD>), grid, dim3(256), 0, stream, a);
        break;
    case EPI_BWD_GELU:
        if (!Y || !Uprev) return fail(-1, "enc_conv: backward needs Y and Uprev");
        if (wide) hipLaunchKernelGGL((k_conv_gemm<128, EPI_BWD_GELU>), grid, dim3(256), 0, stream, a);
        else hipLaunchKernelGGL((k_conv_gemm<64, EPI_BWD_GELU>), grid, dim3(256), 0, stream, a);
        break;
    case EPI_BWD_F32:
        if (!Yf) return fail(-1, "enc_conv: backward to the input needs Yf");
        if (wide) hipLaunchKernelGGL((k_conv_gemm<128, EPI_BWD_F32>), grid, dim3(256), 0, stream, a);
        else hipLaunchKernelGGL((k_conv_gemm<64, EPI_BWD_F32>), grid, dim3(256), 0, stream, a);
        break;
    default:
        return fail(-1, "enc_conv: unknown epilogue");
    }
    return check_launch("enc_conv");
}

extern "C" int sseqs_enc_wgrad(const void* X, int Cin, const void* dU, int Cout, int64_t M,
                               float* dW, hipStream_t stream) {
    if (M <= 0 || Cin % 64 || Cout % 64) return fail(-1, "enc_wgrad: bad sizes");
    if (ceil_div(M, WROWS) > 65535) return fail(-1, "enc_wgrad: at most 65535 * 2048 rows");
    const dim3 grid((unsigned)(9 * Cin / 64), (unsigned)(Cout / 64),
                    (unsigned)ceil_div(M, WROWS));
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(256), 0, stream, static_cast<const __bf16*>(X),
                       Cin, static_cast<const __bf16*>(dU), Cout, M, dW);
    return check_launch("enc_wgrad");
}

extern "C" int sseqs_enc_bias_grad(const void* dU, int C, int64_t M, float* db,
                                   hipStream_t stream) {
    if (M <= 0) return 0;
    if (C % 64) return fail(-1, "enc_bias_grad: C must be a multiple of 64");
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)ceil_div(M, 1024), (unsigned)(C / 64)), dim3(256),
                       0, stream, static_cast<const __bf16*>(dU), C, M, db);
    return check_launch("enc_bias_grad");
}

extern "C" int sseqs_enc_embed(const uint8_t* code, int64_t M, const float* E, void* x0,
                               hipStream_t stream) {
    if (M <= 0) return 0;
    hipLaunchKernelGGL(k_embed, dim3((unsigned)ceil_div(M * 8, 256)), dim3(256), 0, stream, code,
                       M, E, static_cast<__bf16*>(x0));
    return check_launch("enc_embed");
}

extern "C" int sseqs_enc_embed_grad(const uint8_t* code, int64_t M, const float* dx0, float* dE,
                                    hipStream_t stream) {
    if (M <= 0) return 0;
    hipLaunchKernelGGL(k_embed_grad, dim3(grid_cap(M, 4 * 64)), dim3(256), 0, stream, code, M,
                       dx0, dE);
    return check_launch("enc_embed_grad");
}

extern "C" int sseqs_enc_pool(const void* h3, const int64_t* row0, const int32_t* len,
                              int64_t nseq, float* e, float* znorm, hipStream_t stream) {
    if (nseq <= 0) return 0;
    hipLaunchKernelGGL(k_pool, dim3(grid_cap(nseq, 1)), dim3(256), 0, stream,
                       static_cast<const __bf16*>(h3), row0, len, nseq, e, znorm);
    return check_launch("enc_pool");
}

extern "C" int sseqs_enc_pool_grad(const float* de, const float* e, const float* znorm,
                                   const void* U3, const int64_t* row0, const int32_t* len,
                                   int64_t nseq, void* dU3, hipStream_t stream) {
    if (nseq <= 0) return 0;
    hipLaunchKernelGGL(k_pool_grad, dim3(grid_cap(nseq, 1)), dim3(256), 0, stream, de, e, znorm,
                       static_cast<const __bf16*>(U3), row0, len, nseq, static_cast<__bf16*>(dU3));
    return check_launch("enc_pool_grad");
}

extern "C" int sseqs_enc_pair_loss(const float* e, int64_t nseq, const int32_t* qi,
                                   const int32_t* ti, const float* y, const float* w,
                                   int64_t npair, float* de, float* cos_out, float* loss,
                                   hipStream_t stream) {
    if (npair <= 0) return 0;
    hipLaunchKernelGGL(k_pair_loss, dim3((unsigned)ceil_div(npair, 4)), dim3(256), 0, stream, e,
                       nseq, qi, ti, y, w, npair, de, cos_out, loss);
    return check_launch("enc_pair_loss");
}

extern "C" int sseqs_enc_infonce(const float* e, int nq, int nt, const float* P, float tau,
                                 float scale, float* G, float* de, float* loss,
                                 hipStream_t stream) {
    if (nq <= 0 || nt <= 0) return fail(-1, "enc_infonce: need at least one query and one target");
    if (nt > INFONCE_MAX_T) return fail(-1, "enc_infonce: more than 8192 targets");
    if (!(tau > 0.f)) return fail(-1, "enc_infonce: tau must be positive");
    hipLaunchKernelGGL(k_infonce, dim3(nq), dim3(256), 0, stream, e, nq, nt, P, 1.f / tau, scale,
                       G, de, loss);
    if (int r = check_launch("enc_infonce")) return r;
    hipLaunchKernelGGL(k_infonce_targets, dim3(nt), dim3(256), 0, stream, e, nq, nt, G, 1.f / tau,
                       de);
    return check_launch("enc_infonce_targets");
}

extern "C" int sseqs_enc_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                              float b1, float b2, float eps, float wd, int step,
                              hipStream_t stream) {
    if (n <= 0) return 0;
    if (step < 1) return fail(-1, "enc_adam: step counts from 1");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, p, g, m,
                       v, n, lr, b1, b2, eps, wd, c1, c2);
    return check_launch("enc_adam");
}

extern "C" int sseqs_enc_pack_weights(const float* W, int Cin, int Cout, void* fwd, void* bwd,
                                      hipStream_t stream) {
    const int64_t n = (int64_t)9 * Cin * Cout;
    hipLaunchKernelGGL(k_pack_weights, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, W,
                       Cin, Cout, static_cast<__bf16*>(fwd), static_cast<__bf16*>(bwd));
    return check_launch("enc_pack_weights");
}

extern "C" int sseqs_enc_db_rows(const uint8_t* packed, const int64_t* seq_off,
                                 const int32_t* seq_len, int64_t s_begin, int64_t nseq,
                                 const int64_t* row0, int max_len, uint8_t* code, uint8_t* mask,
                                 hipStream_t stream) {
    if (nseq <= 0) return 0;
    hipLaunchKernelGGL(k_db_rows, dim3(grid_cap(nseq, 1)), dim3(256), 0, stream, packed, seq_off,
                       seq_len, s_begin, nseq, row0, max_len, code, mask);
    return check_launch("enc_db_rows");
}

extern "C" int sseqs_enc_to_fp8(const float* e, int64_t n, float scale, uint8_t* out,
                                hipStream_t stream) {
    if (n <= 0) return 0;
    if (n % 4) return fail(-1, "enc_to_fp8: n must be a multiple of 4");
    hipLaunchKernelGGL(k_to_fp8, dim3((unsigned)ceil_div(n / 4, 256)), dim3(256), 0, stream, e, n,
                       scale, out);
    return check_launch("enc_to_fp8");
}

extern "C" int sseqs_rows_max(const void* S, int64_t ld, int nrows, int64_t n, void* out,
                              hipStream_t stream) {
    if (n <= 0) return 0;
    if (nrows < 1) return fail(-1, "rows_max: need at least one row");
    if (ld < n || ld % 8) return fail(-1, "rows_max: ld must be >= n and a multiple of 8");
    if ((reinterpret_cast<uintptr_t>(S) | reinterpret_cast<uintptr_t>(out)) & 15)
        return fail(-1, "rows_max: S and out must be 16-byte aligned");
    hipLaunchKernelGGL(k_rows_max, dim3(grid_cap(ceil_div(n, 8), 256)), dim3(256), 0, stream,
                       static_cast<const _Float16*>(S), ld, nrows, n, static_cast<_Float16*>(out));
    return check_launch("rows_max");
}

extern "C" int sseqs_propose(const uint8_t* table, int64_t n, const uint8_t* q8, int G,
                             void* scores, int64_t ld, hipStream_t stream) {
    if (n <= 0) return 0;
    if (G <= 0 || G % 16 || G > 128) return fail(-1, "propose: G must be 16..128, a multiple of 16");
    if (ld < n) return fail(-1, "propose: ld smaller than n");
    if ((reinterpret_cast<uintptr_t>(table) | reinterpret_cast<uintptr_t>(q8)) & 15)
        return fail(-1, "propose: table and query rows must be 16-byte aligned");
    const unsigned grid =
        (unsigned)(ceil_div(n, PROP_ROWS) < 8192 ? ceil_div(n, PROP_ROWS) : 8192);
    hipLaunchKernelGGL(k_propose<false>, dim3(grid), dim3(256), 0, stream, table, n, q8, G,
                       PropSlots{}, 0, static_cast<_Float16*>(scores), ld);
    return check_launch("propose");
}

extern "C" int sseqs_propose_max(const uint8_t* table, int64_t n, const uint8_t* q8, int G,
                                 const int32_t* row_off, int nslot, void* scores, int64_t ld,
                                 hipStream_t stream) {
    if (n <= 0) return 0;
    if (G <= 0 || G % 16 || G > 64) return fail(-1, "propose_max: G must be 16..64, a multiple of 16");
    if (nslot < 1 || nslot > PROP_MAX_SLOTS) return fail(-1, "propose_max: 1..32 output rows");
    if (!row_off) return fail(-1, "propose_max: no row offsets");
    if (row_off[0] < 0 || row_off[nslot] > G)
        return fail(-1, "propose_max: row offsets outside the G query rows");
    PropSlots sl{0, row_off[0], row_off[nslot]};
    for (int o = 0; o < nslot; ++o) {
        if (row_off[o + 1] <= row_off[o])
            return fail(-1, "propose_max: row offsets must increase (no empty output row)");
        sl.starts |= 1ull << row_off[o];
    }
    if (ld < n) return fail(-1, "propose_max: ld smaller than n");
    if ((reinterpret_cast<uintptr_t>(table) | reinterpret_cast<uintptr_t>(q8)) & 15)
        return fail(-1, "propose_max: table and query rows must be 16-byte aligned");
    const unsigned grid =
        (unsigned)(ceil_div(n, PROP_ROWS) < 8192 ? ceil_div(n, PROP_ROWS) : 8192);
    hipLaunchKernelGGL(k_propose<true>, dim3(grid), dim3(256), 0, stream, table, n, q8, G, sl,
                       nslot, static_cast<_Float16*>(scores), ld);
    return check_launch("propose_max");
}
