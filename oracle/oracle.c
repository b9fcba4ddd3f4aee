/*
 * oracle.c -- CPU restatement of the sseqs MSA-search hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU kernels are
 * compared against; nothing in the product (sseqs_amd/) links, loads or
 * calls it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, through oracle/oracle.py.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Paths are relative to the sseqs repository root.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -shared).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* BLOSUM62 in the order ARNDCQEGHILKMFPSTWYV (sseqs/sw.cu:7-29, identical to
 * sseqs/msa.py:110 and sseqs/sw.py:230-252). */
static const int8_t B62[20][20] = {
    { 4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0},
    {-1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3},
    {-2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3},
    {-2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3},
    { 0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1},
    {-1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2},
    {-1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2},
    { 0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3},
    {-2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3},
    {-1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3},
    {-1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1},
    {-1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2},
    {-1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1},
    {-2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1},
    {-1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2},
    { 1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2},
    { 0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0},
    {-3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3},
    {-2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1},
    { 0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4},
};

/* Residue letter -> 0..19, else -1 (sseqs/sw.cu:31-53 char_to_uint; the
 * Triton table sseqs/sw.py:254-258 is the same map). */
static int aa_index(uint8_t c) {
    switch (c) {
    case 'A': return 0;  case 'R': return 1;  case 'N': return 2;  case 'D': return 3;
    case 'C': return 4;  case 'Q': return 5;  case 'E': return 6;  case 'G': return 7;
    case 'H': return 8;  case 'I': return 9;  case 'L': return 10; case 'K': return 11;
    case 'M': return 12; case 'F': return 13; case 'P': return 14; case 'S': return 15;
    case 'T': return 16; case 'W': return 17; case 'Y': return 18; case 'V': return 19;
    default: return -1;
    }
}

/* Prefilter DB code: 0..19 residues, 20 for X - *, 21 for everything else
 * (sseqs/fft.cu:5-15 ASCII2POS). */
static int corr_code(uint8_t c) {
    int a = aa_index(c);
    if (a >= 0) return a;
    if (c == 'X' || c == '-' || c == '*') return 20;
    return 21;
}

static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* 5-bit codec                                                               */
/* ------------------------------------------------------------------------ */

/* Pack n ASCII symbols (byte-64, low 5 bits) into an MSB-first bit stream of
 * ceil(5n/8) bytes: symbol j of each 8-symbol group sits at bits 39-5j..35-5j
 * of a big-endian 40-bit window (sseqs/xbit.cu:36-62).  Missing symbols of a
 * short final group are packed as 0 ('@'); only bytes inside ceil(5n/8) are
 * written (the reference writes a whole 5-byte group). */
void orc_xbit_pack(const uint8_t* in, int64_t n, uint8_t* out) {
    int64_t nout = (n * 5 + 7) / 8;
    for (int64_t g = 0; g * 8 < n; ++g) {
        uint64_t w = 0;
        for (int k = 0; k < 8; ++k) {
            int64_t i = g * 8 + k;
            uint64_t sym = (i < n) ? (uint64_t)((uint8_t)(in[i] - 64) & 31u) : 0u;
            w |= sym << (35 - 5 * k);
        }
        for (int b = 0; b < 5; ++b) {
            int64_t o = g * 5 + b;
            if (o < nout) out[o] = (uint8_t)(w >> (32 - 8 * b));
        }
    }
}

/* Unpack n symbols starting at byte offset lo, adding back 64
 * (sseqs/xbit.cu:8-34, launcher sseqs/xbit.cu:65-71 offsets the input by lo).
 * Only out[0..n) is written. */
void orc_xbit_unpack(const uint8_t* packed, int64_t n, int64_t lo, uint8_t* out) {
    const uint8_t* p0 = packed + lo;
    for (int64_t g = 0; g * 8 < n; ++g) {
        const uint8_t* p = p0 + g * 5;
        uint64_t w = ((uint64_t)p[0] << 32) | ((uint64_t)p[1] << 24) | ((uint64_t)p[2] << 16) |
                     ((uint64_t)p[3] << 8) | (uint64_t)p[4];
        for (int k = 0; k < 8; ++k) {
            int64_t i = g * 8 + k;
            if (i < n) out[i] = (uint8_t)(((w >> (35 - 5 * k)) & 31u) + 64u);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Smith-Waterman with the reference's 32-stripe rule                        */
/* ------------------------------------------------------------------------ */

/* Width of one lane's query stripe: the kernel is instantiated for the query
 * length rounded up to a multiple of 32 and each of the 32 lanes owns
 * LEN/32 consecutive columns (sseqs/sw.cu:79, 355, 390). */
static int stripe_width(int Lq) { return ((Lq + 31) >> 5); }

/* Substitution score of target letter t against query code qc (0..19) as the
 * CUDA kernels see it: -4 when t is not one of the 20 residues
 * (sseqs/sw.cu:128-130, 258-261).  Query codes >= 20 never come out of the
 * reference's encoders; they are scored -4 here too. */
static int sub_cuda(uint8_t t, uint8_t qc) {
    int c = (t < 128) ? aa_index(t) : -1;
    if (c < 0 || qc >= 20) return -4;
    return B62[c][qc];
}

/* Linear-gap score of one target (sseqs/sw.cu:100-155): uint16 cells,
 * H = max(0, Hdiag+s, Hup+g, Hleft+g) with Hleft = 0 at every stripe start
 * (column j with j % TPT == 0), boundaries 0, result = max cell. */
int32_t orc_sw_linear_tpt(const uint8_t* q, int Lq, const uint8_t* t, int T, int gap, int tpt) {
    if (Lq <= 0 || T <= 0) return 0;
    uint16_t* prev = (uint16_t*)calloc((size_t)Lq, sizeof(uint16_t));
    uint16_t* curr = (uint16_t*)calloc((size_t)Lq, sizeof(uint16_t));
    int best = 0;
    for (int i = 0; i < T; ++i) {
        int left = 0;
        for (int j = 0; j < Lq; ++j) {
            if (j % tpt == 0) left = 0;
            int diag = (j == 0) ? 0 : prev[j - 1];
            int s = sub_cuda(t[i], q[j]);
            int h = imax(0, imax(diag + s, imax((int)prev[j] + gap, left + gap)));
            uint16_t hc = (uint16_t)h;
            curr[j] = hc;
            if ((int)hc > best) best = hc;
            left = hc;
        }
        uint16_t* tmp = prev; prev = curr; curr = tmp;
    }
    free(prev); free(curr);
    return best;
}

/* Affine score of one target (sseqs/sw.cu:227-308), int16 cells:
 *   M = max(0, s + max(M,X,Y)[i-1][j-1])
 *   X = max(0, M[i][j-1]-o, X[i][j-1]-e)   (left operands 0 at stripe starts)
 *   Y = max(0, M[i-1][j]-o, Y[i-1][j]-e)
 * result = max over cells of max(M,X,Y). */
int32_t orc_sw_affine_tpt(const uint8_t* q, int Lq, const uint8_t* t, int T, int o, int e,
                          int tpt) {
    if (Lq <= 0 || T <= 0) return 0;
    int16_t *pM = calloc((size_t)Lq, 2), *pX = calloc((size_t)Lq, 2), *pY = calloc((size_t)Lq, 2);
    int16_t *cM = calloc((size_t)Lq, 2), *cX = calloc((size_t)Lq, 2), *cY = calloc((size_t)Lq, 2);
    int best = 0;
    for (int i = 0; i < T; ++i) {
        int lM = 0, lX = 0;
        for (int j = 0; j < Lq; ++j) {
            if (j % tpt == 0) { lM = 0; lX = 0; }
            int s = sub_cuda(t[i], q[j]);
            int dM = j ? pM[j - 1] : 0, dX = j ? pX[j - 1] : 0, dY = j ? pY[j - 1] : 0;
            int m = imax(0, imax(dM + s, imax(dX + s, dY + s)));
            int x = imax(0, imax(lM - o, lX - e));
            int y = imax(0, imax((int)pM[j] - o, (int)pY[j] - e));
            cM[j] = (int16_t)m; cX[j] = (int16_t)x; cY[j] = (int16_t)y;
            int h = imax(cM[j], imax(cX[j], cY[j]));
            if (h > best) best = h;
            lM = cM[j]; lX = cX[j];
        }
        int16_t* tmp;
        tmp = pM; pM = cM; cM = tmp;
        tmp = pX; pX = cX; cX = tmp;
        tmp = pY; pY = cY; cY = tmp;
    }
    free(pM); free(pX); free(pY); free(cM); free(cX); free(cY);
    return best;
}

/* The reference's kernels always use tpt = ceil(Lq/32); the explicit-tpt
 * forms exist so tests can pin the stripe rule on tiny inputs. */
int32_t orc_sw_linear_one(const uint8_t* q, int Lq, const uint8_t* t, int T, int gap) {
    return orc_sw_linear_tpt(q, Lq, t, T, gap, stripe_width(Lq));
}
int32_t orc_sw_affine_one(const uint8_t* q, int Lq, const uint8_t* t, int T, int o, int e) {
    return orc_sw_affine_tpt(q, Lq, t, T, o, e, stripe_width(Lq));
}

/* Segment selection shared by both kernels (sseqs/sw.cu:67-77, 195-204):
 * target g is ascii[P[gi-1]+1 .. P[gi]) with gi = good_idx[g]; the reference
 * reads P[-1] for gi == 0 (out of bounds); it is taken as -1 here, which makes
 * target 0 the prefix before the first delimiter.  -1 is returned for an
 * empty query, a zero length or stop < start. */
static int select_target(const int32_t* good_idx, const int32_t* P, int g, int64_t* start, int* len) {
    int gi = good_idx[g];
    int64_t prevp = gi > 0 ? (int64_t)P[gi - 1] : -1;
    int64_t st = prevp + 1, stop = (int64_t)P[gi] + 1;
    int64_t length = stop - st - 1;
    *start = st;
    *len = (int)(length > 0 ? length : 0);
    if (length == 0 || stop < st) return 0;
    return 1;
}

void orc_sw_linear(const uint8_t* q, int Lq, const int32_t* good_idx, int M, const uint8_t* ascii,
                   const int32_t* P, int gap, int32_t* out) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int g = 0; g < M; ++g) {
        int64_t st; int len;
        if (Lq == 0 || !select_target(good_idx, P, g, &st, &len)) { out[g] = -1; continue; }
        out[g] = orc_sw_linear_one(q, Lq, ascii + st, len, gap);
    }
}

void orc_sw_affine(const uint8_t* q, int Lq, const int32_t* good_idx, int M, const uint8_t* ascii,
                   const int32_t* P, int o, int e, int32_t* out) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int g = 0; g < M; ++g) {
        int64_t st; int len;
        if (Lq == 0 || !select_target(good_idx, P, g, &st, &len)) { out[g] = -1; continue; }
        out[g] = orc_sw_affine_one(q, Lq, ascii + st, len, o, e);
    }
}

/* Same recurrences over exact DB segments given by (offset, length) into a
 * symbol array; used for the whole-DB cascade and as the CPU baseline. */
void orc_sw_affine_segments(const uint8_t* q, int Lq, const uint8_t* ascii, const int64_t* off,
                            const int32_t* len, int64_t n, int o, int e, int32_t* out) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t g = 0; g < n; ++g) {
        if (Lq == 0 || len[g] <= 0) { out[g] = -1; continue; }
        out[g] = orc_sw_affine_one(q, Lq, ascii + off[g], len[g], o, e);
    }
}

/* Linear-gap counterpart over exact DB segments (sseqs/sw.cu:55-179); the
 * linear-gap run of the whole-DB cascade. */
void orc_sw_linear_segments(const uint8_t* q, int Lq, const uint8_t* ascii, const int64_t* off,
                            const int32_t* len, int64_t n, int gap, int32_t* out) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t g = 0; g < n; ++g) {
        if (Lq == 0 || len[g] <= 0) { out[g] = -1; continue; }
        out[g] = orc_sw_linear_one(q, Lq, ascii + off[g], len[g], gap);
    }
}

/* Textbook Gotoh local alignment with the same recurrences but no stripe
 * rule (horizontal gaps may cross every column).  Not a reference function:
 * used only to pin the stripe-rule oracle (stripe <= textbook, and equal when
 * the horizontal-gap term can never win). */
int32_t orc_gotoh_textbook(const uint8_t* q, int Lq, const uint8_t* t, int T, int o, int e) {
    if (Lq <= 0 || T <= 0) return 0;
    int *pM = calloc((size_t)Lq, 4), *pX = calloc((size_t)Lq, 4), *pY = calloc((size_t)Lq, 4);
    int *cM = calloc((size_t)Lq, 4), *cX = calloc((size_t)Lq, 4), *cY = calloc((size_t)Lq, 4);
    int best = 0;
    for (int i = 0; i < T; ++i) {
        int lM = 0, lX = 0;
        for (int j = 0; j < Lq; ++j) {
            int s = sub_cuda(t[i], q[j]);
            int dH = j ? imax(pM[j - 1], imax(pX[j - 1], pY[j - 1])) : 0;
            int m = imax(0, dH + s);
            int x = imax(0, imax(lM - o, lX - e));
            int y = imax(0, imax(pM[j] - o, pY[j] - e));
            cM[j] = m; cX[j] = x; cY[j] = y;
            if (m > best) best = m;
            lM = m; lX = x;
        }
        int* tmp;
        tmp = pM; pM = cM; cM = tmp;
        tmp = pX; pX = cX; cX = tmp;
        tmp = pY; pY = cY; cY = tmp;
    }
    free(pM); free(pX); free(pY); free(cM); free(cX); free(cY);
    return best;
}

/* ------------------------------------------------------------------------ */
/* Low-rank FFT prefilter, FP64 closed form                                  */
/* ------------------------------------------------------------------------ */

/* Score of every segment b < B (sseqs/filter.cu:92-159 + sseqs/fft.cu:107-412):
 *   a_r[t] = LUT_A[r][qrev[t]]               t < min(Lq, N)
 *   b_r[t] = LUT_Bflat[r*V + code(d_t)]      t < min(len_b, maxlen, N)
 *   corr[m] = (H/N) sum_r sum_t a_r[t] b_r[(m-t) mod N] = 0.5 * (circular conv)
 *   score   = max over even m of corr[m]
 * The forward transform is real-to-complex via an H = N/2 point packed FFT;
 * the inverse is an unnormalised H-point transform scaled by 1/N, hence the
 * factor H/N; only the real lane (even samples) enters the max
 * (sseqs/fft.cu:383-391).  code 21 reads LUT_Bflat[r*V + 21], i.e. the next
 * rank's column 0, exactly as the reference's flat indexing does
 * (sseqs/fft.cu:124-126).  Evaluated as sum_p W[q_s][d_p] with the rank-k
 * product W folded first; exact up to FP64 rounding. */
void orc_corr(const uint8_t* qrev, int Lq, const float* lutA, const float* lutB, int nrows_b, int V,
              int k, const uint8_t* flat, const int32_t* offsets, const int32_t* lengths, int B,
              int maxlen, int N, double* out) {
    /* W[a][c] = sum_r A[r][a] * Bflat[r*V + c] for a <= 20, c <= 21 */
    double W[21][22];
    for (int a = 0; a < 21; ++a)
        for (int c = 0; c < 22; ++c) {
            double acc = 0.0;
            for (int r = 0; r < k; ++r) {
                double av = (a < V) ? (double)lutA[r * V + a] : 0.0;
                int bi = r * V + c;
                double bv = (bi < nrows_b * V) ? (double)lutB[bi] : 0.0;
                acc += av * bv;
            }
            W[a][c] = acc;
        }
    int Lqe = Lq < N ? Lq : N;
    #pragma omp parallel for schedule(dynamic, 64)
    for (int b = 0; b < B; ++b) {
        int L = lengths[b];
        if (L > maxlen) L = maxlen;
        if (L > N) L = N;
        const uint8_t* d = flat + offsets[b];
        double best = -INFINITY;
        for (int m = 0; m < N; m += 2) {
            double acc = 0.0;
            for (int p = 0; p < L; ++p) {
                /* a index t = (m - p) mod N ; contributes when t < Lqe */
                int tt = (m - p) % N;
                if (tt < 0) tt += N;
                if (tt < Lqe) acc += W[qrev[tt] > 20 ? 20 : qrev[tt]][corr_code(d[p])];
            }
            acc *= 0.5;
            if (acc > best) best = acc;
        }
        out[b] = (L >= 0) ? best : 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Full affine DP + stateless traceback (Triton semantics)                   */
/* ------------------------------------------------------------------------ */

/* Substitution as the Triton kernels see it: -5 when either letter is not a
 * residue (sseqs/sw.py:254-260, 498-503, 671-677). */
static int sub_triton(uint8_t qc, uint8_t tc) {
    int a = (qc < 128) ? aa_index(qc) : -1;
    int b = (tc < 128) ? aa_index(tc) : -1;
    if (a < 0 || b < 0) return -5;
    return B62[a][b];
}

/* DP of sseqs/sw.py:474-563 (i = query 1..Lq, j = target 1..T, zero borders):
 *   M  = s + H(i-1,j-1)
 *   Ix = max(H(i-1,j)-o, (i-1==0 ? -16384 : Ix(i-1,j)) - e)
 *   Iy = max(H(i,j-1)-o, (j-1==0 ? -16384 : Iy(i,j-1)) - e)
 *   H  = max(0, M, Ix, Iy); best = max H (int16 cells)
 * then the traceback of sseqs/sw.py:612-768: end cell = first (i, then j)
 * cell with H == best; step while H > 0 preferring M (diagonal), then Ix
 * (query residue against '-'), then Iy ('-' against target residue).
 * Outputs (written only when an alignment exists): q_al/t_al of length
 * info[0]; info = {len, q_start, q_end, t_start, t_end, best}; starts/ends
 * are -1 and len 0 when nothing aligns.  Returns best. */
int orc_backtrack(const uint8_t* q, int Lq, const uint8_t* t, int T, int o, int e,
                  uint8_t* q_al, uint8_t* t_al, int32_t* info) {
    int W = T + 1;
    size_t cells = (size_t)(Lq + 1) * (size_t)W;
    int16_t* M = calloc(cells, 2);
    int16_t* X = calloc(cells, 2);
    int16_t* Y = calloc(cells, 2);
#define IDX(i, j) ((size_t)(i) * (size_t)W + (size_t)(j))
#define HV(i, j) imax(0, imax(M[IDX(i, j)], imax(X[IDX(i, j)], Y[IDX(i, j)])))
    int best = 0;
    for (int i = 1; i <= Lq; ++i)
        for (int j = 1; j <= T; ++j) {
            int s = sub_triton(q[i - 1], t[j - 1]);
            int m = s + HV(i - 1, j - 1);
            int xfrom = (i - 1 == 0) ? -16384 : X[IDX(i - 1, j)];
            int x = imax(HV(i - 1, j) - o, xfrom - e);
            int yfrom = (j - 1 == 0) ? -16384 : Y[IDX(i, j - 1)];
            int y = imax(HV(i, j - 1) - o, yfrom - e);
            M[IDX(i, j)] = (int16_t)m; X[IDX(i, j)] = (int16_t)x; Y[IDX(i, j)] = (int16_t)y;
            int h = HV(i, j);
            if (h > best) best = h;
        }
    info[0] = 0; info[1] = -1; info[2] = -1; info[3] = -1; info[4] = -1; info[5] = (int16_t)best;
    if (best > 0 && Lq > 0 && T > 0) {
        int ei = -1, ej = -1;
        for (int i = 1; i <= Lq && ei < 0; ++i)
            for (int j = 1; j <= T; ++j)
                if (HV(i, j) == best) { ei = i; ej = j; break; }
        if (ei > 0) {
            int i = ei, j = ej, n = 0;
            /* built back to front */
            uint8_t* qa = malloc((size_t)(Lq + T + 2));
            uint8_t* ta = malloc((size_t)(Lq + T + 2));
            for (int it = 0; it < Lq + T + 2; ++it) {
                if (i <= 0 || j <= 0) break;
                int h = HV(i, j);
                if (h == 0) break;
                if (h == M[IDX(i, j)]) {
                    qa[n] = q[i - 1]; ta[n] = t[j - 1]; --i; --j; ++n;
                } else if (h == X[IDX(i, j)]) {
                    qa[n] = q[i - 1]; ta[n] = '-'; --i; ++n;
                } else if (h == Y[IDX(i, j)]) {
                    qa[n] = '-'; ta[n] = t[j - 1]; --j; ++n;
                } else {
                    break;
                }
            }
            for (int k2 = 0; k2 < n; ++k2) { q_al[k2] = qa[n - 1 - k2]; t_al[k2] = ta[n - 1 - k2]; }
            free(qa); free(ta);
            if (n > 0) {
                info[0] = n; info[1] = i; info[2] = ei - 1; info[3] = j; info[4] = ej - 1;
            }
        }
    }
#undef HV
#undef IDX
    free(M); free(X); free(Y);
    return best;
}

/* ------------------------------------------------------------------------ */
/* Synthetic database generator (same integer algorithm as the HIP one)      */
/* ------------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline uint32_t rnd24(uint64_t seed, uint64_t stream, uint64_t ctr) {
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL ^ mix64(stream * 0xD1B54A32D192ED03ULL + ctr));
    return (uint32_t)(h >> 40); /* 24 bits */
}
/* index of the first threshold > r in a cumulative 24-bit table of n entries */
static inline int draw(const int32_t* cdf, int n, uint32_t r) {
    int a = 0;
    while (a < n - 1 && (int32_t)r >= cdf[a]) ++a;
    return a;
}

#define AA "ARNDCQEGHILKMFPSTWYV"

/* Parameters shared with the HIP generator (see sseqs_amd/synth.py):
 *   bg_cdf[20]            background residue CDF (24-bit)
 *   sub_cdf[20*20]        substitution CDF per original residue (row a)
 *   len_class_count[K], len_class_len[K]   sorted length classes
 *   p_indel, p_ins, p_ext 24-bit thresholds; homolog every 20th sequence
 */
typedef struct {
    uint64_t seed;
    int32_t bg_cdf[20];
    int32_t sub_cdf[400];
    int32_t p_indel, p_ins, p_ext, id_lo, id_span; /* id in 1/2^24 units */
    int32_t hom_period;
} orc_gen_params;

/* Length of sequence i: classes are consecutive runs of equal length. */
static int seq_length(const int64_t* cls_start, const int32_t* cls_len, int K, int64_t i, int* cls) {
    int lo = 0, hi = K - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (cls_start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    *cls = lo;
    return cls_len[lo];
}

/* Homolog slot: sequence i is the homolog of its period-block when
 * i % P == rnd(block) % P. */
static int is_homolog(const orc_gen_params* gp, int64_t i, int64_t* h) {
    int64_t blk = i / gp->hom_period;
    int64_t slot = (int64_t)(mix64(gp->seed ^ (0xA5A5A5A5ULL + (uint64_t)blk)) % (uint64_t)gp->hom_period);
    *h = blk;
    return (i % gp->hom_period) == slot;
}

/* Mutated copy of query q for homolog h; returns its length, writing at most
 * cap residues into out (out may be NULL to measure only). */
static int mutate(const orc_gen_params* gp, int64_t h, const uint8_t* q, int Lq, uint8_t* out, int cap) {
    uint64_t S = (1ULL << 61) | (uint64_t)h;
    uint64_t c = 0;
    uint32_t idr = rnd24(gp->seed, S, c++);
    int32_t id = gp->id_lo + (int32_t)(((uint64_t)idr * (uint64_t)gp->id_span) >> 24);
    int pos = 0, n = 0;
    while (pos < Lq && n < cap) {
        if ((int32_t)rnd24(gp->seed, S, c++) < gp->p_indel) {
            int len = 1;
            while (len < 64 && (int32_t)rnd24(gp->seed, S, c++) < gp->p_ext) ++len;
            if ((int32_t)rnd24(gp->seed, S, c++) < gp->p_ins) {
                for (int k = 0; k < len && n < cap; ++k) {
                    int a = draw(gp->bg_cdf, 20, rnd24(gp->seed, S, c++));
                    if (out) out[n] = (uint8_t)AA[a];
                    ++n;
                }
            } else {
                pos += len;
            }
            continue;
        }
        int a = q[pos];
        if ((int32_t)rnd24(gp->seed, S, c++) >= id)
            a = draw(gp->sub_cdf + 20 * a, 20, rnd24(gp->seed, S, c++));
        if (out) out[n] = (uint8_t)AA[a];
        ++n;
        ++pos;
    }
    return n;
}

/* Residue p (0-based) of sequence i (length L).  Homologs embed the mutated
 * query (or a window of it) in background residues. */
static uint8_t gen_residue(const orc_gen_params* gp, int64_t i, int p) {
    int a = draw(gp->bg_cdf, 20, rnd24(gp->seed, (uint64_t)i, (uint64_t)p));
    return (uint8_t)AA[a];
}

/* Generate the packed DB.  queries: concatenated query codes (0..19) with
 * q_off[nq+1]; cls_start[K+1] (cumulative counts) and cls_len[K] describe the
 * length-sorted layout; out_ascii receives sum(L_i+1) symbols (each sequence
 * followed by '@').  Homolog of block h copies query (h * 2654435761) % nq. */
void orc_gen_db(const orc_gen_params* gp, const uint8_t* qcodes, const int64_t* q_off, int nq,
                const int64_t* cls_start, const int32_t* cls_len, const int64_t* cls_sym, int K,
                int64_t n, uint8_t* out_ascii, int32_t* out_hom_query) {
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        int cls;
        int L = seq_length(cls_start, cls_len, K, i, &cls);
        int64_t off = cls_sym[cls] + (i - cls_start[cls]) * (int64_t)(L + 1);
        uint8_t* o = out_ascii + off;
        int64_t h;
        int hom = is_homolog(gp, i, &h) && nq > 0;
        int qi = hom ? (int)((uint64_t)(h * 2654435761ULL) % (uint64_t)nq) : -1;
        if (out_hom_query) out_hom_query[i] = qi;
        for (int p = 0; p < L; ++p) o[p] = gen_residue(gp, i, p);
        if (hom) {
            const uint8_t* q = qcodes + q_off[qi];
            int Lq = (int)(q_off[qi + 1] - q_off[qi]);
            int m = mutate(gp, h, q, Lq, NULL, 4096);
            uint64_t S = (1ULL << 62) | (uint64_t)h;
            if (m <= L) {
                int shift = (int)(((uint64_t)rnd24(gp->seed, S, 0) * (uint64_t)(L - m + 1)) >> 24);
                mutate(gp, h, q, Lq, o + shift, m);
            } else {
                uint8_t* tmp = malloc((size_t)m);
                mutate(gp, h, q, Lq, tmp, m);
                int shift = (int)(((uint64_t)rnd24(gp->seed, S, 0) * (uint64_t)(m - L + 1)) >> 24);
                memcpy(o, tmp + shift, (size_t)L);
                free(tmp);
            }
        }
        o[L] = '@';
    }
}

/* Sequences ids[0..m) only: row r of out (stride 1024) receives sequence
 * ids[r] (ASCII, no delimiter), out_len[r] its length.  Same algorithm as
 * orc_gen_db, for spot checks of very large databases. */
void orc_gen_seqs(const orc_gen_params* gp, const uint8_t* qcodes, const int64_t* q_off, int nq,
                  const int64_t* cls_start, const int32_t* cls_len, int K, const int64_t* ids,
                  int64_t m, uint8_t* out, int32_t* out_len) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < m; ++r) {
        int64_t i = ids[r];
        int cls;
        int L = seq_length(cls_start, cls_len, K, i, &cls);
        uint8_t* o = out + r * 1024;
        out_len[r] = L;
        int64_t h;
        int hom = is_homolog(gp, i, &h) && nq > 0;
        for (int p = 0; p < L; ++p) o[p] = gen_residue(gp, i, p);
        if (!hom) continue;
        int qi = (int)((uint64_t)(h * 2654435761ULL) % (uint64_t)nq);
        const uint8_t* q = qcodes + q_off[qi];
        int Lq = (int)(q_off[qi + 1] - q_off[qi]);
        int mm = mutate(gp, h, q, Lq, NULL, 4096);
        uint64_t S = (1ULL << 62) | (uint64_t)h;
        if (mm <= L) {
            int shift = (int)(((uint64_t)rnd24(gp->seed, S, 0) * (uint64_t)(L - mm + 1)) >> 24);
            mutate(gp, h, q, Lq, o + shift, mm);
        } else {
            uint8_t* tmp = malloc((size_t)mm);
            mutate(gp, h, q, Lq, tmp, mm);
            int shift = (int)(((uint64_t)rnd24(gp->seed, S, 0) * (uint64_t)(mm - L + 1)) >> 24);
            memcpy(o, tmp + shift, (size_t)L);
            free(tmp);
        }
    }
}

/* Background query of length L (codes 0..19) from stream 2^60 | qid. */
void orc_gen_query(const orc_gen_params* gp, int64_t qid, int L, uint8_t* codes) {
    for (int p = 0; p < L; ++p)
        codes[p] = (uint8_t)draw(gp->bg_cdf, 20, rnd24(gp->seed, (1ULL << 60) | (uint64_t)qid, (uint64_t)p));
}
