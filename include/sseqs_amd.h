/*
 * sseqs_amd.h -- C ABI of the MI355X (gfx950) MSA-search library
 * (sseqs_amd/libsseqs_amd.so).
 *
 * Conventions shared by every entry point:
 *   - All pointers are device pointers unless the parameter name starts with
 *     "h_".  Nothing is synchronised: work is enqueued on `stream`.
 *   - The return value is 0 on success.  A negative value is an argument
 *     error detected on the host (nothing was launched); a positive value is
 *     a hipError_t from a launch.  sseqs_last_error() describes the last
 *     failure on the calling thread.  The reference skipped unsupported
 *     sizes silently (sseqs/sw.cu:357-369, sseqs/filter.cu:320); these
 *     entry points report them instead.
 *   - fp16 buffers are IEEE binary16 bit patterns (uint16_t).
 *   - Sizes are int64_t where the reference used int, so one call can cover
 *     an HBM-resident database of more than 2^31 symbols.
 *
 * Entry points that replace a reference launcher keep its name and argument
 * order, adding the stream where the reference had none; each cites the
 * launcher it replaces.  The sseqs_* entry points have no single reference
 * counterpart: they are the database-resident forms of the same operators.
 */
#ifndef SSEQS_AMD_H
#define SSEQS_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message for the last non-zero return on this thread ("" if none). */
const char* sseqs_last_error(void);
/* Library ABI version (bumped when a signature changes). */
int sseqs_abi_version(void);

/* ---------------------------------------------------------------- xbit -- */

/* Unpack n_syms 5-bit symbols starting at byte `lo` of d_in into ASCII
 * (symbol + 64) at d_out[0..n_syms).
 * Replaces decompress_launcher (sseqs/xbit.cu:65-71), which ran on the
 * legacy default stream and wrote whole 8-symbol groups past n_syms. */
int decompress_launcher(const uint8_t* d_in, uint8_t* d_out, int64_t n_syms, int64_t lo,
                        hipStream_t stream);

/* Pack n_syms ASCII symbols into (5*n_syms+7)/8 bytes, MSB first.
 * Replaces compress_launcher (sseqs/xbit.cu:73-78). */
int compress_launcher(const uint8_t* d_in, uint8_t* d_out, int64_t n_syms, hipStream_t stream);

/* ------------------------------------------------------ prefilter (conv) -- */

/* Low-rank FFT-correlation prefilter score of B segments of `flat` (ASCII),
 * segment b = flat[offsets[b] .. offsets[b]+lengths[b]).  corr[b] (fp16) is
 * overwritten with max over even m of the N-point circular correlation
 * 0.5 * sum_r sum_t A_r[qrev[t]] * B_r[code(seg[(m-t) mod N])]
 * with A = weight_LUTS_A (k rows of vocab_size floats, query side) and
 * B = W_db (database side, read as a flat array: code 21 reads the next
 * row's column 0, so W_db must hold k*vocab_size+1 floats).
 * Replaces launch_warp_corr (sseqs/filter.cu:161-174, kernels
 * sseqs/filter.cu:92-159 and sseqs/fft.cu:17-415).  out_stats and
 * query_rfft are accepted for signature compatibility and not used.
 * N may be any power of two from 16 to 2048 (the reference handled 64..1024
 * and left scores at 0 for 2048). */
int launch_warp_corr(uint16_t* corr, const uint8_t* query_reversed_indices,
                     const float* weight_LUTS_A, const float* W_db, const uint8_t* flat,
                     const int32_t* offsets, const int32_t* lengths, float* out_stats, int B,
                     int vocab_size, int k, int query_len, int maxlen, int N_full,
                     hipStream_t stream, uint16_t* query_rfft);

/* ---------------------------------------------------------- Smith-Waterman -- */

/* Linear-gap local alignment score of M targets against the query
 * (codes 0..19, length 1..2048; the reference stopped at 1024,
 * sseqs/sw.cu:355-369).  Target g is
 * ascii[P[good_idx[g]-1]+1 .. P[good_idx[g]]) with P = `lengths` (the
 * reference's name for the delimiter-position array); P[-1] is taken as -1.
 * out[g] = -1 for an empty query or target, like the reference.
 * Scores follow the reference's 32-stripe rule exactly (no horizontal move
 * into a column j with j % ceil(Lq/32) == 0).  gap_val must be <= 0.
 * Replaces launch_sw_cuda_blosum62 (sseqs/sw.cu:336-370). */
int launch_sw_cuda_blosum62(const uint8_t* query_seq_indices_ptr, int query_length,
                            const int32_t* good_idx, const uint8_t* ascii, const int32_t* lengths,
                            int num_sequences_in_db, int32_t* output_scores_ptr, int gap_val,
                            hipStream_t stream);

/* Affine-gap (open gap_open, extend gap_extend, both >= 0) counterpart with
 * the same target selection and stripe rule.
 * Replaces launch_sw_cuda_affine (sseqs/sw.cu:372-425). */
int launch_sw_cuda_affine(const uint8_t* query_seq_indices_ptr, int query_length,
                          const int32_t* good_idx, const uint8_t* ascii, const int32_t* lengths,
                          int num_sequences_in_db, int32_t* output_scores_ptr, int gap_open,
                          int gap_extend, hipStream_t stream);

/* ------------------------------------------------ HBM-resident database -- */

/* A packed database buffer holds ceil(5*n_syms/8) bytes followed by at least
 * 64 readable bytes: the kernels below read whole 16-byte words and 8-symbol
 * windows, which may extend up to 40 bytes past a sequence's last residue. */

/* Index of a packed database holding n_syms symbols: every maximal run of
 * non-'@' symbols is one sequence.  Writes up to `cap` entries of
 * seq_off (symbol offset of the first residue) and seq_len, and the number
 * of sequences found to *d_nseq (device int64).  Empty runs (consecutive
 * '@') are skipped, matching the `lengths != 0` filter of sseqs/msa.py:253.
 * ws must hold sseqs_db_index_ws_bytes(n_syms) bytes; packed must be 4-byte
 * aligned.  With cap == 0 only the count is written.
 * Database-resident form of the delimiter scan of sseqs/msa.py:227-267. */
size_t sseqs_db_index_ws_bytes(int64_t n_syms);
int sseqs_db_index(const uint8_t* packed, int64_t n_syms, int64_t* seq_off, int32_t* seq_len,
                   int64_t cap, int64_t* d_nseq, void* ws, size_t ws_bytes, hipStream_t stream);

/* launch_warp_corr's score for sequences s in [s_begin, s_end) of the
 * resident database (symbols unpacked on the fly from `packed`), all with
 * correlation size N, written as fp32 to scores[s - s_begin].  qrev (codes
 * 0..20), lutA and lutB (k rows of vocab_size floats, flat indexing as in
 * launch_warp_corr) are device arrays.  Sequences longer than
 * min(maxlen, N) are truncated to it, as in sseqs/fft.cu:124-126.
 * N: power of two, 16..2048; k: 1..20.  Fused unpack + prefilter of the
 * resident database. */
int sseqs_db_prefilter(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                       int64_t s_begin, int64_t s_end, const uint8_t* qrev, int query_len,
                       const float* lutA, const float* lutB, int vocab_size, int k, int N,
                       int maxlen, float* scores, hipStream_t stream);

/* sseqs_db_prefilter for (offset, length) spans that may hold '@' inside:
 * the per-batch cascade's segments of a batch whose `lengths != 0` filter
 * shifted its offsets (sseqs/msa.py:253,266-267), where the reference's
 * ASCII prefilter scores an in-span '@' through LUT code 21
 * (sseqs/fft.cu:5-15,91-93).  Always uses the generic kernels, which do the
 * same; arguments as sseqs_db_prefilter. */
int sseqs_db_prefilter_spans(const uint8_t* packed, const int64_t* seq_off,
                             const int32_t* seq_len, int64_t s_begin, int64_t s_end,
                             const uint8_t* qrev, int query_len, const float* lutA,
                             const float* lutB, int vocab_size, int k, int N, int maxlen,
                             float* scores, hipStream_t stream);

/* Name of the kernel sseqs_db_prefilter launches for these sizes (for
 * profiles and reports), written to name[0..name_len).  Host only. */
int sseqs_db_prefilter_kernel(int query_len, int k, int N, int maxlen, char* name, int name_len);

/* Hit gather: row i of out (n rows of `width` bytes) receives the first
 * min(seq_len[ids[i]], width) residues of database sequence ids[i] as
 * ASCII, and out_len[i] that count.  Database-resident form of the hit
 * gather of sseqs/msa.py:349-370 (width 500 there, sseqs/msa.py:183). */
int sseqs_db_gather(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                    const int64_t* ids, int64_t n, int width, uint8_t* out, int32_t* out_len,
                    hipStream_t stream);

/* Affine score (reference stripe rule) of database sequences cand[0..n) of
 * the resident database against the query (codes 0..19, length 1..2048).
 * out[i] = -1 for an empty sequence. */
int sseqs_db_sw_affine(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                       const int64_t* cand, int64_t n, const uint8_t* query, int query_len,
                       int gap_open, int gap_extend, int32_t* out, hipStream_t stream);

/* Linear-gap counterpart (gap_val <= 0 per gap position, reference stripe
 * rule of launch_sw_cuda_blosum62) of sseqs_db_sw_affine.  Database-resident
 * form of sw.sw_cuda_blosum62 (sseqs/sw.cu:55-179), for the linear-gap run
 * of the at-scale cascade. */
int sseqs_db_sw_linear(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                       const int64_t* cand, int64_t n, const uint8_t* query, int query_len,
                       int gap_val, int32_t* out, hipStream_t stream);

/* Batched affine SW (reference stripe rule, gap_open/gap_extend in
 * 0..16000) of many queries' candidate lists: query q is
 * qcodes[q_off[q] .. q_off[q+1]) (codes 0..19, 1..1024 residues), its
 * candidates cand[c_off[q] .. c_off[q+1]) (database ids), and the score of
 * cand[c] goes to out[c] (-1 for an empty sequence), equal to what
 * sseqs_db_sw_affine gives for that query.  One persistent launch over the
 * work queue of one length bucket: item i is candidates
 * [c_off[q] + item_j[i] * item, min(+ item, c_off[q+1])) of query
 * q = item_q[i], and every item's query must have stripe width
 * ceil(Lq / 32) == tpt (1..32).  grid blocks of 256 threads (0: 1024) take
 * items in order through the device counter `counter` (one uint32, reset
 * by the call's first kernel).  Device arrays: q_off, c_off (int64),
 * item_q, item_j (int32).  The database-resident SW of the batched sweep;
 * no single reference counterpart (the reference scores one query per
 * launch, sseqs/sw.cu:372-425). */
int sseqs_db_sw_batch(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                      const uint8_t* qcodes, const int64_t* q_off, const int64_t* cand,
                      const int64_t* c_off, const int32_t* item_q, const int32_t* item_j,
                      int64_t n_items, int item, int tpt, int grid, int gap_open,
                      int gap_extend, int32_t* out, uint32_t* counter, hipStream_t stream);

/* ---------------------------------------------------------------- top-k -- */

/* dtype codes for sseqs_topk */
#define SSEQS_F32 0
#define SSEQS_F16 1
#define SSEQS_I32 2

/* Indices of the k largest of n scores, ties broken by the smaller index
 * (torch.topk leaves tie order unspecified; this order is the defined one).
 * sorted != 0: out_idx[0..k) in (score desc, index asc) order, k <= 16384;
 * sorted == 0: out_idx[0..k) in ascending index order, any k <= n.
 * NaN ranks below every number.  ws: sseqs_topk_ws_bytes(n) bytes. */
size_t sseqs_topk_ws_bytes(int64_t n);
int sseqs_topk(const void* scores, int dtype, int64_t n, int64_t k, int sorted, int64_t* out_idx,
               void* ws, size_t ws_bytes, hipStream_t stream);

/* sseqs_topk without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference.  Same arguments, order and results as sseqs_topk; the
 * workspace is a library-owned static device array of 512 KiB per GPU
 * (held until the process ends), so n is limited to
 * sseqs_topk_noalloc_max_n() (393 M scores) and larger n is rejected with
 * an error.  It calls no hipMalloc/hipFree.  Calls on one GPU share that
 * array: calls that overlap on one GPU (other streams or threads) race. */
int64_t sseqs_topk_noalloc_max_n(void);
int sseqs_topk_noalloc(const void* scores, int dtype, int64_t n, int64_t k, int sorted,
                       int64_t* out_idx, hipStream_t stream);

/* sseqs_topk over many segments in one launch: the per-batch selects of the
 * reference cascade (sseqs/msa.py:297-310, 324, 335) for all batches of a
 * sweep.  Segment j holds scores[seg_begin[j] .. + seg_n[j]) and selects its
 * seg_k[j] largest (ties to the smaller index; sorted != 0: ordered by
 * (score desc, index asc), seg_k <= 4096; sorted == 0: ascending index).
 * Entries drop_front .. seg_k[j] - drop_back - 1 of that order are written
 * to out_idx[out_base[j] ..], as indices within the segment, plus
 * seg_begin[j] when add_begin != 0.  max_k bounds every seg_k[j] (checked
 * against the sorted limit on the host).  Device arrays: seg_begin,
 * out_base (int64), seg_n, seg_k (int32). */
int sseqs_topk_segments(const void* scores, int dtype, const int64_t* seg_begin,
                        const int32_t* seg_n, const int32_t* seg_k, const int64_t* out_base,
                        int64_t nseg, int max_k, int sorted, int drop_front, int drop_back,
                        int add_begin, int64_t* out_idx, hipStream_t stream);

/* ------------------------------------------------------------ backtrack -- */

/* Full affine DP (gaps open from H, -5 for non-residues) and stateless
 * traceback of B targets (ASCII, target b = targets[t_off[b] .. +t_len[b]))
 * against the ASCII query, with the semantics of the Triton kernels
 * sseqs/sw.py:408-768 (driven by sw_affine_backtrack, sseqs/sw.py:27-219).
 * Per target b, with A = max_align:
 *   out_q[b*A + A-len .. b*A + A) / out_t[...]: aligned query / target
 *   (ASCII, '-' for gaps), right-aligned like the reference;
 *   info[b*6 + 0..5] = {len, q_start, q_end, t_start, t_end, score}
 *   (starts/ends -1 and len 0 when nothing aligns).
 * max_align must be >= query_len + max(t_len) + 1.
 * cell_off[b] = query_len * sum_{b' < b} t_len[b'] and total_cells is the
 * full sum (host value); ws: sseqs_backtrack_ws_bytes(query_len,
 * total_cells, B) bytes. */
size_t sseqs_backtrack_ws_bytes(int query_len, int64_t total_cells, int B);
int sseqs_backtrack(const uint8_t* query, int query_len, const uint8_t* targets,
                    const int64_t* t_off, const int32_t* t_len, const int64_t* cell_off,
                    int64_t total_cells, int B, int gap_open, int gap_extend, int max_align,
                    uint8_t* out_q, uint8_t* out_t, int32_t* info, void* ws, size_t ws_bytes,
                    hipStream_t stream);

/* -------------------------------------------------- learned prefilter -- */

/* The encoder of the learned prefilter and its candidate proposal.  None of
 * these has a reference counterpart: the reference has only the FFT
 * prefilter (sseqs/msa.py:270-310), whose stage-1 selection the proposal
 * replaces.  Encoder: embedding 22 x 64 (codes 0..19 residues, 20 other,
 * 21 pad; rows 20, 21 zero), three conv1d layers of width 9
 * (64 -> 128 -> 128 -> 256) with GELU (tanh form), mean pool, L2 norm.
 * A batch is a flat array of M rows (M a multiple of 128): four pad rows
 * (mask 0) before every sequence and after the last.  Activation buffers
 * are bf16 row-major [rows][C]; a pointer "to row 0" of a buffer that a
 * conv reads as its window source (A, X) must have four readable zero
 * rows before row 0 and after row M - 1.  Scores and gradients are fp32.
 *
 * sseqs_enc_conv: C[p][n] = sum_k A[(p - 4) * lda + k] * Bt[n * K + k] for
 * p < M, n < N (N a multiple of 64, K of 32), then by `epi`:
 *   0: U[p][n] = C + bias[n] (if U), Y[p][n] = gelu(U) * mask[p];
 *   1: Y[p][n] = C * gelu'(Uprev[p][n]) * mask[p];
 *   2: Yf[p][n] = C * mask[p] (fp32).
 * Forward layers use Bt[o][t * Cin + i] = W[t][i][o]; the input gradient
 * uses A = dU and Bt[i][t * Cout + o] = W[8 - t][i][o]
 * (sseqs_enc_pack_weights writes both from W[t][i][o]). */
int sseqs_enc_conv(const void* A, int lda, const void* Bt, int64_t M, int N, int K,
                   const float* bias, const uint8_t* mask, void* U, void* Y, const void* Uprev,
                   float* Yf, int epi, hipStream_t stream);
/* dW[t * Cin + i][o] += sum_{p < M} X[(p - 4) * Cin + t * Cin + i] * dU[p][o]
 * (Cin, Cout multiples of 64). */
int sseqs_enc_wgrad(const void* X, int Cin, const void* dU, int Cout, int64_t M, float* dW,
                    hipStream_t stream);
/* db[c] += sum_{p < M} dU[p][c] (C a multiple of 64). */
int sseqs_enc_bias_grad(const void* dU, int C, int64_t M, float* db, hipStream_t stream);
/* x0[p] = E[code[p]] (E: fp32 22 x 64; x0 bf16 [M][64]). */
int sseqs_enc_embed(const uint8_t* code, int64_t M, const float* E, void* x0, hipStream_t stream);
/* dE[code[p]] += dx0[p] for codes 0..19 (dx0 fp32 [M][64]). */
int sseqs_enc_embed_grad(const uint8_t* code, int64_t M, const float* dx0, float* dE,
                         hipStream_t stream);
/* e[s] = z / |z| with z = mean of h3 rows row0[s] .. + len[s] (h3 bf16
 * [.][256]); znorm[s] = |z| (may be null). */
int sseqs_enc_pool(const void* h3, const int64_t* row0, const int32_t* len, int64_t nseq,
                   float* e, float* znorm, hipStream_t stream);
/* dU3 on each sequence's rows: ((de - e (e . de)) / |z|) / len * gelu'(U3). */
int sseqs_enc_pool_grad(const float* de, const float* e, const float* znorm, const void* U3,
                        const int64_t* row0, const int32_t* len, int64_t nseq, void* dU3,
                        hipStream_t stream);
/* Squared error of cos_j = e[qi[j]] . e[ti[j]] against y[j], weighted by
 * w[j] (1 when w is null) and averaged over npair; its gradient is added to
 * de, cos_j to cos_out (if not null), the loss to *loss (if not null).  e
 * holds nseq rows; a pair naming a row outside 0..nseq-1 is skipped. */
int sseqs_enc_pair_loss(const float* e, int64_t nseq, const int32_t* qi, const int32_t* ti,
                        const float* y, const float* w, int64_t npair, float* de, float* cos_out,
                        float* loss, hipStream_t stream);
/* Soft-target InfoNCE: rows 0..nq-1 of e are queries, rows nq..nq+nt-1
 * targets (nt <= 8192).  Loss -scale * sum_t P[q][t] log softmax_t(e_q . e_t
 * / tau) summed over the queries whose P row is not all zero, added to
 * *loss; its gradient is added to de.  G is an nq x nt float workspace. */
int sseqs_enc_infonce(const float* e, int nq, int nt, const float* P, float tau, float scale,
                      float* G, float* de, float* loss, hipStream_t stream);
/* AdamW step `step` (from 1) over n fp32 parameters. */
int sseqs_enc_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                   float b2, float eps, float wd, int step, hipStream_t stream);
/* bf16 forward and input-gradient layouts of W[t][i][o] (bwd may be null). */
int sseqs_enc_pack_weights(const float* W, int Cin, int Cout, void* fwd, void* bwd,
                           hipStream_t stream);
/* Codes (0..19, 20 other) and mask 1 of database sequences s_begin + s
 * (s < nseq) at rows row0[s] .. + min(len, max_len); other rows untouched. */
int sseqs_enc_db_rows(const uint8_t* packed, const int64_t* seq_off, const int32_t* seq_len,
                      int64_t s_begin, int64_t nseq, const int64_t* row0, int max_len,
                      uint8_t* code, uint8_t* mask, hipStream_t stream);
/* out[i] = fp8 e4m3 (OCP) of scale * e[i], n a multiple of 4. */
int sseqs_enc_to_fp8(const float* e, int64_t n, float scale, uint8_t* out, hipStream_t stream);
/* scores[q * ld + s] (fp16) = sum_c table[s][c] * q8[q][c] for the n rows of
 * the fp8 table (256 bytes each) and G query rows (G a multiple of 16, at
 * most 128), on MFMA fp8.  table and q8 are 16-byte aligned. */
int sseqs_propose(const uint8_t* table, int64_t n, const uint8_t* q8, int G, void* scores,
                  int64_t ld, hipStream_t stream);
/* out[j] = max over r < nrows of S[r * ld + j] for j < n (fp16; ld >= n and a
 * multiple of 8, S and out 16-byte aligned): a query's score over its
 * window rows. */
int sseqs_rows_max(const void* S, int64_t ld, int nrows, int64_t n, void* out,
                   hipStream_t stream);
/* sseqs_propose fused with sseqs_rows_max: output row o (nslot of them, at
 * most 32) is the maximum of the proposal scores of query rows row_off[o] ..
 * row_off[o + 1] - 1, taken in ascending row order as sseqs_rows_max takes
 * them: scores[o * ld + s] (fp16).  G query rows (a multiple of 16, at most
 * 64); rows outside the runs are padding.  row_off is a host array of
 * nslot + 1 strictly increasing offsets within 0 .. G; only output rows
 * 0 .. nslot-1 are written.  A query whose score is the maximum over its
 * window rows gets one row write per table row instead of one per window. */
int sseqs_propose_max(const uint8_t* table, int64_t n, const uint8_t* q8, int G,
                      const int32_t* row_off, int nslot, void* scores, int64_t ld,
                      hipStream_t stream);

/* ------------------------------------------------------- a3m (host) -- */

/* Bytes of output that sseqs_a3m_text needs at most for nrow rows whose
 * aligned lengths sum to aligned_total (-1 for a negative argument). */
int64_t sseqs_a3m_bound(int64_t Lq, int64_t nrow, int64_t aligned_total);
/* A3M text of one query's backtracked hits, host memory only
 * (sseqs/to_a3m.py:16-72, rows as sseqs/to_a3m.py:23-49).  Row b's aligned
 * query and target bytes are h_q / h_t + h_row_off[b], h_row_len[b] long,
 * its score h_score[b]; rows are taken in the order h_order (a permutation
 * of 0..nrow-1, the reference's np.argsort(scores)[::-1]).  Writes the text
 * (not terminated) to h_out and returns its length; -1 when an aligned
 * query fragment is not in the query (the reference raises), -2 when the
 * query or a row holds a byte >= 128 (the caller formats those), -3 when
 * cap < sseqs_a3m_bound(...), -4 for a bad argument. */
int64_t sseqs_a3m_text(const char* h_query, int64_t Lq, const uint8_t* h_q, const uint8_t* h_t,
                       const int64_t* h_row_off, const int32_t* h_row_len,
                       const int32_t* h_score, const int64_t* h_order, int64_t nrow,
                       char* h_out, int64_t cap);

/* --------------------------------------------------- synthetic database -- */

/* Parameters of the synthetic generator; layout shared with
 * sseqs_amd/synth.py (1712 bytes). */
typedef struct {
    uint64_t seed;
    int32_t bg_cdf[20];
    int32_t sub_cdf[400];
    int32_t p_indel, p_ins, p_ext, id_lo, id_span;
    int32_t hom_period;
} sseqs_gen_params;

/* Length-sorted synthetic database of n sequences.  Length classes:
 * cls_start[K+1] (cumulative sequence counts), cls_len[K], cls_sym[K]
 * (symbol offset of each class).  Queries (codes 0..19) concatenated with
 * q_off[nq+1] are planted as mutated homologs, one per block of
 * hom_period sequences.  Writes packed (ceil(5*n_syms/8) bytes), seq_off,
 * seq_len and hom_query (query id of a homolog, -1 otherwise).
 * ws: sseqs_gen_ws_bytes(n, hom_period) bytes. */
size_t sseqs_gen_ws_bytes(int64_t n, int hom_period);
int sseqs_gen_db(const sseqs_gen_params* h_params, const uint8_t* qcodes, const int64_t* q_off,
                 int nq, const int64_t* cls_start, const int32_t* cls_len, const int64_t* cls_sym,
                 int K, int64_t n, int64_t n_syms, uint8_t* packed, int64_t* seq_off,
                 int32_t* seq_len, int32_t* hom_query, void* ws, size_t ws_bytes,
                 hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SSEQS_AMD_H */
