// A3M text of one query's backtracked hits, on the host (sseqs/to_a3m.py:16-72).
//
// Each row is built as sseqs/to_a3m.py:23-49 does:
// - a '-' prefix up to the first occurrence of the aligned query fragment
//   (gaps removed) in the query;
// - the aligned columns, with the target letter upper-cased under a query
//   residue, lower-cased under a query gap, and '-' kept;
// - a '-' suffix to the query's length.
// Rows are emitted in the caller's order, and skipped for score 0, an empty
// alignment, a '@', or a repeat (sseqs/to_a3m.py:56-70). The order is the
// caller's because the reference's is numpy's argsort of the float scores,
// reversed, whose tie order only numpy defines.
//
// The case mapping is ASCII's. A row or query holding a byte >= 128 makes the
// call return -2 so that the caller can use the Python path, whose str.upper
// and str.lower also map Latin-1 letters.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <string_view>
#include <unordered_set>

#include "sseqs_amd.h"

extern "C" int64_t sseqs_a3m_bound(int64_t Lq, int64_t nrow, int64_t aligned_total) {
    if (Lq < 0 || nrow < 0 || aligned_total < 0) return -1;
    // ">original_query\n" + query + "\n", then per row a header of at most
    // ">target_" + 19 digits + " score=" + 11 + ".0\n" (48 bytes) and the
    // row + "\n"; a row is at most Lq plus its aligned length long
    return 17 + (Lq + 1) + nrow * (48 + Lq + 1) + aligned_total;
}

extern "C" int64_t sseqs_a3m_text(const char* h_query, int64_t Lq, const uint8_t* h_q,
                                  const uint8_t* h_t, const int64_t* h_row_off,
                                  const int32_t* h_row_len, const int32_t* h_score,
                                  const int64_t* h_order, int64_t nrow, char* h_out, int64_t cap) {
    const char* query = h_query;
    const uint8_t *q_buf = h_q, *t_buf = h_t;
    const int64_t *row_off = h_row_off, *order = h_order;
    const int32_t *row_len = h_row_len, *score = h_score;
    char* out = h_out;
    if (Lq < 0 || nrow < 0 || (nrow > 0 && (!q_buf || !t_buf || !row_off || !row_len || !score ||
                                            !order)) || !out)
        return -4;
    int64_t total = 0;
    for (int64_t b = 0; b < nrow; ++b) total += row_len[b] > 0 ? row_len[b] : 0;
    if (cap < sseqs_a3m_bound(Lq, nrow, total)) return -3;
    for (int64_t i = 0; i < Lq; ++i)
        if ((uint8_t)query[i] >= 128) return -2;
    const std::string_view q(query, (size_t)Lq);
    char* w = out;
    std::memcpy(w, ">original_query\n", 16);
    w += 16;
    std::memcpy(w, query, (size_t)Lq);
    w += Lq;
    *w++ = '\n';
    // rows already written, as views into the output
    std::unordered_set<std::string_view> seen;
    seen.reserve((size_t)nrow * 2 + 1);
    uint8_t up[256], low[256];
    for (int c = 0; c < 256; ++c) {
        up[c] = (uint8_t)(c >= 'a' && c <= 'z' ? c - 32 : c);
        low[c] = (uint8_t)(c >= 'A' && c <= 'Z' ? c + 32 : c);
    }
    std::string compact, row;
    for (int64_t i = 0; i < nrow; ++i) {
        const int64_t b = order[i];
        if (b < 0 || b >= nrow) return -4;
        const int32_t n = row_len[b];
        if (score[b] == 0 || n <= 0) continue;
        const uint8_t* qa = q_buf + row_off[b];
        const uint8_t* ta = t_buf + row_off[b];
        compact.resize((size_t)n);
        size_t m = 0;
        uint8_t hi = 0;
        for (int32_t c = 0; c < n; ++c) {
            hi |= qa[c] | ta[c];
            compact[m] = (char)qa[c];
            m += qa[c] != '-';
        }
        if (hi >= 128) return -2;
        const size_t start = q.find(std::string_view(compact.data(), m));
        if (start == std::string_view::npos) return -1;
        const int64_t tail = Lq - (int64_t)(start + m);
        const size_t len = start + (size_t)n + (size_t)(tail > 0 ? tail : 0);
        row.resize(len);
        char* r = row.data();
        std::memset(r, '-', start);
        for (int32_t c = 0; c < n; ++c) r[start + c] = (char)(qa[c] != '-' ? up[ta[c]] : low[ta[c]]);
        if (tail > 0) std::memset(r + start + n, '-', (size_t)tail);
        if (std::memchr(r + start, '@', (size_t)n) || seen.count(std::string_view(r, len))) continue;
        // the header's score is Python's repr of the integer-valued float
        w += std::snprintf(w, 64, ">target_%lld score=%d.0\n", (long long)i, (int)score[b]);
        std::memcpy(w, r, len);
        seen.insert(std::string_view(w, len));
        w += len;
        *w++ = '\n';
    }
    return (int64_t)(w - out);
}
