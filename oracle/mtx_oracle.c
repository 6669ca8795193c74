/*
 * mtx_oracle.c -- CPU restatement of the reference's MatrixMarket ingestion.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle library; see hb_oracle.h).
 *
 * Follows, sequentially:
 *   src/graph/cache/parser_mtx/producers.rs:15-75    header and size line
 *   src/graph/cache/parser_mtx/producers.rs:76-234   entry lines: trim, skip blank and
 *                                                    '%' lines, i j 1-based, emit i->j
 *                                                    and j->i when not "general" and i != j
 *   src/graph/cache/parser_mtx/mod.rs:34-162         degree count, prefix sum, scatter
 *   src/graph/cache/impl_sort.rs:10-51               each list sorted ascending
 * Whitespace is ASCII (space, \t, \n, \v, \f, \r); tokens parse as Rust's usize does
 * (optional '+', decimal digits, no overflow).  Every entry line is used: the
 * reference's threads stop early on nnz only by timing, so the two agree whenever the
 * file holds at most nnz entries.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hb_oracle.h"

static int is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

/* one line [b, e): trimmed bounds */
static void trim(const char *t, uint64_t *b, uint64_t *e)
{
    while (*b < *e && is_ws(t[*b])) ++*b;
    while (*e > *b && is_ws(t[*e - 1])) --*e;
}

/* next whitespace-separated token in [*p, e); returns 0 if none */
static int next_tok(const char *t, uint64_t *p, uint64_t e, uint64_t *tb, uint64_t *te)
{
    while (*p < e && is_ws(t[*p])) ++*p;
    if (*p >= e) return 0;
    *tb = *p;
    while (*p < e && !is_ws(t[*p])) ++*p;
    *te = *p;
    return 1;
}

/* Rust usize::from_str: optional '+', then >= 1 decimal digit, no overflow */
static int parse_usize(const char *t, uint64_t b, uint64_t e, uint64_t *out)
{
    if (b < e && t[b] == '+') ++b;
    if (b >= e) return 0;
    uint64_t v = 0;
    for (uint64_t i = b; i < e; ++i) {
        const char c = t[i];
        if (c < '0' || c > '9') return 0;
        const uint64_t d = (uint64_t)(c - '0');
        if (v > (UINT64_MAX - d) / 10) return 0;
        v = v * 10 + d;
    }
    *out = v;
    return 1;
}

static int tok_ieq(const char *t, uint64_t b, uint64_t e, const char *word)
{
    const uint64_t n = strlen(word);
    if (e - b != n) return 0;
    for (uint64_t i = 0; i < n; ++i) {
        char c = t[b + i];
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
        if (c != word[i]) return 0;
    }
    return 1;
}

/* end of the line starting at b (index of '\n' or len) */
static uint64_t line_end(const char *t, uint64_t len, uint64_t b)
{
    const char *nl = (const char *)memchr(t + b, '\n', len - b);
    return nl ? (uint64_t)(nl - t) : len;
}

static int cmp_u32(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

int hbo_mtx_header(const char *t, uint64_t len, hbo_mtx_info *info)
{
    memset(info, 0, sizeof *info);
    uint64_t e = line_end(t, len, 0);
    if (len < 14 || memcmp(t, "%%MatrixMarket", 14) != 0) return HBO_MTX_BAD_HEADER;
    uint64_t p = 0, tb[5], te[5];
    int k = 0;
    while (k < 5 && next_tok(t, &p, e, &tb[k], &te[k])) ++k;
    if (k < 5) return HBO_MTX_MALFORMED_HEADER;
    if (!tok_ieq(t, tb[1], te[1], "matrix") || !tok_ieq(t, tb[2], te[2], "coordinate") ||
        !(tok_ieq(t, tb[3], te[3], "pattern") || tok_ieq(t, tb[3], te[3], "integer")) ||
        !(tok_ieq(t, tb[4], te[4], "symmetric") || tok_ieq(t, tb[4], te[4], "skew-symmetric") ||
          tok_ieq(t, tb[4], te[4], "general")))
        return HBO_MTX_UNSUPPORTED;
    info->symmetric = !tok_ieq(t, tb[4], te[4], "general");
    uint64_t line = 1;
    uint64_t b = e < len ? e + 1 : len;
    for (;;) {
        if (b >= len) return HBO_MTX_NO_SIZE_LINE;
        ++line;
        uint64_t le = line_end(t, len, b), lb = b, lt = le;
        b = le < len ? le + 1 : len;
        trim(t, &lb, &lt);
        if (lb == lt || t[lb] == '%') continue;
        uint64_t q = lb, xb, xe, v[3];
        for (int i = 0; i < 3; ++i) {
            if (!next_tok(t, &q, lt, &xb, &xe) || !parse_usize(t, xb, xe, &v[i])) {
                info->err_line = line;
                return HBO_MTX_BAD_SIZE_LINE;
            }
        }
        if (v[0] != v[1]) {
            info->err_line = line;
            return HBO_MTX_NOT_SQUARE;
        }
        info->n = v[0];
        info->nnz = v[2];
        info->body = b;
        info->body_line = line + 1;
        return 0;
    }
}

int hbo_parse_mtx(const char *t, uint64_t len, hbo_mtx_info *info, uint64_t *offsets, uint32_t *nbrs)
{
    int rc = hbo_mtx_header(t, len, info);
    if (rc) return rc;
    const uint64_t n = info->n;
    uint64_t *deg = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    uint64_t *pos = NULL;
    if (!deg) return HBO_MTX_NOMEM;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            if (!offsets) break;
            uint64_t sum = 0;
            for (uint64_t u = 0; u <= n; ++u) {
                offsets[u] = sum;
                sum += deg[u];
            }
            if (!nbrs) break;
            pos = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
            if (!pos) { free(deg); return HBO_MTX_NOMEM; }
            memcpy(pos, offsets, sizeof(uint64_t) * (n + 1));
        }
        uint64_t b = info->body, line = info->body_line - 1, entries = 0, arcs = 0;
        while (b < len) {
            ++line;
            uint64_t le = line_end(t, len, b), lb = b, lt = le;
            b = le < len ? le + 1 : len;
            trim(t, &lb, &lt);
            if (lb == lt || t[lb] == '%') continue;
            uint64_t q = lb, xb, xe, i1, j1;
            next_tok(t, &q, lt, &xb, &xe);
            if (!parse_usize(t, xb, xe, &i1)) { rc = HBO_MTX_BAD_ENTRY; break; }
            if (!next_tok(t, &q, lt, &xb, &xe)) continue;  /* one token: the line is skipped */
            if (!parse_usize(t, xb, xe, &j1)) { rc = HBO_MTX_BAD_ENTRY; break; }
            if (i1 == 0 || j1 == 0) { rc = HBO_MTX_ZERO_INDEX; break; }
            const uint64_t i = i1 - 1, j = j1 - 1;
            if (i >= n || j >= n) { rc = HBO_MTX_OUT_OF_BOUNDS; break; }
            ++entries;
            if (pass == 0) {
                deg[i] += 1;
                if (info->symmetric && i != j) deg[j] += 1;
            } else {
                nbrs[pos[i]++] = (uint32_t)j;
                if (info->symmetric && i != j) nbrs[pos[j]++] = (uint32_t)i;
            }
            arcs += (info->symmetric && i != j) ? 2 : 1;
        }
        if (rc) {
            info->err_line = line;
            break;
        }
        info->entries = entries;
        info->arcs = arcs;
    }
    if (!rc && nbrs && offsets)
        for (uint64_t u = 0; u < n; ++u)
            qsort(nbrs + offsets[u], offsets[u + 1] - offsets[u], sizeof(uint32_t), cmp_u32);
    free(pos);
    free(deg);
    return rc;
}
