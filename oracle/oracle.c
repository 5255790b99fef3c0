/*
 * oracle.c -- CPU oracle for the AdMatch scan path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (admatch_amd/) never links or calls it.
 *
 * It is a plain-C, single-threaded restatement of:
 *   - the .nfa parser          src/transition_graph.cpp:82-294 (iNFA kernel
 *                              type), with 32-bit state ids instead of the
 *                              reference's unsigned short (src/common.h:6);
 *   - merge_transitions        src/transition_graph.cpp:369-411 (per-symbol
 *                              sort + unique into one CSR list);
 *   - the iNFA step loop       src/iNFA_kernel.cu:41-172 executed with
 *                              thread_count == 1, so every in-place
 *                              word-boundary pass runs in list order;
 *   - the accept extraction    src/iNFA_kernel.cu:286-307 (ascending final
 *                              states, accept-map lookup, sort + unique);
 *   - is_word_char             src/kernel_common.cu:4-7.
 * State bit vectors are u32 words, bit i in word i/32, LSB first, exactly as
 * src/state_vector.cpp:52-62.
 *
 * Pinning (what checks this file, see tests/test_oracle.py):
 *   - known answers on data/crs/rule1.nfa ((text|application)/xml, unanchored);
 *   - brute-force set simulation on tiny random NFAs (independent Python);
 *   - Python re.search on the synthetic generator's regex source;
 *   - the fixture statistics of data/crs and data/snort (state, transition,
 *     persistent, word-boundary and accept counts).
 *
 * Declared deviations (malformed input only; the reference has undefined
 * behaviour or throws there): empty lines are skipped; a transition whose
 * src or dst is >= the state count, a symbol outside [0,256], a negative or
 * non-numeric id is a load error; an empty string runs no head word-boundary
 * pass (the reference reads input[0] out of range, src/iNFA_kernel.cu:63).
 */
#include <errno.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_SYMBOLS 257 /* 256 byte values + word boundary, src/common.h:13 */
#define ORC_WB 256      /* src/common.h:15 */

typedef struct {
    uint32_t src, dst;
} orc_tr;

typedef struct {
    orc_tr *v;
    int64_t n, cap;
} orc_vec;

typedef struct orc_nfa {
    int32_t S;   /* state count */
    int32_t W32; /* u32 words per state vector */
    uint32_t *init, *persis, *accept;
    orc_tr *list;                 /* per-symbol sorted unique transitions */
    int64_t off[ORC_SYMBOLS + 1]; /* CSR offsets, off[257] = total */
    /* accept map (first line per state wins, transition_graph.cpp:282) */
    int32_t *acc_off;   /* S+1 */
    int32_t *acc_rules; /* rules */
    /* parse statistics */
    int64_t n_persis_lines, n_init_lines, n_accept_lines, n_accept_map;
    int64_t n_raw_transitions, n_bad_lines;
} orc_nfa;

static void vec_push(orc_vec *a, uint32_t s, uint32_t d) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 16;
        a->v = (orc_tr *)realloc(a->v, (size_t)a->cap * sizeof(orc_tr));
    }
    a->v[a->n].src = s;
    a->v[a->n].dst = d;
    a->n++;
}

static int is_blank(char c) { return c == ' ' || c == '\t'; }

/* strict unsigned decimal (optional leading '+'), whole token */
static int parse_u32(const char *b, const char *e, uint32_t *out) {
    if (b < e && *b == '+') b++;
    if (b >= e) return 0;
    uint64_t v = 0;
    for (const char *p = b; p < e; p++) {
        if (*p < '0' || *p > '9') return 0;
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xFFFFFFFFull) return 0;
    }
    *out = (uint32_t)v;
    return 1;
}

/* strict signed decimal, whole token */
static int parse_i64(const char *b, const char *e, int64_t *out) {
    int neg = 0;
    if (b < e && (*b == '+' || *b == '-')) {
        neg = *b == '-';
        b++;
    }
    if (b >= e) return 0;
    int64_t v = 0;
    for (const char *p = b; p < e; p++) {
        if (*p < '0' || *p > '9') return 0;
        v = v * 10 + (*p - '0');
        if (v > (int64_t)1 << 40) return 0;
    }
    *out = neg ? -v : v;
    return 1;
}

static int contains(const char *b, const char *e, const char *needle) {
    size_t n = strlen(needle);
    for (const char *p = b; p + n <= e; p++)
        if (memcmp(p, needle, n) == 0) return 1;
    return 0;
}

static void set_bit32(uint32_t *v, int32_t S, uint32_t i) {
    if (i >= (uint32_t)S) return; /* StateVector::set_bit ignores, state_vector.cpp:53 */
    v[i >> 5] |= 1u << (i & 31);
}

static int cmp_tr(const void *a, const void *b) {
    const orc_tr *x = (const orc_tr *)a, *y = (const orc_tr *)b;
    if (x->src != y->src) return x->src < y->src ? -1 : 1;
    if (x->dst != y->dst) return x->dst < y->dst ? -1 : 1;
    return 0;
}

static void seterr(char *err, int errlen, const char *msg, int64_t lineno) {
    if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s (line %lld)", msg, (long long)lineno);
}

typedef struct {
    uint32_t state;
    int32_t first_rule, n_rules;
} acc_entry;

void orc_destroy(orc_nfa *g) {
    if (!g) return;
    free(g->init);
    free(g->persis);
    free(g->accept);
    free(g->list);
    free(g->acc_off);
    free(g->acc_rules);
    free(g);
}

/* Parse a .nfa file.  Returns NULL and fills err on failure. */
orc_nfa *orc_load(const char *path, char *err, int errlen) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        seterr(err, errlen, "cannot open file", 0);
        return NULL;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)sz + 1);
    if (fread(buf, 1, (size_t)sz, f) != (size_t)sz) {
        fclose(f);
        free(buf);
        seterr(err, errlen, "read error", 0);
        return NULL;
    }
    buf[sz] = 0;
    fclose(f);

    orc_nfa *g = (orc_nfa *)calloc(1, sizeof(orc_nfa));
    orc_vec per[ORC_SYMBOLS];
    memset(per, 0, sizeof(per));
    acc_entry *acc = NULL;
    int64_t nacc = 0, capacc = 0;
    int32_t *rules = NULL;
    int64_t nrules = 0, caprules = 0;
    uint8_t *has_acc = NULL;
    int have_count = 0;
    int64_t lineno = 0;
    char *p = buf, *end = buf + sz;
    int ok = 1;

    while (p < end && ok) {
        char *ls = p;
        char *le = memchr(p, '\n', (size_t)(end - p));
        if (!le) le = end;
        p = le + 1;
        lineno++;
        if (!have_count) {
            /* first non-comment line: "iss >> state_count" (transition_graph.cpp:91-103) */
            if (ls < le && ls[0] == '#') continue;
            const char *q = ls;
            while (q < le && (is_blank(*q) || *q == '\r')) q++;
            const char *qb = q;
            while (q < le && *q >= '0' && *q <= '9') q++;
            int64_t s = 0;
            if (!parse_i64(qb, q, &s) || s <= 0 || s > (1 << 26)) {
                seterr(err, errlen, "bad state count", lineno);
                ok = 0;
                break;
            }
            g->S = (int32_t)s;
            g->W32 = (g->S + 31) / 32;
            g->init = (uint32_t *)calloc((size_t)g->W32, 4);
            g->persis = (uint32_t *)calloc((size_t)g->W32, 4);
            g->accept = (uint32_t *)calloc((size_t)g->W32, 4);
            has_acc = (uint8_t *)calloc((size_t)g->S, 1);
            have_count = 1;
            continue;
        }
        if (ls == le) continue;      /* empty line (declared deviation) */
        if (ls[0] == '#') continue;  /* comment, transition_graph.cpp:131 */
        if (le[-1] == '\r') le--;    /* CRLF, :136-138 */
        if (ls == le) continue;
        /* split on [[:blank:]]*:[[:blank:]]* -- exactly two parts (:142-148) */
        int ncolon = 0;
        char *colon = NULL;
        for (char *q = ls; q < le; q++)
            if (*q == ':') {
                ncolon++;
                colon = q;
            }
        if (ncolon != 1) {
            g->n_bad_lines++;
            continue;
        }
        char *a0 = ls, *a1 = colon;
        while (a1 > a0 && is_blank(a1[-1])) a1--;
        char *b0 = colon + 1, *b1 = le;
        while (b0 < b1 && is_blank(*b0)) b0++;
        /* split part 0 on [[:blank:]]*->[[:blank:]]* (:152) */
        int narrow = 0;
        char *arrow = NULL;
        for (char *q = a0; q + 1 < a1; q++)
            if (q[0] == '-' && q[1] == '>') {
                narrow++;
                arrow = q;
            }
        if (narrow == 1) {
            char *s0 = a0, *s1 = arrow, *d0 = arrow + 2, *d1 = a1;
            while (s1 > s0 && is_blank(s1[-1])) s1--;
            while (d0 < d1 && is_blank(*d0)) d0++;
            uint32_t src, dst;
            if (!parse_u32(s0, s1, &src) || !parse_u32(d0, d1, &dst)) {
                seterr(err, errlen, "bad state id", lineno);
                ok = 0;
                break;
            }
            if (src >= (uint32_t)g->S || dst >= (uint32_t)g->S) {
                seterr(err, errlen, "state id out of range", lineno);
                ok = 0;
                break;
            }
            /* atoms: blank-separated "sym" or "lo|hi" (:161-181) */
            char *q = b0;
            while (q < b1 && ok) {
                while (q < b1 && is_blank(*q)) q++;
                if (q >= b1) break;
                char *t0 = q;
                while (q < b1 && !is_blank(*q)) q++;
                char *t1 = q;
                char *bar = memchr(t0, '|', (size_t)(t1 - t0));
                int64_t lo, hi;
                if (!bar) {
                    if (!parse_i64(t0, t1, &lo)) {
                        seterr(err, errlen, "bad symbol", lineno);
                        ok = 0;
                        break;
                    }
                    hi = lo;
                } else {
                    char *bar2 = memchr(bar + 1, '|', (size_t)(t1 - bar - 1));
                    char *h1 = bar2 ? bar2 : t1; /* extra "|x" fields ignored */
                    if (!parse_i64(t0, bar, &lo) || !parse_i64(bar + 1, h1, &hi)) {
                        seterr(err, errlen, "bad symbol range", lineno);
                        ok = 0;
                        break;
                    }
                }
                if (src == dst && lo == 0 && hi == 255) {
                    /* persistent self-loop, excluded from the lists (:184-189) */
                    set_bit32(g->persis, g->S, src);
                    g->n_persis_lines++;
                    continue;
                }
                if (lo <= hi && (lo < 0 || hi > ORC_WB)) {
                    seterr(err, errlen, "symbol out of range", lineno);
                    ok = 0;
                    break;
                }
                for (int64_t sym = lo; sym <= hi; sym++) {
                    vec_push(&per[sym], src, dst); /* :214-221 */
                    g->n_raw_transitions++;
                }
            }
        } else if (contains(b0, b1, "accepting")) {
            /* accept line (:261-283) */
            uint32_t st;
            if (narrow != 0 || !parse_u32(a0, a1, &st)) {
                seterr(err, errlen, "bad accept state", lineno);
                ok = 0;
                break;
            }
            g->n_accept_lines++;
            set_bit32(g->accept, g->S, st);
            /* rules = blank-separated tokens after the first */
            char *q = b0;
            int tok = 0;
            int64_t first = nrules;
            while (q < b1) {
                while (q < b1 && is_blank(*q)) q++;
                if (q >= b1) break;
                char *t0 = q;
                while (q < b1 && !is_blank(*q)) q++;
                if (tok++ == 0) continue;
                int64_t r;
                if (!parse_i64(t0, q, &r)) {
                    seterr(err, errlen, "bad rule id", lineno);
                    ok = 0;
                    break;
                }
                if (nrules == caprules) {
                    caprules = caprules ? 2 * caprules : 64;
                    rules = (int32_t *)realloc(rules, (size_t)caprules * 4);
                }
                rules[nrules++] = (int32_t)r;
            }
            if (!ok) break;
            if (tok <= 1) continue; /* "No accept rules" (:271-274) */
            if (st < (uint32_t)g->S && !has_acc[st]) {
                has_acc[st] = 1;
                if (nacc == capacc) {
                    capacc = capacc ? 2 * capacc : 64;
                    acc = (acc_entry *)realloc(acc, (size_t)capacc * sizeof(acc_entry));
                }
                acc[nacc].state = st;
                acc[nacc].first_rule = (int32_t)first;
                acc[nacc].n_rules = (int32_t)(nrules - first);
                nacc++;
            }
        } else if (contains(b0, b1, "initial")) {
            uint32_t st;
            if (narrow != 0 || !parse_u32(a0, a1, &st)) {
                seterr(err, errlen, "bad initial state", lineno);
                ok = 0;
                break;
            }
            g->n_init_lines++;
            set_bit32(g->init, g->S, st);
        } else {
            g->n_bad_lines++; /* "Cannot parse" (:291-293) */
        }
    }
    free(buf);
    if (ok && !have_count) {
        seterr(err, errlen, "no state count line", lineno);
        ok = 0;
    }
    if (!ok) {
        for (int i = 0; i < ORC_SYMBOLS; i++) free(per[i].v);
        free(acc);
        free(rules);
        free(has_acc);
        orc_destroy(g);
        return NULL;
    }
    /* merge_transitions: per-symbol sort + unique (:369-411) */
    int64_t total = 0;
    for (int i = 0; i < ORC_SYMBOLS; i++) {
        if (per[i].n > 1) qsort(per[i].v, (size_t)per[i].n, sizeof(orc_tr), cmp_tr);
        int64_t w = 0;
        for (int64_t k = 0; k < per[i].n; k++)
            if (w == 0 || cmp_tr(&per[i].v[w - 1], &per[i].v[k]) != 0) per[i].v[w++] = per[i].v[k];
        per[i].n = w;
        total += w;
    }
    g->list = (orc_tr *)malloc((size_t)(total ? total : 1) * sizeof(orc_tr));
    int64_t idx = 0;
    for (int i = 0; i < ORC_SYMBOLS; i++) {
        g->off[i] = idx;
        if (per[i].n) memcpy(g->list + idx, per[i].v, (size_t)per[i].n * sizeof(orc_tr));
        idx += per[i].n;
        free(per[i].v);
    }
    g->off[ORC_SYMBOLS] = idx;
    /* accept map as CSR over states */
    g->acc_off = (int32_t *)calloc((size_t)g->S + 1, 4);
    for (int64_t k = 0; k < nacc; k++) g->acc_off[acc[k].state + 1] = acc[k].n_rules;
    for (int32_t s = 0; s < g->S; s++) g->acc_off[s + 1] += g->acc_off[s];
    g->acc_rules = (int32_t *)malloc((size_t)(g->acc_off[g->S] ? g->acc_off[g->S] : 1) * 4);
    for (int64_t k = 0; k < nacc; k++)
        memcpy(g->acc_rules + g->acc_off[acc[k].state], rules + acc[k].first_rule, (size_t)acc[k].n_rules * 4);
    g->n_accept_map = nacc;
    free(acc);
    free(rules);
    free(has_acc);
    return g;
}

/* kernel_common.cu:4-7 */
static inline int word_char(uint8_t c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_';
}

static inline int tst(const uint32_t *v, uint32_t i) { return (v[i >> 5] >> (i & 31)) & 1u; }
static inline void setb(uint32_t *v, uint32_t i) { v[i >> 5] |= 1u << (i & 31); }

static void wb_pass(const orc_nfa *g, uint32_t *cur) {
    /* in place, in (src,dst) order: iNFA_kernel.cu:145-164 with one thread */
    for (int64_t t = g->off[ORC_WB]; t < g->off[ORC_WB + 1]; t++)
        if (tst(cur, g->list[t].src)) setb(cur, g->list[t].dst);
}

/* iNFA step loop of iNFA_kernel.cu:41-172 at thread_count == 1.
 * cur, fut: caller scratch of W32 words each.  Final vector left in *out. */
static void run_steps(const orc_nfa *g, const uint8_t *x, int64_t L, uint32_t *a, uint32_t *b, uint32_t **out) {
    const int32_t W = g->W32;
    uint32_t *cur = a, *fut = b;
    memcpy(cur, g->init, (size_t)W * 4);
    const int has_wb = g->off[ORC_WB + 1] > g->off[ORC_WB];
    if (has_wb && L > 0 && word_char(x[0])) wb_pass(g, cur); /* head, :59-85 */
    for (int64_t byt = 0; byt < L; byt++) {
        for (int32_t i = 0; i < W; i++) fut[i] = cur[i] & g->persis[i]; /* :92-94 */
        const int c = x[byt];
        for (int64_t t = g->off[c]; t < g->off[c + 1]; t++) /* :102-121 */
            if (tst(cur, g->list[t].src)) setb(fut, g->list[t].dst);
        uint32_t *tmp = cur;
        cur = fut;
        fut = tmp; /* :124-130 */
        if (!has_wb) continue;
        if ((byt < L - 1 && word_char(x[byt]) == word_char(x[byt + 1])) ||
            (byt == L - 1 && !word_char(x[L - 1])))
            continue; /* :139-142 */
        wb_pass(g, cur);
    }
    *out = cur;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return x < y ? -1 : x > y;
}

/* accept extraction of iNFA_kernel.cu:290-307: returns the number of unique
 * rules; writes min(count, cap) of them (sorted) to out. */
static int32_t extract(const orc_nfa *g, const uint32_t *fin, int32_t *out, int32_t cap) {
    int32_t n = 0, ncap = 64;
    int32_t *tmp = (int32_t *)malloc((size_t)ncap * 4);
    for (int32_t w = 0; w < g->W32; w++) {
        uint32_t v = fin[w];
        while (v) {
            int b = __builtin_ctz(v);
            v &= v - 1;
            int32_t s = w * 32 + b;
            for (int32_t k = g->acc_off[s]; k < g->acc_off[s + 1]; k++) {
                if (n == ncap) {
                    ncap *= 2;
                    tmp = (int32_t *)realloc(tmp, (size_t)ncap * 4);
                }
                tmp[n++] = g->acc_rules[k];
            }
        }
    }
    if (n > 1) qsort(tmp, (size_t)n, 4, cmp_i32);
    int32_t u = 0;
    for (int32_t k = 0; k < n; k++)
        if (u == 0 || tmp[u - 1] != tmp[k]) tmp[u++] = tmp[k];
    for (int32_t k = 0; k < u && k < cap; k++) out[k] = tmp[k];
    free(tmp);
    return u;
}

/* ---------------- exported API (ctypes) ---------------- */

int32_t orc_state_count(const orc_nfa *g) { return g->S; }

/* info[0..9]: S, W32, total transitions (after unique), WB transitions,
 * persistent states (popcount), initial states, accept-map entries,
 * persistent lines, raw (pre-unique) transitions, unparsable lines */
void orc_info(const orc_nfa *g, int64_t *info) {
    int64_t np = 0, ni = 0;
    for (int32_t w = 0; w < g->W32; w++) {
        np += __builtin_popcount(g->persis[w]);
        ni += __builtin_popcount(g->init[w]);
    }
    info[0] = g->S;
    info[1] = g->W32;
    info[2] = g->off[ORC_SYMBOLS];
    info[3] = g->off[ORC_WB + 1] - g->off[ORC_WB];
    info[4] = np;
    info[5] = ni;
    info[6] = g->n_accept_map;
    info[7] = g->n_persis_lines;
    info[8] = g->n_raw_transitions;
    info[9] = g->n_bad_lines;
}

void orc_offsets(const orc_nfa *g, int64_t *out258) { memcpy(out258, g->off, sizeof(g->off)); }

void orc_transitions(const orc_nfa *g, uint32_t *pairs) {
    for (int64_t t = 0; t < g->off[ORC_SYMBOLS]; t++) {
        pairs[2 * t] = g->list[t].src;
        pairs[2 * t + 1] = g->list[t].dst;
    }
}

void orc_vectors(const orc_nfa *g, uint32_t *init, uint32_t *persis, uint32_t *accept) {
    memcpy(init, g->init, (size_t)g->W32 * 4);
    memcpy(persis, g->persis, (size_t)g->W32 * 4);
    memcpy(accept, g->accept, (size_t)g->W32 * 4);
}

/* accept map: out_off[S+1], out_rules[acc_off[S]] */
int32_t orc_accept_map(const orc_nfa *g, int32_t *out_off, int32_t *out_rules) {
    if (out_off) memcpy(out_off, g->acc_off, ((size_t)g->S + 1) * 4);
    if (out_rules) memcpy(out_rules, g->acc_rules, (size_t)g->acc_off[g->S] * 4);
    return g->acc_off[g->S];
}

/* final state vector (W32 words) after scanning x[0..L) */
void orc_final_states(const orc_nfa *g, const uint8_t *x, int64_t L, uint32_t *out) {
    uint32_t *a = (uint32_t *)malloc((size_t)g->W32 * 4), *b = (uint32_t *)malloc((size_t)g->W32 * 4), *fin;
    run_steps(g, x, L, a, b, &fin);
    memcpy(out, fin, (size_t)g->W32 * 4);
    free(a);
    free(b);
}

int32_t orc_scan(const orc_nfa *g, const uint8_t *x, int64_t L, int32_t *out, int32_t cap) {
    uint32_t *a = (uint32_t *)malloc((size_t)g->W32 * 4), *b = (uint32_t *)malloc((size_t)g->W32 * 4), *fin;
    run_steps(g, x, L, a, b, &fin);
    int32_t n = extract(g, fin, out, cap);
    free(a);
    free(b);
    return n;
}

typedef struct {
    const orc_nfa *g;
    const uint8_t *buf;
    const int64_t *offs;
    int32_t n, per_cap, tid, nthreads;
    int32_t *counts, *rules;
} job_t;

static void *batch_worker(void *arg) {
    job_t *j = (job_t *)arg;
    uint32_t *a = (uint32_t *)malloc((size_t)j->g->W32 * 4), *b = (uint32_t *)malloc((size_t)j->g->W32 * 4), *fin;
    for (int32_t i = j->tid; i < j->n; i += j->nthreads) {
        run_steps(j->g, j->buf + j->offs[i], j->offs[i + 1] - j->offs[i], a, b, &fin);
        j->counts[i] = extract(j->g, fin, j->rules + (int64_t)i * j->per_cap, j->per_cap);
    }
    free(a);
    free(b);
    return NULL;
}

/* Scan n strings buf[offs[i]..offs[i+1]); counts[i] = unique rule count
 * (may exceed per_cap: then only per_cap are written).  nthreads >= 1. */
void orc_batch_scan(const orc_nfa *g, const uint8_t *buf, const int64_t *offs, int32_t n, int32_t *counts,
                    int32_t *rules, int32_t per_cap, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    job_t jobs[64];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (job_t){g, buf, offs, n, per_cap, t, nthreads, counts, rules};
        if (nthreads == 1)
            batch_worker(&jobs[0]);
        else
            pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}
