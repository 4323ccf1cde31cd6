/*
 * oracle/ps5_oracle.c -- CPU oracle for the pS5 string sample sort.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py load this library: as the checker the
 * GPU output is compared against, or as the timed CPU baseline.  The product
 * library never links, loads or calls it.
 *
 * Contents (each function cites the reference file:line it restates; paths are
 * relative to the reference checkout):
 *   1. key extraction and LCP helpers
 *   2. brute-force stable sort with exact LCP (the mathematical definition)
 *   3. verifier: permutation, order, canonical tie order, exact LCP
 *   4. splitter tree builder and classifier
 *   5. pS5 (parallel super scalar string sample sort) with LCP output
 *
 * Output contract shared with the GPU path: `perm[i]` is the input index of the
 * i-th smallest string; runs of byte-identical strings are ordered by ascending
 * input index (the reference's own tie order depends on a stack-address RNG
 * seed, src/parallel/bingmann-parallel_sample_sort.hpp:1397, so it is
 * canonicalised); `lcp[i]` for i >= 1 is LCP(out[i-1], out[i]); `lcp[0]` is
 * never written (src/psstest.cpp:578-584).
 */
#define _GNU_SOURCE
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef const uint8_t* str_t;
typedef uint64_t skey_t;

/* ------------------------------------------------------------------------- */
/* 1. keys and LCP helpers                                                   */

/* Up to 8 characters at `depth`, big-endian, zero from the first NUL onward:
 * src/tools/stringset.hpp:201-231 (get_char_uint64_simple), 251-255. */
static inline skey_t get_key(str_t s, size_t depth)
{
    s += depth;
    skey_t v = 0;
    for (int i = 0; i < 8; ++i) {
        if (s[i] == 0) return v;
        v |= (skey_t)s[i] << (56 - 8 * i);
    }
    return v;
}

/* count_high_zero_bits / count_low_zero_bits with 0 -> 64:
 * src/tools/stringtools.hpp:302-343 */
static inline unsigned clz64z(skey_t x) { return x ? (unsigned)__builtin_clzll(x) : 64u; }
static inline unsigned ctz64z(skey_t x) { return x ? (unsigned)__builtin_ctzll(x) : 64u; }

/* lcpKeyType: src/parallel/bingmann-parallel_sample_sort.hpp:225-230 */
static inline unsigned lcp_key_type(skey_t a, skey_t b) { return clz64z(a ^ b) / 8; }
/* lcpKeyDepth: src/parallel/bingmann-parallel_sample_sort.hpp:232-237 */
static inline unsigned lcp_key_depth(skey_t a) { return 8 - ctz64z(a) / 8; }

/* calc_lcp: src/tools/stringtools.hpp:77-86 */
static inline size_t calc_lcp(str_t a, str_t b)
{
    size_t h = 0;
    while (a[h] != 0 && a[h] == b[h]) ++h;
    return h;
}

/* unsigned strcmp sign, as used by check_sorted_order (src/tools/checker.hpp:145) */
static inline int str_cmp(str_t a, str_t b)
{
    size_t h = calc_lcp(a, b);
    return (int)a[h] - (int)b[h];
}

/* exported wrappers for the tests (closed-form checks of the helpers) */
uint64_t ps5o_get_key(const uint8_t* s, uint64_t depth) { return get_key(s, depth); }
uint32_t ps5o_lcp_key_type(uint64_t a, uint64_t b) { return lcp_key_type(a, b); }
uint32_t ps5o_lcp_key_depth(uint64_t a) { return lcp_key_depth(a); }

/* ------------------------------------------------------------------------- */
/* 2. brute force: stable merge sort on input indices, exact adjacent LCP     */

typedef struct { const uint8_t* arena; const uint64_t* offs; } strtab;

static void msort_rec(const strtab* t, uint32_t* a, uint32_t* tmp, size_t n)
{
    if (n < 2) return;
    if (n <= 16) { /* stable insertion sort */
        for (size_t i = 1; i < n; ++i) {
            uint32_t x = a[i];
            size_t j = i;
            while (j > 0 && str_cmp(t->arena + t->offs[a[j - 1]], t->arena + t->offs[x]) > 0) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = x;
        }
        return;
    }
    size_t h = n / 2;
    msort_rec(t, a, tmp, h);
    msort_rec(t, a + h, tmp + h, n - h);
    size_t i = 0, j = h, k = 0;
    while (i < h && j < n) {
        /* take from the left on ties: stable */
        if (str_cmp(t->arena + t->offs[a[j]], t->arena + t->offs[a[i]]) < 0) tmp[k++] = a[j++];
        else tmp[k++] = a[i++];
    }
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, n * sizeof(uint32_t));
}

/* Stable sort of strings arena+offs[i]; perm/lcp as in the file header. */
void ps5o_brute_sort(const uint8_t* arena, const uint64_t* offs, uint64_t n,
                     uint32_t* perm, uint32_t* lcp)
{
    strtab t = { arena, offs };
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    uint32_t* tmp = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    msort_rec(&t, perm, tmp, n);
    free(tmp);
    if (lcp)
        for (uint64_t i = 1; i < n; ++i)
            lcp[i] = (uint32_t)calc_lcp(arena + offs[perm[i - 1]], arena + offs[perm[i]]);
}

/* ------------------------------------------------------------------------- */
/* 3. verifier                                                               */

/* Returns 0 when perm is a permutation of [0,n), the strings are in
 * non-decreasing unsigned order, byte-identical neighbours appear in ascending
 * input index, and (if lcp != NULL) lcp[i] == calc_lcp for all i >= 1.
 * Otherwise returns 1 + the failing position and writes a message.
 * Together these fix the output uniquely, so the check is exact at any size. */
uint64_t ps5o_verify(const uint8_t* arena, const uint64_t* offs, uint64_t n,
                     const uint32_t* perm, const uint32_t* lcp,
                     char* msg, uint64_t msglen)
{
    if (msg && msglen) msg[0] = 0;
    if (n == 0) return 0;
    uint8_t* seen = (uint8_t*)calloc(n, 1);
    for (uint64_t i = 0; i < n; ++i) {
        if (perm[i] >= n || seen[perm[i]]) {
            if (msg) snprintf(msg, msglen, "not a permutation at %llu (value %u)",
                              (unsigned long long)i, perm[i]);
            free(seen);
            return i + 1;
        }
        seen[perm[i]] = 1;
    }
    free(seen);
    uint64_t bad = UINT64_MAX;
    int kind = 0;
    #pragma omp parallel for schedule(static) reduction(min : bad)
    for (uint64_t i = 1; i < n; ++i) {
        str_t a = arena + offs[perm[i - 1]], b = arena + offs[perm[i]];
        size_t h = calc_lcp(a, b);
        int c = (int)a[h] - (int)b[h];
        if (c > 0 || (c == 0 && perm[i - 1] > perm[i]) || (lcp && lcp[i] != h))
            if (i < bad) bad = i;
    }
    if (bad == UINT64_MAX) return 0;
    {
        str_t a = arena + offs[perm[bad - 1]], b = arena + offs[perm[bad]];
        size_t h = calc_lcp(a, b);
        int c = (int)a[h] - (int)b[h];
        kind = c > 0 ? 1 : (c == 0 && perm[bad - 1] > perm[bad]) ? 2 : 3;
        if (msg)
            snprintf(msg, msglen, "%s at position %llu (lcp got %u want %zu)",
                     kind == 1 ? "order violated" : kind == 2 ? "tie order violated" : "lcp mismatch",
                     (unsigned long long)bad, lcp ? lcp[bad] : 0u, h);
    }
    return bad + 1;
}

/* Distinguishing characters of a verified output, as verify_lcp_cache
 * checks them (stringptr.hpp:1015-1021): cache[i] = get_char(s_i, lcp[i])
 * for i >= 1 -- the character of s_i right after the prefix it shares with
 * s_{i-1} (NUL for equal strings).  Returns 1 + the first i whose cache byte
 * differs from the given one when `check` is non-zero, else fills cache. */
uint64_t ps5o_dist_chars(const uint8_t* arena, const uint64_t* offs, uint64_t n,
                         const uint32_t* perm, const uint32_t* lcp, uint8_t* cache, int check)
{
    for (uint64_t i = 1; i < n; ++i) {
        const uint8_t c = arena[offs[perm[i]] + lcp[i]];
        if (check) {
            if (cache[i] != c) return i + 1;
        } else {
            cache[i] = c;
        }
    }
    return 0;
}

static int cmp_u32(const void* x, const void* y)
{
    uint32_t a = *(const uint32_t*)x, b = *(const uint32_t*)y;
    return a < b ? -1 : a > b;
}

/* Put every run of byte-identical strings in ascending input index. */
void ps5o_canonicalize(const uint8_t* arena, const uint64_t* offs, uint64_t n, uint32_t* perm)
{
    uint64_t i = 0;
    while (i < n) {
        uint64_t j = i + 1;
        while (j < n && str_cmp(arena + offs[perm[i]], arena + offs[perm[j]]) == 0) ++j;
        if (j - i > 1) {
            if (j - i > 64) {
                qsort(perm + i, j - i, sizeof(uint32_t), cmp_u32);
            } else {
                for (uint64_t a = i + 1; a < j; ++a) {
                    uint32_t x = perm[a];
                    uint64_t b = a;
                    while (b > i && perm[b - 1] > x) { perm[b] = perm[b - 1]; --b; }
                    perm[b] = x;
                }
            }
        }
        i = j;
    }
}

/* distinguishing prefix and LCP sum: src/tools/checker.hpp:155-180 */
uint64_t ps5o_dprefix(const uint8_t* arena, const uint64_t* offs, uint64_t n,
                      const uint32_t* perm, uint64_t* lcpsum)
{
    uint64_t D = 0, pdepth = 0, ls = 0;
    for (uint64_t i = 1; i < n; ++i) {
        uint64_t depth = calc_lcp(arena + offs[perm[i - 1]], arena + offs[perm[i]]);
        ls += depth;
        depth++;
        if (pdepth < depth) D += depth - pdepth;
        D += depth;
        pdepth = depth;
    }
    if (lcpsum) *lcpsum = ls;
    return D;
}

/* ------------------------------------------------------------------------- */
/* 4. splitter tree: src/sequential/bingmann-sample_sort_tree_builder.hpp     */

enum { TREEBITS = 10, NSPL = (1 << TREEBITS) - 1, BKTNUM = 2 * NSPL + 1,
       OVERSAMPLE = 2, SAMPLESIZE = OVERSAMPLE * NSPL };

typedef struct {
    skey_t splitter[NSPL];          /* in-order */
    skey_t tree[NSPL + 1];          /* level order, root at 1 */
    uint8_t splitter_lcp[NSPL + 1];
} classifier;

typedef struct { skey_t* spl; skey_t* tree; uint8_t* lcp_iter; } tb_state;

/* TreeBuilderPreAndLevelOrder::recurse, tree_builder.hpp:118-183 */
static skey_t tb_recurse(tb_state* st, const skey_t* lo, const skey_t* hi,
                        unsigned treeidx, const skey_t* rec_prevkey)
{
    const skey_t* mid = lo + (hi - lo) / 2;
    skey_t mykey = st->tree[treeidx] = *mid;
    const skey_t* midlo = mid;
    while (lo < midlo && *(midlo - 1) == mykey) midlo--;
    const skey_t* midhi = mid;
    while (midhi + 1 < hi && *midhi == mykey) midhi++;
    if (2 * treeidx < NSPL) {
        skey_t prevkey = tb_recurse(st, lo, midlo, 2 * treeidx + 0, rec_prevkey);
        skey_t x = prevkey ^ mykey;
        *st->spl++ = mykey;
        *st->lcp_iter++ = (uint8_t)((clz64z(x) / 8) | ((mykey & 0xFF) ? 0 : 0x80));
        return tb_recurse(st, midhi, hi, 2 * treeidx + 1, &mykey);
    }
    skey_t x = *rec_prevkey ^ mykey;
    *st->spl++ = mykey;
    *st->lcp_iter++ = (uint8_t)((clz64z(x) / 8) | ((mykey & 0xFF) ? 0 : 0x80));
    return mykey;
}

/* TreeBuilderPreAndLevelOrder ctor, tree_builder.hpp:88-106 */
static void classifier_build(classifier* c, const skey_t* samples, size_t samplesize)
{
    tb_state st = { c->splitter, c->tree, c->splitter_lcp };
    skey_t sentinel = 0;
    tb_recurse(&st, samples, samples + samplesize, 1, &sentinel);
    c->splitter_lcp[0] &= 0x80;
    c->splitter_lcp[NSPL] = 0;
}

/* ClassifyTreeSimple::find_bkt, src/sequential/bingmann-sample_sortBTC.hpp:45-72 */
static inline unsigned find_bkt(const classifier* c, skey_t key)
{
    unsigned i = 1;
    while (i <= NSPL) i = 2 * i + (key <= c->tree[i] ? 0 : 1);
    i -= NSPL + 1;
    unsigned b = i * 2;
    if (i < NSPL && c->splitter[i] == key) b += 1;
    return b;
}

/* ClassifyTreeUnrollInterleave::classify with Rollout=4,
 * src/sequential/bingmann-sample_sortBTC.hpp:297-399 */
static void classify(const classifier* c, str_t* begin, str_t* end, uint16_t* bktout, size_t depth)
{
    while (begin != end) {
        if (begin + 4 < end) {
            skey_t k[4];
            unsigned i[4] = { 1, 1, 1, 1 };
            for (int u = 0; u < 4; ++u) k[u] = get_key(begin[u], depth);
            for (int l = 0; l < TREEBITS; ++l)
                for (int u = 0; u < 4; ++u) i[u] = 2 * i[u] + (k[u] <= c->tree[i[u]] ? 0 : 1);
            for (int u = 0; u < 4; ++u) {
                unsigned j = i[u] - (NSPL + 1);
                unsigned b = 2 * j;
                if (j < NSPL && c->splitter[j] == k[u]) b += 1;
                bktout[u] = (uint16_t)b;
            }
            begin += 4;
            bktout += 4;
        } else {
            *bktout++ = (uint16_t)find_bkt(c, get_key(*begin++, depth));
        }
    }
}

/* exported for tests: build a classifier from sorted samples and classify keys */
void ps5o_classify_keys(const uint64_t* sorted_samples, const uint64_t* keys, uint64_t n,
                        uint16_t* bkt_out, uint64_t* splitters_out, uint8_t* splitter_lcp_out)
{
    classifier* c = (classifier*)malloc(sizeof(classifier));
    classifier_build(c, sorted_samples, SAMPLESIZE);
    for (uint64_t i = 0; i < n; ++i) bkt_out[i] = (uint16_t)find_bkt(c, keys[i]);
    if (splitters_out) memcpy(splitters_out, c->splitter, sizeof(c->splitter));
    if (splitter_lcp_out) memcpy(splitter_lcp_out, c->splitter_lcp, sizeof(c->splitter_lcp));
    free(c);
}

/* ------------------------------------------------------------------------- */
/* 5. pS5                                                                    */

/* LCGRandom: src/tools/lcgrandom.hpp:27-40.  The reference seeds it with a
 * stack address (bingmann-parallel_sample_sort.hpp:1397); the oracle seeds it
 * from the segment so runs are reproducible.  Only splitter choice depends on
 * it, never the output. */
typedef struct { uint64_t xn; } lcg;
static inline uint64_t lcg_next(lcg* r)
{
    r->xn = 0x27BB2EE687B0B0FDULL * r->xn + 0xB504F32DULL;
    return r->xn;
}

/* StringShadowPtr / StringShadowLcpPtr, src/tools/stringptr.hpp:350-486,630-741:
 * `act` holds the strings now, `sha` is the other array, `flipped` says
 * whether act is the shadow; the LCP array is indexed by position. */
typedef struct { str_t* act; str_t* sha; uint64_t* lcp; size_t n; int flipped; } sptr;

static inline sptr sp_sub(sptr s, size_t off, size_t n)
{ sptr r = { s.act + off, s.sha + off, s.lcp ? s.lcp + off : NULL, n, s.flipped }; return r; }
static inline sptr sp_flip(sptr s, size_t off, size_t n)
{ sptr r = { s.sha + off, s.act + off, s.lcp ? s.lcp + off : NULL, n, !s.flipped }; return r; }
static inline sptr sp_original(sptr s)
{ if (!s.flipped) return s; sptr r = { s.sha, s.act, s.lcp, s.n, 0 }; return r; }
static inline sptr sp_copy_back(sptr s)
{
    if (!s.flipped) return s;
    memcpy(s.sha, s.act, s.n * sizeof(str_t));
    sptr r = { s.sha, s.act, s.lcp, s.n, 0 };
    return r;
}
static inline void sp_set_lcp(sptr s, size_t i, uint64_t v) { if (s.lcp) s.lcp[i] = v; }
static inline void sp_fill_lcp(sptr s, uint64_t v)
{ if (s.lcp) for (size_t i = 1; i < s.n; ++i) s.lcp[i] = v; }

/* Context, bingmann-parallel_sample_sort.hpp:120-184 */
enum { MAXPROCS = 2 * 64 + 1 };
static const size_t g_smallsort_threshold = 1024 * 1024;   /* .hpp:103 */
static const size_t g_inssort_threshold = 32;              /* .hpp:104 */
/* the job queue's work sharing (.hpp:678-755,1180-1236) becomes OpenMP tasks
 * for sub-problems at least this large */
static const size_t g_task_threshold = 16384;

typedef struct { size_t totalsize, threadnum; int calc_lcp; uint64_t seed; } ctx_t;

static inline size_t seq_threshold(const ctx_t* c)
{
    size_t t = c->totalsize / c->threadnum;
    return t > g_smallsort_threshold ? t : g_smallsort_threshold;
}

static int cmp_key(const void* a, const void* b)
{
    skey_t x = *(const skey_t*)a, y = *(const skey_t*)b;
    return x < y ? -1 : x > y;
}

static void draw_samples(const ctx_t* c, sptr sp, size_t depth, skey_t* samples)
{
    lcg rng = { c->seed ^ (uint64_t)(uintptr_t)sp.act ^ ((uint64_t)depth << 48) ^ sp.n };
    for (size_t i = 0; i < SAMPLESIZE; ++i)
        samples[i] = get_key(sp.act[lcg_next(&rng) % sp.n], depth);
    qsort(samples, SAMPLESIZE, sizeof(skey_t), cmp_key);
}

/* sample_sort_lcp, .hpp:311-396 (the goto structure restated as one loop) */
static void sample_sort_lcp(const classifier* cl, sptr sp, size_t depth, const size_t* bkt)
{
    if (!sp.lcp) return;
    size_t b = 0;
    skey_t prevkey = 0;
    int have_prev = 0;
    for (; b < BKTNUM; ++b) {
        if (bkt[b] == bkt[b + 1]) continue;
        skey_t thiskey = (b & 1) ? cl->splitter[b / 2] : get_key(sp.act[bkt[b]], depth);
        if (have_prev)
            sp_set_lcp(sp, bkt[b], depth + lcp_key_type(prevkey, thiskey));
        prevkey = (b & 1) ? thiskey : get_key(sp.act[bkt[b + 1] - 1], depth);
        have_prev = 1;
    }
}

static void enqueue(const ctx_t* c, sptr sp, size_t depth);

/* lcp_insertion_sort, src/sequential/bingmann-lcp_inssort.hpp:36-187 (the
 * separate bounds-checked last iteration folded into the loop) */
static void lcp_insertion_sort(str_t* str, uint64_t* lcp, size_t n, size_t depth)
{
    if (n <= 1) return;
    for (size_t j = 0; j < n; ++j) {
        str_t new_str = str[j];
        size_t new_lcp = depth;
        size_t i = j;
        while (i > 0) {
            size_t prev_lcp = new_lcp;
            str_t cur_str = str[i - 1];
            size_t cur_lcp = lcp[i];
            if (cur_lcp < new_lcp) {
                break;
            } else if (cur_lcp == new_lcp) {
                while (new_str[new_lcp] != 0 && new_str[new_lcp] == cur_str[new_lcp]) ++new_lcp;
                if (!(new_str[new_lcp] < cur_str[new_lcp])) {
                    lcp[i] = new_lcp;
                    new_lcp = prev_lcp;
                    break;
                }
            }
            str[i] = cur_str;
            if (i + 1 < n) lcp[i + 1] = cur_lcp;
            --i;
        }
        str[i] = new_str;
        if (i + 1 < n) lcp[i + 1] = new_lcp;
    }
}

/* inssort_generic, src/sequential/inssort.hpp:61-97 */
static void inssort_generic(str_t* s, size_t n, size_t depth)
{
    for (size_t i = 1; i < n; ++i) {
        str_t tmp = s[i];
        size_t j = i;
        while (j > 0) {
            str_t a = s[j - 1] + depth, t = tmp + depth;
            while (*a != 0 && *a == *t) ++a, ++t;
            if (*a <= *t) break;
            s[j] = s[j - 1];
            --j;
        }
        s[j] = tmp;
    }
}

/* insertion_sort type switch, .hpp:249-306 */
static void insertion_sort(sptr sp, size_t depth)
{
    if (sp.lcp) lcp_insertion_sort(sp.act, sp.lcp, sp.n, depth);
    else inssort_generic(sp.act, sp.n, depth);
}

/* insertion_sort_cache_block + insertion_sort_cache, .hpp:787-861 */
static void insertion_sort_cache(sptr sp0, skey_t* cache, size_t depth, int cache_dirty)
{
    sptr sp = sp_copy_back(sp0);
    if (sp.n <= 1) return;
    if (cache_dirty) { insertion_sort(sp, depth); return; }
    str_t* s = sp.act;
    size_t n = sp.n;
    for (size_t pi = 1; pi < n; ++pi) {
        str_t tmps = s[pi];
        skey_t tmpc = cache[pi];
        size_t pj = pi;
        for (; pj > 0; --pj) {
            if (cache[pj - 1] <= tmpc) break;
            s[pj] = s[pj - 1];
            cache[pj] = cache[pj - 1];
        }
        s[pj] = tmps;
        cache[pj] = tmpc;
    }
    size_t start = 0, bktsize = 1;
    for (size_t i = 0; i + 1 <= n; ++i) {
        if (i + 1 < n && cache[i] == cache[i + 1]) { ++bktsize; continue; }
        if (start != 0)
            sp_set_lcp(sp, start, depth + lcp_key_type(cache[start - 1], cache[start]));
        if (bktsize > 1) {
            if (cache[start] & 0xFF) insertion_sort(sp_sub(sp, start, bktsize), depth + 8);
            else sp_fill_lcp(sp_sub(sp, start, bktsize), depth + lcp_key_depth(cache[start]));
        }
        bktsize = 1;
        start = i + 1;
    }
}

static int cmp3(skey_t a, skey_t b) { return a > b ? 1 : a < b ? -1 : 0; }

/* med3, .hpp:769-784 */
static size_t med3(const skey_t* A, size_t i, size_t j, size_t k)
{
    if (A[i] == A[j]) return i;
    if (A[k] == A[i] || A[k] == A[j]) return k;
    if (A[i] < A[j]) {
        if (A[j] < A[k]) return j;
        if (A[i] < A[k]) return k;
        return i;
    }
    if (A[j] > A[k]) return j;
    if (A[i] < A[k]) return i;
    return k;
}

#define SWAPS(a, b) do { str_t _t = (a); (a) = (b); (b) = _t; } while (0)
#define SWAPK(a, b) do { skey_t _t = (a); (a) = (b); (b) = _t; } while (0)

/* MKQSStep (PS5_CALC_LCP_MKQS == 1), .hpp:863-1051, with the explicit stack
 * of sort_mkqs_cache (.hpp:1056-1178) as recursion */
static void mkqs_step(const ctx_t* c, sptr sp, skey_t* cache, size_t depth, int cache_dirty)
{
    size_t n = sp.n;
    str_t* s = sp.act;
    if (cache_dirty)
        for (size_t i = 0; i < n; ++i) cache[i] = get_key(s[i], depth);
    size_t p = med3(cache,
                    med3(cache, 0, n / 8, n / 4),
                    med3(cache, n / 2 - n / 8, n / 2, n / 2 + n / 8),
                    med3(cache, n - 1 - n / 4, n - 1 - n / 8, n - 3));
    SWAPS(s[0], s[p]);
    SWAPK(cache[0], cache[p]);
    skey_t pivot = cache[0];
    skey_t max_lt = 0, min_gt = UINT64_MAX;
    size_t leq = 1, llt = 1, rgt = n - 1, req = n - 1;
    for (;;) {
        while (llt <= rgt) {
            int r = cmp3(cache[llt], pivot);
            if (r > 0) { if (cache[llt] < min_gt) min_gt = cache[llt]; break; }
            else if (r == 0) { SWAPS(s[leq], s[llt]); SWAPK(cache[leq], cache[llt]); leq++; }
            else { if (cache[llt] > max_lt) max_lt = cache[llt]; }
            ++llt;
        }
        while (llt <= rgt) {
            int r = cmp3(cache[rgt], pivot);
            if (r < 0) { if (cache[rgt] > max_lt) max_lt = cache[rgt]; break; }
            else if (r == 0) { SWAPS(s[req], s[rgt]); SWAPK(cache[req], cache[rgt]); req--; }
            else { if (cache[rgt] < min_gt) min_gt = cache[rgt]; }
            --rgt;
        }
        if (llt > rgt) break;
        SWAPS(s[llt], s[rgt]);
        SWAPK(cache[llt], cache[rgt]);
        ++llt;
        --rgt;
    }
    size_t num_leq = leq, num_req = n - 1 - req;
    size_t num_eq = num_leq + num_req;
    size_t num_lt = llt - leq;
    size_t num_gt = req - rgt;
    size_t size1 = num_leq < num_lt ? num_leq : num_lt;
    for (size_t i = 0; i < size1; ++i) { SWAPS(s[i], s[llt - size1 + i]); SWAPK(cache[i], cache[llt - size1 + i]); }
    size_t size2 = num_req < num_gt ? num_req : num_gt;
    for (size_t i = 0; i < size2; ++i) { SWAPS(s[llt + i], s[n - size2 + i]); SWAPK(cache[llt + i], cache[n - size2 + i]); }
    int eq_recurse = (pivot & 0xFF) != 0;
    unsigned lcp_lt = 0, lcp_gt = 0, lcp_eq = lcp_key_depth(pivot);
    if (num_lt > 0) lcp_lt = lcp_key_type(max_lt, pivot);
    if (num_gt > 0) lcp_gt = lcp_key_type(pivot, min_gt);

    /* lt part */
    if (num_lt > 0) {
        sptr sub = sp_sub(sp, 0, num_lt);
        if (num_lt < g_inssort_threshold) insertion_sort_cache(sub, cache, depth, 0);
        else {
            #pragma omp task if (num_lt >= g_task_threshold)
            mkqs_step(c, sub, cache, depth, 0);
        }
    }
    /* eq part */
    {
        sptr sub = sp_sub(sp, num_lt, num_eq);
        if (!eq_recurse) {
            sptr spb = sp_copy_back(sub);
            sp_fill_lcp(spb, depth + lcp_eq);
        } else if (num_eq < g_inssort_threshold) {
            insertion_sort_cache(sub, cache + num_lt, depth + 8, 1);
        } else {
            #pragma omp task if (num_eq >= g_task_threshold)
            mkqs_step(c, sub, cache + num_lt, depth + 8, 1);
        }
    }
    /* gt part */
    if (num_gt > 0) {
        sptr sub = sp_sub(sp, num_lt + num_eq, num_gt);
        if (num_gt < g_inssort_threshold) insertion_sort_cache(sub, cache + num_lt + num_eq, depth, 0);
        else {
            #pragma omp task if (num_gt >= g_task_threshold)
            mkqs_step(c, sub, cache + num_lt + num_eq, depth, 0);
        }
    }
    #pragma omp taskwait
    /* calculate_lcp, .hpp:1012-1026 */
    sptr o = sp_original(sp);
    if (num_lt > 0) sp_set_lcp(o, num_lt, depth + lcp_lt);
    if (num_gt > 0) sp_set_lcp(o, num_lt + num_eq, depth + lcp_gt);
}

/* sort_mkqs_cache entry, .hpp:1056-1090 */
static void sort_mkqs_cache(const ctx_t* c, sptr sp, size_t depth)
{
    if (sp.n < g_inssort_threshold) {
        insertion_sort(sp_copy_back(sp), depth);
        return;
    }
    skey_t* cache = (skey_t*)malloc(sp.n * sizeof(skey_t));
    mkqs_step(c, sp, cache, depth, 1);
    free(cache);
}

/* SeqSampleSortStep + sort_sample_sort, .hpp:431-676 */
static void seq_sample_sort(const ctx_t* c, sptr sp, size_t depth)
{
    size_t n = sp.n;
    classifier* cl = (classifier*)malloc(sizeof(classifier));
    skey_t samples[SAMPLESIZE];
    draw_samples(c, sp, depth, samples);
    classifier_build(cl, samples, SAMPLESIZE);
    uint16_t* bktcache = (uint16_t*)malloc(n * sizeof(uint16_t));
    classify(cl, sp.act, sp.act + n, bktcache, depth);
    size_t* bkt = (size_t*)calloc(BKTNUM + 1, sizeof(size_t));
    for (size_t i = 0; i < n; ++i) ++bkt[bktcache[i]];
    for (unsigned i = 1; i < BKTNUM; ++i) bkt[i] += bkt[i - 1];
    bkt[BKTNUM] = n;
    for (size_t i = 0; i < n; ++i) sp.sha[--bkt[bktcache[i]]] = sp.act[i];
    free(bktcache);
    for (size_t i = 0; i < BKTNUM; ++i) {
        size_t bktsize = bkt[i + 1] - bkt[i];
        if (bktsize == 0) continue;
        sptr sub = sp_flip(sp, bkt[i], bktsize);
        if (i % 2 == 0) {
            size_t d = depth + (cl->splitter_lcp[i / 2] & 0x7F);
            #pragma omp task if (bktsize >= g_task_threshold)
            {
                if (bktsize < g_smallsort_threshold) sort_mkqs_cache(c, sub, d);
                else seq_sample_sort(c, sub, d);
            }
        } else if (cl->splitter_lcp[i / 2] & 0x80) {
            sptr spb = sp_copy_back(sub);
            sp_fill_lcp(spb, depth + lcp_key_depth(cl->splitter[i / 2]));
        } else {
            #pragma omp task if (bktsize >= g_task_threshold)
            {
                if (bktsize < g_smallsort_threshold) sort_mkqs_cache(c, sub, depth + 8);
                else seq_sample_sort(c, sub, depth + 8);
            }
        }
    }
    #pragma omp taskwait
    sample_sort_lcp(cl, sp_original(sp), depth, bkt);
    free(bkt);
    free(cl);
}

/* SmallsortJob::run, .hpp:528-563 */
static void smallsort(const ctx_t* c, sptr sp, size_t depth)
{
    if (sp.n >= g_smallsort_threshold) seq_sample_sort(c, sp, depth);
    else sort_mkqs_cache(c, sp, depth);
}

/* SampleSortStep: sample (.hpp:1384-1410), count (1414-1435), prefix sum
 * (1437-1460), distribute (1464-1490), distribute_finished (1492-1584),
 * substep_all_done -> sample_sort_lcp (1588-1600) */
static void sample_sort_step(const ctx_t* c, sptr sp, size_t depth)
{
    size_t n = sp.n;
    size_t parts = n / seq_threshold(c) * 2;
    if (parts == 0) parts = 1;
    if (parts > MAXPROCS) parts = MAXPROCS;
    size_t psize = (n + parts - 1) / parts;

    classifier* cl = (classifier*)malloc(sizeof(classifier));
    skey_t samples[SAMPLESIZE];
    draw_samples(c, sp, depth, samples);
    classifier_build(cl, samples, SAMPLESIZE);

    size_t** bkt = (size_t**)calloc(parts, sizeof(size_t*));
    uint16_t** bktcache = (uint16_t**)calloc(parts, sizeof(uint16_t*));
    for (size_t p = 0; p < parts; ++p) {
        #pragma omp task firstprivate(p)
        {
            size_t b = p * psize, e = (p + 1) * psize < n ? (p + 1) * psize : n;
            if (e < b) e = b;
            bktcache[p] = (uint16_t*)malloc((e - b + 1) * sizeof(uint16_t));
            classify(cl, sp.act + b, sp.act + e, bktcache[p], depth);
            bkt[p] = (size_t*)calloc(BKTNUM + 1, sizeof(size_t));
            for (size_t i = 0; i < e - b; ++i) ++bkt[p][bktcache[p][i]];
        }
    }
    #pragma omp taskwait
    size_t sum = 0;
    for (unsigned i = 0; i < BKTNUM; ++i)
        for (size_t p = 0; p < parts; ++p) bkt[p][i] = (sum += bkt[p][i]);
    for (size_t p = 0; p < parts; ++p) {
        #pragma omp task firstprivate(p)
        {
            size_t b = p * psize, e = (p + 1) * psize < n ? (p + 1) * psize : n;
            if (e < b) e = b;
            size_t* mybkt = bkt[p];
            const uint16_t* bc = bktcache[p];
            for (size_t i = b; i < e; ++i) sp.sha[--mybkt[*bc++]] = sp.act[i];
        }
    }
    #pragma omp taskwait
    size_t* bk = bkt[0];
    bk[BKTNUM] = n;
    for (size_t p = 0; p < parts; ++p) { free(bktcache[p]); if (p) free(bkt[p]); }
    free(bktcache);
    free(bkt);

    for (size_t i = 0; i < BKTNUM; ++i) {
        size_t bktsize = bk[i + 1] - bk[i];
        if (bktsize == 0) continue;
        sptr sub = sp_flip(sp, bk[i], bktsize);
        if (bktsize == 1) { sp_copy_back(sub); continue; }
        if (i % 2 == 0) {
            size_t d = depth + (cl->splitter_lcp[i / 2] & 0x7F);
            #pragma omp task
            enqueue(c, sub, d);
        } else if (cl->splitter_lcp[i / 2] & 0x80) {
            sptr spb = sp_copy_back(sub);
            sp_fill_lcp(spb, depth + lcp_key_depth(cl->splitter[i / 2]));
        } else {
            #pragma omp task
            enqueue(c, sub, depth + 8);
        }
    }
    #pragma omp taskwait
    sample_sort_lcp(cl, sp_original(sp), depth, bk);
    free(bk);
    free(cl);
}

/* Enqueue, .hpp:1613-1634 */
static void enqueue(const ctx_t* c, sptr sp, size_t depth)
{
    if (sp.n > seq_threshold(c)) sample_sort_step(c, sp, depth);
    else smallsort(c, sp, depth);
}

/* parallel_sample_sort_base / parallel_sample_sort_lcp_base, .hpp:1640-1748:
 * the reference's own string-pointer interface.  `lcp` may be NULL (plain
 * parallel_sample_sortBTCUI); lcp[0] is left untouched. */
void ps5o_ps5_sort_ptrs(const uint8_t** strings, uint64_t n, uint64_t* lcp,
                        int nthreads, uint64_t seed)
{
    if (n <= 1) return;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    str_t* shadow = (str_t*)malloc(n * sizeof(str_t));
    ctx_t c = { n, (size_t)nthreads, lcp != NULL, seed };
    sptr sp = { (str_t*)strings, shadow, lcp, n, 0 };
    #pragma omp parallel num_threads(nthreads)
    #pragma omp single
    enqueue(&c, sp, 0);
    free(shadow);
}

static int cmp_u64pair(const void* a, const void* b)
{
    uint64_t x = ((const uint64_t*)a)[0], y = ((const uint64_t*)b)[0];
    return x < y ? -1 : x > y;
}

/* Input index of every sorted string pointer s[i] (offsets are distinct),
 * then runs of identical strings in ascending input index. */
static void ptrs_to_perm(const uint8_t* arena, const uint64_t* offs, uint64_t n, str_t* s, uint32_t* perm)
{
    uint64_t* tab = (uint64_t*)malloc(2 * n * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) { tab[2 * i] = offs[i]; tab[2 * i + 1] = i; }
    qsort(tab, n, 2 * sizeof(uint64_t), cmp_u64pair);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t o = (uint64_t)(s[i] - arena);
        uint64_t lo = 0, hi = n;
        while (lo + 1 < hi) { uint64_t m = (lo + hi) / 2; if (tab[2 * m] <= o) lo = m; else hi = m; }
        perm[i] = (uint32_t)tab[2 * lo + 1];
    }
    free(tab);
    ps5o_canonicalize(arena, offs, n, perm);
}

/* pS5 over arena+offs, returning the canonical permutation and, if lcp is
 * not NULL, the LCP array.  With lcp == NULL the restatement runs the plain
 * contestant's branches (parallel_sample_sortBTCUI: no LCP array, generic
 * insertion sort, .hpp:1685-1700, inssort.hpp:61-97). */
void ps5o_ps5_sort(const uint8_t* arena, const uint64_t* offs, uint64_t n,
                   uint32_t* perm, uint32_t* lcp, int nthreads, uint64_t seed)
{
    if (n == 0) return;
    str_t* s = (str_t*)malloc(n * sizeof(str_t));
    for (uint64_t i = 0; i < n; ++i) s[i] = arena + offs[i];
    uint64_t* l = lcp ? (uint64_t*)malloc(n * sizeof(uint64_t)) : NULL;
    ps5o_ps5_sort_ptrs(s, n, l, nthreads, seed);
    ptrs_to_perm(arena, offs, n, s, perm);
    if (lcp) for (uint64_t i = 1; i < n; ++i) lcp[i] = (uint32_t)l[i];
    free(l);
    free(s);
}

/* Timed CPU baseline: only the sort call is timed (src/psstest.cpp:593-603),
 * with_lcp selecting parallel_sample_sort_lcp_base over the plain contestant.
 * Afterwards, outside the timed region, the canonical permutation (and the
 * LCP array when with_lcp) go to perm/lcp if they are not NULL, so that the
 * timed sort's own output can be verified. */
double ps5o_time_ps5(const uint8_t* arena, const uint64_t* offs, uint64_t n,
                     int nthreads, int with_lcp, uint32_t* perm, uint32_t* lcp)
{
    str_t* s = (str_t*)malloc((n ? n : 1) * sizeof(str_t));
    for (uint64_t i = 0; i < n; ++i) s[i] = arena + offs[i];
    uint64_t* l = with_lcp ? (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t)) : NULL;
    double t0 = omp_get_wtime();
    ps5o_ps5_sort_ptrs(s, n, l, nthreads, 0x5EED);
    double t1 = omp_get_wtime();
    if (perm && n) ptrs_to_perm(arena, offs, n, s, perm);
    if (lcp && l) for (uint64_t i = 1; i < n; ++i) lcp[i] = (uint32_t)l[i];
    free(l);
    free(s);
    return t1 - t0;
}

int ps5o_max_threads(void) { return omp_get_max_threads(); }
