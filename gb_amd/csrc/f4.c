/* f4.c -- the F4 host around libgbgpu's linear algebra (C11).
 *
 * The loop of gb's f4_julia (f4.c:190-372): select the pairs of minimal
 * degree, build the matrix by symbolic preprocessing, reduce it, turn the
 * new rows into basis elements, update the pairs (Gebauer-Moeller); at the
 * end drop redundant elements and, if asked, interreduce.  Own design
 * throughout; where the matrix has to look like gb's (row format, known
 * pivot = first row per lead, columns = pivot columns first), it does.
 *
 *  * Monomials live in one table: exponent vectors (16 bits per variable,
 *    at most 32767 each, the range of gb's int16 exp_t, data.h:63;
 *    a product leaving that range ends the run with GBGPU_ERANGE),
 *    degree, a divisibility mask (bit i set when x_i occurs) and an
 *    additive hash h(e) = sum_i r_i e_i, so h(a b) = h(a) + h(b) and a
 *    product is looked up without hashing its exponents again.
 *  * A basis element keeps its terms as monomial ids in decreasing order
 *    and its coefficients (leading one 1).  A matrix row is (element,
 *    multiplier): its coefficient array is the element's, shared by all
 *    its multiples -- as in gb (hash.c:903-926).
 *  * Pairs: the Gebauer-Moeller installation of new elements (Becker and
 *    Weispfenning, Groebner Bases, 1993, algorithm UPDATE).
 *  * Symbolic preprocessing: rows of the selected pairs (both generators
 *    of every pair, one row per distinct (element, multiplier)), then a
 *    reducer for every column monomial divisible by a basis lead term --
 *    the first element in basis order whose lead divides it, the rule of
 *    gb's find_multiplied_reducer (symbol.c:151-200).
 *  * Columns: pivot columns (leads of rows) first, then the others, each
 *    block in decreasing monomial order; rows sorted by (lead, length)
 *    (order.c:98-131). */
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <emmintrin.h>
#include <omp.h>

#include "../../include/gbgpu_f4.h"

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

/* ------------------------------------------------------------ monomials */

/* exponents: 16 bits each, at most EX_MAX; four of them fill a 64-bit
 * word, and the sum of two in-range words has no carry between lanes */
typedef uint16_t ex_t;
#define EX_MAX 32767
#define EX_HIGH 0x8000800080008000ull

typedef struct {
    int nv, nvp, order;
    int erange;      /* set when a product left [0, EX_MAX] */
    ex_t *ev;        /* nvp exponents per monomial */
    uint32_t *deg;
    uint64_t *hv;
    uint64_t *dm;
    int32_t n, cap;
    uint64_t *slot;  /* high 32 bits of the hash, then id + 1 (0 = empty) */
    uint64_t smask;
    uint64_t *rnd;   /* nv hash weights */
    ex_t *tmp;       /* nvp scratch exponents */
} mt_t;

static int mt_init(mt_t *t, int nv, int order) {
    memset(t, 0, sizeof(*t));
    t->nv = nv;
    t->nvp = (nv + 3) & ~3;
    t->order = order;
    t->cap = 1 << 12;
    t->ev = (ex_t *)calloc((size_t)t->cap * (size_t)t->nvp, sizeof(ex_t));
    t->deg = (uint32_t *)malloc((size_t)t->cap * sizeof(uint32_t));
    t->hv = (uint64_t *)malloc((size_t)t->cap * sizeof(uint64_t));
    t->dm = (uint64_t *)malloc((size_t)t->cap * sizeof(uint64_t));
    t->smask = (1u << 13) - 1;
    t->slot = (uint64_t *)calloc((size_t)t->smask + 1, sizeof(uint64_t));
    t->tmp = (ex_t *)calloc((size_t)t->nvp, sizeof(ex_t));
    t->rnd = (uint64_t *)malloc((size_t)nv * sizeof(uint64_t));
    if (!t->rnd) return -1;
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < nv; ++i) {
        /* splitmix64 */
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        t->rnd[i] = z ^ (z >> 31);
    }
    return (t->ev && t->deg && t->hv && t->dm && t->slot && t->tmp) ? 0 : -1;
}

static void mt_free(mt_t *t) {
    free(t->ev); free(t->deg); free(t->hv); free(t->dm); free(t->slot); free(t->tmp);
    free(t->rnd);
    memset(t, 0, sizeof(*t));
}

static inline uint64_t mt_pos(uint64_t h) { return h ^ (h >> 29) ^ (h >> 43); }

/* a slot holds the monomial's hash tag beside its id: a probe passes over
 * slots of other monomials without reading their exponents */
#define SLOT_TAG 0xFFFFFFFF00000000ull
#define SLOT_CLAIMED (~0ull)
static inline uint64_t slot_of(uint64_t h, int32_t id) { return (h & SLOT_TAG) | (uint32_t)(id + 1); }
static inline int32_t slot_id(uint64_t s) { return (int32_t)(uint32_t)s - 1; }

/* exponent vectors are nvp 16-bit lanes, nvp a multiple of 4: compared as
 * 64-bit words */
static inline int ev_eq(const ex_t *a, const ex_t *b, int nvp) {
    const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
    uint64_t d = 0;
    for (int w = 0; w < nvp / 4; ++w) d |= x[w] ^ y[w];
    return d == 0;
}

/* a fresh slot array of nm + 1 slots holding the first n monomials */
static uint64_t *slots_build(const uint64_t *hv, int32_t n, uint64_t nm) {
    uint64_t *sl = (uint64_t *)calloc((size_t)nm + 1, sizeof(uint64_t));
    if (!sl) return NULL;
    for (int32_t i = 0; i < n; ++i) {
        uint64_t p = mt_pos(hv[i]) & nm;
        while (sl[p]) p = (p + 1) & nm;
        sl[p] = slot_of(hv[i], i);
    }
    return sl;
}

static int mt_grow(mt_t *t) {
    if (t->n + 1 >= t->cap) {
        int32_t nc = t->cap * 2;
        ex_t *ev = (ex_t *)realloc(t->ev, (size_t)nc * (size_t)t->nvp * sizeof(ex_t));
        uint32_t *dg = (uint32_t *)realloc(t->deg, (size_t)nc * sizeof(uint32_t));
        uint64_t *hv = (uint64_t *)realloc(t->hv, (size_t)nc * sizeof(uint64_t));
        uint64_t *dm = (uint64_t *)realloc(t->dm, (size_t)nc * sizeof(uint64_t));
        if (ev) t->ev = ev;
        if (dg) t->deg = dg;
        if (hv) t->hv = hv;
        if (dm) t->dm = dm;
        if (!ev || !dg || !hv || !dm) return -1;
        t->cap = nc;
    }
    if ((uint64_t)t->n * 2 >= t->smask) {
        const uint64_t nm = t->smask * 2 + 1;
        uint64_t *sl = slots_build(t->hv, t->n, nm);
        if (!sl) return -1;
        free(t->slot);
        t->slot = sl;
        t->smask = nm;
    }
    return 0;
}

/* id of the monomial with exponents e (nvp of them, zero padded) and hash h */
static int32_t mt_get(mt_t *t, const ex_t *e, uint64_t h) {
    uint64_t p = mt_pos(h) & t->smask;
    for (;;) {
        const uint64_t s = t->slot[p];
        if (!s) break;
        if (((s ^ h) & SLOT_TAG) == 0 && ev_eq(t->ev + (size_t)slot_id(s) * t->nvp, e, t->nvp))
            return slot_id(s);
        p = (p + 1) & t->smask;
    }
    if (mt_grow(t)) return -1;
    if ((uint64_t)t->n * 2 >= t->smask) return -1;
    p = mt_pos(h) & t->smask;
    while (t->slot[p]) p = (p + 1) & t->smask;
    const int32_t id = t->n++;
    memcpy(t->ev + (size_t)id * t->nvp, e, (size_t)t->nvp * sizeof(ex_t));
    uint32_t d = 0;
    uint64_t m = 0;
    for (int i = 0; i < t->nv; ++i) {
        d += e[i];
        if (e[i]) m |= 1ull << (i & 63);
    }
    t->deg[id] = d;
    t->hv[id] = h;
    t->dm[id] = m;
    t->slot[p] = slot_of(h, id);
    return id;
}

/* id of an existing monomial, -1 when absent (no insertion: safe to call
 * from many threads while nothing is inserted) */
static inline int32_t mt_find(const mt_t *t, const ex_t *e, uint64_t h) {
    uint64_t p = mt_pos(h) & t->smask;
    for (;;) {
        const uint64_t s = t->slot[p];
        if (!s) return -1;
        if (((s ^ h) & SLOT_TAG) == 0 && ev_eq(t->ev + (size_t)slot_id(s) * t->nvp, e, t->nvp))
            return slot_id(s);
        p = (p + 1) & t->smask;
    }
}

/* room for `extra` more monomials with neither reallocation nor rehash */
static int mt_reserve(mt_t *t, int64_t extra) {
    const int64_t need = (int64_t)t->n + extra + 1;
    if (need > INT32_MAX / 2) return -1;
    if (need > t->cap) {
        int32_t nc = t->cap;
        while (nc < need) nc *= 2;
        ex_t *ev = (ex_t *)realloc(t->ev, (size_t)nc * (size_t)t->nvp * sizeof(ex_t));
        uint32_t *dg = (uint32_t *)realloc(t->deg, (size_t)nc * sizeof(uint32_t));
        uint64_t *hv = (uint64_t *)realloc(t->hv, (size_t)nc * sizeof(uint64_t));
        uint64_t *dm = (uint64_t *)realloc(t->dm, (size_t)nc * sizeof(uint64_t));
        if (ev) t->ev = ev;
        if (dg) t->deg = dg;
        if (hv) t->hv = hv;
        if (dm) t->dm = dm;
        if (!ev || !dg || !hv || !dm) return -1;
        t->cap = nc;
    }
    if ((uint64_t)need * 2 > t->smask) {
        uint64_t nm = t->smask;
        while ((uint64_t)need * 2 > nm) nm = nm * 2 + 1;
        uint64_t *sl = slots_build(t->hv, t->n, nm);
        if (!sl) return -1;
        free(t->slot);
        t->slot = sl;
        t->smask = nm;
    }
    return 0;
}

/* Insertion from many threads at once, after mt_reserve: a slot is
 * claimed by compare-and-swap (0 -> SLOT_CLAIMED), filled, then published;
 * a thread meeting a claimed slot waits for it.  Ids come from an atomic
 * counter; one at or above lim is not stored (the slot is released and
 * *over set) -- the caller falls back to the serial path. */
static int32_t mt_get_par(mt_t *t, const ex_t *e, uint64_t h, int32_t lim, int *over) {
    uint64_t p = mt_pos(h) & t->smask;
    for (;;) {
        uint64_t s = __atomic_load_n(&t->slot[p], __ATOMIC_ACQUIRE);
        if (s == 0) {
            uint64_t expect = 0;
            if (__atomic_compare_exchange_n(&t->slot[p], &expect, SLOT_CLAIMED, 0, __ATOMIC_ACQ_REL,
                                            __ATOMIC_ACQUIRE)) {
                const int32_t id = __atomic_fetch_add(&t->n, 1, __ATOMIC_RELAXED);
                if (id >= lim) {
                    __atomic_store_n(&t->slot[p], 0, __ATOMIC_RELEASE);
                    __atomic_store_n(over, 1, __ATOMIC_RELAXED);
                    return -1;
                }
                memcpy(t->ev + (size_t)id * t->nvp, e, (size_t)t->nvp * sizeof(ex_t));
                uint32_t d = 0;
                uint64_t m = 0;
                for (int i = 0; i < t->nv; ++i) {
                    d += e[i];
                    if (e[i]) m |= 1ull << (i & 63);
                }
                t->deg[id] = d;
                t->hv[id] = h;
                t->dm[id] = m;
                __atomic_store_n(&t->slot[p], slot_of(h, id), __ATOMIC_RELEASE);
                return id;
            }
            s = expect;
        }
        while (s == SLOT_CLAIMED) s = __atomic_load_n(&t->slot[p], __ATOMIC_ACQUIRE);
        if (s == 0) continue;  /* released by an overflowing claim: look again */
        if (((s ^ h) & SLOT_TAG) == 0 && ev_eq(t->ev + (size_t)slot_id(s) * t->nvp, e, t->nvp))
            return slot_id(s);
        p = (p + 1) & t->smask;
    }
}

static uint64_t mt_hash(const mt_t *t, const ex_t *e) {
    uint64_t h = 0;
    for (int i = 0; i < t->nv; ++i) h += t->rnd[i] * e[i];
    return h;
}

static inline const ex_t *EV(const mt_t *t, int32_t a) { return t->ev + (size_t)a * t->nvp; }

/* a b; -1 with t->erange set when an exponent leaves [0, EX_MAX] */
static int32_t mt_mul(mt_t *t, int32_t a, int32_t b) {
    const ex_t *ea = EV(t, a), *eb = EV(t, b);
    int hi = 0;
    for (int i = 0; i < t->nvp; ++i) {
        const unsigned x = (unsigned)ea[i] + eb[i];
        hi |= x > EX_MAX;
        t->tmp[i] = (ex_t)x;
    }
    if (hi) {
        t->erange = 1;
        return -1;
    }
    return mt_get(t, t->tmp, t->hv[a] + t->hv[b]);
}

/* b / a (a divides b) */
static int32_t mt_div(mt_t *t, int32_t b, int32_t a) {
    const ex_t *ea = EV(t, a), *eb = EV(t, b);
    for (int i = 0; i < t->nvp; ++i) t->tmp[i] = (ex_t)(eb[i] - ea[i]);
    return mt_get(t, t->tmp, t->hv[b] - t->hv[a]);
}

static inline int mt_divides(const mt_t *t, int32_t a, int32_t b) {
    if (t->dm[a] & ~t->dm[b]) return 0;
    if (t->deg[a] > t->deg[b]) return 0;
    const ex_t *ea = EV(t, a), *eb = EV(t, b);
    for (int i = 0; i < t->nv; ++i)
        if (ea[i] > eb[i]) return 0;
    return 1;
}

/* > 0 when a > b in the monomial order */
static inline int mt_cmp(const mt_t *t, int32_t a, int32_t b) {
    if (a == b) return 0;
    const ex_t *ea = EV(t, a), *eb = EV(t, b);
    if (t->order == 0) {
        if (t->deg[a] != t->deg[b]) return t->deg[a] > t->deg[b] ? 1 : -1;
        for (int i = t->nv - 1; i >= 0; --i)
            if (ea[i] != eb[i]) return ea[i] < eb[i] ? 1 : -1;
        return 0;
    }
    for (int i = 0; i < t->nv; ++i)
        if (ea[i] != eb[i]) return ea[i] > eb[i] ? 1 : -1;
    return 0;
}

/* > 0 when gb's initial_input_cmp puts a after b (order.c:38-95) */
static inline int input_cmp(const mt_t *t, int32_t a, int32_t b) {
    if (t->order != 0) return mt_cmp(t, a, b);
    if (t->deg[a] != t->deg[b]) return t->deg[a] > t->deg[b] ? 1 : -1;
    const ex_t *ea = EV(t, a), *eb = EV(t, b);
    for (int i = t->nv - 1; i >= 0; --i)
        if (ea[i] != eb[i]) return ea[i] > eb[i] ? 1 : -1;
    return 0;
}

/* ---------------------------------------------------------------- basis */

typedef struct {
    int32_t len;
    int32_t *m;      /* monomials, decreasing */
    uint32_t *c;     /* coefficients, c[0] = 1 */
} poly_t;

typedef struct {
    poly_t *e;
    int32_t n, cap;
    uint8_t *red;    /* redundant */
    int32_t *lm;     /* non-redundant elements in insertion order */
    int32_t nlm;
} basis_t;

typedef struct {
    int32_t i, j;    /* i < j, basis indices */
    int32_t lcm;
    uint32_t deg;
} pair_t;

typedef struct {
    pair_t *p;
    int64_t n, cap;
} pairs_t;

static int pairs_push(pairs_t *P, pair_t x) {
    if (P->n == P->cap) {
        int64_t nc = P->cap ? 2 * P->cap : 1024;
        pair_t *np = (pair_t *)realloc(P->p, (size_t)nc * sizeof(pair_t));
        if (!np) return -1;
        P->p = np;
        P->cap = nc;
    }
    P->p[P->n++] = x;
    return 0;
}

static int basis_push(basis_t *B, poly_t f) {
    if (B->n == B->cap) {
        int32_t nc = B->cap ? 2 * B->cap : 256;
        poly_t *ne = (poly_t *)realloc(B->e, (size_t)nc * sizeof(poly_t));
        uint8_t *nr = (uint8_t *)realloc(B->red, (size_t)nc);
        int32_t *nl = (int32_t *)realloc(B->lm, (size_t)nc * sizeof(int32_t));
        if (ne) B->e = ne;
        if (nr) B->red = nr;
        if (nl) B->lm = nl;
        if (!ne || !nr || !nl) return -1;
        B->cap = nc;
    }
    B->e[B->n] = f;
    B->red[B->n] = 0;
    B->n++;
    return 0;
}

static uint32_t inv_mod(uint32_t a, uint32_t p) {
    int64_t r0 = p, r1 = a % p, s0 = 0, s1 = 1;
    while (r1) {
        const int64_t q = r0 / r1, t = r0 - q * r1;
        r0 = r1; r1 = t;
        const int64_t u = s0 - q * s1;
        s0 = s1; s1 = u;
    }
    return (uint32_t)(s0 < 0 ? s0 + p : s0);
}

/* Gebauer-Moeller installation of element k (its lead not yet in B->lm).
 * Elements lo .. k-1 came from the same batch: when the lead of one of
 * them divides k's lead, k is redundant from the start and only the pair
 * with that element is kept (its S-polynomial reduces k further) -- gb's
 * rule in insert_and_update_spairs (update.c:96-112). */
static inline int lcm_is(const mt_t *T, int32_t a, int32_t b, int32_t l) {
    /* lcm(a, b) == l, on the exponents */
    const ex_t *ea = EV(T, a), *eb = EV(T, b), *el = EV(T, l);
    for (int v = 0; v < T->nv; ++v)
        if ((ea[v] > eb[v] ? ea[v] : eb[v]) != el[v]) return 0;
    return 1;
}

/* lcm of two exponent vectors (nvp lanes, each <= EX_MAX) into e; returns
 * its degree, and in *up bit v & 63 set where a[v] > b[v].  SSE2: eight
 * lanes at a time; the signed 16-bit max and compare are exact below 2^15. */
static inline uint32_t lcm_lanes(const ex_t *a, const ex_t *b, ex_t *e, int nvp, uint64_t *up) {
    const __m128i one = _mm_set1_epi16(1), zero = _mm_setzero_si128();
    __m128i acc = zero;
    uint64_t m = 0;
    int v = 0;
    for (; v + 8 <= nvp; v += 8) {
        const __m128i x = _mm_loadu_si128((const __m128i *)(a + v));
        const __m128i y = _mm_loadu_si128((const __m128i *)(b + v));
        const __m128i mx = _mm_max_epi16(x, y);
        _mm_storeu_si128((__m128i *)(e + v), mx);
        acc = _mm_add_epi32(acc, _mm_madd_epi16(mx, one));
        const __m128i gt = _mm_cmpgt_epi16(x, y);
        m |= (uint64_t)(_mm_movemask_epi8(_mm_packs_epi16(gt, zero)) & 0xFF) << (v & 63);
    }
    if (v < nvp) {  /* nvp is a multiple of 4: one half vector left */
        const __m128i x = _mm_loadl_epi64((const __m128i *)(a + v));
        const __m128i y = _mm_loadl_epi64((const __m128i *)(b + v));
        const __m128i mx = _mm_max_epi16(x, y);
        _mm_storel_epi64((__m128i *)(e + v), mx);
        acc = _mm_add_epi32(acc, _mm_madd_epi16(mx, one));
        const __m128i gt = _mm_cmpgt_epi16(x, y);
        m |= (uint64_t)(_mm_movemask_epi8(_mm_packs_epi16(gt, zero)) & 0xF) << (v & 63);
    }
    acc = _mm_add_epi32(acc, _mm_shuffle_epi32(acc, 0x4E));
    acc = _mm_add_epi32(acc, _mm_shuffle_epi32(acc, 0xB1));
    *up = m;
    return (uint32_t)_mm_cvtsi128_si32(acc);
}

static inline int ediv(const ex_t *a, const ex_t *b, int nv) {
    for (int v = 0; v < nv; ++v)
        if (a[v] > b[v]) return 0;
    return 1;
}

static int update_pairs(mt_t *T, basis_t *B, pairs_t *P, int32_t k, int32_t lo) {
    const int32_t lk = B->e[k].m[0];
    const int32_t nl = B->nlm;
    const int nvp = T->nvp, nv = T->nv;
    int rc = 0;
    for (int32_t i = lo; i < k; ++i) {
        if (B->red[i] || !mt_divides(T, B->e[i].m[0], lk)) continue;
        pair_t x = {i, k, lk, T->deg[lk]};
        B->red[k] = 1;
        return pairs_push(P, x);
    }
    /* candidate pairs (i, k): lcm exponents and degree kept locally, and
     * the variables of lcm / lead(k) (bit v & 63); only the pairs that
     * survive get a monomial id */
    const size_t nn = nl ? (size_t)nl : 1;
    ex_t *le = (ex_t *)malloc(nn * (size_t)nvp * sizeof(ex_t));
    uint32_t *ld = (uint32_t *)malloc(nn * sizeof(uint32_t));
    uint64_t *am = (uint64_t *)malloc(nn * sizeof(uint64_t));
    uint8_t *st = (uint8_t *)malloc(nn);  /* 0 dropped, 1 kept, 2 coprime */
    const uint32_t dk = T->deg[lk];
    if (!le || !ld || !am || !st) { rc = -1; goto out; }
    {
        const ex_t *ek = EV(T, lk);
        for (int32_t q = 0; q < nl; ++q) {
            const int32_t i = B->lm[q];
            const int32_t li = B->e[i].m[0];
            const ex_t *ei = EV(T, li);
            uint64_t m;
            const uint32_t d = lcm_lanes(ei, ek, le + (size_t)q * nvp, nvp, &m);
            ld[q] = d;
            am[q] = m;
            /* pairs with redundant elements (kept as reducers in lex) are not formed */
            st[q] = B->red[i] ? 0 : (d == T->deg[li] + dk) ? 2 : 1;
        }
    }
    /* old pairs: chain criterion (marks, then compaction) */
    {
        const int64_t np = P->n;
        pair_t *pp = P->p;
        if (np > 0) {
            uint8_t *drop = (uint8_t *)malloc((size_t)np);
            if (!drop) { rc = -1; goto out; }
            for (int64_t q = 0; q < np; ++q) {
                const pair_t x = pp[q];
                drop[q] = mt_divides(T, lk, x.lcm) && !lcm_is(T, B->e[x.i].m[0], lk, x.lcm) &&
                          !lcm_is(T, B->e[x.j].m[0], lk, x.lcm);
            }
            int64_t w = 0;
            for (int64_t q = 0; q < np; ++q)
                if (!drop[q]) pp[w++] = pp[q];
            P->n = w;
            free(drop);
        }
    }
    /* new pairs, gb's rule (update.c:157-195): with the candidates sorted
     * by lcm and, for equal lcms, by basis index, a pair is dropped when
     * an earlier one's lcm divides its lcm (hash.c:351-389) -- so of
     * several pairs with one lcm the one with the oldest element stays --
     * and the survivors whose leads are coprime are dropped afterwards
     * (hash.c:553-557).  Coprime pairs still drop the pairs above them.
     * (Dropped pairs need not be left out as divisors: a dropped divisor
     * has an earlier divisor of its own.)  Each candidate is checked on its
     * own against the candidates before it in (degree, index) order -- the
     * only ones that can drop it -- lowest degree first.  (One element's
     * update is a few hundred microseconds at most: it stays on one thread;
     * a fork and join per element costs more than it saves.) */
    {
        uint8_t *keep = (uint8_t *)malloc(nn);
        int32_t *byd = (int32_t *)malloc(nn * sizeof(int32_t));
        int32_t *pos = (int32_t *)malloc(nn * sizeof(int32_t));
        if (!keep || !byd || !pos) { free(keep); free(byd); free(pos); rc = -1; goto out; }
        uint32_t dlo = UINT32_MAX, dhi = 0;
        for (int32_t q = 0; q < nl; ++q) {
            if (ld[q] < dlo) dlo = ld[q];
            if (ld[q] > dhi) dhi = ld[q];
        }
        /* counting sort by degree, stable in the index */
        const uint32_t nd = nl ? dhi - dlo + 2 : 1;
        int32_t *cnt = (int32_t *)calloc(nd, sizeof(int32_t));
        if (!cnt) { free(keep); free(byd); free(pos); rc = -1; goto out; }
        for (int32_t q = 0; q < nl; ++q) cnt[ld[q] - dlo + 1]++;
        for (uint32_t d = 1; d < nd; ++d) cnt[d] += cnt[d - 1];
        for (int32_t q = 0; q < nl; ++q) {
            pos[q] = cnt[ld[q] - dlo]++;
            byd[pos[q]] = q;
        }
        free(cnt);
        /* lcm(r) | lcm(q) iff lcm(r) / lead(k) divides lcm(q) / lead(k):
         * its variables are a subset of q's.  A quotient of degree 1 is one
         * variable x_v, which divides every quotient of higher degree that
         * holds x_v: sv collects those variables (exact while nv <= 64, when
         * bit v stands for x_v alone). */
        uint64_t sv = 0;
        if (nv <= 64)
            for (int32_t q = 0; q < nl; ++q)
                if (st[q] != 0 && ld[q] == dk + 1) sv |= am[q];
        for (int32_t q = 0; q < nl; ++q) {
            keep[q] = st[q];
            if (st[q] != 1) continue;
            const uint64_t mq = am[q];
            if (ld[q] > dk + 1 && (mq & sv)) {
                keep[q] = 0;
                continue;
            }
            const ex_t *eq = le + (size_t)q * nvp;
            for (int32_t a = 0; a < pos[q]; ++a) {
                const int32_t r = byd[a];
                if (st[r] == 0 || (am[r] & ~mq)) continue;
                const ex_t *er = le + (size_t)r * nvp;
                if (ld[r] == ld[q] ? ev_eq(er, eq, nvp) : ediv(er, eq, nv)) {
                    keep[q] = 0;
                    break;
                }
            }
        }
        memcpy(st, keep, nn);
        free(keep);
        free(byd);
        free(pos);
    }
    for (int32_t q = 0; q < nl; ++q) {
        if (st[q] != 1) continue;
        const ex_t *eq = le + (size_t)q * nvp;
        const int32_t i = B->lm[q];
        const int32_t l = mt_get(T, eq, mt_hash(T, eq));
        if (l < 0) { rc = -1; goto out; }
        pair_t x = {i < k ? i : k, i < k ? k : i, l, T->deg[l]};
        if (pairs_push(P, x)) { rc = -1; goto out; }
    }
    /* redundant elements: dropped from the reducer list for DRL; for lex
     * they stay reducers, as in gb (update.c:247-257 compacts the list for
     * mon_order 0 only) */
    {
        int32_t w = 0;
        for (int32_t q = 0; q < nl; ++q) {
            const int32_t i = B->lm[q];
            if (!B->red[i] && mt_divides(T, lk, B->e[i].m[0])) B->red[i] = 1;
            if (!B->red[i] || T->order != 0) B->lm[w++] = i;
        }
        B->nlm = w;
        B->lm[B->nlm++] = k;
    }
out:
    free(le);
    free(ld);
    free(am);
    free(st);
    return rc;
}

/* pairs whose elements went redundant stay: their S-polynomials are still
 * needed (the chain criterion only removes pairs covered by others) */

/* ------------------------------------------------------ one F4 round */

typedef struct {
    int32_t e;    /* basis element */
    int32_t mul;  /* multiplier monomial */
    int32_t lead; /* lead monomial of the row */
    int32_t len;
    int64_t off;  /* its monomials: rows_t.mon[off .. off + len) */
} row_t;

typedef struct {
    row_t *r;
    int64_t n, cap;
    int32_t *mon;  /* the rows' monomials, found once by symbolic preprocessing */
    int64_t nmon, capmon;
} rows_t;

static int rows_push(rows_t *R, row_t x) {
    if (R->n == R->cap) {
        int64_t nc = R->cap ? 2 * R->cap : 1024;
        row_t *nr = (row_t *)realloc(R->r, (size_t)nc * sizeof(row_t));
        if (!nr) return -1;
        R->r = nr;
        R->cap = nc;
    }
    R->r[R->n++] = x;
    return 0;
}

/* marks on monomials: round-local, plus the reducer search's memory */
typedef struct {
    int32_t *stamp;   /* round of the last mark */
    int32_t *col;     /* column index in this round; -1 unprocessed */
    uint8_t *piv;     /* has a row with this lead */
    int32_t *red;     /* reducer found by the last search (-1: none) */
    int32_t *red_n;   /* basis size at that search (-1: never searched) */
    int32_t cap;
} marks_t;

static int marks_fit(marks_t *M, int32_t n) {
    if (n <= M->cap) return 0;
    int32_t nc = M->cap ? M->cap : 1024;
    while (nc < n) nc *= 2;
    int32_t *s = (int32_t *)realloc(M->stamp, (size_t)nc * sizeof(int32_t));
    int32_t *c = (int32_t *)realloc(M->col, (size_t)nc * sizeof(int32_t));
    uint8_t *p = (uint8_t *)realloc(M->piv, (size_t)nc);
    int32_t *r = (int32_t *)realloc(M->red, (size_t)nc * sizeof(int32_t));
    int32_t *rn = (int32_t *)realloc(M->red_n, (size_t)nc * sizeof(int32_t));
    if (s) M->stamp = s;
    if (c) M->col = c;
    if (p) M->piv = p;
    if (r) M->red = r;
    if (rn) M->red_n = rn;
    if (!s || !c || !p || !r || !rn) return -1;
    for (int32_t i = M->cap; i < nc; ++i) {
        M->stamp[i] = -1;
        M->red_n[i] = -1;
    }
    M->cap = nc;
    return 0;
}

static _Thread_local const mt_t *g_sort_t;  /* per thread: images may run side by side */
static int cmp_mon_desc(const void *a, const void *b) {
    return -mt_cmp(g_sort_t, *(const int32_t *)a, *(const int32_t *)b);
}

typedef struct {
    int32_t key;  /* lead column */
    int32_t len;
    int64_t idx;
} rkey_t;
static int cmp_rkey(const void *a, const void *b) {
    const rkey_t *x = (const rkey_t *)a, *y = (const rkey_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->len != y->len) return x->len < y->len ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* a round's matrix in gb's row format, kept from round to round so that
 * the large arrays are allocated (and their pages touched) once */
typedef struct {
    int64_t *rp;
    int32_t *ci, *rcf, *eid;
    int64_t *cp;
    uint32_t *cf;
    rkey_t *key;
    int64_t *src;  /* row q's monomial ids start at R->mon + src[q] */
    int64_t crp, cci, crcf, ceid, ccp, ccf, ckey, csrc;
} csrbuf_t;

/* room for n elements of size elt; the old contents are not kept */
static int buf_fit(void **p, int64_t *cap, int64_t n, size_t elt) {
    if (n <= *cap) return 0;
    int64_t nc = *cap ? *cap : 1024;
    while (nc < n) nc += nc / 2;
    free(*p);
    *p = malloc((size_t)nc * elt);
    *cap = *p ? nc : 0;
    return *p ? 0 : -1;
}

/* Products u t of one multiplier u with the monomials t of several
 * elements: the rows of a round share few multipliers and their elements
 * share most monomials, so most products recur.  Each thread keeps, per
 * monomial t, the last product it formed with t and the (pass, u) it
 * belongs to; a pass number is never reused, so nothing is ever cleared. */
typedef struct {
    uint64_t key;   /* pass << 32 | (u + 1); 0 never matches */
    int64_t val;    /* u t */
} pcent_t;
typedef struct {
    pcent_t *e;     /* per monomial t: one 16-byte entry, one cache access */
    int32_t cap;
} pcache_t;

typedef struct {
    mt_t T;
    basis_t B;
    pairs_t P;
    marks_t M;
    rows_t R;     /* the round's rows, reused */
    csrbuf_t W;   /* the round's matrix, reused */
    pcache_t *pc; /* product caches, one per thread */
    int32_t npc;
    uint32_t pass;
    /* GPU back end: the elements' coefficient arrays on the device, element
     * i's at cfoff[i] (cfoff[ncfs] = the store's size) */
    gbgpu_cfstore *cfs;
    int64_t *cfoff;
    int32_t ncfs, cfoff_cap;
    uint32_t p;
    int32_t round;
    gbgpu_f4_la_t la;
    void *user;
    gbgpu_f4_out_t *out;
    int info;
} f4_t;

/* GPU back end: the basis elements' coefficient arrays not yet on the
 * device, appended to the store in one piece */
static int cfstore_sync(f4_t *F) {
    const basis_t *B = &F->B;
    gbgpu_f4_gpu_t *g = (gbgpu_f4_gpu_t *)F->user;
    int rc;
    if (!F->cfs && (rc = gbgpu_cfstore_create(g->ctx, F->p, &F->cfs)) != GBGPU_OK) return rc;
    if (F->cfoff_cap < B->n + 1) {
        int64_t *o = (int64_t *)realloc(F->cfoff, (size_t)(B->cap + 1) * sizeof(int64_t));
        if (!o) return GBGPU_ENOMEM;
        F->cfoff = o;
        F->cfoff_cap = B->cap + 1;
    }
    if (F->ncfs == B->n) return GBGPU_OK;
    int64_t n = 0;
    for (int32_t i = F->ncfs; i < B->n; ++i) n += B->e[i].len;
    uint32_t *tmp = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    if (!tmp) return GBGPU_ENOMEM;
    int64_t w = 0;
    for (int32_t i = F->ncfs; i < B->n; ++i) {
        memcpy(tmp + w, B->e[i].c, (size_t)B->e[i].len * sizeof(uint32_t));
        w += B->e[i].len;
    }
    int64_t off = 0;
    rc = gbgpu_cfstore_append(F->cfs, tmp, n, &off);
    free(tmp);
    if (rc != GBGPU_OK) return rc;
    for (int32_t i = F->ncfs; i < B->n; ++i) {
        F->cfoff[i] = off;
        off += B->e[i].len;
    }
    F->cfoff[B->n] = off;
    F->ncfs = B->n;
    return GBGPU_OK;
}

/* turn rows (element, multiplier) into gb's CSR over the given columns
 * (arrays in F->W).  byid: for the GPU back end, which forms the column
 * indices on the device -- m->ci stays NULL and F->W.src gives each row's
 * monomial ids in R->mon, mapped to columns by F->M.col
 * (gbgpu_mat_upload_rows) -- and takes the coefficient arrays from the
 * device store: row q uses element e's array (rcf = e, cp = F->cfoff, m->cf
 * NULL; cfstore_sync first). */
static int build_csr(f4_t *F, rows_t *R, int32_t nc, int32_t ncl, gbgpu_csr_t *m, int byid) {
    csrbuf_t *W = &F->W;
    int64_t nnz = 0;
    for (int64_t r = 0; r < R->n; ++r) nnz += R->r[r].len;
    /* rows by (lead column, length); coefficient arrays: one per distinct
     * element */
    if (buf_fit((void **)&W->key, &W->ckey, R->n ? R->n : 1, sizeof(rkey_t)) ||
        buf_fit((void **)&W->eid, &W->ceid, F->B.n ? F->B.n : 1, sizeof(int32_t)) ||
        buf_fit((void **)&W->rp, &W->crp, R->n + 1, sizeof(int64_t)) ||
        (byid ? buf_fit((void **)&W->src, &W->csrc, R->n ? R->n : 1, sizeof(int64_t))
              : buf_fit((void **)&W->ci, &W->cci, nnz ? nnz : 1, sizeof(int32_t))) ||
        buf_fit((void **)&W->rcf, &W->crcf, R->n ? R->n : 1, sizeof(int32_t)))
        return -1;
    rkey_t *key = W->key;
    int32_t *eid = W->eid, *ci = W->ci, *rcf = W->rcf;
    int64_t *rp = W->rp;
    int32_t ncf = 0;
    int64_t npool = 0;
    for (int64_t r = 0; r < R->n; ++r) {
        key[r].key = F->M.col[R->r[r].lead];
        key[r].len = R->r[r].len;
        key[r].idx = r;
    }
    qsort(key, (size_t)R->n, sizeof(rkey_t), cmp_rkey);
    if (byid) {
        rp[0] = 0;
        for (int64_t q = 0; q < R->n; ++q) {
            const row_t *x = &R->r[key[q].idx];
            rp[q + 1] = rp[q] + x->len;
            rcf[q] = x->e;
            W->src[q] = x->off;
        }
        m->nr = (int32_t)R->n;
        m->nc = nc;
        m->ncl = ncl;
        m->ncf = F->ncfs;
        m->p = F->p;
        m->rp = rp; m->ci = NULL; m->rcf = rcf; m->cp = F->cfoff; m->cf = NULL;
        return 0;
    }
    for (int32_t i = 0; i < F->B.n; ++i) eid[i] = -1;
    for (int64_t r = 0; r < R->n; ++r) {
        const int32_t e = R->r[r].e;
        if (eid[e] < 0) {
            eid[e] = ncf++;
            npool += F->B.e[e].len;
        }
    }
    if (buf_fit((void **)&W->cp, &W->ccp, ncf + 1, sizeof(int64_t)) ||
        buf_fit((void **)&W->cf, &W->ccf, npool ? npool : 1, sizeof(uint32_t)))
        return -1;
    int64_t *cp = W->cp;
    uint32_t *cf = W->cf;
    cp[0] = 0;
    for (int32_t i = 0; i < F->B.n; ++i) {
        if (eid[i] < 0) continue;
        const poly_t *f = &F->B.e[i];
        cp[eid[i] + 1] = f->len;
    }
    for (int32_t i = 0; i < ncf; ++i) cp[i + 1] += cp[i];
    for (int32_t i = 0; i < F->B.n; ++i) {
        if (eid[i] < 0) continue;
        const poly_t *f = &F->B.e[i];
        memcpy(cf + cp[eid[i]], f->c, (size_t)f->len * sizeof(uint32_t));
    }
    rp[0] = 0;
    for (int64_t q = 0; q < R->n; ++q) {
        const row_t *x = &R->r[key[q].idx];
        rp[q + 1] = rp[q] + x->len;
        rcf[q] = eid[x->e];
    }
    const int32_t *colmap = F->M.col;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < R->n; ++q) {
        const row_t *x = &R->r[key[q].idx];
        int32_t *o = ci + rp[q];
        const int32_t *mon = R->mon + x->off;
        for (int32_t t = 0; t < x->len; ++t) o[t] = colmap[mon[t]];
    }
    m->nr = (int32_t)R->n;
    m->nc = nc;
    m->ncl = ncl;
    m->ncf = ncf;
    m->p = F->p;
    m->rp = rp; m->ci = ci; m->rcf = rcf; m->cp = cp; m->cf = cf;
    return 0;
}

/* The reducer of a monomial: the first element of the reducer list (in
 * installation order) whose lead divides it -- gb's find_multiplied_reducer
 * (symbol.c:151-200).  Monomials recur from round to round, so the answer
 * is remembered with the basis size it was found at: elements only join
 * the list at its end, so a remembered reducer stays the first divisor while
 * it is still in the list (for DRL, until it turns redundant), and a
 * monomial that had none need only be tried against the elements installed
 * since. */
static int find_reducer(f4_t *F, int32_t mon) {
    const mt_t *T = &F->T;
    const basis_t *B = &F->B;
    int32_t q0 = 0;
    const int32_t upto = F->M.red_n[mon];
    if (upto >= 0) {
        const int32_t e = F->M.red[mon];
        if (e >= 0 && (T->order != 0 || !B->red[e])) return e;
        if (e < 0) {
            /* the list is in installation order: its first element >= upto */
            int32_t lo = 0, hi = B->nlm;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (B->lm[mid] < upto) lo = mid + 1;
                else hi = mid;
            }
            q0 = lo;
        }
    }
    int32_t found = -1;
    for (int32_t q = q0; q < B->nlm; ++q) {
        const int32_t i = B->lm[q];
        if (mt_divides(T, B->e[i].m[0], mon)) {
            found = i;
            break;
        }
    }
    F->M.red[mon] = found;
    F->M.red_n[mon] = B->n;
    return found;
}

/* A round's rows are added in batches: stubs (element, multiplier) first,
 * then the monomials of the whole batch -- looked up in parallel while the
 * table is read only, the few new ones inserted afterwards in row order --
 * then one pass marking the columns (new ones queued) and the leads. */
static int row_stub(rows_t *R, const basis_t *B, int32_t e, int32_t mul) {
    const int32_t len = B->e[e].len;
    if (R->nmon + len > R->capmon) {
        int64_t nc = R->capmon ? R->capmon : 1 << 16;
        while (nc < R->nmon + len) nc *= 2;
        int32_t *nm = (int32_t *)realloc(R->mon, (size_t)nc * sizeof(int32_t));
        if (!nm) return -1;
        R->mon = nm;
        R->capmon = nc;
    }
    row_t x = {e, mul, -1, len, R->nmon};
    R->nmon += len;
    return rows_push(R, x);
}

/* a product cache per thread, each with room for every monomial */
static int pcache_fit(f4_t *F, int32_t n) {
    const int nth = omp_get_max_threads();
    if (F->npc < nth) {
        pcache_t *np = (pcache_t *)realloc(F->pc, (size_t)nth * sizeof(pcache_t));
        if (!np) return -1;
        memset(np + F->npc, 0, (size_t)(nth - F->npc) * sizeof(pcache_t));
        F->pc = np;
        F->npc = nth;
    }
    for (int i = 0; i < nth; ++i) {
        pcache_t *c = &F->pc[i];
        if (c->cap >= n) continue;
        int32_t nc = c->cap ? c->cap : 4096;
        while (nc < n) nc *= 2;
        pcent_t *e = (pcent_t *)realloc(c->e, (size_t)nc * sizeof(pcent_t));
        if (!e) return -1;
        c->e = e;
        memset(c->e + c->cap, 0, (size_t)(nc - c->cap) * sizeof(pcent_t));
        c->cap = nc;
    }
    return 0;
}

static int cmp_u64(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

/* marks of first sight, gathered per thread and queued together */
typedef struct {
    int32_t *v;
    int64_t n, cap;
    int bad;
} mlist_t;

/* first sight of monomial m in this round: whichever thread swaps the stamp
 * queues it (the queue order does not matter: columns are sorted, reducers
 * do not depend on the order) */
static inline void mark_seen(f4_t *F, int32_t m, int32_t round, mlist_t *L) {
    int32_t *stamp = F->M.stamp;
    if (__atomic_load_n(&stamp[m], __ATOMIC_RELAXED) == round) return;
    if (__atomic_exchange_n(&stamp[m], round, __ATOMIC_RELAXED) == round) return;
    F->M.col[m] = -1;
    F->M.piv[m] = 0;
    if (L->n == L->cap) {
        const int64_t nc = L->cap ? 2 * L->cap : 1024;
        int32_t *nv = (int32_t *)realloc(L->v, (size_t)nc * sizeof(int32_t));
        if (!nv) {
            L->bad = 1;
            return;
        }
        L->v = nv;
        L->cap = nc;
    }
    L->v[L->n++] = m;
}

static int queue_push(int32_t **queue, int64_t *nq, int64_t *capq, const int32_t *v, int64_t n);
static int rows_mark(f4_t *F, rows_t *R, int64_t r0, int64_t r1, int32_t **queue, int64_t *nq,
                     int64_t *capq);

/* a thread's marks into the queue (inside a parallel region) */
static void mlist_flush(mlist_t *L, int32_t **queue, int64_t *nq, int64_t *capq, int *bad) {
#pragma omp critical
    {
        if (L->bad || (L->n && queue_push(queue, nq, capq, L->v, L->n))) *bad = 1;
    }
    free(L->v);
}

/* The monomials of rows r0 .. r1-1: every product u t looked up in the
 * table while it is read only (rows taken by multiplier, so that the
 * product caches hit), then the missing ones inserted by all threads.  A
 * product met for the first time in this round is marked and queued as it
 * is formed (a product cache hit was marked when it was first formed); the
 * rows' leads are marked as pivots at the end. */
static int rows_products(f4_t *F, rows_t *R, int64_t r0, int64_t r1, int32_t **queue, int64_t *nq,
                         int64_t *capq) {
    mt_t *T = &F->T;
    const basis_t *B = &F->B;
    const int nw = T->nvp / 4;
    const int64_t nr = r1 - r0;
    const int32_t round = F->round;
    if (nr <= 0) return 0;
    /* the rows by multiplier */
    uint64_t *ord = (uint64_t *)malloc((size_t)nr * sizeof(uint64_t));
    if (!ord || pcache_fit(F, T->n) || marks_fit(&F->M, T->n)) {
        free(ord);
        return -1;
    }
    for (int64_t r = r0; r < r1; ++r)
        ord[r - r0] = (uint64_t)(uint32_t)R->r[r].mul << 32 | (uint64_t)(r - r0);
    qsort(ord, (size_t)nr, sizeof(uint64_t), cmp_u64);
    int64_t nmiss = 0;
    uint64_t high = 0;
    int bad = 0;
    const uint64_t pass = (uint64_t)(++F->pass) << 32;
#pragma omp parallel reduction(+ : nmiss) reduction(| : high)
    {
        uint64_t tmp[nw];
        pcache_t *pc = &F->pc[omp_get_thread_num()];
        mlist_t L = {NULL, 0, 0, 0};
#pragma omp for schedule(dynamic, 8)
        for (int64_t q = 0; q < nr; ++q) {
            const row_t *x = &R->r[r0 + (int64_t)(uint32_t)ord[q]];
            const poly_t *f = &B->e[x->e];
            const uint64_t *em = (const uint64_t *)EV(T, x->mul);
            const uint64_t hm = T->hv[x->mul];
            const uint64_t key = pass | (uint32_t)(x->mul + 1);
            int32_t *o = R->mon + x->off;
            for (int32_t t = 0; t < f->len; ++t) {
                const int32_t mt = f->m[t];
                const pcent_t ce = pc->e[mt];
                if (ce.key == key) {
                    o[t] = (int32_t)ce.val;
                    continue;
                }
                const uint64_t *et = (const uint64_t *)EV(T, mt);
                /* four 16-bit exponents per word, each <= EX_MAX: no carries
                 * between lanes; a lane above EX_MAX shows in its top bit */
                uint64_t hi = 0;
                for (int w = 0; w < nw; ++w) {
                    tmp[w] = em[w] + et[w];
                    hi |= tmp[w];
                }
                if (hi & EX_HIGH) {
                    high |= hi;
                    o[t] = 0;
                    continue;
                }
                const int32_t id = mt_find(T, (const ex_t *)tmp, hm + T->hv[mt]);
                o[t] = id;
                pc->e[mt].key = key;
                pc->e[mt].val = id;
                if (id >= 0) mark_seen(F, id, round, &L);
                else ++nmiss;
            }
        }
        mlist_flush(&L, queue, nq, capq, &bad);
    }
    if (high & EX_HIGH) {
        free(ord);
        T->erange = 1;
        return -1;
    }
    int rc = bad ? -1 : 0;
    if (nmiss && rc == 0) {
        /* new monomials: inserted by all threads into reserved room (each
         * is missed at least once, so nmiss bounds their number; the
         * room is capped) */
        int64_t room = nmiss;
        const int64_t cap_room = (int64_t)T->n * 2 > (1 << 22) ? (int64_t)T->n * 2 : (1 << 22);
        if (room > cap_room) room = cap_room;
        int done = 0;
        if (mt_reserve(T, room) == 0 && marks_fit(&F->M, (int32_t)(T->n + room)) == 0) {
            const int32_t lim = (int32_t)(T->n + room);
            const uint64_t pass2 = (uint64_t)(++F->pass) << 32;
            int over = 0;
#pragma omp parallel
            {
                uint64_t tmp[nw];
                pcache_t *pc = &F->pc[omp_get_thread_num()];
                mlist_t L = {NULL, 0, 0, 0};
#pragma omp for schedule(dynamic, 8)
                for (int64_t q = 0; q < nr; ++q) {
                    const row_t *x = &R->r[r0 + (int64_t)(uint32_t)ord[q]];
                    const poly_t *f = &B->e[x->e];
                    const uint64_t *em = (const uint64_t *)EV(T, x->mul);
                    const uint64_t hm = T->hv[x->mul];
                    const uint64_t key = pass2 | (uint32_t)(x->mul + 1);
                    int32_t *o = R->mon + x->off;
                    for (int32_t t = 0; t < f->len; ++t) {
                        if (o[t] >= 0) continue;
                        const int32_t mt = f->m[t];
                        const pcent_t ce = pc->e[mt];
                        if (ce.key == key && ce.val >= 0) {
                            o[t] = (int32_t)ce.val;
                            continue;
                        }
                        const uint64_t *et = (const uint64_t *)EV(T, mt);
                        for (int w = 0; w < nw; ++w) tmp[w] = em[w] + et[w];
                        const int32_t id =
                            mt_get_par(T, (const ex_t *)tmp, hm + T->hv[mt], lim, &over);
                        o[t] = id;
                        pc->e[mt].key = key;
                        pc->e[mt].val = id;
                        if (id >= 0) mark_seen(F, id, round, &L);
                    }
                }
                mlist_flush(&L, queue, nq, capq, &bad);
            }
            if (T->n > lim) T->n = lim;
            done = !over;
            if (bad) rc = -1;
        }
        if (!done && rc == 0) {
            /* whatever is left (no room, or the room ran out): one thread,
             * then a marking pass over the rows */
            for (int64_t r = r0; r < r1 && rc == 0; ++r) {
                const row_t *x = &R->r[r];
                const poly_t *f = &B->e[x->e];
                int32_t *o = R->mon + x->off;
                for (int32_t t = 0; t < f->len; ++t)
                    if (o[t] < 0 && (o[t] = mt_mul(T, x->mul, f->m[t])) < 0) {
                        rc = -1;
                        break;
                    }
            }
            if (rc == 0 && rows_mark(F, R, r0, r1, queue, nq, capq)) rc = -1;
        }
    }
    free(ord);
    if (rc) return rc;
    for (int64_t r = r0; r < r1; ++r) {
        R->r[r].lead = R->mon[R->r[r].off];
        F->M.piv[R->r[r].lead] = 1;
    }
    return 0;
}

static int queue_push(int32_t **queue, int64_t *nq, int64_t *capq, const int32_t *v, int64_t n) {
    if (*nq + n > *capq) {
        int64_t nc = *capq ? *capq : 4096;
        while (nc < *nq + n) nc *= 2;
        int32_t *nqp = (int32_t *)realloc(*queue, (size_t)nc * sizeof(int32_t));
        if (!nqp) return -1;
        *queue = nqp;
        *capq = nc;
    }
    memcpy(*queue + *nq, v, (size_t)n * sizeof(int32_t));
    *nq += n;
    return 0;
}

/* marking pass over rows r0 .. r1-1 (the serial insertion path of
 * rows_products, whose products were not marked as they were formed) */
static int rows_mark(f4_t *F, rows_t *R, int64_t r0, int64_t r1, int32_t **queue, int64_t *nq,
                     int64_t *capq) {
    if (marks_fit(&F->M, F->T.n)) return -1;
    const int32_t round = F->round;
    int32_t *stamp = F->M.stamp;
    int bad = 0;
#pragma omp parallel
    {
        int32_t *loc = NULL;
        int64_t nloc = 0, cap = 0;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = r0; r < r1; ++r) {
            const row_t *x = &R->r[r];
            const int32_t *mon = R->mon + x->off;
            for (int32_t t = 0; t < x->len; ++t) {
                const int32_t m = mon[t];
                if (__atomic_load_n(&stamp[m], __ATOMIC_RELAXED) == round) continue;
                if (__atomic_exchange_n(&stamp[m], round, __ATOMIC_RELAXED) == round) continue;
                F->M.col[m] = -1;
                F->M.piv[m] = 0;
                if (nloc == cap) {
                    cap = cap ? 2 * cap : 1024;
                    int32_t *nl2 = (int32_t *)realloc(loc, (size_t)cap * sizeof(int32_t));
                    if (!nl2) { bad = 1; cap = nloc; continue; }
                    loc = nl2;
                }
                loc[nloc++] = m;
            }
        }
#pragma omp critical
        {
            if (!bad && nloc && queue_push(queue, nq, capq, loc, nloc)) bad = 1;
        }
        free(loc);
    }
    if (bad) return -1;
    for (int64_t r = r0; r < r1; ++r) F->M.piv[R->r[r].lead] = 1;
    return 0;
}

/* symbolic preprocessing: a reducer for every queued column monomial that
 * is no lead yet and is divisible by a basis lead, until no new column
 * appears; the reducers of one generation of columns are searched in
 * parallel (the result does not depend on the order) */
static int symbolic(f4_t *F, rows_t *R, int32_t **queue, int64_t *nq, int64_t *capq) {
    mt_t *T = &F->T;
    const basis_t *B = &F->B;
    int64_t qdone = 0;
    while (qdone < *nq) {
        const int64_t qend = *nq;
        int32_t *red = (int32_t *)malloc((size_t)(qend - qdone) * sizeof(int32_t));
        if (!red) return -1;
        const int32_t *qq = *queue;
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t q = qdone; q < qend; ++q) {
            const int32_t mon = qq[q];
            red[q - qdone] = F->M.piv[mon] ? -1 : find_reducer(F, mon);
        }
        const int64_t r0 = R->n;
        for (int64_t q = qdone; q < qend; ++q) {
            const int32_t e = red[q - qdone];
            if (e < 0) continue;
            const int32_t mul = mt_div(T, qq[q], B->e[e].m[0]);
            if (mul < 0 || row_stub(R, B, e, mul)) { free(red); return -1; }
        }
        free(red);
        if (rows_products(F, R, r0, R->n, queue, nq, capq)) return -1;
        qdone = qend;
    }
    return 0;
}

typedef struct {
    int32_t e, mul;
} em_t;
static int cmp_em(const void *a, const void *b) {
    const em_t *x = (const em_t *)a, *y = (const em_t *)b;
    if (x->e != y->e) return x->e < y->e ? -1 : 1;
    return x->mul < y->mul ? -1 : (x->mul > y->mul);
}

static int cmp_pair_deg(const void *a, const void *b) {
    const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
    if (x->deg != y->deg) return x->deg < y->deg ? -1 : 1;
    return 0;
}

static int run_la(f4_t *F, const gbgpu_csr_t *m, int interreduce, gbgpu_rows_t *out) {
    const double t0 = now_ms();
    const int rc = F->la(F->user, m, interreduce, out);
    F->out->la_ms += now_ms() - t0;
    return rc;
}

/* the GPU back end (gbgpu_f4_gpu_la, la_option 1) on a matrix whose column
 * indices the device forms from the rows' monomial ids (build_csr, byid) */
static int run_la_byid(f4_t *F, const gbgpu_csr_t *m, gbgpu_rows_t *out) {
    const double t0 = now_ms();
    gbgpu_f4_gpu_t *g = (gbgpu_f4_gpu_t *)F->user;
    gbgpu_mat *dm = NULL;
    int32_t np = 0;
    int rc = gbgpu_mat_upload_rows(g->ctx, m, F->R.mon, F->R.nmon, F->W.src, F->M.col, F->T.n,
                                   F->cfs, &dm);
    if (rc == GBGPU_OK) rc = gbgpu_la_sparse_dense(g->ctx, dm, g->form, &np);
    if (rc == GBGPU_OK) {
        g->device_ms += gbgpu_phase_ms(g->ctx, 2);
        rc = gbgpu_rows_download(g->ctx, dm, out);
    }
    gbgpu_mat_free(dm);
    F->out->la_ms += now_ms() - t0;
    return rc;
}

/* new basis element from row k of the LA output (columns -> monomials) */
static int take_row(f4_t *F, const gbgpu_rows_t *o, int32_t k, const int32_t *cols, poly_t *f) {
    const int64_t b = o->rp[k], e = o->rp[k + 1];
    f->len = (int32_t)(e - b);
    f->m = (int32_t *)malloc((size_t)f->len * sizeof(int32_t));
    f->c = (uint32_t *)malloc((size_t)f->len * sizeof(uint32_t));
    if (!f->m || !f->c) return -1;
    for (int64_t q = b; q < e; ++q) {
        f->m[q - b] = cols[o->ci[q]];
        f->c[q - b] = o->cf[q];
    }
    (void)F;
    return 0;
}

/* FNV-1a over the LA result (row lengths, columns, coefficients): the
 * round's echelon form in 64 bits, for comparing runs round by round */
static uint64_t rows_hash(const gbgpu_rows_t *o) {
    uint64_t h = 0xCBF29CE484222325ull;
#define MIX(v) do { h ^= (uint64_t)(uint32_t)(v); h *= 0x100000001B3ull; } while (0)
    MIX(o->np);
    for (int32_t k = 0; k < o->np; ++k) {
        MIX(o->rp[k + 1] - o->rp[k]);
        for (int64_t q = o->rp[k]; q < o->rp[k + 1]; ++q) {
            MIX(o->ci[q]);
            MIX(o->cf[q]);
        }
    }
#undef MIX
    return h;
}

static int f4_round(f4_t *F) {
    mt_t *T = &F->T;
    basis_t *B = &F->B;
    pairs_t *P = &F->P;
    int rc = GBGPU_ENOMEM;
    ++F->round;
    double tp = now_ms();
    /* selection: all pairs of minimal degree */
    qsort(P->p, (size_t)P->n, sizeof(pair_t), cmp_pair_deg);
    const uint32_t dmin = P->p[0].deg;
    int64_t nsel = 0;
    while (nsel < P->n && P->p[nsel].deg == dmin) ++nsel;
    /* generators of the selected pairs, one row per distinct (e, mul) */
    em_t *gen = (em_t *)malloc((size_t)(2 * nsel) * sizeof(em_t));
    rows_t *R = &F->R;
    R->n = 0;
    R->nmon = 0;
    int32_t *queue = NULL;
    int64_t nq = 0, capq = 0;
    int32_t *cols = NULL;
    gbgpu_rows_t out = {0};
    if (!gen) goto done;
    for (int64_t q = 0; q < nsel; ++q) {
        const pair_t x = P->p[q];
        gen[2 * q].e = x.i;
        gen[2 * q].mul = mt_div(T, x.lcm, B->e[x.i].m[0]);
        gen[2 * q + 1].e = x.j;
        gen[2 * q + 1].mul = mt_div(T, x.lcm, B->e[x.j].m[0]);
        if (gen[2 * q].mul < 0 || gen[2 * q + 1].mul < 0) goto done;
    }
    memmove(P->p, P->p + nsel, (size_t)(P->n - nsel) * sizeof(pair_t));
    P->n -= nsel;
    qsort(gen, (size_t)(2 * nsel), sizeof(em_t), cmp_em);
    for (int64_t q = 0; q < 2 * nsel; ++q) {
        if (q > 0 && gen[q].e == gen[q - 1].e && gen[q].mul == gen[q - 1].mul) continue;
        if (row_stub(R, B, gen[q].e, gen[q].mul)) goto done;
    }
    if (rows_products(F, R, 0, R->n, &queue, &nq, &capq)) goto done;
    const int64_t nspair_rows = R->n;
    F->out->select_ms += now_ms() - tp;
    tp = now_ms();
    if (symbolic(F, R, &queue, &nq, &capq)) goto done;
    F->out->symbolic_ms += now_ms() - tp;
    tp = now_ms();
    /* columns: pivot columns, then the rest, each decreasing */
    cols = (int32_t *)malloc((size_t)(nq ? nq : 1) * sizeof(int32_t));
    if (!cols) goto done;
    int32_t ncl = 0, nc = (int32_t)nq;
    for (int64_t q = 0; q < nq; ++q)
        if (F->M.piv[queue[q]]) cols[ncl++] = queue[q];
    {
        int32_t w = ncl;
        for (int64_t q = 0; q < nq; ++q)
            if (!F->M.piv[queue[q]]) cols[w++] = queue[q];
    }
    g_sort_t = T;
    qsort(cols, (size_t)ncl, sizeof(int32_t), cmp_mon_desc);
    qsort(cols + ncl, (size_t)(nc - ncl), sizeof(int32_t), cmp_mon_desc);
    for (int32_t c = 0; c < nc; ++c) F->M.col[cols[c]] = c;
    gbgpu_csr_t m;
    /* the GPU back end takes the rows as monomial ids */
    const int byid = F->la == gbgpu_f4_gpu_la && F->user;
    if (byid && (rc = cfstore_sync(F)) != GBGPU_OK) goto done;
    rc = GBGPU_ENOMEM;
    if (build_csr(F, R, nc, ncl, &m, byid)) goto done;
    int64_t nnz = m.rp[m.nr];
    F->out->build_ms += now_ms() - tp;
    if (m.nr > F->out->max_rows) F->out->max_rows = m.nr;
    if (nc > F->out->max_cols) F->out->max_cols = nc;
    if (nnz > F->out->max_nnz) F->out->max_nnz = nnz;
    F->out->rows_total += m.nr - ncl;
    const double t0 = now_ms();
    if ((rc = byid ? run_la_byid(F, &m, &out) : run_la(F, &m, 0, &out)) != 0) goto done;
    const double tla = now_ms() - t0;
    /* new elements in the order the LA returns them (decreasing lead) */
    rc = GBGPU_ENOMEM;
    tp = now_ms();
    const int32_t lo = B->n;
    for (int32_t k = 0; k < out.np; ++k) {
        poly_t f;
        if (take_row(F, &out, k, cols, &f)) goto done;
        if (basis_push(B, f)) goto done;
        if (update_pairs(T, B, P, B->n - 1, lo)) goto done;
    }
    F->out->update_ms += now_ms() - tp;
    {
        /* the round's record: what gb's info_level 2 line shows */
        const int32_t r = F->out->rounds;
        int64_t *tr = (int64_t *)realloc(F->out->trace,
                                         (size_t)(r + 1) * GBGPU_F4_TRACE_W * sizeof(int64_t));
        if (!tr) goto done;
        F->out->trace = tr;
        int64_t *x = tr + (size_t)r * GBGPU_F4_TRACE_W;
        x[0] = dmin; x[1] = nsel; x[2] = m.nr; x[3] = nc; x[4] = ncl; x[5] = nnz; x[6] = out.np;
        x[7] = (int64_t)rows_hash(&out);
    }
    if (F->info > 1)
        printf("%3u  %6lld sel  %8lld pairs  %7d x %-7d (%lld spair rows)  %6.2f%%  %5d new  %8.1f ms LA\n",
               dmin, (long long)nsel, (long long)P->n, m.nr, nc, (long long)nspair_rows,
               nc && m.nr ? 100.0 * (double)nnz / ((double)m.nr * nc) : 0.0, out.np, tla);
    if (F->info > 1) fflush(stdout);
    rc = GBGPU_OK;
done:
    gbgpu_rows_free(&out);
    free(gen); free(queue); free(cols);
    return rc;
}

/* interreduction of the non-redundant elements (reduce_basis, f4.c:90-179) */
static int f4_reduce(f4_t *F) {
    mt_t *T = &F->T;
    basis_t *B = &F->B;
    int rc = GBGPU_ENOMEM;
    ++F->round;
    rows_t *R = &F->R;
    R->n = 0;
    R->nmon = 0;
    int32_t *queue = NULL, *cols = NULL;
    int64_t nq = 0, capq = 0;
    gbgpu_rows_t out = {0};
    int32_t one;
    {
        ex_t *z = (ex_t *)calloc((size_t)T->nvp, sizeof(ex_t));
        if (!z) goto done;
        one = mt_get(T, z, 0);
        free(z);
        if (one < 0) goto done;
    }
    for (int32_t q = 0; q < B->nlm; ++q)
        if (row_stub(R, B, B->lm[q], one)) goto done;
    if (rows_products(F, R, 0, R->n, &queue, &nq, &capq)) goto done;
    /* reducers for the tail monomials divisible by a lead, so that the
     * reduced rows have no such monomial left */
    if (symbolic(F, R, &queue, &nq, &capq)) goto done;
    cols = (int32_t *)malloc((size_t)(nq ? nq : 1) * sizeof(int32_t));
    if (!cols) goto done;
    memcpy(cols, queue, (size_t)nq * sizeof(int32_t));
    g_sort_t = T;
    qsort(cols, (size_t)nq, sizeof(int32_t), cmp_mon_desc);
    for (int64_t c = 0; c < nq; ++c) F->M.col[cols[c]] = (int32_t)c;
    gbgpu_csr_t m;
    if (build_csr(F, R, (int32_t)nq, 0, &m, 0)) goto done;
    if ((rc = run_la(F, &m, 1, &out)) != 0) goto done;
    rc = GBGPU_ENOMEM;
    /* the reduced basis replaces the old one: the rows whose lead no
     * other row's lead divides (the others are the reducers' rows), by
     * decreasing lead as gb lists them (f4.c:151-162) */
    B->nlm = 0;
    for (int32_t k = out.np - 1; k >= 0; --k) {
        const int32_t lead = cols[out.ci[out.rp[k]]];
        int mult = 0;
        for (int32_t j = 0; j < out.np && !mult; ++j)
            if (j != k && mt_divides(T, cols[out.ci[out.rp[j]]], lead)) mult = 1;
        if (mult) continue;
        poly_t f;
        if (take_row(F, &out, k, cols, &f)) goto done;
        if (basis_push(B, f)) goto done;
        B->lm[B->nlm++] = B->n - 1;
    }
    rc = GBGPU_OK;
done:
    gbgpu_rows_free(&out);
    free(queue); free(cols);
    return rc;
}

static int cmp_i32(const void *a, const void *b) {
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return x < y ? -1 : (x > y);
}

int gbgpu_f4(const gbgpu_f4_in_t *in, gbgpu_f4_la_t la, void *user, gbgpu_f4_out_t *out) {
    if (!in || !la || !out) return GBGPU_EINVAL;
    memset(out, 0, sizeof(*out));
    if (in->nr_vars < 1 || in->nr_vars > GBGPU_F4_MAX_VARS || in->nr_gens < 0 || in->field_char < 2 ||
        (in->mon_order != 0 && in->mon_order != 1) || !in->lens || !in->exps || !in->cfs)
        return GBGPU_EINVAL;
    const double t_start = now_ms();
    f4_t F;
    memset(&F, 0, sizeof(F));
    F.p = (uint32_t)in->field_char;
    F.la = la;
    F.user = user;
    F.out = out;
    F.info = in->info_level;
    int rc = GBGPU_ENOMEM;
    if (mt_init(&F.T, in->nr_vars, in->mon_order)) goto done;
    /* input: terms sorted decreasing, equal monomials merged, zero terms
     * dropped, made monic; zero generators dropped */
    {
        int64_t off = 0;
        const int nv = in->nr_vars;
        for (int32_t g = 0; g < in->nr_gens; ++g) {
            const int32_t len = in->lens[g];
            if (len < 0) { rc = GBGPU_EINVAL; goto done; }
            int32_t *mm = (int32_t *)malloc((size_t)(len ? len : 1) * sizeof(int32_t));
            uint32_t *cc = (uint32_t *)malloc((size_t)(len ? len : 1) * sizeof(uint32_t));
            int32_t *ord = (int32_t *)malloc((size_t)(len ? len : 1) * sizeof(int32_t));
            if (!mm || !cc || !ord) { free(mm); free(cc); free(ord); goto done; }
            for (int32_t t = 0; t < len; ++t) {
                ex_t *e = F.T.tmp;
                memset(e, 0, (size_t)F.T.nvp * sizeof(ex_t));
                for (int i = 0; i < nv; ++i) {
                    const int32_t x = in->exps[(off + t) * nv + i];
                    if (x < 0 || x > EX_MAX) { free(mm); free(cc); free(ord); rc = GBGPU_ERANGE; goto done; }
                    e[i] = (ex_t)x;
                }
                ex_t ebuf[F.T.nvp];
                memcpy(ebuf, e, (size_t)F.T.nvp * sizeof(ex_t));
                mm[t] = mt_get(&F.T, ebuf, mt_hash(&F.T, ebuf));
                if (mm[t] < 0) { free(mm); free(cc); free(ord); goto done; }
                int64_t c = (int64_t)in->cfs[off + t] % (int64_t)F.p;
                if (c < 0) c += F.p;
                cc[t] = (uint32_t)c;
                ord[t] = t;
            }
            off += len;
            /* insertion sort by monomial, decreasing */
            for (int32_t a = 1; a < len; ++a) {
                const int32_t x = ord[a];
                int32_t b = a - 1;
                while (b >= 0 && mt_cmp(&F.T, mm[ord[b]], mm[x]) < 0) {
                    ord[b + 1] = ord[b];
                    --b;
                }
                ord[b + 1] = x;
            }
            poly_t f = {0, NULL, NULL};
            f.m = (int32_t *)malloc((size_t)(len ? len : 1) * sizeof(int32_t));
            f.c = (uint32_t *)malloc((size_t)(len ? len : 1) * sizeof(uint32_t));
            if (!f.m || !f.c) { free(f.m); free(f.c); free(mm); free(cc); free(ord); goto done; }
            for (int32_t a = 0; a < len; ++a) {
                const int32_t t = ord[a];
                if (f.len > 0 && f.m[f.len - 1] == mm[t]) {
                    f.c[f.len - 1] = (uint32_t)(((uint64_t)f.c[f.len - 1] + cc[t]) % F.p);
                } else {
                    f.m[f.len] = mm[t];
                    f.c[f.len] = cc[t];
                    f.len++;
                }
            }
            free(mm); free(cc); free(ord);
            int32_t w = 0;
            for (int32_t a = 0; a < f.len; ++a)
                if (f.c[a]) { f.m[w] = f.m[a]; f.c[w] = f.c[a]; ++w; }
            f.len = w;
            if (f.len == 0) { free(f.m); free(f.c); continue; }
            const uint32_t inv = inv_mod(f.c[0], F.p);
            for (int32_t a = 0; a < f.len; ++a) f.c[a] = (uint32_t)((uint64_t)f.c[a] * inv % F.p);
            if (basis_push(&F.B, f)) { free(f.m); free(f.c); goto done; }
        }
    }
    /* the input order of gb's initial_input_cmp (order.c:38-95, sorted at
     * f4.c:258-259), then installed one by one (f4.c:272-275): for lex,
     * increasing lead; for DRL, increasing lead degree and, within a
     * degree, DECREASING lead (that comparator returns e_a - e_b at the
     * last differing variable, the opposite of the DRL comparison) */
    {
        const int32_t n0 = F.B.n;
        int32_t *ord = (int32_t *)malloc((size_t)(n0 ? n0 : 1) * sizeof(int32_t));
        poly_t *tmp = (poly_t *)malloc((size_t)(n0 ? n0 : 1) * sizeof(poly_t));
        if (!ord || !tmp) { free(ord); free(tmp); goto done; }
        for (int32_t i = 0; i < n0; ++i) ord[i] = i;
        for (int32_t a = 1; a < n0; ++a) {
            const int32_t x = ord[a];
            int32_t b = a - 1;
            while (b >= 0 && input_cmp(&F.T, F.B.e[ord[b]].m[0], F.B.e[x].m[0]) > 0) {
                ord[b + 1] = ord[b];
                --b;
            }
            ord[b + 1] = x;
        }
        for (int32_t i = 0; i < n0; ++i) tmp[i] = F.B.e[ord[i]];
        memcpy(F.B.e, tmp, (size_t)n0 * sizeof(poly_t));
        free(ord);
        free(tmp);
        for (int32_t k = 0; k < n0; ++k)
            if (update_pairs(&F.T, &F.B, &F.P, k, 0)) goto done;
    }
    while (F.P.n > 0) {
        if ((rc = f4_round(&F)) != GBGPU_OK) goto done;
        out->rounds++;
    }
    /* drop the redundant elements (f4.c:330-338); the reducer search
     * forgets what it found, since the list now changes in the middle */
    {
        int32_t w = 0;
        for (int32_t q = 0; q < F.B.nlm; ++q)
            if (!F.B.red[F.B.lm[q]]) F.B.lm[w++] = F.B.lm[q];
        F.B.nlm = w;
        for (int32_t i = 0; i < F.M.cap; ++i) F.M.red_n[i] = -1;
    }
    if (in->reduce_gb == 1 && F.B.nlm > 0) {
        if ((rc = f4_reduce(&F)) != GBGPU_OK) goto done;
    }
    /* export the non-redundant elements (in installation order; the
     * reduced basis by increasing lead) */
    {
        rc = GBGPU_ENOMEM;
        int32_t *lm = (int32_t *)malloc((size_t)(F.B.nlm ? F.B.nlm : 1) * sizeof(int32_t));
        if (!lm) goto done;
        memcpy(lm, F.B.lm, (size_t)F.B.nlm * sizeof(int32_t));
        if (in->reduce_gb != 1) qsort(lm, (size_t)F.B.nlm, sizeof(int32_t), cmp_i32);
        int64_t nterms = 0;
        for (int32_t q = 0; q < F.B.nlm; ++q) nterms += F.B.e[lm[q]].len;
        out->bld = F.B.nlm;
        out->nterms = nterms;
        const size_t nl = F.B.nlm > 0 ? (size_t)F.B.nlm : 1;
        out->blen = (int32_t *)malloc(nl * sizeof(int32_t));
        out->bexp = (int32_t *)malloc((size_t)(nterms ? nterms : 1) * (size_t)in->nr_vars * sizeof(int32_t));
        out->bcf = (uint32_t *)malloc((size_t)(nterms ? nterms : 1) * sizeof(uint32_t));
        if (!out->blen || !out->bexp || !out->bcf) { free(lm); gbgpu_f4_out_free(out); goto done; }
        int64_t *toff = (int64_t *)malloc(nl * sizeof(int64_t));
        if (!toff) { free(lm); gbgpu_f4_out_free(out); goto done; }
        int64_t t = 0;
        for (int32_t q = 0; q < F.B.nlm; ++q) {
            toff[q] = t;
            out->blen[q] = F.B.e[lm[q]].len;
            t += out->blen[q];
        }
        const int nv = in->nr_vars;
#pragma omp parallel for schedule(dynamic, 4)
        for (int32_t q = 0; q < F.B.nlm; ++q) {
            const poly_t *f = &F.B.e[lm[q]];
            uint32_t *oc = out->bcf + toff[q];
            int32_t *oe = out->bexp + toff[q] * nv;
            for (int32_t a = 0; a < f->len; ++a) {
                oc[a] = f->c[a];
                const ex_t *e = EV(&F.T, f->m[a]);
                for (int i = 0; i < nv; ++i) oe[(int64_t)a * nv + i] = e[i];
            }
        }
        free(toff);
        free(lm);
    }
    rc = GBGPU_OK;
done:
    /* an exponent past gb's int16 range is reported, never wrapped */
    if (rc != GBGPU_OK && F.T.erange) rc = GBGPU_ERANGE;
    for (int32_t i = 0; i < F.B.n; ++i) {
        free(F.B.e[i].m);
        free(F.B.e[i].c);
    }
    free(F.B.e); free(F.B.red); free(F.B.lm);
    free(F.P.p);
    free(F.M.stamp); free(F.M.col); free(F.M.piv); free(F.M.red); free(F.M.red_n);
    free(F.R.r); free(F.R.mon);
    free(F.W.rp); free(F.W.ci); free(F.W.rcf); free(F.W.eid); free(F.W.cp); free(F.W.cf);
    free(F.W.key); free(F.W.src);
    for (int32_t i = 0; i < F.npc; ++i) free(F.pc[i].e);
    free(F.pc);
    gbgpu_cfstore_free(F.cfs);
    free(F.cfoff);
    mt_free(&F.T);
    out->total_ms = now_ms() - t_start;
    return rc;
}

void gbgpu_f4_out_free(gbgpu_f4_out_t *out) {
    if (!out) return;
    free(out->blen);
    free(out->bexp);
    free(out->bcf);
    free(out->trace);
    out->blen = NULL;
    out->bexp = NULL;
    out->bcf = NULL;
    out->trace = NULL;
    out->bld = 0;
    out->nterms = 0;
}

/* ------------------------------------------------------- GPU back end */

int gbgpu_f4_gpu_la(void *user, const gbgpu_csr_t *m, int interreduce, gbgpu_rows_t *out) {
    gbgpu_f4_gpu_t *g = (gbgpu_f4_gpu_t *)user;
    if (!g || !g->ctx) return GBGPU_EINVAL;
    if (!interreduce) {
        const int rc = gbgpu_la_ff(g->ctx, m, 1, g->form, out);
        if (rc == GBGPU_OK) g->device_ms += gbgpu_phase_ms(g->ctx, 2);
        return rc;
    }
    gbgpu_mat *dm = NULL;
    int32_t np = 0;
    int rc = gbgpu_mat_upload(g->ctx, m, &dm);
    if (rc == GBGPU_OK) rc = gbgpu_interreduce(g->ctx, dm, g->form, &np);
    if (rc == GBGPU_OK) rc = gbgpu_rows_download(g->ctx, dm, out);
    gbgpu_mat_free(dm);
    return rc;
}

int64_t gbgpu_f4_julia(int32_t *bld, int32_t **blen, int32_t **bexp, void **bcf,
                       const int32_t *lens, const int32_t *exps, const void *cfs,
                       const int32_t field_char, const int32_t mon_order,
                       const int32_t nr_vars, const int32_t nr_gens, const int32_t ht_size,
                       const int32_t nr_threads, const int32_t max_nr_pairs,
                       const int32_t reset_ht, const int32_t la_option,
                       const int32_t reduce_gb, const int32_t pbm_file,
                       const int32_t info_level) {
    (void)ht_size; (void)nr_threads; (void)max_nr_pairs; (void)reset_ht; (void)pbm_file;
    if (!bld || !blen || !bexp || !bcf) return 0;
    const char *dev = getenv("GBGPU_DEVICE");
    gbgpu_ctx *ctx = NULL;
    if (gbgpu_init(dev ? atoi(dev) : 0, &ctx) != GBGPU_OK) {
        fprintf(stderr, "gbgpu_f4_julia: no usable GPU\n");
        return 0;
    }
    gbgpu_f4_gpu_t g = {ctx, la_option == 101 ? GBGPU_FORM_PLAIN : GBGPU_FORM_MFMA, 0.0};
    gbgpu_f4_in_t in = {nr_vars, nr_gens, lens, exps, (const int32_t *)cfs, field_char,
                        mon_order, 1, reduce_gb, info_level};
    gbgpu_f4_out_t out;
    const int rc = gbgpu_f4(&in, gbgpu_f4_gpu_la, &g, &out);
    if (rc != GBGPU_OK) {
        fprintf(stderr, "gbgpu_f4_julia: failed (%d): %s\n", rc, gbgpu_last_error(ctx));
        gbgpu_finalize(ctx);
        return 0;
    }
    gbgpu_finalize(ctx);
    free(out.trace);
    *bld = out.bld;
    *blen = out.blen;
    *bexp = out.bexp;
    *bcf = (void *)out.bcf;
    return out.nterms;
}
