/*
 * an_oracle.c -- CPU restatement of AHEAD's AN-hardened operators (scalar semantics).
 *
 * TEST INFRASTRUCTURE ONLY (see an_oracle.h).  Straight-line scalar loops that follow the
 * reference's per-tuple loops one to one, so that a reader can hold each function next to the
 * cited reference lines.  Speed is not a goal; the loops are single-threaded and, in the
 * "ref-scalar" build, compiled with -fno-tree-vectorize like tools/CMakeLists.txt:112.
 */
#include "an_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------------------------- */
/* small helpers                                                                                  */
/* ---------------------------------------------------------------------------------------------- */

static void vec_push(or_vec *v, uint64_t x) {
    if (v == NULL) return;
    if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 64;
        uint64_t *nv = (uint64_t *)realloc(v->v, nc * sizeof(uint64_t));
        if (nv == NULL) abort();
        v->v = nv;
        v->cap = nc;
    }
    v->v[v->n++] = x;
}

void or_vec_free(or_vec *vec) {
    if (vec == NULL) return;
    free(vec->v);
    vec->v = NULL;
    vec->n = vec->cap = 0;
}

uint64_t or_mask(int width_bytes) {
    return width_bytes >= 8 ? ~0ull : ((1ull << (8 * width_bytes)) - 1ull);
}

static uint64_t mask_bits(int bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull); }

uint64_t or_inverse(uint64_t a, int bits) {
    if ((a & 1ull) == 0) return 0;
    /* x0 = a is correct to 3 bits (a*a == 1 mod 8 for odd a); every Newton step doubles it. */
    uint64_t x = a;
    for (int i = 0; i < 6; ++i) x *= 2ull - a * x;
    return x & mask_bits(bits);
}

uint64_t or_get(const or_col *c, uint64_t i) {
    if (c->data == NULL) return (uint64_t)(uint32_t)(c->seqbase + i);
    switch (c->width) {
        case 1: return ((const uint8_t *)c->data)[i];
        case 2: return ((const uint16_t *)c->data)[i];
        case 4: return ((const uint32_t *)c->data)[i];
        default: return ((const uint64_t *)c->data)[i];
    }
}

/* ANhelper0<...>::mulIfEncoded (lib/column_operators/AN/ANhelper.tcc:57-67, 93-105):
 * static_cast<T>(code * A^-1) for encoded columns, the value itself otherwise. */
uint64_t or_dec(const or_col *c, uint64_t i) {
    uint64_t x = or_get(c, i);
    if (!c->encoded) return x;
    return (x * c->a_inv) & or_mask(c->width);
}

int or_is_corrupt(const or_col *c, uint64_t i) {
    return c->encoded && or_dec(c, i) > (c->max_u & or_mask(c->width));
}

/* ---------------------------------------------------------------------------------------------- */
/* hash map: uint64 key -> uint64 value, open addressing, linear probing                          */
/* ---------------------------------------------------------------------------------------------- */

typedef struct {
    uint64_t *keys;
    uint64_t *vals;
    uint8_t *used;
    uint64_t cap; /* power of two */
    uint64_t n;
} hmap;

static uint64_t hmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
}

static void hmap_init(hmap *m, uint64_t expect) {
    uint64_t cap = 16;
    while (cap < 2 * expect + 16) cap <<= 1;
    m->keys = (uint64_t *)calloc(cap, sizeof(uint64_t));
    m->vals = (uint64_t *)calloc(cap, sizeof(uint64_t));
    m->used = (uint8_t *)calloc(cap, 1);
    if (!m->keys || !m->vals || !m->used) abort();
    m->cap = cap;
    m->n = 0;
}

static void hmap_free(hmap *m) {
    free(m->keys);
    free(m->vals);
    free(m->used);
}

/* returns pointer to value slot; *found = 1 if key was present */
static uint64_t *hmap_find_or_insert(hmap *m, uint64_t key, int *found);

static void hmap_grow(hmap *m) {
    hmap nm;
    hmap_init(&nm, m->cap);
    for (uint64_t i = 0; i < m->cap; ++i) {
        if (m->used[i]) {
            int f;
            *hmap_find_or_insert(&nm, m->keys[i], &f) = m->vals[i];
        }
    }
    hmap_free(m);
    *m = nm;
}

static uint64_t *hmap_find_or_insert(hmap *m, uint64_t key, int *found) {
    if (2 * (m->n + 1) > m->cap) hmap_grow(m);
    uint64_t s = hmix(key) & (m->cap - 1);
    while (m->used[s]) {
        if (m->keys[s] == key) {
            *found = 1;
            return &m->vals[s];
        }
        s = (s + 1) & (m->cap - 1);
    }
    m->used[s] = 1;
    m->keys[s] = key;
    m->vals[s] = 0;
    m->n++;
    *found = 0;
    return &m->vals[s];
}

static const uint64_t *hmap_find(const hmap *m, uint64_t key) {
    uint64_t s = hmix(key) & (m->cap - 1);
    while (m->used[s]) {
        if (m->keys[s] == key) return &m->vals[s];
        s = (s + 1) & (m->cap - 1);
    }
    return NULL;
}

/* ---------------------------------------------------------------------------------------------- */
/* selectAN                                                                                        */
/* ---------------------------------------------------------------------------------------------- */

/* Op<tail_select_t>(t, th): unsigned comparison in the code width */
static int pred(int op, uint64_t t, uint64_t th) {
    switch (op) {
        case OR_LT: return t < th;
        case OR_LE: return t <= th;
        case OR_GT: return t > th;
        case OR_GE: return t >= th;
        case OR_EQ: return t == th;
        default: return t != th;
    }
}

uint64_t or_select_an(const or_col *col, int op1, uint64_t th1, int op2, uint64_t th2, int comb,
                      uint64_t a_head, uint64_t a_oid, or_vec *oids, or_vec *err) {
    const uint64_t m = or_mask(col->width);
    const uint64_t maxu = col->max_u & m; /* static_cast<tail_select_t>(AN_unencMaxU) :77 */
    th1 &= m;                             /* typename Tail::type_t threshold            */
    th2 &= m;
    uint64_t sum = 0;
    for (uint64_t pos = 0; pos < col->n; ++pos) {
        const uint64_t t = or_get(col, pos);
        const uint64_t d = (t * col->a_inv) & m;
        const int bad = d > maxu;
        if (bad) vec_push(err, pos * a_oid); /* selectAN_scalar.tcc:86-88 */
        int hit;
        if (comb == OR_COMB_NONE) {
            hit = pred(op1, t, th1); /* :89 */
        } else if (comb == OR_COMB_AND) {
            hit = pred(op1, t, th1) && pred(op2, t, th2); /* :175, ahead::and_is */
        } else {
            hit = pred(op1, t, th1) || pred(op2, t, th2); /* ahead::or_is */
        }
        if (hit) {
            /* append_head(iter->head() * AHead): head() is the oid_t (uint32) virtual head */
            vec_push(oids, (uint64_t)(uint32_t)(col->seqbase + pos) * a_head);
            if (!bad) sum += d;
        }
    }
    return sum;
}

/* ---------------------------------------------------------------------------------------------- */
/* encode / check / decode                                                                         */
/* ---------------------------------------------------------------------------------------------- */

void or_check_an(const or_col *col, uint64_t a_oid, or_vec *err) {
    for (uint64_t i = 0; i < col->n; ++i) /* encdecAN_scalar.tcc:87-91 */
        if (or_is_corrupt(col, i)) vec_push(err, i * a_oid);
}

void or_decode_an(const or_col *col, uint64_t a_oid, or_vec *dec, or_vec *err) {
    const uint64_t um = mask_bits(col->unenc_bits);
    for (uint64_t pos = 0; pos < col->n; ++pos) { /* encdecAN_scalar.tcc:117-123 */
        const uint64_t t = or_dec(col, pos);
        vec_push(dec, t & um);
        if (t > (col->max_u & or_mask(col->width))) vec_push(err, pos * a_oid);
    }
}

void or_check_and_decode_an(const or_col *col, uint64_t a_oid, or_vec *dec, or_vec *err) {
    const uint64_t um = mask_bits(col->unenc_bits);
    for (uint64_t i = 0; i < col->n; ++i) { /* encdecAN_scalar.tcc:200-207 */
        const uint64_t t = or_dec(col, i);
        if (t > (col->max_u & or_mask(col->width))) {
            vec_push(err, i * a_oid);
        } else {
            vec_push(dec, t & um);
        }
    }
}

void or_check_and_decode_an2(const or_col *head, const or_col *tail, uint64_t a_oid,
                             or_vec *dec_h, or_vec *dec_t, or_vec *err_h, or_vec *err_t) {
    const uint64_t hm = mask_bits(head->unenc_bits), tm = mask_bits(tail->unenc_bits);
    for (uint64_t pos = 0; pos < tail->n; ++pos) { /* encdecAN_scalar.tcc:158-172 */
        const uint64_t h = or_dec(head, pos), t = or_dec(tail, pos);
        const int hok = !head->encoded || h <= (head->max_u & or_mask(head->width));
        const int tok = !tail->encoded || t <= (tail->max_u & or_mask(tail->width));
        if (!hok) vec_push(err_h, pos * a_oid);
        if (!tok) vec_push(err_t, pos * a_oid);
        if (hok && tok) {
            vec_push(dec_h, h & hm);
            vec_push(dec_t, t & tm);
        }
    }
}

void or_encode_an(const or_col *col, uint64_t a, int out_width, or_vec *out) {
    const uint64_t m = or_mask(out_width);
    for (uint64_t i = 0; i < col->n; ++i) /* encdecAN_scalar.tcc:62-64 */
        vec_push(out, (or_get(col, i) * a) & m);
}

/* ---------------------------------------------------------------------------------------------- */
/* fetchjoinAN / matchjoinAN positional                                                            */
/* ---------------------------------------------------------------------------------------------- */

/* width of TypeMap<T>::v2_encoded_t::type_t for a column of the given kind */
static int enc_width(const or_col *c) {
    if (c->encoded) return c->width;
    return c->width >= 4 ? 8 : 2 * c->width;
}

void or_fetchjoin_an(const or_col *oids, const or_col *col, int reenc, uint64_t a_reenc,
                     uint64_t a_oid, or_vec *out, or_vec *err1, or_vec *err2) {
    const int ow = enc_width(col);
    const uint64_t om = or_mask(ow);
    const uint64_t at1inv = oids->a_inv; /* resoid_t, 64-bit */
    uint64_t f = 1;
    if (reenc) { /* matchjoinAN.tcc:323-327 */
        f = col->encoded ? (a_reenc * or_inverse(col->a, 8 * ow)) & om : a_reenc & om;
    }
    for (uint64_t i = 0; i < oids->n; ++i) { /* matchjoinAN.tcc:335-350 */
        const uint64_t t1 = or_get(oids, i) * at1inv;
        if (t1 >= col->n) {
            vec_push(err1, i * a_oid);
        } else {
            const uint64_t t2 = or_get(col, t1) & om;
            if (col->encoded && ((t2 * col->a_inv) & om) > (col->max_u & om)) {
                vec_push(err2, t1 * a_oid);
            } else {
                vec_push(out, reenc ? (t2 * f) & om : t2);
            }
        }
    }
}

void or_matchjoin_pos_an(const or_col *h1, const or_col *t1, const or_col *t2, int reenc_h,
                         uint64_t f_h1, int wout_h, int reenc_t, uint64_t f_t2, int wout_t,
                         uint64_t a_oid, or_vec *out_h, or_vec *out_t, or_vec *err1,
                         or_vec *err2, or_vec *err4) {
    const uint64_t mh = or_mask(wout_h), mt = or_mask(wout_t);
    for (uint64_t pos1 = 0; pos1 < t1->n; ++pos1) { /* matchjoinAN.tcc:276-297 */
        const uint64_t h = or_get(h1, pos1);
        if (or_is_corrupt(h1, pos1)) vec_push(err1, pos1 * a_oid);
        const uint64_t pos2 = or_get(t1, pos1) * t1->a_inv;
        if (pos2 >= t2->n) {
            vec_push(err2, pos1 * a_oid);
        } else if (or_is_corrupt(t2, pos2)) {
            vec_push(err4, pos2 * a_oid);
        } else {
            const uint64_t v = or_get(t2, pos2);
            vec_push(out_h, reenc_h ? (h * f_h1) & mh : h);
            vec_push(out_t, reenc_t ? (v * f_t2) & mt : v);
        }
    }
}

/* The reference's loop nest, restated: an outer loop over (pos1, pos2) while both lists have
 * rows; inside it, a skip loop over list 1 while t1 < h2 and a skip loop over list 2 while
 * t1 > h2 (matchjoinAN.tcc:95-134).  t1 / h2 are decoded in their own widths and compared as
 * unsigned integers. */
void or_matchjoin_merge_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                           int reenc_h, uint64_t f_h1, int wout_h, int reenc_t, uint64_t f_t2,
                           int wout_t, uint64_t a_oid, or_vec *out_h, or_vec *out_t,
                           or_vec *err1, or_vec *err2, or_vec *err3, or_vec *err4) {
    const uint64_t n1 = t1->n, n2 = h2->n;
    const uint64_t mh = or_mask(wout_h), mt = or_mask(wout_t);
    uint64_t p1 = 0, p2 = 0;
    while (p1 < n1 && p2 < n2) {
        uint64_t hv = or_dec(h2, p2);
        if (or_is_corrupt(h2, p2)) vec_push(err3, p2 * a_oid);
        if (or_is_corrupt(t2, p2)) vec_push(err4, p2 * a_oid);
        uint64_t tv = or_dec(t1, p1);
        int valid;
        for (; (valid = p1 < n1) && tv < hv; ++p1, tv = p1 < n1 ? or_dec(t1, p1) : 0) {
            if (or_is_corrupt(h1, p1)) vec_push(err1, p1 * a_oid);
            if (or_is_corrupt(t1, p1)) vec_push(err2, p1 * a_oid);
        }
        if (!valid) continue;
        for (; (valid = p2 < n2) && tv > hv; ++p2, hv = p2 < n2 ? or_dec(h2, p2) : 0) {
            if (or_is_corrupt(h2, p2)) vec_push(err3, p2 * a_oid);
            if (or_is_corrupt(t2, p2)) vec_push(err4, p2 * a_oid);
        }
        if (valid && tv == hv) {
            const uint64_t h = or_get(h1, p1), t = or_get(t2, p2);
            vec_push(out_h, reenc_h && h1->encoded ? (h * f_h1) & mh : h);
            vec_push(out_t, reenc_t && t2->encoded ? (t * f_t2) & mt : t);
            ++p1;
            ++p2;
        }
    }
}

void or_matchjoin_seq_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                         int reenc_h, uint64_t f_h1, int wout_h, int reenc_t, uint64_t f_t2, int wout_t,
                         uint64_t a_oid, or_vec *out_h, or_vec *out_t, or_vec *err1, or_vec *err4) {
    const uint64_t n1 = t1->n, n2 = t2->n;
    const uint64_t mh = or_mask(wout_h), mt = or_mask(wout_t);
    uint64_t p1 = 0, p2 = 0;
    if (n1 == 0 || n2 == 0) return; /* matchjoinAN.tcc:187 */
    const uint32_t sb1 = (uint32_t)t1->seqbase, sb2 = (uint32_t)h2->seqbase;
    if (sb1 < sb2) { /* :188-194: skip list-1 rows whose virtual OID is below list 2's first */
        for (; p1 < n1 && p2 < n2 && (uint32_t)(sb1 + p1) < sb2; ++p1)
            if (or_is_corrupt(h1, p1)) vec_push(err1, p1 * a_oid);
    } else if (sb1 > sb2) { /* :195-201 */
        for (; p1 < n1 && p2 < n2 && (uint32_t)(sb2 + p2) < sb1; ++p2)
            if (or_is_corrupt(t2, p2)) vec_push(err4, p2 * a_oid);
    }
    for (; p1 < n1 && p2 < n2; ++p1, ++p2) { /* :203-215 */
        if (or_is_corrupt(h1, p1)) vec_push(err1, p1 * a_oid);
        if (or_is_corrupt(t2, p2)) vec_push(err4, p2 * a_oid);
        const uint64_t h = or_get(h1, p1), t = or_get(t2, p2);
        vec_push(out_h, reenc_h && h1->encoded ? (h * f_h1) & mh : h);
        vec_push(out_t, reenc_t && t2->encoded ? (t * f_t2) & mt : t);
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* hashjoinAN (unencoded hash map)                                                                 */
/* ---------------------------------------------------------------------------------------------- */

/* build side: key -> singly linked list of payload rows in insertion order */
typedef struct {
    hmap map;        /* key -> (list index + 1) */
    uint64_t *first; /* per list: first node */
    uint64_t *last;  /* per list: last node  */
    uint64_t *next;  /* per node: next node (UINT64_MAX ends) */
    uint64_t lists;
} hlists;

static void hlists_build(hlists *hl, const or_col *key_col, int key_bits) {
    const uint64_t n = key_col->n;
    hmap_init(&hl->map, n);
    hl->first = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    hl->last = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    hl->next = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    hl->lists = 0;
    const uint64_t km = mask_bits(key_bits);
    for (uint64_t j = 0; j < n; ++j) {
        const uint64_t key = or_dec(key_col, j) & km; /* static_cast<h2unenc_t>(h2) */
        int found;
        uint64_t *slot = hmap_find_or_insert(&hl->map, key, &found);
        hl->next[j] = UINT64_MAX;
        if (!found) {
            *slot = hl->lists;
            hl->first[hl->lists] = j;
            hl->last[hl->lists] = j;
            hl->lists++;
        } else {
            hl->next[hl->last[*slot]] = j;
            hl->last[*slot] = j;
        }
    }
}

static void hlists_free(hlists *hl) {
    hmap_free(&hl->map);
    free(hl->first);
    free(hl->last);
    free(hl->next);
}

static int imax(int a, int b) { return a > b ? a : b; }

void or_hashjoin_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                    int build_right, int reenc_h, uint64_t f_h1, int wout_h, int reenc_t,
                    uint64_t f_t2, int wout_t, uint64_t a_oid, or_vec *out_h, or_vec *out_t,
                    or_vec *e1, or_vec *e2, or_vec *e3, or_vec *e4) {
    const uint64_t mh = or_mask(wout_h), mt = or_mask(wout_t);
    if (t1->n == 0 || t2->n == 0) return; /* hashjoinAN.tcc:128 */
    if (build_right) {
        /* build over the right BAT (:177-187) */
        for (uint64_t j = 0; j < t2->n; ++j) {
            if (or_is_corrupt(h2, j)) vec_push(e3, j * a_oid);
            if (or_is_corrupt(t2, j)) vec_push(e4, j * a_oid);
        }
        hlists hl;
        hlists_build(&hl, h2, h2->unenc_bits);
        const int larger = imax(t1->unenc_bits, h2->unenc_bits);
        const uint64_t h2max = mask_bits(h2->unenc_bits);
        /* probe with the left BAT (:190-215) */
        for (uint64_t i = 0; i < t1->n; ++i) {
            const uint64_t hv = or_get(h1, i);
            if (or_is_corrupt(h1, i)) vec_push(e1, i * a_oid);
            const uint64_t tv = or_dec(t1, i);
            if (t1->encoded && tv > (t1->max_u & or_mask(t1->width))) vec_push(e2, i * a_oid);
            if ((tv & mask_bits(larger)) <= h2max) { /* :199, vacuous when larger == h2 bits */
                const uint64_t *li = hmap_find(&hl.map, tv & h2max);
                if (li != NULL) {
                    for (uint64_t j = hl.first[*li]; j != UINT64_MAX; j = hl.next[j]) {
                        const uint64_t m = or_get(t2, j);
                        vec_push(out_h, reenc_h ? (hv * f_h1) & mh : hv);
                        vec_push(out_t, reenc_t ? (m * f_t2) & mt : m);
                    }
                }
            }
        }
        hlists_free(&hl);
    } else {
        /* build over the left BAT (:131-144) */
        for (uint64_t i = 0; i < t1->n; ++i) {
            if (or_is_corrupt(h1, i)) vec_push(e1, i * a_oid);
            if (or_is_corrupt(t1, i)) vec_push(e2, i * a_oid);
        }
        hlists hl;
        hlists_build(&hl, t1, t1->unenc_bits);
        const int larger = imax(t1->unenc_bits, h2->unenc_bits);
        const uint64_t t1max = mask_bits(t1->unenc_bits);
        /* probe with the right BAT (:147-171).  The tail check of :153 is taken with the
         * width cast every other check applies (the reference omits it there). */
        for (uint64_t j = 0; j < t2->n; ++j) {
            const uint64_t hv = or_dec(h2, j);
            if (h2->encoded && hv > (h2->max_u & or_mask(h2->width))) vec_push(e3, j * a_oid);
            const uint64_t tv = or_get(t2, j);
            if (or_is_corrupt(t2, j)) vec_push(e4, j * a_oid);
            if ((hv & mask_bits(larger)) <= t1max) {
                const uint64_t *li = hmap_find(&hl.map, hv & t1max);
                if (li != NULL) {
                    for (uint64_t i = hl.first[*li]; i != UINT64_MAX; i = hl.next[i]) {
                        const uint64_t m = or_get(h1, i);
                        vec_push(out_h, reenc_h ? (m * f_h1) & mh : m);
                        vec_push(out_t, reenc_t ? (tv * f_t2) & mt : tv);
                    }
                }
            }
        }
        hlists_free(&hl);
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* groupbyAN                                                                                       */
/* ---------------------------------------------------------------------------------------------- */

void or_groupby_an(const or_col *head, const or_col *tail, uint64_t a_oid,
                   or_vec *voidtorgid, or_vec *vgidtoroid, or_vec *e_head, or_vec *e_tail) {
    hmap dict;
    hmap_init(&dict, 64);
    const uint64_t km = mask_bits(tail->unenc_bits);
    uint64_t ngroups = 0;
    for (uint64_t i = 0; i < tail->n; ++i) { /* groupbyAN.tcc:70-91 */
        if (or_is_corrupt(head, i)) vec_push(e_head, i * a_oid);
        const uint64_t t1 = or_dec(tail, i);
        if (tail->encoded && t1 > (tail->max_u & or_mask(tail->width))) vec_push(e_tail, i * a_oid);
        const uint64_t key = t1 & km; /* static_cast<tunenc_t>(t1) */
        int found;
        uint64_t *gid = hmap_find_or_insert(&dict, key, &found);
        if (found) {
            vec_push(voidtorgid, *gid * a_oid);
        } else {
            *gid = ngroups;
            vec_push(vgidtoroid, or_get(head, i) * a_oid); /* iter->head() * AOID: raw head */
            vec_push(voidtorgid, ngroups * a_oid);
            ngroups++;
        }
    }
    hmap_free(&dict);
}

int or_groupby2_an(const or_col *head, const or_col *tail, const or_col *grouping,
                   uint64_t num_groups, int larger_bits, uint64_t a_oid,
                   or_vec *voidtorgid, or_vec *vgidtoroid, or_vec *e_head, or_vec *e_tail,
                   or_vec *e_group) {
    if (tail->n != grouping->n) return -1; /* groupbyAN.tcc:105-107 */
    hmap dict;
    hmap_init(&dict, 64);
    const uint64_t lm = mask_bits(larger_bits);
    uint64_t ngroups = 0;
    for (uint64_t i = 0; i < tail->n; ++i) { /* groupbyAN.tcc:131-160 */
        const uint64_t h = or_dec(head, i);
        if (head->encoded && h > (head->max_u & or_mask(head->width))) vec_push(e_head, i * a_oid);
        const uint64_t t = or_dec(tail, i);
        if (tail->encoded && t > (tail->max_u & or_mask(tail->width))) vec_push(e_tail, i * a_oid);
        const uint64_t g = or_get(grouping, i) * grouping->a_inv;
        if (g > grouping->max_u) vec_push(e_group, i * a_oid);
        const uint64_t cur = ((t & lm) * (num_groups & lm) + (g & lm)) & lm;
        int found;
        uint64_t *rg = hmap_find_or_insert(&dict, cur, &found);
        if (found) {
            vec_push(voidtorgid, *rg); /* stored as oid_t: the dictionary value type */
        } else {
            const uint64_t new_rgid = ngroups * a_oid;
            vec_push(vgidtoroid, h * a_oid); /* decoded head * AOID */
            *rg = (uint32_t)new_rgid;        /* dense_hash_map<largerTail_t, oid_t> truncates */
            vec_push(voidtorgid, new_rgid);
            ngroups++;
        }
    }
    hmap_free(&dict);
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* aggregates and arithmetic                                                                       */
/* ---------------------------------------------------------------------------------------------- */

int or_aggregate_sum_grouped_an(const or_col *head, const or_col *tail, const or_col *grouping,
                                uint64_t num_groups, uint64_t a_result, uint64_t a_oid,
                                uint64_t *sums, or_vec *e_head, or_vec *e_tail, or_vec *e_group) {
    if (tail->n != grouping->n) return -1; /* aggregateAN.tcc:54-56 */
    const uint64_t enc = a_result * (tail->encoded ? or_inverse(tail->a, 64) : 1ull); /* :74-75 */
    for (uint64_t g = 0; g < num_groups; ++g) sums[g] = 0;
    for (uint64_t i = 0; i < tail->n; ++i) { /* :84-112 */
        if (or_is_corrupt(head, i)) vec_push(e_head, i * a_oid);
        const uint64_t t = or_get(tail, i);
        if (or_is_corrupt(tail, i)) vec_push(e_tail, i * a_oid);
        const uint64_t gid = or_get(grouping, i) * grouping->a_inv;
        if (gid > grouping->max_u || gid >= num_groups) {
            vec_push(e_group, i * a_oid);
        } else {
            sums[gid] += t * enc;
        }
    }
    return 0;
}

uint64_t or_aggregate_mul_sum_an(const or_col *t1, const or_col *t2, uint64_t init,
                                 int result_encoded, uint64_t ra, uint64_t a_oid,
                                 or_vec *e1, or_vec *e2) {
    /* aggregateAN_scalar.tcc:83-85 */
    const uint64_t enc = (result_encoded ? ra : 1ull) *
                         (t1->encoded ? or_inverse(t1->a, 64) : 1ull) *
                         (t2->encoded ? or_inverse(t2->a, 64) : 1ull);
    uint64_t total = init;
    const uint64_t n = t1->n < t2->n ? t1->n : t2->n;
    for (uint64_t i = 0; i < n; ++i) { /* :91-120 */
        if (or_is_corrupt(t1, i)) vec_push(e1, i * a_oid);
        if (or_is_corrupt(t2, i)) vec_push(e2, i * a_oid);
        total += or_get(t1, i) * or_get(t2, i);
    }
    return total * enc; /* :121 */
}

int or_arithmetic_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                     int op, uint64_t a_result, int result_width, uint64_t a_oid, or_vec *out,
                     or_vec *eh1, or_vec *et1, or_vec *eh2, or_vec *et2) {
    if (t1->n != t2->n) return -1; /* arithmeticAN_scalar.tcc:70-72 */
    const uint64_t rm = or_mask(result_width);
    for (uint64_t pos = 0; pos < t1->n; ++pos) { /* :90-108 */
        if (or_is_corrupt(h1, pos)) vec_push(eh1, pos * a_oid);
        const uint64_t d1 = or_dec(t1, pos);
        if (or_is_corrupt(t1, pos)) vec_push(et1, pos * a_oid);
        if (or_is_corrupt(h2, pos)) vec_push(eh2, pos * a_oid);
        const uint64_t d2 = or_dec(t2, pos);
        if (or_is_corrupt(t2, pos)) vec_push(et2, pos * a_oid);
        const uint64_t x = (d1 * a_result) & rm, y = (d2 * a_result) & rm;
        uint64_t r;
        switch (op) {
            case 0: r = x + y; break;
            case 1: r = x - y; break;
            case 2: r = x * y; break;
            default: r = y ? x / y : 0; break;
        }
        vec_push(out, r & rm);
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* composed star-join pipeline (one dimension, two attributes, one measure)                        */
/* ---------------------------------------------------------------------------------------------- */

static or_col col_of(const uint64_t *v, uint64_t n, int encoded, uint64_t a, int unenc_bits,
                     uint64_t max_u) {
    or_col c;
    memset(&c, 0, sizeof c);
    c.data = v;
    c.n = n;
    c.width = 8;
    c.encoded = encoded;
    c.unenc_bits = unenc_bits;
    c.a = a;
    c.a_inv = or_inverse(a, 64);
    c.max_u = max_u;
    return c;
}

int64_t or_star1_pipeline(const or_col *fact_fk, const or_col *fact_m, const or_col *dim_key,
                          const or_col *dim_a1, const or_col *dim_a2, uint64_t a_oid_join,
                          uint64_t a_oid_err, uint64_t a_oid_grp, uint64_t a_result,
                          or_vec *sums, or_vec *groid, or_vec *key1, or_vec *key2,
                          or_vec *e_fk, or_vec *e_dk, or_vec *e_a1, or_vec *e_a2, or_vec *e_m,
                          or_vec *e_grp) {
    const uint64_t resoid_maxu = 0x00FFFFFFFFFFFFFFull;
    if (fact_fk->n != fact_m->n || dim_a1->n != dim_key->n || dim_a2->n != dim_key->n) return -1;
    /* 1. hashjoinAN(fact FK, reverse(dim key)) with re-encoded heads and tails, as the
     *    re-encoding overload used at ssbm-q21_continuous.cpp:97-99.  Void sides are not
     *    encoded, so mulIfEncoded(A_R, A_inv) = A_R for both factors. */
    or_col fact_void, dim_void;
    memset(&fact_void, 0, sizeof fact_void);
    fact_void.n = fact_fk->n;
    fact_void.width = 4;
    fact_void.unenc_bits = 32;
    dim_void = fact_void;
    dim_void.n = dim_key->n;
    or_vec jh = {0}, jt = {0};
    or_hashjoin_an(&fact_void, fact_fk, dim_key, &dim_void, 1, 1, a_oid_join, 8, 1, a_oid_join, 8,
                   a_oid_err, &jh, &jt, NULL, e_fk, e_dk, NULL);
    /* 2. fetchjoinAN of both attributes and the measure over the joined OIDs (:151-173).
     *    Row-consistent policy: a row survives only if all three fetched codes are valid. */
    const uint64_t inv_join = or_inverse(a_oid_join, 64);
    or_vec a1 = {0}, a2 = {0}, mv = {0};
    for (uint64_t r = 0; r < jh.n; ++r) {
        const uint64_t d = jt.v[r] * inv_join; /* dimension row, always < |dim| */
        const uint64_t f = jh.v[r] * inv_join; /* fact row */
        int ok = 1;
        if (d >= dim_a1->n || or_is_corrupt(dim_a1, d)) {
            vec_push(e_a1, d * a_oid_err);
            ok = 0;
        }
        if (d >= dim_a2->n || or_is_corrupt(dim_a2, d)) {
            vec_push(e_a2, d * a_oid_err);
            ok = 0;
        }
        if (f >= fact_m->n || or_is_corrupt(fact_m, f)) {
            vec_push(e_m, f * a_oid_err);
            ok = 0;
        }
        if (ok) {
            vec_push(&a1, or_get(dim_a1, d));
            vec_push(&a2, or_get(dim_a2, d));
            vec_push(&mv, or_get(fact_m, f));
        }
    }
    or_vec_free(&jh);
    or_vec_free(&jt);
    /* 3. groupbyAN(attr1), groupbyAN(attr2, grouping, |groups|) (:166-168) */
    or_col c1 = col_of(a1.v, a1.n, 1, dim_a1->a, dim_a1->unenc_bits, dim_a1->max_u);
    c1.a_inv = dim_a1->a_inv;
    c1.width = 8; /* codes held in uint64 slots; the mask below restores the code width */
    or_col c2 = c1;
    c2.data = a2.v;
    c2.a = dim_a2->a;
    c2.a_inv = dim_a2->a_inv;
    c2.max_u = dim_a2->max_u;
    /* narrow codes: decode in their own width */
    uint16_t *a1w = NULL, *a2w = NULL;
    uint32_t *a1w4 = NULL, *a2w4 = NULL;
    if (dim_a1->width == 2) {
        a1w = (uint16_t *)malloc((a1.n + 1) * sizeof(uint16_t));
        a2w = (uint16_t *)malloc((a2.n + 1) * sizeof(uint16_t));
        for (uint64_t i = 0; i < a1.n; ++i) {
            a1w[i] = (uint16_t)a1.v[i];
            a2w[i] = (uint16_t)a2.v[i];
        }
        c1.data = a1w;
        c2.data = a2w;
        c1.width = c2.width = 2;
    } else if (dim_a1->width == 4) {
        a1w4 = (uint32_t *)malloc((a1.n + 1) * sizeof(uint32_t));
        a2w4 = (uint32_t *)malloc((a2.n + 1) * sizeof(uint32_t));
        for (uint64_t i = 0; i < a1.n; ++i) {
            a1w4[i] = (uint32_t)a1.v[i];
            a2w4[i] = (uint32_t)a2.v[i];
        }
        c1.data = a1w4;
        c2.data = a2w4;
        c1.width = c2.width = 4;
    }
    or_col kept_void;
    memset(&kept_void, 0, sizeof kept_void);
    kept_void.n = a1.n;
    kept_void.width = 4;
    kept_void.unenc_bits = 32;
    or_vec g1 = {0}, g1roid = {0}, g2 = {0}, g2roid = {0};
    or_groupby_an(&kept_void, &c1, a_oid_grp, &g1, &g1roid, NULL, NULL);
    or_col g1c = col_of(g1.v, g1.n, 1, a_oid_grp, 32, resoid_maxu);
    const int larger = dim_a2->width == 2 ? 32 : 64; /* restiny -> resshort, else 64-bit */
    or_groupby2_an(&kept_void, &c2, &g1c, g1roid.n, larger, a_oid_grp, &g2, &g2roid, NULL, NULL,
                   NULL);
    /* 4. aggregate_sum_groupedAN<v2_resbigint_t>(measure, grouping, |groups|) (:174) */
    or_col mc = col_of(mv.v, mv.n, 1, fact_m->a, fact_m->unenc_bits, fact_m->max_u);
    mc.a_inv = fact_m->a_inv;
    or_col g2c = col_of(g2.v, g2.n, 1, a_oid_grp, 32, resoid_maxu);
    const uint64_t ng = g2roid.n;
    uint64_t *s = (uint64_t *)calloc(ng + 1, sizeof(uint64_t));
    or_aggregate_sum_grouped_an(&kept_void, &mc, &g2c, ng, a_result, a_oid_grp, s, NULL, NULL,
                                e_grp);
    for (uint64_t g = 0; g < ng; ++g) vec_push(sums, s[g]);
    free(s);
    /* 5. group keys: fetchjoinAN(VGIDtoROID, attr) (:177-181) */
    or_col roidc = col_of(g2roid.v, g2roid.n, 1, a_oid_grp, 32, resoid_maxu);
    or_fetchjoin_an(&roidc, &c1, 0, 0, a_oid_err, key1, NULL, NULL);
    or_fetchjoin_an(&roidc, &c2, 0, 0, a_oid_err, key2, NULL, NULL);
    for (uint64_t g = 0; g < ng; ++g) vec_push(groid, g2roid.v[g]);
    or_vec_free(&g1);
    or_vec_free(&g1roid);
    or_vec_free(&g2);
    or_vec_free(&g2roid);
    or_vec_free(&a1);
    or_vec_free(&a2);
    or_vec_free(&mv);
    free(a1w);
    free(a2w);
    free(a1w4);
    free(a2w4);
    return (int64_t)ng;
}

/* ---------------------------------------------------------------------------------------------- */
/* fused star query: one SSB operator chain evaluated row by row                                    */
/* ---------------------------------------------------------------------------------------------- */

/* One selected dimension row, found by its truncated key. */
typedef struct {
    uint64_t key;
    uint64_t row;
} sq_entry;

static int sq_entry_cmp(const void *x, const void *y) {
    const sq_entry *a = (const sq_entry *)x, *b = (const sq_entry *)y;
    return a->key < b->key ? -1 : (a->key > b->key ? 1 : (a->row < b->row ? -1 : (a->row > b->row)));
}

/* The build side of the dimension's hash join (hashjoinAN.tcc:175-187): the selected rows keyed on
 * their decoded key truncated to the key's unencoded width, sorted for binary search.  The fused
 * form applies only when every key of the dimension (selected or not: the chains also join the
 * whole key column, e.g. ssbm-q42_continuous.cpp:226-231) is valid and unique. */
static int sq_build(const or_sq_dim *d, sq_entry **out, uint64_t *n_out) {
    const uint64_t nk = d->key.n;
    sq_entry *all = (sq_entry *)malloc((nk + 1) * sizeof(sq_entry));
    if (all == NULL) abort();
    for (uint64_t j = 0; j < nk; ++j) {
        if (or_is_corrupt(&d->key, j)) {
            free(all);
            return 2;
        }
        all[j].key = or_dec(&d->key, j) & mask_bits(d->key.unenc_bits);
        all[j].row = j;
    }
    qsort(all, nk, sizeof(sq_entry), sq_entry_cmp);
    for (uint64_t j = 1; j < nk; ++j) {
        if (all[j].key == all[j - 1].key) {
            free(all);
            return 3;
        }
    }
    if (d->sel == NULL) {
        *out = all;
        *n_out = nk;
        return 0;
    }
    free(all);
    sq_entry *e = (sq_entry *)malloc((d->n_sel + 1) * sizeof(sq_entry));
    if (e == NULL) abort();
    for (uint64_t j = 0; j < d->n_sel; ++j) {
        const uint64_t row = d->sel[j] * d->sel_a_inv;
        if (row >= nk) {
            free(e);
            return 1;
        }
        e[j].key = or_dec(&d->key, row) & mask_bits(d->key.unenc_bits);
        e[j].row = row;
    }
    qsort(e, d->n_sel, sizeof(sq_entry), sq_entry_cmp);
    *out = e;
    *n_out = d->n_sel;
    return 0;
}

static int64_t sq_find(const sq_entry *e, uint64_t n, uint64_t key) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (e[mid].key < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && e[lo].key == key) ? (int64_t)e[lo].row : -1;
}

static int sq_pred_pass(const or_sq_pred *p, uint64_t code) {
    const uint64_t m = or_mask(p->col.width);
    const uint64_t th1 = p->th1 & m, th2 = p->th2 & m;
    if (p->comb == OR_COMB_NONE) return pred(p->op1, code, th1);
    if (p->comb == OR_COMB_AND) return pred(p->op1, code, th1) && pred(p->op2, code, th2);
    return pred(p->op1, code, th1) || pred(p->op2, code, th2);
}

void or_sq_out_free(or_sq_out *o) {
    or_vec_free(&o->sums);
    or_vec_free(&o->first_row);
    for (int d = 0; d < 4; ++d) or_vec_free(&o->keys[d]);
    for (int k = 0; k < OR_SQ_KINDS; ++k) or_vec_free(&o->err[k]);
}

int or_star_query(const or_sq_args *a, or_sq_out *o) {
    memset(o, 0, sizeof *o);
    sq_entry *build[4] = {NULL, NULL, NULL, NULL};
    uint64_t nb[4] = {0, 0, 0, 0};
    for (int d = 0; d < a->n_dims; ++d) {
        const int st = sq_build(&a->dims[d], &build[d], &nb[d]);
        if (st != 0) {
            for (int k = 0; k < d; ++k) free(build[k]);
            o->status = st;
            return st;
        }
    }
    /* groups: attribute-value tuples in first-occurrence order (linear dictionary; small) */
    uint64_t *gkey = NULL; /* [groups][4] */
    uint64_t gcap = 0, ng = 0;
    uint64_t *gsum = NULL, *gfirst = NULL;
    const uint64_t aoe = a->a_oid_err;
    for (uint64_t i = 0; i < a->n_fact; ++i) {
        /* 1. fact predicates (selectAN over the whole column: every row checked) */
        int alive = 1;
        for (int p = 0; p < a->n_preds; ++p) {
            const or_sq_pred *pr = &a->preds[p];
            if (or_is_corrupt(&pr->col, i)) vec_push(&o->err[OR_SQ_E_PRED + p], i * aoe);
            if (!sq_pred_pass(pr, or_get(&pr->col, i))) alive = 0;
        }
        if (!alive) continue;
        int dropped = 0;
        /* 2. measures fetched right after the predicates */
        for (int m = 0; m < a->n_meas; ++m) {
            if (((a->meas_pre_mask >> m) & 1) && or_is_corrupt(&a->meas[m], i)) {
                vec_push(&o->err[OR_SQ_E_MEAS + m], i * aoe);
                dropped = 1;
            }
        }
        /* 3. the dimensions in chain order */
        int64_t drow[4] = {-1, -1, -1, -1};
        for (int d = 0; d < a->n_dims && alive; ++d) {
            const or_sq_dim *dm = &a->dims[d];
            if (or_is_corrupt(&dm->fk, i)) {
                vec_push(&o->err[OR_SQ_E_FK + d], i * aoe);
                if (!(dm->flags & OR_SQ_PROBE)) {
                    alive = 0;
                    break;
                }
            }
            const uint64_t key = or_dec(&dm->fk, i) & mask_bits(dm->fk.unenc_bits);
            drow[d] = sq_find(build[d], nb[d], key);
            if (drow[d] < 0) alive = 0;
        }
        if (!alive) continue;
        /* 4. final fetches: measures, re-fetched FKs, group attributes */
        for (int m = 0; m < a->n_meas; ++m) {
            if (!((a->meas_pre_mask >> m) & 1) && or_is_corrupt(&a->meas[m], i)) {
                vec_push(&o->err[OR_SQ_E_MEAS + m], i * aoe);
                dropped = 1;
            }
        }
        for (int d = 0; d < a->n_dims; ++d) {
            const or_sq_dim *dm = &a->dims[d];
            if ((dm->flags & OR_SQ_REFETCH) && or_is_corrupt(&dm->fk, i)) {
                vec_push(&o->err[OR_SQ_E_REFETCH + d], i * aoe);
                dropped = 1;
            }
        }
        for (int d = 0; d < a->n_dims; ++d) {
            const or_sq_dim *dm = &a->dims[d];
            if (dm->attr.n > 0 && or_is_corrupt(&dm->attr, (uint64_t)drow[d])) {
                vec_push(&o->err[OR_SQ_E_ATTR + d], (uint64_t)drow[d] * aoe);
                dropped = 1;
            }
        }
        if (dropped) continue;
        /* 5. group and measure */
        uint64_t k[4] = {0, 0, 0, 0};
        for (int d = 0; d < a->n_dims; ++d)
            if (a->dims[d].attr.n > 0) k[d] = or_dec(&a->dims[d].attr, (uint64_t)drow[d]) &
                                              mask_bits(a->dims[d].attr.unenc_bits);
        uint64_t g = 0;
        for (; g < ng; ++g)
            if (gkey[4 * g] == k[0] && gkey[4 * g + 1] == k[1] && gkey[4 * g + 2] == k[2] && gkey[4 * g + 3] == k[3])
                break;
        if (g == ng) {
            if (ng == gcap) {
                gcap = gcap ? 2 * gcap : 64;
                gkey = (uint64_t *)realloc(gkey, gcap * 4 * sizeof(uint64_t));
                gsum = (uint64_t *)realloc(gsum, gcap * sizeof(uint64_t));
                gfirst = (uint64_t *)realloc(gfirst, gcap * sizeof(uint64_t));
                if (gkey == NULL || gsum == NULL || gfirst == NULL) abort();
            }
            for (int d = 0; d < 4; ++d) gkey[4 * g + d] = k[d];
            gsum[g] = 0;
            gfirst[g] = i;
            ++ng;
            for (int d = 0; d < a->n_dims; ++d)
                if (a->dims[d].attr.n > 0) vec_push(&o->keys[d], or_get(&a->dims[d].attr, (uint64_t)drow[d]));
        }
        const uint64_t d0 = or_dec(&a->meas[0], i);
        uint64_t v = d0;
        if (a->meas_kind == OR_SQ_SUB) v = d0 - or_dec(&a->meas[1], i);
        if (a->meas_kind == OR_SQ_MULSUM) v = d0 * or_dec(&a->meas[1], i);
        gsum[g] += v;
        ++o->kept;
    }
    for (uint64_t g = 0; g < ng; ++g) {
        vec_push(&o->sums, gsum[g] * a->a_result);
        vec_push(&o->first_row, gfirst[g]);
    }
    free(gkey);
    free(gsum);
    free(gfirst);
    for (int d = 0; d < a->n_dims; ++d) free(build[d]);
    return 0;
}
