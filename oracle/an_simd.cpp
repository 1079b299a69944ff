// an_simd.cpp -- all-core, vectorised CPU restatement of the hardened operators' scalar semantics.
//
// TEST INFRASTRUCTURE AND CPU BASELINE ONLY (like an_oracle.c): bench.py times it beside the GPU as
// the CPU path of the same work, and tests hold it bit-exact against the scalar oracle.  Nothing in
// the product path links, loads or falls back to it.
//
// The reference's CPU SIMD operators are SSE4.2 (lib/column_operators/AN/selectAN_SSE.tcc:46-329,
// SSEAN.hpp:58-123, encdecAN_SSE.tcc:123-291, aggregateAN_SSE.tcc:128-221) and diverge from the
// scalar semantics in places (signed detection and comparisons, error positions that never
// advance, kept corrupt rows).  This restatement follows the scalar semantics instead, so its
// outputs equal the oracle's and the GPU's bit for bit:
//   - the check is the unsigned `code * A^-1 mod 2^k > maxU` (selectAN_scalar.tcc:86);
//   - predicates compare raw codes unsigned against encoded thresholds (:89);
//   - error positions are pos * A_OID in scan order.
// Work is split into contiguous row ranges, one per OpenMP thread, and merged in range order.
// Inside a range, rows go through fixed-size blocks: the per-row decode / check / predicate loops
// have no data-dependent control flow, so the compiler vectorises them (AVX-512: 8 x 64-bit
// multiplies per vpmullq; AVX2 emulates the 64-bit product).  The file is compiled three times,
// for x86-64-v4 (AVX-512), x86-64-v3 (AVX2) and the baseline, with SIMD_LEVEL naming the entry
// points (or_star_query_simd_v4 ...); an_simd_dispatch.cpp picks one at run time.
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

extern "C" {
#include "an_oracle.h"
}

#ifndef SIMD_LEVEL
#define SIMD_LEVEL v0
#endif
#define SIMD_CAT2(a, b) a##_##b
#define SIMD_CAT(a, b) SIMD_CAT2(a, b)
#define SIMD_NAME(x) SIMD_CAT(x, SIMD_LEVEL)

namespace {

// bad[i] |= (code[i] * inv mod 2^k) > maxu; dec[i] = that product (vectorised)
void check_block(const uint64_t *code, int cnt, uint64_t inv, uint64_t m, uint64_t maxu, uint64_t *dec,
                             uint8_t *bad) {
    for (int i = 0; i < cnt; ++i) {
        const uint64_t d = (code[i] * inv) & m;
        dec[i] = d;
        bad[i] |= static_cast<uint8_t>(d > maxu);
    }
}

// hit[i] = Comb(Op1(code, th1), Op2(code, th2)) for one (op1, op2, comb) triple (vectorised per op)
void pred_block(const uint64_t *code, int cnt, int op, uint64_t th, uint8_t *hit) {
    switch (op) {
        case OR_LT: for (int i = 0; i < cnt; ++i) hit[i] = code[i] < th; break;
        case OR_LE: for (int i = 0; i < cnt; ++i) hit[i] = code[i] <= th; break;
        case OR_GT: for (int i = 0; i < cnt; ++i) hit[i] = code[i] > th; break;
        case OR_GE: for (int i = 0; i < cnt; ++i) hit[i] = code[i] >= th; break;
        case OR_EQ: for (int i = 0; i < cnt; ++i) hit[i] = code[i] == th; break;
        default: for (int i = 0; i < cnt; ++i) hit[i] = code[i] != th; break;
    }
}

constexpr int kBlock = 2048;

inline uint64_t wmask(int w) { return w >= 8 ? ~0ull : ((1ull << (8 * w)) - 1ull); }
inline uint64_t bmask(int b) { return b >= 64 ? ~0ull : ((1ull << b) - 1ull); }

template <typename T>
inline uint64_t at(const void *p, uint64_t i) {
    return static_cast<const T *>(p)[i];
}

inline uint64_t get(const or_col &c, uint64_t i) {
    switch (c.width) {
        case 1: return at<uint8_t>(c.data, i);
        case 2: return at<uint16_t>(c.data, i);
        case 4: return at<uint32_t>(c.data, i);
        default: return at<uint64_t>(c.data, i);
    }
}

inline bool cmp(int op, uint64_t t, uint64_t th) {
    switch (op) {
        case OR_LT: return t < th;
        case OR_LE: return t <= th;
        case OR_GT: return t > th;
        case OR_GE: return t >= th;
        case OR_EQ: return t == th;
        default: return t != th;
    }
}

// Codes of one block widened to 64 bits (the vector loops below run over these).
inline void widen(const or_col &c, uint64_t r0, int cnt, uint64_t *out) {
    switch (c.width) {
        case 1: for (int i = 0; i < cnt; ++i) out[i] = at<uint8_t>(c.data, r0 + i); break;
        case 2: for (int i = 0; i < cnt; ++i) out[i] = at<uint16_t>(c.data, r0 + i); break;
        case 4: for (int i = 0; i < cnt; ++i) out[i] = at<uint32_t>(c.data, r0 + i); break;
        default: memcpy(out, static_cast<const uint64_t *>(c.data) + r0, sizeof(uint64_t) * cnt); break;
    }
}

void pred_mask(const uint64_t *code, int cnt, int op1, uint64_t th1, int op2, uint64_t th2, int comb, uint8_t *hit,
               uint8_t *tmp) {
    pred_block(code, cnt, op1, th1, hit);
    if (comb == OR_COMB_NONE) return;
    pred_block(code, cnt, op2, th2, tmp);
    if (comb == OR_COMB_AND)
        for (int i = 0; i < cnt; ++i) hit[i] &= tmp[i];
    else
        for (int i = 0; i < cnt; ++i) hit[i] |= tmp[i];
}

void push(or_vec *v, uint64_t x) {
    if (v->n == v->cap) {
        const uint64_t nc = v->cap ? 2 * v->cap : 64;
        uint64_t *nv = static_cast<uint64_t *>(realloc(v->v, nc * sizeof(uint64_t)));
        if (nv == nullptr) abort();
        v->v = nv;
        v->cap = nc;
    }
    v->v[v->n++] = x;
}

void append(or_vec *dst, const std::vector<uint64_t> &src) {
    for (uint64_t x : src) push(dst, x);
}

int threads_or_max(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

}  // namespace

extern "C" {

// selectAN (1 or 2 predicates) with the fused check-and-decode sum, on `threads` threads
// (0: all).  Same outputs as or_select_an.
uint64_t SIMD_NAME(or_select_an_simd)(const or_col *col, int op1, uint64_t th1, int op2, uint64_t th2, int comb,
                           uint64_t a_head, uint64_t a_oid, or_vec *oids, or_vec *err, int threads) {
    const uint64_t n = col->n, m = wmask(col->width), maxu = col->max_u & m;
    th1 &= m;
    th2 &= m;
    const int nt = threads_or_max(threads);
    std::vector<std::vector<uint64_t>> po(nt), pe(nt);
    std::vector<uint64_t> ps(nt, 0);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const uint64_t lo = n * t / T, hi = n * (t + 1) / T;
        uint64_t code[kBlock], dec[kBlock];
        uint8_t bad[kBlock], hit[kBlock], tmp[kBlock];
        uint64_t sum = 0;
        for (uint64_t r0 = lo; r0 < hi; r0 += kBlock) {
            const int cnt = static_cast<int>(std::min<uint64_t>(kBlock, hi - r0));
            widen(*col, r0, cnt, code);
            memset(bad, 0, cnt);
            check_block(code, cnt, col->a_inv, m, maxu, dec, bad);
            pred_mask(code, cnt, op1, th1, op2, th2, comb, hit, tmp);
            for (int i = 0; i < cnt; ++i) {
                const uint64_t pos = r0 + i;
                if (bad[i]) pe[t].push_back(pos * a_oid);
                if (hit[i]) {
                    po[t].push_back(static_cast<uint64_t>(static_cast<uint32_t>(col->seqbase + pos)) * a_head);
                    if (!bad[i]) sum += dec[i];
                }
            }
        }
        ps[t] = sum;
    }
    uint64_t total = 0;
    for (int t = 0; t < nt; ++t) {
        if (oids) append(oids, po[t]);
        if (err) append(err, pe[t]);
        total += ps[t];
    }
    return total;
}

// The fused star query of or_star_query (same arguments, outputs and statuses), on `threads`
// threads (0: all).  Dimensions become dense per-key entry tables over the key span; groups a dense
// table over the product of the attribute spans (status 6 when that exceeds 2^22 entries).
int SIMD_NAME(or_star_query_simd)(const or_sq_args *a, or_sq_out *o, int threads) {
    memset(o, 0, sizeof *o);
    const int nd = a->n_dims;
    // ---- dimensions: key span, validity, uniqueness; entry per key: 0 none, idx + 1, 0xFFFFFFFF
    // (selected row with a corrupt attribute); row per key
    struct DimT {
        uint64_t kmin = 0, span = 0;
        std::vector<uint32_t> entry, row;
        uint64_t amin = 0, aspan = 1, stride = 0;
        bool has_attr = false;
    };
    std::vector<DimT> D(nd);
    uint64_t groups = 1;
    for (int d = 0; d < nd; ++d) {
        const or_sq_dim &dm = a->dims[d];
        DimT &t = D[d];
        const uint64_t nk = dm.key.n, km = bmask(dm.key.unenc_bits);
        uint64_t lo = ~0ull, hi = 0;
        for (uint64_t j = 0; j < nk; ++j) {
            const uint64_t dec = (get(dm.key, j) * dm.key.a_inv) & wmask(dm.key.width);
            if (dec > (dm.key.max_u & wmask(dm.key.width))) {
                o->status = 2;
                return 2;
            }
            const uint64_t k = dec & km;
            lo = std::min(lo, k);
            hi = std::max(hi, k);
        }
        t.kmin = nk ? lo : 0;
        t.span = nk ? hi - lo + 1 : 0;
        if (t.span > (1ull << 32)) {
            o->status = 5;
            return 5;
        }
        t.row.assign(t.span, 0xFFFFFFFFu);
        for (uint64_t j = 0; j < nk; ++j) {
            const uint64_t k = ((get(dm.key, j) * dm.key.a_inv) & wmask(dm.key.width) & km) - t.kmin;
            if (t.row[k] != 0xFFFFFFFFu) {
                o->status = 3;
                return 3;
            }
            t.row[k] = static_cast<uint32_t>(j);
        }
        const uint64_t nsel = dm.sel ? dm.n_sel : nk;
        t.has_attr = dm.attr.n > 0;
        const uint64_t am = wmask(dm.attr.width), amaxu = dm.attr.max_u & am, abits = bmask(dm.attr.unenc_bits);
        uint64_t alo = ~0ull, ahi = 0;
        for (uint64_t j = 0; j < nsel; ++j) {
            const uint64_t r = dm.sel ? dm.sel[j] * dm.sel_a_inv : j;
            if (r >= nk) {
                o->status = 1;
                return 1;
            }
            if (t.has_attr) {
                const uint64_t v = (get(dm.attr, r) * dm.attr.a_inv) & am;
                if (v <= amaxu) {
                    alo = std::min(alo, v & abits);
                    ahi = std::max(ahi, v & abits);
                }
            }
        }
        if (t.has_attr && alo != ~0ull) {
            t.amin = alo;
            t.aspan = ahi - alo + 1;
        }
        t.entry.assign(t.span, 0u);
        for (uint64_t j = 0; j < nsel; ++j) {
            const uint64_t r = dm.sel ? dm.sel[j] * dm.sel_a_inv : j;
            const uint64_t k = ((get(dm.key, r) * dm.key.a_inv) & wmask(dm.key.width) & km) - t.kmin;
            uint32_t e = 1;
            if (t.has_attr) {
                const uint64_t v = (get(dm.attr, r) * dm.attr.a_inv) & am;
                e = v > amaxu ? 0xFFFFFFFFu : static_cast<uint32_t>((v & abits) - t.amin + 1);
            }
            t.entry[k] = e;
        }
        if (t.has_attr) {
            t.stride = groups;
            groups *= t.aspan;
            if (groups > (1ull << 22)) {
                o->status = 6;
                return 6;
            }
        }
    }
    // ---- fact scan
    const uint64_t n = a->n_fact, aoe = a->a_oid_err;
    const int nt = threads_or_max(threads);
    struct Part {
        std::vector<uint64_t> sum;
        std::vector<uint64_t> first;
        std::vector<uint64_t> ev[OR_SQ_KINDS];
        uint64_t kept = 0;
    };
    std::vector<Part> P(nt);
    for (Part &pt : P) {
        pt.sum.assign(groups, 0);
        pt.first.assign(groups, ~0ull);
    }
    const int np = a->n_preds, nm = a->n_meas;
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        Part &pt = P[t];
        const uint64_t lo = n * t / T, hi = n * (t + 1) / T;
        static thread_local uint64_t code[kBlock], dec[4][kBlock], tmpd[kBlock];
        static thread_local uint8_t bad[kBlock], alive[kBlock], hit[kBlock], tmp[kBlock];
        for (uint64_t r0 = lo; r0 < hi; r0 += kBlock) {
            const int cnt = static_cast<int>(std::min<uint64_t>(kBlock, hi - r0));
            memset(bad, 0, cnt);
            memset(alive, 1, cnt);
            for (int p = 0; p < np; ++p) {
                const or_sq_pred &pr = a->preds[p];
                const uint64_t m = wmask(pr.col.width);
                widen(pr.col, r0, cnt, code);
                check_block(code, cnt, pr.col.a_inv, m, pr.col.max_u & m, tmpd, bad);
                pred_mask(code, cnt, pr.op1, pr.th1 & m, pr.op2, pr.th2 & m, pr.comb, hit, tmp);
                for (int i = 0; i < cnt; ++i) alive[i] &= hit[i];
            }
            for (int d = 0; d < nd; ++d) {
                const or_col &fk = a->dims[d].fk;
                widen(fk, r0, cnt, code);
                check_block(code, cnt, fk.a_inv, wmask(fk.width), fk.max_u & wmask(fk.width), dec[d], bad);
            }
            // fast path: rows without any corrupt code
            for (int i = 0; i < cnt; ++i) {
                if (!alive[i] || bad[i]) continue;
                const uint64_t row = r0 + i;
                bool ok = true, err = false;
                uint64_t g = 0;
                for (int d = 0; d < nd && ok; ++d) {
                    const uint64_t k = (dec[d][i] & bmask(a->dims[d].fk.unenc_bits)) - D[d].kmin;
                    const uint32_t e = k < D[d].span ? D[d].entry[k] : 0u;
                    if (e == 0u) ok = false;
                    else if (D[d].has_attr) {
                        if (e == 0xFFFFFFFFu) err = true;
                        else g += (e - 1) * D[d].stride;
                    }
                }
                if (!ok) continue;
                uint64_t dm[2] = {0, 0};
                for (int m = 0; m < nm; ++m) {
                    const or_col &mc = a->meas[m];
                    dm[m] = (get(mc, row) * mc.a_inv) & wmask(mc.width);
                    if (dm[m] > (mc.max_u & wmask(mc.width))) err = true;
                }
                if (err) {
                    bad[i] = 1;  // handled by the exact pass below
                    continue;
                }
                const uint64_t v = a->meas_kind == OR_SQ_SUB ? dm[0] - dm[1]
                                   : a->meas_kind == OR_SQ_MULSUM ? dm[0] * dm[1] : dm[0];
                pt.sum[g] += v;
                if (pt.first[g] == ~0ull) pt.first[g] = row;
                ++pt.kept;
            }
            // exact pass, in row order, over the rows that hold a corrupt code
            for (int i = 0; i < cnt; ++i) {
                if (!bad[i]) continue;
                const uint64_t row = r0 + i;
                bool live = true;
                for (int p = 0; p < np; ++p) {
                    const or_sq_pred &pr = a->preds[p];
                    const uint64_t m = wmask(pr.col.width), c = get(pr.col, row);
                    if (((c * pr.col.a_inv) & m) > (pr.col.max_u & m)) pt.ev[OR_SQ_E_PRED + p].push_back(row * aoe);
                    bool h;
                    if (pr.comb == OR_COMB_NONE) h = cmp(pr.op1, c, pr.th1 & m);
                    else if (pr.comb == OR_COMB_AND) h = cmp(pr.op1, c, pr.th1 & m) && cmp(pr.op2, c, pr.th2 & m);
                    else h = cmp(pr.op1, c, pr.th1 & m) || cmp(pr.op2, c, pr.th2 & m);
                    if (!h) live = false;
                }
                if (!live) continue;
                bool dropped = false;
                uint64_t dm[2] = {0, 0};
                for (int m = 0; m < nm; ++m) {
                    if (!((a->meas_pre_mask >> m) & 1)) continue;
                    const or_col &mc = a->meas[m];
                    dm[m] = (get(mc, row) * mc.a_inv) & wmask(mc.width);
                    if (dm[m] > (mc.max_u & wmask(mc.width))) {
                        pt.ev[OR_SQ_E_MEAS + m].push_back(row * aoe);
                        dropped = true;
                    }
                }
                uint32_t ent[4] = {0, 0, 0, 0};
                uint64_t kk[4] = {0, 0, 0, 0};
                bool fkbad[4] = {false, false, false, false};
                for (int d = 0; d < nd && live; ++d) {
                    const or_sq_dim &dmm = a->dims[d];
                    if (dec[d][i] > (dmm.fk.max_u & wmask(dmm.fk.width))) {
                        fkbad[d] = true;
                        pt.ev[OR_SQ_E_FK + d].push_back(row * aoe);
                        if (!(dmm.flags & OR_SQ_PROBE)) {
                            live = false;
                            break;
                        }
                    }
                    kk[d] = (dec[d][i] & bmask(dmm.fk.unenc_bits)) - D[d].kmin;
                    ent[d] = kk[d] < D[d].span ? D[d].entry[kk[d]] : 0u;
                    if (ent[d] == 0u) live = false;
                }
                if (!live) continue;
                for (int m = 0; m < nm; ++m) {
                    if ((a->meas_pre_mask >> m) & 1) continue;
                    const or_col &mc = a->meas[m];
                    dm[m] = (get(mc, row) * mc.a_inv) & wmask(mc.width);
                    if (dm[m] > (mc.max_u & wmask(mc.width))) {
                        pt.ev[OR_SQ_E_MEAS + m].push_back(row * aoe);
                        dropped = true;
                    }
                }
                for (int d = 0; d < nd; ++d) {
                    if ((a->dims[d].flags & OR_SQ_REFETCH) && fkbad[d]) {
                        pt.ev[OR_SQ_E_REFETCH + d].push_back(row * aoe);
                        dropped = true;
                    }
                }
                uint64_t g = 0;
                for (int d = 0; d < nd; ++d) {
                    if (!D[d].has_attr) continue;
                    if (ent[d] == 0xFFFFFFFFu) {
                        pt.ev[OR_SQ_E_ATTR + d].push_back(static_cast<uint64_t>(D[d].row[kk[d]]) * aoe);
                        dropped = true;
                    } else {
                        g += (ent[d] - 1) * D[d].stride;
                    }
                }
                if (dropped) continue;
                const uint64_t v = a->meas_kind == OR_SQ_SUB ? dm[0] - dm[1]
                                   : a->meas_kind == OR_SQ_MULSUM ? dm[0] * dm[1] : dm[0];
                pt.sum[g] += v;
                if (pt.first[g] == ~0ull || row < pt.first[g]) pt.first[g] = row;
                ++pt.kept;
            }
        }
    }
    // ---- merge: threads own ascending row ranges
    std::vector<uint64_t> sum(groups, 0), first(groups, ~0ull);
    for (int t = 0; t < nt; ++t) {
        o->kept += P[t].kept;
        for (uint64_t g = 0; g < groups; ++g) {
            sum[g] += P[t].sum[g];
            first[g] = std::min(first[g], P[t].first[g]);
        }
        for (int k = 0; k < OR_SQ_KINDS; ++k) append(&o->err[k], P[t].ev[k]);
    }
    std::vector<uint64_t> order;
    for (uint64_t g = 0; g < groups; ++g)
        if (first[g] != ~0ull) order.push_back(g);
    std::sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return first[x] < first[y]; });
    for (uint64_t g : order) {
        push(&o->sums, sum[g] * a->a_result);
        push(&o->first_row, first[g]);
        for (int d = 0; d < nd; ++d) {
            if (!D[d].has_attr) continue;
            const uint64_t idx = (g / D[d].stride) % D[d].aspan;
            push(&o->keys[d], ((D[d].amin + idx) * a->dims[d].attr.a) & wmask(a->dims[d].attr.width));
        }
    }
    return 0;
}

}  // extern "C"
