// starq_an.hip -- fused hardened star query for gfx950: one pass over the fact rows replaces the
// fact side of one SSB operator chain (tools/ssbm/ssbm-q{11..43}_continuous.cpp):
//   selectAN on fact columns (Q1.x), hashjoinAN / positional matchjoinAN + hashjoinAN per
//   dimension, fetchjoinAN of the measures and of the group attributes, groupbyAN over the
//   attributes, aggregate_sum_groupedAN / arithmeticAN(sub) / aggregate_mul_sumAN.
// The dimension selections themselves (selectAN over dimension attributes) run before, with the
// per-operator entry points; their OID lists are this entry point's input.
//
// Per fact row, in fact order (the chain's row semantics; corrupt fetched codes follow the
// row-consistent policy: reported and the row dropped from the sums, where the reference's chain
// throws, groupbyAN.tcc:105-107):
//   1. fact predicates on raw codes, every row checked (selectAN_scalar.tcc:84-90);
//   2. measures fetched right after the predicates (ssbm-q11_continuous.cpp:113), checked;
//   3. dimensions in chain order: the FK code is checked; a corrupt code ends the row unless the
//      dimension is probed straight on the fact column (PROBE: hashjoinAN probes the truncated
//      key anyway, hashjoinAN.tcc:199); the truncated key must belong to a selected row;
//   4. joined rows fetch their final measures, re-fetch PROBE keys (REFETCH) and fetch their
//      group attributes; any corrupt code drops the row;
//   5. kept rows add SUM d0 | SUB d0 - d1 | MULSUM d0 * d1 to their group.
//
// Layout and kernels
//  Per dimension, over the span [kmin, kmax] of its (decoded, truncated) keys: a packed map of
//  1/2/4/8/16-bit entries: 0 = no selected row, v = attribute index + 1 of a selected row,
//  all ones = a selected row whose group attribute is corrupt (entries of dimensions without a
//  group attribute are 1 bit); a bitmap of keys seen (duplicates make the fused form
//  inapplicable); the dimension row of every key (error positions of attribute fetches).
//   k_sq_keystats / k_sq_selstats   key range and validity; selected OIDs, attribute range
//                                   (read back by the host, which plans maps, LDS and groups)
//   k_sq_keys / k_sq_sel            key bitmap + row of key; map entries of the selected rows
//   k_sq_scan<ND,NP>                persistent, one 1024-thread workgroup per CU; the maps that fit
//                                   (most selective first) are copied to LDS, the others are read
//                                   from L2.  Each wave streams one contiguous fact range, 512
//                                   rows per step with 128-bit non-temporal loads of the FK codes;
//                                   dimensions are tested in LDS-first, most-selective-first order,
//                                   so the L2 maps and the measures are read only for rows still
//                                   alive; kept rows add into LDS group sums and first rows.  A row
//                                   with any corrupt code (rare) is re-evaluated in chain order by
//                                   the exact path, which appends its error events.
//   k_sq_final                      one workgroup: orders the groups by first row, writes sums and
//                                   keys; sorts the error events by row and writes each kind's
//                                   vector in order.
#include "host_common.hpp"

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

using namespace ahead_dev;

namespace {

constexpr int kSqBlock = 1024;
constexpr int kSqWaves = kSqBlock / 64;
constexpr int kSqSeg = 512;  // rows per wave step: 4 chunks of 128 rows, 2 rows per lane
constexpr uint32_t kSqMaxGroups = AHEAD_SQ_MAX_GROUPS;
constexpr uint32_t kSqMaxEvents = AHEAD_SQ_MAX_EVENTS;
constexpr uint32_t kSqMapWords = 28416;  // LDS words for maps (with the group tables: < 160 KiB)
constexpr uint32_t kNoLds = 0xFFFFFFFFu;
constexpr int kKinds = AHEAD_SQ_KINDS;
// flags[] words
enum { kFlagDup = 0, kFlagBadKey = 1, kFlagBadOid = 2 };
// stats[] per dimension (uint32): kmin, kmax, amin, amax, selected rows, attribute errors
constexpr int kStatWords = 8;

struct SqDimK {
    const uint64_t *fk;
    uint64_t fk_inv, fk_maxu;
    uint32_t kmask;        // truncation of the decoded key to its unencoded width
    uint32_t kmin, span;   // key range of the map
    uint32_t lb;           // log2 of the entry bits (0..4)
    uint32_t lds_off;      // word offset of the map in LDS, or kNoLds
    uint32_t map_off;      // word offset of the map in the HBM map region
    uint32_t stride;       // group stride of the attribute index (0: no attribute)
    uint32_t errv;         // entry value of a corrupt attribute (0: no attribute)
    uint32_t flags;        // AHEAD_SQ_PROBE | AHEAD_SQ_REFETCH
    uint32_t cpos;         // position in the chain
    const uint32_t *row_of;  // dimension row per key index
};

struct SqPredK {
    const void *col;
    uint32_t width;
    int op1, op2, comb;
    uint64_t inv, maxu, wmask, th1, th2;
};

struct SqMeasK {
    const void *col;
    uint32_t width;
    uint32_t pre;  // fetched right after the predicates
    uint64_t inv, maxu, wmask;
};

struct SqArgs {
    uint64_t n, rows_per_wave;
    uint32_t nd, np, nm, kind, has_pre;
    SqDimK d[AHEAD_SQ_MAX_DIMS];  // in evaluation order
    SqPredK p[AHEAD_SQ_MAX_PREDS];
    SqMeasK m[AHEAD_SQ_MAX_MEAS];
    const uint32_t *maps;  // HBM map region; its first lds_words words are copied to LDS
    uint32_t lds_words;
    uint32_t groups;
    const uint32_t *flags;
    unsigned long long *gsum;
    uint32_t *gfirst;
    unsigned long long *kept;
    uint32_t *ev_count;
    uint64_t *events;  // (row << 32) | (kind << 27) | dimension row
};

__device__ __forceinline__ bool cmp_op(int op, uint64_t t, uint64_t th) {
    switch (op) {
        case AHEAD_LT: return t < th;
        case AHEAD_LE: return t <= th;
        case AHEAD_GT: return t > th;
        case AHEAD_GE: return t >= th;
        case AHEAD_EQ: return t == th;
        default: return t != th;
    }
}

__device__ __forceinline__ bool pred_pass(const SqPredK &p, uint64_t code) {
    if (p.comb == AHEAD_COMB_NONE) return cmp_op(p.op1, code, p.th1);
    if (p.comb == AHEAD_COMB_AND) return cmp_op(p.op1, code, p.th1) && cmp_op(p.op2, code, p.th2);
    return cmp_op(p.op1, code, p.th1) || cmp_op(p.op2, code, p.th2);
}

__device__ __forceinline__ uint64_t load_code(const void *col, uint32_t width, uint64_t i) {
    switch (width) {
        case 1: return static_cast<const uint8_t *>(col)[i];
        case 2: return static_cast<const uint16_t *>(col)[i];
        case 4: return static_cast<const uint32_t *>(col)[i];
        default: return static_cast<const uint64_t *>(col)[i];
    }
}

// Two consecutive codes (rows r0, r0 + 1; r0 even) of a column of any width, one load.
__device__ __forceinline__ void load_pair(const void *col, uint32_t width, uint64_t r0, uint64_t &x0, uint64_t &x1) {
    switch (width) {
        case 1: {
            const uint16_t w = static_cast<const uint16_t *>(col)[r0 >> 1];
            x0 = w & 0xFFu;
            x1 = w >> 8;
            break;
        }
        case 2: {
            const uint32_t w = static_cast<const uint32_t *>(col)[r0 >> 1];
            x0 = w & 0xFFFFu;
            x1 = w >> 16;
            break;
        }
        case 4: {
            const uint64_t w = static_cast<const uint64_t *>(col)[r0 >> 1];
            x0 = w & 0xFFFFFFFFull;
            x1 = w >> 32;
            break;
        }
        default: {
            const u32x4 q = load16_nt(static_cast<const uint64_t *>(col) + r0);
            x0 = (static_cast<uint64_t>(q[1]) << 32) | q[0];
            x1 = (static_cast<uint64_t>(q[3]) << 32) | q[2];
            break;
        }
    }
}

__device__ __forceinline__ uint32_t map_entry(uint32_t w, uint32_t idx, uint32_t lb) {
    const uint32_t sh = 5u - lb;
    const uint32_t bits = 1u << lb;
    const uint32_t emask = bits >= 32u ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    return (w >> ((idx & ((1u << sh) - 1u)) << lb)) & emask;
}

__device__ __forceinline__ uint32_t lookup_global(const SqArgs &a, const SqDimK &d, uint32_t key) {
    const uint32_t idx = key - d.kmin;
    if (idx >= d.span) return 0u;
    return map_entry(a.maps[d.map_off + (idx >> (5u - d.lb))], idx, d.lb);
}

__device__ __forceinline__ void push_event(const SqArgs &a, uint64_t row, uint32_t kind, uint32_t drow) {
    const uint32_t i = atomicAdd(a.ev_count, 1u);
    if (i < kSqMaxEvents)
        a.events[i] = (row << 32) | (static_cast<uint64_t>(kind) << 27) | (drow & 0x7FFFFFFu);
}

__device__ __forceinline__ uint64_t measure_value(const SqArgs &a, uint64_t d0, uint64_t d1) {
    if (a.kind == AHEAD_SQ_SUB) return d0 - d1;
    if (a.kind == AHEAD_SQ_MULSUM) return d0 * d1;
    return d0;
}

// The chain's semantics for one row, with events: used for every row that holds a corrupt code.
__device__ void exact_row(const SqArgs &a, uint64_t row, unsigned long long *s_sum, uint32_t *s_first,
                          uint32_t &kept) {
    bool alive = true;
    for (uint32_t p = 0; p < a.np; ++p) {
        const SqPredK &pr = a.p[p];
        const uint64_t c = load_code(pr.col, pr.width, row);
        if (((c * pr.inv) & pr.wmask) > pr.maxu) push_event(a, row, AHEAD_SQ_E_PRED + p, 0u);
        if (!pred_pass(pr, c)) alive = false;
    }
    if (!alive) return;
    bool dropped = false;
    uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
    for (uint32_t m = 0; m < a.nm; ++m) {
        const SqMeasK &mk = a.m[m];
        if (!mk.pre) continue;
        dm[m] = (load_code(mk.col, mk.width, row) * mk.inv) & mk.wmask;
        if (dm[m] > mk.maxu) {
            push_event(a, row, AHEAD_SQ_E_MEAS + m, 0u);
            dropped = true;
        }
    }
    uint32_t ent[AHEAD_SQ_MAX_DIMS] = {0, 0, 0, 0}, keyv[AHEAD_SQ_MAX_DIMS] = {0, 0, 0, 0};
    bool fk_bad[AHEAD_SQ_MAX_DIMS] = {false, false, false, false};
    for (uint32_t cp = 0; cp < a.nd; ++cp) {
        uint32_t i = 0;
        for (uint32_t j = 0; j < a.nd; ++j)
            if (a.d[j].cpos == cp) i = j;
        const SqDimK &d = a.d[i];
        const uint64_t dec = d.fk[row] * d.fk_inv;
        if (dec > d.fk_maxu) {
            fk_bad[i] = true;
            push_event(a, row, AHEAD_SQ_E_FK + cp, 0u);
            if (!(d.flags & AHEAD_SQ_PROBE)) return;
        }
        keyv[i] = static_cast<uint32_t>(dec) & d.kmask;
        ent[i] = lookup_global(a, d, keyv[i]);
        if (ent[i] == 0u) return;
    }
    for (uint32_t m = 0; m < a.nm; ++m) {
        const SqMeasK &mk = a.m[m];
        if (mk.pre) continue;
        dm[m] = (load_code(mk.col, mk.width, row) * mk.inv) & mk.wmask;
        if (dm[m] > mk.maxu) {
            push_event(a, row, AHEAD_SQ_E_MEAS + m, 0u);
            dropped = true;
        }
    }
    for (uint32_t i = 0; i < a.nd; ++i) {
        if ((a.d[i].flags & AHEAD_SQ_REFETCH) && fk_bad[i]) {
            push_event(a, row, AHEAD_SQ_E_REFETCH + a.d[i].cpos, 0u);
            dropped = true;
        }
    }
    uint32_t g = 0;
    for (uint32_t i = 0; i < a.nd; ++i) {
        const SqDimK &d = a.d[i];
        if (d.errv == 0u) continue;
        if (ent[i] == d.errv) {
            push_event(a, row, AHEAD_SQ_E_ATTR + d.cpos, d.row_of[keyv[i] - d.kmin]);
            dropped = true;
        } else {
            g += (ent[i] - 1u) * d.stride;
        }
    }
    if (dropped || g >= a.groups) return;
    atomicAdd(&s_sum[g], static_cast<unsigned long long>(measure_value(a, dm[0], dm[1])));
    atomicMin(&s_first[g], static_cast<uint32_t>(row));
    ++kept;
}

template <int ND, int NP>
__global__ __launch_bounds__(kSqBlock) void k_sq_scan(SqArgs a) {
    __shared__ uint32_t s_map[kSqMapWords];
    __shared__ unsigned long long s_sum[kSqMaxGroups];
    __shared__ uint32_t s_first[kSqMaxGroups];
    static_assert(sizeof(s_map) + sizeof(s_sum) + sizeof(s_first) <= 160 * 1024 - 1024, "LDS budget");
    if (a.flags[kFlagDup] != 0u) return;
    for (uint32_t i = threadIdx.x; i < a.lds_words; i += kSqBlock) s_map[i] = a.maps[i];
    for (uint32_t g = threadIdx.x; g < a.groups; g += kSqBlock) {
        s_sum[g] = 0ull;
        s_first[g] = 0xFFFFFFFFu;
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * kSqWaves + wave;
    const uint64_t begin = gw * a.rows_per_wave;
    const uint64_t end = begin + a.rows_per_wave < a.n ? begin + a.rows_per_wave : a.n;
    uint32_t kept = 0;

    for (uint64_t seg = begin; seg < end; seg += kSqSeg) {
        const bool full = seg + kSqSeg <= end;
        uint64_t fk[ND][4][2];
        uint64_t pc[NP > 0 ? NP : 1][4][2];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t r0 = seg + c * 128u + lane * 2u;
            if (full) {
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const u32x4 q = load16_nt(a.d[d].fk + r0);
                    fk[d][c][0] = (static_cast<uint64_t>(q[1]) << 32) | q[0];
                    fk[d][c][1] = (static_cast<uint64_t>(q[3]) << 32) | q[2];
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) load_pair(a.p[p].col, a.p[p].width, r0, pc[p][c][0], pc[p][c][1]);
            } else {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const bool in = r0 + k < end;
#pragma unroll
                    for (int d = 0; d < ND; ++d) fk[d][c][k] = in ? a.d[d].fk[r0 + k] : 0ull;
#pragma unroll
                    for (int p = 0; p < NP; ++p) pc[p][c][k] = in ? load_code(a.p[p].col, a.p[p].width, r0 + k) : 0ull;
                }
            }
        }
        uint32_t errbits = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint64_t row = seg + c * 128u + lane * 2u + k;
                const bool live = full || row < end;
                bool err = false, alive = live;
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    const SqPredK &pr = a.p[p];
                    const uint64_t code = pc[p][c][k];
                    err |= ((code * pr.inv) & pr.wmask) > pr.maxu;
                    alive = alive && pred_pass(pr, code);
                }
                uint32_t key[ND];
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const uint64_t dec = fk[d][c][k] * a.d[d].fk_inv;
                    err |= dec > a.d[d].fk_maxu;
                    key[d] = static_cast<uint32_t>(dec) & a.d[d].kmask;
                }
                err = err && live;
                // measures fetched right after the predicates: checked at every row that passes them
                uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
                if (alive && a.has_pre) {
                    for (uint32_t m = 0; m < a.nm; ++m) {
                        const SqMeasK &mk = a.m[m];
                        if (!mk.pre) continue;
                        dm[m] = (load_code(mk.col, mk.width, row) * mk.inv) & mk.wmask;
                        err |= dm[m] > mk.maxu;
                    }
                }
                alive = alive && !err;
                uint32_t g = 0;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const SqDimK &dd = a.d[d];
                    if (alive) {
                        const uint32_t idx = key[d] - dd.kmin;
                        uint32_t e = 0u;
                        if (idx < dd.span) {
                            const uint32_t wi = idx >> (5u - dd.lb);
                            const uint32_t w = dd.lds_off != kNoLds ? s_map[dd.lds_off + wi] : a.maps[dd.map_off + wi];
                            e = map_entry(w, idx, dd.lb);
                        }
                        if (e == 0u) {
                            alive = false;
                        } else if (dd.errv != 0u) {
                            if (e == dd.errv) err = true;
                            g += (e - 1u) * dd.stride;
                        }
                    }
                }
                alive = alive && !err;
                if (alive) {
                    for (uint32_t m = 0; m < a.nm; ++m) {
                        const SqMeasK &mk = a.m[m];
                        if (mk.pre) continue;
                        dm[m] = (load_code(mk.col, mk.width, row) * mk.inv) & mk.wmask;
                        err |= dm[m] > mk.maxu;
                    }
                    if (!err && g < a.groups) {
                        atomicAdd(&s_sum[g], static_cast<unsigned long long>(measure_value(a, dm[0], dm[1])));
                        atomicMin(&s_first[g], static_cast<uint32_t>(row));
                        ++kept;
                    }
                }
                if (err) errbits |= 1u << (c * 2 + k);
            }
        }
        if (__ballot(errbits != 0u) != 0ull) {
            for (int b = 0; b < 8; ++b)
                if ((errbits >> b) & 1u)
                    exact_row(a, seg + (b >> 1) * 128u + lane * 2u + (b & 1), s_sum, s_first, kept);
        }
    }
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < a.groups; g += kSqBlock) {
        if (s_first[g] != 0xFFFFFFFFu) {
            atomicAdd(&a.gsum[g], s_sum[g]);
            atomicMin(&a.gfirst[g], s_first[g]);
        }
    }
    const uint32_t wk = wave_sum(kept);
    if (lane == 0 && wk != 0u) atomicAdd(a.kept, static_cast<unsigned long long>(wk));
}

// ---- dimension side ---------------------------------------------------------------------------

struct DimPrepArgs {
    const uint64_t *key;
    uint64_t n;
    uint64_t key_inv, key_maxu;
    uint32_t kmask;
    const uint64_t *sel;        // selected OIDs, or NULL (every row)
    const uint64_t *sel_count;  // device count of sel (NULL with sel NULL)
    uint64_t sel_inv;
    const void *attr;           // NULL: no group attribute
    uint32_t attr_width;
    uint64_t attr_inv, attr_maxu, attr_wmask;
    uint32_t amask;             // truncation of the decoded attribute to its unencoded width
    uint32_t *stats;            // kStatWords words
    uint32_t *flags;
    // fill (after planning)
    uint32_t kmin, span, lb, amin, errv;
    uint32_t *map, *seen, *row_of;
};

__device__ __forceinline__ uint64_t sel_count_of(const DimPrepArgs &a) {
    if (a.sel == nullptr) return a.n;
    const uint64_t c = *a.sel_count;
    return c < a.n ? c : a.n;  // a selection never holds more OIDs than the dimension has rows
}

__device__ __forceinline__ uint32_t block_min_u32(uint32_t v, uint32_t *s_red) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, static_cast<uint32_t>(__shfl_xor(v, d, 64)));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_red[wid] = v;
    __syncthreads();
    uint32_t r = 0xFFFFFFFFu;
    for (int w = 0; w < kWaves; ++w) r = min(r, s_red[w]);
    __syncthreads();
    return r;
}

// Key range of the valid keys; any corrupt key raises kFlagBadKey.
__global__ __launch_bounds__(kBlock) void k_sq_keystats(DimPrepArgs a) {
    __shared__ uint32_t s_red[kWaves];
    uint32_t lo = 0xFFFFFFFFu, hi_neg = 0xFFFFFFFFu;
    bool bad = false;
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < a.n;
         j += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t dec = a.key[j] * a.key_inv;
        if (dec > a.key_maxu) {
            bad = true;
            continue;
        }
        const uint32_t k = static_cast<uint32_t>(dec) & a.kmask;
        lo = min(lo, k);
        hi_neg = min(hi_neg, ~k);
    }
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&a.flags[kFlagBadKey], 1u);
    lo = block_min_u32(lo, s_red);
    hi_neg = block_min_u32(hi_neg, s_red);
    if (threadIdx.x == 0) {
        atomicMin(&a.stats[0], lo);
        atomicMin(&a.stats[1], hi_neg);  // stored complemented: the host takes ~x
    }
}

// Selected OIDs: range check, attribute range of the valid attributes, attribute errors.
__global__ __launch_bounds__(kBlock) void k_sq_selstats(DimPrepArgs a) {
    __shared__ uint32_t s_red[kWaves];
    const uint64_t cnt = sel_count_of(a);
    uint32_t lo = 0xFFFFFFFFu, hi_neg = 0xFFFFFFFFu, nerr = 0;
    bool bad = false;
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < cnt;
         j += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t row = a.sel ? a.sel[j] * a.sel_inv : j;
        if (row >= a.n) {
            bad = true;
            continue;
        }
        if (a.attr == nullptr) continue;
        const uint64_t dec = (load_code(a.attr, a.attr_width, row) * a.attr_inv) & a.attr_wmask;
        if (dec > a.attr_maxu) {
            ++nerr;
            continue;
        }
        const uint32_t v = static_cast<uint32_t>(dec) & a.amask;
        lo = min(lo, v);
        hi_neg = min(hi_neg, ~v);
    }
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&a.flags[kFlagBadOid], 1u);
    lo = block_min_u32(lo, s_red);
    hi_neg = block_min_u32(hi_neg, s_red);
    const uint32_t ne = wave_sum(nerr);
    if ((threadIdx.x & 63) == 0 && ne != 0u) atomicAdd(&a.stats[5], ne);
    if (threadIdx.x == 0) {
        atomicMin(&a.stats[2], lo);
        atomicMin(&a.stats[3], hi_neg);
        if (blockIdx.x == 0) a.stats[4] = static_cast<uint32_t>(cnt);
    }
}

// Every key: the seen bitmap (a key seen twice raises kFlagDup) and the row of the key.
__global__ __launch_bounds__(kBlock) void k_sq_keys(DimPrepArgs a) {
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < a.n;
         j += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t idx = (static_cast<uint32_t>(a.key[j] * a.key_inv) & a.kmask) - a.kmin;
        if (idx >= a.span) continue;  // cannot happen for valid keys; corrupt keys stop the plan
        const uint32_t bit = 1u << (idx & 31u);
        if (atomicOr(&a.seen[idx >> 5], bit) & bit) atomicOr(&a.flags[kFlagDup], 1u);
        a.row_of[idx] = static_cast<uint32_t>(j);
    }
}

// Map entries of the selected rows.
__global__ __launch_bounds__(kBlock) void k_sq_sel(DimPrepArgs a) {
    const uint64_t cnt = sel_count_of(a);
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < cnt;
         j += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t row = a.sel ? a.sel[j] * a.sel_inv : j;
        if (row >= a.n) continue;
        const uint32_t idx = (static_cast<uint32_t>(a.key[row] * a.key_inv) & a.kmask) - a.kmin;
        if (idx >= a.span) continue;
        uint32_t v = 1u;
        if (a.attr != nullptr) {
            const uint64_t dec = (load_code(a.attr, a.attr_width, row) * a.attr_inv) & a.attr_wmask;
            v = dec > a.attr_maxu ? a.errv : ((static_cast<uint32_t>(dec) & a.amask) - a.amin + 1u);
        }
        const uint32_t sh = 5u - a.lb;
        atomicOr(&a.map[idx >> sh], v << ((idx & ((1u << sh) - 1u)) << a.lb));
    }
}

// ---- finalize ---------------------------------------------------------------------------------

struct SqAttrK {
    uint32_t has, cpos, stride, span, amin;
    uint64_t a, wmask;
};

struct SqFinalArgs {
    uint32_t groups, nd;
    SqAttrK at[AHEAD_SQ_MAX_DIMS];  // in evaluation order
    const unsigned long long *gsum;
    const uint32_t *gfirst;
    const unsigned long long *kept;
    const uint32_t *flags;
    const uint32_t *ev_count;
    const uint64_t *events;
    uint64_t a_result, a_oid_err;
    ahead_sq_out out;
};

__device__ void bitonic_lds(unsigned long long *s, uint32_t p) {
    for (uint32_t k = 2; k <= p; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p; i += kSqBlock) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long x = s[i], y = s[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        s[i] = y;
                        s[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kSqBlock) void k_sq_final(SqFinalArgs a) {
    __shared__ unsigned long long s_key[kSqMaxEvents];
    __shared__ uint32_t s_n;
    __shared__ uint32_t s_scan[kSqWaves];
    __shared__ uint32_t s_base[kKinds];
    const ahead_sq_out &o = a.out;
    if (a.flags[kFlagDup] != 0u) {
        if (threadIdx.x == 0) {
            o.d_counts[0] = 0;
            o.d_counts[1] = 0;
            o.d_counts[2] = AHEAD_SQ_ST_DUP_KEY;
            o.d_counts[3] = 0;
            for (int k = 0; k < kKinds; ++k) o.d_counts[4 + k] = 0;
        }
        return;
    }
    // groups, ordered by first kept row
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < a.groups; g += kSqBlock) {
        const uint32_t f = a.gfirst[g];
        if (f != 0xFFFFFFFFu) s_key[atomicAdd(&s_n, 1u)] = (static_cast<unsigned long long>(f) << 32) | g;
    }
    __syncthreads();
    const uint32_t ng = s_n;
    uint32_t p = 1;
    while (p < ng) p <<= 1;
    for (uint32_t i = ng + threadIdx.x; i < p; i += kSqBlock) s_key[i] = ~0ull;
    __syncthreads();
    bitonic_lds(s_key, p);
    for (uint32_t j = threadIdx.x; j < ng; j += kSqBlock) {
        const uint32_t g = static_cast<uint32_t>(s_key[j] & 0xFFFFFFFFull);
        if (j < o.groups_cap) {
            if (o.sums != nullptr) o.sums[j] = static_cast<uint64_t>(a.gsum[g]) * a.a_result;
            if (o.first_row != nullptr) o.first_row[j] = s_key[j] >> 32;
            for (uint32_t d = 0; d < a.nd; ++d) {
                const SqAttrK &t = a.at[d];
                if (!t.has || o.keys[t.cpos] == nullptr) continue;
                const uint32_t idx = (g / t.stride) % t.span;
                o.keys[t.cpos][j] = (static_cast<uint64_t>(t.amin + idx) * t.a) & t.wmask;
            }
        }
    }
    __syncthreads();
    // error events, ordered by (row, kind)
    const uint32_t total = *a.ev_count;
    const uint32_t ne = total < kSqMaxEvents ? total : kSqMaxEvents;
    p = 1;
    while (p < ne) p <<= 1;
    for (uint32_t i = threadIdx.x; i < p; i += kSqBlock) s_key[i] = i < ne ? a.events[i] : ~0ull;
    if (threadIdx.x < kKinds) s_base[threadIdx.x] = 0;
    __syncthreads();
    bitonic_lds(s_key, p);
    for (uint32_t base = 0; base < ne; base += kSqBlock) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long e = i < ne ? s_key[i] : ~0ull;
        const uint32_t kind = i < ne ? static_cast<uint32_t>((e >> 27) & 31u) : 0xFFu;
        for (uint32_t k = 0; k < static_cast<uint32_t>(kKinds); ++k) {
            uint32_t tot = 0;
            const uint32_t r = block_excl_scan<uint32_t, kSqWaves>(kind == k ? 1u : 0u, tot, s_scan);
            if (kind == k) {
                const uint32_t rank = s_base[k] + r;
                const uint64_t pos = k >= static_cast<uint32_t>(AHEAD_SQ_E_ATTR) ? (e & 0x7FFFFFFull) : (e >> 32);
                if (o.err[k] != nullptr && rank < o.err_cap[k]) o.err[k][rank] = pos * a.a_oid_err;
            }
            __syncthreads();
            if (threadIdx.x == 0) s_base[k] += tot;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        o.d_counts[0] = ng;
        o.d_counts[1] = *a.kept;
        o.d_counts[2] = total > kSqMaxEvents ? AHEAD_SQ_ST_EVENTS : AHEAD_SQ_ST_OK;
        o.d_counts[3] = total;
        for (int k = 0; k < kKinds; ++k) o.d_counts[4 + k] = s_base[k];
    }
}

// Writes a plan-time status (the fused form does not apply) into d_counts.
__global__ void k_sq_status(uint64_t *d_counts, uint64_t status) {
    const int i = threadIdx.x;
    if (i < 4 + kKinds) d_counts[i] = i == 2 ? status : 0ull;
}

}  // namespace

using namespace ahead_host;

namespace {

uint32_t sq_grid() {
    static uint32_t resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        resident = static_cast<uint32_t>(cus);
    }
    return resident;
}

uint32_t prep_grid(uint64_t n) {
    const uint64_t b = (n + kBlock - 1) / kBlock;
    return static_cast<uint32_t>(b == 0 ? 1 : (b < 1024 ? b : 1024));
}

int check_fact_col(const ahead_col *c, uint64_t n, const char *what, bool need64) {
    if (c->data == nullptr) return fail(AHEAD_ERR_INVALID, std::string("star_query: ") + what + " has no data");
    if (c->n != n) return fail(AHEAD_ERR_SIZE_MISMATCH, std::string("star_query: ") + what + " length differs from the fact table");
    if (!c->encoded) return fail(AHEAD_ERR_INVALID, std::string("star_query: ") + what + " is not AN-encoded");
    if (need64 && c->width != 8) return fail(AHEAD_ERR_INVALID, std::string("star_query: ") + what + " must hold 8-byte codes");
    if (c->width != 1 && c->width != 2 && c->width != 4 && c->width != 8)
        return fail(AHEAD_ERR_INVALID, std::string("star_query: ") + what + " has an invalid width");
    const uintptr_t align = c->width == 8 ? 16 : 2 * c->width;
    if (reinterpret_cast<uintptr_t>(c->data) % align != 0)
        return fail(AHEAD_ERR_INVALID, std::string("star_query: ") + what + " is not aligned for paired loads");
    return AHEAD_OK;
}

template <int ND, int NP>
void launch_scan(const SqArgs &a, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL((k_sq_scan<ND, NP>), dim3(grid), dim3(kSqBlock), 0, s, a);
}

template <int ND>
void launch_scan_np(const SqArgs &a, uint32_t grid, hipStream_t s) {
    switch (a.np) {
        case 0: launch_scan<ND, 0>(a, grid, s); break;
        case 1: launch_scan<ND, 1>(a, grid, s); break;
        default: launch_scan<ND, 2>(a, grid, s); break;
    }
}

struct DimPlan {
    DimPrepArgs prep{};
    uint64_t span_cap = 0;  // key span the workspace was sized for
    size_t off_map = 0, off_seen = 0, off_row = 0;
    uint32_t map_words = 0;
    uint32_t kmin = 0, span = 0, amin = 0, aspan = 0, lb = 0;
    double frac = 1.0;  // selected fraction of the dimension
    bool in_lds = false;
};

}  // namespace

extern "C" int ahead_hip_star_query(const ahead_sq_args *q, const ahead_sq_out *out, ahead_hip_ws *ws,
                                    void *stream) {
    if (q == nullptr || out == nullptr || ws == nullptr) return fail(AHEAD_ERR_INVALID, "star_query: NULL argument");
    if (out->d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "star_query: d_counts is NULL");
    if (q->n_dims < 1 || q->n_dims > AHEAD_SQ_MAX_DIMS) return fail(AHEAD_ERR_INVALID, "star_query: 1 to 4 dimensions");
    if (q->n_preds > AHEAD_SQ_MAX_PREDS) return fail(AHEAD_ERR_INVALID, "star_query: at most 2 fact predicates");
    if (q->meas_kind > AHEAD_SQ_MULSUM) return fail(AHEAD_ERR_INVALID, "star_query: unknown measure kind");
    const uint32_t nm = q->meas_kind == AHEAD_SQ_SUM ? 1u : 2u;
    const uint64_t n = q->n_fact;
    if (n >= (1ull << 32)) return fail(AHEAD_ERR_TOO_LARGE, "star_query: more than 2^32-1 fact rows");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st;
    for (uint32_t d = 0; d < q->n_dims; ++d) {
        const ahead_sq_dim &dm = q->dims[d];
        if ((st = check_fact_col(&dm.fk, n, "a dimension's fact FK column", true)) != AHEAD_OK) return st;
        if ((st = check_col(&dm.key, "star_query.key", false, true)) != AHEAD_OK) return st;
        if (dm.key.width != 8) return fail(AHEAD_ERR_INVALID, "star_query: dimension keys must be 8-byte codes");
        if (dm.key.n == 0 || dm.key.n >= (1ull << 27)) return fail(AHEAD_ERR_TOO_LARGE, "star_query: dimension rows must be 1 .. 2^27-1");
        if (dm.fk.unenc_bits > 32 || dm.key.unenc_bits > 32)
            return fail(AHEAD_ERR_INVALID, "star_query: keys wider than 32 bits are not supported");
        if (dm.attr.data != nullptr) {
            if ((st = check_col(&dm.attr, "star_query.attr", false, true)) != AHEAD_OK) return st;
            if (dm.attr.n != dm.key.n) return fail(AHEAD_ERR_SIZE_MISMATCH, "star_query: attribute and key lengths differ");
            if (dm.attr.unenc_bits > 32) return fail(AHEAD_ERR_INVALID, "star_query: attributes wider than 32 bits are not supported");
        }
        if (dm.sel != nullptr && dm.sel_count == nullptr)
            return fail(AHEAD_ERR_INVALID, "star_query: a selection needs its device count");
    }
    for (uint32_t p = 0; p < q->n_preds; ++p)
        if ((st = check_fact_col(&q->preds[p].col, n, "a fact predicate column", false)) != AHEAD_OK) return st;
    for (uint32_t m = 0; m < nm; ++m)
        if ((st = check_fact_col(&q->meas[m], n, "a measure column", false)) != AHEAD_OK) return st;

    // ---- workspace: flags, statistics, group tables, events, then per-dimension maps sized for
    // the largest key span the plan accepts (4 x rows + 1024)
    const uint32_t nd = q->n_dims;
    std::vector<DimPlan> P(nd);
    size_t off = 0;
    const size_t off_flags = off;
    off += 256;
    const size_t off_stats = off;
    off += align_up(static_cast<size_t>(AHEAD_SQ_MAX_DIMS) * kStatWords * 4, 256);
    const size_t off_gsum = off;
    off += static_cast<size_t>(kSqMaxGroups) * 8;
    const size_t off_gfirst = off;
    off += static_cast<size_t>(kSqMaxGroups) * 4;
    const size_t off_kept = off;
    off += 256;  // kept rows (u64) and the event count (u32)
    const size_t clear_end = off;
    const size_t off_events = off;
    off += static_cast<size_t>(kSqMaxEvents) * 8;
    const size_t off_maps = align_up(off, 256);
    size_t maps_words_cap = 0;
    for (uint32_t d = 0; d < nd; ++d) {
        P[d].span_cap = 4ull * q->dims[d].key.n + 1024;
        maps_words_cap += static_cast<size_t>(P[d].span_cap / 2 + 1);  // 16-bit entries at most
    }
    off = off_maps + align_up(maps_words_cap * 4, 256);
    for (uint32_t d = 0; d < nd; ++d) {
        P[d].off_seen = off;
        off += align_up((P[d].span_cap / 32 + 1) * 4, 256);
        P[d].off_row = off;
        off += align_up(P[d].span_cap * 4, 256);
    }
    if ((st = ws_ensure(ws, off)) != AHEAD_OK) return st;
    char *base = ws_region(ws, 0);
    uint32_t *flags = reinterpret_cast<uint32_t *>(base + off_flags);
    uint32_t *stats = reinterpret_cast<uint32_t *>(base + off_stats);
    // clear flags .. event count; the statistics start at all-ones (running minima), count 0
    if ((st = fill_u64(reinterpret_cast<uint64_t *>(base), clear_end / 8, 0, s)) != AHEAD_OK) return st;
    if ((st = fill_u32(stats, static_cast<uint64_t>(AHEAD_SQ_MAX_DIMS) * kStatWords, 0xFFFFFFFFu, s)) != AHEAD_OK) return st;
    if ((st = fill_u32(reinterpret_cast<uint32_t *>(base + off_gfirst), kSqMaxGroups, 0xFFFFFFFFu, s)) != AHEAD_OK) return st;

    for (uint32_t d = 0; d < nd; ++d) {
        const ahead_sq_dim &dm = q->dims[d];
        DimPrepArgs &a = P[d].prep;
        a.key = static_cast<const uint64_t *>(dm.key.data);
        a.n = dm.key.n;
        a.key_inv = dm.key.a_inv;
        a.key_maxu = dm.key.max_u;
        a.kmask = static_cast<uint32_t>(bits_mask(dm.key.unenc_bits));
        a.sel = dm.sel;
        a.sel_count = dm.sel_count;
        a.sel_inv = dm.sel_a_inv;
        a.attr = dm.attr.data;
        a.attr_width = dm.attr.width;
        a.attr_inv = dm.attr.a_inv;
        a.attr_wmask = width_mask(dm.attr.width);
        a.attr_maxu = dm.attr.max_u & a.attr_wmask;
        a.amask = static_cast<uint32_t>(bits_mask(dm.attr.unenc_bits));
        a.stats = stats + d * kStatWords;
        a.flags = flags;
        hipLaunchKernelGGL(k_sq_keystats, dim3(prep_grid(a.n)), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(k_sq_selstats, dim3(prep_grid(a.n)), dim3(kBlock), 0, s, a);
    }
    AHEAD_HIP_TRY(hipGetLastError());
    uint32_t h_stats[AHEAD_SQ_MAX_DIMS * kStatWords];
    uint32_t h_flags[4];
    AHEAD_HIP_TRY(hipMemcpyAsync(h_stats, stats, sizeof h_stats, hipMemcpyDeviceToHost, s));
    AHEAD_HIP_TRY(hipMemcpyAsync(h_flags, flags, sizeof h_flags, hipMemcpyDeviceToHost, s));
    AHEAD_HIP_TRY(hipStreamSynchronize(s));

    // ---- plan
    auto not_fused = [&](uint64_t status) -> int {
        hipLaunchKernelGGL(k_sq_status, dim3(1), dim3(64), 0, s, out->d_counts, status);
        AHEAD_HIP_TRY(hipGetLastError());
        return AHEAD_OK;
    };
    if (h_flags[kFlagBadKey]) return not_fused(AHEAD_SQ_ST_BAD_KEY);
    if (h_flags[kFlagBadOid]) return not_fused(AHEAD_SQ_ST_BAD_OID);
    uint64_t groups = 1;
    size_t map_off = 0;
    for (uint32_t d = 0; d < nd; ++d) {
        const uint32_t *sd = h_stats + d * kStatWords;
        DimPlan &pl = P[d];
        pl.kmin = sd[0];
        const uint32_t kmax = ~sd[1];
        const uint64_t span = static_cast<uint64_t>(kmax) - pl.kmin + 1;
        if (span > pl.span_cap) return not_fused(AHEAD_SQ_ST_KEY_SPAN);
        pl.span = static_cast<uint32_t>(span);
        const uint32_t nsel = sd[4];
        pl.frac = static_cast<double>(nsel) / static_cast<double>(q->dims[d].key.n);
        if (q->dims[d].attr.data != nullptr) {
            const bool any_valid = sd[2] != 0xFFFFFFFFu || ~sd[3] != 0xFFFFFFFFu;
            pl.amin = any_valid ? sd[2] : 0;
            const uint64_t aspan = any_valid ? static_cast<uint64_t>(~sd[3]) - pl.amin + 1 : 1;
            if (aspan > 65534) return not_fused(AHEAD_SQ_ST_GROUPS);
            pl.aspan = static_cast<uint32_t>(aspan);
            uint32_t lb = 1;  // 2-bit entries at least
            while ((1ull << (1u << lb)) < aspan + 2) ++lb;
            pl.lb = lb;
            groups *= aspan;
            if (groups > kSqMaxGroups) return not_fused(AHEAD_SQ_ST_GROUPS);
        } else {
            pl.lb = 0;
            pl.aspan = 0;
        }
        const uint32_t per_word = 32u >> pl.lb;
        pl.map_words = (pl.span + per_word - 1) / per_word;
    }
    // LDS placement: most selective dimension first while the maps fit; evaluation order: LDS
    // maps first, then the L2 maps, each by selected fraction
    std::vector<uint32_t> by_frac(nd);
    std::iota(by_frac.begin(), by_frac.end(), 0u);
    std::stable_sort(by_frac.begin(), by_frac.end(), [&](uint32_t x, uint32_t y) { return P[x].frac < P[y].frac; });
    uint32_t lds_used = 0;
    for (uint32_t d : by_frac) {
        if (lds_used + P[d].map_words <= kSqMapWords) {
            P[d].in_lds = true;
            lds_used += P[d].map_words;
        }
    }
    std::vector<uint32_t> order;
    for (uint32_t d : by_frac)
        if (P[d].in_lds) order.push_back(d);
    for (uint32_t d : by_frac)
        if (!P[d].in_lds) order.push_back(d);
    // maps in the HBM region: LDS maps first (copied as one block), in evaluation order
    for (uint32_t d : order) {
        P[d].off_map = map_off;
        map_off += P[d].map_words;
    }
    uint32_t *maps = reinterpret_cast<uint32_t *>(base + off_maps);
    if ((st = fill_u32(maps, map_off, 0u, s)) != AHEAD_OK) return st;

    SqArgs a{};
    SqFinalArgs f{};
    uint32_t stride = 1;
    for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t d = order[i];
        DimPlan &pl = P[d];
        DimPrepArgs &pa = pl.prep;
        uint32_t *seen = reinterpret_cast<uint32_t *>(base + pl.off_seen);
        if ((st = fill_u32(seen, pl.span / 32 + 1, 0u, s)) != AHEAD_OK) return st;
        pa.kmin = pl.kmin;
        pa.span = pl.span;
        pa.lb = pl.lb;
        pa.amin = pl.amin;
        pa.errv = q->dims[d].attr.data != nullptr ? static_cast<uint32_t>((1ull << (1u << pl.lb)) - 1ull) : 0u;
        pa.map = maps + pl.off_map;
        pa.seen = seen;
        pa.row_of = reinterpret_cast<uint32_t *>(base + pl.off_row);
        hipLaunchKernelGGL(k_sq_keys, dim3(prep_grid(pa.n)), dim3(kBlock), 0, s, pa);
        hipLaunchKernelGGL(k_sq_sel, dim3(prep_grid(pa.n)), dim3(kBlock), 0, s, pa);

        const ahead_sq_dim &dm = q->dims[d];
        SqDimK &k = a.d[i];
        k.fk = static_cast<const uint64_t *>(dm.fk.data);
        k.fk_inv = dm.fk.a_inv;
        k.fk_maxu = dm.fk.max_u;
        k.kmask = static_cast<uint32_t>(bits_mask(dm.fk.unenc_bits));
        k.kmin = pl.kmin;
        k.span = pl.span;
        k.lb = pl.lb;
        k.lds_off = pl.in_lds ? static_cast<uint32_t>(pl.off_map) : kNoLds;
        k.map_off = static_cast<uint32_t>(pl.off_map);
        k.errv = pa.errv;
        k.stride = pa.errv != 0u ? stride : 0u;
        k.flags = dm.flags;
        k.cpos = d;
        k.row_of = pa.row_of;
        SqAttrK &t = f.at[i];
        t.has = pa.errv != 0u;
        t.cpos = d;
        t.stride = stride;
        t.span = pl.aspan ? pl.aspan : 1u;
        t.amin = pl.amin;
        t.a = dm.attr.a;
        t.wmask = width_mask(dm.attr.width);
        if (pa.errv != 0u) stride *= pl.aspan;
    }
    AHEAD_HIP_TRY(hipGetLastError());
    a.n = n;
    a.nd = nd;
    a.np = q->n_preds;
    a.nm = nm;
    a.kind = q->meas_kind;
    for (uint32_t p = 0; p < q->n_preds; ++p) {
        const ahead_sq_pred &pr = q->preds[p];
        SqPredK &k = a.p[p];
        k.col = pr.col.data;
        k.width = pr.col.width;
        k.wmask = width_mask(pr.col.width);
        k.inv = pr.col.a_inv;
        k.maxu = pr.col.max_u & k.wmask;
        k.op1 = pr.op1;
        k.op2 = pr.op2;
        k.comb = pr.comb;
        k.th1 = pr.th1 & k.wmask;
        k.th2 = pr.th2 & k.wmask;
    }
    for (uint32_t m = 0; m < nm; ++m) {
        SqMeasK &k = a.m[m];
        k.col = q->meas[m].data;
        k.width = q->meas[m].width;
        k.wmask = width_mask(k.width);
        k.inv = q->meas[m].a_inv;
        k.maxu = q->meas[m].max_u & k.wmask;
        k.pre = (q->meas_pre_mask >> m) & 1u;
        a.has_pre |= k.pre;
    }
    a.maps = maps;
    a.lds_words = lds_used;
    a.groups = static_cast<uint32_t>(groups);
    a.flags = flags;
    a.gsum = reinterpret_cast<unsigned long long *>(base + off_gsum);
    a.gfirst = reinterpret_cast<uint32_t *>(base + off_gfirst);
    a.kept = reinterpret_cast<unsigned long long *>(base + off_kept);
    a.ev_count = reinterpret_cast<uint32_t *>(base + off_kept + 8);
    a.events = reinterpret_cast<uint64_t *>(base + off_events);
    const uint32_t grid = sq_grid();
    const uint64_t waves = static_cast<uint64_t>(grid) * kSqWaves;
    const uint64_t per = (n + waves - 1) / waves;
    a.rows_per_wave = (per + kSqSeg - 1) / kSqSeg * kSqSeg;
    if (n > 0) {
        switch (nd) {
            case 1: launch_scan_np<1>(a, grid, s); break;
            case 2: launch_scan_np<2>(a, grid, s); break;
            case 3: launch_scan_np<3>(a, grid, s); break;
            default: launch_scan_np<4>(a, grid, s); break;
        }
        AHEAD_HIP_TRY(hipGetLastError());
    }
    f.groups = a.groups;
    f.nd = nd;
    f.gsum = a.gsum;
    f.gfirst = a.gfirst;
    f.kept = a.kept;
    f.flags = flags;
    f.ev_count = a.ev_count;
    f.events = a.events;
    f.a_result = q->a_result;
    f.a_oid_err = q->a_oid_err;
    f.out = *out;
    hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(kSqBlock), 0, s, f);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
