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
#include <type_traits>
#include <string>
#include <vector>

using namespace ahead_dev;

namespace {

constexpr int kSqBlock = 1024;
constexpr int kSqWaves = kSqBlock / 64;
// A wave step covers CH chunks of 128 rows (the registers of a step's codes and of its prefetched
// successor decide; 4 chunks spill even at one dimension with two fact predicates).
constexpr int sq_chunks(int) { return 2; }
constexpr uint32_t kSqMaxGroups = AHEAD_SQ_MAX_GROUPS;
constexpr uint32_t kSqMaxEvents = AHEAD_SQ_MAX_EVENTS;
constexpr uint32_t kSqMapWords = 28416;  // LDS words for maps with 2-chunk steps (with group tables and staging: < 160 KiB)
// map words that fit in LDS beside the group tables and a staging area of 128 * ch rows per wave
constexpr uint32_t sq_map_words(int ch) { return kSqMapWords - static_cast<uint32_t>((ch - 2) * 128 * 16); }
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
    // the set of passing codes as at most three intervals [lo, lo + len] (unused ones repeat the
    // first; `never`: no code passes), so that the scan tests (code - lo) <= len per interval
    uint64_t lo[3], len[3];
    uint32_t never;
};

struct SqMeasK {
    const void *col;
    uint32_t width;
    uint32_t pre;  // fetched right after the predicates
    int pred;      // index of a fact predicate over the same column (checked with it), or -1
    uint64_t inv, maxu, wmask;
};

struct SqArgs {
    uint64_t n, rows_per_wave;
    uint32_t nd, np, nm, kind, has_pre;
    uint32_t pre_load;  // a pre-predicate measure over a column no predicate reads
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

__host__ __device__ __forceinline__ bool pred_pass(const SqPredK &p, uint64_t code) {
    auto cmp = [](int op, uint64_t t, uint64_t th) {
        return op == AHEAD_LT ? t < th : op == AHEAD_LE ? t <= th : op == AHEAD_GT ? t > th
             : op == AHEAD_GE ? t >= th : op == AHEAD_EQ ? t == th : t != th;
    };
    if (p.comb == AHEAD_COMB_NONE) return cmp(p.op1, code, p.th1);
    if (p.comb == AHEAD_COMB_AND) return cmp(p.op1, code, p.th1) && cmp(p.op2, code, p.th2);
    return cmp(p.op1, code, p.th1) || cmp(p.op2, code, p.th2);
}

template <typename U, typename P>
__device__ __forceinline__ bool pred_fast(const P &p, U code) {
    bool in = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) in = in || static_cast<U>(code - static_cast<U>(p.lo[i])) <= static_cast<U>(p.len[i]);
    return in && !p.never;
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
template <typename T>
__device__ __forceinline__ void load_pair(const void *col, uint32_t width, uint64_t r0, T &x0, T &x1) {
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
            x0 = static_cast<T>((static_cast<uint64_t>(q[1]) << 32) | q[0]);
            x1 = static_cast<T>((static_cast<uint64_t>(q[3]) << 32) | q[2]);
            break;
        }
    }
}

// load_pair for a column width known at compile time (PWB bytes)
template <int PWB, typename T>
__device__ __forceinline__ void load_pair_w(const void *col, uint64_t r0, T &x0, T &x1) {
    if constexpr (PWB == 1) {
        const uint16_t w = static_cast<const uint16_t *>(col)[r0 >> 1];
        x0 = static_cast<T>(w & 0xFFu);
        x1 = static_cast<T>(w >> 8);
    } else if constexpr (PWB == 2) {
        const uint32_t w = static_cast<const uint32_t *>(col)[r0 >> 1];
        x0 = static_cast<T>(w & 0xFFFFu);
        x1 = static_cast<T>(w >> 16);
    } else if constexpr (PWB == 4) {
        const uint64_t w = static_cast<const uint64_t *>(col)[r0 >> 1];
        x0 = static_cast<T>(w & 0xFFFFFFFFull);
        x1 = static_cast<T>(w >> 32);
    } else {
        const u32x4 q = load16_nt(static_cast<const uint64_t *>(col) + r0);
        x0 = static_cast<T>((static_cast<uint64_t>(q[1]) << 32) | q[0]);
        x1 = static_cast<T>((static_cast<uint64_t>(q[3]) << 32) | q[2]);
    }
}

template <int PWB>
__device__ __forceinline__ uint64_t load_code_w(const void *col, uint64_t i) {
    if constexpr (PWB == 1) return static_cast<const uint8_t *>(col)[i];
    else if constexpr (PWB == 2) return static_cast<const uint16_t *>(col)[i];
    else if constexpr (PWB == 4) return static_cast<const uint32_t *>(col)[i];
    else return static_cast<const uint64_t *>(col)[i];
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

// One wave step covers CH chunks of 128 rows, each lane holding 2 consecutive rows of every chunk
// (R rows per lane).  Its FK codes (and predicate codes) are loaded into
// registers one step ahead: as soon as a step has decoded its keys, the wave issues the next step's
// loads into the same registers, so they are in flight while the step's dependent reads run (the
// L2 maps, the measures).  Dependent reads are issued for all of a lane's rows at once and waited
// for once: a row whose value is not needed reads a fixed address of the step instead of
// branching, so no read waits behind another row's branch.
// PWB: the byte width shared by the fact predicate columns (1, 2, 4, 8), or 0 when they differ
template <int ND, int NP, int PWB>
__global__ __launch_bounds__(kSqBlock) void k_sq_scan(SqArgs a) {
    constexpr bool PW8 = PWB == 0 || PWB == 8;
    constexpr int CH = sq_chunks(ND);
    constexpr int kSqSeg = CH * 128;  // rows per wave step
    constexpr int R = 2 * CH;         // rows per lane per step
    constexpr uint32_t OFFB = CH == 4 ? 9u : 8u;  // bits of a row's offset in its step
    __shared__ uint32_t s_map[sq_map_words(CH)];
    __shared__ unsigned long long s_sum[kSqMaxGroups];
    __shared__ uint32_t s_first[kSqMaxGroups];
    __shared__ uint32_t s_stage[kSqWaves * kSqSeg];  // per wave: the step's joined rows (offset | group << OFFB)
    static_assert(sizeof(s_map) + sizeof(s_sum) + sizeof(s_first) + sizeof(s_stage) <= 160 * 1024 - 1024,
                  "LDS budget");
    static_assert((1u << OFFB) == kSqSeg && kSqMaxGroups <= (1u << (32 - OFFB)), "staged entry layout");
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

    using PT = typename std::conditional<PW8, uint64_t, uint32_t>::type;  // predicate code registers
    uint64_t fk[ND][R];
    PT pc[NP > 0 ? NP : 1][R];
    // row of register slot r in the step at seg: seg + (r / 2) * 128 + lane * 2 + r % 2
    auto load_step = [&](uint64_t seg) {
        if (seg + kSqSeg <= end) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const uint64_t r0 = seg + c * 128u + lane * 2u;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const u32x4 q = load16_nt(a.d[d].fk + r0);
                    fk[d][2 * c] = (static_cast<uint64_t>(q[1]) << 32) | q[0];
                    fk[d][2 * c + 1] = (static_cast<uint64_t>(q[3]) << 32) | q[2];
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if constexpr (PWB > 0) load_pair_w<PWB>(a.p[p].col, r0, pc[p][2 * c], pc[p][2 * c + 1]);
                    else load_pair(a.p[p].col, a.p[p].width, r0, pc[p][2 * c], pc[p][2 * c + 1]);
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
                const bool in = row < end;
#pragma unroll
                for (int d = 0; d < ND; ++d) fk[d][r] = in ? a.d[d].fk[row] : 0ull;
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if constexpr (PWB > 0) pc[p][r] = in ? static_cast<PT>(load_code_w<PWB>(a.p[p].col, row)) : PT(0);
                    else pc[p][r] = in ? static_cast<PT>(load_code(a.p[p].col, a.p[p].width, row)) : PT(0);
                }
            }
        }
    };

    // the pending batch: one joined row per lane whose measure reads are in flight
    bool pend_act = false;
    uint64_t pend_row = 0;
    uint32_t pend_g = 0;
    uint64_t pend_code[AHEAD_SQ_MAX_MEAS] = {0, 0};
    // finishes the pending batch: kept rows into the group table; returns whether this lane's row
    // holds a corrupt measure (left for the exact path)
    auto finish_pending = [&]() -> bool {
        if (!pend_act) return false;
        pend_act = false;
        uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
        bool bad = pend_g >= a.groups;
#pragma unroll
        for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
            if (m >= static_cast<int>(a.nm)) continue;
            const SqMeasK &mk = a.m[m];
            dm[m] = (pend_code[m] * mk.inv) & mk.wmask;
            bad |= dm[m] > mk.maxu;
        }
        if (bad) return true;
        atomicAdd(&s_sum[pend_g], static_cast<unsigned long long>(measure_value(a, dm[0], dm[1])));
        atomicMin(&s_first[pend_g], static_cast<uint32_t>(pend_row));
        ++kept;
        return false;
    };

    if (begin < end) load_step(begin);
    for (uint64_t seg = begin; seg < end; seg += kSqSeg) {
        // 1. decode and check the codes in registers; predicates on raw codes
        uint32_t key[ND][R];
        uint32_t alive = 0, err = 0;  // bit r: row of slot r
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
            bool e = false, al = row < end;
            const bool live = al;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const SqPredK &pr = a.p[p];
                if constexpr (PW8) {
                    e |= ((pc[p][r] * pr.inv) & pr.wmask) > pr.maxu;
                    al = al && pred_fast<uint64_t>(pr, pc[p][r]);
                } else {  // codes of at most 4 bytes: the low 32 bits of the product decide
                    const uint32_t c32 = static_cast<uint32_t>(pc[p][r]);
                    e |= ((c32 * static_cast<uint32_t>(pr.inv)) & static_cast<uint32_t>(pr.wmask)) >
                         static_cast<uint32_t>(pr.maxu);
                    al = al && pred_fast<uint32_t>(pr, c32);
                }
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const uint64_t dec = fk[d][r] * a.d[d].fk_inv;
                e |= dec > a.d[d].fk_maxu;
                key[d][r] = static_cast<uint32_t>(dec) & a.d[d].kmask;
            }
            if (live && e) err |= 1u << r;
            if (al) alive |= 1u << r;
        }
        // 2. the next step's codes go in flight now
        if (seg + kSqSeg < end) load_step(seg + kSqSeg);
        alive &= ~err;
        const uint64_t dummy = seg;  // a row of this step: reads for rows not needed go here
        // 3. measures fetched right after the predicates: a measure over a predicate's column is
        // already checked with it; any other one is checked here at every row that passed them
        if (a.pre_load) {
#pragma unroll
            for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
                const SqMeasK &mk = a.m[m];
                if (m >= static_cast<int>(a.nm) || !mk.pre || mk.pred >= 0) continue;
                uint64_t code[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
                    code[r] = load_code(mk.col, mk.width, ((alive >> r) & 1u) ? row : dummy);
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (((alive >> r) & 1u) && ((code[r] * mk.inv) & mk.wmask) > mk.maxu) err |= 1u << r;
            }
            alive &= ~err;
        }
        // 4. dimensions in evaluation order (LDS maps first); an L2 map's reads go out together
        uint32_t g[R];
#pragma unroll
        for (int r = 0; r < R; ++r) g[r] = 0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const SqDimK &dd = a.d[d];
            const uint32_t sh = 5u - dd.lb;
            uint32_t w[R], idx[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                idx[r] = key[d][r] - dd.kmin;
                const bool use = ((alive >> r) & 1u) && idx[r] < dd.span;
                if (!use) idx[r] = 0;
                if (!use) alive &= ~(1u << r);
            }
            if (dd.lds_off != kNoLds) {
#pragma unroll
                for (int r = 0; r < R; ++r) w[r] = s_map[dd.lds_off + (idx[r] >> sh)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) w[r] = a.maps[dd.map_off + (idx[r] >> sh)];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t e = map_entry(w[r], idx[r], dd.lb);
                if (e == 0u) alive &= ~(1u << r);
                if (dd.errv != 0u && ((alive >> r) & 1u)) {
                    if (e == dd.errv) err |= 1u << r;
                    g[r] += (e - 1u) * dd.stride;
                }
            }
            alive &= ~err;
        }
        // 5. the rows that joined every dimension (a few per step) are compacted into the wave's
        // staging area; one lane per row reads its measures.  The last 64 rows' reads are left in
        // flight until the next step has decoded its codes (the pending batch); earlier batches
        // (more than 64 rows, rare) are finished at once.  Rows with a corrupt measure go to the
        // exact path.
        const bool prev_bad = finish_pending();
        const uint64_t prev_row = pend_row;
        uint32_t extra_bad = 0;  // bit b: this lane's row of staged batch b holds a corrupt measure
        uint32_t *stage = s_stage + wave * kSqSeg;
        uint32_t total = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool al = (alive >> r) & 1u;
            const uint64_t b = __ballot(al);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
            if (al) stage[total + below] = ((r >> 1) * 128u + lane * 2u + (r & 1)) | (g[r] << OFFB);
            total += static_cast<uint32_t>(__popcll(b));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t i0 = 0; i0 < total; i0 += 64u) {
            const bool act = i0 + lane < total;
            const uint32_t ent = act ? stage[i0 + lane] : 0u;
            pend_act = act;
            pend_row = seg + (ent & ((1u << OFFB) - 1u));
            pend_g = ent >> OFFB;
#pragma unroll
            for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m)
                if (m < static_cast<int>(a.nm)) pend_code[m] = load_code(a.m[m].col, a.m[m].width, act ? pend_row : dummy);
            if (i0 + 64u < total && finish_pending()) extra_bad |= 1u << (i0 >> 6);
        }
        // 6. the exact path (rare), one call site: rows with a corrupt code, the previous step's
        // pending row, rows of extra batches with a corrupt measure
        uint32_t todo = err | (prev_bad ? 1u << R : 0u) | (extra_bad << (R + 1));
        while (__ballot(todo != 0u) != 0ull) {
            if (todo != 0u) {
                const int b = __ffs(todo) - 1;
                todo &= todo - 1u;
                uint64_t row;
                if (b < R) row = seg + (b >> 1) * 128u + lane * 2u + (b & 1);
                else if (b == R) row = prev_row;
                else row = seg + (stage[(b - R - 1) * 64 + lane] & ((1u << OFFB) - 1u));
                exact_row(a, row, s_sum, s_first, kept);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (finish_pending()) exact_row(a, pend_row, s_sum, s_first, kept);
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

// ---- shared scan: several star queries over one fact table in one pass ------------------------
// The batch's fact FK and predicate columns (the union over its queries) are loaded, decoded and
// checked once per wave step; every query then evaluates its predicates, dimension maps,
// compaction and measure batch on those registers, with its own group table in LDS.  A query's
// rows with a corrupt code take the chain-order path with the query's own arguments, exactly as
// in the single-query scan, so each query's results do not depend on its batch.

constexpr uint32_t kSqbMaxQ = AHEAD_SQ_MAX_BATCH;
constexpr uint32_t kSqbGroupPool = 3072;   // group table entries in LDS, over the batch
constexpr uint32_t kSqbMapWords = 26368;   // map words in LDS, over the batch

// a pointer known to address global memory (pointers read from memory are otherwise flat, and a
// flat access is waited for together with every LDS access)
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gbl(const T *p) {
    return (const __attribute__((address_space(1))) T *)(p);
}

__device__ __forceinline__ uint64_t load_code_g(const void *col, uint32_t width, uint64_t i) {
    switch (width) {
        case 1: return gbl(static_cast<const uint8_t *>(col))[i];
        case 2: return gbl(static_cast<const uint16_t *>(col))[i];
        case 4: return gbl(static_cast<const uint32_t *>(col))[i];
        default: return gbl(static_cast<const uint64_t *>(col))[i];
    }
}

struct SqbDimK {  // a query's dimension over one shared FK column
    uint32_t kmin, span, lds_off, map_off, stride, errv;
    uint16_t lb, used;
};

struct SqbPredK {  // a query's predicate over one shared predicate column (codes of 2 or 4 bytes)
    uint32_t used, never;
    uint32_t lo[3], len[3];
};

struct SqbMeasK {
    const void *col;
    uint64_t inv, maxu, wmask;
    uint32_t width, pre;
    int32_t pred;  // the shared predicate column over the same column, or -1
    uint32_t pad;
};

struct SqbQuery {  // what the scan reads: kernel arguments (scalar loads, global pointers)
    SqbDimK d[AHEAD_SQ_MAX_DIMS];    // by shared FK column
    SqbPredK p[AHEAD_SQ_MAX_PREDS];  // by shared predicate column
    SqbMeasK m[AHEAD_SQ_MAX_MEAS];
    const uint32_t *maps;
    uint32_t nm, kind, pre_load, groups, g_lds;
    // the chain-order path: predicates in the query's order, dimensions in chain order
    uint8_t np, nd, pslot[AHEAD_SQ_MAX_PREDS], cslot[AHEAD_SQ_MAX_DIMS], dflags[AHEAD_SQ_MAX_DIMS];
    const uint32_t *row_of[AHEAD_SQ_MAX_DIMS];  // by chain position
    uint32_t *ev_count;
    uint64_t *events;
};

struct SqbCold {  // read once per launch, in device memory
    const uint32_t *flags;
    unsigned long long *gsum;
    uint32_t *gfirst;
    unsigned long long *kept;
    uint32_t lds_base, lds_words;
};

struct SqbColK {
    const void *col;
    uint64_t inv, maxu, wmask;
    uint32_t kmask, width;
};

struct SqbArgs {  // kernel arguments; the shared columns also sit in SqbDesc::cols for the chain-order path
    uint64_t n, rows_per_wave;
    uint32_t nq, pad;
    SqbColK f[AHEAD_SQ_MAX_DIMS];
    SqbColK pc[AHEAD_SQ_MAX_PREDS];
};

// One pass's descriptors in device memory, read through a __restrict__ pointer: nothing in the scan
// writes them, so their uniform loads stay scalar.
struct SqbDesc {
    SqbQuery q[kSqbMaxQ];
    SqbCold c[kSqbMaxQ];
    SqbColK cols[AHEAD_SQ_MAX_DIMS + AHEAD_SQ_MAX_PREDS];  // FK columns, then predicate columns
};
static_assert(sizeof(SqbDesc) <= 4096 - 64, "descriptor block written through kernel arguments");

template <typename T>
__global__ void k_sq_put(T *dst, T v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gblw(T *p) {
    return (__attribute__((address_space(1))) T *)(p);
}

__device__ __forceinline__ void push_event_b(const SqbQuery &Q, uint64_t row, uint32_t kind, uint32_t drow) {
    const uint32_t i = __hip_atomic_fetch_add(gblw(Q.ev_count), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (i < kSqMaxEvents)
        gblw(Q.events)[i] = (row << 32) | (static_cast<uint64_t>(kind) << 27) | (drow & 0x7FFFFFFu);
}

// exact_row over the batch's descriptors: the chain's semantics for one row of query Q, with events
// (predicates in the query's order, dimensions in chain order; the same rules and event kinds).
__device__ void exact_row_b(const SqbColK *__restrict__ cols, const SqbQuery &Q, uint64_t row,
                            unsigned long long *s_sum, uint32_t *s_first, uint32_t &kept) {
    bool alive = true;
#pragma unroll
    for (int p = 0; p < AHEAD_SQ_MAX_PREDS; ++p) {
        if (p >= Q.np) break;
        const uint32_t sl = Q.pslot[p];
        const SqbColK &pc = cols[AHEAD_SQ_MAX_DIMS + sl];
        const uint64_t c = load_code_g(pc.col, pc.width, row);
        if (((c * pc.inv) & pc.wmask) > pc.maxu) push_event_b(Q, row, AHEAD_SQ_E_PRED + p, 0u);
        if (!pred_fast<uint32_t>(Q.p[sl], static_cast<uint32_t>(c))) alive = false;
    }
    if (!alive) return;
    bool dropped = false;
    uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
#pragma unroll
    for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
        const SqbMeasK &mk = Q.m[m];
        if (m >= static_cast<int>(Q.nm) || !mk.pre) continue;
        dm[m] = (load_code_g(mk.col, mk.width, row) * mk.inv) & mk.wmask;
        if (dm[m] > mk.maxu) {
            push_event_b(Q, row, AHEAD_SQ_E_MEAS + m, 0u);
            dropped = true;
        }
    }
    uint32_t ent[AHEAD_SQ_MAX_DIMS] = {0, 0, 0, 0}, keyv[AHEAD_SQ_MAX_DIMS] = {0, 0, 0, 0};
    bool fk_bad[AHEAD_SQ_MAX_DIMS] = {false, false, false, false};
#pragma unroll
    for (int cp = 0; cp < AHEAD_SQ_MAX_DIMS; ++cp) {
        if (cp >= Q.nd) break;
        const uint32_t sl = Q.cslot[cp];
        const SqbColK &f = cols[sl];
        const SqbDimK &d = Q.d[sl];
        const uint64_t dec = gbl(static_cast<const uint64_t *>(f.col))[row] * f.inv;
        if (dec > f.maxu) {
            fk_bad[cp] = true;
            push_event_b(Q, row, AHEAD_SQ_E_FK + cp, 0u);
            if (!(Q.dflags[cp] & AHEAD_SQ_PROBE)) return;
        }
        keyv[cp] = static_cast<uint32_t>(dec) & f.kmask;
        const uint32_t idx = keyv[cp] - d.kmin;
        ent[cp] = idx < d.span ? map_entry(gbl(Q.maps)[d.map_off + (idx >> (5u - d.lb))], idx, d.lb) : 0u;
        if (ent[cp] == 0u) return;
    }
#pragma unroll
    for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
        const SqbMeasK &mk = Q.m[m];
        if (m >= static_cast<int>(Q.nm) || mk.pre) continue;
        dm[m] = (load_code_g(mk.col, mk.width, row) * mk.inv) & mk.wmask;
        if (dm[m] > mk.maxu) {
            push_event_b(Q, row, AHEAD_SQ_E_MEAS + m, 0u);
            dropped = true;
        }
    }
    uint32_t g = 0;
#pragma unroll
    for (int cp = 0; cp < AHEAD_SQ_MAX_DIMS; ++cp) {
        if (cp >= Q.nd) break;
        if ((Q.dflags[cp] & AHEAD_SQ_REFETCH) && fk_bad[cp]) {
            push_event_b(Q, row, AHEAD_SQ_E_REFETCH + cp, 0u);
            dropped = true;
        }
        const SqbDimK &d = Q.d[Q.cslot[cp]];
        if (d.errv == 0u) continue;
        if (ent[cp] == d.errv) {
            push_event_b(Q, row, AHEAD_SQ_E_ATTR + cp, gbl(Q.row_of[cp])[keyv[cp] - d.kmin]);
            dropped = true;
        } else {
            g += (ent[cp] - 1u) * d.stride;
        }
    }
    if (dropped || g >= Q.groups) return;
    const uint64_t v = Q.kind == AHEAD_SQ_SUB ? dm[0] - dm[1] : Q.kind == AHEAD_SQ_MULSUM ? dm[0] * dm[1] : dm[0];
    atomicAdd(&s_sum[g], static_cast<unsigned long long>(v));
    atomicMin(&s_first[g], static_cast<uint32_t>(row));
    ++kept;
}

// NF shared FK columns, NPC shared predicate columns of PWB bytes (0: widths differ)
template <int NF, int NPC, int PWB>
__global__ __launch_bounds__(kSqBlock) void k_sqb_scan(SqbArgs a, const SqbDesc *__restrict__ desc) {
    constexpr int kSeg = 256;  // rows per wave step: 2 chunks of 128
    constexpr int R = 4;       // rows per lane per step
    constexpr uint32_t OFFB = 8u;
    __shared__ uint32_t s_map[kSqbMapWords];
    __shared__ unsigned long long s_sum[kSqbGroupPool];
    __shared__ uint32_t s_first[kSqbGroupPool];
    __shared__ uint32_t s_stage[kSqWaves * kSeg];
    __shared__ uint32_t s_kept[kSqbMaxQ];
    static_assert(sizeof(s_map) + sizeof(s_sum) + sizeof(s_first) + sizeof(s_stage) + sizeof(s_kept) <=
                      160 * 1024 - 1024,
                  "LDS budget");
    static_assert((1u << OFFB) == kSeg && kSqMaxGroups <= (1u << (32 - OFFB)), "staged entry layout");
    uint32_t qmask = 0;  // queries whose plan holds (no repeated key)
    for (uint32_t q = 0; q < a.nq; ++q) {
        const SqbQuery &Q = desc->q[q];
        const SqbCold &C = desc->c[q];
        if (C.flags[kFlagDup] != 0u) continue;
        qmask |= 1u << q;
        for (uint32_t i = threadIdx.x; i < C.lds_words; i += kSqBlock) s_map[C.lds_base + i] = Q.maps[i];
        for (uint32_t g = threadIdx.x; g < Q.groups; g += kSqBlock) {
            s_sum[Q.g_lds + g] = 0ull;
            s_first[Q.g_lds + g] = 0xFFFFFFFFu;
        }
    }
    if (threadIdx.x < kSqbMaxQ) s_kept[threadIdx.x] = 0u;
    __syncthreads();
    if (qmask == 0u) return;

    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * kSqWaves + wave;
    const uint64_t begin = gw * a.rows_per_wave;
    const uint64_t end = begin + a.rows_per_wave < a.n ? begin + a.rows_per_wave : a.n;

    using PT = uint32_t;  // predicate codes of 2 or 4 bytes
    uint64_t fk[NF][R];
    PT pc[NPC > 0 ? NPC : 1][R];
    auto load_step = [&](uint64_t seg) {
        if (seg + kSeg <= end) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint64_t r0 = seg + c * 128u + lane * 2u;
#pragma unroll
                for (int d = 0; d < NF; ++d) {
                    const u32x4 q = load16_nt(static_cast<const uint64_t *>(a.f[d].col) + r0);
                    fk[d][2 * c] = (static_cast<uint64_t>(q[1]) << 32) | q[0];
                    fk[d][2 * c + 1] = (static_cast<uint64_t>(q[3]) << 32) | q[2];
                }
#pragma unroll
                for (int p = 0; p < NPC; ++p) {
                    if constexpr (PWB > 0) load_pair_w<PWB>(a.pc[p].col, r0, pc[p][2 * c], pc[p][2 * c + 1]);
                    else load_pair(a.pc[p].col, a.pc[p].width, r0, pc[p][2 * c], pc[p][2 * c + 1]);
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
                const bool in = row < end;
#pragma unroll
                for (int d = 0; d < NF; ++d) fk[d][r] = in ? static_cast<const uint64_t *>(a.f[d].col)[row] : 0ull;
#pragma unroll
                for (int p = 0; p < NPC; ++p) {
                    if constexpr (PWB > 0) pc[p][r] = in ? static_cast<PT>(load_code_w<PWB>(a.pc[p].col, row)) : PT(0);
                    else pc[p][r] = in ? static_cast<PT>(load_code(a.pc[p].col, a.pc[p].width, row)) : PT(0);
                }
            }
        }
    };

    // the pending batch: one joined row per lane of query pend_q, its measure reads in flight
    bool pend_act = false;
    uint64_t pend_row = 0;
    uint32_t pend_g = 0, pend_q = 0;
    uint64_t pend_code[AHEAD_SQ_MAX_MEAS] = {0, 0};
    // call in wave-uniform control flow; returns whether this lane's row holds a corrupt measure
    auto finish_pending = [&]() -> bool {
        bool bad = false, keep = false;
        if (pend_act) {
            const SqbQuery &P = desc->q[pend_q];
            uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
            bad = pend_g >= P.groups;
#pragma unroll
            for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
                if (m >= static_cast<int>(P.nm)) continue;
                dm[m] = (pend_code[m] * P.m[m].inv) & P.m[m].wmask;
                bad |= dm[m] > P.m[m].maxu;
            }
            if (!bad) {
                const uint64_t v = P.kind == AHEAD_SQ_SUB ? dm[0] - dm[1] : P.kind == AHEAD_SQ_MULSUM ? dm[0] * dm[1] : dm[0];
                atomicAdd(&s_sum[P.g_lds + pend_g], static_cast<unsigned long long>(v));
                atomicMin(&s_first[P.g_lds + pend_g], static_cast<uint32_t>(pend_row));
                keep = true;
            }
        }
        pend_act = false;
        const uint64_t kb = __ballot(keep);
        if (lane == 0 && kb != 0ull) atomicAdd(&s_kept[pend_q], static_cast<uint32_t>(__popcll(kb)));
        return bad;
    };
    auto exact = [&](uint32_t qq, uint64_t row) {
        const SqbQuery &QQ = desc->q[qq];
        uint32_t k = 0;
        exact_row_b(desc->cols, QQ, row, s_sum + QQ.g_lds, s_first + QQ.g_lds, k);
        if (k != 0u) atomicAdd(&s_kept[qq], k);
    };

    if (begin < end) load_step(begin);
    for (uint64_t seg = begin; seg < end; seg += kSeg) {
        // 1. shared: decode and check the FK codes, check the predicate codes
        uint32_t key[NF][R];
        uint32_t ferr[NF];
        uint32_t perr[NPC > 0 ? NPC : 1];
        PT pv[NPC > 0 ? NPC : 1][R];
        uint32_t live = 0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (seg + (r >> 1) * 128u + lane * 2u + (r & 1) < end) live |= 1u << r;
#pragma unroll
        for (int c = 0; c < NF; ++c) {
            ferr[c] = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t dec = fk[c][r] * a.f[c].inv;
                if (dec > a.f[c].maxu) ferr[c] |= 1u << r;
                key[c][r] = static_cast<uint32_t>(dec) & a.f[c].kmask;
            }
        }
#pragma unroll
        for (int p = 0; p < NPC; ++p) {
            perr[p] = 0;
            const SqbColK &pk = a.pc[p];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                pv[p][r] = pc[p][r];
                // codes of at most 4 bytes: the low 32 bits of the product decide
                if (((pc[p][r] * static_cast<uint32_t>(pk.inv)) & static_cast<uint32_t>(pk.wmask)) >
                    static_cast<uint32_t>(pk.maxu))
                    perr[p] |= 1u << r;
            }
        }
        // 2. the next step's codes go in flight now
        if (seg + kSeg < end) load_step(seg + kSeg);
        const uint64_t dummy = seg;
        uint32_t *stage = s_stage + wave * kSeg;
        for (uint32_t q = 0; q < a.nq; ++q) {
            if (((qmask >> q) & 1u) == 0u) continue;
            const SqbQuery &Q = desc->q[q];
            // 3. the query's predicates; its columns' corrupt codes
            uint32_t err = 0, alive = live;
#pragma unroll
            for (int p = 0; p < NPC; ++p) {
                if (!Q.p[p].used) continue;
                err |= perr[p];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (!pred_fast<PT>(Q.p[p], pv[p][r])) alive &= ~(1u << r);
            }
#pragma unroll
            for (int c = 0; c < NF; ++c)
                if (Q.d[c].used) err |= ferr[c];
            err &= live;
            alive &= ~err;
            // 4. measures fetched right after the predicates (not over a predicate's column)
            if (Q.pre_load) {
#pragma unroll
                for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
                    const SqbMeasK &mk = Q.m[m];
                    if (m >= static_cast<int>(Q.nm) || !mk.pre || mk.pred >= 0) continue;
                    uint64_t code[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
                        code[r] = load_code_g(mk.col, mk.width, ((alive >> r) & 1u) ? row : dummy);
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (((alive >> r) & 1u) && ((code[r] * mk.inv) & mk.wmask) > mk.maxu) err |= 1u << r;
                }
                alive &= ~err;
            }
            // 5. the query's dimensions, in the batch's column order
            uint32_t g[R];
#pragma unroll
            for (int r = 0; r < R; ++r) g[r] = 0;
#pragma unroll
            for (int c = 0; c < NF; ++c) {
                const SqbDimK &dd = Q.d[c];
                if (!dd.used) continue;
                const uint32_t sh = 5u - dd.lb;
                uint32_t w[R], idx[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    idx[r] = key[c][r] - dd.kmin;
                    const bool use = ((alive >> r) & 1u) && idx[r] < dd.span;
                    if (!use) idx[r] = 0;
                    if (!use) alive &= ~(1u << r);
                }
                if (dd.lds_off != kNoLds) {
#pragma unroll
                    for (int r = 0; r < R; ++r) w[r] = s_map[dd.lds_off + (idx[r] >> sh)];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) w[r] = gbl(Q.maps)[dd.map_off + (idx[r] >> sh)];
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t e = map_entry(w[r], idx[r], dd.lb);
                    if (e == 0u) alive &= ~(1u << r);
                    if (dd.errv != 0u && ((alive >> r) & 1u)) {
                        if (e == dd.errv) err |= 1u << r;
                        g[r] += (e - 1u) * dd.stride;
                    }
                }
                alive &= ~err;
            }
            // 6. joined rows: compaction, measure reads (the last batch stays pending)
            const bool prev_bad = finish_pending();
            const uint64_t prev_row = pend_row;
            const uint32_t prev_q = pend_q;
            uint32_t extra_bad = 0;
            uint32_t total = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool al = (alive >> r) & 1u;
                const uint64_t b = __ballot(al);
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
                if (al) stage[total + below] = ((r >> 1) * 128u + lane * 2u + (r & 1)) | (g[r] << OFFB);
                total += static_cast<uint32_t>(__popcll(b));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t i0 = 0; i0 < total; i0 += 64u) {
                const bool act = i0 + lane < total;
                const uint32_t ent = act ? stage[i0 + lane] : 0u;
                pend_act = act;
                pend_row = seg + (ent & ((1u << OFFB) - 1u));
                pend_g = ent >> OFFB;
                pend_q = q;
#pragma unroll
                for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m)
                    if (m < static_cast<int>(Q.nm)) pend_code[m] = load_code_g(Q.m[m].col, Q.m[m].width, act ? pend_row : dummy);
                if (i0 + 64u < total && finish_pending()) extra_bad |= 1u << (i0 >> 6);
            }
            // 7. the chain-order path (rare): rows with a corrupt code, the previous pending row,
            // rows of extra batches with a corrupt measure
            uint32_t todo = err | (prev_bad ? 1u << R : 0u) | (extra_bad << (R + 1));
            while (__ballot(todo != 0u) != 0ull) {
                if (todo != 0u) {
                    const int b = __ffs(todo) - 1;
                    todo &= todo - 1u;
                    uint64_t row;
                    uint32_t qq = q;
                    if (b < R) {
                        row = seg + (b >> 1) * 128u + lane * 2u + (b & 1);
                    } else if (b == R) {
                        row = prev_row;
                        qq = prev_q;
                    } else {
                        row = seg + (stage[(b - R - 1) * 64 + lane] & ((1u << OFFB) - 1u));
                    }
                    exact(qq, row);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (finish_pending()) exact(pend_q, pend_row);
    __syncthreads();
    for (uint32_t q = 0; q < a.nq; ++q) {
        if (((qmask >> q) & 1u) == 0u) continue;
        const SqbQuery &Q = desc->q[q];
        const SqbCold &C = desc->c[q];
        for (uint32_t g = threadIdx.x; g < Q.groups; g += kSqBlock) {
            if (s_first[Q.g_lds + g] != 0xFFFFFFFFu) {
                atomicAdd(&C.gsum[g], s_sum[Q.g_lds + g]);
                atomicMin(&C.gfirst[g], s_first[Q.g_lds + g]);
            }
        }
        if (threadIdx.x == 0 && s_kept[q] != 0u) atomicAdd(C.kept, static_cast<unsigned long long>(s_kept[q]));
    }
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

// ---- same-shape shared scan --------------------------------------------------------------------
// Queries of one shape (the same dimensions, FK, key and group attribute columns, join forms, fact
// predicate columns and measures; only their selections and constants differ) share the whole
// per-row work.  Per dimension a presence map holds, per key, a bit per query (the key belongs to
// that query's selection; one bit for all of them when no query selects on the dimension), and
// for a group attribute the key's attribute index over the union of the queries' attribute ranges
// (+ 1; all ones: a corrupt attribute) is either a field of the same entry above the presence bits
// (combined: one lookup) or the entry of a separate attribute map (split: smaller maps, the second
// lookup only where some query still holds the row), or, when the distinct (presence bits, field)
// values of the dimension are few (selections that follow a low-cardinality attribute), a small
// dictionary code per key whose value sits in an LDS class table.  A row costs one or two lookups
// per dimension for all queries.  A fact predicate over values of at most 8 bits is a 256-entry
// table of per-query pass bits, indexed by the decoded value.  What remains per query is the work
// of its joined rows: the group index in its own numbering and its LDS group sums.  A row with a
// corrupt code takes each query's chain-order path over that query's own maps (exact_row_s), as
// in the single-query scan.

constexpr uint32_t kSqsMaxQ = AHEAD_SQ_MAX_SHAPE_PASS;
constexpr uint32_t kSqsQWords = 10;       // per query in LDS: attribute offsets[4], strides[4], groups, group base
constexpr uint32_t kSqsFieldBits = 24;    // a row's union attribute fields, beside its 8-bit offset in the step
// LDS of a same-shape pass: beside the staging area and the small tables, one region split per pass
// between the queries' group tables (12 bytes per group) and the maps (the rest), sized at launch
constexpr uint32_t kSqsLdsBytes = 160 * 1024 - 1024;
constexpr uint32_t kSqsStaticBytes = kSqWaves * 256 * 2 * 4 + AHEAD_SQ_MAX_PREDS * 256 * 4 + kSqsMaxQ * kSqsQWords * 4 + kSqsMaxQ * 4;
constexpr uint32_t kSqsDynBytes = kSqsLdsBytes - kSqsStaticBytes;
constexpr uint32_t kSqsMinMapWords = 1024;
constexpr uint32_t kSqsGroups = (kSqsDynBytes - 4 * kSqsMinMapWords) / 12;  // most groups a pass holds

struct SqsDimK {  // a dimension of the pass, in evaluation order
    const uint64_t *fk;
    uint64_t fk_inv, fk_maxu;
    uint32_t kmask, kmin, span;
    uint32_t lb, lds_off, map_off;        // presence map: log2 entry bits, LDS offset (or kNoLds), offset
    uint32_t all;                         // one presence bit for every query (no query selects here)
    uint32_t pbits;                       // presence bits of an entry (1 or the pass's queries)
    uint32_t comb;                        // the attribute field sits above the presence bits
    uint32_t dict;                        // entries are class codes; their values at s_map[cls_off + code]
    uint32_t cls_off;
    uint32_t alb, a_lds_off, a_map_off;   // separate attribute map (comb 0), likewise
    uint32_t ab;      // bits of the union attribute field (0: no group attribute)
    uint32_t ashift;  // position of the field in a row's packed attribute word
    uint32_t flags;   // AHEAD_SQ_PROBE | AHEAD_SQ_REFETCH
    uint32_t cpos;    // position in the chain
    const uint32_t *row_of;  // dimension row per key index
};

struct SqsPredK {  // a fact predicate column (codes of at most 4 bytes)
    const void *col;
    uint32_t width, table;  // table: the predicate's 256-entry pass-bit table is in LDS
    uint64_t inv, maxu, wmask;
};

struct SqsArgs {
    uint64_t n, rows_per_wave;
    uint32_t nq, nd, np, nm, kind, pre_load;
    SqsDimK d[AHEAD_SQ_MAX_DIMS];
    SqsPredK p[AHEAD_SQ_MAX_PREDS];
    SqMeasK m[AHEAD_SQ_MAX_MEAS];
    const uint32_t *maps;  // presence and attribute maps; the first lds_words words go to LDS
    uint32_t lds_words;
    uint32_t groups;       // group table entries in LDS (the maps follow them)
    const uint32_t *ptab;  // [np][256] pass bits
};

struct SqsQ {  // one query of the pass, by the pass's dimension index (device memory)
    const uint32_t *maps;  // the query's own maps (chain-order path)
    uint32_t map_off[AHEAD_SQ_MAX_DIMS], lb[AHEAD_SQ_MAX_DIMS], errv[AHEAD_SQ_MAX_DIMS];
    uint32_t stride[AHEAD_SQ_MAX_DIMS];  // group strides in the query's own numbering (0: no attribute)
    uint32_t off[AHEAD_SQ_MAX_DIMS];     // the query's attribute minimum minus the union's
    uint32_t groups, goff;               // its groups, their base in the LDS group tables
    SqbPredK p[AHEAD_SQ_MAX_PREDS];      // its predicates as intervals of codes
    const uint32_t *flags;
    uint32_t *ev_count;
    uint64_t *events;
    unsigned long long *gsum;
    uint32_t *gfirst;
    unsigned long long *kept;
};

struct SqsDesc {
    SqsQ q[kSqsMaxQ];
};
constexpr uint32_t kSqsPutQ = 8;  // descriptors written per launch (through kernel arguments)
struct SqsDescPart {
    SqsQ q[kSqsPutQ];
};
static_assert(sizeof(SqsDescPart) <= 4096 - 64 && kSqsMaxQ % kSqsPutQ == 0, "descriptor parts written through kernel arguments");

struct SqsTab {
    uint32_t t[AHEAD_SQ_MAX_PREDS * 256];
};

struct SqsCls {  // a dictionary-coded dimension's classes by code: presence bits, attribute field
    uint32_t v[512];  // (code 0: no query holds the key)
};
constexpr uint32_t kSqsDictSlots = 257;  // per dimension, 64-bit: 256 slots of distinct values, an overflow flag

__device__ __forceinline__ void push_event_s(const SqsQ &Q, uint64_t row, uint32_t kind, uint32_t drow) {
    const uint32_t i = __hip_atomic_fetch_add(gblw(Q.ev_count), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (i < kSqMaxEvents)
        gblw(Q.events)[i] = (row << 32) | (static_cast<uint64_t>(kind) << 27) | (drow & 0x7FFFFFFu);
}

// The chain's semantics for one row of query Q of a same-shape pass: its own maps and constants,
// the pass's columns; dimensions in chain order (the same rules and event kinds as exact_row).
template <int ND, int NP>
__device__ void exact_row_s(const SqsArgs &a, const SqsQ &Q, uint64_t row, unsigned long long *s_sum,
                            uint32_t *s_first, uint32_t &kept) {
    bool alive = true;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const SqsPredK &pc = a.p[p];
        const uint64_t c = load_code_g(pc.col, pc.width, row);
        if (((c * pc.inv) & pc.wmask) > pc.maxu) push_event_s(Q, row, AHEAD_SQ_E_PRED + p, 0u);
        if (!pred_fast<uint32_t>(Q.p[p], static_cast<uint32_t>(c))) alive = false;
    }
    if (!alive) return;
    bool dropped = false;
    uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
#pragma unroll
    for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
        const SqMeasK &mk = a.m[m];
        if (m >= static_cast<int>(a.nm) || !mk.pre) continue;
        dm[m] = (load_code_g(mk.col, mk.width, row) * mk.inv) & mk.wmask;
        if (dm[m] > mk.maxu) {
            push_event_s(Q, row, AHEAD_SQ_E_MEAS + m, 0u);
            dropped = true;
        }
    }
    uint32_t ent[ND], keyv[ND];
    bool fk_bad[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        ent[d] = 0u;
        keyv[d] = 0u;
        fk_bad[d] = false;
    }
    // chain order: position cp is the pass dimension d with cpos == cp (static indices only)
#pragma unroll
    for (int cp = 0; cp < ND; ++cp) {
        bool stop = false;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const SqsDimK &dk = a.d[d];
            if (dk.cpos != static_cast<uint32_t>(cp)) continue;
            const uint64_t dec = gbl(dk.fk)[row] * dk.fk_inv;
            if (dec > dk.fk_maxu) {
                fk_bad[d] = true;
                push_event_s(Q, row, AHEAD_SQ_E_FK + cp, 0u);
                if (!(dk.flags & AHEAD_SQ_PROBE)) stop = true;
            }
            if (!stop) {
                keyv[d] = static_cast<uint32_t>(dec) & dk.kmask;
                const uint32_t idx = keyv[d] - dk.kmin;
                ent[d] = idx < dk.span ? map_entry(gbl(Q.maps)[Q.map_off[d] + (idx >> (5u - Q.lb[d]))], idx, Q.lb[d])
                                       : 0u;
                if (ent[d] == 0u) stop = true;
            }
        }
        if (stop) return;
    }
#pragma unroll
    for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
        const SqMeasK &mk = a.m[m];
        if (m >= static_cast<int>(a.nm) || mk.pre) continue;
        dm[m] = (load_code_g(mk.col, mk.width, row) * mk.inv) & mk.wmask;
        if (dm[m] > mk.maxu) {
            push_event_s(Q, row, AHEAD_SQ_E_MEAS + m, 0u);
            dropped = true;
        }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        if ((a.d[d].flags & AHEAD_SQ_REFETCH) && fk_bad[d]) {
            push_event_s(Q, row, AHEAD_SQ_E_REFETCH + a.d[d].cpos, 0u);
            dropped = true;
        }
    }
    uint32_t g = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        if (Q.errv[d] == 0u) continue;
        if (ent[d] == Q.errv[d]) {
            push_event_s(Q, row, AHEAD_SQ_E_ATTR + a.d[d].cpos, gbl(a.d[d].row_of)[keyv[d] - a.d[d].kmin]);
            dropped = true;
        } else {
            g += (ent[d] - 1u) * Q.stride[d];
        }
    }
    if (dropped || g >= Q.groups) return;
    const uint64_t v = a.kind == AHEAD_SQ_SUB ? dm[0] - dm[1] : a.kind == AHEAD_SQ_MULSUM ? dm[0] * dm[1] : dm[0];
    atomicAdd(&s_sum[g], static_cast<unsigned long long>(v));
    atomicMin(&s_first[g], static_cast<uint32_t>(row));
    ++kept;
}

// ND dimensions, NP fact predicates over codes of PWB bytes (0: widths differ; at most 4 bytes)
template <int ND, int NP, int PWB>
__global__ __launch_bounds__(kSqBlock) void k_sqs_scan(SqsArgs a, const SqsDesc *__restrict__ desc) {
    constexpr int kSeg = 256, R = 4;  // rows per wave step (2 chunks of 128), rows per lane
    constexpr uint32_t OFFB = 8u;     // a staged row: its pass bits, then offset | attribute fields << OFFB
    extern __shared__ __align__(16) unsigned char s_dyn[];  // group sums, first rows, maps
    __shared__ uint32_t s_stage[kSqWaves * kSeg * 2];
    __shared__ uint32_t s_ptab[AHEAD_SQ_MAX_PREDS * 256];
    __shared__ uint32_t s_qd[kSqsMaxQ * kSqsQWords];
    __shared__ uint32_t s_kept[kSqsMaxQ];
    static_assert(sizeof(s_stage) + sizeof(s_ptab) + sizeof(s_qd) + sizeof(s_kept) == kSqsStaticBytes, "LDS budget");
    unsigned long long *s_sum = reinterpret_cast<unsigned long long *>(s_dyn);
    uint32_t *s_first = reinterpret_cast<uint32_t *>(s_dyn + 8u * a.groups);
    uint32_t *s_map = reinterpret_cast<uint32_t *>(s_dyn + 12u * a.groups);
    static_assert((1u << OFFB) == kSeg && kSqsMaxQ <= 32u && kSqsFieldBits <= 32u - OFFB, "staged entry layout");
    uint32_t qmask = 0;  // queries whose plan holds (no repeated key)
    for (uint32_t q = 0; q < a.nq; ++q)
        if (gbl(desc->q[q].flags)[kFlagDup] == 0u) qmask |= 1u << q;
    for (uint32_t i = threadIdx.x; i < a.lds_words; i += kSqBlock) s_map[i] = a.maps[i];
    for (uint32_t i = threadIdx.x; i < a.np * 256u; i += kSqBlock) s_ptab[i] = a.ptab[i];
    for (uint32_t i = threadIdx.x; i < a.groups; i += kSqBlock) {
        s_sum[i] = 0ull;
        s_first[i] = 0xFFFFFFFFu;
    }
    if (threadIdx.x < a.nq * kSqsQWords) {
        const SqsQ &Q = desc->q[threadIdx.x / kSqsQWords];
        const uint32_t k = threadIdx.x % kSqsQWords;
        s_qd[threadIdx.x] = k < 4 ? Q.off[k] : k < 8 ? Q.stride[k - 4] : k == 8 ? Q.groups : Q.goff;
    }
    if (threadIdx.x < kSqsMaxQ) s_kept[threadIdx.x] = 0u;
    __syncthreads();
    if (qmask == 0u) return;

    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * kSqWaves + wave;
    const uint64_t begin = gw * a.rows_per_wave;
    const uint64_t end = begin + a.rows_per_wave < a.n ? begin + a.rows_per_wave : a.n;

    using PT = uint32_t;
    uint64_t fk[ND][R];
    PT pc[NP > 0 ? NP : 1][R];
    auto load_step = [&](uint64_t seg) {
        if (seg + kSeg <= end) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint64_t r0 = seg + c * 128u + lane * 2u;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const u32x4 q = load16_nt(a.d[d].fk + r0);
                    fk[d][2 * c] = (static_cast<uint64_t>(q[1]) << 32) | q[0];
                    fk[d][2 * c + 1] = (static_cast<uint64_t>(q[3]) << 32) | q[2];
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if constexpr (PWB > 0) load_pair_w<PWB>(a.p[p].col, r0, pc[p][2 * c], pc[p][2 * c + 1]);
                    else load_pair(a.p[p].col, a.p[p].width, r0, pc[p][2 * c], pc[p][2 * c + 1]);
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
                const bool in = row < end;
#pragma unroll
                for (int d = 0; d < ND; ++d) fk[d][r] = in ? a.d[d].fk[row] : 0ull;
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if constexpr (PWB > 0) pc[p][r] = in ? static_cast<PT>(load_code_w<PWB>(a.p[p].col, row)) : PT(0);
                    else pc[p][r] = in ? static_cast<PT>(load_code(a.p[p].col, a.p[p].width, row)) : PT(0);
                }
            }
        }
    };

    // the pending batch: one staged row per lane, its measure reads in flight
    bool pend_act = false;
    uint64_t pend_row = 0;
    uint32_t pend_pass = 0, pend_attr = 0;
    uint64_t pend_code[AHEAD_SQ_MAX_MEAS] = {0, 0};
    // finishes the pending batch: the row into each of its queries' group tables; returns the
    // queries whose chain-order path the row needs (a corrupt measure: all of them)
    auto finish_pending = [&]() -> uint32_t {
        if (!pend_act) return 0u;
        pend_act = false;
        uint64_t dm[AHEAD_SQ_MAX_MEAS] = {0, 0};
        bool bad = false;
#pragma unroll
        for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
            if (m >= static_cast<int>(a.nm)) continue;
            dm[m] = (pend_code[m] * a.m[m].inv) & a.m[m].wmask;
            bad |= dm[m] > a.m[m].maxu;
        }
        if (bad) return pend_pass;
        const uint64_t v = a.kind == AHEAD_SQ_SUB ? dm[0] - dm[1] : a.kind == AHEAD_SQ_MULSUM ? dm[0] * dm[1] : dm[0];
        uint32_t rest = pend_pass, redo = 0;
        while (rest != 0u) {
            const uint32_t q = __ffs(rest) - 1;
            rest &= rest - 1u;
            const uint32_t *qs = s_qd + q * kSqsQWords;
            uint32_t g = 0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                if (a.d[d].ab == 0u) continue;
                const uint32_t f = (pend_attr >> a.d[d].ashift) & ((1u << a.d[d].ab) - 1u);
                g += (f - 1u - qs[d]) * qs[4 + d];
            }
            if (g >= qs[8]) {
                redo |= 1u << q;
                continue;
            }
            atomicAdd(&s_sum[qs[9] + g], static_cast<unsigned long long>(v));
            atomicMin(&s_first[qs[9] + g], static_cast<uint32_t>(pend_row));
            atomicAdd(&s_kept[q], 1u);
        }
        return redo;
    };

    if (begin < end) load_step(begin);
    for (uint64_t seg = begin; seg < end; seg += kSeg) {
        // 1. decode and check the codes; each query's predicate bits
        uint32_t key[ND][R];
        uint32_t pass[R], err = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
            bool e = false;
            uint32_t ps = row < end ? qmask : 0u;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const SqsPredK &pk = a.p[p];
                // codes of at most 4 bytes: the low 32 bits of the product decide
                const uint32_t dec = (pc[p][r] * static_cast<uint32_t>(pk.inv)) & static_cast<uint32_t>(pk.wmask);
                e |= dec > static_cast<uint32_t>(pk.maxu);
                if (pk.table) {
                    ps &= s_ptab[p * 256 + (dec & 255u)];
                } else {
                    for (uint32_t q = 0; q < a.nq; ++q)
                        if (!pred_fast<uint32_t>(desc->q[q].p[p], pc[p][r])) ps &= ~(1u << q);
                }
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const uint64_t dec = fk[d][r] * a.d[d].fk_inv;
                e |= dec > a.d[d].fk_maxu;
                key[d][r] = static_cast<uint32_t>(dec) & a.d[d].kmask;
            }
            if (row < end && e) err |= 1u << r;
            pass[r] = e ? 0u : ps;
        }
        // 2. the next step's codes go in flight now
        if (seg + kSeg < end) load_step(seg + kSeg);
        const uint64_t dummy = seg;
        // 3. measures fetched right after the predicates (not over a predicate's column), checked
        // at every row some query's predicates pass; a corrupt one sends the row to every query's
        // chain-order path
        if (a.pre_load) {
#pragma unroll
            for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m) {
                const SqMeasK &mk = a.m[m];
                if (m >= static_cast<int>(a.nm) || !mk.pre || mk.pred >= 0) continue;
                uint64_t code[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t row = seg + (r >> 1) * 128u + lane * 2u + (r & 1);
                    code[r] = load_code(mk.col, mk.width, pass[r] != 0u ? row : dummy);
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (pass[r] != 0u && ((code[r] * mk.inv) & mk.wmask) > mk.maxu) {
                        err |= 1u << r;
                        pass[r] = 0u;
                    }
                }
            }
        }
        // 4. dimensions in evaluation order (LDS maps first): the queries whose selections hold
        // the key, then the union attribute field of rows some query still holds; a corrupt
        // attribute sends the row to the chain-order path of the queries that select its key
        uint32_t attr[R], redo[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            attr[r] = 0u;
            redo[r] = 0u;
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const SqsDimK &dd = a.d[d];
            const uint32_t sh = 5u - dd.lb;
            uint32_t w[R], idx[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                idx[r] = key[d][r] - dd.kmin;
                const bool use = pass[r] != 0u && idx[r] < dd.span;
                if (!use) idx[r] = 0;
                if (!use) pass[r] = 0u;
            }
            if (dd.lds_off != kNoLds) {
#pragma unroll
                for (int r = 0; r < R; ++r) w[r] = s_map[dd.lds_off + (idx[r] >> sh)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) w[r] = a.maps[dd.map_off + (idx[r] >> sh)];
            }
            const uint32_t fmask = (1u << dd.ab) - 1u;
            const uint32_t pmask = dd.pbits >= 32u ? ~0u : (1u << dd.pbits) - 1u;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t e = map_entry(w[r], idx[r], dd.lb);
                uint32_t pres = e & pmask, fd = 0;
                if (dd.dict) {
                    pres = s_map[dd.cls_off + 2u * e];
                    fd = s_map[dd.cls_off + 2u * e + 1u];
                }
                pass[r] &= dd.all ? (pres != 0u ? ~0u : 0u) : pres;
                if (pass[r] == 0u) idx[r] = 0;
                if (dd.comb && dd.ab != 0u && pass[r] != 0u) {
                    const uint32_t f = dd.dict ? fd : (e >> dd.pbits) & fmask;
                    if (f == fmask) {
                        redo[r] |= pass[r];
                        pass[r] = 0u;
                    }
                    attr[r] |= f << dd.ashift;
                }
            }
            if (!dd.comb && dd.ab != 0u) {
                const uint32_t ash = 5u - dd.alb;
                if (dd.a_lds_off != kNoLds) {
#pragma unroll
                    for (int r = 0; r < R; ++r) w[r] = s_map[dd.a_lds_off + (idx[r] >> ash)];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) w[r] = a.maps[dd.a_map_off + (idx[r] >> ash)];
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (pass[r] == 0u) continue;
                    const uint32_t f = map_entry(w[r], idx[r], dd.alb);
                    if (f == fmask) {
                        redo[r] |= pass[r];
                        pass[r] = 0u;
                    }
                    attr[r] |= f << dd.ashift;
                }
            }
        }
        // 5. rows some query joins: compaction into the wave's staging area, one lane per row
        // reads the measures; the last batch of 64 stays pending until the next step's codes are
        // decoded, earlier ones (rare) are finished at once
        const uint32_t prev_redo = finish_pending();
        const uint64_t prev_row = pend_row;
        uint32_t extra_redo[R - 1];
#pragma unroll
        for (int b = 0; b < R - 1; ++b) extra_redo[b] = 0u;
        uint32_t *stage = s_stage + wave * kSeg * 2;
        uint32_t total = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool al = pass[r] != 0u;
            const uint64_t b = __ballot(al);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
            if (al) {
                stage[2 * (total + below)] = pass[r];
                stage[2 * (total + below) + 1] = ((r >> 1) * 128u + lane * 2u + (r & 1)) | (attr[r] << OFFB);
            }
            total += static_cast<uint32_t>(__popcll(b));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const uint32_t i0 = b * 64u;
            if (i0 < total) {
                const bool act = i0 + lane < total;
                const uint32_t ent = act ? stage[2 * (i0 + lane) + 1] : 0u;
                pend_act = act;
                pend_row = seg + (ent & ((1u << OFFB) - 1u));
                pend_pass = act ? stage[2 * (i0 + lane)] : 0u;
                pend_attr = ent >> OFFB;
#pragma unroll
                for (int m = 0; m < AHEAD_SQ_MAX_MEAS; ++m)
                    if (m < static_cast<int>(a.nm)) pend_code[m] = load_code(a.m[m].col, a.m[m].width, act ? pend_row : dummy);
                if (b < R - 1 && i0 + 64u < total) extra_redo[b < R - 1 ? b : 0] = finish_pending();
            }
        }
        // 6. the chain-order path (rare), one call site.  Slots: 0..R-1 this step's rows (a
        // corrupt code: every query; a corrupt attribute: the queries selecting its key), R the
        // previous pending row, R+1.. the rows of this step's earlier batches
        uint32_t slots = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if ((err >> r) & 1u) redo[r] = qmask;
            if (redo[r] != 0u) slots |= 1u << r;
        }
        if (prev_redo != 0u) slots |= 1u << R;
#pragma unroll
        for (int b = 0; b < R - 1; ++b)
            if (extra_redo[b] != 0u) slots |= 1u << (R + 1 + b);
        uint32_t cur = 0;
        uint64_t cur_row = 0;
        while (__ballot(slots != 0u || cur != 0u) != 0ull) {
            if (cur == 0u && slots != 0u) {
                const int b = __ffs(slots) - 1;
                slots &= slots - 1u;
                uint32_t m = prev_redo;
#pragma unroll
                for (int j = 0; j < R; ++j)
                    if (b == j) m = redo[j];
#pragma unroll
                for (int j = 0; j < R - 1; ++j)
                    if (b == R + 1 + j) m = extra_redo[j];
                cur = m;
                if (b < R) cur_row = seg + (b >> 1) * 128u + lane * 2u + (b & 1);
                else if (b == R) cur_row = prev_row;
                else cur_row = seg + (stage[2 * ((b - R - 1) * 64 + lane) + 1] & ((1u << OFFB) - 1u));
            }
            if (cur != 0u) {
                const uint32_t q = __ffs(cur) - 1;
                cur &= cur - 1u;
                uint32_t k = 0;
                const uint32_t goff = s_qd[q * kSqsQWords + 9];
                exact_row_s<ND, NP>(a, desc->q[q], cur_row, s_sum + goff, s_first + goff, k);
                if (k != 0u) atomicAdd(&s_kept[q], k);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    {
        uint32_t cur = finish_pending();
        while (__ballot(cur != 0u) != 0ull) {
            if (cur != 0u) {
                const uint32_t q = __ffs(cur) - 1;
                cur &= cur - 1u;
                uint32_t k = 0;
                const uint32_t goff = s_qd[q * kSqsQWords + 9];
                exact_row_s<ND, NP>(a, desc->q[q], pend_row, s_sum + goff, s_first + goff, k);
                if (k != 0u) atomicAdd(&s_kept[q], k);
            }
        }
    }
    __syncthreads();
    for (uint32_t q = 0; q < a.nq; ++q) {
        if (((qmask >> q) & 1u) == 0u) continue;
        const SqsQ &Q = desc->q[q];
        const uint32_t goff = s_qd[q * kSqsQWords + 9], ng = s_qd[q * kSqsQWords + 8];
        for (uint32_t g = threadIdx.x; g < ng; g += kSqBlock) {
            if (s_first[goff + g] != 0xFFFFFFFFu) {
                atomicAdd(&Q.gsum[g], s_sum[goff + g]);
                atomicMin(&Q.gfirst[g], s_first[goff + g]);
            }
        }
        if (threadIdx.x == 0 && s_kept[q] != 0u) atomicAdd(Q.kept, static_cast<unsigned long long>(s_kept[q]));
    }
}

// A presence map (mode 0: per key the bits of the pass's queries whose selections hold it, or one
// bit for all of them), an attribute map (mode 1: per key its attribute index over the union range
// + 1, all ones for a corrupt attribute, 0 where no query selects it) or both in one entry (mode 2:
// the attribute field above pbits presence bits), from the queries' own maps.  Mode 3 inserts the
// distinct (presence, field) pairs of the selected keys, as field << 32 | presence, into a 256-slot
// table (an overflow sets table[256]; lb must be 5); mode 4 writes each key's class code
// (code_of[slot of its pair], 0 for none).  Queries with a repeated key are left out.
struct SqsCombineArgs {
    uint32_t span, nq, lb, mode, all, pbits, words;
    unsigned long long *table;
    uint8_t code_of[256];
    const uint32_t *qmap[kSqsMaxQ];  // each query's map of this dimension
    const uint32_t *qflags[kSqsMaxQ];
    uint32_t qlb[kSqsMaxQ], qadd[kSqsMaxQ], qerrv[kSqsMaxQ];  // qadd: query's attribute minimum - union's
    uint32_t fall;                   // the corrupt-attribute value (all ones of the field)
    uint32_t *out;
};

__global__ __launch_bounds__(kBlock) void k_sqs_combine(SqsCombineArgs a) {
    uint32_t live = 0;
#pragma unroll
    for (uint32_t q = 0; q < kSqsMaxQ; ++q)
        if (q < a.nq && a.qflags[q][kFlagDup] == 0u) live |= 1u << q;
    const uint32_t per = 32u >> a.lb;
    for (uint32_t w = blockIdx.x * kBlock + threadIdx.x; w < a.words; w += gridDim.x * kBlock) {
        uint32_t word = 0;
        for (uint32_t j = 0; j < per; ++j) {
            const uint32_t idx = w * per + j;
            if (idx >= a.span) break;
            uint32_t pres = 0, f = 0;
#pragma unroll
            for (uint32_t q = 0; q < kSqsMaxQ; ++q) {
                if (((live >> q) & 1u) == 0u) continue;
                const uint32_t v = map_entry(a.qmap[q][idx >> (5u - a.qlb[q])], idx, a.qlb[q]);
                if (v == 0u) continue;
                pres |= 1u << q;
                f = v == a.qerrv[q] ? a.fall : v + a.qadd[q];
            }
            const uint32_t pv = a.all ? (pres != 0u ? 1u : 0u) : pres;
            uint32_t e = a.mode == 1 ? f : a.mode == 0 || a.mode >= 3 || a.pbits >= 32u ? pv : pv | (f << a.pbits);
            if (a.mode >= 3 && pv != 0u) {
                const unsigned long long v = (static_cast<unsigned long long>(f) << 32) | pv;
                uint32_t h = static_cast<uint32_t>((v * 0x9E3779B97F4A7C15ull) >> 56);
                bool done = false;
                for (uint32_t k = 0; k < 256u && !done; ++k, h = (h + 1u) & 255u) {
                    const unsigned long long cur = a.table[h];
                    if (cur == v) {
                        done = true;
                    } else if (cur == 0ull && a.mode == 3) {
                        const unsigned long long old = atomicCAS(&a.table[h], 0ull, v);
                        done = old == 0ull || old == v;
                    }
                }
                if (a.mode == 3 && !done) atomicOr(&a.table[256], 1ull);
                e = done ? a.code_of[(h - 1u) & 255u] : 0u;
            }
            word |= e << (j << a.lb);
        }
        if (a.mode != 3) a.out[w] = word;
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

// The passing codes of a predicate as at most three intervals: its thresholds cut [0, max] into at
// most five segments, each passing or failing as a whole; adjacent passing segments merge.
void pred_intervals(SqPredK &k) {
    const uint64_t mx = k.wmask;
    uint64_t t[2] = {k.th1, k.th2};
    int nt = k.comb == AHEAD_COMB_NONE ? 1 : 2;
    if (nt == 2 && t[1] < t[0]) std::swap(t[0], t[1]);
    if (nt == 2 && t[0] == t[1]) nt = 1;
    struct Seg {
        uint64_t lo, hi;
    };
    Seg seg[5];
    int ns = 0;
    uint64_t next = 0;
    bool more = true;  // next <= mx still to cover
    for (int i = 0; i < nt; ++i) {
        if (t[i] > next) seg[ns++] = {next, t[i] - 1};
        seg[ns++] = {t[i], t[i]};
        if (t[i] == mx) more = false;
        else next = t[i] + 1;
    }
    if (more) seg[ns++] = {next, mx};
    int n = 0;
    for (int i = 0; i < ns; ++i) {
        if (!pred_pass(k, seg[i].lo)) continue;
        if (n > 0 && k.lo[n - 1] + k.len[n - 1] + 1 == seg[i].lo) {
            k.len[n - 1] += seg[i].hi - seg[i].lo + 1;
        } else {
            k.lo[n] = seg[i].lo;
            k.len[n] = seg[i].hi - seg[i].lo;
            ++n;
        }
    }
    k.never = n == 0 ? 1u : 0u;
    if (n == 0) {
        k.lo[0] = 0;
        k.len[0] = 0;
        n = 1;
    }
    for (int i = n; i < 3; ++i) {
        k.lo[i] = k.lo[0];
        k.len[i] = k.len[0];
    }
}

template <int ND, int NP>
void launch_scan(const SqArgs &a, uint32_t grid, hipStream_t s) {
    uint32_t w = a.p[0].width;
    for (uint32_t p = 1; p < a.np; ++p)
        if (a.p[p].width != w) w = 0;
    switch (w) {
        case 2: hipLaunchKernelGGL((k_sq_scan<ND, NP, 2>), dim3(grid), dim3(kSqBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_sq_scan<ND, NP, 4>), dim3(grid), dim3(kSqBlock), 0, s, a); break;
        default: hipLaunchKernelGGL((k_sq_scan<ND, NP, 0>), dim3(grid), dim3(kSqBlock), 0, s, a); break;
    }
}

template <int ND>
void launch_scan_np(const SqArgs &a, uint32_t grid, hipStream_t s) {
    switch (a.np) {
        case 0: hipLaunchKernelGGL((k_sq_scan<ND, 0, 0>), dim3(grid), dim3(kSqBlock), 0, s, a); break;
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

struct ahead_sq_key_index {
    int device = 0;
    const void *key = nullptr;  // the key column it was built from
    uint64_t n = 0;
    uint64_t a_inv = 0;
    uint32_t unenc_bits = 0;
    uint32_t kmin = 0, span = 0;
    int status = AHEAD_SQ_ST_OK;
    uint32_t *row_of = nullptr;  // [span] row of every key (device)
};

extern "C" int ahead_hip_star_key_index_create(const ahead_col *key, ahead_sq_key_index **out, void *stream) {
    if (key == nullptr || out == nullptr) return fail(AHEAD_ERR_INVALID, "star_key_index: NULL argument");
    *out = nullptr;
    int st = check_col(key, "star_key_index.key", false, true);
    if (st != AHEAD_OK) return st;
    if (key->width != 8) return fail(AHEAD_ERR_INVALID, "star_key_index: keys must be 8-byte codes");
    if (key->n == 0 || key->n >= (1ull << 27)) return fail(AHEAD_ERR_TOO_LARGE, "star_key_index: 1 .. 2^27-1 rows");
    if (key->unenc_bits > 32) return fail(AHEAD_ERR_INVALID, "star_key_index: keys wider than 32 bits");
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto *ix = new ahead_sq_key_index();
    AHEAD_HIP_TRY(hipGetDevice(&ix->device));
    ix->key = key->data;
    ix->n = key->n;
    ix->a_inv = key->a_inv;
    ix->unenc_bits = key->unenc_bits;
    const uint64_t span_cap = 4ull * key->n + 1024;
    uint32_t *tmp = nullptr;  // flags (4 words), stats (kStatWords words), seen bitmap
    const size_t seen_words = span_cap / 32 + 1;
    if (hipMalloc(reinterpret_cast<void **>(&tmp), (4 + kStatWords + seen_words) * 4) != hipSuccess) {
        delete ix;
        return fail(AHEAD_ERR_NOMEM, "star_key_index: device allocation failed");
    }
    auto done = [&](int rc) {
        (void)hipFree(tmp);
        if (rc != AHEAD_OK) {
            if (ix->row_of) (void)hipFree(ix->row_of);
            delete ix;
        } else {
            *out = ix;
        }
        return rc;
    };
    uint32_t *flags = tmp, *stats = tmp + 4, *seen = tmp + 4 + kStatWords;
    if ((st = fill_u32(flags, 4, 0u, s)) != AHEAD_OK) return done(st);
    if ((st = fill_u32(stats, kStatWords, 0xFFFFFFFFu, s)) != AHEAD_OK) return done(st);
    DimPrepArgs a{};
    a.key = static_cast<const uint64_t *>(key->data);
    a.n = key->n;
    a.key_inv = key->a_inv;
    a.key_maxu = key->max_u;
    a.kmask = static_cast<uint32_t>(bits_mask(key->unenc_bits));
    a.stats = stats;
    a.flags = flags;
    hipLaunchKernelGGL(k_sq_keystats, dim3(prep_grid(a.n)), dim3(kBlock), 0, s, a);
    uint32_t h[4 + kStatWords];
    if (hipMemcpyAsync(h, tmp, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return done(fail(AHEAD_ERR_HIP, "star_key_index: key statistics"));
    if (h[kFlagBadKey]) {
        ix->status = AHEAD_SQ_ST_BAD_KEY;
        return done(AHEAD_OK);
    }
    ix->kmin = h[4];
    const uint64_t span = static_cast<uint64_t>(~h[5]) - ix->kmin + 1;
    if (span > span_cap) {
        ix->status = AHEAD_SQ_ST_KEY_SPAN;
        return done(AHEAD_OK);
    }
    ix->span = static_cast<uint32_t>(span);
    if (hipMalloc(reinterpret_cast<void **>(&ix->row_of), static_cast<size_t>(ix->span) * 4) != hipSuccess)
        return done(fail(AHEAD_ERR_NOMEM, "star_key_index: device allocation failed"));
    if ((st = fill_u32(seen, ix->span / 32 + 1, 0u, s)) != AHEAD_OK) return done(st);
    a.kmin = ix->kmin;
    a.span = ix->span;
    a.seen = seen;
    a.row_of = ix->row_of;
    hipLaunchKernelGGL(k_sq_keys, dim3(prep_grid(a.n)), dim3(kBlock), 0, s, a);
    if (hipMemcpyAsync(h, flags, 4 * 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return done(fail(AHEAD_ERR_HIP, "star_key_index: key check"));
    if (h[kFlagDup]) ix->status = AHEAD_SQ_ST_DUP_KEY;
    return done(AHEAD_OK);
}

extern "C" int ahead_hip_star_key_index_destroy(ahead_sq_key_index *ix) {
    if (ix == nullptr) return AHEAD_OK;
    if (ix->row_of) AHEAD_HIP_TRY(hipFree(ix->row_of));
    delete ix;
    return AHEAD_OK;
}

extern "C" int ahead_hip_star_key_index_status(const ahead_sq_key_index *ix) {
    return ix == nullptr ? -1 : ix->status;
}

namespace {

// Workspace layout of one star query: flags, statistics, group tables, kept rows and the event
// count, events, then the dimension maps sized for the largest key span the plan accepts
// (4 x rows + 1024, 16-bit entries at most) and, for dimensions without a key index, the key
// bitmap and the row of every key.
struct SqLayout {
    size_t off_flags = 0, off_stats = 0, off_gsum = 0, off_gfirst = 0, off_kept = 0, clear_end = 0;
    size_t off_events = 0, off_maps = 0, total = 0;
    size_t off_seen[AHEAD_SQ_MAX_DIMS] = {}, off_row[AHEAD_SQ_MAX_DIMS] = {};
    uint64_t span_cap[AHEAD_SQ_MAX_DIMS] = {};
};

SqLayout sq_layout(const ahead_sq_args *q) {
    SqLayout L;
    size_t off = 0;
    L.off_flags = off;
    off += 256;
    L.off_stats = off;
    off += align_up(static_cast<size_t>(AHEAD_SQ_MAX_DIMS) * kStatWords * 4, 256);
    L.off_gsum = off;
    off += static_cast<size_t>(kSqMaxGroups) * 8;
    L.off_gfirst = off;
    off += static_cast<size_t>(kSqMaxGroups) * 4;
    L.off_kept = off;
    off += 256;  // kept rows (u64) and the event count (u32)
    L.clear_end = off;
    L.off_events = off;
    off += static_cast<size_t>(kSqMaxEvents) * 8;
    L.off_maps = align_up(off, 256);
    size_t maps_words_cap = 0;
    for (uint32_t d = 0; d < q->n_dims; ++d) {
        L.span_cap[d] = 4ull * q->dims[d].key.n + 1024;
        maps_words_cap += static_cast<size_t>(L.span_cap[d] / 2 + 1);
    }
    off = L.off_maps + align_up(maps_words_cap * 4, 256);
    for (uint32_t d = 0; d < q->n_dims; ++d) {
        if (q->dims[d].key_index != nullptr) continue;
        L.off_seen[d] = off;
        off += align_up((L.span_cap[d] / 32 + 1) * 4, 256);
        L.off_row[d] = off;
        off += align_up(L.span_cap[d] * 4, 256);
    }
    L.total = off;
    return L;
}

int sq_validate(const ahead_sq_args *q, const ahead_sq_out *out) {
    if (q == nullptr || out == nullptr) return fail(AHEAD_ERR_INVALID, "star_query: NULL argument");
    if (out->d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "star_query: d_counts is NULL");
    if (q->n_dims < 1 || q->n_dims > AHEAD_SQ_MAX_DIMS) return fail(AHEAD_ERR_INVALID, "star_query: 1 to 4 dimensions");
    if (q->n_preds > AHEAD_SQ_MAX_PREDS) return fail(AHEAD_ERR_INVALID, "star_query: at most 2 fact predicates");
    if (q->meas_kind > AHEAD_SQ_MULSUM) return fail(AHEAD_ERR_INVALID, "star_query: unknown measure kind");
    const uint32_t nm = q->meas_kind == AHEAD_SQ_SUM ? 1u : 2u;
    const uint64_t n = q->n_fact;
    if (n >= (1ull << 32)) return fail(AHEAD_ERR_TOO_LARGE, "star_query: more than 2^32-1 fact rows");
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
        if (dm.key_index != nullptr &&
            (dm.key_index->key != dm.key.data || dm.key_index->n != dm.key.n || dm.key_index->a_inv != dm.key.a_inv ||
             dm.key_index->unenc_bits != dm.key.unenc_bits))
            return fail(AHEAD_ERR_INVALID, "star_query: a key index built from another key column");
    }
    for (uint32_t p = 0; p < q->n_preds; ++p) {
        const ahead_sq_pred &pr = q->preds[p];
        if ((st = check_fact_col(&pr.col, n, "a fact predicate column", false)) != AHEAD_OK) return st;
        if (pr.op1 < AHEAD_LT || pr.op1 > AHEAD_NE || pr.op2 < AHEAD_LT || pr.op2 > AHEAD_NE ||
            pr.comb < AHEAD_COMB_NONE || pr.comb > AHEAD_COMB_OR)
            return fail(AHEAD_ERR_INVALID, "star_query: unknown predicate or combiner code");
    }
    for (uint32_t m = 0; m < nm; ++m)
        if ((st = check_fact_col(&q->meas[m], n, "a measure column", false)) != AHEAD_OK) return st;
    return AHEAD_OK;
}

// One query between its statistics and its scan.
struct SqQueryPlan {
    const ahead_sq_args *q = nullptr;
    const ahead_sq_out *out = nullptr;
    char *base = nullptr;  // the query's workspace region
    SqLayout L;
    std::vector<DimPlan> P;
    uint64_t status = AHEAD_SQ_ST_OK;  // a plan-time status: the fused form does not apply
    uint64_t groups = 1;
    std::vector<uint32_t> order;  // evaluation order (LDS maps first)
    uint32_t lds_words = 0;       // leading words of the map region that go to LDS
    SqArgs a{};
    SqFinalArgs f{};
    uint32_t *flags() const { return reinterpret_cast<uint32_t *>(base + L.off_flags); }
    uint32_t *stats() const { return reinterpret_cast<uint32_t *>(base + L.off_stats); }
};

// Clears the query's counters and launches its statistics kernels.
int sq_prep(SqQueryPlan &Q, hipStream_t s) {
    const ahead_sq_args *q = Q.q;
    char *base = Q.base;
    const SqLayout &L = Q.L;
    int st;
    // clear flags .. event count; the statistics start at all-ones (running minima), count 0
    if ((st = fill_u64(reinterpret_cast<uint64_t *>(base), L.clear_end / 8, 0, s)) != AHEAD_OK) return st;
    if ((st = fill_u32(Q.stats(), static_cast<uint64_t>(AHEAD_SQ_MAX_DIMS) * kStatWords, 0xFFFFFFFFu, s)) != AHEAD_OK) return st;
    if ((st = fill_u32(reinterpret_cast<uint32_t *>(base + L.off_gfirst), kSqMaxGroups, 0xFFFFFFFFu, s)) != AHEAD_OK) return st;
    Q.P.assign(q->n_dims, DimPlan{});
    for (uint32_t d = 0; d < q->n_dims; ++d) {
        const ahead_sq_dim &dm = q->dims[d];
        DimPlan &pl = Q.P[d];
        pl.span_cap = L.span_cap[d];
        pl.off_seen = L.off_seen[d];
        pl.off_row = L.off_row[d];
        DimPrepArgs &a = pl.prep;
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
        a.stats = Q.stats() + d * kStatWords;
        a.flags = Q.flags();
        if (dm.key_index == nullptr) hipLaunchKernelGGL(k_sq_keystats, dim3(prep_grid(a.n)), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(k_sq_selstats, dim3(prep_grid(a.n)), dim3(kBlock), 0, s, a);
    }
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

// From the statistics read back: key spans, attribute ranges, map sizes, groups; or the status
// that makes the fused form inapplicable (in dimension order, as the oracle reports them).
void sq_plan(SqQueryPlan &Q, const uint32_t *h_stats, const uint32_t *h_flags) {
    const ahead_sq_args *q = Q.q;
    const uint32_t nd = q->n_dims;
    for (uint32_t d = 0; d < nd; ++d) {
        const ahead_sq_key_index *ix = q->dims[d].key_index;
        if (ix != nullptr && ix->status != AHEAD_SQ_ST_OK) {
            Q.status = static_cast<uint64_t>(ix->status);
            return;
        }
    }
    if (h_flags[kFlagBadKey]) {
        Q.status = AHEAD_SQ_ST_BAD_KEY;
        return;
    }
    if (h_flags[kFlagBadOid]) {
        Q.status = AHEAD_SQ_ST_BAD_OID;
        return;
    }
    uint64_t groups = 1;
    for (uint32_t d = 0; d < nd; ++d) {
        const uint32_t *sd = h_stats + d * kStatWords;
        DimPlan &pl = Q.P[d];
        const ahead_sq_key_index *ix = q->dims[d].key_index;
        if (ix != nullptr) {
            pl.kmin = ix->kmin;
            pl.span = ix->span;
        } else {
            pl.kmin = sd[0];
            const uint32_t kmax = ~sd[1];
            const uint64_t span = static_cast<uint64_t>(kmax) - pl.kmin + 1;
            if (span > pl.span_cap) {
                Q.status = AHEAD_SQ_ST_KEY_SPAN;
                return;
            }
            pl.span = static_cast<uint32_t>(span);
        }
        const uint32_t nsel = sd[4];
        pl.frac = static_cast<double>(nsel) / static_cast<double>(q->dims[d].key.n);
        if (q->dims[d].attr.data != nullptr) {
            const bool any_valid = sd[2] != 0xFFFFFFFFu || ~sd[3] != 0xFFFFFFFFu;
            pl.amin = any_valid ? sd[2] : 0;
            const uint64_t aspan = any_valid ? static_cast<uint64_t>(~sd[3]) - pl.amin + 1 : 1;
            if (aspan > 65534) {
                Q.status = AHEAD_SQ_ST_GROUPS;
                return;
            }
            pl.aspan = static_cast<uint32_t>(aspan);
            uint32_t lb = 1;  // 2-bit entries at least
            while ((1ull << (1u << lb)) < aspan + 2) ++lb;
            pl.lb = lb;
            groups *= aspan;
            if (groups > kSqMaxGroups) {
                Q.status = AHEAD_SQ_ST_GROUPS;
                return;
            }
        } else {
            pl.lb = 0;
            pl.aspan = 0;
        }
        const uint32_t per_word = 32u >> pl.lb;
        pl.map_words = (pl.span + per_word - 1) / per_word;
    }
    Q.groups = groups;
}

// Dimensions by selected fraction, most selective first.
std::vector<uint32_t> by_selectivity(const SqQueryPlan &Q) {
    std::vector<uint32_t> v(Q.P.size());
    std::iota(v.begin(), v.end(), 0u);
    std::stable_sort(v.begin(), v.end(), [&](uint32_t x, uint32_t y) { return Q.P[x].frac < Q.P[y].frac; });
    return v;
}

// Fills the maps (LDS maps first in the map region, in evaluation order), builds the key bitmaps
// where no key index holds them, and the scan and finish arguments.  P[d].in_lds is decided.
int sq_build(SqQueryPlan &Q, hipStream_t s) {
    const ahead_sq_args *q = Q.q;
    const uint32_t nd = q->n_dims;
    std::vector<DimPlan> &P = Q.P;
    const std::vector<uint32_t> by_frac = by_selectivity(Q);
    Q.order.clear();
    for (uint32_t d : by_frac)
        if (P[d].in_lds) Q.order.push_back(d);
    for (uint32_t d : by_frac)
        if (!P[d].in_lds) Q.order.push_back(d);
    size_t map_off = 0;
    Q.lds_words = 0;
    for (uint32_t d : Q.order) {
        P[d].off_map = map_off;
        map_off += P[d].map_words;
        if (P[d].in_lds) Q.lds_words += P[d].map_words;
    }
    int st;
    uint32_t *maps = reinterpret_cast<uint32_t *>(Q.base + Q.L.off_maps);
    if ((st = fill_u32(maps, map_off, 0u, s)) != AHEAD_OK) return st;

    SqArgs &a = Q.a;
    SqFinalArgs &f = Q.f;
    a = SqArgs{};
    f = SqFinalArgs{};
    uint32_t stride = 1;
    for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t d = Q.order[i];
        DimPlan &pl = P[d];
        DimPrepArgs &pa = pl.prep;
        const ahead_sq_key_index *ix = q->dims[d].key_index;
        pa.kmin = pl.kmin;
        pa.span = pl.span;
        pa.lb = pl.lb;
        pa.amin = pl.amin;
        pa.errv = q->dims[d].attr.data != nullptr ? static_cast<uint32_t>((1ull << (1u << pl.lb)) - 1ull) : 0u;
        pa.map = maps + pl.off_map;
        if (ix != nullptr) {
            pa.row_of = ix->row_of;
        } else {
            uint32_t *seen = reinterpret_cast<uint32_t *>(Q.base + pl.off_seen);
            if ((st = fill_u32(seen, pl.span / 32 + 1, 0u, s)) != AHEAD_OK) return st;
            pa.seen = seen;
            pa.row_of = reinterpret_cast<uint32_t *>(Q.base + pl.off_row);
            hipLaunchKernelGGL(k_sq_keys, dim3(prep_grid(pa.n)), dim3(kBlock), 0, s, pa);
        }
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
    const uint32_t nm = q->meas_kind == AHEAD_SQ_SUM ? 1u : 2u;
    a.n = q->n_fact;
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
        pred_intervals(k);
    }
    for (uint32_t m = 0; m < nm; ++m) {
        SqMeasK &k = a.m[m];
        k.col = q->meas[m].data;
        k.width = q->meas[m].width;
        k.wmask = width_mask(k.width);
        k.inv = q->meas[m].a_inv;
        k.maxu = q->meas[m].max_u & k.wmask;
        k.pre = (q->meas_pre_mask >> m) & 1u;
        k.pred = -1;
        for (uint32_t p = 0; p < q->n_preds; ++p)
            if (q->preds[p].col.data == k.col && q->preds[p].col.width == k.width && q->preds[p].col.a_inv == k.inv &&
                (q->preds[p].col.max_u & k.wmask) == k.maxu)
                k.pred = static_cast<int>(p);
        a.has_pre |= k.pre;
        if (k.pre && k.pred < 0) a.pre_load = 1;
    }
    a.maps = maps;
    a.lds_words = Q.lds_words;
    a.groups = static_cast<uint32_t>(Q.groups);
    a.flags = Q.flags();
    a.gsum = reinterpret_cast<unsigned long long *>(Q.base + Q.L.off_gsum);
    a.gfirst = reinterpret_cast<uint32_t *>(Q.base + Q.L.off_gfirst);
    a.kept = reinterpret_cast<unsigned long long *>(Q.base + Q.L.off_kept);
    a.ev_count = reinterpret_cast<uint32_t *>(Q.base + Q.L.off_kept + 8);
    a.events = reinterpret_cast<uint64_t *>(Q.base + Q.L.off_events);
    f.groups = a.groups;
    f.nd = nd;
    f.gsum = a.gsum;
    f.gfirst = a.gfirst;
    f.kept = a.kept;
    f.flags = a.flags;
    f.ev_count = a.ev_count;
    f.events = a.events;
    f.a_result = q->a_result;
    f.a_oid_err = q->a_oid_err;
    f.out = *Q.out;
    return AHEAD_OK;
}

int sq_not_fused(const SqQueryPlan &Q, hipStream_t s) {
    hipLaunchKernelGGL(k_sq_status, dim3(1), dim3(64), 0, s, Q.out->d_counts, Q.status);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

uint64_t rows_per_wave(uint64_t n, uint32_t grid) {
    const uint64_t waves = static_cast<uint64_t>(grid) * kSqWaves;
    const uint64_t per = (n + waves - 1) / waves;
    const uint64_t seg_rows = 128ull * sq_chunks(0);
    return (per + seg_rows - 1) / seg_rows * seg_rows;
}

}  // namespace

extern "C" size_t ahead_hip_star_query_ws_bytes(const ahead_sq_args *q) {
    if (q == nullptr || q->n_dims < 1 || q->n_dims > AHEAD_SQ_MAX_DIMS) return 0;
    return kHeaderBytes + sq_layout(q).total;
}

extern "C" int ahead_hip_star_query(const ahead_sq_args *q, const ahead_sq_out *out, ahead_hip_ws *ws,
                                    void *stream) {
    if (ws == nullptr) return fail(AHEAD_ERR_INVALID, "star_query: NULL argument");
    int st = sq_validate(q, out);
    if (st != AHEAD_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    SqQueryPlan Q;
    Q.q = q;
    Q.out = out;
    Q.L = sq_layout(q);
    if ((st = ws_ensure(ws, Q.L.total)) != AHEAD_OK) return st;
    Q.base = ws_region(ws, 0);
    if ((st = sq_prep(Q, s)) != AHEAD_OK) return st;
    uint32_t h_stats[AHEAD_SQ_MAX_DIMS * kStatWords];
    uint32_t h_flags[4];
    AHEAD_HIP_TRY(hipMemcpyAsync(h_stats, Q.stats(), sizeof h_stats, hipMemcpyDeviceToHost, s));
    AHEAD_HIP_TRY(hipMemcpyAsync(h_flags, Q.flags(), sizeof h_flags, hipMemcpyDeviceToHost, s));
    AHEAD_HIP_TRY(hipStreamSynchronize(s));
    sq_plan(Q, h_stats, h_flags);
    if (Q.status != AHEAD_SQ_ST_OK) return sq_not_fused(Q, s);
    // LDS placement: most selective dimension first while the maps fit
    uint32_t lds_used = 0;
    const uint32_t lds_cap = sq_map_words(sq_chunks(static_cast<int>(q->n_dims)));
    for (uint32_t d : by_selectivity(Q)) {
        if (lds_used + Q.P[d].map_words <= lds_cap) {
            Q.P[d].in_lds = true;
            lds_used += Q.P[d].map_words;
        }
    }
    if ((st = sq_build(Q, s)) != AHEAD_OK) return st;
    SqArgs &a = Q.a;
    const uint32_t grid = sq_grid();
    a.rows_per_wave = rows_per_wave(q->n_fact, grid);
    if (q->ev_scan[0] != nullptr) AHEAD_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(q->ev_scan[0]), s));
    if (q->n_fact > 0) {
        switch (q->n_dims) {
            case 1: launch_scan_np<1>(a, grid, s); break;
            case 2: launch_scan_np<2>(a, grid, s); break;
            case 3: launch_scan_np<3>(a, grid, s); break;
            default: launch_scan_np<4>(a, grid, s); break;
        }
        AHEAD_HIP_TRY(hipGetLastError());
    }
    if (q->ev_scan[1] != nullptr) AHEAD_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(q->ev_scan[1]), s));
    hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(kSqBlock), 0, s, Q.f);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

namespace {

// Batch workspace: every query's own region (as ahead_hip_star_query lays it out), then one
// descriptor block per pass (at most nq).
struct SqbLayout {
    std::vector<size_t> off;
    size_t off_qdesc = 0, off_sqs = 0, total = 0;
};

size_t sqs_region_bytes(const ahead_sq_args *const *qs, uint32_t nq);

SqbLayout sqb_layout(const ahead_sq_args *const *qs, uint32_t nq) {
    SqbLayout B;
    size_t off = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        B.off.push_back(off);
        off += align_up(sq_layout(qs[i]).total, 256);
    }
    B.off_qdesc = off;
    off += static_cast<size_t>(nq) * align_up(sizeof(SqbDesc), 256);
    B.off_sqs = off;
    off += nq > 1 ? sqs_region_bytes(qs, nq) : 0;
    B.total = off;
    return B;
}

bool same_col(const ahead_col &x, const ahead_col &y) {
    return x.data == y.data && x.width == y.width && x.a_inv == y.a_inv && x.max_u == y.max_u &&
           x.unenc_bits == y.unenc_bits;
}

// One shared scan: its queries, the shared columns, each query's dimension and predicate slots.
struct SqbPass {
    std::vector<uint32_t> qi;
    std::vector<const ahead_col *> fcols, pcols;
    std::vector<std::vector<int>> dim_slot, pred_slot;
    uint32_t groups = 0;
    bool alone = false;  // a query the shared scan cannot take: it runs by itself
};

// Adds the query to the pass if its columns and group table fit.
bool sqb_try_add(SqbPass &ps, uint32_t qi, const SqQueryPlan &Q) {
    const ahead_sq_args *q = Q.q;
    bool shareable = true;  // the shared scan reads predicate codes of 2 or 4 bytes
    for (uint32_t k = 0; k < q->n_preds; ++k)
        if (q->preds[k].col.width != 2 && q->preds[k].col.width != 4) shareable = false;
    if (ps.alone || (!shareable && !ps.qi.empty())) return false;
    if (ps.qi.size() >= kSqbMaxQ || ps.groups + Q.groups > kSqbGroupPool) return false;
    std::vector<const ahead_col *> f = ps.fcols, p = ps.pcols;
    std::vector<int> ds(q->n_dims, -1), pslot(q->n_preds, -1);
    auto slot_of = [](std::vector<const ahead_col *> &cols, const ahead_col &c, const std::vector<int> &taken) {
        for (size_t j = 0; j < cols.size(); ++j)
            if (same_col(*cols[j], c) && std::find(taken.begin(), taken.end(), static_cast<int>(j)) == taken.end())
                return static_cast<int>(j);
        cols.push_back(&c);
        return static_cast<int>(cols.size() - 1);
    };
    for (uint32_t d = 0; d < q->n_dims; ++d) ds[d] = slot_of(f, q->dims[d].fk, ds);
    for (uint32_t k = 0; k < q->n_preds; ++k) pslot[k] = slot_of(p, q->preds[k].col, pslot);
    if (f.size() > AHEAD_SQ_MAX_DIMS || p.size() > AHEAD_SQ_MAX_PREDS) return false;
    ps.fcols = f;
    ps.pcols = p;
    ps.alone = !shareable;
    ps.qi.push_back(qi);
    ps.dim_slot.push_back(ds);
    ps.pred_slot.push_back(pslot);
    ps.groups += static_cast<uint32_t>(Q.groups);
    return true;
}

template <int NF, int NPC>
void launch_sqb_w(const SqbArgs &a, const SqbDesc *d, uint32_t pwb, uint32_t grid, hipStream_t s) {
    if (pwb == 2) hipLaunchKernelGGL((k_sqb_scan<NF, NPC, 2>), dim3(grid), dim3(kSqBlock), 0, s, a, d);
    else hipLaunchKernelGGL((k_sqb_scan<NF, NPC, 0>), dim3(grid), dim3(kSqBlock), 0, s, a, d);
}

template <int NF>
void launch_sqb_np(const SqbArgs &a, const SqbDesc *d, uint32_t npc, uint32_t pwb, uint32_t grid, hipStream_t s) {
    switch (npc) {
        case 0: hipLaunchKernelGGL((k_sqb_scan<NF, 0, 0>), dim3(grid), dim3(kSqBlock), 0, s, a, d); break;
        case 1: launch_sqb_w<NF, 1>(a, d, pwb, grid, s); break;
        default: launch_sqb_w<NF, 2>(a, d, pwb, grid, s); break;
    }
}

void launch_sqb(const SqbArgs &a, const SqbDesc *d, uint32_t nf, uint32_t npc, uint32_t pwb, uint32_t grid,
                hipStream_t s) {
    switch (nf) {
        case 1: launch_sqb_np<1>(a, d, npc, pwb, grid, s); break;
        case 2: launch_sqb_np<2>(a, d, npc, pwb, grid, s); break;
        case 3: launch_sqb_np<3>(a, d, npc, pwb, grid, s); break;
        default: launch_sqb_np<4>(a, d, npc, pwb, grid, s); break;
    }
}

void launch_single_scan(const SqArgs &a, uint32_t nd, uint32_t grid, hipStream_t s) {
    switch (nd) {
        case 1: launch_scan_np<1>(a, grid, s); break;
        case 2: launch_scan_np<2>(a, grid, s); break;
        case 3: launch_scan_np<3>(a, grid, s); break;
        default: launch_scan_np<4>(a, grid, s); break;
    }
}

// Builds and launches one pass (its queries' maps are filled here).
int sqb_run_pass(std::vector<SqQueryPlan> &plans, SqbPass &ps, SqbDesc *ddesc, uint32_t grid, hipStream_t s) {
    int st;
    if (ps.qi.size() == 1) {  // one query: the single-query scan, with its LDS placement
        SqQueryPlan &Q = plans[ps.qi[0]];
        uint32_t lds_used = 0;
        const uint32_t lds_cap = sq_map_words(sq_chunks(static_cast<int>(Q.q->n_dims)));
        for (uint32_t d : by_selectivity(Q)) {
            if (lds_used + Q.P[d].map_words <= lds_cap) {
                Q.P[d].in_lds = true;
                lds_used += Q.P[d].map_words;
            }
        }
        if ((st = sq_build(Q, s)) != AHEAD_OK) return st;
        Q.a.rows_per_wave = rows_per_wave(Q.q->n_fact, grid);
        if (Q.q->n_fact > 0) launch_single_scan(Q.a, Q.q->n_dims, grid, s);
        AHEAD_HIP_TRY(hipGetLastError());
        return AHEAD_OK;
    }
    // LDS: over all of the pass's dimensions, most selective first, while the maps fit
    struct Cand {
        uint32_t k, d;
        double frac;
    };
    std::vector<Cand> cand;
    for (uint32_t k = 0; k < ps.qi.size(); ++k) {
        const SqQueryPlan &Q = plans[ps.qi[k]];
        for (uint32_t d = 0; d < Q.q->n_dims; ++d) cand.push_back({k, d, Q.P[d].frac});
    }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand &x, const Cand &y) { return x.frac < y.frac; });
    uint32_t lds_used = 0;
    for (const Cand &c : cand) {
        DimPlan &pl = plans[ps.qi[c.k]].P[c.d];
        if (lds_used + pl.map_words <= kSqbMapWords) {
            pl.in_lds = true;
            lds_used += pl.map_words;
        }
    }
    // shared FK columns in order of their most selective dimension
    const uint32_t nf = static_cast<uint32_t>(ps.fcols.size()), npc = static_cast<uint32_t>(ps.pcols.size());
    std::vector<double> fmin(nf, 2.0);
    for (uint32_t k = 0; k < ps.qi.size(); ++k) {
        const SqQueryPlan &Q = plans[ps.qi[k]];
        for (uint32_t d = 0; d < Q.q->n_dims; ++d) {
            const int c = ps.dim_slot[k][d];
            fmin[c] = std::min(fmin[c], Q.P[d].frac + (Q.P[d].in_lds ? 0.0 : 1.0));
        }
    }
    std::vector<uint32_t> fo(nf), rank(nf);
    std::iota(fo.begin(), fo.end(), 0u);
    std::stable_sort(fo.begin(), fo.end(), [&](uint32_t x, uint32_t y) { return fmin[x] < fmin[y]; });
    for (uint32_t i = 0; i < nf; ++i) rank[fo[i]] = i;

    SqbArgs b{};
    b.n = plans[ps.qi[0]].q->n_fact;
    b.rows_per_wave = rows_per_wave(b.n, grid);
    b.nq = static_cast<uint32_t>(ps.qi.size());
    for (uint32_t i = 0; i < nf; ++i) {
        const ahead_col &c = *ps.fcols[fo[i]];
        b.f[i] = SqbColK{c.data, c.a_inv, c.max_u, ~0ull, static_cast<uint32_t>(bits_mask(c.unenc_bits)), 8u};
    }
    uint32_t pwb = npc > 0 ? ps.pcols[0]->width : 0u;
    for (uint32_t p = 0; p < npc; ++p) {
        const ahead_col &c = *ps.pcols[p];
        const uint64_t wm = width_mask(c.width);
        b.pc[p] = SqbColK{c.data, c.a_inv, c.max_u & wm, wm, 0u, c.width};
        if (c.width != pwb) pwb = 0;
    }
    SqbDesc hd{};
    for (uint32_t i = 0; i < nf; ++i) hd.cols[i] = b.f[i];
    for (uint32_t p = 0; p < npc; ++p) hd.cols[AHEAD_SQ_MAX_DIMS + p] = b.pc[p];
    uint32_t lds_base = 0, g_lds = 0;
    for (uint32_t k = 0; k < ps.qi.size(); ++k) {
        SqQueryPlan &Q = plans[ps.qi[k]];
        if ((st = sq_build(Q, s)) != AHEAD_OK) return st;
        const SqArgs &a = Q.a;
        SqbQuery &x = hd.q[k];
        for (uint32_t i = 0; i < a.nd; ++i) {
            const SqDimK &k2 = a.d[i];
            SqbDimK &dd = x.d[rank[ps.dim_slot[k][k2.cpos]]];
            dd.used = 1;
            dd.kmin = k2.kmin;
            dd.span = k2.span;
            dd.lb = static_cast<uint16_t>(k2.lb);
            dd.lds_off = k2.lds_off != kNoLds ? lds_base + k2.lds_off : kNoLds;
            dd.map_off = k2.map_off;
            dd.stride = k2.stride;
            dd.errv = k2.errv;
        }
        for (uint32_t p = 0; p < a.np; ++p) {
            SqbPredK &pp = x.p[ps.pred_slot[k][p]];
            pp.used = 1;
            pp.never = a.p[p].never;
            for (int j = 0; j < 3; ++j) {  // codes of at most 4 bytes: the intervals fit 32 bits
                pp.lo[j] = static_cast<uint32_t>(a.p[p].lo[j]);
                pp.len[j] = static_cast<uint32_t>(a.p[p].len[j]);
            }
        }
        for (uint32_t m = 0; m < a.nm; ++m) {
            const SqMeasK &mk = a.m[m];
            x.m[m] = SqbMeasK{mk.col, mk.inv, mk.maxu, mk.wmask, mk.width, mk.pre,
                              mk.pred >= 0 ? ps.pred_slot[k][mk.pred] : -1, 0u};
        }
        x.maps = a.maps;
        x.nm = a.nm;
        x.kind = a.kind;
        x.pre_load = a.pre_load;
        x.groups = a.groups;
        x.g_lds = g_lds;
        x.np = static_cast<uint8_t>(a.np);
        x.nd = static_cast<uint8_t>(a.nd);
        for (uint32_t p = 0; p < a.np; ++p) x.pslot[p] = static_cast<uint8_t>(ps.pred_slot[k][p]);
        for (uint32_t i = 0; i < a.nd; ++i) {
            const uint32_t cp = a.d[i].cpos;
            x.cslot[cp] = static_cast<uint8_t>(rank[ps.dim_slot[k][cp]]);
            x.dflags[cp] = static_cast<uint8_t>(a.d[i].flags);
            x.row_of[cp] = a.d[i].row_of;
        }
        x.ev_count = a.ev_count;
        x.events = a.events;
        hd.c[k] = SqbCold{a.flags, a.gsum, a.gfirst, a.kept, lds_base, a.lds_words};
        lds_base += a.lds_words;
        g_lds += a.groups;
    }
    hipLaunchKernelGGL((k_sq_put<SqbDesc>), dim3(1), dim3(64), 0, s, ddesc, hd);
    AHEAD_HIP_TRY(hipGetLastError());
    if (b.n > 0) launch_sqb(b, ddesc, nf, npc, pwb, grid, s);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

// ---- same-shape passes ----

bool same_opt_col(const ahead_col &x, const ahead_col &y) {
    if (x.data == nullptr || y.data == nullptr) return x.data == y.data;
    return same_col(x, y);
}

// Whether two planned queries have one shape: the same fact length, dimensions (FK, key and group
// attribute columns, join forms, key ranges), fact predicate columns and measures.
bool same_shape(const SqQueryPlan &X, const SqQueryPlan &Y) {
    const ahead_sq_args *x = X.q, *y = Y.q;
    if (x->n_fact != y->n_fact || x->n_dims != y->n_dims || x->n_preds != y->n_preds ||
        x->meas_kind != y->meas_kind || x->meas_pre_mask != y->meas_pre_mask)
        return false;
    for (uint32_t d = 0; d < x->n_dims; ++d) {
        const ahead_sq_dim &a = x->dims[d], &b = y->dims[d];
        if (!same_col(a.fk, b.fk) || !same_col(a.key, b.key) || !same_opt_col(a.attr, b.attr) || a.flags != b.flags)
            return false;
        if (X.P[d].kmin != Y.P[d].kmin || X.P[d].span != Y.P[d].span) return false;
    }
    for (uint32_t p = 0; p < x->n_preds; ++p)
        if (!same_col(x->preds[p].col, y->preds[p].col)) return false;
    const uint32_t nm = x->meas_kind == AHEAD_SQ_SUM ? 1u : 2u;
    for (uint32_t m = 0; m < nm; ++m)
        if (!same_col(x->meas[m], y->meas[m])) return false;
    return true;
}

// One same-shape pass: its queries and, per chain dimension, the union of their attribute ranges.
struct SqsPass {
    std::vector<uint32_t> qi;
    uint32_t groups = 0;
    uint32_t umin[AHEAD_SQ_MAX_DIMS] = {}, umax[AHEAD_SQ_MAX_DIMS] = {};
    bool closed = false;  // a query that no same-shape pass takes: it runs in a union pass
};

uint32_t field_bits(uint64_t uspan) {  // bits of a field holding 1 .. uspan and an all-ones error value
    uint32_t b = 1;
    while ((1ull << b) < uspan + 2) ++b;
    return b;
}

// Adds the query to the pass if its group table, presence bits and union attribute fields fit.
bool sqs_try_add(SqsPass &ps, uint32_t qi, const std::vector<SqQueryPlan> &plans) {
    const SqQueryPlan &Q = plans[qi];
    const ahead_sq_args *q = Q.q;
    for (uint32_t p = 0; p < q->n_preds; ++p)
        if (q->preds[p].col.width > 4) return false;  // the shared scan reads codes of at most 4 bytes
    if (ps.closed || (!ps.qi.empty() && !same_shape(plans[ps.qi[0]], Q))) return false;
    if (ps.qi.size() >= kSqsMaxQ || ps.groups + Q.groups > kSqsGroups) return false;
    uint32_t umin[AHEAD_SQ_MAX_DIMS], umax[AHEAD_SQ_MAX_DIMS], fields = 0;
    for (uint32_t d = 0; d < q->n_dims; ++d) {
        const DimPlan &pl = Q.P[d];
        const uint32_t hi = pl.amin + (pl.aspan ? pl.aspan - 1 : 0);
        umin[d] = ps.qi.empty() ? pl.amin : std::min(ps.umin[d], pl.amin);
        umax[d] = ps.qi.empty() ? hi : std::max(ps.umax[d], hi);
        if (q->dims[d].attr.data == nullptr) continue;
        const uint32_t ab = field_bits(static_cast<uint64_t>(umax[d]) - umin[d] + 1);
        if (ab > 24) return false;
        fields += ab;
    }
    if (fields > kSqsFieldBits) return false;  // a staged row's union attribute fields share one word
    ps.qi.push_back(qi);
    ps.groups += static_cast<uint32_t>(Q.groups);
    std::copy(umin, umin + q->n_dims, ps.umin);
    std::copy(umax, umax + q->n_dims, ps.umax);
    return true;
}

// Workspace of the same-shape passes (one at a time, in stream order): descriptors, predicate
// tables, presence maps (at most 8 bits per key) and attribute maps (at most 32) sized for the
// largest query's key spans.
size_t sqs_region_bytes(const ahead_sq_args *const *qs, uint32_t nq) {
    size_t maps = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        const SqLayout L = sq_layout(qs[i]);
        size_t m = 0;
        for (uint32_t d = 0; d < qs[i]->n_dims; ++d)
            m += align_up(static_cast<size_t>(L.span_cap[d]) + 4, 256) + align_up(static_cast<size_t>(L.span_cap[d]) * 4 + 4, 256);
        maps = std::max(maps, m);
    }
    return align_up(sizeof(SqsDesc), 256) + align_up(sizeof(SqsTab), 256) +
           align_up(AHEAD_SQ_MAX_DIMS * kSqsDictSlots * 8, 256) + AHEAD_SQ_MAX_DIMS * sizeof(SqsCls) + maps;
}

// One instantiation with its group tables and maps in dynamic LDS (the limit raised once).
template <int ND, int NP, int PWB>
int launch_sqs_one(const SqsArgs &a, const SqsDesc *d, uint32_t grid, hipStream_t s) {
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sqs_scan<ND, NP, PWB>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, kSqsDynBytes);
    AHEAD_HIP_TRY(attr);
    const size_t dyn = 12u * static_cast<size_t>(a.groups) + 4u * static_cast<size_t>(a.lds_words);
    hipLaunchKernelGGL((k_sqs_scan<ND, NP, PWB>), dim3(grid), dim3(kSqBlock), dyn, s, a, d);
    return AHEAD_OK;
}

template <int ND, int NP>
int launch_sqs_w(const SqsArgs &a, const SqsDesc *d, uint32_t pwb, uint32_t grid, hipStream_t s) {
    return pwb == 2 ? launch_sqs_one<ND, NP, 2>(a, d, grid, s) : launch_sqs_one<ND, NP, 0>(a, d, grid, s);
}

template <int ND>
int launch_sqs_np(const SqsArgs &a, const SqsDesc *d, uint32_t pwb, uint32_t grid, hipStream_t s) {
    switch (a.np) {
        case 0: return launch_sqs_one<ND, 0, 0>(a, d, grid, s);
        case 1: return launch_sqs_w<ND, 1>(a, d, pwb, grid, s);
        default: return launch_sqs_w<ND, 2>(a, d, pwb, grid, s);
    }
}

// Builds the pass's queries' own maps, the combined maps, descriptors and predicate tables, and
// launches the shared scan.
int sqs_run_pass(std::vector<SqQueryPlan> &plans, const SqsPass &ps, char *region, uint32_t grid, hipStream_t s) {
    int st;
    const uint32_t nq = static_cast<uint32_t>(ps.qi.size());
    for (uint32_t i : ps.qi) {  // own maps in HBM: the chain-order path and the combine kernels read them
        SqQueryPlan &Q = plans[i];
        for (DimPlan &pl : Q.P) pl.in_lds = false;
        if ((st = sq_build(Q, s)) != AHEAD_OK) return st;
    }
    const SqQueryPlan &Q0 = plans[ps.qi[0]];
    const ahead_sq_args *q0 = Q0.q;
    const uint32_t nd = q0->n_dims;
    SqsDesc *ddesc = reinterpret_cast<SqsDesc *>(region);
    SqsTab *dtab = reinterpret_cast<SqsTab *>(region + align_up(sizeof(SqsDesc), 256));
    unsigned long long *ddict =
        reinterpret_cast<unsigned long long *>(region + align_up(sizeof(SqsDesc), 256) + align_up(sizeof(SqsTab), 256));
    uint32_t *cmaps = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(ddict) +
                                                   align_up(AHEAD_SQ_MAX_DIMS * kSqsDictSlots * 8, 256));

    // map geometry per chain dimension: presence (one bit for all queries where none selects),
    // union attribute field; combined entries (presence | field) or split maps
    uint32_t ab[AHEAD_SQ_MAX_DIMS] = {}, pbits[AHEAD_SQ_MAX_DIMS] = {};
    uint32_t lb[AHEAD_SQ_MAX_DIMS] = {}, alb[AHEAD_SQ_MAX_DIMS] = {}, clb[AHEAD_SQ_MAX_DIMS] = {};
    uint32_t pwords[AHEAD_SQ_MAX_DIMS] = {}, awords[AHEAD_SQ_MAX_DIMS] = {}, cwords[AHEAD_SQ_MAX_DIMS] = {};
    bool all[AHEAD_SQ_MAX_DIMS] = {}, comb[AHEAD_SQ_MAX_DIMS] = {};
    double frac[AHEAD_SQ_MAX_DIMS] = {};
    auto words_of = [](uint32_t span, uint32_t l) { return (span + (32u >> l) - 1) / (32u >> l); };
    for (uint32_t d = 0; d < nd; ++d) {
        all[d] = true;
        for (uint32_t i : ps.qi) all[d] = all[d] && plans[i].q->dims[d].sel == nullptr;
        pbits[d] = all[d] ? 1u : nq;
        while ((1u << lb[d]) < pbits[d]) ++lb[d];
        pwords[d] = words_of(Q0.P[d].span, lb[d]);
        if (q0->dims[d].attr.data != nullptr) {
            ab[d] = field_bits(static_cast<uint64_t>(ps.umax[d]) - ps.umin[d] + 1);
            while ((1u << alb[d]) < ab[d]) ++alb[d];
            while ((1u << clb[d]) < pbits[d] + ab[d]) ++clb[d];
            awords[d] = words_of(Q0.P[d].span, alb[d]);
            cwords[d] = clb[d] <= 5 ? words_of(Q0.P[d].span, clb[d]) : 0u;
        }
        for (uint32_t i : ps.qi) frac[d] += plans[i].P[d].frac;
        frac[d] = std::min(frac[d], 1.0);
    }
    // the queries' maps merged per key into (presence bits | field << pbits) values, as the
    // combine kernel writes them
    auto combine_args = [&](uint32_t d) {
        SqsCombineArgs c{};
        c.span = Q0.P[d].span;
        c.nq = nq;
        c.all = all[d] ? 1u : 0u;
        c.pbits = pbits[d];
        for (uint32_t j = 0; j < nq; ++j) {
            const SqQueryPlan &Q = plans[ps.qi[j]];
            c.qmap[j] = Q.a.maps + Q.P[d].off_map;
            c.qflags[j] = Q.flags();
            c.qlb[j] = Q.P[d].lb;
            c.qadd[j] = Q.P[d].amin - ps.umin[d];
            c.qerrv[j] = Q.P[d].prep.errv;
        }
        c.fall = ab[d] != 0u ? (1u << ab[d]) - 1u : 0u;
        return c;
    };
    // this pass's LDS: its group tables, then the class tables and maps in kSqsMapWords words
    const uint32_t lds_groups = (ps.groups + 63u) / 64u * 64u;
    const uint32_t kSqsMapWords = (kSqsDynBytes - 12u * lds_groups) / 4u;
    // dictionary codes where the distinct (presence, field) pairs are few: each candidate
    // dimension's pairs go into a 256-slot table on the device, read back with one wait for the pass
    bool dict[AHEAD_SQ_MAX_DIMS] = {};
    uint32_t dlb[AHEAD_SQ_MAX_DIMS] = {}, ndict = 0;
    uint8_t code_of[AHEAD_SQ_MAX_DIMS][256] = {};
    SqsCls cls[AHEAD_SQ_MAX_DIMS] = {};
    bool cand[AHEAD_SQ_MAX_DIMS] = {}, any_cand = false;
    for (uint32_t d = 0; d < nd; ++d) {
        cand[d] = pbits[d] + ab[d] > 4;
        any_cand = any_cand || cand[d];
    }
    if (any_cand) {
        if ((st = fill_u64(reinterpret_cast<uint64_t *>(ddict), AHEAD_SQ_MAX_DIMS * kSqsDictSlots, 0u, s)) != AHEAD_OK)
            return st;
        for (uint32_t d = 0; d < nd; ++d) {
            if (!cand[d]) continue;
            SqsCombineArgs c = combine_args(d);
            c.mode = 3;
            c.lb = 5;
            c.words = c.span;
            c.table = ddict + d * kSqsDictSlots;
            hipLaunchKernelGGL(k_sqs_combine, dim3(prep_grid(c.words)), dim3(kBlock), 0, s, c);
        }
        AHEAD_HIP_TRY(hipGetLastError());
        unsigned long long h[AHEAD_SQ_MAX_DIMS * kSqsDictSlots];
        AHEAD_HIP_TRY(hipMemcpyAsync(h, ddict, sizeof h, hipMemcpyDeviceToHost, s));
        AHEAD_HIP_TRY(hipStreamSynchronize(s));
        for (uint32_t d = 0; d < nd; ++d) {
            const unsigned long long *t = h + d * kSqsDictSlots;
            if (!cand[d] || t[256] != 0ull) continue;
            uint32_t n = 0;
            for (uint32_t k = 0; k < 256; ++k)
                if (t[k] != 0ull) n++;
            if (n > 255) continue;
            uint32_t l = 0;
            while ((1u << (1u << l)) <= n) ++l;  // codes 0 .. n
            // codes narrower than the presence entry (no attribute) or than the combined entry
            // (an attribute: one lookup where the split maps take two)
            if (ab[d] == 0u ? l >= lb[d] : (cwords[d] != 0u && l >= clb[d])) continue;
            if ((ndict + 1u) * 512u > kSqsMapWords) continue;  // class tables stay in LDS
            dict[d] = true;
            dlb[d] = l;
            uint32_t code = 0;
            for (uint32_t k = 0; k < 256; ++k) {
                if (t[k] == 0ull) continue;
                code_of[d][k] = static_cast<uint8_t>(++code);
                cls[d].v[2 * code] = static_cast<uint32_t>(t[k]);
                cls[d].v[2 * code + 1] = static_cast<uint32_t>(t[k] >> 32);
            }
            ++ndict;
        }
    }
    for (uint32_t d = 0; d < nd; ++d) {
        if (!dict[d]) continue;
        lb[d] = dlb[d];
        pwords[d] = words_of(Q0.P[d].span, dlb[d]);
    }
    // LDS, most selective dimension (over the union) first while maps fit: a dimension with an
    // attribute takes the combined map if it fits (one LDS lookup), else the split maps if they
    // fit; in L2 it takes combined entries of at most 16 bits (one lookup, the map at most twice
    // the presence map), else split maps (the attribute lookup only where some query holds the
    // row).  Evaluation order: dimensions with an LDS map first, each group by selectivity.
    std::vector<uint32_t> by(nd);
    std::iota(by.begin(), by.end(), 0u);
    std::stable_sort(by.begin(), by.end(), [&](uint32_t x, uint32_t y) { return frac[x] < frac[y]; });
    bool plds[AHEAD_SQ_MAX_DIMS] = {}, alds[AHEAD_SQ_MAX_DIMS] = {};
    uint32_t lds_words = ndict * 512u;  // the class tables of dictionary-coded dimensions come first
    for (uint32_t d : by) {
        if (dict[d]) {
            comb[d] = true;
            if (lds_words + pwords[d] <= kSqsMapWords) plds[d] = true, lds_words += pwords[d];
        } else if (ab[d] == 0u) {
            if (lds_words + pwords[d] <= kSqsMapWords) plds[d] = true, lds_words += pwords[d];
        } else if (cwords[d] != 0u && lds_words + cwords[d] <= kSqsMapWords) {
            comb[d] = plds[d] = true;
            lds_words += cwords[d];
        } else if (lds_words + pwords[d] + awords[d] <= kSqsMapWords) {
            plds[d] = alds[d] = true;
            lds_words += pwords[d] + awords[d];
        } else {
            comb[d] = cwords[d] != 0u && clb[d] <= 4;
        }
    }
    for (uint32_t d : by) {  // a presence map that did not fit with its attribute map may fit alone
        if (ab[d] != 0u && !comb[d] && !plds[d] && lds_words + pwords[d] <= kSqsMapWords) {
            plds[d] = true;
            lds_words += pwords[d];
        }
    }
    for (uint32_t d = 0; d < nd; ++d) {
        if (comb[d] && !dict[d]) {
            lb[d] = clb[d];
            pwords[d] = cwords[d];
        }
        if (comb[d]) awords[d] = 0;
    }
    std::vector<uint32_t> order;
    for (uint32_t d : by)
        if (plds[d]) order.push_back(d);
    for (uint32_t d : by)
        if (!plds[d]) order.push_back(d);
    // map offsets: class tables, LDS maps, then the others
    uint32_t poff[AHEAD_SQ_MAX_DIMS] = {}, aoff[AHEAD_SQ_MAX_DIMS] = {}, coff[AHEAD_SQ_MAX_DIMS] = {}, words = 0;
    for (uint32_t d = 0; d < nd; ++d)
        if (dict[d]) coff[d] = words, words += 512u;
    for (uint32_t d : by) {
        if (plds[d]) poff[d] = words, words += pwords[d];
        if (alds[d]) aoff[d] = words, words += awords[d];
    }
    for (uint32_t d : by) {
        if (!plds[d]) poff[d] = words, words += pwords[d];
        if (awords[d] != 0u && !alds[d]) aoff[d] = words, words += awords[d];
    }

    SqsArgs a{};
    a.n = q0->n_fact;
    a.rows_per_wave = rows_per_wave(a.n, grid);
    a.nq = nq;
    a.nd = nd;
    a.np = q0->n_preds;
    a.nm = Q0.a.nm;
    a.kind = Q0.a.kind;
    a.pre_load = Q0.a.pre_load;
    uint32_t ashift = 0;
    for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t d = order[i];
        const ahead_sq_dim &dm = q0->dims[d];
        SqsDimK &k = a.d[i];
        k.fk = static_cast<const uint64_t *>(dm.fk.data);
        k.fk_inv = dm.fk.a_inv;
        k.fk_maxu = dm.fk.max_u;
        k.kmask = static_cast<uint32_t>(bits_mask(dm.fk.unenc_bits));
        k.kmin = Q0.P[d].kmin;
        k.span = Q0.P[d].span;
        k.lb = lb[d];
        k.lds_off = plds[d] ? poff[d] : kNoLds;
        k.map_off = poff[d];
        k.all = all[d] ? 1u : 0u;
        k.pbits = pbits[d];
        k.comb = comb[d] ? 1u : 0u;
        k.dict = dict[d] ? 1u : 0u;
        k.cls_off = coff[d];
        k.alb = alb[d];
        k.a_lds_off = alds[d] ? aoff[d] : kNoLds;
        k.a_map_off = aoff[d];
        k.ab = ab[d];
        k.ashift = ashift;
        k.flags = dm.flags;
        k.cpos = d;
        k.row_of = Q0.P[d].prep.row_of;
        SqsCombineArgs c = combine_args(d);
        if (dict[d]) {
            c.table = ddict + d * kSqsDictSlots;
            std::copy(code_of[d], code_of[d] + 256, c.code_of);
            hipLaunchKernelGGL((k_sq_put<SqsCls>), dim3(1), dim3(64), 0, s, reinterpret_cast<SqsCls *>(cmaps + coff[d]),
                               cls[d]);
        }
        c.mode = dict[d] ? 4u : comb[d] ? 2u : 0u;
        c.lb = lb[d];
        c.words = pwords[d];
        c.out = cmaps + poff[d];
        hipLaunchKernelGGL(k_sqs_combine, dim3(prep_grid(c.words)), dim3(kBlock), 0, s, c);
        if (ab[d] != 0u && !comb[d]) {
            c.mode = 1;
            c.lb = alb[d];
            c.words = awords[d];
            c.out = cmaps + aoff[d];
            hipLaunchKernelGGL(k_sqs_combine, dim3(prep_grid(c.words)), dim3(kBlock), 0, s, c);
        }
        ashift += ab[d];
    }
    AHEAD_HIP_TRY(hipGetLastError());
    SqsTab ht{};
    uint32_t pwb = a.np > 0 ? q0->preds[0].col.width : 0u;
    for (uint32_t p = 0; p < a.np; ++p) {
        const SqPredK &k0 = Q0.a.p[p];
        SqsPredK &k = a.p[p];
        k.col = k0.col;
        k.width = k0.width;
        k.inv = k0.inv;
        k.maxu = k0.maxu;
        k.wmask = k0.wmask;
        k.table = k0.maxu <= 255 ? 1u : 0u;
        if (k0.width != pwb) pwb = 0;
        if (!k.table) continue;
        const uint64_t A = q0->preds[p].col.a;
        for (uint32_t v = 0; v <= k0.maxu; ++v) {
            const uint64_t code = (static_cast<uint64_t>(v) * A) & k0.wmask;
            uint32_t bits = 0;
            for (uint32_t j = 0; j < nq; ++j)
                if (pred_pass(plans[ps.qi[j]].a.p[p], code)) bits |= 1u << j;
            ht.t[p * 256 + v] = bits;
        }
    }
    for (uint32_t m = 0; m < a.nm; ++m) a.m[m] = Q0.a.m[m];
    a.maps = cmaps;
    a.lds_words = lds_words;
    a.groups = lds_groups;
    a.ptab = dtab->t;

    SqsDesc hd{};
    uint32_t goff = 0;
    for (uint32_t j = 0; j < nq; ++j) {
        const SqQueryPlan &Q = plans[ps.qi[j]];
        SqsQ &x = hd.q[j];
        x.maps = Q.a.maps;
        for (uint32_t i = 0; i < nd; ++i) {
            const uint32_t d = order[i];
            x.map_off[i] = static_cast<uint32_t>(Q.P[d].off_map);
            x.lb[i] = Q.P[d].lb;
            x.errv[i] = Q.P[d].prep.errv;
            for (uint32_t e = 0; e < nd; ++e)
                if (Q.a.d[e].cpos == d) x.stride[i] = Q.a.d[e].stride;
            x.off[i] = ab[d] != 0u ? Q.P[d].amin - ps.umin[d] : 0u;
        }
        x.groups = static_cast<uint32_t>(Q.groups);
        x.goff = goff;
        goff += x.groups;
        for (uint32_t p = 0; p < a.np; ++p) {
            SqbPredK &pp = x.p[p];
            pp.used = 1;
            pp.never = Q.a.p[p].never;
            for (int i = 0; i < 3; ++i) {  // codes of at most 4 bytes: the intervals fit 32 bits
                pp.lo[i] = static_cast<uint32_t>(Q.a.p[p].lo[i]);
                pp.len[i] = static_cast<uint32_t>(Q.a.p[p].len[i]);
            }
        }
        x.flags = Q.a.flags;
        x.ev_count = Q.a.ev_count;
        x.events = Q.a.events;
        x.gsum = Q.a.gsum;
        x.gfirst = Q.a.gfirst;
        x.kept = Q.a.kept;
    }
    for (uint32_t j = 0; j < nq; j += kSqsPutQ) {
        SqsDescPart part{};
        std::copy(hd.q + j, hd.q + j + kSqsPutQ, part.q);
        hipLaunchKernelGGL((k_sq_put<SqsDescPart>), dim3(1), dim3(64), 0, s,
                           reinterpret_cast<SqsDescPart *>(ddesc->q + j), part);
    }
    hipLaunchKernelGGL((k_sq_put<SqsTab>), dim3(1), dim3(64), 0, s, dtab, ht);
    AHEAD_HIP_TRY(hipGetLastError());
    if (a.n > 0) {
        switch (nd) {
            case 1: st = launch_sqs_np<1>(a, ddesc, pwb, grid, s); break;
            case 2: st = launch_sqs_np<2>(a, ddesc, pwb, grid, s); break;
            case 3: st = launch_sqs_np<3>(a, ddesc, pwb, grid, s); break;
            default: st = launch_sqs_np<4>(a, ddesc, pwb, grid, s); break;
        }
        if (st != AHEAD_OK) return st;
        AHEAD_HIP_TRY(hipGetLastError());
    }
    return AHEAD_OK;
}

}  // namespace

extern "C" size_t ahead_hip_star_query_batch_ws_bytes(const ahead_sq_args *const *qs, uint32_t nq) {
    if (qs == nullptr || nq == 0) return 0;
    for (uint32_t i = 0; i < nq; ++i)
        if (qs[i] == nullptr || qs[i]->n_dims < 1 || qs[i]->n_dims > AHEAD_SQ_MAX_DIMS) return 0;
    return kHeaderBytes + sqb_layout(qs, nq).total;
}

extern "C" int ahead_hip_star_query_batch(const ahead_sq_args *const *qs, const ahead_sq_out *const *outs, uint32_t nq,
                                          ahead_hip_ws *ws, void *stream) {
    if (qs == nullptr || outs == nullptr || ws == nullptr) return fail(AHEAD_ERR_INVALID, "star_query_batch: NULL argument");
    if (nq == 0) return AHEAD_OK;
    int st;
    for (uint32_t i = 0; i < nq; ++i) {
        if ((st = sq_validate(qs[i], outs[i])) != AHEAD_OK) return st;
        if (qs[i]->n_fact != qs[0]->n_fact)
            return fail(AHEAD_ERR_SIZE_MISMATCH, "star_query_batch: queries over fact tables of different lengths");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const SqbLayout B = sqb_layout(qs, nq);
    if ((st = ws_ensure(ws, B.total)) != AHEAD_OK) return st;
    char *base = ws_region(ws, 0);
    std::vector<SqQueryPlan> plans(nq);
    for (uint32_t i = 0; i < nq; ++i) {
        SqQueryPlan &Q = plans[i];
        Q.q = qs[i];
        Q.out = outs[i];
        Q.L = sq_layout(qs[i]);
        Q.base = base + B.off[i];
        if ((st = sq_prep(Q, s)) != AHEAD_OK) return st;
    }
    // every query's statistics, one wait
    std::vector<uint32_t> h(static_cast<size_t>(nq) * (AHEAD_SQ_MAX_DIMS * kStatWords + 4));
    for (uint32_t i = 0; i < nq; ++i) {
        uint32_t *hs = h.data() + static_cast<size_t>(i) * (AHEAD_SQ_MAX_DIMS * kStatWords + 4);
        AHEAD_HIP_TRY(hipMemcpyAsync(hs, plans[i].stats(), AHEAD_SQ_MAX_DIMS * kStatWords * 4, hipMemcpyDeviceToHost, s));
        AHEAD_HIP_TRY(hipMemcpyAsync(hs + AHEAD_SQ_MAX_DIMS * kStatWords, plans[i].flags(), 16, hipMemcpyDeviceToHost, s));
    }
    AHEAD_HIP_TRY(hipStreamSynchronize(s));
    // queries of one shape share a scan (same-shape passes of two or more); the others are packed
    // in order into passes over the union of their columns
    std::vector<SqsPass> spasses;
    for (uint32_t i = 0; i < nq; ++i) {
        SqQueryPlan &Q = plans[i];
        const uint32_t *hs = h.data() + static_cast<size_t>(i) * (AHEAD_SQ_MAX_DIMS * kStatWords + 4);
        sq_plan(Q, hs, hs + AHEAD_SQ_MAX_DIMS * kStatWords);
        if (Q.status != AHEAD_SQ_ST_OK) {
            if ((st = sq_not_fused(Q, s)) != AHEAD_OK) return st;
            continue;
        }
        bool added = false;
        for (SqsPass &sp : spasses)
            if ((added = sqs_try_add(sp, i, plans))) break;
        if (!added) {
            spasses.emplace_back();
            if (!sqs_try_add(spasses.back(), i, plans)) {
                spasses.back().qi.push_back(i);
                spasses.back().closed = true;
            }
        }
    }
    std::vector<uint32_t> rest;
    for (const SqsPass &sp : spasses)
        if (sp.qi.size() == 1) rest.push_back(sp.qi[0]);
    std::sort(rest.begin(), rest.end());
    spasses.erase(std::remove_if(spasses.begin(), spasses.end(), [](const SqsPass &sp) { return sp.qi.size() < 2; }),
                  spasses.end());
    std::vector<SqbPass> passes;
    for (uint32_t i : rest) {
        SqQueryPlan &Q = plans[i];
        if (passes.empty() || !sqb_try_add(passes.back(), i, Q)) {
            passes.emplace_back();
            if (!sqb_try_add(passes.back(), i, Q)) return fail(AHEAD_ERR_DEVICE, "star_query_batch: a query fits no pass");
        }
    }
    const uint32_t grid = sq_grid();
    const ahead_sq_args *q0 = qs[0];
    if (q0->ev_scan[0] != nullptr) AHEAD_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(q0->ev_scan[0]), s));
    for (const SqsPass &sp : spasses)
        if ((st = sqs_run_pass(plans, sp, base + B.off_sqs, grid, s)) != AHEAD_OK) return st;
    size_t desc = 0;  // descriptor blocks of the passes, one after another
    for (SqbPass &ps : passes) {
        SqbDesc *d = reinterpret_cast<SqbDesc *>(base + B.off_qdesc + desc * align_up(sizeof(SqbDesc), 256));
        if ((st = sqb_run_pass(plans, ps, d, grid, s)) != AHEAD_OK) return st;
        ++desc;
    }
    if (q0->ev_scan[1] != nullptr) AHEAD_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(q0->ev_scan[1]), s));
    for (const SqsPass &sp : spasses)
        for (uint32_t i : sp.qi) hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(kSqBlock), 0, s, plans[i].f);
    for (const SqbPass &ps : passes)
        for (uint32_t i : ps.qi) hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(kSqBlock), 0, s, plans[i].f);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
