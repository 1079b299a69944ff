// select_tile.hpp -- one look-back tile of selectAN (selectAN_scalar.tcc:66-95, :149-181), shared
// by the per-call kernel (select_an.hip) and the batched persistent kernel (select_batch_an.hip).
#pragma once

#include "host_common.hpp"

namespace ahead_dev {

struct SelectArgs {
    const void *codes;
    uint64_t n;
    uint64_t seqbase;
    uint64_t a_inv;
    uint64_t max_u;
    // canonical predicates: p_k(x) = neg_k ^ (valid_k && ((x - lo_k) mod 2^k) <= span_k)
    uint64_t lo1, span1, lo2, span2;
    uint32_t neg1, neg2, valid1, valid2;
    uint64_t a_head;
    uint64_t a_oid;
    uint64_t *oids;
    uint64_t oids_cap;
    uint64_t *err;
    uint64_t err_cap;
    uint64_t *d_counts;
    uint64_t *partials;  // per-tile decoded sums (fused sum only)
    uint32_t iters;      // 16 KiB sub-tiles per look-back tile (1 .. kMaxSubtiles)
    LbArgs lb;
};

template <typename CodeT>
__device__ __forceinline__ bool in_range(CodeT x, CodeT lo, CodeT span, uint32_t valid, uint32_t neg) {
    const bool in = valid && static_cast<CodeT>(x - lo) <= span;
    return in != static_cast<bool>(neg);
}

// Loads the 4 x 16 bytes of one thread in sub-tile `sb` and evaluates predicate and check.
template <typename CodeT, int COMB, bool SUM>
__device__ __forceinline__ void eval_subtile(const SelectArgs &a, const CodeT *codes, bool aligned, uint64_t sb,
                                             uint32_t &m, uint32_t &e, uint64_t &lsum) {
    constexpr int VEC = 16 / sizeof(CodeT);
    constexpr int CH = 4;
    constexpr int CHUNK = kBlock * VEC;
    const CodeT ainv = static_cast<CodeT>(a.a_inv);
    const CodeT maxu = static_cast<CodeT>(a.max_u);
    const CodeT lo1 = static_cast<CodeT>(a.lo1), sp1 = static_cast<CodeT>(a.span1);
    const CodeT lo2 = static_cast<CodeT>(a.lo2), sp2 = static_cast<CodeT>(a.span2);
    CodeT v[CH][VEC];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const uint64_t r0 = sb + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC;
        if (aligned && r0 + VEC <= a.n) {
            const u32x4 q = load16_nt(codes + r0);
            __builtin_memcpy(&v[j][0], &q, 16);
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) v[j][k] = (r0 + k < a.n) ? codes[r0 + k] : CodeT(0);
        }
    }
    m = 0;
    e = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint64_t row = sb + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC + k;
            const bool live = row < a.n;
            const CodeT x = v[j][k];
            const CodeT d = an_mul<CodeT>(x, ainv);
            const bool bad = live && d > maxu;
            bool hit = in_range<CodeT>(x, lo1, sp1, a.valid1, a.neg1);
            if constexpr (COMB == AHEAD_COMB_AND) hit = hit && in_range<CodeT>(x, lo2, sp2, a.valid2, a.neg2);
            if constexpr (COMB == AHEAD_COMB_OR) hit = hit || in_range<CodeT>(x, lo2, sp2, a.valid2, a.neg2);
            hit = hit && live;
            m |= static_cast<uint32_t>(hit) << (j * VEC + k);
            e |= static_cast<uint32_t>(bad) << (j * VEC + k);
            if constexpr (SUM) lsum += (hit && !bad) ? static_cast<uint64_t>(d) : 0ull;
        }
    }
}

// LDS of one select tile.  ROWS: rows of one 16 KiB sub-tile of the narrowest code the kernel
// handles (8192 for 16-bit codes).
template <int ROWS>
struct SelectShared {
    uint64_t scan[kWaves];
    uint32_t excl[2];
    uint32_t mb[kMaxSubtiles * kBlock];  // match bits of every row of the tile
    uint16_t rows[ROWS];                 // ranked row offsets of one sub-tile
};

// One look-back tile of a select: phase 1, resolve, phase 2 (see the file comment of
// select_an.hip).  `tile` indexes the tiles of a.lb; a.partials[tile] receives the decoded sum.
template <typename CodeT, int COMB, bool SUM, int ROWS>
__device__ __forceinline__ void select_tile(const SelectArgs &a, uint32_t tile, SelectShared<ROWS> &sh) {
    constexpr int VEC = 16 / sizeof(CodeT);
    constexpr int CH = 4;
    constexpr int CHUNK = kBlock * VEC;
    constexpr int SUB = CHUNK * CH;  // rows per 16 KiB sub-tile
    static_assert(SUB <= 65536, "LDS row index must fit 16 bits");
    static_assert(SUB <= ROWS, "LDS row staging too small");
    uint64_t *s_scan = sh.scan;
    uint32_t *s_excl = sh.excl;
    uint32_t *s_mb = sh.mb;
    uint16_t *s_rows = sh.rows;
    const uint32_t iters = a.iters;
    const uint64_t base = static_cast<uint64_t>(tile) * iters * SUB;
    const CodeT *codes = static_cast<const CodeT *>(a.codes);
    const bool aligned = (reinterpret_cast<uintptr_t>(codes) & 15u) == 0;

    // phase 1: stream the tile once; keep one match bit per row (LDS) and which sub-tiles hold
    // an error (register bit mask)
    uint32_t cnt_m = 0, cnt_e = 0, err_subs = 0;
    uint64_t lsum = 0;
#pragma unroll 2
    for (uint32_t it = 0; it < iters; ++it) {
        uint32_t m, e;
        eval_subtile<CodeT, COMB, SUM>(a, codes, aligned, base + static_cast<uint64_t>(it) * SUB, m, e, lsum);
        s_mb[it * kBlock + threadIdx.x] = m;
        cnt_m += __popc(m);
        cnt_e += __popc(e);
        err_subs |= (e != 0u ? 1u : 0u) << it;
    }

    // tile aggregates -> look-back -> exclusive prefixes of both streams
    const uint32_t agg_m = block_sum<uint32_t>(cnt_m, reinterpret_cast<uint32_t *>(s_scan));
    const uint32_t agg_e = block_sum<uint32_t>(cnt_e, reinterpret_cast<uint32_t *>(s_scan));
    const uint32_t agg[2] = {agg_m, agg_e};
    lb_resolve<2>(a.lb, tile, agg, s_excl);
    uint64_t xm = s_excl[0], xe = s_excl[1];

    // phase 2: per sub-tile, rank the matching rows, stage them in LDS, write one coalesced run
    for (uint32_t it = 0; it < iters; ++it) {
        const uint64_t sb = base + static_cast<uint64_t>(it) * SUB;
        const uint32_t m = s_mb[it * kBlock + threadIdx.x];
        uint64_t tot_m = 0;
        const uint64_t ex_m = block_excl_scan<uint64_t>(pack_counts<CH, VEC>(m), tot_m, s_scan);
        for_each_ranked<CH, VEC>(m, ex_m, tot_m, [&](uint32_t r, int j, int k) {
            s_rows[r] = static_cast<uint16_t>(j * CHUNK + threadIdx.x * VEC + k);
        });
        if (agg_e != 0 && __syncthreads_or((err_subs >> it) & 1u)) {
            // rare: recompute this sub-tile's error bits and write its positions in order
            uint32_t mm, e;
            uint64_t dummy = 0;
            eval_subtile<CodeT, COMB, false>(a, codes, aligned, sb, mm, e, dummy);
            uint64_t tot_e = 0;
            const uint64_t ex_e = block_excl_scan<uint64_t>(pack_counts<CH, VEC>(e), tot_e, s_scan);
            for_each_ranked<CH, VEC>(e, ex_e, tot_e, [&](uint32_t r, int j, int k) {
                const uint64_t idx = xe + r;
                const uint64_t pos = sb + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC + k;
                if (idx < a.err_cap) a.err[idx] = pos * a.a_oid;
            });
            xe += fields_sum(tot_e);
        }
        __syncthreads();
        const uint32_t nm = fields_sum(tot_m);
        for (uint32_t k = threadIdx.x; k < nm; k += kBlock) {
            const uint64_t idx = xm + k;
            if (idx < a.oids_cap) {
                const uint32_t head = static_cast<uint32_t>(a.seqbase + sb + s_rows[k]);
                a.oids[idx] = static_cast<uint64_t>(head) * a.a_head;
            }
        }
        xm += nm;
    }
    if constexpr (SUM) {
        const uint64_t ts = block_sum<uint64_t>(lsum, s_scan);
        if (threadIdx.x == 0) a.partials[tile] = ts;
    }
    if (tile == a.lb.num_tiles - 1 && threadIdx.x == 0) {
        a.d_counts[0] = xm;
        a.d_counts[1] = xe;
    }
}


}  // namespace ahead_dev
