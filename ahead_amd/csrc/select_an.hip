// select_an.hip -- selectAN on AN-encoded codes for gfx950, with the fused
// check-and-decode + sum of the select+sum workload.
//
// Semantics (lib/column_operators/AN/selectAN_scalar.tcc:66-95 and :149-181): for every row
// `pos` of the encoded tail, in order,
//   (a) if (code * A^-1 mod 2^k) > maxU       -> err  += pos * A_OID
//   (b) if Op(code, th) [Comb(Op1, Op2)]      -> oids += (uint32)(seqbase + pos) * A_head
// with unsigned comparisons on the raw codes against thresholds the caller already encoded.
// Corrupt rows still take part in (b).
//
// Kernel shape: one 256-thread workgroup per tile of ITERS sub-tiles of 16 KiB (4 x 16-byte
// loads per thread; 8192 rows of 16-bit codes, 4096 of 32-bit, 2048 of 64-bit).
//  phase 1  streams the tile once with coalesced 128-bit non-temporal loads and keeps only one
//           match bit per row (LDS) and one "has an error" bit per sub-tile;
//  resolve  publishes the tile's counts and takes its prefix by decoupled look-back
//           (an_device.hpp) -- one ticket and one look-back per ITERS sub-tiles;
//  phase 2  per sub-tile, packs the per-chunk counts into one uint64 (16-bit fields) so that one
//           block scan ranks every row in order, stages matching row indices in LDS and writes
//           them as one contiguous, coalesced run of 64-bit OIDs.  OIDs depend only on the row
//           index, so the input is never read twice.  Sub-tiles holding an error (rare) are
//           re-read to rank their error positions, which are scattered.
// The fused decoded sum leaves per-tile partials that k_sum_partials adds up; the kernel
// boundary orders them without any device-wide fence.
#include "host_common.hpp"

using namespace ahead_dev;

namespace {

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

template <typename CodeT, int COMB, bool SUM>
__global__ __launch_bounds__(kBlock) void k_select(SelectArgs a) {
    constexpr int VEC = 16 / sizeof(CodeT);
    constexpr int CH = 4;
    constexpr int CHUNK = kBlock * VEC;
    constexpr int SUB = CHUNK * CH;  // rows per 16 KiB sub-tile
    static_assert(SUB <= 65536, "LDS row index must fit 16 bits");

    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_scan[kWaves];
    __shared__ uint32_t s_excl[2];
    __shared__ uint32_t s_mb[kMaxSubtiles * kBlock];  // match bits of every row of the tile
    __shared__ uint16_t s_rows[SUB];

    const uint32_t tile = take_tile(a.lb, &s_tile);
    if (tile >= a.lb.num_tiles) return;
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

// Adds the per-tile partial sums of a k_select launch (kernel order makes them visible).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const uint64_t *partials, uint32_t n, uint64_t *out) {
    __shared__ uint64_t s_scan[kWaves];
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) s += partials[i];
    s = block_sum<uint64_t>(s, s_scan);
    if (threadIdx.x == 0) *out = s;
}

// canonical interval form of one predicate in a code width with mask m
struct Canon {
    uint64_t lo = 0, span = 0;
    uint32_t neg = 0, valid = 1;
};
bool canon(int op, uint64_t th, uint64_t m, Canon *c) {
    th &= m;
    switch (op) {
        case AHEAD_LT:
            if (th == 0) c->valid = 0; else { c->lo = 0; c->span = th - 1; }
            return true;
        case AHEAD_LE: c->lo = 0; c->span = th; return true;
        case AHEAD_GT:
            if (th == m) c->valid = 0; else { c->lo = th + 1; c->span = m - th - 1; }
            return true;
        case AHEAD_GE: c->lo = th; c->span = m - th; return true;
        case AHEAD_EQ: c->lo = th; c->span = 0; return true;
        case AHEAD_NE: c->lo = th; c->span = 0; c->neg = 1; return true;
        default: return false;
    }
}

template <typename CodeT, int COMB, bool SUM>
void launch(const SelectArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_select<CodeT, COMB, SUM>), dim3(a.lb.num_tiles), dim3(kBlock), 0, s, a);
}

template <typename CodeT>
void dispatch(const SelectArgs &a, int comb, bool sum, hipStream_t s) {
    if (comb == AHEAD_COMB_NONE) {
        sum ? launch<CodeT, AHEAD_COMB_NONE, true>(a, s) : launch<CodeT, AHEAD_COMB_NONE, false>(a, s);
    } else if (comb == AHEAD_COMB_AND) {
        sum ? launch<CodeT, AHEAD_COMB_AND, true>(a, s) : launch<CodeT, AHEAD_COMB_AND, false>(a, s);
    } else {
        sum ? launch<CodeT, AHEAD_COMB_OR, true>(a, s) : launch<CodeT, AHEAD_COMB_OR, false>(a, s);
    }
}

}  // namespace

using namespace ahead_host;

extern "C" int ahead_hip_select_an(const ahead_col *col, int op1, uint64_t th1, int op2, uint64_t th2,
                                   int comb, uint64_t a_head, uint64_t a_oid, uint64_t *oids,
                                   uint64_t oids_cap, uint64_t *err, uint64_t err_cap,
                                   uint64_t *d_counts, uint64_t *d_sum, ahead_hip_ws *ws,
                                   void *stream) {
    int st = check_col(col, "select.col", false, true);
    if (st != AHEAD_OK) return st;
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "select: d_counts is NULL");
    if ((oids == nullptr && oids_cap) || (err == nullptr && err_cap))
        return fail(AHEAD_ERR_INVALID, "select: NULL output with nonzero capacity");
    if (comb < AHEAD_COMB_NONE || comb > AHEAD_COMB_OR) return fail(AHEAD_ERR_INVALID, "select: bad combiner");
    const uint64_t m = width_mask(col->width);
    Canon c1, c2;
    if (!canon(op1, th1, m, &c1)) return fail(AHEAD_ERR_INVALID, "select: bad op1");
    if (comb != AHEAD_COMB_NONE && !canon(op2, th2, m, &c2)) return fail(AHEAD_ERR_INVALID, "select: bad op2");
    // Sub-tiles of 16 KiB; one look-back ticket covers `iters` of them.  Fewer tickets keep the
    // contended ticket atomic and the look-back off the critical path; iters stays small enough
    // that the grid still holds >= min_tiles tiles (default 4096, 16 per CU) when it can.
    const uint64_t sub_rows = 4ull * kBlock * (16u / col->width);
    const uint64_t subs = (col->n + sub_rows - 1) / sub_rows;
    const uint32_t iters = subtiles_per_tile(subs);
    const LbPlan plan = lb_plan(col->n, static_cast<uint32_t>(sub_rows * iters), 2);
    const size_t partial_off = plan.status_bytes;
    const size_t need = partial_off + align_up(static_cast<size_t>(plan.num_tiles) * 8, 256);
    if ((st = ws_ensure(ws, need)) != AHEAD_OK) return st;

    SelectArgs a{};
    a.codes = col->data;
    a.n = col->n;
    a.seqbase = col->seqbase;
    a.a_inv = col->a_inv;
    a.max_u = col->max_u;
    a.lo1 = c1.lo; a.span1 = c1.span; a.neg1 = c1.neg; a.valid1 = c1.valid;
    a.lo2 = c2.lo; a.span2 = c2.span; a.neg2 = c2.neg; a.valid2 = c2.valid;
    a.a_head = a_head;
    a.a_oid = a_oid;
    a.oids = oids;
    a.oids_cap = oids == nullptr ? 0 : oids_cap;
    a.err = err;
    a.err_cap = err == nullptr ? 0 : err_cap;
    a.d_counts = d_counts;
    a.partials = reinterpret_cast<uint64_t *>(ws_region(ws, partial_off));
    a.iters = iters;
    if ((st = lb_args(ws, plan, 0, &a.lb)) != AHEAD_OK) return st;

    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool sum = d_sum != nullptr;
    switch (col->width) {
        case 2: dispatch<uint16_t>(a, comb, sum, s); break;
        case 4: dispatch<uint32_t>(a, comb, sum, s); break;
        default: dispatch<uint64_t>(a, comb, sum, s); break;
    }
    AHEAD_HIP_TRY(hipGetLastError());
    if (sum) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, a.partials, plan.num_tiles, d_sum);
        AHEAD_HIP_TRY(hipGetLastError());
    }
    return AHEAD_OK;
}
