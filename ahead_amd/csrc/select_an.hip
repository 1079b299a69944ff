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
// Kernel shape: one 256-thread workgroup per 16 KiB tile, 4 x 16-byte loads per thread
// (8192 rows of 16-bit codes, 4096 of 32-bit, 2048 of 64-bit).  Each thread loads its 4 chunks
// with coalesced 128-bit loads, evaluates the predicate and the check per row, and packs its
// per-chunk match counts into one uint64 (16-bit fields) so that a single block scan yields the
// in-order rank of every row.  The tile's prefix over all earlier tiles comes from the
// decoupled look-back (an_device.hpp).  The fused decoded sum is reduced by the last tile to
// finish (last_tile_sum), so one call is one kernel launch.  Matching row indices are staged in LDS by rank and
// written out as one contiguous, coalesced run of 64-bit OIDs; error positions (rare) are
// scattered straight to HBM.
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
    uint64_t *d_sum;
    LbArgs lb;
};

template <typename CodeT>
__device__ __forceinline__ bool in_range(CodeT x, CodeT lo, CodeT span, uint32_t valid, uint32_t neg) {
    const bool in = valid && static_cast<CodeT>(x - lo) <= span;
    return in != static_cast<bool>(neg);
}

template <typename CodeT, int COMB, bool SUM>
__global__ __launch_bounds__(kBlock) void k_select(SelectArgs a) {
    constexpr int VEC = 16 / sizeof(CodeT);
    constexpr int CH = 4;
    constexpr int CHUNK = kBlock * VEC;
    constexpr int TILE = CHUNK * CH;
    static_assert(TILE <= 65536, "LDS row index must fit 16 bits");

    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_scan[kWaves];
    __shared__ uint32_t s_excl[2];
    __shared__ uint16_t s_rows[TILE];

    const uint32_t tile = take_tile(a.lb, &s_tile);
    if (tile >= a.lb.num_tiles) return;
    const uint64_t base = static_cast<uint64_t>(tile) * TILE;
    const CodeT *codes = static_cast<const CodeT *>(a.codes);
    const bool aligned = (reinterpret_cast<uintptr_t>(codes) & 15u) == 0;

    CodeT v[CH][VEC];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const uint64_t r0 = base + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC;
        if (aligned && r0 + VEC <= a.n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(codes + r0);
            __builtin_memcpy(&v[j][0], &q, 16);
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) v[j][k] = (r0 + k < a.n) ? codes[r0 + k] : CodeT(0);
        }
    }

    const CodeT ainv = static_cast<CodeT>(a.a_inv);
    const CodeT maxu = static_cast<CodeT>(a.max_u);
    const CodeT lo1 = static_cast<CodeT>(a.lo1), sp1 = static_cast<CodeT>(a.span1);
    const CodeT lo2 = static_cast<CodeT>(a.lo2), sp2 = static_cast<CodeT>(a.span2);

    uint32_t mbits = 0, ebits = 0;
    uint64_t lsum = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const uint64_t row = base + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC + k;
            const bool live = row < a.n;
            const CodeT x = v[j][k];
            const CodeT d = an_mul<CodeT>(x, ainv);
            const bool bad = live && d > maxu;
            bool hit = in_range<CodeT>(x, lo1, sp1, a.valid1, a.neg1);
            if constexpr (COMB == AHEAD_COMB_AND) hit = hit && in_range<CodeT>(x, lo2, sp2, a.valid2, a.neg2);
            if constexpr (COMB == AHEAD_COMB_OR) hit = hit || in_range<CodeT>(x, lo2, sp2, a.valid2, a.neg2);
            hit = hit && live;
            mbits |= static_cast<uint32_t>(hit) << (j * VEC + k);
            ebits |= static_cast<uint32_t>(bad) << (j * VEC + k);
            if constexpr (SUM) lsum += (hit && !bad) ? static_cast<uint64_t>(d) : 0ull;
        }
    }

    const uint64_t pm = pack_counts<CH, VEC>(mbits), pe = pack_counts<CH, VEC>(ebits);
    uint64_t tot_m = 0, tot_e = 0, ex_e = 0;
    const uint64_t ex_m = block_excl_scan<uint64_t>(pm, tot_m, s_scan);
    const int any_err = __syncthreads_or(ebits != 0u);
    if (any_err) ex_e = block_excl_scan<uint64_t>(pe, tot_e, s_scan);

    const uint32_t agg[2] = {fields_sum(tot_m), fields_sum(tot_e)};
    lb_resolve<2>(a.lb, tile, agg, s_excl);
    const uint64_t xm = s_excl[0], xe = s_excl[1];

    // stage matching rows in LDS at their in-tile rank; error positions go straight to HBM
    for_each_ranked<CH, VEC>(mbits, ex_m, tot_m, [&](uint32_t r, int j, int k) {
        s_rows[r] = static_cast<uint16_t>(j * CHUNK + threadIdx.x * VEC + k);
    });
    if (any_err) {
        for_each_ranked<CH, VEC>(ebits, ex_e, tot_e, [&](uint32_t r, int j, int k) {
            const uint64_t idx = xe + r;
            const uint64_t pos = base + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC + k;
            if (idx < a.err_cap) a.err[idx] = pos * a.a_oid;
        });
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < agg[0]; k += kBlock) {
        const uint64_t idx = xm + k;
        if (idx < a.oids_cap) {
            const uint32_t head = static_cast<uint32_t>(a.seqbase + base + s_rows[k]);
            a.oids[idx] = static_cast<uint64_t>(head) * a.a_head;
        }
    }
    if constexpr (SUM) {
        __shared__ uint32_t s_last;
        const uint64_t ts = block_sum<uint64_t>(lsum, s_scan);
        last_tile_sum(a.lb.hdr, a.partials, a.lb.num_tiles, tile, ts, a.d_sum, s_scan, &s_last);
    }
    if (tile == a.lb.num_tiles - 1 && threadIdx.x == 0) {
        a.d_counts[0] = xm + agg[0];
        a.d_counts[1] = xe + agg[1];
    }
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
    const uint32_t tile_rows = 4u * kBlock * (16u / col->width);
    const LbPlan plan = lb_plan(col->n, tile_rows, 2);
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
    a.d_sum = d_sum;
    if ((st = lb_args(ws, plan, 0, &a.lb)) != AHEAD_OK) return st;

    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool sum = d_sum != nullptr;
    switch (col->width) {
        case 2: dispatch<uint16_t>(a, comb, sum, s); break;
        case 4: dispatch<uint32_t>(a, comb, sum, s); break;
        default: dispatch<uint64_t>(a, comb, sum, s); break;
    }
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
