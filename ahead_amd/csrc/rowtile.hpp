// rowtile.hpp -- the generic tile engine of the per-operator kernels (fetch, positional match,
// hash-join probe, group-id assignment, grouped sum, product sum, arithmetic).
//
// A tile is G groups of 256 consecutive rows; row = base + g*256 + threadIdx.x, so every global
// load and store of a row group is one coalesced run.  Per row the operator's eval() returns the
// number of entries the row emits into the main output stream (0, 1, or more for a hash-join
// probe) and a bit set of the error streams it reports into.  Ranks are taken with wave ballots
// (0/1 rows) or wave scans (multi-entry rows) plus a block scan over the G x 4 wave totals, and
// the tile's place among all tiles comes from the decoupled look-back of an_device.hpp.  Phase 2
// calls emit_main() / emit_err() with each entry's final position, in row order.
//
// Op interface:
//   static constexpr int kErr;          // error streams (0..3)
//   static constexpr bool kMulti;       // rows may emit more than one main entry
//   static constexpr bool kMain;        // the operator has a main (compacted) output stream
//   using State = ...;                  // what phase 2 needs of a row (kept in registers)
//   uint64_t n;                         // rows
//   __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &acc) const;
//   __device__ void emit_main(uint64_t row, const State &st, uint64_t pos, uint32_t cnt) const;
//   __device__ void emit_err(int e, uint64_t row, const State &st, uint64_t pos) const;
//   __device__ void tile_done(uint32_t tile, uint64_t acc) const;   // per-tile reduction result
//   __device__ void finish(const uint64_t *totals) const;          // last tile: stream totals
#pragma once

#include <type_traits>
#include <utility>

#include "host_common.hpp"

namespace ahead_dev {

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    const uint32_t lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return __popcll(mask & below);
}

// An operator may evaluate all G rows of a thread at once (eval_n): its loads then go out as one
// batch per input column instead of one row after another.
template <class Op, class = void>
struct has_eval_n : std::false_type {};
template <class Op>
struct has_eval_n<Op, std::void_t<decltype(Op::kBatched)>> : std::true_type {};

// An operator may stand a whole launch down (skip(): true in every workgroup, decided from device
// state before any tile is taken).
template <class Op, class = void>
struct has_skip : std::false_type {};
template <class Op>
struct has_skip<Op, std::void_t<decltype(std::declval<const Op &>().skip())>> : std::true_type {};

template <class Op, int G>
__global__ __launch_bounds__(kBlock) void k_rowtile(Op op, LbArgs lb) {
    if constexpr (has_skip<Op>::value) {
        if (op.skip()) return;
    }
    constexpr int NS = 1 + Op::kErr;  // stream 0 = main
    constexpr int NW = G * kWaves;    // wave slots per tile, in row order
    static_assert(NW <= kBlock, "one thread per wave slot in the block scan");
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[NS][NW];
    __shared__ uint32_t s_excl[NS];
    __shared__ uint64_t s_red[kWaves];

    const uint32_t tile = take_tile(lb, &s_tile);
    if (tile >= lb.num_tiles) return;
    const uint64_t base = static_cast<uint64_t>(tile) * (G * kBlock);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;

    typename Op::State st[G];
    uint32_t cnt[G], errs[G];
    uint64_t acc = 0;
    // All G rows are evaluated before any wave-wide step, so the loads of different rows (and
    // their dependent gathers) are in flight together rather than one row after another.
    if constexpr (has_eval_n<Op>::value) {
        uint64_t rows[G];
        bool live[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            rows[g] = base + static_cast<uint64_t>(g) * kBlock + threadIdx.x;
            live[g] = rows[g] < op.n;
        }
        op.template eval_n<G>(rows, live, st, errs, cnt, acc);
    } else if (base + G * kBlock <= op.n) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            errs[g] = 0;
            cnt[g] = op.eval(base + static_cast<uint64_t>(g) * kBlock + threadIdx.x, st[g], errs[g], acc);
        }
    } else {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint64_t row = base + static_cast<uint64_t>(g) * kBlock + threadIdx.x;
            cnt[g] = 0;
            errs[g] = 0;
            if (row < op.n) cnt[g] = op.eval(row, st[g], errs[g], acc);
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        uint32_t wt = 0;
        if constexpr (Op::kMain) {
            if constexpr (Op::kMulti) wt = wave_sum(cnt[g]);
            else wt = __popcll(__ballot(cnt[g] != 0u));
        }
        if (lane == 0) s_cnt[0][g * kWaves + wave] = wt;
#pragma unroll
        for (int e = 0; e < Op::kErr; ++e) {
            const uint32_t we = __popcll(__ballot((errs[g] >> e) & 1u));
            if (lane == 0) s_cnt[1 + e][g * kWaves + wave] = we;
        }
    }
    __syncthreads();
    uint32_t agg[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const uint32_t v = threadIdx.x < NW ? s_cnt[s][threadIdx.x] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<uint32_t>(v, tot, reinterpret_cast<uint32_t *>(s_red));
        if (threadIdx.x < NW) s_cnt[s][threadIdx.x] = ex;
        agg[s] = tot;
    }
    lb_resolve<NS>(lb, tile, agg, s_excl);

#pragma unroll
    for (int g = 0; g < G; ++g) {
        const uint64_t row = base + static_cast<uint64_t>(g) * kBlock + threadIdx.x;
        if constexpr (Op::kMain) {
            uint32_t below;
            if constexpr (Op::kMulti) below = wave_incl_scan(cnt[g]) - cnt[g];
            else below = lanes_below(__ballot(cnt[g] != 0u));
            if (cnt[g] != 0u)
                op.emit_main(row, st[g], static_cast<uint64_t>(s_excl[0]) + s_cnt[0][g * kWaves + wave] + below, cnt[g]);
        }
#pragma unroll
        for (int e = 0; e < Op::kErr; ++e) {
            if (agg[1 + e] == 0u) continue;
            const bool on = (errs[g] >> e) & 1u;
            const uint32_t below = lanes_below(__ballot(on));
            if (on) op.emit_err(e, row, st[g], static_cast<uint64_t>(s_excl[1 + e]) + s_cnt[1 + e][g * kWaves + wave] + below);
        }
    }
    if constexpr (Op::kAcc) {
        const uint64_t t = block_sum<uint64_t>(acc, s_red);
        if (threadIdx.x == 0) op.tile_done(tile, t);
    }
    if (tile == lb.num_tiles - 1 && threadIdx.x == 0) {
        uint64_t totals[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) totals[s] = static_cast<uint64_t>(s_excl[s]) + agg[s];
        op.finish(totals);
    }
}

// element access by runtime width
__device__ __forceinline__ uint64_t load_w(const void *p, uint64_t i, uint32_t w) {
    switch (w) {
        case 1: return static_cast<const uint8_t *>(p)[i];
        case 2: return static_cast<const uint16_t *>(p)[i];
        case 4: return static_cast<const uint32_t *>(p)[i];
        default: return static_cast<const uint64_t *>(p)[i];
    }
}
__device__ __forceinline__ void store_w(void *p, uint64_t i, uint32_t w, uint64_t v) {
    switch (w) {
        case 1: static_cast<uint8_t *>(p)[i] = static_cast<uint8_t>(v); break;
        case 2: static_cast<uint16_t *>(p)[i] = static_cast<uint16_t>(v); break;
        case 4: static_cast<uint32_t *>(p)[i] = static_cast<uint32_t>(v); break;
        default: static_cast<uint64_t *>(p)[i] = v; break;
    }
}

// G elements at once, width chosen once: within each case the G loads are straight-line,
// unconditional code (a row with ok[g] false loads element 0 and yields 0), so they are all issued
// before the first is waited for.  The caller guarantees element 0 exists.
template <int G, typename T>
__device__ __forceinline__ void load_rows_t(const void *p, const uint64_t (&i)[G], const bool (&ok)[G],
                                            uint64_t (&v)[G]) {
    const T *q = static_cast<const T *>(p);
    T x[G];
#pragma unroll
    for (int g = 0; g < G; ++g) x[g] = q[ok[g] ? i[g] : 0ull];
#pragma unroll
    for (int g = 0; g < G; ++g) v[g] = ok[g] ? static_cast<uint64_t>(x[g]) : 0ull;
}
template <int G>
__device__ __forceinline__ void load_rows(const void *p, uint32_t w, const uint64_t (&i)[G], const bool (&ok)[G],
                                          uint64_t (&v)[G]) {
    switch (w) {
        case 1: load_rows_t<G, uint8_t>(p, i, ok, v); break;
        case 2: load_rows_t<G, uint16_t>(p, i, ok, v); break;
        case 4: load_rows_t<G, uint32_t>(p, i, ok, v); break;
        default: load_rows_t<G, uint64_t>(p, i, ok, v); break;
    }
}

// One BAT side on the device: raw element, decode (mulIfEncoded) and check.
struct DCol {
    const void *data;  // NULL: void column (seqbase + i as oid_t)
    uint64_t n;
    uint32_t width;
    uint32_t encoded;
    uint64_t a_inv;
    uint64_t max_u;
    uint64_t mask;     // code-width mask
    uint64_t seqbase;

    __device__ __forceinline__ uint64_t raw(uint64_t i) const {
        return data ? load_w(data, i, width) : static_cast<uint64_t>(static_cast<uint32_t>(seqbase + i));
    }
    // raw() of G rows at once (void columns compute seqbase + i)
    template <int G>
    __device__ __forceinline__ void raw_n(const uint64_t (&i)[G], const bool (&ok)[G], uint64_t (&v)[G]) const {
        if (data && n > 0) {
            load_rows<G>(data, width, i, ok, v);
        } else if (data) {  // empty column: no row can be in range
#pragma unroll
            for (int g = 0; g < G; ++g) v[g] = 0ull;
        } else {
#pragma unroll
            for (int g = 0; g < G; ++g) v[g] = static_cast<uint64_t>(static_cast<uint32_t>(seqbase + i[g]));
        }
    }
    __device__ __forceinline__ uint64_t dec_of(uint64_t x) const { return encoded ? (x * a_inv) & mask : x; }
    __device__ __forceinline__ bool bad_of(uint64_t x) const { return encoded && ((x * a_inv) & mask) > max_u; }
};

}  // namespace ahead_dev

namespace ahead_host {

ahead_dev::DCol dcol(const ahead_col &c);
// width of TypeMap<T>::v2_encoded_t::type_t for a column of the given kind
uint32_t encoded_width(const ahead_col &c);

// Launches one row-tile operator over n rows with `streams` look-back streams (1 + Op::kErr).
// It uses no scratch region of the workspace: an operator's own regions stay intact.
template <class Op, int G = 8>
int launch_rowtile(const Op &op, uint64_t n, int streams, ahead_hip_ws *ws, hipStream_t s,
                   uint32_t *num_tiles = nullptr) {
    const LbPlan plan = lb_plan(n, G * ahead_dev::kBlock, streams);
    ahead_dev::LbArgs lb{};
    int st;
    if ((st = lb_args(ws, plan, &lb)) != AHEAD_OK) return st;
    hipLaunchKernelGGL((ahead_dev::k_rowtile<Op, G>), dim3(plan.num_tiles), dim3(ahead_dev::kBlock), 0, s, op, lb);
    AHEAD_HIP_TRY(hipGetLastError());
    if (num_tiles) *num_tiles = plan.num_tiles;
    return AHEAD_OK;
}

inline size_t rowtile_status_bytes(uint64_t n, int streams, int G = 8) {
    return lb_plan(n, G * ahead_dev::kBlock, streams).status_bytes;
}

}  // namespace ahead_host
