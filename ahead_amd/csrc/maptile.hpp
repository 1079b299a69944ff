// maptile.hpp -- the streaming engine of the per-operator kernels whose main output is 1:1 with
// their input (a row emits one entry or is dropped) or absent: fetch, positional and seqbase
// match joins, check-and-decode of pairs, arithmetic, product sum, the grouped sum's checks.
//
// The row-tile engine (rowtile.hpp) orders its outputs with a decoupled look-back per 2048-row
// tile; at tens of millions of rows the chain of tile prefixes, not HBM, sets its pace.  This
// engine has no chain.  A persistent grid gives every wave one contiguous range of rows:
//
//  k_map      each wave streams its range (all G rows of a lane evaluated before any wave-wide
//             step, so their loads are in flight together), writes every kept row's entry at the
//             wave's range start + its rank in the wave ("staged": final when no earlier wave
//             dropped a row), counts its rows per stream, records its first kErrList error rows
//             in row order and leaves its accumulator partial (kAcc);
//  k_map_scan one workgroup: exclusive prefixes over the waves of every count, the totals
//             (Op::finish) and the number of dropped rows before every wave;
//  k_map_save waves that will move (rows dropped before them, at most kMaxShift in all) copy the
//             staged entries their right neighbour will overwrite into a side buffer;
//  k_map_fix  per wave: the error positions (from the recorded rows, or by re-evaluating the range
//             when they did not fit), and the move of the staged entries to their final place
//             (from the side buffer where a neighbour overwrote them).  When more than kMaxShift
//             rows were dropped the waves instead copy their staged entries to their final places
//             in a scratch array (no staged entry is overwritten before it is read) ...
//  k_map_back ... which this launch copies back over the main outputs (only then).
//
// Op interface: that of rowtile.hpp (kMulti must be false), plus, for kMain operators,
//   static constexpr int kOuts;                                  // main output arrays (1 or 2)
//   __device__ void out_arrays(void *(&p)[2], uint32_t (&w)[2]) const;
// The host guarantees that every main output array holds n entries.
#pragma once

#include <type_traits>

#include "rowtile.hpp"

namespace ahead_dev {

constexpr uint32_t kMapErrList = 16;  // error rows a wave records in row order
constexpr uint32_t kMaxShift = 64;    // dropped rows the move path handles (else: compaction through scratch)
constexpr int kMapG = 4;              // rows per lane per iteration (unless the operator sets kMapRows)

struct MapArgs {
    uint64_t rows_per_wave;
    uint32_t waves;
    uint32_t *wcnt;    // [NS + 1][waves]: main entries, error streams, live rows
    uint32_t *wpre;    // same shape: exclusive prefixes over the waves
    uint32_t *werows;  // [waves] rows with any error
    uint64_t *elist;   // [waves][kMapErrList] error bits << 32 | row
    uint64_t *side;    // [waves][kMaxShift] staged entries saved before a move
    uint32_t *flags;   // [0] dropped rows in all (k_map_scan)
    uint64_t *totals;  // [NS + 1] per stream (k_map_scan)
    uint8_t *scratch[2];  // kMain: per main output array, n entries of 8 bytes (the compaction)
};

template <class Op>
struct MapTraits {
    static constexpr int NS = 1 + Op::kErr;
};

__device__ __forceinline__ uint32_t map_below(uint64_t ballot) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(ballot >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ballot), 0u));
}

// An operator may bring a read-only table to LDS (kFilterWords 64-bit words): filter_init copies it
// at the start of k_map and k_map_fix, and eval_f (eval_n with the table) evaluates the rows.
template <class Op, class = void>
struct has_filter : std::false_type {};
template <class Op>
struct has_filter<Op, std::void_t<decltype(Op::kFilterWords)>> : std::true_type {};
template <class Op, bool = has_filter<Op>::value>
struct FilterWords {
    static constexpr int value = 1;
};
template <class Op>
struct FilterWords<Op, true> {
    static constexpr int value = Op::kFilterWords;
};

// Rows per lane and iteration: kMapG, or the operator's kMapRows (an operator whose rows bring
// few bytes keeps more loads in flight with more rows).
template <class Op, class = void>
struct MapG {
    static constexpr int value = kMapG;
};
template <class Op>
struct MapG<Op, std::void_t<decltype(Op::kMapRows)>> {
    static constexpr int value = Op::kMapRows;
};

template <class Op, int G>
__device__ __forceinline__ void map_eval(const Op &op, const uint64_t (&rows)[G], const bool (&live)[G],
                                         typename Op::State (&st)[G], uint32_t (&errs)[G],
                                         uint32_t (&cnt)[G], uint64_t &acc, const uint64_t *filt = nullptr) {
    if constexpr (has_filter<Op>::value) {
        op.template eval_f<G>(rows, live, st, errs, cnt, acc, filt);
    } else if constexpr (has_eval_n<Op>::value) {
        op.template eval_n<G>(rows, live, st, errs, cnt, acc);
    } else {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            cnt[g] = 0;
            errs[g] = 0;
            if (live[g]) cnt[g] = op.eval(rows[g], st[g], errs[g], acc);
        }
    }
}

// An operator may keep per-workgroup accumulators in LDS (kSharedWords 64-bit words):
// shared_init / shared_flush bracket the scan, and eval_shared replaces eval_n.
template <class Op, class = void>
struct has_shared : std::false_type {};
template <class Op>
struct has_shared<Op, std::void_t<decltype(Op::kSharedWords)>> : std::true_type {};
template <class Op, bool = has_shared<Op>::value>
struct SharedWords {
    static constexpr int value = 1;
};
template <class Op>
struct SharedWords<Op, true> {
    static constexpr int value = Op::kSharedWords;
};

// Rows of one wave's range; lane l handles rows seg + g * 64 + l (one coalesced run per g).
template <class Op>
__global__ __launch_bounds__(kBlock) void k_map(Op op, MapArgs m) {
    static_assert(!Op::kMulti, "the map engine takes operators whose rows emit at most one entry");
    constexpr int NS = MapTraits<Op>::NS;
    constexpr int G = MapG<Op>::value;
    __shared__ unsigned long long s_acc[SharedWords<Op>::value];
    __shared__ uint64_t s_filt[FilterWords<Op>::value];
    if constexpr (has_skip<Op>::value) {
        if (op.skip()) return;
    }
    if constexpr (has_shared<Op>::value) {
        op.shared_init(s_acc);
        __syncthreads();
    }
    if constexpr (has_filter<Op>::value) {
        op.filter_init(s_filt);
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint64_t begin = static_cast<uint64_t>(gw) * m.rows_per_wave;
    const uint64_t end = begin + m.rows_per_wave < op.n ? begin + m.rows_per_wave : op.n;
    uint32_t kept = 0, err_rows = 0;  // wave-uniform
    uint32_t ecnt[Op::kErr > 0 ? Op::kErr : 1] = {};
    uint64_t acc = 0;
    for (uint64_t seg = begin; seg < end; seg += 64u * G) {
        uint64_t rows[G];
        bool live[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            rows[g] = seg + static_cast<uint64_t>(g) * 64u + lane;
            live[g] = rows[g] < end;
        }
        typename Op::State st[G];
        uint32_t errs[G], cnt[G];
        if constexpr (has_shared<Op>::value) {
            op.template eval_shared<G>(rows, live, st, errs, cnt, acc, s_acc);
        } else {
            map_eval(op, rows, live, st, errs, cnt, acc, s_filt);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if constexpr (Op::kMain) {
                const uint64_t b = __ballot(cnt[g] != 0u);
                if (cnt[g] != 0u) op.emit_main(rows[g], st[g], begin + kept + map_below(b), 1u);
                kept += static_cast<uint32_t>(__popcll(b));
            }
#pragma unroll
            for (int e = 0; e < Op::kErr; ++e) ecnt[e] += (errs[g] >> e) & 1u;
            const uint64_t x = __ballot(errs[g] != 0u);
            if (x != 0ull) {  // rare: recorded in row order while the list has room
                if (errs[g] != 0u) {
                    const uint32_t i = err_rows + map_below(x);
                    if (i < kMapErrList)
                        m.elist[static_cast<size_t>(gw) * kMapErrList + i] =
                            (static_cast<uint64_t>(errs[g]) << 32) | static_cast<uint32_t>(rows[g]);
                }
                err_rows += static_cast<uint32_t>(__popcll(x));
            }
        }
    }
    if constexpr (has_shared<Op>::value) {
        __syncthreads();
        op.shared_flush(s_acc);
    }
#pragma unroll
    for (int e = 0; e < Op::kErr; ++e) ecnt[e] = wave_sum(ecnt[e]);
    if constexpr (Op::kAcc) {
        const uint64_t t = wave_sum(acc);
        if (lane == 0) op.tile_done(gw, t);
    }
    if (lane == 0) {
        m.wcnt[gw] = kept;
#pragma unroll
        for (int e = 0; e < Op::kErr; ++e) m.wcnt[static_cast<size_t>(1 + e) * m.waves + gw] = ecnt[e];
        m.wcnt[static_cast<size_t>(NS) * m.waves + gw] = end > begin ? static_cast<uint32_t>(end - begin) : 0u;
        m.werows[gw] = err_rows;
    }
}

// One workgroup of 1024 threads per count stream (NS + 1 of them): exclusive prefixes over the
// waves and the stream total.
template <class Op>
__global__ __launch_bounds__(1024) void k_map_scan(Op op, MapArgs m) {
    __shared__ uint32_t s_w[16];
    if constexpr (has_skip<Op>::value) {
        if (op.skip()) return;
    }
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const size_t s = blockIdx.x;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m.waves; base += 8192u) {
        uint32_t v[8], t = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t i = base + threadIdx.x * 8u + q;
            v[q] = i < m.waves ? m.wcnt[s * m.waves + i] : 0u;
            t += v[q];
        }
        const uint32_t inc = wave_incl_scan(t);
        if (lane == 63u) s_w[wv] = inc;
        __syncthreads();
        uint32_t pre = carry, tot = 0;
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            if (w < static_cast<int>(wv)) pre += s_w[w];
            tot += s_w[w];
        }
        __syncthreads();
        pre += inc - t;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t i = base + threadIdx.x * 8u + q;
            if (i < m.waves) m.wpre[s * m.waves + i] = pre;
            pre += v[q];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) m.totals[s] = carry;
}

// Dropped rows in all (kMain operators): live rows minus main entries.
template <class Op>
__device__ __forceinline__ uint32_t map_total_drops(const MapArgs &m) {
    constexpr int NS = MapTraits<Op>::NS;
    return Op::kMain ? static_cast<uint32_t>(m.totals[NS] - m.totals[0]) : 0u;
}

// Rows dropped before wave gw (kMain operators).
template <class Op>
__device__ __forceinline__ uint32_t map_drops_before(const MapArgs &m, uint32_t gw) {
    constexpr int NS = MapTraits<Op>::NS;
    return m.wpre[static_cast<size_t>(NS) * m.waves + gw] - m.wpre[gw];
}

__device__ __forceinline__ uint64_t map_load_entry(const void *p, uint32_t w, uint64_t i) { return load_w(p, i, w); }

// Waves that will move save the staged entries their right neighbour's move overwrites: the last
// min(D, kept) of their own (D = rows dropped before the wave; the neighbour's entries land at
// most D slots left of its range).  side[(gw * 2 + array) * kMaxShift + k] holds entry k of them.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_map_save(Op op, MapArgs m) {
    if constexpr (has_skip<Op>::value) {
        if (op.skip()) return;
    }
    if constexpr (Op::kMain) {
        const uint32_t total = map_total_drops<Op>(m);
        if (total == 0u || total > kMaxShift) return;
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
        if (gw >= m.waves) return;
        const uint32_t d = map_drops_before<Op>(m, gw);
        const uint32_t kept = m.wcnt[gw];
        const uint32_t keep = d < kept ? d : kept;
        if (keep == 0u) return;
        void *p[2];
        uint32_t w[2];
        op.out_arrays(p, w);
        const uint64_t begin = static_cast<uint64_t>(gw) * m.rows_per_wave;
        // keep <= kMaxShift = 64 entries: one per lane
        if (lane < keep) {
            const uint64_t i = begin + kept - keep + lane;
#pragma unroll
            for (int a = 0; a < Op::kOuts; ++a)
                m.side[(static_cast<size_t>(gw) * 2 + a) * kMaxShift + lane] = map_load_entry(p[a], w[a], i);
        }
    }
}

// Error positions, then the move (or the compaction) of the staged entries.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_map_fix(Op op, MapArgs m) {
    constexpr int NS = MapTraits<Op>::NS;
    if constexpr (has_skip<Op>::value) {
        if (op.skip()) return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (gw >= m.waves) return;
    const uint64_t begin = static_cast<uint64_t>(gw) * m.rows_per_wave;
    const uint64_t end = begin + m.rows_per_wave < op.n ? begin + m.rows_per_wave : op.n;
    const uint32_t total_drops = map_total_drops<Op>(m);
    const bool compact = total_drops > kMaxShift;
    if (gw == 0 && lane == 0) op.finish(m.totals);

    // error positions
    const uint32_t nrows = m.werows[gw];
    if (nrows != 0u && Op::kErr > 0) {
        uint64_t idx[Op::kErr > 0 ? Op::kErr : 1];
#pragma unroll
        for (int e = 0; e < Op::kErr; ++e) idx[e] = m.wpre[static_cast<size_t>(1 + e) * m.waves + gw];
        const bool listed = nrows <= kMapErrList;
        const uint64_t span = listed ? nrows : end - begin;
        for (uint64_t j0 = 0; j0 < span; j0 += 64u) {
            const uint64_t j = j0 + lane;
            uint32_t errs = 0;
            uint64_t row = 0, acc = 0;
            typename Op::State st{};
            if (j < span) {
                if (listed) {
                    const uint64_t ent = m.elist[static_cast<size_t>(gw) * kMapErrList + j];
                    row = static_cast<uint32_t>(ent);
                    (void)op.eval(row, st, errs, acc);  // the row's state for emit_err
                } else {
                    row = begin + j;
                    (void)op.eval(row, st, errs, acc);
                }
            }
#pragma unroll
            for (int e = 0; e < Op::kErr; ++e) {
                const bool on = (errs >> e) & 1u;
                const uint64_t b = __ballot(on);
                if (on) op.emit_err(e, row, st, idx[e] + map_below(b));
                idx[e] += static_cast<uint64_t>(__popcll(b));
            }
        }
    }

    if constexpr (Op::kMain) {
        if (total_drops == 0u) return;
        const uint32_t d = map_drops_before<Op>(m, gw);
        const uint64_t first = m.wpre[gw];  // final position of the wave's first entry
        if (compact) {
            // too many drops to move in place: the staged entries [begin, begin + kept) go to
            // scratch[first, first + kept); k_map_back copies the scratch over the outputs
            const uint32_t kept = m.wcnt[gw];
            void *p[2];
            uint32_t w[2];
            op.out_arrays(p, w);
            constexpr int kPer = 8;
            for (uint32_t k0 = 0; k0 < kept; k0 += 64u * kPer) {
#pragma unroll
                for (int a = 0; a < Op::kOuts; ++a) {
                    uint64_t idx[kPer], v[kPer];
                    bool in[kPer];
#pragma unroll
                    for (int q = 0; q < kPer; ++q) {
                        const uint32_t k = k0 + q * 64u + lane;
                        in[q] = k < kept;
                        idx[q] = begin + k;
                    }
                    load_rows<kPer>(p[a], w[a], idx, in, v);
#pragma unroll
                    for (int q = 0; q < kPer; ++q) {
                        if (in[q]) store_w(m.scratch[a], first + k0 + q * 64u + lane, w[a], v[q]);
                    }
                }
            }
            return;
        }
        if (d == 0u) return;
        // move: staged [begin, begin + kept) -> [first, first + kept), ascending, 64 entries per step;
        // entry k >= kept - keep comes from the side buffer (the right neighbour overwrote it)
        const uint32_t kept = m.wcnt[gw];
        const uint32_t keep = d < kept ? d : kept;
        void *p[2];
        uint32_t w[2];
        op.out_arrays(p, w);
        // 512 entries per step (8 per lane): all are read before any is written; a step's writes
        // land on entries of this step or earlier ones only (they move left)
        constexpr int kPer = 8;
        for (uint32_t k0 = 0; k0 < kept; k0 += 64u * kPer) {
            uint64_t v[2][kPer];
#pragma unroll
            for (int a = 0; a < Op::kOuts; ++a) {
                uint64_t idx[kPer];
                bool from_stage[kPer];
#pragma unroll
                for (int q = 0; q < kPer; ++q) {
                    const uint32_t k = k0 + q * 64u + lane;
                    from_stage[q] = k < kept - keep;
                    idx[q] = begin + k;
                }
                load_rows<kPer>(p[a], w[a], idx, from_stage, v[a]);
#pragma unroll
                for (int q = 0; q < kPer; ++q) {
                    const uint32_t k = k0 + q * 64u + lane;
                    if (k < kept && !from_stage[q])
                        v[a][q] = m.side[(static_cast<size_t>(gw) * 2 + a) * kMaxShift + (k - (kept - keep))];
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const uint32_t k = k0 + q * 64u + lane;
                if (k < kept) {
#pragma unroll
                    for (int a = 0; a < Op::kOuts; ++a) store_w(p[a], first + k, w[a], v[a][q]);
                }
            }
        }
    }
}

// After a compaction (more than kMaxShift rows dropped): the main entries [0, kept) from the
// scratch arrays back over the outputs.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_map_back(Op op, MapArgs m) {
    if constexpr (has_skip<Op>::value) {
        if (op.skip()) return;
    }
    if constexpr (Op::kMain) {
        if (map_total_drops<Op>(m) <= kMaxShift) return;
        const uint64_t kept = m.totals[0];
        void *p[2];
        uint32_t w[2];
        op.out_arrays(p, w);
        const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
        const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
#pragma unroll
        for (int a = 0; a < Op::kOuts; ++a) {
            // the first kept * w bytes: 16 per thread and step where the output is 16-byte aligned
            // (the scratch always is), the tail and unaligned outputs entry by entry
            const uint64_t bytes = kept * w[a];
            uint64_t done = 0;
            if ((reinterpret_cast<uintptr_t>(p[a]) & 15u) == 0u) {
                const uint64_t n16 = bytes / 16u;
                const uint4 *src = reinterpret_cast<const uint4 *>(m.scratch[a]);
                uint4 *dst = reinterpret_cast<uint4 *>(p[a]);
                for (uint64_t i = t0; i < n16; i += stride) dst[i] = src[i];
                done = n16 * 16u / w[a];
            }
            for (uint64_t i = done + t0; i < kept; i += stride) store_w(p[a], i, w[a], load_w(m.scratch[a], i, w[a]));
        }
    }
}

}  // namespace ahead_dev

namespace ahead_host {

// Main output arrays of an operator (0 without a main output).
template <class Op>
constexpr int map_outs() {
    if constexpr (Op::kMain) {
        return Op::kOuts;
    } else {
        return 0;
    }
}

// Persistent grid of the map engine: as many workgroups as can be resident (registers decide),
// at most one wave per 64 rows.
template <class Op>
uint32_t map_grid(uint64_t n) {
    static uint32_t resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(ahead_dev::k_map<Op>), ahead_dev::kBlock, 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 2;
        resident = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
    }
    const uint64_t by_rows = (n + 64u * ahead_dev::kWaves - 1) / (64u * ahead_dev::kWaves);
    const uint64_t g = by_rows < resident ? by_rows : resident;
    return static_cast<uint32_t>(g == 0 ? 1 : g);
}

// Workspace bytes of the map engine's per-wave tables for n rows.
template <class Op>
size_t map_scratch_bytes(uint64_t n) {
    using namespace ahead_dev;
    constexpr int NS = MapTraits<Op>::NS;
    const size_t waves = static_cast<size_t>(map_grid<Op>(n)) * kWaves;
    const size_t compact = map_outs<Op>() * align_up(static_cast<size_t>(n) * 8, 256);
    return 2 * align_up((NS + 1) * waves * 4, 256) + align_up(waves * 4, 256) + align_up(waves * kMapErrList * 8, 256) +
           align_up(waves * 2 * kMaxShift * 8, 256) + 256 + compact;
}

// Launches the map engine for op over op.n rows.  Scratch: per-wave tables in the workspace region
// at `offset`.  A caller that keeps regions of its own before `offset` must have grown the
// workspace to offset + map_scratch_bytes<Op>(n) beforehand (growing moves the workspace).
template <class Op>
int launch_map(const Op &op, ahead_hip_ws *ws, size_t offset, hipStream_t s) {
    using namespace ahead_dev;
    constexpr int NS = MapTraits<Op>::NS;
    const uint32_t grid = map_grid<Op>(op.n);
    const uint32_t waves = grid * kWaves;
    uint64_t per = (op.n + waves - 1) / waves;
    constexpr uint64_t step = 64u * MapG<Op>::value;
    per = (per + step - 1) / step * step;
    const size_t o_cnt = offset, o_pre = o_cnt + align_up(static_cast<size_t>(NS + 1) * waves * 4, 256),
                 o_erows = o_pre + align_up(static_cast<size_t>(NS + 1) * waves * 4, 256),
                 o_elist = o_erows + align_up(static_cast<size_t>(waves) * 4, 256),
                 o_side = o_elist + align_up(static_cast<size_t>(waves) * kMapErrList * 8, 256),
                 o_flags = o_side + align_up(static_cast<size_t>(waves) * 2 * kMaxShift * 8, 256),
                 o_scratch = o_flags + 256,  // flags, then totals; then the compaction's scratch arrays
                 scratch_each = Op::kMain ? align_up(static_cast<size_t>(op.n) * 8, 256) : 0,
                 o_end = o_scratch + map_outs<Op>() * scratch_each;
    int st;
    if (offset > 0 && ws->bytes < kHeaderBytes + o_end)
        return fail(AHEAD_ERR_INVALID, "map engine: the workspace was not grown before its regions were placed");
    if ((st = ws_ensure(ws, o_end)) != AHEAD_OK) return st;
    MapArgs m{};
    m.rows_per_wave = per;
    m.waves = waves;
    m.wcnt = reinterpret_cast<uint32_t *>(ws_region(ws, o_cnt));
    m.wpre = reinterpret_cast<uint32_t *>(ws_region(ws, o_pre));
    m.werows = reinterpret_cast<uint32_t *>(ws_region(ws, o_erows));
    m.elist = reinterpret_cast<uint64_t *>(ws_region(ws, o_elist));
    m.side = reinterpret_cast<uint64_t *>(ws_region(ws, o_side));
    m.flags = reinterpret_cast<uint32_t *>(ws_region(ws, o_flags));
    m.totals = reinterpret_cast<uint64_t *>(ws_region(ws, o_flags + 64));
    m.scratch[0] = Op::kMain ? reinterpret_cast<uint8_t *>(ws_region(ws, o_scratch)) : nullptr;
    m.scratch[1] = map_outs<Op>() > 1 ? reinterpret_cast<uint8_t *>(ws_region(ws, o_scratch + scratch_each)) : nullptr;
    static_assert(64 + 8 * (NS + 1) <= 256, "totals fit the flags block");
    hipLaunchKernelGGL(k_map<Op>, dim3(grid), dim3(kBlock), 0, s, op, m);
    hipLaunchKernelGGL(k_map_scan<Op>, dim3(NS + 1), dim3(1024), 0, s, op, m);
    if constexpr (Op::kMain) hipLaunchKernelGGL(k_map_save<Op>, dim3(grid), dim3(kBlock), 0, s, op, m);
    hipLaunchKernelGGL(k_map_fix<Op>, dim3(grid), dim3(kBlock), 0, s, op, m);
    if constexpr (Op::kMain) hipLaunchKernelGGL(k_map_back<Op>, dim3(grid), dim3(kBlock), 0, s, op, m);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

// Waves of the map engine's grid for n rows (callers size per-wave partial tables with it).
template <class Op>
uint32_t map_waves(uint64_t n) {
    return map_grid<Op>(n) * ahead_dev::kWaves;
}

}  // namespace ahead_host
