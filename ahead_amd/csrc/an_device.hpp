// an_device.hpp -- device-side building blocks shared by the gfx950 AN operator kernels.
//
//  * AN arithmetic on 16/32/64-bit codes (the check is `code * A^-1 mod 2^k > maxU`).
//  * wave64 and 256-thread block scans / reductions.
//  * An ordered, single-pass, multi-stream compaction: every tile (one workgroup) counts the
//    rows it emits into S output streams, publishes its per-stream aggregate, and resolves its
//    exclusive prefix by looking back over its predecessors (decoupled look-back).  Tile ids
//    come from an atomic ticket, so every predecessor of a tile is already resident and the
//    chain always drains.  Status words are self-contained 64-bit values read and written with
//    relaxed agent-scope atomics: [epoch:30 | flag:2 | count:32], flag 1 = aggregate,
//    2 = inclusive prefix.  The epoch (one per launch) makes stale words of earlier launches
//    unreadable without clearing the status array.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ahead_dev {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
// Upper bound on polls of one predecessor status word.  A healthy predecessor publishes within
// microseconds; the bound only guarantees that no wave can spin forever if the workspace was
// misused (e.g. shared by two streams).  On expiry the kernel raises the workspace fail flag.
constexpr uint32_t kSpinLimit = 1u << 22;
// Most 16 KiB sub-tiles one look-back tile may cover.
constexpr uint32_t kMaxSubtiles = 16;

// ---- workspace header (device) -----------------------------------------------------------------
struct WsHeader {
    uint32_t ticket;  // tile ticket counter; reset to 0 by the last ticket holder
    uint32_t fail;    // set to nonzero by a kernel that hit an internal limit
    uint32_t done;    // finished-tile counter of last-tile-done reductions; reset by the last tile
    uint32_t pad[61];
};

// ---- AN arithmetic ----------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T an_mul(T x, T y) {
    // product modulo 2^(8*sizeof(T)); 16-bit operands are widened so no signed promotion occurs
    if constexpr (sizeof(T) < 4) {
        return static_cast<T>(static_cast<uint32_t>(x) * static_cast<uint32_t>(y));
    } else {
        return x * y;
    }
}

// ---- 128-bit streaming loads --------------------------------------------------------------------------
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte load with the non-temporal hint: column data is read exactly once.
__device__ __forceinline__ u32x4 load16_nt(const void *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
}

// ---- wave / block primitives -----------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Exclusive block scan of one value per thread; `total` receives the block sum.  `scratch`
// must hold NW elements of LDS (NW: waves of the block).  Contains two barriers.
template <typename T, int NW = kWaves>
__device__ __forceinline__ T block_excl_scan(T v, T &total, T *scratch) {
    const T inc = wave_incl_scan(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 63) scratch[wid] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const T x = scratch[w];
        if (w < wid) pre += x;
        tot += x;
    }
    total = tot;
    __syncthreads();
    return pre + inc - v;
}

template <typename T, int NW = kWaves>
__device__ __forceinline__ T block_sum(T v, T *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    T tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += scratch[w];
    __syncthreads();
    return tot;
}

// Packed per-chunk counters: field j (16 bits) of a uint64 counts rows of chunk j.
__device__ __forceinline__ uint32_t field16(uint64_t p, int j) {
    return static_cast<uint32_t>((p >> (16 * j)) & 0xFFFFu);
}
__device__ __forceinline__ uint32_t fields_sum(uint64_t p) {
    return field16(p, 0) + field16(p, 1) + field16(p, 2) + field16(p, 3);
}
__device__ __forceinline__ uint32_t fields_before(uint64_t p, int j) {
    uint32_t s = 0;
    for (int k = 0; k < j; ++k) s += field16(p, k);
    return s;
}

// Calls f(rank, j, k) for every set bit (chunk j, lane-local item k) of a thread's row mask, where
// rank is the row's in-tile rank among the set rows of the whole tile (rows ordered by chunk,
// then thread, then item).  ex / tot are the thread's exclusive and the tile's total packed
// per-chunk counts from block_excl_scan.
template <int CH, int VEC, typename F>
__device__ __forceinline__ void for_each_ranked(uint32_t bits_all, uint64_t ex, uint64_t tot, F &&f) {
    constexpr uint32_t VMASK = (VEC >= 32) ? 0xFFFFFFFFu : ((1u << VEC) - 1u);
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        uint32_t r = fields_before(tot, j) + field16(ex, j);
        uint32_t bits = (bits_all >> (j * VEC)) & VMASK;
        while (bits) {
            const int k = __ffs(bits) - 1;
            f(r++, j, k);
            bits &= bits - 1u;
        }
    }
}

template <int CH, int VEC>
__device__ __forceinline__ uint64_t pack_counts(uint32_t bits_all) {
    constexpr uint32_t VMASK = (VEC >= 32) ? 0xFFFFFFFFu : ((1u << VEC) - 1u);
    uint64_t p = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j) p |= static_cast<uint64_t>(__popc((bits_all >> (j * VEC)) & VMASK)) << (16 * j);
    return p;
}

// ---- ordered multi-stream compaction (decoupled look-back) ---------------------------------------
__device__ __forceinline__ uint64_t lb_load(const uint64_t *w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint64_t *w, uint64_t v) {
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_word(uint32_t epoch, uint32_t flag, uint32_t v) {
    return (static_cast<uint64_t>((epoch << 2) | flag) << 32) | v;
}

struct LbArgs {
    WsHeader *hdr;
    uint64_t *status;  // [S][num_tiles]
    uint32_t num_tiles;
    uint32_t epoch;    // 1 .. 2^30-1
};

// Takes a tile ticket (thread 0) and broadcasts it.  The holder of the last ticket resets the
// counter: every other ticket of this launch has been handed out at that point.  A ticket
// beyond the launch (only possible if a workspace is shared between concurrent launches) raises
// the fail flag; callers return immediately when the result is >= num_tiles.
__device__ __forceinline__ uint32_t take_tile(const LbArgs &lb, uint32_t *s_tile) {
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(&lb.hdr->ticket, 1u);
        if (t == lb.num_tiles - 1) {
            __hip_atomic_store(&lb.hdr->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (t >= lb.num_tiles) atomicOr(&lb.hdr->fail, 2u);
        *s_tile = t;
    }
    __syncthreads();
    return *s_tile;
}

// Publishes the tile aggregates (tile 0 publishes its inclusive prefix directly), resolves the
// exclusive prefixes by look-back in wave 0, publishes the inclusive prefixes and leaves the
// exclusive prefixes in s_excl[S].  Must be called by all threads of the block.
//
// One look-back round examines a window of 64 x kLbPerLane predecessors: lane l owns the
// kLbPerLane consecutive tiles [end - W + l*kLbPerLane, ...), and all S x kLbPerLane status loads
// of a lane are issued together, so a round costs one device-scope load latency.  Measured on
// MI355X, one status word per lane (a 64-tile window) gives the fastest scans.
constexpr int kLbPerLane = 1;

template <int S>
__device__ __forceinline__ void lb_resolve(const LbArgs &lb, uint32_t tile, const uint32_t (&agg)[S],
                                           uint32_t *s_excl) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            lb_store(&lb.status[static_cast<size_t>(s) * lb.num_tiles + tile],
                     lb_word(lb.epoch, tile == 0 ? 2u : 1u, agg[s]));
        }
    }
    if (threadIdx.x < 64) {
        uint32_t excl[S];
#pragma unroll
        for (int s = 0; s < S; ++s) excl[s] = 0;
        if (tile > 0) {
            constexpr int L = kLbPerLane;
            constexpr int W = 64 * L;
            const int lane = lane_id();
            uint32_t done = 0;
            int64_t end = tile;
            bool failed = false;
            while (done != (1u << S) - 1u) {
                const int64_t first = end - W + static_cast<int64_t>(lane) * L;
                uint64_t w[S][L];
                // issue every load of this round before waiting on any of them
#pragma unroll
                for (int s = 0; s < S; ++s) {
#pragma unroll
                    for (int i = 0; i < L; ++i) {
                        const int64_t pred = first + i;
                        w[s][i] = (done & (1u << s)) || pred < 0
                                      ? lb_word(lb.epoch, 2u, 0u)
                                      : lb_load(&lb.status[static_cast<size_t>(s) * lb.num_tiles + pred]);
                    }
                }
                // wait for predecessors that have not published yet
#pragma unroll
                for (int s = 0; s < S; ++s) {
#pragma unroll
                    for (int i = 0; i < L; ++i) {
                        uint32_t spins = 0;
                        while (static_cast<uint32_t>(w[s][i] >> 34) != lb.epoch) {
                            if (++spins >= kSpinLimit) {
                                failed = true;
                                w[s][i] = lb_word(lb.epoch, 2u, 0u);
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                            w[s][i] = lb_load(&lb.status[static_cast<size_t>(s) * lb.num_tiles + first + i]);
                        }
                    }
                }
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (done & (1u << s)) continue;
                    // lane-local: sum from the highest inclusive prefix among my L tiles (if any)
                    uint32_t part = 0;
                    bool has_p = false;
#pragma unroll
                    for (int i = L - 1; i >= 0; --i) {
                        if (!has_p) {
                            part += static_cast<uint32_t>(w[s][i]);
                            has_p = ((w[s][i] >> 32) & 3u) == 2u;
                        }
                    }
                    const uint64_t pmask = __ballot(has_p);
                    if (pmask != 0) {
                        const int hi = 63 - __clzll(static_cast<long long>(pmask));
                        excl[s] += wave_sum(lane >= hi ? part : 0u);
                        done |= 1u << s;
                    } else {
                        excl[s] += wave_sum(part);
                    }
                }
                end -= W;
            }
            if (__ballot(failed) != 0 && lane == 0) atomicOr(&lb.hdr->fail, 1u);
            if (lane == 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    lb_store(&lb.status[static_cast<size_t>(s) * lb.num_tiles + tile],
                             lb_word(lb.epoch, 2u, excl[s] + agg[s]));
                }
            }
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) s_excl[s] = excl[s];
        }
    }
    __syncthreads();
}

// Sums one uint64 per tile over all tiles of a launch: every tile stores its partial, makes it
// visible device-wide and counts itself done; the tile that completes the count adds all
// partials and writes *out.  Must be called by all threads of the block.
__device__ __forceinline__ void last_tile_sum(WsHeader *hdr, uint64_t *partials, uint32_t num_tiles,
                                              uint32_t tile, uint64_t partial, uint64_t *out,
                                              uint64_t *s_scan, uint32_t *s_flag) {
    if (threadIdx.x == 0) {
        partials[tile] = partial;
        __threadfence();
        const uint32_t prev = atomicAdd(&hdr->done, 1u);
        *s_flag = (prev == num_tiles - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (*s_flag) {
        __threadfence();
        uint64_t s = 0;
        for (uint32_t i = threadIdx.x; i < num_tiles; i += kBlock) {
            s += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s = block_sum<uint64_t>(s, s_scan);
        if (threadIdx.x == 0) {
            *out = s;
            __hip_atomic_store(&hdr->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace ahead_dev
