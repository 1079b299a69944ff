// star_an.hip -- fused hardened star join over one dimension for gfx950:
//   hashjoinAN(fact FK, reverse(dim key)) -> fetchjoinAN(attr1), fetchjoinAN(attr2),
//   fetchjoinAN(measure) -> groupbyAN(attr1) -> groupbyAN(attr2, g1) ->
//   aggregate_sum_groupedAN(measure)
// in one pass over the fact columns (the operator chain of tools/ssbm/ssbm-q21_continuous.cpp:97-181
// applied to one dimension with two group attributes).
//
// Row-consistent policy: a joined row whose fetched attribute or measure code is corrupt is
// reported in that fetch's error vector and dropped from group-by and sum (the reference's
// separate operators would throw on the resulting size mismatch, groupbyAN.tcc:105-107).
//
// Kernels
//  k_star_build   inserts every dimension row under its decoded, width-truncated key
//                 (hashjoinAN.tcc:177-187) into an open-addressing table of
//                 (key << 32 | row + 1) words; a repeated key raises the "not fused" flag
//                 (the fused path needs unique build keys).
//  k_star_dim     decodes and checks both attribute codes of every dimension row and marks the
//                 (attr1, attr2) combinations of the valid rows;
//  k_star_cid     numbers the marked combinations densely (combination id, cid);
//  k_star_info    stores per dimension row its cid and its two attribute-error bits.
//  k_star_reset   clears the build's workspace (flags, tables, bitmap, infos, key map) in one launch.
//  k_star_range / k_star_map_fill
//                 when the valid keys span a range of at most 4 x |dim| + 1024 values, a direct
//                 map of 16-bit entries (attr-error bits | cid + 1) indexed by key - kmin, so a
//                 fact row costs one 2-byte gather; keys outside the range use the table.
//  k_star_probe   persistent: 4 workgroups per CU take tiles of ITERS x 2048 fact rows in ticket
//                 order.  Phase 1 streams FK and measure codes with 128-bit loads, decodes and
//                 checks the FK, probes the map (or table) with the truncated decoded key even
//                 when the FK is corrupt (hashjoinAN.tcc:199), checks the measure, and adds the
//                 measure code of every kept row into per-wave LDS accumulators indexed by cid,
//                 which live for the whole launch.  Kept rows form an LDS bitmap built from wave
//                 ballots; error rows (rare) set bits with LDS atomics.  Resolve: five look-back
//                 streams (kept rows, FK errors, attr1 errors, attr2 errors, measure errors).
//                 Phase 2 turns the first row of every combination new to the workgroup into its
//                 global kept-row rank and writes the error positions in order.  At the end each
//                 workgroup writes its sums and first ranks to a partial table.
//  k_star_reduce  one workgroup per cid adds the partial sums and takes the smallest first rank.
//  k_star_final   orders the combinations by first kept-row rank (= binary group ids, since
//                 (attr2, g1) <-> (attr1, attr2) is one-to-one), writes sums, VGIDtoROID and key
//                 codes, and clears the global table for the next launch.
#include "host_common.hpp"

using namespace ahead_dev;

namespace {

constexpr int kStarRowsPerSub = kBlock * 2 * 4;  // 2 x 64-bit codes per 16-byte load, 4 loads
constexpr int kStarMaxIters = 8;
constexpr int kStarTileRows = kStarRowsPerSub * kStarMaxIters;  // 16384
constexpr int kStarWords = kStarTileRows / 32;
constexpr uint32_t kLdsCids = 1024;         // combination ids accumulated in LDS per workgroup
constexpr uint32_t kWaveCids = kLdsCids / kWaves;  // per-wave accumulators when cids fit
constexpr uint32_t kCombos = 65536;         // 8-bit attr1 x 8-bit attr2
constexpr uint32_t kFinalMaxGroups = 8192;  // groups the single-workgroup finalize can order
enum StarStream { kKept = 0, kEFk = 1, kEA1 = 2, kEA2 = 3, kEM = 4, kStreams = 5 };
constexpr int kOverflow = kStreams;  // extra bitmap: kept rows whose combination missed the LDS table
// d_counts index of each error stream
__host__ __device__ constexpr int count_index(int e) { return e == kEFk ? 1 : (e == kEA1 ? 3 : (e == kEA2 ? 4 : 5)); }

// The fast probe (k_star_probe2) runs when the build found it applicable (flags[5]) and the host
// conditions hold; the general kernels run otherwise.  Both are launched; each returns at once
// when it is not the one that applies.
__device__ __forceinline__ bool fast_path(const uint32_t *flags, uint32_t fast_ok) {
    return fast_ok != 0u && flags[5] != 0u;
}

__device__ __forceinline__ uint32_t table_slot(uint32_t key, uint32_t bits) {
    // xor-fold: dense small keys map to themselves (no collisions), wide keys are folded in
    return (key ^ (key >> bits) ^ (key >> 2 * bits)) & ((1u << bits) - 1u);
}

struct BuildArgs {
    const uint64_t *key;  // encoded dimension keys (64-bit codes)
    uint32_t n;
    uint64_t a_inv;
    uint64_t key_mask;    // truncation to the unencoded key width (h2unenc_t)
    uint64_t *table;
    uint32_t table_bits;
    uint32_t *flags;      // [0] != 0: the fused path does not apply
};

__global__ __launch_bounds__(kBlock) void k_star_build(BuildArgs a) {
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t mask = (1u << a.table_bits) - 1u;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < a.n; j += stride) {
        const uint32_t key = static_cast<uint32_t>((a.key[j] * a.a_inv) & a.key_mask);
        const uint64_t entry = (static_cast<uint64_t>(key) << 32) | (j + 1u);
        uint32_t s = table_slot(key, a.table_bits);
        for (uint32_t probes = 0; probes <= mask; ++probes) {
            const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long *>(&a.table[s]), 0ull,
                                                      static_cast<unsigned long long>(entry));
            if (prev == 0ull) break;
            if (static_cast<uint32_t>(prev >> 32) == key) {
                atomicOr(&a.flags[0], 1u);  // duplicate build key
                break;
            }
            s = (s + 1u) & mask;
        }
    }
}

struct DimArgs {
    const uint16_t *a1, *a2;
    uint32_t n;
    uint16_t a1_inv, a2_inv, a1_maxu, a2_maxu;
    uint32_t *present;     // [kCombos / 32] bitmap of combinations of valid rows
    uint16_t *combo_cid;   // [kCombos]
    uint16_t *cid_combo;   // [kCombos]
    uint32_t *info;        // [n]: cid | attr1 error << 30 | attr2 error << 31
    uint32_t *flags;       // [1] number of cids, [2] kmin, [3] kmax, [4] dense map in use
    const uint64_t *key;   // dimension key codes
    uint64_t key_inv, key_maxu, key_mask;
    uint16_t *map;         // [map_cap]
    uint32_t map_cap;
};

// Smallest and largest truncated key of the valid dimension keys (one atomic pair per workgroup).
__global__ __launch_bounds__(kBlock) void k_star_range(DimArgs a) {
    __shared__ uint32_t s_lo[kWaves], s_hi[kWaves];
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < a.n; j += stride) {
        const uint64_t d = a.key[j] * a.key_inv;
        if (d > a.key_maxu) continue;
        const uint32_t k = static_cast<uint32_t>(d & a.key_mask);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    if ((threadIdx.x & 63u) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWaves; ++w) {
            lo = s_lo[w] < lo ? s_lo[w] : lo;
            hi = s_hi[w] > hi ? s_hi[w] : hi;
        }
        atomicMin(&a.flags[2], lo);
        atomicMax(&a.flags[3], hi);
    }
}

// Decides whether the valid keys form a range the direct map can hold and clears that range.
// The direct key map: applicable (flags[4]) when the valid keys span at most map_cap values and the
// combinations fit 14-bit entries.  The map starts zeroed (k_star_reset).
__global__ __launch_bounds__(kBlock) void k_star_map_fill(DimArgs a) {
    const uint32_t lo = a.flags[2], hi = a.flags[3];
    const bool dense = lo <= hi && static_cast<uint64_t>(hi) - lo < a.map_cap && a.flags[1] < 0x3FFEu;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.flags[4] = dense ? 1u : 0u;
    if (!dense) return;
    const uint32_t range = hi - lo + 1u;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < a.n; j += stride) {
        const uint32_t k = static_cast<uint32_t>((a.key[j] * a.key_inv) & a.key_mask);
        if (k - lo >= range) continue;
        const uint32_t info = a.info[j];
        const uint32_t e = info >> 30;
        const uint32_t v = (e & 1u ? 0x8000u : 0u) | (e & 2u ? 0x4000u : 0u) | (e ? 0x3FFFu : (info & 0xFFFFu) + 1u);
        a.map[k - lo] = static_cast<uint16_t>(v);
    }
}

// Marks the combinations of valid rows: an LDS bitmap per workgroup, merged with one atomic per
// nonzero word (a handful of distinct combinations would otherwise serialize on a few words).
__global__ __launch_bounds__(kBlock) void k_star_dim(DimArgs a) {
    __shared__ uint32_t s_present[kCombos / 32];
    for (uint32_t i = threadIdx.x; i < kCombos / 32; i += kBlock) s_present[i] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < a.n; j += stride) {
        const uint16_t d1 = an_mul<uint16_t>(a.a1[j], a.a1_inv), d2 = an_mul<uint16_t>(a.a2[j], a.a2_inv);
        if (d1 <= a.a1_maxu && d2 <= a.a2_maxu) {
            const uint32_t combo = (static_cast<uint32_t>(d1 & 0xFFu) << 8) | (d2 & 0xFFu);
            atomicOr(&s_present[combo >> 5], 1u << (combo & 31u));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kCombos / 32; i += kBlock) {
        if (s_present[i]) atomicOr(&a.present[i], s_present[i]);
    }
}

// One workgroup: dense numbering of the present combinations (in combination order).
__global__ __launch_bounds__(kBlock) void k_star_cid(DimArgs a) {
    __shared__ uint32_t s_red[kWaves];
    constexpr int PER = (kCombos / 32) / kBlock;  // 8 bitmap words per thread
    uint32_t w[PER], c = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        w[q] = a.present[threadIdx.x * PER + q];
        c += __popc(w[q]);
    }
    uint32_t tot;
    uint32_t cid = block_excl_scan<uint32_t>(c, tot, s_red);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        uint32_t bits = w[q];
        while (bits) {
            const uint32_t combo = (threadIdx.x * PER + q) * 32u + (__ffs(bits) - 1);
            bits &= bits - 1u;
            a.combo_cid[combo] = static_cast<uint16_t>(cid);
            a.cid_combo[cid] = static_cast<uint16_t>(combo);
            ++cid;
        }
    }
    if (threadIdx.x == 0) a.flags[1] = tot;
}

__global__ __launch_bounds__(kBlock) void k_star_info(DimArgs a) {
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < a.n; j += stride) {
        const uint16_t d1 = an_mul<uint16_t>(a.a1[j], a.a1_inv), d2 = an_mul<uint16_t>(a.a2[j], a.a2_inv);
        const bool e1 = d1 > a.a1_maxu, e2 = d2 > a.a2_maxu;
        uint32_t info = (static_cast<uint32_t>(e1) << 30) | (static_cast<uint32_t>(e2) << 31);
        if (!e1 && !e2) info |= a.combo_cid[(static_cast<uint32_t>(d1 & 0xFFu) << 8) | (d2 & 0xFFu)];
        a.info[j] = info;
    }
}

struct ProbeArgs {
    const uint64_t *fk;
    const uint64_t *m;
    uint32_t n;
    uint32_t iters;
    uint64_t fk_inv, fk_maxu, key_mask;
    uint64_t m_inv, m_maxu;
    const uint32_t *info;       // per dimension row: cid | attr error bits
    const uint16_t *cid_combo;  // cid -> (attr1 << 8 | attr2)
    const uint16_t *map;        // dense key range: key - kmin -> entry (see k_star_map_fill)
    uint32_t n_dim;
    const uint64_t *table;
    uint32_t table_bits;
    const uint32_t *flags;
    uint64_t *g_sum;    // [kCombos] sums of measure codes (mod 2^64)
    uint32_t *g_first;  // [kCombos] 0xFFFFFFFF - first kept-row rank (0 = absent)
    unsigned long long *part_sum;  // [kLdsCids][grid] per-workgroup sums of LDS-accumulated cids
    uint32_t *part_rank;           // [kLdsCids][grid] per-workgroup first kept-row ranks
    uint64_t a_oid_err;
    uint64_t *err[4];   // FK, attr1, attr2, measure error positions
    uint64_t cap[4];
    uint64_t *d_counts; // error counts at count_index(e), [8] kept rows
    LbArgs lb;
    uint32_t fast_ok;   // host conditions of the fast probe (the general probe stands down when it runs)
    // fast probe tables the same launch turns into ordered error positions (k_star_errfix)
    uint64_t rows_per_wave;
    uint32_t waves2;
    const uint32_t *wcnt;      // [kStreams][waves2] per-wave counts
    const uint32_t *wpre;      // [kStreams][waves2] exclusive prefixes over waves
    const uint32_t *werows;    // [waves2] rows with any error per wave
    const uint64_t *elist;     // [waves2][kErrList] the first of them: flags << 32 | row
};

__device__ __forceinline__ int64_t probe_table(const uint64_t *table, uint32_t bits, uint32_t key) {
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = table_slot(key, bits);
    for (uint32_t probes = 0; probes <= mask; ++probes) {
        const uint64_t e = table[s];
        if (e == 0ull) return -1;
        if (static_cast<uint32_t>(e >> 32) == key) return static_cast<int64_t>(static_cast<uint32_t>(e) - 1u);
        s = (s + 1u) & mask;
    }
    return -1;
}

// Everything phase 1 and the rare phase-2 paths need to know about one fact row.
struct RowEval {
    bool joined, kept, e_fk, e_a1, e_a2, e_m;
    uint32_t dim_row, cid;
};

__device__ __forceinline__ RowEval eval_row(const ProbeArgs &a, uint64_t fk, uint64_t m) {
    RowEval r{};
    const uint64_t d = fk * a.fk_inv;
    r.e_fk = d > a.fk_maxu;
    const int64_t j = probe_table(a.table, a.table_bits, static_cast<uint32_t>(d & a.key_mask));
    r.joined = j >= 0;
    if (r.joined) {
        r.dim_row = static_cast<uint32_t>(j);
        const uint32_t info = a.info[j];
        r.e_a1 = (info >> 30) & 1u;
        r.e_a2 = (info >> 31) & 1u;
        r.e_m = (m * a.m_inv) > a.m_maxu;
        r.kept = !(r.e_a1 || r.e_a2 || r.e_m);
        r.cid = info & 0xFFFFu;
    }
    return r;
}

constexpr uint32_t kErrList = 16;  // error rows a fast-probe wave records in row order

// Fast path: one wave per probe wave that saw errors writes that wave's error positions, each
// stream at its exclusive prefix over the earlier waves (positions as in the general probe's
// phase 2c).  From the recorded rows when they all fit the list, else by rescanning the range.
__device__ void star_errfix(const ProbeArgs &a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t per_block = blockDim.x / 64u;
    for (uint32_t gw = blockIdx.x * per_block + (threadIdx.x >> 6); gw < a.waves2; gw += gridDim.x * per_block) {
        const uint32_t nrows = a.werows[gw];  // rows with any error
        if (nrows == 0u) continue;
        uint64_t idx[kStreams];
#pragma unroll
        for (int e = 1; e < kStreams; ++e) idx[e] = a.wpre[e * a.waves2 + gw];
        const uint64_t begin = static_cast<uint64_t>(gw) * a.rows_per_wave;
        const uint64_t end = begin + a.rows_per_wave < a.n ? begin + a.rows_per_wave : a.n;
        const bool listed = nrows <= kErrList;
        const uint64_t span = listed ? nrows : end - begin;
        for (uint64_t j0 = 0; j0 < span; j0 += 64u) {
            const uint64_t j = j0 + lane;
            bool f[kStreams] = {false, false, false, false, false};
            uint64_t row = 0, dim_row = 0;
            if (j < span) {
                if (listed) {
                    const uint64_t ent = a.elist[static_cast<size_t>(gw) * kErrList + j];
                    row = static_cast<uint32_t>(ent);
                    const uint32_t fl = static_cast<uint32_t>(ent >> 32);
#pragma unroll
                    for (int e = 1; e < kStreams; ++e) f[e] = (fl >> (e - 1)) & 1u;
                    if (f[kEA1] || f[kEA2]) dim_row = eval_row(a, a.fk[row], a.m[row]).dim_row;
                } else {
                    row = begin + j;
                    const RowEval r = eval_row(a, a.fk[row], a.m[row]);
                    f[kEFk] = r.e_fk;
                    f[kEA1] = r.joined && r.e_a1;
                    f[kEA2] = r.joined && r.e_a2;
                    f[kEM] = r.joined && r.e_m;
                    dim_row = r.dim_row;
                }
            }
#pragma unroll
            for (int e = 1; e < kStreams; ++e) {
                const uint64_t b = __ballot(f[e]);
                if (f[e]) {
                    const uint64_t i = idx[e] + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
                    const uint64_t pos = (e == kEA1 || e == kEA2) ? dim_row : row;
                    if (i < a.cap[e - 1]) a.err[e - 1][i] = pos * a.a_oid_err;
                }
                idx[e] += static_cast<uint64_t>(__popcll(b));
            }
        }
    }
}

// 16 bits -> even bit positions of 32 (bit interleave)
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_star_probe(ProbeArgs a) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_red[kWaves];
    __shared__ uint32_t s_excl[kStreams];
    __shared__ uint32_t s_bits[kStreams + 1][kStarWords];  // kept / error / overflow bitmaps by row
    __shared__ uint32_t s_wpre[kStarWords];                 // kept rows before each bitmap word
    __shared__ unsigned long long s_sum[kLdsCids];          // per cid (per wave when cids fit)
    __shared__ uint32_t s_tfirst[kLdsCids];                 // first tile-local row of a cid new to the block
    __shared__ uint32_t s_rank[kLdsCids];                   // first kept-row rank of each cid (block)
    __shared__ uint32_t s_new, s_any_overflow;
    __shared__ uint32_t s_flag[5];

    if (fast_path(a.flags, a.fast_ok)) {  // the fast probe ran: only its error positions remain
        star_errfix(a);
        return;
    }
    if (threadIdx.x < 5) s_flag[threadIdx.x] = a.flags[threadIdx.x];
    for (int i = threadIdx.x; i < (kStreams + 1) * kStarWords; i += kBlock) (&s_bits[0][0])[i] = 0u;
    for (uint32_t i = threadIdx.x; i < kLdsCids; i += kBlock) {
        s_sum[i] = 0ull;
        s_tfirst[i] = 0xFFFFFFFFu;
        s_rank[i] = 0xFFFFFFFFu;
    }
    if (threadIdx.x == 0) {
        s_new = 0u;
        s_any_overflow = 0u;
    }
    __syncthreads();
    const bool skip = s_flag[0] != 0u;
    const uint32_t ncid = s_flag[1];
    const bool dense = s_flag[4] != 0u;
    const uint32_t kmin = s_flag[2], krange = s_flag[3] - s_flag[2] + 1u;
    const bool per_wave = ncid <= kWaveCids;  // private accumulators per wave: no cross-wave contention
    const uint32_t nslots = ncid < kLdsCids ? ncid : kLdsCids;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const bool aligned = ((reinterpret_cast<uintptr_t>(a.fk) | reinterpret_cast<uintptr_t>(a.m)) & 15u) == 0;

    // persistent: tiles are taken in ticket order until none is left
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&a.lb.hdr->ticket, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= a.lb.num_tiles) break;
        if (skip) {
            // not applicable (duplicate build keys): publish empty counts so the chain drains
            const uint32_t zero[kStreams] = {0, 0, 0, 0, 0};
            lb_resolve<kStreams>(a.lb, tile, zero, s_excl);
            if (tile == a.lb.num_tiles - 1 && threadIdx.x == 0) {
                for (int e = 1; e < kStreams; ++e) a.d_counts[count_index(e)] = 0;
                a.d_counts[8] = 0;
            }
            __syncthreads();
            continue;
        }
        const uint64_t base = static_cast<uint64_t>(tile) * a.iters * kStarRowsPerSub;
        uint32_t cnt[kStreams] = {0, 0, 0, 0, 0};

        // phase 1
        for (uint32_t it = 0; it < a.iters; ++it) {
            const uint32_t sub = it * kStarRowsPerSub;
            uint64_t fk[4][2], mv[4][2];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint64_t r0 = base + sub + c * (kBlock * 2) + threadIdx.x * 2;
                if (aligned && r0 + 2 <= a.n) {
                    const u32x4 q = load16_nt(a.fk + r0);
                    const u32x4 w = load16_nt(a.m + r0);
                    __builtin_memcpy(&fk[c][0], &q, 16);
                    __builtin_memcpy(&mv[c][0], &w, 16);
                } else {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        fk[c][k] = r0 + k < a.n ? a.fk[r0 + k] : 0ull;
                        mv[c][k] = r0 + k < a.n ? a.m[r0 + k] : 0ull;
                    }
                }
            }
            // batched probe: the map (or table) words of this thread's 8 rows are requested before
            // any is used, so the dependent gathers overlap instead of chaining
            uint32_t key[4][2], slot[4][2];
            uint64_t ent[4][2];
            bool live[4][2], efk[4][2], in_map[4][2];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const uint32_t local = sub + c * (kBlock * 2) + threadIdx.x * 2 + k;
                    live[c][k] = base + local < a.n;
                    const uint64_t d = fk[c][k] * a.fk_inv;
                    efk[c][k] = live[c][k] && d > a.fk_maxu;
                    key[c][k] = static_cast<uint32_t>(d & a.key_mask);
                    in_map[c][k] = dense && key[c][k] - kmin < krange;
                    slot[c][k] = table_slot(key[c][k], a.table_bits);
                    ent[c][k] = !live[c][k] ? 0ull
                              : (in_map[c][k] ? static_cast<uint64_t>(a.map[key[c][k] - kmin]) : a.table[slot[c][k]]);
                }
            }
            bool joined[4][2];
            int64_t row[4][2];
            uint32_t info[4][2];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    row[c][k] = -1;
                    joined[c][k] = false;
                    info[c][k] = 0u;
                    if (!live[c][k] || ent[c][k] == 0ull) continue;
                    if (in_map[c][k]) {
                        const uint32_t e = static_cast<uint32_t>(ent[c][k]);
                        joined[c][k] = true;
                        info[c][k] = ((e >> 15) & 1u) << 30 | ((e >> 14) & 1u) << 31 | ((e & 0x3FFFu) - 1u);
                        continue;
                    }
                    if (static_cast<uint32_t>(ent[c][k] >> 32) == key[c][k]) {
                        row[c][k] = static_cast<uint32_t>(ent[c][k]) - 1u;
                    } else {  // collision: continue the linear probe
                        const uint32_t mask = (1u << a.table_bits) - 1u;
                        uint32_t s2 = (slot[c][k] + 1u) & mask;
                        for (uint32_t probes = 0; probes < mask; ++probes) {
                            const uint64_t e = a.table[s2];
                            if (e == 0ull) break;
                            if (static_cast<uint32_t>(e >> 32) == key[c][k]) {
                                row[c][k] = static_cast<uint32_t>(e) - 1u;
                                break;
                            }
                            s2 = (s2 + 1u) & mask;
                        }
                    }
                    joined[c][k] = row[c][k] >= 0;
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    if (row[c][k] >= 0) info[c][k] = a.info[row[c][k]];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                bool kept[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const uint32_t local = sub + c * (kBlock * 2) + threadIdx.x * 2 + k;
                    kept[k] = false;
                    const uint32_t w = local >> 5, b = 1u << (local & 31u);
                    if (efk[c][k]) { atomicOr(&s_bits[kEFk][w], b); cnt[kEFk]++; }
                    if (!joined[c][k]) continue;
                    const bool e1 = (info[c][k] >> 30) & 1u, e2 = (info[c][k] >> 31) & 1u;
                    const bool em = (mv[c][k] * a.m_inv) > a.m_maxu;
                    if (e1) { atomicOr(&s_bits[kEA1][w], b); cnt[kEA1]++; }
                    if (e2) { atomicOr(&s_bits[kEA2][w], b); cnt[kEA2]++; }
                    if (em) { atomicOr(&s_bits[kEM][w], b); cnt[kEM]++; }
                    if (e1 || e2 || em) continue;
                    kept[k] = true;
                    cnt[kKept]++;
                    const uint32_t cid = info[c][k] & 0xFFFFu;
                    if (cid < kLdsCids) {
                        const uint32_t sl = per_wave ? wave * kWaveCids + cid : cid;
                        atomicAdd(&s_sum[sl], static_cast<unsigned long long>(mv[c][k]));
                        if (s_rank[cid] == 0xFFFFFFFFu && local < s_tfirst[cid]) {
                            atomicMin(&s_tfirst[cid], local);
                            s_new = 1u;
                        }
                    } else {  // more combinations than LDS holds: merged through the global table in phase 2
                        atomicOr(&s_bits[kOverflow][w], b);
                        s_any_overflow = 1u;
                    }
                }
                // kept bitmap words from two ballots (even / odd rows of the wave's 128-row span)
                const uint64_t b0 = __ballot(kept[0]), b1 = __ballot(kept[1]);
                if (lane < 4) {
                    const uint32_t x0 = static_cast<uint32_t>(b0 >> (16 * lane)), x1 = static_cast<uint32_t>(b1 >> (16 * lane));
                    s_bits[kKept][(sub + c * (kBlock * 2) + wave * 128) / 32 + lane] = spread16(x0) | (spread16(x1) << 1);
                }
            }
        }
        // the tail of a tile shorter than the others holds no rows: its kept words must read 0
        if (a.iters < kStarMaxIters) {
            for (uint32_t w = a.iters * (kStarRowsPerSub / 32) + threadIdx.x; w < kStarWords; w += kBlock)
                s_bits[kKept][w] = 0u;
        }

        // resolve: block totals, look-back
        uint32_t agg[kStreams];
#pragma unroll
        for (int s = 0; s < kStreams; ++s) agg[s] = block_sum<uint32_t>(cnt[s], s_red);
        lb_resolve<kStreams>(a.lb, tile, agg, s_excl);
        const uint32_t xk = s_excl[kKept];
        const bool fresh = s_new != 0u;
        const bool overflow = s_any_overflow != 0u;

        // phase 2a (only when the tile holds the block's first row of some combination, or rows that
        // missed the LDS table): kept rows before every bitmap word of the tile
        if (fresh || overflow) {
            constexpr int PER = kStarWords / kBlock;  // 2 words per thread
            uint32_t c0 = 0;
#pragma unroll
            for (int q = 0; q < PER; ++q) c0 += __popc(s_bits[kKept][threadIdx.x * PER + q]);
            uint32_t tot;
            uint32_t ex = block_excl_scan<uint32_t>(c0, tot, s_red);
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                s_wpre[threadIdx.x * PER + q] = ex;
                ex += __popc(s_bits[kKept][threadIdx.x * PER + q]);
            }
            __syncthreads();
        }
        // phase 2b: first kept-row ranks of combinations new to this block
        if (fresh) {
            for (uint32_t cid = threadIdx.x; cid < nslots; cid += kBlock) {
                const uint32_t first = s_tfirst[cid];
                if (first == 0xFFFFFFFFu) continue;
                const uint32_t w = first >> 5;
                s_rank[cid] = xk + s_wpre[w] + __popc(s_bits[kKept][w] & ((1u << (first & 31u)) - 1u));
                s_tfirst[cid] = 0xFFFFFFFFu;
            }
        }
        // phase 2b': rows whose combination found no LDS slot go straight to the global table
        if (overflow) {
            constexpr int PER = kStarWords / kBlock;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const uint32_t w = threadIdx.x * PER + q;
                uint32_t bits = s_bits[kOverflow][w];
                s_bits[kOverflow][w] = 0u;
                while (bits) {
                    const uint32_t b = __ffs(bits) - 1;
                    bits &= bits - 1u;
                    const uint64_t row = base + w * 32u + b;
                    const RowEval r = eval_row(a, a.fk[row], a.m[row]);
                    const uint32_t rank = xk + s_wpre[w] + __popc(s_bits[kKept][w] & ((1u << b) - 1u));
                    const uint32_t combo = a.cid_combo[r.cid];
                    atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_sum[combo]),
                              static_cast<unsigned long long>(a.m[row]));
                    atomicMax(&a.g_first[combo], 0xFFFFFFFFu - rank);
                }
            }
        }
        // phase 2c: ordered error positions (rare): one pass per stream over the bitmap words
#pragma unroll
        for (int e = 1; e < kStreams; ++e) {
            if (agg[e] == 0u) continue;
            constexpr int PER = kStarWords / kBlock;
            uint32_t c0 = 0;
#pragma unroll
            for (int q = 0; q < PER; ++q) c0 += __popc(s_bits[e][threadIdx.x * PER + q]);
            uint32_t tot;
            uint64_t idx = s_excl[e] + block_excl_scan<uint32_t>(c0, tot, s_red);
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                uint32_t bits = s_bits[e][threadIdx.x * PER + q];
                s_bits[e][threadIdx.x * PER + q] = 0u;
                while (bits) {
                    const uint32_t local = (threadIdx.x * PER + q) * 32u + (__ffs(bits) - 1);
                    bits &= bits - 1u;
                    const uint64_t row = base + local;
                    uint64_t pos = row;
                    if (e == kEA1 || e == kEA2) pos = eval_row(a, a.fk[row], a.m[row]).dim_row;
                    if (idx < a.cap[e - 1]) a.err[e - 1][idx] = pos * a.a_oid_err;
                    ++idx;
                }
            }
        }
        if (tile == a.lb.num_tiles - 1 && threadIdx.x == 0) {
#pragma unroll
            for (int e = 1; e < kStreams; ++e) a.d_counts[count_index(e)] = s_excl[e] + agg[e];
            a.d_counts[8] = s_excl[kKept] + agg[kKept];
        }
        __syncthreads();  // bitmaps, s_new and s_tile are reused by the next tile
        if (threadIdx.x == 0) {
            s_new = 0u;
            s_any_overflow = 0u;
        }
    }

    // the block's accumulators -> its column of the partial tables (reduced by k_star_reduce)
    if (!skip) {
        for (uint32_t cid = threadIdx.x; cid < nslots; cid += kBlock) {
            unsigned long long sum = 0ull;
            if (per_wave) {
#pragma unroll
                for (int w = 0; w < kWaves; ++w) sum += s_sum[w * kWaveCids + cid];
            } else {
                sum = s_sum[cid];
            }
            a.part_sum[static_cast<size_t>(cid) * gridDim.x + blockIdx.x] = sum;
            a.part_rank[static_cast<size_t>(cid) * gridDim.x + blockIdx.x] = s_rank[cid];
        }
    }
    if (threadIdx.x == 0) {
        const uint32_t d = atomicAdd(&a.lb.hdr->done, 1u);
        if (d == gridDim.x - 1) {  // every workgroup has taken its last ticket
            __hip_atomic_store(&a.lb.hdr->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.lb.hdr->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

struct Reduce2Args {
    const unsigned long long *part_sum, *part_first;
    uint32_t parts;          // fast probe workgroups
    const uint32_t *wpre;    // [kStreams][waves] exclusive prefixes over waves
    uint32_t waves;
    uint64_t *g_sum;
    uint32_t *g_first;
};

// Fast path: one workgroup per combination id adds the per-workgroup partial sums; the first
// kept-row rank is the kept rows of all earlier waves plus the rank inside the first wave holding it.
__device__ void star_reduce2(const Reduce2Args &a, const uint16_t *cid_combo, uint32_t ncid) {
    __shared__ unsigned long long s_s[kWaves], s_f[kWaves];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    for (uint32_t cid = blockIdx.x; cid < ncid; cid += gridDim.x) {
        unsigned long long sm = 0ull, f = ~0ull;
        for (uint32_t i = threadIdx.x; i < a.parts; i += blockDim.x) {
            sm += a.part_sum[static_cast<size_t>(cid) * a.parts + i];
            const unsigned long long x = a.part_first[static_cast<size_t>(cid) * a.parts + i];
            f = x < f ? x : f;
        }
        sm = wave_sum(sm);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const unsigned long long o = __shfl_xor(f, d);
            f = o < f ? o : f;
        }
        if (lane == 0) {
            s_s[wv] = sm;
            s_f[wv] = f;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            sm = 0ull;
            f = ~0ull;
            for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
                sm += s_s[w];
                f = s_f[w] < f ? s_f[w] : f;
            }
            if (f != ~0ull) {  // some kept row has this combination
                const uint32_t gw = static_cast<uint32_t>(f >> 32);
                const uint32_t rank = a.wpre[kKept * a.waves + gw] + static_cast<uint32_t>(f);
                const uint32_t combo = cid_combo[cid];
                a.g_sum[combo] = sm;
                a.g_first[combo] = 0xFFFFFFFFu - rank;
            }
        }
        __syncthreads();
    }
}

// One workgroup per LDS-accumulated combination id: adds the per-workgroup partial sums and takes
// the smallest first rank into the global per-combination table (cids beyond the LDS table reached
// it directly from k_star_probe).
__global__ __launch_bounds__(kBlock) void k_star_reduce(const unsigned long long *part_sum, const uint32_t *part_rank,
                                                        uint32_t parts, const uint16_t *cid_combo,
                                                        const uint32_t *flags, uint64_t *g_sum, uint32_t *g_first,
                                                        uint32_t fast_ok, Reduce2Args r2) {
    __shared__ unsigned long long s_s[kWaves];
    __shared__ uint32_t s_r[kWaves];
    if (fast_path(flags, fast_ok)) {
        star_reduce2(r2, cid_combo, flags[1]);
        return;
    }
    const uint32_t ncid = flags[1];
    const uint32_t nslots = ncid < kLdsCids ? ncid : kLdsCids;
    if (flags[0] != 0u) return;
    for (uint32_t cid = blockIdx.x; cid < nslots; cid += gridDim.x) {
        unsigned long long s = 0ull;
        uint32_t r = 0xFFFFFFFFu;
        for (uint32_t i = threadIdx.x; i < parts; i += kBlock) {
            s += part_sum[static_cast<size_t>(cid) * parts + i];
            const uint32_t x = part_rank[static_cast<size_t>(cid) * parts + i];
            r = x < r ? x : r;
        }
        s = wave_sum(s);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const uint32_t o = __shfl_xor(r, d);
            r = o < r ? o : r;
        }
        if ((threadIdx.x & 63u) == 0) {
            s_s[threadIdx.x >> 6] = s;
            s_r[threadIdx.x >> 6] = r;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long ts = 0ull;
            uint32_t tr = 0xFFFFFFFFu;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                ts += s_s[w];
                tr = s_r[w] < tr ? s_r[w] : tr;
            }
            if (tr != 0xFFFFFFFFu) {
                const uint32_t combo = cid_combo[cid];
                g_sum[combo] = ts;
                g_first[combo] = 0xFFFFFFFFu - tr;
            }
        }
        __syncthreads();
    }
}

// ---- fast probe: LDS-resident key map, one contiguous fact range per wave ----------------------
//
// Applies when the dense key map is in use, there are at most 126 combinations and no build key
// repeats (flags[5], set by k_star_pack).  Every dimension key of the map gets a 7-bit code:
// 0 = no dimension row, 1..126 = combination id + 1 of a row with clean attributes,
// 127 = look up the full 16-bit map entry (attribute errors).  Nine codes fill a 64-bit word;
// the words of the first kPackKeys keys of the range sit in LDS for the whole launch, the codes
// of the remaining keys in a byte map in HBM (L2-resident), and keys outside the range go to
// the hash table.  Each wave owns one contiguous range of fact rows and keeps, in LDS, its own
// sums per combination and the wave-local kept-row rank of each combination's first row; its
// kept-row and error counts go to a per-wave table.  Nothing in the scan waits on another wave
// or workgroup.  Rows with errors (rare) are recorded per wave in row order.  k_star_scan2 turns
// the per-wave counts into prefixes; the general kernels' launches, which stand down on this path,
// then finish it: k_star_reduce adds the per-workgroup sums and places the first ranks
// (star_reduce2), k_star_probe writes the ordered error positions (star_errfix).
constexpr int kP2Block = 1024;
constexpr int kP2Waves = kP2Block / 64;
constexpr int kP2Seg = 512;                 // rows per wave iteration: 4 x 128 (2 rows per lane)
constexpr int kPackWords = 19200;           // 150 KiB of LDS
constexpr uint32_t kPackKeys = kPackWords * 9u;
constexpr uint32_t kP2Cids = 128;
constexpr uint32_t kFastMaxCids = 126;
constexpr uint32_t kCodeFull = 127;

__device__ __forceinline__ uint32_t code_of_entry(uint32_t e) {
    if (e == 0u) return 0u;
    if (e & 0xC000u) return kCodeFull;
    const uint32_t c1 = e & 0x3FFFu;  // combination id + 1
    return c1 <= kFastMaxCids ? c1 : kCodeFull;
}

struct PackArgs {
    const uint16_t *map;
    uint64_t *packed;  // [kPackWords]
    uint8_t *code8;    // [range - kPackKeys] codes of the keys beyond the packed ones
    uint32_t *flags;
};

// Builds the packed codes and the byte map from the 16-bit map, and decides flags[5].
__global__ __launch_bounds__(kBlock) void k_star_pack(PackArgs a) {
    const bool fast = a.flags[0] == 0u && a.flags[4] != 0u && a.flags[1] <= kFastMaxCids;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.flags[5] = fast ? 1u : 0u;
    if (!fast) return;
    const uint32_t range = a.flags[3] - a.flags[2] + 1u;
    const uint32_t kl = range < kPackKeys ? range : kPackKeys;
    const uint32_t words = (kl + 8u) / 9u;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < words; q += stride) {
        uint64_t w = 0ull;
        for (uint32_t r = 0; r < 9u; ++r) {
            const uint32_t k = q * 9u + r;
            if (k < kl) w |= static_cast<uint64_t>(code_of_entry(a.map[k])) << (7u * r);
        }
        a.packed[q] = w;
    }
    for (uint32_t k = kl + blockIdx.x * kBlock + threadIdx.x; k < range; k += stride)
        a.code8[k - kl] = static_cast<uint8_t>(code_of_entry(a.map[k]));
}

struct Probe2Args {
    const uint64_t *fk, *m;
    uint64_t n;
    uint64_t rows_per_wave;  // multiple of kP2Seg
    uint64_t fk_inv, fk_maxu, key_mask, m_inv, m_maxu;
    const uint64_t *packed;
    const uint8_t *code8;
    const uint16_t *map;
    const uint64_t *table;
    uint32_t table_bits;
    const uint32_t *info;
    const uint32_t *flags;
    uint32_t fast_ok;                 // host-side conditions (16-byte aligned fact columns, n > 0)
    unsigned long long *part_sum;     // [kP2Cids][grid]
    unsigned long long *part_first;   // [kP2Cids][grid]: global wave index << 32 | wave-local rank
    uint32_t *wcnt;                   // [kStreams][waves]: kept, FK, attr1, attr2, measure errors
    uint32_t *werows;                 // [waves] rows with any error
    uint64_t *elist;                  // [waves][kErrList] the first of them, in row order
};

__global__ __launch_bounds__(kP2Block) void k_star_probe2(Probe2Args a) {
    // LDS (of 160 KiB): the packed codes, the workgroup's sums per combination (LDS atomics
    // conflict only within one instruction, so one table serves all waves), per-wave first ranks
    __shared__ uint64_t s_map[kPackWords];
    __shared__ unsigned long long s_sum[kP2Cids];
    __shared__ uint32_t s_first[kP2Waves][kP2Cids];
    static_assert(sizeof(s_map) + sizeof(s_sum) + sizeof(s_first) <= 160 * 1024, "LDS budget");
    if (!fast_path(a.flags, a.fast_ok)) return;
    const uint32_t kmin = a.flags[2], krange = a.flags[3] - a.flags[2] + 1u;
    const uint32_t kl = krange < kPackKeys ? krange : kPackKeys;
    const uint32_t words = (kl + 8u) / 9u;
    for (uint32_t q = threadIdx.x; q < words; q += kP2Block) s_map[q] = a.packed[q];
    for (uint32_t i = threadIdx.x; i < kP2Waves * kP2Cids; i += kP2Block) (&s_first[0][0])[i] = 0xFFFFFFFFu;
    for (uint32_t i = threadIdx.x; i < kP2Cids; i += kP2Block) s_sum[i] = 0ull;
    __syncthreads();

    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * kP2Waves + wave;
    const uint64_t begin = static_cast<uint64_t>(gw) * a.rows_per_wave;
    const uint64_t end = begin + a.rows_per_wave < a.n ? begin + a.rows_per_wave : a.n;
    uint32_t kept_before = 0;           // wave-uniform: kept rows of this wave so far
    uint32_t err_rows = 0;              // wave-uniform: rows with any error so far
    uint32_t unseen = a.flags[1];       // wave-uniform: combinations without a first row in this wave
    uint32_t ecnt[4] = {0, 0, 0, 0};    // wave-uniform: rows with each error kind (counted from ballots)
    uint32_t *wfirst = s_first[wave];

    // The codes of one segment: 4 chunks of 128 rows, 2 rows per lane.  A full segment finds its
    // chunks already loaded and, as it finishes with chunk c, loads chunk c of the next full
    // segment into the same registers, so the wave's next loads are in flight while it computes.
    uint64_t fk[4][2], mv[4][2];
    auto load_chunk = [&](uint64_t seg, int c) {
        const uint64_t r0 = seg + c * 128u + lane * 2u;
        const u32x4 q = load16_nt(a.fk + r0);
        const u32x4 w = load16_nt(a.m + r0);
        __builtin_memcpy(&fk[c][0], &q, 16);
        __builtin_memcpy(&mv[c][0], &w, 16);
    };
    // one segment of kP2Seg rows; FULL: every row is in range (all but a wave's last segment), so
    // the row-range tests compile away; NEXT (full segments only): the next segment is full too
    auto segment = [&](uint64_t seg, auto full_tag, bool next) {
        constexpr bool full = decltype(full_tag)::value;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t r0 = seg + c * 128u + lane * 2u;
            if (!full) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    fk[c][k] = r0 + k < end ? a.fk[r0 + k] : 0ull;
                    mv[c][k] = r0 + k < end ? a.m[r0 + k] : 0ull;
                }
            }
        }
        uint32_t code[4][2];
        bool live[4][2], efk[4][2];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                live[c][k] = full || seg + c * 128u + lane * 2u + k < end;
                const uint64_t d = fk[c][k] * a.fk_inv;
                efk[c][k] = live[c][k] && d > a.fk_maxu;
                const uint32_t rel = static_cast<uint32_t>(d & a.key_mask) - kmin;
                uint32_t cd;
                if (rel < kl) {
                    const uint32_t q = rel / 9u;
                    cd = static_cast<uint32_t>(s_map[q] >> (7u * (rel - q * 9u))) & 127u;
                } else if (rel < krange) {
                    cd = a.code8[rel - kl];
                } else {
                    cd = kCodeFull;
                }
                code[c][k] = live[c][k] ? cd : 0u;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            bool kept[2];
            uint32_t cid[2], eflags[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                uint32_t cd = code[c][k];
                bool e1 = false, e2 = false;
                if (cd == kCodeFull) {  // rare: attribute errors, or a key outside the map range
                    const uint64_t d = fk[c][k] * a.fk_inv;
                    const uint32_t key = static_cast<uint32_t>(d & a.key_mask);
                    uint32_t inf = 0u;
                    bool joined = false;
                    if (key - kmin < krange) {
                        const uint32_t e = a.map[key - kmin];
                        joined = e != 0u;
                        inf = ((e >> 15) & 1u) << 30 | ((e >> 14) & 1u) << 31 | ((e & 0x3FFFu) - 1u);
                    } else {
                        const int64_t j = probe_table(a.table, a.table_bits, key);
                        joined = j >= 0;
                        if (joined) inf = a.info[j];
                    }
                    e1 = joined && ((inf >> 30) & 1u);
                    e2 = joined && ((inf >> 31) & 1u);
                    cd = joined ? ((inf & 0xFFFFu) + 1u) : 0u;
                }
                const bool joined = cd != 0u;
                const bool em = joined && (mv[c][k] * a.m_inv) > a.m_maxu;
                eflags[k] = (efk[c][k] ? 1u : 0u) | (e1 ? 2u : 0u) | (e2 ? 4u : 0u) | (em ? 8u : 0u);
                kept[k] = joined && !e1 && !e2 && !em;
                cid[k] = cd - 1u;
            }
            // wave-local kept-row ranks in row order (row = seg + c * 128 + lane * 2 + k)
            const uint64_t b0 = __ballot(kept[0]), b1 = __ballot(kept[1]);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b0 >> 32),
                                       __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b0), 0u)) +
                                   __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b1 >> 32),
                                       __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b1), 0u));
            // Both rows' "first of its combination in this wave" tests read the table before either
            // row updates it: a lower lane's second row precedes a higher lane's first row.
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (kept[k]) atomicAdd(&s_sum[cid[k]], static_cast<unsigned long long>(mv[c][k]));
            if (unseen != 0u) {
                bool fresh[2], first[2] = {false, false};
#pragma unroll
                for (int k = 0; k < 2; ++k) fresh[k] = kept[k] && wfirst[cid[k]] == 0xFFFFFFFFu;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    // the one lane whose update finds the entry empty counts the combination as seen
                    if (fresh[k])
                        first[k] = atomicMin(&wfirst[cid[k]], kept_before + below + (k == 1 && kept[0] ? 1u : 0u)) ==
                                   0xFFFFFFFFu;
                }
                unseen -= static_cast<uint32_t>(__popcll(__ballot(first[0])) + __popcll(__ballot(first[1])));
            }
            kept_before += static_cast<uint32_t>(__popcll(b0) + __popcll(b1));
            // rows with errors (rare): recorded in row order while the wave's list has room
            const uint64_t x0 = __ballot(eflags[0] != 0u), x1 = __ballot(eflags[1] != 0u);
            if ((x0 | x1) != 0ull) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    ecnt[e] += static_cast<uint32_t>(__popcll(__ballot((eflags[0] >> e) & 1u)) +
                                                     __popcll(__ballot((eflags[1] >> e) & 1u)));
                const uint32_t xb = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(x0 >> 32),
                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(x0), 0u)) +
                                    __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(x1 >> 32),
                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(x1), 0u));
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    if (eflags[k] == 0u) continue;
                    const uint32_t i = err_rows + xb + (k == 1 && eflags[0] != 0u ? 1u : 0u);
                    if (i < kErrList)
                        a.elist[static_cast<size_t>(gw) * kErrList + i] =
                            (static_cast<uint64_t>(eflags[k]) << 32) | (seg + c * 128u + lane * 2u + k);
                }
                err_rows += static_cast<uint32_t>(__popcll(x0) + __popcll(x1));
            }
            if (full && next) load_chunk(seg + kP2Seg, c);
        }
    };
    uint64_t seg = begin;
    if (seg + kP2Seg <= end) {
#pragma unroll
        for (int c = 0; c < 4; ++c) load_chunk(seg, c);
    }
    for (; seg + kP2Seg <= end; seg += kP2Seg) segment(seg, std::true_type{}, seg + 2 * kP2Seg <= end);
    if (seg < end) segment(seg, std::false_type{}, false);
    // per-wave counts
    const uint32_t waves = gridDim.x * kP2Waves;
    if (lane == 0) {
        a.werows[gw] = err_rows;
        a.wcnt[kKept * waves + gw] = kept_before;
#pragma unroll
        for (int e = 0; e < 4; ++e) a.wcnt[(e + 1) * waves + gw] = ecnt[e];
    }
    __syncthreads();
    // the workgroup's column of the partial tables: sums, and its first wave holding each cid
    if (threadIdx.x < kP2Cids) {
        const uint32_t cid = threadIdx.x;
        const unsigned long long sum = s_sum[cid];
        unsigned long long first = ~0ull;
#pragma unroll
        for (int w = 0; w < kP2Waves; ++w) {
            if (first == ~0ull && s_first[w][cid] != 0xFFFFFFFFu)
                first = (static_cast<unsigned long long>(blockIdx.x * kP2Waves + w) << 32) | s_first[w][cid];
        }
        a.part_sum[static_cast<size_t>(cid) * gridDim.x + blockIdx.x] = sum;
        a.part_first[static_cast<size_t>(cid) * gridDim.x + blockIdx.x] = first;
    }
}

// One workgroup of 1024 threads per count stream: exclusive prefixes over the waves (the error
// positions' offsets and the kept rows before each wave) and the stream's total.
__global__ __launch_bounds__(1024) void k_star_scan2(const uint32_t *wcnt, uint32_t *wpre, uint32_t waves,
                                                      const uint32_t *flags, uint32_t fast_ok, uint64_t *d_counts) {
    __shared__ uint32_t s_w[16];
    if (!fast_path(flags, fast_ok)) return;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    {
        const int s = static_cast<int>(blockIdx.x);
        uint32_t carry = 0;
        for (uint32_t base = 0; base < waves; base += 4096u) {
            // 4 consecutive waves per thread, loaded together
            uint32_t v[4], t = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t i = base + threadIdx.x * 4u + q;
                v[q] = i < waves ? wcnt[s * waves + i] : 0u;
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
            for (int q = 0; q < 4; ++q) {
                const uint32_t i = base + threadIdx.x * 4u + q;
                if (i < waves) wpre[s * waves + i] = pre;
                pre += v[q];
            }
            carry += tot;
        }
        if (threadIdx.x == 0) d_counts[s == kKept ? 8 : count_index(s)] = carry;
    }
}

struct FinalArgs {
    uint64_t *g_sum;
    uint32_t *g_first;
    const uint16_t *cid_combo;
    const uint32_t *flags;
    uint64_t sum_factor;   // A_result * A_measure^-1 (mod 2^64)
    uint64_t a_oid_grp;
    uint16_t a1_a, a2_a;   // attribute A (to re-form the attribute codes)
    uint64_t *sums, *groid, *key1, *key2;
    uint64_t cap;
    uint64_t *d_counts;    // [0] groups, [7] status
};

// One workgroup of 1024 threads: collect present combinations, order them by first rank.
__global__ __launch_bounds__(1024) void k_star_final(FinalArgs a) {
    __shared__ unsigned long long s_key[kFinalMaxGroups];  // (rank << 16) | combo
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    const bool skip = a.flags[0] != 0u;
    const uint32_t ncid = a.flags[1];
    for (uint32_t cid = threadIdx.x; cid < ncid; cid += blockDim.x) {
        const uint32_t c = a.cid_combo[cid];
        const uint32_t f = a.g_first[c];
        if (f == 0u) continue;
        const uint32_t rank = 0xFFFFFFFFu - f;
        const uint32_t slot = atomicAdd(&s_n, 1u);
        if (slot < kFinalMaxGroups) s_key[slot] = (static_cast<unsigned long long>(rank) << 16) | c;
    }
    __syncthreads();
    const uint32_t n = s_n;
    if (skip || n > kFinalMaxGroups) {
        // status 1: not applicable (duplicate build keys) / 2: too many groups for this finalize
        for (uint32_t cid = threadIdx.x; cid < ncid; cid += blockDim.x) {
            const uint32_t c = a.cid_combo[cid];
            a.g_first[c] = 0u;
            a.g_sum[c] = 0ull;
        }
        if (threadIdx.x == 0) {
            a.d_counts[0] = 0;
            a.d_counts[6] = 0;
            a.d_counts[7] = skip ? 1u : 2u;
        }
        return;
    }
    if (n <= blockDim.x) {
        // few groups: each key's place is the number of smaller keys (ranks are distinct)
        if (threadIdx.x < n) {
            const unsigned long long key = s_key[threadIdx.x];
            uint32_t g = 0;
            for (uint32_t j = 0; j < n; ++j) g += s_key[j] < key ? 1u : 0u;
            const uint32_t c = static_cast<uint32_t>(key & 0xFFFFu);
            if (g < a.cap) {
                a.sums[g] = a.g_sum[c] * a.sum_factor;
                a.groid[g] = (key >> 16) * a.a_oid_grp;
                a.key1[g] = static_cast<uint16_t>(static_cast<uint32_t>(c >> 8) * a.a1_a);
                a.key2[g] = static_cast<uint16_t>(static_cast<uint32_t>(c & 0xFFu) * a.a2_a);
            }
            a.g_first[c] = 0u;
            a.g_sum[c] = 0ull;
        }
        if (threadIdx.x == 0) {
            a.d_counts[0] = n;
            a.d_counts[6] = 0;  // the grouped sum's group-id check: every kept row has a valid group
            a.d_counts[7] = 0;
        }
        return;
    }
    uint32_t p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (uint32_t i = n + threadIdx.x; i < p2; i += blockDim.x) s_key[i] = ~0ull;
    __syncthreads();
    // bitonic sort of p2 keys (ranks are distinct)
    for (uint32_t k = 2; k <= p2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long x = s_key[i], y = s_key[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        s_key[i] = y;
                        s_key[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t g = threadIdx.x; g < n; g += blockDim.x) {
        const uint32_t c = static_cast<uint32_t>(s_key[g] & 0xFFFFu);
        const uint64_t rank = s_key[g] >> 16;
        if (g < a.cap) {
            a.sums[g] = a.g_sum[c] * a.sum_factor;
            a.groid[g] = rank * a.a_oid_grp;
            a.key1[g] = static_cast<uint16_t>(static_cast<uint32_t>(c >> 8) * a.a1_a);
            a.key2[g] = static_cast<uint16_t>(static_cast<uint32_t>(c & 0xFFu) * a.a2_a);
        }
        a.g_first[c] = 0u;
        a.g_sum[c] = 0ull;
    }
    if (threadIdx.x == 0) {
        a.d_counts[0] = n;
        a.d_counts[6] = 0;  // the grouped sum's group-id check: every kept row has a valid group
        a.d_counts[7] = 0;
    }
}

// The build's starting state in one launch: the workspace from its start (flags, group tables,
// dimension hash table, combination bitmap, combination ids, row infos, direct key map) is
// zeroed, except flags[2] (the smallest key seen), which starts at 0xFFFFFFFF.
__global__ void k_star_reset(uint64_t *ws, uint64_t words) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < words; i += stride)
        ws[i] = i == 1u ? 0xFFFFFFFFull : 0ull;  // word 1 = flags[2] | flags[3] << 32
}

}  // namespace

using namespace ahead_host;

namespace {

int star_validate(const ahead_star1_args *in, const ahead_star1_out *out) {
    if (in == nullptr || out == nullptr) return fail(AHEAD_ERR_INVALID, "star1: NULL arguments");
    int st;
    if ((st = check_col(&in->fact_fk, "star1.fact_fk", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(&in->fact_m, "star1.fact_m", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(&in->dim_key, "star1.dim_key", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(&in->dim_a1, "star1.dim_a1", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(&in->dim_a2, "star1.dim_a2", false, true)) != AHEAD_OK) return st;
    if (in->fact_fk.n != in->fact_m.n) return fail(AHEAD_ERR_SIZE_MISMATCH, "star1: fact columns differ in size");
    if (in->dim_a1.n != in->dim_key.n || in->dim_a2.n != in->dim_key.n)
        return fail(AHEAD_ERR_SIZE_MISMATCH, "star1: dimension columns differ in size");
    if (in->fact_fk.width != 8 || in->fact_m.width != 8 || in->dim_key.width != 8 || in->dim_a1.width != 2 ||
        in->dim_a2.width != 2 || in->dim_a1.unenc_bits != 8 || in->dim_a2.unenc_bits != 8)
        return fail(AHEAD_ERR_INVALID,
                    "star1: the fused path takes 64-bit key/measure codes and restiny (8-bit) attributes");
    if (in->fact_fk.unenc_bits > 32 || in->dim_key.unenc_bits != in->fact_fk.unenc_bits)
        return fail(AHEAD_ERR_INVALID, "star1: the fused path joins keys of one unencoded width <= 32 bits");
    if (out->d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "star1: d_counts is NULL");
    if ((in->a_oid_grp & 1ull) == 0 || (in->fact_m.a & 1ull) == 0)
        return fail(AHEAD_ERR_INVALID, "star1: A values must be odd");
    return AHEAD_OK;
}

// Workspace layout of one star join: [flags 256][g_sum][g_first][table]...[partials].  The
// dimension table survives between ahead_hip_star1_build and ahead_hip_star1_probe as long as no
// other operator runs on the workspace in between.
struct StarLayout {
    uint32_t tbits = 4, iters = 1;
    LbPlan plan;
    size_t off_flags = 0, off_gsum = 0, off_gfirst = 0, off_table = 0, off_present = 0, off_combo_cid = 0,
           off_cid_combo = 0, off_info = 0, off_map = 0, off_psum = 0, off_prank = 0, total = 0;
    size_t off_packed = 0, off_code8 = 0, off_psum2 = 0, off_pfirst2 = 0, off_wcnt = 0, off_wpre = 0,
           off_werows = 0, off_elist = 0, off_check = 0;
    uint32_t grid2 = 0, waves2 = 0;  // fast probe workgroups and waves
    uint64_t rows_per_wave = 0;
    size_t status_bytes = 0;  // scan state of the probe and of the build-side check
    uint32_t map_cap = 0;
    uint32_t grid = 0;  // persistent probe workgroups
};

// Fast probe grid: one workgroup per resident slot (its LDS map allows one per CU).
uint32_t probe2_grid() {
    static uint32_t resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_star_probe2),
                                                         kP2Block, 0) != hipSuccess || per_cu <= 0)
            per_cu = 1;
        resident = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
    }
    return resident;
}

// Persistent probe grid: as many workgroups as can be resident at once (registers and LDS decide,
// queried from the runtime), never more than there are tiles.
uint32_t probe_grid(uint32_t tiles) {
    static uint32_t resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_star_probe), kBlock, 0) !=
                hipSuccess || per_cu <= 0)
            per_cu = 2;
        resident = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
    }
    return tiles < resident ? tiles : resident;
}

StarLayout star_layout(const ahead_star1_args *in) {
    StarLayout L;
    const uint64_t n_dim = in->dim_key.n;
    while ((1ull << L.tbits) < 2ull * n_dim) ++L.tbits;
    const uint64_t subs = (in->fact_fk.n + kStarRowsPerSub - 1) / kStarRowsPerSub;
    L.iters = subtiles_per_tile(subs);
    if (L.iters > static_cast<uint32_t>(kStarMaxIters)) L.iters = kStarMaxIters;
    L.plan = lb_plan(in->fact_fk.n, kStarRowsPerSub * L.iters, kStreams);
    L.off_flags = 0;
    L.off_gsum = 256;
    L.off_gfirst = L.off_gsum + align_up(kCombos * 8ull, 256);
    L.off_table = L.off_gfirst + align_up(kCombos * 4ull, 256);
    L.off_present = L.off_table + align_up((1ull << L.tbits) * 8, 256);
    L.off_combo_cid = L.off_present + align_up(kCombos / 8, 256);
    L.off_cid_combo = L.off_combo_cid + align_up(kCombos * 2ull, 256);
    L.off_info = L.off_cid_combo + align_up(kCombos * 2ull, 256);
    L.off_map = L.off_info + align_up(n_dim * 4ull + 4, 256);
    L.map_cap = static_cast<uint32_t>(4ull * n_dim + 1024);
    L.grid = probe_grid(L.plan.num_tiles);
    L.off_psum = L.off_map + align_up(L.map_cap * 2ull, 256);
    L.off_prank = L.off_psum + align_up(static_cast<size_t>(kLdsCids) * L.grid * 8, 256);
    L.grid2 = probe2_grid();
    L.waves2 = L.grid2 * kP2Waves;
    const uint64_t per = (in->fact_fk.n + L.waves2 - 1) / L.waves2;
    L.rows_per_wave = (per + kP2Seg - 1) / kP2Seg * kP2Seg;
    L.off_packed = L.off_prank + align_up(static_cast<size_t>(kLdsCids) * L.grid * 4, 256);
    L.off_code8 = L.off_packed + align_up(kPackWords * 8ull, 256);
    L.off_psum2 = L.off_code8 + align_up(L.map_cap, 256);
    L.off_pfirst2 = L.off_psum2 + align_up(static_cast<size_t>(kP2Cids) * L.grid2 * 8, 256);
    L.off_wcnt = L.off_pfirst2 + align_up(static_cast<size_t>(kP2Cids) * L.grid2 * 8, 256);
    L.off_wpre = L.off_wcnt + align_up(static_cast<size_t>(kStreams) * L.waves2 * 4, 256);
    L.off_werows = L.off_wpre + align_up(static_cast<size_t>(kStreams) * L.waves2 * 4, 256);
    L.off_elist = L.off_werows + align_up(static_cast<size_t>(L.waves2) * 4, 256);
    L.off_check = L.off_elist + align_up(static_cast<size_t>(L.waves2) * kErrList * 8, 256);
    // the build-side key check's scratch (the streaming engine's per-wave tables) comes last
    L.total = L.off_check + check_an_scratch_bytes(in->dim_key.n, in->dim_key.width);
    const size_t chk = check_an_status_bytes(in->dim_key.n, in->dim_key.width);
    L.status_bytes = L.plan.status_bytes > chk ? L.plan.status_bytes : chk;
    return L;
}

}  // namespace

extern "C" int ahead_hip_star1_build(const ahead_star1_args *in, const ahead_star1_out *out, ahead_hip_ws *ws,
                                     void *stream) {
    int st = star_validate(in, out);
    if (st != AHEAD_OK) return st;
    const StarLayout L = star_layout(in);
    if ((st = ws_ensure(ws, L.total)) != AHEAD_OK) return st;
    // the probe's scan state too, so that the probe allocates nothing
    if ((st = ws_ensure_status(ws, L.status_bytes)) != AHEAD_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    BuildArgs b{};
    b.key = static_cast<const uint64_t *>(in->dim_key.data);
    b.n = static_cast<uint32_t>(in->dim_key.n);
    b.a_inv = in->dim_key.a_inv;
    b.key_mask = bits_mask(in->dim_key.unenc_bits);
    b.table = reinterpret_cast<uint64_t *>(ws_region(ws, L.off_table));
    b.table_bits = L.tbits;
    b.flags = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_flags));
    // flags, the group tables, the hash table, the combination bitmap and ids, the row infos and
    // the direct key map lie one after another from the workspace's start: one launch clears them
    // all and sets the smallest-key flag
    hipLaunchKernelGGL(k_star_reset, dim3(256), dim3(256), 0, s, reinterpret_cast<uint64_t *>(ws_region(ws, 0)),
                       align_up(L.off_map + L.map_cap * 2ull, 8) / 8);
    AHEAD_HIP_TRY(hipGetLastError());
    DimArgs d{};
    d.a1 = static_cast<const uint16_t *>(in->dim_a1.data);
    d.a2 = static_cast<const uint16_t *>(in->dim_a2.data);
    d.n = b.n;
    d.a1_inv = static_cast<uint16_t>(in->dim_a1.a_inv);
    d.a2_inv = static_cast<uint16_t>(in->dim_a2.a_inv);
    d.a1_maxu = static_cast<uint16_t>(in->dim_a1.max_u);
    d.a2_maxu = static_cast<uint16_t>(in->dim_a2.max_u);
    d.present = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_present));
    d.combo_cid = reinterpret_cast<uint16_t *>(ws_region(ws, L.off_combo_cid));
    d.cid_combo = reinterpret_cast<uint16_t *>(ws_region(ws, L.off_cid_combo));
    d.info = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_info));
    d.flags = b.flags;
    d.key = b.key;
    d.key_inv = in->dim_key.a_inv;
    d.key_maxu = in->dim_key.max_u;
    d.key_mask = b.key_mask;
    d.map = reinterpret_cast<uint16_t *>(ws_region(ws, L.off_map));
    d.map_cap = L.map_cap;
    const uint32_t g = b.n ? (b.n + kBlock - 1) / kBlock : 1;
    const dim3 gd(g < 4096 ? g : 4096);
    if (b.n) {
        hipLaunchKernelGGL(k_star_build, gd, dim3(kBlock), 0, s, b);
        hipLaunchKernelGGL(k_star_dim, dim3(g < 64 ? g : 64), dim3(kBlock), 0, s, d);
    }
    hipLaunchKernelGGL(k_star_cid, dim3(1), dim3(kBlock), 0, s, d);
    if (b.n) {
        hipLaunchKernelGGL(k_star_info, gd, dim3(kBlock), 0, s, d);
        hipLaunchKernelGGL(k_star_range, dim3(g < 128 ? g : 128), dim3(kBlock), 0, s, d);
    }
    if (b.n) hipLaunchKernelGGL(k_star_map_fill, gd, dim3(kBlock), 0, s, d);
    PackArgs pk{};
    pk.map = d.map;
    pk.packed = reinterpret_cast<uint64_t *>(ws_region(ws, L.off_packed));
    pk.code8 = reinterpret_cast<uint8_t *>(ws_region(ws, L.off_code8));
    pk.flags = b.flags;
    hipLaunchKernelGGL(k_star_pack, dim3(64), dim3(kBlock), 0, s, pk);
    AHEAD_HIP_TRY(hipGetLastError());
    // the build side's error vector: checkAN over the dimension keys (hashjoinAN.tcc:179-181)
    return check_an_ws(&in->dim_key, in->a_oid_err, out->e_dk, out->e_dk ? out->e_dk_cap : 0,
                       out->d_counts + 2, ws, stream, L.off_check);
}

extern "C" int ahead_hip_star1_probe(const ahead_star1_args *in, const ahead_star1_out *out, ahead_hip_ws *ws,
                                     void *stream) {
    int st = star_validate(in, out);
    if (st != AHEAD_OK) return st;
    const StarLayout L = star_layout(in);
    if (ws->bytes < kHeaderBytes + L.total)
        return fail(AHEAD_ERR_INVALID, "star1_probe: the workspace holds no dimension table (run star1_build)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProbeArgs p{};
    p.fk = static_cast<const uint64_t *>(in->fact_fk.data);
    p.m = static_cast<const uint64_t *>(in->fact_m.data);
    p.n = static_cast<uint32_t>(in->fact_fk.n);
    p.iters = L.iters;
    p.fk_inv = in->fact_fk.a_inv;
    p.fk_maxu = in->fact_fk.max_u;
    p.key_mask = bits_mask(in->dim_key.unenc_bits);
    p.m_inv = in->fact_m.a_inv;
    p.m_maxu = in->fact_m.max_u;
    p.info = reinterpret_cast<const uint32_t *>(ws_region(ws, L.off_info));
    p.cid_combo = reinterpret_cast<const uint16_t *>(ws_region(ws, L.off_cid_combo));
    p.map = reinterpret_cast<const uint16_t *>(ws_region(ws, L.off_map));
    p.n_dim = static_cast<uint32_t>(in->dim_key.n);
    p.table = reinterpret_cast<const uint64_t *>(ws_region(ws, L.off_table));
    p.table_bits = L.tbits;
    p.flags = reinterpret_cast<const uint32_t *>(ws_region(ws, L.off_flags));
    p.g_sum = reinterpret_cast<uint64_t *>(ws_region(ws, L.off_gsum));
    p.g_first = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_gfirst));
    p.part_sum = reinterpret_cast<unsigned long long *>(ws_region(ws, L.off_psum));
    p.part_rank = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_prank));
    p.a_oid_err = in->a_oid_err;
    p.err[0] = out->e_fk; p.cap[0] = out->e_fk ? out->e_fk_cap : 0;
    p.err[1] = out->e_a1; p.cap[1] = out->e_a1 ? out->e_a1_cap : 0;
    p.err[2] = out->e_a2; p.cap[2] = out->e_a2 ? out->e_a2_cap : 0;
    p.err[3] = out->e_m; p.cap[3] = out->e_m ? out->e_m_cap : 0;
    p.d_counts = out->d_counts;
    const bool aligned = ((reinterpret_cast<uintptr_t>(p.fk) | reinterpret_cast<uintptr_t>(p.m)) & 15u) == 0;
    p.fast_ok = aligned && p.n > 0 ? 1u : 0u;
    if ((st = lb_args(ws, L.plan, &p.lb)) != AHEAD_OK) return st;

    Probe2Args q{};
    q.fk = p.fk;
    q.m = p.m;
    q.n = p.n;
    q.rows_per_wave = L.rows_per_wave;
    q.fk_inv = p.fk_inv;
    q.fk_maxu = p.fk_maxu;
    q.key_mask = p.key_mask;
    q.m_inv = p.m_inv;
    q.m_maxu = p.m_maxu;
    q.packed = reinterpret_cast<const uint64_t *>(ws_region(ws, L.off_packed));
    q.code8 = reinterpret_cast<const uint8_t *>(ws_region(ws, L.off_code8));
    q.map = p.map;
    q.table = p.table;
    q.table_bits = p.table_bits;
    q.info = p.info;
    q.flags = p.flags;
    q.fast_ok = p.fast_ok;
    q.part_sum = reinterpret_cast<unsigned long long *>(ws_region(ws, L.off_psum2));
    q.part_first = reinterpret_cast<unsigned long long *>(ws_region(ws, L.off_pfirst2));
    q.wcnt = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_wcnt));
    q.werows = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_werows));
    q.elist = reinterpret_cast<uint64_t *>(ws_region(ws, L.off_elist));
    uint32_t *wpre = reinterpret_cast<uint32_t *>(ws_region(ws, L.off_wpre));
    p.rows_per_wave = L.rows_per_wave;
    p.waves2 = L.waves2;
    p.wcnt = q.wcnt;
    p.wpre = wpre;
    p.werows = q.werows;
    p.elist = q.elist;
    Reduce2Args r2{};
    r2.part_sum = q.part_sum;
    r2.part_first = q.part_first;
    r2.parts = L.grid2;
    r2.wpre = wpre;
    r2.waves = L.waves2;
    r2.g_sum = p.g_sum;
    r2.g_first = p.g_first;

    // fast path: probe2 + scan2, then the general kernels' launches finish it (or, when it does not
    // apply, probe2 and scan2 return at once and the general kernels run)
    if (p.fast_ok) {
        hipLaunchKernelGGL(k_star_probe2, dim3(L.grid2), dim3(kP2Block), 0, s, q);
        hipLaunchKernelGGL(k_star_scan2, dim3(kStreams), dim3(1024), 0, s, q.wcnt, wpre, L.waves2, p.flags, p.fast_ok,
                           p.d_counts);
    }
    hipLaunchKernelGGL(k_star_probe, dim3(L.grid), dim3(kBlock), 0, s, p);
    hipLaunchKernelGGL(k_star_reduce, dim3(256), dim3(kBlock), 0, s, p.part_sum, p.part_rank, L.grid, p.cid_combo,
                       p.flags, p.g_sum, p.g_first, p.fast_ok, r2);
    AHEAD_HIP_TRY(hipGetLastError());

    FinalArgs f{};
    f.g_sum = p.g_sum;
    f.g_first = p.g_first;
    f.cid_combo = p.cid_combo;
    f.flags = p.flags;
    f.sum_factor = in->a_result * inverse_mod2k(in->fact_m.a, 64);
    f.a_oid_grp = in->a_oid_grp;
    f.a1_a = static_cast<uint16_t>(in->dim_a1.a);
    f.a2_a = static_cast<uint16_t>(in->dim_a2.a);
    f.sums = out->sums;
    f.groid = out->groid;
    f.key1 = out->key1;
    f.key2 = out->key2;
    f.cap = (out->sums && out->groid && out->key1 && out->key2) ? out->groups_cap : 0;
    f.d_counts = out->d_counts;
    hipLaunchKernelGGL(k_star_final, dim3(1), dim3(1024), 0, s, f);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

extern "C" int ahead_hip_star1_pipeline(const ahead_star1_args *in, const ahead_star1_out *out,
                                        ahead_hip_ws *ws, void *stream) {
    int st = ahead_hip_star1_build(in, out, ws, stream);
    if (st != AHEAD_OK) return st;
    return ahead_hip_star1_probe(in, out, ws, stream);
}
