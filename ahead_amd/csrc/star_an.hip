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
//  k_star_probe   one workgroup per tile of ITERS x 2048 fact rows.  Phase 1 streams FK and
//                 measure codes with 128-bit loads, decodes and checks the FK, probes the table
//                 with the truncated decoded key even when the FK is corrupt (hashjoinAN.tcc:199),
//                 fetches and checks both attribute codes and the measure, and accumulates the
//                 measure code into an LDS table of (attr1, attr2) combinations together with the
//                 combination's first tile-local row.  Per-row kept/error bits live in LDS
//                 bitmaps.  Resolve: five look-back streams (kept rows, FK errors, attr1 errors,
//                 attr2 errors, measure errors).  Phase 2 turns first rows into global kept-row
//                 ranks, merges the LDS table into the global table with 64-bit atomics and writes
//                 the error positions in order.
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
constexpr int kComboSlots = 1024;           // LDS combination table per tile
constexpr uint32_t kCombos = 65536;         // 8-bit attr1 x 8-bit attr2
constexpr uint32_t kFinalMaxGroups = 8192;  // groups the single-workgroup finalize can order
enum StarStream { kKept = 0, kEFk = 1, kEA1 = 2, kEA2 = 3, kEM = 4, kStreams = 5 };
constexpr int kOverflow = kStreams;  // extra bitmap: kept rows whose combination missed the LDS table
// d_counts index of each error stream
__host__ __device__ constexpr int count_index(int e) { return e == kEFk ? 1 : (e == kEA1 ? 3 : (e == kEA2 ? 4 : 5)); }

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

struct ProbeArgs {
    const uint64_t *fk;
    const uint64_t *m;
    uint32_t n;
    uint32_t iters;
    uint64_t fk_inv, fk_maxu, key_mask;
    uint64_t m_inv, m_maxu;
    const uint16_t *a1;
    const uint16_t *a2;
    uint32_t n_dim;
    uint16_t a1_inv, a2_inv, a1_maxu, a2_maxu;
    const uint64_t *table;
    uint32_t table_bits;
    const uint32_t *flags;
    uint64_t *g_sum;    // [kCombos] sums of measure codes (mod 2^64)
    uint32_t *g_first;  // [kCombos] 0xFFFFFFFF - first kept-row rank (0 = absent)
    uint64_t a_oid_err;
    uint64_t *err[4];   // FK, attr1, attr2, measure error positions
    uint64_t cap[4];
    uint64_t *d_counts; // error counts at count_index(e), [8] kept rows
    LbArgs lb;
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

// Everything phase 2 needs to re-derive about one fact row (rare error paths only).
struct RowEval {
    bool joined, kept, e_fk, e_a1, e_a2, e_m;
    uint32_t dim_row, combo;
};

__device__ __forceinline__ RowEval eval_row(const ProbeArgs &a, uint64_t fk, uint64_t m) {
    RowEval r{};
    const uint64_t d = fk * a.fk_inv;
    r.e_fk = d > a.fk_maxu;
    const int64_t j = probe_table(a.table, a.table_bits, static_cast<uint32_t>(d & a.key_mask));
    r.joined = j >= 0;
    if (r.joined) {
        r.dim_row = static_cast<uint32_t>(j);
        const uint16_t d1 = an_mul<uint16_t>(a.a1[j], a.a1_inv);
        const uint16_t d2 = an_mul<uint16_t>(a.a2[j], a.a2_inv);
        r.e_a1 = d1 > a.a1_maxu;
        r.e_a2 = d2 > a.a2_maxu;
        r.e_m = (m * a.m_inv) > a.m_maxu;
        r.kept = !(r.e_a1 || r.e_a2 || r.e_m);
        r.combo = (static_cast<uint32_t>(d1 & 0xFFu) << 8) | (d2 & 0xFFu);
    }
    return r;
}

__global__ __launch_bounds__(kBlock) void k_star_probe(ProbeArgs a) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_red[kWaves];
    __shared__ uint32_t s_excl[kStreams];
    __shared__ uint32_t s_bits[kStreams + 1][kStarWords];  // kept / error / overflow bitmaps by row
    __shared__ uint32_t s_wpre[kStarWords];             // kept rows before each bitmap word
    __shared__ uint32_t s_ckey[kComboSlots];            // combination + 1 (0 = free)
    __shared__ uint32_t s_cfirst[kComboSlots];          // first tile-local row
    __shared__ unsigned long long s_csum[kComboSlots];
    __shared__ uint32_t s_any_overflow;
    __shared__ uint32_t s_flag;

    const uint32_t tile = take_tile(a.lb, &s_tile);
    if (tile >= a.lb.num_tiles) return;
    if (threadIdx.x == 0) s_flag = a.flags[0];
    __syncthreads();
    if (s_flag != 0u) {
        // not applicable (duplicate build keys): publish empty counts so the chain drains
        const uint32_t zero[kStreams] = {0, 0, 0, 0, 0};
        lb_resolve<kStreams>(a.lb, tile, zero, s_excl);
        if (tile == a.lb.num_tiles - 1 && threadIdx.x == 0) {
            for (int e = 1; e < kStreams; ++e) a.d_counts[count_index(e)] = 0;
            a.d_counts[8] = 0;
        }
        return;
    }
    for (int i = threadIdx.x; i < (kStreams + 1) * kStarWords; i += kBlock) (&s_bits[0][0])[i] = 0u;
    for (int i = threadIdx.x; i < kComboSlots; i += kBlock) {
        s_ckey[i] = 0u;
        s_cfirst[i] = 0xFFFFFFFFu;
        s_csum[i] = 0ull;
    }
    if (threadIdx.x == 0) s_any_overflow = 0u;
    __syncthreads();

    const uint64_t base = static_cast<uint64_t>(tile) * a.iters * kStarRowsPerSub;
    const bool aligned = ((reinterpret_cast<uintptr_t>(a.fk) | reinterpret_cast<uintptr_t>(a.m)) & 15u) == 0;
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
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint32_t local = sub + c * (kBlock * 2) + threadIdx.x * 2 + k;
                if (base + local >= a.n) continue;
                const RowEval r = eval_row(a, fk[c][k], mv[c][k]);
                const uint32_t w = local >> 5, b = 1u << (local & 31u);
                if (r.e_fk) { atomicOr(&s_bits[kEFk][w], b); cnt[kEFk]++; }
                if (!r.joined) continue;
                if (r.e_a1) { atomicOr(&s_bits[kEA1][w], b); cnt[kEA1]++; }
                if (r.e_a2) { atomicOr(&s_bits[kEA2][w], b); cnt[kEA2]++; }
                if (r.e_m) { atomicOr(&s_bits[kEM][w], b); cnt[kEM]++; }
                if (!r.kept) continue;
                atomicOr(&s_bits[kKept][w], b);
                cnt[kKept]++;
                // LDS combination table: open addressing on combo + 1
                const uint32_t key = r.combo + 1u;
                uint32_t s = (r.combo * 0x9E3779B1u) >> 22;  // 10-bit slot
                bool placed = false;
                for (int p = 0; p < kComboSlots; ++p) {
                    const uint32_t prev = atomicCAS(&s_ckey[s], 0u, key);
                    if (prev == 0u || prev == key) {
                        atomicAdd(&s_csum[s], static_cast<unsigned long long>(mv[c][k]));
                        atomicMin(&s_cfirst[s], local);
                        placed = true;
                        break;
                    }
                    s = (s + 1u) & (kComboSlots - 1u);
                }
                if (!placed) {  // table full: this row merges into the global table in phase 2
                    atomicOr(&s_bits[kOverflow][w], b);
                    s_any_overflow = 1u;
                }
            }
        }
    }

    // resolve: block totals, look-back
    uint32_t agg[kStreams];
#pragma unroll
    for (int s = 0; s < kStreams; ++s) agg[s] = block_sum<uint32_t>(cnt[s], s_red);
    lb_resolve<kStreams>(a.lb, tile, agg, s_excl);

    // phase 2a: kept rows before every bitmap word of the tile
    {
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
    }
    __syncthreads();
    // phase 2b: merge the LDS combination table into the global table
    const uint32_t xk = s_excl[kKept];
    for (int i = threadIdx.x; i < kComboSlots; i += kBlock) {
        const uint32_t key = s_ckey[i];
        if (key == 0u) continue;
        const uint32_t combo = key - 1u, first = s_cfirst[i];
        const uint32_t w = first >> 5;
        const uint32_t rank = xk + s_wpre[w] + __popc(s_bits[kKept][w] & ((1u << (first & 31u)) - 1u));
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_sum[combo]), s_csum[i]);
        atomicMax(&a.g_first[combo], 0xFFFFFFFFu - rank);
    }
    // phase 2b': rows whose combination found no LDS slot go straight to the global table
    if (s_any_overflow) {
        constexpr int PER = kStarWords / kBlock;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t w = threadIdx.x * PER + q;
            uint32_t bits = s_bits[kOverflow][w];
            while (bits) {
                const uint32_t b = __ffs(bits) - 1;
                bits &= bits - 1u;
                const uint64_t row = base + w * 32u + b;
                const RowEval r = eval_row(a, a.fk[row], a.m[row]);
                const uint32_t rank = xk + s_wpre[w] + __popc(s_bits[kKept][w] & ((1u << b) - 1u));
                atomicAdd(reinterpret_cast<unsigned long long *>(&a.g_sum[r.combo]),
                          static_cast<unsigned long long>(a.m[row]));
                atomicMax(&a.g_first[r.combo], 0xFFFFFFFFu - rank);
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
}

struct FinalArgs {
    uint64_t *g_sum;
    uint32_t *g_first;
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
    for (uint32_t c = threadIdx.x; c < kCombos; c += blockDim.x) {
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
        for (uint32_t c = threadIdx.x; c < kCombos; c += blockDim.x) {
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

__global__ void k_clear_u64(uint64_t *p, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

}  // namespace

using namespace ahead_host;

extern "C" int ahead_hip_star1_pipeline(const ahead_star1_args *in, const ahead_star1_out *out,
                                        ahead_hip_ws *ws, void *stream) {
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

    const uint32_t n_dim = static_cast<uint32_t>(in->dim_key.n);
    uint32_t tbits = 4;
    while ((1ull << tbits) < 2ull * n_dim) ++tbits;
    const size_t table_bytes = align_up((1ull << tbits) * 8, 256);
    const size_t gsum_bytes = align_up(kCombos * 8ull, 256), gfirst_bytes = align_up(kCombos * 4ull, 256);
    const uint64_t subs = (in->fact_fk.n + kStarRowsPerSub - 1) / kStarRowsPerSub;
    uint32_t iters = subtiles_per_tile(subs);
    if (iters > static_cast<uint32_t>(kStarMaxIters)) iters = kStarMaxIters;
    const LbPlan plan = lb_plan(in->fact_fk.n, kStarRowsPerSub * iters, kStreams);
    // layout: [status][flags 256][g_sum][g_first][table]
    const size_t off_flags = plan.status_bytes, off_gsum = off_flags + 256, off_gfirst = off_gsum + gsum_bytes,
                 off_table = off_gfirst + gfirst_bytes;
    if ((st = ws_ensure(ws, off_table + table_bytes)) != AHEAD_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);

    // dimension key check = hashjoinAN build-side error vector (hashjoinAN.tcc:179-181)
    uint64_t *dk_counts = out->d_counts + 2;
    BuildArgs b{};
    b.key = static_cast<const uint64_t *>(in->dim_key.data);
    b.n = n_dim;
    b.a_inv = in->dim_key.a_inv;
    b.key_mask = bits_mask(in->dim_key.unenc_bits);
    b.table = reinterpret_cast<uint64_t *>(ws_region(ws, off_table));
    b.table_bits = tbits;
    b.flags = reinterpret_cast<uint32_t *>(ws_region(ws, off_flags));
    hipLaunchKernelGGL(k_clear_u64, dim3(256), dim3(256), 0, s, b.table, 1ull << tbits);
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(32), 0, s, reinterpret_cast<uint64_t *>(b.flags), 32ull);
    AHEAD_HIP_TRY(hipGetLastError());
    if (n_dim) {
        const uint32_t g = (n_dim + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_star_build, dim3(g < 4096 ? g : 4096), dim3(kBlock), 0, s, b);
        AHEAD_HIP_TRY(hipGetLastError());
    }

    ProbeArgs p{};
    p.fk = static_cast<const uint64_t *>(in->fact_fk.data);
    p.m = static_cast<const uint64_t *>(in->fact_m.data);
    p.n = static_cast<uint32_t>(in->fact_fk.n);
    p.iters = iters;
    p.fk_inv = in->fact_fk.a_inv;
    p.fk_maxu = in->fact_fk.max_u;
    p.key_mask = bits_mask(in->dim_key.unenc_bits);
    p.m_inv = in->fact_m.a_inv;
    p.m_maxu = in->fact_m.max_u;
    p.a1 = static_cast<const uint16_t *>(in->dim_a1.data);
    p.a2 = static_cast<const uint16_t *>(in->dim_a2.data);
    p.n_dim = n_dim;
    p.a1_inv = static_cast<uint16_t>(in->dim_a1.a_inv);
    p.a2_inv = static_cast<uint16_t>(in->dim_a2.a_inv);
    p.a1_maxu = static_cast<uint16_t>(in->dim_a1.max_u);
    p.a2_maxu = static_cast<uint16_t>(in->dim_a2.max_u);
    p.table = b.table;
    p.table_bits = tbits;
    p.flags = b.flags;
    p.g_sum = reinterpret_cast<uint64_t *>(ws_region(ws, off_gsum));
    p.g_first = reinterpret_cast<uint32_t *>(ws_region(ws, off_gfirst));
    p.a_oid_err = in->a_oid_err;
    p.err[0] = out->e_fk; p.cap[0] = out->e_fk ? out->e_fk_cap : 0;
    p.err[1] = out->e_a1; p.cap[1] = out->e_a1 ? out->e_a1_cap : 0;
    p.err[2] = out->e_a2; p.cap[2] = out->e_a2 ? out->e_a2_cap : 0;
    p.err[3] = out->e_m; p.cap[3] = out->e_m ? out->e_m_cap : 0;
    p.d_counts = out->d_counts;
    if ((st = lb_args(ws, plan, 0, &p.lb)) != AHEAD_OK) return st;
    hipLaunchKernelGGL(k_star_probe, dim3(plan.num_tiles), dim3(kBlock), 0, s, p);
    AHEAD_HIP_TRY(hipGetLastError());

    FinalArgs f{};
    f.g_sum = p.g_sum;
    f.g_first = p.g_first;
    f.flags = b.flags;
    f.sum_factor = in->a_result * (in->fact_m.a_inv != 0 ? inverse_mod2k(in->fact_m.a, 64) : 1ull);
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

    // dimension-key error vector (checkAN semantics over the build keys)
    st = ahead_hip_check_an(&in->dim_key, in->a_oid_err, out->e_dk, out->e_dk ? out->e_dk_cap : 0, dk_counts,
                            ws, stream);
    return st;
}
