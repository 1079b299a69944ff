// hashjoin_an.hip -- hashjoinAN for gfx950 (hashjoinAN.tcc:77-222, entries :552-585).
//
// The reference builds google::dense_hash_map<unencoded key, vector<payload>> over one side and
// probes it with the other.  Here:
//  * build  : every build row is inserted under its decoded, width-truncated key
//             (static_cast<h2unenc_t>(h2), :186) into an HBM hash table (htable.hpp) that counts
//             the rows per key and keeps the first row.  Keys with several rows additionally get
//             a row list: counts are scanned into offsets, rows are scattered into the lists and
//             each list is sorted back into build (insertion) order (in LDS, and merged in HBM
//             for lists of more than 8192 rows).
//  * dense  : when the build keys span a range of at most 4 x |build| + 1024 values, a direct
//             map over that range (u32 per key: 0 = no row, first row + 1 for a key with one row,
//             bit 31 | table slot for a key with several) lets a probe row make one gather into a
//             compact, cache-resident array instead of probing the table and its count and
//             first-row arrays.
//  * checks : the build side's head/tail checks are an ordered two-stream pass on the row-tile
//             engine (positions in build-scan order).
//  * probe  : row-tile engine over the probe side: per row the head/tail checks, the key
//             (decoded, guarded by `(larger_t)key <= max(build unencoded type)` and truncated to the
//             build key width, :156-157 and :199-200 -- even for a corrupt probe key), and the
//             matching rows' count; phase 2 writes the (probe payload, build payload) pairs of
//             every probe row in probe order, build rows in insertion order.
// Both hash sides (hash_side_t::right / ::left) and the re-encoding variants are supported.
#include "htable.hpp"
#include "maptile.hpp"

using namespace ahead_dev;

namespace {

constexpr uint32_t kMaxSortedList = 8192;  // longest per-key row list ordered in LDS alone (longer: merged in HBM)

constexpr uint32_t kDmapMulti = 0x80000000u;  // direct-map entry: the key has several rows (low bits: slot)

struct HjTable {
    HTable ht;
    uint32_t *dmap;   // [map_cap] direct map (flags[4] != 0: in use)
    uint32_t map_cap;
    uint32_t *cnt;    // [size + 1] rows per key
    uint32_t *first;  // [size + 1] first (lowest) build row per key
    uint32_t *off;    // [size + 1] list offsets (keys with several rows)
    uint32_t *cur;    // [size + 1] fill cursors
    uint32_t *rows;   // [n_build] row lists
    uint32_t *flags;  // [0] several rows share a key, [1] table overflow, [2] unused,
                      // [3] smallest build key, [5] largest build key, [4] direct map in use,
                      // [6] some sampled probe row finds no key, [7] some build row r does not
                      // hold key smallest + r (0 with the direct map: the map is the identity)
    uint64_t *filt;   // [kHjFiltWords] probe filter: the bit of every build key set (filt_bit)
};

// The probe filter: kHjFiltBits bits with the bit of every build key set, so a probe key whose bit
// is clear has no build row.  Exact (bit = key - smallest) when the build keys are of 32 bits or
// less and span fewer than kHjFiltBits values; otherwise the top bits of a multiplicative hash.
// 32 KiB: a probe that misses keeps it in LDS and gathers only for keys it passes.
constexpr uint32_t kHjFiltWords = 4096, kHjFiltBits = kHjFiltWords * 64;
__device__ __forceinline__ uint32_t filt_bit(const uint32_t *flags, uint64_t k) {
    const uint32_t lo = flags[3], hi = flags[5];
    if (lo <= hi) {  // the range of 32-bit build keys is known
        if (k < lo || k > hi) return UINT32_MAX;
        if (hi - lo < kHjFiltBits) return static_cast<uint32_t>(k - lo);
    }
    return static_cast<uint32_t>(((k ^ (k >> 29)) * 0x9E3779B97F4A7C15ull) >> 46);
}

struct BuildArgs {
    HjTable t;
    DCol key;
    uint64_t key_mask;
    uint64_t n;
};

__global__ __launch_bounds__(kBlock) void k_hj_insert(BuildArgs a) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; j < a.n; j += stride) {
        const uint64_t k = a.key.dec_of(a.key.raw(j)) & a.key_mask;
        const uint32_t s = a.t.ht.insert(k);
        if (s == UINT32_MAX) continue;
        if (atomicAdd(&a.t.cnt[s], 1u) != 0u) atomicOr(&a.t.flags[0], 1u);
        atomicMin(&a.t.first[s], static_cast<uint32_t>(j));
    }
}

// Smallest and largest (truncated, decoded) build key: one atomic pair per workgroup.  Keys of
// 32 bits or less only (the direct map is indexed by key - smallest).
__global__ __launch_bounds__(kBlock) void k_hj_range(BuildArgs a) {
    __shared__ uint32_t s_lo[kWaves], s_hi[kWaves];
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; j < a.n; j += stride) {
        const uint32_t k = static_cast<uint32_t>(a.key.dec_of(a.key.raw(j)) & a.key_mask);
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
        atomicMin(&a.t.flags[3], lo);
        atomicMax(&a.t.flags[5], hi);
    }
}

// The flags, the filter and the table's per-slot words, cleared in one launch: flags 0 except [3]
// (smallest build key: all ones), keys / counts / cursors 0, first rows all ones.
__global__ __launch_bounds__(kBlock) void k_hj_clear(HjTable t, uint64_t slots) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < slots || i < kHjFiltWords;
         i += stride) {
        if (i < 64) t.flags[i] = i == 3 ? 0xFFFFFFFFu : 0u;
        if (i < kHjFiltWords) t.filt[i] = 0ull;
        if (i < slots) {
            t.ht.keys[i] = 0ull;
            t.cnt[i] = 0u;
            t.first[i] = 0xFFFFFFFFu;
            t.cur[i] = 0u;
        }
    }
}

// Sets the filter bit of every build key (the filter is zeroed before).  Only the probe variant with
// the filter reads it, and that one runs when some sampled probe row found no key (flags[6]).
__global__ __launch_bounds__(kBlock) void k_hj_filter(BuildArgs a) {
    if (a.t.flags[6] == 0u) return;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    uint32_t *w = reinterpret_cast<uint32_t *>(a.t.filt);
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; j < a.n; j += stride) {
        const uint64_t k = a.key.dec_of(a.key.raw(j)) & a.key_mask;
        const uint32_t b = filt_bit(a.t.flags, k);
        if (b != UINT32_MAX) atomicOr(&w[b >> 5], 1u << (b & 31u));
    }
}

// Decides whether the direct map applies and clears it over the key range.
__global__ __launch_bounds__(kBlock) void k_hj_dmap_clear(HjTable t) {
    const uint32_t lo = t.flags[3], hi = t.flags[5];
    const bool dense = t.flags[1] == 0u && lo <= hi && static_cast<uint64_t>(hi) - lo < t.map_cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) t.flags[4] = dense ? 1u : 0u;
    if (!dense) return;
    const uint32_t range = hi - lo + 1u;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < range; i += gridDim.x * kBlock) t.dmap[i] = 0u;
}

// dmap[key - smallest] = first row + 1 (one row) or kDmapMulti | slot (several rows); flags[7]
// records a row r whose key is not smallest + r (a build side of dense keys in row order makes
// the map the identity, and the probe then computes the row instead of reading it).
__global__ __launch_bounds__(kBlock) void k_hj_dmap_fill(BuildArgs a) {
    if (a.t.flags[4] == 0u) return;
    const uint32_t lo = a.t.flags[3];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    bool moved = false;
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; j < a.n; j += stride) {
        const uint64_t k = a.key.dec_of(a.key.raw(j)) & a.key_mask;
        moved = moved || static_cast<uint64_t>(static_cast<uint32_t>(k) - lo) != j;
        const uint32_t s = a.t.ht.find(k);
        if (s == UINT32_MAX) continue;  // not reached: every build key is in the table
        const uint32_t e = a.t.cnt[s] == 1u ? a.t.first[s] + 1u : (kDmapMulti | s);
        a.t.dmap[static_cast<uint32_t>(k) - lo] = e;
    }
    if (__syncthreads_or(moved) && threadIdx.x == 0) atomicOr(&a.t.flags[7], 1u);
}

// offsets of the row lists: a multi-entry scan over the table slots on the row-tile engine
struct OffsetOp {
    static constexpr int kErr = 0;
    static constexpr bool kMulti = true, kMain = true, kAcc = false;
    using State = uint8_t;
    uint64_t n;
    HjTable t;
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &, uint64_t &) const {
        if (t.flags[0] == 0u) return 0;
        const uint32_t c = t.cnt[row];
        return c > 1u ? c : 0u;
    }
    __device__ void emit_main(uint64_t row, const State &, uint64_t pos, uint32_t) const {
        t.off[row] = static_cast<uint32_t>(pos);
    }
    __device__ void emit_err(int, uint64_t, const State &, uint64_t) const {}
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *) const {}
};

__global__ __launch_bounds__(kBlock) void k_hj_fill(BuildArgs a) {
    if (a.t.flags[0] == 0u) return;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; j < a.n; j += stride) {
        const uint64_t k = a.key.dec_of(a.key.raw(j)) & a.key_mask;
        const uint32_t s = a.t.ht.find(k);
        if (s == UINT32_MAX || a.t.cnt[s] < 2u) continue;
        a.t.rows[a.t.off[s] + atomicAdd(&a.t.cur[s], 1u)] = static_cast<uint32_t>(j);
    }
}

// Sorts every row list back into build order: one workgroup per list (grid-stride over slots).
// A list of up to kMaxSortedList rows is sorted by a bitonic sort in LDS; a longer one (a left
// build over a small dimension puts every fact row of a key in one list) has each of its
// kMaxSortedList-row chunks sorted that way, then the sorted runs are merged pairwise in HBM, run
// length doubling each pass (merge path: every thread finds where its share of the output starts
// in both runs by a binary search, then merges sequentially), ping-ponging with `tmp` (one u32
// per build row).  Row numbers are distinct, so the order is total.
__device__ void hj_bitonic_lds(uint32_t *s_v, uint32_t p2) {
    for (uint32_t k = 2; k <= p2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p2; i += kBlock) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t x = s_v[i], y = s_v[l];
                    if ((x > y) == ((i & k) == 0)) {
                        s_v[i] = y;
                        s_v[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Merges the sorted runs [p0, p0 + w) and [p0 + w, p0 + 2w) (clipped to c) of src into dst, the
// outputs [d0, d1) relative to p0.
__device__ void hj_merge_part(const uint32_t *src, uint32_t *dst, uint64_t p0, uint32_t w, uint32_t c,
                              uint32_t d0, uint32_t d1) {
    const uint32_t la = w < c - p0 ? w : static_cast<uint32_t>(c - p0);
    const uint32_t lb = c - p0 - la < w ? static_cast<uint32_t>(c - p0 - la) : w;
    const uint32_t *A = src + p0, *Bv = src + p0 + la;
    // i: how many of the first d0 outputs come from A
    uint32_t lo = d0 > lb ? d0 - lb : 0u, hi = d0 < la ? d0 : la;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A[mid] < Bv[d0 - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    uint32_t i = lo, j = d0 - lo;
    for (uint32_t d = d0; d < d1; ++d) {
        const bool takeA = i < la && (j >= lb || A[i] < Bv[j]);
        dst[p0 + d] = takeA ? A[i++] : Bv[j++];
    }
}

__global__ __launch_bounds__(kBlock) void k_hj_sort(HjTable t, uint32_t slots, uint32_t *tmp) {
    __shared__ uint32_t s_v[kMaxSortedList];
    if (t.flags[0] == 0u) return;
    for (uint32_t s = blockIdx.x; s < slots; s += gridDim.x) {
        const uint32_t c = t.cnt[s];
        if (c < 2u) continue;
        uint32_t *list = t.rows + t.off[s];
        // sorted chunks of at most kMaxSortedList rows
        for (uint32_t c0 = 0; c0 < c; c0 += kMaxSortedList) {
            const uint32_t cc = c - c0 < kMaxSortedList ? c - c0 : kMaxSortedList;
            uint32_t p2 = 1;
            while (p2 < cc) p2 <<= 1;
            for (uint32_t i = threadIdx.x; i < p2; i += kBlock) s_v[i] = i < cc ? list[c0 + i] : 0xFFFFFFFFu;
            __syncthreads();
            hj_bitonic_lds(s_v, p2);
            for (uint32_t i = threadIdx.x; i < cc; i += kBlock) list[c0 + i] = s_v[i];
            __syncthreads();
        }
        if (c <= kMaxSortedList) continue;
        // merge passes in HBM
        uint32_t *src = list, *dst = tmp + t.off[s];
        const uint32_t per = (c + kBlock - 1) / kBlock;
        for (uint32_t w = kMaxSortedList; w < c; w <<= 1) {
            __threadfence();
            __syncthreads();
            uint32_t pos = threadIdx.x * per;
            const uint32_t end = pos + per < c ? pos + per : c;
            while (pos < end) {
                const uint64_t p0 = static_cast<uint64_t>(pos) / (2ull * w) * (2ull * w);
                const uint64_t pair_end = p0 + 2ull * w < c ? p0 + 2ull * w : c;
                const uint32_t seg_end = end < pair_end ? end : static_cast<uint32_t>(pair_end);
                hj_merge_part(src, dst, p0, w, c, static_cast<uint32_t>(pos - p0), static_cast<uint32_t>(seg_end - p0));
                pos = seg_end;
            }
            uint32_t *x = src;
            src = dst;
            dst = x;
        }
        __threadfence();
        __syncthreads();
        if (src != list)
            for (uint32_t i = threadIdx.x; i < c; i += kBlock) list[i] = src[i];
        __threadfence();
        __syncthreads();
    }
}

// ordered checks of two columns (the build side's head and tail)
struct Check2Op {
    static constexpr int kErr = 2;
    static constexpr bool kMulti = false, kMain = false, kAcc = false;
    using State = uint8_t;
    uint64_t n;
    DCol c0, c1;
    uint64_t a_oid;
    uint64_t *err[2];
    uint64_t cap[2];
    uint64_t *counts[2];
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &errs, uint64_t &) const {
        if (c0.bad_of(c0.raw(row))) errs |= 1u;
        if (c1.bad_of(c1.raw(row))) errs |= 2u;
        return 0;
    }
    __device__ void emit_main(uint64_t, const State &, uint64_t, uint32_t) const {}
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        *counts[0] = t[1];
        *counts[1] = t[2];
    }
};

struct ProbeState {
    uint64_t pay;   // probe-side payload (raw)
    uint32_t slot;  // table slot (keys with several rows)
    uint32_t brow;  // the build row (keys with one row)
};

// The lean probe's row: the build row and, for 8-byte build payloads, the payload code itself,
// gathered while the row is evaluated (before any store of the step).
struct LeanState {
    uint64_t pay;   // probe-side payload (raw)
    uint64_t bval;  // build-side payload code (8-byte build payloads)
    uint32_t brow;  // the build row
};

struct ProbeOp {
    static constexpr int kErr = 2;  // probe head check, probe tail check
    static constexpr bool kMulti = true, kMain = true, kAcc = false;
    using State = ProbeState;
    uint64_t n;
    HjTable t;
    DCol phead, ptail;        // probe BAT (head, tail)
    DCol key;                 // the probe key column (decoded): ptail for a right build, phead for a left
    DCol pay;                 // the probe payload column: phead for a right build, ptail for a left
    DCol bpay;                // build payload column
    uint64_t guard_mask, key_mask, build_max;
    uint32_t build_right;
    uint32_t reenc_p, reenc_b; // re-encode probe / build payload
    uint64_t f_p, f_b, m_p, m_b;
    uint32_t w_p, w_b;
    void *out_h, *out_t;
    uint64_t out_cap;
    uint64_t a_oid;
    uint64_t *err[2];
    uint64_t cap[2];
    uint64_t *counts;         // [0] pairs, error counts at [e_index[0]], [e_index[1]]
    uint32_t e_index[2];
    uint32_t map_launched;    // a streaming-engine probe (ProbeMapOp, or ProbeLeanOp when lean) was launched before
    uint32_t lean;            // ... and it was the ProbeLeanOp pair, which takes probes that miss too

    // the tile engine runs this probe when some build key has several rows, when some sampled probe
    // row finds no key, or when the streaming probe could not be launched
    __device__ bool skip() const {
        return map_launched != 0u && t.flags[0] == 0u && (lean != 0u || t.flags[6] == 0u);
    }

    static constexpr bool kBatched = true;
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        uint64_t h[G], tl[G], k[G];
        bool ok[G];
        phead.raw_n<G>(row, live, h);
        ptail.raw_n<G>(row, live, tl);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            errs[g] = (live[g] && phead.bad_of(h[g]) ? 1u : 0u) | (live[g] && ptail.bad_of(tl[g]) ? 2u : 0u);
            const uint64_t kv = key.dec_of(build_right ? tl[g] : h[g]);
            st[g].pay = build_right ? h[g] : tl[g];
            ok[g] = live[g] && (kv & guard_mask) <= build_max;
            k[g] = kv & key_mask;
            cnt[g] = 0;
        }
        if (t.flags[4] != 0u) {  // direct map: one gather per row, all G issued together
            const uint64_t lo = t.flags[3], span = t.flags[5] - t.flags[3];
            uint32_t e[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint64_t rel = k[g] - lo;
                ok[g] = ok[g] && rel <= span;
                e[g] = t.dmap[ok[g] ? rel : 0ull];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (!ok[g] || e[g] == 0u) continue;
                if ((e[g] & kDmapMulti) == 0u) {
                    st[g].brow = e[g] - 1u;
                    cnt[g] = 1;
                } else {
                    st[g].slot = e[g] & ~kDmapMulti;
                    cnt[g] = t.cnt[st[g].slot];
                }
            }
            return;
        }
        uint32_t sl[G];
        t.ht.find_n<G>(k, ok, sl);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (sl[g] == UINT32_MAX) continue;
            st[g].slot = sl[g];
            cnt[g] = t.cnt[sl[g]];
            if (cnt[g] == 1u) st[g].brow = t.first[sl[g]];
        }
    }

    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &) const {
        const uint64_t h = phead.raw(row), tl = ptail.raw(row);
        if (phead.bad_of(h)) errs |= 1u;
        if (ptail.bad_of(tl)) errs |= 2u;
        const uint64_t kv = key.dec_of(build_right ? tl : h);
        st.pay = build_right ? h : tl;
        if ((kv & guard_mask) > build_max) return 0;
        const uint64_t k = kv & key_mask;
        if (t.flags[4] != 0u) {  // direct map over [flags[3], flags[5]] (set by the build)
            const uint64_t rel = k - t.flags[3];
            if (rel > static_cast<uint64_t>(t.flags[5] - t.flags[3])) return 0;
            const uint32_t e = t.dmap[rel];
            if (e == 0u) return 0;
            if ((e & kDmapMulti) == 0u) {
                st.brow = e - 1u;
                return 1;
            }
            st.slot = e & ~kDmapMulti;
            return t.cnt[st.slot];
        }
        const uint32_t s = t.ht.find(k);
        if (s == UINT32_MAX) return 0;
        st.slot = s;
        if (t.cnt[s] == 1u) {
            st.brow = t.first[s];
            return 1;
        }
        return t.cnt[s];
    }
    __device__ void emit_main(uint64_t, const State &st, uint64_t pos, uint32_t c) const {
        const uint64_t pv = reenc_p ? (st.pay * f_p) & m_p : st.pay;
        for (uint32_t k = 0; k < c; ++k) {
            if (pos + k >= out_cap) break;
            const uint32_t brow = c == 1u ? st.brow : t.rows[t.off[st.slot] + k];
            const uint64_t b = bpay.raw(brow);
            const uint64_t bv = reenc_b ? (b * f_b) & m_b : b;
            if (build_right) {
                store_w(out_h, pos + k, w_p, pv);
                store_w(out_t, pos + k, w_b, bv);
            } else {
                store_w(out_h, pos + k, w_b, bv);
                store_w(out_t, pos + k, w_p, pv);
            }
        }
    }
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *tot) const {
        counts[0] = tot[0];
        counts[e_index[0]] = tot[1];
        counts[e_index[1]] = tot[2];
    }
};

// The probe on the streaming engine: when no build key has several rows, a probe row emits at
// most one pair.  The engine stages each wave's pairs in place and moves them once the offsets are
// known; when rows emit nothing the waves leave gaps that, beyond kMaxShift rows in all, cannot be
// closed in place and every row is evaluated twice, so a probe that misses runs on the tile engine.
struct ProbeMapOp : ProbeOp {
    static constexpr bool kMulti = false;
    static constexpr int kOuts = 2;
    __device__ bool skip() const { return t.flags[0] != 0u || t.flags[6] != 0u; }
    __device__ void out_arrays(void *(&p)[2], uint32_t (&w)[2]) const {
        p[0] = out_h;
        p[1] = out_t;
        w[0] = build_right ? w_p : w_b;
        w[1] = build_right ? w_b : w_p;
    }
};

// The probe for its common shapes, on the streaming engine: every build key has one row, the probe
// key column holds KW-byte codes, the probe payload is void (PV) or of 8-byte codes, the build
// payload is void (BV) or of 8-byte codes, and both outputs are 4 or 8 bytes wide.  The same results as
// ProbeOp (hashjoinAN.tcc:150-220 probe loop), with the widths fixed at compile time: a row costs
// its loads, one or two products, the lookup and, when it joins, two stores.
template <bool FILT>
struct LeanFilter {};
template <>
struct LeanFilter<true> {
    static constexpr int kFilterWords = static_cast<int>(kHjFiltWords);
};

template <int KW, bool PV, bool BV, bool FILT>
struct ProbeLeanOp : LeanFilter<FILT> {
    static constexpr int kErr = 2;  // probe head check, probe tail check
    static constexpr bool kMulti = false, kMain = true, kAcc = false, kBatched = true;
    static constexpr int kOuts = 2;
    // 8 rows per lane and step: half the steps, so half the waits of a step's loads for the
    // previous step's stores (one counter covers both)
    static constexpr int kMapRows = 8;
    using State = LeanState;
    using KT = std::conditional_t<KW == 8, uint64_t, uint32_t>;
    uint64_t n;
    HjTable t;
    const KT *key;                            // probe key codes
    uint64_t key_inv, key_cmask, key_maxu;    // decode (key_enc != 0) and the code-width mask
    uint32_t key_enc, key_err;                // error bit of the key side: 0 probe head, 1 probe tail
    const uint64_t *pay;                      // probe payload codes (PV: void, pay_seq + row)
    uint64_t pay_seq, pay_inv, pay_maxu;
    uint32_t pay_enc;
    const uint64_t *bpay;                     // build payload codes (BV: void, bpay_seq + row)
    uint64_t bpay_seq;
    uint64_t guard_mask, key_mask, build_max;
    uint32_t reenc_p, reenc_b;
    uint64_t f_p, f_b;
    void *out_p, *out_b;                      // probe / build payload outputs
    uint32_t ow;                              // their width: 4 or 8 bytes
    void *out_h, *out_t;
    uint64_t out_cap;
    uint64_t a_oid;
    uint64_t *err[2];
    uint64_t cap[2];
    uint64_t *counts;
    uint32_t e_index[2];
    uint32_t pair_launched;  // !FILT: the FILT variant was launched too (it takes probes that miss)

    // a build key with several rows: the tile engine's ProbeOp runs instead.  FILT: only probes in
    // which some sampled row finds no key (flags[6]); its twin without the filter takes the others.
    __device__ bool skip() const {
        if (t.flags[0] != 0u) return true;
        return FILT ? t.flags[6] == 0u : (pair_launched != 0u && t.flags[6] != 0u);
    }
    __device__ void filter_init(uint64_t *s) const {
        for (uint32_t i = threadIdx.x; i < kHjFiltWords; i += kBlock) s[i] = t.filt[i];
    }
    template <int G>
    __device__ __forceinline__ void eval_f(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G],
                                           uint32_t (&errs)[G], uint32_t (&cnt)[G], uint64_t &acc,
                                           const uint64_t *filt) const {
        eval_core<G>(row, live, st, errs, cnt, filt);
    }
    __device__ void out_arrays(void *(&p)[2], uint32_t (&w)[2]) const {
        p[0] = out_h;
        p[1] = out_t;
        w[0] = ow;
        w[1] = ow;
    }
    template <int G>
    __device__ __forceinline__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G],
                                           uint32_t (&errs)[G], uint32_t (&cnt)[G], uint64_t &) const {
        eval_core<G>(row, live, st, errs, cnt, nullptr);
    }
    // filt: the probe filter in LDS, or NULL
    template <int G>
    __device__ __forceinline__ void eval_core(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G],
                                              uint32_t (&errs)[G], uint32_t (&cnt)[G], const uint64_t *filt) const {
        KT x[G];
        uint64_t p[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            x[g] = key[live[g] ? row[g] : 0ull];
            p[g] = PV ? static_cast<uint64_t>(static_cast<uint32_t>(pay_seq + row[g])) : pay[live[g] ? row[g] : 0ull];
        }
        bool ok[G];
        uint64_t k[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint64_t c = x[g];
            const uint64_t d = key_enc ? (c * key_inv) & key_cmask : c;
            const bool kb = key_enc && d > key_maxu;
            const bool pb = !PV && pay_enc && p[g] * pay_inv > pay_maxu;
            errs[g] = live[g] ? ((kb ? 1u : 0u) << key_err) | ((pb ? 1u : 0u) << (1u - key_err)) : 0u;
            st[g].pay = p[g];
            ok[g] = live[g] && (d & guard_mask) <= build_max;
            k[g] = d & key_mask;
            cnt[g] = 0;
        }
        if (filt != nullptr) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint32_t b = filt_bit(t.flags, k[g]);
                ok[g] = ok[g] && b != UINT32_MAX && ((filt[b >> 6] >> (b & 63u)) & 1ull) != 0ull;
            }
        }
        if (t.flags[4] != 0u && t.flags[7] == 0u) {  // identity map: build row = key - smallest
            const uint64_t lo = t.flags[3], span = t.flags[5] - t.flags[3];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint64_t rel = k[g] - lo;
                if (ok[g] && rel <= span) {
                    st[g].brow = static_cast<uint32_t>(rel);
                    cnt[g] = 1;
                }
            }
        } else if (t.flags[4] != 0u) {  // direct map: one gather per row (one row per key: no flag bit)
            const uint64_t lo = t.flags[3], span = t.flags[5] - t.flags[3];
            uint32_t e[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint64_t rel = k[g] - lo;
                ok[g] = ok[g] && rel <= span;
                e[g] = t.dmap[ok[g] ? rel : 0ull];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (ok[g] && e[g] != 0u) {
                    st[g].brow = e[g] - 1u;
                    cnt[g] = 1;
                }
            }
        } else {
            uint32_t sl[G];
            t.ht.find_n<G>(k, ok, sl);
            uint32_t f[G];
#pragma unroll
            for (int g = 0; g < G; ++g) f[g] = t.first[sl[g] != UINT32_MAX ? sl[g] : 0u];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (sl[g] != UINT32_MAX) {
                    st[g].brow = f[g];
                    cnt[g] = 1;
                }
            }
        }
        if constexpr (!BV) {  // the build payloads of the joined rows, all before the step's stores
#pragma unroll
            for (int g = 0; g < G; ++g) {
                st[g].bval = 0;
                if (cnt[g] != 0u) st[g].bval = bpay[st[g].brow];
            }
        }
    }
    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &acc) const {
        const uint64_t r[1] = {row};
        const bool l[1] = {row < n};
        State s1[1];
        uint32_t e1[1], c1[1];
        eval_n<1>(r, l, s1, e1, c1, acc);
        st = s1[0];
        errs = e1[0];
        return c1[0];
    }
    __device__ void emit_main(uint64_t, const State &st, uint64_t pos, uint32_t) const {
        if (pos >= out_cap) return;
        const uint64_t b = BV ? static_cast<uint64_t>(static_cast<uint32_t>(bpay_seq + st.brow)) : st.bval;
        const uint64_t pv = reenc_p ? st.pay * f_p : st.pay, bv = reenc_b ? b * f_b : b;
        if (ow == 8) {
            static_cast<uint64_t *>(out_p)[pos] = pv;
            static_cast<uint64_t *>(out_b)[pos] = bv;
        } else {  // 4-byte outputs keep the low half, as store_w does
            static_cast<uint32_t *>(out_p)[pos] = static_cast<uint32_t>(pv);
            static_cast<uint32_t *>(out_b)[pos] = static_cast<uint32_t>(bv);
        }
    }
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *tot) const {
        counts[0] = tot[0];
        counts[e_index[0]] = tot[1];
        counts[e_index[1]] = tot[2];
    }
};

// Calls f with each ProbeLeanOp variant of the given shape: without the filter, and (8-byte keys)
// with it.  Returns whether the shape has a variant with the filter.
template <class F>
bool lean_shape(uint32_t kw, bool pv, bool bv, F &&f) {
#define AHEAD_LEAN(K, P, B, FL)                     \
    if (kw == K && pv == P && bv == B) {            \
        f(ProbeLeanOp<K, P, B, false>{});           \
        if (FL) f(ProbeLeanOp<K, P, B, FL>{});      \
        return FL;                                  \
    }
    AHEAD_LEAN(8, true, true, true)
    AHEAD_LEAN(8, true, false, true)
    AHEAD_LEAN(8, false, true, true)
    AHEAD_LEAN(8, false, false, true)
    AHEAD_LEAN(4, true, true, false)
    AHEAD_LEAN(4, true, false, false)
    AHEAD_LEAN(4, false, true, false)
    AHEAD_LEAN(4, false, false, false)
#undef AHEAD_LEAN
    return false;
}

// Whether every one of kHjSamples evenly spaced probe rows finds a build key: flags[6] = 1 when
// some do not.  The streaming engine closes the gaps of dropped rows in place only while they total
// at most kMaxShift rows, so it suits probes where (nearly) every row finds its key, as foreign
// keys into a whole dimension do.  It only picks the engine; both give the same results.
constexpr uint32_t kHjSamples = 4096;  // one per thread of kHjSamples / kBlock workgroups
__global__ __launch_bounds__(kBlock) void k_hj_sample(ProbeOp p, uint64_t np) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    const uint64_t row = np * k / kHjSamples;
    ProbeOp::State st{};
    uint32_t errs = 0;
    uint64_t aux = 0;
    if (row < np && p.eval(row, st, errs, aux) == 0u) atomicOr(&p.t.flags[6], 1u);
}

__global__ void k_hj_status(const uint32_t *flags, uint64_t *d_counts) {
    d_counts[5] = flags[1] ? 1u : 0u;
}

}  // namespace

using namespace ahead_host;

extern "C" int ahead_hip_hashjoin_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2, const ahead_col *t2,
                                     int build_right, int reenc_h, uint64_t f_h1, uint32_t wout_h, int reenc_t,
                                     uint64_t f_t2, uint32_t wout_t, uint64_t a_oid, ahead_buf out_h, ahead_buf out_t,
                                     ahead_buf e1, ahead_buf e2, ahead_buf e3, ahead_buf e4, uint64_t *d_counts,
                                     ahead_hip_ws *ws, void *stream) {
    int st;
    if ((st = check_col(h1, "hashjoin.h1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t1, "hashjoin.t1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(h2, "hashjoin.h2", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t2, "hashjoin.t2", true, false)) != AHEAD_OK) return st;
    if (h1->n != t1->n || h2->n != t2->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "hashjoin: head and tail sizes differ");
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "hashjoin: d_counts is NULL");
    for (uint32_t w : {wout_h, wout_t})
        if (w != 1 && w != 2 && w != 4 && w != 8) return fail(AHEAD_ERR_INVALID, "hashjoin: output widths 1, 2, 4, 8");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (t1->n == 0 || t2->n == 0) {  // hashjoinAN.tcc:128: nothing is built, probed or checked
        return fill_u64(d_counts, 6, 0, s);
    }
    const bool right = build_right != 0;
    const ahead_col &bkey = right ? *h2 : *t1, &bpay = right ? *t2 : *h1;
    const ahead_col &phead = right ? *h1 : *h2, &ptail = right ? *t1 : *t2;
    const ahead_col &pkey = right ? *t1 : *h2;
    const uint64_t nb = bkey.n, np = phead.n;
    // the table holds 2^bits >= 2 x build rows slots with bits <= 31 (32-bit slot arithmetic)
    if (nb > (1ull << 30)) return fail(AHEAD_ERR_TOO_LARGE, "hashjoin: build side above 2^30 rows");

    uint32_t bits = 4;
    while ((1ull << bits) < 2ull * nb) ++bits;
    const uint64_t slots = (1ull << bits) + 1;
    // layout: [flags 256][keys][cnt][first][off][cur][rows][direct map]
    const uint32_t key_bits = bkey.data ? bkey.unenc_bits : 32;
    // direct map only for <= 32-bit keys; at least 4 Mi entries (16 MiB), so a filtered dimension
    // of a few million keys (the probe then gathers one entry per row, never a probe chain) fits
    const uint64_t map_cap = key_bits <= 32 ? (4ull * nb + 1024 > (1ull << 22) ? 4ull * nb + 1024 : 1ull << 22) : 0;
    const size_t o_flags = 0, o_keys = 256, o_cnt = o_keys + align_up(slots * 8, 256),
                 o_first = o_cnt + align_up(slots * 4, 256), o_off = o_first + align_up(slots * 4, 256),
                 o_cur = o_off + align_up(slots * 4, 256), o_rows = o_cur + align_up(slots * 4, 256),
                 o_dmap = o_rows + align_up(nb * 4, 256), o_filt = o_dmap + align_up(map_cap * 4, 256),
                 o_end = o_filt + kHjFiltWords * 8;
    // the lean probe's shapes (ProbeLeanOp)
    const ahead_col &ppay = right ? phead : ptail;
    const uint32_t lean_kw = pkey.data ? pkey.width : 0u;
    const bool lean = (lean_kw == 8 || lean_kw == 4) && (!ppay.data || ppay.width == 8) &&
                      (!bpay.data || bpay.width == 8) && wout_h == wout_t && (wout_h == 8 || wout_h == 4) && out_h.data && out_t.data &&
                      (out_h.cap < out_t.cap ? out_h.cap : out_t.cap) >= np;
    // the streaming probe's tables follow the build's
    size_t need = o_end + map_scratch_bytes<ProbeMapOp>(np);
    if (need < o_end + align_up(nb * 4, 256)) need = o_end + align_up(nb * 4, 256);  // long-list merge scratch
    if (lean) {
        lean_shape(lean_kw, !ppay.data, !bpay.data, [&](auto proto) {
            const size_t b = o_end + map_scratch_bytes<decltype(proto)>(np);
            if (b > need) need = b;
        });
    }
    if ((st = ws_ensure(ws, need)) != AHEAD_OK) return st;

    HjTable t{};
    t.flags = reinterpret_cast<uint32_t *>(ws_region(ws, o_flags));
    t.ht.keys = reinterpret_cast<unsigned long long *>(ws_region(ws, o_keys));
    t.ht.bits = bits;
    t.ht.overflow = t.flags + 1;
    t.cnt = reinterpret_cast<uint32_t *>(ws_region(ws, o_cnt));
    t.first = reinterpret_cast<uint32_t *>(ws_region(ws, o_first));
    t.off = reinterpret_cast<uint32_t *>(ws_region(ws, o_off));
    t.cur = reinterpret_cast<uint32_t *>(ws_region(ws, o_cur));
    t.rows = reinterpret_cast<uint32_t *>(ws_region(ws, o_rows));
    t.dmap = reinterpret_cast<uint32_t *>(ws_region(ws, o_dmap));
    t.map_cap = static_cast<uint32_t>(map_cap);
    t.filt = reinterpret_cast<uint64_t *>(ws_region(ws, o_filt));
    {
        const uint64_t span = slots > kHjFiltWords ? slots : kHjFiltWords;
        const uint64_t blocks = (span + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_hj_clear, dim3(static_cast<uint32_t>(blocks < 4096 ? blocks : 4096)), dim3(kBlock), 0, s,
                           t, slots);
    }

    BuildArgs b{};
    b.t = t;
    b.key = dcol(bkey);
    b.key_mask = bits_mask(bkey.data ? bkey.unenc_bits : 32);
    b.n = nb;
    const uint64_t gb = (nb + kBlock - 1) / kBlock;
    const dim3 grid(static_cast<uint32_t>(gb < 4096 ? gb : 4096));
    hipLaunchKernelGGL(k_hj_insert, grid, dim3(kBlock), 0, s, b);
    if (map_cap) {
        hipLaunchKernelGGL(k_hj_range, dim3(gb < 128 ? static_cast<uint32_t>(gb) : 128u), dim3(kBlock), 0, s, b);
        hipLaunchKernelGGL(k_hj_dmap_clear, dim3(512), dim3(kBlock), 0, s, t);
        hipLaunchKernelGGL(k_hj_dmap_fill, grid, dim3(kBlock), 0, s, b);
    }
    AHEAD_HIP_TRY(hipGetLastError());
    OffsetOp off{};
    off.n = slots;
    off.t = t;
    if ((st = launch_rowtile(off, slots, 1, ws, s)) != AHEAD_OK) return st;
    hipLaunchKernelGGL(k_hj_fill, grid, dim3(kBlock), 0, s, b);
    // the merge scratch of long row lists shares the probe's scratch region (the build is done first)
    hipLaunchKernelGGL(k_hj_sort, dim3(1024), dim3(kBlock), 0, s, t, static_cast<uint32_t>(slots),
                       reinterpret_cast<uint32_t *>(ws_region(ws, o_end)));
    AHEAD_HIP_TRY(hipGetLastError());

    // build-side checks, in build-scan order: e3/e4 for a right build, e1/e2 for a left build
    Check2Op chk{};
    chk.n = nb;
    chk.c0 = dcol(right ? *h2 : *h1);
    chk.c1 = dcol(right ? *t2 : *t1);
    chk.a_oid = a_oid;
    const ahead_buf be0 = right ? e3 : e1, be1 = right ? e4 : e2;
    chk.err[0] = static_cast<uint64_t *>(be0.data);
    chk.cap[0] = be0.data ? be0.cap : 0;
    chk.err[1] = static_cast<uint64_t *>(be1.data);
    chk.cap[1] = be1.data ? be1.cap : 0;
    chk.counts[0] = d_counts + (right ? 3 : 1);
    chk.counts[1] = d_counts + (right ? 4 : 2);
    if ((st = launch_rowtile(chk, nb, 3, ws, s)) != AHEAD_OK) return st;

    ProbeOp p{};
    p.n = np;
    p.t = t;
    p.phead = dcol(phead);
    p.ptail = dcol(ptail);
    p.key = dcol(pkey);
    p.pay = dcol(right ? phead : ptail);
    p.bpay = dcol(bpay);
    const uint32_t pbits = pkey.data ? pkey.unenc_bits : 32, bbits = bkey.data ? bkey.unenc_bits : 32;
    p.guard_mask = bits_mask(pbits > bbits ? pbits : bbits);  // larger_type<t1unenc_t, h2unenc_t>
    p.build_max = bits_mask(bbits);                           // numeric_limits<h2unenc_t>::max()
    p.key_mask = bits_mask(bbits);
    p.build_right = right;
    // probe payload = h1 (right build) / t2 (left build); build payload = t2 / h1
    p.reenc_p = right ? reenc_h != 0 : reenc_t != 0;
    p.reenc_b = right ? reenc_t != 0 : reenc_h != 0;
    p.f_p = right ? f_h1 : f_t2;
    p.f_b = right ? f_t2 : f_h1;
    p.w_p = right ? wout_h : wout_t;
    p.w_b = right ? wout_t : wout_h;
    p.m_p = width_mask(p.w_p);
    p.m_b = width_mask(p.w_b);
    p.out_h = out_h.data;
    p.out_t = out_t.data;
    p.out_cap = (out_h.data && out_t.data) ? (out_h.cap < out_t.cap ? out_h.cap : out_t.cap) : 0;
    p.a_oid = a_oid;
    const ahead_buf pe0 = right ? e1 : e3, pe1 = right ? e2 : e4;
    p.err[0] = static_cast<uint64_t *>(pe0.data);
    p.cap[0] = pe0.data ? pe0.cap : 0;
    p.err[1] = static_cast<uint64_t *>(pe1.data);
    p.cap[1] = pe1.data ? pe1.cap : 0;
    p.counts = d_counts;
    p.e_index[0] = right ? 1 : 3;
    p.e_index[1] = right ? 2 : 4;
    // the streaming engine needs room for a staged pair per probe row
    const bool map_ok = out_h.data && out_t.data && p.out_cap >= np;
    if (lean) {
        int lst = AHEAD_OK;
        // the variant with the filter exists for 8-byte keys: the sample picks between the two
        const bool pair = lean_kw == 8;
        if (pair) {  // the sample, then the filter (built only when the sample found a missing key)
            hipLaunchKernelGGL(k_hj_sample, dim3(kHjSamples / kBlock), dim3(kBlock), 0, s, p, np);
            hipLaunchKernelGGL(k_hj_filter, grid, dim3(kBlock), 0, s, b);
        }
        lean_shape(lean_kw, !ppay.data, !bpay.data, [&](auto proto) {
            if (lst != AHEAD_OK) return;
            auto q = proto;
            q.pair_launched = pair ? 1u : 0u;
            q.n = np;
            q.t = t;
            q.key = static_cast<const typename decltype(q)::KT *>(pkey.data);
            q.key_enc = p.key.encoded;
            q.key_inv = p.key.a_inv;
            q.key_cmask = p.key.mask;
            q.key_maxu = p.key.max_u;
            q.key_err = right ? 1u : 0u;
            q.pay = static_cast<const uint64_t *>(ppay.data);
            q.pay_seq = ppay.seqbase;
            q.pay_enc = p.pay.encoded;
            q.pay_inv = p.pay.a_inv;
            q.pay_maxu = p.pay.max_u;
            q.bpay = static_cast<const uint64_t *>(bpay.data);
            q.bpay_seq = bpay.seqbase;
            q.guard_mask = p.guard_mask;
            q.key_mask = p.key_mask;
            q.build_max = p.build_max;
            q.reenc_p = p.reenc_p;
            q.reenc_b = p.reenc_b;
            q.f_p = p.f_p;
            q.f_b = p.f_b;
            q.out_h = out_h.data;
            q.out_t = out_t.data;
            q.out_p = right ? out_h.data : out_t.data;
            q.out_b = right ? out_t.data : out_h.data;
            q.ow = wout_h;
            q.out_cap = p.out_cap;
            q.a_oid = a_oid;
            q.err[0] = p.err[0];
            q.err[1] = p.err[1];
            q.cap[0] = p.cap[0];
            q.cap[1] = p.cap[1];
            q.counts = d_counts;
            q.e_index[0] = p.e_index[0];
            q.e_index[1] = p.e_index[1];
            lst = launch_map(q, ws, o_end, s);
        });
        if (lst != AHEAD_OK) return lst;
        p.map_launched = 1u;
        p.lean = 1u;
    } else if (map_ok) {
        hipLaunchKernelGGL(k_hj_sample, dim3(kHjSamples / kBlock), dim3(kBlock), 0, s, p, np);
        ProbeMapOp pm{};
        static_cast<ProbeOp &>(pm) = p;
        if ((st = launch_map(pm, ws, o_end, s)) != AHEAD_OK) return st;
        p.map_launched = 1u;
    }
    if ((st = launch_rowtile(p, np, 3, ws, s)) != AHEAD_OK) return st;
    hipLaunchKernelGGL(k_hj_status, dim3(1), dim3(1), 0, s, t.flags, d_counts);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
