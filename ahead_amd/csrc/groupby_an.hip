// groupby_an.hip -- groupbyAN, unary and binary, for gfx950 (groupbyAN.tcc:51-94 and :99-165).
//
// Group ids are numbered by first occurrence in scan order:
//  k_grp_insert  every row's key enters an HBM hash table (htable.hpp) that keeps the key's
//                first row (atomicMin); a workgroup first de-duplicates its 8192 rows in an LDS
//                table, so a column with few distinct keys costs one global atomic per key and
//                workgroup instead of one per row.  A row reads the LDS table before it writes:
//                once a key and an earlier first row are there, it makes no atomic at all.
//  k_grp_collect / k_grp_sort   (up to kGrpFastMax distinct keys) the occupied slots are
//                gathered and ordered by first row in one workgroup: the position of a key is its
//                group id.  They write VGIDtoROID and the slot -> group-id map.
//  FirstsOp      (row-tile engine; only with more distinct keys) the same by an ordered
//                compaction of the rows that are their key's first row.
//  AssignOp      (streaming engine) writes VOIDtoRGID for every row and the ordered error vectors.
// Keys: unary static_cast<tunenc_t>(decoded tail) (:79); binary
// (largerTail_t)t * numGroups + g in the larger type's width (:144).  The binary dictionary
// stores RGIDs as oid_t (:122), so a repeated key yields (uint32)(gid * A_OID) (:152).
#include "htable.hpp"
#include "maptile.hpp"

using namespace ahead_dev;

namespace {

constexpr uint32_t kGrpRowsPerBlock = 8192;
constexpr uint32_t kGrpLdsSlots = 2048;
constexpr uint32_t kGrpFastMax = 8192;  // distinct keys the one-workgroup ordering takes
// flags: [0] table overflow, [8] distinct keys collected, [9] more than kGrpFastMax (compaction path)

struct GrpKey {
    DCol tail, grp;
    uint32_t binary;
    uint64_t tail_mask;  // unary: unencoded width of the tail type
    uint64_t lmask;      // binary: width of the larger type
    uint64_t num_groups;
    __device__ __forceinline__ uint64_t key_of(uint64_t tail_code, uint64_t grp_code) const {
        const uint64_t t = tail.dec_of(tail_code);
        if (!binary) return t & tail_mask;
        const uint64_t g = grp_code * grp.a_inv;
        return ((t & lmask) * (num_groups & lmask) + (g & lmask)) & lmask;
    }
    __device__ __forceinline__ uint64_t key(uint64_t row) const {
        return key_of(tail.raw(row), binary ? grp.raw(row) : 0ull);
    }
    // keys of G rows, their loads issued together
    template <int G>
    __device__ __forceinline__ void keys(const uint64_t (&row)[G], const bool (&live)[G], uint64_t (&out)[G]) const {
        uint64_t tc[G], gc[G];
        tail.raw_n<G>(row, live, tc);
        if (binary) grp.raw_n<G>(row, live, gc);
#pragma unroll
        for (int g = 0; g < G; ++g) out[g] = key_of(tc[g], binary ? gc[g] : 0ull);
    }
};

// Global insert of a key with its (block's) first row; reads first, so that a key already
// present with an earlier first row costs no atomic (blocks run roughly in row order, and
// otherwise every workgroup would serialise on the same few slots).
__device__ __forceinline__ void grp_global_insert(const HTable &ht, uint32_t *first, uint64_t key, uint32_t row) {
    uint32_t gs = ht.find(key);
    if (gs == UINT32_MAX) gs = ht.insert(key);
    if (gs != UINT32_MAX && first[gs] > row) atomicMin(&first[gs], row);
}

struct GrpTable {
    HTable ht;
    uint32_t *first;  // [size + 1]
    uint32_t *gid;    // [size + 1]
};

__global__ __launch_bounds__(kBlock) void k_grp_insert(GrpKey k, GrpTable t, uint64_t n) {
    __shared__ unsigned long long s_key[kGrpLdsSlots];  // key + 1, 0 = free
    __shared__ uint32_t s_first[kGrpLdsSlots];
    for (uint32_t i = threadIdx.x; i < kGrpLdsSlots; i += kBlock) {
        s_key[i] = 0ull;
        s_first[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * kGrpRowsPerBlock;
    const uint64_t r1 = r0 + kGrpRowsPerBlock < n ? r0 + kGrpRowsPerBlock : n;
    constexpr int B = 8;  // rows per thread whose key loads go out together
    for (uint64_t base = r0 + threadIdx.x; base < r1; base += B * kBlock) {
      uint64_t rows[B], keys[B];
      bool live[B];
#pragma unroll
      for (int q = 0; q < B; ++q) {
          rows[q] = base + static_cast<uint64_t>(q) * kBlock;
          live[q] = rows[q] < r1;
      }
      k.keys<B>(rows, live, keys);
#pragma unroll
      for (int q = 0; q < B; ++q) {
        if (!live[q]) continue;
        const uint64_t row = rows[q];
        const uint64_t key = keys[q];
        bool placed = false;
        if (key != ~0ull) {
            const unsigned long long w = static_cast<unsigned long long>(key) + 1ull;
            uint32_t s = static_cast<uint32_t>((key * 0x9E3779B97F4A7C15ull) >> 53);  // 11-bit slot
            for (uint32_t p = 0; p < 16; ++p) {
                unsigned long long prev = s_key[s];
                if (prev == 0ull) prev = atomicCAS(&s_key[s], 0ull, w);
                if (prev == 0ull || prev == w) {
                    // rows of a thread come in increasing order: most find an earlier first row
                    if (s_first[s] > static_cast<uint32_t>(row)) atomicMin(&s_first[s], static_cast<uint32_t>(row));
                    placed = true;
                    break;
                }
                s = (s + 1u) & (kGrpLdsSlots - 1u);
            }
        }
        if (!placed)  // LDS table crowded (or the all-ones key): go to the global table directly
            grp_global_insert(t.ht, t.first, key, static_cast<uint32_t>(row));
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kGrpLdsSlots; i += kBlock) {
        const unsigned long long w = s_key[i];
        if (w == 0ull) continue;
        grp_global_insert(t.ht, t.first, static_cast<uint64_t>(w) - 1ull, s_first[i]);
    }
}

// Occupied table slots -> list[] as (first row << 32 | slot), while at most kGrpFastMax.
__global__ __launch_bounds__(kBlock) void k_grp_collect(GrpTable t, uint32_t *flags, uint64_t *list) {
    const uint32_t slots = t.ht.size() + 1u;
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < slots; s += gridDim.x * kBlock) {
        if (t.ht.keys[s] == 0ull) continue;
        const uint32_t i = atomicAdd(&flags[8], 1u);
        if (i < kGrpFastMax) list[i] = (static_cast<uint64_t>(t.first[s]) << 32) | s;
    }
}

struct GrpSortArgs {
    GrpTable t;
    uint32_t *flags;
    const uint64_t *list;
    DCol head;
    uint32_t binary;
    uint64_t a_oid;
    uint64_t *groid;
    uint64_t groid_cap;
    uint64_t *counts;
};

// One workgroup: orders the collected keys by first row (bitonic sort in LDS); a key's position
// is its group id.  More than kGrpFastMax keys: raises flags[9] for the compaction path.
__global__ __launch_bounds__(1024) void k_grp_sort(GrpSortArgs a) {
    __shared__ unsigned long long s_v[kGrpFastMax];
    const uint32_t d = a.flags[8];
    if (d > kGrpFastMax) {
        if (threadIdx.x == 0) a.flags[9] = 1u;
        return;
    }
    uint32_t p2 = 1;
    while (p2 < d) p2 <<= 1;
    for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) s_v[i] = i < d ? a.list[i] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= p2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long x = s_v[i], y = s_v[l];
                    if ((x > y) == ((i & k) == 0)) {
                        s_v[i] = y;
                        s_v[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t g = threadIdx.x; g < d; g += blockDim.x) {
        const uint32_t slot = static_cast<uint32_t>(s_v[g]);
        const uint64_t row = s_v[g] >> 32;
        a.t.gid[slot] = g;
        // unary: iter->head() * AOID (raw head, groupbyAN.tcc:87); binary: decoded head * AOID (:155)
        const uint64_t h = a.binary ? a.head.dec_of(a.head.raw(row)) : a.head.raw(row);
        if (g < a.groid_cap) a.groid[g] = h * a.a_oid;
    }
    if (threadIdx.x == 0) a.counts[1] = d;
}

struct FirstsOp {
    static constexpr int kErr = 0;
    static constexpr bool kMulti = false, kMain = true, kAcc = false;
    using State = uint32_t;
    uint64_t n;
    GrpKey k;
    GrpTable t;
    DCol head;
    uint64_t a_oid;
    uint64_t *groid;
    uint64_t groid_cap;
    uint64_t *counts;
    const uint32_t *flags;
    // runs only when the keys were too many for k_grp_sort
    __device__ bool skip() const { return flags[9] == 0u; }
    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &, uint64_t &) const {
        const uint32_t s = t.ht.find(k.key(row));
        st = s;
        return (s != UINT32_MAX && t.first[s] == static_cast<uint32_t>(row)) ? 1u : 0u;
    }
    __device__ void emit_main(uint64_t row, const State &st, uint64_t pos, uint32_t) const {
        t.gid[st] = static_cast<uint32_t>(pos);
        // unary: iter->head() * AOID (raw head, :87); binary: decoded head * AOID (:155)
        const uint64_t h = k.binary ? head.dec_of(head.raw(row)) : head.raw(row);
        if (pos < groid_cap) groid[pos] = h * a_oid;
    }
    __device__ void emit_err(int, uint64_t, const State &, uint64_t) const {}
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *tot) const { counts[1] = tot[0]; }
};

struct AssignOp {
    static constexpr int kErr = 3;  // head, tail, grouping (the last only for the binary form)
    static constexpr bool kMulti = false, kMain = false, kAcc = false;
    using State = uint8_t;
    uint64_t n;
    GrpKey k;
    GrpTable t;
    DCol head;
    uint64_t a_oid;
    uint64_t *vgid;
    uint64_t *err[3];
    uint64_t cap[3];
    uint64_t *counts;
    static constexpr bool kBatched = true;
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        uint64_t tc[G], gc[G], hc[G], key[G], w[G];
        k.tail.raw_n<G>(row, live, tc);
        if (k.binary) k.grp.raw_n<G>(row, live, gc);
        if (head.encoded) head.raw_n<G>(row, live, hc);
        uint32_t slot[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            key[g] = k.key_of(tc[g], k.binary ? gc[g] : 0ull);
            slot[g] = key[g] == ~0ull ? t.ht.size() : t.ht.home(key[g]);
        }
        // the home slots of all G keys are read together; a key found elsewhere (a collision)
        // continues the probe on its own
#pragma unroll
        for (int g = 0; g < G; ++g) w[g] = live[g] ? t.ht.keys[slot[g]] : 0ull;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const bool at_home = key[g] == ~0ull ? w[g] != 0ull : w[g] == key[g] + 1ull;
            if (live[g] && !at_home) slot[g] = t.ht.find(key[g]);
        }
        uint32_t gid[G], fr[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const bool ok = live[g] && slot[g] != UINT32_MAX;
            gid[g] = ok ? t.gid[slot[g]] : 0u;
            fr[g] = ok ? t.first[slot[g]] : 0u;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            cnt[g] = 0;
            errs[g] = 0;
            if (!live[g]) continue;
            if (head.encoded && head.bad_of(hc[g])) errs[g] |= 1u;
            if (k.tail.bad_of(tc[g])) errs[g] |= 2u;
            if (k.binary && (gc[g] * k.grp.a_inv) > k.grp.max_u) errs[g] |= 4u;
            if (slot[g] == UINT32_MAX) continue;  // only after a table overflow (reported in the status)
            const uint64_t rg = static_cast<uint64_t>(gid[g]) * a_oid;
            const bool first = fr[g] == static_cast<uint32_t>(row[g]);
            vgid[row[g]] = (k.binary && !first) ? static_cast<uint64_t>(static_cast<uint32_t>(rg)) : rg;
        }
    }
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &errs, uint64_t &) const {
        if (head.bad_of(head.raw(row))) errs |= 1u;
        if (k.tail.bad_of(k.tail.raw(row))) errs |= 2u;
        if (k.binary && (k.grp.raw(row) * k.grp.a_inv) > k.grp.max_u) errs |= 4u;
        const uint32_t s = t.ht.find(k.key(row));
        if (s == UINT32_MAX) return 0;  // only after a table overflow (reported in the status)
        const uint64_t rg = static_cast<uint64_t>(t.gid[s]) * a_oid;
        const bool first = t.first[s] == static_cast<uint32_t>(row);
        vgid[row] = (k.binary && !first) ? static_cast<uint64_t>(static_cast<uint32_t>(rg)) : rg;
        return 0;
    }
    __device__ void emit_main(uint64_t, const State &, uint64_t, uint32_t) const {}
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *tot) const {
        counts[0] = n;
        counts[2] = tot[1];
        counts[3] = tot[2];
        counts[4] = tot[3];
    }
};

__global__ void k_grp_status(const uint32_t *overflow, uint64_t *d_counts) { d_counts[5] = *overflow ? 1u : 0u; }

int run_groupby(const ahead_col *head, const ahead_col *tail, const ahead_col *grouping, uint64_t num_groups,
                uint32_t larger_bits, uint32_t table_bits, uint64_t a_oid, uint64_t *voidtorgid, ahead_buf vgidtoroid,
                ahead_buf e_head, ahead_buf e_tail, ahead_buf e_group, uint64_t *d_counts, ahead_hip_ws *ws,
                void *stream) {
    using namespace ahead_host;
    const bool binary = grouping != nullptr;
    int st;
    if ((st = check_col(head, "groupby.head", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(tail, "groupby.tail", false, false)) != AHEAD_OK) return st;
    if (binary) {
        if ((st = check_col(grouping, "groupby.grouping", false, true)) != AHEAD_OK) return st;
        if (tail->n != grouping->n)
            return fail(AHEAD_ERR_SIZE_MISMATCH, "input BAT and existing grouping have different sizes!");
        if (grouping->width != 8) return fail(AHEAD_ERR_INVALID, "groupby: group ids are resoid (8-byte) codes");
        if (larger_bits != 16 && larger_bits != 32 && larger_bits != 64)
            return fail(AHEAD_ERR_INVALID, "groupby: larger_bits must be 16, 32 or 64");
    }
    if (head->n != tail->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "groupby: head and tail sizes differ");
    if (d_counts == nullptr || (voidtorgid == nullptr && tail->n)) return fail(AHEAD_ERR_INVALID, "groupby: NULL output");
    if (tail->n > 0xFFFFFFFEull) return fail(AHEAD_ERR_TOO_LARGE, "groupby: too many rows");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t n = tail->n;
    uint32_t bits = 4;
    if (table_bits) {
        bits = table_bits < 4 ? 4 : (table_bits > 28 ? 28 : table_bits);
    } else {
        while ((1ull << bits) < 2ull * n && bits < 22) ++bits;
    }
    const uint64_t slots = (1ull << bits) + 1;
    const size_t o_flags = 0, o_keys = 256, o_first = o_keys + align_up(slots * 8, 256),
                 o_gid = o_first + align_up(slots * 4, 256), o_list = o_gid + align_up(slots * 4, 256),
                 o_end = o_list + align_up(kGrpFastMax * 8ull, 256);
    // the streaming engine's tables (AssignOp) follow the group table
    if ((st = ws_ensure(ws, o_end + map_scratch_bytes<AssignOp>(n))) != AHEAD_OK) return st;
    GrpTable t{};
    uint32_t *flags = reinterpret_cast<uint32_t *>(ws_region(ws, o_flags));
    t.ht.keys = reinterpret_cast<unsigned long long *>(ws_region(ws, o_keys));
    t.ht.bits = bits;
    t.ht.overflow = flags;
    t.first = reinterpret_cast<uint32_t *>(ws_region(ws, o_first));
    t.gid = reinterpret_cast<uint32_t *>(ws_region(ws, o_gid));
    if ((st = fill_u32(flags, 64, 0, s)) != AHEAD_OK) return st;
    if ((st = fill_u64(reinterpret_cast<uint64_t *>(t.ht.keys), slots, 0, s)) != AHEAD_OK) return st;
    if ((st = fill_u32(t.first, slots, 0xFFFFFFFFu, s)) != AHEAD_OK) return st;

    GrpKey k{};
    k.tail = dcol(*tail);
    k.binary = binary;
    k.tail_mask = bits_mask(tail->unenc_bits);
    if (binary) {
        k.grp = dcol(*grouping);
        k.grp.max_u = grouping->max_u;
        k.lmask = bits_mask(larger_bits);
        k.num_groups = num_groups;
    }
    if (n) {
        const uint64_t blocks = (n + kGrpRowsPerBlock - 1) / kGrpRowsPerBlock;
        hipLaunchKernelGGL(k_grp_insert, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, s, k, t, n);
        AHEAD_HIP_TRY(hipGetLastError());
    }
    uint64_t *list = reinterpret_cast<uint64_t *>(ws_region(ws, o_list));
    hipLaunchKernelGGL(k_grp_collect, dim3(1024), dim3(kBlock), 0, s, t, flags, list);
    GrpSortArgs ga{};
    ga.t = t;
    ga.flags = flags;
    ga.list = list;
    ga.head = dcol(*head);
    ga.binary = binary;
    ga.a_oid = a_oid;
    ga.groid = static_cast<uint64_t *>(vgidtoroid.data);
    ga.groid_cap = vgidtoroid.data ? vgidtoroid.cap : 0;
    ga.counts = d_counts;
    hipLaunchKernelGGL(k_grp_sort, dim3(1), dim3(1024), 0, s, ga);
    AHEAD_HIP_TRY(hipGetLastError());
    FirstsOp f{};
    f.n = n;
    f.k = k;
    f.t = t;
    f.head = dcol(*head);
    f.a_oid = a_oid;
    f.groid = ga.groid;
    f.groid_cap = ga.groid_cap;
    f.counts = d_counts;
    f.flags = flags;
    if ((st = launch_rowtile(f, n, 1, ws, s)) != AHEAD_OK) return st;
    AssignOp a{};
    a.n = n;
    a.k = k;
    a.t = t;
    a.head = f.head;
    a.a_oid = a_oid;
    a.vgid = voidtorgid;
    const ahead_buf e[3] = {e_head, e_tail, e_group};
    for (int i = 0; i < 3; ++i) {
        a.err[i] = static_cast<uint64_t *>(e[i].data);
        a.cap[i] = e[i].data ? e[i].cap : 0;
    }
    a.counts = d_counts;
    if ((st = launch_map(a, ws, o_end, s)) != AHEAD_OK) return st;
    hipLaunchKernelGGL(k_grp_status, dim3(1), dim3(1), 0, s, flags, d_counts);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

}  // namespace

extern "C" int ahead_hip_groupby_an(const ahead_col *head, const ahead_col *tail, uint64_t a_oid, uint32_t table_bits,
                                    uint64_t *voidtorgid, ahead_buf vgidtoroid, ahead_buf e_head, ahead_buf e_tail,
                                    uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    return run_groupby(head, tail, nullptr, 0, 0, table_bits, a_oid, voidtorgid, vgidtoroid, e_head, e_tail,
                       ahead_buf{nullptr, 0}, d_counts, ws, stream);
}

extern "C" int ahead_hip_groupby2_an(const ahead_col *head, const ahead_col *tail, const ahead_col *grouping,
                                     uint64_t num_groups, uint32_t larger_bits, uint64_t a_oid, uint32_t table_bits,
                                     uint64_t *voidtorgid, ahead_buf vgidtoroid, ahead_buf e_head, ahead_buf e_tail,
                                     ahead_buf e_group, uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    return run_groupby(head, tail, grouping, num_groups, larger_bits, table_bits, a_oid, voidtorgid, vgidtoroid,
                       e_head, e_tail, e_group, d_counts, ws, stream);
}
