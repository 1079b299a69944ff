// groupby_an.hip -- groupbyAN, unary and binary, for gfx950 (groupbyAN.tcc:51-94 and :99-165).
//
// Group ids are numbered by first occurrence in scan order.  Three passes:
//  k_grp_insert  every row's key enters an HBM hash table (htable.hpp) that keeps the key's
//                first row (atomicMin); a workgroup first de-duplicates its 8192 rows in an LDS
//                table, so a column with few distinct keys costs one global atomic per key and
//                workgroup instead of one per row.
//  FirstsOp      (row-tile engine) rows that are their key's first row form an ordered
//                compaction: the rank of a first row is its group id.  It writes VGIDtoROID and
//                the slot -> group-id map.
//  AssignOp      (row-tile engine) writes VOIDtoRGID for every row and the ordered error vectors.
// Keys: unary static_cast<tunenc_t>(decoded tail) (:79); binary
// (largerTail_t)t * numGroups + g in the larger type's width (:144).  The binary dictionary
// stores RGIDs as oid_t (:122), so a repeated key yields (uint32)(gid * A_OID) (:152).
#include "htable.hpp"
#include "rowtile.hpp"

using namespace ahead_dev;

namespace {

constexpr uint32_t kGrpRowsPerBlock = 8192;
constexpr uint32_t kGrpLdsSlots = 2048;

struct GrpKey {
    DCol tail, grp;
    uint32_t binary;
    uint64_t tail_mask;  // unary: unencoded width of the tail type
    uint64_t lmask;      // binary: width of the larger type
    uint64_t num_groups;
    __device__ __forceinline__ uint64_t key(uint64_t row) const {
        const uint64_t t = tail.dec_of(tail.raw(row));
        if (!binary) return t & tail_mask;
        const uint64_t g = grp.raw(row) * grp.a_inv;
        return ((t & lmask) * (num_groups & lmask) + (g & lmask)) & lmask;
    }
};

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
    for (uint64_t row = r0 + threadIdx.x; row < r1; row += kBlock) {
        const uint64_t key = k.key(row);
        bool placed = false;
        if (key != ~0ull) {
            const unsigned long long w = static_cast<unsigned long long>(key) + 1ull;
            uint32_t s = static_cast<uint32_t>((key * 0x9E3779B97F4A7C15ull) >> 53);  // 11-bit slot
            for (uint32_t p = 0; p < 16; ++p) {
                const unsigned long long prev = atomicCAS(&s_key[s], 0ull, w);
                if (prev == 0ull || prev == w) {
                    atomicMin(&s_first[s], static_cast<uint32_t>(row));
                    placed = true;
                    break;
                }
                s = (s + 1u) & (kGrpLdsSlots - 1u);
            }
        }
        if (!placed) {  // LDS table crowded (or the all-ones key): go to the global table directly
            const uint32_t gs = t.ht.insert(key);
            if (gs != UINT32_MAX) atomicMin(&t.first[gs], static_cast<uint32_t>(row));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kGrpLdsSlots; i += kBlock) {
        const unsigned long long w = s_key[i];
        if (w == 0ull) continue;
        const uint32_t gs = t.ht.insert(static_cast<uint64_t>(w) - 1ull);
        if (gs != UINT32_MAX) atomicMin(&t.first[gs], s_first[i]);
    }
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
                 o_gid = o_first + align_up(slots * 4, 256), o_end = o_gid + align_up(slots * 4, 256);
    if ((st = ws_ensure(ws, o_end)) != AHEAD_OK) return st;
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
    FirstsOp f{};
    f.n = n;
    f.k = k;
    f.t = t;
    f.head = dcol(*head);
    f.a_oid = a_oid;
    f.groid = static_cast<uint64_t *>(vgidtoroid.data);
    f.groid_cap = vgidtoroid.data ? vgidtoroid.cap : 0;
    f.counts = d_counts;
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
    if ((st = launch_rowtile(a, n, 4, ws, s)) != AHEAD_OK) return st;
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
