// groupby_an.hip -- groupbyAN, unary and binary, for gfx950 (groupbyAN.tcc:51-94 and :99-165).
//
// Group ids are numbered by first occurrence in scan order:
//  k_grp_first   every wave streams one contiguous range of rows, 8 or 16 consecutive rows per
//                lane and batch (one vector load each), against a wave-private LDS set of the keys it
//                has listed.  A key the wave has listed costs one LDS read.  A key new to the wave
//                is listed with its row; the rows a wave visits only ascend, so that row is the
//                key's first row in the range.  Every 64 listed keys (and at the end) the wave's
//                lanes publish them to the HBM hash table (htable.hpp): a key's first row is the
//                atomicMin (after a read) over the waves, and the lane that claims a key's slot
//                appends the slot to the list of distinct keys.
//  k_grp_sort    (up to kGrpFastMax distinct keys) one workgroup orders the listed slots by first
//                row: the position of a key is its group id.  It writes VGIDtoROID, the slot ->
//                group-id map and a compact dictionary (key, RGID, first row) in group order.
//  FirstsOp      (row-tile engine; only with more distinct keys) the same by an ordered
//                compaction of the rows that are their key's first row.
//  AssignOp      (streaming engine) writes VOIDtoRGID for every row and the ordered error vectors.
//                With at most kDictMax (512) groups every workgroup first copies the dictionary into an
//                LDS hash table, so a row costs its code load, one LDS lookup and its store;
//                otherwise a row looks its key up in the HBM table.
// Keys: unary static_cast<tunenc_t>(decoded tail) (:79); binary
// (largerTail_t)t * numGroups + g in the larger type's width (:144).  The binary dictionary
// stores RGIDs as oid_t (:122), so a repeated key yields (uint32)(gid * A_OID) (:152).
#include "htable.hpp"
#include "maptile.hpp"

using namespace ahead_dev;

namespace {

constexpr uint32_t kGrpFastMax = 8192;      // distinct keys the one-workgroup ordering takes
constexpr uint32_t kDictMax = 512;          // groups the LDS dictionary of AssignOp takes
constexpr uint32_t kDictBits = 10;
constexpr uint32_t kDictSlots = 1u << kDictBits;  // at most half full; slot kDictSlots: the all-ones key
constexpr uint32_t kGfBits = 10;
constexpr uint32_t kGfSlots = 1u << kGfBits;  // per-wave key set of k_grp_first (key + 1, 0 = free)
constexpr uint32_t kGfFill = kGfSlots / 2;    // keys in the set before it is cleared
constexpr uint32_t kGfDirectWords = 2048;     // bitmap set of unary keys below 2^16 (one bit a key)
constexpr uint32_t kDirectKeys = 1024;        // AssignOp's LDS table indexed by the key: keys below 2^10
// flags: [0] table overflow, [8] distinct keys claimed, [9] more than kGrpFastMax (compaction path)

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
    // decoded tail of a code of element type T: 32-bit arithmetic for codes of up to 4 bytes
    template <class T>
    __device__ __forceinline__ uint64_t dec_t(T code) const {
        if constexpr (sizeof(T) <= 2) {  // code and inverse below 2^16: the 24-bit multiply is exact
            const uint32_t v = static_cast<uint32_t>(code);
            return tail.encoded ? (__umul24(v, static_cast<uint32_t>(tail.a_inv)) & static_cast<uint32_t>(tail.mask)) : v;
        } else if constexpr (sizeof(T) == 4) {
            const uint32_t v = static_cast<uint32_t>(code);
            return tail.encoded ? ((v * static_cast<uint32_t>(tail.a_inv)) & static_cast<uint32_t>(tail.mask)) : v;
        } else {
            return tail.dec_of(code);
        }
    }
    template <class T, bool BIN>
    __device__ __forceinline__ uint64_t key_t(T code, uint64_t grp_code) const {
        const uint64_t t = dec_t<T>(code);
        if constexpr (!BIN) {
            if constexpr (sizeof(T) <= 4) return static_cast<uint32_t>(t) & static_cast<uint32_t>(tail_mask);
            return t & tail_mask;
        } else {
            const uint64_t g = grp_code * grp.a_inv;
            return ((t & lmask) * (num_groups & lmask) + (g & lmask)) & lmask;
        }
    }
    template <class T>
    __device__ __forceinline__ bool tail_bad_t(T code) const {
        if constexpr (sizeof(T) <= 4) {
            return tail.encoded && static_cast<uint32_t>(dec_t<T>(code)) > static_cast<uint32_t>(tail.max_u);
        } else {
            return tail.bad_of(code);
        }
    }
};

struct GrpTable {
    HTable ht;
    uint32_t *first;  // [size + 1]
    uint32_t *gid;    // [size + 1]
};

// R consecutive elements of type T, held as the 32-bit words they were loaded as: vector loads
// (the column 16-byte aligned, the run in range, r0 a multiple of R) or element loads (0 past end).
template <class T, int R>
struct Run {
    static constexpr int W = R * static_cast<int>(sizeof(T)) / 4;
    static_assert(W == 2 || W % 4 == 0, "runs of 8 or a multiple of 16 bytes");
    uint32_t w[W];
    __device__ __forceinline__ void load_vec(const T *p, uint64_t r0) {
        if constexpr (W == 2) {
            const uint2 x = *reinterpret_cast<const uint2 *>(p + r0);
            w[0] = x.x;
            w[1] = x.y;
        } else {
#pragma unroll
            for (int i = 0; i < W / 4; ++i) {
                const uint4 x = reinterpret_cast<const uint4 *>(p + r0)[i];
                w[4 * i] = x.x;
                w[4 * i + 1] = x.y;
                w[4 * i + 2] = x.z;
                w[4 * i + 3] = x.w;
            }
        }
    }
    __device__ __forceinline__ void load_elem(const T *p, uint64_t r0, uint64_t end) {
#pragma unroll
        for (int i = 0; i < W; ++i) w[i] = 0u;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint64_t v = r0 + j < end ? static_cast<uint64_t>(p[r0 + j]) : 0ull;
            if constexpr (sizeof(T) == 8) {
                w[2 * j] = static_cast<uint32_t>(v);
                w[2 * j + 1] = static_cast<uint32_t>(v >> 32);
            } else {
                constexpr int per = 4 / static_cast<int>(sizeof(T));
                w[j / per] |= static_cast<uint32_t>(v) << (8 * sizeof(T) * (j % per));
            }
        }
    }
    __device__ __forceinline__ T get(int j) const {
        if constexpr (sizeof(T) == 8) {
            return (static_cast<uint64_t>(w[2 * j + 1]) << 32) | w[2 * j];
        } else {
            constexpr int per = 4 / static_cast<int>(sizeof(T));
            return static_cast<T>(w[j / per] >> (8 * sizeof(T) * (j % per)));
        }
    }
};

// Key word of the per-wave set: 32 bits for unary keys of codes up to 4 bytes.  Runs: 64 bytes of
// codes per lane (at most 32 rows), 8 rows where a row brings 8 bytes or more.
template <class T, bool BIN>
struct GfTraits {
    using K = std::conditional_t<(!BIN && sizeof(T) <= 4), uint32_t, uint64_t>;
    static constexpr int R = (BIN || sizeof(T) == 8) ? 8 : (sizeof(T) == 4 ? 16 : 32);
};

__device__ __forceinline__ uint32_t gf_slot(uint64_t key) {
    const uint32_t x = static_cast<uint32_t>(key) ^ static_cast<uint32_t>(key >> 32);
    return (x * 0x9E3779B1u) >> (32 - kGfBits);
}

__device__ __forceinline__ uint32_t dict_slot(uint64_t key) {
    const uint32_t x = static_cast<uint32_t>(key) ^ static_cast<uint32_t>(key >> 32);
    return (x * 0x9E3779B1u) >> (32 - kDictBits);
}

template <class K>
__device__ __forceinline__ K readlane_k(K v, uint32_t l) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), l));
    if constexpr (sizeof(K) == 4) {
        return lo;
    } else {
        const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), l));
        return (static_cast<uint64_t>(hi) << 32) | lo;
    }
}

// A key with a first row into the HBM table.  Reads before the atomic: waves run roughly in row
// order, so most find an earlier first row and make no atomic at all.
__device__ __forceinline__ void grp_publish(const GrpTable &t, uint32_t *flags, uint32_t *list, uint64_t key,
                                            uint32_t row) {
    bool fresh = false;
    uint32_t gs = t.ht.find(key);
    if (gs == UINT32_MAX) gs = t.ht.insert(key, &fresh);
    if (gs == UINT32_MAX) return;  // table overflow (flags[0], reported in the status)
    if (fresh) {
        const uint32_t i = atomicAdd(&flags[8], 1u);
        if (i < kGrpFastMax) list[i] = gs;
    }
    if (t.first[gs] > row) atomicMin(&t.first[gs], row);
}

// DIRECT (unary keys below 2^16): the wave's set is a bitmap over every key, so a row costs one
// LDS read and a bit test, and the set never fills.
template <class T, bool BIN, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_grp_first(GrpKey k, GrpTable t, uint32_t *flags, uint32_t *list, uint64_t n,
                                                      uint64_t per_wave, uint32_t vec) {
    using K = typename GfTraits<T, BIN>::K;
    constexpr int R = GfTraits<T, BIN>::R;
    static_assert(!DIRECT || sizeof(K) == 4, "direct sets take 32-bit keys");
    constexpr K kOnes = static_cast<K>(~static_cast<K>(0));  // the key whose key + 1 is 0 (free)
    constexpr uint32_t kWords = DIRECT ? kGfDirectWords : kGfSlots;
    __shared__ K s_set[kWaves * kWords];
    const uint32_t lane = threadIdx.x & 63u;
    K *set = s_set + (threadIdx.x >> 6) * kWords;
    for (uint32_t i = lane; i < kWords; i += 64u) set[i] = 0;
    const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    const uint64_t begin = gw * per_wave;
    const uint64_t end = begin + per_wave < n ? begin + per_wave : n;
    const T *tail = static_cast<const T *>(k.tail.data);
    const uint64_t *grp = static_cast<const uint64_t *>(k.grp.data);
    constexpr uint64_t kStep = 64u * R;
    // unary keys of codes up to 4 bytes: key = code * ka & km (ka = 1 for plain codes)
    const uint32_t ka = k.tail.encoded ? static_cast<uint32_t>(k.tail.a_inv) : 1u;
    const uint32_t km = (k.tail.encoded ? static_cast<uint32_t>(k.tail.mask) : ~0u) & static_cast<uint32_t>(k.tail_mask);
    uint32_t fill = 0, listed = 0;  // wave-uniform
    bool ones = false;              // wave-uniform: kOnes is listed
    K my_key = 0;                   // lane i holds the wave's i-th listed key and its first row
    uint32_t my_row = 0;
    auto keys_of = [&](const Run<T, R> &c, const Run<uint64_t, R> &g, K (&key)[R]) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if constexpr (!BIN && sizeof(T) <= 2) {  // both factors below 2^16: the 24-bit multiply is exact
                key[j] = static_cast<K>(__umul24(static_cast<uint32_t>(c.get(j)), ka) & km);
            } else if constexpr (!BIN && sizeof(T) == 4) {
                key[j] = static_cast<K>((static_cast<uint32_t>(c.get(j)) * ka) & km);
            } else if constexpr (BIN) {
                key[j] = static_cast<K>(k.key_t<T, BIN>(c.get(j), g.get(j)));
            } else {
                key[j] = static_cast<K>(k.key_t<T, BIN>(c.get(j), 0ull));
            }
        }
    };
    // one batch: the lanes' keys against the set, then the keys new to the wave
    auto examine = [&](const K (&key)[R], uint64_t base, uint64_t r0) {
        // the common case first: a whole run whose keys the wave has all listed.  Every set word is
        // read before any is tested (no per-row branches); nothing of the keys outlives this test.
        bool all = r0 + R <= end;
        if constexpr (DIRECT) {
            uint32_t w[R];
#pragma unroll
            for (int j = 0; j < R; ++j) w[j] = set[static_cast<uint32_t>(key[j]) >> 5];
            uint32_t acc = 1u;  // bit 0 stays set while every key's bit is
#pragma unroll
            for (int j = 0; j < R; ++j) acc &= w[j] >> (static_cast<uint32_t>(key[j]) & 31u);
            all = all && (acc & 1u) != 0u;
        } else {
            K w[R];
#pragma unroll
            for (int j = 0; j < R; ++j) w[j] = set[gf_slot(key[j])];
#pragma unroll
            for (int j = 0; j < R; ++j) all = all && (key[j] == kOnes ? ones : w[j] == static_cast<K>(key[j] + 1));
        }
        if (__ballot(!all) == 0ull) return;
        // exact pass: the batch's codes again (cached) and its keys; bit j: row r0 + j
        K k2[R];
        {
            Run<T, R> c;
            Run<uint64_t, R> g;
            if (vec != 0u && r0 + R <= end) {
                c.load_vec(tail, r0);
                if constexpr (BIN) g.load_vec(grp, r0);
            } else {
                c.load_elem(tail, r0, end);
                if constexpr (BIN) g.load_elem(grp, r0, end);
            }
            keys_of(c, g, k2);
        }
        const uint32_t live = r0 >= end ? 0u : (end - r0 >= static_cast<uint64_t>(R) ? ~0u >> (32 - R)
                                                                                       : (1u << (end - r0)) - 1u);
        uint32_t pend = 0, probe = 0;
        if constexpr (DIRECT) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint32_t w = set[static_cast<uint32_t>(k2[j]) >> 5];
                pend |= (~(w >> (static_cast<uint32_t>(k2[j]) & 31u)) & 1u) << j;
            }
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const K w = set[gf_slot(k2[j])];
                const bool hit = k2[j] == kOnes ? ones : w == static_cast<K>(k2[j] + 1);
                pend |= static_cast<uint32_t>(!hit) << j;
                probe |= static_cast<uint32_t>(k2[j] != kOnes && w != 0) << j;
            }
        }
        pend &= live;
        probe &= pend;
        if (__ballot(pend != 0u) == 0ull) return;
        // keys stored past their home slot
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (((probe >> j) & 1u) == 0u) continue;
            uint32_t s = gf_slot(k2[j]);
            for (uint32_t p = 1; p < kGfSlots; ++p) {
                s = (s + 1u) & (kGfSlots - 1u);
                const K w = set[s];
                if (w == static_cast<K>(k2[j] + 1)) {
                    pend &= ~(1u << j);
                    break;
                }
                if (w == 0) break;
            }
        }
        // keys new to the wave, one per step: the lowest lane with a pending row holds the batch's
        // smallest pending row (lane l's rows all precede lane l + 1's), and a key is pending in
        // every row of the batch that holds it, so that row is the key's first in the batch.
        while (true) {
            const uint64_t m = __ballot(pend != 0u);
            if (m == 0ull) break;
            const uint32_t L = static_cast<uint32_t>(__builtin_ctzll(m));
            const uint32_t mj = pend != 0u ? static_cast<uint32_t>(__builtin_ctz(pend)) : 0u;
            K mk = key[0];
#pragma unroll
            for (int j = 1; j < R; ++j) mk = mj == static_cast<uint32_t>(j) ? k2[j] : mk;
            const K k0 = readlane_k<K>(mk, L);
            const uint32_t row0 = static_cast<uint32_t>(base) + L * static_cast<uint32_t>(R) +
                                  static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mj), L));
#pragma unroll
            for (int j = 0; j < R; ++j) pend &= k2[j] == k0 ? ~(1u << j) : ~0u;
            if constexpr (DIRECT) {
                if (lane == 0) set[static_cast<uint32_t>(k0) >> 5] |= 1u << (static_cast<uint32_t>(k0) & 31u);
            } else if (k0 == kOnes) {
                ones = true;
            } else {
                if (fill == kGfFill) {  // the set is half full: start it again (keys may then be listed twice)
                    for (uint32_t i = lane; i < kGfSlots; i += 64u) set[i] = 0;
                    fill = 0;
                }
                if (lane == 0) {
                    uint32_t s = gf_slot(k0);
                    while (set[s] != 0) s = (s + 1u) & (kGfSlots - 1u);
                    set[s] = static_cast<K>(k0 + 1);
                }
                ++fill;
            }
            if (lane == listed) {
                my_key = k0;
                my_row = row0;
            }
            if (++listed == 64u) {
                grp_publish(t, flags, list, static_cast<uint64_t>(my_key), my_row);
                listed = 0;
            }
        }
    };
    const uint64_t rl = begin + static_cast<uint64_t>(lane) * R;  // this lane's first row
    // whole batches of an aligned column: vector loads, the next batch's issued as soon as this
    // batch's keys are taken.  The reload is unconditional (past the last whole batch it reads that
    // batch again), so that no path through the loop leaves a different set of loads in flight.
    const uint64_t nfull = (vec != 0u && end > begin) ? (end - begin) / kStep : 0u;
    Run<T, R> ca;
    Run<uint64_t, R> ga;
    if (nfull > 0) {
        ca.load_vec(tail, rl);
        if constexpr (BIN) ga.load_vec(grp, rl);
    }
    uint64_t b = 0;
    K key[R];
    for (; b < nfull; ++b) {
        keys_of(ca, ga, key);
        const uint64_t nb = b + 1 < nfull ? b + 1 : b;
        ca.load_vec(tail, rl + nb * kStep);
        if constexpr (BIN) ga.load_vec(grp, rl + nb * kStep);
        examine(key, begin + b * kStep, rl + b * kStep);
    }
    // the last, partial batch (every batch of an unaligned column): element loads
    for (uint64_t base = begin + b * kStep; base < end; base += kStep) {
        const uint64_t r0 = base + static_cast<uint64_t>(lane) * R;
        ca.load_elem(tail, r0, end);
        if constexpr (BIN) ga.load_elem(grp, r0, end);
        keys_of(ca, ga, key);
        examine(key, base, r0);
    }
    if (lane < listed) grp_publish(t, flags, list, static_cast<uint64_t>(my_key), my_row);
}

struct GrpSortArgs {
    GrpTable t;
    uint32_t *flags;
    const uint32_t *list;
    DCol head;
    uint32_t binary;
    uint64_t a_oid;
    uint64_t *groid;
    uint64_t groid_cap;
    uint64_t *counts;
    uint64_t *dkey, *drg;  // dictionary in group order: key, RGID (gid * A_OID), first row
    uint32_t *dfirst;
};

// One workgroup: orders the listed keys by first row (bitonic sort in LDS); a key's position is
// its group id.  More than kGrpFastMax keys: raises flags[9] for the compaction path.
__global__ __launch_bounds__(1024) void k_grp_sort(GrpSortArgs a) {
    __shared__ unsigned long long s_v[kGrpFastMax];
    const uint32_t d = a.flags[8];
    if (threadIdx.x == 0) a.counts[5] = a.flags[0] != 0u ? 1u : 0u;  // the table overflowed: retry larger
    if (d > kGrpFastMax) {
        if (threadIdx.x == 0) a.flags[9] = 1u;
        return;
    }
    uint32_t p2 = 1;
    while (p2 < d) p2 <<= 1;
    for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
        unsigned long long v = ~0ull;
        if (i < d) {
            const uint32_t slot = a.list[i];
            v = (static_cast<unsigned long long>(a.t.first[slot]) << 32) | slot;
        }
        s_v[i] = v;
    }
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
        a.dkey[g] = slot == a.t.ht.size() ? ~0ull : static_cast<uint64_t>(a.t.ht.keys[slot]) - 1ull;
        a.drg[g] = static_cast<uint64_t>(g) * a.a_oid;
        a.dfirst[g] = static_cast<uint32_t>(row);
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

// VOIDtoRGID of every row and the error vectors.  LDS table (kFilterWords): words 2s and 2s + 1
// hold slot s's key + 1 (0 = free; slot kDictSlots: 1 when the all-ones key is a group) and its
// RGID; binary: then the first rows, two per word.
// DIRECT (unary keys below kDirectKeys): word `key` holds the key's RGID.
template <class T, bool BIN, bool DIRECT>
struct AssignOp {
    static_assert(!DIRECT || !BIN, "direct tables take unary keys");
    static constexpr int kErr = 3;  // head, tail, grouping (the last only for the binary form)
    static constexpr bool kMulti = false, kMain = false, kAcc = false;
    static constexpr int kFilterWords =
        DIRECT ? static_cast<int>(kDirectKeys) : 2 * (kDictSlots + 1) + (BIN ? (kDictSlots + 2) / 2 : 0);
    static constexpr int kMapRows = 8;  // more loads in flight than the default: a row reads few bytes
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
    const uint32_t *flags;
    const uint64_t *dkey, *drg;
    const uint32_t *dfirst;

    __device__ __forceinline__ bool use_dict() const {
        return DIRECT || (flags[9] == 0u && flags[8] <= kDictMax);
    }

    __device__ void filter_init(uint64_t *f) const {
        if constexpr (DIRECT) {  // every key of the column is a group, below kDirectKeys
            const uint32_t d = flags[8];
            for (uint32_t i = threadIdx.x; i < d; i += kBlock) f[dkey[i]] = drg[i];
            return;
        }
        if (!use_dict()) return;  // the same for the whole grid
        for (uint32_t i = threadIdx.x; i < 2u * (kDictSlots + 1u); i += kBlock) f[i] = 0ull;
        __syncthreads();
        const uint32_t d = flags[8];
        uint32_t *fr = reinterpret_cast<uint32_t *>(f + 2 * (kDictSlots + 1));
        for (uint32_t i = threadIdx.x; i < d; i += kBlock) {
            const uint64_t key = dkey[i];
            uint32_t s = kDictSlots;
            if (key == ~0ull) {
                f[2 * s] = 1ull;
            } else {
                s = dict_slot(key);
                while (atomicCAS(reinterpret_cast<unsigned long long *>(&f[2 * s]), 0ull,
                                 static_cast<unsigned long long>(key) + 1ull) != 0ull)
                    s = (s + 1u) & (kDictSlots - 1u);
            }
            f[2 * s + 1] = drg[i];
            if constexpr (BIN) fr[s] = dfirst[i];
        }
    }

    template <int G>
    __device__ void eval_f(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &acc, const uint64_t *f) const {
        constexpr int C = 4;  // rows per lookup step (and per step of the HBM-table path)
        static_assert(G % C == 0, "whole steps");
        if (!use_dict()) {
            uint64_t out[G];
            bool put[G];
#pragma unroll
            for (int q = 0; q < G; q += C) {
                uint64_t r[C], o4[C];
                bool l[C], p4[C];
                uint32_t e4[C];
#pragma unroll
                for (int i = 0; i < C; ++i) {
                    r[i] = row[q + i];
                    l[i] = live[q + i];
                }
                lookup_n<C>(r, l, e4, o4, p4);
#pragma unroll
                for (int i = 0; i < C; ++i) {
                    errs[q + i] = e4[i];
                    cnt[q + i] = 0;
                    out[q + i] = o4[i];
                    put[q + i] = p4[i];
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (put[g]) vgid[row[g]] = out[g];
            }
            return;
        }
        // every row's loads go out first (clamped rows: no per-row branches), every store last: a
        // store counts in the same counter as the loads, so a store before a later row's lookup
        // would make that lookup wait for the store
        const T *tp = static_cast<const T *>(k.tail.data);
        T tc[G];
        uint64_t gc[BIN ? G : 1];
#pragma unroll
        for (int g = 0; g < G; ++g) tc[g] = tp[live[g] ? row[g] : 0ull];
        if constexpr (BIN) {
            const uint64_t *gp = static_cast<const uint64_t *>(k.grp.data);
#pragma unroll
            for (int g = 0; g < G; ++g) gc[g] = gp[live[g] ? row[g] : 0ull];
        }
        uint64_t out[G];
        bool put[G];
        const uint32_t *fr = reinterpret_cast<const uint32_t *>(f + 2 * (kDictSlots + 1));
#pragma unroll
        for (int q = 0; q < G; q += C) {
            uint64_t hc[C];
            if (head.encoded) {
                uint64_t r[C];
                bool l[C];
#pragma unroll
                for (int i = 0; i < C; ++i) {
                    r[i] = row[q + i];
                    l[i] = live[q + i];
                }
                head.raw_n<C>(r, l, hc);
            }
            uint64_t key[C], w[C], rg[C];
            uint32_t s[C];
#pragma unroll
            for (int i = 0; i < C; ++i) {
                if constexpr (BIN) {
                    key[i] = k.key_t<T, BIN>(tc[q + i], gc[q + i]);
                } else {
                    key[i] = k.key_t<T, BIN>(tc[q + i], 0ull);
                }
                if constexpr (DIRECT) {
                    s[i] = static_cast<uint32_t>(key[i]);
                    w[i] = 1ull;
                    rg[i] = f[s[i]];
                } else {
                    s[i] = key[i] == ~0ull ? kDictSlots : dict_slot(key[i]);
                    w[i] = f[2 * s[i]];
                    rg[i] = f[2 * s[i] + 1];
                }
            }
            if constexpr (!DIRECT) {
#pragma unroll
                for (int i = 0; i < C; ++i) {  // a key past its home slot
                    if (key[i] == ~0ull) continue;
                    while (w[i] != key[i] + 1ull && w[i] != 0ull) {
                        s[i] = (s[i] + 1u) & (kDictSlots - 1u);
                        w[i] = f[2 * s[i]];
                        rg[i] = f[2 * s[i] + 1];
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const int g = q + i;
                cnt[g] = 0;
                uint32_t e = 0;
                if (head.encoded && head.bad_of(hc[i])) e |= 1u;
                if (k.tail_bad_t<T>(tc[g])) e |= 2u;
                if constexpr (BIN) {
                    if ((gc[g] * k.grp.a_inv) > k.grp.max_u) e |= 4u;
                }
                errs[g] = live[g] ? e : 0u;
                put[g] = live[g] && w[i] != 0ull;  // a key is absent only after a table overflow (in the status)
                bool first = true;
                if constexpr (BIN) first = fr[s[i]] == static_cast<uint32_t>(row[g]);
                out[g] = first ? rg[i] : static_cast<uint64_t>(static_cast<uint32_t>(rg[i]));
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (put[g]) vgid[row[g]] = out[g];
        }
    }

    // HBM-table lookups of G rows: error bits, and VOIDtoRGID (put: the row's key was found)
    template <int G>
    __device__ void lookup_n(const uint64_t (&row)[G], const bool (&live)[G], uint32_t (&errs)[G],
                             uint64_t (&out)[G], bool (&put)[G]) const {
        uint64_t tc[G], gc[G], hc[G], key[G], w[G];
        k.tail.raw_n<G>(row, live, tc);
        if (BIN) k.grp.raw_n<G>(row, live, gc);
        if (head.encoded) head.raw_n<G>(row, live, hc);
        uint32_t slot[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            key[g] = k.key_of(tc[g], BIN ? gc[g] : 0ull);
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
            errs[g] = 0;
            put[g] = false;
            out[g] = 0;
            if (!live[g]) continue;
            if (head.encoded && head.bad_of(hc[g])) errs[g] |= 1u;
            if (k.tail.bad_of(tc[g])) errs[g] |= 2u;
            if (BIN && (gc[g] * k.grp.a_inv) > k.grp.max_u) errs[g] |= 4u;
            if (slot[g] == UINT32_MAX) continue;  // only after a table overflow (reported in the status)
            const uint64_t rg = static_cast<uint64_t>(gid[g]) * a_oid;
            const bool first = fr[g] == static_cast<uint32_t>(row[g]);
            put[g] = true;
            out[g] = (BIN && !first) ? static_cast<uint64_t>(static_cast<uint32_t>(rg)) : rg;
        }
    }
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &errs, uint64_t &) const {
        if (head.bad_of(head.raw(row))) errs |= 1u;
        if (k.tail.bad_of(k.tail.raw(row))) errs |= 2u;
        if (BIN && (k.grp.raw(row) * k.grp.a_inv) > k.grp.max_u) errs |= 4u;
        const uint32_t s = t.ht.find(k.key(row));
        if (s == UINT32_MAX) return 0;  // only after a table overflow (reported in the status)
        const uint64_t rg = static_cast<uint64_t>(t.gid[s]) * a_oid;
        const bool first = t.first[s] == static_cast<uint32_t>(row);
        vgid[row] = (BIN && !first) ? static_cast<uint64_t>(static_cast<uint32_t>(rg)) : rg;
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

// The group table and the flags, cleared in one launch.
__global__ __launch_bounds__(kBlock) void k_grp_clear(uint32_t *flags, unsigned long long *keys, uint32_t *first,
                                                      uint64_t slots) {
    const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i0 < 64) flags[i0] = 0u;
    for (uint64_t i = i0; i < slots; i += static_cast<uint64_t>(gridDim.x) * kBlock) {
        keys[i] = 0ull;
        first[i] = 0xFFFFFFFFu;
    }
}

struct GrpPlan {
    const ahead_col *head;
    uint64_t n;
    uint64_t a_oid;
    uint64_t *voidtorgid;
    ahead_buf vgidtoroid;
    ahead_buf e[3];
    uint64_t *d_counts;
    ahead_hip_ws *ws;
    hipStream_t s;
    GrpKey k;
    GrpTable t;
    uint32_t *flags, *list;
    uint64_t *dkey, *drg;
    uint32_t *dfirst;
    uint64_t slots;
    size_t o_flags, o_keys, o_first, o_gid, o_list, o_dkey, o_drg, o_dfirst, o_end;
    uint32_t vec;
};

// Persistent grid of k_grp_first: resident workgroups (LDS decides), ranges of whole batches.
template <class T, bool BIN, bool DIRECT>
uint32_t first_grid(uint64_t n, uint64_t *per_wave) {
    static uint32_t resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_grp_first<T, BIN, DIRECT>),
                                                         kBlock, 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 2;
        resident = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
    }
    constexpr uint64_t step = 64u * GfTraits<T, BIN>::R;
    const uint64_t batches = (n + step - 1) / step;
    const uint64_t waves_max = static_cast<uint64_t>(resident) * kWaves;
    const uint64_t waves = batches < waves_max ? (batches == 0 ? 1 : batches) : waves_max;
    *per_wave = (batches + waves - 1) / waves * step;
    const uint64_t g = (waves + kWaves - 1) / kWaves;
    return static_cast<uint32_t>(g);
}

// D1: k_grp_first's direct bitmap; D2: AssignOp's direct table
template <class T, bool BIN, bool D1, bool D2>
int run_typed(GrpPlan p) {
    using namespace ahead_host;
    int st;
    // the streaming engine's tables (AssignOp) follow the group table; the workspace grows (and
    // may move) before any region pointer is taken
    if ((st = ws_ensure(p.ws, p.o_end + map_scratch_bytes<AssignOp<T, BIN, D2>>(p.n))) != AHEAD_OK) return st;
    ahead_hip_ws *ws = p.ws;
    p.flags = reinterpret_cast<uint32_t *>(ws_region(ws, p.o_flags));
    p.t.ht.keys = reinterpret_cast<unsigned long long *>(ws_region(ws, p.o_keys));
    p.t.ht.overflow = p.flags;
    p.t.first = reinterpret_cast<uint32_t *>(ws_region(ws, p.o_first));
    p.t.gid = reinterpret_cast<uint32_t *>(ws_region(ws, p.o_gid));
    p.list = reinterpret_cast<uint32_t *>(ws_region(ws, p.o_list));
    p.dkey = reinterpret_cast<uint64_t *>(ws_region(ws, p.o_dkey));
    p.drg = reinterpret_cast<uint64_t *>(ws_region(ws, p.o_drg));
    p.dfirst = reinterpret_cast<uint32_t *>(ws_region(ws, p.o_dfirst));
    {
        const uint64_t blocks = (p.slots + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_grp_clear, dim3(static_cast<uint32_t>(blocks < 2048 ? blocks : 2048)), dim3(kBlock), 0,
                           p.s, p.flags, p.t.ht.keys, p.t.first, p.slots);
        AHEAD_HIP_TRY(hipGetLastError());
    }
    if (p.n) {
        uint64_t per_wave = 0;
        const uint32_t grid = first_grid<T, BIN, D1>(p.n, &per_wave);
        hipLaunchKernelGGL((k_grp_first<T, BIN, D1>), dim3(grid), dim3(kBlock), 0, p.s, p.k, p.t, p.flags, p.list, p.n,
                           per_wave, p.vec);
        AHEAD_HIP_TRY(hipGetLastError());
    }
    GrpSortArgs ga{};
    ga.t = p.t;
    ga.flags = p.flags;
    ga.list = p.list;
    ga.head = dcol(*p.head);
    ga.binary = BIN;
    ga.a_oid = p.a_oid;
    ga.groid = static_cast<uint64_t *>(p.vgidtoroid.data);
    ga.groid_cap = p.vgidtoroid.data ? p.vgidtoroid.cap : 0;
    ga.counts = p.d_counts;
    ga.dkey = p.dkey;
    ga.drg = p.drg;
    ga.dfirst = p.dfirst;
    hipLaunchKernelGGL(k_grp_sort, dim3(1), dim3(1024), 0, p.s, ga);
    AHEAD_HIP_TRY(hipGetLastError());
    FirstsOp f{};
    f.n = p.n;
    f.k = p.k;
    f.t = p.t;
    f.head = ga.head;
    f.a_oid = p.a_oid;
    f.groid = ga.groid;
    f.groid_cap = ga.groid_cap;
    f.counts = p.d_counts;
    f.flags = p.flags;
    // 8192-row tiles: the launch is empty (skip) unless the keys exceeded kGrpFastMax
    if ((st = launch_rowtile<FirstsOp, 32>(f, p.n, 1, p.ws, p.s)) != AHEAD_OK) return st;
    AssignOp<T, BIN, D2> a{};
    a.n = p.n;
    a.k = p.k;
    a.t = p.t;
    a.head = ga.head;
    a.a_oid = p.a_oid;
    a.vgid = p.voidtorgid;
    for (int i = 0; i < 3; ++i) {
        a.err[i] = static_cast<uint64_t *>(p.e[i].data);
        a.cap[i] = p.e[i].data ? p.e[i].cap : 0;
    }
    a.counts = p.d_counts;
    a.flags = p.flags;
    a.dkey = p.dkey;
    a.drg = p.drg;
    a.dfirst = p.dfirst;
    if ((st = launch_map(a, p.ws, p.o_end, p.s)) != AHEAD_OK) return st;
    return AHEAD_OK;
}

template <bool BIN>
int run_width(const GrpPlan &p, uint32_t width, uint32_t key_bits) {
    if constexpr (!BIN) {  // unary keys of codes up to 4 bytes: direct sets and tables where the keys allow
        if (width <= 4 && key_bits <= 10) {
            switch (width) {
                case 1: return run_typed<uint8_t, false, true, true>(p);
                case 2: return run_typed<uint16_t, false, true, true>(p);
                default: return run_typed<uint32_t, false, true, true>(p);
            }
        }
        if (width <= 4 && key_bits <= 16) {
            switch (width) {
                case 1: return run_typed<uint8_t, false, true, false>(p);
                case 2: return run_typed<uint16_t, false, true, false>(p);
                default: return run_typed<uint32_t, false, true, false>(p);
            }
        }
    }
    switch (width) {
        case 1: return run_typed<uint8_t, BIN, false, false>(p);
        case 2: return run_typed<uint16_t, BIN, false, false>(p);
        case 4: return run_typed<uint32_t, BIN, false, false>(p);
        default: return run_typed<uint64_t, BIN, false, false>(p);
    }
}

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
    GrpPlan p{};
    p.s = static_cast<hipStream_t>(stream);
    p.n = tail->n;
    p.head = head;
    p.a_oid = a_oid;
    p.voidtorgid = voidtorgid;
    p.vgidtoroid = vgidtoroid;
    p.e[0] = e_head;
    p.e[1] = e_tail;
    p.e[2] = e_group;
    p.d_counts = d_counts;
    p.ws = ws;
    const uint32_t width = tail->data ? tail->width : 4;
    // distinct keys are bounded by the key's width: unary keys are decoded tails of the code's
    // width cut to the unencoded type, binary keys values of the larger type
    const uint32_t key_bits = binary ? larger_bits : (tail->unenc_bits < 8 * width ? tail->unenc_bits : 8 * width);
    uint32_t bits = 4;
    if (table_bits) {
        bits = table_bits < 4 ? 4 : (table_bits > 28 ? 28 : table_bits);
    } else {
        while ((1ull << bits) < 2ull * p.n && bits < 22) ++bits;
        if (key_bits + 1 < bits) bits = key_bits + 1 < 4 ? 4 : key_bits + 1;
    }
    p.slots = (1ull << bits) + 1;
    p.o_flags = 0;
    p.o_keys = 256;
    p.o_first = p.o_keys + align_up(p.slots * 8, 256);
    p.o_gid = p.o_first + align_up(p.slots * 4, 256);
    p.o_list = p.o_gid + align_up(p.slots * 4, 256);
    p.o_dkey = p.o_list + align_up(kGrpFastMax * 4ull, 256);
    p.o_drg = p.o_dkey + align_up(kGrpFastMax * 8ull, 256);
    p.o_dfirst = p.o_drg + align_up(kGrpFastMax * 8ull, 256);
    p.o_end = p.o_dfirst + align_up(kGrpFastMax * 4ull, 256);
    p.t.ht.bits = bits;

    p.k.tail = dcol(*tail);
    p.k.binary = binary;
    p.k.tail_mask = bits_mask(tail->unenc_bits);
    if (binary) {
        p.k.grp = dcol(*grouping);
        p.k.grp.max_u = grouping->max_u;
        p.k.lmask = bits_mask(larger_bits);
        p.k.num_groups = num_groups;
    }
    // vector loads of whole runs need 16-byte aligned columns
    p.vec = (reinterpret_cast<uintptr_t>(tail->data) % 16 == 0 &&
             (!binary || reinterpret_cast<uintptr_t>(grouping->data) % 16 == 0)) ? 1u : 0u;
    return binary ? run_width<true>(p, width, key_bits) : run_width<false>(p, width, key_bits);
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
