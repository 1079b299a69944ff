// merge_an.hip -- matchjoinAN, merge form <H1, T1, H2, T2> (matchjoinAN.tcc:46-143).
//
// The reference walks two lists ascending in decoded T1 / H2 with two cursors.  Its loop nest is
// a three-state machine:
//   TOP    start of an outer iteration: stop if either list is exhausted, check list-2 row j;
//   SKIP1  while t1[i] < h2[j]: check list-1 row i, i++ (stop when list 1 ends);
//   SKIP2  while t1[i] > h2[j]: check list-2 row j, j++ (stop when list 2 ends); then on
//          t1[i] == h2[j] emit (h1[i], t2[j]) and advance both; back to TOP.
// A list-2 row is therefore checked at TOP and again when SKIP2 passes it, and a matched list-1
// row is never checked; the error vectors record exactly these visits.
//
// The GPU cuts the walk into sub-tiles along the merge path of the two lists: sub-tile b starts at
// the list-1 row i_b where the diagonal i + j = b * D of the merged order (ties: list 1 first)
// crosses, so a sub-tile spans about D rows of both lists together whatever their length ratio.
// For sorted inputs the state in which the walk first reaches row i follows from searches: with
// v = t1[i-1], r its rank in its run of equal keys in list 1 and [lb2, ub2) the run of v in list 2,
// the walk arrives by a match at (i, lb2 + r + 1) in TOP if r < ub2 - lb2, else by a skip at
// (i, ub2) in SKIP1, or never (ub2 = |list 2|: the walk ended first).  k_mm_split computes i_b and
// that state for every sub-tile.  k_mm_count stages a sub-tile's list-1 rows and the list-2 rows
// its walk reads in LDS with coalesced loads (decoded keys and error bits), gives each thread a run
// of the list-1 rows with the entry state computed the same way (searches in LDS), walks them,
// checks that every thread leaves its rows in exactly the state the next one starts from (the last
// thread: the next sub-tile's state) and writes the sub-tile's output pairs, coalesced, to a
// provisional slot at min(i_b, j_b): no earlier sub-tile emits more pairs than it consumes rows of
// either list, so the slots are disjoint and each lies at or after the pairs' final place.  A scan
// of the per-sub-tile counts gives the final offsets and k_mm_move copies every slot there.
// Sub-tiles with an error position to write (rare) or whose walk left its LDS window walk again in
// k_mm_move, writing at the final offsets directly.  Sub-tile 0 starts at (0, 0) in TOP, so if
// every check holds the pieces are the reference's walk.  If any fails (the inputs are not sorted
// in decoded value, e.g. a corrupted key), k_mm_gscan's first thread walks the whole input
// instead: slower, same results.
#include "rowtile.hpp"

namespace {

using namespace ahead_dev;

constexpr uint32_t kMmDiag = 2560;                   // merged rows per sub-tile
constexpr uint32_t kMmRowsLds = 3072;                 // staged rows of both lists (slack for runs of equal keys)
constexpr uint32_t kMmLdsBytes = kMmRowsLds * 9;      // decoded key + error bits per staged row
constexpr uint32_t kMmGroup = 256;                    // sub-tiles per group of the offset scan
constexpr uint64_t kPhTop = 0, kPhSkip1 = 1, kPhEnd = 2, kPhSkip2 = 3;  // kPhSkip2: inside a walk only
constexpr uint64_t kJMask = (1ull << 62) - 1ull;
constexpr int kMmStreams = 5;     // pairs, e1 (h1), e2 (t1), e3 (h2), e4 (t2)

__device__ __forceinline__ uint64_t st_pack(uint64_t j, uint64_t ph) { return ph == kPhEnd ? (kPhEnd << 62) : ((ph << 62) | j); }

struct MergeArgs {
    DCol h1, t1, h2, t2;
    uint64_t n1, n2;
    uint64_t f_h1, f_t2;
    uint32_t reenc_h, reenc_t, wout_h, wout_t;
    uint64_t a_oid;
    void *out_h, *out_t;
    uint64_t out_cap;
    uint64_t *err[4];
    uint64_t err_cap[4];
    const uint64_t *ti, *tstate, *tj0;  // per sub-tile: first list-1 row, entry state, first list-2 row read
    uint64_t num_sub;   // sub-tiles
    uint64_t cap1;      // list-1 rows a sub-tile may hold
    uint32_t h1_is_t1, t2_is_h2;  // a list whose two sides are one column
    uint32_t h1_from_t1, t2_from_h2;  // ... and whose codes re-encode exactly from the decoded keys
    uint64_t a_t1, a_h2;  // A of the key columns (masked to the code width)
    uint32_t *sub_cnt;  // [kMmStreams][num_sub] output of each sub-tile per stream
    uint32_t *sub_slow; // [num_sub] 1: the sub-tile writes in k_mm_move by walking again
    uint64_t *grp;      // [kMmStreams][groups] sums per group of kMmGroup sub-tiles, then their exclusive prefixes
    uint64_t num_grp;
    void *prov_h, *prov_t;  // provisional pair slots
    uint64_t prov_cap;
    uint32_t *flags;    // [0]: a piece of the walk disagreed with its successor's entry state
    uint64_t *d_counts;
    const uint32_t *fast;  // nonzero: the intersection path wrote every output (the sub-tile kernels return)
};

__device__ __forceinline__ uint64_t dec_at(const DCol &c, uint64_t i) { return c.dec_of(c.raw(i)); }

// first index in [lo, hi) whose decoded value is >= v (strict: > v); hi if there is none
__device__ uint64_t first_ge(const DCol &c, uint64_t lo, uint64_t hi, uint64_t v, bool strict) {
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        const uint64_t d = dec_at(c, mid);
        if (strict ? d <= v : d < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
// the same over decoded codes held in LDS
__device__ uint32_t first_ge_lds(const DCol &c, const uint64_t *w, uint32_t n, uint64_t v, bool strict) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint64_t d = c.dec_of(w[mid]);
        if (strict ? d <= v : d < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
// First index in [0, p] whose value is >= v, for p whose value is >= v (or p = n): gallops back
// from p, so a short run costs a load or two.
__device__ uint64_t gallop_lb(const DCol &c, uint64_t p, uint64_t v) {
    uint64_t hi = p, step = 1;
    for (;;) {
        if (hi < step) return first_ge(c, 0, hi, v, false);
        const uint64_t q = hi - step;
        if (dec_at(c, q) < v) return first_ge(c, q + 1, hi, v, false);
        hi = q;
        step <<= 1;
    }
}
// First index in [p, n) whose value is > v, galloping forward from p.
__device__ uint64_t gallop_ub(const DCol &c, uint64_t p, uint64_t n, uint64_t v) {
    uint64_t lo = p, step = 1;
    for (;;) {
        if (lo >= n) return n;
        const uint64_t q = lo + step - 1 < n ? lo + step - 1 : n - 1;
        if (dec_at(c, q) > v) return first_ge(c, lo, q, v, true);
        lo = q + 1;
        step <<= 1;
    }
}
// Entry state at list-1 row i >= 1 from the run of v = t1[i-1]: [lb1, i) in list 1, [lb2, ub2) in
// list 2.
__device__ __forceinline__ uint64_t entry_state(uint64_t i, uint64_t lb1, uint64_t lb2, uint64_t ub2, uint64_t n2) {
    const uint64_t r = i - 1 - lb1;
    if (r < ub2 - lb2) return st_pack(lb2 + r + 1, kPhTop);
    if (ub2 >= n2) return st_pack(0, kPhEnd);
    return st_pack(ub2, kPhSkip1);
}

struct SplitArgs {
    uint64_t diag;    // merged rows per tile
    uint32_t tiles;
    uint64_t *i;      // [tiles + 1] first list-1 row of each tile (i[tiles] = |list 1|)
    uint64_t *state;  // [tiles + 1] entry state at that row (state[tiles]: end)
    uint64_t *j0;     // [tiles + 1] lb2 of t1[i-1]: the first list-2 row the tile's walk can read
};

__global__ __launch_bounds__(kBlock) void k_mm_split(MergeArgs a, SplitArgs sp) {
    if (*a.fast != 0u) return;
    const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // the sub-tile kernels' flag and group sums start at zero
    if (b == 0) a.flags[0] = 0u;
    for (uint64_t k = b; k < a.num_grp * kMmStreams; k += static_cast<uint64_t>(gridDim.x) * blockDim.x) a.grp[k] = 0ull;
    if (b > sp.tiles) return;
    uint64_t i, st, j0;
    if (b == 0 || a.n1 == 0 || a.n2 == 0) {
        i = b == 0 ? 0 : a.n1;
        st = b == 0 && a.n1 && a.n2 ? st_pack(0, kPhTop) : st_pack(0, kPhEnd);
        j0 = 0;
        if (b == sp.tiles) st = st_pack(0, kPhEnd);
    } else if (b == sp.tiles) {
        i = a.n1;
        st = st_pack(0, kPhEnd);
        j0 = a.n2;
    } else {
        // merge path: the number of list-1 rows among the first d rows of the merged order
        const uint64_t d = b * sp.diag;
        uint64_t lo = d > a.n2 ? d - a.n2 : 0, hi = d < a.n1 ? d : a.n1;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (dec_at(a.t1, mid) <= dec_at(a.h2, d - mid - 1)) lo = mid + 1;
            else hi = mid;
        }
        i = lo;
        if (i == 0) {
            st = st_pack(0, kPhTop);
            j0 = 0;
        } else {
            const uint64_t v = dec_at(a.t1, i - 1);
            const uint64_t lb1 = gallop_lb(a.t1, i - 1, v);
            const uint64_t jb = d - i;  // lb2(v) <= jb for sorted inputs
            const uint64_t lb2 = gallop_lb(a.h2, jb, v);
            const uint64_t ub2 = gallop_ub(a.h2, lb2, a.n2, v);
            st = entry_state(i, lb1, lb2, ub2, a.n2);
            j0 = lb2;
        }
    }
    sp.i[b] = i;
    sp.state[b] = st;
    sp.j0[b] = j0;
}

// The walk straight from the columns: the single-thread walk, and the pieces of a sub-tile whose
// list-2 rows leave its LDS window.
struct GlobalSrc {
    const MergeArgs &a;
    __device__ uint64_t t1(uint64_t i) const { return a.t1.raw(i); }
    __device__ uint64_t h1(uint64_t i) const { return a.h1.raw(i); }
    __device__ uint64_t h2(uint64_t j) const { return a.h2.raw(j); }
    __device__ uint64_t t2(uint64_t j) const { return a.t2.raw(j); }
};
// A walk piece from the columns with ORDERED set stops (returning kAbort, which matches no entry
// state) at the first key that is smaller than the one read before it in the same list: on such
// inputs the pieces cannot be the reference's walk, and without the stop a piece could wander
// through most of list 2.
constexpr uint64_t kAbort = ~0ull;

template <bool WRITE, class Src, bool ORDERED = false>
__device__ __forceinline__ uint64_t walk(const MergeArgs &a, const Src &src, uint64_t i, uint64_t iend, uint64_t st,
                         uint32_t (&cnt)[kMmStreams], const uint64_t (&base)[kMmStreams]) {
    uint64_t ph = st >> 62, j = st & kJMask;
    if (ph == kPhEnd) return st_pack(0, kPhEnd);
    uint64_t last1 = 0, last2 = 0;
    auto err = [&](int e, uint64_t pos) {
        if constexpr (WRITE) {
            const uint64_t p = base[1 + e] + cnt[1 + e];
            if (p < a.err_cap[e]) a.err[e][p] = pos * a.a_oid;
        }
        ++cnt[1 + e];
    };
    auto check2 = [&](uint64_t jj, uint64_t hcode) {
        if (a.h2.bad_of(hcode)) err(2, jj);
        if (a.t2.encoded && a.t2.bad_of(src.t2(jj))) err(3, jj);
    };
    uint64_t hcode = 0, h2v = 0;
    if (ph == kPhSkip1) {
        hcode = src.h2(j);
        h2v = a.h2.dec_of(hcode);
        last2 = h2v;
    }
    for (;;) {
        if (ph == kPhTop) {
            if (j >= a.n2) return st_pack(0, kPhEnd);
            hcode = src.h2(j);
            h2v = a.h2.dec_of(hcode);
            if (ORDERED && h2v < last2) return kAbort;
            last2 = h2v;
            check2(j, hcode);
        }
        uint64_t tcode = src.t1(i);
        uint64_t t1v = a.t1.dec_of(tcode);
        if (ORDERED && t1v < last1) return kAbort;
        last1 = t1v;
        while (t1v < h2v) {  // SKIP1
            if (a.h1.encoded && a.h1.bad_of(src.h1(i))) err(0, i);
            if (a.t1.bad_of(tcode)) err(1, i);
            if (++i == iend) return iend == a.n1 ? st_pack(0, kPhEnd) : st_pack(j, kPhSkip1);
            tcode = src.t1(i);
            t1v = a.t1.dec_of(tcode);
            if (ORDERED && t1v < last1) return kAbort;
            last1 = t1v;
        }
        while (t1v > h2v) {  // SKIP2
            check2(j, hcode);
            if (++j >= a.n2) return st_pack(0, kPhEnd);
            hcode = src.h2(j);
            h2v = a.h2.dec_of(hcode);
            if (ORDERED && h2v < last2) return kAbort;
            last2 = h2v;
        }
        if (t1v == h2v) {
            if constexpr (WRITE) {
                const uint64_t p = base[0] + cnt[0];
                if (p < a.out_cap) {
                    const uint64_t h = src.h1(i), t = src.t2(j);
                    store_w(a.out_h, p, a.wout_h, a.reenc_h ? h * a.f_h1 : h);
                    store_w(a.out_t, p, a.wout_t, a.reenc_t ? t * a.f_t2 : t);
                }
            }
            ++cnt[0];
            ++j;
            if (++i == iend) return iend == a.n1 ? st_pack(0, kPhEnd) : st_pack(j, kPhTop);
        }
        ph = kPhTop;
    }
}

constexpr uint32_t kMmThreads = 256;              // threads of the sub-tile kernels
constexpr int kMmWaves = static_cast<int>(kMmThreads / 64);
constexpr uint32_t kMmMoveSubs = 32;              // sub-tiles per block of k_mm_move (divides kMmGroup)

// One sub-tile in LDS: decoded keys of its list-1 rows [I0, I1) and list-2 rows [J0, J0 + W), and
// per row the error bits of both sides (list 1: bit 0 t1, bit 1 h1; list 2: bit 0 h2, bit 1 t2).
struct SubTile {
    uint64_t *L1, *L2;
    uint8_t *F1, *F2;
    uint64_t I0, I1, J0;
    uint32_t n1t, W;
    bool ok;
};

__device__ __forceinline__ void sub_view(const MergeArgs &a, uint64_t sub, uint8_t *lds, SubTile &T) {
    T.I0 = a.ti[sub];
    const uint64_t I1r = a.ti[sub + 1];
    T.ok = I1r >= T.I0 && I1r - T.I0 <= a.cap1;  // false only for unsorted inputs
    T.I1 = T.ok ? I1r : T.I0;
    T.n1t = static_cast<uint32_t>(T.I1 - T.I0);
    T.J0 = a.tj0[sub] < a.n2 ? a.tj0[sub] : a.n2;
    const uint64_t room = kMmRowsLds - T.n1t;
    T.W = static_cast<uint32_t>(room < a.n2 - T.J0 ? room : a.n2 - T.J0);
    T.L1 = reinterpret_cast<uint64_t *>(lds);
    T.L2 = T.L1 + T.n1t;
    T.F1 = reinterpret_cast<uint8_t *>(T.L2 + T.W);
    T.F2 = T.F1 + T.n1t;
}

// Stages n <= kMmRowsLds rows of a key column: decoded key and its error bit, plus the error bit of
// the other side of the list when it is the same column (same code).  Each thread issues all its
// loads (clamped to the last row, so unconditional) before the first LDS store waits on them.  The
// stagers return what the calling thread saw: kSumErr if it set an error bit, kSumWide if a key
// needs more than 32 bits.
constexpr int kStageU = static_cast<int>((kMmRowsLds + kMmThreads - 1) / kMmThreads);
constexpr uint32_t kSumErr = 1u, kSumWide = 2u;
template <typename T>
__device__ __forceinline__ uint32_t stage_keys_t(const DCol &key, const DCol &other, bool same, const T *p,
                                                 uint64_t base, uint32_t n, uint64_t *L, uint8_t *F) {
    T x[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = u * kMmThreads + threadIdx.x;
        x[u] = p[base + (k < n ? k : n - 1)];
    }
    uint32_t sum = 0;
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = u * kMmThreads + threadIdx.x;
        if (k < n) {
            const uint64_t c = x[u];
            const uint64_t d = key.dec_of(c);
            const uint32_t f = (key.bad_of(c) ? 1u : 0u) | (same && other.bad_of(c) ? 2u : 0u);
            L[k] = d;
            F[k] = static_cast<uint8_t>(f);
            sum |= (f ? kSumErr : 0u) | ((d >> 32) ? kSumWide : 0u);
        }
    }
    return sum;
}
// ORs the error bit of the other side (a column of its own) into F.
template <typename T>
__device__ __forceinline__ uint32_t stage_other_t(const DCol &other, const T *p, uint64_t base, uint32_t n, uint8_t *F) {
    T x[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = u * kMmThreads + threadIdx.x;
        x[u] = p[base + (k < n ? k : n - 1)];
    }
    uint32_t sum = 0;
#pragma unroll
    for (int u = 0; u < kStageU; ++u) {
        const uint32_t k = u * kMmThreads + threadIdx.x;
        if (k < n && other.bad_of(x[u])) {
            F[k] |= 2u;
            sum = kSumErr;
        }
    }
    return sum;
}
__device__ __forceinline__ uint32_t stage_list(const DCol &key, const DCol &other, bool same, uint64_t base,
                                               uint32_t n, uint64_t *L, uint8_t *F) {
    if (n == 0) return 0u;
    uint32_t sum = 0;
    if (key.data == nullptr) {  // void: seqbase + i, 32-bit oids, no error bit
        for (uint32_t k = threadIdx.x; k < n; k += kMmThreads) {
            L[k] = key.raw(base + k);
            F[k] = 0;
        }
    } else {
        switch (key.width) {
            case 1: sum = stage_keys_t(key, other, same, static_cast<const uint8_t *>(key.data), base, n, L, F); break;
            case 2: sum = stage_keys_t(key, other, same, static_cast<const uint16_t *>(key.data), base, n, L, F); break;
            case 4: sum = stage_keys_t(key, other, same, static_cast<const uint32_t *>(key.data), base, n, L, F); break;
            default: sum = stage_keys_t(key, other, same, static_cast<const uint64_t *>(key.data), base, n, L, F); break;
        }
    }
    if (same || !other.encoded) return sum;  // a void or plain side has no error bit
    switch (other.width) {
        case 1: sum |= stage_other_t(other, static_cast<const uint8_t *>(other.data), base, n, F); break;
        case 2: sum |= stage_other_t(other, static_cast<const uint16_t *>(other.data), base, n, F); break;
        case 4: sum |= stage_other_t(other, static_cast<const uint32_t *>(other.data), base, n, F); break;
        default: sum |= stage_other_t(other, static_cast<const uint64_t *>(other.data), base, n, F); break;
    }
    return sum;
}
__device__ __forceinline__ uint32_t stage_sub(const MergeArgs &a, const SubTile &T) {
    return stage_list(a.t1, a.h1, a.h1_is_t1 != 0u, T.I0, T.n1t, T.L1, T.F1) |
           stage_list(a.h2, a.t2, a.t2_is_h2 != 0u, T.J0, T.W, T.L2, T.F2);
}

// Number of keys among L[0, n) below v (STRICT: at most v), for keys in ascending order; n <= cap.
// The step count depends on cap only, so lanes of a wave with different n stay in step.
template <bool STRICT>
__device__ __forceinline__ uint32_t lds_rank(const uint64_t *L, uint32_t n, uint32_t cap, uint64_t v) {
    uint32_t pos = 0;
    for (uint32_t step = cap ? (1u << (31 - __clz(cap))) : 0u; step != 0u; step >>= 1) {
        const uint32_t k = pos + step - 1;
        const uint64_t x = L[k < n ? k : 0u];
        if (k < n && (STRICT ? x <= v : x < v)) pos += step;
    }
    return pos;
}

// Entry state at list-1 row I0 + li (1 <= li < n1t) from the staged keys; the columns are read only
// when a run of equal keys reaches past the window.
__device__ __forceinline__ uint64_t lds_state(const MergeArgs &a, const SubTile &T, uint32_t li) {
    const uint64_t v = T.L1[li - 1];
    uint64_t lb1 = T.I0 + li - 1;
    if (li < 2 || !(T.L1[li - 2] < v)) {  // v repeats: the run's first row
        const uint32_t k = lds_rank<false>(T.L1, li - 1, T.n1t, v);
        lb1 = k == 0 && T.I0 > 0 ? gallop_lb(a.t1, T.I0, v) : T.I0 + k;
    }
    const uint32_t lo = lds_rank<false>(T.L2, T.W, T.W, v);
    const bool past = T.J0 + T.W < a.n2;  // list 2 continues beyond the window
    uint64_t lb2 = T.J0 + lo, ub2 = lb2;
    const uint64_t at = lo < T.W ? T.L2[lo] : 0ull;
    const uint64_t nx = lo + 1 < T.W ? T.L2[lo + 1] : 0ull;
    if (lo == T.W) {
        if (past) {
            lb2 = first_ge(a.h2, T.J0 + T.W, a.n2, v, false);
            ub2 = gallop_ub(a.h2, lb2, a.n2, v);
        }
    } else if (at == v) {
        if (lo + 1 < T.W && nx > v) {
            ub2 = lb2 + 1;
        } else {
            const uint32_t k = lds_rank<true>(T.L2, T.W, T.W, v);
            ub2 = k == T.W && past ? first_ge(a.h2, T.J0 + T.W, a.n2, v, true) : T.J0 + k;
        }
    }
    return entry_state(T.I0 + li, lb1, lb2, ub2, a.n2);
}

// Rows per thread and threads with rows, for a sub-tile of n1t list-1 rows.
__device__ __forceinline__ void thread_rows(uint32_t n1t, uint32_t &q, uint32_t &active) {
    q = (n1t + kMmThreads - 1) / kMmThreads;
    active = q ? (n1t + q - 1) / q : 0u;
}

// Entry state of the calling thread's rows (st0) and of the next thread's (st1; the last thread
// with rows: the next sub-tile's).  Contains a barrier.
__device__ __forceinline__ void thread_states(const MergeArgs &a, const SubTile &T, uint64_t sub, uint32_t q,
                                              uint32_t active, uint64_t *s_bound, uint64_t &st0, uint64_t &st1) {
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    st0 = 0;
    if (t == 0) st0 = a.tstate[sub];
    else if (t < active) st0 = lds_state(a, T, t * q);
    st1 = __shfl_down(st0, 1, 64);
    if (lane == 0) s_bound[w] = st0;
    __syncthreads();
    if (lane == 63 && w + 1 < static_cast<uint32_t>(kMmWaves)) st1 = s_bound[w + 1];
    if (t + 1 == active) st1 = a.tstate[sub + 1];
}

__device__ __forceinline__ uint64_t enc_key(const DCol &c, uint64_t a_code, uint64_t v) {
    return c.encoded ? (v * a_code) & c.mask : v;
}

constexpr int kWalkOk = 0, kWalkOutside = 1, kWalkUnsorted = 2;

// The walk over staged rows, one step (an advance of i, of j, or of both) per iteration; the same
// state machine as walk(), with SKIP2 as a phase of its own.  Every lane runs the same step code:
// the comparison picks the advances and the next phase, error bits (rare) take the one branch, and
// the loop has a single exit.  Returns kWalkOutside if the walk needs a list-2 row outside the
// window (the caller redoes the piece from the columns, rewriting what this one wrote), and
// kWalkUnsorted if a list-1 key is smaller than the one before it.  With REC, every match also
// records its local list-2 row in mj[local list-1 row] (mj has a spare last slot that takes the
// writes of the other steps) and sets the bit of its list-1 row (from the piece's first row) in
// mbits.
template <bool WRITE, bool REC = false>
__device__ __forceinline__ int walk_lds(const MergeArgs &a, const SubTile &T, uint64_t i, uint64_t iend, uint64_t st,
                                        uint32_t (&cnt)[kMmStreams], const uint64_t (&base)[kMmStreams], uint64_t &end,
                                        uint16_t *mj = nullptr, uint32_t *mbits = nullptr) {
    const uint64_t ph0 = st >> 62, j0 = st & kJMask;
    if (ph0 == kPhEnd || (ph0 != kPhSkip1 && j0 >= a.n2)) {
        end = st_pack(0, kPhEnd);
        return kWalkOk;
    }
    if (j0 < T.J0 || j0 - T.J0 >= T.W) return kWalkOutside;
    uint32_t c[kMmStreams];
#pragma unroll
    for (int s = 0; s < kMmStreams; ++s) c[s] = cnt[s];
    auto err = [&](int e, uint64_t pos) {
        if constexpr (WRITE) {
            const uint64_t p = base[1 + e] + c[1 + e];
            if (p < a.err_cap[e]) a.err[e][p] = pos * a.a_oid;
        }
        ++c[1 + e];
    };
    const uint64_t lim2 = a.n2 - T.J0;  // local index of the end of list 2
    uint32_t lj = static_cast<uint32_t>(j0 - T.J0);
    uint32_t li = static_cast<uint32_t>(i - T.I0);
    const uint32_t li0 = li;
    const uint32_t lend = static_cast<uint32_t>(iend - T.I0);
    uint32_t ph = static_cast<uint32_t>(ph0);
    uint32_t bits = 0;
    uint64_t v1 = T.L1[li];
    bool fin, uns, endn;
    for (;;) {
        const uint64_t v2 = T.L2[lj];
        const uint32_t f1 = T.F1[li], f2 = T.F2[lj];
        const bool lt = v1 < v2, gt = v2 < v1, eq = !lt && !gt;
        const bool topchk = ph == kPhTop || (ph == kPhSkip2 && lt);
        if (((topchk || gt) && (f2 & 3u)) || (lt && (f1 & 3u))) {  // error positions, in walk order
            const uint64_t jg = T.J0 + lj, ig = T.I0 + li;
            if (topchk) {
                if (f2 & 1u) err(2, jg);
                if (f2 & 2u) err(3, jg);
            }
            if (lt) {
                if (f1 & 2u) err(0, ig);
                if (f1 & 1u) err(1, ig);
            } else if (gt) {
                if (f2 & 1u) err(2, jg);
                if (f2 & 2u) err(3, jg);
            }
        }
        if constexpr (WRITE) {
            if (eq) {
                const uint64_t p = base[0] + c[0];
                if (p < a.out_cap) {
                    const uint64_t h = a.h1_from_t1 ? enc_key(a.t1, a.a_t1, v1) : a.h1.raw(T.I0 + li);
                    const uint64_t t = a.t2_from_h2 ? enc_key(a.h2, a.a_h2, v2) : a.t2.raw(T.J0 + lj);
                    store_w(a.out_h, p, a.wout_h, a.reenc_h ? h * a.f_h1 : h);
                    store_w(a.out_t, p, a.wout_t, a.reenc_t ? t * a.f_t2 : t);
                }
            }
        }
        if constexpr (REC) {
            mj[eq ? li : kMmDiag] = static_cast<uint16_t>(lj);
            bits |= (eq ? 1u : 0u) << (li - li0);
        }
        c[0] += eq ? 1u : 0u;
        li += gt ? 0u : 1u;
        lj += lt ? 0u : 1u;
        ph = eq ? static_cast<uint32_t>(kPhTop) : (lt ? static_cast<uint32_t>(kPhSkip1) : static_cast<uint32_t>(kPhSkip2));
        fin = li == lend;
        const uint64_t nv = T.L1[fin ? li - 1 : li];
        uns = !fin && !gt && nv < v1;
        v1 = gt ? v1 : nv;
        endn = !fin && ph != kPhSkip1 && lj >= lim2;
        if (fin || uns || endn || lj >= T.W) break;
    }
    if (uns) return kWalkUnsorted;
    if (!fin && !endn) return kWalkOutside;
    end = fin ? (iend == a.n1 ? st_pack(0, kPhEnd) : st_pack(T.J0 + lj, ph)) : st_pack(0, kPhEnd);
#pragma unroll
    for (int s = 0; s < kMmStreams; ++s) cnt[s] = c[s];
    if constexpr (REC) *mbits = bits;
    return kWalkOk;
}

// Hands a per-lane integer over in a vector register as it is: the compiler can no longer see that
// it holds 0 or 1, so it does not turn it back into a lane mask (which costs scalar instructions
// in every iteration of a loop it is carried through).
__device__ __forceinline__ uint32_t vreg(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

// walk_lds<false, true> for a sub-tile whose staged rows carry no error bit (the common case): no
// error can be recorded, so the step needs neither the error bits nor the phase, and keys of K bits
// (the staged keys all fit) are compared.  Same results and return codes.
template <typename K>
__device__ __forceinline__ int walk_lean(const MergeArgs &a, const SubTile &T, uint64_t i, uint64_t iend, uint64_t st,
                                         uint32_t &matches, uint64_t &end, uint16_t *mj, uint32_t &mbits) {
    const uint64_t ph0 = st >> 62, j0 = st & kJMask;
    if (ph0 == kPhEnd || (ph0 != kPhSkip1 && j0 >= a.n2)) {
        end = st_pack(0, kPhEnd);
        return kWalkOk;
    }
    if (j0 < T.J0 || j0 - T.J0 >= T.W) return kWalkOutside;
    constexpr uint32_t S = sizeof(uint64_t) / sizeof(K);  // K words per staged key (low word first)
    const K *L1 = reinterpret_cast<const K *>(T.L1);
    const K *L2 = reinterpret_cast<const K *>(T.L2);
    const uint64_t rest2 = a.n2 - T.J0;
    const uint32_t lim2 = static_cast<uint32_t>(rest2 < 0xFFFFFFFFull ? rest2 : 0xFFFFFFFFull);
    const uint32_t W = T.W;
    uint32_t lj = static_cast<uint32_t>(j0 - T.J0);
    uint32_t li = static_cast<uint32_t>(i - T.I0);
    const uint32_t li0 = li;
    const uint32_t lend = static_cast<uint32_t>(iend - T.I0);
    uint32_t m = 0, bits = 0;
    K v1 = L1[li * S];
    // The loop state lives in integers: a per-lane flag that stays a lane mask across the loop
    // costs scalar instructions every iteration.
    uint32_t code, ltv;
    do {
        const K v2 = L2[lj * S];
        const uint32_t lt = v1 < v2 ? 1u : 0u;
        const uint32_t gt = v2 < v1 ? 1u : 0u;
        const uint32_t eq = 1u - lt - gt;
        mj[eq ? li : kMmDiag] = static_cast<uint16_t>(lj);
        bits |= eq << (li - li0);
        m += eq;
        li += 1u - gt;
        lj += 1u - lt;
        const uint32_t fin = li == lend ? 1u : 0u;
        const K nv = L1[(li - fin) * S];
        const uint32_t uns = (fin | gt) == 0u && nv < v1 ? 1u : 0u;
        v1 = gt ? v1 : nv;
        const uint32_t endn = (fin | lt) == 0u && lj >= lim2 ? 1u : 0u;
        code = vreg(fin | (uns << 1) | (endn << 2) | (lj >= W ? 8u : 0u));
        ltv = vreg(lt);
    } while (code == 0u);
    if (code & 2u) return kWalkUnsorted;
    if ((code & 5u) == 0u) return kWalkOutside;
    end = (code & 1u) ? (iend == a.n1 ? st_pack(0, kPhEnd) : st_pack(T.J0 + lj, ltv ? kPhSkip1 : kPhTop))
                      : st_pack(0, kPhEnd);
    matches += m;
    mbits = bits;
    return kWalkOk;
}

// One thread's rows of a sub-tile, from the staged rows or (outside the window, or unsorted) the
// columns, in order-checked form.
template <bool WRITE>
__device__ __forceinline__ uint64_t walk_piece(const MergeArgs &a, const SubTile &T, uint64_t i, uint64_t iend,
                                               uint64_t st, uint32_t (&cnt)[kMmStreams],
                                               const uint64_t (&base)[kMmStreams]) {
    uint64_t end;
    if (walk_lds<WRITE>(a, T, i, iend, st, cnt, base, end) == kWalkOk) return end;
    return walk<WRITE, GlobalSrc, true>(a, GlobalSrc{a}, i, iend, st, cnt, base);
}

__device__ __forceinline__ uint64_t slot_of(uint64_t I0, uint64_t st) {
    const uint64_t j = st & kJMask;
    return I0 < j ? I0 : j;
}

// One block per sub-tile: stage, derive the entry states, walk (checking the order of the list-1
// keys on the way), count, and write the pairs to the sub-tile's provisional slot.
__global__ __launch_bounds__(kMmThreads) void k_mm_count(MergeArgs a) {
    __shared__ __align__(16) uint8_t s_lds[kMmLdsBytes];
    __shared__ uint16_t s_mj[kMmDiag + 1];  // local list-2 row matched by each local list-1 row (+ a spare)
    __shared__ uint16_t s_os[kMmDiag];  // local list-1 row of each output pair
    __shared__ uint64_t s_bound[kMmWaves];
    __shared__ uint32_t s_red[kMmWaves];
    if (*a.fast != 0u) return;
    const uint64_t sub = blockIdx.x;
    const uint32_t t = threadIdx.x;
    SubTile T;
    sub_view(a, sub, s_lds, T);
    uint32_t q, active;
    thread_rows(T.n1t, q, active);
    const uint32_t sum = stage_sub(a, T);
    const bool dirty = __syncthreads_or((sum & kSumErr) != 0u) != 0;
    const bool wide = __syncthreads_or((sum & kSumWide) != 0u) != 0;
    uint64_t st0, st1;
    thread_states(a, T, sub, q, active, s_bound, st0, st1);
    bool bad = t == 0 && (!T.ok || (active == 0 && a.tstate[sub] != a.tstate[sub + 1]));
    const uint64_t zero[kMmStreams] = {0, 0, 0, 0, 0};
    uint32_t cnt[kMmStreams] = {0, 0, 0, 0, 0};
    uint32_t mbits = 0;
    bool over = false;
    if (t < active) {
        const uint64_t i = T.I0 + static_cast<uint64_t>(t) * q;
        const uint64_t iend = i + q < T.I1 ? i + q : T.I1;
        const uint64_t prev = t > 0 ? T.L1[t * q - 1] : (T.I0 > 0 ? dec_at(a.t1, T.I0 - 1) : 0ull);
        uint64_t end = kAbort;
        int r;
        if (T.L1[t * q] < prev) r = kWalkUnsorted;
        else if (dirty) r = walk_lds<false, true>(a, T, i, iend, st0, cnt, zero, end, s_mj, &mbits);
        else if (wide) r = walk_lean<uint64_t>(a, T, i, iend, st0, cnt[0], end, s_mj, mbits);
        else r = walk_lean<uint32_t>(a, T, i, iend, st0, cnt[0], end, s_mj, mbits);
        if (r == kWalkOutside) {
            cnt[0] = 0;
            over = true;
            end = walk<false, GlobalSrc, true>(a, GlobalSrc{a}, i, iend, st0, cnt, zero);
        }
        if (r == kWalkUnsorted || end != st1) bad = true;
    }
    if (bad) atomicOr(&a.flags[0], 1u);
    uint32_t tot[kMmStreams] = {0, 0, 0, 0, 0};
    const uint32_t ex0 = block_excl_scan<uint32_t, kMmWaves>(cnt[0], tot[0], s_red);
    const bool slow = __syncthreads_or((cnt[1] | cnt[2] | cnt[3] | cnt[4]) != 0u || over) != 0;
    if (slow) {
#pragma unroll
        for (int s = 1; s < kMmStreams; ++s) tot[s] = block_sum<uint32_t, kMmWaves>(cnt[s], s_red);
    }
    if (t == 0) {
#pragma unroll
        for (int s = 0; s < kMmStreams; ++s) {
            a.sub_cnt[static_cast<uint64_t>(s) * a.num_sub + sub] = tot[s];
            if (tot[s]) atomicAdd(reinterpret_cast<unsigned long long *>(&a.grp[s * a.num_grp + sub / kMmGroup]),
                                  static_cast<unsigned long long>(tot[s]));
        }
        a.sub_slow[sub] = slow ? 1u : 0u;
    }
    if (slow || tot[0] == 0) return;
    // pairs in walk order: the matched list-1 rows, ascending
    for (uint32_t k = ex0, m = mbits; m != 0u; m &= m - 1u, ++k)
        s_os[k] = static_cast<uint16_t>(t * q + __ffs(static_cast<int>(m)) - 1);
    __syncthreads();
    const uint64_t slot = slot_of(T.I0, a.tstate[sub]);
    if (a.h1_from_t1 && a.t2_from_h2 && a.wout_h == 8 && a.wout_t == 8 && !a.reenc_h && !a.reenc_t) {
        // both sides of both lists are their key columns: the pair is the matched key, encoded
        uint64_t *ph = static_cast<uint64_t *>(a.prov_h), *pt = static_cast<uint64_t *>(a.prov_t);
        for (uint32_t p = t; p < tot[0]; p += kMmThreads) {
            const uint64_t pos = slot + p;
            if (pos >= a.prov_cap) break;
            const uint64_t v = T.L1[s_os[p]];
            ph[pos] = enc_key(a.t1, a.a_t1, v);
            pt[pos] = enc_key(a.h2, a.a_h2, v);
        }
        return;
    }
    for (uint32_t p = t; p < tot[0]; p += kMmThreads) {
        const uint64_t pos = slot + p;
        if (pos >= a.prov_cap) break;
        const uint32_t li = s_os[p], lj = s_mj[li];
        const uint64_t h = a.h1_from_t1 ? enc_key(a.t1, a.a_t1, T.L1[li]) : a.h1.raw(T.I0 + li);
        const uint64_t v = a.t2_from_h2 ? enc_key(a.h2, a.a_h2, T.L2[lj]) : a.t2.raw(T.J0 + lj);
        store_w(a.prov_h, pos, a.wout_h, a.reenc_h ? h * a.f_h1 : h);
        store_w(a.prov_t, pos, a.wout_t, a.reenc_t ? v * a.f_t2 : v);
    }
}

// Exclusive prefixes of the group sums, per stream (one block), and the totals.
__global__ __launch_bounds__(kBlock) void k_mm_gscan(MergeArgs a) {
    __shared__ uint64_t s_red[kWaves];
    if (*a.fast != 0u) return;
    for (int s = 0; s < kMmStreams; ++s) {
        uint64_t *g = a.grp + s * a.num_grp;
        uint64_t carry = 0;
        for (uint64_t r = 0; r < a.num_grp; r += kBlock) {
            const uint64_t k = r + threadIdx.x;
            const uint64_t v = k < a.num_grp ? g[k] : 0ull;
            uint64_t tot;
            const uint64_t ex = block_excl_scan<uint64_t>(v, tot, s_red);
            if (k < a.num_grp) g[k] = carry + ex;
            carry += tot;
        }
        if (threadIdx.x == 0 && a.flags[0] == 0u) a.d_counts[s] = carry;
    }
    // inputs whose sub-tile entry states could not be derived: the whole walk on one thread (the
    // sub-tile outputs are not moved then: k_mm_move returns)
    if (threadIdx.x == 0 && a.flags[0] != 0u) {
        uint32_t cnt[kMmStreams] = {0, 0, 0, 0, 0};
        const uint64_t zero[kMmStreams] = {0, 0, 0, 0, 0};
        if (a.n1 && a.n2) (void) walk<true>(a, GlobalSrc{a}, 0, a.n1, st_pack(0, kPhTop), cnt, zero);
#pragma unroll
        for (int st = 0; st < kMmStreams; ++st) a.d_counts[st] = cnt[st];
    }
}

// Copies m provisional pairs from slot src to the final offset dst (one wave, 4 pairs per lane in
// flight).
template <typename TH, typename TT>
__device__ __forceinline__ void copy_pairs(const MergeArgs &a, uint64_t src, uint64_t dst, uint32_t m, uint32_t lane) {
    const TH *ph = static_cast<const TH *>(a.prov_h);
    const TT *pt = static_cast<const TT *>(a.prov_t);
    TH *oh = static_cast<TH *>(a.out_h);
    TT *ot = static_cast<TT *>(a.out_t);
    for (uint32_t p0 = 0; p0 < m; p0 += 256) {
        TH h[4];
        TT v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t p = p0 + u * 64 + lane;
            const bool ok = p < m && src + p < a.prov_cap;
            h[u] = ph[ok ? src + p : 0];
            v[u] = pt[ok ? src + p : 0];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t p = p0 + u * 64 + lane;
            if (p < m && src + p < a.prov_cap && dst + p < a.out_cap) {
                oh[dst + p] = h[u];
                ot[dst + p] = v[u];
            }
        }
    }
}
__device__ __forceinline__ void copy_pairs_w(const MergeArgs &a, uint64_t src, uint64_t dst, uint32_t m, uint32_t lane) {
    if (a.wout_h == 8 && a.wout_t == 8) {
        copy_pairs<uint64_t, uint64_t>(a, src, dst, m, lane);
        return;
    }
    for (uint32_t p = lane; p < m; p += 64) {
        if (src + p >= a.prov_cap || dst + p >= a.out_cap) break;
        store_w(a.out_h, dst + p, a.wout_h, load_w(a.prov_h, src + p, a.wout_h));
        store_w(a.out_t, dst + p, a.wout_t, load_w(a.prov_t, src + p, a.wout_t));
    }
}

// One block per kMmMoveSubs sub-tiles: final offsets from the group prefixes, then each wave copies
// whole provisional slots; sub-tiles marked slow are staged and walked again by the whole block,
// writing pairs and error positions at their final offsets.
__global__ __launch_bounds__(kMmThreads) void k_mm_move(MergeArgs a) {
    __shared__ __align__(16) uint8_t s_lds[kMmLdsBytes];
    __shared__ uint64_t s_off[kMmStreams][kMmMoveSubs];
    __shared__ uint32_t s_cnt0[kMmMoveSubs], s_slow[kMmMoveSubs];
    __shared__ uint64_t s_red[kMmWaves];
    __shared__ uint32_t s_red32[kMmWaves];
    __shared__ uint64_t s_bound[kMmWaves];
    if (a.flags[0] != 0u || *a.fast != 0u) return;  // the one-thread walk (or the intersection path) writes everything
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint64_t s0 = static_cast<uint64_t>(blockIdx.x) * kMmMoveSubs;
    const uint32_t nsub = static_cast<uint32_t>(a.num_sub - s0 < kMmMoveSubs ? a.num_sub - s0 : kMmMoveSubs);
    const uint64_t g = s0 / kMmGroup, gstart = g * kMmGroup;
    for (int s = 0; s < kMmStreams; ++s) {
        const uint32_t *cs = a.sub_cnt + static_cast<uint64_t>(s) * a.num_sub;
        const uint64_t v = gstart + t < s0 ? cs[gstart + t] : 0ull;
        const uint64_t pre = a.grp[s * a.num_grp + g] + block_sum<uint64_t, kMmWaves>(v, s_red);
        if (t < 64) {
            const uint64_t c = t < nsub ? cs[s0 + t] : 0ull;
            const uint64_t inc = wave_incl_scan(c);
            if (t < nsub) s_off[s][t] = pre + inc - c;
        }
    }
    if (t < nsub) {
        s_cnt0[t] = a.sub_cnt[s0 + t];
        s_slow[t] = a.sub_slow[s0 + t];
    }
    __syncthreads();
    for (uint32_t k = wv; k < nsub; k += kMmWaves) {
        if (s_slow[k] || s_cnt0[k] == 0u) continue;
        copy_pairs_w(a, slot_of(a.ti[s0 + k], a.tstate[s0 + k]), s_off[0][k], s_cnt0[k], lane);
    }
    for (uint32_t k = 0; k < nsub; ++k) {
        if (!s_slow[k]) continue;
        const uint64_t sub = s0 + k;
        __syncthreads();  // the previous slow sub-tile's walks are done with the staged rows
        SubTile T;
        sub_view(a, sub, s_lds, T);
        uint32_t q, active;
        thread_rows(T.n1t, q, active);
        stage_sub(a, T);
        __syncthreads();
        uint64_t st0, st1;
        thread_states(a, T, sub, q, active, s_bound, st0, st1);
        const uint64_t i = T.I0 + static_cast<uint64_t>(t) * q;
        const uint64_t iend = i + q < T.I1 ? i + q : T.I1;
        const uint64_t zero[kMmStreams] = {0, 0, 0, 0, 0};
        uint32_t cnt[kMmStreams] = {0, 0, 0, 0, 0};
        if (t < active) (void) walk_piece<false>(a, T, i, iend, st0, cnt, zero);
        uint64_t base[kMmStreams];
#pragma unroll
        for (int s = 0; s < kMmStreams; ++s) {
            uint32_t tot_s;
            base[s] = s_off[s][k] + block_excl_scan<uint32_t, kMmWaves>(cnt[s], tot_s, s_red32);
        }
        if (t < active) {
            uint32_t c2[kMmStreams] = {0, 0, 0, 0, 0};
            (void) walk_piece<true>(a, T, i, iend, st0, c2, base);
        }
    }
}


// ---- the intersection path -------------------------------------------------------------------
// When both key lists are strictly ascending in decoded value and every code of h1, t1, h2 and t2
// is valid, no check of the walk fails and each list-1 key meets at most one list-2 key: the
// walk's output is the sorted intersection, (h1[i], t2[j]) for every t1[i] == h2[j] in ascending
// i, with empty error vectors.  Such inputs take streaming passes instead of the sub-tile walk:
//  k_mi_clear  list 2's first and last keys give the key range [lo2, hi2] (the path applies when
//              it spans at most kMiBitsPerRow bits per list-2 row), whose bitmap words it clears;
//  k_mi_build  a bitmap of list 2's keys over the range (bit k - lo2) and, when the pairs' tails
//              are gathered from t2, per 64-key word the list-2 row of its first key (the row of
//              a key is that plus the set bits below it);
//  k_mi_tails  the words the build waves share, ORed in;
//  k_mi_count  per wave of list 1: the rows whose key is set;
//  k_mi_scan   the waves' prefixes and the counts;
//  k_mi_write  the pairs at the waves' prefixes.
// The build and count passes check the conditions on every row they read (valid codes, each key
// above the one before it); any failure clears the path's flag, the later passes return, and the
// sub-tile kernels (which return while the flag is set) compute the outputs instead.
constexpr uint64_t kMiBitsPerRow = 16;
constexpr uint32_t kMiNone = 0xFFFFFFFFu;  // no bitmap word (words < 2^30)
constexpr int kMiG = 4;  // rows per lane per step

struct MiArgs {
    uint32_t *ok;       // [0] nonzero while the path applies
    uint64_t *range;    // [0] lo2, [1] bitmap words in use
    uint64_t *bits;     // [words_cap]
    uint32_t *first;    // [words_cap] list-2 row of the first key in each word that holds one
    uint64_t words_cap;
    uint32_t *wcnt;     // [waves1] matches per wave of list 1
    uint32_t *wpre;     // [waves1] their exclusive prefixes
    uint32_t *tail_w;   // [waves2] the last bitmap word of each build wave (kMiNone: none) ...
    uint64_t *tail_m;   // [waves2] ... and the wave's bits of it, ORed in by k_mi_tails
    uint32_t waves1, waves2;
    uint64_t per1, per2;  // rows per wave of list 1 / list 2 (multiples of 64 * kMiG)
};

// The key range of list 2 from its first and last codes (every thread of k_mi_clear computes it;
// thread 0 publishes it), and the bitmap words in use cleared.
__global__ __launch_bounds__(kBlock) void k_mi_clear(MergeArgs a, MiArgs mi, uint32_t eligible) {
    uint32_t ok = eligible && a.n1 > 0 && a.n2 > 0 ? 1u : 0u;
    uint64_t lo = 0, words = 0;
    if (ok) {
        const uint64_t c0 = a.h2.raw(0), c1 = a.h2.raw(a.n2 - 1);
        lo = a.h2.dec_of(c0);
        const uint64_t hi = a.h2.dec_of(c1);
        if (a.h2.bad_of(c0) || a.h2.bad_of(c1) || hi < lo) {
            ok = 0;
        } else {
            words = ((hi - lo) >> 6) + 1;
            if (words > mi.words_cap) ok = 0;
        }
    }
    if (!ok) words = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        mi.range[0] = lo;
        mi.range[1] = words;
        mi.ok[0] = ok;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < words; i += stride) mi.bits[i] = 0ull;
}

__device__ __forceinline__ uint64_t mi_rows_end(uint64_t begin, uint64_t per, uint64_t n) {
    return begin + per < n ? begin + per : n;
}

// Lanes below the calling one whose ballot bit is set.
__device__ __forceinline__ uint32_t mi_below(uint64_t ballot) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(ballot >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ballot), 0u));
}

// Codes of N rows of a column of K codes, loaded as K (widened where they are used, so a load's
// wait sits at its first use).  Rows that are not ok read row 0.
template <typename K, int N>
__device__ __forceinline__ void mi_load(const void *p, const uint64_t (&i)[N], const bool (&ok)[N], K (&x)[N]) {
    const K *q = static_cast<const K *>(p);
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = q[ok[k] ? i[k] : 0ull];
}

// x needs a check of its own: it is encoded and not the column y with y's check.
__device__ __forceinline__ bool mi_own_check(const DCol &x, const DCol &y) {
    return x.encoded && !(x.data == y.data && x.width == y.width && y.encoded && x.a_inv == y.a_inv &&
                          x.max_u == y.max_u && x.mask == y.mask);
}

// List 2: order and validity checks, bitmap words and first rows.  Each lane takes kMiR
// consecutive rows of a step.  A wave's keys ascend, so it holds the last word its keys reached
// (the open word) in registers until a key passes it, then stores it: every word a wave stores
// holds all of its keys from that wave.  A word can hold keys of two waves only as the first's
// last word and the second's first: the first leaves its last word to k_mi_tails (an atomic OR
// after this kernel), so no word is stored twice.  A step's words [wf, wl] are ORed together in
// the wave's LDS slice when they fit; a step spanning more words ORs the bits of the words between
// the open word and wl (the wave's alone) into HBM with atomics.
constexpr uint32_t kMiLdsWords = 512;
constexpr int kMiR = kMiG;

__device__ __forceinline__ uint64_t mi_wave_or(uint64_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x |= __shfl_xor(x, d, 64);
    return x;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_mi_build(MergeArgs a, MiArgs mi) {
    __shared__ unsigned long long s_bits[kWaves][kMiLdsWords];
    if (mi.ok[0] == 0u) return;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * kWaves + wv;
    unsigned long long *sb = s_bits[wv];
    for (uint32_t k = lane; k < kMiLdsWords; k += 64u) sb[k] = 0ull;
    const uint64_t begin = static_cast<uint64_t>(gw) * mi.per2;
    if (gw >= mi.waves2 || begin >= a.n2) return;
    const uint64_t end = mi_rows_end(begin, mi.per2, a.n2);
    const uint64_t lo = mi.range[0], words = mi.range[1];
    const bool chk_t2 = mi_own_check(a.t2, a.h2);  // t2's codes are checked too
    const bool need_first = !a.t2_from_h2;  // the write pass gathers t2 at the rows of the keys
    bool bad = false;
    bool has_prev = begin > 0;
    uint64_t prev = has_prev ? a.h2.dec_of(a.h2.raw(begin - 1)) : 0ull;  // key of the row before the step
    uint32_t ow = kMiNone;  // the open word (wave-uniform) and the wave's bits of it
    uint64_t om = 0;
    // a step's codes are loaded before the previous step's stores are issued (loads and stores
    // complete in order, so a load issued after a store waits for it)
    K cn[kMiR];
    uint64_t cn2[kMiR] = {};
    {
        uint64_t rows[kMiR];
        bool live[kMiR];
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            rows[k] = begin + static_cast<uint64_t>(lane) * kMiR + k;
            live[k] = rows[k] < end;
        }
        mi_load<K, kMiR>(a.h2.data, rows, live, cn);
        if (chk_t2) a.t2.raw_n<kMiR>(rows, live, cn2);
    }
    for (uint64_t seg = begin; seg < end; seg += 64u * kMiR) {
        uint64_t rows[kMiR], c[kMiR], c2[kMiR], v[kMiR];
        uint32_t w[kMiR];
        bool live[kMiR], in[kMiR];
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            rows[k] = seg + static_cast<uint64_t>(lane) * kMiR + k;
            live[k] = rows[k] < end;
            c[k] = cn[k];
            c2[k] = cn2[k];
        }
        {
            uint64_t nrows[kMiR];
            bool nlive[kMiR];
#pragma unroll
            for (int k = 0; k < kMiR; ++k) {
                nrows[k] = seg + 64u * kMiR + static_cast<uint64_t>(lane) * kMiR + k;
                nlive[k] = nrows[k] < end;
            }
            mi_load<K, kMiR>(a.h2.data, nrows, nlive, cn);
            if (chk_t2) a.t2.raw_n<kMiR>(nrows, nlive, cn2);
        }
        const uint64_t last_row = (end < seg + 64u * kMiR ? end : seg + 64u * kMiR) - 1u;  // wave-uniform
        const uint32_t last_lane = static_cast<uint32_t>((last_row - seg) / kMiR);
        const int last_k = static_cast<int>((last_row - seg) % kMiR);
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            v[k] = a.h2.dec_of(c[k]);
            const uint64_t off = v[k] - lo;
            in[k] = live[k] && v[k] >= lo && (off >> 6) < words;
            w[k] = in[k] ? static_cast<uint32_t>(off >> 6) : 0u;
        }
        // the key before each row: the lane's previous row, lane - 1's last row, or the step's predecessor
        const uint64_t up = __shfl_up(v[kMiR - 1], 1, 64);
        const uint64_t pv0 = lane == 0 ? prev : up;
        const bool hp0 = lane > 0 || has_prev;
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            if (!live[k]) continue;
            const uint64_t pv = k == 0 ? pv0 : v[k - 1];
            if (a.h2.bad_of(c[k]) || (chk_t2 && a.t2.bad_of(c2[k])) || !in[k]) bad = true;
            if ((k > 0 || hp0) && !(pv < v[k])) bad = true;
            const bool first_of_word = k == 0 ? (!hp0 || pv < lo || ((pv - lo) >> 6) != w[0]) : w[k - 1] != w[k];
            if (need_first && in[k] && first_of_word) mi.first[w[k]] = static_cast<uint32_t>(rows[k]);
        }
        uint32_t wl_own = w[0];
#pragma unroll
        for (int k = 1; k < kMiR; ++k) wl_own = k == last_k ? w[k] : wl_own;
        const uint32_t wf = __builtin_amdgcn_readfirstlane(w[0]);
        const uint32_t wl = __builtin_amdgcn_readlane(wl_own, static_cast<int>(last_lane));
        const uint32_t span = wl >= wf ? wl - wf + 1u : ~0u;
        if (span <= kMiLdsWords) {
#pragma unroll
            for (int k = 0; k < kMiR; ++k) {
                const uint32_t o = w[k] - wf;
                if (in[k] && o < span) atomicOr(sb + o, 1ull << ((v[k] - lo) & 63u));
            }
            const uint64_t x0 = sb[0], xl = sb[span - 1u];
            for (uint32_t o = lane; o < span; o += 64u) {
                const unsigned long long x = sb[o];
                sb[o] = 0ull;
                if (o > 0u && o + 1u < span && x != 0ull) mi.bits[wf + o] = x;  // all of the word's keys are in this step
            }
            if (ow == wf) {
                om |= x0;
            } else {
                if (ow != kMiNone && lane == 0) mi.bits[ow] = om;
                ow = wf;
                om = x0;
            }
            if (span > 1u) {
                if (lane == 0) mi.bits[ow] = om;
                ow = wl;
                om = xl;
            }
        } else if (wl > wf) {
            uint64_t mo = 0, ml = 0;
#pragma unroll
            for (int k = 0; k < kMiR; ++k) {
                if (!in[k]) continue;
                const uint64_t bit = 1ull << ((v[k] - lo) & 63u);
                if (w[k] == ow) mo |= bit;
                else if (w[k] == wl) ml |= bit;
                else atomicOr(reinterpret_cast<unsigned long long *>(mi.bits + w[k]), bit);
            }
            mo = mi_wave_or(mo);
            ml = mi_wave_or(ml);
            if (ow != kMiNone && lane == 0) mi.bits[ow] = om | mo;
            ow = wl;
            om = ml;
        } else {  // out of order (the path does not apply): nothing more to store
            bad = true;
        }
        uint64_t vl_own = v[0];
#pragma unroll
        for (int k = 1; k < kMiR; ++k) vl_own = k == last_k ? v[k] : vl_own;
        prev = __shfl(vl_own, static_cast<int>(last_lane), 64);
        has_prev = true;
    }
    if (lane == 0) {
        mi.tail_w[gw] = ow;
        mi.tail_m[gw] = om;
    }
    if (bad) mi.ok[0] = 0u;
}

// The build waves' last words, ORed into the bitmap (one thread per wave).
__global__ __launch_bounds__(kBlock) void k_mi_tails(MergeArgs a, MiArgs mi) {
    if (mi.ok[0] == 0u) return;
    const uint32_t gw = blockIdx.x * kBlock + threadIdx.x;
    if (gw >= mi.waves2 || static_cast<uint64_t>(gw) * mi.per2 >= a.n2) return;
    const uint32_t w = mi.tail_w[gw];
    if (w < mi.range[1]) atomicOr(reinterpret_cast<unsigned long long *>(mi.bits + w), static_cast<unsigned long long>(mi.tail_m[gw]));
}

// List 1, count pass: order and validity checks and, per wave, the rows whose key is set.  Each
// lane takes kMiR consecutive rows of a step.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_mi_count(MergeArgs a, MiArgs mi) {
    if (mi.ok[0] == 0u) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (gw >= mi.waves1) return;
    const uint64_t begin = static_cast<uint64_t>(gw) * mi.per1;
    const uint64_t end = begin < a.n1 ? mi_rows_end(begin, mi.per1, a.n1) : begin;
    const uint64_t lo = mi.range[0], words = mi.range[1];
    const bool chk_h1 = mi_own_check(a.h1, a.t1);  // h1's codes are checked too
    bool bad = false;
    bool has_prev = begin > 0 && begin < end;
    uint64_t prev = has_prev ? a.t1.dec_of(a.t1.raw(begin - 1)) : 0ull;
    uint32_t cnt = 0;
    K cn[kMiR];
    uint64_t chn[kMiR] = {};
    {
        uint64_t rows[kMiR];
        bool live[kMiR];
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            rows[k] = begin + static_cast<uint64_t>(lane) * kMiR + k;
            live[k] = rows[k] < end;
        }
        mi_load<K, kMiR>(a.t1.data, rows, live, cn);
        if (chk_h1) a.h1.raw_n<kMiR>(rows, live, chn);
    }
    for (uint64_t seg = begin; seg < end; seg += 64u * kMiR) {
        uint64_t rows[kMiR], c[kMiR], ch[kMiR], v[kMiR], word[kMiR];
        bool live[kMiR], in[kMiR];
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            rows[k] = seg + static_cast<uint64_t>(lane) * kMiR + k;
            live[k] = rows[k] < end;
            c[k] = cn[k];
            ch[k] = chn[k];
        }
        {  // the next step's codes, in flight with this step's bitmap reads
            uint64_t nrows[kMiR];
            bool nlive[kMiR];
#pragma unroll
            for (int k = 0; k < kMiR; ++k) {
                nrows[k] = seg + 64u * kMiR + static_cast<uint64_t>(lane) * kMiR + k;
                nlive[k] = nrows[k] < end;
            }
            mi_load<K, kMiR>(a.t1.data, nrows, nlive, cn);
            if (chk_h1) a.h1.raw_n<kMiR>(nrows, nlive, chn);
        }
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            v[k] = a.t1.dec_of(c[k]);
            in[k] = live[k] && v[k] >= lo && ((v[k] - lo) >> 6) < words;
        }
#pragma unroll
        for (int k = 0; k < kMiR; ++k) word[k] = in[k] ? mi.bits[(v[k] - lo) >> 6] : 0ull;
        const uint64_t last_row = (end < seg + 64u * kMiR ? end : seg + 64u * kMiR) - 1u;
        const uint32_t last_lane = static_cast<uint32_t>((last_row - seg) / kMiR);
        const int last_k = static_cast<int>((last_row - seg) % kMiR);
        const uint64_t up = __shfl_up(v[kMiR - 1], 1, 64);
        const uint64_t pv0 = lane == 0 ? prev : up;
        const bool hp0 = lane > 0 || has_prev;
#pragma unroll
        for (int k = 0; k < kMiR; ++k) {
            if (!live[k]) continue;
            const uint64_t pv = k == 0 ? pv0 : v[k - 1];
            if (a.t1.bad_of(c[k]) || (chk_h1 && a.h1.bad_of(ch[k]))) bad = true;
            if ((k > 0 || hp0) && !(pv < v[k])) bad = true;
            if (in[k] && ((word[k] >> ((v[k] - lo) & 63u)) & 1ull) != 0ull) ++cnt;
        }
        uint64_t vl_own = v[0];
#pragma unroll
        for (int k = 1; k < kMiR; ++k) vl_own = k == last_k ? v[k] : vl_own;
        prev = __shfl(vl_own, static_cast<int>(last_lane), 64);
        has_prev = true;
    }
    cnt = wave_sum(cnt);
    if (lane == 0) mi.wcnt[gw] = cnt;
    if (bad) mi.ok[0] = 0u;
}

// Exclusive prefixes of the per-wave matches (one workgroup of 1024 threads) and the counts.
__global__ __launch_bounds__(1024) void k_mi_scan(MergeArgs a, MiArgs mi) {
    __shared__ uint32_t s_w[16];
    if (mi.ok[0] == 0u) return;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < mi.waves1; base += 8192u) {
        uint32_t v[8], t = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t i = base + threadIdx.x * 8u + q;
            v[q] = i < mi.waves1 ? mi.wcnt[i] : 0u;
            t += v[q];
        }
        const uint32_t inc = wave_incl_scan(t);
        if (lane == 63u) s_w[wv] = inc;
        __syncthreads();
        uint64_t pre = carry, tot = 0;
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
            if (i < mi.waves1) mi.wpre[i] = static_cast<uint32_t>(pre);
            pre += v[q];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        a.d_counts[0] = carry;
#pragma unroll
        for (int s = 1; s < kMmStreams; ++s) a.d_counts[s] = 0ull;
    }
}

// List 1, write pass: the pairs of the rows whose key is set, at the wave's prefix + their rank
// (rows seg + g * 64 + lane, ranks from ballots).  The next step's keys are loaded while this
// step's bitmap words are read.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_mi_write(MergeArgs a, MiArgs mi) {
    if (mi.ok[0] == 0u) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (gw >= mi.waves1) return;
    const uint64_t begin = static_cast<uint64_t>(gw) * mi.per1;
    if (begin >= a.n1) return;
    const uint64_t end = mi_rows_end(begin, mi.per1, a.n1);
    const uint64_t lo = mi.range[0], words = mi.range[1];
    uint64_t pos = mi.wpre[gw];
    K c[kMiG];
    {
        uint64_t rows[kMiG];
        bool live[kMiG];
#pragma unroll
        for (int g = 0; g < kMiG; ++g) {
            rows[g] = begin + static_cast<uint64_t>(g) * 64u + lane;
            live[g] = rows[g] < end;
        }
        mi_load<K, kMiG>(a.t1.data, rows, live, c);
    }
    for (uint64_t seg = begin; seg < end; seg += 64u * kMiG) {
        uint64_t rows[kMiG], v[kMiG], word[kMiG];
        bool live[kMiG], in[kMiG];
#pragma unroll
        for (int g = 0; g < kMiG; ++g) {
            rows[g] = seg + static_cast<uint64_t>(g) * 64u + lane;
            live[g] = rows[g] < end;
            v[g] = a.t1.dec_of(static_cast<uint64_t>(c[g]));
            in[g] = live[g] && v[g] >= lo && ((v[g] - lo) >> 6) < words;
        }
#pragma unroll
        for (int g = 0; g < kMiG; ++g) word[g] = in[g] ? mi.bits[(v[g] - lo) >> 6] : 0ull;
        {  // the next step's keys
            uint64_t nrows[kMiG];
            bool nlive[kMiG];
#pragma unroll
            for (int g = 0; g < kMiG; ++g) {
                nrows[g] = seg + 64u * kMiG + static_cast<uint64_t>(g) * 64u + lane;
                nlive[g] = nrows[g] < end;
            }
            mi_load<K, kMiG>(a.t1.data, nrows, nlive, c);
        }
#pragma unroll
        for (int g = 0; g < kMiG; ++g) {
            const uint32_t b = static_cast<uint32_t>((v[g] - lo) & 63u);
            const bool hit = in[g] && ((word[g] >> b) & 1ull) != 0ull;
            const uint64_t bal = __ballot(hit);
            if (hit) {
                const uint64_t p = pos + mi_below(bal);
                if (p < a.out_cap) {
                    const uint64_t h = a.h1_from_t1 ? enc_key(a.t1, a.a_t1, v[g]) : a.h1.raw(rows[g]);
                    uint64_t t;
                    if (a.t2_from_h2) {
                        t = enc_key(a.h2, a.a_h2, v[g]);
                    } else {
                        const uint64_t j = static_cast<uint64_t>(mi.first[(v[g] - lo) >> 6]) +
                                           static_cast<uint64_t>(__popcll(word[g] & ((1ull << b) - 1ull)));
                        t = a.t2.raw(j);
                    }
                    store_w(a.out_h, p, a.wout_h, a.reenc_h ? h * a.f_h1 : h);
                    store_w(a.out_t, p, a.wout_t, a.reenc_t ? t * a.f_t2 : t);
                }
            }
            pos += static_cast<uint64_t>(__popcll(bal));
        }
    }
}

}  // namespace

using namespace ahead_host;

namespace {

// Waves of an intersection kernel that can be resident at once (its persistent grid's cap).
uint32_t mi_resident(const void *kernel) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlock, 0) != hipSuccess || per <= 0) per = 2;
    return static_cast<uint32_t>(cus) * static_cast<uint32_t>(per) * kWaves;
}

// Waves of an intersection pass over n rows: at most `cap`, whole workgroups, at least 64 * kMiG
// rows per wave.
uint32_t mi_waves(uint64_t n, uint32_t cap) {
    const uint64_t step = 64u * kMiG;
    const uint64_t by_rows = (n + step - 1) / step;
    uint64_t w = by_rows < cap ? by_rows : cap;
    w = (w + kWaves - 1) / kWaves * kWaves;
    return static_cast<uint32_t>(w ? w : kWaves);
}

// Calls f with a value of the unsigned type of a code width (1, 2, 4 or 8 bytes).
template <class F>
void by_width(uint32_t w, F f) {
    switch (w) {
        case 1: f(uint8_t{}); break;
        case 2: f(uint16_t{}); break;
        case 4: f(uint32_t{}); break;
        default: f(uint64_t{}); break;
    }
}

template <typename K>
uint32_t mi_cap1() {
    static const uint32_t c = std::min(mi_resident(reinterpret_cast<const void *>(k_mi_count<K>)),
                                       mi_resident(reinterpret_cast<const void *>(k_mi_write<K>)));
    return c;
}

template <typename K>
uint32_t mi_cap2() {
    static const uint32_t c = mi_resident(reinterpret_cast<const void *>(k_mi_build<K>));
    return c;
}

uint64_t mi_per(uint64_t n, uint32_t waves) {
    const uint64_t step = 64u * kMiG;
    const uint64_t per = (n + waves - 1) / waves;
    const uint64_t r = (per + step - 1) / step * step;
    return r ? r : step;
}

}  // namespace

extern "C" int ahead_hip_matchjoin_merge_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2,
                                            const ahead_col *t2, int reenc_h, uint64_t f_h1, uint32_t wout_h,
                                            int reenc_t, uint64_t f_t2, uint32_t wout_t, uint64_t a_oid,
                                            ahead_buf out_h, ahead_buf out_t, ahead_buf e1, ahead_buf e2,
                                            ahead_buf e3, ahead_buf e4, uint64_t *d_counts, ahead_hip_ws *ws,
                                            void *stream) {
    int st;
    if ((st = check_col(h1, "matchjoin_merge.h1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t1, "matchjoin_merge.t1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(h2, "matchjoin_merge.h2", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t2, "matchjoin_merge.t2", true, false)) != AHEAD_OK) return st;
    if (d_counts == nullptr || ws == nullptr) return fail(AHEAD_ERR_INVALID, "matchjoin_merge: NULL counts or workspace");
    if (h1->n != t1->n || h2->n != t2->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "matchjoin_merge: head and tail sizes differ");
    if ((wout_h != 1 && wout_h != 2 && wout_h != 4 && wout_h != 8) || (wout_t != 1 && wout_t != 2 && wout_t != 4 && wout_t != 8))
        return fail(AHEAD_ERR_INVALID, "matchjoin_merge: output widths are 1, 2, 4 or 8 bytes");
    if (t1->n > 0xFFFFFFFFull || h2->n > 0xFFFFFFFFull) return fail(AHEAD_ERR_TOO_LARGE, "matchjoin_merge: more than 2^32-1 rows");
    if ((out_h.cap && (!out_h.data || !out_t.data)))
        return fail(AHEAD_ERR_INVALID, "matchjoin_merge: NULL output with nonzero capacity");
    hipStream_t s = static_cast<hipStream_t>(stream);

    MergeArgs a{};
    a.h1 = dcol(*h1);
    a.t1 = dcol(*t1);
    a.h2 = dcol(*h2);
    a.t2 = dcol(*t2);
    a.n1 = t1->n;
    a.n2 = h2->n;
    a.f_h1 = f_h1;
    a.f_t2 = f_t2;
    a.reenc_h = reenc_h && h1->encoded && h1->data ? 1u : 0u;
    a.reenc_t = reenc_t && t2->encoded && t2->data ? 1u : 0u;
    a.wout_h = wout_h;
    a.wout_t = wout_t;
    a.a_oid = a_oid;
    a.out_h = out_h.data;
    a.out_t = out_t.data;
    a.out_cap = out_h.cap < out_t.cap ? out_h.cap : out_t.cap;
    const ahead_buf eb[4] = {e1, e2, e3, e4};
    for (int e = 0; e < 4; ++e) {
        a.err[e] = static_cast<uint64_t *>(eb[e].data);
        a.err_cap[e] = eb[e].data ? eb[e].cap : 0;
    }
    a.d_counts = d_counts;

    // a list whose two sides are one column is staged once; its output codes are the keys
    // re-encoded when A * A^-1 = 1 at the code width (always, for consistent metadata)
    a.h1_is_t1 = h1->data != nullptr && h1->data == t1->data && a.h1.width == a.t1.width ? 1u : 0u;
    a.t2_is_h2 = t2->data != nullptr && t2->data == h2->data && a.t2.width == a.h2.width ? 1u : 0u;
    a.a_t1 = t1->a & a.t1.mask;
    a.a_h2 = h2->a & a.h2.mask;
    const bool inv1 = !a.t1.encoded || ((a.a_t1 * a.t1.a_inv) & a.t1.mask) == 1ull;
    const bool inv2 = !a.h2.encoded || ((a.a_h2 * a.h2.a_inv) & a.h2.mask) == 1ull;
    a.h1_from_t1 = a.h1_is_t1 && inv1 ? 1u : 0u;
    a.t2_from_h2 = a.t2_is_h2 && inv2 ? 1u : 0u;
    const uint64_t diag = kMmDiag;
    a.cap1 = diag;
    const uint64_t subs64 = (a.n1 + a.n2 + diag - 1) / diag;
    const uint32_t subs = static_cast<uint32_t>(subs64 ? subs64 : 1);
    a.num_sub = subs;
    a.num_grp = (subs + kMmGroup - 1) / kMmGroup;
    a.prov_cap = a.n1 < a.n2 ? a.n1 : a.n2;

    const size_t arr = align_up((static_cast<size_t>(subs) + 1) * 8, 256);
    const size_t o_flags = 0, o_i = 256, o_state = o_i + arr, o_j0 = o_state + arr, o_cnt = o_j0 + arr;
    const size_t o_slow = o_cnt + align_up(static_cast<size_t>(subs) * 4 * kMmStreams, 256);
    const size_t o_grp = o_slow + align_up(static_cast<size_t>(subs) * 4, 256);
    const size_t o_ph = o_grp + align_up(a.num_grp * 8 * kMmStreams, 256);
    const size_t o_pt = o_ph + align_up(a.prov_cap * wout_h, 256);
    const size_t o_end = o_pt + align_up(a.prov_cap * wout_t, 256);
    // the intersection path's regions lie after the sub-tile path's, whose writes must leave the
    // path's flag intact
    MiArgs mi{};
    mi.words_cap = kMiBitsPerRow * a.n2 / 64u + 2u;
    // the path streams the key columns' codes: void key columns take the sub-tile path
    const bool mi_eligible = t1->data != nullptr && h2->data != nullptr;
    by_width(a.t1.width, [&](auto k) { mi.waves1 = mi_waves(a.n1, mi_cap1<decltype(k)>()); });
    by_width(a.h2.width, [&](auto k) { mi.waves2 = mi_waves(a.n2, mi_cap2<decltype(k)>()); });
    mi.per1 = mi_per(a.n1, mi.waves1);
    mi.per2 = mi_per(a.n2, mi.waves2);
    const size_t o_mi = o_end, o_bits = o_mi + 256, o_first = o_bits + align_up(mi.words_cap * 8, 256),
                 o_wcnt = o_first + align_up(mi.words_cap * 4, 256),
                 o_wpre = o_wcnt + align_up(static_cast<size_t>(mi.waves1) * 4, 256),
                 o_tw = o_wpre + align_up(static_cast<size_t>(mi.waves1) * 4, 256),
                 o_tm = o_tw + align_up(static_cast<size_t>(mi.waves2) * 4, 256),
                 o_all = o_tm + align_up(static_cast<size_t>(mi.waves2) * 8, 256);
    if ((st = ws_ensure(ws, o_all)) != AHEAD_OK) return st;
    mi.ok = reinterpret_cast<uint32_t *>(ws_region(ws, o_mi));
    mi.range = reinterpret_cast<uint64_t *>(ws_region(ws, o_mi + 64));
    mi.bits = reinterpret_cast<uint64_t *>(ws_region(ws, o_bits));
    mi.first = reinterpret_cast<uint32_t *>(ws_region(ws, o_first));
    mi.wcnt = reinterpret_cast<uint32_t *>(ws_region(ws, o_wcnt));
    mi.wpre = reinterpret_cast<uint32_t *>(ws_region(ws, o_wpre));
    mi.tail_w = reinterpret_cast<uint32_t *>(ws_region(ws, o_tw));
    mi.tail_m = reinterpret_cast<uint64_t *>(ws_region(ws, o_tm));
    a.fast = mi.ok;
    a.flags = reinterpret_cast<uint32_t *>(ws_region(ws, o_flags));
    SplitArgs sp{};
    sp.diag = diag;
    sp.tiles = subs;
    sp.i = reinterpret_cast<uint64_t *>(ws_region(ws, o_i));
    sp.state = reinterpret_cast<uint64_t *>(ws_region(ws, o_state));
    sp.j0 = reinterpret_cast<uint64_t *>(ws_region(ws, o_j0));
    a.ti = sp.i;
    a.tstate = sp.state;
    a.tj0 = sp.j0;
    a.sub_cnt = reinterpret_cast<uint32_t *>(ws_region(ws, o_cnt));
    a.sub_slow = reinterpret_cast<uint32_t *>(ws_region(ws, o_slow));
    a.grp = reinterpret_cast<uint64_t *>(ws_region(ws, o_grp));
    a.prov_h = ws_region(ws, o_ph);
    a.prov_t = ws_region(ws, o_pt);

    const uint64_t clear_blocks = (mi.words_cap + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_mi_clear, dim3(static_cast<uint32_t>(mi_eligible ? (clear_blocks < 1024 ? clear_blocks : 1024) : 1)),
                       dim3(kBlock), 0, s, a, mi, mi_eligible ? 1u : 0u);
    if (mi_eligible) {
        by_width(a.h2.width, [&](auto k) {
            hipLaunchKernelGGL(k_mi_build<decltype(k)>, dim3(mi.waves2 / kWaves), dim3(kBlock), 0, s, a, mi);
        });
        hipLaunchKernelGGL(k_mi_tails, dim3((mi.waves2 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, mi);
        by_width(a.t1.width, [&](auto k) {
            hipLaunchKernelGGL(k_mi_count<decltype(k)>, dim3(mi.waves1 / kWaves), dim3(kBlock), 0, s, a, mi);
        });
        hipLaunchKernelGGL(k_mi_scan, dim3(1), dim3(1024), 0, s, a, mi);
        by_width(a.t1.width, [&](auto k) {
            hipLaunchKernelGGL(k_mi_write<decltype(k)>, dim3(mi.waves1 / kWaves), dim3(kBlock), 0, s, a, mi);
        });
    }
    // the sub-tile path: its kernels return when the intersection path applied
    hipLaunchKernelGGL(k_mm_split, dim3((subs + kBlock) / kBlock), dim3(kBlock), 0, s, a, sp);
    hipLaunchKernelGGL(k_mm_count, dim3(subs), dim3(kMmThreads), 0, s, a);
    hipLaunchKernelGGL(k_mm_gscan, dim3(1), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_mm_move, dim3((subs + kMmMoveSubs - 1) / kMmMoveSubs), dim3(kMmThreads), 0, s, a);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
