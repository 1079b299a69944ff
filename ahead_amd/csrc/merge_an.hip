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
// The GPU splits list 1 into chunks of kMmRows rows, one chunk per thread.  For sorted inputs the
// state in which the walk first reaches row i0 = c * kMmRows follows from three binary searches
// (k_mm_start): with v = t1[i0-1], r its rank in its run of equal keys in list 1 and [lb2, ub2)
// the run of v in list 2, the walk arrives by a match at (i0, lb2 + r + 1) in TOP if r < ub2 - lb2,
// else by a skip at (i0, ub2) in SKIP1, or never (ub2 = |list 2|: the walk ended first).  Each
// thread walks its chunk from that state (k_mm_count), counts what it emits into the five output
// streams, and checks that it leaves the chunk in exactly the state computed for the next chunk.
// Chunk 0 starts at (0, 0) in TOP, so if every chunk agrees the chunk walks are the reference's
// walk, and k_mm_write repeats them, writing at offsets from a decoupled look-back scan.  If any
// chunk disagrees (the inputs are not sorted in decoded value, e.g. a corrupted key),
// k_mm_fallback walks the whole input with one thread instead: slower, same results.
#include "rowtile.hpp"

namespace {

using namespace ahead_dev;

constexpr uint32_t kMmRows = 16;  // list-1 rows per chunk
constexpr uint64_t kPhTop = 0, kPhSkip1 = 1, kPhEnd = 2;
constexpr uint64_t kJMask = (1ull << 62) - 1ull;
constexpr int kMmStreams = 5;     // pairs, e1 (h1), e2 (t1), e3 (h2), e4 (t2)

__device__ __forceinline__ uint64_t st_pack(uint64_t j, uint64_t ph) { return ph == kPhEnd ? (kPhEnd << 62) : ((ph << 62) | j); }

struct MergeArgs {
    DCol h1, t1, h2, t2;
    uint64_t n1, n2;
    uint64_t num_chunks;
    uint64_t f_h1, f_t2;
    uint32_t reenc_h, reenc_t, wout_h, wout_t;
    uint64_t a_oid;
    void *out_h, *out_t;
    uint64_t out_cap;
    uint64_t *err[4];
    uint64_t err_cap[4];
    uint64_t *start;    // [num_chunks] entry state of each chunk
    uint32_t *offs;     // [kMmStreams][num_chunks] exclusive output offsets
    uint32_t *flags;    // [0]: a chunk disagreed with its successor's entry state
    uint64_t *d_counts;
};

__device__ __forceinline__ uint64_t dec_at(const DCol &c, uint64_t i) { return c.dec_of(c.raw(i)); }

// first index in [0, hi) whose decoded value is >= v (strict = false) or > v (strict = true)
__device__ uint64_t bound_dec(const DCol &c, uint64_t hi, uint64_t v, bool strict) {
    uint64_t lo = 0;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        const uint64_t d = dec_at(c, mid);
        if (strict ? d <= v : d < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_mm_start(MergeArgs a) {
    const uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= a.num_chunks) return;
    if (c == 0) {
        a.start[0] = st_pack(0, a.n1 && a.n2 ? kPhTop : kPhEnd);
        return;
    }
    const uint64_t i0 = c * kMmRows;
    const uint64_t v = dec_at(a.t1, i0 - 1);
    const uint64_t lb1 = bound_dec(a.t1, i0, v, false);
    const uint64_t r = i0 - 1 - lb1;
    const uint64_t lb2 = bound_dec(a.h2, a.n2, v, false);
    const uint64_t ub2 = bound_dec(a.h2, a.n2, v, true);
    uint64_t s;
    if (r < ub2 - lb2) s = st_pack(lb2 + r + 1, kPhTop);
    else if (ub2 >= a.n2) s = st_pack(0, kPhEnd);
    else s = st_pack(ub2, kPhSkip1);
    a.start[c] = s;
}

// Walks list-1 rows [i, iend) from entry state `st`; returns the state in which the walk reaches
// row iend (kPhEnd if it stops first or iend is the end of list 1).  WRITE: emits at base[] +
// running counts; otherwise only counts.
template <bool WRITE>
__device__ uint64_t walk(const MergeArgs &a, uint64_t i, uint64_t iend, uint64_t st, uint32_t (&cnt)[kMmStreams],
                         const uint64_t (&base)[kMmStreams]) {
    uint64_t ph = st >> 62, j = st & kJMask;
    if (ph == kPhEnd) return st_pack(0, kPhEnd);
    auto err = [&](int e, uint64_t pos) {
        if constexpr (WRITE) {
            const uint64_t p = base[1 + e] + cnt[1 + e];
            if (p < a.err_cap[e]) a.err[e][p] = pos * a.a_oid;
        }
        ++cnt[1 + e];
    };
    auto check2 = [&](uint64_t jj, uint64_t hcode) {
        if (a.h2.bad_of(hcode)) err(2, jj);
        if (a.t2.encoded && a.t2.bad_of(a.t2.raw(jj))) err(3, jj);
    };
    uint64_t hcode = 0, h2v = 0;
    if (ph == kPhSkip1) {
        hcode = a.h2.raw(j);
        h2v = a.h2.dec_of(hcode);
    }
    for (;;) {
        if (ph == kPhTop) {
            if (j >= a.n2) return st_pack(0, kPhEnd);
            hcode = a.h2.raw(j);
            h2v = a.h2.dec_of(hcode);
            check2(j, hcode);
        }
        uint64_t tcode = a.t1.raw(i);
        uint64_t t1v = a.t1.dec_of(tcode);
        while (t1v < h2v) {  // SKIP1
            if (a.h1.encoded && a.h1.bad_of(a.h1.raw(i))) err(0, i);
            if (a.t1.bad_of(tcode)) err(1, i);
            if (++i == iend) return iend == a.n1 ? st_pack(0, kPhEnd) : st_pack(j, kPhSkip1);
            tcode = a.t1.raw(i);
            t1v = a.t1.dec_of(tcode);
        }
        while (t1v > h2v) {  // SKIP2
            check2(j, hcode);
            if (++j >= a.n2) return st_pack(0, kPhEnd);
            hcode = a.h2.raw(j);
            h2v = a.h2.dec_of(hcode);
        }
        if (t1v == h2v) {
            if constexpr (WRITE) {
                const uint64_t p = base[0] + cnt[0];
                if (p < a.out_cap) {
                    const uint64_t h = a.h1.raw(i), t = a.t2.raw(j);
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

__global__ __launch_bounds__(kBlock) void k_mm_count(MergeArgs a, LbArgs lb) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_excl[kMmStreams];
    __shared__ uint32_t s_red[kWaves];
    const uint32_t tile = take_tile(lb, &s_tile);
    if (tile >= lb.num_tiles) return;
    const uint64_t c = static_cast<uint64_t>(tile) * kBlock + threadIdx.x;
    uint32_t cnt[kMmStreams] = {0, 0, 0, 0, 0};
    if (c < a.num_chunks) {
        const uint64_t zero[kMmStreams] = {0, 0, 0, 0, 0};
        const uint64_t i = c * kMmRows;
        const uint64_t iend = i + kMmRows < a.n1 ? i + kMmRows : a.n1;
        const uint64_t end = walk<false>(a, i, iend, a.start[c], cnt, zero);
        if (c + 1 < a.num_chunks && end != a.start[c + 1]) atomicOr(&a.flags[0], 1u);
    }
    uint32_t ex[kMmStreams], agg[kMmStreams];
#pragma unroll
    for (int s = 0; s < kMmStreams; ++s) {
        ex[s] = block_excl_scan<uint32_t>(cnt[s], agg[s], s_red);
        __syncthreads();
    }
    lb_resolve<kMmStreams>(lb, tile, agg, s_excl);
    if (c < a.num_chunks) {
#pragma unroll
        for (int s = 0; s < kMmStreams; ++s) a.offs[static_cast<uint64_t>(s) * a.num_chunks + c] = s_excl[s] + ex[s];
    }
    if (tile == lb.num_tiles - 1 && threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < kMmStreams; ++s) a.d_counts[s] = static_cast<uint64_t>(s_excl[s]) + agg[s];
    }
}

__global__ __launch_bounds__(kBlock) void k_mm_write(MergeArgs a) {
    const uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= a.num_chunks || a.flags[0] != 0u) return;
    uint64_t base[kMmStreams];
#pragma unroll
    for (int s = 0; s < kMmStreams; ++s) base[s] = a.offs[static_cast<uint64_t>(s) * a.num_chunks + c];
    uint32_t cnt[kMmStreams] = {0, 0, 0, 0, 0};
    const uint64_t i = c * kMmRows;
    const uint64_t iend = i + kMmRows < a.n1 ? i + kMmRows : a.n1;
    (void) walk<true>(a, i, iend, a.start[c], cnt, base);
}

// The whole walk on one thread, for inputs whose chunk entry states could not be derived.
__global__ void k_mm_fallback(MergeArgs a) {
    if (threadIdx.x != 0 || a.flags[0] == 0u) return;
    uint32_t cnt[kMmStreams] = {0, 0, 0, 0, 0};
    const uint64_t zero[kMmStreams] = {0, 0, 0, 0, 0};
    if (a.n1 && a.n2) (void) walk<true>(a, 0, a.n1, st_pack(0, kPhTop), cnt, zero);
#pragma unroll
    for (int s = 0; s < kMmStreams; ++s) a.d_counts[s] = cnt[s];
}

}  // namespace

using namespace ahead_host;

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
    a.num_chunks = (a.n1 + kMmRows - 1) / kMmRows;
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

    const uint64_t K = a.num_chunks ? a.num_chunks : 1;
    const size_t o_flags = 0, o_start = 256, o_offs = o_start + align_up(K * 8, 256);
    const size_t o_end = o_offs + align_up(K * 4 * kMmStreams, 256);
    const LbPlan plan = lb_plan(a.num_chunks, kBlock, kMmStreams);
    if ((st = ws_ensure(ws, o_end)) != AHEAD_OK) return st;
    a.flags = reinterpret_cast<uint32_t *>(ws_region(ws, o_flags));
    a.start = reinterpret_cast<uint64_t *>(ws_region(ws, o_start));
    a.offs = reinterpret_cast<uint32_t *>(ws_region(ws, o_offs));
    LbArgs lb{};
    if ((st = lb_args(ws, plan, &lb)) != AHEAD_OK) return st;

    if ((st = fill_u32(a.flags, 1, 0u, s)) != AHEAD_OK) return st;
    const uint32_t grid = static_cast<uint32_t>((K + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_mm_start, dim3(grid), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_mm_count, dim3(plan.num_tiles), dim3(kBlock), 0, s, a, lb);
    hipLaunchKernelGGL(k_mm_write, dim3(grid), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_mm_fallback, dim3(1), dim3(64), 0, s, a);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
