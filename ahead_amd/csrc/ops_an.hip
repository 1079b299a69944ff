// ops_an.hip -- per-operator gfx950 kernels on the row-tile engine (rowtile.hpp):
//   fetchjoinAN                 (lib/column_operators/AN/matchjoinAN.tcc:312-353)
//   matchjoinAN, positional     (matchjoinAN.tcc:240-301)
//   aggregate_sum_groupedAN     (aggregateAN.tcc:41-119)
//   aggregate_mul_sumAN         (aggregateAN_scalar.tcc:58-131)
//   arithmeticAN                (arithmeticAN_scalar.tcc:60-112)
// Every check applies the code-width cast (static_cast<T>(code * A^-1)).  A few reference lines
// compare the int-promoted product of 16-bit codes without it (matchjoinAN.tcc:278 and :342,
// hashjoinAN.tcc:153); no query path feeds 16-bit codes to them.
#include "maptile.hpp"

using namespace ahead_dev;

namespace {

constexpr int kG = 8;  // row groups of 256 per tile (2048 rows)

// ---------------------------------------------------------------------------------------------------
struct FetchOp {
    static constexpr int kErr = 2;
    static constexpr bool kMulti = false, kMain = true, kAcc = false;
    using State = uint64_t;
    uint64_t n;
    DCol oids, col;
    uint32_t ow;
    uint64_t omask;
    uint32_t reenc;
    uint64_t f, a_oid;
    void *out;
    uint64_t out_cap;
    uint64_t *e1, *e2;
    uint64_t e1_cap, e2_cap;
    uint64_t *d_counts;

    static constexpr bool kBatched = true;
    static constexpr int kOuts = 1;
    static constexpr int kMapRows = 8;  // half the steps whose loads wait for the previous step's stores
    __device__ void out_arrays(void *(&p)[2], uint32_t (&w)[2]) const {
        p[0] = out;
        w[0] = ow;
        p[1] = nullptr;
        w[1] = 0;
    }
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        uint64_t o[G], t[G], v[G];
        bool in[G];
        oids.raw_n<G>(row, live, o);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            t[g] = o[g] * oids.a_inv;
            in[g] = live[g] && t[g] < col.n;
        }
        col.raw_n<G>(t, in, v);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            errs[g] = 0;
            cnt[g] = 0;
            st[g] = 0;
            if (!live[g]) continue;
            if (!in[g]) {  // normative bound (the reference reads one past the end at t == |col|)
                errs[g] = 1u;
                continue;
            }
            const uint64_t x = v[g] & omask;
            if (col.encoded && ((x * col.a_inv) & omask) > col.max_u) {
                errs[g] = 2u;
                st[g] = t[g];
                continue;
            }
            st[g] = reenc ? (x * f) & omask : x;
            cnt[g] = 1;
        }
    }
    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &) const {
        const uint64_t t = oids.raw(row) * oids.a_inv;
        if (t >= col.n) {  // normative bound (the reference reads one past the end at t == |col|)
            errs = 1u;
            st = 0;
            return 0;
        }
        const uint64_t v = col.raw(t) & omask;
        if (col.encoded && ((v * col.a_inv) & omask) > col.max_u) {
            errs = 2u;
            st = t;
            return 0;
        }
        st = reenc ? (v * f) & omask : v;
        return 1;
    }
    __device__ void emit_main(uint64_t, const State &st, uint64_t pos, uint32_t) const {
        if (pos < out_cap) store_w(out, pos, ow, st);
    }
    __device__ void emit_err(int e, uint64_t row, const State &st, uint64_t pos) const {
        if (e == 0) {
            if (pos < e1_cap) e1[pos] = row * a_oid;
        } else if (pos < e2_cap) {
            e2[pos] = st * a_oid;
        }
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        d_counts[0] = t[0];
        d_counts[1] = t[1];
        d_counts[2] = t[2];
    }
};

// ---------------------------------------------------------------------------------------------------
struct MatchPosState {
    uint64_t h;  // head value (already re-encoded)
    uint64_t v;  // fetched value, or pos2 for an error
};

struct MatchPosOp {
    static constexpr int kErr = 3;  // e1: corrupt head, e2: position out of range, e4: corrupt fetched tail
    static constexpr bool kMulti = false, kMain = true, kAcc = false;
    using State = MatchPosState;
    uint64_t n;
    DCol h1, t1, t2;
    uint32_t reenc_h, reenc_t, wh, wt;
    uint64_t f_h1, f_t2, mh, mt, a_oid;
    void *out_h, *out_t;
    uint64_t out_cap;
    uint64_t *err[3];
    uint64_t cap[3];
    uint64_t *d_counts;

    static constexpr bool kBatched = true;
    static constexpr int kOuts = 2;
    static constexpr int kMapRows = 8;  // half the steps whose loads wait for the previous step's stores
    __device__ void out_arrays(void *(&p)[2], uint32_t (&w)[2]) const {
        p[0] = out_h;
        p[1] = out_t;
        w[0] = wh;
        w[1] = wt;
    }
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        uint64_t h[G], o[G], p2[G], v[G];
        bool in[G];
        h1.raw_n<G>(row, live, h);
        t1.raw_n<G>(row, live, o);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            p2[g] = o[g] * t1.a_inv;
            in[g] = live[g] && p2[g] < t2.n;
        }
        t2.raw_n<G>(p2, in, v);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            errs[g] = 0;
            cnt[g] = 0;
            if (!live[g]) continue;
            if (h1.bad_of(h[g])) errs[g] |= 1u;
            st[g].h = reenc_h ? (h[g] * f_h1) & mh : h[g];
            if (!in[g]) {
                errs[g] |= 2u;
                continue;
            }
            if (t2.bad_of(v[g])) {
                errs[g] |= 4u;
                st[g].v = p2[g];
                continue;
            }
            st[g].v = reenc_t ? (v[g] * f_t2) & mt : v[g];
            cnt[g] = 1;
        }
    }

    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &) const {
        const uint64_t h = h1.raw(row);
        if (h1.bad_of(h)) errs |= 1u;
        st.h = reenc_h ? (h * f_h1) & mh : h;
        const uint64_t pos2 = t1.raw(row) * t1.a_inv;
        if (pos2 >= t2.n) {
            errs |= 2u;
            return 0;
        }
        const uint64_t v = t2.raw(pos2);
        if (t2.bad_of(v)) {
            errs |= 4u;
            st.v = pos2;
            return 0;
        }
        st.v = reenc_t ? (v * f_t2) & mt : v;
        return 1;
    }
    __device__ void emit_main(uint64_t, const State &st, uint64_t pos, uint32_t) const {
        if (pos < out_cap) {
            store_w(out_h, pos, wh, st.h);
            store_w(out_t, pos, wt, st.v);
        }
    }
    __device__ void emit_err(int e, uint64_t row, const State &st, uint64_t pos) const {
        if (pos >= cap[e]) return;
        err[e][pos] = (e == 2 ? st.v : row) * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        for (int i = 0; i < 4; ++i) d_counts[i] = t[i];
    }
};

// ---------------------------------------------------------------------------------------------------
struct ArithOp {
    static constexpr int kErr = 4;  // head1, tail1, head2, tail2
    static constexpr bool kMulti = false, kMain = false, kAcc = false;
    static constexpr int kMapRows = 8;  // half the steps whose loads wait for the previous step's stores
    using State = uint8_t;
    uint64_t n;
    DCol h1, t1, h2, t2;
    int op;
    uint64_t a_result, rmask, a_oid;
    uint32_t rw;
    void *out;
    uint64_t *err[4];
    uint64_t cap[4];
    uint64_t *d_counts;

    __device__ uint64_t combine(uint64_t c1, uint64_t c2) const {
        const uint64_t x = (t1.dec_of(c1) * a_result) & rmask, y = (t2.dec_of(c2) * a_result) & rmask;
        uint64_t r;
        switch (op) {
            case 0: r = x + y; break;
            case 1: r = x - y; break;
            case 2: r = x * y; break;
            default: r = y ? x / y : 0; break;
        }
        return r & rmask;
    }
    static constexpr bool kBatched = true;
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        uint64_t c1[G], c2[G], a1[G], a2[G];
        t1.raw_n<G>(row, live, c1);
        t2.raw_n<G>(row, live, c2);
        if (h1.encoded) h1.raw_n<G>(row, live, a1);
        if (h2.encoded) h2.raw_n<G>(row, live, a2);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            cnt[g] = 0;
            errs[g] = 0;
            if (!live[g]) continue;
            if (h1.encoded && h1.bad_of(a1[g])) errs[g] |= 1u;
            if (t1.bad_of(c1[g])) errs[g] |= 2u;
            if (h2.encoded && h2.bad_of(a2[g])) errs[g] |= 4u;
            if (t2.bad_of(c2[g])) errs[g] |= 8u;
            store_w(out, row[g], rw, combine(c1[g], c2[g]));
        }
    }
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &errs, uint64_t &) const {
        if (h1.bad_of(h1.raw(row))) errs |= 1u;
        const uint64_t c1 = t1.raw(row), c2 = t2.raw(row);
        if (t1.bad_of(c1)) errs |= 2u;
        if (h2.bad_of(h2.raw(row))) errs |= 4u;
        if (t2.bad_of(c2)) errs |= 8u;
        const uint64_t x = (t1.dec_of(c1) * a_result) & rmask, y = (t2.dec_of(c2) * a_result) & rmask;
        uint64_t r;
        switch (op) {
            case 0: r = x + y; break;
            case 1: r = x - y; break;
            case 2: r = x * y; break;
            default: r = y ? x / y : 0; break;
        }
        store_w(out, row, rw, r & rmask);
        return 0;
    }
    __device__ void emit_main(uint64_t, const State &, uint64_t, uint32_t) const {}
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        d_counts[0] = n;
        for (int i = 1; i < 5; ++i) d_counts[i] = t[i];
    }
};

// ---------------------------------------------------------------------------------------------------
struct MulSumOp {
    static constexpr int kErr = 2;
    static constexpr bool kMulti = false, kMain = false, kAcc = true;
    using State = uint8_t;
    uint64_t n;
    DCol t1, t2;
    uint64_t a_oid;
    uint64_t *partials;
    uint64_t *err[2];
    uint64_t cap[2];
    uint64_t *d_counts;

    static constexpr bool kBatched = true;
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &acc) const {
        uint64_t c1[G], c2[G];
        t1.raw_n<G>(row, live, c1);
        t2.raw_n<G>(row, live, c2);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            cnt[g] = 0;
            errs[g] = (live[g] && t1.bad_of(c1[g]) ? 1u : 0u) | (live[g] && t2.bad_of(c2[g]) ? 2u : 0u);
            acc += c1[g] * c2[g];  // rows past the end load 0
        }
    }
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &errs, uint64_t &acc) const {
        const uint64_t c1 = t1.raw(row), c2 = t2.raw(row);
        if (t1.bad_of(c1)) errs |= 1u;
        if (t2.bad_of(c2)) errs |= 2u;
        acc += c1 * c2;
        return 0;
    }
    __device__ void emit_main(uint64_t, const State &, uint64_t, uint32_t) const {}
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t tile, uint64_t t) const { partials[tile] = t; }
    __device__ void finish(const uint64_t *t) const {
        d_counts[0] = 1;
        d_counts[1] = t[1];
        d_counts[2] = t[2];
    }
};

__global__ __launch_bounds__(kBlock) void k_mulsum_final(const uint64_t *partials, uint32_t n, uint64_t init,
                                                         uint64_t enc, uint64_t *out) {
    __shared__ uint64_t s_red[kWaves];
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) s += partials[i];
    s = block_sum<uint64_t>(s, s_red);
    if (threadIdx.x == 0) *out = (init + s) * enc;
}

// ---------------------------------------------------------------------------------------------------
// Grouped sum: LDS-private accumulators per workgroup when the groups fit, merged with 64-bit
// atomics; otherwise straight 64-bit atomics on the output.
constexpr uint32_t kAggLds = 2048;

struct AggSumArgs {
    uint64_t n;
    DCol head, tail, grp;
    uint64_t num_groups, enc, a_oid;
    uint64_t *sums;
    uint64_t *err[3];
    uint64_t cap[3];
    uint64_t *d_counts;
};

struct AggSumOp {
    // checks and sums in one pass: per-workgroup LDS sums when the groups fit (merged into the
    // output with 64-bit atomics at the end), else 64-bit atomics on the output directly
    static constexpr int kErr = 3;
    static constexpr int kSharedWords = kAggLds;
    static constexpr bool kMulti = false, kMain = false, kAcc = false;
    using State = uint8_t;
    AggSumArgs a;
    uint64_t n;
    static constexpr bool kBatched = true;
    __device__ void shared_init(unsigned long long *s) const {
        const uint32_t ng = a.num_groups <= kAggLds ? static_cast<uint32_t>(a.num_groups) : 0u;
        for (uint32_t i = threadIdx.x; i < ng; i += blockDim.x) s[i] = 0ull;
    }
    __device__ void shared_flush(unsigned long long *s) const {
        if (a.num_groups > kAggLds) return;
        for (uint32_t i = threadIdx.x; i < a.num_groups; i += blockDim.x)
            if (s[i] != 0ull) atomicAdd(reinterpret_cast<unsigned long long *>(&a.sums[i]), s[i]);
    }
    // aggregateAN.tcc:84-112: a corrupt tail is still added (and reported); a row whose group id
    // is corrupt or out of range is reported and skipped
    template <int G>
    __device__ void eval_shared(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G], uint32_t (&errs)[G],
                                uint32_t (&cnt)[G], uint64_t &acc, unsigned long long *s) const {
        uint64_t t[G], gr[G];
        eval_rows<G>(row, live, errs, t, gr);
        const bool lds = a.num_groups <= kAggLds;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            cnt[g] = 0;
            if (!live[g] || (errs[g] & 4u)) continue;
            const uint64_t gi = gr[g] * a.grp.a_inv;
            const unsigned long long add = t[g] * a.enc;
            if (lds) atomicAdd(&s[gi], add);
            else atomicAdd(reinterpret_cast<unsigned long long *>(&a.sums[gi]), add);
        }
        (void)st;
        (void)acc;
    }
    template <int G>
    __device__ void eval_rows(const uint64_t (&row)[G], const bool (&live)[G], uint32_t (&errs)[G], uint64_t (&t)[G],
                              uint64_t (&gr)[G]) const {
        uint64_t h[G];
        if (a.head.encoded) a.head.raw_n<G>(row, live, h);
        a.tail.raw_n<G>(row, live, t);
        a.grp.raw_n<G>(row, live, gr);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            errs[g] = 0;
            if (!live[g]) continue;
            if (a.head.encoded && a.head.bad_of(h[g])) errs[g] |= 1u;
            if (a.tail.bad_of(t[g])) errs[g] |= 2u;
            const uint64_t gi = gr[g] * a.grp.a_inv;
            if (gi > a.grp.max_u || gi >= a.num_groups) errs[g] |= 4u;
        }
    }
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        uint64_t t[G], gr[G];
        eval_rows<G>(row, live, errs, t, gr);
#pragma unroll
        for (int g = 0; g < G; ++g) cnt[g] = 0;
    }
    __device__ uint32_t eval(uint64_t row, State &, uint32_t &errs, uint64_t &) const {
        if (a.head.bad_of(a.head.raw(row))) errs |= 1u;
        if (a.tail.bad_of(a.tail.raw(row))) errs |= 2u;
        const uint64_t g = a.grp.raw(row) * a.grp.a_inv;
        if (g > a.grp.max_u || g >= a.num_groups) errs |= 4u;
        return 0;
    }
    __device__ void emit_main(uint64_t, const State &, uint64_t, uint32_t) const {}
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < a.cap[e]) a.err[e][pos] = row * a.a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        a.d_counts[0] = a.num_groups;
        for (int i = 1; i < 4; ++i) a.d_counts[i] = t[i];
    }
};


__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace

namespace ahead_host {

DCol dcol(const ahead_col &c) {
    DCol d{};
    d.data = c.data;
    d.n = c.n;
    d.width = c.data ? c.width : 4;
    d.encoded = c.encoded && c.data != nullptr;
    d.a_inv = c.a_inv;
    d.mask = width_mask(d.width);
    d.max_u = c.max_u & d.mask;
    d.seqbase = c.seqbase;
    return d;
}

uint32_t encoded_width(const ahead_col &c) {
    if (c.encoded) return c.width;
    const uint32_t w = c.data ? c.width : 4;
    return w >= 4 ? 8 : 2 * w;
}

}  // namespace ahead_host

using namespace ahead_host;

extern "C" int ahead_hip_fetchjoin_an(const ahead_col *oids, const ahead_col *col, int reenc, uint64_t a_reenc,
                                      uint64_t a_oid, ahead_buf out, ahead_buf err1, ahead_buf err2,
                                      uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    int st;
    if ((st = check_col(oids, "fetchjoin.oids", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(col, "fetchjoin.col", false, false)) != AHEAD_OK) return st;
    if (oids->width != 8) return fail(AHEAD_ERR_INVALID, "fetchjoin: OIDs are resoid (8-byte) codes");
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "fetchjoin: d_counts is NULL");
    FetchOp op{};
    op.n = oids->n;
    op.oids = dcol(*oids);
    op.col = dcol(*col);
    op.ow = encoded_width(*col);
    op.omask = width_mask(op.ow);
    op.reenc = reenc != 0;
    // matchjoinAN.tcc:323-327: re-encoding factor A_R * inv(A_col) (A_R alone for a plain column)
    op.f = reenc ? (col->encoded ? (a_reenc * inverse_mod2k(col->a, 8 * op.ow)) & op.omask : a_reenc & op.omask) : 1;
    op.a_oid = a_oid;
    op.out = out.data;
    op.out_cap = out.data ? out.cap : 0;
    op.e1 = static_cast<uint64_t *>(err1.data);
    op.e1_cap = err1.data ? err1.cap : 0;
    op.e2 = static_cast<uint64_t *>(err2.data);
    op.e2_cap = err2.data ? err2.cap : 0;
    op.d_counts = d_counts;
    // the streaming engine stages entries at their row's place: it needs room for every row
    if (op.out != nullptr && op.out_cap >= op.n) return launch_map(op, ws, 0, static_cast<hipStream_t>(stream));
    return launch_rowtile(op, op.n, 3, ws, static_cast<hipStream_t>(stream));
}

extern "C" int ahead_hip_matchjoin_pos_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *t2, int reenc_h,
                                          uint64_t f_h1, uint32_t wout_h, int reenc_t, uint64_t f_t2, uint32_t wout_t,
                                          uint64_t a_oid, ahead_buf out_h, ahead_buf out_t, ahead_buf err1,
                                          ahead_buf err2, ahead_buf err4, uint64_t *d_counts, ahead_hip_ws *ws,
                                          void *stream) {
    int st;
    if ((st = check_col(h1, "matchjoin.h1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t1, "matchjoin.t1", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(t2, "matchjoin.t2", false, false)) != AHEAD_OK) return st;
    if (h1->n != t1->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "matchjoin: head and tail of the left BAT differ");
    if (t1->width != 8) return fail(AHEAD_ERR_INVALID, "matchjoin: the left tail holds resoid (8-byte) codes");
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "matchjoin: d_counts is NULL");
    if ((wout_h != 1 && wout_h != 2 && wout_h != 4 && wout_h != 8) || (wout_t != 1 && wout_t != 2 && wout_t != 4 && wout_t != 8))
        return fail(AHEAD_ERR_INVALID, "matchjoin: output widths must be 1, 2, 4 or 8");
    MatchPosOp op{};
    op.n = t1->n;
    op.h1 = dcol(*h1);
    op.t1 = dcol(*t1);
    op.t2 = dcol(*t2);
    op.reenc_h = reenc_h != 0;
    op.reenc_t = reenc_t != 0;
    op.wh = wout_h;
    op.wt = wout_t;
    op.f_h1 = f_h1;
    op.f_t2 = f_t2;
    op.mh = width_mask(wout_h);
    op.mt = width_mask(wout_t);
    op.a_oid = a_oid;
    op.out_h = out_h.data;
    op.out_t = out_t.data;
    op.out_cap = (out_h.data && out_t.data) ? (out_h.cap < out_t.cap ? out_h.cap : out_t.cap) : 0;
    const ahead_buf e[3] = {err1, err2, err4};
    for (int i = 0; i < 3; ++i) {
        op.err[i] = static_cast<uint64_t *>(e[i].data);
        op.cap[i] = e[i].data ? e[i].cap : 0;
    }
    op.d_counts = d_counts;
    // the streaming engine stages pairs at their row's place: it needs room for every row
    if (op.out_h != nullptr && op.out_t != nullptr && op.out_cap >= op.n)
        return launch_map(op, ws, 0, static_cast<hipStream_t>(stream));
    return launch_rowtile(op, op.n, 4, ws, static_cast<hipStream_t>(stream));
}

extern "C" int ahead_hip_arithmetic_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2, const ahead_col *t2,
                                       int op_code, uint64_t a_result, uint32_t result_width, uint64_t a_oid, void *out,
                                       ahead_buf eh1, ahead_buf et1, ahead_buf eh2, ahead_buf et2, uint64_t *d_counts,
                                       ahead_hip_ws *ws, void *stream) {
    int st;
    if ((st = check_col(h1, "arithmetic.h1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t1, "arithmetic.t1", false, true)) != AHEAD_OK) return st;
    if ((st = check_col(h2, "arithmetic.h2", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t2, "arithmetic.t2", false, true)) != AHEAD_OK) return st;
    if (t1->n != t2->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "arithmetic: bat1->size() != bat2->size()");
    if (op_code < 0 || op_code > 3) return fail(AHEAD_ERR_INVALID, "arithmetic: op must be 0..3 (add sub mul div)");
    if (result_width != 2 && result_width != 4 && result_width != 8)
        return fail(AHEAD_ERR_INVALID, "arithmetic: result width must be 2, 4 or 8");
    if (d_counts == nullptr || (out == nullptr && t1->n)) return fail(AHEAD_ERR_INVALID, "arithmetic: NULL output");
    ArithOp op{};
    op.n = t1->n;
    op.h1 = dcol(*h1);
    op.t1 = dcol(*t1);
    op.h2 = dcol(*h2);
    op.t2 = dcol(*t2);
    op.op = op_code;
    op.a_result = a_result;
    op.rw = result_width;
    op.rmask = width_mask(result_width);
    op.a_oid = a_oid;
    op.out = out;
    const ahead_buf e[4] = {eh1, et1, eh2, et2};
    for (int i = 0; i < 4; ++i) {
        op.err[i] = static_cast<uint64_t *>(e[i].data);
        op.cap[i] = e[i].data ? e[i].cap : 0;
    }
    op.d_counts = d_counts;
    return launch_map(op, ws, 0, static_cast<hipStream_t>(stream));
}

extern "C" int ahead_hip_aggregate_mul_sum_an(const ahead_col *t1, const ahead_col *t2, uint64_t init,
                                              int result_encoded, uint64_t ra, uint64_t a_oid, uint64_t *d_total,
                                              ahead_buf err1, ahead_buf err2, uint64_t *d_counts, ahead_hip_ws *ws,
                                              void *stream) {
    int st;
    if ((st = check_col(t1, "mul_sum.t1", false, false)) != AHEAD_OK) return st;
    if ((st = check_col(t2, "mul_sum.t2", false, false)) != AHEAD_OK) return st;
    if (d_total == nullptr || d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "mul_sum: NULL output");
    MulSumOp op{};
    op.n = t1->n < t2->n ? t1->n : t2->n;  // aggregateAN_scalar.tcc:91 stops at the shorter input
    op.t1 = dcol(*t1);
    op.t2 = dcol(*t2);
    op.a_oid = a_oid;
    const uint32_t waves = map_waves<MulSumOp>(op.n);  // one partial per wave of the streaming engine
    const size_t o_map = align_up(waves * 8ull, 256);
    if ((st = ws_ensure(ws, o_map + map_scratch_bytes<MulSumOp>(op.n))) != AHEAD_OK) return st;
    op.partials = reinterpret_cast<uint64_t *>(ws_region(ws, 0));
    const ahead_buf e[2] = {err1, err2};
    for (int i = 0; i < 2; ++i) {
        op.err[i] = static_cast<uint64_t *>(e[i].data);
        op.cap[i] = e[i].data ? e[i].cap : 0;
    }
    op.d_counts = d_counts;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = launch_map(op, ws, o_map, s)) != AHEAD_OK) return st;
    // aggregateAN_scalar.tcc:83-85: RA * inv(A_T1) * inv(A_T2) over 64 bits
    const uint64_t enc = (result_encoded ? ra : 1ull) * (t1->encoded ? inverse_mod2k(t1->a, 64) : 1ull) *
                         (t2->encoded ? inverse_mod2k(t2->a, 64) : 1ull);
    hipLaunchKernelGGL(k_mulsum_final, dim3(1), dim3(kBlock), 0, s, op.partials, waves, init, enc, d_total);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

extern "C" int ahead_hip_aggregate_sum_grouped_an(const ahead_col *head, const ahead_col *tail, const ahead_col *grouping,
                                                  uint64_t num_groups, uint64_t a_result, uint64_t a_oid, uint64_t *sums,
                                                  ahead_buf e_head, ahead_buf e_tail, ahead_buf e_group,
                                                  uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    int st;
    if ((st = check_col(head, "sum_grouped.head", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(tail, "sum_grouped.tail", false, false)) != AHEAD_OK) return st;
    if ((st = check_col(grouping, "sum_grouped.grouping", false, true)) != AHEAD_OK) return st;
    if (tail->n != grouping->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "bat and grouping must have the same size!");
    if (grouping->width != 8) return fail(AHEAD_ERR_INVALID, "sum_grouped: group ids are resoid (8-byte) codes");
    if (d_counts == nullptr || (sums == nullptr && num_groups)) return fail(AHEAD_ERR_INVALID, "sum_grouped: NULL output");
    AggSumArgs a{};
    a.n = tail->n;
    a.head = dcol(*head);
    a.tail = dcol(*tail);
    a.grp = dcol(*grouping);
    a.grp.max_u = grouping->max_u;
    a.num_groups = num_groups;
    // aggregateAN.tcc:74-75: A_result * inv(A_tail) mod 2^64
    a.enc = a_result * (tail->encoded ? inverse_mod2k(tail->a, 64) : 1ull);
    a.a_oid = a_oid;
    a.sums = sums;
    const ahead_buf e[3] = {e_head, e_tail, e_group};
    for (int i = 0; i < 3; ++i) {
        a.err[i] = static_cast<uint64_t *>(e[i].data);
        a.cap[i] = e[i].data ? e[i].cap : 0;
    }
    a.d_counts = d_counts;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_groups) {
        hipLaunchKernelGGL(k_fill_u64, dim3(64), dim3(256), 0, s, sums, num_groups, 0ull);
        AHEAD_HIP_TRY(hipGetLastError());
    }
    AggSumOp op{};
    op.a = a;
    op.n = a.n;
    return launch_map(op, ws, 0, s);
}
