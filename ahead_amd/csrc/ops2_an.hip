// ops2_an.hip -- two more row-tile operators (rowtile.hpp):
//   checkAndDecodeAN<Head, Tail>           (lib/column_operators/AN/encdecAN_scalar.tcc:129-176)
//   matchjoinAN <H1, void, void, T2>       (matchjoinAN.tcc:146-224), seqbase alignment
#include "rowtile.hpp"

using namespace ahead_dev;

namespace {

constexpr int kG = 8;

// Per row: decode head and tail, report each side's corruption, keep the pair only if both are
// valid (encdecAN_scalar.tcc:158-172).
struct CheckDecode2Op {
    static constexpr int kErr = 2;
    static constexpr bool kMulti = false, kMain = true, kAcc = false;
    struct State {
        uint64_t h, t;
    };
    uint64_t n;
    DCol head, tail;
    uint32_t wh, wt;  // unencoded widths of the outputs
    void *dec_h, *dec_t;
    uint64_t cap, a_oid;
    uint64_t *err[2];
    uint64_t err_cap[2];
    uint64_t *d_counts;

    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &) const {
        const uint64_t ch = head.raw(row), ct = tail.raw(row);
        st.h = head.dec_of(ch);
        st.t = tail.dec_of(ct);
        const bool hbad = head.encoded && st.h > head.max_u, tbad = tail.encoded && st.t > tail.max_u;
        errs = (hbad ? 1u : 0u) | (tbad ? 2u : 0u);
        return (hbad || tbad) ? 0u : 1u;
    }
    __device__ void emit_main(uint64_t, const State &st, uint64_t pos, uint32_t) const {
        if (pos < cap) {
            store_w(dec_h, pos, wh, st.h);
            store_w(dec_t, pos, wt, st.t);
        }
    }
    __device__ void emit_err(int e, uint64_t row, const State &, uint64_t pos) const {
        if (pos < err_cap[e]) err[e][pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        d_counts[0] = t[0];
        d_counts[1] = t[1];
        d_counts[2] = t[2];
    }
};

// Row r of the launch stands for list-1 row r (checked while r < s1 + m), list-2 row r (checked
// while r < s2 + m) and, for r >= s1 + s2, the pair p = r - s1 - s2 of list-1 row s1 + p with list-2
// row s2 + p.  s1 / s2 are the rows the reference's first loop skips on one side to align the
// virtual OIDs; m is the number of pairs.
struct MatchSeqOp {
    static constexpr int kErr = 2;
    static constexpr bool kMulti = false, kMain = true, kAcc = false;
    using State = uint8_t;
    uint64_t n;  // s1 + s2 + m
    uint64_t s1, s2, m;
    DCol h1, t2;
    uint32_t reenc_h, reenc_t, wh, wt;
    uint64_t f_h1, f_t2, a_oid;
    void *out_h, *out_t;
    uint64_t cap;
    uint64_t *err[2];
    uint64_t err_cap[2];
    uint64_t *d_counts;

    __device__ uint32_t eval(uint64_t r, State &, uint32_t &errs, uint64_t &) const {
        if (r < s1 + m && h1.bad_of(h1.raw(r))) errs |= 1u;
        if (r < s2 + m && t2.bad_of(t2.raw(r))) errs |= 2u;
        return r >= s1 + s2 ? 1u : 0u;
    }
    __device__ void emit_main(uint64_t r, const State &, uint64_t pos, uint32_t) const {
        if (pos >= cap) return;
        const uint64_t p = r - s1 - s2;
        const uint64_t h = h1.raw(s1 + p), t = t2.raw(s2 + p);
        store_w(out_h, pos, wh, reenc_h ? h * f_h1 : h);
        store_w(out_t, pos, wt, reenc_t ? t * f_t2 : t);
    }
    __device__ void emit_err(int e, uint64_t r, const State &, uint64_t pos) const {
        if (pos < err_cap[e]) err[e][pos] = r * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        d_counts[0] = t[0];
        d_counts[1] = t[1];
        d_counts[2] = t[2];
    }
};

bool out_width_ok(uint32_t w) { return w == 1 || w == 2 || w == 4 || w == 8; }

}  // namespace

using namespace ahead_host;

extern "C" int ahead_hip_check_and_decode2_an(const ahead_col *head, const ahead_col *tail, uint64_t a_oid,
                                              uint32_t wout_h, uint32_t wout_t, void *dec_h, void *dec_t,
                                              uint64_t dec_cap, ahead_buf err_h, ahead_buf err_t,
                                              uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    int st;
    if ((st = check_col(head, "check_and_decode2.head", false, false)) != AHEAD_OK) return st;
    if ((st = check_col(tail, "check_and_decode2.tail", false, false)) != AHEAD_OK) return st;
    if (!head->encoded && !tail->encoded)
        return fail(AHEAD_ERR_INVALID, "check_and_decode2: at least one of head and tail must be AN-encoded");
    if (head->n != tail->n) return fail(AHEAD_ERR_SIZE_MISMATCH, "check_and_decode2: head and tail sizes differ");
    if (!out_width_ok(wout_h) || !out_width_ok(wout_t)) return fail(AHEAD_ERR_INVALID, "check_and_decode2: bad width");
    if (d_counts == nullptr || (dec_cap && (!dec_h || !dec_t)))
        return fail(AHEAD_ERR_INVALID, "check_and_decode2: NULL output");
    CheckDecode2Op op{};
    op.n = tail->n;
    op.head = dcol(*head);
    op.tail = dcol(*tail);
    op.wh = wout_h;
    op.wt = wout_t;
    op.dec_h = dec_h;
    op.dec_t = dec_t;
    op.cap = dec_cap;
    op.a_oid = a_oid;
    op.err[0] = static_cast<uint64_t *>(err_h.data);
    op.err[1] = static_cast<uint64_t *>(err_t.data);
    op.err_cap[0] = err_h.data ? err_h.cap : 0;
    op.err_cap[1] = err_t.data ? err_t.cap : 0;
    op.d_counts = d_counts;
    return launch_rowtile<CheckDecode2Op, kG>(op, op.n, 3, ws, static_cast<hipStream_t>(stream));
}

extern "C" int ahead_hip_matchjoin_seq_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2,
                                          const ahead_col *t2, int reenc_h, uint64_t f_h1, uint32_t wout_h,
                                          int reenc_t, uint64_t f_t2, uint32_t wout_t, uint64_t a_oid,
                                          ahead_buf out_h, ahead_buf out_t, ahead_buf e1, ahead_buf e4,
                                          uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    int st;
    if ((st = check_col(h1, "matchjoin_seq.h1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t1, "matchjoin_seq.t1", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(h2, "matchjoin_seq.h2", true, false)) != AHEAD_OK) return st;
    if ((st = check_col(t2, "matchjoin_seq.t2", true, false)) != AHEAD_OK) return st;
    if (t1->data != nullptr || h2->data != nullptr)
        return fail(AHEAD_ERR_INVALID, "matchjoin_seq: list 1's tail and list 2's head are void columns");
    if (!out_width_ok(wout_h) || !out_width_ok(wout_t)) return fail(AHEAD_ERR_INVALID, "matchjoin_seq: bad width");
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "matchjoin_seq: d_counts is NULL");
    const uint64_t n1 = t1->n, n2 = h2->n, sb1 = t1->seqbase, sb2 = h2->seqbase;
    if (sb1 + n1 > 0xFFFFFFFFull || sb2 + n2 > 0xFFFFFFFFull)
        return fail(AHEAD_ERR_TOO_LARGE, "matchjoin_seq: virtual OIDs beyond 2^32-1");
    MatchSeqOp op{};
    if (n1 && n2) {
        if (sb1 < sb2) op.s1 = sb2 - sb1 < n1 ? sb2 - sb1 : n1;
        if (sb1 > sb2) op.s2 = sb1 - sb2 < n2 ? sb1 - sb2 : n2;
        const uint64_t r1 = n1 - op.s1, r2 = n2 - op.s2;
        op.m = r1 < r2 ? r1 : r2;
    }
    op.n = op.s1 + op.s2 + op.m;
    op.h1 = dcol(*h1);
    op.t2 = dcol(*t2);
    op.reenc_h = reenc_h && h1->encoded && h1->data ? 1u : 0u;
    op.reenc_t = reenc_t && t2->encoded && t2->data ? 1u : 0u;
    op.wh = wout_h;
    op.wt = wout_t;
    op.f_h1 = f_h1;
    op.f_t2 = f_t2;
    op.a_oid = a_oid;
    op.out_h = out_h.data;
    op.out_t = out_t.data;
    op.cap = (out_h.data && out_t.data) ? (out_h.cap < out_t.cap ? out_h.cap : out_t.cap) : 0;
    op.err[0] = static_cast<uint64_t *>(e1.data);
    op.err[1] = static_cast<uint64_t *>(e4.data);
    op.err_cap[0] = e1.data ? e1.cap : 0;
    op.err_cap[1] = e4.data ? e4.cap : 0;
    op.d_counts = d_counts;
    return launch_rowtile<MatchSeqOp, kG>(op, op.n, 3, ws, static_cast<hipStream_t>(stream));
}
