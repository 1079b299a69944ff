// test_operators_an.cpp -- the C++ operator templates (include/ahead/column_operators/
// OperatorsAN.hpp) driven the way the reference's query drivers drive them
// (tools/ssbm/ssbm-q11_continuous.cpp:69-121, ssbm-q21_continuous.cpp:97-181), on device-resident
// BATs, with every result compared bit for bit against the CPU oracle (oracle/an_oracle.h).
//
// Exit status 0 when every check passes; each check prints one line.
#include <column_operators/OperatorsAN.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../oracle/an_oracle.h"

using namespace ahead;
using namespace ahead::bat::ops;
using namespace ahead::bat::ops::scalar;

namespace {

int g_failures = 0;

uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template<typename T>
std::vector<T> uniform(size_t n, uint64_t seed, uint64_t lo, uint64_t hi, uint64_t a, double flip_rate = 0.0) {
    std::vector<T> v(n);
    for (size_t i = 0; i < n; ++i) {
        const uint64_t r = mix(seed * 0x100000001B3ull + i);
        const uint64_t d = lo + (r >> 11) % (hi - lo + 1);
        T c = static_cast<T>(d * a);
        if (flip_rate > 0 && static_cast<double>(mix(r) >> 11) / 9007199254740992.0 < flip_rate)
            c ^= static_cast<T>(T(1) << (mix(r + 1) % (8 * sizeof(T))));
        v[i] = c;
    }
    return v;
}

template<typename V2>
or_col ocol_of(ColumnDescriptor<V2> & cd, uint64_t n) {
    or_col c {};
    c.seqbase = cd.metaData.seqbase;
    if constexpr (std::is_same_v<V2, v2_void_t>) {
        c.data = nullptr;
        c.n = n;
        c.width = 4;
        c.encoded = 0;
        c.unenc_bits = 32;
        c.a = c.a_inv = 1;
        c.max_u = 0xFFFFFFFFull;
    } else {
        c.data = cd.container->host().data();
        c.n = cd.container->size();
        c.width = sizeof(typename V2::type_t);
        c.encoded = V2::is_encoded;
        c.unenc_bits = 8 * sizeof(typename TypeMap<V2>::v2_base_t::type_t);
        c.a = cd.metaData.AN_A;
        c.a_inv = cd.metaData.AN_Ainv;
        c.max_u = cd.metaData.AN_unencMaxU;
    }
    return c;
}

struct Vec {
    or_vec v {};
    ~Vec() { or_vec_free(&v); }
    or_vec *operator&() { return &v; }
};

template<typename T>
void expect_eq(const char *what, const std::vector<T> & got, const or_vec & ref) {
    bool ok = got.size() == ref.n;
    for (size_t i = 0; ok && i < got.size(); ++i)
        ok = static_cast<uint64_t>(got[i]) == ref.v[i];
    std::printf("%-58s %s (%zu vs %llu)\n", what, ok ? "ok" : "FAILED", got.size(), static_cast<unsigned long long>(ref.n));
    if (!ok)
        ++g_failures;
}

void expect_vec(const char *what, AN_indicator_vector *got, const or_vec & ref, bool expect_null = false) {
    if (expect_null) {
        const bool ok = got == nullptr;
        std::printf("%-58s %s\n", what, ok ? "ok (nullptr)" : "FAILED (expected nullptr)");
        if (!ok)
            ++g_failures;
        return;
    }
    if (!got) {
        std::printf("%-58s FAILED (nullptr)\n", what);
        ++g_failures;
        return;
    }
    expect_eq(what, *got, ref);
    delete got;
}

template<typename H, typename T>
std::vector<typename T::type_t> tail_of(BAT<H, T> *b) { return b->tail.container->host(); }
template<typename H, typename T>
std::vector<typename H::type_t> head_of(BAT<H, T> *b) { return b->head.container->host(); }

template<typename V2>
TempBAT<v2_void_t, V2> *column(const std::vector<typename V2::type_t> & v, uint16_t A) {
    typedef typename V2::type_t code_t;
    ColumnMetaData md(size_bytes<code_t>, A, ext_euclidean(A, 8 * sizeof(code_t)), V2::UNENC_MAX_U, 0);
    return make_bat<V2>(v, md);
}

void test_select_codec() {
    const size_t n = 200003;
    auto *b = column<v2_resshort_t>(uniform<uint32_t>(n, 1, 1, 50, 233, 1e-4), 233);
    or_col c = ocol_of(b->tail, n);
    {
        auto r = selectAN<std::less>(b, 25u * b->tail.metaData.AN_A);
        Vec oids, err;
        or_select_an(&c, OR_LT, 25u * 233u, OR_LT, 25u * 233u, OR_COMB_NONE, 65117, 16089, &oids, &err);
        expect_eq("selectAN<less> OIDs", head_of(r.first), oids.v);
        expect_vec("selectAN<less> errors", r.second, err.v);
        delete r.first;
    }
    {
        auto r = selectAN<std::greater_equal, std::less_equal, ahead::and_is>(b, 1u * 233u, 3u * 233u);
        Vec oids, err;
        or_select_an(&c, OR_GE, 233u, OR_LE, 3u * 233u, OR_COMB_AND, 65117, 16089, &oids, &err);
        expect_eq("selectAN<ge, le, and_is> OIDs", head_of(r.first), oids.v);
        expect_vec("selectAN<ge, le, and_is> errors", r.second, err.v);
        delete r.first;
    }
    {
        AN_indicator_vector *e = checkAN(b, static_cast<resshort_t>(b->tail.metaData.AN_Ainv), static_cast<resshort_t>(0xFFFF));
        Vec err;
        or_check_an(&c, 16089, &err);
        expect_vec("checkAN errors", e, err.v);
    }
    {
        auto r = decodeAN(b);
        Vec dec, err;
        or_decode_an(&c, 16089, &dec, &err);
        expect_eq("decodeAN values", tail_of(r.first), dec.v);
        expect_vec("decodeAN errors", r.second, err.v);
        delete r.first;
    }
    {
        auto r = checkAndDecodeAN(b);
        Vec dec, err;
        or_check_and_decode_an(&c, 16089, &dec, &err);
        expect_eq("checkAndDecodeAN values", tail_of(std::get<0>(r)), dec.v);
        expect_vec("checkAndDecodeAN head errors", std::get<1>(r), dec.v, true);
        expect_vec("checkAndDecodeAN tail errors", std::get<2>(r), err.v);
        delete std::get<0>(r);
    }
    {
        auto *plain = make_bat<v2_shortint_t>(uniform<uint16_t>(n, 5, 0, 0xFFFF, 1), ColumnMetaData(size_bytes<uint16_t>));
        auto *enc = encodeAN(plain, 63877u);
        or_col pc = ocol_of(plain->tail, n);
        Vec out;
        or_encode_an(&pc, 63877u, 4, &out);
        expect_eq("encodeAN codes", tail_of(enc), out.v);
        const bool md_ok = enc->tail.metaData.AN_A == 63877 && enc->tail.metaData.AN_Ainv == 0xD142174Dull;
        std::printf("%-58s %s\n", "encodeAN metadata (A, A^-1 mod 2^32)", md_ok ? "ok" : "FAILED");
        g_failures += !md_ok;
        delete enc;
        delete plain;
    }
    delete b;
}

// Q1.1 shape (ssbm-q11_continuous.cpp:69-121): two selections intersected by the merge join, then
// positional fetches through the intersected OID list and the price x discount product sum.
void test_q11_chain() {
    const size_t n = 300007;
    auto *lq = column<v2_restiny_t>(uniform<uint16_t>(n, 11, 1, 50, 233, 1e-5), 233);
    auto *ld = column<v2_restiny_t>(uniform<uint16_t>(n, 12, 0, 10, 233, 1e-5), 233);
    auto *le = column<v2_resint_t>(uniform<uint64_t>(n, 13, 90000, 10000000, 65117, 1e-5), 65117);

    auto pair1 = selectAN<std::less>(lq, 25u * lq->tail.metaData.AN_A);
    auto pair2 = selectAN<std::greater_equal, std::less_equal, ahead::and_is>(ld, 1u * ld->tail.metaData.AN_A, 3u * ld->tail.metaData.AN_A);
    auto *bat3 = pair1.first->mirror_head();
    auto *batZ = pair2.first->mirror_head();
    auto tuple4 = matchjoinAN(bat3, batZ);
    auto *bat5 = std::get<0>(tuple4)->mirror_head();
    auto tupleD = matchjoinAN(bat5, le);
    auto tupleE = matchjoinAN(bat5, ld);
    auto tupleG = aggregate_mul_sumAN<v2_resbigint_t>(std::get<0>(tupleD), std::get<0>(tupleE));

    // oracle chain on the same columns
    or_col cq = ocol_of(lq->tail, n), cd = ocol_of(ld->tail, n), ce = ocol_of(le->tail, n);
    Vec o1, e1, o2, e2;
    or_select_an(&cq, OR_LT, 25u * 233u, OR_LT, 25u * 233u, OR_COMB_NONE, 65117, 16089, &o1, &e1);
    or_select_an(&cd, OR_GE, 233u, OR_LE, 3u * 233u, OR_COMB_AND, 65117, 16089, &o2, &e2);
    expect_eq("Q1.1 select quantity", head_of(pair1.first), o1.v);
    expect_eq("Q1.1 select discount", head_of(pair2.first), o2.v);
    const uint64_t oid_inv = ext_euclidean(65117, 64);
    auto oid_col = [&](const or_vec & v) {
        or_col c {};
        c.data = v.v;
        c.n = v.n;
        c.width = 8;
        c.encoded = 1;
        c.unenc_bits = 32;
        c.a = 65117;
        c.a_inv = oid_inv;
        c.max_u = v2_resoid_t::UNENC_MAX_U;
        return c;
    };
    or_col s1 = oid_col(o1.v), s2 = oid_col(o2.v);
    Vec mh, mt, m1, m2, m3, m4;
    or_matchjoin_merge_an(&s1, &s1, &s2, &s2, 0, 0, 8, 0, 0, 8, 16089, &mh, &mt, &m1, &m2, &m3, &m4);
    expect_eq("Q1.1 merge matchjoin heads", head_of(std::get<0>(tuple4)), mh.v);
    expect_eq("Q1.1 merge matchjoin tails", tail_of(std::get<0>(tuple4)), mt.v);
    expect_vec("Q1.1 merge matchjoin e1", std::get<1>(tuple4), m1.v);
    expect_vec("Q1.1 merge matchjoin e2", std::get<2>(tuple4), m2.v);
    expect_vec("Q1.1 merge matchjoin e3", std::get<3>(tuple4), m3.v);
    expect_vec("Q1.1 merge matchjoin e4", std::get<4>(tuple4), m4.v);
    or_col j = oid_col(mh.v);
    Vec dh, dt, d1, d2, d4, fh, ft, f1, f2, f4;
    or_matchjoin_pos_an(&j, &j, &ce, 0, 0, 8, 0, 0, 8, 16089, &dh, &dt, &d1, &d2, &d4);
    or_matchjoin_pos_an(&j, &j, &cd, 0, 0, 8, 0, 0, 2, 16089, &fh, &ft, &f1, &f2, &f4);
    expect_eq("Q1.1 fetch extendedprice heads", head_of(std::get<0>(tupleD)), dh.v);
    expect_eq("Q1.1 fetch extendedprice tails", tail_of(std::get<0>(tupleD)), dt.v);
    expect_vec("Q1.1 fetch extendedprice e4", std::get<4>(tupleD), d4.v);
    expect_eq("Q1.1 fetch discount tails", tail_of(std::get<0>(tupleE)), ft.v);
    or_col pd = ocol_of(std::get<0>(tupleD)->tail, 0), pe = ocol_of(std::get<0>(tupleE)->tail, 0);
    Vec g1, g2;
    const uint64_t total = or_aggregate_mul_sum_an(&pd, &pe, 0, 1, 65117, 16089, &g1, &g2);
    const std::vector<uint64_t> got_total = tail_of(std::get<0>(tupleG));
    const bool ok = got_total.size() == 1 && got_total[0] == total;
    std::printf("%-58s %s\n", "Q1.1 aggregate_mul_sumAN<resbigint>", ok ? "ok" : "FAILED");
    g_failures += !ok;
    expect_vec("Q1.1 aggregate_mul_sumAN e1", std::get<1>(tupleG), g1.v);
    expect_vec("Q1.1 aggregate_mul_sumAN e2", std::get<2>(tupleG), g2.v);

    delete pair1.second;
    delete pair2.second;
    delete pair1.first;
    delete pair2.first;
    delete bat3;
    delete batZ;
    delete std::get<0>(tuple4);
    delete bat5;
    for (AN_indicator_vector *v : {std::get<1>(tupleD), std::get<2>(tupleD), std::get<1>(tupleE), std::get<2>(tupleE), std::get<4>(tupleE)})
        delete v;
    delete std::get<0>(tupleD);
    delete std::get<0>(tupleE);
    delete std::get<0>(tupleG);
    delete lq;
    delete ld;
    delete le;
}

// Q2.1 shape (ssbm-q21_continuous.cpp:97-181): fact FK hash-joined to the dimension key, attributes
// fetched, grouped by two attributes, measure summed per group.
void test_star_chain() {
    const size_t nf = 250001, nd = 2000;
    auto *fk = column<v2_resint_t>(uniform<uint64_t>(nf, 21, 0, nd - 1, 63877, 1e-4), 63877);
    auto *fm = column<v2_resint_t>(uniform<uint64_t>(nf, 22, 90000, 10500000, 65117, 1e-4), 65117);
    std::vector<uint64_t> keys(nd);
    for (size_t i = 0; i < nd; ++i) keys[i] = i * 63877ull;
    auto *dk = column<v2_resint_t>(keys, 63877);
    auto *da = column<v2_restiny_t>(uniform<uint16_t>(nd, 23, 0, 24, 29, 1e-3), 29);

    auto *rev = dk->reverse();
    auto tj = hashjoinAN(fk, rev, 65117ull, ext_euclidean(65117, 64), 65117ull, ext_euclidean(65117, 64));
    or_col cfk = ocol_of(fk->tail, nf), cdk = ocol_of(dk->tail, nd);
    or_col vf {}, vd {};
    vf.n = nf;
    vd.n = nd;
    vf.width = vd.width = 4;
    vf.unenc_bits = vd.unenc_bits = 32;
    vf.a = vf.a_inv = vd.a = vd.a_inv = 1;
    vf.max_u = vd.max_u = 0xFFFFFFFFull;
    Vec jh, jt, j1, j2, j3, j4;
    or_hashjoin_an(&vf, &cfk, &cdk, &vd, 1, 1, 65117, 8, 1, 65117, 8, 16089, &jh, &jt, &j1, &j2, &j3, &j4);
    expect_eq("Q2.1 hashjoinAN heads (fact OIDs)", head_of(std::get<0>(tj)), jh.v);
    expect_eq("Q2.1 hashjoinAN tails (dim OIDs)", tail_of(std::get<0>(tj)), jt.v);
    expect_vec("Q2.1 hashjoinAN e1 (void head)", std::get<1>(tj), j1.v, true);
    expect_vec("Q2.1 hashjoinAN e2 (fact FK)", std::get<2>(tj), j2.v);
    expect_vec("Q2.1 hashjoinAN e3 (dim key)", std::get<3>(tj), j3.v);
    expect_vec("Q2.1 hashjoinAN e4 (void tail)", std::get<4>(tj), j4.v, true);

    // fetch the attribute through the dimension OIDs, group by it, sum the fetched measure
    auto *dim_oids = std::get<0>(tj)->clear_head();
    auto tf = fetchjoinAN(dim_oids, da);
    or_col co = ocol_of(dim_oids->tail, 0), cda = ocol_of(da->tail, nd);
    Vec fo, f1, f2;
    or_fetchjoin_an(&co, &cda, 0, 0, 16089, &fo, &f1, &f2);
    expect_eq("Q2.1 fetchjoinAN attribute", tail_of(std::get<0>(tf)), fo.v);
    expect_vec("Q2.1 fetchjoinAN e1", std::get<1>(tf), f1.v);
    expect_vec("Q2.1 fetchjoinAN e2", std::get<2>(tf), f2.v);

    auto tg = groupbyAN(std::get<0>(tf));
    or_col cg = ocol_of(std::get<0>(tf)->tail, 0), vg {};
    vg.n = cg.n;
    vg.width = 4;
    vg.unenc_bits = 32;
    vg.a = vg.a_inv = 1;
    vg.max_u = 0xFFFFFFFFull;
    Vec v2r, g2r, ge1, ge2;
    or_groupby_an(&vg, &cg, 65117, &v2r, &g2r, &ge1, &ge2);
    expect_eq("Q2.1 groupbyAN VOIDtoRGID", tail_of(std::get<0>(tg)), v2r.v);
    expect_eq("Q2.1 groupbyAN VGIDtoROID", tail_of(std::get<1>(tg)), g2r.v);
    expect_vec("Q2.1 groupbyAN tail errors", std::get<3>(tg), ge2.v);

    // a second grouping attribute on the same rows: the binary group-by
    auto *a2 = column<v2_restiny_t>(uniform<uint16_t>(cg.n, 24, 0, 4, 29, 1e-3), 29);
    const size_t ng = std::get<1>(tg)->size();
    auto tg2 = groupbyAN(a2, std::get<0>(tg), ng);
    or_col c2 = ocol_of(a2->tail, 0), cgr = ocol_of(std::get<0>(tg)->tail, 0);
    Vec b2r, bg2r, be1, be2, be3;
    or_groupby2_an(&vg, &c2, &cgr, ng, 32, 65117, &b2r, &bg2r, &be1, &be2, &be3);
    expect_eq("Q2.1 groupbyAN binary VOIDtoRGID", tail_of(std::get<0>(tg2)), b2r.v);
    expect_eq("Q2.1 groupbyAN binary VGIDtoROID", tail_of(std::get<1>(tg2)), bg2r.v);
    expect_vec("Q2.1 groupbyAN binary grouping errors", std::get<4>(tg2), be3.v);

    // grouped sum over the first cg.n measure codes
    auto *m = column<v2_resint_t>(std::vector<uint64_t>(fm->tail.container->host().begin(), fm->tail.container->host().begin() + cg.n), 65117);
    const size_t ng2 = std::get<1>(tg2)->size();
    auto ts = aggregate_sum_groupedAN<v2_resbigint_t>(m, std::get<0>(tg2), ng2);
    or_col cm = ocol_of(m->tail, 0), cg2 = ocol_of(std::get<0>(tg2)->tail, 0);
    std::vector<uint64_t> ref_sums(ng2);
    Vec s1, s2, s3;
    (void) or_aggregate_sum_grouped_an(&vg, &cm, &cg2, ng2, 65117, 65117, ref_sums.data(), &s1, &s2, &s3);
    bool ok = tail_of(std::get<0>(ts)) == ref_sums;
    std::printf("%-58s %s\n", "Q2.1 aggregate_sum_groupedAN<resbigint>", ok ? "ok" : "FAILED");
    g_failures += !ok;
    expect_vec("Q2.1 aggregate_sum_groupedAN tail errors", std::get<2>(ts), s2.v);
    expect_vec("Q2.1 aggregate_sum_groupedAN group errors", std::get<3>(ts), s3.v);

    // arithmeticAN (profit = revenue - supplycost shape) on two measure columns
    auto *m2 = column<v2_resint_t>(uniform<uint64_t>(cg.n, 25, 0, 90000, 63877, 1e-4), 63877);
    auto ta = arithmeticAN<std::minus, v2_resbigint_t>(m, m2);
    or_col cm2 = ocol_of(m2->tail, 0);
    Vec ar, a1, a2v, a3, a4;
    (void) or_arithmetic_an(&vg, &cm, &vg, &cm2, 1, 65117, 8, 16089, &ar, &a1, &a2v, &a3, &a4);
    expect_eq("arithmeticAN<minus, resbigint>", tail_of(std::get<0>(ta)), ar.v);
    expect_vec("arithmeticAN e2", std::get<2>(ta), a2v.v);
    expect_vec("arithmeticAN e4", std::get<4>(ta), a4.v);

    bool threw = false;
    try {
        auto bad = aggregate_sum_groupedAN<v2_resbigint_t>(fm, std::get<0>(tg2), ng2);
        (void) bad;
    } catch (const std::runtime_error &) {
        threw = true;
    }
    std::printf("%-58s %s\n", "aggregate_sum_groupedAN size mismatch throws", threw ? "ok" : "FAILED");
    g_failures += !threw;

    delete std::get<0>(ta);
    delete std::get<0>(ts);
    delete std::get<0>(tg2);
    delete std::get<1>(tg2);
    delete std::get<0>(tg);
    delete std::get<1>(tg);
    delete std::get<0>(tf);
    delete std::get<0>(tj);
    delete dim_oids;
    delete rev;
    delete m;
    delete m2;
    delete a2;
    delete fk;
    delete fm;
    delete dk;
    delete da;
}

// checkAndDecodeAN on a BAT with an encoded head (encdecAN_scalar.tcc:129-176) and the seqbase-
// aligned matchjoinAN <H1, void, void, T2> (matchjoinAN.tcc:146-224).
void test_general_forms() {
    const size_t n = 100003;
    auto *a = column<v2_resint_t>(uniform<uint64_t>(n, 31, 0, 1000000, 65117, 1e-3), 65117);
    auto *b = column<v2_resshort_t>(uniform<uint32_t>(n, 32, 0, 60000, 233, 1e-3), 233);
    // <resint, void> from reverse, then mirror a <resint, resshort> BAT by pairing the two columns
    auto pr = matchjoinAN(a->reverse(), b);    // <H1 = resint, void> x <void, resshort>: seqbase form
    or_col ca = ocol_of(a->tail, n), cb = ocol_of(b->tail, n), v1 {}, v2 {};
    v1.n = v2.n = n;
    v1.width = v2.width = 4;
    v1.unenc_bits = v2.unenc_bits = 32;
    v1.a = v1.a_inv = v2.a = v2.a_inv = 1;
    v1.max_u = v2.max_u = 0xFFFFFFFFull;
    Vec mh, mt, m1, m4;
    or_matchjoin_seq_an(&ca, &v1, &v2, &cb, 0, 0, 8, 0, 0, 4, 16089, &mh, &mt, &m1, &m4);
    expect_eq("matchjoinAN <H1, void, void, T2> heads", head_of(std::get<0>(pr)), mh.v);
    expect_eq("matchjoinAN <H1, void, void, T2> tails", tail_of(std::get<0>(pr)), mt.v);
    expect_vec("matchjoinAN <H1, void, void, T2> e1", std::get<1>(pr), m1.v);
    expect_vec("matchjoinAN <H1, void, void, T2> e4", std::get<4>(pr), m4.v);
    auto cd = checkAndDecodeAN(std::get<0>(pr));
    or_col ph = ocol_of(std::get<0>(pr)->head, 0), pt = ocol_of(std::get<0>(pr)->tail, 0);
    Vec dh, dt, eh, et;
    or_check_and_decode_an2(&ph, &pt, 16089, &dh, &dt, &eh, &et);
    expect_eq("checkAndDecodeAN<resint, resshort> heads", head_of(std::get<0>(cd)), dh.v);
    expect_eq("checkAndDecodeAN<resint, resshort> tails", tail_of(std::get<0>(cd)), dt.v);
    expect_vec("checkAndDecodeAN<resint, resshort> head errors", std::get<1>(cd), eh.v);
    expect_vec("checkAndDecodeAN<resint, resshort> tail errors", std::get<2>(cd), et.v);
    delete std::get<0>(cd);
    delete std::get<2>(pr);
    delete std::get<3>(pr);
    delete std::get<0>(pr);
    delete a;
    delete b;
}

}  // namespace

int main() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        std::printf("no GPU\n");
        return 2;
    }
    test_select_codec();
    test_q11_chain();
    test_star_chain();
    test_general_forms();
    std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "PASSED", g_failures);
    return g_failures ? 1 : 0;
}
