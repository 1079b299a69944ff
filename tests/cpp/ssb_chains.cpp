// ssb_chains.cpp -- the thirteen SSB queries of tools/ssbm/ssbm-q11 .. q43_continuous.cpp as
// operator chains written against the C++ operator templates (include/ahead/column_operators/
// OperatorsAN.hpp), the way the reference's drivers call them, over device-resident BATs.
//
// usage: ssb_chains <dir>
//   <dir>/tables.txt   one line per column: name kind(tiny|short|int) A rows; codes in <dir>/<name>.bin
//   <dir>/queries.txt  the query specifications (written by tests/test_cpp_ssb.py from ahead_amd.ssb)
// writes <dir>/out_<query>.txt: rows, the Q1.x result or the groups (sums, keys), and every
// non-empty error vector by the name the Python chains give it (operator tag + ".e" + position);
// and <dir>/out_<query>_fused.txt: the same query through the fused star query of
// include/ahead/column_operators/StarQueryAN.hpp after the chain's dimension selections (the
// chain's result where the fused form does not apply), its error vectors named as ahead_amd.ssb's
// fused form names them.
// Exit status 0 when every query ran; the comparison with the oracle is the test's.
#include <column_operators/OperatorsAN.hpp>
#include <column_operators/StarQueryAN.hpp>

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <variant>
#include <vector>

using namespace ahead;
using namespace ahead::bat::ops;
using namespace ahead::bat::ops::scalar;

namespace {

using TinyB = BAT<v2_void_t, v2_restiny_t>;
using ShortB = BAT<v2_void_t, v2_resshort_t>;
using IntB = BAT<v2_void_t, v2_resint_t>;
using OidB = BAT<v2_resoid_t, v2_resoid_t>;
using RowsB = BAT<v2_void_t, v2_resoid_t>;
using Col = std::variant<TinyB *, ShortB *, IntB *>;

constexpr resoid_t A_OIDX = 65117;                    // v2_resoid_t::As.back()
const resoid_t A_OIDX_INV = ext_euclidean(A_OIDX, 64);
constexpr resint_t A_MEAS = 65117;                    // the profit's A (RESINT load default)

enum { LT = 0, LE = 1, GT = 2, GE = 3, EQ = 4, NE = 5 };
enum { NONE_ = 0, AND = 1, OR = 2 };

struct Pred {
    int op1 = 0, op2 = 0, comb = NONE_;
    uint64_t v1 = 0, v2 = 0;
    bool two = false;
};

struct DimQ {
    std::string name, attr;
    uint32_t flags = 0;
    std::vector<std::pair<std::string, Pred>> sels;
};

struct Query {
    std::string name, family;
    std::vector<std::pair<std::string, Pred>> preds;
    std::vector<DimQ> dims;
    // the fused form's measure columns, kind and early-fetch mask; the dimensions grouped by
    uint32_t kind = 0, pre_mask = 0;
    std::vector<std::string> meas;
    std::vector<int> groups;
};

std::map<std::string, Col> T;

// Error vectors in call order, named as ahead_amd.ssb.Errors names them.
struct Errors {
    std::vector<std::pair<std::string, std::vector<uint64_t>>> items;

    void put(const std::string &name, int k, AN_indicator_vector *v) {
        if (v != nullptr && !v->empty())
            items.push_back({name + ".e" + std::to_string(k), std::vector<uint64_t>(v->begin(), v->end())});
        delete v;
    }
    void add_vec(const std::string &name, const std::vector<uint64_t> &v) {
        if (!v.empty())
            items.push_back({name + ".e1", v});
    }
    template<typename... V>
    void add(const std::string &name, V *... vs) {
        int k = 0;
        ((++k, put(name, k, vs)), ...);
    }
};

// The first element of an operator's result tuple; the error vectors go to E under `name`.
template<typename Tup, size_t... I>
void add_rest(Errors &E, const std::string &name, Tup &t, std::index_sequence<I...>) {
    E.add(name, std::get<I + 1>(t)...);
}
template<typename Tup>
auto take(Errors &E, const std::string &name, Tup t) {
    add_rest(E, name, t, std::make_index_sequence<std::tuple_size_v<Tup> - 1> {});
    return std::get<0>(t);
}

template<typename B>
B *col(const std::string &name) {
    return std::get<B *>(T.at(name));
}

// selectAN on a column, as its mirror_head (the OID list the chains go on with).
template<typename X>
OidB *select_mirror(BAT<v2_void_t, X> *b, const Pred &p, Errors &E, const std::string &attr) {
    typedef typename X::type_t code_t;
    const uint64_t A = b->tail.metaData.AN_A;
    const auto th = [&](uint64_t v) { return static_cast<code_t>(v * A); };
    const auto fin = [&](auto pr) {
        E.add("select_" + attr, pr.second);
        OidB *m = pr.first->mirror_head();
        delete pr.first;
        return m;
    };
    if (!p.two) {
        switch (p.op1) {
            case LT: return fin(selectAN<std::less>(b, th(p.v1)));
            case LE: return fin(selectAN<std::less_equal>(b, th(p.v1)));
            case GT: return fin(selectAN<std::greater>(b, th(p.v1)));
            case GE: return fin(selectAN<std::greater_equal>(b, th(p.v1)));
            case EQ: return fin(selectAN<std::equal_to>(b, th(p.v1)));
            case NE: return fin(selectAN<std::not_equal_to>(b, th(p.v1)));
        }
    } else if (p.op1 == GE && p.op2 == LE && p.comb == AND) {
        return fin(selectAN<std::greater_equal, std::less_equal, ahead::and_is>(b, th(p.v1), th(p.v2)));
    } else if (p.op1 == EQ && p.op2 == EQ && p.comb == OR) {
        return fin(selectAN<std::equal_to, std::equal_to, ahead::or_is>(b, th(p.v1), th(p.v2)));
    }
    throw std::runtime_error("select: predicate form not used by the SSB queries");
}

OidB *sel(const std::string &attr, const Pred &p, Errors &E) {
    return std::visit([&](auto *b) { return select_mirror(b, p, E, attr); }, T.at(attr));
}

// selectAN on a dimension attribute, then matchjoinAN(reverse(key), mirror_head(sel)): key | OID
// dimension (ssbm-q21_continuous.cpp:87-93).
auto sel_dim(const std::string &key, const std::pair<std::string, Pred> &s, Errors &E) {
    OidB *m = sel(s.first, s.second, E);
    return take(E, "match_" + s.first, matchjoinAN(col<IntB>(key)->reverse(), m));
}

// fetchjoinAN of a column through an OID list; the result keeps the column's type.
Col fetch_attr(RowsB *oids, const std::string &name, Errors &E, const std::string &tag) {
    return std::visit([&](auto *b) -> Col { return take(E, tag, fetchjoinAN(oids, b)); }, T.at(name));
}

// fetchjoinAN(rows, fact FK), then hashjoinAN with the dimension's (key | OID) BAT, re-encoding
// both sides: VOID | OID dimension for every row (ssbm-q31_continuous.cpp:153-160).
template<typename Build>
RowsB *join_dim_oids(Errors &E, RowsB *rows, const std::string &fact_fk, Build *build, const std::string &tag) {
    auto *fk = take(E, "fetch_" + tag, fetchjoinAN(rows, col<IntB>(fact_fk)));
    auto *j = take(E, "join_" + tag, hashjoinAN(fk, build, A_OIDX, A_OIDX_INV, A_OIDX, A_OIDX_INV));
    return j->clear_head();
}

// hashjoinAN of the whole fact FK column with the dimension's (key | OID) BAT.
template<typename Dim>
OidB *probe_dim(Errors &E, const std::string &fact_fk, Dim *dim) {
    auto *r = take(E, "join_" + fact_fk,
                   hashjoinAN(col<IntB>(fact_fk), dim, A_OIDX, A_OIDX_INV, dim->tail.metaData.AN_A, dim->tail.metaData.AN_Ainv));
    return r->mirror_head();
}

// matchjoinAN(rows, fact FK), then hashjoinAN with the dimension: the rows that join.
template<typename Rows, typename Dim>
OidB *fetch_join(Errors &E, Rows *rows, const std::string &fact_fk, Dim *dim, const std::string &tag) {
    auto *f = take(E, "fetch_" + tag, matchjoinAN(rows, col<IntB>(fact_fk)));
    auto *j = take(E, "join_" + tag, hashjoinAN(f, dim));
    return j->mirror_head();
}

IntB *profit(Errors &E, RowsB *rows) {
    auto *ar = take(E, "fetch_revenue", fetchjoinAN(rows, col<IntB>("lo_revenue")));
    auto *asc = take(E, "fetch_supplycost", fetchjoinAN(rows, col<IntB>("lo_supplycost")));
    return take(E, "profit", arithmeticAN<std::minus, v2_resint_t>(ar, asc, A_MEAS, ext_euclidean(A_MEAS, 64)));
}

struct Result {
    uint64_t rows = 0, groups = 0;
    bool scalar = false;
    uint64_t result = 0;
    std::vector<uint64_t> sums;
    std::vector<std::vector<uint64_t>> keys;
    Errors E;
};

// groupbyAN over the attributes (unary, then binary), grouped sum, and the group keys.
void groups(Result &R, const std::vector<Col> &attrs, IntB *measure) {
    Errors &E = R.E;
    auto g = std::visit([&](auto *a) { return groupbyAN(a); }, attrs[0]);
    E.add("group_0", std::get<2>(g), std::get<3>(g));
    RowsB *v2r = std::get<0>(g), *g2r = std::get<1>(g);
    for (size_t k = 1; k < attrs.size(); ++k) {
        auto g2 = std::visit([&](auto *a) { return groupbyAN(a, v2r, g2r->size()); }, attrs[k]);
        E.add("group_" + std::to_string(k), std::get<2>(g2), std::get<3>(g2), std::get<4>(g2));
        v2r = std::get<0>(g2);
        g2r = std::get<1>(g2);
    }
    auto *s = take(E, "grouped_sum", aggregate_sum_groupedAN<v2_resbigint_t>(measure, v2r, g2r->size()));
    for (size_t k = 0; k < attrs.size(); ++k) {
        R.keys.push_back(std::visit([&](auto *a) {
            auto *r = take(E, "group_key_" + std::to_string(k), fetchjoinAN(g2r, a));
            typedef typename std::remove_pointer_t<decltype(r)>::tail_t code_t;
            const uint64_t inv = r->tail.metaData.AN_Ainv;
            const uint64_t mask = sizeof(code_t) >= 8 ? ~0ull : ((1ull << (8 * sizeof(code_t))) - 1);
            std::vector<uint64_t> keys;
            for (code_t c : r->tail.container->host())
                keys.push_back((static_cast<uint64_t>(c) * inv) & mask);
            return keys;
        }, attrs[k]));
    }
    const uint64_t sinv = s->tail.metaData.AN_Ainv;
    for (uint64_t v : s->tail.container->host())
        R.sums.push_back(v * sinv);
    R.groups = g2r->size();
}

Result chain_q1(const Query &q) {
    Result R;
    Errors &E = R.E;
    const auto &[qc, qp] = q.preds[0];
    const auto &[dc, dp] = q.preds[1];
    OidB *s1 = sel(qc, qp, E);
    OidB *s2 = sel(dc, dp, E);
    auto *r4 = take(E, "match_selections", matchjoinAN(s1, s2));
    auto *bat5 = r4->mirror_head();
    auto *r6 = take(E, "fetch_orderdate", matchjoinAN(bat5, col<IntB>("lo_orderdate")));
    const auto &sels = q.dims[0].sels;
    OidB *s7 = sel(sels[0].first, sels[0].second, E);
    if (sels.size() > 1) {  // d_year and d_weeknuminyear (ssbm-q13_continuous.cpp:100-110)
        OidB *s8 = sel(sels[1].first, sels[1].second, E);
        auto *rA = take(E, "match_date_selections", matchjoinAN(s7, s8));
        s7 = rA->mirror_head();
    }
    auto *r9 = take(E, "fetch_datekey", matchjoinAN(s7, col<IntB>("d_datekey")));
    auto *rB = take(E, "join_date", hashjoinAN(r6, r9->reverse()));
    auto *batC = rB->mirror_head();
    auto *rD = take(E, "fetch_price", matchjoinAN(batC, col<IntB>("lo_extendedprice")));
    auto *rE = take(E, "fetch_discount", matchjoinAN(bat5, col<TinyB>(dc)));
    auto *rF = take(E, "match_discount", matchjoinAN(batC, rE));
    auto *rG = take(E, "mul_sum", aggregate_mul_sumAN<v2_resbigint_t>(rD, rF));
    R.scalar = true;
    R.result = rG->tail.container->host()[0] * rG->tail.metaData.AN_Ainv;
    R.rows = rD->size();
    return R;
}

Result chain_q2(const Query &q) {
    Result R;
    Errors &E = R.E;
    auto *supp = sel_dim("s_suppkey", q.dims[0].sels[0], E);
    auto *r5 = take(E, "join_supplier",
                    hashjoinAN(col<IntB>("lo_suppkey"), supp, A_OIDX, A_OIDX_INV, supp->tail.metaData.AN_A, supp->tail.metaData.AN_Ainv));
    auto *r7 = take(E, "fetch_partkey", matchjoinAN(r5->mirror_head(), col<IntB>("lo_partkey")));
    auto *part = sel_dim("p_partkey", q.dims[1].sels[0], E);
    auto *rC = take(E, "join_part", hashjoinAN(r7, part));
    auto *rF = take(E, "fetch_orderdate", matchjoinAN(rC->mirror_head(), col<IntB>("lo_orderdate")));
    auto *rI = take(E, "join_date", hashjoinAN(rF, col<IntB>("d_datekey")->reverse(), rF->head.metaData.AN_A,
                                               rF->head.metaData.AN_Ainv, A_OIDX, A_OIDX_INV));
    auto *batW = rI->mirror_head();
    auto *rX = take(E, "fetch_partkey_2", matchjoinAN(batW, col<IntB>("lo_partkey")));
    auto *rZ = take(E, "join_part_2", hashjoinAN(rX, col<IntB>("p_partkey")->reverse(), rX->head.metaData.AN_A,
                                                 rX->head.metaData.AN_Ainv, A_OIDX, A_OIDX_INV));
    Col ay = fetch_attr(rI->clear_head(), "d_year", E, "fetch_year");
    Col ab = fetch_attr(rZ->clear_head(), "p_brand", E, "fetch_brand");
    auto *ar = take(E, "fetch_revenue", fetchjoinAN(batW->clear_head(), col<IntB>("lo_revenue")));
    groups(R, {ay, ab}, ar);
    R.rows = ar->size();
    return R;
}

Result chain_q3(const Query &q) {
    Result R;
    Errors &E = R.E;
    const DimQ &ds = q.dims[0], &dc = q.dims[1], &dd = q.dims[2];
    auto *supp = sel_dim("s_suppkey", ds.sels[0], E);
    auto *r5 = take(E, "join_supplier",
                    hashjoinAN(col<IntB>("lo_suppkey"), supp, A_OIDX, A_OIDX_INV, supp->tail.metaData.AN_A, supp->tail.metaData.AN_Ainv));
    auto *cust = sel_dim("c_custkey", dc.sels[0], E);
    auto *r11 = take(E, "fetch_custkey", matchjoinAN(r5->mirror_head(), col<IntB>("lo_custkey")));
    auto *r12 = take(E, "join_customer", hashjoinAN(r11, cust));
    auto *date = sel_dim("d_datekey", dd.sels[0], E);
    auto *r18 = take(E, "fetch_orderdate", matchjoinAN(r12->mirror_head(), col<IntB>("lo_orderdate")));
    auto *r19 = take(E, "join_date", hashjoinAN(r18, date));
    RowsB *rows = r19->mirror_head()->clear_head();
    auto *ar = take(E, "fetch_revenue", fetchjoinAN(rows, col<IntB>("lo_revenue")));
    Col as = fetch_attr(join_dim_oids(E, rows, "lo_suppkey", supp, "suppkey"), ds.attr, E, "fetch_" + ds.attr);
    Col ac = fetch_attr(join_dim_oids(E, rows, "lo_custkey", cust, "custkey"), dc.attr, E, "fetch_" + dc.attr);
    Col ad = fetch_attr(join_dim_oids(E, rows, "lo_orderdate", date, "orderdate"), "d_year", E, "fetch_year");
    groups(R, {ad, as, ac}, ar);
    R.rows = ar->size();
    return R;
}

Result chain_q41(const Query &q) {
    Result R;
    Errors &E = R.E;
    const DimQ &dp = q.dims[0], &dsu = q.dims[1], &dcu = q.dims[2];
    auto *part = sel_dim("p_partkey", dp.sels[0], E);
    OidB *bat6 = probe_dim(E, "lo_partkey", part);
    auto *supp = sel_dim("s_suppkey", dsu.sels[0], E);
    OidB *r12 = fetch_join(E, bat6, "lo_suppkey", supp, "suppkey");
    auto *cust = sel_dim("c_custkey", dcu.sels[0], E);
    OidB *r19 = fetch_join(E, r12, "lo_custkey", cust, "custkey");
    RowsB *rows = r19->clear_head();
    IntB *pr = profit(E, rows);
    Col an = fetch_attr(join_dim_oids(E, rows, "lo_custkey", cust, "custkey_2"), dcu.attr, E, "fetch_" + dcu.attr);
    Col ay = fetch_attr(join_dim_oids(E, rows, "lo_orderdate", col<IntB>("d_datekey")->reverse(), "orderdate"), "d_year", E,
                        "fetch_year");
    groups(R, {ay, an}, pr);
    R.rows = pr->size();
    return R;
}

Result chain_q42(const Query &q) {
    Result R;
    Errors &E = R.E;
    const DimQ &dp = q.dims[0], &dsu = q.dims[1], &dcu = q.dims[2], &dd = q.dims[3];
    auto *part = sel_dim("p_partkey", dp.sels[0], E);
    OidB *bat6 = probe_dim(E, "lo_partkey", part);
    auto *supp = sel_dim("s_suppkey", dsu.sels[0], E);
    OidB *bat13 = fetch_join(E, bat6, "lo_suppkey", supp, "suppkey");
    auto *cust = sel_dim("c_custkey", dcu.sels[0], E);
    OidB *bat20 = fetch_join(E, bat13, "lo_custkey", cust, "custkey");
    auto *date = sel_dim("d_datekey", dd.sels[0], E);
    OidB *bat27 = fetch_join(E, bat20, "lo_orderdate", date, "orderdate");
    RowsB *rows = bat27->clear_head();
    IntB *pr = profit(E, rows);
    Col an = fetch_attr(join_dim_oids(E, rows, "lo_suppkey", supp, "suppkey_2"), dsu.attr, E, "fetch_" + dsu.attr);
    Col ay = fetch_attr(join_dim_oids(E, rows, "lo_orderdate", date, "orderdate_2"), "d_year", E, "fetch_year");
    Col ac = fetch_attr(join_dim_oids(E, rows, "lo_partkey", col<IntB>("p_partkey")->reverse(), "partkey_2"), dp.attr, E,
                        "fetch_" + dp.attr);
    groups(R, {ay, an, ac}, pr);
    R.rows = pr->size();
    return R;
}

Result chain_q43(const Query &q) {
    Result R;
    Errors &E = R.E;
    const DimQ &dp = q.dims[0], &dsu = q.dims[1], &dd = q.dims[2];
    auto *part = sel_dim("p_partkey", dp.sels[0], E);
    OidB *bat6 = probe_dim(E, "lo_partkey", part);
    auto *supp = sel_dim("s_suppkey", dsu.sels[0], E);
    OidB *bat13 = fetch_join(E, bat6, "lo_suppkey", supp, "suppkey");
    auto *date = sel_dim("d_datekey", dd.sels[0], E);
    OidB *bat27 = fetch_join(E, bat13, "lo_orderdate", date, "orderdate");
    RowsB *rows = bat27->clear_head();
    IntB *pr = profit(E, rows);
    Col ac = fetch_attr(join_dim_oids(E, rows, "lo_suppkey", supp, "suppkey_2"), dsu.attr, E, "fetch_" + dsu.attr);
    Col ay = fetch_attr(join_dim_oids(E, rows, "lo_orderdate", date, "orderdate_2"), "d_year", E, "fetch_year");
    Col ab = fetch_attr(join_dim_oids(E, rows, "lo_partkey", col<IntB>("p_partkey")->reverse(), "partkey_2"), dp.attr, E,
                        "fetch_" + dp.attr);
    groups(R, {ay, ac, ab}, pr);
    R.rows = pr->size();
    return R;
}

template<typename V2>
TempBAT<v2_void_t, V2> *load_column(const std::string &path, uint64_t A, size_t n) {
    typedef typename V2::type_t code_t;
    std::vector<code_t> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(code_t))))
        throw std::runtime_error("cannot read " + path);
    ColumnMetaData md(size_bytes<code_t>, static_cast<A_t>(A), ext_euclidean(A, 8 * sizeof(code_t)), V2::UNENC_MAX_U, 0);
    return make_bat<V2>(v, md);
}

const std::map<std::string, std::pair<std::string, std::string>> DIM_KEYS = {
    {"date", {"lo_orderdate", "d_datekey"}}, {"cust", {"lo_custkey", "c_custkey"}},
    {"supp", {"lo_suppkey", "s_suppkey"}}, {"part", {"lo_partkey", "p_partkey"}}};
const char *FUSED_KINDS[AHEAD_SQ_KINDS] = {"pred0", "pred1", "meas0", "meas1", "fk0", "fk1", "fk2", "fk3", "refetch0",
                                           "refetch1", "refetch2", "refetch3", "attr0", "attr1", "attr2", "attr3"};

ahead_col desc_of(const std::string &name) {
    return std::visit([](auto *b) { return device::col(b->tail, device::rows(b)); }, T.at(name));
}

// The query as a fused star query (StarQueryAN.hpp) after the chain's dimension selections, with
// the chain as the fallback; the result in the chains' form.
Result fused_query(const Query &q, Result (*chain)(const Query &)) {
    Result R;
    Errors &E = R.E;
    std::vector<StarDimAN> dims;
    for (const DimQ &d : q.dims) {
        const auto &[fk, key] = DIM_KEYS.at(d.name);
        StarDimAN sd = d.attr.empty() ? starDim(col<IntB>(fk), col<IntB>(key), d.flags)
                                      : std::visit([&](auto *a) { return starDim(col<IntB>(fk), col<IntB>(key), a, d.flags); },
                                                   T.at(d.attr));
        if (!d.sels.empty()) {
            OidB *s = sel(d.sels[0].first, d.sels[0].second, E);
            if (d.sels.size() > 1) {  // two selections on one dimension, intersected (ssbm-q13_continuous.cpp:100-110)
                OidB *s2 = sel(d.sels[1].first, d.sels[1].second, E);
                s = take(E, "match_" + d.name + "_selections", matchjoinAN(s, s2))->mirror_head();
            }
            starSelect(sd, s);
        }
        dims.push_back(sd);
    }
    std::vector<StarPredAN> preds;
    for (const auto &[c, p] : q.preds) {
        StarPredAN sp;
        sp.col = desc_of(c);
        const uint64_t A = sp.col.a;
        sp.op1 = p.op1;
        sp.th1 = p.v1 * A;
        if (p.two) {
            sp.op2 = p.op2;
            sp.th2 = p.v2 * A;
            sp.comb = p.comb;
        }
        preds.push_back(sp);
    }
    std::vector<ahead_col> meas;
    for (const std::string &m : q.meas)
        meas.push_back(desc_of(m));
    auto out = starQueryOrChainAN(meas[0].n, dims, preds, meas, q.kind, q.pre_mask, [&] { return chain(q); });
    if (out.chained)
        return *out.chained;
    const StarResultAN &F = out.fused;
    for (int k = 0; k < AHEAD_SQ_KINDS; ++k)
        E.add_vec(k < static_cast<int>(q.preds.size()) ? "select_" + q.preds[k].first : std::string("fused_") + FUSED_KINDS[k],
                  F.err[k]);
    const uint64_t inv = ext_euclidean(static_cast<uint64_t>(65117), 64);
    R.rows = F.rows;
    if (q.family == "q1") {
        R.scalar = true;
        R.result = F.sums.empty() ? 0 : F.sums[0] * inv;
        return R;
    }
    R.groups = F.groups;
    for (uint64_t v : F.sums)
        R.sums.push_back(v * inv);
    for (int d : q.groups) {
        const ahead_col a = desc_of(q.dims[d].attr);
        const uint64_t mask = a.width >= 8 ? ~0ull : ((1ull << (8 * a.width)) - 1);
        std::vector<uint64_t> keys;
        for (uint64_t c : F.keys[d])
            keys.push_back((c * a.a_inv) & mask);
        R.keys.push_back(keys);
    }
    return R;
}

Pred parse_pred(std::istringstream &in) {
    Pred p;
    in >> p.op1 >> p.v1;
    if (in >> p.op2 >> p.v2 >> p.comb)
        p.two = true;
    return p;
}

std::vector<Query> load_queries(const std::string &path) {
    std::ifstream f(path);
    std::vector<Query> qs;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string tag;
        in >> tag;
        if (tag == "query") {
            qs.emplace_back();
            in >> qs.back().name >> qs.back().family;
        } else if (tag == "pred") {
            std::string c;
            in >> c;
            qs.back().preds.push_back({c, parse_pred(in)});
        } else if (tag == "dim") {
            DimQ d;
            in >> d.name >> d.attr >> d.flags;
            if (d.attr == "-")
                d.attr.clear();
            qs.back().dims.push_back(d);
        } else if (tag == "fused") {
            in >> qs.back().kind >> qs.back().pre_mask;
            std::string m;
            while (in >> m)
                qs.back().meas.push_back(m);
        } else if (tag == "groups") {
            int g;
            while (in >> g)
                qs.back().groups.push_back(g);
        } else if (tag == "sel") {
            std::string c;
            in >> c;
            qs.back().dims.back().sels.push_back({c, parse_pred(in)});
        }
    }
    return qs;
}

void write_result(const std::string &path, const Result &R) {
    std::ofstream o(path);
    o << "rows " << R.rows << "\n";
    if (R.scalar) {
        o << "result " << R.result << "\n";
    } else {
        o << "groups " << R.groups << "\nsums";
        for (uint64_t v : R.sums)
            o << ' ' << v;
        o << "\n";
        for (const auto &k : R.keys) {
            o << "keys";
            for (uint64_t v : k)
                o << ' ' << v;
            o << "\n";
        }
    }
    for (const auto &[name, v] : R.E.items) {
        o << "err " << name;
        for (uint64_t x : v)
            o << ' ' << x;
        o << "\n";
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        std::printf("no GPU\n");
        return 2;
    }
    const std::string dir = argv[1];
    std::ifstream tf(dir + "/tables.txt");
    std::string name, kind;
    uint64_t A = 0;
    size_t n = 0;
    while (tf >> name >> kind >> A >> n) {
        const std::string path = dir + "/" + name + ".bin";
        if (kind == "tiny")
            T[name] = load_column<v2_restiny_t>(path, A, n);
        else if (kind == "short")
            T[name] = load_column<v2_resshort_t>(path, A, n);
        else
            T[name] = load_column<v2_resint_t>(path, A, n);
    }
    const std::map<std::string, Result (*)(const Query &)> chains = {
        {"q1", chain_q1}, {"q2", chain_q2}, {"q3", chain_q3}, {"q41", chain_q41}, {"q42", chain_q42}, {"q43", chain_q43}};
    int failed = 0;
    for (const Query &q : load_queries(dir + "/queries.txt")) {
        try {
            const Result R = chains.at(q.family)(q);
            write_result(dir + "/out_" + q.name + ".txt", R);
            std::printf("%s (%s): %llu rows, %zu error vectors\n", q.name.c_str(), q.family.c_str(),
                        static_cast<unsigned long long>(R.rows), R.E.items.size());
        } catch (const std::exception &e) {
            std::printf("%s (%s): FAILED: %s\n", q.name.c_str(), q.family.c_str(), e.what());
            ++failed;
        }
        try {
            const Result F = fused_query(q, chains.at(q.family));
            write_result(dir + "/out_" + q.name + "_fused.txt", F);
            std::printf("%s fused: %llu rows, %zu error vectors\n", q.name.c_str(), static_cast<unsigned long long>(F.rows),
                        F.E.items.size());
        } catch (const std::exception &e) {
            std::printf("%s fused: FAILED: %s\n", q.name.c_str(), e.what());
            ++failed;
        }
        std::fflush(stdout);
    }
    std::printf("%s\n", failed ? "FAILED" : "done");
    return failed ? 1 : 0;
}
