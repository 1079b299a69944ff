// serve_client.cpp -- the served form from a C++ host: the SSB tables of <dir> loaded into device
// BATs, every query of <dir>/queries.txt submitted to the library's request loop (ahead_serve_*,
// include/ahead_hip.h) at once, and each result written as <dir>/serve_<name>.txt in the form
// ssb_chains writes (rows, the Q1.x result or the groups, the error vectors named as
// ahead_amd.ssb's fused form names them).
//
// usage: serve_client <dir> <slots>
//   <dir>/tables.txt and <dir>/queries.txt as for ssb_chains (one selection per dimension).
// Exit status 0 when every query was served with status 0; the comparison with the oracle is the
// test's (tests/test_cpp_serve.py).
#include <column_operators/OperatorsAN.hpp>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

using namespace ahead;
using namespace ahead::bat::ops;

namespace {

using TinyB = BAT<v2_void_t, v2_restiny_t>;
using ShortB = BAT<v2_void_t, v2_resshort_t>;
using IntB = BAT<v2_void_t, v2_resint_t>;
using Col = std::variant<TinyB *, ShortB *, IntB *>;

std::map<std::string, Col> T;

const std::map<std::string, std::pair<std::string, std::string>> DIM_KEYS = {
    {"date", {"lo_orderdate", "d_datekey"}}, {"cust", {"lo_custkey", "c_custkey"}},
    {"supp", {"lo_suppkey", "s_suppkey"}}, {"part", {"lo_partkey", "p_partkey"}}};
const char *FUSED_KINDS[AHEAD_SQ_KINDS] = {"pred0", "pred1", "meas0", "meas1", "fk0", "fk1", "fk2", "fk3", "refetch0",
                                           "refetch1", "refetch2", "refetch3", "attr0", "attr1", "attr2", "attr3"};
constexpr uint64_t A_RESULT = 65117;  // v2_resbigint_t's A of the sums

struct Pred {
    int op1 = AHEAD_LT, op2 = AHEAD_LT, comb = AHEAD_COMB_NONE;
    uint64_t v1 = 0, v2 = 0;
};

struct DimQ {
    std::string name, attr;
    uint32_t flags = 0;
    std::vector<std::pair<std::string, Pred>> sels;
};

struct Query {
    std::string name, family;
    uint32_t kind = 0, pre_mask = 0;
    std::vector<std::string> meas;
    std::vector<int> groups;
    std::vector<std::pair<std::string, Pred>> preds;
    std::vector<DimQ> dims;
};

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

ahead_col desc_of(const std::string &name) {
    return std::visit([](auto *b) { return device::col(b->tail, device::rows(b)); }, T.at(name));
}

Pred parse_pred(std::istringstream &in) {
    Pred p;
    in >> p.op1 >> p.v1;
    if (!(in >> p.op2 >> p.v2 >> p.comb)) {
        p.op2 = AHEAD_LT;
        p.v2 = 0;
        p.comb = AHEAD_COMB_NONE;
    }
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
        } else if (tag == "fused") {
            in >> qs.back().kind >> qs.back().pre_mask;
            std::string m;
            while (in >> m)
                qs.back().meas.push_back(m);
        } else if (tag == "groups") {
            int g;
            while (in >> g)
                qs.back().groups.push_back(g);
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
        } else if (tag == "sel") {
            std::string c;
            in >> c;
            qs.back().dims.back().sels.push_back({c, parse_pred(in)});
        }
    }
    return qs;
}

// The query as the request loop takes it: the star query's columns and, per dimension, the
// selection the server runs.
ahead_serve_query serve_query(const Query &q) {
    ahead_serve_query s {};
    ahead_sq_args &a = s.q;
    const ahead_col m0 = desc_of(q.meas.at(0));
    a.n_fact = m0.n;
    a.n_dims = static_cast<uint32_t>(q.dims.size());
    a.n_preds = static_cast<uint32_t>(q.preds.size());
    a.meas_kind = q.kind;
    a.meas_pre_mask = q.pre_mask;
    a.a_result = A_RESULT;
    a.a_oid_err = device::A_OID_ERR;
    for (size_t d = 0; d < q.dims.size(); ++d) {
        const DimQ &dq = q.dims[d];
        if (dq.sels.size() > 1)
            throw std::runtime_error(q.name + ": the request loop runs one selection per dimension");
        const auto &[fk, key] = DIM_KEYS.at(dq.name);
        a.dims[d].fk = desc_of(fk);
        a.dims[d].key = desc_of(key);
        if (!dq.attr.empty())
            a.dims[d].attr = desc_of(dq.attr);
        a.dims[d].flags = dq.flags;
        if (dq.sels.empty())
            continue;
        const auto &[attr, p] = dq.sels[0];
        ahead_serve_sel &x = s.sel[d];
        x.col = desc_of(attr);
        x.op1 = p.op1;
        x.op2 = p.op2;
        x.comb = p.comb;
        x.th1 = p.v1 * x.col.a;
        x.th2 = p.v2 * x.col.a;
        x.a_oid = device::A_OID_LARGEST;
        x.a_oid_inv = device::A_OID_LARGEST_INV;
        s.has_sel[d] = 1;
    }
    for (size_t k = 0; k < q.preds.size(); ++k) {
        const auto &[c, p] = q.preds[k];
        ahead_sq_pred &x = a.preds[k];
        x.col = desc_of(c);
        x.op1 = p.op1;
        x.op2 = p.op2;
        x.comb = p.comb;
        x.th1 = p.v1 * x.col.a;
        x.th2 = p.v2 * x.col.a;
    }
    for (size_t m = 0; m < q.meas.size(); ++m)
        a.meas[m] = desc_of(q.meas[m]);
    return s;
}

void put(std::ofstream &o, const std::string &name, const uint64_t *v, uint64_t n) {
    if (n == 0)
        return;
    o << "err " << name << ".e1";
    for (uint64_t i = 0; i < n; ++i)
        o << ' ' << v[i];
    o << "\n";
}

void write_result(const std::string &path, const Query &q, const ahead_serve_result &r) {
    std::ofstream o(path);
    o << "status " << r.status << "\nrows " << r.kept << "\n";
    const uint64_t inv = ext_euclidean(A_RESULT, 64);
    if (q.family == "q1") {
        o << "result " << (r.groups ? r.sums[0] * inv : 0) << "\n";
    } else {
        o << "groups " << r.groups << "\nsums";
        for (uint64_t g = 0; g < r.groups; ++g)
            o << ' ' << r.sums[g] * inv;
        o << "\n";
        for (int d : q.groups) {
            const ahead_col a = desc_of(q.dims[d].attr);
            const uint64_t mask = a.width >= 8 ? ~0ull : ((1ull << (8 * a.width)) - 1);
            o << "keys";
            for (uint64_t g = 0; g < r.groups; ++g)
                o << ' ' << ((r.keys[d][g] * a.a_inv) & mask);
            o << "\n";
        }
    }
    for (size_t d = 0; d < q.dims.size(); ++d)
        if (!q.dims[d].sels.empty())
            put(o, "select_" + q.dims[d].sels[0].first, r.sel_err[d], r.n_sel_err[d]);
    for (int k = 0; k < AHEAD_SQ_KINDS; ++k)
        put(o, k < static_cast<int>(q.preds.size()) ? "select_" + q.preds[k].first : std::string("fused_") + FUSED_KINDS[k],
            r.err[k], r.n_err[k]);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <dir> <slots>\n", argv[0]);
        return 2;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        std::printf("no GPU\n");
        return 2;
    }
    const std::string dir = argv[1];
    const uint32_t slots = static_cast<uint32_t>(std::stoul(argv[2]));
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
    const std::vector<Query> qs = load_queries(dir + "/queries.txt");
    std::vector<ahead_serve_query> reqs;
    for (const Query &q : qs)
        reqs.push_back(serve_query(q));
    device::raise(hipDeviceSynchronize() == hipSuccess ? AHEAD_OK : AHEAD_ERR_HIP);
    ahead_serve *srv = nullptr;
    device::raise(ahead_serve_create(0, slots, AHEAD_SQ_MAX_SHAPE_PASS, 0, &srv));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t> tickets(qs.size());
    for (size_t i = 0; i < qs.size(); ++i)
        device::raise(ahead_serve_submit(srv, &reqs[i], &tickets[i]));
    int bad = 0;
    double p_max = 0;
    for (size_t i = 0; i < qs.size(); ++i) {
        ahead_serve_result r {};
        device::raise(ahead_serve_wait(srv, tickets[i], &r));
        write_result(dir + "/serve_" + qs[i].name + ".txt", qs[i], r);
        if (r.status != AHEAD_SQ_ST_OK)
            ++bad;
        p_max = std::max(p_max, r.t_finish - r.t_arrival);
        device::raise(ahead_serve_release(srv, tickets[i]));
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    device::raise(ahead_serve_destroy(srv));
    std::printf("%zu queries on %u slot(s): %.1f ms, %.1f queries/s, longest latency %.1f ms; %d with a status\n",
                qs.size(), slots, s * 1e3, qs.size() / s, p_max * 1e3, bad);
    return bad ? 1 : 0;
}
