// serve.hip -- the served form's request loop in C++ over the C ABI: a device-resident star schema,
// star queries submitted from any host thread, and slots (host threads, each with its own HIP
// stream, workspace and device arena) that take the query at the head of the arrival queue
// together with the queued queries of its shape, admit them first come, first served by the
// device bytes they need, run their dimension selections (ahead_hip_select_an, whose OID counts
// stay on the device where the star queries read them) and one ahead_hip_star_query_batch call,
// and hand each query its outputs in host memory after one copy per batch.
//
// No counterpart in the reference: its drivers run one query at a time
// (tools/ssbm/ssbm-q11_continuous.cpp:69-121).
#include "host_common.hpp"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace ahead_host;

namespace {

constexpr uint64_t kGroupsCap = AHEAD_SQ_MAX_GROUPS;  // the fused form's group limit
constexpr uint64_t kErrCap = 256;                      // per error kind, first try
constexpr uint64_t kSelErrCap = 1024;                  // per selection, first try
constexpr size_t kAlign = 256;

struct Req {
    uint64_t ticket = 0;
    ahead_serve_query q{};
    double t_arrival = 0, t_start = 0, t_finish = 0;
    bool done = false;
    int rc = AHEAD_OK;
    std::string error;
    uint64_t counts[4 + AHEAD_SQ_KINDS] = {};
    std::vector<uint64_t> sums, first, keys[AHEAD_SQ_MAX_DIMS], err[AHEAD_SQ_KINDS], sel_err[AHEAD_SQ_MAX_DIMS];
};

// A device allocation that only grows (freeing waits for the device).
struct DevBuf {
    char *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes) return AHEAD_OK;
        if (p != nullptr) AHEAD_HIP_TRY(hipFree(p));
        p = nullptr;
        bytes = 0;
        const size_t b = align_up(need + need / 4, 1 << 20);
        AHEAD_HIP_TRY(hipMalloc(&p, b));
        bytes = b;
        return AHEAD_OK;
    }
    void release() {
        if (p != nullptr) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct HostBuf {  // pinned host memory (the one read-back copy per batch)
    char *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes) return AHEAD_OK;
        if (p != nullptr) AHEAD_HIP_TRY(hipHostFree(p));
        p = nullptr;
        bytes = 0;
        const size_t b = align_up(need + need / 4, 1 << 20);
        AHEAD_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&p), b, hipHostMallocDefault));
        bytes = b;
        return AHEAD_OK;
    }
    void release() {
        if (p != nullptr) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

bool same_col(const ahead_col &x, const ahead_col &y) {
    return x.data == y.data && x.width == y.width && x.a_inv == y.a_inv && x.max_u == y.max_u &&
           x.unenc_bits == y.unenc_bits;
}

// Queries that read the same columns in the same roles (fact length, FK / key / attribute
// columns, join flags, selection columns, fact predicate columns, measures): only their
// constants differ, so the batch entry point runs them in one same-shape pass.
bool same_shape(const ahead_serve_query &x, const ahead_serve_query &y) {
    const ahead_sq_args &a = x.q, &b = y.q;
    if (a.n_fact != b.n_fact || a.n_dims != b.n_dims || a.n_preds != b.n_preds || a.meas_kind != b.meas_kind ||
        a.meas_pre_mask != b.meas_pre_mask)
        return false;
    for (uint32_t d = 0; d < a.n_dims; ++d) {
        const ahead_sq_dim &p = a.dims[d], &r = b.dims[d];
        if (!same_col(p.fk, r.fk) || !same_col(p.key, r.key) || p.attr.data != r.attr.data || p.flags != r.flags ||
            p.key_index != r.key_index || x.has_sel[d] != y.has_sel[d])
            return false;
        if (x.has_sel[d] && !same_col(x.sel[d].col, y.sel[d].col)) return false;
    }
    for (uint32_t k = 0; k < a.n_preds; ++k)
        if (!same_col(a.preds[k].col, b.preds[k].col)) return false;
    const uint32_t nm = a.meas_kind == AHEAD_SQ_SUM ? 1u : 2u;
    for (uint32_t m = 0; m < nm; ++m)
        if (!same_col(a.meas[m], b.meas[m])) return false;
    return true;
}

uint32_t n_attr(const ahead_sq_args &a) {
    uint32_t k = 0;
    for (uint32_t d = 0; d < a.n_dims; ++d) k += a.dims[d].attr.data != nullptr ? 1u : 0u;
    return k;
}

// The read-back area of one query: per selected dimension its counts and error positions, then
// the star query's counts, sums, first rows, attribute codes and error vectors.
struct ReqLayout {
    size_t sel[AHEAD_SQ_MAX_DIMS] = {};
    size_t out = 0, end = 0;
    size_t oids[AHEAD_SQ_MAX_DIMS] = {};  // offsets in the OID area
};

size_t sq_out_bytes(const ahead_sq_args &a, uint64_t err_cap) {
    return align_up(8 * (4 + AHEAD_SQ_KINDS) + 8 * kGroupsCap * (2 + n_attr(a)) + 8 * err_cap * AHEAD_SQ_KINDS, kAlign);
}

}  // namespace

struct ahead_serve {
    int device = 0;
    uint32_t slots = 1, max_batch = 1;
    uint64_t budget = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_adm;
    std::deque<Req *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_ticket = 0;
    uint64_t adm_next = 0, adm_serving = 0, in_use = 0;  // first-come-first-served admission
    bool stop = false;
    std::vector<std::thread> threads;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

struct Slot {
    ahead_serve *S = nullptr;
    hipStream_t stream = nullptr;
    ahead_hip_ws *ws = nullptr;
    DevBuf small, oids;
    HostBuf host;
};

// Runs the batch with the given capacities; fills each request's results, or marks the requests
// whose selection errors or outputs did not fit (trunc[i]) for a retry with larger capacities.
int run_batch(Slot &sl, const std::vector<Req *> &b, uint64_t sel_err_cap, uint64_t err_cap,
              std::vector<char> &trunc) {
    const size_t n = b.size();
    std::vector<ReqLayout> L(n);
    size_t off = 0, ooff = 0;
    for (size_t i = 0; i < n; ++i) {
        const ahead_serve_query &q = b[i]->q;
        for (uint32_t d = 0; d < q.q.n_dims; ++d) {
            if (!q.has_sel[d]) continue;
            L[i].sel[d] = off;
            off += align_up(16 + 8 * sel_err_cap, kAlign);
            L[i].oids[d] = ooff;
            ooff += align_up(8 * (q.q.dims[d].key.n + 1), kAlign);
        }
        L[i].out = off;
        off += sq_out_bytes(q.q, err_cap);
        L[i].end = off;
    }
    int st;
    if ((st = sl.small.ensure(off)) != AHEAD_OK) return st;
    if ((st = sl.oids.ensure(ooff + kAlign)) != AHEAD_OK) return st;
    if ((st = sl.host.ensure(off)) != AHEAD_OK) return st;
    std::vector<ahead_sq_args> args(n);
    std::vector<ahead_sq_out> outs(n);
    std::vector<const ahead_sq_args *> argv(n);
    std::vector<const ahead_sq_out *> outv(n);
    for (size_t i = 0; i < n; ++i) {
        const ahead_serve_query &q = b[i]->q;
        args[i] = q.q;
        for (uint32_t d = 0; d < q.q.n_dims; ++d) {
            if (!q.has_sel[d]) {
                args[i].dims[d].sel = nullptr;
                args[i].dims[d].sel_count = nullptr;
                continue;
            }
            const ahead_serve_sel &s = q.sel[d];
            uint64_t *cnt = reinterpret_cast<uint64_t *>(sl.small.p + L[i].sel[d]);
            uint64_t *oid = reinterpret_cast<uint64_t *>(sl.oids.p + L[i].oids[d]);
            if ((st = ahead_hip_select_an(&s.col, s.op1, s.th1, s.op2, s.th2, s.comb, s.a_oid, q.q.a_oid_err, oid,
                                          q.q.dims[d].key.n + 1, cnt + 2, sel_err_cap, cnt, nullptr, sl.ws,
                                          sl.stream)) != AHEAD_OK)
                return st;
            args[i].dims[d].sel = oid;
            args[i].dims[d].sel_count = cnt;
            args[i].dims[d].sel_a_inv = s.a_oid_inv;
        }
        args[i].ev_scan[0] = args[i].ev_scan[1] = nullptr;
        ahead_sq_out &o = outs[i];
        o = ahead_sq_out{};
        char *p = sl.small.p + L[i].out;
        o.d_counts = reinterpret_cast<uint64_t *>(p);
        p += 8 * (4 + AHEAD_SQ_KINDS);
        o.sums = reinterpret_cast<uint64_t *>(p);
        p += 8 * kGroupsCap;
        o.first_row = reinterpret_cast<uint64_t *>(p);
        p += 8 * kGroupsCap;
        for (uint32_t d = 0; d < q.q.n_dims; ++d) {
            if (q.q.dims[d].attr.data == nullptr) continue;
            o.keys[d] = reinterpret_cast<uint64_t *>(p);
            p += 8 * kGroupsCap;
        }
        o.groups_cap = kGroupsCap;
        for (int k = 0; k < AHEAD_SQ_KINDS; ++k) {
            o.err[k] = reinterpret_cast<uint64_t *>(p);
            o.err_cap[k] = err_cap;
            p += 8 * err_cap;
        }
        argv[i] = &args[i];
        outv[i] = &outs[i];
    }
    if ((st = ahead_hip_star_query_batch(argv.data(), outv.data(), static_cast<uint32_t>(n), sl.ws, sl.stream)) !=
        AHEAD_OK)
        return st;
    AHEAD_HIP_TRY(hipMemcpyAsync(sl.host.p, sl.small.p, off, hipMemcpyDeviceToHost, sl.stream));
    AHEAD_HIP_TRY(hipStreamSynchronize(sl.stream));
    trunc.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        Req &r = *b[i];
        const ahead_serve_query &q = r.q;
        const char *h = sl.host.p;
        for (uint32_t d = 0; d < q.q.n_dims; ++d) {
            if (!q.has_sel[d]) continue;
            const uint64_t *c = reinterpret_cast<const uint64_t *>(h + L[i].sel[d]);
            if (c[1] > sel_err_cap) trunc[i] = 1;
            r.sel_err[d].assign(c + 2, c + 2 + std::min(c[1], sel_err_cap));
        }
        const uint64_t *c = reinterpret_cast<const uint64_t *>(h + L[i].out);
        std::memcpy(r.counts, c, sizeof r.counts);
        for (int k = 0; k < AHEAD_SQ_KINDS; ++k)
            if (c[4 + k] > err_cap) trunc[i] = 1;
        if (c[0] > kGroupsCap) trunc[i] = 1;
        if (trunc[i]) continue;
        const uint64_t ng = c[0];
        const uint64_t *p = c + 4 + AHEAD_SQ_KINDS;
        r.sums.assign(p, p + ng);
        p += kGroupsCap;
        r.first.assign(p, p + ng);
        p += kGroupsCap;
        for (uint32_t d = 0; d < q.q.n_dims; ++d) {
            r.keys[d].clear();
            if (q.q.dims[d].attr.data == nullptr) continue;
            r.keys[d].assign(p, p + ng);
            p += kGroupsCap;
        }
        for (int k = 0; k < AHEAD_SQ_KINDS; ++k) {
            r.err[k].assign(p, p + c[4 + k]);
            p += err_cap;
        }
    }
    return AHEAD_OK;
}

// Bytes the batch holds on the device while it runs: the workspace the batch entry point needs,
// the read-back area and the selections' OID lists.
uint64_t batch_need(const std::vector<Req *> &b) {
    std::vector<const ahead_sq_args *> argv;
    uint64_t need = 0;
    for (Req *r : b) {
        argv.push_back(&r->q.q);
        need += sq_out_bytes(r->q.q, kErrCap);
        for (uint32_t d = 0; d < r->q.q.n_dims; ++d)
            if (r->q.has_sel[d]) need += align_up(16 + 8 * kSelErrCap, kAlign) + 8 * (r->q.q.dims[d].key.n + 1);
    }
    return need + ahead_hip_star_query_batch_ws_bytes(argv.data(), static_cast<uint32_t>(argv.size()));
}

void execute(Slot &sl, std::vector<Req *> &b) {
    std::vector<char> trunc;
    int st = run_batch(sl, b, kSelErrCap, kErrCap, trunc);
    std::string msg = st != AHEAD_OK ? std::string(ahead_hip_last_error()) : std::string();
    if (st == AHEAD_OK) {
        for (size_t i = 0; i < b.size() && st == AHEAD_OK; ++i) {
            if (!trunc[i]) continue;
            // the query alone, with room for every selection error and event it reported
            Req &r = *b[i];
            uint64_t se = kSelErrCap, ee = kErrCap;
            for (uint32_t d = 0; d < r.q.q.n_dims; ++d) se = std::max<uint64_t>(se, r.sel_err[d].size());
            for (int k = 0; k < AHEAD_SQ_KINDS; ++k) ee = std::max<uint64_t>(ee, r.counts[4 + k]);
            std::vector<Req *> one{&r};
            std::vector<char> t2;
            // a selection's true error count is only known from its counts: take the list length
            // bound (every row of the dimension)
            uint64_t sel_rows = 0;
            for (uint32_t d = 0; d < r.q.q.n_dims; ++d)
                if (r.q.has_sel[d]) sel_rows = std::max<uint64_t>(sel_rows, r.q.sel[d].col.n);
            st = run_batch(sl, one, std::max(se, sel_rows), std::max<uint64_t>(ee, AHEAD_SQ_MAX_EVENTS), t2);
            if (st != AHEAD_OK) msg = ahead_hip_last_error();
            else if (t2[0]) st = fail(AHEAD_ERR_DEVICE, "serve: outputs still truncated after a retry");
        }
    }
    std::lock_guard<std::mutex> lk(sl.S->mu);
    const double t = sl.S->now();
    for (Req *r : b) {
        r->rc = st;
        r->error = msg;
        r->t_finish = t;
        r->done = true;
    }
    sl.S->cv_done.notify_all();
}

void slot_main(ahead_serve *S) {
    Slot sl;
    sl.S = S;
    bool ok = hipSetDevice(S->device) == hipSuccess && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
              ahead_hip_ws_create(S->device, &sl.ws) == AHEAD_OK;
    std::string init_error = ok ? std::string() : std::string(ahead_hip_last_error());
    for (;;) {
        std::vector<Req *> batch;
        uint64_t adm = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty()) break;
            Req *head = S->queue.front();
            S->queue.pop_front();
            batch.push_back(head);
            for (auto it = S->queue.begin(); it != S->queue.end() && batch.size() < S->max_batch;) {
                if (same_shape(head->q, (*it)->q)) {
                    batch.push_back(*it);
                    it = S->queue.erase(it);
                } else {
                    ++it;
                }
            }
            adm = S->adm_next++;
        }
        if (!ok) {
            // the batch's admission ticket passes in its turn, so that no earlier ticket is skipped
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_adm.wait(lk, [&] { return adm == S->adm_serving; });
            for (Req *r : batch) {
                r->rc = AHEAD_ERR_HIP;
                r->error = "serve: slot initialisation failed: " + init_error;
                r->done = true;
            }
            S->adm_serving++;
            S->cv_adm.notify_all();
            S->cv_done.notify_all();
            continue;
        }
        const uint64_t need = std::min<uint64_t>(batch_need(batch), S->budget);
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_adm.wait(lk, [&] { return adm == S->adm_serving && S->in_use + need <= S->budget; });
            S->adm_serving++;
            S->in_use += need;
            const double t = S->now();
            for (Req *r : batch) r->t_start = t;
            S->cv_adm.notify_all();
        }
        execute(sl, batch);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->in_use -= need;
            S->cv_adm.notify_all();
        }
    }
    sl.small.release();
    sl.oids.release();
    sl.host.release();
    if (sl.ws != nullptr) (void)ahead_hip_ws_destroy(sl.ws);
    if (sl.stream != nullptr) (void)hipStreamDestroy(sl.stream);
}

}  // namespace

extern "C" int ahead_serve_create(int device, uint32_t slots, uint32_t max_batch, uint64_t hbm_budget,
                                  ahead_serve **out) {
    if (out == nullptr) return fail(AHEAD_ERR_INVALID, "serve: NULL out-pointer");
    if (slots == 0 || slots > 64) return fail(AHEAD_ERR_INVALID, "serve: 1 to 64 slots");
    if (max_batch == 0) return fail(AHEAD_ERR_INVALID, "serve: max_batch must be at least 1");
    int count = 0;
    AHEAD_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(AHEAD_ERR_INVALID, "serve: no such device");
    ahead_serve *S = new (std::nothrow) ahead_serve();
    if (S == nullptr) return fail(AHEAD_ERR_NOMEM, "serve: host allocation failed");
    S->device = device;
    S->slots = slots;
    S->max_batch = max_batch;
    S->budget = hbm_budget == 0 ? ~0ull : hbm_budget;
    S->t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < slots; ++i) S->threads.emplace_back(slot_main, S);
    *out = S;
    return AHEAD_OK;
}

extern "C" int ahead_serve_submit(ahead_serve *S, const ahead_serve_query *q, uint64_t *ticket) {
    if (S == nullptr || q == nullptr || ticket == nullptr) return fail(AHEAD_ERR_INVALID, "serve: NULL argument");
    if (q->q.n_dims < 1 || q->q.n_dims > AHEAD_SQ_MAX_DIMS) return fail(AHEAD_ERR_INVALID, "serve: 1 to 4 dimensions");
    for (uint32_t d = 0; d < q->q.n_dims; ++d) {
        if (!q->has_sel[d]) continue;
        const ahead_col &c = q->sel[d].col;
        if (c.data == nullptr || c.n != q->q.dims[d].key.n)
            return fail(AHEAD_ERR_SIZE_MISMATCH, "serve: a selection column must be a column of its dimension");
    }
    auto r = std::make_unique<Req>();
    r->q = *q;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(AHEAD_ERR_INVALID, "serve: the server is shutting down");
    r->ticket = S->next_ticket++;
    r->t_arrival = S->now();
    *ticket = r->ticket;
    S->queue.push_back(r.get());
    S->reqs.emplace(r->ticket, std::move(r));
    S->cv_work.notify_one();
    return AHEAD_OK;
}

extern "C" int ahead_serve_wait(ahead_serve *S, uint64_t ticket, ahead_serve_result *out) {
    if (S == nullptr || out == nullptr) return fail(AHEAD_ERR_INVALID, "serve: NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(AHEAD_ERR_INVALID, "serve: unknown or released ticket");
    Req &r = *it->second;
    S->cv_done.wait(lk, [&] { return r.done; });
    if (r.rc != AHEAD_OK) return fail(r.rc, r.error);
    *out = ahead_serve_result{};
    out->status = r.counts[2];
    out->groups = r.counts[0];
    out->kept = r.counts[1];
    out->events = r.counts[3];
    out->sums = r.sums.data();
    out->first_row = r.first.data();
    for (uint32_t d = 0; d < AHEAD_SQ_MAX_DIMS; ++d) {
        out->keys[d] = r.keys[d].empty() ? nullptr : r.keys[d].data();
        out->sel_err[d] = r.sel_err[d].data();
        out->n_sel_err[d] = r.sel_err[d].size();
    }
    for (int k = 0; k < AHEAD_SQ_KINDS; ++k) {
        out->err[k] = r.err[k].data();
        out->n_err[k] = r.err[k].size();
    }
    out->t_arrival = r.t_arrival;
    out->t_start = r.t_start;
    out->t_finish = r.t_finish;
    return AHEAD_OK;
}

extern "C" int ahead_serve_release(ahead_serve *S, uint64_t ticket) {
    if (S == nullptr) return fail(AHEAD_ERR_INVALID, "serve: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(AHEAD_ERR_INVALID, "serve: unknown or released ticket");
    if (!it->second->done) return fail(AHEAD_ERR_INVALID, "serve: the query is still running");
    S->reqs.erase(it);
    return AHEAD_OK;
}

extern "C" double ahead_serve_clock(const ahead_serve *S) { return S == nullptr ? 0.0 : S->now(); }

extern "C" int ahead_serve_destroy(ahead_serve *S) {
    if (S == nullptr) return AHEAD_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (std::thread &t : S->threads) t.join();
    delete S;
    return AHEAD_OK;
}
