/*
 * StarQueryAN.hpp -- the fused star query behind the C++ operator API: one pass over the fact
 * rows (ahead_hip_star_query) in place of the fact side of an SSB operator chain
 * (tools/ssbm/ssbm-q11 .. q43_continuous.cpp: the fact selectAN predicates, the per-dimension
 * hashjoinAN / matchjoinAN, the fetchjoinAN of measures and group attributes, groupbyAN and the
 * grouped or multiplied sum), over the same device BATs the chain takes.  The dimension
 * selections stay the chain's own operators (selectAN, matchjoinAN); their OID lists are the
 * fused query's input.
 *
 * No counterpart in the reference's API (its drivers chain the operators).  starQueryOrChainAN
 * runs the chain the caller gives when the fused form does not apply (a corrupt or repeated
 * dimension key, a key span above 4 x rows + 1024, more than 2048 groups or more than 8192 error
 * events): the caller always gets a result.
 */
#pragma once

#include "OperatorsAN.hpp"

#include <optional>
#include <type_traits>
#include <vector>

namespace ahead {
    namespace bat {
        namespace ops {

            /* One dimension of a star query, in the chain's order. */
            struct StarDimAN {
                ahead_col fk {};    /* the fact FK column (8-byte codes)                              */
                ahead_col key {};   /* the dimension's key column (8-byte codes)                      */
                ahead_col attr {};  /* the group attribute column; data == nullptr: no attribute       */
                const uint64_t *sel = nullptr;  /* the selected dimension OIDs (encoded), or nullptr: every row */
                uint64_t n_sel = 0;
                uint64_t sel_a_inv = 0;
                uint32_t flags = 0; /* AHEAD_SQ_PROBE | AHEAD_SQ_REFETCH                              */
            };

            /* A fact predicate with ENCODED thresholds (selectAN's). */
            struct StarPredAN {
                ahead_col col {};
                int op1 = AHEAD_LT, op2 = AHEAD_LT, comb = AHEAD_COMB_NONE;
                uint64_t th1 = 0, th2 = 0;
            };

            /* The fused query's outputs in host memory: sums encoded with A_RESULT, group attribute
             * codes per dimension (empty: no attribute), the error positions by kind. */
            struct StarResultAN {
                uint64_t status = AHEAD_SQ_ST_OK;
                uint64_t rows = 0, groups = 0, events = 0;
                std::vector<uint64_t> sums, first_row;
                std::vector<std::vector<uint64_t>> keys;
                std::vector<std::vector<uint64_t>> err;
            };

            /* The dimension's BATs: fact FK <void, FK>, key <void, Key>, optional attribute
             * <void, Attr>, optional selection (a BAT whose head holds the selected OIDs). */
            template<typename Fk, typename Key>
            StarDimAN starDim(BAT<v2_void_t, Fk> *fk, BAT<v2_void_t, Key> *key, uint32_t flags = 0) {
                StarDimAN d;
                d.fk = device::col(fk->tail, device::rows(fk));
                d.key = device::col(key->tail, device::rows(key));
                d.flags = flags;
                return d;
            }
            template<typename Fk, typename Key, typename Attr>
            StarDimAN starDim(BAT<v2_void_t, Fk> *fk, BAT<v2_void_t, Key> *key, BAT<v2_void_t, Attr> *attr,
                              uint32_t flags = 0) {
                StarDimAN d = starDim(fk, key, flags);
                d.attr = device::col(attr->tail, device::rows(attr));
                return d;
            }
            template<typename SelT>
            void starSelect(StarDimAN &d, BAT<v2_resoid_t, SelT> *sel) {
                d.sel = sel->head.container->data();
                d.n_sel = sel->head.container->size();
                d.sel_a_inv = sel->head.metaData.AN_Ainv;
            }

            /* The fused star query (ahead_hip_star_query) on this thread's operator context; a
             * status other than AHEAD_SQ_ST_OK leaves the outputs empty. */
            inline StarResultAN starQueryAN(uint64_t n_fact, const std::vector<StarDimAN> &dims,
                                            const std::vector<StarPredAN> &preds, const std::vector<ahead_col> &meas,
                                            uint32_t kind, uint32_t pre_mask = 0, uint64_t a_result = 65117,
                                            uint64_t a_oid_err = device::A_OID_ERR) {
                auto &cx = device::Context::get();
                ahead_sq_args a {};
                a.n_fact = n_fact;
                a.n_dims = static_cast<uint32_t>(dims.size());
                a.n_preds = static_cast<uint32_t>(preds.size());
                a.meas_kind = kind;
                a.meas_pre_mask = pre_mask;
                a.a_result = a_result;
                a.a_oid_err = a_oid_err;
                device_vector<uint64_t> sel_counts(dims.size() + 1);
                {
                    std::vector<uint64_t> h(dims.size() + 1, 0);
                    for (size_t d = 0; d < dims.size(); ++d)
                        h[d] = dims[d].n_sel;
                    ahead::hip_check(hipMemcpy(sel_counts.data(), h.data(), h.size() * sizeof(uint64_t), hipMemcpyHostToDevice), "hipMemcpy");
                }
                for (size_t d = 0; d < dims.size() && d < AHEAD_SQ_MAX_DIMS; ++d) {
                    ahead_sq_dim &x = a.dims[d];
                    x.fk = dims[d].fk;
                    x.key = dims[d].key;
                    x.attr = dims[d].attr;
                    x.sel = dims[d].sel;
                    x.sel_count = dims[d].sel != nullptr ? sel_counts.data() + d : nullptr;
                    x.sel_a_inv = dims[d].sel_a_inv;
                    x.flags = dims[d].flags;
                }
                for (size_t p = 0; p < preds.size() && p < AHEAD_SQ_MAX_PREDS; ++p) {
                    ahead_sq_pred &x = a.preds[p];
                    x.col = preds[p].col;
                    x.op1 = preds[p].op1;
                    x.op2 = preds[p].op2;
                    x.comb = preds[p].comb;
                    x.th1 = preds[p].th1;
                    x.th2 = preds[p].th2;
                }
                for (size_t m = 0; m < meas.size() && m < AHEAD_SQ_MAX_MEAS; ++m)
                    a.meas[m] = meas[m];
                uint64_t groups_cap = AHEAD_SQ_MAX_GROUPS, err_cap = 1024;
                for (;;) {
                    device_vector<uint64_t> buf(groups_cap * (2 + dims.size()) + err_cap * AHEAD_SQ_KINDS + 4 + AHEAD_SQ_KINDS);
                    uint64_t *p = buf.data();
                    ahead_sq_out o {};
                    o.d_counts = p;
                    p += 4 + AHEAD_SQ_KINDS;
                    o.sums = p;
                    p += groups_cap;
                    o.first_row = p;
                    p += groups_cap;
                    for (size_t d = 0; d < dims.size(); ++d) {
                        o.keys[d] = dims[d].attr.data != nullptr ? p : nullptr;
                        p += groups_cap;
                    }
                    o.groups_cap = groups_cap;
                    for (int k = 0; k < AHEAD_SQ_KINDS; ++k) {
                        o.err[k] = p;
                        o.err_cap[k] = err_cap;
                        p += err_cap;
                    }
                    device::raise(ahead_hip_star_query(&a, &o, cx.ws, cx.stream));
                    ahead::hip_check(hipStreamSynchronize(cx.stream), "hipStreamSynchronize");
                    std::vector<uint64_t> h(buf.size());
                    ahead::hip_check(hipMemcpy(h.data(), buf.data(), h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
                    const uint64_t *c = h.data();
                    uint64_t need_e = 0;
                    for (int k = 0; k < AHEAD_SQ_KINDS; ++k)
                        need_e = std::max(need_e, c[4 + k]);
                    if (c[0] > groups_cap || need_e > err_cap) {  // outputs truncated: again with room
                        groups_cap = std::max(groups_cap, c[0]);
                        err_cap = std::max(err_cap, need_e);
                        continue;
                    }
                    StarResultAN r;
                    r.status = c[2];
                    if (r.status != AHEAD_SQ_ST_OK)
                        return r;
                    r.groups = c[0];
                    r.rows = c[1];
                    r.events = c[3];
                    const uint64_t *q = c + 4 + AHEAD_SQ_KINDS;
                    r.sums.assign(q, q + r.groups);
                    q += groups_cap;
                    r.first_row.assign(q, q + r.groups);
                    q += groups_cap;
                    r.keys.resize(dims.size());
                    for (size_t d = 0; d < dims.size(); ++d) {
                        if (dims[d].attr.data != nullptr)
                            r.keys[d].assign(q, q + r.groups);
                        q += groups_cap;
                    }
                    r.err.resize(AHEAD_SQ_KINDS);
                    for (int k = 0; k < AHEAD_SQ_KINDS; ++k) {
                        r.err[k].assign(q, q + c[4 + k]);
                        q += err_cap;
                    }
                    return r;
                }
            }

            /* The fused form, or, where it does not apply, the caller's operator chain: `fused`
             * holds the fused outputs (status AHEAD_SQ_ST_OK) or the status that sent the query
             * to `chain`, whose result is then in `chained`. */
            template<typename Chain>
            struct StarOutcomeAN {
                StarResultAN fused;
                std::optional<std::invoke_result_t<Chain>> chained;
            };

            template<typename Chain>
            StarOutcomeAN<Chain> starQueryOrChainAN(uint64_t n_fact, const std::vector<StarDimAN> &dims,
                                                    const std::vector<StarPredAN> &preds, const std::vector<ahead_col> &meas,
                                                    uint32_t kind, uint32_t pre_mask, Chain chain) {
                StarOutcomeAN<Chain> out;
                out.fused = starQueryAN(n_fact, dims, preds, meas, kind, pre_mask);
                if (out.fused.status != AHEAD_SQ_ST_OK)
                    out.chained.emplace(chain());
                return out;
            }

        }  // namespace ops
    }  // namespace bat
}  // namespace ahead
