/*
 * column_operators/OperatorsAN.hpp -- the AN-hardened column operators on device-resident BATs.
 *
 * Same namespaces, names, template parameters, argument lists and default arguments as the
 * reference's include/column_operators/OperatorsAN.hpp:33-298: the join / group / grouped-sum
 * operators in ahead::bat::ops, and encode / check / decode / select / arithmetic / product-sum
 * in both ahead::bat::ops::scalar and ahead::bat::ops::simd::sse (callers choose one of the two
 * with a using-directive, tools/ssbm/ssb.hpp:62-72; here both name the same gfx950 kernels, which
 * follow the scalar semantics).  Each template marshals its BATs into ahead_col descriptors and
 * calls the extern "C" entry point of include/ahead_hip.h on the calling thread's context
 * (workspace + stream), then waits for it, as the reference's synchronous operators do.
 *
 * Ownership is the reference's: every returned BAT and AN_indicator_vector is allocated with
 * new and owned by the caller; an error vector is nullptr for a side whose type is not
 * AN-encoded (ANhelper.tcc:107-109).  Error conditions throw what the reference throws:
 * std::runtime_error for mismatched sizes, std::bad_alloc when device memory runs out.
 */
#ifndef AHEAD_COLUMN_OPERATORS_OPERATORSAN_HPP
#define AHEAD_COLUMN_OPERATORS_OPERATORSAN_HPP

#include <algorithm>
#include <functional>
#include <initializer_list>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>

#include <ahead_hip.h>
#include <column_storage/Storage.hpp>
#include <column_operators/ANbase.hpp>
#include <column_operators/functors.hpp>

namespace ahead {
    namespace bat {
        namespace ops {

            namespace device {

                inline void raise(int rc) {
                    if (rc == AHEAD_OK)
                        return;
                    if (rc == AHEAD_ERR_NOMEM)
                        throw std::bad_alloc();
                    throw std::runtime_error(std::string("ahead_hip: ") + ahead_hip_last_error());
                }

                /* Per-thread device state: a workspace, the stream the operators run on, a
                 * device count array with its pinned host copy, and four reusable device
                 * buffers for error positions.  Deliberately never destroyed: it lives until
                 * the process ends, so no HIP call runs during static destruction. */
                class Context {
                public:
                    ahead_hip_ws *ws = nullptr;
                    hipStream_t stream = nullptr;
                    uint64_t err_cap = 4096;
                    device_vector<uint64_t> counts;
                    device_vector<uint64_t> err[4];
                    uint64_t *h_counts = nullptr;

                    Context() {
                        int dev = 0;
                        hip_check(hipGetDevice(&dev), "hipGetDevice");
                        raise(ahead_hip_ws_create(dev, &ws));
                        counts.overwrite(16);
                        for (auto & e : err)
                            e.overwrite(err_cap);
                        hip_check(hipHostMalloc(reinterpret_cast<void **>(&h_counts), 16 * sizeof(uint64_t), hipHostMallocDefault), "hipHostMalloc");
                    }
                    static Context & get() {
                        thread_local Context *c = new Context();
                        return *c;
                    }
                    uint64_t *d_counts() { return counts.data(); }
                    ahead_buf errbuf(int k) { return ahead_buf {err[k].data(), err_cap}; }
                    /* Waits for the stream and returns the first k counts. */
                    const uint64_t *read_counts(size_t k) {
                        hip_check(hipMemcpyAsync(h_counts, counts.data(), k * sizeof(uint64_t), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
                        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
                        return h_counts;
                    }
                    /* True if every error count fits the error buffers; otherwise grows them
                     * (the caller then runs the operator again: operators are pure). */
                    bool errors_fit(std::initializer_list<uint64_t> cnts) {
                        uint64_t m = 0;
                        for (uint64_t c : cnts)
                            m = std::max(m, c);
                        if (m <= err_cap)
                            return true;
                        err_cap = std::max(m, 2 * err_cap);
                        for (auto & e : err)
                            e.overwrite(err_cap);
                        return false;
                    }
                    /* Copies error buffer k (cnt entries) into a new AN_indicator_vector, or
                     * returns nullptr for a side that is not AN-encoded. */
                    AN_indicator_vector *take(int k, uint64_t cnt, bool encoded) {
                        if (!encoded)
                            return nullptr;
                        auto *v = new AN_indicator_vector(cnt);
                        if (cnt)
                            hip_check(hipMemcpy(v->data(), err[k].data(), cnt * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
                        return v;
                    }
                };

                template<typename V2>
                inline constexpr bool is_void = std::is_same_v<V2, v2_void_t>;

                template<typename V2>
                inline constexpr bool is_enc = !is_void<V2> && V2::is_encoded;

                /* ahead_col descriptor of one BAT side with `n` rows (n only matters for a void side). */
                template<typename V2>
                ahead_col col(const ColumnDescriptor<V2> & cd, uint64_t n) {
                    ahead_col c {};
                    const ColumnMetaData & md = cd.metaData;
                    c.seqbase = md.seqbase;
                    if constexpr (is_void<V2>) {
                        c.data = nullptr;
                        c.n = n;
                        c.width = sizeof(oid_t);
                        c.encoded = 0;
                        c.unenc_bits = 32;
                        c.a = 1;
                        c.a_inv = 1;
                        c.max_u = 0xFFFFFFFFull;
                    } else {
                        const auto & v = *cd.container;
                        c.data = v.data();
                        c.n = v.size();
                        c.width = sizeof(typename V2::type_t);
                        c.encoded = V2::is_encoded ? 1u : 0u;
                        c.unenc_bits = 8u * sizeof(typename TypeMap<V2>::v2_base_t::type_t);
                        c.a = md.AN_A;
                        c.a_inv = md.AN_Ainv;
                        c.max_u = md.AN_unencMaxU;
                    }
                    return c;
                }

                template<typename H, typename T>
                uint64_t rows(BAT<H, T> *b) { return b->size(); }

                template<typename T>
                struct dependent_false : std::false_type {
                };

                /* std::less ... std::not_equal_to -> the C-ABI predicate code. */
                template<template<typename> class Op>
                constexpr int pred_code() {
                    typedef uint64_t U;
                    if constexpr (std::is_same_v<Op<U>, std::less<U>>)
                        return AHEAD_LT;
                    else if constexpr (std::is_same_v<Op<U>, std::less_equal<U>>)
                        return AHEAD_LE;
                    else if constexpr (std::is_same_v<Op<U>, std::greater<U>>)
                        return AHEAD_GT;
                    else if constexpr (std::is_same_v<Op<U>, std::greater_equal<U>>)
                        return AHEAD_GE;
                    else if constexpr (std::is_same_v<Op<U>, std::equal_to<U>>)
                        return AHEAD_EQ;
                    else if constexpr (std::is_same_v<Op<U>, std::not_equal_to<U>>)
                        return AHEAD_NE;
                    else
                        static_assert(dependent_false<Op<U>>::value, "selectAN: unsupported predicate");
                }

                template<template<typename> class C>
                constexpr int comb_code() {
                    if constexpr (std::is_same_v<C<void>, ahead::and_is<void>>)
                        return AHEAD_COMB_AND;
                    else if constexpr (std::is_same_v<C<void>, ahead::or_is<void>>)
                        return AHEAD_COMB_OR;
                    else
                        static_assert(dependent_false<C<void>>::value, "selectAN: unsupported combiner");
                }

                /* std::plus / minus / multiplies / divides -> the arithmeticAN op code. */
                template<template<typename> class Op>
                constexpr int arith_code() {
                    typedef uint64_t U;
                    if constexpr (std::is_same_v<Op<U>, std::plus<U>>)
                        return 0;
                    else if constexpr (std::is_same_v<Op<U>, std::minus<U>>)
                        return 1;
                    else if constexpr (std::is_same_v<Op<U>, std::multiplies<U>>)
                        return 2;
                    else if constexpr (std::is_same_v<Op<U>, std::divides<U>>)
                        return 3;
                    else
                        static_assert(dependent_false<Op<U>>::value, "arithmeticAN: unsupported operator");
                }

                inline constexpr resoid_t A_OID_ERR = (*v2_resoid_t::AsBFW)[v2_resoid_t::AsBFW->size() - 1];
                inline constexpr resoid_t A_OID_LARGEST = (*v2_resoid_t::As)[v2_resoid_t::As->size() - 1];
                inline constexpr resoid_t A_OID_LARGEST_INV = (*v2_resoid_t::Ainvs)[v2_resoid_t::Ainvs->size() - 1];

                inline ColumnMetaData resoid_md(resoid_t a, resoid_t a_inv) {
                    return ColumnMetaData(size_bytes<resoid_t>, static_cast<A_t>(a), a_inv, v2_resoid_t::UNENC_MAX_U, 0);
                }

                /* selectAN on a <void, Tail> BAT (selectAN_scalar.tcc:66-95, 149-181). */
                template<typename Head, typename Tail>
                std::pair<BAT<typename TypeMap<Head>::v2_encoded_t::v2_select_t, typename Tail::v2_select_t> *, AN_indicator_vector *>
                select(BAT<Head, Tail> *arg, int op1, uint64_t th1, int op2, uint64_t th2, int comb, resoid_t AOID) {
                    static_assert(is_void<Head>, "selectAN: the reference instantiates it for <void, T> BATs");
                    static_assert(is_enc<Tail>, "selectAN: Tail must be an AN-encoded type");
                    typedef typename TypeMap<Head>::v2_encoded_t::v2_select_t HSel;
                    typedef typename Tail::v2_select_t TSel;
                    auto & cx = Context::get();
                    const uint64_t n = rows(arg);
                    const ahead_col c = col(arg->tail, n);
                    auto *r = new TempBAT<HSel, TSel>(ColumnDescriptor<HSel>(resoid_md(A_OID_LARGEST, A_OID_LARGEST_INV)), ColumnDescriptor<TSel>(arg->tail.metaData));
                    uint64_t *oids = r->head.container->overwrite(n);
                    const uint64_t *k = nullptr;
                    do {
                        raise(ahead_hip_select_an(&c, op1, th1, op2, th2, comb, A_OID_LARGEST, AOID, oids, n, cx.err[0].data(), cx.err_cap,
                                cx.d_counts(), nullptr, cx.ws, cx.stream));
                        k = cx.read_counts(2);
                    } while (!cx.errors_fit( {k[1]}));
                    const uint64_t n_oids = k[0], n_err = k[1];
                    r->head.container->resize(n_oids);
                    return std::make_pair(r, cx.take(0, n_err, true));
                }

            }  // namespace device

            /* ---- operators on both scalar and simd::sse names (OperatorsAN.hpp:136-298) ---- */
            namespace hip {

                template<typename Head, typename Tail>
                BAT<Head, typename TypeMap<Tail>::v2_encoded_t> *
                encodeAN(BAT<Head, Tail> *arg, typename TypeMap<Tail>::v2_encoded_t::type_t A = std::get<Tail::As->size() - 1>(*Tail::As)) {
                    typedef typename TypeMap<Tail>::v2_encoded_t Enc;
                    typedef typename Enc::type_t enc_t;
                    const uint64_t n = device::rows(arg);
                    const ahead_col c = device::col(arg->tail, n);
                    ColumnMetaData md(size_bytes<enc_t>, static_cast<A_t>(A), ext_euclidean(A, 8 * sizeof(enc_t)), Enc::UNENC_MAX_U, 0);
                    auto *r = new TempBAT<Head, Enc>(arg->head, ColumnDescriptor<Enc>(md));
                    void *out = r->tail.container->overwrite(n);
                    auto & cx = device::Context::get();
                    device::raise(ahead_hip_encode_an(&c, A, sizeof(enc_t), out, cx.stream));
                    hip_check(hipStreamSynchronize(cx.stream), "hipStreamSynchronize");
                    return r;
                }

                template<typename Head, typename ResTail>
                AN_indicator_vector *
                checkAN(BAT<Head, ResTail> *arg, typename ResTail::type_t aInv = ResTail::A_INV, typename ResTail::type_t unEncMaxU = ResTail::A_UNENC_MAX_U,
                        resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    ahead_col c = device::col(arg->tail, device::rows(arg));
                    c.a_inv = aInv;
                    c.max_u = unEncMaxU;
                    auto & cx = device::Context::get();
                    const uint64_t *k = nullptr;
                    do {
                        device::raise(ahead_hip_check_an(&c, AOID, cx.err[0].data(), cx.err_cap, cx.d_counts(), cx.ws, cx.stream));
                        k = cx.read_counts(1);
                    } while (!cx.errors_fit( {k[0]}));
                    return cx.take(0, k[0], true);
                }

                template<typename Head, typename ResTail>
                std::pair<TempBAT<Head, typename ResTail::v2_unenc_t> *, AN_indicator_vector *>
                decodeAN(BAT<Head, ResTail> *arg, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    typedef typename ResTail::v2_unenc_t Unenc;
                    const uint64_t n = device::rows(arg);
                    const ahead_col c = device::col(arg->tail, n);
                    auto *r = new TempBAT<Head, Unenc>(arg->head, ColumnDescriptor<Unenc>(default_metadata<Unenc>()));
                    void *dec = r->tail.container->overwrite(n);
                    auto & cx = device::Context::get();
                    const uint64_t *k = nullptr;
                    do {
                        device::raise(ahead_hip_decode_an(&c, AOID, dec, cx.err[0].data(), cx.err_cap, cx.d_counts(), cx.ws, cx.stream));
                        k = cx.read_counts(2);
                    } while (!cx.errors_fit( {k[1]}));
                    return std::make_pair(r, cx.take(0, k[1], true));
                }

                template<typename Head, typename Tail>
                std::tuple<BAT<typename Head::v2_unenc_t, typename Tail::v2_unenc_t> *, AN_indicator_vector *, AN_indicator_vector *>
                checkAndDecodeAN(BAT<Head, Tail> *arg, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    typedef typename Head::v2_unenc_t UH;
                    typedef typename Tail::v2_unenc_t UT;
                    const uint64_t n = device::rows(arg);
                    const ahead_col c = device::col(arg->tail, n);
                    auto & cx = device::Context::get();
                    const uint64_t *k = nullptr;
                    if constexpr (device::is_void<Head>) {
                        // encdecAN_scalar.tcc:185-210: corrupt rows dropped
                        auto *r = new TempBAT<v2_void_t, UT>(ColumnDescriptor<v2_void_t>(arg->head.metaData), ColumnDescriptor<UT>(default_metadata<UT>()));
                        void *dec = r->tail.container->overwrite(n);
                        do {
                            device::raise(ahead_hip_check_and_decode_an(&c, AOID, dec, n, cx.err[0].data(), cx.err_cap, cx.d_counts(), cx.ws, cx.stream));
                            k = cx.read_counts(2);
                        } while (!cx.errors_fit( {k[1]}));
                        const uint64_t n_dec = k[0], n_err = k[1];
                        r->tail.container->resize(n_dec);
                        return std::make_tuple(r, nullptr, cx.take(0, n_err, device::is_enc<Tail>));
                    } else {
                        // encdecAN_scalar.tcc:129-176: pairs kept only if both sides are valid
                        const ahead_col h = device::col(arg->head, n);
                        auto *r = new TempBAT<UH, UT>(ColumnDescriptor<UH>(default_metadata<UH>()), ColumnDescriptor<UT>(default_metadata<UT>()));
                        void *dh = r->head.container->overwrite(n);
                        void *dt = r->tail.container->overwrite(n);
                        do {
                            device::raise(ahead_hip_check_and_decode2_an(&h, &c, AOID, sizeof(typename UH::type_t), sizeof(typename UT::type_t), dh, dt, n,
                                    cx.errbuf(0), cx.errbuf(1), cx.d_counts(), cx.ws, cx.stream));
                            k = cx.read_counts(3);
                        } while (!cx.errors_fit( {k[1], k[2]}));
                        const uint64_t np = k[0], e1 = k[1], e2 = k[2];
                        r->head.container->resize(np);
                        r->tail.container->resize(np);
                        return std::make_tuple(r, cx.take(0, e1, device::is_enc<Head>), cx.take(1, e2, device::is_enc<Tail>));
                    }
                }

                template<template<typename > class Op, typename Head, typename Tail>
                std::pair<BAT<typename TypeMap<Head>::v2_encoded_t::v2_select_t, typename Tail::v2_select_t> *, AN_indicator_vector *>
                selectAN(BAT<Head, Tail> *arg, typename Tail::type_t threshold, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    return device::select(arg, device::pred_code<Op>(), threshold, device::pred_code<Op>(), threshold, AHEAD_COMB_NONE, AOID);
                }

                template<template<typename > class Op1, template<typename > class Op2, template<typename > class OpCombine, typename Head, typename Tail>
                std::pair<BAT<typename TypeMap<Head>::v2_encoded_t::v2_select_t, typename Tail::v2_select_t> *, AN_indicator_vector *>
                selectAN(BAT<Head, Tail> *arg, typename Tail::type_t threshold1, typename Tail::type_t threshold2,
                        resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    return device::select(arg, device::pred_code<Op1>(), threshold1, device::pred_code<Op2>(), threshold2, device::comb_code<OpCombine>(), AOID);
                }

                /* The re-encoding overloads accept and ignore ATR / ATInvR, as the reference's
                 * dispatch does (selectAN.tcc:51-70). */
                template<template<typename > class Op, typename Head, typename Tail>
                std::pair<BAT<typename TypeMap<Head>::v2_encoded_t::v2_select_t, typename Tail::v2_select_t> *, AN_indicator_vector *>
                selectAN(BAT<Head, Tail> *arg, typename Tail::type_t threshold, typename Tail::v2_select_t::type_t ATR, typename Tail::v2_select_t::type_t ATInvR,
                        resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    (void) ATR;
                    (void) ATInvR;
                    return selectAN<Op>(arg, threshold, AOID);
                }

                template<template<typename > class Op1, template<typename > class Op2, template<typename > class OpCombine, typename Head, typename Tail>
                std::pair<BAT<typename TypeMap<Head>::v2_encoded_t::v2_select_t, typename Tail::v2_select_t> *, AN_indicator_vector *>
                selectAN(BAT<Head, Tail> *arg, typename Tail::type_t threshold1, typename Tail::type_t threshold2, typename Tail::v2_select_t::type_t ATR,
                        typename Tail::v2_select_t::type_t ATInvR, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    (void) ATR;
                    (void) ATInvR;
                    return selectAN<Op1, Op2, OpCombine>(arg, threshold1, threshold2, AOID);
                }

                template<template<typename > class Op, typename Result, typename Head1, typename Tail1, typename Head2, typename Tail2>
                std::tuple<BAT<v2_void_t, Result> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
                arithmeticAN(BAT<Head1, Tail1> *bat1, BAT<Head2, Tail2> *bat2, typename Result::type_t AResult = std::get<Result::As->size() - 1>(*Result::As),
                        typename Result::type_t AResultInv = std::get<Result::Ainvs->size() - 1>(*Result::Ainvs),
                        resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    typedef typename Result::type_t result_t;
                    const uint64_t n1 = device::rows(bat1), n2 = device::rows(bat2);
                    if (n1 != n2)
                        throw std::runtime_error("arithmeticAN: bat1->size() != bat2->size()");
                    const ahead_col h1 = device::col(bat1->head, n1), t1 = device::col(bat1->tail, n1);
                    const ahead_col h2 = device::col(bat2->head, n2), t2 = device::col(bat2->tail, n2);
                    ColumnMetaData md(size_bytes<result_t>, static_cast<A_t>(AResult), AResultInv, Result::UNENC_MAX_U, 0);
                    auto *r = new TempBAT<v2_void_t, Result>(ColumnDescriptor<v2_void_t>(oid_t(0)), ColumnDescriptor<Result>(md));
                    void *out = r->tail.container->overwrite(n1);
                    auto & cx = device::Context::get();
                    const uint64_t *k = nullptr;
                    do {
                        device::raise(ahead_hip_arithmetic_an(&h1, &t1, &h2, &t2, device::arith_code<Op>(), AResult, sizeof(result_t), AOID, out, cx.errbuf(0),
                                cx.errbuf(1), cx.errbuf(2), cx.errbuf(3), cx.d_counts(), cx.ws, cx.stream));
                        k = cx.read_counts(5);
                    } while (!cx.errors_fit( {k[1], k[2], k[3], k[4]}));
                    const uint64_t e1 = k[1], e2 = k[2], e3 = k[3], e4 = k[4];
                    return std::make_tuple(r, cx.take(0, e1, device::is_enc<Head1>), cx.take(1, e2, device::is_enc<Tail1>), cx.take(2, e3, device::is_enc<Head2>),
                            cx.take(3, e4, device::is_enc<Tail2>));
                }

                template<typename Result, typename Head1, typename Tail1, typename Head2, typename Tail2, typename ResEnc = typename TypeMap<Result>::v2_encoded_t,
                        typename T1Enc = typename TypeMap<Tail1>::v2_encoded_t, typename T2Enc = typename TypeMap<Tail2>::v2_encoded_t>
                std::tuple<BAT<v2_void_t, Result> *, AN_indicator_vector *, AN_indicator_vector *>
                aggregate_mul_sumAN(BAT<Head1, Tail1> *arg1, BAT<Head2, Tail2> *arg2, typename Result::type_t init = typename Result::type_t(0),
                        typename ResEnc::type_t RA = std::get<ResEnc::As->size() - 1>(*ResEnc::As),
                        typename ResEnc::type_t RAInv = std::get<ResEnc::Ainvs->size() - 1>(*ResEnc::Ainvs),
                        resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                    static_assert(sizeof(typename Result::type_t) == 8, "aggregate_mul_sumAN: 64-bit results on the device path");
                    const ahead_col t1 = device::col(arg1->tail, device::rows(arg1));
                    const ahead_col t2 = device::col(arg2->tail, device::rows(arg2));
                    ColumnMetaData md(size_bytes<typename Result::type_t>, static_cast<A_t>(RA), RAInv, Result::UNENC_MAX_U, 0);
                    auto *r = new TempBAT<v2_void_t, Result>(ColumnDescriptor<v2_void_t>(oid_t(0)), ColumnDescriptor<Result>(md));
                    uint64_t *total = reinterpret_cast<uint64_t *>(r->tail.container->overwrite(1));
                    auto & cx = device::Context::get();
                    const uint64_t *k = nullptr;
                    do {
                        device::raise(ahead_hip_aggregate_mul_sum_an(&t1, &t2, init, device::is_enc<Result> ? 1 : 0, RA, AOID, total, cx.errbuf(0), cx.errbuf(1),
                                cx.d_counts(), cx.ws, cx.stream));
                        k = cx.read_counts(3);
                    } while (!cx.errors_fit( {k[1], k[2]}));
                    const uint64_t e1 = k[1], e2 = k[2];
                    return std::make_tuple(r, cx.take(0, e1, device::is_enc<Tail1>), cx.take(1, e2, device::is_enc<Tail2>));
                }

            }  // namespace hip

            namespace scalar {
                using hip::encodeAN;
                using hip::checkAN;
                using hip::decodeAN;
                using hip::checkAndDecodeAN;
                using hip::selectAN;
                using hip::arithmeticAN;
                using hip::aggregate_mul_sumAN;
            }
            namespace simd {
                namespace sse {
                    using hip::encodeAN;
                    using hip::checkAN;
                    using hip::decodeAN;
                    using hip::checkAndDecodeAN;
                    using hip::selectAN;
                    using hip::arithmeticAN;
                    using hip::aggregate_mul_sumAN;
                }
            }

            /* ---- joins (OperatorsAN.hpp:33-106) -------------------------------------------- */
            namespace device {

                /* matchjoinAN: the positional form <H1, v2_resoid_t, v2_void_t, T2>
                 * (matchjoinAN.tcc:240-301) and the merge form on two ascending lists
                 * (matchjoinAN.tcc:46-143). */
                template<typename RH, typename RT, bool reencode, typename H1, typename T1, typename H2, typename T2>
                std::tuple<BAT<RH, RT> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
                matchjoin(BAT<H1, T1> *arg1, BAT<H2, T2> *arg2, resoid_t AOID, uint64_t AH1R, uint64_t AH1InvR, uint64_t AT2R, uint64_t AT2InvR) {
                    typedef typename H1::type_t h1_t;
                    typedef typename T2::type_t t2_t;
                    const uint64_t n1 = rows(arg1), n2 = rows(arg2);
                    const ahead_col h1 = col(arg1->head, n1), t1 = col(arg1->tail, n1);
                    const ahead_col h2 = col(arg2->head, n2), t2 = col(arg2->tail, n2);
                    ColumnMetaData mh = arg1->head.metaData, mt = arg2->tail.metaData;
                    if (reencode) {
                        mh = ColumnMetaData(arg1->head.metaData.width, static_cast<A_t>(AH1R), AH1InvR, arg1->head.metaData.AN_unencMaxU, arg1->head.metaData.AN_unencMinS);
                        mt = ColumnMetaData(arg2->tail.metaData.width, static_cast<A_t>(AT2R), AT2InvR, arg2->tail.metaData.AN_unencMaxU, arg2->tail.metaData.AN_unencMinS);
                    }
                    // reencFactor = mulIfEncoded(A_R, A^-1) in the side's own width; an unencoded side is copied
                    const uint64_t f_h1 = static_cast<h1_t>(AH1R * static_cast<h1_t>(arg1->head.metaData.AN_Ainv));
                    const uint64_t f_t2 = static_cast<t2_t>(AT2R * static_cast<t2_t>(arg2->tail.metaData.AN_Ainv));
                    const int reenc_h = reencode && is_enc<H1>, reenc_t = reencode && is_enc<T2>;
                    auto *r = new TempBAT<RH, RT>(ColumnDescriptor<RH>(mh), ColumnDescriptor<RT>(mt));
                    auto & cx = Context::get();
                    if constexpr (std::is_same_v<T1, v2_resoid_t> && is_void<H2>) {
                        void *oh = r->head.container->overwrite(n1);
                        void *ot = r->tail.container->overwrite(n1);
                        const uint64_t *k = nullptr;
                        do {
                            raise(ahead_hip_matchjoin_pos_an(&h1, &t1, &t2, reenc_h, f_h1, sizeof(typename RH::type_t), reenc_t, f_t2, sizeof(typename RT::type_t), AOID,
                                    ahead_buf {oh, n1}, ahead_buf {ot, n1}, cx.errbuf(0), cx.errbuf(1), cx.errbuf(3), cx.d_counts(), cx.ws, cx.stream));
                            k = cx.read_counts(4);
                        } while (!cx.errors_fit( {k[1], k[2], k[3]}));
                        const uint64_t np = k[0], e1 = k[1], e2 = k[2], e4 = k[3];
                        r->head.container->resize(np);
                        r->tail.container->resize(np);
                        return std::make_tuple(r, cx.take(0, e1, is_enc<H1>), cx.take(1, e2, true), nullptr, cx.take(3, e4, is_enc<T2>));
                    } else if constexpr (!(is_void<T1> && is_void<H2>)) {
                        // the merge form; a void side is its virtual ascending OID sequence (the
                        // drivers' matchjoinAN(reverse(key), mirror_head(selection)),
                        // ssbm-q21_continuous.cpp:92)
                        const uint64_t cap = std::min(n1, n2);
                        void *oh = r->head.container->overwrite(cap);
                        void *ot = r->tail.container->overwrite(cap);
                        const uint64_t *k = nullptr;
                        do {
                            raise(ahead_hip_matchjoin_merge_an(&h1, &t1, &h2, &t2, reenc_h, f_h1, sizeof(typename RH::type_t), reenc_t, f_t2, sizeof(typename RT::type_t),
                                    AOID, ahead_buf {oh, cap}, ahead_buf {ot, cap}, cx.errbuf(0), cx.errbuf(1), cx.errbuf(2), cx.errbuf(3), cx.d_counts(), cx.ws, cx.stream));
                            k = cx.read_counts(5);
                        } while (!cx.errors_fit( {k[1], k[2], k[3], k[4]}));
                        const uint64_t np = k[0], e1 = k[1], e2 = k[2], e3 = k[3], e4 = k[4];
                        r->head.container->resize(np);
                        r->tail.container->resize(np);
                        return std::make_tuple(r, cx.take(0, e1, is_enc<H1>), cx.take(1, e2, is_enc<T1>), cx.take(2, e3, is_enc<H2>), cx.take(3, e4, is_enc<T2>));
                    } else {
                        // <H1, void, void, T2>: virtual OIDs aligned by seqbase (matchjoinAN.tcc:146-224)
                        const uint64_t cap = std::max<uint64_t>(std::min(n1, n2), 1);
                        void *oh = r->head.container->overwrite(cap);
                        void *ot = r->tail.container->overwrite(cap);
                        const uint64_t *k = nullptr;
                        do {
                            raise(ahead_hip_matchjoin_seq_an(&h1, &t1, &h2, &t2, reenc_h, f_h1, sizeof(typename RH::type_t), reenc_t, f_t2, sizeof(typename RT::type_t),
                                    AOID, ahead_buf {oh, cap}, ahead_buf {ot, cap}, cx.errbuf(0), cx.errbuf(3), cx.d_counts(), cx.ws, cx.stream));
                            k = cx.read_counts(3);
                        } while (!cx.errors_fit( {k[1], k[2]}));
                        const uint64_t np = k[0], e1 = k[1], e4 = k[2];
                        r->head.container->resize(np);
                        r->tail.container->resize(np);
                        return std::make_tuple(r, cx.take(0, e1, is_enc<H1>), nullptr, nullptr, cx.take(3, e4, is_enc<T2>));
                    }
                }

                template<typename RH, typename RT, bool reencH, bool reencT, typename H1, typename T1, typename H2, typename T2>
                std::tuple<TempBAT<RH, RT> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
                hashjoin(BAT<H1, T1> *arg1, BAT<H2, T2> *arg2, hash_side_t side, resoid_t AOID, uint64_t AH1R, uint64_t AH1InvR, uint64_t AT2R, uint64_t AT2InvR) {
                    typedef typename TypeMap<H1>::v2_encoded_t::type_t h1enc_t;
                    typedef typename TypeMap<T2>::v2_encoded_t::type_t t2enc_t;
                    const uint64_t n1 = rows(arg1), n2 = rows(arg2);
                    const ahead_col h1 = col(arg1->head, n1), t1 = col(arg1->tail, n1);
                    const ahead_col h2 = col(arg2->head, n2), t2 = col(arg2->tail, n2);
                    const auto & hmd = arg1->head.metaData;
                    const auto & tmd = arg2->tail.metaData;
                    // a re-encoded side takes its result type's metadata: a void side becomes OIDs
                    // encoded with A_R (resoid width and maxU), not the void column's
                    const auto reenc_md = [](const ColumnMetaData &md, bool enc, uint64_t a, uint64_t ainv, const bytes_t &width, uint64_t maxu) {
                        return enc ? ColumnMetaData(md.width, static_cast<A_t>(a), ainv, md.AN_unencMaxU, md.AN_unencMinS)
                                   : ColumnMetaData(width, static_cast<A_t>(a), ainv, maxu, 0);
                    };
                    ColumnMetaData mh = hmd, mt = tmd;
                    if constexpr (reencH)
                        mh = reenc_md(hmd, is_enc<H1>, AH1R, AH1InvR, size_bytes<typename RH::type_t>, RH::UNENC_MAX_U);
                    if constexpr (reencT)
                        mt = reenc_md(tmd, is_enc<T2>, AT2R, AT2InvR, size_bytes<typename RT::type_t>, RT::UNENC_MAX_U);
                    // hashjoinAN.tcc:94-95: reencFactor = mulIfEncoded(A_R, A^-1), applied to every output value
                    const uint64_t f_h1 = is_enc<H1> ? static_cast<h1enc_t>(AH1R * static_cast<h1enc_t>(hmd.AN_Ainv)) : static_cast<h1enc_t>(AH1R);
                    const uint64_t f_t2 = is_enc<T2> ? static_cast<t2enc_t>(AT2R * static_cast<t2enc_t>(tmd.AN_Ainv)) : static_cast<t2enc_t>(AT2R);
                    auto *r = new TempBAT<RH, RT>(ColumnDescriptor<RH>(mh), ColumnDescriptor<RT>(mt));
                    auto & cx = Context::get();
                    uint64_t cap = std::max<uint64_t>(std::max(n1, n2), 1);
                    const uint64_t *k = nullptr;
                    for (;;) {
                        void *oh = r->head.container->overwrite(cap);
                        void *ot = r->tail.container->overwrite(cap);
                        raise(ahead_hip_hashjoin_an(&h1, &t1, &h2, &t2, side == hash_side_t::right ? 1 : 0, reencH ? 1 : 0, f_h1, sizeof(typename RH::type_t),
                                reencT ? 1 : 0, f_t2, sizeof(typename RT::type_t), AOID, ahead_buf {oh, cap}, ahead_buf {ot, cap}, cx.errbuf(0), cx.errbuf(1),
                                cx.errbuf(2), cx.errbuf(3), cx.d_counts(), cx.ws, cx.stream));
                        k = cx.read_counts(6);
                        if (k[5] != 0)
                            throw std::runtime_error("hashjoinAN: the build table overflowed (internal limit)");
                        const bool err_ok = cx.errors_fit( {k[1], k[2], k[3], k[4]});
                        if (k[0] > cap) {
                            cap = k[0];
                            continue;
                        }
                        if (err_ok)
                            break;
                    }
                    const uint64_t np = k[0], e1 = k[1], e2 = k[2], e3 = k[3], e4 = k[4];
                    r->head.container->resize(np);
                    r->tail.container->resize(np);
                    return std::make_tuple(r, cx.take(0, e1, is_enc<H1>), cx.take(1, e2, is_enc<T1>), cx.take(2, e3, is_enc<H2>), cx.take(3, e4, is_enc<T2>));
                }

                template<typename Ret, bool reencode, typename T2>
                std::tuple<BAT<v2_void_t, Ret> *, AN_indicator_vector *, AN_indicator_vector *>
                fetchjoin(BAT<v2_void_t, v2_resoid_t> *arg1, BAT<v2_void_t, T2> *arg2, resoid_t AOID, uint64_t ATReenc, uint64_t ATReencInv) {
                    typedef typename TypeMap<T2>::v2_encoded_t::type_t tailenc_t;
                    const uint64_t n1 = rows(arg1), n2 = rows(arg2);
                    const ahead_col o = col(arg1->tail, n1), c = col(arg2->tail, n2);
                    ColumnMetaData mt = arg2->tail.metaData;
                    if (reencode) {
                        mt.AN_A = static_cast<A_t>(ATReenc);
                        mt.AN_Ainv = ATReencInv;
                    }
                    auto *r = new TempBAT<v2_void_t, Ret>(ColumnDescriptor<v2_void_t>(arg1->head.metaData), ColumnDescriptor<Ret>(mt));
                    void *out = r->tail.container->overwrite(n1);
                    auto & cx = Context::get();
                    const uint64_t *k = nullptr;
                    do {
                        raise(ahead_hip_fetchjoin_an(&o, &c, reencode ? 1 : 0, static_cast<tailenc_t>(ATReenc), AOID, ahead_buf {out, n1}, cx.errbuf(0), cx.errbuf(1),
                                cx.d_counts(), cx.ws, cx.stream));
                        k = cx.read_counts(3);
                    } while (!cx.errors_fit( {k[1], k[2]}));
                    const uint64_t nout = k[0], e1 = k[1], e2 = k[2];
                    r->tail.container->resize(nout);
                    return std::make_tuple(r, cx.take(0, e1, true), cx.take(1, e2, is_enc<T2>));
                }

            }  // namespace device

            template<typename H1, typename T1, typename H2, typename T2>
            std::tuple<BAT<typename H1::v2_select_t, typename T2::v2_select_t> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            matchjoinAN(BAT<H1, T1> *arg1, BAT<H2, T2> *arg2, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::matchjoin<typename H1::v2_select_t, typename T2::v2_select_t, false>(arg1, arg2, AOID, 0, 0, 0, 0);
            }

            /* Head-only re-encoding: the reference passes A_T2R = 0 to the shared implementation
             * (matchjoinAN.tcc:384-392), so an encoded right tail comes out as 0; kept as is. */
            template<typename H1, typename T1, typename H2, typename T2>
            std::tuple<BAT<typename H1::v2_select_t, typename T2::v2_select_t> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            matchjoinAN(BAT<H1, T1> *arg1, BAT<H2, T2> *arg2, typename TypeMap<H1>::v2_encoded_t::type_t AH1reenc, typename TypeMap<H1>::v2_encoded_t::type_t AH1InvReenc,
                    resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::matchjoin<typename H1::v2_select_t, typename T2::v2_select_t, true>(arg1, arg2, AOID, AH1reenc, AH1InvReenc, 0, 0);
            }

            template<typename H1, typename T1, typename H2, typename T2>
            std::tuple<BAT<typename H1::v2_select_t, typename T2::v2_select_t> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            matchjoinAN(BAT<H1, T1> *arg1, BAT<H2, T2> *arg2, typename TypeMap<H1>::v2_encoded_t::type_t AH1reenc, typename TypeMap<H1>::v2_encoded_t::type_t AH1InvReenc,
                    typename TypeMap<T2>::v2_encoded_t::type_t AT2Reenc, typename TypeMap<T2>::v2_encoded_t::type_t AT2InvReenc,
                    resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::matchjoin<typename H1::v2_select_t, typename T2::v2_select_t, true>(arg1, arg2, AOID, AH1reenc, AH1InvReenc, AT2Reenc, AT2InvReenc);
            }

            template<typename Head1, typename Tail1, typename Head2, typename Tail2>
            std::tuple<TempBAT<typename Head1::v2_select_t, typename Tail2::v2_select_t> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            hashjoinAN(BAT<Head1, Tail1> *arg1, BAT<Head2, Tail2> *arg2, hash_side_t hashside = hash_side_t::right,
                    resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::hashjoin<typename Head1::v2_select_t, typename Tail2::v2_select_t, false, false>(arg1, arg2, hashside, AOID, 1, 1, 1, 1);
            }

            template<typename Head1, typename Tail1, typename Head2, typename Tail2>
            std::tuple<TempBAT<typename TypeMap<Head1>::v2_encoded_t::v2_select_t, typename Tail2::v2_select_t> *, AN_indicator_vector *, AN_indicator_vector *,
                    AN_indicator_vector *, AN_indicator_vector *>
            hashjoinAN(BAT<Head1, Tail1> *arg1, BAT<Head2, Tail2> *arg2, typename TypeMap<Head1>::v2_encoded_t::type_t AH1reenc,
                    typename TypeMap<Head1>::v2_encoded_t::type_t AH1InvReenc, hash_side_t hashside = hash_side_t::right,
                    resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::hashjoin<typename TypeMap<Head1>::v2_encoded_t::v2_select_t, typename Tail2::v2_select_t, true, false>(arg1, arg2, hashside, AOID,
                        AH1reenc, AH1InvReenc, 1, 1);
            }

            template<typename Head1, typename Tail1, typename Head2, typename Tail2>
            std::tuple<TempBAT<typename TypeMap<Head1>::v2_encoded_t::v2_select_t, typename TypeMap<Tail2>::v2_encoded_t::v2_select_t> *, AN_indicator_vector *,
                    AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            hashjoinAN(BAT<Head1, Tail1> *arg1, BAT<Head2, Tail2> *arg2, typename TypeMap<Head1>::v2_encoded_t::type_t AH1Reenc,
                    typename TypeMap<Head1>::v2_encoded_t::type_t AH1InvReenc, typename TypeMap<Tail2>::v2_encoded_t::type_t AT2Reenc,
                    typename TypeMap<Tail2>::v2_encoded_t::type_t AT2InvReenc, hash_side_t hashside = hash_side_t::right,
                    resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::hashjoin<typename TypeMap<Head1>::v2_encoded_t::v2_select_t, typename TypeMap<Tail2>::v2_encoded_t::v2_select_t, true, true>(arg1, arg2,
                        hashside, AOID, AH1Reenc, AH1InvReenc, AT2Reenc, AT2InvReenc);
            }

            template<typename T2>
            std::tuple<BAT<v2_void_t, typename TypeMap<T2>::v2_encoded_t> *, AN_indicator_vector *, AN_indicator_vector *>
            fetchjoinAN(BAT<v2_void_t, v2_resoid_t> *arg1, BAT<v2_void_t, T2> *arg2, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::fetchjoin<typename TypeMap<T2>::v2_encoded_t, false>(arg1, arg2, AOID, 1, 1);
            }

            template<typename T2>
            std::tuple<BAT<v2_void_t, typename TypeMap<T2>::v2_encoded_t> *, AN_indicator_vector *, AN_indicator_vector *>
            fetchjoinAN(BAT<v2_void_t, v2_resoid_t> *arg1, BAT<v2_void_t, T2> *arg2, typename TypeMap<T2>::v2_encoded_t::type_t ATReenc,
                    typename TypeMap<T2>::v2_encoded_t::type_t ATReencInv, resoid_t AOID = std::get<v2_resoid_t::AsBFW->size() - 1>(*v2_resoid_t::AsBFW)) {
                return device::fetchjoin<typename TypeMap<T2>::v2_encoded_t, true>(arg1, arg2, AOID, ATReenc, ATReencInv);
            }

            /* ---- grouping and grouped sum (OperatorsAN.hpp:108-134) ----------------------- */
            namespace device {

                /* Runs a group-by entry point, enlarging the dictionary while it reports overflow
                 * (status 1) and the error buffers while they are too small. */
                template<typename Launch>
                const uint64_t *run_groupby(Launch && launch, std::initializer_list<int> err_idx) {
                    auto & cx = Context::get();
                    uint32_t bits = 0;
                    for (;;) {
                        launch(bits);
                        const uint64_t *k = cx.read_counts(6);
                        if (k[5] == 1) {
                            bits = bits ? bits + 2 : 24;
                            if (bits > 30)
                                throw std::runtime_error("groupbyAN: too many groups for the device dictionary");
                            continue;
                        }
                        uint64_t m = 0;
                        for (int i : err_idx)
                            m = std::max(m, k[i]);
                        if (cx.errors_fit( {m}))
                            return k;
                    }
                }

                inline std::pair<TempBAT<v2_void_t, v2_resoid_t> *, TempBAT<v2_void_t, v2_resoid_t> *> group_outputs(uint64_t n, resoid_t AOID, resoid_t AOIDinv) {
                    auto *v2r = new TempBAT<v2_void_t, v2_resoid_t>(ColumnDescriptor<v2_void_t>(oid_t(0)), ColumnDescriptor<v2_resoid_t>(resoid_md(AOID, AOIDinv)));
                    auto *g2r = new TempBAT<v2_void_t, v2_resoid_t>(ColumnDescriptor<v2_void_t>(oid_t(0)), ColumnDescriptor<v2_resoid_t>(resoid_md(AOID, AOIDinv)));
                    v2r->tail.container->overwrite(n);
                    g2r->tail.container->overwrite(std::max<uint64_t>(n, 1));
                    return {v2r, g2r};
                }

            }  // namespace device

            template<typename Head, typename Tail>
            std::tuple<BAT<v2_void_t, v2_resoid_t> *, BAT<v2_void_t, v2_resoid_t> *, AN_indicator_vector *, AN_indicator_vector *>
            groupbyAN(BAT<Head, Tail> *bat, typename v2_resoid_t::type_t AOID = std::get<v2_resoid_t::As->size() - 1>(*v2_resoid_t::As),
                    typename v2_resoid_t::type_t AOIDinv = std::get<v2_resoid_t::Ainvs->size() - 1>(*v2_resoid_t::Ainvs)) {
                const uint64_t n = device::rows(bat);
                const ahead_col h = device::col(bat->head, n), t = device::col(bat->tail, n);
                auto outs = device::group_outputs(n, AOID, AOIDinv);
                auto & cx = device::Context::get();
                const uint64_t *k = device::run_groupby([&](uint32_t bits) {
                    device::raise(ahead_hip_groupby_an(&h, &t, AOID, bits, outs.first->tail.container->data(),
                            ahead_buf {outs.second->tail.container->data(), std::max<uint64_t>(n, 1)}, cx.errbuf(0), cx.errbuf(1), cx.d_counts(), cx.ws, cx.stream));
                }, {2, 3});
                const uint64_t ng = k[1], e1 = k[2], e2 = k[3];
                outs.second->tail.container->resize(ng);
                return std::make_tuple(outs.first, outs.second, cx.take(0, e1, device::is_enc<Head>), cx.take(1, e2, device::is_enc<Tail>));
            }

            template<typename Head, typename Tail>
            std::tuple<BAT<v2_void_t, v2_resoid_t> *, BAT<v2_void_t, v2_resoid_t> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            groupbyAN(BAT<Head, Tail> *bat, BAT<v2_void_t, v2_resoid_t> *grouping, size_t numGroups,
                    typename v2_resoid_t::type_t AOID = std::get<v2_resoid_t::As->size() - 1>(*v2_resoid_t::As),
                    typename v2_resoid_t::type_t AOIDinv = std::get<v2_resoid_t::Ainvs->size() - 1>(*v2_resoid_t::Ainvs)) {
                const uint64_t n = device::rows(bat);
                if (n != device::rows(grouping))
                    throw std::runtime_error("groupbyAN: bat->size() != grouping->size()");
                const ahead_col h = device::col(bat->head, n), t = device::col(bat->tail, n), g = device::col(grouping->tail, n);
                const uint32_t larger_bits = 8u * sizeof(typename Tail::v2_larger_t::type_t);
                auto outs = device::group_outputs(n, AOID, AOIDinv);
                auto & cx = device::Context::get();
                const uint64_t *k = device::run_groupby([&](uint32_t bits) {
                    device::raise(ahead_hip_groupby2_an(&h, &t, &g, numGroups, larger_bits, AOID, bits, outs.first->tail.container->data(),
                            ahead_buf {outs.second->tail.container->data(), std::max<uint64_t>(n, 1)}, cx.errbuf(0), cx.errbuf(1), cx.errbuf(2), cx.d_counts(), cx.ws,
                            cx.stream));
                }, {2, 3, 4});
                const uint64_t ng = k[1], e1 = k[2], e2 = k[3], e3 = k[4];
                outs.second->tail.container->resize(ng);
                return std::make_tuple(outs.first, outs.second, cx.take(0, e1, device::is_enc<Head>), cx.take(1, e2, device::is_enc<Tail>), cx.take(2, e3, true));
            }

            template<typename Result, typename Head, typename Tail>
            std::tuple<BAT<v2_void_t, Result> *, AN_indicator_vector *, AN_indicator_vector *, AN_indicator_vector *>
            aggregate_sum_groupedAN(BAT<Head, Tail> *bat, BAT<v2_void_t, v2_resoid_t> *grouping, size_t numGroups,
                    typename Result::type_t AResult = std::get<TypeMap<Result>::v2_encoded_t::As->size() - 1>(*TypeMap<Result>::v2_encoded_t::As),
                    typename Result::type_t AResultInv = std::get<TypeMap<Result>::v2_encoded_t::Ainvs->size() - 1>(*TypeMap<Result>::v2_encoded_t::Ainvs),
                    resoid_t AOID = std::get<v2_resoid_t::As->size() - 1>(*v2_resoid_t::As)) {
                static_assert(sizeof(typename Result::type_t) == 8, "aggregate_sum_groupedAN: 64-bit results on the device path");
                const uint64_t n = device::rows(bat);
                if (n != device::rows(grouping))
                    throw std::runtime_error("aggregate_sum_groupedAN: bat->size() != grouping->size()");
                const ahead_col h = device::col(bat->head, n), t = device::col(bat->tail, n), g = device::col(grouping->tail, n);
                ColumnMetaData md(size_bytes<typename Result::type_t>, static_cast<A_t>(AResult), AResultInv, Result::UNENC_MAX_U, 0);
                auto *r = new TempBAT<v2_void_t, Result>(ColumnDescriptor<v2_void_t>(oid_t(0)), ColumnDescriptor<Result>(md));
                uint64_t *sums = reinterpret_cast<uint64_t *>(r->tail.container->overwrite(numGroups));
                auto & cx = device::Context::get();
                const uint64_t *k = nullptr;
                do {
                    device::raise(ahead_hip_aggregate_sum_grouped_an(&h, &t, &g, numGroups, AResult, AOID, sums, cx.errbuf(0), cx.errbuf(1), cx.errbuf(2), cx.d_counts(),
                            cx.ws, cx.stream));
                    k = cx.read_counts(4);
                } while (!cx.errors_fit( {k[1], k[2], k[3]}));
                const uint64_t e1 = k[1], e2 = k[2], e3 = k[3];
                return std::make_tuple(r, cx.take(0, e1, device::is_enc<Head>), cx.take(1, e2, device::is_enc<Tail>), cx.take(2, e3, true));
            }

        }
    }
}

#endif
