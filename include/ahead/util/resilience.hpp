/*
 * util/resilience.hpp -- AHEAD's scalar and v2 type tags with their AN parameters, for the
 * device-resident C++ operator layer.
 *
 * Same names and members as the reference's include/util/v2types.hpp:259-460 (plain types) and
 * include/util/resilience.hpp:56-432 (AN-encoded types, ANParameters, TypeMap), so caller code
 * that names v2_resint_t::As, TypeMap<T>::v2_encoded_t or T::v2_select_t compiles unchanged.
 * The A / A^-1 tables are the reference's constants (resilience.hpp:63-108); every pair
 * satisfies A * A^-1 == 1 mod 2^k, which the test suite checks.  UNENC_MAX_U follows
 * lib/util/resilience.cpp:43-71.
 */
#ifndef AHEAD_UTIL_RESILIENCE_HPP
#define AHEAD_UTIL_RESILIENCE_HPP

#include <array>
#include <cstddef>
#include <cstdint>

namespace ahead {

    typedef uint8_t tinyint_t;
    typedef uint16_t shortint_t;
    typedef uint32_t int_t;
    typedef uint64_t bigint_t;
    typedef uint32_t oid_t;
    typedef uint32_t id_t;
    typedef uint16_t restiny_t;
    typedef uint32_t resshort_t;
    typedef uint64_t resint_t;
    typedef uint64_t resbigint_t;
    typedef uint64_t resoid_t;

    struct v2_base_t {
    };
    struct v2_tinyint_t;
    struct v2_shortint_t;
    struct v2_int_t;
    struct v2_bigint_t;
    struct v2_oid_t;
    struct v2_void_t;
    struct v2_restiny_t;
    struct v2_resshort_t;
    struct v2_resint_t;
    struct v2_resbigint_t;
    struct v2_resoid_t;

    /* Plain (unencoded) types: v2types.hpp:289-425. */
#define AHEAD_V2_PLAIN(NAME, T, LARGER, SELECT)                                                     \
    struct NAME : public v2_base_t {                                                               \
        typedef T type_t;                                                                          \
        typedef NAME v2_unenc_t;                                                                   \
        typedef NAME v2_copy_t;                                                                    \
        typedef SELECT v2_select_t;                                                                \
        typedef SELECT v2_compare_t;                                                               \
        typedef LARGER v2_larger_t;                                                                \
        static constexpr bool is_encoded = false;                                                  \
    };
    AHEAD_V2_PLAIN(v2_tinyint_t, tinyint_t, v2_shortint_t, v2_tinyint_t)
    AHEAD_V2_PLAIN(v2_shortint_t, shortint_t, v2_int_t, v2_shortint_t)
    AHEAD_V2_PLAIN(v2_int_t, int_t, v2_bigint_t, v2_int_t)
    AHEAD_V2_PLAIN(v2_bigint_t, bigint_t, v2_bigint_t, v2_bigint_t)
    AHEAD_V2_PLAIN(v2_oid_t, oid_t, v2_bigint_t, v2_oid_t)
    AHEAD_V2_PLAIN(v2_void_t, oid_t, v2_bigint_t, v2_oid_t)
#undef AHEAD_V2_PLAIN

    /* ANParameters<V2, CodeT>: the candidate A values (ascending), their inverses mod 2^k and
     * the minimum-bit-flip-weight subset (resilience.hpp:56-109). */
    template<typename V2, typename CodeT>
    struct ANParameters;

    template<>
    struct ANParameters<v2_restiny_t, uint16_t> {
        static constexpr std::array<uint16_t, 8> As {{1, 3, 7, 13, 29, 59, 115, 233}};
        static constexpr std::array<uint16_t, 8> AInvs {{0x0001, 0xAAAB, 0x6DB7, 0x4EC5, 0xC235, 0xD8F3, 0xA4BB, 0xD759}};
        static constexpr std::array<uint16_t, 4> AsBFW {{1, 3, 29, 233}};
        static constexpr std::array<uint16_t, 4> AInvsBFW {{0x0001, 0xAAAB, 0xC235, 0xD759}};
    };

    template<>
    struct ANParameters<v2_resshort_t, uint32_t> {
        static constexpr std::array<uint32_t, 16> As {{1, 3, 7, 13, 29, 61, 119, 233, 463, 947, 1939, 3349, 7785, 14781, 28183, 63877}};
        static constexpr std::array<uint32_t, 16> AInvs {{0x00000001, 0xAAAAAAAB, 0xB6DB6DB7, 0xC4EC4EC5, 0x4F72C235, 0xC10C9715,
            0x46FDD947, 0x1FDCD759, 0xAB67652F, 0xFF30637B, 0xBC452E9B, 0x21B5DA3D, 0x392F51D9, 0x1ABDC995, 0xAB2DA9A7, 0xD142174D}};
        static constexpr std::array<uint32_t, 6> AsBFW {{1, 3, 61, 463, 7785, 63877}};
        static constexpr std::array<uint32_t, 6> AInvsBFW {{0x00000001, 0xAAAAAAAB, 0xC10C9715, 0xAB67652F, 0x392F51D9, 0xD142174D}};
    };

    /* resint, resbigint and resoid share the 64-bit table (resilience.hpp:104-108). */
    struct ANParameters64 {
        static constexpr std::array<uint64_t, 16> As {{1, 3, 7, 13, 29, 61, 125, 225, 445, 881, 2029, 3973, 7841, 16089, 32417, 65117}};
        static constexpr std::array<uint64_t, 16> AInvs {{1ull, 0xAAAAAAAAAAAAAAABull, 0x6DB6DB6DB6DB6DB7ull, 0x4EC4EC4EC4EC4EC5ull,
            0x34F72C234F72C235ull, 0x4FBCDA3AC10C9715ull, 0x1CAC083126E978D5ull, 0x0FEDCBA987654321ull, 0x64194FF6CBA64195ull,
            0xC87FDACE4F9E5D91ull, 0x49AEFF9F19DD6DE5ull, 0xA6B1742D1ABD914Dull, 0x3434A14919D34561ull, 0x227F016A87066169ull,
            0x73B9563A744AE561ull, 0x3E36D3BC8D9BC5F5ull}};
        static constexpr std::array<uint64_t, 5> AsBFW {{1, 3, 125, 881, 16089}};
        static constexpr std::array<uint64_t, 5> AInvsBFW {{1ull, 0xAAAAAAAAAAAAAAABull, 0x1CAC083126E978D5ull, 0xC87FDACE4F9E5D91ull,
            0x227F016A87066169ull}};
    };
    template<>
    struct ANParameters<v2_resint_t, uint64_t> : ANParameters64 {
    };
    template<>
    struct ANParameters<v2_resbigint_t, uint64_t> : ANParameters64 {
    };
    template<>
    struct ANParameters<v2_resoid_t, uint64_t> : ANParameters64 {
    };

    /* AN-encoded types (resilience.hpp:114-276).  A_INV / A_UNENC_MAX_U are the defaults the
     * check operators take (the largest A of the table). */
#define AHEAD_V2_RES(NAME, T, UNENC, LARGER, MAXU)                                                 \
    struct NAME : public v2_base_t {                                                               \
        typedef T type_t;                                                                          \
        typedef UNENC v2_unenc_t;                                                                  \
        typedef NAME v2_copy_t;                                                                    \
        typedef NAME v2_select_t;                                                                  \
        typedef NAME v2_compare_t;                                                                 \
        typedef LARGER v2_larger_t;                                                                \
        typedef ANParameters<NAME, T> Params;                                                      \
        static constexpr bool is_encoded = true;                                                   \
        static constexpr type_t UNENC_MAX_U = MAXU;                                                \
        static constexpr type_t A_UNENC_MAX_U = MAXU;                                              \
        static constexpr const auto *As = &Params::As;                                             \
        static constexpr const auto *Ainvs = &Params::AInvs;                                       \
        static constexpr const auto *AsBFW = &Params::AsBFW;                                       \
        static constexpr const auto *AinvsBFW = &Params::AInvsBFW;                                 \
        static constexpr type_t A = Params::As[Params::As.size() - 1];                             \
        static constexpr type_t A_INV = Params::AInvs[Params::AInvs.size() - 1];                   \
    };
    AHEAD_V2_RES(v2_restiny_t, restiny_t, v2_tinyint_t, v2_resshort_t, 0x00FFu)
    AHEAD_V2_RES(v2_resshort_t, resshort_t, v2_shortint_t, v2_resint_t, 0xFFFFu)
    AHEAD_V2_RES(v2_resint_t, resint_t, v2_int_t, v2_resbigint_t, 0xFFFFFFFFull)
    AHEAD_V2_RES(v2_resbigint_t, resbigint_t, v2_bigint_t, v2_resbigint_t, 0x00FFFFFFFFFFFFFFull)
    AHEAD_V2_RES(v2_resoid_t, resoid_t, v2_oid_t, v2_resbigint_t, 0x00FFFFFFFFFFFFFFull)
#undef AHEAD_V2_RES
    typedef v2_restiny_t v2_restinyint_t;
    typedef v2_resshort_t v2_resshortint_t;

    /* TypeMap (resilience.hpp:278-432): plain <-> encoded counterparts. */
    template<typename V2>
    struct TypeMap;
#define AHEAD_TYPEMAP(V2, BASE, ENC, ACTUAL)                                                       \
    template<>                                                                                     \
    struct TypeMap<V2> {                                                                           \
        typedef BASE v2_base_t;                                                                    \
        typedef ENC v2_encoded_t;                                                                  \
        typedef ACTUAL v2_actual_t;                                                                \
    };
    AHEAD_TYPEMAP(v2_tinyint_t, v2_tinyint_t, v2_restiny_t, v2_tinyint_t)
    AHEAD_TYPEMAP(v2_shortint_t, v2_shortint_t, v2_resshort_t, v2_shortint_t)
    AHEAD_TYPEMAP(v2_int_t, v2_int_t, v2_resint_t, v2_int_t)
    AHEAD_TYPEMAP(v2_bigint_t, v2_bigint_t, v2_resbigint_t, v2_bigint_t)
    AHEAD_TYPEMAP(v2_oid_t, v2_oid_t, v2_resoid_t, v2_oid_t)
    AHEAD_TYPEMAP(v2_void_t, v2_void_t, v2_resoid_t, v2_oid_t)
    AHEAD_TYPEMAP(v2_restiny_t, v2_tinyint_t, v2_restiny_t, v2_restiny_t)
    AHEAD_TYPEMAP(v2_resshort_t, v2_shortint_t, v2_resshort_t, v2_resshort_t)
    AHEAD_TYPEMAP(v2_resint_t, v2_int_t, v2_resint_t, v2_resint_t)
    AHEAD_TYPEMAP(v2_resbigint_t, v2_bigint_t, v2_resbigint_t, v2_resbigint_t)
    AHEAD_TYPEMAP(v2_resoid_t, v2_oid_t, v2_resoid_t, v2_resoid_t)
#undef AHEAD_TYPEMAP

    /* Multiplicative inverse of odd b modulo 2^bits (0 for even b): the value
     * ext_euclidean(b, bits) returns in the reference (resilience.hpp:481-516), computed here by
     * Newton's iteration x <- x * (2 - b * x), which doubles the number of correct low bits. */
    inline constexpr uint64_t ext_euclidean(uint64_t b, unsigned bits) {
        if ((b & 1u) == 0)
            return 0;
        uint64_t x = b;
        for (int i = 0; i < 6; ++i)
            x *= 2u - b * x;
        return bits >= 64 ? x : (x & ((uint64_t(1) << bits) - 1u));
    }

}

#endif
