/*
 * column_storage/ColumnMetaData.hpp -- per-column operator parameters.
 *
 * Same members and constructors as the reference's ColumnMetaData
 * (include/column_storage/ColumnMetaData.hpp:150-210): width, seqbase, AN_A (16-bit),
 * AN_Ainv, AN_unencMaxU, AN_unencMinS.  The width is kept as a byte count behind the
 * bits_t / bytes_t tags the reference uses to construct it (:40-148).
 */
#ifndef AHEAD_COLUMN_STORAGE_COLUMNMETADATA_HPP
#define AHEAD_COLUMN_STORAGE_COLUMNMETADATA_HPP

#include <climits>
#include <cstdint>

#include <util/resilience.hpp>

namespace ahead {

    typedef uint16_t A_t;

    struct bits_t {
        uint32_t value;
        constexpr bits_t() : value(0) {}
        constexpr bits_t(uint32_t v) : value(v) {}
    };
    struct bytes_t {
        uint32_t value;
        constexpr bytes_t() : value(0) {}
        constexpr bytes_t(uint32_t v) : value(v) {}
    };
    template<typename T>
    inline constexpr bytes_t size_bytes = bytes_t(sizeof(T));
    template<typename T>
    inline constexpr bits_t size_bits = bits_t(sizeof(T) * CHAR_BIT);

    /* A column width given in bits or bytes; stored in bytes (all widths here are whole bytes). */
    struct data_width_t {
        uint32_t bytes;
        constexpr data_width_t(bits_t b) : bytes(b.value / CHAR_BIT) {}
        constexpr data_width_t(bytes_t b) : bytes(b.value) {}
    };

    struct ColumnMetaData {
        data_width_t width;
        uint64_t seqbase;
        A_t AN_A;
        uint64_t AN_Ainv;
        uint64_t AN_unencMaxU;
        int64_t AN_unencMinS;

        ColumnMetaData()
                : width(bits_t(0)), seqbase(0), AN_A(1), AN_Ainv(1), AN_unencMaxU(0), AN_unencMinS(0) {}
        ColumnMetaData(data_width_t w)
                : width(w), seqbase(0), AN_A(1), AN_Ainv(1), AN_unencMaxU(0), AN_unencMinS(0) {}
        ColumnMetaData(oid_t sb)
                : width(bits_t(0)), seqbase(sb), AN_A(1), AN_Ainv(1), AN_unencMaxU(0), AN_unencMinS(0) {}
        ColumnMetaData(data_width_t w, A_t a, uint64_t a_inv, uint64_t max_u, int64_t min_s)
                : width(w), seqbase(0), AN_A(a), AN_Ainv(a_inv), AN_unencMaxU(max_u), AN_unencMinS(min_s) {}
        ColumnMetaData(const ColumnMetaData &) = default;
        ColumnMetaData(ColumnMetaData &&) = default;
        ColumnMetaData & operator=(const ColumnMetaData &) = default;
    };

    /* Metadata a freshly created column of type V2 carries: the type's default A for encoded
     * types (the largest of its table), plain width otherwise. */
    template<typename V2>
    inline ColumnMetaData default_metadata() {
        if constexpr (V2::is_encoded) {
            return ColumnMetaData(size_bytes<typename V2::type_t>, static_cast<A_t>(V2::A), V2::A_INV, V2::A_UNENC_MAX_U, 0);
        } else {
            return ColumnMetaData(size_bytes<typename V2::type_t>);
        }
    }

}

#endif
