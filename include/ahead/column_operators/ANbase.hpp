/*
 * column_operators/ANbase.hpp -- AN_indicator_vector and hash_side_t
 * (include/column_operators/ANbase.hpp:32 and Operators.hpp:71 of the reference).
 */
#ifndef AHEAD_COLUMN_OPERATORS_ANBASE_HPP
#define AHEAD_COLUMN_OPERATORS_ANBASE_HPP

#include <vector>

#include <util/resilience.hpp>

namespace ahead {
    namespace bat {
        namespace ops {
            /* Error positions pos * A_OID (mod 2^64), in the order the operator scans its input. */
            typedef std::vector<resoid_t> AN_indicator_vector;
        }
    }

    /* Which join input the hash join builds its table on; right keeps the left input's order. */
    enum class hash_side_t {
        left, right
    };
}

#endif
