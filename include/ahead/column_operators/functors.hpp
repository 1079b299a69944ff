/*
 * column_operators/functors.hpp -- the predicate combiners of two-predicate selections:
 * ahead::and_is / ahead::or_is (include/column_operators/functors.hpp:223-318 of the
 * reference), used as the OpCombine template argument of selectAN<Op1, Op2, OpCombine>.
 */
#ifndef AHEAD_COLUMN_OPERATORS_FUNCTORS_HPP
#define AHEAD_COLUMN_OPERATORS_FUNCTORS_HPP

namespace ahead {

    template<typename Op = void>
    struct and_is {
        constexpr bool operator()(bool a, bool b) const { return a & b; }
    };

    template<typename Op = void>
    struct or_is {
        constexpr bool operator()(bool a, bool b) const { return a | b; }
    };

}

#endif
