"""AN-code parameters and column types of AHEAD, as plain data for the host side.

The tables restate the reference's ANParameters specialisations
(include/util/resilience.hpp:56-109) and the UNENC_MAX_U constants
(lib/util/resilience.cpp:43-71).  A code of an AN-encoded type stores d * A mod 2^k; it is
corrupt iff (code * A^-1 mod 2^k) > maxU.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

MASK64 = (1 << 64) - 1

# ANParameters<v2_restiny_t, uint16_t> (resilience.hpp:63-64)
RESTINY_AS = (1, 3, 7, 13, 29, 59, 115, 233)
RESTINY_AINVS = (0x0001, 0xAAAB, 0x6DB7, 0x4EC5, 0xC235, 0xD8F3, 0xA4BB, 0xD759)
# ANParameters<v2_restiny_t, uint32_t> (resilience.hpp:76-78)
RESTINY32_AS = (1, 3, 7, 13, 29, 59, 115, 233, 487, 857, 1939, 3813, 7463, 13963, 27247, 55831)
RESTINY32_AINVS = (0x00000001, 0xAAAAAAAB, 0xB6DB6DB7, 0xC4EC4EC5, 0x4F72C235, 0xA08AD8F3, 0x61F2A4BB,
                   0x1FDCD759, 0xB98F81D7, 0x484A14E9, 0xBC452E9B, 0xEECFFEED, 0xFC778297, 0xB9860123,
                   0x833A348F, 0x4DFFFDA7)
# ANParameters<v2_resshort_t, uint32_t> (resilience.hpp:90-92)
RESSHORT_AS = (1, 3, 7, 13, 29, 61, 119, 233, 463, 947, 1939, 3349, 7785, 14781, 28183, 63877)
RESSHORT_AINVS = (0x00000001, 0xAAAAAAAB, 0xB6DB6DB7, 0xC4EC4EC5, 0x4F72C235, 0xC10C9715, 0x46FDD947,
                  0x1FDCD759, 0xAB67652F, 0xFF30637B, 0xBC452E9B, 0x21B5DA3D, 0x392F51D9, 0x1ABDC995,
                  0xAB2DA9A7, 0xD142174D)
# ANParameters<v2_resint_t, uint64_t> (resilience.hpp:104-106); also used by resbigint / resoid
RESINT_AS = (1, 3, 7, 13, 29, 61, 125, 225, 445, 881, 2029, 3973, 7841, 16089, 32417, 65117)
RESINT_AINVS = (1, 0xAAAAAAAAAAAAAAAB, 0x6DB6DB6DB6DB6DB7, 0x4EC4EC4EC4EC4EC5, 0x34F72C234F72C235,
                0x4FBCDA3AC10C9715, 0x1CAC083126E978D5, 0x0FEDCBA987654321, 0x64194FF6CBA64195,
                0xC87FDACE4F9E5D91, 0x49AEFF9F19DD6DE5, 0xA6B1742D1ABD914D, 0x3434A14919D34561,
                0x227F016A87066169, 0x73B9563A744AE561, 0x3E36D3BC8D9BC5F5)
# min-bit-flip-weight subsets (resilience.hpp:65,93,107)
RESTINY_AS_BFW = (RESTINY_AS[0], RESTINY_AS[1], RESTINY_AS[4], RESTINY_AS[7])
RESSHORT_AS_BFW = (RESSHORT_AS[0], RESSHORT_AS[1], RESSHORT_AS[5], RESSHORT_AS[8], RESSHORT_AS[12],
                   RESSHORT_AS[15])
RESINT_AS_BFW = (RESINT_AS[0], RESINT_AS[1], RESINT_AS[6], RESINT_AS[9], RESINT_AS[13])

# Operator defaults baked into include/column_operators/OperatorsAN.hpp
A_OID_ERR = RESINT_AS_BFW[-1]          # 16089: error positions (OperatorsAN.hpp:38 and others)
A_OID = RESINT_AS[-1]                  # 65117: select heads, group ids, grouped-sum result
A_OID_INV = RESINT_AINVS[-1]
A_RESULT = RESINT_AS[-1]               # aggregate_sum_groupedAN / aggregate_mul_sumAN result A


def inverse(a: int, bits: int) -> int:
    """Multiplicative inverse of odd a modulo 2**bits (Newton iteration), or 0 for even a."""
    if a % 2 == 0:
        return 0
    m = (1 << bits) - 1
    x = a
    for _ in range(7):
        x = (x * (2 - a * x)) & MASK64
    return x & m


@dataclass(frozen=True)
class ColType:
    """One BAT side's type: code width, unencoded width and maxU (ColumnMetaData)."""

    name: str
    width: int          # bytes per stored element
    encoded: bool
    unenc_bits: int     # TypeMap<T>::v2_base_t::type_t bits
    max_u: int          # AN_unencMaxU
    larger_bits: int    # T::v2_larger_t::type_t bits (binary group-by hash width)


# resilience.cpp:43-71 / v2types.hpp
RESTINY = ColType("restiny", 2, True, 8, 0x00FF, 32)
RESSHORT = ColType("resshort", 4, True, 16, 0xFFFF, 64)
RESINT = ColType("resint", 8, True, 32, 0xFFFFFFFF, 64)
RESBIGINT = ColType("resbigint", 8, True, 64, 0x00FFFFFFFFFFFFFF, 64)
RESOID = ColType("resoid", 8, True, 32, 0x00FFFFFFFFFFFFFF, 64)
TINYINT = ColType("tinyint", 1, False, 8, 0xFF, 16)
SHORTINT = ColType("shortint", 2, False, 16, 0xFFFF, 32)
INT = ColType("int", 4, False, 32, 0xFFFFFFFF, 64)
BIGINT = ColType("bigint", 8, False, 64, MASK64, 64)
OID = ColType("oid", 4, False, 32, 0xFFFFFFFF, 64)
VOID = ColType("void", 4, False, 32, 0xFFFFFFFF, 64)


@dataclass(frozen=True)
class ColMeta:
    """ColumnMetaData of one BAT side plus its type."""

    type: ColType
    a: int = 1
    a_inv: int = 1
    seqbase: int = 0

    @property
    def width(self) -> int:
        return self.type.width

    @property
    def max_u(self) -> int:
        return self.type.max_u

    def with_seqbase(self, seqbase: int) -> "ColMeta":
        return replace(self, seqbase=seqbase)


def an_meta(t: ColType, a: int) -> ColMeta:
    """Metadata of an AN-encoded column of type t with multiplier a."""
    return ColMeta(t, a, inverse(a, 8 * t.width))


def plain_meta(t: ColType, seqbase: int = 0) -> ColMeta:
    return ColMeta(t, 1, 1, seqbase)


_UNCHECKED: dict = {}


def unchecked_type(t: ColType, raw: bool = False) -> ColType:
    """Values stored at the code width of t with no check: every code is valid (maxU = all ones).
    With A = 1 the hardened operators over such columns are the plain operators of Operators.hpp
    (the special case A = 1 of every AN operator).  Decoded values keep t's unencoded width as
    their key width; raw AN codes (raw=True) use the whole code width."""
    for suffix in ("!unchecked", "!raw"):
        if t.name.endswith(suffix):
            t = _UNCHECKED[t.name]
    bits = 8 * t.width if raw else t.unenc_bits
    u = ColType(f"{t.name}{'!raw' if raw else '!unchecked'}", t.width, True, bits, (1 << (8 * t.width)) - 1,
                max(t.larger_bits, bits))
    _UNCHECKED[u.name] = t
    return u


def checked_of(t: ColType) -> ColType:
    """The AN type an unchecked type was made from."""
    return _UNCHECKED.get(t.name, t)


def unchecked_meta(t: ColType, a: int = 1) -> ColMeta:
    """An unchecked column of t's code width; a = 1 for values, a = A for raw AN codes (thresholds
    are still encoded with A, and the operators compare, join and group the codes themselves)."""
    return ColMeta(unchecked_type(t, raw=a != 1), a, 1)
