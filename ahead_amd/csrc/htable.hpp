// htable.hpp -- open-addressing hash table of 64-bit keys in HBM for the hash-join build and the
// group-by dictionary.
//
// A slot word holds key + 1 (0 = empty) and is claimed with one 64-bit compare-and-swap, so no
// lane ever waits for another lane's write.  The single key that key + 1 cannot represent
// (2^64 - 1) lives in an extra slot at index `size`.  Linear probing; the table is sized to at
// most half full, so every probe sequence ends at an empty slot.
#pragma once

#include "an_device.hpp"

namespace ahead_dev {

struct HTable {
    unsigned long long *keys;  // [size + 1]
    uint32_t bits;             // size = 2^bits
    uint32_t *overflow;        // set when an insert found no free slot within the probe bound

    __device__ __forceinline__ uint32_t size() const { return 1u << bits; }
    __device__ __forceinline__ uint32_t home(uint64_t key) const {
        // xor-fold: dense small keys map to themselves, wider keys are folded in
        const uint64_t x = key ^ (key >> bits) ^ (key >> (2 * bits)) ^ (key >> 48);
        return static_cast<uint32_t>(x) & (size() - 1u);
    }

    // Inserts key (if absent) and returns its slot, or UINT32_MAX if the table is full.  *fresh
    // (if given) tells whether this call claimed the slot.
    __device__ __forceinline__ uint32_t insert(uint64_t key, bool *fresh = nullptr) const {
        if (key == ~0ull) {
            const unsigned long long prev = atomicCAS(&keys[size()], 0ull, 1ull);
            if (fresh) *fresh = prev == 0ull;
            return size();
        }
        const unsigned long long word = static_cast<unsigned long long>(key) + 1ull;
        uint32_t s = home(key);
        for (uint32_t probes = 0; probes < size(); ++probes) {
            const unsigned long long prev = atomicCAS(&keys[s], 0ull, word);
            if (fresh) *fresh = prev == 0ull;
            if (prev == 0ull || prev == word) return s;
            s = (s + 1u) & (size() - 1u);
        }
        atomicOr(overflow, 1u);
        return UINT32_MAX;
    }

    // Returns the slot of key, or UINT32_MAX if absent.
    __device__ __forceinline__ uint32_t find(uint64_t key) const {
        if (key == ~0ull) return keys[size()] != 0ull ? size() : UINT32_MAX;
        const unsigned long long word = static_cast<unsigned long long>(key) + 1ull;
        uint32_t s = home(key);
        for (uint32_t probes = 0; probes < size(); ++probes) {
            const unsigned long long k = keys[s];
            if (k == word) return s;
            if (k == 0ull) return UINT32_MAX;
            s = (s + 1u) & (size() - 1u);
        }
        return UINT32_MAX;
    }

    // find() for G keys at once (ok[g] false: no lookup, UINT32_MAX).  Each round loads the current
    // slot of every key still open in one batch, so a wave waits once per round, not once per key
    // and slot.
    template <int G>
    __device__ __forceinline__ void find_n(const uint64_t (&key)[G], const bool (&ok)[G], uint32_t (&slot)[G]) const {
        uint32_t s[G];
        bool open[G];
        const uint32_t mask = size() - 1u;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            slot[g] = UINT32_MAX;
            open[g] = ok[g];
            if (ok[g] && key[g] == ~0ull) {
                slot[g] = keys[size()] != 0ull ? size() : UINT32_MAX;
                open[g] = false;
            }
            s[g] = home(key[g]);
        }
        for (uint32_t round = 0; round < size(); ++round) {
            unsigned long long k[G];
#pragma unroll
            for (int g = 0; g < G; ++g) k[g] = keys[open[g] ? s[g] : 0u];
            bool more = false;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (!open[g]) continue;
                if (k[g] == static_cast<unsigned long long>(key[g]) + 1ull) {
                    slot[g] = s[g];
                    open[g] = false;
                } else if (k[g] == 0ull) {
                    open[g] = false;
                } else {
                    s[g] = (s[g] + 1u) & mask;
                    more = true;
                }
            }
            if (!more) break;
        }
    }
};

}  // namespace ahead_dev
