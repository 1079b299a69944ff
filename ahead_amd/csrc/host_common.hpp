// host_common.hpp -- host-side helpers shared by the entry points of libahead_hip.so:
// error reporting, argument validation, the workspace and launch geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/ahead_hip.h"
#include "an_device.hpp"

// A workspace holds two device allocations.  `base` is the WsHeader followed by the operators'
// scratch regions (hash tables, partials, offsets), laid out per operator.  `status` holds only
// look-back status words: a word whose epoch equals the current launch's must have been written
// by that launch, so no other data may ever be stored there (scratch data of an earlier operator
// could carry any bit pattern, the current epoch included).
struct ahead_hip_ws {
    int device = 0;
    void *base = nullptr;    // WsHeader + scratch regions
    size_t bytes = 0;
    void *status = nullptr;  // look-back status words only
    size_t status_bytes = 0;
    uint32_t epoch = 0;      // last look-back epoch handed out
};

namespace ahead_host {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define AHEAD_HIP_TRY(expr)                                    \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return ahead_host::hip_fail(_e, #expr); \
    } while (0)

constexpr size_t kHeaderBytes = sizeof(ahead_dev::WsHeader);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Ensures the scratch allocation holds `bytes` beyond the header.  Growing frees the old
// allocation (hipFree waits for the device) and zero-fills the new one.
int ws_ensure(ahead_hip_ws *ws, size_t bytes);

// Ensures the status allocation holds `bytes`; grown the same way.  The epoch keeps counting
// (a zero word carries epoch 0, which is never handed out).
int ws_ensure_status(ahead_hip_ws *ws, size_t bytes);

// Next look-back epoch (1 .. 2^30-1).  On wrap-around the status allocation is cleared.
int ws_next_epoch(ahead_hip_ws *ws, uint32_t *epoch);

inline char *ws_region(ahead_hip_ws *ws, size_t offset) {
    return static_cast<char *>(ws->base) + kHeaderBytes + offset;
}
inline ahead_dev::WsHeader *ws_header(ahead_hip_ws *ws) {
    return static_cast<ahead_dev::WsHeader *>(ws->base);
}

int check_col(const ahead_col *c, const char *name, bool allow_void, bool need_encoded);

// Multiplicative inverse of odd a modulo 2^bits (Newton iteration x <- x(2 - a x)); 0 for even a.
inline uint64_t inverse_mod2k(uint64_t a, uint32_t bits) {
    if ((a & 1ull) == 0) return 0;
    uint64_t x = a;
    for (int i = 0; i < 6; ++i) x *= 2ull - a * x;
    return bits >= 64 ? x : (x & ((1ull << bits) - 1ull));
}

inline uint64_t width_mask(uint32_t w) { return w >= 8 ? ~0ull : ((1ull << (8 * w)) - 1ull); }
inline uint64_t bits_mask(uint32_t b) { return b >= 64 ? ~0ull : ((1ull << b) - 1ull); }

// Look-back launch description: status words for S streams over num_tiles tiles.
struct LbPlan {
    uint32_t num_tiles = 0;
    size_t status_bytes = 0;
};
inline LbPlan lb_plan(uint64_t n, uint32_t tile_rows, int streams) {
    LbPlan p;
    const uint64_t t = (n + tile_rows - 1) / tile_rows;
    p.num_tiles = static_cast<uint32_t>(t == 0 ? 1 : t);
    p.status_bytes = align_up(static_cast<size_t>(p.num_tiles) * streams * sizeof(uint64_t), 256);
    return p;
}
// Look-back arguments of one launch: a fresh epoch and the status allocation (grown to the plan).
int lb_args(ahead_hip_ws *ws, const LbPlan &p, ahead_dev::LbArgs *out);

// Device fill kernels (graph-capture safe replacements for hipMemsetAsync).
int fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s);
int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);

// Process-wide tiling knobs (ahead_hip_set_tiling).
using ahead_dev::kMaxSubtiles;
struct Tiling {
    uint32_t max_subtiles = kMaxSubtiles;
    uint32_t min_tiles = 512;
};
Tiling tiling();
// Sub-tiles per look-back tile for a stream of `subs` sub-tiles: the largest power of two
// <= max_subtiles that still leaves >= min_tiles tiles.
uint32_t subtiles_per_tile(uint64_t subs);

// checkAN for composite operators: its scratch tables go at `offset` of the workspace, after the
// caller's own regions; the caller grows the workspace to offset + check_an_scratch_bytes first.
int check_an_ws(const ahead_col *col, uint64_t a_oid, uint64_t *err, uint64_t err_cap, uint64_t *d_counts,
                ahead_hip_ws *ws, void *stream, size_t offset);
size_t check_an_status_bytes(uint64_t n, uint32_t width);
size_t check_an_scratch_bytes(uint64_t n, uint32_t width);

}  // namespace ahead_host
