// select_an.hip -- selectAN on AN-encoded codes for gfx950, with the fused
// check-and-decode + sum of the select+sum workload.
//
// Semantics (lib/column_operators/AN/selectAN_scalar.tcc:66-95 and :149-181): for every row
// `pos` of the encoded tail, in order,
//   (a) if (code * A^-1 mod 2^k) > maxU       -> err  += pos * A_OID
//   (b) if Op(code, th) [Comb(Op1, Op2)]      -> oids += (uint32)(seqbase + pos) * A_head
// with unsigned comparisons on the raw codes against thresholds the caller already encoded.
// Corrupt rows still take part in (b).
//
// Kernel shape: one 256-thread workgroup per tile of ITERS sub-tiles of 16 KiB (4 x 16-byte
// loads per thread; 8192 rows of 16-bit codes, 4096 of 32-bit, 2048 of 64-bit).
//  phase 1  streams the tile once with coalesced 128-bit non-temporal loads and keeps only one
//           match bit per row (LDS) and one "has an error" bit per sub-tile;
//  resolve  publishes the tile's counts and takes its prefix by decoupled look-back
//           (an_device.hpp) -- one ticket and one look-back per ITERS sub-tiles;
//  phase 2  per sub-tile, packs the per-chunk counts into one uint64 (16-bit fields) so that one
//           block scan ranks every row in order, stages matching row indices in LDS and writes
//           them as one contiguous, coalesced run of 64-bit OIDs.  OIDs depend only on the row
//           index, so the input is never read twice.  Sub-tiles holding an error (rare) are
//           re-read to rank their error positions, which are scattered.
// The fused decoded sum leaves per-tile partials that k_sum_partials adds up; the kernel
// boundary orders them without any device-wide fence.
#include <mutex>
#include <string>

#include "select_tile.hpp"

using namespace ahead_dev;

namespace {

template <typename CodeT, int COMB, bool SUM>
__global__ __launch_bounds__(kBlock) void k_select(SelectArgs a) {
    __shared__ uint32_t s_tile;
    __shared__ SelectShared<kBlock * 4 * (16 / sizeof(CodeT))> sh;
    const uint32_t tile = take_tile(a.lb, &s_tile);
    if (tile >= a.lb.num_tiles) return;
    select_tile<CodeT, COMB, SUM>(a, tile, sh);
}

// Adds the per-tile partial sums of a k_select launch (kernel order makes them visible).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const uint64_t *partials, uint32_t n, uint64_t *out) {
    __shared__ uint64_t s_scan[kWaves];
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) s += partials[i];
    s = block_sum<uint64_t>(s, s_scan);
    if (threadIdx.x == 0) *out = s;
}

// canonical interval form of one predicate in a code width with mask m
struct Canon {
    uint64_t lo = 0, span = 0;
    uint32_t neg = 0, valid = 1;
};
bool canon(int op, uint64_t th, uint64_t m, Canon *c) {
    th &= m;
    switch (op) {
        case AHEAD_LT:
            if (th == 0) c->valid = 0; else { c->lo = 0; c->span = th - 1; }
            return true;
        case AHEAD_LE: c->lo = 0; c->span = th; return true;
        case AHEAD_GT:
            if (th == m) c->valid = 0; else { c->lo = th + 1; c->span = m - th - 1; }
            return true;
        case AHEAD_GE: c->lo = th; c->span = m - th; return true;
        case AHEAD_EQ: c->lo = th; c->span = 0; return true;
        case AHEAD_NE: c->lo = th; c->span = 0; c->neg = 1; return true;
        default: return false;
    }
}

template <typename CodeT, int COMB, bool SUM>
void launch(const SelectArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_select<CodeT, COMB, SUM>), dim3(a.lb.num_tiles), dim3(kBlock), 0, s, a);
}

template <typename CodeT>
void dispatch(const SelectArgs &a, int comb, bool sum, hipStream_t s) {
    if (comb == AHEAD_COMB_NONE) {
        sum ? launch<CodeT, AHEAD_COMB_NONE, true>(a, s) : launch<CodeT, AHEAD_COMB_NONE, false>(a, s);
    } else if (comb == AHEAD_COMB_AND) {
        sum ? launch<CodeT, AHEAD_COMB_AND, true>(a, s) : launch<CodeT, AHEAD_COMB_AND, false>(a, s);
    } else {
        sum ? launch<CodeT, AHEAD_COMB_OR, true>(a, s) : launch<CodeT, AHEAD_COMB_OR, false>(a, s);
    }
}

}  // namespace

using namespace ahead_host;

namespace {

// Arguments common to both select entry points, validated; the caller supplies the partial-sum
// array and the look-back state.
struct SelectPlan {
    SelectArgs a{};
    LbPlan plan{};
    int comb = 0;
};

int select_prepare(const ahead_col *col, int op1, uint64_t th1, int op2, uint64_t th2, int comb, uint64_t a_head,
                   uint64_t a_oid, uint64_t *oids, uint64_t oids_cap, uint64_t *err, uint64_t err_cap,
                   uint64_t *d_counts, SelectPlan *p) {
    int st = check_col(col, "select.col", false, true);
    if (st != AHEAD_OK) return st;
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "select: d_counts is NULL");
    if ((oids == nullptr && oids_cap) || (err == nullptr && err_cap))
        return fail(AHEAD_ERR_INVALID, "select: NULL output with nonzero capacity");
    if (comb < AHEAD_COMB_NONE || comb > AHEAD_COMB_OR) return fail(AHEAD_ERR_INVALID, "select: bad combiner");
    const uint64_t m = width_mask(col->width);
    Canon c1, c2;
    if (!canon(op1, th1, m, &c1)) return fail(AHEAD_ERR_INVALID, "select: bad op1");
    if (comb != AHEAD_COMB_NONE && !canon(op2, th2, m, &c2)) return fail(AHEAD_ERR_INVALID, "select: bad op2");
    // Sub-tiles of 16 KiB; one look-back ticket covers `iters` of them.  Fewer tickets keep the
    // contended ticket atomic and the look-back off the critical path; iters stays small enough
    // that the grid still holds >= min_tiles tiles (default 4096, 16 per CU) when it can.
    const uint64_t sub_rows = 4ull * kBlock * (16u / col->width);
    const uint64_t subs = (col->n + sub_rows - 1) / sub_rows;
    const uint32_t iters = subtiles_per_tile(subs);
    p->plan = lb_plan(col->n, static_cast<uint32_t>(sub_rows * iters), 2);
    p->comb = comb;
    SelectArgs &a = p->a;
    a.codes = col->data;
    a.n = col->n;
    a.seqbase = col->seqbase;
    a.a_inv = col->a_inv;
    a.max_u = col->max_u;
    a.lo1 = c1.lo; a.span1 = c1.span; a.neg1 = c1.neg; a.valid1 = c1.valid;
    a.lo2 = c2.lo; a.span2 = c2.span; a.neg2 = c2.neg; a.valid2 = c2.valid;
    a.a_head = a_head;
    a.a_oid = a_oid;
    a.oids = oids;
    a.oids_cap = oids == nullptr ? 0 : oids_cap;
    a.err = err;
    a.err_cap = err == nullptr ? 0 : err_cap;
    a.d_counts = d_counts;
    a.iters = iters;
    return AHEAD_OK;
}

int select_launch(const SelectPlan &p, uint64_t width, uint64_t *d_sum, hipStream_t s) {
    const bool sum = d_sum != nullptr;
    switch (width) {
        case 2: dispatch<uint16_t>(p.a, p.comb, sum, s); break;
        case 4: dispatch<uint32_t>(p.a, p.comb, sum, s); break;
        default: dispatch<uint64_t>(p.a, p.comb, sum, s); break;
    }
    AHEAD_HIP_TRY(hipGetLastError());
    if (sum) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, p.a.partials, p.plan.num_tiles, d_sum);
        AHEAD_HIP_TRY(hipGetLastError());
    }
    return AHEAD_OK;
}

}  // namespace

// ---- the allocation-free entry point's scratch ------------------------------------------------
// One array per GPU the library runs on, held until the process ends: the workspace header (tile
// ticket), then the per-tile partial sums, then the look-back status words of the two output
// streams.  Host code never passes `scratch` itself to a kernel: it is a host-side stand-in for the
// symbol.  Its device address, per device, comes from hipGetSymbolAddress at the first call on that
// device.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
struct NoAllocState {
    uint8_t *ptr = nullptr;  // device address of `scratch` on this device
    uint32_t epoch = 0;      // last look-back epoch handed out on this device
};
NoAllocState g_noalloc[kNoAllocMaxDevices];
std::mutex g_noalloc_mu;

// Layout of the scratch for a plan of `tiles` look-back tiles: header, partials, status words.
size_t noalloc_partials_off() { return kHeaderBytes; }
size_t noalloc_status_off(uint32_t tiles) { return align_up(kHeaderBytes + size_t(tiles) * 8, 256); }
size_t noalloc_bytes(uint32_t tiles) { return noalloc_status_off(tiles) + size_t(tiles) * 2 * 8; }

}  // namespace

extern "C" int ahead_hip_select_an(const ahead_col *col, int op1, uint64_t th1, int op2, uint64_t th2,
                                   int comb, uint64_t a_head, uint64_t a_oid, uint64_t *oids,
                                   uint64_t oids_cap, uint64_t *err, uint64_t err_cap,
                                   uint64_t *d_counts, uint64_t *d_sum, ahead_hip_ws *ws,
                                   void *stream) {
    SelectPlan p;
    int st = select_prepare(col, op1, th1, op2, th2, comb, a_head, a_oid, oids, oids_cap, err, err_cap, d_counts, &p);
    if (st != AHEAD_OK) return st;
    if ((st = ws_ensure(ws, static_cast<size_t>(p.plan.num_tiles) * 8)) != AHEAD_OK) return st;
    p.a.partials = reinterpret_cast<uint64_t *>(ws_region(ws, 0));
    if ((st = lb_args(ws, p.plan, &p.a.lb)) != AHEAD_OK) return st;
    return select_launch(p, col->width, d_sum, static_cast<hipStream_t>(stream));
}

extern "C" int ahead_hip_select_an_noalloc(const ahead_col *col, int op1, uint64_t th1, int op2, uint64_t th2,
                                           int comb, uint64_t a_head, uint64_t a_oid, uint64_t *oids,
                                           uint64_t oids_cap, uint64_t *err, uint64_t err_cap,
                                           uint64_t *d_counts, uint64_t *d_sum, void *stream) {
    SelectPlan p;
    int st = select_prepare(col, op1, th1, op2, th2, comb, a_head, a_oid, oids, oids_cap, err, err_cap, d_counts, &p);
    if (st != AHEAD_OK) return st;
    const size_t need = noalloc_bytes(p.plan.num_tiles);
    if (need > kNoAllocScratchBytes)
        return fail(AHEAD_ERR_TOO_LARGE, "select_an_noalloc: " + std::to_string(col->n) + " rows need " +
                                             std::to_string(need) + " bytes of scan state; the static scratch holds " +
                                             std::to_string(kNoAllocScratchBytes));
    int dev = 0;
    AHEAD_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(AHEAD_ERR_INVALID, "select_an_noalloc: device index out of range");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t epoch = 0;
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        NoAllocState &g = g_noalloc[dev];
        if (g.ptr == nullptr) {
            // set-up, once per device with that device current: the array's device address, then
            // a clear of the whole array (ticket counter and status words must start at zero)
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || ptr == nullptr) return hip_fail(e, "hipGetSymbolAddress(scratch)");
            if ((st = fill_u64(static_cast<uint64_t *>(ptr), kNoAllocScratchBytes / 8, 0, s)) != AHEAD_OK) return st;
            g.ptr = static_cast<uint8_t *>(ptr);
            g.epoch = 0;
        }
        if (g.epoch + 1 >= (1u << 30)) {
            // status words of every earlier call carry an epoch below 2^30: clear them (stream-ordered)
            if ((st = fill_u64(reinterpret_cast<uint64_t *>(g.ptr + kHeaderBytes), (kNoAllocScratchBytes - kHeaderBytes) / 8, 0,
                               s)) != AHEAD_OK)
                return st;
            g.epoch = 0;
        }
        epoch = ++g.epoch;
        base = g.ptr;
    }
    p.a.partials = reinterpret_cast<uint64_t *>(base + noalloc_partials_off());
    p.a.lb.hdr = reinterpret_cast<WsHeader *>(base);
    p.a.lb.status = reinterpret_cast<uint64_t *>(base + noalloc_status_off(p.plan.num_tiles));
    p.a.lb.num_tiles = p.plan.num_tiles;
    p.a.lb.epoch = epoch;
    return select_launch(p, col->width, d_sum, s);
}
