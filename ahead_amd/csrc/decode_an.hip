// decode_an.hip -- encodeAN, checkAN, decodeAN and checkAndDecodeAN<void,Tail> for gfx950.
//
// Semantics (lib/column_operators/AN/encdecAN_scalar.tcc):
//   encodeAN          :44-67   out[i] = (code_t)v_i * A
//   checkAN           :69-94   err += i*A_OID for every i with (c_i * A^-1 mod 2^k) > maxU
//   decodeAN          :96-126  every row decoded (truncated to the unencoded width), err as above
//   checkAndDecodeAN  :185-210 corrupt rows are dropped from the decoded output, err as above
// Two engines: k_decode (tiles with look-back, decoded values staged in LDS by rank; used when a
// check-and-decode output holds fewer entries than the column has rows) and the streaming map
// engine of maptile.hpp (DecodeMapOp: every other call).
#include "maptile.hpp"

using namespace ahead_dev;

namespace {

enum Mode { kCheck = 0, kDecode = 1, kCheckDecode = 2 };

struct DecodeArgs {
    const void *codes;
    uint64_t n;
    uint64_t a_inv;
    uint64_t max_u;
    uint64_t a_oid;
    void *dec;
    uint64_t dec_cap;
    uint64_t *err;
    uint64_t err_cap;
    uint64_t *d_counts;
    LbArgs lb;
};

template <typename CodeT, typename OutT, int MODE>
__global__ __launch_bounds__(kBlock) void k_decode(DecodeArgs a) {
    constexpr int VEC = 16 / sizeof(CodeT);
    constexpr int CH = 4;
    constexpr int CHUNK = kBlock * VEC;
    constexpr int TILE = CHUNK * CH;

    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_scan[kWaves];
    __shared__ uint32_t s_excl[2];
    __shared__ OutT s_vals[MODE == kCheckDecode ? TILE : 1];

    const uint32_t tile = take_tile(a.lb, &s_tile);
    if (tile >= a.lb.num_tiles) return;
    const uint64_t base = static_cast<uint64_t>(tile) * TILE;
    const CodeT *codes = static_cast<const CodeT *>(a.codes);
    const bool aligned = (reinterpret_cast<uintptr_t>(codes) & 15u) == 0;
    const CodeT ainv = static_cast<CodeT>(a.a_inv);
    const CodeT maxu = static_cast<CodeT>(a.max_u);
    OutT *dec = static_cast<OutT *>(a.dec);

    CodeT d[CH][VEC];
    uint32_t okbits = 0, ebits = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const uint64_t r0 = base + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC;
        CodeT v[VEC];
        if (aligned && r0 + VEC <= a.n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(codes + r0);
            __builtin_memcpy(&v[0], &q, 16);
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) v[k] = (r0 + k < a.n) ? codes[r0 + k] : CodeT(0);
        }
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const bool live = r0 + k < a.n;
            d[j][k] = an_mul<CodeT>(v[k], ainv);
            const bool bad = live && d[j][k] > maxu;
            ebits |= static_cast<uint32_t>(bad) << (j * VEC + k);
            okbits |= static_cast<uint32_t>(live && !bad) << (j * VEC + k);
            if constexpr (MODE == kDecode) {
                if (live) dec[r0 + k] = static_cast<OutT>(d[j][k]);
            }
        }
    }

    uint64_t tot_m = 0, ex_m = 0, tot_e = 0, ex_e = 0;
    if constexpr (MODE == kCheckDecode) ex_m = block_excl_scan<uint64_t>(pack_counts<CH, VEC>(okbits), tot_m, s_scan);
    const int any_err = __syncthreads_or(ebits != 0u);
    if (any_err) ex_e = block_excl_scan<uint64_t>(pack_counts<CH, VEC>(ebits), tot_e, s_scan);

    const uint32_t agg[2] = {fields_sum(tot_m), fields_sum(tot_e)};
    lb_resolve<2>(a.lb, tile, agg, s_excl);
    const uint64_t xm = s_excl[0], xe = s_excl[1];

    if constexpr (MODE == kCheckDecode) {
        for_each_ranked<CH, VEC>(okbits, ex_m, tot_m, [&](uint32_t r, int j, int k) {
            s_vals[r] = static_cast<OutT>(d[j][k]);
        });
    }
    if (any_err) {
        for_each_ranked<CH, VEC>(ebits, ex_e, tot_e, [&](uint32_t r, int j, int k) {
            const uint64_t idx = xe + r;
            const uint64_t pos = base + static_cast<uint64_t>(j) * CHUNK + threadIdx.x * VEC + k;
            if (idx < a.err_cap) a.err[idx] = pos * a.a_oid;
        });
    }
    if constexpr (MODE == kCheckDecode) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < agg[0]; k += kBlock) {
            const uint64_t idx = xm + k;
            if (idx < a.dec_cap) dec[idx] = s_vals[k];
        }
    }
    if (tile == a.lb.num_tiles - 1 && threadIdx.x == 0) {
        if constexpr (MODE == kCheckDecode) {
            a.d_counts[0] = xm + agg[0];
            a.d_counts[1] = xe + agg[1];
        } else if constexpr (MODE == kDecode) {
            a.d_counts[0] = a.n;
            a.d_counts[1] = xe + agg[1];
        } else {
            a.d_counts[0] = xe + agg[1];
        }
    }
}

template <typename CodeT, int MODE>
int launch_mode(const DecodeArgs &a, uint32_t unenc_bits, hipStream_t s) {
    const dim3 g(a.lb.num_tiles), b(kBlock);
    switch (unenc_bits) {
        case 8: hipLaunchKernelGGL((k_decode<CodeT, uint8_t, MODE>), g, b, 0, s, a); break;
        case 16: hipLaunchKernelGGL((k_decode<CodeT, uint16_t, MODE>), g, b, 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_decode<CodeT, uint32_t, MODE>), g, b, 0, s, a); break;
        case 64: hipLaunchKernelGGL((k_decode<CodeT, uint64_t, MODE>), g, b, 0, s, a); break;
        default: return ahead_host::fail(AHEAD_ERR_INVALID, "decode: unenc_bits must be 8, 16, 32 or 64");
    }
    return AHEAD_OK;
}

// One row of checkAN / decodeAN / checkAndDecodeAN for the streaming engine: codes of CodeT, decoded
// values stored at the runtime width ow (the unencoded width).
template <typename CodeT, int MODE>
struct DecodeMapOp {
    static constexpr int kErr = 1;
    static constexpr bool kMulti = false, kMain = MODE == kCheckDecode, kAcc = false;
    static constexpr bool kBatched = true;
    static constexpr int kOuts = 1;
    using State = uint64_t;  // the decoded value
    uint64_t n;
    const CodeT *codes;
    CodeT a_inv, max_u;
    uint32_t ow;
    uint64_t a_oid;
    void *dec;
    uint64_t dec_cap;
    uint64_t *err;
    uint64_t err_cap;
    uint64_t *d_counts;

    __device__ void out_arrays(void *(&p)[2], uint32_t (&w)[2]) const {
        p[0] = dec;
        w[0] = ow;
        p[1] = nullptr;
        w[1] = 0;
    }
    template <int G>
    __device__ void eval_n(const uint64_t (&row)[G], const bool (&live)[G], State (&st)[G], uint32_t (&errs)[G],
                           uint32_t (&cnt)[G], uint64_t &) const {
        CodeT c[G];
#pragma unroll
        for (int g = 0; g < G; ++g) c[g] = codes[live[g] ? row[g] : 0ull];  // n > 0 whenever a row is live
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const CodeT d = an_mul<CodeT>(c[g], a_inv);
            const bool bad = live[g] && d > max_u;
            errs[g] = bad ? 1u : 0u;
            st[g] = d;
            cnt[g] = (MODE == kCheckDecode && live[g] && !bad) ? 1u : 0u;
            if (MODE == kDecode && live[g]) store_w(dec, row[g], ow, d);
        }
    }
    __device__ uint32_t eval(uint64_t row, State &st, uint32_t &errs, uint64_t &) const {
        const CodeT d = an_mul<CodeT>(codes[row], a_inv);
        const bool bad = d > max_u;
        errs = bad ? 1u : 0u;
        st = d;
        if (MODE == kDecode) store_w(dec, row, ow, d);
        return (MODE == kCheckDecode && !bad) ? 1u : 0u;
    }
    __device__ void emit_main(uint64_t, const State &st, uint64_t pos, uint32_t) const {
        if (pos < dec_cap) store_w(dec, pos, ow, st);
    }
    __device__ void emit_err(int, uint64_t row, const State &, uint64_t pos) const {
        if (pos < err_cap) err[pos] = row * a_oid;
    }
    __device__ void tile_done(uint32_t, uint64_t) const {}
    __device__ void finish(const uint64_t *t) const {
        if (MODE == kCheckDecode) {
            d_counts[0] = t[0];
            d_counts[1] = t[1];
        } else if (MODE == kDecode) {
            d_counts[0] = n;
            d_counts[1] = t[1];
        } else {
            d_counts[0] = t[1];
        }
    }
};

template <typename CodeT, int MODE>
int launch_map_mode(const ahead_col *col, uint64_t a_oid, void *dec, uint64_t dec_cap, uint64_t *err, uint64_t err_cap,
                    uint64_t *d_counts, ahead_hip_ws *ws, hipStream_t s, size_t offset = 0) {
    DecodeMapOp<CodeT, MODE> op{};
    op.n = col->n;
    op.codes = static_cast<const CodeT *>(col->data);
    op.a_inv = static_cast<CodeT>(col->a_inv);
    op.max_u = static_cast<CodeT>(col->max_u);
    op.ow = col->unenc_bits / 8;
    op.a_oid = a_oid;
    op.dec = dec;
    op.dec_cap = dec == nullptr ? 0 : dec_cap;
    op.err = err;
    op.err_cap = err == nullptr ? 0 : err_cap;
    op.d_counts = d_counts;
    return ahead_host::launch_map(op, ws, offset, s);
}

template <int MODE>
int launch_map_width(const ahead_col *col, uint64_t a_oid, void *dec, uint64_t dec_cap, uint64_t *err, uint64_t err_cap,
                     uint64_t *d_counts, ahead_hip_ws *ws, hipStream_t s, size_t offset = 0) {
    switch (col->width) {
        case 2: return launch_map_mode<uint16_t, MODE>(col, a_oid, dec, dec_cap, err, err_cap, d_counts, ws, s, offset);
        case 4: return launch_map_mode<uint32_t, MODE>(col, a_oid, dec, dec_cap, err, err_cap, d_counts, ws, s, offset);
        default: return launch_map_mode<uint64_t, MODE>(col, a_oid, dec, dec_cap, err, err_cap, d_counts, ws, s, offset);
    }
}

int run(int mode, const ahead_col *col, uint64_t a_oid, void *dec, uint64_t dec_cap, uint64_t *err,
        uint64_t err_cap, uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    using namespace ahead_host;
    int st = check_col(col, "decode.col", false, true);
    if (st != AHEAD_OK) return st;
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "decode: d_counts is NULL");
    if (mode != kCheck && dec == nullptr && col->n) return fail(AHEAD_ERR_INVALID, "decode: NULL output");
    if (err == nullptr && err_cap) return fail(AHEAD_ERR_INVALID, "decode: NULL err with capacity");
    if (col->unenc_bits > 8 * col->width) return fail(AHEAD_ERR_INVALID, "decode: unenc_bits wider than the code");
    if (col->unenc_bits != 8 && col->unenc_bits != 16 && col->unenc_bits != 32 && col->unenc_bits != 64)
        return fail(AHEAD_ERR_INVALID, "decode: unenc_bits must be 8, 16, 32 or 64");
    {
        // the streaming engine, unless a check-and-decode output cannot hold a staged entry per row
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (mode == kCheck) return launch_map_width<kCheck>(col, a_oid, nullptr, 0, err, err_cap, d_counts, ws, s);
        if (mode == kDecode) return launch_map_width<kDecode>(col, a_oid, dec, col->n, err, err_cap, d_counts, ws, s);
        if (dec != nullptr && dec_cap >= col->n)
            return launch_map_width<kCheckDecode>(col, a_oid, dec, dec_cap, err, err_cap, d_counts, ws, s);
    }
    const uint32_t tile_rows = 4u * kBlock * (16u / col->width);
    const LbPlan plan = lb_plan(col->n, tile_rows, 2);
    DecodeArgs a{};
    a.codes = col->data;
    a.n = col->n;
    a.a_inv = col->a_inv;
    a.max_u = col->max_u;
    a.a_oid = a_oid;
    a.dec = dec;
    a.dec_cap = dec == nullptr ? 0 : dec_cap;
    a.err = err;
    a.err_cap = err == nullptr ? 0 : err_cap;
    a.d_counts = d_counts;
    if ((st = lb_args(ws, plan, &a.lb)) != AHEAD_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (col->width) {
        case 2:
            st = mode == kCheck ? launch_mode<uint16_t, kCheck>(a, col->unenc_bits, s)
               : mode == kDecode ? launch_mode<uint16_t, kDecode>(a, col->unenc_bits, s)
                                 : launch_mode<uint16_t, kCheckDecode>(a, col->unenc_bits, s);
            break;
        case 4:
            st = mode == kCheck ? launch_mode<uint32_t, kCheck>(a, col->unenc_bits, s)
               : mode == kDecode ? launch_mode<uint32_t, kDecode>(a, col->unenc_bits, s)
                                 : launch_mode<uint32_t, kCheckDecode>(a, col->unenc_bits, s);
            break;
        default:
            st = mode == kCheck ? launch_mode<uint64_t, kCheck>(a, col->unenc_bits, s)
               : mode == kDecode ? launch_mode<uint64_t, kDecode>(a, col->unenc_bits, s)
                                 : launch_mode<uint64_t, kCheckDecode>(a, col->unenc_bits, s);
            break;
    }
    if (st != AHEAD_OK) return st;
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

template <typename InT, typename OutT>
__global__ __launch_bounds__(256) void k_encode(const InT *in, uint64_t n, uint64_t seqbase, OutT *out, OutT a) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
        const OutT v = in ? static_cast<OutT>(in[i]) : static_cast<OutT>(static_cast<uint32_t>(seqbase + i));
        out[i] = an_mul<OutT>(v, a);
    }
}

template <typename InT>
int launch_encode(const ahead_col *col, uint64_t a, uint32_t ow, void *out, hipStream_t s) {
    const uint64_t blocks = (col->n + 255) / 256;
    const dim3 g(static_cast<uint32_t>(blocks < 8192 ? blocks : 8192)), b(256);
    const InT *in = static_cast<const InT *>(col->data);
    switch (ow) {
        case 2: hipLaunchKernelGGL((k_encode<InT, uint16_t>), g, b, 0, s, in, col->n, col->seqbase, static_cast<uint16_t *>(out), static_cast<uint16_t>(a)); break;
        case 4: hipLaunchKernelGGL((k_encode<InT, uint32_t>), g, b, 0, s, in, col->n, col->seqbase, static_cast<uint32_t *>(out), static_cast<uint32_t>(a)); break;
        case 8: hipLaunchKernelGGL((k_encode<InT, uint64_t>), g, b, 0, s, in, col->n, col->seqbase, static_cast<uint64_t *>(out), a); break;
        default: return ahead_host::fail(AHEAD_ERR_INVALID, "encode: out_width must be 2, 4 or 8");
    }
    return AHEAD_OK;
}

}  // namespace

using namespace ahead_host;

namespace ahead_host {
int check_an_ws(const ahead_col *col, uint64_t a_oid, uint64_t *err, uint64_t err_cap, uint64_t *d_counts,
                ahead_hip_ws *ws, void *stream, size_t offset) {
    int st = check_col(col, "check.col", false, true);
    if (st != AHEAD_OK) return st;
    if (d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "check: d_counts is NULL");
    if (err == nullptr && err_cap) return fail(AHEAD_ERR_INVALID, "check: NULL err with capacity");
    return launch_map_width<kCheck>(col, a_oid, nullptr, 0, err, err_cap, d_counts, ws, static_cast<hipStream_t>(stream),
                                    offset);
}
size_t check_an_status_bytes(uint64_t n, uint32_t width) {
    return lb_plan(n, 4u * kBlock * (16u / width), 2).status_bytes;
}
size_t check_an_scratch_bytes(uint64_t n, uint32_t width) {
    switch (width) {
        case 2: return map_scratch_bytes<DecodeMapOp<uint16_t, kCheck>>(n);
        case 4: return map_scratch_bytes<DecodeMapOp<uint32_t, kCheck>>(n);
        default: return map_scratch_bytes<DecodeMapOp<uint64_t, kCheck>>(n);
    }
}
}  // namespace ahead_host

extern "C" int ahead_hip_check_an(const ahead_col *col, uint64_t a_oid, uint64_t *err, uint64_t err_cap,
                                  uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    return run(kCheck, col, a_oid, nullptr, 0, err, err_cap, d_counts, ws, stream);
}

extern "C" int ahead_hip_decode_an(const ahead_col *col, uint64_t a_oid, void *dec, uint64_t *err,
                                   uint64_t err_cap, uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    return run(kDecode, col, a_oid, dec, col ? col->n : 0, err, err_cap, d_counts, ws, stream);
}

extern "C" int ahead_hip_check_and_decode_an(const ahead_col *col, uint64_t a_oid, void *dec,
                                             uint64_t dec_cap, uint64_t *err, uint64_t err_cap,
                                             uint64_t *d_counts, ahead_hip_ws *ws, void *stream) {
    return run(kCheckDecode, col, a_oid, dec, dec_cap, err, err_cap, d_counts, ws, stream);
}

extern "C" int ahead_hip_encode_an(const ahead_col *col, uint64_t a, uint32_t out_width, void *out,
                                   void *stream) {
    int st = check_col(col, "encode.col", true, false);
    if (st != AHEAD_OK) return st;
    if (col->n == 0) return AHEAD_OK;
    if (out == nullptr) return fail(AHEAD_ERR_INVALID, "encode: NULL output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (col->data == nullptr) st = launch_encode<uint32_t>(col, a, out_width, out, s);
    else if (col->width == 1) st = launch_encode<uint8_t>(col, a, out_width, out, s);
    else if (col->width == 2) st = launch_encode<uint16_t>(col, a, out_width, out, s);
    else if (col->width == 4) st = launch_encode<uint32_t>(col, a, out_width, out, s);
    else st = launch_encode<uint64_t>(col, a, out_width, out, s);
    if (st != AHEAD_OK) return st;
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
