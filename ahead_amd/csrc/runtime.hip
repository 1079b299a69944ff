// runtime.hip -- error reporting and the workspace of libahead_hip.so.
#include <stdio.h>

#include <new>
#include <string>

#include "host_common.hpp"

namespace ahead_host {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? AHEAD_ERR_NOMEM : AHEAD_ERR_HIP;
}

int ws_ensure(ahead_hip_ws *ws, size_t bytes) {
    if (ws == nullptr) return fail(AHEAD_ERR_INVALID, "workspace is NULL");
    const size_t need = kHeaderBytes + align_up(bytes, 256);
    if (ws->base != nullptr && ws->bytes >= need) return AHEAD_OK;
    size_t grow = ws->bytes + ws->bytes / 2;
    const size_t target = need > grow ? need : grow;
    int cur = 0;
    AHEAD_HIP_TRY(hipGetDevice(&cur));
    if (cur != ws->device) AHEAD_HIP_TRY(hipSetDevice(ws->device));
    if (ws->base != nullptr) {
        AHEAD_HIP_TRY(hipFree(ws->base));
        ws->base = nullptr;
        ws->bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, target);
    if (e != hipSuccess) {
        if (cur != ws->device) (void)hipSetDevice(cur);
        return hip_fail(e, "hipMalloc(workspace)");
    }
    e = hipMemset(p, 0, target);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (cur != ws->device) (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "hipMemset(workspace)");
    }
    ws->base = p;
    ws->bytes = target;
    return AHEAD_OK;
}

int ws_ensure_status(ahead_hip_ws *ws, size_t bytes) {
    if (ws == nullptr) return fail(AHEAD_ERR_INVALID, "workspace is NULL");
    const size_t need = align_up(bytes > 0 ? bytes : 1, 256);
    if (ws->status != nullptr && ws->status_bytes >= need) return AHEAD_OK;
    const size_t grow = ws->status_bytes + ws->status_bytes / 2;
    const size_t target = need > grow ? need : grow;
    int cur = 0;
    AHEAD_HIP_TRY(hipGetDevice(&cur));
    if (cur != ws->device) AHEAD_HIP_TRY(hipSetDevice(ws->device));
    if (ws->status != nullptr) {
        AHEAD_HIP_TRY(hipFree(ws->status));
        ws->status = nullptr;
        ws->status_bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, target);
    if (e != hipSuccess) {
        if (cur != ws->device) (void)hipSetDevice(cur);
        return hip_fail(e, "hipMalloc(workspace status)");
    }
    e = hipMemset(p, 0, target);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (cur != ws->device) (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "hipMemset(workspace status)");
    }
    ws->status = p;
    ws->status_bytes = target;
    return AHEAD_OK;
}

int ws_next_epoch(ahead_hip_ws *ws, uint32_t *epoch) {
    if (ws->epoch + 1 >= (1u << 30)) {
        // Status words of every earlier launch carry an epoch below 2^30; clear them all.
        if (ws->status != nullptr) {
            AHEAD_HIP_TRY(hipMemset(ws->status, 0, ws->status_bytes));
            AHEAD_HIP_TRY(hipDeviceSynchronize());
        }
        ws->epoch = 0;
    }
    *epoch = ++ws->epoch;
    return AHEAD_OK;
}

int lb_args(ahead_hip_ws *ws, const LbPlan &p, ahead_dev::LbArgs *out) {
    int st = ws_ensure_status(ws, p.status_bytes);
    if (st != AHEAD_OK) return st;
    uint32_t epoch = 0;
    if ((st = ws_next_epoch(ws, &epoch)) != AHEAD_OK) return st;
    out->hdr = ws_header(ws);
    out->status = static_cast<uint64_t *>(ws->status);
    out->num_tiles = p.num_tiles;
    out->epoch = epoch;
    return AHEAD_OK;
}

__global__ void k_fill64(uint64_t *p, uint64_t n, uint64_t v) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) p[i] = v;
}
__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) p[i] = v;
}

static uint32_t fill_grid(uint64_t n) {
    const uint64_t b = (n + 255) / 256;
    return static_cast<uint32_t>(b == 0 ? 1 : (b < 2048 ? b : 2048));
}

int fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s) {
    if (n == 0) return AHEAD_OK;
    hipLaunchKernelGGL(k_fill64, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return AHEAD_OK;
    hipLaunchKernelGGL(k_fill32, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

static Tiling g_tiling;

Tiling tiling() { return g_tiling; }

uint32_t subtiles_per_tile(uint64_t subs) {
    const Tiling t = g_tiling;
    uint32_t it = 1;
    while (2 * it <= t.max_subtiles && subs >= 2ull * it * t.min_tiles) it *= 2;
    return it;
}

int check_col(const ahead_col *c, const char *name, bool allow_void, bool need_encoded) {
    if (c == nullptr) return fail(AHEAD_ERR_INVALID, std::string(name) + ": NULL column");
    if (c->n > 0xFFFFFFFFull)
        return fail(AHEAD_ERR_TOO_LARGE, std::string(name) + ": more than 2^32-1 rows");
    if (c->data == nullptr && !allow_void && c->n > 0)
        return fail(AHEAD_ERR_INVALID, std::string(name) + ": void column not allowed here");
    if (c->data != nullptr && c->width != 1 && c->width != 2 && c->width != 4 && c->width != 8)
        return fail(AHEAD_ERR_INVALID, std::string(name) + ": width must be 1, 2, 4 or 8");
    if (need_encoded && !c->encoded)
        return fail(AHEAD_ERR_INVALID, std::string(name) + ": must be an AN-encoded column");
    if (c->encoded && c->data != nullptr && c->width < 2)
        return fail(AHEAD_ERR_INVALID, std::string(name) + ": AN codes are 2, 4 or 8 bytes");
    return AHEAD_OK;
}

}  // namespace ahead_host

using namespace ahead_host;

extern "C" {

const char *ahead_hip_last_error(void) { return g_last_error.c_str(); }

const char *ahead_hip_version(void) { return "ahead_hip 0.1 gfx950"; }

int ahead_hip_ws_create(int device, ahead_hip_ws **ws) {
    if (ws == nullptr) return fail(AHEAD_ERR_INVALID, "ws out-pointer is NULL");
    int count = 0;
    AHEAD_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(AHEAD_ERR_INVALID, "no such device");
    ahead_hip_ws *w = new (std::nothrow) ahead_hip_ws();
    if (w == nullptr) return fail(AHEAD_ERR_NOMEM, "host allocation failed");
    w->device = device;
    int st = ws_ensure(w, 1 << 20);
    if (st == AHEAD_OK) st = ws_ensure_status(w, 1 << 16);
    if (st != AHEAD_OK) {
        (void)ahead_hip_ws_destroy(w);
        return st;
    }
    *ws = w;
    return AHEAD_OK;
}

int ahead_hip_ws_destroy(ahead_hip_ws *ws) {
    if (ws == nullptr) return AHEAD_OK;
    hipError_t e = hipSuccess;
    if (ws->base != nullptr) e = hipFree(ws->base);
    if (ws->status != nullptr) {
        const hipError_t e2 = hipFree(ws->status);
        if (e == hipSuccess) e = e2;
    }
    delete ws;
    if (e != hipSuccess) return hip_fail(e, "hipFree(workspace)");
    return AHEAD_OK;
}

int ahead_hip_ws_reserve(ahead_hip_ws *ws, size_t bytes) {
    const int st = ws_ensure(ws, bytes);
    return st != AHEAD_OK ? st : ws_ensure_status(ws, bytes);
}

size_t ahead_hip_ws_bytes(const ahead_hip_ws *ws) {
    return ws == nullptr ? 0 : ws->bytes + ws->status_bytes;
}

int ahead_hip_set_tiling(uint32_t max_subtiles, uint32_t min_tiles) {
    if (max_subtiles < 1 || max_subtiles > kMaxSubtiles || (max_subtiles & (max_subtiles - 1)) != 0)
        return fail(AHEAD_ERR_INVALID, "set_tiling: max_subtiles must be a power of two in [1, 16]");
    if (min_tiles < 1) return fail(AHEAD_ERR_INVALID, "set_tiling: min_tiles must be >= 1");
    g_tiling.max_subtiles = max_subtiles;
    g_tiling.min_tiles = min_tiles;
    return AHEAD_OK;
}

int ahead_hip_get_tiling(uint32_t *max_subtiles, uint32_t *min_tiles) {
    if (max_subtiles) *max_subtiles = g_tiling.max_subtiles;
    if (min_tiles) *min_tiles = g_tiling.min_tiles;
    return AHEAD_OK;
}

}  // extern "C"
