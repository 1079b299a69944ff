// select_batch_an.hip -- many selectAN (+ fused check-and-decode + sum) instances in one
// persistent launch.
//
// A batch is built once (ahead_hip_select_batch_create): every job's columns, predicates and
// outputs are fixed, its rows are cut into look-back tiles of `max_subtiles` 16 KiB sub-tiles, and
// the tiles of all jobs are numbered consecutively.  A launch (ahead_hip_select_batch_launch) runs
// `blocks` resident workgroups that take tile tickets from one counter until every tile is done;
// the ticket's job is found by binary search over the jobs' first tiles (held in LDS), and the
// tile runs select_tile for the job's code width and predicate combiner.  Tickets are handed out
// in order, so every job's tiles are taken in order and each job's decoupled look-back sees its
// predecessors already running.  A second kernel adds each job's per-tile decoded sums.
//
// A launch allocates nothing, sets no memory from the host and reads its look-back epoch from
// device memory (the last workgroup to finish advances it), so it can be captured in a hipGraph
// and replayed: every replay recomputes all outputs from the (unchanged) inputs.
#include <vector>

#include "select_tile.hpp"

using namespace ahead_dev;

namespace {

constexpr uint32_t kMaxBatchJobs = 1024;

struct BatchJob {
    SelectArgs a;       // a.lb: the job's status words and tile count; a.partials: its partial sums
    uint32_t width;     // code bytes: 2, 4, 8
    uint32_t comb;      // AHEAD_COMB_*
    uint64_t *d_sum;    // decoded sum of the job (nullable)
};

struct BatchArgs {
    const BatchJob *jobs;
    const uint32_t *first_tile;  // [njobs + 1]: first global tile of every job, then the total
    uint32_t njobs;
    uint32_t total_tiles;
    uint32_t grid;
    WsHeader *hdr;               // ticket, fail, done; pad[0] = current look-back epoch
};

template <typename CodeT>
__device__ __forceinline__ void run_tile(const SelectArgs &a, uint32_t comb, uint32_t tile, SelectShared<8192> &sh) {
    if (comb == AHEAD_COMB_AND) select_tile<CodeT, AHEAD_COMB_AND, true>(a, tile, sh);
    else if (comb == AHEAD_COMB_OR) select_tile<CodeT, AHEAD_COMB_OR, true>(a, tile, sh);
    else select_tile<CodeT, AHEAD_COMB_NONE, true>(a, tile, sh);
}

__global__ __launch_bounds__(kBlock) void k_select_batch(BatchArgs b) {
    __shared__ uint32_t s_ticket;
    __shared__ uint32_t s_epoch;
    __shared__ uint32_t s_first[kMaxBatchJobs + 1];
    __shared__ SelectShared<8192> sh;

    for (uint32_t i = threadIdx.x; i <= b.njobs; i += kBlock) s_first[i] = b.first_tile[i];
    if (threadIdx.x == 0) s_epoch = __hip_atomic_load(&b.hdr->pad[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t epoch = s_epoch;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(&b.hdr->ticket, 1u);
        __syncthreads();
        const uint32_t t = s_ticket;
        if (t >= b.total_tiles) break;
        // the job holding tile t: the last j with s_first[j] <= t
        uint32_t lo = 0, hi = b.njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_first[mid] <= t) lo = mid;
            else hi = mid;
        }
        const uint32_t j = __builtin_amdgcn_readfirstlane(lo);
        SelectArgs a = b.jobs[j].a;
        const uint32_t width = b.jobs[j].width, comb = b.jobs[j].comb;
        a.lb.epoch = epoch;
        const uint32_t local = t - s_first[j];
        if (width == 2) run_tile<uint16_t>(a, comb, local, sh);
        else if (width == 4) run_tile<uint32_t>(a, comb, local, sh);
        else run_tile<uint64_t>(a, comb, local, sh);
        __syncthreads();  // LDS and s_ticket are reused by the next tile
    }
    if (threadIdx.x == 0) {
        const uint32_t d = atomicAdd(&b.hdr->done, 1u);
        if (d == b.grid - 1) {
            // every workgroup has taken its last ticket: reset for the next launch
            __hip_atomic_store(&b.hdr->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&b.hdr->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t next = epoch + 1u >= (1u << 30) ? 1u : epoch + 1u;
            __hip_atomic_store(&b.hdr->pad[0], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// One workgroup per job: adds the job's per-tile decoded sums.
__global__ __launch_bounds__(kBlock) void k_select_batch_sums(BatchArgs b) {
    __shared__ uint64_t s_scan[kWaves];
    const uint32_t j = blockIdx.x;
    const BatchJob &job = b.jobs[j];
    if (job.d_sum == nullptr) return;
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < job.a.lb.num_tiles; i += kBlock) s += job.a.partials[i];
    s = block_sum<uint64_t>(s, s_scan);
    if (threadIdx.x == 0) *job.d_sum = s;
}

bool canon_pred(int op, uint64_t th, uint64_t m, uint64_t *lo, uint64_t *span, uint32_t *neg, uint32_t *valid) {
    th &= m;
    *lo = 0;
    *span = 0;
    *neg = 0;
    *valid = 1;
    switch (op) {
        case AHEAD_LT: if (th == 0) *valid = 0; else *span = th - 1; return true;
        case AHEAD_LE: *span = th; return true;
        case AHEAD_GT: if (th == m) *valid = 0; else { *lo = th + 1; *span = m - th - 1; } return true;
        case AHEAD_GE: *lo = th; *span = m - th; return true;
        case AHEAD_EQ: *lo = th; return true;
        case AHEAD_NE: *lo = th; *neg = 1; return true;
        default: return false;
    }
}

}  // namespace

struct ahead_hip_select_batch {
    int device = 0;
    void *mem = nullptr;   // header | jobs | first_tile | status | partials
    BatchArgs args{};
    uint32_t blocks = 0;
};

using namespace ahead_host;

extern "C" int ahead_hip_select_batch_create(const ahead_select_job *jobs, uint32_t njobs, uint32_t max_subtiles,
                                             uint32_t blocks, ahead_hip_select_batch **out) {
    if (out == nullptr || (jobs == nullptr && njobs)) return fail(AHEAD_ERR_INVALID, "select_batch: NULL argument");
    *out = nullptr;
    if (njobs == 0 || njobs > kMaxBatchJobs) return fail(AHEAD_ERR_INVALID, "select_batch: 1 .. 1024 jobs per batch");
    if (max_subtiles == 0 || max_subtiles > kMaxSubtiles || (max_subtiles & (max_subtiles - 1)))
        return fail(AHEAD_ERR_INVALID, "select_batch: max_subtiles is a power of two in 1 .. 16");
    int dev = 0;
    AHEAD_HIP_TRY(hipGetDevice(&dev));
    if (blocks == 0) {
        int cus = 0;
        AHEAD_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        blocks = static_cast<uint32_t>(cus) * 4u;
    }
    std::vector<BatchJob> hj(njobs);
    std::vector<uint32_t> first(njobs + 1);
    std::vector<size_t> status_off(njobs), partial_off(njobs);
    uint64_t tiles = 0;
    size_t status_bytes = 0, partial_bytes = 0;
    for (uint32_t j = 0; j < njobs; ++j) {
        const ahead_select_job &q = jobs[j];
        int st = check_col(&q.col, "select_batch.col", false, true);
        if (st != AHEAD_OK) return st;
        if (q.d_counts == nullptr) return fail(AHEAD_ERR_INVALID, "select_batch: d_counts is NULL");
        if ((q.oids == nullptr && q.oids_cap) || (q.err == nullptr && q.err_cap))
            return fail(AHEAD_ERR_INVALID, "select_batch: NULL output with nonzero capacity");
        if (q.comb < AHEAD_COMB_NONE || q.comb > AHEAD_COMB_OR) return fail(AHEAD_ERR_INVALID, "select_batch: bad combiner");
        BatchJob &b = hj[j];
        SelectArgs &a = b.a;
        const uint64_t m = width_mask(q.col.width);
        if (!canon_pred(q.op1, q.th1, m, &a.lo1, &a.span1, &a.neg1, &a.valid1))
            return fail(AHEAD_ERR_INVALID, "select_batch: bad op1");
        if (q.comb != AHEAD_COMB_NONE && !canon_pred(q.op2, q.th2, m, &a.lo2, &a.span2, &a.neg2, &a.valid2))
            return fail(AHEAD_ERR_INVALID, "select_batch: bad op2");
        a.codes = q.col.data;
        a.n = q.col.n;
        a.seqbase = q.col.seqbase;
        a.a_inv = q.col.a_inv;
        a.max_u = q.col.max_u;
        a.a_head = q.a_head;
        a.a_oid = q.a_oid;
        a.oids = q.oids;
        a.oids_cap = q.oids ? q.oids_cap : 0;
        a.err = q.err;
        a.err_cap = q.err ? q.err_cap : 0;
        a.d_counts = q.d_counts;
        a.iters = max_subtiles;
        const uint64_t tile_rows = 4ull * kBlock * (16u / q.col.width) * max_subtiles;
        const LbPlan plan = lb_plan(q.col.n, static_cast<uint32_t>(tile_rows), 2);
        a.lb.num_tiles = plan.num_tiles;
        b.width = q.col.width;
        b.comb = static_cast<uint32_t>(q.comb);
        b.d_sum = q.d_sum;
        first[j] = static_cast<uint32_t>(tiles);
        tiles += plan.num_tiles;
        if (tiles >= (1ull << 31)) return fail(AHEAD_ERR_TOO_LARGE, "select_batch: too many tiles");
        status_off[j] = status_bytes;
        status_bytes += plan.status_bytes;
        partial_off[j] = partial_bytes;
        partial_bytes += align_up(static_cast<size_t>(plan.num_tiles) * 8, 256);
    }
    first[njobs] = static_cast<uint32_t>(tiles);
    const size_t o_jobs = align_up(sizeof(WsHeader), 256);
    const size_t o_first = o_jobs + align_up(njobs * sizeof(BatchJob), 256);
    const size_t o_status = o_first + align_up((njobs + 1) * sizeof(uint32_t), 256);
    const size_t o_part = o_status + status_bytes;
    const size_t total = o_part + partial_bytes;

    auto *bt = new ahead_hip_select_batch();
    bt->device = dev;
    bt->blocks = blocks;
    hipError_t e = hipMalloc(&bt->mem, total);
    if (e != hipSuccess) {
        delete bt;
        return e == hipErrorOutOfMemory ? fail(AHEAD_ERR_NOMEM, "select_batch: out of device memory")
                                        : hip_fail(e, "hipMalloc");
    }
    char *base = static_cast<char *>(bt->mem);
    WsHeader *hdr = reinterpret_cast<WsHeader *>(base);
    for (uint32_t j = 0; j < njobs; ++j) {
        hj[j].a.lb.hdr = hdr;
        hj[j].a.lb.status = reinterpret_cast<uint64_t *>(base + o_status + status_off[j]);
        hj[j].a.lb.epoch = 0;  // taken from the header at launch
        hj[j].a.partials = reinterpret_cast<uint64_t *>(base + o_part + partial_off[j]);
    }
    WsHeader h0{};
    h0.pad[0] = 1;  // first launch epoch
    if ((e = hipMemset(bt->mem, 0, total)) != hipSuccess ||
        (e = hipMemcpy(base, &h0, sizeof(h0), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(base + o_jobs, hj.data(), njobs * sizeof(BatchJob), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(base + o_first, first.data(), (njobs + 1) * sizeof(uint32_t), hipMemcpyHostToDevice)) != hipSuccess) {
        (void) hipFree(bt->mem);
        delete bt;
        return hip_fail(e, "select_batch: upload");
    }
    bt->args.jobs = reinterpret_cast<const BatchJob *>(base + o_jobs);
    bt->args.first_tile = reinterpret_cast<const uint32_t *>(base + o_first);
    bt->args.njobs = njobs;
    bt->args.total_tiles = static_cast<uint32_t>(tiles);
    bt->args.grid = blocks;
    bt->args.hdr = hdr;
    *out = bt;
    return AHEAD_OK;
}

extern "C" int ahead_hip_select_batch_launch(ahead_hip_select_batch *b, void *stream) {
    if (b == nullptr) return fail(AHEAD_ERR_INVALID, "select_batch: NULL batch");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_select_batch, dim3(b->blocks), dim3(kBlock), 0, s, b->args);
    hipLaunchKernelGGL(k_select_batch_sums, dim3(b->args.njobs), dim3(kBlock), 0, s, b->args);
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

extern "C" int ahead_hip_select_batch_destroy(ahead_hip_select_batch *b) {
    if (b == nullptr) return AHEAD_OK;
    hipError_t e = hipSuccess;
    if (b->mem) e = hipFree(b->mem);
    delete b;
    if (e != hipSuccess) return hip_fail(e, "hipFree");
    return AHEAD_OK;
}

extern "C" int ahead_hip_select_batch_info(const ahead_hip_select_batch *b, uint32_t *total_tiles, uint32_t *blocks) {
    if (b == nullptr) return fail(AHEAD_ERR_INVALID, "select_batch: NULL batch");
    if (total_tiles) *total_tiles = b->args.total_tiles;
    if (blocks) *blocks = b->blocks;
    return AHEAD_OK;
}
