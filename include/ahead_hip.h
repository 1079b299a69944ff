/*
 * ahead_hip.h -- C ABI of the MI355X (gfx950) AN-hardened column operators.
 *
 * This is the drop-in boundary under the reference's operator templates
 * (include/column_operators/OperatorsAN.hpp:33-298): every template of that header that lies
 * on the hardened hot path has one extern "C" entry point here, and the C++ templates in
 * include/ahead/column_operators/OperatorsAN.hpp marshal BAT<Head,Tail> arguments to them.
 *
 * Conventions
 *  - All column and output pointers are DEVICE pointers (HBM).  No entry point copies data to
 *    or from the host, and none allocates except through the workspace (ahead_hip_ws), which
 *    grows on demand.
 *  - Every entry point is asynchronous on `stream` (a hipStream_t; NULL = the null stream).
 *  - Variable-length outputs are (pointer, capacity) pairs.  The operator writes element k of
 *    an output only if k < capacity, and always writes the full element count to d_counts[]
 *    (device memory).  A count larger than the capacity means "truncated": the caller sizes
 *    the buffer from the count and calls again (operators are pure functions of their inputs).
 *  - Error-position vectors (AN_indicator_vector, include/column_operators/ANbase.hpp:32) are
 *    uint64 arrays of pos * A_OID, in the order the reference appends them.
 *  - Return value: AHEAD_OK or an ahead_status error code; ahead_hip_last_error() gives text.
 *  - A workspace must not be used by two streams at the same time.
 */
#ifndef AHEAD_HIP_H
#define AHEAD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AHEAD_OK = 0,
    AHEAD_ERR_INVALID = 1,       /* bad argument (width, operator code, NULL pointer ...)      */
    AHEAD_ERR_SIZE_MISMATCH = 2, /* the reference throws std::runtime_error here               */
    AHEAD_ERR_HIP = 3,           /* HIP runtime error                                          */
    AHEAD_ERR_NOMEM = 4,         /* device allocation failed (std::bad_alloc in the reference) */
    AHEAD_ERR_TOO_LARGE = 5,     /* more than 2^32-1 rows (BAT sizes are oid_t, uint32)        */
    AHEAD_ERR_DEVICE = 6         /* a kernel reported an internal failure                      */
} ahead_status;

/* Column descriptor: one BAT side (ColumnDescriptor + ColumnMetaData,
 * include/column_storage/ColumnDescriptor.hpp:43-102, ColumnMetaData.hpp:150-210). */
typedef struct {
    const void *data;    /* device array, or NULL for a void (virtual dense OID) column   */
    uint64_t n;          /* number of elements                                            */
    uint32_t width;      /* bytes per element: 1, 2, 4, 8 (void: 4, oid_t)                */
    uint32_t encoded;    /* 1: AN-encoded type (v2_res*_t), 0: plain                      */
    uint32_t unenc_bits; /* bits of TypeMap<T>::v2_base_t::type_t (8, 16, 32, 64)         */
    uint32_t reserved;
    uint64_t a;          /* ColumnMetaData::AN_A                                          */
    uint64_t a_inv;      /* ColumnMetaData::AN_Ainv                                       */
    uint64_t max_u;      /* ColumnMetaData::AN_unencMaxU                                  */
    uint64_t seqbase;    /* ColumnMetaData::seqbase (void columns)                        */
} ahead_col;

/* Predicate and combiner codes of selectAN (std::less ... std::not_equal_to; ahead::and_is,
 * ahead::or_is from include/column_operators/functors.hpp). */
enum { AHEAD_LT = 0, AHEAD_LE = 1, AHEAD_GT = 2, AHEAD_GE = 3, AHEAD_EQ = 4, AHEAD_NE = 5 };
enum { AHEAD_COMB_NONE = 0, AHEAD_COMB_AND = 1, AHEAD_COMB_OR = 2 };

typedef struct ahead_hip_ws ahead_hip_ws;

/* A device output buffer: entries at index >= cap are not written (see Conventions). */
typedef struct {
    void *data;
    uint64_t cap;
} ahead_buf;

/* ---- runtime --------------------------------------------------------------------------------- */
const char *ahead_hip_last_error(void);
const char *ahead_hip_version(void);
/* Creates a workspace on `device` (device scratch for scan state, hash tables, partials). */
int ahead_hip_ws_create(int device, ahead_hip_ws **ws);
int ahead_hip_ws_destroy(ahead_hip_ws *ws);
/* Grows the workspace now (so later calls do not allocate, e.g. under stream capture): its
 * scratch to at least `bytes` and its scan-state area to at least `bytes` too.  A reservation
 * covers every call whose scratch plus scan state fits in `bytes`. */
int ahead_hip_ws_reserve(ahead_hip_ws *ws, size_t bytes);
/* Device bytes the workspace holds (scratch + scan state). */
size_t ahead_hip_ws_bytes(const ahead_hip_ws *ws);

/* Tiling of the streaming operators: a look-back tile covers up to `max_subtiles` sub-tiles of
 * 16 KiB, as many as keep at least `min_tiles` tiles in the grid.  Process-wide; defaults 16 and
 * 512.  Used by the autotuner and by tests that need multi-sub-tile tiles at small sizes.      */
int ahead_hip_set_tiling(uint32_t max_subtiles, uint32_t min_tiles);
int ahead_hip_get_tiling(uint32_t *max_subtiles, uint32_t *min_tiles);

/* ---- selectAN ---------------------------------------------------------------------------------
 * Replaces ahead::bat::ops::scalar::selectAN<Op> and selectAN<Op1,Op2,OpCombine>
 * (OperatorsAN.hpp:163-195; lib/column_operators/AN/selectAN_scalar.tcc:66-95, 149-181).
 * col: the encoded tail of a BAT<void, v2_res*_t>.  th1/th2 are ENCODED thresholds.
 * Outputs: oids[] = (uint32)(seqbase+pos) * a_head, err[] = pos * a_oid (every row checked),
 * d_counts[0] = #oids, d_counts[1] = #err.  d_sum (nullable): the sum of decoded values of
 * matching, uncorrupted rows (select + check-and-decode + sum, fused), mod 2^64.            */
int ahead_hip_select_an(const ahead_col *col, int op1, uint64_t th1, int op2, uint64_t th2,
                        int comb, uint64_t a_head, uint64_t a_oid, uint64_t *oids,
                        uint64_t oids_cap, uint64_t *err, uint64_t err_cap, uint64_t *d_counts,
                        uint64_t *d_sum, ahead_hip_ws *ws, void *stream);

/* ---- selectAN without allocation ---------------------------------------------------------------
 * No counterpart in the reference's API: for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as ahead_hip_select_an, without a workspace: the scan state (tile ticket,
 * per-tile partial sums, look-back status words) lives in a fixed 4 MiB device array of the
 * library, one per GPU, held until the process ends.  The call allocates nothing.  A column whose
 * look-back tiles do not fit the array is rejected with AHEAD_ERR_TOO_LARGE (with the default
 * tiling that never happens below 2^32 rows).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, or two host threads) race and must not be issued. */
int ahead_hip_select_an_noalloc(const ahead_col *col, int op1, uint64_t th1, int op2, uint64_t th2,
                                int comb, uint64_t a_head, uint64_t a_oid, uint64_t *oids,
                                uint64_t oids_cap, uint64_t *err, uint64_t err_cap,
                                uint64_t *d_counts, uint64_t *d_sum, void *stream);

/* ---- batched selectAN ---------------------------------------------------------------------------
 * No counterpart in the reference's API: many independent selectAN (+ fused decoded sum)
 * instances, each with the semantics of ahead_hip_select_an, run by one persistent launch.  A
 * batch is created once from its jobs (columns, predicates and outputs are fixed; creation
 * allocates and uploads, synchronously) and launched any number of times; a launch allocates
 * nothing and sets no memory from the host, so it may be captured in a hipGraph and replayed.
 * Tuning: max_subtiles (1, 2, 4, 8, 16) sub-tiles of 16 KiB per look-back tile, `blocks`
 * resident workgroups (0 = 4 per CU).  At most 1024 jobs per batch.  Launches of one batch
 * must not overlap.                                                                           */
typedef struct {
    ahead_col col;         /* encoded tail of a BAT<void, v2_res*_t>                         */
    int32_t op1, op2, comb, reserved;
    uint64_t th1, th2;     /* encoded thresholds                                              */
    uint64_t a_head, a_oid;
    uint64_t *oids;
    uint64_t oids_cap;
    uint64_t *err;
    uint64_t err_cap;
    uint64_t *d_counts;    /* device, 2 entries: #oids, #err                                  */
    uint64_t *d_sum;       /* device, decoded sum of matching valid rows (nullable)          */
} ahead_select_job;

typedef struct ahead_hip_select_batch ahead_hip_select_batch;
int ahead_hip_select_batch_create(const ahead_select_job *jobs, uint32_t njobs, uint32_t max_subtiles,
                                  uint32_t blocks, ahead_hip_select_batch **out);
int ahead_hip_select_batch_launch(ahead_hip_select_batch *batch, void *stream);
int ahead_hip_select_batch_destroy(ahead_hip_select_batch *batch);
int ahead_hip_select_batch_info(const ahead_hip_select_batch *batch, uint32_t *total_tiles, uint32_t *blocks);

/* ---- encode / check / decode ------------------------------------------------------------------
 * encodeAN (OperatorsAN.hpp:137-141; encdecAN_scalar.tcc:44-67): out[i] = (code)v_i * a,
 * out_width bytes per code.                                                                  */
int ahead_hip_encode_an(const ahead_col *col, uint64_t a, uint32_t out_width, void *out,
                        void *stream);
/* checkAN (OperatorsAN.hpp:143-149; encdecAN_scalar.tcc:69-94): d_counts[0] = #err */
int ahead_hip_check_an(const ahead_col *col, uint64_t a_oid, uint64_t *err, uint64_t err_cap,
                       uint64_t *d_counts, ahead_hip_ws *ws, void *stream);
/* decodeAN (OperatorsAN.hpp:151-155; encdecAN_scalar.tcc:96-126): all rows decoded into
 * dec[] (unenc_bits/8 bytes each), d_counts[0] = n, d_counts[1] = #err                       */
int ahead_hip_decode_an(const ahead_col *col, uint64_t a_oid, void *dec, uint64_t *err,
                        uint64_t err_cap, uint64_t *d_counts, ahead_hip_ws *ws, void *stream);
/* checkAndDecodeAN<void,Tail> (OperatorsAN.hpp:157-161; encdecAN_scalar.tcc:185-210):
 * corrupt rows dropped; d_counts[0] = #dec, d_counts[1] = #err                               */
int ahead_hip_check_and_decode_an(const ahead_col *col, uint64_t a_oid, void *dec,
                                  uint64_t dec_cap, uint64_t *err, uint64_t err_cap,
                                  uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* checkAndDecodeAN<Head, Tail> with a non-void head (OperatorsAN.hpp:157-161;
 * encdecAN_scalar.tcc:129-176): pairs decoded to wout_h / wout_t bytes, kept only if both sides
 * are valid; err_h / err_t hold the positions of corrupt heads / tails.
 * d_counts[0..2] = #pairs, #err_h, #err_t.                                                    */
int ahead_hip_check_and_decode2_an(const ahead_col *head, const ahead_col *tail, uint64_t a_oid,
                                   uint32_t wout_h, uint32_t wout_t, void *dec_h, void *dec_t,
                                   uint64_t dec_cap, ahead_buf err_h, ahead_buf err_t,
                                   uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* ---- joins -----------------------------------------------------------------------------------
 * fetchjoinAN (OperatorsAN.hpp:92-106; matchjoinAN.tcc:312-353).  oids: tail of a
 * BAT<void, v2_resoid_t>; col: tail of BAT<void, T2>.  out: kept codes (TypeMap<T2>::
 * v2_encoded_t width), re-encoded with a_reenc * inv(A_col) when reenc != 0.  err1: i * a_oid
 * for OIDs beyond the column (t >= |col|), err2: t * a_oid for corrupt fetched codes (dropped).
 * d_counts[0..2] = #out, #err1, #err2.                                                       */
int ahead_hip_fetchjoin_an(const ahead_col *oids, const ahead_col *col, int reenc, uint64_t a_reenc,
                           uint64_t a_oid, ahead_buf out, ahead_buf err1, ahead_buf err2,
                           uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* matchjoinAN, positional form <H1, v2_resoid_t, v2_void_t, T2> (OperatorsAN.hpp:33-58;
 * matchjoinAN.tcc:240-301).  Pairs (h1 [* f_h1], t2[pos] [* f_t2]) with pos = t1 * A^-1;
 * err1: corrupt h1 (pos1), err2: pos >= |t2| (pos1), err4: corrupt t2 (pos2, dropped).
 * d_counts[0..3] = #pairs, #err1, #err2, #err4.                                              */
int ahead_hip_matchjoin_pos_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *t2,
                               int reenc_h, uint64_t f_h1, uint32_t wout_h, int reenc_t,
                               uint64_t f_t2, uint32_t wout_t, uint64_t a_oid, ahead_buf out_h,
                               ahead_buf out_t, ahead_buf err1, ahead_buf err2, ahead_buf err4,
                               uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* matchjoinAN, merge form <H1, T1, H2, T2> (OperatorsAN.hpp:33-58; matchjoinAN.tcc:46-143):
 * a two-cursor walk over decoded t1 (list 1) and h2 (list 2), both ascending, emitting
 * (h1 [* f_h1], t2 [* f_t2]) on equal keys.  e1..e4 check h1, t1, h2, t2 where the reference's
 * loops visit the rows (a list-2 row skipped after the top of an iteration is reported twice;
 * matched list-1 rows are not checked).  Inputs that are not ascending are walked exactly as the
 * reference does, on a slower single-thread path.  d_counts[0..4] = #pairs, #e1, #e2, #e3, #e4. */
int ahead_hip_matchjoin_merge_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2,
                                 const ahead_col *t2, int reenc_h, uint64_t f_h1, uint32_t wout_h,
                                 int reenc_t, uint64_t f_t2, uint32_t wout_t, uint64_t a_oid,
                                 ahead_buf out_h, ahead_buf out_t, ahead_buf e1, ahead_buf e2,
                                 ahead_buf e3, ahead_buf e4, uint64_t *d_counts, ahead_hip_ws *ws,
                                 void *stream);

/* matchjoinAN, form <H1, v2_void_t, v2_void_t, T2> (OperatorsAN.hpp:33-58;
 * matchjoinAN.tcc:146-224): the virtual OIDs of list 1's tail (t1->seqbase + i) and list 2's
 * head (h2->seqbase + j) are aligned by skipping the leading rows of the side that starts lower,
 * then rows are paired one to one.  e1 checks h1 over the rows visited on list 1, e4 checks t2
 * over those visited on list 2.  d_counts[0..2] = #pairs, #e1, #e4.                         */
int ahead_hip_matchjoin_seq_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2,
                               const ahead_col *t2, int reenc_h, uint64_t f_h1, uint32_t wout_h,
                               int reenc_t, uint64_t f_t2, uint32_t wout_t, uint64_t a_oid,
                               ahead_buf out_h, ahead_buf out_t, ahead_buf e1, ahead_buf e4,
                               uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* hashjoinAN (OperatorsAN.hpp:60-90; hashjoinAN.tcc:77-222, :552-585) with the unencoded-key
 * hash map.  build_right != 0 is hash_side_t::right (the default).  Pairs (h1, t2) in probe
 * order, build rows in insertion order per key; f_h1 / f_t2 are the re-encoding factors
 * mulIfEncoded(A_R, A^-1) of the re-encoding overloads.  e1..e4 check h1, t1, h2, t2 by scan
 * position.  Build keys up to 64 bits, any number of build rows per key (lists of more than
 * 8192 rows are merged in the workspace, 4 bytes per build row).
 * d_counts[0..5] = #pairs, #e1, #e2, #e3, #e4, status (0 ok, 1 = the build table overflowed,
 * an internal limit).                                                                          */
int ahead_hip_hashjoin_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2,
                          const ahead_col *t2, int build_right, int reenc_h, uint64_t f_h1,
                          uint32_t wout_h, int reenc_t, uint64_t f_t2, uint32_t wout_t,
                          uint64_t a_oid, ahead_buf out_h, ahead_buf out_t, ahead_buf e1,
                          ahead_buf e2, ahead_buf e3, ahead_buf e4, uint64_t *d_counts,
                          ahead_hip_ws *ws, void *stream);

/* ---- grouping and aggregation -----------------------------------------------------------------
 * groupbyAN unary (OperatorsAN.hpp:108-113; groupbyAN.tcc:51-94): voidtorgid[i] = gid * a_oid
 * (n entries), vgidtoroid[g] = head(first row of g) * a_oid, ids by first occurrence.
 * table_bits: log2 of the dictionary slots (0 = automatic: twice the rows, at most 2^22 slots
 * and at most twice the number of keys the key's width allows).
 * d_counts[0..5] = n, #groups, #e_head, #e_tail, 0, status (0 ok, 1 = dictionary overflow:
 * call again with a larger table_bits).                                                      */
int ahead_hip_groupby_an(const ahead_col *head, const ahead_col *tail, uint64_t a_oid,
                         uint32_t table_bits, uint64_t *voidtorgid, ahead_buf vgidtoroid,
                         ahead_buf e_head, ahead_buf e_tail, uint64_t *d_counts,
                         ahead_hip_ws *ws, void *stream);
/* groupbyAN binary (OperatorsAN.hpp:115-123; groupbyAN.tcc:99-165): key
 * t * num_groups + g in larger_bits (Tail::v2_larger_t) width; repeated keys return the RGID
 * as stored in the reference's oid_t dictionary.  d_counts[4] = #e_group.                    */
int ahead_hip_groupby2_an(const ahead_col *head, const ahead_col *tail, const ahead_col *grouping,
                          uint64_t num_groups, uint32_t larger_bits, uint64_t a_oid,
                          uint32_t table_bits, uint64_t *voidtorgid, ahead_buf vgidtoroid,
                          ahead_buf e_head, ahead_buf e_tail, ahead_buf e_group,
                          uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* aggregate_sum_groupedAN (OperatorsAN.hpp:125-134; aggregateAN.tcc:41-119): sums[g] =
 * sum of tail codes * (a_result * inv(A_tail)) mod 2^64 over rows with a valid group id
 * (corrupt tails are added and reported).  A decoded group id above maxU or >= num_groups is
 * reported in e_group and skipped.  d_counts[0..3] = num_groups, #e_head, #e_tail, #e_group. */
int ahead_hip_aggregate_sum_grouped_an(const ahead_col *head, const ahead_col *tail,
                                       const ahead_col *grouping, uint64_t num_groups,
                                       uint64_t a_result, uint64_t a_oid, uint64_t *sums,
                                       ahead_buf e_head, ahead_buf e_tail, ahead_buf e_group,
                                       uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* aggregate_mul_sumAN (OperatorsAN.hpp:206-215; aggregateAN_scalar.tcc:58-131):
 * *d_total = (init + sum t1*t2) * (ra? * inv(A_t1) * inv(A_t2)) mod 2^64 over the shorter input.
 * d_counts[0..2] = 1, #err1, #err2.                                                         */
int ahead_hip_aggregate_mul_sum_an(const ahead_col *t1, const ahead_col *t2, uint64_t init,
                                   int result_encoded, uint64_t ra, uint64_t a_oid,
                                   uint64_t *d_total, ahead_buf err1, ahead_buf err2,
                                   uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* arithmeticAN (OperatorsAN.hpp:197-204; arithmeticAN_scalar.tcc:60-112): out[i] =
 * op(d1_i * a_result, d2_i * a_result) in result_width bytes; op 0 add, 1 sub, 2 mul, 3 div
 * (division by zero yields 0).  d_counts[0..4] = n, #e_h1, #e_t1, #e_h2, #e_t2.              */
int ahead_hip_arithmetic_an(const ahead_col *h1, const ahead_col *t1, const ahead_col *h2,
                            const ahead_col *t2, int op, uint64_t a_result, uint32_t result_width,
                            uint64_t a_oid, void *out, ahead_buf eh1, ahead_buf et1, ahead_buf eh2,
                            ahead_buf et2, uint64_t *d_counts, ahead_hip_ws *ws, void *stream);

/* ---- fused star join over one dimension ------------------------------------------------------
 * Replaces, for one dimension with two 8-bit group attributes, the hardened operator chain of
 * tools/ssbm/ssbm-q21_continuous.cpp:97-181:
 *   hashjoinAN(fact_fk, reverse(dim_key))                       (OperatorsAN.hpp:79-90)
 *   fetchjoinAN(dim OIDs, dim_a1), fetchjoinAN(dim OIDs, dim_a2), fetchjoinAN(fact OIDs, fact_m)
 *                                                               (OperatorsAN.hpp:92-106)
 *   groupbyAN(a1), groupbyAN(a2, g1, |g1|)                      (OperatorsAN.hpp:108-123)
 *   aggregate_sum_groupedAN<v2_resbigint_t>(m, g2, |g2|)        (OperatorsAN.hpp:125-134)
 *   fetchjoinAN(VGIDtoROID, a1 / a2) for the group keys
 * with the row-consistent policy: a joined row whose fetched attribute or measure is corrupt is
 * reported by that fetch and dropped from group-by and sum.  Fused path: 64-bit FK / key /
 * measure codes, restiny (16-bit code, 8-bit value) attributes, unique build keys.
 * d_counts[0] = #groups, [1] = #e_fk, [2] = #e_dk, [3] = #e_a1, [4] = #e_a2, [5] = #e_m,
 * [6] = #e_grp, [7] = status (0 ok; 1 = duplicate build keys and 2 = more than 8192 groups: the
 * fused path does not apply and the caller runs the per-operator chain), [8] = kept rows.      */
typedef struct {
    ahead_col fact_fk;    /* BAT<void, resint> foreign-key codes                 */
    ahead_col fact_m;     /* BAT<void, resint> measure codes                     */
    ahead_col dim_key;    /* BAT<void, resint> dimension key codes (build side)  */
    ahead_col dim_a1;     /* BAT<void, restiny> first group attribute            */
    ahead_col dim_a2;     /* BAT<void, restiny> second group attribute           */
    uint64_t a_oid_join;  /* A of the re-encoded join OIDs                       */
    uint64_t a_oid_err;   /* A of error positions (16089 by default)             */
    uint64_t a_oid_grp;   /* A of group ids / VGIDtoROID (65117 by default)      */
    uint64_t a_result;    /* A of the grouped sums (65117 by default)            */
} ahead_star1_args;

typedef struct {
    uint64_t *sums, *groid, *key1, *key2; /* per group, first-occurrence order      */
    uint64_t groups_cap;
    uint64_t *e_fk, *e_dk, *e_a1, *e_a2, *e_m;
    uint64_t e_fk_cap, e_dk_cap, e_a1_cap, e_a2_cap, e_m_cap;
    uint64_t *d_counts;                   /* device, 9 entries                      */
} ahead_star1_out;

int ahead_hip_star1_pipeline(const ahead_star1_args *in, const ahead_star1_out *out,
                             ahead_hip_ws *ws, void *stream);
/* The two halves of ahead_hip_star1_pipeline: build (dimension hash table + dimension-key error
 * vector, d_counts[2]) and probe (fact scan + group finalize).  A built table stays valid for any
 * number of probes with the same dimension as long as no other operator uses the workspace.  */
int ahead_hip_star1_build(const ahead_star1_args *in, const ahead_star1_out *out,
                          ahead_hip_ws *ws, void *stream);
int ahead_hip_star1_probe(const ahead_star1_args *in, const ahead_star1_out *out,
                          ahead_hip_ws *ws, void *stream);

/* ---- fused star query ---------------------------------------------------------------------------
 * Replaces the fact side of one hardened SSB operator chain (tools/ssbm/ssbm-q11 .. q43_continuous.cpp)
 * by one pass over the fact rows: the fact selectAN calls (Q1.x), per dimension the hashjoinAN probe
 * (ssbm-q21_continuous.cpp:95) or positional matchjoinAN + hashjoinAN (:102-115), the fetchjoinAN
 * calls of measures and group attributes, arithmeticAN<sub> (ssbm-q41_continuous.cpp:195),
 * groupbyAN over the attributes and aggregate_sum_groupedAN / aggregate_mul_sumAN
 * (OperatorsAN.hpp:79-134, 280-298).  The dimension-side selectAN calls (and their error vectors)
 * stay with the caller: `sel` is their OID list.
 * Row semantics, in fact order, with the row-consistent policy for corrupt fetched codes
 * (reported, row dropped from the sums; the reference's chain throws there, groupbyAN.tcc:105-107):
 * fact predicates on raw codes (every row checked); measures flagged in meas_pre_mask checked at
 * the rows that pass them; dimensions in chain order (a corrupt FK is reported and ends the row,
 * except for AHEAD_SQ_PROBE dimensions, which probe the truncated key anyway, hashjoinAN.tcc:199);
 * joined rows check their other measures, re-fetch AHEAD_SQ_REFETCH keys and fetch their group
 * attributes; kept rows add SUM d0 | SUB d0-d1 | MULSUM d0*d1 (decoded, mod 2^64) to their group.
 * Groups are numbered by their first kept fact row.  Outputs: sums (x a_result), the first row
 * and the attribute codes of every group; error vectors by kind (positions x a_oid_err: fact rows,
 * dimension rows for attribute fetches).  d_counts: [0] groups, [1] kept rows, [2] status,
 * [3] error events, [4 + kind] entries of each error vector.  A status other than OK means the
 * fused form does not apply (a corrupt or repeated dimension key, a selected OID out of range, a
 * key span above 4 x rows + 1024, more than AHEAD_SQ_MAX_GROUPS groups, more than
 * AHEAD_SQ_MAX_EVENTS error events) and the caller runs the operator chain.  The call waits for
 * the dimension statistics (a few hundred bytes) before it plans and launches the fact scan. */
#define AHEAD_SQ_MAX_DIMS 4
#define AHEAD_SQ_MAX_PREDS 2
#define AHEAD_SQ_MAX_MEAS 2
#define AHEAD_SQ_MAX_GROUPS 2048
#define AHEAD_SQ_MAX_EVENTS 8192
#define AHEAD_SQ_KINDS 16
enum { AHEAD_SQ_SUM = 0, AHEAD_SQ_SUB = 1, AHEAD_SQ_MULSUM = 2 };
enum { AHEAD_SQ_PROBE = 1, AHEAD_SQ_REFETCH = 2 };
/* error vector kinds: + predicate / measure / dimension index (chain order) */
enum { AHEAD_SQ_E_PRED = 0, AHEAD_SQ_E_MEAS = 2, AHEAD_SQ_E_FK = 4, AHEAD_SQ_E_REFETCH = 8,
       AHEAD_SQ_E_ATTR = 12 };
enum { AHEAD_SQ_ST_OK = 0, AHEAD_SQ_ST_BAD_OID = 1, AHEAD_SQ_ST_BAD_KEY = 2, AHEAD_SQ_ST_DUP_KEY = 3,
       AHEAD_SQ_ST_EVENTS = 4, AHEAD_SQ_ST_KEY_SPAN = 5, AHEAD_SQ_ST_GROUPS = 6 };

/* Key index of a dimension: the validity, uniqueness and range of the keys of its key column and
 * the row of every key, built once for a resident dimension and then passed to any number of star
 * queries over it (which then skip those checks).  The key column must not change while an index
 * built from it is in use; create/destroy allocate and free device memory and synchronise. */
typedef struct ahead_sq_key_index ahead_sq_key_index;
int ahead_hip_star_key_index_create(const ahead_col *key, ahead_sq_key_index **out, void *stream);
int ahead_hip_star_key_index_destroy(ahead_sq_key_index *ix);
/* the index's status: AHEAD_SQ_ST_OK, _BAD_KEY, _DUP_KEY or _KEY_SPAN (the fused form does not
 * apply to queries over this dimension) */
int ahead_hip_star_key_index_status(const ahead_sq_key_index *ix);

typedef struct {
    ahead_col fk;               /* fact FK codes (8-byte codes, keys of at most 32 bits)        */
    ahead_col key;              /* dimension key codes (8-byte codes)                           */
    ahead_col attr;             /* group attribute codes; data NULL: no group attribute         */
    const uint64_t *sel;        /* selected dimension OIDs (encoded); NULL: every row           */
    const uint64_t *sel_count;  /* device: number of OIDs in sel (a selectAN's d_counts[0])     */
    uint64_t sel_a_inv;         /* A^-1 of the OIDs in sel                                      */
    uint32_t flags;             /* AHEAD_SQ_PROBE | AHEAD_SQ_REFETCH                            */
    uint32_t reserved;
    const ahead_sq_key_index *key_index; /* optional index of `key` (NULL: checked per call)   */
} ahead_sq_dim;

typedef struct {
    ahead_col col;              /* fact column; th1 / th2 are ENCODED thresholds, as selectAN's */
    int op1, op2, comb;
    uint32_t reserved;
    uint64_t th1, th2;
} ahead_sq_pred;

typedef struct {
    uint64_t n_fact;
    uint32_t n_dims, n_preds;
    uint32_t meas_kind;         /* AHEAD_SQ_SUM (meas[0]), _SUB, _MULSUM (meas[0], meas[1])     */
    uint32_t meas_pre_mask;     /* bit m: measure m is fetched right after the fact predicates  */
    ahead_sq_dim dims[AHEAD_SQ_MAX_DIMS];  /* in the chain's order                               */
    ahead_sq_pred preds[AHEAD_SQ_MAX_PREDS];
    ahead_col meas[AHEAD_SQ_MAX_MEAS];
    uint64_t a_result;          /* A of the sums (65117: v2_resbigint_t)                        */
    uint64_t a_oid_err;         /* A of error positions (16089)                                 */
    void *ev_scan[2];           /* optional hipEvent_t pair recorded on `stream` around the fact
                                   scan kernel (NULL: none), for timing that kernel alone        */
} ahead_sq_args;

typedef struct {
    uint64_t *sums, *first_row;             /* per group (NULL: not written)                    */
    uint64_t *keys[AHEAD_SQ_MAX_DIMS];      /* per group, attribute codes of dimension d        */
    uint64_t groups_cap;
    uint64_t *err[AHEAD_SQ_KINDS];
    uint64_t err_cap[AHEAD_SQ_KINDS];
    uint64_t *d_counts;                     /* device, 4 + AHEAD_SQ_KINDS entries               */
} ahead_sq_out;

int ahead_hip_star_query(const ahead_sq_args *q, const ahead_sq_out *out, ahead_hip_ws *ws, void *stream);
/* Device bytes of workspace a star query with these arguments uses (dimension maps sized for the
 * largest key span the fused form accepts), for admission by memory; 0 for invalid arguments.  */
size_t ahead_hip_star_query_ws_bytes(const ahead_sq_args *q);

/* Several star queries over one fact table (same n_fact) with shared fact scans: the queries'
 * dimension statistics are read back together (one wait).  Queries of one shape (the same
 * dimensions, FK / key / attribute columns, join flags, predicate columns of at most 4 bytes and
 * measures) are grouped into passes of at most AHEAD_SQ_MAX_SHAPE_PASS whose group tables fit one LDS
 * table: a row takes one lookup per dimension for all of a pass's queries (per-key presence bits
 * and a union attribute index merged from the queries' own maps).  The other queries are packed
 * in order into passes whose fact FK columns (at most 4 distinct) and predicate columns (at most 2
 * distinct) fit one scan; each such pass reads those columns once for all its queries.  Each query's outputs, error vectors and
 * status are exactly those of ahead_hip_star_query with the same arguments; a query whose fused
 * form does not apply gets its status and takes no part in the scans.  No counterpart in the
 * reference (its drivers run one query at a time): it serves concurrent queries of a
 * device-resident schema.  ev_scan of qs[0] (if set) brackets all of the batch's scans. */
#define AHEAD_SQ_MAX_BATCH 8
#define AHEAD_SQ_MAX_SHAPE_PASS 32
int ahead_hip_star_query_batch(const ahead_sq_args *const *qs, const ahead_sq_out *const *outs, uint32_t nq,
                               ahead_hip_ws *ws, void *stream);
/* Device bytes of workspace ahead_hip_star_query_batch uses for these queries; 0 for invalid. */
size_t ahead_hip_star_query_batch_ws_bytes(const ahead_sq_args *const *qs, uint32_t nq);

/* ---- served star queries (request loop) -------------------------------------------------------
 * No counterpart in the reference (its drivers run one query at a time,
 * tools/ssbm/ssbm-q11_continuous.cpp:69-121): the served form over a device-resident schema.
 * `slots` host threads, each with its own HIP stream, workspace and device arena, take the query at
 * the head of the arrival queue together with the queued queries of its shape (the same columns
 * in the same roles; at most max_batch), admit them first come, first served once the device bytes
 * they need (the batch's workspace, read-back area and selection lists) fit hbm_budget beside the
 * running batches (0: no limit), run each query's dimension selections (ahead_hip_select_an on the
 * given attribute column; the OID count stays on the device, where the star query reads it) and
 * one ahead_hip_star_query_batch call, and copy every query's outputs to host memory in one copy.
 * Results are exactly those of the selections and ahead_hip_star_query run one by one.          */
typedef struct ahead_serve ahead_serve;
typedef struct {
    ahead_col col;              /* the dimension's attribute column the selection reads          */
    int op1, op2, comb;
    uint32_t reserved;
    uint64_t th1, th2;          /* ENCODED thresholds, as selectAN's                             */
    uint64_t a_oid, a_oid_inv;  /* A (and A^-1 mod 2^64) of the selected OIDs                     */
} ahead_serve_sel;
typedef struct {
    ahead_sq_args q;            /* dims[d].sel / sel_count / sel_a_inv are filled by the server  */
    uint32_t has_sel[AHEAD_SQ_MAX_DIMS];  /* 1: dimension d is selected by sel[d]; 0: every row   */
    ahead_serve_sel sel[AHEAD_SQ_MAX_DIMS];
} ahead_serve_query;
typedef struct {
    uint64_t status, groups, kept, events;  /* as ahead_hip_star_query's d_counts[2], [0], [1], [3] */
    const uint64_t *sums, *first_row;       /* `groups` entries each                              */
    const uint64_t *keys[AHEAD_SQ_MAX_DIMS];  /* attribute codes per group (NULL: no attribute)   */
    const uint64_t *err[AHEAD_SQ_KINDS];
    uint64_t n_err[AHEAD_SQ_KINDS];
    const uint64_t *sel_err[AHEAD_SQ_MAX_DIMS];  /* the selections' error positions              */
    uint64_t n_sel_err[AHEAD_SQ_MAX_DIMS];
    double t_arrival, t_start, t_finish;    /* seconds on the server's clock (ahead_serve_clock)  */
} ahead_serve_result;
int ahead_serve_create(int device, uint32_t slots, uint32_t max_batch, uint64_t hbm_budget, ahead_serve **out);
/* Copies the query; the columns it names must stay valid until its result has been waited for. */
int ahead_serve_submit(ahead_serve *s, const ahead_serve_query *q, uint64_t *ticket);
/* Blocks until the query has run; the result's arrays stay valid until ahead_serve_release.     */
int ahead_serve_wait(ahead_serve *s, uint64_t ticket, ahead_serve_result *result);
int ahead_serve_release(ahead_serve *s, uint64_t ticket);
double ahead_serve_clock(const ahead_serve *s);
/* Serves what is queued, then joins the slots.                                                   */
int ahead_serve_destroy(ahead_serve *s);

/* ---- data generation (seeded, counter-based; the same streams as ahead_amd/datagen.py) ------
 * Not part of the reference's operator API: the device-side generator for the synthetic
 * columns of the benchmark workloads (the reference loads SSB tables from dbgen instead).
 * out[i] = encode(lo + ((rand(seed, stream, i) >> 32) * (hi - lo + 1)) >> 32), a (0 = plain). */
int ahead_hip_gen_uniform(void *out, uint32_t width, uint64_t n, uint64_t seed, uint64_t stream,
                          uint64_t lo, uint64_t hi, uint64_t a, void *hstream);
/* Flips one bit of each code i with rand(seed, stream, i) < threshold; the bit is
 * ((rand(seed, stream+1, i) >> 32) * 8*width) >> 32.  Optionally appends nothing: the flipped
 * positions follow from the same streams on the host.                                         */
int ahead_hip_inject_flips(void *codes, uint32_t width, uint64_t n, uint64_t seed,
                           uint64_t stream, uint64_t threshold, void *hstream);

#ifdef __cplusplus
}
#endif

#endif /* AHEAD_HIP_H */
