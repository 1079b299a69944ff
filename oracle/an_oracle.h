/*
 * an_oracle.h -- CPU restatement of AHEAD's AN-hardened column operators (scalar semantics).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP operators in ahead_amd/:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  Nothing in
 * the product path links, loads or falls back to it.
 *
 * Every function restates one reference operator, cited as file:line relative to the reference
 * repository (brics-db/AHEAD).  All arithmetic is exact unsigned modular integer arithmetic:
 * "mul mod 2^k" is a 64-bit multiply followed by masking to the code width k.
 *
 * Columns are described by or_col.  A "void" column has data == NULL and yields the virtual
 * head value (uint32_t)(seqbase + i), as TempBATIterator<v2_void_t, Tail>::head() does
 * (include/column_storage/TempBatIterator.h:442-444, oid_t arithmetic).
 *
 * Output vectors (or_vec) are malloc'ed by the oracle and released with or_vec_free().
 */
#ifndef AHEAD_AN_ORACLE_H
#define AHEAD_AN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- column descriptor -------------------------------------------------------------------- */
typedef struct {
    const void *data;   /* element array, or NULL for a void (virtual dense OID) column         */
    uint64_t n;         /* number of elements                                                   */
    int width;          /* bytes per element: 1, 2, 4 or 8 (void columns: 4, oid_t)            */
    int encoded;        /* 1: AN-encoded (checked and decoded with a_inv), 0: plain             */
    int unenc_bits;     /* bits of the unencoded base type (TypeMap<T>::v2_base_t::type_t)      */
    uint64_t a;         /* AN_A (ColumnMetaData.hpp:153)                                        */
    uint64_t a_inv;     /* AN_Ainv (ColumnMetaData.hpp:154)                                     */
    uint64_t max_u;     /* AN_unencMaxU (ColumnMetaData.hpp:155)                                */
    uint64_t seqbase;   /* ColumnMetaData.seqbase, used by void columns                         */
} or_col;

typedef struct {
    uint64_t *v;
    uint64_t n;
    uint64_t cap;
} or_vec;

void or_vec_free(or_vec *vec);

/* ---- AN codec ----------------------------------------------------------------------------- */
uint64_t or_mask(int width_bytes);
/* Multiplicative inverse of odd a modulo 2^bits (Newton/Hensel iteration x <- x(2 - a x)).
 * Replaces ext_euclidean (include/util/resilience.hpp:481-516); returns 0 for even a, as
 * ext_euclidean does when gcd(a, 2^bits) != 1. */
uint64_t or_inverse(uint64_t a, int bits);
uint64_t or_get(const or_col *c, uint64_t i);         /* raw element (void: seqbase + i)    */
uint64_t or_dec(const or_col *c, uint64_t i);         /* mulIfEncoded(x, a_inv), width-cast */
int or_is_corrupt(const or_col *c, uint64_t i);       /* encoded && dec > max_u              */

/* ---- operators (scalar semantics) ---------------------------------------------------------- */

/* Predicate codes: the std:: functors the reference instantiates selectAN with
 * (tools/ssbm/ssbm-q*_continuous.cpp) and the combiners of include/column_operators/functors.hpp */
enum { OR_LT = 0, OR_LE = 1, OR_GT = 2, OR_GE = 3, OR_EQ = 4, OR_NE = 5 };
enum { OR_COMB_NONE = 0, OR_COMB_AND = 1, OR_COMB_OR = 2 };

/* selectAN, 1 or 2 predicates on encoded codes vs encoded thresholds
 * (lib/column_operators/AN/selectAN_scalar.tcc:66-95 and :149-181).
 * oids: (uint64)(uint32)(seqbase+pos) * a_head; err: pos * a_oid.  Returns the decoded-value sum
 * of rows that match and are not corrupt (the fused check-and-decode + sum of the select+sum
 * workload), mod 2^64. */
uint64_t or_select_an(const or_col *col, int op1, uint64_t th1, int op2, uint64_t th2, int comb,
                      uint64_t a_head, uint64_t a_oid, or_vec *oids, or_vec *err);

/* checkAN (encdecAN_scalar.tcc:69-94) */
void or_check_an(const or_col *col, uint64_t a_oid, or_vec *err);
/* decodeAN (encdecAN_scalar.tcc:96-126): every row decoded (kept), truncated to unenc_bits */
void or_decode_an(const or_col *col, uint64_t a_oid, or_vec *dec, or_vec *err);
/* checkAndDecodeAN<void,Tail> (encdecAN_scalar.tcc:185-210): corrupt rows dropped */
void or_check_and_decode_an(const or_col *col, uint64_t a_oid, or_vec *dec, or_vec *err);
/* checkAndDecodeAN<Head,Tail> (encdecAN_scalar.tcc:138-175): pairs kept only if both sides OK */
void or_check_and_decode_an2(const or_col *head, const or_col *tail, uint64_t a_oid,
                             or_vec *dec_h, or_vec *dec_t, or_vec *err_h, or_vec *err_t);
/* encodeAN (encdecAN_scalar.tcc:44-67): code = (code_t)v * a, masked to out_width */
void or_encode_an(const or_col *col, uint64_t a, int out_width, or_vec *out);

/* fetchjoinAN (lib/column_operators/AN/matchjoinAN.tcc:312-353).  oids: BAT<void,resoid> tail,
 * col: the fetched tail.  reenc != 0 multiplies kept codes by a_reenc * inv(A_col) (:323-327).
 * Normative bound: t >= |col| goes to err1 (the reference tests t > |col| and reads one past the
 * end for t == |col|, matchjoinAN.tcc:337). */
void or_fetchjoin_an(const or_col *oids, const or_col *col, int reenc, uint64_t a_reenc,
                     uint64_t a_oid, or_vec *out, or_vec *err1, or_vec *err2);

/* matchjoinAN positional <H1, resoid, void, T2> (matchjoinAN.tcc:240-301) */
void or_matchjoin_pos_an(const or_col *h1, const or_col *t1, const or_col *t2, int reenc_h,
                         uint64_t f_h1, int wout_h, int reenc_t, uint64_t f_t2, int wout_t,
                         uint64_t a_oid,
                         or_vec *out_h, or_vec *out_t, or_vec *err1, or_vec *err2, or_vec *err4);

/* matchjoinAN, merge form <H1, T1, H2, T2> on two lists ascending in T1 / H2
 * (matchjoinAN.tcc:46-143): a two-pointer walk over decoded t1 and h2 that emits (h1, t2) on
 * equality.  Rows are checked where the reference's loops visit them: list-2 rows at the top of
 * every outer iteration and again when skipped, list-1 rows only when skipped.
 * reenc_h / reenc_t multiply encoded outputs by f_h1 / f_t2 (mulIfEncoded). */
void or_matchjoin_merge_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                           int reenc_h, uint64_t f_h1, int wout_h, int reenc_t, uint64_t f_t2,
                           int wout_t, uint64_t a_oid, or_vec *out_h, or_vec *out_t,
                           or_vec *err1, or_vec *err2, or_vec *err3, or_vec *err4);

/* matchjoinAN <H1, void, void, T2> (matchjoinAN.tcc:146-224): align the virtual OIDs
 * t1->seqbase + i and h2->seqbase + j by skipping leading rows of the lower side (checking h1 resp.
 * t2 on the skipped rows), then pair the rest one to one, checking both sides. */
void or_matchjoin_seq_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                         int reenc_h, uint64_t f_h1, int wout_h, int reenc_t, uint64_t f_t2, int wout_t,
                         uint64_t a_oid, or_vec *out_h, or_vec *out_t, or_vec *err1, or_vec *err4);

/* hashjoinAN with an unencoded-key hash map (hashjoinAN.tcc:77-222).  build_right != 0 is
 * hash_side_t::right (the default).  reenc_h / reenc_t enable the re-encoding variants
 * (:561-585): outputs h1 * f_h1 and t2 * f_t2 with f = mulIfEncoded(A_R, A_inv).
 * out_h / out_t receive (h1, t2) pairs in probe order, then build-insertion order per key;
 * wout_h / wout_t are the output code widths in bytes (re-encoded outputs wrap there). */
void or_hashjoin_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                    int build_right, int reenc_h, uint64_t f_h1, int wout_h, int reenc_t,
                    uint64_t f_t2, int wout_t, uint64_t a_oid, or_vec *out_h, or_vec *out_t,
                    or_vec *e1, or_vec *e2, or_vec *e3, or_vec *e4);

/* groupbyAN unary (groupbyAN.tcc:51-94): vgid[i] = gid*a_oid, groid[g] = head(first)*a_oid */
void or_groupby_an(const or_col *head, const or_col *tail, uint64_t a_oid,
                   or_vec *voidtorgid, or_vec *vgidtoroid, or_vec *e_head, or_vec *e_tail);

/* groupbyAN binary (groupbyAN.tcc:99-165).  larger_bits: width of Tail::v2_larger_t::type_t.
 * Returns 0, or -1 on a size mismatch (the reference throws std::runtime_error). */
int or_groupby2_an(const or_col *head, const or_col *tail, const or_col *grouping,
                   uint64_t num_groups, int larger_bits, uint64_t a_oid,
                   or_vec *voidtorgid, or_vec *vgidtoroid, or_vec *e_head, or_vec *e_tail,
                   or_vec *e_group);

/* aggregate_sum_groupedAN (aggregateAN.tcc:41-119).  sums must hold num_groups entries.
 * Normative: a decoded group id <= max_u but >= num_groups is reported in e_group and skipped
 * (the reference writes out of bounds there).  Returns 0, or -1 on a size mismatch. */
int or_aggregate_sum_grouped_an(const or_col *head, const or_col *tail, const or_col *grouping,
                                uint64_t num_groups, uint64_t a_result, uint64_t a_oid,
                                uint64_t *sums, or_vec *e_head, or_vec *e_tail, or_vec *e_group);

/* aggregate_mul_sumAN (aggregateAN_scalar.tcc:58-131): returns the encoded total */
uint64_t or_aggregate_mul_sum_an(const or_col *t1, const or_col *t2, uint64_t init,
                                 int result_encoded, uint64_t ra, uint64_t a_oid,
                                 or_vec *e1, or_vec *e2);

/* arithmeticAN (arithmeticAN_scalar.tcc:60-112); op: 0 add, 1 sub, 2 mul, 3 div.
 * Normative: division by a zero operand yields 0 (undefined in the reference).
 * Returns 0, or -1 on a size mismatch. */
int or_arithmetic_an(const or_col *h1, const or_col *t1, const or_col *h2, const or_col *t2,
                     int op, uint64_t a_result, int result_width, uint64_t a_oid, or_vec *out,
                     or_vec *eh1, or_vec *et1, or_vec *eh2, or_vec *et2);

/* ---- composed workloads ---------------------------------------------------------------------- */

/* Star-join pipeline over one dimension (join + fetch + 2-key group-by + grouped sum), composed
 * from the operators above in the order of tools/ssbm/ssbm-q21_continuous.cpp:97-181 with the
 * row-consistent policy: a joined row whose fetched attribute or measure is corrupt is dropped
 * from group-by and sum (the reference aborts there, groupbyAN.tcc:105-107).
 * Results: per binary group g (first-occurrence order): sums[g] (encoded with a_result),
 * groid[g] (first joined+kept row * a_oid_grp), key1[g], key2[g] (attribute codes).
 * Error vectors: e_fk (fact FK), e_dk (dimension key), e_a1, e_a2, e_m (fetch corruptions),
 * e_grp (the grouped sum's group-id check, aggregateAN.tcc:92-94).
 * Returns the number of groups, or -1 on failure. */
int64_t or_star1_pipeline(const or_col *fact_fk, const or_col *fact_m, const or_col *dim_key,
                          const or_col *dim_a1, const or_col *dim_a2, uint64_t a_oid_join,
                          uint64_t a_oid_err, uint64_t a_oid_grp, uint64_t a_result,
                          or_vec *sums, or_vec *groid, or_vec *key1, or_vec *key2,
                          or_vec *e_fk, or_vec *e_dk, or_vec *e_a1, or_vec *e_a2, or_vec *e_m,
                          or_vec *e_grp);

/* ---- fused star query (one SSB operator chain as one pass over the fact rows) -------------------
 * The chains of tools/ssbm/ssbm-q{11..43}_continuous.cpp evaluated row by row in fact order, with
 * the row-consistent policy for corrupt fetched codes (reported, row dropped from the sums; the
 * reference's chain throws there, groupbyAN.tcc:105-107).  Per fact row i:
 *   1. every fact predicate column p is checked (corrupt: E[PRED p] += i) and its predicate is
 *      evaluated on the raw code (selectAN_scalar.tcc:84-90); the row goes on if all pass;
 *   2. measures fetched right after the predicates (flag bit m of meas_pre_mask; the discount
 *      fetch of ssbm-q11_continuous.cpp:113) are checked: corrupt -> E[MEAS m] += i, dropped;
 *   3. the dimensions in chain order: the FK code is checked; a corrupt code is reported
 *      (E[FK d] += i) and, unless the dimension is probed straight on the fact column (PROBE:
 *      hashjoinAN keeps probing with the truncated key, hashjoinAN.tcc:199), ends the row (the
 *      positional fetch drops it, matchjoinAN.tcc:287-295).  The key truncated to the key's
 *      unencoded width must be a selected dimension row's key, or the row ends;
 *   4. rows that joined every dimension fetch their final-scope measures (corrupt: E[MEAS m] += i),
 *      re-fetch the FK of REFETCH dimensions (corrupt: E[REFETCH d] += i) and fetch the group
 *      attribute of every dimension that has one (corrupt: E[ATTR d] += dimension row); any of
 *      these drops the row;
 *   5. kept rows add their measure to their group: SUM d0, SUB d0 - d1, MULSUM d0 * d1 (decoded
 *      values, mod 2^64); groups are numbered by their first kept fact row.
 * Sums are written encoded (x a_result); keys are the attribute codes; error vectors hold
 * position * a_oid_err.  Status: 0 done; 1 a selected OID is out of range, 2 a dimension key is
 * corrupt, 3 two dimension rows share a key (the fused form does not apply; the caller runs the
 * operator chain). */
enum { OR_SQ_SUM = 0, OR_SQ_SUB = 1, OR_SQ_MULSUM = 2 };
enum { OR_SQ_PROBE = 1, OR_SQ_REFETCH = 2 };
enum { OR_SQ_E_PRED = 0, OR_SQ_E_MEAS = 2, OR_SQ_E_FK = 4, OR_SQ_E_REFETCH = 8, OR_SQ_E_ATTR = 12,
       OR_SQ_KINDS = 16 };
typedef struct {
    or_col fk;            /* fact FK codes                                        */
    or_col key;           /* dimension key codes                                  */
    or_col attr;          /* group attribute codes; data == NULL and n == 0: none */
    const uint64_t *sel;  /* selected dimension OIDs (encoded), NULL: every row   */
    uint64_t n_sel;
    uint64_t sel_a_inv;
    int flags;            /* OR_SQ_PROBE | OR_SQ_REFETCH                          */
} or_sq_dim;
typedef struct {
    or_col col;
    int op1, op2, comb;
    uint64_t th1, th2;    /* encoded thresholds                                   */
} or_sq_pred;
typedef struct {
    uint64_t n_fact;
    int n_dims, n_preds, n_meas, meas_kind, meas_pre_mask;
    or_sq_dim dims[4];
    or_sq_pred preds[2];
    or_col meas[2];
    uint64_t a_result, a_oid_err;
} or_sq_args;
typedef struct {
    or_vec sums, first_row, keys[4];  /* per group, first-occurrence order */
    or_vec err[OR_SQ_KINDS];
    uint64_t kept;
    int status;
} or_sq_out;
int or_star_query(const or_sq_args *a, or_sq_out *o);
void or_sq_out_free(or_sq_out *o);

#ifdef __cplusplus
}
#endif

#endif /* AHEAD_AN_ORACLE_H */
