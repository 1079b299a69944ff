// an_simd_dispatch.cpp -- run-time choice among the three builds of an_simd.cpp (x86-64-v4 /
// x86-64-v3 / baseline) by the host CPU's features.  TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
#include <stdint.h>

extern "C" {
#include "an_oracle.h"

uint64_t or_select_an_simd_v4(const or_col *, int, uint64_t, int, uint64_t, int, uint64_t, uint64_t, or_vec *,
                              or_vec *, int);
uint64_t or_select_an_simd_v3(const or_col *, int, uint64_t, int, uint64_t, int, uint64_t, uint64_t, or_vec *,
                              or_vec *, int);
uint64_t or_select_an_simd_v0(const or_col *, int, uint64_t, int, uint64_t, int, uint64_t, uint64_t, or_vec *,
                              or_vec *, int);
int or_star_query_simd_v4(const or_sq_args *, or_sq_out *, int);
int or_star_query_simd_v3(const or_sq_args *, or_sq_out *, int);
int or_star_query_simd_v0(const or_sq_args *, or_sq_out *, int);

/* The build the dispatch picks on this host: 4 = x86-64-v4 (AVX-512 F/BW/DQ/VL), 3 = x86-64-v3
 * (AVX2, BMI2, FMA), 0 = baseline x86-64. */
int or_simd_level(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512dq") &&
        __builtin_cpu_supports("avx512vl"))
        return 4;
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("bmi2") && __builtin_cpu_supports("fma")) return 3;
    return 0;
}

/* selectAN + fused decoded sum on `threads` threads (0: all); the outputs of or_select_an. */
uint64_t or_select_an_simd(const or_col *col, int op1, uint64_t th1, int op2, uint64_t th2, int comb,
                           uint64_t a_head, uint64_t a_oid, or_vec *oids, or_vec *err, int threads) {
    switch (or_simd_level()) {
        case 4: return or_select_an_simd_v4(col, op1, th1, op2, th2, comb, a_head, a_oid, oids, err, threads);
        case 3: return or_select_an_simd_v3(col, op1, th1, op2, th2, comb, a_head, a_oid, oids, err, threads);
        default: return or_select_an_simd_v0(col, op1, th1, op2, th2, comb, a_head, a_oid, oids, err, threads);
    }
}

/* The fused star query of or_star_query on `threads` threads (0: all); the same outputs. */
int or_star_query_simd(const or_sq_args *a, or_sq_out *o, int threads) {
    switch (or_simd_level()) {
        case 4: return or_star_query_simd_v4(a, o, threads);
        case 3: return or_star_query_simd_v3(a, o, threads);
        default: return or_star_query_simd_v0(a, o, threads);
    }
}
}
