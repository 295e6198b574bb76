/* ORACLE (test infrastructure only).
 *
 * refcpu: a C/OpenMP restatement of NuTRL's proton-neutron J-scheme H*v
 * (multiplyx/multiplya/multiplyb, src/NuTRL/NuCoreX.f90:76-226, with the SubM
 * kernels of src/NuCommonX/NuCommonX.f90:609-1104) and of the setup it needs
 * (Lnz_TRD, src/NuTRL/NuCoreX.f90:807-1367; drdbh :1428-1577).
 *
 * It keeps the reference loop nest: bitmask partition sparsity, sequential
 * value pointers walked in next_ij order, per-thread output accumulators
 * reduced after the parallel loop, OpenMP dynamic schedule.  Two precisions:
 *   mode 0 (REAL*4): storage and arithmetic in float, the reference build;
 *   mode 1 (FP64)  : the same algorithm in double, the oracle.
 * The 1e-7 product thresholds of single_productm/single_product1 are applied
 * when `thresholds` is nonzero (the reference always applies them).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library.  It is never part of the product path.
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct refcpu refcpu_t;

int refcpu_create(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                  const char *intfile, int fp64, int thresholds, refcpu_t **out);
/* NuTRLZ's separable pn part from <dopfile>.dop: prd() merges the densities along
 * the eigen-terms (src/NuTRLZ/NuCoreXZ.f90:1613-1887, coefor layout merge) and
 * drdbh() tabulates them with the Racah factor alone (:1536-1595).  limits = 1
 * applies NuTRLZ's zero limits (densities 1e-6, |W| > 1e-6); limits = 0 keeps
 * NuTRL's (1e-11 / 1e-10, W != 0), which isolates the algebra in tests. */
int refcpu_create_sep(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                      const char *intfile, const char *dopfile, int fp64, int thresholds, int limits,
                      refcpu_t **out);
void refcpu_destroy(refcpu_t *h);
int64_t refcpu_dim(const refcpu_t *h);
int refcpu_basd(const refcpu_t *h);
/* block (2Ja, 2Jb, posnv, dimnv) for b in [0, basd) */
int refcpu_block(const refcpu_t *h, int b, int *ja2, int *jb2, int64_t *pos, int64_t *dim);
/* lcoef parity mask (1 = state of the requested total parity) */
int refcpu_lcoef(const refcpu_t *h, uint8_t *mask);
/* N-hbar-omega window mask trunc (:990-1037), returning 1, or lcoef and 0 when
   the .nus sets no window.  On a truncated basis refcpu_apply skips the blocks
   without a partition pair in the window (bas_log, :86-88, :155-157). */
int refcpu_trunc(const refcpu_t *h, uint8_t *mask);
/* v_out = H v_in (double in/out; converted to float in mode 0) */
int refcpu_apply(refcpu_t *h, const double *v_in, double *v_out);
/* algorithmic work of one H*v: sparse flops of the reference loop nest,
   and operator value counts (a-densities, b-densities, H_a, H_b values, table terms) */
int refcpu_work(const refcpu_t *h, double *flops, int64_t *counts5);
int refcpu_set_threads(refcpu_t *h, int nthreads);
const char *refcpu_last_error(void);
/* restatement of tfind (float bisection) on a sorted key array; returns 1-based
   index or 0, exactly as src/NuCommonX/NuCommonX.f90:147-183 */
int refcpu_tfind(const uint8_t *keys8, int n, const uint8_t *key8);
/* 6j symbol {a b c; d e f} with doubled arguments (binomial-sum form) */
double refcpu_sixj(int a, int b, int c, int d, int e, int f);

#ifdef __cplusplus
}
#endif
#endif
