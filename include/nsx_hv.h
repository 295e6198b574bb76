/* nsx_hv.h -- C ABI of the MI355X proton-neutron J-scheme Hamiltonian (H*v)
 * and thick-restart Lanczos, a drop-in for NuShellX's NuTRL/NuLnz hot path.
 *
 * Plain C: pointers, sizes and return codes only.  Every function returns 0 on
 * success and a negative NSX_E* code on failure, with the message available
 * from nsx_last_error() (thread-local).  Nothing here exits or throws.
 */
#ifndef NSX_HV_H
#define NSX_HV_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  NSX_OK = 0,
  NSX_EIO = -1,     /* a file could not be opened or read */
  NSX_EFORMAT = -2, /* a record has an unexpected size or content */
  NSX_ETAG = -3,    /* .sgt x tag differs from the .tri tag */
  NSX_ECAP = -4,    /* unsupported input or size limit */
  NSX_EGPU = -5,    /* HIP runtime error */
  NSX_EARG = -6     /* invalid argument */
};

/* Arithmetic of the block contractions.  Everything outside them (the Lanczos
 * recurrence, reorthogonalisation, Ritz solve, residuals) is FP64 in all modes.
 * Blocks with a side above 40 run on matrix-core tiles: FP64 MFMA in NSX_FP64,
 * FP32 MFMA in NSX_FP32, BF16 MFMA (the three products) in NSX_BF16X3. */
enum {
  NSX_FP64 = 0,   /* FP64 GEMMs: the plain drop-in */
  NSX_FP32 = 1,   /* FP32 operands, FP32 accumulation */
  NSX_BF16X3 = 2  /* BF16 hi/lo split operands (3 products), FP32 accumulation */
};

/* Flag or-ed into the precision of nsx_hv_create: build the instance on the
 * host only (no device plan, no device memory).  Such a handle can only be
 * used through nsx_batch_create; its own apply/solve calls fail with NSX_EARG.
 * For ensembles of many instances, whose operators are planned once per batch. */
#define NSX_HOST_ONLY 0x100
/* Flag or-ed into the precision of nsx_hv_create: a template of a linear
 * ensemble (nsx_ens_create).  Its pn term structure keeps every key of the
 * .oph file, zero-valued ones too, so that any interaction with the same keys
 * fits it, and it records how each folded b-operator is made from the .oph
 * values.  Usually combined with NSX_HOST_ONLY. */
#define NSX_TEMPLATE 0x200

typedef struct nsx_hv nsx_hv;

/* Reads the files program NuTRL reads for one nucleus and total J and uploads
 * the operators.  Replaces Lnz_TRD's setup, src/NuTRL/NuCoreX.f90:807-1367,
 * with the file set of :1634-1670:
 *   nus_prefix  <Nucleus+Jcode>      -> <nus_prefix>.nus
 *   opfile_a    <Nucleus>a           -> .tri/.trd, and .mat/.sgt per J of side a
 *   opfile_b    <Nucleus>b           -> .tri/.trd, and .mat/.sgt per J of side b
 *   intfile     <Nucleus><xtra>      -> .oph and .spe
 * Names are used as given after trimming blanks (Fortran blank-padded names). */
int nsx_hv_create(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                  const char *intfile, int precision, nsx_hv **h);
/* The same for program NuTRLZ (src/NuTRLZ/NuCoreXZ.f90), whose pn part is the
 * separable operator of <dopfile>.dop (the solver's interaction argument, read by
 * prd() at :1637; written by NuDiop, src/NuDiop/NuDiop.f90:170-280).  The
 * densities are merged along its eigen-terms (prd, :1613-1887) and the term
 * tables carry the Racah factor alone (drdbh, :1536-1595); the NuTRLZ zero
 * limits apply (1e-6 on densities, |W| > 1e-6, :1130,1205,1237).  The like-particle
 * part and <intfile>.oph/.spe are read as by nsx_hv_create.  The handle works
 * with every other call (apply, solve, batch, expect).  NuXpct
 * (src/NuXpct/NuCoreXE.f90) builds its operator the same way. */
int nsx_hv_create_sep(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                      const char *intfile, const char *dopfile, int precision, nsx_hv **h);
void nsx_hv_destroy(nsx_hv *h);

/* N = no_state, basd = number of (2Ja,2Jb) blocks (src/NuTRL/NuCoreX.f90:975-988). */
int nsx_hv_dims(const nsx_hv *h, int64_t *N, int32_t *basd);
/* Block b: 2Ja, 2Jb, offset posnv (0-based) and size dimnv. */
int nsx_hv_block(const nsx_hv *h, int32_t b, int32_t *ja2, int32_t *jb2, int64_t *pos, int64_t *dim);
/* lcoef parity mask of the basis (1 = state of the requested parity), N bytes. */
int nsx_hv_lcoef(const nsx_hv *h, uint8_t *mask);
/* The states a Lanczos solve spans, N bytes.  With an N-hbar-omega window in the
 * .nus (min_MSon/max_MSon, src/NuTRL/NuCoreX.f90:813) this is trunc: the states
 * of the requested parity whose partitions lie inside the window (:990-1037),
 * and the call returns 1; otherwise it is lcoef and the call returns 0.
 * On a truncated basis H*v leaves the blocks with no partition pair inside the
 * window at zero (bas_log, :86-88, :155-157); in the others the pn part reads
 * v masked to trunc, as single_productm gathers the lcoef states and adds only
 * the window ones (src/NuCommonX/NuCommonX.f90:760-812, 885-889), and the
 * like-particle part reads v in full (mulmtx, :1009-1104).  W is not masked:
 * the solvers zero it outside trunc after every H*v, as Lnz_OP does (:307-344). */
int nsx_hv_trunc(const nsx_hv *h, uint8_t *mask);
/* Algorithmic work of one H*v: the reference loop nest's sparse flops (all terms,
 * and the like-particle H_a V + V H_b^T part of them) and FP64 bytes
 * (v read, w written, every operator value read once). */
int nsx_hv_work(const nsx_hv *h, double *flops, double *flops_like, int64_t *bytes);

/* v_out = H v_in on host vectors.  The REAL*4 form is the drop-in for
 *   call multiplyx(); call multiplya(); call multiplyb()
 * with coefR -> v_in and coefL -> v_out (src/NuTRL/NuCoreX.f90:317-319,340-342);
 * v_out is overwritten, as multiplyx does. */
int nsx_hv_apply_f32(nsx_hv *h, const float *v_in, float *v_out);
int nsx_hv_apply_f64(nsx_hv *h, const double *v_in, double *v_out);
/* Device vectors (FP64, N each) on a HIP stream (NULL = the handle's stream). */
int nsx_hv_apply_dev(nsx_hv *h, const double *d_in, double *d_out, void *stream);

/* ---------------------------------------------------------------- batches
 * Many instances (nuclei, J values, interaction samples) in one launch pair;
 * vectors are concatenated in the order given (offsets from nsx_batch_offset).
 * The instance handles must outlive the batch. */
typedef struct nsx_batch nsx_batch;
int nsx_batch_create(nsx_hv *const *hs, int32_t n, int precision, nsx_batch **b);
void nsx_batch_destroy(nsx_batch *b);
int nsx_batch_dims(const nsx_batch *b, int64_t *N_total);
int nsx_batch_offset(const nsx_batch *b, int32_t i, int64_t *off);
int nsx_batch_work(const nsx_batch *b, double *flops, double *flops_like, int64_t *bytes);
int nsx_batch_apply_dev(nsx_batch *b, const double *d_in, double *d_out, void *stream);

/* ---------------------------------------------------------------- tuning
 * Autotuner: rebuilds the work plan for several shapes of the small-block
 * paths (staged: phase capacities, chunk lengths, threads, load mode; stacked:
 * chunk LDS and work; fused: round LDS and rounds) and, when the plan has
 * blocks with a side above 40, of the large-block tiles (stage-1 rows,
 * stage-2 K chunk), and times each with HIP events on the instance's own
 * operators; the fastest is kept.
 * No counterpart in the reference (its loop nest has no tunable shape). */
typedef struct {
  int32_t fat_cap;     /* doubles one phase stages in LDS */
  int32_t max_phases;  /* phases per workgroup */
  int32_t threads;     /* workgroup size of the staged kernel */
  int32_t mode;        /* 0: copy at phase start, 1: direct LDS loads one phase ahead */
  int32_t atomic;      /* 0: deterministic partials + reduce; 1: FP64 atomic accumulation
                          (summation order varies run to run at the 1e-16 relative level) */
  double ms;           /* measured time of one H*v with this shape */
  int32_t stacked;     /* 1: two-stage stacked path (X = V Bhat^T shared across left
                          blocks, then W += A sum r X; atomic accumulation);
                          2: fused column path (X and Y in LDS, column partials
                          summed in a fixed order: deterministic);
                          3: the stacked stages on the matrix cores (16 x 16 tiles in
                          registers, no LDS staging; atomic accumulation);
                          4: both stages fused in one matrix-core kernel (X and Y
                          stay in registers; atomic accumulation, or per-task
                          partial tiles summed in a fixed order when deterministic) */
  int32_t stk_cap;     /* stacked path: doubles a chunk stages in LDS (0 = default);
                          fused path: LDS elements of one round;
                          matrix-core stacked path: a-terms per stage-2 workgroup;
                          fused matrix-core path: steps per wavefront (0 = default 2) */
  int32_t stk_work;    /* stacked path: work units per chunk (0 = default);
                          fused path: flops one workgroup takes, about;
                          matrix-core stacked path: products per stage-1 wavefront;
                          fused matrix-core path: output slots per task (2 or 4; 0 = 4) */
  int32_t f_rounds;    /* fused path: rounds one workgroup takes at most (0 = default) */
  int32_t deterministic; /* 1: bitwise reproducible H*v: for the small blocks the
                            fused matrix-core path with partial tiles (stacked 4, or
                            any stacked value but 2) or the fused column path
                            (stacked 2, and where a slab does not fit the matrix-core
                            tiles), and one stage-2 workgroup per large-block W tile
                            (also set for any plan by NSX_DETERMINISTIC=1) */
  int32_t big_rows1;   /* large-block stage-1 tile rows: 64, or 128 (BF16 hi/lo only);
                          0 = by precision (128 for BF16 hi/lo, else 64) */
  int32_t big_kchunk;  /* large-block stage-2 tiles: active K steps one workgroup takes
                          (split-K chunk; 0 = default 128) */
} nsx_tune_result;
int nsx_hv_autotune(nsx_hv *h, int32_t iters, nsx_tune_result *best);
int nsx_batch_autotune(nsx_batch *b, int32_t iters, nsx_tune_result *best);
int nsx_batch_set_config(nsx_batch *b, const nsx_tune_result *cfg);
/* The small-block pn path the batch's plan runs: 0 staged kernel, 1 stacked
 * two-stage path, 2 fused column path, 3 stacked path on the matrix cores,
 * 8 both stages fused on the matrix cores (plus 16 when that path also carries
 * the like-particle part) (a requested path falls back to the
 * staged kernel where its limits are exceeded), plus 4 when the plan is
 * deterministic.  No counterpart in the reference. */
int nsx_batch_path(const nsx_batch *b, int32_t *path);
/* nsx_batch_create with a plan shape and kernel choice known in advance (for
 * instance tuned on a small representative batch): plans the batch once. */
int nsx_batch_create_config(nsx_hv *const *hs, int32_t n, int precision, const nsx_tune_result *cfg,
                            nsx_batch **b);

/* ---------------------------------------------------------------- Lanczos
 * Thick-restart Lanczos for the lowest eigenpairs (replaces TRL.Lnz_OP,
 * src/NuTRL/NuCoreX.f90:240-542, with device-resident basis vectors).
 * The H*v runs in the handle's precision; everything else is FP64. */
typedef struct {
  int32_t nev;        /* eigenvalues wanted (lowest) */
  int32_t max_basis;  /* Krylov basis size before a restart (e.g. 50) */
  int32_t keep;       /* Ritz vectors kept at a restart (e.g. 10) */
  int32_t max_iter;   /* cap on H*v applications */
  double tol;         /* residual |beta * s_last| bound per wanted pair (MeV) */
  uint64_t seed;      /* random start when no start vector is given */
  int32_t check_every; /* Lanczos steps between convergence tests (0 = 8) */
  int32_t graph;      /* 1: replay each block of steps as a captured hipGraph */
  const char *checkpoint; /* NULL, or a file: the restart state is written there at
                             every thick restart, and a solve that finds a matching
                             file resumes from it (NuTRL's .trl resume,
                             src/NuTRL/NuCoreX.f90:266-292, in FP64) */
  int32_t standard;   /* 1: standard Lanczos as Lnzs (src/NuTRL/NuCoreX.f90:595-751):
                         no restarts (max_basis caps the basis), converged when the
                         lowest nev+1 Ritz values move by <= tol between tests */
} nsx_trl_opts;

typedef struct {
  int32_t matvecs, restarts, converged;
  double seconds;     /* wall time of the solve */
  double host_seconds; /* of which host work on the projected problems */
} nsx_trl_stats;

/* start: optional host start vector (N doubles, NULL = seeded random on lcoef).
 * evals: nev doubles.  evecs: optional host N*nev doubles (column-major). */
int nsx_trl_solve(nsx_hv *h, const nsx_trl_opts *opt, const double *start, double *evals,
                  double *evecs, double *residuals, nsx_trl_stats *stats);

/* Lanczos-vector files for NuVec and the other downstream programs: writes
 * <prefix>.lnz (direct access, one REAL*4 record of N values per vector) and
 * <prefix>.abz (sequential: isave, alpha, beta / imax, no_level, tt, no_state,
 * basd / bas), the layout Lnzs writes (src/NuTRL/NuCoreX.f90:738-743) and NuVec
 * reads (src/NuVec/NuVec.f90:113-131).  The nvec Ritz vectors (vecs: N x nvec,
 * column-major) are the records, alpha their Ritz values and beta zero, so
 * NuVec's tridiagonal step is diagonal and it rebuilds exactly these states.
 * prefix is NuVec's Nucleus // file_code(J, nA, nB) (with its directory).
 * Works on host-only handles too. */
int nsx_hv_write_vectors(const nsx_hv *h, const char *prefix, int32_t nvec, const double *evals,
                         const double *vecs, int32_t no_level, float tt);
/* The thick-restart layout of the same states for NuVec: <prefix>.trl (direct
 * access, one REAL*4 record of N values per vector) and the four-record
 * <prefix>.abz that TRL.Lnz_OP writes (src/NuTRL/NuCoreX.f90:526-534):
 * isave, alpha, betat, matrix(isave, isave) / imax, no_level, tt, no_state,
 * basd, count, no_keep, no_conv, first_iter, second_iter, noiter / bas /
 * alpha_copy(noiter), betat(isave+1), indx(noiter).  The Ritz vectors are the
 * records and matrix = diag(Ritz values), count = 2 (a converged restart
 * state), which NuVec's tred2 + tqli path (src/NuVec/NuVec.f90:181-190) returns
 * exactly.  nvec <= 300. */
int nsx_hv_write_trl(const nsx_hv *h, const char *prefix, int32_t nvec, const double *evals,
                     const double *vecs, int32_t no_level, float tt);
/* NuVec's program step (Vec_OP, src/NuVec/NuVec.f90:99-296): reads <prefix>.abz
 * and <prefix>.lnz or, failing that, <prefix>.trl; diagonalises the projected
 * matrix; forms the lowest min(no_level, no_state, isave) states on the device;
 * writes <prefix>.xvc, .xva (energy, average shell occupations of both sides,
 * percentage per (a partition, b partition, 2Ja, 2Jb)), .lev and .lp.  nlev_out
 * (may be NULL) receives the number of states. */
int nsx_nuvec(nsx_hv *h, const char *prefix, int32_t *nlev_out);
/* <prefix>.xvc, the state file NuVec writes (src/NuVec/NuVec.f90:166,200,291) and
 * NuXpct reads: record (no_state, nvec), then per state (energy, vector), REAL*4. */
int nsx_hv_write_xvc(const nsx_hv *h, const char *prefix, int32_t nvec, const double *evals,
                     const double *vecs);

/* Matrix elements of the handle's operator between nvec states (vecs: N x nvec,
 * column-major, FP64): M(i,j) = (H v_i) . v_j, stored M[i + j*nvec], the matrix
 * NuXpct's matrixs forms (src/NuXpct/NuCoreXE.f90:1441-1448).  H v runs on the
 * device in the handle's precision; the dots are FP64. */
int nsx_hv_expect(nsx_hv *h, int32_t nvec, const double *vecs, double *M);
/* NuXpct's file step (:1431-1470): reads the states of <prefix>.xvc, forms M and
 * writes <prefix><operat>.ovl (text) and <prefix><operat>.ovx (record nlev, then
 * one REAL*4 record per column j).  nlev_out (may be NULL) receives the number of
 * states used: those present, at most the count in the file's header. */
int nsx_hv_xpct_files(nsx_hv *h, const char *prefix, const char *operat, int32_t *nlev_out);
/* nsx_hv_expect without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  The
 * states, their images H v_i and M live in a static 8 MiB device scratch array
 * of the library (one per GPU, held until the process ends), so
 * 2*N*nvec + nvec*nvec must not exceed 1048576 doubles (NSX_ECAP otherwise).
 * Calls that overlap on one GPU share that array and race: serialise them. */
int nsx_hv_expect_noalloc(nsx_hv *h, int32_t nvec, const double *vecs, double *M);

/* ---------------------------------------------------------------- device helpers
 * For hosts without their own HIP binding (Fortran, ctypes): device buffers,
 * copies on the default stream, and HIP-event timing of the H*v kernels.
 * These have no counterpart in the reference (it keeps everything on the host). */
int nsx_dev_alloc(int64_t bytes, void **ptr);
int nsx_dev_free(void *ptr);
int nsx_dev_upload(void *dst, const void *src, int64_t bytes);
int nsx_dev_download(void *dst, const void *src, int64_t bytes);
int nsx_dev_sync(void);
/* warmup launches, then three loops of iters back-to-back launches on the
 * batch stream, each bracketed by one pair of HIP events: the whole H*v, its
 * pn kernels alone, its reduce / like-particle kernels alone.  Average ms per
 * launch of each. */
int nsx_batch_time(nsx_batch *b, const double *d_in, double *d_out, int32_t warmup, int32_t iters,
                   double *ms_hv, double *ms_pn, double *ms_reduce);
/* iters launches of a plain 16-byte-per-lane copy kernel (bytes % 16 == 0) on
 * the null stream; *ms = average per launch (HIP events).  Reference point for
 * HBM counters and bandwidth. */
int nsx_dev_copy(void *dst, const void *src, int64_t bytes, int32_t iters, double *ms);
/* device name and compute-unit count of the current device */
int nsx_device_info(char *name, int32_t len, int32_t *cus);

/* All instances of a batch in lockstep: one batched H*v per Lanczos step, the
 * projected problems solved per instance.  start: optional concatenated start
 * vectors (N_total doubles; warm starts).  evals/residuals: n*nev doubles;
 * converged_at[i]: H*v applications when instance i first had all nev pairs
 * within tol (-1 if never).  evecs: optional N_total*nev doubles, column q of
 * instance i's rows = its q-th Ritz vector (same layout as the start vector). */
int nsx_batch_trl_solve(nsx_batch *b, const nsx_trl_opts *opt, const double *start, double *evals,
                        double *evecs, double *residuals, int32_t *converged_at, nsx_trl_stats *stats);

/* ---------------------------------------------------------------- linear ensembles
 * Many interaction samples theta_s (s < S) of the same nuclei and J values as
 * one batch whose operators are made on the device.  Replaces writing and
 * reading one operator file set per sample: the .mat like-particle matrices
 * (NuOrth, src/NuOrth/NuOrth.f90:206-218) and the .oph pn multipoles (NuOp,
 * src/NuOp/NuOp.f90:352-363) are linear in the P interaction parameters, the
 * densities (.tri/.trd) not at all.  So each template (a handle created with
 * NSX_TEMPLATE from the files of one nucleus and J) takes a basis, the values
 * of its files for each unit parameter vector, and sample s gets
 *   values_s = sum_i theta[s*P + i] basis_t[i*K_t + k]
 * in FP64 on the device, from which its folded b-operators and H(J) blocks are
 * built (.oph values below 1e-10 count as zero, as at load).  The batch then
 * holds n_t*S instances, sample-major (instance s*n_t + t is template t with
 * sample s) and works with nsx_batch_apply_dev, nsx_batch_trl_solve,
 * nsx_batch_time and the other batch calls; autotune and set_config are
 * refused (pass a tuning result at creation).  It runs on the stacked path
 * and the large-block tiles (any block side up to 2048).  No counterpart in
 * the reference, which reads one file set per interaction.
 *
 * Value layout of a template (K_t = n_oph + sum d^2): the .oph record values in
 * file order (the raw me, before renorm), then per H block (side a, then side b,
 * 2J ascending) its d x d matrix column-major.  nsx_ens_layout gives n_oph and,
 * per H block, (side 0/1, 2J, d) in h_desc (3 ints each, up to cap blocks).
 * nsx_ens_values writes the template's own file values in that layout. */
int nsx_ens_layout(const nsx_hv *h, int64_t *n_oph, int32_t *n_h, int32_t *h_desc, int32_t cap);
int nsx_ens_values(const nsx_hv *h, double *values);
/* theta: S x P row-major; basis[t]: P x K_t row-major; cfg: optional tuning
 * result (its stacked-path shape is used). */
int nsx_ens_create(nsx_hv *const *templates, int32_t n_t, int32_t S, int32_t P, const double *theta,
                   const double *const *basis, int precision, const nsx_tune_result *cfg, nsx_batch **b);
/* New parameters for the same S samples (S x P); the operators are rebuilt on
 * the device, the plan is kept. */
int nsx_ens_update(nsx_batch *b, const double *theta);
/* S samples of n_t templates (S = 0 for a batch of separate instances). */
int nsx_ens_samples(const nsx_batch *b, int32_t *S, int32_t *n_t);

/* ---------------------------------------------------------------- learned warm start
 * No counterpart in the reference, whose Lanczos always starts from a random
 * vector (src/NuTRL/NuCoreX.f90:299-312).  A multilayer perceptron
 * n_in -> n_hidden (SiLU) x n_layers -> n_out (linear), trained and evaluated
 * on the GPU: BF16 weights (rounded from FP32 master weights) in every
 * product, FP32 activations and accumulation, loss 1/(2 rows) sum ||y_p - y||^2,
 * Adam (beta 0.9 / 0.999, eps 1e-8).  Host arrays are row-major FP32.
 * Parameter layout: per layer, W (out x in, row-major) then b (out). */
typedef struct nsx_mlp nsx_mlp;
int nsx_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed,
                   nsx_mlp **out);
void nsx_mlp_destroy(nsx_mlp *m);
int nsx_mlp_num_params(const nsx_mlp *m, int64_t *n);
int nsx_mlp_get_params(nsx_mlp *m, float *params);
/* sets the master weights and clears the optimiser state */
int nsx_mlp_set_params(nsx_mlp *m, const float *params);
/* one forward + backward pass over rows (X: rows x n_in, Y: rows x n_out);
 * grad: n_params floats (optional); loss: value of the loss */
int nsx_mlp_grad(nsx_mlp *m, const float *X, const float *Y, int32_t rows, float *grad, double *loss);
/* epochs of shuffled mini-batch Adam; epoch_loss (optional): mean loss per epoch */
int nsx_mlp_train(nsx_mlp *m, const float *X, const float *Y, int32_t rows, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float *epoch_loss);
int nsx_mlp_predict(nsx_mlp *m, const float *X, int32_t rows, float *Y);

const char *nsx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
