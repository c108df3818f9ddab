/*
 * mpstools_amd.h -- C ABI of the MI355X (gfx950) DMRG hot path.
 *
 * Replaces, on the GPU, the block-sparse contractions of MPSTools'
 * src/MPSfunctions.hpp that the Lanczos/DMRG engines call:
 *   contractTwosite       (src/MPSfunctions.hpp:949-952, body :726-804)
 *   contractSingle        (src/MPSfunctions.hpp:857-859, body :808-854)
 *   addContractionLayer   (src/MPSfunctions.hpp:442-444 and Ap variant :499-502,
 *                          bodies add_out_sub_2 :276-352 / add_out_sub :359-435)
 * and the Lanczos vector algebra of TSLanzcosEngine::simulate
 * (src/lanzcos.hpp:165-278: BlockMatrix dot/axpy/scale, BlockTensor.hpp:443-544).
 *
 * Conventions
 *   - A block-sparse tensor (BlockMatrix<Ns,AbKeyType,T>) crosses the ABI as a
 *     key table (mpsa_blocks) plus one packed device buffer: block i is the
 *     rows[i] x cols[i] column-major matrix (ld = rows[i]) at element
 *     offset[i].  Keys are (q[2i], q[2i+1], p[ns*i .. ns*i+ns-1]): packed
 *     AbKeyType charges (src/keyTypes.hpp:25-30, unsigned 64-bit wrap-around
 *     arithmetic) and physical indices (MPS / two-site psi) or the MPO aux
 *     index (environments), exactly TensorKeyType<2,Ns,AbKeyType>.
 *   - An MPO tensor (MPOMat<1,T>) is a list of entries (wl, wr, out, in, value):
 *     OpKeyType(out,in) as in src/keyTypes.hpp:384-421.
 *   - i_to_key is an array of d packed charges (state s -> charge).
 *   - The reference's matvec reads psi at the operator's out index and writes
 *     at its in index (src/MPSfunctions.hpp:760-792); these entry points keep
 *     that convention, and environments keep "upper = conjugated = in".
 *   - Every call returns 0 on success; on failure a non-zero MPSA_E* code and
 *     mpsa_last_error() describes it (the reference prints and aborts).
 *   - Host memory belongs to the caller; a plan owns its device workspace.
 *     Calls on one plan must be serialised on one stream; distinct plans may
 *     run concurrently.
 *   - Summation order inside a plan is fixed (no floating-point atomics), so
 *     results are bitwise reproducible run to run.
 */
#ifndef MPSTOOLS_AMD_H
#define MPSTOOLS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* mpsa_stream; /* hipStream_t */

enum {
  MPSA_OK = 0,
  MPSA_EINVAL = 1,     /* bad argument, size or key table mismatch */
  MPSA_ENOKEY = 2,     /* an output block produced by the contraction is missing from the output table */
  MPSA_EHIP = 3,       /* HIP runtime error */
  MPSA_ENOMEM = 4,     /* device allocation failed */
  MPSA_ENOTFINITE = 5  /* NaN/Inf detected where the call checks for it */
};

/* Arithmetic of the sector GEMMs.  Inputs and outputs at the ABI are always
 * FP64 (or planar complex128, see mpsa_blocks); F32 and BF16 plans
 * keep reduced-precision copies of the environments and convert psi on the fly.
 * BF16 means bf16 inputs with fp32 accumulation on MFMA; BF16X3 splits every
 * operand into hi + lo bf16 parts and accumulates hi*hi + hi*lo + lo*hi in fp32
 * (about 16 significant bits per operand, three bf16 MFMAs per product). */
enum { MPSA_F64 = 0, MPSA_F32 = 1, MPSA_BF16 = 2, MPSA_BF16X3 = 3 };
/* OR-ed into an H_eff plan's precision: keep the reference's schedule of one
 * product pair per tuple (its "plain method", F_ref flops) instead of sharing
 * the (L, psi) products across MPO paths.  For baselines. */
enum { MPSA_PLAIN = 0x100 };

typedef struct mpsa_blocks {
  int64_t n;              /* number of blocks */
  int32_t ns;             /* physical indices per key (1 or 2) */
  int32_t is_complex;     /* 0: real double.  1: complex128, planar: the real
                             parts of the blocks at their offsets, then the
                             imaginary parts at offset + extent, extent =
                             max(offset + rows*cols); the buffer holds
                             2*extent doubles.  H_eff plans (any precision)
                             and environment plans (FP64) take complex tables
                             with a real MPO: every table of a plan complex,
                             or none. */
  const uint64_t* q;      /* [2n] charges: left/upper, right/lower */
  const uint32_t* p;      /* [ns*n] physical indices or MPO aux index */
  const int64_t* rows;    /* [n] */
  const int64_t* cols;    /* [n] */
  const int64_t* offset;  /* [n] element offsets into the packed buffer */
} mpsa_blocks;

typedef struct mpsa_mpo {
  int64_t n;              /* number of nonzero entries */
  uint32_t Dl, Dr;        /* aux dimensions (MPOMat::Dl/Dr) */
  int32_t is_complex;     /* must be 0: MPO values are real doubles */
  const uint32_t* wl;
  const uint32_t* wr;
  const uint32_t* out;
  const uint32_t* in;
  const double* val;
} mpsa_mpo;

typedef struct mpsa_plan mpsa_plan;

const char* mpsa_last_error(void);
const char* mpsa_version(void);

/* ---- two-site H_eff . psi  (contractTwosite, src/MPSfunctions.hpp:949-981) ----
 * Output key set: the keys the reference's tuple loop produces
 * (src/MPSfunctions.hpp:792-798), in key order.  Query with cap = 0 for the
 * count, then again with arrays of that size. */
int mpsa_heff2_outkeys(const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x, const mpsa_mpo* Wl,
                       const mpsa_mpo* Wr, const uint64_t* itokey_l, int dl, const uint64_t* itokey_r, int dr,
                       int64_t cap, int64_t* nout, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols);

/* Build a grouped-GEMM plan for y = H_eff x.  The y table must contain every
 * produced key (blocks it holds that are not produced are written as zeros);
 * x and y may be the same table (a Lanczos vector space).  precision: MPSA_F64,
 * MPSA_F32 or MPSA_BF16. */
int mpsa_heff2_plan(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                    const mpsa_blocks* y, const mpsa_mpo* Wl, const mpsa_mpo* Wr, const uint64_t* itokey_l, int dl,
                    const uint64_t* itokey_r, int dr, int precision);

/* ---- one-site H_eff . A  (contractSingle, src/MPSfunctions.hpp:857-885) ---- */
int mpsa_heff1_outkeys(const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x, const mpsa_mpo* W,
                       const uint64_t* itokey, int d, int64_t cap, int64_t* nout, uint64_t* q, uint32_t* p,
                       int64_t* rows, int64_t* cols);
int mpsa_heff1_plan(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                    const mpsa_blocks* y, const mpsa_mpo* W, const uint64_t* itokey, int d, int precision);

/* Bind the device environments (FP64 packed buffers in the L/R tables' layout).
 * Reduced-precision plans convert them here, once; call again whenever the
 * environment contents change. */
int mpsa_plan_bind_env(mpsa_plan* plan, const double* dL, const double* dR, mpsa_stream stream);

/* y[v] = H_eff x[v] for v < nvec; vector v of x starts at dx + v*x_stride
 * elements (likewise y).  All nvec vectors go through one grouped launch per
 * stage. */
int mpsa_heff_apply(mpsa_plan* plan, const double* dx, int64_t x_stride, double* dy, int64_t y_stride, int nvec,
                    mpsa_stream stream);

/* ---- environment update (addContractionLayer, src/MPSfunctions.hpp:442-555) ----
 * direction > 0: left environment grows by one site; direction <= 0: right.
 * Ap (may be NULL) is the upper/conjugated tensor of the mixed variant
 * (:499-555); when NULL the upper tensor is A itself. */
int mpsa_env_outkeys(const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap, const mpsa_mpo* W,
                     const uint64_t* itokey, int d, int direction, int64_t cap, int64_t* nout, uint64_t* q,
                     uint32_t* p, int64_t* rows, int64_t* cols);
int mpsa_env_plan(mpsa_plan** plan, const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap,
                  const mpsa_mpo* W, const uint64_t* itokey, int d, int direction, const mpsa_blocks* out,
                  int precision);
int mpsa_env_apply(mpsa_plan* plan, const double* dBL, const double* dA, const double* dAp, double* dout,
                   mpsa_stream stream);

/* ---- plan queries ---- */
int mpsa_plan_info(const mpsa_plan* plan, double* flops_alg, double* flops_ref, int64_t* n_gemm1, int64_t* n_gemm2,
                   int64_t* workspace_bytes);
/* Algorithmic flops of the two grouped GEMM stages (one vector) and the tile
 * class each stage launches with.  FP64/FP32: 0 = 128x128, 1 = 64x64,
 * 2 = 32x32.  BF16/BF16X3: 0 = 256x128, 1 = 64x64, 2 = 128x128, 3 = 256x256,
 * 4 = 32x32. */
int mpsa_plan_stage_info(const mpsa_plan* plan, double* flops_stage1, double* flops_stage2, int* tile_class1,
                         int* tile_class2);
/* Per-stage HIP-event timing of mpsa_heff_apply on the call's stream.  When
 * enabled, mpsa_plan_last_times returns, for the last apply, the milliseconds
 * of [convert x, stage-1 GEMM, mix + zero fill, stage-2 GEMM]. */
int mpsa_plan_set_timing(mpsa_plan* plan, int enable);

/* ---- tile-class tuning (no counterpart in the reference, whose tuple loop
 * hands each product to BLAS separately) ---- */
/* Launch stage 1 or 2 of the plan with the given tile class from now on.
 * MPSA_EINVAL if the class does not exist for the plan's precision or the
 * stage does not admit it (the 256x256 class needs unit segment scales and
 * per-operand offsets below 2^31).  Results do not depend on the class in
 * FP64; in reduced precision they agree within the precision's bound. */
int mpsa_plan_set_tile_class(mpsa_plan* plan, int stage, int tile_class);
/* Autotuner of an H_eff plan with bound environments: times every admissible
 * tile class of each grouped GEMM stage on seeded pseudo-random vectors
 * (nvec vectors per launch, one warm-up and `reps` timed launches, HIP events
 * on `stream`) and keeps the fastest per stage (a tie keeps the plan's own
 * class).  ms (optional, 10 doubles): average milliseconds per launch of
 * stage 1 classes 0..4 then stage 2 classes 0..4, -1 for classes not admitted
 * or not defined for the precision.
 * Allocates and frees two vectors; synchronizes the stream. */
int mpsa_plan_autotune(mpsa_plan* plan, int nvec, int reps, double* ms, mpsa_stream stream);
/* Copy out a plan's device schedule for inspection: what = 0/1 stage-1 tasks /
 * segments, 2/3 stage-2 tasks / segments, 4/5 mix tasks / terms, 6 zero-filled
 * output ranges.  Records are the packed structs
 *   task   {int64 c_off; int32 m, n, seg0, nseg, tile0, tiles_m}            (32 B)
 *   seg    {int64 a_off, b_off; int32 k, lda, ldb, pad; double alpha}       (40 B)
 *   mix    {int64 dst_off, len; int32 t0, nt, chunk0, pad}                  (32 B)
 *   term   {int64 src_off; double coef}                                     (16 B)
 *   range  {int64 off, len}                                                 (16 B)
 * *n receives the record count (cap = 0 queries it; cap is in bytes); *extra
 * the stage's transpose code (tasks), tile class (segments) or precision (mix). */
int mpsa_plan_describe(const mpsa_plan* plan, int what, void* buf, int64_t cap, int64_t* n, int* extra);
int mpsa_plan_last_times(const mpsa_plan* plan, float* ms4);
/* Average stage times over every apply timed since timing was enabled or the
 * last summary (at most 512 applies are kept; later ones overwrite the last
 * slot); *napplies receives their count and the count restarts at 0. */
int mpsa_plan_timing_summary(mpsa_plan* plan, float* ms4, int* napplies);
int mpsa_plan_destroy(mpsa_plan* plan);

/* Extents (elements) of the x and y layouts and whether they are the same
 * vector space (same keys, shapes and offsets).  Complex plans report both
 * planes (2 * extent doubles). */
int mpsa_plan_vector_lengths(const mpsa_plan* plan, int64_t* x_len, int64_t* y_len, int* same_layout);

/* ---- Lanczos (TSLanzcosEngine::simulate + getState, src/lanzcos.hpp:165-278, 81-92) ----
 * Plain three-term Lanczos on the device in FP64 with the plan as matvec (the
 * plan's x and y tables must coincide).  Same recurrence, stopping rules
 * (|k| < delta, Ritz-value test every 5 steps when testenergy, Ritz-vector
 * residual otherwise, itmax) and getState(target) as the reference; the
 * tridiagonal eigensolve runs on the host in FP64.  energies receives the
 * sorted Ritz values (at most cap), *nit the iterations, *nritz their count;
 * dstate (may be NULL) the Ritz vector `target`.  The sign of a Ritz vector
 * is that of the tridiagonal eigensolver (as with dstev, arbitrary).
 * With testenergy the run is mpsa_lanczos_batched with one vector and graph
 * replay (coefficients on the device, one host read-back per 5-step test;
 * the plan's stage timing is switched off); the residual test is driven from
 * the host step by step.
 * On a complex plan the recurrence runs on the planar vector of 2 * extent
 * doubles.  For a Hermitian H_eff the Lanczos vectors are real polynomials in
 * H_eff applied to psi0, and Re<x|y> is the real dot product of the planar
 * vectors, so this is the complex recurrence (same alpha, beta, Ritz values
 * and states). */
int mpsa_lanczos(mpsa_plan* plan, const double* dpsi0, double* dstate, int itmax, double delta, int testenergy,
                 int target, double* energies, int cap, int* nit, int* nritz, mpsa_stream stream);
/* mpsa_lanczos with the options of the reference's projected run
 * (TSLanzcosEngine::simulate(itmax, delta, testenergy, state, delta_ortho),
 * src/lanzcos.hpp:281-414) for both convergence tests: nproj states at
 * dproj + p*proj_stride projected out of the start vector and of every new
 * recurrence vector, reortho = 1 for full FP64 reorthogonalisation; see
 * mpsa_lanczos_batched_ex for the semantics (with testenergy the run is that
 * entry point with one vector). */
int mpsa_lanczos_ex(mpsa_plan* plan, const double* dpsi0, double* dstate, int itmax, double delta, int testenergy,
                    int target, double* energies, int cap, int* nit, int* nritz, const double* dproj, int nproj,
                    int64_t proj_stride, int reortho, mpsa_stream stream);
/* nvec independent runs of mpsa_lanczos's recurrence (Ritz-value convergence
 * test only) from the vectors dpsi0 + v*stride, v < nvec (the reference runs
 * one TSLanzcosEngine per vector, src/lanzcos.hpp:165-278).  Every step is one
 * batched matvec over all nvec vectors (mpsa_heff_apply with nvec) plus
 * batched vector algebra whose coefficients stay on the device; with
 * use_graph != 0 a block of 5 steps (the convergence-test period,
 * src/lanzcos.hpp:107,210) is captured once into a hipGraph and replayed, and
 * the plan's stage timing is switched off.  The host reads the 5 new
 * (alpha, beta) pairs of every vector after each block and applies the
 * reference's stopping rules to each run on its own; a run that has stopped
 * idles (its later steps are computed and ignored) until all have stopped or
 * itmax steps were made.  Per vector v: energies[v*cap ...] sorted Ritz
 * values, nit[v], nritz[v]; dstates (may be NULL) receives Ritz state
 * `target` of run v at dstates + v*stride.  stride >= the plan's vector
 * length. */
int mpsa_lanczos_batched(mpsa_plan* plan, const double* dpsi0, int64_t stride, int nvec, double* dstates, int itmax,
                         double delta, int target, double* energies, int cap, int* nit, int* nritz, int use_graph,
                         mpsa_stream stream);
/* mpsa_lanczos_batched with the options of the reference's projected run
 * (TSLanzcosEngine::simulate(itmax, delta, testenergy, state, delta_ortho),
 * src/lanzcos.hpp:281-414, called by DMRGengine::doStepExcited,
 * src/container.hpp:340-347):
 *   nproj > 0: the start vector and every new recurrence vector lose their
 *     component along nproj states, state p at dproj + p*proj_stride (and,
 *     with proj_vstride > 0, vector v of the batch has its own set at
 *     + v*proj_vstride; proj_vstride = 0 shares one set).  The states need
 *     not be normalised or orthogonal: they are orthonormalised once on the
 *     device (FP64, classical Gram-Schmidt twice); for one state this is the
 *     reference's A - s (s*A) / (s*s).  On a complex plan the complex span of
 *     each state is projected out (s and i s).  A state that is zero or
 *     linearly dependent on the earlier ones (relative norm below 1e-12
 *     after the projection) is an error.  The reference squeezes its state
 *     (drops blocks whose entries are all <= 1e-15); such blocks change the
 *     projection by at most 1e-15 relative and are kept here.
 *   reortho = 1: after every recurrence step the new vector is
 *     reorthogonalised against all Krylov vectors of its run (FP64 classical
 *     Gram-Schmidt, two passes, coefficients on the device, inside the
 *     replayed graph).  The reference's delta_ortho Gram-Schmidt is commented
 *     out (src/lanzcos.hpp:342-366); this is the option in its place, off by
 *     default (plain three-term recurrence, as the reference runs).
 * nproj = 0 and reortho = 0 is mpsa_lanczos_batched. */
int mpsa_lanczos_batched_ex(mpsa_plan* plan, const double* dpsi0, int64_t stride, int nvec, double* dstates, int itmax,
                            double delta, int target, double* energies, int cap, int* nit, int* nritz, int use_graph,
                            const double* dproj, int nproj, int64_t proj_stride, int64_t proj_vstride, int reortho,
                            mpsa_stream stream);
/* ---- Chebyshev moments on one bond (the recurrence of
 * ChebychevEngine::simulateSSCompress, src/container.hpp:2343-2470, with the
 * plan's exact H_eff in place of the compressed MPO application) ----
 * t_0 = dt0, t_1 = Ht t_0, t_{k+1} = 2 Ht t_k - t_{k-1}, Ht = (H_eff - shift)
 * / scale (|scale| above the spectral half-width keeps |T_k| <= 1).
 * moments[k] = <t_0|t_k> for k < nmom (the real part: the real dot product of
 * the vectors, planar for complex plans).  recons (may be NULL, 2*nmom - 1
 * doubles): the reference's reconstructed moments (:2390-2394, 2433-2436),
 * recons[2k] = 2<t_k|t_k> - mu_0 and recons[2k-1] = 2<t_k|t_{k-1}> - mu_1,
 * which equal mu_2k and mu_{2k-1}.  The plan's x and y tables must coincide;
 * allocates four vectors from the Lanczos pool; synchronizes per moment. */
int mpsa_chebyshev_moments(mpsa_plan* plan, const double* dt0, int nmom, double scale, double shift,
                           double* moments, double* recons, mpsa_stream stream);
/* Host FP64 symmetric tridiagonal eigensolver used by mpsa_lanczos (replaces
 * dstev, blas/lapack_routines.cpp:435-446): d (n) diagonal -> ascending
 * eigenvalues, e (n-1) off-diagonal, z (n*n, column-major, may be NULL). */
int mpsa_tridiag_eig(int n, double* d, const double* e, double* z);
/* Free the device vector pool kept between Lanczos calls. */
int mpsa_release_pools(void);
/* Idle bytes held by that pool and its cap (default 8 GiB).  Vectors handed
 * back beyond the cap are freed, vectors of other lengths first; a failed
 * device allocation releases every idle vector and is tried once more.
 * (No reference counterpart: the reference allocates per Lanczos run.) */
int mpsa_pool_info(int64_t* idle_bytes, int64_t* cap_bytes);
int mpsa_pool_set_cap(int64_t cap_bytes);

/* ---- split of a two-site tensor (split, src/MPSfunctions.hpp:1142-1278) ----
 * psi (ns = 2 table, FP64 device buffer) -> left site tensor, singular values
 * lambda, right site tensor.  Per centre charge qc = l + q(s1) the blocks are
 * gathered into one dense matrix (rows by s1, columns by s2), decomposed by an
 * FP64 SVD with the FixPhase sign rule (blas/lapack_routines.cpp:614-634,
 * blas/helperfunctions.hpp:36-67) and truncated over all sectors together:
 * do_trunc = 0 keeps at most chimax values above 1e-15
 * (DiagBlockMatrix::truncate(chi), src/BlockTensor.hpp:234-276), do_trunc = 1
 * keeps at most chimax values until the truncated weight falls below
 * tr_weight (src/BlockTensor.hpp:85-136); chimax = 0 means 100 / 200 as in
 * split :1257-1258.  normalize scales lambda to unit norm.  Left keys are
 * (qc - q(s1), qc, s1), right keys (qc, qc + q(s2), s2) (split :1225-1238),
 * blocks packed in key order; sectors with nothing kept are dropped
 * (dropDangling, src/MPSfunctions.hpp:1444-1484).  direction > 0 multiplies
 * lambda into the right tensor, direction <= 0 into the left one
 * (DMRGengine::doStep, src/container.hpp:321-328).  The result owns its
 * device buffers until mpsa_split_destroy.  When chimax is below the number
 * of singular values, sectors of at least 1024 x 1024 are decomposed by the
 * density-matrix eigendecomposition (syevd of the Gram matrix of the side
 * that stays an isometry) instead of the SVD: the kept set agrees with the
 * SVD's except among directions of weight below ~1e-16, and values below
 * 1e-10 of the sector's largest count as zero. */
typedef struct mpsa_split_result mpsa_split_result;
int mpsa_split(const mpsa_blocks* psi, const double* dpsi, const uint64_t* itokey_l, int dl, const uint64_t* itokey_r,
               int dr, int direction, unsigned chimax, int do_trunc, double tr_weight, int normalize,
               mpsa_split_result** out, mpsa_stream stream);
/* truncated weight and kept count (the pair split returns), block counts of
 * the left/right tables, number of lambda sectors, packed buffer lengths */
int mpsa_split_info(const mpsa_split_result* r, double* truncated_weight, int* chi, int64_t* n_left, int64_t* n_right,
                    int64_t* n_lambda, int64_t* left_len, int64_t* right_len);
/* which = 0: left table, 1: right table (arrays sized by mpsa_split_info) */
int mpsa_split_table(const mpsa_split_result* r, int which, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols,
                     int64_t* offsets);
/* device buffer of the left (0) or right (1) tensor, owned by the result */
const double* mpsa_split_data(const mpsa_split_result* r, int which);
/* copy the left (0) or right (1) packed buffer to caller device memory (synchronous) */
int mpsa_split_copy(const mpsa_split_result* r, int which, double* dst, mpsa_stream stream);
/* lambda: sector charges, sizes and the concatenated values (descending per sector) */
int mpsa_split_lambda(const mpsa_split_result* r, uint64_t* q, int64_t* sizes, double* values);
int mpsa_split_destroy(mpsa_split_result* r);

/* ---- two-site DMRG sweep (DMRGengine, src/container.hpp:293-337, 383-438) ----
 * A chain of nsites site tensors (ns = 1 tables, FP64 device buffers copied
 * at creation; the state must be right-normalised, as DMRGengine::simulate
 * makes it with orthonormalize("rl")), MPO tensors and i_to_key arrays
 * (concatenated, d[i] per site).  Creation builds R[N] = getRMax and the right
 * environments R[N-1] .. R[2] on the device.  A sweep is the body of
 * DMRGengine::simulate: doStep(site, +1) for site = 0 .. N-3, then
 * doStep(site, -1) for site = N-2 .. 1, each step = two-site tensor,
 * Lanczos (lanmax, delta, energy test; `precision` matvecs, then for a
 * reduced precision at most refine_it FP64 Lanczos steps from the Ritz
 * vector), getState(0), split (chimax, do_trunc, tr_weight as mpsa_split,
 * normalised lambda into the site ahead) and the environment update.  Returns
 * the energy of the last step, the largest truncated weight, the total number
 * of Lanczos iterations and the wall time of the sweep. */
typedef struct mpsa_dmrg mpsa_dmrg;
int mpsa_dmrg_create(mpsa_dmrg** dmrg, int nsites, const mpsa_blocks* sites, const double* const* dA, const mpsa_mpo* W,
                     const uint64_t* itokey, const int* d, mpsa_stream stream);
int mpsa_dmrg_sweep(mpsa_dmrg* dmrg, int precision, int refine_it, unsigned chimax, int do_trunc, double tr_weight,
                    int lanmax, double delta, double* energy, double* max_truncated_weight, int* lanczos_iterations,
                    double* seconds);
/* Site tensor `site`: block count and packed length; with non-NULL arrays the
 * table, with non-NULL dst a device-to-device copy of the data. */
int mpsa_dmrg_site(const mpsa_dmrg* dmrg, int site, int64_t* n_blocks, int64_t* len, uint64_t* q, uint32_t* p,
                   int64_t* rows, int64_t* cols, int64_t* offsets, double* dst, mpsa_stream stream);
int mpsa_dmrg_destroy(mpsa_dmrg* dmrg);

/* ---- Energy and its adjoint over a whole chain ----
 * Forward: MPO<T>::measure (src/MPOTypes.hpp:990-1017), E = <psi|H|psi>
 * unnormalised, by the left environment sweep of addContractionLayer from
 * getLMin (src/MPSfunctions.hpp:1917-1926) and the trace of the
 * charge-diagonal blocks of the last environment.  Every `checkpoint`-th left
 * environment is kept on the device.
 * Backward (no counterpart in the reference): dE/dA[i] = 2 contractSingle(
 * L[i-1], A[i], W[i], R[i+1]) on A[i]'s blocks (valid for a real symmetric H),
 * and dE/dtheta_p = sum_i <A[i]| contractSingle(L[i-1], A[i], dW_p[i], R[i+1])>
 * for an MPO affine in its parameters, W[i](theta) = W0[i] + sum_p theta_p
 * dW_p[i].  The right environments are built from getRMax
 * (src/MPSfunctions.hpp:1996-2022) during the backward sweep and the left
 * ones recomputed segment by segment from the checkpoints.
 * Chain description: nsites site tables (ns = 1, FP64), their MPO tensors,
 * i_to_key arrays concatenated (d[i] charges for site i), and nparam*nsites
 * derivative tensors dW[p*nsites + i] (entries with value dW_p[i]; n = 0 where
 * W[i] does not depend on theta_p).  Site buffers dA[i] / gradients dgrad[i]
 * are device FP64 buffers in the site tables' layouts.  energy_backward uses
 * the checkpoints of the last energy_forward, which must have been run on the
 * same tensors. */
typedef struct mpsa_chain mpsa_chain;
int mpsa_chain_create(mpsa_chain** chain, int nsites, const mpsa_blocks* sites, const mpsa_mpo* W,
                      const uint64_t* itokey, const int* d, int nparam, const mpsa_mpo* dW, int checkpoint);
int mpsa_energy_forward(mpsa_chain* chain, const double* const* dA, double* E, mpsa_stream stream);
int mpsa_energy_backward(mpsa_chain* chain, const double* const* dA, double* const* dgrad, double* dE_dtheta,
                         mpsa_stream stream);
/* Key table of a chain environment: side > 0 gives L[index] (index -1 is
 * getLMin), side <= 0 gives R[index] (index nsites is getRMax).  cap = 0
 * queries the block count. */
int mpsa_chain_env_table(const mpsa_chain* chain, int side, int index, int64_t cap, int64_t* n, uint64_t* q,
                         uint32_t* p, int64_t* rows, int64_t* cols, int64_t* offsets);
int mpsa_chain_destroy(mpsa_chain* chain);

/* ---- Versioned checkpoint of block-sparse tensors (in place of the Boost
 * archives of MPSBase::save/load, src/MPSTypes.hpp:377-426, and
 * Container::write/read, src/container.hpp:87-146) ----
 * A file holds a list of tensors (key table + packed FP64 data, all real or
 * all complex128 planar; ns = 1 for MPS tensors and environments, 2 for
 * two-site states) and caller metadata text.  Format version 1: magic,
 * version, flags, then per tensor its table, data and an FNV-1a 64 hash; a
 * trailer hash over the whole file.  mpsa_ckpt_open checks all of them and
 * refuses a truncated, corrupt or other-version file; the handle holds the
 * contents until mpsa_ckpt_close.  data_on_device / dst_on_device: the
 * buffers are device (1) or host (0) memory. */
typedef struct mpsa_ckpt mpsa_ckpt;
int mpsa_ckpt_write(const char* path, int ntensors, const mpsa_blocks* tables, const double* const* data,
                    int data_on_device, const char* meta, mpsa_stream stream);
int mpsa_ckpt_open(const char* path, mpsa_ckpt** out, int* ntensors, int* is_complex, int64_t* meta_bytes);
int mpsa_ckpt_meta(const mpsa_ckpt* c, char* dst, int64_t cap);
/* Table of tensor i (arrays sized by a first call with NULL arrays); *len =
 * doubles of its data. */
int mpsa_ckpt_table(const mpsa_ckpt* c, int i, int* ns, int64_t* n_blocks, int64_t* len, uint64_t* q, uint32_t* p,
                    int64_t* rows, int64_t* cols, int64_t* offsets);
int mpsa_ckpt_read(const mpsa_ckpt* c, int i, double* dst, int dst_on_device, mpsa_stream stream);
int mpsa_ckpt_close(mpsa_ckpt* c);

/* ---- Two-site product BlockMatrix<2,KeyType,T>(B1, B2) (src/BlockTensor.hpp:285-301) ----
 * For every block pair whose B1 right charge equals the B2 left charge, the
 * product B1 B2 keyed (B1 left, B2 right, s1 = B1's index, s2 = B2's index);
 * a key produced by two pairs keeps the later pair's product, as the
 * reference's assignment does.  b1, b2: ns = 1 tables; out: an ns = 2 table
 * holding every produced key (mpsa_twosite_outkeys lists them in key order,
 * cap = 0 queries the count).  All three real FP64 or all complex128 planar.
 * Blocks of out that no pair produces are zeroed.  FP64 grouped GEMM, one
 * launch; returns when done. */
int mpsa_twosite_outkeys(const mpsa_blocks* b1, const mpsa_blocks* b2, int64_t cap, int64_t* n, uint64_t* q,
                         uint32_t* p, int64_t* rows, int64_t* cols);
int mpsa_twosite_product(const mpsa_blocks* b1, const double* d1, const mpsa_blocks* b2, const double* d2,
                         const mpsa_blocks* out, double* dout, mpsa_stream stream);

/* ---- Chebyshev moments with variational MPS compression ----
 * ChebychevEngine::simulateSSCompress (src/container.hpp:2343-2470) and the
 * MPS functions under it, for nvec start vectors at once.  The engine holds
 * eight MPS slots; slot 0 is the start state given at creation (nvec vectors,
 * each nsites ns = 1 tables and device buffers: sites[v * nsites + i],
 * dA[v * nsites + i]; all real FP64 or all complex128 planar), sharing the
 * MPO W (nsites real tensors) and i_to_key (concatenated, d[i] per site).
 * h1_precision is the arithmetic of the one-site contraction inside the
 * compression sweeps (MPSA_F64, or MPSA_F32 for complex64 / float); every
 * other step is FP64.  Results are per vector; complex results are (re, im).
 *   apply_mpo:  slot dst = applyMPO(W, slot src)        (MPSfunctions.hpp:2088-2345)
 *   overlap:    computeOverlap(a, b) = <b|a>            (MPSfunctions.hpp:2066-2084)
 *   axpby:      slot dst = xa a + xb b, the MPS sum of MPSBase::operator+=
 *               with the scale on the first site        (MPSTypes.hpp:223-300)
 *   compress_single_site: slot dst ~ W |slot src>, conv per vector
 *                                                       (MPSfunctions.hpp:2670-2789)
 *   compress:   slot compressed in place, conv per vector (MPSfunctions.hpp:2353-2510)
 *   moments:    mu_n = <t_n|t_0> for n < nmom (nvec x nmom x 2), the
 *               reconstructed doubled moments (nvec x (2 nmom - 1) x 2), and
 *               the two convergence measures per moment (nvec x nmom); slot 1
 *               holds the last compressed t afterwards.  opi = opf = identity.
 *   site:       one vector's site tensor of a slot (its own charges, 64-element
 *               aligned offsets, planar when complex); NULL arrays query sizes.
 * The pair sweeps, canonicalisation (prepareSite with a batched Jacobi SVD per
 * charge sector), truncate and dropDangling follow MPSfunctions.hpp:21-269,
 * 1444-1484, 1705-1727.  One engine is one stream; calls on it are serial. */
typedef struct mpsa_cheb mpsa_cheb;
int mpsa_cheb_create(mpsa_cheb** out, int nsites, int nvec, const mpsa_blocks* sites, const double* const* dA,
                     const mpsa_mpo* W, const uint64_t* itokey, const int* d, int h1_precision, mpsa_stream stream);
int mpsa_cheb_copy(mpsa_cheb* c, int dst, int src);
/* orthonormalize(mps, "rl") (src/MPSfunctions.hpp:1757-1793) of a slot:
 * prepareSite "r" from the last site to site 1 (batched sector SVDs), then
 * each vector's first site divided by its norm; the state is unchanged up to
 * its norm (and sign).  With mpsa_fill_gauss this makes seeded, normalised
 * inputs on the device (in place of prepareMatrix's host SVDs,
 * src/MPSfunctions.hpp:21-269). */
int mpsa_cheb_orthonormalize(mpsa_cheb* c, int slot);
int mpsa_cheb_apply_mpo(mpsa_cheb* c, int dst, int src);
int mpsa_cheb_overlap(mpsa_cheb* c, int a, int b, double* out);
int mpsa_cheb_axpby(mpsa_cheb* c, int dst, double xa, int a, double xb, int b);
int mpsa_cheb_compress_single_site(mpsa_cheb* c, int dst, int src, unsigned chimax, int maxstep, double delta,
                                   double* conv);
int mpsa_cheb_compress(mpsa_cheb* c, int slot, unsigned chimax, int maxstep, double delta, double* conv);
int mpsa_cheb_moments(mpsa_cheb* c, int nmom, unsigned chimax, int ncmax, double delta, double* moments,
                      double* recons, double* conv1, double* conv2, double* seconds);
int mpsa_cheb_site(mpsa_cheb* c, int slot, int vec, int site, int64_t* n_blocks, int64_t* len, uint64_t* q,
                   uint32_t* p, int64_t* rows, int64_t* cols, int64_t* offsets, double* dst);
/* Plans built so far and device bytes held by the engine's buffer pool. */
int mpsa_cheb_stats(const mpsa_cheb* c, int64_t* plans_built, int64_t* pool_bytes);
/* Host-side time of the engine since its creation: out5 = {seconds building
 * plans (with their output keys: the engine enumerates each structure once),
 * seconds in separate output-key queries (none in this engine), seconds
 * waiting on the stream, number of waits, plans cached now}. */
int mpsa_cheb_profile(const mpsa_cheb* engine, double* out5);
/* Wall seconds of each moment of the last mpsa_cheb_moments call (moment 0:
 * the overlap; 1: compressSingleSite; n >= 2: compressSingleSite, the sum
 * and compress); *n = their count, at most cap written. */
int mpsa_cheb_moment_times(const mpsa_cheb* engine, double* out, int cap, int* n);
/* Charge sectors with min(m, n) >= k (default 1024, the batched kernels'
 * limit) are factorised by rocSOLVER (gesvd, geqrf/ungqr, gelqf/unglq), one
 * call each; smaller ones by the batched kernels in one launch.  k in
 * [1, 1024]. */
/* Host-side timers of the engine since its creation, the first n of: seconds
 * in the one-site H_eff steps, the environment layers, the site
 * factorisations (prepareSite), stream waits, device allocation and release
 * (hipMalloc / hipFree), then the number of allocations and releases, seconds
 * building plans, the number of stream waits, plans cached (no counterpart in
 * the reference: diagnostics). */
int mpsa_cheb_timers(const mpsa_cheb* engine, double* out, int n);
int mpsa_cheb_set_big_sector(mpsa_cheb* engine, int k);
/* Sectors with r = min(m, n) >= r_min (default 32) and max(m, n) <= 1024
 * are factorised by several workgroups each: the Householder reflections in
 * panels of 8 columns, Q and the SVD's back-transform column by column, and
 * (SVD steps, r <= 512) the blocked Jacobi sweeps, one launch per round of
 * the block tournament and one host wait per sweep.  r_min = 0 keeps every
 * sector in the one-workgroup kernels.  Same factorisation either way (no
 * counterpart in the reference: a scheduling choice). */
int mpsa_cheb_set_block_jacobi(mpsa_cheb* engine, int r_min);
/* Sector-shape log of the canonicalisation launches (a profiling aid with no
 * counterpart in the reference): cap > 0 clears the log and records up to cap
 * launches from now on, cap = 0 clears it and stops.  Each record is 8 int64:
 * kind (0 SVD, 1 QR/LQ in the batched kernels; 2, 3 the same on rocSOLVER),
 * sectors, max m, max n, max k = min(m, n), and the sectors with k >= 64,
 * >= 128, >= 256.  mpsa_cheb_factor_records copies min(cap, *n) records. */
int mpsa_cheb_factor_log(mpsa_cheb* engine, int64_t cap);
int mpsa_cheb_factor_records(const mpsa_cheb* engine, int64_t* out, int64_t cap, int64_t* n);
int mpsa_cheb_destroy(mpsa_cheb* c);

/* ---- Lanczos vector algebra over packed block vectors (BlockTensor.hpp:443-544) ----
 * X*Y = sum conj(Y) X (src/BlockTensor.hpp:520-544), computed in FP64 with a
 * fixed-order two-level reduction; result written to host *out (re, im). */
int mpsa_dot(const double* dx, const double* dy, int64_t n, int is_complex, double* out, mpsa_stream stream);
/* mpsa_dot for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  The partial sums live in a file-scope device
 * array (1024 pairs, 16 KiB per GPU, any n accepted), the result (re, im) is
 * written to the device pointer dout on `stream` with no host synchronisation,
 * bitwise equal to mpsa_dot's.  The array is shared by every call on one GPU:
 * calls that overlap on one GPU race. */
int mpsa_dot_noalloc(const double* dx, const double* dy, int64_t n, int is_complex, double* dout,
                     mpsa_stream stream);
/* Seeded N(0, 1) entries (no reference counterpart: the reference's
 * initialisers are time-seeded, src/BlockTensor.hpp:394-404):
 * x[i] = sqrt(-2 ln u_{2i}) cos(2 pi u_{2i+1}) with u_j = ((SplitMix64(seed +
 * (j + 1) * 0x9e3779b97f4a7c15) >> 11) + 1) * 2^-53, ln and cos evaluated by
 * fixed fma series, so a host restatement of the same steps reproduces the
 * bits. */
int mpsa_fill_gauss(double* dx, int64_t n, uint64_t seed, mpsa_stream stream);
/* y = a*x + b*y (FP64; complex when is_complex, a and b complex as 2 doubles). */
int mpsa_axpby(const double* a, const double* dx, const double* b, double* dy, int64_t n, int is_complex,
               mpsa_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* MPSTOOLS_AMD_H */
