// mpsa_internal.h -- descriptors shared by the plan builder (host) and the
// gfx950 kernels.  All structs are plain data copied to device memory once per
// plan.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits>
#include <functional>
#include <vector>

struct mpsa_plan;
struct mpsa_blocks;
struct mpsa_mpo;

namespace mpsa {

// One output block of a grouped GEMM stage: C (m x n, ld = m) at c_off.
// C = sum over segments s of alpha_s * op(A_s) * op(B_s).
struct GTask {
  int64_t c_off;
  int32_t m, n;
  int32_t seg0, nseg;
  int32_t tile0;    // first tile of this task inside one vector's tile range
  int32_t tiles_m;  // tiles along m (tiles are numbered m-fastest)
};

struct GSeg {
  int64_t a_off, b_off;
  int32_t k;
  int32_t lda, ldb;
  int32_t pad;
  double alpha;
};

// dst block (len elements) = sum_t coef_t * src_t (all of length len).
struct MixTask {
  int64_t dst_off;
  int64_t len;
  int32_t t0, nt;
  int32_t chunk0;  // first 1024-element chunk of this task inside one vector
  int32_t pad;
};
struct MixTerm {
  int64_t src_off;
  double coef;
};

struct Repack {
  int64_t src_off, dst_off;
  int32_t rows, cols;
  int32_t dst_ld, transpose;
  int32_t chunk0, pad;
  double scale = 1.0;  // dst = convert(scale * src)
};

struct Range {
  int64_t off, len;
};

enum Prec { F64 = 0, F32 = 1, BF16 = 2, BF16X3 = 3 };

// Plane offsets (elements) of hi/lo split operands (BF16X3); zero otherwise.
struct Planes {
  int64_t a = 0, b = 0, c = 0;
};

// Launchers (mpsa_kernels.hip).  `tr` packs (transA << 1) | transB.
// a_vs/b_vs/c_vs are per-vector strides (elements) for batched launches.
hipError_t launch_gemm(int prec_in, int out_kind, int tr, int tile_class, const GTask* tasks, int ntasks,
                       const GSeg* segs, int tiles_per_vec, int nvec, const void* A, int64_t a_vs, const void* B,
                       int64_t b_vs, void* C, int64_t c_vs, hipStream_t st, Planes planes = Planes());
// out_kind: 0 = same type as the input precision, 1 = double
hipError_t launch_mix(int prec, const MixTask* tasks, int ntasks, const MixTerm* terms, int chunks_per_vec, int nvec,
                      const void* src, void* dst, int64_t vs, hipStream_t st, int64_t plane = 0);
hipError_t launch_convert(int prec, const double* src, void* dst, int64_t n, hipStream_t st);
hipError_t launch_repack(int prec, const Repack* tasks, int ntasks, int chunks_per_vec, int nvec, const double* src,
                         int64_t s_vs, void* dst, int64_t d_vs, hipStream_t st, int64_t plane = 0);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st);
// p[i] = uniform in [-1, 1) from SplitMix64 of seed + i * golden gamma
hipError_t launch_fill_hash(double* p, int64_t n, uint64_t seed, hipStream_t st);
// p[i] = sqrt(-2 ln u_{2i}) cos(2 pi u_{2i+1}), u_j = ((SplitMix64(seed + (j + 1)
// * golden gamma) >> 11) + 1) 2^-53: seeded N(0, 1) entries, host-reproducible
hipError_t launch_fill_gauss(double* p, int64_t n, uint64_t seed, hipStream_t st);
// p[i * stride] = v for i < n
hipError_t launch_fill_strided(double* p, int64_t n, int64_t stride, double v, hipStream_t st);
hipError_t launch_zero_ranges(const Range* r, int nr, double* base, int64_t vs, int nvec, hipStream_t st);
hipError_t launch_dot(const double* x, const double* y, int64_t n, int cplx, double* partial, int nparts,
                      double* out2, hipStream_t st);
hipError_t launch_axpby(double ar, double ai, const double* x, double br, double bi, double* y, int64_t n, int cplx,
                        hipStream_t st);
int gemm_max_grid(int prec_in, int tile_class);
// Finalise a grouped GEMM task list built outside a plan (mpsa_plan.cpp): the
// plans' own invariants (segments with k > 0, non-empty tasks, longest first,
// tiles numbered in task order for the chosen class).
void finalize_gemm_list(std::vector<GTask>& tasks, std::vector<GSeg>& segs, int prec, int* tiles, int* tile_class);
// whether the plan's operands are complex128 (planar)
bool plan_complex(const mpsa_plan* P);
// One-pass planners (mpsa_plan.cpp) for callers that build the output table
// themselves: the structure's produced keys, in key order with their shapes,
// go to make_out, which returns the output table the plan is then built
// against; one tuple enumeration serves both (the C ABI's *_outkeys + *_plan
// enumerate twice).  No plan (*plan = nullptr) when nothing is produced.
// Returns an MPSA_* code, with the message in mpsa_last_error.
typedef std::function<mpsa_blocks(int64_t n, const uint64_t* q, const uint32_t* p, const int64_t* rows,
                                  const int64_t* cols)>
    OutTableFn;
int env_keys_plan(mpsa_plan** plan, const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap,
                  const mpsa_mpo* W, const uint64_t* itk, int d, int direction, const OutTableFn& make_out);
int heff1_keys_plan(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                    const mpsa_mpo* W, const uint64_t* itk, int d, int precision, const OutTableFn& make_out);
// the library's rocBLAS handle of the current device (a rocblas_handle),
// created on first use; callers set its stream before each use
void* rocblas_handle_current();
// planar complex (re, im: n each) -> interleaved z (2 n)
hipError_t launch_pack_z(const double* re, const double* im, int64_t n, double* z, hipStream_t st);
// rows x cols window of A (real, or interleaved complex when cplx; ld lda) ->
// planar out (ld ldo, imaginary plane at out + plane): mode 0 whole window,
// 1 upper trapezoid (i <= j), 2 lower (i >= j), zeros elsewhere
hipError_t launch_unpack_mat(const double* A, int cplx, int64_t lda, int64_t rows, int64_t cols, int mode,
                             double* out, int64_t ldo, int64_t plane, hipStream_t st);
// dense block copies of the split (column-major; see copy2d_kernel)
struct Copy2D {
  int64_t src_off, dst_off;
  int32_t rows, cols, lds, ldd;
  int64_t rs_off, cs_off;  // offsets into the scale array, -1 for none
  double scale;
  int32_t trans, pad;
};
hipError_t launch_copy2d(const Copy2D* tasks, int ntasks, int64_t max_elems, const double* src, double* dst,
                         const double* scales, hipStream_t st);
hipError_t launch_fixphase(double* U, int64_t m, int64_t ldu, double* Vt, int64_t k, int64_t n, int64_t ldv,
                           hipStream_t st);
// out[j] = |A[:, j]| for the cols columns of A (rows x cols, ld lda)
hipError_t launch_colnorms(const double* A, int64_t rows, int64_t cols, int64_t lda, double* out, hipStream_t st);
// Batched thin SVD of small sector matrices (mpsa_factor.hip): M (m x n, ld
// lda) at a_off -> U (m x k, ld m) at u_off, S (k) at s_off, V^dagger (k x n,
// ld k) at vt_off, k = min(m, n); workspace at w_off of svd_workspace(m, n)
// elements per plane (p = max(m, n), r = min(m, n) <= kSvdMaxR: X and the
// QR's reflection vectors, p x r each, the rotated R^H and the accumulated
// rotations, r x r each, and r betas).  Complex data planar.
constexpr int kSvdMaxR = 1024;
inline int64_t svd_workspace(int64_t m, int64_t n) {
  const int64_t p = m > n ? m : n, r = m > n ? n : m;
  return 2 * p * r + 2 * r * r + r;
}
struct SvdTask {
  int64_t a_off, u_off, vt_off, s_off, w_off;
  int32_t m, n, lda, pad;
};
hipError_t launch_svd_batched(int cplx, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                              double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                              int64_t w_plane, hipStream_t st);
// The same SVD in phases for large sectors (the kPh* bits below), several
// workgroups per sector for the reflections, the sweeps and the
// back-transform: the sweeps are launch_svd_block_round launches (one per
// round of the block tournament over all tasks; svd_block_rounds(m, n) rounds
// per sweep, svd_block_pairs(m, n) workgroups per task, their bases in
// pair_base[ntasks + 1], rotated[task] set when a rotation happened) until a
// sweep rotates nothing.  min(m, n) <= kSvdBlockMaxR.  Same outputs and
// workspace as launch_svd_batched.
constexpr int kSvdBlockMaxR = 512;
hipError_t launch_svd_phase(int cplx, int phases, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                            double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                            int64_t w_plane, hipStream_t st);
// rmax: the largest min(m, n) of the launch's tasks (sizes its LDS); round -1:
// the pairs inside each block, rounds 0 .. svd_block_rounds - 1: the pairs
// across the blocks the tournament pairs in that round
hipError_t launch_svd_block_round(int cplx, const SvdTask* tasks, int ntasks, const int* pair_base, int npairs,
                                  int round, double* W, int64_t w_plane, double* rotated, int rmax, hipStream_t st);
int svd_block_pairs(int m, int n);
// Phases of svd_jacobi_kernel / qr_householder_kernel (bit set; kPhAll: the
// whole factorisation in one launch).  The multi-workgroup path for sectors
// with max(m, n) <= kVecRowsMax runs kPhLoad, launch_house_panels (in place of
// kPhReflect), kPhInit, then for the SVD block rounds (kPhSweep),
// kPhBackInit, and launch_house_apply (in place of kPhApply), kPhOut.
constexpr int kPhLoad = 1, kPhReflect = 2, kPhInit = 4, kPhSweep = 8, kPhBackInit = 16, kPhApply = 32, kPhOut = 64;
constexpr int kPhAll = 127;
constexpr int kVecRowsMax = 1024;
hipError_t launch_qr_phase(int cplx, int right, int phases, const SvdTask* tasks, int ntasks, const double* M,
                           int64_t m_plane, double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S,
                           double* W, int64_t w_plane, hipStream_t st);
// workspace layout: 0 the SVD's, 1 / 2 the left / right QR's; col_base[s] ..
// col_base[s + 1]: sector s's house_chunks(m, n, layout) workgroups; kmax: the largest
// min(m, n)
hipError_t launch_house_panels(int cplx, int layout, const SvdTask* tasks, int ntasks, const int* col_base,
                               int nchunks, int kmax, double* W, int64_t w_plane, hipStream_t st);
hipError_t launch_house_apply(int cplx, int layout, const SvdTask* tasks, int ntasks, const int* col_base,
                              int nchunks, double* W, int64_t w_plane, hipStream_t st);
int house_chunks(int m, int n, int layout);
int svd_block_rounds(int m, int n);
// Batched Householder QR of the same tasks in the SVD's output format with
// S = 1: right = 0: U = Q (m x k), V^dagger = R (k x n); right = 1: M = L Q'
// with U = L (m x k), V^dagger = Q' (k x n, orthonormal rows).  Workspace
// p * r + p * k + k elements per plane at w_off, p = rows of X (m, or n when
// right), r = its columns.
hipError_t launch_qr_batched(int cplx, int right, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                             double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                             int64_t w_plane, hipStream_t st);
// batched Lanczos vector algebra (device-resident coefficients and step counter)
hipError_t launch_bdot(const double* x, const double* y, int64_t n, int64_t vs, int nvec, double* partial, int nparts,
                       double* out, int ld, const int* step, int root, hipStream_t st);
hipError_t launch_lz_normalize(const double* W, double* X, double* const* basis, const double* kbuf, int ld,
                               const int* step, int64_t n, int64_t vs, int nvec, hipStream_t st);
hipError_t launch_lz_update(const double* Y, const double* X, double* const* basis, const double* kbuf,
                            const double* ebuf, int ld, const int* step, double* W, int64_t n, int64_t vs, int nvec,
                            double* wpart, hipStream_t st);
// lz_update with wpart writes nvec x lz_update_parts(n) partial sums of the new
// W squared; launch_bdot_final reduces nparts partials per vector into
// out[v * ld + *step] (sqrt when root), as launch_bdot's second pass
int lz_update_parts(int64_t n);
hipError_t launch_bdot_final(const double* partial, int nparts, int nvec, double* out, int ld, const int* step,
                             int root, hipStream_t st);
hipError_t launch_lz_step(int* step, hipStream_t st);
// W_v -= sum_{i < nb} <Q_i(v), W_v> Q_i(v) with Q_i(v) = Q[i] + v * qvs,
// nb = (cnt ? *cnt : 0) + add (at most nbmax); partial holds
// nvec * nbmax * nparts doubles, c nvec * nbmax
hipError_t launch_project_out(double* W, double* const* Q, int64_t qvs, const int* cnt, int add, int nbmax, int64_t n,
                              int64_t vs, int nvec, double* partial, int nparts, double* c, hipStream_t st);
// X_v *= 1 / sqrt(nrm2[v])
hipError_t launch_scale_rsqrt(double* X, const double* nrm2, int64_t n, int64_t vs, int nvec, hipStream_t st);
hipError_t launch_lz_combine(double* const* basis, const double* c, int ldc, const int* m, double* out, int64_t n,
                             int64_t vs, int nvec, hipStream_t st);

// Tile of each (precision, tile class).  FP64/FP32: 0 = 128x128, 1 = 64x64,
// 2 = 32x32 per workgroup; 3 = 16x16 and 4 = 32x32 per wave (four waves per
// workgroup, no LDS).  BF16/BF16X3: 0 = 256x128, 1 = 64x64, 2 = 128x128,
// 3 = 256x256, 4 = 32x32.
constexpr int kMaxTileClasses = 5;
inline void tile_dims(int prec, int cls, int* bm, int* bn) {
  static const int bf[5][2] = {{256, 128}, {64, 64}, {128, 128}, {256, 256}, {32, 32}};
  static const int fp[5][2] = {{128, 128}, {64, 64}, {32, 32}, {16, 16}, {32, 32}};
  const bool b = prec == BF16 || prec == BF16X3;
  const int c = cls < 0 ? 0 : (cls > 4 ? 4 : cls);
  *bm = b ? bf[c][0] : fp[c][0];
  *bn = b ? bf[c][1] : fp[c][1];
}
inline int num_tile_classes(int) { return 5; }
void set_last_error(const char* msg);

}  // namespace mpsa
