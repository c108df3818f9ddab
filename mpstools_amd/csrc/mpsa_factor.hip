// mpsa_factor.hip -- batched thin SVD of many small charge-sector matrices on
// gfx950, one workgroup per sector: the SVD behind prepareSite / prepareBlock
// (src/MPSfunctions.hpp:21-140, BoostSVD + FixPhase blas/lapack_routines.cpp:
// 614-634) for MPS canonicalisation and truncation, where a site tensor splits
// into hundreds of sectors of tens of rows each.
//
// Method: X = M (m >= n) or X = M^H (m < n), p x r with p >= r, is reduced
// by Householder QR, X = Q R, and one-sided Jacobi (Hestenes) runs on the
// columns of A = R^H (r x r; the Drmac-Veselic preconditioning without
// pivoting).  Column pairs are visited in the round-robin order of a
// tournament (r/2 disjoint pairs per round, r - 1 rounds per sweep), one wave
// per pair, so a wave needs no barrier inside a rotation; the workgroup
// synchronises between rounds.  A pair is rotated when
// |a_i^H a_j| > tol sqrt(|a_i|^2 |a_j|^2) with tol = sqrt(r) eps; sweeps stop
// when a whole sweep rotates nothing.  The 2 x 2 Hermitian Jacobi rotation
// with gamma = |gamma| e^{i phi}, zeta = (b - a) / (2 |gamma|),
// t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), c = 1 / sqrt(1 + t^2), s = c t:
//   a_i' = c a_i - s e^{-i phi} a_j,   a_j' = s e^{i phi} a_i + c a_j
// makes the pair orthogonal; Wm accumulates the same rotations.  Then
// sigma_j = |a_j|, sorted descending (ties by column index), Z = A / sigma,
// and X = (Q Wm) Sigma Z^H:
//   m >= n: U = Q Wm, V^dagger = Z^H;   m < n: U = Z, V^dagger = (Q Wm)^H.
// Singular vectors of exactly zero singular values (on the Z side; Q Wm is
// orthonormal by construction) are completed to an orthonormal basis
// (Gram-Schmidt of unit vectors, twice), as a LAPACK thin SVD returns them.
// Real data: FixPhase makes the first largest-|.| entry of every row of
// V^dagger positive (the column of U flips with it).
//
// Complex data is planar (real parts, then imaginary parts at a plane offset),
// like every complex buffer of the library.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "mpsa_internal.h"

namespace mpsa {

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Column j of a p-row matrix with leading dimension ld, planar complex.
template <bool C>
struct Mat {
  double* re;
  double* im;  // unused when !C
  int64_t ld;
  __device__ __forceinline__ double& r(int i, int j) const { return re[i + (int64_t)j * ld]; }
  __device__ __forceinline__ double& m(int i, int j) const { return im[i + (int64_t)j * ld]; }
};

// Rotate columns (i, j) of a p-row matrix by the Jacobi rotation (c, s, phase
// er + i ei): x_i' = c x_i - s conj(e) x_j, x_j' = s e x_i + c x_j.
template <bool C>
__device__ __forceinline__ void rotate_cols(const Mat<C>& X, int p, int i, int j, double c, double s, double er,
                                            double ei, int lane) {
  for (int row = lane; row < p; row += 64) {
    const double xir = X.r(row, i), xjr = X.r(row, j);
    if (C) {
      const double xii = X.m(row, i), xji = X.m(row, j);
      // conj(e) x_j = (er - i ei)(xjr + i xji) = (er xjr + ei xji) + i (er xji - ei xjr)
      const double cjr = er * xjr + ei * xji, cji = er * xji - ei * xjr;
      // e x_i = (er + i ei)(xir + i xii) = (er xir - ei xii) + i (er xii + ei xir)
      const double eir = er * xir - ei * xii, eii = er * xii + ei * xir;
      X.r(row, i) = c * xir - s * cjr;
      X.m(row, i) = c * xii - s * cji;
      X.r(row, j) = s * eir + c * xjr;
      X.m(row, j) = s * eii + c * xji;
    } else {
      X.r(row, i) = c * xir - s * er * xjr;
      X.r(row, j) = s * er * xir + c * xjr;
    }
  }
}

// Rows per lane of the register-held column paths: a column of up to
// 64 * kRpl rows lives in registers of one wave while it is reduced and
// updated (one global read and one write instead of two reads and a write),
// and a reflection vector of up to kRpl * 64 rows sits in LDS.
constexpr int kRpl = 16;
constexpr int kVecRows = 64 * kRpl;
static_assert(kSvdMaxR <= kVecRows, "a Jacobi column of r <= kSvdMaxR rows fits the register path");

// y <- y - f v for one column y (rows from..p-1) held by one wave, with
// f = beta (v^H y) and v at vl (LDS, planar at + kVecRows, v[0] at row from).
template <bool C>
__device__ __forceinline__ void reflect_col_regs(const Mat<C>& Y, int c, int from, int p, const double* vl,
                                                 double b, int lane) {
  double yr[kRpl], yi[kRpl];
  double wr_ = 0.0, wi_ = 0.0;
#pragma unroll
  for (int q = 0; q < kRpl; ++q) {
    const int i = from + lane + 64 * q;
    yr[q] = yi[q] = 0.0;
    if (i < p) {
      yr[q] = Y.r(i, c);
      if (C) yi[q] = Y.m(i, c);
      const double v_r = vl[i - from];
      wr_ += v_r * yr[q];
      if (C) {
        const double v_i = vl[kVecRows + i - from];
        wr_ += v_i * yi[q];
        wi_ += v_r * yi[q] - v_i * yr[q];  // conj(v) y
      }
    }
  }
  wr_ = wave_sum(wr_);
  if (C) wi_ = wave_sum(wi_);
  const double fr = b * wr_, fi = b * wi_;
#pragma unroll
  for (int q = 0; q < kRpl; ++q) {
    const int i = from + lane + 64 * q;
    if (i < p) {
      const double v_r = vl[i - from];
      const double v_i = C ? vl[kVecRows + i - from] : 0.0;
      Y.r(i, c) = yr[q] - (fr * v_r - fi * v_i);
      if (C) Y.m(i, c) = yi[q] - (fr * v_i + fi * v_r);
    }
  }
}

// The same update with v read from its global column (any p).
template <bool C>
__device__ __forceinline__ void reflect_col_global(const Mat<C>& Y, int c, int from, int p, const Mat<C>& Hv, int j,
                                                   double b, int lane) {
  double wr_ = 0.0, wi_ = 0.0;
  for (int i = from + lane; i < p; i += 64) {
    const double v_r = Hv.r(i, j), y_r = Y.r(i, c);
    wr_ += v_r * y_r;
    if (C) {
      const double v_i = Hv.m(i, j), y_i = Y.m(i, c);
      wr_ += v_i * y_i;
      wi_ += v_r * y_i - v_i * y_r;
    }
  }
  wr_ = wave_sum(wr_);
  if (C) wi_ = wave_sum(wi_);
  const double fr = b * wr_, fi = b * wi_;
  for (int i = from + lane; i < p; i += 64) {
    const double v_r = Hv.r(i, j);
    const double v_i = C ? Hv.m(i, j) : 0.0;
    Y.r(i, c) -= fr * v_r - fi * v_i;
    if (C) Y.m(i, c) -= fr * v_i + fi * v_r;
  }
}

// Householder reflections of X (p x r, planar, ld p) in place, one per
// column j < k = min(p, r): H_j = I - beta_j v_j v_j^H with v_j = x - alpha e_j,
// alpha = -e^{i arg x_j} |x| (x = rows j.. of column j), beta_j = 2 / |v_j|^2
// (0 for a zero column: H_j = I).  Column j becomes (alpha, 0, ...), the
// columns after it are reflected (one wave per column, the column in
// registers and v_j in LDS when p - j <= kVecRows); v_j goes to rows j.. of
// column j of Hv, beta_j to beta[j].  NT threads; red holds NT / 64 doubles,
// hv 3, vl 2 kVecRows (LDS).
// house_reflect_range does reflections jb .. je - 1 and applies each to the
// columns after it up to c_end - 1 (the panel of the multi-workgroup path).
template <bool C, int NT>
__device__ void house_reflect_range(const Mat<C>& X, const Mat<C>& Hv, double* beta, int p, int jb, int je, int c_end,
                                    double* red, double* hv, double* vl) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = jb; j < je; ++j) {
    double a = 0.0;
    for (int i = j + tid; i < p; i += NT) {
      const double x = X.r(i, j);
      a += x * x;
      if (C) {
        const double y = X.m(i, j);
        a += y * y;
      }
    }
    a = wave_sum(a);
    if (lane == 0) red[wave] = a;
    __syncthreads();
    if (tid == 0) {
      double nn = 0.0;
      for (int w = 0; w < NW; ++w) nn += red[w];
      const double x0r = X.r(j, j), x0i = C ? X.m(j, j) : 0.0;
      const double a0 = sqrt(x0r * x0r + x0i * x0i), nrm = sqrt(nn);
      double ar = 0.0, ai = 0.0, b = 0.0;
      if (nn > 0.0) {
        const double er = a0 > 0.0 ? x0r / a0 : 1.0, ei = a0 > 0.0 ? x0i / a0 : 0.0;
        ar = -er * nrm;
        ai = -ei * nrm;
        const double v0r = x0r - ar, v0i = x0i - ai;
        const double vn = nn - a0 * a0 + v0r * v0r + v0i * v0i;
        b = vn > 0.0 ? 2.0 / vn : 0.0;
      }
      hv[0] = ar;
      hv[1] = ai;
      hv[2] = b;
      beta[j] = b;
    }
    __syncthreads();
    const double ar = hv[0], ai = hv[1], b = hv[2];
    const bool in_lds = p - j <= kVecRows;
    for (int i = j + tid; i < p; i += NT) {
      double v_r = X.r(i, j), v_i = C ? X.m(i, j) : 0.0;
      if (i == j) {
        v_r -= ar;
        v_i -= ai;
      }
      Hv.r(i, j) = v_r;
      if (C) Hv.m(i, j) = v_i;
      if (in_lds) {
        vl[i - j] = v_r;
        if (C) vl[kVecRows + i - j] = v_i;
      }
    }
    __syncthreads();
    for (int c = j + 1 + wave; c < c_end; c += NW) {
      if (in_lds) reflect_col_regs<C>(X, c, j, p, vl, b, lane);
      else reflect_col_global<C>(X, c, j, p, Hv, j, b, lane);
    }
    for (int i = j + tid; i < p; i += NT) {
      X.r(i, j) = i == j ? ar : 0.0;
      if (C) X.m(i, j) = i == j ? ai : 0.0;
    }
    __syncthreads();
  }
}

template <bool C, int NT>
__device__ void house_reflect(const Mat<C>& X, const Mat<C>& Hv, double* beta, int p, int r, double* red,
                              double* hv, double* vl) {
  house_reflect_range<C, NT>(X, Hv, beta, p, 0, p < r ? p : r, r, red, hv, vl);
}

// Y <- H_0 H_1 ... H_{k-1} Y for the ncols columns of Y (p rows), applying
// H_j for j = k-1 .. 0; with from_j only columns >= j (Y = [I; 0] there, as
// LAPACK's xORGQR forms Q).  vl: 2 kVecRows doubles of LDS.
template <bool C, int NT>
__device__ void house_apply(const Mat<C>& Y, int ncols, bool from_j, const Mat<C>& Hv, const double* beta, int p,
                            int k, double* vl) {
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = k - 1; j >= 0; --j) {
    const double b = beta[j];
    if (b == 0.0) continue;  // uniform across the workgroup
    const bool in_lds = p - j <= kVecRows;
    if (in_lds) {
      for (int i = j + tid; i < p; i += NT) {
        vl[i - j] = Hv.r(i, j);
        if (C) vl[kVecRows + i - j] = Hv.m(i, j);
      }
      __syncthreads();
    }
    for (int c = (from_j ? j : 0) + wave; c < ncols; c += NW) {
      if (in_lds) reflect_col_regs<C>(Y, c, j, p, vl, b, lane);
      else reflect_col_global<C>(Y, c, j, p, Hv, j, b, lane);
    }
    __syncthreads();
  }
}

}  // namespace

template <bool C>
__global__ __launch_bounds__(kThreads) void svd_jacobi_kernel(const SvdTask* __restrict__ tasks,
                                                              const double* __restrict__ M, int64_t m_plane,
                                                              double* __restrict__ U, int64_t u_plane,
                                                              double* __restrict__ Vt, int64_t vt_plane,
                                                              double* __restrict__ S, double* __restrict__ W,
                                                              int64_t w_plane, int max_sweeps, int phases) {
  __shared__ double sig[kSvdMaxR];
  __shared__ int rank[kSvdMaxR];
  __shared__ int rotated;
  __shared__ double red[kWaves], redc[kWaves];
  __shared__ double hv[3];
  __shared__ double vl[2 * kVecRows];
  const SvdTask t = tasks[blockIdx.x];
  const int m = t.m, n = t.n;
  if (m <= 0 || n <= 0) return;
  const bool wide = m < n;
  const int p = wide ? n : m, r = wide ? m : n, k = r;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // X = Q R (Householder), then the rotations run on the columns of
  // A = R^H (r x r), accumulated in Wm: A Wm = Z Sigma, so R = Wm Sigma Z^H
  // and X = (Q Wm) Sigma Z^H.  On R^H the sweeps of a graded or rank-deficient
  // sector are a third of those on R or X itself (the rows of a triangular
  // factor are closer to orthogonal than its columns).  The p-side singular
  // vectors Q Wm are orthonormal by construction; the r-side ones are the
  // rotated columns of A divided by sigma.  Workspace: X (p x r), Wm (r x r),
  // the reflection vectors (p x r), A (r x r), betas (r).
  Mat<C> X{W + t.w_off, W + t.w_off + w_plane, p};
  Mat<C> Wm{W + t.w_off + (int64_t)p * r, W + t.w_off + (int64_t)p * r + w_plane, r};
  const int64_t o_hv = (int64_t)p * r + (int64_t)r * r, o_a = 2 * (int64_t)p * r + (int64_t)r * r;
  Mat<C> Hv{W + t.w_off + o_hv, W + t.w_off + o_hv + w_plane, p};
  Mat<C> A{W + t.w_off + o_a, W + t.w_off + o_a + w_plane, r};
  double* const beta = W + t.w_off + o_a + (int64_t)r * r;
  // phases (kPh* bits; kPhAll: the whole SVD): the multi-workgroup path runs
  // the load, then house_panel/update kernels for the reflections, the init,
  // block rounds for the sweeps, the back-transform's init, house_apply_kernel,
  // and the outputs, each phase reading what the previous left in the workspace
  // X = M or M^H
  if (phases & kPhLoad) {
    for (int64_t e = tid; e < (int64_t)p * r; e += kThreads) {
      const int i = (int)(e % p), j = (int)(e / p);
      const int64_t src = wide ? (j + (int64_t)i * t.lda) : (i + (int64_t)j * t.lda);
      X.r(i, j) = M[t.a_off + src];
      if (C) X.m(i, j) = wide ? -M[t.a_off + m_plane + src] : M[t.a_off + m_plane + src];
    }
    __syncthreads();
  }
  if (phases & kPhReflect) house_reflect<C, kThreads>(X, Hv, beta, p, r, red, hv, vl);
  // A = R^H (lower triangular), Wm = identity
  if (phases & kPhInit) {
    for (int64_t e = tid; e < (int64_t)r * r; e += kThreads) {
      const int i = (int)(e % r), j = (int)(e / r);
      A.r(i, j) = i >= j ? X.r(j, i) : 0.0;
      if (C) A.m(i, j) = i >= j ? -X.m(j, i) : 0.0;
      Wm.r(i, j) = i == j ? 1.0 : 0.0;
      if (C) Wm.m(i, j) = 0.0;
    }
    __syncthreads();
  }
  if (phases & kPhSweep) {
  const double tol = sqrt((double)r) * DBL_EPSILON;
  const int rr = r + (r & 1);
  for (int sweep = 0; sweep < max_sweeps && r > 1; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < rr - 1; ++round) {
      for (int q = wave; q < rr / 2; q += kWaves) {
        int i = (round + q) % (rr - 1);
        int j = q == 0 ? rr - 1 : (round + rr - 1 - q) % (rr - 1);
        if (i > j) {
          const int x = i;
          i = j;
          j = x;
        }
        if (j >= r) continue;
        // the pair's columns stay in registers (r <= kSvdMaxR = 64 kRpl rows)
        double xir[kRpl], xii[kRpl], xjr[kRpl], xji[kRpl];
        double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
        for (int q = 0; q < kRpl; ++q) {
          const int row = lane + 64 * q;
          xir[q] = xii[q] = xjr[q] = xji[q] = 0.0;
          if (row < r) {
            xir[q] = A.r(row, i);
            xjr[q] = A.r(row, j);
            a += xir[q] * xir[q];
            b += xjr[q] * xjr[q];
            gr += xir[q] * xjr[q];
            if (C) {
              xii[q] = A.m(row, i);
              xji[q] = A.m(row, j);
              a += xii[q] * xii[q];
              b += xji[q] * xji[q];
              gr += xii[q] * xji[q];
              gi += xir[q] * xji[q] - xii[q] * xjr[q];  // a_i^H a_j = sum conj(a_i) a_j
            }
          }
        }
        a = wave_sum(a);
        b = wave_sum(b);
        gr = wave_sum(gr);
        if (C) gi = wave_sum(gi);
        const double g = C ? sqrt(gr * gr + gi * gi) : fabs(gr);
        if (!(g > tol * sqrt(a * b)) || a == 0.0 || b == 0.0) continue;
        const double er = gr / g, ei = C ? gi / g : 0.0;
        const double zeta = (b - a) / (2.0 * g);
        const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
        // a_i' = c a_i - s conj(e) a_j, a_j' = s e a_i + c a_j (as rotate_cols)
#pragma unroll
        for (int q = 0; q < kRpl; ++q) {
          const int row = lane + 64 * q;
          if (row < r) {
            if (C) {
              const double cjr = er * xjr[q] + ei * xji[q], cji = er * xji[q] - ei * xjr[q];
              const double eir = er * xir[q] - ei * xii[q], eii = er * xii[q] + ei * xir[q];
              A.r(row, i) = c * xir[q] - s * cjr;
              A.m(row, i) = c * xii[q] - s * cji;
              A.r(row, j) = s * eir + c * xjr[q];
              A.m(row, j) = s * eii + c * xji[q];
            } else {
              A.r(row, i) = c * xir[q] - s * er * xjr[q];
              A.r(row, j) = s * er * xir[q] + c * xjr[q];
            }
          }
        }
        rotate_cols<C>(Wm, r, i, j, c, s, er, ei, lane);
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  }
  // X = Q [Wm; 0]: the accumulated rotations back through the reflections
  if (phases & kPhBackInit) {
    for (int64_t e = tid; e < (int64_t)p * r; e += kThreads) {
      const int i = (int)(e % p), j = (int)(e / p);
      X.r(i, j) = i < r ? Wm.r(i, j) : 0.0;
      if (C) X.m(i, j) = i < r ? Wm.m(i, j) : 0.0;
    }
    __syncthreads();
  }
  if (phases & kPhApply) house_apply<C, kThreads>(X, r, false, Hv, beta, p, r, vl);
  if (!(phases & kPhOut)) return;
  // singular values (column norms of the rotated A) and their descending order
  for (int j = wave; j < r; j += kWaves) {
    double a = 0.0;
    for (int row = lane; row < r; row += 64) {
      const double x = A.r(row, j);
      a += x * x;
      if (C) {
        const double y = A.m(row, j);
        a += y * y;
      }
    }
    a = wave_sum(a);
    if (lane == 0) sig[j] = sqrt(a);
  }
  __syncthreads();
  for (int j = tid; j < r; j += kThreads) {
    int rk = 0;
    const double sj = sig[j];
    for (int i = 0; i < r; ++i) rk += (sig[i] > sj || (sig[i] == sj && i < j)) ? 1 : 0;
    rank[j] = rk;
    S[t.s_off + rk] = sj;
  }
  __syncthreads();
  // U (m x k, ld m) and V^dagger (k x n, ld k): the p side from Q Wm, the r
  // side from A / sigma
  double* Ur = U + t.u_off;
  double* Ui = U + t.u_off + u_plane;
  double* Vr = Vt + t.vt_off;
  double* Vi = Vt + t.vt_off + vt_plane;
  for (int64_t e = tid; e < (int64_t)p * r; e += kThreads) {
    const int row = (int)(e % p), j = (int)(e / p);
    const int rk = rank[j];
    if (!wide) {  // U[:, rk] = (Q Wm)[:, j]
      Ur[row + (int64_t)rk * m] = X.r(row, j);
      if (C) Ui[row + (int64_t)rk * m] = X.m(row, j);
    } else {  // V^dagger[rk, :] = (Q Wm)[:, j]^H
      Vr[rk + (int64_t)row * k] = X.r(row, j);
      if (C) Vi[rk + (int64_t)row * k] = -X.m(row, j);
    }
  }
  for (int64_t e = tid; e < (int64_t)r * r; e += kThreads) {
    const int row = (int)(e % r), j = (int)(e / r);
    const int rk = rank[j];
    const double sj = sig[j];
    const double zr = sj > 0.0 ? A.r(row, j) / sj : 0.0;
    const double zi = (C && sj > 0.0) ? A.m(row, j) / sj : 0.0;
    if (!wide) {  // V^dagger[rk, row] = conj(Z[row, j])
      Vr[rk + (int64_t)row * k] = zr;
      if (C) Vi[rk + (int64_t)row * k] = -zi;
    } else {  // U[row, rk] = Z[row, j]
      Ur[row + (int64_t)rk * m] = zr;
      if (C) Ui[row + (int64_t)rk * m] = zi;
    }
  }
  __syncthreads();
  // orthonormal completion of the r-side vectors of zero singular values
  // (descending order puts them last): vectors of length r stored at
  // base + idx * vs with stride es between elements (V^dagger rows: es = k,
  // vs = 1; U columns: es = 1, vs = m).  The conjugate of a V^dagger row is
  // the vector.
  const int64_t es = wide ? 1 : k, vs = wide ? m : 1;
  double* Br = wide ? Ur : Vr;
  double* Bi = wide ? Ui : Vi;
  const double csign = wide ? 1.0 : -1.0;  // stored imaginary part = csign * vector imaginary part
  for (int rk = 0; rk < k; ++rk) {
    if (S[t.s_off + rk] > 0.0) continue;
    for (int cand = 0; cand < r; ++cand) {
      // w = e_cand, then two passes of w -= (u_i^H w) u_i over i < rk
      for (int row = tid; row < r; row += kThreads) {
        Br[rk * vs + row * es] = row == cand ? 1.0 : 0.0;
        if (C) Bi[rk * vs + row * es] = 0.0;
      }
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < rk; ++i) {
          double dr = 0.0, di = 0.0;
          for (int row = tid; row < r; row += kThreads) {
            const double ur = Br[i * vs + row * es], wr = Br[rk * vs + row * es];
            dr += ur * wr;
            if (C) {
              const double ui = csign * Bi[i * vs + row * es], wi = csign * Bi[rk * vs + row * es];
              dr += ui * wi;
              di += ur * wi - ui * wr;
            }
          }
          dr = wave_sum(dr);
          if (C) di = wave_sum(di);
          if (lane == 0) {
            red[wave] = dr;
            redc[wave] = di;
          }
          __syncthreads();
          double Dr = 0.0, Di = 0.0;
          for (int w = 0; w < kWaves; ++w) {
            Dr += red[w];
            Di += redc[w];
          }
          for (int row = tid; row < r; row += kThreads) {
            const double ur = Br[i * vs + row * es];
            const double ui = C ? csign * Bi[i * vs + row * es] : 0.0;
            // w -= (Dr + i Di) u
            Br[rk * vs + row * es] -= Dr * ur - Di * ui;
            if (C) Bi[rk * vs + row * es] -= csign * (Dr * ui + Di * ur);
          }
          __syncthreads();
        }
      double nn = 0.0;
      for (int row = tid; row < r; row += kThreads) {
        const double wr = Br[rk * vs + row * es];
        nn += wr * wr;
        if (C) {
          const double wi = Bi[rk * vs + row * es];
          nn += wi * wi;
        }
      }
      nn = wave_sum(nn);
      if (lane == 0) red[wave] = nn;
      __syncthreads();
      double tot = 0.0;
      for (int w = 0; w < kWaves; ++w) tot += red[w];
      __syncthreads();
      if (tot > 0.25) {
        const double inv = 1.0 / sqrt(tot);
        for (int row = tid; row < r; row += kThreads) {
          Br[rk * vs + row * es] *= inv;
          if (C) Bi[rk * vs + row * es] *= inv;
        }
        __syncthreads();
        break;
      }
    }
  }
  // FixPhase (real data): first index of the largest |V^dagger[i, c]|; flip
  // row i of V^dagger and column i of U unless that entry is > 0
  if (!C) {
    for (int i = wave; i < k; i += kWaves) {
      double best = 0.0;
      int bi = 0;
      for (int c = lane; c < n; c += 64) {
        const double a = fabs(Vr[i + (int64_t)c * k]);
        if (a > best) {
          best = a;
          bi = c;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      const double phi = Vr[i + (int64_t)bi * k] > 0.0 ? 1.0 : -1.0;
      if (phi < 0.0) {
        for (int c = lane; c < n; c += 64) Vr[i + (int64_t)c * k] = -Vr[i + (int64_t)c * k];
        for (int row = lane; row < m; row += 64) Ur[row + (int64_t)i * m] = -Ur[row + (int64_t)i * m];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Block Jacobi rounds for large sectors, several workgroups per sector.
// After the reduction and init phases the r columns of A = R^H are
// split into blocks of kBj columns; one launch runs one round of the block
// tournament over every large sector at once, one workgroup per block pair
// (I, J): it loads the 2 kBj columns into LDS, rotates every pair (i in I,
// j in J) once with the scalar kernel's rotation and threshold, writes them
// back and applies the accumulated 2 kBj x 2 kBj rotation to the same columns
// of Wm.  The launch with round < 0 does the pairs inside each block (a
// tournament over its columns, one workgroup per block), so a sweep (that
// launch, then the rounds) visits every pair of columns once.  Sectors have
// their block pairs at pair_base[s] .. pair_base[s + 1] of the grid (twice
// that for the intra-block launch); rotated[s] is set when any pair of the
// sector rotated (the host repeats sweeps until none does, then the
// back-transform and output phases finish).  r <= kBjMaxR.
constexpr int kBj = 8;
constexpr int kBjMaxR = 512;
static_assert(kBj == kWaves, "one wave per column of block I in the rounds across blocks");

template <bool C>
__global__ __launch_bounds__(kThreads) void svd_block_round_kernel(const SvdTask* __restrict__ tasks, int ntasks,
                                                                   const int* __restrict__ pair_base, int round,
                                                                   double* __restrict__ W, int64_t w_plane,
                                                                   double* __restrict__ rotated, int rmax) {
  // the block pair's columns: 2 kBj x rmax per plane of dynamic LDS (rmax: the
  // launch's largest r, so that small launches fit two workgroups per CU)
  extern __shared__ double blk_lds[];
  double* const Lr = blk_lds;
  double* const Li = blk_lds + 2 * kBj * rmax;
  __shared__ double Vr[2 * kBj * 2 * kBj], Vi[2 * kBj * 2 * kBj];
  __shared__ int any;
  // sector of this workgroup: pair_base is ascending, pair_base[0] = 0; the
  // intra-block launch (round < 0) has two workgroups per block pair slot
  int lo = 0, hi = ntasks - 1;
  const bool intra = round < 0;
  const int b = blockIdx.x, bs = intra ? b >> 1 : b;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pair_base[mid] <= bs) lo = mid;
    else hi = mid - 1;
  }
  const int sidx = lo;
  const SvdTask t = tasks[sidx];
  const int q = b - (intra ? 2 : 1) * pair_base[sidx];
  const int m = t.m, n = t.n;
  const int p = m < n ? n : m, r = m < n ? m : n;
  const int nb = (r + kBj - 1) / kBj, nbb = nb + (nb & 1);
  if (r > rmax) return;
  int I, J;
  if (intra) {  // the pairs inside block q
    if (q >= nb) return;
    I = q;
    J = nb;
  } else {  // the pairs across blocks I and J of the tournament's round
    if (round >= nbb - 1 || q >= nbb / 2) return;
    I = (round + q) % (nbb - 1);
    J = q == 0 ? nbb - 1 : (round + nbb - 1 - q) % (nbb - 1);
    if (I > J) {
      const int x = I;
      I = J;
      J = x;
    }
    if (J >= nb) return;  // the bye of an odd block count
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = I * kBj, wI = min(kBj, r - i0), j0 = J * kBj, wJ = J < nb ? min(kBj, r - j0) : 0, w = wI + wJ;
  auto colof = [&](int c) { return c < wI ? i0 + c : j0 + (c - wI); };
  const int64_t o_wm = (int64_t)p * r, o_a = 2 * (int64_t)p * r + (int64_t)r * r;
  Mat<C> Wm{W + t.w_off + o_wm, W + t.w_off + o_wm + w_plane, r};
  Mat<C> A{W + t.w_off + o_a, W + t.w_off + o_a + w_plane, r};
  // the block pair's columns of A into LDS (column c at c * r), V = identity
  for (int e = tid; e < w * r; e += kThreads) {
    const int c = e / r, row = e - c * r;
    Lr[c * r + row] = A.r(row, colof(c));
    if (C) Li[c * r + row] = A.m(row, colof(c));
  }
  for (int e = tid; e < w * w; e += kThreads) {
    Vr[e] = (e % w) == (e / w) ? 1.0 : 0.0;
    Vi[e] = 0.0;
  }
  if (tid == 0) any = 0;
  __syncthreads();
  const double tol = sqrt((double)r) * DBL_EPSILON;
  const int ww = w + (w & 1);
  // intra: a tournament over the block's wI columns; across: kBj sub-rounds,
  // wave a pairs column a of I with column (a + sub) mod kBj of J
  const int nsub = intra ? ww - 1 : kBj;
  for (int sub = 0; sub < nsub; ++sub) {
    for (int pq = wave; pq < (intra ? ww / 2 : kBj); pq += kWaves) {
      int i, j;
      if (intra) {
        i = (sub + pq) % (ww - 1);
        j = pq == 0 ? ww - 1 : (sub + ww - 1 - pq) % (ww - 1);
        if (i > j) {
          const int x = i;
          i = j;
          j = x;
        }
      } else {
        i = pq;
        j = wI + (pq + sub) % kBj;
        if (i >= wI) continue;
      }
      if (j >= w) continue;
      double a = 0.0, bb = 0.0, gr = 0.0, gi = 0.0;
      for (int row = lane; row < r; row += 64) {
        const double xir = Lr[i * r + row], xjr = Lr[j * r + row];
        a += xir * xir;
        bb += xjr * xjr;
        gr += xir * xjr;
        if (C) {
          const double xii = Li[i * r + row], xji = Li[j * r + row];
          a += xii * xii;
          bb += xji * xji;
          gr += xii * xji;
          gi += xir * xji - xii * xjr;
        }
      }
      a = wave_sum(a);
      bb = wave_sum(bb);
      gr = wave_sum(gr);
      if (C) gi = wave_sum(gi);
      const double g = C ? sqrt(gr * gr + gi * gi) : fabs(gr);
      if (!(g > tol * sqrt(a * bb)) || a == 0.0 || bb == 0.0) continue;
      const double er = gr / g, ei = C ? gi / g : 0.0;
      const double zeta = (bb - a) / (2.0 * g);
      const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      const double c = 1.0 / sqrt(1.0 + tt * tt), sn = c * tt;
      // a_i' = c a_i - s conj(e) a_j, a_j' = s e a_i + c a_j (as rotate_cols), on the columns and on V
      for (int row = lane; row < r; row += 64) {
        const double xir = Lr[i * r + row], xjr = Lr[j * r + row];
        if (C) {
          const double xii = Li[i * r + row], xji = Li[j * r + row];
          const double cjr = er * xjr + ei * xji, cji = er * xji - ei * xjr;
          const double eir = er * xir - ei * xii, eii = er * xii + ei * xir;
          Lr[i * r + row] = c * xir - sn * cjr;
          Li[i * r + row] = c * xii - sn * cji;
          Lr[j * r + row] = sn * eir + c * xjr;
          Li[j * r + row] = sn * eii + c * xji;
        } else {
          Lr[i * r + row] = c * xir - sn * er * xjr;
          Lr[j * r + row] = sn * er * xir + c * xjr;
        }
      }
      if (lane < w) {
        const int row = lane;
        const double vir = Vr[i * w + row], vjr = Vr[j * w + row];
        if (C) {
          const double vii = Vi[i * w + row], vji = Vi[j * w + row];
          const double cjr = er * vjr + ei * vji, cji = er * vji - ei * vjr;
          const double eir = er * vir - ei * vii, eii = er * vii + ei * vir;
          Vr[i * w + row] = c * vir - sn * cjr;
          Vi[i * w + row] = c * vii - sn * cji;
          Vr[j * w + row] = sn * eir + c * vjr;
          Vi[j * w + row] = sn * eii + c * vji;
        } else {
          Vr[i * w + row] = c * vir - sn * er * vjr;
          Vr[j * w + row] = sn * er * vir + c * vjr;
        }
      }
      if (lane == 0) any = 1;
    }
    __syncthreads();
  }
  if (!any) return;  // uniform: read after the last barrier; nothing changed
  // columns back to A; Wm[:, cols] <- Wm[:, cols] V (V column c: entries c * w + row)
  for (int e = tid; e < w * r; e += kThreads) {
    const int c = e / r, row = e - c * r;
    A.r(row, colof(c)) = Lr[c * r + row];
    if (C) A.m(row, colof(c)) = Li[c * r + row];
  }
  for (int row = tid; row < r; row += kThreads) {
    double orr[2 * kBj], oi[2 * kBj];
#pragma unroll
    for (int c = 0; c < 2 * kBj; ++c) {
      orr[c] = c < w ? Wm.r(row, colof(c)) : 0.0;
      oi[c] = (C && c < w) ? Wm.m(row, colof(c)) : 0.0;
    }
    for (int c = 0; c < w; ++c) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int c2 = 0; c2 < 2 * kBj; ++c2) {
        if (c2 >= w) break;
        const double vr = Vr[c * w + c2], vi = Vi[c * w + c2];  // V[c2, c]
        sr += orr[c2] * vr - (C ? oi[c2] * vi : 0.0);
        if (C) si += orr[c2] * vi + oi[c2] * vr;
      }
      Wm.r(row, colof(c)) = sr;
      if (C) Wm.m(row, colof(c)) = si;
    }
  }
  if (tid == 0) rotated[sidx] = 1.0;
}

// ---------------------------------------------------------------------------
// Batched Householder QR of the same sector matrices, for the gauge steps
// that need an isometry but no Schmidt values: left (right = 0) M = Q R with
// Q (m x k) orthonormal columns, R (k x n) upper trapezoidal; right
// (right = 1) M = L Q' from the QR of M^H = Q R: L = R^H (m x k), Q' = Q^H
// (k x n) orthonormal rows; k = min(m, n).  Written in the SVD's output
// format with S = 1: U = Q (left) or L (right), V^dagger = R (left) or Q'
// (right), so that U S V^dagger = M as for the SVD.
//
// X = M or M^H (p x r) in the workspace; for j < k the reflection
// H_j = I - beta_j v_j v_j^H (v_j at rows j.. of column j of the vector
// store, beta_j = 2 / |v_j|^2, v_j = x - alpha e_j with alpha = -e^{i arg x_j}
// |x|) zeroes column j below the diagonal, applied to columns j+1.. one wave
// per column; then Q = H_0 ... H_{k-1} [I; 0] is accumulated backwards
// (columns j.. of the partial product, as LAPACK's xORGQR does).
// Workspace per task: p r (X) + p k (vectors) + k (betas) elements per plane.
constexpr int kQrThreads = 512;
constexpr int kQrWaves = kQrThreads / 64;

template <bool C>
__global__ __launch_bounds__(kQrThreads) void qr_householder_kernel(const SvdTask* __restrict__ tasks, int right,
                                                                    const double* __restrict__ M, int64_t m_plane,
                                                                    double* __restrict__ U, int64_t u_plane,
                                                                    double* __restrict__ Vt, int64_t vt_plane,
                                                                    double* __restrict__ S, double* __restrict__ W,
                                                                    int64_t w_plane, int phases) {
  __shared__ double red[kQrWaves];
  __shared__ double hv[3];
  __shared__ double vl[2 * kVecRows];
  const SvdTask t = tasks[blockIdx.x];
  const int m = t.m, n = t.n;
  if (m <= 0 || n <= 0) return;
  const int p = right ? n : m, r = right ? m : n, k = p < r ? p : r;
  const int tid = threadIdx.x;
  // X (p x r) at 0, vectors (p x k) at p r, betas at p r + p k (real plane)
  Mat<C> X{W + t.w_off, W + t.w_off + w_plane, p};
  Mat<C> Hv{W + t.w_off + (int64_t)p * r, W + t.w_off + (int64_t)p * r + w_plane, p};
  double* const beta = W + t.w_off + (int64_t)p * r + (int64_t)p * k;
  // phases: kPhLoad, kPhReflect, kPhInit (R out, Q := [I; 0]), kPhApply, kPhOut
  // (the multi-workgroup path runs the reflections and the apply in
  // house_panel/update_kernel and house_apply_kernel between them)
  if (phases & kPhLoad) {
    for (int64_t e = tid; e < (int64_t)p * r; e += kQrThreads) {
      const int i = (int)(e % p), j = (int)(e / p);
      const int64_t src = right ? (j + (int64_t)i * t.lda) : (i + (int64_t)j * t.lda);
      X.r(i, j) = M[t.a_off + src];
      if (C) X.m(i, j) = right ? -M[t.a_off + m_plane + src] : M[t.a_off + m_plane + src];
    }
    __syncthreads();
  }
  if (phases & kPhReflect) house_reflect<C, kQrThreads>(X, Hv, beta, p, r, red, hv, vl);
  // R (k x r) = the upper trapezoid of X; written before X is reused for Q
  double* Ur = U + t.u_off;
  double* Ui = U + t.u_off + u_plane;
  double* Vr = Vt + t.vt_off;
  double* Vi = Vt + t.vt_off + vt_plane;
  if (phases & kPhInit) {
  for (int64_t e = tid; e < (int64_t)k * r; e += kQrThreads) {
    const int i = (int)(e % k), c = (int)(e / k);
    const double rr = i <= c ? X.r(i, c) : 0.0;
    const double ri = (C && i <= c) ? X.m(i, c) : 0.0;
    if (!right) {  // V^dagger (k x n) = R
      Vr[i + (int64_t)c * k] = rr;
      if (C) Vi[i + (int64_t)c * k] = ri;
    } else {  // U (m x k) = R^H: U[c, i] = conj(R[i, c])
      Ur[c + (int64_t)i * m] = rr;
      if (C) Ui[c + (int64_t)i * m] = -ri;
    }
  }
  for (int i = tid; i < k; i += kQrThreads) S[t.s_off + i] = 1.0;
  __syncthreads();
  // Q (p x k) in X's storage: [I; 0], then H_{k-1} .. H_0 applied to columns j..
  for (int64_t e = tid; e < (int64_t)p * k; e += kQrThreads) {
    const int i = (int)(e % p), c = (int)(e / p);
    X.r(i, c) = i == c ? 1.0 : 0.0;
    if (C) X.m(i, c) = 0.0;
  }
  __syncthreads();
  }
  if (phases & kPhApply) house_apply<C, kQrThreads>(X, k, true, Hv, beta, p, k, vl);
  if (!(phases & kPhOut)) return;
  for (int64_t e = tid; e < (int64_t)p * k; e += kQrThreads) {
    const int i = (int)(e % p), c = (int)(e / p);
    if (!right) {  // U (m x k) = Q
      Ur[i + (int64_t)c * m] = X.r(i, c);
      if (C) Ui[i + (int64_t)c * m] = X.m(i, c);
    } else {  // V^dagger (k x n) = Q^H: Vt[c, i] = conj(Q[i, c])
      Vr[c + (int64_t)i * k] = X.r(i, c);
      if (C) Vi[c + (int64_t)i * k] = -X.m(i, c);
    }
  }
}

// ---------------------------------------------------------------------------
// Multi-workgroup Householder reflections for large sectors (p <= kVecRows).
// The one-workgroup kernels above keep a sector on one CU for its whole
// factorisation; here the reflections run in panels of kHp: per panel one
// launch factors the panel's columns (one workgroup per sector,
// house_reflect_range) and one applies its reflections to the later columns,
// one column per wave held in registers through the kHp reflections, the
// panel's vectors in LDS (house_update_kernel); house_apply_kernel forms
// H_0 ... H_{k-1} Y column by column the same way, every column through all
// reflections in one launch.  Workspace layouts (layout): 0 the SVD's (X, Wm,
// vectors, A, betas; p = max(m, n)), 1 / 2 the QR's left / right (X, vectors,
// betas; p = m / n).  Sector s has its 8-column chunks at
// col_base[s] .. col_base[s + 1] of the grid.
constexpr int kHp = 8;
constexpr int kHcols = kThreads / 64;  // columns per workgroup, one per wave
static_assert(kVecRows == kVecRowsMax, "the host routes sectors of at most kVecRowsMax rows to these kernels");

template <bool C>
struct HView {
  int p, r, k;
  Mat<C> X, Hv;
  double* beta;
};

template <bool C>
__device__ __forceinline__ HView<C> house_view(const SvdTask& t, int layout, double* W, int64_t w_plane) {
  HView<C> v;
  const bool sw = layout == 0 ? t.m < t.n : layout == 2;
  v.p = sw ? t.n : t.m;
  v.r = sw ? t.m : t.n;
  v.k = v.p < v.r ? v.p : v.r;
  const int64_t o_hv = layout == 0 ? (int64_t)v.p * v.r + (int64_t)v.r * v.r : (int64_t)v.p * v.r;
  const int64_t o_b = layout == 0 ? 2 * (int64_t)v.p * v.r + 2 * (int64_t)v.r * v.r
                                  : (int64_t)v.p * v.r + (int64_t)v.p * v.k;
  v.X = Mat<C>{W + t.w_off, W + t.w_off + w_plane, v.p};
  v.Hv = Mat<C>{W + t.w_off + o_hv, W + t.w_off + o_hv + w_plane, v.p};
  v.beta = W + t.w_off + o_b;
  return v;
}

__device__ __forceinline__ int task_of(const int* __restrict__ base, int ntasks, int b) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// vectors j0 .. j1 - 1 into LDS, rows 0 .. p - 1, zero above row j
template <bool C>
__device__ __forceinline__ void panel_to_lds(const HView<C>& v, int j0, int j1, double* pl) {
  for (int e = threadIdx.x; e < (j1 - j0) * v.p; e += kThreads) {
    const int jj = e / v.p, i = e - jj * v.p, j = j0 + jj;
    pl[jj * kVecRows + i] = i >= j ? v.Hv.r(i, j) : 0.0;
    if (C) pl[(kHp + jj) * kVecRows + i] = i >= j ? v.Hv.m(i, j) : 0.0;
  }
}

// y <- H_j y for the register-held column y (rows lane + 64 q), v_j in LDS
template <bool C>
__device__ __forceinline__ void reflect_regs_full(double* yr, double* yi, const double* pl, int jj, double b, int p,
                                                  int lane) {
  double wr_ = 0.0, wi_ = 0.0;
#pragma unroll
  for (int q = 0; q < kRpl; ++q) {
    const int i = lane + 64 * q;
    if (i < p) {
      const double v_r = pl[jj * kVecRows + i];
      wr_ += v_r * yr[q];
      if (C) {
        const double v_i = pl[(kHp + jj) * kVecRows + i];
        wr_ += v_i * yi[q];
        wi_ += v_r * yi[q] - v_i * yr[q];
      }
    }
  }
  wr_ = wave_sum(wr_);
  if (C) wi_ = wave_sum(wi_);
  const double fr = b * wr_, fi = b * wi_;
#pragma unroll
  for (int q = 0; q < kRpl; ++q) {
    const int i = lane + 64 * q;
    if (i < p) {
      const double v_r = pl[jj * kVecRows + i];
      const double v_i = C ? pl[(kHp + jj) * kVecRows + i] : 0.0;
      yr[q] -= fr * v_r - fi * v_i;
      if (C) yi[q] -= fr * v_i + fi * v_r;
    }
  }
}

template <bool C>
__device__ __forceinline__ void col_load(const HView<C>& v, int c, double* yr, double* yi, int lane) {
#pragma unroll
  for (int q = 0; q < kRpl; ++q) {
    const int i = lane + 64 * q;
    yr[q] = i < v.p ? v.X.r(i, c) : 0.0;
    yi[q] = (C && i < v.p) ? v.X.m(i, c) : 0.0;
  }
}

template <bool C>
__device__ __forceinline__ void col_store(const HView<C>& v, int c, const double* yr, const double* yi, int lane) {
#pragma unroll
  for (int q = 0; q < kRpl; ++q) {
    const int i = lane + 64 * q;
    if (i < v.p) {
      v.X.r(i, c) = yr[q];
      if (C) v.X.m(i, c) = yi[q];
    }
  }
}

template <bool C>
__global__ __launch_bounds__(kThreads) void house_panel_kernel(const SvdTask* __restrict__ tasks, int layout, int j0,
                                                               double* __restrict__ W, int64_t w_plane) {
  __shared__ double red[kWaves];
  __shared__ double hv[3];
  __shared__ double vl[2 * kVecRows];
  const HView<C> v = house_view<C>(tasks[blockIdx.x], layout, W, w_plane);
  if (j0 >= v.k) return;
  const int j1 = min(j0 + kHp, v.k);
  house_reflect_range<C, kThreads>(v.X, v.Hv, v.beta, v.p, j0, j1, j1, red, hv, vl);
}

template <bool C>
__global__ __launch_bounds__(kThreads) void house_update_kernel(const SvdTask* __restrict__ tasks, int ntasks,
                                                                const int* __restrict__ col_base, int layout, int j0,
                                                                double* __restrict__ W, int64_t w_plane) {
  __shared__ double pl[(C ? 2 : 1) * kHp * kVecRows];
  const int s = task_of(col_base, ntasks, blockIdx.x);
  const HView<C> v = house_view<C>(tasks[s], layout, W, w_plane);
  const int c0 = (blockIdx.x - col_base[s]) * kHcols;
  if (j0 >= v.k) return;
  const int j1 = min(j0 + kHp, v.k);
  if (c0 + kHcols <= j1 || c0 >= v.r) return;  // no column of this workgroup is after the panel
  panel_to_lds<C>(v, j0, j1, pl);
  __syncthreads();
  const int lane = threadIdx.x & 63, c = c0 + (threadIdx.x >> 6);
  if (c < j1 || c >= v.r) return;
  double yr[kRpl], yi[kRpl];
  col_load<C>(v, c, yr, yi, lane);
  for (int j = j0; j < j1; ++j) {
    const double b = v.beta[j];
    if (b != 0.0) reflect_regs_full<C>(yr, yi, pl, j - j0, b, v.p, lane);
  }
  col_store<C>(v, c, yr, yi, lane);
}

// Y <- H_0 ... H_{k-1} Y on X's first ncols columns (layout 0: all r,
// columns from 0; QR layouts: the k columns of [I; 0], reflection j skipped on
// columns before j, which it leaves unchanged, as LAPACK's xORGQR)
template <bool C>
__global__ __launch_bounds__(kThreads) void house_apply_kernel(const SvdTask* __restrict__ tasks, int ntasks,
                                                               const int* __restrict__ col_base, int layout,
                                                               double* __restrict__ W, int64_t w_plane) {
  __shared__ double pl[(C ? 2 : 1) * kHp * kVecRows];
  const int s = task_of(col_base, ntasks, blockIdx.x);
  const HView<C> v = house_view<C>(tasks[s], layout, W, w_plane);
  const bool from_j = layout != 0;
  const int ncols = from_j ? v.k : v.r;
  const int c0 = (blockIdx.x - col_base[s]) * kHcols;
  if (c0 >= ncols) return;
  const int lane = threadIdx.x & 63, c = c0 + (threadIdx.x >> 6);
  const bool mine = c < ncols;
  double yr[kRpl], yi[kRpl];
  if (mine) col_load<C>(v, c, yr, yi, lane);
  const int cmax = min(c0 + kHcols, ncols) - 1;
  for (int j0 = ((v.k - 1) / kHp) * kHp; j0 >= 0; j0 -= kHp) {
    if (from_j && cmax < j0) continue;  // uniform: every column here is before the panel
    const int j1 = min(j0 + kHp, v.k);
    __syncthreads();  // the previous panel's vectors are no longer read
    panel_to_lds<C>(v, j0, j1, pl);
    __syncthreads();
    if (mine)
      for (int j = j1 - 1; j >= j0; --j) {
        const double b = v.beta[j];
        if (b != 0.0 && !(from_j && c < j)) reflect_regs_full<C>(yr, yi, pl, j - j0, b, v.p, lane);
      }
  }
  if (mine) col_store<C>(v, c, yr, yi, lane);
}

hipError_t launch_qr_batched(int cplx, int right, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                             double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                             int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (cplx)
    qr_householder_kernel<true><<<ntasks, kQrThreads, 0, st>>>(tasks, right, M, m_plane, U, u_plane, Vt, vt_plane, S,
                                                               W, w_plane, kPhAll);
  else
    qr_householder_kernel<false><<<ntasks, kQrThreads, 0, st>>>(tasks, right, M, m_plane, U, u_plane, Vt, vt_plane,
                                                                S, W, w_plane, kPhAll);
  return hipGetLastError();
}

hipError_t launch_qr_phase(int cplx, int right, int phases, const SvdTask* tasks, int ntasks, const double* M,
                           int64_t m_plane, double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S,
                           double* W, int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (cplx)
    qr_householder_kernel<true><<<ntasks, kQrThreads, 0, st>>>(tasks, right, M, m_plane, U, u_plane, Vt, vt_plane, S,
                                                               W, w_plane, phases);
  else
    qr_householder_kernel<false><<<ntasks, kQrThreads, 0, st>>>(tasks, right, M, m_plane, U, u_plane, Vt, vt_plane,
                                                                S, W, w_plane, phases);
  return hipGetLastError();
}

hipError_t launch_svd_phase(int cplx, int phases, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                            double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                            int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (cplx)
    svd_jacobi_kernel<true><<<ntasks, kThreads, 0, st>>>(tasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane,
                                                          60, phases);
  else
    svd_jacobi_kernel<false><<<ntasks, kThreads, 0, st>>>(tasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W,
                                                           w_plane, 60, phases);
  return hipGetLastError();
}

hipError_t launch_svd_block_round(int cplx, const SvdTask* tasks, int ntasks, const int* pair_base, int npairs,
                                  int round, double* W, int64_t w_plane, double* rotated, int rmax, hipStream_t st) {
  if (ntasks <= 0 || npairs <= 0) return hipSuccess;
  if (rmax < 1 || rmax > kBjMaxR) return hipErrorInvalidValue;
  static bool attr[2] = {false, false};  // the dynamic LDS limit, raised once per kernel
  const size_t lds = sizeof(double) * (cplx ? 2 : 1) * 2 * kBj * (size_t)rmax;
  const size_t lds_max = sizeof(double) * 2 * 2 * kBj * (size_t)kBjMaxR;
  if (!attr[cplx ? 1 : 0]) {
    const hipError_t e =
        cplx ? hipFuncSetAttribute((const void*)svd_block_round_kernel<true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max)
             : hipFuncSetAttribute((const void*)svd_block_round_kernel<false>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lds_max / 2));
    if (e != hipSuccess) return e;
    attr[cplx ? 1 : 0] = true;
  }
  const int grid = round < 0 ? 2 * npairs : npairs;
  if (cplx)
    svd_block_round_kernel<true><<<grid, kThreads, lds, st>>>(tasks, ntasks, pair_base, round, W, w_plane, rotated,
                                                              rmax);
  else
    svd_block_round_kernel<false><<<grid, kThreads, lds, st>>>(tasks, ntasks, pair_base, round, W, w_plane, rotated,
                                                               rmax);
  return hipGetLastError();
}

hipError_t launch_house_panels(int cplx, int layout, const SvdTask* tasks, int ntasks, const int* col_base,
                               int nchunks, int kmax, double* W, int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0 || nchunks <= 0) return hipSuccess;
  for (int j0 = 0; j0 < kmax; j0 += kHp) {
    if (cplx) {
      house_panel_kernel<true><<<ntasks, kThreads, 0, st>>>(tasks, layout, j0, W, w_plane);
      house_update_kernel<true><<<nchunks, kThreads, 0, st>>>(tasks, ntasks, col_base, layout, j0, W, w_plane);
    } else {
      house_panel_kernel<false><<<ntasks, kThreads, 0, st>>>(tasks, layout, j0, W, w_plane);
      house_update_kernel<false><<<nchunks, kThreads, 0, st>>>(tasks, ntasks, col_base, layout, j0, W, w_plane);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_house_apply(int cplx, int layout, const SvdTask* tasks, int ntasks, const int* col_base,
                              int nchunks, double* W, int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0 || nchunks <= 0) return hipSuccess;
  if (cplx)
    house_apply_kernel<true><<<nchunks, kThreads, 0, st>>>(tasks, ntasks, col_base, layout, W, w_plane);
  else
    house_apply_kernel<false><<<nchunks, kThreads, 0, st>>>(tasks, ntasks, col_base, layout, W, w_plane);
  return hipGetLastError();
}

int house_chunks(int m, int n, int layout) {
  const int r = layout == 0 ? (m < n ? m : n) : (layout == 1 ? n : m);
  return (r + kHcols - 1) / kHcols;
}

int svd_block_pairs(int m, int n) {
  const int r = m < n ? m : n;
  const int nb = (r + kBj - 1) / kBj;
  return (nb + (nb & 1)) / 2;
}
int svd_block_rounds(int m, int n) {
  const int r = m < n ? m : n;
  const int nb = (r + kBj - 1) / kBj;
  return nb + (nb & 1) - 1;
}

hipError_t launch_svd_batched(int cplx, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                              double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                              int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (cplx)
    svd_jacobi_kernel<true><<<ntasks, kThreads, 0, st>>>(tasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane,
                                                          60, kPhAll);
  else
    svd_jacobi_kernel<false><<<ntasks, kThreads, 0, st>>>(tasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W,
                                                           w_plane, 60, kPhAll);
  return hipGetLastError();
}

}  // namespace mpsa
