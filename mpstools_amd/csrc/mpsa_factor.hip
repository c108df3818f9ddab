// mpsa_factor.hip -- batched thin SVD of many small charge-sector matrices on
// gfx950, one workgroup per sector: the SVD behind prepareSite / prepareBlock
// (src/MPSfunctions.hpp:21-140, BoostSVD + FixPhase blas/lapack_routines.cpp:
// 614-634) for MPS canonicalisation and truncation, where a site tensor splits
// into hundreds of sectors of tens of rows each.
//
// Method: one-sided Jacobi (Hestenes) on the columns of X = M (m >= n) or
// X = M^H (m < n), p x r with p >= r.  Column pairs are visited in the
// round-robin order of a tournament (r/2 disjoint pairs per round, r - 1
// rounds per sweep), one wave per pair, so a wave needs no barrier inside a
// rotation; the workgroup synchronises between rounds.  A pair is rotated when
// |x_i^H x_j| > tol sqrt(|x_i|^2 |x_j|^2) with tol = sqrt(p) eps; sweeps stop
// when a whole sweep rotates nothing.  The 2 x 2 Hermitian Jacobi rotation
// with gamma = |gamma| e^{i phi}, zeta = (b - a) / (2 |gamma|),
// t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), c = 1 / sqrt(1 + t^2), s = c t:
//   x_i' = c x_i - s e^{-i phi} x_j,   x_j' = s e^{i phi} x_i + c x_j
// makes the pair orthogonal; V accumulates the same rotations.  Then
// sigma_j = |x_j|, sorted descending (ties by column index), and
//   m >= n: U = X / sigma, V^dagger = V^H;   m < n: U = V, V^dagger = (X / sigma)^H.
// Singular vectors of exactly zero singular values are completed to an
// orthonormal basis (Gram-Schmidt of unit vectors, twice), as a LAPACK thin
// SVD returns them.  Real data: FixPhase makes the first largest-|.| entry of
// every row of V^dagger positive (the column of U flips with it).
//
// Complex data is planar (real parts, then imaginary parts at a plane offset),
// like every complex buffer of the library.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "mpsa_internal.h"

namespace mpsa {

namespace {

constexpr int kThreads = 256;
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

}  // namespace

template <bool C>
__global__ __launch_bounds__(kThreads) void svd_jacobi_kernel(const SvdTask* __restrict__ tasks,
                                                              const double* __restrict__ M, int64_t m_plane,
                                                              double* __restrict__ U, int64_t u_plane,
                                                              double* __restrict__ Vt, int64_t vt_plane,
                                                              double* __restrict__ S, double* __restrict__ W,
                                                              int64_t w_plane, int max_sweeps) {
  __shared__ double sig[kSvdMaxR];
  __shared__ int rank[kSvdMaxR];
  __shared__ int rotated;
  __shared__ double red[kWaves], redc[kWaves];
  const SvdTask t = tasks[blockIdx.x];
  const int m = t.m, n = t.n;
  if (m <= 0 || n <= 0) return;
  const bool wide = m < n;
  const int p = wide ? n : m, r = wide ? m : n, k = r;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  Mat<C> X{W + t.w_off, W + t.w_off + w_plane, p};
  Mat<C> V{W + t.w_off + (int64_t)p * r, W + t.w_off + (int64_t)p * r + w_plane, r};
  // X = M or M^H, V = identity
  for (int64_t e = tid; e < (int64_t)p * r; e += kThreads) {
    const int i = (int)(e % p), j = (int)(e / p);
    const int64_t src = wide ? (j + (int64_t)i * t.lda) : (i + (int64_t)j * t.lda);
    X.r(i, j) = M[t.a_off + src];
    if (C) X.m(i, j) = wide ? -M[t.a_off + m_plane + src] : M[t.a_off + m_plane + src];
  }
  for (int64_t e = tid; e < (int64_t)r * r; e += kThreads) {
    const int i = (int)(e % r), j = (int)(e / r);
    V.r(i, j) = i == j ? 1.0 : 0.0;
    if (C) V.m(i, j) = 0.0;
  }
  __syncthreads();
  const double tol = sqrt((double)p) * DBL_EPSILON;
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
        double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
        for (int row = lane; row < p; row += 64) {
          const double xir = X.r(row, i), xjr = X.r(row, j);
          a += xir * xir;
          b += xjr * xjr;
          gr += xir * xjr;
          if (C) {
            const double xii = X.m(row, i), xji = X.m(row, j);
            a += xii * xii;
            b += xji * xji;
            gr += xii * xji;
            gi += xir * xji - xii * xjr;  // x_i^H x_j = sum conj(x_i) x_j
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
        rotate_cols<C>(X, p, i, j, c, s, er, ei, lane);
        rotate_cols<C>(V, r, i, j, c, s, er, ei, lane);
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  // singular values and their descending order
  for (int j = wave; j < r; j += kWaves) {
    double a = 0.0;
    for (int row = lane; row < p; row += 64) {
      const double x = X.r(row, j);
      a += x * x;
      if (C) {
        const double y = X.m(row, j);
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
  // U (m x k, ld m) and V^dagger (k x n, ld k); "vec" = the side built from X
  double* Ur = U + t.u_off;
  double* Ui = U + t.u_off + u_plane;
  double* Vr = Vt + t.vt_off;
  double* Vi = Vt + t.vt_off + vt_plane;
  for (int64_t e = tid; e < (int64_t)p * r; e += kThreads) {
    const int row = (int)(e % p), j = (int)(e / p);
    const int rk = rank[j];
    const double sj = sig[j];
    const double xr = sj > 0.0 ? X.r(row, j) / sj : 0.0;
    const double xi = (C && sj > 0.0) ? X.m(row, j) / sj : 0.0;
    if (!wide) {  // U[:, rk] = x_j / sigma_j
      Ur[row + (int64_t)rk * m] = xr;
      if (C) Ui[row + (int64_t)rk * m] = xi;
    } else {  // V^dagger[rk, :] = (x_j / sigma_j)^H
      Vr[rk + (int64_t)row * k] = xr;
      if (C) Vi[rk + (int64_t)row * k] = -xi;
    }
  }
  for (int64_t e = tid; e < (int64_t)r * r; e += kThreads) {
    const int row = (int)(e % r), j = (int)(e / r);
    const int rk = rank[j];
    if (!wide) {  // V^dagger[rk, row] = conj(V[row, j])
      Vr[rk + (int64_t)row * k] = V.r(row, j);
      if (C) Vi[rk + (int64_t)row * k] = -V.m(row, j);
    } else {  // U[row, rk] = V[row, j]
      Ur[row + (int64_t)rk * m] = V.r(row, j);
      if (C) Ui[row + (int64_t)rk * m] = V.m(row, j);
    }
  }
  __syncthreads();
  // orthonormal completion of the vectors of zero singular values (descending
  // order puts them last): vectors of length p stored at base + idx * vs with
  // stride es between elements (U columns: es = 1, vs = m; V^dagger rows:
  // es = k, vs = 1).  The conjugate of a V^dagger row is the vector.
  const int64_t es = wide ? k : 1, vs = wide ? 1 : m;
  double* Br = wide ? Vr : Ur;
  double* Bi = wide ? Vi : Ui;
  const double csign = wide ? -1.0 : 1.0;  // stored imaginary part = csign * vector imaginary part
  for (int rk = 0; rk < k; ++rk) {
    if (S[t.s_off + rk] > 0.0) continue;
    for (int cand = 0; cand < p; ++cand) {
      // w = e_cand, then two passes of w -= (u_i^H w) u_i over i < rk
      for (int row = tid; row < p; row += kThreads) {
        Br[rk * vs + row * es] = row == cand ? 1.0 : 0.0;
        if (C) Bi[rk * vs + row * es] = 0.0;
      }
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < rk; ++i) {
          double dr = 0.0, di = 0.0;
          for (int row = tid; row < p; row += kThreads) {
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
          for (int row = tid; row < p; row += kThreads) {
            const double ur = Br[i * vs + row * es];
            const double ui = C ? csign * Bi[i * vs + row * es] : 0.0;
            // w -= (Dr + i Di) u
            Br[rk * vs + row * es] -= Dr * ur - Di * ui;
            if (C) Bi[rk * vs + row * es] -= csign * (Dr * ui + Di * ur);
          }
          __syncthreads();
        }
      double nn = 0.0;
      for (int row = tid; row < p; row += kThreads) {
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
        for (int row = tid; row < p; row += kThreads) {
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

hipError_t launch_svd_batched(int cplx, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                              double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                              int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (cplx)
    svd_jacobi_kernel<true><<<ntasks, kThreads, 0, st>>>(tasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane,
                                                          60);
  else
    svd_jacobi_kernel<false><<<ntasks, kThreads, 0, st>>>(tasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W,
                                                           w_plane, 60);
  return hipGetLastError();
}

}  // namespace mpsa
