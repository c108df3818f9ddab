// mpsa_split.cpp -- split of a two-site tensor into two site tensors and the
// singular values between them, on the device:
//   split (src/MPSfunctions.hpp:1142-1278) with BoostSVD + FixPhase
//   (blas/lapack_routines.cpp:614-634, blas/helperfunctions.hpp:36-67), the
//   truncations of DiagBlockMatrix (src/BlockTensor.hpp:85-136, 234-276),
//   dropDangling (src/MPSfunctions.hpp:1444-1484) and the multiplication of
//   lambda into one side done by DMRGengine::doStep (src/container.hpp:321-328).
//
// Per centre charge qc = l + q(s1) the blocks (l, r, (s1, s2)) are gathered
// into one dense matrix (rows by s1 ascending, columns by s2 ascending), its
// SVD is taken in FP64 by Jacobi (rocSOLVER gesvdj: the factorisation of the
// reference's dgesvd, backward stable), the sign of every singular pair is
// fixed by FixPhase, and the global truncation runs on the host over all
// sectors' singular values exactly as the reference orders them.
// Sectors with min(m, n) >= kPrecondMin (the dense D = 2048 bond's 4096 x 4096
// sector) of a count-limited split (chimax below the number of singular
// values) take the density-matrix eigendecomposition instead: with E the
// eigenvectors (rocSOLVER syevd) of the Gram matrix of the side that stays a
// bare isometry (M M^T when lambda goes right, M^T M when it goes left),
// P = M^T E (or M E) holds sigma_e times the other side's singular vector in
// column e, and sigma_e = |P[:, e]|.  The isometry is E itself (orthonormal to
// syevd's rounding) and the lambda-carrying tensor is P (M = E P^T exactly, up
// to E's orthogonality).  One level resolves sigma_e to SVD accuracy only
// down to about sqrt(eps) sigma_max, since the Gram matrix squares them; so
// the directions with eigenvalues below kDmTail^2 lambda_max are
// diagonalised a second time from their own block Pc of P (Pc^T Pc has
// entries of the tail's size, not lambda_max's): E_c <- E_c E2, Pc <- Pc E2.
// What then limits the tail is its contamination by the large directions,
// about eps sigma_max^2 / (kDmTail sigma_max) absolute (2e-12 sigma_max).
// Directions outside M's range come out near eps sigma_max^2 / sigma_r
// (sigma_r the smallest nonzero singular value), not an SVD's
// eps sigma_max, so values below kDmZero = 1e-10 of the largest count as
// zero: the rank holds while sigma_r > 2e-6 sigma_max, and a real singular
// value below that level (weight under 1e-20) is dropped where the reference
// would keep it.  An untruncated split, whose kept count is the number of
// values above 1e-15, keeps gesvdj, which reveals the rank.  The route
// replaced Jacobi (gesvdj) on B = M V_e after the same syevd, whose stopping
// test bounds the off-diagonal of B^T B absolutely (backward error 1.6e-9
// relative on the 4096 x 4096 sector) and which took 2.1 of the split's 2.9 s
// there.

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

using namespace mpsa;

namespace {

struct SErr : std::runtime_error {
  int code;
  SErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void hchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw SErr(MPSA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}
void rchk(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw SErr(MPSA_EHIP, std::string(what) + ": rocBLAS/rocSOLVER status " + std::to_string((int)s));
}
int64_t align64(int64_t x) { return (x + 63) / 64 * 64; }
// smallest min(m, n) of a sector split by the density-matrix route (below it
// the direct Jacobi SVD is cheap and converges to rounding level); the CPU
// emulation of the device layer builds with a smaller value so that its tests
// reach this path at small sizes
#ifndef MPSA_PRECOND_MIN
#define MPSA_PRECOND_MIN 1024
#endif
constexpr int64_t kPrecondMin = MPSA_PRECOND_MIN;
// the density-matrix route's zero level, relative to the sector's largest
// singular value, and the level below which its second stage re-resolves
// the singular values
constexpr double kDmZero = 1e-10;
constexpr double kDmTail = 1e-4;

// one rocBLAS handle per device, created on first use
rocblas_handle handle_for_current_device() {
  static std::map<int, rocblas_handle> handles;
  int dev = 0;
  hchk(hipGetDevice(&dev), "hipGetDevice");
  auto it = handles.find(dev);
  if (it != handles.end()) return it->second;
  rocblas_handle h = nullptr;
  rchk(rocblas_create_handle(&h), "rocblas_create_handle");
  handles[dev] = h;
  return h;
}

template <typename T>
struct DevArr {
  T* p = nullptr;
  void alloc(size_t n) {
    if (hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
      p = nullptr;
      throw SErr(MPSA_ENOMEM, "split: device allocation failed");
    }
  }
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
};

struct Sector {
  uint64_t qc;
  std::map<uint32_t, std::pair<int64_t, int64_t>> rx;  // s1 -> (row start, rows)
  std::map<uint32_t, std::pair<int64_t, int64_t>> ry;  // s2 -> (col start, cols)
  std::vector<int64_t> blocks;                          // psi block indices
  int64_t m = 0, n = 0, k = 0;
  int64_t off_M = 0, off_U = 0, off_V = 0, off_S = 0;
  int chi = 0;  // kept after truncation
};

}  // namespace

void* mpsa::rocblas_handle_current() { return (void*)handle_for_current_device(); }

struct mpsa_split_result {
  std::vector<uint64_t> lq, rq;
  std::vector<uint32_t> lp, rp;
  std::vector<int64_t> lrows, lcols, loff, rrows, rcols, roff;
  int64_t ltotal = 0, rtotal = 0;
  double* dleft = nullptr;
  double* dright = nullptr;
  std::vector<uint64_t> lam_q;
  std::vector<int64_t> lam_n;
  std::vector<double> lam;
  double tw = 0.0;
  int chi = 0;
  ~mpsa_split_result() {
    if (dleft) (void)hipFree(dleft);
    if (dright) (void)hipFree(dright);
  }
};

extern "C" {

int mpsa_split(const mpsa_blocks* psi, const double* dpsi, const uint64_t* itokey_l, int dl, const uint64_t* itokey_r,
               int dr, int direction, unsigned chimax, int do_trunc, double tr_weight, int normalize,
               mpsa_split_result** out, mpsa_stream stream) {
  mpsa_split_result* R = nullptr;
  try {
    if (!psi || !out || !itokey_l || !itokey_r || dl < 1 || dr < 1 || psi->ns != 2 || psi->is_complex)
      throw SErr(MPSA_EINVAL, "split: bad arguments");
    if (psi->n > 0 && (!dpsi || !psi->q || !psi->p || !psi->rows || !psi->cols || !psi->offset))
      throw SErr(MPSA_EINVAL, "split: null table array");
    hipStream_t st = (hipStream_t)stream;
    // sectors by centre charge (split :1160-1170), ranges by s1 / s2 (:1172-1192)
    std::map<uint64_t, Sector> secs;
    for (int64_t b = 0; b < psi->n; ++b) {
      const uint32_t s1 = psi->p[2 * b], s2 = psi->p[2 * b + 1];
      if ((int)s1 >= dl || (int)s2 >= dr) throw SErr(MPSA_EINVAL, "split: physical index outside i_to_key");
      const uint64_t qc = psi->q[2 * b] + itokey_l[s1];
      Sector& S = secs[qc];
      S.qc = qc;
      S.blocks.push_back(b);
      S.rx[s1].second = psi->rows[b];
      S.ry[s2].second = psi->cols[b];
    }
    int64_t tot_M = 0, tot_U = 0, tot_V = 0, tot_S = 0, maxk = 0;
    for (auto& kv : secs) {
      Sector& S = kv.second;
      for (auto& r : S.rx) {
        r.second.first = S.m;
        S.m += r.second.second;
      }
      for (auto& c : S.ry) {
        c.second.first = S.n;
        S.n += c.second.second;
      }
      S.k = std::min(S.m, S.n);
      S.off_M = tot_M;
      tot_M += align64(S.m * S.n);
      S.off_U = tot_U;
      tot_U += align64(S.m * S.k);
      S.off_V = tot_V;
      tot_V += align64(S.k * S.n);
      S.off_S = tot_S;
      tot_S += S.k;
      maxk = std::max(maxk, S.k);
    }
    // gather the dense sector matrices (:1200-1210)
    DevArr<double> M, U, V, Sg, scales, resid;
    DevArr<rocblas_int> sweeps;
    DevArr<rocblas_int> info;
    M.alloc((size_t)tot_M);
    U.alloc((size_t)tot_U);
    V.alloc((size_t)tot_V);
    Sg.alloc((size_t)tot_S);
    info.alloc(1);
    resid.alloc(1);
    sweeps.alloc(1);
    hchk(launch_fill(M.p, std::max<int64_t>(tot_M, 1), 0.0, st), "split: zero");
    std::vector<Copy2D> gather;
    int64_t maxel = 0;
    for (auto& kv : secs) {
      Sector& S = kv.second;
      for (int64_t b : S.blocks) {
        const auto& rx = S.rx[psi->p[2 * b]];
        const auto& ry = S.ry[psi->p[2 * b + 1]];
        Copy2D t{};
        t.src_off = psi->offset[b];
        t.dst_off = S.off_M + ry.first * S.m + rx.first;
        t.rows = (int32_t)psi->rows[b];
        t.cols = (int32_t)psi->cols[b];
        t.lds = (int32_t)std::max<int64_t>(psi->rows[b], 1);
        t.ldd = (int32_t)std::max<int64_t>(S.m, 1);
        t.rs_off = t.cs_off = -1;
        t.scale = 1.0;
        if ((int64_t)t.rows * t.cols > 0) gather.push_back(t);
        maxel = std::max(maxel, (int64_t)t.rows * t.cols);
      }
    }
    DevArr<Copy2D> dg;
    dg.alloc(gather.size());
    if (!gather.empty()) {
      hchk(hipMemcpyAsync(dg.p, gather.data(), sizeof(Copy2D) * gather.size(), hipMemcpyHostToDevice, st), "split");
      hchk(launch_copy2d(dg.p, (int)gather.size(), maxel, dpsi, M.p, nullptr, st), "split: gather");
    }
    // SVD + FixPhase per sector (BoostSVD, lapack_routines.cpp:614-634)
    rocblas_handle h = handle_for_current_device();
    rchk(rocblas_set_stream(h, st), "rocblas_set_stream");
    // Sectors with min(m, n) >= kPrecondMin take the density-matrix route:
    // the Gram matrix of the side that stays a bare isometry (rho = M M^T
    // when lambda goes right, G = M^T M when it goes left) is diagonalised by
    // syevd, and P = M^T E (or M E) carries sigma_e times the other side's
    // vector in its column e.  sigma_e = |P[:, e]| is the Rayleigh quotient of
    // the exact Gram at the computed eigenvector, so its error is that of the
    // product P (eps sigma_max absolute) plus a second-order term in the
    // eigenvector's angle; the isometry is syevd's orthonormal E, and the
    // lambda-carrying tensor is P itself (the division by sigma below is
    // undone when lambda is multiplied back in).
    // The route resolves singular values down to about sqrt(eps) sigma_max,
    // not the rank of a rank-deficient sector, so it is taken only when the
    // split is count-limited (chimax below the number of singular values):
    // its kept set is then the top chimax values, while an untruncated split,
    // whose kept count is the number of values above 1e-15, keeps gesvdj.
    unsigned cm = chimax;
    if (!cm) cm = do_trunc ? 100u : 200u;  // split :1257-1258
    const bool count_limited = (int64_t)cm < tot_S;
    auto dm_route = [&](const Sector& S) { return count_limited && std::min(S.m, S.n) >= kPrecondMin; };
    int64_t big_ss = 0, big_mn = 0, big_s = 0;
    const bool left_bare = direction > 0;
    for (auto& kv : secs) {
      const Sector& S = kv.second;
      if (dm_route(S)) {
        const int64_t sd = left_bare ? S.m : S.n;
        big_ss = std::max(big_ss, sd * sd);
        big_mn = std::max(big_mn, S.m * S.n);
        big_s = std::max(big_s, sd);
      }
    }
    // Ge: the Gram matrix, then its eigenvectors E; Pw: P; ev/ework: syevd's
    // eigenvalues and workspace; nrm: the column norms of P
    // Gc, T1, T2: the second level's Gram matrix and products
    DevArr<double> Ge, Pw, ev, ework, nrm, Gc, T1, T2;
    if (big_s) {
      Ge.alloc((size_t)big_ss);
      Pw.alloc((size_t)big_mn);
      ev.alloc((size_t)big_s);
      ework.alloc((size_t)big_s);
      nrm.alloc((size_t)big_s);
      Gc.alloc((size_t)big_ss);
      T1.alloc((size_t)big_mn);
      T2.alloc((size_t)big_ss);
    }
    auto check_info = [&](const char* what) {
      rocblas_int ih = 0;
      hchk(hipMemcpyAsync(&ih, info.p, sizeof(ih), hipMemcpyDeviceToHost, st), "split: info");
      hchk(hipStreamSynchronize(st), "split: sync");
      if (ih != 0) throw SErr(MPSA_EINVAL, std::string("split: ") + what + " did not converge (info " + std::to_string(ih) + ")");
    };
    const double one = 1.0, zero = 0.0;
    std::vector<std::vector<double>> big_sig;    // host sources of the sorted sigmas, alive until the final sync
    std::vector<std::vector<Copy2D>> big_moves;  // host sources of the gather tasks, likewise
    std::vector<std::unique_ptr<DevArr<Copy2D>>> big_dmoves;
    for (auto& kv : secs) {
      Sector& S = kv.second;
      if (S.k == 0) continue;
      const rocblas_int m = (rocblas_int)S.m, n = (rocblas_int)S.n, k = (rocblas_int)S.k;
      double* Ms = M.p + S.off_M;
      if (dm_route(S)) {
        const rocblas_int sd = left_bare ? m : n;     // order of the Gram matrix
        const int64_t prow = left_bare ? S.n : S.m;  // rows of P
        if (left_bare)
          rchk(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, m, m, n, &one, Ms, m, Ms, m,
                             &zero, Ge.p, m),
               "split: M M^T");
        else
          rchk(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, n, n, m, &one, Ms, m, Ms, m,
                             &zero, Ge.p, n),
               "split: M^T M");
        rchk(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, sd, Ge.p, sd, ev.p, ework.p, info.p),
             "rocsolver_dsyevd");
        check_info("syevd");
        if (left_bare)  // P = M^T E (n x m)
          rchk(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, n, m, m, &one, Ms, m, Ge.p, m,
                             &zero, Pw.p, n),
               "split: M^T E");
        else  // P = M E (m x n)
          rchk(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, n, &one, Ms, m, Ge.p, n, &zero,
                             Pw.p, m),
               "split: M E");
        // second level: the directions with eigenvalues below kDmTail^2
        // lambda_max (the first sc columns, syevd's order is ascending) are
        // diagonalised again from Pc = P[:, 0:sc] alone, whose Gram matrix
        // Pc^T Pc has entries of their own size, not lambda_max's
        std::vector<double> hev((size_t)sd);
        hchk(hipMemcpyAsync(hev.data(), ev.p, sizeof(double) * (size_t)sd, hipMemcpyDeviceToHost, st),
             "split: eigenvalues");
        hchk(hipStreamSynchronize(st), "split: sync");
        rocblas_int sc = 0;
        while (sc < sd && hev[(size_t)sc] < kDmTail * kDmTail * hev[(size_t)sd - 1]) ++sc;
        if (sc >= 2) {
          rchk(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, sc, sc, (rocblas_int)prow, &one,
                             Pw.p, (rocblas_int)prow, Pw.p, (rocblas_int)prow, &zero, Gc.p, sc),
               "split: Pc^T Pc");
          rchk(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, sc, Gc.p, sc, ev.p, ework.p, info.p),
               "rocsolver_dsyevd");
          check_info("syevd");
          rchk(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)prow, sc, sc, &one, Pw.p,
                             (rocblas_int)prow, Gc.p, sc, &zero, T1.p, (rocblas_int)prow),
               "split: Pc E2");
          rchk(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, sd, sc, sc, &one, Ge.p, sd, Gc.p, sc,
                             &zero, T2.p, sd),
               "split: Ec E2");
          hchk(hipMemcpyAsync(Pw.p, T1.p, sizeof(double) * (size_t)prow * sc, hipMemcpyDeviceToDevice, st),
               "split: tail");
          hchk(hipMemcpyAsync(Ge.p, T2.p, sizeof(double) * (size_t)sd * sc, hipMemcpyDeviceToDevice, st),
               "split: tail");
        }
        hchk(launch_colnorms(Pw.p, prow, sd, prow, nrm.p, st), "split: sigma");
        std::vector<double> hn((size_t)sd);
        hchk(hipMemcpyAsync(hn.data(), nrm.p, sizeof(double) * (size_t)sd, hipMemcpyDeviceToHost, st), "split: sigma");
        hchk(hipStreamSynchronize(st), "split: sync");
        // rank: a direction outside the range of M gets |P[:, e]| of about
        // eps sigma_max^2 / sigma_r (sigma_r its smallest nonzero singular
        // value), not the eps sigma_max of an SVD, so values up to
        // kDmZero sigma_max count as zero (they then fall under the 1e-15 cut)
        double smax = 0.0;
        for (double x : hn) smax = std::max(smax, x);
        const double zero_below = kDmZero * smax;
        for (double& x : hn)
          if (x <= zero_below) x = 0.0;
        // descending sigma; among equal values the larger eigenvalue index (syevd: ascending) first
        std::vector<int> ord((size_t)sd);
        for (int e = 0; e < sd; ++e) ord[(size_t)e] = sd - 1 - e;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return hn[(size_t)x] > hn[(size_t)y]; });
        big_sig.emplace_back((size_t)k);
        std::vector<double>& sg = big_sig.back();
        std::vector<Copy2D> fromE, fromP;
        for (int rk = 0; rk < k; ++rk) {
          const int e = ord[(size_t)rk];
          const double sig = hn[(size_t)e];
          sg[(size_t)rk] = sig;
          Copy2D ce{}, cp{};
          ce.rs_off = ce.cs_off = cp.rs_off = cp.cs_off = -1;
          ce.scale = 1.0;
          cp.scale = sig > 0.0 ? 1.0 / sig : 0.0;
          if (left_bare) {  // U[:, rk] = E[:, e]; V^T[rk, :] = P[:, e]^T / sigma
            ce.src_off = (int64_t)e * m;
            ce.dst_off = S.off_U + (int64_t)rk * m;
            ce.rows = m;
            ce.cols = 1;
            ce.lds = ce.ldd = m;
            cp.src_off = (int64_t)e * n;
            cp.dst_off = S.off_V + rk;
            cp.rows = 1;
            cp.cols = n;
            cp.lds = 1;
            cp.ldd = k;
          } else {  // U[:, rk] = P[:, e] / sigma; V^T[rk, :] = E[:, e]^T
            cp.src_off = (int64_t)e * m;
            cp.dst_off = S.off_U + (int64_t)rk * m;
            cp.rows = m;
            cp.cols = 1;
            cp.lds = cp.ldd = m;
            ce.src_off = (int64_t)e * n;
            ce.dst_off = S.off_V + rk;
            ce.rows = 1;
            ce.cols = n;
            ce.lds = 1;
            ce.ldd = k;
          }
          fromE.push_back(ce);
          fromP.push_back(cp);
        }
        hchk(hipMemcpyAsync(Sg.p + S.off_S, sg.data(), sizeof(double) * (size_t)k, hipMemcpyHostToDevice, st),
             "split: sigma");
        for (int src = 0; src < 2; ++src) {
          big_moves.push_back(src == 0 ? fromE : fromP);
          const std::vector<Copy2D>& mv = big_moves.back();
          big_dmoves.emplace_back(new DevArr<Copy2D>());
          big_dmoves.back()->alloc(mv.size());
          hchk(hipMemcpyAsync(big_dmoves.back()->p, mv.data(), sizeof(Copy2D) * mv.size(), hipMemcpyHostToDevice, st),
               "split: gather");
          const bool to_u = (src == 0) == left_bare;
          hchk(launch_copy2d(big_dmoves.back()->p, (int)mv.size(), to_u ? S.m : S.n, src == 0 ? Ge.p : Pw.p,
                             to_u ? U.p : V.p, nullptr, st),
               "split: gather");
        }
      } else {
        // abstol 0: converge to machine precision; 100 sweeps is far above what
        // Jacobi needs (quadratic convergence), info reports a miss
        rchk(rocsolver_dgesvdj(h, rocblas_svect_singular, rocblas_svect_singular, m, n, Ms, m, 0.0, resid.p, 100,
                               sweeps.p, Sg.p + S.off_S, U.p + S.off_U, m, V.p + S.off_V, k, info.p),
             "rocsolver_dgesvdj");
        check_info("gesvdj");
      }
      hchk(launch_fixphase(U.p + S.off_U, S.m, S.m, V.p + S.off_V, S.k, S.n, S.k, st), "split: FixPhase");
    }
    std::vector<double> sv((size_t)std::max<int64_t>(tot_S, 1));
    if (tot_S) hchk(hipMemcpyAsync(sv.data(), Sg.p, sizeof(double) * tot_S, hipMemcpyDeviceToHost, st), "split: S");
    hchk(hipStreamSynchronize(st), "split: sync");

    // global truncation over all sectors (src/BlockTensor.hpp:85-136, 234-276):
    // multimap<sigma, qc> filled sector by sector in key order, visited from
    // the largest value (reverse iteration: among equal values the last
    // inserted first).
    std::vector<std::pair<double, uint64_t>> all;
    for (auto& kv : secs)
      for (int64_t i = 0; i < kv.second.k; ++i) all.push_back({sv[(size_t)(kv.second.off_S + i)], kv.first});
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<double, uint64_t>& a, const std::pair<double, uint64_t>& b) {
                       return a.first < b.first;
                     });
    R = new mpsa_split_result();
    std::map<uint64_t, std::vector<double>> kept;
    double tw;
    int cnt = 0;
    if (!do_trunc) {
      tw = 1.0;
      for (auto r = all.rbegin(); r != all.rend() && (unsigned)cnt != cm && std::fabs(r->first) > 1e-15; ++r) {
        tw -= r->first * r->first;
        kept[r->second].push_back(r->first);
        ++cnt;
      }
    } else {
      double nrm = 0.0;
      for (auto& a : all) nrm += a.first * a.first;
      nrm = std::sqrt(nrm);
      tw = nrm;
      for (auto r = all.rbegin(); r != all.rend() && std::fabs(tw) > nrm * tr_weight && (unsigned)cnt < cm; ++r) {
        tw -= r->first * r->first / nrm;
        kept[r->second].push_back(r->first);
        ++cnt;
      }
      if (tw < 0 && std::fabs(tw) < 1e-14) tw = 0.0;
    }
    double lnorm = 0.0;
    for (auto& kv : kept)
      for (double x : kv.second) lnorm += x * x;
    lnorm = std::sqrt(lnorm);
    const double lscale = (normalize && lnorm > 0) ? 1.0 / lnorm : 1.0;
    R->tw = tw;
    R->chi = cnt;
    for (auto& kv : kept) {
      R->lam_q.push_back(kv.first);
      R->lam_n.push_back((int64_t)kv.second.size());
      for (double x : kv.second) R->lam.push_back(x * lscale);
      secs[kv.first].chi = (int)kv.second.size();
    }
    // output tables (dropDangling keeps the leading chi columns / rows)
    std::vector<Copy2D> lt, rt;
    std::vector<double> scal;  // lambda per kept sector, for the side it multiplies
    std::map<uint64_t, int64_t> lam_off;
    {
      size_t o = 0;
      for (size_t i = 0; i < R->lam_q.size(); ++i) {
        lam_off[R->lam_q[i]] = (int64_t)o;
        o += (size_t)R->lam_n[i];
      }
      scal = R->lam;
    }
    // left keys (qc - q(s1), qc, s1), right keys (qc, qc + q(s2), s2), both in key order
    struct LK {
      uint64_t q0, q1;
      uint32_t p;
      const Sector* S;
      bool operator<(const LK& o) const {
        if (q0 != o.q0) return q0 < o.q0;
        if (q1 != o.q1) return q1 < o.q1;
        return p < o.p;
      }
    };
    std::vector<LK> lkeys, rkeys;
    for (auto& kv : secs) {
      const Sector& S = kv.second;
      if (S.chi == 0) continue;
      for (auto& r : S.rx) lkeys.push_back(LK{S.qc - itokey_l[r.first], S.qc, r.first, &S});
      for (auto& c : S.ry) rkeys.push_back(LK{S.qc, S.qc + itokey_r[c.first], c.first, &S});
    }
    std::sort(lkeys.begin(), lkeys.end());
    std::sort(rkeys.begin(), rkeys.end());
    int64_t o = 0, maxl = 0, maxr = 0;
    for (auto& k : lkeys) {
      const Sector& S = *k.S;
      const auto& rx = S.rx.at(k.p);
      o = align64(o);
      R->lq.push_back(k.q0);
      R->lq.push_back(k.q1);
      R->lp.push_back(k.p);
      R->lrows.push_back(rx.second);
      R->lcols.push_back(S.chi);
      R->loff.push_back(o);
      Copy2D t{};
      t.src_off = S.off_U + rx.first;
      t.dst_off = o;
      t.rows = (int32_t)rx.second;
      t.cols = S.chi;
      t.lds = (int32_t)std::max<int64_t>(S.m, 1);
      t.ldd = (int32_t)std::max<int64_t>(rx.second, 1);
      t.rs_off = -1;
      t.cs_off = direction <= 0 ? lam_off[S.qc] : -1;  // U * lambda
      t.scale = 1.0;
      if ((int64_t)t.rows * t.cols > 0) lt.push_back(t);
      maxl = std::max(maxl, (int64_t)t.rows * t.cols);
      o += rx.second * S.chi;
    }
    R->ltotal = o;
    o = 0;
    for (auto& k : rkeys) {
      const Sector& S = *k.S;
      const auto& ry = S.ry.at(k.p);
      o = align64(o);
      R->rq.push_back(k.q0);
      R->rq.push_back(k.q1);
      R->rp.push_back(k.p);
      R->rrows.push_back(S.chi);
      R->rcols.push_back(ry.second);
      R->roff.push_back(o);
      Copy2D t{};
      t.src_off = S.off_V + ry.first * S.k;
      t.dst_off = o;
      t.rows = S.chi;
      t.cols = (int32_t)ry.second;
      t.lds = (int32_t)std::max<int64_t>(S.k, 1);
      t.ldd = std::max(S.chi, 1);
      t.rs_off = direction > 0 ? lam_off[S.qc] : -1;  // lambda * V'
      t.cs_off = -1;
      t.scale = 1.0;
      if ((int64_t)t.rows * t.cols > 0) rt.push_back(t);
      maxr = std::max(maxr, (int64_t)t.rows * t.cols);
      o += S.chi * ry.second;
    }
    R->rtotal = o;
    if (hipMalloc((void**)&R->dleft, sizeof(double) * (size_t)std::max<int64_t>(R->ltotal, 1)) != hipSuccess ||
        hipMalloc((void**)&R->dright, sizeof(double) * (size_t)std::max<int64_t>(R->rtotal, 1)) != hipSuccess)
      throw SErr(MPSA_ENOMEM, "split: output allocation failed");
    hchk(launch_fill(R->dleft, std::max<int64_t>(R->ltotal, 1), 0.0, st), "split: zero");
    hchk(launch_fill(R->dright, std::max<int64_t>(R->rtotal, 1), 0.0, st), "split: zero");
    scales.alloc(scal.size());
    DevArr<Copy2D> dl_, dr_;
    dl_.alloc(lt.size());
    dr_.alloc(rt.size());
    if (!scal.empty())
      hchk(hipMemcpyAsync(scales.p, scal.data(), sizeof(double) * scal.size(), hipMemcpyHostToDevice, st), "split");
    if (!lt.empty()) {
      hchk(hipMemcpyAsync(dl_.p, lt.data(), sizeof(Copy2D) * lt.size(), hipMemcpyHostToDevice, st), "split");
      hchk(launch_copy2d(dl_.p, (int)lt.size(), maxl, U.p, R->dleft, scales.p, st), "split: left");
    }
    if (!rt.empty()) {
      hchk(hipMemcpyAsync(dr_.p, rt.data(), sizeof(Copy2D) * rt.size(), hipMemcpyHostToDevice, st), "split");
      hchk(launch_copy2d(dr_.p, (int)rt.size(), maxr, V.p, R->dright, scales.p, st), "split: right");
    }
    hchk(hipStreamSynchronize(st), "split: sync");
    *out = R;
    return MPSA_OK;
  } catch (const SErr& e) {
    delete R;
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    delete R;
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_split_info(const mpsa_split_result* R, double* truncated_weight, int* chi, int64_t* n_left, int64_t* n_right,
                    int64_t* n_lambda, int64_t* left_len, int64_t* right_len) {
  if (!R) {
    set_last_error("split_info: null result");
    return MPSA_EINVAL;
  }
  if (truncated_weight) *truncated_weight = R->tw;
  if (chi) *chi = R->chi;
  if (n_left) *n_left = (int64_t)R->lrows.size();
  if (n_right) *n_right = (int64_t)R->rrows.size();
  if (n_lambda) *n_lambda = (int64_t)R->lam_q.size();
  if (left_len) *left_len = R->ltotal;
  if (right_len) *right_len = R->rtotal;
  return MPSA_OK;
}

int mpsa_split_table(const mpsa_split_result* R, int which, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols,
                     int64_t* offsets) {
  if (!R || !q || !p || !rows || !cols || !offsets) {
    set_last_error("split_table: bad arguments");
    return MPSA_EINVAL;
  }
  const bool l = which == 0;
  const auto& vq = l ? R->lq : R->rq;
  const auto& vp = l ? R->lp : R->rp;
  const auto& vr = l ? R->lrows : R->rrows;
  const auto& vc = l ? R->lcols : R->rcols;
  const auto& vo = l ? R->loff : R->roff;
  std::memcpy(q, vq.data(), sizeof(uint64_t) * vq.size());
  std::memcpy(p, vp.data(), sizeof(uint32_t) * vp.size());
  std::memcpy(rows, vr.data(), sizeof(int64_t) * vr.size());
  std::memcpy(cols, vc.data(), sizeof(int64_t) * vc.size());
  std::memcpy(offsets, vo.data(), sizeof(int64_t) * vo.size());
  return MPSA_OK;
}

const double* mpsa_split_data(const mpsa_split_result* R, int which) {
  if (!R) return nullptr;
  return which == 0 ? R->dleft : R->dright;
}

int mpsa_split_copy(const mpsa_split_result* R, int which, double* dst, mpsa_stream stream) {
  if (!R || !dst) {
    set_last_error("split_copy: bad arguments");
    return MPSA_EINVAL;
  }
  const int64_t n = which == 0 ? R->ltotal : R->rtotal;
  const double* src = which == 0 ? R->dleft : R->dright;
  if (n <= 0) return MPSA_OK;
  hipError_t e = hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) {
    set_last_error(hipGetErrorString(e));
    return MPSA_EHIP;
  }
  return MPSA_OK;
}

int mpsa_split_lambda(const mpsa_split_result* R, uint64_t* q, int64_t* sizes, double* values) {
  if (!R || !q || !sizes || (!values && !R->lam.empty())) {
    set_last_error("split_lambda: bad arguments");
    return MPSA_EINVAL;
  }
  std::memcpy(q, R->lam_q.data(), sizeof(uint64_t) * R->lam_q.size());
  std::memcpy(sizes, R->lam_n.data(), sizeof(int64_t) * R->lam_n.size());
  if (!R->lam.empty()) std::memcpy(values, R->lam.data(), sizeof(double) * R->lam.size());
  return MPSA_OK;
}

int mpsa_split_destroy(mpsa_split_result* R) {
  delete R;
  return MPSA_OK;
}

}  // extern "C"
