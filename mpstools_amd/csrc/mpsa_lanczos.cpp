// mpsa_lanczos.cpp -- device-resident plain Lanczos of TSLanzcosEngine
// (src/lanzcos.hpp:165-278) and getState (:81-92), with the host-side FP64
// symmetric tridiagonal eigensolver replacing LAPACK dstev
// (blas/lapack_routines.cpp:435-446).
//
// The recurrence is the reference's, step for step: k = ||A1||, A1 /= k,
// A2 = H A1, e = <A1|A2>, A3 = (-e A1 + A2) + (-k A0), every _Ndiag = 5 steps a
// convergence test on the lowest Ritz value (testenergy) or on the Ritz-vector
// residual, stop on |k| < delta or it == itmax.  All vectors live on the device
// in the plan's vector layout; the Krylov basis is kept (the reference keeps
// _As) and the requested Ritz state is assembled on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

namespace mpsa {

// Implicit QL iteration with Wilkinson-type shifts for a symmetric tridiagonal
// matrix (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968) 293),
// accumulating the rotations into z (n x n column-major, starts as identity).
// d: diagonal (eigenvalues on exit, ascending); e: n-1 off-diagonals.
int tridiag_ql(int n, double* d, const double* e_in, double* z) {
  if (n <= 0) return 0;
  std::vector<double> e((size_t)n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[(size_t)i] = e_in[i];
  if (z) {
    std::fill(z, z + (size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  }
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[(size_t)m]) <= eps * dd) break;
      }
      if (m != l) {
        if (++iter > 100) return 1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[(size_t)l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[(size_t)l] / (g + std::copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[(size_t)i];
          const double b = c * e[(size_t)i];
          r = std::hypot(f, g);
          e[(size_t)i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[(size_t)m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (z)
            for (int k = 0; k < n; ++k) {
              f = z[(size_t)(i + 1) * n + k];
              z[(size_t)(i + 1) * n + k] = s * z[(size_t)i * n + k] + c * f;
              z[(size_t)i * n + k] = c * z[(size_t)i * n + k] - s * f;
            }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[(size_t)l] = g;
        e[(size_t)m] = 0.0;
      }
    } while (m != l);
  }
  // ascending order (selection sort keeps it simple for the small n of Lanczos)
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      std::swap(d[i], d[k]);
      if (z)
        for (int r = 0; r < n; ++r) std::swap(z[(size_t)i * n + r], z[(size_t)k * n + r]);
    }
  }
  return 0;
}

// Pool of device vectors (Krylov bases are reused across Lanczos calls).
// Idle vectors are kept up to a byte cap; a vector handed back beyond the cap
// frees idle vectors of other lengths first (the ones a sweep over bonds of
// other sizes left behind), then itself.  A failed allocation releases every
// idle vector and tries once more.  All members lock one mutex, so Lanczos
// runs on different host threads share the pool safely.
struct VecPool {
  std::mutex mu;
  std::map<int64_t, std::vector<double*>> free_;
  int64_t idle_bytes = 0;
  int64_t cap_bytes = int64_t(8) << 30;  // 8 GiB of idle vectors
  static int64_t bytes_of(int64_t n) { return (int64_t)sizeof(double) * std::max<int64_t>(n, 1); }
  void release_locked() {
    for (auto& kv : free_)
      for (double* p : kv.second) (void)hipFree(p);
    free_.clear();
    idle_bytes = 0;
  }
  void trim_locked(int64_t keep_len) {  // free idle vectors of other lengths, then of keep_len, down to the cap
    for (auto it = free_.begin(); it != free_.end() && idle_bytes > cap_bytes;) {
      if (it->first == keep_len) {
        ++it;
        continue;
      }
      for (double* p : it->second) {
        (void)hipFree(p);
        idle_bytes -= bytes_of(it->first);
      }
      it = free_.erase(it);
    }
    auto k = free_.find(keep_len);
    while (k != free_.end() && idle_bytes > cap_bytes && !k->second.empty()) {
      (void)hipFree(k->second.back());
      k->second.pop_back();
      idle_bytes -= bytes_of(keep_len);
    }
  }
  double* get(int64_t n) {
    std::lock_guard<std::mutex> g(mu);
    auto& f = free_[n];
    if (!f.empty()) {
      double* p = f.back();
      f.pop_back();
      idle_bytes -= bytes_of(n);
      return p;
    }
    double* p = nullptr;
    if (hipMalloc((void**)&p, (size_t)bytes_of(n)) == hipSuccess) return p;
    (void)hipGetLastError();
    release_locked();
    p = nullptr;
    if (hipMalloc((void**)&p, (size_t)bytes_of(n)) == hipSuccess) return p;
    (void)hipGetLastError();
    return nullptr;
  }
  void put(int64_t n, double* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    free_[n].push_back(p);
    idle_bytes += bytes_of(n);
    if (idle_bytes > cap_bytes) trim_locked(n);
  }
  void release() {
    std::lock_guard<std::mutex> g(mu);
    release_locked();
  }
  void set_cap(int64_t bytes) {
    std::lock_guard<std::mutex> g(mu);
    cap_bytes = std::max<int64_t>(bytes, 0);
    trim_locked(-1);
  }
  int64_t idle() {
    std::lock_guard<std::mutex> g(mu);
    return idle_bytes;
  }
};
static VecPool g_pool;

}  // namespace mpsa

using namespace mpsa;

namespace {
struct LErr : std::runtime_error {
  int code;
  LErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void chk(int rc) {
  if (rc != 0) throw LErr(rc, mpsa_last_error());
}
void hchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw LErr(MPSA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Vectors handed out zero-filled: layouts can have alignment gaps between
// blocks that no kernel writes, and every dot/axpby runs over the full extent.
struct Vecs {
  std::vector<double*> v;
  int64_t n;
  hipStream_t st;
  ~Vecs() {
    // work queued on st may still use the vectors (an exception path)
    if (!v.empty()) (void)hipStreamSynchronize(st);
    for (double* p : v) g_pool.put(n, p);
  }
  double* get() {
    double* p = g_pool.get(n);
    if (!p) throw LErr(MPSA_ENOMEM, "Lanczos vector allocation failed");
    v.push_back(p);
    hchk(launch_fill(p, n, 0.0, st), "zero vector");
    return p;
  }
};

double ddot(const double* x, const double* y, int64_t n, hipStream_t st) {
  double out[2];
  chk(mpsa_dot(x, y, n, 0, out, (mpsa_stream)st));
  return out[0];
}
void axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st) {
  double av[2] = {a, 0.0}, bv[2] = {b, 0.0};
  chk(mpsa_axpby(av, x, bv, y, n, 0, (mpsa_stream)st));
}
}  // namespace

extern "C" {

int mpsa_tridiag_eig(int n, double* d, const double* e, double* z) {
  if (n < 0 || (n > 0 && !d) || (n > 1 && !e)) return MPSA_EINVAL;
  return tridiag_ql(n, d, e, z) == 0 ? MPSA_OK : MPSA_EINVAL;
}

int mpsa_lanczos(mpsa_plan* plan, const double* dpsi0, double* dstate, int itmax, double delta, int testenergy,
                 int target, double* energies, int cap, int* nit, int* nritz, mpsa_stream stream) {
  return mpsa_lanczos_ex(plan, dpsi0, dstate, itmax, delta, testenergy, target, energies, cap, nit, nritz, nullptr, 0,
                         0, 0, stream);
}

int mpsa_lanczos_ex(mpsa_plan* plan, const double* dpsi0, double* dstate, int itmax, double delta, int testenergy,
                    int target, double* energies, int cap, int* nit, int* nritz, const double* dproj, int nproj,
                    int64_t proj_stride, int reortho, mpsa_stream stream) {
  std::vector<void*> small;  // device tables, freed on exit
  auto cleanup = [&]() {
    if (!small.empty()) (void)hipStreamSynchronize((hipStream_t)stream);
    for (void* p : small) (void)hipFree(p);
    small.clear();
  };
  try {
    if (!plan || !dpsi0 || itmax < 1 || !nit || !nritz) throw LErr(MPSA_EINVAL, "lanczos: bad arguments");
    if (nproj < 0 || (nproj > 0 && !dproj) || (reortho != 0 && reortho != 1))
      throw LErr(MPSA_EINVAL, "lanczos: bad projection arguments");
    int64_t xl = 0, yl = 0;
    int same = 0;
    chk(mpsa_plan_vector_lengths(plan, &xl, &yl, &same));
    if (!same) throw LErr(MPSA_EINVAL, "lanczos: the plan's x and y tables must be the same vector space");
    const int64_t n = xl;
    // The energy test runs as one vector of the batched recurrence: norms and
    // alphas stay on the device, five steps per graph replay, one read-back
    // per test -- the same steps, stops and Ritz vector as the loop below,
    // without a host round trip per step.
    if (testenergy)
      return mpsa_lanczos_batched_ex(plan, dpsi0, std::max<int64_t>(n, 1), 1, dstate, itmax, delta, target,
                                     energies, cap, nit, nritz, 1, dproj, nproj, proj_stride, 0, reortho, stream);
    if (nproj > 0 && proj_stride < n) throw LErr(MPSA_EINVAL, "lanczos: projection stride smaller than the vector");
    hipStream_t st = (hipStream_t)stream;
    Vecs V{{}, n, st};
    auto dalloc = [&](size_t bytes) {
      void* p = nullptr;
      if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) throw LErr(MPSA_ENOMEM, "lanczos alloc");
      small.push_back(p);
      return p;
    };
    // projection set (orthonormalised once: s and i s on a complex plan) and
    // the device copy of the basis table for reorthogonalisation
    const bool cplx = plan_complex(plan);
    const int nq = cplx ? 2 * nproj : nproj;
    const int nbmax = std::max(nq, reortho ? itmax + 1 : 0);
    const int nparts = (int)std::max<int64_t>(1, std::min<int64_t>(256, (n + 8191) / 8192));
    double* ppart = nbmax > 0 ? (double*)dalloc(sizeof(double) * (size_t)nbmax * nparts) : nullptr;
    double* pcoef = nbmax > 0 ? (double*)dalloc(sizeof(double) * (size_t)nbmax) : nullptr;
    double** ptab = nq > 0 ? (double**)dalloc(sizeof(double*) * (size_t)nq) : nullptr;
    double** btab = reortho ? (double**)dalloc(sizeof(double*) * (size_t)(itmax + 1)) : nullptr;
    std::vector<double*> hq((size_t)nq), hb;
    hb.reserve((size_t)itmax + 1);  // stable storage: its entries are copied to btab asynchronously
    if (nproj > 0) {
      const int64_t e = n / 2;
      for (int p = 0; p < nproj; ++p) {
        double* Q = hq[(size_t)(cplx ? 2 * p : p)] = V.get();
        hchk(hipMemcpyAsync(Q, dproj + p * proj_stride, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st),
             "copy projection state");
        if (cplx) {
          double* R = hq[(size_t)(2 * p + 1)] = V.get();
          hchk(launch_axpby(-1.0, 0.0, Q + e, 0.0, 0.0, R, e, 0, st), "i s");
          hchk(launch_axpby(1.0, 0.0, Q, 0.0, 0.0, R + e, e, 0, st), "i s");
        }
      }
      hchk(hipMemcpyAsync(ptab, hq.data(), sizeof(double*) * (size_t)nq, hipMemcpyHostToDevice, st), "state table");
      for (int p = 0; p < nq; ++p) {
        double* Q = hq[(size_t)p];
        const double n0 = ddot(Q, Q, n, st);
        for (int pass = 0; pass < 2 && p > 0; ++pass)
          hchk(launch_project_out(Q, ptab, 0, nullptr, p, nbmax, n, n, 1, ppart, nparts, pcoef, st), "state GS");
        const double n1 = ddot(Q, Q, n, st);
        if (!(n0 > 0.0) || !(n1 > 1e-24 * n0))
          throw LErr(MPSA_EINVAL, "lanczos: projection state " + std::to_string(cplx ? p / 2 : p) +
                                      " is zero or linearly dependent on the earlier ones");
        axpby(0.0, Q, 1.0 / std::sqrt(n1), Q, n, st);
      }
    }
    auto project = [&](double* W) {
      if (nq > 0) hchk(launch_project_out(W, ptab, 0, nullptr, nq, nbmax, n, n, 1, ppart, nparts, pcoef, st), "project");
    };
    std::vector<double*> As;
    double* A0 = nullptr;
    double* A1 = V.get();
    double* A2 = V.get();
    hchk(hipMemcpyAsync(A1, dpsi0, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st), "copy psi0");
    project(A1);
    std::vector<double> es, ks;
    int it = 0;
    double k = 0.0, e = 0.0, eold = 1e100;
    bool first = true;
    const int Ndiag = 5;
    while (true) {
      k = std::sqrt(ddot(A1, A1, n, st));
      if (std::fabs(k) < delta) break;
      if (!first) ks.push_back(k);
      axpby(0.0, A1, 1.0 / k, A1, n, st);                  // A1 /= k
      chk(mpsa_heff_apply(plan, A1, 0, A2, 0, 1, stream));  // A2 = H A1
      As.push_back(A1);                                      // the basis keeps A1
      e = ddot(A2, A1, n, st);
      es.push_back(e);
      if (it % Ndiag == 0 && it != 0) {
        std::vector<double> d(es), sub(ks);
        if (testenergy) {
          if (tridiag_ql((int)d.size(), d.data(), sub.data(), nullptr)) throw LErr(MPSA_EINVAL, "tridiag failed");
          if (std::fabs(d[0] - eold) < delta) break;
          eold = d[0];
        } else {
          std::vector<double> z(d.size() * d.size());
          if (tridiag_ql((int)d.size(), d.data(), sub.data(), z.data())) throw LErr(MPSA_EINVAL, "tridiag failed");
          double* A4 = V.get();
          double* A5 = V.get();
          axpby(z[0], As[0], 0.0, A4, n, st);
          for (size_t i = 1; i < d.size(); ++i) axpby(z[i], As[i], 1.0, A4, n, st);
          chk(mpsa_heff_apply(plan, A4, 0, A5, 0, 1, stream));
          axpby(-d[0], A4, 1.0, A5, n, st);
          const double resi = std::sqrt(ddot(A5, A5, n, st));
          if (resi < delta) break;
        }
      }
      // A3 = (-e A1 + A2) + (-k A0), built in A2's storage
      axpby(-e, A1, 1.0, A2, n, st);
      if (!first) axpby(-k, A0, 1.0, A2, n, st);
      if (reortho) {  // against the basis A_0 .. A_it, twice
        hb.push_back(A1);
        hchk(hipMemcpyAsync(btab + (hb.size() - 1), &hb.back(), sizeof(double*), hipMemcpyHostToDevice, st),
             "basis table");
        for (int pass = 0; pass < 2; ++pass)
          hchk(launch_project_out(A2, btab, 0, nullptr, (int)hb.size(), nbmax, n, n, 1, ppart, nparts, pcoef, st),
               "reorthogonalise");
      }
      project(A2);
      first = false;
      A0 = A1;
      A1 = A2;
      A2 = V.get();
      ++it;
      if (it >= itmax) break;
    }
    *nit = it;
    *nritz = (int)es.size();
    if (!es.empty()) {
      const int m = (int)es.size();
      std::vector<double> d(es), sub(ks), z((size_t)m * m);
      if (tridiag_ql(m, d.data(), sub.data(), z.data())) throw LErr(MPSA_EINVAL, "tridiag failed");
      for (int i = 0; i < m && i < cap; ++i) energies[i] = d[(size_t)i];
      if (dstate) {
        if (target < 0 || target >= m) throw LErr(MPSA_EINVAL, "LanzcosEngine::getState(const int j): j<_eigs.size2() failed");
        hchk(launch_fill(dstate, n, 0.0, st), "zero state");
        axpby(z[(size_t)target * m + 0], As[0], 0.0, dstate, n, st);
        for (int i = 1; i < m; ++i) axpby(z[(size_t)target * m + i], As[(size_t)i], 1.0, dstate, n, st);
      }
    }
    hchk(hipStreamSynchronize(st), "lanczos sync");
    cleanup();
    return MPSA_OK;
  } catch (const LErr& e) {
    cleanup();
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    cleanup();
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_lanczos_batched(mpsa_plan* plan, const double* dpsi0, int64_t stride, int nvec, double* dstates, int itmax,
                         double delta, int target, double* energies, int cap, int* nit, int* nritz, int use_graph,
                         mpsa_stream stream) {
  return mpsa_lanczos_batched_ex(plan, dpsi0, stride, nvec, dstates, itmax, delta, target, energies, cap, nit, nritz,
                                 use_graph, nullptr, 0, 0, 0, 0, stream);
}

int mpsa_lanczos_batched_ex(mpsa_plan* plan, const double* dpsi0, int64_t stride, int nvec, double* dstates, int itmax,
                            double delta, int target, double* energies, int cap, int* nit, int* nritz, int use_graph,
                            const double* dproj, int nproj, int64_t proj_stride, int64_t proj_vstride, int reortho,
                            mpsa_stream stream) {
  hipStream_t cs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<void*> small;  // device scalars and tables, freed on exit
  auto cleanup = [&]() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cs) (void)hipStreamSynchronize(cs);
    for (void* p : small) (void)hipFree(p);
    if (cs) (void)hipStreamDestroy(cs);
  };
  try {
    if (!plan || !dpsi0 || nvec < 1 || itmax < 1 || !nit || !nritz || (cap > 0 && !energies))
      throw LErr(MPSA_EINVAL, "lanczos_batched: bad arguments");
    if (nproj < 0 || (nproj > 0 && !dproj) || proj_vstride < 0 || (reortho != 0 && reortho != 1))
      throw LErr(MPSA_EINVAL, "lanczos_batched: bad projection arguments");
    int64_t xl = 0, yl = 0;
    int same = 0;
    chk(mpsa_plan_vector_lengths(plan, &xl, &yl, &same));
    if (!same) throw LErr(MPSA_EINVAL, "lanczos_batched: the plan's x and y tables must be the same vector space");
    const int64_t n = xl, vs = stride;
    if (vs < n) throw LErr(MPSA_EINVAL, "lanczos_batched: stride smaller than the vector length");
    if (nproj > 0 && (proj_stride < n || (proj_vstride > 0 && proj_vstride < n)))
      throw LErr(MPSA_EINVAL, "lanczos_batched: projection strides smaller than the vector length");
    // Capture needs a stream other than the legacy default stream: all work
    // runs on a private stream after the caller's stream has drained.
    hchk(hipStreamSynchronize((hipStream_t)stream), "lanczos_batched: caller stream");
    hchk(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "lanczos_batched: stream");
    if (use_graph) chk(mpsa_plan_set_timing(plan, 0));
    const int ld = itmax + 1;
    const int nparts = (int)std::max<int64_t>(1, std::min<int64_t>(256, (n + 8191) / 8192));
    // the recurrence's own dot products: up to 1024 partial sums (4
    // workgroups per CU on the large vectors)
    const int dparts = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 8191) / 8192));
    auto dalloc = [&](size_t bytes) {
      void* p = nullptr;
      if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) throw LErr(MPSA_ENOMEM, "lanczos_batched alloc");
      small.push_back(p);
      return p;
    };
    int* dstep = (int*)dalloc(sizeof(int));
    double* kbuf = (double*)dalloc(sizeof(double) * (size_t)nvec * ld);
    double* ebuf = (double*)dalloc(sizeof(double) * (size_t)nvec * ld);
    double* partial = (double*)dalloc(sizeof(double) * (size_t)nvec * std::max(nparts, dparts));
    double** table = (double**)dalloc(sizeof(double*) * (size_t)ld);
    const int zero = 0;
    hchk(hipMemcpy(dstep, &zero, sizeof(int), hipMemcpyHostToDevice), "step counter");
    const int64_t tot = vs * nvec;
    Vecs V{{}, tot, cs};
    double* W = V.get();
    double* X = V.get();
    double* Y = V.get();
    hchk(hipMemcpyAsync(W, dpsi0, sizeof(double) * (size_t)tot, hipMemcpyDeviceToDevice, cs), "copy psi0");
    std::vector<double*> basis;
    basis.reserve((size_t)ld);
    auto ensure_basis = [&](int upto) {
      const size_t old = basis.size();
      while ((int)basis.size() < upto) basis.push_back(V.get());
      if (basis.size() > old)
        hchk(hipMemcpy(table + old, basis.data() + old, sizeof(double*) * (basis.size() - old), hipMemcpyHostToDevice),
             "basis table");
    };
    // Projection states (src/lanzcos.hpp:281-414): copied once, then
    // orthonormalised in FP64 (classical Gram-Schmidt, twice); for one state
    // W - q <q, W> with q = s / |s| is the reference's W - s (s*W) / (s*s).
    // With proj_vstride = 0 one set serves every vector (qvs = 0).
    // On a complex plan the recurrence runs on planar vectors with the real
    // dot product; the complex span of a state s is the real span of s and
    // i s, so both are projected out (nq = 2 nproj directions).
    const int nvp = proj_vstride > 0 ? nvec : 1;
    const int64_t qvs = proj_vstride > 0 ? vs : 0;
    const bool cplx = plan_complex(plan);
    const int nq = cplx ? 2 * nproj : nproj;
    const int nbmax = std::max(nq, reortho ? ld : 0);
    double* ppart = nbmax > 0 ? (double*)dalloc(sizeof(double) * (size_t)nvec * nbmax * nparts) : nullptr;
    double* pcoef = nbmax > 0 ? (double*)dalloc(sizeof(double) * (size_t)nvec * nbmax) : nullptr;
    double** ptab = nullptr;
    Vecs P{{}, vs * nvp, cs};
    if (nproj > 0) {
      ptab = (double**)dalloc(sizeof(double*) * (size_t)nq);
      std::vector<double*> hp((size_t)nq);
      double* nrm = (double*)dalloc(sizeof(double) * 2 * (size_t)nvp);
      std::vector<double> hn(2 * (size_t)nvp);
      const int64_t e = n / 2;  // complex extent
      for (int p = 0; p < nproj; ++p) {
        double* Q = hp[(size_t)(cplx ? 2 * p : p)] = P.get();
        for (int v = 0; v < nvp; ++v)
          hchk(hipMemcpyAsync(Q + (int64_t)v * vs, dproj + p * proj_stride + v * proj_vstride,
                              sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, cs),
               "copy projection state");
        if (cplx) {  // i s: (re, im) -> (-im, re)
          double* R = hp[(size_t)(2 * p + 1)] = P.get();
          for (int v = 0; v < nvp; ++v) {
            hchk(launch_axpby(-1.0, 0.0, Q + (int64_t)v * vs + e, 0.0, 0.0, R + (int64_t)v * vs, e, 0, cs), "i s");
            hchk(launch_axpby(1.0, 0.0, Q + (int64_t)v * vs, 0.0, 0.0, R + (int64_t)v * vs + e, e, 0, cs), "i s");
          }
        }
      }
      hchk(hipMemcpyAsync(ptab, hp.data(), sizeof(double*) * (size_t)nq, hipMemcpyHostToDevice, cs), "state table");
      for (int p = 0; p < nq; ++p) {
        double* Q = hp[(size_t)p];
        hchk(launch_bdot(Q, Q, n, vs, nvp, partial, nparts, nrm, 1, dstep, 0, cs), "state norm");
        for (int pass = 0; pass < 2 && p > 0; ++pass)
          hchk(launch_project_out(Q, ptab, vs, nullptr, p, nq, n, vs, nvp, ppart, nparts, pcoef, cs), "state GS");
        hchk(launch_bdot(Q, Q, n, vs, nvp, partial, nparts, nrm + nvp, 1, dstep, 0, cs), "state norm");
        hchk(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * hn.size(), hipMemcpyDeviceToHost, cs), "read norms");
        hchk(hipStreamSynchronize(cs), "state norms");
        for (int v = 0; v < nvp; ++v)
          if (!(hn[(size_t)v] > 0.0) || !(hn[(size_t)nvp + v] > 1e-24 * hn[(size_t)v]))
            throw LErr(MPSA_EINVAL, "lanczos_batched: projection state " + std::to_string(cplx ? p / 2 : p) +
                                        " is zero or linearly dependent on the earlier ones");
        hchk(launch_scale_rsqrt(Q, nrm + nvp, n, vs, nvp, cs), "state normalise");
      }
    }
    // Without projections or reorthogonalisation nothing changes W between
    // the recurrence update and the next step's norm, so the update sums
    // |W|^2 as it writes W and the norm pass only adds the partial sums.
    const bool fuse_norm = nproj == 0 && !reortho;
    const int wparts = lz_update_parts(n);
    double* wpart = fuse_norm ? (double*)dalloc(sizeof(double) * (size_t)nvec * wparts) : nullptr;
    bool norm_ready = false;
    auto step_once = [&]() {
      if (nproj > 0)
        hchk(launch_project_out(W, ptab, qvs, nullptr, nq, nbmax, n, vs, nvec, ppart, nparts, pcoef, cs), "project");
      if (norm_ready)  // |W|^2 summed by the previous step's recurrence update
        hchk(launch_bdot_final(wpart, wparts, nvec, kbuf, ld, dstep, 1, cs), "norm");
      else
        hchk(launch_bdot(W, W, n, vs, nvec, partial, dparts, kbuf, ld, dstep, 1, cs), "norm");
      hchk(launch_lz_normalize(W, X, table, kbuf, ld, dstep, n, vs, nvec, cs), "normalise");
      chk(mpsa_heff_apply(plan, X, vs, Y, vs, nvec, (mpsa_stream)cs));
      hchk(launch_bdot(Y, X, n, vs, nvec, partial, dparts, ebuf, ld, dstep, 0, cs), "alpha");
      hchk(launch_lz_update(Y, X, table, kbuf, ebuf, ld, dstep, W, n, vs, nvec, fuse_norm ? wpart : nullptr, cs),
           "recurrence");
      norm_ready = fuse_norm;
      // full reorthogonalisation against the basis 0..step, twice
      for (int pass = 0; reortho && pass < 2; ++pass)
        hchk(launch_project_out(W, table, vs, dstep, 1, nbmax, n, vs, nvec, ppart, nparts, pcoef, cs), "reorthogonalise");
      hchk(launch_lz_step(dstep, cs), "step");
    };
    const int Ndiag = 5;
    // per-run state of the reference's control flow (src/lanzcos.hpp:184-257)
    struct Run {
      std::vector<double> es, ks;
      double eold = 1e100;
      int it = 0;
      bool first = true, done = false;
    };
    std::vector<Run> runs((size_t)nvec);
    std::vector<double> hk((size_t)nvec * ld), he((size_t)nvec * ld);
    int launched = 0, consumed = 0;
    while (true) {
      // the first block ends at it = 5 (6 steps), later blocks are 5 steps
      int cnt = launched == 0 ? 1 : Ndiag;
      cnt = std::min(cnt, itmax - launched);
      if (cnt <= 0) break;
      ensure_basis(launched + cnt);
      if (use_graph && cnt == Ndiag) {
        if (!gexec) {
          hchk(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
          try {
            for (int i = 0; i < Ndiag; ++i) step_once();
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cs, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          hchk(hipStreamEndCapture(cs, &graph), "end capture");
          hchk(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "instantiate");
        }
        hchk(hipGraphLaunch(gexec, cs), "graph launch");
      } else {
        for (int i = 0; i < cnt; ++i) step_once();
      }
      launched += cnt;
      if (launched < 6 && launched < itmax) continue;  // complete the first test block before reading back
      hchk(hipMemcpyAsync(hk.data(), kbuf, sizeof(double) * hk.size(), hipMemcpyDeviceToHost, cs), "read beta");
      hchk(hipMemcpyAsync(he.data(), ebuf, sizeof(double) * he.size(), hipMemcpyDeviceToHost, cs), "read alpha");
      hchk(hipStreamSynchronize(cs), "lanczos_batched sync");
      bool all_done = true;
      for (int v = 0; v < nvec; ++v) {
        Run& r = runs[(size_t)v];
        for (int j = consumed; j < launched && !r.done; ++j) {
          const double k = hk[(size_t)v * ld + j];
          if (!(std::fabs(k) >= delta)) {  // |k| < delta (a NaN stops the run too)
            r.done = true;
            break;
          }
          if (!r.first) r.ks.push_back(k);
          const double e = he[(size_t)v * ld + j];
          r.es.push_back(e);
          if (r.it % Ndiag == 0 && r.it != 0) {
            std::vector<double> d(r.es), sub(r.ks);
            if (tridiag_ql((int)d.size(), d.data(), sub.data(), nullptr)) throw LErr(MPSA_EINVAL, "tridiag failed");
            if (std::fabs(d[0] - r.eold) < delta) {
              r.done = true;
              break;
            }
            r.eold = d[0];
          }
          r.first = false;
          ++r.it;
          if (r.it >= itmax) r.done = true;
        }
        all_done = all_done && r.done;
      }
      consumed = launched;
      if (all_done) break;
    }
    // Ritz values and states (tridiagonal eigensolve per run, src/lanzcos.hpp:258-268)
    int mmax = 0;
    for (auto& r : runs) mmax = std::max(mmax, (int)r.es.size());
    std::vector<double> coef((size_t)nvec * std::max(mmax, 1), 0.0);
    std::vector<int> ms((size_t)nvec, 0);
    for (int v = 0; v < nvec; ++v) {
      Run& r = runs[(size_t)v];
      nit[v] = r.it;
      const int m = (int)r.es.size();
      nritz[v] = m;
      ms[(size_t)v] = m;
      if (m == 0) continue;
      std::vector<double> d(r.es), sub(r.ks), z((size_t)m * m);
      if (tridiag_ql(m, d.data(), sub.data(), z.data())) throw LErr(MPSA_EINVAL, "tridiag failed");
      for (int i = 0; i < m && i < cap; ++i) energies[(size_t)v * cap + i] = d[(size_t)i];
      if (dstates) {
        if (target < 0 || target >= m)
          throw LErr(MPSA_EINVAL, "LanzcosEngine::getState(const int j): j<_eigs.size2() failed");
        for (int i = 0; i < m; ++i) coef[(size_t)v * std::max(mmax, 1) + i] = z[(size_t)target * m + i];
      }
    }
    if (dstates && mmax > 0) {
      double* dcoef = (double*)dalloc(sizeof(double) * coef.size());
      int* dm = (int*)dalloc(sizeof(int) * ms.size());
      hchk(hipMemcpy(dcoef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice), "coefficients");
      hchk(hipMemcpy(dm, ms.data(), sizeof(int) * ms.size(), hipMemcpyHostToDevice), "lengths");
      hchk(launch_fill(dstates, tot, 0.0, cs), "zero states");
      hchk(launch_lz_combine(table, dcoef, std::max(mmax, 1), dm, dstates, n, vs, nvec, cs), "states");
    }
    hchk(hipStreamSynchronize(cs), "lanczos_batched sync");
    cleanup();
    return MPSA_OK;
  } catch (const LErr& e) {
    cleanup();
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    cleanup();
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_release_pools(void) {
  g_pool.release();
  return MPSA_OK;
}

int mpsa_pool_info(int64_t* idle_bytes, int64_t* cap_bytes) {
  if (idle_bytes) *idle_bytes = g_pool.idle();
  if (cap_bytes) {
    std::lock_guard<std::mutex> g(g_pool.mu);
    *cap_bytes = g_pool.cap_bytes;
  }
  return MPSA_OK;
}

int mpsa_pool_set_cap(int64_t cap_bytes) {
  if (cap_bytes < 0) {
    set_last_error("pool_set_cap: negative cap");
    return MPSA_EINVAL;
  }
  g_pool.set_cap(cap_bytes);
  return MPSA_OK;
}

// Chebyshev recurrence of ChebychevEngine::simulateSSCompress
// (src/container.hpp:2343-2470) on one bond, with the exact H_eff of the plan
// in place of the compressed MPO application: t_1 = Ht t_0,
// t_{k+1} = 2 Ht t_k - t_{k-1} (:2423-2428), Ht = (H_eff - shift) / scale;
// moments mu_k = <t_0|t_k> (:2368, 2388, 2431), and with recons the
// reconstruction of the doubled moments (:2390-2394, 2433-2436):
// 2<t_k|t_k> - mu_0 = mu_2k and 2<t_k|t_{k-1}> - mu_1 = mu_{2k-1}.
int mpsa_chebyshev_moments(mpsa_plan* plan, const double* dt0, int nmom, double scale, double shift,
                           double* moments, double* recons, mpsa_stream stream) {
  try {
    if (!plan || !dt0 || nmom < 1 || !moments || !(std::fabs(scale) > 0.0) || !std::isfinite(scale) ||
        !std::isfinite(shift))
      throw LErr(MPSA_EINVAL, "chebyshev: bad arguments");
    int64_t xl = 0, yl = 0;
    int same = 0;
    chk(mpsa_plan_vector_lengths(plan, &xl, &yl, &same));
    if (!same) throw LErr(MPSA_EINVAL, "chebyshev: the plan's x and y tables must be the same vector space");
    const int64_t n = xl;
    hipStream_t st = (hipStream_t)stream;
    Vecs V{{}, n, st};
    double* t0 = V.get();
    double* a = V.get();  // t_{k-2}
    double* b = V.get();  // t_{k-1}
    double* y = V.get();  // t_k
    hchk(hipMemcpyAsync(t0, dt0, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st), "copy t0");
    hchk(hipMemcpyAsync(a, dt0, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st), "copy t0");
    const double mu0 = ddot(t0, t0, n, st);
    moments[0] = mu0;
    if (recons) recons[0] = mu0;
    if (nmom == 1) return MPSA_OK;
    chk(mpsa_heff_apply(plan, a, n, b, n, 1, stream));
    axpby(-shift / scale, a, 1.0 / scale, b, n, st);  // t_1 = (H t_0 - shift t_0) / scale
    const double mu1 = ddot(t0, b, n, st);
    moments[1] = mu1;
    if (recons) {
      recons[1] = 2.0 * ddot(b, a, n, st) - mu1;
      recons[2] = 2.0 * ddot(b, b, n, st) - mu0;
    }
    for (int k = 2; k < nmom; ++k) {
      chk(mpsa_heff_apply(plan, b, n, y, n, 1, stream));
      axpby(-2.0 * shift / scale, b, 2.0 / scale, y, n, st);  // 2 Ht t_{k-1}
      axpby(-1.0, a, 1.0, y, n, st);                          // - t_{k-2}
      moments[k] = ddot(t0, y, n, st);
      if (recons) {
        recons[2 * k - 1] = 2.0 * ddot(y, b, n, st) - mu1;
        recons[2 * k] = 2.0 * ddot(y, y, n, st) - mu0;
      }
      double* r = a;
      a = b;
      b = y;
      y = r;
    }
    return MPSA_OK;
  } catch (const LErr& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

}  // extern "C"
