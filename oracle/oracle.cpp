// oracle.cpp -- C ABI over the CPU restatement in oracle.hpp, for ctypes.
//
// TEST INFRASTRUCTURE ONLY (the parity checker and the timed CPU baseline).
// Block matrices cross this ABI as flat tables: per block two packed charges
// (uint64, AbKeyType packing), Ns physical indices (uint32), rows, cols; the
// data of all blocks is concatenated column-major in table order.  Complex
// data is interleaved (re, im).
#include "oracle.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <memory>

namespace oracle {

static Blas* g_blas = nullptr;

static void* try_open(const char* name) { return dlopen(name, RTLD_NOW | RTLD_GLOBAL); }

Blas& blas() {
  if (g_blas) return *g_blas;
  Blas* b = new Blas();
  const char* env = getenv("ORACLE_BLAS");
  void* h = nullptr;
  if (!(env && std::string(env) == "loops")) {
    const char* cands[] = {env, "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1",
                           "libopenblas.so.0", "libblas.so.3"};
    for (const char* c : cands) {
      if (!c) continue;
      h = try_open(c);
      if (h) {
        b->name = c;
        break;
      }
    }
  }
  if (h) {
    // MKL: use the GNU OpenMP threading layer so its threads and the
    // OpenMP-over-charges loops share one runtime (libgomp).
    typedef int (*set_layer_fn)(int);
    set_layer_fn sl = (set_layer_fn)dlsym(h, "MKL_Set_Threading_Layer");
    const int layer = sl ? sl(3 /* MKL_THREADING_GNU */) : 3;
    // A process that initialised MKL before us keeps its layer.  The Intel
    // layer (libiomp5) beside the GNU runtime of the charge loops gives racy
    // results; MKL then runs on one thread and the charge loops keep theirs.
    if (layer != 3 && layer != 1 /* MKL_THREADING_SEQUENTIAL */) {
      typedef void (*set_threads_fn)(int);
      set_threads_fn st = (set_threads_fn)dlsym(h, "MKL_Set_Num_Threads");
      if (st) st(1);
      b->name = std::string(b->name) + " (one thread: MKL threading layer already chosen)";
    }
    b->set_threads = (set_threads_fn)dlsym(h, "MKL_Set_Num_Threads");
    if (!b->set_threads) b->set_threads = (set_threads_fn)dlsym(h, "openblas_set_num_threads");
    b->dgemm = (dgemm_fn)dlsym(h, "dgemm_");
    b->zgemm = (zgemm_fn)dlsym(h, "zgemm_");
    b->dstev = (dstev_fn)dlsym(h, "dstev_");
    if (!b->dgemm) b->name = "loops";
  }
  g_blas = b;
  return *g_blas;
}

// Cyclic Jacobi on the dense symmetric tridiagonal matrix (fallback when no
// LAPACK is loaded).  Eigenvalues ascending, eigenvectors sign-normalised like
// nothing in particular (callers only use invariant quantities).
static void jacobi_eig(int n, std::vector<double>& A, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[(size_t)p + (size_t)q * n] * A[(size_t)p + (size_t)q * n];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[(size_t)p + (size_t)q * n];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p + (size_t)p * n], aqq = A[(size_t)q + (size_t)q * n];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[(size_t)k + (size_t)p * n], akq = A[(size_t)k + (size_t)q * n];
          A[(size_t)k + (size_t)p * n] = c * akp - s * akq;
          A[(size_t)k + (size_t)q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[(size_t)p + (size_t)k * n], aqk = A[(size_t)q + (size_t)k * n];
          A[(size_t)p + (size_t)k * n] = c * apk - s * aqk;
          A[(size_t)q + (size_t)k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[(size_t)k + (size_t)p * n], vkq = V[(size_t)k + (size_t)q * n];
          V[(size_t)k + (size_t)p * n] = c * vkp - s * vkq;
          V[(size_t)k + (size_t)q * n] = s * vkp + c * vkq;
        }
      }
  }
}

int tridiag(std::vector<double>& d, std::vector<double> e, std::vector<double>* z) {
  int n = (int)d.size();
  if (n == 0) return 0;
  e.resize((size_t)std::max(1, n - 1));
  if (blas().dstev) {
    char jobz = z ? 'V' : 'N';
    std::vector<double> work((size_t)std::max(1, 2 * n - 2));
    std::vector<double> zz(z ? (size_t)n * n : 1);
    int ldz = n, info = 0;
    blas().dstev(&jobz, &n, d.data(), e.data(), zz.data(), &ldz, work.data(), &info);
    if (z) *z = zz;
    return info;
  }
  std::vector<double> A((size_t)n * n, 0.0), V;
  for (int i = 0; i < n; ++i) A[(size_t)i * n + i] = d[(size_t)i];
  for (int i = 0; i + 1 < n; ++i) A[(size_t)i + (size_t)(i + 1) * n] = A[(size_t)(i + 1) + (size_t)i * n] = e[(size_t)i];
  jacobi_eig(n, A, V);
  std::vector<int> idx((size_t)n);
  for (int i = 0; i < n; ++i) idx[(size_t)i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  std::vector<double> dd((size_t)n);
  for (int i = 0; i < n; ++i) dd[(size_t)i] = A[(size_t)idx[(size_t)i] * n + idx[(size_t)i]];
  d = dd;
  if (z) {
    z->assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) (*z)[(size_t)i + (size_t)j * n] = V[(size_t)i + (size_t)idx[(size_t)j] * n];
  }
  return 0;
}

}  // namespace oracle

using namespace oracle;

// ------------------------------------------------------------------ handles
struct orc_bm {
  int ns = 1;
  int is_c = 0;
  BlockMatrix<1, double> d1;
  BlockMatrix<2, double> d2;
  BlockMatrix<1, cplx> z1;
  BlockMatrix<2, cplx> z2;
};
struct orc_mpo {
  int is_c = 0;
  MPOMat<double> d;
  MPOMat<cplx> z;
};

template <int Ns, typename T>
static void load_bm(BlockMatrix<Ns, T>& bm, int64_t n, const uint64_t* q, const uint32_t* p, const int64_t* rows,
                    const int64_t* cols, const double* data) {
  bm.clear();
  const T* src = reinterpret_cast<const T*>(data);
  size_t off = 0;
  for (int64_t i = 0; i < n; ++i) {
    TKey<Ns> k = mk<Ns>(q[2 * i], q[2 * i + 1], p + Ns * i);
    Mat<T> m(rows[i], cols[i]);
    std::memcpy(m.a.data(), src + off, sizeof(T) * m.a.size());
    off += m.a.size();
    bm[k] = m;
  }
}
template <int Ns, typename T>
static void dump_bm(const BlockMatrix<Ns, T>& bm, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols,
                    double* data) {
  T* dst = reinterpret_cast<T*>(data);
  size_t off = 0;
  int64_t i = 0;
  for (auto& kv : bm) {
    if (q) {
      q[2 * i] = kv.first.q[0];
      q[2 * i + 1] = kv.first.q[1];
    }
    if (p)
      for (int s = 0; s < Ns; ++s) p[Ns * i + s] = kv.first.p[s];
    if (rows) rows[i] = kv.second.r;
    if (cols) cols[i] = kv.second.c;
    if (dst) std::memcpy(dst + off, kv.second.a.data(), sizeof(T) * kv.second.a.size());
    off += kv.second.a.size();
    ++i;
  }
}

static thread_local std::string g_err;

#define ORC_TRY(...)                  \
  try {                               \
    __VA_ARGS__;                      \
    return 0;                         \
  } catch (const std::exception& e) { \
    g_err = e.what();                 \
    return 1;                         \
  }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
const char* orc_blas_name() { return blas().name.c_str(); }
// Seeded Gaussian entries as the library's generator defines them (SplitMix64
// of seed + (j + 1) * golden gamma, top 53 bits + 1 over 2^53 -> u_j in (0, 1];
// x_i = sqrt(-2 ln u_{2i}) cos(2 pi u_{2i+1})), with ln from the atanh series
// ln m = 2 s sum_k s^2k / (2k + 1), s = (m - 1) / (m + 1), m in [sqrt(1/2),
// sqrt(2)), and cos(2 pi u) from the Taylor series of cos / sin on
// [-pi/4, pi/4] after folding u into an octant; every multiply-add an fma.
// Restated here for the host so the device bits can be checked.
static double orc_unit(uint64_t seed, uint64_t j) {
  uint64_t z = seed + (j + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 1.0) * 0x1.0p-53;
}
static double orc_ln(double u) {
  int e = 0;
  double m = std::frexp(u, &e);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = 2.0 * m;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double acc = 1.0 / 23.0;
  for (int k = 10; k >= 0; --k) acc = std::fma(acc, z, 1.0 / (double)(2 * k + 1));
  return std::fma((double)e, 0x1.62e42fefa39efp-1, 2.0 * s * acc);
}
static double orc_cos2pi(double u) {
  const double x = u > 0.5 ? u - 1.0 : u;
  const double a4 = 4.0 * std::fabs(x);
  const double q = a4 < 0.5 ? 0.0 : (a4 < 1.5 ? 1.0 : 2.0);
  const double t = (a4 - q) * 0x1.921fb54442d18p+0, t2 = t * t;
  double fact[23];
  fact[0] = 1.0;
  for (int j = 1; j < 23; ++j) fact[j] = fact[j - 1] * (double)j;
  if (q == 1.0) {  // cos(pi/2 + t) = -sin t
    double sn = -1.0 / fact[21];
    for (int k = 9; k >= 0; --k) sn = std::fma(sn, t2, ((k & 1) ? -1.0 : 1.0) / fact[2 * k + 1]);
    return -(t * sn);
  }
  double c = -1.0 / fact[22];
  for (int k = 10; k >= 0; --k) c = std::fma(c, t2, ((k & 1) ? -1.0 : 1.0) / fact[2 * k]);
  return q == 0.0 ? c : -c;  // cos(pi + t) = -cos t
}
void orc_fill_gauss(double* p, int64_t n, uint64_t seed) {
  for (int64_t i = 0; i < n; ++i)
    p[i] = std::sqrt(-2.0 * orc_ln(orc_unit(seed, 2 * (uint64_t)i))) * orc_cos2pi(orc_unit(seed, 2 * (uint64_t)i + 1));
}
// Threads of the BLAS library (the CPU baseline's 1-thread and all-core runs);
// the charge loops take their thread count per call.
int orc_set_blas_threads(int n) {
  Blas& b = blas();
  if (!b.set_threads) return 1;
  b.set_threads(n < 1 ? 1 : n);
  return 0;
}
// C (m x n) = A (m x k) * B (k x n), column-major, through the loaded BLAS:
// the product the reference issues per tuple (blas/blasroutines.cpp:837-848).
int orc_dgemm(int m, int n, int k, const double* A, const double* B, double* C) {
  ORC_TRY({
    gemm_raw('n', 'n', m, n, k, 1.0, A, m, B, k, C, m);
  })
}

orc_bm* orc_bm_new(int ns, int is_complex) {
  orc_bm* b = new orc_bm();
  b->ns = ns;
  b->is_c = is_complex;
  return b;
}
void orc_bm_free(orc_bm* b) { delete b; }

int orc_bm_load(orc_bm* b, int64_t n, const uint64_t* q, const uint32_t* p, const int64_t* rows,
                const int64_t* cols, const double* data) {
  ORC_TRY({
    if (b->ns == 1 && !b->is_c) load_bm(b->d1, n, q, p, rows, cols, data);
    else if (b->ns == 2 && !b->is_c) load_bm(b->d2, n, q, p, rows, cols, data);
    else if (b->ns == 1) load_bm(b->z1, n, q, p, rows, cols, data);
    else load_bm(b->z2, n, q, p, rows, cols, data);
  })
}
int64_t orc_bm_nblocks(const orc_bm* b) {
  if (b->ns == 1) return b->is_c ? (int64_t)b->z1.size() : (int64_t)b->d1.size();
  return b->is_c ? (int64_t)b->z2.size() : (int64_t)b->d2.size();
}
int64_t orc_bm_nelem(const orc_bm* b) {
  int64_t n = 0;
  if (b->ns == 1 && !b->is_c) for (auto& kv : b->d1) n += (int64_t)kv.second.a.size();
  if (b->ns == 2 && !b->is_c) for (auto& kv : b->d2) n += (int64_t)kv.second.a.size();
  if (b->ns == 1 && b->is_c) for (auto& kv : b->z1) n += (int64_t)kv.second.a.size();
  if (b->ns == 2 && b->is_c) for (auto& kv : b->z2) n += (int64_t)kv.second.a.size();
  return n;
}
int orc_bm_dump(const orc_bm* b, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols, double* data) {
  ORC_TRY({
    if (b->ns == 1 && !b->is_c) dump_bm(b->d1, q, p, rows, cols, data);
    else if (b->ns == 2 && !b->is_c) dump_bm(b->d2, q, p, rows, cols, data);
    else if (b->ns == 1) dump_bm(b->z1, q, p, rows, cols, data);
    else dump_bm(b->z2, q, p, rows, cols, data);
  })
}

// MPO tensor: n entries (wl, wr, out, in, value); value is re or (re,im).
orc_mpo* orc_mpo_new(int is_complex, uint32_t Dl, uint32_t Dr, int64_t n, const uint32_t* wl, const uint32_t* wr,
                     const uint32_t* out, const uint32_t* in, const double* val) {
  orc_mpo* m = new orc_mpo();
  m->is_c = is_complex;
  m->d.Dl = m->z.Dl = Dl;
  m->d.Dr = m->z.Dr = Dr;
  for (int64_t i = 0; i < n; ++i) {
    UKey u{wl[i], wr[i]};
    OpKey o{in[i], out[i]};
    if (is_complex) m->z.m[u][o] = cplx(val[2 * i], val[2 * i + 1]);
    else m->d.m[u][o] = val[i];
  }
  return m;
}
void orc_mpo_free(orc_mpo* m) { delete m; }

static ItoKey mk_itk(const uint64_t* q, int d) { return ItoKey(q, q + d); }

int orc_contract_twosite(orc_bm* out, const orc_bm* L, const orc_bm* R, const orc_bm* psi, const orc_mpo* wl,
                         const orc_mpo* wr, const uint64_t* itl, int dl, const uint64_t* itr, int dr,
                         int nthreads) {
  ORC_TRY({
    if (psi->is_c) contractTwosite(out->z2, L->z1, R->z1, psi->z2, wl->z, wr->z, mk_itk(itl, dl), mk_itk(itr, dr), nthreads);
    else contractTwosite(out->d2, L->d1, R->d1, psi->d2, wl->d, wr->d, mk_itk(itl, dl), mk_itk(itr, dr), nthreads);
  })
}

int orc_contract_single(orc_bm* out, const orc_bm* L, const orc_bm* A, const orc_mpo* w, const orc_bm* R,
                        const uint64_t* itk, int d, int nthreads) {
  ORC_TRY({
    if (A->is_c) contractSingle(out->z1, L->z1, A->z1, w->z, R->z1, mk_itk(itk, d), nthreads);
    else contractSingle(out->d1, L->d1, A->d1, w->d, R->d1, mk_itk(itk, d), nthreads);
  })
}

int orc_add_layer(orc_bm* out, const orc_bm* BL, const orc_bm* A, const orc_mpo* w, const uint64_t* itk, int d,
                  int direction, int nthreads, const orc_bm* Ap) {
  ORC_TRY({
    if (A->is_c) out->z1 = addContractionLayer(BL->z1, A->z1, w->z, mk_itk(itk, d), direction, nthreads, Ap ? &Ap->z1 : nullptr);
    else out->d1 = addContractionLayer(BL->d1, A->d1, w->d, mk_itk(itk, d), direction, nthreads, Ap ? &Ap->d1 : nullptr);
  })
}

int orc_dot(const orc_bm* X, const orc_bm* Y, double* re, double* im) {
  ORC_TRY({
    cplx v;
    if (X->ns == 1) v = X->is_c ? dot(X->z1, Y->z1) : cplx(dot(X->d1, Y->d1));
    else v = X->is_c ? dot(X->z2, Y->z2) : cplx(dot(X->d2, Y->d2));
    *re = v.real();
    *im = v.imag();
  })
}

int orc_twosite(orc_bm* out, const orc_bm* B1, const orc_bm* B2) {
  ORC_TRY({
    if (B1->is_c) out->z2 = twosite(B1->z1, B2->z1);
    else out->d2 = twosite(B1->d1, B2->d1);
  })
}

int orc_lmin(orc_bm* out) { ORC_TRY({ out->d1 = getLMin<double>(); out->z1 = getLMin<cplx>(); }) }
int orc_rmax(orc_bm* out, const orc_bm* Alast) {
  ORC_TRY({
    if (Alast->is_c) out->z1 = getRMax(Alast->z1);
    else out->d1 = getRMax(Alast->d1);
  })
}

// <psi|H|psi> over a chain of n sites (real MPS only).
int orc_measure(int n, const orc_bm* const* mps, const orc_mpo* const* mpo, const uint64_t* itk, int d, int nthreads,
                double* E) {
  ORC_TRY({
    std::vector<BlockMatrix<1, double>> A;
    std::vector<MPOMat<double>> W;
    std::vector<ItoKey> I;
    for (int s = 0; s < n; ++s) {
      A.push_back(mps[s]->d1);
      W.push_back(mpo[s]->d);
      I.push_back(mk_itk(itk + (size_t)s * d, d));
    }
    *E = measure(A, W, I, nthreads);
  })
}

// Two-site Lanczos (TSLanzcosEngine::simulate) on real data.  energies gets
// the final (sorted) Ritz values, at most cap of them; *nit the iteration count,
// *nritz the number of Ritz values; gs receives getState(target).
int orc_lanczos_twosite_state(const orc_bm* L, const orc_bm* R, const orc_bm* psi0, const orc_mpo* wl,
                              const orc_mpo* wr, const uint64_t* itl, int dl, const uint64_t* itr, int dr, int itmax,
                              double delta, int testenergy, int nthreads, int target, double* energies, int cap,
                              int* nit, int* nritz, orc_bm* gs, const orc_bm* state);
int orc_lanczos_twosite(const orc_bm* L, const orc_bm* R, const orc_bm* psi0, const orc_mpo* wl, const orc_mpo* wr,
                        const uint64_t* itl, int dl, const uint64_t* itr, int dr, int itmax, double delta,
                        int testenergy, int nthreads, int target, double* energies, int cap, int* nit, int* nritz,
                        orc_bm* gs) {
  return orc_lanczos_twosite_state(L, R, psi0, wl, wr, itl, dl, itr, dr, itmax, delta, testenergy, nthreads, target,
                                   energies, cap, nit, nritz, gs, nullptr);
}

// The projected run (src/lanzcos.hpp:281-414) when state is given.
int orc_lanczos_twosite_state(const orc_bm* L, const orc_bm* R, const orc_bm* psi0, const orc_mpo* wl,
                              const orc_mpo* wr, const uint64_t* itl, int dl, const uint64_t* itr, int dr, int itmax,
                              double delta, int testenergy, int nthreads, int target, double* energies, int cap,
                              int* nit, int* nritz, orc_bm* gs, const orc_bm* state) {
  ORC_TRY({
    ItoKey il = mk_itk(itl, dl), ir = mk_itk(itr, dr);
    auto mv = [&](BlockMatrix<2, double>& y, const BlockMatrix<2, double>& x) {
      contractTwosite(y, L->d1, R->d1, x, wl->d, wr->d, il, ir, nthreads);
    };
    LanczosResult<double> r =
        lanczos<double>(psi0->d2, mv, itmax, delta, testenergy != 0, state ? &state->d2 : nullptr);
    *nit = r.it;
    *nritz = (int)r.es.size();
    for (int i = 0; i < (int)r.es.size() && i < cap; ++i) energies[i] = r.es[(size_t)i];
    if (gs && !r.es.empty()) gs->d2 = getState(r, target);
  })
}

int orc_tridiag(int n, double* d, const double* e, double* z) {
  ORC_TRY({
    std::vector<double> dd(d, d + n), ee(e, e + std::max(0, n - 1)), zz;
    int info = tridiag(dd, ee, z ? &zz : nullptr);
    if (info) throw std::runtime_error("dstev failed");
    std::memcpy(d, dd.data(), sizeof(double) * (size_t)n);
    if (z) std::memcpy(z, zz.data(), sizeof(double) * (size_t)n * n);
  })
}

}  // extern "C"
