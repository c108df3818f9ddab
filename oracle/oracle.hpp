// oracle.hpp -- CPU restatement of the MPSTools DMRG hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker: it restates the
// reference algorithm (block-sparse two-site/one-site H_eff contraction,
// environment updates, plain Lanczos) literally, with the reference's loop
// order, std::map key order, OpenMP-over-charges partition and one BLAS call
// per tuple.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may call it.  The product library (mpstools_amd) never links it.
//
// Every function cites the reference file:line it follows.
#pragma once
#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

typedef std::complex<double> cplx;

// ---------------------------------------------------------------- charges
// AbKeyType (src/keyTypes.hpp:20-124): up to two U(1) charges packed as
// _qn = (q2<<16) + q1 in an unsigned long; + and - act on the packed word
// with unsigned wrap-around (keyTypes.hpp:46-65); order is the packed-integer
// order (keyTypes.hpp:89-116).  A plain uint64_t has exactly these semantics.
typedef uint64_t QN;
inline QN pack_qn(uint64_t q1, uint64_t q2) { return (q2 << 16) + q1; }  // keyTypes.hpp:25-30

// TensorKeyType<2,Ns,AbKeyType> (keyTypes.hpp:167-226): two charges
// (q[0] = left/upper, q[1] = right/lower) and Ns physical (or MPO aux)
// indices; operator< compares charges first, then physical indices,
// lexicographically (keyTypes.hpp:313-343).
template <int Ns>
struct TKey {
  QN q[2];
  uint32_t p[Ns];
  bool operator<(const TKey& o) const {
    if (q[0] != o.q[0]) return q[0] < o.q[0];
    if (q[1] != o.q[1]) return q[1] < o.q[1];
    for (int i = 0; i < Ns; ++i)
      if (p[i] != o.p[i]) return p[i] < o.p[i];
    return false;
  }
  bool operator==(const TKey& o) const { return !(*this < o) && !(o < *this); }
};
template <int Ns>
inline TKey<Ns> mk(QN a, QN b, const uint32_t* p) {
  TKey<Ns> k;
  k.q[0] = a;
  k.q[1] = b;
  for (int i = 0; i < Ns; ++i) k.p[i] = p[i];
  return k;
}
inline TKey<1> mk1(QN a, QN b, uint32_t p) {
  TKey<1> k;
  k.q[0] = a;
  k.q[1] = b;
  k.p[0] = p;
  return k;
}

// OpKeyType<1> (keyTypes.hpp:384-421): OpKeyType(out,in) stores in in [0]
// and out in (0); the TensorKeyType order therefore sorts by (in, out).
struct OpKey {
  uint32_t in, out;
  bool operator<(const OpKey& o) const {
    if (in != o.in) return in < o.in;
    return out < o.out;
  }
};
// ukey<2> (keyTypes.hpp:424-483): MPO aux pair (wl, wr), lexicographic.
struct UKey {
  uint32_t wl, wr;
  bool operator<(const UKey& o) const {
    if (wl != o.wl) return wl < o.wl;
    return wr < o.wr;
  }
};

// ---------------------------------------------------------------- dense
template <typename T>
struct Mat {  // boost::numeric::ublas::matrix<T,column_major>
  int64_t r = 0, c = 0;
  std::vector<T> a;
  Mat() {}
  Mat(int64_t r_, int64_t c_) : r(r_), c(c_), a((size_t)(r_ * c_), T(0)) {}
  T& operator()(int64_t i, int64_t j) { return a[(size_t)(i + j * r)]; }
  const T& operator()(int64_t i, int64_t j) const { return a[(size_t)(i + j * r)]; }
};

inline double conj_(double x) { return x; }
inline cplx conj_(cplx x) { return std::conj(x); }
template <typename T>
Mat<T> conjm(const Mat<T>& m) {  // boost conj() copy used at MPSfunctions.hpp:336,339
  Mat<T> o = m;
  for (auto& x : o.a) x = conj_(x);
  return o;
}

// BLAS hooks (set by blas.cpp: MKL when present, reference loops otherwise).
typedef void (*dgemm_fn)(const char*, const char*, const int*, const int*, const int*, const double*,
                         const double*, const int*, const double*, const int*, const double*, double*,
                         const int*);
typedef void (*zgemm_fn)(const char*, const char*, const int*, const int*, const int*, const cplx*,
                         const cplx*, const int*, const cplx*, const int*, const cplx*, cplx*,
                         const int*);
typedef void (*dstev_fn)(const char*, const int*, double*, double*, double*, const int*, double*,
                         int*);
typedef void (*set_threads_fn)(int);
struct Blas {
  dgemm_fn dgemm = nullptr;
  zgemm_fn zgemm = nullptr;
  dstev_fn dstev = nullptr;
  set_threads_fn set_threads = nullptr;  // MKL_Set_Num_Threads / openblas_set_num_threads
  std::string name = "loops";
};
Blas& blas();

template <typename T>
void gemm_loops(char ta, char tb, int M, int N, int K, T alpha, const T* A, int lda, const T* B,
                int ldb, T* C, int ldc) {
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) {
      T s = T(0);
      for (int k = 0; k < K; ++k) {
        T a = (ta == 'n') ? A[i + (size_t)k * lda] : A[k + (size_t)i * lda];
        T b = (tb == 'n') ? B[k + (size_t)j * ldb] : B[j + (size_t)k * ldb];
        s += a * b;
      }
      C[i + (size_t)j * ldc] = alpha * s;
    }
}
inline void gemm_raw(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
                     const double* B, int ldb, double* C, int ldc) {
  if (M == 0 || N == 0) return;
  if (blas().dgemm && K > 0) {
    double beta = 0.0;
    blas().dgemm(&ta, &tb, &M, &N, &K, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
  } else {
    gemm_loops<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, C, ldc);
  }
}
inline void gemm_raw(char ta, char tb, int M, int N, int K, cplx alpha, const cplx* A, int lda,
                     const cplx* B, int ldb, cplx* C, int ldc) {
  if (M == 0 || N == 0) return;
  if (blas().zgemm && K > 0) {
    cplx beta = 0.0;
    blas().zgemm(&ta, &tb, &M, &N, &K, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
  } else {
    gemm_loops<cplx>(ta, tb, M, N, K, alpha, A, lda, B, ldb, C, ldc);
  }
}

// MatMatProd (blas/blasroutines.cpp:837-848 -> mat_mat_prod :1239-1274):
// out is resized to op(A) x op(B), beta = 0, ld = leading dims of the blocks.
template <typename T>
void MatMatProd(T alpha, const Mat<T>& A, const Mat<T>& B, Mat<T>& out, char ta, char tb) {
  const int M = (int)(ta == 'n' ? A.r : A.c);
  const int Ka = (int)(ta == 'n' ? A.c : A.r);
  const int N = (int)(tb == 'n' ? B.c : B.r);
  const int Kb = (int)(tb == 'n' ? B.r : B.c);
  if (Ka != Kb) throw std::runtime_error("MatMatProd: inner dimensions differ");
  out = Mat<T>(M, N);
  gemm_raw(ta, tb, M, N, Ka, alpha, A.a.data(), std::max<int>(1, (int)A.r), B.a.data(),
           std::max<int>(1, (int)B.r), out.a.data(), std::max(1, M));
}

// ---------------------------------------------------------------- block tensors
template <int Ns, typename T>
using BlockMatrix = std::map<TKey<Ns>, Mat<T>>;  // BlockTensor.hpp:279-379

template <typename T>
using SparseOperator = std::map<OpKey, T>;  // MPOTypes.hpp:50-129

template <typename T>
struct MPOMat {  // MPOTypes.hpp:839-899
  std::map<UKey, SparseOperator<T>> m;
  uint32_t Dl = 0, Dr = 0;
};

typedef std::vector<QN> ItoKey;  // i_to_key (tensormap.hpp:280-295): local state -> charge

// merge (tensormap.hpp:130-135): overwrite duplicate keys.
template <int Ns, typename T>
void merge(BlockMatrix<Ns, T>& out, const BlockMatrix<Ns, T>& in) {
  for (auto& kv : in) out[kv.first] = kv.second;
}

// Block-matrix scalar product X*Y = sum conj(Y).X over matching keys
// (BlockTensor.hpp:520-544 with blasroutines.cpp:183-209).
template <int Ns, typename T>
T dot(const BlockMatrix<Ns, T>& X, const BlockMatrix<Ns, T>& Y) {
  T val = T(0);
  for (auto& kv : X) {
    auto it = Y.find(kv.first);
    if (it == Y.end()) continue;
    if (it->second.a.size() != kv.second.a.size())
      throw std::runtime_error("BlockMatrix::operator*: size1==size2 failed");
    T s = T(0);
    for (size_t n = 0; n < kv.second.a.size(); ++n) s += conj_(it->second.a[n]) * kv.second.a[n];
    val = val + s;
  }
  return val;
}

// X += a*Y with the reference's union-of-keys semantics (BlockTensor.hpp:443-518).
template <int Ns, typename T>
void axpy(BlockMatrix<Ns, T>& X, T a, const BlockMatrix<Ns, T>& Y) {
  for (auto& kv : Y) {
    auto it = X.find(kv.first);
    if (it == X.end()) {
      Mat<T> m = kv.second;
      for (auto& x : m.a) x *= a;
      X[kv.first] = m;
    } else {
      for (size_t n = 0; n < kv.second.a.size(); ++n) it->second.a[n] += a * kv.second.a[n];
    }
  }
}
template <int Ns, typename T>
void scale(BlockMatrix<Ns, T>& X, T a) {  // BlockTensor.hpp:471-477
  for (auto& kv : X)
    for (auto& x : kv.second.a) x *= a;
}

// Two-site wave function from two MPS tensors (BlockTensor.hpp:285-301).
template <typename T>
BlockMatrix<2, T> twosite(const BlockMatrix<1, T>& B1, const BlockMatrix<1, T>& B2) {
  BlockMatrix<2, T> out;
  for (auto& a : B1)
    for (auto& b : B2)
      if (a.first.q[1] == b.first.q[0]) {
        uint32_t p[2] = {a.first.p[0], b.first.p[0]};
        Mat<T> tmp;
        MatMatProd(T(1.0), a.second, b.second, tmp, 'n', 'n');
        out[mk<2>(a.first.q[0], b.first.q[1], p)] = tmp;
      }
  return out;
}

// ---------------------------------------------------------------- environments
// add_out_sub_2 / add_out_sub (src/MPSfunctions.hpp:276-352, 359-435).  The
// Ap variant (mixed bra/ket) takes the upper matrix from Ap instead of A.
template <typename T>
BlockMatrix<1, T> add_out_sub(const std::map<uint32_t, std::vector<const std::pair<const TKey<1>, Mat<T>>*>>& BLits,
                              const BlockMatrix<1, T>& A, const MPOMat<T>& mpo, const BlockMatrix<1, T>& Aup,
                              QN p, const ItoKey& itokey, int direction) {
  const int idir = (direction > 0 ? 1 : 0);
  const int out_a = idir;
  BlockMatrix<1, T> B;
  for (auto& a : A)
    if (a.first.q[out_a] == p) B[a.first] = a.second;  // :284-287
  BlockMatrix<1, T> Bout;
  for (auto& o : mpo.m) {  // :290
    const UKey& key_o = o.first;
    const uint32_t ko[2] = {key_o.wl, key_o.wr};
    const uint32_t o_in = ko[1 - idir];
    for (auto& oi : o.second) {  // :296
      const QN kyout = itokey.at(oi.first.out);
      TKey<1> key_b = direction > 0 ? mk1(p - kyout, p, oi.first.out) : mk1(p, p + kyout, oi.first.out);
      auto posb = B.find(key_b);
      if (posb == B.end()) continue;  // :306
      auto bl = BLits.find(o_in);
      if (bl == BLits.end()) continue;
      for (auto* itBL : bl->second) {  // :311
        const TKey<1>& kBL = itBL->first;
        if (direction > 0 && kBL.p[0] != ko[0]) continue;
        if (direction <= 0 && kBL.p[0] != ko[1]) continue;
        if (direction > 0 && kBL.q[1] != key_b.q[0]) continue;
        if (direction <= 0 && kBL.q[1] != key_b.q[1]) continue;
        const QN kyin = itokey.at(oi.first.in);
        TKey<1> key_a = direction > 0 ? mk1(kBL.q[0], kBL.q[0] + kyin, oi.first.in)
                                      : mk1(kBL.q[0] - kyin, kBL.q[0], oi.first.in);
        auto posa = Aup.find(key_a);
        if (posa == Aup.end()) continue;  // :327
        const Mat<T>& Bm = itBL->second;
        const Mat<T>& au = posa->second;
        const Mat<T>& ad = posb->second;
        Mat<T> Tmp1, Tmp2;
        if (direction > 0) {  // :335-341
          MatMatProd(T(1.0), Bm, conjm(au), Tmp1, 't', 'n');
          MatMatProd(T(1.0), Tmp1, ad, Tmp2, 't', 'n');
        } else {
          MatMatProd(T(1.0), conjm(au), Bm, Tmp1, 'n', 'n');
          MatMatProd(T(1.0), Tmp1, ad, Tmp2, 'n', 't');
        }
        for (auto& x : Tmp2.a) x *= oi.second;  // :342
        TKey<1> ky = mk1(key_a.q[idir], key_b.q[idir], ko[idir]);
        auto ypos = Bout.find(ky);
        if (ypos != Bout.end())
          for (size_t n = 0; n < Tmp2.a.size(); ++n) Tmp2.a[n] += ypos->second.a[n];
        Bout[ky] = Tmp2;
      }
    }
  }
  return Bout;
}

// addContractionLayer (MPSfunctions.hpp:442-493; Ap variant :499-555).
template <typename T>
BlockMatrix<1, T> addContractionLayer(const BlockMatrix<1, T>& BL, const BlockMatrix<1, T>& A,
                                      const MPOMat<T>& mpo, const ItoKey& itokey, int direction,
                                      int maxthreads, const BlockMatrix<1, T>* Ap = nullptr) {
  if (A.empty()) throw std::runtime_error("Error in addContractionLayer: A is empty.");
  if (Ap && Ap->empty()) throw std::runtime_error("Error in addContractionLayer: Ap is empty.");
  const int idir = (direction > 0 ? 1 : 0);
  std::set<QN> outgoing_set;
  for (auto& a : A) outgoing_set.insert(a.first.q[idir]);
  std::vector<QN> outgoing(outgoing_set.begin(), outgoing_set.end());
  const int n_out = (int)outgoing.size();
  std::map<uint32_t, std::vector<const std::pair<const TKey<1>, Mat<T>>*>> BLits;
  for (auto& b : BL) BLits[b.first.p[0]].push_back(&b);  // :475-478
  if (maxthreads < 1) maxthreads = 1;
  std::vector<BlockMatrix<1, T>> bouts((size_t)maxthreads);
  const BlockMatrix<1, T>& Aup = Ap ? *Ap : A;
  // With one outgoing charge the region stays inactive, so the BLAS calls
  // inside run multi-threaded instead of on the one thread of the loop.
#pragma omp parallel for num_threads(maxthreads) schedule(static) if (n_out > 1)
  for (int i = 0; i < n_out; ++i) {
    int jthr = 0;
#ifdef _OPENMP
    jthr = omp_get_thread_num();
#endif
    BlockMatrix<1, T> Bi = add_out_sub(BLits, A, mpo, Aup, outgoing[i], itokey, direction);
    merge(bouts[(size_t)jthr], Bi);
  }
  BlockMatrix<1, T> Bout;
  for (auto& b : bouts) merge(Bout, b);
  return Bout;
}

// getLMin (MPSfunctions.hpp:1917-1926): 1x1 identity at charge 0, aux 0.
template <typename T>
BlockMatrix<1, T> getLMin() {
  BlockMatrix<1, T> L;
  Mat<T> m(1, 1);
  m(0, 0) = T(1);
  L[mk1(0, 0, 0)] = m;
  return L;
}
// getRMax (MPSfunctions.hpp:1996-2022): identity per right charge of the last tensor, aux 0.
template <typename T>
BlockMatrix<1, T> getRMax(const BlockMatrix<1, T>& Alast) {
  std::set<QN> ps;
  for (auto& a : Alast) ps.insert(a.first.q[1]);
  BlockMatrix<1, T> R;
  for (QN p : ps) {
    int64_t m = 0;
    for (auto& a : Alast)
      if (a.first.q[1] == p) {
        m = a.second.c;
        break;
      }
    Mat<T> id(m, m);
    for (int64_t i = 0; i < m; ++i) id(i, i) = T(1);
    R[mk1(p, p, 0)] = id;
  }
  return R;
}

// MPO<T>::measure (MPOTypes.hpp:990-1017): <psi|H|psi> (unnormalised) from a
// left sweep; trace over charge-diagonal blocks of the final environment.
template <typename T>
T measure(const std::vector<BlockMatrix<1, T>>& mps, const std::vector<MPOMat<T>>& mpo,
          const std::vector<ItoKey>& itokey, int maxthreads) {
  BlockMatrix<1, T> L = getLMin<T>();
  for (size_t s = 0; s < mps.size(); ++s)
    L = addContractionLayer(L, mps[s], mpo[s], itokey[s], 1, maxthreads);
  T out = T(0);
  for (auto& l : L)
    if (l.first.q[0] == l.first.q[1]) {
      int64_t n = std::min(l.second.r, l.second.c);
      for (int64_t i = 0; i < n; ++i) out += l.second(i, i);
    }
  return out;
}

// ---------------------------------------------------------------- H_eff . psi
// Two-site doContraction (MPSfunctions.hpp:726-804).
template <typename T>
void doContraction2(BlockMatrix<2, T>& Aout, const std::vector<const std::pair<const TKey<1>, Mat<T>>*>& BLits,
                    const BlockMatrix<1, T>& BR, const BlockMatrix<2, T>& Ain, const MPOMat<T>& mpol,
                    const MPOMat<T>& mpor, const ItoKey& itl, const ItoKey& itr) {
  for (auto* itBL : BLits) {                                    // :738
    for (auto& o1 : mpol.m) {                                   // :739
      for (auto& o2 : mpor.m) {                                 // :740
        if (o1.first.wl != itBL->first.p[0]) continue;          // :741
        if (o2.first.wl != o1.first.wr) continue;               // :742
        const uint32_t o_out = o2.first.wr;                     // :745
        for (auto& oil : o1.second) {                           // :748
          for (auto& oir : o2.second) {                         // :749
            const QN kylo = itl.at(oil.first.out) + itr.at(oir.first.out);  // :756-758
            const uint32_t out[2] = {oil.first.out, oir.first.out};
            const uint32_t in[2] = {oil.first.in, oir.first.in};
            auto key_mat = mk<2>(itBL->first.q[1], itBL->first.q[1] + kylo, out);  // :761
            auto pos = Ain.find(key_mat);
            if (pos == Ain.end()) continue;
            const QN kyup = itl.at(oil.first.in) + itr.at(oir.first.in);  // :770-772
            auto kyR = mk1(itBL->first.q[0] + kyup, itBL->first.q[1] + kylo, o_out);  // :775
            auto posR = BR.find(kyR);
            if (posR == BR.end()) continue;
            Mat<T> Tmp1, Tmp2;
            T num = oil.second * oir.second;                    // :787
            auto key_out = mk<2>(itBL->first.q[0], posR->first.q[0], in);  // :792
            MatMatProd(num, itBL->second, pos->second, Tmp1, 'n', 'n');  // :793
            MatMatProd(T(1.0), Tmp1, posR->second, Tmp2, 'n', 't');     // :794
            auto ypos = Aout.find(key_out);
            if (ypos != Aout.end())
              for (size_t n = 0; n < Tmp2.a.size(); ++n) Tmp2.a[n] += ypos->second.a[n];
            Aout[key_out] = Tmp2;                               // :798
          }
        }
      }
    }
  }
}

// contractTwosite (MPSfunctions.hpp:949-981): OpenMP over left charges of BL,
// thread-private outputs merged.
template <typename T>
void contractTwosite(BlockMatrix<2, T>& Aout, const BlockMatrix<1, T>& BL, const BlockMatrix<1, T>& BR,
                     const BlockMatrix<2, T>& A, const MPOMat<T>& mpol, const MPOMat<T>& mpor,
                     const ItoKey& itl, const ItoKey& itr, int maxthreads) {
  Aout.clear();
  std::map<QN, std::vector<const std::pair<const TKey<1>, Mat<T>>*>> BLits;
  std::set<QN> pleft;
  for (auto& l : BL) {
    BLits[l.first.q[0]].push_back(&l);
    pleft.insert(l.first.q[0]);
  }
  std::vector<QN> lefts(pleft.begin(), pleft.end());
  const int n = (int)lefts.size();
  if (maxthreads < 1) maxthreads = 1;
  std::vector<BlockMatrix<2, T>> aout((size_t)maxthreads);
  // one left charge (dense case): inactive region, multi-threaded BLAS inside
#pragma omp parallel for num_threads(maxthreads) schedule(static) if (n > 1)
  for (int i = 0; i < n; ++i) {
    int j = 0;
#ifdef _OPENMP
    j = omp_get_thread_num();
#endif
    doContraction2(aout[(size_t)j], BLits.at(lefts[i]), BR, A, mpol, mpor, itl, itr);
  }
  for (auto& a : aout) merge(Aout, a);
}

// One-site doContraction (MPSfunctions.hpp:808-854) and contractSingle (:857-885).
template <typename T>
void doContraction1(BlockMatrix<1, T>& Aout, const std::vector<const std::pair<const TKey<1>, Mat<T>>*>& BLits,
                    const BlockMatrix<1, T>& BR, const BlockMatrix<1, T>& Ain, const MPOMat<T>& mpo,
                    const ItoKey& itokey) {
  for (auto* itBL : BLits) {
    for (auto& o : mpo.m) {
      const uint32_t o_out = o.first.wr, o_in = o.first.wl;
      if (o_in != itBL->first.p[0]) continue;
      for (auto& oi : o.second) {
        const QN kylo = itokey.at(oi.first.out);
        auto key_mat = mk1(itBL->first.q[1], itBL->first.q[1] + kylo, oi.first.out);
        auto pos = Ain.find(key_mat);
        if (pos == Ain.end()) continue;
        const QN kyup = itokey.at(oi.first.in);
        auto kyR = mk1(itBL->first.q[0] + kyup, itBL->first.q[1] + kylo, o_out);
        auto posR = BR.find(kyR);
        if (posR == BR.end()) continue;
        Mat<T> Tmp1, Tmp2;
        T num = oi.second;
        auto key_out = mk1(itBL->first.q[0], posR->first.q[0], oi.first.in);
        MatMatProd(num, itBL->second, pos->second, Tmp1, 'n', 'n');
        MatMatProd(T(1.), Tmp1, posR->second, Tmp2, 'n', 't');
        auto ypos = Aout.find(key_out);
        if (ypos != Aout.end())
          for (size_t n = 0; n < Tmp2.a.size(); ++n) Tmp2.a[n] += ypos->second.a[n];
        Aout[key_out] = Tmp2;
      }
    }
  }
}
template <typename T>
void contractSingle(BlockMatrix<1, T>& Aout, const BlockMatrix<1, T>& BL, const BlockMatrix<1, T>& A,
                    const MPOMat<T>& mpo, const BlockMatrix<1, T>& BR, const ItoKey& itokey, int maxthreads) {
  Aout.clear();
  std::map<QN, std::vector<const std::pair<const TKey<1>, Mat<T>>*>> BLits;
  std::set<QN> pleft;
  for (auto& l : BL) {
    BLits[l.first.q[0]].push_back(&l);
    pleft.insert(l.first.q[0]);
  }
  std::vector<QN> lefts(pleft.begin(), pleft.end());
  const int n = (int)lefts.size();
  if (maxthreads < 1) maxthreads = 1;
  std::vector<BlockMatrix<1, T>> aout((size_t)maxthreads);
#pragma omp parallel for num_threads(maxthreads) schedule(static) if (n > 1)
  for (int i = 0; i < n; ++i) {
    int j = 0;
#ifdef _OPENMP
    j = omp_get_thread_num();
#endif
    doContraction1(aout[(size_t)j], BLits.at(lefts[i]), BR, A, mpo, itokey);
  }
  for (auto& a : aout) merge(Aout, a);
}

// ---------------------------------------------------------------- Lanczos
// Symmetric tridiagonal eigensolve (tridiag, blas/lapack_routines.cpp:435-446):
// LAPACK dstev when a BLAS/LAPACK library is loaded, else cyclic Jacobi on the
// dense tridiagonal matrix (small n only).  Eigenvalues ascending; eigenvectors
// column-major n x n.
int tridiag(std::vector<double>& d, std::vector<double> e, std::vector<double>* z);

// TSLanzcosEngine::simulate (src/lanzcos.hpp:165-278) with the reference's
// exact recurrence order and both convergence tests (testenergy: lowest Ritz
// value every _Ndiag=5 steps; otherwise the residual of the Ritz vector).  The
// matvec is contractTwosite with the given environments and MPO tensors.
template <typename T>
struct LanczosResult {
  int it = 0;
  std::vector<T> es, ks;
  std::vector<double> eigs;  // column-major size x size
  std::vector<BlockMatrix<2, T>> As;
};

// BlockMatrix::squeeze (BlockTensor.hpp:417-441): drop every block whose
// entries are all at most thresh in magnitude.
template <int Ns, typename T>
void squeeze(BlockMatrix<Ns, T>& X, double thresh) {
  for (auto it = X.begin(); it != X.end();) {
    bool er = true;
    for (auto& x : it->second.a)
      if (std::abs(x) > thresh) {
        er = false;
        break;
      }
    it = er ? X.erase(it) : std::next(it);
  }
}

// With `state` the run is the projected one of
// TSLanzcosEngine::simulate(itmax, delta, testenergy, state, delta_ortho)
// (src/lanzcos.hpp:281-414): state is squeezed at 1e-15, the start vector
// and every new recurrence vector lose their component along state,
// A <- A - state ((state*A) / (state*state)); delta_ortho is unused there
// (its Gram-Schmidt block is commented out).
template <typename T, typename MV>
LanczosResult<T> lanczos(const BlockMatrix<2, T>& Ainit, MV matvec, int itmax, double delta, bool testenergy,
                         const BlockMatrix<2, T>* state_in = nullptr) {
  LanczosResult<T> res;
  int it = 0;
  T k = 0, e, eold = T(1e100);
  bool converged = false, first = true;
  std::vector<T> estmp, kstmp;
  BlockMatrix<2, T> A0, A1 = Ainit, A2, A3;
  double resi = 1e5;
  const int Ndiag = 5;  // lanzcos.hpp:107
  BlockMatrix<2, T> state;
  T state_norm = T(0);
  if (state_in) {
    state = *state_in;
    squeeze(state, 1e-15);     // :296
    state_norm = dot(state, state);
  }
  auto project = [&](BlockMatrix<2, T>& A) {  // :298, :340, :372
    if (!state_in) return;
    const T c = dot(state, A) / state_norm;
    axpy(A, -c, state);
  };
  project(A1);
  while (!converged) {
    k = std::sqrt(dot(A1, A1));                     // :187
    if (std::abs(k) < delta) { converged = true; break; }
    if (!first) res.ks.push_back(k);
    scale(A1, T(1.0) / k);                          // :190 (operator/= multiplies by 1/num)
    matvec(A2, A1);                                 // :197
    res.As.push_back(A1);                           // :204-207
    e = dot(A2, A1);                                // :208
    res.es.push_back(e);
    if (it % Ndiag == 0 && it != 0) {               // :210
      if (testenergy) {
        kstmp = res.ks;
        estmp = res.es;
        std::vector<double> d(estmp.size()), sub(kstmp.size());
        for (size_t i = 0; i < d.size(); ++i) d[i] = std::real(estmp[i]);
        for (size_t i = 0; i < sub.size(); ++i) sub[i] = std::real(kstmp[i]);
        tridiag(d, sub, nullptr);
        if (std::abs(d[0] - std::real(eold)) < delta) { converged = true; break; }
        eold = d[0];
      } else {
        kstmp = res.ks;
        estmp = res.es;
        std::vector<double> d(estmp.size()), sub(kstmp.size()), z;
        for (size_t i = 0; i < d.size(); ++i) d[i] = std::real(estmp[i]);
        for (size_t i = 0; i < sub.size(); ++i) sub[i] = std::real(kstmp[i]);
        tridiag(d, sub, &z);
        BlockMatrix<2, T> A4 = res.As[0];
        scale(A4, T(z[0]));
        for (size_t i = 1; i < estmp.size(); ++i) axpy(A4, T(z[i]), res.As[i]);
        BlockMatrix<2, T> A5;
        matvec(A5, A4);
        BlockMatrix<2, T> A6 = A5;
        axpy(A6, T(-d[0]), A4);
        resi = std::abs(std::sqrt(dot(A6, A6)));
        if (std::abs(resi) < delta) { converged = true; break; }
      }
    }
    if (!first) {                                   // :237-242
      A3 = A1;
      scale(A3, -e);
      axpy(A3, T(1.0), A2);
      BlockMatrix<2, T> t0 = A0;
      scale(t0, -k);
      axpy(A3, T(1.0), t0);
      project(A3);
      A0 = A1;
      A1 = A3;
    } else {                                        // :243-248
      A0 = A1;
      A3 = A1;
      scale(A3, -e);
      axpy(A3, T(1.0), A2);
      project(A3);
      A1 = A3;
      first = false;
    }
    it++;
    if (it >= itmax) { converged = true; break; }   // :249-256
  }
  if (converged && !res.es.empty()) {               // :258-268
    const int size = (int)res.es.size();
    std::vector<double> d(size), sub(res.ks.size());
    for (int i = 0; i < size; ++i) d[i] = std::real(res.es[i]);
    for (size_t i = 0; i < sub.size(); ++i) sub[i] = std::real(res.ks[i]);
    // dstev overwrites the diagonal with eigenvalues; the reference passes
    // _es itself, so E(0) afterwards is the lowest Ritz value.
    tridiag(d, sub, &res.eigs);
    for (int i = 0; i < size; ++i) res.es[i] = T(d[i]);
  }
  res.it = it;
  return res;
}

// LanzcosEngine::getState (lanzcos.hpp:81-92): sum_i eig(i,j) A_i.
template <typename T>
BlockMatrix<2, T> getState(const LanczosResult<T>& r, int j) {
  const int size = (int)r.es.size();
  if (j >= size) throw std::runtime_error("LanzcosEngine::getState(const int j): j<_eigs.size2() failed");
  BlockMatrix<2, T> A = r.As[0];
  scale(A, T(r.eigs[(size_t)(0 + j * size)]));
  for (int i = 1; i < size; ++i) axpy(A, T(r.eigs[(size_t)(i + j * size)]), r.As[(size_t)i]);
  return A;
}

}  // namespace oracle
