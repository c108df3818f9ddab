// mpsa_cheb.cpp -- Chebyshev moments with variational MPS compression on the
// device: the reference's ChebychevEngine::simulateSSCompress
// (src/container.hpp:2343-2470) and the MPS functions under it,
//   applyMPO            src/MPSfunctions.hpp:2088-2345
//   computeOverlap      src/MPSfunctions.hpp:2066-2084
//   prepareSite         src/MPSTypes.hpp:586-620 (prepareMatrix/prepareBlock,
//                       src/MPSfunctions.hpp:21-269)
//   truncate            src/MPSfunctions.hpp:1705-1727 with
//                       DiagBlockMatrix::truncate src/BlockTensor.hpp:85-136
//   dropDangling        src/MPSfunctions.hpp:1444-1484
//   compress            src/MPSfunctions.hpp:2353-2510
//   compressSingleSite  src/MPSfunctions.hpp:2670-2789
//   MPS *=, +=, -=      src/MPSTypes.hpp:223-300 (outersum, fixSizes)
//
// Batching: the nvec start vectors share one set of block tables.  Vector v's
// charges carry v in bits 40..55 of the packed AbKeyType word (charges and
// i_to_key values of the models stay far below 2^40), so blocks of different
// vectors never meet in a contraction: every H_eff, environment, SVD and copy
// of a step covers all vectors in one grouped launch, while truncation,
// overlaps and the convergence tests are evaluated per vector.
//
// Device work per step: the contractions go through the plans of
// mpsa_plan.cpp (one-site H_eff and environment updates with a distinct upper
// tensor, cached by block structure); canonicalisation gathers each charge
// sector into a dense matrix (copy2d), factors all sectors with the batched
// Jacobi SVD (mpsa_factor.hip) and multiplies S V^dagger or U S into the
// neighbour with one grouped GEMM.  Host work is table bookkeeping only.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <list>
#include <unordered_map>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_host.h"
#include "mpsa_internal.h"

using namespace mpsa;
using namespace mpsa::host;

#ifndef MPSA_BLOCK_JACOBI_MIN
#define MPSA_BLOCK_JACOBI_MIN 32
#endif
#ifndef MPSA_BIG_SECTOR
#define MPSA_BIG_SECTOR 1024
#endif

namespace {

constexpr int kNsShift = 40;
inline uint64_t ns_of(uint64_t q) { return q >> kNsShift; }
typedef std::complex<double> cplx;
typedef std::array<uint64_t, 3> Key;  // (left charge, right charge, physical index)

// ------------------------------------------------------------ device memory
// Buffers are recycled within one stream: a buffer handed back is reused only
// by work queued later on the same stream.
struct Pool {
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> all;  // every allocation and its size
  size_t idle = 0;
  double t_alloc = 0.0;  // host seconds in hipMalloc / hipFree, and their count
  int64_t n_alloc = 0;
  void* get(size_t bytes) {
    bytes = std::max<size_t>((bytes + 255) / 256 * 256, 256);
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes) {
      void* p = it->second;
      idle -= it->first;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      release_idle();
      if (hipMalloc(&p, bytes) != hipSuccess) throw HErr(MPSA_ENOMEM, "cheb: device allocation failed");
    }
    t_alloc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++n_alloc;
    all[p] = bytes;
    return p;
  }
  size_t size_of(void* p) const {
    auto f = all.find(p);
    return f == all.end() ? 0 : f->second;
  }
  void put(void* p) {
    const size_t b = size_of(p);
    free_.insert({b, p});
    idle += b;
  }
  void release_idle() {
    const auto t0 = std::chrono::steady_clock::now();
    for (auto& kv : free_) {
      (void)hipFree(kv.second);
      all.erase(kv.second);
      ++n_alloc;
    }
    free_.clear();
    idle = 0;
    t_alloc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  ~Pool() {
    for (auto& a : all) (void)hipFree(a.first);
  }
};

struct Buf {
  Pool* pool = nullptr;
  double* p = nullptr;
  int64_t n = 0;
  Buf() {}
  Buf(Pool& pl, int64_t nd, hipStream_t st, bool zero = true) : pool(&pl), n(std::max<int64_t>(nd, 1)) {
    p = (double*)pl.get(sizeof(double) * (size_t)n);
    if (zero) hchk(launch_fill(p, n, 0.0, st), "cheb: zero");
  }
  Buf(Buf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      reset();
      pool = o.pool;
      p = o.p;
      n = o.n;
      o.p = nullptr;
    }
    return *this;
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  void reset() {
    if (p && pool) pool->put(p);
    p = nullptr;
  }
  ~Buf() { reset(); }
};

// An MPS of nvec packed vectors: tables, device buffers, lambdas per bond.
struct DMps {
  std::vector<HTab> A;
  std::vector<Buf> d;
  std::vector<std::map<uint64_t, std::vector<double>>> lam;
  bool valid() const { return !A.empty(); }
};

HTab table_from(const std::map<Key, std::pair<int64_t, int64_t>>& m, int cplx_) {
  HTab t;
  t.ns = 1;
  t.cplx = cplx_;
  for (auto& kv : m) {
    t.q.push_back(kv.first[0]);
    t.q.push_back(kv.first[1]);
    t.p.push_back((uint32_t)kv.first[2]);
    t.rows.push_back(kv.second.first);
    t.cols.push_back(kv.second.second);
  }
  t.pack();
  return t;
}

// Plan-cache keys: a 128-bit digest of the operands' structure (keys, block
// dimensions, MPO entries) instead of the structure itself, which for a
// batch of vectors is megabytes per table and was copied and hashed on every
// call.  Two independent 64-bit lanes, each a multiply-rotate mix of every
// 8-byte word with a final avalanche (the murmur3 finaliser), plus the
// lengths: a collision between two of the engine's few thousand structures
// has probability of order 1e-31.
struct Digest {
  uint64_t a = 0x243f6a8885a308d3ull, b = 0x13198a2e03707344ull;
  static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
  void word(uint64_t w) {
    a = rotl(a ^ (w * 0x87c37b91114253d5ull), 31) * 0x4cf5ad432745937full + 0x52dce729ull;
    b = rotl(b + (w * 0x9e3779b97f4a7c15ull), 27) * 0xc2b2ae3d27d4eb4full ^ 0x38495ab5ull;
  }
  void bytes(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, c + i, 8);
      word(w);
    }
    uint64_t w = 0;
    if (i < n) std::memcpy(&w, c + i, n - i);
    word(w ^ ((uint64_t)n << 56));
  }
  static uint64_t fmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    return k ^ (k >> 33);
  }
  std::string str() const {
    const uint64_t x[2] = {fmix(a ^ rotl(b, 17)), fmix(b ^ rotl(a, 41))};
    return std::string((const char*)x, sizeof(x));
  }
};
std::string sig_tab(const HTab& t) {
  Digest d;
  d.bytes(t.q.data(), t.q.size() * 8);
  d.bytes(t.p.data(), t.p.size() * 4);
  d.bytes(t.rows.data(), t.rows.size() * 8);
  d.bytes(t.cols.data(), t.cols.size() * 8);
  d.word((uint64_t)t.n());
  d.word((uint64_t)t.cplx);
  return d.str();
}
std::string sig_mpo(const HMpo& w) {
  Digest d;
  d.word((uint64_t)w.Dl);
  d.word((uint64_t)w.Dr);
  d.bytes(w.wl.data(), w.wl.size() * 4);
  d.bytes(w.wr.data(), w.wr.size() * 4);
  d.bytes(w.out.data(), w.out.size() * 4);
  d.bytes(w.in.data(), w.in.size() * 4);
  d.bytes(w.val.data(), w.val.size() * 8);
  d.word((uint64_t)w.val.size());
  return d.str();
}

}  // namespace

struct mpsa_cheb {
  hipStream_t st = nullptr;
  int N = 0, nvec = 0, planes = 1, h1prec = MPSA_F64;
  std::vector<HMpo> W, Id;
  std::vector<std::vector<uint64_t>> itk;
  Pool pool;
  std::vector<std::shared_ptr<void>> keep;  // host copies of uploaded descriptors, until the next sync
  // plans by structure key (tables and MPO, not data): a moment loop meets the
  // same few hundred structures again and again once the bond dimensions
  // have settled.  Bounded by count and by workspace bytes; eviction waits
  // for the stream once and then drops the least recently used quarter.
  struct PlanEntry {
    std::string key;
    mpsa_plan* plan;
    int64_t bytes;
    HTab out;  // the plan's output table (the output-key query of its structure)
  };
  std::list<PlanEntry> plans;
  std::unordered_map<std::string, std::list<PlanEntry>::iterator> plan_index;
  // One compression sweep of 16 vectors at D = 1024 meets ~320 structures of
  // ~80 MB of workspace each, so the byte cap holds several sweeps' worth: a
  // cap at one sweep's size makes the LRU rebuild every plan of the next.
  static constexpr size_t kPlanCap = 8192;
  static constexpr int64_t kPlanBytesCap = int64_t(96) << 30;
  int64_t plan_bytes = 0;
  std::vector<DMps> slot;
  int64_t plans_built = 0;
  int64_t big_sector = MPSA_BIG_SECTOR;  // sectors with min(m, n) at least this go to rocSOLVER
  // sectors with min(m, n) at least this (and at most kSvdBlockMaxR) take the
  // blocked Jacobi SVD, several workgroups per sector; 0: never
  int64_t block_min_r = MPSA_BLOCK_JACOBI_MIN;
  // host time (seconds): plan builds, output-key queries, waits on the stream
  double t_plan = 0.0, t_keys = 0.0, t_sync = 0.0;
  double t_heff = 0.0, t_env = 0.0, t_prep = 0.0;  // host seconds in heff1, env_layer, prepare_site
  std::vector<double> moment_s;  // wall seconds of each moment of the last moment loop
  int64_t n_sync = 0;
  // sector shapes of every factorisation launch, while logging is on (up to
  // flog_cap records of kFlogWidth values: kind, sectors, max m, max n, max k,
  // sectors with k >= 64, >= 128, >= 256)
  static constexpr int kFlogWidth = 8;
  std::vector<int64_t> flog;
  int64_t flog_cap = 0;
  // Pinned staging ring for uploads: a copy from pinned memory is queued
  // without waiting, where a pageable source makes the runtime stage it
  // synchronously.  A region is rewritten only after a stream wait that
  // follows its copy (the ring wraps through sync()).
  static constexpr size_t kRing = size_t(64) << 20;
  char* ring = nullptr;
  size_t ring_at = 0;
  ~mpsa_cheb() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& e : plans) mpsa_plan_destroy(e.plan);
    slot.clear();
    if (ring) (void)hipHostFree(ring);
  }
  void sync() {
    const auto t0 = std::chrono::steady_clock::now();
    hchk(hipStreamSynchronize(st), "cheb: sync");
    t_sync += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++n_sync;
    keep.clear();
    ring_at = 0;
  }
  template <typename T>
  Buf upload(const std::vector<T>& v) {
    Buf b(pool, (int64_t)((sizeof(T) * std::max<size_t>(v.size(), 1) + 7) / 8), st, false);
    if (v.empty()) return b;
    const size_t bytes = sizeof(T) * v.size();
    if (!ring && hipHostMalloc((void**)&ring, kRing, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      ring = nullptr;
    }
    if (ring && bytes <= kRing / 4) {
      size_t at = (ring_at + 255) / 256 * 256;
      if (at + bytes > kRing) {
        sync();  // every earlier copy out of the ring has completed
        at = 0;
      }
      std::memcpy(ring + at, v.data(), bytes);
      hchk(hipMemcpyAsync(b.p, ring + at, bytes, hipMemcpyHostToDevice, st), "cheb: upload");
      ring_at = at + bytes;
      return b;
    }
    auto h = std::make_shared<std::vector<T>>(v);
    hchk(hipMemcpyAsync(b.p, h->data(), bytes, hipMemcpyHostToDevice, st), "cheb: upload");
    keep.push_back(h);
    return b;
  }
  const PlanEntry* find_plan(const std::string& k) {
    auto f = plan_index.find(k);
    if (f == plan_index.end()) return nullptr;
    plans.splice(plans.begin(), plans, f->second);
    return &*f->second;
  }
  void put_plan(const std::string& k, mpsa_plan* p, const HTab& out) {
    int64_t ws = 0;
    (void)mpsa_plan_info(p, nullptr, nullptr, nullptr, nullptr, &ws);
    plans.push_front(PlanEntry{k, p, ws, out});
    plan_index[k] = plans.begin();
    plan_bytes += ws;
    ++plans_built;
    if (plans.size() > kPlanCap || plan_bytes > kPlanBytesCap) {
      sync();  // the evicted plans' last launches must have finished
      while (plans.size() > 1 && (plans.size() > kPlanCap * 3 / 4 || plan_bytes > kPlanBytesCap * 3 / 4)) {
        plan_bytes -= plans.back().bytes;
        plan_index.erase(plans.back().key);
        mpsa_plan_destroy(plans.back().plan);
        plans.pop_back();
      }
    }
  }
};

namespace {

typedef mpsa_cheb Eng;

// adds the wall time of its scope to an accumulator
struct ScopeTimer {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit ScopeTimer(double& a) : acc(a) {}
  ~ScopeTimer() { acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// ------------------------------------------------------------ copies, GEMMs
void add_copy(std::vector<Copy2D>& v, const Eng& E, int64_t src_off, int64_t src_plane, int64_t dst_off,
              int64_t dst_plane, int64_t rows, int64_t cols, int64_t lds, int64_t ldd, double scale = 1.0,
              int64_t rs = -1, int64_t cs = -1, int trans = 0) {
  if (rows <= 0 || cols <= 0) return;
  for (int pl = 0; pl < E.planes; ++pl) {
    Copy2D t{};
    t.src_off = src_off + pl * src_plane;
    t.dst_off = dst_off + pl * dst_plane;
    t.rows = (int32_t)rows;
    t.cols = (int32_t)cols;
    t.lds = (int32_t)std::max<int64_t>(lds, 1);
    t.ldd = (int32_t)std::max<int64_t>(ldd, 1);
    t.rs_off = rs;
    t.cs_off = cs;
    t.scale = scale;
    t.trans = trans;
    v.push_back(t);
  }
}

void run_copies(Eng& E, const std::vector<Copy2D>& v, const double* src, double* dst, const double* scales) {
  if (v.empty()) return;
  int64_t mx = 0;
  for (auto& t : v) mx = std::max<int64_t>(mx, (int64_t)t.rows * t.cols);
  Buf d = E.upload(v);
  const int64_t kMaxY = 65535;  // one task per grid row of copy2d_kernel
  for (int64_t first = 0; first < (int64_t)v.size(); first += kMaxY) {
    const int cnt = (int)std::min<int64_t>(kMaxY, (int64_t)v.size() - first);
    hchk(launch_copy2d((const Copy2D*)d.p + first, cnt, mx, src, dst, scales, E.st), "cheb: copy");
  }
}

// C (m x n, ld m) = A (m x k, ld lda) * B (k x n, ld ldb), planar complex when
// the engine is complex: Re C = Ar Br - Ai Bi, Im C = Ar Bi + Ai Br.
struct GemmList {
  std::vector<GTask> tasks;
  std::vector<GSeg> segs;
  void add(const Eng& E, int64_t c_off, int64_t c_plane, int64_t m, int64_t n, int64_t k, int64_t a_off,
           int64_t a_plane, int64_t lda, int64_t b_off, int64_t b_plane, int64_t ldb) {
    if (m <= 0 || n <= 0) return;
    auto seg = [&](int64_t ao, int64_t bo, double al) {
      GSeg g{};
      g.a_off = ao;
      g.b_off = bo;
      g.k = (int32_t)k;
      g.lda = (int32_t)std::max<int64_t>(lda, 1);
      g.ldb = (int32_t)std::max<int64_t>(ldb, 1);
      g.alpha = al;
      segs.push_back(g);
    };
    for (int pl = 0; pl < E.planes; ++pl) {
      GTask t{};
      t.c_off = c_off + pl * c_plane;
      t.m = (int32_t)m;
      t.n = (int32_t)n;
      t.seg0 = (int32_t)segs.size();
      if (E.planes == 1) {
        seg(a_off, b_off, 1.0);
      } else if (pl == 0) {
        seg(a_off, b_off, 1.0);
        seg(a_off + a_plane, b_off + b_plane, -1.0);
      } else {
        seg(a_off, b_off + b_plane, 1.0);
        seg(a_off + a_plane, b_off, 1.0);
      }
      t.nseg = (int32_t)segs.size() - t.seg0;
      tasks.push_back(t);
    }
  }
  void run(Eng& E, const double* A, const double* B, double* C) {
    if (tasks.empty()) return;
    int tiles = 0, cls = 0;
    finalize_gemm_list(tasks, segs, F64, &tiles, &cls);
    if (tasks.empty() || tiles <= 0) return;
    Buf dt = E.upload(tasks), ds = E.upload(segs);
    hchk(launch_gemm(F64, 0, 0, cls, (const GTask*)dt.p, (int)tasks.size(), (const GSeg*)ds.p, tiles, 1, A, 0, B, 0,
                     (void*)C, 0, E.st),
         "cheb: grouped GEMM");
  }
};

// ------------------------------------------------------------ basic tensors
DMps new_mps(Eng& E) {
  DMps M;
  M.A.resize((size_t)E.N);
  M.d.resize((size_t)E.N);
  M.lam.resize((size_t)std::max(E.N - 1, 0));
  return M;
}

DMps copy_mps(Eng& E, const DMps& a) {
  DMps M = new_mps(E);
  for (int s = 0; s < E.N; ++s) {
    M.A[(size_t)s] = a.A[(size_t)s];
    M.d[(size_t)s] = Buf(E.pool, a.A[(size_t)s].len(), E.st, false);
    if (a.A[(size_t)s].len() > 0)
      hchk(hipMemcpyAsync(M.d[(size_t)s].p, a.d[(size_t)s].p, sizeof(double) * (size_t)a.A[(size_t)s].len(),
                          hipMemcpyDeviceToDevice, E.st),
           "cheb: copy");
  }
  M.lam = a.lam;
  return M;
}

Key key_of(const HTab& t, int64_t i) { return Key{t.q[2 * i], t.q[2 * i + 1], (uint64_t)t.p[(size_t)i]}; }

// getLMin per vector: a 1 x 1 identity at each vector's left boundary charge
HTab lmin_tab_of(const DMps& M, int cplx_) {
  std::map<Key, std::pair<int64_t, int64_t>> m;
  for (int64_t i = 0; i < M.A[0].n(); ++i) m[Key{M.A[0].q[2 * i], M.A[0].q[2 * i], 0}] = {1, 1};
  return table_from(m, cplx_);
}
// getRMax: identity per right charge of the last site
HTab rmax_tab_of(const DMps& M, int cplx_) {
  const HTab& a = M.A.back();
  std::map<Key, std::pair<int64_t, int64_t>> m;
  for (int64_t i = 0; i < a.n(); ++i) m[Key{a.q[2 * i + 1], a.q[2 * i + 1], 0}] = {a.cols[(size_t)i], a.cols[(size_t)i]};
  return table_from(m, cplx_);
}
Buf identity_buf(Eng& E, const HTab& t) {
  Buf b(E.pool, t.len(), E.st);
  for (int64_t k = 0; k < t.n(); ++k)
    hchk(launch_fill_strided(b.p + t.off[(size_t)k], t.rows[(size_t)k], t.rows[(size_t)k] + 1, 1.0, E.st),
         "cheb: identity");
  return b;
}

// ------------------------------------------------------------ contractions
// E' = addContractionLayer(E, A, W, dir) with the upper (conjugated) tensor Ap
void env_layer(Eng& E, const HTab& Et, const double* dE, const HTab& At, const double* dA, const HTab& Apt,
               const double* dAp, const HMpo& W, const std::vector<uint64_t>& itk, int dir, HTab& out_t, Buf& out_b) {
  ScopeTimer tt(E.t_env);
  mpsa_blocks be = Et.view(), ba = At.view(), bp = Apt.view();
  mpsa_mpo mw = W.view();
  // a plan and its output table are fixed by the structure of the operands
  const std::string key = "E" + std::to_string(dir > 0) + sig_tab(Et) + sig_tab(At) + sig_tab(Apt) + sig_mpo(W);
  const auto* pe = E.find_plan(key);
  mpsa_plan* pl = pe ? pe->plan : nullptr;
  if (pe) {
    out_t = pe->out;
  } else {
    // keys and plan in one pass; the output table takes this engine's layout
    const int cplx_ = Et.cplx;
    auto make_out = [&](int64_t n, const uint64_t* q, const uint32_t* p, const int64_t* r, const int64_t* c) {
      std::vector<uint64_t> vq(q, q + 2 * n);
      std::vector<uint32_t> vp(p, p + n);
      std::vector<int64_t> vr(r, r + n), vc(c, c + n);
      out_t = from_query(n, 1, vq, vp, vr, vc);
      out_t.cplx = cplx_;
      return out_t.view();
    };
    {
      ScopeTimer tp(E.t_plan);
      chk(env_keys_plan(&pl, &be, &ba, &bp, &mw, itk.data(), (int)itk.size(), dir, make_out));
    }
    if (pl) E.put_plan(key, pl, out_t);
  }
  out_b = Buf(E.pool, out_t.len(), E.st);
  if (out_t.n() && pl) chk(mpsa_env_apply(pl, dE, dA, dAp, out_b.p, (mpsa_stream)E.st));
}

// contractSingle(L, A, W, R): the one-site H_eff applied to A; the output table
// is the produced key set in key order
void heff1(Eng& E, const HTab& Lt, const double* dL, const HTab& At, const double* dA, const HTab& Rt,
           const double* dR, const HMpo& W, const std::vector<uint64_t>& itk, HTab& out_t, Buf& out_b) {
  ScopeTimer tt(E.t_heff);
  mpsa_blocks bl = Lt.view(), br = Rt.view(), ba = At.view();
  mpsa_mpo mw = W.view();
  const std::string key = "H" + std::to_string(E.h1prec) + sig_tab(Lt) + sig_tab(Rt) + sig_tab(At) + sig_mpo(W);
  const auto* pe = E.find_plan(key);
  mpsa_plan* pl = pe ? pe->plan : nullptr;
  if (pe) {
    out_t = pe->out;
  } else {
    // keys and plan in one pass; the output table takes this engine's layout
    const int cplx_ = At.cplx;
    auto make_out = [&](int64_t n, const uint64_t* q, const uint32_t* p, const int64_t* r, const int64_t* c) {
      std::vector<uint64_t> vq(q, q + 2 * n);
      std::vector<uint32_t> vp(p, p + n);
      std::vector<int64_t> vr(r, r + n), vc(c, c + n);
      out_t = from_query(n, 1, vq, vp, vr, vc);
      out_t.cplx = cplx_;
      return out_t.view();
    };
    {
      ScopeTimer tp(E.t_plan);
      chk(heff1_keys_plan(&pl, &bl, &br, &ba, &mw, itk.data(), (int)itk.size(), E.h1prec, make_out));
    }
    if (pl) E.put_plan(key, pl, out_t);
  }
  out_b = Buf(E.pool, out_t.len(), E.st);
  if (!out_t.n() || !pl) return;
  chk(mpsa_plan_bind_env(pl, dL, dR, (mpsa_stream)E.st));
  chk(mpsa_heff_apply(pl, dA, 0, out_b.p, 0, 1, (mpsa_stream)E.st));
}

// the 1 x 1 blocks of a closed left environment, one per vector
std::vector<cplx> env_scalars(Eng& E, const HTab& Lt, const Buf& Lb) {
  const int cp = E.planes == 2;
  std::vector<double> h((size_t)std::max<int64_t>(Lt.len(), 1));
  if (Lt.len() > 0)
    hchk(hipMemcpyAsync(h.data(), Lb.p, sizeof(double) * (size_t)Lt.len(), hipMemcpyDeviceToHost, E.st),
         "cheb: overlap");
  E.sync();
  std::vector<cplx> out((size_t)E.nvec, cplx(0.0, 0.0));
  for (int64_t i = 0; i < Lt.n(); ++i) {
    const uint64_t v = ns_of(Lt.q[2 * i]);
    if (v >= (uint64_t)E.nvec || Lt.rows[(size_t)i] < 1 || Lt.cols[(size_t)i] < 1) continue;
    const int64_t o = Lt.off[(size_t)i];
    out[(size_t)v] = cplx(h[(size_t)o], cp ? h[(size_t)(o + Lt.total)] : 0.0);
  }
  return out;
}

// computeOverlap(a, b) per vector: <b|a> (the upper tensor b is conjugated)
std::vector<cplx> overlap(Eng& E, const DMps& a, const DMps& b) {
  const int cp = E.planes == 2;
  HTab Lt = lmin_tab_of(a, cp);
  Buf Lb = identity_buf(E, Lt);
  for (int s = 0; s < E.N; ++s) {
    HTab nt;
    Buf nb;
    env_layer(E, Lt, Lb.p, a.A[(size_t)s], a.d[(size_t)s].p, b.A[(size_t)s], b.d[(size_t)s].p, E.Id[(size_t)s],
              E.itk[(size_t)s], 1, nt, nb);
    Lt = nt;
    Lb = std::move(nb);
  }
  return env_scalars(E, Lt, Lb);
}

// ------------------------------------------------------------ large sectors
// Sectors with k = min(m, n) >= the engine's big_sector (default
// MPSA_BIG_SECTOR = 1024, the batched kernels' limit) are factorised by
// rocSOLVER, one call each, on the engine stream; the others go to the batched
// launch.  A lower threshold hands more sectors to rocSOLVER, which pays only
// for sectors far larger than the Hubbard D = 1024 workload's (k < 512 there):
// at k >= 96 or >= 192 the serial rocSOLVER calls cost more than the batched
// launch they shorten.
static_assert(MPSA_BIG_SECTOR >= 1 && MPSA_BIG_SECTOR <= kSvdMaxR, "sectors below the threshold go to the batched SVD");

void rschk(rocblas_status s, const char* what) {
  if (s != rocblas_status_success) throw HErr(MPSA_EHIP, std::string("cheb: ") + what + ": rocBLAS/rocSOLVER status " +
                                                            std::to_string((int)s));
}

// One sector M (m x n, ld m, planar) into the batched factorisation's output
// format: SVD (qr = false): U (m x k), S, V^dagger (k x n), FixPhase on real
// data; QR (left) U = Q, V^dagger = R, or LQ (right) U = L, V^dagger = Q, S = 1.
// M is destroyed.  info: one int of device memory for rocSOLVER's status.
void big_factor(Eng& E, bool qr, bool left, int64_t m, int64_t n, double* M, int64_t m_plane, double* U,
                int64_t u_plane, double* Vt, int64_t vt_plane, double* Sv, rocblas_int* info) {
  const int64_t k = std::min(m, n);
  const bool c = E.planes == 2;
  rocblas_handle h = (rocblas_handle)rocblas_handle_current();
  rschk(rocblas_set_stream(h, E.st), "set stream");
  const rocblas_int mi = (rocblas_int)m, ni = (rocblas_int)n, ki = (rocblas_int)k;
  Buf Az, tau(E.pool, 2 * k, E.st, false), ework(E.pool, std::max<int64_t>(k, 1), E.st, false);
  double* A = M;
  if (c) {
    Az = Buf(E.pool, 2 * m * n, E.st, false);
    hchk(launch_pack_z(M, M + m_plane, m * n, Az.p, E.st), "cheb: pack sector");
    A = Az.p;
  }
  auto* Az_ = reinterpret_cast<rocblas_double_complex*>(A);
  auto* tz = reinterpret_cast<rocblas_double_complex*>(tau.p);
  if (!qr) {
    if (!c) {
      rschk(rocsolver_dgesvd(h, rocblas_svect_singular, rocblas_svect_singular, mi, ni, A, mi, Sv, U, mi, Vt, ki,
                             ework.p, rocblas_outofplace, info),
            "dgesvd");
      hchk(launch_fixphase(U, m, m, Vt, k, n, k, E.st), "cheb: fixphase");
    } else {
      Buf uz(E.pool, 2 * m * k, E.st, false), vz(E.pool, 2 * k * n, E.st, false);
      rschk(rocsolver_zgesvd(h, rocblas_svect_singular, rocblas_svect_singular, mi, ni, Az_, mi, Sv,
                             reinterpret_cast<rocblas_double_complex*>(uz.p), mi,
                             reinterpret_cast<rocblas_double_complex*>(vz.p), ki, ework.p, rocblas_outofplace, info),
            "zgesvd");
      hchk(launch_unpack_mat(uz.p, 1, m, m, k, 0, U, m, u_plane, E.st), "cheb: unpack U");
      hchk(launch_unpack_mat(vz.p, 1, k, k, n, 0, Vt, k, vt_plane, E.st), "cheb: unpack Vt");
    }
    return;
  }
  if (left) {  // M = Q R
    if (c) rschk(rocsolver_zgeqrf(h, mi, ni, Az_, mi, tz), "zgeqrf");
    else rschk(rocsolver_dgeqrf(h, mi, ni, A, mi, tau.p), "dgeqrf");
    hchk(launch_unpack_mat(A, c, m, k, n, 1, Vt, k, vt_plane, E.st), "cheb: unpack R");
    if (c) rschk(rocsolver_zungqr(h, mi, ki, ki, Az_, mi, tz), "zungqr");
    else rschk(rocsolver_dorgqr(h, mi, ki, ki, A, mi, tau.p), "dorgqr");
    hchk(launch_unpack_mat(A, c, m, m, k, 0, U, m, u_plane, E.st), "cheb: unpack Q");
  } else {  // M = L Q
    if (c) rschk(rocsolver_zgelqf(h, mi, ni, Az_, mi, tz), "zgelqf");
    else rschk(rocsolver_dgelqf(h, mi, ni, A, mi, tau.p), "dgelqf");
    hchk(launch_unpack_mat(A, c, m, m, k, 2, U, m, u_plane, E.st), "cheb: unpack L");
    if (c) rschk(rocsolver_zunglq(h, ki, ni, ki, Az_, mi, tz), "zunglq");
    else rschk(rocsolver_dorglq(h, ki, ni, ki, A, mi, tau.p), "dorglq");
    hchk(launch_unpack_mat(A, c, m, k, n, 0, Vt, k, vt_plane, E.st), "cheb: unpack Q'");
  }
  hchk(launch_fill(Sv, k, 1.0, E.st), "cheb: unit S");
}

// Chunk table of launch_house_panels / launch_house_apply: sector s has its
// workgroups at base[s] .. base[s + 1]; returns the largest min(m, n).
int house_table(const std::vector<SvdTask>& ts, int layout, std::vector<int>& base) {
  base.assign(ts.size() + 1, 0);
  int kmax = 0;
  for (size_t i = 0; i < ts.size(); ++i) {
    base[i + 1] = base[i] + house_chunks(ts[i].m, ts[i].n, layout);
    kmax = std::max(kmax, (int)std::min(ts[i].m, ts[i].n));
  }
  return kmax;
}

// The SVD of the large sectors in phases (mpsa_factor.hip): the load, the
// reflections in panels over several workgroups per sector, the init, sweeps
// of block-tournament rounds until a sweep rotates nothing (one host wait per
// sweep), the back-transform over several workgroups, the outputs.
void svd_blocked(Eng& E, bool cplx, const std::vector<SvdTask>& big, const double* M, int64_t m_plane, double* U,
                 int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W, int64_t w_plane) {
  const int nb = (int)big.size();
  std::vector<int> hbase;
  const int kmax = house_table(big, 0, hbase);
  Buf dt = E.upload(big), dh = E.upload(hbase);
  Buf flags(E.pool, nb, E.st);
  const SvdTask* t = (const SvdTask*)dt.p;
  auto phase = [&](int ph, const char* what) {
    hchk(launch_svd_phase(cplx, ph, t, nb, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane, E.st), what);
  };
  phase(kPhLoad, "cheb: sector SVD load");
  hchk(launch_house_panels(cplx, 0, t, nb, (const int*)dh.p, hbase.back(), kmax, W, w_plane, E.st),
       "cheb: sector SVD reflections");
  phase(kPhInit, "cheb: sector SVD init");
  // sweeps over the sectors still rotating: a sweep that rotates nothing in a
  // sector leaves it as it is, so later sweeps would not rotate it either
  std::vector<SvdTask> act = big;
  std::vector<double> hf((size_t)nb);
  for (int sweep = 0; sweep < 60 && !act.empty(); ++sweep) {
    const int na = (int)act.size();
    std::vector<int> base((size_t)na + 1, 0);
    int rounds = 0, rmax = 1;
    for (int i = 0; i < na; ++i) {
      const SvdTask& a = act[(size_t)i];
      base[(size_t)i + 1] = base[(size_t)i] + svd_block_pairs(a.m, a.n);
      rounds = std::max(rounds, svd_block_rounds(a.m, a.n));
      rmax = std::max(rmax, (int)std::min(a.m, a.n));
    }
    Buf da = E.upload(act), db = E.upload(base);
    for (int round = -1; round < rounds; ++round)
      hchk(launch_svd_block_round(cplx, (const SvdTask*)da.p, na, (const int*)db.p, base.back(), round, W, w_plane,
                                  flags.p, rmax, E.st),
           "cheb: sector SVD round");
    hchk(hipMemcpyAsync(hf.data(), flags.p, sizeof(double) * (size_t)na, hipMemcpyDeviceToHost, E.st),
         "cheb: sector SVD flags");
    E.sync();
    std::vector<SvdTask> next;
    for (int i = 0; i < na; ++i)
      if (hf[(size_t)i] != 0.0) next.push_back(act[(size_t)i]);
    act.swap(next);
    if (!act.empty()) hchk(launch_fill(flags.p, na, 0.0, E.st), "cheb: sector SVD flags");
  }
  phase(kPhBackInit, "cheb: sector SVD back-transform");
  hchk(launch_house_apply(cplx, 0, t, nb, (const int*)dh.p, hbase.back(), W, w_plane, E.st),
       "cheb: sector SVD back-transform");
  phase(kPhOut, "cheb: sector SVD outputs");
}

// The QR (right = 0) or LQ (right = 1) of the large sectors the same way:
// load, reflections in panels, R and Q := [I; 0], Q formed over several
// workgroups per sector, outputs.
void qr_blocked(Eng& E, bool cplx, int right, const std::vector<SvdTask>& big, const double* M, int64_t m_plane,
                double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W, int64_t w_plane) {
  const int nb = (int)big.size();
  std::vector<int> hbase;
  const int kmax = house_table(big, 1 + right, hbase);
  Buf dt = E.upload(big), dh = E.upload(hbase);
  const SvdTask* t = (const SvdTask*)dt.p;
  auto phase = [&](int ph, const char* what) {
    hchk(launch_qr_phase(cplx, right, ph, t, nb, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane, E.st), what);
  };
  phase(kPhLoad, "cheb: sector QR load");
  hchk(launch_house_panels(cplx, 1 + right, t, nb, (const int*)dh.p, hbase.back(), kmax, W, w_plane, E.st),
       "cheb: sector QR reflections");
  phase(kPhInit, "cheb: sector QR init");
  hchk(launch_house_apply(cplx, 1 + right, t, nb, (const int*)dh.p, hbase.back(), W, w_plane, E.st),
       "cheb: sector QR formation");
  phase(kPhOut, "cheb: sector QR outputs");
}

// ------------------------------------------------------------ prepareSite
// prepareSite(mps, site, "l" or "r", multiply = true) for every sector of the
// site at once; lambdas kept on the host when asked (for a later truncate).
// The sector factorisation is the SVD (U, S V^dagger) when the Schmidt values
// are kept, else a Householder QR (Q, R) or LQ: the site is the same
// isometry up to a unitary of the k kept directions and the neighbour takes
// the rest, so the state and everything computed from it (overlaps, norms,
// the compression's optimum, later Schmidt values) are the reference's; only
// the gauge of the two tensors differs.
void prepare_site(Eng& E, DMps& M, int site, bool left, bool keep_lambda) {
  ScopeTimer tt(E.t_prep);
  const bool use_qr = !keep_lambda;
  const int sa = left ? site : site - 1, sb = left ? site + 1 : site;  // (A, B) = (site, site+1) or (site-1, site)
  const HTab &At = M.A[(size_t)sa], &Bt = M.A[(size_t)sb];
  const HTab& X = left ? At : Bt;  // the tensor factorised
  struct Sec {
    std::vector<int64_t> blk;  // blocks of X in key order
    std::vector<int64_t> off;  // their row (l) or column (r) offsets in the sector matrix
    int64_t m = 0, n = 0, k = 0, mo = 0, uo = 0, vo = 0, so = 0, wo = 0;
    bool big = false;  // factorised by rocSOLVER
  };
  std::map<uint64_t, Sec> secs;
  for (int64_t i = 0; i < X.n(); ++i) {
    const uint64_t q = left ? X.q[2 * i + 1] : X.q[2 * i];
    Sec& S = secs[q];
    S.blk.push_back(i);
  }
  int64_t tm = 0, tu = 0, tv = 0, ts = 0, tw = 0;
  for (auto& kv : secs) {
    Sec& S = kv.second;
    for (int64_t i : S.blk) {
      if (left) {
        S.off.push_back(S.m);
        S.m += X.rows[(size_t)i];
        S.n = X.cols[(size_t)i];
      } else {
        S.off.push_back(S.n);
        S.n += X.cols[(size_t)i];
        S.m = X.rows[(size_t)i];
      }
    }
    S.k = std::min(S.m, S.n);
    S.big = S.k >= E.big_sector;
    S.mo = tm;
    tm += align64(S.m * S.n);
    S.uo = tu;
    tu += align64(S.m * S.k);
    S.vo = tv;
    tv += align64(S.k * S.n);
    S.so = ts;
    ts += S.k;
    S.wo = tw;
    const int64_t qp = left ? S.m : S.n;  // rows of the factorised X = M or M^H
    if (!S.big) tw += use_qr ? align64(S.m * S.n + qp * S.k + S.k) : align64(svd_workspace(S.m, S.n));
  }
  const int pl = E.planes;
  Buf Mb(E.pool, pl * tm, E.st), Ub(E.pool, pl * tu, E.st), Vb(E.pool, pl * tv, E.st), Sb(E.pool, ts, E.st),
      Wb(E.pool, pl * tw, E.st, false);
  // gather
  std::vector<Copy2D> g;
  const int64_t xoff_plane = X.total;
  const double* dX = left ? M.d[(size_t)sa].p : M.d[(size_t)sb].p;
  for (auto& kv : secs) {
    Sec& S = kv.second;
    for (size_t j = 0; j < S.blk.size(); ++j) {
      const int64_t i = S.blk[j];
      const int64_t dst = left ? S.mo + S.off[j] : S.mo + S.off[j] * S.m;
      add_copy(g, E, X.off[(size_t)i], xoff_plane, dst, tm, X.rows[(size_t)i], X.cols[(size_t)i], X.rows[(size_t)i],
               S.m);
    }
  }
  run_copies(E, g, dX, Mb.p, nullptr);
  std::vector<SvdTask> tasks;
  int nbig = 0;
  for (auto& kv : secs) {
    const Sec& S = kv.second;
    if (S.big) {
      ++nbig;
      continue;
    }
    SvdTask t{};
    t.a_off = S.mo;
    t.u_off = S.uo;
    t.vt_off = S.vo;
    t.s_off = S.so;
    t.w_off = S.wo;
    t.m = (int32_t)S.m;
    t.n = (int32_t)S.n;
    t.lda = (int32_t)std::max<int64_t>(S.m, 1);
    tasks.push_back(t);
  }
  // largest sectors first: a launch lasts as long as its largest sector, so
  // the big ones must not wait for a free CU behind the small ones
  std::stable_sort(tasks.begin(), tasks.end(), [](const SvdTask& a, const SvdTask& b) {
    const int64_t ka = std::min(a.m, a.n), kb = std::min(b.m, b.n);
    return (int64_t)a.m * a.n * ka > (int64_t)b.m * b.n * kb;
  });
  if (E.flog_cap > 0 && (int64_t)E.flog.size() < E.flog_cap * Eng::kFlogWidth) {
    // kind: 0 SVD, 1 QR/LQ in the batched kernels; 2, 3 the same on rocSOLVER
    for (int big = 0; big < 2; ++big) {
      int64_t rec[Eng::kFlogWidth] = {(use_qr ? 1 : 0) + 2 * big, 0, 0, 0, 0, 0, 0, 0};
      for (auto& kv : secs) {
        const Sec& S = kv.second;
        if (S.big != (big == 1) || S.k < 1) continue;
        ++rec[1];
        rec[2] = std::max(rec[2], S.m);
        rec[3] = std::max(rec[3], S.n);
        rec[4] = std::max(rec[4], S.k);
        rec[5] += S.k >= 64;
        rec[6] += S.k >= 128;
        rec[7] += S.k >= 256;
      }
      if (rec[1] && (int64_t)E.flog.size() < E.flog_cap * Eng::kFlogWidth) E.flog.insert(E.flog.end(), rec, rec + 8);
    }
  }
  std::vector<SvdTask> blocked;  // large sectors: several workgroups each
  if (E.block_min_r > 0) {
    std::vector<SvdTask> rest;
    for (const SvdTask& t : tasks) {
      const int64_t r = std::min(t.m, t.n), p = std::max(t.m, t.n);
      const bool ok = r >= E.block_min_r && p <= kVecRowsMax && (use_qr || r <= kSvdBlockMaxR);
      (ok ? blocked : rest).push_back(t);
    }
    tasks.swap(rest);
  }
  if (!blocked.empty()) {
    if (use_qr)
      qr_blocked(E, pl == 2, left ? 0 : 1, blocked, Mb.p, tm, Ub.p, tu, Vb.p, tv, Sb.p, Wb.p, tw);
    else
      svd_blocked(E, pl == 2, blocked, Mb.p, tm, Ub.p, tu, Vb.p, tv, Sb.p, Wb.p, tw);
  }
  {
    Buf dt = E.upload(tasks);
    if (use_qr)
      hchk(launch_qr_batched(pl == 2, left ? 0 : 1, (const SvdTask*)dt.p, (int)tasks.size(), Mb.p, tm, Ub.p, tu, Vb.p,
                             tv, Sb.p, Wb.p, tw, E.st),
           "cheb: sector QR");
    else
      hchk(launch_svd_batched(pl == 2, (const SvdTask*)dt.p, (int)tasks.size(), Mb.p, tm, Ub.p, tu, Vb.p, tv, Sb.p,
                              Wb.p, tw, E.st),
           "cheb: sector SVD");
  }
  if (nbig) {
    Buf dinfo(E.pool, nbig, E.st);  // one rocblas_int per sector (zero-filled doubles)
    rocblas_int* info = reinterpret_cast<rocblas_int*>(dinfo.p);
    int b = 0;
    for (auto& kv : secs) {
      const Sec& S = kv.second;
      if (!S.big) continue;
      big_factor(E, use_qr, left, S.m, S.n, Mb.p + S.mo, tm, Ub.p + S.uo, tu, Vb.p + S.vo, tv, Sb.p + S.so,
                 info + 2 * b);
      ++b;
    }
    if (!use_qr) {  // gesvd reports non-convergence in info
      std::vector<rocblas_int> hi((size_t)2 * nbig);
      hchk(hipMemcpyAsync(hi.data(), info, sizeof(rocblas_int) * hi.size(), hipMemcpyDeviceToHost, E.st),
           "cheb: sector info");
      E.sync();
      for (int i = 0; i < nbig; ++i)
        if (hi[(size_t)2 * i] != 0) throw HErr(MPSA_EINVAL, "cheb: rocSOLVER gesvd did not converge on a sector");
    }
  }
  // new tables: the factorised tensor keeps U (l) or V^dagger (r) with k
  // columns (rows); the neighbour's blocks of the sector charge take k rows
  // (columns)
  const HTab& Y = left ? Bt : At;  // the neighbour
  auto x_dims = [&](int64_t i) {
    const Sec& S = secs[left ? X.q[2 * i + 1] : X.q[2 * i]];
    return left ? std::make_pair(X.rows[(size_t)i], S.k) : std::make_pair(S.k, X.cols[(size_t)i]);
  };
  auto y_dims = [&](int64_t i) {
    auto it = secs.find(left ? Y.q[2 * i] : Y.q[2 * i + 1]);
    if (it == secs.end()) return std::make_pair(Y.rows[(size_t)i], Y.cols[(size_t)i]);
    return left ? std::make_pair(it->second.k, Y.cols[(size_t)i]) : std::make_pair(Y.rows[(size_t)i], it->second.k);
  };
  // The new tables have the keys of X and Y.  Tables in key order (the
  // engine's own) give them in the same order, so block i of X is block i of
  // NX and block t of NY is block t of Y; otherwise go through key maps.
  auto in_order = [](const HTab& t) {
    for (int64_t i = 1; i < t.n(); ++i)
      if (!(key_of(t, i - 1) < key_of(t, i))) return false;
    return true;
  };
  const bool ordered = in_order(X) && in_order(Y);
  HTab NX, NY;
  std::map<Key, int64_t> nx_index, y_index;
  if (ordered) {
    auto rebuild = [&](const HTab& src, auto dims_of) {
      HTab t;
      t.ns = 1;
      t.cplx = pl == 2;
      t.q = src.q;
      t.p = src.p;
      t.rows.resize((size_t)src.n());
      t.cols.resize((size_t)src.n());
      for (int64_t i = 0; i < src.n(); ++i) {
        const auto rc = dims_of(i);
        t.rows[(size_t)i] = rc.first;
        t.cols[(size_t)i] = rc.second;
      }
      t.pack();
      return t;
    };
    NX = rebuild(X, x_dims);
    NY = rebuild(Y, y_dims);
  } else {
    std::map<Key, std::pair<int64_t, int64_t>> nx, ny;
    for (int64_t i = 0; i < X.n(); ++i) nx[key_of(X, i)] = x_dims(i);
    for (int64_t i = 0; i < Y.n(); ++i) ny[key_of(Y, i)] = y_dims(i);
    NX = table_from(nx, pl == 2);
    NY = table_from(ny, pl == 2);
    for (int64_t t = 0; t < NX.n(); ++t) nx_index[key_of(NX, t)] = t;
    for (int64_t t = 0; t < Y.n(); ++t) y_index[key_of(Y, t)] = t;
  }
  Buf dNX(E.pool, NX.len(), E.st), dNY(E.pool, NY.len(), E.st);
  // factorised tensor from U / V^dagger
  std::vector<Copy2D> c1;
  for (auto& kv : secs) {
    const Sec& S = kv.second;
    for (size_t j = 0; j < S.blk.size(); ++j) {
      const int64_t i = S.blk[j];
      const int64_t b = ordered ? i : nx_index.at(key_of(X, i));
      if (left)
        add_copy(c1, E, S.uo + S.off[j], tu, NX.off[(size_t)b], NX.total, X.rows[(size_t)i], S.k, S.m,
                 X.rows[(size_t)i]);
      else
        add_copy(c1, E, S.vo + S.off[j] * S.k, tv, NX.off[(size_t)b], NX.total, S.k, X.cols[(size_t)i], S.k, S.k);
    }
  }
  run_copies(E, c1, left ? Ub.p : Vb.p, dNX.p, nullptr);
  // S V^dagger (l) or U S (r), then the neighbour's product
  Buf F(E.pool, left ? pl * tv : pl * tu, E.st);
  std::vector<Copy2D> c2;
  for (auto& kv : secs) {
    const Sec& S = kv.second;
    if (left) add_copy(c2, E, S.vo, tv, S.vo, tv, S.k, S.n, S.k, S.k, 1.0, S.so, -1);
    else add_copy(c2, E, S.uo, tu, S.uo, tu, S.m, S.k, S.m, S.m, 1.0, -1, S.so);
  }
  run_copies(E, c2, left ? Vb.p : Ub.p, F.p, Sb.p);
  GemmList gl;
  std::vector<Copy2D> c3;  // neighbour blocks outside every sector are copied
  const double* dY = left ? M.d[(size_t)sb].p : M.d[(size_t)sa].p;
  for (int64_t t = 0; t < NY.n(); ++t) {
    const Key kk = key_of(NY, t);
    const int64_t i = ordered ? t : y_index.at(kk);
    const uint64_t q = left ? kk[0] : kk[1];
    auto it = secs.find(q);
    if (it == secs.end()) {
      add_copy(c3, E, Y.off[(size_t)i], Y.total, NY.off[(size_t)t], NY.total, Y.rows[(size_t)i], Y.cols[(size_t)i],
               Y.rows[(size_t)i], Y.rows[(size_t)i]);
      continue;
    }
    const Sec& S = it->second;
    if (left)  // (k x n) (n x cols)
      gl.add(E, NY.off[(size_t)t], NY.total, S.k, Y.cols[(size_t)i], S.n, S.vo, tv, S.k, Y.off[(size_t)i], Y.total,
             Y.rows[(size_t)i]);
    else  // (rows x m) (m x k)
      gl.add(E, NY.off[(size_t)t], NY.total, Y.rows[(size_t)i], S.k, S.m, Y.off[(size_t)i], Y.total,
             Y.rows[(size_t)i], S.uo, tu, S.m);
  }
  if (left) gl.run(E, F.p, dY, dNY.p);
  else gl.run(E, dY, F.p, dNY.p);
  run_copies(E, c3, dY, dNY.p, nullptr);
  if (keep_lambda) {
    std::vector<double> h((size_t)std::max<int64_t>(ts, 1));
    if (ts) hchk(hipMemcpyAsync(h.data(), Sb.p, sizeof(double) * (size_t)ts, hipMemcpyDeviceToHost, E.st), "cheb: S");
    E.sync();
    auto& L = M.lam[(size_t)(left ? site : site - 1)];
    L.clear();
    for (auto& kv : secs) L[kv.first].assign(h.begin() + kv.second.so, h.begin() + kv.second.so + kv.second.k);
  }
  if (left) {
    M.A[(size_t)sa] = NX;
    M.d[(size_t)sa] = std::move(dNX);
    M.A[(size_t)sb] = NY;
    M.d[(size_t)sb] = std::move(dNY);
  } else {
    M.A[(size_t)sb] = NX;
    M.d[(size_t)sb] = std::move(dNX);
    M.A[(size_t)sa] = NY;
    M.d[(size_t)sa] = std::move(dNY);
  }
}

// ------------------------------------------------------------ truncate
// DiagBlockMatrix::truncate(tr_weight, chimax, false) per vector: the largest
// values over the vector's sectors while |tw| > norm * tr_weight and fewer
// than chimax are kept (multimap<value, q> visited from the end: among equal
// values the last inserted first).  Returns the kept count per charge.
std::map<uint64_t, int64_t> lambda_keep(const std::map<uint64_t, std::vector<double>>& lam, unsigned chimax,
                                        double tr_weight) {
  std::map<uint64_t, std::vector<std::pair<double, uint64_t>>> by_vec;
  for (auto& kv : lam)
    for (double x : kv.second) by_vec[ns_of(kv.first)].push_back({x, kv.first});
  std::map<uint64_t, int64_t> keep;
  for (auto& vv : by_vec) {
    auto& all = vv.second;
    double nrm = 0.0;
    for (auto& a : all) nrm += a.first * a.first;
    nrm = std::sqrt(nrm);
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<double, uint64_t>& a, const std::pair<double, uint64_t>& b) {
                       return a.first < b.first;
                     });
    double tw = nrm;
    unsigned chi = 0;
    for (auto r = all.rbegin(); r != all.rend() && std::fabs(tw) > nrm * tr_weight && chi < chimax; ++r) {
      tw -= r->first * r->first / nrm;
      ++keep[r->second];
      ++chi;
    }
  }
  return keep;
}

// truncate(mps, chimax, delta) + dropDangling on every bond
void truncate_mps(Eng& E, DMps& M, unsigned chimax, double delta) {
  if (!chimax) return;
  std::vector<std::map<uint64_t, int64_t>> keep((size_t)E.N - 1);
  for (int b = 0; b + 1 < E.N; ++b) keep[(size_t)b] = lambda_keep(M.lam[(size_t)b], chimax, delta);
  for (int s = 0; s < E.N; ++s) {
    const HTab& T = M.A[(size_t)s];
    std::map<Key, std::pair<int64_t, int64_t>> nm;
    std::vector<int64_t> src;
    for (int64_t i = 0; i < T.n(); ++i) {
      int64_t r = T.rows[(size_t)i], c = T.cols[(size_t)i];
      if (s > 0) {
        auto it = keep[(size_t)s - 1].find(T.q[2 * i]);
        if (it == keep[(size_t)s - 1].end()) continue;
        r = std::min(r, it->second);
      }
      if (s + 1 < E.N) {
        auto it = keep[(size_t)s].find(T.q[2 * i + 1]);
        if (it == keep[(size_t)s].end()) continue;
        c = std::min(c, it->second);
      }
      nm[key_of(T, i)] = {r, c};
    }
    HTab NT = table_from(nm, T.cplx);
    Buf nb(E.pool, NT.len(), E.st);
    std::vector<Copy2D> cp;
    std::map<Key, int64_t> t_index;
    for (int64_t i = 0; i < T.n(); ++i) t_index[key_of(T, i)] = i;
    for (int64_t t = 0; t < NT.n(); ++t) {
      const int64_t i = t_index.at(key_of(NT, t));
      add_copy(cp, E, T.off[(size_t)i], T.total, NT.off[(size_t)t], NT.total, NT.rows[(size_t)t], NT.cols[(size_t)t],
               T.rows[(size_t)i], NT.rows[(size_t)t]);
    }
    run_copies(E, cp, M.d[(size_t)s].p, nb.p, nullptr);
    M.A[(size_t)s] = NT;
    M.d[(size_t)s] = std::move(nb);
  }
  for (int b = 0; b + 1 < E.N; ++b) {
    auto& L = M.lam[(size_t)b];
    std::map<uint64_t, std::vector<double>> nl;
    for (auto& kv : keep[(size_t)b]) {
      auto it = L.find(kv.first);
      if (it != L.end()) nl[kv.first].assign(it->second.begin(), it->second.begin() + kv.second);
    }
    L.swap(nl);
  }
}

// ------------------------------------------------------------ applyMPO
// The left bond of the applied state carried from one site to the next:
// (a, wl) -> new charge, and the sizes of the coarse blocks per new charge.
struct ApplyBond {
  std::map<std::pair<uint64_t, uint32_t>, uint64_t> new_left;
  std::map<uint64_t, std::vector<int64_t>> left_sizes;
};

ApplyBond apply_start(const DMps& M, const std::vector<HMpo>& W) {
  ApplyBond b;
  for (int64_t i = 0; i < M.A[0].n(); ++i) {  // site 0: each vector's boundary charge keeps its value
    b.new_left[{M.A[0].q[2 * i], 0u}] = M.A[0].q[2 * i];
    b.left_sizes[M.A[0].q[2 * i]] = std::vector<int64_t>(W[0].Dl, 1);
  }
  return b;
}

// site s of applyMPO(W, M) into (NT, nb); advances the bond state
void apply_site(Eng& E, ApplyBond& bond, const DMps& M, const std::vector<HMpo>& W, int s, HTab& NT_out,
                Buf& nb_out) {
  auto& new_left = bond.new_left;
  auto& left_sizes = bond.left_sizes;
  {
    const HTab& T = M.A[(size_t)s];
    const HMpo& w = W[(size_t)s];
    const auto& lmap = E.itk[(size_t)s];
    // five-leg pieces: ((a, wl), (b, wr), out) -> (block, value)
    typedef std::pair<uint64_t, uint32_t> DK;
    struct P {
      int64_t blk;
      double v;
    };
    std::map<std::tuple<DK, DK, uint32_t>, P> F;
    for (size_t e = 0; e < w.val.size(); ++e)
      for (int64_t i = 0; i < T.n(); ++i) {
        if (T.p[(size_t)i] != w.in[e]) continue;
        auto k = std::make_tuple(DK{T.q[2 * i], w.wl[e]}, DK{T.q[2 * i + 1], w.wr[e]}, w.out[e]);
        if (F.count(k)) throw HErr(MPSA_EINVAL, "cheb: applyMPO with two MPO entries for one block (MPO not supported)");
        F[k] = P{i, w.val[e]};
      }
    std::map<DK, uint64_t> new_right;
    for (auto& kv : F) {
      const DK& kl = std::get<0>(kv.first);
      auto it = new_left.find(kl);
      if (it == new_left.end()) continue;
      const DK& kr = std::get<1>(kv.first);
      if (!new_right.count(kr)) new_right[kr] = it->second + lmap[std::get<2>(kv.first)];
    }
    std::map<uint64_t, std::vector<DK>> inv_left, inv_right;
    for (auto& kv : new_left) inv_left[kv.second].push_back(kv.first);
    for (auto& kv : new_right) inv_right[kv.second].push_back(kv.first);
    std::map<uint64_t, std::vector<int64_t>> right_sizes;
    for (auto& il : inv_left)
      for (uint32_t sp = 0; sp < lmap.size(); ++sp) {
        const uint64_t qr = il.first + lmap[sp];
        auto ir = inv_right.find(qr);
        if (ir == inv_right.end()) continue;
        for (auto& kl : il.second)
          for (auto& kr : ir->second) {
            auto f = F.find(std::make_tuple(kl, kr, sp));
            if (f == F.end()) continue;
            auto& v = right_sizes[qr];
            if (v.empty()) v.assign(w.Dr, 0);
            v[kr.second] = std::max(v[kr.second], T.cols[(size_t)f->second.blk]);
          }
      }
    // blocks (ql, qr, s) of the coarse grid
    struct Piece {
      Key key;
      int64_t blk, r0, c0;
      double v;
    };
    std::vector<Piece> pieces;
    std::map<Key, std::pair<int64_t, int64_t>> nm;
    for (auto& il : inv_left) {
      const uint64_t ql = il.first;
      auto lsit = left_sizes.find(ql);
      std::vector<int64_t> ls = lsit != left_sizes.end() ? lsit->second : std::vector<int64_t>(w.Dl, 0);
      ls.resize(w.Dl, 0);
      for (uint32_t sp = 0; sp < lmap.size(); ++sp) {
        const uint64_t qr = ql + lmap[sp];
        auto ir = inv_right.find(qr);
        if (ir == inv_right.end()) continue;
        std::vector<std::pair<DK, DK>> found;
        for (auto& kl : il.second)
          for (auto& kr : ir->second)
            if (F.count(std::make_tuple(kl, kr, sp))) found.push_back({kl, kr});
        if (found.empty()) continue;
        const auto& rs = right_sizes[qr];
        std::vector<int64_t> sx(w.Dl + 1, 0), sy(w.Dr + 1, 0);
        for (uint32_t a = 0; a < w.Dl; ++a) sx[a + 1] = sx[a] + ls[a];
        for (uint32_t b = 0; b < w.Dr; ++b) sy[b + 1] = sy[b] + rs[b];
        const Key nk{ql, qr, sp};
        nm[nk] = {sx[w.Dl], sy[w.Dr]};
        for (auto& fr : found) {
          const P& pc = F[std::make_tuple(fr.first, fr.second, sp)];
          pieces.push_back(Piece{nk, pc.blk, sx[fr.first.second], sy[fr.second.second], pc.v});
        }
      }
    }
    HTab NT = table_from(nm, T.cplx);
    Buf nb(E.pool, NT.len(), E.st);
    std::map<Key, int64_t> where;
    for (int64_t t = 0; t < NT.n(); ++t) where[key_of(NT, t)] = t;
    std::vector<Copy2D> cp;
    for (auto& pc : pieces) {
      const int64_t t = where[pc.key];
      const int64_t rows = T.rows[(size_t)pc.blk], cols = T.cols[(size_t)pc.blk];
      add_copy(cp, E, T.off[(size_t)pc.blk], T.total, NT.off[(size_t)t] + pc.r0 + pc.c0 * NT.rows[(size_t)t], NT.total,
               rows, cols, rows, NT.rows[(size_t)t], pc.v);
    }
    run_copies(E, cp, M.d[(size_t)s].p, nb.p, nullptr);
    NT_out = NT;
    nb_out = std::move(nb);
    left_sizes = right_sizes;
    new_left = new_right;
  }
}

DMps apply_mpo(Eng& E, const DMps& M, const std::vector<HMpo>& W) {
  DMps O = new_mps(E);
  ApplyBond bond = apply_start(M, W);
  for (int s = 0; s < E.N; ++s) apply_site(E, bond, M, W, s, O.A[(size_t)s], O.d[(size_t)s]);
  return O;
}

std::vector<double> real_parts(const std::vector<cplx>& v);

// computeOverlap(t, t) of t = applyMPO(W, M) per vector, with t made one site
// at a time and folded into the running left environment at once: the
// applied state (bond w D) is never held whole.  Same launches, in the same
// order, as overlap(E, t, t) on the whole applied state.
std::vector<double> applied_norm_sq(Eng& E, const DMps& M, const std::vector<HMpo>& W) {
  // the applied state's boundary charges are the state's own
  HTab Lt = lmin_tab_of(M, E.planes == 2), Bt;
  Buf Lb = identity_buf(E, Lt), Bb;
  ApplyBond bond = apply_start(M, W);
  for (int s = 0; s < E.N; ++s) {
    apply_site(E, bond, M, W, s, Bt, Bb);
    HTab nt;
    Buf nb;
    env_layer(E, Lt, Lb.p, Bt, Bb.p, Bt, Bb.p, E.Id[(size_t)s], E.itk[(size_t)s], 1, nt, nb);
    Lt = nt;
    Lb = std::move(nb);
  }
  return real_parts(env_scalars(E, Lt, Lb));
}

// ------------------------------------------------------------ MPS sums
// xa a + xb b as the reference's += (first site side by side, last site
// stacked, bulk outersum, sizes fixed): a direct sum of the bond spaces with
// a's part first; the scale of each summand sits on its first site.
DMps axpby_mps(Eng& E, double xa, const DMps& a, double xb, const DMps& b) {
  const int N = E.N;
  // bond dims per charge of each summand (bond s between sites s and s+1)
  auto dims = [&](const DMps& M, int bond) {
    std::map<uint64_t, int64_t> d;
    const HTab& L = M.A[(size_t)bond];
    for (int64_t i = 0; i < L.n(); ++i) d[L.q[2 * i + 1]] = std::max(d[L.q[2 * i + 1]], L.cols[(size_t)i]);
    const HTab& R = M.A[(size_t)bond + 1];
    for (int64_t i = 0; i < R.n(); ++i) d[R.q[2 * i]] = std::max(d[R.q[2 * i]], R.rows[(size_t)i]);
    return d;
  };
  std::vector<std::map<uint64_t, int64_t>> da, db;
  for (int bnd = 0; bnd + 1 < N; ++bnd) {
    da.push_back(dims(a, bnd));
    db.push_back(dims(b, bnd));
  }
  auto get = [](const std::map<uint64_t, int64_t>& m, uint64_t q) {
    auto it = m.find(q);
    return it == m.end() ? (int64_t)0 : it->second;
  };
  DMps O = new_mps(E);
  for (int s = 0; s < N; ++s) {
    const HTab &Ta = a.A[(size_t)s], &Tb = b.A[(size_t)s];
    std::map<Key, std::pair<int64_t, int64_t>> nm;
    auto shape = [&](const Key& k) {
      const int64_t r = s == 0 ? 1 : get(da[(size_t)s - 1], k[0]) + get(db[(size_t)s - 1], k[0]);
      const int64_t c = s == N - 1 ? 1 : get(da[(size_t)s], k[1]) + get(db[(size_t)s], k[1]);
      return std::make_pair(r, c);
    };
    for (int64_t i = 0; i < Ta.n(); ++i) nm[key_of(Ta, i)] = shape(key_of(Ta, i));
    for (int64_t i = 0; i < Tb.n(); ++i) nm[key_of(Tb, i)] = shape(key_of(Tb, i));
    HTab NT = table_from(nm, Ta.cplx);
    Buf nb(E.pool, NT.len(), E.st);
    std::map<Key, int64_t> where;
    for (int64_t t = 0; t < NT.n(); ++t) where[key_of(NT, t)] = t;
    std::vector<Copy2D> ca, cb;
    for (int64_t i = 0; i < Ta.n(); ++i) {
      const int64_t t = where[key_of(Ta, i)];
      add_copy(ca, E, Ta.off[(size_t)i], Ta.total, NT.off[(size_t)t], NT.total, Ta.rows[(size_t)i], Ta.cols[(size_t)i],
               Ta.rows[(size_t)i], NT.rows[(size_t)t], s == 0 ? xa : 1.0);
    }
    for (int64_t i = 0; i < Tb.n(); ++i) {
      const Key k = key_of(Tb, i);
      const int64_t t = where[k];
      const int64_t r0 = s == 0 ? 0 : get(da[(size_t)s - 1], k[0]);
      const int64_t c0 = s == N - 1 ? 0 : get(da[(size_t)s], k[1]);
      add_copy(cb, E, Tb.off[(size_t)i], Tb.total, NT.off[(size_t)t] + r0 + c0 * NT.rows[(size_t)t], NT.total,
               Tb.rows[(size_t)i], Tb.cols[(size_t)i], Tb.rows[(size_t)i], NT.rows[(size_t)t], s == 0 ? xb : 1.0);
    }
    run_copies(E, ca, a.d[(size_t)s].p, nb.p, nullptr);
    run_copies(E, cb, b.d[(size_t)s].p, nb.p, nullptr);
    O.A[(size_t)s] = NT;
    O.d[(size_t)s] = std::move(nb);
  }
  return O;
}

// ------------------------------------------------------------ compression
double conv_of(double n1sq, double n2sq, cplx ov) { return (n1sq + n2sq - 2.0 * std::abs(ov)) / n1sq; }

// Copy the blocks of the vectors marked in `mask` from `from` to `to` (both
// in layout T).
void copy_vector_blocks(Eng& E, const HTab& T, const double* from, double* to, const std::vector<char>& mask) {
  std::vector<Copy2D> cp;
  for (int64_t i = 0; i < T.n(); ++i) {
    if (!mask[(size_t)ns_of(T.q[2 * i])]) continue;
    add_copy(cp, E, T.off[(size_t)i], T.total, T.off[(size_t)i], T.total, T.rows[(size_t)i], T.cols[(size_t)i],
             T.rows[(size_t)i], T.rows[(size_t)i]);
  }
  run_copies(E, cp, from, to, nullptr);
}

bool same_tables(const HTab& a, const HTab& b) {
  return a.q == b.q && a.p == b.p && a.rows == b.rows && a.cols == b.cols;
}
bool same_tables(const DMps& a, const DMps& b) {
  for (size_t s = 0; s < a.A.size(); ++s)
    if (!same_tables(a.A[s], b.A[s])) return false;
  return true;
}

// The variational sweeps shared by compressSingleSite (MPO W) and compress
// (identity MPO): comp is optimised against `mps` with mixed environments
// (mps below, comp above).  Every vector follows the reference's loop: stop
// when conv grows, converge when step >= maxstep or |conv| < delta; a vector
// that has stopped keeps its state (and its BR[1]) while the others sweep on.
// With final_site0 the left-most tensor is solved once more from BR[1] (the
// end of compress).
void sweeps(Eng& E, DMps& comp, const DMps& mps, const std::vector<HMpo>& Wv, int maxstep, double delta,
            const std::vector<double>& n1sq, bool final_site0) {
  const int N = E.N, cp = E.planes == 2;
  if (maxstep == 0) return;
  std::vector<HTab> BLt((size_t)N + 1), BRt((size_t)N + 1);  // BLt[s+1] = BL[s]; BRt[s] = BR[s]
  std::vector<Buf> BLb((size_t)N + 1), BRb((size_t)N + 1);
  BLt[0] = lmin_tab_of(mps, cp);
  BLb[0] = identity_buf(E, BLt[0]);
  BRt[(size_t)N] = rmax_tab_of(mps, cp);
  BRb[(size_t)N] = identity_buf(E, BRt[(size_t)N]);
  for (int s = N - 1; s > 0; --s)
    env_layer(E, BRt[(size_t)s + 1], BRb[(size_t)s + 1].p, mps.A[(size_t)s], mps.d[(size_t)s].p, comp.A[(size_t)s],
              comp.d[(size_t)s].p, Wv[(size_t)s], E.itk[(size_t)s], -1, BRt[(size_t)s], BRb[(size_t)s]);
  std::vector<char> done((size_t)E.nvec, 0);
  std::vector<double> conv_old((size_t)E.nvec, 1.0);
  HTab br1_t;   // BR[1] of every vector as of its last sweep
  Buf br1_b;
  int step = 1;
  while (true) {
    bool any = false, frozen_any = false;
    for (char d : done) {
      any = any || !d;
      frozen_any = frozen_any || d;
    }
    if (!any) break;
    DMps before;
    if (frozen_any) before = copy_mps(E, comp);
    for (int s = 0; s < N - 1; ++s) {
      HTab t;
      Buf b;
      heff1(E, BLt[(size_t)s], BLb[(size_t)s].p, mps.A[(size_t)s], mps.d[(size_t)s].p, BRt[(size_t)s + 1],
            BRb[(size_t)s + 1].p, Wv[(size_t)s], E.itk[(size_t)s], t, b);
      comp.A[(size_t)s] = t;
      comp.d[(size_t)s] = std::move(b);
      prepare_site(E, comp, s, true, false);
      env_layer(E, BLt[(size_t)s], BLb[(size_t)s].p, mps.A[(size_t)s], mps.d[(size_t)s].p, comp.A[(size_t)s],
                comp.d[(size_t)s].p, Wv[(size_t)s], E.itk[(size_t)s], 1, BLt[(size_t)s + 1], BLb[(size_t)s + 1]);
    }
    for (int s = N - 1; s > 0; --s) {
      HTab t;
      Buf b;
      heff1(E, BLt[(size_t)s], BLb[(size_t)s].p, mps.A[(size_t)s], mps.d[(size_t)s].p, BRt[(size_t)s + 1],
            BRb[(size_t)s + 1].p, Wv[(size_t)s], E.itk[(size_t)s], t, b);
      comp.A[(size_t)s] = t;
      comp.d[(size_t)s] = std::move(b);
      prepare_site(E, comp, s, false, false);
      env_layer(E, BRt[(size_t)s + 1], BRb[(size_t)s + 1].p, mps.A[(size_t)s], mps.d[(size_t)s].p, comp.A[(size_t)s],
                comp.d[(size_t)s].p, Wv[(size_t)s], E.itk[(size_t)s], -1, BRt[(size_t)s], BRb[(size_t)s]);
    }
    if (frozen_any) {
      if (!same_tables(before, comp) || (final_site0 && br1_b.p && !same_tables(br1_t, BRt[1])))
        throw HErr(MPSA_EINVAL, "cheb: block structure changed between compression sweeps");
      for (int s = 0; s < N; ++s)
        copy_vector_blocks(E, comp.A[(size_t)s], before.d[(size_t)s].p, comp.d[(size_t)s].p, done);
    }
    const std::vector<cplx> n2 = overlap(E, comp, comp), ov = overlap(E, comp, mps);
    std::vector<char> newly((size_t)E.nvec, 0);
    for (int v = 0; v < E.nvec; ++v) {
      if (done[(size_t)v]) continue;
      const double conv = conv_of(n1sq[(size_t)v], n2[(size_t)v].real(), ov[(size_t)v]);
      if (conv_old[(size_t)v] < conv) {
        done[(size_t)v] = newly[(size_t)v] = 1;
        continue;
      }
      conv_old[(size_t)v] = conv;
      if (step >= maxstep || std::fabs(conv) < delta) done[(size_t)v] = newly[(size_t)v] = 1;
    }
    if (final_site0) {
      if (!br1_b.p) {
        br1_t = BRt[1];
        br1_b = Buf(E.pool, br1_t.len(), E.st);
      }
      copy_vector_blocks(E, br1_t, BRb[1].p, br1_b.p, newly);
    }
    ++step;
  }
  if (final_site0) {
    HTab t;
    Buf b;
    heff1(E, BLt[0], BLb[0].p, mps.A[0], mps.d[0].p, br1_t, br1_b.p, Wv[0], E.itk[0], t, b);
    comp.A[0] = t;
    comp.d[0] = std::move(b);
  }
}

std::vector<double> real_parts(const std::vector<cplx>& v) {
  std::vector<double> r;
  for (auto& x : v) r.push_back(x.real());
  return r;
}

// compressSingleSite(mps, W, chimax, maxstep, delta): returns comp ~ H mps
DMps compress_single_site(Eng& E, const DMps& mps, unsigned chimax, int maxstep, double delta,
                          std::vector<double>& conv) {
  const int N = E.N;
  // n1^2 = computeOverlap(applyMPO(W, mps), applyMPO(W, mps)), site by site
  const std::vector<double> n1sq = applied_norm_sq(E, mps, E.W);
  DMps tmps = copy_mps(E, mps);
  // the right-to-left pass sets every bond's Schmidt values: the left pass
  // only brings the state to left-canonical form
  for (int s = 0; s < N - 1; ++s) prepare_site(E, tmps, s, true, false);
  for (int s = N - 1; s > 0; --s) prepare_site(E, tmps, s, false, true);
  const unsigned D = N > 1 ? E.W[1].Dr : 1;
  truncate_mps(E, tmps, (unsigned)std::floor((double)chimax / (double)D), delta);
  DMps comp = apply_mpo(E, tmps, E.W);
  tmps = DMps();
  for (int s = 0; s < N - 1; ++s) prepare_site(E, comp, s, true, false);
  for (int s = N - 1; s > 0; --s) prepare_site(E, comp, s, false, false);
  sweeps(E, comp, mps, E.W, maxstep, delta, n1sq, false);
  const std::vector<cplx> n2 = overlap(E, comp, comp), ov = overlap(E, comp, mps);
  conv.assign((size_t)E.nvec, 0.0);
  for (int v = 0; v < E.nvec; ++v) conv[(size_t)v] = conv_of(n1sq[(size_t)v], n2[(size_t)v].real(), ov[(size_t)v]);
  return comp;
}

// compress(mps, chimax, maxstep, delta, normalize = false): mps is
// canonicalised in place, the compressed state is returned
DMps compress(Eng& E, DMps& mps, unsigned chimax, int maxstep, double delta, std::vector<double>& conv) {
  const int N = E.N;
  // the right-to-left pass starts at site N - 2, so bond N - 2 keeps the
  // Schmidt values of the left pass (as in the reference)
  for (int s = 0; s < N - 1; ++s) prepare_site(E, mps, s, true, s == N - 2);
  for (int s = N - 2; s > 0; --s) prepare_site(E, mps, s, false, true);
  DMps comp = copy_mps(E, mps);
  truncate_mps(E, comp, chimax, delta);
  for (int s = 0; s < N - 1; ++s) prepare_site(E, comp, s, true, false);
  for (int s = N - 1; s > 0; --s) prepare_site(E, comp, s, false, false);
  const std::vector<double> n1sq = real_parts(overlap(E, mps, mps));
  sweeps(E, comp, mps, E.Id, maxstep, delta, n1sq, true);  // ends with the site-0 solve
  for (int s = 0; s < N - 1; ++s) prepare_site(E, comp, s, true, false);
  for (int s = N - 1; s > 0; --s) prepare_site(E, comp, s, false, false);
  const std::vector<cplx> n2 = overlap(E, comp, comp), ov = overlap(E, comp, mps);
  conv.assign((size_t)E.nvec, 0.0);
  for (int v = 0; v < E.nvec; ++v) conv[(size_t)v] = conv_of(n1sq[(size_t)v], n2[(size_t)v].real(), ov[(size_t)v]);
  return comp;
}

// orthonormalize(mps, "rl") (src/MPSfunctions.hpp:1757-1793): prepareSite
// "r" from the last site to site 1, then the first site divided by each
// vector's norm (the reference's prepareMatrix into the left dummy keeps
// V^dagger there; this differs at most by the sign of the whole state).
void orthonormalize_rl(Eng& E, DMps& M) {
  for (int s = E.N - 1; s > 0; --s) prepare_site(E, M, s, false, false);
  const std::vector<cplx> nrm = overlap(E, M, M);
  const HTab& T = M.A[0];
  std::vector<Copy2D> cp;
  for (int64_t i = 0; i < T.n(); ++i) {
    const double n2 = nrm[(size_t)ns_of(T.q[2 * i])].real();
    if (!(n2 > 0.0)) throw HErr(MPSA_EINVAL, "cheb: orthonormalize of a vector with zero norm");
    add_copy(cp, E, T.off[(size_t)i], T.total, T.off[(size_t)i], T.total, T.rows[(size_t)i], T.cols[(size_t)i],
             T.rows[(size_t)i], T.rows[(size_t)i], 1.0 / std::sqrt(n2));
  }
  run_copies(E, cp, M.d[0].p, M.d[0].p, nullptr);
}

DMps& slot_at(Eng& E, int i, bool must_exist) {
  if (i < 0 || i >= (int)E.slot.size()) throw HErr(MPSA_EINVAL, "cheb: slot index out of range");
  if (must_exist && !E.slot[(size_t)i].valid()) throw HErr(MPSA_EINVAL, "cheb: empty slot");
  return E.slot[(size_t)i];
}

int fail(const std::exception& e, int code) {
  set_last_error(e.what());
  return code;
}

}  // namespace

extern "C" {

int mpsa_cheb_create(mpsa_cheb** out, int nsites, int nvec, const mpsa_blocks* sites, const double* const* dA,
                     const mpsa_mpo* W, const uint64_t* itokey, const int* d, int h1_precision, mpsa_stream stream) {
  mpsa_cheb* E = nullptr;
  try {
    if (!out || nsites < 2 || nvec < 1 || nvec > 4096 || !sites || !dA || !W || !itokey || !d)
      throw HErr(MPSA_EINVAL, "cheb_create: bad arguments");
    if (h1_precision != MPSA_F64 && h1_precision != MPSA_F32)
      throw HErr(MPSA_EINVAL, "cheb_create: the one-site contraction runs in FP64 or FP32");
    E = new mpsa_cheb();
    E->st = (hipStream_t)stream;
    E->N = nsites;
    E->nvec = nvec;
    E->h1prec = h1_precision;
    E->planes = sites[0].is_complex ? 2 : 1;
    const uint64_t* ik = itokey;
    for (int i = 0; i < nsites; ++i) {
      if (d[i] < 1) throw HErr(MPSA_EINVAL, "cheb_create: d < 1");
      E->itk.emplace_back(ik, ik + d[i]);
      ik += d[i];
      for (auto x : E->itk.back())
        if (x >> kNsShift) throw HErr(MPSA_EINVAL, "cheb_create: i_to_key value above 2^40");
      E->W.push_back(copy_mpo(&W[i]));
      HMpo id;
      id.Dl = id.Dr = 1;
      for (int s = 0; s < d[i]; ++s) {
        id.wl.push_back(0);
        id.wr.push_back(0);
        id.out.push_back((uint32_t)s);
        id.in.push_back((uint32_t)s);
        id.val.push_back(1.0);
      }
      E->Id.push_back(id);
    }
    E->slot.resize(8);
    DMps M = new_mps(*E);
    for (int s = 0; s < nsites; ++s) {
      std::map<Key, std::pair<int64_t, int64_t>> nm;
      struct Src {
        int v;
        int64_t i;
      };
      std::map<Key, Src> src;
      for (int v = 0; v < nvec; ++v) {
        const mpsa_blocks& b = sites[(size_t)v * nsites + s];
        if (b.ns != 1 || b.n < 0 || (b.is_complex != 0) != (E->planes == 2))
          throw HErr(MPSA_EINVAL, "cheb_create: site tables must be ns = 1 and all real or all complex");
        for (int64_t i = 0; i < b.n; ++i) {
          if ((b.q[2 * i] >> kNsShift) || (b.q[2 * i + 1] >> kNsShift))
            throw HErr(MPSA_EINVAL, "cheb_create: charge above 2^40");
          const Key k{b.q[2 * i] + ((uint64_t)v << kNsShift), b.q[2 * i + 1] + ((uint64_t)v << kNsShift),
                      (uint64_t)b.p[i]};
          nm[k] = {b.rows[i], b.cols[i]};
          src[k] = Src{v, i};
        }
      }
      HTab T = table_from(nm, E->planes == 2);
      Buf buf(E->pool, T.len(), E->st);
      // one batched copy per vector: its blocks into the shared table
      std::vector<std::vector<Copy2D>> per_vec((size_t)nvec);
      std::vector<int64_t> tot((size_t)nvec, 0);
      for (int v = 0; v < nvec; ++v) {
        const mpsa_blocks& b = sites[(size_t)v * nsites + s];
        for (int64_t i = 0; i < b.n; ++i)
          tot[(size_t)v] = std::max<int64_t>(tot[(size_t)v], b.offset[i] + b.rows[i] * b.cols[i]);
      }
      for (int64_t t = 0; t < T.n(); ++t) {
        const Src& sr = src[key_of(T, t)];
        const mpsa_blocks& b = sites[(size_t)sr.v * nsites + s];
        add_copy(per_vec[(size_t)sr.v], *E, b.offset[sr.i], tot[(size_t)sr.v], T.off[(size_t)t], T.total,
                 b.rows[sr.i], b.cols[sr.i], b.rows[sr.i], b.rows[sr.i]);
      }
      for (int v = 0; v < nvec; ++v) {
        if (per_vec[(size_t)v].empty()) continue;
        const double* base = dA[(size_t)v * nsites + s];
        if (!base) throw HErr(MPSA_EINVAL, "cheb_create: null site buffer");
        run_copies(*E, per_vec[(size_t)v], base, buf.p, nullptr);
      }
      M.A[(size_t)s] = T;
      M.d[(size_t)s] = std::move(buf);
    }
    E->slot[0] = std::move(M);
    E->sync();
    *out = E;
    return MPSA_OK;
  } catch (const HErr& e) {
    delete E;
    return fail(e, e.code);
  } catch (const std::exception& e) {
    delete E;
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_copy(mpsa_cheb* E, int dst, int src) {
  try {
    if (!E) throw HErr(MPSA_EINVAL, "cheb: null engine");
    DMps c = copy_mps(*E, slot_at(*E, src, true));
    slot_at(*E, dst, false) = std::move(c);
    E->sync();
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_orthonormalize(mpsa_cheb* E, int slot) {
  try {
    if (!E) throw HErr(MPSA_EINVAL, "cheb: null engine");
    orthonormalize_rl(*E, slot_at(*E, slot, true));
    E->sync();
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_apply_mpo(mpsa_cheb* E, int dst, int src) {
  try {
    if (!E) throw HErr(MPSA_EINVAL, "cheb: null engine");
    DMps o = apply_mpo(*E, slot_at(*E, src, true), E->W);
    slot_at(*E, dst, false) = std::move(o);
    E->sync();
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_overlap(mpsa_cheb* E, int a, int b, double* out) {
  try {
    if (!E || !out) throw HErr(MPSA_EINVAL, "cheb: bad arguments");
    const auto v = overlap(*E, slot_at(*E, a, true), slot_at(*E, b, true));
    for (size_t i = 0; i < v.size(); ++i) {
      out[2 * i] = v[i].real();
      out[2 * i + 1] = v[i].imag();
    }
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_axpby(mpsa_cheb* E, int dst, double xa, int a, double xb, int b) {
  try {
    if (!E) throw HErr(MPSA_EINVAL, "cheb: null engine");
    DMps o = axpby_mps(*E, xa, slot_at(*E, a, true), xb, slot_at(*E, b, true));
    slot_at(*E, dst, false) = std::move(o);
    E->sync();
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_compress_single_site(mpsa_cheb* E, int dst, int src, unsigned chimax, int maxstep, double delta,
                                   double* conv) {
  try {
    if (!E || maxstep < 0) throw HErr(MPSA_EINVAL, "cheb: bad arguments");
    std::vector<double> c;
    DMps o = compress_single_site(*E, slot_at(*E, src, true), chimax, maxstep, delta, c);
    slot_at(*E, dst, false) = std::move(o);
    E->sync();
    if (conv) std::copy(c.begin(), c.end(), conv);
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_compress(mpsa_cheb* E, int slot, unsigned chimax, int maxstep, double delta, double* conv) {
  try {
    if (!E || maxstep < 0) throw HErr(MPSA_EINVAL, "cheb: bad arguments");
    std::vector<double> c;
    DMps o = compress(*E, slot_at(*E, slot, true), chimax, maxstep, delta, c);
    slot_at(*E, slot, false) = std::move(o);
    E->sync();
    if (conv) std::copy(c.begin(), c.end(), conv);
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

// simulateSSCompress with the start vectors of slot 0 as t0 and as the final
// state (opi = opf = identity): mu_n = <t_n|t_0> per vector (re, im),
// reconstructed doubled moments, and both convergence measures per moment.
int mpsa_cheb_moments(mpsa_cheb* E, int nmom, unsigned chimax, int ncmax, double delta, double* moments,
                      double* recons, double* conv1, double* conv2, double* seconds) {
  try {
    if (!E || nmom < 1 || ncmax < 0 || !moments) throw HErr(MPSA_EINVAL, "cheb_moments: bad arguments");
    const auto t0c = std::chrono::steady_clock::now();
    const int nv = E->nvec;
    auto put = [&](double* arr, int ld, int idx, int v, cplx x) {
      if (!arr) return;
      arr[((size_t)v * ld + idx) * 2] = x.real();
      arr[((size_t)v * ld + idx) * 2 + 1] = x.imag();
    };
    auto put_r = [&](double* arr, int idx, int v, double x) {
      if (arr) arr[(size_t)v * nmom + idx] = x;
    };
    const int nrec = 2 * nmom - 1;
    // wall time of every moment (each ends with overlaps read back to the host)
    E->moment_s.assign((size_t)nmom, 0.0);
    auto lap = [&, t = t0c](int idx) mutable {
      const auto now = std::chrono::steady_clock::now();
      E->moment_s[(size_t)idx] = std::chrono::duration<double>(now - t).count();
      t = now;
    };
    DMps& f = slot_at(*E, 0, true);
    DMps t0 = copy_mps(*E, f);
    auto mu0 = overlap(*E, f, t0);
    for (int v = 0; v < nv; ++v) {
      put(moments, nmom, 0, v, mu0[(size_t)v]);
      put_r(conv1, 0, v, 0.0);
      put_r(conv2, 0, v, 0.0);
    }
    lap(0);
    if (nmom > 1) {
      std::vector<double> c1;
      DMps t1 = compress_single_site(*E, t0, chimax, ncmax, delta, c1);
      auto mu1 = overlap(*E, f, t1);
      auto a = overlap(*E, t0, t0), b = overlap(*E, t1, t0), c = overlap(*E, t1, t1);
      for (int v = 0; v < nv; ++v) {
        put(moments, nmom, 1, v, mu1[(size_t)v]);
        put(recons, nrec, 0, v, 2.0 * a[(size_t)v] - mu0[(size_t)v]);
        put(recons, nrec, 1, v, 2.0 * b[(size_t)v] - mu1[(size_t)v]);
        put(recons, nrec, 2, v, 2.0 * c[(size_t)v] - mu0[(size_t)v]);
        put_r(conv1, 1, v, c1[(size_t)v]);
        put_r(conv2, 1, v, 0.0);
      }
      lap(1);
      for (int n = 2; n < nmom; ++n) {
        DMps tm = compress_single_site(*E, t1, chimax, ncmax, delta, c1);
        DMps t2 = axpby_mps(*E, 2.0, tm, -1.0, t0);
        tm = DMps();
        auto nsq = overlap(*E, t2, t2);
        auto mun = overlap(*E, f, t2);
        auto x = overlap(*E, t2, t1);
        std::vector<double> c2;
        DMps t2c = compress(*E, t2, chimax, ncmax, delta, c2);
        for (int v = 0; v < nv; ++v) {
          put(moments, nmom, n, v, mun[(size_t)v]);
          put(recons, nrec, 2 * n - 1, v, 2.0 * x[(size_t)v] - mu1[(size_t)v]);
          put(recons, nrec, 2 * n, v, 2.0 * nsq[(size_t)v] - mu0[(size_t)v]);
          put_r(conv1, n, v, c1[(size_t)v]);
          put_r(conv2, n, v, c2[(size_t)v]);
        }
        t0 = std::move(t1);
        t1 = std::move(t2c);
        lap(n);
      }
      slot_at(*E, 1, false) = std::move(t1);  // the last moment's state
    }
    E->sync();
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0c).count();
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

// Site tensor `site` of vector `vec` in slot `slot`: its own charges (the
// vector's bits removed), packed at 64-element offsets; planar when complex.
int mpsa_cheb_site(mpsa_cheb* E, int slot, int vec, int site, int64_t* n_blocks, int64_t* len, uint64_t* q,
                   uint32_t* p, int64_t* rows, int64_t* cols, int64_t* offsets, double* dst) {
  try {
    if (!E || vec < 0 || vec >= E->nvec || site < 0 || site >= E->N || !n_blocks || !len)
      throw HErr(MPSA_EINVAL, "cheb_site: bad arguments");
    const DMps& M = slot_at(*E, slot, true);
    const HTab& T = M.A[(size_t)site];
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < T.n(); ++i)
      if (ns_of(T.q[2 * i]) == (uint64_t)vec) idx.push_back(i);
    int64_t o = 0;
    std::vector<int64_t> offs;
    for (int64_t i : idx) {
      o = align64(o);
      offs.push_back(o);
      o += T.rows[(size_t)i] * T.cols[(size_t)i];
    }
    *n_blocks = (int64_t)idx.size();
    *len = o * E->planes;
    const uint64_t base = (uint64_t)vec << kNsShift;
    if (q && p && rows && cols && offsets)
      for (size_t j = 0; j < idx.size(); ++j) {
        q[2 * j] = T.q[2 * idx[j]] - base;
        q[2 * j + 1] = T.q[2 * idx[j] + 1] - base;
        p[j] = T.p[(size_t)idx[j]];
        rows[j] = T.rows[(size_t)idx[j]];
        cols[j] = T.cols[(size_t)idx[j]];
        offsets[j] = offs[j];
      }
    if (dst && o > 0) {
      std::vector<Copy2D> cp;
      for (size_t j = 0; j < idx.size(); ++j) {
        const int64_t i = idx[j];
        add_copy(cp, *E, T.off[(size_t)i], T.total, offs[j], o, T.rows[(size_t)i], T.cols[(size_t)i],
                 T.rows[(size_t)i], T.rows[(size_t)i]);
      }
      run_copies(*E, cp, M.d[(size_t)site].p, dst, nullptr);
      E->sync();
    }
    return MPSA_OK;
  } catch (const HErr& e) {
    return fail(e, e.code);
  } catch (const std::exception& e) {
    return fail(e, MPSA_EINVAL);
  }
}

int mpsa_cheb_stats(const mpsa_cheb* E, int64_t* plans_built, int64_t* pool_bytes) {
  if (!E) {
    set_last_error("cheb_stats: null engine");
    return MPSA_EINVAL;
  }
  if (plans_built) *plans_built = E->plans_built;
  if (pool_bytes) {
    int64_t b = 0;
    for (auto& a : E->pool.all) b += (int64_t)a.second;
    *pool_bytes = b;
  }
  return MPSA_OK;
}

int mpsa_cheb_profile(const mpsa_cheb* E, double* out5) {
  if (!E || !out5) {
    set_last_error("cheb_profile: null argument");
    return MPSA_EINVAL;
  }
  out5[0] = E->t_plan;
  out5[1] = E->t_keys;
  out5[2] = E->t_sync;
  out5[3] = (double)E->n_sync;
  out5[4] = (double)E->plans.size();
  return MPSA_OK;
}

int mpsa_cheb_timers(const mpsa_cheb* E, double* out, int n) {
  if (!E || !out || n < 0) {
    set_last_error("cheb_timers: null argument or n < 0");
    return MPSA_EINVAL;
  }
  const double v[9] = {E->t_heff, E->t_env, E->t_prep, E->t_sync, E->pool.t_alloc, (double)E->pool.n_alloc,
                       E->t_plan, (double)E->n_sync, (double)E->plans.size()};
  for (int i = 0; i < n && i < 9; ++i) out[i] = v[i];
  return MPSA_OK;
}

int mpsa_cheb_set_big_sector(mpsa_cheb* E, int k) {
  if (!E || k < 1 || k > kSvdMaxR) {
    set_last_error("cheb_set_big_sector: k must be in [1, 1024]");
    return MPSA_EINVAL;
  }
  E->big_sector = k;
  return MPSA_OK;
}

int mpsa_cheb_moment_times(const mpsa_cheb* E, double* out, int cap, int* n) {
  if (!E || !n || cap < 0 || (cap > 0 && !out)) {
    set_last_error("cheb_moment_times: bad arguments");
    return MPSA_EINVAL;
  }
  *n = (int)E->moment_s.size();
  for (int i = 0; i < cap && i < *n; ++i) out[i] = E->moment_s[(size_t)i];
  return MPSA_OK;
}

int mpsa_cheb_set_block_jacobi(mpsa_cheb* E, int r) {
  if (!E || r < 0) {
    set_last_error("cheb_set_block_jacobi: r must be >= 0");
    return MPSA_EINVAL;
  }
  E->block_min_r = r;
  return MPSA_OK;
}

int mpsa_cheb_factor_log(mpsa_cheb* E, int64_t cap) {
  if (!E || cap < 0) {
    set_last_error("cheb_factor_log: bad arguments");
    return MPSA_EINVAL;
  }
  E->flog.clear();
  E->flog_cap = cap;
  return MPSA_OK;
}

int mpsa_cheb_factor_records(const mpsa_cheb* E, int64_t* out, int64_t cap, int64_t* n) {
  if (!E || !n || cap < 0 || (cap > 0 && !out)) {
    set_last_error("cheb_factor_records: bad arguments");
    return MPSA_EINVAL;
  }
  const int64_t w = mpsa_cheb::kFlogWidth;
  *n = (int64_t)E->flog.size() / w;
  for (int64_t i = 0; i < std::min(cap, *n) * w; ++i) out[i] = E->flog[(size_t)i];
  return MPSA_OK;
}

int mpsa_cheb_destroy(mpsa_cheb* E) {
  delete E;
  return MPSA_OK;
}

}  // extern "C"
