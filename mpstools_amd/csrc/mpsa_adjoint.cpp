// mpsa_adjoint.cpp -- energy of an MPS under an MPO, E = <psi|H|psi>, and its
// adjoint: dE/dA[i] for every site and dE/dtheta for MPO parameters.
//
// Forward: MPO<T>::measure (src/MPOTypes.hpp:990-1017) -- a left environment
// sweep L[i] = addContractionLayer(L[i-1], A[i], W[i], +1) from getLMin
// (src/MPSfunctions.hpp:1917-1926), then the trace of the charge-diagonal
// blocks of L[N-1].  Every K-th environment is kept (a checkpoint); the others
// are dropped.
//
// Backward (no reference counterpart): E is quadratic in each A[i] with the
// others fixed, E = a_i^T M_i a_i with M_i the one-site effective matrix
// built from L[i-1], W[i], R[i+1].  For a real symmetric H, M_i is symmetric
// and dE/dA[i] = 2 M_i a_i = 2 contractSingle(L[i-1], A[i], W[i], R[i+1])
// restricted to A[i]'s blocks.  E is affine in every MPO parameter theta_p,
// W[i](theta) = W0[i] + sum_p theta_p G_p[i], so
//   dE/dtheta_p = sum_i <A[i] | contractSingle(L[i-1], A[i], G_p[i], R[i+1])>
// with the environments of the undifferentiated W.  The right sweep builds
// R[i+1] from getRMax (src/MPSfunctions.hpp:1996-2022) site by site; the
// left environments of each K-site segment are recomputed from its
// checkpoint when the sweep reaches the segment.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

using namespace mpsa;

namespace {

struct AErr : std::runtime_error {
  int code;
  AErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void chk(int rc) {
  if (rc != 0) throw AErr(rc, mpsa_last_error());
}
void hchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw AErr(MPSA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int64_t align64(int64_t x) { return (x + 63) / 64 * 64; }

// A key table owned on the host.
struct HTab {
  int ns = 1;
  std::vector<uint64_t> q;
  std::vector<uint32_t> p;
  std::vector<int64_t> rows, cols, off;
  int64_t total = 0;
  int64_t n() const { return (int64_t)rows.size(); }
  mpsa_blocks view() const {
    mpsa_blocks b;
    b.n = n();
    b.ns = ns;
    b.is_complex = 0;
    b.q = q.data();
    b.p = p.data();
    b.rows = rows.data();
    b.cols = cols.data();
    b.offset = off.data();
    return b;
  }
  void pack() {  // 64-element aligned offsets in key order
    off.assign(rows.size(), 0);
    int64_t o = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
      o = align64(o);
      off[i] = o;
      o += rows[i] * cols[i];
    }
    total = o;
  }
};

HTab copy_tab(const mpsa_blocks* b, int ns) {
  if (!b || b->ns != ns || b->is_complex || b->n < 0) throw AErr(MPSA_EINVAL, "chain: bad site table");
  HTab t;
  t.ns = ns;
  t.q.assign(b->q, b->q + 2 * b->n);
  t.p.assign(b->p, b->p + ns * b->n);
  t.rows.assign(b->rows, b->rows + b->n);
  t.cols.assign(b->cols, b->cols + b->n);
  t.off.assign(b->offset, b->offset + b->n);
  for (int64_t i = 0; i < b->n; ++i) t.total = std::max(t.total, t.off[i] + t.rows[i] * t.cols[i]);
  return t;
}

HTab from_query(int64_t n, int ns, std::vector<uint64_t>& q, std::vector<uint32_t>& p, std::vector<int64_t>& r,
                std::vector<int64_t>& c) {
  HTab t;
  t.ns = ns;
  q.resize((size_t)2 * n);
  p.resize((size_t)ns * n);
  r.resize((size_t)n);
  c.resize((size_t)n);
  t.q = q;
  t.p = p;
  t.rows = r;
  t.cols = c;
  t.pack();
  return t;
}

// An MPO tensor owned on the host.
struct HMpo {
  std::vector<uint32_t> wl, wr, out, in;
  std::vector<double> val;
  uint32_t Dl = 0, Dr = 0;
  mpsa_mpo view() const {
    mpsa_mpo m;
    m.n = (int64_t)val.size();
    m.Dl = Dl;
    m.Dr = Dr;
    m.is_complex = 0;
    m.wl = wl.data();
    m.wr = wr.data();
    m.out = out.data();
    m.in = in.data();
    m.val = val.data();
    return m;
  }
};

HMpo copy_mpo(const mpsa_mpo* w) {
  if (!w || w->is_complex || w->n < 0 || (w->n > 0 && (!w->wl || !w->wr || !w->out || !w->in || !w->val)))
    throw AErr(MPSA_EINVAL, "chain: bad MPO tensor");
  HMpo m;
  m.Dl = w->Dl;
  m.Dr = w->Dr;
  m.wl.assign(w->wl, w->wl + w->n);
  m.wr.assign(w->wr, w->wr + w->n);
  m.out.assign(w->out, w->out + w->n);
  m.in.assign(w->in, w->in + w->n);
  m.val.assign(w->val, w->val + w->n);
  return m;
}

// The one-site output table of a plan for site i: A[i]'s blocks at A[i]'s own
// offsets, followed by every other produced block, so the first A.total
// elements of the output are in A[i]'s layout.
HTab heff1_y(const HTab& L, const HTab& R, const HTab& A, const HMpo& W, const std::vector<uint64_t>& itk) {
  mpsa_blocks bl = L.view(), br = R.view(), ba = A.view();
  mpsa_mpo mw = W.view();
  int64_t n = 0;
  chk(mpsa_heff1_outkeys(&bl, &br, &ba, &mw, itk.data(), (int)itk.size(), 0, &n, nullptr, nullptr, nullptr,
                         nullptr));
  std::vector<uint64_t> q((size_t)2 * n + 2);
  std::vector<uint32_t> p((size_t)n + 1);
  std::vector<int64_t> r((size_t)n + 1), c((size_t)n + 1);
  if (n)
    chk(mpsa_heff1_outkeys(&bl, &br, &ba, &mw, itk.data(), (int)itk.size(), n, &n, q.data(), p.data(), r.data(),
                           c.data()));
  // key -> (rows, cols, offset); A's blocks first
  struct K {
    uint64_t q0, q1;
    uint32_t p;
    bool operator<(const K& o) const {
      if (q0 != o.q0) return q0 < o.q0;
      if (q1 != o.q1) return q1 < o.q1;
      return p < o.p;
    }
  };
  std::map<K, std::array<int64_t, 3>> m;
  for (int64_t i = 0; i < A.n(); ++i) m[K{A.q[2 * i], A.q[2 * i + 1], A.p[i]}] = {A.rows[i], A.cols[i], A.off[i]};
  int64_t o = A.total;
  for (int64_t i = 0; i < n; ++i) {
    K k{q[2 * i], q[2 * i + 1], p[i]};
    if (m.count(k)) continue;
    o = align64(o);
    m[k] = {r[i], c[i], o};
    o += r[i] * c[i];
  }
  HTab t;
  t.ns = 1;
  for (auto& kv : m) {
    t.q.push_back(kv.first.q0);
    t.q.push_back(kv.first.q1);
    t.p.push_back(kv.first.p);
    t.rows.push_back(kv.second[0]);
    t.cols.push_back(kv.second[1]);
    t.off.push_back(kv.second[2]);
  }
  t.total = o;
  return t;
}

HTab env_out(const HTab& E, const HTab& A, const HMpo& W, const std::vector<uint64_t>& itk, int dir) {
  mpsa_blocks be = E.view(), ba = A.view();
  mpsa_mpo mw = W.view();
  int64_t n = 0;
  chk(mpsa_env_outkeys(&be, &ba, nullptr, &mw, itk.data(), (int)itk.size(), dir, 0, &n, nullptr, nullptr, nullptr,
                       nullptr));
  std::vector<uint64_t> q((size_t)2 * n + 2);
  std::vector<uint32_t> p((size_t)n + 1);
  std::vector<int64_t> r((size_t)n + 1), c((size_t)n + 1);
  if (n)
    chk(mpsa_env_outkeys(&be, &ba, nullptr, &mw, itk.data(), (int)itk.size(), dir, n, &n, q.data(), p.data(),
                         r.data(), c.data()));
  return from_query(n, 1, q, p, r, c);
}

struct DBuf {
  double* p = nullptr;
  int64_t n = 0;
  void alloc(int64_t len, hipStream_t st) {
    if (p && n >= len) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = std::max<int64_t>(len, 1);
    if (hipMalloc((void**)&p, sizeof(double) * (size_t)n) != hipSuccess) {
      p = nullptr;
      throw AErr(MPSA_ENOMEM, "chain: device allocation failed");
    }
    hchk(launch_fill(p, n, 0.0, st), "chain: zero");
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct mpsa_chain {
  int N = 0, nparam = 0, K = 8;
  std::vector<HTab> A;                      // site tables
  std::vector<HMpo> W;                      // MPO tensors
  std::vector<std::vector<HMpo>> G;         // G[p][i] = dW[i]/dtheta_p
  std::vector<std::vector<uint64_t>> itk;   // i_to_key per site
  std::vector<HTab> Lt;                     // Lt[i+1] = table of L[i]; Lt[0] = getLMin
  std::vector<HTab> Rt;                     // Rt[i] = table of R[i]; Rt[N] = getRMax
  std::vector<mpsa_plan*> EL, ER;           // env plans per site
  std::vector<mpsa_plan*> H1;               // one-site plans per site (W)
  std::vector<std::vector<mpsa_plan*>> HG;  // one-site plans per (param, site) (G), null when G is empty
  std::vector<HTab> Y1;                     // one-site output tables per site (W)
  std::vector<std::vector<HTab>> YG;        // per (param, site)
  std::map<int, DBuf> ckpt;                 // kept left environments, by index into Lt
  std::vector<DBuf> seg;                    // recomputed left environments of one segment
  DBuf rA, rB, y, dotpart;                  // right environment ping-pong, one-site output
  bool forward_done = false;
  ~mpsa_chain() {
    for (auto* v : {&EL, &ER, &H1})
      for (mpsa_plan* p : *v)
        if (p) mpsa_plan_destroy(p);
    for (auto& v : HG)
      for (mpsa_plan* p : v)
        if (p) mpsa_plan_destroy(p);
    for (auto& kv : ckpt) kv.second.free();
    for (auto& b : seg) b.free();
    rA.free();
    rB.free();
    y.free();
    dotpart.free();
  }
};

namespace {

HTab lmin_tab() {
  HTab t;
  t.q = {0, 0};
  t.p = {0};
  t.rows = {1};
  t.cols = {1};
  t.pack();
  return t;
}

// getRMax: identity per right charge of the last site tensor, aux 0.
HTab rmax_tab(const HTab& Alast) {
  std::map<uint64_t, int64_t> m;
  for (int64_t i = 0; i < Alast.n(); ++i) m[Alast.q[2 * i + 1]] = Alast.cols[i];
  HTab t;
  for (auto& kv : m) {
    t.q.push_back(kv.first);
    t.q.push_back(kv.first);
    t.p.push_back(0);
    t.rows.push_back(kv.second);
    t.cols.push_back(kv.second);
  }
  t.pack();
  return t;
}

// identity blocks of a boundary table (every block square)
void fill_identity(const HTab& t, DBuf& b, hipStream_t st) {
  b.alloc(t.total, st);
  hchk(launch_fill(b.p, b.n, 0.0, st), "identity: zero");
  for (int64_t k = 0; k < t.n(); ++k)
    hchk(launch_fill_strided(b.p + t.off[k], t.rows[k], t.rows[k] + 1, 1.0, st), "identity");
}

double dot_dev(const double* x, const double* y, int64_t n, hipStream_t st) {
  double out[2] = {0, 0};
  chk(mpsa_dot(x, y, n, 0, out, (mpsa_stream)st));
  return out[0];
}

}  // namespace

extern "C" {

int mpsa_chain_create(mpsa_chain** chain, int nsites, const mpsa_blocks* sites, const mpsa_mpo* W,
                      const uint64_t* itokey, const int* d, int nparam, const mpsa_mpo* dW, int checkpoint) {
  mpsa_chain* C = nullptr;
  try {
    if (!chain || nsites < 1 || !sites || !W || !itokey || !d || nparam < 0 || (nparam > 0 && !dW) || checkpoint < 1)
      throw AErr(MPSA_EINVAL, "chain_create: bad arguments");
    C = new mpsa_chain();
    C->N = nsites;
    C->nparam = nparam;
    C->K = checkpoint;
    const uint64_t* ik = itokey;
    for (int i = 0; i < nsites; ++i) {
      if (d[i] < 1) throw AErr(MPSA_EINVAL, "chain_create: d < 1");
      C->A.push_back(copy_tab(&sites[i], 1));
      C->W.push_back(copy_mpo(&W[i]));
      C->itk.emplace_back(ik, ik + d[i]);
      ik += d[i];
    }
    C->G.resize((size_t)nparam);
    for (int p = 0; p < nparam; ++p)
      for (int i = 0; i < nsites; ++i) C->G[(size_t)p].push_back(copy_mpo(&dW[(size_t)p * nsites + i]));
    // environment tables and plans
    C->Lt.push_back(lmin_tab());
    for (int i = 0; i < nsites; ++i) {
      HTab out = env_out(C->Lt.back(), C->A[(size_t)i], C->W[(size_t)i], C->itk[(size_t)i], 1);
      mpsa_blocks be = C->Lt.back().view(), ba = C->A[(size_t)i].view(), bo = out.view();
      mpsa_mpo mw = C->W[(size_t)i].view();
      mpsa_plan* pl = nullptr;
      chk(mpsa_env_plan(&pl, &be, &ba, nullptr, &mw, C->itk[(size_t)i].data(), (int)C->itk[(size_t)i].size(), 1, &bo,
                        MPSA_F64));
      C->EL.push_back(pl);
      C->Lt.push_back(out);
    }
    C->Rt.assign((size_t)nsites + 1, HTab());
    C->ER.assign((size_t)nsites, nullptr);
    C->Rt[(size_t)nsites] = rmax_tab(C->A[(size_t)nsites - 1]);
    for (int i = nsites - 1; i >= 0; --i) {
      HTab out = env_out(C->Rt[(size_t)i + 1], C->A[(size_t)i], C->W[(size_t)i], C->itk[(size_t)i], 0);
      mpsa_blocks be = C->Rt[(size_t)i + 1].view(), ba = C->A[(size_t)i].view(), bo = out.view();
      mpsa_mpo mw = C->W[(size_t)i].view();
      mpsa_plan* pl = nullptr;
      chk(mpsa_env_plan(&pl, &be, &ba, nullptr, &mw, C->itk[(size_t)i].data(), (int)C->itk[(size_t)i].size(), 0, &bo,
                        MPSA_F64));
      C->ER[(size_t)i] = pl;
      C->Rt[(size_t)i] = out;
    }
    // one-site plans: L[i-1] = Lt[i], R[i+1] = Rt[i+1]
    auto h1 = [&](int i, const HMpo& w, HTab& yt) -> mpsa_plan* {
      if (w.val.empty()) return nullptr;
      yt = heff1_y(C->Lt[(size_t)i], C->Rt[(size_t)i + 1], C->A[(size_t)i], w, C->itk[(size_t)i]);
      mpsa_blocks bl = C->Lt[(size_t)i].view(), br = C->Rt[(size_t)i + 1].view(), bx = C->A[(size_t)i].view(),
                  by = yt.view();
      mpsa_mpo mw = w.view();
      mpsa_plan* pl = nullptr;
      chk(mpsa_heff1_plan(&pl, &bl, &br, &bx, &by, &mw, C->itk[(size_t)i].data(), (int)C->itk[(size_t)i].size(),
                          MPSA_F64));
      return pl;
    };
    C->Y1.resize((size_t)nsites);
    for (int i = 0; i < nsites; ++i) C->H1.push_back(h1(i, C->W[(size_t)i], C->Y1[(size_t)i]));
    C->HG.assign((size_t)nparam, std::vector<mpsa_plan*>((size_t)nsites, nullptr));
    C->YG.assign((size_t)nparam, std::vector<HTab>((size_t)nsites));
    for (int p = 0; p < nparam; ++p)
      for (int i = 0; i < nsites; ++i)
        C->HG[(size_t)p][(size_t)i] = h1(i, C->G[(size_t)p][(size_t)i], C->YG[(size_t)p][(size_t)i]);
    *chain = C;
    return MPSA_OK;
  } catch (const AErr& e) {
    delete C;
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    delete C;
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_chain_destroy(mpsa_chain* chain) {
  delete chain;
  return MPSA_OK;
}

int mpsa_chain_env_table(const mpsa_chain* C, int side, int index, int64_t cap, int64_t* n, uint64_t* q, uint32_t* p,
                         int64_t* rows, int64_t* cols, int64_t* offsets) {
  try {
    if (!C || !n) throw AErr(MPSA_EINVAL, "chain_env_table: bad arguments");
    const HTab* t = nullptr;
    if (side > 0) {  // L[index], index in [-1, N-1]
      if (index < -1 || index >= C->N) throw AErr(MPSA_EINVAL, "chain_env_table: index out of range");
      t = &C->Lt[(size_t)index + 1];
    } else {  // R[index], index in [0, N]
      if (index < 0 || index > C->N) throw AErr(MPSA_EINVAL, "chain_env_table: index out of range");
      t = &C->Rt[(size_t)index];
    }
    *n = t->n();
    if (cap <= 0) return MPSA_OK;
    if (cap < t->n() || !q || !p || !rows || !cols || !offsets)
      throw AErr(MPSA_EINVAL, "chain_env_table: arrays too small");
    std::memcpy(q, t->q.data(), sizeof(uint64_t) * t->q.size());
    std::memcpy(p, t->p.data(), sizeof(uint32_t) * t->p.size());
    std::memcpy(rows, t->rows.data(), sizeof(int64_t) * t->rows.size());
    std::memcpy(cols, t->cols.data(), sizeof(int64_t) * t->cols.size());
    std::memcpy(offsets, t->off.data(), sizeof(int64_t) * t->off.size());
    return MPSA_OK;
  } catch (const AErr& e) {
    set_last_error(e.what());
    return e.code;
  }
}

int mpsa_energy_forward(mpsa_chain* C, const double* const* dA, double* E, mpsa_stream stream) {
  try {
    if (!C || !dA || !E) throw AErr(MPSA_EINVAL, "energy_forward: bad arguments");
    for (int i = 0; i < C->N; ++i)
      if (!dA[i] && C->A[(size_t)i].total > 0) throw AErr(MPSA_EINVAL, "energy_forward: null site buffer");
    hipStream_t st = (hipStream_t)stream;
    C->forward_done = false;
    // L[-1] = getLMin is checkpoint 0 (index into Lt); ping-pong the others
    fill_identity(C->Lt[0], C->ckpt[0], st);
    DBuf a, b;
    const double* cur = C->ckpt[0].p;
    try {
      for (int i = 0; i < C->N; ++i) {
        const int li = i + 1;  // index of L[i] in Lt
        const bool keep = (li % C->K == 0) || li == C->N;
        double* out;
        if (keep) {
          C->ckpt[li].alloc(C->Lt[(size_t)li].total, st);
          out = C->ckpt[li].p;
        } else {
          DBuf& t = (cur == a.p) ? b : a;
          t.alloc(C->Lt[(size_t)li].total, st);
          out = t.p;
        }
        chk(mpsa_env_apply(C->EL[(size_t)i], cur, dA[i], nullptr, out, stream));
        cur = out;
      }
      hchk(hipStreamSynchronize(st), "energy_forward sync");
    } catch (...) {
      a.free();
      b.free();
      throw;
    }
    a.free();
    b.free();
    // trace of the charge-diagonal blocks of L[N-1]
    const HTab& t = C->Lt[(size_t)C->N];
    std::vector<double> h((size_t)std::max<int64_t>(t.total, 1));
    hchk(hipMemcpy(h.data(), C->ckpt[C->N].p, sizeof(double) * (size_t)t.total, hipMemcpyDeviceToHost), "trace");
    double e = 0.0;
    for (int64_t k = 0; k < t.n(); ++k)
      if (t.q[2 * k] == t.q[2 * k + 1]) {
        const int64_t m = std::min(t.rows[k], t.cols[k]);
        for (int64_t j = 0; j < m; ++j) e += h[(size_t)(t.off[k] + j * t.rows[k] + j)];
      }
    *E = e;
    C->forward_done = true;
    return MPSA_OK;
  } catch (const AErr& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_energy_backward(mpsa_chain* C, const double* const* dA, double* const* dgrad, double* dE_dtheta,
                         mpsa_stream stream) {
  try {
    if (!C || !dA || !dgrad || (C->nparam > 0 && !dE_dtheta)) throw AErr(MPSA_EINVAL, "energy_backward: bad arguments");
    if (!C->forward_done) throw AErr(MPSA_EINVAL, "energy_backward: run energy_forward on the same tensors first");
    hipStream_t st = (hipStream_t)stream;
    const int N = C->N, K = C->K;
    for (int p = 0; p < C->nparam; ++p) dE_dtheta[p] = 0.0;
    // right boundary R[N]
    DBuf* rcur = &C->rA;
    DBuf* rnext = &C->rB;
    fill_identity(C->Rt[(size_t)N], *rcur, st);
    if ((int)C->seg.size() < K) C->seg.resize((size_t)K);
    // segments of Lt indices [s0, s0 + K) with s0 a checkpoint; site i needs Lt[i]
    int seg0 = -1;
    for (int i = N - 1; i >= 0; --i) {
      const int li = i;  // L[i-1] = Lt[i]
      const int s0 = (li / K) * K;
      if (s0 != seg0) {  // recompute Lt[s0 .. li] from checkpoint s0
        seg0 = s0;
        if (!C->ckpt.count(s0)) throw AErr(MPSA_EINVAL, "energy_backward: missing checkpoint");
        const double* prev = C->ckpt[s0].p;
        for (int j = s0 + 1; j <= li; ++j) {
          DBuf& b = C->seg[(size_t)(j - s0)];
          b.alloc(C->Lt[(size_t)j].total, st);
          chk(mpsa_env_apply(C->EL[(size_t)j - 1], prev, dA[j - 1], nullptr, b.p, stream));
          prev = b.p;
        }
      }
      const double* Lbuf = li == s0 ? C->ckpt[s0].p : C->seg[(size_t)(li - s0)].p;
      const HTab& At = C->A[(size_t)i];
      // dE/dA[i] = 2 (H_eff A[i]) on A[i]'s blocks
      if (C->H1[(size_t)i]) {
        C->y.alloc(C->Y1[(size_t)i].total, st);
        hchk(launch_fill(C->y.p, C->y.n, 0.0, st), "zero output");
        chk(mpsa_plan_bind_env(C->H1[(size_t)i], Lbuf, rcur->p, stream));
        chk(mpsa_heff_apply(C->H1[(size_t)i], dA[i], 0, C->y.p, 0, 1, stream));
        const double two[2] = {2.0, 0.0}, zero[2] = {0.0, 0.0};
        if (At.total > 0) chk(mpsa_axpby(two, C->y.p, zero, dgrad[i], At.total, 0, stream));
      } else if (At.total > 0) {
        hchk(launch_fill(dgrad[i], At.total, 0.0, st), "zero gradient");
      }
      for (int p = 0; p < C->nparam; ++p) {
        mpsa_plan* pl = C->HG[(size_t)p][(size_t)i];
        if (!pl) continue;
        C->y.alloc(C->YG[(size_t)p][(size_t)i].total, st);
        hchk(launch_fill(C->y.p, C->y.n, 0.0, st), "zero output");
        chk(mpsa_plan_bind_env(pl, Lbuf, rcur->p, stream));
        chk(mpsa_heff_apply(pl, dA[i], 0, C->y.p, 0, 1, stream));
        if (At.total > 0) dE_dtheta[p] += dot_dev(C->y.p, dA[i], At.total, st);
      }
      // R[i] = addContractionLayer(R[i+1], A[i], W[i], 0) for the next site
      if (i > 0) {
        rnext->alloc(C->Rt[(size_t)i].total, st);
        chk(mpsa_env_apply(C->ER[(size_t)i], rcur->p, dA[i], nullptr, rnext->p, stream));
        std::swap(rcur, rnext);
      }
    }
    hchk(hipStreamSynchronize(st), "energy_backward sync");
    return MPSA_OK;
  } catch (const AErr& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

}  // extern "C"
