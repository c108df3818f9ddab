// mpsa_host.h -- host-side tables, MPO tensors and device buffers shared by
// the chain-level drivers (energy/adjoint, DMRG sweep) built on the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

namespace mpsa {
namespace host {

struct HErr : std::runtime_error {
  int code;
  HErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void chk(int rc) {
  if (rc != 0) throw HErr(rc, mpsa_last_error());
}
inline void hchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HErr(MPSA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

inline int64_t align64(int64_t x) { return (x + 63) / 64 * 64; }

// A key table owned on the host.
struct HTab {
  int ns = 1;
  int cplx = 0;  // complex128, planar: real parts, then imaginary parts at + total
  std::vector<uint64_t> q;
  std::vector<uint32_t> p;
  std::vector<int64_t> rows, cols, off;
  int64_t total = 0;
  int64_t n() const { return (int64_t)rows.size(); }
  int64_t len() const { return total * (cplx ? 2 : 1); }  // doubles of a packed buffer
  mpsa_blocks view() const {
    mpsa_blocks b;
    b.n = n();
    b.ns = ns;
    b.is_complex = cplx;
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

inline HTab copy_tab(const mpsa_blocks* b, int ns) {
  if (!b || b->ns != ns || b->is_complex || b->n < 0) throw HErr(MPSA_EINVAL, "chain: bad site table");
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

inline HTab from_query(int64_t n, int ns, std::vector<uint64_t>& q, std::vector<uint32_t>& p, std::vector<int64_t>& r,
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

inline HMpo copy_mpo(const mpsa_mpo* w) {
  if (!w || w->is_complex || w->n < 0 || (w->n > 0 && (!w->wl || !w->wr || !w->out || !w->in || !w->val)))
    throw HErr(MPSA_EINVAL, "chain: bad MPO tensor");
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
inline HTab heff1_y(const HTab& L, const HTab& R, const HTab& A, const HMpo& W, const std::vector<uint64_t>& itk) {
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

inline HTab env_out(const HTab& E, const HTab& A, const HMpo& W, const std::vector<uint64_t>& itk, int dir) {
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
      throw HErr(MPSA_ENOMEM, "chain: device allocation failed");
    }
    hchk(launch_fill(p, n, 0.0, st), "chain: zero");
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};


inline HTab lmin_tab() {
  HTab t;
  t.q = {0, 0};
  t.p = {0};
  t.rows = {1};
  t.cols = {1};
  t.pack();
  return t;
}

// getRMax: identity per right charge of the last site tensor, aux 0.
inline HTab rmax_tab(const HTab& Alast) {
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
inline void fill_identity(const HTab& t, DBuf& b, hipStream_t st) {
  b.alloc(t.total, st);
  hchk(launch_fill(b.p, b.n, 0.0, st), "identity: zero");
  for (int64_t k = 0; k < t.n(); ++k)
    hchk(launch_fill_strided(b.p + t.off[k], t.rows[k], t.rows[k] + 1, 1.0, st), "identity");
}


}  // namespace host
}  // namespace mpsa
