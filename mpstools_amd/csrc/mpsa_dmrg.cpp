// mpsa_dmrg.cpp -- device-resident two-site DMRG sweep:
//   DMRGengine::doStep / simulate (src/container.hpp:293-337, 383-438).
//
// Per bond (site, site+1), as doStep does it:
//   psi   = BlockMatrix<2>(A[site], A[site+1])      (src/BlockTensor.hpp:285-301)
//   Lanczos on H_eff(L[site-1], W[site], W[site+1], R[site+2]) from psi, the
//           ground state by getState(0)              (src/lanzcos.hpp:165-278)
//   split  with the truncation of the caller's choice (src/MPSfunctions.hpp:1142)
//   lambda into A[site+1] (dir > 0) or A[site] (dir <= 0)   (container.hpp:321-328)
//   L[site] = addContractionLayer(L[site-1], A[site], W[site], +1)       (dir > 0)
//   R[site+1] = addContractionLayer(R[site+2], A[site+1], W[site+1], -1) (dir <= 0)
// A sweep is simulate's loop body: doStep(site, +1) for site = 0 .. N-3, then
// doStep(site, -1) for site = N-2 .. 1.  Site tensors and every environment
// stay on the device; plans are cached by the shapes they were built for, so
// the bulk of a dense chain reuses one H_eff plan and one plan per direction.
//
// Precision: the Lanczos matvecs run at the requested precision; for a
// reduced precision the Ritz vector is refined by a short FP64 Lanczos
// (refine_it matvecs at most) before the split.  Everything else is FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <string>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_host.h"
#include "mpsa_internal.h"

using namespace mpsa;
using namespace mpsa::host;

namespace {

// Byte string of a table's keys and shapes (plan cache key part).
std::string sig(const HTab& t) {
  std::string s;
  s.append((const char*)t.q.data(), t.q.size() * sizeof(uint64_t));
  s.append((const char*)t.p.data(), t.p.size() * sizeof(uint32_t));
  s.append((const char*)t.rows.data(), t.rows.size() * sizeof(int64_t));
  s.append((const char*)t.cols.data(), t.cols.size() * sizeof(int64_t));
  s.append((const char*)t.off.data(), t.off.size() * sizeof(int64_t));
  s.push_back('|');
  return s;
}
std::string sig(const HMpo& w) {
  std::string s;
  s.append((const char*)&w.Dl, sizeof(w.Dl));
  s.append((const char*)&w.Dr, sizeof(w.Dr));
  s.append((const char*)w.wl.data(), w.wl.size() * 4);
  s.append((const char*)w.wr.data(), w.wr.size() * 4);
  s.append((const char*)w.out.data(), w.out.size() * 4);
  s.append((const char*)w.in.data(), w.in.size() * 4);
  s.append((const char*)w.val.data(), w.val.size() * 8);
  s.push_back('|');
  return s;
}
std::string sig(const std::vector<uint64_t>& v) {
  return std::string((const char*)v.data(), v.size() * 8) + "|";
}

// A small LRU cache of plans (a plan holds device workspace).
struct PlanCache {
  size_t cap;
  std::list<std::pair<std::string, mpsa_plan*>> items;
  explicit PlanCache(size_t c) : cap(c) {}
  mpsa_plan* find(const std::string& k) {
    for (auto it = items.begin(); it != items.end(); ++it)
      if (it->first == k) {
        items.splice(items.begin(), items, it);
        return it->second;
      }
    return nullptr;
  }
  void put(const std::string& k, mpsa_plan* p) {
    items.emplace_front(k, p);
    while (items.size() > cap) {
      mpsa_plan_destroy(items.back().second);
      items.pop_back();
    }
  }
  ~PlanCache() {
    for (auto& kv : items) mpsa_plan_destroy(kv.second);
  }
};

// Two-site tensor table: keys (l, r, (s1, s2)) of every product of an A1 block
// (l, m, s1) with an A2 block (m, r, s2), in key order.
struct TwoSite {
  HTab t;                                               // ns = 2
  std::vector<std::array<int64_t, 3>> prod;             // (block of t, block of A1, block of A2)
};

TwoSite twosite_table(const HTab& A1, const HTab& A2) {
  std::map<std::array<uint64_t, 4>, std::array<int64_t, 4>> m;  // key -> rows, cols, i1, i2
  for (int64_t i = 0; i < A1.n(); ++i)
    for (int64_t j = 0; j < A2.n(); ++j)
      if (A1.q[2 * i + 1] == A2.q[2 * j]) {
        std::array<uint64_t, 4> k{A1.q[2 * i], A2.q[2 * j + 1], A1.p[i], A2.p[j]};
        if (m.count(k)) throw HErr(MPSA_EINVAL, "dmrg: two middle charges for one two-site key");
        m[k] = {A1.rows[i], A2.cols[j], i, j};
      }
  TwoSite ts;
  ts.t.ns = 2;
  for (auto& kv : m) {
    ts.t.q.push_back(kv.first[0]);
    ts.t.q.push_back(kv.first[1]);
    ts.t.p.push_back((uint32_t)kv.first[2]);
    ts.t.p.push_back((uint32_t)kv.first[3]);
    ts.t.rows.push_back(kv.second[0]);
    ts.t.cols.push_back(kv.second[1]);
    ts.prod.push_back({(int64_t)ts.prod.size(), kv.second[2], kv.second[3]});
  }
  ts.t.pack();
  return ts;
}

// Union of two ns = 2 tables (key order, fresh packing).
HTab union_tab(const HTab& a, const HTab& b) {
  std::map<std::array<uint64_t, 4>, std::pair<int64_t, int64_t>> m;
  for (const HTab* t : {&a, &b})
    for (int64_t i = 0; i < t->n(); ++i)
      m[{t->q[2 * i], t->q[2 * i + 1], t->p[2 * i], t->p[2 * i + 1]}] = {t->rows[i], t->cols[i]};
  HTab u;
  u.ns = 2;
  for (auto& kv : m) {
    u.q.push_back(kv.first[0]);
    u.q.push_back(kv.first[1]);
    u.p.push_back((uint32_t)kv.first[2]);
    u.p.push_back((uint32_t)kv.first[3]);
    u.rows.push_back(kv.second.first);
    u.cols.push_back(kv.second.second);
  }
  u.pack();
  return u;
}

int find_block(const HTab& t, uint64_t q0, uint64_t q1, uint32_t p0, uint32_t p1) {
  for (int64_t i = 0; i < t.n(); ++i)
    if (t.q[2 * i] == q0 && t.q[2 * i + 1] == q1 && t.p[(size_t)t.ns * i] == p0 &&
        (t.ns == 1 || t.p[(size_t)t.ns * i + 1] == p1))
      return (int)i;
  return -1;
}

}  // namespace

struct mpsa_dmrg {
  int N = 0;
  hipStream_t st = nullptr;
  std::vector<HTab> A;
  std::vector<DBuf> dA;
  std::vector<HMpo> W;
  std::vector<std::vector<uint64_t>> itk;
  std::vector<HTab> Lt;  // Lt[i + 1] = table of L[i]; Lt[0] = getLMin
  std::vector<DBuf> Lb;
  std::vector<HTab> Rt;  // Rt[i] = table of R[i]; Rt[N] = getRMax
  std::vector<DBuf> Rb;
  PlanCache env{16}, heff{6};
  DBuf psi, state, refined;
  // device task lists of the two-site product
  GTask* d_tasks = nullptr;
  GSeg* d_segs = nullptr;
  size_t cap_tasks = 0;
  ~mpsa_dmrg() {
    for (auto& b : dA) b.free();
    for (auto& b : Lb) b.free();
    for (auto& b : Rb) b.free();
    psi.free();
    state.free();
    refined.free();
    if (d_tasks) (void)hipFree(d_tasks);
    if (d_segs) (void)hipFree(d_segs);
  }
};

namespace {

// E' = addContractionLayer(E, A[s], W[s], dir) on the device, plan cached.
void env_step(mpsa_dmrg* D, const HTab& Et, const double* dE, int s, int dir, HTab& out_t, DBuf& out_b) {
  out_t = env_out(Et, D->A[(size_t)s], D->W[(size_t)s], D->itk[(size_t)s], dir);
  const std::string key = std::string(dir > 0 ? "L" : "R") + sig(Et) + sig(D->A[(size_t)s]) + sig(D->W[(size_t)s]) +
                          sig(D->itk[(size_t)s]) + sig(out_t);
  mpsa_plan* pl = D->env.find(key);
  if (!pl) {
    mpsa_blocks be = Et.view(), ba = D->A[(size_t)s].view(), bo = out_t.view();
    mpsa_mpo mw = D->W[(size_t)s].view();
    chk(mpsa_env_plan(&pl, &be, &ba, nullptr, &mw, D->itk[(size_t)s].data(), (int)D->itk[(size_t)s].size(), dir, &bo,
                      MPSA_F64));
    D->env.put(key, pl);
  }
  out_b.alloc(out_t.total, D->st);
  chk(mpsa_env_apply(pl, dE, D->dA[(size_t)s].p, nullptr, out_b.p, (mpsa_stream)D->st));
}

// psi (in the layout of vt) = BlockMatrix<2>(A[s], A[s+1]), FP64 grouped GEMM.
void twosite_product(mpsa_dmrg* D, int s, const TwoSite& ts, const HTab& vt, DBuf& out) {
  const HTab &A1 = D->A[(size_t)s], &A2 = D->A[(size_t)s + 1];
  out.alloc(vt.total, D->st);
  hchk(launch_fill(out.p, out.n, 0.0, D->st), "dmrg: zero psi");
  std::vector<GTask> tasks;
  std::vector<GSeg> segs;
  int tile = 0;
  const int BM = 128, BN = 128;  // FP64 tile class 0
  for (auto& pr : ts.prod) {
    const int64_t i = pr[1], j = pr[2];
    const int b = find_block(vt, A1.q[2 * i], A2.q[2 * j + 1], A1.p[i], A2.p[j]);
    if (b < 0) throw HErr(MPSA_EINVAL, "dmrg: two-site key missing from the vector space");
    GTask t{};
    t.c_off = vt.off[(size_t)b];
    t.m = (int32_t)A1.rows[i];
    t.n = (int32_t)A2.cols[j];
    if (t.m == 0 || t.n == 0) continue;
    t.seg0 = (int32_t)segs.size();
    t.nseg = A1.cols[i] > 0 ? 1 : 0;
    t.tiles_m = (t.m + BM - 1) / BM;
    t.tile0 = tile;
    tile += t.tiles_m * ((t.n + BN - 1) / BN);
    if (t.nseg) {
      GSeg g{};
      g.a_off = A1.off[i];
      g.b_off = A2.off[j];
      g.k = (int32_t)A1.cols[i];
      g.lda = (int32_t)std::max<int64_t>(A1.rows[i], 1);
      g.ldb = (int32_t)std::max<int64_t>(A2.rows[j], 1);
      g.alpha = 1.0;
      segs.push_back(g);
    }
    tasks.push_back(t);
  }
  if (tasks.empty()) return;
  if (tasks.size() > D->cap_tasks || segs.size() > D->cap_tasks) {
    if (D->d_tasks) (void)hipFree(D->d_tasks);
    if (D->d_segs) (void)hipFree(D->d_segs);
    D->d_tasks = nullptr;
    D->d_segs = nullptr;
    D->cap_tasks = std::max(tasks.size(), segs.size()) * 2;
    hchk(hipMalloc((void**)&D->d_tasks, sizeof(GTask) * D->cap_tasks), "dmrg: tasks");
    hchk(hipMalloc((void**)&D->d_segs, sizeof(GSeg) * D->cap_tasks), "dmrg: segs");
  }
  hchk(hipMemcpy(D->d_tasks, tasks.data(), sizeof(GTask) * tasks.size(), hipMemcpyHostToDevice), "dmrg: tasks");
  if (!segs.empty())
    hchk(hipMemcpy(D->d_segs, segs.data(), sizeof(GSeg) * segs.size(), hipMemcpyHostToDevice), "dmrg: segs");
  hchk(launch_gemm(F64, 0, 0, 0, D->d_tasks, (int)tasks.size(), D->d_segs, tile, 1, D->dA[(size_t)s].p, 0,
                   D->dA[(size_t)s + 1].p, 0, out.p, 0, D->st),
       "dmrg: two-site product");
}

mpsa_plan* heff_plan(mpsa_dmrg* D, int s, const HTab& vt, int prec) {
  const HTab &Lt = D->Lt[(size_t)s], &Rt = D->Rt[(size_t)s + 2];
  const std::string key = std::to_string(prec) + "|" + sig(Lt) + sig(Rt) + sig(vt) + sig(D->W[(size_t)s]) +
                          sig(D->W[(size_t)s + 1]) + sig(D->itk[(size_t)s]) + sig(D->itk[(size_t)s + 1]);
  mpsa_plan* pl = D->heff.find(key);
  if (!pl) {
    mpsa_blocks bl = Lt.view(), br = Rt.view(), bx = vt.view();
    mpsa_mpo ml = D->W[(size_t)s].view(), mr = D->W[(size_t)s + 1].view();
    chk(mpsa_heff2_plan(&pl, &bl, &br, &bx, &bx, &ml, &mr, D->itk[(size_t)s].data(), (int)D->itk[(size_t)s].size(),
                        D->itk[(size_t)s + 1].data(), (int)D->itk[(size_t)s + 1].size(), prec));
    D->heff.put(key, pl);
  }
  chk(mpsa_plan_bind_env(pl, D->Lb[(size_t)s].p, D->Rb[(size_t)s + 2].p, (mpsa_stream)D->st));
  return pl;
}

// Output table of H_eff on the two-site space.
HTab heff_out(mpsa_dmrg* D, int s, const HTab& xt) {
  mpsa_blocks bl = D->Lt[(size_t)s].view(), br = D->Rt[(size_t)s + 2].view(), bx = xt.view();
  mpsa_mpo ml = D->W[(size_t)s].view(), mr = D->W[(size_t)s + 1].view();
  int64_t n = 0;
  chk(mpsa_heff2_outkeys(&bl, &br, &bx, &ml, &mr, D->itk[(size_t)s].data(), (int)D->itk[(size_t)s].size(),
                         D->itk[(size_t)s + 1].data(), (int)D->itk[(size_t)s + 1].size(), 0, &n, nullptr, nullptr,
                         nullptr, nullptr));
  HTab y;
  y.ns = 2;
  y.q.resize((size_t)2 * n + 2);
  y.p.resize((size_t)2 * n + 2);
  y.rows.resize((size_t)n + 1);
  y.cols.resize((size_t)n + 1);
  if (n)
    chk(mpsa_heff2_outkeys(&bl, &br, &bx, &ml, &mr, D->itk[(size_t)s].data(), (int)D->itk[(size_t)s].size(),
                           D->itk[(size_t)s + 1].data(), (int)D->itk[(size_t)s + 1].size(), n, &n, y.q.data(),
                           y.p.data(), y.rows.data(), y.cols.data()));
  y.q.resize((size_t)2 * n);
  y.p.resize((size_t)2 * n);
  y.rows.resize((size_t)n);
  y.cols.resize((size_t)n);
  y.pack();
  return y;
}

struct StepOut {
  double E = 0, tw = 0;
  int it = 0;
};

StepOut do_step(mpsa_dmrg* D, int s, int dir, int prec, int refine_it, unsigned chimax, int do_trunc, double tr_weight,
                int lanmax, double delta) {
  StepOut o;
  TwoSite ts = twosite_table(D->A[(size_t)s], D->A[(size_t)s + 1]);
  const HTab vt = union_tab(ts.t, heff_out(D, s, ts.t));
  twosite_product(D, s, ts, vt, D->psi);
  D->state.alloc(vt.total, D->st);
  std::vector<double> E((size_t)lanmax + 2);
  int nit = 0, nr = 0;
  mpsa_plan* pl = heff_plan(D, s, vt, prec);
  chk(mpsa_lanczos(pl, D->psi.p, D->state.p, lanmax, delta, 1, 0, E.data(), (int)E.size(), &nit, &nr,
                   (mpsa_stream)D->st));
  o.E = E[0];
  o.it = nit;
  const double* gs = D->state.p;
  if (prec != MPSA_F64 && refine_it > 0) {
    mpsa_plan* p64 = heff_plan(D, s, vt, MPSA_F64);
    D->refined.alloc(vt.total, D->st);
    chk(mpsa_lanczos(p64, D->state.p, D->refined.p, refine_it, delta, 1, 0, E.data(), (int)E.size(), &nit, &nr,
                     (mpsa_stream)D->st));
    o.E = E[0];
    o.it += nit;
    gs = D->refined.p;
  }
  // split, lambda into the site ahead of the sweep
  mpsa_blocks bv = vt.view();
  mpsa_split_result* r = nullptr;
  chk(mpsa_split(&bv, gs, D->itk[(size_t)s].data(), (int)D->itk[(size_t)s].size(), D->itk[(size_t)s + 1].data(),
                 (int)D->itk[(size_t)s + 1].size(), dir > 0 ? 1 : 0, chimax, do_trunc, tr_weight, 1, &r,
                 (mpsa_stream)D->st));
  try {
    int chi = 0;
    int64_t nl = 0, nrgt = 0, nlam = 0, ll = 0, rl = 0;
    chk(mpsa_split_info(r, &o.tw, &chi, &nl, &nrgt, &nlam, &ll, &rl));
    for (int w = 0; w < 2; ++w) {
      const int64_t nb = w == 0 ? nl : nrgt, len = w == 0 ? ll : rl;
      HTab t;
      t.ns = 1;
      t.q.resize((size_t)2 * nb);
      t.p.resize((size_t)nb);
      t.rows.resize((size_t)nb);
      t.cols.resize((size_t)nb);
      t.off.resize((size_t)nb);
      if (nb) chk(mpsa_split_table(r, w, t.q.data(), t.p.data(), t.rows.data(), t.cols.data(), t.off.data()));
      t.total = len;
      DBuf& b = D->dA[(size_t)s + w];
      b.alloc(len, D->st);
      if (len) chk(mpsa_split_copy(r, w, b.p, (mpsa_stream)D->st));
      D->A[(size_t)s + w] = t;
    }
  } catch (...) {
    mpsa_split_destroy(r);
    throw;
  }
  mpsa_split_destroy(r);
  // environment of the sweep direction
  if (dir > 0) env_step(D, D->Lt[(size_t)s], D->Lb[(size_t)s].p, s, 1, D->Lt[(size_t)s + 1], D->Lb[(size_t)s + 1]);
  else env_step(D, D->Rt[(size_t)s + 2], D->Rb[(size_t)s + 2].p, s + 1, 0, D->Rt[(size_t)s + 1], D->Rb[(size_t)s + 1]);
  return o;
}

}  // namespace

extern "C" {

int mpsa_dmrg_create(mpsa_dmrg** out, int nsites, const mpsa_blocks* sites, const double* const* dA, const mpsa_mpo* W,
                     const uint64_t* itokey, const int* d, mpsa_stream stream) {
  mpsa_dmrg* D = nullptr;
  try {
    if (!out || nsites < 3 || !sites || !dA || !W || !itokey || !d) throw HErr(MPSA_EINVAL, "dmrg_create: bad arguments");
#ifndef MPSA_DMRG_SWEEP
    // Switched off together with mpsa_dmrg_sweep (see there): create builds
    // the right environments through the same plan cache and buffer code.
    throw HErr(MPSA_EINVAL, "dmrg_create: disabled in this build (unexplained GPU memory fault under investigation)");
#endif
    D = new mpsa_dmrg();
    D->N = nsites;
    D->st = (hipStream_t)stream;
    const uint64_t* ik = itokey;
    D->dA.resize((size_t)nsites);
    for (int i = 0; i < nsites; ++i) {
      if (d[i] < 1) throw HErr(MPSA_EINVAL, "dmrg_create: d < 1");
      D->A.push_back(copy_tab(&sites[i], 1));
      D->W.push_back(copy_mpo(&W[i]));
      D->itk.emplace_back(ik, ik + d[i]);
      ik += d[i];
      const int64_t len = D->A.back().total;
      D->dA[(size_t)i].alloc(len, D->st);
      if (len) {
        if (!dA[i]) throw HErr(MPSA_EINVAL, "dmrg_create: null site buffer");
        hchk(hipMemcpyAsync(D->dA[(size_t)i].p, dA[i], sizeof(double) * (size_t)len, hipMemcpyDeviceToDevice, D->st),
             "dmrg: copy site");
      }
    }
    // environments: L[-1] = getLMin, R[N] = getRMax, R[i] for i = N-1 .. 2
    D->Lt.assign((size_t)nsites + 1, HTab());
    D->Lb.resize((size_t)nsites + 1);
    D->Rt.assign((size_t)nsites + 1, HTab());
    D->Rb.resize((size_t)nsites + 1);
    D->Lt[0] = lmin_tab();
    fill_identity(D->Lt[0], D->Lb[0], D->st);
    D->Rt[(size_t)nsites] = rmax_tab(D->A[(size_t)nsites - 1]);
    fill_identity(D->Rt[(size_t)nsites], D->Rb[(size_t)nsites], D->st);
    for (int i = nsites - 1; i >= 2; --i)
      env_step(D, D->Rt[(size_t)i + 1], D->Rb[(size_t)i + 1].p, i, 0, D->Rt[(size_t)i], D->Rb[(size_t)i]);
    hchk(hipStreamSynchronize(D->st), "dmrg_create sync");
    *out = D;
    return MPSA_OK;
  } catch (const HErr& e) {
    delete D;
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    delete D;
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_dmrg_sweep(mpsa_dmrg* D, int precision, int refine_it, unsigned chimax, int do_trunc, double tr_weight,
                    int lanmax, double delta, double* energy, double* max_truncated_weight, int* lanczos_iterations,
                    double* seconds) {
  try {
    if (!D || lanmax < 1) throw HErr(MPSA_EINVAL, "dmrg_sweep: bad arguments");
#ifndef MPSA_DMRG_SWEEP
    // Switched off in the gfx950 library: the device sweep faulted the GPU on
    // one test chain (XXZ L = 8 with a field) after passing another, and the
    // cause is not found.  This host logic runs clean, bounds-checked and
    // under AddressSanitizer, on the CPU emulation of the device layer used by
    // the tests (tests/emu), which is the only build that defines
    // MPSA_DMRG_SWEEP.
    throw HErr(MPSA_EINVAL, "dmrg_sweep: disabled in this build (unexplained GPU memory fault under investigation)");
#endif
    const auto t0 = std::chrono::steady_clock::now();
    StepOut o;
    double twmax = 0.0;
    int its = 0;
    for (int site = 0; site < D->N - 2; ++site) {
      o = do_step(D, site, 1, precision, refine_it, chimax, do_trunc, tr_weight, lanmax, delta);
      twmax = std::max(twmax, o.tw);
      its += o.it;
    }
    for (int site = D->N - 2; site > 0; --site) {
      o = do_step(D, site, -1, precision, refine_it, chimax, do_trunc, tr_weight, lanmax, delta);
      twmax = std::max(twmax, o.tw);
      its += o.it;
    }
    hchk(hipStreamSynchronize(D->st), "dmrg_sweep sync");
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (energy) *energy = o.E;
    if (max_truncated_weight) *max_truncated_weight = twmax;
    if (lanczos_iterations) *lanczos_iterations = its;
    if (seconds) *seconds = dt;
    return MPSA_OK;
  } catch (const HErr& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_dmrg_site(const mpsa_dmrg* D, int site, int64_t* n_blocks, int64_t* len, uint64_t* q, uint32_t* p,
                   int64_t* rows, int64_t* cols, int64_t* offsets, double* dst, mpsa_stream stream) {
  try {
    if (!D || site < 0 || site >= D->N || !n_blocks || !len) throw HErr(MPSA_EINVAL, "dmrg_site: bad arguments");
    const HTab& t = D->A[(size_t)site];
    *n_blocks = t.n();
    *len = t.total;
    if (q && p && rows && cols && offsets) {
      std::memcpy(q, t.q.data(), sizeof(uint64_t) * t.q.size());
      std::memcpy(p, t.p.data(), sizeof(uint32_t) * t.p.size());
      std::memcpy(rows, t.rows.data(), sizeof(int64_t) * t.rows.size());
      std::memcpy(cols, t.cols.data(), sizeof(int64_t) * t.cols.size());
      std::memcpy(offsets, t.off.data(), sizeof(int64_t) * t.off.size());
    }
    if (dst && t.total > 0) {
      hchk(hipMemcpyAsync(dst, D->dA[(size_t)site].p, sizeof(double) * (size_t)t.total, hipMemcpyDeviceToDevice,
                          (hipStream_t)stream),
           "dmrg_site copy");
      hchk(hipStreamSynchronize((hipStream_t)stream), "dmrg_site sync");
    }
    return MPSA_OK;
  } catch (const HErr& e) {
    set_last_error(e.what());
    return e.code;
  }
}

int mpsa_dmrg_destroy(mpsa_dmrg* D) {
  delete D;
  return MPSA_OK;
}

}  // extern "C"
