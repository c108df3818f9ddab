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
#include "mpsa_host.h"
#include "mpsa_internal.h"

using namespace mpsa;

namespace {
using AErr = mpsa::host::HErr;
using namespace mpsa::host;
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
