// mpsa_blockops.cpp -- block-sparse tensor algebra outside the contraction
// plans: the two-site product BlockMatrix<2>(B1, B2) that starts every DMRG
// bond step (src/BlockTensor.hpp:285-301).
//
// Every produced block is one GEMM of the grouped FP64 kernel (two tasks of
// two segments each for complex128 planar data: Re = Ar Br - Ai Bi,
// Im = Ar Bi + Ai Br), finalised by the planner's own rules (positive k,
// longest tasks first, tile class by padding) and launched once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

namespace mpsa {
namespace {

struct BErr : std::runtime_error {
  int code;
  BErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void hchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw BErr(MPSA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// extent (elements of one plane) of a table, after checking its fields
int64_t extent_of(const mpsa_blocks* b, int ns, const char* name) {
  if (!b || b->ns != ns || b->n < 0 || (b->is_complex != 0 && b->is_complex != 1) ||
      (b->n > 0 && (!b->q || !b->p || !b->rows || !b->cols || !b->offset)))
    throw BErr(MPSA_EINVAL, std::string("twosite: bad table ") + name);
  int64_t e = 0;
  for (int64_t i = 0; i < b->n; ++i) {
    if (b->rows[i] < 0 || b->cols[i] < 0 || b->offset[i] < 0)
      throw BErr(MPSA_EINVAL, std::string("twosite: negative extent in table ") + name);
    e = std::max(e, b->offset[i] + b->rows[i] * b->cols[i]);
  }
  return e;
}

// the produced keys (l, r, s1, s2) in key order with the block pair that
// makes each one: the reference assigns (*this)[k] = B1 B2 pair by pair, so
// a key met twice keeps the later pair
typedef std::tuple<uint64_t, uint64_t, uint32_t, uint32_t> K2;
std::map<K2, std::pair<int64_t, int64_t>> pairs_of(const mpsa_blocks* b1, const mpsa_blocks* b2) {
  std::map<K2, std::pair<int64_t, int64_t>> out;
  // index B2 by its left charge, keeping the table order within a charge
  std::map<uint64_t, std::vector<int64_t>> by_left;
  for (int64_t j = 0; j < b2->n; ++j) by_left[b2->q[2 * j]].push_back(j);
  for (int64_t i = 0; i < b1->n; ++i) {
    auto f = by_left.find(b1->q[2 * i + 1]);
    if (f == by_left.end()) continue;
    for (int64_t j : f->second) {
      if (b1->cols[i] != b2->rows[j])
        throw BErr(MPSA_EINVAL, "twosite: inner dimensions of a matching block pair differ");
      out[K2{b1->q[2 * i], b2->q[2 * j + 1], b1->p[i], b2->p[j]}] = {i, j};
    }
  }
  return out;
}

}  // namespace
}  // namespace mpsa

using namespace mpsa;

extern "C" {

int mpsa_twosite_outkeys(const mpsa_blocks* b1, const mpsa_blocks* b2, int64_t cap, int64_t* n, uint64_t* q,
                         uint32_t* p, int64_t* rows, int64_t* cols) {
  try {
    extent_of(b1, 1, "b1");
    extent_of(b2, 1, "b2");
    if (!n || cap < 0 || (cap > 0 && (!q || !p || !rows || !cols)))
      throw BErr(MPSA_EINVAL, "twosite_outkeys: bad output arguments");
    const auto pr = pairs_of(b1, b2);
    *n = (int64_t)pr.size();
    if (cap == 0) return MPSA_OK;
    if (cap < *n) throw BErr(MPSA_EINVAL, "twosite_outkeys: cap smaller than the key count");
    int64_t t = 0;
    for (auto& kv : pr) {
      q[2 * t] = std::get<0>(kv.first);
      q[2 * t + 1] = std::get<1>(kv.first);
      p[2 * t] = std::get<2>(kv.first);
      p[2 * t + 1] = std::get<3>(kv.first);
      rows[t] = b1->rows[kv.second.first];
      cols[t] = b2->cols[kv.second.second];
      ++t;
    }
    return MPSA_OK;
  } catch (const BErr& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

int mpsa_twosite_product(const mpsa_blocks* b1, const double* d1, const mpsa_blocks* b2, const double* d2,
                         const mpsa_blocks* out, double* dout, mpsa_stream stream) {
  GTask* dt = nullptr;
  GSeg* ds = nullptr;
  hipStream_t st = (hipStream_t)stream;
  auto cleanup = [&]() {
    if (dt || ds) (void)hipStreamSynchronize(st);
    if (dt) (void)hipFree(dt);
    if (ds) (void)hipFree(ds);
    dt = nullptr;
    ds = nullptr;
  };
  try {
    const int64_t e1 = extent_of(b1, 1, "b1"), e2 = extent_of(b2, 1, "b2"), eo = extent_of(out, 2, "out");
    if (b1->is_complex != b2->is_complex || b1->is_complex != out->is_complex)
      throw BErr(MPSA_EINVAL, "twosite_product: tables must be all real or all complex");
    if ((e1 > 0 && !d1) || (e2 > 0 && !d2) || (eo > 0 && !dout))
      throw BErr(MPSA_EINVAL, "twosite_product: null device buffer");
    const int planes = b1->is_complex ? 2 : 1;
    std::map<K2, int64_t> where;
    for (int64_t t = 0; t < out->n; ++t)
      where[K2{out->q[2 * t], out->q[2 * t + 1], out->p[2 * t], out->p[2 * t + 1]}] = t;
    const auto pr = pairs_of(b1, b2);
    // blocks of out that no pair produces are zero; produced ones are overwritten
    if (eo > 0) hchk(launch_fill(dout, eo * planes, 0.0, st), "twosite_product: zero");
    std::vector<GTask> tasks;
    std::vector<GSeg> segs;
    for (auto& kv : pr) {
      auto f = where.find(kv.first);
      if (f == where.end()) throw BErr(MPSA_EINVAL, "twosite_product: a produced key is missing from the output table");
      const int64_t t = f->second, i = kv.second.first, j = kv.second.second;
      const int64_t m = b1->rows[i], k = b1->cols[i], n = b2->cols[j];
      if (out->rows[t] != m || out->cols[t] != n)
        throw BErr(MPSA_EINVAL, "twosite_product: output block shape differs from the product's");
      if (m == 0 || n == 0 || k == 0) continue;  // empty product: the zeroed block stands
      auto seg = [&](int64_t ao, int64_t bo, double al) {
        GSeg g{};
        g.a_off = ao;
        g.b_off = bo;
        g.k = (int32_t)k;
        g.lda = (int32_t)m;
        g.ldb = (int32_t)k;
        g.alpha = al;
        segs.push_back(g);
      };
      for (int pl = 0; pl < planes; ++pl) {
        GTask g{};
        g.c_off = out->offset[t] + pl * eo;
        g.m = (int32_t)m;
        g.n = (int32_t)n;
        g.seg0 = (int32_t)segs.size();
        const int64_t a = b1->offset[i], b = b2->offset[j];
        if (planes == 1) {
          seg(a, b, 1.0);
        } else if (pl == 0) {
          seg(a, b, 1.0);
          seg(a + e1, b + e2, -1.0);
        } else {
          seg(a, b + e2, 1.0);
          seg(a + e1, b, 1.0);
        }
        g.nseg = (int32_t)segs.size() - g.seg0;
        tasks.push_back(g);
      }
    }
    if (!tasks.empty()) {
      int tiles = 0, cls = 0;
      finalize_gemm_list(tasks, segs, F64, &tiles, &cls);
      if (!tasks.empty() && tiles > 0) {
        hchk(hipMalloc((void**)&dt, sizeof(GTask) * tasks.size()), "twosite_product: tasks");
        hchk(hipMalloc((void**)&ds, sizeof(GSeg) * std::max<size_t>(segs.size(), 1)), "twosite_product: segs");
        hchk(hipMemcpyAsync(dt, tasks.data(), sizeof(GTask) * tasks.size(), hipMemcpyHostToDevice, st),
             "twosite_product: tasks");
        if (!segs.empty())
          hchk(hipMemcpyAsync(ds, segs.data(), sizeof(GSeg) * segs.size(), hipMemcpyHostToDevice, st),
               "twosite_product: segs");
        hchk(launch_gemm(F64, 0, 0, cls, dt, (int)tasks.size(), ds, tiles, 1, d1, 0, d2, 0, dout, 0, st),
             "twosite_product: GEMM");
      }
    }
    cleanup();  // waits for the launch: the host descriptor copies may go
    return MPSA_OK;
  } catch (const BErr& e) {
    cleanup();
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    cleanup();
    set_last_error(e.what());
    return MPSA_EINVAL;
  }
}

}  // extern "C"
