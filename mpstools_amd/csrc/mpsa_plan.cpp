// mpsa_plan.cpp -- sector plans and the C ABI (include/mpstools_amd.h).
//
// A plan flattens the reference's tuple loop into device task lists:
//   contractTwosite / doContraction (src/MPSfunctions.hpp:726-804, 949-981)
//   contractSingle  / doContraction (src/MPSfunctions.hpp:808-885)
//   addContractionLayer / add_out_sub(_2) (src/MPSfunctions.hpp:276-555)
// Each valid tuple (environment block, MPO entries, tensor block) of the
// reference becomes a coefficient on a product; products that share the same
// two operands are computed once (stage 1), the MPO coefficients are applied
// by a mix, and every output block is one segmented-K product (stage 2).  The
// tuples are enumerated in exactly the reference's loop order, so the output
// key set is the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mpstools_amd.h"
#include "mpsa_internal.h"

namespace mpsa {

static thread_local std::string g_err;
void set_last_error(const char* msg) { g_err = msg; }

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) throw Error(MPSA_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ tables
struct BKey {
  uint64_t q0, q1;
  uint32_t p0, p1;
  bool operator<(const BKey& o) const {
    if (q0 != o.q0) return q0 < o.q0;
    if (q1 != o.q1) return q1 < o.q1;
    if (p0 != o.p0) return p0 < o.p0;
    return p1 < o.p1;
  }
  bool operator==(const BKey& o) const { return q0 == o.q0 && q1 == o.q1 && p0 == o.p0 && p1 == o.p1; }
};
// hash for the lookup tables (their iteration order is never used)
struct BKeyHash {
  size_t operator()(const BKey& k) const {
    uint64_t h = k.q0 * 0x9E3779B97F4A7C15ull;
    h ^= (k.q1 + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
    h ^= ((uint64_t)k.p0 << 32 | k.p1) * 0x165667B19E3779F9ull;
    return (size_t)(h ^ (h >> 29));
  }
};
inline uint64_t pair_key(int a, int b) { return (uint64_t)(uint32_t)a << 32 | (uint32_t)b; }

struct Tab {
  int64_t n = 0;
  int ns = 1;
  std::vector<BKey> key;
  std::vector<int64_t> rows, cols, off;
  std::unordered_map<BKey, int, BKeyHash> idx;
  int64_t total = 0;  // extent of one plane: max(offset + rows * cols)
  int cplx = 0;       // complex128, planar: real parts, then imaginary parts at + total
  int find(const BKey& k) const {
    auto it = idx.find(k);
    return it == idx.end() ? -1 : it->second;
  }
};

static Tab parse(const mpsa_blocks* b, int ns_expect, const char* name) {
  if (!b) throw Error(MPSA_EINVAL, std::string(name) + ": null table");
  if (b->n < 0) throw Error(MPSA_EINVAL, std::string(name) + ": negative block count");
  if (b->ns != ns_expect)
    throw Error(MPSA_EINVAL, std::string(name) + ": expected ns=" + std::to_string(ns_expect));
  if (b->is_complex != 0 && b->is_complex != 1) throw Error(MPSA_EINVAL, std::string(name) + ": bad is_complex");
  if (b->n > 0 && (!b->q || !b->p || !b->rows || !b->cols || !b->offset))
    throw Error(MPSA_EINVAL, std::string(name) + ": null array");
  Tab t;
  t.n = b->n;
  t.ns = b->ns;
  t.cplx = b->is_complex;
  t.idx.reserve((size_t)b->n);
  t.key.reserve((size_t)b->n);
  for (int64_t i = 0; i < b->n; ++i) {
    BKey k{b->q[2 * i], b->q[2 * i + 1], b->p[b->ns * i], b->ns == 2 ? b->p[b->ns * i + 1] : 0u};
    if (b->rows[i] < 0 || b->cols[i] < 0 || b->offset[i] < 0 || b->rows[i] > (1 << 30) || b->cols[i] > (1 << 30))
      throw Error(MPSA_EINVAL, std::string(name) + ": bad block shape/offset");
    if (!t.idx.emplace(k, (int)i).second) throw Error(MPSA_EINVAL, std::string(name) + ": duplicate key");
    t.key.push_back(k);
    t.rows.push_back(b->rows[i]);
    t.cols.push_back(b->cols[i]);
    t.off.push_back(b->offset[i]);
    t.total = std::max(t.total, b->offset[i] + b->rows[i] * b->cols[i]);
  }
  return t;
}

struct Entry {
  uint32_t in, out;
  double v;
};
struct Group {
  uint32_t wl, wr;
  std::vector<Entry> e;
};

// MPOMat<1,T> in the reference's iteration order: ukey (wl, wr) ascending,
// entries by OpKeyType order (in, out) (src/keyTypes.hpp:313-343, 471-483).
static std::vector<Group> parse_mpo(const mpsa_mpo* w, int d, const char* name) {
  if (!w) throw Error(MPSA_EINVAL, std::string(name) + ": null MPO");
  if (w->is_complex) throw Error(MPSA_EINVAL, std::string(name) + ": complex MPO not supported by this build");
  if (w->n < 0 || (w->n > 0 && (!w->wl || !w->wr || !w->out || !w->in || !w->val)))
    throw Error(MPSA_EINVAL, std::string(name) + ": bad MPO arrays");
  std::map<std::pair<uint32_t, uint32_t>, std::map<std::pair<uint32_t, uint32_t>, double>> m;
  for (int64_t i = 0; i < w->n; ++i) {
    if ((int)w->out[i] >= d || (int)w->in[i] >= d)
      throw Error(MPSA_EINVAL, std::string(name) + ": physical index outside i_to_key");
    m[{w->wl[i], w->wr[i]}][{w->in[i], w->out[i]}] = w->val[i];
  }
  std::vector<Group> g;
  for (auto& kv : m) {
    Group gr{kv.first.first, kv.first.second, {}};
    for (auto& e : kv.second) gr.e.push_back(Entry{e.first.first, e.first.second, e.second});
    g.push_back(gr);
  }
  return g;
}

// ------------------------------------------------------------------ plan
struct Stage {
  std::vector<GTask> tasks;
  std::vector<GSeg> segs;
  int tiles = 0;
  int tile_class = 0;
  int tr = 0;
  GTask* d_tasks = nullptr;
  GSeg* d_segs = nullptr;
};

struct Product {  // a T1 slot: C(m x n) = op(P) op(Q)
  int ia, ib;
  int64_t m, n, k;
  int64_t ws_off;
};
struct T2Slot {
  int iout, iu;
  int64_t m, n;
  std::vector<std::pair<int, double>> terms;  // (T1 slot, coef)
  int64_t ws_off = -1;                        // -1: single term, read the T1 slot directly
};

}  // namespace mpsa

using namespace mpsa;

struct mpsa_plan {
  int kind = 0;  // 0: heff (two- or one-site), 2: env
  int prec = F64;
  int direction = 1;
  // layout facts
  int64_t L_total = 0, R_total = 0, x_total = 0, y_total = 0;  // heff: L, R, x, y; env: BL, A, Ap, out
  int64_t ws_per_vec = 0;                                     // workspace elements per vector (compute type)
  int64_t xc_per_vec = 0;                                     // converted x elements per vector
  size_t esize = 8;
  Stage s1, s2;
  std::vector<MixTask> mix;
  std::vector<MixTerm> terms;
  int mix_chunks = 0;
  std::vector<Range> zero;
  std::vector<Repack> rpL, rpR, rpx;
  int rpL_chunks = 0, rpR_chunks = 0, rpx_chunks = 0;
  int64_t Lc_elems = 0, Rc_elems = 0;
  int planes = 1;  // 2: hi/lo bf16 planes (BF16X3)
  int cplx = 0;    // complex128 operands in planar layout (FP64 only)
  // device
  MixTask* d_mix = nullptr;
  MixTerm* d_terms = nullptr;
  Range* d_zero = nullptr;
  Repack *d_rpL = nullptr, *d_rpR = nullptr, *d_rpx = nullptr;
  void* ws = nullptr;
  int ws_nvec = 0;
  void* xc = nullptr;
  int xc_nvec = 0;
  void* Lc = nullptr;
  void* Rc = nullptr;
  const double* dL = nullptr;
  const double* dR = nullptr;
  double flops_alg = 0, flops_ref = 0;
  double flops1 = 0, flops2 = 0;  // algorithmic flops of stage 1 / stage 2 (one vector)
  int timing = 0;
  int on_device = 0;
  int xy_same = 0;
  // the launches of one apply replayed from a graph (H_eff plans): valid for
  // the vectors, strides, vector count and buffers it was captured with.  The
  // g_* fields hold the last untimed apply's arguments (g_valid); a graph is
  // recorded on the second apply in a row with the same arguments, so a plan
  // applied once per environment (bind_env drops both) never pays for one.
  hipGraphExec_t gexec = nullptr;
  bool g_valid = false;
  const double* g_dx = nullptr;
  double* g_dy = nullptr;
  int64_t g_xs = 0, g_ys = 0;
  int g_nvec = 0;
  const void *g_ws = nullptr, *g_xc = nullptr;
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    g_valid = false;
  }
  // timing: one set of 5 events per timed apply (up to kTimedApplies kept)
  static constexpr int kTimedApplies = 512;
  std::vector<std::array<hipEvent_t, 5>> evs;
  int nrec = 0;    // applies recorded since the last summary
  int cur = -1;    // event set of the last timed apply
  ~mpsa_plan() {
    drop_graph();
    for (auto& set : evs)
      for (auto& e : set)
        if (e) (void)hipEventDestroy(e);
    for (void* p : {(void*)s1.d_tasks, (void*)s1.d_segs, (void*)s2.d_tasks, (void*)s2.d_segs, (void*)d_mix,
                    (void*)d_terms, (void*)d_zero, (void*)d_rpL, (void*)d_rpR, (void*)d_rpx, ws, xc, Lc, Rc})
      if (p) (void)hipFree(p);
  }
};

namespace mpsa {

template <typename T>
static T* upload(const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* d = nullptr;
  if (hipMalloc((void**)&d, sizeof(T) * v.size()) != hipSuccess) throw Error(MPSA_ENOMEM, "hipMalloc descriptors");
  HIPCHK(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return d;
}

static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }
// bf16 layouts pad every k dimension (and the row-major T leading dimension)
// with zeros to a multiple of 32: the streaming kernel moves whole 32-wide slabs.
static int64_t pad8(int64_t x) { return (x + 31) & ~int64_t(31); }

// Tile class: the largest workgroup tile that wastes at most 20% of its
// MFMA work on padding; the smallest tile otherwise.
// The 256x256 class needs unit scales and per-lane operand offsets
// (row * ld * 2 bytes inside one block) below 2^31.
static bool fits_pipe_tile(const std::vector<GTask>& t, const std::vector<GSeg>& segs) {
  for (auto& x : t)
    for (int j = 0; j < x.nseg; ++j) {
      const GSeg& g = segs[(size_t)(x.seg0 + j)];
      if (g.alpha != 1.0) return false;
      if ((int64_t)x.m * g.lda * 2 >= (int64_t(1) << 31) || (int64_t)x.n * g.ldb * 2 >= (int64_t(1) << 31))
        return false;
    }
  return true;
}

// Tile classes a stage may launch with: every class of the precision, the
// 256x256 class only where fits_pipe_tile holds.
static bool class_admissible(const std::vector<GTask>& t, const std::vector<GSeg>& segs, int prec, int cls) {
  if (cls < 0 || cls >= num_tile_classes(prec)) return false;
  const bool bf = prec == BF16 || prec == BF16X3;
  return !bf || cls != 3 || fits_pipe_tile(t, segs);
}

static int pick_class(const std::vector<GTask>& t, const std::vector<GSeg>& segs, int prec) {
  int best = -1, fallback = 1;
  int best_area = 0;
  double fallback_fill = -1.0;
  const bool pipe_ok = fits_pipe_tile(t, segs);
  const bool bfp = prec == BF16 || prec == BF16X3;
  double fill_cls[kMaxTileClasses] = {0, 0, 0, 0, 0};
  for (int cls = 0; cls < num_tile_classes(prec); ++cls) {
    if (bfp && cls == 3 && !pipe_ok) continue;
    int bm, bn;
    tile_dims(prec, cls, &bm, &bn);
    double useful = 0, padded = 0;
    for (auto& x : t) {
      useful += (double)x.m * x.n;
      padded += (double)((x.m + bm - 1) / bm) * ((x.n + bn - 1) / bn) * (double)bm * bn;
    }
    const double fill = padded > 0 ? useful / padded : 1.0;
    fill_cls[cls] = fill;
    if (!bfp && cls >= 3) continue;  // the wave tiles: decided below (class 4 by the autotuner only)
    if (fill >= 0.8 && bm * bn > best_area) {
      best = cls;
      best_area = bm * bn;
    }
    if (fill > fallback_fill) {  // no tile fills 80 %: the least padded one
      fallback = cls;
      fallback_fill = fill;
    }
  }
  const int pick = best >= 0 ? best : fallback;
  // FP64/FP32 stages of blocks mostly below 16 rows or columns (the 32x32
  // tile under 40 % full): the barrier-free 16x16 wave tile (measured 1.9x
  // on the 16-vector Hubbard D = 1024 bond, slower than 32x32 on the t-V and
  // XXZ bonds whose blocks reach 80 and 196)
  if (!bfp && pick == 2 && fill_cls[2] < 0.4) return 3;
  return pick;
}

// Number the workgroup tiles of a stage for tile class cls, in task order
// (tiles m-fastest inside a task).
static void assign_tiles(Stage& s, int prec, int cls) {
  int BM, BN;
  tile_dims(prec, cls, &BM, &BN);
  int tile = 0;
  for (auto& t : s.tasks) {
    t.tiles_m = (t.m + BM - 1) / BM;
    t.tile0 = tile;
    tile += t.tiles_m * ((t.n + BN - 1) / BN);
  }
  s.tiles = tile;
  s.tile_class = cls;
}

static void finalize_stage(Stage& s, int prec) {
  // drop empty-k segments (kernels assume k > 0); a task left without
  // segments still writes its C block (as zeros)
  {
    std::vector<GSeg> segs;
    for (auto& t : s.tasks) {
      const int first = (int)segs.size();
      for (int j = 0; j < t.nseg; ++j)
        if (s.segs[(size_t)(t.seg0 + j)].k > 0) segs.push_back(s.segs[(size_t)(t.seg0 + j)]);
      t.seg0 = first;
      t.nseg = (int)segs.size() - first;
    }
    s.segs = segs;
  }
  // longest tiles first (static round-robin over the persistent grid)
  std::vector<int64_t> cost(s.tasks.size());
  for (size_t i = 0; i < s.tasks.size(); ++i) {
    int64_t k = 0;
    for (int j = 0; j < s.tasks[i].nseg; ++j) k += s.segs[(size_t)(s.tasks[i].seg0 + j)].k;
    cost[i] = k;
  }
  std::vector<size_t> ord(s.tasks.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
  std::vector<GTask> t2;
  for (size_t i : ord)
    if (s.tasks[i].m > 0 && s.tasks[i].n > 0) t2.push_back(s.tasks[i]);
  s.tasks = t2;
  assign_tiles(s, prec, pick_class(s.tasks, s.segs, prec));
}

// A grouped GEMM assembled outside a plan (block products of the compression
// and the two-site product) goes through the same finalisation: empty-k
// segments dropped, empty tasks dropped, longest tasks first, tile class by
// the same heuristic, tiles numbered per task.
void finalize_gemm_list(std::vector<GTask>& tasks, std::vector<GSeg>& segs, int prec, int* tiles, int* tile_class) {
  Stage s;
  s.tasks.swap(tasks);
  s.segs.swap(segs);
  finalize_stage(s, prec);
  tasks.swap(s.tasks);
  segs.swap(s.segs);
  *tiles = s.tiles;
  *tile_class = s.tile_class;
}

static void finalize_mix(mpsa_plan* P) {
  int chunk = 0;
  std::vector<MixTask> m2;
  for (auto t : P->mix) {
    if (t.len <= 0) continue;
    t.chunk0 = chunk;
    chunk += (int)((t.len + 1023) / 1024);
    m2.push_back(t);
  }
  P->mix = m2;
  P->mix_chunks = chunk;
}

static int finalize_repack(std::vector<Repack>& r) {
  int chunk = 0;
  std::vector<Repack> r2;
  for (auto t : r) {
    if ((int64_t)t.rows * t.cols <= 0) continue;
    // the conversion kernels index a block with 32-bit arithmetic
    if ((int64_t)t.rows * t.cols >= (int64_t(1) << 31)) throw Error(MPSA_EINVAL, "block too large for conversion");
    t.chunk0 = chunk;
    chunk += (int)(((int64_t)t.rows * t.cols + 1023) / 1024);
    r2.push_back(t);
  }
  r = r2;
  return chunk;
}

// The generic two-stage builder.  Tuples come from the enumerators below.
struct Tuple {
  int ia, ib, iu;  // stage-1 operands (P, Q) and stage-2 operand U
  BKey out;
  double coef;
  int64_t m1, n1, k1;  // T1 dims
  int64_t n_out_rows, n_out_cols;
};

struct Built {
  std::vector<Product> prods;
  std::vector<T2Slot> slots;
  std::vector<BKey> outkeys;  // produced keys in key order
  double flops_ref = 0;
};

// The produced keys (sorted) and their shapes.
static void produced_keys(const std::vector<Tuple>& tup, std::vector<BKey>& outkeys,
                          std::map<BKey, std::pair<int64_t, int64_t>>* out_dims) {
  // lookups hashed; the key list sorted
  std::unordered_map<BKey, std::pair<int64_t, int64_t>, BKeyHash> dims;
  dims.reserve(tup.size() / 4 + 16);
  for (auto& t : tup) {
    auto ins = dims.emplace(t.out, std::make_pair(t.n_out_rows, t.n_out_cols));
    if (!ins.second && (ins.first->second.first != t.n_out_rows || ins.first->second.second != t.n_out_cols))
      throw Error(MPSA_EINVAL, "inconsistent output block shapes for one key");
  }
  outkeys.clear();
  outkeys.reserve(dims.size());
  for (auto& kv : dims) outkeys.push_back(kv.first);
  std::sort(outkeys.begin(), outkeys.end());
  for (auto& k : outkeys) {
    auto od = out_dims->find(k);
    const auto& d = dims.at(k);
    if (od == out_dims->end()) out_dims->emplace_hint(out_dims->end(), k, d);
    else if (od->second != d) throw Error(MPSA_EINVAL, "inconsistent output block shapes for one key");
  }
}

// Products (ia, ib) and stage-2 slots (iout, iu) in first-seen order, against
// the output table ytab.  plain = true keeps the reference's schedule: one
// product per tuple (F_ref flops), no sharing of (P, Q) products and no mix.
static void group_products(const std::vector<Tuple>& tup, const Tab* ytab, bool plain, Built& b) {
  std::unordered_map<uint64_t, int> pidx, sidx;
  pidx.reserve(tup.size() / 2 + 16);
  sidx.reserve(tup.size() / 2 + 16);
  int tix = 0;
  for (auto& t : tup) {
    int iout = ytab->find(t.out);
    if (iout < 0) throw Error(MPSA_ENOKEY, "output table lacks a produced key");
    const uint64_t pk = plain ? pair_key(tix, -1) : pair_key(t.ia, t.ib);
    auto it = pidx.emplace(pk, (int)b.prods.size());
    const int pi = it.first->second;
    if (it.second) b.prods.push_back(Product{t.ia, t.ib, t.m1, t.n1, t.k1, 0});
    const uint64_t sk = plain ? pair_key(iout, tix) : pair_key(iout, t.iu);
    ++tix;
    auto jt = sidx.emplace(sk, (int)b.slots.size());
    const int si = jt.first->second;
    if (jt.second) {
      T2Slot s;
      s.iout = iout;
      s.iu = t.iu;
      s.m = t.m1;
      s.n = t.n1;
      b.slots.push_back(s);
    }
    auto& terms = b.slots[(size_t)si].terms;
    bool found = false;
    for (auto& tm : terms)
      if (tm.first == pi) {
        tm.second += t.coef;
        found = true;
      }
    if (!found) terms.push_back({pi, t.coef});
  }
}

static Built group_tuples(const std::vector<Tuple>& tup, const Tab* ytab, bool need_y,
                          std::map<BKey, std::pair<int64_t, int64_t>>* out_dims, bool plain = false) {
  Built b;
  produced_keys(tup, b.outkeys, out_dims);
  if (need_y) group_products(tup, ytab, plain, b);
  return b;
}

// ---- tuple enumerators (reference loop order) ----

// Two-site: doContraction (src/MPSfunctions.hpp:726-804) over left charges in
// ascending order (contractTwosite :958-976).
static std::vector<Tuple> tuples_heff2(const Tab& L, const Tab& R, const Tab& x, const std::vector<Group>& wl,
                                       const std::vector<Group>& wr, const uint64_t* itl, const uint64_t* itr,
                                       double* fref) {
  std::vector<Tuple> out;
  *fref = 0;
  for (int iL = 0; iL < (int)L.n; ++iL) {
    const BKey& kl = L.key[(size_t)iL];
    for (auto& o1 : wl) {
      for (auto& o2 : wr) {
        if (o1.wl != kl.p0) continue;
        if (o2.wl != o1.wr) continue;
        for (auto& a : o1.e) {
          for (auto& b : o2.e) {
            const uint64_t kylo = itl[a.out] + itr[b.out];
            int ix = x.find(BKey{kl.q1, kl.q1 + kylo, a.out, b.out});
            if (ix < 0) continue;
            const uint64_t kyup = itl[a.in] + itr[b.in];
            int iR = R.find(BKey{kl.q0 + kyup, kl.q1 + kylo, o2.wr, 0});
            if (iR < 0) continue;
            if (L.cols[(size_t)iL] != x.rows[(size_t)ix] || R.cols[(size_t)iR] != x.cols[(size_t)ix])
              throw Error(MPSA_EINVAL, "heff2: block dimensions of L, psi, R do not match");
            Tuple t;
            t.ia = iL;
            t.ib = ix;
            t.iu = iR;
            t.out = BKey{kl.q0, R.key[(size_t)iR].q0, a.in, b.in};
            t.coef = a.v * b.v;
            t.m1 = L.rows[(size_t)iL];
            t.n1 = x.cols[(size_t)ix];
            t.k1 = L.cols[(size_t)iL];
            t.n_out_rows = L.rows[(size_t)iL];
            t.n_out_cols = R.rows[(size_t)iR];
            *fref += 2.0 * t.m1 * t.k1 * t.n1 + 2.0 * t.m1 * t.n1 * t.n_out_cols;
            out.push_back(t);
          }
        }
      }
    }
  }
  return out;
}

// One-site: doContraction (src/MPSfunctions.hpp:808-854).
static std::vector<Tuple> tuples_heff1(const Tab& L, const Tab& R, const Tab& x, const std::vector<Group>& w,
                                       const uint64_t* itk, double* fref) {
  std::vector<Tuple> out;
  *fref = 0;
  for (int iL = 0; iL < (int)L.n; ++iL) {
    const BKey& kl = L.key[(size_t)iL];
    for (auto& o : w) {
      if (o.wl != kl.p0) continue;
      for (auto& a : o.e) {
        const uint64_t kylo = itk[a.out];
        int ix = x.find(BKey{kl.q1, kl.q1 + kylo, a.out, 0});
        if (ix < 0) continue;
        const uint64_t kyup = itk[a.in];
        int iR = R.find(BKey{kl.q0 + kyup, kl.q1 + kylo, o.wr, 0});
        if (iR < 0) continue;
        if (L.cols[(size_t)iL] != x.rows[(size_t)ix] || R.cols[(size_t)iR] != x.cols[(size_t)ix])
          throw Error(MPSA_EINVAL, "heff1: block dimensions of L, A, R do not match");
        Tuple t;
        t.ia = iL;
        t.ib = ix;
        t.iu = iR;
        t.out = BKey{kl.q0, R.key[(size_t)iR].q0, a.in, 0};
        t.coef = a.v;
        t.m1 = L.rows[(size_t)iL];
        t.n1 = x.cols[(size_t)ix];
        t.k1 = L.cols[(size_t)iL];
        t.n_out_rows = L.rows[(size_t)iL];
        t.n_out_cols = R.rows[(size_t)iR];
        *fref += 2.0 * t.m1 * t.k1 * t.n1 + 2.0 * t.m1 * t.n1 * t.n_out_cols;
        out.push_back(t);
      }
    }
  }
  return out;
}

// Environment: add_out_sub(_2) (src/MPSfunctions.hpp:276-435) for every
// outgoing charge in ascending order (addContractionLayer :455-488).
static std::vector<Tuple> tuples_env(const Tab& BL, const Tab& A, const Tab& Aup, const std::vector<Group>& w,
                                     const uint64_t* itk, int direction, double* fref) {
  std::vector<Tuple> out;
  *fref = 0;
  const int idir = direction > 0 ? 1 : 0;
  std::map<uint64_t, int> outgoing;
  for (auto& k : A.key) outgoing[idir ? k.q1 : k.q0] = 1;
  // BL blocks by (aux index, inner charge), each list in table order: the
  // loop below visits exactly the blocks the reference's scan matches, in
  // the same order, without scanning the others
  std::map<std::pair<uint32_t, uint64_t>, std::vector<int>> blits;
  for (int i = 0; i < (int)BL.n; ++i) blits[{BL.key[(size_t)i].p0, BL.key[(size_t)i].q1}].push_back(i);
  for (auto& pk : outgoing) {
    const uint64_t p = pk.first;
    for (auto& o : w) {
      const uint32_t ko[2] = {o.wl, o.wr};
      const uint32_t o_in = ko[1 - idir];
      for (auto& oi : o.e) {
        const uint64_t kyout = itk[oi.out];
        BKey kb = direction > 0 ? BKey{p - kyout, p, oi.out, 0} : BKey{p, p + kyout, oi.out, 0};
        int ib = A.find(kb);
        if (ib < 0) continue;
        auto bl = blits.find({o_in, direction > 0 ? kb.q0 : kb.q1});
        if (bl == blits.end()) continue;
        for (int iBL : bl->second) {
          const BKey& k = BL.key[(size_t)iBL];
          const uint64_t kyin = itk[oi.in];
          BKey ka = direction > 0 ? BKey{k.q0, k.q0 + kyin, oi.in, 0} : BKey{k.q0 - kyin, k.q0, oi.in, 0};
          int ia = Aup.find(ka);
          if (ia < 0) continue;
          Tuple t;
          t.ia = iBL;
          t.ib = ib;
          t.iu = ia;
          t.out = BKey{idir ? ka.q1 : ka.q0, idir ? kb.q1 : kb.q0, ko[idir], 0};
          t.coef = oi.v;
          const int64_t ma = BL.rows[(size_t)iBL], mb = BL.cols[(size_t)iBL];
          if (direction > 0) {
            // T1 = BL (ma x mb) * A_low (mb x mb'); out = Aup^T (ma' x ma) * T1
            if (A.rows[(size_t)ib] != mb || Aup.rows[(size_t)ia] != ma)
              throw Error(MPSA_EINVAL, "env: block dimensions do not match");
            t.m1 = ma;
            t.n1 = A.cols[(size_t)ib];
            t.k1 = mb;
            t.n_out_rows = Aup.cols[(size_t)ia];
            t.n_out_cols = A.cols[(size_t)ib];
          } else {
            // T1 = BR (ma x mb) * A_low^T (mb x mb'); out = Aup (ma' x ma) * T1
            if (A.cols[(size_t)ib] != mb || Aup.cols[(size_t)ia] != ma)
              throw Error(MPSA_EINVAL, "env: block dimensions do not match");
            t.m1 = ma;
            t.n1 = A.rows[(size_t)ib];
            t.k1 = mb;
            t.n_out_rows = Aup.rows[(size_t)ia];
            t.n_out_cols = A.rows[(size_t)ib];
          }
          *fref += 2.0 * t.k1 * ma * t.n_out_rows + 2.0 * t.n_out_rows * ma * t.n1;
          out.push_back(t);
        }
      }
    }
  }
  return out;
}

static int64_t write_keys(const std::vector<BKey>& keys, const std::map<BKey, std::pair<int64_t, int64_t>>& dims,
                          int ns, int64_t cap, int64_t* nout, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols) {
  *nout = (int64_t)keys.size();
  if (cap <= 0) return 0;
  if (cap < (int64_t)keys.size()) throw Error(MPSA_EINVAL, "outkeys: capacity too small");
  for (size_t i = 0; i < keys.size(); ++i) {
    q[2 * i] = keys[i].q0;
    q[2 * i + 1] = keys[i].q1;
    p[(size_t)ns * i] = keys[i].p0;
    if (ns == 2) p[(size_t)ns * i + 1] = keys[i].p1;
    auto d = dims.at(keys[i]);
    if (rows) rows[i] = d.first;
    if (cols) cols[i] = d.second;
  }
  return 0;
}

// Build the device plan of a two-stage contraction.
//   heff: stage 1 T1 = P * Q  (P = L env, Q = x),  stage 2 y = T2 * U^T (U = R env)
//   env>0: T1 = P * Q (P = BL, Q = A), out = U^T * T2 (U = Ap)
//   env<=0: T1 = P * Q^T (P = BR, Q = A), out = U * T2 (U = Ap)
// Complex128 plans keep every operand in planar layout: the real parts of its
// blocks at the table's offsets, then the imaginary parts one plane extent
// further on.  The MPO values are real, so each product a * op(A) * op(B) of
// the real schedule becomes four real products on the planes,
//   Re C += a (Ar Br - s Ai Bi),   Im C += a (Ar Bi + s Ai Br),
// with s = -1 where A enters conjugated (the upper tensor of an environment
// update: the reference's conj(A_in), src/MPSfunctions.hpp:335-341) and +1
// otherwise.  Transposes act on both planes alike.  The real and the
// imaginary half of an output block are separate GEMM tasks whose segments
// accumulate in registers, so every output plane is written once.
// nega / negb >= 0: offset of a negated copy of A's / B's imaginary plane
// (reduced-precision plans keep one), which then carries the minus sign of
// Ar Br - Ai Bi so that every segment keeps its scale.
static void complexify_stage(Stage& S, int64_t pa, int64_t pb, int64_t pc, double s, int64_t nega = -1,
                             int64_t negb = -1) {
  std::vector<GTask> t2;
  std::vector<GSeg> s2;
  t2.reserve(2 * S.tasks.size());
  s2.reserve(4 * S.segs.size());
  for (int im = 0; im < 2; ++im)
    for (const GTask& t : S.tasks) {
      GTask u = t;
      u.c_off = t.c_off + (im ? pc : 0);
      u.seg0 = (int32_t)s2.size();
      u.nseg = 2 * t.nseg;
      for (int j = 0; j < t.nseg; ++j) {
        const GSeg& g = S.segs[(size_t)(t.seg0 + j)];
        GSeg x = g, y = g;
        y.a_off += pa;  // the second product always takes A's imaginary plane
        if (!im) {      // Ar Br, -s Ai Bi
          y.b_off += pb;
          y.alpha = -s * g.alpha;
          if (s > 0 && nega >= 0) {
            y.a_off = g.a_off + nega;
            y.alpha = g.alpha;
          } else if (s > 0 && negb >= 0) {
            y.b_off = g.b_off + negb;
            y.alpha = g.alpha;
          }
        } else {        // Ar Bi, s Ai Br
          x.b_off += pb;
          y.alpha = s * g.alpha;
        }
        s2.push_back(x);
        s2.push_back(y);
      }
      t2.push_back(u);
    }
  S.tasks.swap(t2);
  S.segs.swap(s2);
}

// Plane extents: P (stage-1 A: L or the old environment), Q (stage-1 B: x or
// the lower tensor), U (R, or the upper tensor of an environment update), Y
// (the output).  Doubles the plan's buffer extents and its flop counts
// (a complex multiply-add is four real ones).
// Reduced-precision H_eff plans convert both planes: L and R as [re | im |
// -im], x as [re | im], each plane in the real plan's converted layout.
static void dup_repack(std::vector<Repack>& r, int64_t src_plane, int64_t dst_plane, bool neg) {
  const size_t n = r.size();
  for (size_t i = 0; i < n; ++i) {
    Repack im = r[i];
    im.src_off += src_plane;
    im.dst_off += dst_plane;
    r.push_back(im);
    if (neg) {
      Repack ng = im;
      ng.dst_off += dst_plane;
      ng.scale = -im.scale;
      r.push_back(ng);
    }
  }
}

static void complexify(mpsa_plan* P, int64_t pP, int64_t pQ, int64_t pU, int64_t pY, int kind) {
  const int64_t wsr = P->ws_per_vec;
  if (P->prec == F64) {
    complexify_stage(P->s1, pP, pQ, wsr, 1.0);
    if (kind != 2) complexify_stage(P->s2, wsr, pU, pY, 1.0);  // y = T R^T
    else complexify_stage(P->s2, pU, wsr, pY, -1.0);           // out = op(conj(Aup)) T
  } else {  // H_eff only (environment plans are FP64)
    const int64_t Lr = P->Lc_elems, Rr = P->Rc_elems, xr = P->xc_per_vec;
    dup_repack(P->rpL, pP, Lr, true);
    dup_repack(P->rpR, pU, Rr, true);
    dup_repack(P->rpx, pQ, xr, false);
    complexify_stage(P->s1, Lr, xr, wsr, 1.0, 2 * Lr, -1);
    complexify_stage(P->s2, wsr, Rr, pY, 1.0, -1, 2 * Rr);
    P->Lc_elems = 3 * Lr;
    P->Rc_elems = 3 * Rr;
    P->xc_per_vec = 2 * xr;
  }
  const size_t nm = P->mix.size();
  for (size_t i = 0; i < nm; ++i) {  // W mixes are real: the same sums on the imaginary plane
    MixTask m = P->mix[i];
    m.dst_off += wsr;
    const int32_t t0 = (int32_t)P->terms.size();
    for (int j = 0; j < m.nt; ++j) {
      MixTerm x = P->terms[(size_t)(m.t0 + j)];
      x.src_off += wsr;
      P->terms.push_back(x);
    }
    m.t0 = t0;
    P->mix.push_back(m);
  }
  const size_t nz = P->zero.size();
  for (size_t i = 0; i < nz; ++i) P->zero.push_back(Range{P->zero[i].off + pY, P->zero[i].len});
  P->ws_per_vec = 2 * wsr;
  P->L_total = 2 * pP;
  P->x_total = 2 * pQ;
  P->R_total = 2 * pU;
  P->y_total = 2 * pY;
  P->flops_alg *= 4;
  P->flops1 *= 4;
  P->flops2 *= 4;
}

static void build_plan(mpsa_plan* P, Built& b, const Tab& Ptab, const Tab& Qtab, const Tab& Utab, const Tab& Ytab,
                       int kind, int direction) {
  const bool bf = P->prec == BF16 || P->prec == BF16X3;
  // A single-term W mix is elided: stage 2 reads the T1 product directly and
  // the coefficient rides on the segment (FP64/FP32), or, in BF16 plans, on a
  // scaled copy of the R block made once per environment binding, so that
  // every BF16 segment has unit scale.
  // workspace layout
  int64_t off = 0;
  for (auto& pr : b.prods) {
    pr.ws_off = off;
    off = align_up(off + pr.m * (bf ? pad8(pr.n) : pr.n), 64);
  }
  for (auto& s : b.slots) {
    if (s.terms.size() == 1) continue;
    s.ws_off = off;
    off = align_up(off + s.m * (bf ? pad8(s.n) : s.n), 64);
  }
  P->ws_per_vec = std::max<int64_t>(off, 64);
  P->flops_alg = 0;
  for (auto& pr : b.prods) P->flops_alg += 2.0 * pr.m * pr.n * pr.k;
  P->flops1 = P->flops_alg;

  // bf16: converted operands with padded, k-contiguous layouts
  std::vector<int64_t> Loff(Ptab.n), Roff(Utab.n), xoff(Qtab.n);
  std::map<std::pair<int, double>, int64_t> Rscaled;  // (R block, coefficient) -> scaled bf16 copy
  if (bf) {
    int64_t o = 0;
    for (int i = 0; i < (int)Ptab.n; ++i) {  // L^T: rows*pad8(cols) (element (i,k) at i*pad8(K)+k)
      Loff[(size_t)i] = o;
      P->rpL.push_back(Repack{Ptab.off[(size_t)i], o, (int32_t)Ptab.rows[(size_t)i], (int32_t)Ptab.cols[(size_t)i],
                              (int32_t)pad8(Ptab.cols[(size_t)i]), 1, 0, 0});
      o = align_up(o + Ptab.rows[(size_t)i] * pad8(Ptab.cols[(size_t)i]), 64);
    }
    P->Lc_elems = std::max<int64_t>(o, 64);
    o = 0;
    for (int i = 0; i < (int)Utab.n; ++i) {  // R row-major: element (j,k) at j*pad8(cols)+k
      Roff[(size_t)i] = o;
      P->rpR.push_back(Repack{Utab.off[(size_t)i], o, (int32_t)Utab.rows[(size_t)i], (int32_t)Utab.cols[(size_t)i],
                              (int32_t)pad8(Utab.cols[(size_t)i]), 1, 0, 0});
      o = align_up(o + Utab.rows[(size_t)i] * pad8(Utab.cols[(size_t)i]), 64);
    }
    if (kind != 2) {  // scaled R copies for the coefficients of single-term slots
      for (auto& sl : b.slots) {
        if (sl.terms.size() != 1 || sl.terms[0].second == 1.0) continue;
        const auto key = std::make_pair((int)sl.iu, sl.terms[0].second);
        if (Rscaled.count(key)) continue;
        const int i = sl.iu;
        Rscaled[key] = o;
        Repack r{Utab.off[(size_t)i], o, (int32_t)Utab.rows[(size_t)i], (int32_t)Utab.cols[(size_t)i],
                 (int32_t)pad8(Utab.cols[(size_t)i]), 1, 0, 0};
        r.scale = sl.terms[0].second;
        P->rpR.push_back(r);
        o = align_up(o + Utab.rows[(size_t)i] * pad8(Utab.cols[(size_t)i]), 64);
      }
    }
    P->Rc_elems = std::max<int64_t>(o, 64);
    o = 0;
    for (int i = 0; i < (int)Qtab.n; ++i) {  // x col-major with ld pad8(rows)
      xoff[(size_t)i] = o;
      P->rpx.push_back(Repack{Qtab.off[(size_t)i], o, (int32_t)Qtab.rows[(size_t)i], (int32_t)Qtab.cols[(size_t)i],
                              (int32_t)pad8(Qtab.rows[(size_t)i]), 0, 0, 0});
      o = align_up(o + pad8(Qtab.rows[(size_t)i]) * Qtab.cols[(size_t)i], 64);
    }
    P->xc_per_vec = std::max<int64_t>(o, 64);
  } else if (P->prec == F32) {
    for (int i = 0; i < (int)Ptab.n; ++i) {
      Loff[(size_t)i] = Ptab.off[(size_t)i];
      P->rpL.push_back(Repack{Ptab.off[(size_t)i], Ptab.off[(size_t)i], (int32_t)Ptab.rows[(size_t)i],
                              (int32_t)Ptab.cols[(size_t)i], (int32_t)Ptab.rows[(size_t)i], 0, 0, 0});
    }
    for (int i = 0; i < (int)Utab.n; ++i) {
      Roff[(size_t)i] = Utab.off[(size_t)i];
      P->rpR.push_back(Repack{Utab.off[(size_t)i], Utab.off[(size_t)i], (int32_t)Utab.rows[(size_t)i],
                              (int32_t)Utab.cols[(size_t)i], (int32_t)Utab.rows[(size_t)i], 0, 0, 0});
    }
    for (int i = 0; i < (int)Qtab.n; ++i) {
      xoff[(size_t)i] = Qtab.off[(size_t)i];
      P->rpx.push_back(Repack{Qtab.off[(size_t)i], Qtab.off[(size_t)i], (int32_t)Qtab.rows[(size_t)i],
                              (int32_t)Qtab.cols[(size_t)i], (int32_t)Qtab.rows[(size_t)i], 0, 0, 0});
    }
    P->Lc_elems = std::max<int64_t>(P->L_total, 64);
    P->Rc_elems = std::max<int64_t>(P->R_total, 64);
    P->xc_per_vec = std::max<int64_t>(P->x_total, 64);
  } else {
    for (int i = 0; i < (int)Ptab.n; ++i) Loff[(size_t)i] = Ptab.off[(size_t)i];
    for (int i = 0; i < (int)Utab.n; ++i) Roff[(size_t)i] = Utab.off[(size_t)i];
    for (int i = 0; i < (int)Qtab.n; ++i) xoff[(size_t)i] = Qtab.off[(size_t)i];
  }

  // stage 1 (one task and one segment per product; stage 2: one segment per slot)
  P->s1.tasks.reserve(P->s1.tasks.size() + b.prods.size());
  P->s1.segs.reserve(P->s1.segs.size() + b.prods.size());
  P->s2.segs.reserve(P->s2.segs.size() + b.slots.size());
  for (auto& pr : b.prods) {
    GTask t{};
    t.c_off = pr.ws_off;
    t.m = (int32_t)pr.m;
    t.n = (int32_t)pr.n;
    t.seg0 = (int32_t)P->s1.segs.size();
    t.nseg = 1;
    GSeg s{};
    s.a_off = Loff[(size_t)pr.ia];
    s.b_off = xoff[(size_t)pr.ib];
    s.k = (int32_t)pr.k;
    s.alpha = 1.0;
    if (bf) {
      s.lda = (int32_t)pad8(Ptab.cols[(size_t)pr.ia]);
      s.ldb = (int32_t)pad8(Qtab.rows[(size_t)pr.ib]);
    } else {
      s.lda = (int32_t)std::max<int64_t>(1, Ptab.rows[(size_t)pr.ia]);
      s.ldb = (int32_t)std::max<int64_t>(1, Qtab.rows[(size_t)pr.ib]);
    }
    P->s1.segs.push_back(s);
    P->s1.tasks.push_back(t);
  }
  // heff: NN; env left: NN; env right: NT
  P->s1.tr = (kind == 2 && direction <= 0) ? 1 : 0;

  // mix
  for (auto& sl : b.slots) {
    if (sl.terms.size() == 1) continue;
    MixTask mt{};
    mt.dst_off = sl.ws_off;
    mt.len = sl.m * (bf ? pad8(sl.n) : sl.n);
    mt.t0 = (int32_t)P->terms.size();
    mt.nt = (int32_t)sl.terms.size();
    for (auto& tm : sl.terms) P->terms.push_back(MixTerm{b.prods[(size_t)tm.first].ws_off, tm.second});
    P->mix.push_back(mt);
  }

  // stage 2: one task per produced output block; segments in slot order
  std::vector<std::vector<int>> per_out((size_t)Ytab.n);  // slots per output block, output blocks ascending
  for (int i = 0; i < (int)b.slots.size(); ++i) per_out[(size_t)b.slots[(size_t)i].iout].push_back(i);
  std::vector<char> produced((size_t)Ytab.n, 0);
  for (int iy = 0; iy < (int)Ytab.n; ++iy) {
    const std::vector<int>& kvs = per_out[(size_t)iy];
    if (kvs.empty()) continue;
    produced[(size_t)iy] = 1;
    GTask t{};
    t.c_off = Ytab.off[(size_t)iy];
    t.m = (int32_t)Ytab.rows[(size_t)iy];
    t.n = (int32_t)Ytab.cols[(size_t)iy];
    t.seg0 = (int32_t)P->s2.segs.size();
    t.nseg = (int32_t)kvs.size();
    for (int si : kvs) {
      const T2Slot& sl = b.slots[(size_t)si];
      GSeg s{};
      int64_t toff;
      double alpha;
      if (sl.terms.size() == 1) {
        toff = b.prods[(size_t)sl.terms[0].first].ws_off;
        alpha = sl.terms[0].second;
      } else {
        toff = sl.ws_off;
        alpha = 1.0;
      }
      int64_t roff = Roff[(size_t)sl.iu];
      if (bf && kind != 2 && alpha != 1.0) {  // the coefficient lives in a scaled R copy
        roff = Rscaled.at(std::make_pair((int)sl.iu, alpha));
        alpha = 1.0;
      }
      if (kind != 2) {  // y = T * R^T: A = T (m x n_j), B = R (y.cols x n_j), k = n_j
        s.a_off = toff;
        s.b_off = roff;
        s.k = (int32_t)sl.n;
        s.lda = (int32_t)(bf ? pad8(sl.n) : std::max<int64_t>(1, sl.m));
        s.ldb = (int32_t)(bf ? pad8(Utab.cols[(size_t)sl.iu]) : std::max<int64_t>(1, Utab.rows[(size_t)sl.iu]));
        P->flops_alg += 2.0 * t.m * t.n * sl.n;
      } else {  // out = op(U) * T: A = U, B = T (m_a x n), k = m_a
        s.a_off = Utab.off[(size_t)sl.iu];
        s.b_off = toff;
        s.k = (int32_t)sl.m;
        s.lda = (int32_t)std::max<int64_t>(1, Utab.rows[(size_t)sl.iu]);
        s.ldb = (int32_t)std::max<int64_t>(1, sl.m);
        P->flops_alg += 2.0 * t.m * t.n * sl.m;
      }
      s.alpha = alpha;
      P->s2.segs.push_back(s);
    }
    P->s2.tasks.push_back(t);
  }
  P->flops2 = P->flops_alg - P->flops1;
  if (kind != 2) P->s2.tr = 1;                // N, T
  else P->s2.tr = direction > 0 ? 2 : 0;      // T, N  /  N, N
  for (int i = 0; i < (int)Ytab.n; ++i)
    if (!produced[(size_t)i] && Ytab.rows[(size_t)i] * Ytab.cols[(size_t)i] > 0)
      P->zero.push_back(Range{Ytab.off[(size_t)i], Ytab.rows[(size_t)i] * Ytab.cols[(size_t)i]});

  if (P->cplx) complexify(P, Ptab.total, Qtab.total, Utab.total, Ytab.total, kind);
  finalize_stage(P->s1, P->prec);
  finalize_stage(P->s2, P->prec);
  finalize_mix(P);
  P->rpL_chunks = finalize_repack(P->rpL);
  P->rpR_chunks = finalize_repack(P->rpR);
  P->rpx_chunks = finalize_repack(P->rpx);
  P->esize = P->prec == F64 ? 8 : (P->prec == F32 ? 4 : 2);
  P->planes = P->prec == BF16X3 ? 2 : 1;
}

// Plans are built on the host only; descriptors move to the device on first use.
static void ensure_device(mpsa_plan* P) {
  if (P->on_device) return;
  P->s1.d_tasks = upload(P->s1.tasks);
  P->s1.d_segs = upload(P->s1.segs);
  P->s2.d_tasks = upload(P->s2.tasks);
  P->s2.d_segs = upload(P->s2.segs);
  P->d_mix = upload(P->mix);
  P->d_terms = upload(P->terms);
  P->d_zero = upload(P->zero);
  P->d_rpL = upload(P->rpL);
  P->d_rpR = upload(P->rpR);
  P->d_rpx = upload(P->rpx);
  P->on_device = 1;
}

// Zero n bytes of a fresh device buffer with a kernel on the stream the plan is
// used with (allocations are rounded up to whole doubles), so that the zeroing
// is ordered before the first kernel that reads the buffer on that stream, also
// a non-blocking one, and no memset node can end up in a captured graph.
static void zero_on_stream(void* p, size_t bytes, hipStream_t st) {
  HIPCHK(launch_fill((double*)p, (int64_t)((bytes + 7) / 8), 0.0, st));
}
static size_t round8(size_t bytes) { return (bytes + 7) / 8 * 8; }

static void ensure_ws(mpsa_plan* P, int nvec, hipStream_t st) {
  ensure_device(P);
  if (P->ws_nvec < nvec) {
    if (P->ws) HIPCHK(hipFree(P->ws));
    P->ws = nullptr;
    P->ws_nvec = 0;
    const size_t bytes = round8((size_t)P->ws_per_vec * P->planes * P->esize * (size_t)nvec);
    if (hipMalloc(&P->ws, std::max<size_t>(bytes, 8)) != hipSuccess) {
      P->ws = nullptr;
      throw Error(MPSA_ENOMEM, "workspace allocation failed");
    }
    zero_on_stream(P->ws, bytes, st);  // padding lanes must read as zero
    P->ws_nvec = nvec;
  }
  if (P->prec != F64 && P->xc_nvec < nvec) {
    if (P->xc) HIPCHK(hipFree(P->xc));
    P->xc = nullptr;
    P->xc_nvec = 0;
    const size_t bytes = round8((size_t)P->xc_per_vec * P->planes * P->esize * (size_t)nvec);
    if (hipMalloc(&P->xc, std::max<size_t>(bytes, 8)) != hipSuccess) {
      P->xc = nullptr;
      throw Error(MPSA_ENOMEM, "converted-vector allocation failed");
    }
    zero_on_stream(P->xc, bytes, st);
    P->xc_nvec = nvec;
  }
}

bool plan_complex(const mpsa_plan* P) { return P && P->cplx; }



}  // namespace mpsa

// =================================================================== C ABI
// Re-number a stage's tiles for tile class cls and refresh its device task
// list (same number of tasks, so the allocation is reused).  Kernels already
// queued on `st` read the old list, so the stream drains before the copy, and
// the copy itself is queued on `st` ahead of every later launch.
static void set_stage_class(mpsa_plan* P, Stage& S, int prec, int cls, hipStream_t st) {
  if (S.tile_class == cls) return;
  P->drop_graph();  // the captured launches name the old class's kernel and grid
  if (P->on_device && S.d_tasks && !S.tasks.empty()) HIPCHK(hipStreamSynchronize(st));
  assign_tiles(S, prec, cls);
  if (P->on_device && S.d_tasks && !S.tasks.empty()) {
    HIPCHK(hipMemcpyAsync(S.d_tasks, S.tasks.data(), sizeof(GTask) * S.tasks.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
}

// The grouped GEMM launch of stage 1 (T = L x) or stage 2 (y = T R^T) of an
// H_eff plan, with the operands mpsa_heff_apply gives it.
static void launch_stage(mpsa_plan* P, int which, const double* dx, int64_t x_stride, double* dy, int64_t y_stride,
                         int nvec, hipStream_t st) {
  const int64_t xcs = P->xc_per_vec * P->planes, wss = P->ws_per_vec * P->planes;
  if (which == 1) {
    const void* A1 = P->prec == F64 ? (const void*)P->dL : P->Lc;
    const void* B1 = P->prec == F64 ? (const void*)dx : P->xc;
    Planes pl1;
    pl1.a = P->Lc_elems;
    pl1.b = P->xc_per_vec;
    pl1.c = P->ws_per_vec;
    HIPCHK(launch_gemm(P->prec, 0, P->s1.tr, P->s1.tile_class, P->s1.d_tasks, (int)P->s1.tasks.size(),
                       P->s1.d_segs, P->s1.tiles, nvec, A1, 0, B1, P->prec == F64 ? x_stride : xcs, P->ws, wss, st,
                       pl1));
  } else {
    const void* B2 = P->prec == F64 ? (const void*)P->dR : P->Rc;
    Planes pl2;
    pl2.a = P->ws_per_vec;
    pl2.b = P->Rc_elems;
    HIPCHK(launch_gemm(P->prec, 1, P->s2.tr, P->s2.tile_class, P->s2.d_tasks, (int)P->s2.tasks.size(),
                       P->s2.d_segs, P->s2.tiles, nvec, P->ws, wss, B2, 0, dy, y_stride, st, pl2));
  }
}

#define MPSA_TRY(...)                          \
  try {                                        \
    __VA_ARGS__;                               \
    return MPSA_OK;                            \
  } catch (const mpsa::Error& e) {             \
    mpsa::g_err = e.what();                    \
    return e.code;                             \
  } catch (const std::exception& e) {          \
    mpsa::g_err = e.what();                    \
    return MPSA_EINVAL;                        \
  }

extern "C" {

const char* mpsa_last_error(void) { return mpsa::g_err.c_str(); }
const char* mpsa_version(void) { return "mpstools_amd 0.1 gfx950"; }

static void check_itk(const uint64_t* itk, int d, const char* name) {
  if (!itk || d <= 0) throw Error(MPSA_EINVAL, std::string(name) + ": empty i_to_key");
}

int mpsa_heff2_outkeys(const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x, const mpsa_mpo* Wl,
                       const mpsa_mpo* Wr, const uint64_t* itl, int dl, const uint64_t* itr, int dr, int64_t cap,
                       int64_t* nout, uint64_t* q, uint32_t* p, int64_t* rows, int64_t* cols) {
  MPSA_TRY({
    check_itk(itl, dl, "itokey_l");
    check_itk(itr, dr, "itokey_r");
    Tab tl = parse(L, 1, "L"), tr = parse(R, 1, "R"), tx = parse(x, 2, "x");
    auto gl = parse_mpo(Wl, dl, "Wl"), gr = parse_mpo(Wr, dr, "Wr");
    double fref;
    auto tup = tuples_heff2(tl, tr, tx, gl, gr, itl, itr, &fref);
    std::map<BKey, std::pair<int64_t, int64_t>> dims;
    Built b = group_tuples(tup, nullptr, false, &dims);
    write_keys(b.outkeys, dims, 2, cap, nout, q, p, rows, cols);
  })
}

int mpsa_heff1_outkeys(const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x, const mpsa_mpo* W,
                       const uint64_t* itk, int d, int64_t cap, int64_t* nout, uint64_t* q, uint32_t* p,
                       int64_t* rows, int64_t* cols) {
  MPSA_TRY({
    check_itk(itk, d, "itokey");
    Tab tl = parse(L, 1, "L"), tr = parse(R, 1, "R"), tx = parse(x, 1, "x");
    auto g = parse_mpo(W, d, "W");
    double fref;
    auto tup = tuples_heff1(tl, tr, tx, g, itk, &fref);
    std::map<BKey, std::pair<int64_t, int64_t>> dims;
    Built b = group_tuples(tup, nullptr, false, &dims);
    write_keys(b.outkeys, dims, 1, cap, nout, q, p, rows, cols);
  })
}

int mpsa_env_outkeys(const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap, const mpsa_mpo* W,
                     const uint64_t* itk, int d, int direction, int64_t cap, int64_t* nout, uint64_t* q, uint32_t* p,
                     int64_t* rows, int64_t* cols) {
  MPSA_TRY({
    check_itk(itk, d, "itokey");
    Tab tb = parse(BL, 1, "BL"), ta = parse(A, 1, "A");
    if (ta.n == 0) throw Error(MPSA_EINVAL, "Error in addContractionLayer: A is empty.");
    Tab tap = Ap ? parse(Ap, 1, "Ap") : ta;
    if (Ap && tap.n == 0) throw Error(MPSA_EINVAL, "Error in addContractionLayer: Ap is empty.");
    auto g = parse_mpo(W, d, "W");
    double fref;
    auto tup = tuples_env(tb, ta, tap, g, itk, direction, &fref);
    std::map<BKey, std::pair<int64_t, int64_t>> dims;
    Built b = group_tuples(tup, nullptr, false, &dims);
    write_keys(b.outkeys, dims, 1, cap, nout, q, p, rows, cols);
  })
}

// 1 when every table of a plan is complex128, 0 when none is.
static int complex_tables(std::initializer_list<const Tab*> t) {
  int nc = 0;
  for (const Tab* x : t) nc += x->cplx;
  if (nc == 0) return 0;
  if (nc != (int)t.size()) throw Error(MPSA_EINVAL, "complex plans need every table complex128");
  return 1;
}


static int plan_heff(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                     const mpsa_blocks* y, const std::vector<Tuple>& tup, double fref, int nsx, int precision) {
  (void)nsx;
  if (!plan) throw Error(MPSA_EINVAL, "null plan pointer");
  const bool plain = (precision & MPSA_PLAIN) != 0;
  precision &= ~MPSA_PLAIN;
  if (precision < 0 || precision > 3)
    throw Error(MPSA_EINVAL, "precision must be MPSA_F64, MPSA_F32, MPSA_BF16 or MPSA_BF16X3");
  std::unique_ptr<mpsa_plan> P(new mpsa_plan());
  P->kind = 0;
  P->prec = precision;
  Tab tl = parse(L, 1, "L"), tr = parse(R, 1, "R"), tx = parse(x, x->ns, "x"), ty = parse(y, y->ns, "y");
  P->cplx = complex_tables({&tl, &tr, &tx, &ty});
  P->L_total = tl.total;
  P->R_total = tr.total;
  P->x_total = tx.total;
  P->y_total = ty.total;
  std::map<BKey, std::pair<int64_t, int64_t>> dims;
  Built b = group_tuples(tup, &ty, true, &dims, plain);
  for (auto& k : b.outkeys) {
    int iy = ty.find(k);
    auto dd = dims.at(k);
    if (ty.rows[(size_t)iy] != dd.first || ty.cols[(size_t)iy] != dd.second)
      throw Error(MPSA_EINVAL, "y table block shape differs from the produced block");
  }
  build_plan(P.get(), b, tl, tx, tr, ty, 0, 1);
  P->flops_ref = P->cplx ? 4 * fref : fref;
  P->xy_same = (tx.n == ty.n && tx.ns == ty.ns && tx.rows == ty.rows && tx.cols == ty.cols && tx.off == ty.off &&
                tx.key == ty.key) ? 1 : 0;
  *plan = P.release();
  return MPSA_OK;
}

int mpsa_heff2_plan(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                    const mpsa_blocks* y, const mpsa_mpo* Wl, const mpsa_mpo* Wr, const uint64_t* itl, int dl,
                    const uint64_t* itr, int dr, int precision) {
  MPSA_TRY({
    check_itk(itl, dl, "itokey_l");
    check_itk(itr, dr, "itokey_r");
    if (!x || !y || x->ns != 2 || y->ns != 2) throw Error(MPSA_EINVAL, "heff2: x and y must be two-site tables");
    Tab tl = parse(L, 1, "L"), tr = parse(R, 1, "R"), tx = parse(x, 2, "x");
    auto gl = parse_mpo(Wl, dl, "Wl"), gr = parse_mpo(Wr, dr, "Wr");
    double fref;
    auto tup = tuples_heff2(tl, tr, tx, gl, gr, itl, itr, &fref);
    int rc = plan_heff(plan, L, R, x, y, tup, fref, 2, precision);
    if (rc) return rc;
  })
}

int mpsa_heff1_plan(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                    const mpsa_blocks* y, const mpsa_mpo* W, const uint64_t* itk, int d, int precision) {
  MPSA_TRY({
    check_itk(itk, d, "itokey");
    if (!x || !y || x->ns != 1 || y->ns != 1) throw Error(MPSA_EINVAL, "heff1: x and y must be one-site tables");
    Tab tl = parse(L, 1, "L"), tr = parse(R, 1, "R"), tx = parse(x, 1, "x");
    auto g = parse_mpo(W, d, "W");
    double fref;
    auto tup = tuples_heff1(tl, tr, tx, g, itk, &fref);
    int rc = plan_heff(plan, L, R, x, y, tup, fref, 1, precision);
    if (rc) return rc;
  })
}

int mpsa_env_plan(mpsa_plan** plan, const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap,
                  const mpsa_mpo* W, const uint64_t* itk, int d, int direction, const mpsa_blocks* out,
                  int precision) {
  MPSA_TRY({
    if (!plan) throw Error(MPSA_EINVAL, "null plan pointer");
    if (precision != MPSA_F64) throw Error(MPSA_EINVAL, "environment plans run in FP64 only");
    check_itk(itk, d, "itokey");
    Tab tb = parse(BL, 1, "BL"), ta = parse(A, 1, "A");
    if (ta.n == 0) throw Error(MPSA_EINVAL, "Error in addContractionLayer: A is empty.");
    Tab tap = Ap ? parse(Ap, 1, "Ap") : ta;
    if (Ap && tap.n == 0) throw Error(MPSA_EINVAL, "Error in addContractionLayer: Ap is empty.");
    Tab to = parse(out, 1, "out");
    auto g = parse_mpo(W, d, "W");
    double fref;
    auto tup = tuples_env(tb, ta, tap, g, itk, direction, &fref);
    std::unique_ptr<mpsa_plan> P(new mpsa_plan());
    P->kind = 2;
    P->prec = F64;
    P->direction = direction;
    P->cplx = complex_tables({&tb, &ta, &tap, &to});
    P->L_total = tb.total;
    P->x_total = ta.total;
    P->R_total = tap.total;
    P->y_total = to.total;
    std::map<BKey, std::pair<int64_t, int64_t>> dims;
    Built b = group_tuples(tup, &to, true, &dims);
    for (auto& k : b.outkeys) {
      int io = to.find(k);
      auto dd = dims.at(k);
      if (to.rows[(size_t)io] != dd.first || to.cols[(size_t)io] != dd.second)
        throw Error(MPSA_EINVAL, "out table block shape differs from the produced block");
    }
    build_plan(P.get(), b, tb, ta, tap, to, 2, direction);
    P->flops_ref = P->cplx ? 4 * fref : fref;
    *plan = P.release();
  })
}

int mpsa_plan_bind_env(mpsa_plan* P, const double* dL, const double* dR, mpsa_stream stream) {
  MPSA_TRY({
    if (!P || P->kind != 0) throw Error(MPSA_EINVAL, "bind_env: not an H_eff plan");
    if ((P->L_total > 0 && !dL) || (P->R_total > 0 && !dR)) throw Error(MPSA_EINVAL, "bind_env: null environment");
    hipStream_t st = (hipStream_t)stream;
    P->drop_graph();  // the captured launches name the old environment buffers
    ensure_device(P);
    P->dL = dL;
    P->dR = dR;
    if (P->prec != F64) {
      const size_t lb = (size_t)P->Lc_elems * P->planes * P->esize, rb = (size_t)P->Rc_elems * P->planes * P->esize;
      if (!P->Lc) {
        if (hipMalloc(&P->Lc, round8(lb) + 8) != hipSuccess) {
          P->Lc = nullptr;
          throw Error(MPSA_ENOMEM, "env copy");
        }
        zero_on_stream(P->Lc, round8(lb), st);
      }
      if (!P->Rc) {
        if (hipMalloc(&P->Rc, round8(rb) + 8) != hipSuccess) {
          P->Rc = nullptr;
          throw Error(MPSA_ENOMEM, "env copy");
        }
        zero_on_stream(P->Rc, round8(rb), st);
      }
      HIPCHK(launch_repack(P->prec, P->d_rpL, (int)P->rpL.size(), P->rpL_chunks, 1, dL, 0, P->Lc, 0, st, P->Lc_elems));
      HIPCHK(launch_repack(P->prec, P->d_rpR, (int)P->rpR.size(), P->rpR_chunks, 1, dR, 0, P->Rc, 0, st, P->Rc_elems));
    }
  })
}

// The launches of one H_eff apply: repack (reduced precision), stage 1, mix,
// zeroing of unproduced output blocks, stage 2; ev[1..4] recorded between
// them when timing.
static void apply_launches(mpsa_plan* P, const double* dx, int64_t x_stride, double* dy, int64_t y_stride, int nvec,
                           hipStream_t st, hipEvent_t* ev) {
  const int64_t xcs = P->xc_per_vec * P->planes, wss = P->ws_per_vec * P->planes;
  if (P->prec != F64)
    HIPCHK(launch_repack(P->prec, P->d_rpx, (int)P->rpx.size(), P->rpx_chunks, nvec, dx, x_stride, P->xc, xcs, st,
                         P->xc_per_vec));
  if (ev) HIPCHK(hipEventRecord(ev[1], st));
  launch_stage(P, 1, dx, x_stride, dy, y_stride, nvec, st);
  if (ev) HIPCHK(hipEventRecord(ev[2], st));
  HIPCHK(launch_mix(P->prec, P->d_mix, (int)P->mix.size(), P->d_terms, P->mix_chunks, nvec, P->ws, P->ws, wss, st,
                    P->ws_per_vec));
  HIPCHK(launch_zero_ranges(P->d_zero, (int)P->zero.size(), dy, y_stride, nvec, st));
  if (ev) HIPCHK(hipEventRecord(ev[3], st));
  launch_stage(P, 2, dx, x_stride, dy, y_stride, nvec, st);
  if (ev) HIPCHK(hipEventRecord(ev[4], st));
}

int mpsa_heff_apply(mpsa_plan* P, const double* dx, int64_t x_stride, double* dy, int64_t y_stride, int nvec,
                    mpsa_stream stream) {
  MPSA_TRY({
    if (!P || P->kind != 0) throw Error(MPSA_EINVAL, "apply: not an H_eff plan");
    if (nvec < 1) throw Error(MPSA_EINVAL, "apply: nvec must be >= 1");
    if (!P->dL && P->L_total > 0) throw Error(MPSA_EINVAL, "apply: environments not bound");
    if (!dx || !dy) throw Error(MPSA_EINVAL, "apply: null vector");
    if (nvec > 1 && (x_stride < P->x_total || y_stride < P->y_total))
      throw Error(MPSA_EINVAL, "apply: vector stride smaller than the table extent");
    hipStream_t st = (hipStream_t)stream;
    ensure_ws(P, nvec, st);
    // Untimed applies on a stream of the caller's own run from a graph of
    // their launches (recorded on the second apply in a row with the same
    // arguments, replayed from the third on), so
    // that the dependent launches follow each other without host gaps.  A
    // stream that is already being captured (a caller's graph) or the null
    // stream takes the launches directly.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (st) HIPCHK(hipStreamIsCapturing(st, &cap));
    const bool graphable = !P->timing && st && cap == hipStreamCaptureStatusNone;
    const bool same = P->g_valid && P->g_dx == dx && P->g_dy == dy && P->g_xs == x_stride &&
                      P->g_ys == y_stride && P->g_nvec == nvec && P->g_ws == P->ws && P->g_xc == P->xc;
    if (graphable && P->gexec && same) {
      HIPCHK(hipGraphLaunch(P->gexec, st));
      return MPSA_OK;
    }
    if (graphable && !same) {  // first apply with these arguments: direct launches, remembered
      P->drop_graph();
      P->g_valid = true;
      P->g_dx = dx;
      P->g_dy = dy;
      P->g_xs = x_stride;
      P->g_ys = y_stride;
      P->g_nvec = nvec;
      P->g_ws = P->ws;
      P->g_xc = P->xc;
    } else if (graphable) {  // the second in a row: record the graph
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        apply_launches(P, dx, x_stride, dy, y_stride, nvec, st, nullptr);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamEndCapture(st, &g));
      const hipError_t ie = hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        P->gexec = nullptr;
        throw Error(MPSA_EHIP, std::string("apply: graph instantiate: ") + hipGetErrorString(ie));
      }
      HIPCHK(hipGraphLaunch(P->gexec, st));
      return MPSA_OK;
    }
    hipEvent_t* ev = nullptr;
    if (P->timing) {
      const int slot = std::min(P->nrec, mpsa_plan::kTimedApplies - 1);
      while ((int)P->evs.size() <= slot) {
        std::array<hipEvent_t, 5> set{};
        for (auto& e : set) HIPCHK(hipEventCreate(&e));
        P->evs.push_back(set);
      }
      ev = P->evs[(size_t)slot].data();
      P->cur = slot;
      P->nrec = slot + 1;
      HIPCHK(hipEventRecord(ev[0], st));
    }
    apply_launches(P, dx, x_stride, dy, y_stride, nvec, st, ev);
  })
}

int mpsa_plan_set_tile_class(mpsa_plan* P, int stage, int tile_class) {
  MPSA_TRY({
    if (!P || (stage != 1 && stage != 2)) throw Error(MPSA_EINVAL, "set_tile_class: bad arguments");
    Stage& S = stage == 1 ? P->s1 : P->s2;
    const int prec = P->kind == 2 ? F64 : P->prec;
    if (!class_admissible(S.tasks, S.segs, prec, tile_class))
      throw Error(MPSA_EINVAL, "set_tile_class: tile class not available for this stage");
    // no stream argument: every queued kernel on the device may read the old
    // task list, so the whole device drains first
    if (P->on_device && S.tile_class != tile_class) HIPCHK(hipDeviceSynchronize());
    set_stage_class(P, S, prec, tile_class, nullptr);
  })
}

int mpsa_plan_autotune(mpsa_plan* P, int nvec, int reps, double* ms, mpsa_stream stream) {
  double* x = nullptr;
  double* y = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto release = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
  };
  try {
    if (!P || P->kind != 0) throw Error(MPSA_EINVAL, "autotune: not an H_eff plan");
    if (nvec < 1 || reps < 1) throw Error(MPSA_EINVAL, "autotune: nvec and reps must be >= 1");
    if (!P->dL && P->L_total > 0) throw Error(MPSA_EINVAL, "autotune: environments not bound");
    hipStream_t st = (hipStream_t)stream;
    const int64_t xs = std::max<int64_t>(P->x_total, 1), ys = std::max<int64_t>(P->y_total, 1);
    if (hipMalloc((void**)&x, sizeof(double) * (size_t)(xs * nvec)) != hipSuccess ||
        hipMalloc((void**)&y, sizeof(double) * (size_t)(ys * nvec)) != hipSuccess)
      throw Error(MPSA_ENOMEM, "autotune: vector allocation failed");
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(launch_fill_hash(x, xs * nvec, 0x9e3779b97f4a7c15ull, st));
    const int timing = P->timing;
    P->timing = 0;
    // one full apply leaves converted operands and stage-1 products in place
    // for the stage launches timed below
    const int rc = mpsa_heff_apply(P, x, xs, y, ys, nvec, stream);
    P->timing = timing;
    if (rc != MPSA_OK) throw Error(rc, mpsa_last_error());
    for (int stage = 1; stage <= 2; ++stage) {
      Stage& S = stage == 1 ? P->s1 : P->s2;
      const int cur = S.tile_class;
      int best = -1;
      double best_ms = 0.0;
      for (int cls = 0; cls < kMaxTileClasses; ++cls) {
        if (ms) ms[(stage - 1) * kMaxTileClasses + cls] = -1.0;
        if (!class_admissible(S.tasks, S.segs, P->prec, cls)) continue;
        set_stage_class(P, S, P->prec, cls, st);
        launch_stage(P, stage, x, xs, y, ys, nvec, st);  // warm-up
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) launch_stage(P, stage, x, xs, y, ys, nvec, st);
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        const double per = (double)t / reps;
        if (ms) ms[(stage - 1) * kMaxTileClasses + cls] = per;
        // the strictly fastest class wins; on a tie the plan's own class stays
        if (best < 0 || per < best_ms || (per == best_ms && cls == cur)) {
          best = cls;
          best_ms = per;
        }
      }
      set_stage_class(P, S, P->prec, best, st);
    }
    HIPCHK(hipStreamSynchronize(st));
    release();
    return MPSA_OK;
  } catch (const Error& e) {
    release();
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    release();
    g_err = e.what();
    return MPSA_EINVAL;
  }
}

int mpsa_env_apply(mpsa_plan* P, const double* dBL, const double* dA, const double* dAp, double* dout,
                   mpsa_stream stream) {
  MPSA_TRY({
    if (!P || P->kind != 2) throw Error(MPSA_EINVAL, "env_apply: not an environment plan");
    if (!dBL || !dA || !dout) throw Error(MPSA_EINVAL, "env_apply: null buffer");
    const double* up = dAp ? dAp : dA;
    hipStream_t st = (hipStream_t)stream;
    ensure_ws(P, 1, st);
    HIPCHK(launch_gemm(F64, 0, P->s1.tr, P->s1.tile_class, P->s1.d_tasks, (int)P->s1.tasks.size(), P->s1.d_segs,
                       P->s1.tiles, 1, dBL, 0, dA, 0, P->ws, 0, st));
    HIPCHK(launch_mix(F64, P->d_mix, (int)P->mix.size(), P->d_terms, P->mix_chunks, 1, P->ws, P->ws, 0, st));
    HIPCHK(launch_zero_ranges(P->d_zero, (int)P->zero.size(), dout, 0, 1, st));
    HIPCHK(launch_gemm(F64, 1, P->s2.tr, P->s2.tile_class, P->s2.d_tasks, (int)P->s2.tasks.size(), P->s2.d_segs,
                       P->s2.tiles, 1, up, 0, P->ws, 0, dout, 0, st));
  })
}

int mpsa_plan_info(const mpsa_plan* P, double* flops_alg, double* flops_ref, int64_t* n_gemm1, int64_t* n_gemm2,
                   int64_t* workspace_bytes) {
  MPSA_TRY({
    if (!P) throw Error(MPSA_EINVAL, "null plan");
    if (flops_alg) *flops_alg = P->flops_alg;
    if (flops_ref) *flops_ref = P->flops_ref;
    if (n_gemm1) *n_gemm1 = (int64_t)P->s1.tasks.size();
    if (n_gemm2) *n_gemm2 = (int64_t)P->s2.segs.size();
    if (workspace_bytes) *workspace_bytes = P->ws_per_vec * (int64_t)P->esize;
  })
}

int mpsa_plan_vector_lengths(const mpsa_plan* P, int64_t* x_len, int64_t* y_len, int* same_layout) {
  MPSA_TRY({
    if (!P) throw Error(MPSA_EINVAL, "null plan");
    if (x_len) *x_len = P->x_total;
    if (y_len) *y_len = P->y_total;
    if (same_layout) *same_layout = P->xy_same;
  })
}

int mpsa_plan_stage_info(const mpsa_plan* P, double* flops_stage1, double* flops_stage2, int* tile_class1,
                         int* tile_class2) {
  MPSA_TRY({
    if (!P) throw Error(MPSA_EINVAL, "null plan");
    if (flops_stage1) *flops_stage1 = P->flops1;
    if (flops_stage2) *flops_stage2 = P->flops2;
    if (tile_class1) *tile_class1 = P->s1.tile_class;
    if (tile_class2) *tile_class2 = P->s2.tile_class;
  })
}

int mpsa_plan_describe(const mpsa_plan* P, int what, void* buf, int64_t cap, int64_t* n, int* extra) {
  MPSA_TRY({
    if (!P || !n) throw Error(MPSA_EINVAL, "null argument");
    const void* src = nullptr;
    size_t es = 0, cnt = 0;
    int ex = 0;
    switch (what) {
      case 0: src = P->s1.tasks.data(); es = sizeof(GTask); cnt = P->s1.tasks.size(); ex = P->s1.tr; break;
      case 1: src = P->s1.segs.data(); es = sizeof(GSeg); cnt = P->s1.segs.size(); ex = P->s1.tile_class; break;
      case 2: src = P->s2.tasks.data(); es = sizeof(GTask); cnt = P->s2.tasks.size(); ex = P->s2.tr; break;
      case 3: src = P->s2.segs.data(); es = sizeof(GSeg); cnt = P->s2.segs.size(); ex = P->s2.tile_class; break;
      case 4: src = P->mix.data(); es = sizeof(MixTask); cnt = P->mix.size(); ex = P->prec; break;
      case 5: src = P->terms.data(); es = sizeof(MixTerm); cnt = P->terms.size(); break;
      case 6: src = P->zero.data(); es = sizeof(Range); cnt = P->zero.size(); break;
      default: throw Error(MPSA_EINVAL, "describe: unknown table");
    }
    *n = (int64_t)cnt;
    if (extra) *extra = ex;
    if (cap > 0) {
      if (!buf || cap < (int64_t)(cnt * es)) throw Error(MPSA_EINVAL, "describe: buffer too small");
      if (cnt) std::memcpy(buf, src, cnt * es);
    }
  })
}

int mpsa_plan_set_timing(mpsa_plan* P, int enable) {
  MPSA_TRY({
    if (!P) throw Error(MPSA_EINVAL, "null plan");
    P->timing = enable ? 1 : 0;
    P->nrec = 0;
    P->cur = -1;
  })
}

int mpsa_plan_last_times(const mpsa_plan* P, float* ms) {
  MPSA_TRY({
    if (!P || !ms) throw Error(MPSA_EINVAL, "null argument");
    if (P->cur < 0) throw Error(MPSA_EINVAL, "no timed apply on this plan");
    const auto& ev = P->evs[(size_t)P->cur];
    HIPCHK(hipEventSynchronize(ev[4]));
    for (int i = 0; i < 4; ++i) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
  })
}

int mpsa_plan_timing_summary(mpsa_plan* P, float* ms, int* napplies) {
  MPSA_TRY({
    if (!P || !ms || !napplies) throw Error(MPSA_EINVAL, "null argument");
    const int n = P->nrec;
    for (int i = 0; i < 4; ++i) ms[i] = 0.f;
    if (n > 0) HIPCHK(hipEventSynchronize(P->evs[(size_t)(n - 1)][4]));
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < 4; ++i) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, P->evs[(size_t)a][(size_t)i], P->evs[(size_t)a][(size_t)i + 1]));
        ms[i] += t / (float)n;
      }
    *napplies = n;
    P->nrec = 0;
  })
}

int mpsa_plan_destroy(mpsa_plan* P) {
  MPSA_TRY({ delete P; })
}

static double* g_partial = nullptr;
static double* g_dotout = nullptr;
static const int kDotParts = 1024;

int mpsa_dot(const double* dx, const double* dy, int64_t n, int is_complex, double* out, mpsa_stream stream) {
  MPSA_TRY({
    if (n < 0 || !out || (n > 0 && (!dx || !dy))) throw Error(MPSA_EINVAL, "dot: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    if (!g_partial) {
      HIPCHK(hipMalloc((void**)&g_partial, sizeof(double) * 2 * kDotParts));
      HIPCHK(hipMalloc((void**)&g_dotout, sizeof(double) * 2));
    }
    HIPCHK(launch_dot(dx, dy, n, is_complex, g_partial, kDotParts, g_dotout, st));
    HIPCHK(hipMemcpyAsync(out, g_dotout, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  })
}

int mpsa_fill_gauss(double* dx, int64_t n, uint64_t seed, mpsa_stream stream) {
  MPSA_TRY({
    if (n < 0 || (n > 0 && !dx)) throw Error(MPSA_EINVAL, "fill_gauss: bad arguments");
    HIPCHK(launch_fill_gauss(dx, n, seed, (hipStream_t)stream));
  })
}

int mpsa_axpby(const double* a, const double* dx, const double* b, double* dy, int64_t n, int is_complex,
               mpsa_stream stream) {
  MPSA_TRY({
    if (n < 0 || !a || !b || (n > 0 && (!dx || !dy))) throw Error(MPSA_EINVAL, "axpby: bad arguments");
    HIPCHK(launch_axpby(a[0], is_complex ? a[1] : 0.0, dx, b[0], is_complex ? b[1] : 0.0, dy, n, is_complex,
                        (hipStream_t)stream));
  })
}

}  // extern "C"

namespace mpsa {

// The produced keys of a structure handed to the caller, who returns the
// output table (its own block layout) that the plan is then built against;
// one tuple enumeration and one key pass for both (see mpsa_internal.h).
static mpsa_blocks keys_to_table(const std::vector<BKey>& keys,
                                 const std::map<BKey, std::pair<int64_t, int64_t>>& dims, const OutTableFn& make_out) {
  const int64_t n = (int64_t)keys.size();
  std::vector<uint64_t> q((size_t)2 * n + 2);
  std::vector<uint32_t> p((size_t)n + 1);
  std::vector<int64_t> r((size_t)n + 1), c((size_t)n + 1);
  int64_t nout = 0;
  write_keys(keys, dims, 1, n, &nout, q.data(), p.data(), r.data(), c.data());
  return make_out(n, q.data(), p.data(), r.data(), c.data());
}

static void env_keys_plan_impl(mpsa_plan** plan, const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap,
                               const mpsa_mpo* W, const uint64_t* itk, int d, int direction,
                               const OutTableFn& make_out) {
  if (!plan) throw Error(MPSA_EINVAL, "null plan pointer");
  *plan = nullptr;
  check_itk(itk, d, "itokey");
  Tab tb = parse(BL, 1, "BL"), ta = parse(A, 1, "A");
  if (ta.n == 0) throw Error(MPSA_EINVAL, "Error in addContractionLayer: A is empty.");
  Tab tap = Ap ? parse(Ap, 1, "Ap") : ta;
  if (Ap && tap.n == 0) throw Error(MPSA_EINVAL, "Error in addContractionLayer: Ap is empty.");
  auto g = parse_mpo(W, d, "W");
  double fref;
  auto tup = tuples_env(tb, ta, tap, g, itk, direction, &fref);
  std::map<BKey, std::pair<int64_t, int64_t>> dims;
  Built b;
  produced_keys(tup, b.outkeys, &dims);
  const mpsa_blocks out = keys_to_table(b.outkeys, dims, make_out);
  if (b.outkeys.empty()) return;
  Tab to = parse(&out, 1, "out");
  std::unique_ptr<mpsa_plan> P(new mpsa_plan());
  P->kind = 2;
  P->prec = F64;
  P->direction = direction;
  P->cplx = complex_tables({&tb, &ta, &tap, &to});
  P->L_total = tb.total;
  P->x_total = ta.total;
  P->R_total = tap.total;
  P->y_total = to.total;
  group_products(tup, &to, false, b);
  build_plan(P.get(), b, tb, ta, tap, to, 2, direction);
  P->flops_ref = P->cplx ? 4 * fref : fref;
  *plan = P.release();
}

static void heff1_keys_plan_impl(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R,
                                 const mpsa_blocks* x, const mpsa_mpo* W, const uint64_t* itk, int d, int precision,
                                 const OutTableFn& make_out) {
  if (!plan) throw Error(MPSA_EINVAL, "null plan pointer");
  *plan = nullptr;
  check_itk(itk, d, "itokey");
  if (!x || x->ns != 1) throw Error(MPSA_EINVAL, "heff1: x must be a one-site table");
  if (precision < 0 || precision > 3)
    throw Error(MPSA_EINVAL, "precision must be MPSA_F64, MPSA_F32, MPSA_BF16 or MPSA_BF16X3");
  Tab tl = parse(L, 1, "L"), tr = parse(R, 1, "R"), tx = parse(x, 1, "x");
  auto g = parse_mpo(W, d, "W");
  double fref;
  auto tup = tuples_heff1(tl, tr, tx, g, itk, &fref);
  std::map<BKey, std::pair<int64_t, int64_t>> dims;
  Built b;
  produced_keys(tup, b.outkeys, &dims);
  const mpsa_blocks y = keys_to_table(b.outkeys, dims, make_out);
  if (b.outkeys.empty()) return;
  Tab ty = parse(&y, 1, "y");
  std::unique_ptr<mpsa_plan> P(new mpsa_plan());
  P->kind = 0;
  P->prec = precision;
  P->cplx = complex_tables({&tl, &tr, &tx, &ty});
  P->L_total = tl.total;
  P->R_total = tr.total;
  P->x_total = tx.total;
  P->y_total = ty.total;
  group_products(tup, &ty, false, b);
  build_plan(P.get(), b, tl, tx, tr, ty, 0, 1);
  P->flops_ref = P->cplx ? 4 * fref : fref;
  P->xy_same = (tx.n == ty.n && tx.ns == ty.ns && tx.rows == ty.rows && tx.cols == ty.cols && tx.off == ty.off &&
                tx.key == ty.key) ? 1 : 0;
  *plan = P.release();
}

int env_keys_plan(mpsa_plan** plan, const mpsa_blocks* BL, const mpsa_blocks* A, const mpsa_blocks* Ap,
                  const mpsa_mpo* W, const uint64_t* itk, int d, int direction, const OutTableFn& make_out) {
  MPSA_TRY(env_keys_plan_impl(plan, BL, A, Ap, W, itk, d, direction, make_out))
}

int heff1_keys_plan(mpsa_plan** plan, const mpsa_blocks* L, const mpsa_blocks* R, const mpsa_blocks* x,
                    const mpsa_mpo* W, const uint64_t* itk, int d, int precision, const OutTableFn& make_out) {
  MPSA_TRY(heff1_keys_plan_impl(plan, L, R, x, W, itk, d, precision, make_out))
}

}  // namespace mpsa
