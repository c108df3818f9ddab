// emu.cpp -- CPU emulation of the device layer under the host drivers of
// libmpstools_amd (test infrastructure only; never part of the gfx950 library).
//
// Linked with the library's own host sources (mpsa_plan.cpp, mpsa_dmrg.cpp,
// mpsa_split.cpp, mpsa_lanczos.cpp, mpsa_adjoint.cpp) in place of the HIP
// runtime, the gfx950 kernels (mpsa_kernels.hip) and rocSOLVER, so that the
// host logic -- plan building, descriptor upload, buffer lifetimes, the
// sweep's control flow -- runs unchanged on the CPU:
//   * device memory is host memory registered per allocation, filled with a
//     NaN byte pattern (reads of never-written memory poison the results);
//     freed allocations are kept, poisoned, so later accesses are reported as
//     use after free;
//   * each launcher computes what its kernel computes and checks every address
//     the kernel touches (the FP64/FP32 GEMM kernel clamps its loads into the
//     logical block, so its physical and logical extents agree; a segment with
//     k <= 0 or a tile of an empty task, whose clamped address would be -1, is
//     reported);
//   * streams run synchronously; stream capture records the launches into a
//     graph that hipGraphLaunch replays;
//   * rocsolver_dgesvd / dgesvdj check their documented array extents and
//     call LAPACK dgesvd (loaded at run time, like the oracle's BLAS).
// The BF16 kernels are not emulated (their launch is reported).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <dlfcn.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "mpsa_internal.h"

namespace {
struct Alloc {
  size_t bytes;
  bool live;
  long serial;
};
std::map<uintptr_t, Alloc> g_alloc;
long g_serial = 0, g_viol = 0, g_launch = 0;
std::string g_ctx = "-";
std::vector<std::string> g_msgs;

void report(const std::string& m) {
  ++g_viol;
  if (g_msgs.size() < 200) g_msgs.push_back("[launch " + std::to_string(g_launch) + " " + g_ctx + "] " + m);
  if (g_viol <= 20) fprintf(stderr, "EMU VIOLATION [launch %ld %s] %s\n", g_launch, g_ctx.c_str(), m.c_str());
}

// [p, p + bytes) inside one live allocation?
bool check(const void* p, size_t bytes, const char* what) {
  if (bytes == 0) return true;
  const uintptr_t a = (uintptr_t)p;
  auto it = g_alloc.upper_bound(a);
  if (it != g_alloc.begin()) {
    --it;
    const uintptr_t b = it->first;
    if (a >= b && a + bytes <= b + it->second.bytes) {
      if (!it->second.live) {
        report(std::string(what) + ": use after free (alloc #" + std::to_string(it->second.serial) + ")");
        return false;
      }
      return true;
    }
    if (a >= b && a < b + it->second.bytes + 4096) {
      char buf[256];
      snprintf(buf, sizeof buf, "%s: out of bounds, offset %ld bytes %zu in alloc #%ld of %zu bytes%s", what,
               (long)(a - b), bytes, it->second.serial, it->second.bytes, it->second.live ? "" : " (freed)");
      report(buf);
      return false;
    }
  }
  auto nx = g_alloc.lower_bound(a);
  if (nx != g_alloc.end() && nx->first < a + 4096 + bytes) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: before the start of alloc #%ld by %ld bytes", what, nx->second.serial,
             (long)(nx->first - a));
    report(buf);
    return false;
  }
  report(std::string(what) + ": not device memory");
  return false;
}
template <typename T>
bool chk_range(const T* base, int64_t lo, int64_t hi_incl, const char* what) {  // elements [lo, hi]
  if (hi_incl < lo) return true;
  return check(base + lo, sizeof(T) * (size_t)(hi_incl - lo + 1), what);
}
bool is_dev(const void* p) {
  const uintptr_t a = (uintptr_t)p;
  auto it = g_alloc.upper_bound(a);
  if (it == g_alloc.begin()) return false;
  --it;
  return a >= it->first && a < it->first + it->second.bytes;
}

// streams, capture, graphs
struct Stream {
  bool capturing = false;
  std::vector<std::function<void()>> cap;
};
struct Graph {
  std::vector<std::function<void()>> ops;
};
Stream g_null;
Stream* S(hipStream_t s) { return s ? reinterpret_cast<Stream*>(s) : &g_null; }
void submit(hipStream_t st, const char* name, std::function<void()> f) {
  Stream* s = S(st);
  std::string nm = name;
  auto wrapped = [f, nm]() {
    ++g_launch;
    g_ctx = nm;
    f();
    g_ctx = "-";
  };
  if (s->capturing) s->cap.push_back(wrapped);
  else wrapped();
}
}  // namespace

extern "C" {
long emu_violations() { return g_viol; }
const char* emu_message(int i) { return i < (int)g_msgs.size() ? g_msgs[(size_t)i].c_str() : nullptr; }
long emu_launches() { return g_launch; }
long emu_live_bytes() {
  long b = 0;
  for (auto& kv : g_alloc)
    if (kv.second.live) b += (long)kv.second.bytes;
  return b;
}
}

// ---------------------------------------------------------------------------
// HIP runtime
// ---------------------------------------------------------------------------
hipError_t hipMalloc(void** p, size_t bytes) {
  if (!p) return hipErrorInvalidValue;
  const size_t b = bytes ? bytes : 1;
  void* m = aligned_alloc(256, (b + 255) / 256 * 256);
  memset(m, 0xff, (b + 255) / 256 * 256);  // NaN pattern: reads of never-written memory show up
  g_alloc[(uintptr_t)m] = Alloc{b, true, ++g_serial};
  *p = m;
  return hipSuccess;
}
// pinned host memory is ordinary host memory here
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned int) {
  if (!p) return hipErrorInvalidValue;
  *p = aligned_alloc(256, ((bytes ? bytes : 1) + 255) / 256 * 256);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  auto it = g_alloc.find((uintptr_t)p);
  if (it == g_alloc.end()) {
    report("hipFree of a pointer that is not an allocation");
    return hipErrorInvalidValue;
  }
  if (!it->second.live) {
    report("double hipFree");
    return hipErrorInvalidValue;
  }
  it->second.live = false;
  memset(p, 0xfe, it->second.bytes);
  return hipSuccess;
}
static hipError_t emu_copy(void* dst, const void* src, size_t n, hipMemcpyKind k) {
  const bool dd = k == hipMemcpyDeviceToDevice || k == hipMemcpyHostToDevice || (k == hipMemcpyDefault && is_dev(dst));
  const bool sd = k == hipMemcpyDeviceToDevice || k == hipMemcpyDeviceToHost || (k == hipMemcpyDefault && is_dev(src));
  bool ok = true;
  if (dd) ok = check(dst, n, "memcpy dst") && ok;
  if (sd) ok = check(src, n, "memcpy src") && ok;
  if (ok && n) memmove(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind k) {
  if (g_null.capturing) return hipErrorStreamCaptureUnsupported;
  return emu_copy(dst, src, n, k);
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind k, hipStream_t st) {
  submit(st, "memcpyAsync", [=]() { (void)emu_copy(dst, src, n, k); });
  return hipSuccess;
}
hipError_t hipMemset(void* p, int v, size_t n) {
  if (check(p, n, "memset")) memset(p, v, n);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = reinterpret_cast<hipStream_t>(new Stream());
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  if (S(s) != &g_null) delete S(s);
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  S(s)->capturing = true;
  S(s)->cap.clear();
  return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus* st) {
  *st = S(s)->capturing ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  Graph* G = new Graph();
  G->ops = S(s)->cap;
  S(s)->cap.clear();
  S(s)->capturing = false;
  *g = reinterpret_cast<hipGraph_t>(G);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  Graph* G = new Graph(*reinterpret_cast<Graph*>(g));
  *e = reinterpret_cast<hipGraphExec_t>(G);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t st) {
  for (auto& f : reinterpret_cast<Graph*>(e)->ops) {
    if (S(st)->capturing) S(st)->cap.push_back(f);
    else f();
  }
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete reinterpret_cast<Graph*>(e);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete reinterpret_cast<Graph*>(g);
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(new int(0));
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<int*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated HIP error"; }

// ---------------------------------------------------------------------------
// rocBLAS handle and rocSOLVER dgesvd (LAPACK dgesvd underneath)
// ---------------------------------------------------------------------------
typedef void (*dgesvd_fn)(const char*, const char*, const int*, const int*, double*, const int*, double*, double*,
                          const int*, double*, const int*, double*, const int*, int*);
// The first LAPACK library found (MKL or the system LAPACK)
static void* lapack_lib() {
  static void* lib = nullptr;
  static bool tried = false;
  if (!tried) {
    tried = true;
    const char* cands[] = {getenv("EMU_LAPACK"), "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1",
                           "liblapack.so.3", "libopenblas.so.0"};
    for (const char* c : cands) {
      if (!c) continue;
      void* h = dlopen(c, RTLD_NOW | RTLD_LOCAL);
      if (h && dlsym(h, "dgesvd_")) {
        // MKL: the GNU threading layer, as the oracle asks for (the Intel
        // layer beside the GNU OpenMP runtime makes the oracle's loops racy)
        typedef int (*set_layer_fn)(int);
        if (set_layer_fn sl = (set_layer_fn)dlsym(h, "MKL_Set_Threading_Layer")) sl(3);
        lib = h;
        break;
      }
    }
  }
  return lib;
}
static dgesvd_fn lapack_dgesvd() {
  void* h = lapack_lib();
  return h ? (dgesvd_fn)dlsym(h, "dgesvd_") : nullptr;
}
rocblas_status rocblas_create_handle(rocblas_handle* h) {
  *h = reinterpret_cast<rocblas_handle>(new int(1));
  return rocblas_status_success;
}
rocblas_status rocblas_set_stream(rocblas_handle, hipStream_t) { return rocblas_status_success; }
rocblas_status rocsolver_dgesvd(rocblas_handle, const rocblas_svect ls, const rocblas_svect rs, const rocblas_int m,
                                const rocblas_int n, double* A, const rocblas_int lda, double* Sg, double* U,
                                const rocblas_int ldu, double* V, const rocblas_int ldv, double* E,
                                const rocblas_workmode, rocblas_int* info);
// gesvdj: the same arrays as gesvd without E, plus a residual and a sweep
// count on the device (computed here by LAPACK dgesvd)
rocblas_status rocsolver_dgesvdj(rocblas_handle h, const rocblas_svect ls, const rocblas_svect rs,
                                 const rocblas_int m, const rocblas_int n, double* A, const rocblas_int lda,
                                 const double, double* residual, const rocblas_int, rocblas_int* n_sweeps,
                                 double* Sg, double* U, const rocblas_int ldu, double* V, const rocblas_int ldv,
                                 rocblas_int* info) {
  submit(nullptr, "rocsolver_dgesvdj outputs", [=]() {
    if (check(residual, 8, "gesvdj residual")) *residual = 0.0;
    if (check(n_sweeps, sizeof(rocblas_int), "gesvdj sweeps")) *n_sweeps = 1;
  });
  return rocsolver_dgesvd(h, ls, rs, m, n, A, lda, Sg, U, ldu, V, ldv, nullptr, rocblas_outofplace, info);
}
rocblas_status rocsolver_dgesvd(rocblas_handle, const rocblas_svect ls, const rocblas_svect rs, const rocblas_int m,
                                const rocblas_int n, double* A, const rocblas_int lda, double* Sg, double* U,
                                const rocblas_int ldu, double* V, const rocblas_int ldv, double* E,
                                const rocblas_workmode, rocblas_int* info) {
  submit(nullptr, "rocsolver_dgesvd", [=]() {
    const int k = std::min(m, n);
    if (ls != rocblas_svect_singular || rs != rocblas_svect_singular || m < 0 || n < 0 || lda < std::max(1, m) ||
        ldu < std::max(1, m) || ldv < std::max(1, k)) {
      report("gesvd: unsupported or inconsistent arguments");
      return;
    }
    bool ok = chk_range(A, 0, (int64_t)lda * (n - 1) + m - 1, "gesvd A");
    ok = chk_range(Sg, 0, k - 1, "gesvd S") && ok;
    ok = chk_range(U, 0, (int64_t)ldu * (k - 1) + m - 1, "gesvd U") && ok;
    ok = chk_range(V, 0, (int64_t)ldv * (n - 1) + k - 1, "gesvd V") && ok;
    if (k > 1 && E) ok = chk_range(E, 0, k - 2, "gesvd E") && ok;
    ok = check(info, sizeof(rocblas_int), "gesvd info") && ok;
    if (!ok || k == 0) return;
    dgesvd_fn dgesvd_ = lapack_dgesvd();
    if (!dgesvd_) {
      report("gesvd: no LAPACK library found");
      return;
    }
    int M = m, N = n, LDA = lda, LDU = ldu, LDV = ldv, lwork = -1, inf = 0;
    double wq = 0;
    dgesvd_("S", "S", &M, &N, A, &LDA, Sg, U, &LDU, V, &LDV, &wq, &lwork, &inf);
    lwork = (int)wq;
    std::vector<double> work((size_t)std::max(lwork, 1));
    dgesvd_("S", "S", &M, &N, A, &LDA, Sg, U, &LDU, V, &LDV, work.data(), &lwork, &inf);
    *info = inf;
  });
  return rocblas_status_success;
}

// rocSOLVER QR / LQ / complex SVD by LAPACK (the engine's large-sector path);
// every array is checked against its documented extent first
template <typename F>
static F lapack_fn(const char* name) {
  void* h = lapack_lib();
  return h ? (F)dlsym(h, name) : nullptr;
}
typedef void (*geqrf_fn)(const int*, const int*, void*, const int*, void*, void*, const int*, int*);
typedef void (*orgqr_fn)(const int*, const int*, const int*, void*, const int*, const void*, void*, const int*, int*);
static void qr_like(const char* what, const char* fname, int m, int n, void* A, int lda, void* tau, int esz,
                    int ktau) {
  submit(nullptr, what, [=]() {
    if (m < 0 || n < 0 || lda < std::max(1, m)) {
      report("qr: inconsistent arguments");
      return;
    }
    bool ok = check(A, (size_t)esz * ((size_t)lda * (n - 1) + m), "qr A");
    ok = check(tau, (size_t)esz * std::max(ktau, 1), "qr tau") && ok;
    if (!ok || m == 0 || n == 0) return;
    geqrf_fn f = lapack_fn<geqrf_fn>(fname);
    if (!f) {
      report("qr: no LAPACK library found");
      return;
    }
    int M = m, N = n, LDA = lda, lw = -1, info = 0;
    double wq[2] = {0, 0};
    f(&M, &N, A, &LDA, tau, wq, &lw, &info);
    lw = std::max(1, (int)wq[0]);
    std::vector<double> work((size_t)2 * lw);
    f(&M, &N, A, &LDA, tau, work.data(), &lw, &info);
    if (info) report("qr: LAPACK info");
  });
}
static void orgq_like(const char* what, const char* fname, int m, int n, int k, void* A, int lda, void* tau, int esz) {
  submit(nullptr, what, [=]() {
    if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m)) {
      report("orgq: inconsistent arguments");
      return;
    }
    bool ok = check(A, (size_t)esz * ((size_t)lda * (n - 1) + m), "orgq A");
    ok = check(tau, (size_t)esz * std::max(k, 1), "orgq tau") && ok;
    if (!ok || m == 0 || n == 0) return;
    orgqr_fn f = lapack_fn<orgqr_fn>(fname);
    if (!f) {
      report("orgq: no LAPACK library found");
      return;
    }
    int M = m, N = n, K = k, LDA = lda, lw = -1, info = 0;
    double wq[2] = {0, 0};
    f(&M, &N, &K, A, &LDA, tau, wq, &lw, &info);
    lw = std::max(1, (int)wq[0]);
    std::vector<double> work((size_t)2 * lw);
    f(&M, &N, &K, A, &LDA, tau, work.data(), &lw, &info);
    if (info) report("orgq: LAPACK info");
  });
}
rocblas_status rocsolver_dgeqrf(rocblas_handle, const rocblas_int m, const rocblas_int n, double* A,
                                const rocblas_int lda, double* ipiv) {
  qr_like("rocsolver_dgeqrf", "dgeqrf_", m, n, A, lda, ipiv, 8, std::min(m, n));
  return rocblas_status_success;
}
rocblas_status rocsolver_zgeqrf(rocblas_handle, const rocblas_int m, const rocblas_int n, rocblas_double_complex* A,
                                const rocblas_int lda, rocblas_double_complex* ipiv) {
  qr_like("rocsolver_zgeqrf", "zgeqrf_", m, n, A, lda, ipiv, 16, std::min(m, n));
  return rocblas_status_success;
}
rocblas_status rocsolver_dgelqf(rocblas_handle, const rocblas_int m, const rocblas_int n, double* A,
                                const rocblas_int lda, double* ipiv) {
  qr_like("rocsolver_dgelqf", "dgelqf_", m, n, A, lda, ipiv, 8, std::min(m, n));
  return rocblas_status_success;
}
rocblas_status rocsolver_zgelqf(rocblas_handle, const rocblas_int m, const rocblas_int n, rocblas_double_complex* A,
                                const rocblas_int lda, rocblas_double_complex* ipiv) {
  qr_like("rocsolver_zgelqf", "zgelqf_", m, n, A, lda, ipiv, 16, std::min(m, n));
  return rocblas_status_success;
}
rocblas_status rocsolver_dorgqr(rocblas_handle, const rocblas_int m, const rocblas_int n, const rocblas_int k,
                                double* A, const rocblas_int lda, double* ipiv) {
  orgq_like("rocsolver_dorgqr", "dorgqr_", m, n, k, A, lda, ipiv, 8);
  return rocblas_status_success;
}
rocblas_status rocsolver_zungqr(rocblas_handle, const rocblas_int m, const rocblas_int n, const rocblas_int k,
                                rocblas_double_complex* A, const rocblas_int lda, rocblas_double_complex* ipiv) {
  orgq_like("rocsolver_zungqr", "zungqr_", m, n, k, A, lda, ipiv, 16);
  return rocblas_status_success;
}
rocblas_status rocsolver_dorglq(rocblas_handle, const rocblas_int m, const rocblas_int n, const rocblas_int k,
                                double* A, const rocblas_int lda, double* ipiv) {
  orgq_like("rocsolver_dorglq", "dorglq_", m, n, k, A, lda, ipiv, 8);
  return rocblas_status_success;
}
rocblas_status rocsolver_zunglq(rocblas_handle, const rocblas_int m, const rocblas_int n, const rocblas_int k,
                                rocblas_double_complex* A, const rocblas_int lda, rocblas_double_complex* ipiv) {
  orgq_like("rocsolver_zunglq", "zunglq_", m, n, k, A, lda, ipiv, 16);
  return rocblas_status_success;
}
rocblas_status rocsolver_zgesvd(rocblas_handle, const rocblas_svect ls, const rocblas_svect rs, const rocblas_int m,
                                const rocblas_int n, rocblas_double_complex* A, const rocblas_int lda, double* Sg,
                                rocblas_double_complex* U, const rocblas_int ldu, rocblas_double_complex* V,
                                const rocblas_int ldv, double* E, const rocblas_workmode, rocblas_int* info) {
  submit(nullptr, "rocsolver_zgesvd", [=]() {
    const int k = std::min(m, n);
    if (ls != rocblas_svect_singular || rs != rocblas_svect_singular || m < 0 || n < 0 || lda < std::max(1, m) ||
        ldu < std::max(1, m) || ldv < std::max(1, k)) {
      report("zgesvd: unsupported or inconsistent arguments");
      return;
    }
    bool ok = check(A, 16 * ((size_t)lda * (n - 1) + m), "zgesvd A");
    ok = chk_range(Sg, 0, k - 1, "zgesvd S") && ok;
    ok = check(U, 16 * ((size_t)ldu * (k - 1) + m), "zgesvd U") && ok;
    ok = check(V, 16 * ((size_t)ldv * (n - 1) + k), "zgesvd V") && ok;
    if (k > 1 && E) ok = chk_range(E, 0, k - 2, "zgesvd E") && ok;
    ok = check(info, sizeof(rocblas_int), "zgesvd info") && ok;
    if (!ok || k == 0) return;
    typedef void (*zgesvd_fn)(const char*, const char*, const int*, const int*, void*, const int*, double*, void*,
                              const int*, void*, const int*, void*, const int*, double*, int*);
    zgesvd_fn f = lapack_fn<zgesvd_fn>("zgesvd_");
    if (!f) {
      report("zgesvd: no LAPACK library found");
      return;
    }
    int M = m, N = n, LDA = lda, LDU = ldu, LDV = ldv, lw = -1, inf = 0;
    std::vector<double> rwork((size_t)5 * k + 8);
    double wq[2] = {0, 0};
    f("S", "S", &M, &N, A, &LDA, Sg, U, &LDU, V, &LDV, wq, &lw, rwork.data(), &inf);
    lw = std::max(1, (int)wq[0]);
    std::vector<double> work((size_t)2 * lw);
    f("S", "S", &M, &N, A, &LDA, Sg, U, &LDU, V, &LDV, work.data(), &lw, rwork.data(), &inf);
    *info = inf;
  });
  return rocblas_status_success;
}
namespace mpsa {
hipError_t launch_pack_z(const double* re, const double* im, int64_t n, double* z, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "pack_z", [=]() {
    if (!chk_range(re, 0, n - 1, "pack re") || !chk_range(im, 0, n - 1, "pack im") ||
        !chk_range(z, 0, 2 * n - 1, "pack z"))
      return;
    for (int64_t i = 0; i < n; ++i) {
      z[2 * i] = re[i];
      z[2 * i + 1] = im[i];
    }
  });
  return hipSuccess;
}
hipError_t launch_unpack_mat(const double* A, int cplx, int64_t lda, int64_t rows, int64_t cols, int mode,
                             double* out, int64_t ldo, int64_t plane, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  submit(st, "unpack_mat", [=]() {
    const int64_t last = (rows - 1) + (cols - 1) * lda;
    if (!chk_range(A, 0, cplx ? 2 * last + 1 : last, "unpack A") ||
        !chk_range(out, 0, (rows - 1) + (cols - 1) * ldo, "unpack out") ||
        (cplx && !chk_range(out + plane, 0, (rows - 1) + (cols - 1) * ldo, "unpack out im")))
      return;
    for (int64_t j = 0; j < cols; ++j)
      for (int64_t i = 0; i < rows; ++i) {
        const bool keep = mode == 0 || (mode == 1 && i <= j) || (mode == 2 && i >= j);
        const int64_t src = i + j * lda;
        out[i + j * ldo] = keep ? (cplx ? A[2 * src] : A[src]) : 0.0;
        if (cplx) out[plane + i + j * ldo] = keep ? A[2 * src + 1] : 0.0;
      }
  });
  return hipSuccess;
}
}  // namespace mpsa
// rocBLAS dgemm (host pointer mode), computed with plain loops
rocblas_status rocblas_dgemm(rocblas_handle, rocblas_operation ta, rocblas_operation tb, rocblas_int m,
                             rocblas_int n, rocblas_int k, const double* alpha, const double* A, rocblas_int lda,
                             const double* B, rocblas_int ldb, const double* beta, double* C, rocblas_int ldc) {
  const double al = *alpha, be = *beta;
  submit(nullptr, "rocblas_dgemm", [=]() {
    const bool tA = ta != rocblas_operation_none, tB = tb != rocblas_operation_none;
    const int ar = tA ? k : m, ac = tA ? m : k, br = tB ? n : k, bc = tB ? k : n;
    if (m < 0 || n < 0 || k < 0 || lda < std::max(1, ar) || ldb < std::max(1, br) || ldc < std::max(1, m)) {
      report("dgemm: inconsistent arguments");
      return;
    }
    if (m == 0 || n == 0) return;
    bool ok = chk_range(C, 0, (int64_t)ldc * (n - 1) + m - 1, "dgemm C");
    if (k > 0) {
      ok = chk_range(A, 0, (int64_t)lda * (ac - 1) + ar - 1, "dgemm A") && ok;
      ok = chk_range(B, 0, (int64_t)ldb * (bc - 1) + br - 1, "dgemm B") && ok;
    }
    if (!ok) return;
    std::vector<double> out((size_t)m * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        double acc = 0;
        for (int l = 0; l < k; ++l)
          acc += (tA ? A[l + (int64_t)i * lda] : A[i + (int64_t)l * lda]) *
                 (tB ? B[j + (int64_t)l * ldb] : B[l + (int64_t)j * ldb]);
        out[i + (size_t)j * m] = al * acc + (be == 0.0 ? 0.0 : be * C[i + (int64_t)j * ldc]);
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) C[i + (int64_t)j * ldc] = out[i + (size_t)j * m];
  });
  return rocblas_status_success;
}

// rocSOLVER dsyevd (LAPACK dsyev underneath): eigenvalues ascending in D,
// eigenvectors over A; E is the documented n-element workspace
rocblas_status rocsolver_dsyevd(rocblas_handle, const rocblas_evect evect, const rocblas_fill uplo,
                                const rocblas_int n, double* A, const rocblas_int lda, double* D, double* E,
                                rocblas_int* info) {
  submit(nullptr, "rocsolver_dsyevd", [=]() {
    if (evect != rocblas_evect_original || n < 0 || lda < std::max(1, n)) {
      report("syevd: unsupported or inconsistent arguments");
      return;
    }
    bool ok = check(info, sizeof(rocblas_int), "syevd info");
    if (n > 0) {
      ok = chk_range(A, 0, (int64_t)lda * (n - 1) + n - 1, "syevd A") && ok;
      ok = chk_range(D, 0, n - 1, "syevd D") && ok;
      ok = chk_range(E, 0, n - 1, "syevd E") && ok;
    }
    if (!ok || n == 0) return;
    typedef void (*dsyev_fn)(const char*, const char*, const int*, double*, const int*, double*, double*,
                             const int*, int*);
    void* h = lapack_lib();
    dsyev_fn dsyev_ = h ? (dsyev_fn)dlsym(h, "dsyev_") : nullptr;
    if (!dsyev_) {
      report("syevd: no LAPACK library found");
      return;
    }
    const char* ul = uplo == rocblas_fill_upper ? "U" : "L";
    int N = n, LDA = lda, lwork = -1, inf = 0;
    double wq = 0;
    dsyev_("V", ul, &N, A, &LDA, D, &wq, &lwork, &inf);
    lwork = (int)wq;
    std::vector<double> work((size_t)std::max(lwork, 1));
    dsyev_("V", ul, &N, A, &LDA, D, work.data(), &lwork, &inf);
    E[0] = 0.0;
    *info = inf;
  });
  return rocblas_status_success;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace mpsa {

// find_task of the kernels, with its descriptor reads checked
template <typename Task>
static int find_task(const Task* tasks, int ntasks, int t, int Task::*field) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (!check(tasks + mid, sizeof(Task), "task table")) return 0;
    if (tasks[mid].*field <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, typename TOut>
static void gemm16(int BM, int BN, bool TA, bool TB, const GTask* tasks, int ntasks, const GSeg* segs, int tpv,
                   int nvec, const T* A, int64_t a_vs, const T* B, int64_t b_vs, TOut* C, int64_t c_vs) {
  const int total = tpv * nvec;
  for (int g = 0; g < total; ++g) {
    const int v = g / tpv, t = g - v * tpv;
    if (ntasks < 1) {
      report("gemm: tiles but no tasks");
      return;
    }
    const int ti = find_task(tasks, ntasks, t, &GTask::tile0);
    if (!check(tasks + ti, sizeof(GTask), "gemm task")) return;
    const GTask tk = tasks[ti];
    if (tk.tiles_m <= 0) {
      report("gemm: task with tiles_m <= 0 reached");
      return;
    }
    const int local = t - tk.tile0, tm = local % tk.tiles_m, tn = local / tk.tiles_m;
    const int row0 = tm * BM, col0 = tn * BN, m = tk.m, n = tk.n;
    if (m <= 0 || n <= 0) {
      report("gemm: tile of an empty task (clamped address -1)");
      return;
    }
    const int r1 = std::min(row0 + BM, m), c1 = std::min(col0 + BN, n);
    std::vector<double> acc((size_t)BM * BN, 0.0);
    const T* Av = A + (int64_t)v * a_vs;
    const T* Bv = B + (int64_t)v * b_vs;
    for (int s = tk.seg0; s < tk.seg0 + tk.nseg; ++s) {
      if (!check(segs + s, sizeof(GSeg), "gemm segment")) return;
      const GSeg sg = segs[s];
      if (sg.k <= 0) {
        report("gemm: segment with k <= 0 (the kernel reads at clamped index -1)");
        return;
      }
      // physical reads: rows clamp(row0 .. row0+BM-1) in [row0, m-1] (row0 < m), k in [0, k-1]
      const int rlo = std::min(row0, m - 1), rhi = m - 1 < row0 + BM - 1 ? m - 1 : row0 + BM - 1;
      const int clo = std::min(col0, n - 1), chi = n - 1 < col0 + BN - 1 ? n - 1 : col0 + BN - 1;
      const T* Ab = Av + sg.a_off;
      const T* Bb = Bv + sg.b_off;
      bool ok = TA ? chk_range(Ab, (int64_t)rlo * sg.lda, (int64_t)rhi * sg.lda + sg.k - 1, "gemm A")
                   : chk_range(Ab, rlo, (int64_t)(sg.k - 1) * sg.lda + rhi, "gemm A");
      ok = (TB ? chk_range(Bb, clo, (int64_t)(sg.k - 1) * sg.ldb + chi, "gemm B")
               : chk_range(Bb, (int64_t)clo * sg.ldb, (int64_t)chi * sg.ldb + sg.k - 1, "gemm B")) &&
           ok;
      if (!ok) return;
      for (int j = col0; j < c1; ++j)
        for (int i = row0; i < r1; ++i) {
          double a = 0;
          for (int kk = 0; kk < sg.k; ++kk) {
            const double x = TA ? (double)Ab[kk + (int64_t)i * sg.lda] : (double)Ab[i + (int64_t)kk * sg.lda];
            const double y = TB ? (double)Bb[j + (int64_t)kk * sg.ldb] : (double)Bb[kk + (int64_t)j * sg.ldb];
            a += x * y;
          }
          acc[(size_t)(j - col0) * BM + (i - row0)] += sg.alpha * a;
        }
    }
    TOut* Cb = C + (int64_t)v * c_vs + tk.c_off;
    if (!chk_range(Cb, row0 + (int64_t)col0 * m, (r1 - 1) + (int64_t)(c1 - 1) * m, "gemm C")) return;
    for (int j = col0; j < c1; ++j)
      for (int i = row0; i < r1; ++i) Cb[i + (int64_t)j * m] = (TOut)acc[(size_t)(j - col0) * BM + (i - row0)];
  }
}

hipError_t launch_gemm(int prec_in, int out_kind, int tr, int tile_class, const GTask* tasks, int ntasks,
                       const GSeg* segs, int tiles_per_vec, int nvec, const void* A, int64_t a_vs, const void* B,
                       int64_t b_vs, void* C, int64_t c_vs, hipStream_t st, Planes) {
  if (tiles_per_vec * nvec <= 0) return hipSuccess;
  int bm = 64, bn = 64;
  tile_dims(prec_in, tile_class, &bm, &bn);
  const bool TA = (tr >> 1) & 1, TB = tr & 1;
  if (prec_in == F64) {
    submit(st, "gemm f64", [=]() {
      gemm16<double, double>(bm, bm, TA, TB, tasks, ntasks, segs, tiles_per_vec, nvec, (const double*)A, a_vs,
                             (const double*)B, b_vs, (double*)C, c_vs);
    });
    return hipSuccess;
  }
  if (prec_in == F32) {
    submit(st, "gemm f32", [=]() {
      if (out_kind == 1)
        gemm16<float, double>(bm, bm, TA, TB, tasks, ntasks, segs, tiles_per_vec, nvec, (const float*)A, a_vs,
                              (const float*)B, b_vs, (double*)C, c_vs);
      else
        gemm16<float, float>(bm, bm, TA, TB, tasks, ntasks, segs, tiles_per_vec, nvec, (const float*)A, a_vs,
                             (const float*)B, b_vs, (float*)C, c_vs);
    });
    return hipSuccess;
  }
  submit(st, "gemm bf16", []() { report("bf16 GEMM is not emulated"); });
  return hipSuccess;
}

hipError_t launch_mix(int prec, const MixTask* tasks, int ntasks, const MixTerm* terms, int chunks_per_vec, int nvec,
                      const void* src, void* dst, int64_t vs, hipStream_t st, int64_t) {
  const int total = chunks_per_vec * nvec;
  if (total <= 0 || ntasks <= 0) return hipSuccess;
  if (prec != F64 && prec != F32) {
    submit(st, "mix bf16", []() { report("bf16 mix is not emulated"); });
    return hipSuccess;
  }
  submit(st, "mix", [=]() {
    const size_t es = prec == F64 ? 8 : 4;
    for (int g = 0; g < total; ++g) {
      const int v = g / chunks_per_vec, c = g - v * chunks_per_vec;
      const int ti = find_task(tasks, ntasks, c, &MixTask::chunk0);
      if (!check(tasks + ti, sizeof(MixTask), "mix task")) return;
      const MixTask tk = tasks[ti];
      const int64_t base = (int64_t)(c - tk.chunk0) * 1024;
      const int64_t e1 = std::min<int64_t>(base + 1024, tk.len);
      if (e1 <= base) continue;
      if (!chk_range(terms, tk.t0, (int64_t)tk.t0 + tk.nt - 1, "mix terms")) return;
      const char* sv = (const char*)src + es * (size_t)(v * vs);
      char* dv = (char*)dst + es * (size_t)(v * vs + tk.dst_off);
      if (!check(dv + es * base, es * (size_t)(e1 - base), "mix dst")) return;
      for (int q = 0; q < tk.nt; ++q)
        if (!check(sv + es * (size_t)(terms[tk.t0 + q].src_off + base), es * (size_t)(e1 - base), "mix src")) return;
      for (int64_t e = base; e < e1; ++e) {
        double acc = 0;
        for (int q = 0; q < tk.nt; ++q) {
          const MixTerm tm = terms[tk.t0 + q];
          acc += tm.coef * (prec == F64 ? ((const double*)sv)[tm.src_off + e] : (double)((const float*)sv)[tm.src_off + e]);
        }
        if (prec == F64) ((double*)dv)[e] = acc;
        else ((float*)dv)[e] = (float)acc;
      }
    }
  });
  return hipSuccess;
}

hipError_t launch_convert(int prec, const double* src, void* dst, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "convert", [=]() {
    const size_t es = prec == F32 ? 4 : prec == BF16 ? 2 : 8;
    if (!check(src, 8 * (size_t)n, "convert src") || !check(dst, es * (size_t)n, "convert dst")) return;
    for (int64_t i = 0; i < n; ++i) {
      if (prec == F32) ((float*)dst)[i] = (float)src[i];
      else if (prec == F64) ((double*)dst)[i] = src[i];
    }
  });
  return hipSuccess;
}

hipError_t launch_repack(int prec, const Repack* tasks, int ntasks, int chunks_per_vec, int nvec, const double* src,
                         int64_t s_vs, void* dst, int64_t d_vs, hipStream_t st, int64_t plane) {
  const int total = chunks_per_vec * nvec;
  if (total <= 0 || ntasks <= 0) return hipSuccess;
  submit(st, "repack", [=]() {
    const size_t es = prec == F32 ? 4 : (prec == BF16 || prec == BF16X3) ? 2 : 8;
    for (int g = 0; g < total; ++g) {
      const int v = g / chunks_per_vec, c = g - v * chunks_per_vec;
      const int ti = find_task(tasks, ntasks, c, &Repack::chunk0);
      if (!check(tasks + ti, sizeof(Repack), "repack task")) return;
      const Repack tk = tasks[ti];
      const int64_t len = (int64_t)tk.rows * tk.cols, base = (int64_t)(c - tk.chunk0) * 1024;
      const int64_t e1 = std::min<int64_t>(base + 1024, len);
      if (e1 <= base) continue;
      const double* s = src + v * s_vs + tk.src_off;
      char* d = (char*)dst + es * (size_t)(v * d_vs + tk.dst_off);
      if (!chk_range(s, base, e1 - 1, "repack src")) return;
      for (int64_t e = base; e < e1; ++e) {
        const int64_t j = e / tk.rows, i = e - j * tk.rows;
        const int64_t di = tk.transpose ? (j + i * tk.dst_ld) : (i + j * tk.dst_ld);
        if (!check(d + es * (size_t)di, es, "repack dst")) return;
        if (prec == BF16X3 && !check(d + es * (size_t)(di + plane), es, "repack dst lo plane")) return;
        const double x = tk.scale * s[e];
        if (prec == F32) *(float*)(d + 4 * di) = (float)x;
        else if (prec == F64) *(double*)(d + 8 * di) = x;
      }
    }
  });
  return hipSuccess;
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "fill", [=]() {
    if (chk_range(p, 0, n - 1, "fill"))
      for (int64_t i = 0; i < n; ++i) p[i] = v;
  });
  return hipSuccess;
}
hipError_t launch_fill_hash(double* p, int64_t n, uint64_t seed, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "fill_hash", [=]() {
    if (!chk_range(p, 0, n - 1, "fill_hash")) return;
    for (int64_t i = 0; i < n; ++i) {
      uint64_t z = seed + ((uint64_t)i + 1) * 0x9e3779b97f4a7c15ull;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      z ^= z >> 31;
      p[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
  });
  return hipSuccess;
}
// seeded Gaussian fill: the oracle's host restatement (oracle/oracle.cpp,
// compiled into this emulation library)
extern "C" void orc_fill_gauss(double* p, int64_t n, uint64_t seed);
hipError_t launch_fill_gauss(double* p, int64_t n, uint64_t seed, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "fill_gauss", [=]() {
    if (!chk_range(p, 0, n - 1, "fill_gauss")) return;
    orc_fill_gauss(p, n, seed);
  });
  return hipSuccess;
}
hipError_t launch_fill_strided(double* p, int64_t n, int64_t stride, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "fill_strided", [=]() {
    for (int64_t i = 0; i < n; ++i)
      if (check(p + i * stride, 8, "fill_strided")) p[i * stride] = v;
  });
  return hipSuccess;
}
hipError_t launch_zero_ranges(const Range* r, int nr, double* base, int64_t vs, int nvec, hipStream_t st) {
  if (nr <= 0 || nvec <= 0) return hipSuccess;
  submit(st, "zero_ranges", [=]() {
    if (!chk_range(r, 0, nr - 1, "zero_ranges table")) return;
    for (int z = 0; z < nvec; ++z)
      for (int y = 0; y < nr; ++y) {
        double* b = base + z * vs + r[y].off;
        if (chk_range(b, 0, r[y].len - 1, "zero_ranges"))
          for (int64_t i = 0; i < r[y].len; ++i) b[i] = 0.0;
      }
  });
  return hipSuccess;
}
hipError_t launch_dot(const double* x, const double* y, int64_t n, int cplx, double* partial, int nparts, double* out2,
                      hipStream_t st) {
  submit(st, "dot", [=]() {
    const int64_t len = cplx ? 2 * n : n;
    if (!chk_range(x, 0, len - 1, "dot x") || !chk_range(y, 0, len - 1, "dot y") ||
        !chk_range(partial, 0, 2 * (int64_t)nparts - 1, "dot partial") || !chk_range(out2, 0, 1, "dot out"))
      return;
    double ar = 0, ai = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (!cplx) ar += y[i] * x[i];
      else {
        ar += y[2 * i] * x[2 * i] + y[2 * i + 1] * x[2 * i + 1];
        ai += y[2 * i] * x[2 * i + 1] - y[2 * i + 1] * x[2 * i];
      }
    }
    for (int p = 0; p < 2 * nparts; ++p) partial[p] = 0;
    partial[0] = ar;
    partial[1] = ai;
    out2[0] = ar;
    out2[1] = ai;
  });
  return hipSuccess;
}
hipError_t launch_axpby(double ar, double ai, const double* x, double br, double bi, double* y, int64_t n, int cplx,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  submit(st, "axpby", [=]() {
    const int64_t len = cplx ? 2 * n : n;
    if (!chk_range(x, 0, len - 1, "axpby x") || !chk_range(y, 0, len - 1, "axpby y")) return;
    for (int64_t i = 0; i < n; ++i) {
      if (!cplx) y[i] = ar * x[i] + br * y[i];
      else {
        const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
        y[2 * i] = ar * xr - ai * xi + br * yr - bi * yi;
        y[2 * i + 1] = ar * xi + ai * xr + br * yi + bi * yr;
      }
    }
  });
  return hipSuccess;
}
hipError_t launch_bdot(const double* x, const double* y, int64_t n, int64_t vs, int nvec, double* partial, int nparts,
                       double* out, int ld, const int* step, int root, hipStream_t st) {
  submit(st, "bdot", [=]() {
    if (!check(step, 4, "bdot step") || !chk_range(partial, 0, (int64_t)nvec * nparts - 1, "bdot partial")) return;
    const int j = *step;
    for (int v = 0; v < nvec; ++v) {
      if (!chk_range(x + v * vs, 0, n - 1, "bdot x") || !chk_range(y + v * vs, 0, n - 1, "bdot y")) return;
      double a = 0;
      for (int64_t i = 0; i < n; ++i) a += x[v * vs + i] * y[v * vs + i];
      if (!check(out + (int64_t)v * ld + j, 8, "bdot out")) return;
      out[(int64_t)v * ld + j] = root ? std::sqrt(a) : a;
    }
  });
  return hipSuccess;
}
hipError_t launch_lz_normalize(const double* W, double* X, double* const* basis, const double* kbuf, int ld,
                               const int* step, int64_t n, int64_t vs, int nvec, hipStream_t st) {
  submit(st, "lz_normalize", [=]() {
    if (!check(step, 4, "lz step")) return;
    const int j = *step;
    if (!check(basis + j, 8, "lz basis table")) return;
    for (int v = 0; v < nvec; ++v) {
      if (!check(kbuf + (int64_t)v * ld + j, 8, "lz kbuf")) return;
      const double r = 1.0 / kbuf[(int64_t)v * ld + j];
      const int64_t o = (int64_t)v * vs;
      double* Bv = basis[j] + o;
      if (!chk_range(W + o, 0, n - 1, "lz W") || !chk_range(X + o, 0, n - 1, "lz X") ||
          !chk_range(Bv, 0, n - 1, "lz basis"))
        return;
      for (int64_t i = 0; i < n; ++i) X[o + i] = Bv[i] = r * W[o + i];
    }
  });
  return hipSuccess;
}
int lz_update_parts(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return (int)(blocks < 1024 ? (blocks > 0 ? blocks : 1) : 1024);
}
hipError_t launch_bdot_final(const double* partial, int nparts, int nvec, double* out, int ld, const int* step,
                             int root, hipStream_t st) {
  submit(st, "bdot final", [=]() {
    if (!check(step, 4, "bdot step") || !chk_range(partial, 0, (int64_t)nvec * nparts - 1, "bdot partial")) return;
    const int j = *step;
    for (int v = 0; v < nvec; ++v) {
      double a = 0;
      for (int q = 0; q < nparts; ++q) a += partial[(int64_t)v * nparts + q];
      if (!check(out + (int64_t)v * ld + j, 8, "bdot out")) return;
      out[(int64_t)v * ld + j] = root ? std::sqrt(a) : a;
    }
  });
  return hipSuccess;
}
hipError_t launch_lz_update(const double* Y, const double* X, double* const* basis, const double* kbuf,
                            const double* ebuf, int ld, const int* step, double* W, int64_t n, int64_t vs, int nvec,
                            double* wpart, hipStream_t st) {
  const int parts = lz_update_parts(n);
  submit(st, "lz_update", [=]() {
    if (wpart && !chk_range(wpart, 0, (int64_t)nvec * parts - 1, "lz wpart")) return;
    if (!check(step, 4, "lz step")) return;
    const int j = *step;
    if (j > 0 && !check(basis + j - 1, 8, "lz basis table")) return;
    for (int v = 0; v < nvec; ++v) {
      const int64_t o = (int64_t)v * vs;
      if (!check(ebuf + (int64_t)v * ld + j, 8, "lz ebuf") || !check(kbuf + (int64_t)v * ld + j, 8, "lz kbuf")) return;
      const double e = ebuf[(int64_t)v * ld + j], k = kbuf[(int64_t)v * ld + j];
      const double* A0 = j > 0 ? basis[j - 1] + o : nullptr;
      if (!chk_range(Y + o, 0, n - 1, "lz Y") || !chk_range(X + o, 0, n - 1, "lz X") ||
          !chk_range(W + o, 0, n - 1, "lz W") || (A0 && !chk_range(A0, 0, n - 1, "lz A0")))
        return;
      double a = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        double t = -e * X[o + i] + Y[o + i];
        if (A0) t = t + -k * A0[i];
        W[o + i] = t;
        a += t * t;
      }
      if (wpart)  // the sum in the first part, zeros in the others
        for (int q = 0; q < parts; ++q) wpart[(int64_t)v * parts + q] = q == 0 ? a : 0.0;
    }
  });
  return hipSuccess;
}
hipError_t launch_project_out(double* W, double* const* Q, int64_t qvs, const int* cnt, int add, int nbmax, int64_t n,
                              int64_t vs, int nvec, double* partial, int nparts, double* c, hipStream_t st) {
  if (nbmax <= 0 || nvec <= 0) return hipSuccess;
  submit(st, "project_out", [=]() {
    if (cnt && !check(cnt, 4, "project count")) return;
    const int nb = std::min((cnt ? *cnt : 0) + add, nbmax);
    if (nb > 0 && !chk_range(Q, 0, nb - 1, "project table")) return;
    if (!chk_range(partial, 0, (int64_t)nvec * nbmax * nparts - 1, "project partial") ||
        !chk_range(c, 0, (int64_t)nvec * nbmax - 1, "project coefficients"))
      return;
    for (int v = 0; v < nvec; ++v) {
      double* wv = W + (int64_t)v * vs;
      if (!chk_range(wv, 0, n - 1, "project W")) return;
      for (int i = 0; i < nb; ++i) {
        const double* q = Q[i] + (int64_t)v * qvs;
        if (!chk_range(q, 0, n - 1, "project Q")) return;
        double a = 0;
        for (int64_t k = 0; k < n; ++k) a += q[k] * wv[k];
        c[(int64_t)v * nbmax + i] = a;
      }
      for (int64_t k = 0; k < n; ++k) {
        double t = wv[k];
        for (int i = 0; i < nb; ++i) t -= c[(int64_t)v * nbmax + i] * Q[i][(int64_t)v * qvs + k];
        wv[k] = t;
      }
    }
  });
  return hipSuccess;
}
hipError_t launch_scale_rsqrt(double* X, const double* nrm2, int64_t n, int64_t vs, int nvec, hipStream_t st) {
  submit(st, "scale_rsqrt", [=]() {
    if (!chk_range(nrm2, 0, nvec - 1, "rsqrt norms")) return;
    for (int v = 0; v < nvec; ++v) {
      if (!chk_range(X + (int64_t)v * vs, 0, n - 1, "rsqrt X")) return;
      const double r = 1.0 / std::sqrt(nrm2[v]);
      for (int64_t k = 0; k < n; ++k) X[(int64_t)v * vs + k] *= r;
    }
  });
  return hipSuccess;
}
hipError_t launch_lz_step(int* step, hipStream_t st) {
  submit(st, "lz_step", [=]() {
    if (check(step, 4, "lz step")) *step += 1;
  });
  return hipSuccess;
}
hipError_t launch_lz_combine(double* const* basis, const double* c, int ldc, const int* m, double* out, int64_t n,
                             int64_t vs, int nvec, hipStream_t st) {
  submit(st, "lz_combine", [=]() {
    if (!chk_range(m, 0, nvec - 1, "lz m")) return;
    for (int v = 0; v < nvec; ++v) {
      const int mv = m[v];
      const int64_t o = (int64_t)v * vs;
      if (!chk_range(out + o, 0, n - 1, "lz out") || !chk_range(c + (int64_t)v * ldc, 0, mv - 1, "lz coef") ||
          !chk_range(basis, 0, mv - 1, "lz basis table"))
        return;
      for (int b = 0; b < mv; ++b)
        if (!chk_range(basis[b] + o, 0, n - 1, "lz basis")) return;
      for (int64_t i = 0; i < n; ++i) {
        double t = 0;
        for (int b = 0; b < mv; ++b) t += c[(int64_t)v * ldc + b] * basis[b][o + i];
        out[o + i] = t;
      }
    }
  });
  return hipSuccess;
}
hipError_t launch_copy2d(const Copy2D* tasks, int ntasks, int64_t max_elems, const double* src, double* dst,
                         const double* scales, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  submit(st, "copy2d", [=]() {
    if (!chk_range(tasks, 0, ntasks - 1, "copy2d tasks")) return;
    for (int y = 0; y < ntasks; ++y) {
      const Copy2D t = tasks[y];
      const int64_t total = (int64_t)t.rows * t.cols;
      if (total > max_elems) report("copy2d: task larger than max_elems (part of it is never copied)");
      for (int64_t e = 0; e < total; ++e) {
        const int64_t i = e % t.rows, j = e / t.rows;
        const int64_t si = t.src_off + (t.trans ? i * t.lds + j : j * t.lds + i);
        if (!check(src + si, 8, "copy2d src")) return;
        double v = src[si] * t.scale;
        if (t.rs_off >= 0) {
          if (!check(scales + t.rs_off + i, 8, "copy2d row scale")) return;
          v *= scales[t.rs_off + i];
        }
        if (t.cs_off >= 0) {
          if (!check(scales + t.cs_off + j, 8, "copy2d col scale")) return;
          v *= scales[t.cs_off + j];
        }
        if (!check(dst + t.dst_off + j * t.ldd + i, 8, "copy2d dst")) return;
        dst[t.dst_off + j * t.ldd + i] = v;
      }
    }
  });
  return hipSuccess;
}
hipError_t launch_fixphase(double* U, int64_t m, int64_t ldu, double* Vt, int64_t k, int64_t n, int64_t ldv,
                           hipStream_t st) {
  if (k <= 0) return hipSuccess;
  submit(st, "fixphase", [=]() {
    for (int64_t i = 0; i < k; ++i) {
      double best = 0;
      int64_t bi = 0;
      for (int64_t j = 0; j < n; ++j) {
        if (!check(Vt + i + j * ldv, 8, "fixphase Vt")) return;
        if (std::fabs(Vt[i + j * ldv]) > best) {
          best = std::fabs(Vt[i + j * ldv]);
          bi = j;
        }
      }
      if (!check(Vt + i + bi * ldv, 8, "fixphase Vt")) return;
      const double phi = Vt[i + bi * ldv] > 0.0 ? 1.0 : -1.0;
      for (int64_t j = 0; j < n; ++j) Vt[i + j * ldv] *= phi;
      for (int64_t r = 0; r < m; ++r) {
        if (!check(U + r + i * ldu, 8, "fixphase U")) return;
        U[r + i * ldu] *= phi;
      }
    }
  });
  return hipSuccess;
}
hipError_t launch_colnorms(const double* A, int64_t rows, int64_t cols, int64_t lda, double* out, hipStream_t st) {
  if (cols <= 0) return hipSuccess;
  submit(st, "colnorms", [=]() {
    if (!chk_range(A, 0, lda * (cols - 1) + rows - 1, "colnorms A") || !chk_range(out, 0, cols - 1, "colnorms out"))
      return;
    for (int64_t j = 0; j < cols; ++j) {
      double a = 0.0;
      for (int64_t i = 0; i < rows; ++i) a += A[i + j * lda] * A[i + j * lda];
      out[j] = std::sqrt(a);
    }
  });
  return hipSuccess;
}
// batched Householder QR (mpsa_factor.hip's contract): X = M (right = 0) or
// M^H (right = 1), p x r, X = Q R; outputs U = Q, V^dagger = R (left) or
// U = R^H, V^dagger = Q^H (right), S = 1
hipError_t launch_qr_batched(int cplx, int right, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                             double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                             int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  submit(st, "qr batched", [=]() {
    if (!chk_range(tasks, 0, ntasks - 1, "qr tasks")) return;
    typedef std::complex<double> Z;
    const int pl = cplx ? 2 : 1;
    for (int ti = 0; ti < ntasks; ++ti) {
      const SvdTask t = tasks[ti];
      const int m = t.m, n = t.n, k = std::min(m, n);
      if (m <= 0 || n <= 0) continue;
      const int p = right ? n : m, r = right ? m : n;
      if (t.lda < m) {
        report("qr: task outside the kernel's limits");
        return;
      }
      bool ok = true;
      for (int q = 0; q < pl; ++q) {
        ok = chk_range(M, t.a_off + q * m_plane, t.a_off + q * m_plane + (int64_t)t.lda * (n - 1) + m - 1, "qr M") && ok;
        ok = chk_range(U, t.u_off + q * u_plane, t.u_off + q * u_plane + (int64_t)m * k - 1, "qr U") && ok;
        ok = chk_range(Vt, t.vt_off + q * vt_plane, t.vt_off + q * vt_plane + (int64_t)k * n - 1, "qr Vt") && ok;
        ok = chk_range(W, t.w_off + q * w_plane,
                       t.w_off + q * w_plane + (int64_t)p * r + (int64_t)p * k + k - 1, "qr W") && ok;
      }
      ok = chk_range(S, t.s_off, t.s_off + k - 1, "qr S") && ok;
      if (!ok) return;
      auto mat = [&](int i, int j) {  // M[i, j]
        const int64_t o = t.a_off + i + (int64_t)j * t.lda;
        return Z(M[o], cplx ? M[o + m_plane] : 0.0);
      };
      std::vector<Z> X((size_t)p * r), Q((size_t)p * k, Z(0.0));
      for (int j = 0; j < r; ++j)
        for (int i = 0; i < p; ++i) X[(size_t)i + (size_t)j * p] = right ? std::conj(mat(j, i)) : mat(i, j);
      std::vector<std::vector<Z>> vs;
      std::vector<double> betas;
      for (int j = 0; j < k; ++j) {
        double nn = 0.0;
        for (int i = j; i < p; ++i) nn += std::norm(X[(size_t)i + (size_t)j * p]);
        const Z x0 = X[(size_t)j + (size_t)j * p];
        const double a0 = std::abs(x0);
        Z alpha(0.0);
        std::vector<Z> v((size_t)p, Z(0.0));
        double b = 0.0;
        if (nn > 0.0) {
          alpha = -(a0 > 0.0 ? x0 / a0 : Z(1.0)) * std::sqrt(nn);
          for (int i = j; i < p; ++i) v[(size_t)i] = X[(size_t)i + (size_t)j * p];
          v[(size_t)j] -= alpha;
          double vn = 0.0;
          for (int i = j; i < p; ++i) vn += std::norm(v[(size_t)i]);
          b = vn > 0.0 ? 2.0 / vn : 0.0;
        }
        for (int c = j + 1; c < r; ++c) {
          Z w(0.0);
          for (int i = j; i < p; ++i) w += std::conj(v[(size_t)i]) * X[(size_t)i + (size_t)c * p];
          for (int i = j; i < p; ++i) X[(size_t)i + (size_t)c * p] -= b * w * v[(size_t)i];
        }
        for (int i = j; i < p; ++i) X[(size_t)i + (size_t)j * p] = i == j ? alpha : Z(0.0);
        vs.push_back(v);
        betas.push_back(b);
      }
      for (int c = 0; c < k; ++c) Q[(size_t)c + (size_t)c * p] = 1.0;
      for (int j = k - 1; j >= 0; --j)
        for (int c = j; c < k; ++c) {
          Z w(0.0);
          for (int i = j; i < p; ++i) w += std::conj(vs[(size_t)j][(size_t)i]) * Q[(size_t)i + (size_t)c * p];
          for (int i = j; i < p; ++i) Q[(size_t)i + (size_t)c * p] -= betas[(size_t)j] * w * vs[(size_t)j][(size_t)i];
        }
      auto put = [&](double* base, int64_t plane, int64_t off, Z z) {
        base[off] = z.real();
        if (cplx) base[off + plane] = z.imag();
      };
      for (int c = 0; c < r; ++c)
        for (int i = 0; i < k; ++i) {
          const Z rv = i <= c ? X[(size_t)i + (size_t)c * p] : Z(0.0);
          if (!right) put(Vt, vt_plane, t.vt_off + i + (int64_t)c * k, rv);
          else put(U, u_plane, t.u_off + c + (int64_t)i * m, std::conj(rv));
        }
      for (int c = 0; c < k; ++c)
        for (int i = 0; i < p; ++i) {
          const Z qv = Q[(size_t)i + (size_t)c * p];
          if (!right) put(U, u_plane, t.u_off + i + (int64_t)c * m, qv);
          else put(Vt, vt_plane, t.vt_off + c + (int64_t)i * k, std::conj(qv));
        }
      for (int i = 0; i < k; ++i) S[t.s_off + i] = 1.0;
    }
  });
  return hipSuccess;
}
// batched thin SVD (mpsa_factor.hip's contract): LAPACK dgesvd / zgesvd per
// task, singular values descending, FixPhase on real data
hipError_t launch_svd_batched(int cplx, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane, double* U,
                              int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W, int64_t w_plane,
                              hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  submit(st, "svd batched", [=]() {
    if (!chk_range(tasks, 0, ntasks - 1, "svd tasks")) return;
    void* h = lapack_lib();
    typedef void (*zgesvd_fn)(const char*, const char*, const int*, const int*, void*, const int*, double*, void*,
                              const int*, void*, const int*, void*, const int*, double*, int*);
    dgesvd_fn dg = lapack_dgesvd();
    zgesvd_fn zg = h ? (zgesvd_fn)dlsym(h, "zgesvd_") : nullptr;
    if (!dg || (cplx && !zg)) {
      report("svd: no LAPACK library found");
      return;
    }
    const int pl = cplx ? 2 : 1;
    for (int ti = 0; ti < ntasks; ++ti) {
      const SvdTask t = tasks[ti];
      const int m = t.m, n = t.n, k = std::min(m, n), p = std::max(m, n);
      if (m <= 0 || n <= 0) continue;
      if (k > kSvdMaxR || t.lda < m) {
        report("svd: task outside the kernel's limits");
        return;
      }
      bool ok = true;
      for (int q = 0; q < pl; ++q) {
        ok = chk_range(M, t.a_off + q * m_plane, t.a_off + q * m_plane + (int64_t)t.lda * (n - 1) + m - 1, "svd M") && ok;
        ok = chk_range(U, t.u_off + q * u_plane, t.u_off + q * u_plane + (int64_t)m * k - 1, "svd U") && ok;
        ok = chk_range(Vt, t.vt_off + q * vt_plane, t.vt_off + q * vt_plane + (int64_t)k * n - 1, "svd Vt") && ok;
        const int64_t wl = svd_workspace(m, n);
        ok = chk_range(W, t.w_off + q * w_plane, t.w_off + q * w_plane + wl - 1, "svd W") && ok;
      }
      ok = chk_range(S, t.s_off, t.s_off + k - 1, "svd S") && ok;
      if (!ok) return;
      int mm = m, nn = n, lda = m, ldu = m, ldvt = k, lwork = -1, info = 0;
      if (!cplx) {
        std::vector<double> a((size_t)m * n), u((size_t)m * k), vt((size_t)k * n), s((size_t)k);
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < m; ++i) a[(size_t)i + (size_t)j * m] = M[t.a_off + i + (int64_t)j * t.lda];
        double wq = 0;
        dg("S", "S", &mm, &nn, a.data(), &lda, s.data(), u.data(), &ldu, vt.data(), &ldvt, &wq, &lwork, &info);
        lwork = (int)wq;
        std::vector<double> work((size_t)std::max(lwork, 1));
        dg("S", "S", &mm, &nn, a.data(), &lda, s.data(), u.data(), &ldu, vt.data(), &ldvt, work.data(), &lwork,
           &info);
        for (int i = 0; i < k; ++i) {  // FixPhase
          double best = 0.0;
          int bi = 0;
          for (int j = 0; j < n; ++j)
            if (std::fabs(vt[(size_t)i + (size_t)j * k]) > best) {
              best = std::fabs(vt[(size_t)i + (size_t)j * k]);
              bi = j;
            }
          if (!(vt[(size_t)i + (size_t)bi * k] > 0.0)) {
            for (int j = 0; j < n; ++j) vt[(size_t)i + (size_t)j * k] = -vt[(size_t)i + (size_t)j * k];
            for (int r = 0; r < m; ++r) u[(size_t)r + (size_t)i * m] = -u[(size_t)r + (size_t)i * m];
          }
        }
        for (size_t e = 0; e < u.size(); ++e) U[t.u_off + (int64_t)e] = u[e];
        for (size_t e = 0; e < vt.size(); ++e) Vt[t.vt_off + (int64_t)e] = vt[e];
        for (int i = 0; i < k; ++i) S[t.s_off + i] = s[(size_t)i];
      } else {
        std::vector<double> a((size_t)2 * m * n), u((size_t)2 * m * k), vt((size_t)2 * k * n), s((size_t)k),
            rwork((size_t)5 * k + 8);
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < m; ++i) {
            a[2 * ((size_t)i + (size_t)j * m)] = M[t.a_off + i + (int64_t)j * t.lda];
            a[2 * ((size_t)i + (size_t)j * m) + 1] = M[t.a_off + m_plane + i + (int64_t)j * t.lda];
          }
        double wq[2] = {0, 0};
        zg("S", "S", &mm, &nn, a.data(), &lda, s.data(), u.data(), &ldu, vt.data(), &ldvt, wq, &lwork, rwork.data(),
           &info);
        lwork = (int)wq[0];
        std::vector<double> work((size_t)2 * std::max(lwork, 1));
        zg("S", "S", &mm, &nn, a.data(), &lda, s.data(), u.data(), &ldu, vt.data(), &ldvt, work.data(), &lwork,
           rwork.data(), &info);
        for (size_t e = 0; e < (size_t)m * k; ++e) {
          U[t.u_off + (int64_t)e] = u[2 * e];
          U[t.u_off + u_plane + (int64_t)e] = u[2 * e + 1];
        }
        for (size_t e = 0; e < (size_t)k * n; ++e) {
          Vt[t.vt_off + (int64_t)e] = vt[2 * e];
          Vt[t.vt_off + vt_plane + (int64_t)e] = vt[2 * e + 1];
        }
        for (int i = 0; i < k; ++i) S[t.s_off + i] = s[(size_t)i];
      }
      if (info != 0) report("svd: LAPACK did not converge");
    }
  });
  return hipSuccess;
}
// the multi-workgroup factorisation's phases: the phases before the outputs
// leave nothing the emulation needs (they check their tasks), no round
// rotates, and kPhOut is the whole SVD or QR (LAPACK) from M
static void chk_phase_tasks(hipStream_t st, const SvdTask* tasks, int ntasks, int rmax) {
  submit(st, "factor phase", [=]() {
    if (!chk_range(tasks, 0, ntasks - 1, "factor tasks")) return;
    for (int i = 0; i < ntasks; ++i) {
      if (std::min(tasks[i].m, tasks[i].n) > rmax) report("factor phase: task above the blocked path's limit");
      if (std::max(tasks[i].m, tasks[i].n) > kVecRowsMax) report("factor phase: task above the register path's rows");
    }
  });
}
hipError_t launch_svd_phase(int cplx, int phases, const SvdTask* tasks, int ntasks, const double* M, int64_t m_plane,
                            double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S, double* W,
                            int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (!(phases & kPhOut)) {
    chk_phase_tasks(st, tasks, ntasks, kSvdBlockMaxR);
    return hipSuccess;
  }
  return launch_svd_batched(cplx, tasks, ntasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane, st);
}
hipError_t launch_qr_phase(int cplx, int right, int phases, const SvdTask* tasks, int ntasks, const double* M,
                           int64_t m_plane, double* U, int64_t u_plane, double* Vt, int64_t vt_plane, double* S,
                           double* W, int64_t w_plane, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (!(phases & kPhOut)) {
    chk_phase_tasks(st, tasks, ntasks, kVecRowsMax);
    return hipSuccess;
  }
  return launch_qr_batched(cplx, right, tasks, ntasks, M, m_plane, U, u_plane, Vt, vt_plane, S, W, w_plane, st);
}
static void chk_house_table(hipStream_t st, const SvdTask* tasks, int ntasks, const int* col_base, int nchunks,
                            int layout) {
  submit(st, "house chunks", [=]() {
    if (!chk_range(tasks, 0, ntasks - 1, "house tasks") || !chk_range(col_base, 0, ntasks, "house chunk base"))
      return;
    if (layout < 0 || layout > 2) report("house: bad layout");
    if (col_base[0] != 0 || col_base[ntasks] != nchunks) report("house: bad chunk table");
    for (int i = 0; i < ntasks; ++i)
      if (col_base[i + 1] - col_base[i] != house_chunks(tasks[i].m, tasks[i].n, layout))
        report("house: chunk count differs from house_chunks");
  });
}
hipError_t launch_house_panels(int, int layout, const SvdTask* tasks, int ntasks, const int* col_base, int nchunks,
                               int kmax, double*, int64_t, hipStream_t st) {
  if (ntasks <= 0 || nchunks <= 0) return hipSuccess;
  if (kmax < 1) report("house panels: kmax < 1");
  chk_house_table(st, tasks, ntasks, col_base, nchunks, layout);
  return hipSuccess;
}
hipError_t launch_house_apply(int, int layout, const SvdTask* tasks, int ntasks, const int* col_base, int nchunks,
                              double*, int64_t, hipStream_t st) {
  if (ntasks <= 0 || nchunks <= 0) return hipSuccess;
  chk_house_table(st, tasks, ntasks, col_base, nchunks, layout);
  return hipSuccess;
}
int house_chunks(int m, int n, int layout) {
  const int r = layout == 0 ? std::min(m, n) : (layout == 1 ? n : m);
  return (r + 7) / 8;
}
hipError_t launch_svd_block_round(int, const SvdTask* tasks, int ntasks, const int* pair_base, int npairs, int round,
                                  double*, int64_t, double* rotated, int rmax, hipStream_t st) {
  if (rmax < 1 || rmax > kSvdBlockMaxR) return hipErrorInvalidValue;
  submit(st, "svd block round", [=]() {
    for (int i = 0; i < ntasks; ++i)
      if (std::min(tasks[i].m, tasks[i].n) > rmax) report("svd block round: task above rmax");
    if (!chk_range(tasks, 0, ntasks - 1, "svd tasks") || !chk_range(pair_base, 0, ntasks, "svd pair base") ||
        !chk_range(rotated, 0, ntasks - 1, "svd rotated"))
      return;
    if (pair_base[0] != 0 || pair_base[ntasks] != npairs || round < -1) report("svd block round: bad pair table");
  });
  return hipSuccess;
}
int svd_block_pairs(int m, int n) {
  const int r = std::min(m, n), nb = (r + 7) / 8;
  return (nb + (nb & 1)) / 2;
}
int svd_block_rounds(int m, int n) {
  const int r = std::min(m, n), nb = (r + 7) / 8;
  return nb + (nb & 1) - 1;
}
int gemm_max_grid(int, int) { return 512; }

}  // namespace mpsa
