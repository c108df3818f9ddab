// mpsa_kernels.hip -- gfx950 kernels of the DMRG hot path.
//
// * gemm16   : persistent grouped, segmented-K GEMM on the 16x16x4 MFMA
//              (v_mfma_f64_16x16x4_f64 for FP64, v_mfma_f32_16x16x4_f32 for
//              FP32).  One launch covers every sector GEMM of a stage for all
//              batched vectors; tiles are distributed over a persistent grid.
// * gemm_bf16: the same on v_mfma_f32_16x16x32_bf16 (bf16 in, fp32 accumulate),
//              for operands stored k-contiguous with leading dims padded to 8.
// * mix      : the W-mixing between the two GEMM stages (T2 = sum c T1).
// * convert / zero / dot / axpby: vector plumbing for the Lanczos recurrence.
//
// The contraction these kernels carry out is the reference's
// doContraction/add_out_sub tuple loop (src/MPSfunctions.hpp:276-352,
// :726-804) regrouped into unique (L, psi) products, a W mix, and one
// segmented-K product per output block.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <utility>

#include "mpsa_internal.h"

namespace mpsa {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <typename T>
struct Acc4;
template <>
struct Acc4<double> {
  typedef d4 type;
};
template <>
struct Acc4<float> {
  typedef f4 type;
};

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row (within a 16x16 output tile) of accumulator register r in lanes 16*lq..16*lq+15;
// the column is lane % 16.  v_mfma_f32_16x16x4_f32 and the 16x16x32 bf16 MFMA
// give each lane four consecutive rows (4*lq + r); v_mfma_f64_16x16x4_f64
// interleaves them (4*r + lq).
template <typename T>
__device__ __forceinline__ int acc_row(int lq, int r);
template <>
__device__ __forceinline__ int acc_row<double>(int lq, int r) {
  return 4 * r + lq;
}
template <>
__device__ __forceinline__ int acc_row<float>(int lq, int r) {
  return 4 * lq + r;
}

// Find the task owning tile t (tasks sorted by tile0; tile0[0] == 0).
template <typename Task>
__device__ __forceinline__ int find_task(const Task* tasks, int ntasks, int t, int Task::*field) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].*field <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Position of a k-step iterator over the segments of one task, with the
// segment's descriptor held in registers: memory is read only when the
// iterator enters a new segment, never per k-step.
struct SegCursor {
  int si, kk, k, lda, ldb;
  int64_t a_off, b_off;
  double alpha;
};
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
// Plans never emit segments with k <= 0, so entering a segment is one uniform
// (scalar) descriptor read.
__device__ __forceinline__ void cur_enter(SegCursor& c, const GSeg* __restrict__ segs, int si, int send) {
  si = uni(si);
  c.si = si;
  c.kk = 0;
  if (si < send) {
    const GSeg g = *(segs + si);
    c.k = uni(g.k);
    c.lda = uni(g.lda);
    c.ldb = uni(g.ldb);
    c.a_off = g.a_off;
    c.b_off = g.b_off;
    c.alpha = g.alpha;
  }
}
__device__ __forceinline__ void cur_next(SegCursor& c, const GSeg* __restrict__ segs, int send, int bk) {
  c.kk += bk;
  if (c.kk >= c.k) cur_enter(c, segs, uni(c.si) + 1, send);
}

// Persistent-grid mapping that keeps consecutive logical work items on one
// XCD (workgroups are dispatched round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_logical_id() {
  const int G = gridDim.x, b = blockIdx.x;
  if ((G & 7) != 0) return b;
  return (b & 7) * (G >> 3) + (b >> 3);
}

template <typename TO, typename TA>
__device__ __forceinline__ TO cvt(TA x) {
  return (TO)x;
}
template <>
__device__ __forceinline__ __hip_bfloat16 cvt<__hip_bfloat16, float>(float x) {
  return __float2bfloat16(x);
}
template <>
__device__ __forceinline__ __hip_bfloat16 cvt<__hip_bfloat16, double>(double x) {
  return __float2bfloat16((float)x);
}

// ---------------------------------------------------------------------------
// FP64 / FP32 grouped GEMM, 16x16x4 MFMA.  256 threads = 4 waves in 2x2; each
// wave owns a (BM/2)x(BN/2) sub-tile.  A and B tiles go through LDS (k-major);
// the next k-step's global loads are issued into registers before the MFMAs of
// the current step.
// ---------------------------------------------------------------------------
template <typename T, typename TOut, int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm16_kernel(const GTask* __restrict__ tasks, int ntasks,
                                                        const GSeg* __restrict__ segs, int tiles_per_vec,
                                                        int total_tiles, const T* __restrict__ A, int64_t a_vs,
                                                        const T* __restrict__ B, int64_t b_vs,
                                                        TOut* __restrict__ C, int64_t c_vs) {
  constexpr int BK = 16;
  constexpr int PA = BM + 2, PB = BN + 2;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  typedef typename Acc4<T>::type V4;
  __shared__ T As[BK][PA];
  __shared__ T Bs[BK][PB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lq = lane >> 4;

  for (int g = xcd_logical_id(); g < total_tiles; g += gridDim.x) {
    const int v = g / tiles_per_vec;
    const int t = g - v * tiles_per_vec;
    const int ti_ = find_task(tasks, ntasks, t, &GTask::tile0);
    const GTask tk = tasks[ti_];
    const int local = t - tk.tile0;
    const int tm = local % tk.tiles_m, tn = local / tk.tiles_m;
    const int row0 = tm * BM, col0 = tn * BN;
    const int m = tk.m, n = tk.n;
    V4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = V4{0, 0, 0, 0};

    const int send = tk.seg0 + tk.nseg;
    SegCursor cl;  // position of the step held in registers
    cur_enter(cl, segs, tk.seg0, send);
    T ra[NA], rb[NB];
    unsigned oka = 0, okb = 0;  // in-bounds flags, applied at the LDS store so loads stay in flight
    T ralpha = (T)1;
    auto load = [&](const SegCursor& c) {
      const T* Ab = A + (int64_t)v * a_vs + c.a_off;
      const T* Bb = B + (int64_t)v * b_vs + c.b_off;
      ralpha = (T)c.alpha;
      const int kk0 = c.kk;
      oka = 0;
      okb = 0;
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int e = tid + r * 256;
        int i, kk;
        if (!TA) { i = e % BM; kk = e / BM; }
        else { kk = e % BK; i = e / BK; }
        const int gi = row0 + i, gk = kk0 + kk;
        const bool ok = gi < m && gk < c.k;
        const int gic = gi < m ? gi : m - 1, gkc = gk < c.k ? gk : c.k - 1;  // always-valid address
        ra[r] = TA ? Ab[gkc + (int64_t)gic * c.lda] : Ab[gic + (int64_t)gkc * c.lda];
        oka |= (ok ? 1u : 0u) << r;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int e = tid + r * 256;
        int j, kk;
        if (!TB) { kk = e % BK; j = e / BK; }
        else { j = e % BN; kk = e / BN; }
        const int gj = col0 + j, gk = kk0 + kk;
        const bool ok = gj < n && gk < c.k;
        const int gjc = gj < n ? gj : n - 1, gkc = gk < c.k ? gk : c.k - 1;
        rb[r] = TB ? Bb[gjc + (int64_t)gkc * c.ldb] : Bb[gkc + (int64_t)gjc * c.ldb];
        okb |= (ok ? 1u : 0u) << r;
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int e = tid + r * 256;
        int i, kk;
        if (!TA) { i = e % BM; kk = e / BM; }
        else { kk = e % BK; i = e / BK; }
        As[kk][i] = ((oka >> r) & 1u) ? ra[r] * ralpha : (T)0;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int e = tid + r * 256;
        int j, kk;
        if (!TB) { kk = e % BK; j = e / BK; }
        else { j = e % BN; kk = e / BN; }
        Bs[kk][j] = ((okb >> r) & 1u) ? rb[r] : (T)0;
      }
    };
    if (cl.si < send) load(cl);
    while (cl.si < send) {
      __syncthreads();  // previous MFMAs done reading LDS
      store();
      __syncthreads();
      cur_next(cl, segs, send, BK);
      if (cl.si < send) load(cl);
#pragma unroll
      for (int kq = 0; kq < BK / 4; ++kq) {
        T af[TM], bf[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = As[kq * 4 + lq][wm * WM + i * 16 + l16];
#pragma unroll
        for (int j = 0; j < TN; ++j) bf[j] = Bs[kq * 4 + lq][wn * WN + j * 16 + l16];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = mfma16(af[i], bf[j], acc[i][j]);
      }
    }
    // epilogue: see acc_row for the row each accumulator register holds
    TOut* Cb = C + (int64_t)v * c_vs + tk.c_off;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = col0 + wn * WN + j * 16 + l16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + wm * WM + i * 16 + acc_row<T>(lq, r);
          if (row < m && col < n) Cb[row + (int64_t)col * m] = cvt<TOut>(acc[i][j][r]);
        }
      }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// FP64 / FP32 grouped GEMM for small sector blocks: one WTxWT output tile per
// wave (WT = 16 or 32: 1 or 2x2 16x16x4 MFMA tiles), four independent waves
// per workgroup, no LDS and no barrier.  Each MFMA takes its A fragment
// (lane: row l16, k lq) and B fragment (k lq, column l16) straight from
// global memory (the blocks of a small sector stay in L1/L2); four k-steps of
// loads are issued ahead of their MFMAs, and a 32-wide tile reuses each
// fragment for two MFMAs.  Tiles are numbered as for gemm16_kernel
// (m-fastest inside a task) and every output element is accumulated over k
// in the same order, so FP64 results equal gemm16_kernel's bit for bit;
// loads clamp to valid addresses and zero what lies outside the block.
// ---------------------------------------------------------------------------
template <typename T, typename TOut, bool TA, bool TB, int WT>
__global__ __launch_bounds__(256) void gemm_wave_kernel(const GTask* __restrict__ tasks, int ntasks,
                                                        const GSeg* __restrict__ segs, int tiles_per_vec,
                                                        int total_tiles, const T* __restrict__ A, int64_t a_vs,
                                                        const T* __restrict__ B, int64_t b_vs,
                                                        TOut* __restrict__ C, int64_t c_vs) {
  typedef typename Acc4<T>::type V4;
  constexpr int NT = WT / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int nw = gridDim.x * 4;
  for (int g = xcd_logical_id() * 4 + wave; g < total_tiles; g += nw) {
    const int v = g / tiles_per_vec;
    const int t = g - v * tiles_per_vec;
    const GTask tk = tasks[find_task(tasks, ntasks, t, &GTask::tile0)];
    const int local = t - tk.tile0;
    const int row0 = (local % tk.tiles_m) * WT, col0 = (local / tk.tiles_m) * WT;
    const int m = tk.m, n = tk.n;
    bool rok[NT], cok[NT];
    int rowc[NT], colc[NT];  // always-valid addresses
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int row = row0 + 16 * i + l16, col = col0 + 16 * i + l16;
      rok[i] = row < m;
      cok[i] = col < n;
      rowc[i] = rok[i] ? row : m - 1;
      colc[i] = cok[i] ? col : n - 1;
    }
    V4 acc[NT][NT];
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = V4{0, 0, 0, 0};
    const int send = tk.seg0 + tk.nseg;
    for (int si = tk.seg0; si < send; ++si) {
      const GSeg sg = segs[uni(si)];
      const int k = uni(sg.k), lda = uni(sg.lda), ldb = uni(sg.ldb);
      const T* Ab = A + (int64_t)v * a_vs + sg.a_off;
      const T* Bb = B + (int64_t)v * b_vs + sg.b_off;
      const T alpha = (T)sg.alpha;
      for (int k0 = 0; k0 < k; k0 += 16) {
        T af[4][NT], bf[4][NT];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int kk = k0 + 4 * q + lq;
          const bool kok = kk < k;
          const int kc = kok ? kk : k - 1;
#pragma unroll
          for (int i = 0; i < NT; ++i) {
            const T a = TA ? Ab[kc + (int64_t)rowc[i] * lda] : Ab[rowc[i] + (int64_t)kc * lda];
            const T b = TB ? Bb[colc[i] + (int64_t)kc * ldb] : Bb[kc + (int64_t)colc[i] * ldb];
            af[q][i] = (rok[i] && kok) ? a * alpha : (T)0;
            bf[q][i] = (cok[i] && kok) ? b : (T)0;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (k0 + 4 * q < k)  // uniform: k0, q and k are the same in every lane
#pragma unroll
            for (int i = 0; i < NT; ++i)
#pragma unroll
              for (int j = 0; j < NT; ++j) acc[i][j] = mfma16(af[q][i], bf[q][j], acc[i][j]);
      }
    }
    TOut* Cb = C + (int64_t)v * c_vs + tk.c_off;
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int orow = row0 + 16 * i + acc_row<T>(lq, r);
          const int ocol = col0 + 16 * j + l16;
          if (orow < m && ocol < n) Cb[orow + (int64_t)ocol * m] = cvt<TOut>(acc[i][j][r]);
        }
  }
}

// Output store of the bf16 kernels: doubles as-is, bf16 as one plane or a
// hi/lo pair (hi = bf16(x), lo = bf16(x - hi)).
template <typename TOut, bool SPLIT>
__device__ __forceinline__ void store_out(TOut* C, int64_t idx, int64_t plane, float x);
template <>
__device__ __forceinline__ void store_out<double, false>(double* C, int64_t idx, int64_t, float x) {
  C[idx] = (double)x;
}
template <>
__device__ __forceinline__ void store_out<double, true>(double* C, int64_t idx, int64_t, float x) {
  C[idx] = (double)x;
}
template <>
__device__ __forceinline__ void store_out<__hip_bfloat16, false>(__hip_bfloat16* C, int64_t idx, int64_t, float x) {
  C[idx] = __float2bfloat16(x);
}
template <>
__device__ __forceinline__ void store_out<__hip_bfloat16, true>(__hip_bfloat16* C, int64_t idx, int64_t plane,
                                                                float x) {
  const __hip_bfloat16 h = __float2bfloat16(x);
  C[idx] = h;
  C[idx + plane] = __float2bfloat16(x - __bfloat162float(h));
}

// ---------------------------------------------------------------------------
// BF16 grouped GEMM on v_mfma_f32_16x16x32_bf16.  Operands are stored
// k-contiguous: op(A)(i,k) = A[i*lda + k], op(B)(k,j) = B[j*ldb + k], with
// lda, ldb, offsets multiples of 8 and the k padding zero-filled, so every
// fragment is one 16-byte load.  C is written either column-major (ld = m) in
// TOut, or row-major (ld = n, padded to 8) when c_rowmajor is set (the layout
// the next stage reads as a k-contiguous A operand).
//
// SPLIT: every operand is a pair of bf16 planes x = hi + lo (lo at a fixed
// plane offset), and each product is hi*hi + hi*lo + lo*hi on MFMA with fp32
// accumulation: about 16 significant bits per operand instead of 8.  Outputs
// in bf16 are written back as hi/lo planes.
// ---------------------------------------------------------------------------
template <typename TOut, int BM, int BN, bool SPLIT>
__global__ __launch_bounds__(256, 2) void gemm_bf16_kernel(const GTask* __restrict__ tasks, int ntasks,
                                                           const GSeg* __restrict__ segs, int tiles_per_vec,
                                                           int total_tiles, const __hip_bfloat16* __restrict__ A,
                                                           int64_t a_vs, int64_t a_plane,
                                                           const __hip_bfloat16* __restrict__ B, int64_t b_vs,
                                                           int64_t b_plane, TOut* __restrict__ C, int64_t c_vs,
                                                           int64_t c_plane, int c_rowmajor) {
  constexpr int BK = 32, PK = BK + 8;
  constexpr int NP = SPLIT ? 2 : 1;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  // 16-byte vectors per plane of a tile slab, and per thread (a 32-row tile
  // has 128: half the threads load it)
  constexpr int NVA = BM * BK / 8, NVB = BN * BK / 8;
  constexpr int NA = (NVA + 255) / 256, NB = (NVB + 255) / 256;
  __shared__ __attribute__((aligned(16))) unsigned short As[NP][BM][PK];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[NP][BN][PK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lq = lane >> 4;
  const unsigned short* Au = reinterpret_cast<const unsigned short*>(A);
  const unsigned short* Bu = reinterpret_cast<const unsigned short*>(B);

  for (int g = xcd_logical_id(); g < total_tiles; g += gridDim.x) {
    const int v = g / tiles_per_vec;
    const int t = g - v * tiles_per_vec;
    const GTask tk = tasks[find_task(tasks, ntasks, t, &GTask::tile0)];
    const int local = t - tk.tile0;
    const int tm = local % tk.tiles_m, tn = local / tk.tiles_m;
    const int row0 = tm * BM, col0 = tn * BN;
    const int m = tk.m, n = tk.n;
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

    const int send = tk.seg0 + tk.nseg;
    SegCursor cl;  // step held in registers
    cur_enter(cl, segs, tk.seg0, send);
    u16x8 ra[NP][NA], rb[NP][NB];
    unsigned oka = 0, okb = 0;  // in-bounds flags, applied at the LDS store so loads stay in flight
    float alpha = 1.f;
    auto load = [&](const SegCursor& c) {
      oka = 0;
      okb = 0;
      const unsigned short* Ab = Au + (int64_t)v * a_vs + c.a_off;
      const unsigned short* Bb = Bu + (int64_t)v * b_vs + c.b_off;
      const int kpad = (c.k + 7) & ~7;
      const int kk0 = c.kk;
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int e = tid + r * 256;
        if (NVA % 256 != 0 && e >= NVA) continue;
        const int i = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        const int gi = row0 + i, gk = kk0 + kc;
        const bool ok = gi < m && gk < kpad;
        const int gic = gi < m ? gi : m - 1, gkc = gk < kpad ? gk : kpad - 8;  // always-valid address
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          u16x8 val = *reinterpret_cast<const u16x8*>(Ab + p * a_plane + (int64_t)gic * c.lda + gkc);
          ra[p][r] = val;
        }
        oka |= (ok ? 1u : 0u) << r;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int e = tid + r * 256;
        if (NVB % 256 != 0 && e >= NVB) continue;
        const int j = e / (BK / 8), kc = (e % (BK / 8)) * 8;
        const int gj = col0 + j, gk = kk0 + kc;
        const bool ok = gj < n && gk < kpad;
        const int gjc = gj < n ? gj : n - 1, gkc = gk < kpad ? gk : kpad - 8;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          u16x8 val = *reinterpret_cast<const u16x8*>(Bb + p * b_plane + (int64_t)gjc * c.ldb + gkc);
          rb[p][r] = val;
        }
        okb |= (ok ? 1u : 0u) << r;
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int r = 0; r < NA; ++r) {
          const int e = tid + r * 256;
          if (NVA % 256 != 0 && e >= NVA) continue;
          *reinterpret_cast<u16x8*>(&As[p][e / (BK / 8)][(e % (BK / 8)) * 8]) =
              ((oka >> r) & 1u) ? ra[p][r] : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          const int e = tid + r * 256;
          if (NVB % 256 != 0 && e >= NVB) continue;
          *reinterpret_cast<u16x8*>(&Bs[p][e / (BK / 8)][(e % (BK / 8)) * 8]) =
              ((okb >> r) & 1u) ? rb[p][r] : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
      }
    };
    if (cl.si < send) load(cl);
    while (cl.si < send) {
      __syncthreads();
      store();
      __syncthreads();
      alpha = (float)cl.alpha;
      cur_next(cl, segs, send, BK);
      if (cl.si < send) load(cl);
      bf8 ah[TM], bh[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) ah[i] = *reinterpret_cast<const bf8*>(&As[0][wm * WM + i * 16 + l16][lq * 8]);
#pragma unroll
      for (int j = 0; j < TN; ++j) bh[j] = *reinterpret_cast<const bf8*>(&Bs[0][wn * WN + j * 16 + l16][lq * 8]);
      f4 part[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          part[i][j] = (alpha == 1.f) ? acc[i][j] : f4{0, 0, 0, 0};
      if (SPLIT) {
        bf8 al[TM], bl[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
          al[i] = *reinterpret_cast<const bf8*>(&As[NP - 1][wm * WM + i * 16 + l16][lq * 8]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          bl[j] = *reinterpret_cast<const bf8*>(&Bs[NP - 1][wn * WN + j * 16 + l16][lq * 8]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], part[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], part[i][j], 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], part[i][j], 0, 0, 0);
      if (alpha == 1.f) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = part[i][j];
      } else {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] += alpha * part[i][j];
      }
    }
    TOut* Cb = C + (int64_t)v * c_vs + tk.c_off;
    const int ldr = (n + 31) & ~31;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = col0 + wn * WN + j * 16 + l16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + wm * WM + i * 16 + lq * 4 + r;
          if (row < m && col < n) {
            const int64_t idx = c_rowmajor ? ((int64_t)row * ldr + col) : (row + (int64_t)col * m);
            store_out<TOut, SPLIT>(Cb, idx, c_plane, acc[i][j][r]);
          }
        }
      }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Streaming BF16 grouped GEMM: operands go straight from global memory into
// LDS with global_load_lds_dwordx4 (no register staging), NSTAGE k-steps in
// flight, one barrier per k-step.  Workgroup tile 256 x BN: BN = 128 with four
// waves of 128 x 64, BN = 256 with eight waves of 128 x 64.  Requirements (met by the plan's bf16 layouts): k
// dimensions padded with zeros to a multiple of 32, so every k-step is a full
// 32-wide slab.  Rows past m (columns past n) are loaded from a clamped, valid
// address and their garbage results are never stored.  LDS rows are 64 bytes
// with the 16-byte chunk index XOR-ed by (row >> 1) & 3: a ds_read_b128 serves
// 8 lanes (128 B) per cycle, and 8 consecutive rows at one logical chunk then
// land on 8 distinct 16-byte bank groups (rows 2 apart are 128 B apart).  Segment scales are applied
// by rescaling the accumulators when the scale changes.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void* lds_vptr;
typedef __attribute__((address_space(1))) void* gbl_vptr;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Fragment reads from LDS.  The reads are issued from inline asm (the
// compiler's own LDS loads would be made to wait for every LDS-DMA load still
// in flight, since it cannot tell the pipeline stages apart), and the same asm
// statement waits for them (lgkmcnt(0)) before it ends: an asm statement's
// outputs must hold their values when the statement completes.  Each address
// operand starts a group of fragments 1024 B apart (16 rows of 64 B; the row
// swizzle depends only on the lane).
__device__ __forceinline__ void lds_rd_4(unsigned a0, bf8 (&f)[4]) {
  u32x4 r0, r1, r2, r3;
  asm volatile(
      "ds_read_b128 %0, %4\n\t"
      "ds_read_b128 %1, %4 offset:1024\n\t"
      "ds_read_b128 %2, %4 offset:2048\n\t"
      "ds_read_b128 %3, %4 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3)
      : "v"(a0));
  f[0] = __builtin_bit_cast(bf8, r0);
  f[1] = __builtin_bit_cast(bf8, r1);
  f[2] = __builtin_bit_cast(bf8, r2);
  f[3] = __builtin_bit_cast(bf8, r3);
}

__device__ __forceinline__ void lds_rd_4_4(unsigned a0, unsigned a1, bf8 (&f)[8]) {
  u32x4 r0, r1, r2, r3, r4, r5, r6, r7;
  asm volatile(
      "ds_read_b128 %0, %8\n\t"
      "ds_read_b128 %1, %8 offset:1024\n\t"
      "ds_read_b128 %2, %8 offset:2048\n\t"
      "ds_read_b128 %3, %8 offset:3072\n\t"
      "ds_read_b128 %4, %9\n\t"
      "ds_read_b128 %5, %9 offset:1024\n\t"
      "ds_read_b128 %6, %9 offset:2048\n\t"
      "ds_read_b128 %7, %9 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6), "=&v"(r7)
      : "v"(a0), "v"(a1));
  f[0] = __builtin_bit_cast(bf8, r0);
  f[1] = __builtin_bit_cast(bf8, r1);
  f[2] = __builtin_bit_cast(bf8, r2);
  f[3] = __builtin_bit_cast(bf8, r3);
  f[4] = __builtin_bit_cast(bf8, r4);
  f[5] = __builtin_bit_cast(bf8, r5);
  f[6] = __builtin_bit_cast(bf8, r6);
  f[7] = __builtin_bit_cast(bf8, r7);
}

__device__ __forceinline__ void lds_rd_4_8(unsigned a0, unsigned a1, bf8 (&f)[12]) {
  u32x4 r0, r1, r2, r3, r4, r5, r6, r7, r8, r9, r10, r11;
  asm volatile(
      "ds_read_b128 %0, %12\n\t"
      "ds_read_b128 %1, %12 offset:1024\n\t"
      "ds_read_b128 %2, %12 offset:2048\n\t"
      "ds_read_b128 %3, %12 offset:3072\n\t"
      "ds_read_b128 %4, %13\n\t"
      "ds_read_b128 %5, %13 offset:1024\n\t"
      "ds_read_b128 %6, %13 offset:2048\n\t"
      "ds_read_b128 %7, %13 offset:3072\n\t"
      "ds_read_b128 %8, %13 offset:4096\n\t"
      "ds_read_b128 %9, %13 offset:5120\n\t"
      "ds_read_b128 %10, %13 offset:6144\n\t"
      "ds_read_b128 %11, %13 offset:7168\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6), "=&v"(r7), "=&v"(r8), "=&v"(r9), "=&v"(r10), "=&v"(r11)
      : "v"(a0), "v"(a1));
  f[0] = __builtin_bit_cast(bf8, r0);
  f[1] = __builtin_bit_cast(bf8, r1);
  f[2] = __builtin_bit_cast(bf8, r2);
  f[3] = __builtin_bit_cast(bf8, r3);
  f[4] = __builtin_bit_cast(bf8, r4);
  f[5] = __builtin_bit_cast(bf8, r5);
  f[6] = __builtin_bit_cast(bf8, r6);
  f[7] = __builtin_bit_cast(bf8, r7);
  f[8] = __builtin_bit_cast(bf8, r8);
  f[9] = __builtin_bit_cast(bf8, r9);
  f[10] = __builtin_bit_cast(bf8, r10);
  f[11] = __builtin_bit_cast(bf8, r11);
}

__device__ __forceinline__ void lds_rd_4_4_4_4(unsigned a0, unsigned a1, unsigned a2, unsigned a3, bf8 (&f)[16]) {
  u32x4 r0, r1, r2, r3, r4, r5, r6, r7, r8, r9, r10, r11, r12, r13, r14, r15;
  asm volatile(
      "ds_read_b128 %0, %16\n\t"
      "ds_read_b128 %1, %16 offset:1024\n\t"
      "ds_read_b128 %2, %16 offset:2048\n\t"
      "ds_read_b128 %3, %16 offset:3072\n\t"
      "ds_read_b128 %4, %17\n\t"
      "ds_read_b128 %5, %17 offset:1024\n\t"
      "ds_read_b128 %6, %17 offset:2048\n\t"
      "ds_read_b128 %7, %17 offset:3072\n\t"
      "ds_read_b128 %8, %18\n\t"
      "ds_read_b128 %9, %18 offset:1024\n\t"
      "ds_read_b128 %10, %18 offset:2048\n\t"
      "ds_read_b128 %11, %18 offset:3072\n\t"
      "ds_read_b128 %12, %19\n\t"
      "ds_read_b128 %13, %19 offset:1024\n\t"
      "ds_read_b128 %14, %19 offset:2048\n\t"
      "ds_read_b128 %15, %19 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6), "=&v"(r7), "=&v"(r8), "=&v"(r9), "=&v"(r10), "=&v"(r11), "=&v"(r12), "=&v"(r13), "=&v"(r14), "=&v"(r15)
      : "v"(a0), "v"(a1), "v"(a2), "v"(a3));
  f[0] = __builtin_bit_cast(bf8, r0);
  f[1] = __builtin_bit_cast(bf8, r1);
  f[2] = __builtin_bit_cast(bf8, r2);
  f[3] = __builtin_bit_cast(bf8, r3);
  f[4] = __builtin_bit_cast(bf8, r4);
  f[5] = __builtin_bit_cast(bf8, r5);
  f[6] = __builtin_bit_cast(bf8, r6);
  f[7] = __builtin_bit_cast(bf8, r7);
  f[8] = __builtin_bit_cast(bf8, r8);
  f[9] = __builtin_bit_cast(bf8, r9);
  f[10] = __builtin_bit_cast(bf8, r10);
  f[11] = __builtin_bit_cast(bf8, r11);
  f[12] = __builtin_bit_cast(bf8, r12);
  f[13] = __builtin_bit_cast(bf8, r13);
  f[14] = __builtin_bit_cast(bf8, r14);
  f[15] = __builtin_bit_cast(bf8, r15);
}

__device__ __forceinline__ void lds_rd_4_4_8_8(unsigned a0, unsigned a1, unsigned a2, unsigned a3, bf8 (&f)[24]) {
  u32x4 r0, r1, r2, r3, r4, r5, r6, r7, r8, r9, r10, r11, r12, r13, r14, r15, r16, r17, r18, r19, r20, r21, r22, r23;
  asm volatile(
      "ds_read_b128 %0, %24\n\t"
      "ds_read_b128 %1, %24 offset:1024\n\t"
      "ds_read_b128 %2, %24 offset:2048\n\t"
      "ds_read_b128 %3, %24 offset:3072\n\t"
      "ds_read_b128 %4, %25\n\t"
      "ds_read_b128 %5, %25 offset:1024\n\t"
      "ds_read_b128 %6, %25 offset:2048\n\t"
      "ds_read_b128 %7, %25 offset:3072\n\t"
      "ds_read_b128 %8, %26\n\t"
      "ds_read_b128 %9, %26 offset:1024\n\t"
      "ds_read_b128 %10, %26 offset:2048\n\t"
      "ds_read_b128 %11, %26 offset:3072\n\t"
      "ds_read_b128 %12, %26 offset:4096\n\t"
      "ds_read_b128 %13, %26 offset:5120\n\t"
      "ds_read_b128 %14, %26 offset:6144\n\t"
      "ds_read_b128 %15, %26 offset:7168\n\t"
      "ds_read_b128 %16, %27\n\t"
      "ds_read_b128 %17, %27 offset:1024\n\t"
      "ds_read_b128 %18, %27 offset:2048\n\t"
      "ds_read_b128 %19, %27 offset:3072\n\t"
      "ds_read_b128 %20, %27 offset:4096\n\t"
      "ds_read_b128 %21, %27 offset:5120\n\t"
      "ds_read_b128 %22, %27 offset:6144\n\t"
      "ds_read_b128 %23, %27 offset:7168\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6), "=&v"(r7), "=&v"(r8), "=&v"(r9), "=&v"(r10), "=&v"(r11), "=&v"(r12), "=&v"(r13), "=&v"(r14), "=&v"(r15), "=&v"(r16), "=&v"(r17), "=&v"(r18), "=&v"(r19), "=&v"(r20), "=&v"(r21), "=&v"(r22), "=&v"(r23)
      : "v"(a0), "v"(a1), "v"(a2), "v"(a3));
  f[0] = __builtin_bit_cast(bf8, r0);
  f[1] = __builtin_bit_cast(bf8, r1);
  f[2] = __builtin_bit_cast(bf8, r2);
  f[3] = __builtin_bit_cast(bf8, r3);
  f[4] = __builtin_bit_cast(bf8, r4);
  f[5] = __builtin_bit_cast(bf8, r5);
  f[6] = __builtin_bit_cast(bf8, r6);
  f[7] = __builtin_bit_cast(bf8, r7);
  f[8] = __builtin_bit_cast(bf8, r8);
  f[9] = __builtin_bit_cast(bf8, r9);
  f[10] = __builtin_bit_cast(bf8, r10);
  f[11] = __builtin_bit_cast(bf8, r11);
  f[12] = __builtin_bit_cast(bf8, r12);
  f[13] = __builtin_bit_cast(bf8, r13);
  f[14] = __builtin_bit_cast(bf8, r14);
  f[15] = __builtin_bit_cast(bf8, r15);
  f[16] = __builtin_bit_cast(bf8, r16);
  f[17] = __builtin_bit_cast(bf8, r17);
  f[18] = __builtin_bit_cast(bf8, r18);
  f[19] = __builtin_bit_cast(bf8, r19);
  f[20] = __builtin_bit_cast(bf8, r20);
  f[21] = __builtin_bit_cast(bf8, r21);
  f[22] = __builtin_bit_cast(bf8, r22);
  f[23] = __builtin_bit_cast(bf8, r23);
}

__device__ __forceinline__ unsigned lds_u32(const unsigned char* p) {
  return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) unsigned char*)p;
}

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  // gfx9 s_waitcnt: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14]
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | (((N >> 4) & 3) << 14));
}

template <int TN, bool SPLIT>
__device__ __forceinline__ void mfma_rows(f4 (&acc)[8][TN], const bf8* ah, const bf8* al, const bf8* bh,
                                          const bf8* bl, int i0) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
      acc[i0 + i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i0 + i][j], 0, 0, 0);
  if constexpr (SPLIT) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i0 + i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i0 + i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i0 + i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i0 + i][j], 0, 0, 0);
  }
}

// UNIT: every segment has scale 1 (no rescaling code in the k-loop).
// NW waves: 4 (2 x 2 wave grid) or 8 (2 x 4, two waves per SIMD, so one
// wave's LDS reads overlap the other's MFMAs).
template <typename TOut, bool SPLIT, int NSTAGE, int BN, bool UNIT, int NW>
__global__ __launch_bounds__(NW * 64, 1) void gemm_bf16_stream_kernel(
    const GTask* __restrict__ tasks, int ntasks, const GSeg* __restrict__ segs, int tiles_per_vec, int total_tiles,
    const __hip_bfloat16* __restrict__ A, int64_t a_vs, int64_t a_plane, const __hip_bfloat16* __restrict__ B,
    int64_t b_vs, int64_t b_plane, TOut* __restrict__ C, int64_t c_vs, int64_t c_plane, int c_rowmajor) {
  constexpr int BM = 256, BK = 32;
  constexpr int NP = SPLIT ? 2 : 1;
  constexpr int WM = 128, WN = BN / (NW / 2), TM = WM / 16, TN = WN / 16;
  constexpr int ROWB = BK * 2;                        // bytes per LDS row (one plane)
  constexpr int A_BYTES = BM * ROWB, B_BYTES = BN * ROWB;
  constexpr int STAGE_BYTES = NP * (A_BYTES + B_BYTES);
  constexpr int LA = BM / (16 * NW), LB = BN / (16 * NW);  // load instructions per wave and plane
  constexpr int LOADS = NP * (LA + LB);               // per wave per stage
  static_assert(TM == 8 && (TN == 4 || TN == 8), "wave tile 128 x 64 or 128 x 128");
  static_assert(NSTAGE * STAGE_BYTES <= 160 * 1024, "LDS budget");
  __shared__ __attribute__((aligned(1024))) unsigned char smem[NSTAGE * STAGE_BYTES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int l16 = lane & 15, lq = lane >> 4;
  const unsigned short* Au = reinterpret_cast<const unsigned short*>(A);
  const unsigned short* Bu = reinterpret_cast<const unsigned short*>(B);
  // byte offsets of fragment 0 inside a plane; the swizzle (r >> 1) & 3 of row
  // r = base + 16 i + l16 does not depend on i
  const int aoff0 = (wm * WM + l16) * ROWB + ((lq ^ ((l16 >> 1) & 3)) << 4);
  const int boff0 = (wn * WN + l16) * ROWB + ((lq ^ ((l16 >> 1) & 3)) << 4);
  // loader lane mapping: instruction q of this wave covers rows 16*(wave*L + q) + lane/4
  const int lrow = lane >> 2, lslot = lane & 3;

  for (int g = xcd_logical_id(); g < total_tiles; g += gridDim.x) {
    const int v = g / tiles_per_vec;
    const int t = g - v * tiles_per_vec;
    const GTask tk = tasks[find_task(tasks, ntasks, t, &GTask::tile0)];
    const int local = t - tk.tile0;
    const int tm = local % tk.tiles_m, tn = local / tk.tiles_m;
    const int row0 = tm * BM, col0 = tn * BN;
    const int m = tk.m, n = tk.n;
    const int send = tk.seg0 + tk.nseg;
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

    auto issue = [&](const SegCursor& c, int stage) {
      unsigned char* base = smem + stage * STAGE_BYTES;
      const unsigned short* Ab = Au + (int64_t)v * a_vs + c.a_off;
      const unsigned short* Bb = Bu + (int64_t)v * b_vs + c.b_off;
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
          const int rb = 16 * (wave * LA + q);
          const int r = rb + lrow;
          const int gi = row0 + r < m ? row0 + r : m - 1;
          const int chunk = lslot ^ ((r >> 1) & 3);
          const unsigned short* src = Ab + p * a_plane + (int64_t)gi * c.lda + c.kk + chunk * 8;
          __builtin_amdgcn_global_load_lds((gbl_vptr)src, (lds_vptr)(base + p * A_BYTES + rb * ROWB), 16, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
          const int rb = 16 * (wave * LB + q);
          const int r = rb + lrow;
          const int gj = col0 + r < n ? col0 + r : n - 1;
          const int chunk = lslot ^ ((r >> 1) & 3);
          const unsigned short* src = Bb + p * b_plane + (int64_t)gj * c.ldb + c.kk + chunk * 8;
          __builtin_amdgcn_global_load_lds((gbl_vptr)src,
                                           (lds_vptr)(base + NP * A_BYTES + p * B_BYTES + rb * ROWB), 16, 0, 0);
        }
      }
    };

    // cursors of the NSTAGE-1 steps loaded ahead of the computed one
    SegCursor cc;               // step being computed
    cur_enter(cc, segs, tk.seg0, send);
    SegCursor cl = cc;          // next step to issue
    int issued = 0;             // steps issued ahead of cc (including cc)
#pragma unroll
    for (int st = 0; st < NSTAGE - 1; ++st) {
      if (cl.si < send) {
        issue(cl, st);
        cur_next(cl, segs, send, BK);
        ++issued;
      }
    }
    float ascale = (cc.si < send) ? (float)cc.alpha : 1.f;
    int stage = 0;
    while (cc.si < send) {
      // wait for this step's slab: allow the loads of the later issued steps to stay in flight
      const int ahead = issued - 1;
      if (NSTAGE > 2 && ahead >= 2) wait_vmcnt<(2 * LOADS < 63 ? 2 * LOADS : 63)>();
      else if (ahead == 1) wait_vmcnt<LOADS>();
      else wait_vmcnt<0>();
      // Every wave's slab is in LDS (each wave waited for its own loads above) and
      // every wave is done with the previous step (its fragment reads completed
      // inside the reader's asm).  A bare barrier: __syncthreads' release fence
      // would also wait for the loads still in flight for later steps.
      asm volatile("s_barrier" ::: "memory");
      // refill the buffer freed by the previous step
      if (cl.si < send) {
        issue(cl, (stage + NSTAGE - 1) % NSTAGE);
        cur_next(cl, segs, send, BK);
        ++issued;
      }
      const float a = UNIT ? 1.f : (float)cc.alpha;
      if (!UNIT && a != ascale && a != 0.f) {
        const float f = ascale / a;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] *= f;
        ascale = a;
      }
      if (UNIT || a != 0.f) {
        const unsigned char* base = smem + stage * STAGE_BYTES;
        const unsigned aaddr = lds_u32(base + aoff0);
        const unsigned baddr = lds_u32(base + NP * A_BYTES + boff0);
        // first half: A fragments 0..3 (hi, lo) and every B fragment (hi, lo)
        bf8 bh[TN], bl[TN], ah[4], al[4];
        if constexpr (!SPLIT) {
          bf8 f[4 + TN];
          if constexpr (TN == 4) lds_rd_4_4(aaddr, baddr, f);
          else lds_rd_4_8(aaddr, baddr, f);
#pragma unroll
          for (int i = 0; i < 4; ++i) ah[i] = f[i];
#pragma unroll
          for (int j = 0; j < TN; ++j) bh[j] = f[4 + j];
        } else {
          bf8 f[8 + 2 * TN];
          if constexpr (TN == 4) lds_rd_4_4_4_4(aaddr, aaddr + A_BYTES, baddr, baddr + B_BYTES, f);
          else lds_rd_4_4_8_8(aaddr, aaddr + A_BYTES, baddr, baddr + B_BYTES, f);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            ah[i] = f[i];
            al[i] = f[4 + i];
          }
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            bh[j] = f[8 + j];
            bl[j] = f[8 + TN + j];
          }
        }
        mfma_rows<TN, SPLIT>(acc, ah, al, bh, bl, 0);
        // second half: A fragments 4..7; these reads overlap the MFMAs above
        if constexpr (!SPLIT) {
          bf8 f[4];
          lds_rd_4(aaddr + 4096, f);
#pragma unroll
          for (int i = 0; i < 4; ++i) ah[i] = f[i];
        } else {
          bf8 f[8];
          lds_rd_4_4(aaddr + 4096, aaddr + A_BYTES + 4096, f);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            ah[i] = f[i];
            al[i] = f[4 + i];
          }
        }
        mfma_rows<TN, SPLIT>(acc, ah, al, bh, bl, 4);
      }
      cur_next(cc, segs, send, BK);
      --issued;
      stage = (stage + 1) % NSTAGE;
    }
    __syncthreads();  // all LDS reads of this tile are done before the next tile's first loads
    TOut* Cb = C + (int64_t)v * c_vs + tk.c_off;
    const int ldr = (n + 31) & ~31;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = col0 + wn * WN + j * 16 + l16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + wm * WM + i * 16 + lq * 4 + r;
          if (row < m && col < n) {
            const int64_t idx = c_rowmajor ? ((int64_t)row * ldr + col) : (row + (int64_t)col * m);
            store_out<TOut, SPLIT>(Cb, idx, c_plane, UNIT ? acc[i][j][r] : acc[i][j][r] * ascale);
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Pipelined 256 x 256 BF16 tile for unit segment scales.
//
// Four waves, one per SIMD, each accumulate a 128 x 128 block in AGPRs (8 x 8
// MFMA 16x16x32 tiles).  A wave issues at most one instruction per four
// cycles and an MFMA occupies its SIMD's matrix pipe for 16, so with one wave
// per SIMD every other instruction of the k-loop has to sit between MFMAs:
//   * the MFMAs are issued from inline asm, eight per statement, and the same
//     statements issue the LDS fragment reads for later MFMAs (the A
//     fragments of row i + 2 while row i runs, from a ring of four A rows in
//     VGPRs; the B fragment i of step s + 1, B double-buffered in VGPRs) and
//     the global_load_lds DMA of the k-slab D steps ahead (with its address
//     arithmetic);
//   * a fragment read fills its destination registers after its statement
//     has ended.  The registers are outputs of the issuing statement and
//     in/out operands of a later statement that waits (s_waitcnt lgkmcnt);
//     nothing else touches them in between (a test of the build checks the
//     generated code for this);
//   * k-slabs live in two LDS rings of RA and RB slots, A(s + 1) being read
//     (rows 0 and 1) during step s.  Each step issues one group, the A and B
//     slabs of step s + D; past the last slab the cursor stays on the last
//     segment, so the surplus groups re-read valid memory into free slots and
//     every step waits for the same vmcnt.  The step loop is unrolled over
//     the ring period, so every slot address is a compile-time choice.
// Per k-step and wave: 64 MFMAs (192 for the split form), 16 fragment reads
// (32), 8 DMA loads (16).  Accumulators are read only after a statement that
// waits out the MFMA result latency and names them as operands.
// TR swaps the MFMA operands: the tile is computed transposed, so a lane holds
// four consecutive columns of one row (row-major output is stored 8 bytes at
// a time).
// ---------------------------------------------------------------------------
#define MPSA_MF(C, A, B)                                                                \
  ".if %[tr]\n\tv_mfma_f32_16x16x32_bf16 %[" C "], %[" B "], %[" A "], %[" C "]\n\t" \
  ".else\n\tv_mfma_f32_16x16x32_bf16 %[" C "], %[" A "], %[" B "], %[" C "]\n\t.endif\n\t"
#define MPSA_ACC8_OUT(c)                                                                                   \
  [c0] "+a"(c[0]), [c1] "+a"(c[1]), [c2] "+a"(c[2]), [c3] "+a"(c[3]), [c4] "+a"(c[4]), [c5] "+a"(c[5]), \
      [c6] "+a"(c[6]), [c7] "+a"(c[7])
#define MPSA_B8_IN(b)                                                                                       \
  [b0] "v"(b[0]), [b1] "v"(b[1]), [b2] "v"(b[2]), [b3] "v"(b[3]), [b4] "v"(b[4]), [b5] "v"(b[5]),        \
      [b6] "v"(b[6]), [b7] "v"(b[7])
// LDS DMA of 16 bytes per lane from g + v + kk into LDS m0 + mo + OFF + 16 lane:
// the address add and the M0 write come first, an MFMA sits between them and
// the load
#define MPSA_DA(T, V, OFF) "v_add_u32 %[" T "], %[" V "], %[kk]\n\ts_add_u32 m0, %[m0], %[mo]+" OFF "\n\t"
#define MPSA_LD(T) "global_load_lds_dwordx4 %[" T "], %[g]\n\t"

// c[j] += a * b[j] for j = 0..7; two fragment reads (o0 <- LDS[ad0 + O0],
// o1 <- LDS[ad1 + O1]) are issued after the first and second MFMA and land
// asynchronously.
template <bool TR, int O0, int O1>
__device__ __forceinline__ void mm8_rd2(f4 (&c)[8], const bf8& a, const bf8 (&b)[8], bf8& o0, unsigned ad0,
                                        bf8& o1, unsigned ad1) {
  asm volatile(MPSA_MF("c0", "a", "b0")
               "ds_read_b128 %[o0], %[ad0] offset:%[f0]\n\t"
               MPSA_MF("c1", "a", "b1")
               "ds_read_b128 %[o1], %[ad1] offset:%[f1]\n\t"
               MPSA_MF("c2", "a", "b2") MPSA_MF("c3", "a", "b3") MPSA_MF("c4", "a", "b4")
               MPSA_MF("c5", "a", "b5") MPSA_MF("c6", "a", "b6") MPSA_MF("c7", "a", "b7")
               : MPSA_ACC8_OUT(c), [o0] "=&v"(o0), [o1] "=&v"(o1)
               : [a] "v"(a), MPSA_B8_IN(b), [ad0] "v"(ad0), [ad1] "v"(ad1), [f0] "i"(O0), [f1] "i"(O1),
                 [tr] "i"(TR ? 1 : 0));
}

// mm8_rd2 plus one DMA load (LDS m0b + MO, global g + v + kk)
template <bool TR, int O0, int O1, int MO>
__device__ __forceinline__ void mm8_rd2_ld1(f4 (&c)[8], const bf8& a, const bf8 (&b)[8], bf8& o0, unsigned ad0,
                                            bf8& o1, unsigned ad1, unsigned m0b, const char* g, unsigned v,
                                            unsigned kk) {
  unsigned t0;
  asm volatile(MPSA_MF("c0", "a", "b0")
               "ds_read_b128 %[o0], %[ad0] offset:%[f0]\n\t"
               MPSA_MF("c1", "a", "b1")
               "ds_read_b128 %[o1], %[ad1] offset:%[f1]\n\t"
               MPSA_DA("t0", "v0", "0")
               MPSA_MF("c2", "a", "b2")
               MPSA_LD("t0")
               MPSA_MF("c3", "a", "b3") MPSA_MF("c4", "a", "b4")
               MPSA_MF("c5", "a", "b5") MPSA_MF("c6", "a", "b6") MPSA_MF("c7", "a", "b7")
               : MPSA_ACC8_OUT(c), [o0] "=&v"(o0), [o1] "=&v"(o1), [t0] "=&v"(t0)
               : [a] "v"(a), MPSA_B8_IN(b), [ad0] "v"(ad0), [ad1] "v"(ad1), [f0] "i"(O0), [f1] "i"(O1),
                 [m0] "s"(m0b), [mo] "i"(MO), [g] "s"(g), [v0] "v"(v), [kk] "s"(kk), [tr] "i"(TR ? 1 : 0)
               : "m0", "scc");
}

// mm8_rd2 plus two DMA loads (LDS m0b + MO + 1024 q, global g + v_q + kk, q = 0, 1)
template <bool TR, int O0, int O1, int MO>
__device__ __forceinline__ void mm8_rd2_ld2(f4 (&c)[8], const bf8& a, const bf8 (&b)[8], bf8& o0, unsigned ad0,
                                            bf8& o1, unsigned ad1, unsigned m0b, const char* g, unsigned v0,
                                            unsigned v1, unsigned kk) {
  unsigned t0, t1;
  asm volatile(MPSA_MF("c0", "a", "b0")
               "ds_read_b128 %[o0], %[ad0] offset:%[f0]\n\t"
               MPSA_MF("c1", "a", "b1")
               "ds_read_b128 %[o1], %[ad1] offset:%[f1]\n\t"
               MPSA_DA("t0", "v0", "0")
               MPSA_MF("c2", "a", "b2") MPSA_MF("c3", "a", "b3")
               MPSA_LD("t0") MPSA_DA("t1", "v1", "1024")
               MPSA_MF("c4", "a", "b4") MPSA_MF("c5", "a", "b5")
               MPSA_LD("t1")
               MPSA_MF("c6", "a", "b6") MPSA_MF("c7", "a", "b7")
               : MPSA_ACC8_OUT(c), [o0] "=&v"(o0), [o1] "=&v"(o1), [t0] "=&v"(t0), [t1] "=&v"(t1)
               : [a] "v"(a), MPSA_B8_IN(b), [ad0] "v"(ad0), [ad1] "v"(ad1), [f0] "i"(O0), [f1] "i"(O1),
                 [m0] "s"(m0b), [mo] "i"(MO), [g] "s"(g), [v0] "v"(v0), [v1] "v"(v1), [kk] "s"(kk),
                 [tr] "i"(TR ? 1 : 0)
               : "m0", "scc");
}

// mm8_rd2 plus four DMA loads (LDS m0b + MO + 1024 q, global g + v[q] + kk)
template <bool TR, int O0, int O1, int MO>
__device__ __forceinline__ void mm8_rd2_ld4(f4 (&c)[8], const bf8& a, const bf8 (&b)[8], bf8& o0, unsigned ad0,
                                            bf8& o1, unsigned ad1, unsigned m0b, const char* g,
                                            const unsigned (&v)[4], unsigned kk) {
  unsigned t0, t1, t2, t3;
  asm volatile(MPSA_MF("c0", "a", "b0")
               "ds_read_b128 %[o0], %[ad0] offset:%[f0]\n\t"
               MPSA_MF("c1", "a", "b1")
               "ds_read_b128 %[o1], %[ad1] offset:%[f1]\n\t"
               MPSA_DA("t0", "v0", "0")
               MPSA_MF("c2", "a", "b2")
               MPSA_LD("t0") MPSA_DA("t1", "v1", "1024")
               MPSA_MF("c3", "a", "b3")
               MPSA_LD("t1") MPSA_DA("t2", "v2", "2048")
               MPSA_MF("c4", "a", "b4")
               MPSA_LD("t2") MPSA_DA("t3", "v3", "3072")
               MPSA_MF("c5", "a", "b5")
               MPSA_LD("t3")
               MPSA_MF("c6", "a", "b6") MPSA_MF("c7", "a", "b7")
               : MPSA_ACC8_OUT(c), [o0] "=&v"(o0), [o1] "=&v"(o1), [t0] "=&v"(t0), [t1] "=&v"(t1),
                 [t2] "=&v"(t2), [t3] "=&v"(t3)
               : [a] "v"(a), MPSA_B8_IN(b), [ad0] "v"(ad0), [ad1] "v"(ad1), [f0] "i"(O0), [f1] "i"(O1),
                 [m0] "s"(m0b), [mo] "i"(MO), [g] "s"(g), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]),
                 [v3] "v"(v[3]), [kk] "s"(kk), [tr] "i"(TR ? 1 : 0)
               : "m0", "scc");
}

// c[j] += a * b[j] for j = 0..7
template <bool TR>
__device__ __forceinline__ void mm8(f4 (&c)[8], const bf8& a, const bf8 (&b)[8]) {
  asm volatile(MPSA_MF("c0", "a", "b0") MPSA_MF("c1", "a", "b1") MPSA_MF("c2", "a", "b2") MPSA_MF("c3", "a", "b3")
               MPSA_MF("c4", "a", "b4") MPSA_MF("c5", "a", "b5") MPSA_MF("c6", "a", "b6") MPSA_MF("c7", "a", "b7")
               : MPSA_ACC8_OUT(c)
               : [a] "v"(a), MPSA_B8_IN(b), [tr] "i"(TR ? 1 : 0));
}

// mm8 plus one DMA load (LDS m0b + MO, global g + v + kk)
template <bool TR, int MO>
__device__ __forceinline__ void mm8_ld1(f4 (&c)[8], const bf8& a, const bf8 (&b)[8], unsigned m0b, const char* g,
                                        unsigned v, unsigned kk) {
  unsigned t0;
  asm volatile(MPSA_MF("c0", "a", "b0")
               MPSA_DA("t0", "v0", "0")
               MPSA_MF("c1", "a", "b1")
               MPSA_LD("t0")
               MPSA_MF("c2", "a", "b2") MPSA_MF("c3", "a", "b3") MPSA_MF("c4", "a", "b4")
               MPSA_MF("c5", "a", "b5") MPSA_MF("c6", "a", "b6") MPSA_MF("c7", "a", "b7")
               : MPSA_ACC8_OUT(c), [t0] "=&v"(t0)
               : [a] "v"(a), MPSA_B8_IN(b), [m0] "s"(m0b), [mo] "i"(MO), [g] "s"(g), [v0] "v"(v), [kk] "s"(kk),
                 [tr] "i"(TR ? 1 : 0)
               : "m0", "scc");
}

// four DMA loads outside the MFMA stream (tile prologue)
template <int MO>
__device__ __forceinline__ void dma4(unsigned m0b, const char* g, const unsigned (&v)[4], unsigned kk) {
  unsigned t0, t1, t2, t3;
  asm volatile(MPSA_DA("t0", "v0", "0") "s_nop 0\n\t" MPSA_LD("t0") MPSA_DA("t1", "v1", "1024") "s_nop 0\n\t"
               MPSA_LD("t1") MPSA_DA("t2", "v2", "2048") "s_nop 0\n\t" MPSA_LD("t2")
               MPSA_DA("t3", "v3", "3072") "s_nop 0\n\t" MPSA_LD("t3")
               : [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
               : [m0] "s"(m0b), [mo] "i"(MO), [g] "s"(g), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]),
                 [v3] "v"(v[3]), [kk] "s"(kk)
               : "m0", "scc");
}
#undef MPSA_MF
#undef MPSA_DA
#undef MPSA_LD

// Waits until at most N LDS reads are outstanding; x and y hold their data after it.
template <int N>
__device__ __forceinline__ void lgkm_wait(bf8& x, bf8& y) {
  asm volatile("s_waitcnt lgkmcnt(%[n])" : "+v"(x), "+v"(y) : [n] "i"(N));
}
template <int N>
__device__ __forceinline__ void lgkm_wait(bf8& x) {
  asm volatile("s_waitcnt lgkmcnt(%[n])" : "+v"(x) : [n] "i"(N));
}
// Waits for every LDS read; the eight B fragments of one plane and two A rows land.
__device__ __forceinline__ void lgkm_wait_all(bf8 (&b)[8], bf8& x, bf8& y) {
  asm volatile("s_waitcnt lgkmcnt(0)"
               : "+v"(b[0]), "+v"(b[1]), "+v"(b[2]), "+v"(b[3]), "+v"(b[4]), "+v"(b[5]), "+v"(b[6]), "+v"(b[7]),
                 "+v"(x), "+v"(y));
}
// Orders later accumulator reads after the MFMA result latency (nops are issued
// by the first call only).
template <bool NOPS>
__device__ __forceinline__ void acc_fence(f4 (&c)[8]) {
  if constexpr (NOPS)
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" : MPSA_ACC8_OUT(c));
  else
    asm volatile("" : MPSA_ACC8_OUT(c));
}
#undef MPSA_ACC8_OUT
#undef MPSA_B8_IN

// Synchronous reads (the statement waits for them): two fragments.
__device__ __forceinline__ void lds_rd_2(unsigned a0, unsigned a1, bf8& f0, bf8& f1) {
  asm volatile(
      "ds_read_b128 %0, %2\n\t"
      "ds_read_b128 %1, %3\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(f0), "=&v"(f1)
      : "v"(a0), "v"(a1));
}

template <int V>
struct IC {
  static constexpr int value = V;
};

// a wave-uniform pointer, held in SGPRs
__device__ __forceinline__ const char* uni_ptr(const char* p) {
  const uint64_t x = (uint64_t)(uintptr_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
  return (const char*)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

template <typename TOut, bool SPLIT, bool ROWM, int D, int RA, int RB>
__global__ __launch_bounds__(256, 1) void gemm_bf16_pipe_kernel(
    const GTask* __restrict__ tasks, int ntasks, const GSeg* __restrict__ segs, int tiles_per_vec, int total_tiles,
    const __hip_bfloat16* __restrict__ A, int64_t a_vs, int64_t a_plane, const __hip_bfloat16* __restrict__ B,
    int64_t b_vs, int64_t b_plane, TOut* __restrict__ C, int64_t c_vs, int64_t c_plane) {
  constexpr int BM = 256, BN = 256, BK = 32, NP = SPLIT ? 2 : 1;
  constexpr int ROWB = BK * 2;        // bytes per LDS row (one plane)
  constexpr int PLANE = BM * ROWB;    // one plane of one operand slab
  constexpr int SLOT = NP * PLANE;
  constexpr int G = 8 * NP;           // DMA loads per wave and step: A and B, NP planes, 4 each
  // the ring period: every slot index below is a function of the step's phase
  constexpr int P = (RA == RB && RA % 2 == 0) ? RA : (RA == 3 && RB == 2 ? 6 : 0);
  static_assert(P > 0, "ring period");
  static_assert(D >= 2 && RA >= D + 1 && RB >= D && (RA + RB) * SLOT <= 160 * 1024, "LDS rings");
  static_assert((D - 2) * G < 64, "vmcnt range");
  // the rings, and after the k-loop the staging area of a row-major epilogue:
  // 128 rows of 128 bf16 per wave, 16 bytes of padding per row
  constexpr int EP = 272;
  constexpr int SMEM = (RA + RB) * SLOT > 4 * 128 * EP ? (RA + RB) * SLOT : 4 * 128 * EP;
  __shared__ __attribute__((aligned(1024))) unsigned char smem[SMEM];
  const int wave = uni(threadIdx.x >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  // M0 of this wave's first DMA row (16 * 4 * wave) in slot 0 of each ring
  const unsigned m0a0 = (unsigned)uni((int)(lds_u32(smem) + 64 * wave * ROWB));
  const unsigned m0b0 = (unsigned)uni((int)(lds_u32(smem) + RA * SLOT + 64 * wave * ROWB));
  const char* const Ac = reinterpret_cast<const char*>(A);
  const char* const Bc = reinterpret_cast<const char*>(B);

  for (int g = xcd_logical_id(); g < total_tiles; g += gridDim.x) {
    // lane-dependent values are re-derived per tile (an opaque copy of the
    // lane id): nothing lane-dependent stays live across the tile loop
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int l16 = lane & 15, lq = lane >> 4;
    // fragment f of a 16-row group: row l16, 16-byte chunk lq, stored at chunk lq ^ ((row >> 1) & 3)
    const int foff = l16 * ROWB + ((lq ^ ((l16 >> 1) & 3)) << 4);
    // DMA lane mapping: load q of this wave covers rows 16 * (4 * wave + q) + lane / 4 of a plane;
    // lane writes LDS chunk lane & 3 with global chunk (lane & 3) ^ ((row >> 1) & 3)
    const int lrow = lane >> 2;
    const int lchunk = (lane & 3) ^ ((lrow >> 1) & 3);

    const int v = g / tiles_per_vec;
    const int t = g - v * tiles_per_vec;
    const GTask tk = tasks[find_task(tasks, ntasks, t, &GTask::tile0)];
    const int local = t - tk.tile0;
    const int tm = local % tk.tiles_m, tn = local / tk.tiles_m;
    const int row0 = tm * BM, col0 = tn * BN;
    const int m = tk.m, n = tk.n;
    const int send = tk.seg0 + tk.nseg;
    int nsteps = 0;
    for (int si = tk.seg0; si < send; ++si) nsteps += (uni(segs[si].k) + BK - 1) / BK;

    f4 acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = f4{0, 0, 0, 0};

    if (nsteps > 0) {
      // fragment addresses of every ring slot
      unsigned aS[RA], bS[RB];
#pragma unroll
      for (int q = 0; q < RA; ++q) aS[q] = lds_u32(smem) + q * SLOT + wm * 128 * ROWB + foff;
#pragma unroll
      for (int q = 0; q < RB; ++q) bS[q] = lds_u32(smem + RA * SLOT) + q * SLOT + wn * 128 * ROWB + foff;
      // DMA cursor (group = A slab + B slab of one k-step): the tile it loads
      // for, the plane bases of its segment, the per-lane byte offsets of its
      // rows and 2 kk.  Past the last slab of this tile the cursor moves on to
      // the first slabs of this workgroup's next tile: the surplus groups of the
      // last steps then bring those into L2 (their LDS copies land in free slots
      // and are dropped), ahead of the next tile's prologue.  Past the last
      // segment of a tile the cursor keeps that segment's fields and restarts at k = 0.
      int d_row0 = row0, d_col0 = col0, d_m = m, d_n = n, d_v = v, d_send = send;
      bool moved_on = false;
      SegCursor cl;
      cur_enter(cl, segs, tk.seg0, send);
      unsigned offA[4], offB[4], kk2 = 0;
      const char *sa0, *sa1, *sb0, *sb1;
      auto seg_setup = [&]() {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 16 * (4 * wave + q) + lrow;
          const int gi = d_row0 + r < d_m ? d_row0 + r : d_m - 1;
          const int gj = d_col0 + r < d_n ? d_col0 + r : d_n - 1;
          offA[q] = ((unsigned)gi * (unsigned)cl.lda + (unsigned)(lchunk * 8)) * 2u;
          offB[q] = ((unsigned)gj * (unsigned)cl.ldb + (unsigned)(lchunk * 8)) * 2u;
        }
        sa0 = uni_ptr(Ac + 2 * ((int64_t)d_v * a_vs + cl.a_off));
        sb0 = uni_ptr(Bc + 2 * ((int64_t)d_v * b_vs + cl.b_off));
        sa1 = uni_ptr(sa0 + 2 * a_plane);
        sb1 = uni_ptr(sb0 + 2 * b_plane);
        kk2 = 0;
        // an SGPR written by a VALU instruction is read by VMEM only 5 wait states later
        asm volatile("s_nop 4");
      };
      seg_setup();
      auto advance = [&]() {
        cl.kk += BK;
        kk2 += 2 * BK;
        if (cl.kk >= cl.k) {
          cur_enter(cl, segs, uni(cl.si) + 1, d_send);
          if (cl.si >= d_send && !moved_on) {
            moved_on = true;
            const int g2 = g + (int)gridDim.x;
            if (g2 < total_tiles) {
              const int v2 = g2 / tiles_per_vec;
              const int t2 = g2 - v2 * tiles_per_vec;
              const GTask tk2 = tasks[find_task(tasks, ntasks, t2, &GTask::tile0)];
              if (tk2.nseg > 0) {
                const int local2 = t2 - tk2.tile0;
                d_row0 = (local2 % tk2.tiles_m) * BM;
                d_col0 = (local2 / tk2.tiles_m) * BN;
                d_m = tk2.m;
                d_n = tk2.n;
                d_v = v2;
                d_send = tk2.seg0 + tk2.nseg;
                cur_enter(cl, segs, tk2.seg0, d_send);
              }
            }
          }
          // unconditional: the descriptor load is waited for here and not at every later use
          seg_setup();
        }
      };

      // prologue: groups 0 .. D-1 into slots 0 .. D-1
#pragma unroll
      for (int s = 0; s < D; ++s) {
        dma4<0>(m0a0 + s * SLOT, sa0, offA, kk2);
        if constexpr (SPLIT) dma4<PLANE>(m0a0 + s * SLOT, sa1, offA, kk2);
        dma4<0>(m0b0 + s * SLOT, sb0, offB, kk2);
        if constexpr (SPLIT) dma4<PLANE>(m0b0 + s * SLOT, sb1, offB, kk2);
        advance();
      }
      bf8 bh0[8], bl0[8], bh1[8], bl1[8], ah[4], al[4];
      // groups 0 and 1 present: groups 2 .. D-1 may stay in flight
      wait_vmcnt<(D - 2) * G>();
      asm volatile("s_barrier" ::: "memory");
      {
        const unsigned b0 = bS[0], a0 = aS[0];
        lds_rd_2(b0, b0 + 1024, bh0[0], bh0[1]);
        lds_rd_2(b0 + 2048, b0 + 3072, bh0[2], bh0[3]);
        lds_rd_2(b0 + 4096, b0 + 5120, bh0[4], bh0[5]);
        lds_rd_2(b0 + 6144, b0 + 7168, bh0[6], bh0[7]);
        lds_rd_2(a0, a0 + 1024, ah[0], ah[1]);
        if constexpr (SPLIT) {
          lds_rd_2(b0 + PLANE, b0 + PLANE + 1024, bl0[0], bl0[1]);
          lds_rd_2(b0 + PLANE + 2048, b0 + PLANE + 3072, bl0[2], bl0[3]);
          lds_rd_2(b0 + PLANE + 4096, b0 + PLANE + 5120, bl0[4], bl0[5]);
          lds_rd_2(b0 + PLANE + 6144, b0 + PLANE + 7168, bl0[6], bl0[7]);
          lds_rd_2(a0 + PLANE, a0 + PLANE + 1024, al[0], al[1]);
        }
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) acc_fence<false>(acc[i]);
      asm volatile("s_nop 4");

      // step of phase PH (= s mod P): A(s) in slot PH % RA, A(s + 1) in (PH + 1) % RA,
      // B(s + 1) in (PH + 1) % RB; the group of step s + D goes to A slot
      // (PH + D) % RA and B slot (PH + D) % RB
      auto step = [&](auto PH_, bf8 (&bh)[8], bf8 (&bl)[8], bf8 (&nh)[8], bf8 (&nl)[8]) {
        constexpr int PH = decltype(PH_)::value;
        constexpr int KA = PH % RA, KA1 = (PH + 1) % RA, KB1 = (PH + 1) % RB;
        constexpr int IA = (PH + D) % RA, IB = (PH + D) % RB;
        // A(s + 1) and B(s + 1) have landed: groups s + 2 .. s + D - 1 may stay in flight
        wait_vmcnt<(D - 2) * G>();
        // every wave's DMA for step s + 1 has landed and every wave is done
        // with the slots the new group overwrites
        asm volatile("s_barrier" ::: "memory");
        const unsigned a_cur = aS[KA], a_nxt = aS[KA1], b_nxt = bS[KB1];
        const unsigned kk = kk2;
        auto row = [&](auto I_) {
          constexpr int I = decltype(I_)::value;
          constexpr int R = I % 4, RN = (I + 2) % 4;
          constexpr int AO = ((I + 2) & 7) * 1024;
          const unsigned aad = I + 2 < 8 ? a_cur : a_nxt;
          if constexpr (!SPLIT) {
            // load I of the group: A rows (q = I) for I < 4, B rows (q = I - 4) after
            if constexpr (I < 4)
              mm8_rd2_ld1<ROWM, AO, I * 1024, IA * SLOT + I * 1024>(acc[I], ah[R], bh, ah[RN], aad, nh[I], b_nxt,
                                                                  m0a0, sa0, offA[I], kk);
            else
              mm8_rd2_ld1<ROWM, AO, I * 1024, IB * SLOT + (I - 4) * 1024>(acc[I], ah[R], bh, ah[RN], aad, nh[I],
                                                                        b_nxt, m0b0, sb0, offB[I - 4], kk);
            // row I + 1's A fragment came from row I - 1; the reads of row I and
            // the B read of row I - 1 may stay in flight
            if constexpr (I < 7) lgkm_wait<3>(ah[(I + 1) % 4]);
          } else {
            // rows 0 .. 5 carry the group, one load per MFMA statement (load E of
            // the group in chunk E % 3 of row E / 3): the DMA path costs the
            // issuing wave cycles per load, so the loads are spread thin
            constexpr int E0 = 3 * I, E1 = 3 * I + 1, E2 = 3 * I + 2;
            auto ld = [&](auto EC) {
              constexpr int E = decltype(EC)::value;
              constexpr bool LB = E >= 8;
              constexpr int p = (E / 4) % 2, q = E % 4;
              constexpr int MO = (LB ? IB : IA) * SLOT + p * PLANE + q * 1024;
              struct L {
                unsigned m0;
                const char* g;
                unsigned v;
              };
              return std::pair<IC<MO>, L>{IC<MO>{}, L{LB ? m0b0 : m0a0, LB ? (p ? sb1 : sb0) : (p ? sa1 : sa0),
                                                        LB ? offB[q] : offA[q]}};
            };
            if constexpr (E0 < 16) {
              const auto l0 = ld(IC<E0>{});
              mm8_rd2_ld1<ROWM, AO, AO + PLANE, decltype(l0.first)::value>(acc[I], ah[R], bh, ah[RN], aad, al[RN], aad,
                                                                           l0.second.m0, l0.second.g, l0.second.v, kk);
            } else {
              mm8_rd2<ROWM, AO, AO + PLANE>(acc[I], ah[R], bh, ah[RN], aad, al[RN], aad);
            }
            if constexpr (E1 < 16) {
              const auto l1 = ld(IC<E1>{});
              mm8_rd2_ld1<ROWM, I * 1024, I * 1024 + PLANE, decltype(l1.first)::value>(
                  acc[I], al[R], bh, nh[I], b_nxt, nl[I], b_nxt, l1.second.m0, l1.second.g, l1.second.v, kk);
            } else {
              mm8_rd2<ROWM, I * 1024, I * 1024 + PLANE>(acc[I], al[R], bh, nh[I], b_nxt, nl[I], b_nxt);
            }
            if constexpr (E2 < 16) {
              const auto l2 = ld(IC<E2>{});
              mm8_ld1<ROWM, decltype(l2.first)::value>(acc[I], ah[R], bl, l2.second.m0, l2.second.g, l2.second.v, kk);
            } else {
              mm8<ROWM>(acc[I], ah[R], bl);
            }
            // row I + 1's A fragments came from row I - 1; the reads of row I and
            // the B reads of row I - 1 may stay in flight
            if constexpr (I < 7) lgkm_wait<6>(ah[(I + 1) % 4], al[(I + 1) % 4]);
          }
        };
        row(IC<0>{});
        row(IC<1>{});
        row(IC<2>{});
        row(IC<3>{});
        row(IC<4>{});
        row(IC<5>{});
        row(IC<6>{});
        row(IC<7>{});
        lgkm_wait_all(nh, ah[0], ah[1]);
        if constexpr (SPLIT) lgkm_wait_all(nl, al[0], al[1]);
        advance();
      };
      // The loop runs whole ring periods and has its only exit at the bottom
      // (an exit between steps defeats register coalescing of the
      // accumulators); the remaining steps follow it.  Even phases read B
      // from bh0/bl0, odd phases from bh1/bl1.
      int s = 0;
      for (; s + P <= nsteps; s += P) {
        step(IC<0>{}, bh0, bl0, bh1, bl1);
        step(IC<1>{}, bh1, bl1, bh0, bl0);
        if constexpr (P > 2) {
          step(IC<2>{}, bh0, bl0, bh1, bl1);
          step(IC<3>{}, bh1, bl1, bh0, bl0);
        }
        if constexpr (P > 4) {
          step(IC<4>{}, bh0, bl0, bh1, bl1);
          step(IC<5>{}, bh1, bl1, bh0, bl0);
        }
      }
      if (s < nsteps) step(IC<0>{}, bh0, bl0, bh1, bl1);
      if (s + 1 < nsteps) step(IC<1>{}, bh1, bl1, bh0, bl0);
      if constexpr (P > 2) {
        if (s + 2 < nsteps) step(IC<2>{}, bh0, bl0, bh1, bl1);
        if (s + 3 < nsteps) step(IC<3>{}, bh1, bl1, bh0, bl0);
      }
      if constexpr (P > 4) {
        if (s + 4 < nsteps) step(IC<4>{}, bh0, bl0, bh1, bl1);
      }
      // the surplus groups are still in flight: drain them before the rings are reused
      wait_vmcnt<0>();
    }
    __syncthreads();  // every wave is done with the LDS rings before the next tile's loads

    TOut* Cb = C + (int64_t)v * c_vs + tk.c_off;
    if constexpr (ROWM && sizeof(TOut) == 2) {
      // Row-major bf16 output staged through LDS, one plane at a time: each
      // wave writes its 128 x 128 block (8 bytes per lane and tile) and reads
      // it back as 4 rows x 256 bytes per instruction, so every global store
      // writes whole 256-byte row segments.  A wave reads back only what it
      // wrote itself; LDS operations of one wave complete in order.
      unsigned char* const wst = smem + wave * 128 * EP;
      const int ldr = (n + 31) & ~31;
      // the accumulators are read only after this point (the first fence also
      // waits out the MFMA latency)
      acc_fence<true>(acc[0]);
#pragma unroll
      for (int i = 1; i < 8; ++i) acc_fence<false>(acc[i]);
#pragma unroll 1
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          // keeps row i's conversions after row i - 1's LDS writes (few live registers)
          acc_fence<false>(acc[i]);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            unsigned short h[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const __hip_bfloat16 hi = __float2bfloat16(acc[i][j][r]);
              h[r] = __builtin_bit_cast(unsigned short,
                                        p == 0 ? hi : __float2bfloat16(acc[i][j][r] - __bfloat162float(hi)));
            }
            const uint2 u = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
            *reinterpret_cast<uint2*>(wst + (16 * i + l16) * EP + (16 * j + 4 * lq) * 2) = u;
          }
        }
        unsigned short* const Cp = reinterpret_cast<unsigned short*>(Cb) + p * c_plane;
#pragma unroll 4
        for (int it = 0; it < 32; ++it) {
          const int r = it * 4 + (lane >> 4), c8 = (lane & 15) * 8;
          const uint4 val = *reinterpret_cast<const uint4*>(wst + r * EP + c8 * 2);
          const int grow = row0 + wm * 128 + r, gcol = col0 + wn * 128 + c8;
          if (grow < m) {
            unsigned short* q = Cp + (int64_t)grow * ldr + gcol;
            if (gcol + 7 < n) {
              *reinterpret_cast<uint4*>(q) = val;
            } else {
              const uint32_t w[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
              for (int e = 0; e < 8; ++e)
                if (gcol + e < n) q[e] = (unsigned short)(w[e >> 1] >> (16 * (e & 1)));
            }
          }
        }
      }
      // every wave is done with the staging area before the next tile's DMA
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else if constexpr (!ROWM && sizeof(TOut) == 8) {
      // Column-major double output staged through LDS, 32 columns at a time:
      // each wave writes 32 columns x 128 rows (32 bytes per lane and tile;
      // columns 1040 bytes apart) and reads a whole column (1 KB) back per
      // instruction, so every global store writes a 1 KB run of one column.
      constexpr int CP = 1040;
      unsigned char* const wst = smem + wave * 32 * CP;
      acc_fence<true>(acc[0]);
#pragma unroll
      for (int i = 1; i < 8; ++i) acc_fence<false>(acc[i]);
      double* const Cd = reinterpret_cast<double*>(Cb);
#pragma unroll
      for (int cg = 0; cg < 4; ++cg) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          // keeps row i's conversions after the previous LDS writes (few live registers)
          acc_fence<false>(acc[i]);
#pragma unroll
          for (int jj = 0; jj < 2; ++jj) {
            const f4 x = acc[i][2 * cg + jj];
            double* q = reinterpret_cast<double*>(wst + (16 * jj + l16) * CP + (16 * i + 4 * lq) * 8);
            q[0] = x[0];
            q[1] = x[1];
            q[2] = x[2];
            q[3] = x[3];
          }
        }
        const int grow = row0 + wm * 128 + 2 * lane;
#pragma unroll 4
        for (int c = 0; c < 32; ++c) {
          const double2 val = *reinterpret_cast<const double2*>(wst + c * CP + lane * 16);
          const int gcol = col0 + wn * 128 + 32 * cg + c;
          if (gcol < n) {
            double* q = Cd + grow + (int64_t)gcol * m;
            if (grow + 1 < m) {
              if ((((uintptr_t)q) & 15) == 0) {
                *reinterpret_cast<double2*>(q) = val;
              } else {
                q[0] = val.x;
                q[1] = val.y;
              }
            } else if (grow < m) {
              q[0] = val.x;
            }
          }
        }
      }
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        // row i of the accumulators is read only after this point (the first
        // fence also waits out the MFMA latency)
        if (i == 0) acc_fence<true>(acc[0]);
        else acc_fence<false>(acc[i]);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if constexpr (ROWM) {
            // transposed tile: row row0 + 128 wm + 16 i + l16, columns col0 + 128 wn + 16 j + 4 lq + r
            const int ldr = (n + 31) & ~31;
            const int row = row0 + wm * 128 + i * 16 + l16;
            const int col = col0 + wn * 128 + j * 16 + lq * 4;
            if (row < m) {
              TOut* p = Cb + (int64_t)row * ldr + col;
#pragma unroll
              for (int r = 0; r < 4; ++r)
                if (col + r < n) store_out<TOut, SPLIT>(p, r, c_plane, acc[i][j][r]);
            }
          } else {
            // column col0 + 128 wn + 16 j + l16, rows row0 + 128 wm + 16 i + 4 lq + r (contiguous)
            const int col = col0 + wn * 128 + j * 16 + l16;
            const int row = row0 + wm * 128 + i * 16 + lq * 4;
            if (col < n) {
              TOut* p = Cb + row + (int64_t)col * m;
              if (row + 3 < m) {
#pragma unroll
                for (int r = 0; r < 4; ++r) store_out<TOut, SPLIT>(p, r, c_plane, acc[i][j][r]);
              } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  if (row + r < m) store_out<TOut, SPLIT>(p, r, c_plane, acc[i][j][r]);
              }
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// W mix: dst = sum_t coef_t * src_t over 1024-element chunks (4 per thread).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float ldf(const T* p) {
  return (float)*p;
}
template <>
__device__ __forceinline__ float ldf<__hip_bfloat16>(const __hip_bfloat16* p) {
  return __bfloat162float(*p);
}

template <typename T, typename TC>
__global__ __launch_bounds__(256) void mix_kernel(const MixTask* __restrict__ tasks, int ntasks,
                                                  const MixTerm* __restrict__ terms, int chunks_per_vec,
                                                  int total_chunks, const T* __restrict__ src, T* __restrict__ dst,
                                                  int64_t vs) {
  for (int g = blockIdx.x; g < total_chunks; g += gridDim.x) {
    const int v = g / chunks_per_vec;
    const int c = g - v * chunks_per_vec;
    const MixTask tk = tasks[find_task(tasks, ntasks, c, &MixTask::chunk0)];
    const int64_t base = (int64_t)(c - tk.chunk0) * 1024;
    const T* sv = src + (int64_t)v * vs;
    T* dv = dst + (int64_t)v * vs + tk.dst_off;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = base + threadIdx.x + r * 256;
      if (e < tk.len) {
        TC acc = (TC)0;
        for (int q = 0; q < tk.nt; ++q) {
          const MixTerm tm = terms[tk.t0 + q];
          acc += (TC)tm.coef * (TC)ldf(sv + tm.src_off + e);
        }
        dv[e] = cvt<T>(acc);
      }
    }
  }
}
template <>
__global__ __launch_bounds__(256) void mix_kernel<double, double>(const MixTask* __restrict__ tasks, int ntasks,
                                                                  const MixTerm* __restrict__ terms,
                                                                  int chunks_per_vec, int total_chunks,
                                                                  const double* __restrict__ src,
                                                                  double* __restrict__ dst, int64_t vs) {
  for (int g = blockIdx.x; g < total_chunks; g += gridDim.x) {
    const int v = g / chunks_per_vec;
    const int c = g - v * chunks_per_vec;
    const MixTask tk = tasks[find_task(tasks, ntasks, c, &MixTask::chunk0)];
    const int64_t base = (int64_t)(c - tk.chunk0) * 1024;
    const double* sv = src + (int64_t)v * vs;
    double* dv = dst + (int64_t)v * vs + tk.dst_off;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = base + threadIdx.x + r * 256;
      if (e < tk.len) {
        double acc = 0.0;
        for (int q = 0; q < tk.nt; ++q) {
          const MixTerm tm = terms[tk.t0 + q];
          acc += tm.coef * sv[tm.src_off + e];
        }
        dv[e] = acc;
      }
    }
  }
}

// W mix on hi/lo bf16 planes: dst = split(sum_t coef_t * (hi_t + lo_t)).
// BF16 mix, 8 elements (16 bytes) per thread and plane; a 1024-element chunk
// per 128-thread block.  BF16 mix lengths are m * pad32(n) and every offset a
// multiple of 64 elements, so all accesses are aligned 16-byte vectors.
// SPLIT: sources and destination are hi/lo plane pairs (x = hi + lo).
typedef unsigned short u16x8v __attribute__((ext_vector_type(8)));
__device__ __forceinline__ float bf_bits_to_f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
template <bool SPLIT>
__global__ __launch_bounds__(128) void mix_bf16_kernel(const MixTask* __restrict__ tasks, int ntasks,
                                                       const MixTerm* __restrict__ terms, int chunks_per_vec,
                                                       int total_chunks, const __hip_bfloat16* __restrict__ src,
                                                       __hip_bfloat16* __restrict__ dst, int64_t vs, int64_t plane) {
  const unsigned short* su = reinterpret_cast<const unsigned short*>(src);
  unsigned short* du = reinterpret_cast<unsigned short*>(dst);
  for (int g = blockIdx.x; g < total_chunks; g += gridDim.x) {
    const int v = g / chunks_per_vec;
    const int c = g - v * chunks_per_vec;
    const MixTask tk = tasks[find_task(tasks, ntasks, c, &MixTask::chunk0)];
    const int64_t e = (int64_t)(c - tk.chunk0) * 1024 + threadIdx.x * 8;
    if (e >= tk.len) continue;
    const unsigned short* sv = su + (int64_t)v * vs;
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.f;
    for (int q = 0; q < tk.nt; ++q) {
      const MixTerm tm = terms[tk.t0 + q];
      const float cf = (float)tm.coef;
      const u16x8v h = *reinterpret_cast<const u16x8v*>(sv + tm.src_off + e);
      if (SPLIT) {
        const u16x8v l = *reinterpret_cast<const u16x8v*>(sv + tm.src_off + e + plane);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += cf * (bf_bits_to_f(h[k]) + bf_bits_to_f(l[k]));
      } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += cf * bf_bits_to_f(h[k]);
      }
    }
    u16x8v oh, ol;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const __hip_bfloat16 hb = __float2bfloat16(acc[k]);
      oh[k] = __builtin_bit_cast(unsigned short, hb);
      if (SPLIT) ol[k] = __builtin_bit_cast(unsigned short, __float2bfloat16(acc[k] - __bfloat162float(hb)));
    }
    unsigned short* dv = du + (int64_t)v * vs + tk.dst_off + e;
    *reinterpret_cast<u16x8v*>(dv) = oh;
    if (SPLIT) *reinterpret_cast<u16x8v*>(dv + plane) = ol;
  }
}

// ---------------------------------------------------------------------------
template <typename TO>
__global__ void convert_kernel(const double* __restrict__ src, TO* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = cvt<TO>(src[i]);
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014; constants as
// published with Vigna's splitmix64.c) of seed + i * 0x9e3779b97f4a7c15,
// top 53 bits mapped to [-1, 1).
__device__ __host__ __forceinline__ double hash_unit(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
__global__ void fill_hash_kernel(double* __restrict__ p, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = hash_unit(seed, (uint64_t)i);
}

// Gaussian entries for seeded inputs: SplitMix64 (as hash_unit) -> two
// uniforms in (0, 1] -> Box-Muller, z_i = sqrt(-2 ln u_{2i}) cos(2 pi u_{2i+1}).
// ln and cos are this library's own series (atanh series of ln m on
// [sqrt(1/2), sqrt(2)); Taylor series of sin / cos on [-pi/4, pi/4] after an
// exact octant reduction), every multiply-add an explicit fma and every other
// step a correctly rounded IEEE operation, so a host build of the same steps
// gives the same bits.
__device__ __forceinline__ double unit_open(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 1.0) * 0x1.0p-53;
}
__device__ __forceinline__ double ln_unit(double u) {
  int e = 0;
  double m = frexp(u, &e);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = 2.0 * m;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double p = 1.0 / 23.0;
  for (int k = 10; k >= 0; --k) p = fma(p, z, 1.0 / (double)(2 * k + 1));
  return fma((double)e, 0x1.62e42fefa39efp-1, 2.0 * s * p);
}
__device__ __forceinline__ double cos_2pi(double u) {
  const double x = u > 0.5 ? u - 1.0 : u;
  const double a4 = 4.0 * fabs(x);
  const double q = a4 < 0.5 ? 0.0 : (a4 < 1.5 ? 1.0 : 2.0);
  const double t = (a4 - q) * 0x1.921fb54442d18p+0, t2 = t * t;
  double f[23];
  f[0] = 1.0;
  for (int j = 1; j < 23; ++j) f[j] = f[j - 1] * (double)j;
  if (q == 1.0) {
    double sn = -1.0 / f[21];
    for (int k = 9; k >= 0; --k) sn = fma(sn, t2, ((k & 1) ? -1.0 : 1.0) / f[2 * k + 1]);
    return -(t * sn);
  }
  double c = -1.0 / f[22];
  for (int k = 10; k >= 0; --k) c = fma(c, t2, ((k & 1) ? -1.0 : 1.0) / f[2 * k]);
  return q == 0.0 ? c : -c;
}
__global__ void fill_gauss_kernel(double* __restrict__ p, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double u1 = unit_open(seed, 2 * (uint64_t)i), u2 = unit_open(seed, 2 * (uint64_t)i + 1);
    p[i] = sqrt(-2.0 * ln_unit(u1)) * cos_2pi(u2);
  }
}

__global__ void fill_strided_kernel(double* __restrict__ p, int64_t n, int64_t stride, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i * stride] = v;
}

__global__ void zero_ranges_kernel(const Range* __restrict__ r, int nr, double* __restrict__ base, int64_t vs) {
  const Range rg = r[blockIdx.y];
  double* b = base + (int64_t)blockIdx.z * vs + rg.off;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rg.len; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = 0.0;
}

// Deterministic two-level dot: fixed partition, fixed tree inside a block,
// partials summed in order by one block.
__global__ __launch_bounds__(256) void dot_partial_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                          int64_t n, int cplx, double* __restrict__ partial) {
  __shared__ double sr[256], si[256];
  double ar = 0.0, ai = 0.0;
  if (!cplx) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
      ar += y[i] * x[i];
  } else {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
      const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
      ar += yr * xr + yi * xi;  // conj(y) * x
      ai += yr * xi - yi * xr;
    }
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr[0];
    partial[2 * blockIdx.x + 1] = si[0];
  }
}
__global__ void dot_final_kernel(const double* __restrict__ partial, int nparts, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double r = 0.0, i = 0.0;
    for (int p = 0; p < nparts; ++p) {
      r += partial[2 * p];
      i += partial[2 * p + 1];
    }
    out[0] = r;
    out[1] = i;
  }
}

__global__ void axpby_kernel(double ar, double ai, const double* __restrict__ x, double br, double bi,
                             double* __restrict__ y, int64_t n, int cplx) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (!cplx) {
      y[i] = ar * x[i] + br * y[i];
    } else {
      const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
      y[2 * i] = ar * xr - ai * xi + br * yr - bi * yi;
      y[2 * i + 1] = ar * xi + ai * xr + br * yi + bi * yr;
    }
  }
}


// ---------------------------------------------------------------------------
// Batched Lanczos vector algebra.  nvec independent recurrences, vector v at
// base + v * vs.  The step number j and the table of Krylov basis vectors are
// read from device memory, so the launches of a block of steps are identical
// from block to block and can be captured once into a graph and replayed.
//   kbuf[v * ld + j] = ||W_v||            (beta of step j)
//   X_v = W_v * (1 / kbuf),  basis[j]_v = X_v
//   ebuf[v * ld + j] = <Y_v, X_v>         (alpha of step j, Y = H X)
//   W_v = (-e X_v + Y_v) + (-k basis[j-1]_v)   (the second term from j = 1 on)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bdot_partial_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                           int64_t n, int64_t vs, double* __restrict__ partial) {
  __shared__ double sr[256];
  const int v = blockIdx.y;
  const double* xv = x + (int64_t)v * vs;
  const double* yv = y + (int64_t)v * vs;
  // two independent chains per thread (more loads in flight), fixed order
  double a0 = 0.0, a1 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + stride < n; i += 2 * stride) {
    a0 += yv[i] * xv[i];
    a1 += yv[i + stride] * xv[i + stride];
  }
  if (i < n) a0 += yv[i] * xv[i];
  sr[threadIdx.x] = a0 + a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sr[threadIdx.x] += sr[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)v * gridDim.x + blockIdx.x] = sr[0];
}

__global__ void bdot_final_kernel(const double* __restrict__ partial, int nparts, double* __restrict__ out, int ld,
                                  const int* __restrict__ step, int root) {
  if (threadIdx.x != 0) return;
  const int v = blockIdx.x;
  double r = 0.0;
  for (int p = 0; p < nparts; ++p) r += partial[(int64_t)v * nparts + p];
  out[(int64_t)v * ld + *step] = root ? sqrt(r) : r;
}

__global__ void lz_normalize_kernel(const double* __restrict__ W, double* __restrict__ X, double* const* __restrict__ basis,
                                    const double* __restrict__ kbuf, int ld, const int* __restrict__ step, int64_t n,
                                    int64_t vs) {
  const int v = blockIdx.y;
  const int j = *step;
  const double r = 1.0 / kbuf[(int64_t)v * ld + j];
  const int64_t o = (int64_t)v * vs;
  double* B = basis[j] + o;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double t = r * W[o + i];
    X[o + i] = t;
    B[i] = t;
  }
}

// W = Y - alpha X - beta basis[j-1]; with wpart, also the per-workgroup sums
// of W^2 (the next step's norm, finished by bdot_final_kernel over gridDim.x
// parts).  256 threads.
__global__ __launch_bounds__(256) void lz_update_kernel(const double* __restrict__ Y, const double* __restrict__ X,
                                                        double* const* __restrict__ basis,
                                                        const double* __restrict__ kbuf,
                                                        const double* __restrict__ ebuf, int ld,
                                                        const int* __restrict__ step, double* __restrict__ W,
                                                        int64_t n, int64_t vs, double* __restrict__ wpart) {
  __shared__ double sr[256];
  const int v = blockIdx.y;
  const int j = *step;
  const double e = ebuf[(int64_t)v * ld + j], k = kbuf[(int64_t)v * ld + j];
  const int64_t o = (int64_t)v * vs;
  const double* A0 = j > 0 ? basis[j - 1] + o : nullptr;
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double t = -e * X[o + i] + Y[o + i];
    if (A0) t = t + -k * A0[i];
    W[o + i] = t;
    a += t * t;
  }
  if (!wpart) return;  // uniform across the workgroup
  sr[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sr[threadIdx.x] += sr[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) wpart[(int64_t)v * gridDim.x + blockIdx.x] = sr[0];
}

// Projection of the recurrence vector against a set of orthonormal vectors
// Q_0 .. Q_{nb-1}; vector v of Q_i lies at Q[i] + v * qvs (qvs = 0: one set
// shared by all vectors).  nb = (*cnt) + add when cnt is given (the Krylov
// basis grows with the device step counter), else add; at most nbmax.
//   c_{v,i} = <Q_i(v), W_v>   (fixed partition and fixed-order sums)
//   W_v    -= sum_i c_{v,i} Q_i(v)
// The dot products run eight basis vectors per pass over W.
__global__ __launch_bounds__(256) void multidot_partial_kernel(const double* __restrict__ W,
                                                               double* const* __restrict__ Q, int64_t qvs,
                                                               const int* __restrict__ cnt, int add, int nbmax,
                                                               int64_t n, int64_t vs, double* __restrict__ partial) {
  __shared__ double sr[8][256];
  const int v = blockIdx.y;
  const int nb = min((cnt ? *cnt : 0) + add, nbmax);
  const double* wv = W + (int64_t)v * vs;
  for (int i0 = 0; i0 < nb; i0 += 8) {
    const int g = min(8, nb - i0);
    const double* q[8];
    for (int t = 0; t < 8; ++t) q[t] = Q[i0 + min(t, g - 1)] + (int64_t)v * qvs;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
      const double w = wv[k];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (t < g) a[t] += q[t][k] * w;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) sr[t][threadIdx.x] = a[t];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        for (int t = 0; t < g; ++t) sr[t][threadIdx.x] += sr[t][threadIdx.x + w];
      __syncthreads();
    }
    if ((int)threadIdx.x < g) partial[((int64_t)v * nbmax + i0 + threadIdx.x) * gridDim.x + blockIdx.x] = sr[threadIdx.x][0];
    __syncthreads();
  }
}
__global__ void multidot_final_kernel(const double* __restrict__ partial, int nparts, const int* __restrict__ cnt,
                                      int add, int nbmax, double* __restrict__ c) {
  const int v = blockIdx.x;
  const int nb = min((cnt ? *cnt : 0) + add, nbmax);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    double r = 0.0;
    for (int p = 0; p < nparts; ++p) r += partial[((int64_t)v * nbmax + i) * nparts + p];
    c[(int64_t)v * nbmax + i] = r;
  }
}
__global__ void multisub_kernel(double* __restrict__ W, double* const* __restrict__ Q, int64_t qvs,
                                const int* __restrict__ cnt, int add, int nbmax, const double* __restrict__ c, int64_t n,
                                int64_t vs) {
  const int v = blockIdx.y;
  const int nb = min((cnt ? *cnt : 0) + add, nbmax);
  double* wv = W + (int64_t)v * vs;
  const double* cv = c + (int64_t)v * nbmax;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double t = wv[k];
    for (int i = 0; i < nb; ++i) t -= cv[i] * Q[i][(int64_t)v * qvs + k];
    wv[k] = t;
  }
}
// X_v *= 1 / sqrt(nrm2[v])
__global__ void scale_rsqrt_kernel(double* __restrict__ X, const double* __restrict__ nrm2, int64_t n, int64_t vs) {
  const int v = blockIdx.y;
  const double r = 1.0 / sqrt(nrm2[v]);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    X[(int64_t)v * vs + k] *= r;
}

__global__ void lz_step_kernel(int* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// state_v = sum_{i < m_v} c[v * ldc + i] * basis[i]_v
__global__ void lz_combine_kernel(double* const* __restrict__ basis, const double* __restrict__ c, int ldc,
                                  const int* __restrict__ m, double* __restrict__ out, int64_t n, int64_t vs) {
  const int v = blockIdx.y;
  const int mv = m[v];
  const int64_t o = (int64_t)v * vs;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double t = 0.0;
    for (int b = 0; b < mv; ++b) t += c[(int64_t)v * ldc + b] * basis[b][o + i];
    out[o + i] = t;
  }
}

// ---------------------------------------------------------------------------
// Split of a two-site tensor: block gather/scatter between packed block
// buffers and per-sector dense matrices, and the FixPhase sign rule.
// ---------------------------------------------------------------------------
// dst(i, j) = scale * src(i, j) [* rs[i]] [* cs[j]]  (column-major; src
// transposed when trans: src(j, i) with ld lds)
__global__ __launch_bounds__(256) void copy2d_kernel(const Copy2D* __restrict__ tasks, const double* __restrict__ src,
                                                     double* __restrict__ dst, const double* __restrict__ sc) {
  const Copy2D t = tasks[blockIdx.y];
  const int64_t total = (int64_t)t.rows * t.cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % t.rows, j = e / t.rows;
    double v = t.trans ? src[t.src_off + i * t.lds + j] : src[t.src_off + j * t.lds + i];
    v *= t.scale;
    if (t.rs_off >= 0) v *= sc[t.rs_off + i];
    if (t.cs_off >= 0) v *= sc[t.cs_off + j];
    dst[t.dst_off + j * t.ldd + i] = v;
  }
}

// FixPhase (real case): vector i = row i of Vt (k x n, ld ldv).  Find the first
// index of the largest |entry|, take phi = +1 if that entry is > 0 else -1,
// and multiply row i of Vt and column i of U (m x k, ld ldu) by phi.
__global__ __launch_bounds__(256) void fixphase_kernel(double* __restrict__ U, int64_t m, int64_t ldu,
                                                       double* __restrict__ Vt, int64_t n, int64_t ldv) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  const int64_t i = blockIdx.x;
  double best = 0.0;
  int64_t bi = 0;
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    const double a = fabs(Vt[i + j * ldv]);
    if (a > best) {  // strided scan: the first index of each thread's maximum
      best = a;
      bi = j;
    }
  }
  // threads that saw no entry above 0 keep (0, 0), which never beats a true entry
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double a = sv[threadIdx.x + w];
      const int64_t ai = si[threadIdx.x + w];
      if (a > sv[threadIdx.x] || (a == sv[threadIdx.x] && ai < si[threadIdx.x])) {
        sv[threadIdx.x] = a;
        si[threadIdx.x] = ai;
      }
    }
    __syncthreads();
  }
  const double phi = Vt[i + si[0] * ldv] > 0.0 ? 1.0 : -1.0;
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) Vt[i + j * ldv] *= phi;
  for (int64_t r = threadIdx.x; r < m; r += blockDim.x) U[r + i * ldu] *= phi;
}

// ---------------------------------------------------------------------------
// Repack FP64 blocks into the compute layout (type conversion, optional
// transpose, padded leading dimension).  One task per block, 1024-element
// chunks of the source block.
// ---------------------------------------------------------------------------
template <typename TO>
__global__ __launch_bounds__(256) void repack_kernel(const Repack* __restrict__ tasks, int ntasks, int chunks_per_vec,
                                                     int total, const double* __restrict__ src, int64_t s_vs,
                                                     TO* __restrict__ dst, int64_t d_vs) {
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int v = g / chunks_per_vec;
    const int c = g - v * chunks_per_vec;
    const Repack tk = tasks[find_task(tasks, ntasks, c, &Repack::chunk0)];
    const int64_t len = (int64_t)tk.rows * tk.cols;
    const int64_t base = (int64_t)(c - tk.chunk0) * 1024;
    const double* s = src + (int64_t)v * s_vs + tk.src_off;
    TO* d = dst + (int64_t)v * d_vs + tk.dst_off;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = base + threadIdx.x + r * 256;
      if (e < len) {
        // blocks hold fewer than 2^31 elements (checked when the plan is built): 32-bit index math
        const uint32_t e32 = (uint32_t)e, rows = (uint32_t)tk.rows;
        const uint32_t j = e32 / rows, i = e32 - j * rows;
        const int64_t di = tk.transpose ? ((int64_t)j + (int64_t)i * tk.dst_ld) : ((int64_t)i + (int64_t)j * tk.dst_ld);
        d[di] = cvt<TO>(tk.scale * s[e]);
      }
    }
  }
}

// Repack into hi/lo bf16 planes: hi = bf16(x), lo = bf16(x - hi).
__global__ __launch_bounds__(256) void repack_split_kernel(const Repack* __restrict__ tasks, int ntasks,
                                                           int chunks_per_vec, int total, const double* __restrict__ src,
                                                           int64_t s_vs, __hip_bfloat16* __restrict__ dst,
                                                           int64_t d_vs, int64_t plane) {
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int v = g / chunks_per_vec;
    const int c = g - v * chunks_per_vec;
    const Repack tk = tasks[find_task(tasks, ntasks, c, &Repack::chunk0)];
    const int64_t len = (int64_t)tk.rows * tk.cols;
    const int64_t base = (int64_t)(c - tk.chunk0) * 1024;
    const double* sp = src + (int64_t)v * s_vs + tk.src_off;
    __hip_bfloat16* d = dst + (int64_t)v * d_vs + tk.dst_off;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = base + threadIdx.x + r * 256;
      if (e < len) {
        const uint32_t e32 = (uint32_t)e, rows = (uint32_t)tk.rows;
        const uint32_t j = e32 / rows, i = e32 - j * rows;
        const int64_t di = tk.transpose ? ((int64_t)j + (int64_t)i * tk.dst_ld) : ((int64_t)i + (int64_t)j * tk.dst_ld);
        const double x = tk.scale * sp[e];
        const __hip_bfloat16 h = __float2bfloat16((float)x);
        d[di] = h;
        d[di + plane] = __float2bfloat16((float)(x - (double)__bfloat162float(h)));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int g_cu = 0;
static int num_cu() {
  if (!g_cu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) g_cu = p.multiProcessorCount;
    if (g_cu <= 0) g_cu = 256;
  }
  return g_cu;
}

int gemm_max_grid(int prec_in, int tile_class) {
  (void)prec_in;
  (void)tile_class;
  return num_cu() * 2;
}

template <typename T, typename TOut, int WT>
static hipError_t runwave(int tr, const GTask* tasks, int ntasks, const GSeg* segs, int tpv, int nvec, const void* A,
                          int64_t a_vs, const void* B, int64_t b_vs, void* C, int64_t c_vs, hipStream_t st) {
  const int total = tpv * nvec;
  if (total <= 0) return hipSuccess;
  int grid = num_cu() * (WT == 16 ? 8 : 4);  // 32 (16 per wave) or 16 waves per CU
  const int need = (total + 3) / 4;
  if (grid > need) grid = need;
  if (grid >= 8) grid &= ~7;
  const T* a = (const T*)A;
  const T* b = (const T*)B;
  TOut* c = (TOut*)C;
  switch (tr) {
    case 0: gemm_wave_kernel<T, TOut, false, false, WT><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
    case 1: gemm_wave_kernel<T, TOut, false, true, WT><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
    case 2: gemm_wave_kernel<T, TOut, true, false, WT><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
    default: gemm_wave_kernel<T, TOut, true, true, WT><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
  }
  return hipGetLastError();
}

template <typename T, typename TOut, int BM, int BN>
static hipError_t run16(int tr, const GTask* tasks, int ntasks, const GSeg* segs, int tpv, int nvec, const void* A,
                        int64_t a_vs, const void* B, int64_t b_vs, void* C, int64_t c_vs, hipStream_t st) {
  const int total = tpv * nvec;
  if (total <= 0) return hipSuccess;
  // persistent workgroups per CU: the small tiles (64x64: 92 VGPRs in FP64,
  // 64 in FP32; 32x32 fewer) run several per CU so that the short-k tiles of
  // small sectors overlap their load latency; 128x128 keeps two
  int grid = num_cu() * (BM == 32 ? 8 : (BM == 64 ? (sizeof(T) == 4 ? 8 : 5) : 2));
  if (grid > total) grid = total;
  if (grid >= 8) grid &= ~7;
  const T* a = (const T*)A;
  const T* b = (const T*)B;
  TOut* c = (TOut*)C;
  switch (tr) {
    case 0: gemm16_kernel<T, TOut, BM, BN, false, false><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
    case 1: gemm16_kernel<T, TOut, BM, BN, false, true><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
    case 2: gemm16_kernel<T, TOut, BM, BN, true, false><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
    default: gemm16_kernel<T, TOut, BM, BN, true, true><<<grid, 256, 0, st>>>(tasks, ntasks, segs, tpv, total, a, a_vs, b, b_vs, c, c_vs); break;
  }
  return hipGetLastError();
}

// BM == 256 selects the streaming large-tile kernel (BN 128 or 256).
template <typename TOut, int BM, int BN, bool SPLIT>
static hipError_t runbf(const GTask* tasks, int ntasks, const GSeg* segs, int tpv, int nvec, const void* A,
                        int64_t a_vs, int64_t a_pl, const void* B, int64_t b_vs, int64_t b_pl, void* C, int64_t c_vs,
                        int64_t c_pl, int c_rowmajor, hipStream_t st) {
  const int total = tpv * nvec;
  if (total <= 0) return hipSuccess;
  int grid = num_cu() * (BM == 256 ? 1 : (BM == 64 ? 5 : (BM == 32 ? 8 : 2)));  // 64x64: at most 86 VGPRs
  if (grid > total) grid = total;
  if (grid >= 8) grid &= ~7;
  if constexpr (BM == 256) {
    // 256x128: stream kernel, 4 stages (3 split) in the 160 KB LDS
    if constexpr (BN == 256) {
      // DMA distance 2 steps, rings 3 + 2 slots of 32 KB (split); 3 steps, 4 + 4 slots of 16 KB (plain)
      constexpr int D = SPLIT ? 2 : 3, RA = SPLIT ? 3 : 4, RB = SPLIT ? 2 : 4;
      if (c_rowmajor)
        gemm_bf16_pipe_kernel<TOut, SPLIT, true, D, RA, RB><<<grid, 256, 0, st>>>(
            tasks, ntasks, segs, tpv, total, (const __hip_bfloat16*)A, a_vs, a_pl, (const __hip_bfloat16*)B, b_vs,
            b_pl, (TOut*)C, c_vs, c_pl);
      else
        gemm_bf16_pipe_kernel<TOut, SPLIT, false, D, RA, RB><<<grid, 256, 0, st>>>(
            tasks, ntasks, segs, tpv, total, (const __hip_bfloat16*)A, a_vs, a_pl, (const __hip_bfloat16*)B, b_vs,
            b_pl, (TOut*)C, c_vs, c_pl);
    } else {
      constexpr int NST = SPLIT ? 3 : 4;
      gemm_bf16_stream_kernel<TOut, SPLIT, NST, BN, false, 4><<<grid, 256, 0, st>>>(
          tasks, ntasks, segs, tpv, total, (const __hip_bfloat16*)A, a_vs, a_pl, (const __hip_bfloat16*)B, b_vs,
          b_pl, (TOut*)C, c_vs, c_pl, c_rowmajor);
    }
  } else {
    gemm_bf16_kernel<TOut, BM, BN, SPLIT><<<grid, 256, 0, st>>>(
        tasks, ntasks, segs, tpv, total, (const __hip_bfloat16*)A, a_vs, a_pl, (const __hip_bfloat16*)B, b_vs, b_pl,
        (TOut*)C, c_vs, c_pl, c_rowmajor);
  }
  return hipGetLastError();
}

hipError_t launch_gemm(int prec_in, int out_kind, int tr, int tile_class, const GTask* tasks, int ntasks,
                       const GSeg* segs, int tiles_per_vec, int nvec, const void* A, int64_t a_vs, const void* B,
                       int64_t b_vs, void* C, int64_t c_vs, hipStream_t st, Planes planes) {
#define MPSA_RUN16(T_, TO_)                                                                                    \
  do {                                                                                                       \
    if (tile_class == 0)                                                                                     \
      return run16<T_, TO_, 128, 128>(tr, tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, B, b_vs, C, c_vs, st); \
    if (tile_class == 2)                                                                                     \
      return run16<T_, TO_, 32, 32>(tr, tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, B, b_vs, C, c_vs, st);   \
    if (tile_class == 3)                                                                                     \
      return runwave<T_, TO_, 16>(tr, tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, B, b_vs, C, c_vs, st);     \
    if (tile_class == 4)                                                                                     \
      return runwave<T_, TO_, 32>(tr, tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, B, b_vs, C, c_vs, st);     \
    return run16<T_, TO_, 64, 64>(tr, tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, B, b_vs, C, c_vs, st);     \
  } while (0)
  if (prec_in == F64) MPSA_RUN16(double, double);
  if (prec_in == F32) {
    if (out_kind == 1) MPSA_RUN16(float, double);
    MPSA_RUN16(float, float);
  }
#undef MPSA_RUN16
  // BF16 / BF16X3: operands in the k-contiguous layout; out_kind 0 -> bf16
  // row-major (hi/lo planes for BF16X3), 1 -> double column-major.
#define MPSA_BF_DISPATCH(TOUT, SPLIT_, APL, BPL, CPL, ROWM)                                                           \
  do {                                                                                                               \
    if (tile_class == 3)                                                                                             \
      return runbf<TOUT, 256, 256, SPLIT_>(tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, APL, B, b_vs, BPL, C, \
                                           c_vs, CPL, ROWM, st);                                                     \
    if (tile_class == 0)                                                                                             \
      return runbf<TOUT, 256, 128, SPLIT_>(tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, APL, B, b_vs, BPL, C, \
                                           c_vs, CPL, ROWM, st);                                                     \
    if (tile_class == 2)                                                                                             \
      return runbf<TOUT, 128, 128, SPLIT_>(tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, APL, B, b_vs, BPL, C, \
                                           c_vs, CPL, ROWM, st);                                                     \
    if (tile_class == 4)                                                                                             \
      return runbf<TOUT, 32, 32, SPLIT_>(tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, APL, B, b_vs, BPL, C,   \
                                         c_vs, CPL, ROWM, st);                                                       \
    return runbf<TOUT, 64, 64, SPLIT_>(tasks, ntasks, segs, tiles_per_vec, nvec, A, a_vs, APL, B, b_vs, BPL, C,     \
                                       c_vs, CPL, ROWM, st);                                                         \
  } while (0)
  if (prec_in == BF16) {
    if (out_kind == 1) MPSA_BF_DISPATCH(double, false, 0, 0, 0, 0);
    MPSA_BF_DISPATCH(__hip_bfloat16, false, 0, 0, 0, 1);
  }
  if (out_kind == 1) MPSA_BF_DISPATCH(double, true, planes.a, planes.b, 0, 0);
  MPSA_BF_DISPATCH(__hip_bfloat16, true, planes.a, planes.b, planes.c, 1);
#undef MPSA_BF_DISPATCH
}

hipError_t launch_mix(int prec, const MixTask* tasks, int ntasks, const MixTerm* terms, int chunks_per_vec, int nvec,
                      const void* src, void* dst, int64_t vs, hipStream_t st, int64_t plane) {
  const int total = chunks_per_vec * nvec;
  if (total <= 0 || ntasks <= 0) return hipSuccess;
  int grid = total < 8192 ? total : 8192;
  if (prec == BF16X3 || prec == BF16) {
    if (prec == BF16X3)
      mix_bf16_kernel<true><<<grid, 128, 0, st>>>(tasks, ntasks, terms, chunks_per_vec, total,
                                                  (const __hip_bfloat16*)src, (__hip_bfloat16*)dst, vs, plane);
    else
      mix_bf16_kernel<false><<<grid, 128, 0, st>>>(tasks, ntasks, terms, chunks_per_vec, total,
                                                   (const __hip_bfloat16*)src, (__hip_bfloat16*)dst, vs, plane);
    return hipGetLastError();
  }
  if (prec == F64)
    mix_kernel<double, double><<<grid, 256, 0, st>>>(tasks, ntasks, terms, chunks_per_vec, total, (const double*)src, (double*)dst, vs);
  else if (prec == F32)
    mix_kernel<float, float><<<grid, 256, 0, st>>>(tasks, ntasks, terms, chunks_per_vec, total, (const float*)src, (float*)dst, vs);
  return hipGetLastError();
}

hipError_t launch_convert(int prec, const double* src, void* dst, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  int grid = (int)(blocks < 16384 ? blocks : 16384);
  if (prec == F32) convert_kernel<float><<<grid, 256, 0, st>>>(src, (float*)dst, n);
  else if (prec == BF16) convert_kernel<__hip_bfloat16><<<grid, 256, 0, st>>>(src, (__hip_bfloat16*)dst, n);
  else convert_kernel<double><<<grid, 256, 0, st>>>(src, (double*)dst, n);
  return hipGetLastError();
}

hipError_t launch_repack(int prec, const Repack* tasks, int ntasks, int chunks_per_vec, int nvec, const double* src,
                         int64_t s_vs, void* dst, int64_t d_vs, hipStream_t st, int64_t plane) {
  const int total = chunks_per_vec * nvec;
  if (total <= 0 || ntasks <= 0) return hipSuccess;
  int grid = total < 8192 ? total : 8192;
  if (prec == BF16X3) {
    repack_split_kernel<<<grid, 256, 0, st>>>(tasks, ntasks, chunks_per_vec, total, src, s_vs, (__hip_bfloat16*)dst,
                                              d_vs, plane);
    return hipGetLastError();
  }
  if (prec == F32)
    repack_kernel<float><<<grid, 256, 0, st>>>(tasks, ntasks, chunks_per_vec, total, src, s_vs, (float*)dst, d_vs);
  else if (prec == BF16)
    repack_kernel<__hip_bfloat16><<<grid, 256, 0, st>>>(tasks, ntasks, chunks_per_vec, total, src, s_vs, (__hip_bfloat16*)dst, d_vs);
  else
    repack_kernel<double><<<grid, 256, 0, st>>>(tasks, ntasks, chunks_per_vec, total, src, s_vs, (double*)dst, d_vs);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  int grid = (int)(blocks < 16384 ? blocks : 16384);
  fill_kernel<<<grid, 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_hash(double* p, int64_t n, uint64_t seed, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  int grid = (int)(blocks < 16384 ? blocks : 16384);
  fill_hash_kernel<<<grid, 256, 0, st>>>(p, n, seed);
  return hipGetLastError();
}

hipError_t launch_fill_gauss(double* p, int64_t n, uint64_t seed, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  int grid = (int)(blocks < 16384 ? blocks : 16384);
  fill_gauss_kernel<<<grid, 256, 0, st>>>(p, n, seed);
  return hipGetLastError();
}

hipError_t launch_fill_strided(double* p, int64_t n, int64_t stride, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  int grid = (int)(blocks < 1024 ? blocks : 1024);
  fill_strided_kernel<<<grid, 256, 0, st>>>(p, n, stride, v);
  return hipGetLastError();
}

hipError_t launch_zero_ranges(const Range* r, int nr, double* base, int64_t vs, int nvec, hipStream_t st) {
  if (nr <= 0 || nvec <= 0) return hipSuccess;
  dim3 grid(64, nr, nvec);
  zero_ranges_kernel<<<grid, 256, 0, st>>>(r, nr, base, vs);
  return hipGetLastError();
}

hipError_t launch_dot(const double* x, const double* y, int64_t n, int cplx, double* partial, int nparts, double* out2,
                      hipStream_t st) {
  dot_partial_kernel<<<nparts, 256, 0, st>>>(x, y, n, cplx, partial);
  dot_final_kernel<<<1, 64, 0, st>>>(partial, nparts, out2);
  return hipGetLastError();
}

hipError_t launch_axpby(double ar, double ai, const double* x, double br, double bi, double* y, int64_t n, int cplx,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  int grid = (int)(blocks < 16384 ? blocks : 16384);
  axpby_kernel<<<grid, 256, 0, st>>>(ar, ai, x, br, bi, y, n, cplx);
  return hipGetLastError();
}

static int vec_grid(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return (int)(blocks < 1024 ? (blocks > 0 ? blocks : 1) : 1024);
}

hipError_t launch_bdot(const double* x, const double* y, int64_t n, int64_t vs, int nvec, double* partial, int nparts,
                       double* out, int ld, const int* step, int root, hipStream_t st) {
  bdot_partial_kernel<<<dim3(nparts, nvec), 256, 0, st>>>(x, y, n, vs, partial);
  bdot_final_kernel<<<nvec, 64, 0, st>>>(partial, nparts, out, ld, step, root);
  return hipGetLastError();
}

hipError_t launch_lz_normalize(const double* W, double* X, double* const* basis, const double* kbuf, int ld,
                               const int* step, int64_t n, int64_t vs, int nvec, hipStream_t st) {
  lz_normalize_kernel<<<dim3(vec_grid(n), nvec), 256, 0, st>>>(W, X, basis, kbuf, ld, step, n, vs);
  return hipGetLastError();
}

hipError_t launch_lz_update(const double* Y, const double* X, double* const* basis, const double* kbuf,
                            const double* ebuf, int ld, const int* step, double* W, int64_t n, int64_t vs, int nvec,
                            double* wpart, hipStream_t st) {
  lz_update_kernel<<<dim3(vec_grid(n), nvec), 256, 0, st>>>(Y, X, basis, kbuf, ebuf, ld, step, W, n, vs, wpart);
  return hipGetLastError();
}

int lz_update_parts(int64_t n) { return vec_grid(n); }

hipError_t launch_bdot_final(const double* partial, int nparts, int nvec, double* out, int ld, const int* step,
                             int root, hipStream_t st) {
  bdot_final_kernel<<<nvec, 64, 0, st>>>(partial, nparts, out, ld, step, root);
  return hipGetLastError();
}

hipError_t launch_project_out(double* W, double* const* Q, int64_t qvs, const int* cnt, int add, int nbmax, int64_t n,
                              int64_t vs, int nvec, double* partial, int nparts, double* c, hipStream_t st) {
  if (nbmax <= 0 || nvec <= 0) return hipSuccess;
  multidot_partial_kernel<<<dim3(nparts, nvec), 256, 0, st>>>(W, Q, qvs, cnt, add, nbmax, n, vs, partial);
  multidot_final_kernel<<<nvec, 256, 0, st>>>(partial, nparts, cnt, add, nbmax, c);
  multisub_kernel<<<dim3(vec_grid(n), nvec), 256, 0, st>>>(W, Q, qvs, cnt, add, nbmax, c, n, vs);
  return hipGetLastError();
}

hipError_t launch_scale_rsqrt(double* X, const double* nrm2, int64_t n, int64_t vs, int nvec, hipStream_t st) {
  scale_rsqrt_kernel<<<dim3(vec_grid(n), nvec), 256, 0, st>>>(X, nrm2, n, vs);
  return hipGetLastError();
}

hipError_t launch_lz_step(int* step, hipStream_t st) {
  lz_step_kernel<<<1, 64, 0, st>>>(step);
  return hipGetLastError();
}

hipError_t launch_lz_combine(double* const* basis, const double* c, int ldc, const int* m, double* out, int64_t n,
                             int64_t vs, int nvec, hipStream_t st) {
  lz_combine_kernel<<<dim3(vec_grid(n), nvec), 256, 0, st>>>(basis, c, ldc, m, out, n, vs);
  return hipGetLastError();
}

hipError_t launch_copy2d(const Copy2D* tasks, int ntasks, int64_t max_elems, const double* src, double* dst,
                         const double* scales, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  int64_t blocks = (max_elems + 255) / 256;
  int gx = (int)(blocks < 256 ? (blocks > 0 ? blocks : 1) : 256);
  copy2d_kernel<<<dim3(gx, ntasks), 256, 0, st>>>(tasks, src, dst, scales);
  return hipGetLastError();
}

// Library-factorisation plumbing (sectors handed to rocSOLVER): planar
// complex -> interleaved, and a rows x cols window of a real or interleaved
// complex matrix (ld lda) -> planar (ld ldo, imaginary plane at + plane),
// keeping the whole window (mode 0), its upper (1: i <= j) or lower (2:
// i >= j) trapezoid, zeros elsewhere.
__global__ void pack_z_kernel(const double* __restrict__ re, const double* __restrict__ im, int64_t n,
                              double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    z[2 * i] = re[i];
    z[2 * i + 1] = im[i];
  }
}
__global__ void unpack_mat_kernel(const double* __restrict__ A, int cplx, int64_t lda, int64_t rows, int64_t cols,
                                  int mode, double* __restrict__ out, int64_t ldo, int64_t plane) {
  const int64_t n = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    const bool keep = mode == 0 || (mode == 1 && i <= j) || (mode == 2 && i >= j);
    const int64_t src = i + j * lda;
    out[i + j * ldo] = keep ? (cplx ? A[2 * src] : A[src]) : 0.0;
    if (cplx) out[plane + i + j * ldo] = keep ? A[2 * src + 1] : 0.0;
  }
}

hipError_t launch_pack_z(const double* re, const double* im, int64_t n, double* z, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  pack_z_kernel<<<vec_grid(n), 256, 0, st>>>(re, im, n, z);
  return hipGetLastError();
}

hipError_t launch_unpack_mat(const double* A, int cplx, int64_t lda, int64_t rows, int64_t cols, int mode,
                             double* out, int64_t ldo, int64_t plane, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  unpack_mat_kernel<<<vec_grid(rows * cols), 256, 0, st>>>(A, cplx, lda, rows, cols, mode, out, ldo, plane);
  return hipGetLastError();
}

hipError_t launch_fixphase(double* U, int64_t m, int64_t ldu, double* Vt, int64_t k, int64_t n, int64_t ldv,
                           hipStream_t st) {
  if (k <= 0) return hipSuccess;
  fixphase_kernel<<<(unsigned)k, 256, 0, st>>>(U, m, ldu, Vt, n, ldv);
  return hipGetLastError();
}

// Euclidean norm of every column of A (rows x cols, ld lda), one workgroup per
// column: per-thread sums in a fixed stride, then a fixed-order tree.
__global__ __launch_bounds__(256) void colnorm_kernel(const double* __restrict__ A, int64_t rows, int64_t lda,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  const int64_t j = blockIdx.x;
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < rows; i += 256) {
    const double x = A[i + j * lda];
    a += x * x;
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sqrt(red[0]);
}

hipError_t launch_colnorms(const double* A, int64_t rows, int64_t cols, int64_t lda, double* out, hipStream_t st) {
  if (cols <= 0) return hipSuccess;
  colnorm_kernel<<<(unsigned)cols, 256, 0, st>>>(A, rows, lda, out);
  return hipGetLastError();
}

}  // namespace mpsa
