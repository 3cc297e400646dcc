// Dense FP64 building blocks on gfx950: MFMA GEMM, blocked Cholesky,
// triangular inverse, the Cholesky flush and elementwise/reduction kernels.
//
// These replace the LAPACK/BLAS calls of the reference: potrf
// (src/blocks.jl:97, src/schur_solvers.jl:17), potri (src/blocks.jl:118),
// trmm/trsm/gemm/syrk (src/directions.jl:50-98, src/step.jl:160-162).
#include "common.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

namespace msdp {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

bool& shared_device() {
  static thread_local bool on = false;
  return on;
}

Recorder*& rec_active() {
  static thread_local Recorder* r = nullptr;
  return r;
}

namespace {
struct SyncEvent {
  hipEvent_t ev = nullptr;
  int device = -1;
  ~SyncEvent() {
    if (ev) (void)hipEventDestroy(ev);
  }
};
thread_local bool t_blocking = false;
thread_local SyncEvent t_sync;
}  // namespace

void set_blocking_sync(bool on) { t_blocking = on; }

int stream_sync(hipStream_t s) {
  if (!t_blocking) {
    MSDP_HIP(hipStreamSynchronize(s));
    return 0;
  }
  int dev = 0;
  MSDP_HIP(hipGetDevice(&dev));
  if (!t_sync.ev || t_sync.device != dev) {
    if (t_sync.ev) MSDP_HIP(hipEventDestroy(t_sync.ev));
    t_sync.ev = nullptr;
    MSDP_HIP(hipEventCreateWithFlags(&t_sync.ev, hipEventBlockingSync | hipEventDisableTiming));
    t_sync.device = dev;
  }
  MSDP_HIP(hipEventRecord(t_sync.ev, s));
  MSDP_HIP(hipEventSynchronize(t_sync.ev));
  return 0;
}
const char* last_error() { return g_err.c_str(); }
int fail_hip(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
           what, file, line);
  set_error(buf);
  return -1;
}

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ==========================================================================
// GEMM: 64x64 output tile per workgroup of 4 waves; each wave owns a 32x32
// quadrant made of 2x2 v_mfma_f64_16x16x4 tiles.  K advances 16 at a time
// through LDS.  Operand layouts of the f64 16x16x4 MFMA (wave64):
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   D (16x16): lane l, register r holds D[l/16 + 4*r][l%16]
//              (measured on gfx950: rows 0, 5, 10, 15 are the only rows on
//              which this agrees with the f32 layout 4*(l/16)+r)
// ==========================================================================
constexpr int GT = 64;   // tile edge
constexpr int GK = 16;   // k depth per LDS stage

// K range [kb, ke) of output tile (m0, n0) (edge T) where the triangular
// operands named by `tri` can both be nonzero; kb and ke are multiples of T
// (or ke = K).
__host__ __device__ inline void tri_krange(int tri, int m0, int n0, int T, int K, int& kb,
                                           int& ke) {
  kb = 0;
  ke = K;
  if (tri & TRI_AL) ke = min(ke, m0 + T);
  if (tri & TRI_AU) kb = max(kb, m0);
  if (tri & TRI_BL) kb = max(kb, n0);
  if (tri & TRI_BU) ke = min(ke, n0 + T);
}

// One 64 x 64 output tile (tm, tn) of C = alpha op(A) op(B) + beta C.
template <bool TA, bool TB>
__device__ __attribute__((always_inline)) inline void gemm64_tile(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ B, int64_t ldb, double beta, double* __restrict__ C, int64_t ldc,
    int lower, int tm, int tn, int tri = 0) {
  if (lower && tn > tm) return;
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  v4d acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
  const int m0 = tm * GT, n0 = tn * GT;
  int kb, ke;
  tri_krange(tri, m0, n0, GT, K, kb, ke);
  for (int k0 = kb; k0 < ke; k0 += GK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      int mm, kk;
      if (!TA) { mm = e & 63; kk = e >> 6; } else { kk = e & 15; mm = e >> 4; }
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < M && gk < K) v = TA ? A[gk + (int64_t)gm * lda] : A[gm + (int64_t)gk * lda];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      int nn, kk;
      if (!TB) { kk = e & 15; nn = e >> 4; } else { nn = e & 63; kk = e >> 6; }
      const int gn = n0 + nn, gk = k0 + kk;
      double v = 0.0;
      if (gn < N && gk < K) v = TB ? B[gn + (int64_t)gk * ldb] : B[gk + (int64_t)gn * ldb];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const double a0 = As[kk + lk][wm + li];
      const double a1 = As[kk + lk][wm + 16 + li];
      const double b0 = Bs[kk + lk][wn + li];
      const double b1 = Bs[kk + lk][wn + 16 + li];
      acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
      acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
    }
    __syncthreads();
  }
  const v4d* accs[4] = {&acc00, &acc01, &acc10, &acc11};
  // C's old values first, all loads issued together at clamped (in-bounds)
  // addresses: loads under the store guards were each waited for before the
  // next one was issued
  double old[4][4];
  if (beta != 0.0) {  // uniform
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gn = min(n0 + wn + 16 * (q & 1) + li, N - 1);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = min(m0 + wm + 16 * (q >> 1) + lk + 4 * r, M - 1);
        old[q][r] = C[gm + (int64_t)gn * ldc];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i0 = m0 + wm + 16 * (q >> 1) + lk;
    const int gn = n0 + wn + 16 * (q & 1) + li;
    if (gn >= N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gm = i0 + 4 * r;
      if (gm >= M || (lower && gm < gn)) continue;
      const double v = alpha * (*accs[q])[r];
      C[gm + (int64_t)gn * ldc] = beta == 0.0 ? v : v + beta * old[q][r];
    }
  }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda, int64_t sA,
    const double* __restrict__ B, int64_t ldb, int64_t sB, double beta, double* __restrict__ C,
    int64_t ldc, int64_t sC, int lower, const int* __restrict__ guard, int tri) {
  if (guard && guard[blockIdx.z] != 0) return;  // one guard per batch entry
  const int64_t bz = blockIdx.z;
  gemm64_tile<TA, TB>(M, N, K, alpha, A + bz * sA, lda, B + bz * sB, ldb, beta, C + bz * sC, ldc,
                      lower, blockIdx.x, blockIdx.y, tri);
}

// Variable-size batch of 64-tile products (recorded per-block GEMMs): the
// workgroup's item is found in the prefix of the items' tile counts.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_vb_kernel(const GemmItem* __restrict__ items,
                                                      const int* __restrict__ start, int count) {
  int lo = 0, hi = count - 1;
  const int w = blockIdx.x;
  while (lo < hi) {  // last item with start <= w
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= w) lo = mid; else hi = mid - 1;
  }
  const GemmItem it = items[lo];
  int t = w - start[lo];
  const int ntm = (it.M + GT - 1) / GT;
  const int per = ntm * ((it.N + GT - 1) / GT);
  const int64_t bz = t / per;  // the item's product (batched items)
  t -= (int)bz * per;
  gemm64_tile<TA, TB>(it.M, it.N, it.K, it.alpha, it.A + bz * it.sA, it.lda, it.B + bz * it.sB,
                      it.ldb, it.beta, it.C + bz * it.sC, it.ldc, it.lower, t % ntm, t / ntm,
                      it.tri);
}

// Large-tile GEMM: 128x128 output tile per workgroup of 4 waves, each wave
// a 64x64 quadrant of 4x4 v_mfma_f64_16x16x4 tiles (16 accumulators).  The
// next 16-deep K slab is fetched into registers while the current one is
// multiplied out of LDS (double-buffered, one barrier per slab).
constexpr int BT = 128;
constexpr int BK = 16;

// The K slabs [s0, s1) of one 128 x 128 tile (m0, n0) of op(A) op(B), added
// to acc.  The next 16-deep slab is fetched into registers while the current
// one is multiplied out of LDS (double-buffered, one barrier per slab); the
// last barrier leaves the LDS free for the caller's next tile.
template <bool TA, bool TB>
__device__ __attribute__((always_inline)) inline void gemm128_slabs(
    int M, int N, int K, const double* __restrict__ A, int64_t lda, const double* __restrict__ B,
    int64_t ldb, int m0, int n0, int s0, int s1, double (*As)[BK][BT + 2],
    double (*Bs)[BK][BT + 2], v4d (&acc)[4][4]) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
  double ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + r * 256;
      int mm, kk;
      if (!TA) { mm = e & 127; kk = e >> 7; } else { kk = e & 15; mm = e >> 4; }
      const int gm = m0 + mm, gk = k0 + kk;
      ra[r] = (gm < M && gk < K) ? (TA ? A[gk + (int64_t)gm * lda] : A[gm + (int64_t)gk * lda]) : 0.0;
      int nn;
      if (!TB) { kk = e & 15; nn = e >> 4; } else { nn = e & 127; kk = e >> 7; }
      const int gn = n0 + nn, gk2 = k0 + kk;
      rb[r] = (gn < N && gk2 < K) ? (TB ? B[gn + (int64_t)gk2 * ldb] : B[gk2 + (int64_t)gn * ldb]) : 0.0;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + r * 256;
      int mm, kk;
      if (!TA) { mm = e & 127; kk = e >> 7; } else { kk = e & 15; mm = e >> 4; }
      As[buf][kk][mm] = ra[r];
      int nn;
      if (!TB) { kk = e & 15; nn = e >> 4; } else { nn = e & 127; kk = e >> 7; }
      Bs[buf][kk][nn] = rb[r];
    }
  };
  gload(s0 * BK);
  sstore(0);
  __syncthreads();
  for (int kt = s0; kt < s1; ++kt) {
    const int buf = (kt - s0) & 1;
    if (kt + 1 < s1) gload((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = As[buf][kk + lk][wm + 16 * q + li];
        bv[q] = Bs[buf][kk + lk][wn + 16 * q + li];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (kt + 1 < s1) sstore(buf ^ 1);
    __syncthreads();
  }
}

// C tile (m0, n0) = alpha acc + beta C (lower: the entries on or below the
// diagonal only)
__device__ __attribute__((always_inline)) inline void gemm128_store(
    int M, int N, double alpha, double beta, double* __restrict__ C, int64_t ldc, int m0, int n0,
    int lower, const v4d (&acc)[4][4]) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    // this row group's old C values first, the 16 loads issued together at
    // clamped (in-bounds) addresses: loads under the store guards were each
    // waited for before the next one was issued
    double old[4][4];
    if (beta != 0.0) {  // uniform
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int gn = min(n0 + wn + 16 * b + li, N - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = min(m0 + wm + 16 * a + lk + 4 * r, M - 1);
          old[b][r] = C[gm + (int64_t)gn * ldc];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gn = n0 + wn + 16 * b + li;
      if (gn >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * a + lk + 4 * r;
        if (gm >= M || (lower && gm < gn)) continue;
        const double v = alpha * acc[a][b][r];
        C[gm + (int64_t)gn * ldc] = beta == 0.0 ? v : v + beta * old[b][r];
      }
    }
  }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm128_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda, int64_t sA,
    const double* __restrict__ B, int64_t ldb, int64_t sB, double beta, double* __restrict__ C,
    int64_t ldc, int64_t sC, int lower, const int* __restrict__ guard, int tri) {
  if (guard && guard[blockIdx.z] != 0) return;  // one guard per batch entry
  int tm = blockIdx.x, tn = blockIdx.y;
  if (lower == 2) {  // square lower triangle on a 1-D, XCD-grouped grid
    if (!lower_tile_xcd((M + BT - 1) / BT, tm, tn)) return;
  } else if (lower && tn > tm) {
    return;
  }
  const int64_t bz = blockIdx.z;
  A += bz * sA;
  B += bz * sB;
  C += bz * sC;
  __shared__ double As[2][BK][BT + 2];
  __shared__ double Bs[2][BK][BT + 2];
  v4d acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (v4d){0, 0, 0, 0};
  int kb, ke;
  tri_krange(tri, tm * BT, tn * BT, BT, K, kb, ke);
  if (kb < ke)
    gemm128_slabs<TA, TB>(M, N, K, A, lda, B, ldb, tm * BT, tn * BT, kb / BK, (ke + BK - 1) / BK,
                          As, Bs, acc);
  gemm128_store(M, N, alpha, beta, C, ldc, tm * BT, tn * BT, lower, acc);
}

// --------------------------------------------------------------------------
// Stream-K GEMM for large plain products.  The T = ntm x ntn tiles of 128 x
// 128 are cut into U = T x I slab units (I = K slabs of 16 per tile) and the
// G workgroups (one per CU) take equal contiguous runs of units, so a tile
// grid that is not a multiple of the CU count leaves no partial last wave.
// A workgroup whose run covers a whole tile stores it into C; a run that
// covers part of a tile stores its accumulators into slot 2r (the run's first
// tile) or 2r + 1 (its last) of the partials buffer P, in the accumulator
// layout (register-major, thread-minor: coalesced), and the fixup kernel
// adds the parts of each split tile in K order (deterministic) into C.  Run
// r goes to workgroup r's XCD-contiguous position: consecutive workgroups sit
// on different XCDs, so each XCD's L2 serves a contiguous range of tiles.
// Tiles are visited in groups of SK_GM tile rows, column by column.
// --------------------------------------------------------------------------
constexpr int SK_GM = 8;

struct SkGrid {
  int ntm, ntn, I, G;
  int64_t U;
  int tri = 0, lower = 0;  // the variable-unit partition (gemm_skv_*)
  __host__ __device__ int64_t first(int r) const { return (int64_t)r * U / G; }
  __host__ __device__ int owner(int64_t u) const { return (int)(((u + 1) * G - 1) / U); }
  __host__ __device__ void tile(int t, int& tm, int& tn) const {
    const int per = SK_GM * ntn;
    const int g = t / per, in = t % per;
    const int rows = min(SK_GM, ntm - g * SK_GM);
    tn = in / rows;
    tm = g * SK_GM + in % rows;
  }
};

__device__ inline int sk_run(const SkGrid& sg) {
  const int w = blockIdx.x;
  const int per = sg.G / 8, rem = sg.G % 8, x = w & 7, q = w >> 3;
  // XCD x holds runs [x * per + min(x, rem), ...) when G is not a multiple of 8
  return x * per + min(x, rem) + q;
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_sk_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ B, int64_t ldb, double beta, double* __restrict__ C, int64_t ldc,
    SkGrid sg, double* __restrict__ P) {
  const int r = sk_run(sg);
  if (r >= sg.G) return;
  __shared__ double As[2][BK][BT + 2];
  __shared__ double Bs[2][BK][BT + 2];
  int64_t u = sg.first(r);
  const int64_t u1 = sg.first(r + 1);
  const int tfirst = (int)(u / sg.I);
  while (u < u1) {
    const int t = (int)(u / sg.I);
    const int s0 = (int)(u - (int64_t)t * sg.I);
    const int s1 = (int)min<int64_t>(sg.I, s0 + (u1 - u));
    int tm, tn;
    sg.tile(t, tm, tn);
    v4d acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (v4d){0, 0, 0, 0};
    gemm128_slabs<TA, TB>(M, N, K, A, lda, B, ldb, tm * BT, tn * BT, s0, s1, As, Bs, acc);
    if (s0 == 0 && s1 == sg.I) {
      gemm128_store(M, N, alpha, beta, C, ldc, tm * BT, tn * BT, 0, acc);
    } else {
      double* slot = P + (int64_t)(2 * r + (t == tfirst ? 0 : 1)) * (BT * BT);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int q = 0; q < 4; ++q) slot[((a * 4 + b) * 4 + q) * 256 + threadIdx.x] = acc[a][b][q];
    }
    u += s1 - s0;
  }
}

__global__ __launch_bounds__(256) void gemm_sk_fixup_kernel(int M, int N, double alpha,
                                                            double beta, double* __restrict__ C,
                                                            int64_t ldc, SkGrid sg,
                                                            const double* __restrict__ P) {
  const int t = blockIdx.x;
  const int64_t ua = (int64_t)t * sg.I, ub = ua + sg.I - 1;
  const int ra = sg.owner(ua), rb = sg.owner(ub);
  if (ra == rb) return;  // the whole tile was one run's: already in C
  v4d acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (v4d){0, 0, 0, 0};
  for (int r = ra; r <= rb; ++r) {
    const int tfirst = (int)(sg.first(r) / sg.I);
    const double* slot = P + (int64_t)(2 * r + (t == tfirst ? 0 : 1)) * (BT * BT);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[a][b][q] += slot[((a * 4 + b) * 4 + q) * 256 + threadIdx.x];
  }
  int tm, tn;
  sg.tile(t, tm, tn);
  gemm128_store(M, N, alpha, beta, C, ldc, tm * BT, tn * BT, 0, acc);
}

// --------------------------------------------------------------------------
// Stream-K over a variable number of slabs per tile: triangular operands
// (each tile runs only the K range where both can be nonzero) and lower-
// triangle outputs (tiles above the diagonal run none).  pre[t] is the
// prefix of the tiles' slab counts in the grouped tile order (pre[T] = U,
// from sk_prefix_kernel); run r covers units [r U / G, (r + 1) U / G) as in
// the uniform partition, a unit being one slab of one tile.
// --------------------------------------------------------------------------
__host__ __device__ inline void sk_tile_slabs(const SkGrid& sg, int tm, int tn, int& lo,
                                              int& hi) {
  if (sg.lower && tn > tm) {
    lo = hi = 0;
    return;
  }
  int kb, ke;
  tri_krange(sg.tri, tm * BT, tn * BT, BT, sg.I * BK, kb, ke);
  lo = kb / BK;
  hi = kb < ke ? ke / BK : lo;
}

__global__ __launch_bounds__(256) void sk_prefix_kernel(SkGrid sg, long long* __restrict__ pre) {
  __shared__ long long part[256];
  const int T = sg.ntm * sg.ntn, tid = threadIdx.x;
  const int per = (T + 255) / 256, t0 = tid * per, t1 = min(T, t0 + per);
  long long sum = 0;
  for (int t = t0; t < t1; ++t) {
    int tm, tn, lo, hi;
    sg.tile(t, tm, tn);
    sk_tile_slabs(sg, tm, tn, lo, hi);
    sum += hi - lo;
  }
  part[tid] = sum;
  __syncthreads();
  if (tid == 0) {  // exclusive scan of the 256 chunk sums
    long long run = 0;
    for (int q = 0; q < 256; ++q) {
      const long long v = part[q];
      part[q] = run;
      run += v;
    }
    pre[T] = run;
  }
  __syncthreads();
  long long run = part[tid];
  for (int t = t0; t < t1; ++t) {
    int tm, tn, lo, hi;
    sg.tile(t, tm, tn);
    sk_tile_slabs(sg, tm, tn, lo, hi);
    pre[t] = run;
    run += hi - lo;
  }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_skv_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ B, int64_t ldb, double beta, double* __restrict__ C, int64_t ldc,
    SkGrid sg, const long long* __restrict__ pre, double* __restrict__ P) {
  const int r = sk_run(sg);
  if (r >= sg.G) return;
  const int T = sg.ntm * sg.ntn;
  const long long U = pre[T];
  long long u = (long long)r * U / sg.G;
  const long long u1 = (long long)(r + 1) * U / sg.G;
  if (u >= u1) return;
  const long long ufirst = u;
  int t = 0, hi_t = T - 1;  // the last tile with pre[t] <= u: it holds unit u
  while (t < hi_t) {
    const int mid = (t + hi_t + 1) >> 1;
    if (pre[mid] <= u) t = mid; else hi_t = mid - 1;
  }
  __shared__ double As[2][BK][BT + 2];
  __shared__ double Bs[2][BK][BT + 2];
  while (u < u1) {
    while (pre[t + 1] <= u) ++t;  // past tiles without slabs
    int tm, tn, slo, shi;
    sg.tile(t, tm, tn);
    sk_tile_slabs(sg, tm, tn, slo, shi);
    const long long pt = pre[t];
    const int s0 = slo + (int)(u - pt);
    const int s1 = slo + (int)min<long long>(shi - slo, u1 - pt);
    v4d acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (v4d){0, 0, 0, 0};
    gemm128_slabs<TA, TB>(M, N, K, A, lda, B, ldb, tm * BT, tn * BT, s0, s1, As, Bs, acc);
    if (s0 == slo && s1 == shi) {
      gemm128_store(M, N, alpha, beta, C, ldc, tm * BT, tn * BT, sg.lower, acc);
    } else {
      double* slot = P + (int64_t)(2 * r + (pt <= ufirst ? 0 : 1)) * (BT * BT);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int q = 0; q < 4; ++q) slot[((a * 4 + b) * 4 + q) * 256 + threadIdx.x] = acc[a][b][q];
    }
    u = pt + (s1 - slo);
  }
}

// One workgroup per tile: the parts of a split tile added in K order into C;
// a tile without slabs gets beta C (none above the diagonal of a lower
// product).
__global__ __launch_bounds__(256) void gemm_skv_fixup_kernel(int M, int N, double alpha,
                                                             double beta, double* __restrict__ C,
                                                             int64_t ldc, SkGrid sg,
                                                             const long long* __restrict__ pre,
                                                             const double* __restrict__ P) {
  const int t = blockIdx.x;
  int tm, tn;
  sg.tile(t, tm, tn);
  const long long pa = pre[t], pb = pre[t + 1];
  v4d acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (v4d){0, 0, 0, 0};
  if (pa == pb) {
    if (sg.lower && tn > tm) return;
    gemm128_store(M, N, alpha, beta, C, ldc, tm * BT, tn * BT, sg.lower, acc);
    return;
  }
  const long long U = pre[sg.ntm * sg.ntn];
  const int ra = (int)(((pa + 1) * sg.G - 1) / U), rb = (int)((pb * sg.G - 1) / U);
  if (ra == rb) return;  // the whole tile was one run's: already in C
  for (int r = ra; r <= rb; ++r) {
    const long long fr = (long long)r * U / sg.G;
    const double* slot = P + (int64_t)(2 * r + (fr >= pa ? 0 : 1)) * (BT * BT);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[a][b][q] += slot[((a * 4 + b) * 4 + q) * 256 + threadIdx.x];
  }
  gemm128_store(M, N, alpha, beta, C, ldc, tm * BT, tn * BT, sg.lower, acc);
}

// Split-K GEMM for skinny outputs with a long K (few output tiles): each
// blockIdx.z computes the partial product over one K chunk into P; a second
// kernel sums the chunks (fixed order, deterministic) into C.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_splitk_kernel(int M, int N, int K, int kc,
                                                          const double* __restrict__ A,
                                                          int64_t lda,
                                                          const double* __restrict__ B,
                                                          int64_t ldb, double* __restrict__ P,
                                                          int lower, int tri) {
  const int tm = blockIdx.x, tn = blockIdx.y, part = blockIdx.z;
  if (lower && tn > tm) return;  // wholly above the diagonal: never read back
  int kb, ke;
  tri_krange(tri, tm * GT, tn * GT, GT, K, kb, ke);
  const int kbeg = max(part * kc, kb), kend = min(min(K, part * kc + kc), ke);
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  v4d acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
  const int m0 = tm * GT, n0 = tn * GT;
  for (int k0 = kbeg; k0 < kend; k0 += GK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      int mm, kk;
      if (!TA) { mm = e & 63; kk = e >> 6; } else { kk = e & 15; mm = e >> 4; }
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < M && gk < kend) v = TA ? A[gk + (int64_t)gm * lda] : A[gm + (int64_t)gk * lda];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      int nn, kk;
      if (!TB) { kk = e & 15; nn = e >> 4; } else { nn = e & 63; kk = e >> 6; }
      const int gn = n0 + nn, gk = k0 + kk;
      double v = 0.0;
      if (gn < N && gk < kend) v = TB ? B[gn + (int64_t)gk * ldb] : B[gk + (int64_t)gn * ldb];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const double a0 = As[kk + lk][wm + li];
      const double a1 = As[kk + lk][wm + 16 + li];
      const double b0 = Bs[kk + lk][wn + li];
      const double b1 = Bs[kk + lk][wn + 16 + li];
      acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
      acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
    }
    __syncthreads();
  }
  double* Pp = P + (int64_t)part * M * N;
  const v4d* accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i0 = m0 + wm + 16 * (q >> 1) + lk;
    const int gn = n0 + wn + 16 * (q & 1) + li;
    if (gn >= N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gm = i0 + 4 * r;
      if (gm < M) Pp[gm + (int64_t)gn * M] = (*accs[q])[r];
    }
  }
}

__global__ void splitk_reduce_kernel(int M, int N, int parts, const double* P, double alpha,
                                     double beta, double* C, int64_t ldc, int lower) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)M * N) return;
  const int i = (int)(e % M), j = (int)(e / M);
  if (lower && i < j) return;
  // four interleaved sums keep four loads in flight (deep splits are
  // latency-bound otherwise), combined in a fixed order
  const int64_t mn = (int64_t)M * N;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int q = 0;
  for (; q + 4 <= parts; q += 4) {
    a0 += P[e + (int64_t)q * mn];
    a1 += P[e + (int64_t)(q + 1) * mn];
    a2 += P[e + (int64_t)(q + 2) * mn];
    a3 += P[e + (int64_t)(q + 3) * mn];
  }
  if (q < parts) a0 += P[e + (int64_t)q * mn];
  if (q + 1 < parts) a1 += P[e + (int64_t)(q + 1) * mn];
  if (q + 2 < parts) a2 += P[e + (int64_t)(q + 2) * mn];
  const double v = (a0 + a1) + (a2 + a3);
  double* c = C + i + (int64_t)j * ldc;
  *c = beta == 0.0 ? alpha * v : alpha * v + beta * (*c);
}

// --------------------------------------------------------------------------
// Kernel-selection parameters.  Defaults are compiled in; an offline sweep
// on the target (madsdp_amd/autotune.py) writes a table that the host layer
// applies once, before any solver runs, through msdp_set_tuning.  Within a
// process they do not change, so results stay deterministic.
// --------------------------------------------------------------------------
namespace {
struct TuneParam {
  const char* name;
  std::atomic<int> value;
  int lo, hi;
};
TuneParam g_tune[] = {
    {"gemm_wide_min_tiles", {256}, 1, 1 << 20},   // 128-tiles used when this many run ...
    {"gemm_wide_deep_tiles", {96}, 1, 1 << 20},   // ... or this many when K >= 1024
    {"splitk_workgroups", {1024}, 64, 1 << 16},   // split-K: workgroups aimed at
    {"splitk_min_k", {256}, 64, 1 << 20},         // split-K only for K at least this
    {"streamk_occupancy", {2}, 1, 2},              // stream-K workgroups per CU
    {"streamk_min_slabs", {8}, 1, 1024},           // stream-K: slabs per workgroup at least
    {"streamk_max_waves", {4}, 1, 1 << 20},        // stream-K: tiles below this many waves
};
enum { T_WIDE_MIN = 0, T_WIDE_DEEP, T_SPLITK_WGS, T_SPLITK_MIN_K, T_SK_OCC, T_SK_MIN_UNITS,
       T_SK_MAX_WAVES };
inline int tune(int k) { return g_tune[k].value.load(std::memory_order_relaxed); }
}  // namespace

int set_tuning(const char* name, double value) {
  for (auto& p : g_tune) {
    if (std::strcmp(p.name, name) != 0) continue;
    if (!(value >= p.lo && value <= p.hi) || value != std::floor(value)) {
      set_error(std::string("tuning value out of range for ") + name);
      return -2;
    }
    p.value.store((int)value, std::memory_order_relaxed);
    return 0;
  }
  set_error(std::string("no tuning parameter ") + name);
  return -2;
}

int get_tuning(const char* name, double* value) {
  for (auto& p : g_tune)
    if (std::strcmp(p.name, name) == 0) {
      *value = p.value.load(std::memory_order_relaxed);
      return 0;
    }
  set_error(std::string("no tuning parameter ") + name);
  return -2;
}

constexpr int SPLITK_MIN_CHUNK = 64;

// Compute units of the current device (the stream-K grid: one workgroup per
// CU, so the partition -- and with it the summation order -- is fixed per
// device).
static int device_cus() {
  static std::mutex mu;
  static std::vector<int> cus;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)cus.size() <= dev) cus.resize(dev + 1, 0);
  if (cus[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

// Plain products (no batch, no triangle, no guard) whose 128-tile grid would
// leave a partial last wave go through the stream-K kernel: at least
// streamk_min_slabs slabs per workgroup, more tiles than CUs, and fewer than
// streamk_max_waves waves of them (with many waves the last one's idle
// share is small and the fixup's traffic is not).

static bool stream_k_shape(int M, int N, int K, int batch, bool lower, const int* guard, int G,
                           SkGrid& sg) {
  if (batch != 1 || lower || guard || M < BT || N < BT) return false;
  sg.ntm = cdiv(M, BT);
  sg.ntn = cdiv(N, BT);
  sg.I = cdiv(K, BK);
  sg.G = G;
  sg.U = (int64_t)sg.ntm * sg.ntn * sg.I;
  const int64_t T = (int64_t)sg.ntm * sg.ntn;
  return T > G && T % G != 0 && sg.U >= (int64_t)tune(T_SK_MIN_UNITS) * G &&
         T < (int64_t)tune(T_SK_MAX_WAVES) * G;
}

// Variable-unit stream-K (triangular operands or a lower-triangle output):
// the tiles' slab counts decide; at least streamk_min_slabs slabs per
// workgroup, at least half as many tiles with slabs as workgroups (fewer
// would be split many ways), fewer than streamk_max_waves waves of tiles.
static bool stream_kv_shape(int M, int N, int K, int batch, bool lower, int tri, const int* guard,
                            int G, SkGrid& sg) {
  if (batch != 1 || guard || M < BT || N < BT || (!lower && !tri)) return false;
  sg.ntm = cdiv(M, BT);
  sg.ntn = cdiv(N, BT);
  sg.I = cdiv(K, BK);
  sg.G = G;
  sg.tri = tri;
  sg.lower = lower ? 1 : 0;
  const int64_t T = (int64_t)sg.ntm * sg.ntn;
  if (T >= (int64_t)tune(T_SK_MAX_WAVES) * G) return false;
  int64_t U = 0, live = 0;
  for (int tm = 0; tm < sg.ntm; ++tm)
    for (int tn = 0; tn < sg.ntn; ++tn) {
      int lo, hi;
      sk_tile_slabs(sg, tm, tn, lo, hi);
      U += hi - lo;
      live += hi > lo;
    }
  sg.U = U;
  return U >= (int64_t)tune(T_SK_MIN_UNITS) * G && 2 * live >= G;
}

// stream-K partials at the largest occupancy (two of the 256-thread,
// 67 KB-LDS workgroups fit per CU), and the variable partition's tile prefix
size_t gemm_sk_partials_bytes() {
  const size_t G = (size_t)2 * device_cus();
  return sizeof(double) * (2 * G * BT * BT + (size_t)tune(T_SK_MAX_WAVES) * G + 1);
}

// Partials of the stream-K and split-K paths: the caller's workspace `ws` of
// `ws_elems` doubles, or (ws == nullptr) a block of the stream-ordered pool.
static int partials(hipStream_t s, size_t elems, double* ws, size_t ws_elems, double** out) {
  if (ws) {
    if (elems > ws_elems) {
      char buf[160];
      snprintf(buf, sizeof buf, "GEMM partials need %zu bytes, the scratch holds %zu",
               elems * sizeof(double), ws_elems * sizeof(double));
      set_error(buf);
      return -2;
    }
    *out = ws;
    return 0;
  }
  void* P = nullptr;
  MSDP_HIP(hipMallocAsync(&P, sizeof(double) * elems, s));
  *out = static_cast<double*>(P);
  return 0;
}

static int gemm_stream_k(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                         const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                         double* C, int64_t ldc, const SkGrid& sg, double* ws, size_t ws_elems) {
  double* Pd = nullptr;
  MSDP_TRY(partials(s, 2 * (size_t)sg.G * BT * BT, ws, ws_elems, &Pd));
  const dim3 g(sg.G);
  if (!ta && !tb) gemm_sk_f64_kernel<false, false><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, Pd);
  else if (ta && !tb) gemm_sk_f64_kernel<true, false><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, Pd);
  else if (!ta && tb) gemm_sk_f64_kernel<false, true><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, Pd);
  else gemm_sk_f64_kernel<true, true><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, Pd);
  gemm_sk_fixup_kernel<<<sg.ntm * sg.ntn, 256, 0, s>>>(M, N, alpha, beta, C, ldc, sg, Pd);
  MSDP_HIP(hipGetLastError());
  if (!ws) MSDP_HIP(hipFreeAsync(Pd, s));
  return 0;
}

static int gemm_stream_kv(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                          const double* A, int64_t lda, const double* B, int64_t ldb,
                          double beta, double* C, int64_t ldc, const SkGrid& sg, double* ws,
                          size_t ws_elems) {
  const int T = sg.ntm * sg.ntn;
  const size_t slots = 2 * (size_t)sg.G * BT * BT;
  double* Pd = nullptr;
  MSDP_TRY(partials(s, slots + T + 1, ws, ws_elems, &Pd));
  long long* pre = reinterpret_cast<long long*>(Pd + slots);
  sk_prefix_kernel<<<1, 256, 0, s>>>(sg, pre);
  const dim3 g(sg.G);
  if (!ta && !tb) gemm_skv_f64_kernel<false, false><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, pre, Pd);
  else if (ta && !tb) gemm_skv_f64_kernel<true, false><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, pre, Pd);
  else if (!ta && tb) gemm_skv_f64_kernel<false, true><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, pre, Pd);
  else gemm_skv_f64_kernel<true, true><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, pre, Pd);
  gemm_skv_fixup_kernel<<<T, 256, 0, s>>>(M, N, alpha, beta, C, ldc, sg, pre, Pd);
  MSDP_HIP(hipGetLastError());
  if (!ws) MSDP_HIP(hipFreeAsync(Pd, s));
  return 0;
}

// Multiply-adds of a product whose triangular operands (tri) and lower
// output skip work, counted over 64 x 64 tiles (the profiler's flop count).
static double product_flops(int M, int N, int K, bool lower, int tri) {
  if (!tri) return 2.0 * M * N * (double)K * (lower ? 0.5 : 1.0);
  double f = 0.0;
  for (int tm = 0; tm < cdiv(M, GT); ++tm)
    for (int tn = 0; tn < cdiv(N, GT); ++tn) {
      if (lower && tn > tm) continue;
      int kb, ke;
      tri_krange(tri, tm * GT, tn * GT, GT, K, kb, ke);
      if (ke > kb)
        f += 2.0 * std::min(GT, M - tm * GT) * std::min(GT, N - tn * GT) * (double)(ke - kb);
    }
  return f;
}

static int gemm_guarded(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc, int batch, int64_t sA, int64_t sB, int64_t sC,
                        bool lower, const int* guard, double* ws = nullptr,
                        size_t ws_elems = 0, int tri = 0) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  if (K <= 0) {
    // C = beta C
    if (beta == 1.0) return 0;
    for (int b = 0; b < batch; ++b) {
      for (int j = 0; j < N; ++j) {
        // rare path: scale columns by beta via axpby
        MSDP_TRY(axpby(s, M, beta, C + b * sC + (int64_t)j * ldc, 0.0, C + b * sC + (int64_t)j * ldc,
                       C + b * sC + (int64_t)j * ldc));
      }
    }
    return 0;
  }
  const bool prof = prof_active(PROF_GEMM, s);
  SkGrid sg;
  // a device shared with other solves: one launch per product
  const bool whole_chip = !shared_device();
  if (whole_chip && !tri && stream_k_shape(M, N, K, batch, lower, guard, tune(T_SK_OCC) * device_cus(), sg)) {
    if (prof) prof_before(s);
    MSDP_TRY(gemm_stream_k(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, ws,
                           ws_elems));
    if (prof) prof_after(s, 2.0 * M * N * (double)K);
    return 0;
  }
  if (whole_chip &&
      stream_kv_shape(M, N, K, batch, lower, tri, guard, tune(T_SK_OCC) * device_cus(), sg)) {
    if (prof) prof_before(s);
    MSDP_TRY(gemm_stream_kv(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, sg, ws,
                            ws_elems));
    if (prof) prof_after(s, product_flops(M, N, K, lower, tri));
    return 0;
  }
  const int64_t tiles64 = (int64_t)cdiv(M, GT) * cdiv(N, GT) * batch;
  if (whole_chip && !guard && batch == 1 && tiles64 < 128 && K >= tune(T_SPLITK_MIN_K)) {
    // split K so that about splitk_workgroups workgroups run, chunks of at
    // least SPLITK_MIN_CHUNK
    int parts = (int)std::min<int64_t>(std::max<int64_t>(2, tune(T_SPLITK_WGS) / tiles64),
                                       K / SPLITK_MIN_CHUNK);
    const int kc = ((K + parts - 1) / parts + GK - 1) / GK * GK;
    parts = (K + kc - 1) / kc;
    double* Pd = nullptr;
    MSDP_TRY(partials(s, (size_t)M * N * parts, ws, ws_elems, &Pd));
    if (prof) prof_before(s);
    dim3 g3(cdiv(M, GT), cdiv(N, GT), parts);
    if (!ta && !tb) gemm_splitk_kernel<false, false><<<g3, 256, 0, s>>>(M, N, K, kc, A, lda, B, ldb, Pd, lower, tri);
    else if (ta && !tb) gemm_splitk_kernel<true, false><<<g3, 256, 0, s>>>(M, N, K, kc, A, lda, B, ldb, Pd, lower, tri);
    else if (!ta && tb) gemm_splitk_kernel<false, true><<<g3, 256, 0, s>>>(M, N, K, kc, A, lda, B, ldb, Pd, lower, tri);
    else gemm_splitk_kernel<true, true><<<g3, 256, 0, s>>>(M, N, K, kc, A, lda, B, ldb, Pd, lower, tri);
    splitk_reduce_kernel<<<cdiv((int64_t)M * N, 256), 256, 0, s>>>(M, N, parts, Pd, alpha, beta, C, ldc,
                                                                     lower);
    if (prof) prof_after(s, product_flops(M, N, K, lower, tri));
    MSDP_HIP(hipGetLastError());
    if (!ws) MSDP_HIP(hipFreeAsync(Pd, s));
    return 0;
  }
  if (prof) prof_before(s);
  // 128 x 128 tiles only when they still fill the chip (lower: about half
  // the tiles run), or at least a third of it for deep products; otherwise
  // the 64 x 64 kernel's four times as many workgroups finish sooner
  const int64_t tiles128 = (int64_t)cdiv(M, BT) * cdiv(N, BT) * batch;
  const int64_t run128 = lower ? tiles128 / 2 : tiles128;
  if (M >= BT && N >= BT &&
      (run128 >= tune(T_WIDE_MIN) || (K >= 1024 && run128 >= tune(T_WIDE_DEEP)))) {
    dim3 g2(cdiv(M, BT), cdiv(N, BT), batch);
    int lmode = lower ? 1 : 0;
    if (lower && batch == 1 && M == N) {
      g2 = dim3(lower_tile_blocks(cdiv(M, BT)), 1, 1);
      lmode = 2;
    }
    if (!ta && !tb)
      gemm128_f64_kernel<false, false><<<g2, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode, guard, tri);
    else if (ta && !tb)
      gemm128_f64_kernel<true, false><<<g2, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode, guard, tri);
    else if (!ta && tb)
      gemm128_f64_kernel<false, true><<<g2, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode, guard, tri);
    else
      gemm128_f64_kernel<true, true><<<g2, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode, guard, tri);
    if (prof) prof_after(s, product_flops(M, N, K, lower, tri) * batch);
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  dim3 grid(cdiv(M, GT), cdiv(N, GT), batch);
  if (!ta && !tb)
    gemm_f64_kernel<false, false><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard, tri);
  else if (ta && !tb)
    gemm_f64_kernel<true, false><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard, tri);
  else if (!ta && tb)
    gemm_f64_kernel<false, true><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard, tri);
  else
    gemm_f64_kernel<true, true><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard, tri);
  if (prof) prof_after(s, product_flops(M, N, K, lower, tri) * batch);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha, const double* A,
         int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
         int batch, int64_t sA, int64_t sB, int64_t sC, bool lower, int tri) {
  if (rec_active()) {
    RecOp o;
    // batched only where a lone launch takes the 64-tile kernel as well
    if (batch == 1 && M > 0 && N > 0 && K > 0 && std::max(M, std::max(N, K)) < 512) {
      o.kind = ta ? (tb ? REC_GEMM_TT : REC_GEMM_TN) : (tb ? REC_GEMM_NT : REC_GEMM_NN);
      o.g = GemmItem{M, N, K, lower ? 1 : 0, tri, alpha, beta, A, B, C, lda, ldb, ldc};
    }
    o.run = [=] {
      return gemm_guarded(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB,
                          sC, lower, nullptr, nullptr, 0, tri);
    };
    return rec_add(std::move(o));
  }
  return gemm_guarded(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB, sC,
                      lower, nullptr, nullptr, 0, tri);
}

// --------------------------------------------------------------------------
// GEMM that allocates nothing: the partials of the stream-K and split-K
// paths live in one file-scope device array (NOALLOC_SCRATCH_BYTES per
// device, held until the process ends).  Its address on each device is
// looked up once, with that device current, and kept; host code never
// passes the array's own name (a host-side stand-in) to a kernel.  Every
// call on a device shares that device's array: calls that overlap on one
// device race.  Nothing in it needs to start at zero (the kernels write
// every partial they read).
// --------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

static int noalloc_scratch(double** out) {
  static std::mutex mu;
  static std::vector<double*> ptrs;
  int dev = 0;
  MSDP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    ptrs[dev] = static_cast<double*>(p);
  }
  *out = ptrs[dev];
  return 0;
}

size_t gemm_noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

int gemm_noalloc(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc) {
  double* ws = nullptr;
  MSDP_TRY(noalloc_scratch(&ws));
  return gemm_guarded(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, 0, 0, 0, false,
                      nullptr, ws, NOALLOC_SCRATCH_BYTES / sizeof(double));
}

// ==========================================================================
// Cholesky
// ==========================================================================
// --------------------------------------------------------------------------
// One-wavefront 64 x 64 kernels.  Lane i keeps row i (factorisation) or
// column i (inverse) of the block in 64 registers; the loops over the block
// are unrolled so the register arrays are indexed statically, the column
// being eliminated is published in LDS, and no workgroup barrier is needed
// (a wavefront's LDS accesses complete in program order).  The dependent
// chain per column step is one square root / reciprocal and one FMA.
// --------------------------------------------------------------------------
constexpr int W64_LD = 65;  // LDS leading dimension of a 64 x 64 block

// Right-looking Cholesky of the n x n (n <= 64) block whose row `lane`
// (lower part; zeros right of the diagonal and at rows >= n) is in r.
// Column j: the pivot's square root, the column divided by it, then
// r[c] -= L[lane][j] L[c][j] for c > j, entries updated in increasing j as
// in LAPACK's right-looking potrf.  Column j of L is left in
// Lc[j * W64_LD + i] (zeros above the diagonal and at rows >= n).  The
// column steps are template instances (compile-time j), so every register
// index is static; the published column is read in chunks of 32 whose loads
// are issued together (at most two LDS round trips per step; the reads are
// broadcasts, every lane the same address).
template <int J>
__device__ __attribute__((always_inline)) inline void chol64_step_one(double (&r)[64], int n,
                                                                      double* Lc, int& fail) {
  const int lane = threadIdx.x & 63;
  if (J < n && fail == 0) {  // uniform
    const double d = readlane_f64(r[J], J);
    if (!(d > 0.0) || !isfinite(d)) {
      fail = J + 1;
    } else {
      const double ljj = sqrt(d);
      const double v = lane == J ? ljj : (lane > J ? r[J] / ljj : 0.0);
      r[J] = v;
      Lc[J * W64_LD + lane] = v;
      wave_lds_sync();
#pragma unroll
      for (int g = ((J + 1) / 32) * 32; g < 64; g += 32) {
        double lc[32];
#pragma unroll
        for (int q = 0; q < 32; ++q)
          if (g + q > J) lc[q] = Lc[J * W64_LD + g + q];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 32; ++q)
          if (g + q > J) r[g + q] -= v * lc[q];
      }
    }
  }
}
template <int J>
__device__ __attribute__((always_inline)) inline void chol64_step(double (&r)[64], int n,
                                                                  double* Lc, int& fail) {
  chol64_step_one<J>(r, n, Lc, fail);
  if constexpr (J + 1 < 64) chol64_step<J + 1>(r, n, Lc, fail);
}

__device__ __attribute__((always_inline)) inline int chol64_rows(double (&r)[64], int n,
                                                                 double* Lc) {
  int fail = 0;
  chol64_step<0>(r, n, Lc, fail);
  return fail;
}

// Column `lane` of L^-1 for the n x n lower factor in Lc (column-major, ld
// W64_LD), right-looking: x starts at e_lane; step k scales x_k by 1 / L_kk
// and subtracts L_ik x_k from every x_i below it.
template <int K>
__device__ __attribute__((always_inline)) inline void trinv64_step(const double* Lc, int n,
                                                                   double (&x)[64]) {
  if (K < n) {
    const double rd = 1.0 / Lc[K * W64_LD + K];
    x[K] *= rd;
#pragma unroll
    for (int g = ((K + 1) / 16) * 16; g < 64; g += 16) {
      double lc[16];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (g + q > K) lc[q] = Lc[K * W64_LD + g + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (g + q > K) x[g + q] -= lc[q] * x[K];
    }
  }
  if constexpr (K + 1 < 64) trinv64_step<K + 1>(Lc, n, x);
}

__device__ __attribute__((always_inline)) inline void trinv64_cols(const double* Lc, int n,
                                                                   double (&x)[64]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 64; ++i) x[i] = i == lane ? 1.0 : 0.0;
  trinv64_step<0>(Lc, n, x);
}

// Factor one diagonal block (n <= 64) in place (lower triangle; the upper is
// not touched) with one wavefront.  With `dinv`, also L^-1 (64 x 64,
// column-major, ld 64, zero outside the n x n lower triangle).
// info[b] = info_base + first failing column.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void chol_diag_kernel(int n, double* A, int64_t lda,
                                                       int64_t stride, int* info, int info_base,
                                                       int per_batch_info, double* dinv) {
  int* inf = info + (per_batch_info ? blockIdx.x : 0);
  if (*inf != 0) return;  // an earlier block of this matrix failed
  A += (int64_t)blockIdx.x * stride;
  __shared__ double Lc[64 * W64_LD];
  const int lane = threadIdx.x;
  double r[64];
#pragma unroll
  for (int c = 0; c < 64; ++c)
    r[c] = (lane < n && c <= lane) ? A[lane + (int64_t)c * lda] : 0.0;
  const int fail = chol64_rows(r, n, Lc);
  if (fail) {
    if (lane == 0) atomicCAS(inf, 0, info_base + fail);
    return;
  }
#pragma unroll
  for (int c = 0; c < 64; ++c)
    if (lane < n && c <= lane) A[lane + (int64_t)c * lda] = r[c];
  if (!dinv) return;
  wave_lds_sync();
  double x[64];
  trinv64_cols(Lc, n, x);
  wave_lds_sync();
  // transpose through LDS so that the stores of dinv are coalesced
#pragma unroll
  for (int i = 0; i < 64; ++i) Lc[i * W64_LD + lane] = x[i];
  wave_lds_sync();
#pragma unroll 8
  for (int c = 0; c < 64; ++c)
    dinv[lane + 64 * c] = (lane < n && c < n && c <= lane) ? Lc[lane * W64_LD + c] : 0.0;
}

__global__ void diag_inv_kernel(int n, const double* L, int64_t ldl, double* Di, const int* info);

// A21 <- A21 L11^-T in place by substitution (the rows are independent):
// one wavefront per 64 rows, lane = row, the row in registers; for k = 0..63
// x_k *= 1 / L_kk, then x_j -= x_k L_jk for j > k (BLAS dtrsm, right side,
// which also scales by the reciprocal).  Column k of L11 is read from LDS in
// chunks of 32 whose loads are issued together.  The step guard (always true
// for the 64-wide panels) ends a basic block per step: without it the
// scheduler hoists the LDS reads of later steps and the row spills.
template <int K>
__device__ __attribute__((always_inline)) inline void chol_panel_step_one(const double* Lt,
                                                                          double rdk, int nb,
                                                                          double (&x)[64]) {
  if (K >= nb) return;
  x[K] *= rdk;
#pragma unroll
  for (int g = ((K + 1) / 32) * 32; g < 64; g += 32) {
    double lc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q)
      if (g + q > K) lc[q] = Lt[K * W64_LD + g + q];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 32; ++q)
      if (g + q > K) x[g + q] -= x[K] * lc[q];
  }
}
template <int K>
__device__ __attribute__((always_inline)) inline void chol_panel_step(const double* Lt,
                                                                      const double* rd, int nb,
                                                                      double (&x)[64]) {
  chol_panel_step_one<K>(Lt, rd[K], nb, x);
  if constexpr (K + 1 < 64) chol_panel_step<K + 1>(Lt, rd, nb, x);
}

// The factorisation of a 64 x 64 diagonal block (wavefront 0, rows in lanes)
// and the substitution of 64 rows below it (wavefront 1) as one pipeline:
// step J factors column J while the substitution applies column J - 1, one
// workgroup barrier between steps.  Each wavefront does exactly the
// arithmetic of chol64_step / chol_panel_step (the reciprocal 1 / L_KK is
// the same quotient), so the bits are those of the two separate phases.
template <int J>
__device__ __attribute__((always_inline)) inline void chol_pipe_step(bool factor, double (&v)[64],
                                                                     int jb, double* Lt,
                                                                     int& fail) {
  if (factor) {
    chol64_step_one<J>(v, jb, Lt, fail);
  } else if constexpr (J > 0) {
    chol_panel_step_one<J - 1>(Lt, 1.0 / Lt[(J - 1) * (W64_LD + 1)], jb, v);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  if constexpr (J + 1 < 64) {
    chol_pipe_step<J + 1>(factor, v, jb, Lt, fail);
  } else if (!factor) {
    chol_panel_step_one<63>(Lt, 1.0 / Lt[63 * (W64_LD + 1)], jb, v);
  }
}


// A full 64 x 64 diagonal block and the rows below it in one launch: every
// workgroup factors the diagonal block itself (chol_diag_kernel's arithmetic,
// so the same bits in every workgroup, the factor straight into the LDS
// layout of the substitution; wavefront 0) and solves its 64 rows by
// substitution (x_k /= L_kk, then x_j -= x_k L_jk for j > k; wavefront 1),
// the two pipelined one column apart (chol_pipe_step); workgroup 0 records a
// failure and writes the block's factor to `dout` (64 x 64, ld 64) -- not
// over A11, which the other workgroups may still be reading -- and
// diag_store_kernel moves the factors into place after the last step.  One
// dependent launch per column step fewer than a diagonal launch followed by
// a panel launch.
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1))) void
chol_diag_panel_kernel(int rows, double* A11, int64_t lda, int64_t sA, int* info, int info_base,
                       int jb, double* dout, int64_t sW) {
  // jb (64) arrives at run time: the per-step guards then end a basic block
  // per step, which keeps the scheduler from hoisting later steps' LDS reads
  // (the row would spill)
  A11 += blockIdx.y * sA;
  dout += blockIdx.y * sW;
  int* inf = info + blockIdx.y;
  __shared__ double Lt[64 * W64_LD];  // L11[j][k] at Lt[k * W64_LD + j]
  __shared__ int s_skip, s_fail;
  const int lane = threadIdx.x & 63;
  // wavefront 0 factors, wavefront 1 substitutes (a wave-uniform, scalar branch)
  const bool factor = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0;
  // an earlier step's failure ends the factorisation: one read, shared, so
  // that both wavefronts take the same branch
  if (threadIdx.x == 0) {
    s_skip = *inf;
    s_fail = 0;
  }
  __syncthreads();
  if (s_skip != 0) return;
  double* A21 = A11 + 64;
  const int r = blockIdx.x * 64 + lane;
  double v[64];
  if (factor) {
#pragma unroll
    for (int c = 0; c < 64; ++c) v[c] = c <= lane ? A11[lane + (int64_t)c * lda] : 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < 64; ++k) v[k] = r < rows ? A21[r + (int64_t)k * lda] : 0.0;
  }
  int fail = 0;
  chol_pipe_step<0>(factor, v, jb, Lt, fail);
  if (factor) {
    if (fail) {
      if (lane == 0) s_fail = fail;
    } else if (blockIdx.x == 0) {
#pragma unroll
      for (int c = 0; c < 64; ++c) dout[lane + 64 * c] = c <= lane ? v[c] : 0.0;
    }
  }
  __syncthreads();
  if (s_fail != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicCAS(inf, 0, info_base + s_fail);
    return;
  }
  if (factor || r >= rows) return;
#pragma unroll
  for (int k = 0; k < 64; ++k) A21[r + (int64_t)k * lda] = v[k];
}

// The diagonal blocks 0..nblk-1 that chol_diag_panel_kernel factored into
// W (block b at W + 4096 b, ld 64) copied into A's lower triangle.
__global__ __launch_bounds__(256) void diag_store_kernel(double* A, int64_t lda, int64_t sA,
                                                         const double* W, int64_t sW,
                                                         const int* info) {
  const int b = blockIdx.x;
  A += blockIdx.y * sA;
  W += blockIdx.y * sW + (int64_t)b * 4096;
  if (info[blockIdx.y] != 0) return;
  double* D = A + (int64_t)64 * b * (lda + 1);
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e & 63, c = e >> 6;
    if (r >= c) D[r + (int64_t)c * lda] = W[e];
  }
}

size_t potrf_work_elems(int n) { return (size_t)cdiv(std::max(n, 1), 64) * 4096; }

// Right-looking blocked Cholesky in 256-column panels for large n: inside a
// panel the 64-column steps (diagonal block, rows below by substitution)
// update only the panel's own later columns; the rest of the matrix takes one
// rank-256 update per panel.  A rank-64 update per step re-reads and rewrites
// the whole trailing triangle four times as often with a quarter of the
// depth (measured 0.23 of the FP64 MFMA peak at n = 8000).
constexpr int POTRF_NB = 256;
constexpr int POTRF_BLOCKED_MIN = 4096;

static int potrf_blocked(hipStream_t s, int n, double* A, int64_t lda, int64_t sA, int count,
                         int* info, double* work, int64_t sW) {
  for (int J0 = 0; J0 < n; J0 += POTRF_NB) {
    const int J1 = std::min(n, J0 + POTRF_NB);
    for (int j0 = J0; j0 < J1; j0 += 64) {
      const int jb = std::min(64, J1 - j0);
      double* A11 = A + j0 + (int64_t)j0 * lda;
      const int rest = n - j0 - jb;
      if (rest <= 0) {
        chol_diag_kernel<<<count, 64, 0, s>>>(jb, A11, lda, sA, info, j0, 1, nullptr);
        MSDP_HIP(hipGetLastError());
        continue;
      }
      double* A21 = A11 + jb;  // jb == 64 here
      chol_diag_panel_kernel<<<dim3(cdiv(rest, 64), count), 128, 0, s>>>(
          rest, A11, lda, sA, info, j0, jb, work + (int64_t)(j0 / 64) * 4096, sW);
      MSDP_HIP(hipGetLastError());
      const int w = J1 - j0 - jb;  // the panel's later columns
      if (w <= 0) continue;
      double* A22 = A21 + (int64_t)jb * lda;
      // the panel's later columns, all rows below the step: one launch (the
      // lower mask skips the square part's upper triangle)
      MSDP_TRY(gemm_guarded(s, false, true, rest, w, jb, -1.0, A21, lda, A21, lda, 1.0, A22, lda,
                            count, sA, sA, sA, true, info));
    }
    if (n > J1) {
      double* P = A + J1 + (int64_t)J0 * lda;
      MSDP_TRY(gemm_guarded(s, false, true, n - J1, n - J1, J1 - J0, -1.0, P, lda, P, lda, 1.0,
                            A + J1 + (int64_t)J1 * lda, lda, count, sA, sA, sA, true, info));
    }
  }
  return 0;
}

static int potrf_steps(hipStream_t s, int n, double* A, int64_t lda, int64_t sA, int count,
                       int* info, double* work, int64_t sW);

// count (1 or 2) matrices: the second at A + sA, its info at info[1]; work
// (potrf_work_elems(n) per matrix, the second at work + sW) receives the
// factors of the diagonal blocks on the way
static int potrf_count(hipStream_t s, int n, double* A, int64_t lda, int64_t sA, int count,
                       int* info, double* work, int64_t sW) {
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * count, s));
  if (n >= POTRF_BLOCKED_MIN && !shared_device())
    MSDP_TRY(potrf_blocked(s, n, A, lda, sA, count, info, work, sW));
  else
    MSDP_TRY(potrf_steps(s, n, A, lda, sA, count, info, work, sW));
  // every block but the last went through a fused diagonal-and-panel step
  const int nfused = cdiv(n, 64) - 1;
  if (nfused > 0) {
    diag_store_kernel<<<dim3(nfused, count), 256, 0, s>>>(A, lda, sA, work, sW, info);
    MSDP_HIP(hipGetLastError());
  }
  return 0;
}

static int potrf_steps(hipStream_t s, int n, double* A, int64_t lda, int64_t sA, int count,
                       int* info, double* work, int64_t sW) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int jb = std::min(64, n - j0);
    double* A11 = A + j0 + (int64_t)j0 * lda;
    const int rest = n - j0 - jb;
    if (rest == 0) {
      chol_diag_kernel<<<count, 64, 0, s>>>(jb, A11, lda, sA, info, j0, 1, nullptr);
      MSDP_HIP(hipGetLastError());
    }
    if (rest > 0) {  // jb == 64 here
      double* A21 = A11 + jb;
      chol_diag_panel_kernel<<<dim3(cdiv(rest, 64), count), 128, 0, s>>>(
          rest, A11, lda, sA, info, j0, jb, work + (int64_t)(j0 / 64) * 4096, sW);
      MSDP_HIP(hipGetLastError());
      double* A22 = A21 + (int64_t)jb * lda;
      MSDP_TRY(gemm_guarded(s, false, true, rest, rest, jb, -1.0, A21, lda, A21, lda, 1.0, A22, lda,
                            count, sA, sA, sA, true, info));
    }
  }
  return 0;
}

int potrf(hipStream_t s, int n, double* A, int64_t lda, int* info, double* work, bool inverses) {
  MSDP_TRY(rec_forbid("potrf"));
  MSDP_TRY(potrf_count(s, n, A, lda, 0, 1, info, work, 0));
  if (inverses) {  // all diagonal-block inverses in one launch
    diag_inv_kernel<<<cdiv(n, 64), 64, 0, s>>>(n, A, lda, work, info);
    MSDP_HIP(hipGetLastError());
  }
  return 0;
}

int potrf2(hipStream_t s, int n, double* A0, double* A1, int64_t lda, int* info, double* work) {
  MSDP_TRY(rec_forbid("potrf2"));
  const intptr_t d = reinterpret_cast<intptr_t>(A1) - reinterpret_cast<intptr_t>(A0);
  if (d % (intptr_t)sizeof(double) != 0) {
    set_error("potrf2: the two matrices are not a whole number of doubles apart");
    return -2;
  }
  return potrf_count(s, n, A0, lda, (int64_t)(d / (intptr_t)sizeof(double)), 2, info, work,
                     (int64_t)potrf_work_elems(n));
}

int potrf_small_batched(hipStream_t s, int n, double* A, int64_t lda, int64_t stride, int batch,
                        int* info) {
  if (n > 64) { set_error("potrf_small_batched: n > 64"); return -2; }
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * batch, s));
  chol_diag_kernel<<<batch, 64, 0, s>>>(n, A, lda, stride, info, 0, 1, nullptr);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// Triangular inverse (lower), recursive: diagonal 64-blocks inverted in LDS,
// off-diagonal blocks X21 = -L22^-1 L21 L11^-1 by two GEMMs.
// ==========================================================================
template <int K>
__device__ __attribute__((always_inline)) inline void trtri64_step(const double* T, int nb,
                                                                   double (&x)[64]) {
  const int lane = threadIdx.x & 63;
  if (K < nb) {
    const double tkk = T[K * W64_LD + K];
    x[K] = K == lane ? 1.0 / tkk : x[K] / tkk;
#pragma unroll
    for (int g = ((K + 1) / 16) * 16; g < 64; g += 16) {
      double lc[16];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (g + q > K) lc[q] = T[K * W64_LD + g + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (g + q > K) x[g + q] -= lc[q] * x[K];
    }
  }
  if constexpr (K + 1 < 64) trtri64_step<K + 1>(T, nb, x);
}

template <bool DENSE64 = false>
__device__ __attribute__((always_inline)) inline void trtri64_body(int n, const double* L,
                                                                   int64_t ldl, double* Li,
                                                                   int64_t ldi, int b0) {
  // Column c of T^-1 by forward substitution in lane c, right-looking: x_k
  // is final at step k (1 / T_kk for k = c, x_k / T_kk below), then
  // x_i -= T_ik x_k for i > k.
  const int nb = min(64, n - b0);
  __shared__ double T[64 * W64_LD];
  const int lane = threadIdx.x;
  double x[64];
  // all 64 loads in flight before the first LDS store
#pragma unroll
  for (int j = 0; j < 64; ++j)
    x[j] = (lane < nb && j < nb && lane >= j) ? L[(b0 + lane) + (int64_t)(b0 + j) * ldl] : 0.0;
#pragma unroll
  for (int j = 0; j < 64; ++j) T[j * W64_LD + lane] = x[j];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 64; ++i) x[i] = 0.0;
  trtri64_step<0>(T, nb, x);
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 64; ++i) T[i * W64_LD + lane] = x[i];
  wave_lds_sync();
  if constexpr (DENSE64) {  // the whole 64 x 64 tile, zeros outside the nb x nb block
#pragma unroll 8
    for (int c = 0; c < 64; ++c)
      Li[lane + (int64_t)c * ldi] = (lane < nb && c < nb) ? T[lane * W64_LD + c] : 0.0;
    return;
  }
  if (lane >= nb) return;
#pragma unroll 8
  for (int c = 0; c < 64; ++c)
    if (c < nb) Li[(b0 + lane) + (int64_t)(b0 + c) * ldi] = T[lane * W64_LD + c];
}

// inverses of the diagonal 64-blocks of a Cholesky factor, block b into
// Di + 4096 b (64 x 64, ld 64, zeros outside the lower triangle of the block)
__global__ __launch_bounds__(64) void diag_inv_kernel(int n, const double* L, int64_t ldl,
                                                      double* Di, const int* info) {
  if (*info != 0) return;
  const int b0 = blockIdx.x * 64;
  trtri64_body<true>(n - b0, L + b0 + (int64_t)b0 * ldl, ldl, Di + (size_t)blockIdx.x * 4096, 64,
                     0);
}

// the diagonal 64-blocks (blockIdx.x) of a batch of matrices (blockIdx.y)
__global__ __launch_bounds__(64) void trtri_diag_kernel(int n, const double* L, int64_t ldl,
                                                        int64_t sL, double* Li, int64_t ldi,
                                                        int64_t sLi) {
  trtri64_body(n, L + (int64_t)blockIdx.y * sL, ldl, Li + (int64_t)blockIdx.y * sLi, ldi,
               blockIdx.x * 64);
}

// recorded n <= 64 inverses of different blocks, one wavefront each
__global__ __launch_bounds__(64) void trtri64_vb_kernel(const TrtriItem* __restrict__ items) {
  const TrtriItem it = items[blockIdx.x];
  trtri64_body(it.n, it.L, it.ldl, it.Li, it.ldi, 0);
}

// L^-1 for 64 < n <= 128 in one workgroup of two wavefronts: the inverses
// of the two diagonal blocks (wavefront 0: L11, wavefront 1: L22, each by
// trtri64_body's column substitution), then W = L21 L11^-1 and
// X21 = -L22^-1 W (each wavefront half of the 64 columns, a row per lane),
// and the whole factor written with a zero upper triangle.  One launch
// where the recursive path takes four.
__device__ __attribute__((always_inline)) inline void trtri_mid_body(int n, const double* L,
                                                                     int64_t ldl, double* Li,
                                                                     int64_t ldi) {
  __shared__ double T1[64 * W64_LD];  // L11, then L11^-1 by rows: T1[r * W64_LD + c]
  __shared__ double T2[64 * W64_LD];  // L22, then L22^-1 by rows
  __shared__ double Wl[64 * W64_LD];  // W by columns: Wl[k * W64_LD + i] = W[i][k]
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n2 = n - 64;
  const int b0 = w == 0 ? 0 : 64, nb = w == 0 ? 64 : n2;
  double* T = w == 0 ? T1 : T2;
  {
    double x[64];
#pragma unroll
    for (int j = 0; j < 64; ++j)
      x[j] = (lane < nb && j < nb && lane >= j) ? L[(b0 + lane) + (int64_t)(b0 + j) * ldl] : 0.0;
#pragma unroll
    for (int j = 0; j < 64; ++j) T[j * W64_LD + lane] = x[j];
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 64; ++i) x[i] = 0.0;
    trtri64_step<0>(T, nb, x);
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 64; ++i) T[i * W64_LD + lane] = x[i];
  }
  __syncthreads();
  // W[i][k] = sum_j L21[i][j] L11^-1[j][k] for this wavefront's columns k
  const int k0 = 32 * w;
  {
    double l[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) l[j] = lane < n2 ? L[(64 + lane) + (int64_t)j * ldl] : 0.0;
    double acc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.0;
#pragma unroll
    for (int j = 0; j < 64; ++j)
#pragma unroll
      for (int q = 0; q < 32; ++q) acc[q] = fma(l[j], T1[j * W64_LD + k0 + q], acc[q]);
#pragma unroll
    for (int q = 0; q < 32; ++q) Wl[(k0 + q) * W64_LD + lane] = acc[q];
  }
  __syncthreads();
  // X21[i][k] = -sum_j L22^-1[i][j] W[j][k]
  if (lane < n2) {
    double acc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.0;
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      const double a = T2[lane * W64_LD + j];
#pragma unroll
      for (int q = 0; q < 32; ++q) acc[q] = fma(a, Wl[(k0 + q) * W64_LD + j], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < 32; ++q) Li[(64 + lane) + (int64_t)(k0 + q) * ldi] = -acc[q];
  }
  // the diagonal blocks and the zero upper triangle: wavefront 0 rows 0..63,
  // wavefront 1 rows 64..n-1
  if (w == 0) {
    for (int c = 0; c < n; ++c)
      Li[lane + (int64_t)c * ldi] = (c < 64 && c <= lane) ? T1[lane * W64_LD + c] : 0.0;
  } else if (lane < n2) {
    for (int c = 0; c < n2; ++c)
      Li[(64 + lane) + (int64_t)(64 + c) * ldi] = c <= lane ? T2[lane * W64_LD + c] : 0.0;
  }
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1))) void trtri_mid_kernel(
    int n, const double* L, int64_t ldl, int64_t sL, double* Li, int64_t ldi, int64_t sLi) {
  trtri_mid_body(n, L + (int64_t)blockIdx.x * sL, ldl, Li + (int64_t)blockIdx.x * sLi, ldi);
}

// recorded 64 < n <= 128 inverses of different blocks, one workgroup each
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1))) void
trtri_mid_vb_kernel(const TrtriItem* __restrict__ items) {
  const TrtriItem it = items[blockIdx.x];
  trtri_mid_body(it.n, it.L, it.ldl, it.Li, it.ldi);
}

__global__ void zero_upper_kernel(int n, double* A, int64_t lda, int64_t sA) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  A += (int64_t)blockIdx.z * sA;
  if (i < n && i < j) A[i + (int64_t)j * lda] = 0.0;
}

// All nodes of one depth of the recursion (count <= TRL_MAX, each smaller
// than 512 on both sides) in one launch: stage 1 forms every node's
// work = L21 L11^-1 (the node's own slice of the work buffer), stage 2 its
// X21 = -L22^-1 work.  The nodes are independent; each tile is the 64-tile
// kernel's, so the arithmetic is that of one product per node.
constexpr int TRL_MAX = 64;
struct TrtriLevel {
  int count;
  int lo[TRL_MAX], mid[TRL_MAX], hi[TRL_MAX], start[TRL_MAX + 1];
  int64_t woff[TRL_MAX];
};

template <int STAGE>
__global__ __launch_bounds__(256) void trtri_level_kernel(TrtriLevel lv, const double* L,
                                                          int64_t ldl, double* Li, int64_t ldi,
                                                          double* work, int64_t sL, int64_t sLi,
                                                          int64_t sW) {
  // matrix blockIdx.y of a batch at the given strides
  L += blockIdx.y * sL;
  Li += blockIdx.y * sLi;
  work += blockIdx.y * sW;
  int lo = 0, hi = lv.count - 1;
  const int w = blockIdx.x;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (lv.start[m] <= w) lo = m; else hi = m - 1;
  }
  const int k = lo;
  const int a = lv.lo[k], b = lv.mid[k], c = lv.hi[k];
  const int r = c - b, cc = b - a;
  const int t = w - lv.start[k];
  const int ntm = (r + GT - 1) / GT;
  double* wk = work + lv.woff[k];
  if (STAGE == 1)
    gemm64_tile<false, false>(r, cc, cc, 1.0, L + b + (int64_t)a * ldl, ldl,
                              Li + a + (int64_t)a * ldi, ldi, 0.0, wk, r, 0, t % ntm, t / ntm,
                              TRI_BL);
  else
    gemm64_tile<false, false>(r, cc, r, -1.0, Li + b + (int64_t)b * ldi, ldi, wk, r, 0.0,
                              Li + b + (int64_t)a * ldi, ldi, 0, t % ntm, t / ntm, TRI_AL);
}

static void trtri_tree(int lo, int hi, int depth, std::vector<std::vector<std::array<int, 3>>>& lv) {
  const int nblk = cdiv(hi - lo, 64);
  if (nblk <= 1) return;
  const int mid = lo + 64 * ((nblk + 1) / 2);
  trtri_tree(lo, mid, depth + 1, lv);
  trtri_tree(mid, hi, depth + 1, lv);
  if ((int)lv.size() <= depth) lv.resize(depth + 1);
  lv[depth].push_back({lo, mid, hi});
}

// The recursion of trtri_rec for one matrix, depth by depth from the leaves.
static int trtri_levels(hipStream_t s, int n, const double* L, int64_t ldl, double* Li,
                        int64_t ldi, double* work, int count = 1, int64_t sL = 0,
                        int64_t sLi = 0, int64_t sW = 0) {
  std::vector<std::vector<std::array<int, 3>>> lv;
  trtri_tree(0, n, 0, lv);
  for (int dpt = (int)lv.size() - 1; dpt >= 0; --dpt) {
    const auto& nodes = lv[dpt];
    bool small = (int)nodes.size() <= TRL_MAX;
    for (const auto& nd : nodes) small = small && nd[2] - nd[1] < 512 && nd[1] - nd[0] < 512;
    if (!small) {
      for (const auto& nd : nodes) {
        const int lo = nd[0], mid = nd[1], hi = nd[2];
        const int r = hi - mid, c = mid - lo;
        MSDP_TRY(gemm(s, false, false, r, c, c, 1.0, L + mid + (int64_t)lo * ldl, ldl,
                      Li + lo + (int64_t)lo * ldi, ldi, 0.0, work, r, count, sL, sLi, sW));
        MSDP_TRY(gemm(s, false, false, r, c, r, -1.0, Li + mid + (int64_t)mid * ldi, ldi, work, r,
                      0.0, Li + mid + (int64_t)lo * ldi, ldi, count, sLi, sW, sLi));
      }
      continue;
    }
    TrtriLevel p{};
    p.count = (int)nodes.size();
    int tiles = 0;
    int64_t woff = 0;
    for (int k = 0; k < p.count; ++k) {
      p.lo[k] = nodes[k][0];
      p.mid[k] = nodes[k][1];
      p.hi[k] = nodes[k][2];
      const int r = p.hi[k] - p.mid[k], c = p.mid[k] - p.lo[k];
      p.start[k] = tiles;
      p.woff[k] = woff;
      tiles += cdiv(r, GT) * cdiv(c, GT);
      woff += (int64_t)r * c;
    }
    p.start[p.count] = tiles;
    trtri_level_kernel<1><<<dim3(tiles, count), 256, 0, s>>>(p, L, ldl, Li, ldi, work, sL, sLi, sW);
    trtri_level_kernel<2><<<dim3(tiles, count), 256, 0, s>>>(p, L, ldl, Li, ldi, work, sL, sLi, sW);
    MSDP_HIP(hipGetLastError());
  }
  return 0;
}

// One level of the recursion for every matrix of the batch at once (the
// batch strides sL, sLi, sW apply to L, Li and the work buffer).
static int trtri_rec(hipStream_t s, int lo, int hi, const double* L, int64_t ldl, int64_t sL,
                     double* Li, int64_t ldi, int64_t sLi, double* work, int64_t sW, int count) {
  const int nblk = cdiv(hi - lo, 64);
  if (nblk <= 1) return 0;
  const int mid = lo + 64 * ((nblk + 1) / 2);
  MSDP_TRY(trtri_rec(s, lo, mid, L, ldl, sL, Li, ldi, sLi, work, sW, count));
  MSDP_TRY(trtri_rec(s, mid, hi, L, ldl, sL, Li, ldi, sLi, work, sW, count));
  const int r = hi - mid, c = mid - lo;
  // work = L21 * L11inv   (r x c)
  MSDP_TRY(gemm(s, false, false, r, c, c, 1.0, L + mid + (int64_t)lo * ldl, ldl,
                Li + lo + (int64_t)lo * ldi, ldi, 0.0, work, r, count, sL, sLi, sW));
  // X21 = -L22inv * work
  MSDP_TRY(gemm(s, false, false, r, c, r, -1.0, Li + mid + (int64_t)mid * ldi, ldi, work, r, 0.0,
                Li + mid + (int64_t)lo * ldi, ldi, count, sLi, sW, sLi));
  return 0;
}

size_t trtri_work_elems(int n) { return (size_t)(n / 2 + 64) * (size_t)(n / 2 + 64); }

int trtri_lower_batched(hipStream_t s, int n, int count, const double* L, int64_t ldl,
                        int64_t sL, double* Li, int64_t ldi, int64_t sLi, double* work,
                        int64_t sW) {
  if (rec_active() && count == 1 && n > 0 && n <= 64) {
    // no recursion, no workspace: the zeroed upper triangle and the one
    // diagonal-block inverse as two recorded ops
    RecOp z;
    z.kind = REC_EW;
    z.e = EwItem{EW_ZERO_UPPER, n, n, 0, 0, ldi, 0, 0, 0, 0, nullptr, nullptr, nullptr, Li};
    z.run = [=]() -> int {
      zero_upper_kernel<<<dim3(cdiv(n, 256), n, 1), 256, 0, s>>>(n, Li, ldi, 0);
      MSDP_HIP(hipGetLastError());
      return 0;
    };
    MSDP_TRY(rec_add(std::move(z)));
    RecOp t;
    t.kind = REC_TRTRI64;
    t.tr = TrtriItem{n, 0, L, Li, ldl, ldi};
    t.run = [=]() -> int {
      trtri_diag_kernel<<<dim3(1, 1), 64, 0, s>>>(n, L, ldl, 0, Li, ldi, 0);
      MSDP_HIP(hipGetLastError());
      return 0;
    };
    return rec_add(std::move(t));
  }
  if (n > 64 && n <= 128) {  // one workgroup per matrix
    if (rec_active() && count == 1) {
      RecOp t;
      t.kind = REC_TRTRI128;
      t.tr = TrtriItem{n, 0, L, Li, ldl, ldi};
      t.run = [=]() -> int {
        trtri_mid_kernel<<<1, 128, 0, s>>>(n, L, ldl, 0, Li, ldi, 0);
        MSDP_HIP(hipGetLastError());
        return 0;
      };
      return rec_add(std::move(t));
    }
    MSDP_TRY(rec_forbid("trtri"));
    if (count <= 0) return 0;
    trtri_mid_kernel<<<count, 128, 0, s>>>(n, L, ldl, sL, Li, ldi, sLi);
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  MSDP_TRY(rec_forbid("trtri"));
  if (n <= 0 || count <= 0) return 0;
  dim3 g(cdiv(n, 256), n, count);
  zero_upper_kernel<<<g, 256, 0, s>>>(n, Li, ldi, sLi);
  MSDP_HIP(hipGetLastError());
  trtri_diag_kernel<<<dim3(cdiv(n, 64), count), 64, 0, s>>>(n, L, ldl, sL, Li, ldi, sLi);
  MSDP_HIP(hipGetLastError());
  // every depth of the recursion in one pair of launches for all matrices
  if (count <= 65535) return trtri_levels(s, n, L, ldl, Li, ldi, work, count, sL, sLi, sW);
  return trtri_rec(s, 0, n, L, ldl, sL, Li, ldi, sLi, work, sW, count);
}

int trtri_lower(hipStream_t s, int n, const double* L, int64_t ldl, double* Li, int64_t ldi,
                double* work) {
  return trtri_lower_batched(s, n, 1, L, ldl, 0, Li, ldi, 0, work, 0);
}

// ==========================================================================
// reductions (deterministic: fixed grid, partials added in a fixed order)
// ==========================================================================
constexpr int RED_BLOCKS = 512;

// One launch: every workgroup reduces its grid-strided share into a partial,
// and the last one to finish (an integer counter in the scratch after the
// partials, reset by that workgroup) adds the partials in index order.  The
// grid is fixed by the count, so the order, and the result, is too.
template <int OP>
__global__ __launch_bounds__(256) void reduce_one_kernel(int64_t count, const double* X,
                                                         const double* Y, double* partial,
                                                         int* counter, double* out) {
  __shared__ double sh[4];
  __shared__ int last;
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256) {
    if (OP == 0) v += X[i] * Y[i];
    else v = fmax(v, fabs(X[i]));
  }
  v = block_reduce<OP == 0 ? 0 : 1>(v, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = v;
    __threadfence();
    last = atomicAdd(counter, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  const volatile double* pv = partial;
  double w = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += 256) w = OP == 0 ? w + pv[i] : fmax(w, pv[i]);
  w = block_reduce<OP == 0 ? 0 : 1>(w, sh);
  if (threadIdx.x == 0) {
    *out = w;
    *counter = 0;
  }
}

// the partials (RED_BLOCKS) and, after them, the counter of reduce_one_kernel
// (zero when the scratch is handed out)
size_t reduce_scratch_elems(int64_t) { return RED_BLOCKS; }

int dot(hipStream_t s, int64_t count, const double* X, const double* Y, double* out,
        double* scratch) {
  if (rec_active()) {
    RecOp o;
    if (count > 0 && count <= (int64_t)RED_BLOCKS * 1024) {
      o.kind = REC_RED;
      o.r = RedItem{RED_DOT, 0, count, X, Y, out, nullptr};
    }
    o.run = [=] { return dot(s, count, X, Y, out, scratch); };
    return rec_add(std::move(o));
  }
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (count + 1023) / 1024));
  reduce_one_kernel<0><<<g, 256, 0, s>>>(count, X, Y, scratch,
                                         reinterpret_cast<int*>(scratch + RED_BLOCKS), out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int max_abs(hipStream_t s, int64_t count, const double* X, double* out, double* scratch) {
  if (rec_active()) {
    RecOp o;
    if (count > 0 && count <= (int64_t)RED_BLOCKS * 1024) {
      o.kind = REC_RED;
      o.r = RedItem{RED_MAXABS, 0, count, X, nullptr, out, nullptr};
    }
    o.run = [=] { return max_abs(s, count, X, out, scratch); };
    return rec_add(std::move(o));
  }
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (count + 1023) / 1024));
  reduce_one_kernel<1><<<g, 256, 0, s>>>(count, X, nullptr, scratch,
                                         reinterpret_cast<int*>(scratch + RED_BLOCKS), out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// Cholesky flush (src/blocks.jl:101-111)
// ==========================================================================
__global__ __launch_bounds__(256) void diag_maxabs_kernel(int n, const double* L, int64_t ld,
                                                          double* out) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v = fmax(v, fabs(L[i + (int64_t)i * ld]));
  v = block_reduce<1>(v, sh);
  if (threadIdx.x == 0) *out = v;
}

__global__ void flush_kernel(int n, double* L, int64_t ld, const double* dmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const double eps = 2.220446049250313e-16;
  const double tol = eps * eps * (*dmax);
  double* p = L + i + (int64_t)j * ld;
  if (i < j) *p = 0.0;
  else if (i > j && fabs(*p) < tol) *p = 0.0;
}

int chol_flush(hipStream_t s, int n, double* L, int64_t ld, double* scratch) {
  MSDP_TRY(rec_forbid("chol_flush"));
  if (n <= 0) return 0;
  diag_maxabs_kernel<<<1, 256, 0, s>>>(n, L, ld, scratch);
  dim3 g(cdiv(n, 256), n);
  flush_kernel<<<g, 256, 0, s>>>(n, L, ld, scratch);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// _cholesky_factor! for n <= 64 in one launch of one wavefront: the
// symmetrised copy, the factorisation (chol64_rows: one row per lane in
// registers), the flush and the zeroed upper triangle.  Same arithmetic per
// entry as the blocked path's diagonal kernel.
__device__ __attribute__((always_inline)) inline void chol_small_body(int n, const double* M,
                                                                      int64_t ldm, double* L,
                                                                      int64_t ldl, int* info) {
  __shared__ double Lc[64 * W64_LD];
  const int lane = threadIdx.x;
  double r[64];
#pragma unroll
  for (int c = 0; c < 64; ++c)
    r[c] = (lane < n && c <= lane)
               ? (M[lane + (int64_t)c * ldm] + M[c + (int64_t)lane * ldm]) / 2
               : 0.0;
  const int fail = chol64_rows(r, n, Lc);
  if (lane == 0) *info = fail;
  if (fail) return;
  wave_lds_sync();
  double dm = lane < n ? fabs(Lc[lane * W64_LD + lane]) : 0.0;
  dm = wave_max(dm);
  const double eps = 2.220446049250313e-16;
  const double tol = eps * eps * dm;
  if (lane >= n) return;
#pragma unroll
  for (int c = 0; c < 64; ++c) {
    if (c < n) {
      double v = 0.0;
      if (lane == c) v = r[c];
      else if (lane > c) v = fabs(r[c]) < tol ? 0.0 : r[c];
      L[lane + (int64_t)c * ldl] = v;
    }
  }
}

__global__ __launch_bounds__(64) void chol_small_kernel(int n, const double* M, int64_t ldm,
                                                        double* L, int64_t ldl, int* info) {
  chol_small_body(n, M, ldm, L, ldl, info);
}

// a batch of blocks of different sizes, one wavefront each
__global__ __launch_bounds__(64) void chol_small_vb_kernel(const CholItem* __restrict__ items) {
  const CholItem it = items[blockIdx.x];
  chol_small_body(it.n, it.M, it.ldm, it.L, it.ldl, it.info);
}

// _cholesky_factor! for 64 < n <= 128 in one workgroup of two wavefronts:
// the symmetrised copy; the first 64 columns factored (wavefront 0) with the
// rows below substituted one column behind (wavefront 1), as in
// chol_diag_panel_kernel; the trailing block updated (A22 -= L21 L21',
// rank-1 steps in k order) and factored by wavefront 1; then the flush of
// entries below eps^2 max|L_ii| and the zeroed upper triangle.  One launch
// where the blocked path takes nine.
__device__ __attribute__((always_inline)) inline void chol_mid_body(int n, const double* M,
                                                                    int64_t ldm, double* L,
                                                                    int64_t ldl, int* info) {
  __shared__ double Lt1[64 * W64_LD];  // L11 by columns: Lt1[k * W64_LD + j] = L[j][k]
  __shared__ double Lt2[64 * W64_LD];  // L22, the same layout
  __shared__ double R21[64 * W64_LD];  // L21 by rows: R21[i * W64_LD + k] = L[64 + i][k]
  __shared__ double s_dm[2];
  __shared__ int s_fail;
  const int lane = threadIdx.x & 63;
  const bool w0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0;
  const int n2 = n - 64;     // order of the trailing block, 1..64
  const int i2 = 64 + lane;  // wavefront 1's row
  double v[64];
  if (w0) {
#pragma unroll
    for (int c = 0; c < 64; ++c)
      v[c] = c <= lane ? (M[lane + (int64_t)c * ldm] + M[c + (int64_t)lane * ldm]) / 2 : 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < 64; ++k)
      v[k] = lane < n2 ? (M[i2 + (int64_t)k * ldm] + M[k + (int64_t)i2 * ldm]) / 2 : 0.0;
  }
  if (threadIdx.x == 0) s_fail = 0;
  int fail = 0;
  chol_pipe_step<0>(w0, v, 64, Lt1, fail);
  if (w0) {
    if (fail && lane == 0) s_fail = fail;
  } else {
#pragma unroll
    for (int k = 0; k < 64; ++k) R21[lane * W64_LD + k] = v[k];
  }
  __syncthreads();
  if (s_fail) {
    if (threadIdx.x == 0) *info = s_fail;
    return;
  }
  if (!w0) {
#pragma unroll
    for (int c = 0; c < 64; ++c)
      v[c] = (lane < n2 && c <= lane)
                 ? (M[i2 + (int64_t)(64 + c) * ldm] + M[(64 + c) + (int64_t)i2 * ldm]) / 2
                 : 0.0;
    for (int k = 0; k < 64; ++k) {
      const double x = R21[lane * W64_LD + k];
#pragma unroll
      for (int c = 0; c < 64; ++c) v[c] -= x * R21[c * W64_LD + k];
    }
    fail = chol64_rows(v, n2, Lt2);
    if (fail && lane == 0) s_fail = 64 + fail;
  }
  __syncthreads();
  if (s_fail) {
    if (threadIdx.x == 0) *info = s_fail;
    return;
  }
  if (threadIdx.x == 0) *info = 0;
  double dm = w0 ? fabs(Lt1[lane * W64_LD + lane]) : (lane < n2 ? fabs(Lt2[lane * W64_LD + lane]) : 0.0);
  dm = wave_max(dm);
  if (lane == 0) s_dm[threadIdx.x >> 6] = dm;
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  const double tol = eps * eps * fmax(s_dm[0], s_dm[1]);
  auto flushed = [tol](double x, bool diag) { return diag ? x : (fabs(x) < tol ? 0.0 : x); };
  if (w0) {
    for (int c = 0; c < 64; ++c) {
      const double x = Lt1[c * W64_LD + lane];
      L[lane + (int64_t)c * ldl] = lane >= c ? flushed(x, lane == c) : 0.0;
    }
    for (int c = 64; c < n; ++c) L[lane + (int64_t)c * ldl] = 0.0;
  } else if (lane < n2) {
    for (int c = 0; c < 64; ++c)
      L[i2 + (int64_t)c * ldl] = flushed(R21[lane * W64_LD + c], false);
    for (int c = 0; c < n2; ++c) {
      const double x = Lt2[c * W64_LD + lane];
      L[i2 + (int64_t)(64 + c) * ldl] = lane >= c ? flushed(x, lane == c) : 0.0;
    }
  }
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1))) void chol_mid_kernel(
    int n, const double* M, int64_t ldm, double* L, int64_t ldl, int* info) {
  chol_mid_body(n, M, ldm, L, ldl, info);
}

// a batch of blocks of different sizes, one workgroup each
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1))) void
chol_mid_vb_kernel(const CholItem* __restrict__ items) {
  const CholItem it = items[blockIdx.x];
  chol_mid_body(it.n, it.M, it.ldm, it.L, it.ldl, it.info);
}

int cholesky_mid(hipStream_t s, int n, const double* M, int64_t ldm, double* L, int64_t ldl,
                 int* info) {
  if (n <= 64 || n > 128) {
    set_error("cholesky_mid: n must be in 65..128");
    return -2;
  }
  if (rec_active()) {
    RecOp o;
    o.kind = REC_CHOL128;
    o.ch = CholItem{n, 0, M, L, info, ldm, ldl};
    o.run = [=] { return cholesky_mid(s, n, M, ldm, L, ldl, info); };
    return rec_add(std::move(o));
  }
  chol_mid_kernel<<<1, 128, 0, s>>>(n, M, ldm, L, ldl, info);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int cholesky_small(hipStream_t s, int n, const double* M, int64_t ldm, double* L, int64_t ldl,
                   int* info) {
  if (n <= 0 || n > 64) {
    set_error("cholesky_small: n must be in 1..64");
    return -2;
  }
  if (rec_active()) {
    RecOp o;
    o.kind = REC_CHOL64;
    o.ch = CholItem{n, 0, M, L, info, ldm, ldl};
    o.run = [=] { return cholesky_small(s, n, M, ldm, L, ldl, info); };
    return rec_add(std::move(o));
  }
  chol_small_kernel<<<1, 64, 0, s>>>(n, M, ldm, L, ldl, info);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// elementwise
// ==========================================================================
// (A + A')/2 (mode 0) or the lower triangle copied over the upper (mode 1),
// in place, by 64 x 64 tile pairs: workgroup b takes the tiles (I, J) and
// (J, I), I >= J, b = I (I + 1) / 2 + J.  Both tiles are read and written
// along their columns and the transpose goes through LDS; each pair (x > y)
// gets (A[x,y] + A[y,x]) / 2, the arithmetic of the batched form.
__global__ __launch_bounds__(256) void sym_tiles_kernel(int n, double* A, int64_t lda, int mode,
                                                        int64_t sA = 0) {
  __shared__ double Ut[64][65];  // Ut[yy][xx] = A[y0 + yy, x0 + xx]
  A += blockIdx.y * sA;  // matrix blockIdx.y of a batch
  const int b = blockIdx.x;
  int I = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= b) ++I;
  while (I * (I + 1) / 2 > b) --I;
  const int J = b - I * (I + 1) / 2;
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int x0 = 64 * I, y0 = 64 * J;
#pragma unroll
  for (int q = 0; q < 16; ++q) {  // the upper tile
    const int xx = 16 * g + q, row = y0 + r, col = x0 + xx;
    Ut[r][xx] = (row < n && col < n) ? A[row + (int64_t)col * lda] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {  // the lower tile: x = x0 + r, y = y0 + yy
    const int yy = 16 * g + q, x = x0 + r, y = y0 + yy;
    if (x < n && y < n && x > y) {
      double* a = A + x + (int64_t)y * lda;
      double v = *a;
      if (mode == 0) {
        v = (v + Ut[yy][r]) / 2;
        *a = v;
      }
      Ut[yy][r] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {  // the upper tile back
    const int xx = 16 * g + q, row = y0 + r, col = x0 + xx;
    if (row < n && col < n && col > row) A[row + (int64_t)col * lda] = Ut[r][xx];
  }
}

static int sym_tiles(hipStream_t s, int n, double* A, int64_t lda, int mode) {
  const int nt = cdiv(n, 64);
  sym_tiles_kernel<<<nt * (nt + 1) / 2, 256, 0, s>>>(n, A, lda, mode);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// (A + A')/2 of two matrices of the same side in one launch (A1 at A0 + sA)
int symmetrize2(hipStream_t s, int n, double* A0, int64_t sA, int64_t lda) {
  MSDP_TRY(rec_forbid("symmetrize2"));
  if (n <= 1) return 0;
  const int nt = cdiv(n, 64);
  sym_tiles_kernel<<<dim3(nt * (nt + 1) / 2, 2), 256, 0, s>>>(n, A0, lda, 0, sA);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int symmetrize(hipStream_t s, int n, double* A, int64_t lda) {
  if (n <= 1) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_SYM, n, n, 0, 0, lda, 0, 0, 0, 0, nullptr, nullptr, nullptr, A};
    o.run = [=] { return symmetrize(s, n, A, lda); };
    return rec_add(std::move(o));
  }
  return sym_tiles(s, n, A, lda, 0);
}

int mirror_lower(hipStream_t s, int n, double* A, int64_t lda) {
  if (n <= 1) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_MIRROR, n, n, 0, 0, lda, 0, 0, 0, 0, nullptr, nullptr, nullptr, A};
    o.run = [=] { return mirror_lower(s, n, A, lda); };
    return rec_add(std::move(o));
  }
  return sym_tiles(s, n, A, lda, 1);
}

__global__ void copy_kernel(int rows, const double* A, int64_t lda, double* B, int64_t ldb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < rows) B[i + (int64_t)j * ldb] = A[i + (int64_t)j * lda];
}

int copy_mat(hipStream_t s, int rows, int cols, const double* A, int64_t lda, double* B,
             int64_t ldb) {
  if (rows <= 0 || cols <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_COPY, rows, cols, 0, 0, lda, ldb, 0, 0, 0, A, nullptr, nullptr, B};
    o.run = [=] { return copy_mat(s, rows, cols, A, lda, B, ldb); };
    return rec_add(std::move(o));
  }
  dim3 g(cdiv(rows, 256), cols);
  copy_kernel<<<g, 256, 0, s>>>(rows, A, lda, B, ldb);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void transpose_kernel(int rows, int cols, const double* A, int64_t lda, double* B,
                                 int64_t ldb) {
  __shared__ double t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int i = bx + tx, j = by + k;
    if (i < rows && j < cols) t[k][tx] = A[i + (int64_t)j * lda];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int j = by + tx, i = bx + k;
    if (i < rows && j < cols) B[j + (int64_t)i * ldb] = t[tx][k];
  }
}

int transpose(hipStream_t s, int rows, int cols, const double* A, int64_t lda, double* B,
              int64_t ldb) {
  if (rows <= 0 || cols <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_TRANSPOSE, rows, cols, 0, 0, lda, ldb, 0, 0, 0, A, nullptr, nullptr, B};
    o.run = [=] { return transpose(s, rows, cols, A, lda, B, ldb); };
    return rec_add(std::move(o));
  }
  dim3 g(cdiv(rows, 32), cdiv(cols, 32));
  transpose_kernel<<<g, 256, 0, s>>>(rows, cols, A, lda, B, ldb);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void axpby_kernel(int64_t count, double a, const double* X, double b, const double* Y,
                             double* Z) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    Z[i] = a * X[i] + b * Y[i];
}

__global__ void axpbypcz_kernel(int64_t count, double a, const double* X, double b,
                                const double* Y, double c, const double* V, double* Z) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    Z[i] = a * X[i] + b * Y[i] + c * V[i];
}

static int ew_grid(int64_t count) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(8192, (count + 255) / 256));
}

int axpby(hipStream_t s, int64_t count, double a, const double* X, double b, const double* Y,
          double* Z) {
  if (count <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_AXPBY, 0, 0, 0, count, 0, 0, a, b, 0, X, Y, nullptr, Z};
    o.run = [=] { return axpby(s, count, a, X, b, Y, Z); };
    return rec_add(std::move(o));
  }
  axpby_kernel<<<ew_grid(count), 256, 0, s>>>(count, a, X, b, Y, Z);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int axpbypcz(hipStream_t s, int64_t count, double a, const double* X, double b, const double* Y,
             double c, const double* V, double* Z) {
  if (count <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_AXPBYPCZ, 0, 0, 0, count, 0, 0, a, b, c, X, Y, V, Z};
    o.run = [=] { return axpbypcz(s, count, a, X, b, Y, c, V, Z); };
    return rec_add(std::move(o));
  }
  axpbypcz_kernel<<<ew_grid(count), 256, 0, s>>>(count, a, X, b, Y, c, V, Z);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void identity_kernel(int n, double* A, int64_t lda, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < n) A[i + (int64_t)j * lda] = i == j ? scale : 0.0;
}

int set_identity(hipStream_t s, int n, double* A, int64_t lda, double scale) {
  if (n <= 0) return 0;
  dim3 g(cdiv(n, 256), n);
  identity_kernel<<<g, 256, 0, s>>>(n, A, lda, scale);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void fill_kernel(int64_t count, double v, double* X) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    X[i] = v;
}

int gemm_vb_record(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                   double* C, int64_t ldc, int batch, int64_t sA, int64_t sB, int64_t sC,
                   bool lower) {
  if (!rec_active()) {
    set_error("gemm_vb_record: no recorded region");
    return -2;
  }
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return 0;
  RecOp o;
  o.kind = ta ? (tb ? REC_GEMM_TT : REC_GEMM_TN) : (tb ? REC_GEMM_NT : REC_GEMM_NN);
  o.g = GemmItem{M, N, K, lower ? 1 : 0, 0, alpha, beta, A, B, C, lda, ldb, ldc, batch, sA, sB,
                 sC};
  o.run = [=] {
    return gemm(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB, sC, lower);
  };
  return rec_add(std::move(o));
}

int svec_record(hipStream_t s, int n, int count, const double* P, double* B, int64_t ldb) {
  if (!rec_active()) {
    set_error("svec_record: no recorded region");
    return -2;
  }
  const int64_t d = (int64_t)n * (n + 1) / 2;
  if (d <= 0 || count <= 0) return 0;
  RecOp o;
  o.kind = REC_EW;
  o.e = EwItem{EW_SVEC, (int)d, count, n, 0, 0, ldb, 0, 0, 0, P, nullptr, nullptr, B};
  o.run = [=] { return svec_pack(s, n, count, P, B, ldb); };
  return rec_add(std::move(o));
}

int fill(hipStream_t s, int64_t count, double v, double* X) {
  if (count <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_FILL, 0, 0, 0, count, 0, 0, v, 0, 0, nullptr, nullptr, nullptr, X};
    o.run = [=] { return fill(s, count, v, X); };
    return rec_add(std::move(o));
  }
  fill_kernel<<<ew_grid(count), 256, 0, s>>>(count, v, X);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void add_diag_kernel(int n, double* A, int64_t lda, const double* dv, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i + (int64_t)i * lda] += dv ? *dv : v;
}

int add_diag(hipStream_t s, int n, double* A, int64_t lda, const double* dv, double v) {
  if (n <= 0) return 0;
  add_diag_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, A, lda, dv, v);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void scale_cols_kernel(int rows, double* A, int64_t lda, const double* d, int p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= rows) return;
  const double sc = sqrt(d[j]);
  double* a = A + i + (int64_t)j * lda;
  if (p > 0) *a *= sc; else *a /= sc;
}

__global__ void scale_rows_kernel(int rows, double* A, int64_t lda, const double* d, int p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= rows) return;
  const double sc = sqrt(d[i]);
  double* a = A + i + (int64_t)j * lda;
  if (p > 0) *a *= sc; else *a /= sc;
}

int scale_cols(hipStream_t s, int rows, int cols, double* A, int64_t lda, const double* d, int p) {
  if (rows <= 0 || cols <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_SCALE_COLS, rows, cols, p, 0, lda, 0, 0, 0, 0, nullptr, d, nullptr, A};
    o.run = [=] { return scale_cols(s, rows, cols, A, lda, d, p); };
    return rec_add(std::move(o));
  }
  dim3 g(cdiv(rows, 256), cols);
  scale_cols_kernel<<<g, 256, 0, s>>>(rows, A, lda, d, p);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int scale_rows(hipStream_t s, int rows, int cols, double* A, int64_t lda, const double* d, int p) {
  if (rows <= 0 || cols <= 0) return 0;
  if (rec_active()) {
    RecOp o;
    o.kind = REC_EW;
    o.e = EwItem{EW_SCALE_ROWS, rows, cols, p, 0, lda, 0, 0, 0, 0, nullptr, d, nullptr, A};
    o.run = [=] { return scale_rows(s, rows, cols, A, lda, d, p); };
    return rec_add(std::move(o));
  }
  dim3 g(cdiv(rows, 256), cols);
  scale_rows_kernel<<<g, 256, 0, s>>>(rows, A, lda, d, p);
  MSDP_HIP(hipGetLastError());
  return 0;
}


// ==========================================================================
// Batched launches of recorded per-block ops (see Recorder in common.h).
// Each item's arithmetic is that of the single-op kernel it replaces.
// ==========================================================================
__device__ inline int vb_item(const int* __restrict__ start, int count) {
  int lo = 0, hi = count - 1;
  const int w = blockIdx.x;
  while (lo < hi) {  // last item with start <= w
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void ew_vb_kernel(const EwItem* __restrict__ items,
                                                    const int* __restrict__ start, int count) {
  const int k = vb_item(start, count);
  const EwItem it = items[k];
  const int64_t e = (int64_t)(blockIdx.x - start[k]) * 256 + threadIdx.x;
  switch (it.op) {
    case EW_AXPBY:
      if (e < it.count) it.Z[e] = it.a * it.X[e] + it.b * it.Y[e];
      break;
    case EW_AXPBYPCZ:
      if (e < it.count) it.Z[e] = it.a * it.X[e] + it.b * it.Y[e] + it.c * it.V[e];
      break;
    case EW_FILL:
      if (e < it.count) it.Z[e] = it.a;
      break;
    default: {
      const int i = (int)(e % it.rows), j = (int)(e / it.rows);
      if (j >= it.cols) break;
      switch (it.op) {
        case EW_COPY: it.Z[i + (int64_t)j * it.ldb] = it.X[i + (int64_t)j * it.lda]; break;
        case EW_TRANSPOSE: it.Z[j + (int64_t)i * it.ldb] = it.X[i + (int64_t)j * it.lda]; break;
        case EW_SYM:
          if (i > j) {
            double* a = it.Z + i + (int64_t)j * it.lda;
            double* b = it.Z + j + (int64_t)i * it.lda;
            const double v = (*a + *b) / 2;
            *a = v;
            *b = v;
          }
          break;
        case EW_MIRROR:
          if (i < j) it.Z[i + (int64_t)j * it.lda] = it.Z[j + (int64_t)i * it.lda];
          break;
        case EW_ZERO_UPPER:
          if (i < j) it.Z[i + (int64_t)j * it.lda] = 0.0;
          break;
        case EW_SCALE_COLS: {
          const double sc = sqrt(it.Y[j]);
          double* a = it.Z + i + (int64_t)j * it.lda;
          if (it.p > 0) *a *= sc; else *a /= sc;
          break;
        }
        case EW_SCALE_ROWS: {
          const double sc = sqrt(it.Y[i]);
          double* a = it.Z + i + (int64_t)j * it.lda;
          if (it.p > 0) *a *= sc; else *a /= sc;
          break;
        }
        case EW_RNT: {
          const int n = it.rows;
          it.Z[i + (int64_t)j * n] =
              -(it.X[i + (int64_t)j * n] + it.X[j + (int64_t)i * n]) / (it.Y[i] + it.Y[j]);
          break;
        }
        case EW_SCALE_BY_DIAG: {
          const int n = it.rows;
          it.Z[i + (int64_t)j * n] = it.X[i + (int64_t)j * n] * it.Y[j + (int64_t)j * n];
          break;
        }
        case EW_SVEC: {  // entry i of svec(X_j): column c, row r >= c of the lower triangle
          const int64_t n = it.p;
          auto off = [n](int64_t c) { return c * n - c * (c - 1) / 2; };
          const double b2 = 2.0 * n + 1.0;
          int64_t c = (int64_t)((b2 - sqrt(fmax(b2 * b2 - 8.0 * (double)i, 0.0))) * 0.5);
          c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
          while (c + 1 < n && off(c + 1) <= i) ++c;
          while (off(c) > i) --c;
          const int64_t r = c + (i - off(c));
          const double v = it.X[(int64_t)j * n * n + r + c * n];
          it.Z[i + (int64_t)j * it.ldb] = r == c ? v : 1.4142135623730951 * v;
          break;
        }
        default: break;
      }
    }
  }
}

// One workgroup per reduction.  The dot product and the maximum keep the
// order of dot / max_abs (a partial per workgroup of their fixed grid, the
// partials added in index order), so the sums are the same bits.
__global__ __launch_bounds__(256) void red_vb_kernel(const RedItem* __restrict__ items) {
  const RedItem it = items[blockIdx.x];
  __shared__ double sh[4];
  __shared__ double part[RED_BLOCKS];
  const int tid = threadIdx.x;
  if (it.op == RED_SQRTMIN) {
    double v = INFINITY;
    for (int64_t k = tid; k < it.count; k += 256) {
      v = fmin(v, it.X[k]);
      it.D[k] = sqrt(it.X[k]);
    }
    v = block_reduce<2>(v, sh);
    if (tid == 0) *it.out = v;
    return;
  }
  const int g = (int)max((int64_t)1, min((int64_t)RED_BLOCKS, (it.count + 1023) / 1024));
  for (int b = 0; b < g; ++b) {
    double v = 0.0;
    for (int64_t i = (int64_t)b * 256 + tid; i < it.count; i += (int64_t)g * 256) {
      if (it.op == RED_DOT) v += it.X[i] * it.Y[i];
      else v = fmax(v, fabs(it.X[i]));
    }
    v = it.op == RED_DOT ? block_reduce<0>(v, sh) : block_reduce<1>(v, sh);
    if (tid == 0) part[b] = v;
  }
  __syncthreads();
  double v = 0.0;
  for (int i = tid; i < g; i += 256) v = it.op == RED_DOT ? v + part[i] : fmax(v, part[i]);
  v = it.op == RED_DOT ? block_reduce<0>(v, sh) : block_reduce<1>(v, sh);
  if (tid == 0) *it.out = v;
}

static size_t rec_item_bytes(int kind) {
  switch (kind) {
    case REC_EW: return sizeof(EwItem);
    case REC_RED: return sizeof(RedItem);
    case REC_CHOL64: return sizeof(CholItem);
    case REC_CHOL128: return sizeof(CholItem);
    case REC_TRTRI64: return sizeof(TrtriItem);
    case REC_TRTRI128: return sizeof(TrtriItem);
    default: return sizeof(GemmItem);
  }
}

static size_t rec_align(size_t b) { return (b + 255) & ~(size_t)255; }

size_t rec_bytes(int kind, size_t count) {
  return rec_align(count * rec_item_bytes(kind)) + rec_align((count + 1) * sizeof(int));
}

size_t rec_pack(int kind, const std::vector<const RecOp*>& items, char* host, int* grid) {
  const size_t count = items.size();
  const size_t ib = rec_item_bytes(kind);
  int* start = reinterpret_cast<int*>(host + rec_align(count * ib));
  int total = 0;
  for (size_t k = 0; k < count; ++k) {
    const RecOp& o = *items[k];
    int wgs = 1;
    switch (kind) {
      case REC_EW: {
        std::memcpy(host + k * ib, &o.e, ib);
        const int64_t work = (o.e.op == EW_AXPBY || o.e.op == EW_AXPBYPCZ || o.e.op == EW_FILL)
                                 ? o.e.count
                                 : (int64_t)o.e.rows * o.e.cols;
        wgs = (int)std::max<int64_t>(1, (work + 255) / 256);
        break;
      }
      case REC_RED: std::memcpy(host + k * ib, &o.r, ib); break;
      case REC_CHOL64:
      case REC_CHOL128: std::memcpy(host + k * ib, &o.ch, ib); break;
      case REC_TRTRI64:
      case REC_TRTRI128: std::memcpy(host + k * ib, &o.tr, ib); break;
      default:
        std::memcpy(host + k * ib, &o.g, ib);
        wgs = cdiv(o.g.M, GT) * cdiv(o.g.N, GT) * std::max(1, o.g.batch);
        break;
    }
    start[k] = total;
    total += wgs;
  }
  start[count] = total;
  *grid = total;
  return rec_bytes(kind, count);
}

int rec_launch(hipStream_t s, int kind, int count, int grid, const char* dev) {
  if (count <= 0) return 0;
  const size_t ib = rec_item_bytes(kind);
  const int* start = reinterpret_cast<const int*>(dev + rec_align(count * ib));
  switch (kind) {
    case REC_EW:
      ew_vb_kernel<<<grid, 256, 0, s>>>(reinterpret_cast<const EwItem*>(dev), start, count);
      break;
    case REC_RED:
      red_vb_kernel<<<count, 256, 0, s>>>(reinterpret_cast<const RedItem*>(dev));
      break;
    case REC_CHOL64:
      chol_small_vb_kernel<<<count, 64, 0, s>>>(reinterpret_cast<const CholItem*>(dev));
      break;
    case REC_CHOL128:
      chol_mid_vb_kernel<<<count, 128, 0, s>>>(reinterpret_cast<const CholItem*>(dev));
      break;
    case REC_TRTRI64:
      trtri64_vb_kernel<<<count, 64, 0, s>>>(reinterpret_cast<const TrtriItem*>(dev));
      break;
    case REC_TRTRI128:
      trtri_mid_vb_kernel<<<count, 128, 0, s>>>(reinterpret_cast<const TrtriItem*>(dev));
      break;
    case REC_GEMM_NN:
      gemm_vb_kernel<false, false><<<grid, 256, 0, s>>>(reinterpret_cast<const GemmItem*>(dev), start, count);
      break;
    case REC_GEMM_TN:
      gemm_vb_kernel<true, false><<<grid, 256, 0, s>>>(reinterpret_cast<const GemmItem*>(dev), start, count);
      break;
    case REC_GEMM_NT:
      gemm_vb_kernel<false, true><<<grid, 256, 0, s>>>(reinterpret_cast<const GemmItem*>(dev), start, count);
      break;
    case REC_GEMM_TT:
      gemm_vb_kernel<true, true><<<grid, 256, 0, s>>>(reinterpret_cast<const GemmItem*>(dev), start, count);
      break;
    default:
      set_error("rec_launch: not a batchable kind");
      return -2;
  }
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
