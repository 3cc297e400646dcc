// FP32 MFMA GEMM for the mixed-precision Schur assembly: FP32 operands,
// FP32 accumulation (v_mfma_f32_16x16x4f32), output written as FP32 or
// widened to FP64 (C = alpha op(A) op(B) + beta C).  128 x 128 tile per
// workgroup of 4 waves, 64 x 64 per wave, register-prefetched
// double-buffered LDS slabs of K = 16.
//   f32 16x16x4 operand layout (wave64): A lane l holds A[l%16][l/16],
//   B lane l holds B[l/16][l%16]; the output layout is MSDP_F32_DROW below.
#include "common.h"

#include <algorithm>

namespace msdp {

typedef float v4f __attribute__((ext_vector_type(4)));

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// row of D held by lane l, register r
__device__ inline int f32_drow(int lk, int r) { return 4 * lk + r; }

template <bool TA, bool TB, typename OutT>
__global__ __launch_bounds__(256, 2) void gemm128_f32_kernel(
    int M, int N, int K, float alpha, const float* __restrict__ A, int64_t lda, int64_t sA,
    const float* __restrict__ B, int64_t ldb, int64_t sB, double beta, OutT* __restrict__ C,
    int64_t ldc, int64_t sC, int lower) {
  int tm = blockIdx.x, tn = blockIdx.y;
  if (lower == 2) {  // square lower triangle on a 1-D, XCD-grouped grid
    if (!lower_tile_xcd((M + 127) / 128, tm, tn)) return;
  } else if (lower && tn > tm) {
    return;
  }
  const int64_t bz = blockIdx.z;
  A += bz * sA;
  B += bz * sB;
  C += bz * sC;
  constexpr int BT = 128, BK = 16;
  __shared__ float As[2][BK][BT + 4];
  __shared__ float Bs[2][BK][BT + 4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = tm * BT, n0 = tn * BT;
  float ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + r * 256;
      int mm, kk;
      if (!TA) { mm = e & 127; kk = e >> 7; } else { kk = e & 15; mm = e >> 4; }
      const int gm = m0 + mm, gk = k0 + kk;
      ra[r] = (gm < M && gk < K) ? (TA ? A[gk + (int64_t)gm * lda] : A[gm + (int64_t)gk * lda]) : 0.f;
      int nn;
      if (!TB) { kk = e & 15; nn = e >> 4; } else { nn = e & 127; kk = e >> 7; }
      const int gn = n0 + nn, gk2 = k0 + kk;
      rb[r] = (gn < N && gk2 < K) ? (TB ? B[gn + (int64_t)gk2 * ldb] : B[gk2 + (int64_t)gn * ldb]) : 0.f;
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
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
  const int nk = (K + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = As[buf][kk + lk][wm + 16 * q + li];
        bv[q] = Bs[buf][kk + lk][wn + 16 * q + li];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (kt + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gn = n0 + wn + 16 * b + li;
      if (gn >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * a + f32_drow(lk, r);
        if (gm >= M || (lower && gm < gn)) continue;
        OutT* c = C + gm + (int64_t)gn * ldc;
        const double v = (double)(alpha * acc[a][b][r]);
        *c = (OutT)(beta == 0.0 ? v : v + beta * (double)(*c));
      }
    }
  }
}

template <typename OutT>
static int gemm32_t(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha,
                    const float* A, int64_t lda, const float* B, int64_t ldb, double beta,
                    OutT* C, int64_t ldc, int batch, int64_t sA, int64_t sB, int64_t sC,
                    bool lower) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return 0;
  dim3 g(cdiv(M, 128), cdiv(N, 128), batch);
  int lmode = lower ? 1 : 0;
  if (lower && batch == 1 && M == N) {
    g = dim3(lower_tile_blocks(cdiv(M, 128)), 1, 1);
    lmode = 2;
  }
  if (!ta && !tb)
    gemm128_f32_kernel<false, false, OutT><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode);
  else if (ta && !tb)
    gemm128_f32_kernel<true, false, OutT><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode);
  else if (!ta && tb)
    gemm128_f32_kernel<false, true, OutT><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode);
  else
    gemm128_f32_kernel<true, true, OutT><<<g, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lmode);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int gemm32(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha, const float* A,
           int64_t lda, const float* B, int64_t ldb, double beta, float* C, int64_t ldc,
           int batch, int64_t sA, int64_t sB, int64_t sC, bool lower) {
  const bool prof = prof_active(PROF_GEMM32, s);
  if (prof) prof_before(s);
  MSDP_TRY(gemm32_t<float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB, sC, lower));
  if (prof) prof_after(s, 2.0 * M * N * (double)K * batch * (lower ? 0.5 : 1.0));
  return 0;
}

int gemm32_to64(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha,
                const float* A, int64_t lda, const float* B, int64_t ldb, double beta,
                double* C, int64_t ldc, bool lower) {
  const bool prof = prof_active(PROF_GEMM32, s);
  if (prof) prof_before(s);
  MSDP_TRY(gemm32_t<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, 0, 0, 0, lower));
  if (prof) prof_after(s, 2.0 * M * N * (double)K * (lower ? 0.5 : 1.0));
  return 0;
}

__global__ void d2f_kernel(int64_t count, const double* x, float* y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

int d2f(hipStream_t s, int64_t count, const double* x, float* y) {
  if (count <= 0) return 0;
  d2f_kernel<<<(int)std::min<int64_t>(8192, (count + 255) / 256), 256, 0, s>>>(count, x, y);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// svec pack of a batch of FP32 matrices (sqrt(2) off the diagonal)
__global__ void svec_pack32_kernel(int n, int cb, const float* P, float* B, int64_t ldb) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= d || j >= cb) return;
  int c = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
  while ((int64_t)(c + 1) * (c + 2) / 2 <= k) ++c;
  while ((int64_t)c * (c + 1) / 2 > k) --c;
  const int r = (int)(k - (int64_t)c * (c + 1) / 2);
  const float v = P[(int64_t)j * n * n + r + (int64_t)c * n];
  B[k + (int64_t)j * ldb] = r == c ? v : 1.41421356f * v;
}

int svec_pack32(hipStream_t s, int n, int cb, const float* P, float* B, int64_t ldb) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  svec_pack32_kernel<<<dim3(cdiv(d, 256), cb), 256, 0, s>>>(n, cb, P, B, ldb);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
