// Dense FP64 building blocks on gfx950: MFMA GEMM, blocked Cholesky,
// triangular inverse, the Cholesky flush and elementwise/reduction kernels.
//
// These replace the LAPACK/BLAS calls of the reference: potrf
// (src/blocks.jl:97, src/schur_solvers.jl:17), potri (src/blocks.jl:118),
// trmm/trsm/gemm/syrk (src/directions.jl:50-98, src/step.jl:160-162).
#include "common.h"

#include <algorithm>
#include <cstdio>
#include <mutex>

namespace msdp {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
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

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda, int64_t sA,
    const double* __restrict__ B, int64_t ldb, int64_t sB, double beta, double* __restrict__ C,
    int64_t ldc, int64_t sC, int lower, const int* __restrict__ guard) {
  if (guard && *guard != 0) return;
  const int tm = blockIdx.x, tn = blockIdx.y;
  if (lower && tn > tm) return;
  const int64_t bz = blockIdx.z;
  A += bz * sA;
  B += bz * sB;
  C += bz * sC;
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  v4d acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
  const int m0 = tm * GT, n0 = tn * GT;
  for (int k0 = 0; k0 < K; k0 += GK) {
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
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i0 = m0 + wm + 16 * (q >> 1) + lk;
    const int gn = n0 + wn + 16 * (q & 1) + li;
    if (gn >= N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gm = i0 + 4 * r;
      if (gm >= M || (lower && gm < gn)) continue;
      double* c = C + gm + (int64_t)gn * ldc;
      const double v = alpha * (*accs[q])[r];
      *c = beta == 0.0 ? v : v + beta * (*c);
    }
  }
}

static int gemm_guarded(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc, int batch, int64_t sA, int64_t sB, int64_t sC,
                        bool lower, const int* guard) {
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
  dim3 grid(cdiv(M, GT), cdiv(N, GT), batch);
  const bool prof = prof_active(PROF_GEMM, s);
  if (prof) prof_before(s);
  if (!ta && !tb)
    gemm_f64_kernel<false, false><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard);
  else if (ta && !tb)
    gemm_f64_kernel<true, false><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard);
  else if (!ta && tb)
    gemm_f64_kernel<false, true><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard);
  else
    gemm_f64_kernel<true, true><<<grid, 256, 0, s>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, lower, guard);
  if (prof) prof_after(s, 2.0 * M * N * (double)K * batch * (lower ? 0.5 : 1.0));
  MSDP_HIP(hipGetLastError());
  return 0;
}

int gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha, const double* A,
         int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
         int batch, int64_t sA, int64_t sB, int64_t sC, bool lower) {
  return gemm_guarded(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB, sC,
                      lower, nullptr);
}

// ==========================================================================
// Cholesky
// ==========================================================================
// Factor one diagonal block (n <= 64) in LDS; one workgroup per batch entry.
// info[b] gets info_base + (1-based failing column) on failure.
__global__ __launch_bounds__(256) void potrf_small_kernel(int n, double* A, int64_t lda,
                                                          int64_t stride, int* info,
                                                          int info_base, int per_batch_info) {
  int* inf = info + (per_batch_info ? blockIdx.x : 0);
  if (!per_batch_info && *inf != 0) return;
  A += (int64_t)blockIdx.x * stride;
  __shared__ double Ls[64][65];
  __shared__ int fail;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += 256) {
    const int i = e % n, j = e / n;
    Ls[i][j] = i >= j ? A[i + (int64_t)j * lda] : 0.0;
  }
  if (tid == 0) fail = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      const double d = Ls[j][j];
      if (!(d > 0.0) || !isfinite(d)) fail = j + 1;
      else Ls[j][j] = sqrt(d);
    }
    __syncthreads();
    if (fail) break;
    const double djj = Ls[j][j];
    for (int i = j + 1 + tid; i < n; i += 256) Ls[i][j] /= djj;
    __syncthreads();
    const int r = n - j - 1;
    for (int e = tid; e < r * r; e += 256) {
      const int i = j + 1 + e % r, k = j + 1 + e / r;
      if (k <= i) Ls[i][k] -= Ls[i][j] * Ls[k][j];
    }
    __syncthreads();
  }
  if (fail) {
    if (tid == 0) atomicCAS(inf, 0, info_base + fail);
    return;
  }
  for (int e = tid; e < n * n; e += 256) {
    const int i = e % n, j = e / n;
    if (i >= j) A[i + (int64_t)j * lda] = Ls[i][j];
  }
}

// A21 <- A21 * L11^-T for a 64-wide panel; one thread per row.
__global__ __launch_bounds__(128) void trsm_panel_kernel(int rows, double* A21, int64_t lda,
                                                         const double* L11, const int* info) {
  if (*info != 0) return;
  __shared__ double Ls[64][65];
  __shared__ double dinv[64];
  for (int e = threadIdx.x; e < 64 * 64; e += 128) {
    const int i = e & 63, j = e >> 6;
    Ls[i][j] = i >= j ? L11[i + (int64_t)j * lda] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < 64) dinv[threadIdx.x] = 1.0 / Ls[threadIdx.x][threadIdx.x];
  __syncthreads();
  const int r = blockIdx.x * 128 + threadIdx.x;
  if (r >= rows) return;
  double x[64];
#pragma unroll
  for (int c = 0; c < 64; ++c) x[c] = A21[r + (int64_t)c * lda];
#pragma unroll
  for (int c = 0; c < 64; ++c) {
    double v = x[c];
#pragma unroll
    for (int k = 0; k < c; ++k) v -= x[k] * Ls[c][k];
    x[c] = v * dinv[c];
  }
#pragma unroll
  for (int c = 0; c < 64; ++c) A21[r + (int64_t)c * lda] = x[c];
}

size_t potrf_work_elems(int) { return 0; }

int potrf(hipStream_t s, int n, double* A, int64_t lda, int* info, double*) {
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int jb = std::min(64, n - j0);
    double* A11 = A + j0 + (int64_t)j0 * lda;
    potrf_small_kernel<<<1, 256, 0, s>>>(jb, A11, lda, 0, info, j0, 0);
    MSDP_HIP(hipGetLastError());
    const int rest = n - j0 - jb;
    if (rest > 0) {
      double* A21 = A11 + jb;
      trsm_panel_kernel<<<cdiv(rest, 128), 128, 0, s>>>(rest, A21, lda, A11, info);
      MSDP_HIP(hipGetLastError());
      double* A22 = A21 + (int64_t)jb * lda;
      MSDP_TRY(gemm_guarded(s, false, true, rest, rest, jb, -1.0, A21, lda, A21, lda, 1.0, A22, lda,
                            1, 0, 0, 0, true, info));
    }
  }
  return 0;
}

int potrf_small_batched(hipStream_t s, int n, double* A, int64_t lda, int64_t stride, int batch,
                        int* info) {
  if (n > 64) { set_error("potrf_small_batched: n > 64"); return -2; }
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * batch, s));
  potrf_small_kernel<<<batch, 256, 0, s>>>(n, A, lda, stride, info, 0, 1);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// Triangular inverse (lower), recursive: diagonal 64-blocks inverted in LDS,
// off-diagonal blocks X21 = -L22^-1 L21 L11^-1 by two GEMMs.
// ==========================================================================
__global__ __launch_bounds__(64) void trtri_diag_kernel(int n, const double* L, int64_t ldl,
                                                        double* Li, int64_t ldi) {
  const int b0 = blockIdx.x * 64;
  const int nb = min(64, n - b0);
  __shared__ double T[64][65];
  __shared__ double X[64][65];
  const int c = threadIdx.x;
  for (int e = c; e < 64 * 64; e += 64) {
    const int i = e & 63, j = e >> 6;
    T[i][j] = (i < nb && j < nb && i >= j) ? L[(b0 + i) + (int64_t)(b0 + j) * ldl] : 0.0;
  }
  __syncthreads();
  if (c < nb) {
    // column c of T^-1 by forward substitution
    for (int i = 0; i < c; ++i) X[i][c] = 0.0;
    X[c][c] = 1.0 / T[c][c];
    for (int i = c + 1; i < nb; ++i) {
      double v = 0.0;
      for (int k = c; k < i; ++k) v -= T[i][k] * X[k][c];
      X[i][c] = v / T[i][i];
    }
  }
  __syncthreads();
  for (int e = c; e < nb * nb; e += 64) {
    const int i = e % nb, j = e / nb;
    Li[(b0 + i) + (int64_t)(b0 + j) * ldi] = X[i][j];
  }
}

__global__ void zero_upper_kernel(int n, double* A, int64_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < n && i < j) A[i + (int64_t)j * lda] = 0.0;
}

static int trtri_rec(hipStream_t s, int lo, int hi, const double* L, int64_t ldl, double* Li,
                     int64_t ldi, double* work) {
  const int nblk = cdiv(hi - lo, 64);
  if (nblk <= 1) return 0;
  const int mid = lo + 64 * ((nblk + 1) / 2);
  MSDP_TRY(trtri_rec(s, lo, mid, L, ldl, Li, ldi, work));
  MSDP_TRY(trtri_rec(s, mid, hi, L, ldl, Li, ldi, work));
  const int r = hi - mid, c = mid - lo;
  // work = L21 * L11inv   (r x c)
  MSDP_TRY(gemm(s, false, false, r, c, c, 1.0, L + mid + (int64_t)lo * ldl, ldl,
                Li + lo + (int64_t)lo * ldi, ldi, 0.0, work, r));
  // X21 = -L22inv * work
  MSDP_TRY(gemm(s, false, false, r, c, r, -1.0, Li + mid + (int64_t)mid * ldi, ldi, work, r, 0.0,
                Li + mid + (int64_t)lo * ldi, ldi));
  return 0;
}

size_t trtri_work_elems(int n) { return (size_t)(n / 2 + 64) * (size_t)(n / 2 + 64); }

int trtri_lower(hipStream_t s, int n, const double* L, int64_t ldl, double* Li, int64_t ldi,
                double* work) {
  if (n <= 0) return 0;
  dim3 g(cdiv(n, 256), n);
  zero_upper_kernel<<<g, 256, 0, s>>>(n, Li, ldi);
  MSDP_HIP(hipGetLastError());
  trtri_diag_kernel<<<cdiv(n, 64), 64, 0, s>>>(n, L, ldl, Li, ldi);
  MSDP_HIP(hipGetLastError());
  return trtri_rec(s, 0, n, L, ldl, Li, ldi, work);
}

// ==========================================================================
// reductions (deterministic: fixed grid, two passes)
// ==========================================================================
constexpr int RED_BLOCKS = 512;

template <int OP>
__global__ __launch_bounds__(256) void reduce_pass1(int64_t count, const double* X,
                                                    const double* Y, double* partial) {
  __shared__ double sh[4];
  double v = OP == 0 ? 0.0 : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256) {
    if (OP == 0) v += X[i] * Y[i];
    else v = fmax(v, fabs(X[i]));
  }
  v = block_reduce<OP == 0 ? 0 : 1>(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

template <int OP>
__global__ __launch_bounds__(256) void reduce_pass2(int nparts, const double* partial,
                                                    double* out) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) v = OP == 0 ? v + partial[i] : fmax(v, partial[i]);
  v = block_reduce<OP == 0 ? 0 : 1>(v, sh);
  if (threadIdx.x == 0) *out = v;
}

size_t reduce_scratch_elems(int64_t) { return RED_BLOCKS; }

int dot(hipStream_t s, int64_t count, const double* X, const double* Y, double* out,
        double* scratch) {
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (count + 1023) / 1024));
  reduce_pass1<0><<<g, 256, 0, s>>>(count, X, Y, scratch);
  reduce_pass2<0><<<1, 256, 0, s>>>(g, scratch, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int max_abs(hipStream_t s, int64_t count, const double* X, double* out, double* scratch) {
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (count + 1023) / 1024));
  reduce_pass1<1><<<g, 256, 0, s>>>(count, X, nullptr, scratch);
  reduce_pass2<1><<<1, 256, 0, s>>>(g, scratch, out);
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
  if (n <= 0) return 0;
  diag_maxabs_kernel<<<1, 256, 0, s>>>(n, L, ld, scratch);
  dim3 g(cdiv(n, 256), n);
  flush_kernel<<<g, 256, 0, s>>>(n, L, ld, scratch);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// elementwise
// ==========================================================================
__global__ void symmetrize_kernel(int n, double* A, int64_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n || i <= j) return;  // each pair (i > j) handled once
  double* a = A + i + (int64_t)j * lda;
  double* b = A + j + (int64_t)i * lda;
  const double v = (*a + *b) / 2;
  *a = v;
  *b = v;
}

int symmetrize(hipStream_t s, int n, double* A, int64_t lda) {
  if (n <= 1) return 0;
  dim3 g(cdiv(n, 256), n);
  symmetrize_kernel<<<g, 256, 0, s>>>(n, A, lda);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void mirror_lower_kernel(int n, double* A, int64_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n || i >= j) return;
  A[i + (int64_t)j * lda] = A[j + (int64_t)i * lda];
}

int mirror_lower(hipStream_t s, int n, double* A, int64_t lda) {
  if (n <= 1) return 0;
  dim3 g(cdiv(n, 256), n);
  mirror_lower_kernel<<<g, 256, 0, s>>>(n, A, lda);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ void copy_kernel(int rows, const double* A, int64_t lda, double* B, int64_t ldb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < rows) B[i + (int64_t)j * ldb] = A[i + (int64_t)j * lda];
}

int copy_mat(hipStream_t s, int rows, int cols, const double* A, int64_t lda, double* B,
             int64_t ldb) {
  if (rows <= 0 || cols <= 0) return 0;
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
  axpby_kernel<<<ew_grid(count), 256, 0, s>>>(count, a, X, b, Y, Z);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int axpbypcz(hipStream_t s, int64_t count, double a, const double* X, double b, const double* Y,
             double c, const double* V, double* Z) {
  if (count <= 0) return 0;
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

int fill(hipStream_t s, int64_t count, double v, double* X) {
  if (count <= 0) return 0;
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
  dim3 g(cdiv(rows, 256), cols);
  scale_cols_kernel<<<g, 256, 0, s>>>(rows, A, lda, d, p);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int scale_rows(hipStream_t s, int rows, int cols, double* A, int64_t lda, const double* d, int p) {
  if (rows <= 0 || cols <= 0) return 0;
  dim3 g(cdiv(rows, 256), cols);
  scale_rows_kernel<<<g, 256, 0, s>>>(rows, A, lda, d, p);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
