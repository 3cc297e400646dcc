// Symmetric indefinite LDL' factorisation with Bunch-Kaufman pivoting
// (Bunch & Kaufman 1977, alpha = (1 + sqrt 17) / 8) and its solve, on the
// device: the factorisation behind `MadNLPSchur`, which factors the Schur
// complement with MadNLP's LAPACK solver in BUNCHKAUFMAN mode
// (src/kkt.jl:7-32, src/schur_solvers.jl:23-39, src/components.jl:275-283)
// so that a mildly indefinite H near the end of the central path still
// factors.
//
// Storage follows the LAPACK convention of the lower triangle: on exit A
// holds L (unit lower, below the diagonal) and the 1x1 / 2x2 diagonal blocks
// D; ipiv[k] = p + 1 for a 1x1 pivot at k interchanged with p, and
// ipiv[k] = ipiv[k+1] = -(p + 1) for a 2x2 pivot at (k, k+1) whose second row
// was interchanged with p.  info = 1-based index of the first exactly zero
// pivot (the reference's SingularException), 0 otherwise.
//
// One pivot step is two launches: `bk_pivot_kernel` (one workgroup: the
// column and row maxima, the pivot choice, the symmetric interchange, the
// multipliers) and `bk_update_kernel` (the rank-1 / rank-2 update of the
// trailing lower triangle, 64 x 64 tiles over the whole GPU).  The column
// reached is kept on the device, so the host issues n pairs without reading
// anything back; pairs past the end (2x2 pivots advance by two) return at once.
#include "common.h"

#include <algorithm>

namespace msdp {

struct BkCtl {
  int k;       // next column to eliminate
  int cur;     // column of the step just taken
  int kstep;   // 1 or 2 (0: zero pivot, no update)
  int info;
};

constexpr int BK_NT = 1024;

__device__ inline void bk_argmax(double v, int idx, double* sv, int* si) {
  // block-wide max |v| with the smallest index among ties (LAPACK's idamax)
  const int tid = threadIdx.x;
  sv[tid] = v;
  si[tid] = idx;
  __syncthreads();
  for (int off = BK_NT / 2; off > 0; off >>= 1) {
    if (tid < off) {
      const double a = sv[tid], b = sv[tid + off];
      const int ia = si[tid], ib = si[tid + off];
      if (b > a || (b == a && ib < ia)) {
        sv[tid] = b;
        si[tid] = ib;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(BK_NT) void bk_pivot_kernel(int n, double* A, int64_t lda, int* ipiv,
                                                          BkCtl* ctl, double* xbuf) {
  __shared__ double sv[BK_NT];
  __shared__ int si[BK_NT];
  __shared__ int s_kp, s_kstep;
  const int tid = threadIdx.x;
  const int k = ctl->k;
  if (k >= n) return;
  auto a = [&](int i, int j) -> double& { return A[i + (int64_t)j * lda]; };
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  // column maximum below the diagonal
  double v = -1.0;
  int idx = n;
  for (int i = k + 1 + tid; i < n; i += BK_NT) {
    const double t = fabs(a(i, k));
    if (t > v) { v = t; idx = i; }
  }
  bk_argmax(v, idx, sv, si);
  const double colmax = k + 1 < n ? sv[0] : 0.0;
  const int imax = si[0];
  __syncthreads();
  const double absakk = fabs(a(k, k));
  if (tid == 0) {
    s_kstep = 1;
    s_kp = k;
  }
  if (fmax(absakk, colmax) == 0.0) {
    // zero column: no pivot, no update (info records the first one)
    if (tid == 0) {
      if (ctl->info == 0) ctl->info = k + 1;
      ipiv[k] = k + 1;
      ctl->cur = k;
      ctl->kstep = 0;
      ctl->k = k + 1;
    }
    return;
  }
  if (absakk < alpha * colmax) {
    // largest off-diagonal entry of row imax in the trailing matrix
    double r = -1.0;
    for (int j = k + tid; j < n; j += BK_NT) {
      if (j == imax) continue;
      const double t = j < imax ? fabs(a(imax, j)) : fabs(a(j, imax));
      r = fmax(r, t);
    }
    bk_argmax(r, 0, sv, si);
    const double rowmax = sv[0];
    if (tid == 0) {
      if (absakk >= alpha * colmax * (colmax / rowmax)) {
        s_kp = k;
      } else if (fabs(a(imax, imax)) >= alpha * rowmax) {
        s_kp = imax;
      } else {
        s_kp = imax;
        s_kstep = 2;
      }
    }
  }
  __syncthreads();
  const int kp = s_kp, kstep = s_kstep;
  const int kk = k + kstep - 1;
  if (kp != kk) {
    // symmetric interchange of rows / columns kk and kp of the trailing matrix
    for (int i = kp + 1 + tid; i < n; i += BK_NT) {
      const double t = a(i, kk);
      a(i, kk) = a(i, kp);
      a(i, kp) = t;
    }
    for (int j = kk + 1 + tid; j < kp; j += BK_NT) {
      const double t = a(j, kk);
      a(j, kk) = a(kp, j);
      a(kp, j) = t;
    }
    if (tid == 0) {
      const double t = a(kk, kk);
      a(kk, kk) = a(kp, kp);
      a(kp, kp) = t;
      if (kstep == 2) {
        const double u = a(k + 1, k);
        a(k + 1, k) = a(kp, k);
        a(kp, k) = u;
      }
    }
  }
  __syncthreads();
  if (kstep == 1) {
    // x = A(k+1:n, k) kept for the update; L(:, k) = x / d
    const double r1 = 1.0 / a(k, k);
    for (int i = k + 1 + tid; i < n; i += BK_NT) {
      const double x = a(i, k);
      xbuf[i] = x;
      a(i, k) = x * r1;
    }
  } else {
    // D = [a b; b c] at (k, k+1): L(:, k:k+1) = [x y] D^-1
    const double d21 = a(k + 1, k);
    const double d11 = a(k + 1, k + 1) / d21;
    const double d22 = a(k, k) / d21;
    const double t = 1.0 / (d11 * d22 - 1.0);
    const double s21 = t / d21;
    for (int j = k + 2 + tid; j < n; j += BK_NT) {
      const double x = a(j, k), y = a(j, k + 1);
      xbuf[j] = x;
      xbuf[n + j] = y;
      a(j, k) = s21 * (d11 * x - y);
      a(j, k + 1) = s21 * (d22 * y - x);
    }
  }
  if (tid == 0) {
    if (kstep == 1) {
      ipiv[k] = kp + 1;
    } else {
      ipiv[k] = -(kp + 1);
      ipiv[k + 1] = -(kp + 1);
    }
    ctl->cur = k;
    ctl->kstep = kstep;
    ctl->k = k + kstep;
  }
}

// A(i, j) -= x_i L(j, k) (+ y_i L(j, k+1)) for j0 <= j <= i, j0 = cur + kstep,
// over 64 x 64 tiles of the trailing lower triangle; the grid covers the
// largest trailing block this launch can meet and surplus tiles return.
__global__ __launch_bounds__(256) void bk_update_kernel(int n, double* A, int64_t lda,
                                                        const BkCtl* ctl, const double* xbuf) {
  const int kstep = ctl->kstep;
  if (kstep == 0) return;
  const int c = ctl->cur;
  const int j0 = c + kstep;
  if (j0 >= n) return;
  const int nt = (n - j0 + 63) / 64;
  // blockIdx.x -> lower tile (tm >= tn) of the trailing block
  int t = blockIdx.x, tn = 0;
  if (t >= nt * (nt + 1) / 2) return;
  while (t >= nt - tn) {
    t -= nt - tn;
    ++tn;
  }
  const int tm = tn + t;
  const int i0 = j0 + tm * 64, jb = j0 + tn * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int i = i0 + (e & 63), j = jb + (e >> 6);
    if (i >= n || j >= n || i < j) continue;
    double* aij = A + i + (int64_t)j * lda;
    if (kstep == 1) {
      *aij -= xbuf[i] * A[j + (int64_t)c * lda];
    } else {
      *aij -= xbuf[i] * A[j + (int64_t)c * lda] + xbuf[n + i] * A[j + (int64_t)(c + 1) * lda];
    }
  }
}

__global__ void bk_init_kernel(BkCtl* ctl) {
  ctl->k = 0;
  ctl->cur = 0;
  ctl->kstep = 0;
  ctl->info = 0;
}

size_t sytrf_bk_ctl_bytes() { return 256; }

// LDL' of the lower triangle of A (n x n) in place.  work: 2n doubles;
// ctl: sytrf_bk_ctl_bytes() of device memory; *info_dev as described above.
int sytrf_bk(hipStream_t s, int n, double* A, int64_t lda, int* ipiv, void* ctl_mem,
             double* work, int* info_dev) {
  BkCtl* ctl = static_cast<BkCtl*>(ctl_mem);
  bk_init_kernel<<<1, 1, 0, s>>>(ctl);
  for (int step = 0; step < n; ++step) {
    bk_pivot_kernel<<<1, BK_NT, 0, s>>>(n, A, lda, ipiv, ctl, work);
    // the trailing block after this launch's pivot is at most n - step - 1 wide
    const int nt = (n - step - 1 + 63) / 64;
    if (nt > 0) bk_update_kernel<<<nt * (nt + 1) / 2, 256, 0, s>>>(n, A, lda, ctl, work);
  }
  MSDP_HIP(hipMemcpyAsync(info_dev, &ctl->info, sizeof(int), hipMemcpyDeviceToDevice, s));
  MSDP_HIP(hipGetLastError());
  return 0;
}

// x = A^-1 b with the factorisation above (one workgroup): P, L, D, L', P'.
__global__ __launch_bounds__(BK_NT) void bk_solve_kernel(int n, const double* A, int64_t lda,
                                                         const int* ipiv, double* b) {
  __shared__ double red[BK_NT / 64];
  __shared__ double s_bk[2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto a = [&](int i, int j) { return A[i + (int64_t)j * lda]; };
  auto swap_b = [&](int i, int p) {
    if (tid == 0 && i != p) {
      const double t = b[i];
      b[i] = b[p];
      b[p] = t;
    }
  };
  auto block_sum = [&](double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0) {
      for (int w = 0; w < BK_NT / 64; ++w) s += red[w];
    }
    return s;
  };
  // forward: L D y = P b
  for (int k = 0; k < n;) {
    if (ipiv[k] > 0) {
      swap_b(k, ipiv[k] - 1);
      __syncthreads();
      const double bk = b[k];
      for (int i = k + 1 + tid; i < n; i += BK_NT) b[i] -= a(i, k) * bk;
      __syncthreads();
      if (tid == 0) b[k] = bk / a(k, k);
      __syncthreads();
      k += 1;
    } else {
      swap_b(k + 1, -ipiv[k] - 1);
      __syncthreads();
      const double b0 = b[k], b1 = b[k + 1];
      for (int i = k + 2 + tid; i < n; i += BK_NT) b[i] -= a(i, k) * b0 + a(i, k + 1) * b1;
      __syncthreads();
      if (tid == 0) {
        const double d21 = a(k + 1, k);
        const double d11 = a(k, k) / d21, d22 = a(k + 1, k + 1) / d21;
        const double den = d11 * d22 - 1.0;
        const double u = b0 / d21, w = b1 / d21;
        b[k] = (d22 * u - w) / den;
        b[k + 1] = (d11 * w - u) / den;
      }
      __syncthreads();
      k += 2;
    }
  }
  // backward: L' P' x = y
  for (int k = n - 1; k >= 0;) {
    const bool two = ipiv[k] < 0;
    const int k0 = two ? k - 1 : k;
    double s0 = 0.0, s1 = 0.0;
    for (int i = k + 1 + tid; i < n; i += BK_NT) {
      s0 += a(i, k) * b[i];
      if (two) s1 += a(i, k0) * b[i];
    }
    s0 = block_sum(s0);
    __syncthreads();
    if (two) s1 = block_sum(s1);
    if (tid == 0) {
      s_bk[0] = s0;
      s_bk[1] = s1;
    }
    __syncthreads();
    if (tid == 0) {
      b[k] -= s_bk[0];
      if (two) b[k0] -= s_bk[1];
      const int p = two ? -ipiv[k] - 1 : ipiv[k] - 1;
      if (p != k) {
        const double t = b[k];
        b[k] = b[p];
        b[p] = t;
      }
    }
    __syncthreads();
    k = k0 - 1;
  }
}

int sytrs_bk(hipStream_t s, int n, const double* A, int64_t lda, const int* ipiv, double* b) {
  if (n <= 0) return 0;
  bk_solve_kernel<<<1, BK_NT, 0, s>>>(n, A, lda, ipiv, b);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
