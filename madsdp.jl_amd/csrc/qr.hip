// QR Schur solver (reference src/qr_schur.jl): the NT-scaled constraint
// matrix B[:, j] = [svec(G' A_j G) per block; sqrt(x ./ z) .* C_lin[j, :]],
// column-equilibrated, Householder QR, and the projection solve of the
// Newton system.  The reference factors B with SuiteSparse SPQR; here B is
// dense in HBM and factored column by column (one workgroup forms each
// reflector, one workgroup per trailing column applies it).
#include "qr.h"

#include <algorithm>

namespace msdp {

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
constexpr double QEPS = 2.220446049250313e-16;

// svec index of (r, c), r <= c, column-major upper order (qr_schur.jl:46)
__device__ inline void svec_rc(int64_t k, int& r, int& c) {
  int cc = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
  while ((int64_t)(cc + 1) * (cc + 2) / 2 <= k) ++cc;
  while ((int64_t)cc * (cc + 1) / 2 > k) --cc;
  c = cc;
  r = (int)(k - (int64_t)cc * (cc + 1) / 2);
}

// B[off + k, con[j]] = w_k P_j[r, c] for P_j n x n (batch of c_cnt)
__global__ void qr_pack_kernel(int n, int cnt, const double* P, const int* con, double* B,
                               int64_t ldb, int64_t off) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= d || j >= cnt) return;
  int r, c;
  svec_rc(k, r, c);
  const double v = P[(int64_t)j * n * n + r + (int64_t)c * n];
  B[off + k + (int64_t)con[j] * ldb] = r == c ? v : 1.4142135623730951 * v;
}

// COO constraints: (G' A_j G)[a, b] = sum_e val_e G[row_e, a] G[col_e, b]
__global__ void qr_coo_kernel(int n, int ncon, const int* con_list, const int* con_ptr,
                              const int* row, const int* col, const double* val,
                              const double* G, double* B, int64_t ldb, int64_t off) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= d || j >= ncon) return;
  int a, b;
  svec_rc(k, a, b);
  double s = 0.0;
  for (int e = con_ptr[j]; e < con_ptr[j + 1]; ++e)
    s += val[e] * G[row[e] + (int64_t)a * n] * G[col[e] + (int64_t)b * n];
  B[off + k + (int64_t)con_list[j] * ldb] = a == b ? s : 1.4142135623730951 * s;
}

// scalar rows: B[off + l, j] = sqrt(x_l / z_l) C_lin[j, l]
__global__ void qr_scalar_rows_kernel(int m, int ns, const double* Clin, const double* x,
                                      const double* z, double* B, int64_t ldb, int64_t off) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (l >= ns) return;
  B[off + l + (int64_t)j * ldb] = sqrt(x[l] / z[l]) * Clin[j + (int64_t)l * m];
}

// column equilibration: scale_j = 1/||B[:, j]||, flag if zero or not finite
__global__ __launch_bounds__(256) void qr_colscale_kernel(int64_t rows, double* B, int64_t ldb,
                                                          double* scale, int* bad) {
  __shared__ double red[4];
  double* c = B + (int64_t)blockIdx.x * ldb;
  double ss = 0.0;
  for (int64_t i = threadIdx.x; i < rows; i += 256) ss += c[i] * c[i];
  ss = block_reduce<0>(ss, red);
  const double mag = sqrt(ss);
  const bool ok = mag != 0.0 && isfinite(mag);
  if (threadIdx.x == 0) {
    scale[blockIdx.x] = ok ? 1.0 / mag : 1.0;
    if (!ok) atomicExch(bad, 1);
  }
  const double s = ok ? 1.0 / mag : 1.0;
  for (int64_t i = threadIdx.x; i < rows; i += 256) c[i] *= s;
}

// reflector of column k: B[k:, k] -> beta e_1, v stored below the diagonal
// (v_k = 1 implicit), tau[k]; R_kk = beta stays on the diagonal.
__global__ __launch_bounds__(1024) void qr_reflector_kernel(int64_t rows, int k, double* B,
                                                            int64_t ldb, double* tau) {
  __shared__ double red[16];
  __shared__ double s_sc;
  double* c = B + (int64_t)k * ldb;
  double ss = 0.0;
  for (int64_t i = k + 1 + threadIdx.x; i < rows; i += blockDim.x) ss += c[i] * c[i];
  ss = block_reduce<0>(ss, red);
  if (threadIdx.x == 0) {
    const double alpha = c[k];
    double beta, t, sc;
    if (ss == 0.0) {
      beta = alpha; t = 0.0; sc = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + ss), alpha);
      t = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    tau[k] = t;
    c[k] = beta;
    s_sc = sc;
  }
  __syncthreads();
  const double sc = s_sc;
  for (int64_t i = k + 1 + threadIdx.x; i < rows; i += blockDim.x) c[i] *= sc;
}

// apply H_k = I - tau v v' to the columns j > k, one workgroup per column
__global__ __launch_bounds__(256) void qr_apply_kernel(int64_t rows, int k, int m, double* B,
                                                       int64_t ldb, const double* tau) {
  __shared__ double red[4];
  const int j = k + 1 + blockIdx.x;
  if (j >= m) return;
  const double* v = B + (int64_t)k * ldb;
  double* c = B + (int64_t)j * ldb;
  double s = 0.0;
  for (int64_t i = k + threadIdx.x; i < rows; i += 256) s += (i == k ? 1.0 : v[i]) * c[i];
  s = block_reduce<0>(s, red);
  const double f = tau[k] * s;
  for (int64_t i = k + threadIdx.x; i < rows; i += 256) c[i] -= f * (i == k ? 1.0 : v[i]);
}

// R diagonal magnitudes: out[0] = max |R_kk|, out[1] = min |R_kk|
__global__ __launch_bounds__(256) void qr_rdiag_kernel(int m, const double* B, int64_t ldb,
                                                       double* out) {
  __shared__ double red[4];
  double mx = 0.0, mn = INFINITY;
  for (int k = threadIdx.x; k < m; k += 256) {
    const double v = fabs(B[k + (int64_t)k * ldb]);
    mx = fmax(mx, v);
    mn = fmin(mn, isfinite(v) ? v : 0.0);
  }
  mx = block_reduce<1>(mx, red);
  mn = block_reduce<2>(mn, red);
  if (threadIdx.x == 0) { out[0] = mx; out[1] = mn; }
}

// y <- Q' y (trans) or Q y, with the m reflectors stored in B; one workgroup
__global__ __launch_bounds__(1024) void qr_apply_vec_kernel(int64_t rows, int m, const double* B,
                                                            int64_t ldb, const double* tau,
                                                            double* y, int trans) {
  __shared__ double red[16];
  for (int step = 0; step < m; ++step) {
    const int k = trans ? step : m - 1 - step;
    const double* v = B + (int64_t)k * ldb;
    double s = 0.0;
    for (int64_t i = k + threadIdx.x; i < rows; i += blockDim.x) s += (i == k ? 1.0 : v[i]) * y[i];
    s = block_reduce<0>(s, red);
    const double f = tau[k] * s;
    for (int64_t i = k + threadIdx.x; i < rows; i += blockDim.x) y[i] -= f * (i == k ? 1.0 : v[i]);
    __syncthreads();
  }
}

// R x = b (upper, trans = 0) or R' x = b (trans = 1), in place, one workgroup
__global__ __launch_bounds__(1024) void qr_trsv_kernel(int m, const double* B, int64_t ldb,
                                                       double* x, int trans) {
  __shared__ double red[16];
  __shared__ double xk;
  if (!trans) {
    for (int k = m - 1; k >= 0; --k) {
      double s = 0.0;
      for (int c = k + 1 + threadIdx.x; c < m; c += blockDim.x) s += B[k + (int64_t)c * ldb] * x[c];
      s = block_reduce<0>(s, red);
      if (threadIdx.x == 0) xk = (x[k] - s) / B[k + (int64_t)k * ldb];
      __syncthreads();
      if (threadIdx.x == 0) x[k] = xk;
      __syncthreads();
    }
  } else {
    for (int k = 0; k < m; ++k) {
      double s = 0.0;
      for (int r = threadIdx.x; r < k; r += blockDim.x) s += B[r + (int64_t)k * ldb] * x[r];
      s = block_reduce<0>(s, red);
      if (threadIdx.x == 0) xk = (x[k] - s) / B[k + (int64_t)k * ldb];
      __syncthreads();
      if (threadIdx.x == 0) x[k] = xk;
      __syncthreads();
    }
  }
}

// base rows of a block: svec(P) with P = -G' Rd G + diag(mu/D - D) (+ RNT)
__global__ void qr_base_block_kernel(int n, const double* P, const double* D, double mu,
                                     const double* RNT, int corrected, double* base,
                                     int64_t off) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d) return;
  int r, c;
  svec_rc(k, r, c);
  double v = -P[r + (int64_t)c * n];
  if (r == c) v += mu / D[r] - D[r];
  if (corrected) v += RNT[r + (int64_t)c * n];
  base[off + k] = r == c ? v : 1.4142135623730951 * v;
}

__global__ void qr_base_scalar_kernel(int ns, const double* x, const double* z,
                                      const double* Rdlin, const double* corrlin, double mu,
                                      int corrected, double* base, int64_t off) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= ns) return;
  const double w = sqrt(x[l] / z[l]);
  const double p = sqrt(x[l] * z[l]);
  double v = mu / p - p - w * Rdlin[l];
  if (corrected) v += corrlin[l] / w;
  base[off + l] = v;
}

// t = c - q1 ; later dely = scale .* (R^-1 t)
__global__ void qr_combine_kernel(int m, const double* c, const double* q1, double* t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) t[k] = c[k] - q1[k];
}

__global__ void qr_mul_kernel(int m, const double* a, const double* b, double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) out[k] = a[k] * b[k];
}

// smat of the direction rows of a block (off-diagonal / sqrt(2))
__global__ void qr_smat_kernel(int n, const double* dir, int64_t off, double* M) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d) return;
  int r, c;
  svec_rc(k, r, c);
  double v = dir[off + k];
  if (r != c) v /= 1.4142135623730951;
  M[r + (int64_t)c * n] = v;
  M[c + (int64_t)r * n] = v;
}

// delx = sqrt(x ./ z) .* direction[off:]
__global__ void qr_delx_kernel(int ns, const double* x, const double* z, const double* dir,
                               int64_t off, double* delx) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < ns) delx[l] = sqrt(x[l] / z[l]) * dir[off + l];
}

// ---------------------------------------------------------------- host ----
int qr_pack_dense(hipStream_t s, int n, int cnt, const double* P, const int* con, double* B,
                  int64_t ldb, int64_t off) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  qr_pack_kernel<<<dim3(cdiv(d, 256), cnt), 256, 0, s>>>(n, cnt, P, con, B, ldb, off);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_pack_coo(hipStream_t s, const BlockOps& b, const double* G, double* B, int64_t ldb,
                int64_t off) {
  const int n = b.n;
  const int64_t d = (int64_t)n * (n + 1) / 2;
  qr_coo_kernel<<<dim3(cdiv(d, 256), b.ncon), 256, 0, s>>>(n, b.ncon, b.con_list, b.con_ptr,
                                                            b.row, b.col, b.val, G, B, ldb, off);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_scalar_rows(hipStream_t s, int m, int ns, const double* Clin, const double* x,
                   const double* z, double* B, int64_t ldb, int64_t off) {
  if (ns <= 0 || m <= 0) return 0;
  qr_scalar_rows_kernel<<<dim3(cdiv(ns, 256), m), 256, 0, s>>>(m, ns, Clin, x, z, B, ldb, off);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_factor(hipStream_t s, int64_t rows, int m, double* B, double* scale, double* tau,
              int* bad, double* rdiag_out) {
  MSDP_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  qr_colscale_kernel<<<m, 256, 0, s>>>(rows, B, rows, scale, bad);
  for (int k = 0; k < m; ++k) {
    qr_reflector_kernel<<<1, 1024, 0, s>>>(rows, k, B, rows, tau);
    if (k + 1 < m) qr_apply_kernel<<<m - k - 1, 256, 0, s>>>(rows, k, m, B, rows, tau);
  }
  qr_rdiag_kernel<<<1, 256, 0, s>>>(m, B, rows, rdiag_out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_apply_vec(hipStream_t s, int64_t rows, int m, const double* B, const double* tau,
                 double* y, bool trans) {
  qr_apply_vec_kernel<<<1, 1024, 0, s>>>(rows, m, B, rows, tau, y, trans ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_trsv(hipStream_t s, int m, const double* B, int64_t ldb, double* x, bool trans) {
  qr_trsv_kernel<<<1, 1024, 0, s>>>(m, B, ldb, x, trans ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_base_block(hipStream_t s, int n, const double* P, const double* D, double mu,
                  const double* RNT, bool corrected, double* base, int64_t off) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  qr_base_block_kernel<<<cdiv(d, 256), 256, 0, s>>>(n, P, D, mu, RNT, corrected ? 1 : 0, base, off);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_base_scalar(hipStream_t s, int ns, const double* x, const double* z, const double* Rdlin,
                   const double* corrlin, double mu, bool corrected, double* base, int64_t off) {
  if (ns <= 0) return 0;
  qr_base_scalar_kernel<<<cdiv(ns, 256), 256, 0, s>>>(ns, x, z, Rdlin, corrlin, mu,
                                                      corrected ? 1 : 0, base, off);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_combine(hipStream_t s, int m, const double* c, const double* q1, double* t) {
  qr_combine_kernel<<<cdiv(m, 256), 256, 0, s>>>(m, c, q1, t);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_mul(hipStream_t s, int m, const double* a, const double* b, double* out) {
  qr_mul_kernel<<<cdiv(m, 256), 256, 0, s>>>(m, a, b, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_smat(hipStream_t s, int n, const double* dir, int64_t off, double* M) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  qr_smat_kernel<<<cdiv(d, 256), 256, 0, s>>>(n, dir, off, M);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_delx(hipStream_t s, int ns, const double* x, const double* z, const double* dir,
            int64_t off, double* delx) {
  if (ns <= 0) return 0;
  qr_delx_kernel<<<cdiv(ns, 256), 256, 0, s>>>(ns, x, z, dir, off, delx);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
