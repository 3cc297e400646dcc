// QR Schur solver (reference src/qr_schur.jl): the NT-scaled constraint
// matrix B[:, j] = [svec(G' A_j G) per block; sqrt(x ./ z) .* C_lin[j, :]],
// column-equilibrated, Householder QR, and the projection solve of the
// Newton system.  The reference factors B with SuiteSparse SPQR; here B is
// dense in HBM and factored by blocked Householder QR: panels of QR_NB
// columns are factored column by column (one full-chip launch per column
// forms the reflector and applies it to the panel), each panel's reflectors are
// gathered into the compact WY form H_1..H_nb = I - V T V' (T upper
// triangular, LAPACK's larft), and the trailing columns take the panel as
// three MFMA products, C -= V (T' (V' C)).  Q'y, Qy and the solves with R
// run one launch per panel over many workgroups.
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

// Low-rank constraints: G' A_j G = sum_r lam_r g_r g_r' with g_r = G' f_r
// (GF = G' F, n x R):  B[off + k, con_j] = w_k sum_{r in j} lam_r GF[a, r] GF[b, r]
__global__ void qr_lowrank_kernel(int n, int ncon, const int* con_list, const int* con_ptr,
                                  const double* lam, const double* GF, double* B, int64_t ldb,
                                  int64_t off) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= d || j >= ncon) return;
  int a, b;
  svec_rc(k, a, b);
  double s = 0.0;
  for (int r = con_ptr[j]; r < con_ptr[j + 1]; ++r)
    s += lam[r] * GF[a + (int64_t)r * n] * GF[b + (int64_t)r * n];
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

// ---- panel factorisation, one launch per column -------------------------
// Rows are cut into chunks of QR_CH; workgroup g owns chunk g.  Step k of
// the panel [j0, j1) forms reflector k and applies it to the panel's later
// columns: with ss = sum_{i>k} B[i,k]^2 and d_j = sum_{i>k} B[i,k] B[i,j]
// (summed from the previous launch's per-chunk partials) and alpha = B[k,k],
//   beta = -sign(alpha) sqrt(alpha^2 + ss), tau = (beta - alpha) / beta,
//   v = [1; B[k+1:, k] / (alpha - beta)],
//   B[:, j] -= tau (B[k, j] + d_j / (alpha - beta)) v   (j = k+1 .. j1-1),
// (ss = 0: tau = 0, H_k = I).  Each workgroup then writes its chunk's
// partials of ss and d_j for step k+1, which depend on that chunk's updated
// rows only, and the owner of row k+1 copies that row (the next pivot row),
// so that no workgroup reads a row another one is updating.  Partials and
// pivot rows alternate between two buffers.  Launch `k < 0` only writes the
// first step's partials.
__global__ __launch_bounds__(256) void qr_panel_step_kernel(
    int64_t rows, double* B, int64_t ldb, double* tau, int j1, int k, int nk,
    const double* part_in, double* part_out, const double* prow_in, double* prow_out) {
  __shared__ double red[4][QR_NB];
  __shared__ double red2[4][QR_NB];
  __shared__ double f[QR_NB];
  __shared__ double s_beta, s_tau, s_sc;
  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6, G = gridDim.x;
  const int64_t lo = (int64_t)g * QR_CH, hi = std::min<int64_t>(lo + QR_CH, rows);
  const int64_t i = lo + t;
  // rows that take reflector k, and rows below the next pivot (the next
  // step's sums); with k >= 0 the second set lies inside the first
  const bool upd = k >= 0 && i < hi && i > k;
  const bool below = i < hi && i > nk;
  const int ncol = j1 - nk;
  // this thread's entries of column k and of the first eight columns from nk
  // are loaded before the partial sums are reduced (only this thread writes
  // them in this launch)
  const double vraw = upd ? B[i + (int64_t)k * ldb] : 0.0;
  double c[8];
#pragma unroll
  for (int u = 0; u < 8; ++u)
    c[u] = (u < ncol && (upd || below)) ? B[i + (int64_t)(nk + u) * ldb] : 0.0;
  double sc = 0.0;
  if (k >= 0) {
    const int nv = j1 - k;  // ss and the dots with columns k+1 .. j1-1
    const int gk = k + 1 < rows ? (int)((k + 1) / QR_CH) : G;
    double a = 0.0;
    if (lane < nv) {
#pragma unroll 8
      for (int h = gk + wave; h < G; h += 4) a += part_in[(int64_t)h * QR_NB + lane];
    }
    red[wave][lane] = a;
    __syncthreads();
    if (t == 0) {
      const double ss = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
      const double alpha = prow_in[0];
      double beta, tk, scl;
      if (ss == 0.0) {
        beta = alpha; tk = 0.0; scl = 0.0;
      } else {
        beta = -copysign(sqrt(alpha * alpha + ss), alpha);
        tk = (beta - alpha) / beta;
        scl = 1.0 / (alpha - beta);
      }
      s_beta = beta; s_tau = tk; s_sc = scl;
    }
    __syncthreads();
    sc = s_sc;
    if (t > 0 && t < nv) {
      const double d = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
      f[t] = s_tau * (prow_in[t] + sc * d);
    }
    __syncthreads();
    if (i == k && i < hi) {
      B[k + (int64_t)k * ldb] = s_beta;
      tau[k] = s_tau;
      for (int q = 1; q < nv; ++q) B[k + (int64_t)(k + q) * ldb] -= f[q];
    }
  } else if (i == nk && i < hi) {
    for (int q = 0; q < j1 - nk; ++q) prow_out[q] = B[nk + (int64_t)(nk + q) * ldb];
  }
  const double v = sc * vraw;
  if (upd) B[i + (int64_t)k * ldb] = v;
  // columns nk .. j1-1: updated (k >= 0), the new pivot row copied, and each
  // row's terms of ss' = sum x^2 and d'_j = sum x B[., j], x = B[., nk],
  // summed over the wavefront at once; eight columns' loads in flight
  double x = 0.0;
  for (int q0 = 0; q0 < ncol; q0 += 8) {
    if (q0 > 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        c[u] = (q0 + u < ncol && (upd || below)) ? B[i + (int64_t)(nk + q0 + u) * ldb] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + u;
      if (q < ncol) {
        if (upd) {
          c[u] -= f[nk + q - k] * v;
          B[i + (int64_t)(nk + q) * ldb] = c[u];
          if (i == nk) prow_out[q] = c[u];
        }
        if (q == 0) x = c[u];
        double pv = below ? x * c[u] : 0.0;
        for (int o = 32; o > 0; o >>= 1) pv += __shfl_xor(pv, o);
        if (lane == 0) red2[wave][q] = pv;
      }
    }
  }
  if (nk >= j1) return;
  __syncthreads();
  if (hi > nk + 1 && t < ncol)
    part_out[(int64_t)g * QR_NB + t] = (red2[0][t] + red2[1][t]) + (red2[2][t] + red2[3][t]);
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

// ---- blocked factorisation helpers -------------------------------------
// Vp (r x nbe, ld r) = the panel's unit lower trapezoidal V: ones on the
// diagonal, B's sub-diagonal entries below it, zeros above.
__global__ void qr_panel_v_kernel(int64_t r, int nbe, const double* Bp, int64_t ldb, double* Vp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= r || c >= nbe) return;
  Vp[i + c * r] = i == c ? 1.0 : (i > c ? Bp[i + (int64_t)c * ldb] : 0.0);
}

// T (QR_NB x QR_NB, ld QR_NB) of the compact WY form from S = V'V (ld nbe)
// and the panel's tau: T_cc = tau_c, T[0:c, c] = -tau_c T[0:c, 0:c] S[0:c, c]
// (larft, forward, columnwise).  Sixteen lanes share each row's dot product.
__global__ __launch_bounds__(1024) void qr_larft_kernel(int nbe, const double* S,
                                                        const double* tau, double* T) {
  __shared__ double Ts[QR_NB][QR_NB + 1];
  __shared__ double Ss[QR_NB][QR_NB + 1];
  __shared__ double ts[QR_NB];
  const int t = threadIdx.x;
  for (int e = t; e < QR_NB * QR_NB; e += 1024) {
    const int r = e % QR_NB, c = e / QR_NB;
    Ss[r][c] = (r < nbe && c < nbe) ? S[r + (int64_t)c * nbe] : 0.0;
  }
  if (t < QR_NB) ts[t] = t < nbe ? tau[t] : 0.0;
  __syncthreads();
  const int r = t >> 4, part = t & 15;
  for (int c = 0; c < QR_NB; ++c) {
    double v = 0.0;
    if (c < nbe && r < c)
      for (int q = r + part; q < c; q += 16) v += Ts[r][q] * Ss[q][c];
    for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (part == 0) Ts[r][c] = c < nbe ? (r < c ? -ts[c] * v : (r == c ? ts[c] : 0.0)) : 0.0;
    __syncthreads();
  }
  for (int e = t; e < QR_NB * QR_NB; e += 1024) T[e] = Ts[e % QR_NB][e / QR_NB];
}

// ---- Q'y and Qy, one launch per panel ------------------------------------
// Rows are cut into chunks of QR_CH; workgroup g owns chunk g for the whole
// sweep.  Launch p applies panel `cur` to its chunk, y -= V op(T) w with
// w = V'y summed from the previous launch's per-chunk partials, then writes
// its chunk's partials of V'y for panel `nxt`.  A chunk's partials for the
// next panel depend on that chunk's rows of y only, which the same workgroup
// has just updated, so one launch per panel suffices.  Partials alternate
// between two buffers (read one, write the other).
// V[i, c] of the panel starting at column j0 (i an absolute row).
__device__ inline double qr_v(const double* B, int64_t ldb, int j0, int64_t i, int c) {
  const int64_t d = i - j0;
  return d == c ? 1.0 : (d > c ? B[i + (int64_t)(j0 + c) * ldb] : 0.0);
}

__global__ __launch_bounds__(256) void qr_vec_step_kernel(
    int64_t rows, int m, const double* B, int64_t ldb, const double* T, double* y,
    const double* part_in, double* part_out, int cur, int nxt, int trans) {
  __shared__ double ys[QR_CH];
  __shared__ double w[QR_NB], w2[QR_NB];
  __shared__ double red[4][QR_NB];
  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t lo = (int64_t)g * QR_CH, hi = std::min<int64_t>(lo + QR_CH, rows);
  for (int64_t i = lo + t; i < hi; i += 256) ys[i - lo] = y[i];
  __syncthreads();
  if (cur >= 0) {
    const int j0 = cur * QR_NB, nbe = std::min(QR_NB, m - j0);
    const int g0 = j0 / QR_CH, G = (int)((rows + QR_CH - 1) / QR_CH);
    // w[c] = sum of the chunks' partials, in chunk order (4 interleaved
    // sums per column, combined in a fixed order)
    double a = 0.0;
    for (int h = g0 + wave; h < G; h += 4) a += part_in[(int64_t)h * QR_NB + lane];
    red[wave][lane] = a;
    __syncthreads();
    if (t < QR_NB) w[t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    __syncthreads();
    // w2 = T' w (Q'y) or T w (Qy); T is upper triangular
    if (t < QR_NB) {
      double v = 0.0;
      if (t < nbe) {
        if (trans) { for (int r = 0; r <= t; ++r) v += T[(int64_t)cur * QR_NB * QR_NB + r + t * QR_NB] * w[r]; }
        else { for (int c = t; c < nbe; ++c) v += T[(int64_t)cur * QR_NB * QR_NB + t + c * QR_NB] * w[c]; }
      }
      w2[t] = v;
    }
    __syncthreads();
    if (hi > j0) {
      for (int64_t i = std::max<int64_t>(lo, j0) + t; i < hi; i += 256) {
        double s = 0.0;
        for (int c = 0; c < nbe; ++c) s += qr_v(B, ldb, j0, i, c) * w2[c];
        ys[i - lo] -= s;
      }
    }
    __syncthreads();
    for (int64_t i = lo + t; i < hi; i += 256) y[i] = ys[i - lo];
  }
  if (nxt >= 0) {
    const int j0 = nxt * QR_NB, nbe = std::min(QR_NB, m - j0);
    if (hi <= j0) return;
    const int64_t a0 = std::max<int64_t>(lo, j0);
    // wave `wave` sums columns wave, wave + 4, ...; lanes stride the rows
    for (int c = wave; c < nbe; c += 4) {
      double s = 0.0;
      for (int64_t i = a0 + lane; i < hi; i += 64) s += qr_v(B, ldb, j0, i, c) * ys[i - lo];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) part_out[(int64_t)g * QR_NB + c] = s;
    }
  }
}

// ---- solves with R, one launch per panel ---------------------------------
// The diagonal QR_NB x QR_NB block of a panel is solved by one wavefront
// holding the right-hand side in its lanes (lane k <-> row k of the block).
// R x = b, rows/cols [b0, b0 + nbe): back substitution.
__device__ inline void qr_diag_upper(const double* B, int64_t ldb, int b0, int nbe, double* x,
                                     int lane) {
  double v = lane < nbe ? x[b0 + lane] : 0.0;
  for (int k = nbe - 1; k >= 0; --k) {
    const double xk = __shfl(v, k) / B[(b0 + k) + (int64_t)(b0 + k) * ldb];
    if (lane == k) v = xk;
    else if (lane < k) v -= B[(b0 + lane) + (int64_t)(b0 + k) * ldb] * xk;
  }
  if (lane < nbe) x[b0 + lane] = v;
}

// R' x = b on the block: forward substitution.
__device__ inline void qr_diag_lower_t(const double* B, int64_t ldb, int b0, int nbe, double* x,
                                       int lane) {
  double v = lane < nbe ? x[b0 + lane] : 0.0;
  for (int k = 0; k < nbe; ++k) {
    const double xk = __shfl(v, k) / B[(b0 + k) + (int64_t)(b0 + k) * ldb];
    if (lane == k) v = xk;
    else if (lane > k && lane < nbe) v -= B[(b0 + k) + (int64_t)(b0 + lane) * ldb] * xk;
  }
  if (lane < nbe) x[b0 + lane] = v;
}

// R x = b.  Rows [0, u0) take the solved columns [u0, u1):
// x[r] -= R[r, u0:u1] x[u0:u1]; workgroup g owns rows [g QR_NB, (g+1) QR_NB)
// and the last one then solves its diagonal block (the next panel up).
__global__ __launch_bounds__(256) void qr_trsv_upper_kernel(int m, const double* B, int64_t ldb,
                                                            double* x, int u0, int u1) {
  __shared__ double red[4][QR_NB];
  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r0 = g * QR_NB, r1 = std::min(r0 + QR_NB, u0);
  const int r = r0 + lane;
  double s = 0.0;
  if (r < r1)
    for (int c = u0 + wave; c < u1; c += 4) s += B[r + (int64_t)c * ldb] * x[c];
  red[wave][lane] = s;
  __syncthreads();
  if (t < QR_NB && r < r1 && u1 > u0) x[r] -= (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
  if (g == (int)gridDim.x - 1) {
    __syncthreads();
    if (wave == 0) qr_diag_upper(B, ldb, r0, r1 - r0, x, lane);
  }
}

// R' x = b.  Columns [u1, m) take the solved rows [u0, u1):
// x[c] -= R[u0:u1, c]' x[u0:u1]; workgroup g owns columns
// [u1 + g QR_NB, u1 + (g+1) QR_NB) and the first one then solves its
// diagonal block (the next panel down).
__global__ __launch_bounds__(256) void qr_trsv_lower_t_kernel(int m, const double* B, int64_t ldb,
                                                              double* x, int u0, int u1) {
  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int c0 = u1 + g * QR_NB, c1 = std::min(c0 + QR_NB, m);
  const int r = u0 + lane;
  const double xr = r < u1 ? x[r] : 0.0;
  __syncthreads();
  for (int c = c0 + wave; c < c1; c += 4) {
    double s = r < u1 ? B[r + (int64_t)c * ldb] * xr : 0.0;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0 && u1 > u0) x[c] -= s;
  }
  if (g == 0) {
    __syncthreads();
    if (wave == 0) qr_diag_lower_t(B, ldb, c0, c1 - c0, x, lane);
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

int qr_pack_lowrank(hipStream_t s, const BlockOps& b, const double* G, double* B, int64_t ldb,
                    int64_t off) {
  const int n = b.n;
  const int64_t d = (int64_t)n * (n + 1) / 2;
  MSDP_TRY(gemm(s, true, false, n, b.R, n, 1.0, G, n, b.F, n, 0.0, b.FW, n));
  qr_lowrank_kernel<<<dim3(cdiv(d, 256), b.ncon), 256, 0, s>>>(n, b.ncon, b.con_list, b.con_ptr,
                                                                b.lam, b.FW, B, ldb, off);
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

size_t qr_t_elems(int m) { return (size_t)std::max(1, cdiv(m, QR_NB)) * QR_NB * QR_NB; }

size_t qr_work_elems(int64_t rows, int m) {
  const int64_t G = std::max<int64_t>(1, (rows + QR_CH - 1) / QR_CH);
  return (size_t)(rows * QR_NB + 2 * (int64_t)QR_NB * std::max(m, 1) + QR_NB * QR_NB +
                  2 * G * QR_NB + 2 * QR_NB);
}

int qr_factor(hipStream_t s, int64_t rows, int m, double* B, double* scale, double* tau,
              double* T, double* work, int* bad, double* rdiag_out) {
  if (rows < m || rows > INT32_MAX) {
    set_error("qr_factor: needs m <= rows < 2^31");
    return -2;
  }
  MSDP_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  if (m <= 0) return 0;
  qr_colscale_kernel<<<m, 256, 0, s>>>(rows, B, rows, scale, bad);
  double* Vp = work;
  double* W = Vp + rows * QR_NB;
  double* W2 = W + (int64_t)QR_NB * m;
  double* S = W2 + (int64_t)QR_NB * m;
  const int G = cdiv(rows, QR_CH);
  double* part[2] = {S + QR_NB * QR_NB, S + QR_NB * QR_NB + (int64_t)G * QR_NB};
  double* prow[2] = {part[1] + (int64_t)G * QR_NB, part[1] + (int64_t)G * QR_NB + QR_NB};
  for (int j0 = 0, p = 0; j0 < m; j0 += QR_NB, ++p) {
    const int j1 = std::min(j0 + QR_NB, m), nbe = j1 - j0;
    // panel: one launch per column (reflector k applied to columns k+1..j1-1)
    qr_panel_step_kernel<<<G, 256, 0, s>>>(rows, B, rows, tau, j1, -1, j0, part[1], part[0],
                                           prow[1], prow[0]);
    for (int k = j0; k < j1; ++k) {
      const int c = (k - j0) & 1;
      qr_panel_step_kernel<<<G, 256, 0, s>>>(rows, B, rows, tau, j1, k, k + 1, part[c],
                                             part[c ^ 1], prow[c], prow[c ^ 1]);
    }
    const int64_t r = rows - j0;
    double* Bp = B + j0 + (int64_t)j0 * rows;
    qr_panel_v_kernel<<<dim3(cdiv(r, 256), nbe), 256, 0, s>>>(r, nbe, Bp, rows, Vp);
    MSDP_TRY(gemm(s, true, false, nbe, nbe, (int)r, 1.0, Vp, r, Vp, r, 0.0, S, nbe));
    double* Tp = T + (int64_t)p * QR_NB * QR_NB;
    qr_larft_kernel<<<1, 1024, 0, s>>>(nbe, S, tau + j0, Tp);
    if (j1 < m) {
      // C = B[j0:, j1:m] <- (I - V T' V') C = C - V (T' (V' C))
      const int nc = m - j1;
      double* Cp = B + j0 + (int64_t)j1 * rows;
      MSDP_TRY(gemm(s, true, false, nbe, nc, (int)r, 1.0, Vp, r, Cp, rows, 0.0, W, nbe));
      MSDP_TRY(gemm(s, true, false, nbe, nc, nbe, 1.0, Tp, QR_NB, W, nbe, 0.0, W2, nbe));
      MSDP_TRY(gemm(s, false, false, (int)r, nc, nbe, -1.0, Vp, r, W2, nbe, 1.0, Cp, rows));
    }
  }
  qr_rdiag_kernel<<<1, 256, 0, s>>>(m, B, rows, rdiag_out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_apply_vec(hipStream_t s, int64_t rows, int m, const double* B, const double* T,
                 double* work, double* y, bool trans) {
  if (m <= 0) return 0;
  const int P = cdiv(m, QR_NB);
  const int G = cdiv(rows, QR_CH);
  double* part[2] = {work, work + (int64_t)G * QR_NB};
  // panels in order H_1..H_P for Q'y, in reverse for Qy
  auto panel = [&](int k) { return trans ? k : P - 1 - k; };
  qr_vec_step_kernel<<<G, 256, 0, s>>>(rows, m, B, rows, T, y, part[1], part[0], -1, panel(0),
                                       trans ? 1 : 0);
  for (int k = 0; k < P; ++k)
    qr_vec_step_kernel<<<G, 256, 0, s>>>(rows, m, B, rows, T, y, part[k & 1], part[(k + 1) & 1],
                                         panel(k), k + 1 < P ? panel(k + 1) : -1, trans ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int qr_trsv(hipStream_t s, int m, const double* B, int64_t ldb, double* x, bool trans) {
  if (m <= 0) return 0;
  const int P = cdiv(m, QR_NB);
  if (!trans) {
    // last panel's diagonal block, then one launch per panel upwards
    qr_trsv_upper_kernel<<<P, 256, 0, s>>>(m, B, ldb, x, m, m);
    for (int p = P - 1; p >= 1; --p) {
      const int u0 = p * QR_NB, u1 = std::min(u0 + QR_NB, m);
      qr_trsv_upper_kernel<<<p, 256, 0, s>>>(m, B, ldb, x, u0, u1);
    }
  } else {
    qr_trsv_lower_t_kernel<<<P, 256, 0, s>>>(m, B, ldb, x, 0, 0);
    for (int p = 0; p + 1 < P; ++p) {
      const int u0 = p * QR_NB, u1 = u0 + QR_NB;
      qr_trsv_lower_t_kernel<<<P - p - 1, 256, 0, s>>>(m, B, ldb, x, u0, u1);
    }
  }
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
