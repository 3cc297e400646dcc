// Symmetric eigensolvers on gfx950 for the NT scaling (src/directions.jl:54,
// `LRO.eigen!`), the eigenvalue step length (src/step.jl:180-184,
// `LRO.lanczos_eigmin!`) and the proximity eigenvalues of DualCentering
// (src/dual_centering.jl:205-215).
//
//  n <= 64 : parallel cyclic Jacobi, one workgroup per matrix, all in LDS.
//  n  > 64 : Householder tridiagonalisation (two launches per column: a
//            one-workgroup reflector kernel and a fused rank-2-update +
//            symmetric-product kernel over the trailing matrix), then
//            Cuppen's divide and conquer on the tridiagonal with Jacobi
//            leaves, secular equations solved per root by bisection in
//            shifted coordinates, eigenvector recovery through the Loewner
//            (Gu-Eisenstat) formula, MFMA GEMMs for the merges, and a blocked
//            compact-WY back-transformation.
//  lambda_min : tridiagonalisation + Sturm-sequence multisection.
#include "common.h"

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include <hip/hip_ext.h>

#include <algorithm>
#include <cfloat>
#include <vector>

namespace msdp {

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
constexpr double DEPS = 2.220446049250313e-16;

// ==========================================================================
// Jacobi (n <= 64) in LDS
// ==========================================================================
__device__ void jacobi_lds(int n, double (*A)[65], double (*V)[65], bool vectors,
                           double* wout, double* vout, int64_t ldv) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = n + (n & 1);
  __shared__ double cs[32], sn[32];
  __shared__ int pp[32], qq[32];
  __shared__ int rotated;
  __shared__ double fro;
  __shared__ double red[16];
  // Frobenius norm for the absolute threshold
  double v = 0.0;
  for (int e = tid; e < n * n; e += nt) {
    const double a = A[e % n][e / n];
    v += a * a;
  }
  v = block_reduce<0>(v, red);
  if (tid == 0) fro = sqrt(v);
  if (vectors)
    for (int e = tid; e < n * n; e += nt) V[e % n][e / n] = (e % n) == (e / n) ? 1.0 : 0.0;
  __syncthreads();
  const double absthr = 1e-3 * DEPS * fro;
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      if (tid < np / 2) {
        const int i = tid;
        const int ka = i, kb = np - 1 - i;
        int p = ka == 0 ? 0 : ((ka - 1 + round) % (np - 1)) + 1;
        int q = kb == 0 ? 0 : ((kb - 1 + round) % (np - 1)) + 1;
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[p][q], app = A[p][p], aqq = A[q][q];
          if (fabs(apq) > fmax(DEPS * sqrt(fabs(app) * fabs(aqq)), absthr) && apq != 0.0) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            rotated = 1;
          }
        }
        pp[i] = p; qq[i] = q; cs[i] = c; sn[i] = s;
      }
      __syncthreads();
      const int npairs = np / 2;
      // rows: row_p' = c row_p - s row_q ; row_q' = s row_p + c row_q
      for (int e = tid; e < npairs * n; e += nt) {
        const int i = e / n, k = e % n;
        const double s = sn[i];
        if (s == 0.0) continue;
        const double c = cs[i];
        const int p = pp[i], q = qq[i];
        const double rp = A[p][k], rq = A[q][k];
        A[p][k] = c * rp - s * rq;
        A[q][k] = s * rp + c * rq;
      }
      __syncthreads();
      for (int e = tid; e < npairs * n; e += nt) {
        const int i = e / n, k = e % n;
        const double s = sn[i];
        if (s == 0.0) continue;
        const double c = cs[i];
        const int p = pp[i], q = qq[i];
        const double ap = A[k][p], aq = A[k][q];
        A[k][p] = c * ap - s * aq;
        A[k][q] = s * ap + c * aq;
        if (vectors) {
          const double vp = V[k][p], vq = V[k][q];
          V[k][p] = c * vp - s * vq;
          V[k][q] = s * vp + c * vq;
        }
      }
      __syncthreads();
      if (tid < npairs && sn[tid] != 0.0) {
        A[pp[tid]][qq[tid]] = 0.0;
        A[qq[tid]][pp[tid]] = 0.0;
      }
      __syncthreads();
    }
    if (!rotated) break;
    __syncthreads();
  }
  // sort ascending (rank by value, ties by index)
  for (int i = tid; i < n; i += nt) {
    const double di = A[i][i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j][j];
      r += (dj < di) || (dj == di && j < i);
    }
    wout[r] = di;
    if (vectors)
      for (int k = 0; k < n; ++k) vout[k + (int64_t)r * ldv] = V[k][i];
  }
}

__global__ __launch_bounds__(256) void jacobi_batched_kernel(int n, const double* A, int64_t lda,
                                                             int64_t sA, double* w, int64_t sw,
                                                             double* V, int64_t ldv, int64_t sV,
                                                             int vectors) {
  __shared__ double As[64][65];
  __shared__ double Vs[64][65];
  const int64_t b = blockIdx.x;
  A += b * sA;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) As[e % n][e / n] = A[e % n + (int64_t)(e / n) * lda];
  __syncthreads();
  jacobi_lds(n, As, Vs, vectors != 0, w + b * sw, vectors ? V + b * sV : nullptr, ldv);
}

int jacobi_batched(hipStream_t s, int n, const double* A, int64_t lda, int64_t sA, double* w,
                   int64_t sw, double* V, int64_t ldv, int64_t sV, int batch, bool vectors) {
  if (n <= 0 || batch <= 0) return 0;
  if (n > 64) { set_error("jacobi_batched: n > 64"); return -2; }
  jacobi_batched_kernel<<<batch, 256, 0, s>>>(n, A, lda, sA, w, sw, V, ldv, sV, vectors ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// Variable-size batched Jacobi (n <= 128): one workgroup per matrix of the
// batch, the matrix in LDS, the rotations accumulated in a global scratch
// (Vtmp, n x n) and the sorted vectors written to V.  One launch covers the
// eigenproblems of all small PSD blocks of a solve.
// ==========================================================================
__global__ __launch_bounds__(512) void jacobi_vb_kernel(const EigBatch* items, int vectors) {
  const EigBatch it = items[blockIdx.x];
  const int n = it.n;
  if (n <= 0) return;
  __shared__ double A[128][129];
  __shared__ double cs[64], sn[64];
  __shared__ int pp[64], qq[64];
  __shared__ int rotated;
  __shared__ double fro;
  __shared__ double red[16];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = n + (n & 1);
  double v = 0.0;
  for (int e = tid; e < n * n; e += nt) {
    const int i = e % n, j = e / n;
    const double a = it.A[i + (int64_t)j * it.lda];
    A[i][j] = a;
    v += a * a;
  }
  v = block_reduce<0>(v, red);
  if (tid == 0) fro = sqrt(v);
  double* Vt = it.Vtmp;
  if (vectors)
    for (int e = tid; e < n * n; e += nt) Vt[e] = (e % n) == (e / n) ? 1.0 : 0.0;
  __syncthreads();
  const double absthr = 1e-3 * DEPS * fro;
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      if (tid < np / 2) {
        const int i = tid;
        const int ka = i, kb = np - 1 - i;
        int p = ka == 0 ? 0 : ((ka - 1 + round) % (np - 1)) + 1;
        int q = kb == 0 ? 0 : ((kb - 1 + round) % (np - 1)) + 1;
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[p][q], app = A[p][p], aqq = A[q][q];
          if (fabs(apq) > fmax(DEPS * sqrt(fabs(app) * fabs(aqq)), absthr) && apq != 0.0) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            rotated = 1;
          }
        }
        pp[i] = p; qq[i] = q; cs[i] = c; sn[i] = s;
      }
      __syncthreads();
      const int npairs = np / 2;
      for (int e = tid; e < npairs * n; e += nt) {
        const int i = e / n, k = e % n;
        const double s = sn[i];
        if (s == 0.0) continue;
        const double c = cs[i];
        const int p = pp[i], q = qq[i];
        const double rp = A[p][k], rq = A[q][k];
        A[p][k] = c * rp - s * rq;
        A[q][k] = s * rp + c * rq;
      }
      __syncthreads();
      for (int e = tid; e < npairs * n; e += nt) {
        const int i = e / n, k = e % n;
        const double s = sn[i];
        if (s == 0.0) continue;
        const double c = cs[i];
        const int p = pp[i], q = qq[i];
        const double ap = A[k][p], aq = A[k][q];
        A[k][p] = c * ap - s * aq;
        A[k][q] = s * ap + c * aq;
        if (vectors) {
          double* vp = Vt + k + (int64_t)p * n;
          double* vq = Vt + k + (int64_t)q * n;
          const double x = *vp, y = *vq;
          *vp = c * x - s * y;
          *vq = s * x + c * y;
        }
      }
      __syncthreads();
      if (tid < npairs && sn[tid] != 0.0) {
        A[pp[tid]][qq[tid]] = 0.0;
        A[qq[tid]][pp[tid]] = 0.0;
      }
      __syncthreads();
    }
    if (!rotated) break;
    __syncthreads();
  }
  for (int i = tid; i < n; i += nt) {
    const double di = A[i][i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j][j];
      r += (dj < di) || (dj == di && j < i);
    }
    it.w[r] = di;
    if (r == 0 && it.wmin) *it.wmin = di;
  }
  if (!vectors) return;
  __syncthreads();
  // sorted columns: rank of column i computed again per element (n <= 128)
  for (int e = tid; e < n * n; e += nt) {
    const int k = e % n, i = e / n;
    const double di = A[i][i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j][j];
      r += (dj < di) || (dj == di && j < i);
    }
    it.V[k + (int64_t)r * it.ldv] = Vt[k + (int64_t)i * n];
  }
}

int jacobi_vbatched(hipStream_t s, const EigBatch* items, int count, bool vectors) {
  if (count <= 0) return 0;
  jacobi_vb_kernel<<<count, 512, 0, s>>>(items, vectors ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// Householder tridiagonalisation (lower), A = Q T Q',  Q = H_0 ... H_{n-3}
// ==========================================================================
struct TrdWork {
  double *A, *V, *tau, *d, *e, *w, *p0, *p1, *W, *hv, *part, *ppart;
};

constexpr int TRD_NB = 32;

static size_t trd_work_elems(int n) {
  return 2 * (size_t)n * n + 8 * (size_t)n + 3 * (size_t)n * TRD_NB + 2 * TRD_NB + 64;
}

static TrdWork trd_carve(int n, double* work) {
  TrdWork t;
  t.A = work;
  t.V = t.A + (size_t)n * n;
  t.tau = t.V + (size_t)n * n;
  t.d = t.tau + n;
  t.e = t.d + n;
  t.w = t.e + n;
  t.p0 = t.w + n;
  t.p1 = t.p0 + n;
  // W sits in the middle of an n x 3 TRD_NB buffer: the panel's V is copied
  // into the columns just left and just right of W's, so that [V W] and
  // [W V] are contiguous and A -= V W' + W V' is one GEMM of depth 2 jb
  t.W = t.p1 + n + (size_t)n * TRD_NB;
  t.hv = t.W + 2 * (size_t)n * TRD_NB;  // hv[0:nb], hw[nb:2nb]
  t.part = t.hv + 2 * TRD_NB;       // per-workgroup partial squares (<= n / 256 + 1)
  t.ppart = t.part + n / 256 + 8;   // per-workgroup p_c v_c (<= n + 2 TRD_NB)
  return t;
}

// Blocked tridiagonalisation (LAPACK's latrd scheme).  Within a panel of
// TRD_NB columns the trailing matrix is not updated; column k is corrected on
// the fly by the panel's reflectors V and their W = tau (A v - ...) vectors,
// and after the panel A -= V W' + W V' is applied by two MFMA GEMMs.
//
// Column step of panel column j (global column k = k0 + j), two launches:
//  trd_rows_kernel (grid, one row per thread; the row's panel entries are
//    loaded together, the scalars of row k come from LDS):
//    c2   = tau_{k-1}^2 / 2 * (pc . v_{k-1}), with
//           pc . v_{k-1} = p . v_{k-1} - 2 hv . hw          (hv = V' v, hw = W' v)
//    pc_i = p_i - sum_{t<j-1} V[i,t] hw[t] + W[i,t] hv[t]  (= A v_{k-1} now)
//    w_i  = tau_{k-1} pc_i - c2 v_{k-1,i}                  (column j-1 of W)
//    col_i = A[i,k] - sum_{t<j} V[i,t] W[k,t] + W[i,t] V[k,t]   (column k now)
//    and per-workgroup partial sums of col_i^2 (i >= k+2); d_k = col_k.
//  trd_panel_product_kernel (grid, one workgroup per trailing column): every
//    workgroup forms the Householder reflector of column k from the partial
//    squares (tau_k, e_k; v = [1, col_{k+2:} / (alpha - beta)]), stores its
//    entry of V[:, k], and computes its entry of p = A_trailing v (or of
//    hv = V_panel' v, hw = W_panel' v) and the product p_c v_c.
// The identity for pc . v needs no grid-wide reduction of pc, so one launch
// per column is saved against forming pc first.  Deterministic throughout.
__global__ __launch_bounds__(256) void trd_rows_kernel(int n, int k, int j, int k0, int npp,
                                                       const double* A, const double* V,
                                                       double* W, const double* p,
                                                       const double* hv, const double* ppart,
                                                       double* tau, double* colk, double* part,
                                                       double* d, double* e) {
  __shared__ double rowW[TRD_NB], rowV[TRD_NB], hvs[TRD_NB], hws[TRD_NB];
  __shared__ double red[4], red2[4];
  __shared__ double s_c2, s_wk, s_vk, s_tp;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int prev = j - 1;
  const int kp = k0 + prev;             // column of the previous reflector (j > 0)
  const int ncor = j > 0 ? prev : 0;    // panel columns whose terms pc_i still lacks
  const int start = k >= 0 ? k : k0 + j;
  // every load of this launch is issued before the first wait
  if (tid < ncor) {
    hvs[tid] = hv[tid];
    hws[tid] = hv[TRD_NB + tid];
    if (k >= 0) {
      rowW[tid] = W[k + (int64_t)tid * n];
      rowV[tid] = V[k + (int64_t)(k0 + tid) * n];
    }
  }
  double sp = 0.0;
  if (j > 0) {  // 16 independent sums: up to 4096 partials in one round trip
    double su[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) su[u] = 0.0;
    for (int q0 = tid; q0 < npp; q0 += 256 * 16) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int q = q0 + 256 * u;
        su[u] += q < npp ? ppart[q] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 16; u += 2) su[u] += su[u + 1];
#pragma unroll
    for (int u = 0; u < 16; u += 4) su[u] += su[u + 2];
    sp = (su[0] + su[4]) + (su[8] + su[12]);
  }
  const int i = start + blockIdx.x * 256 + tid;
  const bool live = i < n;
  double pi = 0.0, bv = 0.0, vi = 0.0;
  double vr[TRD_NB], wr[TRD_NB];
  if (live) {
    if (k >= 0) bv = A[i + (int64_t)k * n];
    if (j > 0) {
      pi = p[i];
      vi = V[i + (int64_t)kp * n];
    }
#pragma unroll
    for (int t = 0; t < TRD_NB; ++t) {
      vr[t] = t < ncor ? V[i + (int64_t)(k0 + t) * n] : 0.0;
      wr[t] = t < ncor ? W[i + (int64_t)t * n] : 0.0;
    }
  }
  double pk = 0.0, vpk = 0.0, tp = 0.0;
  if (tid == 0 && j > 0) {
    tp = tau[kp];
    if (k >= 0) {
      pk = p[k];
      vpk = V[k + (int64_t)kp * n];
    }
  }
  // three barriers in all: (A) the panel scalars and the partial sums of
  // p . v are in LDS, (B) the column's scalars are, (C) the partial squares are
  if (j > 0) {
    sp = wave_sum(sp);
    if (lane == 0) red[wv] = sp;
  }
  __syncthreads();  // (A)
  double pcv = pi;
  if (live) {
#pragma unroll
    for (int t = 0; t < TRD_NB; ++t)
      if (t < ncor) {
        pcv -= vr[t] * hws[t] + wr[t] * hvs[t];
        if (k >= 0) bv -= vr[t] * rowW[t] + wr[t] * rowV[t];
      }
  }
  if (j > 0) {
    if (tid < 64) {  // wave 0: the scalars of column k
      sp = red[0];
      for (int q = 1; q < 4; ++q) sp += red[q];
      const int t = tid;
      double hh = t < ncor ? hvs[t] * hws[t] : 0.0;
      double ck = (k >= 0 && t < ncor) ? rowV[t] * hws[t] + rowW[t] * hvs[t] : 0.0;
      hh = wave_sum(hh);
      ck = wave_sum(ck);
      if (tid == 0) {
        const double c2 = 0.5 * tp * tp * (sp - 2.0 * hh);
        s_tp = tp;
        s_c2 = c2;
        if (k >= 0) {
          s_wk = tp * (pk - ck) - c2 * vpk;
          s_vk = vpk;
        }
      }
    }
    __syncthreads();  // (B)
  }
  double ss = 0.0;
  if (live) {
    if (j > 0) {
      const double wi = i > kp ? s_tp * pcv - s_c2 * vi : 0.0;
      W[i + (int64_t)prev * n] = wi;
      if (k >= 0) bv -= vi * s_wk + wi * s_vk;
    }
    if (k >= 0) {
      colk[i] = bv;
      if (i >= k + 2) ss = bv * bv;
      if (i == k) d[k] = bv;
      if (k == n - 2 && i == n - 1) e[k] = bv;
    }
  }
  if (k < 0) return;
  if (k == n - 2) {  // the trailing 2 x 2 block: no reflector
    if (blockIdx.x == 0 && tid == 0) {
      tau[k] = 0.0;
      d[n - 1] = A[(n - 1) + (int64_t)(n - 1) * n];
    }
    return;
  }
  ss = wave_sum(ss);
  if (lane == 0) red2[wv] = ss;
  __syncthreads();  // (C)
  if (tid == 0) {
    ss = red2[0];
    for (int q = 1; q < 4; ++q) ss += red2[q];
    part[blockIdx.x] = ss;
  }
}

// One workgroup per output: p[c] = A[base:, c]' v for the trailing columns
// c >= base (A symmetric, full storage; workgroups 0..L-1), then
// hv = V_panel' v and hw = W_panel' v for the j panel columns before k.
// Each thread keeps TRD_PR rows of the column and of v's numerators in
// flight at once; the partial sums are combined in a fixed order.
constexpr int TRD_PR = 12;

__global__ __launch_bounds__(256) void trd_panel_product_kernel(int n, int k, int j, int k0,
                                                                int nparts, const double* A,
                                                                double* V, const double* W,
                                                                const double* colk,
                                                                const double* part, double* tau,
                                                                double* e, double* p,
                                                                double* hv, double* ppart,
                                                                double* vl, double* vr) {
  __shared__ double red[4], red2[4];
  const int tid = threadIdx.x;
  const int base = k + 1;
  const int L = n - base;
  const int c = blockIdx.x;
  const double* colp;
  double* out;
  if (c < L) {
    colp = A + (int64_t)(base + c) * n;
    out = p + base + c;
  } else if (c < L + j) {
    colp = V + (int64_t)(k0 + c - L) * n;
    out = hv + (c - L);
  } else {
    colp = W + (int64_t)(c - L - j) * n;
    out = hv + TRD_NB + (c - L - j);
  }
  // every load is issued before the first wait: the column's rows, column k,
  // the partial squares and the scalars thread 0 needs
  double av[TRD_PR], cv[TRD_PR];
#pragma unroll
  for (int s = 0; s < TRD_PR; ++s) {
    const int i = base + tid + 256 * s;
    av[s] = i < n ? colp[i] : 0.0;
    cv[s] = i < n ? colk[i] : 0.0;
  }
  double ss = 0.0;
  for (int q = tid; q < nparts; q += 256) ss += part[q];
  double alpha = 0.0, ckc = 0.0;
  if (tid == 0) {
    alpha = colk[base];
    if (c < L) ckc = colk[base + c];
  }
  // p_c = A[base, c] + sc * sum_{i > base} A[i, c] col_k[i]: the unscaled sum
  // needs no reflector, so one reduction (of the pair) serves both
  double acc = 0.0;
#pragma unroll
  for (int s = 0; s < TRD_PR; ++s) {
    const int i = base + tid + 256 * s;
    if (i > base) acc += av[s] * cv[s];
  }
  for (int i = base + tid + 256 * TRD_PR; i < n; i += 256) acc += colp[i] * colk[i];
  const int lane = tid & 63, w = tid >> 6;
  ss = wave_sum(ss);
  acc = wave_sum(acc);
  if (lane == 0) {
    red[w] = ss;
    red2[w] = acc;
  }
  __syncthreads();
  if (tid != 0) return;
  ss = (red[0] + red[1]) + (red[2] + red[3]);
  acc = (red2[0] + red2[1]) + (red2[2] + red2[3]);
  double beta, t, sc;
  if (ss == 0.0) {
    beta = alpha; t = 0.0; sc = 0.0;
  } else {
    beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    t = (beta - alpha) / beta;
    sc = 1.0 / (alpha - beta);
  }
  if (c == 0) {
    tau[k] = t;
    e[k] = beta;
  }
  const double pc = av[0] + sc * acc;  // av[0] = A[base, c] in thread 0
  *out = pc;
  double pv = 0.0;
  if (c < L) {
    const double vc = c == 0 ? 1.0 : ckc * sc;
    V[(base + c) + (int64_t)k * n] = vc;
    vl[base + c] = vc;  // the copies left and right of W (rows > k)
    vr[base + c] = vc;
    pv = pc * vc;
  }
  ppart[c] = pv;
}

// --------------------------------------------------------------------------
// One-workgroup tridiagonalisation for n <= TRD1_MAX: the whole reduction in
// a single launch (3n launches of the blocked scheme otherwise), the working
// matrix in global memory (L2-resident at these sizes).  Unblocked Householder
// (LAPACK's sytd2, lower) with the rank-2 update of step k fused with the
// matrix-vector product of step k+1: column k+1 is updated first, its
// reflector v' formed, then one pass over the trailing block applies
// A -= v w' + w v' and accumulates p' = A v'.  Same outputs as the blocked
// routine (V, tau, d, e).
// --------------------------------------------------------------------------
constexpr int TRD1_MAX = 640;
constexpr int TRD1_NT = 1024;

__device__ inline void trd1_reflector(int n, int c, double* A, double* V, double* tau, double* d,
                                      double* e, double* vdst, double* red, double* s_tau,
                                      double* s_scale) {
  const int tid = threadIdx.x;
  const double* col = A + (int64_t)c * n;
  double ss = 0.0;
  for (int i = c + 2 + tid; i < n; i += TRD1_NT) ss += col[i] * col[i];
  ss = block_reduce<0>(ss, red);
  if (tid == 0) {
    const double alpha = col[c + 1];
    double beta, t, sc;
    if (ss == 0.0) {
      beta = alpha; t = 0.0; sc = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + ss), alpha);
      t = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    d[c] = col[c];
    e[c] = beta;
    tau[c] = t;
    *s_tau = t;
    *s_scale = sc;
  }
  __syncthreads();
  const double sc = *s_scale;
  for (int i = c + 1 + tid; i < n; i += TRD1_NT) {
    const double v = i == c + 1 ? 1.0 : col[i] * sc;
    vdst[i] = v;
    V[i + (int64_t)c * n] = v;
  }
  __syncthreads();
}

// rows/cols >= c0: if FUSED, A -= v w' + w v' first; p = A vn (unscaled).
// Wave g (of 16) takes the columns c0 + g + 16 m, lane = row within a 64-row
// chunk; loads of the next 8 columns are issued before the stores of the
// current 8, so the stores' latency overlaps the loads'.  The 16 partial
// sums of every row are added in a fixed order.
constexpr int TRD1_W = TRD1_NT / 64;

template <bool FUSED>
__device__ inline void trd1_pass(int n, int c0, double* A, const double* v, const double* w,
                                 const double* vn, double* p, double (*part)[TRD1_MAX]) {
  constexpr int U = 8, S = TRD1_W;
  const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
  for (int rb = c0; rb < n; rb += 64) {
    const int i = rb + lane;
    double acc = 0.0;
    if (i < n) {
      const double vi = FUSED ? v[i] : 0.0, wi = FUSED ? w[i] : 0.0;
      double* Ai = A + i;
      int j = c0 + g;
      double a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] = j + S * u < n ? Ai[(int64_t)(j + S * u) * n] : 0.0;
      for (; j < n; j += S * U) {
        const int jn = j + S * U;
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = jn + S * u < n ? Ai[(int64_t)(jn + S * u) * n] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int jj = j + S * u;
          if (jj < n) {
            double x = a[u];
            if (FUSED) {
              x -= vi * w[jj] + wi * v[jj];
              Ai[(int64_t)jj * n] = x;
            }
            acc += x * vn[jj];
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = b[u];
      }
      part[g][i - c0] = acc;
    }
  }
  __syncthreads();
  for (int i = c0 + tid; i < n; i += TRD1_NT) {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < S; ++q) sum += part[q][i - c0];
    p[i] = sum;
  }
  __syncthreads();
}

// Workgroup 1 (when launched) reduces Ain1 into the work arrays wstride
// elements further on: two independent matrices in one launch.
__global__ __launch_bounds__(TRD1_NT) void trd1_kernel(int n, const double* Ain, int64_t lda,
                                                       double* A, double* V, double* tau,
                                                       double* d, double* e, const double* Ain1,
                                                       int64_t wstride) {
  if (blockIdx.x) {
    Ain = Ain1;
    A += wstride; V += wstride; tau += wstride; d += wstride; e += wstride;
  }
  __shared__ double vs[TRD1_MAX], ws[TRD1_MAX], vn[TRD1_MAX], ps[TRD1_MAX];
  __shared__ double part[TRD1_W][TRD1_MAX];
  __shared__ double red[16];
  __shared__ double s_tau, s_scale;
  const int tid = threadIdx.x;
  for (int64_t x = tid; x < (int64_t)n * n; x += TRD1_NT) {
    const int i = (int)(x % n), j = (int)(x / n);
    A[x] = Ain[i + (int64_t)j * lda];
    V[x] = 0.0;
  }
  __syncthreads();
  if (n <= 2) {
    if (tid == 0) {
      d[0] = A[0];
      if (n == 2) {
        e[0] = A[1];
        d[1] = A[3];
        tau[0] = 0.0;
      }
    }
    return;
  }
  trd1_reflector(n, 0, A, V, tau, d, e, vs, red, &s_tau, &s_scale);
  trd1_pass<false>(n, 1, A, nullptr, nullptr, vs, ps, part);
  for (int k = 0;; ++k) {
    const double t = s_tau;
    double pv = 0.0;
    for (int i = k + 1 + tid; i < n; i += TRD1_NT) pv += ps[i] * vs[i];
    pv = block_reduce<0>(pv, red);
    const double c2 = 0.5 * t * t * pv;
    for (int i = k + 1 + tid; i < n; i += TRD1_NT) ws[i] = t * ps[i] - c2 * vs[i];
    __syncthreads();
    if (k == n - 3) {  // the trailing 2 x 2 block, then no reflector
      if (tid == 0) {
        const int a = n - 2, b = n - 1;
        const double aa = A[a + (int64_t)a * n] - 2.0 * vs[a] * ws[a];
        const double ba = A[b + (int64_t)a * n] - (vs[b] * ws[a] + ws[b] * vs[a]);
        const double bb = A[b + (int64_t)b * n] - 2.0 * vs[b] * ws[b];
        d[a] = aa;
        e[a] = ba;
        d[b] = bb;
        tau[a] = 0.0;
      }
      return;
    }
    const int c = k + 1;
    for (int i = c + tid; i < n; i += TRD1_NT)
      A[i + (int64_t)c * n] -= vs[i] * ws[c] + ws[i] * vs[c];
    __syncthreads();
    trd1_reflector(n, c, A, V, tau, d, e, vn, red, &s_tau, &s_scale);
    trd1_pass<true>(n, c + 1, A, vs, ws, vn, ps, part);
    for (int i = c + 1 + tid; i < n; i += TRD1_NT) vs[i] = vn[i];
    __syncthreads();
  }
}

// `count` (1 or 2) one-workgroup reductions in one launch; under live
// profiling the launch is timed on its dispatch, algorithmic work the
// Householder tridiagonalisation's (4/3) n^3 flops per matrix
static int launch_trd1(hipStream_t s, int count, int n, const double* A0, int64_t lda,
                       const TrdWork& t, const double* A1, int64_t wstride) {
  if (prof_active(PROF_TRD1, s)) {
    hipEvent_t e0, e1;
    prof_dispatch_events(&e0, &e1, count * (4.0 / 3.0) * (double)n * n * n);
    hipExtLaunchKernelGGL(trd1_kernel, dim3(count), dim3(TRD1_NT), 0, s, e0, e1, 0, n, A0, lda,
                          t.A, t.V, t.tau, t.d, t.e, A1, wstride);
  } else {
    trd1_kernel<<<count, TRD1_NT, 0, s>>>(n, A0, lda, t.A, t.V, t.tau, t.d, t.e, A1, wstride);
  }
  MSDP_HIP(hipGetLastError());
  return 0;
}

static int sytrd_columns(hipStream_t s, int n, const TrdWork& t);

// The column sequence of the blocked reduction (n > TRD1_MAX) has no host
// read and no data-dependent launch: for a given n and workspace it is the
// same 2n + n/32 kernels every time.  It is captured once per (device, n,
// workspace) as a HIP graph and replayed, which removes the host launch cost
// of its ~6,000 dependent kernels (n = 3000).  Not under live profiling
// (whose events are stamped per launch), not on the legacy null stream
// (which cannot be captured).
namespace {
struct TrdGraphKey {
  int device, n;
  const double* work;
  bool operator==(const TrdGraphKey& o) const {
    return device == o.device && n == o.n && work == o.work;
  }
};
struct TrdGraphCache {
  std::mutex mu;
  std::vector<std::pair<TrdGraphKey, hipGraphExec_t>> entries;
};
TrdGraphCache g_trd_graphs;
constexpr size_t TRD_GRAPHS_MAX = 64;
constexpr int TRD_GRAPH_MAX_N = 4096;
}  // namespace

static int sytrd_columns_graph(hipStream_t s, int n, const TrdWork& t, bool& done) {
  done = false;
  // n <= TRD_GRAPH_MAX_N: the panel updates' products then take the plain
  // tile kernels (no pool allocation inside the capture)
  // MADSDP_HIP_GRAPHS=0 turns the replay off (eager launches, e.g. for
  // kernel tracers that do not follow captured launches)
  static const bool graphs_off = [] {
    const char* v = std::getenv("MADSDP_HIP_GRAPHS");
    return v && std::strcmp(v, "0") == 0;
  }();
  if (graphs_off || s == nullptr || n > TRD_GRAPH_MAX_N || prof_active(PROF_TRD, s) ||
      prof_active(PROF_GEMM, s))
    return 0;
  TrdGraphKey key{0, n, t.A};
  MSDP_HIP(hipGetDevice(&key.device));
  hipGraphExec_t exec = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_trd_graphs.mu);
    for (auto& e : g_trd_graphs.entries)
      if (e.first == key) exec = e.second;
    if (!exec && g_trd_graphs.entries.size() >= TRD_GRAPHS_MAX) return 0;
  }
  if (!exec) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MSDP_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return 0;  // the caller is capturing: launch eagerly
    MSDP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = sytrd_columns(s, n, t);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != 0) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture", __FILE__, __LINE__);
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail_hip(ei, "hipGraphInstantiate", __FILE__, __LINE__);
    std::lock_guard<std::mutex> lk(g_trd_graphs.mu);
    g_trd_graphs.entries.emplace_back(key, exec);
  }
  MSDP_HIP(hipGraphLaunch(exec, s));
  done = true;
  return 0;
}

static int sytrd(hipStream_t s, int n, const double* A, int64_t lda, const TrdWork& t) {
  if (n <= TRD1_MAX) return launch_trd1(s, 1, n, A, lda, t, A, 0);
  MSDP_TRY(copy_mat(s, n, n, A, lda, t.A, n));
  MSDP_TRY(fill(s, (int64_t)n * n, 0.0, t.V));
  bool done = false;
  MSDP_TRY(sytrd_columns_graph(s, n, t, done));
  if (done) return 0;
  return sytrd_columns(s, n, t);
}

static int sytrd_columns(hipStream_t s, int n, const TrdWork& t) {
  // column step: rows (w_{k-1}, column k) then the reflector and products
  auto rows = [&](int k, int j, int k0, int npp) -> int {
    const int start = k >= 0 ? k : k0 + j;
    const int nparts = std::max(1, cdiv(n - start, 256));
    trd_rows_kernel<<<nparts, 256, 0, s>>>(n, k, j, k0, npp, t.A, t.V, t.W, t.p0, t.hv, t.ppart,
                                           t.tau, t.p1, t.part, t.d, t.e);
    MSDP_HIP(hipGetLastError());
    return nparts;
  };
  const int last = n - 3;  // last column with a reflector
  for (int k0 = 0; k0 <= last; k0 += TRD_NB) {
    const int jb = std::min(TRD_NB, last - k0 + 1);
    int npp = 0;
    for (int j = 0; j < jb; ++j) {
      const int k = k0 + j;
      const int nparts = rows(k, j, k0, npp);
      if (nparts < 0) return nparts;
      const int L = n - k - 1;
      npp = L + 2 * j;
      double* vl = t.W - (int64_t)(jb - j) * n;  // [V_panel W_panel V_panel] around W
      double* vr = t.W + (int64_t)(jb + j) * n;
      if (prof_active(PROF_TRD, s)) {
        // timed on the dispatch itself; algorithmic bytes: the trailing
        // L x L matrix read once
        hipEvent_t e0, e1;
        prof_dispatch_events(&e0, &e1, 8.0 * L * (double)L);
        hipExtLaunchKernelGGL(trd_panel_product_kernel, dim3(npp), dim3(256), 0, s, e0, e1, 0, n, k,
                              j, k0, nparts, (const double*)t.A, t.V, (const double*)t.W,
                              (const double*)t.p1, (const double*)t.part, t.tau, t.e, t.p0, t.hv,
                              t.ppart, vl, vr);
      } else {
        trd_panel_product_kernel<<<npp, 256, 0, s>>>(n, k, j, k0, nparts, t.A, t.V, t.W, t.p1,
                                                     t.part, t.tau, t.e, t.p0, t.hv, t.ppart, vl,
                                                     vr);
      }
      MSDP_HIP(hipGetLastError());
    }
    // w of the panel's last column
    MSDP_TRY(rows(-1, jb, k0, npp));
    const int r0 = k0 + jb;
    const int L2 = n - r0;
    if (L2 > 0) {  // A22 -= [V W] [W V]'
      const double* VW = t.W - (int64_t)jb * n + r0;
      const double* WV = t.W + r0;
      double* A22 = t.A + r0 + (int64_t)r0 * n;
      MSDP_TRY(gemm(s, false, true, L2, L2, 2 * jb, -1.0, VW, n, WV, n, 1.0, A22, n));
    }
  }
  // the last 2 x 2 block: d_{n-2}, e_{n-2}, d_{n-1} (no reflector)
  MSDP_TRY(rows(n - 2, 0, std::max(0, n - 2), 0));
  return 0;
}

// ==========================================================================
// Sturm multisection for the smallest eigenvalue of the tridiagonal (d, e)
// ==========================================================================
__device__ inline int sturm_count(int n, const double* d, const double* e, double x,
                                  double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0;
  }
  return cnt;
}

// workgroup 1 (when launched): the tridiagonal ws elements further on, into out1
__global__ __launch_bounds__(256) void tridiag_min_kernel(int n, const double* d, const double* e,
                                                          double* out, int64_t ws, double* out1) {
  if (blockIdx.x) {
    d += ws; e += ws; out = out1;
  }
  __shared__ double red[4];
  __shared__ double lo_s, hi_s;
  __shared__ int first;
  const int tid = threadIdx.x;
  double gl = INFINITY, gu = -INFINITY, emax = 0.0;
  for (int i = tid; i < n; i += 256) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    gl = fmin(gl, d[i] - r);
    gu = fmax(gu, d[i] + r);
    if (i < n - 1) emax = fmax(emax, e[i] * e[i]);
  }
  gl = block_reduce<2>(gl, red);
  gu = block_reduce<1>(gu, red);
  emax = block_reduce<1>(emax, red);
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  if (tid == 0) {
    lo_s = gl - 2 * DEPS * tnorm - 2 * pivmin;
    hi_s = gu + 2 * DEPS * tnorm + 2 * pivmin;
  }
  __syncthreads();
  for (int round = 0; round < 12; ++round) {
    const double lo = lo_s, hi = hi_s;
    if (hi - lo <= 2 * DEPS * fmax(fabs(lo), fabs(hi)) + 4 * pivmin) break;
    if (tid == 0) first = 256;
    __syncthreads();
    const double x = lo + (hi - lo) * (double)(tid + 1) / 257.0;
    const int c = sturm_count(n, d, e, x, pivmin);
    if (c >= 1) atomicMin(&first, tid);
    __syncthreads();
    if (tid == 0) {
      const int f = first;
      const double nlo = f == 0 ? lo : lo + (hi - lo) * (double)f / 257.0;
      const double nhi = f == 256 ? hi : lo + (hi - lo) * (double)(f + 1) / 257.0;
      lo_s = nlo;
      hi_s = nhi;
    }
    __syncthreads();
  }
  if (tid == 0) *out = 0.5 * (lo_s + hi_s);
}

// ==========================================================================
// Eigenvalues only, n <= 128, a batch of matrices of different sizes (one
// workgroup each): Householder tridiagonalisation with the matrix in LDS,
// then Sturm bisection (all eigenvalues, ascending, one thread per index) or
// multisection for the smallest.  Replaces the Jacobi sweeps where no
// eigenvectors are needed (step lengths, the DualCentering predictor).
// ==========================================================================
constexpr int EV_MAX = 128;

// Householder reduction of the n x n (n <= 128) symmetric matrix in As to
// tridiagonal form (d[0..n-1], e[0..n-2]) by 512 threads, four barriers per
// step: wavefront 0 forms the reflector, all threads p = tau A v, wavefront 0
// w = p - (tau/2)(p'v) v, all threads the rank-2 update.  The two sums (the
// reflector's norm, p'v) are wave sums over lanes holding rows k + c + lane
// and k + c + 64 + lane, added in that order: the sums of a block reduction
// over one row per thread.  With Hv, column k receives the reflector (v
// below the diagonal, tau on it).
__device__ __attribute__((always_inline)) inline void trd_lds(double (*As)[EV_MAX + 1], int n,
                                                              double* v, double* p, double* w,
                                                              double* d, double* e, double* Hv,
                                                              double* s_tau) {
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int k = 0; k < n - 2; ++k) {
    if (tid < 64) {
      const int lane = tid, i0 = k + 2 + lane, i1 = i0 + 64;
      const double a0 = i0 < n ? As[i0][k] : 0.0, a1 = i1 < n ? As[i1][k] : 0.0;
      // each term as fma(x, x, 0) (a product, not contracted into the sum)
      const double ss = wave_sum(fma(a0, a0, 0.0)) + wave_sum(fma(a1, a1, 0.0));
      const double alpha = As[k + 1][k];
      double beta, t, sc;
      if (ss == 0.0) {
        beta = alpha; t = 0.0; sc = 0.0;
      } else {
        beta = -copysign(sqrt(alpha * alpha + ss), alpha);
        t = (beta - alpha) / beta;
        sc = 1.0 / (alpha - beta);
      }
      if (lane == 0) {
        d[k] = As[k][k];
        e[k] = beta;
        *s_tau = t;
        if (Hv) Hv[k + (int64_t)k * n] = t;
      }
      for (int i = k + 1 + lane; i < n; i += 64) {
        const double vi = i == k + 1 ? 1.0 : As[i][k] * sc;
        v[i] = vi;
        if (Hv) Hv[i + (int64_t)k * n] = vi;
      }
    }
    __syncthreads();
    const double t = *s_tau;
    if (t == 0.0) continue;  // uniform: H_k = I
    const int L = n - k - 1;
    {  // p = t A22 v, 4 lanes per row
      const int r = tid >> 2, q = tid & 3, i = k + 1 + r;
      double acc = 0.0;
      if (r < L)
        for (int j = k + 1 + q; j < n; j += 4) acc += As[i][j] * v[j];
      acc += __shfl_xor(acc, 1);
      acc += __shfl_xor(acc, 2);
      if (q == 0 && r < L) p[i] = t * acc;
    }
    __syncthreads();
    if (tid < 64) {
      const int lane = tid, i0 = k + 1 + lane, i1 = i0 + 64;
      const double q0 = i0 < n ? fma(p[i0], v[i0], 0.0) : 0.0;
      const double q1 = i1 < n ? fma(p[i1], v[i1], 0.0) : 0.0;
      const double pv = wave_sum(q0) + wave_sum(q1);
      const double c2 = 0.5 * t * pv;
      if (i0 < n) w[i0] = p[i0] - c2 * v[i0];
      if (i1 < n) w[i1] = p[i1] - c2 * v[i1];
    }
    __syncthreads();
    for (int x = tid; x < L * L; x += NT) {
      const int i = k + 1 + x % L, j = k + 1 + x / L;
      As[i][j] -= v[i] * w[j] + w[i] * v[j];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(512) void eigvals_vb_kernel(const EigBatch* items) {
  const EigBatch it = items[blockIdx.x];
  const int n = it.n;
  __shared__ double As[EV_MAX][EV_MAX + 1];
  __shared__ double v[EV_MAX], p[EV_MAX], w[EV_MAX], d[EV_MAX], e[EV_MAX];
  __shared__ double red[16];
  __shared__ double s_tau, lo_s, hi_s;
  __shared__ int first;
  const int tid = threadIdx.x, NT = 512;
  for (int x = tid; x < n * n; x += NT) {
    const int i = x % n, j = x / n;
    As[i][j] = it.A[i + (int64_t)j * it.lda];
  }
  __syncthreads();
  trd_lds(As, n, v, p, w, d, e, nullptr, &s_tau);
  if (tid == 0) {
    if (n >= 2) {
      d[n - 2] = As[n - 2][n - 2];
      e[n - 2] = As[n - 1][n - 2];
    }
    d[n - 1] = As[n - 1][n - 1];
  }
  __syncthreads();
  // Gershgorin bounds
  double gl = INFINITY, gu = -INFINITY, emax = 0.0;
  for (int i = tid; i < n; i += NT) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    gl = fmin(gl, d[i] - r);
    gu = fmax(gu, d[i] + r);
    if (i < n - 1) emax = fmax(emax, e[i] * e[i]);
  }
  gl = block_reduce<2>(gl, red);
  gu = block_reduce<1>(gu, red);
  emax = block_reduce<1>(emax, red);
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double lo0 = gl - 2 * DEPS * tnorm - 2 * pivmin, hi0 = gu + 2 * DEPS * tnorm + 2 * pivmin;
  if (it.w) {  // every eigenvalue by bisection on its index
    if (tid < n) {
      double lo = lo0, hi = hi0;
      for (int s = 0; s < 128; ++s) {
        if (hi - lo <= 2 * DEPS * fmax(fabs(lo), fabs(hi)) + 4 * pivmin) break;
        const double mid = 0.5 * (lo + hi);
        if (sturm_count(n, d, e, mid, pivmin) >= tid + 1) hi = mid;
        else lo = mid;
      }
      const double lam = 0.5 * (lo + hi);
      it.w[tid] = lam;
      if (tid == 0 && it.wmin) *it.wmin = lam;
    }
    return;
  }
  if (tid == 0) { lo_s = lo0; hi_s = hi0; }
  __syncthreads();
  for (int round = 0; round < 12; ++round) {  // 513-way multisection for the smallest
    const double lo = lo_s, hi = hi_s;
    if (hi - lo <= 2 * DEPS * fmax(fabs(lo), fabs(hi)) + 4 * pivmin) break;
    if (tid == 0) first = NT;
    __syncthreads();
    const double x = lo + (hi - lo) * (double)(tid + 1) / (NT + 1.0);
    if (sturm_count(n, d, e, x, pivmin) >= 1) atomicMin(&first, tid);
    __syncthreads();
    if (tid == 0) {
      const int f = first;
      lo_s = f == 0 ? lo : lo + (hi - lo) * (double)f / (NT + 1.0);
      hi_s = f == NT ? hi : lo + (hi - lo) * (double)(f + 1) / (NT + 1.0);
    }
    __syncthreads();
  }
  if (tid == 0 && it.wmin) *it.wmin = 0.5 * (lo_s + hi_s);
}

// --------------------------------------------------------------------------
// Eigen-decomposition with vectors, n <= 128, one workgroup per matrix:
// Householder tridiagonalisation in LDS (reflectors kept in the item's Vtmp,
// tau_k on its diagonal), implicit QL with Wilkinson-type shifts on the
// tridiagonal (wavefront 0 forms each sweep's rotations, every thread applies
// them to its row of Z held in the same LDS), the back-transformation
// V = H_0 ... H_{n-3} Z, and the columns sorted by eigenvalue.  Replaces the
// cyclic Jacobi sweeps for the NT scaling of small blocks.
// --------------------------------------------------------------------------
__device__ void syev_small(const EigBatch& it) {
  const int n = it.n;
  __shared__ double As[EV_MAX][EV_MAX + 1];
  __shared__ double v[EV_MAX], p[EV_MAX], w[EV_MAX], d[EV_MAX], e[EV_MAX];
  __shared__ double cs[2 * EV_MAX], sn[2 * EV_MAX];  // two sweeps' rotations
  __shared__ double s_tau;
  __shared__ int s_cnt[2], s_top[2], s_state;
  const int tid = threadIdx.x, NT = blockDim.x;
  double* Hv = it.Vtmp;
  for (int x = tid; x < n * n; x += NT) {
    const int i = x % n, j = x / n;
    As[i][j] = it.A[i + (int64_t)j * it.lda];
  }
  __syncthreads();
  trd_lds(As, n, v, p, w, d, e, Hv, &s_tau);
  if (tid == 0) {
    if (n >= 2) {
      d[n - 2] = As[n - 2][n - 2];
      e[n - 2] = As[n - 1][n - 2];
    }
    d[n - 1] = As[n - 1][n - 1];
    e[n - 1] = 0.0;
  }
  __syncthreads();
  // Z = I in the same LDS
  for (int x = tid; x < n * n; x += NT) {
    const int i = x % n, j = x / n;
    As[i][j] = i == j ? 1.0 : 0.0;
  }
  // implicit QL (e[i] couples d[i] and d[i+1]).  Wavefront 0 runs the
  // control in every lane (same values): the split search is one ballot per
  // 64 entries, and the chase carries d[i], e[i] in registers, loading the
  // next pair one rotation ahead (d[i] is not written by rotation i).
  int l = 0, iter = 0, sweep = 0;
  for (;;) {
    __syncthreads();
    if (tid < 64) {
      const int lane = tid;
      int state = 0, cnt = 0, top = 0;
      for (;;) {
        if (l >= n) { state = 2; break; }
        // first m >= l with |e[m]| <= eps (|d[m]| + |d[m+1]|), else n - 1
        int m = n - 1;
        for (int b = l; b < n - 1; b += 64) {
          const int mm = b + lane;
          const bool hit = mm < n - 1 && fabs(e[mm]) <= DEPS * (fabs(d[mm]) + fabs(d[mm + 1]));
          const unsigned long long mask = __ballot(hit);
          if (mask) {
            m = b + __ffsll((long long)mask) - 1;
            break;
          }
        }
        if (m == l) {  // d[l] converged
          ++l;
          iter = 0;
          continue;
        }
        if (++iter > 60) { state = 3; break; }
        // The chase below is a dependent chain whose longest latencies were a
        // square root and a quotient per rotation: each rotation instead takes
        // y = 1 / sqrt(f^2 + g^2) (rsqrt_nr), then r = x y, s = f y, c = g y.
        // The sum of squares replaces hypot: the operands are matrix entries
        // far inside the range where squaring overflows or underflows.
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = sqrt(fma(g, g, 1.0));
        g = d[m] - d[l] + e[l] / (g + copysign(r, g));
        double s = 1.0, c = 1.0, pp = 0.0;
        bool split = false;
        double dip1 = d[m], ei = e[m - 1], di = d[m - 1];
        for (int i = m - 1; i >= l; --i) {
          // the next pair, loaded without a branch (a guarded load would
          // be waited for at once); at i == l it is read and not used
          const int ip = i > 0 ? i - 1 : 0;
          const double en = e[ip], dn = d[ip];
          const double f = s * ei, b = c * ei;
          const double x = fma(f, f, g * g);
          if (x == 0.0) {
            e[i + 1] = 0.0;
            d[i + 1] = dip1 - pp;
            e[m] = 0.0;
            split = true;
            break;
          }
          const double ri = rsqrt_nr(x);
          e[i + 1] = x * ri;
          s = f * ri;
          c = g * ri;
          g = dip1 - pp;
          r = (di - g) * s + 2.0 * c * b;
          pp = s * r;
          d[i + 1] = g + pp;
          g = c * r - b;
          cs[EV_MAX * (sweep & 1) + cnt] = c;
          sn[EV_MAX * (sweep & 1) + cnt] = s;
          ++cnt;
          dip1 = di;
          ei = en;
          di = dn;
        }
        top = m - 1;
        if (!split) {
          d[l] -= pp;
          e[l] = g;
          e[m] = 0.0;
        }
        break;  // apply this sweep's rotations
      }
      if (lane == 0) {
        s_state = state;
        s_cnt[sweep & 1] = cnt;
        s_top[sweep & 1] = top;
      }
    } else if (sweep > 0 && tid - 64 < n) {
      // wavefronts 1.. apply the previous sweep's rotations to the rows of Z
      // while wavefront 0 forms this sweep's (they touch d, e only).  A sweep
      // that sets the state forms no rotations, so none are left after it.
      const int pb = (sweep - 1) & 1;
      const int cnt = s_cnt[pb], top = s_top[pb];
      const double* csp = cs + EV_MAX * pb;
      const double* snp = sn + EV_MAX * pb;
      // rotation q maps (z_i, z_i+1) of the row, i = top - q, downwards: the
      // entry it leaves at i is the next rotation's z_i+1, carried in f
      double* zr = As[tid - 64];
      if (cnt > 0) {
        double f = zr[top + 1];
        int q = 0;
        // four rotations per group, every operand loaded at the group's start
        for (; q + 4 <= cnt; q += 4) {
          const int i = top - q;
          double c[4], s[4], a[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            c[u] = csp[q + u];
            s[u] = snp[q + u];
            a[u] = zr[i - u];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            zr[i - u + 1] = s[u] * a[u] + c[u] * f;
            f = c[u] * a[u] - s[u] * f;
          }
        }
        for (; q < cnt; ++q) {
          const int i = top - q;
          const double c = csp[q], s = snp[q];
          const double a = zr[i];
          zr[i + 1] = s * a + c * f;
          f = c * a - s * f;
        }
        zr[top - cnt + 1] = f;
      }
    }
    __syncthreads();
    if (s_state != 0) break;
    ++sweep;
  }
  __syncthreads();
  // V = H_0 ... H_{n-3} Z.  Two barriers per reflector: its vector is
  // staged in LDS during the previous reflector's update (v and p alternate).
  {
    int cur = 0;
    bool staged = false;
    for (int k = n - 3; k >= 0; --k) {
      const double t = Hv[k + (int64_t)k * n];
      if (t == 0.0) {  // uniform
        staged = false;
        continue;
      }
      double* vk = cur ? p : v;
      if (!staged) {
        for (int i = k + 1 + tid; i < n; i += NT) vk[i] = Hv[i + (int64_t)k * n];
        __syncthreads();
      }
      {
        const int c = tid >> 2, q = tid & 3;
        double acc = 0.0;
        if (c < n)
          for (int i = k + 1 + q; i < n; i += 4) acc += vk[i] * As[i][c];
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (q == 0 && c < n) w[c] = t * acc;
      }
      __syncthreads();
      if (k > 0) {
        double* vn = cur ? v : p;
        for (int i = k + tid; i < n; i += NT) vn[i] = Hv[i + (int64_t)(k - 1) * n];
      }
      const int L = n - k - 1;
      for (int x = tid; x < L * n; x += NT) {
        const int i = k + 1 + x % L, c = x / L;
        As[i][c] -= vk[i] * w[c];
      }
      __syncthreads();
      cur ^= 1;
      staged = k > 0;
    }
  }
  // sort by eigenvalue (ties by index) and write
  for (int i = tid; i < n; i += NT) {
    const double di = d[i];
    int r = 0;
    for (int j = 0; j < n; ++j) r += (d[j] < di) || (d[j] == di && j < i);
    it.w[r] = di;
    if (r == 0 && it.wmin) *it.wmin = di;
    p[i] = (double)r;
  }
  __syncthreads();
  for (int x = tid; x < n * n; x += NT) {
    const int row = x % n, col = x / n;
    it.V[row + (int64_t)((int)p[col]) * it.ldv] = As[row][col];
  }
}

__global__ __launch_bounds__(512) void syev_vb_kernel(const EigBatch* items) {
  syev_small(items[blockIdx.x]);
}

__global__ __launch_bounds__(512) void syev_one_kernel(EigBatch it) { syev_small(it); }

int syev_vbatched(hipStream_t s, const EigBatch* items, int count) {
  if (count <= 0) return 0;
  syev_vb_kernel<<<count, 512, 0, s>>>(items);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int eigvals_vbatched(hipStream_t s, const EigBatch* items, int count) {
  if (count <= 0) return 0;
  eigvals_vb_kernel<<<count, 512, 0, s>>>(items);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ==========================================================================
// Divide and conquer on the tridiagonal
// ==========================================================================
struct DcNode {
  int lo, mid, hi;      // children [lo, mid), [mid, hi)
  int64_t qoff;         // offset of this node's K x K scratch matrices
};

// Leaves: dense tridiagonal (with Cuppen's diagonal corrections already
// applied to d) solved by Jacobi; eigenvectors into Z's diagonal block.
__global__ __launch_bounds__(256) void dc_leaf_kernel(const int* leaf_lo, const int* leaf_hi,
                                                      const double* d, const double* e,
                                                      double* lam, double* Z, int64_t ldz) {
  __shared__ double As[64][65];
  __shared__ double Vs[64][65];
  const int lo = leaf_lo[blockIdx.x], hi = leaf_hi[blockIdx.x];
  const int n = hi - lo;
  for (int x = threadIdx.x; x < n * n; x += blockDim.x) {
    const int i = x % n, j = x / n;
    double v = 0.0;
    if (i == j) v = d[lo + i];
    else if (i == j + 1) v = e[lo + j];
    else if (j == i + 1) v = e[lo + i];
    As[i][j] = v;
  }
  __syncthreads();
  jacobi_lds(n, As, Vs, true, lam + lo, Z + lo + (int64_t)lo * ldz, ldz);
}

__global__ void dc_split_kernel(int nsplit, const int* split, const double* e, double* d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsplit) return;
  const int m = split[t];  // boundary between m-1 and m
  const double b = fabs(e[m - 1]);
  d[m - 1] -= b;
  d[m] -= b;
}

// Per-node scratch vectors, each of length K at offset node.lo:
//   ds, zs  sorted d and z;  src sorted -> child column;  fpos sorted -> final
//   column of the gathered matrix; dn, zn compacted non-deflated d, z;
//   orig/tauv secular root (origin index, shift); zh Loewner z.
struct DcVec {
  double *ds, *zs, *dn, *zn, *tauv, *zh, *lamv, *rho;
  int *src, *fpos, *orig, *kp, *nrot, *rot_a, *rot_b, *defl;
  double *rot_c, *rot_s;
};

constexpr int DC_LDS_MAX = 4096;

// M1: gather z, merge-sort the children's eigenvalues, deflate (one thread).
__global__ __launch_bounds__(1024) void dc_deflate_kernel(const DcNode* nodes, const double* lam,
                                                          const double* e, const double* Z,
                                                          int64_t ldz, DcVec v) {
  const DcNode nd = nodes[blockIdx.x];
  const int lo = nd.lo, mid = nd.mid, hi = nd.hi;
  const int K = hi - lo, n1 = mid - lo;
  const int tid = threadIdx.x, nt = blockDim.x;
  const double beta = e[mid - 1];
  const double sgn = beta >= 0 ? 1.0 : -1.0;
  const double rho = 2.0 * fabs(beta);
  const double r2 = 0.7071067811865476;
  double* ds = v.ds + lo;
  double* zs = v.zs + lo;
  int* src = v.src + lo;
  // merged order by rank
  for (int c = tid; c < K; c += nt) {
    const double dc = lam[lo + c];
    int rank;
    if (c < n1) {  // count child-2 values strictly below
      int a = 0, b = K - n1;
      while (a < b) { const int m = (a + b) >> 1; if (lam[mid + m] < dc) a = m + 1; else b = m; }
      rank = c + a;
    } else {      // count child-1 values below or equal
      int a = 0, b = n1;
      while (a < b) { const int m = (a + b) >> 1; if (lam[lo + m] <= dc) a = m + 1; else b = m; }
      rank = (c - n1) + a;
    }
    const double z = c < n1 ? Z[(mid - 1) + (int64_t)(lo + c) * ldz]
                            : sgn * Z[mid + (int64_t)(lo + c) * ldz];
    ds[rank] = dc;
    zs[rank] = z * r2;
    src[rank] = c;
  }
  __syncthreads();
  // the serial scan below reads and rewrites ds/zs; keep them in LDS
  __shared__ double lds_ds[DC_LDS_MAX], lds_zs[DC_LDS_MAX];
  const bool use_lds = K <= DC_LDS_MAX;
  if (use_lds)
    for (int c = tid; c < K; c += nt) { lds_ds[c] = ds[c]; lds_zs[c] = zs[c]; }
  __syncthreads();
  double* gds = ds;
  double* gzs = zs;
  if (use_lds) { ds = lds_ds; zs = lds_zs; }
  // the deflation tolerance: max |d| and max |z| over the whole merge
  __shared__ double red[16];
  __shared__ int s_kp, s_nd;
  double dmx = 0.0, zmx = 0.0;
  for (int r = tid; r < K; r += nt) {
    dmx = fmax(dmx, fabs(ds[r]));
    zmx = fmax(zmx, fabs(zs[r]));
  }
  const double dmax = block_reduce<1>(dmx, red);
  const double zmax = block_reduce<1>(zmx, red);
  int* fpos = v.fpos + lo;
  int* defl = v.defl + lo;       // list of deflated sorted positions
  int* ndl = v.orig + lo;        // list of non-deflated sorted positions
  double* dn = v.dn + lo;
  double* zn = v.zn + lo;
  if (tid == 0) {
  const double tol = 8.0 * DEPS * fmax(dmax, zmax);
  int* rota = v.rot_a + lo;
  int* rotb = v.rot_b + lo;
  double* rotc = v.rot_c + lo;
  double* rots = v.rot_s + lo;
  int kp = 0, nd_ = 0, nr = 0;
  if (rho * zmax <= tol) {
    for (int r = 0; r < K; ++r) defl[nd_++] = r;
  } else {
    int pj = -1;
    for (int j = 0; j < K; ++j) {
      if (rho * fabs(zs[j]) <= tol) { defl[nd_++] = j; continue; }
      if (pj < 0) { pj = j; continue; }
      double s = zs[pj], c = zs[j];
      const double t = ds[j] - ds[pj];
      // |t c s| <= tol for the normalised (c, s) reads |t| |c| |s| <= tol
      // (c^2 + s^2): a pair well outside it (a factor 2 beyond any rounding)
      // is kept without forming hypot and the quotients; the exact test
      // below decides every other pair, so the decisions are unchanged
      if (fabs(t) * fabs(c) * fabs(s) > 2.0 * tol * (c * c + s * s)) {
        ndl[kp++] = pj;
        pj = j;
        continue;
      }
      const double tau = hypot(c, s);
      c /= tau;
      s = -s / tau;
      if (fabs(t * c * s) <= tol) {
        zs[j] = tau;
        zs[pj] = 0.0;
        rota[nr] = pj; rotb[nr] = j; rotc[nr] = c; rots[nr] = s; ++nr;
        const double t2 = ds[pj] * c * c + ds[j] * s * s;
        ds[j] = ds[pj] * s * s + ds[j] * c * c;
        ds[pj] = t2;
        defl[nd_++] = pj;
        pj = j;
      } else {
        ndl[kp++] = pj;
        pj = j;
      }
    }
    if (pj >= 0) ndl[kp++] = pj;
  }
  v.kp[blockIdx.x] = kp;
  v.nrot[blockIdx.x] = nr;
  v.rho[blockIdx.x] = rho;
  s_kp = kp;
  s_nd = nd_;
  }
  __syncthreads();
  // positions of the non-deflated values first, then the deflated ones
  {
    const int kp = s_kp, nd_ = s_nd;
    for (int q = tid; q < kp; q += nt) {
      const int r = ndl[q];
      fpos[r] = q;
      dn[q] = ds[r];
      zn[q] = zs[r];
    }
    for (int q = tid; q < nd_; q += nt) fpos[defl[q]] = kp + q;
  }
  __syncthreads();
  if (use_lds)
    for (int c = tid; c < K; c += nt) { gds[c] = lds_ds[c]; gzs[c] = lds_zs[c]; }
}

// M2: gather the children's eigenvector columns into the node's K x K
// matrix Qg in final order (non-deflated first): one thread per element;
// then the deflation rotations, applied row by row in their scan order.
__global__ __launch_bounds__(256) void dc_gather_kernel(const DcNode* nodes, const double* Z,
                                                        int64_t ldz, DcVec v, double* Qg) {
  const DcNode nd = nodes[blockIdx.z];
  const int lo = nd.lo, K = nd.hi - nd.lo;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= K || r >= K) return;
  Qg[nd.qoff + i + (int64_t)v.fpos[lo + r] * K] = Z[(lo + i) + (int64_t)(lo + v.src[lo + r]) * ldz];
}

__global__ __launch_bounds__(256) void dc_rotate_kernel(const DcNode* nodes, DcVec v, double* Qg) {
  const DcNode nd = nodes[blockIdx.y];
  const int lo = nd.lo, K = nd.hi - nd.lo;
  const int nr = v.nrot[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (nr == 0 || i >= K) return;
  double* Q = Qg + nd.qoff;
  const int* fpos = v.fpos + lo;
  for (int t = 0; t < nr; ++t) {
    const int a = fpos[v.rot_a[lo + t]], b = fpos[v.rot_b[lo + t]];
    const double c = v.rot_c[lo + t], s = v.rot_s[lo + t];
    double* x = Q + i + (int64_t)a * K;
    double* y = Q + i + (int64_t)b * K;
    const double xv = *x, yv = *y;
    *x = c * xv + s * yv;
    *y = c * yv - s * xv;
  }
}

// M3: secular equation 1/rho + sum z_i^2/(d_i - lambda) = 0, one wave per
// root, bisection on the shift tau from the nearer pole.
__global__ __launch_bounds__(256) void dc_secular_kernel(const DcNode* nodes, int nnodes,
                                                         const int* root_node,
                                                         const int* root_local, DcVec v) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int node = root_node[gw < 0 ? 0 : gw];
  if (node < 0) return;
  const int j = root_local[gw];
  const int kp = v.kp[node];
  if (j >= kp) return;
  const int lo = nodes[node].lo;
  const double* dn = v.dn + lo;
  const double* zn = v.zn + lo;
  const double rho = v.rho[node];
  const double irho = 1.0 / rho;
  auto f = [&](int o, double tau) {
    const double dof = dn[o];
    double s = 0.0;
    for (int i = lane; i < kp; i += 64) s += zn[i] * zn[i] / ((dn[i] - dof) - tau);
    return irho + wave_sum(s);
  };
  int o;
  double lo_t, hi_t;
  if (j < kp - 1) {
    const double delta = dn[j + 1] - dn[j];
    const double fm = f(j, 0.5 * delta);
    if (fm >= 0) { o = j; lo_t = 0.0; hi_t = 0.5 * delta; }
    else { o = j + 1; lo_t = -0.5 * delta; hi_t = 0.0; }
  } else {
    double s = 0.0;
    for (int i = lane; i < kp; i += 64) s += zn[i] * zn[i];
    s = wave_sum(s);
    o = j; lo_t = 0.0; hi_t = rho * s;
    // make sure f(hi) > 0
    for (int g = 0; g < 8 && f(o, hi_t) <= 0; ++g) hi_t *= 2.0;
  }
  for (int it = 0; it < 400; ++it) {
    const double m = 0.5 * (lo_t + hi_t);
    if (m == lo_t || m == hi_t) break;
    if (hi_t - lo_t <= 2.0 * DEPS * fmax(fabs(lo_t), fabs(hi_t))) break;
    if (f(o, m) > 0) hi_t = m; else lo_t = m;
  }
  if (lane == 0) {
    v.orig[lo + j] = o;
    v.tauv[lo + j] = 0.5 * (lo_t + hi_t);
  }
}

// M4: Loewner formula for z-hat, one wave per non-deflated index.
__global__ __launch_bounds__(256) void dc_lowner_kernel(const DcNode* nodes,
                                                        const int* root_node,
                                                        const int* root_local, DcVec v) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int node = root_node[gw];
  if (node < 0) return;
  const int i = root_local[gw];
  const int kp = v.kp[node];
  if (i >= kp) return;
  const int lo = nodes[node].lo;
  const double* dn = v.dn + lo;
  const double rho = v.rho[node];
  const double di = dn[i];
  // lambda_j - d_i = (d_{o_j} - d_i) + tau_j
  double prod = 1.0;
  for (int j = lane; j < kp; j += 64) {
    const double lmd = (dn[v.orig[lo + j]] - di) + v.tauv[lo + j];
    double den;
    if (j == kp - 1) den = rho;
    else if (j < i) den = dn[j] - di;
    else den = dn[j + 1] - di;
    prod *= lmd / den;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) prod *= __shfl_xor(prod, o, 64);
  if (lane == 0) {
    const double zi = v.zn[lo + i];
    v.zh[lo + i] = copysign(sqrt(fmax(prod, 0.0)), zi);
    v.lamv[lo + i] = dn[v.orig[lo + i]] + v.tauv[lo + i];
  }
}

// M5: normalized eigenvectors of the rank-one update, one workgroup per
// column j: U[i, j] = zh_i / (d_i - lambda_j).
__global__ __launch_bounds__(256) void dc_vectors_kernel(const DcNode* nodes,
                                                         const int* col_node,
                                                         const int* col_local, DcVec v,
                                                         double* Ubuf) {
  __shared__ double red[4];
  const int node = col_node[blockIdx.x];
  if (node < 0) return;
  const int j = col_local[blockIdx.x];
  const int kp = v.kp[node];
  if (j >= kp) return;
  const int lo = nodes[node].lo;
  const double* dn = v.dn + lo;
  const double* zh = v.zh + lo;
  const double dof = dn[v.orig[lo + j]];
  const double tj = v.tauv[lo + j];
  double* U = Ubuf + nodes[node].qoff + (int64_t)j * kp;
  double ss = 0.0;
  for (int i = threadIdx.x; i < kp; i += 256) {
    const double u = zh[i] / ((dn[i] - dof) - tj);
    U[i] = u;
    ss += u * u;
  }
  ss = block_reduce<0>(ss, red);
  const double inv = 1.0 / sqrt(ss);
  for (int i = threadIdx.x; i < kp; i += 256) U[i] *= inv;
}

// M7: final sort of the node's eigenvalues (rank counting, values staged in
// LDS per workgroup) and scatter of the eigenvectors into Z's diagonal block.
__global__ __launch_bounds__(256) void dc_rank_kernel(const DcNode* nodes, DcVec v, double* lam,
                                                      int* colsrc) {
  const DcNode nd = nodes[blockIdx.y];
  const int lo = nd.lo, K = nd.hi - nd.lo;
  const int kp = v.kp[blockIdx.y];
  __shared__ double vals[DC_LDS_MAX];
  const bool use_lds = K <= DC_LDS_MAX;
  auto value = [&](int t) {
    return t < kp ? v.lamv[lo + t] : v.ds[lo + v.defl[lo + t - kp]];
  };
  if (use_lds)
    for (int t = threadIdx.x; t < K; t += blockDim.x) vals[t] = value(t);
  __syncthreads();
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= K) return;
  const double val = use_lds ? vals[q] : value(q);
  int r = 0;
  for (int t = 0; t < K; ++t) {
    const double vt = use_lds ? vals[t] : value(t);
    r += (vt < val) || (vt == val && t < q);
  }
  v.zs[lo + r] = val;  // zs reused as the sorted eigenvalue buffer
  colsrc[lo + r] = q;
}

__global__ __launch_bounds__(256) void dc_scatter_kernel(const DcNode* nodes, DcVec v,
                                                         const double* Qg, const double* Qn,
                                                         const int* colsrc, double* lam,
                                                         double* Z, int64_t ldz) {
  const DcNode nd = nodes[blockIdx.z];
  const int lo = nd.lo, K = nd.hi - nd.lo;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= K) return;
  const int kp = v.kp[blockIdx.z];
  const int q = colsrc[lo + r];
  if (i < K) {
    const double val = q < kp ? Qn[nd.qoff + i + (int64_t)q * K] : Qg[nd.qoff + i + (int64_t)q * K];
    Z[(lo + i) + (int64_t)(lo + r) * ldz] = val;
  }
  if (i == 0) lam[lo + r] = v.zs[lo + r];
}

struct DcPlan {
  std::vector<int> leaf_lo, leaf_hi, splits;
  std::vector<std::vector<DcNode>> levels;  // bottom-up
};

static int dc_build(int lo, int hi, int depth, DcPlan& P, std::vector<std::vector<DcNode>>& bydepth) {
  if (hi - lo <= 64) {
    P.leaf_lo.push_back(lo);
    P.leaf_hi.push_back(hi);
    return depth;
  }
  const int mid = lo + (hi - lo) / 2;
  P.splits.push_back(mid);
  const int d1 = dc_build(lo, mid, depth + 1, P, bydepth);
  const int d2 = dc_build(mid, hi, depth + 1, P, bydepth);
  if ((int)bydepth.size() <= depth) bydepth.resize(depth + 1);
  bydepth[depth].push_back(DcNode{lo, mid, hi, 0});
  return std::max(d1, d2);
}

static DcPlan dc_plan(int n) {
  DcPlan P;
  std::vector<std::vector<DcNode>> bydepth;
  dc_build(0, n, 0, P, bydepth);
  // merge order: deepest first
  for (int dd = (int)bydepth.size() - 1; dd >= 0; --dd) {
    if (bydepth[dd].empty()) continue;
    int64_t off = 0;
    for (auto& nd : bydepth[dd]) {
      nd.qoff = off;
      off += (int64_t)(nd.hi - nd.lo) * (nd.hi - nd.lo);
    }
    P.levels.push_back(bydepth[dd]);
  }
  return P;
}

// workspace layout (doubles): Z (n^2), Qg (n^2), U (n^2), Qn (n^2), 12 n vectors
static size_t dc_work_elems(int n) { return 4 * (size_t)n * n + 16 * (size_t)n + 64; }
// The level's root maps from its per-node deflation counts: roots of node q
// are rmap = q, rloc = 0..kp[q]-1, in node order; entries past the roots
// (the padding to 4-wave blocks) are -1 / 0.  One workgroup.
constexpr int DC_ROOTMAP_MAX = 1024;
__global__ __launch_bounds__(256) void dc_rootmap_kernel(const DcNode* nodes, int nn, const int* kp,
                                                         int padded, int* rmap, int* rloc) {
  __shared__ int off[DC_ROOTMAP_MAX + 1];
  (void)nodes;
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int q = 0; q < nn; ++q) {
      off[q] = acc;
      acc += kp[q];
    }
    off[nn] = acc;
  }
  __syncthreads();
  const int nroots = off[nn];
  for (int r = threadIdx.x; r < padded; r += 256) {
    if (r >= nroots) {
      rmap[r] = -1;
      rloc[r] = 0;
      continue;
    }
    int lo = 0, hi = nn - 1;  // last q with off[q] <= r
    while (lo < hi) {
      const int m = (lo + hi + 1) >> 1;
      if (off[m] <= r) lo = m; else hi = m - 1;
    }
    rmap[r] = lo;
    rloc[r] = r - off[lo];
  }
}

static size_t dc_iwork_elems(int n) { return 24 * (size_t)n + 64; }

// Eigen-decomposition of the symmetric tridiagonal (d, e): lam ascending,
// Z (n x n, ld n).  d is modified.
static int stedc(hipStream_t s, int n, double* d, const double* e, double* lam, double* Z,
                 double* work, int* iwork) {
  DcPlan P = dc_plan(n);
  double* Qg = work;
  double* Ub = Qg + (size_t)n * n;
  double* Qn = Ub + (size_t)n * n;
  double* vb = Qn + (size_t)n * n;
  DcVec v;
  v.ds = vb; v.zs = vb + n; v.dn = vb + 2 * n; v.zn = vb + 3 * n; v.tauv = vb + 4 * n;
  v.zh = vb + 5 * n; v.lamv = vb + 6 * n; v.rot_c = vb + 7 * n; v.rot_s = vb + 8 * n;
  v.rho = vb + 9 * n;
  int* ib = iwork;
  v.src = ib; v.fpos = ib + n; v.orig = ib + 2 * n; v.kp = ib + 3 * n; v.nrot = ib + 4 * n;
  v.rot_a = ib + 5 * n; v.rot_b = ib + 6 * n; v.defl = ib + 7 * n;
  int* colsrc = ib + 8 * n;
  int* leafbuf = ib + 9 * n;        // leaf_lo, leaf_hi (<= n each)
  int* splitbuf = ib + 11 * n;      // <= n
  int* rootmap = ib + 12 * n;       // root_node, root_local: 2 * (n + 3) ints
  int* nodebuf_i = ib + 16 * n;     // node structs (24 bytes each), 8-byte aligned
  DcNode* nodes_dev = reinterpret_cast<DcNode*>(nodebuf_i);

  MSDP_TRY(fill(s, (int64_t)n * n, 0.0, Z));
  const int nleaf = (int)P.leaf_lo.size();
  const int nsplit = (int)P.splits.size();
  // the plan's static data -- leaves, splits and every level's nodes -- goes
  // up at once; the host vectors live until the first level's read of the
  // deflation counts has synchronised the stream, so no other wait is needed
  std::vector<int> hostbuf(2 * nleaf);
  for (int i = 0; i < nleaf; ++i) { hostbuf[i] = P.leaf_lo[i]; hostbuf[nleaf + i] = P.leaf_hi[i]; }
  std::vector<DcNode> allnodes;
  std::vector<int> level_off;
  for (const auto& level : P.levels) {
    level_off.push_back((int)allnodes.size());
    allnodes.insert(allnodes.end(), level.begin(), level.end());
  }
  if ((allnodes.size() * sizeof(DcNode) + sizeof(int) - 1) / sizeof(int) > 8 * (size_t)n + 64) {
    set_error("stedc: too many merge nodes for the workspace");
    return -2;
  }
  MSDP_HIP(hipMemcpyAsync(leafbuf, hostbuf.data(), sizeof(int) * 2 * nleaf, hipMemcpyHostToDevice, s));
  if (nsplit > 0)
    MSDP_HIP(hipMemcpyAsync(splitbuf, P.splits.data(), sizeof(int) * nsplit, hipMemcpyHostToDevice, s));
  if (!allnodes.empty())
    MSDP_HIP(hipMemcpyAsync(nodes_dev, allnodes.data(), sizeof(DcNode) * allnodes.size(),
                            hipMemcpyHostToDevice, s));
  if (nsplit > 0) dc_split_kernel<<<cdiv(nsplit, 256), 256, 0, s>>>(nsplit, splitbuf, e, d);
  dc_leaf_kernel<<<nleaf, 256, 0, s>>>(leafbuf, leafbuf + nleaf, d, e, lam, Z, n);
  MSDP_HIP(hipGetLastError());
  for (size_t li = 0; li < P.levels.size(); ++li) {
    const auto& level = P.levels[li];
    const int nn = (int)level.size();
    const DcNode* nodes = nodes_dev + level_off[li];
    dc_deflate_kernel<<<nn, 1024, 0, s>>>(nodes, lam, e, Z, n, v);
    int maxK = 0;
    for (const auto& nd : level) maxK = std::max(maxK, nd.hi - nd.lo);
    dc_gather_kernel<<<dim3(cdiv(maxK, 256), maxK, nn), 256, 0, s>>>(nodes, Z, n, v, Qg);
    dc_rotate_kernel<<<dim3(cdiv(maxK, 256), nn), 256, 0, s>>>(nodes, v, Qg);
    MSDP_HIP(hipGetLastError());
    // the deflation counts decide the launch sizes below: the one host wait
    // of the level
    std::vector<int> kp(nn);
    MSDP_HIP(hipMemcpyAsync(kp.data(), v.kp, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
    MSDP_TRY(stream_sync(s));
    int nroots = 0;
    for (int q = 0; q < nn; ++q) nroots += kp[q];
    if (nroots > 0) {
      // root -> (node, local) maps, built on the device from the counts;
      // padded to a whole number of 4-wave blocks
      const int padded = cdiv(nroots, 4) * 4;
      if (nn > DC_ROOTMAP_MAX) {
        set_error("stedc: too many merge nodes in one level");
        return -2;
      }
      dc_rootmap_kernel<<<1, 256, 0, s>>>(nodes, nn, v.kp, padded, rootmap, rootmap + padded);
      dc_secular_kernel<<<padded / 4, 256, 0, s>>>(nodes, nn, rootmap, rootmap + padded, v);
      dc_lowner_kernel<<<padded / 4, 256, 0, s>>>(nodes, rootmap, rootmap + padded, v);
      // reuse the maps: one workgroup per column
      dc_vectors_kernel<<<nroots, 256, 0, s>>>(nodes, rootmap, rootmap + padded, v, Ub);
      MSDP_HIP(hipGetLastError());
      for (int q = 0; q < nn; ++q) {
        const int K = level[q].hi - level[q].lo;
        if (kp[q] == 0) continue;
        MSDP_TRY(gemm(s, false, false, K, kp[q], kp[q], 1.0, Qg + level[q].qoff, K,
                      Ub + level[q].qoff, kp[q], 0.0, Qn + level[q].qoff, K));
      }
    }
    dc_rank_kernel<<<dim3(cdiv(maxK, 256), nn), 256, 0, s>>>(nodes, v, lam, colsrc);
    dc_scatter_kernel<<<dim3(cdiv(maxK, 256), maxK, nn), 256, 0, s>>>(nodes, v, Qg, Qn, colsrc,
                                                                       lam, Z, n);
    MSDP_HIP(hipGetLastError());
  }
  return 0;
}

// ==========================================================================
// back-transformation Z <- Q Z with Q = H_0 ... H_{n-3} (compact WY panels)
// ==========================================================================
// T (nb x nb, upper) of a panel of reflectors from their Gram matrix
// G = V'V: T_ii = tau_i, T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i].
// One wavefront; lane r keeps row r of T in registers (rows are independent:
// T[r, i] = -tau_i sum_{k<i} T[r, k] G[k, i]), the columns of G are read
// from LDS in batches, and the column index is a template parameter so every
// register index is static.
template <int I>
__device__ __attribute__((always_inline)) inline void larft_step(int nb, const double* Gs,
                                                                 double taur, double (&tr)[64]) {
  const int r = threadIdx.x & 63;
  if (I < nb) {
    const double ti = readlane_f64(taur, I);
    double v = 0.0;
#pragma unroll
    for (int g = 0; g < I; g += 16) {
      double gc[16];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (g + q < I) gc[q] = Gs[I * 65 + g + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (g + q < I) v += tr[g + q] * gc[q];
    }
    tr[I] = r < I ? -ti * v : (r == I ? ti : 0.0);
  }
  if constexpr (I + 1 < 64) larft_step<I + 1>(nb, Gs, taur, tr);
}

// One workgroup per panel of 64 reflectors (the last may be narrower): G and
// T of panel p at G + 4096 p, T + 4096 p, leading dimension 64.
__global__ __launch_bounds__(64) void larft_kernel(int nref, const double* Gall,
                                                   const double* tauall, double* Tall) {
  __shared__ double Gs[64 * 65];  // column i of G at Gs[i * 65 + k]
  const int r = threadIdx.x;
  const int k0 = blockIdx.x * 64;
  const int nb = min(64, nref - k0);
  const double* G = Gall + (int64_t)blockIdx.x * 4096;
  const double* tau = tauall + k0;
  double* T = Tall + (int64_t)blockIdx.x * 4096;
  double tr[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) tr[i] = (r < nb && i < nb) ? G[r + (int64_t)i * 64] : 0.0;
  const double taur = r < nb ? tau[r] : 0.0;
#pragma unroll
  for (int i = 0; i < 64; ++i) Gs[i * 65 + r] = tr[i];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 64; ++i) tr[i] = 0.0;
  larft_step<0>(nb, Gs, taur, tr);
  if (r >= nb) return;
#pragma unroll
  for (int i = 0; i < 64; ++i)
    if (i < nb) T[r + (int64_t)i * 64] = tr[i];
}

// Panels of 64 reflectors, applied last to first.  The Gram matrices
// G_p = V_p' V_p of all panels are one batched product and their T factors
// one launch, before the sequential updates of Z.  Column k of V is zero in
// rows <= k, so every full panel's Gram is taken over rows 1..n-1 (the extra
// rows add exact zeros).
static size_t ormtr_work_elems(int n) {
  const size_t npanel = (size_t)cdiv(std::max(n - 2, 1), 64);
  return 2 * 4096 * npanel + 2 * 64 * (size_t)n;
}

static int ormtr(hipStream_t s, int n, const double* V, const double* tau, double* Z, int64_t ldz,
                 double* work) {
  const int nref = n - 2;
  if (nref <= 0) return 0;
  const int nb = 64;
  const int npanel = cdiv(nref, nb);
  double* G = work;                          // npanel x (64 x 64)
  double* T = G + (size_t)4096 * npanel;     // npanel x (64 x 64)
  double* W1 = T + (size_t)4096 * npanel;    // nb x n
  double* W2 = W1 + (size_t)nb * n;          // nb x n
  const int nfull = nref / nb;
  if (nfull > 0)
    MSDP_TRY(gemm(s, true, false, nb, nb, n - 1, 1.0, V + 1, n, V + 1, n, 0.0, G, nb, nfull,
                  (int64_t)nb * n, (int64_t)nb * n, 4096));
  if (nfull < npanel) {
    const int k0 = nfull * nb, kb = nref - k0, r0 = k0 + 1;
    const double* Vp = V + r0 + (int64_t)k0 * n;
    MSDP_TRY(gemm(s, true, false, kb, kb, n - r0, 1.0, Vp, n, Vp, n, 0.0,
                  G + (size_t)4096 * nfull, nb));
  }
  larft_kernel<<<npanel, 64, 0, s>>>(nref, G, tau, T);
  MSDP_HIP(hipGetLastError());
  for (int pi = npanel - 1; pi >= 0; --pi) {
    const int k0 = pi * nb;
    const int kb = std::min(nb, nref - k0);
    const int r0 = k0 + 1;                // first row touched
    const int rows = n - r0;
    const double* Vp = V + r0 + (int64_t)k0 * n;  // rows x kb, ld n
    const double* Tp = T + (size_t)4096 * pi;
    MSDP_TRY(gemm(s, true, false, kb, n, rows, 1.0, Vp, n, Z + r0, ldz, 0.0, W1, kb));
    MSDP_TRY(gemm(s, false, false, kb, n, kb, 1.0, Tp, nb, W1, kb, 0.0, W2, kb));
    MSDP_TRY(gemm(s, false, false, rows, n, kb, -1.0, Vp, n, W2, kb, 1.0, Z + r0, ldz));
  }
  return 0;
}

// ==========================================================================
// public entry points
// ==========================================================================
size_t eig_min_work_elems(int n);
size_t syevd_work_elems(int n) {
  return std::max(trd_work_elems(n) + (size_t)n * n +
                      std::max(dc_work_elems(n), ormtr_work_elems(n)) + 64,
                  2 * eig_min_work_elems(n));  // room for eig_min2
}
size_t syevd_iwork_elems(int n) { return dc_iwork_elems(n); }

int syevd(hipStream_t s, int n, const double* A, int64_t lda, double* w, double* V, int64_t ldv,
          bool vectors, double* work, int* iwork) {
  MSDP_TRY(rec_forbid("syevd"));
  if (n <= 0) return 0;
  if (n <= EV_MAX && V) {
    syev_one_kernel<<<1, 512, 0, s>>>(EigBatch{n, A, lda, w, nullptr, V, ldv, work});
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  if (n <= 64)
    return jacobi_batched(s, n, A, lda, 0, w, 0, V, ldv, 0, 1, vectors);
  TrdWork t = trd_carve(n, work);
  double* rest = work + trd_work_elems(n);
  MSDP_TRY(sytrd(s, n, A, lda, t));
  if (!vectors) {
    // eigenvalues only: divide and conquer still needs the vectors' boundary
    // rows, so the full path is used; the vectors are discarded.
  }
  double* Zt = rest;
  double* dcw = Zt + (size_t)n * n;
  MSDP_TRY(stedc(s, n, t.d, t.e, w, Zt, dcw, iwork));
  if (vectors) {
    MSDP_TRY(copy_mat(s, n, n, Zt, n, V, ldv));
    MSDP_TRY(ormtr(s, n, t.V, t.tau, V, ldv, dcw));
  }
  return 0;
}

// ==========================================================================
// Lanczos for lambda_min of an explicit symmetric matrix (the step-length
// eigenvalue of src/step.jl:180-184, which the reference obtains with
// LowRankOpt's Lanczos).  Full reorthogonalisation (classical Gram-Schmidt
// applied twice); the Ritz value is accepted once the residual bound
// beta_j |y_last| of the smallest Ritz pair is below LZ_RTOL |theta|, when
// the Krylov space becomes invariant, or when it spans the whole space.
// ==========================================================================
constexpr int LZ_MIN_N = 400;     // below this, tridiagonalisation + Sturm
constexpr double LZ_RTOL = 1e-10;
constexpr int LZ_CHECK = 16;

__device__ inline double lz_hash(int64_t i) {
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// The multi-workgroup kernels below run one Lanczos process per instance:
// instance y = blockIdx.y (blockIdx.x for lz_init) + inst0, whose work
// arrays lie y * ws elements after those passed (and whose matrix is A1
// for y = 1), so that two step-length matrices share every launch.
__global__ __launch_bounds__(1024) void lz_init_kernel(int n, double* q, int64_t ws, int inst0) {
  __shared__ double red[16];
  q += (blockIdx.x + inst0) * ws;
  double ss = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = lz_hash(i);
    q[i] = v;
    ss += v * v;
  }
  ss = block_reduce<0>(ss, red);
  const double inv = 1.0 / sqrt(ss);
  for (int i = threadIdx.x; i < n; i += blockDim.x) q[i] *= inv;
}

// w = A q (A symmetric: column c's dot with q gives w[c]).  LZ_GEMV_COLS
// columns per workgroup: every thread keeps its rows of q in registers for
// all of them and has LZ_GEMV_COLS x 4 loads of A in flight; the per-column
// sums are added over the threads in a fixed order.
constexpr int LZ_GEMV_COLS = 8;

__global__ __launch_bounds__(256) void lz_gemv_kernel(int n, const double* A, const double* A1,
                                                      int64_t lda, const double* q, double* w,
                                                      int64_t ws, int inst0) {
  constexpr int C = LZ_GEMV_COLS, U = 4;
  __shared__ double red[C][4];
  const int y = blockIdx.y + inst0;
  if (y) A = A1;
  q += y * ws;
  w += y * ws;
  const int c0 = blockIdx.x * C;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double* cols[C];
#pragma unroll
  for (int k = 0; k < C; ++k) cols[k] = A + (int64_t)min(c0 + k, n - 1) * lda;
  double acc[C];
#pragma unroll
  for (int k = 0; k < C; ++k) acc[k] = 0.0;
  for (int r0 = 0; r0 < n; r0 += 256 * U) {
    double qv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = r0 + tid + 256 * u;
      qv[u] = i < n ? q[i] : 0.0;
    }
    // rows past n read row n - 1 (a finite entry) against q = 0: every load
    // is unconditional, so all of them are in flight together
#pragma unroll
    for (int k = 0; k < C; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u)
        acc[k] = fma(cols[k][min(r0 + tid + 256 * u, n - 1)], qv[u], acc[k]);
  }
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const double t = wave_sum(acc[k]);
    if (lane == 0) red[k][wv] = t;
  }
  __syncthreads();
  if (tid < C && c0 + tid < n)
    w[c0 + tid] = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
}

// h[c] = Q[:, c]' w, one workgroup per column
__global__ __launch_bounds__(256) void lz_dots_kernel(int n, const double* Q, const double* w,
                                                      double* h, int64_t ws, int inst0) {
  __shared__ double red[4];
  const int64_t off = (blockIdx.y + inst0) * ws;
  Q += off; w += off; h += off;
  const int c = blockIdx.x;
  const double* col = Q + (int64_t)c * n;
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += col[i] * w[i];
  v = block_reduce<0>(v, red);
  if (threadIdx.x == 0) h[c] = v;
}

// w -= Q[:, 0:ncols] h: 64 rows per workgroup, 4 lanes per row split the
// columns, partial sums combined in LDS.
__global__ __launch_bounds__(256) void lz_update_kernel(int n, int ncols, const double* Q,
                                                        const double* h, double* w, int64_t ws,
                                                        int inst0) {
  __shared__ double part[4][64];
  const int64_t off = (blockIdx.y + inst0) * ws;
  Q += off; h += off; w += off;
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  double v = 0.0;
  if (i < n)
    for (int c = g; c < ncols; c += 4) v += Q[i + (int64_t)c * n] * h[c];
  part[g][r] = v;
  __syncthreads();
  if (g == 0 && i < n) w[i] -= (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
}

// alpha_j = h1[j] + h2[j], beta_j = ||w||, q_{j+1} = w / beta_j
__global__ __launch_bounds__(1024) void lz_finish_kernel(int n, int j, const double* w,
                                                         const double* h1, const double* h2,
                                                         double* alpha, double* beta, double* qn,
                                                         int64_t ws, int inst0) {
  __shared__ double red[16];
  const int64_t off = (blockIdx.y + inst0) * ws;
  w += off; h1 += off; h2 += off; alpha += off; beta += off; qn += off;
  double ss = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) ss += w[i] * w[i];
  ss = block_reduce<0>(ss, red);
  const double b = sqrt(ss);
  const double inv = b > 0 ? 1.0 / b : 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) qn[i] = w[i] * inv;
  if (threadIdx.x == 0) {
    alpha[j] = h1[j] + h2[j];
    beta[j] = b;
  }
}

// Smallest eigenvalue theta of the k x k Lanczos tridiagonal (alpha, beta,
// in LDS) by Sturm multisection over NT points per round, and the residual
// bound beta_{k-1} |y_{k-1}| of its normalized eigenvector y (inverse
// iteration with partial pivoting, one thread).  Called by every thread of
// the workgroup; out[0] = theta, out[1] = residual bound, out[2] =
// |beta_{k-1}| are written by thread 0 (LDS or global).
struct LzShared {
  double lo, hi;
  int first;
};

template <int NT>
__device__ void lz_ritz_min(int k, const double* alpha, const double* beta, double* scratch,
                            double* red, LzShared& sh, double* out) {
  const int tid = threadIdx.x;
  double gl = INFINITY, gu = -INFINITY, emax = 0.0;
  for (int i = tid; i < k; i += NT) {
    const double r = (i > 0 ? fabs(beta[i - 1]) : 0.0) + (i < k - 1 ? fabs(beta[i]) : 0.0);
    gl = fmin(gl, alpha[i] - r);
    gu = fmax(gu, alpha[i] + r);
    if (i < k - 1) emax = fmax(emax, beta[i] * beta[i]);
  }
  gl = block_reduce<2>(gl, red);
  gu = block_reduce<1>(gu, red);
  emax = block_reduce<1>(emax, red);
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  if (tid == 0) {
    sh.lo = gl - 2 * DEPS * tnorm - 2 * pivmin;
    sh.hi = gu + 2 * DEPS * tnorm + 2 * pivmin;
  }
  __syncthreads();
  for (int round = 0; round < 12; ++round) {
    const double lo = sh.lo, hi = sh.hi;
    if (hi - lo <= 2 * DEPS * fmax(fabs(lo), fabs(hi)) + 4 * pivmin) break;
    __syncthreads();
    if (tid == 0) sh.first = NT;
    __syncthreads();
    const double x = lo + (hi - lo) * (double)(tid + 1) / (double)(NT + 1);
    if (sturm_count(k, alpha, beta, x, pivmin) >= 1) atomicMin(&sh.first, tid);
    __syncthreads();
    if (tid == 0) {
      const int f = sh.first;
      sh.lo = f == 0 ? lo : lo + (hi - lo) * (double)f / (double)(NT + 1);
      sh.hi = f == NT ? hi : lo + (hi - lo) * (double)(f + 1) / (double)(NT + 1);
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double theta = 0.5 * (sh.lo + sh.hi);
    // inverse iteration on (T - theta I) y = b with partial pivoting
    double* u0 = scratch;           // k
    double* u1 = scratch + k;       // k
    double* u2 = scratch + 2 * k;   // k
    double* y = scratch + 3 * k;    // k
    double* bb = scratch + 4 * k;   // k
    double* l = scratch + 5 * k;    // multipliers k
    int* sw = reinterpret_cast<int*>(scratch + 6 * k);  // swaps k
    const double tiny = DEPS * fmax(tnorm, DBL_MIN);
    for (int i = 0; i < k; ++i) y[i] = 1.0;
    for (int pass = 0; pass < 3; ++pass) {
      // factor
      double dcur = alpha[0] - theta, ucur = k > 1 ? beta[0] : 0.0;
      for (int i = 0; i < k; ++i) bb[i] = y[i];
      for (int i = 0; i < k - 1; ++i) {
        const double sub = beta[i];
        const double dn = alpha[i + 1] - theta;
        const double un = i + 1 < k - 1 ? beta[i + 1] : 0.0;
        if (fabs(dcur) >= fabs(sub)) {
          const double f = dcur != 0.0 ? sub / dcur : 0.0;
          u0[i] = dcur; u1[i] = ucur; u2[i] = 0.0; l[i] = f; sw[i] = 0;
          dcur = dn - f * ucur;
          ucur = un;
        } else {
          const double f = dcur / sub;
          u0[i] = sub; u1[i] = dn; u2[i] = un; l[i] = f; sw[i] = 1;
          dcur = ucur - f * dn;
          ucur = -f * un;
        }
      }
      u0[k - 1] = dcur;
      // forward
      for (int i = 0; i < k - 1; ++i) {
        if (sw[i]) { const double t = bb[i]; bb[i] = bb[i + 1]; bb[i + 1] = t - l[i] * bb[i]; }
        else bb[i + 1] -= l[i] * bb[i];
      }
      // backward
      for (int i = k - 1; i >= 0; --i) {
        double v = bb[i];
        if (i + 1 < k) v -= u1[i] * y[i + 1];
        if (i + 2 < k) v -= u2[i] * y[i + 2];
        double piv = u0[i];
        if (fabs(piv) < tiny) piv = piv < 0 ? -tiny : tiny;
        y[i] = v / piv;
      }
      double nrm = 0.0;
      for (int i = 0; i < k; ++i) nrm = fmax(nrm, fabs(y[i]));
      if (!(nrm > 0) || !isfinite(nrm)) break;
      for (int i = 0; i < k; ++i) y[i] /= nrm;
    }
    double n2 = 0.0;
    for (int i = 0; i < k; ++i) n2 += y[i] * y[i];
    const double ylast = fabs(y[k - 1]) / sqrt(n2);
    out[0] = theta;
    out[1] = isfinite(ylast) ? fabs(beta[k - 1]) * ylast : INFINITY;
    out[2] = fabs(beta[k - 1]);
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void lz_check_kernel(int k, const double* alpha_g,
                                                       const double* beta_g, double* scratch_g,
                                                       double* out, int64_t ws, int inst0) {
  __shared__ double red[4];
  // the tridiagonal and the inverse-iteration arrays live in LDS (k <= 1024)
  __shared__ double alpha[1024], beta[1024];
  __shared__ double scratch[7 * 1024];
  __shared__ LzShared sh;
  (void)scratch_g;
  const int64_t off = (blockIdx.y + inst0) * ws;
  alpha_g += off; beta_g += off; out += off;
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    alpha[i] = alpha_g[i];
    beta[i] = beta_g[i];
  }
  __syncthreads();
  lz_ritz_min<256>(k, alpha, beta, scratch, red, sh, out);
}

// The whole Lanczos lambda_min of lanczos_min (same start vector, full
// reorthogonalisation by classical Gram-Schmidt applied twice, the same
// check every LZ_CHECK steps and the same stopping rules) in one workgroup,
// for n <= LZ1_MAX: one launch and no host round trips instead of six
// launches per step and a host read per check.  The Krylov basis Q (n x
// (n + 1)) stays in global memory (L2-resident at these sizes), the current
// vectors and the tridiagonal in LDS.
constexpr int LZ1_MAX = 640;
constexpr int LZ1_NT = 1024;

// workgroup 1 (when launched): A1, its work ws elements further on, into out1
__global__ __launch_bounds__(LZ1_NT) void lz1_kernel(int n, const double* A, int64_t lda,
                                                    double* Q, double* out, const double* A1,
                                                    int64_t ws, double* out1) {
  if (blockIdx.x) {
    A = A1; Q += ws; out = out1;
  }
  __shared__ double w[LZ1_MAX], q[LZ1_MAX], h[LZ1_MAX + 1];
  __shared__ double alpha[LZ1_MAX], beta[LZ1_MAX];
  __shared__ double scratch[7 * LZ1_MAX];
  __shared__ double red[16];
  __shared__ double chk[3];
  __shared__ LzShared sh;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = LZ1_NT / 64;
  const int kmax = n;
  // q_0: the hashed start vector, normalised
  double ss = 0.0;
  for (int i = tid; i < n; i += LZ1_NT) {
    const double v = lz_hash(i);
    q[i] = v;
    ss += v * v;
  }
  ss = block_reduce<0>(ss, red);
  {
    const double inv = 1.0 / sqrt(ss);
    for (int i = tid; i < n; i += LZ1_NT) {
      q[i] *= inv;
      Q[i] = q[i];
    }
  }
  __syncthreads();
  double anorm_est = 0.0;
  double theta = 0.0;
  for (int j = 0; j < kmax; ++j) {
    // w = A q_j (A symmetric: row i = column i), one row per thread
    for (int i = tid; i < n; i += LZ1_NT) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int c = 0;
      for (; c + 3 < n; c += 4) {
        a0 += A[i + (int64_t)c * lda] * q[c];
        a1 += A[i + (int64_t)(c + 1) * lda] * q[c + 1];
        a2 += A[i + (int64_t)(c + 2) * lda] * q[c + 2];
        a3 += A[i + (int64_t)(c + 3) * lda] * q[c + 3];
      }
      for (; c < n; ++c) a0 += A[i + (int64_t)c * lda] * q[c];
      w[i] = (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
    double hj = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      // h = Q[:, 0:j+1]' w, one wavefront per column
      for (int c = wave; c <= j; c += NW) {
        const double* col = Q + (int64_t)c * n;
        double v = 0.0;
        for (int i = lane; i < n; i += 64) v += col[i] * w[i];
        v = wave_sum(v);
        if (lane == 0) h[c] = v;
      }
      __syncthreads();
      // w -= Q h
      for (int i = tid; i < n; i += LZ1_NT) {
        double v = 0.0;
        for (int c = 0; c <= j; ++c) v += Q[i + (int64_t)c * n] * h[c];
        w[i] -= v;
      }
      hj += h[j];
      __syncthreads();
    }
    // alpha_j, beta_j = ||w||, q_{j+1} = w / beta_j
    double s2 = 0.0;
    for (int i = tid; i < n; i += LZ1_NT) s2 += w[i] * w[i];
    s2 = block_reduce<0>(s2, red);
    const double b = sqrt(s2);
    const double inv = b > 0 ? 1.0 / b : 0.0;
    if (tid == 0) {
      alpha[j] = hj;
      beta[j] = b;
    }
    for (int i = tid; i < n; i += LZ1_NT) {
      q[i] = w[i] * inv;
      Q[i + (int64_t)(j + 1) * n] = q[i];
    }
    __syncthreads();
    const int k = j + 1;
    if (k % LZ_CHECK == 0 || k == kmax) {
      lz_ritz_min<LZ1_NT>(k, alpha, beta, scratch, red, sh, chk);
      theta = chk[0];
      anorm_est = fmax(anorm_est, fabs(theta));
      const bool invariant = chk[2] <= 1e-14 * fmax(anorm_est, 1e-300);
      if (chk[1] <= LZ_RTOL * fmax(fabs(theta), 1e-300) || invariant || k == n) break;
    }
  }
  if (tid == 0) out[0] = theta;
}

// count (1 or 2) Lanczos processes in the same launches: matrix A0 (and A1)
// into out0 (and out1), instance y's work at work + y * ws.  A process that
// has converged drops out of the launches.
static int lanczos_min(hipStream_t s, int n, const double* A0, const double* A1, int64_t lda,
                       double* out0, double* out1, double* work, int64_t ws, int count) {
  const int kmax = std::min(n, 1024);
  double* Q = work;                                  // n x (kmax + 1)
  double* w = Q + (size_t)n * (kmax + 1);            // n
  double* h1 = w + n;                                // kmax + 1
  double* h2 = h1 + kmax + 1;                        // kmax + 1
  double* al = h2 + kmax + 1;                        // kmax
  double* be = al + kmax;                            // kmax
  double* chk = be + kmax;                           // 4
  double* scr = chk + 4;                             // 7 kmax
  bool active[2] = {true, count > 1};
  double* outs[2] = {out0, out1};
  double anorm_est[2] = {0.0, 0.0};
  lz_init_kernel<<<count, 1024, 0, s>>>(n, Q, ws, 0);
  MSDP_HIP(hipGetLastError());
  for (int j = 0; j < kmax; ++j) {
    const int inst0 = active[0] ? 0 : 1;
    const int ny = (active[0] && active[1]) ? 2 : 1;
    const double* qj = Q + (size_t)j * n;
    lz_gemv_kernel<<<dim3(cdiv(n, LZ_GEMV_COLS), ny), 256, 0, s>>>(n, A0, A1, lda, qj, w, ws,
                                                                   inst0);
    lz_dots_kernel<<<dim3(j + 1, ny), 256, 0, s>>>(n, Q, w, h1, ws, inst0);
    lz_update_kernel<<<dim3(cdiv(n, 64), ny), 256, 0, s>>>(n, j + 1, Q, h1, w, ws, inst0);
    lz_dots_kernel<<<dim3(j + 1, ny), 256, 0, s>>>(n, Q, w, h2, ws, inst0);
    lz_update_kernel<<<dim3(cdiv(n, 64), ny), 256, 0, s>>>(n, j + 1, Q, h2, w, ws, inst0);
    lz_finish_kernel<<<dim3(1, ny), 1024, 0, s>>>(n, j, w, h1, h2, al, be,
                                                  Q + (size_t)(j + 1) * n, ws, inst0);
    MSDP_HIP(hipGetLastError());
    const int k = j + 1;
    if (k % LZ_CHECK == 0 || k == kmax) {
      lz_check_kernel<<<dim3(1, ny), 256, 0, s>>>(k, al, be, scr, chk, ws, inst0);
      MSDP_HIP(hipGetLastError());
      double hv[2][3] = {{0, 0, 0}, {0, 0, 0}};
      for (int y = 0; y < 2; ++y)
        if (active[y])
          MSDP_HIP(hipMemcpyAsync(hv[y], chk + y * ws, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                  s));
      MSDP_TRY(stream_sync(s));
      for (int y = 0; y < 2; ++y) {
        if (!active[y]) continue;
        anorm_est[y] = std::max(anorm_est[y], std::fabs(hv[y][0]));
        const bool invariant = hv[y][2] <= 1e-14 * std::max(anorm_est[y], 1e-300);
        if (hv[y][1] <= LZ_RTOL * std::max(std::fabs(hv[y][0]), 1e-300) || invariant || k == n ||
            k == kmax) {
          MSDP_HIP(hipMemcpyAsync(outs[y], chk + y * ws, sizeof(double), hipMemcpyDeviceToDevice,
                                  s));
          active[y] = false;
        }
      }
      if (!active[0] && !active[1]) break;
    }
  }
  return 0;
}

size_t eig_min_work_elems(int n) {
  const size_t kmax = std::min(n, 1024);
  return std::max(trd_work_elems(n), (size_t)n * (kmax + 2) + 12 * kmax + 64);
}

// count (1 or 2) matrices of side n: instance y's work at work + y * ws,
// ws = eig_min_work_elems(n); the two share every launch.
static int eig_min_count(hipStream_t s, int n, const double* A0, const double* A1, int64_t lda,
                         double* out0, double* out1, double* work, int count) {
  if (n <= 0) return 0;
  const int64_t ws = (int64_t)eig_min_work_elems(n);
  if (count == 1) {
    A1 = A0;
    out1 = out0;
  }
  if (n <= 64) {
    for (int y = 0; y < count; ++y) {
      double* wv = work + y * ws;
      MSDP_TRY(jacobi_batched(s, n, y ? A1 : A0, lda, 0, wv, 0, nullptr, 0, 0, 1, false));
      MSDP_HIP(hipMemcpyAsync(y ? out1 : out0, wv, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    return 0;
  }
  if (n >= LZ_MIN_N && n <= LZ1_MAX) {
    lz1_kernel<<<count, LZ1_NT, 0, s>>>(n, A0, lda, work, out0, A1, ws, out1);
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  if (n >= LZ_MIN_N) return lanczos_min(s, n, A0, A1, lda, out0, out1, work, ws, count);
  // n < LZ_MIN_N <= TRD1_MAX: the one-workgroup tridiagonalisation, then Sturm
  TrdWork t = trd_carve(n, work);
  MSDP_TRY(launch_trd1(s, count, n, A0, lda, t, A1, ws));
  tridiag_min_kernel<<<count, 256, 0, s>>>(n, t.d, t.e, out0, ws, out1);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int eig_min(hipStream_t s, int n, const double* A, int64_t lda, double* out, double* work, int*) {
  MSDP_TRY(rec_forbid("eig_min"));
  return eig_min_count(s, n, A, nullptr, lda, out, nullptr, work, 1);
}

int eig_min2(hipStream_t s, int n, const double* A0, const double* A1, int64_t lda, double* out0,
             double* out1, double* work) {
  MSDP_TRY(rec_forbid("eig_min2"));
  return eig_min_count(s, n, A0, A1, lda, out0, out1, work, 2);
}

}  // namespace msdp
