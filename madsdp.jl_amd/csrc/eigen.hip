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
// Householder tridiagonalisation (lower), A = Q T Q',  Q = H_0 ... H_{n-3}
// ==========================================================================
// Reflector kernel for step k (k = -1 for the first): forms w_k from p_k,
// then the column j = k+1 of the updated matrix, its reflector v_j (stored
// as column j of V, zero above row j+1) and d_j, e_j, tau_j.  It also zeroes
// the accumulation buffer of the next product.
__global__ __launch_bounds__(1024) void trd_reflector_kernel(int n, int k, const double* A,
                                                             int64_t lda, double* V, double* tau,
                                                             double* d, double* e, double* w,
                                                             const double* p, double* pnext) {
  __shared__ double red[16];
  __shared__ double s_wj, s_vj, s_alpha, s_beta, s_tau, s_scale;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int j = k + 1;
  const double* vk = k >= 0 ? V + (int64_t)k * n : nullptr;
  // w_k = tau p - (tau^2/2)(p'v) v  over indices >= k+1
  if (k >= 0) {
    const double tk = tau[k];
    double sdot = 0.0;
    for (int i = k + 1 + tid; i < n; i += nt) sdot += p[i] * vk[i];
    sdot = block_reduce<0>(sdot, red);
    const double c2 = 0.5 * tk * tk * sdot;
    for (int i = tid; i < n; i += nt) w[i] = i >= k + 1 ? tk * p[i] - c2 * vk[i] : 0.0;
    __syncthreads();
    if (tid == 0) { s_wj = w[j]; s_vj = vk[j]; }
    __syncthreads();
  }
  for (int i = tid; i < n; i += nt) pnext[i] = 0.0;
  // column j of the updated matrix, rows >= j
  auto col = [&](int i) -> double {
    double c = A[i + (int64_t)j * lda];
    if (k >= 0) c -= vk[i] * s_wj + w[i] * s_vj;
    return c;
  };
  if (tid == 0) d[j] = col(j);
  if (j == n - 2) {
    if (tid == 0) {
      e[j] = col(n - 1);
      tau[j] = 0.0;
      double dl = A[(n - 1) + (int64_t)(n - 1) * lda];
      if (k >= 0) dl -= 2.0 * vk[n - 1] * w[n - 1];
      d[n - 1] = dl;
    }
    return;
  }
  if (j >= n - 1) return;
  // Householder of x = col(j+1 .. n-1)
  double ss = 0.0;
  for (int i = j + 2 + tid; i < n; i += nt) {
    const double c = col(i);
    ss += c * c;
  }
  ss = block_reduce<0>(ss, red);
  if (tid == 0) {
    const double alpha = col(j + 1);
    double beta, t, sc;
    if (ss == 0.0) {
      beta = alpha; t = 0.0; sc = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + ss), alpha);
      t = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    s_alpha = alpha; s_beta = beta; s_tau = t; s_scale = sc;
    tau[j] = t;
    e[j] = beta;
  }
  __syncthreads();
  double* vj = V + (int64_t)j * n;
  const double sc = s_scale;
  for (int i = tid; i < n; i += nt) {
    double v;
    if (i <= j) v = 0.0;
    else if (i == j + 1) v = 1.0;
    else v = col(i) * sc;
    vj[i] = v;
  }
}

// Fused kernel of step k: applies the pending update of step k-1,
// A <- A - v w' - w v' on [k+1, n)^2, and accumulates p_k = A v_k there.
// One thread per row, 64 columns per workgroup.
__global__ __launch_bounds__(256) void trd_update_product_kernel(int n, int k, double* A,
                                                                 int64_t lda, const double* V,
                                                                 const double* w, double* p) {
  const int base = k + 1;
  const int L = n - base;
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * 64;
  __shared__ double sv[64], sw[64], sn[64];
  const double* vprev = k >= 1 ? V + (int64_t)(k - 1) * n : nullptr;
  const double* vk = V + (int64_t)k * n;
  if (threadIdx.x < 64) {
    const int c = base + c0 + threadIdx.x;
    const bool ok = c0 + (int)threadIdx.x < L;
    sv[threadIdx.x] = (ok && vprev) ? vprev[c] : 0.0;
    sw[threadIdx.x] = (ok && vprev) ? w[c] : 0.0;
    sn[threadIdx.x] = ok ? vk[c] : 0.0;
  }
  __syncthreads();
  if (r >= L) return;
  const int i = base + r;
  const double vi = vprev ? vprev[i] : 0.0, wi = vprev ? w[i] : 0.0;
  const int cend = min(64, L - c0);
  double acc = 0.0;
  double* colp = A + i + (int64_t)(base + c0) * lda;
  if (vprev) {
    for (int cc = 0; cc < cend; ++cc) {
      double a = colp[(int64_t)cc * lda];
      a -= vi * sw[cc] + wi * sv[cc];
      colp[(int64_t)cc * lda] = a;
      acc += a * sn[cc];
    }
  } else {
    for (int cc = 0; cc < cend; ++cc) acc += colp[(int64_t)cc * lda] * sn[cc];
  }
  atomicAdd(p + i, acc);
}

struct TrdWork {
  double *A, *V, *tau, *d, *e, *w, *p0, *p1;
};

static size_t trd_work_elems(int n) { return 2 * (size_t)n * n + 7 * (size_t)n; }

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
  return t;
}

static int sytrd(hipStream_t s, int n, const double* A, int64_t lda, const TrdWork& t) {
  MSDP_TRY(copy_mat(s, n, n, A, lda, t.A, n));
  MSDP_TRY(fill(s, (int64_t)n * n, 0.0, t.V));
  // k = -1: first reflector from column 0
  trd_reflector_kernel<<<1, 1024, 0, s>>>(n, -1, t.A, n, t.V, t.tau, t.d, t.e, t.w, t.p1, t.p0);
  MSDP_HIP(hipGetLastError());
  for (int k = 0; k <= n - 3; ++k) {
    double* pk = (k & 1) ? t.p1 : t.p0;
    double* pn = (k & 1) ? t.p0 : t.p1;
    const int L = n - k - 1;
    dim3 g(cdiv(L, 256), cdiv(L, 64));
    const bool prof = prof_active(PROF_TRD, s);
    if (prof) prof_before(s);
    trd_update_product_kernel<<<g, 256, 0, s>>>(n, k, t.A, n, t.V, t.w, pk);
    // algorithmic bytes: the trailing L x L matrix read (and written back
    // when an update is pending)
    if (prof) prof_after(s, 8.0 * L * (double)L * (k >= 1 ? 2.0 : 1.0));
    trd_reflector_kernel<<<1, 1024, 0, s>>>(n, k, t.A, n, t.V, t.tau, t.d, t.e, t.w, pk, pn);
  }
  MSDP_HIP(hipGetLastError());
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

__global__ __launch_bounds__(256) void tridiag_min_kernel(int n, const double* d, const double* e,
                                                          double* out) {
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
  if (tid != 0) return;
  double dmax = 0.0, zmax = 0.0;
  for (int r = 0; r < K; ++r) { dmax = fmax(dmax, fabs(ds[r])); zmax = fmax(zmax, fabs(zs[r])); }
  const double tol = 8.0 * DEPS * fmax(dmax, zmax);
  int* fpos = v.fpos + lo;
  int* defl = v.defl + lo;       // list of deflated sorted positions
  int* rota = v.rot_a + lo;
  int* rotb = v.rot_b + lo;
  double* rotc = v.rot_c + lo;
  double* rots = v.rot_s + lo;
  double* dn = v.dn + lo;
  double* zn = v.zn + lo;
  int kp = 0, nd_ = 0, nr = 0;
  // non-deflated positions are collected into fpos[0..kp) temporarily
  int* ndl = v.orig + lo;  // temp list of non-deflated sorted positions
  if (rho * zmax <= tol) {
    for (int r = 0; r < K; ++r) defl[nd_++] = r;
  } else {
    int pj = -1;
    for (int j = 0; j < K; ++j) {
      if (rho * fabs(zs[j]) <= tol) { defl[nd_++] = j; continue; }
      if (pj < 0) { pj = j; continue; }
      double s = zs[pj], c = zs[j];
      const double tau = hypot(c, s);
      const double t = ds[j] - ds[pj];
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
  for (int q = 0; q < kp; ++q) {
    const int r = ndl[q];
    fpos[r] = q;
    dn[q] = ds[r];
    zn[q] = zs[r];
  }
  for (int q = 0; q < nd_; ++q) fpos[defl[q]] = kp + q;
  v.kp[blockIdx.x] = kp;
  v.nrot[blockIdx.x] = nr;
  v.rho[blockIdx.x] = rho;
}

// M2: gather the children's eigenvector columns into the node's K x K
// matrix Qg in final order (non-deflated first), applying the deflation
// rotations.  One thread per row of the node.
__global__ __launch_bounds__(256) void dc_gather_kernel(const DcNode* nodes, const double* Z,
                                                        int64_t ldz, DcVec v, double* Qg) {
  const DcNode nd = nodes[blockIdx.y];
  const int lo = nd.lo, K = nd.hi - nd.lo;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= K) return;
  double* Q = Qg + nd.qoff;
  const int* src = v.src + lo;
  const int* fpos = v.fpos + lo;
  for (int r = 0; r < K; ++r)
    Q[i + (int64_t)fpos[r] * K] = Z[(lo + i) + (int64_t)(lo + src[r]) * ldz];
  const int nr = v.nrot[blockIdx.y];
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

// M7: final sort of the node's eigenvalues (rank counting) and scatter of the
// eigenvectors into Z's diagonal block.  First kernel: ranks.
__global__ __launch_bounds__(1024) void dc_rank_kernel(const DcNode* nodes, DcVec v, double* lam,
                                                       int* colsrc) {
  const DcNode nd = nodes[blockIdx.x];
  const int lo = nd.lo, K = nd.hi - nd.lo;
  const int kp = v.kp[blockIdx.x];
  // value of final column q: q < kp -> lamv[q], else deflated ds
  for (int q = threadIdx.x; q < K; q += blockDim.x) {
    const double val = q < kp ? v.lamv[lo + q] : v.ds[lo + v.defl[lo + q - kp]];
    int r = 0;
    for (int t = 0; t < K; ++t) {
      const double vt = t < kp ? v.lamv[lo + t] : v.ds[lo + v.defl[lo + t - kp]];
      r += (vt < val) || (vt == val && t < q);
    }
    v.zs[lo + r] = val;  // zs reused as the sorted eigenvalue buffer
    colsrc[lo + r] = q;
  }
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
  std::vector<int> hostbuf(2 * nleaf);
  for (int i = 0; i < nleaf; ++i) { hostbuf[i] = P.leaf_lo[i]; hostbuf[nleaf + i] = P.leaf_hi[i]; }
  MSDP_HIP(hipMemcpyAsync(leafbuf, hostbuf.data(), sizeof(int) * 2 * nleaf, hipMemcpyHostToDevice, s));
  if (nsplit > 0) {
    MSDP_HIP(hipMemcpyAsync(splitbuf, P.splits.data(), sizeof(int) * nsplit, hipMemcpyHostToDevice, s));
    dc_split_kernel<<<cdiv(nsplit, 256), 256, 0, s>>>(nsplit, splitbuf, e, d);
  }
  dc_leaf_kernel<<<nleaf, 256, 0, s>>>(leafbuf, leafbuf + nleaf, d, e, lam, Z, n);
  MSDP_HIP(hipGetLastError());
  // the host copies below are stream-ordered from pageable memory; keep the
  // staging vectors alive until the stream is synchronised at each level.
  for (const auto& level : P.levels) {
    const int nn = (int)level.size();
    MSDP_HIP(hipMemcpyAsync(nodes_dev, level.data(), sizeof(DcNode) * nn, hipMemcpyHostToDevice, s));
    dc_deflate_kernel<<<nn, 1024, 0, s>>>(nodes_dev, lam, e, Z, n, v);
    int maxK = 0;
    for (const auto& nd : level) maxK = std::max(maxK, nd.hi - nd.lo);
    dc_gather_kernel<<<dim3(cdiv(maxK, 256), nn), 256, 0, s>>>(nodes_dev, Z, n, v, Qg);
    MSDP_HIP(hipGetLastError());
    std::vector<int> kp(nn);
    MSDP_HIP(hipMemcpyAsync(kp.data(), v.kp, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
    MSDP_HIP(hipStreamSynchronize(s));
    // root -> (node, local) maps
    std::vector<int> rmap;
    std::vector<int> rloc;
    for (int q = 0; q < nn; ++q)
      for (int j = 0; j < kp[q]; ++j) { rmap.push_back(q); rloc.push_back(j); }
    const int nroots = (int)rmap.size();
    if (nroots > 0) {
      // pad to a whole number of 4-wave blocks
      const int padded = cdiv(nroots, 4) * 4;
      rmap.resize(padded, -1);
      rloc.resize(padded, 0);
      std::vector<int> both(2 * padded);
      std::copy(rmap.begin(), rmap.end(), both.begin());
      std::copy(rloc.begin(), rloc.end(), both.begin() + padded);
      MSDP_HIP(hipMemcpyAsync(rootmap, both.data(), sizeof(int) * 2 * padded, hipMemcpyHostToDevice, s));
      dc_secular_kernel<<<padded / 4, 256, 0, s>>>(nodes_dev, nn, rootmap, rootmap + padded, v);
      dc_lowner_kernel<<<padded / 4, 256, 0, s>>>(nodes_dev, rootmap, rootmap + padded, v);
      // reuse the maps: one workgroup per column
      dc_vectors_kernel<<<nroots, 256, 0, s>>>(nodes_dev, rootmap, rootmap + padded, v, Ub);
      MSDP_HIP(hipGetLastError());
      for (int q = 0; q < nn; ++q) {
        const int K = level[q].hi - level[q].lo;
        if (kp[q] == 0) continue;
        MSDP_TRY(gemm(s, false, false, K, kp[q], kp[q], 1.0, Qg + level[q].qoff, K,
                      Ub + level[q].qoff, kp[q], 0.0, Qn + level[q].qoff, K));
      }
      MSDP_HIP(hipStreamSynchronize(s));  // keep `both` alive until consumed
    }
    dc_rank_kernel<<<nn, 1024, 0, s>>>(nodes_dev, v, lam, colsrc);
    dc_scatter_kernel<<<dim3(cdiv(maxK, 256), maxK, nn), 256, 0, s>>>(nodes_dev, v, Qg, Qn, colsrc,
                                                                       lam, Z, n);
    MSDP_HIP(hipGetLastError());
    MSDP_HIP(hipStreamSynchronize(s));  // `level` staging copy consumed
  }
  return 0;
}

// ==========================================================================
// back-transformation Z <- Q Z with Q = H_0 ... H_{n-3} (compact WY panels)
// ==========================================================================
// T (nb x nb, upper) of a panel of reflectors from their Gram matrix
// G = V'V: T_ii = tau_i, T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i].
__global__ __launch_bounds__(64) void larft_kernel(int nb, const double* G, const double* tau,
                                                   double* T) {
  __shared__ double Ts[64][65];
  const int t = threadIdx.x;
  for (int e = t; e < 64 * 64; e += 64) Ts[e & 63][e >> 6] = 0.0;
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    const double ti = tau[i];
    // column i: x = G[0:i, i]; T[0:i, i] = -ti * T[0:i,0:i] x
    if (t < i) {
      double v = 0.0;
      for (int k = t; k < i; ++k) v += Ts[t][k] * G[k + (int64_t)i * nb];
      Ts[t][i] = -ti * v;
    }
    if (t == 0) Ts[i][i] = ti;
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += 64) T[(e % nb) + (int64_t)(e / nb) * nb] = Ts[e % nb][e / nb];
}

static int ormtr(hipStream_t s, int n, const double* V, const double* tau, double* Z, int64_t ldz,
                 double* work) {
  const int nref = n - 2;
  if (nref <= 0) return 0;
  const int nb = 64;
  double* G = work;                       // nb x nb
  double* T = G + nb * nb;                // nb x nb
  double* W1 = T + nb * nb;               // nb x n
  double* W2 = W1 + (size_t)nb * n;       // nb x n
  const int npanel = cdiv(nref, nb);
  for (int pi = npanel - 1; pi >= 0; --pi) {
    const int k0 = pi * nb;
    const int kb = std::min(nb, nref - k0);
    const int r0 = k0 + 1;                // first row touched
    const int rows = n - r0;
    const double* Vp = V + r0 + (int64_t)k0 * n;  // rows x kb, ld n
    MSDP_TRY(gemm(s, true, false, kb, kb, rows, 1.0, Vp, n, Vp, n, 0.0, G, kb));
    larft_kernel<<<1, 64, 0, s>>>(kb, G, tau + k0, T);
    MSDP_HIP(hipGetLastError());
    MSDP_TRY(gemm(s, true, false, kb, n, rows, 1.0, Vp, n, Z + r0, ldz, 0.0, W1, kb));
    MSDP_TRY(gemm(s, false, false, kb, n, kb, 1.0, T, kb, W1, kb, 0.0, W2, kb));
    MSDP_TRY(gemm(s, false, false, rows, n, kb, -1.0, Vp, n, W2, kb, 1.0, Z + r0, ldz));
  }
  return 0;
}

// ==========================================================================
// public entry points
// ==========================================================================
size_t syevd_work_elems(int n) {
  return trd_work_elems(n) + dc_work_elems(n) + (size_t)n * n + 2 * 64 * 64 + 2 * 64 * (size_t)n + 64;
}
size_t syevd_iwork_elems(int n) { return dc_iwork_elems(n); }

int syevd(hipStream_t s, int n, const double* A, int64_t lda, double* w, double* V, int64_t ldv,
          bool vectors, double* work, int* iwork) {
  if (n <= 0) return 0;
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

int eig_min(hipStream_t s, int n, const double* A, int64_t lda, double* out, double* work, int*) {
  if (n <= 0) return 0;
  if (n <= 64) {
    double* wv = work;
    MSDP_TRY(jacobi_batched(s, n, A, lda, 0, wv, 0, nullptr, 0, 0, 1, false));
    MSDP_HIP(hipMemcpyAsync(out, wv, sizeof(double), hipMemcpyDeviceToDevice, s));
    return 0;
  }
  TrdWork t = trd_carve(n, work);
  MSDP_TRY(sytrd(s, n, A, lda, t));
  tridiag_min_kernel<<<1, 256, 0, s>>>(n, t.d, t.e, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
