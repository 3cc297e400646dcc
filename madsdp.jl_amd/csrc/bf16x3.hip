// BF16x3 split-product Schur assembly and the FP64 panel operator of its
// refinement.
//
// The scaled constraint panel B (drow x m_i, FP64: column j = svec or vec of
// P_j = op(Lm) A_j op(F2)) is split once per assembly into hi = bf16(B) and
// lo = bf16(B - hi), stored K-contiguous with the columns padded by zeros to a
// multiple of 32.  H = B'B is then formed as hi'hi + hi'lo + lo'hi on the
// gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation), widened to
// FP64 into the lower triangle.  hi + lo carries about 17 significant bits of
// each entry, so the product is accurate to a few units of 2^-17 relative
// before the FP32 accumulation; the FP64 accuracy of the Newton system is won
// back by iterative refinement against the FP64 operator B'B, applied by two
// panel products (t = B v, then B't) that read the FP64 panel once each.
//
// bf16 16x16x32 operand layout (wave64, checked on the device): lane l holds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7; the
// result D[4 (l / 16) + r][l % 16], r = 0..3.
#include "common.h"

#include <algorithm>

namespace msdp {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

int64_t bf16x3_ld(int64_t rows) { return (rows + 31) / 32 * 32; }

// hi = bf16(x) and lo = bf16(x - hi) (round to nearest even through FP32),
// column j of the rows x cols panel; rows rows..ldh-1 of each column zero.
__global__ __launch_bounds__(256) void bf16_split_kernel(int rows, const double* __restrict__ B,
                                                         int64_t ldb, __bf16* __restrict__ hi,
                                                         __bf16* __restrict__ lo, int64_t ldh) {
  const int64_t j = blockIdx.y;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ldh;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double v = k < rows ? B[k + j * ldb] : 0.0;
    const __bf16 h = (__bf16)(float)v;
    hi[k + j * ldh] = h;
    lo[k + j * ldh] = (__bf16)(float)(v - (double)(float)h);
  }
}

int bf16_split(hipStream_t s, int rows, int cols, const double* B, int64_t ldb, __bf16* hi,
               __bf16* lo, int64_t ldh) {
  if (rows <= 0 || cols <= 0) return 0;
  if (ldh < rows || ldh % 32 != 0) {
    set_error("bf16_split: ldh must be a multiple of 32 and at least rows");
    return -2;
  }
  if (cols > 65535) {
    set_error("bf16_split: at most 65535 columns");
    return -2;
  }
  dim3 g(std::min(cdiv(ldh, 256), 64), cols);
  bf16_split_kernel<<<g, 256, 0, s>>>(rows, B, ldb, hi, lo, ldh);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// C = alpha (hi'hi + hi'lo + lo'hi) + beta C on the lower triangle of the
// M x M result; hi and lo are K x M (K a multiple of 32), column-major with
// leading dimension lda (a multiple of 32).  One 128 x 128 lower tile per
// workgroup (XCD-grouped 1-D grid), four waves of 64 x 64; K-slabs of 32 in
// double-buffered LDS rows of 40 bf16 (16-byte reads of 16 consecutive rows
// fall in distinct bank groups), the next slab's global loads in registers
// while the current one is multiplied.
constexpr int SK_T = 128, SK_KS = 32, SK_LD = 40;

// Lower tiles of an nt x nt tile grid for a 1-D launch of lower_tile_blocks(nt)
// workgroups, grouped so that the tiles one XCD runs together share their
// panels: the tiles are enumerated by 8 x 8 super-tiles (column-major lower
// order of super-tiles, column-major inside each, the diagonal super-tiles'
// lower half only) and each XCD (blockIdx % 8) takes a contiguous range.  The
// 64 workgroups an XCD holds at once then cover about one super-tile, which
// reads 16 row / column panels instead of the 65 of a column run.
constexpr int SK_SUPER = 8;
__device__ inline bool lower_tile_super(int nt, int& tm, int& tn) {
  const int T = nt * (nt + 1) / 2;
  const int per = (T + 7) / 8;
  const int w = blockIdx.x;
  int t = (w & 7) * per + (w >> 3);
  if (t >= T) return false;
  const int ns = (nt + SK_SUPER - 1) / SK_SUPER;
  for (int TN = 0; TN < ns; ++TN) {
    const int b1 = min(SK_SUPER, nt - TN * SK_SUPER);  // tile columns in this super-column
    for (int TM = TN; TM < ns; ++TM) {
      const int a1 = min(SK_SUPER, nt - TM * SK_SUPER);
      const int cnt = TM == TN ? a1 * (a1 + 1) / 2 : a1 * b1;
      if (t >= cnt) {
        t -= cnt;
        continue;
      }
      if (TM != TN) {
        tn = TN * SK_SUPER + t / a1;
        tm = TM * SK_SUPER + t % a1;
      } else {  // lower half of a diagonal super-tile, column by column
        int b = 0;
        while (t >= a1 - b) {
          t -= a1 - b;
          ++b;
        }
        tn = TN * SK_SUPER + b;
        tm = TN * SK_SUPER + b + t;
      }
      return true;
    }
  }
  return false;
}

__global__ __launch_bounds__(256, 2) void syrk_bf16x3_kernel(int M, int K, double alpha,
                                                             const __bf16* __restrict__ Ah,
                                                             const __bf16* __restrict__ Al,
                                                             int64_t lda, double beta,
                                                             double* __restrict__ C, int64_t ldc) {
  int tm, tn;
  if (!lower_tile_super((M + SK_T - 1) / SK_T, tm, tn)) return;
  // [buffer][hi, lo][column][k]: the tile's row panel (I) and column panel (J)
  __shared__ __attribute__((aligned(16))) __bf16 sI[2][2][SK_T * SK_LD];
  __shared__ __attribute__((aligned(16))) __bf16 sJ[2][2][SK_T * SK_LD];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;
  const int m0 = tm * SK_T, n0 = tn * SK_T;
  const bool diag = tm == tn;
  // a diagonal tile's upper-right quadrant holds no lower entries
  const bool idle = diag && wi < wj;
  // the next slab in registers: two 16-byte chunks (8 consecutive k of one
  // column) per thread and operand, hi and lo, for the row and column panels
  uint4 i0h, i0l, i1h, i1l, j0h, j0l, j1h, j1l;
  const int c0 = tid >> 2, c1 = (tid + 256) >> 2, kq = 8 * (tid & 3);
  // rows i >= M of the last row tile read column M - 1 instead: they only
  // reach results C[i][.] with i >= M, which are not written
  const __bf16* pi0h = Ah + kq + (int64_t)min(m0 + c0, M - 1) * lda;
  const __bf16* pi1h = Ah + kq + (int64_t)min(m0 + c1, M - 1) * lda;
  const __bf16* pi0l = Al + kq + (int64_t)min(m0 + c0, M - 1) * lda;
  const __bf16* pi1l = Al + kq + (int64_t)min(m0 + c1, M - 1) * lda;
  // column tiles lie at or left of the row tile: n0 + 127 < M
  const __bf16* pj0h = Ah + kq + (int64_t)(n0 + c0) * lda;
  const __bf16* pj1h = Ah + kq + (int64_t)(n0 + c1) * lda;
  const __bf16* pj0l = Al + kq + (int64_t)(n0 + c0) * lda;
  const __bf16* pj1l = Al + kq + (int64_t)(n0 + c1) * lda;
  const int o0 = c0 * SK_LD + kq, o1 = c1 * SK_LD + kq;
#define SK_GLOAD(k0)                                                     \
  do {                                                                   \
    i0h = *reinterpret_cast<const uint4*>(pi0h + (k0));                  \
    i0l = *reinterpret_cast<const uint4*>(pi0l + (k0));                  \
    i1h = *reinterpret_cast<const uint4*>(pi1h + (k0));                  \
    i1l = *reinterpret_cast<const uint4*>(pi1l + (k0));                  \
    if (!diag) {                                                         \
      j0h = *reinterpret_cast<const uint4*>(pj0h + (k0));                \
      j0l = *reinterpret_cast<const uint4*>(pj0l + (k0));                \
      j1h = *reinterpret_cast<const uint4*>(pj1h + (k0));                \
      j1l = *reinterpret_cast<const uint4*>(pj1l + (k0));                \
    }                                                                    \
  } while (0)
#define SK_SSTORE(buf)                                                   \
  do {                                                                   \
    *reinterpret_cast<uint4*>(&sI[buf][0][o0]) = i0h;                    \
    *reinterpret_cast<uint4*>(&sI[buf][1][o0]) = i0l;                    \
    *reinterpret_cast<uint4*>(&sI[buf][0][o1]) = i1h;                    \
    *reinterpret_cast<uint4*>(&sI[buf][1][o1]) = i1l;                    \
    if (!diag) {                                                         \
      *reinterpret_cast<uint4*>(&sJ[buf][0][o0]) = j0h;                  \
      *reinterpret_cast<uint4*>(&sJ[buf][1][o0]) = j0l;                  \
      *reinterpret_cast<uint4*>(&sJ[buf][0][o1]) = j1h;                  \
      *reinterpret_cast<uint4*>(&sJ[buf][1][o1]) = j1l;                  \
    }                                                                    \
  } while (0)
  v4f acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
  const int nk = K / SK_KS;
  SK_GLOAD(0);
  SK_SSTORE(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) SK_GLOAD((int64_t)(kt + 1) * SK_KS);
    if (!idle) {
      const __bf16* Ih = sI[buf][0];
      const __bf16* Il = sI[buf][1];
      const __bf16* Jh = diag ? Ih : sJ[buf][0];
      const __bf16* Jl = diag ? Il : sJ[buf][1];
      bf8 ih[4], il[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int o = (wi + 16 * a + li) * SK_LD + 8 * lk;
        ih[a] = *reinterpret_cast<const bf8*>(Ih + o);
        il[a] = *reinterpret_cast<const bf8*>(Il + o);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int o = (wj + 16 * b + li) * SK_LD + 8 * lk;
        const bf8 jh = *reinterpret_cast<const bf8*>(Jh + o);
        const bf8 jl = *reinterpret_cast<const bf8*>(Jl + o);
        // D[j][i] = sum_k J[k][j] I[k][i]: the small terms first, four
        // independent accumulators between dependent products
#pragma unroll
        for (int a = 0; a < 4; ++a)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(jl, ih[a], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(jh, il[a], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(jh, ih[a], acc[a][b], 0, 0, 0);
      }
    }
    if (kt + 1 < nk) SK_SSTORE(buf ^ 1);
    __syncthreads();
  }
#undef SK_GLOAD
#undef SK_SSTORE
  if (idle) return;
  // lane (li, lk) holds C[i0 + li][j0 + 4 lk + r]: 16 consecutive rows of
  // one column per group of 16 lanes (coalesced)
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = m0 + wi + 16 * a + li;
    if (i >= M) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = n0 + wj + 16 * b + 4 * lk + r;
        if (j > i) continue;
        double* c = C + i + (int64_t)j * ldc;
        const double v = alpha * (double)acc[a][b][r];
        *c = beta == 0.0 ? v : v + beta * *c;
      }
    }
  }
}

int syrk_bf16x3(hipStream_t s, int M, int K, double alpha, const __bf16* hi, const __bf16* lo,
                int64_t lda, double beta, double* C, int64_t ldc) {
  if (M <= 0) return 0;
  if (K % SK_KS != 0 || lda % SK_KS != 0 || lda < K) {
    set_error("syrk_bf16x3: K and lda must be multiples of 32, lda >= K");
    return -2;
  }
  if (K == 0) {
    set_error("syrk_bf16x3: K must be positive");
    return -2;
  }
  const bool prof = prof_active(PROF_GEMM32, s);
  if (prof) prof_before(s);
  const int nt = cdiv(M, SK_T);
  syrk_bf16x3_kernel<<<lower_tile_blocks(nt), 256, 0, s>>>(M, K, alpha, hi, lo, lda, beta, C, ldc);
  MSDP_HIP(hipGetLastError());
  // three products of the lower triangle
  if (prof) prof_after(s, 3.0 * (double)M * (M + 1) * (double)K);
  return 0;
}

// ---------------------------------------------------------------- panel
// t = B v for the FP64 panel B (rows x cols): column ranges of PANEL_CHUNK
// summed per row into partials, the partials added in range order.
constexpr int PANEL_CHUNK = 512;

__global__ __launch_bounds__(256) void panel_gemv_n_kernel(int rows, int cols,
                                                           const double* __restrict__ B,
                                                           int64_t ldb, const double* __restrict__ v,
                                                           double* __restrict__ part) {
  const int r = blockIdx.x * 256 + threadIdx.x, s = blockIdx.y;
  if (r >= rows) return;
  const int j0 = s * PANEL_CHUNK, j1 = min(cols, j0 + PANEL_CHUNK);
  double acc = 0.0;
#pragma unroll 8
  for (int j = j0; j < j1; ++j) acc = fma(B[r + (int64_t)j * ldb], v[j], acc);
  part[r + (int64_t)s * rows] = acc;
}

__global__ __launch_bounds__(256) void panel_sum_kernel(int rows, int S, const double* __restrict__ part,
                                                        double* __restrict__ t) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  double acc = part[r];
  for (int s = 1; s < S; ++s) acc += part[r + (int64_t)s * rows];
  t[r] = acc;
}

// out[j] (+)= B[:, j]' t, one workgroup per column
__global__ __launch_bounds__(256) void panel_gemv_t_kernel(int rows, const double* __restrict__ B,
                                                           int64_t ldb, const double* __restrict__ t,
                                                           double* __restrict__ out, int accumulate) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const double* col = B + j * ldb;
  double acc = 0.0;
  for (int r = threadIdx.x; r < rows; r += 256) acc = fma(col[r], t[r], acc);
  acc = block_reduce<0>(acc, red);
  if (threadIdx.x == 0) out[j] = accumulate ? out[j] + acc : acc;
}

size_t panel_apply_work_elems(int rows, int cols) {
  return (size_t)rows * (size_t)std::max(1, cdiv(cols, PANEL_CHUNK)) + (size_t)rows;
}

// out (+)= B'(B v): the FP64 operator of the panel, deterministic (fixed
// summation order); work holds panel_apply_work_elems(rows, cols) doubles.
int panel_apply(hipStream_t s, int rows, int cols, const double* B, int64_t ldb, const double* v,
                double* out, bool accumulate, double* work) {
  if (rows <= 0 || cols <= 0) return 0;
  if (cols > 65535 * PANEL_CHUNK) {
    set_error("panel_apply: too many columns");
    return -2;
  }
  const int S = cdiv(cols, PANEL_CHUNK);
  double* part = work;
  double* t = work + (size_t)rows * S;
  panel_gemv_n_kernel<<<dim3(cdiv(rows, 256), S), 256, 0, s>>>(rows, cols, B, ldb, v, part);
  MSDP_HIP(hipGetLastError());
  panel_sum_kernel<<<cdiv(rows, 256), 256, 0, s>>>(rows, S, part, t);
  MSDP_HIP(hipGetLastError());
  panel_gemv_t_kernel<<<cols, 256, 0, s>>>(rows, B, ldb, t, out, accumulate ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// vb[j] = v[con[j]] and out[con[j]] += u[j] (con distinct) for blocks that
// touch a subset of the constraints
__global__ void gather_kernel(int n, const int* __restrict__ con, const double* __restrict__ v,
                              double* __restrict__ vb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) vb[j] = v[con[j]];
}
__global__ void scatter_add_kernel(int n, const int* __restrict__ con, const double* __restrict__ u,
                                   double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[con[j]] += u[j];
}

int gather_vec(hipStream_t s, int n, const int* con, const double* v, double* vb) {
  if (n <= 0) return 0;
  gather_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, con, v, vb);
  MSDP_HIP(hipGetLastError());
  return 0;
}
int scatter_add_vec(hipStream_t s, int n, const int* con, const double* u, double* out) {
  if (n <= 0) return 0;
  scatter_add_kernel<<<cdiv(n, 256), 256, 0, s>>>(n, con, u, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
