// Shared declarations of the MadSDP gfx950 library: device helpers and the
// host-side launchers of every kernel.  All matrices are column-major FP64
// (Julia's `Matrix{Float64}` layout), leading dimension given explicitly.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <functional>
#include <vector>
#include <cstdint>
#include <cmath>
#include <string>

namespace msdp {

// roctx range over one stage of the interior-point iteration, shown by
// rocprofv3 --marker-trace (a no-op when no tool is attached)
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int WAVE = 64;

// --------------------------------------------------------------------------
// error plumbing: C-ABI functions return 0 on success, < 0 on API/HIP error,
// > 0 on a numerical failure (first failing pivot).
// --------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();
int fail_hip(hipError_t e, const char* what, const char* file, int line);
// Host wait for a stream.  By default the calling thread spins (lowest
// latency for one solve); a thread that switched to blocking mode sleeps on a
// blocking-sync event instead, leaving its core to other solves' threads.
int stream_sync(hipStream_t s);
void set_blocking_sync(bool on);
// The calling thread's products share the device with other solves: each
// product is one launch (no split-K / stream-K partition over the chip).
bool& shared_device();
#define MSDP_HIP(x)                                                      \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) return ::msdp::fail_hip(e_, #x, __FILE__, __LINE__); \
  } while (0)
#define MSDP_TRY(x)              \
  do {                           \
    int r_ = (x);                \
    if (r_ < 0) return r_;       \
  } while (0)

// --------------------------------------------------------------------------
// device helpers
// --------------------------------------------------------------------------
__device__ inline double readlane_f64(double v, int l);
// One DPP move of a double (two 32-bit halves); lanes outside row_mask, or
// whose source lane is out of range, receive 0.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
// Sum over the wavefront, returned in every lane: an inclusive scan within
// each row of 16 lanes (row_shr 1, 2, 4, 8), the row totals carried by
// row_bcast 15 and 31, then lane 63's total read back.  DPP moves take a few
// cycles where a shuffle through ds_bpermute takes an LDS round trip; the
// order of the additions is fixed, so the result is deterministic.
__device__ inline double wave_sum(double v) {
  v += dpp_f64<0x111, 0xf>(v);
  v += dpp_f64<0x112, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return readlane_f64(v, 63);
}
// The same DPP move with `id` (the reduction's identity) where the source
// lane is out of range or the row is masked off.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_f64_or(double v, double id) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned long long o = (unsigned long long)__double_as_longlong(id);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, CTRL, ROW_MASK,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32),
                                             CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
// Maximum / minimum over the wavefront (exact in any order), by the scan of
// wave_sum, returned in every lane.
__device__ inline double wave_max(double v) {
  const double id = -INFINITY;
  v = fmax(v, dpp_f64_or<0x111, 0xf>(v, id));
  v = fmax(v, dpp_f64_or<0x112, 0xf>(v, id));
  v = fmax(v, dpp_f64_or<0x114, 0xf>(v, id));
  v = fmax(v, dpp_f64_or<0x118, 0xf>(v, id));
  v = fmax(v, dpp_f64_or<0x142, 0xa>(v, id));
  v = fmax(v, dpp_f64_or<0x143, 0xc>(v, id));
  return readlane_f64(v, 63);
}
__device__ inline double wave_min(double v) {
  const double id = INFINITY;
  v = fmin(v, dpp_f64_or<0x111, 0xf>(v, id));
  v = fmin(v, dpp_f64_or<0x112, 0xf>(v, id));
  v = fmin(v, dpp_f64_or<0x114, 0xf>(v, id));
  v = fmin(v, dpp_f64_or<0x118, 0xf>(v, id));
  v = fmin(v, dpp_f64_or<0x142, 0xa>(v, id));
  v = fmin(v, dpp_f64_or<0x143, 0xc>(v, id));
  return readlane_f64(v, 63);
}
// 1 / sqrt(x) for finite x > 0: the hardware estimate (v_rsq_f64) refined by
// two Newton steps y <- y + y (1 - x y^2) / 2, each doubling the correct
// bits, to within a few ulp.  A short dependent chain where a correctly
// rounded square root followed by a quotient is a long one (the plane
// rotations of the QL chase).  The Cholesky column steps keep the correctly
// rounded operations: with these the dual-centering parity case converged
// one iteration earlier than the FP64 reference.
__device__ inline double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}
// Orders a wavefront's LDS stores before its later LDS loads (no workgroup
// barrier: the wavefront's LDS accesses complete in program order).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// The value of v in lane l (l uniform), through two v_readlane_b32.
__device__ inline double readlane_f64(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Lower-triangle tiles of a square nt x nt tile grid on a 1-D launch of
// lower_tile_blocks(nt) workgroups.  Consecutive workgroups go to different
// XCDs (blockIdx % 8), so each XCD is given a contiguous run of tiles in
// column-major lower order: the runs share their operand panels in that
// XCD's L2 instead of every XCD streaming all of them.  Returns false for
// the padding workgroups.
__host__ __device__ inline int lower_tile_blocks(int nt) {
  const int T = nt * (nt + 1) / 2;
  return 8 * ((T + 7) / 8);
}
__device__ inline bool lower_tile_xcd(int nt, int& tm, int& tn) {
  const int T = nt * (nt + 1) / 2;
  const int per = (T + 7) / 8;
  const int w = blockIdx.x;
  const int t = (w & 7) * per + (w >> 3);
  if (t >= T) return false;
  int c = 0, rem = t;
  while (rem >= nt - c) {
    rem -= nt - c;
    ++c;
  }
  tn = c;
  tm = c + rem;
  return true;
}
// Block-wide reduction for blockDim.x a multiple of 64 (<= 1024).
template <int OP>  // 0 sum, 1 max, 2 min
__device__ inline double block_reduce(double v, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = OP == 0 ? wave_sum(v) : (OP == 1 ? wave_max(v) : wave_min(v));
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int i = 1; i < nw; ++i) r = OP == 0 ? r + sh[i] : (OP == 1 ? fmax(r, sh[i]) : fmin(r, sh[i]));
  __syncthreads();
  return r;
}

// --------------------------------------------------------------------------
// dense FP64 kernels (dense.hip)
// --------------------------------------------------------------------------
// Triangular operands of a product (`tri`, bits): op(A) lower (its entries
// right of the diagonal are zero) / upper, op(B) lower / upper.  Every output
// tile then runs only the K range where both operands can be nonzero; the
// entries skipped must be stored zeros (the result is the full product's).
enum { TRI_AL = 1, TRI_AU = 2, TRI_BL = 4, TRI_BU = 8 };
// C = alpha op(A) op(B) + beta C for a batch of `batch` problems separated by
// the given strides; `lower` restricts the write (and the work) to C's lower
// triangle, i.e. a SYRK-shaped product.
int gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
         const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
         double* C, int64_t ldc, int batch = 1, int64_t sA = 0, int64_t sB = 0,
         int64_t sC = 0, bool lower = false, int tri = 0);
// One recorded op (the recorder must be active) of `batch` products at the
// given strides, whatever their sizes: one 64-tile item of the
// variable-size batched launch per position (the dense Schur assembly of
// many blocks).
int gemm_vb_record(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                   double* C, int64_t ldc, int batch, int64_t sA, int64_t sB, int64_t sC,
                   bool lower);
// One recorded op: B[:, j] = svec(lower triangle of P_j) for j < count
// (P_j n x n at j n^2; the order of svec_pack_kernel)
int svec_record(hipStream_t s, int n, int count, const double* P, double* B, int64_t ldb);
// the same as one launch (ops.hip)
int svec_pack(hipStream_t s, int n, int count, const double* P, double* B, int64_t ldb);
// the same product with its partials in the library's per-device scratch
// array (no allocation; calls that overlap on one device race)
int gemm_noalloc(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc);
size_t gemm_noalloc_scratch_bytes();
// bytes of stream-K partials one product may take from the pool
size_t gemm_sk_partials_bytes();
// kernel-selection parameters by name (dense.hip)
int set_tuning(const char* name, double value);
int get_tuning(const char* name, double* value);
// Bunch-Kaufman LDL' of the lower triangle (ldlt.hip): work 2n doubles, ctl
// sytrf_bk_ctl_bytes(); *info_dev = first zero pivot (1-based) or 0
int sytrf_bk(hipStream_t s, int n, double* A, int64_t lda, int* ipiv, void* ctl, double* work,
             int* info_dev);
size_t sytrf_bk_ctl_bytes();
int sytrs_bk(hipStream_t s, int n, const double* A, int64_t lda, const int* ipiv, double* b);
// Lower Cholesky in place (reads the lower triangle).  *info_dev receives 0 or
// the 1-based index of the first non-positive pivot; kernels after a failure
// return immediately.  work (potrf_work_elems) holds the diagonal blocks'
// factors on the way; with `inverses` it receives the inverses of the
// factor's diagonal 64-blocks (block b at work + 4096 b, 64 x 64, ld 64).
int potrf(hipStream_t s, int n, double* A, int64_t lda, int* info_dev, double* work,
          bool inverses = false);
// Two independent factorisations of the same side in the same launches:
// info_dev[0] and info_dev[1] receive the two statuses; work holds
// 2 potrf_work_elems(n).
int potrf2(hipStream_t s, int n, double* A0, double* A1, int64_t lda, int* info_dev,
           double* work);
size_t potrf_work_elems(int n);
// Batched Cholesky of small matrices (n <= 64), one workgroup each.
int potrf_small_batched(hipStream_t s, int n, double* A, int64_t lda, int64_t stride,
                        int batch, int* info_dev);
// Inverse of a lower-triangular matrix: Linv = L^-1 (lower, upper zeroed).
int trtri_lower(hipStream_t s, int n, const double* L, int64_t ldl, double* Linv,
                int64_t ldi, double* work);
size_t trtri_work_elems(int n);
// count matrices at strides sL / sLi; work holds count x trtri_work_elems(n)
// at stride sW
int trtri_lower_batched(hipStream_t s, int n, int count, const double* L, int64_t ldl,
                        int64_t sL, double* Li, int64_t ldi, int64_t sLi, double* work,
                        int64_t sW);
// Cholesky flush of the factor: zero the upper triangle, zero entries below
// eps^2 * max |L_ii| (src/blocks.jl:101-111).
int chol_flush(hipStream_t s, int n, double* L, int64_t ld, double* scratch);
// elementwise helpers
int symmetrize(hipStream_t s, int n, double* A, int64_t lda);
// two matrices of the same side (the second at A0 + sA) in one launch
int symmetrize2(hipStream_t s, int n, double* A0, int64_t sA, int64_t lda);
// L = chol((M + M')/2) with the flush and a zero upper triangle, n <= 64, one
// launch; *info = 0, or the 1-based first non-positive pivot
int cholesky_small(hipStream_t s, int n, const double* M, int64_t ldm, double* L, int64_t ldl,
                   int* info);
// the same for 64 < n <= 128 in one launch of one workgroup (recorded like
// cholesky_small); info = the 1-based failing column, 0 on success
int cholesky_mid(hipStream_t s, int n, const double* M, int64_t ldm, double* L, int64_t ldl,
                 int* info);
int mirror_lower(hipStream_t s, int n, double* A, int64_t lda);  // upper := lower'
int copy_mat(hipStream_t s, int rows, int cols, const double* A, int64_t lda, double* B,
             int64_t ldb);
int transpose(hipStream_t s, int rows, int cols, const double* A, int64_t lda, double* B,
              int64_t ldb);
// Z = a*X + b*Y (+ c*V if V), elementwise over an n x n matrix; Z may alias.
int axpby(hipStream_t s, int64_t count, double a, const double* X, double b, const double* Y,
          double* Z);
int axpbypcz(hipStream_t s, int64_t count, double a, const double* X, double b,
             const double* Y, double c, const double* V, double* Z);
int set_identity(hipStream_t s, int n, double* A, int64_t lda, double scale);
int fill(hipStream_t s, int64_t count, double v, double* X);
int add_diag(hipStream_t s, int n, double* A, int64_t lda, const double* dev_val, double val);
int scale_cols(hipStream_t s, int rows, int cols, double* A, int64_t lda, const double* d,
               int power_half_sign);  // A[:,k] *= d[k]^(p/2), p in {-1,+1}
int scale_rows(hipStream_t s, int rows, int cols, double* A, int64_t lda, const double* d,
               int power_half_sign);
// Reductions into a device scalar (deterministic two-pass).
int dot(hipStream_t s, int64_t count, const double* X, const double* Y, double* out_dev,
        double* scratch);
int max_abs(hipStream_t s, int64_t count, const double* X, double* out_dev, double* scratch);
size_t reduce_scratch_elems(int64_t count);

// --------------------------------------------------------------------------
// symmetric eigensolvers (eigen.hip)
// --------------------------------------------------------------------------
// Eigenvalues (ascending) and optionally eigenvectors (columns of V) of the
// symmetric matrix A (full storage, read only).
int syevd(hipStream_t s, int n, const double* A, int64_t lda, double* w, double* V,
          int64_t ldv, bool vectors, double* work, int* iwork);
size_t syevd_work_elems(int n);
size_t syevd_iwork_elems(int n);
// Smallest eigenvalue of the symmetric matrix A into *out_dev.
int eig_min(hipStream_t s, int n, const double* A, int64_t lda, double* out_dev,
            double* work, int* iwork);
// The smallest eigenvalues of two symmetric matrices of the same side in the
// same launches; work holds 2 eig_min_work_elems(n) (syevd_work_elems(n)).
int eig_min2(hipStream_t s, int n, const double* A0, const double* A1, int64_t lda,
             double* out0, double* out1, double* work);
size_t eig_min_work_elems(int n);
// One eigenproblem of a variable-size batch (n <= 128): eigenvalues into w
// (ascending), optionally w[0] also into *wmin, vectors into V (ld ldv) with
// Vtmp (n x n) as scratch.
struct EigBatch {
  int n;
  const double* A;
  int64_t lda;
  double* w;
  double* wmin;
  double* V;
  int64_t ldv;
  double* Vtmp;
};
constexpr int JACOBI_VB_MAX = 128;
// --------------------------------------------------------------------------
// Per-block op recorder.  Inside a recorded region the device ops below do
// not launch: each appends itself (a batchable descriptor and a closure that
// launches it alone) to the current block's list.  The region's flush then
// walks the lists position by position: the same kind of op of many blocks
// at one position becomes one variable-size batched launch (64 x 64 MFMA
// tiles, elementwise, reductions, small Cholesky), everything else replays
// per block in order.  A block's ops keep their order and each batched
// launch does exactly the arithmetic of the separate launches, so results
// are bit-identical with or without recording.
// --------------------------------------------------------------------------
enum RecKind {
  REC_OPAQUE = 0, REC_GEMM_NN, REC_GEMM_TN, REC_GEMM_NT, REC_GEMM_TT, REC_EW, REC_RED,
  REC_CHOL64, REC_TRTRI64, REC_CHOL128, REC_TRTRI128, REC_NKINDS
};
struct GemmItem {
  int M, N, K, lower, tri;
  double alpha, beta;
  const double* A;
  const double* B;
  double* C;
  int64_t lda, ldb, ldc;
  int batch = 1;  // `batch` products at the strides below (gemm_vb_record)
  int64_t sA = 0, sB = 0, sC = 0;
};
enum EwOp {
  EW_AXPBY = 0, EW_AXPBYPCZ, EW_FILL, EW_COPY, EW_SYM, EW_MIRROR, EW_TRANSPOSE, EW_SCALE_COLS,
  EW_SCALE_ROWS, EW_RNT, EW_SCALE_BY_DIAG, EW_ZERO_UPPER,
  EW_SVEC  // Z[:, j] = svec of the lower triangle of X_j (n = p, X_j at j n^2, ld ldb)
};
struct EwItem {
  int op, rows, cols, p;
  int64_t count, lda, ldb;
  double a, b, c;
  const double* X;
  const double* Y;
  const double* V;
  double* Z;
};
enum RedOp { RED_DOT = 0, RED_MAXABS, RED_SQRTMIN };
struct RedItem {
  int op, pad;
  int64_t count;
  const double* X;
  const double* Y;
  double* out;
  double* D;
};
struct CholItem {  // _cholesky_factor! of an n <= 64 block (cholesky_small)
  int n, pad;
  const double* M;
  double* L;
  int* info;
  int64_t ldm, ldl;
};
struct TrtriItem {  // L^-1 of an n <= 64 lower factor (trtri_lower)
  int n, pad;
  const double* L;
  double* Li;
  int64_t ldl, ldi;
};
struct RecOp {
  int kind = REC_OPAQUE;
  GemmItem g{};
  EwItem e{};
  RedItem r{};
  CholItem ch{};
  TrtriItem tr{};
  std::function<int()> run;
};
struct Recorder {
  std::vector<std::vector<RecOp>> ops;  // per block
  int cur = 0;
};
// the recorder of the calling thread's current region (null: launch at once)
Recorder*& rec_active();
// ops without a recorded form refuse to run inside a recorded region (the
// caller wraps them as opaque closures instead)
inline int rec_forbid(const char* what) {
  if (!rec_active()) return 0;
  set_error(std::string(what) + " called inside a recorded region");
  return -3;
}
inline int rec_add(RecOp&& op) {
  rec_active()->ops[rec_active()->cur].push_back(std::move(op));
  return 0;
}
// A batched launch's arguments: rec_pack writes the items of one kind (and
// the prefix of their workgroup counts) to host memory and returns the byte
// count and the grid size; rec_launch runs them from the same bytes copied
// to device memory.
size_t rec_bytes(int kind, size_t count);
size_t rec_pack(int kind, const std::vector<const RecOp*>& items, char* host, int* grid);
int rec_launch(hipStream_t s, int kind, int count, int grid, const char* dev);

int jacobi_vbatched(hipStream_t s, const EigBatch* items_dev, int count, bool vectors);
// eigenvalues ascending in w and the matching vectors in V (Vtmp: n x n
// scratch), n <= 128, one workgroup per item
int syev_vbatched(hipStream_t s, const EigBatch* items_dev, int count);
// eigenvalues only (w ascending if non-null, else just *wmin), n <= 128
int eigvals_vbatched(hipStream_t s, const EigBatch* items_dev, int count);
// Batched Jacobi for many small matrices (n <= 64).
int jacobi_batched(hipStream_t s, int n, const double* A, int64_t lda, int64_t strideA,
                   double* w, int64_t stridew, double* V, int64_t ldv, int64_t strideV,
                   int batch, bool vectors);

// --------------------------------------------------------------------------
// FP32 MFMA GEMM and conversions (gemm32.hip), for the mixed-precision
// Schur assembly
// --------------------------------------------------------------------------
int gemm32(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha, const float* A,
           int64_t lda, const float* B, int64_t ldb, double beta, float* C, int64_t ldc,
           int batch = 1, int64_t sA = 0, int64_t sB = 0, int64_t sC = 0, bool lower = false);
int gemm32_to64(hipStream_t s, bool ta, bool tb, int M, int N, int K, float alpha,
                const float* A, int64_t lda, const float* B, int64_t ldb, double beta,
                double* C, int64_t ldc, bool lower = false);
int d2f(hipStream_t s, int64_t count, const double* x, float* y);
int svec_pack32(hipStream_t s, int n, int cb, const float* P, float* B, int64_t ldb);

// --------------------------------------------------------------------------
// BF16x3 split products (bf16x3.hip): the mixed-precision Schur assembly
// H = B'B from hi = bf16(B), lo = bf16(B - hi) on the bf16 MFMA, and the
// FP64 panel operator B'(B v) its refinement applies
// --------------------------------------------------------------------------
// leading dimension of a split panel with `rows` rows (a multiple of 32)
int64_t bf16x3_ld(int64_t rows);
int bf16_split(hipStream_t s, int rows, int cols, const double* B, int64_t ldb, __bf16* hi,
               __bf16* lo, int64_t ldh);
// C = alpha (hi'hi + hi'lo + lo'hi) + beta C, lower triangle of M x M;
// hi, lo: K x M, K and lda multiples of 32
int syrk_bf16x3(hipStream_t s, int M, int K, double alpha, const __bf16* hi, const __bf16* lo,
                int64_t lda, double beta, double* C, int64_t ldc);
size_t panel_apply_work_elems(int rows, int cols);
// out (+)= B'(B v), B rows x cols FP64, deterministic
int panel_apply(hipStream_t s, int rows, int cols, const double* B, int64_t ldb, const double* v,
                double* out, bool accumulate, double* work);
int gather_vec(hipStream_t s, int n, const int* con, const double* v, double* vb);
int scatter_add_vec(hipStream_t s, int n, const int* con, const double* u, double* out);

}  // namespace msdp

namespace msdp {
// --------------------------------------------------------------------------
// Live kernel timing: when enabled for a kernel class on a stream, every
// launch of that class is timed with HIP events on that stream (recorded
// around the launch, or stamped on the dispatch by hipExtLaunchKernel) and
// its algorithmic work (flops or bytes) is accumulated.
// --------------------------------------------------------------------------
enum ProfClass { PROF_GEMM = 0, PROF_TRD = 1, PROF_GEMM32 = 2, PROF_TRD1 = 3, PROF_NONE = -1 };
bool prof_active(int cls, hipStream_t s);
void prof_before(hipStream_t s);
void prof_after(hipStream_t s, double work);
// For a launch through hipExtLaunchKernel: a start/stop event pair that the
// runtime stamps on the dispatch itself (no marker packets around it), with
// the launch's algorithmic work.
void prof_dispatch_events(hipEvent_t* start, hipEvent_t* stop, double work);
// s = nullptr: every stream of the process
int prof_begin(int cls, hipStream_t s);
int prof_end(double* total_ms, int64_t* launches, double* work);
}  // namespace msdp
