// Device kernels of the model callbacks the reference takes from LowRankOpt
// and of the scalar (LP) block:
//   A(X)_j    = sum_i <A_ij, X_i>               LRO.add_jprod!     (src/step.jl:19,75)
//   A*(y)_i   = sum_j y_j A_ij                  LRO.unsafe_jtprod  (src/step.jl:30,97)
//   H_jk     += <A_ij, L A_ik R>                LRO.schur_complement! (src/solver.jl:304)
// Constraint storage per block: COO entries sorted by constraint (with a
// by-position index for A*), dense matrices of the touching constraints, or
// low-rank factors A_j = sum_r lam_r f_r f_r' (LowRankOpt's rank-one
// structure, reference README.md:148): A(X)_j = sum_r lam_r f_r' X f_r,
// A*(y) = F diag(lam y) F', H_jk = sum_{r,s} lam_r lam_s (f_r' L f_s)(f_s' R f_r).
#include "ops.h"

#include <algorithm>

namespace msdp {

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------- A(X) ----
__global__ void jprod_coo_kernel(int ncon, const int* con_list, const int* con_ptr,
                                 const int* row, const int* col, const double* val,
                                 const double* X, int64_t ldx, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncon) return;
  double s = 0.0;
  for (int e = con_ptr[t]; e < con_ptr[t + 1]; ++e) s += val[e] * X[row[e] + (int64_t)col[e] * ldx];
  out[con_list[t]] += s;
}

__global__ __launch_bounds__(256) void jprod_dense_kernel(int64_t n2, const int* dcon,
                                                          const double* mats, const double* X,
                                                          double* out) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const double* Aj = mats + j * n2;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n2; e += 256) s += Aj[e] * X[e];
  s = block_reduce<0>(s, red);
  if (threadIdx.x == 0) out[dcon[j]] += s;
}

// out[con_list[j]] += sum_{r in j} lam_r f_r' Y[:, r]   (Y = X F)
__global__ __launch_bounds__(256) void jprod_lowrank_kernel(int n, const int* con_list,
                                                            const int* con_ptr, const double* F,
                                                            const double* Y, const double* lam,
                                                            double* out) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int r = con_ptr[j]; r < con_ptr[j + 1]; ++r) {
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) t += F[i + (int64_t)r * n] * Y[i + (int64_t)r * n];
    s += lam[r] * block_reduce<0>(t, red);
  }
  if (threadIdx.x == 0) out[con_list[j]] += s;
}

int jprod(hipStream_t st, const BlockOps& b, const double* X, double* out) {
  MSDP_TRY(rec_forbid("jprod"));
  if (b.kind == BlockOps::EMPTY) return 0;
  if (b.kind == BlockOps::LOWRANK) {
    MSDP_TRY(gemm(st, false, false, b.n, b.R, b.n, 1.0, X, b.n, b.F, b.n, 0.0, b.FW, b.n));
    jprod_lowrank_kernel<<<b.ncon, 256, 0, st>>>(b.n, b.con_list, b.con_ptr, b.F, b.FW, b.lam, out);
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  if (b.kind == BlockOps::COO) {
    jprod_coo_kernel<<<cdiv(b.ncon, 256), 256, 0, st>>>(b.ncon, b.con_list, b.con_ptr, b.row,
                                                         b.col, b.val, X, b.n, out);
  } else {
    jprod_dense_kernel<<<b.ncon, 256, 0, st>>>((int64_t)b.n * b.n, b.con_list, b.mats, X, out);
  }
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- A*(y) ---
__global__ void jtprod_coo_kernel(int npos, const int64_t* pos_lin, const int* pos_ptr,
                                  const int* pos_ent, const int* con, const double* val,
                                  const double* y, double* out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npos) return;
  double s = 0.0;
  for (int k = pos_ptr[p]; k < pos_ptr[p + 1]; ++k) {
    const int e = pos_ent[k];
    s += y[con[e]] * val[e];
  }
  out[pos_lin[p]] = s;
}

// out (or part[split]) = sum over the split's constraint range; the ranges
// are added in a fixed order afterwards, so the result is deterministic.
__global__ void jtprod_dense_kernel(int64_t n2, int mi, int per, const int* dcon,
                                    const double* mats, const double* y, double* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2) return;
  const int j0 = blockIdx.y * per, j1 = min(mi, j0 + per);
  double s0 = 0.0, s1 = 0.0;
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    s0 += y[dcon[j]] * mats[(int64_t)j * n2 + e];
    s1 += y[dcon[j + 1]] * mats[(int64_t)(j + 1) * n2 + e];
  }
  if (j < j1) s0 += y[dcon[j]] * mats[(int64_t)j * n2 + e];
  out[(int64_t)blockIdx.y * n2 + e] = s0 + s1;
}

__global__ void jtprod_reduce_kernel(int64_t n2, int split, const double* part, double* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2) return;
  double s = 0.0;
  for (int k = 0; k < split; ++k) s += part[(int64_t)k * n2 + e];
  out[e] = s;
}

// the dense block holding key: the last one whose first pair (tile) is <= key
__device__ inline int dense_blk_of(const DenseBlkDev* b, int nblk, int key, bool tiles) {
  int lo = 0, hi = nblk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((tiles ? b[mid].tile0 : b[mid].pair0) <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// part[p] = <A_j, X_b> for pair p = (b, j): jprod_dense_kernel's sum
__global__ __launch_bounds__(256) void jprod_pairs_kernel(const DenseBlkDev* blks, int nblk,
                                                          int which, double* part) {
  __shared__ double red[4];
  const int p = blockIdx.x;
  const DenseBlkDev& b = blks[dense_blk_of(blks, nblk, p, false)];
  const int64_t n2 = b.n2;
  const double* Aj = b.mats + (int64_t)(p - b.pair0) * n2;
  const double* X = b.in[which];
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n2; e += 256) s += Aj[e] * X[e];
  s = block_reduce<0>(s, red);
  if (threadIdx.x == 0) part[p] = s;
}

__global__ void jprod_gather_kernel(int m, const int* cptr, const int* cpairs,
                                    const double* part, double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const int k0 = cptr[c], k1 = cptr[c + 1];
  if (k0 == k1) return;
  double o = out[c];
  for (int k = k0; k < k1; ++k) o += part[cpairs[k]];
  out[c] = o;
}

// out_b[e] = sum_j y[con_b[j]] A_j[e] for every block b (jtprod_dense_kernel
// over the block's whole constraint range)
__global__ __launch_bounds__(256) void jtprod_tiles_kernel(const DenseBlkDev* blks, int nblk,
                                                           const double* y, int which) {
  const DenseBlkDev& b = blks[dense_blk_of(blks, nblk, blockIdx.x, true)];
  const int64_t n2 = b.n2;
  const int64_t e = (int64_t)(blockIdx.x - b.tile0) * 256 + threadIdx.x;
  if (e >= n2) return;
  const int mi = b.ncon;
  const int* dcon = b.con_list;
  const double* mats = b.mats;
  double s0 = 0.0, s1 = 0.0;
  int j = 0;
  for (; j + 1 < mi; j += 2) {
    s0 += y[dcon[j]] * mats[(int64_t)j * n2 + e];
    s1 += y[dcon[j + 1]] * mats[(int64_t)(j + 1) * n2 + e];
  }
  if (j < mi) s0 += y[dcon[j]] * mats[(int64_t)j * n2 + e];
  b.out[which][e] = s0 + s1;
}

int jprod_dense_batched(hipStream_t st, const DenseBatch& d, int m, int which, double* out) {
  if (d.npairs == 0) return 0;
  jprod_pairs_kernel<<<d.npairs, 256, 0, st>>>(d.blks, d.nblk, which, d.part);
  jprod_gather_kernel<<<cdiv(m, 256), 256, 0, st>>>(m, d.cptr, d.cpairs, d.part, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int jtprod_dense_batched(hipStream_t st, const DenseBatch& d, const double* y, int which) {
  if (d.ntiles == 0) return 0;
  jtprod_tiles_kernel<<<d.ntiles, 256, 0, st>>>(d.blks, d.nblk, y, which);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int jtprod_split(int n, int ncon) {
  const int64_t n2 = (int64_t)n * n;
  const int tiles = cdiv(n2, 256);
  int split = cdiv(4096, tiles);
  split = std::min(split, std::max(1, ncon / 32));
  return std::max(1, std::min(split, 64));
}

// FW[:, r] = lam_r y[fcon[r]] F[:, r]
__global__ void lowrank_scale_kernel(int n, int R, const double* F, const double* lam,
                                     const int* fcon, const double* y, double* FW) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * R) return;
  const int r = (int)(e / n);
  FW[e] = F[e] * (lam[r] * y[fcon[r]]);
}

int jtprod(hipStream_t st, const BlockOps& b, const double* y, double* out) {
  MSDP_TRY(rec_forbid("jtprod"));
  const int64_t n2 = (int64_t)b.n * b.n;
  if (b.kind == BlockOps::EMPTY) return fill(st, n2, 0.0, out);
  if (b.kind == BlockOps::LOWRANK) {
    lowrank_scale_kernel<<<cdiv((int64_t)b.n * b.R, 256), 256, 0, st>>>(b.n, b.R, b.F, b.lam,
                                                                        b.fcon, y, b.FW);
    MSDP_HIP(hipGetLastError());
    MSDP_TRY(gemm(st, false, true, b.n, b.n, b.R, 1.0, b.FW, b.n, b.F, b.n, 0.0, out, b.n));
    return symmetrize(st, b.n, out, b.n);
  }
  if (b.kind == BlockOps::COO) {
    MSDP_TRY(fill(st, n2, 0.0, out));
    jtprod_coo_kernel<<<cdiv(b.npos, 256), 256, 0, st>>>(b.npos, b.pos_lin, b.pos_ptr, b.pos_ent,
                                                          b.con, b.val, y, out);
  } else {
    const int split = b.jt_part ? b.jt_split : 1;
    const int per = cdiv(b.ncon, split);
    jtprod_dense_kernel<<<dim3(cdiv(n2, 256), split), 256, 0, st>>>(
        n2, b.ncon, per, b.con_list, b.mats, y, split > 1 ? b.jt_part : out);
    if (split > 1) {
      MSDP_HIP(hipGetLastError());
      jtprod_reduce_kernel<<<cdiv(n2, 256), 256, 0, st>>>(n2, split, b.jt_part, out);
    }
  }
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------ Schur assembly ----
// One-hot constraints A_j = v_j e_{r_j} e_{r_j}':  H[c_a, c_b] += v_a v_b L[r_a, r_b] R[r_b, r_a]
__global__ void schur_onehot_kernel(int cnt, const int* con, const int* row, const double* val,
                                    const double* L, int64_t ldl, const double* R, int64_t ldr,
                                    double* H, int64_t ldh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (a >= cnt || a < b) return;
  // L and R are symmetric (W, S^-1 or X), so R[r_b, r_a] is read as R[r_a, r_b]
  // to keep the loads of a wavefront contiguous.
  const int ra = row[a], rb = row[b];
  H[con[a] + (int64_t)con[b] * ldh] += val[a] * val[b] * L[ra + (int64_t)rb * ldl] * R[ra + (int64_t)rb * ldr];
}

// General COO: H[j,k] += sum_{e in j, f in k} v_e v_f L[c_e, r_f] R[c_f, r_e]
__global__ void schur_coo_kernel(int ncon, const int* con_list, const int* con_ptr,
                                 const int* row, const int* col, const double* val,
                                 const double* L, int64_t ldl, const double* R, int64_t ldr,
                                 double* H, int64_t ldh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (a >= ncon || a < b) return;
  double s = 0.0;
  for (int e = con_ptr[a]; e < con_ptr[a + 1]; ++e)
    for (int f = con_ptr[b]; f < con_ptr[b + 1]; ++f)
      s += val[e] * val[f] * L[col[e] + (int64_t)row[f] * ldl] * R[col[f] + (int64_t)row[e] * ldr];
  H[con_list[a] + (int64_t)con_list[b] * ldh] += s;
}

// svec with sqrt(2) off the diagonal over the lower triangle in column-major
// order (column c holds rows c..n-1 at c n - c (c - 1) / 2): consecutive
// entries are consecutive rows of P_j, so the reads are coalesced, and the
// lower triangle is the only one the product computes.  H = B'B and the
// refinement operator B'(B v) do not depend on the order of B's rows.
__global__ void svec_pack_kernel(int n, int cb, const double* P, double* B, int64_t ldb) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= d || j >= cb) return;
  // invert k -> (r, c): the last column c whose offset is <= k
  auto off = [n](int64_t c) { return c * n - c * (c - 1) / 2; };
  const double b2 = 2.0 * n + 1.0;
  int64_t c = (int64_t)((b2 - sqrt(fmax(b2 * b2 - 8.0 * (double)k, 0.0))) * 0.5);
  c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
  while (c + 1 < n && off(c + 1) <= k) ++c;
  while (off(c) > k) --c;
  const int64_t r = c + (k - off(c));
  const double v = P[(int64_t)j * n * n + r + c * n];
  B[k + (int64_t)j * ldb] = r == c ? v : 1.4142135623730951 * v;
}

int svec_pack(hipStream_t s, int n, int count, const double* P, double* B, int64_t ldb) {
  const int64_t d = (int64_t)n * (n + 1) / 2;
  if (d <= 0 || count <= 0) return 0;
  svec_pack_kernel<<<dim3(cdiv(d, 256), count), 256, 0, s>>>(n, count, P, B, ldb);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// The dense Schur contribution of one block as recorded ops (the recorder
// batches each across the blocks): T_j = A_j op(F2) and P_j = op(Lm) T_j
// (its lower triangle) for all the block's constraints at once, B[:, j] =
// svec(P_j), Hbuf = B'B (lower triangle), T and P the block's own
// mi n^2 buffers.  NT and dual-centering scalings (svec panels).
int dense_schur_record(hipStream_t st, const BlockOps& b, const SchurScaling& sc, double* T,
                       double* P) {
  const int n = b.n, mi = b.ncon;
  const int64_t n2 = (int64_t)n * n;
  const int64_t drow = dense_panel_rows(n, false);
  if (sc.full_vec || b.kind != BlockOps::DENSE || mi <= 0 || b.dcon_identity) {
    set_error("dense_schur_record: a dense block of constraints with svec panels only");
    return -2;
  }
  MSDP_TRY(gemm_vb_record(st, false, sc.transF2, n, n, n, 1.0, b.mats, n, sc.F2, n, 0.0, T, n, mi,
                          n2, 0, n2, false));
  MSDP_TRY(gemm_vb_record(st, sc.transL, false, n, n, n, 1.0, sc.Lm, n, T, n, 0.0, P, n, mi, 0, n2,
                          n2, true));
  MSDP_TRY(svec_record(st, n, mi, P, b.Bbuf, drow));
  return gemm_vb_record(st, true, false, mi, mi, (int)drow, 1.0, b.Bbuf, drow, b.Bbuf, drow, 0.0,
                        b.Hbuf, mi, 1, 0, 0, 0, true);
}

// Low-rank: H[j,k] += sum_{r in j, s in k} lam_r lam_s PL[r,s] PR[s,r] with
// PL = F' L F and PR = F' R F (R x R, ld R)
__global__ void schur_lowrank_kernel(int ncon, const int* con_list, const int* con_ptr,
                                     const double* lam, const double* PL, const double* PR, int R,
                                     double* H, int64_t ldh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (a >= ncon || a < b) return;
  double s = 0.0;
  for (int r = con_ptr[a]; r < con_ptr[a + 1]; ++r)
    for (int q = con_ptr[b]; q < con_ptr[b + 1]; ++q)
      s += lam[r] * lam[q] * PL[r + (int64_t)q * R] * PR[q + (int64_t)r * R];
  H[con_list[a] + (int64_t)con_list[b] * ldh] += s;
}

__global__ void scatter_lower_kernel(int mi, const int* dcon, const double* Hb, double* H,
                                     int64_t ldh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (a >= mi || a < b) return;
  H[dcon[a] + (int64_t)dcon[b] * ldh] += Hb[a + (int64_t)b * mi];
}

int scatter_lower_add(hipStream_t st, int mi, const int* con, const double* Hb, double* H,
                      int64_t ldh) {
  if (mi <= 0) return 0;
  dim3 g(cdiv(mi, 128), mi);
  scatter_lower_kernel<<<g, 128, 0, st>>>(mi, con, Hb, H, ldh);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int64_t dense_panel_rows(int n, bool full_vec) {
  return full_vec ? (int64_t)n * n : (int64_t)n * (n + 1) / 2;
}

// The scaled constraint panel of a dense block in FP64: P_j = op(Lm) (A_j
// op(F2)) by batched products, b.Bbuf[:, j] = svec(P_j) (vec(P_j) for HKM),
// so that <A_j, L A_k R> = B[:, j]' B[:, k].
int dense_panel(hipStream_t st, const BlockOps& b, const SchurScaling& sc, double* work,
                size_t work_elems) {
  MSDP_TRY(rec_forbid("dense_panel"));
  const int n = b.n;
  const int64_t n2 = (int64_t)n * n;
  const int64_t drow = dense_panel_rows(n, sc.full_vec);
  const int mi = b.ncon;
  double* Bm = b.Bbuf;  // drow x mi
  // per-chunk buffers for T_j and P_j
  const int64_t per = 2 * n2;
  int cb = (int)std::max<int64_t>(1, std::min<int64_t>(mi, (int64_t)(work_elems / per)));
  if (cb > 64 && cb < mi) cb = cb / 64 * 64;  // whole 64-constraint batches (measured faster)
  double* T = work;
  double* P = work + (int64_t)cb * n2;
  for (int j0 = 0; j0 < mi; j0 += cb) {
    const int c = std::min(cb, mi - j0);
    MSDP_TRY(gemm(st, false, sc.transF2, n, n, n, 1.0, b.mats + (int64_t)j0 * n2, n, sc.F2, n, 0.0,
                  T, n, c, n2, 0, n2));
    // P_j is symmetric: svec packs its lower triangle, the only part computed
    double* dst = sc.full_vec ? Bm + (int64_t)j0 * n2 : P;
    MSDP_TRY(gemm(st, sc.transL, false, n, n, n, 1.0, sc.Lm, n, T, n, 0.0, dst, n, c, 0, n2, n2,
                  !sc.full_vec));
    if (!sc.full_vec) {
      dim3 g(cdiv(drow, 256), c);
      svec_pack_kernel<<<g, 256, 0, st>>>(n, c, P, Bm + (int64_t)j0 * drow, drow);
      MSDP_HIP(hipGetLastError());
    }
  }
  return 0;
}

int schur_block(hipStream_t st, const BlockOps& b, const SchurScaling& sc, double* H, int64_t m,
                double* work, size_t work_elems) {
  MSDP_TRY(rec_forbid("schur_block"));
  const int n = b.n;
  if (b.kind == BlockOps::EMPTY || b.ncon == 0) return 0;
  if (b.kind == BlockOps::LOWRANK) {
    // PL = F' (L F); PR = F' (R F), the same buffer when L == R
    MSDP_TRY(gemm(st, false, false, n, b.R, n, 1.0, sc.L, n, b.F, n, 0.0, b.FW, n));
    MSDP_TRY(gemm(st, true, false, b.R, b.R, n, 1.0, b.F, n, b.FW, n, 0.0, b.PL, b.R));
    const double* PR = b.PL;
    if (sc.R != sc.L) {
      MSDP_TRY(gemm(st, false, false, n, b.R, n, 1.0, sc.R, n, b.F, n, 0.0, b.FW, n));
      MSDP_TRY(gemm(st, true, false, b.R, b.R, n, 1.0, b.F, n, b.FW, n, 0.0, b.PR, b.R));
      PR = b.PR;
    }
    dim3 g(cdiv(b.ncon, 128), b.ncon);
    schur_lowrank_kernel<<<g, 128, 0, st>>>(b.ncon, b.con_list, b.con_ptr, b.lam, b.PL, PR, b.R,
                                           H, m);
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  if (b.kind == BlockOps::COO) {
    dim3 g(cdiv(b.ncon, 128), b.ncon);
    if (b.one_hot)
      schur_onehot_kernel<<<g, 128, 0, st>>>(b.ncon, b.con, b.row, b.val, sc.L, n, sc.R, n, H, m);
    else
      schur_coo_kernel<<<g, 128, 0, st>>>(b.ncon, b.con_list, b.con_ptr, b.row, b.col, b.val,
                                          sc.L, n, sc.R, n, H, m);
    MSDP_HIP(hipGetLastError());
    return 0;
  }
  // dense: B[:, j] = svec(P_j) or vec(P_j) (dense_panel); Hb = B'B,
  // scattered into H unless H is null (then the caller scatters b.Hbuf later)
  MSDP_TRY(dense_panel(st, b, sc, work, work_elems));
  const int64_t drow = dense_panel_rows(b.n, sc.full_vec);
  const int mi = b.ncon;
  double* Bm = b.Bbuf;
  if (b.dcon_identity) {
    MSDP_TRY(gemm(st, true, false, mi, mi, (int)drow, 1.0, Bm, drow, Bm, drow, 1.0, H, m, 1, 0, 0,
                  0, true));
  } else {
    double* Hb = b.Hbuf;
    MSDP_TRY(gemm(st, true, false, mi, mi, (int)drow, 1.0, Bm, drow, Bm, drow, 0.0, Hb, mi, 1, 0,
                  0, 0, true));
    if (H) MSDP_TRY(scatter_lower_add(st, mi, b.con_list, Hb, H, m));
  }
  return 0;
}

// ------------------------------------------------------ scalar block ------
__global__ void scalar_kernel(int op, int ns, double mu, const double* x, const double* z,
                              const double* zinv, const double* a, const double* b,
                              double* out, double* out2, int corrected) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  switch (op) {
    case SOP_ZINV: out[k] = 1.0 / z[k]; break;
    case SOP_RHS: {  // v = x - mu zinv + x zinv Rd_lin (- corr_lin)    a = Rd_lin, b = corr_lin
      double v = x[k] - mu * zinv[k] + x[k] * zinv[k] * a[k];
      if (corrected) v -= b[k];
      out[k] = v;
      break;
    }
    case SOP_DIRS: {  // out = delz = Rd_lin - CtDy (a = Rd_lin, b = CtDy) ; out2 = delx
      const double dz = a[k] - b[k];
      out[k] = dz;
      double dx = mu * zinv[k] - x[k] - x[k] * zinv[k] * dz;
      if (corrected) dx += out2[k];  // out2 holds corr_lin on entry
      out2[k] = dx;
      break;
    }
    case SOP_CORR:  // out = -delx delz zinv   (a = delx, b = delz)
      out[k] = -a[k] * b[k] * zinv[k];
      break;
    case SOP_XZ:  // out = x .* zinv (Schur scalar scaling)
      out[k] = x[k] * zinv[k];
      break;
  }
}

int scalar_op(hipStream_t st, int op, int ns, double mu, const double* x, const double* z,
              const double* zinv, const double* a, const double* b, double* out, double* out2,
              bool corrected) {
  if (ns <= 0) return 0;
  scalar_kernel<<<cdiv(ns, 256), 256, 0, st>>>(op, ns, mu, x, z, zinv, a, b, out, out2,
                                                corrected ? 1 : 0);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// min over dx < 0 of -x/dx (+inf if none), and sum (x + a dx)(z + b dz)
__global__ __launch_bounds__(256) void scalar_reduce_kernel(int op, int ns, const double* x,
                                                            const double* dx, const double* z,
                                                            const double* dz, double a,
                                                            double bb, double* out) {
  __shared__ double red[4];
  if (op == 0) {
    double v = INFINITY;
    for (int k = threadIdx.x; k < ns; k += 256)
      if (dx[k] < 0) v = fmin(v, -x[k] / dx[k]);
    v = block_reduce<2>(v, red);
    if (threadIdx.x == 0) *out = v;
  } else {
    double v = 0.0;
    for (int k = threadIdx.x; k < ns; k += 256) v += (x[k] + a * dx[k]) * (z[k] + bb * dz[k]);
    v = block_reduce<0>(v, red);
    if (threadIdx.x == 0) *out = v;
  }
}

int scalar_min_step(hipStream_t st, int ns, const double* x, const double* dx, double* out) {
  scalar_reduce_kernel<<<1, 256, 0, st>>>(0, ns, x, dx, nullptr, nullptr, 0, 0, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

int scalar_affine(hipStream_t st, int ns, const double* x, const double* dx, const double* z,
                  const double* dz, double a, double b, double* out) {
  scalar_reduce_kernel<<<1, 256, 0, st>>>(1, ns, x, dx, z, dz, a, b, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------- NT correction / E ------
// RNT_pq = -(M_pq + M_qp) / (D_p + D_q)   (src/directions.jl:97-99)
__global__ void nt_rnt_kernel(int n, const double* M, const double* D, double* R) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
  if (p >= n) return;
  R[p + (int64_t)q * n] = -(M[p + (int64_t)q * n] + M[q + (int64_t)p * n]) / (D[p] + D[q]);
}

int nt_rnt(hipStream_t st, int n, const double* M, const double* D, double* R) {
  if (rec_active()) {
    RecOp op;
    op.kind = REC_EW;
    op.e = EwItem{EW_RNT, n, n, 0, 0, n, n, 0, 0, 0, M, D, nullptr, R};
    op.run = [=] { return nt_rnt(st, n, M, D, R); };
    return rec_add(std::move(op));
  }
  dim3 g(cdiv(n, 256), n);
  nt_rnt_kernel<<<g, 256, 0, st>>>(n, M, D, R);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// D = sqrt(lambda); *minlam for the positivity check
__global__ __launch_bounds__(256) void sqrt_min_kernel(int n, const double* lam, double* D,
                                                       double* minlam) {
  __shared__ double red[4];
  double v = INFINITY;
  for (int k = threadIdx.x; k < n; k += 256) {
    v = fmin(v, lam[k]);
    D[k] = sqrt(lam[k]);
  }
  v = block_reduce<2>(v, red);
  if (threadIdx.x == 0) *minlam = v;
}

int sqrt_min(hipStream_t st, int n, const double* lam, double* D, double* minlam) {
  if (rec_active()) {
    RecOp op;
    op.kind = REC_RED;
    op.r = RedItem{RED_SQRTMIN, 0, n, lam, nullptr, minlam, D};
    op.run = [=] { return sqrt_min(st, n, lam, D, minlam); };
    return rec_add(std::move(op));
  }
  sqrt_min_kernel<<<1, 256, 0, st>>>(n, lam, D, minlam);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// out = W diag(a) : column scaling by the diagonal of M (one-hot A*(y))
__global__ void scale_by_diag_kernel(int n, const double* W, const double* M, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  out[i + (int64_t)j * n] = W[i + (int64_t)j * n] * M[j + (int64_t)j * n];
}

int scale_by_diag(hipStream_t st, int n, const double* W, const double* M, double* out) {
  if (rec_active()) {
    RecOp op;
    op.kind = REC_EW;
    op.e = EwItem{EW_SCALE_BY_DIAG, n, n, 0, 0, n, n, 0, 0, 0, W, M, nullptr, out};
    op.run = [=] { return scale_by_diag(st, n, W, M, out); };
    return rec_add(std::move(op));
  }
  dim3 g(cdiv(n, 256), n);
  scale_by_diag_kernel<<<g, 256, 0, st>>>(n, W, M, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------- triangular solves ------
// H y = b with H = L L' (potrs).  The factorisation leaves the inverses of
// the 64 x 64 diagonal blocks (dinv) and L' is kept column-major (Lt).  Each
// sweep walks super-blocks of 512 rows: one workgroup solves the 512 x 512
// diagonal triangle in eight 64-row steps (x_j = D_j b_j, then the rows below
// within the super-block), then a grid-wide kernel applies the super-block's
// panel to every remaining row.  32 launches per sweep at m = 8000.
constexpr int TRSV_SB = 512;

__global__ __launch_bounds__(1024) void trsv_super_kernel(int K0, int KB, const double* M,
                                                          int64_t ld, const double* dinv,
                                                          const double* rhs, double* out,
                                                          int upper) {
  __shared__ double Ds[64][65];
  __shared__ double bs[TRSV_SB];
  __shared__ double xs[64];
  __shared__ double part[16][65];
  __shared__ double part2[4][256];
  const int t = threadIdx.x;
  for (int i = t; i < KB; i += 1024) bs[i] = rhs[K0 + i];
  const int nsub = (KB + 63) / 64;
  for (int s = 0; s < nsub; ++s) {
    const int j = upper ? nsub - 1 - s : s;
    const int j0 = 64 * j, jb = min(64, KB - j0);
    const double* D = dinv + (size_t)j * 4096;
    for (int e = t; e < 4096; e += 1024) Ds[e & 63][e >> 6] = D[e];
    __syncthreads();
    {  // x_j = D_j b_j (forward) or D_j' b_j (backward): row r, 4 columns per thread
      const int r = t & 63, g = t >> 6;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * g + q;
        if (c < jb) v += (upper ? Ds[c][r] : Ds[r][c]) * bs[j0 + c];
      }
      part[g][r] = v;
    }
    __syncthreads();
    if (t < 64) {
      double v = 0.0;
#pragma unroll
      for (int g = 0; g < 16; ++g) v += part[g][t];
      xs[t] = t < jb ? v : 0.0;
      if (t < jb) out[K0 + j0 + t] = v;
    }
    __syncthreads();
    // remaining rows of the super-block: below (forward) or above (backward)
    const int lo = upper ? 0 : j0 + jb, hi = upper ? j0 : KB;
    for (int base = lo; base < hi; base += 256) {
      const int ri = t & 255, cg = t >> 8;
      const int i = base + ri;
      double v = 0.0;
      if (i < hi) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int c = 16 * cg + q;
          if (c < jb) v += M[(K0 + i) + (int64_t)(K0 + j0 + c) * ld] * xs[c];
        }
      }
      part2[cg][ri] = v;
      __syncthreads();
      if (t < 256 && i < hi) bs[i] -= (part2[0][ri] + part2[1][ri]) + (part2[2][ri] + part2[3][ri]);
      __syncthreads();
    }
  }
}

// target[r0 + r] -= sum_{c < KB} M[r0 + r, K0 + c] v[K0 + c], r < rows
__global__ __launch_bounds__(256) void trsv_panel_kernel(int r0, int rows, int K0, int KB,
                                                         const double* M, int64_t ld,
                                                         const double* v, double* target) {
  __shared__ double xs[TRSV_SB];
  __shared__ double part[4][64];
  const int t = threadIdx.x;
  for (int i = t; i < KB; i += 256) xs[i] = v[K0 + i];
  __syncthreads();
  const int lane = t & 63, w = t >> 6;
  const int rr = blockIdx.x * 64 + lane;
  const int per = (KB + 3) / 4;
  const int c0 = w * per, c1 = min(KB, c0 + per);
  double a0 = 0.0, a1 = 0.0;
  if (rr < rows) {
    const double* col = M + (r0 + rr) + (int64_t)K0 * ld;
    int c = c0;
    for (; c + 1 < c1; c += 2) {
      a0 += col[(int64_t)c * ld] * xs[c];
      a1 += col[(int64_t)(c + 1) * ld] * xs[c + 1];
    }
    if (c < c1) a0 += col[(int64_t)c * ld] * xs[c];
  }
  part[w][lane] = a0 + a1;
  __syncthreads();
  if (w == 0 && rr < rows)
    target[r0 + rr] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// out[K0 + i] = sum_{j <= i} D[i + 512 j] in[K0 + j], i < 512: the inverse of a
// full 512 x 512 diagonal block applied to its part of the vector.  64 rows
// per workgroup, the 4 waves split the columns, partial sums added in order.
__global__ __launch_bounds__(256) void trsv_blockinv_kernel(int K0, const double* D,
                                                            const double* in, double* out) {
  __shared__ double xs[TRSV_SB];
  __shared__ double part[4][64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int j = t; j < TRSV_SB; j += 256) xs[j] = in[K0 + j];
  __syncthreads();
  const int i = blockIdx.x * 64 + lane;
  double a0 = 0.0, a1 = 0.0;
  const int j1 = min(i + 1, (w + 1) * 128);
  int j = w * 128;
  for (; j + 1 < j1; j += 2) {
    a0 += D[i + (int64_t)j * TRSV_SB] * xs[j];
    a1 += D[i + (int64_t)(j + 1) * TRSV_SB] * xs[j + 1];
  }
  if (j < j1) a0 += D[i + (int64_t)j * TRSV_SB] * xs[j];
  part[w][lane] = a0 + a1;
  __syncthreads();
  if (w == 0) out[K0 + i] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// out[K0 + i] = sum_j D[j + 512 i] in[K0 + j] over j >= i (upper = 1: the
// transposed inverse, D' applied) or j <= i (upper = 0: D holds the transposed
// inverse, so column i is row i of the inverse), one wavefront per output:
// the column is read contiguously and summed by wave_sum.
__global__ __launch_bounds__(256) void trsv_blockinv_t_kernel(int K0, const double* D,
                                                              const double* in, double* out,
                                                              int upper) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const double* col = D + (int64_t)i * TRSV_SB;
  const int j0 = upper ? i : 0, j1 = upper ? TRSV_SB : i + 1;
  double a0 = 0.0, a1 = 0.0;
  int j = j0 + lane;
  for (; j + 64 < j1; j += 128) {
    a0 += col[j] * in[K0 + j];
    a1 += col[j + 64] * in[K0 + j + 64];
  }
  if (j < j1) a0 += col[j] * in[K0 + j];
  const double acc = wave_sum(a0 + a1);
  if (lane == 0) out[K0 + i] = acc;
}

// target[r0 + r] -= sum_{c < KB} Mt[(K0 + c) + (r0 + r) ld] v[K0 + c]: the
// super-block's panel applied through the transposed storage, where each
// row's KB coefficients are one contiguous run.  One wavefront per row (the
// run read in 64-wide coalesced steps, summed by wave_sum).
__global__ __launch_bounds__(256) void trsv_rowdot_kernel(int r0, int rows, int K0, int KB,
                                                          const double* __restrict__ Mt, int64_t ld,
                                                          const double* __restrict__ v,
                                                          double* __restrict__ target) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  const double* run = Mt + K0 + (int64_t)(r0 + r) * ld;
  const double* x = v + K0;
  double a0 = 0.0, a1 = 0.0;
  int c = lane;
  for (; c + 64 < KB; c += 128) {
    a0 += run[c] * x[c];
    a1 += run[c + 64] * x[c + 64];
  }
  if (c < KB) a0 += run[c] * x[c];
  const double acc = wave_sum(a0 + a1);
  if (lane == 0) target[r0 + r] -= acc;
}

int potrs(hipStream_t st, int m, const double* L, const double* Lt, int64_t ld,
          const double* dinv, double* b, double* tmp, const double* di512,
          const double* di512t) {
  if (m <= 0) return 0;
  const int nsb = cdiv(m, TRSV_SB);
  // forward: L x = b (x -> tmp, b overwritten by the updates)
  for (int k = 0; k < nsb; ++k) {
    const int K0 = k * TRSV_SB, KB = std::min(TRSV_SB, m - K0);
    if (di512 && KB == TRSV_SB) {
      if (di512t)
        trsv_blockinv_t_kernel<<<TRSV_SB / 4, 256, 0, st>>>(
            K0, di512t + (size_t)k * TRSV_SB * TRSV_SB, b, tmp, 0);
      else
        trsv_blockinv_kernel<<<TRSV_SB / 64, 256, 0, st>>>(
            K0, di512 + (size_t)k * TRSV_SB * TRSV_SB, b, tmp);
    } else {
      trsv_super_kernel<<<1, 1024, 0, st>>>(K0, KB, L, ld, dinv + (size_t)(K0 / 64) * 4096, b, tmp, 0);
    }
    const int rest = m - K0 - KB;
    if (rest > 0) {
      // rows below: row r of L's columns K0.. is column r of L' (contiguous)
      if (Lt)
        trsv_rowdot_kernel<<<cdiv(rest, 4), 256, 0, st>>>(K0 + KB, rest, K0, KB, Lt, ld, tmp, b);
      else
        trsv_panel_kernel<<<cdiv(rest, 64), 256, 0, st>>>(K0 + KB, rest, K0, KB, L, ld, tmp, b);
    }
  }
  // backward: L' z = x (z -> b, tmp overwritten by the updates)
  for (int k = nsb - 1; k >= 0; --k) {
    const int K0 = k * TRSV_SB, KB = std::min(TRSV_SB, m - K0);
    if (di512 && KB == TRSV_SB)
      trsv_blockinv_t_kernel<<<TRSV_SB / 4, 256, 0, st>>>(K0, di512 + (size_t)k * TRSV_SB * TRSV_SB, tmp, b, 1);
    else
      trsv_super_kernel<<<1, 1024, 0, st>>>(K0, KB, Lt, ld, dinv + (size_t)(K0 / 64) * 4096, tmp, b, 1);
    // rows above: row r of L' 's columns K0.. is column r of L (contiguous)
    if (K0 > 0) trsv_rowdot_kernel<<<cdiv(K0, 4), 256, 0, st>>>(0, K0, K0, KB, L, ld, b, tmp);
  }
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp

namespace msdp {

// A[:, j] *= d[j]
__global__ void scale_cols_vec_kernel(int rows, double* A, int64_t lda, const double* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < rows) A[i + (int64_t)j * lda] *= d[j];
}

int scale_cols_vec(hipStream_t st, int rows, int cols, double* A, int64_t lda, const double* d) {
  if (rows <= 0 || cols <= 0) return 0;
  dim3 g(cdiv(rows, 256), cols);
  scale_cols_vec_kernel<<<g, 256, 0, st>>>(rows, A, lda, d);
  MSDP_HIP(hipGetLastError());
  return 0;
}

__global__ __launch_bounds__(256) void vec_min_kernel(int n, const double* v, double* out) {
  __shared__ double red[4];
  double r = INFINITY;
  for (int k = threadIdx.x; k < n; k += 256) r = fmin(r, v[k]);
  r = block_reduce<2>(r, red);
  if (threadIdx.x == 0) *out = r;
}

int vec_min(hipStream_t st, int n, const double* v, double* out) {
  vec_min_kernel<<<1, 256, 0, st>>>(n, v, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// H[i, j] += L[i, j] R[i, j] for i >= j (one-hot constraints A_j = e_j e_j',
// symmetric L, R): the Hadamard Schur complement of max-cut.
__global__ void schur_onehot_identity_kernel(int m, const double* L, int64_t ldl, const double* R,
                                             int64_t ldr, double* H, int64_t ldh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m || i < j) return;
  H[i + (int64_t)j * ldh] += L[i + (int64_t)j * ldl] * R[i + (int64_t)j * ldr];
}

int schur_onehot_identity(hipStream_t st, int m, const double* L, int64_t ldl, const double* R,
                          int64_t ldr, double* H, int64_t ldh) {
  if (m <= 0) return 0;
  dim3 g(cdiv(m, 256), m);
  schur_onehot_identity_kernel<<<g, 256, 0, st>>>(m, L, ldl, R, ldr, H, ldh);
  MSDP_HIP(hipGetLastError());
  return 0;
}

// DualCentering step length (src/dual_centering.jl:146-201): alpha_max from
// the proximity eigenvalues, then bisection on xi(alpha) = A; one workgroup,
// every xi evaluation a block reduction.
__global__ __launch_bounds__(1024) void dc_alpha_kernel(int total, const double* mu, double A,
                                                        double* out) {
  __shared__ double red[16];
  const double eps = 2.220446049250313e-16;
  double am = 1.0;
  for (int k = threadIdx.x; k < total; k += blockDim.x) {
    const double v = mu[k];
    if (v > 0) am = fmin(am, 1.0 / v);
    if (v < 0) am = fmin(am, 1.0 / (1.0 - v));
  }
  am = block_reduce<2>(am, red);
  auto xi = [&](double a) {
    const double b2 = a / (1 - a);
    double s = 0.0, bad = 0.0;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
      const double u = 1 - a * mu[k], v = 1 + b2 * mu[k];
      if (u > 0 && v > 0) s -= log(u) + log(v);
      else bad = 1.0;
    }
    s = block_reduce<0>(s, red);
    bad = block_reduce<1>(bad, red);
    return bad > 0 ? INFINITY : s;
  };
  double hi = am * 0.99999;
  const double xh = xi(hi);
  double res;
  if (!isfinite(xh) || xh < A) {
    res = hi;
  } else {
    double lo = 0.0;
    for (int it = 0; it < 80; ++it) {
      const double mid = (lo + hi) / 2;
      if (xi(mid) < A) lo = mid; else hi = mid;
      if (hi - lo <= eps * fmax(1.0, hi)) break;
    }
    res = (lo + hi) / 2;
  }
  if (threadIdx.x == 0) *out = res;
}

int dc_solve_alpha(hipStream_t st, int total, const double* mu, double A, double* out) {
  dc_alpha_kernel<<<1, 1024, 0, st>>>(total, mu, A, out);
  MSDP_HIP(hipGetLastError());
  return 0;
}

}  // namespace msdp
