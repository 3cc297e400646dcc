// Constraint operators, Schur assembly, scalar-block and small solver
// kernels (ops.hip).
#pragma once
#include "common.h"

namespace msdp {

// Device view of the constraint matrices touching one PSD block.
struct BlockOps {
  enum Kind { EMPTY = 0, COO = 1, DENSE = 2, LOWRANK = 3 };
  int kind = EMPTY;
  int n = 0;
  bool one_hot = false;       // every constraint is v e_r e_r'
  int ncon = 0;               // constraints touching the block
  int* con_list = nullptr;    // [ncon] sorted constraint indices
  // COO: entries sorted by constraint, con_ptr[ncon+1] ranges
  int64_t nnz = 0;
  int* con_ptr = nullptr;
  int* con = nullptr;
  int* row = nullptr;
  int* col = nullptr;
  double* val = nullptr;
  // A*(y) by position: pos_lin[npos] linear index, pos_ptr[npos+1], pos_ent
  int npos = 0;
  int64_t* pos_lin = nullptr;
  int* pos_ptr = nullptr;
  int* pos_ent = nullptr;
  // DENSE: ncon consecutive n x n matrices
  double* mats = nullptr;
  bool dcon_identity = false;  // con_list == 0..m-1
  double* Bbuf = nullptr;      // svec/vec panel of the scaled constraints
  double* Hbuf = nullptr;      // ncon x ncon block of H before the scatter
  int jt_split = 1;            // A*(y): constraint ranges summed separately
  double* jt_part = nullptr;   // [jt_split x n^2] partial sums
  // LOWRANK: A_j = sum_{r in con_ptr[j]..con_ptr[j+1]} lam_r f_r f_r' with the
  // factor columns f_r of F (n x R, column-major), fcon[r] = the constraint
  // index of column r; FW (n x R) and PL, PR (R x R) are work buffers
  int R = 0;
  double* F = nullptr;
  double* lam = nullptr;
  int* fcon = nullptr;
  double* FW = nullptr;
  double* PL = nullptr;
  double* PR = nullptr;
};

// number of constraint ranges of the dense A*(y) for a block of side n
int jtprod_split(int n, int ncon);

// (L, R) of H_jk = <A_j, L A_k R> and, for dense blocks, the factored form
// P_j = op(Lm) A_j op(F2) with H_jk = <P_j, P_k>.
struct SchurScaling {
  const double* L = nullptr;
  const double* R = nullptr;
  const double* Lm = nullptr;
  bool transL = false;
  const double* F2 = nullptr;
  bool transF2 = false;
  bool full_vec = false;  // HKM: P_j is not symmetric
};

int jprod(hipStream_t st, const BlockOps& b, const double* X, double* out);
int jtprod(hipStream_t st, const BlockOps& b, const double* y, double* out);
int scatter_lower_add(hipStream_t st, int mi, const int* con, const double* Hb, double* H,
                      int64_t ldh);
int schur_block(hipStream_t st, const BlockOps& b, const SchurScaling& sc, double* H, int64_t m,
                double* work, size_t work_elems);
// the block's contribution as recorded ops into its Hbuf (T, P: mi n^2 each)
int dense_schur_record(hipStream_t st, const BlockOps& b, const SchurScaling& sc, double* T,
                       double* P);
// rows of a dense block's scaled constraint panel: n(n+1)/2 (svec) or n^2 (vec)
int64_t dense_panel_rows(int n, bool full_vec);
// b.Bbuf[:, j] = svec(P_j) (vec for HKM) in FP64, P_j = op(Lm) A_j op(F2)
int dense_panel(hipStream_t st, const BlockOps& b, const SchurScaling& sc, double* work,
                size_t work_elems);

// A(.) and A*(.) of every block in two launches each, for problems whose
// blocks are all dense (two or more): one workgroup per (block, constraint)
// pair writes its inner product to part, then each constraint adds its
// pairs in block order (the additions of the per-block launches, in their
// order); A*(y) runs one thread per entry of every block.
struct DenseBlkDev {
  const double* mats;
  int64_t n2;
  const int* con_list;
  int ncon, pair0, tile0, pad;
  const double* in[3];  // the A(.) arguments: X, buf, S^-1
  double* out[2];       // the A*(.) results: tmp2, Dc
};
enum { JP_X = 0, JP_BUF = 1, JP_SI = 2 };
enum { JT_TMP2 = 0, JT_DC = 1 };
struct DenseBatch {
  int nblk = 0, npairs = 0, ntiles = 0;
  DenseBlkDev* blks = nullptr;
  int* cptr = nullptr;    // [m + 1]: constraint c's pairs at cpairs[cptr[c]..cptr[c + 1])
  int* cpairs = nullptr;  // [npairs], block order within a constraint
  double* part = nullptr;  // [npairs]
};
int jprod_dense_batched(hipStream_t st, const DenseBatch& d, int m, int which, double* out);
int jtprod_dense_batched(hipStream_t st, const DenseBatch& d, const double* y, int which);

enum { SOP_ZINV = 0, SOP_RHS = 1, SOP_DIRS = 2, SOP_CORR = 3, SOP_XZ = 4 };
int scalar_op(hipStream_t st, int op, int ns, double mu, const double* x, const double* z,
              const double* zinv, const double* a, const double* b, double* out, double* out2,
              bool corrected);
int scalar_min_step(hipStream_t st, int ns, const double* x, const double* dx, double* out);
int scalar_affine(hipStream_t st, int ns, const double* x, const double* dx, const double* z,
                  const double* dz, double a, double b, double* out);
int nt_rnt(hipStream_t st, int n, const double* M, const double* D, double* R);
int sqrt_min(hipStream_t st, int n, const double* lam, double* D, double* minlam);
int scale_by_diag(hipStream_t st, int n, const double* W, const double* M, double* out);
// H y = b from H = L L' (L and its transpose Lt, the 64 x 64 diagonal-block
// inverses dinv); with di512 (the inverses of the full 512 x 512 diagonal
// blocks, consecutive) those blocks are applied as products, with di512t
// (their transposes) the forward ones read row-wise too.
int potrs(hipStream_t st, int m, const double* L, const double* Lt, int64_t ld,
          const double* dinv, double* b, double* tmp, const double* di512 = nullptr,
          const double* di512t = nullptr);
int scale_cols_vec(hipStream_t st, int rows, int cols, double* A, int64_t lda, const double* d);
int vec_min(hipStream_t st, int n, const double* v, double* out);
int schur_onehot_identity(hipStream_t st, int m, const double* L, int64_t ldl, const double* R,
                          int64_t ldr, double* H, int64_t ldh);
int dc_solve_alpha(hipStream_t st, int total, const double* mu, double A, double* out);

}  // namespace msdp
