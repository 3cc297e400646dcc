// Dense Householder QR Schur solver kernels (qr.hip).
#pragma once
#include "ops.h"

namespace msdp {
int qr_pack_dense(hipStream_t s, int n, int cnt, const double* P, const int* con, double* B,
                  int64_t ldb, int64_t off);
int qr_pack_coo(hipStream_t s, const BlockOps& b, const double* G, double* B, int64_t ldb,
                int64_t off);
int qr_pack_lowrank(hipStream_t s, const BlockOps& b, const double* G, double* B, int64_t ldb,
                    int64_t off);
int qr_scalar_rows(hipStream_t s, int m, int ns, const double* Clin, const double* x,
                   const double* z, double* B, int64_t ldb, int64_t off);
// panel width of the blocked factorisation and rows per workgroup of the
// vector sweeps
constexpr int QR_NB = 64;
constexpr int QR_CH = 256;
// doubles of the per-panel T factors, and of the factorisation's and the
// vector sweeps' workspace
size_t qr_t_elems(int m);
size_t qr_work_elems(int64_t rows, int m);
// B (rows x m, ld rows, rows >= m) <- scaled columns, then V below and R on
// and above the diagonal; tau, T per panel; bad = zero or non-finite column;
// rdiag_out = {max |R_kk|, min |R_kk|}
int qr_factor(hipStream_t s, int64_t rows, int m, double* B, double* scale, double* tau,
              double* T, double* work, int* bad, double* rdiag_out);
// y <- Q'y (trans) or Qy
int qr_apply_vec(hipStream_t s, int64_t rows, int m, const double* B, const double* T,
                 double* work, double* y, bool trans);
int qr_trsv(hipStream_t s, int m, const double* B, int64_t ldb, double* x, bool trans);
int qr_base_block(hipStream_t s, int n, const double* P, const double* D, double mu,
                  const double* RNT, bool corrected, double* base, int64_t off);
int qr_base_scalar(hipStream_t s, int ns, const double* x, const double* z, const double* Rdlin,
                   const double* corrlin, double mu, bool corrected, double* base, int64_t off);
int qr_combine(hipStream_t s, int m, const double* c, const double* q1, double* t);
int qr_mul(hipStream_t s, int m, const double* a, const double* b, double* out);
int qr_smat(hipStream_t s, int n, const double* dir, int64_t off, double* M);
int qr_delx(hipStream_t s, int ns, const double* x, const double* z, const double* dir,
            int64_t off, double* delx);
}  // namespace msdp
