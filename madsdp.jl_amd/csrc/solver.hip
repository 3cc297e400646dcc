// MadSDP interior-point driver on one gfx950 device.
//
// The state of `MadSDPSolver` (reference src/solver.jl:7-49) lives in device
// memory; this file sequences the device kernels exactly as the reference's
// driver and components do (src/solver.jl, src/primal_dual.jl, src/step.jl,
// src/directions.jl, src/blocks.jl, src/schur_solvers.jl,
// src/dual_centering.jl).  Only scalars cross to the host: the step lengths,
// mu, the termination quantities and the factorisation statuses, i.e. the
// values the reference's control flow branches on.
#include "ops.h"
#include "qr.h"
#include "madsdp_hip.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <vector>

using namespace msdp;

namespace {

constexpr double EPS = 2.220446049250313e-16;

enum Status {
  SOLVE_SUCCEEDED = 1,
  SEARCH_DIRECTION_BECOMES_TOO_SMALL = 3,
  MAXIMUM_ITERATIONS_EXCEEDED = 6,
  MAXIMUM_WALLTIME_EXCEEDED = 7,
  INITIAL = 11,
  REGULAR = 12,
  ERROR_IN_STEP_COMPUTATION = -3,
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Per-PSD-block device state: SDPBlock + PrimalDualBlock + NTWorkspace
// (src/blocks.jl:10-63, src/directions.jl:16-43) and the DualCentering block
// workspace (src/dual_centering.jl:48-55).
struct Blk {
  int n = 0;
  BlockOps ops;
  double* C = nullptr;
  double cnorm = 0.0;
  double *X, *S, *Rd, *Ls, *Si, *Lsinv, *delX, *delS, *Lx, *Lxinv, *corr, *buf, *buf2, *tmp,
      *tmp2, *W, *G, *Gi, *D, *RNT, *Mm, *Nm, *Vv, *lam, *Shat, *Dc, *Lh;
  // T_j and P_j of all the block's constraints (the recorded Schur assembly
  // of many dense blocks), mi n^2 each
  double *asmT = nullptr, *asmP = nullptr;
  bool rd_nonzero = true;
  // mixed precision: FP32 copies of the dense constraints, the scaled
  // constraint panel and the scaling factors
  float *mats32 = nullptr, *B32 = nullptr, *Lm32 = nullptr, *F32 = nullptr;
  // BF16x3 assembly: the split FP64 panel (hi, lo: bf16x3_ld(drow) x ncon);
  // panel_fp64 marks that ops.Bbuf holds the FP64 panel of the current
  // factorisation (the refinement applies H through it)
  __bf16 *Bh = nullptr, *Bl = nullptr;
  bool panel_fp64 = false;
  int64_t panel_rows = 0;
  // host copies of the constraint data until finalize
  std::vector<int> h_con, h_row, h_col;
  std::vector<double> h_val;
  std::vector<int> h_dcon;
  std::vector<double> h_mats;
  std::vector<double> h_C;
  // low-rank: constraint indices, column ranges, factors (n x R), weights
  std::vector<int> h_lcon, h_lptr;
  std::vector<double> h_lF, h_lw;
  bool has_coo = false, has_dense = false, has_lowrank = false;
};

}  // namespace

struct msdp_solver {
  int device = 0;
  hipStream_t st = nullptr;
  bool own_stream = true;
  // host copies of b, d_lin, C_lin until finalize
  std::vector<double> h_b, h_dlin, h_Clin;
  // released: device state freed after a solve; the solution stays on the host
  bool released = false;
  std::vector<double> r_y, r_x, r_z;
  std::vector<std::vector<double>> r_X, r_S;
  double* hsc_pin = nullptr;
  // Multi-block problems whose blocks are all small (n <= JACOBI_VB_MAX): the
  // per-block work of a stage runs on NAUX auxiliary streams (block i on
  // stream i % NAUX) between a fork and a join on the main stream, each with
  // its own scratch.  Accumulations shared by blocks stay on the main stream.
  static constexpr int NAUX = 8;
  bool multi = false;
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> aux_done;
  hipEvent_t forked = nullptr;
  std::vector<double*> aux_trwork, aux_red, aux_swork;
  // Recorded regions of the multi-block stages (Recorder in common.h): the
  // per-block ops of a region become one batched launch per op position.
  // Their arguments go through a ring of pinned host slots (an event per
  // slot guards its reuse) into one device buffer, in stream order.
  static constexpr int REC_RING = 4;
  Recorder rec;
  bool rec_on = false;
  char* rec_host[REC_RING] = {};
  hipEvent_t rec_ev[REC_RING] = {};
  char* rec_dev = nullptr;
  size_t rec_cap = 0;
  int64_t alloc_bytes = 0;  // device bytes of the solver's own allocations
  int rec_slot = 0;
  int nb = 0;
  int64_t m = 0, ns = 0;
  std::vector<Blk> blk;
  // all blocks dense (two or more): A(.) / A*(.) of every block in one go
  DenseBatch dbatch;
  static constexpr size_t ASM_BATCH_MAX_BYTES = (size_t)2 << 30;
  std::vector<DenseBlkDev> dbh_blks;
  std::vector<int> dbh_cptr, dbh_cpairs;
  int build_dense_batch();
  // A(.) of every block accumulated into out (which: JP_X, JP_BUF, JP_SI),
  // in block order
  int jprod_blocks(int which, double* out) {
    if (dbatch.nblk > 0) return jprod_dense_batched(st, dbatch, (int)m, which, out);
    for (auto& bk : blk) {
      const double* in = which == JP_X ? bk.X : which == JP_BUF ? bk.buf : bk.Si;
      MSDP_TRY(jprod(st, bk.ops, in, out));
    }
    return 0;
  }
  msdp_options opt;
  bool finalized = false;
  // device vectors
  double *b = nullptr, *dlin = nullptr, *Clin = nullptr, *y = nullptr, *dely = nullptr,
         *Rp = nullptr, *x = nullptr, *z = nullptr, *delx = nullptr, *delz = nullptr,
         *Rdlin = nullptr, *zinv = nullptr, *corrlin = nullptr, *wm = nullptr, *ws1 = nullptr,
         *ws2 = nullptr, *Cs = nullptr;
  double *H = nullptr, *Hs = nullptr, *Ht = nullptr, *Hdinv = nullptr, *wm2 = nullptr;
  // inverses of the factor's full 512 x 512 diagonal blocks (triangular
  // solves apply them as products) and the workspace that forms them
  // and their transposes (the forward sweep reads rows of the inverses as
  // contiguous columns)
  double *Hdi = nullptr, *Hdit = nullptr, *Hdiw = nullptr;
  bool hdi_valid = false;
  // Forms the block inverses once per factorisation (every factorisation
  // serves two or more solves).
  int form_block_inverses() {
    if (!Hdi || hdi_valid) return 0;
    // all full 512-row diagonal blocks in one batched triangular inverse
    MSDP_TRY(trtri_lower_batched(st, 512, (int)(m / 512), H, m, 512 * (m + 1), Hdi, 512,
                                 512 * 512, Hdiw, (int64_t)trtri_work_elems(512)));
    for (int64_t k = 0; k < m / 512; ++k)
      MSDP_TRY(transpose(st, 512, 512, Hdi + k * 512 * 512, 512, Hdit + k * 512 * 512, 512));
    hdi_valid = true;
    return 0;
  }
  int* info = nullptr;    // per-block factorisation infos (4 per block) + 8
  double* sc = nullptr;   // scalar slots
  double* red = nullptr;  // reduction scratch
  double* ework = nullptr;
  int* eiwork = nullptr;
  double* trwork = nullptr;
  double* swork = nullptr;
  size_t swork_elems = 0;
  double* eigall = nullptr;  // DualCentering proximity eigenvalues, all blocks
  // batched Jacobi descriptors of the small (n <= 128) blocks
  EigBatch *eb_nt = nullptr, *eb_step = nullptr, *eb_dc = nullptr;
  int nsmall = 0;
  // mixed precision state
  float* swork32 = nullptr;
  size_t swork32_elems = 0;
  double *rh = nullptr, *rr = nullptr, *rdl = nullptr;
  // BF16x3: panel-operator work (partials, t) and the gathered / product
  // vectors of blocks that touch a subset of the constraints
  double *pwork = nullptr, *pvb = nullptr, *pu = nullptr;
  double reg_used = 0.0;
  bool h_approx = false, force_fp64 = false;
  int lowp_count = 0, fallback_count = 0, refine_steps = 0;
  // QRSchur workspace (qr_schur.jl:5-10): B with the reflectors, scale, tau
  double *qB = nullptr, *qscale = nullptr, *qtau = nullptr, *qbase = nullptr, *qy = nullptr,
         *qc = nullptr, *qt = nullptr, *qrd = nullptr, *qT = nullptr, *qwork = nullptr;
  int* qbad = nullptr;
  int64_t qrows = 0;
  // DualCentering vectors
  double *dc_d = nullptr, *dc_g = nullptr, *dc_dy = nullptr, *dc_yhat = nullptr;
  // host scalars (src/solver.jl:37-48)
  double mu = 0, sigma = 0, alpha = 0, beta = 0, inf_pr = 0, inf_du = 0, inf_compl = 0,
         obj_val = 0, pobj = 0, dobj = 0;
  int iter = 0, status = INITIAL;
  double start_time = 0;
  double bnorm = 0, cnorm_max = 0, dnorm = 0;
  int dc_nu = 0, n_newton = 0, n_predictor = 0;
  double dc_A = 0, dc_t = 0, dc_lambda = 0;
  std::vector<double> ip_X, ip_y, ip_x;
  std::vector<void*> allocs;
  std::vector<std::array<double, 9>> trace;
  int nslots = 0;
  std::vector<double> hsc;

  ~msdp_solver() {
    free_device();
    destroy_aux();
    free_rec();
    if (hsc_pin) (void)hipHostFree(hsc_pin);
    if (st && own_stream) (void)hipStreamDestroy(st);
  }

  // device memory is stream-ordered (hipMallocAsync on the solver's stream),
  // so solves on other streams never wait for this solver's allocations
  int ensure_stream() {
    if (st) return 0;
    MSDP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    own_stream = true;
    return 0;
  }
  template <class T>
  int alloc(T** p, size_t count) {
    MSDP_TRY(ensure_stream());
    void* q = nullptr;
    MSDP_HIP(hipMallocAsync(&q, std::max<size_t>(count, 1) * sizeof(T), st));
    allocs.push_back(q);
    alloc_bytes += (int64_t)(std::max<size_t>(count, 1) * sizeof(T));
    *p = reinterpret_cast<T*>(q);
    return 0;
  }
  void free_device() {
    if (!st || allocs.empty()) return;
    (void)hipStreamSynchronize(st);
    for (void* p : allocs) (void)hipFreeAsync(p, st);
    allocs.clear();
    alloc_bytes = 0;
    (void)hipStreamSynchronize(st);
  }
  int release();
  int keep_solution();
  // the recorded regions' argument buffers (after the stream has drained)
  void free_rec() {
    for (int k = 0; k < REC_RING; ++k) {
      if (rec_ev[k]) (void)hipEventSynchronize(rec_ev[k]);
      if (rec_host[k]) (void)hipHostFree(rec_host[k]);
      if (rec_ev[k]) (void)hipEventDestroy(rec_ev[k]);
      rec_host[k] = nullptr;
      rec_ev[k] = nullptr;
    }
    if (rec_dev) (void)hipFree(rec_dev);
    rec_dev = nullptr;
    rec_cap = 0;
  }
  void destroy_aux() {
    for (hipStream_t q : aux) (void)hipStreamDestroy(q);
    for (hipEvent_t ev : aux_done) (void)hipEventDestroy(ev);
    if (forked) (void)hipEventDestroy(forked);
    aux.clear();
    aux_done.clear();
    aux_trwork.clear();
    aux_red.clear();
    aux_swork.clear();
    forked = nullptr;
    multi = false;
  }
  int setup_aux();
  int fork() {
    if (!multi) return 0;
    MSDP_HIP(hipEventRecord(forked, st));
    for (hipStream_t q : aux) MSDP_HIP(hipStreamWaitEvent(q, forked, 0));
    return 0;
  }
  int join() {
    if (!multi) return 0;
    for (int k = 0; k < NAUX; ++k) {
      MSDP_HIP(hipEventRecord(aux_done[k], aux[k]));
      MSDP_HIP(hipStreamWaitEvent(st, aux_done[k], 0));
    }
    return 0;
  }
  // a recorded region in place of fork() ... join() (multi-block only)
  int rec_begin() {
    if (!multi) return 0;
    rec.ops.assign(nb, {});
    rec.cur = 0;
    rec_active() = &rec;
    rec_on = true;
    return 0;
  }
  int rec_end() {
    if (!rec_on) return 0;
    rec_active() = nullptr;
    rec_on = false;
    return rec_flush();
  }
  int rec_flush();
  // f now, or recorded as one op without a batched form (replayed per block
  // in its OnBlock scope)
  int opaque(std::function<int()> f) {
    if (!rec_active()) return f();
    RecOp o;
    o.run = std::move(f);
    return rec_add(std::move(o));
  }
  // block i's work on its auxiliary stream (and scratch) for the scope
  struct OnBlock {
    msdp_solver* s;
    hipStream_t st0;
    double *tw0, *red0, *sw0;
    OnBlock(msdp_solver* s_, int i) : s(s_), st0(s_->st), tw0(s_->trwork), red0(s_->red), sw0(s_->swork) {
      if (s->rec_on) s->rec.cur = i;
      if (s->multi) {
        const int k = i % NAUX;
        s->st = s->aux[k];
        s->trwork = s->aux_trwork[k];
        s->red = s->aux_red[k];
        s->swork = s->aux_swork[k];
      }
    }
    ~OnBlock() {
      s->st = st0;
      s->trwork = tw0;
      s->red = red0;
      s->swork = sw0;
    }
  };
  int64_t footprint() const;
  // the dense assembly's chunk buffer: 512 MiB, or 64 MiB for problems whose
  // auxiliary streams each hold a copy (setup_aux)
  size_t swork_cap_bytes() const {
    int64_t nmax = 0;
    for (const auto& b : blk) nmax = std::max<int64_t>(nmax, b.n);
    const bool aux_copies = nb >= 2 && nmax <= JACOBI_VB_MAX;
    return aux_copies ? ((size_t)64 << 20) : ((size_t)512 << 20);
  }

  double* slot(int k) { return sc + k; }
  // slots: 0..31 global, 32 + 8*i per block
  int bslot(int i, int k) const { return 32 + 8 * i + k; }
  int read_slots() {
    hsc.resize(nslots);
    MSDP_HIP(hipMemcpyAsync(hsc_pin, sc, sizeof(double) * nslots, hipMemcpyDeviceToHost, st));
    MSDP_TRY(stream_sync(st));
    std::memcpy(hsc.data(), hsc_pin, sizeof(double) * nslots);
    return 0;
  }
  int read_infos(std::vector<int>& out) {
    out.resize(4 * nb + 8);
    MSDP_HIP(hipMemcpyAsync(out.data(), info, sizeof(int) * out.size(), hipMemcpyDeviceToHost, st));
    MSDP_TRY(stream_sync(st));
    return 0;
  }
  // the factorisation infos and the scalar slots behind one wait
  int read_infos_slots(std::vector<int>& out) {
    out.resize(4 * nb + 8);
    hsc.resize(nslots);
    MSDP_HIP(hipMemcpyAsync(out.data(), info, sizeof(int) * out.size(), hipMemcpyDeviceToHost, st));
    MSDP_HIP(hipMemcpyAsync(hsc_pin, sc, sizeof(double) * nslots, hipMemcpyDeviceToHost, st));
    MSDP_TRY(stream_sync(st));
    std::memcpy(hsc.data(), hsc_pin, sizeof(double) * nslots);
    return 0;
  }
  int* binfo(int i, int k) { return info + 8 + 4 * i + k; }

  bool nt() const { return opt.direction == MSDP_DIR_NT; }
  bool dual_centering() const { return opt.algorithm == MSDP_ALG_DUAL_CENTERING; }
  int barrier_dim() const {
    int64_t s = ns;
    for (const auto& b : blk) s += b.n;
    return (int)s;
  }

  int finalize();
  // block kernels
  int cholesky_factor(Blk& b, const double* M, double* L, int* inf);
  int cholesky_factor2(Blk& b, const double* M0, double* L0, const double* M1, double* L1,
                       int* inf);
  int prepare_block(Blk& b, int i, bool factored = false);
  int inverse_product(Blk& b);
  int set_dual_slacks(const double* yv);
  int slack(Blk& b, const double* yv, double* out);
  // PrimalDual
  int update_residuals();
  int prepare(bool& ok);
  int nt_scaling(Blk& bk, int i);
  int ensure_mixed();
  int schur_block_fp32(Blk& bk, const SchurScaling& sc);
  int schur_block_bf16x3(Blk& bk, const SchurScaling& sc);
  bool bf16x3() const { return opt.precision == 2; }
  int apply_H(const double* v, double* out);
  int schur_solve(double* v);
  int factorize_schur(bool& ok);
  double dense_assembly_gflop() const;
  int hsolve(double* v);
  bool ldlt() const { return opt.schur == MSDP_SCHUR_LDLT; }
  // MadNLPSchur: the pivots and the step control live in Ht (unused there)
  int* ldlt_ipiv() const { return reinterpret_cast<int*>(Ht); }
  void* ldlt_ctl() const {
    return reinterpret_cast<char*>(Ht) + ((size_t)m * sizeof(int) + 255) / 256 * 256;
  }
  bool lowp_eligible() const;
  int schur_scaling(Blk& bk, SchurScaling& sc);
  int apply_E(Blk& b, const double* M, bool diag, double* out);
  int solve_directions(double mu_, bool corrected);
  int correction();
  int step_lengths(double& a, double& bt);
  int affine_complementarity(double a, double bt, double& out);
  int take_step(double a, double bt);
  int total_complementarity(double& out);
  int iterate();
  int update_termination();
  int initialize();
  int initial_point();
  int solve();
  // DualCentering
  int refresh_slacks(bool& ok);
  int newton_direction(bool& ok);
  int dc_initialize();
  int dc_iterate();
  int predictor_step(bool& ok);
  // QRSchur
  int qr_factorize(bool& ok);
  int qr_solve_directions(double mu_, bool corrected);
};

static int as_err(const char* msg) {
  set_error(msg);
  return -2;
}

// ==========================================================================
// set-up
// ==========================================================================
int msdp_solver::setup_aux() {
  if (nb < 2) return 0;
  int nmax = 1;
  for (auto& b : blk) {
    if (b.n > JACOBI_VB_MAX) return 0;
    nmax = std::max(nmax, b.n);
  }
  // room for the paired triangular inverses (two work areas)
  const size_t tw = std::max(2 * trtri_work_elems(nmax), potrf_work_elems(nmax));
  for (int k = 0; k < NAUX; ++k) {
    hipStream_t q;
    MSDP_HIP(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    aux.push_back(q);
    hipEvent_t ev;
    MSDP_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    aux_done.push_back(ev);
    double *a, *b, *c = nullptr;
    MSDP_TRY(alloc(&a, tw));
    MSDP_TRY(alloc(&b, reduce_scratch_elems(0) + 64));
    MSDP_TRY(fill(st, 64, 0.0, b + reduce_scratch_elems(0)));  // the reduction counter
    if (swork_elems) MSDP_TRY(alloc(&c, swork_elems));
    aux_trwork.push_back(a);
    aux_red.push_back(b);
    aux_swork.push_back(c);
  }
  MSDP_HIP(hipEventCreateWithFlags(&forked, hipEventDisableTiming));
  multi = true;
  return 0;
}

int msdp_solver::finalize() {
  if (released) return as_err("solver released: its device state was freed after the solve");
  if (finalized) return 0;
  MSDP_HIP(hipSetDevice(device));
  MSDP_TRY(ensure_stream());
  int nmax = 1;
  for (auto& b : blk) nmax = std::max(nmax, b.n);
  const size_t nm2 = (size_t)nmax * nmax;
  for (int i = 0; i < nb; ++i) {
    Blk& b = blk[i];
    const size_t n2 = (size_t)b.n * b.n;
    double** mats[] = {&b.X, &b.S, &b.Rd, &b.Ls, &b.Si, &b.Lsinv, &b.delX, &b.delS, &b.Lx,
                       &b.Lxinv, &b.corr, &b.buf, &b.buf2, &b.tmp, &b.tmp2, &b.W, &b.G, &b.Gi,
                       &b.RNT, &b.Mm, &b.Nm, &b.Vv, &b.Shat, &b.Dc, &b.Lh};
    for (double** p : mats) {
      MSDP_TRY(alloc(p, n2));
      MSDP_HIP(hipMemsetAsync(*p, 0, sizeof(double) * n2, st));
    }
    MSDP_TRY(alloc(&b.D, b.n));
    MSDP_TRY(alloc(&b.lam, b.n));
    MSDP_TRY(alloc(&b.C, n2));
    if (b.h_C.size() == n2)
      MSDP_HIP(hipMemcpyAsync(b.C, b.h_C.data(), sizeof(double) * n2, hipMemcpyHostToDevice, st));
    else
      MSDP_HIP(hipMemsetAsync(b.C, 0, sizeof(double) * n2, st));
    // constraints
    BlockOps& o = b.ops;
    o.n = b.n;
    if (b.has_dense) {
      const int mi = (int)b.h_dcon.size();
      o.kind = mi > 0 ? BlockOps::DENSE : BlockOps::EMPTY;
      o.ncon = mi;
      MSDP_TRY(alloc(&o.con_list, mi));
      MSDP_TRY(alloc(&o.mats, (size_t)mi * n2));
      MSDP_HIP(hipMemcpyAsync(o.con_list, b.h_dcon.data(), sizeof(int) * mi, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.mats, b.h_mats.data(), sizeof(double) * mi * n2, hipMemcpyHostToDevice, st));
      bool ident = mi == m;
      for (int j = 0; ident && j < mi; ++j) ident = b.h_dcon[j] == j;
      o.dcon_identity = ident;
      const size_t drow = n2;  // vec (HKM) is the larger of vec / svec
      MSDP_TRY(alloc(&o.Bbuf, drow * mi));
      if (!ident) MSDP_TRY(alloc(&o.Hbuf, (size_t)mi * mi));
      o.jt_split = jtprod_split(b.n, mi);
      if (o.jt_split > 1) MSDP_TRY(alloc(&o.jt_part, (size_t)o.jt_split * n2));
    } else if (b.has_lowrank && !b.h_lcon.empty()) {
      const int mi = (int)b.h_lcon.size();
      const int R = b.h_lptr[mi];
      std::vector<int> fcon(R);
      for (int j = 0; j < mi; ++j)
        for (int r = b.h_lptr[j]; r < b.h_lptr[j + 1]; ++r) fcon[r] = b.h_lcon[j];
      o.kind = BlockOps::LOWRANK;
      o.ncon = mi;
      o.R = R;
      MSDP_TRY(alloc(&o.con_list, mi));
      MSDP_TRY(alloc(&o.con_ptr, mi + 1));
      MSDP_TRY(alloc(&o.fcon, R));
      MSDP_TRY(alloc(&o.F, (size_t)b.n * R));
      MSDP_TRY(alloc(&o.lam, R));
      MSDP_TRY(alloc(&o.FW, (size_t)b.n * R));
      MSDP_TRY(alloc(&o.PL, (size_t)R * R));
      MSDP_TRY(alloc(&o.PR, (size_t)R * R));
      MSDP_HIP(hipMemcpyAsync(o.con_list, b.h_lcon.data(), sizeof(int) * mi, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.con_ptr, b.h_lptr.data(), sizeof(int) * (mi + 1), hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.fcon, fcon.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.F, b.h_lF.data(), sizeof(double) * b.n * R, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.lam, b.h_lw.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
      MSDP_TRY(stream_sync(st));  // fcon is a local
    } else if (b.has_coo && !b.h_con.empty()) {
      // sort entries by constraint (stable), then build the ranges
      const int64_t nnz = (int64_t)b.h_con.size();
      std::vector<int64_t> perm(nnz);
      for (int64_t e = 0; e < nnz; ++e) perm[e] = e;
      std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t c) { return b.h_con[a] < b.h_con[c]; });
      std::vector<int> con(nnz), row(nnz), col(nnz);
      std::vector<double> val(nnz);
      for (int64_t e = 0; e < nnz; ++e) {
        con[e] = b.h_con[perm[e]]; row[e] = b.h_row[perm[e]]; col[e] = b.h_col[perm[e]]; val[e] = b.h_val[perm[e]];
      }
      std::vector<int> clist, cptr;
      for (int64_t e = 0; e < nnz; ++e) {
        if (clist.empty() || clist.back() != con[e]) { clist.push_back(con[e]); cptr.push_back((int)e); }
      }
      cptr.push_back((int)nnz);
      bool onehot = (int64_t)clist.size() == nnz;
      for (int64_t e = 0; onehot && e < nnz; ++e) onehot = row[e] == col[e];
      // positions for A*(y)
      std::vector<int64_t> lin(nnz), pp(nnz);
      for (int64_t e = 0; e < nnz; ++e) { lin[e] = (int64_t)row[e] + (int64_t)col[e] * b.n; pp[e] = e; }
      std::stable_sort(pp.begin(), pp.end(), [&](int64_t a, int64_t c) { return lin[a] < lin[c]; });
      std::vector<int64_t> plin;
      std::vector<int> pptr, pent(nnz);
      for (int64_t k = 0; k < nnz; ++k) {
        const int64_t e = pp[k];
        pent[k] = (int)e;
        if (plin.empty() || plin.back() != lin[e]) { plin.push_back(lin[e]); pptr.push_back((int)k); }
      }
      pptr.push_back((int)nnz);
      o.kind = BlockOps::COO;
      o.one_hot = onehot;
      o.nnz = nnz;
      o.ncon = onehot ? (int)nnz : (int)clist.size();
      o.npos = (int)plin.size();
      MSDP_TRY(alloc(&o.con, nnz));
      MSDP_TRY(alloc(&o.row, nnz));
      MSDP_TRY(alloc(&o.col, nnz));
      MSDP_TRY(alloc(&o.val, nnz));
      MSDP_TRY(alloc(&o.con_list, clist.size()));
      MSDP_TRY(alloc(&o.con_ptr, cptr.size()));
      MSDP_TRY(alloc(&o.pos_lin, plin.size()));
      MSDP_TRY(alloc(&o.pos_ptr, pptr.size()));
      MSDP_TRY(alloc(&o.pos_ent, pent.size()));
      MSDP_HIP(hipMemcpyAsync(o.con, con.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.row, row.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.col, col.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.val, val.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.con_list, clist.data(), sizeof(int) * clist.size(), hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.con_ptr, cptr.data(), sizeof(int) * cptr.size(), hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.pos_lin, plin.data(), sizeof(int64_t) * plin.size(), hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.pos_ptr, pptr.data(), sizeof(int) * pptr.size(), hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(o.pos_ent, pent.data(), sizeof(int) * pent.size(), hipMemcpyHostToDevice, st));
      MSDP_TRY(stream_sync(st));  // the host arrays above are locals
    } else {
      o.kind = BlockOps::EMPTY;
    }
  }
  // vectors
  const size_t mm = std::max<int64_t>(m, 1), nss = std::max<int64_t>(ns, 1);
  double** mv[] = {&y, &dely, &Rp, &wm, &dc_d, &dc_g, &dc_dy, &dc_yhat};
  for (double** p : mv) { MSDP_TRY(alloc(p, mm)); MSDP_HIP(hipMemsetAsync(*p, 0, sizeof(double) * mm, st)); }
  MSDP_TRY(alloc(&b, mm));
  if ((int64_t)h_b.size() == m && m > 0)
    MSDP_HIP(hipMemcpyAsync(b, h_b.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
  else
    MSDP_HIP(hipMemsetAsync(b, 0, sizeof(double) * mm, st));
  double** sv[] = {&x, &z, &delx, &delz, &Rdlin, &zinv, &corrlin, &ws1, &ws2};
  for (double** p : sv) { MSDP_TRY(alloc(p, nss)); MSDP_HIP(hipMemsetAsync(*p, 0, sizeof(double) * nss, st)); }
  MSDP_TRY(alloc(&dlin, nss));
  MSDP_TRY(alloc(&Clin, mm * nss));
  MSDP_HIP(hipMemsetAsync(dlin, 0, sizeof(double) * nss, st));
  MSDP_HIP(hipMemsetAsync(Clin, 0, sizeof(double) * mm * nss, st));
  if (ns > 0 && (int64_t)h_dlin.size() == ns)
    MSDP_HIP(hipMemcpyAsync(dlin, h_dlin.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st));
  if (ns > 0 && m > 0 && (int64_t)h_Clin.size() == ns * m)
    MSDP_HIP(hipMemcpyAsync(Clin, h_Clin.data(), sizeof(double) * ns * m, hipMemcpyHostToDevice, st));
  MSDP_TRY(alloc(&Cs, mm * nss));
  MSDP_TRY(alloc(&H, mm * mm));
  MSDP_TRY(alloc(&Hs, mm * mm));
  // Ht: the transposed Cholesky factor, or MadNLPSchur's pivots and control
  MSDP_TRY(alloc(&Ht, std::max<size_t>(mm * mm, (mm * sizeof(int) + 1024) / sizeof(double))));
  MSDP_TRY(alloc(&Hdinv, potrf_work_elems((int)mm)));
  if (m >= 512) {
    MSDP_TRY(alloc(&Hdi, (size_t)(m / 512) * 512 * 512));
    MSDP_TRY(alloc(&Hdit, (size_t)(m / 512) * 512 * 512));
    MSDP_TRY(alloc(&Hdiw, (size_t)(m / 512) * trtri_work_elems(512)));
  }
  MSDP_TRY(alloc(&wm2, mm));
  MSDP_TRY(alloc(&rh, mm));
  MSDP_TRY(alloc(&rr, mm));
  MSDP_TRY(alloc(&rdl, mm));
  MSDP_TRY(alloc(&info, 4 * nb + 8));
  nslots = 32 + 8 * nb;
  MSDP_TRY(alloc(&sc, nslots));
  if (!hsc_pin) MSDP_HIP(hipHostMalloc((void**)&hsc_pin, sizeof(double) * nslots, hipHostMallocDefault));
  MSDP_TRY(alloc(&red, reduce_scratch_elems(0) + 64));
  MSDP_TRY(fill(st, 64, 0.0, red + reduce_scratch_elems(0)));  // the reduction counter
  MSDP_TRY(alloc(&ework, syevd_work_elems(nmax)));
  MSDP_TRY(alloc(&eiwork, syevd_iwork_elems(nmax)));
  MSDP_TRY(alloc(&trwork, std::max(std::max(2 * trtri_work_elems(nmax), potrf_work_elems(nmax)),
                                    (size_t)m + 64)));
  int64_t tot = 0;
  for (auto& bb : blk) tot += bb.n;
  MSDP_TRY(alloc(&eigall, std::max<int64_t>(tot, 1)));
  {
    std::vector<EigBatch> nt_items, step_items, dc_items;
    int64_t off = 0;
    for (int i = 0; i < nb; ++i) {
      Blk& bk = blk[i];
      if (bk.n <= JACOBI_VB_MAX) {
        nt_items.push_back(EigBatch{bk.n, bk.Nm, bk.n, bk.lam, nullptr, bk.Vv, bk.n, bk.Mm});
        step_items.push_back(EigBatch{bk.n, bk.tmp2, bk.n, nullptr, sc + bslot(i, 3), nullptr, 0, nullptr});
        step_items.push_back(EigBatch{bk.n, bk.buf2, bk.n, nullptr, sc + bslot(i, 4), nullptr, 0, nullptr});
        dc_items.push_back(EigBatch{bk.n, bk.tmp2, bk.n, eigall + off, nullptr, nullptr, 0, nullptr});
      }
      off += bk.n;
    }
    nsmall = (int)nt_items.size();
    if (nsmall > 0) {
      MSDP_TRY(alloc(&eb_nt, nsmall));
      MSDP_TRY(alloc(&eb_step, 2 * nsmall));
      MSDP_TRY(alloc(&eb_dc, nsmall));
      MSDP_HIP(hipMemcpyAsync(eb_nt, nt_items.data(), sizeof(EigBatch) * nsmall, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(eb_step, step_items.data(), sizeof(EigBatch) * 2 * nsmall, hipMemcpyHostToDevice, st));
      MSDP_HIP(hipMemcpyAsync(eb_dc, dc_items.data(), sizeof(EigBatch) * nsmall, hipMemcpyHostToDevice, st));
      MSDP_TRY(stream_sync(st));
    }
  }
  // chunk buffers of the dense Schur assembly: as many constraints per chunk
  // as fit in 512 MiB (fewer, larger batched products: at n = 250, 512
  // constraints per chunk fill the chip with 2048 tiles), 64 MiB where the
  // auxiliary streams keep a copy each
  size_t want = 0;
  for (auto& bb : blk)
    if (bb.ops.kind == BlockOps::DENSE)
      want = std::max(want, 2 * (size_t)bb.n * bb.n * (size_t)bb.ops.ncon);
  want = std::min(want, swork_cap_bytes() / 8);
  for (auto& bb : blk)
    if (bb.ops.kind == BlockOps::DENSE) want = std::max(want, 2 * (size_t)bb.n * bb.n);
  swork_elems = want;
  if (want) MSDP_TRY(alloc(&swork, want));
  MSDP_TRY(setup_aux());
  MSDP_TRY(build_dense_batch());
  // norms of the data (src/solver.jl:119-120, 217-221, 322, 330)
  for (int i = 0; i < nb; ++i) MSDP_TRY(max_abs(st, (int64_t)blk[i].n * blk[i].n, blk[i].C, slot(bslot(i, 0)), red));
  if (m > 0) MSDP_TRY(max_abs(st, m, b, slot(0), red));
  else MSDP_TRY(fill(st, 1, 0.0, slot(0)));
  if (ns > 0) MSDP_TRY(max_abs(st, ns, dlin, slot(1), red));
  else MSDP_TRY(fill(st, 1, 0.0, slot(1)));
  MSDP_TRY(read_slots());
  bnorm = hsc[0];
  dnorm = hsc[1];
  cnorm_max = 0;
  for (int i = 0; i < nb; ++i) {
    blk[i].cnorm = hsc[bslot(i, 0)];
    cnorm_max = std::max(cnorm_max, blk[i].cnorm);
  }
  // read_slots synchronised the stream: the host staging can go
  for (auto& bb : blk) {
    std::vector<double>().swap(bb.h_mats);
    std::vector<double>().swap(bb.h_C);
    std::vector<double>().swap(bb.h_val);
    std::vector<int>().swap(bb.h_con);
    std::vector<int>().swap(bb.h_row);
    std::vector<int>().swap(bb.h_col);
    std::vector<int>().swap(bb.h_lcon);
    std::vector<int>().swap(bb.h_lptr);
    std::vector<double>().swap(bb.h_lF);
    std::vector<double>().swap(bb.h_lw);
  }
  std::vector<double>().swap(h_b);
  std::vector<double>().swap(h_dlin);
  std::vector<double>().swap(h_Clin);
  std::vector<DenseBlkDev>().swap(dbh_blks);
  std::vector<int>().swap(dbh_cptr);
  std::vector<int>().swap(dbh_cpairs);
  finalized = true;
  return 0;
}

// Device bytes held by a finalized solver (estimate from the sizes alone, so
// a server can admit a request before anything is allocated).
int64_t msdp_solver::footprint() const {
  const int64_t d = sizeof(double);
  const int64_t mm = std::max<int64_t>(m, 1), nss = std::max<int64_t>(ns, 1);
  int64_t bytes = 0, nmax = 1, tot = 0;
  int64_t swork_b = 0;
  for (const auto& b : blk) {
    const int64_t n2 = (int64_t)b.n * b.n;
    nmax = std::max<int64_t>(nmax, b.n);
    tot += b.n;
    bytes += 26 * n2 * d + 2 * b.n * d;
    if (b.has_dense) {
      const int64_t mi = (int64_t)b.h_dcon.size();
      bytes += 2 * mi * n2 * d + mi * 4 + (mi == m ? 0 : mi * mi * d);
      bytes += (int64_t)jtprod_split(b.n, (int)mi) * n2 * d;
      if (opt.precision == 1) bytes += 2 * mi * n2 * 4 + 2 * n2 * 4;
      if (opt.precision == 2) {
        const int64_t drow = dense_panel_rows(b.n, opt.direction != MSDP_DIR_NT);
        bytes += 2 * bf16x3_ld(drow) * mi * 2 + (int64_t)panel_apply_work_elems((int)drow, (int)mi) * d +
                 2 * mi * d;
      }
      swork_b = std::max(swork_b, std::max(2 * n2 * d, std::min<int64_t>(2 * n2 * mi * d,
                                                                       (int64_t)swork_cap_bytes())));
    } else if (b.has_coo) {
      bytes += (int64_t)b.h_con.size() * (4 * 3 + 8 + 8 + 4 + 4) + 64;
    } else if (b.has_lowrank && !b.h_lcon.empty()) {
      const int64_t mi = (int64_t)b.h_lcon.size(), R = b.h_lptr[mi];
      bytes += 2 * b.n * R * d + 2 * R * R * d + R * (d + 4) + (2 * mi + 1) * 4 + 64;
    }
  }
  bytes += swork_b;
  {  // the batched dense A(.) / A*(.) tables (build_dense_batch)
    bool all_dense = nb >= 2 && m > 0;
    int64_t pairs = 0;
    for (const auto& b : blk) {
      all_dense = all_dense && b.has_dense && !b.h_dcon.empty();
      pairs += (int64_t)b.h_dcon.size();
    }
    if (all_dense) {
      bytes += nb * (int64_t)sizeof(DenseBlkDev) + (mm + 1) * 4 + pairs * 12 + 64;
      int64_t asm_b = 0;
      bool small = true;
      for (const auto& b : blk) {
        asm_b += 2 * (int64_t)b.h_dcon.size() * b.n * b.n * d;
        small = small && b.n <= JACOBI_VB_MAX;
      }
      if (small && asm_b <= (int64_t)ASM_BATCH_MAX_BYTES) bytes += asm_b;
    }
  }
  bytes += (3 * mm * mm + 2 * mm * nss + 12 * mm + 9 * nss) * d;
  bytes += (int64_t)potrf_work_elems((int)mm) * d;
  bytes += (mm / 512) * (2 * 512 * 512 + (int64_t)trtri_work_elems(512)) * d;  // Hdi, Hdit, Hdiw
  bytes += ((int64_t)syevd_work_elems((int)nmax) +
            std::max<int64_t>(std::max<int64_t>(2 * (int64_t)trtri_work_elems((int)nmax),
                                                (int64_t)potrf_work_elems((int)nmax)),
                              mm + 64) + 3 * tot) * d;
  // the auxiliary streams' scratch of multi-block problems (setup_aux)
  if (nb >= 2 && nmax <= JACOBI_VB_MAX) {
    const int64_t tw = (int64_t)std::max(2 * trtri_work_elems((int)nmax), potrf_work_elems((int)nmax));
    bytes += NAUX * (tw + (int64_t)reduce_scratch_elems(0) + 64 + swork_b / d) * d;
  }
  // GEMM partials from the stream-ordered pool: split-K (at most 4M doubles)
  // and, when a product's 128-tiles outnumber the CUs, stream-K
  bytes += (int64_t)32 << 20;
  if (std::max<int64_t>(m, nmax) > 2800) bytes += (int64_t)gemm_sk_partials_bytes();
  if (opt.schur == MSDP_SCHUR_QR) {
    // the dense scaled constraint matrix, its T factors and workspace
    int64_t rows = ns;
    for (const auto& b : blk) rows += (int64_t)b.n * (b.n + 1) / 2;
    rows = std::max<int64_t>(rows, 1);
    bytes += (rows * mm + 3 * rows + 5 * mm + 2 + (int64_t)qr_t_elems((int)mm) +
              (int64_t)qr_work_elems(rows, (int)mm)) * d + 4;
  }
  // the recorder's device argument buffer (grown on demand, 1 MiB at least)
  return bytes + ((int64_t)4 << 20);
}

// The batched dense A(.) / A*(.) tables (DenseBatch): every block dense, two
// or more of them.  Pairs (block, constraint) in block order; each
// constraint's pairs listed in block order.
int msdp_solver::build_dense_batch() {
  dbatch = DenseBatch{};
  if (nb < 2 || m <= 0) return 0;
  for (auto& b : blk)
    if (b.ops.kind != BlockOps::DENSE) return 0;
  dbh_blks.assign(nb, DenseBlkDev{});
  std::vector<std::vector<int>> lists((size_t)m);
  int P = 0, T = 0;
  for (int i = 0; i < nb; ++i) {
    Blk& b = blk[i];
    DenseBlkDev& x = dbh_blks[i];
    x.mats = b.ops.mats;
    x.n2 = (int64_t)b.n * b.n;
    x.con_list = b.ops.con_list;
    x.ncon = b.ops.ncon;
    x.pair0 = P;
    x.tile0 = T;
    x.in[0] = b.X;
    x.in[1] = b.buf;
    x.in[2] = b.Si;
    x.out[0] = b.tmp2;
    x.out[1] = b.Dc;
    for (int j = 0; j < x.ncon; ++j) lists[b.h_dcon[j]].push_back(P + j);
    P += x.ncon;
    T += cdiv(x.n2, 256);
  }
  dbh_cptr.assign((size_t)m + 1, 0);
  dbh_cpairs.clear();
  dbh_cpairs.reserve(P);
  for (int64_t c = 0; c < m; ++c) {
    dbh_cpairs.insert(dbh_cpairs.end(), lists[c].begin(), lists[c].end());
    dbh_cptr[c + 1] = (int)dbh_cpairs.size();
  }
  DenseBatch d;
  d.nblk = nb;
  d.npairs = P;
  d.ntiles = T;
  MSDP_TRY(alloc(&d.blks, nb));
  MSDP_TRY(alloc(&d.cptr, (size_t)m + 1));
  MSDP_TRY(alloc(&d.cpairs, (size_t)P));
  MSDP_TRY(alloc(&d.part, (size_t)P));
  MSDP_HIP(hipMemcpyAsync(d.blks, dbh_blks.data(), sizeof(DenseBlkDev) * nb, hipMemcpyHostToDevice, st));
  MSDP_HIP(hipMemcpyAsync(d.cptr, dbh_cptr.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, st));
  if (P > 0)
    MSDP_HIP(hipMemcpyAsync(d.cpairs, dbh_cpairs.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
  dbatch = d;
  // the recorded Schur assembly's per-block products, when the blocks are
  // recorded (multi) and they fit a modest budget
  if (multi) {
    size_t tot = 0;
    for (auto& b : blk) tot += 2 * (size_t)b.ops.ncon * b.n * b.n;
    if (tot * sizeof(double) <= ASM_BATCH_MAX_BYTES)
      for (auto& b : blk) {
        const size_t e = (size_t)b.ops.ncon * b.n * b.n;
        MSDP_TRY(alloc(&b.asmT, e));
        MSDP_TRY(alloc(&b.asmP, e));
      }
  }
  return 0;
}

// Copies the solution (X, S, y, x, z) to the host, frees the device state and
// drops a borrowed stream (its owner may destroy it next).
int msdp_solver::release() {
  if (released) return 0;
  int rc = 0;
  if (finalized) rc = keep_solution();
  free_device();
  destroy_aux();
  free_rec();
  released = finalized;
  if (!own_stream) st = nullptr;
  return rc;
}

int msdp_solver::keep_solution() {
  MSDP_HIP(hipSetDevice(device));
  r_y.resize(m);
  r_x.resize(ns);
  r_z.resize(ns);
  if (m > 0) MSDP_HIP(hipMemcpyAsync(r_y.data(), y, sizeof(double) * m, hipMemcpyDeviceToHost, st));
  if (ns > 0) {
    MSDP_HIP(hipMemcpyAsync(r_x.data(), x, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
    MSDP_HIP(hipMemcpyAsync(r_z.data(), z, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
  }
  r_X.resize(nb);
  r_S.resize(nb);
  for (int i = 0; i < nb; ++i) {
    const size_t n2 = (size_t)blk[i].n * blk[i].n;
    r_X[i].resize(n2);
    r_S[i].resize(n2);
    MSDP_HIP(hipMemcpyAsync(r_X[i].data(), blk[i].X, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
    MSDP_HIP(hipMemcpyAsync(r_S[i].data(), blk[i].S, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
  }
  MSDP_TRY(stream_sync(st));
  return 0;
}

// ==========================================================================
// block kernels (src/blocks.jl)
// ==========================================================================
// _cholesky_factor! (blocks.jl:94-113)
int msdp_solver::cholesky_factor(Blk& bk, const double* M, double* L, int* inf) {
  const int n = bk.n;
  if (n <= 64) return cholesky_small(st, n, M, n, L, n, inf);
  if (n <= 128) return cholesky_mid(st, n, M, n, L, n, inf);
  return opaque([=]() -> int {
    MSDP_TRY(copy_mat(st, n, n, M, n, L, n));
    MSDP_TRY(symmetrize(st, n, L, n));
    MSDP_TRY(potrf(st, n, L, n, inf, trwork));
    MSDP_TRY(chol_flush(st, n, L, n, red));
    return 0;
  });
}

// two _cholesky_factor! of one block (n > 64) in the same launches; the
// statuses into inf[0] and inf[1]
int msdp_solver::cholesky_factor2(Blk& bk, const double* M0, double* L0, const double* M1,
                                  double* L1, int* inf) {
  const int n = bk.n;
  if (n > 64 && n <= 128) {
    MSDP_TRY(cholesky_mid(st, n, M0, n, L0, n, inf));
    return cholesky_mid(st, n, M1, n, L1, n, inf + 1);
  }
  return opaque([=]() -> int {
    MSDP_TRY(copy_mat(st, n, n, M0, n, L0, n));
    MSDP_TRY(symmetrize(st, n, L0, n));
    MSDP_TRY(copy_mat(st, n, n, M1, n, L1, n));
    MSDP_TRY(symmetrize(st, n, L1, n));
    MSDP_TRY(potrf2(st, n, L0, L1, n, inf, trwork));
    MSDP_TRY(chol_flush(st, n, L0, n, red));
    MSDP_TRY(chol_flush(st, n, L1, n, red));
    return 0;
  });
}

// prepare_block! (blocks.jl:126-130): Ls = chol(S), Si = S^-1 (potri);
// `factored`: Ls is already there
int msdp_solver::prepare_block(Blk& bk, int i, bool factored) {
  const int n = bk.n;
  if (!factored) MSDP_TRY(cholesky_factor(bk, bk.S, bk.Ls, binfo(i, 0)));
  Blk* bp = &bk;
  if (n <= 128) MSDP_TRY(trtri_lower(st, n, bk.Ls, n, bk.Lsinv, n, trwork));  // recorded
  else MSDP_TRY(opaque([=] { return trtri_lower(st, n, bp->Ls, n, bp->Lsinv, n, trwork); }));
  return inverse_product(bk);
}

// S^-1 = L_S^-T L_S^-1 (blocks.jl:128-130): the lower triangle, each tile
// over the K range where both triangular factors are nonzero, then mirrored
int msdp_solver::inverse_product(Blk& bk) {
  const int n = bk.n;
  MSDP_TRY(gemm(st, true, false, n, n, n, 1.0, bk.Lsinv, n, bk.Lsinv, n, 0.0, bk.Si, n, 1, 0, 0,
                0, true, TRI_AU | TRI_BL));
  MSDP_TRY(mirror_lower(st, n, bk.Si, n));
  return 0;
}

// out = C_i - A*_i(y), symmetrised (solver.jl:193-198)
int msdp_solver::slack(Blk& bk, const double* yv, double* out) {
  const int64_t n2 = (int64_t)bk.n * bk.n;
  MSDP_TRY(jtprod(st, bk.ops, yv, bk.tmp2));
  MSDP_TRY(axpby(st, n2, 1.0, bk.C, -1.0, bk.tmp2, out));
  MSDP_TRY(symmetrize(st, bk.n, out, bk.n));
  return 0;
}

// _set_dual_slacks! (solver.jl:201-211)
int msdp_solver::set_dual_slacks(const double* yv) {
  for (auto& bk : blk) MSDP_TRY(slack(bk, yv, bk.S));
  if (ns > 0) {
    MSDP_TRY(gemm(st, true, false, (int)ns, 1, (int)m, 1.0, Clin, m, yv, m, 0.0, ws1, ns));
    MSDP_TRY(axpby(st, ns, 1.0, dlin, -1.0, ws1, z));
  }
  return 0;
}

// ==========================================================================
// PrimalDual components
// Executes the recorded region: op position by op position, every batchable
// kind of op present at the position as one batched launch on the main
// stream, the other ops per block on the auxiliary streams (fork ... join).
// All batched arguments are packed into one pinned slot and copied once.
int msdp_solver::rec_flush() {
  size_t L = 0;
  for (auto& v : rec.ops) L = std::max(L, v.size());
  struct Group {
    int kind;
    std::vector<const RecOp*> items;
    size_t off = 0;
    int grid = 0;
  };
  std::vector<std::vector<Group>> groups(L);
  std::vector<std::vector<int>> opq(L);  // blocks whose op at the position is opaque
  size_t bytes = 0;
  for (size_t k = 0; k < L; ++k) {
    std::vector<const RecOp*> bykind[REC_NKINDS];
    for (int i = 0; i < nb; ++i) {
      if (k >= rec.ops[i].size()) continue;
      const RecOp& o = rec.ops[i][k];
      if (o.kind == REC_OPAQUE) opq[k].push_back(i);
      else bykind[o.kind].push_back(&o);
    }
    for (int t = 1; t < REC_NKINDS; ++t) {
      if (bykind[t].empty()) continue;
      Group g;
      g.kind = t;
      g.items = std::move(bykind[t]);
      g.off = bytes;
      bytes += rec_bytes(t, g.items.size());
      groups[k].push_back(std::move(g));
    }
  }
  if (bytes > 0) {
    if (bytes > rec_cap) {  // grow the ring and the device buffer
      MSDP_HIP(hipStreamSynchronize(st));
      for (int r = 0; r < REC_RING; ++r) {
        if (rec_ev[r]) MSDP_HIP(hipEventSynchronize(rec_ev[r]));
        if (rec_host[r]) MSDP_HIP(hipHostFree(rec_host[r]));
        rec_host[r] = nullptr;
      }
      if (rec_dev) MSDP_HIP(hipFree(rec_dev));
      rec_dev = nullptr;
      rec_cap = std::max<size_t>(bytes * 2, 1 << 20);
      for (int r = 0; r < REC_RING; ++r) {
        MSDP_HIP(hipHostMalloc((void**)&rec_host[r], rec_cap, hipHostMallocDefault));
        if (!rec_ev[r]) MSDP_HIP(hipEventCreateWithFlags(&rec_ev[r], hipEventDisableTiming));
      }
      MSDP_HIP(hipMalloc((void**)&rec_dev, rec_cap));
    }
    rec_slot = (rec_slot + 1) % REC_RING;
    char* host = rec_host[rec_slot];
    MSDP_HIP(hipEventSynchronize(rec_ev[rec_slot]));  // its previous copy has completed
    for (auto& gk : groups)
      for (auto& g : gk) rec_pack(g.kind, g.items, host + g.off, &g.grid);
    MSDP_HIP(hipMemcpyAsync(rec_dev, host, bytes, hipMemcpyHostToDevice, st));
    MSDP_HIP(hipEventRecord(rec_ev[rec_slot], st));
  }
  for (size_t k = 0; k < L; ++k) {
    for (auto& g : groups[k])
      MSDP_TRY(rec_launch(st, g.kind, (int)g.items.size(), g.grid, rec_dev + g.off));
    if (!opq[k].empty()) {
      MSDP_TRY(fork());
      for (int i : opq[k]) {
        OnBlock on(this, i);
        MSDP_TRY(rec.ops[i][k].run());
      }
      MSDP_TRY(join());
    }
  }
  for (auto& v : rec.ops) v.clear();
  return 0;
}

// ==========================================================================
// update_residuals! (step.jl:147-182)
int msdp_solver::update_residuals() {
  Range range_("madsdp:update_residuals");
  if (m > 0) {
    MSDP_TRY(fill(st, m, 0.0, Rp));
    MSDP_TRY(jprod_blocks(JP_X, Rp));
    if (ns > 0) MSDP_TRY(gemm(st, false, false, (int)m, 1, (int)ns, 1.0, Clin, m, x, ns, 1.0, Rp, m));
    MSDP_TRY(axpby(st, m, 1.0, b, -1.0, Rp, Rp));
  }
  const bool jt_batched = dbatch.nblk > 0;
  if (jt_batched) MSDP_TRY(jtprod_dense_batched(st, dbatch, y, JT_TMP2));
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    Blk& bk = blk[i];
    const int64_t n2 = (int64_t)bk.n * bk.n;
    Blk* bp = &bk;
    if (!jt_batched) MSDP_TRY(opaque([=] { return jtprod(st, bp->ops, y, bp->tmp2); }));
    MSDP_TRY(axpbypcz(st, n2, 1.0, bk.C, -1.0, bk.S, -1.0, bk.tmp2, bk.Rd));
    MSDP_TRY(max_abs(st, n2, bk.Rd, slot(bslot(i, 1)), red));
  }
  MSDP_TRY(rec_end());
  if (ns > 0) {
    MSDP_TRY(gemm(st, true, false, (int)ns, 1, (int)m, 1.0, Clin, m, y, m, 0.0, ws1, ns));
    MSDP_TRY(axpbypcz(st, ns, 1.0, dlin, -1.0, z, -1.0, ws1, Rdlin));
  }
  MSDP_TRY(read_slots());
  for (int i = 0; i < nb; ++i) {
    const double thr = 100 * EPS * std::max(1.0, blk[i].cnorm);
    blk[i].rd_nonzero = hsc[bslot(i, 1)] > thr;
  }
  return 0;
}

// _prepare! (primal_dual.jl:29-37): prepare_block!, prepare_direction!,
// prepare_step! for every block, then zinv.
int msdp_solver::prepare(bool& ok) {
  Range range_("madsdp:prepare");
  MSDP_TRY(fill(st, 1, 0.0, slot(2)));
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * (4 * nb + 8), st));
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    Blk& bk = blk[i];
    Blk* bp = &bk;
    const int n = bk.n;
    const bool need_lx = nt() || opt.step == MSDP_STEP_EIGEN;
    // chol(S) and chol(X) of a large block share their launches
    const bool pair = need_lx && n > 64;
    if (pair) {
      MSDP_TRY(cholesky_factor2(bk, bk.S, bk.Ls, bk.X, bk.Lx, binfo(i, 0)));
      if (n <= 128) {
        // recorded: batched with the other blocks' inverses at this position
        MSDP_TRY(trtri_lower(st, n, bk.Ls, n, bk.Lsinv, n, trwork));
        MSDP_TRY(trtri_lower(st, n, bk.Lx, n, bk.Lxinv, n, trwork));
      } else {
        // L_S^-1 and L_X^-1 in the same launches (a batch of two at the
        // matrices' distance, like the paired factorisation)
        MSDP_TRY(opaque([=] {
          return trtri_lower_batched(st, n, 2, bp->Ls, n, bp->Lx - bp->Ls, bp->Lsinv, n,
                                     bp->Lxinv - bp->Lsinv, trwork, (int64_t)trtri_work_elems(n));
        }));
      }
      MSDP_TRY(inverse_product(bk));
    } else {
      MSDP_TRY(prepare_block(bk, i, false));
      if (need_lx) {
        MSDP_TRY(cholesky_factor(bk, bk.X, bk.Lx, binfo(i, 1)));
        if (n <= 128) MSDP_TRY(trtri_lower(st, n, bk.Lx, n, bk.Lxinv, n, trwork));  // recorded
        else MSDP_TRY(opaque([=] { return trtri_lower(st, n, bp->Lx, n, bp->Lxinv, n, trwork); }));
      }
    }
    if (nt()) {  // prepare_direction!(::NTWorkspace) directions.jl:45-75
      // M = Ls' Lx (upper times lower), N = M'M (its lower triangle, mirrored)
      MSDP_TRY(gemm(st, true, false, n, n, n, 1.0, bk.Ls, n, bk.Lx, n, 0.0, bk.Mm, n, 1, 0, 0, 0,
                    false, TRI_AU | TRI_BL));
      MSDP_TRY(gemm(st, true, false, n, n, n, 1.0, bk.Mm, n, bk.Mm, n, 0.0, bk.Nm, n, 1, 0, 0, 0,
                    true));
      MSDP_TRY(mirror_lower(st, n, bk.Nm, n));
      if (n > JACOBI_VB_MAX) {
        MSDP_TRY(opaque([=]() -> int {
          MSDP_TRY(syevd(st, n, bp->Nm, n, bp->lam, bp->Vv, n, true, ework, eiwork));
          return nt_scaling(*bp, i);
        }));
      }
    }
  }
  MSDP_TRY(rec_end());
  if (nt() && nsmall > 0) {
    // eigen-decompositions of all small blocks in one launch
    MSDP_TRY(syev_vbatched(st, eb_nt, nsmall));
    MSDP_TRY(rec_begin());
    for (int i = 0; i < nb; ++i) {
      OnBlock on(this, i);
      if (blk[i].n <= JACOBI_VB_MAX) MSDP_TRY(nt_scaling(blk[i], i));
    }
    MSDP_TRY(rec_end());
  }
  if (ns > 0) MSDP_TRY(scalar_op(st, SOP_ZINV, (int)ns, 0, nullptr, z, nullptr, nullptr, nullptr, zinv, nullptr, false));
  std::vector<int> inf;
  MSDP_TRY(read_infos_slots(inf));
  ok = true;
  for (int i = 0; i < nb && ok; ++i) {
    if (inf[8 + 4 * i] != 0 || inf[8 + 4 * i + 1] != 0) ok = false;
    if (ok && nt() && !(hsc[bslot(i, 2)] > 0.0)) ok = false;  // directions.jl:55-58
  }
  return 0;
}

// D = sqrt(lambda), G = Lx V D^-1/2, Gi = D^1/2 V' Lx^-1, W = G G'
// (directions.jl:55-73) from the eigen-decomposition of N in (lam, Vv).
int msdp_solver::nt_scaling(Blk& bk, int i) {
  const int n = bk.n;
  MSDP_TRY(sqrt_min(st, n, bk.lam, bk.D, slot(bslot(i, 2))));
  MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.Lx, n, bk.Vv, n, 0.0, bk.G, n, 1, 0, 0, 0, false,
                TRI_AL));
  MSDP_TRY(scale_cols(st, n, n, bk.G, n, bk.D, -1));
  MSDP_TRY(transpose(st, n, n, bk.Vv, n, bk.tmp, n));
  MSDP_TRY(scale_rows(st, n, n, bk.tmp, n, bk.D, +1));
  MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.tmp, n, bk.Lxinv, n, 0.0, bk.Gi, n, 1, 0, 0, 0,
                false, TRI_BL));
  MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.G, n, bk.G, n, 0.0, bk.W, n, 1, 0, 0, 0, true));
  MSDP_TRY(mirror_lower(st, n, bk.W, n));
  return 0;
}

// (L, R) of the Schur complement and the factored dense form for one block
// (src/directions.jl:77,112; src/dual_centering.jl:66)
int msdp_solver::schur_scaling(Blk& bk, SchurScaling& sc) {
  if (dual_centering()) {
    sc.L = bk.Si; sc.R = bk.Si;
    sc.Lm = bk.Lsinv; sc.transL = false; sc.F2 = bk.Lsinv; sc.transF2 = true; sc.full_vec = false;
  } else if (nt()) {
    sc.L = bk.W; sc.R = bk.W;
    sc.Lm = bk.G; sc.transL = true; sc.F2 = bk.G; sc.transF2 = false; sc.full_vec = false;
  } else {
    sc.L = bk.Si; sc.R = bk.X;
    sc.Lm = bk.Lsinv; sc.transL = false; sc.F2 = bk.Lx; sc.transF2 = false; sc.full_vec = true;
    if (bk.ops.kind == BlockOps::DENSE && opt.step != MSDP_STEP_EIGEN) {
      // HKM needs chol(X) for the factored dense form
      const int i = (int)(&bk - blk.data());
      MSDP_TRY(cholesky_factor(bk, bk.X, bk.Lx, binfo(i, 3)));
    }
  }
  return 0;
}

// Flops (in GFLOP) of one Schur assembly of the dense-constraint blocks:
// P_j = L A_j R for the m_i touching matrices (4 m_i n^3) and the product
// H += B'B of their svec (vec for HKM) panels (m_i^2 d).  Low-precision
// assembly pays only where this work outweighs the refinement it brings.
double msdp_solver::dense_assembly_gflop() const {
  double f = 0.0;
  for (const auto& bk : blk) {
    if (bk.ops.kind != BlockOps::DENSE) continue;
    const double n = bk.n, mi = bk.ops.ncon;
    const double d = nt() ? n * (n + 1) / 2 : n * n;
    f += 4.0 * mi * n * n * n + mi * mi * d;
  }
  return f * 1e-9;
}

// Low-precision assembly needs a Schur complement that can be positive
// definite: with more constraints than the sum of svec dimensions and
// scalars, H is singular, the reference's ladder of absolute shifts
// (1e-12 ... 1e-5) keeps it factorable in FP64 but not past an FP32
// rounding error of H.
bool msdp_solver::lowp_eligible() const {
  double dim = (double)ns;
  for (const auto& bk : blk) dim += 0.5 * (double)bk.n * (bk.n + 1);
  return !ldlt() && (double)m <= dim && dense_assembly_gflop() >= opt.lowp_min_gflop;
}

// factorize_schur! (solver.jl:299-314) + factorize!(DenseCholesky)
int msdp_solver::factorize_schur(bool& ok) {
  Range range_("madsdp:factorize_schur");
  ok = false;
  if (opt.schur == MSDP_SCHUR_QR) return qr_factorize(ok);
  if (m == 0) { ok = true; return 0; }
  // mixed precision: FP32 or BF16x3 assembly of the dense blocks while
  // mu >= mu_switch
  const bool lowp = opt.precision != 0 && !dual_centering() && mu >= opt.mu_switch && !force_fp64 &&
                    lowp_eligible();
  bool any_lowp = false;
  if (lowp) MSDP_TRY(ensure_mixed());
  for (auto& bk : blk) bk.panel_fp64 = false;
  MSDP_TRY(fill(st, m * m, 0.0, Hs));
  // Multi-block: the per-block Schur contributions of dense blocks that touch
  // a subset of the constraints are formed concurrently into their own
  // buffers, then scattered into H in block order on the main stream.
  const bool split = multi && !lowp;
  auto deferred = [&](const Blk& bk) {
    return split && bk.ops.kind == BlockOps::DENSE && !bk.ops.dcon_identity && bk.ops.ncon > 0;
  };
  // every block's products as recorded ops: one batched launch per product
  // for all the blocks (NT and dual-centering scalings: svec panels)
  const bool recorded = split && multi && blk[0].asmT && (nt() || dual_centering());
  if (recorded) {
    MSDP_TRY(rec_begin());
    for (int i = 0; i < nb; ++i) {
      Blk& bk = blk[i];
      if (!deferred(bk)) continue;
      OnBlock on(this, i);
      SchurScaling sc;
      MSDP_TRY(schur_scaling(bk, sc));
      MSDP_TRY(dense_schur_record(st, bk.ops, sc, bk.asmT, bk.asmP));
    }
    MSDP_TRY(rec_end());
  } else {
    MSDP_TRY(fork());
    for (int i = 0; i < nb; ++i) {
      Blk& bk = blk[i];
      if (!deferred(bk)) continue;
      OnBlock on(this, i);
      SchurScaling sc;
      MSDP_TRY(schur_scaling(bk, sc));
      MSDP_TRY(schur_block(st, bk.ops, sc, nullptr, m, swork, swork_elems));
    }
    MSDP_TRY(join());
  }
  for (auto& bk : blk) {
    if (deferred(bk)) {
      MSDP_TRY(scatter_lower_add(st, bk.ops.ncon, bk.ops.con_list, bk.ops.Hbuf, Hs, m));
      continue;
    }
    SchurScaling sc;
    MSDP_TRY(schur_scaling(bk, sc));
    if (lowp && bk.ops.kind == BlockOps::DENSE) {
      MSDP_TRY(bf16x3() ? schur_block_bf16x3(bk, sc) : schur_block_fp32(bk, sc));
      any_lowp = true;
      continue;
    }
    MSDP_TRY(schur_block(st, bk.ops, sc, Hs, m, swork, swork_elems));
  }
  if (ns > 0) {
    // H += C_lin diag(x ./ z) C_lin'
    MSDP_TRY(scalar_op(st, SOP_XZ, (int)ns, 0, x, nullptr, zinv, nullptr, nullptr, ws2, nullptr, false));
    MSDP_TRY(copy_mat(st, (int)m, (int)ns, Clin, m, Cs, m));
    MSDP_TRY(scale_cols_vec(st, (int)m, (int)ns, Cs, m, ws2));
    MSDP_TRY(gemm(st, false, true, (int)m, (int)m, (int)ns, 1.0, Cs, m, Clin, m, 1.0, Hs, m, 1, 0, 0, 0, true));
  }
  double reg = opt.regularize_schur;
  for (int t = 0; t < opt.max_schur_reg; ++t) {
    MSDP_TRY(copy_mat(st, (int)m, (int)m, Hs, m, H, m));
    if (reg > 0) MSDP_TRY(add_diag(st, (int)m, H, m, nullptr, reg));
    if (ldlt()) {
      // MadNLPSchur: Bunch-Kaufman LDL' (src/kkt.jl:7-32); it fails only on
      // an exactly zero pivot (src/schur_solvers.jl:27-37)
      MSDP_TRY(sytrf_bk(st, (int)m, H, m, ldlt_ipiv(), ldlt_ctl(), Hdinv, info));
      int hinfo = 0;
      MSDP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st));
      MSDP_TRY(stream_sync(st));
      if (hinfo == 0) {
        reg_used = reg > 0 ? reg : 0.0;
        h_approx = false;
        ok = true;
        return 0;
      }
      reg = reg == 0 ? 1e-12 : 10 * reg;
      continue;
    }
    MSDP_TRY(potrf(st, (int)m, H, m, info, Hdinv, true));
    int hinfo = 0;
    MSDP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st));
    MSDP_TRY(stream_sync(st));
    if (hinfo == 0) {
      MSDP_TRY(transpose(st, (int)m, (int)m, H, m, Ht, m));
      hdi_valid = false;
      reg_used = reg > 0 ? reg : 0.0;
      h_approx = any_lowp;
      if (any_lowp) ++lowp_count;
      ok = true;
      return 0;
    }
    // an FP32-assembled H that is not numerically PD at the reference's
    // first regularisation is off by its assembly error, which the ladder's
    // absolute shifts (<= 1e-5) do not cover: FP64 from here on at once
    if (any_lowp) break;
    reg = reg == 0 ? 1e-12 : 10 * reg;
  }
  if (any_lowp) {
    // the FP32-assembled H is not numerically PD: FP64 from here on
    ++fallback_count;
    force_fp64 = true;
    return factorize_schur(ok);
  }
  return 0;
}

// apply_E! (directions.jl:80-84 NT, 116-120 HKM); `diag` marks a diagonal M
// (one-hot A*(y)) for which the first product is a column scaling.
int msdp_solver::apply_E(Blk& bk, const double* M, bool diag, double* out) {
  const int n = bk.n;
  const double* P = nt() ? bk.W : bk.X;
  const double* Q = nt() ? bk.W : bk.Si;
  if (diag) MSDP_TRY(scale_by_diag(st, n, P, M, bk.tmp));
  else MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, P, n, M, n, 0.0, bk.tmp, n));
  MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.tmp, n, Q, n, 0.0, out, n));
  MSDP_TRY(symmetrize(st, n, out, n));
  return 0;
}

// solve_directions! (primal_dual.jl:44-54) = rhs! (step.jl:198-224),
// schur_solve! (schur_solvers.jl:21), directions! (step.jl:229-259)
int msdp_solver::solve_directions(double mu_, bool corrected) {
  Range range_("madsdp:solve_directions");
  if (opt.schur == MSDP_SCHUR_QR) return qr_solve_directions(mu_, corrected);
  // rhs
  if (m > 0) {
    MSDP_HIP(hipMemcpyAsync(dely, Rp, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    MSDP_TRY(rec_begin());
    for (int i = 0; i < nb; ++i) {
      OnBlock on(this, i);
      Blk& bk = blk[i];
      const int64_t n2 = (int64_t)bk.n * bk.n;
      MSDP_TRY(axpby(st, n2, 1.0, bk.X, -mu_, bk.Si, bk.buf));
      if (bk.rd_nonzero) {
        MSDP_TRY(apply_E(bk, bk.Rd, false, bk.buf2));
        MSDP_TRY(axpby(st, n2, 1.0, bk.buf, 1.0, bk.buf2, bk.buf));
      }
      if (corrected) MSDP_TRY(axpby(st, n2, 1.0, bk.buf, -1.0, bk.corr, bk.buf));
    }
    MSDP_TRY(rec_end());
    // A(.) accumulates every block into dely: in block order on the main stream
    MSDP_TRY(jprod_blocks(JP_BUF, dely));
    if (ns > 0) {
      MSDP_TRY(scalar_op(st, SOP_RHS, (int)ns, mu_, x, z, zinv, Rdlin, corrlin, ws1, nullptr, corrected));
      MSDP_TRY(gemm(st, false, false, (int)m, 1, (int)ns, 1.0, Clin, m, ws1, ns, 1.0, dely, m));
    }
    MSDP_TRY(schur_solve(dely));
  }
  // directions
  const bool jt_batched = dbatch.nblk > 0;
  if (jt_batched) MSDP_TRY(jtprod_dense_batched(st, dbatch, dely, JT_TMP2));
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    Blk& bk = blk[i];
    const int n = bk.n;
    const int64_t n2 = (int64_t)n * n;
    Blk* bp = &bk;
    if (!jt_batched) MSDP_TRY(opaque([=] { return jtprod(st, bp->ops, dely, bp->tmp2); }));
    MSDP_TRY(axpby(st, n2, 1.0, bk.Rd, -1.0, bk.tmp2, bk.delS));
    MSDP_TRY(symmetrize(st, n, bk.delS, n));
    if (bk.ops.kind == BlockOps::EMPTY) MSDP_TRY(fill(st, n2, 0.0, bk.buf2));
    else MSDP_TRY(apply_E(bk, bk.tmp2, bk.ops.one_hot, bk.buf2));
    MSDP_TRY(axpbypcz(st, n2, mu_, bk.Si, -1.0, bk.X, 1.0, bk.buf2, bk.delX));
    if (bk.rd_nonzero) {
      MSDP_TRY(apply_E(bk, bk.Rd, false, bk.buf2));
      MSDP_TRY(axpby(st, n2, 1.0, bk.delX, -1.0, bk.buf2, bk.delX));
    }
    if (corrected) MSDP_TRY(axpby(st, n2, 1.0, bk.delX, 1.0, bk.corr, bk.delX));
    MSDP_TRY(symmetrize(st, n, bk.delX, n));
  }
  MSDP_TRY(rec_end());
  if (ns > 0) {
    MSDP_TRY(gemm(st, true, false, (int)ns, 1, (int)m, 1.0, Clin, m, dely, m, 0.0, ws1, ns));
    if (corrected) MSDP_HIP(hipMemcpyAsync(delx, corrlin, sizeof(double) * ns, hipMemcpyDeviceToDevice, st));
    MSDP_TRY(scalar_op(st, SOP_DIRS, (int)ns, mu_, x, z, zinv, Rdlin, ws1, delz, delx, corrected));
  }
  return 0;
}

// correction! (step.jl:263-271; directions.jl:90-100 NT, 123-127 HKM)
int msdp_solver::correction() {
  Range range_("madsdp:correction");
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    Blk& bk = blk[i];
    const int n = bk.n;
    if (nt()) {
      MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.Gi, n, bk.delX, n, 0.0, bk.tmp, n));
      MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.tmp, n, bk.delS, n, 0.0, bk.RNT, n));
      MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.RNT, n, bk.G, n, 0.0, bk.Mm, n));
      MSDP_TRY(nt_rnt(st, n, bk.Mm, bk.D, bk.RNT));
      MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.G, n, bk.RNT, n, 0.0, bk.tmp, n));
      MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.tmp, n, bk.G, n, 0.0, bk.corr, n));
    } else {
      MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.delX, n, bk.delS, n, 0.0, bk.tmp, n));
      MSDP_TRY(gemm(st, false, false, n, n, n, -1.0, bk.tmp, n, bk.Si, n, 0.0, bk.corr, n));
    }
    MSDP_TRY(symmetrize(st, n, bk.corr, n));
  }
  MSDP_TRY(rec_end());
  if (ns > 0) MSDP_TRY(scalar_op(st, SOP_CORR, (int)ns, 0, nullptr, nullptr, zinv, delx, delz, corrlin, nullptr, false));
  return 0;
}

// step_lengths (step.jl:186-198 EigenStep, 200-228 CholeskyBacktrack)
int msdp_solver::step_lengths(double& a, double& bt) {
  Range range_("madsdp:step_lengths");
  if (opt.step == MSDP_STEP_EIGEN) {
    MSDP_TRY(rec_begin());
    for (int i = 0; i < nb; ++i) {
      OnBlock on(this, i);
      Blk& bk = blk[i];
      const int n = bk.n;
      const double* Ls[2] = {bk.Lxinv, bk.Lsinv};
      const double* Dl[2] = {bk.delX, bk.delS};
      double* T[2] = {bk.tmp2, bk.buf2};
      if (!rec_on && n <= 1024) {
        // L_X^-1 dX L_X^-T and L_S^-1 dS L_S^-T in the same launches: a batch
        // of two at the matrices' distances (intermediates tmp and buf);
        // larger blocks keep the stream-K products of the whole chip
        const int64_t dL = Ls[1] - Ls[0], dD = Dl[1] - Dl[0];
        MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, Ls[0], n, Dl[0], n, 0.0, bk.tmp, n, 2, dL,
                      dD, bk.buf - bk.tmp, false, TRI_AL));
        MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.tmp, n, Ls[0], n, 0.0, T[0], n, 2,
                      bk.buf - bk.tmp, dL, T[1] - T[0], false, TRI_BU));
        MSDP_TRY(symmetrize2(st, n, T[0], T[1] - T[0], n));
      } else {
        for (int k = 0; k < 2; ++k) {
          MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, Ls[k], n, Dl[k], n, 0.0, bk.tmp, n, 1, 0,
                        0, 0, false, TRI_AL));
          MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.tmp, n, Ls[k], n, 0.0, T[k], n, 1, 0, 0,
                        0, false, TRI_BU));
          MSDP_TRY(symmetrize(st, n, T[k], n));
        }
      }
      if (n > JACOBI_VB_MAX) {  // the primal and dual eigenvalues in the same launches
        double *T0 = T[0], *T1 = T[1];
        double *o0 = slot(bslot(i, 3)), *o1 = slot(bslot(i, 4));
        MSDP_TRY(opaque([=] { return eig_min2(st, n, T0, T1, n, o0, o1, ework); }));
      }
    }
    MSDP_TRY(rec_end());
    if (nsmall > 0) MSDP_TRY(eigvals_vbatched(st, eb_step, 2 * nsmall));
    if (ns > 0) {
      MSDP_TRY(scalar_min_step(st, (int)ns, x, delx, slot(3)));
      MSDP_TRY(scalar_min_step(st, (int)ns, z, delz, slot(4)));
    }
    MSDP_TRY(read_slots());
    const double tau = opt.tau;
    auto step = [&](double lam) { return lam > -std::sqrt(EPS) ? 1.0 : std::min(1.0, -tau / lam); };
    a = bt = 1.0;
    for (int i = 0; i < nb; ++i) {
      a = std::min(a, step(hsc[bslot(i, 3)]));
      bt = std::min(bt, step(hsc[bslot(i, 4)]));
    }
    if (ns > 0) {
      a = std::min(a, tau * hsc[3]);
      bt = std::min(bt, tau * hsc[4]);
    }
    return 0;
  }
  // CholeskyBacktrack: _backtrack for X (alpha) and S (beta)
  auto backtrack = [&](bool primal, double& out) -> int {
    double al = 1.0;
    for (int i = 0; i < nb; ++i) {
      Blk& bk = blk[i];
      const int n = bk.n;
      const double* Xm = primal ? bk.X : bk.S;
      const double* dX = primal ? bk.delX : bk.delS;
      for (;;) {
        MSDP_TRY(axpby(st, (int64_t)n * n, 1.0, Xm, al, dX, bk.buf));
        MSDP_TRY(potrf(st, n, bk.buf, n, info, trwork));
        int hinfo = 0;
        MSDP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st));
        MSDP_TRY(stream_sync(st));
        if (hinfo == 0) break;
        al *= opt.shrink;
        if (al < EPS) { out = 0.0; return 0; }
      }
    }
    if (al != 1.0) al *= opt.boundary;
    if (ns > 0) {
      MSDP_TRY(scalar_min_step(st, (int)ns, primal ? x : z, primal ? delx : delz, slot(3)));
      MSDP_TRY(read_slots());
      al = std::min(al, opt.boundary * hsc[3]);
    }
    out = al;
    return 0;
  };
  MSDP_TRY(backtrack(true, a));
  MSDP_TRY(backtrack(false, bt));
  return 0;
}

// affine_complementarity (step.jl:231-243)
int msdp_solver::affine_complementarity(double a, double bt, double& out) {
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    Blk& bk = blk[i];
    const int64_t n2 = (int64_t)bk.n * bk.n;
    MSDP_TRY(dot(st, n2, bk.X, bk.S, slot(bslot(i, 3)), red));
    MSDP_TRY(dot(st, n2, bk.X, bk.delS, slot(bslot(i, 4)), red));
    MSDP_TRY(dot(st, n2, bk.delX, bk.S, slot(bslot(i, 5)), red));
    MSDP_TRY(dot(st, n2, bk.delX, bk.delS, slot(bslot(i, 6)), red));
  }
  MSDP_TRY(rec_end());
  if (ns > 0) MSDP_TRY(scalar_affine(st, (int)ns, x, delx, z, delz, a, bt, slot(5)));
  MSDP_TRY(read_slots());
  double s = 0.0;
  for (int i = 0; i < nb; ++i) {
    s += hsc[bslot(i, 3)];
    s += bt * hsc[bslot(i, 4)];
    s += a * hsc[bslot(i, 5)];
    s += a * bt * hsc[bslot(i, 6)];
  }
  if (ns > 0) s += hsc[5];
  out = s / std::max(barrier_dim(), 1);
  return 0;
}

int msdp_solver::total_complementarity(double& out) {
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    MSDP_TRY(dot(st, (int64_t)blk[i].n * blk[i].n, blk[i].X, blk[i].S, slot(bslot(i, 3)), red));
  }
  MSDP_TRY(rec_end());
  if (ns > 0) MSDP_TRY(dot(st, ns, x, z, slot(5), red));
  MSDP_TRY(read_slots());
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += hsc[bslot(i, 3)];
  if (ns > 0) s += hsc[5];
  out = s;
  return 0;
}

// _take_step! (primal_dual.jl:56-72)
int msdp_solver::take_step(double a, double bt) {
  Range range_("madsdp:take_step");
  alpha = a;
  beta = bt;
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    Blk& bk = blk[i];
    const int64_t n2 = (int64_t)bk.n * bk.n;
    MSDP_TRY(axpby(st, n2, 1.0, bk.X, a, bk.delX, bk.X));
    MSDP_TRY(symmetrize(st, bk.n, bk.X, bk.n));
    MSDP_TRY(axpby(st, n2, 1.0, bk.S, bt, bk.delS, bk.S));
    MSDP_TRY(symmetrize(st, bk.n, bk.S, bk.n));
  }
  MSDP_TRY(rec_end());
  if (m > 0) MSDP_TRY(axpby(st, m, 1.0, y, bt, dely, y));
  if (ns > 0) {
    MSDP_TRY(axpby(st, ns, 1.0, x, a, delx, x));
    MSDP_TRY(axpby(st, ns, 1.0, z, bt, delz, z));
  }
  iter += 1;
  return update_residuals();
}

// iterate! (primal_dual.jl:85-121)
int msdp_solver::iterate() {
  Range range_("madsdp:iterate");
  if (dual_centering()) return dc_iterate();
  bool ok1 = false, ok2 = false;
  MSDP_TRY(prepare(ok1));
  if (ok1) MSDP_TRY(factorize_schur(ok2));
  if (!ok1 || !ok2) { status = ERROR_IN_STEP_COMPUTATION; return 0; }
  const double nu = std::max(barrier_dim(), 1);
  if (opt.centering == MSDP_CENTER_MEHROTRA) {
    double compl_ = 0;
    MSDP_TRY(total_complementarity(compl_));
    const double mu_ = compl_ / nu;
    mu = mu_;
    MSDP_TRY(solve_directions(0.0, false));
    double aa, ba;
    MSDP_TRY(step_lengths(aa, ba));
    double mu_aff;
    MSDP_TRY(affine_complementarity(aa, ba, mu_aff));
    const double sg = std::min(std::max(std::pow(mu_aff / std::max(mu_, EPS), opt.sigma_exponent), 1e-8), 1.0);
    sigma = sg;
    MSDP_TRY(correction());
    MSDP_TRY(solve_directions(sg * mu_, true));
    double a, bt;
    MSDP_TRY(step_lengths(a, bt));
    return take_step(a, bt);
  }
  MSDP_TRY(solve_directions(mu, false));
  double a, bt;
  MSDP_TRY(step_lengths(a, bt));
  MSDP_TRY(take_step(a, bt));
  double compl_ = 0;
  MSDP_TRY(total_complementarity(compl_));
  double mu_ = opt.ratio * compl_ / nu;
  if (std::min(a, bt) < 0.5) mu_ *= 1.5;
  if (a + bt > 1.6) mu_ *= 0.75;
  if (a + bt > 1.9) mu_ /= 1 + 0.1 * iter;
  mu = mu_;
  return 0;
}

// update_termination! (primal_dual.jl:123-135, dual_centering.jl:293-303)
int msdp_solver::update_termination() {
  Range range_("madsdp:update_termination");
  if (dual_centering()) MSDP_TRY(update_residuals());
  MSDP_TRY(rec_begin());
  for (int i = 0; i < nb; ++i) {
    OnBlock on(this, i);
    MSDP_TRY(dot(st, (int64_t)blk[i].n * blk[i].n, blk[i].C, blk[i].X, slot(bslot(i, 7)), red));
  }
  MSDP_TRY(rec_end());
  if (ns > 0) {
    MSDP_TRY(dot(st, ns, dlin, x, slot(6), red));
    MSDP_TRY(max_abs(st, ns, Rdlin, slot(7), red));
  }
  if (m > 0) {
    MSDP_TRY(dot(st, m, b, y, slot(8), red));
    MSDP_TRY(max_abs(st, m, Rp, slot(9), red));
  }
  MSDP_TRY(read_slots());
  double p = 0.0;
  for (int i = 0; i < nb; ++i) p += hsc[bslot(i, 7)];
  if (ns > 0) p += hsc[6];
  const double d = m > 0 ? hsc[8] : 0.0;
  pobj = p;
  dobj = d;
  obj_val = p;
  inf_pr = (m > 0 ? hsc[9] : 0.0) / std::max(1.0, bnorm);
  double rd = 0.0;
  for (int i = 0; i < nb; ++i) rd = std::max(rd, hsc[bslot(i, 1)]);
  double cn = cnorm_max;
  if (ns > 0) {
    rd = std::max(rd, hsc[7]);
    cn = std::max(cn, dnorm);
  }
  inf_du = rd / std::max(1.0, cn);
  bool converged;
  if (dual_centering()) {
    inf_compl = dc_nu / std::max(dc_t, EPS) / std::max(1.0, std::fabs(p) + std::fabs(d));
    converged = inf_compl <= opt.tol;
  } else {
    inf_compl = std::fabs(p - d) / std::max(1.0, std::fabs(p) + std::fabs(d));
    converged = std::max(inf_pr, std::max(inf_du, inf_compl)) <= opt.tol;
  }
  if (converged) status = SOLVE_SUCCEEDED;
  else if (iter >= opt.max_iter) status = MAXIMUM_ITERATIONS_EXCEEDED;
  else if (now_s() - start_time >= opt.max_wall_time) status = MAXIMUM_WALLTIME_EXCEEDED;
  if (!dual_centering() && status == REGULAR && mu < opt.mu_min)
    status = SEARCH_DIRECTION_BECOMES_TOO_SMALL;
  trace.push_back({p, d, inf_pr, inf_du, inf_compl, mu, alpha, beta, sigma});
  return 0;
}

// initial_point! (solver.jl:214-284)
int msdp_solver::initial_point() {
  if (opt.initial_point == MSDP_INIT_SCALED_IDENTITY) {
    double cscale = cnorm_max;
    if (ns > 0) cscale = std::max(cscale, dnorm);
    double beta0 = opt.identity_scale;
    if (!(beta0 > 0))
      beta0 = std::max(10.0, std::sqrt(std::max(bnorm, 1.0) * std::max(cscale, 1.0)) *
                                 std::sqrt((double)barrier_dim()));
    for (auto& bk : blk) {
      MSDP_TRY(set_identity(st, bk.n, bk.X, bk.n, beta0));
      MSDP_TRY(set_identity(st, bk.n, bk.S, bk.n, beta0));
    }
    if (m > 0) MSDP_TRY(fill(st, m, 0.0, y));
    if (ns > 0) {
      MSDP_TRY(fill(st, ns, beta0, x));
      MSDP_TRY(fill(st, ns, beta0, z));
    }
    return 0;
  }
  if (opt.initial_point == MSDP_INIT_POINT) {
    if ((int64_t)ip_y.size() != m) return as_err("InitialPoint not set");
    if (m > 0) MSDP_HIP(hipMemcpyAsync(y, ip_y.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
    size_t off = 0;
    for (auto& bk : blk) {
      const size_t n2 = (size_t)bk.n * bk.n;
      MSDP_HIP(hipMemcpyAsync(bk.X, ip_X.data() + off, sizeof(double) * n2, hipMemcpyHostToDevice, st));
      off += n2;
    }
    if (ns > 0) {
      if (ip_x.empty()) MSDP_TRY(fill(st, ns, 1.0, x));
      else MSDP_HIP(hipMemcpyAsync(x, ip_x.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st));
    }
    MSDP_TRY(set_dual_slacks(y));
    MSDP_TRY(stream_sync(st));
    return 0;
  }
  // DualShift
  for (auto& bk : blk) MSDP_TRY(set_identity(st, bk.n, bk.X, bk.n, 1.0));
  if (ns > 0) MSDP_TRY(fill(st, ns, 1.0, x));
  double rho = 0.0;
  while (rho <= opt.max_shift) {
    if (m > 0) MSDP_TRY(fill(st, m, -rho, y));
    MSDP_TRY(set_dual_slacks(y));
    // _dual_interior (solver.jl:255-258)
    MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * (4 * nb + 8), st));
    for (int i = 0; i < nb; ++i) MSDP_TRY(cholesky_factor(blk[i], blk[i].S, blk[i].Ls, binfo(i, 0)));
    std::vector<int> inf;
    MSDP_TRY(read_infos(inf));
    bool interior = true;
    for (int i = 0; i < nb; ++i) interior = interior && inf[8 + 4 * i] == 0;
    if (interior && ns > 0) {  // all(z .> 0)
      MSDP_TRY(vec_min(st, (int)ns, z, slot(10)));
      MSDP_TRY(read_slots());
      interior = hsc[10] > 0;
    }
    if (interior) {
      if (rho != 0.0) {
        if (m > 0) MSDP_TRY(fill(st, m, -2 * rho, y));
        MSDP_TRY(set_dual_slacks(y));
      }
      return 0;
    }
    rho = rho == 0.0 ? opt.initial_shift : 2 * rho;
  }
  set_error("MadSDP: could not find a dual interior point `y = -rho 1`; use `InitialPoint` to give one.");
  return -3;
}

// initialize! (solver.jl:286-293) + initialize_algorithm!
int msdp_solver::initialize() {
  Range range_("madsdp:initialize");
  MSDP_TRY(finalize());
  MSDP_TRY(initial_point());
  iter = 0;
  status = REGULAR;
  start_time = now_s();
  trace.clear();
  alpha = beta = sigma = 0.0;
  force_fp64 = false;
  lowp_count = fallback_count = refine_steps = 0;
  if (dual_centering()) return dc_initialize();
  MSDP_TRY(update_residuals());
  double compl_ = 0;
  MSDP_TRY(total_complementarity(compl_));
  const double nu = std::max(barrier_dim(), 1);
  mu = opt.centering == MSDP_CENTER_MEHROTRA ? compl_ / nu : opt.initial_ratio * compl_ / nu;
  return 0;
}

int msdp_solver::solve() {
  MSDP_TRY(initialize());
  MSDP_TRY(update_termination());
  while (status == REGULAR) {
    MSDP_TRY(iterate());
    if (status != REGULAR) break;
    MSDP_TRY(update_termination());
  }
  MSDP_TRY(stream_sync(st));
  return 0;
}

// ==========================================================================
// DualCentering (src/dual_centering.jl)
// ==========================================================================
int msdp_solver::refresh_slacks(bool& ok) {  // :83-86
  MSDP_TRY(set_dual_slacks(y));
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * (4 * nb + 8), st));
  for (int i = 0; i < nb; ++i) MSDP_TRY(prepare_block(blk[i], i));
  std::vector<int> inf;
  MSDP_TRY(read_infos(inf));
  ok = true;
  for (int i = 0; i < nb; ++i) ok = ok && inf[8 + 4 * i] == 0;
  return 0;
}

int msdp_solver::newton_direction(bool& ok) {  // :89-110
  MSDP_TRY(fill(st, m, 0.0, wm));
  MSDP_TRY(jprod_blocks(JP_SI, wm));
  MSDP_TRY(axpby(st, m, 1.0, wm, -dc_t, b, dc_g));
  MSDP_TRY(factorize_schur(ok));
  if (!ok) return 0;
  MSDP_HIP(hipMemcpyAsync(dc_d, dc_g, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
  MSDP_TRY(hsolve(dc_d));
  MSDP_TRY(dot(st, m, dc_g, dc_d, slot(11), red));
  MSDP_TRY(read_slots());
  dc_lambda = std::sqrt(std::max(hsc[11], 0.0));
  return 0;
}

int msdp_solver::dc_initialize() {  // :112-143
  dc_nu = 0;
  for (auto& bk : blk) dc_nu += bk.n;
  if (opt.dc_A < 0) {
    const double tau = 2 * opt.dc_beta / ((1 - opt.dc_beta) * (1 - opt.dc_beta));
    dc_A = opt.dc_delta + (-tau - std::log1p(-tau));
  } else {
    dc_A = opt.dc_A;
  }
  n_predictor = n_newton = 0;
  bool ok = false;
  MSDP_TRY(refresh_slacks(ok));
  if (!ok) { set_error("MadSDP: the initial `y` does not give a dual slack `S > 0`."); return -3; }
  dc_t = 0.0;
  MSDP_TRY(newton_direction(ok));
  if (!ok) { set_error("MadSDP: failed to factorize the dual Hessian at y0."); return -3; }
  MSDP_HIP(hipMemcpyAsync(dc_dy, b, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
  MSDP_TRY(hsolve(dc_dy));
  MSDP_TRY(dot(st, m, b, dc_dy, slot(12), red));
  MSDP_TRY(dot(st, m, b, dc_d, slot(13), red));
  MSDP_TRY(read_slots());
  const double bBb = std::max(hsc[12], EPS);
  const double bBg = hsc[13];
  dc_t = std::max(bBg / bBb, opt.dc_beta / std::sqrt(bBb));
  for (int k = 0; k < opt.dc_centering_max_iter; ++k) {
    MSDP_TRY(newton_direction(ok));
    if (!ok) { set_error("MadSDP: failed to factorize the dual Hessian during the initial centering."); return -3; }
    if (dc_lambda <= opt.dc_beta / 2) break;
    MSDP_TRY(axpby(st, m, 1.0, y, -1.0 / (1 + dc_lambda), dc_d, y));
    iter += 1;
    n_newton += 1;
    MSDP_TRY(refresh_slacks(ok));
    if (!ok) { set_error("MadSDP: damped Newton left the interior during the initial centering."); return -3; }
  }
  mu = 1.0 / dc_t;
  return 0;
}

int msdp_solver::predictor_step(bool& ok) {  // :232-264
  MSDP_TRY(axpby(st, m, 1.0, y, 1.0, dc_d, dc_yhat));
  MSDP_HIP(hipMemcpyAsync(dc_dy, b, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
  MSDP_TRY(hsolve(dc_dy));
  MSDP_TRY(axpby(st, m, dc_t, dc_dy, 0.0, dc_dy, dc_dy));
  MSDP_HIP(hipMemsetAsync(info, 0, sizeof(int) * (4 * nb + 8), st));
  int64_t off = 0;
  for (int i = 0; i < nb; ++i) {
    Blk& bk = blk[i];
    const int n = bk.n;
    const int64_t n2 = (int64_t)n * n;
    MSDP_TRY(slack(bk, dc_yhat, bk.Shat));
    MSDP_TRY(jtprod(st, bk.ops, dc_dy, bk.Dc));
    MSDP_TRY(symmetrize(st, n, bk.Dc, n));
    // proximity eigenvalues (:205-215)
    MSDP_TRY(cholesky_factor(bk, bk.Shat, bk.Lh, binfo(i, 2)));
    MSDP_TRY(trtri_lower(st, n, bk.Lh, n, bk.Lx, n, trwork));  // Lx reused for Lhat^-1
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.Lx, n, bk.Dc, n, 0.0, bk.tmp, n, 1, 0, 0, 0,
                  false, TRI_AL));
    MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.tmp, n, bk.Lx, n, 0.0, bk.tmp2, n, 1, 0, 0, 0,
                  false, TRI_BU));
    MSDP_TRY(symmetrize(st, n, bk.tmp2, n));
    if (n > JACOBI_VB_MAX)
      MSDP_TRY(syevd(st, n, bk.tmp2, n, eigall + off, bk.Vv, n, false, ework, eiwork));
    // ||Ls^-1 Shat Ls^-T||_F^2 (:219-229)
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.Lsinv, n, bk.Shat, n, 0.0, bk.tmp, n, 1, 0, 0,
                  0, false, TRI_AL));
    MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.tmp, n, bk.Lsinv, n, 0.0, bk.buf2, n, 1, 0, 0,
                  0, false, TRI_BU));
    MSDP_TRY(dot(st, n2, bk.buf2, bk.buf2, slot(bslot(i, 6)), red));
    off += n;
  }
  if (nsmall > 0) MSDP_TRY(eigvals_vbatched(st, eb_dc, nsmall));
  std::vector<int> inf;
  MSDP_TRY(read_infos(inf));
  for (int i = 0; i < nb; ++i)
    if (inf[8 + 4 * i + 2] != 0) { ok = false; return 0; }
  MSDP_TRY(dc_solve_alpha(st, (int)off, eigall, dc_A, slot(14)));
  MSDP_TRY(read_slots());
  double shat_n2 = 0.0;
  for (int i = 0; i < nb; ++i) shat_n2 += hsc[bslot(i, 6)];
  const double a = hsc[14];
  alpha = beta = a;
  for (auto& bk : blk) {
    const int n = bk.n;
    const int64_t n2 = (int64_t)n * n;
    MSDP_TRY(axpby(st, n2, 1 - a, bk.Shat, a, bk.Dc, bk.buf));
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.Si, n, bk.buf, n, 0.0, bk.buf2, n));
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0 / dc_t, bk.buf2, n, bk.Si, n, 0.0, bk.X, n));
    MSDP_TRY(symmetrize(st, n, bk.X, n));
  }
  MSDP_TRY(axpby(st, m, 1.0, dc_yhat, a, dc_dy, y));
  dc_t = dc_nu * dc_t / ((1 - a) * shat_n2);
  n_predictor += 1;
  ok = true;
  return 0;
}

int msdp_solver::dc_iterate() {  // :266-289
  Range range_("madsdp:dc_iterate");
  bool ok = false;
  MSDP_TRY(newton_direction(ok));
  if (!ok) { status = ERROR_IN_STEP_COMPUTATION; return 0; }
  if (dc_lambda > opt.dc_beta) {
    MSDP_TRY(axpby(st, m, 1.0, y, -1.0 / (1 + dc_lambda), dc_d, y));
    alpha = beta = 1.0 / (1 + dc_lambda);
    n_newton += 1;
  } else {
    MSDP_TRY(predictor_step(ok));
    if (!ok) { status = ERROR_IN_STEP_COMPUTATION; return 0; }
  }
  iter += 1;
  MSDP_TRY(refresh_slacks(ok));
  if (!ok) { status = ERROR_IN_STEP_COMPUTATION; return 0; }
  mu = 1.0 / dc_t;
  return 0;
}

// ==========================================================================
// Mixed precision: FP32 MFMA Schur assembly + FP64 iterative refinement
// ==========================================================================
int msdp_solver::ensure_mixed() {
  if (bf16x3()) {
    size_t pw = 0, pv = 0;
    bool need_pw = false;
    for (auto& bk : blk) {
      if (bk.ops.kind != BlockOps::DENSE) continue;
      const int mi = bk.ops.ncon;
      const int64_t drow = dense_panel_rows(bk.n, !nt());
      pw = std::max(pw, panel_apply_work_elems((int)drow, mi));
      if (!bk.ops.dcon_identity) pv = std::max(pv, (size_t)mi);
      if (bk.Bh) continue;
      need_pw = true;
      const size_t ld = (size_t)bf16x3_ld(drow);
      MSDP_TRY(alloc(&bk.Bh, ld * mi));
      MSDP_TRY(alloc(&bk.Bl, ld * mi));
    }
    if (need_pw) {
      MSDP_TRY(alloc(&pwork, pw));
      if (pv > 0) {
        MSDP_TRY(alloc(&pvb, pv));
        MSDP_TRY(alloc(&pu, pv));
      }
    }
    return 0;
  }
  size_t want = 0;
  for (auto& bk : blk) {
    if (bk.ops.kind != BlockOps::DENSE || bk.mats32) continue;
    const size_t n2 = (size_t)bk.n * bk.n;
    const int mi = bk.ops.ncon;
    MSDP_TRY(alloc(&bk.mats32, n2 * mi));
    MSDP_TRY(d2f(st, (int64_t)(n2 * mi), bk.ops.mats, bk.mats32));
    MSDP_TRY(alloc(&bk.B32, n2 * mi));  // vec (HKM) is the larger layout
    MSDP_TRY(alloc(&bk.Lm32, n2));
    MSDP_TRY(alloc(&bk.F32, n2));
    want = std::max(want, 2 * n2 * (size_t)std::min(mi, 64));
  }
  if (want > swork32_elems) {
    MSDP_TRY(alloc(&swork32, want));
    swork32_elems = want;
  }
  return 0;
}

// <A_j, L A_k R> of a dense block by BF16x3 split products: the FP64 panel
// B (svec / vec of P_j = op(Lm) A_j op(F2), the FP64 path's own), split into
// hi = bf16(B) and lo = bf16(B - hi), then H += hi'hi + hi'lo + lo'hi on the
// bf16 MFMA with FP32 accumulation.  The FP64 panel stays for the refinement.
int msdp_solver::schur_block_bf16x3(Blk& bk, const SchurScaling& sc) {
  const BlockOps& o = bk.ops;
  const int mi = o.ncon;
  MSDP_TRY(dense_panel(st, o, sc, swork, swork_elems));
  const int64_t drow = dense_panel_rows(bk.n, sc.full_vec);
  const int64_t ld = bf16x3_ld(drow);
  MSDP_TRY(bf16_split(st, (int)drow, mi, o.Bbuf, drow, bk.Bh, bk.Bl, ld));
  if (o.dcon_identity) {
    MSDP_TRY(syrk_bf16x3(st, mi, (int)ld, 1.0, bk.Bh, bk.Bl, ld, 1.0, Hs, m));
  } else {
    MSDP_TRY(syrk_bf16x3(st, mi, (int)ld, 1.0, bk.Bh, bk.Bl, ld, 0.0, o.Hbuf, mi));
    MSDP_TRY(scatter_lower_add(st, mi, o.con_list, o.Hbuf, Hs, m));
  }
  bk.panel_fp64 = true;
  bk.panel_rows = drow;
  return 0;
}

// <A_j, L A_k R> of a dense block with FP32 operands and accumulation:
// P_j = op(Lm) A_j op(F2) by batched FP32 GEMMs, B = svec(P) (vec for HKM),
// H += B' B widened to FP64.
int msdp_solver::schur_block_fp32(Blk& bk, const SchurScaling& sc) {
  const int n = bk.n;
  const int64_t n2 = (int64_t)n * n;
  const BlockOps& o = bk.ops;
  const int mi = o.ncon;
  const int64_t drow = sc.full_vec ? n2 : (int64_t)n * (n + 1) / 2;
  MSDP_TRY(d2f(st, n2, sc.Lm, bk.Lm32));
  MSDP_TRY(d2f(st, n2, sc.F2, bk.F32));
  const int cb = (int)std::max<int64_t>(1, std::min<int64_t>(mi, (int64_t)(swork32_elems / (2 * n2))));
  float* T = swork32;
  float* P = swork32 + (int64_t)cb * n2;
  for (int j0 = 0; j0 < mi; j0 += cb) {
    const int c = std::min(cb, mi - j0);
    MSDP_TRY(gemm32(st, false, sc.transF2, n, n, n, 1.0f, bk.mats32 + (int64_t)j0 * n2, n, bk.F32,
                    n, 0.0, T, n, c, n2, 0, n2));
    float* dst = sc.full_vec ? bk.B32 + (int64_t)j0 * n2 : P;
    MSDP_TRY(gemm32(st, sc.transL, false, n, n, n, 1.0f, bk.Lm32, n, T, n, 0.0, dst, n, c, 0, n2, n2));
    if (!sc.full_vec) MSDP_TRY(svec_pack32(st, n, c, P, bk.B32 + (int64_t)j0 * drow, drow));
  }
  if (o.dcon_identity) {
    MSDP_TRY(gemm32_to64(st, true, false, mi, mi, (int)drow, 1.0f, bk.B32, drow, bk.B32, drow, 1.0,
                         Hs, m, true));
  } else {
    MSDP_TRY(gemm32_to64(st, true, false, mi, mi, (int)drow, 1.0f, bk.B32, drow, bk.B32, drow, 0.0,
                         o.Hbuf, mi, true));
    MSDP_TRY(scatter_lower_add(st, mi, o.con_list, o.Hbuf, Hs, m));
  }
  return 0;
}

// out = (H + reg I) v with the FP64 operator H_jk = <A_j, L A_k R> +
// C_lin diag(x ./ z) C_lin', applied matrix-free: A(L A*(v) R).
int msdp_solver::apply_H(const double* v, double* out) {
  MSDP_TRY(axpby(st, m, reg_used, v, 0.0, v, out));
  for (auto& bk : blk) {
    if (bk.ops.kind == BlockOps::EMPTY) continue;
    if (bk.panel_fp64) {
      // the block's H_jk = B[:, j]' B[:, k] with its FP64 panel
      const BlockOps& o = bk.ops;
      const int drow = (int)bk.panel_rows;
      if (o.dcon_identity) {
        MSDP_TRY(panel_apply(st, drow, o.ncon, o.Bbuf, drow, v, out, true, pwork));
      } else {
        MSDP_TRY(gather_vec(st, o.ncon, o.con_list, v, pvb));
        MSDP_TRY(panel_apply(st, drow, o.ncon, o.Bbuf, drow, pvb, pu, false, pwork));
        MSDP_TRY(scatter_add_vec(st, o.ncon, o.con_list, pu, out));
      }
      continue;
    }
    const int n = bk.n;
    const double *L, *R;
    if (dual_centering()) { L = bk.Si; R = bk.Si; }
    else if (nt()) { L = bk.W; R = bk.W; }
    else { L = bk.Si; R = bk.X; }
    MSDP_TRY(jtprod(st, bk.ops, v, bk.tmp2));
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, L, n, bk.tmp2, n, 0.0, bk.tmp, n));
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.tmp, n, R, n, 0.0, bk.buf, n));
    MSDP_TRY(symmetrize(st, n, bk.buf, n));
    MSDP_TRY(jprod(st, bk.ops, bk.buf, out));
  }
  if (ns > 0) {
    MSDP_TRY(gemm(st, true, false, (int)ns, 1, (int)m, 1.0, Clin, m, v, m, 0.0, ws1, ns));
    MSDP_TRY(scalar_op(st, SOP_XZ, (int)ns, 0, x, nullptr, zinv, nullptr, nullptr, ws2, nullptr, false));
    MSDP_TRY(qr_mul(st, (int)ns, ws1, ws2, ws1));
    MSDP_TRY(gemm(st, false, false, (int)m, 1, (int)ns, 1.0, Clin, m, ws1, ns, 1.0, out, m));
  }
  return 0;
}

// The solve with the last factorization of H: Cholesky (DenseCholesky) or
// Bunch-Kaufman LDL' (MadNLPSchur).
int msdp_solver::hsolve(double* v) {
  if (ldlt()) return sytrs_bk(st, (int)m, H, m, ldlt_ipiv(), v);
  // every factorisation serves at least the predictor and corrector solves:
  // the 512 x 512 diagonal-block inverses (one batched launch sequence) turn
  // each super-block's dependent 64-block steps into one wide product
  MSDP_TRY(form_block_inverses());
  return potrs(st, (int)m, H, Ht, m, Hdinv, v, wm2, hdi_valid ? Hdi : nullptr,
               hdi_valid ? Hdit : nullptr);
}

// schur_solve! (schur_solvers.jl:21).  With an FP32-assembled H, the FP64
// solution is recovered by iterative refinement against the FP64 operator;
// if it stalls above 1e-9 relative, H is re-assembled in FP64.
int msdp_solver::schur_solve(double* v) {
  if (!h_approx) return hsolve(v);
  MSDP_TRY(form_block_inverses());
  const double* di = hdi_valid ? Hdi : nullptr;
  const double* dit = hdi_valid ? Hdit : nullptr;
  MSDP_HIP(hipMemcpyAsync(rh, v, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
  MSDP_TRY(dot(st, m, rh, rh, slot(15), red));
  MSDP_TRY(potrs(st, (int)m, H, Ht, m, Hdinv, v, wm2, di, dit));
  double prev = INFINITY, hn = 0.0, rn = 0.0;
  for (int it = 0;; ++it) {
    MSDP_TRY(apply_H(v, rr));
    MSDP_TRY(axpby(st, m, 1.0, rh, -1.0, rr, rr));
    MSDP_TRY(dot(st, m, rr, rr, slot(16), red));
    MSDP_TRY(read_slots());
    hn = std::sqrt(hsc[15]);
    rn = std::sqrt(hsc[16]);
    if (!(rn > 1e-12 * hn) || it >= opt.refine_max_iter || !(rn < 0.5 * prev)) break;
    prev = rn;
    MSDP_TRY(potrs(st, (int)m, H, Ht, m, Hdinv, rr, wm2, di, dit));
    MSDP_TRY(axpby(st, m, 1.0, v, 1.0, rr, v));
    ++refine_steps;
  }
  if (rn <= 1e-9 * hn || !std::isfinite(hn)) return 0;
  // stalled: FP64 assembly from here on (the conditioning of H only grows
  // as mu decreases)
  ++fallback_count;
  force_fp64 = true;
  bool ok = false;
  MSDP_TRY(factorize_schur(ok));
  if (!ok) { set_error("FP64 re-factorisation of the Schur complement failed"); return -4; }
  MSDP_HIP(hipMemcpyAsync(v, rh, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
  return hsolve(v);
}

// ==========================================================================
// QRSchur (src/qr_schur.jl)
// ==========================================================================
// factorize_schur! (qr_schur.jl:24-88)
int msdp_solver::qr_factorize(bool& ok) {
  Range range_("madsdp:qr_factorize");
  ok = false;
  int64_t rows = ns;
  for (auto& bk : blk) rows += (int64_t)bk.n * (bk.n + 1) / 2;
  if (!qB) {
    if ((double)rows * (double)std::max<int64_t>(m, 1) * 8.0 > 64e9)
      return as_err("QRSchur: the dense scaled constraint matrix would exceed 64 GB");
    const size_t rr = std::max<int64_t>(rows, 1), mm = std::max<int64_t>(m, 1);
    MSDP_TRY(alloc(&qB, rr * mm));
    MSDP_TRY(alloc(&qbase, rr));
    MSDP_TRY(alloc(&qy, rr));
    MSDP_TRY(alloc(&qscale, mm));
    MSDP_TRY(alloc(&qtau, mm));
    MSDP_TRY(alloc(&qT, qr_t_elems((int)mm)));
    MSDP_TRY(alloc(&qwork, qr_work_elems((int64_t)rr, (int)mm)));
    MSDP_TRY(alloc(&qc, mm));
    MSDP_TRY(alloc(&qt, mm));
    MSDP_TRY(alloc(&qrd, 2));
    MSDP_TRY(alloc(&qbad, 1));
    qrows = rows;
  }
  if (m == 0) { ok = true; return 0; }
  MSDP_TRY(fill(st, rows * m, 0.0, qB));
  int64_t off = 0;
  for (auto& bk : blk) {
    const int n = bk.n;
    const int64_t n2 = (int64_t)n * n;
    const BlockOps& o = bk.ops;
    if (o.kind == BlockOps::DENSE) {
      const int cb = (int)std::max<int64_t>(1, std::min<int64_t>(o.ncon, (int64_t)(swork_elems / (2 * n2))));
      for (int j0 = 0; j0 < o.ncon; j0 += cb) {
        const int c = std::min(cb, o.ncon - j0);
        double* T = swork;
        double* P = swork + (int64_t)cb * n2;
        MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, o.mats + (int64_t)j0 * n2, n, bk.G, n, 0.0, T, n, c, n2, 0, n2));
        MSDP_TRY(gemm(st, true, false, n, n, n, 1.0, bk.G, n, T, n, 0.0, P, n, c, 0, n2, n2));
        MSDP_TRY(qr_pack_dense(st, n, c, P, o.con_list + j0, qB, rows, off));
      }
    } else if (o.kind == BlockOps::COO) {
      MSDP_TRY(qr_pack_coo(st, o, bk.G, qB, rows, off));
    } else if (o.kind == BlockOps::LOWRANK) {
      MSDP_TRY(qr_pack_lowrank(st, o, bk.G, qB, rows, off));
    }
    off += (int64_t)n * (n + 1) / 2;
  }
  MSDP_TRY(qr_scalar_rows(st, (int)m, (int)ns, Clin, x, z, qB, rows, off));
  // more multipliers than rows (qr_schur.jl:83-85); the zero-column check
  // (:74-77) reports the same failure, so its order does not matter
  if (rows < m) return 0;
  MSDP_TRY(qr_factor(st, rows, (int)m, qB, qscale, qtau, qT, qwork, qbad, qrd));
  int bad = 0;
  double rd[2] = {0, 0};
  MSDP_HIP(hipMemcpyAsync(&bad, qbad, sizeof(int), hipMemcpyDeviceToHost, st));
  MSDP_HIP(hipMemcpyAsync(rd, qrd, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  MSDP_TRY(stream_sync(st));
  if (bad) return 0;              // zero or non-finite column (qr_schur.jl:74-77)
  // A dense Householder QR leaves roundoff where SPQR's exact zero pivot would
  // be; a pivot below this tolerance counts as zero (:87).
  if (!(rd[1] > 64 * EPS * std::max(1.0, rd[0]))) return 0;
  ok = true;
  return 0;
}

// solve_directions! (qr_schur.jl:90-160)
int msdp_solver::qr_solve_directions(double mu_, bool corrected) {
  Range range_("madsdp:qr_solve_directions");
  const int64_t rows = qrows;
  MSDP_TRY(fill(st, rows, 0.0, qbase));
  int64_t off = 0;
  for (auto& bk : blk) {
    const int n = bk.n;
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.Rd, n, bk.G, n, 0.0, bk.tmp, n));
    MSDP_TRY(gemm(st, true, false, n, n, n, 1.0, bk.G, n, bk.tmp, n, 0.0, bk.tmp2, n));
    MSDP_TRY(qr_base_block(st, n, bk.tmp2, bk.D, mu_, bk.RNT, corrected, qbase, off));
    off += (int64_t)n * (n + 1) / 2;
  }
  MSDP_TRY(qr_base_scalar(st, (int)ns, x, z, Rdlin, corrlin, mu_, corrected, qbase, off));
  if (m > 0) {
    // qbase = Q' base; c = R'^-1 (scale .* Rp); t = c - qbase[1:m]
    MSDP_HIP(hipMemcpyAsync(qy, qbase, sizeof(double) * rows, hipMemcpyDeviceToDevice, st));
    MSDP_TRY(qr_apply_vec(st, rows, (int)m, qB, qT, qwork, qy, true));
    MSDP_TRY(qr_mul(st, (int)m, qscale, Rp, qc));
    MSDP_TRY(qr_trsv(st, (int)m, qB, rows, qc, true));
    MSDP_TRY(qr_combine(st, (int)m, qc, qy, qt));
    // dely = scale .* (R^-1 t)
    MSDP_HIP(hipMemcpyAsync(dely, qt, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    MSDP_TRY(qr_trsv(st, (int)m, qB, rows, dely, false));
    MSDP_TRY(qr_mul(st, (int)m, qscale, dely, dely));
    // direction = base + Q [t; 0]
    MSDP_TRY(fill(st, rows, 0.0, qy));
    MSDP_HIP(hipMemcpyAsync(qy, qt, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    MSDP_TRY(qr_apply_vec(st, rows, (int)m, qB, qT, qwork, qy, false));
    MSDP_TRY(axpby(st, rows, 1.0, qbase, 1.0, qy, qy));
  } else {
    MSDP_HIP(hipMemcpyAsync(qy, qbase, sizeof(double) * rows, hipMemcpyDeviceToDevice, st));
  }
  off = 0;
  for (auto& bk : blk) {
    const int n = bk.n;
    const int64_t n2 = (int64_t)n * n;
    MSDP_TRY(qr_smat(st, n, qy, off, bk.buf));
    MSDP_TRY(gemm(st, false, false, n, n, n, 1.0, bk.G, n, bk.buf, n, 0.0, bk.tmp, n));
    MSDP_TRY(gemm(st, false, true, n, n, n, 1.0, bk.tmp, n, bk.G, n, 0.0, bk.delX, n));
    MSDP_TRY(symmetrize(st, n, bk.delX, n));
    MSDP_TRY(jtprod(st, bk.ops, dely, bk.tmp2));
    MSDP_TRY(axpby(st, n2, 1.0, bk.Rd, -1.0, bk.tmp2, bk.delS));
    MSDP_TRY(symmetrize(st, n, bk.delS, n));
    off += (int64_t)n * (n + 1) / 2;
  }
  if (ns > 0) {
    MSDP_TRY(gemm(st, true, false, (int)ns, 1, (int)m, 1.0, Clin, m, dely, m, 0.0, ws1, ns));
    MSDP_TRY(axpby(st, ns, 1.0, Rdlin, -1.0, ws1, delz));
    MSDP_TRY(qr_delx(st, (int)ns, x, z, qy, off, delx));
  }
  return 0;
}

// ==========================================================================
// C ABI
// ==========================================================================
// The stream-ordered allocator keeps freed memory for reuse instead of
// returning it to the driver, so solves that come and go (a serving loop)
// never pay for, or synchronise on, device allocation.
static void keep_pool(int device) {
  static bool done[64] = {};
  if (device < 0 || device >= 64 || done[device]) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  done[device] = true;
}

// A recorded region that an error return left open ends with the API call:
// the calling thread records nothing afterwards.
struct RecScope {
  msdp_solver* s;
  ~RecScope() {
    if (s && s->rec_on) {
      rec_active() = nullptr;
      s->rec_on = false;
      for (auto& v : s->rec.ops) v.clear();
    }
  }
};
// the solver's shared_device option for the calling thread, for the call
struct SharedScope {
  bool prev;
  explicit SharedScope(const msdp_solver* s) : prev(shared_device()) {
    shared_device() = s->opt.shared_device != 0;
  }
  ~SharedScope() { shared_device() = prev; }
};
#define GUARD0(s) \
  if (!(s)) return as_err("null solver"); \
  RecScope rec_scope_{const_cast<msdp_solver*>(s)}; \
  SharedScope shared_scope_{s};
#define GUARD(s) \
  GUARD0(s)      \
  if ((s)->released) return as_err("solver released: its device state was freed after the solve");
#define GUARD_READY(s) \
  GUARD(s)             \
  if (!(s)->finalized) return as_err("call msdp_solver_initialize first");

extern "C" {

const char* msdp_last_error(void) { return last_error(); }

void msdp_default_options(msdp_options* o) {
  std::memset(o, 0, sizeof *o);
  o->tol = 1e-7;
  o->max_iter = 100;
  o->max_wall_time = 3600;
  o->mu_min = 1e-14;
  o->regularize_schur = 1e-12;
  o->max_schur_reg = 8;
  o->algorithm = MSDP_ALG_PRIMAL_DUAL;
  o->direction = MSDP_DIR_NT;
  o->step = MSDP_STEP_EIGEN;
  o->centering = MSDP_CENTER_MEHROTRA;
  o->initial_point = MSDP_INIT_SCALED_IDENTITY;
  o->schur = MSDP_SCHUR_DENSE_CHOLESKY;
  o->tau = 0.95;
  o->shrink = 0.8;
  o->boundary = 0.95;
  o->sigma_exponent = 3.0;
  o->initial_ratio = 0.25;
  o->ratio = 0.5;
  o->identity_scale = 0.0;
  o->initial_shift = 1.0;
  o->max_shift = 1e10;
  o->dc_beta = 0.2;
  o->dc_delta = 1.0;
  o->dc_A = 2.0;
  o->dc_centering_max_iter = 50;
  o->precision = 0;
  o->mu_switch = 1e-5;
  o->refine_max_iter = 10;
  o->lowp_min_gflop = 500.0;
}

int msdp_solver_create(msdp_solver** out, int device, int32_t nblocks, const int64_t* dims,
                       int64_t m, int64_t n_scalar) {
  if (!out) return as_err("null out");
  *out = nullptr;
  if (nblocks < 0 || m < 0 || n_scalar < 0) return as_err("negative size");
  if (nblocks == 0 && n_scalar == 0) return as_err("MadSDP: model has no variables.");
  int ndev = 0;
  MSDP_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return as_err("no such HIP device");
  MSDP_HIP(hipSetDevice(device));
  auto* s = new msdp_solver();
  s->device = device;
  s->nb = nblocks;
  s->m = m;
  s->ns = n_scalar;
  s->blk.resize(nblocks);
  for (int i = 0; i < nblocks; ++i) {
    if (dims[i] <= 0) { delete s; return as_err("block dimension must be positive"); }
    s->blk[i].n = (int)dims[i];
  }
  msdp_default_options(&s->opt);
  keep_pool(device);
  *out = s;
  return 0;
}

int msdp_solver_destroy(msdp_solver* s) {
  delete s;
  return 0;
}

int msdp_solver_set_cost_block(msdp_solver* s, int32_t block, const double* C) {
  GUARD(s);
  if (block < 0 || block >= s->nb) return as_err("block out of range");
  if (s->released) return as_err("solver released");
  Blk& b = s->blk[block];
  const size_t n2 = (size_t)b.n * b.n;
  if (!s->finalized) {  // staged on the host until finalize
    b.h_C.assign(C, C + n2);
    return 0;
  }
  MSDP_HIP(hipSetDevice(s->device));
  MSDP_HIP(hipMemcpyAsync(b.C, C, sizeof(double) * n2, hipMemcpyHostToDevice, s->st));
  MSDP_TRY(stream_sync(s->st));
  return 0;
}

int msdp_solver_set_constraints_coo(msdp_solver* s, int32_t block, int64_t nnz, const int64_t* con,
                                    const int64_t* row, const int64_t* col, const double* val) {
  GUARD(s);
  if (block < 0 || block >= s->nb) return as_err("block out of range");
  if (s->finalized) return as_err("problem already finalized");
  Blk& b = s->blk[block];
  if (b.has_dense || b.has_lowrank) return as_err("block already has dense or low-rank constraints");
  for (int64_t e = 0; e < nnz; ++e) {
    if (con[e] < 0 || con[e] >= s->m || row[e] < 0 || row[e] >= b.n || col[e] < 0 || col[e] >= b.n)
      return as_err("constraint entry out of range");
    if (val[e] == 0.0) continue;
    b.h_con.push_back((int)con[e]);
    b.h_row.push_back((int)row[e]);
    b.h_col.push_back((int)col[e]);
    b.h_val.push_back(val[e]);
  }
  b.has_coo = true;
  return 0;
}

int msdp_solver_set_constraints_dense(msdp_solver* s, int32_t block, int64_t m_i,
                                      const int64_t* con_idx, const double* mats) {
  GUARD(s);
  if (block < 0 || block >= s->nb) return as_err("block out of range");
  if (s->finalized) return as_err("problem already finalized");
  Blk& b = s->blk[block];
  if (b.has_coo || b.has_lowrank) return as_err("block already has coo or low-rank constraints");
  for (int64_t j = 0; j < m_i; ++j) {
    if (con_idx[j] < 0 || con_idx[j] >= s->m) return as_err("constraint index out of range");
    if (j > 0 && con_idx[j] <= con_idx[j - 1]) return as_err("con_idx must be strictly increasing");
  }
  b.h_dcon.assign(con_idx, con_idx + m_i);
  b.h_mats.assign(mats, mats + (size_t)m_i * b.n * b.n);
  b.has_dense = true;
  return 0;
}

int msdp_solver_set_constraints_lowrank(msdp_solver* s, int32_t block, int64_t m_i,
                                        const int64_t* con_idx, const int64_t* rank_ptr,
                                        const double* factors, const double* weights) {
  GUARD(s);
  if (block < 0 || block >= s->nb) return as_err("block out of range");
  if (s->finalized) return as_err("problem already finalized");
  Blk& b = s->blk[block];
  if (b.has_coo || b.has_dense) return as_err("block already has coo or dense constraints");
  if (m_i < 0 || rank_ptr[0] != 0) return as_err("rank_ptr must start at 0");
  for (int64_t j = 0; j < m_i; ++j) {
    if (con_idx[j] < 0 || con_idx[j] >= s->m) return as_err("constraint index out of range");
    if (j > 0 && con_idx[j] <= con_idx[j - 1]) return as_err("con_idx must be strictly increasing");
    if (rank_ptr[j + 1] <= rank_ptr[j]) return as_err("each constraint needs at least one factor");
  }
  const int64_t R = rank_ptr[m_i];
  if (R > INT32_MAX / 2) return as_err("too many factor columns");
  b.h_lcon.assign(con_idx, con_idx + m_i);
  b.h_lptr.assign(rank_ptr, rank_ptr + m_i + 1);
  b.h_lF.assign(factors, factors + (size_t)R * b.n);
  b.h_lw.assign(weights, weights + R);
  b.has_lowrank = true;
  return 0;
}

int msdp_solver_set_rhs(msdp_solver* s, const double* bv) {
  GUARD(s);
  if (s->released) return as_err("solver released");
  if (!s->finalized) {
    s->h_b.assign(bv, bv + s->m);
    return 0;
  }
  MSDP_HIP(hipSetDevice(s->device));
  if (s->m > 0) {
    MSDP_HIP(hipMemcpyAsync(s->b, bv, sizeof(double) * s->m, hipMemcpyHostToDevice, s->st));
    MSDP_TRY(stream_sync(s->st));
  }
  return 0;
}

int msdp_solver_set_scalar_data(msdp_solver* s, const double* d_lin, const double* C_lin) {
  GUARD(s);
  if (s->released) return as_err("solver released");
  if (!s->finalized) {
    s->h_dlin.assign(d_lin, d_lin + s->ns);
    s->h_Clin.assign(C_lin, C_lin + s->ns * s->m);
    return 0;
  }
  MSDP_HIP(hipSetDevice(s->device));
  if (s->ns > 0) {
    MSDP_HIP(hipMemcpyAsync(s->dlin, d_lin, sizeof(double) * s->ns, hipMemcpyHostToDevice, s->st));
    if (s->m > 0)
      MSDP_HIP(hipMemcpyAsync(s->Clin, C_lin, sizeof(double) * s->ns * s->m, hipMemcpyHostToDevice, s->st));
    MSDP_TRY(stream_sync(s->st));
  }
  return 0;
}

int msdp_solver_configure(msdp_solver* s, const msdp_options* o) {
  GUARD(s);
  if (o->algorithm == MSDP_ALG_DUAL_CENTERING) {
    if (s->ns > 0) return as_err("MadSDP: `DualCentering` does not support scalar (LP) blocks yet.");
    if (s->nb == 0) return as_err("MadSDP: `DualCentering` needs at least one PSD block.");
  }
  if (o->schur == MSDP_SCHUR_QR &&
      !(o->algorithm == MSDP_ALG_PRIMAL_DUAL && o->direction == MSDP_DIR_NT))
    return as_err("QRSchur requires PrimalDual with the NT direction.");
  if (o->precision < 0 || o->precision > 2)
    return as_err("precision must be 0 (FP64), 1 (mixed FP32) or 2 (mixed BF16x3)");
  s->opt = *o;
  return 0;
}

int msdp_solver_set_initial_point(msdp_solver* s, const double* X, const double* yv, const double* xv) {
  GUARD(s);
  size_t tot = 0;
  for (auto& b : s->blk) tot += (size_t)b.n * b.n;
  s->ip_X.assign(X, X + tot);
  s->ip_y.assign(yv, yv + s->m);
  if (xv) s->ip_x.assign(xv, xv + s->ns);
  else s->ip_x.clear();
  return 0;
}

int msdp_solver_initialize(msdp_solver* s) { GUARD(s); MSDP_HIP(hipSetDevice(s->device)); return s->initialize(); }
int msdp_solver_iterate(msdp_solver* s) { GUARD_READY(s); MSDP_HIP(hipSetDevice(s->device)); return s->iterate(); }
int msdp_solver_update_termination(msdp_solver* s) { GUARD_READY(s); MSDP_HIP(hipSetDevice(s->device)); return s->update_termination(); }
int msdp_solver_solve(msdp_solver* s) { GUARD(s); MSDP_HIP(hipSetDevice(s->device)); return s->solve(); }
int msdp_solver_set_mu(msdp_solver* s, double mu) {
  GUARD_READY(s);
  s->mu = mu;
  return 0;
}
int msdp_solver_synchronize(msdp_solver* s) {
  GUARD0(s);
  if (s->st && !s->released) MSDP_TRY(stream_sync(s->st));
  return 0;
}
int msdp_solver_set_stream(msdp_solver* s, void* stream) {
  GUARD(s);
  if (s->finalized) return as_err("set the stream before msdp_solver_initialize");
  if (!stream) return as_err("null stream");
  if (s->st && s->own_stream) MSDP_HIP(hipStreamDestroy(s->st));
  s->st = (hipStream_t)stream;
  s->own_stream = false;
  return 0;
}
int msdp_solver_footprint(const msdp_solver* s, int64_t* bytes) {
  GUARD0(s);
  if (!bytes) return as_err("null bytes");
  *bytes = s->footprint();
  return 0;
}
int msdp_solver_upload(msdp_solver* s) {
  GUARD(s);
  MSDP_HIP(hipSetDevice(s->device));
  return s->finalize();
}
int msdp_solver_release(msdp_solver* s) {
  GUARD0(s);
  return s->release();
}
int msdp_solver_profile_begin(msdp_solver* s, int32_t cls) {
  GUARD_READY(s);
  MSDP_TRY(stream_sync(s->st));
  return prof_begin(cls, s->st);
}
int msdp_solver_profile_end(msdp_solver* s, double* ms, int64_t* launches, double* work) {
  GUARD0(s);
  return prof_end(ms, launches, work);
}
int msdp_set_tuning(const char* name, double value) {
  if (!name) { set_error("null name"); return -2; }
  return set_tuning(name, value);
}
int msdp_get_tuning(const char* name, double* value) {
  if (!name || !value) { set_error("null argument"); return -2; }
  return get_tuning(name, value);
}
int msdp_profile_begin(int32_t cls) {
  if (cls < PROF_GEMM || cls > PROF_TRD1) { set_error("no such kernel class"); return -2; }
  return prof_begin(cls, nullptr);
}
int msdp_profile_end(double* ms, int64_t* launches, double* work) {
  if (!ms || !launches || !work) { set_error("null output"); return -2; }
  return prof_end(ms, launches, work);
}

int msdp_solver_get_info(const msdp_solver* s, msdp_info* o) {
  GUARD0(s);
  o->mu = s->mu; o->sigma = s->sigma; o->alpha = s->alpha; o->beta = s->beta;
  o->inf_pr = s->inf_pr; o->inf_du = s->inf_du; o->inf_compl = s->inf_compl; o->obj_val = s->obj_val;
  o->pobj = s->pobj; o->dobj = s->dobj; o->iter = s->iter; o->status = s->status;
  o->n_newton = s->n_newton; o->n_predictor = s->n_predictor; o->dc_t = s->dc_t; o->dc_lambda = s->dc_lambda;
  o->device_bytes = s->alloc_bytes + (int64_t)s->rec_cap;
  o->lowp_factorizations = s->lowp_count; o->fp64_fallbacks = s->fallback_count;
  o->refine_steps = s->refine_steps; o->pad = 0;
  o->schur_reg = s->reg_used;
  return 0;
}

int msdp_solver_trace(const msdp_solver* s, int32_t maxrec, double* out9, int32_t* nrec) {
  GUARD0(s);
  const int n = (int)std::min<size_t>(s->trace.size(), (size_t)std::max(maxrec, 0));
  for (int k = 0; k < n; ++k) std::memcpy(out9 + 9 * k, s->trace[k].data(), 9 * sizeof(double));
  *nrec = (int)s->trace.size();
  return 0;
}

int msdp_solver_update_residuals(msdp_solver* s) { GUARD(s); MSDP_TRY(s->finalize()); return s->update_residuals(); }
int msdp_solver_prepare(msdp_solver* s, int32_t* ok) {
  GUARD_READY(s);
  bool b = false;
  MSDP_TRY(s->prepare(b));
  *ok = b;
  return 0;
}
int msdp_solver_factorize_schur(msdp_solver* s, int32_t* ok) {
  GUARD_READY(s);
  bool b = false;
  MSDP_TRY(s->factorize_schur(b));
  *ok = b;
  return 0;
}
int msdp_solver_solve_directions(msdp_solver* s, double mu, int32_t corrected) {
  GUARD_READY(s);
  return s->solve_directions(mu, corrected != 0);
}
int msdp_solver_correction(msdp_solver* s) { GUARD_READY(s); return s->correction(); }
int msdp_solver_step_lengths(msdp_solver* s, double* a, double* bt) { GUARD_READY(s); return s->step_lengths(*a, *bt); }
int msdp_solver_affine_complementarity(msdp_solver* s, double a, double bt, double* out) {
  GUARD_READY(s);
  return s->affine_complementarity(a, bt, *out);
}
int msdp_solver_take_step(msdp_solver* s, double a, double bt) { GUARD_READY(s); return s->take_step(a, bt); }
int msdp_solver_schur_solve(msdp_solver* s, double* rhs) {
  GUARD_READY(s);
  MSDP_HIP(hipMemcpyAsync(s->wm, rhs, sizeof(double) * s->m, hipMemcpyHostToDevice, s->st));
  MSDP_TRY(s->hsolve(s->wm));
  MSDP_HIP(hipMemcpyAsync(rhs, s->wm, sizeof(double) * s->m, hipMemcpyDeviceToHost, s->st));
  MSDP_TRY(stream_sync(s->st));
  return 0;
}

static int arr(msdp_solver* s, int which, int block, double** p, size_t* count) {
  auto needb = [&]() { return block >= 0 && block < s->nb; };
  Blk* b = needb() ? &s->blk[block] : nullptr;
  const size_t n2 = b ? (size_t)b->n * b->n : 0;
  switch (which) {
    case MSDP_X: if (!b) return -2; *p = b->X; *count = n2; return 0;
    case MSDP_S: if (!b) return -2; *p = b->S; *count = n2; return 0;
    case MSDP_DELX: if (!b) return -2; *p = b->delX; *count = n2; return 0;
    case MSDP_DELS: if (!b) return -2; *p = b->delS; *count = n2; return 0;
    case MSDP_RD: if (!b) return -2; *p = b->Rd; *count = n2; return 0;
    case MSDP_W: if (!b) return -2; *p = b->W; *count = n2; return 0;
    case MSDP_G: if (!b) return -2; *p = b->G; *count = n2; return 0;
    case MSDP_SI: if (!b) return -2; *p = b->Si; *count = n2; return 0;
    case MSDP_CORR: if (!b) return -2; *p = b->corr; *count = n2; return 0;
    case MSDP_LS: if (!b) return -2; *p = b->Ls; *count = n2; return 0;
    case MSDP_LX: if (!b) return -2; *p = b->Lx; *count = n2; return 0;
    case MSDP_D: if (!b) return -2; *p = b->D; *count = b->n; return 0;
    case MSDP_Y: *p = s->y; *count = s->m; return 0;
    case MSDP_DELY: *p = s->dely; *count = s->m; return 0;
    case MSDP_RP: *p = s->Rp; *count = s->m; return 0;
    case MSDP_XLIN: *p = s->x; *count = s->ns; return 0;
    case MSDP_ZLIN: *p = s->z; *count = s->ns; return 0;
    case MSDP_DELXLIN: *p = s->delx; *count = s->ns; return 0;
    case MSDP_DELZLIN: *p = s->delz; *count = s->ns; return 0;
    case MSDP_RDLIN: *p = s->Rdlin; *count = s->ns; return 0;
    case MSDP_H: *p = s->H; *count = (size_t)s->m * s->m; return 0;
  }
  return -2;
}

int msdp_solver_get(const msdp_solver* cs, int32_t which, int32_t block, double* host) {
  GUARD0(cs);
  auto* s = const_cast<msdp_solver*>(cs);
  if (s->released) {  // the solution kept on the host
    const std::vector<double>* v = nullptr;
    if (which == MSDP_Y) v = &s->r_y;
    else if (which == MSDP_XLIN) v = &s->r_x;
    else if (which == MSDP_ZLIN) v = &s->r_z;
    else if ((which == MSDP_X || which == MSDP_S) && block >= 0 && block < s->nb)
      v = which == MSDP_X ? &s->r_X[block] : &s->r_S[block];
    if (!v) return as_err("solver released: only X, S, y, x and z are kept");
    std::memcpy(host, v->data(), sizeof(double) * v->size());
    return 0;
  }
  MSDP_TRY(s->finalize());
  double* p; size_t c;
  if (arr(s, which, block, &p, &c) != 0) return as_err("bad array selector or block");
  MSDP_HIP(hipMemcpyAsync(host, p, sizeof(double) * c, hipMemcpyDeviceToHost, s->st));
  MSDP_TRY(stream_sync(s->st));
  return 0;
}

int msdp_solver_set(msdp_solver* s, int32_t which, int32_t block, const double* host) {
  GUARD(s);
  MSDP_TRY(s->finalize());
  double* p; size_t c;
  if (arr(s, which, block, &p, &c) != 0) return as_err("bad array selector or block");
  MSDP_HIP(hipMemcpyAsync(p, host, sizeof(double) * c, hipMemcpyHostToDevice, s->st));
  MSDP_TRY(stream_sync(s->st));
  return 0;
}

// ---- kernel layer ----
int msdp_k_gemm(void* stream, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, double alpha,
                const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                int64_t ldc) {
  return gemm((hipStream_t)stream, ta != 0, tb != 0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}
int msdp_k_gemm_tri(void* stream, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K,
                    double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                    double beta, double* C, int64_t ldc, int32_t lower, int32_t tri) {
  if (tri < 0 || tri > 15) {
    set_error("msdp_k_gemm_tri: tri must be a combination of MSDP_TRI_* bits");
    return -2;
  }
  return gemm((hipStream_t)stream, ta != 0, tb != 0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
              1, 0, 0, 0, lower != 0, tri);
}
int msdp_k_gemm_noalloc(void* stream, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K,
                        double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                        double beta, double* C, int64_t ldc) {
  return gemm_noalloc((hipStream_t)stream, ta != 0, tb != 0, M, N, K, alpha, A, lda, B, ldb, beta,
                      C, ldc);
}
size_t msdp_k_gemm_noalloc_scratch_bytes(void) { return gemm_noalloc_scratch_bytes(); }
int msdp_k_syrk_lower(void* stream, int32_t trans, int32_t N, int32_t K, double alpha,
                      const double* A, int64_t lda, double beta, double* C, int64_t ldc) {
  // C = alpha X X' + beta C with X = op(A)' (N x K): trans -> X = A' (A is K x N)
  return gemm((hipStream_t)stream, trans != 0, trans == 0, N, N, K, alpha, A, lda, A, lda, beta, C,
              ldc, 1, 0, 0, 0, true);
}
int msdp_k_potrf(void* stream, int32_t n, double* A, int64_t lda, int32_t* info, double* work) {
  return potrf((hipStream_t)stream, n, A, lda, info, work);
}
size_t msdp_k_potrf_work(int32_t n) { return potrf_work_elems(n); }
size_t msdp_k_cholesky_factor_scratch(int32_t n) {
  return reduce_scratch_elems(0) + 64 + potrf_work_elems(n);
}
int msdp_k_cholesky_factor(void* stream, int32_t n, const double* M, int64_t ldm, double* L, int64_t ldl,
                           int32_t* info, double* scratch) {
  hipStream_t st = (hipStream_t)stream;
  if (n >= 1 && n <= 64) return cholesky_small(st, n, M, ldm, L, ldl, info);
  if (n > 64 && n <= 128) return cholesky_mid(st, n, M, ldm, L, ldl, info);
  MSDP_TRY(copy_mat(st, n, n, M, ldm, L, ldl));
  MSDP_TRY(symmetrize(st, n, L, ldl));
  MSDP_TRY(potrf(st, n, L, ldl, info, scratch + reduce_scratch_elems(0) + 64));
  return chol_flush(st, n, L, ldl, scratch);
}
size_t msdp_k_trtri_work(int32_t n) { return trtri_work_elems(n); }
int msdp_k_trtri(void* stream, int32_t n, const double* L, int64_t ldl, double* Li, int64_t ldi, double* work) {
  return trtri_lower((hipStream_t)stream, n, L, ldl, Li, ldi, work);
}
int msdp_k_sytrf(void* stream, int32_t n, double* A, int64_t lda, int32_t* ipiv, void* ctl,
                 double* work, int32_t* info) {
  return sytrf_bk((hipStream_t)stream, n, A, lda, ipiv, ctl, work, info);
}
size_t msdp_k_sytrf_ctl_bytes(void) { return sytrf_bk_ctl_bytes(); }
int msdp_k_sytrs(void* stream, int32_t n, const double* A, int64_t lda, const int32_t* ipiv,
                 double* b) {
  return sytrs_bk((hipStream_t)stream, n, A, lda, ipiv, b);
}
size_t msdp_k_syevd_work(int32_t n) { return syevd_work_elems(n); }
size_t msdp_k_syevd_iwork(int32_t n) { return syevd_iwork_elems(n); }
int msdp_k_syevd(void* stream, int32_t n, const double* A, int64_t lda, double* w, double* V, int64_t ldv,
                 int32_t vectors, double* work, int32_t* iwork) {
  return syevd((hipStream_t)stream, n, A, lda, w, V, ldv, vectors != 0, work, iwork);
}
int msdp_k_eig_min(void* stream, int32_t n, const double* A, int64_t lda, double* out, double* work) {
  return eig_min((hipStream_t)stream, n, A, lda, out, work, nullptr);
}
int msdp_k_eig_min2(void* stream, int32_t n, const double* A0, const double* A1, int64_t lda,
                    double* out0, double* out1, double* work) {
  return eig_min2((hipStream_t)stream, n, A0, A1, lda, out0, out1, work);
}
int msdp_k_gemm32(void* stream, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K,
                  float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                  double beta, double* C, int64_t ldc) {
  return gemm32_to64((hipStream_t)stream, ta != 0, tb != 0, M, N, K, alpha, A, lda, B, ldb, beta,
                     C, ldc, false);
}
int msdp_k_syrk_bf16x3(void* stream, int32_t N, int32_t K, double alpha, const double* A,
                       int64_t lda, double beta, double* C, int64_t ldc, void* work) {
  hipStream_t st = (hipStream_t)stream;
  if (N <= 0 || K <= 0) return 0;
  const int64_t ld = bf16x3_ld(K);
  __bf16* hi = reinterpret_cast<__bf16*>(work);
  __bf16* lo = hi + ld * N;
  MSDP_TRY(bf16_split(st, K, N, A, lda, hi, lo, ld));
  return syrk_bf16x3(st, N, (int)ld, alpha, hi, lo, ld, beta, C, ldc);
}
size_t msdp_k_syrk_bf16x3_work_bytes(int32_t N, int32_t K) {
  return N <= 0 || K <= 0 ? 0 : (size_t)(2 * bf16x3_ld(K) * N * 2);
}
int msdp_k_panel_apply(void* stream, int32_t rows, int32_t cols, const double* B, int64_t ldb,
                       const double* v, double* out, int32_t accumulate, double* work) {
  return panel_apply((hipStream_t)stream, rows, cols, B, ldb, v, out, accumulate != 0, work);
}
size_t msdp_k_panel_apply_work(int32_t rows, int32_t cols) {
  return rows <= 0 || cols <= 0 ? 0 : panel_apply_work_elems(rows, cols);
}
int msdp_k_schur_onehot(void* stream, int32_t m, const double* L, int64_t ldl, const double* R, int64_t ldr,
                        double* H, int64_t ldh) {
  return schur_onehot_identity((hipStream_t)stream, m, L, ldl, R, ldr, H, ldh);
}

}  // extern "C"
