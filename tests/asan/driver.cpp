// Host-AddressSanitizer exercise of the library's non-kernel code through its
// C ABI: solver set-up and teardown (single and multi-block, one-hot, COO and
// dense constraints, scalars), every component combination's driver loop,
// the serving loop's worker threads with concurrent solves and their release,
// and the API's error paths.  Built with the compiler's -fsanitize=address
// against libmadsdp_hip_asan.so (madsdp_amd.build.build_asan_driver); any
// heap misuse aborts with an ASan report.  Exit status 0 = clean.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "madsdp_hip.h"

namespace {

struct Lcg {
  uint64_t s;
  explicit Lcg(uint64_t seed) : s(seed * 6364136223846793005ULL + 1442695040888963407ULL) {}
  double uniform() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  double normal() {
    const double u = uniform() + 1e-300, v = uniform();
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
  }
};

int fail(const char* what, int rc) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, msdp_last_error());
  return 1;
}

#define CHECK(call)                        \
  do {                                     \
    const int rc_ = (call);                \
    if (rc_ < 0) return fail(#call, rc_);  \
  } while (0)

// max-cut over a random graph: one n x n block, one-hot constraints
int make_maxcut(msdp_solver** out, int n, uint64_t seed) {
  Lcg g(seed);
  std::vector<double> L((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (g.uniform() < 0.2) {
        L[i + (size_t)j * n] -= 0.25;
        L[j + (size_t)i * n] -= 0.25;
        L[i + (size_t)i * n] += 0.25;
        L[j + (size_t)j * n] += 0.25;
      }
  for (auto& v : L) v = -v;  // C = -L/4
  const int64_t dims[1] = {n};
  CHECK(msdp_solver_create(out, 0, 1, dims, n, 0));
  CHECK(msdp_solver_set_cost_block(*out, 0, L.data()));
  std::vector<int64_t> idx(n);
  std::vector<double> one(n, 1.0);
  for (int i = 0; i < n; ++i) idx[i] = i;
  CHECK(msdp_solver_set_constraints_coo(*out, 0, n, idx.data(), idx.data(), idx.data(), one.data()));
  CHECK(msdp_solver_set_rhs(*out, one.data()));
  return 0;
}

// blocks with dense symmetric constraints touching each, scalars; X = I, x = 1
// feasible and y0 dual feasible (C = sum y0_j A_j + I, d = C_lin' y0 + 1)
int make_dense(msdp_solver** out, const std::vector<int>& sizes, int m, int ns, uint64_t seed) {
  Lcg g(seed);
  std::vector<int64_t> dims(sizes.begin(), sizes.end());
  CHECK(msdp_solver_create(out, 0, (int)sizes.size(), dims.data(), m, ns));
  std::vector<double> y0(m), b(m, 0.0), Clin((size_t)m * ns), d(ns, 1.0);
  for (auto& v : y0) v = g.normal();
  for (size_t k = 0; k < sizes.size(); ++k) {
    const int n = sizes[k];
    std::vector<int64_t> con;
    for (int j = 0; j < m; ++j)
      if ((j + (int)k) % 2 == 0 || sizes.size() == 1) con.push_back(j);
    std::vector<double> mats(con.size() * n * n), C((size_t)n * n, 0.0);
    for (size_t c = 0; c < con.size(); ++c) {
      double* A = mats.data() + c * n * n;
      for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) A[i + (size_t)j * n] = A[j + (size_t)i * n] = g.normal();
      for (int i = 0; i < n; ++i) b[con[c]] += A[i + (size_t)i * n];
      for (int e = 0; e < n * n; ++e) C[e] += y0[con[c]] * A[e];
    }
    for (int i = 0; i < n; ++i) C[i + (size_t)i * n] += 1.0;
    CHECK(msdp_solver_set_cost_block(*out, (int)k, C.data()));
    CHECK(msdp_solver_set_constraints_dense(*out, (int)k, (int64_t)con.size(), con.data(), mats.data()));
  }
  for (int l = 0; l < ns; ++l)
    for (int j = 0; j < m; ++j) {
      const double v = g.uniform();
      Clin[j + (size_t)l * m] = v;
      b[j] += v;
      d[l] += v * y0[j];
    }
  CHECK(msdp_solver_set_rhs(*out, b.data()));
  if (ns > 0) CHECK(msdp_solver_set_scalar_data(*out, d.data(), Clin.data()));
  return 0;
}

int solve_one(msdp_solver* s, int direction, int step, int centering, int schur) {
  msdp_options o;
  msdp_default_options(&o);
  o.direction = direction;
  o.step = step;
  o.centering = centering;
  o.schur = schur;
  CHECK(msdp_solver_configure(s, &o));
  CHECK(msdp_solver_solve(s));
  msdp_info inf;
  CHECK(msdp_solver_get_info(s, &inf));
  std::vector<double> tr(9 * 200);
  int32_t nrec = 0;
  CHECK(msdp_solver_trace(s, 200, tr.data(), &nrec));
  if (inf.status != 1) {
    std::fprintf(stderr, "status %d after %d iterations\n", inf.status, inf.iter);
    return 1;
  }
  return 0;
}

}  // namespace

int main() {
  int bad = 0;
  // single solves over the component combinations and Schur solvers
  for (int d = 0; d < 2; ++d)
    for (int st = 0; st < 2; ++st)
      for (int c = 0; c < 2; ++c) {
        msdp_solver* s = nullptr;
        if (make_maxcut(&s, 40, 7 + d + 2 * st + 4 * c)) return 1;
        bad |= solve_one(s, d, st, c, 0);
        CHECK(msdp_solver_destroy(s));
      }
  for (int schur = 0; schur < 3; ++schur) {
    msdp_solver* s = nullptr;
    if (make_dense(&s, {10}, 30, 4, 11 + schur)) return 1;
    bad |= solve_one(s, 0, 0, 0, schur);
    CHECK(msdp_solver_release(s));  // solution kept on the host
    std::vector<double> y(30);
    CHECK(msdp_solver_get(s, MSDP_Y, -1, y.data()));
    CHECK(msdp_solver_destroy(s));
  }
  {  // multi-block: the recorded, batched per-block stages
    msdp_solver* s = nullptr;
    if (make_dense(&s, {5, 17, 9, 30, 64, 70}, 60, 6, 21)) return 1;
    bad |= solve_one(s, 0, 0, 0, 0);
    CHECK(msdp_solver_destroy(s));
  }
  // the serving loop: 24 requests on 6 slots, arrivals spread over 50 ms
  msdp_server* srv = nullptr;
  CHECK(msdp_server_create(&srv, 0, 6, (int64_t)16 << 30));
  std::vector<msdp_solver*> reqs;
  for (int k = 0; k < 24; ++k) {
    msdp_solver* s = nullptr;
    const int rc = (k % 3 == 0) ? make_maxcut(&s, 30 + k, 100 + k)
                 : (k % 3 == 1) ? make_dense(&s, {8 + k % 5}, 20 + k, 3, 200 + k)
                                : make_dense(&s, {4, 9, 13}, 24, 2, 300 + k);
    if (rc) return 1;
    msdp_options o;
    msdp_default_options(&o);
    CHECK(msdp_solver_configure(s, &o));
    int64_t id = 0;
    CHECK(msdp_server_submit(srv, s, 0.002 * k, &id));
    reqs.push_back(s);
  }
  CHECK(msdp_server_start(srv));
  CHECK(msdp_server_wait_all(srv));
  for (size_t k = 0; k < reqs.size(); ++k) {
    msdp_request_info ri;
    CHECK(msdp_server_request(srv, (int64_t)k, &ri));
    msdp_info inf;
    CHECK(msdp_solver_get_info(reqs[k], &inf));
    if (ri.rc != 0 || inf.status != 1) {
      std::fprintf(stderr, "request %zu: rc %d status %d\n", k, ri.rc, inf.status);
      bad = 1;
    }
  }
  CHECK(msdp_server_destroy(srv));
  for (auto* s : reqs) CHECK(msdp_solver_destroy(s));
  // error paths: they must fail cleanly with a message
  msdp_solver* s = nullptr;
  const int64_t dims[1] = {-3};
  if (msdp_solver_create(&s, 0, 1, dims, 2, 0) >= 0) bad = 1;
  if (msdp_solver_solve(nullptr) >= 0) bad = 1;
  if (!msdp_last_error() || !msdp_last_error()[0]) bad = 1;
  std::printf(bad ? "ASAN DRIVER: wrong results\n" : "ASAN DRIVER OK\n");
  std::fflush(stdout);
  // skip the HIP runtime's static teardown: under ASan it trips the
  // sanitizer's device-allocator check after our code has finished
  std::_Exit(bad);
}
