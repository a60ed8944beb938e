// Dense FP64 factorisations on the GPU: thin QR (the project's one-workgroup
// Householder panel kernel up to 64 columns; wider, rocSOLVER Householder
// geqrf + orgqr/ungqr, the LAPACK semantics of Julia's `qr`, with complex
// factors checked and refactored unblocked where the blocked form fails), the
// Hermitian eigendecomposition (real n <= 1024: the project's one-launch
// Householder tridiagonalisation and reflector kernels around a divide and
// conquer tridiagonal solve, eigh.hip; complex: rocSOLVER heevd -- the
// algorithm class of Julia's `eigen(Hermitian)`), and the project's
// one-sided Jacobi kernel for the thin SVD (after a QR reduction of very
// tall inputs).
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "peps_internal.h"

namespace peps {

namespace {

void ck(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw Error(PS_ESOLVER, std::string(what) + " failed with rocblas status " + std::to_string(int(s)));
}

// device-side info word of a rocSOLVER call
struct DevInt {
  std::shared_ptr<Buffer> b;
  rocblas_int* p;
  explicit DevInt(Ctx& c) {
    b = std::make_shared<Buffer>();
    b->pool = c.pool.get();
    b->bytes = 256;
    b->ptr = c.pool->get(256);
    p = static_cast<rocblas_int*>(b->ptr);
  }
  int get(Ctx& c) {
    int h = 0;
    PEPS_HIP(hipMemcpyAsync(&h, p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PEPS_HIP(hipStreamSynchronize(c.stream));
    return h;
  }
};

// rocSOLVER is never handed NaN/Inf (a non-finite input is reported as a
// solver failure instead) nor an all-zero matrix (whose factorisations are
// written down directly).  Returns whether A has a non-zero entry.
bool check_input(Ctx& c, const Tensor& A, const char* what) {
  bool finite, nonzero;
  scan_values(c, A, &finite, &nonzero);
  PEPS_CHECK(finite, PS_ESOLVER, std::string("non-finite input to ") + what);
  return nonzero;
}

void check_int_range(int64_t m, int64_t n) {
  PEPS_CHECK(m < (1ll << 31) && n < (1ll << 31) && m * n < (1ll << 62), PS_EUNSUPPORTED,
             "factorisation: matrix too large");
}

// Q = the first k columns of the m-row factor W (column-major: a prefix).
Tensor q_prefix(Ctx& c, const Tensor& W, int64_t m, int64_t k) {
  if (W.dims[1] == k) return W.reshaped({m, k});
  Tensor Q = c.alloc({m, k}, W.cplx);
  if (Q.bytes()) PEPS_HIP(hipMemcpyAsync(Q.data(), W.data(), Q.bytes(), hipMemcpyDeviceToDevice, c.stream));
  return Q;
}

// Complex QR of a matrix wider than the panel kernel takes.  rocSOLVER's
// blocked zgeqrf returns non-finite values in every trailing column once a
// column is exactly zero when its reflector is formed (tau = 0), which
// charge-structured matrices produce all the time, while the unblocked
// zgeqr2 / zung2r factor the same matrices to rounding.  The blocked form runs
// first; a non-finite R or Q repeats the factorisation of A unblocked.
QR zqr_large(Ctx& c, const Tensor& A, bool want_q) {
  const int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n);
  QR r;
  for (int blocked = 1; blocked >= 0; --blocked) {
    Tensor W = copy_tensor(c, A);
    Tensor tau = c.alloc({std::max<int64_t>(k, 1)}, true);
    auto* w = (rocblas_double_complex*)W.data();
    auto* t = (rocblas_double_complex*)tau.data();
    if (k > 0) {
      if (blocked) ck(rocsolver_zgeqrf(c.blas, int(m), int(n), w, int(m), t), "zgeqrf");
      else ck(rocsolver_zgeqr2(c.blas, int(m), int(n), w, int(m), t), "zgeqr2");
    }
    r.R = upper_trapezoid(c, W, k);
    if (want_q) {
      if (k > 0) {
        if (blocked) ck(rocsolver_zungqr(c.blas, int(m), int(k), int(k), w, int(m), t), "zungqr");
        else ck(rocsolver_zung2r(c.blas, int(m), int(k), int(k), w, int(m), t), "zung2r");
      }
      r.Q = q_prefix(c, W, m, k);
    }
    if (want_q ? all_finite(c, {&r.R, &r.Q}) : all_finite(c, {&r.R})) return r;
    if (blocked) c.count("qr_unblocked", 0.0, 0);
  }
  PEPS_CHECK(false, PS_ESOLVER, "qr: non-finite factors of a finite " + std::to_string(m) + "x" +
                                    std::to_string(n) + " complex matrix");
  return r;
}

// Blocked Householder QR of a real matrix wider than one panel (LAPACK
// dgeqrf / dorgqr with the project's kernels): panels of 64 columns are
// factored by the one-workgroup panel kernel, their reflectors put in the
// compact WY form H = I - V T V^T (householder_v / householder_t), and the
// trailing columns updated as A2 - V (T^T (V^T A2)) with three GEMMs; Q is
// accumulated backwards from the first k columns of I the same way.  The
// same reflectors as the unblocked algorithm, so R and Q are LAPACK's up to
// rounding; about a dozen launches per panel where the library's blocked
// routine makes several hundred small ones per call.
QR dqr_blocked(Ctx& c, const Tensor& A, bool want_q) {
  const int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n), nb = qr_small_max_cols();
  Tensor W = copy_tensor(c, A);
  struct Panel {
    int64_t j0, b;
    Tensor V, T;
  };
  std::vector<Panel> panels;
  for (int64_t j0 = 0; j0 < k; j0 += nb) {
    const int64_t b = std::min(nb, k - j0), mp = m - j0, n2 = n - j0 - b;
    Tensor P = sub_block(c, W, j0, j0, mp, b);
    Tensor taup = c.alloc({b}, false);
    qr_panel(c, P, taup);
    put_block(c, W, j0, j0, P);
    Panel pn;
    pn.j0 = j0;
    pn.b = b;
    pn.V = householder_v(c, P);
    Tensor G = matmul(c, pn.V, pn.V, false, false, true, false);  // V^T V
    pn.T = householder_t(c, G, taup);
    if (n2 > 0) {
      Tensor A2 = sub_block(c, W, j0, j0 + b, mp, n2);
      Tensor X = matmul(c, pn.V, A2, false, false, true, false);  // V^T A2 (b x n2)
      Tensor Y = matmul(c, pn.T, X, false, false, true, false);   // T^T X
      scale(c, Y, -1.0);
      gemm(c, false, mp, n2, b, pn.V.data(), false, false, Y.data(), true, false, A2.data(), 1.0);
      put_block(c, W, j0, j0 + b, A2);
    }
    if (want_q) panels.push_back(std::move(pn));
  }
  QR r;
  r.R = upper_trapezoid(c, W, k);
  if (!want_q) return r;
  r.Q = eye(c, m, k, false);
  for (auto it = panels.rbegin(); it != panels.rend(); ++it) {
    const int64_t mp = m - it->j0, kq = k - it->j0;
    Tensor Q2 = sub_block(c, r.Q, it->j0, it->j0, mp, kq);
    Tensor X = matmul(c, it->V, Q2, false, false, true, false);  // V^T Q2
    Tensor Y = matmul(c, it->T, X);                             // T X
    scale(c, Y, -1.0);
    gemm(c, false, mp, kq, it->b, it->V.data(), false, false, Y.data(), true, false, Q2.data(), 1.0);
    put_block(c, r.Q, it->j0, it->j0, Q2);
  }
  return r;
}

}  // namespace

QR qr_thin(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "qr: matrix expected");
  int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n);
  check_int_range(m, n);
  if (!check_input(c, A, "qr")) {  // A = 0: Q = first k columns of I, R = 0
    QR r;
    r.Q = eye(c, m, k, A.cplx);
    r.R = c.zeros({k, n}, A.cplx);
    return r;
  }
  hipEvent_t ev = nullptr;
  const std::string cls = c.gemm_shapes ? "qr " + std::to_string(m) + "x" + std::to_string(n) : Ctx::bucket("qr", n);
  c.tbegin(cls, &ev);
  c.count(cls, (A.cplx ? 4.0 : 1.0) * (4.0 * m * n * k - 2.0 * (m + n) * k * k + 4.0 * k * k * k / 3.0), 0);
  QR r;
  if (n > qr_small_max_cols() && A.cplx) {
    r = zqr_large(c, A, true);
  } else if (n > qr_small_max_cols() && c.qr_backend == 0) {
    r = dqr_blocked(c, A, true);
  } else {
    Tensor W = copy_tensor(c, A);
    Tensor tau = c.alloc({std::max<int64_t>(k, 1)}, A.cplx);
    if (n <= qr_small_max_cols()) {  // narrow panel: one-workgroup Householder kernel
      if (k > 0) qr_panel(c, W, tau);
      r.R = upper_trapezoid(c, W, k);
      r.Q = c.alloc({m, k}, A.cplx);
      if (k > 0) qr_form_q(c, W, tau, r.Q);
    } else {
      if (k > 0) ck(rocsolver_dgeqrf(c.blas, int(m), int(n), W.data(), int(m), tau.data()), "dgeqrf");
      r.R = upper_trapezoid(c, W, k);
      if (k > 0) ck(rocsolver_dorgqr(c.blas, int(m), int(k), int(k), W.data(), int(m), tau.data()), "dorgqr");
      r.Q = q_prefix(c, W, m, k);
    }
  }
  c.tend(cls, ev);
  return r;
}

Tensor qr_r_only(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "qr: matrix expected");
  int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n);
  check_int_range(m, n);
  if (!check_input(c, A, "qr")) return c.zeros({k, n}, A.cplx);
  hipEvent_t ev = nullptr;
  const std::string cls = c.gemm_shapes ? "qr " + std::to_string(m) + "x" + std::to_string(n) : Ctx::bucket("qr", n);
  c.tbegin(cls, &ev);
  c.count(cls, (A.cplx ? 4.0 : 1.0) * (2.0 * m * n * k - (m + n) * k * k + 2.0 * k * k * k / 3.0), 0);
  Tensor R;
  if (n > qr_small_max_cols() && A.cplx) {
    R = zqr_large(c, A, false).R;
  } else {
    if (n > qr_small_max_cols() && c.qr_backend == 0) {
      R = dqr_blocked(c, A, false).R;
    } else {
      Tensor W = copy_tensor(c, A);
      Tensor tau = c.alloc({std::max<int64_t>(k, 1)}, A.cplx);
      if (k > 0) {
        if (n <= qr_small_max_cols()) qr_panel(c, W, tau);
        else ck(rocsolver_dgeqrf(c.blas, int(m), int(n), W.data(), int(m), tau.data()), "dgeqrf");
      }
      R = upper_trapezoid(c, W, k);
    }
  }
  c.tend(cls, ev);
  return R;
}

Tensor conj_transpose(Ctx& c, const Tensor& A) {
  Tensor T = permute(c, A, {1, 0});
  if (T.data() == A.data()) T = copy_tensor(c, T);  // vector-shaped: permute returned a view
  conj_inplace(c, T);
  return T;
}

namespace {
// The matrix the Jacobi kernel diagonalises for one dense SVD: B = A or A^H
// (tall), reduced to R by a Householder QR where that shrinks the work.
struct DensePlan {
  bool tr = false, useq = false;
  int64_t mb = 0, nb = 0;
  Tensor W, Q, V, sg;
  double tol = 0.0;
};

// One thin SVD: either answered directly (empty or zero input) or a dense
// plan on the compacted block (rows / columns that are exactly zero dropped).
struct ThinPlan {
  bool direct = false;
  SVD out;
  int64_t m = 0, n = 0, k = 0;
  bool compact = false;
  std::vector<int64_t> nzr, nzc;
  DensePlan d;
};

DensePlan dense_prepare(Ctx& c, const Tensor& A) {
  // A has no zero row or column
  TierGuard fp64(c, TIER_FP64);
  DensePlan p;
  const int64_t m = A.dims[0], n = A.dims[1];
  p.tr = m < n;
  Tensor B = p.tr ? conj_transpose(c, A) : A;
  p.mb = B.dims[0];
  p.nb = B.dims[1];
  // a Householder QR first only where it shrinks the Jacobi work markedly
  p.useq = p.mb > 2 * p.nb && p.mb * p.nb > 65536;
  if (p.useq) {
    QR f = qr_thin(c, B);
    p.Q = f.Q;
    p.W = f.R;
  } else {
    p.W = copy_tensor(c, B);
  }
  p.V = c.alloc({p.nb, p.nb}, A.cplx);
  p.sg = c.alloc({p.nb}, false);
  // convergence threshold of LAPACK's dgesvj, sqrt(m) eps on the cosine of a
  // column pair (a fixed 1e-15 lies below the rounding floor for m > ~80 and
  // keeps pairs rotating until the sweep limit)
  p.tol = std::max(1e-15, std::sqrt(double(p.W.dims[0])) * 1.1102230246251565e-16);
  return p;
}

SVD dense_finish(Ctx& c, DensePlan& p, int sweeps) {
  TierGuard fp64(c, TIER_FP64);
  const int64_t mb = p.mb, nb = p.nb;
  c.count("svd[sweeps]", double(sweeps), 0);
  if (sweeps >= c.jacobi_max_sweeps) {
    // the sweep cap was reached with rotations still applied: the factors
    // are returned as they are and the condition is counted as a warning
    c.count("svd[unconverged]", 1.0, 0);
    ++c.warnings;
  }
  std::vector<double> s = c.to_host(p.sg);
  std::vector<int64_t> ord(nb);
  for (int64_t i = 0; i < nb; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return s[x] > s[y]; });
  std::vector<double> inv(nb);
  for (int64_t i = 0; i < nb; ++i) inv[i] = s[i] > 1e-140 ? 1.0 / s[i] : 0.0;  // columns below 1e-140 are left unrotated: zero
  scale_cols(c, p.W, inv);
  Tensor Ub = p.useq ? matmul(c, p.Q, p.W) : p.W;
  std::vector<int64_t> rows_u(mb), rows_v(nb);
  for (int64_t i = 0; i < mb; ++i) rows_u[i] = i;
  for (int64_t i = 0; i < nb; ++i) rows_v[i] = i;
  Ub = gather2d(c, Ub, rows_u, ord);
  Tensor Vb = gather2d(c, p.V, rows_v, ord);
  SVD r;
  r.S.resize(nb);
  for (int64_t i = 0; i < nb; ++i) r.S[i] = s[ord[i]];
  if (!p.tr) {
    r.U = Ub;
    r.Vt = conj_transpose(c, Vb);
  } else {
    r.U = Vb;
    r.Vt = conj_transpose(c, Ub);
  }
  return r;
}

ThinPlan thin_prepare(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "svd: matrix expected");
  ThinPlan t;
  t.m = A.dims[0];
  t.n = A.dims[1];
  t.k = std::min(t.m, t.n);
  check_int_range(t.m, t.n);
  if (t.k == 0) {
    t.direct = true;
    t.out.U = c.alloc({t.m, 0}, A.cplx);
    t.out.Vt = c.alloc({0, t.n}, A.cplx);
    return t;
  }
  if (!check_input(c, A, "svd")) {  // A = 0: S = 0 with identity vectors
    t.direct = true;
    t.out.U = eye(c, t.m, t.k, A.cplx);
    t.out.S.assign(t.k, 0.0);
    t.out.Vt = conj_transpose(c, eye(c, t.n, t.k, A.cplx));
    return t;
  }
  // rows and columns that are exactly zero (Gauss-masked blocks have many)
  // do not change the non-zero singular triplets: factorise the compacted
  // block and pad with zero triplets
  nonzero_lines(c, A, &t.nzr, &t.nzc);
  t.compact = int64_t(t.nzr.size()) < t.m || int64_t(t.nzc.size()) < t.n;
  t.d = dense_prepare(c, t.compact ? gather2d(c, A, t.nzr, t.nzc) : A);
  return t;
}

SVD thin_finish(Ctx& c, ThinPlan& t, int sweeps) {
  if (t.direct) return t.out;
  SVD f = dense_finish(c, t.d, sweeps);
  if (!t.compact) return f;
  const bool cp = f.U.cplx;
  const int64_t kc = int64_t(f.S.size());
  std::vector<int64_t> jk(kc);
  for (int64_t i = 0; i < kc; ++i) jk[i] = i;
  SVD r;
  r.U = c.zeros({t.m, t.k}, cp);
  r.Vt = c.zeros({t.k, t.n}, cp);
  scatter_cols(c, r.U, f.U, t.nzr, jk, jk);
  scatter_rows(c, r.Vt, f.Vt, jk, jk, t.nzc);
  r.S = f.S;
  r.S.resize(t.k, 0.0);
  return r;
}
}  // namespace

// Thin SVD A = U diag(S) Vt with S descending.  Wide inputs are handled
// through their conjugate transpose; inputs much taller than wide are first
// reduced by a Householder QR (rocSOLVER geqrf/orgqr); the one-sided Jacobi
// kernel (kernels.hip) diagonalises what remains.  Left vectors of zero
// singular values are zero, except for A = 0 (identity vectors).
SVD svd_thin(Ctx& c, const Tensor& A) {
  ThinPlan t = thin_prepare(c, A);
  int sweeps = 0;
  if (!t.direct) sweeps = jacobi_svd(c, t.d.W, t.d.V, t.d.sg, t.d.tol, c.jacobi_max_sweeps);
  return thin_finish(c, t, sweeps);
}

// svd_thin of several independent matrices; their Jacobi diagonalisations
// run concurrently (one workgroup each).  Same results as one by one.
std::vector<SVD> svd_thin_batch(Ctx& c, const std::vector<Tensor>& As) {
  std::vector<ThinPlan> plans;
  plans.reserve(As.size());
  for (const Tensor& A : As) plans.push_back(thin_prepare(c, A));
  std::vector<JacobiJob> jobs;
  std::vector<size_t> of;
  for (size_t i = 0; i < plans.size(); ++i)
    if (!plans[i].direct) {
      jobs.push_back({&plans[i].d.W, &plans[i].d.V, &plans[i].d.sg, plans[i].d.tol, 0});
      of.push_back(i);
    }
  std::vector<int> sweeps(plans.size(), 0);
  if (!jobs.empty()) {
    TierGuard fp64(c, TIER_FP64);
    jacobi_svd_batch(c, jobs, c.jacobi_max_sweeps);
    for (size_t j = 0; j < jobs.size(); ++j) sweeps[of[j]] = jobs[j].sweeps;
  }
  std::vector<SVD> out;
  out.reserve(plans.size());
  for (size_t i = 0; i < plans.size(); ++i) out.push_back(thin_finish(c, plans[i], sweeps[i]));
  return out;
}

EIG eigh(Ctx& c, const Tensor& A) {
  // Hermitian eigendecomposition, eigenvalues ascending; only the lower
  // triangle of A is read.
  PEPS_CHECK(A.rank() == 2 && A.dims[0] == A.dims[1], PS_ESHAPE, "eigh: square matrix expected");
  const int64_t n = A.dims[0];
  check_int_range(n, n);
  EIG r;
  if (n == 0) {
    r.V = c.alloc({0, 0}, A.cplx);
    return r;
  }
  if (!check_input(c, A, "eigh")) {  // A = 0
    r.V = eye(c, n, n, A.cplx);
    r.w.assign(n, 0.0);
    return r;
  }
  if (!A.cplx && n >= 2 && n <= sym_tridiag_max_n()) {
    // the project's one-launch tridiagonal reduction, divide and conquer on
    // the tridiagonal matrix, the reflectors applied to its eigenvectors
    Tridiag t = tridiagonalize(c, A);
    tridiag_eig(c, t, &r.w, &r.V);
    apply_q(c, t, r.V, false);
    return r;
  }
  r.V = copy_tensor(c, A);
  Tensor w = c.alloc({n}, false), e = c.alloc({n}, false);
  DevInt info(c);
  hipEvent_t ev = nullptr;
  const std::string cls = c.gemm_shapes ? "eigh " + std::to_string(n) : Ctx::bucket("eigh", n);
  c.tbegin(cls, &ev);
  c.count(cls, (A.cplx ? 4.0 : 1.0) * 9.0 * double(n) * n * n, 0);
  if (A.cplx)
    ck(rocsolver_zheevd(c.blas, rocblas_evect_original, rocblas_fill_lower, int(n),
                        (rocblas_double_complex*)r.V.data(), int(n), w.data(), e.data(), info.p),
       "zheevd");
  else
    ck(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_lower, int(n), r.V.data(),
                        int(n), w.data(), e.data(), info.p),
       "dsyevd");
  c.tend(cls, ev);
  const int inf = info.get(c);
  PEPS_CHECK(inf == 0, PS_ESOLVER, "syevd did not converge (info=" + std::to_string(inf) + ")");
  r.w = c.to_host(w);
  return r;
}

}  // namespace peps
