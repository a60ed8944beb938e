// Two-site DMRG of a matrix-product operator on the GPU (mps_lgt.jl:333-467):
// environments, the two-site effective Hamiltonian, a Lanczos solve of the
// local ground state and the truncated SVD split, all on device tensors.
// Real MPOs and MPS only: the U(1) lattice-gauge Hamiltonian is real
// symmetric, so its ground state is real up to a global phase.
//
// Layouts follow the reference: MPS tensors psi[p] (Dl, Dr, d), MPO tensors
// W[p] (Dl, Dr, k, b) with k the bra and b the ket physical index, all
// column-major.  Environments are (bra bond, MPO bond, ket bond).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "engine.h"
#include "peps_internal.h"

namespace peps {

namespace {

std::vector<int64_t> seq_range(int64_t n) {
  std::vector<int64_t> v(n);
  std::iota(v.begin(), v.end(), 0);
  return v;
}

Tensor ones111(Ctx& c) {
  const double one = 1.0;
  return c.upload({1, 1, 1}, false, &one);
}

// left_env (mps_lgt.jl:363-368): L[a,b,c] = psi[a',a,k] W[b',b,k,q] psi[c',c,q] L'[a',b',c']
Tensor left_env(Ctx& c, const Tensor& Lp, const Tensor& W, const Tensor& psi) {
  Tensor t1 = contract(c, Lp, "xyz", psi, "zcq", "xycq");
  Tensor t2 = contract(c, t1, "xycq", W, "ybkq", "xcbk");
  return contract(c, t2, "xcbk", psi, "xak", "abc");
}

// right_env (mps_lgt.jl:355-361): R[a,b,c] = psi[a,a',k] W[b,b',k,q] psi[c,c',q] R'[a',b',c']
Tensor right_env(Ctx& c, const Tensor& Rn, const Tensor& W, const Tensor& psi) {
  Tensor t1 = contract(c, Rn, "xyz", psi, "czq", "xycq");
  Tensor t2 = contract(c, t1, "xycq", W, "bykq", "xcbk");
  return contract(c, t2, "xcbk", psi, "axk", "abc");
}

// apply_Heff2 (mps_lgt.jl:371-377): the two-site effective Hamiltonian on
// Theta[a, k1, k2, c].
Tensor apply_heff2(Ctx& c, const Tensor& L, const Tensor& W1, const Tensor& W2, const Tensor& R,
                   const Tensor& th) {
  Tensor t1 = contract(c, L, "abz", th, "zqrw", "abqrw");
  Tensor t2 = contract(c, t1, "abqrw", W1, "bmxq", "amxrw");
  Tensor t3 = contract(c, t2, "amxrw", W2, "mpyr", "axypw");
  return contract(c, t3, "axypw", R, "cpw", "axyc");
}

// Eigenpairs of a small real symmetric matrix (row-major n x n, n <= 64) by
// cyclic Jacobi rotations, eigenvalues ascending with their vectors as the
// columns of V (row-major).  The host side of the Lanczos projection.
void small_sym_eig(int n, std::vector<double> A, std::vector<double>* w, std::vector<double>* V) {
  std::vector<double> Q(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) Q[size_t(i) * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double x = A[size_t(i) * n + j] * A[size_t(i) * n + j];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[size_t(p) * n + q];
        if (apq == 0.0) continue;
        const double app = A[size_t(p) * n + p], aqq = A[size_t(q) * n + q];
        const double theta = 0.5 * (aqq - app) / apq;
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {  // A <- J^T A J
          const double akp = A[size_t(k) * n + p], akq = A[size_t(k) * n + q];
          A[size_t(k) * n + p] = cs * akp - sn * akq;
          A[size_t(k) * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[size_t(p) * n + k], aqk = A[size_t(q) * n + k];
          A[size_t(p) * n + k] = cs * apk - sn * aqk;
          A[size_t(q) * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double qkp = Q[size_t(k) * n + p], qkq = Q[size_t(k) * n + q];
          Q[size_t(k) * n + p] = cs * qkp - sn * qkq;
          Q[size_t(k) * n + q] = sn * qkp + cs * qkq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return A[size_t(x) * n + x] < A[size_t(y) * n + y]; });
  w->resize(n);
  V->assign(size_t(n) * n, 0.0);
  for (int j = 0; j < n; ++j) {
    (*w)[j] = A[size_t(ord[j]) * n + ord[j]];
    for (int i = 0; i < n; ++i) (*V)[size_t(i) * n + j] = Q[size_t(i) * n + ord[j]];
  }
}

// Lowest eigenpair of the Hermitian operator `apply` from the start vector
// x0 (KrylovKit.eigsolve(..., 1, :SR) in two_site_update!): a Krylov space
// of up to `krylovdim` vectors with thick restarts in the Krylov-Schur
// manner -- the lowest half of the Ritz vectors is kept, with their images,
// and the space is extended from the lowest pair's residual -- until that
// residual ||H x - theta x|| is at most tol (or maxiter restarts).  The basis
// V and its image HV are device matrices: the projection V^T HV, the two
// Gram-Schmidt passes and the Ritz combinations are GEMMs, and a new vector
// reads back one number (its norm).
template <class F>
double lanczos_lowest(Ctx& c, F&& apply, const Tensor& x0, double tol, int maxiter, int krylovdim, Tensor* xout) {
  const int64_t N = x0.numel();
  const int kmax = int(std::max<int64_t>(2, std::min<int64_t>(krylovdim, N)));
  const int nkeep = std::max(1, kmax / 2);
  Tensor V = c.zeros({N, kmax}, false), HV = c.zeros({N, kmax}, false);
  auto col = [&](Tensor& M, int j) { return M.slice_last(j, j + 1); };  // N x 1 view
  auto put = [&](Tensor& M, int j, const Tensor& v) {
    PEPS_HIP(hipMemcpyAsync(col(M, j).data(), v.data(), size_t(N) * 8, hipMemcpyDeviceToDevice, c.stream));
  };
  Tensor x = copy_tensor(c, x0).reshaped({N, 1});
  {
    const double nx = std::sqrt(norm2sq(c, x));
    if (!(nx > 0.0) || !std::isfinite(nx)) fill_random(c, x, 0x6c616e637a6f73ull);
    scale(c, x, 1.0 / std::sqrt(norm2sq(c, x)));
  }
  put(V, 0, x);
  put(HV, 0, apply(x.reshaped(x0.dims)).reshaped({N, 1}));
  int m = 1;
  Tensor r = copy_tensor(c, col(HV, 0));  // next direction before orthogonalisation
  double theta = 0.0;
  Tensor y = x;
  int applied = 1;
  for (int restart = 0; restart < std::max(1, maxiter); ++restart) {
    while (m < kmax) {
      Tensor Vm = V.slice_last(0, m);
      for (int pass = 0; pass < 2; ++pass)
        axpy_inplace(c, r, matmul(c, Vm, matmul(c, Vm, r, false, false, true, false)), -1.0);
      const double b = std::sqrt(norm2sq(c, r));
      if (!(b > 1e-14) || !std::isfinite(b)) break;  // invariant subspace
      scale(c, r, 1.0 / b);
      put(V, m, r);
      Tensor hr = apply(r.reshaped(x0.dims)).reshaped({N, 1});
      ++applied;
      put(HV, m, hr);
      r = copy_tensor(c, hr);
      ++m;
    }
    Tensor Vm = V.slice_last(0, m), HVm = HV.slice_last(0, m);
    std::vector<double> T = c.to_host(matmul(c, Vm, HVm, false, false, true, false));  // column-major m x m
    std::vector<double> Ts(size_t(m) * m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) Ts[size_t(i) * m + j] = 0.5 * (T[size_t(i) + size_t(j) * m] + T[size_t(j) + size_t(i) * m]);
    std::vector<double> tw, tv;
    small_sym_eig(m, Ts, &tw, &tv);
    theta = tw[0];
    const int keep = std::min(nkeep, m);
    std::vector<double> Sk(size_t(m) * keep);  // column-major m x keep
    for (int j = 0; j < keep; ++j)
      for (int i = 0; i < m; ++i) Sk[size_t(i) + size_t(j) * m] = tv[size_t(i) * m + j];
    Tensor S = c.upload({m, keep}, false, Sk.data());
    Tensor Y = matmul(c, Vm, S), HY = matmul(c, HVm, S);  // N x keep
    y = copy_tensor(c, col(Y, 0));
    Tensor res = copy_tensor(c, col(HY, 0));
    axpy_inplace(c, res, y, -theta);
    const double rn = std::sqrt(norm2sq(c, res));
    c.count("dmrg[lanczos restarts]", 1.0, 0);
    if (rn <= tol || m < kmax) break;  // converged, or the space is invariant
    if (restart == std::max(1, maxiter) - 1) {
      c.count("dmrg[lanczos unconverged]", 1.0, 0);
      break;
    }
    PEPS_HIP(hipMemcpyAsync(V.data(), Y.data(), size_t(N) * keep * 8, hipMemcpyDeviceToDevice, c.stream));
    PEPS_HIP(hipMemcpyAsync(HV.data(), HY.data(), size_t(N) * keep * 8, hipMemcpyDeviceToDevice, c.stream));
    m = keep;
    r = res;
  }
  c.count("dmrg[lanczos applications]", double(applied), 0);
  scale(c, y, 1.0 / std::sqrt(norm2sq(c, y)));
  *xout = y.reshaped(x0.dims);
  return theta;
}

// Smallest block dimension of the two-site split that takes the
// Jordan-Wielandt form instead of the one-workgroup Jacobi SVD.
constexpr int64_t kDmrgJwMinDim = 64;

// two_site_update! (mps_lgt.jl:384-416) at bond (p, p+1).
double two_site_update(Ctx& c, std::vector<Tensor>& psi, const std::vector<Tensor>& W, std::vector<Tensor>& L,
                       std::vector<Tensor>& R, size_t p, int D, bool to_right, double tol, int maxiter,
                       int krylovdim) {
  const int64_t a = psi[p].dims[0], k1 = psi[p].dims[2], cr = psi[p + 1].dims[1], k2 = psi[p + 1].dims[2];
  Tensor th = contract(c, psi[p], "amx", psi[p + 1], "mcy", "axyc");
  auto heff = [&](const Tensor& v) { return apply_heff2(c, L[p], W[p], W[p + 1], R[p + 1], v); };
  Tensor g;
  const double E = lanczos_lowest(c, heff, th, tol, maxiter, krylovdim, &g);
  // the split M = U S V^T of Theta (a k1) x (k2 c), keeping at most D
  // values above 1e-12 s_0: large blocks through the Jordan-Wielandt form
  // (the side that becomes canonical orthonormalised, the other carrying S),
  // small ones through the Jacobi SVD
  const Tensor Mt = g.reshaped({a * k1, k2 * cr});
  Tensor U, Vt;
  if (std::min(a * k1, k2 * cr) >= kDmrgJwMinDim) {
    Tensor Q, P;
    if (to_right) {  // U orthonormal, S V^T = (M^T U)^T
      topk_right_jw(c, permute(c, Mt, {1, 0}), D, 1e-12, true, &Q, &P);
      U = Q;
      Vt = copy_tensor(c, permute(c, P, {1, 0}));
    } else {         // V orthonormal, U S = M V
      topk_right_jw(c, Mt, D, 1e-12, true, &Q, &P);
      U = P;
      Vt = copy_tensor(c, permute(c, Q, {1, 0}));
    }
  } else {
    SVD f = svd_thin(c, Mt);
    int64_t cnt = 0;
    for (double s : f.S) cnt += s > 1e-12 * f.S[0];
    const int64_t k = std::max<int64_t>(1, std::min<int64_t>({int64_t(D), cnt, int64_t(f.S.size())}));
    U = copy_tensor(c, f.U.slice_last(0, k));                       // (a k1) x k
    Vt = gather2d(c, f.Vt, seq_range(k), seq_range(k2 * cr));     // k x (k2 c)
    std::vector<double> S(f.S.begin(), f.S.begin() + k);
    if (to_right) scale_rows(c, Vt, S);
    else scale_cols(c, U, S);
  }
  const int64_t keep = U.dims[1];
  psi[p] = permute(c, U.reshaped({a, k1, keep}), {0, 2, 1});
  psi[p + 1] = permute(c, Vt.reshaped({keep, k2, cr}), {0, 2, 1});
  if (to_right) L[p + 1] = left_env(c, L[p], W[p], psi[p]);
  else R[p] = right_env(c, R[p + 1], W[p + 1], psi[p + 1]);
  return E;
}

}  // namespace

double mpo_expect(Ctx& c, const std::vector<Tensor>& W, const std::vector<Tensor>& psi) {
  // mpo_expect (mps_lgt.jl:308-315): <psi|W|psi> by absorbing sites into the
  // (bra, mpo, ket) sandwich from the left
  PEPS_CHECK(W.size() == psi.size() && !W.empty(), PS_ESHAPE, "mpo_expect: lengths");
  Tensor E = ones111(c);
  for (size_t p = 0; p < W.size(); ++p) E = left_env(c, E, W[p], psi[p]);
  PEPS_CHECK(E.numel() == 1, PS_ESHAPE, "mpo_expect: open boundary bonds expected");
  return c.to_host(E)[0];
}

DmrgResult dmrg_ground_state(Ctx& c, const std::vector<Tensor>& W, std::vector<Tensor> psi, int D, int nsweeps,
                             double tol, int maxiter, int krylovdim, double etol) {
  // dmrg_ground_state (mps_lgt.jl:418-467)
  TierGuard fp64(c, TIER_FP64);
  const size_t n = W.size();
  PEPS_CHECK(n >= 2 && psi.size() == n && D >= 1, PS_EARG, "dmrg: at least two sites, D >= 1");
  for (size_t p = 0; p < n; ++p) {
    PEPS_CHECK(W[p].rank() == 4 && psi[p].rank() == 3 && !W[p].cplx && !psi[p].cplx, PS_ESHAPE,
               "dmrg: real W (Dl, Dr, d, d) and psi (Dl, Dr, d) expected");
    PEPS_CHECK(W[p].dims[2] == W[p].dims[3] && W[p].dims[2] == psi[p].dims[2], PS_ESHAPE, "dmrg: physical dims");
    if (p + 1 < n)
      PEPS_CHECK(W[p].dims[1] == W[p + 1].dims[0] && psi[p].dims[1] == psi[p + 1].dims[0], PS_ESHAPE,
                 "dmrg: bond dims");
  }
  PEPS_CHECK(W[0].dims[0] == 1 && W[n - 1].dims[1] == 1 && psi[0].dims[0] == 1 && psi[n - 1].dims[1] == 1,
             PS_ESHAPE, "dmrg: open boundaries expected");
  // mps_compress!(psi; Dmax = D) (eps 1e-10 relative), then psi[1] /= ||psi||:
  // the boundary-MPS sweep on the (Dl, d, Dr) view leaves every tensor but
  // the first an isometry, so the norm is that of the first tensor
  MPS m(n);
  for (size_t p = 0; p < n; ++p) m[p] = permute(c, psi[p], {0, 2, 1});
  compress_mps(c, m, D, 1e-10, true);
  for (size_t p = 0; p < n; ++p) psi[p] = permute(c, m[p], {0, 2, 1});
  {
    const double nrm = std::sqrt(norm2sq(c, psi[0]));
    PEPS_CHECK(nrm > 0.0 && std::isfinite(nrm), PS_ESOLVER, "dmrg: zero or non-finite start state");
    psi[0] = copy_tensor(c, psi[0]);
    scale(c, psi[0], 1.0 / nrm);
  }
  DmrgResult r;
  double Eprev = HUGE_VAL;
  for (int sweep = 0; sweep < nsweeps; ++sweep) {
    std::vector<Tensor> L(n), R(n);
    R[n - 1] = ones111(c);
    for (size_t p = n - 1; p >= 1; --p) R[p - 1] = right_env(c, R[p], W[p], psi[p]);
    L[0] = ones111(c);
    for (size_t p = 0; p + 1 < n; ++p) r.E = two_site_update(c, psi, W, L, R, p, D, true, tol, maxiter, krylovdim);
    std::vector<Tensor> L2(n), R2(n);
    L2[0] = ones111(c);
    for (size_t p = 1; p < n; ++p) L2[p] = left_env(c, L2[p - 1], W[p - 1], psi[p - 1]);
    R2[n - 1] = ones111(c);
    for (size_t p = n - 1; p >= 1; --p) r.E = two_site_update(c, psi, W, L2, R2, p - 1, D, false, tol, maxiter, krylovdim);
    r.sweeps = sweep + 1;
    c.count("dmrg[sweeps]", 1.0, 0);
    if (c.verbose) {
      int64_t bmax = 1;
      for (const Tensor& t : psi) bmax = std::max(bmax, t.dims[1]);
      c.note("dmrg sweep " + std::to_string(sweep + 1) + ": E = " + std::to_string(r.E) + ", max bond " +
             std::to_string(bmax));
    }
    if (std::fabs(r.E - Eprev) < etol) break;
    Eprev = r.E;
  }
  r.psi = std::move(psi);
  return r;
}

}  // namespace peps
