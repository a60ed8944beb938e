// Two-site DMRG of a matrix-product operator on the GPU (mps_lgt.jl:333-467):
// environments, the two-site effective Hamiltonian, a Lanczos solve of the
// local ground state and the truncated SVD split, all on device tensors.
// DMRG takes real MPOs and MPS only: the U(1) lattice-gauge Hamiltonian is
// real symmetric, so its ground state is real up to a global phase.  The
// real-time part (mps_lgt.jl:204-303, 546-586) -- overlaps, the zip-up MPO
// application, direct sums, compression and the global Krylov step -- works
// on complex MPS.
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
  // (bra, mpo, ket) sandwich from the left; a complex psi takes the complex
  // form of W and the conjugate bra (the real part is returned)
  PEPS_CHECK(W.size() == psi.size() && !W.empty(), PS_ESHAPE, "mpo_expect: lengths");
  const bool cp = psi[0].cplx;
  Tensor E = cp ? as_complex(c, ones111(c)) : ones111(c);
  for (size_t p = 0; p < W.size(); ++p) {
    PEPS_CHECK(psi[p].cplx == cp, PS_EARG, "mpo_expect: mixed real/complex MPS");
    const Tensor Wp = cp ? as_complex(c, W[p]) : W[p];
    Tensor t1 = contract(c, E, "xyz", psi[p], "zcq", "xycq");
    Tensor t2 = contract(c, t1, "xycq", Wp, "ybkq", "xcbk");
    E = contract(c, t2, "xcbk", psi[p], "xak", "abc", false, true);
  }
  PEPS_CHECK(E.numel() == 1, PS_ESHAPE, "mpo_expect: open boundary bonds expected");
  return c.to_host(E)[0];
}

// ── real-time evolution (mps_lgt.jl:204-303, 546-586) ──────────────────────
//
// Inside the evolution the MPS tensors are held as (Dl, d, Dr), so that the
// left matrix (Dl d) x Dr and the right matrix Dl x (d Dr) of every site are
// free reshapes; the boundary-MPS compression sweep takes the same layout.
// All tensors are complex.

namespace {

MPS to_ldr(Ctx& c, const std::vector<Tensor>& psi) {
  MPS m(psi.size());
  for (size_t p = 0; p < psi.size(); ++p) m[p] = permute(c, as_complex(c, psi[p]), {0, 2, 1});
  return m;
}

std::vector<Tensor> from_ldr(Ctx& c, const MPS& m) {
  std::vector<Tensor> psi(m.size());
  for (size_t p = 0; p < m.size(); ++p) psi[p] = permute(c, m[p], {0, 2, 1});
  return psi;
}

// With the context's finite-value checks on, name the first non-finite
// intermediate of the evolution.
void check_finite(Ctx& c, const Tensor& t, const char* what, size_t site) {
  if (!c.debug_finite) return;
  PEPS_CHECK(all_finite(c, t), PS_ESOLVER,
             std::string("krylov_evolve: non-finite ") + what + " at site " + std::to_string(site + 1));
}

Tensor one11c(Ctx& c) {
  const double one[2] = {1.0, 0.0};
  return c.upload({1, 1}, true, one);
}

// mps_overlap (mps_lgt.jl:204-215): <phi|psi>, v(a, b) over (bra, ket) bonds.
cplx overlap(Ctx& c, const MPS& phi, const MPS& psi) {
  Tensor v = one11c(c);
  for (size_t p = 0; p < psi.size(); ++p) {
    Tensor t = contract(c, v, "ab", psi[p], "bqc", "aqc");
    v = contract(c, phi[p], "aqd", t, "aqc", "dc", true, false);
  }
  std::vector<double> h = c.to_host(v);
  return {h[0], h[1]};
}

double mps_norm(Ctx& c, const MPS& psi) { return std::sqrt(std::max(0.0, overlap(c, psi, psi).re)); }

// psi with its first tensor multiplied by s (the other tensors shared).
MPS scaled(Ctx& c, const MPS& psi, cplx s) {
  MPS out = psi;
  out[0] = copy_tensor(c, psi[0]);
  scale(c, out[0], s);
  return out;
}

// The truncated split M ~= U R of a (m x n) site matrix: U (m x k) with
// orthonormal columns spanning the leading k left singular vectors, R = U^H M
// (k x n), k = max(1, min(kmax, #{s > eps s_0})) (the keep rule of
// mps_lgt.jl:250).  Blocks with min(m, n) >= kEvolveJwMinDim take the
// Jordan-Wielandt form of M^H, smaller ones the one-workgroup Jacobi SVD.
constexpr int64_t kEvolveJwMinDim = 32;

void trunc_left(Ctx& c, const Tensor& M, int64_t kmax, double eps, Tensor* U, Tensor* R) {
  const int64_t m = M.dims[0], n = M.dims[1];
  if (std::min(m, n) >= kEvolveJwMinDim) {
    Tensor P;
    topk_right_jw(c, conj_transpose(c, M), kmax, eps, true, U, &P);  // U: leading left vectors of M, P = M^H U
    *R = conj_transpose(c, P);
    return;
  }
  SVD f = svd_thin(c, M);
  int64_t cnt = 0;
  for (double x : f.S) cnt += x > eps * f.S[0];
  const int64_t k = std::max<int64_t>(1, std::min<int64_t>({kmax, cnt, int64_t(f.S.size())}));
  *U = copy_tensor(c, f.U.slice_last(0, k));
  *R = gather2d(c, f.Vt, seq_range(k), seq_range(n));
  scale_rows(c, *R, std::vector<double>(f.S.begin(), f.S.begin() + k));
}

// mpo_apply_mps_zipup (mps_lgt.jl:239-257): one left-to-right sweep that
// contracts L (dl, aH, apsi) with W[p] and psi[p] and splits the result at
// once, so the new bond never exceeds Dmax.
MPS zipup(Ctx& c, const std::vector<Tensor>& W, const MPS& psi, int64_t Dmax, double eps) {
  const size_t n = psi.size();
  MPS out(n);
  Tensor L = as_complex(c, ones111(c));
  for (size_t p = 0; p < n; ++p) {
    Tensor t = contract(c, L, "xhs", psi[p], "sqb", "xhqb");
    Tensor B = contract(c, t, "xhqb", W[p], "hgkq", "xkgb");  // (dl, k, bH, bpsi)
    const int64_t dl = B.dims[0], d = B.dims[1], bH = B.dims[2], bp = B.dims[3];
    check_finite(c, B, "zip-up site product", p);
    if (p + 1 < n) {
      Tensor U, R;
      trunc_left(c, B.reshaped({dl * d, bH * bp}), Dmax, eps, &U, &R);
      check_finite(c, U, "zip-up left factor", p);
      check_finite(c, R, "zip-up carried factor", p);
      const int64_t keep = U.dims[1];
      out[p] = U.reshaped({dl, d, keep});
      L = R.reshaped({keep, bH, bp});
    } else {
      PEPS_CHECK(bH * bp == 1, PS_ESHAPE, "zipup: open right boundary expected");
      out[p] = B.reshaped({dl, d, 1});
    }
  }
  return out;
}

// mps_axpby (mps_lgt.jl:259-281): the direct sum a A + b B, block-diagonal
// in the bonds, with a and b on the first site.
MPS axpby(Ctx& c, cplx a, const MPS& A, cplx b, const MPS& B) {
  const size_t n = A.size();
  MPS C(n);
  for (size_t p = 0; p < n; ++p) {
    const int64_t dla = A[p].dims[0], d = A[p].dims[1], dra = A[p].dims[2];
    const int64_t dlb = B[p].dims[0], drb = B[p].dims[2];
    const bool first = p == 0, last = p + 1 == n;
    const int64_t Dl = first ? 1 : dla + dlb, Dr = last ? 1 : dra + drb;
    Tensor T = c.zeros({Dl, d, Dr}, true);
    Tensor Tm = T.reshaped({Dl, d * Dr});
    Tensor Ap = first ? copy_tensor(c, A[p]) : A[p], Bp = first ? copy_tensor(c, B[p]) : B[p];
    if (first) {
      scale(c, Ap, a);
      scale(c, Bp, b);
    }
    put_block(c, Tm, 0, 0, Ap.reshaped({dla, d * dra}));
    put_block(c, Tm, first ? 0 : dla, last ? 0 : d * dra, Bp.reshaped({dlb, d * drb}));
    check_finite(c, T, "direct sum", p);
    C[p] = T;
  }
  return C;
}

// mps_compress! (mps_lgt.jl:283-303): the boundary-MPS sweep (orthonormal
// left factors, then the truncated right-to-left SVD with the relative cut).
void compress(Ctx& c, MPS& psi, int64_t Dmax, double eps) { compress_mps(c, psi, int(Dmax), eps, true); }

// exp(-i dt T) e_1 for the real symmetric tridiagonal T = tridiag(beta,
// alpha, beta): T = Q diag(w) Q^T, so the coefficients are
// sum_l Q(j, l) e^{-i dt w_l} Q(0, l).
std::vector<cplx> krylov_coeffs(const std::vector<double>& al, const std::vector<double>& be, double dt) {
  const int k = int(al.size());
  std::vector<double> T(size_t(k) * k, 0.0), w, Q;
  for (int i = 0; i < k; ++i) T[size_t(i) * k + i] = al[i];
  for (int i = 0; i + 1 < k; ++i) T[size_t(i) * k + i + 1] = T[size_t(i + 1) * k + i] = be[i];
  small_sym_eig(k, T, &w, &Q);
  std::vector<cplx> out(k, {0.0, 0.0});
  for (int j = 0; j < k; ++j)
    for (int l = 0; l < k; ++l) {
      const double f = Q[size_t(j) * k + l] * Q[size_t(l)];
      out[j].re += f * std::cos(dt * w[l]);
      out[j].im -= f * std::sin(dt * w[l]);
    }
  return out;
}

// krylov_evolve_step (mps_lgt.jl:546-586): psi <- exp(-i dt H) psi on the
// MPS manifold, Lanczos with full reorthogonalisation, renormalised.
MPS krylov_step(Ctx& c, const std::vector<Tensor>& W, const MPS& psi, double dt, int m, int64_t Dmax, double eps,
                int* kdim) {
  const cplx one{1.0, 0.0};
  std::vector<MPS> V;
  V.push_back(scaled(c, psi, {1.0 / mps_norm(c, psi), 0.0}));
  std::vector<double> al, be;
  MPS w = zipup(c, W, V[0], Dmax, eps);
  al.push_back(overlap(c, V[0], w).re);
  w = axpby(c, one, w, {-al[0], 0.0}, V[0]);
  compress(c, w, Dmax, eps);
  for (int j = 2; j <= m; ++j) {
    const double b = mps_norm(c, w);
    if (b < 1e-12) break;
    be.push_back(b);
    V.push_back(scaled(c, w, {1.0 / b, 0.0}));
    const MPS& vj = V.back();
    w = zipup(c, W, vj, Dmax, eps);
    const double aj = overlap(c, vj, w).re;
    al.push_back(aj);
    w = axpby(c, one, w, {-aj, 0.0}, vj);
    w = axpby(c, one, w, {-b, 0.0}, V[V.size() - 2]);
    compress(c, w, Dmax, eps);
    for (const MPS& vk : V) {  // full reorthogonalisation
      const cplx ck = overlap(c, vk, w);
      w = axpby(c, one, w, {-ck.re, -ck.im}, vk);
    }
    compress(c, w, Dmax, eps);
  }
  const std::vector<cplx> co = krylov_coeffs(al, be, dt);
  const int k = int(al.size());
  MPS out = scaled(c, V[0], co[0]);
  for (int j = 1; j < k; ++j) {
    out = axpby(c, one, out, co[j], V[j]);
    compress(c, out, Dmax, eps);
  }
  out = scaled(c, out, {1.0 / mps_norm(c, out), 0.0});
  c.count("mps_evolve[krylov dimension]", double(k), 0);
  if (kdim) *kdim = k;
  return out;
}

void check_chain(const std::vector<Tensor>& W, const std::vector<Tensor>& psi, const char* who) {
  const size_t n = W.size();
  PEPS_CHECK(n >= 2 && psi.size() == n, PS_EARG, std::string(who) + ": at least two sites");
  for (size_t p = 0; p < n; ++p) {
    PEPS_CHECK(W[p].rank() == 4 && psi[p].rank() == 3 && !W[p].cplx, PS_ESHAPE,
               std::string(who) + ": real W (Dl, Dr, d, d) and psi (Dl, Dr, d) expected");
    PEPS_CHECK(W[p].dims[2] == W[p].dims[3] && W[p].dims[2] == psi[p].dims[2], PS_ESHAPE,
               std::string(who) + ": physical dims");
    if (p + 1 < n)
      PEPS_CHECK(W[p].dims[1] == W[p + 1].dims[0] && psi[p].dims[1] == psi[p + 1].dims[0], PS_ESHAPE,
                 std::string(who) + ": bond dims");
  }
  PEPS_CHECK(W[0].dims[0] == 1 && W[n - 1].dims[1] == 1 && psi[0].dims[0] == 1 && psi[n - 1].dims[1] == 1,
             PS_ESHAPE, std::string(who) + ": open boundaries expected");
}

}  // namespace

std::vector<Tensor> krylov_evolve(Ctx& c, const std::vector<Tensor>& W, const std::vector<Tensor>& psi0, double dt,
                                  int nsteps, int m, int Dmax, double eps) {
  TierGuard fp64(c, TIER_FP64);
  check_chain(W, psi0, "krylov_evolve");
  PEPS_CHECK(nsteps >= 0 && m >= 1 && Dmax >= 1 && eps >= 0.0 && std::isfinite(dt), PS_EARG,
             "krylov_evolve: nsteps >= 0, m >= 1, Dmax >= 1, eps >= 0, finite dt");
  std::vector<Tensor> Wc(W.size());
  for (size_t p = 0; p < W.size(); ++p) Wc[p] = as_complex(c, W[p]);
  MPS psi = to_ldr(c, psi0);
  const double nrm = mps_norm(c, psi);
  PEPS_CHECK(nrm > 0.0 && std::isfinite(nrm), PS_ESOLVER, "krylov_evolve: zero or non-finite state");
  for (int s = 0; s < nsteps; ++s) {
    int k = 0;
    psi = krylov_step(c, Wc, psi, dt, m, Dmax, eps, &k);
    c.count("mps_evolve[steps]", 1.0, 0);
    if (c.verbose) {
      int64_t bmax = 1;
      for (const Tensor& t : psi) bmax = std::max(bmax, t.dims[2]);
      c.note("krylov step " + std::to_string(s + 1) + ": Krylov dimension " + std::to_string(k) + ", max bond " +
             std::to_string(bmax));
    }
  }
  return from_ldr(c, psi);
}

std::vector<double> mps_local_expect(Ctx& c, const std::vector<Tensor>& psi0, const std::vector<int>& sites,
                                     const std::vector<Tensor>& ops) {
  // mps_local_expect (mps_lgt.jl:317-330) for a list of (site, operator)
  // pairs: the left and right overlap environments are built once and each
  // <psi|O_p|psi> / <psi|psi> closes at its site
  const size_t n = psi0.size();
  PEPS_CHECK(n >= 1 && sites.size() == ops.size(), PS_EARG, "mps_local_expect: lengths");
  for (size_t p = 0; p < n; ++p) {
    PEPS_CHECK(psi0[p].rank() == 3, PS_ESHAPE, "mps_local_expect: psi (Dl, Dr, d) expected");
    if (p + 1 < n) PEPS_CHECK(psi0[p].dims[1] == psi0[p + 1].dims[0], PS_ESHAPE, "mps_local_expect: bond dims");
  }
  PEPS_CHECK(psi0[0].dims[0] == 1 && psi0[n - 1].dims[1] == 1, PS_ESHAPE, "mps_local_expect: open boundaries");
  MPS psi = to_ldr(c, psi0);
  std::vector<Tensor> EL(n + 1), ER(n + 1);
  EL[0] = one11c(c);
  for (size_t p = 0; p < n; ++p) {
    Tensor t = contract(c, EL[p], "ab", psi[p], "bqc", "aqc");
    EL[p + 1] = contract(c, psi[p], "aqd", t, "aqc", "dc", true, false);
  }
  ER[n] = one11c(c);
  for (size_t p = n; p-- > 0;) {
    Tensor t = contract(c, psi[p], "bqc", ER[p + 1], "dc", "bqd");
    ER[p] = contract(c, psi[p], "aqd", t, "bqd", "ab", true, false);
  }
  const double nrm = c.to_host(EL[n])[0];
  PEPS_CHECK(nrm > 0.0 && std::isfinite(nrm), PS_ESOLVER, "mps_local_expect: zero or non-finite state");
  std::vector<double> out(sites.size());
  for (size_t i = 0; i < sites.size(); ++i) {
    const int p = sites[i];
    PEPS_CHECK(p >= 0 && size_t(p) < n, PS_EARG, "mps_local_expect: site out of range");
    const int64_t d = psi[p].dims[1];
    PEPS_CHECK(ops[i].rank() == 2 && ops[i].dims[0] == d && ops[i].dims[1] == d && !ops[i].cplx, PS_ESHAPE,
               "mps_local_expect: real (d, d) operator expected");
    Tensor Op = contract(c, psi[p], "bqc", as_complex(c, ops[i]), "kq", "bkc");  // (O psi)(b, k, c)
    Tensor t = contract(c, EL[p], "ab", Op, "bkc", "akc");
    Tensor u = contract(c, t, "akc", ER[p + 1], "dc", "akd");
    out[i] = dot(c, u, psi[p], true).re / nrm;
  }
  return out;
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
