// Host-side model: u1_lgt_hamiltonian.jl:81-163, 605-863 restated in C++
// (labels generalised to 2E, see model.h).  All matrices are small
// (<= 2*gd^2 squared per site) and built in FP64 on the host.
#include "model.h"

#include <cmath>
#include <stdexcept>

namespace peps {

SiteDims site_dims(int ix, int iy, int nx, int ny, int gd) {
  // u1_lgt_hamiltonian.jl:605-611
  SiteDims s;
  s.dR = ix < nx ? gd : 1;
  s.dU = iy < ny ? gd : 1;
  s.d = 2 * s.dR * s.dU;
  return s;
}

int64_t boundary_label(int gd) { return (gd % 2 == 1) ? 0 : 1; }

void decode_phys(int p, int dR, int dU, int gd, int64_t* nf, int64_t* lR, int64_t* lU) {
  // finite_peps_gauge_invariant.jl:58-67
  int p0 = p - 1;
  *nf = p0 / (dR * dU);
  int rem = p0 % (dR * dU);
  int iR1 = rem / dU, iU0 = rem % dU;
  int64_t b0 = boundary_label(gd);
  *lR = dR == 1 ? b0 : int64_t(2 * iR1 - (gd - 1));
  *lU = dU == 1 ? b0 : int64_t(2 * iU0 - (gd - 1));
}

void phys_labels(int ix, int iy, int nx, int ny, int gd, std::vector<int64_t>* nf,
                 std::vector<int64_t>* lR, std::vector<int64_t>* lU) {
  SiteDims s = site_dims(ix, iy, nx, ny, gd);
  nf->resize(s.d);
  lR->resize(s.d);
  lU->resize(s.d);
  for (int p = 1; p <= s.d; ++p) decode_phys(p, s.dR, s.dU, gd, &(*nf)[p - 1], &(*lR)[p - 1], &(*lU)[p - 1]);
}

static Mat zeros(int n) {
  Mat m;
  m.n = n;
  m.a.assign(size_t(n) * n, 0.0);
  return m;
}
static Mat eye(int n) {
  Mat m = zeros(n);
  for (int i = 0; i < n; ++i) m(i, i) = 1.0;
  return m;
}
// Julia kron(A, B): (i-1)*p + k, B fastest.
Mat kron(const Mat& A, const Mat& B) {
  Mat C = zeros(A.n * B.n);
  for (int j = 0; j < A.n; ++j)
    for (int i = 0; i < A.n; ++i) {
      double x = A(i, j);
      if (x == 0.0) continue;
      for (int l = 0; l < B.n; ++l)
        for (int k = 0; k < B.n; ++k) C(i * B.n + k, j * B.n + l) = x * B(k, l);
    }
  return C;
}
static Mat kron3(const Mat& A, const Mat& B, const Mat& C) { return kron(kron(A, B), C); }
static void axpy(Mat& Y, double a, const Mat& X) {
  for (size_t i = 0; i < Y.a.size(); ++i) Y.a[i] += a * X.a[i];
}
static Mat transpose(const Mat& A) {
  Mat T = zeros(A.n);
  for (int j = 0; j < A.n; ++j)
    for (int i = 0; i < A.n; ++i) T(j, i) = A(i, j);
  return T;
}
static Mat hsym(const Mat& H) {  // 0.5 (H + H')
  Mat T = transpose(H);
  Mat R = zeros(H.n);
  for (size_t i = 0; i < R.a.size(); ++i) R.a[i] = 0.5 * (H.a[i] + T.a[i]);
  return R;
}

static Mat op_E2(int gd) {  // u1_lgt_hamiltonian.jl:110-120, E = l/2
  Mat E = zeros(gd);
  for (int i = 0; i < gd; ++i) {
    double e = 0.5 * (2 * i - (gd - 1));
    E(i, i) = e * e;
  }
  return E;
}
static Mat op_U(int gd) {  // :123-130
  Mat U = zeros(gd);
  for (int i = 0; i + 1 < gd; ++i) U(i + 1, i) = 1.0;
  return U;
}
static Mat op_c() {  // :138-144
  Mat c = zeros(2);
  c(0, 1) = 1.0;
  return c;
}
static Mat op_nf() {
  Mat n = zeros(2);
  n(1, 1) = 1.0;
  return n;
}

Mat H_onsite_site(int ix, int iy, int nx, int ny, int gd, double g, double m, double mu) {
  // u1_lgt_hamiltonian.jl:658-676
  SiteDims s = site_dims(ix, iy, nx, ny, gd);
  Mat H = zeros(s.d);
  if (ix < nx) axpy(H, g * g / 2, kron3(eye(2), op_E2(gd), eye(s.dU)));
  if (iy < ny) axpy(H, g * g / 2, kron3(eye(2), eye(s.dR), op_E2(gd)));
  int stag = ((ix + iy) % 2 == 0) ? 1 : -1;
  axpy(H, m * stag, kron3(op_nf(), eye(s.dR), eye(s.dU)));
  if (mu != 0.0) axpy(H, mu, kron3(op_nf(), eye(s.dR), eye(s.dU)));
  return hsym(H);
}

Mat H_hop_h_site(int ix, int iy, int nx, int ny, int gd, double t) {
  // u1_lgt_hamiltonian.jl:689-701
  if (ix >= nx) throw std::invalid_argument("H_hop_h_site: no right bond");
  SiteDims L = site_dims(ix, iy, nx, ny, gd), R = site_dims(ix + 1, iy, nx, ny, gd);
  Mat cdagUL = kron3(transpose(op_c()), op_U(gd), eye(L.dU));
  Mat cR = kron3(op_c(), eye(R.dR), eye(R.dU));
  Mat hop = kron(cdagUL, cR);
  Mat H = zeros(hop.n);
  axpy(H, -t, hop);
  axpy(H, -t, transpose(hop));
  return hsym(H);
}

Mat H_hop_v_site(int ix, int iy, int nx, int ny, int gd, double t) {
  // u1_lgt_hamiltonian.jl:789-801
  if (iy >= ny) throw std::invalid_argument("H_hop_v_site: no up bond");
  SiteDims D = site_dims(ix, iy, nx, ny, gd), U = site_dims(ix, iy + 1, nx, ny, gd);
  Mat cdagUD = kron3(transpose(op_c()), eye(D.dR), op_U(gd));
  Mat cU = kron3(op_c(), eye(U.dR), eye(U.dU));
  Mat hop = kron(cdagUD, cU);
  Mat H = zeros(hop.n);
  axpy(H, -t, hop);
  axpy(H, -t, transpose(hop));
  return hsym(H);
}

double onsite_weight(int ix, int iy, int nx, int ny) {
  // u1_lgt_hamiltonian.jl:815-818
  int nb = (ix > 1) + (ix < nx) + (iy > 1) + (iy < ny);
  return 1.0 / nb;
}

Mat H_merged_h_site(int ix, int iy, int nx, int ny, int gd, double g, double t, double m, double mu) {
  // u1_lgt_hamiltonian.jl:829-841
  int dL = site_dims(ix, iy, nx, ny, gd).d, dR = site_dims(ix + 1, iy, nx, ny, gd).d;
  Mat H = H_hop_h_site(ix, iy, nx, ny, gd, t);
  axpy(H, onsite_weight(ix, iy, nx, ny), kron(H_onsite_site(ix, iy, nx, ny, gd, g, m, mu), eye(dR)));
  axpy(H, onsite_weight(ix + 1, iy, nx, ny), kron(eye(dL), H_onsite_site(ix + 1, iy, nx, ny, gd, g, m, mu)));
  return H;
}

Mat H_merged_v_site(int ix, int iy, int nx, int ny, int gd, double g, double t, double m, double mu) {
  // u1_lgt_hamiltonian.jl:851-863
  int dD = site_dims(ix, iy, nx, ny, gd).d, dU = site_dims(ix, iy + 1, nx, ny, gd).d;
  Mat H = H_hop_v_site(ix, iy, nx, ny, gd, t);
  axpy(H, onsite_weight(ix, iy, nx, ny), kron(H_onsite_site(ix, iy, nx, ny, gd, g, m, mu), eye(dU)));
  axpy(H, onsite_weight(ix, iy + 1, nx, ny), kron(eye(dD), H_onsite_site(ix, iy + 1, nx, ny, gd, g, m, mu)));
  return H;
}

Mat H_plaquette_gauge(int gd, double g) {
  // u1_lgt_hamiltonian.jl:283-289: -(1/2g^2)(W + W'), W = Ud (x) U (x) Ud (x) U
  Mat U = op_U(gd), Ud = transpose(op_U(gd));
  Mat W = kron(kron3(Ud, U, Ud), U);
  Mat H = zeros(W.n);
  axpy(H, -1.0 / (2.0 * g * g), W);
  axpy(H, -1.0 / (2.0 * g * g), transpose(W));
  return hsym(H);
}

Mat identity(int n) { return eye(n); }

Mat op_U_gauge(int gd) { return op_U(gd); }

Mat site_nf(int ix, int iy, int nx, int ny, int gd) {
  // embed_f_site(op_nf(), d_gR, d_gU), u1_lgt_hamiltonian.jl:629
  SiteDims s = site_dims(ix, iy, nx, ny, gd);
  return kron3(op_nf(), eye(s.dR), eye(s.dU));
}

Mat site_E2(int ix, int iy, int nx, int ny, int gd) {
  // E_R^2 + E_U^2 over the existing links (finite_peps_quench.jl:151-161)
  SiteDims s = site_dims(ix, iy, nx, ny, gd);
  Mat O = zeros(s.d);
  if (ix < nx) axpy(O, 1.0, kron3(eye(2), op_E2(gd), eye(s.dU)));
  if (iy < ny) axpy(O, 1.0, kron3(eye(2), eye(s.dR), op_E2(gd)));
  return O;
}

// Cyclic Jacobi eigendecomposition of a real symmetric matrix (Golub & Van
// Loan, Alg. 8.5.3, with the classical rotation formulas).
static void jacobi_eig(Mat A, std::vector<double>* w, Mat* V) {
  const int n = A.n;
  *V = eye(n);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += A(i, j) * A(i, j);
        if (i != j) off += A(i, j) * A(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A(p, q);
        if (apq == 0.0) continue;
        double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = (*V)(k, p), vkq = (*V)(k, q);
          (*V)(k, p) = c * vkp - s * vkq;
          (*V)(k, q) = s * vkp + c * vkq;
        }
      }
  }
  w->resize(n);
  for (int i = 0; i < n; ++i) (*w)[i] = A(i, i);
}

void expm_sym_imag(const Mat& H, double dt, Mat* re, Mat* im) {
  // exp(-i dt H) = V diag(cos(dt w) - i sin(dt w)) V^T for real symmetric H
  std::vector<double> w;
  Mat V;
  jacobi_eig(H, &w, &V);
  *re = zeros(H.n);
  *im = zeros(H.n);
  for (int k = 0; k < H.n; ++k) {
    const double cr = std::cos(dt * w[k]), ci = -std::sin(dt * w[k]);
    for (int j = 0; j < H.n; ++j) {
      const double vj = V(j, k);
      if (vj == 0.0) continue;
      for (int i = 0; i < H.n; ++i) {
        (*re)(i, j) += V(i, k) * vj * cr;
        (*im)(i, j) += V(i, k) * vj * ci;
      }
    }
  }
}

Mat expm_sym(const Mat& H, double tau) {
  std::vector<double> w;
  Mat V;
  jacobi_eig(H, &w, &V);
  Mat G = zeros(H.n);
  for (int k = 0; k < H.n; ++k) {
    double e = std::exp(-tau * w[k]);
    for (int j = 0; j < H.n; ++j) {
      double vj = V(j, k) * e;
      if (vj == 0.0) continue;
      for (int i = 0; i < H.n; ++i) G(i, j) += V(i, k) * vj;
    }
  }
  return G;
}

}  // namespace peps
