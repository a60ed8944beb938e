// U(1) lattice-gauge model on an open nx x ny lattice: site Hilbert spaces,
// Gauss-law labels and the merged two-site Hamiltonians / Trotter gates.
//
// Flux labels are integers l = 2E.  A link of dimension gd carries
// l in {-(gd-1), ..., gd-1} step 2; a missing boundary link carries b0
// (0 for odd gd, +1 for even gd).  Gauss law: lR + lU - lL - lD = 2 (n_f + q).
#pragma once
#include <cstdint>
#include <vector>

namespace peps {

struct SiteDims {
  int d, dR, dU;
};

SiteDims site_dims(int ix, int iy, int nx, int ny, int gd);
int64_t boundary_label(int gd);
// decode_phys for a 1-based physical index p: (n_f, lR, lU)
void decode_phys(int p, int dR, int dU, int gd, int64_t* nf, int64_t* lR, int64_t* lU);
void phys_labels(int ix, int iy, int nx, int ny, int gd, std::vector<int64_t>* nf,
                 std::vector<int64_t>* lR, std::vector<int64_t>* lU);

// Dense column-major real matrices (the model's operators are real).
struct Mat {
  int n = 0;
  std::vector<double> a;  // n*n, column-major
  double& operator()(int i, int j) { return a[i + size_t(j) * n]; }
  double operator()(int i, int j) const { return a[i + size_t(j) * n]; }
};

Mat H_onsite_site(int ix, int iy, int nx, int ny, int gd, double g, double m, double mu = 0.0);
Mat H_hop_h_site(int ix, int iy, int nx, int ny, int gd, double t);
Mat H_hop_v_site(int ix, int iy, int nx, int ny, int gd, double t);
double onsite_weight(int ix, int iy, int nx, int ny);
// mu: the optional chemical potential mu n_f of H_onsite_site (0 in every
// full-update and gauge-invariant caller; the dense trotter_step! takes it)
Mat H_merged_h_site(int ix, int iy, int nx, int ny, int gd, double g, double t, double m, double mu = 0.0);
Mat H_merged_v_site(int ix, int iy, int nx, int ny, int gd, double g, double t, double m, double mu = 0.0);
// Pure-gauge plaquette term on the links (b, l, r, t), b slowest
// (u1_lgt_hamiltonian.jl:283-289).
Mat H_plaquette_gauge(int gd, double g);
Mat identity(int n);
// Raising operator U|e> = |e+1> of one link (u1_lgt_hamiltonian.jl:123-130).
Mat op_U_gauge(int gd);
// n_f and E_R^2 + E_U^2 (existing links only) on one site's physical space.
Mat site_nf(int ix, int iy, int nx, int ny, int gd);
Mat site_E2(int ix, int iy, int nx, int ny, int gd);
// Julia kron(A, B): the second factor is the fastest index.
Mat kron(const Mat& A, const Mat& B);
// exp(-tau H) for real symmetric H (Jacobi eigendecomposition, FP64).
Mat expm_sym(const Mat& H, double tau);
// exp(-i dt H) for real symmetric H: real and imaginary parts (the real-time
// gate of realtime_trotter_step!, finite_peps_quench.jl:81-109).
void expm_sym_imag(const Mat& H, double dt, Mat* re, Mat* im);

}  // namespace peps
