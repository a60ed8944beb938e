// Exact diagonalisation of the finite U(1) lattice gauge theory in one
// Gauss-law sector (finite_ed.jl): basis enumeration, the Hamiltonian, its
// ground state and the observables, on the GPU.
#pragma once

#include <memory>
#include <vector>

#include "peps_internal.h"

namespace peps {

struct EDBasis {
  Ctx* c = nullptr;
  int nx = 0, ny = 0, dg = 0, lb = 0;  // lb: bits per link field of the key
  int nsite = 0, nh = 0, nv = 0;
  std::vector<int64_t> g;              // background charges, ix fastest
  int64_t dim = 0;
  std::shared_ptr<Buffer> keys;        // uint64 keys, the reference's order
  std::shared_ptr<Buffer> sorted;      // the keys ascending
  std::shared_ptr<Buffer> perm;        // int32: basis index of each sorted key
  // Hamiltonian: diagonal and, per state, W neighbour indices (-1: none),
  // each coupled by -t (column-major dim x W)
  bool has_h = false;
  int W = 0;
  double t_hop = 0.0;
  Tensor diag;
  std::shared_ptr<Buffer> cols;
  bool has_gs = false;
  double E0 = 0.0;
  Tensor psi0;
};

struct EDObs {
  double nf_mean = 0, nf_even = 0, nf_odd = 0, E2_mean = 0;
  std::vector<double> nf_grid, E2_grid;  // nx*ny, ix fastest
};

// build_basis (finite_ed.jl:125-233); g: nx*ny charges, ix fastest.
std::unique_ptr<EDBasis> ed_build_basis(Ctx& c, int nx, int ny, int dg, const std::vector<int64_t>& g);
// build_hamiltonian (finite_ed.jl:252-326).
void ed_build_hamiltonian(EDBasis& b, double g, double t, double m);
// y = H x (real vectors of the basis dimension).
void ed_apply(EDBasis& b, const Tensor& x, Tensor& y);
// find_ground_state (finite_ed.jl:333-341): restarted Lanczos with full
// reorthogonalisation until ||H psi - E0 psi|| <= tol.
double ed_ground_state(EDBasis& b, double tol, int krylov, int max_restarts);
// measure_ED (finite_ed.jl:369-395) of a real or complex state.
EDObs ed_measure(EDBasis& b, const Tensor& psi);
// y = H x for interleaved complex vectors.
void ed_apply_complex(EDBasis& b, const Tensor& x, Tensor& y);
// exp(-i dt H) psi (complex psi), the KrylovKit exponentiate call of
// evolve_ED! (finite_ed.jl:417-419): Lanczos projections of at most krylov
// vectors, each interval halved until the error estimate is <= tol;
// substeps (may be null) gets the number of projections used.
Tensor ed_exp_step(EDBasis& b, const Tensor& psi, double dt, int krylov, double tol, int* substeps);

}  // namespace peps
