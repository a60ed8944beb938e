// The finite-PEPS contraction engine on the GPU: state, boundary MPS, bond
// environments, ALS truncation, simple/full bond updates and the Trotter
// drivers with exactly cached environments.
#pragma once
#include <map>
#include <string>

#include "model.h"
#include "peps_internal.h"

namespace peps {

constexpr double FU_REG = 1e-10;      // finite_peps_full_update.jl:336
constexpr double GI_REG = 1e-12;      // finite_peps_gauge_invariant.jl:224
constexpr double V4_ALS_REG = 1e-10;  // finite_peps_fullupdate_v4.jl:30
constexpr double CMPS_CUT = 1e-14;    // finite_peps_boundary_mps.jl:93
constexpr double SVD_CUT = 1e-12;     // finite_peps_ground_state.jl:57
constexpr double SU_PEEL = 1e-15;     // finite_peps_ground_state.jl:264-269

using MPS = std::vector<Tensor>;

struct Params {
  double g = 1.0, t = 1.0, m = 0.5, tau = 0.05;
  double mu = 0.0;  // chemical potential mu n_f in H_onsite_site (dense step only)
  int Dmax = 2, chi = 16, n_als = 30;
  double als_tol = 0.0;  // > 0: stop sweeping when the relative cost change < tol
  bool renorm = false;   // normalise every boundary MPS after each row absorption
  bool positive_cost = false;  // accept an ALS refinement only if its cost is >= 0
};

// GaugePEPS (finite_peps_gauge_invariant.jl:82-88) held on the device.
struct State {
  Ctx* c = nullptr;
  int nx = 0, ny = 0, gd = 3;
  bool cplx = false;
  std::vector<int64_t> gch;
  std::vector<Tensor> T;
  std::vector<std::vector<double>> lh, lv;
  std::vector<std::vector<int64_t>> qh, qv;
  // false once a dense (label-free) update has run: the flux labels are
  // then undefined and the gauge-invariant operations refuse the state
  bool labelled = true;

  Tensor& t(int ix, int iy) { return T.at(size_t(ix - 1) + size_t(iy - 1) * nx); }
  std::vector<double>& LH(int ix, int iy) { return lh.at(size_t(ix - 1) + size_t(iy - 1) * (nx - 1)); }
  std::vector<double>& LV(int ix, int iy) { return lv.at(size_t(ix - 1) + size_t(iy - 1) * nx); }
  std::vector<int64_t>& QH(int ix, int iy) { return qh.at(size_t(ix - 1) + size_t(iy - 1) * (nx - 1)); }
  std::vector<int64_t>& QV(int ix, int iy) { return qv.at(size_t(ix - 1) + size_t(iy - 1) * nx); }
  int64_t q(int ix, int iy) const { return gch.at(size_t(ix - 1) + size_t(iy - 1) * nx); }
  void init(Ctx* ctx, int nx_, int ny_, int gd_, const int64_t* g, bool cplx_);
};

struct Env {
  Tensor env, RA, RB, QA, QB;
};

struct ALSInfo {
  double cost0 = 0, cost1 = 0;
  bool accepted = false;
  int sweeps = 0;
};

struct CSVD {
  Tensor U, Vt;
  std::vector<double> S;
  std::vector<int64_t> q;
};

// ── site-level pieces ────────────────────────────────────────────────────────
void apply_mask(State& s, int ix, int iy);
double forbidden_norm(State& s);
Tensor sqrt_absorbed(State& s, int ix, int iy);
Tensor dl_traced(Ctx& c, const Tensor& T);
// _dl_open: (p,l,r,u,d) -> (p, p', l², r², u², d²), bra (') fast (kernels.hip)
Tensor dl_open(Ctx& c, const Tensor& T);
Tensor gate_tensor(Ctx& c, const Mat& G, bool cplx);

// ── boundary MPS ─────────────────────────────────────────────────────────────
MPS row_dl(State& s, int iy);
MPS row_as_mps_down(Ctx& c, const MPS& row);
MPS row_as_mps_up(Ctx& c, const MPS& row);
MPS absorb_row_from_top(Ctx& c, const MPS& mps, const MPS& row);
MPS absorb_row_from_bottom(Ctx& c, const MPS& mps, const MPS& row);
// relative: keep singular values > cutoff * (largest) instead of > cutoff
void compress_mps(Ctx& c, MPS& mps, int chi, double cutoff = CMPS_CUT, bool relative = false);
// Same truncated MPS from left Gram matrices and top-chi eigenvectors.
void compress_mps_gram(Ctx& c, MPS& mps, int chi, double cutoff = CMPS_CUT);

// Truncated right factor of M (m x n, real or complex): Vk (n x k)
// orthonormal, the leading k right singular vectors with k = max(1, min(kmax,
// #{s > thr})), thr = cut s_0 (relative) or cut, and P = M Vk, from the
// Jordan-Wielandt eigenproblem [[0, C], [C^H, 0]] of M's square core C
// (engine.cpp).
void topk_right_jw(Ctx& c, const Tensor& M, int64_t kmax, double cut, bool relative, Tensor* Vk, Tensor* P);

// ── MPS / DMRG (dmrg.cpp; mps_lgt.jl) ──────────────────────────────────────
// MPS tensors (Dl, Dr, d), MPO tensors (Dl, Dr, k, b); DMRG real only.
struct DmrgResult {
  double E = 0.0;
  int sweeps = 0;
  std::vector<Tensor> psi;
};
// dmrg_ground_state (mps_lgt.jl:418-467) with two_site_update! (:384-416)
DmrgResult dmrg_ground_state(Ctx& c, const std::vector<Tensor>& W, std::vector<Tensor> psi, int D, int nsweeps,
                             double tol, int maxiter, int krylovdim, double etol);
// mpo_expect (mps_lgt.jl:308-315): <psi|W|psi> (real part; psi real or complex)
double mpo_expect(Ctx& c, const std::vector<Tensor>& W, const std::vector<Tensor>& psi);
// krylov_evolve_step (mps_lgt.jl:546-586) nsteps times with the zip-up MPO
// application (:239-257), direct sums (:259-281) and compression (:283-303):
// psi <- exp(-i dt W) psi, renormalised; returns a complex MPS.
std::vector<Tensor> krylov_evolve(Ctx& c, const std::vector<Tensor>& W, const std::vector<Tensor>& psi0, double dt,
                                  int nsteps, int m, int Dmax, double eps);
// mps_local_expect (mps_lgt.jl:317-330) for each (site, real operator) pair.
std::vector<double> mps_local_expect(Ctx& c, const std::vector<Tensor>& psi, const std::vector<int>& sites,
                                     const std::vector<Tensor>& ops);
// absorb_row_from_top/bottom + compress_mps!, literal or fused Gram form.
// key_row > 0 names the row just absorbed: the Gram eigensolves of this
// boundary warm-start from the previous computation at the same position.
// normalise: relative compression cut and the result divided by its norm
// (PEPS_STEP_RENORM).
MPS absorb_compress(Ctx& c, const MPS& mps, const MPS& row, bool from_top, int chi, int key_row = 0,
                    bool normalise = false);
MPS trivial_mps(Ctx& c, int n, bool cplx);
MPS boundary_top(State& s, int stop_row, int chi);
MPS boundary_bottom(State& s, int stop_row, int chi);
cplx peps_norm2(State& s, int chi);

// ── bond environments ────────────────────────────────────────────────────────
Tensor spectator_left(Ctx& c, const Tensor& Lenv, const Tensor& top, const Tensor& bot, const Tensor& E);
Tensor spectator_right(Ctx& c, const Tensor& Renv, const Tensor& top, const Tensor& bot, const Tensor& E);
Tensor spectator_left_v(Ctx& c, const Tensor& Lenv, const Tensor& top, const Tensor& bot,
                        const Tensor& E0, const Tensor& E1);
Tensor spectator_right_v(Ctx& c, const Tensor& Renv, const Tensor& top, const Tensor& bot,
                         const Tensor& E0, const Tensor& E1);
Tensor ones_tensor(Ctx& c, int rank, bool cplx);
Env bond_env_h(State& s, int ix, int iy, const MPS& top, const MPS& bot, const Tensor& Lenv,
               const Tensor& Renv);
Env bond_env_v(State& s, int ix, int iy, const MPS& top, const MPS& bot, const Tensor& Lenv,
               const Tensor& Renv);
// Recompute-from-scratch forms (finite_peps_boundary_mps.jl:389, :507).
Env bond_environment_h(State& s, int ix, int iy, int chi);
Env bond_environment_v(State& s, int ix, int iy, int chi);

// ── truncations ──────────────────────────────────────────────────────────────
// leading_only: sectors much larger than D_max get only their leading D_max
// triplets (randomized subspace iteration, residual-checked, full SVD as
// the fallback) -- all that the global top-D_max selection reads.
CSVD charge_svd(Ctx& c, const Tensor& M, const std::vector<int64_t>& rq,
                const std::vector<int64_t>& cq, int Dmax, double reg = GI_REG, bool leading_only = false);
Tensor apply_gate(Ctx& c, const Tensor& th, const Tensor& G);
// _solve_herm (finite_peps_fullupdate_v4.jl:37-45) for the k columns of S.
Tensor solve_herm(Ctx& c, const Tensor& N, const Tensor& S, double reg);
double als_cost(Ctx& c, const Tensor& th, const Tensor& env, const Tensor& X, const Tensor& Y,
                const Tensor* Eth = nullptr);
void als_truncate(Ctx& c, const Tensor& th, const Tensor& env, int Dmax,
                  const std::vector<int64_t>& eR, int n_als, double reg, double tol, Tensor* X,
                  Tensor* Y, std::vector<int64_t>* bflux, ALSInfo* info,
                  const Tensor* X_init = nullptr, const Tensor* Y_init = nullptr,
                  bool positive_cost = false);

// ── bond updates and drivers ────────────────────────────────────────────────
void su_bond_h(State& s, int ix, int iy, const Tensor& G, int Dmax);
void su_bond_v(State& s, int ix, int iy, const Tensor& G, int Dmax);
void fu_bond_h(State& s, int ix, int iy, const Tensor& G, int Dmax, int n_als, double tol,
               const Env& E, ALSInfo* info = nullptr, bool positive_cost = false);
void fu_bond_v(State& s, int ix, int iy, const Tensor& G, int Dmax, int n_als, double tol,
               const Env& E, ALSInfo* info = nullptr, bool positive_cost = false);

struct GateCache {
  std::map<std::string, Tensor> g;
  // exp(-tau_factor tau H_b) of bond b, or exp(-i tau_factor tau H_b) when
  // real_time (complex states only)
  Tensor get(State& s, char orient, int ix, int iy, const Params& p, double tau_factor, bool real_time = false);
};

void trotter_step_gauge(State& s, const Params& p, GateCache& gc);
void trotter_step_v4(State& s, const Params& p, GateCache& gc, std::vector<ALSInfo>* infos = nullptr);
double energy_v4(State& s, const Params& p);
// <n_f> from simple-update weights, averaged over sites (finite_peps_quench.jl:134).
double mean_nf(State& s);

// ── dense simple update with plaquettes (finite_peps_ground_state.jl) ───────
struct PlaquetteGate {
  Tensor GA, GB, GC;  // (gd,gd,gd,gd,Da,Db), (gd,gd,Da), (gd,gd,Db)
  int Da = 1, Db = 1;
};
// The constants that distinguish the label-free simple updates: singular
// values kept above `cut`, lambda = S / max(||S||, norm_floor) (0: plain
// norm), transverse weights divided out with 1 / (sqrt(lambda) + peel).
struct SURule {
  double cut, peel, norm_floor;
};
constexpr SURule SU_DENSE{SVD_CUT, SU_PEEL, 0.0};  // update_bond_h!/v!
constexpr SURule SU_V2{FU_REG, FU_REG, 1e-15};     // full_update_bond_h!/v! (finite_peps_full_update.jl:336-454)
// update_bond_h! / update_bond_v! (:175-382)
void dense_bond_h(State& s, int ix, int iy, const Tensor& G, int Dtrunc, const SURule& rule);
void dense_bond_v(State& s, int ix, int iy, const Tensor& G, int Dtrunc, const SURule& rule);
// decompose_plaquette_LShape(exp(-tau H_plaquette_gauge)) (u1_lgt_hamiltonian.jl:728-768)
PlaquetteGate plaquette_gate(Ctx& c, int gd, double g, double tau, double cutoff, int D_aux_max, bool cplx);
// update_plaquette! (:405-483)
void update_plaquette(State& s, int ix, int iy, const PlaquetteGate& pg, int Dtrunc);
// trotter_step! (:498-541)
void trotter_step_dense(State& s, const Params& p, GateCache& gc);
// trotter_step_full! (finite_peps_full_update.jl:550-575, use_gauss = false)
void trotter_step_full_v2(State& s, const Params& p, GateCache& gc);
// apply_site_op! (finite_peps_quench.jl:228-234); marks the state unlabelled
void apply_site_op(State& s, int ix, int iy, const Tensor& O);
// realtime_trotter_step! (finite_peps_quench.jl:81-109): dt = p.tau
void trotter_step_realtime(State& s, const Params& p, GateCache& gc);
// expect_site (:549-567), O is dp x dp column-major (complex interleaved if o_cplx)
double expect_site(State& s, int ix, int iy, const std::vector<double>& O, bool o_cplx);
// plaquette_energy (:589-684) and total_energy (:690-805)
double plaquette_energy(State& s, int ix, int iy, double g);
double total_energy(State& s, double g, double t, double m);
// <W> on the plaquette with lower-left site (ix, iy) from boundary MPS
// (two-row strip; the boundary-MPS counterpart of plaquette_energy)
cplx plaquette_expect(State& s, int ix, int iy, int chi);
// bond_entropy and measure_all_finite (finite_peps_quench.jl:121-194)
double bond_entropy(const std::vector<double>& lam);
struct Measurements {
  double nf_mean = 0, nf_even = 0, nf_odd = 0, E2_mean = 0, S_h_mean = 0, S_v_mean = 0, S_mean = 0;
  std::vector<double> nf_grid, E2_grid;  // nx x ny, ix fastest
};
Measurements measure_all_finite(State& s);

}  // namespace peps
