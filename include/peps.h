/*
 * peps.h — C ABI of libpeps_hip.so, the MI355X finite-PEPS contraction engine.
 *
 * Each entry point replaces one function of the Julia reference
 * (NoaFeldman/LGT); the reference file:line is given beside it.  A Julia
 * host reaches them through `ccall`; this repository's tests and bench
 * reach them through Python ctypes (lgt_amd/peps.py).
 *
 * Conventions
 *  - Every call returns an int32 status (PEPS_OK == 0).  On failure the
 *    context's message is available from peps_last_error().  No exception or
 *    longjmp crosses the ABI.  Shape @asserts of the reference map to
 *    PEPS_ESHAPE, the charge-SVD "no sector" error() to PEPS_ENOSECTOR.
 *  - Arrays are column-major, first index fastest (Julia layout).  Complex
 *    arrays are interleaved (re, im) doubles: byte-compatible with Julia
 *    ComplexF64 and C `double _Complex`.  Real arrays are plain doubles.  A
 *    state is created real or complex; a real state runs the real fast path.
 *  - Lattice coordinates are 1-based (ix in 1..nx, iy in 1..ny) as in Julia.
 *  - Host buffers are read or written only during the call; the library
 *    never retains host pointers.
 *  - Flux labels are int64 integers 2E.  A link of dimension gd carries
 *    labels -(gd-1), -(gd-3), ..., gd-1; a missing boundary link carries 0
 *    for odd gd (the reference's convention) and +1 for even gd.
 *  - A context is not thread-safe: use one per host thread.
 */
#ifndef LIBPEPS_HIP_PEPS_H_
#define LIBPEPS_HIP_PEPS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct peps_ctx peps_ctx;
typedef struct peps_state peps_state;
typedef struct peps_tensor_list peps_tensor_list;
typedef struct peps_mlp peps_mlp;
typedef struct peps_ed peps_ed;

enum {
  PEPS_OK = 0,
  PEPS_EARG = 1,
  PEPS_ESHAPE = 2,       /* reference @assert on shapes */
  PEPS_ENOSECTOR = 3,    /* charge_svd: no singular value above reg (gauge_invariant.jl:265) */
  PEPS_EHIP = 4,
  PEPS_ENOMEM = 5,
  PEPS_ESOLVER = 6,      /* factorisation did not converge */
  PEPS_EUNSUPPORTED = 7
};

enum { PEPS_TIER_FP64 = 0, PEPS_TIER_FP32 = 1, PEPS_TIER_BF16 = 2 };
enum { PEPS_H = 0, PEPS_V = 1 };
/* PEPS_STEP_REALTIME: realtime_trotter_step! (finite_peps_quench.jl:81-109),
   one second-order real-time step exp(-i dt H) with dt = peps_params.tau
   through the dense update_bond_h!/v! (complex states only). */
enum { PEPS_STEP_V4 = 0, PEPS_STEP_GAUGE = 1, PEPS_STEP_DENSE = 2, PEPS_STEP_FULL_V2 = 3, PEPS_STEP_REALTIME = 4 };
/* peps_params.flags bits.  PEPS_STEP_RENORM has no counterpart in the
   reference: every boundary MPS built by a full-update step or an energy
   evaluation is truncated at 1e-14 times its largest singular value (the
   reference: 1e-14 absolute, compress_mps!) and divided by its norm after
   each row absorption.  The environments enter the ALS solve and the
   energy only through ratios, so the results are the reference's whenever
   no boundary singular value lies between the two thresholds.  The v4
   update rescales the state by 1/||S|| at every bond, so on an 8x8 lattice
   <psi|psi> drifts by ~30 orders of magnitude within one sweep and, without
   the flag, the boundaries collapse to bond dimension 1.  Off by default
   (reference behaviour). */
/* PEPS_STEP_POSITIVE_COST has no counterpart in the reference either: an ALS
   refinement is accepted only if its cost is also >= 0.  With truncated
   boundaries the environment need not be positive semidefinite, the cost
   can then be driven below zero along its negative directions, and the
   reference's test (cost < initial cost, finite_peps_fullupdate_v4.jl:145-153)
   accepts such a bond, which on an 8x8 lattice zeroes the state within a
   sweep.  Off by default (reference behaviour). */
enum { PEPS_STEP_RENORM = 1, PEPS_STEP_POSITIVE_COST = 2 };

typedef struct peps_params {
  double g;        /* gauge coupling g (the reference takes g, not g^2) */
  double t_hop;    /* hopping t */
  double m;        /* staggered mass m */
  double tau;      /* Trotter step tau */
  int32_t D_max;   /* maximal PEPS bond dimension */
  int32_t chi;     /* boundary-MPS bond dimension */
  int32_t n_als;   /* ALS sweeps (exactly n_als when als_tol == 0) */
  int32_t flags;    /* PEPS_STEP_* bits, 0 for the reference's behaviour */
  double als_tol;  /* > 0: stop ALS when the relative cost change < als_tol */
  double mu;       /* chemical potential mu n_f (trotter_step!'s mu,
                      finite_peps_ground_state.jl:500): PEPS_STEP_DENSE and
                      PEPS_STEP_REALTIME only, 0 for every other kind */
} peps_params;

/* ── context ──────────────────────────────────────────────────────────────── */
int32_t peps_ctx_create(int32_t device, peps_ctx** out);
int32_t peps_ctx_destroy(peps_ctx* ctx);
const char* peps_last_error(const peps_ctx* ctx);
/* number of warnings: the reference's @warn conditions (‖Θ‖≈0 in
   _recanonicalise_*!, finite_peps_fullupdate_v4.jl:236-238,336-338) and the
   factorisation iteration caps reached (peps_set_solver_limits) */
int64_t peps_warning_count(const peps_ctx* ctx);
/* contraction precision: FP64 MFMA, FP32 MFMA, BF16 MFMA (FP32 accumulate) */
int32_t peps_set_tier(peps_ctx* ctx, int32_t tier);
/* boundary-MPS compression (compress_mps!, finite_peps_boundary_mps.jl:70):
   0 auto (Gram form when the exact product bond exceeds 1024), 1 the QR/SVD
   sweep on the explicit product, 2 the Gram form: left Gram matrices of the
   factored product and top-chi eigenvectors of Y^H G Y — the same truncated
   MPS without forming the chi D^2 x D^2 x chi D^2 product. */
int32_t peps_set_compress_mode(peps_ctx* ctx, int32_t mode);
/* Singular-value cut of the boundary compression in the steps and energy
   evaluations (compress_mps!'s 1e-14, finite_peps_boundary_mps.jl:93, the
   default; relative to each bond's largest singular value under
   PEPS_STEP_RENORM).  No reference counterpart: a reduced-precision tier
   can cut at its own rounding level instead of keeping noise above 1e-14.
   0 < cut < 1. */
int32_t peps_set_compress_cut(peps_ctx* ctx, double cut);
/* Iteration caps of the factorisations (no reference counterpart; the
   reference's LAPACK calls have their own): Chebyshev subspace iterations
   per top-k eigenproblem of the Gram-form compression (default 25) and
   one-sided Jacobi sweeps per SVD (default 80).  Reaching a cap counts a
   warning (peps_warning_count); an unconverged subspace iteration is then
   answered by a direct eigendecomposition (matrices up to 4096). */
int32_t peps_set_solver_limits(peps_ctx* ctx, int32_t subspace_max_it, int32_t jacobi_max_sweeps);
/* per-kernel-class HIP-event timing and the collected counters: on = 0 off,
   1 every class, 2 only the MFMA GEMM classes of the precision tiers
   (gemm_f64_tiled, gemm_f32_tiled, gemm_bf16): two events per timed launch
   cost about 4% of a full-update step when every class is timed */
int32_t peps_set_timing(peps_ctx* ctx, int32_t on);
int32_t peps_stats_count(peps_ctx* ctx, int32_t* n);
int32_t peps_stats_get(peps_ctx* ctx, int32_t i, char* name, int32_t name_len, double* flops,
                       double* bytes, double* ms, int64_t* launches);
int32_t peps_stats_reset(peps_ctx* ctx);
int32_t peps_sync(peps_ctx* ctx);

/* ── state: GaugePEPS (finite_peps_gauge_invariant.jl:82-88) ─────────────── */
/* g_charges: nx*ny background charges, ix fastest.  gd in [2, 5]
   (PEPS_EUNSUPPORTED otherwise: the interior physical dimension 2 gd^2 is
   limited to 64 by the Gauss-mask kernels). */
int32_t peps_state_create(peps_ctx* ctx, int32_t nx, int32_t ny, int32_t gd,
                          const int64_t* g_charges, int32_t is_complex, peps_state** out);
int32_t peps_state_destroy(peps_state* st);
/* tensors[ix,iy] = T (p,l,r,u,d) */
int32_t peps_state_set_site(peps_state* st, int32_t ix, int32_t iy, const int64_t dims[5],
                            const double* data);
int32_t peps_state_site_dims(peps_state* st, int32_t ix, int32_t iy, int64_t dims[5]);
int32_t peps_state_get_site(peps_state* st, int32_t ix, int32_t iy, double* data);
/* λh/qh (orient PEPS_H) or λv/qv (PEPS_V) of bond (ix,iy) */
int32_t peps_state_set_bond(peps_state* st, int32_t orient, int32_t ix, int32_t iy, int64_t D,
                            const double* lambda, const int64_t* flux);
int32_t peps_state_bond_dim(peps_state* st, int32_t orient, int32_t ix, int32_t iy, int64_t* D);
int32_t peps_state_get_bond(peps_state* st, int32_t orient, int32_t ix, int32_t iy,
                            double* lambda, int64_t* flux);
/* apply_mask! (finite_peps_gauge_invariant.jl:139) */
int32_t peps_apply_mask(peps_state* st, int32_t ix, int32_t iy);
/* forbidden_norm (finite_peps_gauge_invariant.jl:154) */
int32_t peps_forbidden_norm(peps_state* st, double* out);
/* mean_nf with simple-update weights (finite_peps_ground_state.jl:550-578) */
int32_t peps_mean_nf(peps_state* st, double* out);

/* ── result lists (boundary MPS, environment tuples, factors) ────────────── */
int32_t peps_list_len(const peps_tensor_list* l, int32_t* n);
int32_t peps_list_dims(const peps_tensor_list* l, int32_t i, int32_t* rank, int64_t dims[8],
                       int32_t* is_complex);
int32_t peps_list_get(const peps_tensor_list* l, int32_t i, double* data);
int32_t peps_list_destroy(peps_tensor_list* l);

/* ── kernel-level entry points ────────────────────────────────────────────── */
/* _dl_traced (finite_peps_full_update.jl:79-91): T (p,l,r,u,d) -> E (l²,r²,u²,d²) */
int32_t peps_dl_traced(peps_ctx* ctx, int32_t is_complex, const int64_t dims[5], const double* T,
                       double* E);
/* _dl_open (finite_peps_full_update.jl:97-109): T (p,l,r,u,d) ->
   E (p, p', l², r², u², d²) = T[p,..] conj(T[p',..]) with the bra index
   fast in every merged pair (the physical indices left open) */
int32_t peps_dl_open(peps_ctx* ctx, int32_t is_complex, const int64_t dims[5], const double* T, double* E);
/* absorb_row_from_top / _bottom (finite_peps_boundary_mps.jl:128,156).
   mps_dims: 3n, row_dims: 4n. */
int32_t peps_absorb_row(peps_ctx* ctx, int32_t from_top, int32_t is_complex, int32_t n,
                        const int64_t* mps_dims, const double* const* mps,
                        const int64_t* row_dims, const double* const* row,
                        peps_tensor_list** out);
/* dmrg_ground_state (mps_lgt.jl:418-467) with two_site_update! (:384-416):
   two-site DMRG of a real MPO of n >= 2 sites, W[p] (Dl, Dr, d, d) column-major
   (bra index, then ket) with dimensions mpo_dims[4p..4p+3], from the start
   MPS psi0[p] (Dl, Dr, d), dimensions psi_dims[3p..3p+2] (the reference
   draws a random one when none is given; the caller passes it here).  The
   start is compressed to bond D and normalised; bonds grow up to D, at most
   nsweeps sweeps, stopping once a sweep changes the energy by less than etol;
   each local ground state by restarted Lanczos (krylovdim vectors, residual
   <= tol, at most maxiter restarts), split by an SVD keeping values above
   1e-12 of the largest.  Returns the last update's energy, the sweeps run
   (sweeps may be NULL) and the MPS.  Real tensors only. */
int32_t peps_dmrg(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                  const int64_t* psi_dims, const double* const* psi0, int32_t D, int32_t nsweeps, double tol,
                  int32_t maxiter, int32_t krylovdim, double etol, double* energy, int32_t* sweeps,
                  peps_tensor_list** psi_out);
/* mpo_expect (mps_lgt.jl:308-315): <psi|W|psi> of a real MPS and MPO (not
   divided by <psi|psi>) */
int32_t peps_mpo_expect(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                        const int64_t* psi_dims, const double* const* psi, double* out);
/* The same expectation for a real or complex MPS (psi_complex != 0: every
   tensor interleaved re/im); the real part of <psi|W|psi>. */
int32_t peps_mpo_expect_c(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                          int32_t psi_complex, const int64_t* psi_dims, const double* const* psi, double* out);
/* krylov_evolve_step (mps_lgt.jl:546-586), nsteps times, as evolve_mps
   (mps_quench.jl:42-62) calls it: psi <- exp(-i dt H) psi by Lanczos on the
   MPS manifold with krylov_m vectors (full reorthogonalisation), each H v by
   the zip-up MPO application (mpo_apply_mps_zipup, mps_lgt.jl:239-257), sums
   by mps_axpby (:259-281) and compression by mps_compress! (:283-303), every
   truncation keeping at most Dmax values above eps times the largest; the
   state is renormalised after each step.  H real (Dl, Dr, d, d), psi0 real
   or complex (Dl, Dr, d); returns a complex MPS.  nsteps = 0 returns psi0. */
int32_t peps_krylov_evolve(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                           int32_t psi_complex, const int64_t* psi_dims, const double* const* psi0, double dt,
                           int32_t nsteps, int32_t krylov_m, int32_t Dmax, double eps, peps_tensor_list** psi_out);
/* mps_local_expect (mps_lgt.jl:317-330) for nops (site, operator) pairs:
   out[i] = Re <psi|O_i at chain site sites[i] (0-based)|psi> / <psi|psi>,
   O_i real (d, d) with op_dims[2 i], op_dims[2 i + 1] = d. */
int32_t peps_mps_local_expect(peps_ctx* ctx, int32_t n, int32_t psi_complex, const int64_t* psi_dims,
                              const double* const* psi, int32_t nops, const int32_t* sites, const int64_t* op_dims,
                              const double* const* ops, double* out);
/* compress_mps! (finite_peps_boundary_mps.jl:70-105) */
int32_t peps_compress_mps(peps_ctx* ctx, int32_t is_complex, int32_t n, const int64_t* dims,
                          const double* const* mps, int32_t chi, double cutoff,
                          peps_tensor_list** out);
/* boundary_top / boundary_bottom (finite_peps_boundary_mps.jl:183,207) */
int32_t peps_boundary(peps_state* st, int32_t from_top, int32_t stop_row, int32_t chi,
                      peps_tensor_list** out);
/* peps_norm2 (finite_peps_boundary_mps.jl:228): out = (re, im) */
int32_t peps_norm2(peps_state* st, int32_t chi, double out[2]);
/* bond_environment_h / _v (finite_peps_boundary_mps.jl:389,507): list = env, RA, RB, QA, QB */
int32_t peps_bond_env(peps_state* st, int32_t orient, int32_t ix, int32_t iy, int32_t chi,
                      peps_tensor_list** out);
/* charge_svd (finite_peps_gauge_invariant.jl:236): list = U, Vt, S; charges[Dk] */
int32_t peps_charge_svd(peps_ctx* ctx, int32_t is_complex, int64_t nrow, int64_t ncol,
                        const double* M, const int64_t* row_charge, const int64_t* col_charge,
                        int32_t D_max, double reg, int64_t* Dk, int64_t* charges,
                        peps_tensor_list** out);
/* als_truncate (finite_peps_fullupdate_v4.jl:94): theta (nA,dpL,dpR,nB),
   env (nA,nB,nA,nB); list = X, Y; bond_flux[Dp]; info = cost0, cost1, accepted */
int32_t peps_als_truncate(peps_ctx* ctx, int32_t is_complex, const int64_t th_dims[4],
                          const double* theta, const double* env, int32_t D_max,
                          const int64_t* eR_of_pL, int32_t n_als, double reg, int64_t* Dp,
                          int64_t* bond_flux, double info[3], peps_tensor_list** out);
/* als_truncate with an optional initial bond and a tolerance (no reference
   counterpart for these two): X0 (nA, dpL, Dp) and Y0 (nB, dpR, Dp), both
   NULL for the reference's SVD initialisation, otherwise shaped like it
   (the mask of X is applied); als_tol > 0 stops when the relative cost
   change of a sweep is below it.  The accept rule is the reference's: the
   result is kept only if its cost is below that of the SVD initialisation.
   info = cost0 (SVD init), cost1, accepted, sweeps run. */
int32_t peps_als_truncate_init(peps_ctx* ctx, int32_t is_complex, const int64_t th_dims[4],
                               const double* theta, const double* env, int32_t D_max,
                               const int64_t* eR_of_pL, int32_t n_als, double reg, double als_tol,
                               const double* X0, const double* Y0, int64_t* Dp, int64_t* bond_flux,
                               double info[4], peps_tensor_list** out);
/* simple_update_bond_h! / _v! (finite_peps_gauge_invariant.jl:296,412);
   gate: (dL*dR)^2, kron(L,R) ordering (right site fast) */
int32_t peps_su_bond(peps_state* st, int32_t orient, int32_t ix, int32_t iy, const double* gate,
                     int32_t D_max);
/* full_update_bond_h_v4! / _v_v4! (finite_peps_fullupdate_v4.jl:165,274) */
int32_t peps_fu_bond_v4(peps_state* st, int32_t orient, int32_t ix, int32_t iy,
                        const double* gate, int32_t D_max, int32_t chi, int32_t n_als);
/* trotter_step_v4! (finite_peps_fullupdate_v4.jl:415), trotter_step_gauge!
   (finite_peps_gauge_invariant.jl:523) or, for PEPS_STEP_DENSE, the dense
   trotter_step! with plaquettes (finite_peps_ground_state.jl:498-541; uses
   g, t_hop, m, tau and D_max as D_trunc) or, for PEPS_STEP_FULL_V2,
   trotter_step_full! (finite_peps_full_update.jl:550-575, use_gauss=false) */
int32_t peps_trotter_step(peps_state* st, int32_t kind, const peps_params* p);
/* One peps_trotter_step on each of n independent states at once (no
   reference counterpart: the reference runs independent instances as a
   SLURM array).  Every state must belong to its own context: each context's
   stream runs in its own host thread, so the latency-bound factorisations of
   one instance overlap the work of the others on the same GPU.  status[i]
   (optional) receives each state's result; the return value is the first
   non-OK status, or PEPS_EARG for shared contexts. */
int32_t peps_batch_trotter_step(peps_state* const* states, int32_t n, int32_t kind, const peps_params* p,
                                int32_t* status);
/* Imaginary-time ground-state search with the reference's tau annealing:
   ite_ground_state_v4 (finite_peps_fullupdate_v4.jl:440-467, kind
   PEPS_STEP_V4), ite_ground_state_v3 (finite_peps_gauge_invariant.jl:548-577,
   PEPS_STEP_GAUGE) and ite_ground_state_v2 (finite_peps_full_update.jl:
   582-629, PEPS_STEP_FULL_V2).  n_ite steps: n/8 at tau 0.1, n/8 at 0.05,
   n/4 at 0.02, n/6 at 0.01, n/6 at 0.005 and the rest at 0.001 (integer
   division); p->tau is ignored.  The state is the initial PEPS
   (init_gauge_peps / init_finite_peps_v2, built by the caller) and is
   stepped in place; steps_done (optional) receives the steps completed,
   also on failure.  After a PEPS_STEP_V4 run peps_last_als_stats covers the
   whole run.  Progress lines (PEPS_HIP_VERBOSE=1) every 25 steps. */
int32_t peps_ite_ground_state(peps_state* st, int32_t kind, const peps_params* p, int64_t n_ite,
                              int64_t* steps_done);
/* apply_site_op! (finite_peps_quench.jl:228-234): T(ix,iy) <- O * reshape(T,
   dp, :) with O (dp x dp, column-major) in the state's element type.  The
   flux labels are undefined afterwards (the state is then label-free, as
   after a dense update). */
int32_t peps_apply_site_op(peps_state* st, int32_t ix, int32_t iy, const double* O);
/* ALS statistics of the last PEPS_STEP_V4 step: bonds, accepted, total sweeps */
int32_t peps_last_als_stats(peps_state* st, int64_t out[3]);
/* compute_energy_v4 (finite_peps_fullupdate_v4.jl:390): total energy */
int32_t peps_energy_v4(peps_state* st, const peps_params* p, double* E);

/* ── dense simple update with plaquettes (finite_peps_ground_state.jl) ──── */
/* These treat the state as a FinitePEPS: no Gauss mask, no flux labels.
   After any dense update the bond labels are undefined (read back as 0)
   and the gauge-invariant entries (peps_apply_mask, peps_forbidden_norm,
   peps_su_bond, peps_fu_bond_v4, peps_trotter_step V4/GAUGE) return
   PEPS_EARG for that state. */
/* update_bond_h! / update_bond_v! (finite_peps_ground_state.jl:175-382):
   gate (dL*dR)^2 with the right/upper site fastest, state element type */
int32_t peps_su_bond_dense(peps_state* st, int32_t orient, int32_t ix, int32_t iy, const double* gate,
                           int32_t D_trunc);
/* full_update_bond_h! / full_update_bond_v! (finite_peps_full_update.jl:346-543):
   the same label-free update with cut 1e-10, lambda = S/max(||S||, 1e-15)
   and transverse weights divided out with +1e-10.  use_gauss must be 0
   (PEPS_EUNSUPPORTED otherwise: the reference's projector path recovers dg
   from the site dimension incorrectly; its driver never enables it). */
int32_t peps_fu_bond_v2(peps_state* st, int32_t orient, int32_t ix, int32_t iy, const double* gate,
                        int32_t D_trunc, int32_t use_gauss);
/* decompose_plaquette_LShape(exp(-tau H_plaquette_gauge(gd; g)), gd; cutoff)
   (u1_lgt_hamiltonian.jl:283-289,728-768; auxiliary bonds capped at 64).
   Pass GA == NULL to query Da, Db only.  Outputs real column-major
   GA (gd,gd,gd,gd,Da,Db), GB (gd,gd,Da), GC (gd,gd,Db). */
int32_t peps_plaquette_gate(peps_ctx* ctx, int32_t gd, double g, double tau, double cutoff,
                            int32_t* Da, int32_t* Db, double* GA, double* GB, double* GC);
/* update_plaquette! (finite_peps_ground_state.jl:405-483) on the L-shape
   A=(ix,iy), B=(ix+1,iy), C=(ix,iy+1); GA/GB/GC shaped as above, in the
   state's element type */
int32_t peps_update_plaquette(peps_state* st, int32_t ix, int32_t iy, int32_t Da, int32_t Db,
                              const double* GA, const double* GB, const double* GC, int32_t D_trunc);
/* expect_site (finite_peps_ground_state.jl:549-567): O (dp x dp) in the
   state's element type */
int32_t peps_expect_site(peps_state* st, int32_t ix, int32_t iy, const double* O, double* out);
/* plaquette_energy (finite_peps_ground_state.jl:589-684) */
int32_t peps_plaquette_energy(peps_state* st, int32_t ix, int32_t iy, double g, double* out);
/* total_energy (finite_peps_ground_state.jl:690-805): uses g, t_hop, m */
int32_t peps_total_energy(peps_state* st, const peps_params* p, double* out);

/* <W> = <psi| Ud_b U_l Ud_r U_t |psi> / <psi|psi> on the plaquette with
   lower-left site (ix, iy), from boundary-MPS environments of bond
   dimension chi (the two-row strip between boundary_top(iy+1) and
   boundary_bottom(iy)); out = (re, im).  No reference counterpart: the
   boundary-MPS form of plaquette_energy (finite_peps_ground_state.jl:589,
   which uses simple-update weights); the plaquette energy is -Re<W>/g^2. */
int32_t peps_plaquette_expect(peps_state* st, int32_t ix, int32_t iy, int32_t chi, double out[2]);
/* bond_entropy (finite_peps_quench.jl:121-127): von Neumann entropy of
   p_i = lambda_i^2 / sum lambda^2 (0 when the sum is below 1e-30) */
int32_t peps_bond_entropy(const double* lambda, int64_t n, double* out);
/* measure_all_finite (finite_peps_quench.jl:166-194) without the time
   stamp: out = nf_mean, nf_even, nf_odd, E2_mean, S_h_mean, S_v_mean,
   S_mean; nf_grid / E2_grid (nx x ny, ix fastest) may be NULL */
int32_t peps_measure_all(peps_state* st, double out[7], double* nf_grid, double* E2_grid);

/* ── learned ALS warm start (no counterpart in the reference) ──────────────
   A multilayer perceptron in -> width -> width -> out (ReLU hidden layers,
   linear output) trained and evaluated on the GPU by the library's own
   kernels (products in the FP32 MFMA tier).  The batch runs along columns:
   X (n_in x batch), Y and T (n_out x batch), column-major.  Parameters are
   flattened as W1 (width x n_in), b1, W2 (width x width), b2, W3 (n_out x
   width), b3; loss = 1/(2 batch) ||Y - T||^2; training uses Adam (0.9,
   0.999, 1e-8).  It proposes the initial bond of the ALS solve
   (finite_peps_fullupdate_v4.jl:99-111 is the slot it replaces); the
   reference's SVD initialisation stays the fallback. */
int32_t peps_mlp_create(peps_ctx* ctx, int32_t n_in, int32_t width, int32_t n_out, uint64_t seed,
                        peps_mlp** out);
int32_t peps_mlp_destroy(peps_mlp* m);
int32_t peps_mlp_param_count(const peps_mlp* m, int64_t* n);
int32_t peps_mlp_get_params(peps_mlp* m, double* out);
int32_t peps_mlp_set_params(peps_mlp* m, const double* in);
int32_t peps_mlp_forward(peps_mlp* m, int64_t batch, const double* X, double* Y);
/* loss and the gradient of every parameter (flattened as above) */
int32_t peps_mlp_gradients(peps_mlp* m, int64_t batch, const double* X, const double* T, double* loss,
                           double* grad);
/* one Adam step; loss = the loss before the step */
int32_t peps_mlp_train_step(peps_mlp* m, int64_t batch, const double* X, const double* T, double lr,
                            double* loss);

/* The contraction GEMM itself (no counterpart in the reference; exposed so
   callers can validate the MFMA path of the current tier):
   C (M x N) = op(A) op(B); A stored (K x M) if a_kfast else (M x K);
   B stored (K x N) if b_kfast else (N x K); conj flags conjugate an operand. */
int32_t peps_gemm(peps_ctx* ctx, int32_t is_complex, int64_t M, int64_t N, int64_t K,
                  const double* A, int32_t a_kfast, int32_t conjA, const double* B,
                  int32_t b_kfast, int32_t conjB, double* Cout);

/* peps_gemm for callers that cannot allocate device memory: code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc.  No
   counterpart in the reference.  A, B and Cout are DEVICE addresses on the
   context's device, owned by the caller (checked against their allocations:
   a host pointer or a short buffer returns PEPS_EARG); same layouts and
   result as peps_gemm.  The call only enqueues kernels on the context's
   stream (peps_sync waits) and allocates nothing: the split-K partial
   products go to a fixed 64 MiB array that the library holds on each device
   from the context's creation until the process ends (the split count is
   lowered until they fit, which can change the rounding of the sum relative
   to peps_gemm), and the FP32 tier converts operands inside the GEMM kernel
   instead of casting them into buffers first.  Calls that overlap on one
   device share that array and race: serialise them (one stream, or events).
   Extents must stay below 2^31 (PEPS_EUNSUPPORTED otherwise). */
int32_t peps_gemm_noalloc(peps_ctx* ctx, int32_t is_complex, int64_t M, int64_t N, int64_t K,
                          const double* A, int32_t a_kfast, int32_t conjA, const double* B,
                          int32_t b_kfast, int32_t conjB, double* Cout);

/* The engine's tensor permutation (no counterpart in the reference; exposed
   for validation): out.dims[i] = dims[perm[i]], 0-based perm. */
int32_t peps_permute(peps_ctx* ctx, int32_t is_complex, int32_t rank, const int64_t* dims,
                     const int32_t* perm, const double* in, double* out);

/* The same GEMM accumulating into C: C (M x N) = op(A) op(B) + beta C (the
   form the chunked contractions use). */
int32_t peps_gemm_acc(peps_ctx* ctx, int32_t is_complex, int64_t M, int64_t N, int64_t K,
                      const double* A, int32_t a_kfast, int32_t conjA, const double* B,
                      int32_t b_kfast, int32_t conjB, double beta, double* Cinout);

/* Hermitian eigendecomposition used by the ALS solve (_solve_herm,
   finite_peps_fullupdate_v4.jl:37-45): Householder tridiagonalisation and
   divide and conquer, one matrix after the other (real n <= 1024: the
   library's one-launch tridiagonal reduction and reflector kernels around
   rocSOLVER stedc; otherwise rocSOLVER syevd/heevd).
   A: nb contiguous n x n Hermitian matrices (lower triangles read); outputs
   V (same layout) and w (nb*n, ascending per matrix). */
int32_t peps_eigh(peps_ctx* ctx, int32_t is_complex, int32_t n, int32_t nb, const double* A,
                  double* V, double* w);

/* _solve_herm (finite_peps_fullupdate_v4.jl:37-45): X = V diag(inv) V^H S for
   the Hermitian part of N (n x n) and the nrhs columns of S, inv_i = 1/w_i
   where w_i > reg * max(w_max, 1e-30) and 0 otherwise (the ALS solve).  Real
   matrices whose couplings between groups of rows total at most eps ||N||_F
   (the gauge sectors of the ALS normal equations) are solved block by block
   with the floor taken against the largest eigenvalue of all blocks. */
int32_t peps_solve_herm(peps_ctx* ctx, int32_t is_complex, int64_t n, int64_t nrhs, const double* N,
                        const double* S, double reg, double* X);

/* Thin SVD used by compress_mps!/charge_svd (LinearAlgebra.svd in the
   reference): QR reduction + one-sided Jacobi kernel.  A (m x n); outputs
   U (m x k), S (k, descending), Vt (k x n), k = min(m, n). */
int32_t peps_svd(peps_ctx* ctx, int32_t is_complex, int64_t m, int64_t n, const double* A,
                 double* U, double* S, double* Vt);

/* Thin QR used by the boundary compression, the QR reductions of the bond
   environments and the subspace iterations (LinearAlgebra.qr in the
   reference, finite_peps_boundary_mps.jl:78-84, :325-351): Householder
   reflectors with LAPACK's sign convention; real panels wider than 64
   columns in blocked form.  A (m x n); outputs Q (m x k), R (k x n),
   k = min(m, n). */
int32_t peps_qr(peps_ctx* ctx, int32_t is_complex, int64_t m, int64_t n, const double* A, double* Q, double* R);

/* ── model (u1_lgt_hamiltonian.jl:829-863, :396) ─────────────────────────── */
/* H_merged_{h,v}_site (tau == 0) or its gate exp(-tau H); pass out == NULL to
   query n.  out is n*n real, column-major. */
int32_t peps_bond_operator(int32_t orient, int32_t ix, int32_t iy, int32_t nx, int32_t ny,
                           int32_t gd, double g, double t_hop, double m, double tau, int32_t* n,
                           double* out);

/* ── exact diagonalisation in one gauge sector (finite_ed.jl) ─────────────
   H = sum (g^2/2)(E_R^2 + E_U^2) + m sum (-1)^(ix+iy) n_f - t sum (hops with
   the link raised on a forward hop), no plaquette term (finite_ed.jl:6-10).
   States are keyed as in finite_ed.jl:79-95 (n_f bits, then E + dg in
   2-bit link fields for dg = 1, wider fields for dg = 2, 3). */
/* build_basis (finite_ed.jl:125-233): g_charges nx*ny, ix fastest; the
   basis is enumerated on the GPU in the reference's order.  Lattices from
   2x2 to 32 sites, dg in [1, 3], keys within 64 bits. */
int32_t peps_ed_create(peps_ctx* ctx, int32_t nx, int32_t ny, int32_t dg, const int64_t* g_charges,
                       peps_ed** out);
int32_t peps_ed_destroy(peps_ed* ed);
int32_t peps_ed_dim(const peps_ed* ed, int64_t* dim);
/* the encoded basis states, in basis order (dim entries) */
int32_t peps_ed_keys(peps_ed* ed, uint64_t* keys);
/* build_hamiltonian (finite_ed.jl:252-326) */
int32_t peps_ed_hamiltonian(peps_ed* ed, double g, double t, double m);
/* y = H x for real vectors of the basis dimension */
int32_t peps_ed_apply(peps_ed* ed, const double* x, double* y);
/* find_ground_state (finite_ed.jl:333-341): restarted Lanczos (Krylov
   dimension 50, at most 300 restarts) until ||H psi - E0 psi|| <= tol;
   psi (optional, dim doubles) receives the normalised real ground state.
   PEPS_ESOLVER if the tolerance is not reached. */
int32_t peps_ed_ground_state(peps_ed* ed, double tol, double* E0, double* psi);
/* measure_ED (finite_ed.jl:369-395) of psi (NULL: the ground state):
   out = {nf_mean, nf_even, nf_odd, E2_mean}; nf_grid, E2_grid (optional)
   nx*ny, ix fastest. */
int32_t peps_ed_measure(peps_ed* ed, const double* psi, double out[4], double* nf_grid, double* E2_grid);
/* evolve_ED! (finite_ed.jl:405-430): psi <- exp(-i dt H) psi, renormalised,
   nsteps times, each step by KrylovKit's exponentiate as the reference calls
   it (Lanczos projections of at most krylovdim vectors, error estimate <= tol,
   the interval halved until it is); psi0 and psi_out (may be NULL) are
   interleaved complex vectors of the basis dimension.  obs gets (nsteps + 1)
   rows of {nf_mean, nf_even, nf_odd, E2_mean} (t = 0 first); nf_grids and
   E2_grids (may be NULL) (nsteps + 1) site grids of nx*ny (ix fastest);
   projections (may be NULL) the Krylov projections used in all. */
int32_t peps_ed_evolve(peps_ed* ed, const double* psi0, double dt, int32_t nsteps, int32_t krylovdim, double tol,
                       double* psi_out, double* obs, double* nf_grids, double* E2_grids, int32_t* projections);

#ifdef __cplusplus
}
#endif
#endif /* LIBPEPS_HIP_PEPS_H_ */
