// C ABI of libpeps_hip.so (declared in include/peps.h).
#include "peps.h"

#include <cmath>
#include <cstring>
#include <set>
#include <thread>

#include "ed.h"
#include "engine.h"
#include "mlp.h"

// States and result lists hold a reference to the context, so device
// buffers always return to a live allocator whatever order the caller
// releases handles in.  `ref` is declared first: it is released last.
struct peps_ctx {
  explicit peps_ctx(int dev) : c(std::make_shared<peps::Ctx>(dev)) {}
  std::shared_ptr<peps::Ctx> c;
};

struct peps_state {
  std::shared_ptr<peps::Ctx> ref;
  peps::State s;
  peps::GateCache gc;
  int64_t als_bonds = 0, als_accepted = 0, als_sweeps = 0;
};

struct peps_mlp {
  std::shared_ptr<peps::Ctx> ref;
  peps::Mlp m;
  peps_mlp(std::shared_ptr<peps::Ctx> c, int nin, int w, int nout, uint64_t seed)
      : ref(std::move(c)), m(*ref, nin, w, nout, seed) {}
};

struct peps_tensor_list {
  std::shared_ptr<peps::Ctx> ref;
  std::vector<peps::Tensor> t;
};

struct peps_ed {
  std::shared_ptr<peps::Ctx> ref;
  std::unique_ptr<peps::EDBasis> b;  // released before the context
};

namespace {

using peps::Tensor;

template <typename F>
int32_t guard(peps::Ctx* c, F&& f) {
  try {
    f();
    return PEPS_OK;
  } catch (const peps::Error& e) {
    if (c) c->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (c) c->last_error = "host allocation failed";
    return PEPS_ENOMEM;
  } catch (const std::exception& e) {
    if (c) c->last_error = e.what();
    return PEPS_EARG;
  }
}

peps_tensor_list* make_list(const std::shared_ptr<peps::Ctx>& c, std::vector<Tensor> t) {
  auto* l = new peps_tensor_list;
  l->ref = c;
  l->t = std::move(t);
  return l;
}

void check_site(peps::State& s, int ix, int iy) {
  PEPS_CHECK(ix >= 1 && ix <= s.nx && iy >= 1 && iy <= s.ny, peps::PS_EARG, "site out of range");
}

std::vector<double>& bond_lam(peps::State& s, int orient, int ix, int iy) {
  if (orient == PEPS_H) {
    PEPS_CHECK(ix >= 1 && ix < s.nx && iy >= 1 && iy <= s.ny, peps::PS_EARG, "h bond out of range");
    return s.LH(ix, iy);
  }
  PEPS_CHECK(ix >= 1 && ix <= s.nx && iy >= 1 && iy < s.ny, peps::PS_EARG, "v bond out of range");
  return s.LV(ix, iy);
}
std::vector<int64_t>& bond_q(peps::State& s, int orient, int ix, int iy) {
  bond_lam(s, orient, ix, iy);
  return orient == PEPS_H ? s.QH(ix, iy) : s.QV(ix, iy);
}

void require_labels(const peps::State& s) {
  PEPS_CHECK(s.labelled, peps::PS_EARG,
             "the state's flux labels are undefined after a dense (label-free) update");
}

peps::Params to_params(const peps_params* p) {
  PEPS_CHECK(p != nullptr, peps::PS_EARG, "null params");
  PEPS_CHECK(p->D_max >= 1 && p->chi >= 1 && p->n_als >= 0, peps::PS_EARG, "invalid params");
  peps::Params q;
  q.g = p->g;
  q.t = p->t_hop;
  q.m = p->m;
  q.tau = p->tau;
  q.Dmax = p->D_max;
  q.chi = p->chi;
  q.n_als = p->n_als;
  q.als_tol = p->als_tol;
  q.mu = p->mu;
  PEPS_CHECK(std::isfinite(q.mu), peps::PS_EARG, "mu must be finite");
  PEPS_CHECK((p->flags & ~int32_t(PEPS_STEP_RENORM | PEPS_STEP_POSITIVE_COST)) == 0, peps::PS_EARG,
             "unknown step flags");
  q.renorm = (p->flags & PEPS_STEP_RENORM) != 0;
  q.positive_cost = (p->flags & PEPS_STEP_POSITIVE_COST) != 0;
  return q;
}

Tensor gate_from_host(peps::State& s, int orient, int ix, int iy, const double* gate) {
  int jx = orient == PEPS_H ? ix + 1 : ix, jy = orient == PEPS_H ? iy : iy + 1;
  int64_t n = int64_t(peps::site_dims(ix, iy, s.nx, s.ny, s.gd).d) * peps::site_dims(jx, jy, s.nx, s.ny, s.gd).d;
  PEPS_CHECK(gate != nullptr, peps::PS_EARG, "null gate");
  return s.c->upload({n, n}, s.cplx, gate);
}

}  // namespace

extern "C" {

int32_t peps_ctx_create(int32_t device, peps_ctx** out) {
  if (!out) return PEPS_EARG;
  *out = nullptr;
  return guard(nullptr, [&] { *out = new peps_ctx(device); });
}

int32_t peps_ctx_destroy(peps_ctx* ctx) {
  delete ctx;
  return PEPS_OK;
}

const char* peps_last_error(const peps_ctx* ctx) { return ctx ? ctx->c->last_error.c_str() : "null context"; }
int64_t peps_warning_count(const peps_ctx* ctx) { return ctx ? ctx->c->warnings : 0; }

int32_t peps_set_tier(peps_ctx* ctx, int32_t tier) {
  if (!ctx || tier < 0 || tier > 2) return PEPS_EARG;
  ctx->c->tier = peps::Tier(tier);
  return PEPS_OK;
}

int32_t peps_set_compress_mode(peps_ctx* ctx, int32_t mode) {
  if (!ctx || mode < 0 || mode > 2) return PEPS_EARG;
  ctx->c->compress_mode = mode;
  return PEPS_OK;
}

int32_t peps_set_compress_cut(peps_ctx* ctx, double cut) {
  if (!ctx || !(cut > 0.0) || !std::isfinite(cut) || cut >= 1.0) return PEPS_EARG;
  ctx->c->bmps_cut = cut;
  return PEPS_OK;
}

int32_t peps_set_solver_limits(peps_ctx* ctx, int32_t subspace_max_it, int32_t jacobi_max_sweeps) {
  if (!ctx || subspace_max_it < 1 || jacobi_max_sweeps < 1) return PEPS_EARG;
  ctx->c->subspace_max_it = subspace_max_it;
  ctx->c->jacobi_max_sweeps = jacobi_max_sweeps;
  return PEPS_OK;
}

int32_t peps_set_timing(peps_ctx* ctx, int32_t on) {
  if (!ctx) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    PEPS_CHECK(on >= 0 && on <= 2, peps::PS_EARG, "set_timing: mode 0, 1 or 2");
    ctx->c->resolve_timing();
    ctx->c->timing = on != 0;
    ctx->c->timing_gemm_only = on == 2;
  });
}

int32_t peps_stats_count(peps_ctx* ctx, int32_t* n) {
  if (!ctx || !n) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    ctx->c->resolve_timing();
    *n = int32_t(ctx->c->stats.size());
  });
}

int32_t peps_stats_get(peps_ctx* ctx, int32_t i, char* name, int32_t name_len, double* flops,
                       double* bytes, double* ms, int64_t* launches) {
  if (!ctx || i < 0 || i >= int32_t(ctx->c->stats.size())) return PEPS_EARG;
  auto it = ctx->c->stats.begin();
  std::advance(it, i);
  if (name && name_len > 0) {
    std::strncpy(name, it->first.c_str(), size_t(name_len - 1));
    name[name_len - 1] = 0;
  }
  if (flops) *flops = it->second.flops;
  if (bytes) *bytes = it->second.bytes;
  if (ms) *ms = it->second.ms;
  if (launches) *launches = it->second.launches;
  return PEPS_OK;
}

int32_t peps_stats_reset(peps_ctx* ctx) {
  if (!ctx) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    ctx->c->resolve_timing();
    ctx->c->stats.clear();
  });
}

int32_t peps_sync(peps_ctx* ctx) {
  if (!ctx) return PEPS_EARG;
  return guard(ctx->c.get(), [&] { ctx->c->sync(); });
}

// ── state ────────────────────────────────────────────────────────────────────

int32_t peps_state_create(peps_ctx* ctx, int32_t nx, int32_t ny, int32_t gd,
                          const int64_t* g_charges, int32_t is_complex, peps_state** out) {
  if (!ctx || !out || !g_charges) return PEPS_EARG;
  *out = nullptr;
  return guard(ctx->c.get(), [&] {
    auto* st = new peps_state;
    st->ref = ctx->c;
    try {
      st->s.init(ctx->c.get(), nx, ny, gd, g_charges, is_complex != 0);
    } catch (...) {
      delete st;
      throw;
    }
    *out = st;
  });
}

int32_t peps_state_destroy(peps_state* st) {
  delete st;
  return PEPS_OK;
}

int32_t peps_state_set_site(peps_state* st, int32_t ix, int32_t iy, const int64_t dims[5],
                            const double* data) {
  if (!st || !dims || !data) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    check_site(st->s, ix, iy);
    auto sd = peps::site_dims(ix, iy, st->s.nx, st->s.ny, st->s.gd);
    PEPS_CHECK(dims[0] == sd.d, peps::PS_ESHAPE, "site tensor: physical dimension mismatch");
    for (int i = 0; i < 5; ++i) PEPS_CHECK(dims[i] >= 1, peps::PS_ESHAPE, "site tensor: empty leg");
    PEPS_CHECK((ix > 1 || dims[1] == 1) && (ix < st->s.nx || dims[2] == 1) &&
                   (iy < st->s.ny || dims[3] == 1) && (iy > 1 || dims[4] == 1),
               peps::PS_ESHAPE, "site tensor: boundary legs must have dimension 1");
    st->s.t(ix, iy) = st->ref->upload(std::vector<int64_t>(dims, dims + 5), st->s.cplx, data);
  });
}

int32_t peps_state_site_dims(peps_state* st, int32_t ix, int32_t iy, int64_t dims[5]) {
  if (!st || !dims) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    check_site(st->s, ix, iy);
    auto& t = st->s.t(ix, iy);
    PEPS_CHECK(!t.empty(), peps::PS_EARG, "site not set");
    for (int i = 0; i < 5; ++i) dims[i] = t.dims[i];
  });
}

int32_t peps_state_get_site(peps_state* st, int32_t ix, int32_t iy, double* data) {
  if (!st || !data) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    check_site(st->s, ix, iy);
    auto& t = st->s.t(ix, iy);
    PEPS_CHECK(!t.empty(), peps::PS_EARG, "site not set");
    st->ref->download(t, data);
  });
}

int32_t peps_state_set_bond(peps_state* st, int32_t orient, int32_t ix, int32_t iy, int64_t D,
                            const double* lambda, const int64_t* flux) {
  if (!st || !lambda || !flux || D < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    PEPS_CHECK(D <= 64, peps::PS_EUNSUPPORTED, "bond dimension > 64");
    bond_lam(st->s, orient, ix, iy).assign(lambda, lambda + D);
    bond_q(st->s, orient, ix, iy).assign(flux, flux + D);
  });
}

int32_t peps_state_bond_dim(peps_state* st, int32_t orient, int32_t ix, int32_t iy, int64_t* D) {
  if (!st || !D) return PEPS_EARG;
  return guard(st->ref.get(), [&] { *D = int64_t(bond_lam(st->s, orient, ix, iy).size()); });
}

int32_t peps_state_get_bond(peps_state* st, int32_t orient, int32_t ix, int32_t iy,
                            double* lambda, int64_t* flux) {
  if (!st) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    auto& l = bond_lam(st->s, orient, ix, iy);
    auto& q = bond_q(st->s, orient, ix, iy);
    if (lambda) std::memcpy(lambda, l.data(), l.size() * sizeof(double));
    if (flux) std::memcpy(flux, q.data(), q.size() * sizeof(int64_t));
  });
}

int32_t peps_apply_mask(peps_state* st, int32_t ix, int32_t iy) {
  if (!st) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    check_site(st->s, ix, iy);
    require_labels(st->s);
    peps::apply_mask(st->s, ix, iy);
  });
}

int32_t peps_forbidden_norm(peps_state* st, double* out) {
  if (!st || !out) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    require_labels(st->s);
    *out = peps::forbidden_norm(st->s);
  });
}

int32_t peps_mean_nf(peps_state* st, double* out) {
  if (!st || !out) return PEPS_EARG;
  return guard(st->ref.get(), [&] { *out = peps::mean_nf(st->s); });
}

// ── lists ────────────────────────────────────────────────────────────────────

int32_t peps_list_len(const peps_tensor_list* l, int32_t* n) {
  if (!l || !n) return PEPS_EARG;
  *n = int32_t(l->t.size());
  return PEPS_OK;
}

int32_t peps_list_dims(const peps_tensor_list* l, int32_t i, int32_t* rank, int64_t dims[8],
                       int32_t* is_complex) {
  if (!l || i < 0 || i >= int32_t(l->t.size()) || !rank || !dims) return PEPS_EARG;
  const auto& t = l->t[i];
  if (t.rank() > 8) return PEPS_EUNSUPPORTED;
  *rank = t.rank();
  for (int k = 0; k < t.rank(); ++k) dims[k] = t.dims[k];
  if (is_complex) *is_complex = t.cplx ? 1 : 0;
  return PEPS_OK;
}

int32_t peps_list_get(const peps_tensor_list* l, int32_t i, double* data) {
  if (!l || i < 0 || i >= int32_t(l->t.size()) || !data) return PEPS_EARG;
  return guard(l->ref.get(), [&] { l->ref->download(l->t[i], data); });
}

int32_t peps_list_destroy(peps_tensor_list* l) {
  delete l;
  return PEPS_OK;
}

// ── kernel-level ─────────────────────────────────────────────────────────────

int32_t peps_dl_traced(peps_ctx* ctx, int32_t is_complex, const int64_t dims[5], const double* T,
                       double* E) {
  if (!ctx || !dims || !T || !E) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    Tensor t = ctx->c->upload(std::vector<int64_t>(dims, dims + 5), is_complex != 0, T);
    ctx->c->download(peps::dl_traced(*ctx->c, t), E);
  });
}

int32_t peps_dl_open(peps_ctx* ctx, int32_t is_complex, const int64_t dims[5], const double* T, double* E) {
  if (!ctx || !dims || !T || !E) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    for (int i = 0; i < 5; ++i) PEPS_CHECK(dims[i] >= 1, peps::PS_ESHAPE, "dl_open: dims must be >= 1");
    Tensor t = ctx->c->upload(std::vector<int64_t>(dims, dims + 5), is_complex != 0, T);
    ctx->c->download(peps::dl_open(*ctx->c, t), E);
  });
}

static peps::MPS upload_list(peps::Ctx& c, bool cplx, int n, int rank, const int64_t* dims,
                             const double* const* data) {
  peps::MPS m;
  for (int i = 0; i < n; ++i) {
    PEPS_CHECK(data[i] != nullptr, peps::PS_EARG, "null tensor pointer");
    m.push_back(c.upload(std::vector<int64_t>(dims + rank * i, dims + rank * (i + 1)), cplx, data[i]));
  }
  return m;
}

int32_t peps_absorb_row(peps_ctx* ctx, int32_t from_top, int32_t is_complex, int32_t n,
                        const int64_t* mps_dims, const double* const* mps,
                        const int64_t* row_dims, const double* const* row,
                        peps_tensor_list** out) {
  if (!ctx || !mps_dims || !mps || !row_dims || !row || !out || n < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto M = upload_list(*ctx->c, is_complex != 0, n, 3, mps_dims, mps);
    auto R = upload_list(*ctx->c, is_complex != 0, n, 4, row_dims, row);
    auto o = from_top ? peps::absorb_row_from_top(*ctx->c, M, R) : peps::absorb_row_from_bottom(*ctx->c, M, R);
    *out = make_list(ctx->c, o);
  });
}

int32_t peps_compress_mps(peps_ctx* ctx, int32_t is_complex, int32_t n, const int64_t* dims,
                          const double* const* mps, int32_t chi, double cutoff,
                          peps_tensor_list** out) {
  if (!ctx || !dims || !mps || !out || n < 1 || chi < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto M = upload_list(*ctx->c, is_complex != 0, n, 3, dims, mps);
    if (ctx->c->compress_mode == 2) peps::compress_mps_gram(*ctx->c, M, chi, cutoff);
    else peps::compress_mps(*ctx->c, M, chi, cutoff);
    *out = make_list(ctx->c, M);
  });
}

int32_t peps_dmrg(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                  const int64_t* psi_dims, const double* const* psi0, int32_t D, int32_t nsweeps, double tol,
                  int32_t maxiter, int32_t krylovdim, double etol, double* energy, int32_t* sweeps,
                  peps_tensor_list** psi_out) {
  if (!ctx || !mpo_dims || !mpo || !psi_dims || !psi0 || !energy || !psi_out || n < 2 || D < 1 || nsweeps < 1 ||
      maxiter < 1 || krylovdim < 1 || !(tol > 0.0) || !(etol >= 0.0))
    return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto W = upload_list(*ctx->c, false, n, 4, mpo_dims, mpo);
    auto P = upload_list(*ctx->c, false, n, 3, psi_dims, psi0);
    peps::DmrgResult r = peps::dmrg_ground_state(*ctx->c, W, P, D, nsweeps, tol, maxiter, krylovdim, etol);
    *energy = r.E;
    if (sweeps) *sweeps = r.sweeps;
    *psi_out = make_list(ctx->c, r.psi);
  });
}

int32_t peps_mpo_expect(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                        const int64_t* psi_dims, const double* const* psi, double* out) {
  if (!ctx || !mpo_dims || !mpo || !psi_dims || !psi || !out || n < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto W = upload_list(*ctx->c, false, n, 4, mpo_dims, mpo);
    auto P = upload_list(*ctx->c, false, n, 3, psi_dims, psi);
    *out = peps::mpo_expect(*ctx->c, W, P);
  });
}

int32_t peps_mpo_expect_c(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                          int32_t psi_complex, const int64_t* psi_dims, const double* const* psi, double* out) {
  if (!ctx || !mpo_dims || !mpo || !psi_dims || !psi || !out || n < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto W = upload_list(*ctx->c, false, n, 4, mpo_dims, mpo);
    auto P = upload_list(*ctx->c, psi_complex != 0, n, 3, psi_dims, psi);
    *out = peps::mpo_expect(*ctx->c, W, P);
  });
}

int32_t peps_krylov_evolve(peps_ctx* ctx, int32_t n, const int64_t* mpo_dims, const double* const* mpo,
                           int32_t psi_complex, const int64_t* psi_dims, const double* const* psi0, double dt,
                           int32_t nsteps, int32_t krylov_m, int32_t Dmax, double eps, peps_tensor_list** psi_out) {
  if (!ctx || !mpo_dims || !mpo || !psi_dims || !psi0 || !psi_out || n < 2 || nsteps < 0 || krylov_m < 1 ||
      Dmax < 1 || !(eps >= 0.0) || !std::isfinite(dt))
    return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto W = upload_list(*ctx->c, false, n, 4, mpo_dims, mpo);
    auto P = upload_list(*ctx->c, psi_complex != 0, n, 3, psi_dims, psi0);
    *psi_out = make_list(ctx->c, peps::krylov_evolve(*ctx->c, W, P, dt, nsteps, krylov_m, Dmax, eps));
  });
}

int32_t peps_mps_local_expect(peps_ctx* ctx, int32_t n, int32_t psi_complex, const int64_t* psi_dims,
                              const double* const* psi, int32_t nops, const int32_t* sites, const int64_t* op_dims,
                              const double* const* ops, double* out) {
  if (!ctx || !psi_dims || !psi || n < 1 || nops < 0 || (nops > 0 && (!sites || !op_dims || !ops || !out)))
    return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    auto P = upload_list(*ctx->c, psi_complex != 0, n, 3, psi_dims, psi);
    auto O = upload_list(*ctx->c, false, nops, 2, op_dims, ops);
    std::vector<int> st(sites, sites + nops);
    std::vector<double> v = peps::mps_local_expect(*ctx->c, P, st, O);
    std::copy(v.begin(), v.end(), out);
  });
}

int32_t peps_boundary(peps_state* st, int32_t from_top, int32_t stop_row, int32_t chi,
                      peps_tensor_list** out) {
  if (!st || !out || chi < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    auto m = from_top ? peps::boundary_top(st->s, stop_row, chi) : peps::boundary_bottom(st->s, stop_row, chi);
    *out = make_list(st->ref, m);
  });
}

int32_t peps_norm2(peps_state* st, int32_t chi, double out[2]) {
  if (!st || !out || chi < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    peps::cplx v = peps::peps_norm2(st->s, chi);
    out[0] = v.re;
    out[1] = v.im;
  });
}

int32_t peps_bond_env(peps_state* st, int32_t orient, int32_t ix, int32_t iy, int32_t chi,
                      peps_tensor_list** out) {
  if (!st || !out || chi < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    bond_lam(st->s, orient, ix, iy);
    peps::Env E = orient == PEPS_H ? peps::bond_environment_h(st->s, ix, iy, chi)
                                   : peps::bond_environment_v(st->s, ix, iy, chi);
    *out = make_list(st->ref, {E.env, E.RA, E.RB, E.QA, E.QB});
  });
}

int32_t peps_charge_svd(peps_ctx* ctx, int32_t is_complex, int64_t nrow, int64_t ncol,
                        const double* M, const int64_t* row_charge, const int64_t* col_charge,
                        int32_t D_max, double reg, int64_t* Dk, int64_t* charges,
                        peps_tensor_list** out) {
  if (!ctx || !M || !row_charge || !col_charge || !Dk || !charges || !out || D_max < 1)
    return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    Tensor t = ctx->c->upload({nrow, ncol}, is_complex != 0, M);
    std::vector<int64_t> rq(row_charge, row_charge + nrow), cq(col_charge, col_charge + ncol);
    peps::CSVD r = peps::charge_svd(*ctx->c, t, rq, cq, D_max, reg);
    *Dk = int64_t(r.S.size());
    std::memcpy(charges, r.q.data(), r.q.size() * sizeof(int64_t));
    Tensor S = ctx->c->upload({int64_t(r.S.size())}, false, r.S.data());
    *out = make_list(ctx->c, {r.U, r.Vt, S});
  });
}

int32_t peps_als_truncate(peps_ctx* ctx, int32_t is_complex, const int64_t th_dims[4],
                          const double* theta, const double* env, int32_t D_max,
                          const int64_t* eR_of_pL, int32_t n_als, double reg, int64_t* Dp,
                          int64_t* bond_flux, double info[3], peps_tensor_list** out) {
  if (!ctx || !th_dims || !theta || !env || !eR_of_pL || !Dp || !bond_flux || !out || D_max < 1)
    return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    bool cp = is_complex != 0;
    int64_t nA = th_dims[0], nB = th_dims[3];
    Tensor th = ctx->c->upload(std::vector<int64_t>(th_dims, th_dims + 4), cp, theta);
    Tensor en = ctx->c->upload({nA, nB, nA, nB}, cp, env);
    std::vector<int64_t> eR(eR_of_pL, eR_of_pL + th_dims[1]);
    Tensor X, Y;
    std::vector<int64_t> bf;
    peps::ALSInfo inf;
    peps::als_truncate(*ctx->c, th, en, D_max, eR, n_als, reg, 0.0, &X, &Y, &bf, &inf);
    *Dp = int64_t(bf.size());
    std::memcpy(bond_flux, bf.data(), bf.size() * sizeof(int64_t));
    if (info) {
      info[0] = inf.cost0;
      info[1] = inf.cost1;
      info[2] = inf.accepted ? 1.0 : 0.0;
    }
    *out = make_list(ctx->c, {X, Y});
  });
}

int32_t peps_als_truncate_init(peps_ctx* ctx, int32_t is_complex, const int64_t th_dims[4],
                               const double* theta, const double* env, int32_t D_max,
                               const int64_t* eR_of_pL, int32_t n_als, double reg, double als_tol,
                               const double* X0, const double* Y0, int64_t* Dp, int64_t* bond_flux,
                               double info[4], peps_tensor_list** out) {
  if (!ctx || !th_dims || !theta || !env || !eR_of_pL || !Dp || !bond_flux || !out || D_max < 1 || n_als < 0 ||
      als_tol < 0.0 || (!X0) != (!Y0))
    return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    bool cp = is_complex != 0;
    int64_t nA = th_dims[0], nB = th_dims[3];
    Tensor th = ctx->c->upload(std::vector<int64_t>(th_dims, th_dims + 4), cp, theta);
    Tensor en = ctx->c->upload({nA, nB, nA, nB}, cp, env);
    std::vector<int64_t> eR(eR_of_pL, eR_of_pL + th_dims[1]);
    Tensor X, Y, Xi, Yi;
    std::vector<int64_t> bf;
    peps::ALSInfo inf;
    if (X0) {
      // the warm start has the shapes of the SVD initialisation: find them first
      peps::als_truncate(*ctx->c, th, en, D_max, eR, 0, reg, 0.0, &X, &Y, &bf, &inf);
      Xi = ctx->c->upload(X.dims, cp, X0);
      Yi = ctx->c->upload(Y.dims, cp, Y0);
    }
    peps::als_truncate(*ctx->c, th, en, D_max, eR, n_als, reg, als_tol, &X, &Y, &bf, &inf, X0 ? &Xi : nullptr,
                       X0 ? &Yi : nullptr);
    *Dp = int64_t(bf.size());
    std::memcpy(bond_flux, bf.data(), bf.size() * sizeof(int64_t));
    if (info) {
      info[0] = inf.cost0;
      info[1] = inf.cost1;
      info[2] = inf.accepted ? 1.0 : 0.0;
      info[3] = double(inf.sweeps);
    }
    *out = make_list(ctx->c, {X, Y});
  });
}

int32_t peps_su_bond(peps_state* st, int32_t orient, int32_t ix, int32_t iy, const double* gate,
                     int32_t D_max) {
  if (!st || D_max < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    bond_lam(st->s, orient, ix, iy);
    require_labels(st->s);
    Tensor G = gate_from_host(st->s, orient, ix, iy, gate);
    if (orient == PEPS_H) peps::su_bond_h(st->s, ix, iy, G, D_max);
    else peps::su_bond_v(st->s, ix, iy, G, D_max);
  });
}

int32_t peps_fu_bond_v4(peps_state* st, int32_t orient, int32_t ix, int32_t iy,
                        const double* gate, int32_t D_max, int32_t chi, int32_t n_als) {
  if (!st || D_max < 1 || chi < 1 || n_als < 0) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    bond_lam(st->s, orient, ix, iy);
    require_labels(st->s);
    Tensor G = gate_from_host(st->s, orient, ix, iy, gate);
    if (orient == PEPS_H) {
      peps::apply_mask(st->s, ix, iy);
      peps::apply_mask(st->s, ix + 1, iy);
      peps::Env E = peps::bond_environment_h(st->s, ix, iy, chi);
      peps::fu_bond_h(st->s, ix, iy, G, D_max, n_als, 0.0, E);
    } else {
      peps::apply_mask(st->s, ix, iy);
      peps::apply_mask(st->s, ix, iy + 1);
      peps::Env E = peps::bond_environment_v(st->s, ix, iy, chi);
      peps::fu_bond_v(st->s, ix, iy, G, D_max, n_als, 0.0, E);
    }
  });
}

namespace {
void run_step(peps_state* st, int32_t kind, const peps::Params& q) {
    PEPS_CHECK(q.mu == 0.0 || kind == PEPS_STEP_DENSE || kind == PEPS_STEP_REALTIME, peps::PS_EARG,
               "mu is taken by PEPS_STEP_DENSE and PEPS_STEP_REALTIME only");
    if (kind == PEPS_STEP_V4) {
      require_labels(st->s);
      std::vector<peps::ALSInfo> infos;
      peps::trotter_step_v4(st->s, q, st->gc, &infos);
      st->als_bonds = int64_t(infos.size());
      st->als_accepted = 0;
      st->als_sweeps = 0;
      for (auto& i : infos) {
        st->als_accepted += i.accepted;
        st->als_sweeps += i.sweeps;
      }
    } else if (kind == PEPS_STEP_GAUGE) {
      require_labels(st->s);
      peps::trotter_step_gauge(st->s, q, st->gc);
    } else if (kind == PEPS_STEP_DENSE) {
      peps::trotter_step_dense(st->s, q, st->gc);
    } else if (kind == PEPS_STEP_FULL_V2) {
      peps::trotter_step_full_v2(st->s, q, st->gc);
    } else if (kind == PEPS_STEP_REALTIME) {
      PEPS_CHECK(st->s.cplx, peps::PS_EARG, "PEPS_STEP_REALTIME needs a complex state");
      peps::trotter_step_realtime(st->s, q, st->gc);
    } else {
      throw peps::Error(peps::PS_EARG, "unknown step kind");
    }
}

// The imaginary-time tau annealing of ite_ground_state_v4 / _v3 / _v2
// (finite_peps_fullupdate_v4.jl:447-449, finite_peps_gauge_invariant.jl:
// 556-559, finite_peps_full_update.jl:595-608): n/8 steps at 0.1, n/8 at
// 0.05, n/4 at 0.02, n/6 at 0.01, n/6 at 0.005 and the rest at 0.001
// (integer division).
std::vector<std::pair<double, int64_t>> ite_tau_stages(int64_t n) {
  const int64_t s1 = n / 8, s2 = n / 8, s3 = n / 4, s4 = n / 6, s5 = n / 6;
  return {{0.1, s1}, {0.05, s2}, {0.02, s3}, {0.01, s4}, {0.005, s5}, {0.001, n - (s1 + s2 + s3 + s4 + s5)}};
}
}  // namespace

int32_t peps_trotter_step(peps_state* st, int32_t kind, const peps_params* p) {
  if (!st || !p) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    peps::Params q = to_params(p);
    for (auto& t : st->s.T) PEPS_CHECK(!t.empty(), peps::PS_EARG, "state has unset site tensors");
    run_step(st, kind, q);
  });
}

int32_t peps_ite_ground_state(peps_state* st, int32_t kind, const peps_params* p, int64_t n_ite,
                              int64_t* steps_done) {
  if (!st || !p || n_ite < 0) return PEPS_EARG;
  if (steps_done) *steps_done = 0;
  return guard(st->ref.get(), [&] {
    PEPS_CHECK(kind == PEPS_STEP_V4 || kind == PEPS_STEP_GAUGE || kind == PEPS_STEP_FULL_V2, peps::PS_EARG,
               "ite_ground_state: kind must be PEPS_STEP_V4, PEPS_STEP_GAUGE or PEPS_STEP_FULL_V2");
    peps::Params q = to_params(p);
    for (auto& t : st->s.T) PEPS_CHECK(!t.empty(), peps::PS_EARG, "state has unset site tensors");
    int64_t step = 0, bonds = 0, accepted = 0, sweeps = 0;
    for (const auto& stage : ite_tau_stages(n_ite)) {
      q.tau = stage.first;
      for (int64_t i = 0; i < stage.second; ++i) {
        run_step(st, kind, q);
        ++step;
        if (steps_done) *steps_done = step;
        bonds += st->als_bonds;
        accepted += st->als_accepted;
        sweeps += st->als_sweeps;
        if (step % 25 == 0 || step == 1) {
          char buf[96];
          std::snprintf(buf, sizeof buf, "ite step %lld/%lld (tau=%.4f)", (long long)step, (long long)n_ite,
                        stage.first);
          st->ref->note(buf);
        }
      }
    }
    if (kind == PEPS_STEP_V4) {  // ALS statistics over the whole run
      st->als_bonds = bonds;
      st->als_accepted = accepted;
      st->als_sweeps = sweeps;
    }
  });
}

int32_t peps_batch_trotter_step(peps_state* const* states, int32_t n, int32_t kind, const peps_params* p,
                                int32_t* status) {
  if (!states || n < 1 || !p) return PEPS_EARG;
  std::set<const peps::Ctx*> ctxs;
  for (int32_t i = 0; i < n; ++i) {
    if (!states[i]) return PEPS_EARG;
    ctxs.insert(states[i]->ref.get());
  }
  if (int32_t(ctxs.size()) != n) return PEPS_EARG;  // one context (stream, handles, pool) per state
  std::vector<int32_t> st(size_t(n), PEPS_OK);
  std::vector<std::thread> workers;
  workers.reserve(size_t(n));
  for (int32_t i = 0; i < n; ++i)
    workers.emplace_back([&, i] {
      // the device is a per-thread setting
      if (hipSetDevice(states[i]->ref->device) != hipSuccess) {
        st[size_t(i)] = PEPS_EHIP;
        return;
      }
      st[size_t(i)] = peps_trotter_step(states[i], kind, p);
    });
  for (auto& w : workers) w.join();
  int32_t first = PEPS_OK;
  for (int32_t i = 0; i < n; ++i) {
    if (status) status[i] = st[size_t(i)];
    if (first == PEPS_OK) first = st[size_t(i)];
  }
  return first;
}

int32_t peps_last_als_stats(peps_state* st, int64_t out[3]) {
  if (!st || !out) return PEPS_EARG;
  out[0] = st->als_bonds;
  out[1] = st->als_accepted;
  out[2] = st->als_sweeps;
  return PEPS_OK;
}

int32_t peps_energy_v4(peps_state* st, const peps_params* p, double* E) {
  if (!st || !p || !E) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    peps::Params q = to_params(p);
    PEPS_CHECK(q.mu == 0.0, peps::PS_EARG, "compute_energy_v4 takes no mu");
    *E = peps::energy_v4(st->s, q);
  });
}

int32_t peps_apply_site_op(peps_state* st, int32_t ix, int32_t iy, const double* O) {
  if (!st || !O) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    check_site(st->s, ix, iy);
    Tensor& T = st->s.t(ix, iy);
    PEPS_CHECK(!T.empty(), peps::PS_EARG, "site tensor not set");
    const int64_t dp = T.dims[0];
    Tensor Od = st->s.c->upload({dp, dp}, st->s.cplx, O);
    peps::apply_site_op(st->s, ix, iy, Od);
  });
}

// ── dense simple update with plaquettes ─────────────────────────────────────

int32_t peps_su_bond_dense(peps_state* st, int32_t orient, int32_t ix, int32_t iy, const double* gate,
                           int32_t D_trunc) {
  if (!st || D_trunc < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    bond_lam(st->s, orient, ix, iy);
    Tensor G = gate_from_host(st->s, orient, ix, iy, gate);
    if (orient == PEPS_H) peps::dense_bond_h(st->s, ix, iy, G, D_trunc, peps::SU_DENSE);
    else peps::dense_bond_v(st->s, ix, iy, G, D_trunc, peps::SU_DENSE);
  });
}

int32_t peps_fu_bond_v2(peps_state* st, int32_t orient, int32_t ix, int32_t iy, const double* gate,
                        int32_t D_trunc, int32_t use_gauss) {
  if (!st || D_trunc < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    PEPS_CHECK(use_gauss == 0, peps::PS_EUNSUPPORTED,
               "full_update_bond_h!: use_gauss=true is not supported (the reference recovers dg "
               "from the site dimension incorrectly and projects the wrong index group)");
    bond_lam(st->s, orient, ix, iy);
    Tensor G = gate_from_host(st->s, orient, ix, iy, gate);
    if (orient == PEPS_H) peps::dense_bond_h(st->s, ix, iy, G, D_trunc, peps::SU_V2);
    else peps::dense_bond_v(st->s, ix, iy, G, D_trunc, peps::SU_V2);
  });
}

int32_t peps_plaquette_gate(peps_ctx* ctx, int32_t gd, double g, double tau, double cutoff,
                            int32_t* Da, int32_t* Db, double* GA, double* GB, double* GC) {
  if (!ctx || !Da || !Db || gd < 2 || gd > 8 || !(g != 0.0)) return PEPS_EARG;
  if (GA && (!GB || !GC)) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    peps::PlaquetteGate pg = peps::plaquette_gate(*ctx->c, gd, g, tau, cutoff, 64, false);
    *Da = pg.Da;
    *Db = pg.Db;
    if (GA) {
      ctx->c->download(pg.GA, GA);
      ctx->c->download(pg.GB, GB);
      ctx->c->download(pg.GC, GC);
    }
  });
}

int32_t peps_update_plaquette(peps_state* st, int32_t ix, int32_t iy, int32_t Da, int32_t Db,
                              const double* GA, const double* GB, const double* GC, int32_t D_trunc) {
  if (!st || !GA || !GB || !GC || Da < 1 || Db < 1 || D_trunc < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    peps::State& s = st->s;
    const int64_t gd = s.gd;
    for (auto& t : s.T) PEPS_CHECK(!t.empty(), peps::PS_EARG, "state has unset site tensors");
    peps::PlaquetteGate pg;
    pg.Da = Da;
    pg.Db = Db;
    pg.GA = s.c->upload({gd, gd, gd, gd, Da, Db}, s.cplx, GA);
    pg.GB = s.c->upload({gd, gd, Da}, s.cplx, GB);
    pg.GC = s.c->upload({gd, gd, Db}, s.cplx, GC);
    peps::update_plaquette(s, ix, iy, pg, D_trunc);
  });
}

int32_t peps_expect_site(peps_state* st, int32_t ix, int32_t iy, const double* O, double* out) {
  if (!st || !O || !out) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    peps::State& s = st->s;
    check_site(s, ix, iy);
    const int64_t dp = s.t(ix, iy).dims.at(0);
    std::vector<double> o(O, O + dp * dp * (s.cplx ? 2 : 1));
    *out = peps::expect_site(s, ix, iy, o, s.cplx);
  });
}

int32_t peps_plaquette_energy(peps_state* st, int32_t ix, int32_t iy, double g, double* out) {
  if (!st || !out || !(g != 0.0)) return PEPS_EARG;
  return guard(st->ref.get(), [&] { *out = peps::plaquette_energy(st->s, ix, iy, g); });
}

int32_t peps_total_energy(peps_state* st, const peps_params* p, double* out) {
  if (!st || !p || !out) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    for (auto& t : st->s.T) PEPS_CHECK(!t.empty(), peps::PS_EARG, "state has unset site tensors");
    PEPS_CHECK(p->g != 0.0, peps::PS_EARG, "total_energy: g must be non-zero");
    *out = peps::total_energy(st->s, p->g, p->t_hop, p->m);
  });
}

int32_t peps_plaquette_expect(peps_state* st, int32_t ix, int32_t iy, int32_t chi, double out[2]) {
  if (!st || !out || chi < 1) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    for (auto& t : st->s.T) PEPS_CHECK(!t.empty(), peps::PS_EARG, "state has unset site tensors");
    peps::cplx w = peps::plaquette_expect(st->s, ix, iy, chi);
    out[0] = w.re;
    out[1] = w.im;
  });
}

int32_t peps_bond_entropy(const double* lambda, int64_t n, double* out) {
  if ((!lambda && n > 0) || n < 0 || !out) return PEPS_EARG;
  *out = peps::bond_entropy(std::vector<double>(lambda, lambda + n));
  return PEPS_OK;
}

int32_t peps_measure_all(peps_state* st, double out[7], double* nf_grid, double* E2_grid) {
  if (!st || !out) return PEPS_EARG;
  return guard(st->ref.get(), [&] {
    for (auto& t : st->s.T) PEPS_CHECK(!t.empty(), peps::PS_EARG, "state has unset site tensors");
    peps::Measurements m = peps::measure_all_finite(st->s);
    const double v[7] = {m.nf_mean, m.nf_even, m.nf_odd, m.E2_mean, m.S_h_mean, m.S_v_mean, m.S_mean};
    std::copy(v, v + 7, out);
    if (nf_grid) std::copy(m.nf_grid.begin(), m.nf_grid.end(), nf_grid);
    if (E2_grid) std::copy(m.E2_grid.begin(), m.E2_grid.end(), E2_grid);
  });
}

// ── learned ALS warm start: the small MLP ───────────────────────────────────

int32_t peps_mlp_create(peps_ctx* ctx, int32_t n_in, int32_t width, int32_t n_out, uint64_t seed,
                        peps_mlp** out) {
  if (!ctx || !out) return PEPS_EARG;
  *out = nullptr;
  return guard(ctx->c.get(), [&] { *out = new peps_mlp(ctx->c, n_in, width, n_out, seed); });
}

int32_t peps_mlp_destroy(peps_mlp* m) {
  delete m;
  return PEPS_OK;
}

int32_t peps_mlp_param_count(const peps_mlp* m, int64_t* n) {
  if (!m || !n) return PEPS_EARG;
  *n = m->m.param_count();
  return PEPS_OK;
}

int32_t peps_mlp_get_params(peps_mlp* m, double* out) {
  if (!m || !out) return PEPS_EARG;
  return guard(m->ref.get(), [&] { m->m.get_params(out); });
}

int32_t peps_mlp_set_params(peps_mlp* m, const double* in) {
  if (!m || !in) return PEPS_EARG;
  return guard(m->ref.get(), [&] { m->m.set_params(in); });
}

int32_t peps_mlp_forward(peps_mlp* m, int64_t batch, const double* X, double* Y) {
  if (!m || !X || !Y || batch < 1) return PEPS_EARG;
  return guard(m->ref.get(), [&] {
    Tensor x = m->ref->upload({m->m.nin, batch}, false, X);
    m->ref->download(m->m.forward(x), Y);
  });
}

int32_t peps_mlp_gradients(peps_mlp* m, int64_t batch, const double* X, const double* T, double* loss,
                           double* grad) {
  if (!m || !X || !T || !loss || !grad || batch < 1) return PEPS_EARG;
  return guard(m->ref.get(), [&] {
    Tensor x = m->ref->upload({m->m.nin, batch}, false, X);
    Tensor t = m->ref->upload({m->m.nout, batch}, false, T);
    Tensor gW[3], gb[3];
    *loss = m->m.gradients(x, t, gW, gb);
    int64_t off = 0;
    for (int l = 0; l < 3; ++l) {
      m->ref->download(gW[l], grad + off);
      off += gW[l].numel();
      m->ref->download(gb[l], grad + off);
      off += gb[l].numel();
    }
  });
}

int32_t peps_mlp_train_step(peps_mlp* m, int64_t batch, const double* X, const double* T, double lr,
                            double* loss) {
  if (!m || !X || !T || !loss || batch < 1 || !(lr > 0.0)) return PEPS_EARG;
  return guard(m->ref.get(), [&] {
    Tensor x = m->ref->upload({m->m.nin, batch}, false, X);
    Tensor t = m->ref->upload({m->m.nout, batch}, false, T);
    *loss = m->m.train_step(x, t, lr);
  });
}

int32_t peps_gemm(peps_ctx* ctx, int32_t is_complex, int64_t M, int64_t N, int64_t K,
                  const double* A, int32_t a_kfast, int32_t conjA, const double* B,
                  int32_t b_kfast, int32_t conjB, double* Cout) {
  if (!ctx || !A || !B || !Cout || M < 0 || N < 0 || K < 0) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    bool cp = is_complex != 0;
    Tensor a = ctx->c->upload({std::max<int64_t>(1, M * K)}, cp, A);
    Tensor b = ctx->c->upload({std::max<int64_t>(1, K * N)}, cp, B);
    Tensor c = ctx->c->alloc({std::max<int64_t>(1, M * N)}, cp);
    peps::gemm(*ctx->c, cp, M, N, K, a.data(), a_kfast != 0, conjA != 0, b.data(), b_kfast != 0,
               conjB != 0, c.data(), 0.0);
    ctx->c->download(c.reshaped({M * N}), Cout);
  });
}

// [ptr, ptr + bytes) inside one device allocation of the context's device
static void check_device_span(const peps::Ctx& c, const void* ptr, int64_t bytes, const char* what) {
  hipPointerAttribute_t at;
  PEPS_CHECK(hipPointerGetAttributes(&at, ptr) == hipSuccess && at.type == hipMemoryTypeDevice &&
                 at.device == c.device,
             peps::PS_EARG, std::string("gemm_noalloc: ") + what + " is not device memory of the context's device");
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  PEPS_CHECK(hipMemGetAddressRange(&base, &size, const_cast<void*>(ptr)) == hipSuccess, peps::PS_EARG,
             std::string("gemm_noalloc: no allocation holds ") + what);
  const char* b = static_cast<const char*>(base);
  const char* p = static_cast<const char*>(ptr);
  PEPS_CHECK(p >= b && uint64_t(bytes) <= size - size_t(p - b), peps::PS_EARG,
             std::string("gemm_noalloc: ") + what + " extends past the end of its allocation");
}

int32_t peps_gemm_noalloc(peps_ctx* ctx, int32_t is_complex, int64_t M, int64_t N, int64_t K,
                          const double* A, int32_t a_kfast, int32_t conjA, const double* B,
                          int32_t b_kfast, int32_t conjB, double* Cout) {
  if (!ctx || !A || !B || !Cout || M < 0 || N < 0 || K < 0) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    PEPS_CHECK(M < (int64_t(1) << 31) && N < (int64_t(1) << 31) && K < (int64_t(1) << 31), peps::PS_EUNSUPPORTED,
               "gemm_noalloc: extents must stay below 2^31");
    const int64_t esz = is_complex ? 16 : 8;
    if (M * N == 0) return;
    check_device_span(*ctx->c, Cout, M * N * esz, "C");
    if (K > 0) {
      check_device_span(*ctx->c, A, M * K * esz, "A");
      check_device_span(*ctx->c, B, K * N * esz, "B");
    }
    peps::gemm_noalloc(*ctx->c, is_complex != 0, M, N, K, A, a_kfast != 0, conjA != 0, B, b_kfast != 0,
                       conjB != 0, Cout);
  });
}

int32_t peps_permute(peps_ctx* ctx, int32_t is_complex, int32_t rank, const int64_t* dims,
                     const int32_t* perm, const double* in, double* out) {
  if (!ctx || !dims || !perm || !in || !out || rank < 1 || rank > 16) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    std::vector<int64_t> d(dims, dims + rank);
    std::vector<int> pr(perm, perm + rank);
    std::vector<int> seen(rank, 0);
    for (int i = 0; i < rank; ++i) {
      PEPS_CHECK(d[i] >= 1 && pr[i] >= 0 && pr[i] < rank && !seen[pr[i]], peps::PS_EARG, "permute: bad arguments");
      seen[pr[i]] = 1;
    }
    Tensor t = ctx->c->upload(d, is_complex != 0, in);
    Tensor o = peps::permute(*ctx->c, t, pr);
    ctx->c->download(o, out);
  });
}

int32_t peps_gemm_acc(peps_ctx* ctx, int32_t is_complex, int64_t M, int64_t N, int64_t K,
                      const double* A, int32_t a_kfast, int32_t conjA, const double* B,
                      int32_t b_kfast, int32_t conjB, double beta, double* Cinout) {
  if (!ctx || !A || !B || !Cinout || M < 0 || N < 0 || K < 0) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    bool cp = is_complex != 0;
    Tensor a = ctx->c->upload({std::max<int64_t>(1, M * K)}, cp, A);
    Tensor b = ctx->c->upload({std::max<int64_t>(1, K * N)}, cp, B);
    Tensor c = ctx->c->upload({std::max<int64_t>(1, M * N)}, cp, Cinout);
    peps::gemm(*ctx->c, cp, M, N, K, a.data(), a_kfast != 0, conjA != 0, b.data(), b_kfast != 0,
               conjB != 0, c.data(), beta);
    ctx->c->download(c.reshaped({M * N}), Cinout);
  });
}

int32_t peps_eigh(peps_ctx* ctx, int32_t is_complex, int32_t n, int32_t nb, const double* A,
                  double* V, double* w) {
  if (!ctx || !A || !V || !w || n < 1 || nb < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    bool cp = is_complex != 0;
    const int64_t nn = int64_t(n) * n, esz = cp ? 2 : 1;
    for (int32_t b = 0; b < nb; ++b) {
      Tensor a = ctx->c->upload({n, n}, cp, A + b * nn * esz);
      peps::EIG f = peps::eigh(*ctx->c, a);
      ctx->c->download(f.V, V + b * nn * esz);
      std::copy(f.w.begin(), f.w.end(), w + int64_t(b) * n);
    }
  });
}

int32_t peps_solve_herm(peps_ctx* ctx, int32_t is_complex, int64_t n, int64_t nrhs, const double* N,
                        const double* S, double reg, double* X) {
  if (!ctx || !N || !S || !X || n < 1 || nrhs < 1 || !(reg >= 0.0)) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    const bool cp = is_complex != 0;
    Tensor a = ctx->c->upload({n, n}, cp, N);
    Tensor b = ctx->c->upload({n, nrhs}, cp, S);
    ctx->c->download(peps::solve_herm(*ctx->c, a, b, reg), X);
  });
}

int32_t peps_svd(peps_ctx* ctx, int32_t is_complex, int64_t m, int64_t n, const double* A,
                 double* U, double* S, double* Vt) {
  if (!ctx || !A || !U || !S || !Vt || m < 1 || n < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    Tensor a = ctx->c->upload({m, n}, is_complex != 0, A);
    peps::SVD f = peps::svd_thin(*ctx->c, a);
    ctx->c->download(f.U, U);
    ctx->c->download(f.Vt, Vt);
    std::memcpy(S, f.S.data(), f.S.size() * sizeof(double));
  });
}

int32_t peps_qr(peps_ctx* ctx, int32_t is_complex, int64_t m, int64_t n, const double* A, double* Q, double* R) {
  if (!ctx || !A || !Q || !R || m < 1 || n < 1) return PEPS_EARG;
  return guard(ctx->c.get(), [&] {
    Tensor a = ctx->c->upload({m, n}, is_complex != 0, A);
    peps::QR f = peps::qr_thin(*ctx->c, a);
    ctx->c->download(f.Q, Q);
    ctx->c->download(f.R, R);
  });
}

int32_t peps_bond_operator(int32_t orient, int32_t ix, int32_t iy, int32_t nx, int32_t ny,
                           int32_t gd, double g, double t_hop, double m, double tau, int32_t* n,
                           double* out) {
  if (!n) return PEPS_EARG;
  return guard(nullptr, [&] {
    PEPS_CHECK(nx >= 2 && ny >= 2 && gd >= 2, peps::PS_EARG, "bad lattice");
    PEPS_CHECK(orient == PEPS_H ? (ix >= 1 && ix < nx && iy >= 1 && iy <= ny)
                                : (ix >= 1 && ix <= nx && iy >= 1 && iy < ny),
               peps::PS_EARG, "bond out of range");
    peps::Mat H = orient == PEPS_H ? peps::H_merged_h_site(ix, iy, nx, ny, gd, g, t_hop, m)
                                   : peps::H_merged_v_site(ix, iy, nx, ny, gd, g, t_hop, m);
    if (tau != 0.0) H = peps::expm_sym(H, tau);
    *n = H.n;
    if (out) std::memcpy(out, H.a.data(), H.a.size() * sizeof(double));
  });
}

}  // extern "C"

// ── exact diagonalisation (finite_ed.jl) ─────────────────────────────────────

extern "C" {

int32_t peps_ed_create(peps_ctx* ctx, int32_t nx, int32_t ny, int32_t dg, const int64_t* g_charges,
                       peps_ed** out) {
  if (!ctx || !g_charges || !out || nx < 1 || ny < 1) return PEPS_EARG;
  *out = nullptr;
  return guard(ctx->c.get(), [&] {
    auto e = std::make_unique<peps_ed>();
    e->ref = ctx->c;
    e->b = peps::ed_build_basis(*ctx->c, nx, ny, dg, std::vector<int64_t>(g_charges, g_charges + int64_t(nx) * ny));
    *out = e.release();
  });
}

int32_t peps_ed_destroy(peps_ed* ed) {
  delete ed;
  return PEPS_OK;
}

int32_t peps_ed_dim(const peps_ed* ed, int64_t* dim) {
  if (!ed || !dim) return PEPS_EARG;
  *dim = ed->b->dim;
  return PEPS_OK;
}

int32_t peps_ed_keys(peps_ed* ed, uint64_t* keys) {
  if (!ed || !keys) return PEPS_EARG;
  return guard(ed->ref.get(), [&] {
    peps::Ctx& c = *ed->ref;
    PEPS_HIP(hipMemcpyAsync(keys, ed->b->keys->ptr, size_t(ed->b->dim) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            c.stream));
    PEPS_HIP(hipStreamSynchronize(c.stream));
  });
}

int32_t peps_ed_hamiltonian(peps_ed* ed, double g, double t, double m) {
  if (!ed) return PEPS_EARG;
  return guard(ed->ref.get(), [&] { peps::ed_build_hamiltonian(*ed->b, g, t, m); });
}

int32_t peps_ed_apply(peps_ed* ed, const double* x, double* y) {
  if (!ed || !x || !y) return PEPS_EARG;
  return guard(ed->ref.get(), [&] {
    peps::Ctx& c = *ed->ref;
    Tensor xd = c.upload({ed->b->dim}, false, x);
    Tensor yd = c.alloc({ed->b->dim}, false);
    peps::ed_apply(*ed->b, xd, yd);
    c.download(yd, y);
  });
}

int32_t peps_ed_ground_state(peps_ed* ed, double tol, double* E0, double* psi) {
  if (!ed || !E0 || !(tol > 0.0)) return PEPS_EARG;
  return guard(ed->ref.get(), [&] {
    *E0 = peps::ed_ground_state(*ed->b, tol, 50, 300);
    if (psi) ed->ref->download(ed->b->psi0, psi);
  });
}

int32_t peps_ed_evolve(peps_ed* ed, const double* psi0, double dt, int32_t nsteps, int32_t krylovdim, double tol,
                       double* psi_out, double* obs, double* nf_grids, double* E2_grids, int32_t* projections) {
  if (!ed || !psi0 || !obs || nsteps < 0 || krylovdim < 2 || !(tol > 0.0) || !std::isfinite(dt)) return PEPS_EARG;
  return guard(ed->ref.get(), [&] {
    peps::Ctx& c = *ed->ref;
    peps::EDBasis& b = *ed->b;
    const size_t ns = size_t(b.nx) * b.ny;
    Tensor psi = c.upload({b.dim}, true, psi0);
    int total = 0;
    for (int32_t s = 0; s <= nsteps; ++s) {
      if (s > 0) {
        int sub = 0;
        psi = peps::ed_exp_step(b, psi, dt, krylovdim, tol, &sub);
        total += sub;
        const double nrm = std::sqrt(peps::norm2sq(c, psi));  // psi ./= norm(psi)
        PEPS_CHECK(nrm > 0.0 && std::isfinite(nrm), peps::PS_ESOLVER, "ed_evolve: zero or non-finite state");
        peps::scale(c, psi, 1.0 / nrm);
      }
      peps::EDObs o = peps::ed_measure(b, psi);
      obs[4 * size_t(s)] = o.nf_mean;
      obs[4 * size_t(s) + 1] = o.nf_even;
      obs[4 * size_t(s) + 2] = o.nf_odd;
      obs[4 * size_t(s) + 3] = o.E2_mean;
      if (nf_grids) std::copy(o.nf_grid.begin(), o.nf_grid.end(), nf_grids + ns * s);
      if (E2_grids) std::copy(o.E2_grid.begin(), o.E2_grid.end(), E2_grids + ns * s);
    }
    if (psi_out) c.download(psi, psi_out);
    if (projections) *projections = total;
  });
}

int32_t peps_ed_measure(peps_ed* ed, const double* psi, double out[4], double* nf_grid, double* E2_grid) {
  if (!ed || !out) return PEPS_EARG;
  return guard(ed->ref.get(), [&] {
    peps::Ctx& c = *ed->ref;
    PEPS_CHECK(psi || ed->b->has_gs, peps::PS_EARG, "ed_measure: no state given and no ground state computed");
    Tensor p = psi ? c.upload({ed->b->dim}, false, psi) : ed->b->psi0;
    peps::EDObs o = peps::ed_measure(*ed->b, p);
    out[0] = o.nf_mean;
    out[1] = o.nf_even;
    out[2] = o.nf_odd;
    out[3] = o.E2_mean;
    if (nf_grid) std::copy(o.nf_grid.begin(), o.nf_grid.end(), nf_grid);
    if (E2_grid) std::copy(o.E2_grid.begin(), o.E2_grid.end(), E2_grid);
  });
}

}  // extern "C"
