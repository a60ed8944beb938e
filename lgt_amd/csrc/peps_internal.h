// Internal declarations of the MI355X finite-PEPS engine (libpeps_hip.so).
//
// Storage convention: every tensor is column-major with the first index
// fastest, exactly like the Julia reference arrays, and holds either real
// FP64 (8 B/entry) or complex FP64 interleaved (re,im; 16 B/entry, byte
// compatible with Julia ComplexF64 / C double _Complex).  The contraction
// precision (FP64 / FP32 / BF16 MFMA) is a property of the context, not of
// the storage: GEMM tiles are converted while being staged into LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <initializer_list>
#include <vector>

namespace peps {

// ── status codes (mirror include/peps.h) ─────────────────────────────────────
enum Status : int {
  PS_OK = 0,
  PS_EARG = 1,
  PS_ESHAPE = 2,
  PS_ENOSECTOR = 3,
  PS_EHIP = 4,
  PS_ENOMEM = 5,
  PS_ESOLVER = 6,
  PS_EUNSUPPORTED = 7,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PEPS_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::peps::Error(::peps::PS_EHIP, std::string(#call) + ": " +           \
                                            hipGetErrorString(e_));           \
  } while (0)

#define PEPS_CHECK(cond, code, msg)                                            \
  do {                                                                         \
    if (!(cond)) throw ::peps::Error((code), (msg));                           \
  } while (0)

struct cplx {
  double re, im;
};

// Contraction precision tiers.
enum Tier : int { TIER_FP64 = 0, TIER_FP32 = 1, TIER_BF16 = 2 };

// ── device memory: a caching allocator (no hipMalloc on the hot path) ────────
class Pool {
 public:
  explicit Pool(int device) : device_(device) {}
  ~Pool();
  void* get(size_t bytes);
  void put(void* p, size_t bytes);
  void trim();
  size_t bytes_in_use() const { return in_use_; }
  size_t bytes_cached() const { return cached_; }

 private:
  int device_;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> sizes_;
  size_t in_use_ = 0, cached_ = 0;
};

struct Buffer {
  Pool* pool = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
  ~Buffer() {
    if (pool && ptr) pool->put(ptr, bytes);
  }
};

class Ctx;

// A column-major device tensor.
struct Tensor {
  std::shared_ptr<Buffer> buf;
  std::vector<int64_t> dims;
  bool cplx = false;
  int64_t off = 0;  // element offset into buf (views/slices)

  int64_t numel() const {
    int64_t n = 1;
    for (auto d : dims) n *= d;
    return n;
  }
  int esz() const { return cplx ? 16 : 8; }
  size_t bytes() const { return size_t(numel()) * esz(); }
  int rank() const { return int(dims.size()); }
  int64_t dim(int i) const { return dims.at(i); }
  double* data() const {
    return buf ? static_cast<double*>(buf->ptr) + off * (cplx ? 2 : 1) : nullptr;
  }
  bool empty() const { return !buf; }
  // Reinterpret the same storage with other column-major dimensions.
  Tensor reshaped(std::vector<int64_t> nd) const;
  // View of the index range [i0, i1) of the last (slowest) dimension.
  Tensor slice_last(int64_t i0, int64_t i1) const;
};

struct KernelStat {
  double flops = 0;    // algorithmic flops (real-equivalent)
  double bytes = 0;    // algorithmic HBM bytes
  double ms = 0;       // summed launch duration (HIP events), when timing is on
  int64_t launches = 0;
};

class Ctx {
 public:
  explicit Ctx(int device);
  ~Ctx();

  int device;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::unique_ptr<Pool> pool;
  Tier tier = TIER_FP64;
  std::string last_error;
  int64_t warnings = 0;
  int ncu = 256;
  // boundary-MPS compression: 0 auto (Gram form when the exact bond > 1024),
  // 1 QR/SVD sweep on the exact product, 2 Gram form
  int compress_mode = 0;
  // progress lines on stderr, one per bond update (environment variable
  // PEPS_HIP_VERBOSE=1 at context creation)
  int verbose = 0;
  // Hermitian top-k eigenproblems up to this size are solved directly, larger
  // ones by subspace iteration (environment variable PEPS_HIP_EIG_DIRECT_MAX)
  int eig_direct_max = 512;
  // iteration caps of the factorisations: Chebyshev subspace iterations per
  // top-k eigenproblem and one-sided Jacobi sweeps per SVD.  Reaching a cap
  // counts a warning (peps_warning_count); an unconverged subspace iteration
  // is then answered by a direct eigendecomposition (N <= 4096).
  int subspace_max_it = 25;
  // subspace width of top_eigs for `want` pairs: 0 the default
  // max(want + 64, 2 want), else want + this many (environment variable
  // PEPS_HIP_SUBSPACE_OVERSAMPLE)
  int subspace_oversample = 0;
  // tridiagonal reduction above the register kernel's size (n > 288): 1
  // rocSOLVER sytrd (2-4x faster from n = 512 on), 0 the library's
  // one-workgroup global-memory kernel (environment variable PEPS_HIP_TRD=own)
  int trd_backend = 1;
  // leading-triplet (randomized, residual-checked) SVD of large charge
  // sectors in the ALS initialisation (environment variable
  // PEPS_HIP_LEADING_SVD=0 turns it off: every sector takes the full SVD)
  bool leading_svd = true;
  // ALS normal equations split into independent blocks where the couplings
  // between them are within the matrix's rounding level (environment
  // variable PEPS_HIP_ALS_BLOCKS=0 turns it off)
  bool als_blocks = true;
  // check every intermediate of the boundary compression for non-finite
  // values and name the first step that produced one (environment variable
  // PEPS_HIP_DEBUG_FINITE=1; one device scan per factor)
  bool debug_finite = false;
  int jacobi_max_sweeps = 80;
  // account GEMMs per exact shape ("gemm_f32 MxNxK ...") instead of per
  // precision (environment variable PEPS_HIP_GEMM_SHAPES=1)
  bool gemm_shapes = false;
  // singular-value cut of the boundary-MPS compression of the drivers
  // (compress_mps!'s 1e-14, finite_peps_boundary_mps.jl:93; relative to the
  // largest singular value under PEPS_STEP_RENORM); peps_set_compress_cut
  double bmps_cut = 1e-14;
  // real QR wider than one panel: 1 rocSOLVER geqrf/orgqr (default), 0 the
  // project's blocked Householder QR (PEPS_HIP_QR=blocked); the blocked form
  // is ~5x slower on the tall panels of the subspace iteration (5184 x 288:
  // 58 ms against 11 ms), whose one-workgroup panel kernel is the limit
  int qr_backend = 1;
  // last Ritz basis of each boundary-MPS column position: the warm start of
  // the next subspace iteration at the same position
  std::map<std::string, Tensor> ritz_cache;
  // one progress line on stderr when verbose
  void note(const std::string& what);

  // Per kernel-class accounting ("gemm", "permute", "svd", ...).
  std::map<std::string, KernelStat> stats;
  bool timing = false;
  bool timing_gemm_only = false;  // time only the MFMA GEMM classes of the tiers
  struct PendingEv {
    std::string cls;
    hipEvent_t a, b;
  };
  std::vector<PendingEv> pending;
  std::vector<hipEvent_t> ev_free;

  // Begin/end a timed region for kernel class `cls` (no-ops unless timing).
  void tbegin(const std::string& cls, hipEvent_t* a);
  void tend(const std::string& cls, hipEvent_t a);
  void resolve_timing();
  // class name with a size bucket, e.g. "qr[<=64]" (factorisation statistics)
  static std::string bucket(const char* cls, int64_t n) {
    const char* b = n <= 16 ? "<=16" : n <= 64 ? "<=64" : n <= 256 ? "<=256" : n <= 1024 ? "<=1024" : ">1024";
    return std::string(cls) + "[" + b + "]";
  }
  void count(const std::string& cls, double flops, double bytes) {
    auto& s = stats[cls];
    s.flops += flops;
    s.bytes += bytes;
    s.launches += 1;
  }

  Tensor alloc(std::vector<int64_t> dims, bool cplx);
  Tensor zeros(std::vector<int64_t> dims, bool cplx);
  Tensor upload(std::vector<int64_t> dims, bool cplx, const double* host);
  void download(const Tensor& t, double* host);
  std::vector<double> to_host(const Tensor& t);
  void sync();

  // Device scratch for small host→device uploads (label tables, weights).
  void* scratch_upload(const void* host, size_t bytes, std::shared_ptr<Buffer>* keep);
  // Host -> device copy enqueued on the stream without waiting for it: the
  // host bytes are first copied into a pinned staging ring (kStageBytes), so
  // the caller's array may go away at once.  Copies larger than a quarter of
  // the ring go synchronously.  When the ring wraps, the copies staged on its
  // previous lap are waited for (one event, recorded after every copy).
  void h2d_async(void* dst, const void* host, size_t bytes);
  static constexpr size_t kStageBytes = size_t(16) << 20;
  char* stage_host = nullptr;
  size_t stage_head = 0;
  hipEvent_t stage_ev = nullptr;
  bool stage_used = false;
};

// Runs the enclosed GEMMs at a fixed precision tier (the factorisations and
// their internal products stay FP64 whatever tier the contractions use).
struct TierGuard {
  Ctx& c;
  Tier old;
  TierGuard(Ctx& ctx, Tier t) : c(ctx), old(ctx.tier) { c.tier = t; }
  ~TierGuard() { c.tier = old; }
};

// ── kernels (kernels.hip) ────────────────────────────────────────────────────
// out = permute(in, perm): out.dims[i] = in.dims[perm[i]] (0-based perm).
Tensor permute(Ctx& c, const Tensor& in, const std::vector<int>& perm);
// C = op(A) * op(B), matrices column-major contiguous.
//   A is (M x K) if !a_kfast (A[m + k*M]) or stored (K x M) if a_kfast (A[k + m*K])
//   B is (K x N) if b_kfast (B[k + n*K]) or stored (N x K) if !b_kfast (B[n + k*N])
void gemm(Ctx& c, bool cplx, int64_t M, int64_t N, int64_t K, const double* A, bool a_kfast,
          bool conjA, const double* B, bool b_kfast, bool conjB, double* C, double beta);
// The same product with no allocation: split-K partial products in the
// library's fixed per-device array of kGemmScratchBytes (the split count is
// lowered until they fit), no FP32 operand cast, no timing events; C =
// op(A) op(B).  scratch_attach(device) must have run on the device (the
// context's constructor does it).
constexpr size_t kGemmScratchBytes = size_t(64) << 20;
void scratch_attach(int device);
void gemm_noalloc(Ctx& c, bool cplx, int64_t M, int64_t N, int64_t K, const double* A, bool a_kfast,
                  bool conjA, const double* B, bool b_kfast, bool conjB, double* C);
void scale_legs(Ctx& c, Tensor& T, const std::vector<double>* w /*[4]: l,r,u,d*/);
void scale(Ctx& c, Tensor& T, double s);
void axpy_inplace(Ctx& c, Tensor& Y, const Tensor& X, double a);  // Y += a X
void conj_inplace(Ctx& c, Tensor& T);
// T *= s for a complex factor (a real tensor accepts only s.im == 0).
void scale(Ctx& c, Tensor& T, cplx s);
// A complex copy of a real tensor (imaginary parts zero); complex T itself.
Tensor as_complex(Ctx& c, const Tensor& T);
// Zero the entries of a (p,l,r,u,d) tensor that violate the Gauss law.
void gauss_mask(Ctx& c, Tensor& T, const std::vector<int64_t>& nf, const std::vector<int64_t>& lR,
                const std::vector<int64_t>& lU, const std::vector<int64_t>& fl,
                const std::vector<int64_t>& fr, const std::vector<int64_t>& fu,
                const std::vector<int64_t>& fd, int64_t q);
// sum |entries of T outside the Gauss mask|^2
double gauss_violation(Ctx& c, const Tensor& T, const std::vector<int64_t>& nf,
                       const std::vector<int64_t>& lR, const std::vector<int64_t>& lU,
                       const std::vector<int64_t>& fl, const std::vector<int64_t>& fr,
                       const std::vector<int64_t>& fu, const std::vector<int64_t>& fd, int64_t q);
// sum_i a_i * conj?(b_i), deterministic two-pass reduction, result (re, im).
cplx dot(Ctx& c, const Tensor& a, const Tensor& b, bool conj_b);
double norm2sq(Ctx& c, const Tensor& a);
// M (nrow x ncol): per column, index of the first row with maximal |M|^2.
std::vector<int64_t> col_argmax(Ctx& c, const Tensor& M);
// M (nrow x ncol): zero entries with row_charge[r] != col_charge[c].
void zero_cross_sector(Ctx& c, Tensor& M, const std::vector<int64_t>& rq,
                       const std::vector<int64_t>& cq);
// out(i,j) = M(rows[i], cols[j]).
Tensor gather2d(Ctx& c, const Tensor& M, const std::vector<int64_t>& rows,
                const std::vector<int64_t>& cols);
// M(rows[i], col) = v(i) for column-vector sources; dst must be zeroed.
void scatter_cols(Ctx& c, Tensor& dst, const Tensor& src, const std::vector<int64_t>& rows,
                  const std::vector<int64_t>& src_cols, const std::vector<int64_t>& dst_cols);
void scatter_rows(Ctx& c, Tensor& dst, const Tensor& src, const std::vector<int64_t>& src_rows,
                  const std::vector<int64_t>& dst_rows, const std::vector<int64_t>& cols);
// Zero X(aA, pL, beta) where keep(pL, beta) == 0.
void mask_x(Ctx& c, Tensor& X, const std::vector<int8_t>& keep);
// out = upper trapezoid of A (rows 0..k-1) as a dense (k x n) matrix.
Tensor upper_trapezoid(Ctx& c, const Tensor& A, int64_t k);
// Column scaling M(:, j) *= s[j]; row scaling M(i, :) *= s[i].
void scale_cols(Ctx& c, Tensor& M, const std::vector<double>& s);
void scale_rows(Ctx& c, Tensor& M, const std::vector<double>& s);
// Hermitian part 0.5 (N + N^H) in place.
void hermitize(Ctx& c, Tensor& N);
// Copy a sub-block: dst(:, j0:j0+n) = src (column block).
Tensor copy_tensor(Ctx& c, const Tensor& t);
bool all_finite(Ctx& c, const Tensor& t);
// Squared 2-norms of the columns of M (synchronises).
std::vector<double> col_norm2(Ctx& c, const Tensor& M);
// Indices of the rows / columns of M that hold a non-zero entry (synchronises).
void nonzero_lines(Ctx& c, const Tensor& M, std::vector<int64_t>* rows, std::vector<int64_t>* cols);
// Householder QR of a panel with at most qr_small_max_cols() columns (one
// workgroup): W overwritten LAPACK-style (R above, reflectors below), tau
// their scalars; qr_form_q builds the thin Q (m x min(m,n)).
int qr_small_max_cols();
void qr_panel(Ctx& c, Tensor& W, Tensor& tau);
// The reflectors of a factored real panel as an explicit unit lower
// trapezoid V, and the triangular factor T of I - V T V^T from G = V^T V.
Tensor householder_v(Ctx& c, const Tensor& P);
Tensor householder_t(Ctx& c, const Tensor& G, const Tensor& tau);
// Contiguous sub-block copies of a matrix (no index tables): M(r0:r0+rows,
// c0:c0+cols) out, and B back into M at (r0, c0).
Tensor sub_block(Ctx& c, const Tensor& M, int64_t r0, int64_t c0, int64_t rows, int64_t cols);
void put_block(Ctx& c, Tensor& M, int64_t r0, int64_t c0, const Tensor& B);
void qr_form_q(Ctx& c, const Tensor& W, const Tensor& tau, Tensor& Q);
// One pass: are all entries finite, is any entry non-zero (synchronises).
void scan_values(Ctx& c, const Tensor& t, bool* finite, bool* nonzero);
// every entry of every tensor finite (one stream synchronisation)
bool all_finite(Ctx& c, std::initializer_list<const Tensor*> ts);
// The first k columns of the m x m identity.
Tensor eye(Ctx& c, int64_t m, int64_t k, bool cplx);
// Small-MLP pieces (real, batch in columns): Z += b per row (+ ReLU);
// dH *= (H > 0); db = row sums of dZ; D = (Y - T) scale; one Adam step.
void mlp_bias_act(Ctx& c, Tensor& Z, const Tensor& b, bool relu);
void mlp_relu_grad(Ctx& c, Tensor& dH, const Tensor& H);
void mlp_row_sum(Ctx& c, const Tensor& dZ, Tensor& db);
void mlp_scaled_diff(Ctx& c, const Tensor& Y, const Tensor& T, Tensor& D, double scale);
void mlp_adam(Ctx& c, Tensor& p, const Tensor& g, Tensor& m, Tensor& v, double lr, double b1, double b2, double eps,
              int64_t t);
// Deterministic uniform(-1,1) entries (complex: both parts).
void fill_random(Ctx& c, Tensor& T, uint64_t seed);

// ── contraction front end (contract.cpp) ────────────────────────────────────
// Einstein-summation contraction of two tensors with single-character labels,
// e.g. contract(c, A, "aub", B, "lrud", "ladrb").  Repeated labels on both
// operands are summed; every output label appears in exactly one operand.
Tensor contract(Ctx& c, const Tensor& A, const std::string& la, const Tensor& B,
                const std::string& lb, const std::string& lc, bool conjA = false,
                bool conjB = false);
// Plain matrix product of column-major matrices: C = op(A) op(B).
Tensor matmul(Ctx& c, const Tensor& A, const Tensor& B, bool conjA = false, bool conjB = false,
              bool transA = false, bool transB = false);

// ── dense linear algebra (linalg.cpp, rocSOLVER) ────────────────────────────
struct QR {
  Tensor Q, R;  // thin: Q (m x k), R (k x n), k = min(m, n)
};
QR qr_thin(Ctx& c, const Tensor& A);
// A^H as a new tensor (never a view of A).
Tensor conj_transpose(Ctx& c, const Tensor& A);
Tensor qr_r_only(Ctx& c, const Tensor& A);
struct SVD {
  Tensor U, Vt;            // U (m x k), Vt (k x n)
  std::vector<double> S;   // descending
};
SVD svd_thin(Ctx& c, const Tensor& A);
std::vector<SVD> svd_thin_batch(Ctx& c, const std::vector<Tensor>& As);
struct EIG {
  std::vector<double> w;   // ascending
  Tensor V;
};
EIG eigh(Ctx& c, const Tensor& A);

// Real symmetric tridiagonal reduction and its pieces (eigh.hip), n <= 1024:
// A = Q T Q^T with T = tridiag(d, e); R holds the reflectors (LAPACK dsytrd
// lower layout) and tau their scalars.
struct Tridiag {
  int64_t n = 0;
  Tensor R, tau, d, e;            // d, e: tridiagonal of 2^-ex A
  std::shared_ptr<Buffer> ex;     // device int: the scaling exponent ex
};
int64_t sym_tridiag_max_n();
// Reads the lower triangle of A (real, n x n); one kernel launch.
Tridiag tridiagonalize(Ctx& c, const Tensor& A);
// Z (n x k) <- Q Z, or Q^T Z when trans.
void apply_q(Ctx& c, const Tridiag& t, Tensor& Z, bool trans);
// Eigenvalues of T (ascending, host) and its eigenvectors Z (n x n, device).
void tridiag_eig(Ctx& c, const Tridiag& t, std::vector<double>* w, Tensor* Z);

// One-sided Jacobi SVD kernel (kernels.hip): A (m x n, m >= n) is overwritten
// by its orthogonalised columns u_j sigma_j, V gets the right vectors and sig
// the column norms (unsorted).
int jacobi_svd(Ctx& c, Tensor& A, Tensor& V, Tensor& sig, double tol, int max_sweeps);
// One one-sided Jacobi SVD per job (W m x n with m >= n, V n x n, sig n):
// the matrices the register kernels take run concurrently, one workgroup
// each; sweeps of each job are returned in its `sweeps`.
struct JacobiJob {
  Tensor* W;
  Tensor* V;
  Tensor* sig;
  double tol;
  int sweeps;
};
void jacobi_svd_batch(Ctx& c, std::vector<JacobiJob>& jobs, int max_sweeps);

}  // namespace peps
