// Real symmetric eigensolver for n <= 1024 on one MI355X:
//   1. sytrd_lower_kernel: Householder reduction to tridiagonal form in ONE
//      workgroup (1024 threads) and one launch.  The trailing matrix is read
//      and written once per reflector: each pass applies the previous
//      step's rank-2 update and forms the next matrix-vector product in the
//      same sweep over memory, and the column the next reflector needs is
//      kept in LDS, so every step touches global memory exactly once.
//   2. the symmetric tridiagonal eigenproblem (divide and conquer, rocSOLVER
//      stedc on the n-vectors d, e);
//   3. apply_q_kernel: Z <- Q Z or Q^T Z with the n - 1 reflectors, each
//      column of Z held in the registers of 32 lanes (no LDS, no barriers),
//      so only the columns a caller needs are transformed.
// The reflectors follow LAPACK dsytrd with uplo = 'L': H(k) = I - tau_k v v^T
// with v(0..k) = 0, v(k+1) = 1 and v(k+2..n-1) stored in A(k+2..n-1, k).
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "peps_internal.h"

namespace peps {

namespace {

constexpr int kTrdThreads = 1024;
constexpr int kTrdBatch = 8;     // loads in flight per thread in the trailing-block sweep
constexpr int kQChunkBytes = 65536;  // LDS for the reflectors staged by apply_q_kernel

// Sum over the workgroup; every thread gets the same value (fixed order).
__device__ __forceinline__ double trd_block_sum(double x, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ double trd_block_max(double x, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s = fmax(s, red[w]);
  return s;
}

// A (n x n, leading dimension lda): the lower triangle is read; on exit it
// holds the reflectors below the subdiagonal.  d (n), e (n-1), tau (n-1).
// The matrix is first scaled by the power of two 2^-ex that brings its
// largest entry into [1, 2) (exact); d and e are those of the scaled matrix
// (so that the tridiagonal solver sees O(1) numbers) and ex goes to *exo.
// Dynamic LDS: 4 n + 1024 + 64 doubles.
__global__ __launch_bounds__(kTrdThreads) void sytrd_lower_kernel(double* __restrict__ A, int n, int lda,
                                                                  double* __restrict__ dd, double* __restrict__ ee,
                                                                  double* __restrict__ tt, int* __restrict__ exo) {
  extern __shared__ double sm[];
  double* bufA = sm;
  double* bufB = sm + n;
  double* wp = sm + 2 * n;     // w of the pending rank-2 update
  double* col = sm + 3 * n;    // current column (rows >= k)
  double* ypart = sm + 4 * n;  // per-thread partial products
  double* red0 = ypart + kTrdThreads;
  double* red1 = red0 + 32;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t nn = int64_t(n) * n;
  // scale to [1, 2) by a power of two, and fill the upper triangle
  double mx = 0.0;
  for (int64_t q = tid; q < nn; q += nt) {
    const int i = int(q % n), j = int(q / n);
    if (i >= j) mx = fmax(mx, fabs(A[i + int64_t(j) * lda]));
  }
  mx = trd_block_max(mx, red0);
  const int ex = mx > 0.0 ? ilogb(mx) : 0;
  const double s = ldexp(1.0, -ex);
  if (threadIdx.x == 0) *exo = ex;
  __syncthreads();  // every thread has read A's lower triangle for the maximum
  for (int64_t q = tid; q < nn; q += nt) {
    const int i = int(q % n), j = int(q / n);
    if (i >= j) {
      const double a = A[i + int64_t(j) * lda] * s;
      A[i + int64_t(j) * lda] = a;
      A[j + int64_t(i) * lda] = a;
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += nt) {
    col[i] = A[i];
    bufB[i] = 0.0;
    wp[i] = 0.0;
  }
  __syncthreads();
  double* vk = bufA;  // reflector of this step
  double* vp = bufB;  // reflector of the pending update
  for (int k = 0; k < n - 1; ++k) {
    const int m = n - k - 1;
    double part = 0.0;
    for (int i = k + 2 + tid; i < n; i += nt) part += col[i] * col[i];
    const double xn2 = trd_block_sum(part, red0);
    // reflector scalars (dlarfg), computed identically by every thread
    const double alpha = col[k + 1];
    double tau = 0.0, beta = alpha, scl = 0.0;
    if (xn2 > 0.0) {
      const double nrm = sqrt(alpha * alpha + xn2);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scl = 1.0 / (alpha - beta);
    }
    if (tid == 0) {
      dd[k] = col[k];
      ee[k] = beta;
      tt[k] = tau;
    }
    for (int i = tid; i < n; i += nt) {
      const double v = i <= k ? 0.0 : (i == k + 1 ? 1.0 : col[i] * scl);
      vk[i] = v;
      if (i >= k + 2) A[i + int64_t(k) * lda] = v;
    }
    __syncthreads();
    // one sweep over the trailing block (rows, columns k+1..n-1): apply the
    // pending update, store, accumulate y = A v_k, keep column k+1 in LDS
    const int Mp = (m + 63) & ~63;
    const int G = nt / Mp;
    const int r = tid % Mp, g = tid / Mp;
    double y = 0.0;
    if (g < G && r < m) {
      const int i = k + 1 + r;
      const double vpi = vp[i], wpi = wp[i];
      double* Ai = A + i;
      // kTrdBatch independent loads in flight per thread, then their updates
      for (int jj0 = g; jj0 < m; jj0 += kTrdBatch * G) {
        double a[kTrdBatch];
#pragma unroll
        for (int u = 0; u < kTrdBatch; ++u) {
          const int jj = jj0 + u * G;
          a[u] = jj < m ? Ai[int64_t(k + 1 + jj) * lda] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kTrdBatch; ++u) {
          const int jj = jj0 + u * G;
          if (jj < m) {
            const int j = k + 1 + jj;
            const double x = a[u] - (vpi * wp[j] + wpi * vp[j]);
            Ai[int64_t(j) * lda] = x;
            y += x * vk[j];
            if (jj == 0) col[i] = x;
          }
        }
      }
    }
    ypart[tid] = y;
    __syncthreads();
    double yi = 0.0, pdot = 0.0;
    if (tid < m) {
      for (int q = 0; q < G; ++q) yi += ypart[q * Mp + tid];
      yi *= tau;
      pdot = yi * vk[k + 1 + tid];
    }
    const double K = trd_block_sum(pdot, red1);
    const double half = -0.5 * tau * K;
    if (tid < m) wp[k + 1 + tid] = yi + half * vk[k + 1 + tid];
    __syncthreads();
    // column k+1 after this step's update (rows k+1..n-1)
    const double wk1 = wp[k + 1];
    for (int i = k + 1 + tid; i < n; i += nt) col[i] -= vk[i] * wk1 + wp[i];
    double* t = vp;
    vp = vk;
    vk = t;
    __syncthreads();
  }
  if (tid == 0) dd[n - 1] = col[n - 1];
}

// Z (n x ncols, leading dimension ldz) <- Q Z (trans = 0) or Q^T Z (trans =
// 1), Q = H(0) H(1) ... H(n-2).  Column c of Z lives in the registers of
// the 32 lanes (tid & 31) of half-wave c, NR rows per lane; the dot product
// with each reflector is a butterfly over those 32 lanes.  The reflectors
// are staged through LDS in chunks of ch (ch * n doubles, loaded by the
// whole workgroup with many loads in flight), so the per-reflector work
// reads only LDS.
template <int NR>
__global__ __launch_bounds__(1024) void apply_q_kernel(const double* __restrict__ R, int lda,
                                                       const double* __restrict__ tau, int n,
                                                       double* __restrict__ Z, int ldz, int ncols, int trans,
                                                       int ch) {
  extern __shared__ double vs[];  // ch reflectors of length n, then ch scalars
  double* ts = vs + int64_t(ch) * n;
  const int hl = threadIdx.x & 31;
  const int colid = blockIdx.x * (blockDim.x >> 5) + (threadIdx.x >> 5);
  const bool active = colid < ncols;
  double* zc = Z + int64_t(active ? colid : 0) * ldz;
  double z[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int i = hl + 32 * s;
    z[s] = (active && i < n) ? zc[i] : 0.0;
  }
  const int nref = n - 1;
  for (int t0 = 0; t0 < nref; t0 += ch) {
    const int cnt = min(ch, nref - t0);
    __syncthreads();  // the previous chunk is consumed
    for (int q = threadIdx.x; q < cnt * n; q += blockDim.x) {
      const int c = q / n, i = q - c * n;
      const int t = t0 + c;
      const int k = trans ? t : nref - 1 - t;
      vs[q] = i <= k ? 0.0 : (i == k + 1 ? 1.0 : R[i + int64_t(k) * lda]);
    }
    for (int c = threadIdx.x; c < cnt; c += blockDim.x) {
      const int t = t0 + c;
      ts[c] = tau[trans ? t : nref - 1 - t];
    }
    __syncthreads();
    for (int c = 0; c < cnt; ++c) {
      const double tk = ts[c];
      if (tk == 0.0) continue;
      const double* v = vs + int64_t(c) * n;
      double part = 0.0;
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int i = hl + 32 * s;
        part += (i < n ? v[i] : 0.0) * z[s];
      }
      for (int o = 16; o > 0; o >>= 1) part += __shfl_xor(part, o);
      const double f = tk * part;
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int i = hl + 32 * s;
        if (i < n) z[s] -= f * v[i];
      }
    }
  }
  if (active) {
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int i = hl + 32 * s;
      if (i < n) zc[i] = z[s];
    }
  }
}

void ck_solver(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    throw Error(PS_ESOLVER, std::string(what) + " failed with rocblas status " + std::to_string(int(st)));
}

}  // namespace

int64_t sym_tridiag_max_n() { return 1024; }

Tridiag tridiagonalize(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2 && A.dims[0] == A.dims[1] && !A.cplx, PS_ESHAPE, "tridiagonalize: real square matrix");
  const int64_t n = A.dims[0];
  PEPS_CHECK(n >= 1 && n <= sym_tridiag_max_n(), PS_EUNSUPPORTED, "tridiagonalize: 1 <= n <= 1024");
  Tridiag t;
  t.n = n;
  t.R = copy_tensor(c, A);
  t.d = c.alloc({n}, false);
  t.e = c.alloc({std::max<int64_t>(n - 1, 1)}, false);
  t.tau = c.alloc({std::max<int64_t>(n - 1, 1)}, false);
  t.ex = std::make_shared<Buffer>();
  t.ex->pool = c.pool.get();
  t.ex->bytes = 256;
  t.ex->ptr = c.pool->get(256);
  const std::string cls = Ctx::bucket("eigh_trd", n);
  hipEvent_t ev = nullptr;
  c.tbegin(cls, &ev);
  c.count(cls, 4.0 / 3.0 * double(n) * n * n, 16.0 * double(n) * n * n / 3.0);
  const size_t lds = size_t(4 * n + kTrdThreads + 64) * sizeof(double);
  hipLaunchKernelGGL(sytrd_lower_kernel, dim3(1), dim3(kTrdThreads), lds, c.stream, t.R.data(), int(n), int(n),
                     t.d.data(), t.e.data(), t.tau.data(), static_cast<int*>(t.ex->ptr));
  PEPS_HIP(hipGetLastError());
  c.tend(cls, ev);
  return t;
}

void apply_q(Ctx& c, const Tridiag& t, Tensor& Z, bool trans) {
  PEPS_CHECK(Z.rank() == 2 && Z.dims[0] == t.n && !Z.cplx, PS_ESHAPE, "apply_q: Z must be real n x k");
  const int64_t n = t.n, k = Z.dims[1];
  if (n <= 1 || k == 0) return;
  const std::string cls = Ctx::bucket("eigh_q", n);
  hipEvent_t ev = nullptr;
  c.tbegin(cls, &ev);
  c.count(cls, 4.0 * double(n) * n * k, 8.0 * (double(n) * n / 2.0 + 2.0 * double(n) * k));
  const int cols_per_block = int(std::min<int64_t>(32, (k + 1) & ~int64_t(1)));
  const dim3 grid(unsigned((k + cols_per_block - 1) / cols_per_block)), block(unsigned(32 * cols_per_block));
  const int tr = trans ? 1 : 0;
  // reflectors per LDS chunk: ch (n + 1) doubles within kQChunkBytes
  const int ch = int(std::max<int64_t>(1, std::min<int64_t>(n - 1, kQChunkBytes / (8 * (n + 1)))));
  const size_t lds = size_t(ch) * size_t(n + 1) * sizeof(double);
  if (n <= 256)
    hipLaunchKernelGGL(apply_q_kernel<8>, grid, block, lds, c.stream, t.R.data(), int(n), t.tau.data(), int(n),
                       Z.data(), int(n), int(k), tr, ch);
  else if (n <= 512)
    hipLaunchKernelGGL(apply_q_kernel<16>, grid, block, lds, c.stream, t.R.data(), int(n), t.tau.data(), int(n),
                       Z.data(), int(n), int(k), tr, ch);
  else
    hipLaunchKernelGGL(apply_q_kernel<32>, grid, block, lds, c.stream, t.R.data(), int(n), t.tau.data(), int(n),
                       Z.data(), int(n), int(k), tr, ch);
  PEPS_HIP(hipGetLastError());
  c.tend(cls, ev);
}

void tridiag_eig(Ctx& c, const Tridiag& t, std::vector<double>* w, Tensor* Z) {
  const int64_t n = t.n;
  Tensor d = copy_tensor(c, t.d), e = copy_tensor(c, t.e);
  *Z = c.alloc({n, n}, false);
  auto ib = std::make_shared<Buffer>();
  ib->pool = c.pool.get();
  ib->bytes = 256;
  ib->ptr = c.pool->get(256);
  rocblas_int* info = static_cast<rocblas_int*>(ib->ptr);
  const std::string cls = Ctx::bucket("eigh_stedc", n);
  hipEvent_t ev = nullptr;
  c.tbegin(cls, &ev);
  c.count(cls, 4.0 * double(n) * n * n, 0);
  ck_solver(rocsolver_dstedc(c.blas, rocblas_evect_tridiagonal, int(n), d.data(), e.data(), Z->data(), int(n), info),
            "dstedc");
  c.tend(cls, ev);
  int hinfo = 0, ex = 0;
  PEPS_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipMemcpyAsync(&ex, t.ex->ptr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  *w = c.to_host(d);  // synchronises the stream
  PEPS_CHECK(hinfo == 0, PS_ESOLVER, "stedc did not converge (info=" + std::to_string(hinfo) + ")");
  for (double& x : *w) x = std::ldexp(x, ex);  // undo the power-of-two scaling
}

}  // namespace peps
