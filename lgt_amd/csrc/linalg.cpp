// Dense FP64 factorisations on the GPU: thin QR (rocSOLVER Householder
// geqrf + orgqr/ungqr, the LAPACK semantics of Julia's `qr`), thin SVD
// (rocSOLVER gesvd) and the Hermitian eigendecomposition (the project's
// Jacobi kernel).
#include <rocsolver/rocsolver.h>

#include <algorithm>

#include "peps_internal.h"

namespace peps {

namespace {

void ck(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw Error(PS_ESOLVER, std::string(what) + " failed with rocblas status " + std::to_string(int(s)));
}

struct DevInt {
  std::shared_ptr<Buffer> b;
  rocblas_int* p;
  explicit DevInt(Ctx& c) {
    b = std::make_shared<Buffer>();
    b->pool = c.pool.get();
    b->bytes = 256;
    b->ptr = c.pool->get(256);
    p = static_cast<rocblas_int*>(b->ptr);
  }
  int get(Ctx& c) {
    int h = 0;
    PEPS_HIP(hipMemcpyAsync(&h, p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PEPS_HIP(hipStreamSynchronize(c.stream));
    return h;
  }
};

// rocSOLVER is never handed NaN/Inf: a non-finite input is reported as a
// solver failure instead.
void check_finite(Ctx& c, const Tensor& A, const char* what) {
  PEPS_CHECK(all_finite(c, A), PS_ESOLVER, std::string("non-finite input to ") + what);
}

void check_int_range(int64_t m, int64_t n) {
  PEPS_CHECK(m < (1ll << 31) && n < (1ll << 31) && m * n < (1ll << 62), PS_EUNSUPPORTED,
             "factorisation: matrix too large");
}

}  // namespace

QR qr_thin(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "qr: matrix expected");
  int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n);
  check_int_range(m, n);
  check_finite(c, A, "qr");
  Tensor W = copy_tensor(c, A);
  Tensor tau = c.alloc({std::max<int64_t>(k, 1)}, A.cplx);
  hipEvent_t ev = nullptr;
  c.tbegin("qr", &ev);
  c.count("qr", (A.cplx ? 4.0 : 1.0) * (4.0 * m * n * k - 2.0 * (m + n) * k * k + 4.0 * k * k * k / 3.0), 0);
  if (k > 0) {
    if (A.cplx)
      ck(rocsolver_zgeqrf(c.blas, int(m), int(n), (rocblas_double_complex*)W.data(), int(m),
                          (rocblas_double_complex*)tau.data()), "zgeqrf");
    else
      ck(rocsolver_dgeqrf(c.blas, int(m), int(n), W.data(), int(m), tau.data()), "dgeqrf");
  }
  QR r;
  r.R = upper_trapezoid(c, W, k);
  if (k > 0) {
    if (A.cplx)
      ck(rocsolver_zungqr(c.blas, int(m), int(k), int(k), (rocblas_double_complex*)W.data(), int(m),
                          (rocblas_double_complex*)tau.data()), "zungqr");
    else
      ck(rocsolver_dorgqr(c.blas, int(m), int(k), int(k), W.data(), int(m), tau.data()), "dorgqr");
  }
  // Q = first k columns of W (column-major: a prefix of the buffer)
  if (k == n) {
    r.Q = W.reshaped({m, n});
  } else {
    r.Q = c.alloc({m, k}, A.cplx);
    if (r.Q.bytes())
      PEPS_HIP(hipMemcpyAsync(r.Q.data(), W.data(), r.Q.bytes(), hipMemcpyDeviceToDevice, c.stream));
  }
  c.tend("qr", ev);
  return r;
}

Tensor qr_r_only(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "qr: matrix expected");
  int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n);
  check_int_range(m, n);
  check_finite(c, A, "qr");
  Tensor W = copy_tensor(c, A);
  Tensor tau = c.alloc({std::max<int64_t>(k, 1)}, A.cplx);
  hipEvent_t ev = nullptr;
  c.tbegin("qr", &ev);
  c.count("qr", (A.cplx ? 4.0 : 1.0) * (2.0 * m * n * k - (m + n) * k * k + 2.0 * k * k * k / 3.0), 0);
  if (k > 0) {
    if (A.cplx)
      ck(rocsolver_zgeqrf(c.blas, int(m), int(n), (rocblas_double_complex*)W.data(), int(m),
                          (rocblas_double_complex*)tau.data()), "zgeqrf");
    else
      ck(rocsolver_dgeqrf(c.blas, int(m), int(n), W.data(), int(m), tau.data()), "dgeqrf");
  }
  Tensor R = upper_trapezoid(c, W, k);
  c.tend("qr", ev);
  return R;
}

SVD svd_thin(Ctx& c, const Tensor& A) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "svd: matrix expected");
  int64_t m = A.dims[0], n = A.dims[1], k = std::min(m, n);
  check_int_range(m, n);
  SVD r;
  r.U = c.alloc({m, k}, A.cplx);
  r.Vt = c.alloc({k, n}, A.cplx);
  if (k == 0) return r;
  check_finite(c, A, "svd");
  Tensor W = copy_tensor(c, A);
  Tensor S = c.alloc({k}, false);
  Tensor E = c.alloc({std::max<int64_t>(k, 1)}, false);
  DevInt info(c);
  hipEvent_t ev = nullptr;
  c.tbegin("svd", &ev);
  c.count("svd", (A.cplx ? 4.0 : 1.0) * (4.0 * m * n * k + 8.0 * k * k * k), 0);
  if (A.cplx)
    ck(rocsolver_zgesvd(c.blas, rocblas_svect_singular, rocblas_svect_singular, int(m), int(n),
                        (rocblas_double_complex*)W.data(), int(m), S.data(),
                        (rocblas_double_complex*)r.U.data(), int(m),
                        (rocblas_double_complex*)r.Vt.data(), int(k), E.data(),
                        rocblas_outofplace, info.p), "zgesvd");
  else
    ck(rocsolver_dgesvd(c.blas, rocblas_svect_singular, rocblas_svect_singular, int(m), int(n),
                        W.data(), int(m), S.data(), r.U.data(), int(m), r.Vt.data(), int(k),
                        E.data(), rocblas_outofplace, info.p), "dgesvd");
  c.tend("svd", ev);
  int inf = info.get(c);
  PEPS_CHECK(inf == 0, PS_ESOLVER, "gesvd did not converge (info=" + std::to_string(inf) + ")");
  r.S = c.to_host(S);
  return r;
}

EIG eigh(Ctx& c, const Tensor& A) {
  // Hermitian eigendecomposition by the project's batched Jacobi kernel
  // (kernels.hip); eigenvalues are returned unsorted.
  PEPS_CHECK(A.rank() == 2 && A.dims[0] == A.dims[1], PS_ESHAPE, "eigh: square matrix expected");
  int64_t n = A.dims[0];
  EIG r;
  r.V = c.alloc({n, n}, A.cplx);
  if (n == 0) return r;
  check_finite(c, A, "eigh");
  Tensor W = copy_tensor(c, A);
  Tensor w = c.alloc({n}, false);
  jacobi_eigh(c, W, r.V, w, int(n), 1, 1e-15, 60);
  r.w = c.to_host(w);
  return r;
}

}  // namespace peps
