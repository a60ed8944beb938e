// Context, caching device allocator and host<->device transfer helpers.
#include "peps_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace peps {

Pool::~Pool() { trim(); }

void* Pool::get(size_t bytes) {
  bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
  auto it = free_.lower_bound(bytes);
  // best fit, reusing a block at most 2x the request
  if (it != free_.end() && it->first <= 2 * bytes + (1 << 20)) {
    void* p = it->second;
    cached_ -= it->first;
    in_use_ += it->first;
    free_.erase(it);
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(PS_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
  }
  sizes_[p] = bytes;
  in_use_ += bytes;
  return p;
}

void Pool::put(void* p, size_t /*bytes*/) {
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  size_t b = it->second;
  in_use_ -= b;
  cached_ += b;
  free_.emplace(b, p);
}

void Pool::trim() {
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    sizes_.erase(kv.second);
  }
  free_.clear();
  cached_ = 0;
}

Tensor Tensor::reshaped(std::vector<int64_t> nd) const {
  int64_t n = 1;
  for (auto d : nd) n *= d;
  PEPS_CHECK(n == numel(), PS_ESHAPE, "reshape: element count mismatch");
  Tensor t = *this;
  t.dims = std::move(nd);
  return t;
}

Tensor Tensor::slice_last(int64_t i0, int64_t i1) const {
  PEPS_CHECK(rank() >= 1 && 0 <= i0 && i0 <= i1 && i1 <= dims.back(), PS_ESHAPE,
             "slice_last: range out of bounds");
  Tensor t = *this;
  int64_t inner = numel() / std::max<int64_t>(1, dims.back());
  t.off = off + i0 * inner;
  t.dims.back() = i1 - i0;
  return t;
}

Ctx::Ctx(int dev) : device(dev) {
  PEPS_HIP(hipSetDevice(dev));
  scratch_attach(dev);  // the device address of the no-allocation GEMM's fixed workspace
  PEPS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (rocblas_create_handle(&blas) != rocblas_status_success)
    throw Error(PS_EHIP, "rocblas_create_handle failed");
  rocblas_set_stream(blas, stream);
  pool.reset(new Pool(dev));
  PEPS_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage_host), kStageBytes, hipHostMallocDefault));
  PEPS_HIP(hipEventCreateWithFlags(&stage_ev, hipEventDisableTiming));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  if (const char* v = std::getenv("PEPS_HIP_VERBOSE")) verbose = std::atoi(v);
  if (const char* v = std::getenv("PEPS_HIP_SUBSPACE_OVERSAMPLE")) subspace_oversample = std::max(0, std::atoi(v));
  if (const char* v = std::getenv("PEPS_HIP_TRD")) trd_backend = std::string(v) == "own" ? 0 : 1;
  if (const char* v = std::getenv("PEPS_HIP_LEADING_SVD")) leading_svd = std::atoi(v) != 0;
  if (const char* v = std::getenv("PEPS_HIP_DEBUG_FINITE")) debug_finite = std::atoi(v) != 0;
  if (const char* v = std::getenv("PEPS_HIP_ALS_BLOCKS")) als_blocks = std::atoi(v) != 0;
  if (const char* v = std::getenv("PEPS_HIP_EIG_DIRECT_MAX")) eig_direct_max = std::max(0, std::atoi(v));
  if (const char* v = std::getenv("PEPS_HIP_GEMM_SHAPES")) gemm_shapes = std::atoi(v) != 0;
  if (const char* v = std::getenv("PEPS_HIP_QR")) qr_backend = std::string(v) == "blocked" ? 0 : 1;
}

Ctx::~Ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& p : pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ev_free) (void)hipEventDestroy(e);
  // cached tensors return their buffers to the pool: release them first
  ritz_cache.clear();
  pool.reset();
  if (blas) rocblas_destroy_handle(blas);
  if (stream) (void)hipStreamDestroy(stream);
  if (stage_ev) (void)hipEventDestroy(stage_ev);
  if (stage_host) (void)hipHostFree(stage_host);
}

static hipEvent_t take_event(Ctx& c) {
  if (!c.ev_free.empty()) {
    hipEvent_t e = c.ev_free.back();
    c.ev_free.pop_back();
    return e;
  }
  hipEvent_t e;
  PEPS_HIP(hipEventCreate(&e));
  return e;
}

void Ctx::tbegin(const std::string& cls, hipEvent_t* a) {
  *a = nullptr;
  if (!timing) return;
  if (timing_gemm_only && cls != "gemm_f64_tiled" && cls != "gemm_f32_tiled" && cls != "gemm_bf16") return;
  *a = take_event(*this);
  PEPS_HIP(hipEventRecord(*a, stream));
}

void Ctx::tend(const std::string& cls, hipEvent_t a) {
  if (!timing || !a) return;
  hipEvent_t b = take_event(*this);
  PEPS_HIP(hipEventRecord(b, stream));
  pending.push_back({cls, a, b});
  if (pending.size() > 4096) resolve_timing();
}

void Ctx::resolve_timing() {
  if (pending.empty()) return;
  PEPS_HIP(hipStreamSynchronize(stream));
  for (auto& p : pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) stats[p.cls].ms += ms;
    ev_free.push_back(p.a);
    ev_free.push_back(p.b);
  }
  pending.clear();
}

void Ctx::note(const std::string& what) {
  if (!verbose) return;
  static const auto t0 = std::chrono::steady_clock::now();
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(stderr, "[peps %9.2fs] %s\n", t, what.c_str());
  std::fflush(stderr);
}

Tensor Ctx::alloc(std::vector<int64_t> dims, bool cplx) {
  Tensor t;
  t.dims = std::move(dims);
  t.cplx = cplx;
  t.buf = std::make_shared<Buffer>();
  t.buf->pool = pool.get();
  t.buf->bytes = std::max<size_t>(t.bytes(), 8);
  t.buf->ptr = pool->get(t.buf->bytes);
  return t;
}

Tensor Ctx::zeros(std::vector<int64_t> dims, bool cplx) {
  Tensor t = alloc(std::move(dims), cplx);
  if (t.bytes()) PEPS_HIP(hipMemsetAsync(t.data(), 0, t.bytes(), stream));
  return t;
}

void Ctx::h2d_async(void* dst, const void* host, size_t bytes) {
  if (!bytes) return;
  const size_t need = (bytes + 255) & ~size_t(255);
  if (!stage_host || need > kStageBytes / 4) {
    PEPS_HIP(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, stream));
    PEPS_HIP(hipStreamSynchronize(stream));
    return;
  }
  if (stage_head + need > kStageBytes) {
    // wrap: every copy staged so far is on the stream before this record, so
    // once it completes the whole ring is free (one wait per ~16 MiB staged)
    if (stage_used) {
      PEPS_HIP(hipEventRecord(stage_ev, stream));
      PEPS_HIP(hipEventSynchronize(stage_ev));
    }
    stage_head = 0;
  }
  std::memcpy(stage_host + stage_head, host, bytes);
  PEPS_HIP(hipMemcpyAsync(dst, stage_host + stage_head, bytes, hipMemcpyHostToDevice, stream));
  stage_used = true;
  stage_head += need;
}

Tensor Ctx::upload(std::vector<int64_t> dims, bool cplx, const double* host) {
  Tensor t = alloc(std::move(dims), cplx);
  h2d_async(t.data(), host, t.bytes());
  return t;
}

void Ctx::download(const Tensor& t, double* host) {
  if (!t.bytes()) return;
  PEPS_HIP(hipMemcpyAsync(host, t.data(), t.bytes(), hipMemcpyDeviceToHost, stream));
  PEPS_HIP(hipStreamSynchronize(stream));
}

std::vector<double> Ctx::to_host(const Tensor& t) {
  std::vector<double> h(size_t(t.numel()) * (t.cplx ? 2 : 1));
  download(t, h.data());
  return h;
}

void Ctx::sync() { PEPS_HIP(hipStreamSynchronize(stream)); }

void* Ctx::scratch_upload(const void* host, size_t bytes, std::shared_ptr<Buffer>* keep) {
  auto b = std::make_shared<Buffer>();
  b->pool = pool.get();
  b->bytes = std::max<size_t>(bytes, 8);
  b->ptr = pool->get(b->bytes);
  h2d_async(b->ptr, host, bytes);
  *keep = b;
  return b->ptr;
}

}  // namespace peps
