// Exact diagonalisation of the finite U(1) lattice gauge theory in one
// Gauss-law sector (finite_ed.jl), on the GPU.
//
// Basis (build_basis, finite_ed.jl:125-233).  A basis state is a fermion
// mask with N_f = -sum(g) bits plus the free E_U values of rows 1..ny-1
// (columns 1..nx-1); the last E_U of each row and every E_R follow from the
// Gauss law and must stay within [-dg, dg].  Candidates are numbered mask
// by mask (masks ascending) and, within a mask, by the free E_U values read
// as a base-(2dg+1) number with E_U(1,1) the most significant digit: the
// reference's recursion order.  One workgroup per mask counts its valid
// candidates; a second pass writes each valid state's key at the mask's
// offset plus its rank among the mask's valid candidates (wave ballots and
// an LDS prefix over the four waves), so the basis comes out in the
// reference's order.
//
// Keys follow finite_ed.jl:79-95: n_f in bits 0..N-1 (iy outer, ix inner),
// then E_R + dg and E_U + dg in lb-bit fields (lb = 2 for dg = 1, as in the
// reference; wider for larger dg).  A sorted copy of the keys (hipCUB radix
// sort, with each key's basis index) answers the lookups of the Hamiltonian.
//
// Hamiltonian (build_hamiltonian, finite_ed.jl:252-326), ELL form: one
// thread per state writes its diagonal (g^2/2 sum E^2 + m sum (-1)^(ix+iy)
// n_f) and the basis indices of the states one hop away -- the forward hop
// (left/lower site filled from the right/upper one, link raised) and its
// reverse -- each coupled by -t.  The hopped key is the state's key plus or
// minus the three changed fields, looked up by binary search.
//
// Ground state (find_ground_state, finite_ed.jl:333-341): Lanczos with full
// reorthogonalisation (twice per step, as GEMVs), restarted from the lowest
// Ritz vector until ||H psi - E0 psi|| <= tol; the small tridiagonal
// eigenproblem is solved on the host by cyclic Jacobi.
//
// Observables (measure_ED, finite_ed.jl:369-395): per-block LDS sums of
// |psi_i|^2 n_f and |psi_i|^2 (E_R^2 + E_U^2) per site, reduced on the host.
#include "ed.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>

namespace peps {
namespace {

constexpr int kEdMaxSites = 32;

struct EdGeom {
  int nx, ny, dg, lb, nsite, nh, nv, F, B;
  int64_t K;  // candidates per mask: B^F
  int g[kEdMaxSites];
};

std::shared_ptr<Buffer> dev_bytes(Ctx& c, size_t bytes) {
  auto b = std::make_shared<Buffer>();
  b->pool = c.pool.get();
  b->bytes = std::max<size_t>(bytes, 8);
  b->ptr = c.pool->get(b->bytes);
  return b;
}

// Row-cumulative divergence S[iy] = sum_{rows <= iy} sum_ix (n_f + g).
__device__ void ed_row_sums(const EdGeom& G, uint32_t mask, int* S) {
  int acc = 0;
  for (int iy = 0; iy < G.ny; ++iy) {
    for (int ix = 0; ix < G.nx; ++ix) acc += int((mask >> (iy * G.nx + ix)) & 1u) + G.g[iy * G.nx + ix];
    S[iy] = acc;
  }
}

// Candidate cidx of mask: whether the Gauss law closes with every field in
// [-dg, dg], and its key.
__device__ bool ed_candidate(const EdGeom& G, uint32_t mask, const int* S, int64_t cidx, uint64_t* key) {
  int eU[kEdMaxSites];  // eU[iy * nx + ix], rows iy < ny - 1
  int64_t rem = cidx;
  for (int f = G.F - 1; f >= 0; --f) {
    const int iy = f / (G.nx - 1), ix = f % (G.nx - 1);
    eU[iy * G.nx + ix] = int(rem % G.B) - G.dg;
    rem /= G.B;
  }
  for (int iy = 0; iy < G.ny - 1; ++iy) {
    int sum = 0;
    for (int ix = 0; ix < G.nx - 1; ++ix) sum += eU[iy * G.nx + ix];
    const int last = S[iy] - sum;
    if (last < -G.dg || last > G.dg) return false;
    eU[iy * G.nx + G.nx - 1] = last;
  }
  uint64_t k = mask;
  for (int iy = 0; iy < G.ny; ++iy) {
    int acc = 0;
    for (int ix = 0; ix < G.nx; ++ix) {
      const int nf = int((mask >> (iy * G.nx + ix)) & 1u);
      const int now = iy < G.ny - 1 ? eU[iy * G.nx + ix] : 0;
      const int prev = iy > 0 ? eU[(iy - 1) * G.nx + ix] : 0;
      acc += nf + G.g[iy * G.nx + ix] - now + prev;
      if (ix < G.nx - 1) {
        if (acc < -G.dg || acc > G.dg) return false;
        k |= uint64_t(acc + G.dg) << (G.nsite + G.lb * (iy * (G.nx - 1) + ix));
      } else if (acc != 0) {
        return false;
      }
    }
  }
  int off = G.nsite + G.lb * G.nh;
  for (int iy = 0; iy < G.ny - 1; ++iy)
    for (int ix = 0; ix < G.nx; ++ix) {
      k |= uint64_t(eU[iy * G.nx + ix] + G.dg) << off;
      off += G.lb;
    }
  *key = k;
  return true;
}

__global__ __launch_bounds__(256) void ed_count_kernel(EdGeom G, const uint32_t* __restrict__ masks,
                                                       int64_t* __restrict__ counts) {
  __shared__ int S[kEdMaxSites];
  __shared__ long long wsum[4];
  const uint32_t mask = masks[blockIdx.x];
  if (threadIdx.x == 0) ed_row_sums(G, mask, S);
  __syncthreads();
  long long cnt = 0;
  for (int64_t c = threadIdx.x; c < G.K; c += blockDim.x) {
    uint64_t k;
    cnt += ed_candidate(G, mask, S, c, &k) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void ed_write_kernel(EdGeom G, const uint32_t* __restrict__ masks,
                                                       const int64_t* __restrict__ offsets,
                                                       uint64_t* __restrict__ keys) {
  __shared__ int S[kEdMaxSites];
  __shared__ int wcount[4];
  const uint32_t mask = masks[blockIdx.x];
  if (threadIdx.x == 0) ed_row_sums(G, mask, S);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t base = offsets[blockIdx.x];
  for (int64_t c0 = 0; c0 < G.K; c0 += 256) {  // uniform trip count: the barriers are safe
    const int64_t c = c0 + threadIdx.x;
    uint64_t k = 0;
    const bool v = c < G.K && ed_candidate(G, mask, S, c, &k);
    const unsigned long long bal = __ballot(v);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; ++w) woff += wcount[w];
    const int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (v) keys[base + woff + before] = k;
    base += total;
    __syncthreads();
  }
}

__global__ void ed_iota_kernel(int* __restrict__ v, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    v[i] = int(i);
}

__device__ int ed_lookup(const uint64_t* __restrict__ sorted, const int* __restrict__ perm, int64_t n,
                         uint64_t key) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const uint64_t v = sorted[mid];
    if (v == key) return perm[mid];
    if (v < key) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

__global__ void ed_h_kernel(EdGeom G, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ sorted,
                            const int* __restrict__ perm, int64_t dim, double ge, double m, int W,
                            double* __restrict__ diag, int* __restrict__ cols) {
  const uint64_t fmask = (1ull << G.lb) - 1ull;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < dim; i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    double e2 = 0.0, mass = 0.0;
    int slot = 0;
    for (int iy = 0; iy < G.ny; ++iy)
      for (int ix = 0; ix < G.nx; ++ix)
        if ((k >> (iy * G.nx + ix)) & 1ull) mass += ((ix + iy) & 1) ? -1.0 : 1.0;
    for (int iy = 0; iy < G.ny; ++iy)
      for (int ix = 0; ix < G.nx - 1; ++ix) {
        const int off = G.nsite + G.lb * (iy * (G.nx - 1) + ix);
        const int e = int((k >> off) & fmask) - G.dg;
        e2 += double(e) * e;
        const int pL = iy * G.nx + ix, pR = pL + 1;
        const int nL = int((k >> pL) & 1ull), nR = int((k >> pR) & 1ull);
        if (nL == 0 && nR == 1 && e < G.dg)
          cols[int64_t(slot++) * dim + i] = ed_lookup(sorted, perm, dim, k + (1ull << pL) - (1ull << pR) + (1ull << off));
        if (nL == 1 && nR == 0 && e > -G.dg)
          cols[int64_t(slot++) * dim + i] = ed_lookup(sorted, perm, dim, k - (1ull << pL) + (1ull << pR) - (1ull << off));
      }
    for (int iy = 0; iy < G.ny - 1; ++iy)
      for (int ix = 0; ix < G.nx; ++ix) {
        const int off = G.nsite + G.lb * (G.nh + iy * G.nx + ix);
        const int e = int((k >> off) & fmask) - G.dg;
        e2 += double(e) * e;
        const int pD = iy * G.nx + ix, pU = pD + G.nx;
        const int nD = int((k >> pD) & 1ull), nU = int((k >> pU) & 1ull);
        if (nD == 0 && nU == 1 && e < G.dg)
          cols[int64_t(slot++) * dim + i] = ed_lookup(sorted, perm, dim, k + (1ull << pD) - (1ull << pU) + (1ull << off));
        if (nD == 1 && nU == 0 && e > -G.dg)
          cols[int64_t(slot++) * dim + i] = ed_lookup(sorted, perm, dim, k - (1ull << pD) + (1ull << pU) - (1ull << off));
      }
    for (; slot < W; ++slot) cols[int64_t(slot) * dim + i] = -1;
    diag[i] = 0.5 * ge * ge * e2 + m * mass;
  }
}

__global__ void ed_spmv_kernel(int64_t dim, int W, const double* __restrict__ diag, const int* __restrict__ cols,
                               double t, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < dim; i += int64_t(gridDim.x) * blockDim.x) {
    double h = 0.0;
    for (int w = 0; w < W; ++w) {
      const int c = cols[int64_t(w) * dim + i];
      if (c >= 0) h += x[c];
    }
    y[i] = diag[i] * x[i] - t * h;
  }
}

// The same product for an interleaved complex x (H is real).
__global__ void ed_spmv_cplx_kernel(int64_t dim, int W, const double* __restrict__ diag,
                                    const int* __restrict__ cols, double t, const double* __restrict__ x,
                                    double* __restrict__ y) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < dim; i += int64_t(gridDim.x) * blockDim.x) {
    double hr = 0.0, hi = 0.0;
    for (int w = 0; w < W; ++w) {
      const int c = cols[int64_t(w) * dim + i];
      if (c >= 0) {
        hr += x[2 * int64_t(c)];
        hi += x[2 * int64_t(c) + 1];
      }
    }
    y[2 * i] = diag[i] * x[2 * i] - t * hr;
    y[2 * i + 1] = diag[i] * x[2 * i + 1] - t * hi;
  }
}

// |x_i| of an interleaved complex vector.
__global__ void ed_cabs_kernel(int64_t dim, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < dim; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = sqrt(x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1]);
}

__global__ __launch_bounds__(256) void ed_measure_kernel(EdGeom G, const uint64_t* __restrict__ keys,
                                                         const double* __restrict__ psi, int64_t dim,
                                                         double* __restrict__ part) {
  __shared__ double acc[2 * kEdMaxSites];
  const int ns = G.nsite;
  for (int j = threadIdx.x; j < 2 * ns; j += blockDim.x) acc[j] = 0.0;
  __syncthreads();
  const uint64_t fmask = (1ull << G.lb) - 1ull;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < dim; i += int64_t(gridDim.x) * blockDim.x) {
    const double p = psi[i] * psi[i];
    if (p == 0.0) continue;
    const uint64_t k = keys[i];
    for (int iy = 0; iy < G.ny; ++iy)
      for (int ix = 0; ix < G.nx; ++ix) {
        const int s = iy * G.nx + ix;
        if ((k >> s) & 1ull) atomicAdd(&acc[s], p);
        double e2 = 0.0;
        if (ix < G.nx - 1) {
          const int e = int((k >> (G.nsite + G.lb * (iy * (G.nx - 1) + ix))) & fmask) - G.dg;
          e2 += double(e) * e;
        }
        if (iy < G.ny - 1) {
          const int e = int((k >> (G.nsite + G.lb * (G.nh + iy * G.nx + ix))) & fmask) - G.dg;
          e2 += double(e) * e;
        }
        if (e2 != 0.0) atomicAdd(&acc[ns + s], p * e2);
      }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 2 * ns; j += blockDim.x) part[int64_t(blockIdx.x) * 2 * ns + j] = acc[j];
}

EdGeom geom_of(const EDBasis& b) {
  EdGeom G{};
  G.nx = b.nx;
  G.ny = b.ny;
  G.dg = b.dg;
  G.lb = b.lb;
  G.nsite = b.nsite;
  G.nh = b.nh;
  G.nv = b.nv;
  G.F = (b.nx - 1) * (b.ny - 1);
  G.B = 2 * b.dg + 1;
  G.K = 1;
  for (int f = 0; f < G.F; ++f) G.K *= G.B;
  for (int s = 0; s < b.nsite; ++s) G.g[s] = int(b.g[s]);
  return G;
}

unsigned grid_for(int64_t n, int64_t cap = 16384) {
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)));
}

// Cyclic Jacobi eigensolver of a small dense symmetric matrix (row-major
// k x k, overwritten); V receives the eigenvectors as columns (Golub & Van
// Loan, algorithm 8.5.3).
void jacobi_host(std::vector<double>& A, int k, std::vector<double>& w, std::vector<double>& V) {
  V.assign(size_t(k) * k, 0.0);
  for (int i = 0; i < k; ++i) V[size_t(i) * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        tot += A[size_t(i) * k + j] * A[size_t(i) * k + j];
        if (i != j) off += A[size_t(i) * k + j] * A[size_t(i) * k + j];
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[size_t(p) * k + q];
        if (apq == 0.0) continue;
        const double app = A[size_t(p) * k + p], aqq = A[size_t(q) * k + q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (int r = 0; r < k; ++r) {  // A <- A J
          const double arp = A[size_t(r) * k + p], arq = A[size_t(r) * k + q];
          A[size_t(r) * k + p] = cs * arp - sn * arq;
          A[size_t(r) * k + q] = sn * arp + cs * arq;
        }
        for (int r = 0; r < k; ++r) {  // A <- J^T A
          const double apr = A[size_t(p) * k + r], aqr = A[size_t(q) * k + r];
          A[size_t(p) * k + r] = cs * apr - sn * aqr;
          A[size_t(q) * k + r] = sn * apr + cs * aqr;
        }
        for (int r = 0; r < k; ++r) {  // V <- V J
          const double vrp = V[size_t(r) * k + p], vrq = V[size_t(r) * k + q];
          V[size_t(r) * k + p] = cs * vrp - sn * vrq;
          V[size_t(r) * k + q] = sn * vrp + cs * vrq;
        }
      }
  }
  w.resize(k);
  for (int i = 0; i < k; ++i) w[i] = A[size_t(i) * k + i];
}

}  // namespace

std::unique_ptr<EDBasis> ed_build_basis(Ctx& c, int nx, int ny, int dg, const std::vector<int64_t>& g) {
  PEPS_CHECK(nx >= 2 && ny >= 2 && nx * ny <= kEdMaxSites, PS_EUNSUPPORTED,
             "ed: lattices from 2x2 up to 32 sites");
  PEPS_CHECK(dg >= 1 && dg <= 3, PS_EUNSUPPORTED, "ed: dg in [1, 3]");
  PEPS_CHECK(int64_t(g.size()) == int64_t(nx) * ny, PS_EARG, "ed: nx*ny background charges expected");
  auto b = std::make_unique<EDBasis>();
  b->c = &c;
  b->nx = nx;
  b->ny = ny;
  b->dg = dg;
  b->lb = 2;
  while ((1 << b->lb) < 2 * dg + 1) ++b->lb;
  b->nsite = nx * ny;
  b->nh = (nx - 1) * ny;
  b->nv = nx * (ny - 1);
  b->g = g;
  PEPS_CHECK(b->nsite + b->lb * (b->nh + b->nv) <= 64, PS_EUNSUPPORTED, "ed: the state key exceeds 64 bits");
  int64_t nf = 0;
  for (int64_t q : g) nf -= q;
  PEPS_CHECK(nf >= 0 && nf <= b->nsite, PS_EARG, "ed: inconsistent sector: N_f = -sum(g) outside [0, nx*ny]");
  const EdGeom G = geom_of(*b);
  // fermion masks with N_f bits, ascending
  std::vector<uint32_t> masks;
  for (uint64_t mk = 0; mk < (uint64_t(1) << b->nsite); ++mk)
    if (__builtin_popcountll(mk) == nf) masks.push_back(uint32_t(mk));
  PEPS_CHECK(!masks.empty() && double(masks.size()) * double(G.K) <= double(int64_t(1) << 38), PS_EUNSUPPORTED,
             "ed: more than 2^38 candidate states");
  PEPS_CHECK(masks.size() < (size_t(1) << 31), PS_EUNSUPPORTED, "ed: too many fermion masks");
  std::shared_ptr<Buffer> km;
  auto* dmasks = (const uint32_t*)c.scratch_upload(masks.data(), masks.size() * sizeof(uint32_t), &km);
  auto counts = dev_bytes(c, masks.size() * sizeof(int64_t));
  hipLaunchKernelGGL(ed_count_kernel, dim3(unsigned(masks.size())), dim3(256), 0, c.stream, G, dmasks,
                     (int64_t*)counts->ptr);
  PEPS_HIP(hipGetLastError());
  std::vector<int64_t> hc(masks.size());
  PEPS_HIP(hipMemcpyAsync(hc.data(), counts->ptr, hc.size() * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  std::vector<int64_t> offs(masks.size());
  int64_t dim = 0;
  for (size_t i = 0; i < masks.size(); ++i) {
    offs[i] = dim;
    dim += hc[i];
  }
  PEPS_CHECK(dim > 0, PS_EARG, "ed: the sector holds no state");
  PEPS_CHECK(dim < (int64_t(1) << 31), PS_EUNSUPPORTED, "ed: basis dimension beyond 2^31");
  b->dim = dim;
  std::shared_ptr<Buffer> ko;
  auto* doffs = (const int64_t*)c.scratch_upload(offs.data(), offs.size() * sizeof(int64_t), &ko);
  b->keys = dev_bytes(c, size_t(dim) * sizeof(uint64_t));
  hipLaunchKernelGGL(ed_write_kernel, dim3(unsigned(masks.size())), dim3(256), 0, c.stream, G, dmasks, doffs,
                     (uint64_t*)b->keys->ptr);
  PEPS_HIP(hipGetLastError());
  // the sorted keys and the basis index of each
  auto idx = dev_bytes(c, size_t(dim) * sizeof(int));
  hipLaunchKernelGGL(ed_iota_kernel, dim3(grid_for(dim)), dim3(256), 0, c.stream, (int*)idx->ptr, dim);
  PEPS_HIP(hipGetLastError());
  b->sorted = dev_bytes(c, size_t(dim) * sizeof(uint64_t));
  b->perm = dev_bytes(c, size_t(dim) * sizeof(int));
  size_t tmp_bytes = 0;
  PEPS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (const uint64_t*)b->keys->ptr,
                                              (uint64_t*)b->sorted->ptr, (const int*)idx->ptr, (int*)b->perm->ptr,
                                              int(dim), 0, 64, c.stream));
  auto tmp = dev_bytes(c, tmp_bytes);
  PEPS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp->ptr, tmp_bytes, (const uint64_t*)b->keys->ptr,
                                              (uint64_t*)b->sorted->ptr, (const int*)idx->ptr, (int*)b->perm->ptr,
                                              int(dim), 0, 64, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));  // the uploaded masks and offsets are released on return
  return b;
}

void ed_build_hamiltonian(EDBasis& b, double g, double t, double m) {
  Ctx& c = *b.c;
  const EdGeom G = geom_of(b);
  b.W = b.nh + b.nv;  // one hop per bond at most: forward and reverse exclude each other
  b.t_hop = t;
  b.diag = c.alloc({b.dim}, false);
  b.cols = dev_bytes(c, size_t(b.dim) * b.W * sizeof(int));
  hipLaunchKernelGGL(ed_h_kernel, dim3(grid_for(b.dim)), dim3(256), 0, c.stream, G, (const uint64_t*)b.keys->ptr,
                     (const uint64_t*)b.sorted->ptr, (const int*)b.perm->ptr, b.dim, g, m, b.W, b.diag.data(),
                     (int*)b.cols->ptr);
  PEPS_HIP(hipGetLastError());
  b.has_h = true;
  b.has_gs = false;
}

void ed_apply(EDBasis& b, const Tensor& x, Tensor& y) {
  PEPS_CHECK(b.has_h, PS_EARG, "ed: build the Hamiltonian first");
  PEPS_CHECK(x.numel() == b.dim && y.numel() == b.dim && !x.cplx && !y.cplx, PS_ESHAPE, "ed: vector size");
  Ctx& c = *b.c;
  hipLaunchKernelGGL(ed_spmv_kernel, dim3(grid_for(b.dim)), dim3(256), 0, c.stream, b.dim, b.W, b.diag.data(),
                     (const int*)b.cols->ptr, b.t_hop, x.data(), y.data());
  PEPS_HIP(hipGetLastError());
}

void ed_apply_complex(EDBasis& b, const Tensor& x, Tensor& y) {
  PEPS_CHECK(b.has_h, PS_EARG, "ed: build the Hamiltonian first");
  PEPS_CHECK(x.numel() == b.dim && y.numel() == b.dim && x.cplx && y.cplx, PS_ESHAPE, "ed: complex vector size");
  Ctx& c = *b.c;
  hipLaunchKernelGGL(ed_spmv_cplx_kernel, dim3(grid_for(b.dim)), dim3(256), 0, c.stream, b.dim, b.W, b.diag.data(),
                     (const int*)b.cols->ptr, b.t_hop, x.data(), y.data());
  PEPS_HIP(hipGetLastError());
}

Tensor ed_exp_step(EDBasis& b, const Tensor& psi0, double dt, int krylov, double tol, int* substeps) {
  // exponentiate(H, -i dt, psi; krylovdim, tol, ishermitian = true), the
  // call of evolve_ED! (finite_ed.jl:417-419): a Lanczos basis of psi with
  // full reorthogonalisation, psi(h) = |psi| V exp(-i h T) e_1, and the
  // error estimate |psi| beta_k |e_k^T exp(-i h T) e_1| of the truncated
  // projection; while it exceeds tol the interval is halved (the basis does
  // not depend on h), and the rest of dt is covered by further projections.
  PEPS_CHECK(b.has_h, PS_EARG, "ed: build the Hamiltonian first");
  PEPS_CHECK(psi0.cplx && psi0.numel() == b.dim && krylov >= 2 && tol > 0.0, PS_EARG, "ed_evolve: bad input");
  Ctx& c = *b.c;
  TierGuard fp64(c, TIER_FP64);
  const int64_t n = b.dim;
  const int m = int(std::min<int64_t>(krylov, n));
  Tensor V = c.alloc({n, int64_t(m)}, true);
  Tensor w = c.alloc({n}, true);
  Tensor psi = copy_tensor(c, psi0.reshaped({n}));
  double left = dt;
  int nsub = 0;
  while (left > 0.0) {
    const double nu = std::sqrt(norm2sq(c, psi));
    PEPS_CHECK(nu > 0.0 && std::isfinite(nu), PS_ESOLVER, "ed_evolve: zero or non-finite state");
    Tensor v0 = V.slice_last(0, 1);
    PEPS_HIP(hipMemcpyAsync(v0.data(), psi.data(), size_t(n) * 16, hipMemcpyDeviceToDevice, c.stream));
    scale(c, v0, 1.0 / nu);
    std::vector<double> alpha, beta;
    int k = 0;
    double bk = 0.0;
    for (int j = 0; j < m; ++j) {
      Tensor vj = V.slice_last(j, j + 1).reshaped({n});
      ed_apply_complex(b, vj, w);
      double a = dot(c, w, vj, true).re;
      axpy_inplace(c, w, vj, -a);
      if (j > 0) axpy_inplace(c, w, V.slice_last(j - 1, j).reshaped({n}), -beta[j - 1]);
      Tensor Vj = V.slice_last(0, j + 1);
      Tensor wm = w.reshaped({n, 1});
      for (int pass = 0; pass < 2; ++pass) {  // full reorthogonalisation
        Tensor h = matmul(c, Vj, wm, true, false, true, false);
        a += c.to_host(h)[2 * size_t(j)];
        axpy_inplace(c, wm, matmul(c, Vj, h), -1.0);
      }
      bk = std::sqrt(norm2sq(c, w));
      alpha.push_back(a);
      k = j + 1;
      if (j + 1 == m || bk <= 1e-14 * std::max(1.0, std::fabs(a))) break;
      beta.push_back(bk);
      Tensor vn = V.slice_last(j + 1, j + 2);
      PEPS_HIP(hipMemcpyAsync(vn.data(), w.data(), size_t(n) * 16, hipMemcpyDeviceToDevice, c.stream));
      scale(c, vn, 1.0 / bk);
    }
    const bool exhausted = bk <= 1e-14 * std::max(1.0, std::fabs(alpha.back()));
    std::vector<double> T(size_t(k) * k, 0.0), ev, Z;
    for (int i = 0; i < k; ++i) {
      T[size_t(i) * k + i] = alpha[i];
      if (i + 1 < k) T[size_t(i) * k + i + 1] = T[size_t(i + 1) * k + i] = beta[i];
    }
    jacobi_host(T, k, ev, Z);
    auto coeffs = [&](double h) {
      std::vector<double> co(2 * size_t(k), 0.0);
      for (int i = 0; i < k; ++i)
        for (int l = 0; l < k; ++l) {
          const double f = Z[size_t(i) * k + l] * Z[size_t(l)];
          co[2 * i] += f * std::cos(h * ev[l]);
          co[2 * i + 1] -= f * std::sin(h * ev[l]);
        }
      return co;
    };
    double h = left;
    std::vector<double> co = coeffs(h);
    int halvings = 0;
    while (!exhausted && nu * bk * std::hypot(co[2 * (k - 1)], co[2 * (k - 1) + 1]) > tol) {
      PEPS_CHECK(++halvings <= 40, PS_ESOLVER, "ed_evolve: the Krylov error estimate does not fall below tol");
      h *= 0.5;
      co = coeffs(h);
    }
    for (double& x : co) x *= nu;
    Tensor cv = c.upload({int64_t(k), 1}, true, co.data());
    psi = matmul(c, V.slice_last(0, k), cv).reshaped({n});
    left = (h >= left) ? 0.0 : left - h;
    ++nsub;
  }
  if (substeps) *substeps = nsub;
  c.count("ed_evolve[projections]", double(nsub), 0);
  return psi;
}

double ed_ground_state(EDBasis& b, double tol, int krylov, int max_restarts) {
  PEPS_CHECK(b.has_h, PS_EARG, "ed: build the Hamiltonian first");
  PEPS_CHECK(tol > 0.0 && krylov >= 2 && max_restarts >= 1, PS_EARG, "ed: bad Lanczos parameters");
  Ctx& c = *b.c;
  TierGuard fp64(c, TIER_FP64);
  const int64_t n = b.dim;
  const int m = int(std::min<int64_t>(krylov, n));
  Tensor V = c.alloc({n, int64_t(m) + 1}, false);
  Tensor w = c.alloc({n, 1}, false);
  Tensor y = c.alloc({n, 1}, false);
  fill_random(c, y, 0x1a2c05ull + uint64_t(n));
  double theta = 0.0;
  for (int restart = 0; restart < max_restarts; ++restart) {
    scale(c, y, 1.0 / std::sqrt(norm2sq(c, y)));
    Tensor v0 = V.slice_last(0, 1);
    PEPS_HIP(hipMemcpyAsync(v0.data(), y.data(), size_t(n) * 8, hipMemcpyDeviceToDevice, c.stream));
    std::vector<double> alpha, beta;
    int k = 0;
    for (int j = 0; j < m; ++j) {
      Tensor vj = V.slice_last(j, j + 1);
      ed_apply(b, vj, w);
      double a = dot(c, w, vj, false).re;
      axpy_inplace(c, w, vj, -a);
      if (j > 0) axpy_inplace(c, w, V.slice_last(j - 1, j), -beta[j - 1]);
      Tensor Vj = V.slice_last(0, j + 1);
      for (int pass = 0; pass < 2; ++pass) {  // full reorthogonalisation
        Tensor h = matmul(c, Vj, w, false, false, true, false);
        const std::vector<double> hh = c.to_host(h);
        a += hh[j];
        axpy_inplace(c, w, matmul(c, Vj, h), -1.0);
      }
      const double bn = std::sqrt(norm2sq(c, w));
      alpha.push_back(a);
      beta.push_back(bn);
      k = j + 1;
      if (j + 1 == m || bn <= 1e-14 * std::max(1.0, std::fabs(a))) break;
      Tensor vn = V.slice_last(j + 1, j + 2);
      PEPS_HIP(hipMemcpyAsync(vn.data(), w.data(), size_t(n) * 8, hipMemcpyDeviceToDevice, c.stream));
      scale(c, vn, 1.0 / bn);
    }
    std::vector<double> T(size_t(k) * k, 0.0), ev, Z;
    for (int i = 0; i < k; ++i) {
      T[size_t(i) * k + i] = alpha[i];
      if (i + 1 < k) T[size_t(i) * k + i + 1] = T[size_t(i + 1) * k + i] = beta[i];
    }
    jacobi_host(T, k, ev, Z);
    int lo = int(std::min_element(ev.begin(), ev.end()) - ev.begin());
    theta = ev[lo];
    std::vector<double> s(k);
    for (int i = 0; i < k; ++i) s[i] = Z[size_t(i) * k + lo];
    Tensor sv = c.upload({k, 1}, false, s.data());
    y = matmul(c, V.slice_last(0, k), sv);
    scale(c, y, 1.0 / std::sqrt(norm2sq(c, y)));
    ed_apply(b, y, w);
    axpy_inplace(c, w, y, -theta);
    const double res = std::sqrt(norm2sq(c, w));
    if (res <= tol) {
      b.E0 = theta;
      b.psi0 = y;
      b.has_gs = true;
      return theta;
    }
  }
  c.warnings += 1;
  c.count("ed_lanczos[unconverged]", 1.0, 0);
  PEPS_CHECK(false, PS_ESOLVER, "ed: the ground-state Lanczos did not reach the tolerance");
  return theta;
}

EDObs ed_measure(EDBasis& b, const Tensor& psi_in) {
  PEPS_CHECK(psi_in.numel() == b.dim, PS_ESHAPE, "ed: state size");
  Ctx& c = *b.c;
  // the observables are diagonal in the basis: a complex state is measured
  // through |psi|
  Tensor psi = psi_in;
  if (psi_in.cplx) {
    psi = c.alloc({b.dim}, false);
    hipLaunchKernelGGL(ed_cabs_kernel, dim3(grid_for(b.dim)), dim3(256), 0, c.stream, b.dim, psi_in.data(),
                       psi.data());
    PEPS_HIP(hipGetLastError());
  }
  const EdGeom G = geom_of(b);
  const unsigned nb = grid_for(b.dim, 1024);
  const int ns = b.nsite;
  Tensor part = c.alloc({int64_t(nb) * 2 * ns}, false);
  hipLaunchKernelGGL(ed_measure_kernel, dim3(nb), dim3(256), 0, c.stream, G, (const uint64_t*)b.keys->ptr,
                     psi.data(), b.dim, part.data());
  PEPS_HIP(hipGetLastError());
  const std::vector<double> h = c.to_host(part);
  EDObs o;
  o.nf_grid.assign(ns, 0.0);
  o.E2_grid.assign(ns, 0.0);
  for (unsigned blk = 0; blk < nb; ++blk)
    for (int s = 0; s < ns; ++s) {
      o.nf_grid[s] += h[size_t(blk) * 2 * ns + s];
      o.E2_grid[s] += h[size_t(blk) * 2 * ns + ns + s];
    }
  double se = 0, so = 0, st = 0, e2 = 0;
  int ne = 0, no = 0;
  for (int iy = 0; iy < b.ny; ++iy)
    for (int ix = 0; ix < b.nx; ++ix) {
      const int s = iy * b.nx + ix;
      st += o.nf_grid[s];
      e2 += o.E2_grid[s];
      if ((ix + iy) % 2 == 0) {
        se += o.nf_grid[s];
        ++ne;
      } else {
        so += o.nf_grid[s];
        ++no;
      }
    }
  o.nf_mean = st / ns;
  o.nf_even = ne ? se / ne : 0.0;
  o.nf_odd = no ? so / no : 0.0;
  o.E2_mean = e2 / ns;
  return o;
}

}  // namespace peps
