// Real symmetric eigensolver for n <= 1024 on one MI355X:
//   1. Householder reduction to tridiagonal form in ONE workgroup and one
//      launch: for n <= 288 sytrd_lower_reg_kernel keeps the lower triangle
//      in registers (no global memory in the reflector loop); above that
//      sytrd_lower_kernel reads and writes the trailing matrix once per
//      reflector: each pass applies the previous step's rank-2 update and
//      forms the next matrix-vector product in the same sweep over memory,
//      and the column the next reflector needs is kept in LDS.
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

// Cross-lane sums of a double within DPP rows (VALU moves, no LDS):
// over the 4 lanes of a quad (lane bits 0, 1), and over the lanes l, l+4,
// l+8, l+12 of a 16-lane row (lane bits 2, 3).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double quad_sum(double x) {
  x += dpp_f64<0xB1>(x);  // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);  // quad_perm [2,3,0,1]
  return x;
}
__device__ __forceinline__ double row4_sum(double x) {
  x += dpp_f64<0x124>(x);  // row_ror:4
  x += dpp_f64<0x128>(x);  // row_ror:8
  return x;
}
// Sum over the 64 lanes of a wave, the same value in every lane: DPP sums
// within each 16-lane row, then the four row sums read from lanes 0, 16, 32
// and 48 and added in that order.  Every lane of the wave must be active.
__device__ __forceinline__ double wave_sum(double x) {
  x = row4_sum(quad_sum(x));
  auto lane_val = [&](int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
  };
  return ((lane_val(0) + lane_val(16)) + lane_val(32)) + lane_val(48);
}

// Register-resident variant of the same reduction for n <= 16 * NT (NT tile
// rows), one workgroup of NW waves.  The lower triangle of A lives in
// registers as 16x16 tiles (tile (I, J), I >= J, numbered column by column;
// a diagonal tile holds its full symmetric block).  Each 16-lane group owns
// one tile per slot (tile 4 (NW s + wave) + group), each lane a 4x4 block of
// it: rows 4 (l >> 2) + {0..3}, columns 4 (l & 3) + {0..3} of the tile (l
// the lane within the group).  With LDS0 the first slot -- the first 4 NW
// tiles, all in the first block columns and so the first to be finished --
// is kept in LDS instead of registers.
// Per reflector k, four barriers:
//   1. column k (already updated) and the partial norms of its rows >= k+2,
//      one per tile row, are in LDS: every thread forms the reflector
//      scalars and thread i stores v_i;
//   2. y = A v is accumulated from the tiles -- row sums over the 4
//      lanes of a quad, column sums of the off-diagonal tiles over the 4
//      lanes 4 apart, both with DPP -- and each tile writes its 16-row
//      pieces into its own slots part[I][J] (rows) and part[J][I] (columns);
//   3. thread i sums part[i >> 4][q] over the live q in a fixed order, forms
//      y_i and its share of y.v;
//   4. the rank-2 update A -= v w^T + w v^T runs in registers (w formed
//      from y and v as it is read), and the owners of column k + 1 write
//      it, with the partial norms of its rows >= k + 3, into the other
//      column buffer.
// No global memory is touched between the initial load and the stores of
// the reflectors, d, e and tau.
// 4x4 block of a tile as seen by one lane: registers, or LDS strided by the
// workgroup size (the tile slot kept in LDS).
struct RegBlk {
  double (&b)[4][4];
  __device__ double& operator()(int ri, int ci) const { return b[ri][ci]; }
};
struct LdsBlk {
  double* p;
  int stride;
  __device__ double& operator()(int ri, int ci) const { return p[(4 * ri + ci) * stride]; }
};

template <int NT, int NW, int SLOTS, bool LDS0>
__global__ __launch_bounds__(64 * NW) void sytrd_lower_reg_kernel(const double* __restrict__ A, int n, int lda,
                                                                   double* __restrict__ R, double* __restrict__ dd,
                                                                   double* __restrict__ ee, double* __restrict__ tt,
                                                                   int* __restrict__ exo) {
  constexpr int NP = 16 * NT;  // padded size
  constexpr int NTH = 64 * NW;
  constexpr int NTILES = NT * (NT + 1) / 2;
  constexpr int RS = LDS0 ? SLOTS - 1 : SLOTS;  // slots held in registers
  static_assert((NTILES + 4 * NW - 1) / (4 * NW) <= SLOTS, "tile slots");
  static_assert(NP <= NTH, "one index per thread");
  static_assert(NT <= 32 && NW <= 32, "norm and reduction buffers");
  extern __shared__ double dsm[];
  // [2][4][NP]: the 4-column blocks holding column k and column k + 1
  double (*colb)[4][NP] = reinterpret_cast<double (*)[4][NP]>(dsm);
  double* ybuf = dsm + 8 * NP;                                 // [NP]: y = tau A v
  double* vbuf = ybuf + NP;                                    // [NP]: v of reflector k
  double (*part)[NT][16] = reinterpret_cast<double (*)[NT][16]>(vbuf + NP);  // [NT][NT][16]
  double (*nbuf)[32] = reinterpret_cast<double (*)[32]>(vbuf + NP + NT * NT * 16);  // [4][32]
  double* red = vbuf + NP + NT * NT * 16 + 128;                // [32]: per-wave partial sums
  double* slot0 = red + 32;                                    // [16][NTH] when LDS0
  __shared__ unsigned char tI[NTILES], tJ[NTILES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int grp = lane >> 4, l16 = lane & 15;
  const int rb = 4 * (l16 >> 2), cb = 4 * (l16 & 3);
  if (tid == 0) {
    int t = 0;
    for (int J = 0; J < NT; ++J)
      for (int I = J; I < NT; ++I, ++t) {
        tI[t] = (unsigned char)I;
        tJ[t] = (unsigned char)J;
      }
  }
  __syncthreads();
  auto tile_of = [&](int s) { return 4 * (NW * s + wave) + grp; };
  // load: lower triangle, mirrored into diagonal tiles; padding is zero
  double a[RS > 0 ? RS : 1][4][4];
  const LdsBlk b{slot0 + tid, NTH};
  double mx = 0.0;
  auto load_tile = [&](int t, auto blk) {
#pragma unroll
    for (int ri = 0; ri < 4; ++ri)
#pragma unroll
      for (int ci = 0; ci < 4; ++ci) {
        double x = 0.0;
        if (t < NTILES) {
          const int i = 16 * tI[t] + rb + ri, j = 16 * tJ[t] + cb + ci;
          if (i < n && j < n) x = i >= j ? A[i + int64_t(j) * lda] : A[j + int64_t(i) * lda];
        }
        blk(ri, ci) = x;
        mx = fmax(mx, fabs(x));
      }
  };
#pragma unroll
  for (int s = 0; s < RS; ++s) load_tile(tile_of(s + (LDS0 ? 1 : 0)), RegBlk{a[s]});
  if (LDS0) load_tile(tile_of(0), b);
  mx = trd_block_max(mx, red);
  const int ex = mx > 0.0 ? ilogb(mx) : 0;
  const double sc = ldexp(1.0, -ex);
  if (tid == 0) *exo = ex;
#pragma unroll
  for (int s = 0; s < RS; ++s)
#pragma unroll
    for (int ri = 0; ri < 4; ++ri)
#pragma unroll
      for (int ci = 0; ci < 4; ++ci) a[s][ri][ci] *= sc;
  if (LDS0) {
#pragma unroll
    for (int e = 0; e < 16; ++e) slot0[e * NTH + tid] *= sc;
  }
  // a tile takes part in reflector k while its columns reach k + 1
  auto active = [&](int t, int k) { return t < NTILES && 16 * tJ[t] + 15 > k; };
  // owners of column c (tiles (I, c >> 4), the lanes whose 4 columns hold
  // it) write their 4-column block to dst[ci][row] and, per tile row, the
  // sums of squares of each of the 4 columns over the rows >= (its index) + 2
  // to nbuf[ci][I]; only column c & 3 of them is used.  Constant register
  // indices only (a selected column would put the tile array in scratch);
  // the whole 16-lane group runs this (the DPP sums read lanes 4 apart),
  // only the owners store.
  auto gather = [&](int t, auto blk, int c, double (*dst)[NP]) {
    if (t < NTILES && tJ[t] == (c >> 4)) {
      const int i0 = 16 * tI[t] + rb, c0 = c & ~3;
      double s[4];
#pragma unroll
      for (int ci = 0; ci < 4; ++ci) {
        s[ci] = 0.0;
#pragma unroll
        for (int ri = 0; ri < 4; ++ri)
          if (i0 + ri >= c0 + ci + 2) s[ci] += blk(ri, ci) * blk(ri, ci);
        s[ci] = row4_sum(s[ci]);
      }
      if (cb == (c & 12)) {
#pragma unroll
        for (int ri = 0; ri < 4; ++ri)
#pragma unroll
          for (int ci = 0; ci < 4; ++ci) dst[ci][i0 + ri] = blk(ri, ci);
        if (rb == 0) {
#pragma unroll
          for (int ci = 0; ci < 4; ++ci) nbuf[ci][tI[t]] = s[ci];
        }
      }
    }
  };
  // column 0
  __syncthreads();  // the load's block maximum has read red
  if (LDS0) gather(tile_of(0), b, 0, colb[0]);
#pragma unroll
  for (int s = 0; s < RS; ++s) gather(tile_of(s + (LDS0 ? 1 : 0)), RegBlk{a[s]}, 0, colb[0]);
  __syncthreads();
  for (int k = 0; k < n - 1; ++k) {
    const double* col = colb[k & 1][k & 3];
    // reflector scalars (dlarfg), computed identically by every thread
    // (fixed trip counts: the loads are issued together, the order is fixed)
    double xn2 = 0.0;
    {
      const double* nb = nbuf[k & 3];
      const int I0 = k >> 4;
#pragma unroll
      for (int I = 0; I < NT; ++I) xn2 += I >= I0 ? nb[I] : 0.0;
    }
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
    if (tid < NP) {
      const double x = col[tid];
      const double v = tid == k + 1 ? 1.0 : (tid >= k + 2 && tid < n ? x * scl : 0.0);
      vbuf[tid] = v;
      if (tid >= k + 2 && tid < n) R[tid + int64_t(k) * lda] = v;
    }
    __syncthreads();
    auto vof = [&](int j) { return vbuf[j]; };
    // y = A v into part[][][]
    auto matvec = [&](int t, auto blk) {
      const int I = tI[t], J = tJ[t];
      const int I0 = 16 * I + rb, J0 = 16 * J + cb;
      double vj[4], vi[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vj[q] = vof(J0 + q);
        vi[q] = vof(I0 + q);
      }
#pragma unroll
      for (int ri = 0; ri < 4; ++ri) {
        double y = 0.0;
#pragma unroll
        for (int ci = 0; ci < 4; ++ci) y += blk(ri, ci) * vj[ci];
        y = quad_sum(y);
        if (cb == 0) part[I][J][rb + ri] = y;
      }
      if (I != J) {
#pragma unroll
        for (int ci = 0; ci < 4; ++ci) {
          double y = 0.0;
#pragma unroll
          for (int ri = 0; ri < 4; ++ri) y += blk(ri, ci) * vi[ri];
          y = row4_sum(y);
          if (rb == 0) part[J][I][cb + ci] = y;
        }
      }
    };
    if (LDS0 && active(tile_of(0), k)) matvec(tile_of(0), b);
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      const int t = tile_of(s + (LDS0 ? 1 : 0));
      if (active(t, k)) matvec(t, RegBlk{a[s]});
      __builtin_amdgcn_sched_barrier(0);  // one tile's operands live at a time
    }
    __syncthreads();
    // thread i owns index i: y_i = tau (A v)_i, and y.v over the workgroup;
    // part[I][q] is live for q >= (k + 1) >> 4 (the tiles still active)
    double yi = 0.0, pd = 0.0;
    if (tid > k && tid < n) {
      const int I = tid >> 4, r = tid & 15, q0 = (k + 1) >> 4;
#pragma unroll
      for (int q = 0; q < NT; ++q) yi += q >= q0 ? part[I][q][r] : 0.0;
      yi *= tau;
      pd = yi * vof(tid);
    }
    if (tid < NP) ybuf[tid] = yi;
    pd = wave_sum(pd);
    if (lane == 0) red[wave] = pd;
    __syncthreads();
    double K = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) K += red[w];
    // w = y - (tau/2)(y.v) v
    const double half = -0.5 * tau * K;
    auto wof = [&](int j) { return ybuf[j] + half * vof(j); };
    auto update = [&](int t, auto blk) {
      const int I0 = 16 * tI[t] + rb, J0 = 16 * tJ[t] + cb;
      double vj[4], wj[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vj[q] = vof(J0 + q);
        wj[q] = wof(J0 + q);
      }
#pragma unroll
      for (int ri = 0; ri < 4; ++ri) {
        const double vi = vof(I0 + ri), wi = wof(I0 + ri);
#pragma unroll
        for (int ci = 0; ci < 4; ++ci) blk(ri, ci) -= vi * wj[ci] + wi * vj[ci];
      }
    };
    double (*nxt)[NP] = colb[(k + 1) & 1];
    if (LDS0 && active(tile_of(0), k)) {
      update(tile_of(0), b);
      gather(tile_of(0), b, k + 1, nxt);
    }
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      const int t = tile_of(s + (LDS0 ? 1 : 0));
      if (active(t, k)) {
        update(t, RegBlk{a[s]});
        gather(t, RegBlk{a[s]}, k + 1, nxt);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  if (tid == 0) dd[n - 1] = colb[(n - 1) & 1][(n - 1) & 3][n - 1];
}

template <int NT, int NW, bool LDS0>
constexpr size_t trd_reg_lds_bytes() {
  return sizeof(double) * (size_t(10) * 16 * NT + size_t(NT) * NT * 16 + 160 + (LDS0 ? size_t(16) * 64 * NW : 0));
}

// Z (n x ncols, leading dimension ldz) <- Q Z (trans = 0) or Q^T Z (trans =
// 1), Q = H(0) H(1) ... H(n-2).  Each half-wave owns CPH columns of Z, held
// in the registers of its 32 lanes (NR rows per lane); the dot products with
// each reflector are butterflies over those 32 lanes.  The reflectors are
// staged through LDS in chunks of ch (ch * n doubles, loaded by the whole
// workgroup), and each lane reads a reflector's entries once for all its
// columns.
template <int NR, int CPH>
__global__ __launch_bounds__(256) void apply_q_kernel(const double* __restrict__ R, int lda,
                                                      const double* __restrict__ tau, int n,
                                                      double* __restrict__ Z, int ldz, int ncols, int trans,
                                                      int ch) {
  extern __shared__ double vs[];  // ch reflectors of length n, then ch scalars
  double* ts = vs + int64_t(ch) * n;
  const int hl = threadIdx.x & 31;
  const int col0 = (blockIdx.x * (blockDim.x >> 5) + (threadIdx.x >> 5)) * CPH;
  double z[CPH][NR];
#pragma unroll
  for (int c = 0; c < CPH; ++c) {
    const bool act = col0 + c < ncols;
    const double* zc = Z + int64_t(act ? col0 + c : 0) * ldz;
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int i = hl + 32 * s;
      z[c][s] = (act && i < n) ? zc[i] : 0.0;
    }
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
      double vv[NR];
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int i = hl + 32 * s;
        vv[s] = i < n ? v[i] : 0.0;
      }
#pragma unroll
      for (int cc = 0; cc < CPH; ++cc) {
        double part = 0.0;
#pragma unroll
        for (int s = 0; s < NR; ++s) part += vv[s] * z[cc][s];
        for (int o = 16; o > 0; o >>= 1) part += __shfl_xor(part, o);
        const double f = tk * part;
#pragma unroll
        for (int s = 0; s < NR; ++s) z[cc][s] -= f * vv[s];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CPH; ++c) {
    if (col0 + c >= ncols) break;
    double* zc = Z + int64_t(col0 + c) * ldz;
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int i = hl + 32 * s;
      if (i < n) zc[i] = z[c][s];
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
  const std::string cls = c.gemm_shapes ? "eigh_trd " + std::to_string(n) : Ctx::bucket("eigh_trd", n);
  hipEvent_t ev = nullptr;
  c.tbegin(cls, &ev);
  c.count(cls, 4.0 / 3.0 * double(n) * n * n, 16.0 * double(n) * n * n / 3.0);
  int* exo = static_cast<int*>(t.ex->ptr);
  const int ni = int(n);
#define PEPS_TRD_REG(NT, NW, SLOTS, L0)                                                                       \
  do {                                                                                                      \
    const size_t lds = trd_reg_lds_bytes<NT, NW, L0>();                                                     \
    static bool attr = false;                                                                               \
    if (!attr) {                                                                                            \
      PEPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&sytrd_lower_reg_kernel<NT, NW, SLOTS, L0>), \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));                  \
      attr = true;                                                                                          \
    }                                                                                                       \
    hipLaunchKernelGGL((sytrd_lower_reg_kernel<NT, NW, SLOTS, L0>), dim3(1), dim3(64 * NW), lds, c.stream,   \
                       t.R.data(), ni, ni, t.R.data(), t.d.data(), t.e.data(), t.tau.data(), exo);          \
  } while (0)
  if (n <= 64)
    PEPS_TRD_REG(4, 8, 1, false);
  else if (n <= 96)
    PEPS_TRD_REG(6, 8, 1, false);
  else if (n <= 144)
    PEPS_TRD_REG(9, 8, 2, false);
  else if (n <= 208)
    PEPS_TRD_REG(13, 8, 3, false);
  else if (n <= 256)
    PEPS_TRD_REG(16, 12, 3, true);
  else if (n <= 288)
    PEPS_TRD_REG(18, 11, 4, true);
  else if (c.trd_backend == 1) {
#undef PEPS_TRD_REG
    // rocSOLVER sytrd (same LAPACK reflector layout) on the matrix scaled by
    // the power of two that brings its Frobenius norm into [1, 2)
    const double fro = std::sqrt(norm2sq(c, t.R));
    const int ex = fro > 0.0 ? std::ilogb(fro) : 0;
    if (ex != 0) scale(c, t.R, std::ldexp(1.0, -ex));
    PEPS_HIP(hipMemcpyAsync(exo, &ex, sizeof(int), hipMemcpyHostToDevice, c.stream));
    PEPS_HIP(hipStreamSynchronize(c.stream));  // ex lives on this stack frame
    ck_solver(rocsolver_dsytrd(c.blas, rocblas_fill_lower, ni, t.R.data(), ni, t.d.data(), t.e.data(),
                               t.tau.data()),
              "dsytrd");
  } else {
    const size_t lds = size_t(4 * n + kTrdThreads + 64) * sizeof(double);
    hipLaunchKernelGGL(sytrd_lower_kernel, dim3(1), dim3(kTrdThreads), lds, c.stream, t.R.data(), ni, ni,
                       t.d.data(), t.e.data(), t.tau.data(), exo);
  }
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
  const int tr = trans ? 1 : 0;
  // reflectors per LDS chunk: ch (n + 1) doubles within kQChunkBytes
  const int ch = int(std::max<int64_t>(1, std::min<int64_t>(n - 1, kQChunkBytes / (8 * (n + 1)))));
  const size_t lds = size_t(ch) * size_t(n + 1) * sizeof(double);
  // one column per half-wave for a few right-hand sides, four for many
  const int cph = k <= 32 ? 1 : 4;
  const int64_t halfwaves = (k + cph - 1) / cph;
  const int hw_per_block = int(std::min<int64_t>(8, halfwaves + (halfwaves & 1)));
  const dim3 grid(unsigned((halfwaves + hw_per_block - 1) / hw_per_block)), block(unsigned(32 * hw_per_block));
#define PEPS_APPLY_Q(NR, CPH)                                                                                   \
  hipLaunchKernelGGL((apply_q_kernel<NR, CPH>), grid, block, lds, c.stream, t.R.data(), int(n), t.tau.data(), \
                     int(n), Z.data(), int(n), int(k), tr, ch)
  if (cph == 1) {
    if (n <= 256) PEPS_APPLY_Q(8, 1);
    else if (n <= 512) PEPS_APPLY_Q(16, 1);
    else PEPS_APPLY_Q(32, 1);
  } else {
    if (n <= 256) PEPS_APPLY_Q(8, 4);
    else if (n <= 512) PEPS_APPLY_Q(16, 4);
    else PEPS_APPLY_Q(32, 4);
  }
#undef PEPS_APPLY_Q
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
  const std::string cls = c.gemm_shapes ? "eigh_stedc " + std::to_string(n) : Ctx::bucket("eigh_stedc", n);
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
