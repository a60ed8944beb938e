// HIP kernels of the finite-PEPS engine, written for CDNA4 (gfx950).
//
//   gemm      LDS-tiled MFMA GEMM.  Storage is FP64 (real or interleaved
//             complex); the compute tier picks the MFMA: f64 16x16x4,
//             f32 16x16x4 or bf16 16x16x16 (FP32 accumulation).  Tiles are
//             converted while being staged into LDS.  Split-K with a
//             deterministic (fixed-order) reduction fills the chip when the
//             output is small and K is huge (environment closures).
//   permute   HBM-streaming N-d transpose (coalesced writes).
//   the rest  elementwise weights/masks and wave reductions.
#include "peps_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace peps {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

static inline unsigned cdiv(int64_t a, int64_t b) { return unsigned((a + b - 1) / b); }

// ════════════════════════════════════════════════════════════════════════════
//  GEMM
// ════════════════════════════════════════════════════════════════════════════

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}

template <int PREC> struct LdsT;
template <> struct LdsT<0> { typedef double T; };
template <> struct LdsT<1> { typedef float T; };
template <> struct LdsT<2> { typedef unsigned short T; };

template <int PREC> struct AccT;
template <> struct AccT<0> { typedef v4d T; };
template <> struct AccT<1> { typedef v4f T; };
template <> struct AccT<2> { typedef v4f T; };

template <int PREC>
__device__ __forceinline__ typename LdsT<PREC>::T cvt(double x) {
  if constexpr (PREC == 0) return x;
  else if constexpr (PREC == 1) return float(x);
  else return f2bf(float(x));
}

// Block of 256 threads = 4 waves in a 2x2 arrangement; each wave owns
// (16*WM) x (16*WN) outputs as WM x WN MFMA 16x16 tiles.  BK = 16.
//
// MFMA 16x16 operand layout (lane l): A row = l%16, k-group = l/16;
// B col = l%16, k-group = l/16; output row = 4*(l/16)+v, col = l%16.
template <bool CPLX, int PREC, bool AK, bool BKF, int WM, int WN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_kernel(int64_t M, int64_t N, int64_t K,
                                                   const double* __restrict__ A, int64_t lda,
                                                   const double* __restrict__ B, int64_t ldb,
                                                   double* __restrict__ C, int64_t ldc, int conjA,
                                                   int conjB, double beta, int64_t kchunk,
                                                   int64_t wstride) {
  constexpr int BM = 32 * WM, BN = 32 * WN, BK = 16;
  constexpr int NC = CPLX ? 2 : 1;
  typedef typename LdsT<PREC>::T LT;
  typedef typename AccT<PREC>::T ACC;
  // PREC 0/1 store tiles k-major ([k][m]); bf16 stores m-major ([m][k]) so a
  // lane's 4 consecutive k are one 8-byte read.
  constexpr int APAD = (PREC == 2) ? (BK + 4) : (BM + 1);
  constexpr int BPAD = (PREC == 2) ? (BK + 4) : (BN + 1);
  constexpr int ASZ = (PREC == 2) ? BM * APAD : BK * APAD;
  constexpr int BSZ = (PREC == 2) ? BN * BPAD : BK * BPAD;
  __shared__ __align__(16) LT As[NC][ASZ];
  __shared__ __align__(16) LT Bs[NC][BSZ];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * (16 * WM);
  const int wn = (wave >> 1) * (16 * WN);
  // tile of this block: the linear block index is remapped so that the
  // blocks the hardware deals to one XCD (round robin over 8) are
  // consecutive, then tiles are visited in groups of 8 tile rows: the blocks
  // resident on one XCD share a few A and B panels, whose current k-slices
  // stay in that XCD's L2 while the blocks advance through K together
  int64_t m0, n0;
  {
    const int tilesM = int((M + BM - 1) / BM), tilesN = int((N + BN - 1) / BN);
    const int nt = tilesM * tilesN;
    int w = blockIdx.x;
    const int q = nt >> 3, r = nt & 7, x = w & 7, i = w >> 3;
    w = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
    constexpr int GM = 8;
    const int per_group = GM * tilesN;
    const int first = (w / per_group) * GM;
    const int gs = min(tilesM - first, GM);
    m0 = int64_t(first + (w % per_group) % gs) * BM;
    n0 = int64_t((w % per_group) / gs) * BN;
  }
  const int64_t kbeg = int64_t(blockIdx.z) * kchunk;
  const int64_t kend = min(K, kbeg + kchunk);
  double* Cz = C + int64_t(blockIdx.z) * wstride;

  ACC acc[NC][WM][WN];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j) acc[c][i][j] = ACC{0, 0, 0, 0};

  constexpr int AE = BM * BK / 256;  // elements per thread per tile
  constexpr int BE = BN * BK / 256;
  double ra[AE][NC], rb[BE][NC];

  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      int idx = tid + e * 256;
      int kk, mm;
      if (AK) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
      int64_t gm = m0 + mm, gk = k0 + kk;
      bool ok = gm < M && gk < kend;
      int64_t off = AK ? (gk + gm * lda) : (gm + gk * lda);
      if (CPLX) {
        ra[e][0] = ok ? A[2 * off] : 0.0;
        ra[e][NC - 1] = ok ? A[2 * off + 1] : 0.0;
      } else {
        ra[e][0] = ok ? A[off] : 0.0;
      }
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      int idx = tid + e * 256;
      int kk, nn;
      if (BKF) { kk = idx % BK; nn = idx / BK; } else { nn = idx % BN; kk = idx / BN; }
      int64_t gn = n0 + nn, gk = k0 + kk;
      bool ok = gn < N && gk < kend;
      int64_t off = BKF ? (gk + gn * ldb) : (gn + gk * ldb);
      if (CPLX) {
        rb[e][0] = ok ? B[2 * off] : 0.0;
        rb[e][NC - 1] = ok ? B[2 * off + 1] : 0.0;
      } else {
        rb[e][0] = ok ? B[off] : 0.0;
      }
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      int idx = tid + e * 256;
      int kk, mm;
      if (AK) { kk = idx % BK; mm = idx / BK; } else { mm = idx % BM; kk = idx / BM; }
      int li = (PREC == 2) ? (mm * APAD + kk) : (kk * APAD + mm);
      As[0][li] = cvt<PREC>(ra[e][0]);
      if (CPLX) As[NC - 1][li] = cvt<PREC>(conjA ? -ra[e][NC - 1] : ra[e][NC - 1]);
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      int idx = tid + e * 256;
      int kk, nn;
      if (BKF) { kk = idx % BK; nn = idx / BK; } else { nn = idx % BN; kk = idx / BN; }
      int li = (PREC == 2) ? (nn * BPAD + kk) : (kk * BPAD + nn);
      Bs[0][li] = cvt<PREC>(rb[e][0]);
      if (CPLX) Bs[NC - 1][li] = cvt<PREC>(conjB ? -rb[e][NC - 1] : rb[e][NC - 1]);
    }
  };

  const int lr = lane & 15;
  const int lg = lane >> 4;

  if (kbeg < kend) {
    gload(kbeg);
    lstore();
  }
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = (k0 + BK) < kend;
    if (more) gload(k0 + BK);
    if constexpr (PREC == 2) {
      v4s a[NC][WM], b[NC][WN];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int i = 0; i < WM; ++i)
          a[c][i] = *reinterpret_cast<const v4s*>(&As[c][(wm + i * 16 + lr) * APAD + 4 * lg]);
#pragma unroll
        for (int j = 0; j < WN; ++j)
          b[c][j] = *reinterpret_cast<const v4s*>(&Bs[c][(wn + j * 16 + lr) * BPAD + 4 * lg]);
      }
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) {
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][i], b[0][j], acc[0][i][j], 0, 0, 0);
          if constexpr (CPLX) {
            // bf16 negation flips the sign bit
            const short sb = short(0x8000);
            v4s nai = a[1][i] ^ v4s{sb, sb, sb, sb};
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nai, b[1][j], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][i], b[1][j], acc[1][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1][i], b[0][j], acc[1][i][j], 0, 0, 0);
          }
        }
    } else {
#pragma unroll
      for (int kq = 0; kq < BK / 4; ++kq) {
        const int kr = kq * 4 + lg;
        LT a[NC][WM], b[NC][WN];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
          for (int i = 0; i < WM; ++i) a[c][i] = As[c][kr * APAD + wm + i * 16 + lr];
#pragma unroll
          for (int j = 0; j < WN; ++j) b[c][j] = Bs[c][kr * BPAD + wn + j * 16 + lr];
        }
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
          for (int j = 0; j < WN; ++j) {
            if constexpr (PREC == 0) {
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i], b[0][j], acc[0][i][j], 0, 0, 0);
              if constexpr (CPLX) {
                acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[NC - 1][i], b[NC - 1][j], acc[0][i][j], 0, 0, 0);
                acc[NC - 1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i], b[NC - 1][j], acc[NC - 1][i][j], 0, 0, 0);
                acc[NC - 1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[NC - 1][i], b[0][j], acc[NC - 1][i][j], 0, 0, 0);
              }
            } else {
              acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i], b[0][j], acc[0][i][j], 0, 0, 0);
              if constexpr (CPLX) {
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[NC - 1][i], b[NC - 1][j], acc[0][i][j], 0, 0, 0);
                acc[NC - 1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i], b[NC - 1][j], acc[NC - 1][i][j], 0, 0, 0);
                acc[NC - 1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[NC - 1][i], b[0][j], acc[NC - 1][i][j], 0, 0, 0);
              }
            }
          }
      }
    }
    __syncthreads();
    if (more) {
      lstore();
      __syncthreads();
    }
  }

  // epilogue: C = acc + beta * C
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      int64_t col = n0 + wn + j * 16 + lr;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // f64 16x16x4 keeps rows 4v + l/16 per lane (group size 1); the
        // f32/bf16 instructions keep rows 4(l/16) + v (group size 4)
        int64_t row = m0 + wm + i * 16 + (PREC == 0 ? (lg + 4 * v) : (4 * lg + v));
        if (row >= M) continue;
        int64_t off = row + col * ldc;
        if (CPLX) {
          double re = double(acc[0][i][j][v]), im = double(acc[NC - 1][i][j][v]);
          if (beta != 0.0) {
            re += beta * Cz[2 * off];
            im += beta * Cz[2 * off + 1];
          }
          Cz[2 * off] = re;
          Cz[2 * off + 1] = im;
        } else {
          double x = double(acc[0][i][j][v]);
          if (beta != 0.0) x += beta * Cz[off];
          Cz[off] = x;
        }
      }
    }
}

__global__ void splitk_reduce(double* __restrict__ C, const double* __restrict__ W, int64_t n,
                              int splits, int64_t wstride, double beta) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += W[z * wstride + i];
    C[i] = (beta != 0.0 ? beta * C[i] : 0.0) + s;
  }
}

// Real GEMM of the FP32 tier for large outputs (M, N >= 128).
//
// 128x128 output tile per 256-thread block (2x2 waves of 64x64, 4x4 MFMA
// f32 16x16x4 each), k-tiles of 32 staged FP64 -> FP32 through two LDS
// buffers (one barrier per k-tile; the next tile's global loads are in
// flight during the MFMAs).  LDS rows hold 32 consecutive k of one m (or n):
// lane group g supplies k = 8g + s in MFMA step s, so a lane reads its eight
// operands of a row as two float4.  Tiles are visited in groups of 8 tile
// rows (L2 reuse of A and B panels) and the block index is remapped so that
// the blocks the hardware sends to one XCD (round robin over 8) are
// consecutive in that order.  The epilogue transposes the accumulators
// through LDS so that C is read and written in whole column segments.
// TIN is the operand storage type: double (cast to FP32 on the way into LDS)
// or float (operands cast once by cast_f32_kernel, half the L2 -> CU bytes
// per k-tile); both give the same bits, since the cast is the same
// round-to-nearest conversion and the accumulation order does not change.
template <typename TIN, bool AK, bool BKF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_f32_tiled(int64_t M, int64_t N, int64_t K,
                                                      const TIN* __restrict__ A, int64_t lda,
                                                      const TIN* __restrict__ B, int64_t ldb,
                                                      double* __restrict__ C, int64_t ldc, double beta,
                                                      int64_t kchunk, int64_t wstride, int tilesM,
                                                      int tilesN) {
  constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 4, LDC = BM + 1;
  constexpr int BUF = (BM + BN) * LDK;
  static_assert(BN * LDC <= 2 * BUF, "epilogue staging must fit the operand buffers");
  __shared__ __align__(16) float smem[2 * BUF];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;

  // tile of this block: XCD-contiguous remap, then groups of 8 tile rows
  const int nt = tilesM * tilesN;
  int w = blockIdx.x;
  {
    const int q = nt >> 3, r = nt & 7, x = w & 7, i = w >> 3;
    w = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
  }
  constexpr int GM = 8;
  const int per_group = GM * tilesN;
  const int first = (w / per_group) * GM;
  const int gs = min(tilesM - first, GM);
  const int tm = first + (w % per_group) % gs;
  const int tn = (w % per_group) / gs;
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;

  const int64_t kbeg = int64_t(blockIdx.z) * kchunk;
  const int64_t kend = min(K, kbeg + kchunk);
  double* Cz = C + int64_t(blockIdx.z) * wstride;

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // Each thread moves 4 groups of 4 consecutive k of one row (m or n) per
  // operand and k-tile: four coalesced 8-byte loads per group (lanes run
  // along the contiguous global dim) and one 16-byte LDS store.
  TIN ra[4][4], rb[4][4];
  // group e of this thread in a 128 x 32 operand tile: (row, first k)
  auto arow = [&](int e) { return AK ? (tid >> 3) + 32 * e : (tid & 127); };
  auto acol = [&](int e) { return AK ? 4 * (tid & 7) : 4 * ((tid >> 7) + 2 * e); };
  auto brow = [&](int e) { return BKF ? (tid >> 3) + 32 * e : (tid & 127); };
  auto bcol = [&](int e) { return BKF ? 4 * (tid & 7) : 4 * ((tid >> 7) + 2 * e); };
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t gm = m0 + arow(e), gn = n0 + brow(e);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t gk = k0 + acol(e) + j, gkb = k0 + bcol(e) + j;
        ra[e][j] = (gm < M && gk < kend) ? A[AK ? gk + gm * lda : gm + gk * lda] : TIN(0);
        rb[e][j] = (gn < N && gkb < kend) ? B[BKF ? gkb + gn * ldb : gn + gkb * ldb] : TIN(0);
      }
    }
  };
  // Interior tiles (all 128 rows of A and B in range) and full k-tiles take
  // a branch-free path: the tile base is uniform, each thread adds one
  // constant lane offset, and the (e, j) steps are uniform multiples of the
  // leading dimension (lane group tid >> 7 is the same across a wave).
  const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
  const int kgrp = __builtin_amdgcn_readfirstlane(tid >> 7);
  const int64_t laneA = AK ? int64_t(tid >> 3) * lda + 4 * (tid & 7) : int64_t(tid & 127);
  const int64_t laneB = BKF ? int64_t(tid >> 3) * ldb + 4 * (tid & 7) : int64_t(tid & 127);
  auto gload_fast = [&](int64_t k0) {
    const TIN* Ab = A + (AK ? m0 * lda + k0 : m0 + k0 * lda) + laneA;
    const TIN* Bb = B + (BKF ? n0 * ldb + k0 : n0 + k0 * ldb) + laneB;
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ra[e][j] = Ab[AK ? int64_t(32 * e) * lda + j : int64_t(4 * (kgrp + 2 * e) + j) * lda];
        rb[e][j] = Bb[BKF ? int64_t(32 * e) * ldb + j : int64_t(4 * (kgrp + 2 * e) + j) * ldb];
      }
  };
  auto load = [&](int64_t k0) {
    if (interior && k0 + BK <= kend) gload_fast(k0);
    else gload(k0);
  };
  auto lstore = [&](int buf) {
    float* As = smem + buf * BUF;
    float* Bs = As + BM * LDK;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      *reinterpret_cast<v4f*>(As + arow(e) * LDK + acol(e)) =
          v4f{float(ra[e][0]), float(ra[e][1]), float(ra[e][2]), float(ra[e][3])};
      *reinterpret_cast<v4f*>(Bs + brow(e) * LDK + bcol(e)) =
          v4f{float(rb[e][0]), float(rb[e][1]), float(rb[e][2]), float(rb[e][3])};
    }
  };

  const int64_t ntk = (kend - kbeg + BK - 1) / BK;
  if (ntk > 0) {
    load(kbeg);
    lstore(0);
  }
  __syncthreads();
  for (int64_t t = 0; t < ntk; ++t) {
    const bool more = t + 1 < ntk;
    if (more) load(kbeg + (t + 1) * BK);
    const float* As = smem + (t & 1) * BUF;
    const float* Bs = As + BM * LDK;
    v4f a[4][2], b[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float* pa = As + (wm + i * 16 + lr) * LDK + 8 * lg;
      a[i][0] = *reinterpret_cast<const v4f*>(pa);
      a[i][1] = *reinterpret_cast<const v4f*>(pa + 4);
      const float* pb = Bs + (wn + i * 16 + lr) * LDK + 8 * lg;
      b[i][0] = *reinterpret_cast<const v4f*>(pb);
      b[i][1] = *reinterpret_cast<const v4f*>(pb + 4);
    }
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][s >> 2][s & 3], b[j][s >> 2][s & 3], acc[i][j],
                                                             0, 0, 0);
    if (more) lstore((t + 1) & 1);
    __syncthreads();
  }

  // epilogue: accumulators -> LDS (column-major 128 x 128, padded) -> C
  float* Cs = smem;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) Cs[(wn + j * 16 + lr) * LDC + wm + i * 16 + 4 * lg + v] = acc[i][j][v];
  __syncthreads();
  for (int idx = tid; idx < BM * BN; idx += 256) {
    const int mm = idx & (BM - 1), nn = idx / BM;
    const int64_t gm = m0 + mm, gn = n0 + nn;
    if (gm >= M || gn >= N) continue;
    const int64_t off = gm + gn * ldc;
    double x = double(Cs[nn * LDC + mm]);
    if (beta != 0.0) x += beta * Cz[off];
    Cz[off] = x;
  }
}

// Real GEMM of the FP64 tier for outputs with M, N >= 128.
//
// 128x128 output tile per 256-thread block (2x2 waves of 64x64, 4x4
// v_mfma_f64_16x16x4f64 each), k-tiles of 16 doubles through two LDS buffers
// (one barrier per k-tile; the next tile's global loads are in flight during
// the 64 MFMAs of the current one).  LDS rows hold the 16 consecutive k of
// one m (or n), padded to 18 doubles (144 bytes, the FP32 kernel's row
// geometry): lane group g supplies k = 4g + s in MFMA step s, so a lane
// reads the four operands of one row for the whole k-tile as two 16-byte
// LDS reads.  Operands are loaded four consecutive k at a time per row
// (coalesced along the contiguous global dim), stored as two 16-byte LDS
// writes; interior tiles and full k-tiles load without per-element guards.
// Tiles are visited in groups of 8 tile rows with the XCD-contiguous block
// remap of the FP32 kernel.  The accumulators are written straight to C
// (each lane holds rows 4v + lane/16 of column lane%16 of a 16x16 tile).
template <bool AK, bool BKF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_f64_tiled(int64_t M, int64_t N, int64_t K,
                                                      const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t ldb,
                                                      double* __restrict__ C, int64_t ldc, double beta,
                                                      int64_t kchunk, int64_t wstride, int tilesM,
                                                      int tilesN) {
  constexpr int BM = 128, BN = 128, BK = 16, LDK = BK + 2;
  constexpr int BUF = (BM + BN) * LDK;
  __shared__ __align__(16) double smem[2 * BUF];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;

  const int nt = tilesM * tilesN;
  int w = blockIdx.x;
  {
    const int q = nt >> 3, r = nt & 7, x = w & 7, i = w >> 3;
    w = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
  }
  constexpr int GM = 8;
  const int per_group = GM * tilesN;
  const int first = (w / per_group) * GM;
  const int gs = min(tilesM - first, GM);
  const int64_t m0 = int64_t(first + (w % per_group) % gs) * BM;
  const int64_t n0 = int64_t((w % per_group) / gs) * BN;

  const int64_t kbeg = int64_t(blockIdx.z) * kchunk;
  const int64_t kend = min(K, kbeg + kchunk);
  double* Cz = C + int64_t(blockIdx.z) * wstride;

  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  // two groups of 4 consecutive k of one row per operand and thread
  double ra[2][4], rb[2][4];
  auto arow = [&](int e) { return AK ? (tid >> 2) + 64 * e : (tid & 127); };
  auto acol = [&](int e) { return AK ? 4 * (tid & 3) : 4 * ((tid >> 7) + 2 * e); };
  auto brow = [&](int e) { return BKF ? (tid >> 2) + 64 * e : (tid & 127); };
  auto bcol = [&](int e) { return BKF ? 4 * (tid & 3) : 4 * ((tid >> 7) + 2 * e); };
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int64_t gm = m0 + arow(e), gn = n0 + brow(e);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t gk = k0 + acol(e) + j, gkb = k0 + bcol(e) + j;
        ra[e][j] = (gm < M && gk < kend) ? A[AK ? gk + gm * lda : gm + gk * lda] : 0.0;
        rb[e][j] = (gn < N && gkb < kend) ? B[BKF ? gkb + gn * ldb : gn + gkb * ldb] : 0.0;
      }
    }
  };
  const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
  const int kgrp = __builtin_amdgcn_readfirstlane(tid >> 7);
  const int64_t laneA = AK ? int64_t(tid >> 2) * lda + 4 * (tid & 3) : int64_t(tid & 127);
  const int64_t laneB = BKF ? int64_t(tid >> 2) * ldb + 4 * (tid & 3) : int64_t(tid & 127);
  auto gload_fast = [&](int64_t k0) {
    const double* Ab = A + (AK ? m0 * lda + k0 : m0 + k0 * lda) + laneA;
    const double* Bb = B + (BKF ? n0 * ldb + k0 : n0 + k0 * ldb) + laneB;
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ra[e][j] = Ab[AK ? int64_t(64 * e) * lda + j : int64_t(4 * (kgrp + 2 * e) + j) * lda];
        rb[e][j] = Bb[BKF ? int64_t(64 * e) * ldb + j : int64_t(4 * (kgrp + 2 * e) + j) * ldb];
      }
  };
  auto load = [&](int64_t k0) {
    if (interior && k0 + BK <= kend) gload_fast(k0);
    else gload(k0);
  };
  auto lstore = [&](int buf) {
    double* As = smem + buf * BUF;
    double* Bs = As + BM * LDK;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      double* pa = As + arow(e) * LDK + acol(e);
      double* pb = Bs + brow(e) * LDK + bcol(e);
      *reinterpret_cast<double2*>(pa) = double2{ra[e][0], ra[e][1]};
      *reinterpret_cast<double2*>(pa + 2) = double2{ra[e][2], ra[e][3]};
      *reinterpret_cast<double2*>(pb) = double2{rb[e][0], rb[e][1]};
      *reinterpret_cast<double2*>(pb + 2) = double2{rb[e][2], rb[e][3]};
    }
  };

  const int64_t ntk = (kend - kbeg + BK - 1) / BK;
  if (ntk > 0) {
    load(kbeg);
    lstore(0);
  }
  __syncthreads();
  for (int64_t t = 0; t < ntk; ++t) {
    const bool more = t + 1 < ntk;
    if (more) load(kbeg + (t + 1) * BK);
    const double* As = smem + (t & 1) * BUF;
    const double* Bs = As + BM * LDK;
    // two halves of the k-tile (steps s = 2h, 2h + 1): one 16-byte LDS read
    // per operand row and half
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double2 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = *reinterpret_cast<const double2*>(As + (wm + i * 16 + lr) * LDK + 4 * lg + 2 * h);
        b[i] = *reinterpret_cast<const double2*>(Bs + (wn + i * 16 + lr) * LDK + 4 * lg + 2 * h);
      }
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(s ? a[i].y : a[i].x, s ? b[j].y : b[j].x, acc[i][j],
                                                             0, 0, 0);
    }
    if (more) lstore((t + 1) & 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = n0 + wn + j * 16 + lr;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = m0 + wm + i * 16 + lg + 4 * v;
        if (row >= M) continue;
        const int64_t off = row + col * ldc;
        double x = acc[i][j][v];
        if (beta != 0.0) x += beta * Cz[off];
        Cz[off] = x;
      }
    }
}

static void launch_gemm_f64_tiled(bool ak, bool bk, hipStream_t st, int64_t M, int64_t N, int64_t K,
                                  const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                                  int64_t ldc, double beta, int splits, int64_t kchunk, int64_t wstride) {
  const int tm = int(cdiv(M, 128)), tn = int(cdiv(N, 128));
  dim3 grid(unsigned(tm) * unsigned(tn), 1, splits);
  if (ak && bk)
    hipLaunchKernelGGL((gemm_f64_tiled<true, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
  else if (ak)
    hipLaunchKernelGGL((gemm_f64_tiled<true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
  else if (bk)
    hipLaunchKernelGGL((gemm_f64_tiled<false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
  else
    hipLaunchKernelGGL((gemm_f64_tiled<false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
}

template <typename TIN>
static void launch_gemm_f32_tiled(bool ak, bool bk, hipStream_t st, int64_t M, int64_t N, int64_t K,
                                  const TIN* A, int64_t lda, const TIN* B, int64_t ldb, double* C,
                                  int64_t ldc, double beta, int splits, int64_t kchunk, int64_t wstride) {
  const int tm = int(cdiv(M, 128)), tn = int(cdiv(N, 128));
  dim3 grid(unsigned(tm) * unsigned(tn), 1, splits);
  if (ak && bk)
    hipLaunchKernelGGL((gemm_f32_tiled<TIN, true, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
  else if (ak)
    hipLaunchKernelGGL((gemm_f32_tiled<TIN, true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
  else if (bk)
    hipLaunchKernelGGL((gemm_f32_tiled<TIN, false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
  else
    hipLaunchKernelGGL((gemm_f32_tiled<TIN, false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk, wstride, tm, tn);
}

// FP64 -> FP32 (round to nearest) copy of a packed operand.
__global__ void cast_f32_kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = float(x[i]);
}

// Operands of a tiled GEMM with both extents >= this are cast to FP32 once:
// each operand element is then read by >= 4 tiles, so the cast (12 bytes of
// HBM traffic per element) costs less than the FP64 bytes it saves.
constexpr int64_t kCastMin = 512;

template <bool CPLX, int PREC, bool AK, bool BKF, int WM, int WN>
static void launch_gemm_t(hipStream_t st, int64_t M, int64_t N, int64_t K, const double* A,
                          int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                          int cA, int cB, double beta, int splits, int64_t kchunk,
                          int64_t wstride) {
  // one linear block index per output tile (the kernel maps it to a tile)
  dim3 grid(unsigned(int64_t(cdiv(M, 32 * WM)) * cdiv(N, 32 * WN)), 1, splits);
  hipLaunchKernelGGL((gemm_kernel<CPLX, PREC, AK, BKF, WM, WN>), grid, dim3(256), 0, st, M, N, K,
                     A, lda, B, ldb, C, ldc, cA, cB, beta, kchunk, wstride);
}

template <bool CPLX, int PREC, int WM, int WN>
static void dispatch_layout(bool ak, bool bk, hipStream_t st, int64_t M, int64_t N, int64_t K,
                            const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                            int64_t ldc, int cA, int cB, double beta, int splits, int64_t kchunk,
                            int64_t wstride) {
  if (ak && bk)
    launch_gemm_t<CPLX, PREC, true, true, WM, WN>(st, M, N, K, A, lda, B, ldb, C, ldc, cA, cB, beta, splits, kchunk, wstride);
  else if (ak && !bk)
    launch_gemm_t<CPLX, PREC, true, false, WM, WN>(st, M, N, K, A, lda, B, ldb, C, ldc, cA, cB, beta, splits, kchunk, wstride);
  else if (!ak && bk)
    launch_gemm_t<CPLX, PREC, false, true, WM, WN>(st, M, N, K, A, lda, B, ldb, C, ldc, cA, cB, beta, splits, kchunk, wstride);
  else
    launch_gemm_t<CPLX, PREC, false, false, WM, WN>(st, M, N, K, A, lda, B, ldb, C, ldc, cA, cB, beta, splits, kchunk, wstride);
}

template <bool CPLX, int PREC>
static void dispatch_tile(bool big, bool ak, bool bk, hipStream_t st, int64_t M, int64_t N,
                          int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb,
                          double* C, int64_t ldc, int cA, int cB, double beta, int splits,
                          int64_t kchunk, int64_t wstride) {
  // complex outputs always take 64x64 tiles (a 128x128 complex tile does
  // not fit the registers of two waves per SIMD)
  if constexpr (!CPLX) {
    if (big) {
      dispatch_layout<CPLX, PREC, 4, 4>(ak, bk, st, M, N, K, A, lda, B, ldb, C, ldc, cA, cB, beta, splits, kchunk,
                                        wstride);
      return;
    }
  }
  dispatch_layout<CPLX, PREC, 2, 2>(ak, bk, st, M, N, K, A, lda, B, ldb, C, ldc, cA, cB, beta, splits, kchunk, wstride);
}

__global__ void zero_kernel(double* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] = 0.0;
}

namespace {
// Where the split-K partial products of one GEMM go: ws_fixed == nullptr
// takes them from the context's pool (at most ws_cap bytes), otherwise they
// are written to ws_fixed, which holds ws_cap bytes.  cast allows the FP32
// tier's one-time operand cast (two pool buffers); timed brackets the
// kernels with the context's timing events.
struct GemmWorkspace {
  double* ws_fixed = nullptr;
  double ws_cap = double(1ull << 30);
  bool cast = true;
  bool timed = true;
};

void gemm_impl(Ctx& c, bool cplx, int64_t M, int64_t N, int64_t K, const double* A, bool a_kfast,
               bool conjA, const double* B, bool b_kfast, bool conjB, double* C, double beta,
               const GemmWorkspace& wsp) {
  if (M <= 0 || N <= 0) return;
  const int64_t lda = a_kfast ? K : M;
  const int64_t ldb = b_kfast ? K : N;
  const int esz = cplx ? 2 : 1;
  if (K <= 0) {
    if (beta == 0.0) {
      if (wsp.ws_fixed) {  // a kernel, not a memset node, for callers under graph capture
        const int64_t n = M * N * esz;
        hipLaunchKernelGGL(zero_kernel, dim3(unsigned(std::min<int64_t>(cdiv(n, 256), 8192))), dim3(256), 0,
                           c.stream, C, n);
        PEPS_HIP(hipGetLastError());
      } else {
        PEPS_HIP(hipMemsetAsync(C, 0, size_t(M * N) * esz * 8, c.stream));
      }
    } else if (beta != 1.0) {
      Tensor t;  // scale existing C
      t.buf = std::make_shared<Buffer>();
      t.buf->ptr = C;
      t.dims = {M, N};
      t.cplx = cplx;
      scale(c, t, beta);
      t.buf->ptr = nullptr;
    }
    return;
  }
  const bool big = !cplx && M >= 128 && N >= 128;
  const int BMt = big ? 128 : 64, BNt = big ? 128 : 64;
  const int64_t tiles = int64_t(cdiv(M, BMt)) * cdiv(N, BNt);
  int splits = 1;
  const int64_t target = int64_t(c.ncu) * 2;
  if (tiles < target && K >= 1024) {
    splits = int(std::min<int64_t>((target + tiles - 1) / tiles, K / 256));
    splits = std::max(1, std::min(splits, 512));
    // bound the workspace
    while (splits > 1 && double(splits) * M * N * esz * 8 > wsp.ws_cap) splits /= 2;
  }
  int64_t kchunk = K;
  if (splits > 1) {
    kchunk = ((K + splits - 1) / splits + 15) / 16 * 16;
    splits = int((K + kchunk - 1) / kchunk);
  }
  const double flops = (cplx ? 8.0 : 2.0) * double(M) * double(N) * double(K);
  const double bytes = double(esz * 8) * (double(M) * K + double(K) * N + double(M) * N);
  const bool tiled = !cplx && c.tier == TIER_FP32 && M >= 128 && N >= 128 &&
                     int64_t(cdiv(M, 128)) * cdiv(N, 128) < (int64_t(1) << 31);
  const bool tiled64 = !cplx && c.tier == TIER_FP64 && M >= 128 && N >= 128 &&
                       int64_t(cdiv(M, 128)) * cdiv(N, 128) < (int64_t(1) << 31);
  std::string cls;
  const bool cast = wsp.cast && tiled && M >= kCastMin && N >= kCastMin;
  if (wsp.timed) {
    cls = tiled ? "gemm_f32_tiled"
                : tiled64 ? "gemm_f64_tiled"
                          : (c.tier == TIER_FP64 ? "gemm_f64" : (c.tier == TIER_FP32 ? "gemm_f32" : "gemm_bf16"));
    if (c.gemm_shapes)
      cls += " " + std::to_string(M) + "x" + std::to_string(N) + "x" + std::to_string(K) + (cplx ? "c" : "") +
             (a_kfast ? " ak" : " am") + (b_kfast ? " bk" : " bn") + " s" + std::to_string(splits);
    // algorithmic bytes of the GEMM kernel itself: FP32 operands when cast
    c.count(cls, flops, cast ? 4.0 * (double(M) * K + double(K) * N) + 8.0 * double(M) * N : bytes);
  }
  std::shared_ptr<Buffer> a32, b32;
  if (cast) {
    // A and B are packed (lda = M or K, ldb = K or N): cast M*K and K*N
    // contiguous elements; the cast kernel is timed as its own class
    a32 = std::make_shared<Buffer>();
    b32 = std::make_shared<Buffer>();
    a32->pool = b32->pool = c.pool.get();
    a32->bytes = size_t(M * K) * sizeof(float);
    b32->bytes = size_t(K * N) * sizeof(float);
    a32->ptr = c.pool->get(a32->bytes);
    b32->ptr = c.pool->get(b32->bytes);
    hipEvent_t evc = nullptr;
    c.tbegin("cast_f32", &evc);
    c.count("cast_f32", 0.0, 12.0 * double(M * K + K * N));
    const unsigned gc = unsigned(std::min<int64_t>(cdiv(std::max(M, N) * K, 256), 8192));
    hipLaunchKernelGGL(cast_f32_kernel, dim3(gc), dim3(256), 0, c.stream, A, (float*)a32->ptr, M * K);
    hipLaunchKernelGGL(cast_f32_kernel, dim3(gc), dim3(256), 0, c.stream, B, (float*)b32->ptr, K * N);
    PEPS_HIP(hipGetLastError());
    c.tend("cast_f32", evc);
  }
  hipEvent_t ev = nullptr;
  if (wsp.timed) c.tbegin(cls, &ev);
  std::shared_ptr<Buffer> ws;
  double* out = C;
  int64_t wstride = 0;
  double kbeta = beta;
  if (splits > 1) {
    wstride = M * N * esz;
    const size_t need = size_t(splits) * wstride * 8;
    if (wsp.ws_fixed) {
      PEPS_CHECK(double(need) <= wsp.ws_cap, PS_EUNSUPPORTED, "gemm: split-K workspace exceeds the fixed buffer");
      out = wsp.ws_fixed;
    } else {
      ws = std::make_shared<Buffer>();
      ws->pool = c.pool.get();
      ws->bytes = need;
      ws->ptr = c.pool->get(ws->bytes);
      out = static_cast<double*>(ws->ptr);
    }
    kbeta = 0.0;
  }
  PEPS_CHECK(int64_t(cdiv(M, 64)) * cdiv(N, 64) < (int64_t(1) << 31), PS_EUNSUPPORTED,
             "gemm: more than 2^31 output tiles");
  auto go = [&](auto cplx_tag, auto prec_tag) {
    constexpr bool CP = decltype(cplx_tag)::value;
    constexpr int PR = decltype(prec_tag)::value;
    dispatch_tile<CP, PR>(big, a_kfast, b_kfast, c.stream, M, N, K, A, lda, B, ldb, out, M,
                          conjA, conjB, kbeta, splits, kchunk, wstride);
  };
  using T0 = std::integral_constant<int, 0>;
  using T1 = std::integral_constant<int, 1>;
  using T2 = std::integral_constant<int, 2>;
  if (a32) {
    launch_gemm_f32_tiled<float>(a_kfast, b_kfast, c.stream, M, N, K, (const float*)a32->ptr, lda,
                                 (const float*)b32->ptr, ldb, out, M, kbeta, splits, kchunk, wstride);
  } else if (tiled) {
    launch_gemm_f32_tiled<double>(a_kfast, b_kfast, c.stream, M, N, K, A, lda, B, ldb, out, M, kbeta, splits,
                                  kchunk, wstride);
  } else if (tiled64) {
    launch_gemm_f64_tiled(a_kfast, b_kfast, c.stream, M, N, K, A, lda, B, ldb, out, M, kbeta, splits, kchunk,
                          wstride);
  } else if (cplx) {
    if (c.tier == TIER_FP64) go(std::true_type{}, T0{});
    else if (c.tier == TIER_FP32) go(std::true_type{}, T1{});
    else go(std::true_type{}, T2{});
  } else {
    if (c.tier == TIER_FP64) go(std::false_type{}, T0{});
    else if (c.tier == TIER_FP32) go(std::false_type{}, T1{});
    else go(std::false_type{}, T2{});
  }
  PEPS_HIP(hipGetLastError());
  if (wsp.timed) c.tend(cls, ev);  // the GEMM kernel alone (the split-K reduce is its own class)
  if (splits > 1) {
    int64_t n = M * N * esz;
    unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 8192));
    hipEvent_t ev2 = nullptr;
    if (wsp.timed) {
      c.tbegin("splitk_reduce", &ev2);
      c.count("splitk_reduce", 0.0, double(n) * 8.0 * (splits + 2));
    }
    hipLaunchKernelGGL(splitk_reduce, dim3(g), dim3(256), 0, c.stream, C, out, n, splits, wstride, beta);
    PEPS_HIP(hipGetLastError());
    if (wsp.timed) c.tend("splitk_reduce", ev2);
  }
}
}  // namespace

void gemm(Ctx& c, bool cplx, int64_t M, int64_t N, int64_t K, const double* A, bool a_kfast,
          bool conjA, const double* B, bool b_kfast, bool conjB, double* C, double beta) {
  gemm_impl(c, cplx, M, N, K, A, a_kfast, conjA, B, b_kfast, conjB, C, beta, GemmWorkspace{});
}

// ── the GEMM without allocation ─────────────────────────────────────────────
//
// The split-K partial products go to `scratch`, one array per device that the
// library's code object holds from its first use on that device until the
// process ends.  Host code never passes the array's name to a kernel (on the
// host it is a stand-in, not a device address): scratch_attach, called once
// per device by the context's constructor with that device current, looks up
// the device address with hipGetSymbolAddress and keeps it per device.
// Calls that overlap on one device share the array and race.
static __device__ __align__(256) double scratch[kGemmScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

void scratch_attach(int device) {
  PEPS_CHECK(device >= 0 && device < kMaxDevices, PS_EARG, "scratch_attach: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  PEPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PEPS_CHECK(p != nullptr, PS_EHIP, "scratch_attach: no device address for the scratch array");
  g_scratch[device] = static_cast<double*>(p);
}

void gemm_noalloc(Ctx& c, bool cplx, int64_t M, int64_t N, int64_t K, const double* A, bool a_kfast,
                  bool conjA, const double* B, bool b_kfast, bool conjB, double* C) {
  double* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ws = c.device >= 0 && c.device < kMaxDevices ? g_scratch[c.device] : nullptr;
  }
  PEPS_CHECK(ws != nullptr, PS_EHIP, "gemm_noalloc: the device's scratch array is not attached");
  GemmWorkspace wsp;
  wsp.ws_fixed = ws;
  wsp.ws_cap = double(kGemmScratchBytes);
  wsp.cast = false;
  wsp.timed = false;
  gemm_impl(c, cplx, M, N, K, A, a_kfast, conjA, B, b_kfast, conjB, C, 0.0, wsp);
}

// ════════════════════════════════════════════════════════════════════════════
//  permute
// ════════════════════════════════════════════════════════════════════════════

constexpr int PMAX = 16;
struct PermArgs {
  int nd;
  int64_t odims[PMAX];
  int64_t istr[PMAX];  // input stride (elements) of output dim d
};

template <typename E>
__global__ void permute_kernel(const E* __restrict__ in, E* __restrict__ out, int64_t n,
                               PermArgs p) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < n;
       j += int64_t(gridDim.x) * blockDim.x) {
    int64_t rem = j, off = 0;
#pragma unroll 4
    for (int d = 0; d < p.nd; ++d) {
      int64_t q = rem / p.odims[d];
      off += (rem - q * p.odims[d]) * p.istr[d];
      rem = q;
    }
    out[j] = in[off];
  }
}

// 2-D tiled transpose through LDS for the common "swap two blocks" case:
// in is (R x C) col-major, out is (C x R) col-major, batched over `nb`.
template <typename E>
__global__ void transpose_kernel(const E* __restrict__ in, E* __restrict__ out, int64_t R,
                                 int64_t Cc, int64_t nb) {
  __shared__ E tile[32][33];
  const int64_t b = blockIdx.z;
  const E* inb = in + b * R * Cc;
  E* outb = out + b * R * Cc;
  for (int64_t r0 = int64_t(blockIdx.x) * 32; r0 < R; r0 += int64_t(gridDim.x) * 32)
    for (int64_t c0 = int64_t(blockIdx.y) * 32; c0 < Cc; c0 += int64_t(gridDim.y) * 32) {
      for (int k = threadIdx.y; k < 32; k += 8) {
        int64_t r = r0 + threadIdx.x, cc = c0 + k;
        if (r < R && cc < Cc) tile[k][threadIdx.x] = inb[r + cc * R];
      }
      __syncthreads();
      for (int k = threadIdx.y; k < 32; k += 8) {
        int64_t cc = c0 + threadIdx.x, r = r0 + k;
        if (r < R && cc < Cc) outb[cc + r * Cc] = tile[threadIdx.x][k];
      }
      __syncthreads();
    }
}

// General permutation as a batch of 2-D transposes through LDS: output dim 0
// (output-contiguous, input stride s0) against the output dim whose input
// stride is 1 (output stride o1); every other dim is a batch index whose
// input/output offsets are decoded once per tile, not per element.  Reads
// are coalesced along the input-contiguous dim, writes along output dim 0.
struct PermTiled {
  int nb;
  int64_t bdims[PMAX], bistr[PMAX], bostr[PMAX];
  int64_t n0, n1, s0, o1, nbatch;
};

template <typename E>
__global__ __launch_bounds__(256) void permute_tiled_kernel(const E* __restrict__ in, E* __restrict__ out,
                                                            PermTiled p) {
  __shared__ E tile[32][33];
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int64_t b = blockIdx.z; b < p.nbatch; b += gridDim.z) {
    int64_t rem = b, ioff = 0, ooff = 0;
    for (int d = 0; d < p.nb; ++d) {
      const int64_t q = rem / p.bdims[d], r = rem - q * p.bdims[d];
      ioff += r * p.bistr[d];
      ooff += r * p.bostr[d];
      rem = q;
    }
    for (int64_t t0 = int64_t(blockIdx.x) * 32; t0 < p.n0; t0 += int64_t(gridDim.x) * 32)
      for (int64_t t1 = int64_t(blockIdx.y) * 32; t1 < p.n1; t1 += int64_t(gridDim.y) * 32) {
        for (int k = ty; k < 32; k += 8) {
          const int64_t i0 = t0 + k, i1 = t1 + tx;
          if (i0 < p.n0 && i1 < p.n1) tile[k][tx] = in[ioff + i0 * p.s0 + i1];
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
          const int64_t i0 = t0 + tx, i1 = t1 + k;
          if (i0 < p.n0 && i1 < p.n1) out[ooff + i0 + i1 * p.o1] = tile[tx][k];
        }
        __syncthreads();
      }
  }
}

Tensor permute(Ctx& c, const Tensor& in, const std::vector<int>& perm) {
  const int nd = in.rank();
  PEPS_CHECK(int(perm.size()) == nd, PS_ESHAPE, "permute: rank mismatch");
  std::vector<int64_t> od(nd);
  for (int i = 0; i < nd; ++i) od[i] = in.dims[perm[i]];
  bool ident = true;
  for (int i = 0; i < nd; ++i) ident = ident && (perm[i] == i || in.dims[perm[i]] == 1);
  // identity up to unit dims: pure reshape
  {
    std::vector<int> nz;
    for (int i = 0; i < nd; ++i)
      if (in.dims[perm[i]] != 1) nz.push_back(perm[i]);
    bool sorted = true;
    for (size_t i = 1; i < nz.size(); ++i) sorted = sorted && nz[i] > nz[i - 1];
    if (sorted) return in.reshaped(od);
  }
  (void)ident;
  Tensor out = c.alloc(od, in.cplx);
  const int64_t n = in.numel();
  if (n == 0) return out;
  std::vector<int64_t> istride(nd);
  {
    int64_t s = 1;
    for (int i = 0; i < nd; ++i) {
      istride[i] = s;
      s *= in.dims[i];
    }
  }
  // collapse: drop unit dims, merge output dims that are contiguous in input
  std::vector<int64_t> cd, cs;
  for (int i = 0; i < nd; ++i) {
    if (od[i] == 1) continue;
    int64_t st = istride[perm[i]];
    if (!cd.empty() && cs.back() * cd.back() == st) {
      cd.back() *= od[i];
    } else {
      cd.push_back(od[i]);
      cs.push_back(st);
    }
  }
  std::string pcls = "permute";
  if (c.gemm_shapes) {
    pcls += in.cplx ? " c" : " r";
    for (size_t i = 0; i < cd.size(); ++i) pcls += " " + std::to_string(cd[i]) + "@" + std::to_string(cs[i]);
  }
  c.count(pcls, 0, 2.0 * in.bytes());
  hipEvent_t ev = nullptr;
  c.tbegin(pcls, &ev);
  // two-block swap => tiled transpose: output (d0, d1) with d0 contiguous at
  // input stride d1-extent.. i.e. input is (d1 x d0) col-major.
  bool tr = false;
  int64_t R = 0, Cc = 0, nb = 1;
  if (cd.size() == 2 && cs[1] == 1 && cs[0] == cd[1]) {
    tr = true; R = cd[1]; Cc = cd[0];
  } else if (cd.size() == 3 && cs[1] == 1 && cs[0] == cd[1] && cs[2] == cd[0] * cd[1]) {
    tr = true; R = cd[1]; Cc = cd[0]; nb = cd[2];
  }
  // input-contiguous output dim, for the batched tiled transpose
  int kc = -1;
  for (size_t i = 1; i < cs.size(); ++i)
    if (cs[i] == 1) kc = int(i);
  const bool tiled = !tr && kc > 0 && cs[0] != 1 && cd[0] >= 16 && cd[size_t(kc)] >= 16 &&
                     int(cd.size()) - 2 <= PMAX;
  if (tiled) {
    PermTiled p{};
    p.n0 = cd[0];
    p.s0 = cs[0];
    p.n1 = cd[size_t(kc)];
    int64_t ostr = 1;
    p.nbatch = 1;
    for (size_t i = 0; i < cd.size(); ++i) {
      if (int(i) == kc) p.o1 = ostr;
      if (i != 0 && int(i) != kc) {
        p.bdims[p.nb] = cd[i];
        p.bistr[p.nb] = cs[i];
        p.bostr[p.nb] = ostr;
        ++p.nb;
        p.nbatch *= cd[i];
      }
      ostr *= cd[i];
    }
    dim3 grid(unsigned(std::min<int64_t>(cdiv(p.n0, 32), 64)), unsigned(std::min<int64_t>(cdiv(p.n1, 32), 64)),
              unsigned(std::min<int64_t>(p.nbatch, 16384)));
    if (in.cplx)
      hipLaunchKernelGGL(permute_tiled_kernel<double2>, grid, dim3(32, 8), 0, c.stream,
                         (const double2*)in.data(), (double2*)out.data(), p);
    else
      hipLaunchKernelGGL(permute_tiled_kernel<double>, grid, dim3(32, 8), 0, c.stream, in.data(), out.data(), p);
  } else if (tr && nb < 65536) {
    dim3 grid(std::min<unsigned>(cdiv(R, 32), 1024u), std::min<unsigned>(cdiv(Cc, 32), 1024u), unsigned(nb));
    if (in.cplx)
      hipLaunchKernelGGL(transpose_kernel<double2>, grid, dim3(32, 8), 0, c.stream,
                         (const double2*)in.data(), (double2*)out.data(), R, Cc, nb);
    else
      hipLaunchKernelGGL(transpose_kernel<double>, grid, dim3(32, 8), 0, c.stream, in.data(),
                         out.data(), R, Cc, nb);
  } else {
    PEPS_CHECK(int(cd.size()) <= PMAX, PS_EUNSUPPORTED, "permute: too many dims");
    PermArgs p{};
    p.nd = int(cd.size());
    for (int i = 0; i < p.nd; ++i) {
      p.odims[i] = cd[i];
      p.istr[i] = cs[i];
    }
    unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 65536));
    if (in.cplx)
      hipLaunchKernelGGL(permute_kernel<double2>, dim3(g), dim3(256), 0, c.stream,
                         (const double2*)in.data(), (double2*)out.data(), n, p);
    else
      hipLaunchKernelGGL(permute_kernel<double>, dim3(g), dim3(256), 0, c.stream, in.data(),
                         out.data(), n, p);
  }
  PEPS_HIP(hipGetLastError());
  c.tend(pcls, ev);
  return out;
}

// ════════════════════════════════════════════════════════════════════════════
//  elementwise
// ════════════════════════════════════════════════════════════════════════════

constexpr int WMAX = 64;
struct LegW {
  int64_t d[5];
  double w[4][WMAX];
};

__global__ void scale_legs_kernel(double* T, int esz, int64_t n, LegW a) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i / a.d[0];
    int64_t l = r % a.d[1]; r /= a.d[1];
    int64_t rr = r % a.d[2]; r /= a.d[2];
    int64_t u = r % a.d[3]; r /= a.d[3];
    int64_t d = r;
    double s = a.w[0][l] * a.w[1][rr] * a.w[2][u] * a.w[3][d];
    T[i * esz] *= s;
    if (esz == 2) T[i * esz + 1] *= s;
  }
}

void scale_legs(Ctx& c, Tensor& T, const std::vector<double>* w) {
  PEPS_CHECK(T.rank() == 5, PS_ESHAPE, "scale_legs: rank-5 tensor expected");
  LegW a{};
  for (int i = 0; i < 5; ++i) a.d[i] = T.dims[i];
  for (int k = 0; k < 4; ++k) {
    PEPS_CHECK(int64_t(w[k].size()) == T.dims[k + 1], PS_ESHAPE, "scale_legs: weight length");
    PEPS_CHECK(w[k].size() <= size_t(WMAX), PS_EUNSUPPORTED, "scale_legs: bond dimension > 64");
    for (size_t j = 0; j < w[k].size(); ++j) a.w[k][j] = w[k][j];
  }
  int64_t n = T.numel();
  if (!n) return;
  c.count("elementwise", 0, 2.0 * T.bytes());
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(scale_legs_kernel, dim3(g), dim3(256), 0, c.stream, T.data(),
                     T.cplx ? 2 : 1, n, a);
  PEPS_HIP(hipGetLastError());
}

__global__ void scale_kernel(double* x, int64_t n, double s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] *= s;
}
void scale(Ctx& c, Tensor& T, double s) {
  int64_t n = T.numel() * (T.cplx ? 2 : 1);
  if (!n) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(scale_kernel, dim3(g), dim3(256), 0, c.stream, T.data(), n, s);
  PEPS_HIP(hipGetLastError());
}

__global__ void axpy_kernel(double* y, const double* x, int64_t n, double a) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] += a * x[i];
}
void axpy_inplace(Ctx& c, Tensor& Y, const Tensor& X, double a) {
  PEPS_CHECK(Y.numel() == X.numel() && Y.cplx == X.cplx, PS_ESHAPE, "axpy: shape");
  int64_t n = Y.numel() * (Y.cplx ? 2 : 1);
  if (!n) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(axpy_kernel, dim3(g), dim3(256), 0, c.stream, Y.data(), X.data(), n, a);
  PEPS_HIP(hipGetLastError());
}

__global__ void scale_cplx_kernel(double* x, int64_t n, double re, double im) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double a = x[2 * i], b = x[2 * i + 1];
    x[2 * i] = a * re - b * im;
    x[2 * i + 1] = a * im + b * re;
  }
}
void scale(Ctx& c, Tensor& T, cplx s) {
  if (!T.cplx) {
    PEPS_CHECK(s.im == 0.0, PS_EARG, "scale: complex factor on a real tensor");
    scale(c, T, s.re);
    return;
  }
  if (!T.numel()) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(T.numel(), 256), 16384));
  hipLaunchKernelGGL(scale_cplx_kernel, dim3(g), dim3(256), 0, c.stream, T.data(), T.numel(), s.re, s.im);
  PEPS_HIP(hipGetLastError());
}

__global__ void real_to_cplx_kernel(double* out, const double* in, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    out[2 * i] = in[i];
    out[2 * i + 1] = 0.0;
  }
}
Tensor as_complex(Ctx& c, const Tensor& T) {
  if (T.cplx) return T;
  Tensor out = c.alloc(T.dims, true);
  if (!T.numel()) return out;
  unsigned g = unsigned(std::min<int64_t>(cdiv(T.numel(), 256), 16384));
  hipLaunchKernelGGL(real_to_cplx_kernel, dim3(g), dim3(256), 0, c.stream, out.data(), T.data(), T.numel());
  PEPS_HIP(hipGetLastError());
  return out;
}

__global__ void conj_kernel(double* x, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[2 * i + 1] = -x[2 * i + 1];
}
void conj_inplace(Ctx& c, Tensor& T) {
  if (!T.cplx || !T.numel()) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(T.numel(), 256), 16384));
  hipLaunchKernelGGL(conj_kernel, dim3(g), dim3(256), 0, c.stream, T.data(), T.numel());
  PEPS_HIP(hipGetLastError());
}

// ── Gauss mask ───────────────────────────────────────────────────────────────
struct MaskArgs {
  int64_t d[5];
  int64_t q;
  int64_t nf[WMAX], lR[WMAX], lU[WMAX];
  int64_t fl[WMAX], fr[WMAX], fu[WMAX], fd[WMAX];
};

__device__ __forceinline__ bool mask_ok(const MaskArgs& a, int64_t i) {
  int64_t p = i % a.d[0];
  int64_t r = i / a.d[0];
  int64_t l = r % a.d[1]; r /= a.d[1];
  int64_t rr = r % a.d[2]; r /= a.d[2];
  int64_t u = r % a.d[3]; r /= a.d[3];
  int64_t d = r;
  int64_t R = a.lR[p], U = a.lU[p];
  return a.fr[rr] == R && a.fu[u] == U && (R + U - a.fl[l] - a.fd[d] == 2 * (a.nf[p] + a.q));
}

__global__ void mask_kernel(double* T, int esz, int64_t n, MaskArgs a) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (!mask_ok(a, i)) {
      T[i * esz] = 0.0;
      if (esz == 2) T[i * esz + 1] = 0.0;
    }
  }
}

__global__ void mask_viol_kernel(const double* T, int esz, int64_t n, MaskArgs a, double* part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (!mask_ok(a, i)) {
      double x = T[i * esz];
      s += x * x;
      if (esz == 2) {
        double y = T[i * esz + 1];
        s += y * y;
      }
    }
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static MaskArgs make_mask_args(const Tensor& T, const std::vector<int64_t>& nf,
                               const std::vector<int64_t>& lR, const std::vector<int64_t>& lU,
                               const std::vector<int64_t>& fl, const std::vector<int64_t>& fr,
                               const std::vector<int64_t>& fu, const std::vector<int64_t>& fd,
                               int64_t q) {
  PEPS_CHECK(T.rank() == 5, PS_ESHAPE, "mask: rank-5 tensor expected");
  MaskArgs a{};
  for (int i = 0; i < 5; ++i) a.d[i] = T.dims[i];
  PEPS_CHECK(int64_t(nf.size()) == T.dims[0] && lR.size() == nf.size() && lU.size() == nf.size() &&
                 int64_t(fl.size()) == T.dims[1] &&
                 int64_t(fr.size()) == T.dims[2] && int64_t(fu.size()) == T.dims[3] &&
                 int64_t(fd.size()) == T.dims[4],
             PS_ESHAPE, "mask: label length mismatch");
  PEPS_CHECK(T.dims[0] <= WMAX && T.dims[1] <= WMAX && T.dims[2] <= WMAX && T.dims[3] <= WMAX &&
                 T.dims[4] <= WMAX,
             PS_EUNSUPPORTED, "mask: dimension > 64");
  a.q = q;
  for (size_t i = 0; i < nf.size(); ++i) { a.nf[i] = nf[i]; a.lR[i] = lR[i]; a.lU[i] = lU[i]; }
  for (size_t i = 0; i < fl.size(); ++i) a.fl[i] = fl[i];
  for (size_t i = 0; i < fr.size(); ++i) a.fr[i] = fr[i];
  for (size_t i = 0; i < fu.size(); ++i) a.fu[i] = fu[i];
  for (size_t i = 0; i < fd.size(); ++i) a.fd[i] = fd[i];
  return a;
}

void gauss_mask(Ctx& c, Tensor& T, const std::vector<int64_t>& nf, const std::vector<int64_t>& lR,
                const std::vector<int64_t>& lU, const std::vector<int64_t>& fl,
                const std::vector<int64_t>& fr, const std::vector<int64_t>& fu,
                const std::vector<int64_t>& fd, int64_t q) {
  MaskArgs a = make_mask_args(T, nf, lR, lU, fl, fr, fu, fd, q);
  int64_t n = T.numel();
  if (!n) return;
  c.count("elementwise", 0, 2.0 * T.bytes());
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(mask_kernel, dim3(g), dim3(256), 0, c.stream, T.data(), T.cplx ? 2 : 1, n, a);
  PEPS_HIP(hipGetLastError());
}

double gauss_violation(Ctx& c, const Tensor& T, const std::vector<int64_t>& nf,
                       const std::vector<int64_t>& lR, const std::vector<int64_t>& lU,
                       const std::vector<int64_t>& fl, const std::vector<int64_t>& fr,
                       const std::vector<int64_t>& fu, const std::vector<int64_t>& fd, int64_t q) {
  MaskArgs a = make_mask_args(T, nf, lR, lU, fl, fr, fu, fd, q);
  int64_t n = T.numel();
  if (!n) return 0.0;
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 1024));
  Tensor part = c.alloc({int64_t(g)}, false);
  hipLaunchKernelGGL(mask_viol_kernel, dim3(g), dim3(256), 0, c.stream, T.data(),
                     T.cplx ? 2 : 1, n, a, part.data());
  PEPS_HIP(hipGetLastError());
  auto h = c.to_host(part);
  double s = 0.0;
  for (double x : h) s += x;
  return s;
}

// ── reductions ───────────────────────────────────────────────────────────────
__global__ void dot_kernel(const double* a, const double* b, int64_t n, int cplx, int conjb,
                           double* part) {
  __shared__ double sre[256], sim[256];
  double re = 0.0, im = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (cplx) {
      double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
      if (conjb) bi = -bi;
      re += ar * br - ai * bi;
      im += ar * bi + ai * br;
    } else {
      re += a[i] * b[i];
    }
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      sre[threadIdx.x] += sre[threadIdx.x + o];
      sim[threadIdx.x] += sim[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sre[0];
    part[2 * blockIdx.x + 1] = sim[0];
  }
}

cplx dot(Ctx& c, const Tensor& a, const Tensor& b, bool conj_b) {
  PEPS_CHECK(a.numel() == b.numel() && a.cplx == b.cplx, PS_ESHAPE, "dot: shape mismatch");
  int64_t n = a.numel();
  if (!n) return cplx{0, 0};
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 1024));
  Tensor part = c.alloc({2 * int64_t(g)}, false);
  c.count("reduce", 0, 2.0 * a.bytes());
  hipLaunchKernelGGL(dot_kernel, dim3(g), dim3(256), 0, c.stream, a.data(), b.data(), n,
                     a.cplx ? 1 : 0, conj_b ? 1 : 0, part.data());
  PEPS_HIP(hipGetLastError());
  auto h = c.to_host(part);
  cplx s{0, 0};
  for (unsigned i = 0; i < g; ++i) {
    s.re += h[2 * i];
    s.im += h[2 * i + 1];
  }
  return s;
}

double norm2sq(Ctx& c, const Tensor& a) { return dot(c, a, a, true).re; }

// one wave per column: first row index of max |M|^2
__global__ void col_argmax_kernel(const double* M, int64_t nrow, int64_t ncol, int cplx,
                                  int64_t* out) {
  const int lane = threadIdx.x & 63;
  const int64_t col = blockIdx.x * int64_t(blockDim.x / 64) + (threadIdx.x >> 6);
  if (col >= ncol) return;
  double best = -1.0;
  int64_t bi = 0;
  for (int64_t r = lane; r < nrow; r += 64) {
    int64_t off = r + col * nrow;
    double v = cplx ? (M[2 * off] * M[2 * off] + M[2 * off + 1] * M[2 * off + 1]) : M[off] * M[off];
    if (v > best) {
      best = v;
      bi = r;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    double ob = __shfl_xor(best, o);
    int64_t oi = __shfl_xor(bi, o);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  if (lane == 0) out[col] = bi;
}

std::vector<int64_t> col_argmax(Ctx& c, const Tensor& M) {
  PEPS_CHECK(M.rank() == 2, PS_ESHAPE, "col_argmax: matrix expected");
  int64_t nrow = M.dims[0], ncol = M.dims[1];
  std::vector<int64_t> h(ncol);
  if (!ncol) return h;
  auto buf = std::make_shared<Buffer>();
  buf->pool = c.pool.get();
  buf->bytes = size_t(ncol) * 8;
  buf->ptr = c.pool->get(buf->bytes);
  hipLaunchKernelGGL(col_argmax_kernel, dim3(cdiv(ncol, 4)), dim3(256), 0, c.stream, M.data(),
                     nrow, ncol, M.cplx ? 1 : 0, static_cast<int64_t*>(buf->ptr));
  PEPS_HIP(hipGetLastError());
  PEPS_HIP(hipMemcpyAsync(h.data(), buf->ptr, buf->bytes, hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  return h;
}

__global__ void cross_zero_kernel(double* M, int64_t nrow, int64_t ncol, int esz,
                                  const int64_t* rq, const int64_t* cq) {
  int64_t n = nrow * ncol;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % nrow, cc = i / nrow;
    if (rq[r] != cq[cc]) {
      M[i * esz] = 0.0;
      if (esz == 2) M[i * esz + 1] = 0.0;
    }
  }
}

void zero_cross_sector(Ctx& c, Tensor& M, const std::vector<int64_t>& rq,
                       const std::vector<int64_t>& cq) {
  int64_t nrow = M.dims[0], ncol = M.dims[1];
  PEPS_CHECK(M.rank() == 2 && int64_t(rq.size()) == nrow && int64_t(cq.size()) == ncol, PS_ESHAPE,
             "cross: labels");
  if (nrow == 0 || ncol == 0) return;
  std::shared_ptr<Buffer> k1, k2;
  auto* drq = (const int64_t*)c.scratch_upload(rq.data(), rq.size() * 8, &k1);
  auto* dcq = (const int64_t*)c.scratch_upload(cq.data(), cq.size() * 8, &k2);
  unsigned g = unsigned(std::min<int64_t>(cdiv(nrow * ncol, 256), 16384));
  hipLaunchKernelGGL(cross_zero_kernel, dim3(g), dim3(256), 0, c.stream, M.data(), nrow, ncol,
                     M.cplx ? 2 : 1, drq, dcq);
  PEPS_HIP(hipGetLastError());
}

// An index list as a kernel sees it: the uploaded entries, or, for a range
// off, off+1, ... (p == nullptr), the entry computed.
struct Idx {
  const int64_t* p;
  int64_t off;
  __device__ __forceinline__ int64_t operator[](int64_t i) const { return p ? p[i] : off + i; }
};

// out(i, j) = M(rows[i], cols[j]); M column-major nrow x ncol
__global__ void gather2d_kernel(const double* M, int64_t nrow, double* out, int64_t nr, int64_t nc, int esz,
                                Idx rows, Idx cols) {
  int64_t n = nr * nc;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % nr, cc = i / nr;
    int64_t src = rows[r] + cols[cc] * nrow;
    out[i * esz] = M[src * esz];
    if (esz == 2) out[i * esz + 1] = M[src * esz + 1];
  }
}

// Host-side bounds check of an index list (the kernels below index with it).
static void check_index(const std::vector<int64_t>& ix, int64_t bound, const char* what) {
  for (int64_t v : ix)
    PEPS_CHECK(v >= 0 && v < bound, PS_EARG, std::string(what) + ": index out of range");
}

static int64_t mat_cols(const Tensor& T) { return T.dims[0] ? T.numel() / T.dims[0] : 0; }

// The index lists of one gather/scatter launch: ranges travel as their first
// entry, the others are staged to the device together in one copy.
struct IdxSet {
  std::shared_ptr<Buffer> keep;
  Idx ix[3] = {{nullptr, 0}, {nullptr, 0}, {nullptr, 0}};
};
static IdxSet device_indices(Ctx& c, std::initializer_list<const std::vector<int64_t>*> lists) {
  IdxSet s;
  std::vector<int64_t> packed;
  size_t at[3] = {0, 0, 0};
  bool up[3] = {false, false, false};
  int k = 0;
  for (const std::vector<int64_t>* v : lists) {
    bool range = !v->empty();
    for (size_t i = 1; range && i < v->size(); ++i) range = (*v)[i] == (*v)[0] + int64_t(i);
    if (range) {
      s.ix[k].off = (*v)[0];
    } else {
      up[k] = true;
      at[k] = packed.size();
      packed.insert(packed.end(), v->begin(), v->end());
    }
    ++k;
  }
  if (!packed.empty()) {
    auto* base = static_cast<const int64_t*>(c.scratch_upload(packed.data(), packed.size() * 8, &s.keep));
    for (int j = 0; j < 3; ++j)
      if (up[j]) s.ix[j].p = base + at[j];
  }
  return s;
}

Tensor gather2d(Ctx& c, const Tensor& M, const std::vector<int64_t>& rows,
                const std::vector<int64_t>& cols) {
  PEPS_CHECK(M.rank() >= 1, PS_ESHAPE, "gather2d: matrix expected");
  check_index(rows, M.dims[0], "gather2d rows");
  check_index(cols, mat_cols(M), "gather2d cols");
  Tensor out = c.alloc({int64_t(rows.size()), int64_t(cols.size())}, M.cplx);
  if (rows.empty() || cols.empty()) return out;
  IdxSet ix = device_indices(c, {&rows, &cols});
  int64_t n = int64_t(rows.size() * cols.size());
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(gather2d_kernel, dim3(g), dim3(256), 0, c.stream, M.data(), M.dims[0],
                     out.data(), int64_t(rows.size()), int64_t(cols.size()), M.cplx ? 2 : 1, ix.ix[0], ix.ix[1]);
  PEPS_HIP(hipGetLastError());
  return out;
}

// dst(rows[i], dst_cols[j]) = src(i, src_cols[j])
__global__ void scatter_cols_kernel(double* dst, int64_t dnrow, const double* src, int64_t snrow,
                                    int64_t nr, int64_t nc, int esz, Idx rows, Idx scols, Idx dcols) {
  int64_t n = nr * nc;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % nr, j = i / nr;
    int64_t so = r + scols[j] * snrow;
    int64_t d = rows[r] + dcols[j] * dnrow;
    dst[d * esz] = src[so * esz];
    if (esz == 2) dst[d * esz + 1] = src[so * esz + 1];
  }
}

void scatter_cols(Ctx& c, Tensor& dst, const Tensor& src, const std::vector<int64_t>& rows,
                  const std::vector<int64_t>& src_cols, const std::vector<int64_t>& dst_cols) {
  PEPS_CHECK(src_cols.size() == dst_cols.size(), PS_EARG, "scatter_cols");
  PEPS_CHECK(int64_t(rows.size()) <= src.dims[0], PS_ESHAPE, "scatter_cols: rows");
  check_index(rows, dst.dims[0], "scatter_cols rows");
  check_index(src_cols, mat_cols(src), "scatter_cols src");
  check_index(dst_cols, mat_cols(dst), "scatter_cols dst");
  if (rows.empty() || src_cols.empty()) return;
  IdxSet ix = device_indices(c, {&rows, &src_cols, &dst_cols});
  int64_t n = int64_t(rows.size() * src_cols.size());
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(scatter_cols_kernel, dim3(g), dim3(256), 0, c.stream, dst.data(), dst.dims[0],
                     src.data(), src.dims[0], int64_t(rows.size()), int64_t(src_cols.size()),
                     dst.cplx ? 2 : 1, ix.ix[0], ix.ix[1], ix.ix[2]);
  PEPS_HIP(hipGetLastError());
}

// dst(dst_rows[i], cols[j]) = src(src_rows[i], j)
__global__ void scatter_rows_kernel(double* dst, int64_t dnrow, const double* src, int64_t snrow,
                                    int64_t nr, int64_t nc, int esz, Idx srows, Idx drows, Idx cols) {
  int64_t n = nr * nc;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % nr, j = i / nr;
    int64_t so = srows[r] + j * snrow;
    int64_t d = drows[r] + cols[j] * dnrow;
    dst[d * esz] = src[so * esz];
    if (esz == 2) dst[d * esz + 1] = src[so * esz + 1];
  }
}

void scatter_rows(Ctx& c, Tensor& dst, const Tensor& src, const std::vector<int64_t>& src_rows,
                  const std::vector<int64_t>& dst_rows, const std::vector<int64_t>& cols) {
  PEPS_CHECK(src_rows.size() == dst_rows.size(), PS_EARG, "scatter_rows");
  check_index(src_rows, src.dims[0], "scatter_rows src");
  check_index(dst_rows, dst.dims[0], "scatter_rows dst");
  check_index(cols, mat_cols(dst), "scatter_rows cols");
  PEPS_CHECK(int64_t(cols.size()) <= mat_cols(src), PS_ESHAPE, "scatter_rows: cols");
  if (src_rows.empty() || cols.empty()) return;
  IdxSet ix = device_indices(c, {&src_rows, &dst_rows, &cols});
  int64_t n = int64_t(src_rows.size() * cols.size());
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(scatter_rows_kernel, dim3(g), dim3(256), 0, c.stream, dst.data(), dst.dims[0],
                     src.data(), src.dims[0], int64_t(src_rows.size()), int64_t(cols.size()),
                     dst.cplx ? 2 : 1, ix.ix[0], ix.ix[1], ix.ix[2]);
  PEPS_HIP(hipGetLastError());
}

__global__ void mask_x_kernel(double* X, int64_t nA, int64_t dp, int64_t Dp, int esz,
                              const int8_t* keep) {
  int64_t n = nA * dp * Dp;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i / nA;
    int64_t p = r % dp, b = r / dp;
    if (!keep[p + b * dp]) {
      X[i * esz] = 0.0;
      if (esz == 2) X[i * esz + 1] = 0.0;
    }
  }
}

void mask_x(Ctx& c, Tensor& X, const std::vector<int8_t>& keep) {
  PEPS_CHECK(X.rank() == 3 && int64_t(keep.size()) == X.dims[1] * X.dims[2], PS_ESHAPE, "mask_x");
  std::shared_ptr<Buffer> k1;
  auto* dk = (const int8_t*)c.scratch_upload(keep.data(), keep.size(), &k1);
  int64_t n = X.numel();
  if (!n) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(mask_x_kernel, dim3(g), dim3(256), 0, c.stream, X.data(), X.dims[0],
                     X.dims[1], X.dims[2], X.cplx ? 2 : 1, dk);
  PEPS_HIP(hipGetLastError());
}

__global__ void upper_trap_kernel(const double* A, int64_t m, double* R, int64_t k, int64_t n,
                                  int esz) {
  int64_t tot = k * n;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < tot;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % k, cc = i / k;
    double re = 0.0, im = 0.0;
    if (r <= cc) {
      re = A[(r + cc * m) * esz];
      if (esz == 2) im = A[(r + cc * m) * esz + 1];
    }
    R[i * esz] = re;
    if (esz == 2) R[i * esz + 1] = im;
  }
}

Tensor upper_trapezoid(Ctx& c, const Tensor& A, int64_t k) {
  int64_t m = A.dims[0], n = A.dims[1];
  Tensor R = c.alloc({k, n}, A.cplx);
  if (!k || !n) return R;
  unsigned g = unsigned(std::min<int64_t>(cdiv(k * n, 256), 16384));
  hipLaunchKernelGGL(upper_trap_kernel, dim3(g), dim3(256), 0, c.stream, A.data(), m, R.data(), k,
                     n, A.cplx ? 2 : 1);
  PEPS_HIP(hipGetLastError());
  return R;
}

__global__ void scale_rc_kernel(double* M, int64_t nrow, int64_t ncol, int esz, const double* s,
                                int by_col) {
  int64_t n = nrow * ncol;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double f = by_col ? s[i / nrow] : s[i % nrow];
    M[i * esz] *= f;
    if (esz == 2) M[i * esz + 1] *= f;
  }
}

static void scale_rc(Ctx& c, Tensor& M, const std::vector<double>& s, bool by_col) {
  int64_t nrow = M.dims[0], ncol = M.numel() / std::max<int64_t>(1, nrow);
  PEPS_CHECK(int64_t(s.size()) == (by_col ? ncol : nrow), PS_ESHAPE, "scale_rc: length");
  if (!M.numel()) return;
  std::shared_ptr<Buffer> k1;
  auto* ds = (const double*)c.scratch_upload(s.data(), s.size() * 8, &k1);
  unsigned g = unsigned(std::min<int64_t>(cdiv(M.numel(), 256), 16384));
  hipLaunchKernelGGL(scale_rc_kernel, dim3(g), dim3(256), 0, c.stream, M.data(), nrow, ncol,
                     M.cplx ? 2 : 1, ds, by_col ? 1 : 0);
  PEPS_HIP(hipGetLastError());
}
void scale_cols(Ctx& c, Tensor& M, const std::vector<double>& s) { scale_rc(c, M, s, true); }
void scale_rows(Ctx& c, Tensor& M, const std::vector<double>& s) { scale_rc(c, M, s, false); }

__global__ void hermitize_kernel(double* N, int64_t n, int esz) {
  int64_t tot = n * n;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < tot;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % n, cc = i / n;
    if (r > cc) continue;
    int64_t j = cc + r * n;
    if (esz == 1) {
      double v = 0.5 * (N[i] + N[j]);
      N[i] = v;
      N[j] = v;
    } else {
      double re = 0.5 * (N[2 * i] + N[2 * j]);
      double im = 0.5 * (N[2 * i + 1] - N[2 * j + 1]);
      N[2 * i] = re;
      N[2 * i + 1] = im;
      N[2 * j] = re;
      N[2 * j + 1] = -im;
    }
  }
}
void hermitize(Ctx& c, Tensor& N) {
  int64_t n = N.dims[0];
  PEPS_CHECK(N.rank() == 2 && N.dims[1] == n, PS_ESHAPE, "hermitize: square matrix expected");
  if (!n) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(n * n, 256), 16384));
  hipLaunchKernelGGL(hermitize_kernel, dim3(g), dim3(256), 0, c.stream, N.data(), n,
                     N.cplx ? 2 : 1);
  PEPS_HIP(hipGetLastError());
}

Tensor copy_tensor(Ctx& c, const Tensor& t) {
  Tensor o = c.alloc(t.dims, t.cplx);
  if (t.bytes())
    PEPS_HIP(hipMemcpyAsync(o.data(), t.data(), t.bytes(), hipMemcpyDeviceToDevice, c.stream));
  return o;
}

// flags[0] = 1 if any entry is non-finite, flags[1] = 1 if any entry is non-zero.
__global__ void scan_kernel(const double* x, int64_t n, int* flags) {
  int bad = 0, nz = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double v = x[i];
    if (!isfinite(v)) bad = 1;
    else if (v != 0.0) nz = 1;
  }
  if (bad) flags[0] = 1;
  if (nz) flags[1] = 1;
}

void scan_values(Ctx& c, const Tensor& t, bool* finite, bool* nonzero) {
  int64_t n = t.numel() * (t.cplx ? 2 : 1);
  *finite = true;
  *nonzero = false;
  if (!n) return;
  auto buf = std::make_shared<Buffer>();
  buf->pool = c.pool.get();
  buf->bytes = 8;
  buf->ptr = c.pool->get(8);
  PEPS_HIP(hipMemsetAsync(buf->ptr, 0, 8, c.stream));
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 4096));
  hipLaunchKernelGGL(scan_kernel, dim3(g), dim3(256), 0, c.stream, t.data(), n, (int*)buf->ptr);
  PEPS_HIP(hipGetLastError());
  int h[2] = {0, 0};
  PEPS_HIP(hipMemcpyAsync(h, buf->ptr, 8, hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  *finite = h[0] == 0;
  *nonzero = h[1] != 0;
}

bool all_finite(Ctx& c, std::initializer_list<const Tensor*> ts) {
  auto buf = std::make_shared<Buffer>();
  buf->pool = c.pool.get();
  buf->bytes = 8;
  buf->ptr = c.pool->get(8);
  PEPS_HIP(hipMemsetAsync(buf->ptr, 0, 8, c.stream));
  for (const Tensor* t : ts) {
    int64_t n = t->numel() * (t->cplx ? 2 : 1);
    if (!n) continue;
    unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 4096));
    hipLaunchKernelGGL(scan_kernel, dim3(g), dim3(256), 0, c.stream, t->data(), n, (int*)buf->ptr);
    PEPS_HIP(hipGetLastError());
  }
  int h[2] = {0, 0};
  PEPS_HIP(hipMemcpyAsync(h, buf->ptr, 8, hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  return h[0] == 0;
}

// out = the first columns of the identity (m x k).
__global__ void eye_kernel(double* out, int64_t m, int64_t k, int esz) {
  int64_t n = m * k;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = i % m, cc = i / m;
    out[i * esz] = r == cc ? 1.0 : 0.0;
    if (esz == 2) out[i * esz + 1] = 0.0;
  }
}

Tensor eye(Ctx& c, int64_t m, int64_t k, bool cplx) {
  Tensor out = c.alloc({m, k}, cplx);
  if (m * k == 0) return out;
  unsigned g = unsigned(std::min<int64_t>(cdiv(m * k, 256), 16384));
  hipLaunchKernelGGL(eye_kernel, dim3(g), dim3(256), 0, c.stream, out.data(), m, k, cplx ? 2 : 1);
  PEPS_HIP(hipGetLastError());
  return out;
}

bool all_finite(Ctx& c, const Tensor& t) {
  bool f, nz;
  scan_values(c, t, &f, &nz);
  return f;
}


// ════════════════════════════════════════════════════════════════════════════
//  Complex/real helpers of the one-sided Jacobi SVD kernel
// ════════════════════════════════════════════════════════════════════════════

namespace {
template <bool CP> struct JT;
template <> struct JT<false> {
  typedef double T;
  static __device__ double re(T x) { return x; }
  static __device__ double abs2(T x) { return x * x; }
  static __device__ double mag(T x) { return fabs(x); }
};
template <> struct JT<true> {
  typedef double2 T;
  static __device__ double re(T x) { return x.x; }
  static __device__ double abs2(T x) { return x.x * x.x + x.y * x.y; }
  static __device__ double mag(T x) { return hypot(x.x, x.y); }
};
__device__ __forceinline__ double mulc(double a, double b) { return a * b; }
__device__ __forceinline__ double2 mulc(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double conjv(double a) { return a; }
__device__ __forceinline__ double2 conjv(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double scl(double a, double s) { return a * s; }
__device__ __forceinline__ double2 scl(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double addv(double a, double b) { return a + b; }
__device__ __forceinline__ double2 addv(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double unitv(double b, double ab) { return b >= 0.0 ? 1.0 : -1.0; }
__device__ __forceinline__ double2 unitv(double2 b, double ab) {
  return ab > 0 ? make_double2(b.x / ab, b.y / ab) : make_double2(1.0, 0.0);
}
__device__ __forceinline__ void set_one(double& x) { x = 1.0; }
__device__ __forceinline__ void set_one(double2& x) { x = make_double2(1.0, 0.0); }
__device__ __forceinline__ void set_zero(double& x) { x = 0.0; }
__device__ __forceinline__ void set_zero(double2& x) { x = make_double2(0.0, 0.0); }
}  // namespace

}  // namespace peps

namespace peps {

// Deterministic uniform(-1,1) fill (splitmix64 hash of the index).
__global__ void fill_random_kernel(double* x, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * uint64_t(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    x[i] = double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

void fill_random(Ctx& c, Tensor& T, uint64_t seed) {
  int64_t n = T.numel() * (T.cplx ? 2 : 1);
  if (!n) return;
  unsigned g = unsigned(std::min<int64_t>(cdiv(n, 256), 16384));
  hipLaunchKernelGGL(fill_random_kernel, dim3(g), dim3(256), 0, c.stream, T.data(), n, seed);
  PEPS_HIP(hipGetLastError());
}

}  // namespace peps

namespace peps {

// ════════════════════════════════════════════════════════════════════════════
//  One-sided (Hestenes) Jacobi SVD, one workgroup per matrix
// ════════════════════════════════════════════════════════════════════════════
//
// A (m x n, m >= n, column-major) is orthogonalised column pair by column
// pair: the Gram block [[|a_p|^2, g],[conj g, |a_q|^2]], g = a_p^H a_q, is
// diagonalised by the Jacobi rotation of the Hermitian eigensolver and the
// rotation is applied to the columns of A and V.  Pairs follow the same
// round-robin schedule; one wave owns one pair at a time, lanes stride the
// rows, sums are reduced with an XOR butterfly (identical in every lane, so
// the rotate/skip decision is wave-uniform).  Sweeps stop when a sweep
// applies no rotation (|g| <= tol |a_p| |a_q|, or a column of norm below
// 1e-140, which is treated as zero).  On exit the columns of A are
// u_j sigma_j, V holds the right vectors and sig the column norms.

template <bool CP>
__global__ __launch_bounds__(1024) void jacobi_svd_kernel(typename JT<CP>::T* __restrict__ A,
                                                          typename JT<CP>::T* __restrict__ V,
                                                          double* __restrict__ sig, int m, int n,
                                                          double tol, int max_sweeps,
                                                          int* __restrict__ sweeps_out) {
  typedef typename JT<CP>::T T;
  __shared__ int rot;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const int mm = (n + 1) & ~1;
  const int half = mm / 2;
  const int64_t nn = int64_t(n) * n;
  for (int64_t i = tid; i < nn; i += blockDim.x) {
    int64_t r = i % n, c = i / n;
    if (r == c) set_one(V[i]); else set_zero(V[i]);
  }
  if (tid == 0) rot = 0;
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    for (int rnd = 0; rnd < mm - 1; ++rnd) {
      for (int k = wave; k < half; k += nw) {
        int i0 = k, i1 = mm - 1 - k;
        int p = (i0 == 0) ? 0 : ((i0 - 1 + rnd) % (mm - 1)) + 1;
        int q = ((i1 - 1 + rnd) % (mm - 1)) + 1;
        if (p > q) { int t = p; p = q; q = t; }
        if (q >= n) continue;
        T* ap = A + int64_t(p) * m;
        T* aq = A + int64_t(q) * m;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int i = lane; i < m; i += 64) {
          T x = ap[i], y = aq[i];
          al += JT<CP>::abs2(x);
          be += JT<CP>::abs2(y);
          T g = mulc(conjv(x), y);
          if constexpr (CP) { gr += g.x; gi += g.y; } else { gr += g; }
        }
        for (int o = 32; o > 0; o >>= 1) {
          al += __shfl_xor(al, o);
          be += __shfl_xor(be, o);
          gr += __shfl_xor(gr, o);
          gi += __shfl_xor(gi, o);
        }
        // |g| without squaring (gr^2 can underflow), so that the phase g/|g|
        // is unimodular to rounding; columns of norm below 1e-140 count as
        // zero and pairs orthogonal to working precision are skipped
        const double na = sqrt(al), nb = sqrt(be);
        const double ag = CP ? hypot(gr, gi) : fabs(gr);
        if (!(na > 1e-140 && nb > 1e-140 && ag > tol * na * nb)) continue;
        const double th = (be - al) / (2.0 * ag);
        const double t = fabs(th) > 1e150 ? 0.5 / th
                                          : (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        if (t == 0.0) continue;
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        T ec;
        if constexpr (CP) ec = make_double2(gr / ag, -gi / ag); else ec = gr >= 0.0 ? 1.0 : -1.0;
        for (int i = lane; i < m; i += 64) {
          T x = ap[i], y = mulc(aq[i], ec);
          ap[i] = addv(scl(x, c), scl(y, -s));
          aq[i] = addv(scl(x, s), scl(y, c));
        }
        T* vp = V + int64_t(p) * n;
        T* vq = V + int64_t(q) * n;
        for (int i = lane; i < n; i += 64) {
          T x = vp[i], y = mulc(vq[i], ec);
          vp[i] = addv(scl(x, c), scl(y, -s));
          vq[i] = addv(scl(x, s), scl(y, c));
        }
        if (lane == 0) atomicAdd(&rot, 1);
      }
      __syncthreads();
    }
    const int r = rot;
    __syncthreads();
    if (r == 0) break;
    if (tid == 0) rot = 0;
    __syncthreads();
  }
  for (int j = wave; j < n; j += nw) {
    const T* aj = A + int64_t(j) * m;
    double s2 = 0.0;
    for (int i = lane; i < m; i += 64) s2 += JT<CP>::abs2(aj[i]);
    for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
    if (lane == 0) sig[j] = sqrt(s2);
  }
  if (tid == 0 && sweeps_out) *sweeps_out = sweep;
}

// Same one-sided Jacobi sweeps with each column pair handled by L lanes
// (64/L pairs per wave) and the pair's four columns (A_p, A_q, V_p, V_q)
// held in registers, R entries per lane: one load and one store per
// rotation instead of three passes over global memory.  Needs m, n <= L*R.
template <bool CP, int L, int R>
__device__ __forceinline__ void jacobi_reg_body(typename JT<CP>::T* __restrict__ A,
                                                typename JT<CP>::T* __restrict__ V,
                                                double* __restrict__ sig, int m, int n, double tol,
                                                int max_sweeps, int* __restrict__ sweeps_out) {
  typedef typename JT<CP>::T T;
  __shared__ int rot;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const int sub = lane & (L - 1);                          // lane within the pair group
  const int slot = wave * (64 / L) + lane / L;             // pair slot of this group
  const int nslots = nw * (64 / L);
  const int mm = (n + 1) & ~1;
  const int half = mm / 2;
  const int64_t nn = int64_t(n) * n;
  for (int64_t i = tid; i < nn; i += blockDim.x) {
    int64_t r = i % n, c = i / n;
    if (r == c) set_one(V[i]); else set_zero(V[i]);
  }
  if (tid == 0) rot = 0;
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    for (int rnd = 0; rnd < mm - 1; ++rnd) {
      for (int k = slot; k < half; k += nslots) {
        int i0 = k, i1 = mm - 1 - k;
        int p = (i0 == 0) ? 0 : ((i0 - 1 + rnd) % (mm - 1)) + 1;
        int q = ((i1 - 1 + rnd) % (mm - 1)) + 1;
        if (p > q) { int t = p; p = q; q = t; }
        if (q >= n) continue;
        T* ap = A + int64_t(p) * m;
        T* aq = A + int64_t(q) * m;
        T* vp = V + int64_t(p) * n;
        T* vq = V + int64_t(q) * n;
        T xa[R], xb[R], ya[R], yb[R];
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = sub + L * r;
          if (i < m) { xa[r] = ap[i]; xb[r] = aq[i]; } else { set_zero(xa[r]); set_zero(xb[r]); }
          if (i < n) { ya[r] = vp[i]; yb[r] = vq[i]; } else { set_zero(ya[r]); set_zero(yb[r]); }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          al += JT<CP>::abs2(xa[r]);
          be += JT<CP>::abs2(xb[r]);
          T g = mulc(conjv(xa[r]), xb[r]);
          if constexpr (CP) { gr += g.x; gi += g.y; } else { gr += g; }
        }
        for (int o = L / 2; o > 0; o >>= 1) {
          al += __shfl_xor(al, o);
          be += __shfl_xor(be, o);
          gr += __shfl_xor(gr, o);
          gi += __shfl_xor(gi, o);
        }
        const double na = sqrt(al), nb = sqrt(be);
        const double ag = CP ? hypot(gr, gi) : fabs(gr);
        if (!(na > 1e-140 && nb > 1e-140 && ag > tol * na * nb)) continue;
        const double th = (be - al) / (2.0 * ag);
        const double t = fabs(th) > 1e150 ? 0.5 / th
                                          : (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        if (t == 0.0) continue;
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        T ec;
        if constexpr (CP) ec = make_double2(gr / ag, -gi / ag); else ec = gr >= 0.0 ? 1.0 : -1.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = sub + L * r;
          T x = xa[r], y = mulc(xb[r], ec);
          if (i < m) {
            ap[i] = addv(scl(x, c), scl(y, -s));
            aq[i] = addv(scl(x, s), scl(y, c));
          }
          T u = ya[r], w = mulc(yb[r], ec);
          if (i < n) {
            vp[i] = addv(scl(u, c), scl(w, -s));
            vq[i] = addv(scl(u, s), scl(w, c));
          }
        }
        if (sub == 0) atomicAdd(&rot, 1);
      }
      __syncthreads();
    }
    const int r = rot;
    __syncthreads();
    if (r == 0) break;
    if (tid == 0) rot = 0;
    __syncthreads();
  }
  for (int j = wave; j < n; j += nw) {
    const T* aj = A + int64_t(j) * m;
    double s2 = 0.0;
    for (int i = lane; i < m; i += 64) s2 += JT<CP>::abs2(aj[i]);
    for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
    if (lane == 0) sig[j] = sqrt(s2);
  }
  if (tid == 0 && sweeps_out) *sweeps_out = sweep;
}

template <bool CP, int L, int R>
__global__ __launch_bounds__(1024) void jacobi_svd_reg_kernel(typename JT<CP>::T* __restrict__ A,
                                                              typename JT<CP>::T* __restrict__ V,
                                                              double* __restrict__ sig, int m, int n,
                                                              double tol, int max_sweeps,
                                                              int* __restrict__ sweeps_out) {
  jacobi_reg_body<CP, L, R>(A, V, sig, m, n, tol, max_sweeps, sweeps_out);
}

// One matrix of a batched Jacobi launch (workgroup blockIdx.x takes item
// blockIdx.x); the pointers are FP64 (real) or interleaved complex storage.
struct JacobiItem {
  double* A;
  double* V;
  double* sig;
  int* sweeps;
  int m, n;
  double tol;
};

template <bool CP, int L, int R>
__global__ __launch_bounds__(1024) void jacobi_svd_reg_batched_kernel(const JacobiItem* __restrict__ items,
                                                                      int max_sweeps) {
  typedef typename JT<CP>::T T;
  const JacobiItem it = items[blockIdx.x];
  jacobi_reg_body<CP, L, R>(reinterpret_cast<T*>(it.A), reinterpret_cast<T*>(it.V), it.sig, it.m, it.n, it.tol,
                            max_sweeps, it.sweeps);
}

// Register-resident variant for a real or complex matrix with m, n >= 2:
// 0: real m, n <= 160 (16 lanes x 10 rows per pair, up to 64 pairs at once),
// 1: real <= 256 (32 x 8), 2: real <= 512 (64 x 8), 3: complex <= 256
// (64 x 4); -1: the global-memory kernel.
static int jacobi_variant(bool cplx, int64_t m, int64_t n) {
  const int64_t mx = std::max(m, n);
  if (n < 2) return -1;
  if (!cplx && mx <= 160) return 0;
  if (!cplx && mx <= 256) return 1;
  if (!cplx && mx <= 512) return 2;
  if (cplx && mx <= 256) return 3;
  return -1;
}

// lanes per column pair of a variant
static int jacobi_lanes(int v) { return v == 0 ? 16 : v == 1 ? 32 : 64; }

// threads per workgroup for a matrix with n columns under variant v: enough
// pair slots for the ceil(n/2) pairs of a round, at most 1024 threads
static int jacobi_threads(int v, int64_t n) {
  const int half = int((n + 1) / 2);
  const int per_wave = 64 / jacobi_lanes(v);
  return 64 * std::max(1, std::min(16, int(cdiv(half, per_wave))));
}

static void launch_jacobi_batched(int v, hipStream_t st, unsigned nb, unsigned threads, const JacobiItem* items,
                                  int max_sweeps) {
  switch (v) {
    case 0:
      hipLaunchKernelGGL((jacobi_svd_reg_batched_kernel<false, 16, 10>), dim3(nb), dim3(threads), 0, st, items,
                         max_sweeps);
      break;
    case 1:
      hipLaunchKernelGGL((jacobi_svd_reg_batched_kernel<false, 32, 8>), dim3(nb), dim3(threads), 0, st, items,
                         max_sweeps);
      break;
    case 2:
      hipLaunchKernelGGL((jacobi_svd_reg_batched_kernel<false, 64, 8>), dim3(nb), dim3(threads), 0, st, items,
                         max_sweeps);
      break;
    default:
      hipLaunchKernelGGL((jacobi_svd_reg_batched_kernel<true, 64, 4>), dim3(nb), dim3(threads), 0, st, items,
                         max_sweeps);
      break;
  }
}

void jacobi_svd_batch(Ctx& c, std::vector<JacobiJob>& jobs, int max_sweeps) {
  // matrices the register kernels take run together, one launch per variant
  // (one workgroup each); the rest run one by one on the global-memory kernel
  std::vector<int> var(jobs.size());
  for (size_t i = 0; i < jobs.size(); ++i) {
    JacobiJob& j = jobs[i];
    const int64_t m = j.W->dims[0], n = j.W->dims[1];
    PEPS_CHECK(m >= n && n >= 1 && n <= 4096 && m < (1ll << 31), PS_EUNSUPPORTED,
               "jacobi_svd: need m >= n, 1 <= n <= 4096");
    PEPS_CHECK(j.V->numel() == n * n && j.V->cplx == j.W->cplx && j.sig->numel() == n && !j.sig->cplx, PS_ESHAPE,
               "jacobi_svd: buffer sizes");
    var[i] = jacobi_variant(j.W->cplx, m, n);
    if (var[i] < 0) j.sweeps = jacobi_svd(c, *j.W, *j.V, *j.sig, j.tol, max_sweeps);
  }
  auto sw = std::make_shared<Buffer>();
  sw->pool = c.pool.get();
  sw->bytes = std::max<size_t>(jobs.size(), 1) * sizeof(int);
  sw->ptr = c.pool->get(sw->bytes);
  int* swd = static_cast<int*>(sw->ptr);
  bool any = false;
  std::vector<std::shared_ptr<Buffer>> keep;
  for (int v = 0; v < 4; ++v) {
    std::vector<JacobiItem> items;
    int threads = 64, nmax = 0;
    double flops = 0.0;
    for (size_t i = 0; i < jobs.size(); ++i) {
      if (var[i] != v) continue;
      JacobiJob& j = jobs[i];
      const int64_t m = j.W->dims[0], n = j.W->dims[1];
      items.push_back({j.W->data(), j.V->data(), j.sig->data(), swd + i, int(m), int(n), j.tol});
      threads = std::max(threads, jacobi_threads(v, n));
      nmax = std::max(nmax, int(n));
      flops += (j.W->cplx ? 4.0 : 1.0) * 10.0 * 6.0 * double(m) * n * n;
    }
    if (items.empty()) continue;
    std::shared_ptr<Buffer> kb;
    const JacobiItem* dev =
        static_cast<const JacobiItem*>(c.scratch_upload(items.data(), items.size() * sizeof(JacobiItem), &kb));
    keep.push_back(kb);
    const std::string cls = Ctx::bucket("svd", nmax);
    c.count(cls, flops, 0);
    hipEvent_t ev = nullptr;
    c.tbegin(cls, &ev);
    launch_jacobi_batched(v, c.stream, unsigned(items.size()), unsigned(threads), dev, max_sweeps);
    PEPS_HIP(hipGetLastError());
    c.tend(cls, ev);
    any = true;
  }
  if (!any) return;
  std::vector<int> h(jobs.size(), 0);
  PEPS_HIP(hipMemcpyAsync(h.data(), swd, jobs.size() * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  for (size_t i = 0; i < jobs.size(); ++i)
    if (var[i] >= 0) jobs[i].sweeps = h[i];
}

int jacobi_svd(Ctx& c, Tensor& A, Tensor& V, Tensor& sig, double tol, int max_sweeps) {
  PEPS_CHECK(A.rank() == 2, PS_ESHAPE, "jacobi_svd: matrix expected");
  const int64_t m = A.dims[0], n = A.dims[1];
  PEPS_CHECK(m >= n && n >= 1 && n <= 4096 && m < (1ll << 31), PS_EUNSUPPORTED,
             "jacobi_svd: need m >= n, 1 <= n <= 4096");
  PEPS_CHECK(V.numel() == n * n && V.cplx == A.cplx && sig.numel() == n && !sig.cplx, PS_ESHAPE,
             "jacobi_svd: buffer sizes");
  auto buf = std::make_shared<Buffer>();
  buf->pool = c.pool.get();
  buf->bytes = sizeof(int);
  buf->ptr = c.pool->get(buf->bytes);
  const int half = int((n + 1) / 2);
  const int waves = std::max(1, std::min(16, half));
  hipEvent_t ev = nullptr;
  const std::string cls = c.gemm_shapes ? "svd " + std::to_string(m) + "x" + std::to_string(n) : Ctx::bucket("svd", n);
  c.tbegin(cls, &ev);
  c.count(cls, (A.cplx ? 4.0 : 1.0) * 10.0 * 6.0 * double(m) * n * n, 0);
  // register-resident variants (4 columns x R entries per lane in VGPRs,
  // within the 128-VGPR budget of 1024-thread blocks), see jacobi_variant
  const int v = jacobi_variant(A.cplx, m, n);
  const unsigned th = v >= 0 ? unsigned(jacobi_threads(v, n)) : 64u;
  int* swp = (int*)buf->ptr;
  if (v == 0)
    hipLaunchKernelGGL((jacobi_svd_reg_kernel<false, 16, 10>), dim3(1), dim3(th), 0, c.stream, A.data(), V.data(),
                       sig.data(), int(m), int(n), tol, max_sweeps, swp);
  else if (v == 1)
    hipLaunchKernelGGL((jacobi_svd_reg_kernel<false, 32, 8>), dim3(1), dim3(th), 0, c.stream, A.data(), V.data(),
                       sig.data(), int(m), int(n), tol, max_sweeps, swp);
  else if (v == 2)
    hipLaunchKernelGGL((jacobi_svd_reg_kernel<false, 64, 8>), dim3(1), dim3(th), 0, c.stream, A.data(), V.data(),
                       sig.data(), int(m), int(n), tol, max_sweeps, swp);
  else if (v == 3)
    hipLaunchKernelGGL((jacobi_svd_reg_kernel<true, 64, 4>), dim3(1), dim3(th), 0, c.stream, (double2*)A.data(),
                       (double2*)V.data(), sig.data(), int(m), int(n), tol, max_sweeps, swp);
  else if (A.cplx)
    hipLaunchKernelGGL(jacobi_svd_kernel<true>, dim3(1), dim3(64 * waves), 0, c.stream,
                       (double2*)A.data(), (double2*)V.data(), sig.data(), int(m), int(n), tol,
                       max_sweeps, (int*)buf->ptr);
  else
    hipLaunchKernelGGL(jacobi_svd_kernel<false>, dim3(1), dim3(64 * waves), 0, c.stream, A.data(),
                       V.data(), sig.data(), int(m), int(n), tol, max_sweeps, (int*)buf->ptr);
  PEPS_HIP(hipGetLastError());
  c.tend(cls, ev);
  int sw = 0;
  PEPS_HIP(hipMemcpyAsync(&sw, buf->ptr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  return sw;
}

}  // namespace peps

namespace peps {

// ════════════════════════════════════════════════════════════════════════════
//  Zero lines of a matrix; Householder QR of narrow panels
// ════════════════════════════════════════════════════════════════════════════

// rf[r] = 1 / cf[c] = 1 when row r / column c of M (m x n) has a non-zero
// entry (every writer stores the same value).
__global__ void nonzero_lines_kernel(const double* __restrict__ M, int64_t m, int64_t n, int esz,
                                     int* __restrict__ rf, int* __restrict__ cf) {
  const int64_t tot = m * n;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < tot;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool nz = M[i * esz] != 0.0 || (esz == 2 && M[i * esz + 1] != 0.0);
    if (nz) {
      rf[i % m] = 1;
      cf[i / m] = 1;
    }
  }
}

void nonzero_lines(Ctx& c, const Tensor& M, std::vector<int64_t>* rows, std::vector<int64_t>* cols) {
  PEPS_CHECK(M.rank() == 2, PS_ESHAPE, "nonzero_lines: matrix expected");
  const int64_t m = M.dims[0], n = M.dims[1];
  rows->clear();
  cols->clear();
  if (m == 0 || n == 0) return;
  auto buf = std::make_shared<Buffer>();
  buf->pool = c.pool.get();
  buf->bytes = size_t(m + n) * sizeof(int);
  buf->ptr = c.pool->get(buf->bytes);
  PEPS_HIP(hipMemsetAsync(buf->ptr, 0, buf->bytes, c.stream));
  int* rf = static_cast<int*>(buf->ptr);
  unsigned g = unsigned(std::min<int64_t>(cdiv(m * n, 256), 16384));
  hipLaunchKernelGGL(nonzero_lines_kernel, dim3(g), dim3(256), 0, c.stream, M.data(), m, n,
                     M.cplx ? 2 : 1, rf, rf + m);
  PEPS_HIP(hipGetLastError());
  std::vector<int> h(m + n);
  PEPS_HIP(hipMemcpyAsync(h.data(), buf->ptr, buf->bytes, hipMemcpyDeviceToHost, c.stream));
  PEPS_HIP(hipStreamSynchronize(c.stream));
  for (int64_t i = 0; i < m; ++i)
    if (h[i]) rows->push_back(i);
  for (int64_t j = 0; j < n; ++j)
    if (h[m + j]) cols->push_back(j);
}

// Householder QR of a narrow panel A (m x n, n <= QR_SMALL_N), one workgroup
// of QR_THREADS threads, A in global memory and overwritten LAPACK-style: R in
// the upper triangle, the reflectors v_j (v_j[j] = 1 implicit) below it,
// tau[j] their scalars.  Reflector j (Golub & Van Loan 5.1.2 in the
// convention of LAPACK's larfg): beta = -sign(Re a_jj) ||A[j:, j]||,
// tau = (beta - a_jj) / beta, v = A[j:, j] / (a_jj - beta); tau = 0 when the
// column below the diagonal is already zero and a_jj is real.  The trailing
// columns are updated with H^H = I - conj(tau) v v^H.
constexpr int QR_SMALL_N = 64;
constexpr int QR_THREADS = 256;

template <bool CP>
__global__ __launch_bounds__(QR_THREADS) void qr_panel_kernel(typename JT<CP>::T* __restrict__ A,
                                                              typename JT<CP>::T* __restrict__ tau,
                                                              int64_t m, int n) {
  typedef typename JT<CP>::T T;
  __shared__ double part[QR_THREADS / 64];
  __shared__ T s_tau, s_scale;
  __shared__ int s_skip;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = QR_THREADS / 64;
  const int k = int(std::min<int64_t>(m, n));
  for (int j = 0; j < k; ++j) {
    T* aj = A + int64_t(j) * m;
    // sigma = sum_{i>j} |a_ij|^2
    double sg = 0.0;
    for (int64_t i = j + 1 + tid; i < m; i += QR_THREADS) sg += JT<CP>::abs2(aj[i]);
    for (int o = 32; o > 0; o >>= 1) sg += __shfl_xor(sg, o);
    if (lane == 0) part[wave] = sg;
    __syncthreads();
    if (tid == 0) {
      double sigma = 0.0;
      for (int w = 0; w < nw; ++w) sigma += part[w];
      const T alpha = aj[j];
      const double ar = JT<CP>::re(alpha);
      double ai = 0.0;
      if constexpr (CP) ai = alpha.y;
      if (sigma == 0.0 && ai == 0.0) {
        s_skip = 1;
        set_zero(s_tau);
      } else {
        const double nrm = sqrt(ar * ar + ai * ai + sigma);
        const double beta = ar >= 0.0 ? -nrm : nrm;
        // tau = (beta - alpha) / beta, scale = 1 / (alpha - beta)
        if constexpr (CP) {
          s_tau = make_double2((beta - ar) / beta, -ai / beta);
          const double dr = ar - beta, di = ai, dd = dr * dr + di * di;
          s_scale = make_double2(dr / dd, -di / dd);
          aj[j] = make_double2(beta, 0.0);
        } else {
          s_tau = (beta - ar) / beta;
          s_scale = 1.0 / (ar - beta);
          aj[j] = beta;
        }
        s_skip = 0;
      }
      tau[j] = s_tau;
    }
    __syncthreads();
    if (s_skip) continue;  // uniform: H_j = I
    const T scale = s_scale;
    for (int64_t i = j + 1 + tid; i < m; i += QR_THREADS) aj[i] = mulc(aj[i], scale);
    __syncthreads();
    // trailing columns: w = v^H a_c ; a_c -= conj(tau) v w
    const T ctau = conjv(s_tau);
    for (int cix = j + 1 + wave; cix < n; cix += nw) {
      T* ac = A + int64_t(cix) * m;
      double wr = 0.0, wi = 0.0;
      for (int64_t i = j + lane; i < m; i += 64) {
        T v = i == j ? T{} : aj[i];
        if (i == j) set_one(v);
        const T p = mulc(conjv(v), ac[i]);
        if constexpr (CP) { wr += p.x; wi += p.y; } else { wr += p; }
      }
      for (int o = 32; o > 0; o >>= 1) {
        wr += __shfl_xor(wr, o);
        wi += __shfl_xor(wi, o);
      }
      T w;
      if constexpr (CP) w = make_double2(wr, wi); else w = wr;
      const T f = mulc(ctau, w);
      for (int64_t i = j + lane; i < m; i += 64) {
        T v = i == j ? T{} : aj[i];
        if (i == j) set_one(v);
        const T d = mulc(v, f);
        if constexpr (CP) ac[i] = make_double2(ac[i].x - d.x, ac[i].y - d.y); else ac[i] = ac[i] - d;
      }
    }
    __syncthreads();
  }
}

// Q (m x k) = H_0 H_1 ... H_{k-1} [I; 0] from the reflectors left in A by
// qr_panel_kernel (accumulated backwards, H_j = I - tau_j v_j v_j^H).
template <bool CP>
__global__ __launch_bounds__(QR_THREADS) void qr_form_q_kernel(const typename JT<CP>::T* __restrict__ A,
                                                               const typename JT<CP>::T* __restrict__ tau,
                                                               typename JT<CP>::T* __restrict__ Q,
                                                               int64_t m, int k) {
  typedef typename JT<CP>::T T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = QR_THREADS / 64;
  const int64_t mk = m * k;
  for (int64_t i = tid; i < mk; i += QR_THREADS) {
    const int64_t r = i % m, cc = i / m;
    if (r == cc) set_one(Q[i]); else set_zero(Q[i]);
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const T tj = tau[j];
    bool zero_tau;
    if constexpr (CP) zero_tau = tj.x == 0.0 && tj.y == 0.0; else zero_tau = tj == 0.0;
    if (zero_tau) continue;  // uniform
    const T* aj = A + int64_t(j) * m;
    for (int cix = j + wave; cix < k; cix += nw) {
      T* qc = Q + int64_t(cix) * m;
      double wr = 0.0, wi = 0.0;
      for (int64_t i = j + lane; i < m; i += 64) {
        T v = i == j ? T{} : aj[i];
        if (i == j) set_one(v);
        const T p = mulc(conjv(v), qc[i]);
        if constexpr (CP) { wr += p.x; wi += p.y; } else { wr += p; }
      }
      for (int o = 32; o > 0; o >>= 1) {
        wr += __shfl_xor(wr, o);
        wi += __shfl_xor(wi, o);
      }
      T w;
      if constexpr (CP) w = make_double2(wr, wi); else w = wr;
      const T f = mulc(tj, w);
      for (int64_t i = j + lane; i < m; i += 64) {
        T v = i == j ? T{} : aj[i];
        if (i == j) set_one(v);
        const T d = mulc(v, f);
        if constexpr (CP) qc[i] = make_double2(qc[i].x - d.x, qc[i].y - d.y); else qc[i] = qc[i] - d;
      }
    }
    __syncthreads();
  }
}

int qr_small_max_cols() { return QR_SMALL_N; }

// V (m x b): the unit lower trapezoid of a factored panel P (m x b, real):
// V(i, j) = P(i, j) below the diagonal, 1 on it, 0 above.
__global__ void householder_v_kernel(const double* __restrict__ P, double* __restrict__ V, int64_t m, int b) {
  const int64_t n = m * b;
  for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = q % m, j = q / m;
    V[q] = i > j ? P[q] : (i == j ? 1.0 : 0.0);
  }
}

// T (b x b, upper triangular) of the compact WY form H_0 H_1 ... H_{b-1} =
// I - V T V^T (LAPACK dlarft, forward, columnwise) from G = V^T V and the
// reflector scalars: T(i, i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).
// One workgroup of QR_SMALL_N threads; T is built in LDS column by column.
__global__ __launch_bounds__(QR_SMALL_N) void householder_t_kernel(const double* __restrict__ G,
                                                                   const double* __restrict__ tau,
                                                                   double* __restrict__ T, int b) {
  __shared__ double Ts[QR_SMALL_N * QR_SMALL_N];
  __shared__ double Gs[QR_SMALL_N * QR_SMALL_N];
  const int r = threadIdx.x;
  for (int q = r; q < QR_SMALL_N * QR_SMALL_N; q += QR_SMALL_N) {
    Ts[q] = 0.0;
    Gs[q] = q < b * b ? G[q] : 0.0;  // G (b x b) column-major, leading dimension b
  }
  __syncthreads();
  for (int i = 0; i < b; ++i) {
    double w = 0.0;
    if (r < i)
      for (int s2 = r; s2 < i; ++s2) w += Ts[r + s2 * QR_SMALL_N] * Gs[s2 + i * b];
    // column i is written only after every thread has read columns < i
    __syncthreads();
    if (r < i) Ts[r + i * QR_SMALL_N] = -tau[i] * w;
    if (r == i) Ts[i + i * QR_SMALL_N] = tau[i];
    __syncthreads();
  }
  for (int q = r; q < b * b; q += QR_SMALL_N) T[q] = Ts[(q % b) + (q / b) * QR_SMALL_N];
}

// dst(i, j) = src(i, j) for i < rows, j < cols, both column-major with their
// own leading dimensions (esz doubles per element).
__global__ void copy_block_kernel(double* __restrict__ dst, int64_t ldd, const double* __restrict__ src,
                                  int64_t lds, int64_t rows, int64_t cols, int esz) {
  const int64_t n = rows * cols * esz;
  for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = q % esz, ij = q / esz, i = ij % rows, j = ij / rows;
    dst[(i + j * ldd) * esz + e] = src[(i + j * lds) * esz + e];
  }
}

static void copy_block(Ctx& c, double* dst, int64_t ldd, const double* src, int64_t lds, int64_t rows,
                       int64_t cols, bool cplx) {
  if (rows <= 0 || cols <= 0) return;
  const int esz = cplx ? 2 : 1;
  const unsigned g = unsigned(std::min<int64_t>(cdiv(rows * cols * esz, 256), 16384));
  hipLaunchKernelGGL(copy_block_kernel, dim3(g), dim3(256), 0, c.stream, dst, ldd, src, lds, rows, cols, esz);
  PEPS_HIP(hipGetLastError());
}

Tensor sub_block(Ctx& c, const Tensor& M, int64_t r0, int64_t c0, int64_t rows, int64_t cols) {
  PEPS_CHECK(M.rank() == 2 && r0 >= 0 && c0 >= 0 && r0 + rows <= M.dims[0] && c0 + cols <= M.dims[1], PS_ESHAPE,
             "sub_block: range out of bounds");
  Tensor B = c.alloc({rows, cols}, M.cplx);
  const int esz = M.cplx ? 2 : 1;
  copy_block(c, B.data(), rows, M.data() + (r0 + c0 * M.dims[0]) * esz, M.dims[0], rows, cols, M.cplx);
  return B;
}

void put_block(Ctx& c, Tensor& M, int64_t r0, int64_t c0, const Tensor& B) {
  PEPS_CHECK(M.rank() == 2 && B.rank() == 2 && M.cplx == B.cplx && r0 >= 0 && c0 >= 0 &&
                 r0 + B.dims[0] <= M.dims[0] && c0 + B.dims[1] <= M.dims[1],
             PS_ESHAPE, "put_block: range out of bounds");
  const int esz = M.cplx ? 2 : 1;
  copy_block(c, M.data() + (r0 + c0 * M.dims[0]) * esz, M.dims[0], B.data(), B.dims[0], B.dims[0], B.dims[1],
             M.cplx);
}

Tensor householder_v(Ctx& c, const Tensor& P) {
  PEPS_CHECK(P.rank() == 2 && !P.cplx && P.dims[1] <= QR_SMALL_N, PS_ESHAPE, "householder_v: real m x b panel");
  const int64_t m = P.dims[0];
  const int b = int(P.dims[1]);
  Tensor V = c.alloc({m, int64_t(b)}, false);
  if (m * b == 0) return V;
  const unsigned g = unsigned(std::min<int64_t>(cdiv(m * b, 256), 16384));
  hipLaunchKernelGGL(householder_v_kernel, dim3(g), dim3(256), 0, c.stream, P.data(), V.data(), m, b);
  PEPS_HIP(hipGetLastError());
  return V;
}

Tensor householder_t(Ctx& c, const Tensor& G, const Tensor& tau) {
  const int b = int(G.dims[0]);
  PEPS_CHECK(G.rank() == 2 && G.dims[1] == b && b <= QR_SMALL_N && !G.cplx && tau.numel() >= b, PS_ESHAPE,
             "householder_t: real b x b Gram matrix, b <= 64");
  Tensor T = c.alloc({int64_t(b), int64_t(b)}, false);
  if (b == 0) return T;
  hipLaunchKernelGGL(householder_t_kernel, dim3(1), dim3(QR_SMALL_N), 0, c.stream, G.data(), tau.data(), T.data(),
                     b);
  PEPS_HIP(hipGetLastError());
  return T;
}

void qr_panel(Ctx& c, Tensor& W, Tensor& tau) {
  PEPS_CHECK(W.rank() == 2 && W.dims[1] <= QR_SMALL_N && W.dims[0] < (1ll << 40), PS_EUNSUPPORTED,
             "qr_panel: at most 64 columns");
  const int64_t m = W.dims[0];
  const int n = int(W.dims[1]);
  PEPS_CHECK(tau.numel() >= std::min<int64_t>(m, n) && tau.cplx == W.cplx, PS_ESHAPE, "qr_panel: tau");
  if (m == 0 || n == 0) return;
  if (W.cplx)
    hipLaunchKernelGGL(qr_panel_kernel<true>, dim3(1), dim3(QR_THREADS), 0, c.stream, (double2*)W.data(),
                       (double2*)tau.data(), m, n);
  else
    hipLaunchKernelGGL(qr_panel_kernel<false>, dim3(1), dim3(QR_THREADS), 0, c.stream, W.data(),
                       tau.data(), m, n);
  PEPS_HIP(hipGetLastError());
}

void qr_form_q(Ctx& c, const Tensor& W, const Tensor& tau, Tensor& Q) {
  const int64_t m = W.dims[0];
  const int k = int(std::min<int64_t>(m, W.dims[1]));
  PEPS_CHECK(Q.numel() == m * k && Q.cplx == W.cplx && k <= QR_SMALL_N, PS_ESHAPE, "qr_form_q: Q");
  if (m == 0 || k == 0) return;
  if (W.cplx)
    hipLaunchKernelGGL(qr_form_q_kernel<true>, dim3(1), dim3(QR_THREADS), 0, c.stream,
                       (const double2*)W.data(), (const double2*)tau.data(), (double2*)Q.data(), m, k);
  else
    hipLaunchKernelGGL(qr_form_q_kernel<false>, dim3(1), dim3(QR_THREADS), 0, c.stream, W.data(),
                       tau.data(), Q.data(), m, k);
  PEPS_HIP(hipGetLastError());
}

}  // namespace peps

namespace peps {

// out[j] = ||M[:, j]||^2 for a column-major matrix M (m x n): one wave per
// column, an XOR-butterfly wave sum.
__global__ void col_norm2_kernel(const double* __restrict__ M, int64_t m, int64_t n, int esz,
                                 double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t col = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (col >= n) return;
  const double* p = M + col * m * esz;
  double s = 0.0;
  for (int64_t i = lane; i < m * esz; i += 64) s += p[i] * p[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[col] = s;
}

std::vector<double> col_norm2(Ctx& c, const Tensor& M) {
  PEPS_CHECK(M.rank() == 2, PS_ESHAPE, "col_norm2: matrix expected");
  const int64_t m = M.dims[0], n = M.dims[1];
  std::vector<double> h(n, 0.0);
  if (n == 0 || m == 0) return h;
  Tensor out = c.alloc({n}, false);
  const unsigned blocks = unsigned(cdiv(n, 4));
  hipLaunchKernelGGL(col_norm2_kernel, dim3(blocks), dim3(256), 0, c.stream, M.data(), m, n,
                     M.cplx ? 2 : 1, out.data());
  PEPS_HIP(hipGetLastError());
  return c.to_host(out);
}

}  // namespace peps

namespace peps {

// ════════════════════════════════════════════════════════════════════════════
//  Small-MLP kernels (learned ALS warm start): batch in columns, column-major
// ════════════════════════════════════════════════════════════════════════════

// Z (n x B) += b (per row); ReLU when relu != 0.
__global__ void bias_act_kernel(double* __restrict__ Z, const double* __restrict__ b, int64_t n, int64_t B,
                                int relu) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n * B; i += int64_t(gridDim.x) * blockDim.x) {
    double z = Z[i] + b[i % n];
    Z[i] = relu ? (z > 0.0 ? z : 0.0) : z;
  }
}

// dH (n x B) *= (H > 0)
__global__ void relu_grad_kernel(double* __restrict__ dH, const double* __restrict__ H, int64_t nB) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nB; i += int64_t(gridDim.x) * blockDim.x)
    if (!(H[i] > 0.0)) dH[i] = 0.0;
}

// db[i] = sum_j dZ[i + j n]: one wave per row, fixed-order lane sums
__global__ void row_sum_kernel(const double* __restrict__ dZ, double* __restrict__ db, int64_t n, int64_t B) {
  const int lane = threadIdx.x & 63;
  const int64_t row = blockIdx.x * int64_t(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  double s = 0.0;
  for (int64_t j = lane; j < B; j += 64) s += dZ[row + j * n];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) db[row] = s;
}

// D = (Y - T) * scale
__global__ void scaled_diff_kernel(const double* __restrict__ Y, const double* __restrict__ T, double* __restrict__ D,
                                   int64_t n, double scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    D[i] = (Y[i] - T[i]) * scale;
}

// Adam (Kingma & Ba) with bias correction c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void adam_kernel(double* __restrict__ p, const double* __restrict__ g, double* __restrict__ m,
                            double* __restrict__ v, int64_t n, double lr, double b1, double b2, double eps, double c1,
                            double c2) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double gi = g[i];
    const double mi = b1 * m[i] + (1.0 - b1) * gi;
    const double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrt(vi / c2) + eps);
  }
}

static unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>(cdiv(n, 256), 8192))); }

void mlp_bias_act(Ctx& c, Tensor& Z, const Tensor& b, bool relu) {
  const int64_t n = Z.dims[0], B = Z.dims[1];
  PEPS_CHECK(!Z.cplx && b.numel() == n, PS_ESHAPE, "mlp_bias_act: shapes");
  hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for(n * B)), dim3(256), 0, c.stream, Z.data(), b.data(), n, B,
                     relu ? 1 : 0);
  PEPS_HIP(hipGetLastError());
}

void mlp_relu_grad(Ctx& c, Tensor& dH, const Tensor& H) {
  PEPS_CHECK(dH.numel() == H.numel(), PS_ESHAPE, "mlp_relu_grad: shapes");
  hipLaunchKernelGGL(relu_grad_kernel, dim3(grid_for(H.numel())), dim3(256), 0, c.stream, dH.data(), H.data(),
                     H.numel());
  PEPS_HIP(hipGetLastError());
}

void mlp_row_sum(Ctx& c, const Tensor& dZ, Tensor& db) {
  const int64_t n = dZ.dims[0], B = dZ.dims[1];
  PEPS_CHECK(db.numel() == n, PS_ESHAPE, "mlp_row_sum: shapes");
  hipLaunchKernelGGL(row_sum_kernel, dim3(unsigned(cdiv(n, 4))), dim3(256), 0, c.stream, dZ.data(), db.data(), n, B);
  PEPS_HIP(hipGetLastError());
}

void mlp_scaled_diff(Ctx& c, const Tensor& Y, const Tensor& T, Tensor& D, double scale) {
  PEPS_CHECK(Y.numel() == T.numel() && D.numel() == Y.numel(), PS_ESHAPE, "mlp_scaled_diff: shapes");
  hipLaunchKernelGGL(scaled_diff_kernel, dim3(grid_for(Y.numel())), dim3(256), 0, c.stream, Y.data(), T.data(),
                     D.data(), Y.numel(), scale);
  PEPS_HIP(hipGetLastError());
}

void mlp_adam(Ctx& c, Tensor& p, const Tensor& g, Tensor& m, Tensor& v, double lr, double b1, double b2, double eps,
              int64_t t) {
  PEPS_CHECK(p.numel() == g.numel() && m.numel() == p.numel() && v.numel() == p.numel(), PS_ESHAPE,
             "mlp_adam: shapes");
  const double c1 = 1.0 - std::pow(b1, double(t)), c2 = 1.0 - std::pow(b2, double(t));
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(p.numel())), dim3(256), 0, c.stream, p.data(), g.data(), m.data(),
                     v.data(), p.numel(), lr, b1, b2, eps, c1, c2);
  PEPS_HIP(hipGetLastError());
}

}  // namespace peps

namespace peps {

// _dl_open (finite_peps_full_update.jl:97-109): the double layer with the
// physical indices left open, E[p1, p2, (l1 l2), (r1 r2), (u1 u2), (d1 d2)] =
// T[p1, l1, r1, u1, d1] conj(T[p2, l2, r2, u2, d2]), the bra index (2) fast
// in every merged pair.  One output element per thread (grid-stride): an
// outer product, bound by the HBM write of E.
template <bool CP>
__global__ void dl_open_kernel(const typename JT<CP>::T* __restrict__ T, typename JT<CP>::T* __restrict__ E, int dp,
                               int Dl, int Dr, int Du, int Dd, int64_t total) {
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    int64_t q = idx;
    const int p1 = int(q % dp); q /= dp;
    const int p2 = int(q % dp); q /= dp;
    const int l2 = int(q % Dl); q /= Dl;
    const int l1 = int(q % Dl); q /= Dl;
    const int r2 = int(q % Dr); q /= Dr;
    const int r1 = int(q % Dr); q /= Dr;
    const int u2 = int(q % Du); q /= Du;
    const int u1 = int(q % Du); q /= Du;
    const int d2 = int(q % Dd); q /= Dd;
    const int d1 = int(q);
    const int64_t a = p1 + int64_t(dp) * (l1 + int64_t(Dl) * (r1 + int64_t(Dr) * (u1 + int64_t(Du) * d1)));
    const int64_t b = p2 + int64_t(dp) * (l2 + int64_t(Dl) * (r2 + int64_t(Dr) * (u2 + int64_t(Du) * d2)));
    E[idx] = mulc(T[a], conjv(T[b]));
  }
}

Tensor dl_open(Ctx& c, const Tensor& T) {
  PEPS_CHECK(T.rank() == 5, PS_ESHAPE, "dl_open: (p,l,r,u,d) tensor expected");
  const int64_t dp = T.dims[0], Dl = T.dims[1], Dr = T.dims[2], Du = T.dims[3], Dd = T.dims[4];
  Tensor E = c.alloc({dp, dp, Dl * Dl, Dr * Dr, Du * Du, Dd * Dd}, T.cplx);
  const int64_t total = E.numel();
  if (total == 0) return E;
  hipEvent_t ev = nullptr;
  c.tbegin("elementwise", &ev);
  c.count("elementwise", 0, double(E.bytes()) + double(T.bytes()));
  const unsigned blocks = unsigned(std::min<int64_t>(cdiv(total, 256), 4096));
  if (T.cplx)
    hipLaunchKernelGGL(dl_open_kernel<true>, dim3(blocks), dim3(256), 0, c.stream, (const double2*)T.data(),
                       (double2*)E.data(), int(dp), int(Dl), int(Dr), int(Du), int(Dd), total);
  else
    hipLaunchKernelGGL(dl_open_kernel<false>, dim3(blocks), dim3(256), 0, c.stream, T.data(), E.data(), int(dp),
                       int(Dl), int(Dr), int(Du), int(Dd), total);
  PEPS_HIP(hipGetLastError());
  c.tend("elementwise", ev);
  return E;
}

}  // namespace peps
