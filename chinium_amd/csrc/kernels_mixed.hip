// Mixed-precision form: the two dense contractions on split-BF16 MFMA.
//
// Every FP32 operand x is held as a pair of bf16 numbers, hi = bf16(x) and
// lo = bf16(x - hi), so that hi + lo carries 16 significant bits. A product
// x*y is formed as hi*hi' + hi*lo' + lo*hi' on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation (the dropped lo*lo' term is below 2^-16 |x y|). The
// functional, the E_xc / N_e sums and the final V accumulation stay FP64.
//
// The AO tables are cached at Grid::getAO in this split form (as the FP64
// cached path keeps them in FP64, src/Grid/Grid.cpp:22-33 caches them in the
// reference), 32 KB + 8 KB per component of the transposed copy per slot:
//   AOg [c=0..3][g=0..63][hl][mu=0..31]  (c = phi, d/dx, d/dy, d/dz; K = mu contiguous)
//   AOt [c=0..nT-1][hl][mu=0..31][g=0..63]  (K = g contiguous; nT = 1, or 4 for meta-GGA)
// since the MFMA wants the contraction index contiguous in both operands: mu
// for the density X = Phi D, g for Vxc M = Phi^T B.
// The MFMA operand layouts used below: lane l, element e of the 8-vector holds
//   A[i = l % 16][k = 8 (l / 16) + e],  B[k = 8 (l / 16) + e][j = l % 16],
// and result register r of lane l is C[4 (l / 16) + r][l % 16].
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace chx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int kMxRow = 2 * kTile;             // one AOg row: 32 hi then 32 lo
constexpr int kMxG = 4 * kBatch * kMxRow;     // AOg elements per slot
constexpr int kMxT = 2 * kTile * kBatch;      // AOt / Bt elements per component
constexpr int kMxSup = 4;                     // tiles per Vxc super-tile (128 AOs)
constexpr int kLdsRow = 5;                    // uint4 per LDS row: 4 data + 1 pad (no bank conflicts)

__device__ __forceinline__ u16 f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ void split_bf(float x, u16& hi, u16& lo) {
  hi = f2bf(x);
  lo = f2bf(x - bf2f(hi));
}
__device__ __forceinline__ bf16x8 as_bf(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ f32x4 mfma3(uint4 ah, uint4 al, uint4 bh, uint4 bl, f32x4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(al), as_bf(bh), c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(ah), as_bf(bl), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(ah), as_bf(bh), c, 0, 0, 0);
}
// value of the pair stored at p (hi) and p + off (lo)
__device__ __forceinline__ float pair_at(const u16* p, int off) { return bf2f(p[0]) + bf2f(p[off]); }

// ---------------------------------------------------------------------------
// AO cache in split form: one block per slot. Values are evaluated in FP64,
// rounded to FP32 and split.
__global__ __launch_bounds__(256) void k_ao_cache_mx(const int* __restrict__ slot_batch, const int* __restrict__ slot_tile,
                                                     const int* __restrict__ tile_nao, const double* __restrict__ bx,
                                                     const double* __restrict__ by, const double* __restrict__ bz,
                                                     BasisView bv, int nT, int sstride, u16* __restrict__ cache) {
  __shared__ float E[4][kBatch][kTile + 1];
  __shared__ double px[kBatch], py[kBatch], pz[kBatch];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int b = slot_batch[s], t = slot_tile[s];
  for (int e = tid; e < 4 * kBatch * (kTile + 1); e += 256) (&E[0][0][0])[e] = 0.0f;
  if (tid < kBatch) {
    px[tid] = bx[b * kBatch + tid];
    py[tid] = by[b * kBatch + tid];
    pz[tid] = bz[b * kBatch + tid];
  }
  __syncthreads();
  (void)tile_nao;  // padding AOs stay zero from the clear above
  const int s0 = bv.tile_seg_off[t], s1 = bv.tile_seg_off[t + 1];
  for (int item = tid; item < (s1 - s0) * kBatch; item += 256) {
    const int g = item % kBatch;
    const int4 sg = bv.seg[s0 + item / kBatch];
    double v[7], gx[7], gy[7], gz[7];
    shell_eval<true>(bv, sg.x, px[g], py[g], pz[g], v, gx, gy, gz);
#pragma unroll
    for (int m = 0; m < 7; ++m)
      if (m < sg.z) {
        const int col = sg.w + m;
        E[0][g][col] = (float)v[sg.y + m];
        E[1][g][col] = (float)gx[sg.y + m];
        E[2][g][col] = (float)gy[sg.y + m];
        E[3][g][col] = (float)gz[sg.y + m];
      }
  }
  __syncthreads();
  u16* out = cache + (size_t)s * sstride;
  for (int e = tid; e < 4 * kBatch * kTile; e += 256) {
    const int c = e / (kBatch * kTile), g = (e / kTile) % kBatch, mu = e % kTile;
    u16 hi, lo;
    split_bf(E[c][g][mu], hi, lo);
    out[(c * kBatch + g) * kMxRow + mu] = hi;
    out[(c * kBatch + g) * kMxRow + kTile + mu] = lo;
  }
  for (int e = tid; e < nT * kTile * kBatch; e += 256) {
    const int c = e / (kTile * kBatch), mu = (e / kBatch) % kTile, g = e % kBatch;
    u16 hi, lo;
    split_bf(E[c][g][mu], hi, lo);
    out[kMxG + c * kMxT + mu * kBatch + g] = hi;
    out[kMxG + c * kMxT + kTile * kBatch + mu * kBatch + g] = lo;
  }
}

void launch_ao_cache_mx(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* tile_nao,
                        const double* bx, const double* by, const double* bz, BasisView bv, int nT, int sstride,
                        unsigned short* cache) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_ao_cache_mx, dim3(nslots), dim3(256), 0, s, slot_batch, slot_tile, tile_nao, bx, by, bz, bv, nT,
                     sstride, cache);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// D (reference order, column-major) -> internal order, split: Ds[hl][mu][nu].
__global__ void k_split_d(int n, int npad, const double* __restrict__ D, const int* __restrict__ int2ref,
                          u16* __restrict__ Ds) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long np2 = (long)npad * npad;
  if (idx >= np2) return;
  const int r = (int)(idx / npad), c = (int)(idx % npad);
  const int rr = int2ref[r], cc = int2ref[c];
  const double v = (rr >= 0 && cc >= 0) ? D[(size_t)cc * n + rr] : 0.0;
  u16 hi, lo;
  split_bf((float)v, hi, lo);
  Ds[idx] = hi;
  Ds[np2 + idx] = lo;
}

void launch_split_d(hipStream_t s, int n, int npad, const double* D, const int* int2ref, unsigned short* Ds) {
  const long tot = (long)npad * npad;
  hipLaunchKernelGGL(k_split_d, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, npad, D, int2ref, Ds);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Density. Block = (batch, group of <= 4 alive tiles as output columns), 4
// waves in a 2 x 2 grid: wave (wr, wc) owns rows g = 32 wr..+31 and columns
// 64 wc..+63 of X = Phi_b D_b (64 x 128). K loop over the batch's alive tiles
// (32 AOs = one MFMA K step): the A operand (phi rows of the wave) comes
// straight from the cache into registers; the D block (128 nu x 32 mu, hi and
// lo) is staged in LDS, double-buffered, the next step prefetched into
// registers while the MFMAs of the current one run.
struct DensMxSmem {
  uint4 B[2][2][kGroup * kTile][kLdsRow];  // [buf][hl][nu][mu chunk]
  float red[2][4][kBatch];
};

// MODE 0: A = phi; epilogue rho and grad rho / 2 (4 components into part).
// MODE r = 1..3: A = d phi / d x_r; epilogue sum X_r d_r phi (tau partial of direction r).
template <int MODE>
__global__ __launch_bounds__(256) void k_density_mx(const int* __restrict__ grp_batch, const int* __restrict__ grp_tiles,
                                                    const int* __restrict__ grp_slots, const int* __restrict__ grp_ntiles,
                                                    const int* __restrict__ bt_off, const int* __restrict__ bt_list,
                                                    const u16* __restrict__ cache, int sstride,
                                                    const u16* __restrict__ Ds, int npad, double* __restrict__ part,
                                                    int grp_stride) {
  constexpr int NOUT = MODE == 0 ? 4 : 1;
  __shared__ DensMxSmem sm;
  const int grp = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  const int b = grp_batch[grp];
  const int nt = grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = grp_tiles[grp * kGroup + q];
  const int it0 = bt_off[b], it1 = bt_off[b + 1];
  const size_t np2 = (size_t)npad * npad;

  uint4 an[2][2], ac[2][2], rb[4];  // A next / current [row block][hl]; D staging
  auto load = [&](int it) {
    const u16* src = cache + (size_t)it * sstride + (size_t)MODE * kBatch * kMxRow;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const u16* row = src + (32 * wr + 16 * r + lrow) * kMxRow + 8 * lk;
      an[r][0] = *reinterpret_cast<const uint4*>(row);
      an[r][1] = *reinterpret_cast<const uint4*>(row + kTile);
    }
    const int ti = bt_list[it];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      const int hl = e >> 9, n = (e >> 2) & 127, qc = e & 3, q = n >> 5;
      rb[j] = q < nt ? *reinterpret_cast<const uint4*>(Ds + hl * np2 + (size_t)(jt[q] * kTile + (n & 31)) * npad +
                                                       ti * kTile + 8 * qc)
                     : uint4{0u, 0u, 0u, 0u};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      sm.B[buf][e >> 9][(e >> 2) & 127][e & 3] = rb[j];
    }
  };

  f32x4 acc[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (it0 < it1) load(it0);
  int buf = 0;
  for (int it = it0; it < it1; ++it) {
    store(buf);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      ac[r][0] = an[r][0];
      ac[r][1] = an[r][1];
    }
    __syncthreads();
    if (it + 1 < it1) load(it + 1);
    // columns of absent tiles hold zeros in LDS, so every wave runs all blocks
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int n = 64 * wc + 16 * cb + lrow;
      const uint4 bh = sm.B[buf][0][n][lk], bl = sm.B[buf][1][n][lk];
#pragma unroll
      for (int r = 0; r < 2; ++r) acc[r][cb] = mfma3(ac[r][0], ac[r][1], bh, bl, acc[r][cb]);
    }
    buf ^= 1;
  }

  // Epilogue: X[g = 32 wr + 16 r + 4 lk + i][n = 64 wc + 16 cb + lrow] against phi / grad phi of tile q = n / 32.
  float s[NOUT][2][4];
#pragma unroll
  for (int c = 0; c < NOUT; ++c)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int i = 0; i < 4; ++i) s[c][r][i] = 0.f;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int q = 2 * wc + (cb >> 1);  // wave-uniform
    if (q < nt) {
      const u16* src = cache + (size_t)grp_slots[grp * kGroup + q] * sstride;
      const int col = 16 * (cb & 1) + lrow;
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int g = 32 * wr + 16 * r + 4 * lk + i;
          const float x = acc[r][cb][i];
#pragma unroll
          for (int c = 0; c < NOUT; ++c)
            s[c][r][i] += x * pair_at(src + ((MODE == 0 ? c : MODE) * kBatch + g) * kMxRow + col, kTile);
        }
    }
  }
#pragma unroll
  for (int c = 0; c < NOUT; ++c)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = s[c][r][i];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        s[c][r][i] = v;
      }
  if (lrow == 0) {
#pragma unroll
    for (int c = 0; c < NOUT; ++c)
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) sm.red[wc][c][32 * wr + 16 * r + 4 * lk + i] = s[c][r][i];
  }
  __syncthreads();
  if (tid < NOUT * kBatch) {
    const int c = tid / kBatch, g = tid % kBatch;
    part[(size_t)grp * grp_stride + c * kBatch + g] = (double)sm.red[0][c][g] + (double)sm.red[1][c][g];
  }
}

void launch_density_mx(hipStream_t s, int mode, int ngroups, const int* grp_batch, const int* grp_tiles,
                       const int* grp_slots, const int* grp_ntiles, const int* bt_off, const int* bt_list,
                       const unsigned short* cache, int sstride, const unsigned short* Ds, int npad, double* part) {
  if (ngroups == 0) return;
  const int stride = (mode == 0 ? 4 : 3) * kBatch;  // part: [group][4][64]; tau partials: [group][3][64]
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_density_mx<0>, dim3(ngroups), dim3(256), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, sstride, Ds, npad, part, stride); break;
    case 1: hipLaunchKernelGGL(k_density_mx<1>, dim3(ngroups), dim3(256), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, sstride, Ds, npad, part, stride); break;
    case 2: hipLaunchKernelGGL(k_density_mx<2>, dim3(ngroups), dim3(256), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, sstride, Ds, npad, part, stride); break;
    default: hipLaunchKernelGGL(k_density_mx<3>, dim3(ngroups), dim3(256), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, sstride, Ds, npad, part, stride); break;
  }
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// B^T in split form for every slot of a batch (one block per batch):
// Bt[slot][0][hl][nu][g] = split(a phi + c . grad phi); meta-GGA (NC = 4) adds
// Bt[slot][r][hl][nu][g] = split(t d_r phi), t = w v_tau / 4.
template <int NC>
__global__ __launch_bounds__(256) void k_make_b_mx(const int* __restrict__ bt_off, const double* __restrict__ coef,
                                                   const u16* __restrict__ cache, int sstride, u16* __restrict__ bt) {
  __shared__ float ca[kBatch], cx[kBatch], cy[kBatch], cz[kBatch], ct[kBatch];
  const int b = blockIdx.x, tid = threadIdx.x;
  if (tid < kBatch) {
    const size_t slot = (size_t)b * kBatch + tid;
    ca[tid] = (float)coef[kCoef * slot];
    cx[tid] = (float)coef[kCoef * slot + 1];
    cy[tid] = (float)coef[kCoef * slot + 2];
    cz[tid] = (float)coef[kCoef * slot + 3];
    ct[tid] = (float)coef[kCoef * slot + 4];
  }
  __syncthreads();
  const int g = tid & 63, m8 = tid >> 6;  // point, 8-AO chunk
  for (int it = bt_off[b]; it < bt_off[b + 1]; ++it) {
    const u16* src = cache + (size_t)it * sstride;
    float v[4][8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const u16* row = src + (c * kBatch + g) * kMxRow + 8 * m8;
      const uint4 hv = *reinterpret_cast<const uint4*>(row), lv = *reinterpret_cast<const uint4*>(row + kTile);
      const u16* h = reinterpret_cast<const u16*>(&hv);
      const u16* l = reinterpret_cast<const u16*>(&lv);
#pragma unroll
      for (int e = 0; e < 8; ++e) v[c][e] = bf2f(h[e]) + bf2f(l[e]);
    }
    u16* dst = bt + (size_t)it * NC * kMxT;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int o = (8 * m8 + e) * kBatch + g;
      u16 hi, lo;
      split_bf(ca[g] * v[0][e] + cx[g] * v[1][e] + cy[g] * v[2][e] + cz[g] * v[3][e], hi, lo);
      dst[o] = hi;
      dst[kTile * kBatch + o] = lo;
      if (NC == 4) {
#pragma unroll
        for (int r = 1; r < 4; ++r) {
          split_bf(ct[g] * v[r][e], hi, lo);
          dst[r * kMxT + o] = hi;
          dst[r * kMxT + kTile * kBatch + o] = lo;
        }
      }
    }
  }
}

void launch_make_b_mx(hipStream_t s, int ncomp, int nbatches, const int* bt_off, const double* coef,
                      const unsigned short* cache, int sstride, unsigned short* bt) {
  if (nbatches == 0) return;
  if (ncomp == 4)
    hipLaunchKernelGGL(k_make_b_mx<4>, dim3(nbatches), dim3(256), 0, s, bt_off, coef, cache, sstride, bt);
  else
    hipLaunchKernelGGL(k_make_b_mx<1>, dim3(nbatches), dim3(256), 0, s, bt_off, coef, cache, sstride, bt);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Vxc: M_IJ += sum_g Phi_I[g]^T B_J[g] over the work item's batches, for
// 128-AO super-tiles I, J. A stage is 32 points of one component: Phi^T_I
// (128 mu x 32 g) and B^T_J (128 nu x 32 g), hi and lo, staged in LDS and
// double-buffered. 4 waves in a 2 x 2 grid, each owning a 64 x 64 block of M.
// One FP64 atomic add per element per work item.
struct VxcMxSmem {
  uint4 A[2][2][kMxSup * kTile][kLdsRow];  // [buf][hl][mu][g chunk]
  uint4 B[2][2][kMxSup * kTile][kLdsRow];
};

template <int NC>
__global__ __launch_bounds__(256) void k_vxc_mx(const int4* __restrict__ work, const int* __restrict__ pair_batches,
                                                const int* __restrict__ slot_of, int ntiles, const u16* __restrict__ cache,
                                                int sstride, const u16* __restrict__ bt, double* __restrict__ Mint,
                                                int npad) {
  __shared__ VxcMxSmem sm;
  const int4 wk = work[blockIdx.x];
  const int I = wk.x, J = wk.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  // Thread t moves, for tile t / 64 of each super-tile, the 16-byte chunks
  // e = t + 256 j (j = 0..3): hl = (e >> 7) & 1, row (e >> 2) & 31, chunk e & 3.
  const int tj = tid >> 6;
  uint4 ra[4], rq[4];
  auto load = [&](int k) {  // stage k -> (batch, half, component)
    const int it = wk.z + k / (2 * NC), h = (k / NC) % 2, c = k % NC;
    const int b = pair_batches[it];
    const int si = slot_of[(size_t)b * ntiles + I * kMxSup + tj];
    const int sj = slot_of[(size_t)b * ntiles + J * kMxSup + tj];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = (tid & 63) + 64 * j;  // chunk within the tile: 256 chunks = [hl 2][row 32][chunk 4]
      const int off = ((e >> 7) & 1) * (kTile * kBatch) + ((e >> 2) & 31) * kBatch + h * 32 + 8 * (e & 3);
      ra[j] = si >= 0 ? *reinterpret_cast<const uint4*>(cache + (size_t)si * sstride + kMxG + c * kMxT + off)
                      : uint4{0u, 0u, 0u, 0u};
      rq[j] = sj >= 0 ? *reinterpret_cast<const uint4*>(bt + ((size_t)sj * NC + c) * kMxT + off)
                      : uint4{0u, 0u, 0u, 0u};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = (tid & 63) + 64 * j;
      const int hl = (e >> 7) & 1, row = tj * kTile + ((e >> 2) & 31), ch = e & 3;
      sm.A[buf][hl][row][ch] = ra[j];
      sm.B[buf][hl][row][ch] = rq[j];
    }
  };
  f32x4 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nstage = (wk.w - wk.z) * 2 * NC;
  if (nstage > 0) load(0);
  int buf = 0;
  for (int k = 0; k < nstage; ++k) {
    store(buf);
    __syncthreads();
    if (k + 1 < nstage) load(k + 1);
    uint4 ah[4], al[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ah[r] = sm.A[buf][0][64 * wr + 16 * r + lrow][lk];
      al[r] = sm.A[buf][1][64 * wr + 16 * r + lrow][lk];
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const uint4 bh = sm.B[buf][0][64 * wc + 16 * cb + lrow][lk];
      const uint4 bl = sm.B[buf][1][64 * wc + 16 * cb + lrow][lk];
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r][cb] = mfma3(ah[r], al[r], bh, bl, acc[r][cb]);
    }
    buf ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int mu = I * kMxSup * kTile + 64 * wr + 16 * r + 4 * lk + i;
        const int nu = J * kMxSup * kTile + 64 * wc + 16 * cb + lrow;
        unsafeAtomicAdd(&Mint[(size_t)mu * npad + nu], (double)acc[r][cb][i]);
      }
}

void launch_vxc_mx(hipStream_t s, int ncomp, int nwork, const int4* work, const int* pair_batches, const int* slot_of,
                   int ntiles, const unsigned short* cache, int sstride, const unsigned short* bt, double* Mint,
                   int npad) {
  if (nwork == 0) return;
  if (ncomp == 4)
    hipLaunchKernelGGL(k_vxc_mx<4>, dim3(nwork), dim3(256), 0, s, work, pair_batches, slot_of, ntiles, cache, sstride,
                       bt, Mint, npad);
  else
    hipLaunchKernelGGL(k_vxc_mx<1>, dim3(nwork), dim3(256), 0, s, work, pair_batches, slot_of, ntiles, cache, sstride,
                       bt, Mint, npad);
  CHX_HIP(hipGetLastError());
}

}  // namespace chx
