// FP64 plain form with the AO tables cached in HBM.
//
// phi and grad phi do not change between SCF iterations, so they are computed
// once per grid (at Grid::getAO, as the reference caches its AO tables,
// src/Grid/Grid.cpp:22-33) for every alive (batch, 32-AO tile) slot and kept in
// HBM (~80 KB per slot: 4.4 GB at (H2O)16/def2-SVP/SG-2). Each iteration then
// streams them through two MFMA contractions:
//   density: X = Phi_b D_b (A = phi tiles, B = D blocks, both staged in LDS,
//            double-buffered, next tile prefetched into registers), then
//            rho = sum X phi and grad rho / 2 = sum X grad phi from the cache;
//   Vxc:     M_IJ = Phi_I^T B_J over 128-AO super-tiles, B = a phi + c . grad phi
//            built once per iteration by k_make_b; V = M + M^T at finalisation.
// Meta-GGA adds tau = 1/2 sum_r (d_r Phi D) . d_r phi (three more density
// contractions) and sum_r (d_r Phi)^T T_r to M (three more K components).
// Layouts: cache[slot][c][g][32] (c = phi, d/dx, d/dy, d/dz), bmat[slot][NC][g][32].
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace chx {

constexpr int kSlotAO = 4 * kBatch * kTile;  // doubles per cached slot
constexpr int kSlotB = kBatch * kTile;
constexpr int kSupC = 4;                    // tiles per Vxc super-tile (128 AOs)
constexpr int kSupCAO = kSupC * kTile;
constexpr int kHalf = 32;                   // points per Vxc LDS stage

// ---------------------------------------------------------------------------
// AO cache: one block per slot.
__global__ __launch_bounds__(256) void k_ao_cache(const int* __restrict__ slot_batch, const int* __restrict__ slot_tile,
                                                  const int* __restrict__ tile_nao, const double* __restrict__ bx,
                                                  const double* __restrict__ by, const double* __restrict__ bz,
                                                  BasisView bv, double* __restrict__ cache) {
  __shared__ double px[kBatch], py[kBatch], pz[kBatch];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int b = slot_batch[s], t = slot_tile[s];
  double* out = cache + (size_t)s * kSlotAO;
  if (tid < kBatch) {
    px[tid] = bx[b * kBatch + tid];
    py[tid] = by[b * kBatch + tid];
    pz[tid] = bz[b * kBatch + tid];
  }
  // padding AOs of the last tile are zero
  const int nao = tile_nao[t];
  for (int e = tid; e < 4 * kBatch * (kTile - nao); e += 256) {
    const int w = kTile - nao;
    const int c = e / (kBatch * w), r = e % (kBatch * w);
    out[((size_t)c * kBatch + r / w) * kTile + nao + r % w] = 0.0;
  }
  __syncthreads();
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
        out[(0 * kBatch + g) * kTile + col] = v[sg.y + m];
        out[(1 * kBatch + g) * kTile + col] = gx[sg.y + m];
        out[(2 * kBatch + g) * kTile + col] = gy[sg.y + m];
        out[(3 * kBatch + g) * kTile + col] = gz[sg.y + m];
      }
  }
}

void launch_ao_cache(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* tile_nao,
                     const double* bx, const double* by, const double* bz, BasisView bv, double* cache) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_ao_cache, dim3(nslots), dim3(256), 0, s, slot_batch, slot_tile, tile_nao, bx, by, bz, bv, cache);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Density from the cache. Block = (batch, group of <= 4 alive tiles as output
// columns), 8 waves: wave w owns rows g = 16(w%4)..+15 and output columns
// 64(w/4)..+63 (two tiles). K loop over the batch's alive tiles: phi tile
// [g][mu] and D block [mu][nu] double-buffered in LDS, the next K step
// prefetched into registers while the MFMAs of the current one run.
struct DensCSmem {
  double A[2][kBatch][kTile + 1];
  double B[2][kTile][kGroup * kTile];
  double red[2][4][kBatch];
};

// MODE 0: A = phi, epilogue rho and grad rho / 2 (4 components into part).
// MODE r = 1..3: A = d phi / d x_r, epilogue sum X_r d_r phi into part (one
// component: the tau partial of direction r).
template <int MODE>
__global__ __launch_bounds__(512) void k_density_c64(const int* __restrict__ grp_batch, const int* __restrict__ grp_tiles,
                                                     const int* __restrict__ grp_slots, const int* __restrict__ grp_ntiles,
                                                     const int* __restrict__ bt_off, const int* __restrict__ bt_list,
                                                     const double* __restrict__ cache, const double* __restrict__ Dint,
                                                     int npad, double* __restrict__ part, int grp_stride) {
  constexpr int NOUT = MODE == 0 ? 4 : 1;
  __shared__ DensCSmem sm;
  const int grp = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform, keeps branches scalar
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 3, wc = wave >> 2;  // row block, column half
  const int b = grp_batch[grp];
  const int nt = grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = grp_tiles[grp * kGroup + q];
  const int it0 = bt_off[b], it1 = bt_off[b + 1];

  double2 ra[2], rb[4];
  auto load = [&](int it) {
    const double2* src = reinterpret_cast<const double2*>(cache + (size_t)it * kSlotAO + MODE * kSlotB);
#pragma unroll
    for (int j = 0; j < 2; ++j) ra[j] = src[tid + 512 * j];
    const int ti = bt_list[it];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = 2 * (tid + 512 * j);
      const int r = e / (kGroup * kTile), cc = e % (kGroup * kTile), q = cc / kTile;
      rb[j] = (q < nt) ? *reinterpret_cast<const double2*>(Dint + (size_t)(ti * kTile + r) * npad + jt[q] * kTile + (cc % kTile))
                       : double2{0.0, 0.0};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int e = 2 * (tid + 512 * j);
      sm.A[buf][e / kTile][e % kTile] = ra[j].x;
      sm.A[buf][e / kTile][e % kTile + 1] = ra[j].y;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = 2 * (tid + 512 * j);
      *reinterpret_cast<double2*>(&sm.B[buf][e / (kGroup * kTile)][e % (kGroup * kTile)]) = rb[j];
    }
  };

  dbl4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = dbl4{0, 0, 0, 0};
  const int ncb = min(4, max(0, 2 * nt - 4 * wc));  // valid column blocks of this wave
  if (it0 < it1) load(it0);
  int buf = 0;
  for (int it = it0; it < it1; ++it) {
    store(buf);
    __syncthreads();
    if (it + 1 < it1) load(it + 1);
#pragma unroll
    for (int ks = 0; ks < kTile / 4; ++ks) {
      const double a = sm.A[buf][16 * wr + lrow][4 * ks + lk];
      // columns of absent tiles hold zeros in LDS, so every wave runs all four
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[cb] = mfma_f64(a, sm.B[buf][4 * ks + lk][64 * wc + 16 * cb + lrow], acc[cb]);
    }
    buf ^= 1;
  }

  // Epilogue: X[g = 16wr + 4r + lk][nu = 16h + lrow] of tile q = 2wc + cb/2, against phi and grad phi.
  double s[4][4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) s[c][r] = 0.0;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    if (cb < ncb) {
      const int q = 2 * wc + cb / 2, h = cb % 2;
      const double* src = cache + (size_t)grp_slots[grp * kGroup + q] * kSlotAO;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int g = 16 * wr + 4 * r + lk;
        const double x = acc[cb][r];
#pragma unroll
        for (int c = 0; c < NOUT; ++c)
          s[c][r] += x * src[((size_t)(MODE == 0 ? c : MODE) * kBatch + g) * kTile + 16 * h + lrow];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NOUT; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = s[c][r];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      s[c][r] = v;
    }
  double out = 0;
#pragma unroll
  for (int c = 0; c < NOUT; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (lrow == 4 * c + r) out = s[c][r];
  const int c = lrow >> 2, r = lrow & 3;
  if (c < NOUT) sm.red[wc][c][16 * wr + 4 * r + lk] = out;
  __syncthreads();
  if (tid < NOUT * kBatch) {
    const int cc = tid / kBatch, g = tid % kBatch;
    part[(size_t)grp * grp_stride + cc * kBatch + g] = sm.red[0][cc][g] + sm.red[1][cc][g];
  }
}

void launch_density_c64(hipStream_t s, int mode, int ngroups, const int* grp_batch, const int* grp_tiles,
                        const int* grp_slots, const int* grp_ntiles, const int* bt_off, const int* bt_list,
                        const double* cache, const double* Dint, int npad, double* part) {
  if (ngroups == 0) return;
  const int stride = (mode == 0 ? 4 : 3) * kBatch;  // part: [group][4][64]; tau partials: [group][3][64]
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_density_c64<0>, dim3(ngroups), dim3(512), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, Dint, npad, part, stride); break;
    case 1: hipLaunchKernelGGL(k_density_c64<1>, dim3(ngroups), dim3(512), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, Dint, npad, part, stride); break;
    case 2: hipLaunchKernelGGL(k_density_c64<2>, dim3(ngroups), dim3(512), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, Dint, npad, part, stride); break;
    default: hipLaunchKernelGGL(k_density_c64<3>, dim3(ngroups), dim3(512), 0, s, grp_batch, grp_tiles, grp_slots, grp_ntiles, bt_off, bt_list, cache, Dint, npad, part, stride); break;
  }
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// B = a phi + c . grad phi for every slot of a batch (one block per batch),
// one component per slot for every family: the meta-GGA operand
// T_r = t d_r phi (t = w v_tau / 4, so that V = M + M^T carries
// (1/2) w v_tau grad phi_mu . grad phi_nu) is formed from the cached d_r phi
// while the Vxc kernel loads it, instead of being stored.
__global__ __launch_bounds__(256) void k_make_b(const int* __restrict__ bt_off, const double* __restrict__ coef,
                                                const double* __restrict__ cache, double* __restrict__ bmat) {
  __shared__ double ca[kBatch], cx[kBatch], cy[kBatch], cz[kBatch];
  const int b = blockIdx.x, tid = threadIdx.x;
  if (tid < kBatch) {
    const size_t slot = (size_t)b * kBatch + tid;
    ca[tid] = coef[kCoef * slot];
    cx[tid] = coef[kCoef * slot + 1];
    cy[tid] = coef[kCoef * slot + 2];
    cz[tid] = coef[kCoef * slot + 3];
  }
  __syncthreads();
  for (int it = bt_off[b]; it < bt_off[b + 1]; ++it) {
    const double* src = cache + (size_t)it * kSlotAO;
    double* dst = bmat + (size_t)it * kSlotB;
    for (int e = tid; e < kSlotB; e += 256) {
      const int g = e / kTile;
      const double dx = src[kSlotB + e], dy = src[2 * kSlotB + e], dz = src[3 * kSlotB + e];
      dst[e] = ca[g] * src[e] + cx[g] * dx + cy[g] * dy + cz[g] * dz;
    }
  }
}

void launch_make_b(hipStream_t s, int nbatches, const int* bt_off, const double* coef, const double* cache,
                   double* bmat) {
  if (nbatches == 0) return;
  hipLaunchKernelGGL(k_make_b, dim3(nbatches), dim3(256), 0, s, bt_off, coef, cache, bmat);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Vxc from the cache: M_IJ += sum_g phi_I[g]^T B_J[g] over the work item's
// batches, 32 points per LDS stage (double-buffered), 8 waves: wave w owns rows
// mu = 16w..16w+15 of the 128 x 128 block. One atomic add per element per item.
struct VxcCSmem {
  double P[2][kHalf][kSupCAO + 1];
  double Bt[2][kHalf][kSupCAO + 1];
};

// NC = 1: M = phi^T B. NC = 4 (meta-GGA): also sum_r (d_r phi)^T T_r, each
// component streamed as further K stages, T_r = t d_r phi formed from the
// cached d_r phi and the point's t = coef[4] as it is loaded.
template <int NC>
__global__ __launch_bounds__(512) void k_vxc_c64(const int4* __restrict__ work, const int* __restrict__ pair_batches,
                                                 const int* __restrict__ slot_of, int ntiles,
                                                 const double* __restrict__ cache, const double* __restrict__ bmat,
                                                 const double* __restrict__ coef, double* __restrict__ Mint,
                                                 int npad) {
  __shared__ VxcCSmem sm;
  const int4 wk = work[blockIdx.x];
  const int I = wk.x, J = wk.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  // Per stage and operand: 4 tiles x 32 points x 32 AOs; thread t moves the
  // double2 at element 2 (t + 512 j) of tile j.
  double2 rp[kSupC], rq[kSupC];
  const int r0 = 2 * tid;  // element within a tile's 32 x 32 block
  auto load = [&](int k) {  // stage k -> (batch, half, component)
    const int it = wk.z + k / (2 * NC), h = (k / NC) % 2, c = k % NC;
    const int b = pair_batches[it];
    const int g = h * kHalf + r0 / kTile, mu = r0 % kTile;
    const double t = c ? coef[kCoef * ((size_t)b * kBatch + g) + 4] : 0.0;
#pragma unroll
    for (int j = 0; j < kSupC; ++j) {
      const int si = slot_of[(size_t)b * ntiles + I * kSupC + j];
      const int sj = slot_of[(size_t)b * ntiles + J * kSupC + j];
      rp[j] = si >= 0 ? *reinterpret_cast<const double2*>(cache + (size_t)si * kSlotAO + (size_t)c * kSlotB + g * kTile + mu)
                      : double2{0.0, 0.0};
      if (sj < 0) {
        rq[j] = double2{0.0, 0.0};
      } else if (c == 0) {
        rq[j] = *reinterpret_cast<const double2*>(bmat + (size_t)sj * kSlotB + g * kTile + mu);
      } else {
        const double2 d = *reinterpret_cast<const double2*>(cache + (size_t)sj * kSlotAO + (size_t)c * kSlotB +
                                                             g * kTile + mu);
        rq[j] = double2{t * d.x, t * d.y};
      }
    }
  };
  auto store = [&](int buf) {
    const int g = r0 / kTile, m = r0 % kTile;
#pragma unroll
    for (int j = 0; j < kSupC; ++j) {
      const int col = j * kTile + m;
      sm.P[buf][g][col] = rp[j].x;
      sm.P[buf][g][col + 1] = rp[j].y;
      sm.Bt[buf][g][col] = rq[j].x;
      sm.Bt[buf][g][col + 1] = rq[j].y;
    }
  };
  dbl4 acc[8];
#pragma unroll
  for (int cb = 0; cb < 8; ++cb) acc[cb] = dbl4{0, 0, 0, 0};
  const int nstage = (wk.w - wk.z) * (kBatch / kHalf) * NC;
  if (nstage > 0) load(0);
  int buf = 0;
  for (int k = 0; k < nstage; ++k) {
    store(buf);
    __syncthreads();
    if (k + 1 < nstage) load(k + 1);
#pragma unroll
    for (int ks = 0; ks < kHalf / 4; ++ks) {
      const int g = 4 * ks + lk;
      const double a = sm.P[buf][g][16 * wave + lrow];
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) acc[cb] = mfma_f64(a, sm.Bt[buf][g][16 * cb + lrow], acc[cb]);
    }
    buf ^= 1;
  }
#pragma unroll
  for (int cb = 0; cb < 8; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = I * kSupCAO + 16 * wave + 4 * r + lk;
      const int nu = J * kSupCAO + 16 * cb + lrow;
      unsafeAtomicAdd(&Mint[(size_t)mu * npad + nu], acc[cb][r]);
    }
}

void launch_vxc_c64(hipStream_t s, int ncomp, int nwork, const int4* work, const int* pair_batches, const int* slot_of,
                    int ntiles, const double* cache, const double* bmat, const double* coef, double* Mint, int npad) {
  if (nwork == 0) return;
  if (ncomp == 4)
    hipLaunchKernelGGL(k_vxc_c64<4>, dim3(nwork), dim3(512), 0, s, work, pair_batches, slot_of, ntiles, cache, bmat,
                       coef, Mint, npad);
  else
    hipLaunchKernelGGL(k_vxc_c64<1>, dim3(nwork), dim3(512), 0, s, work, pair_batches, slot_of, ntiles, cache, bmat,
                       coef, Mint, npad);
  CHX_HIP(hipGetLastError());
}

// V (reference order, column-major) = M + M^T.
__global__ void k_finalize_sym(int n, int npad, const double* __restrict__ Mint, const int* __restrict__ ref2int,
                               double* __restrict__ V) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int col = (int)(idx / n), row = (int)(idx % n);
  const int mu = ref2int[row], nu = ref2int[col];
  V[idx] = Mint[(size_t)mu * npad + nu] + Mint[(size_t)nu * npad + mu];
}

void launch_finalize_sym(hipStream_t s, int n, int npad, const double* Mint, const int* ref2int, double* V) {
  const long tot = (long)n * n;
  hipLaunchKernelGGL(k_finalize_sym, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, npad, Mint, ref2int, V);
  CHX_HIP(hipGetLastError());
}

}  // namespace chx
