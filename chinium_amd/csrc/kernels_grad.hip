// XC nuclear gradient, FP64, skeleton form with the grid held fixed
// (src/HartreeFockKohnSham/Restricted/Grad.cpp:72-84 -> SubGrid::getDensitySkeleton,
// src/Grid/GridDensity.cpp:79-130 -> SubGrid::getEnergyGrad, src/Grid/GridEnergy.cpp:15-30).
//
// With a = w v_rho / 2, c = 2 w v_sigma grad rho and B = a phi + c . grad phi
// (the Vxc integrand of the same point), the reference's
//   E[t,A] = sum_g w (v_rho dRho/dA_t + v_sigma dSigma/dA_t)
// reduces to
//   E[t,A] = -2 sum_{mu in A} sum_g [ d_t phi_mu (Z_mu + a X_mu) + (c . grad) d_t phi_mu X_mu ],
//   X_mu = sum_nu D_mu,nu phi_nu,   Z_mu = sum_nu D_mu,nu B_nu,
// two contractions of the density's shape (k_xgrad_xz, FP64 MFMA over the
// cached phi and B tables) and a per-slot epilogue with the AO Hessian
// (k_xgrad_epi, one wave per slot, lane = point).
//
// The skeleton XC response (Grid::getFockU<s_t>, src/Grid/Grid.cpp:193-216)
// takes the same skeleton density derivatives per point, for every nuclear
// coordinate (GridDensity.cpp:90-128):
//   rho^{A,t}     = -2 sum_{mu in A} d_t phi_mu X_mu,
//   d_r rho^{A,t} = -2 sum_{mu in A} (d_r d_t phi_mu X_mu + d_t phi_mu Y^r_mu),
//   Y^r_mu = sum_nu D_mu,nu d_r phi_nu:
// the contraction kernel run with (phi, d_x phi) and (d_y phi, d_z phi) as its
// two A operands, then k_skel_rho per slot.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace chx {

constexpr int kGSlotAO = 4 * kBatch * kTile;  // FP64 cache: doubles per slot
constexpr int kGSlotB = kBatch * kTile;
constexpr int kGCols = kGroup * kTile;        // 128 columns per group

// ---------------------------------------------------------------------------
// Two products of one density work group (batch, <= 4 column tiles) with the
// group's D block: A1 = cache component c1 (0 phi, r d_r phi), A2 = component
// c2, or the B table when c2 < 0. Written as xz[grp][0 = A1 D, 1 = A2 D][column][g];
// (c1, c2) = (0, -1) gives the gradient's X and Z. 4 waves, wave w owns rows
// 16w..16w+15.
struct XgradSmem {
  double Ap[kBatch][kTile + 1];  // A1 of the K tile, [g][nu]
  double Ab[kBatch][kTile + 1];  // A2 of the K tile
  double Bd[kTile][kGCols];      // D block [nu][mu]
};

__global__ __launch_bounds__(256) void k_xgrad_xz(const int* __restrict__ grp_batch, const int* __restrict__ grp_tiles,
                                                  const int* __restrict__ grp_ntiles, const int* __restrict__ bt_off,
                                                  const int* __restrict__ bt_list, const double* __restrict__ cache,
                                                  const double* __restrict__ bmat, const double* __restrict__ Dint,
                                                  int npad, int c1, int c2, double* __restrict__ xz) {
  __shared__ XgradSmem sm;
  const int grp = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int b = grp_batch[grp];
  const int nt = grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = grp_tiles[grp * kGroup + q];
  dbl4 ax[8], az[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) ax[c] = az[c] = dbl4{0, 0, 0, 0};
  for (int it = bt_off[b]; it < bt_off[b + 1]; ++it) {
    const int ti = bt_list[it];
    const double* p = cache + (size_t)it * kGSlotAO + (size_t)c1 * kBatch * kTile;  // [g][nu]
    const double* q = c2 < 0 ? bmat + (size_t)it * kGSlotB : cache + (size_t)it * kGSlotAO + (size_t)c2 * kBatch * kTile;
    for (int e = tid; e < kBatch * kTile; e += 256) {
      sm.Ap[e / kTile][e % kTile] = p[e];
      sm.Ab[e / kTile][e % kTile] = q[e];
    }
    for (int e = tid; e < kTile * kGCols; e += 256) {
      const int r = e / kGCols, c = e % kGCols, qq = c / kTile;
      sm.Bd[r][c] = qq < nt ? Dint[(size_t)(ti * kTile + r) * npad + jt[qq] * kTile + (c % kTile)] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kTile / 4; ++ks) {
      const double a1 = sm.Ap[16 * wave + lrow][4 * ks + lk];
      const double a2 = sm.Ab[16 * wave + lrow][4 * ks + lk];
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) {
        const double bb = sm.Bd[4 * ks + lk][16 * cb + lrow];
        ax[cb] = mfma_f64(a1, bb, ax[cb]);
        az[cb] = mfma_f64(a2, bb, az[cb]);
      }
    }
    __syncthreads();
  }
  // result register r of lane l: row g = 16 wave + 4 r + l / 16, column 16 cb + l % 16
  double* out = xz + (size_t)grp * 2 * kGCols * kBatch;
#pragma unroll
  for (int cb = 0; cb < 8; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int g = 16 * wave + 4 * r + lk, col = 16 * cb + lrow;
      out[(size_t)col * kBatch + g] = ax[cb][r];
      out[(size_t)(kGCols + col) * kBatch + g] = az[cb][r];
    }
}

void launch_xgrad_xz(hipStream_t s, int ngroups, const int* grp_batch, const int* grp_tiles, const int* grp_ntiles,
                     const int* bt_off, const int* bt_list, const double* cache, const double* bmat,
                     const double* Dint, int npad, int c1, int c2, double* xz) {
  if (ngroups == 0) return;
  hipLaunchKernelGGL(k_xgrad_xz, dim3(ngroups), dim3(256), 0, s, grp_batch, grp_tiles, grp_ntiles, bt_off, bt_list,
                     cache, bmat, Dint, npad, c1, c2, xz);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Angular parts with first and second derivatives, one shell type per
// instantiation (conventions of src/Grid/AO/{S,CartP,PureP,PureD,PureF}.hpp).
// Hessian order (xx, xy, xz, yy, yz, zz).
template <int TYPE>
__device__ __forceinline__ void angular_h(double x, double y, double z, double* q, double (*g)[3], double (*h)[6]) {
  const double r3 = 1.7320508075688772, r6 = 2.449489742783178, r10 = 3.1622776601683795, r15 = 3.872983346207417;
  auto set = [&](int m, double v, double gx, double gy, double gz, double hxx, double hxy, double hxz, double hyy,
                 double hyz, double hzz) {
    q[m] = v;
    g[m][0] = gx; g[m][1] = gy; g[m][2] = gz;
    h[m][0] = hxx; h[m][1] = hxy; h[m][2] = hxz; h[m][3] = hyy; h[m][4] = hyz; h[m][5] = hzz;
  };
  if (TYPE == 0) {
    set(0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0);
  } else if (TYPE == 1) {  // x, y, z
    set(0, x, 1, 0, 0, 0, 0, 0, 0, 0, 0);
    set(1, y, 0, 1, 0, 0, 0, 0, 0, 0, 0);
    set(2, z, 0, 0, 1, 0, 0, 0, 0, 0, 0);
  } else if (TYPE == -1) {  // y, z, x
    set(0, y, 0, 1, 0, 0, 0, 0, 0, 0, 0);
    set(1, z, 0, 0, 1, 0, 0, 0, 0, 0, 0);
    set(2, x, 1, 0, 0, 0, 0, 0, 0, 0, 0);
  } else if (TYPE == -2) {
    set(0, r3 * x * y, r3 * y, r3 * x, 0, 0, r3, 0, 0, 0, 0);
    set(1, r3 * y * z, 0, r3 * z, r3 * y, 0, 0, 0, 0, r3, 0);
    set(2, z * z - 0.5 * (x * x + y * y), -x, -y, 2 * z, -1, 0, 0, -1, 0, 2);
    set(3, r3 * x * z, r3 * z, 0, r3 * x, 0, 0, r3, 0, 0, 0);
    set(4, 0.5 * r3 * (x * x - y * y), r3 * x, -r3 * y, 0, r3, 0, 0, -r3, 0, 0);
  } else {  // pure f
    const double xx = x * x, yy = y * y, zz = z * z;
    const double c10 = 0.25 * r10, c6 = 0.25 * r6;
    // y (3x^2 - y^2)
    set(0, c10 * y * (3 * xx - yy), c10 * 6 * x * y, c10 * 3 * (xx - yy), 0,
        c10 * 6 * y, c10 * 6 * x, 0, -c10 * 6 * y, 0, 0);
    // xyz
    set(1, r15 * x * y * z, r15 * y * z, r15 * x * z, r15 * x * y, 0, r15 * z, r15 * y, 0, r15 * x, 0);
    // y (4z^2 - x^2 - y^2)
    set(2, c6 * y * (4 * zz - xx - yy), -c6 * 2 * x * y, c6 * (4 * zz - xx - 3 * yy), c6 * 8 * y * z,
        -c6 * 2 * y, -c6 * 2 * x, 0, -c6 * 6 * y, c6 * 8 * z, c6 * 8 * y);
    // z (z^2 - 3/2 (x^2 + y^2))
    set(3, z * (zz - 1.5 * (xx + yy)), -3 * x * z, -3 * y * z, 3 * zz - 1.5 * (xx + yy),
        -3 * z, 0, -3 * x, -3 * z, -3 * y, 6 * z);
    // x (4z^2 - x^2 - y^2)
    set(4, c6 * x * (4 * zz - xx - yy), c6 * (4 * zz - 3 * xx - yy), -c6 * 2 * x * y, c6 * 8 * x * z,
        -c6 * 6 * x, -c6 * 2 * y, c6 * 8 * z, -c6 * 2 * x, 0, c6 * 8 * x);
    // (x^2 - y^2) z
    set(5, 0.5 * r15 * (xx - yy) * z, r15 * x * z, -r15 * y * z, 0.5 * r15 * (xx - yy),
        r15 * z, 0, r15 * x, -r15 * z, -r15 * y, 0);
    // x (x^2 - 3y^2)
    set(6, c10 * x * (xx - 3 * yy), c10 * 3 * (xx - yy), -c10 * 6 * x * y, 0,
        c10 * 6 * x, -c10 * 6 * y, 0, -c10 * 6 * x, 0, 0);
  }
}

// Gradient contribution of components first..first+cnt-1 of one shell at this
// lane's point: s_t += d_t phi (Z + a X) + (c . grad) d_t phi X.
template <int TYPE>
__device__ __forceinline__ void xgrad_shell(double x, double y, double z, double A0, const double* dA, const double* HA,
                                            int first, int cnt, const double* xcol, const double* zcol, double a,
                                            const double* c, double* st) {
  constexpr int n = TYPE == 0 ? 1 : TYPE == -2 ? 5 : TYPE == -3 ? 7 : 3;
  double q[n], g[n][3], h[n][6];
  angular_h<TYPE>(x, y, z, q, g, h);
  const int hi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
  for (int m = 0; m < n; ++m) {
    const int k = m - first;
    if (k < 0 || k >= cnt) continue;
    const double X = xcol[k * kBatch], Zv = zcol[k * kBatch];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const double dt = dA[t] * q[m] + A0 * g[m][t];  // d_t phi
      double cd = 0;                                  // (c . grad) d_t phi
#pragma unroll
      for (int r = 0; r < 3; ++r)
        cd += c[r] * (HA[hi[r][t]] * q[m] + dA[r] * g[m][t] + dA[t] * g[m][r] + A0 * h[m][hi[r][t]]);
      st[t] += dt * (Zv + a * X) + cd * X;
    }
  }
}

// One wave per slot (batch, tile), lane = point; per shell segment the three
// sums are reduced over the wave and added to grad[atom][t] (scaled by -2).
__global__ __launch_bounds__(256) void k_xgrad_epi(int nslots, const int* __restrict__ slot_batch,
                                                   const int* __restrict__ slot_tile, const int* __restrict__ slot_grp,
                                                   const int* __restrict__ slot_q, const double* __restrict__ bx,
                                                   const double* __restrict__ by, const double* __restrict__ bz,
                                                   BasisView bv, const double* __restrict__ coef,
                                                   const double* __restrict__ xz, double* __restrict__ grad) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), g = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= nslots) return;
  const int b = slot_batch[slot], t = slot_tile[slot];
  const size_t p = (size_t)b * kBatch + g;
  const double px = bx[p], py = by[p], pz = bz[p];
  const double a = coef[kCoef * p];
  const double c[3] = {coef[kCoef * p + 1], coef[kCoef * p + 2], coef[kCoef * p + 3]};
  const double* X = xz + (size_t)slot_grp[slot] * 2 * kGCols * kBatch + (size_t)slot_q[slot] * kTile * kBatch + g;
  const double* Z = X + (size_t)kGCols * kBatch;
  const int s0 = bv.tile_seg_off[t], s1 = bv.tile_seg_off[t + 1];
  for (int k = s0; k < s1; ++k) {
    const int4 sg = bv.seg[k];
    const int sh = sg.x;
    const double x = px - bv.sh_xyz[3 * sh], y = py - bv.sh_xyz[3 * sh + 1], z = pz - bv.sh_xyz[3 * sh + 2];
    const double r2 = x * x + y * y + z * z;
    double A0 = 0, A1 = 0, A2 = 0;
    for (int i = 0, po = bv.sh_poff[sh]; i < bv.sh_nprim[sh]; ++i) {
      const double e = bv.p_exp[po + i];
      const double v = bv.p_coef[po + i] * ::exp(-e * r2);
      A0 += v;
      A1 += e * v;
      A2 += e * e * v;
    }
    const double d[3] = {x, y, z};
    double dA[3], HA[6];
    for (int r = 0; r < 3; ++r) dA[r] = -2 * d[r] * A1;
    HA[0] = 4 * x * x * A2 - 2 * A1; HA[1] = 4 * x * y * A2; HA[2] = 4 * x * z * A2;
    HA[3] = 4 * y * y * A2 - 2 * A1; HA[4] = 4 * y * z * A2; HA[5] = 4 * z * z * A2 - 2 * A1;
    double st[3] = {0, 0, 0};
    const double* xc = X + (size_t)sg.w * kBatch;
    const double* zc = Z + (size_t)sg.w * kBatch;
    switch (bv.sh_type[sh]) {
      case 0: xgrad_shell<0>(x, y, z, A0, dA, HA, sg.y, sg.z, xc, zc, a, c, st); break;
      case 1: xgrad_shell<1>(x, y, z, A0, dA, HA, sg.y, sg.z, xc, zc, a, c, st); break;
      case -1: xgrad_shell<-1>(x, y, z, A0, dA, HA, sg.y, sg.z, xc, zc, a, c, st); break;
      case -2: xgrad_shell<-2>(x, y, z, A0, dA, HA, sg.y, sg.z, xc, zc, a, c, st); break;
      default: xgrad_shell<-3>(x, y, z, A0, dA, HA, sg.y, sg.z, xc, zc, a, c, st); break;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double v = st[r];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      st[r] = v;
    }
    if (g == 0) {
      const int atom = bv.sh_atom[sh];
#pragma unroll
      for (int r = 0; r < 3; ++r) unsafeAtomicAdd(&grad[3 * atom + r], -2.0 * st[r]);
    }
  }
}

void launch_xgrad_epi(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* slot_grp,
                      const int* slot_q, const double* bx, const double* by, const double* bz, BasisView bv,
                      const double* coef, const double* xz, double* grad) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_xgrad_epi, dim3((nslots + 3) / 4), dim3(256), 0, s, nslots, slot_batch, slot_tile, slot_grp,
                     slot_q, bx, by, bz, bv, coef, xz, grad);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Skeleton density derivatives of one shell's components at this lane's
// point: acc[t] += (d_t phi X, sum over r of d_r d_t phi X + d_t phi Y^r for r = x, y, z).
template <int TYPE>
__device__ __forceinline__ void skel_shell(double x, double y, double z, double A0, const double* dA, const double* HA,
                                           int first, int cnt, const double* c1, const double* c2,
                                           double (*acc)[4]) {
  constexpr int n = TYPE == 0 ? 1 : TYPE == -2 ? 5 : TYPE == -3 ? 7 : 3;
  double q[n], g[n][3], h[n][6];
  angular_h<TYPE>(x, y, z, q, g, h);
  const int hi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
  for (int m = 0; m < n; ++m) {
    const int k = m - first;
    if (k < 0 || k >= cnt) continue;
    const double X = c1[k * kBatch];
    const double Y[3] = {c1[(kGCols + k) * kBatch], c2[k * kBatch], c2[(kGCols + k) * kBatch]};
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const double dt = dA[t] * q[m] + A0 * g[m][t];  // d_t phi
      acc[t][0] += dt * X;
#pragma unroll
      for (int r = 0; r < 3; ++r)
        acc[t][1 + r] += (HA[hi[r][t]] * q[m] + dA[r] * g[m][t] + dA[t] * g[m][r] + A0 * h[m][hi[r][t]]) * X +
                         dt * Y[r];
    }
  }
}

// One wave per slot (batch, tile), lane = point. xz1 / xz2: the group products
// (phi D, d_x phi D) and (d_y phi D, d_z phi D). The sums of consecutive shells
// of one atom are added, scaled by -2, to rhos[3 A + t][point][rho, d_x, d_y, d_z]
// (npt points per coordinate) by FP64 atomics.
__global__ __launch_bounds__(256) void k_skel_rho(int nslots, const int* __restrict__ slot_batch,
                                                  const int* __restrict__ slot_tile, const int* __restrict__ slot_grp,
                                                  const int* __restrict__ slot_q, const double* __restrict__ bx,
                                                  const double* __restrict__ by, const double* __restrict__ bz,
                                                  BasisView bv, const double* __restrict__ xz1,
                                                  const double* __restrict__ xz2, long npt, double* __restrict__ rhos) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), g = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= nslots) return;
  const int b = slot_batch[slot], t = slot_tile[slot];
  const size_t p = (size_t)b * kBatch + g;
  const double px = bx[p], py = by[p], pz = bz[p];
  const size_t col0 = (size_t)slot_grp[slot] * 2 * kGCols * kBatch + (size_t)slot_q[slot] * kTile * kBatch + g;
  const double* X1 = xz1 + col0;
  const double* X2 = xz2 + col0;
  double acc[3][4] = {};
  int atom = -1;
  auto flush = [&]() {
    if (atom < 0) return;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        unsafeAtomicAdd(&rhos[((size_t)(3 * atom + d) * npt + p) * 4 + k], -2.0 * acc[d][k]);
        acc[d][k] = 0.0;
      }
  };
  const int s0 = bv.tile_seg_off[t], s1 = bv.tile_seg_off[t + 1];
  for (int k = s0; k < s1; ++k) {
    const int4 sg = bv.seg[k];
    const int sh = sg.x;
    if (bv.sh_atom[sh] != atom) {
      flush();
      atom = bv.sh_atom[sh];
    }
    const double x = px - bv.sh_xyz[3 * sh], y = py - bv.sh_xyz[3 * sh + 1], z = pz - bv.sh_xyz[3 * sh + 2];
    const double r2 = x * x + y * y + z * z;
    double A0 = 0, A1 = 0, A2 = 0;
    for (int i = 0, po = bv.sh_poff[sh]; i < bv.sh_nprim[sh]; ++i) {
      const double e = bv.p_exp[po + i];
      const double v = bv.p_coef[po + i] * ::exp(-e * r2);
      A0 += v;
      A1 += e * v;
      A2 += e * e * v;
    }
    const double d[3] = {x, y, z};
    double dA[3], HA[6];
    for (int r = 0; r < 3; ++r) dA[r] = -2 * d[r] * A1;
    HA[0] = 4 * x * x * A2 - 2 * A1; HA[1] = 4 * x * y * A2; HA[2] = 4 * x * z * A2;
    HA[3] = 4 * y * y * A2 - 2 * A1; HA[4] = 4 * y * z * A2; HA[5] = 4 * z * z * A2 - 2 * A1;
    const double* c1 = X1 + (size_t)sg.w * kBatch;
    const double* c2 = X2 + (size_t)sg.w * kBatch;
    switch (bv.sh_type[sh]) {
      case 0: skel_shell<0>(x, y, z, A0, dA, HA, sg.y, sg.z, c1, c2, acc); break;
      case 1: skel_shell<1>(x, y, z, A0, dA, HA, sg.y, sg.z, c1, c2, acc); break;
      case -1: skel_shell<-1>(x, y, z, A0, dA, HA, sg.y, sg.z, c1, c2, acc); break;
      case -2: skel_shell<-2>(x, y, z, A0, dA, HA, sg.y, sg.z, c1, c2, acc); break;
      default: skel_shell<-3>(x, y, z, A0, dA, HA, sg.y, sg.z, c1, c2, acc); break;
    }
  }
  flush();
}

void launch_skel_rho(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* slot_grp,
                     const int* slot_q, const double* bx, const double* by, const double* bz, BasisView bv,
                     const double* xz1, const double* xz2, long npt, double* rhos) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_skel_rho, dim3((nslots + 3) / 4), dim3(256), 0, s, nslots, slot_batch, slot_tile, slot_grp,
                     slot_q, bx, by, bz, bv, xz1, xz2, npt, rhos);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Skeleton derivative of the XC Fock matrix with the potential held fixed
// (Grid::getFockSkeleton, src/Grid/Grid.cpp:173-191 -> GridFock.cpp:51-97).
// With a, c and B = a phi + c . grad phi as above, for direction t:
//   N^t[mu][nu] = sum_g (d_t phi_mu B_nu + E^t_mu phi_nu),  E^t = a d_t phi + (c . grad) d_t phi,
//   F[3 A + t] = -(P_A N^t + (P_A N^t)^T), P_A keeping the rows of atom A's AOs.
// The two products run through the cached Vxc contraction: Q^t = (d_t Phi)^T B
// and P^t = Phi^T E^t, so N^t[mu][nu] = Q^t[mu][nu] + P^t[nu][mu].
template <int TYPE>
__device__ __forceinline__ void etab_shell(double x, double y, double z, double A0, const double* dA, const double* HA,
                                           int first, int cnt, double a, const double* c, double* out,
                                           size_t tstride) {
  constexpr int n = TYPE == 0 ? 1 : TYPE == -2 ? 5 : TYPE == -3 ? 7 : 3;
  double q[n], g[n][3], h[n][6];
  angular_h<TYPE>(x, y, z, q, g, h);
  const int hi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
  for (int m = 0; m < n; ++m) {
    const int k = m - first;
    if (k < 0 || k >= cnt) continue;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const double dt = dA[t] * q[m] + A0 * g[m][t];  // d_t phi
      double cd = 0;                                  // (c . grad) d_t phi
#pragma unroll
      for (int r = 0; r < 3; ++r)
        cd += c[r] * (HA[hi[r][t]] * q[m] + dA[r] * g[m][t] + dA[t] * g[m][r] + A0 * h[m][hi[r][t]]);
      out[t * tstride + k] = a * dt + cd;
    }
  }
}

// One wave per slot (batch, tile), lane = point: E^t of the tile's AOs into
// etab[t][slot][g][column] (columns past the tile's AOs keep their zeros).
__global__ __launch_bounds__(256) void k_skel_etab(int nslots, const int* __restrict__ slot_batch,
                                                   const int* __restrict__ slot_tile, const double* __restrict__ bx,
                                                   const double* __restrict__ by, const double* __restrict__ bz,
                                                   BasisView bv, const double* __restrict__ coef,
                                                   double* __restrict__ etab) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), g = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= nslots) return;
  const int b = slot_batch[slot], t = slot_tile[slot];
  const size_t p = (size_t)b * kBatch + g;
  const double px = bx[p], py = by[p], pz = bz[p];
  const double a = coef[kCoef * p];
  const double c[3] = {coef[kCoef * p + 1], coef[kCoef * p + 2], coef[kCoef * p + 3]};
  double* out0 = etab + (size_t)slot * kGSlotB + (size_t)g * kTile;
  const size_t tstride = (size_t)nslots * kGSlotB;
  const int s0 = bv.tile_seg_off[t], s1 = bv.tile_seg_off[t + 1];
  for (int k = s0; k < s1; ++k) {
    const int4 sg = bv.seg[k];
    const int sh = sg.x;
    const double x = px - bv.sh_xyz[3 * sh], y = py - bv.sh_xyz[3 * sh + 1], z = pz - bv.sh_xyz[3 * sh + 2];
    const double r2 = x * x + y * y + z * z;
    double A0 = 0, A1 = 0, A2 = 0;
    for (int i = 0, po = bv.sh_poff[sh]; i < bv.sh_nprim[sh]; ++i) {
      const double e = bv.p_exp[po + i];
      const double v = bv.p_coef[po + i] * ::exp(-e * r2);
      A0 += v;
      A1 += e * v;
      A2 += e * e * v;
    }
    const double d[3] = {x, y, z};
    double dA[3], HA[6];
    for (int r = 0; r < 3; ++r) dA[r] = -2 * d[r] * A1;
    HA[0] = 4 * x * x * A2 - 2 * A1; HA[1] = 4 * x * y * A2; HA[2] = 4 * x * z * A2;
    HA[3] = 4 * y * y * A2 - 2 * A1; HA[4] = 4 * y * z * A2; HA[5] = 4 * z * z * A2 - 2 * A1;
    double* cols = out0 + sg.w;  // the segment's first column
    switch (bv.sh_type[sh]) {
      case 0: etab_shell<0>(x, y, z, A0, dA, HA, sg.y, sg.z, a, c, cols, tstride); break;
      case 1: etab_shell<1>(x, y, z, A0, dA, HA, sg.y, sg.z, a, c, cols, tstride); break;
      case -1: etab_shell<-1>(x, y, z, A0, dA, HA, sg.y, sg.z, a, c, cols, tstride); break;
      case -2: etab_shell<-2>(x, y, z, A0, dA, HA, sg.y, sg.z, a, c, cols, tstride); break;
      default: etab_shell<-3>(x, y, z, A0, dA, HA, sg.y, sg.z, a, c, cols, tstride); break;
    }
  }
}

void launch_skel_etab(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const double* bx,
                      const double* by, const double* bz, BasisView bv, const double* coef, double* etab) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_skel_etab, dim3((nslots + 3) / 4), dim3(256), 0, s, nslots, slot_batch, slot_tile, bx, by, bz,
                     bv, coef, etab);
  CHX_HIP(hipGetLastError());
}

// F[3 A + t] in the reference's order (column-major), from the internal Q^t
// (qp[t]) and P^t (qp[3 + t]), npad x npad row-major [mu][nu]:
//   F[row, col] = -( [atom(row) = A] N^t[mu][nu] + [atom(col) = A] N^t[nu][mu] ),
//   N^t[mu][nu] = Q^t[mu][nu] + P^t[nu][mu], mu = ref2int[row], nu = ref2int[col].
__global__ void k_fock_skel_fin(int n, int npad, int natoms, const double* __restrict__ qp,
                                const int* __restrict__ ref2int, const int* __restrict__ ao_atom,
                                double* __restrict__ F) {
  const size_t nn = (size_t)n * n, tot = (size_t)3 * natoms * nn, pp = (size_t)npad * npad;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(idx / nn), A = m / 3, t = m % 3;
    const size_t e = idx % nn;
    const int col = (int)(e / n), row = (int)(e % n);
    const size_t mu = ref2int[row], nu = ref2int[col];
    const double* Q = qp + (size_t)t * pp;
    const double* P = qp + (size_t)(3 + t) * pp;
    double v = 0.0;
    if (ao_atom[row] == A) v += Q[mu * npad + nu] + P[nu * npad + mu];
    if (ao_atom[col] == A) v += Q[nu * npad + mu] + P[mu * npad + nu];
    F[idx] = -v;
  }
}

void launch_fock_skel_fin(hipStream_t s, int n, int npad, int natoms, const double* qp, const int* ref2int,
                          const int* ao_atom, double* F) {
  const size_t tot = (size_t)3 * natoms * n * n;
  if (tot == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fock_skel_fin, dim3(blocks), dim3(256), 0, s, n, npad, natoms, qp, ref2int, ao_atom, F);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// XC energy Hessian, skeleton form with the grid held fixed
// (Restricted/Hess.cpp:78-90 -> getDensitySkeleton2, GridDensity.cpp:132-209,
// getEnergyHess, GridEnergy.cpp:32-45). With a = (A, t), b = (B, s):
//   H[a][b] = H1 + H2 + H3,
//   H1 = sum_g (2 c0^a rho^b + c^a . grad rho^b): the response coefficients of
//        the first-order skeleton density (rho^a, grad rho^a) contracted with rho^b;
//   H2 = 2 sum_{mu in A, nu in B} D_mu,nu K^{ts}_mu,nu,
//        K^{ts} = sum_g (d_t phi_mu E^s_nu + E^t_mu d_s phi_nu), E^s = a d_s phi + (c . grad) d_s phi;
//   H3 = [A = B] 2 sum_{mu in A} sum_g (2 a d_t d_s phi_mu X_mu
//        + sum_r c_r (d_r d_t d_s phi_mu X_mu + d_t d_s phi_mu Y^r_mu)).

// Third derivatives of the angular parts (xxx, xxy, xxz, xyy, xyz, xzz, yyy,
// yyz, yzz, zzz): zero up to d; constants for pure f (angular_h's order).
template <int TYPE>
__device__ __forceinline__ void angular_3(double (*k)[10]) {
  constexpr int n = TYPE == 0 ? 1 : TYPE == -2 ? 5 : TYPE == -3 ? 7 : 3;
#pragma unroll
  for (int m = 0; m < n; ++m)
#pragma unroll
    for (int j = 0; j < 10; ++j) k[m][j] = 0.0;
  if (TYPE == -3) {
    const double r15 = 3.872983346207417, c10 = 0.25 * 3.1622776601683795, c6 = 0.25 * 2.449489742783178;
    k[0][1] = 6 * c10; k[0][6] = -6 * c10;                    // y (3x^2 - y^2)
    k[1][4] = r15;                                            // xyz
    k[2][1] = -2 * c6; k[2][6] = -6 * c6; k[2][8] = 8 * c6;   // y (4z^2 - x^2 - y^2)
    k[3][9] = 6; k[3][2] = -3; k[3][7] = -3;                  // z (z^2 - 3/2 (x^2 + y^2))
    k[4][5] = 8 * c6; k[4][0] = -6 * c6; k[4][3] = -2 * c6;   // x (4z^2 - x^2 - y^2)
    k[5][2] = r15; k[5][7] = -r15;                            // (x^2 - y^2) z
    k[6][0] = 6 * c10; k[6][3] = -6 * c10;                    // x (x^2 - 3y^2)
  }
}

// H3 of components first..first+cnt-1 of one shell at this lane's point:
// acc[tu] += 2 (2 a phi_tu X + sum_r c_r (phi_rtu X + phi_tu Y^r)), tu over (xx, xy, xz, yy, yz, zz).
template <int TYPE>
__device__ __forceinline__ void hess3_shell(double x, double y, double z, double A0, const double* dA, const double* HA,
                                            const double* TA, int first, int cnt, const double* c1, const double* c2,
                                            double a, const double* c, double* acc) {
  constexpr int n = TYPE == 0 ? 1 : TYPE == -2 ? 5 : TYPE == -3 ? 7 : 3;
  double q[n], g[n][3], h[n][6], k3[n][10];
  angular_h<TYPE>(x, y, z, q, g, h);
  angular_3<TYPE>(k3);
  const int hi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  const int ti[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}, {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                           {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
  const int pt[6] = {0, 0, 0, 1, 1, 2}, pu[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int m = 0; m < n; ++m) {
    const int kk = m - first;
    if (kk < 0 || kk >= cnt) continue;
    const double X = c1[kk * kBatch];
    const double Y[3] = {c1[(kGCols + kk) * kBatch], c2[kk * kBatch], c2[(kGCols + kk) * kBatch]};
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const int t = pt[p], u = pu[p], tu = hi[t][u];
      const double ptu = HA[tu] * q[m] + dA[t] * g[m][u] + dA[u] * g[m][t] + A0 * h[m][tu];
      double v = 2 * a * ptu * X;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double prtu = TA[ti[r][t][u]] * q[m] + HA[hi[r][t]] * g[m][u] + HA[hi[r][u]] * g[m][t] +
                            HA[tu] * g[m][r] + dA[r] * h[m][tu] + dA[t] * h[m][hi[r][u]] +
                            dA[u] * h[m][hi[r][t]] + A0 * k3[m][ti[r][t][u]];
        v += c[r] * (prtu * X + ptu * Y[r]);
      }
      acc[p] += 2 * v;
    }
  }
}

// One wave per slot, lane = point: H3 summed over the wave per run of one
// atom's shells and added to H[3 A + t][3 A + u] (np = 3 natoms).
__global__ __launch_bounds__(256) void k_hess_diag(int nslots, const int* __restrict__ slot_batch,
                                                   const int* __restrict__ slot_tile, const int* __restrict__ slot_grp,
                                                   const int* __restrict__ slot_q, const double* __restrict__ bx,
                                                   const double* __restrict__ by, const double* __restrict__ bz,
                                                   BasisView bv, const double* __restrict__ coef,
                                                   const double* __restrict__ xz1, const double* __restrict__ xz2,
                                                   int np, double* __restrict__ H) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), g = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= nslots) return;
  const int b = slot_batch[slot], t = slot_tile[slot];
  const size_t p = (size_t)b * kBatch + g;
  const double px = bx[p], py = by[p], pz = bz[p];
  const double a = coef[kCoef * p];
  const double c[3] = {coef[kCoef * p + 1], coef[kCoef * p + 2], coef[kCoef * p + 3]};
  const size_t col0 = (size_t)slot_grp[slot] * 2 * kGCols * kBatch + (size_t)slot_q[slot] * kTile * kBatch + g;
  const double* X1 = xz1 + col0;
  const double* X2 = xz2 + col0;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int atom = -1;
  const int pt[6] = {0, 0, 0, 1, 1, 2}, pu[6] = {0, 1, 2, 1, 2, 2};
  auto flush = [&]() {
    if (atom < 0) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      acc[k] = 0.0;
      if (g == 0) {
        const int r0 = 3 * atom + pt[k], r1 = 3 * atom + pu[k];
        unsafeAtomicAdd(&H[(size_t)r0 * np + r1], v);
        if (r0 != r1) unsafeAtomicAdd(&H[(size_t)r1 * np + r0], v);
      }
    }
  };
  const int s0 = bv.tile_seg_off[t], s1 = bv.tile_seg_off[t + 1];
  for (int k = s0; k < s1; ++k) {
    const int4 sg = bv.seg[k];
    const int sh = sg.x;
    if (bv.sh_atom[sh] != atom) {
      flush();
      atom = bv.sh_atom[sh];
    }
    const double x = px - bv.sh_xyz[3 * sh], y = py - bv.sh_xyz[3 * sh + 1], z = pz - bv.sh_xyz[3 * sh + 2];
    const double r2 = x * x + y * y + z * z;
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0;
    for (int i = 0, po = bv.sh_poff[sh]; i < bv.sh_nprim[sh]; ++i) {
      const double e = bv.p_exp[po + i];
      const double v = bv.p_coef[po + i] * ::exp(-e * r2);
      A0 += v;
      A1 += e * v;
      A2 += e * e * v;
      A3 += e * e * e * v;
    }
    const double d[3] = {x, y, z};
    double dA[3], HA[6], TA[10];
    for (int r = 0; r < 3; ++r) dA[r] = -2 * d[r] * A1;
    HA[0] = 4 * x * x * A2 - 2 * A1; HA[1] = 4 * x * y * A2; HA[2] = 4 * x * z * A2;
    HA[3] = 4 * y * y * A2 - 2 * A1; HA[4] = 4 * y * z * A2; HA[5] = 4 * z * z * A2 - 2 * A1;
    {
      const int ix[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                             {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
#pragma unroll
      for (int j = 0; j < 10; ++j) {
        const int i0 = ix[j][0], i1 = ix[j][1], i2 = ix[j][2];
        TA[j] = -8 * d[i0] * d[i1] * d[i2] * A3 +
                4 * ((i0 == i1 ? d[i2] : 0.0) + (i0 == i2 ? d[i1] : 0.0) + (i1 == i2 ? d[i0] : 0.0)) * A2;
      }
    }
    const double* c1 = X1 + (size_t)sg.w * kBatch;
    const double* c2 = X2 + (size_t)sg.w * kBatch;
    switch (bv.sh_type[sh]) {
      case 0: hess3_shell<0>(x, y, z, A0, dA, HA, TA, sg.y, sg.z, c1, c2, a, c, acc); break;
      case 1: hess3_shell<1>(x, y, z, A0, dA, HA, TA, sg.y, sg.z, c1, c2, a, c, acc); break;
      case -1: hess3_shell<-1>(x, y, z, A0, dA, HA, TA, sg.y, sg.z, c1, c2, a, c, acc); break;
      case -2: hess3_shell<-2>(x, y, z, A0, dA, HA, TA, sg.y, sg.z, c1, c2, a, c, acc); break;
      default: hess3_shell<-3>(x, y, z, A0, dA, HA, TA, sg.y, sg.z, c1, c2, a, c, acc); break;
    }
  }
  flush();
}

// H1: block (16 a's x 16 b's, chunk of points), thread (a, b); the response
// coefficients of rho^a (rhos[a][p][0..3]) with fx (v2rho2, v2rhosigma,
// v2sigma2, v_sigma, grad rho) contracted with (2 rho^b, grad rho^b).
__global__ __launch_bounds__(256) void k_hess_fx(long npt, int np, const double* __restrict__ rhos,
                                                 const double* __restrict__ fx, const double* __restrict__ bw,
                                                 int family, long chunk, double* __restrict__ H) {
  const int a = blockIdx.x * 16 + (threadIdx.x >> 4), b = blockIdx.y * 16 + (threadIdx.x & 15);
  if (a >= np || b >= np) return;
  const long p0 = (long)blockIdx.z * chunk, p1 = p0 + chunk < npt ? p0 + chunk : npt;
  double e = 0.0;
  for (long p = p0; p < p1; ++p) {
    const double w = bw[p];
    if (w == 0.0) continue;
    const double* f = fx + (size_t)kFx * p;
    const double* u = rhos + ((size_t)a * npt + p) * 4;
    const double* v = rhos + ((size_t)b * npt + p) * 4;
    const double sx = 2.0 * (f[4] * u[1] + f[5] * u[2] + f[6] * u[3]);
    double t = w * (f[0] * u[0] + f[1] * sx) * v[0];  // 2 c0 rho^b
    if (family >= 1) {
      const double q = 2.0 * w * (f[1] * u[0] + f[2] * sx), r = 2.0 * w * f[3];
      t += (q * f[4] + r * u[1]) * v[1] + (q * f[5] + r * u[2]) * v[2] + (q * f[6] + r * u[3]) * v[3];
    }
    e += t;
  }
  unsafeAtomicAdd(&H[(size_t)a * np + b], e);
}

// H2: one thread per internal element (mu, nu) of D (npad^2, internal order),
// nine products added to H[3 A + t][3 B + s] (atom: atom of each internal AO, -1 padding).
__global__ void k_hess_dk(int npad, int np, const double* __restrict__ Dint, const double* __restrict__ rr,
                          const int* __restrict__ atom, double* __restrict__ H) {
  const size_t pp = (size_t)npad * npad;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < pp; idx += (size_t)gridDim.x * blockDim.x) {
    const int mu = (int)(idx / npad), nu = (int)(idx % npad);
    const int A = atom[mu], B = atom[nu];
    if (A < 0 || B < 0) continue;
    const double dd = Dint[idx];
    if (dd == 0.0) continue;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const double k = rr[(size_t)(3 * t + s) * pp + (size_t)mu * npad + nu] +
                         rr[(size_t)(3 * s + t) * pp + (size_t)nu * npad + mu];
        unsafeAtomicAdd(&H[(size_t)(3 * A + t) * np + 3 * B + s], 2.0 * dd * k);
      }
  }
}

void launch_hess_terms(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* slot_grp,
                       const int* slot_q, const double* bx, const double* by, const double* bz, BasisView bv,
                       const double* coef, const double* xz1, const double* xz2, long npt, const double* rhos,
                       const double* fx, const double* bw, int family, int npad, const double* Dint,
                       const double* rr, const int* atom, int np, double* H) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_hess_diag, dim3((nslots + 3) / 4), dim3(256), 0, s, nslots, slot_batch, slot_tile, slot_grp,
                       slot_q, bx, by, bz, bv, coef, xz1, xz2, np, H);
  const long chunk = 4096;
  const dim3 gf((np + 15) / 16, (np + 15) / 16, (unsigned)((npt + chunk - 1) / chunk));
  if (npt > 0) hipLaunchKernelGGL(k_hess_fx, gf, dim3(256), 0, s, npt, np, rhos, fx, bw, family, chunk, H);
  const size_t pp = (size_t)npad * npad;
  const unsigned blocks = (unsigned)std::min<size_t>((pp + 255) / 256, 65536);
  hipLaunchKernelGGL(k_hess_dk, dim3(blocks), dim3(256), 0, s, npad, np, Dint, rr, atom, H);
  CHX_HIP(hipGetLastError());
}

}  // namespace chx
