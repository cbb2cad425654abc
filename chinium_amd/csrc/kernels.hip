// HIP kernels of the XC path for gfx950 (CDNA4). FP64 plain form.
//
// Data layout in HBM
//   points   : batch-major slots, kBatch (64) slots per batch, SoA x/y/z/w;
//              padding slots repeat the batch's first point with w = 0.
//   basis    : internal AO order (atoms sorted along a space-filling order),
//              split into 32-AO tiles; per tile a list of shell segments.
//   D, V     : npad x npad, row-major, internal order (npad = nbasis rounded
//              up to 64). D is symmetric; V holds the upper 64x64 blocks.
//   partials : density per (batch, tile group): 4 x 64 doubles (rho, drho/2).
//   coef     : per slot (a, c_x, c_y, c_z) with the Vxc integrand
//              B_nu = a phi_nu + c . grad phi_nu, a = w v_rho / 2, c = 2 w v_sigma grad rho.
//
// The two dense contractions, X = Phi.D (density) and V = Phi^T B + B^T Phi
// (Vxc), run on v_mfma_f64_16x16x4_f64. AO values are recomputed from the
// shell data inside each kernel (never stored in HBM); the functional and
// the reductions run on the vector ALUs.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "internal.hpp"
#include "xc_functionals.hpp"

namespace chx {

// ---------------------------------------------------------------------------
// Becke partition (GridGeneration.cpp:136-153): w *= P_owner / sum_B P_B,
// P_B = prod_{C != B} s(mu_BC), s = (1 - p(p(p(mu))))/2, p(m) = 1.5m - 0.5m^3.
__global__ __launch_bounds__(256) void k_becke(long npts, const double* __restrict__ px, const double* __restrict__ py,
                                               const double* __restrict__ pz, double* __restrict__ w,
                                               const int* __restrict__ owner, int natoms,
                                               const double* __restrict__ axyz, const double* __restrict__ inv_rab) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const double x = px[i], y = py[i], z = pz[i];
  const int me = owner[i];
  double total = 0, mine = 0;
  for (int b = 0; b < natoms; ++b) {
    const double dxb = x - axyz[3 * b], dyb = y - axyz[3 * b + 1], dzb = z - axyz[3 * b + 2];
    const double rb = ::sqrt(dxb * dxb + dyb * dyb + dzb * dzb);
    double pb = 1.0;
    for (int c = 0; c < natoms && pb != 0.0; ++c) {
      if (c == b) continue;
      const double dxc = x - axyz[3 * c], dyc = y - axyz[3 * c + 1], dzc = z - axyz[3 * c + 2];
      const double rc = ::sqrt(dxc * dxc + dyc * dyc + dzc * dzc);
      double m = (rb - rc) / inv_rab[b * natoms + c];  // inv_rab holds R_BC (division as in s_p_mu)
      m = 1.5 * m - 0.5 * m * m * m;
      m = 1.5 * m - 0.5 * m * m * m;
      m = 1.5 * m - 0.5 * m * m * m;
      pb *= 0.5 * (1.0 - m);
    }
    total += pb;
    if (b == me) mine = pb;
  }
  w[i] *= mine / total;
}

void launch_becke(hipStream_t s, long npts, const double* px, const double* py, const double* pz, double* w,
                  const int* owner, int natoms, const double* atom_xyz, const double* inv_rab) {
  if (npts == 0) return;
  const int blocks = (int)((npts + 255) / 256);
  hipLaunchKernelGGL(k_becke, dim3(blocks), dim3(256), 0, s, npts, px, py, pz, w, owner, natoms, atom_xyz, inv_rab);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// AO screening: one wavefront per batch, one lane per slot. A tile is alive
// when some AO in it has |phi| or |grad phi| above thr at some real point.
__global__ __launch_bounds__(64) void k_screen(int ntiles, const double* __restrict__ bx, const double* __restrict__ by,
                                               const double* __restrict__ bz, const int* __restrict__ bn, BasisView bv,
                                               double thr, unsigned char* __restrict__ alive) {
  const int b = blockIdx.x;
  const int g = threadIdx.x;
  const bool real = g < bn[b];
  const double x = bx[b * kBatch + g], y = by[b * kBatch + g], z = bz[b * kBatch + g];
  for (int t = 0; t < ntiles; ++t) {
    bool any = false;
    for (int s = bv.tile_seg_off[t]; s < bv.tile_seg_off[t + 1] && !any; ++s) {
      const int4 sg = bv.seg[s];
      double v[7], gx[7], gy[7], gz[7];
      shell_eval<true>(bv, sg.x, x, y, z, v, gx, gy, gz);
      double mx = 0;
      for (int m = 0; m < 7; ++m) {
        if (m >= sg.y && m < sg.y + sg.z) {
          mx = fmax(mx, fabs(v[m]));
          mx = fmax(mx, fmax(fabs(gx[m]), fmax(fabs(gy[m]), fabs(gz[m]))));
        }
      }
      any = __ballot(real && mx > thr) != 0;
    }
    if (g == 0) alive[(size_t)b * ntiles + t] = any ? 1 : 0;
  }
}

void launch_screen(hipStream_t s, int nbatches, int ntiles, const double* bx, const double* by, const double* bz,
                   const int* bn, BasisView bv, double thr, unsigned char* alive) {
  if (nbatches == 0) return;
  hipLaunchKernelGGL(k_screen, dim3(nbatches), dim3(64), 0, s, ntiles, bx, by, bz, bn, bv, thr, alive);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// D (reference order, column-major n x n) -> internal order, npad x npad, zero padded.
__global__ void k_permute_d(int n, int npad, const double* __restrict__ D, const int* __restrict__ int2ref,
                            double* __restrict__ Dint) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npad * npad) return;
  const int r = (int)(idx / npad), c = (int)(idx % npad);
  const int rr = int2ref[r], cc = int2ref[c];
  Dint[idx] = (rr >= 0 && cc >= 0) ? D[(size_t)cc * n + rr] : 0.0;
}

void launch_permute_d(hipStream_t s, int n, int npad, const double* D, const int* int2ref, double* Dint) {
  const long tot = (long)npad * npad;
  hipLaunchKernelGGL(k_permute_d, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, npad, D, int2ref, Dint);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Density. One block (4 waves) per work group = (batch b, up to 4 alive AO
// tiles j as output columns). X[g, nu] = sum_{alive tiles i} A_i[g, :] D[i, nu]
// accumulates on MFMA over K = the batch's alive tiles, then
// MODE 0 (A = phi): rho[g] += sum_nu X phi_nu, drho[g]/2 += sum_nu X grad phi_nu
//   (4 components per point into part, [group][4][64]);
// MODE r = 1..3 (A = d_r phi): sum_nu X d_r phi_nu, the tau partial of
//   direction r (part points at component r - 1 of [group][3][64]).
// Wave w owns rows g = 16w..16w+15 and all 128 output columns. AO values are
// recomputed here (no tables in HBM).
struct DensSmem {
  union {
    struct {
      double A[kTile][kBatch];           // Phi_i, [mu][g]
      double B[kTile][kGroup * kTile];   // D block, [mu][nu]
    } k;
    double E[4][kBatch][kTile + 1];      // phi_j and grad phi_j, [c][g][nu]
  };
  double px[kBatch], py[kBatch], pz[kBatch];
};

template <int MODE>
__global__ __launch_bounds__(256) void k_density_f64(const int* __restrict__ grp_batch, const int* __restrict__ grp_tiles,
                                                     const int* __restrict__ grp_ntiles, const int* __restrict__ bt_off,
                                                     const int* __restrict__ bt_list, const double* __restrict__ bx,
                                                     const double* __restrict__ by, const double* __restrict__ bz,
                                                     BasisView bv, const double* __restrict__ Dint, int npad,
                                                     double* __restrict__ part) {
  __shared__ DensSmem sm;
  const int grp = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = grp_batch[grp];
  const int nt = grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = grp_tiles[grp * kGroup + q];

  if (tid < kBatch) {
    sm.px[tid] = bx[b * kBatch + tid];
    sm.py[tid] = by[b * kBatch + tid];
    sm.pz[tid] = bz[b * kBatch + tid];
  }
  dbl4 acc[2 * kGroup];
#pragma unroll
  for (int c = 0; c < 2 * kGroup; ++c) acc[c] = dbl4{0, 0, 0, 0};
  __syncthreads();

  const int lrow = lane & 15, lk = lane >> 4;
  for (int it = bt_off[b]; it < bt_off[b + 1]; ++it) {
    const int ti = bt_list[it];
    // zero the Phi tile, then fill the AOs this tile holds
    for (int e = tid; e < kTile * kBatch; e += 256) (&sm.k.A[0][0])[e] = 0.0;
    __syncthreads();
    const int s0 = bv.tile_seg_off[ti], s1 = bv.tile_seg_off[ti + 1];
    for (int item = tid; item < (s1 - s0) * kBatch; item += 256) {
      const int g = item % kBatch;
      const int4 sg = bv.seg[s0 + item / kBatch];
      double v[7], gx[7], gy[7], gz[7];
      shell_eval<MODE != 0>(bv, sg.x, sm.px[g], sm.py[g], sm.pz[g], v, gx, gy, gz);
      const double* a = MODE == 0 ? v : MODE == 1 ? gx : MODE == 2 ? gy : gz;
#pragma unroll
      for (int m = 0; m < 7; ++m)
        if (m < sg.z) sm.k.A[sg.w + m][g] = a[sg.y + m];
    }
    // D block rows of tile i, columns of the group's tiles
    for (int e = tid; e < kTile * kGroup * kTile; e += 256) {
      const int r = e / (kGroup * kTile), c = e % (kGroup * kTile);
      const int q = c / kTile;
      sm.k.B[r][c] = (q < nt) ? Dint[(size_t)(ti * kTile + r) * npad + jt[q] * kTile + (c % kTile)] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kTile / 4; ++ks) {
      const double a = sm.k.A[4 * ks + lk][16 * wave + lrow];
#pragma unroll
      for (int cb = 0; cb < 2 * kGroup; ++cb) {
        if (cb < 2 * nt) {
          const double bb = sm.k.B[4 * ks + lk][16 * cb + lrow];
          acc[cb] = mfma_f64(a, bb, acc[cb]);
        }
      }
    }
    __syncthreads();
  }

  // Epilogue: contract X with phi_j and grad phi_j (MODE 0) or d_r phi_j
  // (MODE r), tile by tile.
  constexpr int NOUT = MODE == 0 ? 4 : 1;
  double s[4][4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) s[c][r] = 0.0;
#pragma unroll
  for (int q = 0; q < kGroup; ++q) {
    if (q < nt) {
      for (int e = tid; e < 4 * kBatch * (kTile + 1); e += 256) (&sm.E[0][0][0])[e] = 0.0;
      __syncthreads();
      const int tj = jt[q];
      const int s0 = bv.tile_seg_off[tj], s1 = bv.tile_seg_off[tj + 1];
      for (int item = tid; item < (s1 - s0) * kBatch; item += 256) {
        const int g = item % kBatch;
        const int4 sg = bv.seg[s0 + item / kBatch];
        double v[7], gx[7], gy[7], gz[7];
        shell_eval<true>(bv, sg.x, sm.px[g], sm.py[g], sm.pz[g], v, gx, gy, gz);
#pragma unroll
        for (int m = 0; m < 7; ++m)
          if (m < sg.z) {
            sm.E[0][g][sg.w + m] = v[sg.y + m];
            sm.E[1][g][sg.w + m] = gx[sg.y + m];
            sm.E[2][g][sg.w + m] = gy[sg.y + m];
            sm.E[3][g][sg.w + m] = gz[sg.y + m];
          }
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const dbl4 x = acc[2 * q + h];
        const int col = 16 * h + lrow;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int g = 16 * wave + 4 * r + lk;
#pragma unroll
          for (int c = 0; c < NOUT; ++c) s[c][r] += x[r] * sm.E[MODE == 0 ? c : MODE][g][col];
        }
      }
      __syncthreads();
    }
  }
  // Reduce over the 16 lanes that share rows, then lane k writes (c = k/4, r = k%4).
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
  constexpr int grp_stride = (MODE == 0 ? 4 : 3) * kBatch;
  if (c < NOUT) part[(size_t)grp * grp_stride + c * kBatch + 16 * wave + 4 * r + lk] = out;
}

void launch_density_f64(hipStream_t s, int mode, int ngroups, const int* grp_batch, const int* grp_tiles,
                        const int* grp_ntiles, const int* bt_off, const int* bt_list, const double* bx,
                        const double* by, const double* bz, BasisView bv, const double* Dint, int npad, double* part) {
  if (ngroups == 0) return;
#define CHX_DF64(M)                                                                                       \
  hipLaunchKernelGGL(k_density_f64<M>, dim3(ngroups), dim3(256), 0, s, grp_batch, grp_tiles, grp_ntiles, \
                     bt_off, bt_list, bx, by, bz, bv, Dint, npad, part);
  switch (mode) {
    case 0: CHX_DF64(0); break;
    case 1: CHX_DF64(1); break;
    case 2: CHX_DF64(2); break;
    default: CHX_DF64(3); break;
  }
#undef CHX_DF64
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Functional: one thread per slot. Sums the density partials of the batch's
// work groups, evaluates the functional and writes the Vxc integrand
// coefficients; block partial sums of E = sum w rho eps and N = sum w rho.
__device__ __forceinline__ void functional_slot(int slot, int nslots, const int* __restrict__ grp_off,
                                                const double* __restrict__ part, const double* __restrict__ part_tau,
                                                const double* __restrict__ bw, const int* __restrict__ codes,
                                                int ncodes, int family, int want_e, int want_v,
                                                double* __restrict__ rho_out, double* __restrict__ coef, double& e_acc,
                                                double& n_acc) {
  e_acc = n_acc = 0.0;
  if (slot >= nslots) return;
  const int b = slot / kBatch, g = slot % kBatch;
  double rho = 0, gx = 0, gy = 0, gz = 0, tau = 0;
  for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
    const double* p = part + (size_t)k * 4 * kBatch;
    rho += p[g];
    gx += p[kBatch + g];
    gy += p[2 * kBatch + g];
    gz += p[3 * kBatch + g];
    if (family >= 2) {
      const double* q = part_tau + (size_t)k * 3 * kBatch;
      tau += q[g] + q[kBatch + g] + q[2 * kBatch + g];
    }
  }
  gx *= 2; gy *= 2; gz *= 2;
  tau *= 0.5;
  const double w = bw[slot];
  double sigma = family >= 1 ? gx * gx + gy * gy + gz * gz : 0.0;
  double eps, vr, vs, vt;
  eval_codes(codes, ncodes, rho, sigma, tau, &eps, &vr, &vs, &vt);
  rho_out[slot] = rho;
  if (want_v) {
    coef[kCoef * (size_t)slot + 0] = 0.5 * w * vr;
    const double f = family >= 1 ? 2.0 * w * vs : 0.0;
    coef[kCoef * (size_t)slot + 1] = f * gx;
    coef[kCoef * (size_t)slot + 2] = f * gy;
    coef[kCoef * (size_t)slot + 3] = f * gz;
    coef[kCoef * (size_t)slot + 4] = family >= 2 ? 0.25 * w * vt : 0.0;
  }
  if (want_e) e_acc = w * rho * eps;
  n_acc = w * rho;
}

// Sums of (e, n) over a 256-thread block: within each wave by xor shuffles,
// then the four wave sums in a fixed order. Thread 0 holds the result.
__device__ __forceinline__ void block_sum2(double& e, double& n) {
  __shared__ double red[2][4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    n += __shfl_xor(n, off, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wave] = e;
    red[1][wave] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    e = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    n = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

__global__ __launch_bounds__(256) void k_functional(int nslots, const int* __restrict__ grp_off,
                                                    const double* __restrict__ part, const double* __restrict__ part_tau,
                                                    const double* __restrict__ bw, const int* __restrict__ codes,
                                                    int ncodes, int family, int want_e, int want_v,
                                                    double* __restrict__ rho_out, double* __restrict__ coef,
                                                    double* __restrict__ en_part) {
  double e, n;
  functional_slot(blockIdx.x * 256 + threadIdx.x, nslots, grp_off, part, part_tau, bw, codes, ncodes, family, want_e,
                  want_v, rho_out, coef, e, n);
  block_sum2(e, n);
  if (threadIdx.x == 0) {
    en_part[2 * blockIdx.x] = e;
    en_part[2 * blockIdx.x + 1] = n;
  }
}

void launch_functional(hipStream_t s, int nbatches, const int* grp_off, const double* part, const double* part_tau,
                       const double* bw, const int* codes, int ncodes, int family, int want_e, int want_v,
                       double* rho_out, double* coef, double* en_part, int en_blocks) {
  const int nslots = nbatches * kBatch;
  hipLaunchKernelGGL(k_functional, dim3(en_blocks), dim3(256), 0, s, nslots, grp_off, part, part_tau, bw, codes, ncodes,
                     family, want_e, want_v, rho_out, coef, en_part);
  CHX_HIP(hipGetLastError());
}

// The functional of several molecules in one launch (packs): block g belongs
// to molecule m with fm[m].blk0 <= g < fm[m + 1].blk0 and runs that molecule's
// block g - blk0 exactly as k_functional does. Each molecule's buffers are
// addressed as base + byte offset, base a kernel argument, so that every access
// compiles to a global (not flat) memory instruction. Then one block per
// molecule reduces its block partial sums into en[2 m], en[2 m + 1].
__global__ __launch_bounds__(256) void k_functional_multi(const char* __restrict__ base,
                                                          const FnMolecule* __restrict__ fm, int nmol,
                                                          const int* __restrict__ codes, int ncodes, int family) {
  const int gb = blockIdx.x;
  int lo = 0, hi = nmol;  // the last molecule whose blk0 <= gb
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (fm[mid].blk0 <= gb) lo = mid; else hi = mid;
  }
  const FnMolecule m = fm[lo];
  const int blk = gb - m.blk0;
  double e, n;
  functional_slot(blk * 256 + threadIdx.x, m.nslots, reinterpret_cast<const int*>(base + m.grp_off),
                  reinterpret_cast<const double*>(base + m.part), reinterpret_cast<const double*>(base + m.part_tau),
                  reinterpret_cast<const double*>(base + m.w), codes, ncodes, family, 1, 1,
                  reinterpret_cast<double*>(const_cast<char*>(base) + m.rho),
                  reinterpret_cast<double*>(const_cast<char*>(base) + m.coef), e, n);
  block_sum2(e, n);
  if (threadIdx.x == 0) {
    double* en_part = reinterpret_cast<double*>(const_cast<char*>(base) + m.en_part);
    en_part[2 * blk] = e;
    en_part[2 * blk + 1] = n;
  }
}

__global__ __launch_bounds__(256) void k_reduce_en_multi(const char* __restrict__ base,
                                                         const FnMolecule* __restrict__ fm, double* __restrict__ en) {
  const FnMolecule m = fm[blockIdx.x];
  const double* en_part = reinterpret_cast<const double*>(base + m.en_part);
  const int nparts = fm[blockIdx.x + 1].blk0 - m.blk0;
  double e = 0, n = 0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    e += en_part[2 * i];
    n += en_part[2 * i + 1];
  }
  block_sum2(e, n);
  if (threadIdx.x == 0) {
    en[2 * m.slot_en] = e;
    en[2 * m.slot_en + 1] = n;
  }
}

void launch_functional_multi(hipStream_t s, const char* base, const FnMolecule* fm, int nmol, int nblocks,
                             const int* codes, int ncodes, int family, double* en) {
  if (nmol == 0) return;
  hipLaunchKernelGGL(k_functional_multi, dim3(nblocks), dim3(256), 0, s, base, fm, nmol, codes, ncodes, family);
  CHX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_en_multi, dim3(nmol), dim3(256), 0, s, base, fm, en);
  CHX_HIP(hipGetLastError());
}

// Spin-polarised functional, one thread per slot: the density partials of
// each spin's density matrix (part / part_tau: alpha, part_b / part_tau_b:
// beta), the polarised functional, and the Vxc integrand coefficients of each
// spin (coef: alpha, coef_b: beta):
//   a_u = w v_rho_u / 2,  c_u = w (2 v_sigma_uu grad rho_u + v_sigma_ab grad rho_u'),
//   t_u = w v_tau_u / 4 (meta-GGA),
// the scaled sigma derivatives of src/Grid/GridFock.cpp:21-30 and the tau
// term of :32-40 at nspin = 2. Block partial sums of E = sum w (rho_a + rho_b) eps and N.
__global__ __launch_bounds__(256) void k_functional_pol(int nslots, const int* __restrict__ grp_off,
                                                        const double* __restrict__ part,
                                                        const double* __restrict__ part_b,
                                                        const double* __restrict__ part_tau,
                                                        const double* __restrict__ part_tau_b,
                                                        const double* __restrict__ bw, const int* __restrict__ codes,
                                                        int ncodes, int family, int want_e, int want_v,
                                                        double* __restrict__ rho_out, double* __restrict__ coef,
                                                        double* __restrict__ coef_b, double* __restrict__ en_part) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  double e_acc = 0, n_acc = 0;
  if (slot < nslots) {
    const int b = slot / kBatch, g = slot % kBatch;
    double ra = 0, rb = 0, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0}, ta = 0, tb = 0;
    for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
      const double* p = part + (size_t)k * 4 * kBatch;
      const double* q = part_b + (size_t)k * 4 * kBatch;
      ra += p[g];
      rb += q[g];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        ga[r] += p[(r + 1) * kBatch + g];
        gb[r] += q[(r + 1) * kBatch + g];
      }
      if (family >= 2) {
        const double* pt = part_tau + (size_t)k * 3 * kBatch;
        const double* qt = part_tau_b + (size_t)k * 3 * kBatch;
        ta += pt[g] + pt[kBatch + g] + pt[2 * kBatch + g];
        tb += qt[g] + qt[kBatch + g] + qt[2 * kBatch + g];
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      ga[r] *= 2;
      gb[r] *= 2;
    }
    ta *= 0.5;
    tb *= 0.5;
    const double w = bw[slot];
    double saa = 0, sab = 0, sbb = 0;
    if (family >= 1) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        saa += ga[r] * ga[r];
        sab += ga[r] * gb[r];
        sbb += gb[r] * gb[r];
      }
    }
    double eps, vr[2], vs[3], vt[2] = {0, 0};
    if (family >= 2)
      eval_codes_mpol(codes, ncodes, ra, rb, saa, sab, sbb, ta, tb, &eps, vr, vs, vt);
    else
      eval_codes_pol(codes, ncodes, ra, rb, saa, sab, sbb, &eps, vr, vs);
    rho_out[slot] = ra + rb;
    if (want_v) {
      const double fa = family >= 1 ? w : 0.0;
      coef[kCoef * (size_t)slot + 0] = 0.5 * w * vr[0];
      coef_b[kCoef * (size_t)slot + 0] = 0.5 * w * vr[1];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        coef[kCoef * (size_t)slot + 1 + r] = fa * (2.0 * vs[0] * ga[r] + vs[1] * gb[r]);
        coef_b[kCoef * (size_t)slot + 1 + r] = fa * (2.0 * vs[2] * gb[r] + vs[1] * ga[r]);
      }
      coef[kCoef * (size_t)slot + 4] = family >= 2 ? 0.25 * w * vt[0] : 0.0;
      coef_b[kCoef * (size_t)slot + 4] = family >= 2 ? 0.25 * w * vt[1] : 0.0;
    }
    if (want_e) e_acc = w * (ra + rb) * eps;
    n_acc = w * (ra + rb);
  }
  block_sum2(e_acc, n_acc);
  if (threadIdx.x == 0) {
    en_part[2 * blockIdx.x] = e_acc;
    en_part[2 * blockIdx.x + 1] = n_acc;
  }
}

void launch_functional_pol(hipStream_t s, int nbatches, const int* grp_off, const double* part, const double* part_b,
                           const double* part_tau, const double* part_tau_b, const double* bw, const int* codes,
                           int ncodes, int family, int want_e, int want_v, double* rho_out, double* coef,
                           double* coef_b, double* en_part, int en_blocks) {
  const int nslots = nbatches * kBatch;
  hipLaunchKernelGGL(k_functional_pol, dim3(en_blocks), dim3(256), 0, s, nslots, grp_off, part, part_b, part_tau,
                     part_tau_b, bw, codes, ncodes, family, want_e, want_v, rho_out, coef, coef_b, en_part);
  CHX_HIP(hipGetLastError());
}

// Spin-polarised functional at arbitrary points, libxc XC_POLARIZED layout;
// meta-GGA (tau non-null): tau[2n] in, vtau[2n] out.
__global__ void k_eval_functional_pol(long n, const int* __restrict__ codes, int ncodes, const double* __restrict__ rho,
                                      const double* __restrict__ sigma, const double* __restrict__ tau,
                                      double* __restrict__ exc, double* __restrict__ vrho, double* __restrict__ vsigma,
                                      double* __restrict__ vtau) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double eps, vr[2], vs[3], vt[2] = {0, 0};
  const double saa = sigma ? sigma[3 * i] : 0.0, sab = sigma ? sigma[3 * i + 1] : 0.0, sbb = sigma ? sigma[3 * i + 2] : 0.0;
  if (tau)
    eval_codes_mpol(codes, ncodes, rho[2 * i], rho[2 * i + 1], saa, sab, sbb, tau[2 * i], tau[2 * i + 1], &eps, vr, vs,
                    vt);
  else
    eval_codes_pol(codes, ncodes, rho[2 * i], rho[2 * i + 1], saa, sab, sbb, &eps, vr, vs);
  exc[i] = eps;
  vrho[2 * i] = vr[0];
  vrho[2 * i + 1] = vr[1];
  if (vsigma)
    for (int k = 0; k < 3; ++k) vsigma[3 * i + k] = vs[k];
  if (vtau) {
    vtau[2 * i] = vt[0];
    vtau[2 * i + 1] = vt[1];
  }
}

void launch_eval_functional_pol(hipStream_t s, long n, const int* codes, int ncodes, const double* rho,
                                const double* sigma, const double* tau, double* exc, double* vrho, double* vsigma,
                                double* vtau) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_functional_pol, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, codes, ncodes, rho,
                     sigma, tau, exc, vrho, vsigma, vtau);
  CHX_HIP(hipGetLastError());
}

// Functional values and derivatives at arbitrary (rho, sigma, tau), as libxc's
// xc_*_exc_vxc returns them (src/ExchangeCorrelation.cpp:98,122,157).
__global__ void k_eval_functional(long n, const int* __restrict__ codes, int ncodes, const double* __restrict__ rho,
                                  const double* __restrict__ sigma, const double* __restrict__ tau,
                                  double* __restrict__ exc, double* __restrict__ vrho, double* __restrict__ vsigma,
                                  double* __restrict__ vtau) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e, vr, vs, vt;
  eval_codes(codes, ncodes, rho[i], sigma[i], tau[i], &e, &vr, &vs, &vt);
  exc[i] = e;
  vrho[i] = vr;
  vsigma[i] = vs;
  vtau[i] = vt;
}

void launch_eval_functional(hipStream_t s, long n, const int* codes, int ncodes, const double* rho, const double* sigma,
                            const double* tau, double* exc, double* vrho, double* vsigma, double* vtau) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_functional, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, codes, ncodes, rho,
                     sigma, tau, exc, vrho, vsigma, vtau);
  CHX_HIP(hipGetLastError());
}

// ---- Response path (closed shell, LDA/GGA). Replaces Evaluate("f")
// (src/ExchangeCorrelation.cpp:101-104,126-133), SubGrid::getDensityU
// (src/Grid/GridDensity.cpp:47-77) and SubGrid::getFockU<u_t>
// (src/Grid/GridFock.cpp:99-160) for nspin = 1.

// f_xc at the grid's ground-state density: per slot (v2rho2, v2rhosigma,
// v2sigma2, v_sigma, grad rho), from the density partials.
__global__ __launch_bounds__(256) void k_functional_f(int nslots, const int* __restrict__ grp_off,
                                                      const double* __restrict__ part, const int* __restrict__ codes,
                                                      int ncodes, int family, double* __restrict__ fx) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= nslots) return;
  const int b = slot / kBatch, g = slot % kBatch;
  double rho = 0, gx = 0, gy = 0, gz = 0;
  for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
    const double* p = part + (size_t)k * 4 * kBatch;
    rho += p[g];
    gx += p[kBatch + g];
    gy += p[2 * kBatch + g];
    gz += p[3 * kBatch + g];
  }
  gx *= 2; gy *= 2; gz *= 2;
  const double sigma = family >= 1 ? gx * gx + gy * gy + gz * gz : 0.0;
  double vs, frr, frs, fss;
  eval_codes_f(codes, ncodes, rho, sigma, &vs, &frr, &frs, &fss);
  double* f = fx + (size_t)kFx * slot;
  f[0] = frr;
  f[1] = family >= 1 ? frs : 0.0;
  f[2] = family >= 1 ? fss : 0.0;
  f[3] = family >= 1 ? vs : 0.0;
  f[4] = gx;
  f[5] = gy;
  f[6] = gz;
}

// rho^x and grad rho^x of one perturbed density matrix, per slot, from its partials.
__global__ __launch_bounds__(256) void k_point_rho(int nslots, const int* __restrict__ grp_off,
                                                   const double* __restrict__ part, double* __restrict__ rhou) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= nslots) return;
  const int b = slot / kBatch, g = slot % kBatch;
  double rho = 0, gx = 0, gy = 0, gz = 0;
  for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
    const double* p = part + (size_t)k * 4 * kBatch;
    rho += p[g];
    gx += p[kBatch + g];
    gy += p[2 * kBatch + g];
    gz += p[3 * kBatch + g];
  }
  rhou[4 * (size_t)slot + 0] = rho;
  rhou[4 * (size_t)slot + 1] = 2 * gx;
  rhou[4 * (size_t)slot + 2] = 2 * gy;
  rhou[4 * (size_t)slot + 3] = 2 * gz;
}

// Coefficients of the response Fock matrix for one perturbation, in the
// layout of k_functional's (a phi + c . grad phi, then M + M^T):
//   sigma^x = 2 grad rho . grad rho^x,
//   a = 1/2 w (v2rho2 rho^x + v2rhosigma sigma^x),
//   c = w (2 (v2rhosigma rho^x + v2sigma2 sigma^x) grad rho + 2 v_sigma grad rho^x).
__global__ __launch_bounds__(256) void k_response_coef(int nslots, const double* __restrict__ rhou,
                                                       const double* __restrict__ fx, const double* __restrict__ bw,
                                                       int family, double* __restrict__ coef) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= nslots) return;
  const double* f = fx + (size_t)kFx * slot;
  const double* u = rhou + 4 * (size_t)slot;
  const double w = bw[slot];
  const double sx = 2.0 * (f[4] * u[1] + f[5] * u[2] + f[6] * u[3]);
  double* c = coef + kCoef * (size_t)slot;
  c[0] = 0.5 * w * (f[0] * u[0] + f[1] * sx);
  if (family >= 1) {
    const double q = 2.0 * w * (f[1] * u[0] + f[2] * sx), r = 2.0 * w * f[3];
    c[1] = q * f[4] + r * u[1];
    c[2] = q * f[5] + r * u[2];
    c[3] = q * f[6] + r * u[3];
  } else {
    c[1] = c[2] = c[3] = 0.0;
  }
  c[4] = 0.0;
}

// ---- Spin-polarised response (nspin = 2, LDA/GGA): Evaluate("f") with the
// functional polarised, getDensityU / getFockU<u_t> at nspin = 2.

// Polarised f_xc at the grid's ground-state density (alpha partials in part,
// beta in part_b): per slot kFxPol values, eval_codes_fpol's 18 (vsigma,
// v2rho2, v2rhosigma, v2sigma2), then grad rho_a (3) and grad rho_b (3).
__global__ __launch_bounds__(256) void k_functional_fpol(int nslots, const int* __restrict__ grp_off,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ part_b,
                                                         const int* __restrict__ codes, int ncodes, int family,
                                                         double* __restrict__ fx) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= nslots) return;
  const int b = slot / kBatch, g = slot % kBatch;
  double ra = 0, rb = 0, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
  for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
    const double* p = part + (size_t)k * 4 * kBatch;
    const double* q = part_b + (size_t)k * 4 * kBatch;
    ra += p[g];
    rb += q[g];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      ga[r] += p[(r + 1) * kBatch + g];
      gb[r] += q[(r + 1) * kBatch + g];
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    ga[r] = family >= 1 ? 2 * ga[r] : 0.0;
    gb[r] = family >= 1 ? 2 * gb[r] : 0.0;
  }
  const double saa = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2];
  const double sab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
  const double sbb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
  double* f = fx + (size_t)kFxPol * slot;
  eval_codes_fpol(codes, ncodes, ra, rb, saa, sab, sbb, f);
  if (family < 1) {  // LDA: no sigma terms
#pragma unroll
    for (int k = 0; k < 3; ++k) f[k] = 0.0;
#pragma unroll
    for (int k = 6; k < 18; ++k) f[k] = 0.0;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    f[18 + r] = ga[r];
    f[21 + r] = gb[r];
  }
}

// Coefficients of spin u's response matrix for one perturbation (rho^x_s,
// grad rho^x_s of both spins per slot in rhou_a / rhou_b), in k_functional's
// layout (a phi + c . grad phi, then M + M^T):
//   sigma^x = (2 grad rho_a . grad rho^x_a, grad rho^x_a . grad rho_b + grad rho_a . grad rho^x_b,
//              2 grad rho_b . grad rho^x_b),
//   dv_rho_u = sum_v f_{rho_u rho_v} rho^x_v + sum_x f_{rho_u sigma_x} sigma^x_x,
//   dv_sigma_x = sum_w f_{rho_w sigma_x} rho^x_w + sum_y f_{sigma_x sigma_y} sigma^x_y,
//   a = 1/2 w dv_rho_u,
//   c = w (2 dv_sigma_uu grad rho_u + dv_sigma_ab grad rho_o + 2 v_sigma_uu grad rho^x_u
//          + v_sigma_ab grad rho^x_o)
// (the derivative of the polarised ground-state potential along D^x).
__global__ __launch_bounds__(256) void k_response_coef_pol(int nslots, const double* __restrict__ rhou_a,
                                                           const double* __restrict__ rhou_b,
                                                           const double* __restrict__ fx,
                                                           const double* __restrict__ bw, int family, int spin,
                                                           double* __restrict__ coef) {
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= nslots) return;
  const double* f = fx + (size_t)kFxPol * slot;
  const double* xa = rhou_a + 4 * (size_t)slot;
  const double* xb = rhou_b + 4 * (size_t)slot;
  const double* ga = f + 18;
  const double* gb = f + 21;
  const double w = bw[slot];
  double sx[3] = {0, 0, 0};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    sx[0] += 2.0 * ga[r] * xa[1 + r];
    sx[1] += xa[1 + r] * gb[r] + ga[r] * xb[1 + r];
    sx[2] += 2.0 * gb[r] * xb[1 + r];
  }
  const double rx[2] = {xa[0], xb[0]};
  const int u = spin;
  // f layout: vsigma 0..2, v2rho2 3..5 (aa, ab, bb), v2rhosigma 6 + 3 w + x, v2sigma2 12.. (upper triangle)
  double dvr = f[3 + u] * rx[0] + f[4 + u] * rx[1];
#pragma unroll
  for (int x = 0; x < 3; ++x) dvr += f[6 + 3 * u + x] * sx[x];
  double* c = coef + kCoef * (size_t)slot;
  c[0] = 0.5 * w * dvr;
  if (family >= 1) {
    const int fss[3][3] = {{12, 13, 14}, {13, 15, 16}, {14, 16, 17}};
    double dvs[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      dvs[x] = f[6 + x] * rx[0] + f[9 + x] * rx[1];
#pragma unroll
      for (int y = 0; y < 3; ++y) dvs[x] += f[fss[x][y]] * sx[y];
    }
    const int uu = u == 0 ? 0 : 2;
    const double* gu = u == 0 ? ga : gb;
    const double* go = u == 0 ? gb : ga;
    const double* xu = u == 0 ? xa : xb;
    const double* xo = u == 0 ? xb : xa;
#pragma unroll
    for (int r = 0; r < 3; ++r)
      c[1 + r] = w * (2.0 * dvs[uu] * gu[r] + dvs[1] * go[r] + 2.0 * f[uu] * xu[1 + r] + f[1] * xo[1 + r]);
  } else {
    c[1] = c[2] = c[3] = 0.0;
  }
  c[4] = 0.0;
}

// Polarised f_xc at arbitrary points (libxc XC_POLARIZED layout): rho[2n],
// sigma[3n] (null: LDA) -> v2rho2[3n], v2rhosigma[6n], v2sigma2[6n].
__global__ void k_eval_functional_fpol(long n, const int* __restrict__ codes, int ncodes,
                                       const double* __restrict__ rho, const double* __restrict__ sigma,
                                       double* __restrict__ v2rho2, double* __restrict__ v2rhosigma,
                                       double* __restrict__ v2sigma2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[18];
  eval_codes_fpol(codes, ncodes, rho[2 * i], rho[2 * i + 1], sigma ? sigma[3 * i] : 0.0, sigma ? sigma[3 * i + 1] : 0.0,
                  sigma ? sigma[3 * i + 2] : 0.0, f);
#pragma unroll
  for (int k = 0; k < 3; ++k) v2rho2[3 * i + k] = f[3 + k];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v2rhosigma[6 * i + k] = sigma ? f[6 + k] : 0.0;
    v2sigma2[6 * i + k] = sigma ? f[12 + k] : 0.0;
  }
}

void launch_functional_fpol(hipStream_t s, int nbatches, const int* grp_off, const double* part, const double* part_b,
                            const int* codes, int ncodes, int family, double* fx) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_functional_fpol, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, grp_off, part, part_b,
                     codes, ncodes, family, fx);
  CHX_HIP(hipGetLastError());
}

void launch_response_coef_pol(hipStream_t s, int nbatches, const double* rhou_a, const double* rhou_b,
                              const double* fx, const double* bw, int family, int spin, double* coef) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_response_coef_pol, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, rhou_a, rhou_b, fx, bw,
                     family, spin, coef);
  CHX_HIP(hipGetLastError());
}

void launch_eval_functional_fpol(hipStream_t s, long n, const int* codes, int ncodes, const double* rho,
                                 const double* sigma, double* v2rho2, double* v2rhosigma, double* v2sigma2) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_functional_fpol, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, n, codes, ncodes,
                     rho, sigma, v2rho2, v2rhosigma, v2sigma2);
  CHX_HIP(hipGetLastError());
}

// f_xc at arbitrary (rho, sigma), as libxc's xc_{lda,gga}_fxc returns it.
__global__ void k_eval_functional_f(long n, const int* __restrict__ codes, int ncodes, const double* __restrict__ rho,
                                    const double* __restrict__ sigma, double* __restrict__ v2rho2,
                                    double* __restrict__ v2rhosigma, double* __restrict__ v2sigma2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double vs;
  eval_codes_f(codes, ncodes, rho[i], sigma[i], &vs, v2rho2 + i, v2rhosigma + i, v2sigma2 + i);
}

void launch_functional_f(hipStream_t s, int nbatches, const int* grp_off, const double* part, const int* codes,
                         int ncodes, int family, double* fx) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_functional_f, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, grp_off, part, codes, ncodes,
                     family, fx);
  CHX_HIP(hipGetLastError());
}

void launch_point_rho(hipStream_t s, int nbatches, const int* grp_off, const double* part, double* rhou) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_point_rho, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, grp_off, part, rhou);
  CHX_HIP(hipGetLastError());
}

void launch_response_coef(hipStream_t s, int nbatches, const double* rhou, const double* fx, const double* bw,
                          int family, double* coef) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_response_coef, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, rhou, fx, bw, family, coef);
  CHX_HIP(hipGetLastError());
}

void launch_eval_functional_f(hipStream_t s, long n, const int* codes, int ncodes, const double* rho,
                              const double* sigma, double* v2rho2, double* v2rhosigma, double* v2sigma2) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_functional_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, codes, ncodes, rho,
                     sigma, v2rho2, v2rhosigma, v2sigma2);
  CHX_HIP(hipGetLastError());
}

// Deterministic final sum of the block partials (one block): E, N, and a
// cleared slot en[2] for the mixed form's energy correction, which the
// potential's build fills.
__global__ __launch_bounds__(256) void k_reduce_en(int nparts, const double* __restrict__ en_part, double* __restrict__ en) {
  double e = 0, n = 0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    e += en_part[2 * i];
    n += en_part[2 * i + 1];
  }
  block_sum2(e, n);
  if (threadIdx.x == 0) {
    en[0] = e;
    en[1] = n;
    en[2] = 0.0;
  }
}

void launch_reduce_en(hipStream_t s, int nparts, const double* en_part, double* en) {
  hipLaunchKernelGGL(k_reduce_en, dim3(1), dim3(256), 0, s, nparts, en_part, en);
  CHX_HIP(hipGetLastError());
}

__global__ void k_fill(double* p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill(hipStream_t s, double* p, size_t n, double v) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  CHX_HIP(hipGetLastError());
}

// v_rho supplied by the caller (SAP guess): coefficients a = w v_rho / 2, c = 0.
__global__ void k_vrho_coef(long nslots, const double* __restrict__ vrho_w, double* __restrict__ coef) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  coef[kCoef * i] = 0.5 * vrho_w[i];
  coef[kCoef * i + 1] = 0;
  coef[kCoef * i + 2] = 0;
  coef[kCoef * i + 3] = 0;
  coef[kCoef * i + 4] = 0;
}

void launch_vrho_coef(hipStream_t s, long nslots, const double* vrho_w, double* coef) {
  hipLaunchKernelGGL(k_vrho_coef, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s, nslots, vrho_w, coef);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Vxc. One block per work item (super-tiles I <= J of 64 AOs, range of
// batches where both are alive). V_IJ += sum_g P_I^T B_J + B_I^T P_J on
// MFMA, K = points, 32 points per LDS stage. Pass 0 has P = phi and
// B = a phi + c . grad phi; meta-GGA (NC = 4) adds passes r = 1..3 with
// P = d_r phi and B = t d_r phi (t = w v_tau / 4, so the pair of products
// gives w v_tau / 2 d_r phi_mu d_r phi_nu). AO values are recomputed for each
// pass (no tables in HBM). Wave w owns output rows 16w..16w+15 and all 64
// columns. The block result is added to V with hardware FP64 atomics once per
// work item.
constexpr int kVxcK = 32;
constexpr int kSupAO = kSuper * kTile;  // 64

struct VxcSmem {
  double PI[kVxcK][kSupAO + 1];
  double BI[kVxcK][kSupAO + 1];
  double PJ[kVxcK][kSupAO + 1];
  double BJ[kVxcK][kSupAO + 1];
  double px[kVxcK], py[kVxcK], pz[kVxcK], ca[kVxcK], cx[kVxcK], cy[kVxcK], cz[kVxcK], ct[kVxcK];
};

__device__ __forceinline__ void vxc_fill_super(VxcSmem& sm, double (*P)[kSupAO + 1], double (*B)[kSupAO + 1],
                                               int sup, const unsigned char* alive_b, const BasisView& bv, int tid,
                                               int pass) {
  for (int half = 0; half < kSuper; ++half) {
    const int t = sup * kSuper + half;
    if (!alive_b[t]) continue;
    const int s0 = bv.tile_seg_off[t], s1 = bv.tile_seg_off[t + 1];
    for (int item = tid; item < (s1 - s0) * kVxcK; item += 256) {
      const int g = item % kVxcK;
      const int4 sg = bv.seg[s0 + item / kVxcK];
      double v[7], gx[7], gy[7], gz[7];
      shell_eval<true>(bv, sg.x, sm.px[g], sm.py[g], sm.pz[g], v, gx, gy, gz);
      const double a = sm.ca[g], cx = sm.cx[g], cy = sm.cy[g], cz = sm.cz[g], t = sm.ct[g];
      const double* d = pass == 1 ? gx : pass == 2 ? gy : gz;
#pragma unroll
      for (int m = 0; m < 7; ++m)
        if (m < sg.z) {
          const int col = half * kTile + sg.w + m;
          if (pass == 0) {
            const double ph = v[sg.y + m];
            P[g][col] = ph;
            B[g][col] = a * ph + cx * gx[sg.y + m] + cy * gy[sg.y + m] + cz * gz[sg.y + m];
          } else {
            P[g][col] = d[sg.y + m];
            B[g][col] = t * d[sg.y + m];
          }
        }
    }
  }
}

template <int NC>
__global__ __launch_bounds__(256) void k_vxc_f64(const int4* __restrict__ work, const int* __restrict__ pair_batches,
                                                 const unsigned char* __restrict__ alive, int ntiles,
                                                 const double* __restrict__ bx, const double* __restrict__ by,
                                                 const double* __restrict__ bz, const double* __restrict__ coef,
                                                 BasisView bv, double* __restrict__ Vint, int npad) {
  __shared__ VxcSmem sm;
  const int4 wk = work[blockIdx.x];
  const int I = wk.x, J = wk.y;
  const bool diag = (I == J);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lk = lane >> 4;
  dbl4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = dbl4{0, 0, 0, 0};

  for (int it = wk.z; it < wk.w; ++it) {
    const int b = pair_batches[it];
    const unsigned char* ab = alive + (size_t)b * ntiles;
    for (int h = 0; h < kBatch / kVxcK; ++h) {
      // load the points and their coefficients
      if (tid < kVxcK) {
        const size_t slot = (size_t)b * kBatch + h * kVxcK + tid;
        sm.px[tid] = bx[slot];
        sm.py[tid] = by[slot];
        sm.pz[tid] = bz[slot];
        sm.ca[tid] = coef[kCoef * slot];
        sm.cx[tid] = coef[kCoef * slot + 1];
        sm.cy[tid] = coef[kCoef * slot + 2];
        sm.cz[tid] = coef[kCoef * slot + 3];
        sm.ct[tid] = NC == 4 ? coef[kCoef * slot + 4] : 0.0;
      }
      for (int pass = 0; pass < NC; ++pass) {
      // zero the tiles (AOs outside the alive tiles stay zero), then fill them
      for (int e = tid; e < kVxcK * (kSupAO + 1); e += 256) {
        (&sm.PI[0][0])[e] = 0.0;
        (&sm.BI[0][0])[e] = 0.0;
        if (!diag) {
          (&sm.PJ[0][0])[e] = 0.0;
          (&sm.BJ[0][0])[e] = 0.0;
        }
      }
      __syncthreads();
      vxc_fill_super(sm, sm.PI, sm.BI, I, ab, bv, tid, pass);
      if (!diag) vxc_fill_super(sm, sm.PJ, sm.BJ, J, ab, bv, tid, pass);
      __syncthreads();
      double (*PJ)[kSupAO + 1] = diag ? sm.PI : sm.PJ;
      double (*BJ)[kSupAO + 1] = diag ? sm.BI : sm.BJ;
#pragma unroll
      for (int ks = 0; ks < kVxcK / 4; ++ks) {
        const int g = 4 * ks + lk;
        const double a1 = sm.PI[g][16 * wave + lrow];
        const double a2 = sm.BI[g][16 * wave + lrow];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          acc[cb] = mfma_f64(a1, BJ[g][16 * cb + lrow], acc[cb]);
          acc[cb] = mfma_f64(a2, PJ[g][16 * cb + lrow], acc[cb]);
        }
      }
      __syncthreads();
      }
    }
  }
  // C layout: acc[cb][r] = V[row = 16w + 4r + lk][col = 16 cb + lrow]
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = I * kSupAO + 16 * wave + 4 * r + lk;
      const int nu = J * kSupAO + 16 * cb + lrow;
      unsafeAtomicAdd(&Vint[(size_t)mu * npad + nu], acc[cb][r]);
    }
}

void launch_vxc_f64(hipStream_t s, int ncomp, int nwork, const int4* work, const int* pair_batches,
                    const unsigned char* alive, int ntiles, const double* bx, const double* by, const double* bz,
                    const double* coef, BasisView bv, double* Vint, int npad) {
  if (nwork == 0) return;
  if (ncomp == 4)
    hipLaunchKernelGGL(k_vxc_f64<4>, dim3(nwork), dim3(256), 0, s, work, pair_batches, alive, ntiles, bx, by, bz,
                       coef, bv, Vint, npad);
  else
    hipLaunchKernelGGL(k_vxc_f64<1>, dim3(nwork), dim3(256), 0, s, work, pair_batches, alive, ntiles, bx, by, bz,
                       coef, bv, Vint, npad);
  CHX_HIP(hipGetLastError());
}

// Internal upper 64x64 blocks -> full symmetric V in reference order, column-major.
__global__ void k_finalize_v(int n, int npad, const double* __restrict__ Vint, const int* __restrict__ ref2int,
                             double* __restrict__ V) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int col = (int)(idx / n), row = (int)(idx % n);
  const int mu = ref2int[row], nu = ref2int[col];
  const double v = (mu / kSupAO <= nu / kSupAO) ? Vint[(size_t)mu * npad + nu] : Vint[(size_t)nu * npad + mu];
  V[idx] = v;
}

void launch_finalize_v(hipStream_t s, int n, int npad, const double* Vint, const int* ref2int, double* V) {
  const long tot = (long)n * n;
  hipLaunchKernelGGL(k_finalize_v, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, npad, Vint, ref2int, V);
  CHX_HIP(hipGetLastError());
}

}  // namespace chx

namespace chx {
// Self-test of the FP64 MFMA operand/result layout the kernels assume:
// A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, C[4r + l/16][l%16] in reg r.
__global__ void k_selftest(double* out) {
  __shared__ double A[16][4], B[4][16];
  const int l = threadIdx.x;
  A[l % 16][l / 16] = 1.0 + 0.25 * l;
  B[l / 16][l % 16] = 0.5 - 0.125 * l;
  __syncthreads();
  dbl4 c = {0, 0, 0, 0};
  c = mfma_f64(A[l % 16][l / 16], B[l / 16][l % 16], c);
  double err = 0;
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * r + l / 16, j = l % 16;
    double ref = 0;
    for (int k = 0; k < 4; ++k) ref += A[i][k] * B[k][j];
    err = fmax(err, fabs(c[r] - ref));
  }
  for (int o = 32; o > 0; o >>= 1) err = fmax(err, __shfl_xor(err, o, 64));
  if (l == 0) out[0] = err;
}

void launch_selftest(hipStream_t s, double* out) {
  hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, s, out);
  CHX_HIP(hipGetLastError());
}
}  // namespace chx
