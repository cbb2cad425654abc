// Mixed-precision form: the two dense contractions on split-BF16 MFMA.
//
// Every FP32 operand x is held as a pair of bf16 numbers, hi = bf16(x) and
// lo = bf16(x - hi), so that hi + lo carries 16 significant bits. A product
// x*y is formed as hi*hi' + hi*lo' + lo*hi' on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation (the dropped lo*lo' term is below 2^-16 |x y|). The
// functional, the E_xc / N_e sums and the final V accumulation stay FP64.
//
// The MFMA operands are cached at Grid::getAO in this split form (as the
// reference caches its AO tables once, src/Grid/Grid.cpp:22-33), nT = 1
// component (phi), or 4 (phi, d/dx, d/dy, d/dz) for meta-GGA, in two layouts:
//   AOg [c][g=0..63][hl][mu=0..31]  (K = mu contiguous: A of the density X = Phi D)
//   AOt [c][hl][mu=0..31][g=0..63]  (K = g contiguous: A of Vxc M = Phi^T B)
// 8 KB per component and layout per slot. Values needed outside the MFMA
// operands (the density epilogue, the Vxc integrand B) are recomputed in FP32
// from the basis: the hardware exp makes that cheaper than reading them back.
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
// 16 bytes of operand data; a native vector (HIP's u32x4 is a struct whose
// copies the compiler may keep in private memory)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMxRow = 2 * kTile;             // one AOg row: 32 hi then 32 lo
constexpr int kMxT = 2 * kTile * kBatch;      // elements per component of AOg, AOt and Bt
constexpr int kMxSup = 4;                     // tiles per Vxc super-tile (128 AOs)
// LDS operand tiles: rows of 4 16-byte chunks (32 bf16), chunk c of row r
// stored at position c ^ ((r >> 2) & 3). A wave's 16-byte accesses then hit
// distinct banks both when 16 lanes read one chunk of 16 consecutive rows
// (MFMA operands) and when they write all 4 chunks of 4 rows (staging).
__device__ __forceinline__ int swz(int r, int c) { return c ^ ((r >> 2) & 3); }

// float -> bf16, round to nearest even: the hardware conversion
// (v_cvt_pk_bf16_f32). k_bf16_check compares it with the bit-level rule
// f2bf_rne on the device.
__device__ __forceinline__ u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }
__device__ __forceinline__ u16 f2bf_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ void split_bf(float x, u16& hi, u16& lo) {
  hi = f2bf(x);
  lo = f2bf(x - bf2f(hi));
}

// Device check of the conversion: n pseudo-random finite floats spread over
// many binades; counts values where the hardware result differs from f2bf_rne.
__global__ void k_bf16_check(unsigned n, unsigned* __restrict__ bad) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = i * 2654435761u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  const unsigned e = 64u + (h >> 25);  // biased exponent 64..191
  const float f = __uint_as_float((h & 0x807fffffu) | (e << 23));
  if (f2bf(f) != f2bf_rne(f)) atomicAdd(bad, 1u);
}

__global__ void k_fill_u32(unsigned* p, size_t n, unsigned v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_u32(hipStream_t s, unsigned* p, size_t n, unsigned v) {
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  CHX_HIP(hipGetLastError());
}

void launch_bf16_check(hipStream_t s, unsigned n, unsigned* bad) {
  hipLaunchKernelGGL(k_bf16_check, dim3((n + 255) / 256), dim3(256), 0, s, n, bad);
  CHX_HIP(hipGetLastError());
}
__device__ __forceinline__ bf16x8 as_bf(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ f32x4 mfma3(u32x4 ah, u32x4 al, u32x4 bh, u32x4 bl, f32x4 c) {
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
                                                     BasisView bv, int nT, int sstride, u16* __restrict__ cache,
                                                     const int* __restrict__ slot_core, u16* __restrict__ cache3) {
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
  for (int e = tid; e < nT * kBatch * kTile; e += 256) {
    const int c = e / (kBatch * kTile), g = (e / kTile) % kBatch, mu = e % kTile;
    u16 hi, lo;
    split_bf(E[c][g][mu], hi, lo);
    out[(c * kBatch + g) * kMxRow + mu] = hi;
    out[(c * kBatch + g) * kMxRow + kTile + mu] = lo;
  }
  u16* outt = out + nT * kMxT;
  for (int e = tid; e < nT * kTile * kBatch; e += 256) {
    const int c = e / (kTile * kBatch), mu = (e / kBatch) % kTile, g = e % kBatch;
    u16 hi, lo;
    split_bf(E[c][g][mu], hi, lo);
    outt[c * kMxT + mu * kBatch + g] = hi;
    outt[c * kMxT + kTile * kBatch + mu * kBatch + g] = lo;
  }
  const int cs = slot_core ? slot_core[s] : -1;
  if (cs >= 0) {  // core slot: the third term x - hi - lo, AOg order without the hl plane
    u16* o3 = cache3 + (size_t)cs * nT * kBatch * kTile;
    for (int e = tid; e < nT * kBatch * kTile; e += 256) {
      const int c = e / (kBatch * kTile), g = (e / kTile) % kBatch, mu = e % kTile;
      const float x = E[c][g][mu];
      u16 hi, lo;
      split_bf(x, hi, lo);
      o3[e] = f2bf((x - bf2f(hi)) - bf2f(lo));
    }
  }
}

void launch_ao_cache_mx(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* tile_nao,
                        const double* bx, const double* by, const double* bz, BasisView bv, int nT, int sstride,
                        unsigned short* cache, const int* slot_core, unsigned short* cache3) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_ao_cache_mx, dim3(nslots), dim3(256), 0, s, slot_batch, slot_tile, tile_nao, bx, by, bz, bv, nT,
                     sstride, cache, slot_core, cache3);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// D (reference order, column-major) -> internal order, split: Ds[hl][mu][nu].
// R (if set): what the split leaves of D, D - hi - lo (its FP32 rounding
// included), at D's own position. The products see hi + lo; sum V . R, with V
// the potential of this density, is the first-order energy of the remainder
// (dE/dD = V), which the energy takes back when the potential is built
// (k_finalize_corr).
__global__ void k_split_d(int n, int npad, const double* __restrict__ D, const int* __restrict__ int2ref,
                          u16* __restrict__ Ds, float* __restrict__ R) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long np2 = (long)npad * npad;
  if (idx >= np2) return;
  const int r = (int)(idx / npad), c = (int)(idx % npad);
  const int rr = int2ref[r], cc = int2ref[c];
  const bool real = rr >= 0 && cc >= 0;
  const double v = real ? D[(size_t)cc * n + rr] : 0.0;
  u16 hi, lo;
  const float x = (float)v;
  split_bf(x, hi, lo);
  Ds[idx] = hi;
  Ds[np2 + idx] = lo;
  if (R && real) R[(size_t)cc * n + rr] = (float)(v - (double)bf2f(hi) - (double)bf2f(lo));
}

void launch_split_d(hipStream_t s, int n, int npad, const double* D, const int* int2ref, unsigned short* Ds,
                    float* R) {
  const long tot = (long)npad * npad;
  hipLaunchKernelGGL(k_split_d, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, npad, D, int2ref, Ds, R);
  CHX_HIP(hipGetLastError());
}

// V = M + M^T (reference order, as k_finalize_sym) and the block's partial sum
// of V . R, the split remainder of D: grid-stride over the n^2 elements with a
// fixed block count, so the partials have one fixed order of summation.
constexpr int kCorrBlocks = 1024;
int corr_blocks(int n) { return (int)std::min<long>(kCorrBlocks, ((long)n * n + 255) / 256); }

__device__ __forceinline__ double block_sum_d(double v) {
  __shared__ double red[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_finalize_corr(int n, int npad, const double* __restrict__ Mint,
                                                       const int* __restrict__ ref2int, double* __restrict__ V,
                                                       const float* __restrict__ R, double* __restrict__ cpart) {
  const long tot = (long)n * n;
  double acc = 0.0;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < tot; idx += (long)gridDim.x * 256) {
    const int col = (int)(idx / n), row = (int)(idx % n);
    const int mu = ref2int[row], nu = ref2int[col];
    const double v = Mint[(size_t)mu * npad + nu] + Mint[(size_t)nu * npad + mu];
    V[idx] = v;
    acc += v * (double)R[idx];
  }
  acc = block_sum_d(acc);
  if (threadIdx.x == 0) cpart[blockIdx.x] = acc;
}

void launch_finalize_corr(hipStream_t s, int n, int npad, const double* Mint, const int* ref2int, double* V,
                          const float* R, double* cpart) {
  hipLaunchKernelGGL(k_finalize_corr, dim3((unsigned)corr_blocks(n)), dim3(256), 0, s, n, npad, Mint, ref2int, V, R,
                     cpart);
  CHX_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_reduce_corr(int nparts, const double* __restrict__ cpart,
                                                     double* __restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) acc += cpart[i];
  acc = block_sum_d(acc);
  if (threadIdx.x == 0) *out = acc;
}

void launch_reduce_corr(hipStream_t s, int nparts, const double* cpart, double* out) {
  hipLaunchKernelGGL(k_reduce_corr, dim3(1), dim3(256), 0, s, nparts, cpart, out);
  CHX_HIP(hipGetLastError());
}

// One block per molecule of a pack lane: its partials (at base + cpart) added
// to its energy, after the functional's sums have written it.
__global__ __launch_bounds__(256) void k_corr_multi(const char* __restrict__ base, const FnMolecule* __restrict__ fm,
                                                    double* __restrict__ en) {
  const FnMolecule m = fm[blockIdx.x];
  const double* cpart = reinterpret_cast<const double*>(base + m.cpart);
  double acc = 0.0;
  for (int i = threadIdx.x; i < m.ncpart; i += 256) acc += cpart[i];
  acc = block_sum_d(acc);
  if (threadIdx.x == 0) en[2 * m.slot_en] += acc;
}

void launch_corr_multi(hipStream_t s, const char* base, const FnMolecule* fm, int nmol, double* en) {
  if (nmol <= 0) return;
  hipLaunchKernelGGL(k_corr_multi, dim3((unsigned)nmol), dim3(256), 0, s, base, fm, en);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Angular parts of one shell in FP32 (same conventions as shell_eval:
// src/Grid/AO/{S,CartP,PureP,PureD,PureF}.hpp). phi_m = R Q_m and
// grad phi_m = grad R Q_m + R grad Q_m with R = sum_p c_p exp(-a_p r^2), so an
// operation sees the radial part once per shell (op.begin / op.end) and, per
// component, op.comp(column, Q, dQ/dx, dQ/dy, dQ/dz). The shell type is a
// template argument: each type's polynomials are straight-line code, the
// constant derivatives fold away, and components first..first+cnt-1 are
// emitted with compile-time indices.
template <int TYPE, class Op>
__device__ __forceinline__ void emit_shell_f32(float x, float y, float z, int first, int cnt, int col, Op& op) {
  constexpr int n = TYPE == 0 ? 1 : TYPE == -2 ? 5 : TYPE == -3 ? 7 : 3;
  float q[n], qx[n], qy[n], qz[n];
  const float r3 = 1.7320508075688772f, r6 = 2.449489742783178f, r10 = 3.1622776601683795f, r15 = 3.872983346207417f;
  if (TYPE == 0) {
    q[0] = 1.f; qx[0] = qy[0] = qz[0] = 0.f;
  } else if (TYPE == 1) {  // Cartesian p: x, y, z
    q[0] = x; qx[0] = 1.f; qy[0] = 0.f; qz[0] = 0.f;
    q[1] = y; qx[1] = 0.f; qy[1] = 1.f; qz[1] = 0.f;
    q[2] = z; qx[2] = 0.f; qy[2] = 0.f; qz[2] = 1.f;
  } else if (TYPE == -1) {  // pure p: y, z, x
    q[0] = y; qx[0] = 0.f; qy[0] = 1.f; qz[0] = 0.f;
    q[1] = z; qx[1] = 0.f; qy[1] = 0.f; qz[1] = 1.f;
    q[2] = x; qx[2] = 1.f; qy[2] = 0.f; qz[2] = 0.f;
  } else if (TYPE == -2) {
    q[0] = r3 * x * y;                     qx[0] = r3 * y; qy[0] = r3 * x;  qz[0] = 0.f;
    q[1] = r3 * y * z;                     qx[1] = 0.f;    qy[1] = r3 * z;  qz[1] = r3 * y;
    q[2] = z * z - 0.5f * (x * x + y * y); qx[2] = -x;     qy[2] = -y;      qz[2] = 2.f * z;
    q[3] = r3 * x * z;                     qx[3] = r3 * z; qy[3] = 0.f;     qz[3] = r3 * x;
    q[4] = 0.5f * r3 * (x * x - y * y);    qx[4] = r3 * x; qy[4] = -r3 * y; qz[4] = 0.f;
  } else {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float c10 = 0.25f * r10, c6 = 0.25f * r6;
    q[0] = c10 * y * (3.f * xx - yy);     qx[0] = c10 * 6.f * x * y;  qy[0] = c10 * 3.f * (xx - yy);             qz[0] = 0.f;
    q[1] = r15 * x * y * z;               qx[1] = r15 * y * z;        qy[1] = r15 * x * z;                       qz[1] = r15 * x * y;
    q[2] = c6 * y * (4.f * zz - xx - yy); qx[2] = -c6 * 2.f * x * y;  qy[2] = c6 * (4.f * zz - xx - 3.f * yy);  qz[2] = c6 * 8.f * y * z;
    q[3] = z * (zz - 1.5f * (xx + yy));   qx[3] = -3.f * x * z;       qy[3] = -3.f * y * z;                      qz[3] = 3.f * zz - 1.5f * (xx + yy);
    q[4] = c6 * x * (4.f * zz - xx - yy); qx[4] = c6 * (4.f * zz - 3.f * xx - yy); qy[4] = -c6 * 2.f * x * y;   qz[4] = c6 * 8.f * x * z;
    q[5] = 0.5f * r15 * (xx - yy) * z;    qx[5] = r15 * x * z;        qy[5] = -r15 * y * z;                      qz[5] = 0.5f * r15 * (xx - yy);
    q[6] = c10 * x * (xx - 3.f * yy);     qx[6] = c10 * 3.f * (xx - yy); qy[6] = -c10 * 6.f * x * y;             qz[6] = 0.f;
  }
  if (first == 0 && cnt == n) {  // the usual case: the whole shell lies in this tile
#pragma unroll
    for (int m = 0; m < n; ++m) op.comp(col + m, q[m], qx[m], qy[m], qz[m]);
  } else {
#pragma unroll
    for (int m = 0; m < n; ++m) {
      const int k = m - first;  // component m of the shell is AO col + k of the tile
      if (k >= 0 && k < cnt) op.comp(col + k, q[m], qx[m], qy[m], qz[m]);
    }
  }
}

// R = sum_p c_p exp(-a_p r^2) and sum_p a_p c_p exp(-a_p r^2) over NP primitives.
template <int NP>
__device__ __forceinline__ void radial_f32(const SegRec& rec, float r2, float& A0, float& A1) {
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const float e = rec.c[p] * __expf(-rec.e[p] * r2);
    A0 += e;
    A1 += rec.e[p] * e;
  }
}

// One wave walks one tile at 64 points, lane = point. Every segment record
// comes in through scalar loads and the shell loop is uniform across the
// wave: no divergence, no LDS, no block barrier. The offset from the centre
// is taken in FP64 (points and centres are tens of bohr apart, the offsets
// that matter are small), everything after it in FP32 with the hardware exp.
// The basis tables are read through the constant address space: they never
// change while a kernel runs, and so are always read with scalar loads (a
// kernel that also issues LDS DMA loads would otherwise get vector loads).

__device__ __forceinline__ SegRec load_segrec(const SegRec* recs, int k) {
  using cvec = const __attribute__((address_space(4))) u32x4;
  static_assert(sizeof(SegRec) == 8 * sizeof(u32x4), "SegRec is 128 bytes");
  u32x4 w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = ((cvec*)(recs + k))[i];
  SegRec r;
  __builtin_memcpy(&r, w, sizeof(SegRec));
  return r;
}

// PREFETCH: the next record is loaded while the current one is evaluated
// (for an Op that reads no LDS, whose waits would take the record load along).
template <bool PREFETCH = false, class Op>
__device__ __forceinline__ void walk_tile_f32(const BasisView& bv, int t, double px, double py, double pz, Op& op) {
  using cint = const __attribute__((address_space(4))) int;
  const int s0 = ((cint*)bv.tile_seg_off)[t], s1 = ((cint*)bv.tile_seg_off)[t + 1];
  SegRec nxt;
  if (PREFETCH && s0 < s1) nxt = load_segrec(bv.segrec, s0);
  for (int k = s0; k < s1; ++k) {
    SegRec rec;
    if (PREFETCH) {
      rec = nxt;
      nxt = load_segrec(bv.segrec, k + 1 < s1 ? k + 1 : k);
    } else {
      rec = load_segrec(bv.segrec, k);
    }
    const float x = (float)(px - rec.ax), y = (float)(py - rec.ay), z = (float)(pz - rec.az);
    const float r2 = x * x + y * y + z * z;
    float A0 = 0.f, A1 = 0.f;
    switch ((rec.meta >> 8) & 0xff) {  // primitive count: straight-line code per count
      case 1: radial_f32<1>(rec, r2, A0, A1); break;
      case 2: radial_f32<2>(rec, r2, A0, A1); break;
      case 3: radial_f32<3>(rec, r2, A0, A1); break;
      case 4: radial_f32<4>(rec, r2, A0, A1); break;
      case 5: radial_f32<5>(rec, r2, A0, A1); break;
      case 6: radial_f32<6>(rec, r2, A0, A1); break;
      case 7: radial_f32<7>(rec, r2, A0, A1); break;
      case 8: radial_f32<8>(rec, r2, A0, A1); break;
      default: radial_f32<kSegPrim>(rec, r2, A0, A1); break;  // unused slots hold c = 0
    }
    // R = A0, grad R = -2 r A1
    op.begin(A0, -2.f * x * A1, -2.f * y * A1, -2.f * z * A1);
    const int first = (rec.meta >> 16) & 0xff, cnt = rec.meta >> 24;
    switch ((int)(signed char)(rec.meta & 0xff)) {
      case 0: emit_shell_f32<0>(x, y, z, first, cnt, rec.col, op); break;
      case 1: emit_shell_f32<1>(x, y, z, first, cnt, rec.col, op); break;
      case -1: emit_shell_f32<-1>(x, y, z, first, cnt, rec.col, op); break;
      case -2: emit_shell_f32<-2>(x, y, z, first, cnt, rec.col, op); break;
      default: emit_shell_f32<-3>(x, y, z, first, cnt, rec.col, op); break;
    }
    op.end();
  }
}

// Density epilogue on one tile: sum_nu X[g][nu] phi_nu and sum_nu X[g][nu] grad phi_nu
// (MODE 0), or sum_nu X_r[g][nu] d_r phi_nu (MODE r), as
// R sum X Q + grad R sum X Q and R sum X grad Q per shell.
typedef float f32x32 __attribute__((ext_vector_type(32)));
template <int MODE>
struct DensEpi {
  f32x32 xv;  // X of this tile at this lane's point, in registers (uniform column index: no memory access)
  float o[4] = {0.f, 0.f, 0.f, 0.f};
  float R, Rx, Ry, Rz, S0, Sx, Sy, Sz;
  // xq: X^T of this tile at this lane's point, xq[col * (kBatch + 1)]
  __device__ void load(const float* xq) {
#pragma unroll
    for (int c = 0; c < kTile; ++c) xv[c] = xq[c * (kBatch + 1)];
  }
  __device__ void begin(float a, float ax, float ay, float az) {
    R = a; Rx = ax; Ry = ay; Rz = az;
    S0 = Sx = Sy = Sz = 0.f;
  }
  __device__ void comp(int col, float q, float qx, float qy, float qz) {
    const float x = xv[col];
    S0 += x * q;
    if (MODE == 0 || MODE == 1) Sx += x * qx;
    if (MODE == 0 || MODE == 2) Sy += x * qy;
    if (MODE == 0 || MODE == 3) Sz += x * qz;
  }
  __device__ void end() {
    if (MODE == 0) {
      o[0] += R * S0;
      o[1] += Rx * S0 + R * Sx;
      o[2] += Ry * S0 + R * Sy;
      o[3] += Rz * S0 + R * Sz;
    } else if (MODE == 1) {
      o[0] += Rx * S0 + R * Sx;
    } else if (MODE == 2) {
      o[0] += Ry * S0 + R * Sy;
    } else {
      o[0] += Rz * S0 + R * Sz;
    }
  }
};

// ---------------------------------------------------------------------------
// Density. Block = (batch, group of <= 4 alive tiles as output columns), 4
// waves in a 2 x 2 grid: wave (wr, wc) owns rows g = 32 wr..+31 and columns
// 64 wc..+63 of X = Phi_b D_b (64 x 128). K loop over the batch's alive tiles
// (32 AOs = one MFMA K step). A K step's operands (A: 64 points x 32 AOs, D:
// 128 columns x 32 AOs, hi and lo) are 24 fragments of 1 KB that the waves'
// DMA loads (global_load_lds_dwordx4: lane i -> bytes 16 i of the fragment)
// bring from memory straight into LDS (no VGPR staging), coalesced, in a
// swizzled order from which the MFMA lanes read without bank conflicts.
// Per step: wait for the step's fragments, read the wave's 12 into registers,
// barrier, start the DMA of the next step into the same buffer, run the
// wave's 24 MFMAs while it lands (one 24 KB buffer, so three blocks share a
// CU and cover each other's barriers). Epilogue: X goes to LDS, wave q loads
// tile q's columns at its lane's point into registers and contracts them with
// phi / grad phi recomputed in FP32.
constexpr int kDxFrag = 64;  // u32x4 per fragment
constexpr int kDxA = 8;      // A fragments per step: [row block][hl]
constexpr int kDxB = 16;     // D fragments per step: [column block][hl]
constexpr int kDxA3 = 4;     // P3: third-term A fragments per step: [row block]
template <bool P3>
struct DensMxSmem {
  union {
    u32x4 F[(kDxA + kDxB + (P3 ? kDxA3 : 0)) * kDxFrag];  // one K step's fragments
    float X[kGroup * kTile][kBatch + 1];                          // epilogue: X^T [nu][g]
  };
  float red[kGroup][4][kBatch];
  double px[kBatch], py[kBatch], pz[kBatch];
  int tl[kMxMaxTiles];  // the batch's alive tiles (K steps)
};

// MODE 0: A = phi; epilogue rho and grad rho / 2 (4 components into part).
// MODE r = 1..3 (A = d phi / d x_r, epilogue sum X_r d_r phi, the tau partial
// of direction r) is the single-direction form of k_density_mx_tau below.
// P3 (core batches only): the AO operand carries a third bf16 term (x - hi -
// lo, 24 significant bits in all) from the core slots' side cache, and a
// product takes the five terms hi hi, lo hi, hi lo, lo2 hi and lo lo: the AO
// values at FP32 level where the density is largest, and D as hi + lo exactly
// (its remainder D - hi - lo, the same in every batch, is taken back in the
// energy to first order once the potential is built).
// Blocks walk the core group list; the others run the two-term kernel.
// The grid's descriptor comes by value (a kernel argument: every pointer in it
// is known to address global memory, and uniform table reads become scalar
// loads; a descriptor read from device memory would leave them flat).
// r: the block's item (after the XCD-aware remap) within the molecule's launch.
template <int MODE, bool P3>
__device__ __forceinline__ void density_mx_body(const MolDev& m, int r) {
  using lds_ptr = __attribute__((address_space(3))) void*;
  constexpr int NOUT = MODE == 0 ? 4 : 1;
  constexpr int grp_stride = (MODE == 0 ? 4 : 3) * kBatch;  // part: [group][4][64]; tau partials: [group][3][64]
  __shared__ DensMxSmem<P3> sm;
  const int grp = P3 ? m.core_grps[r] : (MODE == 0 && m.main_grps) ? m.main_grps[r] : r;
  const u16* __restrict__ cache = m.cache + (size_t)MODE * kMxT;
  const u16* __restrict__ Ds = m.Ds;
  const size_t sstride = (size_t)m.sstride;
  const int npad = m.npad;
  double* __restrict__ part = MODE == 0 ? m.part : m.part_tau + (MODE - 1) * kBatch;
  const BasisView bv = m.bv;  // descriptor fields in registers, read once
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  const int b = m.grp_batch[grp];
  const int nt = m.grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = m.grp_tiles[grp * kGroup + q];
  const int it0 = m.bt_off[b], nk = m.bt_off[b + 1] - it0;  // the batch's slots it0.. are its K steps
  if (tid < kBatch) {
    sm.px[tid] = m.bx[b * kBatch + tid];
    sm.py[tid] = m.by[b * kBatch + tid];
    sm.pz[tid] = m.bz[b * kBatch + tid];
  }
  for (int j = tid; j < nk; j += 256) sm.tl[j] = m.bt_list[it0 + j];
  // The wave's fragments are f = wave + 4 j. A (j = 0, 1): row block f >> 1,
  // hl f & 1. D (j = 2..5, f' = f - 8): column block c = f' >> 1 (tile c >> 1,
  // half c & 1), hl f' & 1; columns past the group's tiles read a present
  // tile's rows (their X columns are never used). A fragment is 16 rows of
  // 64 bytes; DMA lane i fetches row i >> 2, 16-byte chunk (i & 3) ^ sw(row)
  // (four neighbouring lanes read one contiguous 64-byte run: coalesced), and
  // the MFMA lane that needs (row, chunk) reads it back at frag_idx(row, chunk)
  // (the XOR keeps those LDS reads free of bank conflicts). P3 adds the AO
  // operand's third term: A3 row block = wave.
  const int drow = lane >> 2, dch = (lane & 3) ^ ((drow >> 2) & 3);
  const int fidx = 4 * lrow + (lk ^ ((lrow >> 2) & 3));  // this lane's MFMA operand in a fragment
  int offa[2], offb[4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int f = wave + 4 * j;
    offa[j] = (16 * (f >> 1) + drow) * kMxRow + (f & 1) * kTile + 8 * dch;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = wave + 4 * j, c = f >> 1, q = c >> 1;
    const int t = q == 1 ? jt[1] : q == 2 ? jt[2] : q == 3 ? jt[3] : jt[0];
    offb[j] = (f & 1) * npad * npad + ((q < nt ? t : jt[0]) * kTile + 16 * (c & 1) + drow) * npad + 8 * dch;
  }
  const u16* __restrict__ c3 = nullptr;
  if (P3) {
    // this batch's core slots in the side cache: [core slot][c][g][mu]
    c3 = m.cache3 + ((size_t)m.core_off[b] * (m.toff / 2) + (size_t)MODE * kBatch * kTile) +
         (16 * wave + drow) * kTile + 8 * dch;
  }
  __syncthreads();
  auto dma = [&](int k) {
    const u16* a = cache + (size_t)(it0 + k) * sstride;
#pragma unroll
    for (int j = 0; j < 2; ++j)
      __builtin_amdgcn_global_load_lds(a + offa[j], (lds_ptr)(sm.F + (wave + 4 * j) * kDxFrag), 16, 0, 0);
    const int ti = sm.tl[k] * kTile;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      __builtin_amdgcn_global_load_lds(Ds + offb[j] + ti, (lds_ptr)(sm.F + (kDxA + wave + 4 * j) * kDxFrag), 16, 0, 0);
    if (P3)
      __builtin_amdgcn_global_load_lds(c3 + (size_t)k * (m.toff / 2),
                                       (lds_ptr)(sm.F + (kDxA + kDxB + wave) * kDxFrag), 16, 0, 0);
  };

  f32x4 acc[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ncb = 2 * (nt - 2 * wc);  // the wave's column blocks that hold group tiles (may be <= 0)
  dma(0);
  for (int k = 0; k < nk; ++k) {
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's fragments of step k are in LDS
    __syncthreads();                     // ... and every other wave's
    u32x4 a[2][2], bq[4][2], a3[2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int h = 0; h < 2; ++h) a[r][h] = sm.F[((2 * wr + r) * 2 + h) * kDxFrag + fidx];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) bq[c][h] = sm.F[(kDxA + (4 * wc + c) * 2 + h) * kDxFrag + fidx];
    if (P3) {
#pragma unroll
      for (int r = 0; r < 2; ++r) a3[r] = sm.F[(kDxA + kDxB + 2 * wr + r) * kDxFrag + fidx];
    }
    __syncthreads();  // every wave holds its fragments: the buffer is free
    if (k + 1 < nk) dma(k + 1);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < ncb) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          f32x4 v = acc[r][c];
          if (P3) {  // the small terms first
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a[r][1]), as_bf(bq[c][1]), v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a3[r]), as_bf(bq[c][0]), v, 0, 0, 0);
          }
          acc[r][c] = mfma3(a[r][0], a[r][1], bq[c][0], bq[c][1], v);
        }
      }
  }

  // Epilogue. X[g = 32 wr + 16 r + 4 lk + i][n = 64 wc + 16 c + lrow] -> LDS as X^T.
  __syncthreads();  // every wave is done with the fragments
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) sm.X[64 * wc + 16 * c + lrow][32 * wr + 16 * r + 4 * lk + i] = acc[r][c][i];
  __syncthreads();
  DensEpi<MODE> epi;
  if (wave < nt) {  // wave q = tile q of the group; lane = point
    const int t = wave == 1 ? jt[1] : wave == 2 ? jt[2] : wave == 3 ? jt[3] : jt[0];
    epi.load(&sm.X[kTile * wave][lane]);
    walk_tile_f32<true>(bv, t, sm.px[lane], sm.py[lane], sm.pz[lane], epi);
  }
#pragma unroll
  for (int c = 0; c < NOUT; ++c) sm.red[wave][c][lane] = epi.o[c];
  __syncthreads();
  if (tid < NOUT * kBatch) {
    const int c = tid / kBatch, g = tid % kBatch;
    double v = 0.0;
    for (int q = 0; q < nt; ++q) v += (double)sm.red[q][c][g];
    part[(size_t)grp * grp_stride + c * kBatch + g] = v;
  }
}
template <int MODE, bool P3>
__global__ __launch_bounds__(256, P3 ? 2 : 4) void k_density_mx(MolDev m) {
  density_mx_body<MODE, P3>(m, xcd_remap(blockIdx.x, gridDim.x));
}

// The three tau partials in one block: per K step the three A = d_r phi
// operands (24 fragments) and the D block (16 fragments, shared by the three
// products) go to LDS together, so D is staged and read once instead of three
// times and a step takes two barriers instead of six. Three accumulator sets;
// the epilogue runs the MODE r contraction for r = 1, 2, 3 in turn through the
// X staging area. Same indexing as k_density_mx.
struct DensMxTauSmem {
  union {
    u32x4 F[(3 * kDxA + kDxB) * kDxFrag];  // one K step's fragments: A_1, A_2, A_3, D
    float X[kGroup * kTile][kBatch + 1];   // epilogue: X_r^T [nu][g]
  };
  float red[kGroup][3][kBatch];
  double px[kBatch], py[kBatch], pz[kBatch];
  int tl[kMxMaxTiles];
};

template <int R>
__device__ __forceinline__ void dens_tau_epilogue(DensMxTauSmem& sm, const f32x4 (&acc)[2][4], const BasisView& bv,
                                                  const int (&jt)[kGroup], int nt, int wave, int lane, int wr,
                                                  int wc, int lrow, int lk) {
  __syncthreads();  // the previous user of the X area (fragments or X_{R-1}) is done
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) sm.X[64 * wc + 16 * c + lrow][32 * wr + 16 * r + 4 * lk + i] = acc[r][c][i];
  __syncthreads();
  DensEpi<R> epi;
  if (wave < nt) {
    const int t = wave == 1 ? jt[1] : wave == 2 ? jt[2] : wave == 3 ? jt[3] : jt[0];
    epi.load(&sm.X[kTile * wave][lane]);
    walk_tile_f32<true>(bv, t, sm.px[lane], sm.py[lane], sm.pz[lane], epi);
  }
  sm.red[wave][R - 1][lane] = epi.o[0];
}

__global__ __launch_bounds__(256, 2) void k_density_mx_tau(MolDev m) {
  using lds_ptr = __attribute__((address_space(3))) void*;
  __shared__ DensMxTauSmem sm;
  const int grp = xcd_remap(blockIdx.x, gridDim.x);
  const u16* __restrict__ cache = m.cache + (size_t)kMxT;  // component 1 (d_x phi); d_y, d_z follow
  const u16* __restrict__ Ds = m.Ds;
  const size_t sstride = (size_t)m.sstride;
  const int npad = m.npad;
  double* __restrict__ part = m.part_tau;  // [group][3][64]
  const BasisView bv = m.bv;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  const int b = m.grp_batch[grp];
  const int nt = m.grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = m.grp_tiles[grp * kGroup + q];
  const int it0 = m.bt_off[b], nk = m.bt_off[b + 1] - it0;
  if (tid < kBatch) {
    sm.px[tid] = m.bx[b * kBatch + tid];
    sm.py[tid] = m.by[b * kBatch + tid];
    sm.pz[tid] = m.bz[b * kBatch + tid];
  }
  for (int j = tid; j < nk; j += 256) sm.tl[j] = m.bt_list[it0 + j];
  const int drow = lane >> 2, dch = (lane & 3) ^ ((drow >> 2) & 3);
  const int fidx = 4 * lrow + (lk ^ ((lrow >> 2) & 3));
  int offa[2], offb[4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int f = wave + 4 * j;
    offa[j] = (16 * (f >> 1) + drow) * kMxRow + (f & 1) * kTile + 8 * dch;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = wave + 4 * j, c = f >> 1, q = c >> 1;
    const int t = q == 1 ? jt[1] : q == 2 ? jt[2] : q == 3 ? jt[3] : jt[0];
    offb[j] = (f & 1) * npad * npad + ((q < nt ? t : jt[0]) * kTile + 16 * (c & 1) + drow) * npad + 8 * dch;
  }
  __syncthreads();
  auto dma = [&](int k) {
    const u16* a = cache + (size_t)(it0 + k) * sstride;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        __builtin_amdgcn_global_load_lds(a + (size_t)r * kMxT + offa[j],
                                         (lds_ptr)(sm.F + (r * kDxA + wave + 4 * j) * kDxFrag), 16, 0, 0);
    const int ti = sm.tl[k] * kTile;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      __builtin_amdgcn_global_load_lds(Ds + offb[j] + ti, (lds_ptr)(sm.F + (3 * kDxA + wave + 4 * j) * kDxFrag), 16,
                                       0, 0);
  };

  f32x4 acc[3][2][4];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[d][r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ncb = 2 * (nt - 2 * wc);
  dma(0);
  for (int k = 0; k < nk; ++k) {
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's fragments of step k are in LDS
    __syncthreads();                     // ... and every other wave's
    u32x4 a[3][2][2], bq[4][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int h = 0; h < 2; ++h) a[d][r][h] = sm.F[(d * kDxA + (2 * wr + r) * 2 + h) * kDxFrag + fidx];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) bq[c][h] = sm.F[(3 * kDxA + (4 * wc + c) * 2 + h) * kDxFrag + fidx];
    __syncthreads();  // every wave holds its fragments: the buffer is free
    if (k + 1 < nk) dma(k + 1);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < ncb) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int r = 0; r < 2; ++r)
            acc[d][r][c] = mfma3(a[d][r][0], a[d][r][1], bq[c][0], bq[c][1], acc[d][r][c]);
      }
  }
  dens_tau_epilogue<1>(sm, acc[0], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  dens_tau_epilogue<2>(sm, acc[1], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  dens_tau_epilogue<3>(sm, acc[2], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  __syncthreads();
  if (tid < 3 * kBatch) {
    const int c = tid / kBatch, g = tid % kBatch;
    double v = 0.0;
    for (int q = 0; q < nt; ++q) v += (double)sm.red[q][c][g];
    part[(size_t)grp * 3 * kBatch + c * kBatch + g] = v;
  }
}

// Meta-GGA density in one launch: the four products Phi D and d_r Phi D share
// one D block per K step (48 fragments: A_0..A_3, then D), so D is staged
// once per build instead of once for rho and once for the tau partials.
// Epilogues: rho and grad rho / 2 into part (as k_density_mx<0>), then the
// three tau partials into part_tau (as k_density_mx_tau). Same indexing as
// k_density_mx.
struct DensMxMggaSmem {
  union {
    u32x4 F[(4 * kDxA + kDxB) * kDxFrag];  // one K step's fragments: A_0..A_3, D
    float X[kGroup * kTile][kBatch + 1];
  };
  float red[kGroup][4][kBatch];
  double px[kBatch], py[kBatch], pz[kBatch];
  int tl[kMxMaxTiles];
};

template <int R>
__device__ __forceinline__ void dens_mgga_epilogue(DensMxMggaSmem& sm, const f32x4 (&acc)[2][4], const BasisView& bv,
                                                   const int (&jt)[kGroup], int nt, int wave, int lane, int wr, int wc,
                                                   int lrow, int lk) {
  __syncthreads();  // the previous user of the X area is done
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) sm.X[64 * wc + 16 * c + lrow][32 * wr + 16 * r + 4 * lk + i] = acc[r][c][i];
  __syncthreads();
  DensEpi<R> epi;
  if (wave < nt) {
    const int t = wave == 1 ? jt[1] : wave == 2 ? jt[2] : wave == 3 ? jt[3] : jt[0];
    epi.load(&sm.X[kTile * wave][lane]);
    walk_tile_f32<true>(bv, t, sm.px[lane], sm.py[lane], sm.pz[lane], epi);
  }
  if (R == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) sm.red[wave][c][lane] = epi.o[c];
  } else {
    sm.red[wave][R - 1][lane] = epi.o[0];
  }
}

__global__ __launch_bounds__(256, 2) void k_density_mx_mgga(MolDev m) {
  using lds_ptr = __attribute__((address_space(3))) void*;
  __shared__ DensMxMggaSmem sm;
  const int grp = xcd_remap(blockIdx.x, gridDim.x);
  const u16* __restrict__ cache = m.cache;  // components 0 (phi), 1..3 (d_r phi)
  const u16* __restrict__ Ds = m.Ds;
  const size_t sstride = (size_t)m.sstride;
  const int npad = m.npad;
  const BasisView bv = m.bv;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  const int b = m.grp_batch[grp];
  const int nt = m.grp_ntiles[grp];
  int jt[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) jt[q] = m.grp_tiles[grp * kGroup + q];
  const int it0 = m.bt_off[b], nk = m.bt_off[b + 1] - it0;
  if (tid < kBatch) {
    sm.px[tid] = m.bx[b * kBatch + tid];
    sm.py[tid] = m.by[b * kBatch + tid];
    sm.pz[tid] = m.bz[b * kBatch + tid];
  }
  for (int j = tid; j < nk; j += 256) sm.tl[j] = m.bt_list[it0 + j];
  const int drow = lane >> 2, dch = (lane & 3) ^ ((drow >> 2) & 3);
  const int fidx = 4 * lrow + (lk ^ ((lrow >> 2) & 3));
  int offa[2], offb[4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int f = wave + 4 * j;
    offa[j] = (16 * (f >> 1) + drow) * kMxRow + (f & 1) * kTile + 8 * dch;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = wave + 4 * j, c = f >> 1, q = c >> 1;
    const int t = q == 1 ? jt[1] : q == 2 ? jt[2] : q == 3 ? jt[3] : jt[0];
    offb[j] = (f & 1) * npad * npad + ((q < nt ? t : jt[0]) * kTile + 16 * (c & 1) + drow) * npad + 8 * dch;
  }
  __syncthreads();
  auto dma = [&](int k) {
    const u16* a = cache + (size_t)(it0 + k) * sstride;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        __builtin_amdgcn_global_load_lds(a + (size_t)r * kMxT + offa[j],
                                         (lds_ptr)(sm.F + (r * kDxA + wave + 4 * j) * kDxFrag), 16, 0, 0);
    const int ti = sm.tl[k] * kTile;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      __builtin_amdgcn_global_load_lds(Ds + offb[j] + ti, (lds_ptr)(sm.F + (4 * kDxA + wave + 4 * j) * kDxFrag), 16,
                                       0, 0);
  };

  f32x4 acc[4][2][4];
#pragma unroll
  for (int d = 0; d < 4; ++d)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[d][r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ncb = 2 * (nt - 2 * wc);
  dma(0);
  for (int k = 0; k < nk; ++k) {
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's fragments of step k are in LDS
    __syncthreads();                     // ... and every other wave's
    u32x4 a[4][2][2], bq[4][2];
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int h = 0; h < 2; ++h) a[d][r][h] = sm.F[(d * kDxA + (2 * wr + r) * 2 + h) * kDxFrag + fidx];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) bq[c][h] = sm.F[(4 * kDxA + (4 * wc + c) * 2 + h) * kDxFrag + fidx];
    __syncthreads();  // every wave holds its fragments: the buffer is free
    if (k + 1 < nk) dma(k + 1);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < ncb) {
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
          for (int r = 0; r < 2; ++r)
            acc[d][r][c] = mfma3(a[d][r][0], a[d][r][1], bq[c][0], bq[c][1], acc[d][r][c]);
      }
  }
  dens_mgga_epilogue<0>(sm, acc[0], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  __syncthreads();
  if (tid < 4 * kBatch) {
    const int c = tid / kBatch, g = tid % kBatch;
    double v = 0.0;
    for (int q = 0; q < nt; ++q) v += (double)sm.red[q][c][g];
    m.part[(size_t)grp * 4 * kBatch + c * kBatch + g] = v;
  }
  dens_mgga_epilogue<1>(sm, acc[1], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  dens_mgga_epilogue<2>(sm, acc[2], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  dens_mgga_epilogue<3>(sm, acc[3], bv, jt, nt, wave, lane, wr, wc, lrow, lk);
  __syncthreads();
  if (tid < 3 * kBatch) {
    const int c = tid / kBatch, g = tid % kBatch;
    double v = 0.0;
    for (int q = 0; q < nt; ++q) v += (double)sm.red[q][c][g];
    m.part_tau[(size_t)grp * 3 * kBatch + c * kBatch + g] = v;
  }
}

void launch_density_mx_core(hipStream_t s, int mode, int ncore_groups, const MolDev& one) {
  if (ncore_groups == 0) return;
  const dim3 g(ncore_groups), b(256);
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_density_mx<0, true>), g, b, 0, s, one); break;
    case 1: hipLaunchKernelGGL((k_density_mx<1, true>), g, b, 0, s, one); break;
    case 2: hipLaunchKernelGGL((k_density_mx<2, true>), g, b, 0, s, one); break;
    case 3: hipLaunchKernelGGL((k_density_mx<3, true>), g, b, 0, s, one); break;
    default: throw Error(CHX_ERR_ARG, "core density mode must be 0..3");
  }
  CHX_HIP(hipGetLastError());
}

void launch_density_mx(hipStream_t s, int mode, int nitems, const MolDev& one) {
  if (nitems == 0) return;
  if (mode == 0)
    hipLaunchKernelGGL((k_density_mx<0, false>), dim3(nitems), dim3(256), 0, s, one);
  else if (mode == 4)  // the three tau partials in one launch
    hipLaunchKernelGGL(k_density_mx_tau, dim3(nitems), dim3(256), 0, s, one);
  else if (mode == 5)  // rho, grad rho and the tau partials in one launch
    hipLaunchKernelGGL(k_density_mx_mgga, dim3(nitems), dim3(256), 0, s, one);
  else
    throw Error(CHX_ERR_ARG, "density mode must be 0 (rho, grad rho), 4 (tau partials) or 5 (both)");
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// B^T in split form, one wave per slot (lane = point), from phi and grad phi
// recomputed in FP32:
//   Bt[slot][0][hl][nu][g] = split(a phi + c . grad phi),
//   meta-GGA (NC = 4) adds Bt[slot][r][hl][nu][g] = split(t d_r phi), t = w v_tau / 4.
// Each lane stores its point's values straight into the slot's [c][hl][nu][g]
// block (for each AO, the wave's 64 lanes write one contiguous 128-byte row);
// columns past the tile's last AO are written as zeros. No LDS stage and no
// barrier, so occupancy is set by registers alone.
template <int NC>
struct MakeB {
  u16* sv;  // this lane's column of the [c][hl][nu][g] block
  float ca, cx, cy, cz, ct, R, Rx, Ry, Rz, s1;
  __device__ void begin(float a, float ax, float ay, float az) {
    R = a; Rx = ax; Ry = ay; Rz = az;
    s1 = ca * a + cx * ax + cy * ay + cz * az;  // (a + c . grad) R
  }
  __device__ void comp(int col, float q, float qx, float qy, float qz) {
    u16 hi, lo;
    split_bf(q * s1 + R * (cx * qx + cy * qy + cz * qz), hi, lo);  // a phi + c . grad phi
    sv[col * kBatch] = hi;
    sv[kTile * kBatch + col * kBatch] = lo;
    if (NC == 4) {
      split_bf(ct * (Rx * q + R * qx), hi, lo);
      sv[kMxT + col * kBatch] = hi;
      sv[kMxT + kTile * kBatch + col * kBatch] = lo;
      split_bf(ct * (Ry * q + R * qy), hi, lo);
      sv[2 * kMxT + col * kBatch] = hi;
      sv[2 * kMxT + kTile * kBatch + col * kBatch] = lo;
      split_bf(ct * (Rz * q + R * qz), hi, lo);
      sv[3 * kMxT + col * kBatch] = hi;
      sv[3 * kMxT + kTile * kBatch + col * kBatch] = lo;
    }
  }
  __device__ void end() {}
};

// The wave of slot index nslots clears the all-zero slot that the Vxc items
// read for tiles that are not alive in a batch: it sits at nslots * NC
// components, so its place moves with the potential's component count and is
// rewritten on every build.
template <int NC>
__device__ __forceinline__ void make_b_slot(const MolDev& m, int slot, int g) {
  if (slot > m.nslots) return;
  u16* __restrict__ col0 = m.bt + (size_t)slot * NC * kMxT + g;  // [c][hl][nu][g], this lane's point
  if (slot == m.nslots) {
    for (int col = 0; col < kTile; ++col)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        col0[c * kMxT + col * kBatch] = 0;
        col0[c * kMxT + kTile * kBatch + col * kBatch] = 0;
      }
    return;
  }
  const BasisView bv = m.bv;  // descriptor fields in registers, read once
  const int b = m.slot_batch[slot], t = m.slot_tile[slot];
  const size_t p = (size_t)b * kBatch + g;
  const double* __restrict__ coef = m.coef;
  MakeB<NC> op;
  op.sv = col0;
  op.ca = (float)coef[kCoef * p];
  op.cx = (float)coef[kCoef * p + 1];
  op.cy = (float)coef[kCoef * p + 2];
  op.cz = (float)coef[kCoef * p + 3];
  op.ct = (float)coef[kCoef * p + 4];
  walk_tile_f32<true>(bv, t, m.bx[p], m.by[p], m.bz[p], op);
  // zeros past the tile's last AO (segments hold (shell, first component, count, local offset))
  int nao = 0;
  for (int k = bv.tile_seg_off[t]; k < bv.tile_seg_off[t + 1]; ++k) {
    const int4 sg = bv.seg[k];
    nao = max(nao, sg.w + sg.z);
  }
  for (int col = nao; col < kTile; ++col)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      col0[c * kMxT + col * kBatch] = 0;
      col0[c * kMxT + kTile * kBatch + col * kBatch] = 0;
    }
}
template <int NC>
__global__ __launch_bounds__(NC == 4 ? 64 : 256) void k_make_b_mx(MolDev m) {
  constexpr int W = NC == 4 ? 1 : 4;  // waves per block
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  make_b_slot<NC>(m, blockIdx.x * W + wave, threadIdx.x & 63);
}

void launch_make_b_mx(hipStream_t s, int ncomp, const MolDev& one) {
  const int waves = one.nslots + 1;  // every slot and the all-zero slot
  if (ncomp == 4)
    hipLaunchKernelGGL((k_make_b_mx<4>), dim3(waves), dim3(64), 0, s, one);
  else
    hipLaunchKernelGGL((k_make_b_mx<1>), dim3((waves + 3) / 4), dim3(256), 0, s, one);
  CHX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Vxc: M_IJ += sum_g Phi_I[g]^T B_J[g] over the work item's batches, for
// 128-AO super-tiles I, J. A K step is 32 points of one component: Phi^T_I
// (128 mu x 32 g) and B^T_J (128 nu x 32 g), hi and lo, 32 fragments of 1 KB
// (16 rows of 64 bytes) that the waves' DMA loads bring straight into LDS,
// coalesced and swizzled as in the density kernel. One buffer: wait for the
// step's fragments, read the wave's 16 into registers, barrier, start the DMA
// of the next step, run the wave's 48 MFMAs. 4 waves in a 2 x 2 grid, each
// owning a 64 x 64 block of M. One FP64 atomic add per element per work item.
// Meta-GGA (NC = 4): the tau part sum_r (d_r Phi)^T (t d_r Phi) is symmetric
// and V = M + M^T, so an item with I > J skips it, an item with I < J takes it
// twice (its tau steps run first, then the accumulators are doubled before the
// phi steps), and I = J takes it once.
constexpr int kVxA = 16;  // Phi^T fragments per step: [row block (8)][hl]
struct VxcMxSmem {
  u32x4 F[2 * kVxA * kDxFrag];  // one K step's fragments: Phi^T_I, then B^T_J
  int sl[kMxMaxChunk][8];       // per batch of the item: slots of I's and J's tiles
};

template <int NC>
__device__ __forceinline__ void vxc_mx_body(const MolDev& m, int item) {
  using lds_ptr = __attribute__((address_space(3))) void*;
  __shared__ VxcMxSmem sm;
  const int4 wk = m.cwork[item];
  const int I = wk.x, J = wk.y;
  const u16* __restrict__ cache = m.cache + m.toff;
  const u16* __restrict__ bt = m.bt;
  const size_t sstride = (size_t)m.sstride;
  const int npad = m.npad;
  double* __restrict__ Mint = m.Mint;
  {  // the item's slot table (absent tiles point at the all-zero slot)
    const int* __restrict__ cs = m.cslot + (size_t)wk.z * 8;
    for (int e = threadIdx.x; e < (wk.w - wk.z) * 8; e += 256) sm.sl[e >> 3][e & 7] = cs[e];
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  // The wave's fragments are f = wave + 4 j, j = 0..3 of Phi^T_I and of B^T_J:
  // row block f >> 1 (tile j = f >> 2, rows 16 ((f >> 1) & 1)..+15 of it), hl
  // f & 1. DMA lane i fetches row i >> 2, chunk (i & 3) ^ sw(row); the MFMA
  // lane that needs (row, chunk) reads it back at fidx.
  const int drow = lane >> 2, dch = (lane & 3) ^ ((drow >> 2) & 3);
  const int fidx = 4 * lrow + (lk ^ ((lrow >> 2) & 3));
  int off[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = wave + 4 * j;
    off[j] = (f & 1) * (kTile * kBatch) + (16 * ((f >> 1) & 1) + drow) * kBatch + 8 * dch;
  }
  __syncthreads();
  const int nbat = wk.w - wk.z;                       // items hold >= 1 batch
  const int ntau = (NC == 4 && I <= J) ? 6 * nbat : 0;  // tau steps, first
  const int nstage = ntau + 2 * nbat;
  // K step k -> (batch of the item, half of its points, component): tau steps
  // (it, h, c = 1..3) for k < ntau, then the phi steps (it, h, c = 0)
  auto dma = [&](int k) {
    int it, h, c;
    if (k < ntau) {
      it = k / 6; h = (k / 3) % 2; c = 1 + k % 3;
    } else {
      it = (k - ntau) >> 1; h = (k - ntau) & 1; c = 0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const u16* pa = cache + (size_t)sm.sl[it][j] * sstride + c * kMxT + h * 32;
      const u16* pb = bt + ((size_t)sm.sl[it][4 + j] * NC + c) * kMxT + h * 32;
      __builtin_amdgcn_global_load_lds(pa + off[j], (lds_ptr)(sm.F + (wave + 4 * j) * kDxFrag), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(pb + off[j], (lds_ptr)(sm.F + (kVxA + wave + 4 * j) * kDxFrag), 16, 0, 0);
    }
  };
  f32x4 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  dma(0);
  for (int k = 0; k < nstage; ++k) {
    if (NC == 4 && k == ntau && I < J) {  // 2 M_tau for the item's (I, J) and (J, I) blocks
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] *= 2.f;
    }
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's fragments of step k are in LDS
    __syncthreads();                     // ... and every other wave's
    u32x4 a[4][2], bq[4][2];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        a[r][h] = sm.F[((4 * wr + r) * 2 + h) * kDxFrag + fidx];
        bq[r][h] = sm.F[(kVxA + (4 * wc + r) * 2 + h) * kDxFrag + fidx];
      }
    __syncthreads();  // every wave holds its fragments: the buffer is free
    if (k + 1 < nstage) dma(k + 1);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r][c] = mfma3(a[r][0], a[r][1], bq[c][0], bq[c][1], acc[r][c]);
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
template <int NC>
__global__ __launch_bounds__(256, 3) void k_vxc_mx(MolDev m) {
  vxc_mx_body<NC>(m, xcd_remap(blockIdx.x, gridDim.x));
}

// Vxc for large bases: M_IJ += Phi_I^T B_J over 256 x 256 blocks whose rows
// and columns are groups of 8 tiles of a batch chunk's alive-tile union, so
// every operand byte staged in LDS feeds twice the MFMAs of the 128 x 128
// kernel (64 instead of 32 FLOP per byte). An item is a pair of groups I <= J
// of one chunk and covers both blocks: first M_IJ (the tau part, symmetric,
// taken twice for I < J and once for I = J, then the phi part), then for
// I < J the mirror block M_JI (phi part only), so every item does about the
// same work per batch and the items that run together on an XCD stream the
// chunk's slots in step. A K step is 32 points of one component: Phi^T of the
// row group and B^T of the column group, 256 rows each, hi and lo, 64
// fragments of 1 KB brought in by the waves' DMA (same coalesced lane order
// and swizzle). 8 waves in a 2 x 4 grid, each owning a 128 x 64 block (up to
// 96 MFMAs per step). Per batch a 16-bit mask (I's 8 tiles, then J's) marks
// the tiles alive in it: a dead tile's fragments are not loaded and the
// products that involve it are skipped, so the MFMAs follow each batch's
// screening.
constexpr int kMxSupW = 8;   // tiles per group (256 AOs)
constexpr int kVxAW = 32;    // Phi^T fragments per step: [row block (16)][hl]
constexpr int kVxWStep = 2 * kVxAW * kDxFrag;  // u32x4 per K step: Phi^T of the row group, then B^T of the column group
struct VxcMxWideSmem {
  u32x4 F[2 * kVxWStep];  // two K steps' fragments: step k in buffer k & 1
  int sl[kMxMaxChunk][16];       // per batch: slots of I's 8 tiles, then J's
  unsigned mk[kMxMaxChunk];      // per batch: alive bits of those 16 tiles
  int tile[16];                  // I's tiles, then J's (-1: padding)
};

// Two LDS buffers and one barrier per K step: after it, step k's fragments
// are in LDS and every wave has read step k - 1's buffer, so the DMA of step
// k + 1 goes there before step k's fragment reads and products (measured at
// (H2O)128 / r2SCAN: 114.5 ms per launch against 125.5 ms with one buffer and
// a second barrier after the fragment reads).
template <int NC>
__device__ __forceinline__ void vxc_mx_wide_body(const MolDev& m, int item) {
  using lds_ptr = __attribute__((address_space(3))) void*;
  __shared__ VxcMxWideSmem sm;
  const int4 wk = m.cwork[item];
  const int I = wk.x, J = wk.y;    // I <= J
  const int4 ch = m.vchunk[wk.z];  // (table offset, batches, table width, tile-list offset)
  const int nbat = ch.y;
  const int zero = m.nslots;       // the all-zero slot: tile not alive in that batch
  const u16* __restrict__ cache = m.cache + m.toff;
  const u16* __restrict__ bt = m.bt;
  const size_t sstride = (size_t)m.sstride;
  const int npad = m.npad;
  double* __restrict__ Mint = m.Mint;
  {
    const int* __restrict__ tab = m.vtab + ch.x;
    for (int e = threadIdx.x; e < nbat * 16; e += 512) {
      const int it = e >> 4, t = e & 15;
      sm.sl[it][t] = tab[(size_t)it * ch.z + (t < 8 ? I * kMxSupW + t : J * kMxSupW + t - 8)];
    }
    if (threadIdx.x < 16) {
      const int t = threadIdx.x;
      sm.tile[t] = m.vtile[ch.w + (t < 8 ? I * kMxSupW + t : J * kMxSupW + t - 8)];
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < nbat; it += 512) {
    unsigned b = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) b |= (sm.sl[it][t] != zero ? 1u : 0u) << t;
    sm.mk[it] = b;
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lrow = lane & 15, lk = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  // The wave's fragments are f = wave + 8 j, j = 0..3, of the row group's Phi^T
  // and of the column group's B^T: hl f & 1 = wave & 1, rows
  // 16 ((wave >> 1) & 1)..+15 of tile (wave >> 2) + 2 j.
  const int drow = lane >> 2, dch = (lane & 3) ^ ((drow >> 2) & 3);
  const int fidx = 4 * lrow + (lk ^ ((lrow >> 2) & 3));
  const int off = (wave & 1) * (kTile * kBatch) + (16 * ((wave >> 1) & 1) + drow) * kBatch + 8 * dch;
  const int t0 = wave >> 2;
  __syncthreads();
  const int ntau = NC == 4 ? 6 * nbat : 0;  // tau steps of M_IJ, first
  const int kmirror = ntau + 2 * nbat;      // first step of M_JI (I < J)
  const int nstage = kmirror + (I < J ? 2 * nbat : 0);
  // K step k -> (batch, half of its points, component, mirror block?)
  auto decode = [&](int k, int& it, int& h, int& c, bool& mir) {
    mir = k >= kmirror;
    if (k < ntau) {
      it = k / 6; h = (k / 3) % 2; c = 1 + k % 3;
    } else {
      const int q = k - (mir ? kmirror : ntau);
      it = q >> 1; h = q & 1; c = 0;
    }
  };
  // issue step k's loads into buffer buf
  auto dma = [&](int k, int buf) {
    int it, h, c;
    bool mir;
    decode(k, it, h, c, mir);
    const int ra = mir ? 8 : 0, cb = mir ? 0 : 8;  // table columns of the row and column groups
    const unsigned mk = __builtin_amdgcn_readfirstlane(sm.mk[it]);
    u32x4* F = sm.F + buf * kVxWStep;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = t0 + 2 * j;
      if ((mk >> (ra + t)) & 1u) {
        const u16* pa = cache + (size_t)sm.sl[it][ra + t] * sstride + c * kMxT + h * 32;
        __builtin_amdgcn_global_load_lds(pa + off, (lds_ptr)(F + (wave + 8 * j) * kDxFrag), 16, 0, 0);
      }
      if ((mk >> (cb + t)) & 1u) {
        const u16* pb = bt + ((size_t)sm.sl[it][cb + t] * NC + c) * kMxT + h * 32;
        __builtin_amdgcn_global_load_lds(pb + off, (lds_ptr)(F + (kVxAW + wave + 8 * j) * kDxFrag), 16, 0, 0);
      }
    }
  };
  f32x4 acc[8][4];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  // one FP64 atomic add per element of the wave's 128 x 64 block; rows from
  // table columns ro.., columns from co..
  auto flush = [&](int ro, int co) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int ti = sm.tile[ro + 4 * wr + (r >> 1)];
      if (ti < 0) continue;
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const int tj = sm.tile[co + 2 * wc + (cb >> 1)];
        if (tj < 0) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int mu = ti * kTile + 16 * (r & 1) + 4 * lk + i;
          const int nu = tj * kTile + 16 * (cb & 1) + lrow;
          unsafeAtomicAdd(&Mint[(size_t)mu * npad + nu], (double)acc[r][cb][i]);
        }
      }
    }
  };
  dma(0, 0);
  for (int k = 0; k < nstage; ++k) {
    if (NC == 4 && k == ntau && I < J) {  // 2 M_tau: the item's (I, J) block and, by symmetry, (J, I)
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] *= 2.f;
    }
    if (k == kmirror) {  // M_IJ done: write it and start M_JI
      flush(0, 8);
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    int it, h, c;
    bool mir;
    decode(k, it, h, c, mir);
    // this wave's row tiles 4 wr..4 wr+3 (bits of rmk) and column tiles 2 wc, 2 wc+1 (cmk) alive in the batch
    const unsigned mk = __builtin_amdgcn_readfirstlane(sm.mk[it]);
    const unsigned rmk = (mk >> ((mir ? 8 : 0) + 4 * wr)) & 15u, cmk = (mk >> ((mir ? 0 : 8) + 2 * wc)) & 3u;
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's fragments of step k are in LDS
    __syncthreads();                     // ... and every other wave's, and step k - 1's buffer is read
    if (k + 1 < nstage) dma(k + 1, (k + 1) & 1);
    const u32x4* F = sm.F + (k & 1) * kVxWStep;
    u32x4 a[8][2], bq[4][2];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) a[r][hh] = F[((8 * wr + r) * 2 + hh) * kDxFrag + fidx];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) bq[r][hh] = F[(kVxAW + (4 * wc + r) * 2 + hh) * kDxFrag + fidx];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
      if ((cmk >> (cc >> 1)) & 1u) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if ((rmk >> (r >> 1)) & 1u) acc[r][cc] = mfma3(a[r][0], a[r][1], bq[cc][0], bq[cc][1], acc[r][cc]);
      }
  }
  if (I < J) flush(8, 0);
  else flush(0, 8);
}
template <int NC>
__global__ __launch_bounds__(512, 1) void k_vxc_mx_wide(MolDev m) {
  vxc_mx_wide_body<NC>(m, xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// Persistent stage kernels over many molecules (packs). A table holds one
// descriptor per molecule whose pointer fields are byte offsets from a base
// kernel argument; rebase() adds the base, so every table address the kernels
// form derives from a kernel argument and stays a global-memory access (a
// descriptor read from device memory with its own pointers would leave them
// flat). starts[k] is the first item of molecule k in the stage's flat item
// list (starts[nmol] = total). A fixed grid of blocks walks the list: virtual
// block v = blockIdx.x + j gridDim.x takes item xcd_remap(v, total) (gridDim.x
// a multiple of 8, so each XCD keeps a contiguous range), finds its molecule by
// bisection, and runs the single-molecule body on it.
template <class T>
__device__ __forceinline__ T* rb(const char* base, T* off) {
  return reinterpret_cast<T*>(const_cast<char*>(base) + reinterpret_cast<long>(off));
}
__device__ __forceinline__ MolDev rebase(const char* base, const MolDev& o) {
  MolDev m = o;
  m.grp_batch = rb(base, o.grp_batch); m.grp_tiles = rb(base, o.grp_tiles); m.grp_ntiles = rb(base, o.grp_ntiles);
  m.bt_off = rb(base, o.bt_off); m.bt_list = rb(base, o.bt_list);
  m.bx = rb(base, o.bx); m.by = rb(base, o.by); m.bz = rb(base, o.bz);
  m.bv.sh_xyz = rb(base, o.bv.sh_xyz); m.bv.sh_type = rb(base, o.bv.sh_type);
  m.bv.sh_nprim = rb(base, o.bv.sh_nprim); m.bv.sh_poff = rb(base, o.bv.sh_poff);
  m.bv.p_exp = rb(base, o.bv.p_exp); m.bv.p_coef = rb(base, o.bv.p_coef);
  m.bv.tile_seg_off = rb(base, o.bv.tile_seg_off); m.bv.seg = rb(base, o.bv.seg);
  m.bv.segrec = rb(base, o.bv.segrec); m.bv.sh_atom = rb(base, o.bv.sh_atom);
  m.cache = rb(base, o.cache); m.Ds = rb(base, o.Ds);
  m.part = rb(base, o.part); m.part_tau = rb(base, o.part_tau);
  m.slot_batch = rb(base, o.slot_batch); m.slot_tile = rb(base, o.slot_tile);
  m.coef = rb(base, o.coef); m.bt = rb(base, o.bt);
  m.cwork = rb(base, o.cwork); m.pair_batches = rb(base, o.pair_batches); m.slot_of = rb(base, o.slot_of);
  m.cslot = rb(base, o.cslot); m.Mint = rb(base, o.Mint);
  m.vchunk = rb(base, o.vchunk); m.vtab = rb(base, o.vtab); m.vtile = rb(base, o.vtile);
  m.cache3 = rb(base, o.cache3);
  m.core_off = rb(base, o.core_off); m.core_grps = rb(base, o.core_grps); m.main_grps = rb(base, o.main_grps);
  return m;
}
__device__ __forceinline__ int item_owner(const int* __restrict__ starts, int nmol, int item) {
  int lo = 0, hi = nmol;  // the last molecule whose first item is <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (starts[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}
// A block's items come in increasing order (xcd_remap keeps each block's XCD),
// so after the first bisection its molecule cursor only moves forward.
__device__ __forceinline__ int item_owner_from(const int* __restrict__ starts, int nmol, int item, int k) {
  if (k < 0) return item_owner(starts, nmol, item);
  while (k + 1 < nmol && starts[k + 1] <= item) ++k;
  return k;
}

template <int MODE, bool P3>
__global__ __launch_bounds__(256, P3 ? 2 : 4) void k_density_mx_multi(const char* base, const MolDev* __restrict__ tab,
                                                                      const int* __restrict__ starts, int nmol,
                                                                      int total) {
  int k = -1;
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int item = xcd_remap(v, total);
    k = item_owner_from(starts, nmol, item, k);
    density_mx_body<MODE, P3>(rebase(base, tab[k]), item - starts[k]);
    __syncthreads();  // the next item reuses the block's LDS
  }
}

template <int NC>
__global__ __launch_bounds__(256) void k_make_b_mx_multi(const char* base, const MolDev* __restrict__ tab,
                                                         const int* __restrict__ starts, int nmol, int total) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int k = -1;
  for (int v = blockIdx.x * 4 + wave; v < total; v += gridDim.x * 4) {  // one wave per slot, no LDS
    k = item_owner_from(starts, nmol, v, k);
    make_b_slot<NC>(rebase(base, tab[k]), v - starts[k], threadIdx.x & 63);
  }
}

template <int NC>
__global__ __launch_bounds__(256, 3) void k_vxc_mx_multi(const char* base, const MolDev* __restrict__ tab,
                                                         const int* __restrict__ starts, int nmol, int total) {
  int k = -1;
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int item = xcd_remap(v, total);
    k = item_owner_from(starts, nmol, item, k);
    vxc_mx_body<NC>(rebase(base, tab[k]), item - starts[k]);
    __syncthreads();
  }
}

template <int NC>
__global__ __launch_bounds__(512, 1) void k_vxc_mx_wide_multi(const char* base, const MolDev* __restrict__ tab,
                                                              const int* __restrict__ starts, int nmol, int total) {
  int k = -1;
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int item = xcd_remap(v, total);
    k = item_owner_from(starts, nmol, item, k);
    vxc_mx_wide_body<NC>(rebase(base, tab[k]), item - starts[k]);
    __syncthreads();
  }
}

// Blocks of a persistent launch: the kernel's resident blocks per CU times the
// CUs, a multiple of 8, at most the items (rounded up to 8).
template <class K>
static int persistent_blocks(K kernel, int threads, int total) {
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    CHX_HIP(hipGetDevice(&dev));
    CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  int per_cu = 0;
  CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  const long full = (long)std::max(1, per_cu) * ncu;
  const long want = std::min<long>(full, ((long)total + 7) / 8 * 8);
  return (int)std::max<long>(8, want / 8 * 8);
}

void launch_stage_multi(hipStream_t s, int stage, int ncomp, const char* base, const MolDev* tab, const int* starts,
                        int nmol, int total) {
  if (total <= 0) return;
  switch (stage) {
    case kStageDensity: {
      auto k = k_density_mx_multi<0, false>;
      hipLaunchKernelGGL(k, dim3(persistent_blocks(k, 256, total)), dim3(256), 0, s, base, tab, starts, nmol, total);
      break;
    }
    case kStageCore: {
      auto k = k_density_mx_multi<0, true>;
      hipLaunchKernelGGL(k, dim3(persistent_blocks(k, 256, total)), dim3(256), 0, s, base, tab, starts, nmol, total);
      break;
    }
    case kStageMakeB: {
      if (ncomp != 1) throw Error(CHX_ERR_UNSUPPORTED, "persistent packs are built for LDA and GGA");
      auto k = k_make_b_mx_multi<1>;
      const int blocks = persistent_blocks(k, 256, (total + 3) / 4);
      hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, base, tab, starts, nmol, total);
      break;
    }
    case kStageVxc: {
      if (ncomp != 1) throw Error(CHX_ERR_UNSUPPORTED, "persistent packs are built for LDA and GGA");
      auto k = k_vxc_mx_multi<1>;
      hipLaunchKernelGGL(k, dim3(persistent_blocks(k, 256, total)), dim3(256), 0, s, base, tab, starts, nmol, total);
      break;
    }
    case kStageVxcWide: {
      if (ncomp != 1) throw Error(CHX_ERR_UNSUPPORTED, "persistent packs are built for LDA and GGA");
      auto k = k_vxc_mx_wide_multi<1>;
      hipLaunchKernelGGL(k, dim3(persistent_blocks(k, 512, total)), dim3(512), 0, s, base, tab, starts, nmol, total);
      break;
    }
    default:
      throw Error(CHX_ERR_ARG, "unknown pack stage");
  }
  CHX_HIP(hipGetLastError());
}

void launch_vxc_mx_wide(hipStream_t s, int ncomp, int nitems, const MolDev& one) {
  if (nitems == 0) return;
  if (ncomp == 4)
    hipLaunchKernelGGL((k_vxc_mx_wide<4>), dim3(nitems), dim3(512), 0, s, one);
  else
    hipLaunchKernelGGL((k_vxc_mx_wide<1>), dim3(nitems), dim3(512), 0, s, one);
  CHX_HIP(hipGetLastError());
}

void launch_vxc_mx(hipStream_t s, int ncomp, int nitems, const MolDev& one) {
  if (nitems == 0) return;
  if (ncomp == 4)
    hipLaunchKernelGGL((k_vxc_mx<4>), dim3(nitems), dim3(256), 0, s, one);
  else
    hipLaunchKernelGGL((k_vxc_mx<1>), dim3(nitems), dim3(256), 0, s, one);
  CHX_HIP(hipGetLastError());
}

}  // namespace chx
