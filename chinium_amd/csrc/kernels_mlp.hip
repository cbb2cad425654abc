// Learned stand-in for the meta-GGA functional kernel: a small MLP on the
// spin-restricted descriptors (ln rho, s, alpha) that predicts the enhancement
// factor F = e_xc / e_x^unif(rho); e_xc's rho / sigma / tau derivatives come
// from the network's own input gradient. Trained on the GPU by the kernels
// below (forward + backward with the weight gradient, Adam); evaluated with
// FP32 weights (split into two BF16 terms for the MFMA products) and FP32
// arithmetic. With an analytic base (MlpBox::base, e.g. Slater + VWN5) the
// network predicts F - F_base and the base is added back. Points outside the
// training box, and the low-density tail, take the analytic functional
// (FP64). No reference
// counterpart: the reference evaluates functionals through libxc
// (src/ExchangeCorrelation.cpp:86-163), whose result this replaces.
//
// Network: 3 -> 64 -> 64 -> 64 -> 1, tanh hidden units, linear output. One
// wave handles 64 samples (lane = sample). Training runs the network per lane
// in FP32 with the weights uniform across the wave, read through the constant
// address space (scalar loads) from a run copy that holds W2 transposed, so
// that every weight row a loop walks is contiguous. Inference (Evaluate and
// point evaluation) runs the three hidden products and the two backward ones
// as transposed split-BF16 MFMA GEMMs over the wave's 64 points.
#include "internal.hpp"
#include "device_common.hpp"
#include "xc_functionals.hpp"

namespace chx {

namespace {

constexpr int kLanes = 64;
constexpr double kCxUnif = -0.7385587663820224;  // e_x^unif = kCxUnif rho^(4/3) = -(3/4)(3/pi)^(1/3) rho^(4/3)
constexpr double kSFac = 0.16162045967399548;    // s = kSFac |grad rho| / rho^(4/3); 1 / (2 (3 pi^2)^(1/3))
constexpr double kTauUnif = 2.8712340001881915;  // tau_unif = kTauUnif rho^(5/3); (3/10) (3 pi^2)^(2/3)

__device__ __forceinline__ float ldp(const float* P, int i) {
  using cf = const __attribute__((address_space(4))) float;
  return ((cf*)P)[i];
}

// Network value and input gradient of one sample, activations in registers.
// Forward: a = tanh(W1 u + b1), b = tanh(W2 a + b2), h = tanh(W3 b + b3),
// y = w4 . h + b4. Backward from dy = 1: d3 = w4 (1 - h^2),
// d2 = (W3^T d3)(1 - b^2), d1 = (W2^T d2)(1 - a^2), du = W1^T d1. Layer-1
// activations are recomputed in the backward loop rather than held.
__device__ __forceinline__ float mlp_value_grad(const float* R, const float u[3], float du[3]) {
  float z2[kMlpH];
#pragma unroll
  for (int i = 0; i < kMlpH; ++i) z2[i] = ldp(R, kMlpB2 + i);
  for (int j = 0; j < kMlpH; ++j) {  // z2 += W2[:, j] a_j (row j of W2^T)
    const float aj = ::tanhf(ldp(R, kMlpB1 + j) + ldp(R, kMlpW1 + 3 * j) * u[0] + ldp(R, kMlpW1 + 3 * j + 1) * u[1] +
                             ldp(R, kMlpW1 + 3 * j + 2) * u[2]);
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) z2[i] += ldp(R, kMlpW2 + kMlpH * j + i) * aj;
  }
#pragma unroll
  for (int i = 0; i < kMlpH; ++i) z2[i] = ::tanhf(z2[i]);  // b
  float acc[kMlpH];
#pragma unroll
  for (int j = 0; j < kMlpH; ++j) acc[j] = 0.f;
  float y = ldp(R, kMlpB4);
  for (int i = 0; i < kMlpH; ++i) {  // row i of W3: forward dot, then W3^T d3 as an axpy with the same row
    float z = ldp(R, kMlpB3 + i);
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) z += ldp(R, kMlpW3 + kMlpH * i + j) * z2[j];
    const float h = ::tanhf(z), w = ldp(R, kMlpW4 + i);
    y += w * h;
    const float d3 = w * (1.f - h * h);
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) acc[j] += ldp(R, kMlpW3 + kMlpH * i + j) * d3;
  }
#pragma unroll
  for (int j = 0; j < kMlpH; ++j) acc[j] *= 1.f - z2[j] * z2[j];  // d2
  du[0] = du[1] = du[2] = 0.f;
  for (int k = 0; k < kMlpH; ++k) {  // (W2^T d2)_k: row k of the transposed copy
    float t = 0.f;
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) t += ldp(R, kMlpW2 + kMlpH * k + j) * acc[j];
    const float w0 = ldp(R, kMlpW1 + 3 * k), w1 = ldp(R, kMlpW1 + 3 * k + 1), w2 = ldp(R, kMlpW1 + 3 * k + 2);
    const float ak = ::tanhf(ldp(R, kMlpB1 + k) + w0 * u[0] + w1 * u[1] + w2 * u[2]);
    const float d1 = t * (1.f - ak * ak);
    du[0] += w0 * d1;
    du[1] += w1 * d1;
    du[2] += w2 * d1;
  }
  return y;
}

// ---- MFMA inference (one wave = 64 points) -------------------------------
// Every layer is computed transposed, C[unit][point] = W X, on
// v_mfma_f32_16x16x32_bf16 (operand layout: A[m = l%16][k = 8(l/16)+e],
// B[k = 8(l/16)+e][n = l%16], C register r = C[4(l/16)+r][l%16]). Lane
// l = 16 g + c holds X[mt][nt][r] = activation of unit 16 mt + 4 g + r at
// point 16 nt + c: the C registers of a layer. The K slots of every product
// are permuted so that these registers are also the next product's B operand:
// K step ks, slot 8 g + e carries unit U(ks, 8g+e) = 16 (2 ks + e/4) + 4 g + e%4,
// i.e. B element e of (nt, ks) is X[2 ks + e/4][nt][e%4]. The packed weights
// (k_mlp_pack_mx) hold A[m][slot] in that order, split in two BF16 terms
// (hi = bf16(w), lo = bf16(w - hi)), and activations enter split the same way,
// so each tile is three MFMAs per K step (lo.hi, hi.lo, then hi.hi): FP32-level
// products with FP32 accumulation. Forward: W2, W3; backward for dF/du: W3^T, W2^T.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float fast_tanh(float x) {  // 1 - 2 / (e^(2x) + 1)
  return 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * x) + 1.f);
}
__device__ __forceinline__ unsigned short bf_bits(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float bf_val(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

// B operand of K step ks for point tile nt, split: hi and lo halves.
__device__ __forceinline__ void pack_b(const float (&X)[4][4][4], int nt, int ks, u32x4_t& hi, u32x4_t& lo) {
  unsigned h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float x0 = X[2 * ks + (2 * q) / 4][nt][(2 * q) % 4], x1 = X[2 * ks + (2 * q + 1) / 4][nt][(2 * q + 1) % 4];
    const unsigned short h0 = bf_bits(x0), h1 = bf_bits(x1);
    const unsigned short l0 = bf_bits(x0 - bf_val(h0)), l1 = bf_bits(x1 - bf_val(h1));
    h[q] = (unsigned)h0 | ((unsigned)h1 << 16);
    l[q] = (unsigned)l0 | ((unsigned)l1 << 16);
  }
  hi = u32x4_t{h[0], h[1], h[2], h[3]};
  lo = u32x4_t{l[0], l[1], l[2], l[3]};
}

// acc = W_mat X (64 x 64 times 64 x 64 points), packed weights Wpk.
__device__ __forceinline__ void mx_layer(const u32x4_t* __restrict__ Wpk, int mat, const float (&X)[4][4][4],
                                         f32x4_t (&acc)[4][4], int lane) {
  u32x4_t bh[4][2], bl[4][2];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) pack_b(X, nt, ks, bh[nt][ks], bl[nt][ks]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    // fragment (mat, mt, ks, hl): Wpk[(((mat * 4 + mt) * 2 + ks) * 2 + hl) * 64 + lane]
    const int f0 = ((mat * 4 + mt) * 2) * 2;
    const bf16x8_t a0h = __builtin_bit_cast(bf16x8_t, Wpk[(f0 + 0) * 64 + lane]);
    const bf16x8_t a0l = __builtin_bit_cast(bf16x8_t, Wpk[(f0 + 1) * 64 + lane]);
    const bf16x8_t a1h = __builtin_bit_cast(bf16x8_t, Wpk[(f0 + 2) * 64 + lane]);
    const bf16x8_t a1l = __builtin_bit_cast(bf16x8_t, Wpk[(f0 + 3) * 64 + lane]);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4_t c = f32x4_t{0.f, 0.f, 0.f, 0.f};  // the small terms first
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0l, __builtin_bit_cast(bf16x8_t, bh[nt][0]), c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1l, __builtin_bit_cast(bf16x8_t, bh[nt][1]), c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0h, __builtin_bit_cast(bf16x8_t, bl[nt][0]), c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1h, __builtin_bit_cast(bf16x8_t, bl[nt][1]), c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0h, __builtin_bit_cast(bf16x8_t, bh[nt][0]), c, 0, 0, 0);
      acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1h, __builtin_bit_cast(bf16x8_t, bh[nt][1]), c, 0, 0, 0);
    }
  }
}

// Network value and input gradient of the wave's 64 points (lane = point);
// every lane of the wave must call it. R: run parameters (FP32, for W1,
// biases, w4); Wpk: packed split-BF16 W2, W3, W3^T, W2^T fragments.
__device__ __forceinline__ float mlp_value_grad_mx(const u32x4_t* __restrict__ Wpk, const float* __restrict__ R,
                                                   const float u[3], float du[3]) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  float un[4][3];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int k = 0; k < 3; ++k) un[nt][k] = __shfl(u[k], 16 * nt + c);
  float X1[4][4][4], X2[4][4][4];
  f32x4_t acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int unit = 16 * mt + 4 * g + r;
      const float w0 = R[kMlpW1 + 3 * unit], w1 = R[kMlpW1 + 3 * unit + 1], w2 = R[kMlpW1 + 3 * unit + 2];
      const float b = R[kMlpB1 + unit];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) X1[mt][nt][r] = fast_tanh(b + w0 * un[nt][0] + w1 * un[nt][1] + w2 * un[nt][2]);
    }
  mx_layer(Wpk, 0, X1, acc, lane);  // W2
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float b = R[kMlpB2 + 16 * mt + 4 * g + r];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) X2[mt][nt][r] = fast_tanh(acc[mt][nt][r] + b);
    }
  mx_layer(Wpk, 1, X2, acc, lane);  // W3
  float yp[4] = {0.f, 0.f, 0.f, 0.f};
  float D[4][4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int unit = 16 * mt + 4 * g + r;
      const float b = R[kMlpB3 + unit], w = R[kMlpW4 + unit];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const float h = fast_tanh(acc[mt][nt][r] + b);
        yp[nt] += w * h;
        D[mt][nt][r] = w * (1.f - h * h);  // d3
      }
    }
  mx_layer(Wpk, 2, D, acc, lane);  // W3^T d3
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int r = 0; r < 4; ++r) D[mt][nt][r] = acc[mt][nt][r] * (1.f - X2[mt][nt][r] * X2[mt][nt][r]);  // d2
  mx_layer(Wpk, 3, D, acc, lane);  // W2^T d2
  float dp[4][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int unit = 16 * mt + 4 * g + r;
      const float w0 = R[kMlpW1 + 3 * unit], w1 = R[kMlpW1 + 3 * unit + 1], w2 = R[kMlpW1 + 3 * unit + 2];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const float d1 = acc[mt][nt][r] * (1.f - X1[mt][nt][r] * X1[mt][nt][r]);
        dp[nt][0] += w0 * d1;
        dp[nt][1] += w1 * d1;
        dp[nt][2] += w2 * d1;
      }
    }
  // sums over the four lane groups (units 4 g + r); lane 16 g + c then reads point tile nt = g
  float y = 0.f;
  du[0] = du[1] = du[2] = 0.f;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float v = yp[nt];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    float d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      d[k] = dp[nt][k];
      d[k] += __shfl_xor(d[k], 16);
      d[k] += __shfl_xor(d[k], 32);
    }
    if (nt == g) {
      y = v;
      du[0] = d[0];
      du[1] = d[1];
      du[2] = d[2];
    }
  }
  return y + R[kMlpB4];
}

// Descriptors t = (ln rho, s/(1+s), alpha/(1+alpha)) and their derivatives
// dt[k] = (dt_k/drho, dt_k/dsigma, dt_k/dtau). s and alpha are formed as the
// analytic meta-GGA sees them: tau floored at 1e-20, sigma floored at 1e-40
// and capped at 8 rho tau (where the cap binds alpha = 0 and sigma follows
// 8 rho tau).
__device__ __forceinline__ void mlp_descriptors(double rho, double sigma, double tau, double t[3], double dt[3][3]) {
  const double r = rho;
  const double r13 = ::cbrt(r);
  const bool tfloor = tau < 1e-20;
  const double tk = tfloor ? 1e-20 : tau;
  const double sfl = 1e-40;
  double sg = sigma > sfl ? sigma : sfl;
  double dsg[3] = {0.0, sigma > sfl ? 1.0 : 0.0, 0.0};  // d sigma_eff / d(rho, sigma, tau)
  if (sg > 8.0 * r * tk) {
    sg = 8.0 * r * tk;
    dsg[0] = 8.0 * tk;
    dsg[1] = 0.0;
    dsg[2] = tfloor ? 0.0 : 8.0 * r;
  }
  const double dtk = tfloor ? 0.0 : 1.0;
  t[0] = ::log(r);
  dt[0][0] = 1.0 / r;
  dt[0][1] = dt[0][2] = 0.0;
  const double s = kSFac * ::sqrt(sg) / (r * r13);
  const double ds_dsg = s / (2.0 * sg), ds_dr = -4.0 / 3.0 * s / r;
  const double ps = 1.0 / ((1.0 + s) * (1.0 + s));
  t[1] = s / (1.0 + s);
  dt[1][0] = ps * (ds_dr + ds_dsg * dsg[0]);
  dt[1][1] = ps * ds_dsg * dsg[1];
  dt[1][2] = ps * ds_dsg * dsg[2];
  const double tu = kTauUnif * r * r13 * r13;
  double al = (tk - sg / (8.0 * r)) / tu;
  double dal[3];
  dal[0] = (sg / (8.0 * r * r) - dsg[0] / (8.0 * r)) / tu - 5.0 / 3.0 * al / r;
  dal[1] = -dsg[1] / (8.0 * r * tu);
  dal[2] = (dtk - dsg[2] / (8.0 * r)) / tu;
  if (al < 0.0) {  // rounding below the cap
    al = 0.0;
    dal[0] = dal[1] = dal[2] = 0.0;
  }
  const double pa = 1.0 / ((1.0 + al) * (1.0 + al));
  t[2] = al / (1.0 + al);
  for (int k = 0; k < 3; ++k) dt[2][k] = pa * dal[k];
}

// Surrogate at one point: eps (per particle) and v_rho, v_sigma, v_tau, or the
// analytic functional where the point is below rho_min or outside the box
// (returns false then). Every lane runs the network.
template <bool MX>
__device__ __forceinline__ bool mlp_point(const float* R, const u32x4_t* Wpk, const MlpBox& box, const int* codes,
                                          int ncodes, double rho, double sigma, double tau, double* eps, double* vr,
                                          double* vs, double* vt) {
  const double r = rho > 1e-300 ? rho : 1e-300;
  double t[3], dt[3][3];
  mlp_descriptors(r, sigma, tau, t, dt);
  bool in = rho >= box.rho_min;
  float u[3];
  double hinv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lo = box.lo[k], hi = box.hi[k];
    in = in && t[k] >= lo && t[k] <= hi;
    hinv[k] = 2.0 / (hi - lo);
    u[k] = (float)((t[k] - 0.5 * (lo + hi)) * hinv[k]);
  }
  float du[3];
  const float F = MX ? mlp_value_grad_mx(Wpk, R, u, du) : mlp_value_grad(R, u, du);
  if (in) {
    double g[3];  // dF/dt
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = (double)du[k] * hinv[k];
    const double r13 = ::cbrt(r);
    const double eu = kCxUnif * r * r13;
    double dF[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dF[c] = g[0] * dt[0][c] + g[1] * dt[1][c] + g[2] * dt[2][c];
    double e0 = 0.0, r0 = 0.0, s0 = 0.0, t0 = 0.0;  // the analytic base, if any
    if (box.nbase > 0) eval_codes(box.base, box.nbase, rho, sigma, tau, &e0, &r0, &s0, &t0);
    *eps = e0 + eu * (double)F / r;
    *vr = r0 + (4.0 / 3.0) * kCxUnif * r13 * (double)F + eu * dF[0];
    *vs = s0 + eu * dF[1];
    *vt = t0 + eu * dF[2];
    return true;
  }
  eval_codes(codes, ncodes, rho, sigma, tau, eps, vr, vs, vt);
  return false;
}

__device__ __forceinline__ int run_index(int k) {
  if (k >= kMlpW2 && k < kMlpW2 + kMlpH * kMlpH) {
    const int o = k - kMlpW2;
    return kMlpW2 + kMlpH * (o % kMlpH) + o / kMlpH;
  }
  return k;
}

__device__ __forceinline__ float bf16_round(float x) {
  return __builtin_bit_cast(float, (unsigned)__builtin_bit_cast(unsigned short, (__bf16)x) << 16);
}

}  // namespace

// Run copies of a canonical parameter block: FP32 (training) and BF16-rounded
// (inference), both with W2 transposed.
__global__ void k_mlp_pack(const float* __restrict__ P, float* __restrict__ Prun, float* __restrict__ Pinf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kMlpParams) return;
  const float p = P[k];
  Prun[run_index(k)] = p;
  Pinf[run_index(k)] = bf16_round(p);
}

// Raw network on normalised inputs U (n x 3): F (n) and dF/dU (n x 3); the
// per-lane FP32 network, or with Wpk the MFMA network of the inference path.
__global__ __launch_bounds__(64) void k_mlp_forward(long n, const float* __restrict__ U, const float* R,
                                                    const u32x4_t* __restrict__ Wpk, float* __restrict__ F,
                                                    float* __restrict__ dF) {
  const long i0 = (long)blockIdx.x * kLanes + threadIdx.x;
  const long i = i0 < n ? i0 : n - 1;  // the MFMA network needs every lane of the wave
  const float u[3] = {U[3 * i], U[3 * i + 1], U[3 * i + 2]};
  float du[3];
  const float y = Wpk ? mlp_value_grad_mx(Wpk, R, u, du) : mlp_value_grad(R, u, du);
  if (i0 >= n) return;
  F[i] = y;
  if (dF) {
    dF[3 * i] = du[0];
    dF[3 * i + 1] = du[1];
    dF[3 * i + 2] = du[2];
  }
}

// Packed split-BF16 A-operand fragments of the MFMA network: for matrix mat
// (0 W2, 1 W3, 2 W3^T, 3 W2^T), m tile mt, K step ks, term hl (0 hi, 1 lo),
// lane l = 16 g + c, element e: A[m = 16 mt + c][slot 8 g + e] with the slot's
// unit U = 16 (2 ks + e/4) + 4 g + e%4.
__global__ void k_mlp_pack_mx(const float* __restrict__ P, unsigned short* __restrict__ Wpk) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // ((((mat * 4 + mt) * 2 + ks) * 2 + hl) * 64 + lane) * 8 + e
  if (idx >= 4 * 4 * 2 * 2 * 64 * 8) return;
  const int e = idx & 7, lane = (idx >> 3) & 63, hl = (idx >> 9) & 1, ks = (idx >> 10) & 1, mt = (idx >> 11) & 3,
            mat = idx >> 13;
  const int g = lane >> 4, c = lane & 15;
  const int m = 16 * mt + c, un = 16 * (2 * ks + e / 4) + 4 * g + e % 4;
  float w;
  switch (mat) {
    case 0: w = P[kMlpW2 + kMlpH * m + un]; break;
    case 1: w = P[kMlpW3 + kMlpH * m + un]; break;
    case 2: w = P[kMlpW3 + kMlpH * un + m]; break;
    default: w = P[kMlpW2 + kMlpH * un + m]; break;
  }
  const unsigned short h = bf_bits(w);
  Wpk[idx] = hl ? bf_bits(w - bf_val(h)) : h;
}

// Loss L = cs sum_i C_i (y_i - T_i)^2 over the samples [0, n) (C_i = 1 when C
// is null) and its gradient with respect to the canonical parameters, added
// to G (and L to *loss). A block takes chunks of 64 samples. Per chunk, each
// lane runs its sample's forward pass and layer deltas, and lane t then
// reduces row t of every weight gradient over the chunk's samples
// (dW = D A^T). Activations are held sample-major [sample][unit], deltas
// unit-major [unit][sample], both with a 65-float pitch, so the per-lane and
// the per-row accesses are conflict-free. The rows accumulate in registers
// over the block's chunks and are added to G once.
__global__ __launch_bounds__(64) void k_mlp_grad(long n, const float* __restrict__ U, const float* __restrict__ T,
                                                 const float* __restrict__ C, float cs, const float* R,
                                                 float* __restrict__ G, float* __restrict__ loss) {
  __shared__ float A1[kLanes][kMlpH + 1];  // layer-1 activations, [sample][unit]
  __shared__ float A2[kLanes][kMlpH + 1];  // layer-2 activations, [sample][unit]
  __shared__ float D[kMlpH][kLanes + 1];   // h3, then d3, d2, d1: [unit][sample]
  __shared__ float Us[kLanes][4];
  __shared__ float gs[kLanes];
  const int lane = threadIdx.x;
  float gW2[kMlpH], gW3[kMlpH];
#pragma unroll
  for (int j = 0; j < kMlpH; ++j) gW2[j] = gW3[j] = 0.f;
  float gW1[3] = {0.f, 0.f, 0.f}, gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f, lsum = 0.f;
  const long nchunks = (n + kLanes - 1) / kLanes;
  for (long c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const long i = c * kLanes + lane;
    const bool live = i < n;
    const long ic = live ? i : n - 1;
    const float u[3] = {U[3 * ic], U[3 * ic + 1], U[3 * ic + 2]};
    __syncthreads();  // the previous chunk's readers are done
    Us[lane][0] = u[0];
    Us[lane][1] = u[1];
    Us[lane][2] = u[2];
    // forward, own sample
    float z2[kMlpH];
#pragma unroll
    for (int q = 0; q < kMlpH; ++q) z2[q] = ldp(R, kMlpB2 + q);
    for (int j = 0; j < kMlpH; ++j) {
      const float aj = ::tanhf(ldp(R, kMlpB1 + j) + ldp(R, kMlpW1 + 3 * j) * u[0] +
                               ldp(R, kMlpW1 + 3 * j + 1) * u[1] + ldp(R, kMlpW1 + 3 * j + 2) * u[2]);
      A1[lane][j] = aj;
#pragma unroll
      for (int q = 0; q < kMlpH; ++q) z2[q] += ldp(R, kMlpW2 + kMlpH * j + q) * aj;
    }
#pragma unroll
    for (int q = 0; q < kMlpH; ++q) {
      z2[q] = ::tanhf(z2[q]);
      A2[lane][q] = z2[q];
    }
    float y = ldp(R, kMlpB4);
    for (int q = 0; q < kMlpH; ++q) {
      float z = ldp(R, kMlpB3 + q);
#pragma unroll
      for (int j = 0; j < kMlpH; ++j) z += ldp(R, kMlpW3 + kMlpH * q + j) * z2[j];
      const float h = ::tanhf(z);
      D[q][lane] = h;
      y += ldp(R, kMlpW4 + q) * h;
    }
    const float cw = live ? cs * (C ? C[ic] : 1.f) : 0.f;
    const float e = live ? y - T[ic] : 0.f;
    lsum += cw * e * e;
    const float g = 2.f * cw * e;  // dL/dy
    gs[lane] = g;
    __syncthreads();
    // w4, b4: row `lane` of h3 against dL/dy
    {
      float t = 0.f, b = 0.f;
      for (int q = 0; q < kLanes; ++q) {
        t += gs[q] * D[lane][q];
        b += gs[q];
      }
      gw4 += t;
      gb4 += b;  // every lane holds the same sum; lane 0 adds it to G
    }
    __syncthreads();
    // d3, own column
#pragma unroll
    for (int q = 0; q < kMlpH; ++q) {
      const float h = D[q][lane];
      D[q][lane] = g * ldp(R, kMlpW4 + q) * (1.f - h * h);
    }
    __syncthreads();
    // row `lane` of dW3 = d3 a2^T, db3
    for (int q = 0; q < kLanes; ++q) {
      const float d = D[lane][q];
      gb3 += d;
#pragma unroll
      for (int j = 0; j < kMlpH; ++j) gW3[j] += d * A2[q][j];
    }
    __syncthreads();
    // d2 = (W3^T d3)(1 - a2^2), own column
    {
      float acc[kMlpH];
#pragma unroll
      for (int j = 0; j < kMlpH; ++j) acc[j] = 0.f;
      for (int q = 0; q < kMlpH; ++q) {
        const float d3 = D[q][lane];
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) acc[j] += ldp(R, kMlpW3 + kMlpH * q + j) * d3;
      }
#pragma unroll
      for (int j = 0; j < kMlpH; ++j) {
        const float b = A2[lane][j];
        D[j][lane] = acc[j] * (1.f - b * b);
      }
    }
    __syncthreads();
    // row `lane` of dW2 = d2 a1^T, db2
    for (int q = 0; q < kLanes; ++q) {
      const float d = D[lane][q];
      gb2 += d;
#pragma unroll
      for (int j = 0; j < kMlpH; ++j) gW2[j] += d * A1[q][j];
    }
    __syncthreads();
    // d1 = (W2^T d2)(1 - a1^2), own column
    {
      float d2[kMlpH];
#pragma unroll
      for (int j = 0; j < kMlpH; ++j) d2[j] = D[j][lane];
      for (int k = 0; k < kMlpH; ++k) {
        float t = 0.f;
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) t += ldp(R, kMlpW2 + kMlpH * k + j) * d2[j];
        const float a = A1[lane][k];
        D[k][lane] = t * (1.f - a * a);
      }
    }
    __syncthreads();
    // row `lane` of dW1 = d1 u^T, db1
    for (int q = 0; q < kLanes; ++q) {
      const float d = D[lane][q];
      gb1 += d;
      gW1[0] += d * Us[q][0];
      gW1[1] += d * Us[q][1];
      gW1[2] += d * Us[q][2];
    }
  }
  // canonical layout: W[out][in], row `lane` of each matrix
#pragma unroll
  for (int j = 0; j < kMlpH; ++j) {
    atomicAdd(&G[kMlpW2 + kMlpH * lane + j], gW2[j]);
    atomicAdd(&G[kMlpW3 + kMlpH * lane + j], gW3[j]);
  }
  atomicAdd(&G[kMlpW1 + 3 * lane], gW1[0]);
  atomicAdd(&G[kMlpW1 + 3 * lane + 1], gW1[1]);
  atomicAdd(&G[kMlpW1 + 3 * lane + 2], gW1[2]);
  atomicAdd(&G[kMlpB1 + lane], gb1);
  atomicAdd(&G[kMlpB2 + lane], gb2);
  atomicAdd(&G[kMlpB3 + lane], gb3);
  atomicAdd(&G[kMlpW4 + lane], gw4);
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if (lane == 0) {
    atomicAdd(&G[kMlpB4], gb4);
    atomicAdd(loss, lsum);
  }
}

// Adam (Kingma and Ba, ICLR 2015) on the canonical FP32 weights; clears the
// gradient and refreshes both run copies.
__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ G, float* __restrict__ M,
                           float* __restrict__ V, float* __restrict__ Prun, float* __restrict__ Pinf, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kMlpParams) return;
  const float g = G[k];
  const float m = b1 * M[k] + (1.f - b1) * g, v = b2 * V[k] + (1.f - b2) * g * g;
  M[k] = m;
  V[k] = v;
  const float p = P[k] - lr * (m / c1) / (::sqrtf(v / c2) + eps);
  P[k] = p;
  G[k] = 0.f;
  Prun[run_index(k)] = p;
  Pinf[run_index(k)] = bf16_round(p);
}

// Training samples: descriptors t (n x 3) and target F = e_xc / e_x^unif of
// the analytic functional (codes) at each (rho, sigma, tau).
__global__ void k_mlp_features(long n, const double* __restrict__ rho, const double* __restrict__ sigma,
                               const double* __restrict__ tau, const int* __restrict__ codes, int ncodes,
                               float* __restrict__ tout, float* __restrict__ Fout) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i] > 1e-300 ? rho[i] : 1e-300;
  double t[3], dt[3][3];
  mlp_descriptors(r, sigma[i], tau[i], t, dt);
  tout[3 * i] = (float)t[0];
  tout[3 * i + 1] = (float)t[1];
  tout[3 * i + 2] = (float)t[2];
  double e, vr, vs, vt;
  eval_codes(codes, ncodes, rho[i], sigma[i], tau[i], &e, &vr, &vs, &vt);
  Fout[i] = (float)(e / (kCxUnif * ::cbrt(r)));  // (rho eps) / (kCxUnif rho^(4/3))
}

// Surrogate at n arbitrary points; *nfall counts the analytic fallbacks.
__global__ __launch_bounds__(64) void k_mlp_eval(long n, const float* R, const u32x4_t* __restrict__ Wpk, MlpBox box,
                                                 const int* __restrict__ codes,
                                                 int ncodes, const double* __restrict__ rho,
                                                 const double* __restrict__ sigma, const double* __restrict__ tau,
                                                 double* __restrict__ exc, double* __restrict__ vrho,
                                                 double* __restrict__ vsigma, double* __restrict__ vtau,
                                                 int* __restrict__ nfall) {
  const long i0 = (long)blockIdx.x * kLanes + threadIdx.x;
  const bool live = i0 < n;
  const long i = live ? i0 : n - 1;  // tail lanes run a copy of the last point and store nothing
  double e, vr, vs, vt;
  const bool net = mlp_point<true>(R, Wpk, box, codes, ncodes, rho[i], sigma[i], tau[i], &e, &vr, &vs, &vt);
  if (live) {
    exc[i] = e;
    vrho[i] = vr;
    vsigma[i] = vs;
    vtau[i] = vt;
    if (!net) atomicAdd(nfall, 1);
  }
}

// The Evaluate step of a meta-GGA grid with the surrogate (the counterpart of
// k_functional): rho, grad rho and tau from the density partials, the network
// (or the analytic fallback) per point, the Vxc integrand coefficients, and
// per-block sums of E = sum w rho eps and N = sum w rho.
__global__ __launch_bounds__(64) void k_functional_mlp(int nslots, const int* __restrict__ grp_off,
                                                       const double* __restrict__ part,
                                                       const double* __restrict__ part_tau,
                                                       const double* __restrict__ bw, const float* R,
                                                       const u32x4_t* __restrict__ Wpk, MlpBox box,
                                                       const int* __restrict__ codes, int ncodes, int want_e,
                                                       int want_v, double* __restrict__ rho_out,
                                                       double* __restrict__ coef, double* __restrict__ en_part,
                                                       int* __restrict__ nfall) {
  const int slot0 = blockIdx.x * kLanes + threadIdx.x;
  const bool live = slot0 < nslots;
  const int slot = live ? slot0 : nslots - 1;
  const int b = slot / kBatch, g = slot % kBatch;
  double rho = 0, gx = 0, gy = 0, gz = 0, tau = 0;
  for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
    const double* p = part + (size_t)k * 4 * kBatch;
    rho += p[g];
    gx += p[kBatch + g];
    gy += p[2 * kBatch + g];
    gz += p[3 * kBatch + g];
    const double* q = part_tau + (size_t)k * 3 * kBatch;
    tau += q[g] + q[kBatch + g] + q[2 * kBatch + g];
  }
  gx *= 2; gy *= 2; gz *= 2;
  tau *= 0.5;
  const double w = bw[slot];
  const double sigma = gx * gx + gy * gy + gz * gz;
  double eps, vr, vs, vt;
  const bool net = mlp_point<true>(R, Wpk, box, codes, ncodes, rho, sigma, tau, &eps, &vr, &vs, &vt);
  double e_acc = 0, n_acc = 0;
  if (live) {
    rho_out[slot] = rho;
    if (want_v) {
      coef[kCoef * (size_t)slot + 0] = 0.5 * w * vr;
      const double f = 2.0 * w * vs;
      coef[kCoef * (size_t)slot + 1] = f * gx;
      coef[kCoef * (size_t)slot + 2] = f * gy;
      coef[kCoef * (size_t)slot + 3] = f * gz;
      coef[kCoef * (size_t)slot + 4] = 0.25 * w * vt;
    }
    if (want_e) e_acc = w * rho * eps;
    n_acc = w * rho;
    if (!net && rho >= kDensThreshold && w != 0.0) atomicAdd(nfall, 1);
  }
  for (int o = 32; o > 0; o >>= 1) {
    e_acc += __shfl_xor(e_acc, o);
    n_acc += __shfl_xor(n_acc, o);
  }
  if (threadIdx.x == 0) {
    en_part[2 * blockIdx.x] = e_acc;
    en_part[2 * blockIdx.x + 1] = n_acc;
  }
}

// Per-point density of a grid after getDensity: rho, sigma, tau, weight (slot order).
__global__ void k_point_density(int nslots, const int* __restrict__ grp_off, const double* __restrict__ part,
                                const double* __restrict__ part_tau, int has_tau, const double* __restrict__ bw,
                                double* __restrict__ rho_o, double* __restrict__ sig_o, double* __restrict__ tau_o,
                                double* __restrict__ w_o) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int b = slot / kBatch, g = slot % kBatch;
  double rho = 0, gx = 0, gy = 0, gz = 0, tau = 0;
  for (int k = grp_off[b]; k < grp_off[b + 1]; ++k) {
    const double* p = part + (size_t)k * 4 * kBatch;
    rho += p[g];
    gx += p[kBatch + g];
    gy += p[2 * kBatch + g];
    gz += p[3 * kBatch + g];
    if (has_tau) {
      const double* q = part_tau + (size_t)k * 3 * kBatch;
      tau += q[g] + q[kBatch + g] + q[2 * kBatch + g];
    }
  }
  rho_o[slot] = rho;
  sig_o[slot] = 4.0 * (gx * gx + gy * gy + gz * gz);
  tau_o[slot] = 0.5 * tau;
  w_o[slot] = bw[slot];
}

// ---- launchers
void launch_mlp_pack(hipStream_t st, const float* P, float* Prun, float* Pinf) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, P, Prun, Pinf);
  CHX_HIP(hipGetLastError());
}

void launch_mlp_forward(hipStream_t st, long n, const float* U, const float* R, const unsigned* Wpk, float* F,
                        float* dF) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + kLanes - 1) / kLanes)), dim3(kLanes), 0, st, n, U, R,
                     reinterpret_cast<const u32x4_t*>(Wpk), F, dF);
  CHX_HIP(hipGetLastError());
}

void launch_mlp_pack_mx(hipStream_t st, const float* P, unsigned* Wpk) {
  hipLaunchKernelGGL(k_mlp_pack_mx, dim3(4 * 4 * 2 * 2 * 64 * 8 / 256), dim3(256), 0, st, P,
                     reinterpret_cast<unsigned short*>(Wpk));
  CHX_HIP(hipGetLastError());
}

void launch_mlp_grad(hipStream_t st, long n, const float* U, const float* T, const float* C, float cs, const float* R,
                     float* G, float* loss, int max_blocks) {
  if (n <= 0) return;
  const long nchunks = (n + kLanes - 1) / kLanes;
  const unsigned nb = (unsigned)(nchunks < max_blocks ? nchunks : max_blocks);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(kLanes), 0, st, n, U, T, C, cs, R, G, loss);
  CHX_HIP(hipGetLastError());
}

void launch_mlp_adam(hipStream_t st, float* P, float* G, float* M, float* V, float* Prun, float* Pinf, float lr,
                     int step) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, P, G, M, V, Prun, Pinf, lr, b1, b2,
                     eps, c1, c2);
  CHX_HIP(hipGetLastError());
}

void launch_mlp_features(hipStream_t st, long n, const double* rho, const double* sigma, const double* tau,
                         const int* codes, int ncodes, float* t, float* F) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, rho, sigma, tau, codes,
                     ncodes, t, F);
  CHX_HIP(hipGetLastError());
}

void launch_mlp_eval(hipStream_t st, long n, const float* R, const unsigned* Wpk, const MlpBox& box, const int* codes,
                     int ncodes, const double* rho, const double* sigma, const double* tau, double* exc, double* vrho,
                     double* vsigma, double* vtau, int* nfall) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)((n + kLanes - 1) / kLanes)), dim3(kLanes), 0, st, n, R,
                     reinterpret_cast<const u32x4_t*>(Wpk), box, codes, ncodes, rho, sigma, tau, exc, vrho, vsigma,
                     vtau, nfall);
  CHX_HIP(hipGetLastError());
}

void launch_functional_mlp(hipStream_t st, int nbatches, const int* grp_off, const double* part,
                           const double* part_tau, const double* bw, const float* R, const unsigned* Wpk,
                           const MlpBox& box, const int* codes, int ncodes, int want_e, int want_v, double* rho_out,
                           double* coef, double* en_part, int* nfall) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_functional_mlp, dim3((unsigned)((nslots + kLanes - 1) / kLanes)), dim3(kLanes), 0, st, nslots,
                     grp_off, part, part_tau, bw, R, reinterpret_cast<const u32x4_t*>(Wpk), box, codes, ncodes, want_e,
                     want_v, rho_out, coef, en_part, nfall);
  CHX_HIP(hipGetLastError());
}

void launch_point_density(hipStream_t st, int nbatches, const int* grp_off, const double* part,
                          const double* part_tau, int has_tau, const double* bw, double* rho, double* sigma,
                          double* tau, double* w) {
  const int nslots = nbatches * kBatch;
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_point_density, dim3((nslots + 255) / 256), dim3(256), 0, st, nslots, grp_off, part, part_tau,
                     has_tau, bw, rho, sigma, tau, w);
  CHX_HIP(hipGetLastError());
}

}  // namespace chx
