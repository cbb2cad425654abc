// FP64 CPU restatement of the learned functional surrogate (test
// infrastructure: the checker of the GPU training and inference kernels; the
// shipped library never links or calls it). The surrogate has no counterpart
// in the reference, whose functionals come from libxc
// (src/ExchangeCorrelation.cpp:86-163); the analytic fallback below is the
// oracle's own restatement of those functionals (functionals.hpp).
//
// Network: u (3) -> tanh(W1 u + b1) -> tanh(W2 a + b2) -> tanh(W3 b + b3)
// -> w4 . h + b4; canonical parameter block, matrices [out][in]:
// W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4.
#include <cmath>
#include <vector>

#include "functionals.hpp"
#include "oracle.h"

namespace {

constexpr int H = 64;
constexpr int oW1 = 0, oB1 = 192, oW2 = 256, oB2 = 4352, oW3 = 4416, oB3 = 8512, oW4 = 8576, oB4 = 8640, NP = 8641;

struct Acts {
  double a[H], b[H], h[H], y;
};

void forward(const double* P, const double* u, Acts& s) {
  for (int i = 0; i < H; ++i) {
    double z = P[oB1 + i];
    for (int k = 0; k < 3; ++k) z += P[oW1 + 3 * i + k] * u[k];
    s.a[i] = std::tanh(z);
  }
  for (int i = 0; i < H; ++i) {
    double z = P[oB2 + i];
    for (int j = 0; j < H; ++j) z += P[oW2 + H * i + j] * s.a[j];
    s.b[i] = std::tanh(z);
  }
  s.y = P[oB4];
  for (int i = 0; i < H; ++i) {
    double z = P[oB3 + i];
    for (int j = 0; j < H; ++j) z += P[oW3 + H * i + j] * s.b[j];
    s.h[i] = std::tanh(z);
    s.y += P[oW4 + i] * s.h[i];
  }
}

// Back-propagates dL/dy = g: layer deltas, optional parameter gradient (added
// to G) and input gradient du.
void backward(const double* P, const double* u, const Acts& s, double g, double* G, double* du) {
  double d3[H], d2[H], d1[H];
  for (int i = 0; i < H; ++i) d3[i] = g * P[oW4 + i] * (1.0 - s.h[i] * s.h[i]);
  for (int j = 0; j < H; ++j) {
    double t = 0;
    for (int i = 0; i < H; ++i) t += P[oW3 + H * i + j] * d3[i];
    d2[j] = t * (1.0 - s.b[j] * s.b[j]);
  }
  for (int k = 0; k < H; ++k) {
    double t = 0;
    for (int j = 0; j < H; ++j) t += P[oW2 + H * j + k] * d2[j];
    d1[k] = t * (1.0 - s.a[k] * s.a[k]);
  }
  if (G) {
    G[oB4] += g;
    for (int i = 0; i < H; ++i) {
      G[oW4 + i] += g * s.h[i];
      G[oB3 + i] += d3[i];
      G[oB2 + i] += d2[i];
      G[oB1 + i] += d1[i];
      for (int j = 0; j < H; ++j) {
        G[oW3 + H * i + j] += d3[i] * s.b[j];
        G[oW2 + H * i + j] += d2[i] * s.a[j];
      }
      for (int k = 0; k < 3; ++k) G[oW1 + 3 * i + k] += d1[i] * u[k];
    }
  }
  if (du)
    for (int k = 0; k < 3; ++k) {
      double t = 0;
      for (int i = 0; i < H; ++i) t += P[oW1 + 3 * i + k] * d1[i];
      du[k] = t;
    }
}

constexpr double kCx = -0.7385587663820224;      // -(3/4)(3/pi)^(1/3)
constexpr double kSFac = 0.16162045967399548;    // 1 / (2 (3 pi^2)^(1/3))
constexpr double kTauUnif = 2.8712340001881915;  // (3/10)(3 pi^2)^(2/3)

// Descriptors t = (ln rho, s/(1+s), alpha/(1+alpha)) of (rho, sigma, tau)
// after the meta-GGA input rules (tau >= 1e-20, 1e-40 <= sigma <= 8 rho tau),
// with dt[k][c] = d t_k / d (rho, sigma, tau)_c.
void descriptors(double rho, double sigma, double tau, double t[3], double dt[3][3]) {
  const bool tfloor = tau < 1e-20;
  const double tk = tfloor ? 1e-20 : tau;
  double sg = sigma > 1e-40 ? sigma : 1e-40;
  double dsg[3] = {0.0, sigma > 1e-40 ? 1.0 : 0.0, 0.0};
  if (sg > 8.0 * rho * tk) {
    sg = 8.0 * rho * tk;
    dsg[0] = 8.0 * tk;
    dsg[1] = 0.0;
    dsg[2] = tfloor ? 0.0 : 8.0 * rho;
  }
  t[0] = std::log(rho);
  dt[0][0] = 1.0 / rho;
  dt[0][1] = dt[0][2] = 0.0;
  const double r43 = std::pow(rho, 4.0 / 3.0);
  const double s = kSFac * std::sqrt(sg) / r43;
  const double ps = 1.0 / ((1.0 + s) * (1.0 + s));
  const double ds[3] = {-4.0 / 3.0 * s / rho + s / (2.0 * sg) * dsg[0], s / (2.0 * sg) * dsg[1],
                        s / (2.0 * sg) * dsg[2]};
  t[1] = s / (1.0 + s);
  for (int c = 0; c < 3; ++c) dt[1][c] = ps * ds[c];
  const double tu = kTauUnif * std::pow(rho, 5.0 / 3.0);
  double al = (tk - sg / (8.0 * rho)) / tu;
  // d alpha: numerator N = tk - sg/(8 rho), alpha = N / tu, d tu / d rho = 5/3 tu / rho
  double dal[3] = {(sg / (8.0 * rho * rho) - dsg[0] / (8.0 * rho)) / tu - 5.0 / 3.0 * al / rho,
                   -dsg[1] / (8.0 * rho * tu), ((tfloor ? 0.0 : 1.0) - dsg[2] / (8.0 * rho)) / tu};
  if (al < 0) {
    al = 0;
    dal[0] = dal[1] = dal[2] = 0;
  }
  const double pa = 1.0 / ((1.0 + al) * (1.0 + al));
  t[2] = al / (1.0 + al);
  for (int c = 0; c < 3; ++c) dt[2][c] = pa * dal[c];
}

}  // namespace

extern "C" {

int ox_mlp_forward(long n, const double* U, const double* P, double* F, double* dU) {
  for (long i = 0; i < n; ++i) {
    Acts s;
    forward(P, U + 3 * i, s);
    F[i] = s.y;
    if (dU) backward(P, U + 3 * i, s, 1.0, nullptr, dU + 3 * i);
  }
  return 0;
}

int ox_mlp_loss_grad(long n, const double* U, const double* T, const double* C, double cs, const double* P,
                     double* G, double* loss) {
  for (int k = 0; k < NP; ++k) G[k] = 0;
  double L = 0;
  for (long i = 0; i < n; ++i) {
    Acts s;
    forward(P, U + 3 * i, s);
    const double c = cs * (C ? C[i] : 1.0), e = s.y - T[i];
    L += c * e * e;
    backward(P, U + 3 * i, s, 2.0 * c * e, G, nullptr);
  }
  if (loss) *loss = L;
  return 0;
}

int ox_mlp_descriptors(long n, const double* rho, const double* sigma, const double* tau, double* t, double* dt) {
  for (long i = 0; i < n; ++i) {
    double ti[3], d[3][3];
    descriptors(rho[i], sigma[i], tau[i], ti, d);
    for (int k = 0; k < 3; ++k) {
      t[3 * i + k] = ti[k];
      if (dt)
        for (int c = 0; c < 3; ++c) dt[9 * i + 3 * k + c] = d[k][c];
    }
  }
  return 0;
}

int ox_mlp_point(const double* P, const double* lo, const double* hi, double rho_min, int ncodes, const int* codes,
                 int nbase, const int* base, long n, const double* rho, const double* sigma, const double* tau,
                 double* exc, double* vrho, double* vsigma, double* vtau, int* used_net) {
  for (long i = 0; i < n; ++i) {
    bool in = rho[i] >= rho_min;
    double t[3], dt[3][3], u[3];
    if (in) {
      descriptors(rho[i], sigma[i], tau[i], t, dt);
      for (int k = 0; k < 3; ++k) {
        in = in && t[k] >= lo[k] && t[k] <= hi[k];
        u[k] = (t[k] - 0.5 * (lo[k] + hi[k])) * 2.0 / (hi[k] - lo[k]);
      }
    }
    if (in) {
      Acts s;
      forward(P, u, s);
      double du[3];
      backward(P, u, s, 1.0, nullptr, du);
      const double eu = kCx * std::pow(rho[i], 4.0 / 3.0);
      double dF[3] = {0, 0, 0};
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) dF[c] += du[k] * 2.0 / (hi[k] - lo[k]) * dt[k][c];
      // the analytic base (the network predicts F - F_base), then the network's part
      double e0 = 0, r0 = 0, s0 = 0, t0 = 0;
      for (int k = 0; k < nbase; ++k) oxf::eval_point(base[k], rho[i], sigma[i], tau[i], &e0, &r0, &s0, &t0);
      exc[i] = e0 + eu * s.y / rho[i];
      vrho[i] = r0 + 4.0 / 3.0 * kCx * std::cbrt(rho[i]) * s.y + eu * dF[0];
      vsigma[i] = s0 + eu * dF[1];
      vtau[i] = t0 + eu * dF[2];
    } else {
      double e = 0, vr = 0, vs = 0, vt = 0;
      for (int k = 0; k < ncodes; ++k) oxf::eval_point(codes[k], rho[i], sigma[i], tau[i], &e, &vr, &vs, &vt);
      exc[i] = e;
      vrho[i] = vr;
      vsigma[i] = vs;
      vtau[i] = vt;
    }
    if (used_net) used_net[i] = in ? 1 : 0;
  }
  return 0;
}

}  // extern "C"
