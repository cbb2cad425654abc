// Device-side exchange-correlation functionals, spin-unpolarised, FP64.
//
// Replaces the libxc calls of src/ExchangeCorrelation.cpp:86-163 (summed over
// the functional codes of a .df file, :202-217). Each functional is coded as
// its energy density e = rho*eps(rho, sigma, tau); v_rho, v_sigma and v_tau
// are exact derivatives obtained by forward-mode dual numbers, so there is no
// hand-differentiated code to drift from the energy.
//
// libxc conventions: eps per particle; sigma = |grad rho|^2;
// tau = 1/2 sum_i |grad psi_i|^2. Below the density threshold a code
// contributes nothing; sigma is floored at thr^(8/3).
#pragma once
#include <hip/hip_runtime.h>

namespace chx {

struct D3 {
  double v, d0, d1, d2;
};
__device__ __forceinline__ D3 dconst(double x) { return {x, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ D3 dchain(const D3& a, double f, double df) { return {f, df * a.d0, df * a.d1, df * a.d2}; }
__device__ __forceinline__ D3 operator+(const D3& a, const D3& b) { return {a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
__device__ __forceinline__ D3 operator-(const D3& a, const D3& b) { return {a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
__device__ __forceinline__ D3 operator-(const D3& a) { return {-a.v, -a.d0, -a.d1, -a.d2}; }
__device__ __forceinline__ D3 operator*(const D3& a, const D3& b) {
  return {a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
__device__ __forceinline__ D3 operator/(const D3& a, const D3& b) {
  const double r = 1.0 / b.v;
  const double q = a.v * r;
  return {q, (a.d0 - q * b.d0) * r, (a.d1 - q * b.d1) * r, (a.d2 - q * b.d2) * r};
}
__device__ __forceinline__ D3 operator+(const D3& a, double b) { return {a.v + b, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 operator+(double a, const D3& b) { return {a + b.v, b.d0, b.d1, b.d2}; }
__device__ __forceinline__ D3 operator-(const D3& a, double b) { return {a.v - b, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 operator-(double a, const D3& b) { return {a - b.v, -b.d0, -b.d1, -b.d2}; }
__device__ __forceinline__ D3 operator*(const D3& a, double b) { return {a.v * b, a.d0 * b, a.d1 * b, a.d2 * b}; }
__device__ __forceinline__ D3 operator*(double a, const D3& b) { return {a * b.v, a * b.d0, a * b.d1, a * b.d2}; }
__device__ __forceinline__ D3 operator/(const D3& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ D3 operator/(double a, const D3& b) { return dconst(a) / b; }
__device__ __forceinline__ D3 dexp(const D3& a) { const double e = ::exp(a.v); return dchain(a, e, e); }
__device__ __forceinline__ D3 dlog(const D3& a) { return dchain(a, ::log(a.v), 1.0 / a.v); }
__device__ __forceinline__ D3 dsqrt(const D3& a) { const double s = ::sqrt(a.v); return dchain(a, s, 0.5 / s); }
__device__ __forceinline__ D3 dpow(const D3& a, double p) { const double f = ::pow(a.v, p); return dchain(a, f, p * f / a.v); }
__device__ __forceinline__ D3 datan(const D3& a) { return dchain(a, ::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
__device__ __forceinline__ D3 dasinh(const D3& a) { return dchain(a, ::asinh(a.v), 1.0 / ::sqrt(1.0 + a.v * a.v)); }

constexpr double kPi = 3.14159265358979323846;

// Second-order forward mode in the two variables (rho, sigma): value,
// gradient and the three Hessian entries, for the kernel f_xc of the response
// path (Evaluate("f")). The LDA/GGA energy densities below are templates on
// the number type, so f_xc is the exact second derivative of the same code.
struct H2 {
  double v, g0, g1, h00, h01, h11;
};
__device__ __forceinline__ H2 hchain(const H2& a, double f, double f1, double f2) {
  return {f, f1 * a.g0, f1 * a.g1, f1 * a.h00 + f2 * a.g0 * a.g0, f1 * a.h01 + f2 * a.g0 * a.g1,
          f1 * a.h11 + f2 * a.g1 * a.g1};
}
__device__ __forceinline__ H2 operator+(const H2& a, const H2& b) {
  return {a.v + b.v, a.g0 + b.g0, a.g1 + b.g1, a.h00 + b.h00, a.h01 + b.h01, a.h11 + b.h11};
}
__device__ __forceinline__ H2 operator-(const H2& a, const H2& b) {
  return {a.v - b.v, a.g0 - b.g0, a.g1 - b.g1, a.h00 - b.h00, a.h01 - b.h01, a.h11 - b.h11};
}
__device__ __forceinline__ H2 operator-(const H2& a) { return {-a.v, -a.g0, -a.g1, -a.h00, -a.h01, -a.h11}; }
__device__ __forceinline__ H2 operator*(const H2& a, const H2& b) {
  return {a.v * b.v,
          a.g0 * b.v + a.v * b.g0,
          a.g1 * b.v + a.v * b.g1,
          a.h00 * b.v + 2.0 * a.g0 * b.g0 + a.v * b.h00,
          a.h01 * b.v + a.g0 * b.g1 + a.g1 * b.g0 + a.v * b.h01,
          a.h11 * b.v + 2.0 * a.g1 * b.g1 + a.v * b.h11};
}
__device__ __forceinline__ H2 hinv(const H2& a) {
  const double r = 1.0 / a.v;
  return hchain(a, r, -r * r, 2.0 * r * r * r);
}
__device__ __forceinline__ H2 operator/(const H2& a, const H2& b) { return a * hinv(b); }
__device__ __forceinline__ H2 operator+(const H2& a, double b) { return {a.v + b, a.g0, a.g1, a.h00, a.h01, a.h11}; }
__device__ __forceinline__ H2 operator+(double a, const H2& b) { return b + a; }
__device__ __forceinline__ H2 operator-(const H2& a, double b) { return {a.v - b, a.g0, a.g1, a.h00, a.h01, a.h11}; }
__device__ __forceinline__ H2 operator-(double a, const H2& b) { return (-b) + a; }
__device__ __forceinline__ H2 operator*(const H2& a, double b) {
  return {a.v * b, a.g0 * b, a.g1 * b, a.h00 * b, a.h01 * b, a.h11 * b};
}
__device__ __forceinline__ H2 operator*(double a, const H2& b) { return b * a; }
__device__ __forceinline__ H2 operator/(const H2& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ H2 operator/(double a, const H2& b) { return hinv(b) * a; }
__device__ __forceinline__ H2 dexp(const H2& a) { const double e = ::exp(a.v); return hchain(a, e, e, e); }
__device__ __forceinline__ H2 dlog(const H2& a) { const double r = 1.0 / a.v; return hchain(a, ::log(a.v), r, -r * r); }
__device__ __forceinline__ H2 dsqrt(const H2& a) {
  const double s = ::sqrt(a.v);
  return hchain(a, s, 0.5 / s, -0.25 / (s * a.v));
}
__device__ __forceinline__ H2 datan(const H2& a) {
  const double q = 1.0 / (1.0 + a.v * a.v);
  return hchain(a, ::atan(a.v), q, -2.0 * a.v * q * q);
}
__device__ __forceinline__ H2 dasinh(const H2& a) {
  const double q = 1.0 / ::sqrt(1.0 + a.v * a.v);
  return hchain(a, ::asinh(a.v), q, -a.v * q * q * q);
}


// Slater exchange (Dirac): e = -(3/4)(3/pi)^(1/3) rho^(4/3).
template <class T>
__device__ __forceinline__ T e_lda_x(const T& rho, const T& rho13) {
  const double cx = -0.7385587663820224;  // -(3/4)(3/pi)^(1/3)
  return cx * rho * rho13;
}

// Wigner-Seitz radius from rho^(1/3): rs = (3/(4 pi))^(1/3) / rho^(1/3).
template <class T>
__device__ __forceinline__ T rs_from_rho13(const T& rho13) { return 0.6203504908994000 / rho13; }

// VWN paramagnetic interpolation formula (Vosko-Wilk-Nusair 1980, eq. 4.4), Hartree.
template <class T>
__device__ __forceinline__ T vwn_eps(const T& rs, double A, double b, double c, double x0) {
  const T x = dsqrt(rs);
  const T X = x * x + b * x + c;
  const double Q = ::sqrt(4.0 * c - b * b);
  const double X0 = x0 * x0 + b * x0 + c;
  const T at = datan(Q / (2.0 * x + b));
  const T xm = x - x0;
  return A * (dlog(x * x / X) + (2.0 * b / Q) * at - (b * x0 / X0) * (dlog(xm * xm / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
}
template <class T>
__device__ __forceinline__ T vwn5_eps(const T& rs) { return vwn_eps(rs, 0.0310907, 3.72744, 12.9352, -0.10498); }
template <class T>
__device__ __forceinline__ T vwn_rpa_eps(const T& rs) { return vwn_eps(rs, 0.0310907, 13.0720, 42.7198, -0.409286); }

// PW92 paramagnetic correlation (Perdew-Wang 1992, eq. 10), libxc PW_MOD parameters.
template <class T>
__device__ __forceinline__ T pw92_eps(const T& rs) {
  const double A = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const T srs = dsqrt(rs);
  const T den = (2.0 * A) * (srs * (b1 + b3 * rs) + rs * (b2 + b4 * rs));
  return (-2.0 * A) * (1.0 + a1 * rs) * dlog(1.0 + 1.0 / den);
}

// PBE exchange: e = e_x^LDA * (1 + kappa - kappa / (1 + mu s^2 / kappa)).
template <class T>
__device__ __forceinline__ T e_pbe_x(const T& rho, const T& rho13, const T& sigma, double kappa, double mu) {
  // s^2 = sigma / (4 (3 pi^2)^(2/3) rho^(8/3))
  const T r83 = rho * rho * rho13 * rho13;
  const T s2 = sigma / (38.28312000250922 * r83);
  const T F = (1.0 + kappa) - kappa / (1.0 + (mu / kappa) * s2);
  return e_lda_x(rho, rho13) * F;
}

// PBE correlation with zeta = 0: e = rho (eps_c^PW92 + H(rs, t)).
template <class T>
__device__ __forceinline__ T e_pbe_c(const T& rho, const T& rho13, const T& sigma, double beta) {
  const double gamma = 0.031090690869654895;  // (1 - ln 2) / pi^2
  const T rs = rs_from_rho13(rho13);
  const T ec = pw92_eps(rs);
  // t^2 = sigma / (4 ks^2 rho^2), ks^2 = 4 kF / pi, kF = (3 pi^2 rho)^(1/3)
  const T kf = 3.0936677262801355 * rho13;
  const T t2 = sigma / ((16.0 / kPi) * kf * rho * rho);
  const T A = (beta / gamma) / (dexp(-ec / gamma) - 1.0);
  const T At2 = A * t2;
  const T H = gamma * dlog(1.0 + (beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}

// Becke 88 exchange, spin-unpolarised: two equal spin channels rho/2.
template <class T>
__device__ __forceinline__ T e_b88_x(const T& rho, const T& rho13, const T& sigma) {
  const double beta = 0.0042;
  // rho_s^(4/3) = (rho/2)^(4/3); x = |grad rho_s| / rho_s^(4/3)
  const T rs43 = 0.39685026299204984 * rho * rho13;  // 2^(-4/3) rho^(4/3)
  const T x = 0.5 * dsqrt(sigma) / rs43;
  const double cx = 0.9305257363491000;  // (3/2) (3/(4 pi))^(1/3)
  return -2.0 * rs43 * (cx + beta * x * x / (1.0 + 6.0 * beta * x * dasinh(x)));
}

// LYP correlation, closed shell (Miehlich-Savin-Stoll-Preuss 1989, eq. 2 at
// rho_a = rho_b = rho/2): e = -a rho/(1 + d rho^-1/3)
//   - a b exp(-c rho^-1/3)/(1 + d rho^-1/3) [CF rho - rho^-5/3 sigma (1/24 + 7 delta/72)].
template <class T>
__device__ __forceinline__ T e_lyp_c(const T& rho, const T& rho13, const T& sigma) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 2.8712340001881915;  // (3/10)(3 pi^2)^(2/3)
  const T rm13 = 1.0 / rho13;
  const T inv_den = 1.0 / (1.0 + d * rm13);
  const T delta = rm13 * (c + d * inv_den);
  const T ex = dexp(-c * rm13);
  const T rm53 = rm13 * rm13 / rho;
  return -a * rho * inv_den - (a * b) * ex * inv_den * (CF * rho - rm53 * sigma * (1.0 / 24.0 + (7.0 / 72.0) * delta));
}

// ---- r2SCAN (Furness, Kaplan, Ning, Perdew, Sun, JPCL 11, 8208 (2020) and
// erratum), spin-unpolarised: p = s^2, alpha_bar = (tau - tau_W)/(tau_unif + eta tau_W).
struct R2Poly { double c[8]; double c1, c2, d; };
template <class T>
__device__ __forceinline__ T r2_interp(const T& a, const R2Poly& f) {
  if (a.v <= 0.0) return dexp(-f.c1 * a / (1.0 - a));
  if (a.v <= 2.5) {
    T r = a * 0.0 + f.c[7];
#pragma unroll
    for (int i = 6; i >= 0; --i) r = r * a + f.c[i];
    return r;
  }
  return -f.d * dexp(f.c2 / (1.0 - a));
}
__device__ __forceinline__ double r2_slope(const R2Poly& f) {
  double s = 0;
#pragma unroll
  for (int i = 1; i < 8; ++i) s += i * f.c[i];
  return s;
}
__device__ __forceinline__ R2Poly r2_fx() {
  return {{1.0, -0.667, -0.4445555, -0.663086601049, 1.451297044490, -0.887998041597, 0.234528941479, -0.023185843322},
          0.667, 0.8, 1.24};
}
__device__ __forceinline__ R2Poly r2_fc() {
  return {{1.0, -0.64, -0.4352, -1.535685604549, 3.061560252175, -1.915710236206, 0.516884468372, -0.051848879792},
          0.64, 1.5, 0.7};
}
constexpr double kR2Eta = 0.001;
constexpr double kR2Dp2_4 = 0.361 * 0.361 * 0.361 * 0.361;

template <class T>
__device__ __forceinline__ void r2_descriptors(const T& rho, const T& rho13, const T& sigma, const T& tau, T* p,
                                               T* abar) {
  const double k2 = 9.570780000627305;  // (3 pi^2)^(2/3)
  const T r53 = rho * rho13 * rho13;
  *p = sigma / ((4.0 * k2) * rho * r53);
  const T tw = sigma / (8.0 * rho);
  const T tu = (0.3 * k2) * r53;
  *abar = (tau - tw) / (tu + kR2Eta * tw);
}

template <class T>
__device__ __forceinline__ T e_r2scan_x(const T& rho, const T& rho13, const T& sigma, const T& tau) {
  const double h0x = 1.174, k1 = 0.065, muak = 10.0 / 81.0, a1 = 4.9479;
  const R2Poly fxp = r2_fx();
  const double c2 = -r2_slope(fxp) * (1.0 - h0x);
  const double ceta = 20.0 / 27.0 + kR2Eta * 5.0 / 3.0;
  T p, ab;
  r2_descriptors(rho, rho13, sigma, tau, &p, &ab);
  const T x = (ceta * c2 * dexp(-(p * p) / kR2Dp2_4) + muak) * p;
  const T h1x = (1.0 + k1) - k1 / (1.0 + x / k1);
  const T fx = r2_interp(ab, fxp);
  const T gx = p.v > 0.0 ? 1.0 - dexp(-a1 / dpow(p, 0.25)) : p * 0.0 + 1.0;
  return e_lda_x(rho, rho13) * (h1x + fx * (h0x - h1x)) * gx;
}

__device__ __forceinline__ D3 e_r2scan_c(const D3& rho, const D3& rho13, const D3& sigma, const D3& tau) {
  const double gamma = 0.031090690869654895;
  const double b1c = 0.0285764, b2c = 0.0889, b3c = 0.125541, chi_inf = 0.128026;
  const R2Poly fcp = r2_fc();
  D3 p, ab;
  r2_descriptors(rho, rho13, sigma, tau, &p, &ab);
  const D3 rs = rs_from_rho13(rho13);
  const D3 srs = dsqrt(rs);
  const double A = 0.0310907, a1 = 0.21370, B1 = 7.5957, B2 = 3.5876, B3 = 1.6382, B4 = 0.49294;
  const D3 Q = (2.0 * A) * (srs * (B1 + B3 * rs) + rs * (B2 + B4 * rs));
  const D3 dQ = (2.0 * A) * ((0.5 * B1) / srs + B2 + (1.5 * B3) * srs + (2.0 * B4) * rs);
  const D3 L = dlog(1.0 + 1.0 / Q);
  const D3 e1 = (-2.0 * A) * (1.0 + a1 * rs) * L;
  const D3 de1 = (-2.0 * A * a1) * L + (2.0 * A) * (1.0 + a1 * rs) * dQ / (Q * Q + Q);
  const D3 den0 = 1.0 + b2c * srs + b3c * rs;
  const D3 e0 = -b1c / den0;
  const D3 de0 = b1c * ((0.5 * b2c) / srs + b3c) / (den0 * den0);
  const D3 beta = 0.066725 * (1.0 + 0.1 * rs) / (1.0 + 0.1778 * rs);
  const D3 w1 = dexp(-e1 / gamma) - 1.0;
  const D3 t2 = 1.5073033983379013 * p / rs;  // (3 pi^2 / 16)^(2/3) p / rs
  const D3 y = beta * t2 / (gamma * w1);
  const D3 dy = (r2_slope(fcp) / (27.0 * gamma)) / w1 * (20.0 * rs * (de0 - de1) - (45.0 * kR2Eta) * (e0 - e1)) * p *
                dexp(-(p * p) / kR2Dp2_4);
  const D3 g = dpow(1.0 + 4.0 * (y - dy), -0.25);
  const D3 ec1 = e1 + gamma * dlog(1.0 + w1 * (1.0 - g));
  const D3 w0 = dexp(-e0 / b1c) - 1.0;
  const D3 ginf = dpow(1.0 + (4.0 * chi_inf) * p, -0.25);
  const D3 ec0 = e0 + b1c * dlog(1.0 + w0 * (1.0 - ginf));
  const D3 fc = r2_interp(ab, fcp);
  return rho * (ec1 + fc * (ec0 - ec1));
}

// Functional codes handled by the kernels (libxc IDs).
enum : int {
  kLdaX = 1, kLdaCVwn = 7, kLdaCVwnRpa = 8, kGgaXPbe = 101, kGgaCPbe = 130, kGgaXB88 = 106,
  kGgaCLyp = 131, kGgaXPbeG = 321, kGgaCPbeG = 322, kHybB3lyp = 402, kMggaXR2scan = 497, kMggaCR2scan = 498
};

__host__ __device__ inline int code_family(int code) {
  switch (code) {
    case kLdaX: case kLdaCVwn: case kLdaCVwnRpa: return 0;
    case kGgaXPbe: case kGgaCPbe: case kGgaXB88: case kGgaCLyp: case kGgaXPbeG: case kGgaCPbeG: case kHybB3lyp: return 1;
    case kMggaXR2scan: case kMggaCR2scan: return 2;
    default: return -1;
  }
}
__host__ __device__ inline double code_exx(int code) { return code == kHybB3lyp ? 0.2 : 0.0; }
constexpr double kDensThreshold = 1e-15;

// LDA and GGA codes, on any number type (first or second order).
template <class T>
__device__ __forceinline__ T energy_density_lg(int code, const T& rho, const T& rho13, const T& sigma) {
  switch (code) {
    case kLdaX: return e_lda_x(rho, rho13);
    case kLdaCVwn: return rho * vwn5_eps(rs_from_rho13(rho13));
    case kLdaCVwnRpa: return rho * vwn_rpa_eps(rs_from_rho13(rho13));
    case kGgaXPbe: return e_pbe_x(rho, rho13, sigma, 0.804, 0.2195149727645171);
    case kGgaXPbeG: return e_pbe_x(rho, rho13, sigma, 0.804, 0.219510);
    case kGgaCPbe: return e_pbe_c(rho, rho13, sigma, 0.06672455060314922);
    case kGgaCPbeG: return e_pbe_c(rho, rho13, sigma, 0.06672632268006);
    case kGgaXB88: return e_b88_x(rho, rho13, sigma);
    case kGgaCLyp: return e_lyp_c(rho, rho13, sigma);
    case kHybB3lyp:
      return 0.08 * e_lda_x(rho, rho13) + 0.72 * e_b88_x(rho, rho13, sigma) +
             0.19 * (rho * vwn_rpa_eps(rs_from_rho13(rho13))) + 0.81 * e_lyp_c(rho, rho13, sigma);
    default: return 0.0 * rho;
  }
}

__device__ __forceinline__ D3 energy_density(int code, const D3& rho, const D3& rho13, const D3& sigma, const D3& tau) {
  switch (code) {
    case kMggaXR2scan: return e_r2scan_x(rho, rho13, sigma, tau);
    case kMggaCR2scan: return e_r2scan_c(rho, rho13, sigma, tau);
    default: return energy_density_lg(code, rho, rho13, sigma);
  }
}

// Sum over codes at one point; returns eps (per particle) and derivatives.
// Meta-GGA codes see tau floored at 1e-20 and sigma capped at 8 rho tau.
__device__ __forceinline__ void eval_codes(const int* codes, int ncodes, double rho, double sigma, double tau,
                                           double* eps, double* vrho, double* vsigma, double* vtau) {
  double e = 0, vr = 0, vs = 0, vt = 0;
  if (rho >= kDensThreshold) {
    const double sfloor = 1.0000000000000000e-40;  // (thr^(4/3))^2 for thr = 1e-15
    const double s = sigma > sfloor ? sigma : sfloor;
    const D3 r = {rho, 1.0, 0.0, 0.0};
    const double r13v = ::cbrt(rho);
    const D3 rho13 = {r13v, r13v / (3.0 * rho), 0.0, 0.0};
    for (int k = 0; k < ncodes; ++k) {
      const int code = codes[k];
      double sk = s, tk = tau;
      if (code_family(code) == 2) {
        tk = tau > 1e-20 ? tau : 1e-20;
        if (sk > 8.0 * rho * tk) sk = 8.0 * rho * tk;
      }
      const D3 ed = energy_density(code, r, rho13, D3{sk, 0.0, 1.0, 0.0}, D3{tk, 0.0, 0.0, 1.0});
      e += ed.v;
      vr += ed.d0;
      vs += ed.d1;
      vt += ed.d2;
    }
    e /= rho;
  }
  *eps = e;
  *vrho = vr;
  *vsigma = vs;
  *vtau = vt;
}

// Second derivatives of the summed LDA/GGA energy density at one point
// (libxc's xc_*_fxc, src/ExchangeCorrelation.cpp:101-104,126-133): v2rho2,
// v2rhosigma, v2sigma2, and v_sigma with them. Same threshold and sigma floor
// as eval_codes; meta-GGA codes are not allowed on this path.
__device__ __forceinline__ void eval_codes_f(const int* codes, int ncodes, double rho, double sigma, double* vsigma,
                                             double* v2rho2, double* v2rhosigma, double* v2sigma2) {
  double vs = 0, frr = 0, frs = 0, fss = 0;
  if (rho >= kDensThreshold) {
    const double sfloor = 1.0000000000000000e-40;
    const double sg = sigma > sfloor ? sigma : sfloor;
    const H2 r = {rho, 1.0, 0.0, 0.0, 0.0, 0.0};
    const double r13 = ::cbrt(rho);
    const H2 rho13 = hchain(r, r13, r13 / (3.0 * rho), -2.0 * r13 / (9.0 * rho * rho));
    const H2 s = {sg, 0.0, 1.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < ncodes; ++k) {
      const H2 ed = energy_density_lg(codes[k], r, rho13, s);
      vs += ed.g1;
      frr += ed.h00;
      frs += ed.h01;
      fss += ed.h11;
    }
  }
  *vsigma = vs;
  *v2rho2 = frr;
  *v2rhosigma = frs;
  *v2sigma2 = fss;
}

// ---- spin-polarised LDA/GGA (libxc XC_POLARIZED conventions) --------------
// Inputs rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb; outputs eps per particle
// of the total density and the five first derivatives of e = rho eps. The
// number type P5 carries those five partials.
struct P5 {
  double v, d[5];
};
__device__ __forceinline__ P5 pconst(double x) { return {x, {0.0, 0.0, 0.0, 0.0, 0.0}}; }
__device__ __forceinline__ P5 pvar(double x, int k) {
  P5 r = pconst(x);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ P5 pchain(const P5& a, double f, double df) {
  P5 r;
  r.v = f;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = df * a.d[k];
  return r;
}
__device__ __forceinline__ P5 operator+(const P5& a, const P5& b) {
  P5 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ P5 operator-(const P5& a) { return pchain(a, -a.v, -1.0); }
__device__ __forceinline__ P5 operator-(const P5& a, const P5& b) { return a + (-b); }
__device__ __forceinline__ P5 operator*(const P5& a, const P5& b) {
  P5 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ P5 operator/(const P5& a, const P5& b) {
  const double ib = 1.0 / b.v, q = a.v * ib;
  P5 r;
  r.v = q;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - q * b.d[k]) * ib;
  return r;
}
__device__ __forceinline__ P5 operator+(const P5& a, double b) { P5 r = a; r.v += b; return r; }
__device__ __forceinline__ P5 operator+(double a, const P5& b) { return b + a; }
__device__ __forceinline__ P5 operator-(const P5& a, double b) { return a + (-b); }
__device__ __forceinline__ P5 operator-(double a, const P5& b) { return (-b) + a; }
__device__ __forceinline__ P5 operator*(const P5& a, double b) { return pchain(a, a.v * b, b); }
__device__ __forceinline__ P5 operator*(double a, const P5& b) { return b * a; }
__device__ __forceinline__ P5 operator/(const P5& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ P5 operator/(double a, const P5& b) { return pconst(a) / b; }
__device__ __forceinline__ P5 dexp(const P5& a) { const double e = ::exp(a.v); return pchain(a, e, e); }
__device__ __forceinline__ P5 dlog(const P5& a) { return pchain(a, ::log(a.v), 1.0 / a.v); }
__device__ __forceinline__ P5 dsqrt(const P5& a) { const double s = ::sqrt(a.v); return pchain(a, s, 0.5 / s); }
__device__ __forceinline__ P5 datan(const P5& a) { return pchain(a, ::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
__device__ __forceinline__ P5 dasinh(const P5& a) { return pchain(a, ::asinh(a.v), 1.0 / ::sqrt(1.0 + a.v * a.v)); }
__device__ __forceinline__ P5 dcbrt(const P5& a) { const double c = ::cbrt(a.v); return pchain(a, c, c / (3.0 * a.v)); }
// x^p for x > 0 (a power of 1 +- zeta)
__device__ __forceinline__ P5 dpowp(const P5& a, double p) {
  const double f = ::pow(a.v, p);
  return pchain(a, f, p * f / a.v);
}

// Second-order forward mode in the five polarised variables (rho_a, rho_b,
// sigma_aa, sigma_ab, sigma_bb): value, gradient and the Hessian's upper
// triangle (row-major, h[hix(i, j)] for i <= j), for the polarised f_xc of the
// response path. The polarised energy densities below are templates on the
// number type, so f_xc is the exact second derivative of the same code.
struct H5 {
  double v, g[5], h[15];
};
__device__ __forceinline__ constexpr int hix(int i, int j) { return i * 5 - i * (i - 1) / 2 + (j - i); }
__device__ __forceinline__ H5 h5const(double x) {
  H5 r;
  r.v = x;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.g[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 15; ++k) r.h[k] = 0.0;
  return r;
}
__device__ __forceinline__ H5 h5var(double x, int k) {
  H5 r = h5const(x);
  r.g[k] = 1.0;
  return r;
}
__device__ __forceinline__ H5 h5chain(const H5& a, double f, double f1, double f2) {
  H5 r;
  r.v = f;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.g[i] = f1 * a.g[i];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = i; j < 5; ++j) r.h[hix(i, j)] = f1 * a.h[hix(i, j)] + f2 * a.g[i] * a.g[j];
  return r;
}
__device__ __forceinline__ H5 operator+(const H5& a, const H5& b) {
  H5 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.g[k] = a.g[k] + b.g[k];
#pragma unroll
  for (int k = 0; k < 15; ++k) r.h[k] = a.h[k] + b.h[k];
  return r;
}
__device__ __forceinline__ H5 operator-(const H5& a) { return h5chain(a, -a.v, -1.0, 0.0); }
__device__ __forceinline__ H5 operator-(const H5& a, const H5& b) { return a + (-b); }
__device__ __forceinline__ H5 operator*(const H5& a, const H5& b) {
  H5 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = i; j < 5; ++j)
      r.h[hix(i, j)] = a.h[hix(i, j)] * b.v + a.g[i] * b.g[j] + a.g[j] * b.g[i] + a.v * b.h[hix(i, j)];
  return r;
}
__device__ __forceinline__ H5 h5inv(const H5& a) {
  const double q = 1.0 / a.v;
  return h5chain(a, q, -q * q, 2.0 * q * q * q);
}
__device__ __forceinline__ H5 operator/(const H5& a, const H5& b) { return a * h5inv(b); }
__device__ __forceinline__ H5 operator+(const H5& a, double b) { H5 r = a; r.v += b; return r; }
__device__ __forceinline__ H5 operator+(double a, const H5& b) { return b + a; }
__device__ __forceinline__ H5 operator-(const H5& a, double b) { return a + (-b); }
__device__ __forceinline__ H5 operator-(double a, const H5& b) { return (-b) + a; }
__device__ __forceinline__ H5 operator*(const H5& a, double b) { return h5chain(a, a.v * b, b, 0.0); }
__device__ __forceinline__ H5 operator*(double a, const H5& b) { return b * a; }
__device__ __forceinline__ H5 operator/(const H5& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ H5 operator/(double a, const H5& b) { return h5inv(b) * a; }
__device__ __forceinline__ H5 dexp(const H5& a) { const double e = ::exp(a.v); return h5chain(a, e, e, e); }
__device__ __forceinline__ H5 dlog(const H5& a) { const double q = 1.0 / a.v; return h5chain(a, ::log(a.v), q, -q * q); }
__device__ __forceinline__ H5 dsqrt(const H5& a) {
  const double q = ::sqrt(a.v);
  return h5chain(a, q, 0.5 / q, -0.25 / (q * a.v));
}
__device__ __forceinline__ H5 datan(const H5& a) {
  const double q = 1.0 / (1.0 + a.v * a.v);
  return h5chain(a, ::atan(a.v), q, -2.0 * a.v * q * q);
}
__device__ __forceinline__ H5 dasinh(const H5& a) {
  const double q = 1.0 / ::sqrt(1.0 + a.v * a.v);
  return h5chain(a, ::asinh(a.v), q, -a.v * q * q * q);
}
__device__ __forceinline__ H5 dcbrt(const H5& a) {
  const double c = ::cbrt(a.v), q = 1.0 / a.v;
  return h5chain(a, c, c * q / 3.0, -2.0 * c * q * q / 9.0);
}
__device__ __forceinline__ H5 dpowp(const H5& a, double p) {
  const double f = ::pow(a.v, p), q = 1.0 / a.v;
  return h5chain(a, f, p * f * q, p * (p - 1.0) * f * q * q);
}

// Constants in either polarised number type.
template <class T> __device__ __forceinline__ T tconst(double x);
template <> __device__ __forceinline__ P5 tconst<P5>(double x) { return pconst(x); }
template <> __device__ __forceinline__ H5 tconst<H5>(double x) { return h5const(x); }

// f(zeta) of the spin interpolation and its curvature at 0, 4 / (9 (2^(1/3) - 1)).
constexpr double kFpp0 = 1.709920934161365617563962776245;
template <class T>
__device__ __forceinline__ T f_zeta(const T& opz, const T& omz) {  // opz = 1 + zeta, omz = 1 - zeta
  return (opz * dcbrt(opz) + omz * dcbrt(omz) - 2.0) / (2.5198420997897464 - 2.0);  // 2^(4/3) - 2
}

// VWN with spin (Vosko-Wilk-Nusair 1980, eq. 4.4 with their spin-stiffness
// interpolation, as libxc's lda_c_vwn): eps = eP + a f (1 - z^4)/f''(0) + (eF - eP) f z^4.
template <class T>
__device__ __forceinline__ T vwn5_eps_pol(const T& rs, const T& z, const T& opz, const T& omz) {
  const T eP = vwn_eps(rs, 0.0310907, 3.72744, 12.9352, -0.10498);
  const T eF = vwn_eps(rs, 0.01554535, 7.06042, 18.0578, -0.32500);
  const T ea = vwn_eps(rs, -1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584);
  const T f = f_zeta(opz, omz);
  const T z4 = z * z * z * z;
  return eP + ea * f * (1.0 - z4) / kFpp0 + (eF - eP) * f * z4;
}
// VWN RPA fit with the plain f(zeta) interpolation (libxc lda_c_vwn_rpa).
template <class T>
__device__ __forceinline__ T vwn_rpa_eps_pol(const T& rs, const T& opz, const T& omz) {
  const T eP = vwn_eps(rs, 0.0310907, 13.0720, 42.7198, -0.409286);
  const T eF = vwn_eps(rs, 0.01554535, 20.1231, 101.578, -0.743294);
  return eP + (eF - eP) * f_zeta(opz, omz);
}

// PW92 G(rs) of Perdew-Wang 1992 eq. 10.
template <class T>
__device__ __forceinline__ T pw92_g(const T& rs, double A, double a1, double b1, double b2, double b3, double b4) {
  const T srs = dsqrt(rs);
  const T den = (2.0 * A) * (srs * (b1 + b3 * rs) + rs * (b2 + b4 * rs));
  return (-2.0 * A) * (1.0 + a1 * rs) * dlog(1.0 + 1.0 / den);
}
// dG/drs of the same G.
template <class T>
__device__ __forceinline__ T pw92_g_drs(const T& rs, double A, double a1, double b1, double b2, double b3, double b4) {
  const T srs = dsqrt(rs);
  const T den = (2.0 * A) * (srs * (b1 + b3 * rs) + rs * (b2 + b4 * rs));
  const T dden = (2.0 * A) * ((0.5 * b1) / srs + b2 + (1.5 * b3) * srs + (2.0 * b4) * rs);
  return (-2.0 * A * a1) * dlog(1.0 + 1.0 / den) + (2.0 * A) * (1.0 + a1 * rs) * dden / (den * den + den);
}
// PW92 with spin (eq. 8), libxc PW_MOD parameters; -alpha_c = G(rs; A_alpha, ...).
template <class T>
__device__ __forceinline__ T pw92_eps_pol(const T& rs, const T& z, const T& opz, const T& omz) {
  const T e0 = pw92_g(rs, 0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const T e1 = pw92_g(rs, 0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const T ma = pw92_g(rs, 0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const T f = f_zeta(opz, omz);
  const T z4 = z * z * z * z;
  return e0 - ma * f * (1.0 - z4) / kFpp0 + (e1 - e0) * f * z4;
}

// PBE correlation with spin: phi = ((1+z)^(2/3) + (1-z)^(2/3))/2,
// t^2 = |grad rho|^2 / (4 phi^2 ks^2 rho^2), H = gamma phi^3 ln(1 + beta/gamma t^2 (1+At^2)/(1+At^2+A^2t^4)).
template <class T>
__device__ __forceinline__ T e_pbe_c_pol(const T& rho, const T& rho13, const T& gg, const T& z, const T& opz,
                                         const T& omz, double beta) {
  const double gamma = 0.031090690869654895;
  const T rs = rs_from_rho13(rho13);
  const T ec = pw92_eps_pol(rs, z, opz, omz);
  const T c1 = dcbrt(opz), c2 = dcbrt(omz);
  const T phi = 0.5 * (c1 * c1 + c2 * c2);
  const T phi3 = phi * phi * phi;
  const T kf = 3.0936677262801355 * rho13;
  const T t2 = gg / ((16.0 / kPi) * kf * rho * rho * phi * phi);
  const T A = (beta / gamma) / (dexp(-ec / (gamma * phi3)) - 1.0);
  const T At2 = A * t2;
  const T H = gamma * phi3 * dlog(1.0 + (beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}

// LYP correlation, general spin form (Miehlich, Savin, Stoll, Preuss 1989, eq. 2).
template <class T>
__device__ __forceinline__ T e_lyp_c_pol(const T& ra, const T& rb, const T& rho, const T& rho13, const T& gaa,
                                         const T& gab, const T& gbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 2.8712340001881915;
  const T rm13 = 1.0 / rho13;
  const T inv_den = 1.0 / (1.0 + d * rm13);
  const T delta = rm13 * (c + d * inv_den);
  const T omega = dexp(-c * rm13) * inv_den * (rm13 * rm13 / (rho * rho * rho));  // rho^(-11/3)
  const T gg = gaa + 2.0 * gab + gbb;
  auto p83 = [](const T& x) {  // x^(8/3), 0 for an empty channel
    if (!(x.v > 0.0)) return tconst<T>(0.0);
    const T c13 = dcbrt(x);
    return x * x * c13 * c13;
  };
  const T ra83 = p83(ra), rb83 = p83(rb);
  const T inner = ra * rb * (12.699208415745595 * CF * (ra83 + rb83) + (47.0 / 18.0 - (7.0 / 18.0) * delta) * gg -
                              (2.5 - delta / 18.0) * (gaa + gbb) - (delta - 11.0) / 9.0 * (ra / rho * gaa + rb / rho * gbb)) -
                   (2.0 / 3.0) * rho * rho * gg + ((2.0 / 3.0) * rho * rho - ra * ra) * gbb +
                   ((2.0 / 3.0) * rho * rho - rb * rb) * gaa;
  return -4.0 * a * inv_den * ra * rb / rho - (a * b) * omega * inner;
}

// Exchange by spin scaling: E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2.
template <class T>
__device__ __forceinline__ T e_x_channel(int code, const T& r2, const T& s4) {
  const T r13 = dcbrt(r2);
  switch (code) {
    case kLdaX: return 0.5 * e_lda_x(r2, r13);
    case kGgaXPbe: return 0.5 * e_pbe_x(r2, r13, s4, 0.804, 0.2195149727645171);
    case kGgaXPbeG: return 0.5 * e_pbe_x(r2, r13, s4, 0.804, 0.219510);
    case kGgaXB88: return 0.5 * e_b88_x(r2, r13, s4);
    default: return tconst<T>(0.0);
  }
}

// Energy density of one LDA/GGA code with spin. ra_x / rb_x: the exchange
// channels (zero below the threshold, then skipped).
template <class T>
__device__ __forceinline__ T energy_density_pol(int code, const T& ra, const T& rb, const T& saa, const T& sab,
                                                const T& sbb, bool a_on, bool b_on) {
  auto ex = [&](int xc) {
    T e = tconst<T>(0.0);
    if (a_on) e = e + e_x_channel(xc, 2.0 * ra, 4.0 * saa);
    if (b_on) e = e + e_x_channel(xc, 2.0 * rb, 4.0 * sbb);
    return e;
  };
  const T rho = ra + rb;
  const T rho13 = dcbrt(rho);
  const T z = (ra - rb) / rho;
  T opz = 1.0 + z, omz = 1.0 - z;
  if (opz.v < 1e-15) opz = tconst<T>(1e-15);  // zeta threshold: fully polarised limit
  if (omz.v < 1e-15) omz = tconst<T>(1e-15);
  switch (code) {
    case kLdaX: case kGgaXPbe: case kGgaXPbeG: case kGgaXB88: return ex(code);
    case kLdaCVwn: return rho * vwn5_eps_pol(rs_from_rho13(rho13), z, opz, omz);
    case kLdaCVwnRpa: return rho * vwn_rpa_eps_pol(rs_from_rho13(rho13), opz, omz);
    case kGgaCPbe: return e_pbe_c_pol(rho, rho13, saa + 2.0 * sab + sbb, z, opz, omz, 0.06672455060314922);
    case kGgaCPbeG: return e_pbe_c_pol(rho, rho13, saa + 2.0 * sab + sbb, z, opz, omz, 0.06672632268006);
    case kGgaCLyp: return e_lyp_c_pol(ra, rb, rho, rho13, saa, sab, sbb);
    case kHybB3lyp:
      return 0.08 * ex(kLdaX) + 0.72 * ex(kGgaXB88) + 0.19 * (rho * vwn_rpa_eps_pol(rs_from_rho13(rho13), opz, omz)) +
             0.81 * e_lyp_c_pol(ra, rb, rho, rho13, saa, sab, sbb);
    default: return tconst<T>(0.0);
  }
}

// Sum over codes at one point, spin-polarised LDA/GGA: eps (per particle of
// rho_a + rho_b), v_rho (a, b), v_sigma (aa, ab, bb). A code contributes
// nothing where rho_a + rho_b is below the threshold; an exchange channel
// contributes nothing where its own density is. sigma_aa and sigma_bb are
// floored at 1e-40, |sigma_ab| capped at sqrt(sigma_aa sigma_bb).
__device__ __forceinline__ void eval_codes_pol(const int* codes, int ncodes, double ra, double rb, double saa,
                                               double sab, double sbb, double* eps, double* vr, double* vs) {
  double e = 0, d[5] = {0, 0, 0, 0, 0};
  ra = ra > 0 ? ra : 0.0;  // negative or NaN channel: empty
  rb = rb > 0 ? rb : 0.0;
  const double rho = ra + rb;
  if (rho >= kDensThreshold) {
    const double sfl = 1e-40;
    saa = saa > sfl ? saa : sfl;
    sbb = sbb > sfl ? sbb : sfl;
    const double cap = ::sqrt(saa * sbb);
    sab = sab > cap ? cap : (sab < -cap ? -cap : sab);
    const P5 pa = pvar(ra, 0), pb = pvar(rb, 1), paa = pvar(saa, 2), pab = pvar(sab, 3), pbb = pvar(sbb, 4);
    const bool a_on = ra >= kDensThreshold, b_on = rb >= kDensThreshold;
    for (int k = 0; k < ncodes; ++k) {
      const P5 ed = energy_density_pol(codes[k], pa, pb, paa, pab, pbb, a_on, b_on);
      e += ed.v;
#pragma unroll
      for (int j = 0; j < 5; ++j) d[j] += ed.d[j];
    }
    e /= rho;
  }
  *eps = e;
  vr[0] = d[0];
  vr[1] = d[1];
  vs[0] = d[2];
  vs[1] = d[3];
  vs[2] = d[4];
}

// Second derivatives of the summed polarised LDA/GGA energy density at one
// point (libxc xc_{lda,gga}_fxc with XC_POLARIZED, as Evaluate("f") copies them,
// src/ExchangeCorrelation.cpp:126-133,232-245), with the same thresholds,
// floors and caps as eval_codes_pol. f[18]: vsigma (aa, ab, bb), v2rho2 (a a,
// a b, b b), v2rhosigma (a aa, a ab, a bb, b aa, b ab, b bb), v2sigma2 (aa aa,
// aa ab, aa bb, ab ab, ab bb, bb bb).
__device__ __forceinline__ void eval_codes_fpol(const int* codes, int ncodes, double ra, double rb, double saa,
                                                double sab, double sbb, double* f) {
#pragma unroll
  for (int k = 0; k < 18; ++k) f[k] = 0.0;
  ra = ra > 0 ? ra : 0.0;
  rb = rb > 0 ? rb : 0.0;
  if (!(ra + rb >= kDensThreshold)) return;
  const double sfl = 1e-40;
  saa = saa > sfl ? saa : sfl;
  sbb = sbb > sfl ? sbb : sfl;
  const double cap = ::sqrt(saa * sbb);
  sab = sab > cap ? cap : (sab < -cap ? -cap : sab);
  const H5 pa = h5var(ra, 0), pb = h5var(rb, 1), paa = h5var(saa, 2), pab = h5var(sab, 3), pbb = h5var(sbb, 4);
  const bool a_on = ra >= kDensThreshold, b_on = rb >= kDensThreshold;
  for (int k = 0; k < ncodes; ++k) {
    const H5 e = energy_density_pol(codes[k], pa, pb, paa, pab, pbb, a_on, b_on);
#pragma unroll
    for (int x = 0; x < 3; ++x) f[x] += e.g[2 + x];
    f[3] += e.h[hix(0, 0)];
    f[4] += e.h[hix(0, 1)];
    f[5] += e.h[hix(1, 1)];
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
      for (int x = 0; x < 3; ++x) f[6 + 3 * w + x] += e.h[hix(w, 2 + x)];
    int q = 12;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = x; y < 3; ++y) f[q++] += e.h[hix(2 + x, 2 + y)];
  }
}

// ---- spin-polarised r2SCAN (libxc 497 / 498, XC_POLARIZED conventions) -----
// Variables rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb, tau_a, tau_b; the
// number type P7 carries the seven partials. Exchange by spin scaling,
// E_x = (E_x[2 ra, 4 saa, 2 ta] + E_x[2 rb, 4 sbb, 2 tb]) / 2. Correlation with
// SCAN's spin dependence (Sun, Ruzsinszky, Perdew, PRL 115, 036402 (2015)):
// phi, d_s = ((1+z)^(5/3) + (1-z)^(5/3))/2 in tau_unif, G_c(z) =
// (1 - 2.3631 (d_x - 1))(1 - z^12) on eps_c^0, the polarised PW92 in eps_c^1,
// and r2SCAN's Delta y with eps_LSDA0 = eps_LDA0 G_c and 1 / (d_s phi^3).
struct P7 {
  double v, d[7];
};
__device__ __forceinline__ P7 qconst(double x) { return {x, {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}}; }
__device__ __forceinline__ P7 qvar(double x, int k) {
  P7 r = qconst(x);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ P7 qchain(const P7& a, double f, double df) {
  P7 r;
  r.v = f;
#pragma unroll
  for (int k = 0; k < 7; ++k) r.d[k] = df * a.d[k];
  return r;
}
__device__ __forceinline__ P7 operator+(const P7& a, const P7& b) {
  P7 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 7; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ P7 operator-(const P7& a) { return qchain(a, -a.v, -1.0); }
__device__ __forceinline__ P7 operator-(const P7& a, const P7& b) { return a + (-b); }
__device__ __forceinline__ P7 operator*(const P7& a, const P7& b) {
  P7 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 7; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ P7 operator/(const P7& a, const P7& b) {
  const double ib = 1.0 / b.v, q = a.v * ib;
  P7 r;
  r.v = q;
#pragma unroll
  for (int k = 0; k < 7; ++k) r.d[k] = (a.d[k] - q * b.d[k]) * ib;
  return r;
}
__device__ __forceinline__ P7 operator+(const P7& a, double b) { P7 r = a; r.v += b; return r; }
__device__ __forceinline__ P7 operator+(double a, const P7& b) { return b + a; }
__device__ __forceinline__ P7 operator-(const P7& a, double b) { return a + (-b); }
__device__ __forceinline__ P7 operator-(double a, const P7& b) { return (-b) + a; }
__device__ __forceinline__ P7 operator*(const P7& a, double b) { return qchain(a, a.v * b, b); }
__device__ __forceinline__ P7 operator*(double a, const P7& b) { return b * a; }
__device__ __forceinline__ P7 operator/(const P7& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ P7 operator/(double a, const P7& b) { return qconst(a) / b; }
__device__ __forceinline__ P7 dexp(const P7& a) { const double e = ::exp(a.v); return qchain(a, e, e); }
__device__ __forceinline__ P7 dlog(const P7& a) { return qchain(a, ::log(a.v), 1.0 / a.v); }
__device__ __forceinline__ P7 dsqrt(const P7& a) { const double s = ::sqrt(a.v); return qchain(a, s, 0.5 / s); }
__device__ __forceinline__ P7 dcbrt(const P7& a) { const double c = ::cbrt(a.v); return qchain(a, c, c / (3.0 * a.v)); }
__device__ __forceinline__ P7 dpow(const P7& a, double p) { const double f = ::pow(a.v, p); return qchain(a, f, p * f / a.v); }

__device__ __forceinline__ P7 e_r2scan_c_pol(const P7& rho, const P7& rho13, const P7& sigma, const P7& tau,
                                             const P7& z, const P7& opz, const P7& omz) {
  const double gamma = 0.031090690869654895;
  const double b1c = 0.0285764, b2c = 0.0889, b3c = 0.125541, chi_inf = 0.128026;
  const double k2 = 9.570780000627305;  // (3 pi^2)^(2/3)
  const R2Poly fcp = r2_fc();
  const P7 c1 = dcbrt(opz), c2 = dcbrt(omz);
  const P7 phi = 0.5 * (c1 * c1 + c2 * c2);
  const P7 ds = 0.5 * (opz * c1 * c1 + omz * c2 * c2);
  const P7 dx = 0.5 * (opz * c1 + omz * c2);
  const P7 z2 = z * z, z4 = z2 * z2;
  const P7 Gc = (1.0 - 2.3631 * (dx - 1.0)) * (1.0 - z4 * z4 * z4);
  const P7 r53 = rho * rho13 * rho13;
  const P7 p = sigma / ((4.0 * k2) * rho * r53);
  const P7 tw = sigma / (8.0 * rho);
  const P7 ab = (tau - tw) / ((0.3 * k2) * r53 * ds + kR2Eta * tw);
  const P7 rs = rs_from_rho13(rho13);
  const P7 srs = dsqrt(rs);
  // polarised PW92 and its rs-derivative at fixed zeta
  const double q0[6] = {0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
  const double q1[6] = {0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
  const double qa[6] = {0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};
  const P7 g0 = pw92_g(rs, q0[0], q0[1], q0[2], q0[3], q0[4], q0[5]);
  const P7 g1 = pw92_g(rs, q1[0], q1[1], q1[2], q1[3], q1[4], q1[5]);
  const P7 ga = pw92_g(rs, qa[0], qa[1], qa[2], qa[3], qa[4], qa[5]);
  const P7 dg0 = pw92_g_drs(rs, q0[0], q0[1], q0[2], q0[3], q0[4], q0[5]);
  const P7 dg1 = pw92_g_drs(rs, q1[0], q1[1], q1[2], q1[3], q1[4], q1[5]);
  const P7 dga = pw92_g_drs(rs, qa[0], qa[1], qa[2], qa[3], qa[4], qa[5]);
  const P7 f = f_zeta(opz, omz);
  const P7 fs = f * (1.0 - z4) / kFpp0, fz = f * z4;
  const P7 e1 = g0 - ga * fs + (g1 - g0) * fz;
  const P7 de1 = dg0 - dga * fs + (dg1 - dg0) * fz;
  const P7 den0 = 1.0 + b2c * srs + b3c * rs;
  const P7 e0 = -b1c / den0;
  const P7 de0 = b1c * ((0.5 * b2c) / srs + b3c) / (den0 * den0);
  const P7 phi3 = phi * phi * phi;
  const P7 beta = 0.066725 * (1.0 + 0.1 * rs) / (1.0 + 0.1778 * rs);
  const P7 w1 = dexp(-e1 / (gamma * phi3)) - 1.0;
  const P7 t2 = 1.5073033983379013 * p / (phi * phi * rs);  // (3 pi^2 / 16)^(2/3) p / (phi^2 rs)
  const P7 y = beta * t2 / (gamma * w1);
  const P7 dy = (r2_slope(fcp) / (27.0 * gamma)) / (ds * phi3 * w1) *
                (20.0 * rs * (de0 * Gc - de1) - (45.0 * kR2Eta) * (e0 * Gc - e1)) * p * dexp(-(p * p) / kR2Dp2_4);
  const P7 g = dpow(1.0 + 4.0 * (y - dy), -0.25);
  const P7 ec1 = e1 + gamma * phi3 * dlog(1.0 + w1 * (1.0 - g));
  const P7 w0 = dexp(-e0 / b1c) - 1.0;
  const P7 ginf = dpow(1.0 + (4.0 * chi_inf) * p, -0.25);
  const P7 ec0 = (e0 + b1c * dlog(1.0 + w0 * (1.0 - ginf))) * Gc;
  const P7 fc = r2_interp(ab, fcp);
  return rho * (ec1 + fc * (ec0 - ec1));
}

// Sum over meta-GGA codes at one point, spin-polarised: eps (per particle of
// rho_a + rho_b), v_rho (a, b), v_sigma (aa, ab, bb), v_tau (a, b). Density
// threshold on the total; an exchange channel below it contributes nothing;
// per spin tau_s >= 1e-20, 1e-40 <= sigma_ss <= 8 rho_s tau_s, and
// |sigma_ab| <= sqrt(sigma_aa sigma_bb).
__device__ __forceinline__ void eval_codes_mpol(const int* codes, int ncodes, double ra, double rb, double saa,
                                                double sab, double sbb, double ta, double tb, double* eps, double* vr,
                                                double* vs, double* vt) {
  double e = 0, d[7] = {0, 0, 0, 0, 0, 0, 0};
  ra = ra > 0 ? ra : 0.0;
  rb = rb > 0 ? rb : 0.0;
  const double rho = ra + rb;
  if (rho >= kDensThreshold) {
    ta = ta > 1e-20 ? ta : 1e-20;
    tb = tb > 1e-20 ? tb : 1e-20;
    const double ca = 8.0 * ra * ta > 1e-40 ? 8.0 * ra * ta : 1e-40, cb = 8.0 * rb * tb > 1e-40 ? 8.0 * rb * tb : 1e-40;
    saa = saa > 1e-40 ? (saa < ca ? saa : ca) : 1e-40;
    sbb = sbb > 1e-40 ? (sbb < cb ? sbb : cb) : 1e-40;
    const double cap = ::sqrt(saa * sbb);
    sab = sab > cap ? cap : (sab < -cap ? -cap : sab);
    const P7 pa = qvar(ra, 0), pb = qvar(rb, 1), paa = qvar(saa, 2), pab = qvar(sab, 3), pbb = qvar(sbb, 4);
    const P7 pta = qvar(ta, 5), ptb = qvar(tb, 6);
    const bool a_on = ra >= kDensThreshold, b_on = rb >= kDensThreshold;
    const P7 prho = pa + pb;
    const P7 r13 = dcbrt(prho);
    P7 z = (pa - pb) / prho;
    if (z.v > 1.0 - 1e-15) z = qconst(1.0 - 1e-15);  // zeta threshold: fully polarised limit
    if (z.v < -1.0 + 1e-15) z = qconst(-1.0 + 1e-15);
    const P7 opz = 1.0 + z, omz = 1.0 - z;
    for (int k = 0; k < ncodes; ++k) {
      P7 ed = qconst(0.0);
      if (codes[k] == kMggaXR2scan) {
        if (a_on) {
          const P7 r2 = 2.0 * pa;
          ed = ed + 0.5 * e_r2scan_x(r2, dcbrt(r2), 4.0 * paa, 2.0 * pta);
        }
        if (b_on) {
          const P7 r2 = 2.0 * pb;
          ed = ed + 0.5 * e_r2scan_x(r2, dcbrt(r2), 4.0 * pbb, 2.0 * ptb);
        }
      } else if (codes[k] == kMggaCR2scan) {
        ed = e_r2scan_c_pol(prho, r13, paa + 2.0 * pab + pbb, pta + ptb, z, opz, omz);
      }
      e += ed.v;
#pragma unroll
      for (int j = 0; j < 7; ++j) d[j] += ed.d[j];
    }
    e /= rho;
  }
  *eps = e;
  vr[0] = d[0];
  vr[1] = d[1];
  vs[0] = d[2];
  vs[1] = d[3];
  vs[2] = d[4];
  vt[0] = d[5];
  vt[1] = d[6];
}

}  // namespace chx