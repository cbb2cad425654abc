// FP64 exchange-correlation functionals for the CPU oracle (test infrastructure).
//
// The reference evaluates these through libxc (src/ExchangeCorrelation.cpp:86-163,
// per-code sum at :202-217). libxc is not present, so each functional is written
// here from its published form, spin-unpolarised (the reference's closed-shell
// SCF sets Spin=1, src/HartreeFockKohnSham/Restricted/Gateway.cpp:62).
// Derivatives come from forward-mode dual numbers, so v_rho, v_sigma and v_tau
// are exact derivatives of the energy density rho*eps that is coded below.
//
// libxc conventions kept: eps is per particle; sigma = |grad rho|^2;
// tau = 1/2 sum_i |grad psi_i|^2; a point with rho below the density threshold
// contributes nothing; sigma is floored at thr^(8/3) and, for meta-GGAs, capped
// at 8 rho tau.
#pragma once
#include <algorithm>
#include <cmath>

namespace oxf {

template <int N> struct Dual {
  double v;
  double d[N];
  Dual() : v(0) { for (int i = 0; i < N; ++i) d[i] = 0; }
  Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0; }
  static Dual var(double x, int k) { Dual r(x); r.d[k] = 1; return r; }
};
template <int N> Dual<N> chain(const Dual<N>& a, double f, double df) {
  Dual<N> r; r.v = f; for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i]; return r;
}
template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v / b.v; for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v); return r; }
template <int N> Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) { return a * Dual<N>(b); }
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return Dual<N>(a) * b; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> Dual<N> exp(const Dual<N>& a) { double e = std::exp(a.v); return chain(a, e, e); }
template <int N> Dual<N> log(const Dual<N>& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
template <int N> Dual<N> sqrt(const Dual<N>& a) { double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int N> Dual<N> pow(const Dual<N>& a, double p) { double f = std::pow(a.v, p); return chain(a, f, p * std::pow(a.v, p - 1.0)); }
template <int N> Dual<N> atan(const Dual<N>& a) { return chain(a, std::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
template <int N> Dual<N> asinh(const Dual<N>& a) { return chain(a, std::asinh(a.v), 1.0 / std::sqrt(1.0 + a.v * a.v)); }
template <int N> Dual<N> tanh(const Dual<N>& a) { double t = std::tanh(a.v); return chain(a, t, 1.0 - t * t); }
template <int N> Dual<N> expm1(const Dual<N>& a) { return chain(a, std::expm1(a.v), std::exp(a.v)); }
template <int N> Dual<N> log1p(const Dual<N>& a) { return chain(a, std::log1p(a.v), 1.0 / (1.0 + a.v)); }
inline double exp(double a) { return std::exp(a); }
inline double log(double a) { return std::log(a); }
inline double sqrt(double a) { return std::sqrt(a); }
inline double pow(double a, double p) { return std::pow(a, p); }
inline double atan(double a) { return std::atan(a); }
inline double asinh(double a) { return std::asinh(a); }

constexpr double kPi = 3.14159265358979323846;

// Second-order forward mode in N variables (value, gradient, Hessian), for
// f_xc (libxc's xc_*_fxc, src/ExchangeCorrelation.cpp:101-104,126-133).
template <int N> struct Hyper {
  double v, g[N], h[N][N];
  Hyper() : v(0) { for (int i = 0; i < N; ++i) { g[i] = 0; for (int j = 0; j < N; ++j) h[i][j] = 0; } }
  Hyper(double x) : Hyper() { v = x; }
  static Hyper var(double x, int k) { Hyper r(x); r.g[k] = 1; return r; }
};
template <int N> Hyper<N> chain(const Hyper<N>& a, double f, double f1, double f2) {
  Hyper<N> r; r.v = f;
  for (int i = 0; i < N; ++i) {
    r.g[i] = f1 * a.g[i];
    for (int j = 0; j < N; ++j) r.h[i][j] = f1 * a.h[i][j] + f2 * a.g[i] * a.g[j];
  }
  return r;
}
template <int N> Hyper<N> operator+(const Hyper<N>& a, const Hyper<N>& b) {
  Hyper<N> r; r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) { r.g[i] = a.g[i] + b.g[i]; for (int j = 0; j < N; ++j) r.h[i][j] = a.h[i][j] + b.h[i][j]; }
  return r;
}
template <int N> Hyper<N> operator-(const Hyper<N>& a) {
  Hyper<N> r; r.v = -a.v;
  for (int i = 0; i < N; ++i) { r.g[i] = -a.g[i]; for (int j = 0; j < N; ++j) r.h[i][j] = -a.h[i][j]; }
  return r;
}
template <int N> Hyper<N> operator-(const Hyper<N>& a, const Hyper<N>& b) { return a + (-b); }
template <int N> Hyper<N> operator*(const Hyper<N>& a, const Hyper<N>& b) {
  Hyper<N> r; r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) {
    r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    for (int j = 0; j < N; ++j) r.h[i][j] = a.h[i][j] * b.v + a.g[i] * b.g[j] + a.g[j] * b.g[i] + a.v * b.h[i][j];
  }
  return r;
}
template <int N> Hyper<N> recip(const Hyper<N>& a) { const double q = 1.0 / a.v; return chain(a, q, -q * q, 2 * q * q * q); }
template <int N> Hyper<N> operator/(const Hyper<N>& a, const Hyper<N>& b) { return a * recip(b); }
template <int N> Hyper<N> operator+(const Hyper<N>& a, double b) { return a + Hyper<N>(b); }
template <int N> Hyper<N> operator+(double a, const Hyper<N>& b) { return Hyper<N>(a) + b; }
template <int N> Hyper<N> operator-(const Hyper<N>& a, double b) { return a - Hyper<N>(b); }
template <int N> Hyper<N> operator-(double a, const Hyper<N>& b) { return Hyper<N>(a) - b; }
template <int N> Hyper<N> operator*(const Hyper<N>& a, double b) { return a * Hyper<N>(b); }
template <int N> Hyper<N> operator*(double a, const Hyper<N>& b) { return Hyper<N>(a) * b; }
template <int N> Hyper<N> operator/(const Hyper<N>& a, double b) { return a * (1.0 / b); }
template <int N> Hyper<N> operator/(double a, const Hyper<N>& b) { return recip(b) * a; }
template <int N> Hyper<N> exp(const Hyper<N>& a) { const double e = std::exp(a.v); return chain(a, e, e, e); }
template <int N> Hyper<N> log(const Hyper<N>& a) { return chain(a, std::log(a.v), 1.0 / a.v, -1.0 / (a.v * a.v)); }
template <int N> Hyper<N> sqrt(const Hyper<N>& a) { const double q = std::sqrt(a.v); return chain(a, q, 0.5 / q, -0.25 / (q * a.v)); }
template <int N> Hyper<N> pow(const Hyper<N>& a, double p) {
  return chain(a, std::pow(a.v, p), p * std::pow(a.v, p - 1.0), p * (p - 1.0) * std::pow(a.v, p - 2.0));
}
template <int N> Hyper<N> atan(const Hyper<N>& a) {
  const double q = 1.0 / (1.0 + a.v * a.v);
  return chain(a, std::atan(a.v), q, -2.0 * a.v * q * q);
}
template <int N> Hyper<N> asinh(const Hyper<N>& a) {
  const double q = 1.0 / std::sqrt(1.0 + a.v * a.v);
  return chain(a, std::asinh(a.v), q, -a.v * q * q * q);
}


// ---- LDA exchange (libxc 1, Slater/Dirac): e = -(3/4)(3/pi)^(1/3) rho^(4/3)
template <class T> T lda_x(const T& rho) {
  const double cx = -0.75 * std::cbrt(3.0 / kPi);
  return cx * pow(rho, 4.0 / 3.0);
}

// ---- VWN paramagnetic interpolation, eps_c(rs) in Hartree.
// Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200 (1980), eq. 4.4.
struct VwnPar { double A, b, c, x0; };
constexpr VwnPar kVwn5 = {0.0310907, 3.72744, 12.9352, -0.10498};       // libxc 7 (fit to Ceperley-Alder)
constexpr VwnPar kVwnRpa = {0.0310907, 13.0720, 42.7198, -0.409286};    // libxc 8 (fit to RPA)
template <class T> T vwn_eps(const T& rs, const VwnPar& p) {
  const T x = sqrt(rs);
  const T X = x * x + p.b * x + p.c;
  const double Q = std::sqrt(4.0 * p.c - p.b * p.b);
  const double X0 = p.x0 * p.x0 + p.b * p.x0 + p.c;
  const T at = atan(Q / (2.0 * x + p.b));
  return p.A * (log(x * x / X) + (2.0 * p.b / Q) * at
                - (p.b * p.x0 / X0) * (log((x - p.x0) * (x - p.x0) / X) + (2.0 * (p.b + 2.0 * p.x0) / Q) * at));
}
template <class T> T rs_of(const T& rho) { return pow(3.0 / (4.0 * kPi) / rho, 1.0 / 3.0); }
template <class T> T lda_c_vwn(const T& rho, const VwnPar& p) { return rho * vwn_eps(rs_of(rho), p); }

// ---- PW92 paramagnetic correlation with libxc's PW_MOD parameters.
// Perdew, Wang, PRB 45, 13244 (1992), eq. 10; G(rs; A, a1, b1..b4, p=1).
template <class T> T pw92_eps(const T& rs) {
  const double A = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const T srs = sqrt(rs);
  const T den = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * log(1.0 + 1.0 / den);
}

// ---- PBE exchange. Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996), eq. 14.
template <class T> T gga_x_pbe(const T& rho, const T& sigma, double kappa, double mu) {
  const T ex_lda = lda_x(rho);
  const double c2 = 4.0 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const T s2 = sigma / (c2 * pow(rho, 8.0 / 3.0));
  const T F = 1.0 + kappa - kappa / (1.0 + mu * s2 / kappa);
  return ex_lda * F;
}

// ---- PBE correlation (zeta = 0, phi = 1). PRL 77, 3865, eqs. 7-8.
template <class T> T gga_c_pbe(const T& rho, const T& sigma, double beta) {
  const double gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
  const T rs = rs_of(rho);
  const T ec = pw92_eps(rs);
  const T kf = pow(3.0 * kPi * kPi * rho, 1.0 / 3.0);
  const T ks2 = 4.0 * kf / kPi;
  const T t2 = sigma / (4.0 * ks2 * rho * rho);
  const T A = (beta / gamma) / (exp(-ec / gamma) - 1.0);
  const T At2 = A * t2;
  const T H = gamma * log(1.0 + (beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return rho * (ec + H);
}

// ---- Becke 88 exchange, spin-unpolarised. Becke, PRA 38, 3098 (1988), eq. 8.
template <class T> T gga_x_b88(const T& rho, const T& sigma) {
  const double beta = 0.0042;
  const T rs_ = 0.5 * rho;                         // rho_sigma
  const T gs = 0.5 * sqrt(sigma);                  // |grad rho_sigma|
  const T r43 = pow(rs_, 4.0 / 3.0);
  const T x = gs / r43;
  const double cx = 1.5 * std::cbrt(3.0 / (4.0 * kPi));
  const T per_spin = -r43 * (cx + beta * x * x / (1.0 + 6.0 * beta * x * asinh(x)));
  return 2.0 * per_spin;
}

// ---- LYP correlation, closed shell. Miehlich, Savin, Stoll, Preuss,
// CPL 157, 200 (1989), eq. 2 with rho_a = rho_b = rho/2, |grad rho_a|^2 = sigma/4.
template <class T> T gga_c_lyp(const T& rho, const T& sigma) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const T rm13 = pow(rho, -1.0 / 3.0);
  const T den = 1.0 + d * rm13;
  const T delta = c * rm13 + d * rm13 / den;
  const T ex = exp(-c * rm13);
  const T t1 = -a * rho / den;
  const T t2 = -a * b * ex / den * (CF * rho - pow(rho, -5.0 / 3.0) * sigma * (1.0 / 24.0 + 7.0 * delta / 72.0));
  return t1 + t2;
}

// ---- r2SCAN (Furness, Kaplan, Ning, Perdew, Sun, JPCL 11, 8208 (2020) and
// erratum JPCL 11, 9248), spin-unpolarised. Shared pieces:
//   p = |grad rho|^2 / (4 (3 pi^2)^(2/3) rho^(8/3)),
//   alpha_bar = (tau - tau_W) / (tau_unif + eta tau_W), eta = 0.001,
//   tau_W = sigma / (8 rho), tau_unif = (3/10)(3 pi^2)^(2/3) rho^(5/3).
constexpr double kR2Eta = 0.001;
constexpr double kR2Dp2 = 0.361;
constexpr double kR2Cx[8] = {1.0, -0.667, -0.4445555, -0.663086601049, 1.451297044490, -0.887998041597,
                             0.234528941479, -0.023185843322};
constexpr double kR2Cc[8] = {1.0, -0.64, -0.4352, -1.535685604549, 3.061560252175, -1.915710236206,
                             0.516884468372, -0.051848879792};

// Interpolation f(alpha_bar): exp(-c1 a / (1 - a)) for a <= 0, the degree-7
// polynomial for 0 < a <= 2.5, -d exp(c2 / (1 - a)) for a > 2.5.
template <class T> T r2_interp(const T& a, const double* c, double c1, double c2, double d) {
  if (a.v <= 0.0) return exp(-c1 * a / (1.0 - a));
  if (a.v <= 2.5) {
    T r = T(c[7]);
    for (int i = 6; i >= 0; --i) r = r * a + c[i];
    return r;
  }
  return -d * exp(c2 / (1.0 - a));
}
inline double r2_poly_slope(const double* c) {  // sum_{i>=1} i c_i
  double s = 0;
  for (int i = 1; i < 8; ++i) s += i * c[i];
  return s;
}

template <class T> void r2_reduced(const T& rho, const T& sigma, const T& tau, T* p, T* abar) {
  const double k2 = std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const T r83 = pow(rho, 8.0 / 3.0);
  *p = sigma / (4.0 * k2 * r83);
  const T tw = sigma / (8.0 * rho);
  const T tu = 0.3 * k2 * pow(rho, 5.0 / 3.0);
  *abar = (tau - tw) / (tu + kR2Eta * tw);
}

template <class T> T mgga_x_r2scan(const T& rho, const T& sigma, const T& tau) {
  const double h0x = 1.174, k1 = 0.065, muak = 10.0 / 81.0, a1 = 4.9479;
  const double ceta = 20.0 / 27.0 + kR2Eta * 5.0 / 3.0;
  const double c2 = -r2_poly_slope(kR2Cx) * (1.0 - h0x);
  T p, ab;
  r2_reduced(rho, sigma, tau, &p, &ab);
  const T x = (ceta * c2 * exp(-p * p / std::pow(kR2Dp2, 4)) + muak) * p;
  const T h1x = 1.0 + k1 - k1 / (1.0 + x / k1);
  const T fx = r2_interp(ab, kR2Cx, 0.667, 0.8, 1.24);
  const T gx = p.v > 0 ? 1.0 - exp(-a1 / pow(p, 0.25)) : T(1.0);
  return lda_x(rho) * (h1x + fx * (h0x - h1x)) * gx;
}

template <class T> T mgga_c_r2scan(const T& rho, const T& sigma, const T& tau) {
  const double gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
  const double b1c = 0.0285764, b2c = 0.0889, b3c = 0.125541, chi_inf = 0.128026;
  T p, ab;
  r2_reduced(rho, sigma, tau, &p, &ab);
  const T rs = rs_of(rho);
  const T srs = sqrt(rs);
  // LSDA1 = PW92 and its rs-derivative; LSDA0 = -b1c / (1 + b2c sqrt(rs) + b3c rs)
  const double A = 0.0310907, a1 = 0.21370, B1 = 7.5957, B2 = 3.5876, B3 = 1.6382, B4 = 0.49294;
  const T Q = 2.0 * A * (B1 * srs + B2 * rs + B3 * rs * srs + B4 * rs * rs);
  const T dQ = 2.0 * A * (0.5 * B1 / srs + B2 + 1.5 * B3 * srs + 2.0 * B4 * rs);
  const T L = log(1.0 + 1.0 / Q);
  const T e1 = -2.0 * A * (1.0 + a1 * rs) * L;
  const T de1 = -2.0 * A * a1 * L + 2.0 * A * (1.0 + a1 * rs) * dQ / (Q * Q + Q);
  const T den0 = 1.0 + b2c * srs + b3c * rs;
  const T e0lda = -b1c / den0;
  const T de0 = b1c * (0.5 * b2c / srs + b3c) / (den0 * den0);
  // eps_c^1 = eps_LSDA1 + H1 with the r2SCAN gradient correction Delta y
  const T beta = 0.066725 * (1.0 + 0.1 * rs) / (1.0 + 0.1778 * rs);
  const T w1 = exp(-e1 / gamma) - 1.0;
  const T t2 = std::pow(3.0 * kPi * kPi / 16.0, 2.0 / 3.0) * p / rs;
  const T y = beta * t2 / (gamma * w1);
  const double dfc2 = r2_poly_slope(kR2Cc);
  const T dy = dfc2 / (27.0 * gamma * w1) * (20.0 * rs * (de0 - de1) - 45.0 * kR2Eta * (e0lda - e1)) * p *
               exp(-p * p / std::pow(kR2Dp2, 4));
  const T g = pow(1.0 + 4.0 * (y - dy), -0.25);
  const T H1 = gamma * log(1.0 + w1 * (1.0 - g));
  const T ec1 = e1 + H1;
  // eps_c^0 = eps_LDA0 + H0 (G_c(zeta = 0) = 1)
  const T w0 = exp(-e0lda / b1c) - 1.0;
  const T ginf = pow(1.0 + 4.0 * chi_inf * p, -0.25);
  const T ec0 = e0lda + b1c * log(1.0 + w0 * (1.0 - ginf));
  const T fc = r2_interp(ab, kR2Cc, 0.64, 1.5, 0.7);
  return rho * (ec1 + fc * (ec0 - ec1));
}

// ---- Functional table (libxc IDs used by the reference's .df files).
struct Component { int kind; double coef; };
// kind: 1 lda_x, 7 vwn5, 8 vwn_rpa, 101/321 pbe_x, 130/322 pbe_c, 106 b88, 131 lyp
inline int family_of(int code) {
  switch (code) {
    case 1: case 7: case 8: return 0;
    case 101: case 130: case 321: case 322: case 106: case 131: case 402: return 1;
    case 497: case 498: return 2;
    default: return -1;
  }
}
inline double density_threshold(int /*code*/) { return 1e-15; }

// Energy density rho*eps of one libxc code, on dual inputs (rho, sigma, tau).
template <class T> T energy_density(int code, const T& rho, const T& sigma, const T& tau) {
  switch (code) {
    case 497: return mgga_x_r2scan(rho, sigma, tau);
    case 498: return mgga_c_r2scan(rho, sigma, tau);
    case 1: return lda_x(rho);
    case 7: return lda_c_vwn(rho, kVwn5);
    case 8: return lda_c_vwn(rho, kVwnRpa);
    case 101: return gga_x_pbe(rho, sigma, 0.804, 0.2195149727645171);
    case 321: return gga_x_pbe(rho, sigma, 0.804, 0.219510);
    case 130: return gga_c_pbe(rho, sigma, 0.06672455060314922);
    case 322: return gga_c_pbe(rho, sigma, 0.06672632268006);
    case 106: return gga_x_b88(rho, sigma);
    case 131: return gga_c_lyp(rho, sigma);
    case 402:  // B3LYP (XC part; 20% exact exchange handled by the caller)
      return 0.08 * lda_x(rho) + 0.72 * gga_x_b88(rho, sigma) + 0.19 * lda_c_vwn(rho, kVwnRpa) + 0.81 * gga_c_lyp(rho, sigma);
    default: return T(0.0);
  }
}

// Evaluate one code at one point, accumulating like src/ExchangeCorrelation.cpp:202-217.
inline void eval_point(int code, double rho, double sigma, double tau,
                       double* eps, double* vrho, double* vsigma, double* vtau) {
  const double thr = density_threshold(code);
  if (!(rho >= thr)) return;
  const double sthr = std::pow(thr, 4.0 / 3.0);
  double s = sigma > sthr * sthr ? sigma : sthr * sthr;
  double t = tau;
  if (family_of(code) == 2) {
    t = tau > 1e-20 ? tau : 1e-20;
    if (s > 8.0 * rho * t) s = 8.0 * rho * t;
  }
  typedef Dual<3> D3;
  const D3 e = energy_density(code, D3::var(rho, 0), D3::var(s, 1), D3::var(t, 2));
  *eps += e.v / rho;
  *vrho += e.d[0];
  *vsigma += e.d[1];
  *vtau += e.d[2];
}

// Second derivatives of one LDA/GGA code at one point (libxc xc_*_fxc):
// accumulates v_sigma, v2rho2, v2rhosigma, v2sigma2 with the same threshold
// and sigma floor as eval_point.
inline void eval_point_f(int code, double rho, double sigma, double* vsigma, double* v2rho2, double* v2rhosigma,
                         double* v2sigma2) {
  const double thr = density_threshold(code);
  if (!(rho >= thr)) return;
  const double sthr = std::pow(thr, 4.0 / 3.0);
  const double s = sigma > sthr * sthr ? sigma : sthr * sthr;
  typedef Hyper<2> H;
  const H e = energy_density(code, H::var(rho, 0), H::var(s, 1), H(0.0));
  *vsigma += e.g[1];
  *v2rho2 += e.h[0][0];
  *v2rhosigma += e.h[0][1];
  *v2sigma2 += e.h[1][1];
}

// ---- Spin-polarised LDA/GGA (libxc XC_POLARIZED): variables rho_a, rho_b,
// sigma_aa, sigma_ab, sigma_bb (Dual<5>). Exchange by spin scaling
// E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2; correlation with the published
// spin interpolations.
typedef Dual<5> P5d;

// 1 + zeta and 1 - zeta enter each power clamped at kZetaThr from below (a
// constant there, as libxc's piecewise opz_pow_n), so a fully polarised point
// keeps the derivatives of the other term.
constexpr double kZetaThr = 1e-15;
template <class T> T opz_pow(const T& x, double n) {
  return x.v > kZetaThr ? pow(x, n) : T(std::pow(kZetaThr, n));
}
template <class T> T spin_f(const T& z) {  // f(zeta) = ((1+z)^(4/3) + (1-z)^(4/3) - 2) / (2^(4/3) - 2)
  const T a = 1.0 + z, b = 1.0 - z;
  auto p43 = [](const T& x) { return opz_pow(x, 4.0 / 3.0); };
  return (p43(a) + p43(b) - 2.0) / (std::pow(2.0, 4.0 / 3.0) - 2.0);
}
inline double spin_fpp0() { return 4.0 / (9.0 * (std::cbrt(2.0) - 1.0)); }

// VWN with spin: Vosko, Wilk, Nusair (1980); libxc 7 uses the spin-stiffness
// interpolation, libxc 8 (RPA) the plain f(zeta) one.
constexpr VwnPar kVwn5F = {0.01554535, 7.06042, 18.0578, -0.32500};
constexpr VwnPar kVwnRpaF = {0.01554535, 20.1231, 101.578, -0.743294};
template <class T> T vwn5_pol_eps(const T& rs, const T& z) {
  const VwnPar stiff = {-1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584};
  const T eP = vwn_eps(rs, kVwn5), eF = vwn_eps(rs, kVwn5F), ac = vwn_eps(rs, stiff);
  const T f = spin_f(z), z4 = z * z * z * z;
  return eP + ac * f * (1.0 - z4) / spin_fpp0() + (eF - eP) * f * z4;
}
template <class T> T vwnrpa_pol_eps(const T& rs, const T& z) {
  const T eP = vwn_eps(rs, kVwnRpa), eF = vwn_eps(rs, kVwnRpaF);
  return eP + (eF - eP) * spin_f(z);
}

// PW92 with spin, PW_MOD parameters (eqs. 8, 10).
template <class T> T pw92_G(const T& rs, const double* q) {  // q = A, a1, b1, b2, b3, b4
  const T srs = sqrt(rs);
  const T den = 2.0 * q[0] * (q[2] * srs + q[3] * rs + q[4] * rs * srs + q[5] * rs * rs);
  return -2.0 * q[0] * (1.0 + q[1] * rs) * log(1.0 + 1.0 / den);
}
// dG/drs of the same G (closed form).
template <class T> T pw92_G_drs(const T& rs, const double* q) {
  const T srs = sqrt(rs);
  const T den = 2.0 * q[0] * (q[2] * srs + q[3] * rs + q[4] * rs * srs + q[5] * rs * rs);
  const T dden = 2.0 * q[0] * (0.5 * q[2] / srs + q[3] + 1.5 * q[4] * srs + 2.0 * q[5] * rs);
  return -2.0 * q[0] * q[1] * log(1.0 + 1.0 / den) + 2.0 * q[0] * (1.0 + q[1] * rs) * dden / (den * den + den);
}
constexpr double kPw92Q0[6] = {0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
constexpr double kPw92Q1[6] = {0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
constexpr double kPw92Qa[6] = {0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};
// PW92 eps_c(rs, zeta); with drs, its partial derivative in rs at fixed zeta.
template <class T> T pw92_pol_eps(const T& rs, const T& z, T* drs = nullptr) {
  const T e0 = pw92_G(rs, kPw92Q0), e1 = pw92_G(rs, kPw92Q1), alpha_c = -pw92_G(rs, kPw92Qa);
  const T f = spin_f(z), z4 = z * z * z * z;
  if (drs) {
    const T d0 = pw92_G_drs(rs, kPw92Q0), d1 = pw92_G_drs(rs, kPw92Q1), da = -pw92_G_drs(rs, kPw92Qa);
    *drs = d0 + da * f * (1.0 - z4) / spin_fpp0() + (d1 - d0) * f * z4;
  }
  return e0 + alpha_c * f * (1.0 - z4) / spin_fpp0() + (e1 - e0) * f * z4;
}

// PBE correlation with spin (PRL 77, 3865, eqs. 3, 7-8).
template <class T> T pbe_c_pol(const T& rho, const T& z, const T& gg, double beta) {
  const double gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
  const T ec = pw92_pol_eps(rs_of(rho), z);
  auto p23 = [](const T& x) { return opz_pow(x, 2.0 / 3.0); };
  const T phi = 0.5 * (p23(1.0 + z) + p23(1.0 - z));
  const T kf = pow(3.0 * kPi * kPi * rho, 1.0 / 3.0);
  const T ks2 = 4.0 * kf / kPi;
  const T t2 = gg / (4.0 * phi * phi * ks2 * rho * rho);
  const T g3 = gamma * phi * phi * phi;
  const T A = (beta / gamma) / (exp(-ec / g3) - 1.0);
  const T At2 = A * t2;
  return rho * (ec + g3 * log(1.0 + (beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2)));
}

// LYP, general spin form (CPL 157, 200, eq. 2).
template <class T> T lyp_pol(const T& ra, const T& rb, const T& gaa, const T& gab, const T& gbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const T rho = ra + rb;
  const T rm13 = pow(rho, -1.0 / 3.0);
  const T den = 1.0 + d * rm13;
  const T omega = exp(-c * rm13) / den * pow(rho, -11.0 / 3.0);
  const T delta = c * rm13 + d * rm13 / den;
  const T g = gaa + gbb + 2.0 * gab;
  auto p83 = [](const T& x) { return x.v > 0 ? pow(x, 8.0 / 3.0) : T(0.0); };
  const T bracket = ra * rb * (std::pow(2.0, 11.0 / 3.0) * CF * (p83(ra) + p83(rb)) +
                                 (47.0 / 18.0 - 7.0 / 18.0 * delta) * g - (2.5 - delta / 18.0) * (gaa + gbb) -
                                 (delta - 11.0) / 9.0 * (ra / rho * gaa + rb / rho * gbb)) -
                      2.0 / 3.0 * rho * rho * g + (2.0 / 3.0 * rho * rho - ra * ra) * gbb +
                      (2.0 / 3.0 * rho * rho - rb * rb) * gaa;
  return -a * 4.0 / den * ra * rb / rho - a * b * omega * bracket;
}

// One exchange channel of spin density r_s with |grad r_s|^2 = g_s: E_x[2 r_s, 4 g_s] / 2.
template <class T> T x_channel(int code, const T& r_s, const T& g_s) {
  const T r2 = 2.0 * r_s, g4 = 4.0 * g_s;
  switch (code) {
    case 1: return 0.5 * lda_x(r2);
    case 101: return 0.5 * gga_x_pbe(r2, g4, 0.804, 0.2195149727645171);
    case 321: return 0.5 * gga_x_pbe(r2, g4, 0.804, 0.219510);
    case 106: return 0.5 * gga_x_b88(r2, g4);
    default: return T(0.0);
  }
}

template <class T> T energy_density_pol(int code, const T& ra, const T& rb, const T& gaa, const T& gab, const T& gbb,
                                        bool a_on, bool b_on) {
  auto xs = [&](int c) {
    T e(0.0);
    if (a_on) e = e + x_channel(c, ra, gaa);
    if (b_on) e = e + x_channel(c, rb, gbb);
    return e;
  };
  const T rho = ra + rb;
  const T z = (ra - rb) / rho;  // |z| <= 1; each power of 1 +- z is clamped (opz_pow)
  const T g = gaa + gbb + 2.0 * gab;
  switch (code) {
    case 1: case 101: case 321: case 106: return xs(code);
    case 7: return rho * vwn5_pol_eps(rs_of(rho), z);
    case 8: return rho * vwnrpa_pol_eps(rs_of(rho), z);
    case 130: return pbe_c_pol(rho, z, g, 0.06672455060314922);
    case 322: return pbe_c_pol(rho, z, g, 0.06672632268006);
    case 131: return lyp_pol(ra, rb, gaa, gab, gbb);
    case 402:
      return 0.08 * xs(1) + 0.72 * xs(106) + 0.19 * (rho * vwnrpa_pol_eps(rs_of(rho), z)) + 0.81 * lyp_pol(ra, rb, gaa, gab, gbb);
    default: return T(0.0);
  }
}

// One code at one point, spin-polarised, accumulating eps (per particle of the
// total density), vrho[2] and vsigma[3]. Same thresholds as the unpolarised
// path on the total density; an exchange channel below the threshold is
// skipped; sigma_aa, sigma_bb >= 1e-40 and |sigma_ab| <= sqrt(sigma_aa sigma_bb).
inline void eval_point_pol(int code, double ra, double rb, double saa, double sab, double sbb, double* eps,
                           double* vrho, double* vsigma) {
  ra = ra > 0 ? ra : 0;
  rb = rb > 0 ? rb : 0;
  const double thr = density_threshold(code);
  if (!(ra + rb >= thr)) return;
  saa = std::max(saa, 1e-40);
  sbb = std::max(sbb, 1e-40);
  const double cap = std::sqrt(saa * sbb);
  sab = std::min(std::max(sab, -cap), cap);
  const P5d e = energy_density_pol(code, P5d::var(ra, 0), P5d::var(rb, 1), P5d::var(saa, 2), P5d::var(sab, 3),
                                   P5d::var(sbb, 4), ra >= thr, rb >= thr);
  *eps += e.v / (ra + rb);
  vrho[0] += e.d[0];
  vrho[1] += e.d[1];
  vsigma[0] += e.d[2];
  vsigma[1] += e.d[3];
  vsigma[2] += e.d[4];
}

// Second derivatives of one code at one point, spin-polarised (libxc
// xc_{lda,gga}_fxc with XC_POLARIZED, the layout Evaluate("f") copies,
// src/ExchangeCorrelation.cpp:126-133,232-245): accumulates vsigma[3] (aa, ab,
// bb), v2rho2[3] (a a, a b, b b), v2rhosigma[6] (a aa, a ab, a bb, b aa, b ab,
// b bb) and v2sigma2[6] (aa aa, aa ab, aa bb, ab ab, ab bb, bb bb). Same
// thresholds, floors and caps as eval_point_pol.
inline void eval_point_fpol(int code, double ra, double rb, double saa, double sab, double sbb, double* vsigma,
                            double* v2rho2, double* v2rhosigma, double* v2sigma2) {
  ra = ra > 0 ? ra : 0;
  rb = rb > 0 ? rb : 0;
  const double thr = density_threshold(code);
  if (!(ra + rb >= thr)) return;
  saa = std::max(saa, 1e-40);
  sbb = std::max(sbb, 1e-40);
  const double cap = std::sqrt(saa * sbb);
  sab = std::min(std::max(sab, -cap), cap);
  typedef Hyper<5> H5;
  const H5 e = energy_density_pol(code, H5::var(ra, 0), H5::var(rb, 1), H5::var(saa, 2), H5::var(sab, 3),
                                  H5::var(sbb, 4), ra >= thr, rb >= thr);
  for (int x = 0; x < 3; ++x) vsigma[x] += e.g[2 + x];
  v2rho2[0] += e.h[0][0];
  v2rho2[1] += e.h[0][1];
  v2rho2[2] += e.h[1][1];
  for (int w = 0; w < 2; ++w)
    for (int x = 0; x < 3; ++x) v2rhosigma[3 * w + x] += e.h[w][2 + x];
  for (int x = 0, k = 0; x < 3; ++x)
    for (int y = x; y < 3; ++y, ++k) v2sigma2[k] += e.h[2 + x][2 + y];
}

// ---- Spin-polarised r2SCAN (libxc 497 / 498 with XC_POLARIZED): variables
// rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb, tau_a, tau_b (Dual<7>).
// Exchange by spin scaling, E_x[ra, rb, ta, tb] = (E_x[2 ra, 4 saa, 2 ta] +
// E_x[2 rb, 4 sbb, 2 tb]) / 2. Correlation (Furness et al. 2020 with the
// spin dependence of SCAN, Sun, Ruzsinszky, Perdew, PRL 115, 036402 (2015)):
//   zeta = (ra - rb) / rho, phi = ((1+z)^(2/3) + (1-z)^(2/3)) / 2,
//   d_s = ((1+z)^(5/3) + (1-z)^(5/3)) / 2, d_x = ((1+z)^(4/3) + (1-z)^(4/3)) / 2,
//   alpha_bar = (tau - tau_W) / (tau_unif d_s + eta tau_W) on the totals,
//   eps_c^1 = eps_PW92(rs, z) + gamma phi^3 ln(1 + w1 (1 - g(y - dy))),
//     w1 = exp(-eps_PW92 / (gamma phi^3)) - 1, y = beta t^2 / (gamma w1),
//     t^2 = (3 pi^2 / 16)^(2/3) p / (phi^2 rs),
//     dy = Dfc2 / (27 gamma d_s phi^3 w1) (20 rs (d eps_0 / d rs - d eps_PW92 / d rs)
//          - 45 eta (eps_0 - eps_PW92)) p exp(-p^2 / dp2^4), eps_0 = eps_LDA0 G_c,
//   eps_c^0 = (eps_LDA0 + b1c ln(1 + w0 (1 - g_inf(p)))) G_c(z),
//     G_c = (1 - 2.3631 (d_x - 1)) (1 - z^12),
//   eps_c = eps_c^1 + f_c(alpha_bar) (eps_c^0 - eps_c^1).
// At z = 0 this is the unpolarised mgga_c_r2scan above.
typedef Dual<7> P7d;

template <class T> T mgga_c_r2scan_pol(const T& rho, const T& z, const T& sigma, const T& tau) {
  const double gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
  const double b1c = 0.0285764, b2c = 0.0889, b3c = 0.125541, chi_inf = 0.128026;
  const double k2 = std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  const T opz = 1.0 + z, omz = 1.0 - z;
  auto pw = [](const T& x, double e) { return x.v > 0 ? pow(x, e) : T(0.0); };
  const T phi = 0.5 * (pw(opz, 2.0 / 3.0) + pw(omz, 2.0 / 3.0));
  const T ds = 0.5 * (pw(opz, 5.0 / 3.0) + pw(omz, 5.0 / 3.0));
  const T dx = 0.5 * (pw(opz, 4.0 / 3.0) + pw(omz, 4.0 / 3.0));
  const T z2 = z * z, z4 = z2 * z2, z12 = z4 * z4 * z4;
  const T Gc = (1.0 - 2.3631 * (dx - 1.0)) * (1.0 - z12);
  const T p = sigma / (4.0 * k2 * pow(rho, 8.0 / 3.0));
  const T tw = sigma / (8.0 * rho);
  const T ab = (tau - tw) / (0.3 * k2 * pow(rho, 5.0 / 3.0) * ds + kR2Eta * tw);
  const T rs = rs_of(rho);
  const T srs = sqrt(rs);
  T de1;
  const T e1 = pw92_pol_eps(rs, z, &de1);
  const T den0 = 1.0 + b2c * srs + b3c * rs;
  const T e0lda = -b1c / den0;
  const T de0lda = b1c * (0.5 * b2c / srs + b3c) / (den0 * den0);
  const T phi3 = phi * phi * phi;
  const T beta = 0.066725 * (1.0 + 0.1 * rs) / (1.0 + 0.1778 * rs);
  const T w1 = exp(-e1 / (gamma * phi3)) - 1.0;
  const T t2 = std::pow(3.0 * kPi * kPi / 16.0, 2.0 / 3.0) * p / (phi * phi * rs);
  const T y = beta * t2 / (gamma * w1);
  const double dfc2 = r2_poly_slope(kR2Cc);
  const T dy = dfc2 / (27.0 * gamma * ds * phi3 * w1) *
               (20.0 * rs * (de0lda * Gc - de1) - 45.0 * kR2Eta * (e0lda * Gc - e1)) * p *
               exp(-p * p / std::pow(kR2Dp2, 4));
  const T g = pow(1.0 + 4.0 * (y - dy), -0.25);
  const T ec1 = e1 + gamma * phi3 * log(1.0 + w1 * (1.0 - g));
  const T w0 = exp(-e0lda / b1c) - 1.0;
  const T ginf = pow(1.0 + 4.0 * chi_inf * p, -0.25);
  const T ec0 = (e0lda + b1c * log(1.0 + w0 * (1.0 - ginf))) * Gc;
  const T fc = r2_interp(ab, kR2Cc, 0.64, 1.5, 0.7);
  return rho * (ec1 + fc * (ec0 - ec1));
}

// Energy density of one code with spin, meta-GGA variables; LDA/GGA codes take
// the P5 path's forms (their energy does not depend on tau).
inline P7d energy_density_mpol(int code, const P7d& ra, const P7d& rb, const P7d& saa, const P7d& sab,
                               const P7d& sbb, const P7d& ta, const P7d& tb, bool a_on, bool b_on) {
  const P7d rho = ra + rb;
  P7d z = (ra - rb) / rho;
  if (z.v > 1.0 - 1e-15) z = P7d(1.0 - 1e-15);  // zeta threshold
  if (z.v < -1.0 + 1e-15) z = P7d(-1.0 + 1e-15);
  switch (code) {
    case 497: {
      P7d e(0.0);
      if (a_on) e = e + 0.5 * mgga_x_r2scan(2.0 * ra, 4.0 * saa, 2.0 * ta);
      if (b_on) e = e + 0.5 * mgga_x_r2scan(2.0 * rb, 4.0 * sbb, 2.0 * tb);
      return e;
    }
    case 498: return mgga_c_r2scan_pol(rho, z, saa + 2.0 * sab + sbb, ta + tb);
    default: return P7d(0.0);
  }
}

// One meta-GGA code at one point, spin-polarised, accumulating eps, vrho[2],
// vsigma[3], vtau[2]. Density threshold on the total; an exchange channel
// below it is skipped; per spin tau_s >= 1e-20 and sigma_ss <= 8 rho_s tau_s,
// sigma_ss >= 1e-40, |sigma_ab| <= sqrt(sigma_aa sigma_bb).
inline void eval_point_mpol(int code, double ra, double rb, double saa, double sab, double sbb, double ta, double tb,
                            double* eps, double* vrho, double* vsigma, double* vtau) {
  ra = ra > 0 ? ra : 0;
  rb = rb > 0 ? rb : 0;
  const double thr = density_threshold(code);
  if (!(ra + rb >= thr)) return;
  ta = std::max(ta, 1e-20);
  tb = std::max(tb, 1e-20);
  saa = std::min(std::max(saa, 1e-40), 8.0 * ra * ta > 1e-40 ? 8.0 * ra * ta : 1e-40);
  sbb = std::min(std::max(sbb, 1e-40), 8.0 * rb * tb > 1e-40 ? 8.0 * rb * tb : 1e-40);
  const double cap = std::sqrt(saa * sbb);
  sab = std::min(std::max(sab, -cap), cap);
  const P7d e = energy_density_mpol(code, P7d::var(ra, 0), P7d::var(rb, 1), P7d::var(saa, 2), P7d::var(sab, 3),
                                    P7d::var(sbb, 4), P7d::var(ta, 5), P7d::var(tb, 6), ra >= thr, rb >= thr);
  *eps += e.v / (ra + rb);
  vrho[0] += e.d[0];
  vrho[1] += e.d[1];
  for (int k = 0; k < 3; ++k) vsigma[k] += e.d[2 + k];
  vtau[0] += e.d[5];
  vtau[1] += e.d[6];
}

}  // namespace oxf
