// Density-fitted Coulomb matrix (RI-J) and one-electron integrals on the
// device, for the SCF loop around the XC path.
//
// The reference builds J from four-centre integrals (Int4C2E::ContractInts,
// src/Integral/Int4C2E.cpp:673-683) and S, T, V from libint2
// (src/Integral/Representation.cpp:62-63). Here J is fitted in an auxiliary
// basis with the Coulomb metric:
//   gamma_P = sum_mn (P|mn) D_mn,   (P|Q) c = gamma,   J_mn = sum_P (mn|P) c_P,
// with every integral over Hermite Gaussians (McMurchie and Davidson, J.
// Comput. Phys. 26, 218 (1978); Helgaker, Jorgensen, Olsen, Molecular
// Electronic-Structure Theory, ch. 9):
//   (ab|c) = 2 pi^5/2 / (p q sqrt(p + q)) sum_tuv E^ab_tuv sum_t'u'v' (-1)^(t'+u'+v') E^c_t'u'v' R_t+t',u+u',v+v'.
// The work is organised as a J engine: the density is contracted into each
// primitive pair's Hermite coefficients h_tuv = sum_mn D_mn E^mn_tuv once per
// iteration, and the fit coefficients into each auxiliary primitive's
// Hermite coefficients, so the inner loops run over (pair, auxiliary
// primitive) with one Boys evaluation, one R table and a short contraction.
// A pure auxiliary function S_lm(r - C) exp(-g |r - C|^2) is (2g)^-l times the
// Hermite Gaussians of its harmonic polynomial's monomials (Hobson's theorem),
// so only the top Hermite degree l of an auxiliary function enters.
//
// Threads: the gamma pass gives each lane one auxiliary primitive and walks a
// chunk of pairs in step across the wave (pair data uniform: scalar loads);
// the Coulomb pass gives each lane one pair and walks all auxiliary
// primitives in step. Pairs are sorted by class L = la + lb and primitives by
// l, so every wave runs one compile-time class at a time. All arithmetic is
// FP64.
#include <hip/hip_runtime.h>

#include "coulomb.hpp"

namespace chx {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi52 = 34.98683665524972;  // 2 pi^(5/2)

// Boys function F_n(T), n = 0..L: F_L from an 8-term Taylor expansion about
// the nearest tabulated point, then downward recursion
// F_n = (2T F_{n+1} + exp(-T)) / (2n + 1); above kBoysTmax the asymptotic
// F_0 = sqrt(pi / T) / 2 and upward recursion F_{n+1} = ((2n+1) F_n - exp(-T)) / 2T.
template <int L>
__device__ __forceinline__ void boys(const double* __restrict__ tab, double T, double (&F)[L + 1]) {
  if (T < kBoysTmax) {
    const int k = (int)(T * (1.0 / kBoysStep) + 0.5);
    const double y = k * kBoysStep - T;
    const double* row = tab + (size_t)k * kBoysOrders + L;
    double f = row[kBoysTaylor - 1];
#pragma unroll
    for (int j = kBoysTaylor - 2; j >= 0; --j) f = row[j] + f * y / (double)(j + 1);
    F[L] = f;
    if (L > 0) {
      const double e = exp(-T);
#pragma unroll
      for (int n = L - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) / (double)(2 * n + 1);
    }
  } else {
    F[0] = 0.5 * sqrt(kPi / T);
    if (L > 0) {
      const double e = exp(-T), r = 0.5 / T;
#pragma unroll
      for (int n = 0; n < L; ++n) F[n + 1] = ((double)(2 * n + 1) * F[n] - e) * r;
    }
  }
}

// Hermite Coulomb integrals R_tuv = R^0_tuv(alpha, X, Y, Z), t + u + v <= L:
// R^n_000 = (-2 alpha)^n F_n(alpha |XYZ|^2), R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{tuv}
// (and likewise in u with Y, v with Z), level by level from n = L down to 0.
template <int L>
__device__ __forceinline__ void hermite_r(const double* __restrict__ tab, double alpha, double X, double Y, double Z,
                                          double (&R)[jnherm(L)]) {
  double F[L + 1];
  boys<L>(tab, alpha * (X * X + Y * Y + Z * Z), F);
  double lev[L + 1][jnherm(L)];
  double m2a = 1.0;
#pragma unroll
  for (int n = 0; n <= L; ++n) {
    lev[n][0] = m2a * F[n];
    m2a *= -2.0 * alpha;
  }
#pragma unroll
  for (int n = L - 1; n >= 0; --n)
#pragma unroll
    for (int s = 1; s <= L - n; ++s)
#pragma unroll
      for (int t = s; t >= 0; --t)
#pragma unroll
        for (int u = s - t; u >= 0; --u) {
          const int v = s - t - u;
          double val;
          if (t > 0)
            val = (t > 1 ? (t - 1) * lev[n + 1][jhidx(t - 2, u, v)] : 0.0) + X * lev[n + 1][jhidx(t - 1, u, v)];
          else if (u > 0)
            val = (u > 1 ? (u - 1) * lev[n + 1][jhidx(t, u - 2, v)] : 0.0) + Y * lev[n + 1][jhidx(t, u - 1, v)];
          else
            val = (v > 1 ? (v - 1) * lev[n + 1][jhidx(t, u, v - 2)] : 0.0) + Z * lev[n + 1][jhidx(t, u, v - 1)];
          lev[n][jhidx(t, u, v)] = val;
        }
#pragma unroll
  for (int i = 0; i < jnherm(L); ++i) R[i] = lev[0][i];
}

// 1-D Hermite expansion coefficients E[i][j][t] of x_A^i x_B^j exp(...) about
// P, without the exponential prefactor (carried by K): E^00_0 = 1,
// E^{i+1,j}_t = E^ij_{t-1} / 2p + X_PA E^ij_t + (t+1) E^ij_{t+1}, and likewise
// in j with X_PB.
template <int LA, int LB>
struct Herm1 {
  double e[LA + 1][LB + 1][LA + LB + 1];
  __device__ __forceinline__ void build(double o2p, double xpa, double xpb) {
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
      for (int j = 0; j <= LB; ++j)
#pragma unroll
        for (int t = 0; t <= LA + LB; ++t) e[i][j][t] = 0.0;
    e[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
      for (int j = 0; j <= LB; ++j) {
        if (i == 0 && j == 0) continue;
#pragma unroll
        for (int t = 0; t <= i + j; ++t) {
          if (i > 0) {
            double v = xpa * e[i - 1][j][t];
            if (t > 0) v += o2p * e[i - 1][j][t - 1];
            if (t + 1 <= i - 1 + j) v += (t + 1) * e[i - 1][j][t + 1];
            e[i][j][t] = v;
          } else {
            double v = xpb * e[0][j - 1][t];
            if (t > 0) v += o2p * e[0][j - 1][t - 1];
            if (t + 1 <= j - 1) v += (t + 1) * e[0][j - 1][t + 1];
            e[0][j][t] = v;
          }
        }
      }
  }
};

__device__ __forceinline__ int type_index(int type) { return type == 0 ? 0 : type == 1 ? 1 : type == -1 ? 2 : 3; }
// components and angular momentum of a shell type (0 s, 1 Cartesian p, -1 pure p, -2 pure d)
__device__ __forceinline__ int ncomp_of(int type) { return type == 1 ? 3 : 2 * (type < 0 ? -type : type) + 1; }
__device__ __forceinline__ int l_of(int type) { return type == 1 ? 1 : (type < 0 ? -type : type); }

// Cartesian exponents of component c of a degree-l shell
template <int L>
struct CartOf {
  int i[jncart(L)], j[jncart(L)], k[jncart(L)];
  __device__ __forceinline__ CartOf() {
    int n = 0;
#pragma unroll
    for (int a = L; a >= 0; --a)
#pragma unroll
      for (int b = L - a; b >= 0; --b) {
        i[n] = a;
        j[n] = b;
        k[n] = L - a - b;
        ++n;
      }
  }
};

// ---------------------------------------------------------------------------
// Hermite density of each pair: h_tuv = f K sum_{ij} Dc_ij Ex Ey Ez, Dc the
// Cartesian image of the D block (Dc = Ta^T D_ab Tb), f = 2 for a pair of two
// different shells (it stands for both orders), 1 within one shell.
template <int LA, int LB>
__device__ __forceinline__ void herm_density(const JDev& J, int k, const double* __restrict__ D, double* h) {
  constexpr int NA = jncart(LA), NB = jncart(LB), NH = jnherm(LA + LB);
  const double p = J.p[k], o2p = 0.5 / p;
  Herm1<LA, LB> ex, ey, ez;
  ex.build(o2p, J.pax[k], J.pbx[k]);
  ey.build(o2p, J.pay[k], J.pby[k]);
  ez.build(o2p, J.paz[k], J.pbz[k]);
  const int ta = J.ta[k], tb = J.tb[k], aoa = J.aoa[k], aob = J.aob[k], n = J.n;
  const int ma = ncomp_of(ta), mb = ncomp_of(tb);
  const double* Ta = J.orbT + type_index(ta) * 30;
  const double* Tb = J.orbT + type_index(tb) * 30;
  double Dc[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) Dc[i][j] = 0.0;
  for (int a = 0; a < ma; ++a)
    for (int b = 0; b < mb; ++b) {
      const double d = D[(size_t)(aob + b) * n + (aoa + a)];
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) Dc[i][j] += Ta[a * 6 + i] * d * Tb[b * 6 + j];
    }
  const double f = (J.diag[k] ? 1.0 : 2.0) * J.K[k];
  const CartOf<LA> ca;
  const CartOf<LB> cb;
  double acc[NH];
#pragma unroll
  for (int x = 0; x < NH; ++x) acc[x] = 0.0;
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int t = 0; t <= ca.i[i] + cb.i[j]; ++t)
#pragma unroll
        for (int u = 0; u <= ca.j[i] + cb.j[j]; ++u)
#pragma unroll
          for (int v = 0; v <= ca.k[i] + cb.k[j]; ++v)
            acc[jhidx(t, u, v)] +=
                Dc[i][j] * ex.e[ca.i[i]][cb.i[j]][t] * ey.e[ca.j[i]][cb.j[j]][u] * ez.e[ca.k[i]][cb.k[j]][v];
#pragma unroll
  for (int x = 0; x < NH; ++x) h[x] = f * acc[x];
}

__global__ __launch_bounds__(256) void k_jherm(JDev J, const double* __restrict__ D) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int npairs = J.cls[kJMaxLOrb * 2 + 1];
  if (k >= npairs) return;
  int L = 0;
  while (k >= J.cls[L + 1]) ++L;
  double* h = J.h + J.hoff[L] + (long)(k - J.cls[L]) * jnherm(L);
  const int la = l_of(J.ta[k]), lb = L - la;
  switch (la * 3 + lb) {
    case 0: herm_density<0, 0>(J, k, D, h); break;
    case 3: herm_density<1, 0>(J, k, D, h); break;
    case 4: herm_density<1, 1>(J, k, D, h); break;
    case 6: herm_density<2, 0>(J, k, D, h); break;
    case 7: herm_density<2, 1>(J, k, D, h); break;
    default: herm_density<2, 2>(J, k, D, h); break;
  }
}

// ---------------------------------------------------------------------------
// Gamma pass: lane = auxiliary primitive of class LC; the wave walks the
// chunk's pairs (class LAB) in step. acc[e] collects, for each top-degree
// Hermite index e of the primitive, sum_pairs pref sum_tuv h_tuv R_{tuv + e}.
template <int LAB, int LC>
__device__ __forceinline__ void gamma_rows(const JDev& J, int k0, int k1, double q, double cx, double cy, double cz,
                                           double (&acc)[jncart(LC)]) {
  constexpr int NH = jnherm(LAB);
  const CartOf<LC> ce;
  const double* __restrict__ hb = J.h + J.hoff[LAB];
  for (int k = k0; k < k1; ++k) {
    const double p = J.p[k];
    const double* h = hb + (long)(k - J.cls[LAB]) * NH;
    const double pq = p + q, alpha = p * q / pq;
    double R[jnherm(LAB + LC)];
    hermite_r<LAB + LC>(J.boys, alpha, J.px[k] - cx, J.py[k] - cy, J.pz[k] - cz, R);
    const double pref = kTwoPi52 / (p * q * sqrt(pq));
    double hv[NH];
#pragma unroll
    for (int x = 0; x < NH; ++x) hv[x] = h[x];
#pragma unroll
    for (int e = 0; e < jncart(LC); ++e) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t <= LAB; ++t)
#pragma unroll
        for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
          for (int v = 0; v <= LAB - t - u; ++v)
            s += hv[jhidx(t, u, v)] * R[jhidx(t + ce.i[e], u + ce.j[e], v + ce.k[e])];
      acc[e] += pref * s;
    }
  }
}

template <int LC>
__global__ __launch_bounds__(64) void k_jgamma(JDev J) {
  const int slot = J.aoff[LC] + blockIdx.x * 64 + threadIdx.x;  // within the class's padded range
  const int row = blockIdx.y;
  const int4 ch = J.chunks[row];
  const double q = J.ae[slot], cx = J.ax[slot], cy = J.ay[slot], cz = J.az[slot];
  double acc[jncart(LC)];
#pragma unroll
  for (int e = 0; e < jncart(LC); ++e) acc[e] = 0.0;
  switch (ch.z) {
    case 0: gamma_rows<0, LC>(J, ch.x, ch.y, q, cx, cy, cz, acc); break;
    case 1: gamma_rows<1, LC>(J, ch.x, ch.y, q, cx, cy, cz, acc); break;
    case 2: gamma_rows<2, LC>(J, ch.x, ch.y, q, cx, cy, cz, acc); break;
    case 3: gamma_rows<3, LC>(J, ch.x, ch.y, q, cx, cy, cz, acc); break;
    default: gamma_rows<4, LC>(J, ch.x, ch.y, q, cx, cy, cz, acc); break;
  }
  double* out = J.gpart + ((size_t)row * J.nslots + slot) * kJCart;
#pragma unroll
  for (int e = 0; e < jncart(LC); ++e) out[e] = acc[e];
}

// gamma of each contracted auxiliary function f = (shell, m): the sum over its
// primitives of coef sum_e S_lm[e] sum_rows part (fixed order: reproducible).
__global__ void k_jgamma_reduce(JDev J, const int* __restrict__ fn_slot0, const int* __restrict__ fn_nprim,
                                const int* __restrict__ fn_l, const int* __restrict__ fn_m, double* __restrict__ gamma) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= J.naux) return;
  const int l = fn_l[f], m = fn_m[f], s0 = fn_slot0[f], np = fn_nprim[f];
  const double* S = J.auxS + (l * 9 + m) * kJCart;
  double g = 0.0;
  for (int s = s0; s < s0 + np; ++s) {
    double v = 0.0;
    for (int e = 0; e < jncart(l); ++e) {
      double sum = 0.0;
      for (int r = 0; r < J.nrows; ++r) sum += J.gpart[((size_t)r * J.nslots + s) * kJCart + e];
      v += S[e] * sum;
    }
    g += J.acoef[s] * v;
  }
  gamma[f] = g;
}

// Hermite coefficients of the fitted density per auxiliary primitive:
// g[e] = coef sum_m c_(fn+m) S_lm[e].
__global__ void k_jgcoef(JDev J, const int* __restrict__ slot_l, const double* __restrict__ cfit, double* __restrict__ gco) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= J.nslots) return;
  const int l = slot_l[s];
  double* out = gco + (size_t)s * kJCart;
  for (int e = 0; e < kJCart; ++e) out[e] = 0.0;
  if (l < 0) return;  // padding slot
  for (int m = 0; m < 2 * l + 1; ++m) {
    const double c = J.acoef[s] * cfit[J.afn[s] + m];
    const double* S = J.auxS + (l * 9 + m) * kJCart;
    for (int e = 0; e < jncart(l); ++e) out[e] += c * S[e];
  }
}

// ---------------------------------------------------------------------------
// Contraction of a pair's Hermite integrals X_tuv into its Cartesian block,
// C_ij = K sum_tuv Ex Ey Ez X_tuv, then the pure transform and the adds into
// the n x n column-major matrix (both triangles for two different shells).
template <int LA, int LB>
__device__ __forceinline__ void pair_out(const JDev& J, int k, const double (&X)[jnherm(LA + LB)], double scale,
                                         double* __restrict__ M) {
  constexpr int NA = jncart(LA), NB = jncart(LB);
  const double p = J.p[k], o2p = 0.5 / p;
  Herm1<LA, LB> ex, ey, ez;
  ex.build(o2p, J.pax[k], J.pbx[k]);
  ey.build(o2p, J.pay[k], J.pby[k]);
  ez.build(o2p, J.paz[k], J.pbz[k]);
  const CartOf<LA> ca;
  const CartOf<LB> cb;
  double C[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t <= ca.i[i] + cb.i[j]; ++t)
#pragma unroll
        for (int u = 0; u <= ca.j[i] + cb.j[j]; ++u)
#pragma unroll
          for (int v = 0; v <= ca.k[i] + cb.k[j]; ++v)
            s += ex.e[ca.i[i]][cb.i[j]][t] * ey.e[ca.j[i]][cb.j[j]][u] * ez.e[ca.k[i]][cb.k[j]][v] * X[jhidx(t, u, v)];
      C[i][j] = s;
    }
  const int ta = J.ta[k], tb = J.tb[k], aoa = J.aoa[k], aob = J.aob[k], n = J.n;
  const int ma = ncomp_of(ta), mb = ncomp_of(tb);
  const double* Ta = J.orbT + type_index(ta) * 30;
  const double* Tb = J.orbT + type_index(tb) * 30;
  const double f = scale * J.K[k];
  const bool both = !J.diag[k];
  for (int a = 0; a < ma; ++a)
    for (int b = 0; b < mb; ++b) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) s += Ta[a * 6 + i] * C[i][j] * Tb[b * 6 + j];
      s *= f;
      unsafeAtomicAdd(&M[(size_t)(aob + b) * n + (aoa + a)], s);
      if (both) unsafeAtomicAdd(&M[(size_t)(aoa + a) * n + (aob + b)], s);
    }
}

template <int LAB>
__device__ __forceinline__ void pair_out_any(const JDev& J, int k, const double (&X)[jnherm(LAB)], double scale,
                                             double* M) {
  const int la = l_of(J.ta[k]);
  if constexpr (LAB == 0) {
    pair_out<0, 0>(J, k, X, scale, M);
  } else if constexpr (LAB == 1) {
    pair_out<1, 0>(J, k, X, scale, M);
  } else if constexpr (LAB == 2) {
    if (la == 2) pair_out<2, 0>(J, k, X, scale, M);
    else pair_out<1, 1>(J, k, X, scale, M);
  } else if constexpr (LAB == 3) {
    pair_out<2, 1>(J, k, X, scale, M);
  } else {
    pair_out<2, 2>(J, k, X, scale, M);
  }
}

// Coulomb pass: lane = pair of class LAB; the wave walks every auxiliary
// primitive in step (its data uniform: scalar loads), X_tuv += pref sum_e g_e R_{tuv + e}.
template <int LAB, int LC>
__device__ __forceinline__ void coulomb_aux(const JDev& J, double p, double px, double py, double pz,
                                            double (&X)[jnherm(LAB)]) {
  const CartOf<LC> ce;
  const int s1 = J.aoff[LC] + J.anum[LC];
  for (int s = J.aoff[LC]; s < s1; ++s) {
    const double q = J.ae[s], pq = p + q, alpha = p * q / pq;
    double R[jnherm(LAB + LC)];
    hermite_r<LAB + LC>(J.boys, alpha, px - J.ax[s], py - J.ay[s], pz - J.az[s], R);
    const double pref = kTwoPi52 / (p * q * sqrt(pq));
    const double* g = J.gco + (size_t)s * kJCart;
#pragma unroll
    for (int e = 0; e < jncart(LC); ++e) {
      const double ge = pref * g[e];
#pragma unroll
      for (int t = 0; t <= LAB; ++t)
#pragma unroll
        for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
          for (int v = 0; v <= LAB - t - u; ++v)
            X[jhidx(t, u, v)] += ge * R[jhidx(t + ce.i[e], u + ce.j[e], v + ce.k[e])];
    }
  }
}

template <int LAB>
__global__ __launch_bounds__(256) void k_jcoulomb(JDev J, double* __restrict__ Jm) {
  const int c0 = J.cls[LAB], c1 = J.cls[LAB + 1];
  const int k = c0 + blockIdx.x * 256 + threadIdx.x;
  const bool live = k < c1;
  const int kk = live ? k : c0;
  const double p = J.p[kk], px = J.px[kk], py = J.py[kk], pz = J.pz[kk];
  double X[jnherm(LAB)];
#pragma unroll
  for (int x = 0; x < jnherm(LAB); ++x) X[x] = 0.0;
  coulomb_aux<LAB, 0>(J, p, px, py, pz, X);
  coulomb_aux<LAB, 1>(J, p, px, py, pz, X);
  coulomb_aux<LAB, 2>(J, p, px, py, pz, X);
  coulomb_aux<LAB, 3>(J, p, px, py, pz, X);
  coulomb_aux<LAB, 4>(J, p, px, py, pz, X);
  if (live) pair_out_any<LAB>(J, k, X, 1.0, Jm);
}

// ---------------------------------------------------------------------------
// One-electron integrals per pair: overlap and kinetic energy from the 1-D
// overlaps S_ij = E^ij_0 sqrt(pi / p) (kinetic: -1/2 [j (j-1) S_{i,j-2} -
// 2b (2j+1) S_ij + 4 b^2 S_{i,j+2}] per direction), nuclear attraction from
// X_tuv = sum_C -Z_C (2 pi / p) R_tuv(p, P - C).
template <int LA, int LB>
__device__ __forceinline__ void onee_pair(const JDev& J, int k, double b_exp, double* S, double* T, double* V) {
  constexpr int NA = jncart(LA), NB = jncart(LB), NH = jnherm(LA + LB);
  const double p = J.p[k], o2p = 0.5 / p;
  Herm1<LA, LB + 2> ex, ey, ez;
  ex.build(o2p, J.pax[k], J.pbx[k]);
  ey.build(o2p, J.pay[k], J.pby[k]);
  ez.build(o2p, J.paz[k], J.pbz[k]);
  const double r = sqrt(kPi / p);
  const CartOf<LA> ca;
  const CartOf<LB> cb;
  double X[NH];
#pragma unroll
  for (int x = 0; x < NH; ++x) X[x] = 0.0;
  const double px = J.px[k], py = J.py[k], pz = J.pz[k];
  for (int c = 0; c < J.natoms; ++c) {
    const double* nc = J.nuc + 4 * c;
    double R[NH];
    hermite_r<LA + LB>(J.boys, p, px - nc[0], py - nc[1], pz - nc[2], R);
    const double z = -nc[3] * 2.0 * kPi / p;
#pragma unroll
    for (int x = 0; x < NH; ++x) X[x] += z * R[x];
  }
  double Cs[NA][NB], Ct[NA][NB], Cv[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int ix = ca.i[i], iy = ca.j[i], iz = ca.k[i], jx = cb.i[j], jy = cb.j[j], jz = cb.k[j];
      auto s1 = [&](const Herm1<LA, LB + 2>& e, int a, int bb) { return bb < 0 ? 0.0 : e.e[a][bb][0] * r; };
      auto t1 = [&](const Herm1<LA, LB + 2>& e, int a, int bb) {
        return -0.5 * (bb * (bb - 1) * s1(e, a, bb - 2) - 2.0 * b_exp * (2 * bb + 1) * s1(e, a, bb) +
                       4.0 * b_exp * b_exp * s1(e, a, bb + 2));
      };
      const double sx = s1(ex, ix, jx), sy = s1(ey, iy, jy), sz = s1(ez, iz, jz);
      Cs[i][j] = sx * sy * sz;
      Ct[i][j] = t1(ex, ix, jx) * sy * sz + sx * t1(ey, iy, jy) * sz + sx * sy * t1(ez, iz, jz);
      double s = 0.0;
#pragma unroll
      for (int t = 0; t <= ix + jx; ++t)
#pragma unroll
        for (int u = 0; u <= iy + jy; ++u)
#pragma unroll
          for (int v = 0; v <= iz + jz; ++v) s += ex.e[ix][jx][t] * ey.e[iy][jy][u] * ez.e[iz][jz][v] * X[jhidx(t, u, v)];
      Cv[i][j] = s;
    }
  const int ta = J.ta[k], tb = J.tb[k], aoa = J.aoa[k], aob = J.aob[k], n = J.n;
  const int ma = ncomp_of(ta), mb = ncomp_of(tb);
  const double* Ta = J.orbT + type_index(ta) * 30;
  const double* Tb = J.orbT + type_index(tb) * 30;
  const double f = J.K[k];
  const bool both = !J.diag[k];
  for (int a = 0; a < ma; ++a)
    for (int b = 0; b < mb; ++b) {
      double s = 0.0, t = 0.0, v = 0.0;
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const double w = Ta[a * 6 + i] * Tb[b * 6 + j];
          s += w * Cs[i][j];
          t += w * Ct[i][j];
          v += w * Cv[i][j];
        }
      const size_t i0 = (size_t)(aob + b) * n + (aoa + a), i1 = (size_t)(aoa + a) * n + (aob + b);
      unsafeAtomicAdd(&S[i0], f * s);
      unsafeAtomicAdd(&T[i0], f * t);
      unsafeAtomicAdd(&V[i0], f * v);
      if (both) {
        unsafeAtomicAdd(&S[i1], f * s);
        unsafeAtomicAdd(&T[i1], f * t);
        unsafeAtomicAdd(&V[i1], f * v);
      }
    }
}

__global__ __launch_bounds__(256) void k_joneint(JDev J, const double* __restrict__ bexp, double* S, double* T,
                                                 double* V) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int npairs = J.cls[kJMaxLOrb * 2 + 1];
  if (k >= npairs) return;
  int L = 0;
  while (k >= J.cls[L + 1]) ++L;
  const int la = l_of(J.ta[k]), lb = L - la;
  const double b = bexp[k];
  switch (la * 3 + lb) {
    case 0: onee_pair<0, 0>(J, k, b, S, T, V); break;
    case 3: onee_pair<1, 0>(J, k, b, S, T, V); break;
    case 4: onee_pair<1, 1>(J, k, b, S, T, V); break;
    case 6: onee_pair<2, 0>(J, k, b, S, T, V); break;
    case 7: onee_pair<2, 1>(J, k, b, S, T, V); break;
    default: onee_pair<2, 2>(J, k, b, S, T, V); break;
  }
}

// ---------------------------------------------------------------------------
// Coulomb metric (P|Q) between contracted auxiliary functions: lane = P
// primitive (class LP), the wave walks a range of Q primitives (class LQ);
// each (P, Q) primitive pair adds its (2LP+1) x (2LQ+1) block.
template <int LP, int LQ>
__device__ __forceinline__ void metric_rows(const JDev& J, int slotP, bool live, int q0, int q1, double* __restrict__ Vm) {
  const CartOf<LP> cp;
  const CartOf<LQ> cq;
  const double a = J.ae[slotP], ax = J.ax[slotP], ay = J.ay[slotP], az = J.az[slotP];
  // (2 a)^-l (-1)^l c is in acoef; the metric's sign factor is (-1)^LQ
  const double cP = J.acoef[slotP] * ((LP & 1) ? -1.0 : 1.0);
  const int fP = J.afn[slotP];
  const double* SP = J.auxS + LP * 9 * kJCart;
  const double* SQ = J.auxS + LQ * 9 * kJCart;
  for (int s = q0; s < q1; ++s) {
    const double b = J.ae[s], ab = a + b, alpha = a * b / ab;
    double R[jnherm(LP + LQ)];
    hermite_r<LP + LQ>(J.boys, alpha, ax - J.ax[s], ay - J.ay[s], az - J.az[s], R);
    const double pref = kTwoPi52 / (a * b * sqrt(ab)) * cP * J.acoef[s];
    const int fQ = J.afn[s];
    double W[jncart(LP)][jncart(LQ)];
#pragma unroll
    for (int i = 0; i < jncart(LP); ++i)
#pragma unroll
      for (int j = 0; j < jncart(LQ); ++j) W[i][j] = R[jhidx(cp.i[i] + cq.i[j], cp.j[i] + cq.j[j], cp.k[i] + cq.k[j])];
    if (!live) continue;
    for (int mP = 0; mP < 2 * LP + 1; ++mP)
      for (int mQ = 0; mQ < 2 * LQ + 1; ++mQ) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < jncart(LP); ++i)
#pragma unroll
          for (int j = 0; j < jncart(LQ); ++j) v += SP[mP * kJCart + i] * W[i][j] * SQ[mQ * kJCart + j];
        unsafeAtomicAdd(&Vm[(size_t)(fQ + mQ) * J.naux + (fP + mP)], pref * v);
      }
  }
}

template <int LP>
__global__ __launch_bounds__(64) void k_jmetric(JDev J, double* Vm) {
  const int slot = J.aoff[LP] + blockIdx.x * 64 + threadIdx.x;
  const bool live = slot < J.aoff[LP] + J.anum[LP];
  const int sP = live ? slot : J.aoff[LP];
  // Q range: blockIdx.y of gridDim.y equal parts of every class
#pragma unroll
  for (int LQ = 0; LQ <= kJMaxLAux; ++LQ) {
    const int n = J.anum[LQ], per = (n + gridDim.y - 1) / gridDim.y;
    const int q0 = J.aoff[LQ] + min(n, (int)blockIdx.y * per), q1 = J.aoff[LQ] + min(n, (int)(blockIdx.y + 1) * per);
    switch (LQ) {
      case 0: metric_rows<LP, 0>(J, sP, live, q0, q1, Vm); break;
      case 1: metric_rows<LP, 1>(J, sP, live, q0, q1, Vm); break;
      case 2: metric_rows<LP, 2>(J, sP, live, q0, q1, Vm); break;
      case 3: metric_rows<LP, 3>(J, sP, live, q0, q1, Vm); break;
      default: metric_rows<LP, 4>(J, sP, live, q0, q1, Vm); break;
    }
  }
}

}  // namespace

// ---- launchers --------------------------------------------------------------
void launch_jherm(hipStream_t s, const JDev& J, const double* D) {
  const int np = J.cls[kJMaxLOrb * 2 + 1];
  if (np == 0) return;
  hipLaunchKernelGGL(k_jherm, dim3((np + 255) / 256), dim3(256), 0, s, J, D);
  CHX_HIP(hipGetLastError());
}

void launch_jgamma(hipStream_t s, const JDev& J, const int* fn_slot0, const int* fn_nprim, const int* fn_l,
                   const int* fn_m, double* gamma) {
  for (int LC = 0; LC <= kJMaxLAux; ++LC) {
    const int waves = (J.aoff[LC + 1] - J.aoff[LC]) / 64;
    if (waves == 0 || J.nrows == 0) continue;
    const dim3 grid(waves, J.nrows);
    switch (LC) {
      case 0: hipLaunchKernelGGL(k_jgamma<0>, grid, dim3(64), 0, s, J); break;
      case 1: hipLaunchKernelGGL(k_jgamma<1>, grid, dim3(64), 0, s, J); break;
      case 2: hipLaunchKernelGGL(k_jgamma<2>, grid, dim3(64), 0, s, J); break;
      case 3: hipLaunchKernelGGL(k_jgamma<3>, grid, dim3(64), 0, s, J); break;
      default: hipLaunchKernelGGL(k_jgamma<4>, grid, dim3(64), 0, s, J); break;
    }
    CHX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_jgamma_reduce, dim3((J.naux + 255) / 256), dim3(256), 0, s, J, fn_slot0, fn_nprim, fn_l, fn_m,
                     gamma);
  CHX_HIP(hipGetLastError());
}

void launch_jcoulomb(hipStream_t s, const JDev& J, const int* slot_l, const double* cfit, double* gco, double* Jm) {
  hipLaunchKernelGGL(k_jgcoef, dim3((J.nslots + 255) / 256), dim3(256), 0, s, J, slot_l, cfit, gco);
  CHX_HIP(hipGetLastError());
  for (int L = 0; L <= 2 * kJMaxLOrb; ++L) {
    const int np = J.cls[L + 1] - J.cls[L];
    if (np == 0) continue;
    const dim3 grid((np + 255) / 256);
    switch (L) {
      case 0: hipLaunchKernelGGL(k_jcoulomb<0>, grid, dim3(256), 0, s, J, Jm); break;
      case 1: hipLaunchKernelGGL(k_jcoulomb<1>, grid, dim3(256), 0, s, J, Jm); break;
      case 2: hipLaunchKernelGGL(k_jcoulomb<2>, grid, dim3(256), 0, s, J, Jm); break;
      case 3: hipLaunchKernelGGL(k_jcoulomb<3>, grid, dim3(256), 0, s, J, Jm); break;
      default: hipLaunchKernelGGL(k_jcoulomb<4>, grid, dim3(256), 0, s, J, Jm); break;
    }
    CHX_HIP(hipGetLastError());
  }
}

void launch_joneint(hipStream_t s, const JDev& J, const double* bexp, double* S, double* T, double* V) {
  const int np = J.cls[kJMaxLOrb * 2 + 1];
  if (np == 0) return;
  hipLaunchKernelGGL(k_joneint, dim3((np + 255) / 256), dim3(256), 0, s, J, bexp, S, T, V);
  CHX_HIP(hipGetLastError());
}

void launch_jmetric(hipStream_t s, const JDev& J, double* Vm) {
  for (int LP = 0; LP <= kJMaxLAux; ++LP) {
    const int waves = (J.aoff[LP + 1] - J.aoff[LP]) / 64;
    if (waves == 0) continue;
    const dim3 grid(waves, 16);
    switch (LP) {
      case 0: hipLaunchKernelGGL(k_jmetric<0>, grid, dim3(64), 0, s, J, Vm); break;
      case 1: hipLaunchKernelGGL(k_jmetric<1>, grid, dim3(64), 0, s, J, Vm); break;
      case 2: hipLaunchKernelGGL(k_jmetric<2>, grid, dim3(64), 0, s, J, Vm); break;
      case 3: hipLaunchKernelGGL(k_jmetric<3>, grid, dim3(64), 0, s, J, Vm); break;
      default: hipLaunchKernelGGL(k_jmetric<4>, grid, dim3(64), 0, s, J, Vm); break;
    }
    CHX_HIP(hipGetLastError());
  }
}

}  // namespace chx
