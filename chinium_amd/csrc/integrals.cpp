// One- and two-electron integrals over the molecule's AOs, for the SCF loop
// around the XC path on small molecules (host code, McMurchie-Davidson).
//
// The reference takes these from libint2: Int2C1E (overlap, kinetic, nuclear
// attraction; src/Integral/Representation.cpp:62-63) and Int4C2E, whose
// ContractInts (src/Integral/Int4C2E.cpp:673-683) returns J and EXX-scaled K
// for the restricted update (src/HartreeFockKohnSham/Restricted/SP.cpp:43-56).
// Each AO is written exactly as a sum of Cartesian primitives
// c x^i y^j z^k exp(-a r^2) about its centre: the normalised contraction
// coefficients of the shell times the monomials of its component polynomial
// (s, Cartesian p, pure p/d/f in the reference's component order, as the grid
// kernels evaluate them), so these integrals and the XC path share one AO
// convention. Products of primitives are expanded in Hermite Gaussians
// (E coefficients); Coulomb-type integrals use the Hermite integrals R_tuv
// over the Boys function (McMurchie, Davidson, J. Comput. Phys. 26, 218
// (1978); Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure Theory,
// ch. 9). The full ERI tensor is formed, so this is for small bases.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace chx {
std::string& last_error();
}
using namespace chx;

namespace {

constexpr int kMaxL = 3;                // f
constexpr int kMaxHerm = 4 * kMaxL + 1;  // t + u + v of an ERI
constexpr long kMaxEriBasis = 96;        // n^4 doubles are formed

struct Prim {
  double a, c;
  int l[3];
  const double* A;
};

struct Mono {
  double f;
  int l[3];
};

// Monomials of each component polynomial, in the component order of the grid
// kernels (src/Grid/AO/{S,CartP,PureP,PureD,PureF}.hpp).
std::vector<std::vector<Mono>> monomials(int type) {
  const double r3 = std::sqrt(3.0), r6 = std::sqrt(6.0), r10 = std::sqrt(10.0), r15 = std::sqrt(15.0);
  const double c10 = 0.25 * r10, c6 = 0.25 * r6;
  switch (type) {
    case 0: return {{{1, {0, 0, 0}}}};
    case 1: return {{{1, {1, 0, 0}}}, {{1, {0, 1, 0}}}, {{1, {0, 0, 1}}}};
    case -1: return {{{1, {0, 1, 0}}}, {{1, {0, 0, 1}}}, {{1, {1, 0, 0}}}};
    case -2:
      return {{{r3, {1, 1, 0}}},
              {{r3, {0, 1, 1}}},
              {{1, {0, 0, 2}}, {-0.5, {2, 0, 0}}, {-0.5, {0, 2, 0}}},
              {{r3, {1, 0, 1}}},
              {{0.5 * r3, {2, 0, 0}}, {-0.5 * r3, {0, 2, 0}}}};
    case -3:
      return {{{3 * c10, {2, 1, 0}}, {-c10, {0, 3, 0}}},
              {{r15, {1, 1, 1}}},
              {{4 * c6, {0, 1, 2}}, {-c6, {2, 1, 0}}, {-c6, {0, 3, 0}}},
              {{1, {0, 0, 3}}, {-1.5, {2, 0, 1}}, {-1.5, {0, 2, 1}}},
              {{4 * c6, {1, 0, 2}}, {-c6, {3, 0, 0}}, {-c6, {1, 2, 0}}},
              {{0.5 * r15, {2, 0, 1}}, {-0.5 * r15, {0, 2, 1}}},
              {{c10, {3, 0, 0}}, {-3 * c10, {1, 2, 0}}}};
    default: throw Error(CHX_ERR_UNSUPPORTED, "integrals are built for s, p, d and f shells");
  }
}

std::vector<std::vector<Prim>> ao_prims(const chx_mol* m) {
  std::vector<std::vector<Prim>> ao(m->nbasis);
  for (const ShellH& sh : m->shells) {
    const auto q = monomials(sh.type);
    for (size_t k = 0; k < q.size(); ++k)
      for (int p = 0; p < sh.nprim; ++p)
        for (const Mono& t : q[k])
          ao[sh.ao_off + k].push_back(Prim{m->exps[sh.prim_off + p], m->coefs[sh.prim_off + p] * t.f,
                                           {t.l[0], t.l[1], t.l[2]}, &m->xyz[3 * sh.center]});
  }
  return ao;
}

// Hermite expansion coefficients E[i][j][t] of x_A^i x_B^j exp(-a x_A^2 - b x_B^2)
// = sum_t E_t Lambda_t(x_P), for i <= imax, j <= jmax (X = A - B).
constexpr int kE = kMaxL + 3;
struct Herm {
  double e[kE][kE][2 * kE];
};

void hermite(int imax, int jmax, double a, double b, double X, Herm& H) {
  const double p = a + b, mu = a * b / p, xpa = -b / p * X, xpb = a / p * X, h = 0.5 / p;
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j)
      for (int t = 0; t <= i + j + 1; ++t) H.e[i][j][t] = 0.0;
  H.e[0][0][0] = std::exp(-mu * X * X);
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        const double (*s)[2 * kE] = i > 0 ? H.e[i - 1] : nullptr;
        if (i > 0) {
          const double* prev = s[j];
          H.e[i][j][t] = (t > 0 ? h * prev[t - 1] : 0.0) + xpa * prev[t] + (t + 1 <= i - 1 + j ? (t + 1) * prev[t + 1] : 0.0);
        } else {
          const double* prev = H.e[0][j - 1];
          H.e[0][j][t] = (t > 0 ? h * prev[t - 1] : 0.0) + xpb * prev[t] + (t + 1 <= j - 1 ? (t + 1) * prev[t + 1] : 0.0);
        }
      }
    }
}

// Boys function F_n(T), n = 0..nmax: the series for F_nmax and downward
// recursion, or the asymptotic form where exp(-T) is below rounding.
void boys(int nmax, double T, double* F) {
  if (T < 1e-15) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1);
    return;
  }
  if (T > 36.0) {
    F[0] = 0.5 * std::sqrt(M_PI / T);
    for (int n = 1; n <= nmax; ++n) F[n] = F[n - 1] * (2 * n - 1) / (2 * T);
    return;
  }
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int k = 1; k < 400; ++k) {
    term *= 2 * T / (2 * nmax + 2 * k + 1);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  const double ex = std::exp(-T);
  F[nmax] = ex * sum;
  for (int n = nmax - 1; n >= 0; --n) F[n] = (2 * T * F[n + 1] + ex) / (2 * n + 1);
}

// Hermite Coulomb integrals R^0_tuv(alpha, X, Y, Z) for t + u + v <= L.
struct RTab {
  int L = 0;
  std::vector<double> r;  // [n][t][u][v], each index 0..L
  double& at(int n, int t, int u, int v) { return r[(((size_t)n * (L + 1) + t) * (L + 1) + u) * (L + 1) + v]; }
};

void hermite_coulomb(int L, double alpha, double X, double Y, double Z, RTab& R) {
  R.L = L;
  R.r.assign((size_t)(L + 1) * (L + 1) * (L + 1) * (L + 1), 0.0);
  double F[kMaxHerm + 3];
  boys(L, alpha * (X * X + Y * Y + Z * Z), F);
  double m2a = 1.0;
  for (int n = 0; n <= L; ++n) {
    R.at(n, 0, 0, 0) = m2a * F[n];
    m2a *= -2 * alpha;
  }
  for (int n = L - 1; n >= 0; --n)
    for (int s = 1; s <= L - n; ++s)
      for (int t = 0; t <= s; ++t)
        for (int u = 0; u <= s - t; ++u) {
          const int v = s - t - u;
          double val;
          if (t > 0)
            val = (t > 1 ? (t - 1) * R.at(n + 1, t - 2, u, v) : 0.0) + X * R.at(n + 1, t - 1, u, v);
          else if (u > 0)
            val = (u > 1 ? (u - 1) * R.at(n + 1, t, u - 2, v) : 0.0) + Y * R.at(n + 1, t, u - 1, v);
          else
            val = (v > 1 ? (v - 1) * R.at(n + 1, t, u, v - 2) : 0.0) + Z * R.at(n + 1, t, u, v - 1);
          R.at(n, t, u, v) = val;
        }
}

// Product of two primitives in Hermite form: p, P and the nonzero
// E^x_t E^y_u E^z_v (times both coefficients).
struct PairTerm {
  double p, P[3];
  int L;
  std::vector<std::array<int, 3>> tuv;
  std::vector<double> e;
};

PairTerm pair_term(const Prim& A, const Prim& B) {
  PairTerm pt;
  pt.p = A.a + B.a;
  Herm H[3];
  for (int d = 0; d < 3; ++d) {
    pt.P[d] = (A.a * A.A[d] + B.a * B.A[d]) / pt.p;
    hermite(A.l[d], B.l[d], A.a, B.a, A.A[d] - B.A[d], H[d]);
  }
  const int lx = A.l[0] + B.l[0], ly = A.l[1] + B.l[1], lz = A.l[2] + B.l[2];
  pt.L = lx + ly + lz;
  for (int t = 0; t <= lx; ++t)
    for (int u = 0; u <= ly; ++u)
      for (int v = 0; v <= lz; ++v) {
        const double e = A.c * B.c * H[0].e[A.l[0]][B.l[0]][t] * H[1].e[A.l[1]][B.l[1]][u] *
                         H[2].e[A.l[2]][B.l[2]][v];
        if (e != 0.0) {
          pt.tuv.push_back({t, u, v});
          pt.e.push_back(e);
        }
      }
  return pt;
}

template <class F>
int iguard(F&& f) {
  try {
    f();
    return CHX_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return CHX_ERR_ARG;
  }
}

}  // namespace

extern "C" {

int chx_mol_nuclear_repulsion(const chx_mol* m, double* E) {
  return iguard([&] {
    if (!m || !E) throw Error(CHX_ERR_ARG, "bad arguments");
    double e = 0;
    for (int a = 0; a < m->natoms; ++a)
      for (int b = 0; b < a; ++b) {
        const double dx = m->xyz[3 * a] - m->xyz[3 * b], dy = m->xyz[3 * a + 1] - m->xyz[3 * b + 1],
                     dz = m->xyz[3 * a + 2] - m->xyz[3 * b + 2];
        e += m->Z[a] * m->Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    *E = e;
  });
}

int chx_mol_int1e(const chx_mol* m, double* S, double* T, double* V) {
  return iguard([&] {
    if (!m || !S || !T || !V) throw Error(CHX_ERR_ARG, "bad arguments");
    if (m->nbasis == 0) throw Error(CHX_ERR_STATE, "the molecule has no basis set");
    const int n = m->nbasis;
    const auto ao = ao_prims(m);
    RTab R;
    for (int mu = 0; mu < n; ++mu)
      for (int nu = 0; nu <= mu; ++nu) {
        double s = 0, t = 0, v = 0;
        for (const Prim& A : ao[mu])
          for (const Prim& B : ao[nu]) {
            const double p = A.a + B.a;
            Herm H[3];
            double sd[3], td[3];
            for (int d = 0; d < 3; ++d) {
              hermite(A.l[d], B.l[d] + 2, A.a, B.a, A.A[d] - B.A[d], H[d]);
              const int i = A.l[d], j = B.l[d];
              const double r = std::sqrt(M_PI / p);
              auto s1 = [&](int jj) { return jj < 0 ? 0.0 : H[d].e[i][jj][0] * r; };
              sd[d] = s1(j);
              td[d] = -0.5 * (j * (j - 1) * s1(j - 2) - 2 * B.a * (2 * j + 1) * s1(j) + 4 * B.a * B.a * s1(j + 2));
            }
            const double cc = A.c * B.c;
            s += cc * sd[0] * sd[1] * sd[2];
            t += cc * (td[0] * sd[1] * sd[2] + sd[0] * td[1] * sd[2] + sd[0] * sd[1] * td[2]);
            const PairTerm pt = pair_term(A, B);
            for (int c = 0; c < m->natoms; ++c) {
              const double* C = &m->xyz[3 * c];
              hermite_coulomb(pt.L, pt.p, pt.P[0] - C[0], pt.P[1] - C[1], pt.P[2] - C[2], R);
              double acc = 0;
              for (size_t k = 0; k < pt.e.size(); ++k) acc += pt.e[k] * R.at(0, pt.tuv[k][0], pt.tuv[k][1], pt.tuv[k][2]);
              v -= m->Z[c] * 2 * M_PI / pt.p * acc;
            }
          }
        S[(size_t)mu * n + nu] = S[(size_t)nu * n + mu] = s;
        T[(size_t)mu * n + nu] = T[(size_t)nu * n + mu] = t;
        V[(size_t)mu * n + nu] = V[(size_t)nu * n + mu] = v;
      }
  });
}

int chx_mol_eri(const chx_mol* m, double* eri) {
  return iguard([&] {
    if (!m || !eri) throw Error(CHX_ERR_ARG, "bad arguments");
    if (m->nbasis == 0) throw Error(CHX_ERR_STATE, "the molecule has no basis set");
    const long n = m->nbasis;
    if (n > kMaxEriBasis)
      throw Error(CHX_ERR_UNSUPPORTED, "the full ERI tensor is formed only up to 96 basis functions");
    const auto ao = ao_prims(m);
    // AO pairs mu >= nu in Hermite form
    std::vector<std::vector<PairTerm>> pairs((size_t)n * (n + 1) / 2);
    for (long mu = 0; mu < n; ++mu)
      for (long nu = 0; nu <= mu; ++nu) {
        auto& v = pairs[(size_t)mu * (mu + 1) / 2 + nu];
        for (const Prim& A : ao[mu])
          for (const Prim& B : ao[nu]) v.push_back(pair_term(A, B));
      }
    const double pre = 2 * std::pow(M_PI, 2.5);
    const long np = (long)pairs.size();
    // rows ij of the unique quartets (ij >= kl) handed out to host threads;
    // each quartet writes its own eight entries, so the threads never collide
    std::atomic<long> next{0};
    auto worker = [&] {
      RTab R;
      for (long ij = next++; ij < np; ij = next++)
        for (long kl = 0; kl <= ij; ++kl) {
        double val = 0;
        for (const PairTerm& a : pairs[ij])
          for (const PairTerm& b : pairs[kl]) {
            const double alpha = a.p * b.p / (a.p + b.p);
            hermite_coulomb(a.L + b.L, alpha, a.P[0] - b.P[0], a.P[1] - b.P[1], a.P[2] - b.P[2], R);
            double acc = 0;
            for (size_t x = 0; x < a.e.size(); ++x)
              for (size_t y = 0; y < b.e.size(); ++y) {
                const auto& s = a.tuv[x];
                const auto& q = b.tuv[y];
                const double sg = ((q[0] + q[1] + q[2]) & 1) ? -1.0 : 1.0;
                acc += a.e[x] * b.e[y] * sg * R.at(0, s[0] + q[0], s[1] + q[1], s[2] + q[2]);
              }
            val += pre / (a.p * b.p * std::sqrt(a.p + b.p)) * acc;
          }
        long i = 0;
        while ((i + 1) * (i + 2) / 2 <= ij) ++i;
        const long j = ij - i * (i + 1) / 2;
        long k = 0;
        while ((k + 1) * (k + 2) / 2 <= kl) ++k;
        const long l = kl - k * (k + 1) / 2;
        const long idx[8][4] = {{i, j, k, l}, {j, i, k, l}, {i, j, l, k}, {j, i, l, k},
                                {k, l, i, j}, {l, k, i, j}, {k, l, j, i}, {l, k, j, i}};
        for (const auto& q : idx) eri[q[0] + n * (q[1] + n * (q[2] + n * q[3]))] = val;
      }
    };
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; ++t) pool.emplace_back(worker);
    worker();
    for (std::thread& t : pool) t.join();
  });
}

}  // extern "C"
