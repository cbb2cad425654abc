// FP64 CPU oracle: restatement of the reference's XC path, loop nest by loop nest.
// TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled like the reference
// (makefile:41-42: -O3 -fopenmp -march=native) so it doubles as the CPU baseline.
#include "oracle.h"
#include "functionals.hpp"

#include <omp.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

const char* kElem[] = {"X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
                       "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni",
                       "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
const int kNElem = 36;

struct Shell {
  int type;      // 0 s, 1 Cartesian p, -l pure
  int center;
  std::vector<double> exps, coefs;  // coefs normalised
};

// One angular function as a polynomial: sum of c * x^i y^j z^k.
struct Mono { double c; int i, j, k; };
typedef std::vector<Mono> Poly;

std::vector<Poly> angular_functions(int type) {
  const double s3 = std::sqrt(3.0), s6 = std::sqrt(6.0), s10 = std::sqrt(10.0), s15 = std::sqrt(15.0);
  switch (type) {
    case 0: return {{{1, 0, 0, 0}}};
    case 1: return {{{1, 1, 0, 0}}, {{1, 0, 1, 0}}, {{1, 0, 0, 1}}};                 // AO/CartP.hpp
    case -1: return {{{1, 0, 1, 0}}, {{1, 0, 0, 1}}, {{1, 1, 0, 0}}};                // AO/PureP.hpp (y,z,x)
    case -2: return {                                                                 // AO/PureD.hpp
        {{s3, 1, 1, 0}},
        {{s3, 0, 1, 1}},
        {{1, 0, 0, 2}, {-0.5, 2, 0, 0}, {-0.5, 0, 2, 0}},
        {{s3, 1, 0, 1}},
        {{s3 / 2, 2, 0, 0}, {-s3 / 2, 0, 2, 0}}};
    case -3: return {                                                                 // AO/PureF.hpp
        {{3 * s10 / 4, 2, 1, 0}, {-s10 / 4, 0, 3, 0}},
        {{s15, 1, 1, 1}},
        {{s6, 0, 1, 2}, {-s6 / 4, 2, 1, 0}, {-s6 / 4, 0, 3, 0}},
        {{1, 0, 0, 3}, {-1.5, 2, 0, 1}, {-1.5, 0, 2, 1}},
        {{s6, 1, 0, 2}, {-s6 / 4, 3, 0, 0}, {-s6 / 4, 1, 2, 0}},
        {{s15 / 2, 2, 0, 1}, {-s15 / 2, 0, 2, 1}},
        {{s10 / 4, 3, 0, 0}, {-3 * s10 / 4, 1, 2, 0}}};
    default: throw std::runtime_error("oracle: angular momentum above f is not supported");
  }
}

inline double ipow(double x, int n) { double r = 1; for (int i = 0; i < n; ++i) r *= x; return r; }

// Value, gradient and second derivatives (xx,yy,zz) of a polynomial.
void poly_eval(const Poly& p, double x, double y, double z, double* q, double* qg, double* qll) {
  double v = 0, gx = 0, gy = 0, gz = 0, lxx = 0, lyy = 0, lzz = 0;
  for (const Mono& m : p) {
    const double xi = ipow(x, m.i), yj = ipow(y, m.j), zk = ipow(z, m.k);
    v += m.c * xi * yj * zk;
    if (m.i > 0) gx += m.c * m.i * ipow(x, m.i - 1) * yj * zk;
    if (m.j > 0) gy += m.c * m.j * xi * ipow(y, m.j - 1) * zk;
    if (m.k > 0) gz += m.c * m.k * xi * yj * ipow(z, m.k - 1);
    if (m.i > 1) lxx += m.c * m.i * (m.i - 1) * ipow(x, m.i - 2) * yj * zk;
    if (m.j > 1) lyy += m.c * m.j * (m.j - 1) * xi * ipow(y, m.j - 2) * zk;
    if (m.k > 1) lzz += m.c * m.k * (m.k - 1) * xi * yj * ipow(z, m.k - 2);
  }
  *q = v; qg[0] = gx; qg[1] = gy; qg[2] = gz; *qll = lxx + lyy + lzz;
}

// Value, gradient and Hessian (xx, xy, xz, yy, yz, zz) of a polynomial.
void poly_eval_h(const Poly& p, double x, double y, double z, double* q, double* qg, double* qh) {
  double v = 0, g[3] = {0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0};
  auto pw = [](double b, int n) { return n < 0 ? 0.0 : ipow(b, n); };
  for (const Mono& m : p) {
    const int e[3] = {m.i, m.j, m.k};
    const double b[3] = {x, y, z};
    double f[3], d1[3], d2[3];
    for (int r = 0; r < 3; ++r) {
      f[r] = pw(b[r], e[r]);
      d1[r] = e[r] * pw(b[r], e[r] - 1);
      d2[r] = e[r] * (e[r] - 1) * pw(b[r], e[r] - 2);
    }
    v += m.c * f[0] * f[1] * f[2];
    g[0] += m.c * d1[0] * f[1] * f[2];
    g[1] += m.c * f[0] * d1[1] * f[2];
    g[2] += m.c * f[0] * f[1] * d1[2];
    h[0] += m.c * d2[0] * f[1] * f[2];
    h[1] += m.c * d1[0] * d1[1] * f[2];
    h[2] += m.c * d1[0] * f[1] * d1[2];
    h[3] += m.c * f[0] * d2[1] * f[2];
    h[4] += m.c * f[0] * d1[1] * d1[2];
    h[5] += m.c * f[0] * f[1] * d2[2];
  }
  *q = v;
  for (int r = 0; r < 3; ++r) qg[r] = g[r];
  for (int k = 0; k < 6; ++k) qh[k] = h[k];
}

// Third derivatives of a polynomial, (xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz).
void poly_eval_3(const Poly& p, double x, double y, double z, double* q3) {
  static const int idx[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                                 {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
  const double b[3] = {x, y, z};
  for (int k = 0; k < 10; ++k) {
    double v = 0;
    for (const Mono& m : p) {
      int e[3] = {m.i, m.j, m.k};
      double c = m.c;
      for (int d = 0; d < 3; ++d) {  // differentiate once along each index
        c *= e[idx[k][d]];
        --e[idx[k][d]];
      }
      if (c == 0) continue;
      v += c * ipow(b[0], e[0]) * ipow(b[1], e[1]) * ipow(b[2], e[2]);
    }
    q3[k] = v;
  }
}

// (r, s, t) -> position in (xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz)
inline int tidx(int r, int s, int t) {
  int a[3] = {r, s, t};
  std::sort(a, a + 3);
  static const int map[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}},
                                   {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                                   {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
  return map[a[0]][a[1]][a[2]];
}

inline int hidx(int r, int t) {  // (r, t) -> position in (xx, xy, xz, yy, yz, zz)
  static const int map[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  return map[r][t];
}

struct SubGrid {
  int n;      // unpadded
  int npad;   // padded to a multiple of 8 (GridGeneration.cpp:214)
  long first = 0;         // first point in generation order
  bool weighted = false;  // w holds the Becke-partitioned weights (computed on first use)
  std::vector<double> x, y, z, w;
  // AO tables, [mu][g] (value), [r][mu][g] (gradient), [mu][g] (Laplacian)
  std::vector<double> ao, ao1, ao2l;
  std::vector<double> rho, rho1, sigma, tau, lapl;           // per point
  std::vector<double> eps, vrho, vsigma, vlapl, vtau;
};

}  // namespace

struct ox_system {
  int natoms = 0;
  std::vector<int> Z;
  std::vector<double> xyz;
  std::vector<Shell> shells;
  int nbasis = 0;
  std::vector<std::vector<Poly>> shell_ang;
  std::vector<double> points;   // 4*G generation order
  std::vector<double> raw_w;    // radial x angular weight before the Becke partition
  std::vector<long> atom_start; // first point of each atom's grid (natoms + 1)
  std::vector<SubGrid> subgrids;
  std::vector<int> active;
  std::vector<int> codes;
  int family = -1;
  int nthreads = 1;
  bool ao_ready = false;
};

namespace {

std::string elem_name(int z) {
  if (z < 1 || z > kNElem) throw std::runtime_error("oracle: unsupported element");
  return kElem[z];
}

int elem_z(std::string s) {
  for (auto& ch : s) ch = (char)std::toupper((unsigned char)ch);
  for (int z = 1; z <= kNElem; ++z) {
    std::string e = kElem[z];
    for (auto& ch : e) ch = (char)std::toupper((unsigned char)ch);
    if (e == s) return z;
  }
  return -1;
}

double parse_fortran_double(std::string t) {
  for (auto& ch : t) if (ch == 'D' || ch == 'd') ch = 'E';
  return std::stod(t);
}

// ---- Basis: .gbs reader (restates legacy/MwfnIO/Mwfn.cpp:606-680) + normalisation
// (libint2 Shell::renorm as called by src/Integral/Normalization.cpp:11-18).
std::map<int, std::vector<Shell>> read_gbs(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Basis set file is missing: " + path);
  std::map<int, std::vector<Shell>> out;
  std::string line;
  int cur = -1;
  std::vector<Shell> cur_shells;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    std::string w;
    if (!(ss >> w)) continue;
    if (w[0] == '!') continue;
    if (w[0] == '-') {
      cur = elem_z(w.substr(1));
      cur_shells.clear();
    } else if (w == "****") {
      if (cur > 0) out[cur] = cur_shells;
      cur = -1;
    } else if (w == "S" || w == "SP" || w == "P" || w == "D" || w == "F" || w == "G") {
      std::string ns; ss >> ns;
      const int n = std::stoi(ns);
      Shell a, b;
      a.type = (w == "S" || w == "SP") ? 0 : (w == "P" ? 1 : (w == "D" ? -2 : (w == "F" ? -3 : -4)));
      b.type = 1;
      for (int i = 0; i < n; ++i) {
        std::getline(f, line);
        std::stringstream ls(line);
        std::string e, c1, c2;
        ls >> e >> c1;
        a.exps.push_back(parse_fortran_double(e));
        a.coefs.push_back(parse_fortran_double(c1));
        if (w == "SP") {
          ls >> c2;
          b.exps.push_back(parse_fortran_double(e));
          b.coefs.push_back(parse_fortran_double(c2));
        }
      }
      cur_shells.push_back(a);
      if (w == "SP") cur_shells.push_back(b);
    }
  }
  return out;
}

double dfact(int n) { double r = 1; for (int k = n; k > 1; k -= 2) r *= k; return r; }  // n!!

void normalise(Shell& s) {
  const int l = std::abs(s.type);
  const double pi32 = std::pow(M_PI, 1.5);
  const double dfl = dfact(2 * l - 1);
  for (size_t p = 0; p < s.exps.size(); ++p) {
    const double a2 = 2 * s.exps[p];
    s.coefs[p] *= std::sqrt(std::pow(2.0, l) * std::pow(a2, l + 1.5) / (pi32 * dfl));
  }
  double norm = 0;
  for (size_t p = 0; p < s.exps.size(); ++p)
    for (size_t q = 0; q < s.exps.size(); ++q)
      norm += dfl * pi32 * s.coefs[p] * s.coefs[q] / (std::pow(2.0, l) * std::pow(s.exps[p] + s.exps[q], l + 1.5));
  const double f = 1.0 / std::sqrt(norm);
  for (double& c : s.coefs) c *= f;
}

int shell_size(int type) { return type == 1 ? 3 : 2 * std::abs(type) + 1; }

// ---- Lebedev: octahedral expansion of orbit generators.
std::map<int, std::vector<std::array<double, 4>>> g_lebedev;

void oh_orbit(int type, double a, double b, double v, std::vector<std::array<double, 4>>& out) {
  auto signs3 = [&](double x, double y, double z) {
    for (int sx = -1; sx <= 1; sx += 2)
      for (int sy = -1; sy <= 1; sy += 2)
        for (int sz = -1; sz <= 1; sz += 2) {
          if ((x == 0 && sx < 0) || (y == 0 && sy < 0) || (z == 0 && sz < 0)) continue;
          out.push_back({sx * x, sy * y, sz * z, v});
        }
  };
  auto perms = [&](double p, double q, double r, bool distinct) {
    // all distinct permutations of (p,q,r)
    double t[3] = {p, q, r};
    std::vector<std::array<double, 3>> seen;
    int idx[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (auto& ix : idx) {
      std::array<double, 3> c = {t[ix[0]], t[ix[1]], t[ix[2]]};
      bool dup = false;
      for (auto& s : seen) if (s == c) dup = true;
      if (dup) continue;
      seen.push_back(c);
      signs3(c[0], c[1], c[2]);
    }
    (void)distinct;
  };
  switch (type) {
    case 1: perms(1, 0, 0, false); break;
    case 2: { double r = std::sqrt(0.5); perms(0, r, r, false); break; }
    case 3: { double r = std::sqrt(1.0 / 3.0); perms(r, r, r, false); break; }
    case 4: perms(a, a, std::sqrt(1 - 2 * a * a), false); break;
    case 5: perms(a, std::sqrt(1 - a * a), 0, true); break;
    case 6: perms(a, b, std::sqrt(1 - a * a - b * b), true); break;
    default: throw std::runtime_error("oracle: bad Lebedev orbit type");
  }
}

void load_lebedev(const std::string& path) {
  if (!g_lebedev.empty()) return;
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Lebedev table missing: " + path);
  std::string line;
  int cur = -1;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::stringstream ss(line);
    std::string w; ss >> w;
    if (w == "rule") { int n, k; ss >> n >> k; cur = n; g_lebedev[cur] = {}; continue; }
    int type = std::stoi(w);
    std::string sa, sb, sv; ss >> sa >> sb >> sv;
    oh_orbit(type, std::stod(sa), std::stod(sb), std::stod(sv), g_lebedev[cur]);
  }
  for (auto& kv : g_lebedev)
    if ((int)kv.second.size() != kv.first) throw std::runtime_error("oracle: Lebedev expansion size mismatch");
}

// s(mu) with three Becke smoothing iterations (GridGeneration.cpp:50-71).
double becke_s(double x0, double y0, double z0, const double* b, const double* c) {
  const double r01 = std::sqrt((x0 - b[0]) * (x0 - b[0]) + (y0 - b[1]) * (y0 - b[1]) + (z0 - b[2]) * (z0 - b[2]));
  const double r02 = std::sqrt((x0 - c[0]) * (x0 - c[0]) + (y0 - c[1]) * (y0 - c[1]) + (z0 - c[2]) * (z0 - c[2]));
  const double r12 = std::sqrt((b[0] - c[0]) * (b[0] - c[0]) + (b[1] - c[1]) * (b[1] - c[1]) + (b[2] - c[2]) * (b[2] - c[2]));
  const double mu = (r01 - r02) / r12;
  double p = mu;
  for (int k = 0; k < 3; ++k) p = 1.5 * p - 0.5 * p * p * p;
  return 0.5 * (1 - p);
}

struct AtomGrid { std::vector<double> r, wr; std::vector<int> order; };

AtomGrid read_atom_grid(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Missing element grid file in folder!");
  std::string line;
  std::getline(f, line);
  int ngroups, ntot;
  { std::stringstream ss(line); ss >> ngroups >> ntot; }
  std::getline(f, line);
  std::stringstream ss(line);
  std::string formula; ss >> formula;
  AtomGrid g;
  g.r.resize(ntot); g.wr.resize(ntot);
  if (formula == "de2") {  // GridGeneration.cpp:90-100
    double alpha, x0, x1; ss >> alpha >> x0 >> x1;
    const double h = (x1 - x0) / (double)(ntot - 1);
    for (int i = 0; i < ntot; ++i) {
      const double xi = h * i + x0;
      g.r[i] = std::exp(alpha * xi - std::exp(-xi));
      g.wr[i] = std::exp(3 * alpha * xi - 3 * std::exp(-xi)) * (alpha + std::exp(-xi)) * h;
    }
  } else if (formula == "em") {  // GridGeneration.cpp:103-113
    double R; ss >> R;
    for (int i = 0; i < ntot; ++i) {
      const double ii = i + 1;
      g.r[i] = R * std::pow(ii / (ntot + 1 - ii), 2);
      g.wr[i] = 2 * std::pow(R, 3) * (ntot + 1) * std::pow(ii, 5) / std::pow(ntot + 1 - ii, 7);
    }
  } else {
    throw std::runtime_error("Unrecognized radial formula!");
  }
  for (int k = 0; k < ngroups; ++k) {
    std::getline(f, line);
    std::stringstream gs(line);
    int npts, nsh; gs >> npts >> nsh;
    for (int s = 0; s < nsh; ++s) g.order.push_back(npts);
  }
  if ((int)g.order.size() != ntot) throw std::runtime_error("Corrupted grid file");
  return g;
}

// Contiguous recursive halving into <=128-point sub-grids. The reference's Cut
// (GridGeneration.cpp:177-201) reorders only the first 4 points of a parent and
// then splits by column index, so membership is exactly this halving.
void recursive_halving(long lo, long n, std::vector<std::pair<long, long>>& out) {
  if (n <= 128) { out.push_back({lo, n}); return; }
  const long left = n / 2;
  recursive_halving(lo, left, out);
  recursive_halving(lo + left, n - left, out);
}

// AO tables for one sub-grid (GridAO.cpp:14-215).
void compute_ao(ox_system* s, SubGrid& sg, int family) {
  const int ng = sg.npad, nb = s->nbasis;
  sg.ao.assign((size_t)ng * nb, 0);
  sg.ao1.assign((size_t)3 * ng * nb, 0);
  if (family >= 2) sg.ao2l.assign((size_t)ng * nb, 0);
  else sg.ao2l.clear();
  int mu0 = 0;
  for (size_t is = 0; is < s->shells.size(); ++is) {
    const Shell& sh = s->shells[is];
    const double* c = &s->xyz[3 * sh.center];
    const std::vector<Poly>& ang = s->shell_ang[is];
    for (int g = 0; g < ng; ++g) {
      const double x = sg.x[g] - c[0], y = sg.y[g] - c[1], z = sg.z[g] - c[2];
      const double r2 = x * x + y * y + z * z;
      double A0 = 0, A1 = 0, A2 = 0;
      for (size_t p = 0; p < sh.exps.size(); ++p) {
        const double t = sh.coefs[p] * std::exp(-sh.exps[p] * r2);
        A0 += t; A1 += t * sh.exps[p]; A2 += t * sh.exps[p] * sh.exps[p];
      }
      const double Ag[3] = {-2 * x * A1, -2 * y * A1, -2 * z * A1};
      const double Alap = -6 * A1 + 4 * r2 * A2;
      for (size_t m = 0; m < ang.size(); ++m) {
        double q, qg[3], ql;
        poly_eval(ang[m], x, y, z, &q, qg, &ql);
        const int mu = mu0 + (int)m;
        sg.ao[(size_t)mu * ng + g] = A0 * q;
        for (int r = 0; r < 3; ++r) sg.ao1[((size_t)r * nb + mu) * ng + g] = Ag[r] * q + A0 * qg[r];
        if (family >= 2)
          sg.ao2l[(size_t)mu * ng + g] = Alap * q + 2 * (Ag[0] * qg[0] + Ag[1] * qg[1] + Ag[2] * qg[2]) + A0 * ql;
      }
    }
    mu0 += (int)ang.size();
  }
}

// Density on one sub-grid (GridDensity.cpp:12-41, EinSum nests Rho, Rho1, Sigma, Tau, Lapl).
void density_subgrid(const ox_system* s, SubGrid& sg, const double* D, int family, std::vector<double>& tmp) {
  const int ng = sg.npad, nb = s->nbasis;
  tmp.assign(ng, 0);
  sg.rho.assign(ng, 0);
  // Rho(g) = sum_nu (sum_mu D(mu,nu) AO(g,mu)) AO(g,nu)
  for (int nu = 0; nu < nb; ++nu) {
    std::fill(tmp.begin(), tmp.end(), 0.0);
    for (int mu = 0; mu < nb; ++mu) {
      const double d = D[(size_t)nu * nb + mu];
      const double* a = &sg.ao[(size_t)mu * ng];
      for (int g = 0; g < ng; ++g) tmp[g] += d * a[g];
    }
    const double* a = &sg.ao[(size_t)nu * ng];
    for (int g = 0; g < ng; ++g) sg.rho[g] += tmp[g] * a[g];
  }
  if (family >= 1) {
    sg.rho1.assign((size_t)3 * ng, 0);
    for (int r = 0; r < 3; ++r)
      for (int nu = 0; nu < nb; ++nu) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        for (int mu = 0; mu < nb; ++mu) {
          const double d = D[(size_t)nu * nb + mu];
          const double* a1 = &sg.ao1[((size_t)r * nb + mu) * ng];
          for (int g = 0; g < ng; ++g) tmp[g] += d * a1[g];
        }
        const double* a = &sg.ao[(size_t)nu * ng];
        for (int g = 0; g < ng; ++g) sg.rho1[(size_t)r * ng + g] += tmp[g] * a[g];
      }
    for (double& v : sg.rho1) v *= 2;
    sg.sigma.assign(ng, 0);
    for (int g = 0; g < ng; ++g)
      for (int r = 0; r < 3; ++r) sg.sigma[g] += sg.rho1[(size_t)r * ng + g] * sg.rho1[(size_t)r * ng + g];
  }
  if (family >= 2) {
    sg.tau.assign(ng, 0);
    for (int r = 0; r < 3; ++r)
      for (int nu = 0; nu < nb; ++nu) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        for (int mu = 0; mu < nb; ++mu) {
          const double d = D[(size_t)nu * nb + mu];
          const double* a1 = &sg.ao1[((size_t)r * nb + mu) * ng];
          for (int g = 0; g < ng; ++g) tmp[g] += d * a1[g];
        }
        const double* a1n = &sg.ao1[((size_t)r * nb + nu) * ng];
        for (int g = 0; g < ng; ++g) sg.tau[g] += tmp[g] * a1n[g];
      }
    for (double& v : sg.tau) v *= 0.5;
    sg.lapl.assign(ng, 0);
    for (int g = 0; g < ng; ++g) sg.lapl[g] = 4 * sg.tau[g];
    for (int nu = 0; nu < nb; ++nu) {
      std::fill(tmp.begin(), tmp.end(), 0.0);
      for (int mu = 0; mu < nb; ++mu) {
        const double d = D[(size_t)nu * nb + mu];
        const double* a = &sg.ao[(size_t)mu * ng];
        for (int g = 0; g < ng; ++g) tmp[g] += d * a[g];
      }
      const double* l = &sg.ao2l[(size_t)nu * ng];
      for (int g = 0; g < ng; ++g) sg.lapl[g] += tmp[g] * l[g];
    }
    for (int nu = 0; nu < nb; ++nu) {
      std::fill(tmp.begin(), tmp.end(), 0.0);
      for (int mu = 0; mu < nb; ++mu) {
        const double d = D[(size_t)nu * nb + mu];
        const double* l = &sg.ao2l[(size_t)mu * ng];
        for (int g = 0; g < ng; ++g) tmp[g] += d * l[g];
      }
      const double* a = &sg.ao[(size_t)nu * ng];
      for (int g = 0; g < ng; ++g) sg.lapl[g] += tmp[g] * a[g];
    }
  }
}

void evaluate_subgrid(const ox_system* s, SubGrid& sg) {
  const int ng = sg.npad;
  sg.eps.assign(ng, 0); sg.vrho.assign(ng, 0); sg.vsigma.assign(ng, 0);
  sg.vlapl.assign(ng, 0); sg.vtau.assign(ng, 0);
  for (int code : s->codes)
    for (int g = 0; g < ng; ++g) {
      const double sig = s->family >= 1 ? sg.sigma[g] : 0.0;
      const double tau = s->family >= 2 ? sg.tau[g] : 0.0;
      oxf::eval_point(code, sg.rho[g], sig, tau, &sg.eps[g], &sg.vrho[g], &sg.vsigma[g], &sg.vtau[g]);
    }
}

// Vxc contribution of one sub-grid (GridFock.cpp:11-49), added to F_ (nb x nb column-major).
void fock_subgrid(const ox_system* s, const SubGrid& sg, double* Fout, std::vector<double>& F,
                  std::vector<double>& tmp, std::vector<double>& tmp3) {
  const int ng = sg.npad, nb = s->nbasis;
  F.assign((size_t)nb * nb, 0);
  tmp.assign(ng, 0);
  // F0(mu,nu) = sum_g W v_rho AO(g,mu) AO(g,nu); F += 0.5 F0
  for (int g = 0; g < ng; ++g) tmp[g] = sg.w[g] * sg.vrho[g];
  for (int mu = 0; mu < nb; ++mu) {
    std::vector<double> wa(ng);
    const double* am = &sg.ao[(size_t)mu * ng];
    for (int g = 0; g < ng; ++g) wa[g] = tmp[g] * am[g];
    for (int nu = 0; nu < nb; ++nu) {
      const double* an = &sg.ao[(size_t)nu * ng];
      double acc = 0;
      for (int g = 0; g < ng; ++g) acc += wa[g] * an[g];
      F[(size_t)nu * nb + mu] += 0.5 * acc;
    }
  }
  if (s->family >= 1) {
    // F1(mu,nu) = sum_g W (2 v_sigma) sum_r Rho1(g,r) AO1(g,mu,r) AO(g,nu)
    tmp3.assign((size_t)3 * ng, 0);
    for (int r = 0; r < 3; ++r)
      for (int g = 0; g < ng; ++g) tmp3[(size_t)r * ng + g] = sg.w[g] * 2.0 * sg.vsigma[g] * sg.rho1[(size_t)r * ng + g];
    std::vector<double> wa(ng);
    for (int mu = 0; mu < nb; ++mu) {
      std::fill(wa.begin(), wa.end(), 0.0);
      for (int r = 0; r < 3; ++r) {
        const double* a1 = &sg.ao1[((size_t)r * nb + mu) * ng];
        for (int g = 0; g < ng; ++g) wa[g] += tmp3[(size_t)r * ng + g] * a1[g];
      }
      for (int nu = 0; nu < nb; ++nu) {
        const double* an = &sg.ao[(size_t)nu * ng];
        double acc = 0;
        for (int g = 0; g < ng; ++g) acc += wa[g] * an[g];
        F[(size_t)nu * nb + mu] += acc;
      }
    }
  }
  if (s->family >= 2) {
    // F2 = sum_g W (v_tau/2 + 2 v_lapl) AO1.AO1; F += 0.5 F2; then v_lapl AO AO2L
    for (int g = 0; g < ng; ++g) tmp[g] = sg.w[g] * (0.5 * sg.vtau[g] + 2 * sg.vlapl[g]);
    std::vector<double> wa(ng);
    for (int mu = 0; mu < nb; ++mu)
      for (int nu = 0; nu < nb; ++nu) {
        double acc = 0;
        for (int r = 0; r < 3; ++r) {
          const double* am = &sg.ao1[((size_t)r * nb + mu) * ng];
          const double* an = &sg.ao1[((size_t)r * nb + nu) * ng];
          for (int g = 0; g < ng; ++g) acc += tmp[g] * am[g] * an[g];
        }
        F[(size_t)nu * nb + mu] += 0.5 * acc;
      }
    for (int g = 0; g < ng; ++g) tmp[g] = sg.w[g] * sg.vlapl[g];
    bool any = false;
    for (int g = 0; g < ng; ++g) any |= (tmp[g] != 0);
    if (any)
      for (int mu = 0; mu < nb; ++mu)
        for (int nu = 0; nu < nb; ++nu) {
          const double* am = &sg.ao[(size_t)mu * ng];
          const double* ln = &sg.ao2l[(size_t)nu * ng];
          double acc = 0;
          for (int g = 0; g < ng; ++g) acc += tmp[g] * am[g] * ln[g];
          F[(size_t)nu * nb + mu] += acc;
        }
  }
  // F += F^T, then accumulate (all AOs kept: BasisList is the identity).
  for (int nu = 0; nu < nb; ++nu)
    for (int mu = 0; mu < nb; ++mu)
      Fout[(size_t)nu * nb + mu] += F[(size_t)nu * nb + mu] + F[(size_t)mu * nb + nu];
}

template <class Fn> int guarded(Fn fn) {
  try { fn(); return 0; } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

}  // namespace

extern "C" {

const char* ox_last_error(void) { return g_err.c_str(); }

// Plain Becke partition (GridGeneration.cpp:136-153): w = w_rad w_ang P_A / sum_B P_B
// for the points of the given sub-grids that are not weighted yet, written
// into the sub-grid and into the point list. Done lazily so that a timing
// sample of a few sub-grids of a large system does not pay for all points.
static void weigh_subgrids(ox_system* s, const std::vector<int>& ids) {
  std::vector<long> pts;
  for (int i : ids) {
    const SubGrid& sg = s->subgrids[i];
    if (sg.weighted) continue;
    for (int g = 0; g < sg.n; ++g) pts.push_back(sg.first + g);
  }
  const int nat = s->natoms;
  const double* X = s->xyz.data();
  const long np = (long)pts.size();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(s->nthreads)
  for (long k = 0; k < np; ++k) {
    const long p = pts[k];
    int owner = 0;
    while (p >= s->atom_start[owner + 1]) ++owner;
    double* q = &s->points[4 * p];
    double total = 0, mine = 0;
    for (int b = 0; b < nat; ++b) {
      double pb = 1;
      for (int c = 0; c < nat; ++c) {
        if (b == c) continue;
        pb *= becke_s(q[0], q[1], q[2], X + 3 * b, X + 3 * c);
      }
      total += pb;
      if (b == owner) mine = pb;
    }
    q[3] = s->raw_w[p] * (mine / total);
  }
  for (int i : ids) {
    SubGrid& sg = s->subgrids[i];
    if (sg.weighted) continue;
    for (int g = 0; g < sg.n; ++g) sg.w[g] = s->points[4 * (sg.first + g) + 3];
    sg.weighted = true;
  }
}

static void weigh_all(ox_system* s) {
  std::vector<int> ids(s->subgrids.size());
  for (size_t i = 0; i < ids.size(); ++i) ids[i] = (int)i;
  weigh_subgrids(s, ids);
}

ox_system* ox_create(const char* data_root, const char* grid_name, const char* basis_name,
                     int natoms, const int* Z, const double* xyz_bohr, int nthreads) {
  ox_system* s = new ox_system;
  int rc = guarded([&] {
    s->natoms = natoms;
    s->Z.assign(Z, Z + natoms);
    s->xyz.assign(xyz_bohr, xyz_bohr + 3 * natoms);
    s->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
    const std::string root(data_root);
    load_lebedev(root + "/lebedev_orbits.txt");
    // basis
    auto lib = read_gbs(root + "/BasisSets/" + basis_name + ".gbs");
    for (int a = 0; a < natoms; ++a) {
      auto it = lib.find(Z[a]);
      if (it == lib.end()) throw std::runtime_error("Basis set file does not include this element!");
      for (Shell sh : it->second) {
        sh.center = a;
        normalise(sh);
        s->shells.push_back(sh);
      }
    }
    s->nbasis = 0;
    for (auto& sh : s->shells) {
      s->shell_ang.push_back(angular_functions(sh.type));
      s->nbasis += shell_size(sh.type);
    }
    // grid (GridGeneration.cpp:73-162)
    std::vector<AtomGrid> ag(natoms);
    long G = 0;
    for (int a = 0; a < natoms; ++a) {
      ag[a] = read_atom_grid(root + "/Grids/" + grid_name + "/" + elem_name(Z[a]) + ".grid");
      for (int n : ag[a].order) {
        if (!g_lebedev.count(n)) throw std::runtime_error("oracle: Lebedev order not tabulated");
        G += n;
      }
    }
    s->points.assign((size_t)4 * G, 0);
    std::vector<long> start(natoms + 1, 0);
    for (int a = 0; a < natoms; ++a) {
      long n = 0; for (int k : ag[a].order) n += k;
      start[a + 1] = start[a] + n;
    }
    for (int a = 0; a < natoms; ++a) {
      long p = start[a];
      const double* ca = &s->xyz[3 * a];
      for (size_t ish = 0; ish < ag[a].order.size(); ++ish) {
        const auto& leb = g_lebedev[ag[a].order[ish]];
        for (const auto& l : leb) {
          double* q = &s->points[4 * p];
          q[0] = l[0] * ag[a].r[ish] + ca[0];
          q[1] = l[1] * ag[a].r[ish] + ca[1];
          q[2] = l[2] * ag[a].r[ish] + ca[2];
          q[3] = ag[a].wr[ish] * 4 * M_PI * l[3];
          ++p;
        }
      }
    }
    s->raw_w.resize(G);
    for (long p = 0; p < G; ++p) s->raw_w[p] = s->points[4 * p + 3];
    s->atom_start = start;
    // sub-grids (the Becke partition of their weights is applied on first use,
    // see weigh_subgrids; membership depends only on the generation order)
    std::vector<std::pair<long, long>> parts;
    recursive_halving(0, G, parts);
    s->subgrids.resize(parts.size());
    for (size_t i = 0; i < parts.size(); ++i) {
      SubGrid& sg = s->subgrids[i];
      sg.n = (int)parts[i].second;
      sg.npad = (sg.n + 7) & ~7;
      sg.first = parts[i].first;
      sg.x.resize(sg.npad); sg.y.resize(sg.npad); sg.z.resize(sg.npad); sg.w.assign(sg.npad, 0);
      for (int g = 0; g < sg.npad; ++g) {
        const long src = parts[i].first + (g < sg.n ? g : g - sg.n);
        sg.x[g] = s->points[4 * src]; sg.y[g] = s->points[4 * src + 1]; sg.z[g] = s->points[4 * src + 2];
      }
    }
    s->active.clear();
    for (size_t i = 0; i < s->subgrids.size(); ++i) s->active.push_back((int)i);
  });
  if (rc) { delete s; return nullptr; }
  return s;
}

void ox_destroy(ox_system* s) { delete s; }
int ox_nbasis(const ox_system* s) { return s->nbasis; }
long ox_npoints(const ox_system* s) { return (long)(s->points.size() / 4); }
int ox_nsubgrids(const ox_system* s) { return (int)s->subgrids.size(); }
int ox_nshells(const ox_system* s) { return (int)s->shells.size(); }

int ox_get_raw_weights(const ox_system* s, double* w) {
  std::memcpy(w, s->raw_w.data(), s->raw_w.size() * sizeof(double));
  return 0;
}

int ox_get_points(ox_system* s, double* xyzw) {
  weigh_all(s);
  std::memcpy(xyzw, s->points.data(), s->points.size() * sizeof(double));
  return 0;
}

int ox_get_subgrid_sizes(const ox_system* s, int* sizes) {
  for (size_t i = 0; i < s->subgrids.size(); ++i) sizes[i] = s->subgrids[i].n;
  return 0;
}

int ox_get_shells(const ox_system* s, int* type, int* nprim, int* center, double* exps, double* coefs) {
  size_t k = 0;
  for (size_t i = 0; i < s->shells.size(); ++i) {
    type[i] = s->shells[i].type;
    nprim[i] = (int)s->shells[i].exps.size();
    center[i] = s->shells[i].center;
    for (size_t p = 0; p < s->shells[i].exps.size(); ++p, ++k) {
      exps[k] = s->shells[i].exps[p];
      coefs[k] = s->shells[i].coefs[p];
    }
  }
  return 0;
}

int ox_lebedev(const char* data_root, int n, double* xyzw) {
  return guarded([&] {
    load_lebedev(std::string(data_root) + "/lebedev_orbits.txt");
    auto it = g_lebedev.find(n);
    if (it == g_lebedev.end()) throw std::runtime_error("oracle: Lebedev order not tabulated");
    for (size_t i = 0; i < it->second.size(); ++i)
      for (int k = 0; k < 4; ++k) xyzw[4 * i + k] = it->second[i][k];
  });
}

int ox_set_functional_codes(ox_system* s, int ncodes, const int* codes) {
  return guarded([&] {
    s->codes.assign(codes, codes + ncodes);
    int fam = 0;
    for (int c : s->codes) {
      const int f = oxf::family_of(c);
      if (f < 0) throw std::runtime_error("oracle: functional code not implemented");
      fam = std::max(fam, f);
    }
    if (s->family != fam) s->ao_ready = false;
    s->family = fam;
  });
}

int ox_set_functional_file(ox_system* s, const char* data_root, const char* df_name) {
  std::string path = std::string(data_root) + "/DensityFunctionals/" + df_name + ".df";
  std::ifstream f(path);
  if (!f.good()) { g_err = "Density functional file is missing!"; return -1; }
  std::string line;
  std::getline(f, line);
  std::stringstream ss(line);
  std::vector<int> codes;
  int c;
  while (ss >> c) codes.push_back(c);
  if (codes.empty()) { g_err = "Corrupted density functional file!"; return -1; }
  return ox_set_functional_codes(s, (int)codes.size(), codes.data());
}

int ox_family(const ox_system* s) { return s->family; }

int ox_set_subgrid_subset(ox_system* s, int n, const int* ids) {
  s->active.clear();
  if (n <= 0) {
    for (size_t i = 0; i < s->subgrids.size(); ++i) s->active.push_back((int)i);
  } else {
    for (int i = 0; i < n; ++i) {
      if (ids[i] < 0 || ids[i] >= (int)s->subgrids.size()) { g_err = "bad sub-grid id"; return -1; }
      s->active.push_back(ids[i]);
    }
  }
  s->ao_ready = false;
  return 0;
}

int ox_get_ao(ox_system* s) {
  return guarded([&] {
    if (s->family < 0) throw std::runtime_error("oracle: set a functional first");
    // free AO tables of inactive sub-grids
    std::vector<char> act(s->subgrids.size(), 0);
    for (int i : s->active) act[i] = 1;
    for (size_t i = 0; i < s->subgrids.size(); ++i)
      if (!act[i]) { s->subgrids[i].ao.clear(); s->subgrids[i].ao.shrink_to_fit(); s->subgrids[i].ao1.clear(); s->subgrids[i].ao1.shrink_to_fit(); s->subgrids[i].ao2l.clear(); }
    weigh_subgrids(s, s->active);
    const int n = (int)s->active.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
    for (int i = 0; i < n; ++i) compute_ao(s, s->subgrids[s->active[i]], s->family);
    s->ao_ready = true;
  });
}

int ox_density(ox_system* s, const double* D) {
  if (!s->ao_ready) { g_err = "oracle: AO tables not ready"; return -1; }
  const int n = (int)s->active.size();
#pragma omp parallel num_threads(s->nthreads)
  {
    std::vector<double> tmp;
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) density_subgrid(s, s->subgrids[s->active[i]], D, s->family, tmp);
  }
  return 0;
}

int ox_evaluate(ox_system* s) {
  const int n = (int)s->active.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
  for (int i = 0; i < n; ++i) evaluate_subgrid(s, s->subgrids[s->active[i]]);
  return 0;
}

int ox_energy(ox_system* s, double* Exc) {
  double E = 0;
  const int n = (int)s->active.size();
#pragma omp parallel for reduction(+ : E) schedule(static) num_threads(s->nthreads)
  for (int i = 0; i < n; ++i) {
    const SubGrid& sg = s->subgrids[s->active[i]];
    for (int g = 0; g < sg.npad; ++g) E += sg.w[g] * sg.rho[g] * sg.eps[g];
  }
  *Exc = E;
  return 0;
}

int ox_num_electrons(ox_system* s, double* Nel) {
  double N = 0;
  const int n = (int)s->active.size();
#pragma omp parallel for reduction(+ : N) schedule(static) num_threads(s->nthreads)
  for (int i = 0; i < n; ++i) {
    const SubGrid& sg = s->subgrids[s->active[i]];
    for (int g = 0; g < sg.npad; ++g) N += sg.w[g] * sg.rho[g];
  }
  *Nel = N;
  return 0;
}

int ox_fock(ox_system* s, double* V) {
  const int nb = s->nbasis;
  const int n = (int)s->active.size();
  std::memset(V, 0, sizeof(double) * nb * nb);
  const int nt = s->nthreads;
  std::vector<std::vector<double>> part(nt, std::vector<double>((size_t)nb * nb, 0.0));
#pragma omp parallel num_threads(nt)
  {
    std::vector<double> F, tmp, tmp3;
    double* mine = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) fock_subgrid(s, s->subgrids[s->active[i]], mine, F, tmp, tmp3);
  }
  for (int t = 0; t < nt; ++t)
    for (size_t k = 0; k < (size_t)nb * nb; ++k) V[k] += part[t][k];
  return 0;
}

// One perturbation's response matrix on one sub-grid (GridFock.cpp:99-160 at
// nspin = 1), from the perturbed density rho^x, grad rho^x at its points
// (rx[g], r1x[r][g]; may alias sg.rho / sg.rho1) and the ground state's
// grad rho (rho10) and f_xc (fx[4g..]: v_sigma, v2rho2, v2rhosigma, v2sigma2):
//   sigma^x = 2 grad rho . grad rho^x,
//   F0 weight: v2rho2 rho^x + v2rhosigma sigma^x,
//   F1 vector: (v2rhosigma rho^x + v2sigma2 sigma^x) grad rho + v_sigma grad rho^x,
// evaluated through fock_subgrid (vrho := the F0 weight, vsigma := 1,
// rho1 := the F1 vector); adds into Fout. Leaves sg.vrho, sg.vsigma, sg.rho1 changed.
static void response_subgrid(const ox_system* s, SubGrid& sg, const std::vector<double>& fx,
                             const std::vector<double>& rho10, const double* rx, const double* r1x, double* Fout,
                             std::vector<double>& F, std::vector<double>& tmp, std::vector<double>& tmp3) {
  const int ng = sg.npad;
  std::vector<double> vr(ng), g1(s->family >= 1 ? (size_t)3 * ng : 0);
  for (int g = 0; g < ng; ++g) {
    double sx = 0.0;
    if (s->family >= 1)
      for (int r = 0; r < 3; ++r) sx += 2.0 * rho10[(size_t)r * ng + g] * r1x[(size_t)r * ng + g];
    vr[g] = fx[4 * g + 1] * rx[g] + fx[4 * g + 2] * sx;
    if (s->family >= 1) {
      const double q = fx[4 * g + 2] * rx[g] + fx[4 * g + 3] * sx;
      for (int r = 0; r < 3; ++r) g1[(size_t)r * ng + g] = q * rho10[(size_t)r * ng + g] + fx[4 * g] * r1x[(size_t)r * ng + g];
    }
  }
  sg.vrho = vr;
  sg.vsigma.assign(ng, 1.0);
  if (s->family >= 1) sg.rho1 = g1;
  fock_subgrid(s, sg, Fout, F, tmp, tmp3);
}

// Response Fock matrices (Evaluate("f") + Grid::getDensityU + Grid::getFockU<u_t>,
// src/ExchangeCorrelation.cpp:101-104,126-133, GridDensity.cpp:47-77,
// GridFock.cpp:99-160), closed shell, LDA/GGA, after ox_density at the
// reference density. For each perturbed density matrix D^x:
//   RhoU = sum D^x phi phi, Rho1U = 2 sum D^x grad phi phi, SigmaU = 2 Rho1 . Rho1U,
//   TMP0 = v2rho2 RhoU + v2rhosigma SigmaU                 -> F0 (weight 1/2),
//   TMP4 = 2 (v2rhosigma RhoU + v2sigma2 SigmaU) Rho1 + 2 v_sigma Rho1U -> F1,
//   F = 1/2 F0 + F1, F += F^T.
// Evaluated through fock_subgrid with vrho := TMP0, vsigma := 1 and
// rho1 := TMP4 / 2 on each sub-grid.
int ox_fock_u(ox_system* s, int nmats, const double* Dx, double* Fx) {
  return guarded([&] {
    if (s->family > 1) throw std::runtime_error("oracle: the response path is LDA/GGA");
    const int nb = s->nbasis, n = (int)s->active.size(), nt = s->nthreads;
    const size_t nb2 = (size_t)nb * nb;
    std::vector<std::vector<double>> part(nt, std::vector<double>((size_t)nmats * nb2, 0.0));
#pragma omp parallel num_threads(nt)
    {
      std::vector<double> F, tmp, tmp3, fx, rho0, rho10, sig0, vr0, vs0;
      double* mine = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        rho0 = sg.rho; rho10 = sg.rho1; sig0 = sg.sigma; vr0 = sg.vrho; vs0 = sg.vsigma;
        fx.assign((size_t)4 * ng, 0.0);  // v_sigma, v2rho2, v2rhosigma, v2sigma2
        for (int code : s->codes)
          for (int g = 0; g < ng; ++g)
            oxf::eval_point_f(code, rho0[g], s->family >= 1 ? sig0[g] : 0.0, &fx[4 * g], &fx[4 * g + 1],
                              &fx[4 * g + 2], &fx[4 * g + 3]);
        for (int m = 0; m < nmats; ++m) {
          density_subgrid(s, sg, Dx + m * nb2, s->family, tmp);
          response_subgrid(s, sg, fx, rho10, sg.rho.data(), sg.rho1.data(), mine + m * nb2, F, tmp, tmp3);
        }
        sg.rho = rho0; sg.rho1 = rho10; sg.sigma = sig0; sg.vrho = vr0; sg.vsigma = vs0;
      }
    }
    std::memset(Fx, 0, sizeof(double) * nmats * nb2);
    for (int t = 0; t < nt; ++t)
      for (size_t k = 0; k < (size_t)nmats * nb2; ++k) Fx[k] += part[t][k];
  });
}

// Spin-polarised XC build (Unrestricted/SP.cpp:65-68 -> Grid::getDensity({Da, Db}),
// Evaluate("ev") with the functional polarised, getEnergy, getFock), LDA/GGA
// and r2SCAN: per sub-grid, rho_s, grad rho_s (and tau_s) of each spin
// (GridDensity.cpp:22-41 at nspin = 2, sigma packed aa, ab, bb), the polarised
// functional per point, then per spin u (GridFock.cpp:16-40): F = 1/2 F0[w v_rho_u]
// + F1 with the scaled sigma derivatives (2 on aa / bb, 1 on ab) against
// grad rho_v (+ 1/2 F2[w v_tau_u / 2]), F += F^T. Evaluated through fock_subgrid
// with vrho := v_rho_u, vsigma := 1, rho1 := (2 v_sigma_uu grad rho_u +
// v_sigma_ab grad rho_other) / 2 and vtau := v_tau_u.
int ox_xc_pol(ox_system* s, const double* Da, const double* Db, double* Va, double* Vb, double* Exc, double* Nel) {
  return guarded([&] {
    if (!s->ao_ready) throw std::runtime_error("oracle: AO tables not ready");
    const bool mgga = s->family >= 2;
    for (int code : s->codes)
      if (mgga && code != 497 && code != 498)
        throw std::runtime_error("oracle: the spin-polarised meta-GGA path is r2SCAN");
    const int nb = s->nbasis, n = (int)s->active.size(), nt = s->nthreads;
    const size_t nb2 = (size_t)nb * nb;
    std::vector<std::vector<double>> part(nt, std::vector<double>(2 * nb2, 0.0));
    double E = 0, N = 0;
#pragma omp parallel num_threads(nt) reduction(+ : E, N)
    {
      std::vector<double> F, tmp, tmp3, ra, rb, ga, gb, ta, tb, eps, vr, vs, vt;
      double* mine = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        density_subgrid(s, sg, Da, s->family, tmp);
        ra = sg.rho;
        ga = s->family >= 1 ? sg.rho1 : std::vector<double>((size_t)3 * ng, 0.0);
        if (mgga) ta = sg.tau;
        density_subgrid(s, sg, Db, s->family, tmp);
        rb = sg.rho;
        gb = s->family >= 1 ? sg.rho1 : std::vector<double>((size_t)3 * ng, 0.0);
        if (mgga) tb = sg.tau;
        eps.assign(ng, 0.0);
        vr.assign((size_t)2 * ng, 0.0);
        vs.assign((size_t)3 * ng, 0.0);
        vt.assign((size_t)2 * ng, 0.0);
        for (int g = 0; g < ng; ++g) {
          double saa = 0, sab = 0, sbb = 0;
          for (int r = 0; r < 3; ++r) {
            const double a = ga[(size_t)r * ng + g], b = gb[(size_t)r * ng + g];
            saa += a * a; sab += a * b; sbb += b * b;
          }
          for (int code : s->codes)
            if (mgga)
              oxf::eval_point_mpol(code, ra[g], rb[g], saa, sab, sbb, ta[g], tb[g], &eps[g], &vr[2 * g], &vs[3 * g],
                                   &vt[2 * g]);
            else
              oxf::eval_point_pol(code, ra[g], rb[g], saa, sab, sbb, &eps[g], &vr[2 * g], &vs[3 * g]);
          E += sg.w[g] * (ra[g] + rb[g]) * eps[g];
          N += sg.w[g] * (ra[g] + rb[g]);
        }
        for (int u = 0; u < 2; ++u) {
          const std::vector<double>& gu = u == 0 ? ga : gb;
          const std::vector<double>& go = u == 0 ? gb : ga;
          sg.vrho.assign(ng, 0.0);
          sg.vsigma.assign(ng, 1.0);
          sg.rho1.assign((size_t)3 * ng, 0.0);
          if (mgga) {  // the tau term of spin u (GridFock.cpp:32-40 with Eps1Tau of spin u); no Laplacian
            sg.vtau.assign(ng, 0.0);
            sg.vlapl.assign(ng, 0.0);
            for (int g = 0; g < ng; ++g) sg.vtau[g] = vt[2 * g + u];
          }
          for (int g = 0; g < ng; ++g) {
            sg.vrho[g] = vr[2 * g + u];
            const double suu = vs[3 * g + (u == 0 ? 0 : 2)], sab = vs[3 * g + 1];
            for (int r = 0; r < 3; ++r)
              sg.rho1[(size_t)r * ng + g] = 0.5 * (2.0 * suu * gu[(size_t)r * ng + g] + sab * go[(size_t)r * ng + g]);
          }
          fock_subgrid(s, sg, mine + u * nb2, F, tmp, tmp3);
        }
      }
    }
    std::memset(Va, 0, sizeof(double) * nb2);
    std::memset(Vb, 0, sizeof(double) * nb2);
    for (int t = 0; t < nt; ++t)
      for (size_t k = 0; k < nb2; ++k) {
        Va[k] += part[t][k];
        Vb[k] += part[t][nb2 + k];
      }
    if (Exc) *Exc = E;
    if (Nel) *Nel = N;
  });
}

int ox_eval_functional_mpol(int ncodes, const int* codes, long n, const double* rho, const double* sigma,
                            const double* tau, double* exc, double* vrho, double* vsigma, double* vtau) {
  for (long i = 0; i < n; ++i) {
    double e = 0, vr[2] = {0, 0}, vs[3] = {0, 0, 0}, vt[2] = {0, 0};
    for (int k = 0; k < ncodes; ++k) {
      if (codes[k] != 497 && codes[k] != 498) { g_err = "oracle: the spin-polarised meta-GGA path is r2SCAN"; return -1; }
      oxf::eval_point_mpol(codes[k], rho[2 * i], rho[2 * i + 1], sigma[3 * i], sigma[3 * i + 1], sigma[3 * i + 2],
                           tau[2 * i], tau[2 * i + 1], &e, vr, vs, vt);
    }
    exc[i] = e;
    vrho[2 * i] = vr[0]; vrho[2 * i + 1] = vr[1];
    for (int k = 0; k < 3; ++k) vsigma[3 * i + k] = vs[k];
    vtau[2 * i] = vt[0]; vtau[2 * i + 1] = vt[1];
  }
  return 0;
}

// Spin-polarised response (Evaluate("f") with the functional polarised, then
// Grid::getDensityU / Grid::getFockU<u_t> at nspin = 2: GridDensity.cpp:47-77,
// GridFock.cpp:99-160), LDA/GGA, ground state Da, Db. Per perturbation m and
// spin u, from rho^x_s, grad rho^x_s of Dxa[m], Dxb[m]:
//   sigma^x_aa = 2 grad rho_a . grad rho^x_a, sigma^x_ab = grad rho^x_a . grad rho_b
//     + grad rho_a . grad rho^x_b, sigma^x_bb = 2 grad rho_b . grad rho^x_b (SigmaU);
//   dv_rho_u = sum_v f_{rho_u rho_v} rho^x_v + sum_x f_{rho_u sigma_x} sigma^x_x;
//   dv_sigma_x = sum_w f_{rho_w sigma_x} rho^x_w + sum_y f_{sigma_x sigma_y} sigma^x_y;
//   F^x_u = 1/2 F0[w dv_rho_u] + F1[w (S(uu) dv_sigma_uu grad rho_u + dv_sigma_ab grad rho_o
//           + S(uu) v_sigma_uu grad rho^x_u + v_sigma_ab grad rho^x_o)], F += F^T,
// with S = 2 on aa and bb and 1 on ab: the derivative of the polarised
// ground-state Fock matrix (GridFock.cpp:16-41 at nspin = 2) along D^x. The
// reference scales f_{sigma sigma} by S of the index it contracts with
// SigmaU (FockEinSum/S_x...Eps2Sigma2_g,w,x---ScaledEps2Sigma2_g,w,x.hpp with
// ScaledEps2Sigma2_g,u,v...SigmaG_g,mat,v---TMP1_g,mat,u.hpp), which differs
// from the derivative in the aa-ab and ab-bb cross terms; this follows the
// derivative. Evaluated through fock_subgrid (vrho := dv_rho_u, vsigma := 1,
// rho1 := the F1 vector / 2).
int ox_fock_u_pol(ox_system* s, const double* Da, const double* Db, int nmats, const double* Dxa, const double* Dxb,
                  double* Fa, double* Fb) {
  return guarded([&] {
    if (!s->ao_ready) throw std::runtime_error("oracle: AO tables not ready");
    if (s->family > 1) throw std::runtime_error("oracle: the response path is LDA/GGA");
    const bool gga = s->family >= 1;
    const int nb = s->nbasis, n = (int)s->active.size(), nt = s->nthreads;
    const size_t nb2 = (size_t)nb * nb;
    std::vector<std::vector<double>> part(nt, std::vector<double>((size_t)2 * nmats * nb2, 0.0));
#pragma omp parallel num_threads(nt)
    {
      std::vector<double> F, tmp, tmp3, ga, gb, fx, xa, xb, gxa, gxb;
      double* mine = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        const std::vector<double> zero3((size_t)3 * ng, 0.0);
        density_subgrid(s, sg, Da, s->family, tmp);
        const std::vector<double> ra = sg.rho;
        ga = gga ? sg.rho1 : zero3;
        density_subgrid(s, sg, Db, s->family, tmp);
        const std::vector<double> rb = sg.rho;
        gb = gga ? sg.rho1 : zero3;
        // per point: vsigma[3], v2rho2[3], v2rhosigma[6], v2sigma2[6]
        fx.assign((size_t)18 * ng, 0.0);
        for (int g = 0; g < ng; ++g) {
          double saa = 0, sab = 0, sbb = 0;
          for (int r = 0; r < 3; ++r) {
            const double a = ga[(size_t)r * ng + g], b = gb[(size_t)r * ng + g];
            saa += a * a; sab += a * b; sbb += b * b;
          }
          double* f = &fx[(size_t)18 * g];
          for (int code : s->codes) oxf::eval_point_fpol(code, ra[g], rb[g], saa, sab, sbb, f, f + 3, f + 6, f + 12);
          if (!gga) std::fill(f, f + 3, 0.0), std::fill(f + 6, f + 18, 0.0);
        }
        for (int m = 0; m < nmats; ++m) {
          density_subgrid(s, sg, Dxa + m * nb2, s->family, tmp);
          xa = sg.rho;
          gxa = gga ? sg.rho1 : zero3;
          density_subgrid(s, sg, Dxb + m * nb2, s->family, tmp);
          xb = sg.rho;
          gxb = gga ? sg.rho1 : zero3;
          std::vector<double> dvr((size_t)2 * ng), vec((size_t)2 * 3 * ng, 0.0);
          for (int g = 0; g < ng; ++g) {
            const double* f = &fx[(size_t)18 * g];
            double sx[3] = {0, 0, 0};
            for (int r = 0; r < 3; ++r) {
              const size_t k = (size_t)r * ng + g;
              sx[0] += 2.0 * ga[k] * gxa[k];
              sx[1] += gxa[k] * gb[k] + ga[k] * gxb[k];
              sx[2] += 2.0 * gb[k] * gxb[k];
            }
            const double rx[2] = {xa[g], xb[g]};
            auto frr = [&](int u, int v) { return f[3 + (u + v)]; };  // aa, ab, bb
            auto frs = [&](int w, int x) { return f[6 + 3 * w + x]; };
            auto fss = [&](int x, int y) {
              const int a = std::min(x, y), b = std::max(x, y);
              static const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
              return f[12 + idx[a][b]];
            };
            double dvs[3];
            for (int x = 0; x < 3; ++x) {
              dvs[x] = frs(0, x) * rx[0] + frs(1, x) * rx[1];
              for (int y = 0; y < 3; ++y) dvs[x] += fss(x, y) * sx[y];
            }
            for (int u = 0; u < 2; ++u) {
              double d = frr(u, 0) * rx[0] + frr(u, 1) * rx[1];
              for (int x = 0; x < 3; ++x) d += frs(u, x) * sx[x];
              dvr[(size_t)u * ng + g] = d;
              if (!gga) continue;
              const int uu = u == 0 ? 0 : 2;
              const std::vector<double>& gu = u == 0 ? ga : gb;
              const std::vector<double>& go = u == 0 ? gb : ga;
              const std::vector<double>& gxu = u == 0 ? gxa : gxb;
              const std::vector<double>& gxo = u == 0 ? gxb : gxa;
              for (int r = 0; r < 3; ++r) {
                const size_t k = (size_t)r * ng + g;
                vec[(size_t)u * 3 * ng + k] =
                    0.5 * (2.0 * dvs[uu] * gu[k] + dvs[1] * go[k] + 2.0 * f[uu] * gxu[k] + f[1] * gxo[k]);
              }
            }
          }
          for (int u = 0; u < 2; ++u) {
            sg.vrho.assign(dvr.begin() + (size_t)u * ng, dvr.begin() + (size_t)(u + 1) * ng);
            sg.vsigma.assign(ng, 1.0);
            if (gga) sg.rho1.assign(vec.begin() + (size_t)u * 3 * ng, vec.begin() + (size_t)(u + 1) * 3 * ng);
            fock_subgrid(s, sg, mine + ((size_t)u * nmats + m) * nb2, F, tmp, tmp3);
          }
        }
      }
    }
    std::memset(Fa, 0, sizeof(double) * nmats * nb2);
    std::memset(Fb, 0, sizeof(double) * nmats * nb2);
    for (int t = 0; t < nt; ++t)
      for (size_t k = 0; k < (size_t)nmats * nb2; ++k) {
        Fa[k] += part[t][k];
        Fb[k] += part[t][(size_t)nmats * nb2 + k];
      }
  });
}

// Polarised f_xc of the summed LDA/GGA codes at n points (libxc XC_POLARIZED
// layout): rho[2n], sigma[3n] in; vsigma[3n], v2rho2[3n], v2rhosigma[6n],
// v2sigma2[6n] out (each may be null).
int ox_eval_functional_fpol(int ncodes, const int* codes, long n, const double* rho, const double* sigma,
                            double* vsigma, double* v2rho2, double* v2rhosigma, double* v2sigma2) {
  for (long i = 0; i < n; ++i) {
    double f[18] = {0};
    const double s0 = sigma ? sigma[3 * i] : 0.0, s1 = sigma ? sigma[3 * i + 1] : 0.0, s2 = sigma ? sigma[3 * i + 2] : 0.0;
    for (int k = 0; k < ncodes; ++k) {
      if (oxf::family_of(codes[k]) > 1) { g_err = "oracle: the spin-polarised f_xc is LDA/GGA"; return -1; }
      oxf::eval_point_fpol(codes[k], rho[2 * i], rho[2 * i + 1], s0, s1, s2, f, f + 3, f + 6, f + 12);
    }
    if (vsigma) for (int k = 0; k < 3; ++k) vsigma[3 * i + k] = f[k];
    if (v2rho2) for (int k = 0; k < 3; ++k) v2rho2[3 * i + k] = f[3 + k];
    if (v2rhosigma) for (int k = 0; k < 6; ++k) v2rhosigma[6 * i + k] = f[6 + k];
    if (v2sigma2) for (int k = 0; k < 6; ++k) v2sigma2[6 * i + k] = f[12 + k];
  }
  return 0;
}

int ox_eval_functional_pol(int ncodes, const int* codes, long n, const double* rho, const double* sigma, double* exc,
                           double* vrho, double* vsigma) {
  for (long i = 0; i < n; ++i) {
    double e = 0, vr[2] = {0, 0}, vs[3] = {0, 0, 0};
    const double s0 = sigma ? sigma[3 * i] : 0.0, s1 = sigma ? sigma[3 * i + 1] : 0.0, s2 = sigma ? sigma[3 * i + 2] : 0.0;
    for (int k = 0; k < ncodes; ++k) {
      if (oxf::family_of(codes[k]) > 1) { g_err = "oracle: the spin-polarised functionals are LDA/GGA"; return -1; }
      oxf::eval_point_pol(codes[k], rho[2 * i], rho[2 * i + 1], s0, s1, s2, &e, vr, vs);
    }
    if (exc) exc[i] = e;
    if (vrho) { vrho[2 * i] = vr[0]; vrho[2 * i + 1] = vr[1]; }
    if (vsigma) for (int k = 0; k < 3; ++k) vsigma[3 * i + k] = vs[k];
  }
  return 0;
}

// phi, grad phi and the Hessian (xx, xy, xz, yy, yz, zz) of every AO at one
// sub-grid's points: ao[mu][g], ao1[r][mu][g], ao2[hidx][mu][g] (the reference's
// AO, AO1, AO2 of Grid::getAO(2, .), GridAO.cpp).
static void ao_hessian_tables(const ox_system* s, const SubGrid& sg, std::vector<double>& ao, std::vector<double>& ao1,
                              std::vector<double>& ao2) {
  const int nb = s->nbasis, ng = sg.npad;
  ao.assign((size_t)nb * ng, 0); ao1.assign((size_t)3 * nb * ng, 0); ao2.assign((size_t)6 * nb * ng, 0);
  int mu0 = 0;
  for (size_t is = 0; is < s->shells.size(); ++is) {
    const Shell& sh = s->shells[is];
    const double* c = &s->xyz[3 * sh.center];
    for (int g = 0; g < ng; ++g) {
      const double d[3] = {sg.x[g] - c[0], sg.y[g] - c[1], sg.z[g] - c[2]};
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double A0 = 0, A1 = 0, A2 = 0;
      for (size_t p = 0; p < sh.exps.size(); ++p) {
        const double t = sh.coefs[p] * std::exp(-sh.exps[p] * r2);
        A0 += t; A1 += t * sh.exps[p]; A2 += t * sh.exps[p] * sh.exps[p];
      }
      double Rg[3], Rh[6];
      for (int r = 0; r < 3; ++r) Rg[r] = -2 * d[r] * A1;
      for (int r = 0; r < 3; ++r)
        for (int t = r; t < 3; ++t) Rh[hidx(r, t)] = 4 * d[r] * d[t] * A2 - (r == t ? 2 * A1 : 0.0);
      for (size_t m = 0; m < s->shell_ang[is].size(); ++m) {
        double q, qg[3], qh[6];
        poly_eval_h(s->shell_ang[is][m], d[0], d[1], d[2], &q, qg, qh);
        const size_t mu = mu0 + m;
        ao[mu * ng + g] = A0 * q;
        for (int r = 0; r < 3; ++r) ao1[((size_t)r * nb + mu) * ng + g] = Rg[r] * q + A0 * qg[r];
        for (int r = 0; r < 3; ++r)
          for (int t = r; t < 3; ++t)
            ao2[((size_t)hidx(r, t) * nb + mu) * ng + g] =
                Rh[hidx(r, t)] * q + Rg[r] * qg[t] + Rg[t] * qg[r] + A0 * qh[hidx(r, t)];
      }
    }
    mu0 += (int)s->shell_ang[is].size();
  }
}

// XC nuclear gradient in the reference's skeleton form, the grid held fixed
// (Restricted/Grad.cpp:72-84 -> SubGrid::getDensitySkeleton,
// GridDensity.cpp:79-130 -> SubGrid::getEnergyGrad, GridEnergy.cpp:15-30):
//   RhoGrad[g,t,A]    = -2 sum_{mu in A, nu} D_mu,nu d_t phi_mu phi_nu
//   Rho1Grad[g,r,t,A] = -2 sum_{mu in A, nu} D_mu,nu (d_r d_t phi_mu phi_nu + d_t phi_mu d_r phi_nu)
//   SigmaGrad[g,t,A]  = 2 sum_r Rho1[g,r] Rho1Grad[g,r,t,A]
//   E[t,A] = sum_g W (Eps1Rho RhoGrad + Eps1Sigma SigmaGrad)
// with Eps1Rho, Eps1Sigma, Rho1 from the preceding density / evaluation of
// the same D. LDA and GGA (the reference has no tau term here).
int ox_xc_grad(ox_system* s, const double* D, double* grad) {
  return guarded([&] {
    if (s->family > 1) throw std::runtime_error("oracle: the XC gradient is built for LDA and GGA");
    if (ox_density(s, D) || ox_evaluate(s)) throw std::runtime_error(g_err);
    const int nb = s->nbasis, na = s->natoms;
    std::vector<int> ao_atom(nb);
    for (size_t is = 0, mu = 0; is < s->shells.size(); ++is)
      for (size_t m = 0; m < s->shell_ang[is].size(); ++m) ao_atom[mu++] = s->shells[is].center;
    std::vector<double> G((size_t)3 * na, 0.0);
    const int n = (int)s->active.size();
#pragma omp parallel num_threads(s->nthreads)
    {
      std::vector<double> Gl((size_t)3 * na, 0.0), ao, ao1, ao2, X, U;
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        const SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        ao_hessian_tables(s, sg, ao, ao1, ao2);
        // X_mu = sum_nu D phi_nu, U_mu,r = sum_nu D d_r phi_nu
        X.assign((size_t)nb * ng, 0); U.assign((size_t)3 * nb * ng, 0);
        for (int mu = 0; mu < nb; ++mu)
          for (int nu = 0; nu < nb; ++nu) {
            const double dd = D[(size_t)nu * nb + mu];
            if (dd == 0.0) continue;
            for (int g = 0; g < ng; ++g) X[(size_t)mu * ng + g] += dd * ao[(size_t)nu * ng + g];
            if (s->family >= 1)
              for (int r = 0; r < 3; ++r)
                for (int g = 0; g < ng; ++g)
                  U[((size_t)r * nb + mu) * ng + g] += dd * ao1[((size_t)r * nb + nu) * ng + g];
          }
        for (int mu = 0; mu < nb; ++mu) {
          const int A = ao_atom[mu];
          for (int t = 0; t < 3; ++t) {
            double acc = 0;
            for (int g = 0; g < ng; ++g) {
              const double dt = ao1[((size_t)t * nb + mu) * ng + g], x = X[(size_t)mu * ng + g];
              double e = sg.vrho[g] * (-2 * dt * x);  // RhoGrad
              if (s->family >= 1) {
                double sg_ = 0;  // SigmaGrad = 2 sum_r Rho1 Rho1Grad
                for (int r = 0; r < 3; ++r) {
                  const double r1g = -2 * (ao2[((size_t)hidx(r, t) * nb + mu) * ng + g] * x +
                                           dt * U[((size_t)r * nb + mu) * ng + g]);
                  sg_ += 2 * sg.rho1[(size_t)r * ng + g] * r1g;
                }
                e += sg.vsigma[g] * sg_;
              }
              acc += sg.w[g] * e;
            }
            Gl[(size_t)3 * A + t] += acc;
          }
        }
      }
#pragma omp critical
      for (size_t k = 0; k < G.size(); ++k) G[k] += Gl[k];
    }
    for (size_t k = 0; k < G.size(); ++k) grad[k] = G[k];
  });
}

// XC response to the skeleton density derivatives (Restricted/Hess.cpp:79-97:
// Evaluate("f"), getDensitySkeleton2({D}), then Grid::getFockU<s_t>,
// Grid.cpp:193-216 with GridFock.cpp:99-160): for each nuclear coordinate
// (atom A, direction t), the response matrix of GridFock.cpp:99-160 with the
// perturbed density taken from getDensitySkeleton (GridDensity.cpp:90-128):
//   RhoGrad[g,t,A]    = -2 sum_{mu in A, nu} D_mu,nu d_t phi_mu phi_nu,
//   Rho1Grad[g,r,t,A] = -2 sum_{mu in A, nu} D_mu,nu (d_r d_t phi_mu phi_nu + d_t phi_mu d_r phi_nu),
//   SigmaGrad = 2 grad rho . Rho1Grad (formed inside the response step).
// D is the ground-state density (f_xc is evaluated at it). Fx: 3 natoms
// column-major nbasis^2 matrices, [3 A + t]. Closed shell, LDA/GGA.
int ox_fock_s(ox_system* s, const double* D, double* Fx) {
  return guarded([&] {
    if (s->family > 1) throw std::runtime_error("oracle: the skeleton response is LDA/GGA");
    if (ox_density(s, D)) throw std::runtime_error(g_err);
    const int nb = s->nbasis, na = s->natoms, n = (int)s->active.size(), nt = s->nthreads;
    const size_t nb2 = (size_t)nb * nb, npert = (size_t)3 * na;
    std::vector<int> ao_atom(nb);
    for (size_t is = 0, mu = 0; is < s->shells.size(); ++is)
      for (size_t m = 0; m < s->shell_ang[is].size(); ++m) ao_atom[mu++] = s->shells[is].center;
    std::vector<std::vector<double>> part(nt, std::vector<double>(npert * nb2, 0.0));
#pragma omp parallel num_threads(nt)
    {
      std::vector<double> F, tmp, tmp3, fx, rho0, rho10, sig0, vr0, vs0, ao, ao1, ao2, X, U, rx, r1x;
      double* mine = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        rho0 = sg.rho; rho10 = sg.rho1; sig0 = sg.sigma; vr0 = sg.vrho; vs0 = sg.vsigma;
        fx.assign((size_t)4 * ng, 0.0);
        for (int code : s->codes)
          for (int g = 0; g < ng; ++g)
            oxf::eval_point_f(code, rho0[g], s->family >= 1 ? sig0[g] : 0.0, &fx[4 * g], &fx[4 * g + 1],
                              &fx[4 * g + 2], &fx[4 * g + 3]);
        ao_hessian_tables(s, sg, ao, ao1, ao2);
        // X_mu = sum_nu D phi_nu, U_mu,r = sum_nu D d_r phi_nu
        X.assign((size_t)nb * ng, 0); U.assign((size_t)3 * nb * ng, 0);
        for (int mu = 0; mu < nb; ++mu)
          for (int nu = 0; nu < nb; ++nu) {
            const double dd = D[(size_t)nu * nb + mu];
            if (dd == 0.0) continue;
            for (int g = 0; g < ng; ++g) X[(size_t)mu * ng + g] += dd * ao[(size_t)nu * ng + g];
            for (int r = 0; r < 3; ++r)
              for (int g = 0; g < ng; ++g)
                U[((size_t)r * nb + mu) * ng + g] += dd * ao1[((size_t)r * nb + nu) * ng + g];
          }
        for (int A = 0; A < na; ++A)
          for (int t = 0; t < 3; ++t) {
            rx.assign(ng, 0.0); r1x.assign((size_t)3 * ng, 0.0);
            for (int mu = 0; mu < nb; ++mu) {
              if (ao_atom[mu] != A) continue;
              for (int g = 0; g < ng; ++g) {
                const double dt = ao1[((size_t)t * nb + mu) * ng + g], x = X[(size_t)mu * ng + g];
                rx[g] += -2 * dt * x;
                for (int r = 0; r < 3; ++r)
                  r1x[(size_t)r * ng + g] += -2 * (ao2[((size_t)hidx(r, t) * nb + mu) * ng + g] * x +
                                                   dt * U[((size_t)r * nb + mu) * ng + g]);
              }
            }
            response_subgrid(s, sg, fx, rho10, rx.data(), r1x.data(), mine + (size_t)(3 * A + t) * nb2, F, tmp, tmp3);
          }
        sg.rho = rho0; sg.rho1 = rho10; sg.sigma = sig0; sg.vrho = vr0; sg.vsigma = vs0;
      }
    }
    std::memset(Fx, 0, sizeof(double) * npert * nb2);
    for (int t = 0; t < nt; ++t)
      for (size_t k = 0; k < npert * nb2; ++k) Fx[k] += part[t][k];
  });
}

// Third derivatives of every AO at a sub-grid's points, ao3[tidx][mu][g]:
// phi = R Q with R = sum c exp(-alpha r^2), the product rule over R's and Q's
// derivatives, R_ijk = -8 x_i x_j x_k A3 + 4 (d_ij x_k + d_ik x_j + d_jk x_i) A2.
static void ao_third_tables(const ox_system* s, const SubGrid& sg, std::vector<double>& ao3) {
  const int nb = s->nbasis, ng = sg.npad;
  ao3.assign((size_t)10 * nb * ng, 0);
  static const int idx[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                                 {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
  int mu0 = 0;
  for (size_t is = 0; is < s->shells.size(); ++is) {
    const Shell& sh = s->shells[is];
    const double* c = &s->xyz[3 * sh.center];
    for (int g = 0; g < ng; ++g) {
      const double d[3] = {sg.x[g] - c[0], sg.y[g] - c[1], sg.z[g] - c[2]};
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double A[4] = {0, 0, 0, 0};
      for (size_t p = 0; p < sh.exps.size(); ++p) {
        const double t = sh.coefs[p] * std::exp(-sh.exps[p] * r2);
        A[0] += t; A[1] += t * sh.exps[p]; A[2] += t * sh.exps[p] * sh.exps[p];
        A[3] += t * sh.exps[p] * sh.exps[p] * sh.exps[p];
      }
      auto R1 = [&](int i) { return -2 * d[i] * A[1]; };
      auto R2 = [&](int i, int j) { return 4 * d[i] * d[j] * A[2] - (i == j ? 2 * A[1] : 0.0); };
      auto R3 = [&](int i, int j, int k) {
        return -8 * d[i] * d[j] * d[k] * A[3] +
               4 * ((i == j ? d[k] : 0.0) + (i == k ? d[j] : 0.0) + (j == k ? d[i] : 0.0)) * A[2];
      };
      for (size_t m = 0; m < s->shell_ang[is].size(); ++m) {
        double q, qg[3], qh[6], q3[10];
        poly_eval_h(s->shell_ang[is][m], d[0], d[1], d[2], &q, qg, qh);
        poly_eval_3(s->shell_ang[is][m], d[0], d[1], d[2], q3);
        const size_t mu = mu0 + m;
        for (int k = 0; k < 10; ++k) {
          const int i = idx[k][0], j = idx[k][1], l = idx[k][2];
          const double v = R3(i, j, l) * q + R2(i, j) * qg[l] + R2(i, l) * qg[j] + R2(j, l) * qg[i] +
                           R1(i) * qh[hidx(j, l)] + R1(j) * qh[hidx(i, l)] + R1(l) * qh[hidx(i, j)] + A[0] * q3[k];
          ao3[((size_t)k * nb + mu) * ng + g] = v;
        }
      }
    }
    mu0 += (int)s->shell_ang[is].size();
  }
}

// XC energy Hessian in the reference's skeleton form, the grid held fixed
// (Restricted/Hess.cpp:78-90 -> SubGrid::getDensitySkeleton2, GridDensity.cpp:132-209,
// and SubGrid::getEnergyHess, GridEnergy.cpp:32-45), closed shell, LDA/GGA: with
// a = (A, t), b = (B, s) and the skeleton density derivatives
//   rho^a = -2 sum_{mu in A} d_t phi_mu X_mu,   X = D phi, Y^r = D d_r phi,
//   rho^ab = 2 [ [A = B] sum_{mu in A} d_t d_s phi_mu X_mu
//                + sum_{mu in A, nu in B} D_mu,nu d_t phi_mu d_s phi_nu ]
// and their gradients, sigma^a = 2 grad rho . grad rho^a,
// sigma^ab = 2 (grad rho . grad rho^ab + grad rho^a . grad rho^b):
//   H[a][b] = sum_g W (Eps2Rho2 rho^a rho^b + Eps2RhoSigma (rho^a sigma^b + sigma^a rho^b)
//             + Eps2Sigma2 sigma^a sigma^b + Eps1Rho rho^ab + Eps1Sigma sigma^ab).
// H: 3 natoms x 3 natoms, row-major, [3 A + t][3 B + s].
int ox_xc_hess(ox_system* s, const double* D, double* H) {
  return guarded([&] {
    if (s->family > 1) throw std::runtime_error("oracle: the XC Hessian is built for LDA and GGA");
    if (ox_density(s, D) || ox_evaluate(s)) throw std::runtime_error(g_err);
    const int nb = s->nbasis, na = s->natoms, n = (int)s->active.size(), nt = s->nthreads, np = 3 * na;
    const bool gga = s->family >= 1;
    std::vector<int> ao_atom(nb);
    for (size_t is = 0, mu = 0; is < s->shells.size(); ++is)
      for (size_t m = 0; m < s->shell_ang[is].size(); ++m) ao_atom[mu++] = s->shells[is].center;
    std::vector<std::vector<double>> part(nt, std::vector<double>((size_t)np * np, 0.0));
#pragma omp parallel num_threads(nt)
    {
      std::vector<double> ao, ao1, ao2, ao3, X, U;
      std::vector<double> r1((size_t)np * 4), r2((size_t)np * np * 4);
      double* Hl = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        const SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        ao_hessian_tables(s, sg, ao, ao1, ao2);
        if (gga) ao_third_tables(s, sg, ao3);
        X.assign((size_t)nb * ng, 0); U.assign((size_t)3 * nb * ng, 0);
        for (int mu = 0; mu < nb; ++mu)
          for (int nu = 0; nu < nb; ++nu) {
            const double dd = D[(size_t)nu * nb + mu];
            if (dd == 0.0) continue;
            for (int g = 0; g < ng; ++g) X[(size_t)mu * ng + g] += dd * ao[(size_t)nu * ng + g];
            for (int r = 0; r < 3; ++r)
              for (int g = 0; g < ng; ++g)
                U[((size_t)r * nb + mu) * ng + g] += dd * ao1[((size_t)r * nb + nu) * ng + g];
          }
        for (int g = 0; g < ng; ++g) {
          if (sg.w[g] == 0.0) continue;
          auto A1 = [&](int t, int mu) { return ao1[((size_t)t * nb + mu) * ng + g]; };
          auto A2 = [&](int r, int t, int mu) { return ao2[((size_t)hidx(r, t) * nb + mu) * ng + g]; };
          auto A3 = [&](int r, int t, int u, int mu) { return ao3[((size_t)tidx(r, t, u) * nb + mu) * ng + g]; };
          std::fill(r1.begin(), r1.end(), 0.0);
          std::fill(r2.begin(), r2.end(), 0.0);
          // first order: r1[a][0] = rho^a, r1[a][1 + r] = d_r rho^a
          for (int mu = 0; mu < nb; ++mu) {
            const int A = ao_atom[mu];
            const double x = X[(size_t)mu * ng + g];
            for (int t = 0; t < 3; ++t) {
              double* q = &r1[(size_t)(3 * A + t) * 4];
              q[0] += -2 * A1(t, mu) * x;
              if (gga)
                for (int r = 0; r < 3; ++r)
                  q[1 + r] += -2 * (A2(r, t, mu) * x + A1(t, mu) * U[((size_t)r * nb + mu) * ng + g]);
            }
          }
          // second order, diagonal atom blocks
          for (int mu = 0; mu < nb; ++mu) {
            const int A = ao_atom[mu];
            const double x = X[(size_t)mu * ng + g];
            for (int t = 0; t < 3; ++t)
              for (int u = 0; u < 3; ++u) {
                double* q = &r2[((size_t)(3 * A + t) * np + 3 * A + u) * 4];
                q[0] += 2 * A2(t, u, mu) * x;
                if (gga)
                  for (int r = 0; r < 3; ++r)
                    q[1 + r] += 2 * (A3(r, t, u, mu) * x + A2(t, u, mu) * U[((size_t)r * nb + mu) * ng + g]);
              }
          }
          // second order, D_mu,nu d_t phi_mu d_s phi_nu
          for (int mu = 0; mu < nb; ++mu)
            for (int nu = 0; nu < nb; ++nu) {
              const double dd = D[(size_t)nu * nb + mu];
              if (dd == 0.0) continue;
              const int A = ao_atom[mu], B = ao_atom[nu];
              for (int t = 0; t < 3; ++t)
                for (int u = 0; u < 3; ++u) {
                  double* q = &r2[((size_t)(3 * A + t) * np + 3 * B + u) * 4];
                  q[0] += 2 * dd * A1(t, mu) * A1(u, nu);
                  if (gga)
                    for (int r = 0; r < 3; ++r)
                      q[1 + r] += 2 * dd * (A2(r, t, mu) * A1(u, nu) + A1(t, mu) * A2(r, u, nu));
                }
            }
          // the functional's derivatives at the point (summed over the codes)
          double vs = 0, frr = 0, frs = 0, fss = 0;
          for (int code : s->codes) {
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            oxf::eval_point_f(code, sg.rho[g], gga ? sg.sigma[g] : 0.0, &a0, &a1, &a2, &a3);
            vs += a0; frr += a1; frs += a2; fss += a3;
          }
          (void)vs;
          double gr[3] = {0, 0, 0};  // rho1 is filled for GGA only
          if (gga)
            for (int r = 0; r < 3; ++r) gr[r] = sg.rho1[(size_t)r * ng + g];
          const double w = sg.w[g], vr = sg.vrho[g], vsg = gga ? sg.vsigma[g] : 0.0;
          for (int a = 0; a < np; ++a) {
            const double* qa = &r1[(size_t)a * 4];
            const double sa = gga ? 2 * (gr[0] * qa[1] + gr[1] * qa[2] + gr[2] * qa[3]) : 0.0;
            for (int b = 0; b < np; ++b) {
              const double* qb = &r1[(size_t)b * 4];
              const double* qab = &r2[((size_t)a * np + b) * 4];
              double e = frr * qa[0] * qb[0] + vr * qab[0];
              if (gga) {
                const double sb = 2 * (gr[0] * qb[1] + gr[1] * qb[2] + gr[2] * qb[3]);
                const double sab = 2 * (gr[0] * qab[1] + gr[1] * qab[2] + gr[2] * qab[3] + qa[1] * qb[1] +
                                        qa[2] * qb[2] + qa[3] * qb[3]);
                e += frs * (qa[0] * sb + sa * qb[0]) + fss * sa * sb + vsg * sab;
              }
              Hl[(size_t)a * np + b] += w * e;
            }
          }
        }
      }
    }
    std::memset(H, 0, sizeof(double) * np * np);
    for (int t = 0; t < nt; ++t)
      for (size_t k = 0; k < (size_t)np * np; ++k) H[k] += part[t][k];
  });
}

// Skeleton derivative of the XC Fock matrix, the potential held fixed
// (Restricted/Hess.cpp:94 -> Grid::getFockSkeleton, Grid.cpp:173-191 ->
// SubGrid::getFockSkeleton, GridFock.cpp:51-97): for each atom A and direction t,
//   M[mu,nu] = sum_g W (Eps1Rho d_t phi_mu phi_nu
//              + 2 Eps1Sigma sum_r Rho1[g,r] (d_r d_t phi_mu phi_nu + d_t phi_mu d_r phi_nu))
// over mu in A (GridFock.cpp:57-84; 2 Eps1Sigma is ScaledEps1Sigma at nspin = 1),
// F[3 A + t] = -(M + M^T) (GridFock.cpp:64, 83, 86). Eps1Rho, Eps1Sigma and Rho1
// come from the density of D and its evaluation. Closed shell, LDA/GGA.
int ox_fock_skeleton(ox_system* s, const double* D, double* Fx) {
  return guarded([&] {
    if (s->family > 1) throw std::runtime_error("oracle: the skeleton Fock derivative is LDA/GGA");
    if (ox_density(s, D) || ox_evaluate(s)) throw std::runtime_error(g_err);
    const int nb = s->nbasis, na = s->natoms, n = (int)s->active.size(), nt = s->nthreads;
    const size_t nb2 = (size_t)nb * nb, npert = (size_t)3 * na;
    std::vector<int> ao_atom(nb);
    for (size_t is = 0, mu = 0; is < s->shells.size(); ++is)
      for (size_t m = 0; m < s->shell_ang[is].size(); ++m) ao_atom[mu++] = s->shells[is].center;
    // M^t[mu][nu] for every mu (each AO belongs to one atom), summed over the sub-grids
    std::vector<std::vector<double>> part(nt, std::vector<double>(3 * nb2, 0.0));
#pragma omp parallel num_threads(nt)
    {
      std::vector<double> ao, ao1, ao2, c;
      double* M = part[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < n; ++i) {
        const SubGrid& sg = s->subgrids[s->active[i]];
        const int ng = sg.npad;
        ao_hessian_tables(s, sg, ao, ao1, ao2);
        for (int t = 0; t < 3; ++t)
          for (int mu = 0; mu < nb; ++mu) {
            // per point: the coefficient of phi_nu (c0) and of d_r phi_nu (c1..3)
            c.assign(4 * (size_t)ng, 0.0);
            for (int g = 0; g < ng; ++g) {
              const double dt = ao1[((size_t)t * nb + mu) * ng + g];
              double c0 = sg.vrho[g] * dt;
              if (s->family >= 1)
                for (int r = 0; r < 3; ++r) {
                  const double s2 = 2 * sg.vsigma[g] * sg.rho1[(size_t)r * ng + g];
                  c0 += s2 * ao2[((size_t)hidx(r, t) * nb + mu) * ng + g];
                  c[4 * (size_t)g + 1 + r] = s2 * dt;
                }
              c[4 * (size_t)g] = c0;
            }
            for (int nu = 0; nu < nb; ++nu) {
              double acc = 0;
              for (int g = 0; g < ng; ++g) {
                double v = c[4 * (size_t)g] * ao[(size_t)nu * ng + g];
                if (s->family >= 1)
                  for (int r = 0; r < 3; ++r) v += c[4 * (size_t)g + 1 + r] * ao1[((size_t)r * nb + nu) * ng + g];
                acc += sg.w[g] * v;
              }
              M[(size_t)t * nb2 + (size_t)nu * nb + mu] += acc;
            }
          }
      }
    }
    std::vector<double> Mt(3 * nb2, 0.0);
    for (int k = 0; k < nt; ++k)
      for (size_t q = 0; q < Mt.size(); ++q) Mt[q] += part[k][q];
    std::memset(Fx, 0, sizeof(double) * npert * nb2);
    for (int A = 0; A < na; ++A)
      for (int t = 0; t < 3; ++t) {
        double* F = Fx + (size_t)(3 * A + t) * nb2;
        const double* Ms = Mt.data() + (size_t)t * nb2;  // column-major: Ms[nu * nb + mu] = M[mu][nu]
        for (int nu = 0; nu < nb; ++nu)
          for (int mu = 0; mu < nb; ++mu) {
            double v = 0;
            if (ao_atom[mu] == A) v += Ms[(size_t)nu * nb + mu];
            if (ao_atom[nu] == A) v += Ms[(size_t)mu * nb + nu];
            F[(size_t)nu * nb + mu] = -v;
          }
      }
  });
}

// Moves the basis functions of one atom (not its grid); the AO tables must be
// rebuilt with ox_get_ao. Used by finite-difference checks of ox_xc_grad.
int ox_shift_basis(ox_system* s, int atom, double dx, double dy, double dz) {
  return guarded([&] {
    if (atom < 0 || atom >= s->natoms) throw std::runtime_error("oracle: bad atom");
    s->xyz[3 * atom] += dx; s->xyz[3 * atom + 1] += dy; s->xyz[3 * atom + 2] += dz;
    s->ao_ready = false;
  });
}

int ox_xc(ox_system* s, const double* D, double* V, double* Exc, double* Nel) {
  if (ox_density(s, D)) return -1;
  if (ox_evaluate(s)) return -1;
  if (ox_energy(s, Exc)) return -1;
  if (Nel && ox_num_electrons(s, Nel)) return -1;
  return ox_fock(s, V);
}

int ox_eval_functional(int ncodes, const int* codes, long n,
                       const double* rho, const double* sigma, const double* lapl, const double* tau,
                       double* exc, double* vrho, double* vsigma, double* vlapl, double* vtau) {
  (void)lapl;
  for (long i = 0; i < n; ++i) {
    double e = 0, vr = 0, vs = 0, vt = 0;
    for (int k = 0; k < ncodes; ++k)
      oxf::eval_point(codes[k], rho[i], sigma ? sigma[i] : 0.0, tau ? tau[i] : 0.0, &e, &vr, &vs, &vt);
    if (exc) exc[i] = e;
    if (vrho) vrho[i] = vr;
    if (vsigma) vsigma[i] = vs;
    if (vlapl) vlapl[i] = 0;
    if (vtau) vtau[i] = vt;
  }
  return 0;
}

int ox_eval_functional_f(int ncodes, const int* codes, long n, const double* rho, const double* sigma,
                         double* vsigma, double* v2rho2, double* v2rhosigma, double* v2sigma2) {
  for (long i = 0; i < n; ++i) {
    double vs = 0, frr = 0, frs = 0, fss = 0;
    for (int k = 0; k < ncodes; ++k) {
      if (oxf::family_of(codes[k]) > 1) { g_err = "oracle: second derivatives are LDA/GGA"; return -1; }
      oxf::eval_point_f(codes[k], rho[i], sigma ? sigma[i] : 0.0, &vs, &frr, &frs, &fss);
    }
    if (vsigma) vsigma[i] = vs;
    if (v2rho2) v2rho2[i] = frr;
    if (v2rhosigma) v2rhosigma[i] = frs;
    if (v2sigma2) v2sigma2[i] = fss;
  }
  return 0;
}

int ox_ao_values(const ox_system* s, long n, const double* xyz, double* ao, double* grad, double* lapl) {
  return guarded([&] {
    SubGrid sg;
    sg.npad = (int)n;
    sg.x.resize(n); sg.y.resize(n); sg.z.resize(n);
    for (long i = 0; i < n; ++i) { sg.x[i] = xyz[3 * i]; sg.y[i] = xyz[3 * i + 1]; sg.z[i] = xyz[3 * i + 2]; }
    compute_ao(const_cast<ox_system*>(s), sg, 2);
    std::memcpy(ao, sg.ao.data(), sizeof(double) * sg.ao.size());
    if (grad) std::memcpy(grad, sg.ao1.data(), sizeof(double) * sg.ao1.size());
    if (lapl) std::memcpy(lapl, sg.ao2l.data(), sizeof(double) * sg.ao2l.size());
  });
}

long ox_active_points(const ox_system* s) {
  long n = 0;
  for (int i : s->active) n += s->subgrids[i].n;
  return n;
}

int ox_get_active_xyz(const ox_system* s, double* xyz) {
  long k = 0;
  for (int i : s->active) {
    const SubGrid& sg = s->subgrids[i];
    for (int g = 0; g < sg.n; ++g, ++k) {
      xyz[3 * k] = sg.x[g];
      xyz[3 * k + 1] = sg.y[g];
      xyz[3 * k + 2] = sg.z[g];
    }
  }
  return 0;
}

int ox_get_active_raw_weights(const ox_system* s, double* w) {
  long k = 0;
  for (int i : s->active) {
    const SubGrid& sg = s->subgrids[i];
    for (int g = 0; g < sg.n; ++g, ++k) w[k] = s->raw_w[sg.first + g];
  }
  return 0;
}

int ox_get_density(const ox_system* s, double* rho, double* grad, double* sigma, double* tau, double* w) {
  long k = 0;
  for (int i : s->active) {
    const SubGrid& sg = s->subgrids[i];
    for (int g = 0; g < sg.n; ++g, ++k) {
      if (rho) rho[k] = sg.rho[g];
      if (grad && !sg.rho1.empty()) for (int r = 0; r < 3; ++r) grad[3 * k + r] = sg.rho1[(size_t)r * sg.npad + g];
      if (sigma && !sg.sigma.empty()) sigma[k] = sg.sigma[g];
      if (tau && !sg.tau.empty()) tau[k] = sg.tau[g];
      if (w) w[k] = sg.w[g];
    }
  }
  return 0;
}

}  // extern "C"
