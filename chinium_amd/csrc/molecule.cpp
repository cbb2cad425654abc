// Molecule and basis: .gbs reading, libint2-convention normalisation and the
// analytic overlap. These replace the Mwfn/libint2 pieces the XC path reads
// (legacy/MwfnIO/Mwfn.cpp:606-712; src/Integral/Normalization.cpp:11-18).
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>

#include "internal.hpp"

namespace chx {

int shell_size(int type) { return type == 1 ? 3 : 2 * std::abs(type) + 1; }

namespace {

const char* kSymbols[] = {"", "H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE", "NA", "MG", "AL", "SI", "P",
                          "S", "CL", "AR", "K", "CA", "SC", "TI", "V", "CR", "MN", "FE", "CO", "NI", "CU", "ZN",
                          "GA", "GE", "AS", "SE", "BR", "KR"};
constexpr int kMaxZ = 36;

int symbol_to_z(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  for (int z = 1; z <= kMaxZ; ++z)
    if (s == kSymbols[z]) return z;
  return -1;
}

double to_double(std::string t) {
  for (char& c : t)
    if (c == 'D' || c == 'd') c = 'E';
  return std::stod(t);
}

struct RawShell {
  int type;
  std::vector<double> e, c;
};

// Gaussian-94 basis file: "-El 0" opens an element, "<L> <nprim> <scale>"
// opens a shell (SP gives an s and a Cartesian p shell sharing exponents, P is
// Cartesian, D and above are pure), "****" closes the element.
std::map<int, std::vector<RawShell>> parse_gbs(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw Error(CHX_ERR_FILE, "Basis set file is missing in $CHINIUM_PATH/BasisSets/ ! (" + path + ")");
  std::map<int, std::vector<RawShell>> lib;
  std::string line;
  int z = -1;
  std::vector<RawShell> cur;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    std::string w;
    if (!(ss >> w) || w[0] == '!') continue;
    if (w[0] == '-') {
      z = symbol_to_z(w.substr(1));
      cur.clear();
      continue;
    }
    if (w == "****") {
      if (z > 0) lib[z] = cur;
      z = -1;
      continue;
    }
    int type;
    if (w == "S" || w == "SP") type = 0;
    else if (w == "P") type = 1;
    else if (w == "D") type = -2;
    else if (w == "F") type = -3;
    else if (w == "G") type = -4;
    else if (w == "H") type = -5;
    else if (w == "I") type = -6;
    else throw Error(CHX_ERR_FILE, "Unrecognized symbol of angular momentum!");
    std::string ntok;
    ss >> ntok;
    const int n = std::stoi(ntok);
    RawShell a{type, {}, {}}, b{1, {}, {}};
    for (int i = 0; i < n; ++i) {
      if (!std::getline(f, line)) throw Error(CHX_ERR_FILE, "Corrupted basis set file");
      std::stringstream ls(line);
      std::string te, tc, tc2;
      ls >> te >> tc;
      a.e.push_back(to_double(te));
      a.c.push_back(to_double(tc));
      if (w == "SP") {
        ls >> tc2;
        b.e.push_back(a.e.back());
        b.c.push_back(to_double(tc2));
      }
    }
    cur.push_back(a);
    if (w == "SP") cur.push_back(b);
  }
  return lib;
}

double double_factorial(int n) {
  double r = 1;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// Primitive factors [2^l (2a)^(l+3/2) / (pi^(3/2) (2l-1)!!)]^(1/2), then unit
// self-overlap of the contracted x^l component (the libint2 convention).
void normalise(int l, const std::vector<double>& e, std::vector<double>& c) {
  const double pi32 = std::pow(M_PI, 1.5), df = double_factorial(2 * l - 1), two_l = std::ldexp(1.0, l);
  for (size_t p = 0; p < e.size(); ++p) c[p] *= std::sqrt(two_l * std::pow(2 * e[p], l + 1.5) / (pi32 * df));
  double s = 0;
  for (size_t p = 0; p < e.size(); ++p)
    for (size_t q = 0; q < e.size(); ++q) s += c[p] * c[q] * df * pi32 / (two_l * std::pow(e[p] + e[q], l + 1.5));
  const double k = 1.0 / std::sqrt(s);
  for (double& v : c) v *= k;
}

// Angular functions as sums of monomials c x^i y^j z^k, in the reference's
// component order (AO/S.hpp, CartP.hpp, PureP.hpp, PureD.hpp, PureF.hpp).
struct Mono { double c; int i, j, k; };
std::vector<std::vector<Mono>> monomials(int type) {
  const double r3 = std::sqrt(3.0), r6 = std::sqrt(6.0), r10 = std::sqrt(10.0), r15 = std::sqrt(15.0);
  switch (type) {
    case 0: return {{{1, 0, 0, 0}}};
    case 1: return {{{1, 1, 0, 0}}, {{1, 0, 1, 0}}, {{1, 0, 0, 1}}};
    case -1: return {{{1, 0, 1, 0}}, {{1, 0, 0, 1}}, {{1, 1, 0, 0}}};
    case -2:
      return {{{r3, 1, 1, 0}}, {{r3, 0, 1, 1}}, {{1, 0, 0, 2}, {-0.5, 2, 0, 0}, {-0.5, 0, 2, 0}},
              {{r3, 1, 0, 1}}, {{r3 / 2, 2, 0, 0}, {-r3 / 2, 0, 2, 0}}};
    case -3:
      return {{{0.75 * r10, 2, 1, 0}, {-0.25 * r10, 0, 3, 0}},
              {{r15, 1, 1, 1}},
              {{r6, 0, 1, 2}, {-0.25 * r6, 2, 1, 0}, {-0.25 * r6, 0, 3, 0}},
              {{1, 0, 0, 3}, {-1.5, 2, 0, 1}, {-1.5, 0, 2, 1}},
              {{r6, 1, 0, 2}, {-0.25 * r6, 3, 0, 0}, {-0.25 * r6, 1, 2, 0}},
              {{0.5 * r15, 2, 0, 1}, {-0.5 * r15, 0, 2, 1}},
              {{0.25 * r10, 3, 0, 0}, {-0.75 * r10, 1, 2, 0}}};
    default: throw Error(CHX_ERR_UNSUPPORTED, "angular momentum above f is not supported");
  }
}

// 1-D Obara-Saika overlap table S[i][j] of (x-A)^i and (x-B)^j Gaussians.
void os_1d(double a, double b, double A, double B, int li, int lj, double* S /* (li+1) x (lj+1) */) {
  const double p = a + b, P = (a * A + b * B) / p, XPA = P - A, XPB = P - B, o2p = 0.5 / p;
  const int W = lj + 1;
  S[0] = std::sqrt(M_PI / p) * std::exp(-a * b / p * (A - B) * (A - B));
  for (int i = 0; i <= li; ++i)
    for (int j = 0; j <= lj; ++j) {
      if (i == 0 && j == 0) continue;
      double v;
      if (i > 0) {
        v = XPA * S[(i - 1) * W + j];
        if (i > 1) v += o2p * (i - 1) * S[(i - 2) * W + j];
        if (j > 0) v += o2p * j * S[(i - 1) * W + j - 1];
      } else {
        v = XPB * S[i * W + j - 1];
        if (j > 1) v += o2p * (j - 1) * S[i * W + j - 2];
      }
      S[i * W + j] = v;
    }
}

}  // namespace

std::vector<std::vector<MonoTerm>> component_monomials(int type) {
  std::vector<std::vector<MonoTerm>> out;
  for (const auto& comp : monomials(type)) {
    out.emplace_back();
    for (const Mono& t : comp) out.back().push_back(MonoTerm{t.c, t.i, t.j, t.k});
  }
  return out;
}

std::vector<AuxShell> aux_basis_shells(const chx_mol* m, const std::string& data_root, const std::string& basis) {
  auto lib = parse_gbs(data_root + "/BasisSets/" + basis + ".gbs");
  std::vector<AuxShell> out;
  for (int a = 0; a < m->natoms; ++a) {
    auto it = lib.find(m->Z[a]);
    if (it == lib.end()) throw Error(CHX_ERR_FILE, "Basis set file does not include this element!");
    for (const RawShell& rs : it->second) {
      const int l = rs.type == 1 ? 1 : std::abs(rs.type);
      if (l > 4) throw Error(CHX_ERR_UNSUPPORTED, "auxiliary shells above g are not supported");
      std::vector<double> c = rs.c;
      normalise(l, rs.e, c);
      out.push_back(AuxShell{l, a, rs.e, c});
    }
  }
  return out;
}

}  // namespace chx

using namespace chx;

void mol_set_basis_file(chx_mol* m, const std::string& data_root, const std::string& basis) {
  auto lib = parse_gbs(data_root + "/BasisSets/" + basis + ".gbs");
  m->shells.clear();
  m->exps.clear();
  m->coefs.clear();
  m->nbasis = 0;
  for (int a = 0; a < m->natoms; ++a) {
    auto it = lib.find(m->Z[a]);
    if (it == lib.end()) throw Error(CHX_ERR_FILE, "Basis set file does not include this element!");
    for (const RawShell& rs : it->second) {
      if (rs.type < -3) throw Error(CHX_ERR_UNSUPPORTED, "angular momentum above f is not supported");
      std::vector<double> c = rs.c;
      normalise(std::abs(rs.type), rs.e, c);
      ShellH sh{rs.type, a, (int)rs.e.size(), (int)m->exps.size(), m->nbasis};
      m->exps.insert(m->exps.end(), rs.e.begin(), rs.e.end());
      m->coefs.insert(m->coefs.end(), c.begin(), c.end());
      m->shells.push_back(sh);
      m->nbasis += shell_size(rs.type);
    }
  }
}

void mol_overlap(const chx_mol* m, double* S) {
  const int n = m->nbasis;
  for (size_t a = 0; a < m->shells.size(); ++a)
    for (size_t b = 0; b < m->shells.size(); ++b) {
      const ShellH& A = m->shells[a];
      const ShellH& B = m->shells[b];
      const auto ma = monomials(A.type), mb = monomials(B.type);
      const int la = A.type == 1 ? 1 : std::abs(A.type), lb = B.type == 1 ? 1 : std::abs(B.type);
      const double* ca = &m->xyz[3 * A.center];
      const double* cb = &m->xyz[3 * B.center];
      std::vector<double> blk(ma.size() * mb.size(), 0.0);
      std::vector<double> Sx((la + 1) * (lb + 1)), Sy(Sx.size()), Sz(Sx.size());
      for (int p = 0; p < A.nprim; ++p)
        for (int q = 0; q < B.nprim; ++q) {
          const double ea = m->exps[A.prim_off + p], eb = m->exps[B.prim_off + q];
          const double cc = m->coefs[A.prim_off + p] * m->coefs[B.prim_off + q];
          os_1d(ea, eb, ca[0], cb[0], la, lb, Sx.data());
          os_1d(ea, eb, ca[1], cb[1], la, lb, Sy.data());
          os_1d(ea, eb, ca[2], cb[2], la, lb, Sz.data());
          const int W = lb + 1;
          for (size_t u = 0; u < ma.size(); ++u)
            for (size_t v = 0; v < mb.size(); ++v) {
              double s = 0;
              for (const Mono& x : ma[u])
                for (const Mono& y : mb[v])
                  s += x.c * y.c * Sx[x.i * W + y.i] * Sy[x.j * W + y.j] * Sz[x.k * W + y.k];
              blk[u * mb.size() + v] += cc * s;
            }
        }
      for (size_t u = 0; u < ma.size(); ++u)
        for (size_t v = 0; v < mb.size(); ++v)
          S[(size_t)(B.ao_off + v) * n + (A.ao_off + u)] = blk[u * mb.size() + v];
    }
}
