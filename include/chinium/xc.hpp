// chinium/xc.hpp - header-only C++ drop-in for the reference's Grid and
// ExchangeCorrelation classes (src/Grid/Grid.h:90-116, src/ExchangeCorrelation.h:8-17),
// implemented over the C ABI in chinium_xc.h (libchinium_xc.so, gfx950).
//
// Same member names, argument meaning, call order and failure behaviour:
// failures throw std::runtime_error carrying the reference's message
// (e.g. "Density functional file is missing!"), data files are found through
// $CHINIUM_PATH. The per-iteration protocol is the reference's
// (Restricted/SP.cpp:52-55):
//     grid.getDensity({2 * D}); xc.Evaluate("ev", grid);
//     double E = grid.getEnergy(); EigenMatrix V = grid.getFock()[0];
//
// Matrices: with Eigen present, EigenMatrix is Eigen::MatrixXd as in the
// reference (src/Macro/Abbreviation.h:4); otherwise a minimal column-major
// matrix with the same data()/rows()/cols() accessors.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../chinium_xc.h"

#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
namespace chinium {
using EigenMatrix = Eigen::MatrixXd;
inline EigenMatrix make_matrix(int r, int c) { return EigenMatrix::Zero(r, c); }
}  // namespace chinium
#else
namespace chinium {
class EigenMatrix {
 public:
  EigenMatrix() = default;
  EigenMatrix(int r, int c) : r_(r), c_(c), d_((size_t)r * c, 0.0) {}
  int rows() const { return r_; }
  int cols() const { return c_; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }
  double& operator()(int i, int j) { return d_[(size_t)j * r_ + i]; }
  double operator()(int i, int j) const { return d_[(size_t)j * r_ + i]; }

 private:
  int r_ = 0, c_ = 0;
  std::vector<double> d_;
};
inline EigenMatrix make_matrix(int r, int c) { return EigenMatrix(r, c); }
}  // namespace chinium
#endif

namespace chinium {

inline void check(int rc) {
  if (rc != CHX_OK) throw std::runtime_error(chx_last_error());
}

// Which perturbed density Grid::getFockU contracts (src/Grid/Grid.h): the
// skeleton derivatives of getDensitySkeleton (s_t) or getDensityU's (u_t).
enum D_t { s_t, u_t };

inline std::string data_root() {
  const char* p = std::getenv("CHINIUM_PATH");
  if (!p) throw std::runtime_error("CHINIUM_PATH is not set");
  return p;
}

// The Mwfn fields the XC path reads: centres (Z, bohr) and normalised shells.
class Molecule {
 public:
  Molecule(const std::vector<int>& Z, const std::vector<double>& xyz_bohr) : natoms_((int)Z.size()) {
    check(chx_mol_create((int)Z.size(), Z.data(), xyz_bohr.data(), &h_));
  }
  int getNumCenters() const { return natoms_; }
  // Mwfn::setBasis + Normalize (legacy/MwfnIO/Mwfn.cpp:686; src/Integral/Normalization.cpp:11).
  void setBasis(const std::string& name) { check(chx_mol_set_basis_file(h_, data_root().c_str(), name.c_str())); }
  // One MwfnShell: Type, Exponents, NormalizedCoefficients.
  void addShell(int center, int type, const std::vector<double>& exps, const std::vector<double>& norm_coefs) {
    check(chx_mol_add_shell(h_, center, type, (int)exps.size(), exps.data(), norm_coefs.data()));
  }
  int getNumBasis() const {
    int n = 0;
    check(chx_mol_nbasis(h_, &n));
    return n;
  }
  const chx_mol* handle() const { return h_; }
  ~Molecule() { chx_mol_destroy(h_); }
  Molecule(const Molecule&) = delete;
  Molecule& operator=(const Molecule&) = delete;

 private:
  chx_mol* h_ = nullptr;
  int natoms_ = 0;
};

class Grid {
 public:
  Grid() = default;
  // Grid::Grid(Mwfn*, std::string grid, int nthreads, int output)
  Grid(const Molecule* mol, std::string grid, int nthreads, int output)
      : nbasis_(mol->getNumBasis()), natoms_(mol->getNumCenters()) {
    if (grid.empty()) return;  // as the reference: an empty name builds nothing
    check(chx_grid_create(mol->handle(), grid.c_str(), data_root().c_str(), nthreads, output, &h_));
  }
  Grid(const Grid& o) : nbasis_(o.nbasis_), natoms_(o.natoms_) {
    if (o.h_) check(chx_grid_copy(o.h_, &h_));
  }
  Grid& operator=(const Grid& o) {
    if (this != &o) {
      Grid tmp(o);
      std::swap(h_, tmp.h_);
      std::swap(nbasis_, tmp.nbasis_);
      std::swap(natoms_, tmp.natoms_);
    }
    return *this;
  }
  ~Grid() { chx_grid_destroy(h_); }

  int getNumThreads() const {
    int n = 0;
    check(chx_grid_num_threads(h_, &n));
    return n;
  }
  void setType(int type) { check(chx_grid_set_type(h_, type)); }
  void getAO(int derivative, int output) { check(chx_grid_get_ao(h_, derivative, output)); }
  void getDensity(std::vector<EigenMatrix> D) {
    const int nspin = (int)D.size();
    std::vector<double> buf;
    buf.reserve((size_t)nspin * nbasis_ * nbasis_);
    for (const EigenMatrix& m : D) buf.insert(buf.end(), m.data(), m.data() + (size_t)nbasis_ * nbasis_);
    check(chx_grid_get_density(h_, buf.data(), nspin, nbasis_));
    nspin_ = nspin;
  }
  double getNumElectrons() {
    double n = 0;
    check(chx_grid_num_electrons(h_, &n));
    return n;
  }
  double getEnergy() {
    double e = 0;
    check(chx_grid_energy(h_, &e));
    return e;
  }
  // One matrix per spin of the last getDensity (Grid.cpp:156-171).
  std::vector<EigenMatrix> getFock() {
    std::vector<double> buf((size_t)nspin_ * nbasis_ * nbasis_);
    check(chx_grid_fock(h_, buf.data(), nspin_, nbasis_));
    std::vector<EigenMatrix> F(nspin_, make_matrix(nbasis_, nbasis_));
    for (int s = 0; s < nspin_; ++s)
      std::copy(buf.begin() + (size_t)s * nbasis_ * nbasis_, buf.begin() + (size_t)(s + 1) * nbasis_ * nbasis_,
                F[s].data());
    return F;
  }
  // SubGrid::getDensitySkeleton / Grid::getEnergyGrad (GridDensity.cpp:79-130,
  // GridEnergy.cpp:15-30): XC nuclear gradient, 3 * natoms, [atom][x, y, z].
  void getDensitySkeleton(std::vector<EigenMatrix> D) {
    const int nspin = (int)D.size();
    std::vector<double> buf;
    buf.reserve((size_t)nspin * nbasis_ * nbasis_);
    for (const EigenMatrix& m : D) buf.insert(buf.end(), m.data(), m.data() + (size_t)nbasis_ * nbasis_);
    check(chx_grid_get_density_skeleton(h_, buf.data(), nspin, nbasis_));
  }
  std::vector<double> getEnergyGrad() {
    std::vector<double> g(3 * (size_t)natoms_);
    check(chx_grid_energy_grad(h_, g.data()));
    return g;
  }
  // Response path, LDA/GGA, FP64 form, nspin 1 or 2: Grid::getDensityU
  // (src/Grid/Grid.cpp:63-78) takes D[spin][mat]; Grid::getFockU<u_t>
  // (src/Grid/Grid.cpp:193-216) returns F[spin][mat]. Evaluate("f") first.
  void getDensityU(std::vector<std::vector<EigenMatrix>> D) {
    const int nspin = (int)D.size(), nmats = nspin ? (int)D[0].size() : 0;
    for (const auto& spin : D)
      if ((int)spin.size() != nmats) throw std::runtime_error("Inconsistent number of spin types!");
    std::vector<double> buf;
    buf.reserve((size_t)nspin * nmats * nbasis_ * nbasis_);
    for (const auto& spin : D)
      for (const EigenMatrix& m : spin) buf.insert(buf.end(), m.data(), m.data() + (size_t)nbasis_ * nbasis_);
    check(chx_grid_get_density_u(h_, buf.data(), nmats, nspin, nbasis_));
    nmats_u_ = nmats;
    nspin_u_ = nspin;
  }
  // Grid::getDensitySkeleton2 / getEnergyHess (Restricted/Hess.cpp:83-88).
  void getDensitySkeleton2(std::vector<EigenMatrix> D) {
    check(chx_grid_get_density_skeleton2(h_, D.at(0).data(), 1, nbasis_));
  }
  std::vector<std::vector<double>> getEnergyHess() {
    const int m = 3 * natoms_;
    std::vector<double> buf((size_t)m * m);
    check(chx_grid_energy_hess(h_, buf.data()));
    std::vector<std::vector<double>> H(m);
    for (int k = 0; k < m; ++k) H[k].assign(buf.begin() + (size_t)k * m, buf.begin() + (size_t)(k + 1) * m);
    return H;
  }
  // Grid::getFockSkeleton (Restricted/Hess.cpp:94): one matrix per nuclear
  // coordinate, [spin][3 atom + t], the potential held fixed.
  std::vector<std::vector<EigenMatrix>> getFockSkeleton() {
    const int nmats = 3 * natoms_;
    std::vector<double> buf((size_t)nmats * nbasis_ * nbasis_);
    check(chx_grid_fock_skeleton(h_, buf.data(), 1, nbasis_));
    std::vector<std::vector<EigenMatrix>> F(1);
    for (int m = 0; m < nmats; ++m) {
      F[0].push_back(make_matrix(nbasis_, nbasis_));
      std::copy(buf.begin() + (size_t)m * nbasis_ * nbasis_, buf.begin() + (size_t)(m + 1) * nbasis_ * nbasis_,
                F[0].back().data());
    }
    return F;
  }
  // D_t = u_t: the matrices of the last getDensityU. D_t = s_t (the Hessian's
  // skeleton term, Restricted/Hess.cpp:96): one matrix per nuclear coordinate,
  // [3 atom + t], from the last getDensitySkeleton.
  template <D_t d_t = u_t>
  std::vector<std::vector<EigenMatrix>> getFockU() {
    const int nmats = d_t == s_t ? 3 * natoms_ : nmats_u_;
    const int nspin = d_t == s_t ? 1 : nspin_u_;
    std::vector<double> buf((size_t)nspin * nmats * nbasis_ * nbasis_);
    if (d_t == s_t)
      check(chx_grid_fock_s(h_, buf.data(), 1, nbasis_));
    else
      check(chx_grid_fock_u(h_, buf.data(), nmats, nspin, nbasis_));
    std::vector<std::vector<EigenMatrix>> F(nspin);
    for (int u = 0; u < nspin; ++u)
      for (int m = 0; m < nmats; ++m) {
        const size_t k = (size_t)u * nmats + m;
        F[u].push_back(make_matrix(nbasis_, nbasis_));
        std::copy(buf.begin() + k * nbasis_ * nbasis_, buf.begin() + (k + 1) * nbasis_ * nbasis_,
                  F[u].back().data());
      }
    return F;
  }
  // GPU options (no reference counterpart): CHX_PREC_FP64 or CHX_PREC_MIXED.
  void setPrecision(int precision) { check(chx_grid_set_precision(h_, precision)); }
  chx_grid* handle() { return h_; }

 private:
  chx_grid* h_ = nullptr;
  int nbasis_ = 0;
  int natoms_ = 0;
  int nmats_u_ = 0, nspin_u_ = 1;
  int nspin_ = 1;  // spin count of the last getDensity
};

class ExchangeCorrelation {
 public:
  std::vector<int> Codes;
  double EXX = 1;
  int Spin = 1;  // 1 - Unpolarized, 2 - Polarized
  std::string Family;

  void Read(std::string df, bool output) {
    chx_xc* h = nullptr;
    check(chx_xc_read(data_root().c_str(), df.c_str(), Spin, output ? 1 : 0, &h));
    reset(h);
  }
  void Evaluate(std::string order, Grid& grid) { check(chx_xc_evaluate(h_, grid.handle(), order.c_str())); }
  explicit operator bool() const { return Codes.size() > 0; }

  ExchangeCorrelation() = default;
  ExchangeCorrelation(const ExchangeCorrelation& o) : Codes(o.Codes), EXX(o.EXX), Spin(o.Spin), Family(o.Family) {
    if (!Codes.empty()) {
      chx_xc* h = nullptr;
      check(chx_xc_create((int)Codes.size(), Codes.data(), Spin, &h));
      h_ = h;
    }
  }
  ExchangeCorrelation& operator=(const ExchangeCorrelation& o) {
    if (this != &o) {
      ExchangeCorrelation tmp(o);
      std::swap(h_, tmp.h_);
      Codes = o.Codes;
      EXX = o.EXX;
      Spin = o.Spin;
      Family = o.Family;
    }
    return *this;
  }
  ~ExchangeCorrelation() { chx_xc_destroy(h_); }

 private:
  chx_xc* h_ = nullptr;
  void reset(chx_xc* h) {
    chx_xc_destroy(h_);
    h_ = h;
    int family = 0, n = 0;
    check(chx_xc_info(h_, &family, &EXX, &n, nullptr));
    Codes.assign(n, 0);
    check(chx_xc_info(h_, nullptr, nullptr, nullptr, Codes.data()));
    static const char* names[] = {"LDA", "GGA", "mGGA"};
    Family = names[family];
  }
};

}  // namespace chinium
