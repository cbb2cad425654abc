// A caller written against the reference's class interface, as the SCF does
// on every iteration (src/HartreeFockKohnSham/Restricted/SP.cpp:52-55), built
// on include/chinium/xc.hpp. Usage:
//   xc_step <grid> <basis> <functional> <geometry.txt> <D.bin> <out.bin>
// geometry.txt: one "Z x y z" line per atom (bohr); D.bin: nbasis^2 doubles
// (column-major); out.bin receives E_xc, N_e, then V (column-major); for
// LDA/GGA also the response matrix of D^x = D, then the spin-unrestricted
// E_xc and V_alpha, V_beta of Da = 0.6 D, Db = 0.4 D.
#include <cstdio>
#include <fstream>
#include <vector>

#include "chinium/xc.hpp"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: xc_step grid basis functional geometry D out\n");
    return 2;
  }
  std::vector<int> Z;
  std::vector<double> xyz;
  std::ifstream g(argv[4]);
  int z;
  double x, y, w;
  while (g >> z >> x >> y >> w) {
    Z.push_back(z);
    xyz.insert(xyz.end(), {x, y, w});
  }
  try {
    chinium::Molecule mol(Z, xyz);
    mol.setBasis(argv[2]);
    const int n = mol.getNumBasis();
    chinium::ExchangeCorrelation xc;
    xc.Read(argv[3], false);
    chinium::Grid grid(&mol, argv[1], 1, 0);
    grid.setType(xc.Family == "LDA" ? 0 : (xc.Family == "GGA" ? 1 : 2));
    grid.getAO(0, 0);
    chinium::EigenMatrix D = chinium::make_matrix(n, n);
    std::ifstream df(argv[5], std::ios::binary);
    df.read(reinterpret_cast<char*>(D.data()), sizeof(double) * n * n);
    grid.getDensity({D});
    xc.Evaluate("ev", grid);
    const double E = grid.getEnergy();
    const double N = grid.getNumElectrons();
    chinium::EigenMatrix V = grid.getFock()[0];
    std::ofstream out(argv[6], std::ios::binary);
    out.write(reinterpret_cast<const char*>(&E), sizeof(double));
    out.write(reinterpret_cast<const char*>(&N), sizeof(double));
    out.write(reinterpret_cast<const char*>(V.data()), sizeof(double) * n * n);
    if (xc.Family != "mGGA") {  // response to the perturbation D^x = D, as RestrictedFinite/SP.cpp:444-455 does
      xc.Evaluate("f", grid);
      grid.getDensityU({{D}});
      chinium::EigenMatrix F = grid.getFockU()[0][0];
      out.write(reinterpret_cast<const char*>(F.data()), sizeof(double) * n * n);
      // skeleton response per nuclear coordinate, in the Hessian's order (Restricted/Hess.cpp:78-96)
      grid.getAO(2, 1);
      grid.getDensitySkeleton({D});
      const std::vector<std::vector<chinium::EigenMatrix>> Fs = grid.getFockU<chinium::s_t>();
      for (const chinium::EigenMatrix& m : Fs[0]) out.write(reinterpret_cast<const char*>(m.data()), sizeof(double) * n * n);
      // the Hessian's other XC terms (Restricted/Hess.cpp:83-94)
      grid.getDensitySkeleton2({D});
      const std::vector<std::vector<double>> H = grid.getEnergyHess();
      for (const std::vector<double>& row : H) out.write(reinterpret_cast<const char*>(row.data()), sizeof(double) * row.size());
      const std::vector<std::vector<chinium::EigenMatrix>> Fk = grid.getFockSkeleton();
      for (const chinium::EigenMatrix& m : Fk[0]) out.write(reinterpret_cast<const char*>(m.data()), sizeof(double) * n * n);
      // spin-unrestricted call sequence (Unrestricted/SP.cpp:65-68) with Da = 0.6 D, Db = 0.4 D
      chinium::ExchangeCorrelation uxc;
      uxc.Spin = 2;
      uxc.Read(argv[3], false);
      chinium::EigenMatrix Da = chinium::make_matrix(n, n), Db = chinium::make_matrix(n, n);
      for (int k = 0; k < n * n; ++k) {
        Da.data()[k] = 0.6 * D.data()[k];
        Db.data()[k] = 0.4 * D.data()[k];
      }
      grid.getDensity({Da, Db});
      uxc.Evaluate("ev", grid);
      const double Eu = grid.getEnergy();
      const std::vector<chinium::EigenMatrix> Vu = grid.getFock();
      out.write(reinterpret_cast<const char*>(&Eu), sizeof(double));
      for (const chinium::EigenMatrix& m : Vu) out.write(reinterpret_cast<const char*>(m.data()), sizeof(double) * n * n);
    }
    std::printf("E_xc %.12f N_e %.10f\n", E, N);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
