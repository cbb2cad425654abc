// Density-fitted Coulomb (RI-J) engine: host/device shared layout.
#pragma once
#include "internal.hpp"

namespace chx {

constexpr int kJMaxLOrb = 2;                          // orbital shells up to d
constexpr int kJMaxLAux = 4;                          // auxiliary shells up to g
constexpr int kJMaxL = 2 * kJMaxLOrb + kJMaxLAux;     // highest Hermite order of a 3-centre class
constexpr int kJMaxLMetric = 2 * kJMaxLAux;           // highest order of a 2-centre (metric) class
constexpr int kBoysTaylor = 8;                        // Taylor terms of the tabulated Boys function
constexpr int kBoysOrders = kJMaxLMetric + kBoysTaylor + 1;  // F_0 .. F_16 tabulated
constexpr double kBoysStep = 0.05;
constexpr double kBoysTmax = 40.0;                    // asymptotic form above
constexpr int kBoysPts = 802;                         // T = 0, 0.05, ..., 40.05
constexpr int kJCart = 15;                            // Cartesian components of a g shell
constexpr int kJChunkRows = 128;                      // pair chunks of the gamma pass (partial sums)

__host__ __device__ constexpr int jncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int jnherm(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }
// flat index of the Hermite (or Cartesian) exponent triple (t, u, v): shells of
// s = t + u + v in increasing order, t descending, then u descending
__host__ __device__ constexpr int jhidx(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
// position of (i, j, k), i + j + k = l, within its degree-l shell
__host__ __device__ constexpr int jcidx(int i, int j, int k) { return (j + k) * (j + k + 1) / 2 + k; }

// Everything the kernels read, by value.
struct JDev {
  // orbital primitive pairs, sorted by class L = la + lb (la >= lb)
  const double *px, *py, *pz, *p, *K;   // Gaussian product centre, exponent sum, c_a c_b exp(-ab/p |A-B|^2)
  const double *pax, *pay, *paz, *pbx, *pby, *pbz;  // P - A, P - B
  const int *ta, *tb;                    // shell types (0 s, 1 Cartesian p, -1 pure p, -2 pure d)
  const int *aoa, *aob;                  // first AO of each shell (reference order)
  const unsigned char* diag;             // both primitives from one shell
  int cls[kJMaxLOrb * 2 + 2];            // first pair of each class, npairs at the end
  long hoff[kJMaxLOrb * 2 + 1];          // first Hermite-density entry of each class
  double* h;                             // Hermite density per pair (jnherm(L) entries)
  int n;                                 // basis functions
  // auxiliary primitives, sorted by l, each class padded to a multiple of 64 slots
  const double *ae, *ax, *ay, *az;       // exponent, centre
  const double* acoef;                   // (-1)^l c (2 e)^-l: Hermite coefficient of the pure components
  const int* afn;                        // first contracted function of the primitive's shell
  int aoff[kJMaxLAux + 2];               // first slot of each class, padded; total slots at the end
  int anum[kJMaxLAux + 1];               // real primitives of each class
  int nslots, naux;
  const double* auxS;                    // harmonic polynomials [l][m < 9][cart < 15]
  const double* orbT;                    // orbital component polynomials [type index][m < 5][cart < 6]
  const double* boys;                    // F_n(T_k), [k][n]
  // gamma pass
  const int4* chunks;                    // (first pair, end, class, 0) of each partial-sum row
  int nrows;
  double* gpart;                         // [row][slot][kJCart]
  // coulomb pass
  const double* gco;                     // [slot][kJCart] Hermite coefficients of the fitted density
  // nuclei (one-electron pass)
  const double* nuc;                     // x, y, z, Z
  int natoms;
};

}  // namespace chx
