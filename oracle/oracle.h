/*
 * FP64 CPU oracle for the Kohn-Sham exchange-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it; the shipped
 * library (chinium_amd) never links, loads or calls it.
 *
 * It restates, loop nest by loop nest, what the reference does on one SCF
 * iteration (src/Grid + src/ExchangeCorrelation of Chinium):
 *   grid generation      src/Grid/GridGeneration.cpp:27-162,203-230
 *   basis normalisation  src/Integral/Normalization.cpp:11-18 (libint2 renorm)
 *   AO tables            src/Grid/GridAO.cpp:14-290
 *   density              src/Grid/GridDensity.cpp:12-45
 *   functional           src/ExchangeCorrelation.cpp:165-248 (libxc replaced
 *                        by the project's own FP64 functionals, oracle side)
 *   energy               src/Grid/GridEnergy.cpp:11-13
 *   Vxc                  src/Grid/GridFock.cpp:11-49, src/Grid/Grid.cpp:156-171
 *
 * Parity status of each part is recorded in oracle/README.md.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ox_system ox_system;

/* Build molecule + basis + grid. Coordinates in bohr, row-major natoms x 3.
 * Reads <data_root>/Grids/<grid>/<El>.grid and <data_root>/BasisSets/<basis>.gbs.
 * Returns NULL on failure (message via ox_last_error). */
ox_system* ox_create(const char* data_root, const char* grid_name, const char* basis_name,
                     int natoms, const int* Z, const double* xyz_bohr, int nthreads);
void ox_destroy(ox_system* s);
const char* ox_last_error(void);

int ox_nbasis(const ox_system* s);
long ox_npoints(const ox_system* s);      /* unpadded points, generation order */
int ox_nsubgrids(const ox_system* s);
/* Points in generation order: xyzw[4*i + {0,1,2,3}]. */
int ox_get_points(ox_system* s, double* xyzw);
/* Radial x angular weights before the Becke partition, generation order. */
int ox_get_raw_weights(const ox_system* s, double* w);
/* Sub-grid sizes (unpadded) in order. */
int ox_get_subgrid_sizes(const ox_system* s, int* sizes);
/* Normalised contraction coefficients, flattened shell by shell (for tests). */
int ox_nshells(const ox_system* s);
int ox_get_shells(const ox_system* s, int* type, int* nprim, int* center, double* exps, double* coefs);

/* Lebedev rule with n points (x, y, z, w), expanded from <data_root>/lebedev_orbits.txt. */
int ox_lebedev(const char* data_root, int n, double* xyzw);

/* Functional: codes are libxc IDs. Family: 0 LDA, 1 GGA, 2 mGGA. */
int ox_set_functional_codes(ox_system* s, int ncodes, const int* codes);
int ox_set_functional_file(ox_system* s, const char* data_root, const char* df_name);
int ox_family(const ox_system* s);

/* Restrict all following work to a subset of sub-grids (for timing samples).
 * n <= 0 restores the full set. */
int ox_set_subgrid_subset(ox_system* s, int n, const int* ids);

/* Cache AO tables (value, gradient, and Laplacian when mGGA) for the active
 * sub-grids, as the reference does once per SCF (Grid::getAO(0,1)). */
int ox_get_ao(ox_system* s);

/* One XC build: getDensity(D) -> Evaluate("ev") -> getEnergy -> getFock.
 * D and V are nbasis x nbasis column-major (symmetric). */
int ox_xc(ox_system* s, const double* D, double* V, double* Exc, double* Nel);
/* XC nuclear gradient, skeleton form with the grid fixed (Restricted/Grad.cpp:72-84,
 * GridDensity.cpp:79-130, GridEnergy.cpp:15-30); grad: natoms x 3, LDA/GGA. */
int ox_xc_grad(ox_system* s, const double* D, double* grad);
/* Moves one atom's basis functions, not its grid points (finite-difference checks). */
int ox_shift_basis(ox_system* s, int atom, double dx, double dy, double dz);

/* Individual phases, for timing. */
int ox_density(ox_system* s, const double* D);
int ox_evaluate(ox_system* s);
int ox_energy(ox_system* s, double* Exc);
int ox_num_electrons(ox_system* s, double* Nel);
int ox_fock(ox_system* s, double* V);

/* Per-point functional, libxc conventions (e per particle, derivatives of rho*e).
 * Arrays may be NULL when not needed by the family. */
int ox_eval_functional(int ncodes, const int* codes, long n,
                       const double* rho, const double* sigma, const double* lapl, const double* tau,
                       double* exc, double* vrho, double* vsigma, double* vlapl, double* vtau);

/* AO values at arbitrary points (n points): ao[mu*n + i], grad[(r*nbasis + mu)*n + i],
 * lapl[mu*n + i]. grad/lapl may be NULL. */
int ox_ao_values(const ox_system* s, long n, const double* xyz, double* ao, double* grad, double* lapl);
/* Second derivatives of the summed LDA/GGA codes at n points (libxc xc_*_fxc):
 * v_sigma, v2rho2, v2rhosigma, v2sigma2 (any output may be NULL). */
int ox_eval_functional_f(int ncodes, const int* codes, long n, const double* rho, const double* sigma,
                         double* vsigma, double* v2rho2, double* v2rhosigma, double* v2sigma2);
/* Response Fock matrices of nmats perturbed density matrices (column-major
 * nbasis^2 each, back to back), closed shell, LDA/GGA, after ox_density at the
 * reference density: Evaluate("f"), getDensityU, getFockU<u_t>. */
int ox_fock_u(ox_system* s, int nmats, const double* Dx, double* Fx);
/* Skeleton XC response (getDensitySkeleton, Evaluate("f"), Grid::getFockU<s_t>):
 * the response matrix of each nuclear coordinate's skeleton density derivative,
 * closed shell, LDA/GGA; D is the ground-state density. Fx: 3 natoms
 * column-major nbasis^2 matrices, [3 atom + t]. */
int ox_fock_s(ox_system* s, const double* D, double* Fx);
/* Skeleton derivative of the XC Fock matrix with the potential held fixed
 * (Grid::getFockSkeleton), closed shell, LDA/GGA, at the density D. Fx: 3 natoms
 * column-major nbasis^2 matrices, [3 atom + t]. */
int ox_fock_skeleton(ox_system* s, const double* D, double* Fx);
/* XC energy Hessian in the skeleton form (getDensitySkeleton2, getEnergyHess), the
 * grid held fixed, closed shell, LDA/GGA, at the density D: H[3 A + t][3 B + s],
 * row-major 3 natoms x 3 natoms. */
int ox_xc_hess(ox_system* s, const double* D, double* H);

/* Density quantities on the active sub-grids after ox_density, in the
 * sub-grid order with padding removed: rho, grad (3 per point), sigma, tau. */
long ox_active_points(const ox_system* s);
int ox_get_density(const ox_system* s, double* rho, double* grad, double* sigma, double* tau, double* w);
/* Coordinates of the active sub-grids' unpadded points, in the same order. */
int ox_get_active_xyz(const ox_system* s, double* xyz);
/* radial x angular weights (before the Becke partition) of the active sub-grids' points */
int ox_get_active_raw_weights(const ox_system* s, double* w);

/* Spin-polarised XC build, LDA/GGA and r2SCAN: getDensity({Da, Db}) ->
 * Evaluate("ev") (polarised functional) -> getEnergy -> getFock, Va / Vb per spin. */
int ox_xc_pol(ox_system* s, const double* Da, const double* Db, double* Va, double* Vb, double* Exc, double* Nel);
/* Spin-polarised functional at n points, libxc XC_POLARIZED layout: rho[2n]
 * (a, b), sigma[3n] (aa, ab, bb), vrho[2n], vsigma[3n]; LDA/GGA codes. */
/* Spin-polarised r2SCAN (497, 498) at n points, libxc XC_POLARIZED layout:
 * rho[2n], sigma[3n] (aa, ab, bb), tau[2n] -> exc[n], vrho[2n], vsigma[3n], vtau[2n]. */
int ox_eval_functional_mpol(int ncodes, const int* codes, long n, const double* rho, const double* sigma,
                            const double* tau, double* exc, double* vrho, double* vsigma, double* vtau);
/* Spin-polarised response, LDA/GGA: ground state Da, Db; nmats perturbed matrices per spin
 * (Dxa, Dxb, each nmats x nbasis^2, column-major); response Fock matrices Fa, Fb in the same
 * layout. The derivative of the polarised Fock matrix along D^x. */
int ox_fock_u_pol(ox_system* s, const double* Da, const double* Db, int nmats, const double* Dxa, const double* Dxb,
                  double* Fa, double* Fb);
/* Polarised f_xc at n points: rho[2n], sigma[3n] -> vsigma[3n], v2rho2[3n], v2rhosigma[6n], v2sigma2[6n]. */
int ox_eval_functional_fpol(int ncodes, const int* codes, long n, const double* rho, const double* sigma,
                            double* vsigma, double* v2rho2, double* v2rhosigma, double* v2sigma2);
int ox_eval_functional_pol(int ncodes, const int* codes, long n, const double* rho, const double* sigma, double* exc,
                           double* vrho, double* vsigma);

/* Learned functional surrogate (mlp.cpp): the network in FP64 on the
 * canonical parameter block (8641 doubles, matrices [out][in]); forward value
 * and input gradient, loss L = cs sum C_i (y_i - T_i)^2 (C may be NULL) with
 * its parameter gradient, the descriptors t(rho, sigma, tau) with
 * dt[9i + 3k + c] = d t_k / d(rho, sigma, tau)_c, and the surrogate point
 * evaluation with the analytic fallback. */
int ox_mlp_forward(long n, const double* U, const double* P, double* F, double* dU);
int ox_mlp_loss_grad(long n, const double* U, const double* T, const double* C, double cs, const double* P,
                     double* G, double* loss);
int ox_mlp_descriptors(long n, const double* rho, const double* sigma, const double* tau, double* t, double* dt);
/* The surrogate at n points (nbase analytic base codes added to the network's
 * part; 0: the network is the whole F), with the analytic fallback of codes. */
int ox_mlp_point(const double* P, const double* lo, const double* hi, double rho_min, int ncodes, const int* codes,
                 int nbase, const int* base, long n, const double* rho, const double* sigma, const double* tau,
                 double* exc, double* vrho, double* vsigma, double* vtau, int* used_net);

#ifdef __cplusplus
}
#endif
