/*
 * chinium_xc.h - C ABI of the MI355X (gfx950) exchange-correlation path.
 *
 * Drop-in replacement for the reference's per-SCF-iteration grid/XC calls.
 * Each entry point names the reference interface it replaces (file:line,
 * relative to the Chinium source tree). All functions return CHX_OK (0) or a
 * negative status and never throw; chx_last_error() gives the message, which
 * for the reference's failure cases is the reference's own text.
 *
 * Matrices (D, V, S) are nbasis x nbasis, column-major (Eigen::MatrixXd
 * layout, src/Macro/Abbreviation.h:4), in the reference's AO order
 * (centres in input order -> shells in basis-file order -> components,
 * src/Grid/GridAO.cpp:72-75). Coordinates are in bohr.
 *
 * Threading: a chx_grid owns one HIP stream and its device buffers; calls on
 * one handle are blocking and must not overlap (the reference is not
 * re-entrant on one Grid either). Different handles may be used from
 * different host threads.
 */
#ifndef CHINIUM_XC_H
#define CHINIUM_XC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHX_OK 0
#define CHX_ERR_ARG -1       /* bad argument / wrong size or spin count */
#define CHX_ERR_FILE -2      /* missing or corrupt data file */
#define CHX_ERR_STATE -3     /* call out of protocol order */
#define CHX_ERR_DEVICE -4    /* HIP runtime error / no device */
#define CHX_ERR_UNSUPPORTED -5

typedef struct chx_mol chx_mol;   /* molecule + normalised basis (the Mwfn fields the path reads) */
typedef struct chx_grid chx_grid; /* molecular grid, device resident (class Grid) */
typedef struct chx_xc chx_xc;     /* functional (class ExchangeCorrelation) */

/* Precision of the two dense contractions (Phi.D and Phi^T.B). */
#define CHX_PREC_FP64 0          /* plain form: everything FP64 */
/* Chosen form: Phi.D and Phi^T.B on MFMA as split-BF16 products (hi*hi +
 * hi*lo + lo*hi, 16 significant bits per operand) with FP32 accumulation, AO
 * tables cached in that split form; Becke weights, the functional and the
 * E_xc / N_e / V sums stay FP64. */
#define CHX_PREC_MIXED 1

const char* chx_last_error(void);
int chx_device_count(int* n);

/* ---- Molecule and basis --------------------------------------------------
 * Replaces Mwfn::setCenters / Mwfn::setBasis (legacy/MwfnIO/Mwfn.cpp:606-712)
 * and Normalize (src/Integral/Normalization.cpp:11-18). */
int chx_mol_create(int natoms, const int* Z, const double* xyz_bohr, chx_mol** out);
/* Read <data_root>/BasisSets/<basis>.gbs and normalise (libint2 convention). */
int chx_mol_set_basis_file(chx_mol* m, const char* data_root, const char* basis_name);
/* Append one shell with already-normalised coefficients (MwfnShell fields
 * Type, Exponents, NormalizedCoefficients; legacy/MwfnIO/MwfnShell.h:1-9).
 * type: 0 s, 1 Cartesian p, -1 pure p, -2 pure d, -3 pure f. */
int chx_mol_add_shell(chx_mol* m, int center, int type, int nprim, const double* exps, const double* norm_coefs);
int chx_mol_nbasis(const chx_mol* m, int* nbasis);
int chx_mol_nshells(const chx_mol* m, int* nshells);
/* Copy out shell data: arrays sized nshells (type, nprim, center) and
 * sum(nprim) (exps, coefs). Any pointer may be NULL. */
int chx_mol_get_shells(const chx_mol* m, int* type, int* nprim, int* center, double* exps, double* coefs);
/* Analytic AO overlap S (column-major), used to build S-orthonormal inputs. */
int chx_mol_overlap(const chx_mol* m, double* S);
/* Integrals for the SCF loop around the XC path (small molecules; host code,
 * McMurchie-Davidson over the same AOs as the grid). They replace libint2's
 * Int2C1E (src/Integral/Representation.cpp:62-63) and the integrals behind
 * Int4C2E::ContractInts (src/Integral/Int4C2E.cpp:673-683).
 * chx_mol_int1e: overlap S, kinetic T and nuclear attraction V (n x n each).
 * chx_mol_eri: (mu nu|la si) at mu + n (nu + n (la + n si)), n <= 96.
 * chx_mol_nuclear_repulsion: sum over atom pairs of Z_A Z_B / R_AB. */
int chx_mol_int1e(const chx_mol* m, double* S, double* T, double* V);
int chx_mol_eri(const chx_mol* m, double* eri);
int chx_mol_nuclear_repulsion(const chx_mol* m, double* E);
void chx_mol_destroy(chx_mol* m);

/* ---- Grid ------------------------------------------------------------------
 * Grid::Grid(Mwfn*, std::string grid, int nthreads, int output)
 *   src/Grid/GridGeneration.cpp:232-310 (points, Becke weights, batches).
 * Reads <data_root>/Grids/<grid_name>/<El>.grid. nthreads is recorded for
 * chx_grid_num_threads (Grid::getNumThreads, src/Grid/Grid.h:98) only. */
int chx_grid_create(const chx_mol* m, const char* grid_name, const char* data_root,
                    int nthreads, int output, chx_grid** out);
/* Grid(const Grid&) / operator= (GridGeneration.cpp:312-331): deep copy. */
int chx_grid_copy(const chx_grid* src, chx_grid** out);
void chx_grid_destroy(chx_grid* g);
int chx_grid_num_threads(const chx_grid* g, int* n);
/* Number of retained quadrature points (points with zero weight are dropped:
 * they contribute exactly nothing to any sum) and their coordinates/weights
 * in the library's point order (xyzw[4*i+k]). */
int chx_grid_num_points(const chx_grid* g, long* n);
int chx_grid_get_points(const chx_grid* g, double* xyzw);
/* Grid::setType (src/Grid/Grid.cpp:14-20): -1 none, 0 LDA, 1 GGA, 2 meta-GGA.
 * Meta-GGA (tau) runs on the AO tables kept in HBM (the default when they fit). */
int chx_grid_set_type(chx_grid* g, int type);
/* Grid::getAO (src/Grid/Grid.cpp:22-33). The GPU path recomputes AO values
 * inside the contraction kernels; this call builds the per-batch AO screening
 * that those kernels use and must precede the first density call. */
int chx_grid_get_ao(chx_grid* g, int derivative, int output);
/* Grid::getDensity(std::vector<EigenMatrix> D) (src/Grid/Grid.cpp:35-49):
 * D is nspin column-major nbasis x nbasis matrices, back to back (host memory).
 * nspin = 2 (alpha, beta; Unrestricted/SP.cpp:65-68) is built for LDA, GGA
 * and meta-GGA grids (meta-GGA: r2SCAN); the functional must then be
 * polarised (spin = 2). */
int chx_grid_get_density(chx_grid* g, const double* D, int nspin, int nbasis);
/* Same, with D already in device memory (HBM); used to time the path with
 * inputs resident on the GPU. */
int chx_grid_get_density_device(chx_grid* g, const double* D_dev, int nspin, int nbasis);
/* Grid::getNumElectrons (src/Grid/Grid.cpp:51-61). */
int chx_grid_num_electrons(chx_grid* g, double* N);
/* Grid::getEnergy (src/Grid/Grid.cpp:112-122). Mixed form after
 * Evaluate("ev"): the contractions see each D_s as its two bf16 terms hi + lo,
 * and E adds back the first-order energy of the remainder,
 * sum_s V_s . (D_s - hi - lo) (dE/dD_s = V_s), which makes E a smooth function
 * of D (without it E moves by up to ~2^-17 |V . D| between nearby densities,
 * 5e-5 Eh at (H2O)128). The correction needs the potential: if no Fock call
 * followed the Evaluate, this call builds it into the grid's own buffer, and
 * the next chx_grid_fock / chx_grid_fock_device returns that copy. After
 * Evaluate("e") (no potential) E is returned without the correction. */
int chx_grid_energy(chx_grid* g, double* E);
/* Grid::getFock (src/Grid/Grid.cpp:156-171): writes nspin matrices (host),
 * nspin equal to the evaluated density's. */
int chx_grid_fock(chx_grid* g, double* F, int nspin, int nbasis);
/* Same, into device memory. */
int chx_grid_fock_device(chx_grid* g, double* F_dev, int nspin, int nbasis);
/* XC nuclear gradient in the reference's skeleton form, the grid held fixed:
 * chx_grid_get_ao(g, 1, .) (a no-op on a planned grid, which keeps the last
 * iteration's density and potentials, Restricted/Grad.cpp:73), then
 * chx_grid_get_density_skeleton replaces SubGrid::getDensitySkeleton
 * (GridDensity.cpp:79-130) and chx_grid_energy_grad replaces
 * Grid::getEnergyGrad / SubGrid::getEnergyGrad (GridEnergy.cpp:15-30), using
 * v_rho, v_sigma and grad rho of the last Evaluate("ev"). grad: natoms x 3,
 * row-major [atom][x, y, z]. FP64 form with the AO tables in HBM; LDA and
 * GGA (the reference's getEnergyGrad has no tau term). */
int chx_grid_get_density_skeleton(chx_grid* g, const double* D, int nspin, int nbasis);
int chx_grid_energy_grad(chx_grid* g, double* grad);
/* Response path (coupled-perturbed KS, src/HartreeFockKohnSham/CoupledPerturbed.cpp:133-134,
 * Universal.h:92-118), LDA/GGA, FP64 form, nspin = 1 or 2 (the polarised
 * f_xc from a polarised Evaluate("f") after getDensity({Da, Db})):
 * chx_xc_evaluate(xc, g, "f") after the ground-state density takes f_xc
 * (Evaluate("f"), src/ExchangeCorrelation.cpp:101-104,126-133);
 * chx_grid_get_density_u replaces Grid::getDensityU (src/Grid/Grid.cpp:63-78,
 * SubGrid::getDensityU, src/Grid/GridDensity.cpp:47-77): D holds nspin x nmats
 * column-major nbasis^2 matrices, back to back, [spin][mat] (host memory);
 * chx_grid_fock_u replaces Grid::getFockU<u_t> (src/Grid/Grid.cpp:193-216,
 * SubGrid::getFockU, src/Grid/GridFock.cpp:99-160) and writes the nspin x
 * nmats response matrices in the same order. At nspin = 2 the matrix of spin
 * u for perturbation m is the derivative of the polarised ground-state Fock
 * matrix of spin u along (Dx_a[m], Dx_b[m]): GridFock.cpp:99-160 with the
 * sigma-sigma term scaled by the factor of the differentiated sigma (2 on aa
 * and bb, 1 on ab), where the reference applies the factor of the contracted
 * sigma (FockEinSum/S_x...Eps2Sigma2_g,w,x---ScaledEps2Sigma2_g,w,x.hpp); the
 * two agree for LDA and whenever f_{sigma sigma} has no aa-ab or ab-bb
 * coupling (B88, LYP), and differ for PBE correlation. */
int chx_grid_get_density_u(chx_grid* g, const double* D, int nmats, int nspin, int nbasis);
int chx_grid_fock_u(chx_grid* g, double* F, int nmats, int nspin, int nbasis);
/* Grid::getFockU<s_t> (src/Grid/Grid.cpp:193-216, SubGrid::getFockU with the
 * skeleton derivatives, GridFock.cpp:99-160), as the Hessian calls it
 * (src/HartreeFockKohnSham/Restricted/Hess.cpp:78-97): after
 * chx_xc_evaluate(xc, g, "f") and chx_grid_get_density_skeleton(g, D, ...),
 * the XC response matrix of each nuclear coordinate's skeleton density
 * derivative (rho^{A,t} = -2 sum_{mu in A} D d_t phi_mu phi, and its gradient,
 * GridDensity.cpp:90-128). Writes 3 natoms column-major nbasis^2 matrices,
 * [3 A + t] (the reference's Fs[0][jpert]). Closed shell, LDA/GGA, FP64 form
 * with the AO tables in HBM. */
int chx_grid_fock_s(chx_grid* g, double* F, int nspin, int nbasis);
/* Grid::getFockSkeleton (src/Grid/Grid.cpp:173-191, SubGrid::getFockSkeleton,
 * src/Grid/GridFock.cpp:51-97), as the Hessian calls it
 * (src/HartreeFockKohnSham/Restricted/Hess.cpp:94): the derivative of the XC
 * Fock matrix along each nuclear coordinate with the potential (v_rho,
 * v_sigma, grad rho of the last getDensity / Evaluate("ev")) and the grid
 * held fixed; with chx_grid_fock_s it makes the total derivative. Writes 3
 * natoms column-major nbasis^2 matrices, [3 A + t] (the reference's
 * Fs[0][jpert]). Closed shell, LDA/GGA, FP64 form with the AO tables in HBM. */
int chx_grid_fock_skeleton(chx_grid* g, double* F, int nspin, int nbasis);
/* Grid::getDensitySkeleton2 (src/Grid/Grid.cpp:96-111, GridDensity.cpp:132-209)
 * and Grid::getEnergyHess (Grid.cpp:139-154, GridEnergy.cpp:32-45), as the
 * Hessian calls them (src/HartreeFockKohnSham/Restricted/Hess.cpp:78-90): the
 * XC energy's second derivatives along pairs of nuclear coordinates with D and
 * the grid held fixed. chx_grid_get_density_skeleton2 takes D (as
 * chx_grid_get_density_skeleton); chx_grid_energy_hess, after it,
 * chx_xc_evaluate(xc, g, "ev") at D and chx_xc_evaluate(xc, g, "f"), writes H
 * (3 natoms x 3 natoms, row-major, [3 A + t][3 B + s], the reference's
 * hess[pert]). Closed shell, LDA/GGA, FP64 form with the AO tables in HBM. */
int chx_grid_get_density_skeleton2(chx_grid* g, const double* D, int nspin, int nbasis);
int chx_grid_energy_hess(chx_grid* g, double* H);
/* Supply v_rho directly, in the library's point order, as the SAP guess does
 * when it writes SubGrid::Eps1Rho before Grid::getFock
 * (src/HartreeFockKohnSham/GuessSCF.cpp:37-60). Other potentials are zeroed. */
int chx_grid_set_vrho(chx_grid* g, const double* vrho, long npoints);
/* Device stream the handle launches on (hipStream_t), for event timing. */
int chx_grid_stream(const chx_grid* g, void** stream);
/* Options of the GPU path (no reference counterpart). */
int chx_grid_set_precision(chx_grid* g, int precision);
/* Mixed form: batches with a point within `radius` bohr of a nucleus with
 * Z > 2 (the atomic cores, where the density is largest; the grid batches
 * the points within 0.3 bohr of such a nucleus apart) form rho and grad rho
 * with the AO operand split in three bf16 terms (FP32-level) instead of two
 * (D keeps two: its remainder is taken back in the energy, chx_grid_energy);
 * 0 turns this off, at most 1. Default 0.3; set before getAO. No reference
 * counterpart. */
int chx_grid_set_core_radius(chx_grid* g, double radius);
/* number of such core batches in the current plan (0 for the FP64 form) */
int chx_grid_core_batches(const chx_grid* g, long* n);
/* algorithmic FLOPs of the core batches' launch per build (2 sum g_b n_b^2
 * over those batches, the (rho, grad rho) contraction); the LDA/GGA density
 * launch does the rest of that stage's work */
int chx_grid_core_flops(const chx_grid* g, double* flops);
/* AO screening threshold: an AO tile is dropped from a batch when every |phi|
 * and |grad phi| in it stays at or below the threshold at every point of the
 * batch. Takes effect at the next chx_grid_get_ao. Default 1e-14. */
int chx_grid_set_screening(chx_grid* g, double threshold);
/* Autotuner (no reference counterpart): times `reps` XC builds (D_dev ->
 * V_dev, device pointers, column-major nbasis^2) for each Vxc work-item
 * length (8, 16, 32, 128 batches), keeps the fastest; chunk_out gets it,
 * ms_out[4] the times. Needs the AO tables in HBM (cached or mixed form). */
int chx_grid_autotune(chx_grid* g, const chx_xc* xc, const double* D_dev, double* V_dev, int reps, int* chunk_out,
                      float* ms_out);
/* Mixed form's Vxc block shape (no reference counterpart): 0 = 128 x 128
 * blocks over 128-AO super-tile pairs and the batches where both are alive;
 * 1 = 256 x 256 blocks whose rows and columns are groups of 8 tiles of each
 * batch chunk's alive-tile union, dead tiles skipped per batch; -1 (default)
 * = 1 from npad 1024 on, else 0. chx_grid_vxc_shape reports the shape in use. */
int chx_grid_set_vxc_shape(chx_grid* g, int shape);
/* Vxc work-item length in batches (0: the form's default, at most 128); the
 * autotuners set it from their timings. */
int chx_grid_set_vxc_chunk(chx_grid* g, int chunk);
int chx_grid_vxc_shape(const chx_grid* g, int* shape);
/* chx_grid_autotune over both shapes (mixed form): times the four work
 * lengths for shape 0, then for shape 1, keeps the fastest pair;
 * ms_out[8] = [shape][length]. */
int chx_grid_autotune_tiles(chx_grid* g, const chx_xc* xc, const double* D_dev, double* V_dev, int reps,
                            int* shape_out, int* chunk_out, float* ms_out);
/* Statistics of the batch layout: batches, points, AO tiles per batch (mean),
 * algorithmic FLOPs of one XC build (dense-contraction count, k*2*sum g_b n_b^2). */
int chx_grid_stats(const chx_grid* g, long* nbatches, long* npoints, double* mean_tiles, double* flops_per_build);

/* AO tables in HBM (no reference counterpart; the reference caches its own,
 * src/Grid/Grid.cpp:22-33): -1 (default) caches phi and grad phi of every
 * alive (batch, tile) at the next chx_grid_get_ao when they fit in 80% of the
 * free device memory, 0 never (recompute inside the kernels), 1 always. */
int chx_grid_set_ao_cache(chx_grid* g, int mode);
int chx_grid_ao_cached(const chx_grid* g, int* cached);
/* Per-kernel HIP-event timing on the handle's stream (no reference counterpart).
 * Stages: 0 D permute, 1 density contraction, 2 functional + E/N sums,
 * 3 Vxc contraction, 4 V finalise, 5 Vxc integrand B (cached path),
 * 6 the core batches' three-term density launch (mixed form).
 * Times accumulate until read with reset. */
int chx_grid_profile(chx_grid* g, int enable);
int chx_grid_kernel_times(chx_grid* g, int reset, double* ms /*[7]*/, long* launches /*[7]*/);

/* ---- Functional ------------------------------------------------------------
 * ExchangeCorrelation::Read(std::string df, bool output)
 *   src/ExchangeCorrelation.cpp:18-84: reads <data_root>/DensityFunctionals/<df>.df */
int chx_xc_read(const char* data_root, const char* df_name, int spin, int output, chx_xc** out);
/* Same from explicit libxc IDs (the codes a .df line holds). spin: 1
 * unpolarised, 2 polarised (ExchangeCorrelation::Spin, libxc XC_POLARIZED);
 * of the meta-GGA codes, r2SCAN (497, 498) is built polarised, the others
 * return CHX_ERR_UNSUPPORTED. */
int chx_xc_create(int ncodes, const int* codes, int spin, chx_xc** out);
void chx_xc_destroy(chx_xc* xc);
/* Members Family (0 LDA, 1 GGA, 2 mGGA), EXX, Codes (ExchangeCorrelation.h:9-12). */
int chx_xc_info(const chx_xc* xc, int* family, double* exx, int* ncodes, int* codes);
/* ExchangeCorrelation::Evaluate(std::string order, Grid&)
 *   src/ExchangeCorrelation.cpp:165-248. order: "ev", "e", "v", or "f" (second
 *   derivatives for the response path; LDA/GGA, FP64 form). */
int chx_xc_evaluate(const chx_xc* xc, chx_grid* g, const char* order);

/* Lebedev-Laikov angular rule with npoints points (x, y, z, w; sum w = 1), as
 * ld_by_order (src/Grid/sphere_lebedev_rule.cpp:671-846) returns it; the grid
 * scales w by 4 pi (GridGeneration.cpp:133). Host only. */
int chx_lebedev_rule(int npoints, double* xyzw);

/* Functional values at n points on the device, with libxc's conventions as the
 * reference receives them from xc_{lda,gga,mgga}_exc_vxc
 * (src/ExchangeCorrelation.cpp:98,122,157): exc per particle and the summed
 * derivatives over the functional's codes. Host arrays; sigma/tau may be NULL.
 * A polarised functional (spin = 2) takes libxc's XC_POLARIZED layout: rho[2n]
 * (a, b), sigma[3n] (aa, ab, bb), tau[2n] (a, b; meta-GGA, else unused), and
 * writes vrho[2n], vsigma[3n], vtau[2n] (meta-GGA). */
int chx_xc_eval_points(const chx_xc* xc, long n, const double* rho, const double* sigma, const double* tau,
                       double* exc, double* vrho, double* vsigma, double* vtau);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (serving loops, code under allocation-free policies): the same
 * evaluation as chx_xc_eval_points, same arguments and results, with its
 * device staging (inputs, outputs, code list) in a fixed array the library
 * holds on each device until the process ends (about 5.8 MB) instead of
 * buffers allocated per call. It calls no hipMalloc / hipFree and creates no
 * stream (it runs on the null stream). At most CHX_NOALLOC_MAX_POINTS points
 * per call (CHX_ERR_ARG beyond); spin-polarised meta-GGA functionals are
 * not taken (CHX_ERR_UNSUPPORTED). Calls that overlap on one device share
 * that array and race. */
#define CHX_NOALLOC_MAX_POINTS 65536
int chx_xc_eval_points_noalloc(const chx_xc* xc, long n, const double* rho, const double* sigma, const double* tau,
                               double* exc, double* vrho, double* vsigma, double* vtau);
/* Second derivatives at n points, as xc_{lda,gga}_fxc returns them
 * (src/ExchangeCorrelation.cpp:101-104,126-133): v2rho2, v2rhosigma, v2sigma2
 * summed over the codes. LDA/GGA; host arrays; sigma may be NULL (LDA). A
 * polarised functional (spin 2) takes libxc's XC_POLARIZED layout: rho[2n]
 * (a, b), sigma[3n] (aa, ab, bb) in; v2rho2[3n] (a a, a b, b b),
 * v2rhosigma[6n] (a aa, a ab, a bb, b aa, b ab, b bb), v2sigma2[6n]
 * (aa aa, aa ab, aa bb, ab ab, ab bb, bb bb) out. */
int chx_xc_eval_points_f(const chx_xc* xc, long n, const double* rho, const double* sigma, double* v2rho2,
                         double* v2rhosigma, double* v2sigma2);

/* ---- Coulomb matrix and one-electron integrals for the SCF loop ---------------
 * The SCF update around the XC path (src/HartreeFockKohnSham/Restricted/SP.cpp:
 * 43-56) takes J from Int4C2E::ContractInts (src/Integral/Int4C2E.cpp:673-683)
 * and S, T, V from Int2C1E (src/Integral/Representation.cpp:62-63). This engine
 * computes them on the device: S, T, V exactly, and J by density fitting in an
 * auxiliary basis read from $CHINIUM_PATH/BasisSets/<aux>.gbs (e.g. def2-svp-j)
 * with the Coulomb metric: gamma_P = sum (P|mn) D_mn, (P|Q) c = gamma,
 * J_mn = sum_P (mn|P) c_P. Orbital shells up to d, auxiliary shells up to g.
 * Matrices are column-major nbasis x nbasis (reference AO order), host memory;
 * the metric is naux x naux. The fitting solve itself (a Cholesky factor of the
 * metric, once) is the caller's. */
typedef struct chx_jfit chx_jfit;
int chx_jfit_create(const chx_mol* m, const char* data_root, const char* aux_basis, chx_jfit** out);
/* auxiliary functions, primitive pairs kept and before screening, auxiliary primitives */
int chx_jfit_info(const chx_jfit* j, int* naux, long* npairs, long* npairs_unscreened, int* naux_prims);
int chx_jfit_int1e(chx_jfit* j, double* S, double* T, double* V);
int chx_jfit_metric(chx_jfit* j, double* Vpq);
int chx_jfit_gamma(chx_jfit* j, const double* D, double* gamma);
int chx_jfit_coulomb(chx_jfit* j, const double* cfit, double* J);
void chx_jfit_destroy(chx_jfit* j);

/* ---- Packs: many molecules per XC build (no reference counterpart) -----------
 * Chinium runs one molecule per process (SelfConsistentField.cpp:27-33). A
 * pack takes grids that are each built, set to CHX_PREC_MIXED, typed and
 * planned (chx_grid_get_ao) on their own, and runs one XC build of all of them:
 * every molecule's single-grid kernels, on staggered lanes (see
 * chx_pack_set_lanes); the sequence is captured as a hipGraph and replayed
 * while the D / V pointers and the functional stay the same. The grids must
 * outlive the pack and must not be used on their own while a pack call runs.
 * A member grid that is re-planned (chx_grid_get_ao), re-tuned or whose
 * buffers were reallocated between pack calls is picked up at the next pack
 * call: the lanes are cut again and the graph is recaptured. A member grid
 * switched to another form or type makes the pack call fail (CHX_ERR_STATE).
 * D_dev[m] / V_dev[m]: device arrays, column-major nbasis_m^2, reference
 * order. E, N (host, may be NULL): E_xc (with the correction for the split
 * of D, as chx_grid_energy) and N_e per molecule. */
typedef struct chx_pack chx_pack;
int chx_pack_create(chx_grid* const* grids, int n, chx_pack** out);
int chx_pack_xc(chx_pack* p, const chx_xc* xc, const double* const* D_dev, double* const* V_dev, double* E,
                double* N);
/* 1 (default) replays a captured hipGraph, 0 launches eagerly. */
int chx_pack_set_graph(chx_pack* p, int enable);
/* Autotuner: times `reps` builds for each Vxc work-item length (8, 16, 32,
 * 128 batches), keeps the fastest (chunk_out gets it, ms_out[4] the times);
 * then the Vxc tile shape per molecule (chx_grid_set_vxc_shape) by a
 * basis-size threshold T (256-AO union groups for npad >= T: none, 1024, 512,
 * all), keeps the fastest; then the form and lanes (per-molecule kernels or
 * per-lane stage launches on 4, 8 or 16 lanes, or persistent stages; the last
 * two for LDA/GGA), keeps the fastest. */
int chx_pack_autotune(chx_pack* p, const chx_xc* xc, const double* const* D_dev, double* const* V_dev, int reps,
                      int* chunk_out, float* ms_out);
/* How the build is launched. CHX_PACK_LANES (default): every molecule's
 * single-grid kernels on the lanes. CHX_PACK_PERSISTENT: one launch per stage
 * over every molecule's work items (density, core batches, functional, B^T,
 * Vxc of each block shape), each a fixed grid of resident blocks that walks
 * the flat (molecule, item) list, with each molecule's descriptor read from a
 * table of offsets. CHX_PACK_LANE_STAGES: the same stage launches per lane,
 * over the lane's molecules, the lanes staggered as with CHX_PACK_LANES. The
 * last two are built for LDA/GGA grids (CHX_ERR_UNSUPPORTED on meta-GGA grids;
 * a meta-GGA functional on the pack runs CHX_PACK_LANES). The whole build is
 * one hipGraph in every form; results do not depend on the form. */
#define CHX_PACK_LANES 0
#define CHX_PACK_PERSISTENT 1
#define CHX_PACK_LANE_STAGES 2
int chx_pack_set_form(chx_pack* p, int form);
int chx_pack_form(const chx_pack* p, int* form);
/* The threshold the last autotune chose (-1: no molecule on union groups; -2: not tuned). */
int chx_pack_vxc_wide_from(const chx_pack* p, int* npad_min);
/* Lanes (1..16, default 16, at most one per molecule): the molecules are
 * split into that many contiguous groups of about equal work, each on a stream
 * of its own; a group's density launches wait for the previous group's, so the
 * later stages of one group run beside the density of the next. Results do not
 * depend on it. chx_pack_stream stays the stream the whole build is ordered
 * on. Setting lanes leaves CHX_PACK_PERSISTENT for CHX_PACK_LANES. */
int chx_pack_set_lanes(chx_pack* p, int lanes);
/* Lanes in use (after clamping to the molecule count). */
int chx_pack_lanes(const chx_pack* p, int* lanes);
int chx_pack_stream(const chx_pack* p, void** stream);
void chx_pack_destroy(chx_pack* p);

/* ---- Learned functional surrogate (no reference counterpart) ----------------
 * A 3 -> 64 -> 64 -> 64 -> 1 tanh network on the spin-restricted descriptors
 * t = (ln rho, s/(1+s), alpha/(1+alpha)) that predicts F = e_xc / e_x^unif,
 * e_x^unif = -(3/4)(3/pi)^(1/3) rho^(4/3); v_rho, v_sigma, v_tau are its input
 * gradient chained through t. It stands in for the functional evaluation the
 * reference does with libxc's xc_mgga_exc_vxc (src/ExchangeCorrelation.cpp:
 * 146-160) on meta-GGA grids. Network input u = (t - c)/h, c and h the centre
 * and half width of the training box [lo, hi]; points outside the box or with
 * rho < rho_min take the analytic functional (the out-of-distribution
 * fallback). Parameters: CHX_MLP_PARAMS floats, matrices [out][in]:
 * W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4. Training
 * runs in FP32 with Adam; inference runs the hidden products on MFMA with
 * weights and activations split in two BF16 terms each (FP32-level products)
 * and FP32 accumulation. With an analytic base (chx_mlp_set_base) the network
 * predicts F - F_base, and Evaluate adds the base's own values back. */
#define CHX_MLP_PARAMS 8641
typedef struct chx_mlp chx_mlp;
int chx_mlp_create(unsigned long long seed, chx_mlp** out); /* Glorot-uniform init, seeded */
void chx_mlp_destroy(chx_mlp* m);
int chx_mlp_get_params(const chx_mlp* m, float* P);
int chx_mlp_set_params(chx_mlp* m, const float* P);
int chx_mlp_set_box(chx_mlp* m, const float* lo, const float* hi, double rho_min);
int chx_mlp_get_box(const chx_mlp* m, float* lo, float* hi, double* rho_min);
/* Analytic base: 0 to 2 LDA / GGA codes (e.g. 1, 7: Slater + VWN5); 0 codes:
 * the network predicts the whole F. Training targets must then be
 * F_functional - F_base at the same points. */
int chx_mlp_set_base(chx_mlp* m, int ncodes, const int* codes);
int chx_mlp_get_base(const chx_mlp* m, int* ncodes, int* codes);
/* Training samples from the analytic functional of xc at n points: t (n x 3)
 * and F (n). Host arrays. */
int chx_mlp_features(const chx_xc* xc, long n, const double* rho, const double* sigma, const double* tau, float* t,
                     float* F);
/* Training set (normalised inputs U n x 3, targets T, sample weights C or
 * NULL for uniform), copied to the device. */
int chx_mlp_set_data(chx_mlp* m, long n, const float* U, const float* T, const float* C);
/* Adam over the training set in minibatches of `batch` samples (loss
 * sum C_i (y_i - T_i)^2 / sum C_i per minibatch), learning rate cosine-decayed
 * from lr to lr_final over all steps. epoch_loss (may be NULL): mean minibatch
 * loss of each epoch. */
int chx_mlp_train(chx_mlp* m, int epochs, long batch, double lr, double lr_final, double* epoch_loss);
/* One loss and gradient evaluation at the current parameters (FP32):
 * L = cs sum_i C_i (y_i - T_i)^2, cs = 1 / sum C_i; G gets dL/dP. */
int chx_mlp_loss_grad(chx_mlp* m, long n, const float* U, const float* T, const float* C, float* G, double* loss);
/* Raw network at n normalised inputs: F and dF/dU (n x 3, may be NULL).
 * bf16 = 0: FP32 training weights, per-lane FP32 network; 1: the BF16
 * inference weights on that network; 2: the inference path itself (BF16
 * weights, split-BF16 activations on MFMA). */
int chx_mlp_forward(const chx_mlp* m, long n, const float* U, int bf16, float* F, float* dF);
/* The surrogate at n points (host arrays), with the analytic fallback of xc's
 * codes; *nfallback gets the number of points that took it. */
int chx_mlp_eval_points(const chx_mlp* m, const chx_xc* xc, long n, const double* rho, const double* sigma,
                        const double* tau, double* exc, double* vrho, double* vsigma, double* vtau, long* nfallback);
/* Attach (m != NULL) or detach the surrogate: Evaluate on a meta-GGA grid
 * then runs the network. The network must stay alive while attached and
 * have a box. Packs reject a functional with a surrogate. */
int chx_xc_set_surrogate(chx_xc* xc, const chx_mlp* m);
/* Per-point rho, sigma, tau and weight of the last getDensity, in slot order
 * (npoints_padded = batches x 64 entries; padding slots have weight 0). Any
 * output may be NULL; *n gets the slot count. tau is 0 below meta-GGA type. */
int chx_grid_point_density(chx_grid* g, double* rho, double* sigma, double* tau, double* w, long* n);
/* Points of the last Evaluate that took the analytic fallback (surrogate). */
int chx_grid_fallback_count(const chx_grid* g, long* n);

/* ---- Diagnostics (no reference counterpart) ---------------------------------
 * Runs one FP64 MFMA tile on the device and returns the largest deviation from
 * the scalar product, plus the number of 2^20 sample floats whose hardware
 * bf16 conversion differs from round-to-nearest-even; 0 means the operand
 * layout and the rounding the kernels assume hold. */
int chx_selftest(double* max_err);

#ifdef __cplusplus
}
#endif
#endif /* CHINIUM_XC_H */
