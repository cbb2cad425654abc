// Internal host-side structures of the MI355X XC library (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "chinium_xc.h"

namespace chx {

// Batch and tile geometry shared by host planning and the kernels.
constexpr int kBatch = 64;     // quadrature points per batch (MFMA K for Vxc, M for density)
constexpr int kTile = 32;      // AOs per screening tile
constexpr int kGroup = 4;      // AO tiles per density work group (128 output columns)
constexpr int kSuper = 2;      // AO tiles per Vxc super-tile (64 x 64 output block)
constexpr int kVxcChunk = 32;  // batches per Vxc work item
constexpr int kPadAO = 128;    // nbasis is padded to a multiple of this (4 tiles)
constexpr int kCacheSuper = 4; // tiles per Vxc super-tile in the cached path
constexpr int kCacheChunk = 16;  // batches per cached-path Vxc work item (FP64 form default)
constexpr int kCoef = 5;       // Vxc integrand coefficients per point: a, c_x, c_y, c_z, t
constexpr double kCoreRadius = 0.3;  // bohr: points this close to a nucleus with Z > 2 are batched apart

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define CHX_HIP(x) ::chx::hip_check((x), #x)

struct ShellH {
  int type;    // 0 s, 1 Cartesian p, -1 pure p, -2 pure d, -3 pure f
  int center;
  int nprim;
  int prim_off;
  int ao_off;  // offset in the reference AO order
};

int shell_size(int type);

// A component polynomial of a shell as monomials c x^i y^j z^k, in the
// reference's component order (molecule.cpp).
struct MonoTerm {
  double c;
  int i, j, k;
};
std::vector<std::vector<MonoTerm>> component_monomials(int type);
// A shell of an auxiliary (fitting) basis: angular momentum l (pure), centre,
// exponents, normalised contraction coefficients.
struct AuxShell {
  int l, center;
  std::vector<double> e, c;
};

// A device buffer owned by a handle.
template <class T> struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    CHX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  T* alloc_get(size_t count) {
    alloc(count);
    return p;
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    alloc(h.size());
    if (!h.empty()) CHX_HIP(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void copy_from(const DevBuf& o, hipStream_t s) {
    alloc(o.n);
    if (o.n) CHX_HIP(hipMemcpyAsync(p, o.p, o.n * sizeof(T), hipMemcpyDeviceToDevice, s));
  }
};

// One shell segment of a tile with everything its evaluation reads, so that a
// wave fetches it with two scalar loads (mixed-precision form).
constexpr int kSegPrim = 12;  // primitives per record
struct SegRec {
  double ax, ay, az;  // shell centre
  int meta;           // type & 0xff | nprim << 8 | first component << 16 | count << 24
  int col;            // AO offset in the tile
  float e[kSegPrim], c[kSegPrim];
};
static_assert(sizeof(SegRec) == 128, "segment record is two 64-byte scalar loads");

// Kernel-side view of the basis in the internal AO order.
struct BasisView {
  const double* sh_xyz;   // 3 per shell
  const int* sh_type;
  const int* sh_nprim;
  const int* sh_poff;
  const double* p_exp;
  const double* p_coef;
  const int* tile_seg_off;  // ntiles + 1
  const int4* seg;          // (shell, first component, count, local AO offset in tile)
  const SegRec* segrec;     // per segment, same index as seg (mixed-precision form)
  const int* sh_atom;       // atom of each shell (XC gradient)
};

// Launchers (kernels.hip).
void launch_becke(hipStream_t s, long npts, const double* px, const double* py, const double* pz, double* w,
                  const int* owner, int natoms, const double* atom_xyz, const double* inv_rab);
void launch_screen(hipStream_t s, int nbatches, int ntiles, const double* bx, const double* by, const double* bz,
                   const int* bn, BasisView bv, double thr, unsigned char* alive);
void launch_permute_d(hipStream_t s, int n, int npad, const double* D, const int* int2ref, double* Dint);
// mode 0: rho, grad rho / 2 into part ([group][4][64]); mode r = 1..3: the tau
// partial of direction r (part = component r - 1 of [group][3][64]).
void launch_density_f64(hipStream_t s, int mode, int ngroups, const int* grp_batch, const int* grp_tiles,
                        const int* grp_ntiles, const int* bt_off, const int* bt_list, const double* bx,
                        const double* by, const double* bz, BasisView bv, const double* Dint, int npad, double* part);
// One molecule of a multi-molecule functional launch: byte offsets of its
// buffers from the launch's base pointer, its slot count, its first block
// (entry nmol is a sentinel whose blk0 is the total block count) and its E/N slot.
struct FnMolecule {
  long grp_off, part, part_tau, w, rho, coef, en_part, cpart;
  int nslots, blk0, slot_en, ncpart;
};
void launch_functional_multi(hipStream_t s, const char* base, const FnMolecule* fm, int nmol, int nblocks,
                             const int* codes, int ncodes, int family, double* en);
void launch_functional(hipStream_t s, int nbatches, const int* grp_off, const double* part, const double* part_tau,
                       const double* bw, const int* codes, int ncodes, int family, int want_e, int want_v,
                       double* rho_out, double* coef, double* en_part, int en_blocks);
void launch_eval_functional(hipStream_t s, long n, const int* codes, int ncodes, const double* rho, const double* sigma,
                            const double* tau, double* exc, double* vrho, double* vsigma, double* vtau);
void launch_reduce_en(hipStream_t s, int nparts, const double* en_part, double* en);
void launch_fill(hipStream_t s, double* p, size_t n, double v);
// ncomp 1: a phi + c . grad phi; ncomp 4 adds the tau term (t = coef[4]).
void launch_vxc_f64(hipStream_t s, int ncomp, int nwork, const int4* work, const int* pair_batches,
                    const unsigned char* alive, int ntiles, const double* bx, const double* by, const double* bz,
                    const double* coef, BasisView bv, double* Vint, int npad);
void launch_finalize_v(hipStream_t s, int n, int npad, const double* Vint, const int* ref2int, double* V);
void launch_vrho_coef(hipStream_t s, long nslots, const double* vrho_w, double* coef);
// Cached-AO FP64 path (kernels_cached.hip).
void launch_ao_cache(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* tile_nao,
                     const double* bx, const double* by, const double* bz, BasisView bv, double* cache);
// response path (closed shell, LDA/GGA)
constexpr int kFx = 8;  // per slot: v2rho2, v2rhosigma, v2sigma2, v_sigma, grad rho (3), -
// polarised (nspin = 2): vsigma (3), v2rho2 (3), v2rhosigma (6), v2sigma2 (6), grad rho_a (3), grad rho_b (3)
constexpr int kFxPol = 24;
void launch_functional_fpol(hipStream_t s, int nbatches, const int* grp_off, const double* part, const double* part_b,
                            const int* codes, int ncodes, int family, double* fx);
void launch_response_coef_pol(hipStream_t s, int nbatches, const double* rhou_a, const double* rhou_b,
                              const double* fx, const double* bw, int family, int spin, double* coef);
void launch_eval_functional_fpol(hipStream_t s, long n, const int* codes, int ncodes, const double* rho,
                                 const double* sigma, double* v2rho2, double* v2rhosigma, double* v2sigma2);
void launch_functional_f(hipStream_t s, int nbatches, const int* grp_off, const double* part, const int* codes,
                         int ncodes, int family, double* fx);
void launch_point_rho(hipStream_t s, int nbatches, const int* grp_off, const double* part, double* rhou);
void launch_response_coef(hipStream_t s, int nbatches, const double* rhou, const double* fx, const double* bw,
                          int family, double* coef);
void launch_eval_functional_f(hipStream_t s, long n, const int* codes, int ncodes, const double* rho,
                              const double* sigma, double* v2rho2, double* v2rhosigma, double* v2sigma2);
// spin-polarised (nspin = 2) functional, LDA/GGA and r2SCAN (part_tau / part_tau_b: meta-GGA)
void launch_functional_pol(hipStream_t s, int nbatches, const int* grp_off, const double* part, const double* part_b,
                           const double* part_tau, const double* part_tau_b, const double* bw, const int* codes,
                           int ncodes, int family, int want_e, int want_v, double* rho_out, double* coef,
                           double* coef_b, double* en_part, int en_blocks);
// at points, libxc's polarised layout; tau / vtau (meta-GGA) may be null
void launch_eval_functional_pol(hipStream_t s, long n, const int* codes, int ncodes, const double* rho,
                                const double* sigma, const double* tau, double* exc, double* vrho, double* vsigma,
                                double* vtau);
// Surrogate network (kernels_mlp.hip): 3 -> 64 -> 64 -> 64 -> 1, tanh. Offsets
// of the parameter block (floats); matrices are [out][in] (y = W x + b). The
// run copies the kernels read hold W2 transposed ([in][out]) at the same offset.
constexpr int kMlpH = 64;
constexpr int kMlpW1 = 0, kMlpB1 = 192, kMlpW2 = 256, kMlpB2 = 4352, kMlpW3 = 4416, kMlpB3 = 8512, kMlpW4 = 8576,
              kMlpB4 = 8640, kMlpParams = 8641;
// Descriptor box of the training data: t = (ln rho, s/(1+s), alpha/(1+alpha));
// the network input is u = (t - c) / h with c, h the box centre and half width.
struct MlpBox {
  float lo[3], hi[3];
  double rho_min;  // below: analytic functional (low-density tail)
  int nbase;       // the network predicts F - F_base of these analytic codes (0: the whole F)
  int base[2];
};
void launch_mlp_pack(hipStream_t s, const float* P, float* Prun, float* Pinf);
void launch_mlp_forward(hipStream_t s, long n, const float* U, const float* R, const unsigned* Wpk, float* F,
                        float* dF);
constexpr int kMlpPacked = 4 * 4 * 2 * 2 * 64 * 4;  // u32 words of the packed split-BF16 (hi, lo) MFMA weight fragments
void launch_mlp_pack_mx(hipStream_t s, const float* P, unsigned* Wpk);
void launch_mlp_grad(hipStream_t s, long n, const float* U, const float* T, const float* C, float cs, const float* R,
                     float* G, float* loss, int max_blocks);
void launch_mlp_adam(hipStream_t s, float* P, float* G, float* M, float* V, float* Prun, float* Pinf, float lr,
                     int step);
void launch_mlp_features(hipStream_t s, long n, const double* rho, const double* sigma, const double* tau,
                         const int* codes, int ncodes, float* t, float* F);
void launch_mlp_eval(hipStream_t s, long n, const float* R, const unsigned* Wpk, const MlpBox& box, const int* codes,
                     int ncodes,
                     const double* rho, const double* sigma, const double* tau, double* exc, double* vrho,
                     double* vsigma, double* vtau, int* nfall);
void launch_functional_mlp(hipStream_t s, int nbatches, const int* grp_off, const double* part,
                           const double* part_tau, const double* bw, const float* R, const unsigned* Wpk,
                           const MlpBox& box, const int* codes, int ncodes, int want_e, int want_v, double* rho_out,
                           double* coef, double* en_part, int* nfall);
void launch_point_density(hipStream_t s, int nbatches, const int* grp_off, const double* part,
                          const double* part_tau, int has_tau, const double* bw, double* rho, double* sigma,
                          double* tau, double* w);
void launch_density_c64(hipStream_t s, int mode, int ngroups, const int* grp_batch, const int* grp_tiles,
                        const int* grp_slots, const int* grp_ntiles, const int* bt_off, const int* bt_list,
                        const double* cache, const double* Dint, int npad, double* part);
// B = a phi + c . grad phi per slot (one component for every family)
void launch_make_b(hipStream_t s, int nbatches, const int* bt_off, const double* coef, const double* cache,
                   double* bmat);
// ncomp 4 adds the tau term, its operand t d_r phi formed from the cache and coef[4] as it is loaded
void launch_vxc_c64(hipStream_t s, int ncomp, int nwork, const int4* work, const int* pair_batches, const int* slot_of,
                    int ntiles, const double* cache, const double* bmat, const double* coef, double* Mint, int npad);
void launch_finalize_sym(hipStream_t s, int n, int npad, const double* Mint, const int* ref2int, double* V);
// Device descriptor of one grid for the mixed-precision kernels. A kernel
// launch covers a list of (grid, item) pairs, so one launch can serve many
// molecules (a pack) as well as a single grid.
struct MolDev {
  const int *grp_batch, *grp_tiles, *grp_ntiles, *bt_off, *bt_list;  // density work groups
  const double *bx, *by, *bz;                                          // points
  BasisView bv;
  const unsigned short* cache;  // split operand cache
  int sstride, toff;            // slot stride; offset of the transposed layout
  const unsigned short* Ds;     // split D
  int npad, ntiles, nslots, pad_;
  double *part, *part_tau;      // density partials
  const int *slot_batch, *slot_tile;
  const double* coef;           // Vxc integrand coefficients per point
  unsigned short* bt;           // B^T per slot
  const int4* cwork;            // Vxc work items
  const int *pair_batches, *slot_of;
  const int* cslot;             // per pair_batches entry: slots of I's 4 tiles, then J's (zero slot if absent)
  double* Mint;                 // M accumulator, npad^2
  // Vxc over chunk tile groups (large bases): per chunk of consecutive batches
  // (table offset, batches, table width = padded alive-tile union, tile-list
  // offset); per (chunk batch, union position) the slot (zero slot if the tile
  // is not alive in that batch); per union position the tile (-1: padding).
  const int4* vchunk;
  const int* vtab;
  const int* vtile;
  // core batches (points within the core radius of a nucleus with Z > 2): their
  // density runs with the AO operands split in three bf16 terms; cache3 holds
  // the third term of their slots' AOg blocks [core slot][c][g][mu],
  // core_off the first core slot of each batch (-1: not a core batch),
  // core_grps the density groups of the core batches, main_grps those of the
  // other batches (the (rho, grad rho) launch's groups; null: every group)
  const unsigned short* cache3;
  const int *core_off, *core_grps, *main_grps;
};
constexpr int kMxMaxTiles = 512;  // alive tiles per batch the mixed density kernel stages (npad <= 16384)
constexpr int kMxMaxChunk = 128;  // batches per Vxc work item in the mixed form

// XC nuclear gradient (kernels_grad.hip), FP64 cached form.
void launch_xgrad_xz(hipStream_t s, int ngroups, const int* grp_batch, const int* grp_tiles, const int* grp_ntiles,
                     const int* bt_off, const int* bt_list, const double* cache, const double* bmat,
                     const double* Dint, int npad, int c1, int c2, double* xz);
void launch_xgrad_epi(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* slot_grp,
                      const int* slot_q, const double* bx, const double* by, const double* bz, BasisView bv,
                      const double* coef, const double* xz, double* grad);
// skeleton density derivatives (rho, grad rho per nuclear coordinate) into
// rhos[3 atom + t][npt points][4], accumulated (clear it first)
// Skeleton Fock derivative: E^t = a d_t phi + (c . grad) d_t phi per slot into
// etab[t][slot][g][column] (zero it first), then F[3 A + t] in the reference's
// order from Q^t = (d_t Phi)^T B (qp[t]) and P^t = Phi^T E^t (qp[3 + t]).
void launch_skel_etab(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const double* bx,
                      const double* by, const double* bz, BasisView bv, const double* coef, double* etab);
void launch_fock_skel_fin(hipStream_t s, int n, int npad, int natoms, const double* qp, const int* ref2int,
                          const int* ao_atom, double* F);
// XC energy Hessian terms added to H (np x np, np = 3 natoms; clear it first):
// the per-slot diagonal-atom term (AO third derivatives, X and Y^r from xz1 /
// xz2), the f_xc term of the first-order skeleton densities rhos, and the D
// term from rr[3 t + s] = (d_t Phi)^T E^s (npad^2 each, internal order).
void launch_hess_terms(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* slot_grp,
                       const int* slot_q, const double* bx, const double* by, const double* bz, BasisView bv,
                       const double* coef, const double* xz1, const double* xz2, long npt, const double* rhos,
                       const double* fx, const double* bw, int family, int npad, const double* Dint,
                       const double* rr, const int* atom, int np, double* H);
void launch_skel_rho(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* slot_grp,
                     const int* slot_q, const double* bx, const double* by, const double* bz, BasisView bv,
                     const double* xz1, const double* xz2, long npt, double* rhos);

// Mixed-precision form (kernels_mixed.hip): MFMA operands cached as bf16
// hi/lo pairs, split-BF16 MFMA contractions with FP32 accumulation, FP32
// recomputation of the AO values outside the MFMA operands.
constexpr int kMxComp = 2 * kTile * kBatch;  // elements per component and layout of a slot
// slot_core / cache3 (may be null): the third bf16 term of the AOg block of each
// slot with slot_core >= 0, at cache3 + slot_core * nT * kBatch * kTile
void launch_ao_cache_mx(hipStream_t s, int nslots, const int* slot_batch, const int* slot_tile, const int* tile_nao,
                        const double* bx, const double* by, const double* bz, BasisView bv, int nT, int sstride,
                        unsigned short* cache, const int* slot_core, unsigned short* cache3);
// R (may be null): what the two terms leave of D, D - hi - lo, in FP32 and in
// D's own (reference, column-major) order
void launch_split_d(hipStream_t s, int n, int npad, const double* D, const int* int2ref, unsigned short* Ds,
                    float* R);
// First-order energy correction for the split of D: V = M + M^T as
// launch_finalize_sym, and per block the partial sum of V . R into cpart
// (corr_blocks(n) blocks); launch_reduce_corr sums nparts partials into *out.
int corr_blocks(int n);
void launch_finalize_corr(hipStream_t s, int n, int npad, const double* Mint, const int* ref2int, double* V,
                          const float* R, double* cpart);
void launch_reduce_corr(hipStream_t s, int nparts, const double* cpart, double* out);
// the packed form: per molecule of the table, en[2 slot_en] += its ncpart partials
void launch_corr_multi(hipStream_t s, const char* base, const FnMolecule* fm, int nmol, double* en);
void launch_bf16_check(hipStream_t s, unsigned n, unsigned* bad);
void launch_fill_u32(hipStream_t s, unsigned* p, size_t n, unsigned v);
// One launch over the grid `one` (descriptor by value): density work groups,
// B^T slot blocks, Vxc work items. Density mode 0: rho, grad rho; 4: the three
// tau partials in one launch.
void launch_density_mx(hipStream_t s, int mode, int nitems, const MolDev& one);
// the core groups again with three-term operands: mode 0 (rho, grad rho) or r = 1..3 (tau partial r)
void launch_density_mx_core(hipStream_t s, int mode, int ncore_groups, const MolDev& one);
// B^T of every slot plus the all-zero slot (index nslots) for ncomp components
void launch_make_b_mx(hipStream_t s, int ncomp, const MolDev& one);
void launch_vxc_mx(hipStream_t s, int ncomp, int nitems, const MolDev& one);
// Persistent stage launches over many molecules (packs, LDA/GGA): tab holds one
// descriptor per molecule with byte offsets from base in place of pointers;
// starts (nmol + 1) the first item of each molecule in the stage's flat list.
enum { kStageDensity = 0, kStageCore = 1, kStageMakeB = 2, kStageVxc = 3, kStageVxcWide = 4 };
void launch_stage_multi(hipStream_t s, int stage, int ncomp, const char* base, const MolDev* tab, const int* starts,
                        int nmol, int total);
// 256 x 256 output blocks over groups of 8 tiles of each batch chunk's alive-tile
// union (vchunk / vtab / vtile); items (group I, group J, chunk)
void launch_vxc_mx_wide(hipStream_t s, int ncomp, int nitems, const MolDev& one);

}  // namespace chx

// Opaque handle definitions.
namespace chx {
std::vector<AuxShell> aux_basis_shells(const chx_mol* m, const std::string& data_root, const std::string& basis);
}

struct chx_mol {
  int natoms = 0;
  std::vector<int> Z;
  std::vector<double> xyz;
  std::vector<chx::ShellH> shells;
  std::vector<double> exps, coefs;
  int nbasis = 0;
};

struct chx_mlp {
  hipStream_t stream = nullptr;
  chx::DevBuf<float> P, Prun, Pinf, G, M, V, loss;  // canonical, run (FP32), run (BF16-rounded); Adam state
  chx::DevBuf<unsigned> Wpk;                         // packed BF16 MFMA fragments of the inference weights
  chx::DevBuf<float> U, T, C;                        // training set: inputs, targets, sample weights
  std::vector<double> cprefix;                       // prefix sums of C (minibatch normalisation)
  long ndata = 0;
  int step = 0;
  bool have_box = false;
  chx::MlpBox box{};
  ~chx_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct chx_xc {
  std::vector<int> codes;
  int family = 0;
  double exx = 0;
  int spin = 1;
  const chx_mlp* surrogate = nullptr;  // when set, Evaluate runs the network (meta-GGA grids)
};

struct chx_grid {
  int device = 0;
  hipStream_t stream = nullptr;
  int nthreads = 1;
  int type = 0;
  int precision = CHX_PREC_FP64;
  double screen_thr = 1e-14;
  chx_mol mol;

  // AO layout: internal order groups AOs of spatially sorted atoms.
  int nbasis = 0, npad = 0, ntiles = 0;
  std::vector<int> int2ref;  // npad, -1 for padding
  std::vector<int> ref2int;  // nbasis
  std::vector<double> sh_xyz;
  std::vector<int> sh_type, sh_nprim, sh_poff, sh_atom;
  std::vector<double> p_exp, p_coef;
  std::vector<int> tile_seg_off;
  std::vector<int4> seg;

  // Points: nbatches x kBatch slots (padding slots carry w = 0).
  long npoints = 0;
  int nbatches = 0;
  std::vector<int> batch_n;        // real points per batch
  std::vector<double> hx, hy, hz, hw;  // slot arrays (host copy)

  // Screening products (built by get_ao).
  bool ao_ready = false;
  // Bumped whenever the plan, the work lists or a buffer a captured pack graph
  // addresses may change (grid_plan, set_vxc_chunk, reallocations).
  unsigned long plan_gen = 0;
  std::vector<unsigned char> alive;  // nbatches x ntiles
  std::vector<int> bt_off, bt_list;  // alive tiles per batch (CSR)
  std::vector<int> grp_batch, grp_tiles, grp_ntiles, grp_off;  // density work groups
  std::vector<int4> vwork;           // Vxc work items (tile I, tile J, begin, end)
  std::vector<int> pair_batches;
  // Cached-AO path: phi and grad phi per alive (batch, tile) slot kept in HBM.
  bool cached = false;
  int cache_mode = -1;               // -1 automatic, 0 never, 1 always
  std::vector<int> slot_batch, slot_tile, slot_of, tile_nao, grp_slots;
  int vxc_chunk = 0;                 // batches per Vxc work item; 0 = the form's default
  int vxc_shape = -1;                // mixed form's Vxc blocks: -1 by basis size, 0 128-AO super-tiles, 1 256-AO union groups
  int mx_wide = 0;                   // mixed form: Vxc over 256-AO super-tiles (large bases)
  std::vector<int4> cwork;
  std::vector<int> cpair_batches;
  std::vector<int> cslot;              // mixed form: slot table per cpair_batches entry
  chx::DevBuf<int> d_cslot;
  // mixed form, large bases: Vxc items over groups of each batch chunk's alive-tile union
  std::vector<int4> vchunk;
  std::vector<int> vtab, vtile;
  chx::DevBuf<int4> d_vchunk;
  chx::DevBuf<int> d_vtab, d_vtile;
  double vxc_exec_frac = 1.0;          // useful / executed tile-pair products of the mixed Vxc plan
  // mixed form: core batches (density with three-term operands)
  double core_radius = chx::kCoreRadius;  // bohr; 0 turns the core treatment off
  int ncore_batches = 0;
  std::vector<int> core_grps, main_grps;
  chx::DevBuf<unsigned short> d_cmx3;
  chx::DevBuf<int> d_core_off, d_core_grps, d_main_grps, d_slot_core;
  chx::DevBuf<double> d_cache, d_bmat, d_Mint;
  chx::DevBuf<int> d_slot_batch, d_slot_tile, d_slot_of, d_tile_nao, d_grp_slots, d_cpair_batches;
  chx::DevBuf<int4> d_cwork;
  double flops_build = 0, mean_tiles = 0;
  std::vector<double> batch_work;  // g_b n_b^2 per batch (points x kept AOs squared)
  double core_flops = 0;           // the core batches' three-term (rho, grad rho) launch: 2 sum g_b n_b^2

  // Device mirrors.
  chx::DevBuf<double> d_sh_xyz, d_p_exp, d_p_coef;
  chx::DevBuf<chx::SegRec> d_segrec;
  int max_nprim = 0;
  // Mixed form: split AO cache (AOg + nT transposed components per slot), B^T, split D.
  chx::DevBuf<unsigned short> d_cmx, d_bmx, d_Dsplit;
  chx::MolDev hmol{};                 // descriptor of this grid (mixed form), passed to kernels by value
  int mx_sstride = 0, mx_nT = 0, bmx_comps = 0;
  chx::DevBuf<int> d_sh_type, d_sh_nprim, d_sh_poff, d_tile_seg_off, d_sh_atom;
  chx::DevBuf<int4> d_seg;
  chx::DevBuf<int> d_int2ref, d_ref2int;
  chx::DevBuf<double> d_x, d_y, d_z, d_w;
  chx::DevBuf<int> d_bn;
  chx::DevBuf<unsigned char> d_alive;
  chx::DevBuf<int> d_bt_off, d_bt_list, d_grp_batch, d_grp_tiles, d_grp_ntiles, d_grp_off;
  chx::DevBuf<int4> d_vwork;
  chx::DevBuf<int> d_pair_batches;
  chx::DevBuf<int> d_codes;

  // Per-iteration state.
  // d_en: E, N and (mixed form) the first-order correction for the split of D
  chx::DevBuf<double> d_Din, d_Dint, d_part, d_part_tau, d_rho, d_coef, d_en_part, d_en, d_Vint, d_Vout;
  // mixed form: D - hi - lo per spin (FP32, reference order) and the partial sums of V . (D - hi - lo)
  chx::DevBuf<float> d_Dres;
  chx::DevBuf<double> d_cpart;
  // corr_pending: the last Evaluate computed E and v in the mixed form, so the potential's build
  // adds the correction to E; corr_done: a build since then has; vout_ready: d_Vout holds that
  // potential (vout_nspin matrices)
  bool corr_pending = false, corr_done = false, vout_ready = false;
  int vout_nspin = 0;
  int v_family = 0;    // family of the potential on the grid (set by Evaluate / set_vrho)
  int en_blocks = 0;
  // XC gradient (Grid::getDensitySkeleton / getEnergyGrad)
  chx::DevBuf<double> d_Dskel;
  std::vector<int> slot_grp, slot_q;
  chx::DevBuf<int> d_slot_grp, d_slot_q;
  bool have_skeleton = false;
  bool have_density = false, have_eval = false, have_v = false;
  // response path: f_xc at the ground-state density, perturbed densities
  chx::DevBuf<double> d_fx, d_part_u, d_rhou, d_coef_u;
  int fx_family = -1;  // family of the f_xc on the grid (-1: none)
  int fx_nspin = 1;    // spin count of that f_xc (2: kFxPol values per slot)
  int nmats_u = 0;     // perturbed densities held in d_rhou (per spin)
  int nspin_u = 1;     // spin count of those densities ([spin][mat] in d_rhou)
  // spin-polarised densities (nspin = 2): beta partials and coefficients; alpha in d_part / d_coef
  int nspin = 1;    // spin count of the density on the grid
  int v_nspin = 1;  // spin count of the potential on the grid
  chx::DevBuf<double> d_part_b, d_part_tau_b, d_coef_b, d_coef_save;
  // surrogate functional: points that fell back to the analytic functional in the last Evaluate
  chx::DevBuf<int> d_nfall;
  bool surrogate_used = false;

  // Optional per-kernel HIP-event timing (stage ids in chx_grid_kernel_times).
  bool profile = false;
  struct Pending { int stage; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double stage_ms[8] = {0};
  long stage_n[8] = {0};
  void mark_begin(int stage, hipEvent_t* a);
  void mark_end(int stage, hipEvent_t a);
  void collect_times();
  int eval_family = -1;
  double last_E = 0, last_N = 0;

  chx::BasisView view() const;
};
