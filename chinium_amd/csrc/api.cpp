// C ABI of the XC path (include/chinium_xc.h). Every entry point catches
// exceptions and returns a status; the message is kept per host thread.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "internal.hpp"
#include "xc_functionals.hpp"

void mol_set_basis_file(chx_mol* m, const std::string& data_root, const std::string& basis);
void mol_overlap(const chx_mol* m, double* S);

namespace chx {
void grid_build(chx_grid* g, const std::string& grid_name, const std::string& data_root);
void grid_plan(chx_grid* g);
void set_vxc_chunk(chx_grid* g, int chunk);
void set_vxc_shape(chx_grid* g, int shape);
void ensure_stream(chx_grid* g);
void launch_selftest(hipStream_t s, double* out);
}  // namespace chx
int chx_lebedev_rule_impl(int n, double* xyzw);
namespace chx {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(CHX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace chx

using namespace chx;

void chx_grid::mark_begin(int stage, hipEvent_t* a) {
  (void)stage;
  *a = nullptr;
  if (!profile) return;
  CHX_HIP(hipEventCreate(a));
  CHX_HIP(hipEventRecord(*a, stream));
}

void chx_grid::mark_end(int stage, hipEvent_t a) {
  if (!profile || !a) return;
  hipEvent_t b;
  CHX_HIP(hipEventCreate(&b));
  CHX_HIP(hipEventRecord(b, stream));
  pending.push_back({stage, a, b});
}

void chx_grid::collect_times() {
  if (pending.empty()) return;
  CHX_HIP(hipStreamSynchronize(stream));
  for (const Pending& p : pending) {
    float ms = 0;
    CHX_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    stage_ms[p.stage] += ms;
    stage_n[p.stage] += 1;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  pending.clear();
}

namespace {
thread_local std::string g_last_error;
}  // namespace
namespace chx {
std::string& last_error() { return g_last_error; }
}  // namespace chx
namespace {

template <class F> int guard(F&& f) {
  try {
    f();
    return CHX_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CHX_ERR_ARG;
  }
}

void require(bool ok, int code, const char* msg) {
  if (!ok) throw Error(code, msg);
}

void make_stream(chx_grid* g) { ensure_stream(g); }

void run_functional(chx_grid* g, const chx_xc* xc, bool want_e, bool want_v) {
  if (xc->spin != g->nspin)
    throw Error(CHX_ERR_STATE, "the functional's spin count differs from the density's (ExchangeCorrelation::Spin)");
  g->d_codes.upload(xc->codes, g->stream);
  hipEvent_t e = nullptr;
  g->surrogate_used = false;
  g->vout_ready = g->corr_done = false;
  g->corr_pending = g->precision == CHX_PREC_MIXED && want_e && want_v;
  if (g->nspin == 2) {  // spin-polarised: alpha coefficients in d_coef, beta in d_coef_b
    if (xc->family > g->type)
      throw Error(CHX_ERR_STATE, "the grid's type holds no tau for this meta-GGA functional (Grid::setType)");
    g->d_coef_b.alloc(g->d_coef.n);
    g->mark_begin(2, &e);
    launch_functional_pol(g->stream, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_part_b.p, g->d_part_tau.p,
                          xc->family >= 2 ? g->d_part_tau_b.p : nullptr, g->d_w.p, g->d_codes.p, (int)xc->codes.size(),
                          xc->family, want_e ? 1 : 0, want_v ? 1 : 0, g->d_rho.p, g->d_coef.p, g->d_coef_b.p,
                          g->d_en_part.p, g->en_blocks);
    launch_reduce_en(g->stream, g->en_blocks, g->d_en_part.p, g->d_en.p);
    g->mark_end(2, e);
    return;
  }
  if (xc->surrogate && xc->family == 2) {  // learned functional, one partial sum per 64-slot block
    const chx_mlp* m = xc->surrogate;
    if (g->d_en_part.n < 2 * (size_t)g->nbatches) {
      g->d_en_part.alloc(2 * (size_t)g->nbatches);
      ++g->plan_gen;
    }
    g->d_nfall.alloc(1);
    launch_fill_u32(g->stream, (unsigned*)g->d_nfall.p, 1, 0u);
    g->mark_begin(2, &e);
    launch_functional_mlp(g->stream, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_part_tau.p, g->d_w.p, m->Prun.p,
                          m->Wpk.p, m->box, g->d_codes.p, (int)xc->codes.size(), want_e ? 1 : 0, want_v ? 1 : 0, g->d_rho.p,
                          g->d_coef.p, g->d_en_part.p, g->d_nfall.p);
    launch_reduce_en(g->stream, g->nbatches, g->d_en_part.p, g->d_en.p);
    g->mark_end(2, e);
    g->surrogate_used = true;
    return;
  }
  g->mark_begin(2, &e);
  launch_functional(g->stream, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_part_tau.p, g->d_w.p, g->d_codes.p,
                    (int)xc->codes.size(), xc->family, want_e ? 1 : 0, want_v ? 1 : 0, g->d_rho.p, g->d_coef.p,
                    g->d_en_part.p, g->en_blocks);
  launch_reduce_en(g->stream, g->en_blocks, g->d_en_part.p, g->d_en.p);
  g->mark_end(2, e);
}

// Stage ids of chx_grid_kernel_times.
enum { kStPermute = 0, kStDensity = 1, kStFunctional = 2, kStVxc = 3, kStFinalize = 4, kStMakeB = 5, kStCore = 6 };

// spin: which of the (up to two) split remainders of D this density's is
void density_from_device(chx_grid* g, const double* dD, int spin = 0) {
  hipEvent_t e = nullptr;
  const int ng = (int)g->grp_batch.size();
  g->corr_pending = g->corr_done = g->vout_ready = false;
  if (g->precision == CHX_PREC_MIXED) {
    g->mark_begin(kStPermute, &e);
    const int ncg = (int)g->core_grps.size();
    launch_split_d(g->stream, g->nbasis, g->npad, dD, g->d_int2ref.p, g->d_Dsplit.p,
                   g->d_Dres.p + (size_t)spin * g->nbasis * g->nbasis);
    g->mark_end(kStPermute, e);
    g->mark_begin(kStDensity, &e);
    // meta-GGA: rho, grad rho and the tau partials (A = d_r phi) in one launch that stages D once
    // (rho, grad rho alone: the core batches' groups are left to the core launch)
    launch_density_mx(g->stream, g->type >= 2 ? 5 : 0, g->type >= 2 ? ng : ng - ncg, g->hmol);
    g->mark_end(kStDensity, e);
    // core batches' rho and grad rho with three-term AO operands (meta-GGA: overwriting their
    // partials; tau keeps two terms: near the nuclei its error is far below rho's effect on E_xc)
    if (ncg) {
      g->mark_begin(kStCore, &e);
      launch_density_mx_core(g->stream, 0, ncg, g->hmol);
      g->mark_end(kStCore, e);
    }
    g->have_density = true;
    g->have_eval = false;
    g->have_v = false;
    return;
  }
  g->mark_begin(kStPermute, &e);
  launch_permute_d(g->stream, g->nbasis, g->npad, dD, g->d_int2ref.p, g->d_Dint.p);
  g->mark_end(kStPermute, e);
  g->mark_begin(kStDensity, &e);
  if (g->cached) {
    launch_density_c64(g->stream, 0, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_slots.p, g->d_grp_ntiles.p,
                       g->d_bt_off.p, g->d_bt_list.p, g->d_cache.p, g->d_Dint.p, g->npad, g->d_part.p);
    if (g->type >= 2)  // tau: three more contractions with A = d_r phi
      for (int r = 1; r <= 3; ++r)
        launch_density_c64(g->stream, r, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_slots.p, g->d_grp_ntiles.p,
                           g->d_bt_off.p, g->d_bt_list.p, g->d_cache.p, g->d_Dint.p, g->npad,
                           g->d_part_tau.p + (size_t)(r - 1) * kBatch);
  } else {  // AO values recomputed in the kernels
    launch_density_f64(g->stream, 0, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                       g->d_bt_list.p, g->d_x.p, g->d_y.p, g->d_z.p, g->view(), g->d_Dint.p, g->npad, g->d_part.p);
    if (g->type >= 2)
      for (int r = 1; r <= 3; ++r)
        launch_density_f64(g->stream, r, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                           g->d_bt_list.p, g->d_x.p, g->d_y.p, g->d_z.p, g->view(), g->d_Dint.p, g->npad,
                           g->d_part_tau.p + (size_t)(r - 1) * kBatch);
  }
  g->mark_end(kStDensity, e);
  g->have_density = true;
  g->have_eval = false;
  g->have_v = false;
}

// spin: the potential's spin index, whose split remainder of D enters the
// energy correction (mixed form, when the last Evaluate computed E and v)
void fock_to_device(chx_grid* g, double* dV, int spin = 0) {
  hipEvent_t e = nullptr;
  if (g->precision == CHX_PREC_MIXED) {
    const int nc = g->v_family >= 2 ? 4 : 1;
    require(nc <= g->mx_nT, CHX_ERR_STATE, "a meta-GGA potential needs a grid of type 2 before getAO");
    g->mark_begin(kStMakeB, &e);
    launch_make_b_mx(g->stream, nc, g->hmol);
    g->mark_end(kStMakeB, e);
    launch_fill(g->stream, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
    g->mark_begin(kStVxc, &e);
    if (g->mx_wide)
      launch_vxc_mx_wide(g->stream, nc, (int)g->cwork.size(), g->hmol);
    else
      launch_vxc_mx(g->stream, nc, (int)g->cwork.size(), g->hmol);
    g->mark_end(kStVxc, e);
    g->mark_begin(kStFinalize, &e);
    if (g->corr_pending)
      launch_finalize_corr(g->stream, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, dV,
                           g->d_Dres.p + (size_t)spin * g->nbasis * g->nbasis,
                           g->d_cpart.p + (size_t)spin * corr_blocks(g->nbasis));
    else
      launch_finalize_sym(g->stream, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, dV);
    g->mark_end(kStFinalize, e);
    return;
  }
  if (g->cached) {
    const int nc = g->v_family >= 2 ? 4 : 1;
    g->mark_begin(kStMakeB, &e);
    launch_make_b(g->stream, g->nbatches, g->d_bt_off.p, g->d_coef.p, g->d_cache.p, g->d_bmat.p);
    g->mark_end(kStMakeB, e);
    launch_fill(g->stream, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
    g->mark_begin(kStVxc, &e);
    launch_vxc_c64(g->stream, nc, (int)g->cwork.size(), g->d_cwork.p, g->d_cpair_batches.p, g->d_slot_of.p, g->ntiles,
                   g->d_cache.p, g->d_bmat.p, g->d_coef.p, g->d_Mint.p, g->npad);
    g->mark_end(kStVxc, e);
    g->mark_begin(kStFinalize, &e);
    launch_finalize_sym(g->stream, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, dV);
    g->mark_end(kStFinalize, e);
    return;
  }
  launch_fill(g->stream, g->d_Vint.p, (size_t)g->npad * g->npad, 0.0);
  g->mark_begin(kStVxc, &e);
  launch_vxc_f64(g->stream, g->v_family >= 2 ? 4 : 1, (int)g->vwork.size(), g->d_vwork.p, g->d_pair_batches.p,
                 g->d_alive.p, g->ntiles, g->d_x.p, g->d_y.p, g->d_z.p, g->d_coef.p, g->view(), g->d_Vint.p, g->npad);
  g->mark_end(kStVxc, e);
  g->mark_begin(kStFinalize, &e);
  launch_finalize_v(g->stream, g->nbasis, g->npad, g->d_Vint.p, g->d_ref2int.p, dV);
  g->mark_end(kStFinalize, e);
}

void check_dm_args(const chx_grid* g, int nspin, int nbasis, int max_spin = 1) {
  require(g->ao_ready, CHX_ERR_STATE, "Grid::getAO must be called before the density");
  require(nspin >= 1 && nspin <= 2, CHX_ERR_ARG, "Wrong number of spin types!");
  require(nspin <= max_spin, CHX_ERR_UNSUPPORTED, "this call is built for the closed-shell (nspin = 1) path");
  require(nbasis == g->nbasis, CHX_ERR_ARG, "density matrix size does not match the basis");
}

// Density of one or two spin matrices at dD (device, back to back): beta first,
// its partials kept in d_part_b, then alpha in d_part (GridDensity.cpp:12-31
// per spin index).
void density_spin(chx_grid* g, const double* dD, int nspin) {
  if (nspin == 2) {
    density_from_device(g, dD + (size_t)g->nbasis * g->nbasis, 1);
    g->d_part_b.alloc(g->d_part.n);
    CHX_HIP(hipMemcpyAsync(g->d_part_b.p, g->d_part.p, sizeof(double) * g->d_part.n, hipMemcpyDeviceToDevice,
                           g->stream));
    if (g->type >= 2) {  // the tau partials of beta too
      g->d_part_tau_b.alloc(g->d_part_tau.n);
      CHX_HIP(hipMemcpyAsync(g->d_part_tau_b.p, g->d_part_tau.p, sizeof(double) * g->d_part_tau.n,
                             hipMemcpyDeviceToDevice, g->stream));
    }
  }
  density_from_device(g, dD);
  g->nspin = nspin;
}

// Vxc of each spin: alpha from d_coef, then beta with d_coef_b swapped in
// (the alpha coefficients are restored afterwards). Mixed form after an
// energy evaluation: the energy's correction for the split of D,
// sum_s V_s . (D_s - hi - lo), into d_en[2].
void fock_spin(chx_grid* g, double* dV, int nspin) {
  require(nspin == g->v_nspin, CHX_ERR_ARG, "Fock spin count differs from the evaluated density's");
  fock_to_device(g, dV, 0);
  if (nspin == 2) {
    const size_t nc = g->d_coef.n, nb2 = (size_t)g->nbasis * g->nbasis;
    g->d_coef_save.alloc(nc);
    CHX_HIP(hipMemcpyAsync(g->d_coef_save.p, g->d_coef.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, g->stream));
    CHX_HIP(hipMemcpyAsync(g->d_coef.p, g->d_coef_b.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, g->stream));
    fock_to_device(g, dV + nb2, 1);
    CHX_HIP(hipMemcpyAsync(g->d_coef.p, g->d_coef_save.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, g->stream));
  }
  if (g->precision == CHX_PREC_MIXED && g->corr_pending) {
    launch_reduce_corr(g->stream, nspin * corr_blocks(g->nbasis), g->d_cpart.p, g->d_en.p + 2);
    g->corr_done = true;
  }
}

// The potential of the last Evaluate into d_Vout (mixed form, before the
// energy is read: the energy's correction needs it); later Fock calls copy it.
void build_vout(chx_grid* g) {
  const size_t n = (size_t)g->v_nspin * g->nbasis * g->nbasis;
  if (g->d_Vout.n < n) g->d_Vout.alloc(n);
  fock_spin(g, g->d_Vout.p, g->v_nspin);
  g->vout_ready = true;
  g->vout_nspin = g->v_nspin;
}
}  // namespace

extern "C" {

const char* chx_last_error(void) { return g_last_error.c_str(); }

int chx_device_count(int* n) {
  return guard([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
  });
}

// ---- molecule -------------------------------------------------------------
int chx_mol_create(int natoms, const int* Z, const double* xyz, chx_mol** out) {
  return guard([&] {
    require(natoms > 0 && Z && xyz && out, CHX_ERR_ARG, "bad molecule arguments");
    chx_mol* m = new chx_mol;
    m->natoms = natoms;
    m->Z.assign(Z, Z + natoms);
    m->xyz.assign(xyz, xyz + 3 * natoms);
    *out = m;
  });
}

int chx_mol_set_basis_file(chx_mol* m, const char* data_root, const char* basis_name) {
  return guard([&] {
    require(m && data_root && basis_name, CHX_ERR_ARG, "bad arguments");
    mol_set_basis_file(m, data_root, basis_name);
  });
}

int chx_mol_add_shell(chx_mol* m, int center, int type, int nprim, const double* exps, const double* coefs) {
  return guard([&] {
    require(m && center >= 0 && center < m->natoms && nprim > 0 && exps && coefs, CHX_ERR_ARG, "bad shell");
    require(type >= -3 && type <= 1, CHX_ERR_UNSUPPORTED, "angular momentum above f is not supported");
    ShellH s{type, center, nprim, (int)m->exps.size(), m->nbasis};
    m->exps.insert(m->exps.end(), exps, exps + nprim);
    m->coefs.insert(m->coefs.end(), coefs, coefs + nprim);
    m->shells.push_back(s);
    m->nbasis += shell_size(type);
  });
}

int chx_mol_nbasis(const chx_mol* m, int* n) {
  return guard([&] { *n = m->nbasis; });
}

int chx_mol_nshells(const chx_mol* m, int* n) {
  return guard([&] { *n = (int)m->shells.size(); });
}

int chx_mol_get_shells(const chx_mol* m, int* type, int* nprim, int* center, double* exps, double* coefs) {
  return guard([&] {
    for (size_t i = 0; i < m->shells.size(); ++i) {
      if (type) type[i] = m->shells[i].type;
      if (nprim) nprim[i] = m->shells[i].nprim;
      if (center) center[i] = m->shells[i].center;
    }
    if (exps) std::memcpy(exps, m->exps.data(), sizeof(double) * m->exps.size());
    if (coefs) std::memcpy(coefs, m->coefs.data(), sizeof(double) * m->coefs.size());
  });
}

int chx_mol_overlap(const chx_mol* m, double* S) {
  return guard([&] { mol_overlap(m, S); });
}

void chx_mol_destroy(chx_mol* m) { delete m; }

// ---- grid -----------------------------------------------------------------
int chx_grid_create(const chx_mol* m, const char* grid_name, const char* data_root, int nthreads, int output,
                    chx_grid** out) {
  chx_grid* g = nullptr;
  int rc = guard([&] {
    require(m && grid_name && data_root && out, CHX_ERR_ARG, "bad grid arguments");
    g = new chx_grid;
    g->mol = *m;
    g->nthreads = nthreads > 0 ? nthreads : 1;
    grid_build(g, grid_name, data_root);  // reads the grid files, then creates the stream
    if (output)
      std::printf("Grid %s: %ld points in %d batches of %d; %d AOs in %d tiles\n", grid_name, g->npoints, g->nbatches,
                  kBatch, g->nbasis, g->ntiles);
  });
  if (rc != CHX_OK) {
    if (g) chx_grid_destroy(g);
    return rc;
  }
  *out = g;
  return CHX_OK;
}

int chx_grid_copy(const chx_grid* s, chx_grid** out) {
  chx_grid* g = nullptr;
  int rc = guard([&] {
    require(s && out, CHX_ERR_ARG, "bad arguments");
    g = new chx_grid;
    g->nthreads = s->nthreads; g->type = s->type; g->precision = s->precision; g->screen_thr = s->screen_thr;
    g->cache_mode = s->cache_mode;
    g->core_radius = s->core_radius;
    g->mol = s->mol;
    g->nbasis = s->nbasis; g->npad = s->npad; g->ntiles = s->ntiles;
    g->int2ref = s->int2ref; g->ref2int = s->ref2int;
    g->sh_xyz = s->sh_xyz; g->sh_type = s->sh_type; g->sh_nprim = s->sh_nprim; g->sh_poff = s->sh_poff;
    g->p_exp = s->p_exp; g->p_coef = s->p_coef; g->tile_seg_off = s->tile_seg_off; g->seg = s->seg;
    g->npoints = s->npoints; g->nbatches = s->nbatches; g->batch_n = s->batch_n;
    g->hx = s->hx; g->hy = s->hy; g->hz = s->hz; g->hw = s->hw;
    make_stream(g);
    hipStream_t st = g->stream;
    CHX_HIP(hipStreamSynchronize(s->stream));
    g->d_sh_xyz.copy_from(s->d_sh_xyz, st); g->d_p_exp.copy_from(s->d_p_exp, st); g->d_p_coef.copy_from(s->d_p_coef, st);
    g->d_segrec.copy_from(s->d_segrec, st); g->max_nprim = s->max_nprim;
    g->sh_atom = s->sh_atom; g->d_sh_atom.copy_from(s->d_sh_atom, st);
    g->d_sh_type.copy_from(s->d_sh_type, st); g->d_sh_nprim.copy_from(s->d_sh_nprim, st);
    g->d_sh_poff.copy_from(s->d_sh_poff, st); g->d_tile_seg_off.copy_from(s->d_tile_seg_off, st);
    g->d_seg.copy_from(s->d_seg, st); g->d_int2ref.copy_from(s->d_int2ref, st); g->d_ref2int.copy_from(s->d_ref2int, st);
    g->d_x.copy_from(s->d_x, st); g->d_y.copy_from(s->d_y, st); g->d_z.copy_from(s->d_z, st);
    g->d_w.copy_from(s->d_w, st); g->d_bn.copy_from(s->d_bn, st);
    CHX_HIP(hipStreamSynchronize(st));
    if (s->ao_ready) grid_plan(g);
  });
  if (rc != CHX_OK) {
    if (g) chx_grid_destroy(g);
    return rc;
  }
  *out = g;
  return CHX_OK;
}

void chx_grid_destroy(chx_grid* g) {
  if (!g) return;
  try {
    g->collect_times();
  } catch (...) {
  }
  if (g->stream) {
    (void)hipStreamSynchronize(g->stream);
    (void)hipStreamDestroy(g->stream);
  }
  delete g;
}

int chx_grid_num_threads(const chx_grid* g, int* n) {
  return guard([&] { *n = g->nthreads; });
}

int chx_grid_num_points(const chx_grid* g, long* n) {
  return guard([&] { *n = g->npoints; });
}

int chx_grid_get_points(const chx_grid* g, double* xyzw) {
  return guard([&] {
    long k = 0;
    for (int b = 0; b < g->nbatches; ++b)
      for (int i = 0; i < g->batch_n[b]; ++i, ++k) {
        const size_t s = (size_t)b * kBatch + i;
        xyzw[4 * k] = g->hx[s]; xyzw[4 * k + 1] = g->hy[s]; xyzw[4 * k + 2] = g->hz[s]; xyzw[4 * k + 3] = g->hw[s];
      }
  });
}

int chx_grid_set_type(chx_grid* g, int type) {
  return guard([&] {
    require(type >= -1 && type <= 2, CHX_ERR_ARG, "grid type must be -1, 0, 1 or 2");
    if (g->type != type) g->ao_ready = false;  // tau buffers are sized at getAO
    g->type = type;
  });
}

int chx_grid_get_ao(chx_grid* g, int derivative, int output) {
  return guard([&] {
    require(derivative >= 0 && derivative <= 2, CHX_ERR_UNSUPPORTED, "AO derivatives above second order are not built");
    // derivative 1 (before the XC gradient, Restricted/Grad.cpp:73) or 2
    // (before the Hessian's skeleton terms, Restricted/Hess.cpp:78): the tables
    // of a planned grid already serve them (the kernels that need second AO
    // derivatives form them in place), and the density and potentials of the
    // last iteration must survive, as in the reference.
    if (derivative >= 1 && g->ao_ready) return;
    grid_plan(g);
    if (output)
      std::printf("AO screening: %.2f of %d tiles alive per batch, %.3e FLOP per XC build\n", g->mean_tiles, g->ntiles,
                  g->flops_build);
  });
}

int chx_grid_get_density(chx_grid* g, const double* D, int nspin, int nbasis) {
  return guard([&] {
    check_dm_args(g, nspin, nbasis, 2);
    require(D != nullptr, CHX_ERR_ARG, "null density matrix");
    const size_t n = (size_t)nspin * nbasis * nbasis;
    if (g->d_Din.n < n) g->d_Din.alloc(n);
    CHX_HIP(hipMemcpyAsync(g->d_Din.p, D, sizeof(double) * n, hipMemcpyHostToDevice, g->stream));
    density_spin(g, g->d_Din.p, nspin);
  });
}

int chx_grid_get_density_skeleton(chx_grid* g, const double* D, int nspin, int nbasis) {
  return guard([&] {
    check_dm_args(g, nspin, nbasis);
    require(D != nullptr, CHX_ERR_ARG, "null density matrix");
    require(g->precision == CHX_PREC_FP64 && g->cached, CHX_ERR_UNSUPPORTED,
            "the XC gradient is built for the FP64 form with the AO tables in HBM");
    CHX_HIP(hipMemcpyAsync(g->d_Din.p, D, sizeof(double) * (size_t)nbasis * nbasis, hipMemcpyHostToDevice, g->stream));
    launch_permute_d(g->stream, g->nbasis, g->npad, g->d_Din.p, g->d_int2ref.p, g->d_Dskel.alloc_get((size_t)g->npad * g->npad));
    g->have_skeleton = true;
  });
}

int chx_grid_get_density_u(chx_grid* g, const double* D, int nmats, int nspin, int nbasis) {
  return guard([&] {
    check_dm_args(g, nspin, nbasis, 2);
    require(D != nullptr && nmats >= 1, CHX_ERR_ARG, "bad perturbed density matrices");
    require(g->precision == CHX_PREC_FP64, CHX_ERR_UNSUPPORTED, "the response path is built for the FP64 form");
    const size_t nslots = (size_t)g->nbatches * kBatch, nb2 = (size_t)nbasis * nbasis;
    const int ng = (int)g->grp_batch.size();
    double* part = g->d_part_u.alloc_get(std::max<size_t>(1, (size_t)ng * 4 * kBatch));
    double* rhou = g->d_rhou.alloc_get(std::max<size_t>(1, (size_t)nspin * nmats * 4 * nslots));
    for (int m = 0; m < nspin * nmats; ++m) {  // rho^x and grad rho^x of each matrix ([spin][mat]), per point
      CHX_HIP(hipMemcpyAsync(g->d_Din.p, D + m * nb2, sizeof(double) * nb2, hipMemcpyHostToDevice, g->stream));
      launch_permute_d(g->stream, g->nbasis, g->npad, g->d_Din.p, g->d_int2ref.p, g->d_Dint.p);
      if (g->cached)
        launch_density_c64(g->stream, 0, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_slots.p, g->d_grp_ntiles.p,
                           g->d_bt_off.p, g->d_bt_list.p, g->d_cache.p, g->d_Dint.p, g->npad, part);
      else
        launch_density_f64(g->stream, 0, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                           g->d_bt_list.p, g->d_x.p, g->d_y.p, g->d_z.p, g->view(), g->d_Dint.p, g->npad, part);
      launch_point_rho(g->stream, g->nbatches, g->d_grp_off.p, part, rhou + (size_t)m * 4 * nslots);
    }
    CHX_HIP(hipStreamSynchronize(g->stream));
    g->nmats_u = nmats;
    g->nspin_u = nspin;
  });
}

// Response matrices of nmats perturbed densities held per point in rhou
// ([spin][mat][point][rho, grad rho], point = batch-major slot): per matrix
// (and spin, nspin = 2) the response coefficients (k_response_coef /
// k_response_coef_pol), then the ground-state Vxc pipeline on them
// (B = a phi + c . grad phi, M = Phi^T B, V = M + M^T), V to F (host,
// [spin][mat]).
static void response_fock(chx_grid* g, const double* rhou, int nmats, double* F, int nspin = 1) {
  const size_t nslots = (size_t)g->nbatches * kBatch, nb2 = (size_t)g->nbasis * g->nbasis;
  g->vout_ready = false;  // d_Vout stages the response matrices
  double* coef = g->d_coef_u.alloc_get(std::max<size_t>(kCoef, kCoef * nslots));
  for (int m = 0; m < nspin * nmats; ++m) {
    if (nspin == 2)  // output m = spin u, perturbation m % nmats (both spins' rho^x enter)
      launch_response_coef_pol(g->stream, g->nbatches, rhou + (size_t)(m % nmats) * 4 * nslots,
                               rhou + (size_t)(nmats + m % nmats) * 4 * nslots, g->d_fx.p, g->d_w.p, g->fx_family,
                               m / nmats, coef);
    else
      launch_response_coef(g->stream, g->nbatches, rhou + (size_t)m * 4 * nslots, g->d_fx.p, g->d_w.p, g->fx_family,
                           coef);
    if (g->cached) {  // B = a phi + c . grad phi from the tables, M = Phi^T B, V = M + M^T
      launch_make_b(g->stream, g->nbatches, g->d_bt_off.p, coef, g->d_cache.p, g->d_bmat.p);
      launch_fill(g->stream, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
      launch_vxc_c64(g->stream, 1, (int)g->cwork.size(), g->d_cwork.p, g->d_cpair_batches.p, g->d_slot_of.p,
                     g->ntiles, g->d_cache.p, g->d_bmat.p, coef, g->d_Mint.p, g->npad);
      launch_finalize_sym(g->stream, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, g->d_Vout.p);
    } else {
      launch_fill(g->stream, g->d_Vint.p, (size_t)g->npad * g->npad, 0.0);
      launch_vxc_f64(g->stream, 1, (int)g->vwork.size(), g->d_vwork.p, g->d_pair_batches.p, g->d_alive.p,
                     g->ntiles, g->d_x.p, g->d_y.p, g->d_z.p, coef, g->view(), g->d_Vint.p, g->npad);
      launch_finalize_v(g->stream, g->nbasis, g->npad, g->d_Vint.p, g->d_ref2int.p, g->d_Vout.p);
    }
    CHX_HIP(hipMemcpyAsync(F + m * nb2, g->d_Vout.p, sizeof(double) * nb2, hipMemcpyDeviceToHost, g->stream));
  }
  CHX_HIP(hipStreamSynchronize(g->stream));
}

int chx_grid_fock_u(chx_grid* g, double* F, int nmats, int nspin, int nbasis) {
  return guard([&] {
    require(F != nullptr, CHX_ERR_ARG, "null output");
    require(nspin == 1 || nspin == 2, CHX_ERR_ARG, "Wrong number of spin types!");
    require(nbasis == g->nbasis, CHX_ERR_ARG, "Fock matrix size does not match the basis");
    require(g->fx_family >= 0, CHX_ERR_STATE, "Evaluate(\"f\") must precede Grid::getFockU");
    require(g->nmats_u > 0, CHX_ERR_STATE, "Grid::getDensityU must precede Grid::getFockU");
    require(nmats == g->nmats_u, CHX_ERR_ARG, "number of matrices differs from the last Grid::getDensityU");
    require(nspin == g->nspin_u && nspin == g->fx_nspin, CHX_ERR_ARG,
            "spin count differs from the last Grid::getDensityU / Evaluate(\"f\")");
    response_fock(g, g->d_rhou.p, nmats, F, nspin);
  });
}

// slot -> (density group, column tile in the group), for the per-slot epilogues
static void ensure_slot_groups(chx_grid* g) {
  const int ng = (int)g->grp_batch.size(), ns = (int)g->bt_list.size();
  if (g->slot_grp.size() == (size_t)ns) return;
  g->slot_grp.assign(ns, 0);
  g->slot_q.assign(ns, 0);
  for (int k = 0; k < ng; ++k)
    for (int q = 0; q < g->grp_ntiles[k]; ++q) {
      const int sl = g->grp_slots[(size_t)k * kGroup + q];
      g->slot_grp[sl] = k;
      g->slot_q[sl] = q;
    }
  g->d_slot_grp.upload(g->slot_grp, g->stream);
  g->d_slot_q.upload(g->slot_q, g->stream);
}

int chx_grid_fock_s(chx_grid* g, double* F, int nspin, int nbasis) {
  return guard([&] {
    require(F != nullptr, CHX_ERR_ARG, "null output");
    require(nspin == 1, CHX_ERR_UNSUPPORTED, "only the closed-shell path is built");
    require(nbasis == g->nbasis, CHX_ERR_ARG, "Fock matrix size does not match the basis");
    require(g->have_skeleton, CHX_ERR_STATE, "Grid::getDensitySkeleton must precede Grid::getFockU<s_t>");
    require(g->fx_family >= 0, CHX_ERR_STATE, "Evaluate(\"f\") must precede Grid::getFockU<s_t>");
    require(g->fx_nspin == 1, CHX_ERR_UNSUPPORTED, "the skeleton response is built for the closed-shell path");
    require(g->precision == CHX_PREC_FP64 && g->cached, CHX_ERR_UNSUPPORTED,
            "the skeleton response is built for the FP64 form with the AO tables in HBM");
    const int na = g->mol.natoms, ng = (int)g->grp_batch.size(), ns = (int)g->bt_list.size();
    const long npt = (long)g->nbatches * kBatch;
    ensure_slot_groups(g);
    // per point and nuclear coordinate: rho^{A,t}, grad rho^{A,t}
    const size_t nxz = std::max<size_t>(1, (size_t)ng * 2 * kGroup * kTile * kBatch);
    DevBuf<double> xz1, xz2;
    xz1.alloc(nxz);
    xz2.alloc(nxz);
    double* rhos = g->d_rhou.alloc_get(std::max<size_t>(1, (size_t)3 * na * 4 * npt));
    launch_fill(g->stream, rhos, (size_t)3 * na * 4 * npt, 0.0);
    launch_xgrad_xz(g->stream, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                    g->d_bt_list.p, g->d_cache.p, g->d_bmat.p, g->d_Dskel.p, g->npad, 0, 1, xz1.p);
    launch_xgrad_xz(g->stream, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                    g->d_bt_list.p, g->d_cache.p, g->d_bmat.p, g->d_Dskel.p, g->npad, 2, 3, xz2.p);
    launch_skel_rho(g->stream, ns, g->d_slot_batch.p, g->d_slot_tile.p, g->d_slot_grp.p, g->d_slot_q.p, g->d_x.p,
                    g->d_y.p, g->d_z.p, g->view(), xz1.p, xz2.p, npt, rhos);
    g->nmats_u = 0;  // rhou now holds the skeleton densities, not getDensityU's
    response_fock(g, rhos, 3 * na, F);
  });
}

int chx_grid_fock_skeleton(chx_grid* g, double* F, int nspin, int nbasis) {
  return guard([&] {
    require(F != nullptr, CHX_ERR_ARG, "null output");
    require(nspin == 1, CHX_ERR_UNSUPPORTED, "only the closed-shell path is built");
    require(nbasis == g->nbasis, CHX_ERR_ARG, "Fock matrix size does not match the basis");
    require(g->have_v, CHX_ERR_STATE, "ExchangeCorrelation::Evaluate(\"ev\") must precede Grid::getFockSkeleton");
    require(g->v_family <= 1 && g->type <= 1, CHX_ERR_UNSUPPORTED,
            "the skeleton Fock derivative is built for LDA and GGA");
    require(g->v_nspin == 1, CHX_ERR_UNSUPPORTED, "the skeleton Fock derivative is built for the closed-shell path");
    require(g->precision == CHX_PREC_FP64 && g->cached, CHX_ERR_UNSUPPORTED,
            "the skeleton Fock derivative is built for the FP64 form with the AO tables in HBM");
    const int na = g->mol.natoms, ns = (int)g->bt_list.size(), n = g->nbasis;
    const size_t pp = (size_t)g->npad * g->npad, sb = (size_t)kBatch * kTile;
    // atom of each AO, reference order
    std::vector<int> ao_atom(n, 0);
    for (const chx::ShellH& sh : g->mol.shells)
      for (int k = 0; k < chx::shell_size(sh.type); ++k) ao_atom[sh.ao_off + k] = sh.center;
    DevBuf<int> d_atom;
    d_atom.upload(ao_atom, g->stream);
    launch_make_b(g->stream, g->nbatches, g->d_bt_off.p, g->d_coef.p, g->d_cache.p, g->d_bmat.p);
    DevBuf<double> etab, qp, out;
    etab.alloc(std::max<size_t>(1, 3 * (size_t)ns * sb));
    launch_fill(g->stream, etab.p, 3 * (size_t)ns * sb, 0.0);
    launch_skel_etab(g->stream, ns, g->d_slot_batch.p, g->d_slot_tile.p, g->d_x.p, g->d_y.p, g->d_z.p, g->view(),
                     g->d_coef.p, etab.p);
    qp.alloc(6 * pp);
    launch_fill(g->stream, qp.p, 6 * pp, 0.0);
    for (int t = 0; t < 3; ++t) {
      // Q^t = (d_t Phi)^T B: the cached component t + 1 as the row operand
      launch_vxc_c64(g->stream, 1, (int)g->cwork.size(), g->d_cwork.p, g->d_cpair_batches.p, g->d_slot_of.p,
                     g->ntiles, g->d_cache.p + (size_t)(t + 1) * sb, g->d_bmat.p, g->d_coef.p, qp.p + t * pp,
                     g->npad);
      // P^t = Phi^T E^t: E^t in the B table's place
      launch_vxc_c64(g->stream, 1, (int)g->cwork.size(), g->d_cwork.p, g->d_cpair_batches.p, g->d_slot_of.p,
                     g->ntiles, g->d_cache.p, etab.p + (size_t)t * ns * sb, g->d_coef.p, qp.p + (3 + t) * pp,
                     g->npad);
    }
    const size_t tot = (size_t)3 * na * n * n;
    out.alloc(std::max<size_t>(1, tot));
    launch_fock_skel_fin(g->stream, n, g->npad, na, qp.p, g->d_ref2int.p, d_atom.p, out.p);
    CHX_HIP(hipMemcpyAsync(F, out.p, sizeof(double) * tot, hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
  });
}

int chx_grid_get_density_skeleton2(chx_grid* g, const double* D, int nspin, int nbasis) {
  // the second-order skeleton densities are formed inside chx_grid_energy_hess from the same D
  return chx_grid_get_density_skeleton(g, D, nspin, nbasis);
}

int chx_grid_energy_hess(chx_grid* g, double* H) {
  return guard([&] {
    require(H != nullptr, CHX_ERR_ARG, "null output");
    require(g->have_skeleton, CHX_ERR_STATE, "Grid::getDensitySkeleton2 must precede Grid::getEnergyHess");
    require(g->have_v, CHX_ERR_STATE, "ExchangeCorrelation::Evaluate(\"ev\") must precede Grid::getEnergyHess");
    require(g->fx_family >= 0, CHX_ERR_STATE, "Evaluate(\"f\") must precede Grid::getEnergyHess");
    require(g->v_family <= 1 && g->type <= 1 && g->fx_family <= 1, CHX_ERR_UNSUPPORTED,
            "the XC Hessian is built for LDA and GGA");
    require(g->v_nspin == 1 && g->fx_nspin == 1, CHX_ERR_UNSUPPORTED, "the XC Hessian is built for the closed-shell path");
    require(g->precision == CHX_PREC_FP64 && g->cached, CHX_ERR_UNSUPPORTED,
            "the XC Hessian is built for the FP64 form with the AO tables in HBM");
    const int na = g->mol.natoms, np = 3 * na, ng = (int)g->grp_batch.size(), ns = (int)g->bt_list.size();
    const long npt = (long)g->nbatches * kBatch;
    const size_t pp = (size_t)g->npad * g->npad, sb = (size_t)kBatch * kTile;
    ensure_slot_groups(g);
    // atom of each internal AO (-1: padding)
    std::vector<int> ao_atom(g->nbasis, 0), atom_int(g->npad, -1);
    for (const chx::ShellH& sh : g->mol.shells)
      for (int k = 0; k < chx::shell_size(sh.type); ++k) ao_atom[sh.ao_off + k] = sh.center;
    for (int i = 0; i < g->npad; ++i)
      if (g->int2ref[i] >= 0) atom_int[i] = ao_atom[g->int2ref[i]];
    DevBuf<int> d_atom;
    d_atom.upload(atom_int, g->stream);
    // X = D phi, Y^x = D d_x phi (xz1); Y^y, Y^z (xz2); first-order skeleton densities
    const size_t nxz = std::max<size_t>(1, (size_t)ng * 2 * kGroup * kTile * kBatch);
    DevBuf<double> xz1, xz2, rhos, etab, rr, dH;
    xz1.alloc(nxz);
    xz2.alloc(nxz);
    launch_xgrad_xz(g->stream, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                    g->d_bt_list.p, g->d_cache.p, g->d_bmat.p, g->d_Dskel.p, g->npad, 0, 1, xz1.p);
    launch_xgrad_xz(g->stream, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p,
                    g->d_bt_list.p, g->d_cache.p, g->d_bmat.p, g->d_Dskel.p, g->npad, 2, 3, xz2.p);
    rhos.alloc(std::max<size_t>(1, (size_t)np * 4 * npt));
    launch_fill(g->stream, rhos.p, (size_t)np * 4 * npt, 0.0);
    launch_skel_rho(g->stream, ns, g->d_slot_batch.p, g->d_slot_tile.p, g->d_slot_grp.p, g->d_slot_q.p, g->d_x.p,
                    g->d_y.p, g->d_z.p, g->view(), xz1.p, xz2.p, npt, rhos.p);
    // rr[3 t + s] = (d_t Phi)^T E^s over the cached contraction
    launch_make_b(g->stream, g->nbatches, g->d_bt_off.p, g->d_coef.p, g->d_cache.p, g->d_bmat.p);
    etab.alloc(std::max<size_t>(1, 3 * (size_t)ns * sb));
    launch_fill(g->stream, etab.p, 3 * (size_t)ns * sb, 0.0);
    launch_skel_etab(g->stream, ns, g->d_slot_batch.p, g->d_slot_tile.p, g->d_x.p, g->d_y.p, g->d_z.p, g->view(),
                     g->d_coef.p, etab.p);
    rr.alloc(9 * pp);
    launch_fill(g->stream, rr.p, 9 * pp, 0.0);
    for (int t = 0; t < 3; ++t)
      for (int u = 0; u < 3; ++u)
        launch_vxc_c64(g->stream, 1, (int)g->cwork.size(), g->d_cwork.p, g->d_cpair_batches.p, g->d_slot_of.p,
                       g->ntiles, g->d_cache.p + (size_t)(t + 1) * sb, etab.p + (size_t)u * ns * sb, g->d_coef.p,
                       rr.p + (size_t)(3 * t + u) * pp, g->npad);
    dH.alloc((size_t)np * np);
    launch_fill(g->stream, dH.p, (size_t)np * np, 0.0);
    launch_hess_terms(g->stream, ns, g->d_slot_batch.p, g->d_slot_tile.p, g->d_slot_grp.p, g->d_slot_q.p, g->d_x.p,
                      g->d_y.p, g->d_z.p, g->view(), g->d_coef.p, xz1.p, xz2.p, npt, rhos.p, g->d_fx.p, g->d_w.p,
                      g->fx_family, g->npad, g->d_Dskel.p, rr.p, d_atom.p, np, dH.p);
    CHX_HIP(hipMemcpyAsync(H, dH.p, sizeof(double) * np * np, hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
  });
}

int chx_grid_energy_grad(chx_grid* g, double* grad) {
  return guard([&] {
    require(grad != nullptr, CHX_ERR_ARG, "null output");
    require(g->have_skeleton, CHX_ERR_STATE, "Grid::getDensitySkeleton must precede Grid::getEnergyGrad");
    require(g->have_v, CHX_ERR_STATE, "ExchangeCorrelation::Evaluate(\"ev\") must precede Grid::getEnergyGrad");
    require(g->v_family <= 1 && g->type <= 1, CHX_ERR_UNSUPPORTED, "the XC gradient is built for LDA and GGA");
    require(g->v_nspin == 1, CHX_ERR_UNSUPPORTED, "the XC gradient is built for the closed-shell path");
    const int na = g->mol.natoms, ng = (int)g->grp_batch.size(), ns = (int)g->bt_list.size();
    ensure_slot_groups(g);
    launch_make_b(g->stream, g->nbatches, g->d_bt_off.p, g->d_coef.p, g->d_cache.p, g->d_bmat.p);
    DevBuf<double> xz, dg;
    xz.alloc(std::max<size_t>(1, (size_t)ng * 2 * kGroup * kTile * kBatch));
    dg.alloc(3 * (size_t)na);
    launch_fill(g->stream, dg.p, 3 * (size_t)na, 0.0);
    launch_xgrad_xz(g->stream, ng, g->d_grp_batch.p, g->d_grp_tiles.p, g->d_grp_ntiles.p, g->d_bt_off.p, g->d_bt_list.p,
                    g->d_cache.p, g->d_bmat.p, g->d_Dskel.p, g->npad, 0, -1, xz.p);
    launch_xgrad_epi(g->stream, ns, g->d_slot_batch.p, g->d_slot_tile.p, g->d_slot_grp.p, g->d_slot_q.p, g->d_x.p,
                     g->d_y.p, g->d_z.p, g->view(), g->d_coef.p, xz.p, dg.p);
    CHX_HIP(hipMemcpyAsync(grad, dg.p, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
  });
}

int chx_grid_autotune(chx_grid* g, const chx_xc* xc, const double* D_dev, double* V_dev, int reps, int* chunk_out,
                      float* ms_out) {
  return guard([&] {
    require(g && xc && D_dev && V_dev && reps > 0, CHX_ERR_ARG, "bad arguments");
    require(g->ao_ready && g->cached, CHX_ERR_STATE, "Grid::getAO (with the AO tables in HBM) must precede tuning");
    require(g->type >= xc->family, CHX_ERR_STATE, "grid type is below the functional family (call setType)");
    static const int cand[] = {8, 16, 32, 128};
    hipEvent_t a, b;
    CHX_HIP(hipEventCreate(&a));
    CHX_HIP(hipEventCreate(&b));
    const bool prof = g->profile;
    g->profile = false;
    auto build = [&] {
      density_from_device(g, D_dev);
      g->nspin = 1;
      run_functional(g, xc, true, true);
      g->v_family = xc->family;
      g->have_eval = g->have_v = true;
      fock_to_device(g, V_dev);
    };
    int best = -1;
    float best_ms = 0.f;
    for (int k = 0; k < 4; ++k) {
      set_vxc_chunk(g, cand[k]);
      build();  // warm-up
      CHX_HIP(hipEventRecord(a, g->stream));
      for (int r = 0; r < reps; ++r) build();
      CHX_HIP(hipEventRecord(b, g->stream));
      CHX_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      CHX_HIP(hipEventElapsedTime(&ms, a, b));
      ms /= (float)reps;
      if (ms_out) ms_out[k] = ms;
      if (best < 0 || ms < best_ms) {
        best = k;
        best_ms = ms;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    g->profile = prof;
    set_vxc_chunk(g, cand[best]);
    if (chunk_out) *chunk_out = cand[best];
  });
}

int chx_grid_set_vxc_chunk(chx_grid* g, int chunk) {
  return guard([&] {
    require(g && chunk >= 0 && chunk <= kMxMaxChunk, CHX_ERR_ARG, "Vxc work length must be 0 (default) to 128 batches");
    set_vxc_chunk(g, chunk);
  });
}

int chx_grid_set_vxc_shape(chx_grid* g, int shape) {
  return guard([&] {
    require(g && shape >= -1 && shape <= 1, CHX_ERR_ARG, "Vxc shape must be -1 (by basis size), 0 or 1");
    set_vxc_shape(g, shape);
  });
}

int chx_grid_vxc_shape(const chx_grid* g, int* shape) {
  return guard([&] {
    require(g && shape, CHX_ERR_ARG, "bad arguments");
    *shape = g->mx_wide;
  });
}

int chx_grid_autotune_tiles(chx_grid* g, const chx_xc* xc, const double* D_dev, double* V_dev, int reps,
                            int* shape_out, int* chunk_out, float* ms_out) {
  return guard([&] {
    require(g && xc && D_dev && V_dev && reps > 0, CHX_ERR_ARG, "bad arguments");
    require(g->precision == CHX_PREC_MIXED && g->ao_ready, CHX_ERR_STATE,
            "tile shapes are tuned for the mixed form after Grid::getAO");
    int best_shape = -1, best_chunk = 0;
    float best_ms = 0.f;
    for (int shape = 0; shape < 2; ++shape) {
      set_vxc_shape(g, shape);
      int chunk = 0;
      float ms[4];
      const int rc = chx_grid_autotune(g, xc, D_dev, V_dev, reps, &chunk, ms);
      if (rc != CHX_OK) throw Error(rc, g_last_error);
      float m = ms[0];
      for (int k = 0; k < 4; ++k) {
        if (ms_out) ms_out[4 * shape + k] = ms[k];
        m = std::min(m, ms[k]);
      }
      if (best_shape < 0 || m < best_ms) {
        best_shape = shape;
        best_chunk = chunk;
        best_ms = m;
      }
    }
    set_vxc_shape(g, best_shape);
    set_vxc_chunk(g, best_chunk);
    if (shape_out) *shape_out = best_shape;
    if (chunk_out) *chunk_out = best_chunk;
  });
}

int chx_grid_get_density_device(chx_grid* g, const double* dD, int nspin, int nbasis) {
  return guard([&] {
    check_dm_args(g, nspin, nbasis, 2);
    require(dD != nullptr, CHX_ERR_ARG, "null density matrix");
    density_spin(g, dD, nspin);
  });
}

int chx_grid_num_electrons(chx_grid* g, double* N) {
  return guard([&] {
    require(g->have_density, CHX_ERR_STATE, "no density on the grid");
    if (!g->have_eval) {
      chx_xc none;
      none.family = g->type < 0 ? 0 : g->type;
      if (g->nspin == 2) none.family = std::min(none.family, 1);
      none.spin = g->nspin;
      run_functional(g, &none, false, false);
    }
    double en[2];
    CHX_HIP(hipMemcpyAsync(en, g->d_en.p, sizeof(en), hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
    *N = en[1];
  });
}

// Mixed form: E carries the first-order correction for the split of D
// (sum_s V_s . (D_s - hi - lo)), which needs the potential; if Evaluate("ev")
// was not yet followed by a Fock call, the potential is built here (into the
// grid's own buffer, which the next Fock call returns without rebuilding);
// after a Fock call the correction is already in place.
int chx_grid_energy(chx_grid* g, double* E) {
  return guard([&] {
    require(g->have_eval, CHX_ERR_STATE, "ExchangeCorrelation::Evaluate must precede Grid::getEnergy");
    if (g->corr_pending && g->have_v && !g->corr_done) build_vout(g);
    double en[3];
    CHX_HIP(hipMemcpyAsync(en, g->d_en.p, sizeof(en), hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
    g->last_E = en[0] + en[2];
    g->last_N = en[1];
    *E = en[0] + en[2];
  });
}

int chx_grid_fock(chx_grid* g, double* F, int nspin, int nbasis) {
  return guard([&] {
    require(g->have_v, CHX_ERR_STATE, "no potential on the grid (Evaluate with \"v\" or \"ev\" first)");
    require(nbasis == g->nbasis, CHX_ERR_ARG, "Fock matrix size does not match the basis");
    require(nspin == g->v_nspin, CHX_ERR_ARG, "Fock spin count differs from the evaluated density's");
    const size_t n = (size_t)nspin * nbasis * nbasis;
    if (!g->vout_ready) build_vout(g);
    CHX_HIP(hipMemcpyAsync(F, g->d_Vout.p, sizeof(double) * n, hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
  });
}

int chx_grid_fock_device(chx_grid* g, double* dF, int nspin, int nbasis) {
  return guard([&] {
    require(g->have_v, CHX_ERR_STATE, "no potential on the grid (Evaluate with \"v\" or \"ev\" first)");
    require(nbasis == g->nbasis, CHX_ERR_ARG, "Fock matrix size does not match the basis");
    require(nspin == g->v_nspin, CHX_ERR_ARG, "Fock spin count differs from the evaluated density's");
    if (g->vout_ready)
      CHX_HIP(hipMemcpyAsync(dF, g->d_Vout.p, sizeof(double) * (size_t)nspin * nbasis * nbasis,
                             hipMemcpyDeviceToDevice, g->stream));
    else
      fock_spin(g, dF, nspin);
  });
}

int chx_grid_set_vrho(chx_grid* g, const double* vrho, long npoints) {
  return guard([&] {
    require(g->ao_ready, CHX_ERR_STATE, "Grid::getAO must be called first");
    require(npoints == g->npoints && vrho, CHX_ERR_ARG, "v_rho must have one value per grid point");
    std::vector<double> vw((size_t)g->nbatches * kBatch, 0.0);
    long k = 0;
    for (int b = 0; b < g->nbatches; ++b)
      for (int i = 0; i < g->batch_n[b]; ++i, ++k) {
        const size_t s = (size_t)b * kBatch + i;
        vw[s] = vrho[k] * g->hw[s];
      }
    DevBuf<double> d;
    d.upload(vw, g->stream);
    launch_vrho_coef(g->stream, (long)vw.size(), d.p, g->d_coef.p);
    CHX_HIP(hipStreamSynchronize(g->stream));
    g->corr_pending = g->corr_done = g->vout_ready = false;  // a potential of its own: no energy correction
    g->have_v = true;
    g->v_family = 0;
    g->v_nspin = 1;
  });
}

int chx_grid_stream(const chx_grid* g, void** stream) {
  return guard([&] { *stream = (void*)g->stream; });
}

int chx_grid_set_precision(chx_grid* g, int precision) {
  return guard([&] {
    require(precision == CHX_PREC_FP64 || precision == CHX_PREC_MIXED, CHX_ERR_ARG,
            "precision must be CHX_PREC_FP64 or CHX_PREC_MIXED");
    if (g->precision != precision) g->ao_ready = false;  // the AO tables are laid out per form
    g->precision = precision;
  });
}

int chx_grid_set_core_radius(chx_grid* g, double radius) {
  return guard([&] {
    require(g && radius >= 0 && radius <= 1.0, CHX_ERR_ARG, "core radius must be in [0, 1] bohr");
    if (g->core_radius != radius) g->ao_ready = false;  // the plan holds the core batches
    g->core_radius = radius;
  });
}

int chx_grid_core_batches(const chx_grid* g, long* n) {
  return guard([&] {
    require(g && n, CHX_ERR_ARG, "bad arguments");
    *n = g->precision == CHX_PREC_MIXED && g->ao_ready ? g->ncore_batches : 0;
  });
}

int chx_grid_core_flops(const chx_grid* g, double* flops) {
  return guard([&] {
    require(g && flops, CHX_ERR_ARG, "bad arguments");
    *flops = g->precision == CHX_PREC_MIXED && g->ao_ready ? g->core_flops : 0.0;
  });
}

int chx_grid_set_screening(chx_grid* g, double thr) {
  return guard([&] {
    require(thr >= 0, CHX_ERR_ARG, "screening threshold must be >= 0");
    g->screen_thr = thr;
    g->ao_ready = false;
  });
}

int chx_grid_profile(chx_grid* g, int enable) {
  return guard([&] {
    g->collect_times();
    g->profile = enable != 0;
  });
}

int chx_grid_kernel_times(chx_grid* g, int reset, double* ms, long* launches) {
  return guard([&] {
    g->collect_times();
    for (int k = 0; k < 7; ++k) {
      if (ms) ms[k] = g->stage_ms[k];
      if (launches) launches[k] = g->stage_n[k];
      if (reset) { g->stage_ms[k] = 0; g->stage_n[k] = 0; }
    }
  });
}

int chx_grid_set_ao_cache(chx_grid* g, int mode) {
  return guard([&] {
    require(mode >= -1 && mode <= 1, CHX_ERR_ARG, "cache mode must be -1, 0 or 1");
    g->cache_mode = mode;
    g->ao_ready = false;
  });
}

int chx_grid_ao_cached(const chx_grid* g, int* cached) {
  return guard([&] { *cached = g->cached ? 1 : 0; });
}

int chx_grid_stats(const chx_grid* g, long* nbatches, long* npoints, double* mean_tiles, double* flops) {
  return guard([&] {
    if (nbatches) *nbatches = g->nbatches;
    if (npoints) *npoints = g->npoints;
    if (mean_tiles) *mean_tiles = g->mean_tiles;
    if (flops) *flops = g->flops_build;
  });
}

// ---- functional -----------------------------------------------------------
int chx_xc_create(int ncodes, const int* codes, int spin, chx_xc** out) {
  return guard([&] {
    require(ncodes > 0 && codes && out, CHX_ERR_ARG, "bad functional arguments");
    require(spin == 1 || spin == 2, CHX_ERR_ARG, "spin must be 1 (unpolarised) or 2 (polarised)");
    chx_xc* x = new chx_xc;
    x->spin = spin;
    for (int i = 0; i < ncodes; ++i) {
      const int f = code_family(codes[i]);
      if (f < 0) {
        delete x;
        throw Error(CHX_ERR_UNSUPPORTED, "functional code " + std::to_string(codes[i]) + " is not implemented");
      }
      if (spin == 2 && f > 1 && codes[i] != kMggaXR2scan && codes[i] != kMggaCR2scan) {
        delete x;
        throw Error(CHX_ERR_UNSUPPORTED, "of the meta-GGA functionals, r2SCAN is built spin-polarised");
      }
      x->codes.push_back(codes[i]);
      x->family = std::max(x->family, f);
      x->exx = std::max(x->exx, code_exx(codes[i]));
    }
    if (spin == 2 && x->family >= 2)
      for (int c : x->codes)
        if (c != kMggaXR2scan && c != kMggaCR2scan) {
          delete x;
          throw Error(CHX_ERR_UNSUPPORTED, "a spin-polarised meta-GGA functional is built from r2SCAN codes only");
        }
    *out = x;
  });
}

int chx_xc_read(const char* data_root, const char* df_name, int spin, int output, chx_xc** out) {
  std::vector<int> codes;
  int rc = guard([&] {
    require(data_root && df_name && out, CHX_ERR_ARG, "bad arguments");
    std::ifstream f(std::string(data_root) + "/DensityFunctionals/" + df_name + ".df");
    if (!f.good()) throw Error(CHX_ERR_FILE, "Density functional file is missing!");
    std::string line;
    std::getline(f, line);
    std::stringstream ss(line);
    std::string tok;
    while (ss >> tok) {
      try {
        size_t used = 0;
        const int c = std::stoi(tok, &used);
        if (used != tok.size()) throw std::invalid_argument(tok);
        codes.push_back(c);
      } catch (const std::exception&) {
        throw Error(CHX_ERR_FILE, "Corrupted density functional file!");
      }
    }
    if (codes.empty()) throw Error(CHX_ERR_FILE, "Corrupted density functional file!");
  });
  if (rc != CHX_OK) return rc;
  rc = chx_xc_create((int)codes.size(), codes.data(), spin, out);
  if (rc == CHX_OK && output) {
    std::printf("Density functional %s: %zu code(s), family %d, EXX %.2f\n", df_name, codes.size(), (*out)->family,
                (*out)->exx);
  }
  return rc;
}

void chx_xc_destroy(chx_xc* xc) { delete xc; }

int chx_xc_info(const chx_xc* xc, int* family, double* exx, int* ncodes, int* codes) {
  return guard([&] {
    if (family) *family = xc->family;
    if (exx) *exx = xc->exx;
    if (ncodes) *ncodes = (int)xc->codes.size();
    if (codes) std::memcpy(codes, xc->codes.data(), sizeof(int) * xc->codes.size());
  });
}

int chx_xc_evaluate(const chx_xc* xc, chx_grid* g, const char* order) {
  return guard([&] {
    require(xc && g && order, CHX_ERR_ARG, "bad arguments");
    const std::string o(order);
    require(o == "ev" || o == "e" || o == "v" || o == "f", CHX_ERR_UNSUPPORTED,
            "only the orders \"ev\", \"e\", \"v\" and \"f\" are built");
    require(g->have_density, CHX_ERR_STATE, "Grid::getDensity must precede Evaluate");
    require(g->type >= xc->family, CHX_ERR_STATE, "grid type is below the functional family (call setType)");
    if (o == "f") {  // f_xc at the ground-state density, for the response path
      require(xc->family <= 1, CHX_ERR_UNSUPPORTED, "Evaluate(\"f\") is built for LDA and GGA functionals");
      require(g->nspin == xc->spin, CHX_ERR_STATE,
              "the functional's spin count differs from the density's (ExchangeCorrelation::Spin)");
      require(g->precision == CHX_PREC_FP64, CHX_ERR_UNSUPPORTED, "the response path is built for the FP64 form");
      g->d_codes.upload(xc->codes, g->stream);
      const size_t nslots = (size_t)g->nbatches * kBatch;
      if (g->nspin == 2)  // alpha partials in d_part, beta in d_part_b
        launch_functional_fpol(g->stream, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_part_b.p, g->d_codes.p,
                               (int)xc->codes.size(), xc->family,
                               g->d_fx.alloc_get(std::max<size_t>(1, kFxPol * nslots)));
      else
        launch_functional_f(g->stream, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_codes.p, (int)xc->codes.size(),
                            xc->family, g->d_fx.alloc_get(std::max<size_t>(1, kFx * nslots)));
      CHX_HIP(hipStreamSynchronize(g->stream));
      g->fx_family = xc->family;
      g->fx_nspin = g->nspin;
      return;
    }
    const bool e = o.find('e') != std::string::npos, v = o.find('v') != std::string::npos;
    run_functional(g, xc, e, v);
    if (v) {
      g->v_family = xc->family;
      g->v_nspin = g->nspin;
    }
    g->have_eval = e || g->have_eval;
    g->have_v = v;
  });
}

int chx_lebedev_rule(int npoints, double* xyzw) {
  return guard([&] {
    require(xyzw != nullptr, CHX_ERR_ARG, "null output");
    chx_lebedev_rule_impl(npoints, xyzw);
  });
}

int chx_xc_eval_points(const chx_xc* xc, long n, const double* rho, const double* sigma, const double* tau,
                       double* exc, double* vrho, double* vsigma, double* vtau) {
  if (xc && xc->spin == 2)
    return guard([&] {  // libxc XC_POLARIZED layout: rho[2n], sigma[3n], tau[2n], vrho[2n], vsigma[3n], vtau[2n]
      require(n >= 0 && rho && exc && vrho, CHX_ERR_ARG, "bad arguments");
      const bool mgga = xc->family >= 2;
      require(!mgga || (sigma && tau), CHX_ERR_ARG, "a meta-GGA functional needs sigma and tau");
      if (n == 0) return;
      hipStream_t s;
      CHX_HIP(hipStreamCreate(&s));
      DevBuf<double> dr, ds, dt, de, dvr, dvs, dvt;
      DevBuf<int> dc;
      dr.upload(std::vector<double>(rho, rho + 2 * n), s);
      if (sigma) ds.upload(std::vector<double>(sigma, sigma + 3 * n), s);
      if (mgga) dt.upload(std::vector<double>(tau, tau + 2 * n), s);
      dc.upload(xc->codes, s);
      de.alloc(n); dvr.alloc(2 * n); dvs.alloc(3 * n); dvt.alloc(2 * n);
      launch_eval_functional_pol(s, n, dc.p, (int)xc->codes.size(), dr.p, sigma ? ds.p : nullptr, mgga ? dt.p : nullptr,
                                 de.p, dvr.p, dvs.p, dvt.p);
      CHX_HIP(hipMemcpyAsync(exc, de.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      CHX_HIP(hipMemcpyAsync(vrho, dvr.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
      if (vsigma) CHX_HIP(hipMemcpyAsync(vsigma, dvs.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
      if (vtau && mgga) CHX_HIP(hipMemcpyAsync(vtau, dvt.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
      CHX_HIP(hipStreamSynchronize(s));
      CHX_HIP(hipStreamDestroy(s));
    });
  return guard([&] {
    require(xc && n >= 0 && rho && exc && vrho, CHX_ERR_ARG, "bad arguments");
    hipStream_t s;
    CHX_HIP(hipStreamCreate(&s));
    std::vector<double> zero(n, 0.0);
    const std::vector<double> hr(rho, rho + n), hs(sigma ? std::vector<double>(sigma, sigma + n) : zero),
        ht(tau ? std::vector<double>(tau, tau + n) : zero);
    DevBuf<double> dr, ds, dt, de, dvr, dvs, dvt;
    DevBuf<int> dc;
    dr.upload(hr, s); ds.upload(hs, s); dt.upload(ht, s); dc.upload(xc->codes, s);
    de.alloc(std::max<long>(1, n)); dvr.alloc(std::max<long>(1, n)); dvs.alloc(std::max<long>(1, n));
    dvt.alloc(std::max<long>(1, n));
    launch_eval_functional(s, n, dc.p, (int)xc->codes.size(), dr.p, ds.p, dt.p, de.p, dvr.p, dvs.p, dvt.p);
    auto down = [&](double* h, const DevBuf<double>& d) {
      if (h && n) CHX_HIP(hipMemcpyAsync(h, d.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    };
    down(exc, de); down(vrho, dvr); down(vsigma, dvs); down(vtau, dvt);
    CHX_HIP(hipStreamSynchronize(s));
    CHX_HIP(hipStreamDestroy(s));
  });
}

int chx_xc_eval_points_f(const chx_xc* xc, long n, const double* rho, const double* sigma, double* v2rho2,
                         double* v2rhosigma, double* v2sigma2) {
  return guard([&] {
    require(xc && n >= 0 && rho && v2rho2, CHX_ERR_ARG, "bad arguments");
    require(xc->family <= 1, CHX_ERR_UNSUPPORTED, "second derivatives are built for LDA and GGA functionals");
    const bool pol = xc->spin == 2;  // libxc XC_POLARIZED: rho 2, sigma 3; v2rho2 3, v2rhosigma 6, v2sigma2 6 per point
    const long nr = pol ? 2 : 1, ns = pol ? 3 : 1, o0 = pol ? 3 : 1, o1 = pol ? 6 : 1, o2 = pol ? 6 : 1;
    hipStream_t s;
    CHX_HIP(hipStreamCreate(&s));
    const std::vector<double> hr(rho, rho + nr * n),
        hs(sigma ? std::vector<double>(sigma, sigma + ns * n) : std::vector<double>(ns * n, 0.0));
    DevBuf<double> dr, ds, f0, f1, f2;
    DevBuf<int> dc;
    dr.upload(hr, s); ds.upload(hs, s); dc.upload(xc->codes, s);
    f0.alloc(std::max<long>(1, o0 * n)); f1.alloc(std::max<long>(1, o1 * n)); f2.alloc(std::max<long>(1, o2 * n));
    if (pol)
      launch_eval_functional_fpol(s, n, dc.p, (int)xc->codes.size(), dr.p, sigma ? ds.p : nullptr, f0.p, f1.p, f2.p);
    else
      launch_eval_functional_f(s, n, dc.p, (int)xc->codes.size(), dr.p, ds.p, f0.p, f1.p, f2.p);
    auto down = [&](double* h, const DevBuf<double>& d, long k) {
      if (h && n) CHX_HIP(hipMemcpyAsync(h, d.p, sizeof(double) * k * n, hipMemcpyDeviceToHost, s));
    };
    down(v2rho2, f0, o0); down(v2rhosigma, f1, o1); down(v2sigma2, f2, o2);
    CHX_HIP(hipStreamSynchronize(s));
    CHX_HIP(hipStreamDestroy(s));
  });
}

// ---- learned functional surrogate ----------------------------------------
static_assert(CHX_MLP_PARAMS == kMlpParams, "parameter block size");

namespace {
void mlp_require_stream(chx_mlp* m) {
  if (!m->stream) CHX_HIP(hipStreamCreate(&m->stream));
}

void mlp_upload_params(chx_mlp* m, const float* P) {
  CHX_HIP(hipMemcpyAsync(m->P.p, P, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
  launch_mlp_pack(m->stream, m->P.p, m->Prun.p, m->Pinf.p);
  launch_mlp_pack_mx(m->stream, m->P.p, m->Wpk.p);
}

// Glorot-uniform weights (tanh units), zero biases, output bias 1 (F of the
// uniform gas exchange); splitmix64 for the seeded stream.
void mlp_init(unsigned long long seed, std::vector<float>& P) {
  P.assign(kMlpParams, 0.f);
  unsigned long long x = seed;
  auto next = [&] {
    x += 0x9E3779B97F4A7C15ull;
    unsigned long long z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int nout, int nin) {
    const double a = std::sqrt(6.0 / (nin + nout));
    for (int k = 0; k < nout * nin; ++k) P[off + k] = (float)((2.0 * next() - 1.0) * a);
  };
  fill(kMlpW1, kMlpH, 3);
  fill(kMlpW2, kMlpH, kMlpH);
  fill(kMlpW3, kMlpH, kMlpH);
  fill(kMlpW4, 1, kMlpH);
  P[kMlpB4] = 1.f;
}
}  // namespace

int chx_mlp_create(unsigned long long seed, chx_mlp** out) {
  return guard([&] {
    require(out != nullptr, CHX_ERR_ARG, "null output");
    chx_mlp* m = new chx_mlp;
    try {
      mlp_require_stream(m);
      m->P.alloc(kMlpParams);
      m->Prun.alloc(kMlpParams);
      m->Pinf.alloc(kMlpParams);
      m->G.alloc(kMlpParams);
      m->M.alloc(kMlpParams);
      m->V.alloc(kMlpParams);
      m->loss.alloc(1);
      m->Wpk.alloc(kMlpPacked);
      std::vector<float> P;
      mlp_init(seed, P);
      mlp_upload_params(m, P.data());
      CHX_HIP(hipMemsetAsync(m->G.p, 0, sizeof(float) * kMlpParams, m->stream));
      CHX_HIP(hipMemsetAsync(m->M.p, 0, sizeof(float) * kMlpParams, m->stream));
      CHX_HIP(hipMemsetAsync(m->V.p, 0, sizeof(float) * kMlpParams, m->stream));
      CHX_HIP(hipStreamSynchronize(m->stream));
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

void chx_mlp_destroy(chx_mlp* m) { delete m; }

int chx_mlp_get_params(const chx_mlp* m, float* P) {
  return guard([&] {
    require(m && P, CHX_ERR_ARG, "bad arguments");
    CHX_HIP(hipMemcpyAsync(P, m->P.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
    CHX_HIP(hipStreamSynchronize(m->stream));
  });
}

int chx_mlp_set_params(chx_mlp* m, const float* P) {
  return guard([&] {
    require(m && P, CHX_ERR_ARG, "bad arguments");
    mlp_upload_params(m, P);
    CHX_HIP(hipMemsetAsync(m->M.p, 0, sizeof(float) * kMlpParams, m->stream));  // fresh optimizer state
    CHX_HIP(hipMemsetAsync(m->V.p, 0, sizeof(float) * kMlpParams, m->stream));
    CHX_HIP(hipStreamSynchronize(m->stream));
    m->step = 0;
  });
}

int chx_mlp_set_box(chx_mlp* m, const float* lo, const float* hi, double rho_min) {
  return guard([&] {
    require(m && lo && hi, CHX_ERR_ARG, "bad arguments");
    for (int k = 0; k < 3; ++k) {
      require(std::isfinite(lo[k]) && std::isfinite(hi[k]) && hi[k] > lo[k], CHX_ERR_ARG, "empty descriptor box");
      m->box.lo[k] = lo[k];
      m->box.hi[k] = hi[k];
    }
    require(rho_min > 0, CHX_ERR_ARG, "rho_min must be positive");
    m->box.rho_min = rho_min;
    m->have_box = true;
  });
}

int chx_mlp_set_base(chx_mlp* m, int ncodes, const int* codes) {
  return guard([&] {
    require(m && ncodes >= 0 && ncodes <= 2 && (ncodes == 0 || codes), CHX_ERR_ARG,
            "the analytic base takes 0 to 2 functional codes");
    for (int k = 0; k < ncodes; ++k)
      require(code_family(codes[k]) == 0 || code_family(codes[k]) == 1, CHX_ERR_UNSUPPORTED,
              "the analytic base is an LDA or GGA functional");
    m->box.nbase = ncodes;
    for (int k = 0; k < 2; ++k) m->box.base[k] = k < ncodes ? codes[k] : 0;
  });
}

int chx_mlp_get_base(const chx_mlp* m, int* ncodes, int* codes) {
  return guard([&] {
    require(m && ncodes, CHX_ERR_ARG, "bad arguments");
    *ncodes = m->box.nbase;
    if (codes)
      for (int k = 0; k < m->box.nbase; ++k) codes[k] = m->box.base[k];
  });
}

int chx_mlp_get_box(const chx_mlp* m, float* lo, float* hi, double* rho_min) {
  return guard([&] {
    require(m && m->have_box, CHX_ERR_STATE, "the network has no descriptor box");
    for (int k = 0; k < 3; ++k) {
      if (lo) lo[k] = m->box.lo[k];
      if (hi) hi[k] = m->box.hi[k];
    }
    if (rho_min) *rho_min = m->box.rho_min;
  });
}

int chx_mlp_features(const chx_xc* xc, long n, const double* rho, const double* sigma, const double* tau, float* t,
                     float* F) {
  return guard([&] {
    require(xc && n >= 0 && rho && sigma && tau && t && F, CHX_ERR_ARG, "bad arguments");
    if (n == 0) return;
    hipStream_t s;
    CHX_HIP(hipStreamCreate(&s));
    DevBuf<double> dr, ds, dt;
    DevBuf<int> dc;
    DevBuf<float> dtt, dF;
    dr.upload(std::vector<double>(rho, rho + n), s);
    ds.upload(std::vector<double>(sigma, sigma + n), s);
    dt.upload(std::vector<double>(tau, tau + n), s);
    dc.upload(xc->codes, s);
    dtt.alloc(3 * (size_t)n);
    dF.alloc(n);
    launch_mlp_features(s, n, dr.p, ds.p, dt.p, dc.p, (int)xc->codes.size(), dtt.p, dF.p);
    CHX_HIP(hipMemcpyAsync(t, dtt.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    CHX_HIP(hipMemcpyAsync(F, dF.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    CHX_HIP(hipStreamDestroy(s));
  });
}

int chx_mlp_set_data(chx_mlp* m, long n, const float* U, const float* T, const float* C) {
  return guard([&] {
    require(m && n > 0 && U && T, CHX_ERR_ARG, "bad arguments");
    m->U.alloc(3 * (size_t)n);
    m->T.alloc(n);
    CHX_HIP(hipMemcpyAsync(m->U.p, U, sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->stream));
    CHX_HIP(hipMemcpyAsync(m->T.p, T, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    m->cprefix.assign(n + 1, 0.0);
    if (C) {
      m->C.alloc(n);
      CHX_HIP(hipMemcpyAsync(m->C.p, C, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
      for (long i = 0; i < n; ++i) {
        require(C[i] >= 0.f && std::isfinite(C[i]), CHX_ERR_ARG, "sample weights must be finite and >= 0");
        m->cprefix[i + 1] = m->cprefix[i] + C[i];
      }
    } else {
      m->C.release();
      for (long i = 0; i < n; ++i) m->cprefix[i + 1] = (double)(i + 1);
    }
    CHX_HIP(hipStreamSynchronize(m->stream));
    m->ndata = n;
  });
}

int chx_mlp_train(chx_mlp* m, int epochs, long batch, double lr, double lr_final, double* epoch_loss) {
  return guard([&] {
    require(m && epochs >= 0 && batch > 0 && lr > 0 && lr_final >= 0, CHX_ERR_ARG, "bad arguments");
    require(m->ndata > 0, CHX_ERR_STATE, "no training set (chx_mlp_set_data)");
    const long n = m->ndata;
    const long nb = (n + batch - 1) / batch;
    const long total = (long)epochs * nb;
    long it = 0;
    for (int ep = 0; ep < epochs; ++ep) {
      launch_fill_u32(m->stream, (unsigned*)m->loss.p, 1, 0u);  // 0.0f
      for (long b = 0; b < nb; ++b, ++it) {
        const long s0 = b * batch, s1 = std::min(n, s0 + batch);
        const double csum = m->cprefix[s1] - m->cprefix[s0];
        if (!(csum > 0)) continue;
        const double cosw = 0.5 * (1.0 + std::cos(M_PI * (double)it / (double)std::max<long>(1, total)));
        const double lr_t = lr_final + (lr - lr_final) * cosw;
        launch_mlp_grad(m->stream, s1 - s0, m->U.p + 3 * s0, m->T.p + s0, m->C.p ? m->C.p + s0 : nullptr,
                        (float)(1.0 / csum), m->Prun.p, m->G.p, m->loss.p, 512);
        m->step += 1;
        launch_mlp_adam(m->stream, m->P.p, m->G.p, m->M.p, m->V.p, m->Prun.p, m->Pinf.p, (float)lr_t, m->step);
      }
      float l = 0.f;
      CHX_HIP(hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
      CHX_HIP(hipStreamSynchronize(m->stream));
      if (epoch_loss) epoch_loss[ep] = (double)l / (double)nb;
    }
    launch_mlp_pack_mx(m->stream, m->P.p, m->Wpk.p);
    CHX_HIP(hipStreamSynchronize(m->stream));
  });
}

int chx_mlp_loss_grad(chx_mlp* m, long n, const float* U, const float* T, const float* C, float* G, double* loss) {
  return guard([&] {
    require(m && n > 0 && U && T && G, CHX_ERR_ARG, "bad arguments");
    DevBuf<float> dU, dT, dC;
    dU.alloc(3 * (size_t)n);
    dT.alloc(n);
    CHX_HIP(hipMemcpyAsync(dU.p, U, sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->stream));
    CHX_HIP(hipMemcpyAsync(dT.p, T, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    double csum = (double)n;
    if (C) {
      dC.alloc(n);
      CHX_HIP(hipMemcpyAsync(dC.p, C, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
      csum = 0;
      for (long i = 0; i < n; ++i) csum += C[i];
    }
    require(csum > 0, CHX_ERR_ARG, "sample weights sum to zero");
    DevBuf<float> dG;
    dG.alloc(kMlpParams);
    launch_fill_u32(m->stream, (unsigned*)dG.p, kMlpParams, 0u);
    launch_fill_u32(m->stream, (unsigned*)m->loss.p, 1, 0u);
    launch_mlp_grad(m->stream, n, dU.p, dT.p, C ? dC.p : nullptr, (float)(1.0 / csum), m->Prun.p, dG.p, m->loss.p,
                    512);
    float l = 0.f;
    CHX_HIP(hipMemcpyAsync(G, dG.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
    CHX_HIP(hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    CHX_HIP(hipStreamSynchronize(m->stream));
    if (loss) *loss = l;
  });
}

int chx_mlp_forward(const chx_mlp* m, long n, const float* U, int bf16, float* F, float* dF) {
  return guard([&] {
    require(m && n >= 0 && U && F, CHX_ERR_ARG, "bad arguments");
    if (n == 0) return;
    DevBuf<float> dU, dFv, dG;
    dU.alloc(3 * (size_t)n);
    dFv.alloc(n);
    if (dF) dG.alloc(3 * (size_t)n);
    CHX_HIP(hipMemcpyAsync(dU.p, U, sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->stream));
    // 0: FP32 weights, per-lane network; 1: BF16-rounded weights, per-lane network; 2: the inference
    // path (FP32 weights as split-BF16 MFMA operands)
    launch_mlp_forward(m->stream, n, dU.p, bf16 == 1 ? m->Pinf.p : m->Prun.p, bf16 == 2 ? m->Wpk.p : nullptr, dFv.p,
                       dF ? dG.p : nullptr);
    CHX_HIP(hipMemcpyAsync(F, dFv.p, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    if (dF) CHX_HIP(hipMemcpyAsync(dF, dG.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, m->stream));
    CHX_HIP(hipStreamSynchronize(m->stream));
  });
}

int chx_mlp_eval_points(const chx_mlp* m, const chx_xc* xc, long n, const double* rho, const double* sigma,
                        const double* tau, double* exc, double* vrho, double* vsigma, double* vtau, long* nfallback) {
  return guard([&] {
    require(m && xc && n >= 0 && rho && sigma && tau && exc && vrho && vsigma && vtau, CHX_ERR_ARG, "bad arguments");
    require(m->have_box, CHX_ERR_STATE, "the network has no descriptor box");
    if (n == 0) {
      if (nfallback) *nfallback = 0;
      return;
    }
    hipStream_t s = m->stream;
    DevBuf<double> dr, ds, dt, de, dvr, dvs, dvt;
    DevBuf<int> dc, dn;
    dr.upload(std::vector<double>(rho, rho + n), s);
    ds.upload(std::vector<double>(sigma, sigma + n), s);
    dt.upload(std::vector<double>(tau, tau + n), s);
    dc.upload(xc->codes, s);
    de.alloc(n); dvr.alloc(n); dvs.alloc(n); dvt.alloc(n);
    dn.alloc(1);
    launch_fill_u32(s, (unsigned*)dn.p, 1, 0u);
    launch_mlp_eval(s, n, m->Prun.p, m->Wpk.p, m->box, dc.p, (int)xc->codes.size(), dr.p, ds.p, dt.p, de.p, dvr.p, dvs.p, dvt.p,
                    dn.p);
    int nf = 0;
    CHX_HIP(hipMemcpyAsync(exc, de.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    CHX_HIP(hipMemcpyAsync(vrho, dvr.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    CHX_HIP(hipMemcpyAsync(vsigma, dvs.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    CHX_HIP(hipMemcpyAsync(vtau, dvt.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    CHX_HIP(hipMemcpyAsync(&nf, dn.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    if (nfallback) *nfallback = nf;
  });
}

int chx_xc_set_surrogate(chx_xc* xc, const chx_mlp* m) {
  return guard([&] {
    require(xc != nullptr, CHX_ERR_ARG, "null functional");
    if (m) {
      require(xc->family == 2, CHX_ERR_UNSUPPORTED, "the surrogate stands in for a meta-GGA functional");
      require(m->have_box, CHX_ERR_STATE, "the network has no descriptor box");
    }
    xc->surrogate = m;
  });
}

int chx_grid_point_density(chx_grid* g, double* rho, double* sigma, double* tau, double* w, long* n) {
  return guard([&] {
    require(g != nullptr, CHX_ERR_ARG, "null grid");
    require(g->have_density, CHX_ERR_STATE, "Grid::getDensity must precede the point density");
    const size_t ns = (size_t)g->nbatches * kBatch;
    if (n) *n = (long)ns;
    if (!rho && !sigma && !tau && !w) return;
    DevBuf<double> buf;
    buf.alloc(std::max<size_t>(1, 4 * ns));
    launch_point_density(g->stream, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_part_tau.p, g->type >= 2 ? 1 : 0,
                         g->d_w.p, buf.p, buf.p + ns, buf.p + 2 * ns, buf.p + 3 * ns);
    double* outs[4] = {rho, sigma, tau, w};
    for (int k = 0; k < 4; ++k)
      if (outs[k] && ns) CHX_HIP(hipMemcpyAsync(outs[k], buf.p + k * ns, sizeof(double) * ns, hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
  });
}

int chx_grid_fallback_count(const chx_grid* g, long* n) {
  return guard([&] {
    require(g && n, CHX_ERR_ARG, "bad arguments");
    require(g->surrogate_used, CHX_ERR_STATE, "the last Evaluate did not run the surrogate");
    int nf = 0;
    CHX_HIP(hipMemcpyAsync(&nf, g->d_nfall.p, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
    *n = nf;
  });
}

int chx_selftest(double* max_err) {
  return guard([&] {
    DevBuf<double> d;
    d.alloc(1);
    hipStream_t s;
    CHX_HIP(hipStreamCreate(&s));
    launch_selftest(s, d.p);
    DevBuf<unsigned> bad;
    bad.alloc(1);
    launch_fill_u32(s, bad.p, 1, 0u);
    launch_bf16_check(s, 1u << 20, bad.p);
    unsigned nbad = 0;
    CHX_HIP(hipMemcpyAsync(max_err, d.p, sizeof(double), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    CHX_HIP(hipStreamDestroy(s));
    *max_err += (double)nbad;
  });
}

}  // extern "C"
