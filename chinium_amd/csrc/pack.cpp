// Many molecules per XC build (mixed-precision form).
//
// A pack holds grids that were each built and planned on their own (points,
// Becke weights, screening, split operand cache, Vxc work items). One XC build
// of all of them runs the single-grid kernels of every molecule (descriptor
// passed by value, so each kernel addresses its tables as global memory) on a
// few lanes: the molecules are cut into contiguous groups of about equal work,
// each group on a stream of its own, and a group's density launches wait for
// the previous group's. One group's functional, B^T and Vxc stages then run
// beside the next group's density, which keeps the GPU filled between the
// small per-molecule launches. The whole build is captured once as a hipGraph
// (streams forked from and joined back to the pack's stream) and replayed each
// SCF iteration; a capture is redone when the D/V pointers, the functional,
// the lanes or the Vxc work length change. LDA/GGA packs can instead run each
// stage as one persistent launch over every molecule's items (the molecules'
// descriptors in a table of offsets from a base pointer, so the kernels keep
// their global-memory instructions). chx_pack_autotune picks the Vxc work-item
// length, the Vxc block shape by basis size, and lanes or persistent stages by
// timing.
//
// No reference counterpart: Chinium runs one molecule per process
// (src/HartreeFockKohnSham/SelfConsistentField.cpp:27-33).
#include <array>
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace chx {
void set_vxc_chunk(chx_grid* g, int chunk);
void set_vxc_shape(chx_grid* g, int shape);
std::string& last_error();
}
using namespace chx;

struct chx_pack {
  std::vector<chx_grid*> grids;  // not owned
  hipStream_t stream = nullptr;
  int type = 0;
  DevBuf<int> d_codes;
  DevBuf<double> d_en;  // E, N per molecule
  std::vector<int> codes;
  int family = -1;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const double*> gD;  // pointers the graph was captured with
  std::vector<double*> gV;
  std::vector<unsigned long> gen;  // each grid's plan generation when the lanes and graph were built
  std::vector<double> h_en;
  int lanes = 16;
  int wide_from = -2;    // Vxc shape of the last autotune: 256-AO groups for npad >= wide_from (-1: none; -2: not tuned)
  std::vector<int> cut;  // lane k holds molecules [cut[k], cut[k + 1])
  std::vector<hipStream_t> lane_st;
  std::vector<hipEvent_t> ev_dens, ev_done;
  hipEvent_t ev_fork = nullptr;
  // per lane: the multi-molecule functional launch's table (molecules of the
  // lane, then a sentinel) and base pointer, and its total block count
  DevBuf<FnMolecule> fn_tab;  // every lane's table, back to back from fn_off[lane]
  std::vector<int> fn_off;
  std::vector<const char*> fn_base;
  std::vector<int> fn_blocks;
  // persistent stages (LDA/GGA packs): one launch per stage over every
  // molecule's items; descriptors with byte offsets from mol_base, per stage the
  // first item of each molecule (kStages x (n + 1)), and the functional's table
  // over all molecules
  // form: 0 per-molecule kernels on the lanes, 1 persistent stages over every
  // molecule, 2 one launch per stage per lane (each over the lane's molecules)
  int form = 0;
  DevBuf<MolDev> mol_tab;
  const char* mol_base = nullptr;
  DevBuf<int> d_starts, d_ident;
  std::vector<int> totals;
  // form 2: per lane, the stages' first items of its molecules (5 x (molecules + 1), from lane_soff)
  DevBuf<int> d_lane_starts;
  std::vector<int> lane_soff;
  std::vector<std::array<int, 5>> lane_tot;
  DevBuf<FnMolecule> fn_all;
  const char* fn_all_base = nullptr;
  int fn_all_blocks = 0;
  ~chx_pack() {
    drop_graph();
    free_lanes();
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void free_lanes() {
    for (hipStream_t t : lane_st) (void)hipStreamDestroy(t);
    for (hipEvent_t e : ev_dens) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_done) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    lane_st.clear();
    ev_dens.clear();
    ev_done.clear();
    ev_fork = nullptr;
  }
};

namespace {
template <class F> int pguard(F&& f) {
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

void prequire(bool ok, int code, const std::string& msg) {
  if (!ok) throw Error(code, msg);
}

// The functional launch's table for molecules [m0, m1): their buffers as byte
// offsets from the lowest address among them, then a sentinel; its block count.
void fn_table(const chx_pack* p, int m0, int m1, std::vector<FnMolecule>& tab, const char*& base, int& blocks) {
  auto ptrs = [&](const chx_grid* g) {
    return std::array<const char*, 8>{(const char*)g->d_grp_off.p, (const char*)g->d_part.p,
                                      (const char*)g->d_part_tau.p, (const char*)g->d_w.p,
                                      (const char*)g->d_rho.p, (const char*)g->d_coef.p,
                                      (const char*)g->d_en_part.p, (const char*)g->d_cpart.p};
  };
  base = nullptr;
  for (int m = m0; m < m1; ++m)
    for (const char* q : ptrs(p->grids[m]))
      if (!base || q < base) base = q;
  tab.clear();
  int blk = 0;
  for (int m = m0; m < m1; ++m) {
    const chx_grid* g = p->grids[m];
    const auto q = ptrs(g);
    FnMolecule f{};
    f.grp_off = q[0] - base; f.part = q[1] - base; f.part_tau = q[2] - base; f.w = q[3] - base;
    f.rho = q[4] - base; f.coef = q[5] - base; f.en_part = q[6] - base; f.cpart = q[7] - base;
    f.ncpart = corr_blocks(g->nbasis);
    f.nslots = g->nbatches * kBatch;
    f.blk0 = blk;
    f.slot_en = m;
    blk += g->en_blocks;
    tab.push_back(f);
  }
  FnMolecule end{};
  end.blk0 = blk;
  tab.push_back(end);
  blocks = blk;
}

// Tables of the persistent stages (LDA/GGA packs): each molecule's descriptor
// with its pointers as byte offsets from the lowest of them (a molecule without
// core batches gets an identity list for the (rho, grad rho) launch's groups),
// the items of each stage per molecule, and the functional's table over all.
void build_persistent(chx_pack* p) {
  const int n = (int)p->grids.size();
  int maxng = 1;
  for (const chx_grid* g : p->grids) maxng = std::max(maxng, (int)g->grp_batch.size());
  std::vector<int> ident(maxng);
  for (int k = 0; k < maxng; ++k) ident[k] = k;
  p->d_ident.upload(ident, p->stream);
  std::vector<MolDev> mols(n);
  for (int m = 0; m < n; ++m) {
    mols[m] = p->grids[m]->hmol;
    if (!mols[m].main_grps) mols[m].main_grps = p->d_ident.p;
  }
  // every pointer field of a descriptor (the fields rebase() offsets), read and
  // written as addresses through memcpy
  auto fields = [](MolDev& d) {
    return std::array<void*, 38>{
        &d.grp_batch, &d.grp_tiles, &d.grp_ntiles, &d.bt_off, &d.bt_list, &d.bx, &d.by, &d.bz,
        &d.bv.sh_xyz, &d.bv.sh_type, &d.bv.sh_nprim, &d.bv.sh_poff, &d.bv.p_exp, &d.bv.p_coef,
        &d.bv.tile_seg_off, &d.bv.seg, &d.bv.segrec, &d.bv.sh_atom, &d.cache, &d.Ds, &d.part, &d.part_tau,
        &d.slot_batch, &d.slot_tile, &d.coef, &d.bt, &d.cwork, &d.pair_batches, &d.slot_of, &d.cslot, &d.Mint,
        &d.vchunk, &d.vtab, &d.vtile, &d.cache3, &d.core_off, &d.core_grps, &d.main_grps};
  };
  auto get = [](void* f) {
    uintptr_t v;
    std::memcpy(&v, f, sizeof v);
    return v;
  };
  uintptr_t base = 0;
  for (MolDev& d : mols)
    for (void* f : fields(d)) {
      const uintptr_t v = get(f);
      if (v && (!base || v < base)) base = v;
    }
  for (MolDev& d : mols)  // null fields (never read by that molecule's items) map to the base
    for (void* f : fields(d)) {
      const uintptr_t v = get(f), off = v ? v - base : 0;
      std::memcpy(f, &off, sizeof off);
    }
  p->mol_base = (const char*)base;
  p->mol_tab.upload(mols, p->stream);
  // per stage: density (the groups outside the core batches), core groups, B^T slots (+ the zero
  // slot), Vxc items of 128-AO molecules, Vxc items of 256-AO molecules
  std::vector<int> starts(5 * (size_t)(n + 1), 0);
  p->totals.assign(5, 0);
  for (int m = 0; m < n; ++m) {
    const chx_grid* g = p->grids[m];
    const int ng = (int)g->grp_batch.size(), ncg = (int)g->core_grps.size(), nw = (int)g->cwork.size();
    const int cnt[5] = {ng - ncg, ncg, (int)g->bt_list.size() + 1, g->mx_wide ? 0 : nw, g->mx_wide ? nw : 0};
    for (int st = 0; st < 5; ++st) starts[(size_t)st * (n + 1) + m + 1] = starts[(size_t)st * (n + 1) + m] + cnt[st];
  }
  for (int st = 0; st < 5; ++st) p->totals[st] = starts[(size_t)st * (n + 1) + n];
  p->d_starts.upload(starts, p->stream);
  // the same per lane (form 2)
  const int L = (int)p->cut.size() - 1;
  std::vector<int> ls;
  p->lane_soff.assign(std::max(L, 0), 0);
  p->lane_tot.assign(std::max(L, 0), std::array<int, 5>{0, 0, 0, 0, 0});
  for (int k = 0; k < L; ++k) {
    const int m0 = p->cut[k], m1 = p->cut[k + 1], nl = m1 - m0;
    p->lane_soff[k] = (int)ls.size();
    for (int st = 0; st < 5; ++st) {
      const int* a = &starts[(size_t)st * (n + 1)];
      for (int i = 0; i <= nl; ++i) ls.push_back(a[m0 + i] - a[m0]);
      p->lane_tot[k][st] = a[m1] - a[m0];
    }
  }
  p->d_lane_starts.upload(ls.empty() ? std::vector<int>{0} : ls, p->stream);
  std::vector<FnMolecule> tab;
  fn_table(p, 0, n, tab, p->fn_all_base, p->fn_all_blocks);
  p->fn_all.upload(tab, p->stream);
  CHX_HIP(hipStreamSynchronize(p->stream));  // the host vectors go out of scope
}

// Lane cut points (equal shares of density groups + Vxc work items, at least
// one molecule per lane) and the lanes' streams and events.
void build_lanes(chx_pack* p) {
  const int n = (int)p->grids.size();
  std::vector<double> cum(n + 1, 0.0);
  for (int m = 0; m < n; ++m)
    cum[m + 1] = cum[m] + (double)p->grids[m]->grp_batch.size() + (double)p->grids[m]->cwork.size();
  const int L = std::max(1, std::min(p->lanes, n));
  p->cut.assign(L + 1, n);
  int c = 0;
  for (int k = 0; k < L; ++k) {
    p->cut[k] = c;
    if (k == L - 1) break;
    const double goal = cum[n] * (k + 1) / L;
    const int m0 = c;
    while (c < n - (L - 1 - k) && (c == m0 || cum[c + 1] <= goal)) ++c;
  }
  p->drop_graph();
  p->free_lanes();
  p->gen.resize(n);
  for (int m = 0; m < n; ++m) p->gen[m] = p->grids[m]->plan_gen;
  // the functional of each lane's molecules in one launch
  std::vector<FnMolecule> all;
  p->fn_off.assign(L, 0);
  p->fn_base.assign(L, nullptr);
  p->fn_blocks.assign(L, 0);
  for (int k = 0; k < L; ++k) {
    std::vector<FnMolecule> tab;
    fn_table(p, p->cut[k], p->cut[k + 1], tab, p->fn_base[k], p->fn_blocks[k]);
    p->fn_off[k] = (int)all.size();
    all.insert(all.end(), tab.begin(), tab.end());
  }
  p->fn_tab.upload(all, p->stream);
  CHX_HIP(hipStreamSynchronize(p->stream));  // the table's host vector goes out of scope
  if (p->type <= 1) build_persistent(p);  // after the cut: form 2's tables are per lane
  if (L > 1) {
    p->lane_st.resize(L);
    p->ev_dens.resize(L);
    p->ev_done.resize(L);
    for (int k = 0; k < L; ++k) {
      CHX_HIP(hipStreamCreate(&p->lane_st[k]));
      CHX_HIP(hipEventCreateWithFlags(&p->ev_dens[k], hipEventDisableTiming));
      CHX_HIP(hipEventCreateWithFlags(&p->ev_done[k], hipEventDisableTiming));
    }
    CHX_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
  }
}

// The molecules of a lane on stream s: split D and clear M first; the density
// launches wait for `after` (if set) and signal `dens_done` (if set); then the
// functional and E/N sums of all the lane's molecules in one launch each, per
// molecule B^T, Vxc and V = M + M^T (with the partial sums of V . (D - hi -
// lo)), and last the energies' correction for the split of D, one launch for
// the lane. Apart from the functional and the correction, the kernels and their
// arguments are those of the single-grid build.
bool stages_ok(const chx_pack* p) { return p->type <= 1 && p->family <= 1; }

// Form 2: the lane's molecules in one launch per stage (the persistent stage
// kernels over the lane's part of the descriptor table).
void enqueue_lane_stages(chx_pack* p, int lane, hipStream_t s, hipEvent_t after, hipEvent_t dens_done,
                         const double* const* D, double* const* V) {
  const int m0 = p->cut[lane], m1 = p->cut[lane + 1], nl = m1 - m0;
  for (int m = m0; m < m1; ++m) {
    chx_grid* g = p->grids[m];
    launch_split_d(s, g->nbasis, g->npad, D[m], g->d_int2ref.p, g->d_Dsplit.p, g->d_Dres.p);
    launch_fill(s, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
  }
  if (after) CHX_HIP(hipStreamWaitEvent(s, after, 0));
  const int* st = p->d_lane_starts.p + p->lane_soff[lane];
  auto stage = [&](int k) {
    launch_stage_multi(s, k, 1, p->mol_base, p->mol_tab.p + m0, st + (size_t)k * (nl + 1), nl, p->lane_tot[lane][k]);
  };
  stage(kStageDensity);
  stage(kStageCore);
  if (dens_done) CHX_HIP(hipEventRecord(dens_done, s));
  launch_functional_multi(s, p->fn_base[lane], p->fn_tab.p + p->fn_off[lane], nl, p->fn_blocks[lane], p->d_codes.p,
                          (int)p->codes.size(), p->family, p->d_en.p);
  stage(kStageMakeB);
  stage(kStageVxc);
  stage(kStageVxcWide);
  for (int m = m0; m < m1; ++m) {
    chx_grid* g = p->grids[m];
    launch_finalize_corr(s, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, V[m], g->d_Dres.p, g->d_cpart.p);
  }
  launch_corr_multi(s, p->fn_base[lane], p->fn_tab.p + p->fn_off[lane], nl, p->d_en.p);
}

void enqueue_lane(chx_pack* p, int lane, hipStream_t s, hipEvent_t after, hipEvent_t dens_done,
                  const double* const* D, double* const* V) {
  if (p->form == 2 && stages_ok(p)) {
    enqueue_lane_stages(p, lane, s, after, dens_done, D, V);
    return;
  }
  const int m0 = p->cut[lane], m1 = p->cut[lane + 1];
  for (int m = m0; m < m1; ++m) {
    chx_grid* g = p->grids[m];
    launch_split_d(s, g->nbasis, g->npad, D[m], g->d_int2ref.p, g->d_Dsplit.p, g->d_Dres.p);
    launch_fill(s, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
  }
  if (after) CHX_HIP(hipStreamWaitEvent(s, after, 0));
  for (int m = m0; m < m1; ++m) {
    chx_grid* g = p->grids[m];
    const int ng = (int)g->grp_batch.size(), ncg = (int)g->core_grps.size();
    launch_density_mx(s, p->type >= 2 ? 5 : 0, p->type >= 2 ? ng : ng - ncg, g->hmol);
    // core batches' rho and grad rho again with three-term AO operands
    launch_density_mx_core(s, 0, (int)g->core_grps.size(), g->hmol);
  }
  if (dens_done) CHX_HIP(hipEventRecord(dens_done, s));
  const int nc = p->family >= 2 ? 4 : 1;
  // the functional and the E/N sums of all the lane's molecules: one launch each
  launch_functional_multi(s, p->fn_base[lane], p->fn_tab.p + p->fn_off[lane], m1 - m0, p->fn_blocks[lane], p->d_codes.p,
                          (int)p->codes.size(), p->family, p->d_en.p);
  for (int m = m0; m < m1; ++m) {
    chx_grid* g = p->grids[m];
    launch_make_b_mx(s, nc, g->hmol);
    if (g->mx_wide)
      launch_vxc_mx_wide(s, nc, (int)g->cwork.size(), g->hmol);
    else
      launch_vxc_mx(s, nc, (int)g->cwork.size(), g->hmol);
    launch_finalize_corr(s, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, V[m], g->d_Dres.p, g->d_cpart.p);
  }
  launch_corr_multi(s, p->fn_base[lane], p->fn_tab.p + p->fn_off[lane], m1 - m0, p->d_en.p);
}

// One XC build of every molecule as persistent stages on the pack's stream
// (LDA/GGA): per molecule split D and clear M; then one launch per stage over
// every molecule's items (density, core batches, functional, B^T, Vxc of each
// block shape); per molecule V = M + M^T with the partial sums of
// V . (D - hi - lo), and one launch for the energies' corrections.
void enqueue_persistent(chx_pack* p, const double* const* D, double* const* V) {
  const int n = (int)p->grids.size();
  hipStream_t s = p->stream;
  for (int m = 0; m < n; ++m) {
    chx_grid* g = p->grids[m];
    launch_split_d(s, g->nbasis, g->npad, D[m], g->d_int2ref.p, g->d_Dsplit.p, g->d_Dres.p);
    launch_fill(s, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
  }
  const int* st = p->d_starts.p;
  auto stage = [&](int k) {
    launch_stage_multi(s, k, 1, p->mol_base, p->mol_tab.p, st + (size_t)k * (n + 1), n, p->totals[k]);
  };
  stage(kStageDensity);
  stage(kStageCore);
  launch_functional_multi(s, p->fn_all_base, p->fn_all.p, n, p->fn_all_blocks, p->d_codes.p, (int)p->codes.size(),
                          p->family, p->d_en.p);
  stage(kStageMakeB);
  stage(kStageVxc);
  stage(kStageVxcWide);
  for (int m = 0; m < n; ++m) {
    chx_grid* g = p->grids[m];
    launch_finalize_corr(s, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, V[m], g->d_Dres.p, g->d_cpart.p);
  }
  launch_corr_multi(s, p->fn_all_base, p->fn_all.p, n, p->d_en.p);
}


// One XC build of every molecule: persistent stages, or lanes (one lane runs on
// the pack's stream; more lanes fork from it, are staggered by their density
// launches and join back).
void enqueue(chx_pack* p, const double* const* D, double* const* V) {
  if (p->form == 1 && stages_ok(p)) {
    enqueue_persistent(p, D, V);
    return;
  }
  const int L = (int)p->cut.size() - 1;
  if (L == 1) {
    enqueue_lane(p, 0, p->stream, nullptr, nullptr, D, V);
    return;
  }
  CHX_HIP(hipEventRecord(p->ev_fork, p->stream));
  for (int k = 0; k < L; ++k) {
    CHX_HIP(hipStreamWaitEvent(p->lane_st[k], p->ev_fork, 0));
    enqueue_lane(p, k, p->lane_st[k], k ? p->ev_dens[k - 1] : nullptr, p->ev_dens[k], D, V);
    CHX_HIP(hipEventRecord(p->ev_done[k], p->lane_st[k]));
  }
  for (int k = 0; k < L; ++k) CHX_HIP(hipStreamWaitEvent(p->stream, p->ev_done[k], 0));
}

void run(chx_pack* p, const chx_xc* xc, const double* const* D, double* const* V, double* E, double* N) {
  const int n = (int)p->grids.size();
  prequire(xc->family <= p->type, CHX_ERR_STATE, "grid type is below the functional family (call setType)");
  prequire(!xc->surrogate, CHX_ERR_UNSUPPORTED, "packs run the analytic functional; detach the surrogate");
  prequire(xc->spin == 1, CHX_ERR_UNSUPPORTED, "packs run the closed-shell (nspin = 1) path");
  if (xc->codes != p->codes || xc->family != p->family) {
    p->codes = xc->codes;
    p->family = xc->family;
    p->d_codes.upload(p->codes, p->stream);
    CHX_HIP(hipStreamSynchronize(p->stream));
    p->drop_graph();
  }
  // A member grid that was re-planned, re-tuned or had a buffer reallocated on
  // its own since the lanes were cut: its work counts and device pointers in
  // any captured graph are stale, so cut the lanes again (which drops the graph).
  for (int m = 0; m < n; ++m)
    if (p->grids[m]->plan_gen != p->gen[m]) {
      prequire(p->grids[m]->ao_ready && p->grids[m]->precision == CHX_PREC_MIXED &&
                   p->grids[m]->type == p->type,
               CHX_ERR_STATE, "a packed grid was changed to another form or type; create a new pack");
      build_lanes(p);
      break;
    }
  const std::vector<const double*> vd(D, D + n);
  const std::vector<double*> vv(V, V + n);
  if (p->use_graph) {
    if (!p->exec || vd != p->gD || vv != p->gV) {
      p->drop_graph();
      CHX_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(p, D, V);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(p->stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      CHX_HIP(hipStreamEndCapture(p->stream, &p->graph));
      CHX_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      p->gD = vd;
      p->gV = vv;
    }
    CHX_HIP(hipGraphLaunch(p->exec, p->stream));
  } else {
    enqueue(p, D, V);
  }
  p->h_en.resize(2 * (size_t)n);
  CHX_HIP(hipMemcpyAsync(p->h_en.data(), p->d_en.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, p->stream));
  CHX_HIP(hipStreamSynchronize(p->stream));
  for (int m = 0; m < n; ++m) {
    if (E) E[m] = p->h_en[2 * m];
    if (N) N[m] = p->h_en[2 * m + 1];
  }
}
}  // namespace

extern "C" {

int chx_pack_create(chx_grid* const* grids, int n, chx_pack** out) {
  chx_pack* p = nullptr;
  int rc = pguard([&] {
    prequire(grids && n > 0 && out, CHX_ERR_ARG, "bad pack arguments");
    p = new chx_pack;
    for (int m = 0; m < n; ++m) {
      chx_grid* g = grids[m];
      prequire(g != nullptr, CHX_ERR_ARG, "null grid in pack");
      prequire(g->precision == CHX_PREC_MIXED, CHX_ERR_UNSUPPORTED, "a pack takes grids in the mixed-precision form");
      prequire(g->ao_ready, CHX_ERR_STATE, "Grid::getAO must be called on every grid before packing");
      prequire(m == 0 || g->type == grids[0]->type, CHX_ERR_ARG, "all grids of a pack need the same type");
      for (int k = 0; k < m; ++k) prequire(grids[k] != g, CHX_ERR_ARG, "a grid appears twice in the pack");
      CHX_HIP(hipStreamSynchronize(g->stream));
      p->grids.push_back(g);
    }
    p->type = grids[0]->type;
    CHX_HIP(hipSetDevice(grids[0]->device));
    CHX_HIP(hipStreamCreate(&p->stream));
    p->d_en.alloc(2 * (size_t)n);
    build_lanes(p);
  });
  if (rc != CHX_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return CHX_OK;
}

int chx_pack_xc(chx_pack* p, const chx_xc* xc, const double* const* D_dev, double* const* V_dev, double* E, double* N) {
  return pguard([&] {
    prequire(p && xc && D_dev && V_dev, CHX_ERR_ARG, "bad arguments");
    run(p, xc, D_dev, V_dev, E, N);
  });
}

int chx_pack_set_graph(chx_pack* p, int enable) {
  return pguard([&] {
    prequire(p != nullptr, CHX_ERR_ARG, "bad arguments");
    p->use_graph = enable != 0;
    p->drop_graph();
  });
}

int chx_pack_autotune(chx_pack* p, const chx_xc* xc, const double* const* D_dev, double* const* V_dev, int reps,
                      int* chunk_out, float* ms_out) {
  return pguard([&] {
    prequire(p && xc && D_dev && V_dev && reps > 0, CHX_ERR_ARG, "bad arguments");
    static const int cand[] = {8, 16, 32, 128};
    hipEvent_t a, b;
    CHX_HIP(hipEventCreate(&a));
    CHX_HIP(hipEventCreate(&b));
    int best = -1;
    float best_ms = 0.f;
    for (int k = 0; k < 4; ++k) {
      for (chx_grid* g : p->grids) set_vxc_chunk(g, cand[k]);
      build_lanes(p);
      run(p, xc, D_dev, V_dev, nullptr, nullptr);  // capture and warm up
      CHX_HIP(hipEventRecord(a, p->stream));
      for (int r = 0; r < reps; ++r) run(p, xc, D_dev, V_dev, nullptr, nullptr);
      CHX_HIP(hipEventRecord(b, p->stream));
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
    for (chx_grid* g : p->grids) set_vxc_chunk(g, cand[best]);
    // then the Vxc tile shape per molecule, by a basis-size threshold: 256-AO union groups for
    // npad >= T, 128-AO super-tiles below (T = 0: all groups; -1: none)
    static const int thr[] = {-1, 1024, 512, 0};
    int best_thr = -1;
    float best_thr_ms = 0.f;
    for (int T : thr) {
      for (chx_grid* g : p->grids) set_vxc_shape(g, T >= 0 && g->npad >= T ? 1 : 0);
      build_lanes(p);
      run(p, xc, D_dev, V_dev, nullptr, nullptr);
      CHX_HIP(hipEventRecord(a, p->stream));
      for (int r = 0; r < reps; ++r) run(p, xc, D_dev, V_dev, nullptr, nullptr);
      CHX_HIP(hipEventRecord(b, p->stream));
      CHX_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      CHX_HIP(hipEventElapsedTime(&ms, a, b));
      if (best_thr_ms == 0.f || ms < best_thr_ms) {
        best_thr = T;
        best_thr_ms = ms;
      }
    }
    for (chx_grid* g : p->grids) set_vxc_shape(g, best_thr >= 0 && g->npad >= best_thr ? 1 : 0);
    p->wide_from = best_thr;
    // then the form and lane count at that length: per-molecule kernels or per-lane stage launches
    // on 4, 8 or 16 lanes (at most one lane per molecule), or persistent stages (LDA/GGA)
    const int n = (int)p->grids.size();
    const bool st_ok = p->type <= 1 && xc->family <= 1;
    int best_form = -1, best_lanes = 0;
    float best_form_ms = 0.f;
    const int cands[7][2] = {{0, 4}, {0, 8}, {0, 16}, {2, 4}, {2, 8}, {2, 16}, {1, 16}};
    for (const auto& c : cands) {
      if (c[0] != 0 && !st_ok) continue;
      if (c[0] != 1 && c[1] > n && c[1] != 4) continue;
      p->form = c[0];
      p->lanes = std::min(c[1], n);
      build_lanes(p);
      run(p, xc, D_dev, V_dev, nullptr, nullptr);
      CHX_HIP(hipEventRecord(a, p->stream));
      for (int r = 0; r < reps; ++r) run(p, xc, D_dev, V_dev, nullptr, nullptr);
      CHX_HIP(hipEventRecord(b, p->stream));
      CHX_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      CHX_HIP(hipEventElapsedTime(&ms, a, b));
      if (best_form < 0 || ms < best_form_ms) {
        best_form = p->form;
        best_lanes = p->lanes;
        best_form_ms = ms;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    p->form = best_form;
    p->lanes = best_lanes;
    build_lanes(p);
    if (chunk_out) *chunk_out = cand[best];
  });
}

int chx_pack_vxc_wide_from(const chx_pack* p, int* npad_min) {
  return pguard([&] {
    prequire(p && npad_min, CHX_ERR_ARG, "bad arguments");
    *npad_min = p->wide_from;
  });
}

int chx_pack_lanes(const chx_pack* p, int* lanes) {
  return pguard([&] {
    prequire(p && lanes, CHX_ERR_ARG, "bad arguments");
    *lanes = (int)p->cut.size() - 1;
  });
}

int chx_pack_set_lanes(chx_pack* p, int lanes) {
  return pguard([&] {
    prequire(p != nullptr && lanes >= 1 && lanes <= 16, CHX_ERR_ARG, "lanes must be in 1..16");
    CHX_HIP(hipStreamSynchronize(p->stream));
    for (hipStream_t t : p->lane_st) CHX_HIP(hipStreamSynchronize(t));
    p->lanes = lanes;
    if (p->form == 1) p->form = 0;  // lanes, not persistent stages
    build_lanes(p);
  });
}

int chx_pack_set_form(chx_pack* p, int form) {
  return pguard([&] {
    prequire(p != nullptr && form >= 0 && form <= 2, CHX_ERR_ARG, "pack form must be 0, 1 or 2");
    prequire(form == 0 || p->type <= 1, CHX_ERR_UNSUPPORTED, "stage launches are built for LDA and GGA grids");
    CHX_HIP(hipStreamSynchronize(p->stream));
    for (hipStream_t t : p->lane_st) CHX_HIP(hipStreamSynchronize(t));
    p->form = form;
    p->drop_graph();
  });
}

int chx_pack_form(const chx_pack* p, int* form) {
  return pguard([&] {
    prequire(p && form, CHX_ERR_ARG, "bad arguments");
    *form = p->form;
  });
}

int chx_pack_stream(const chx_pack* p, void** stream) {
  return pguard([&] {
    prequire(p && stream, CHX_ERR_ARG, "bad arguments");
    *stream = (void*)p->stream;
  });
}

void chx_pack_destroy(chx_pack* p) { delete p; }

}  // extern "C"
