// Many molecules in one launch per stage (mixed-precision form).
//
// A pack holds grids that were each built and planned on their own (points,
// Becke weights, screening, split operand cache). Per XC build it runs, on one
// stream: split D per molecule -> one density launch over every molecule's
// work groups -> functional per molecule -> one B^T launch over every slot ->
// one Vxc launch over every work item -> V = M + M^T per molecule. The big
// kernels take a list of (molecule, item) pairs and a device array of
// per-molecule descriptors, so small molecules fill the GPU together with
// large ones. The whole sequence is captured once as a hipGraph and replayed
// each SCF iteration (the small per-molecule kernels then cost no launch
// overhead); a capture is redone when the D/V pointers or the functional
// change. chx_pack_autotune picks the Vxc work-item length by timing.
//
// No reference counterpart: Chinium runs one molecule per process
// (src/HartreeFockKohnSham/SelfConsistentField.cpp:27-33).
#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "internal.hpp"

namespace chx {
void set_vxc_chunk(chx_grid* g, int chunk);
std::string& last_error();
}
using namespace chx;

struct chx_pack {
  std::vector<chx_grid*> grids;  // not owned
  hipStream_t stream = nullptr;
  int type = 0;
  DevBuf<MolDev> d_mols;
  DevBuf<int2> d_dens, d_mb1, d_mb4, d_vxc;
  DevBuf<int> d_codes;
  DevBuf<double> d_en;  // E, N per molecule
  std::vector<int> codes;
  int family = -1;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const double*> gD;  // pointers the graph was captured with
  std::vector<double*> gV;
  std::vector<double> h_en;
  ~chx_pack() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
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

// Work lists over all molecules: molecules in order, each one's items in its
// own (XCD-friendly) order.
void build_items(chx_pack* p) {
  std::vector<MolDev> mols;
  std::vector<int2> dens, mb1, mb4, vxc;
  for (size_t m = 0; m < p->grids.size(); ++m) {
    const chx_grid* g = p->grids[m];
    mols.push_back(g->hmol);
    const int mi = (int)m;
    for (int k = 0; k < (int)g->grp_batch.size(); ++k) dens.push_back(make_int2(mi, k));
    const int ns = (int)g->bt_list.size();
    for (int k = 0; k < ns; k += 4) mb1.push_back(make_int2(mi, k));
    for (int k = 0; k < ns; ++k) mb4.push_back(make_int2(mi, k));
    for (int k = 0; k < (int)g->cwork.size(); ++k) vxc.push_back(make_int2(mi, k));
  }
  p->d_mols.upload(mols, p->stream);
  p->d_dens.upload(dens, p->stream);
  p->d_mb1.upload(mb1, p->stream);
  p->d_mb4.upload(mb4, p->stream);
  p->d_vxc.upload(vxc, p->stream);
  CHX_HIP(hipStreamSynchronize(p->stream));
  p->drop_graph();
}

// One XC build of every molecule, enqueued on the pack's stream.
void enqueue(chx_pack* p, const double* const* D, double* const* V) {
  hipStream_t s = p->stream;
  const int n = (int)p->grids.size();
  for (int m = 0; m < n; ++m) {
    chx_grid* g = p->grids[m];
    launch_split_d(s, g->nbasis, g->npad, D[m], g->d_int2ref.p, g->d_Dsplit.p);
  }
  const int nd = (int)p->d_dens.n;
  for (int r = 0; r <= (p->type >= 2 ? 3 : 0); ++r) launch_density_mx(s, r, nd, p->grids[0]->hmol, p->d_mols.p, p->d_dens.p);
  for (int m = 0; m < n; ++m) {
    chx_grid* g = p->grids[m];
    launch_functional(s, g->nbatches, g->d_grp_off.p, g->d_part.p, g->d_part_tau.p, g->d_w.p, p->d_codes.p,
                      (int)p->codes.size(), p->family, 1, 1, g->d_rho.p, g->d_coef.p, g->d_en_part.p, g->en_blocks);
    launch_reduce_en(s, g->en_blocks, g->d_en_part.p, p->d_en.p + 2 * m);
  }
  const int nc = p->family >= 2 ? 4 : 1;
  if (nc == 4)
    launch_make_b_mx(s, 4, (int)p->d_mb4.n, p->grids[0]->hmol, p->d_mols.p, p->d_mb4.p);
  else
    launch_make_b_mx(s, 1, (int)p->d_mb1.n, p->grids[0]->hmol, p->d_mols.p, p->d_mb1.p);
  for (int m = 0; m < n; ++m) {
    chx_grid* g = p->grids[m];
    launch_fill(s, g->d_Mint.p, (size_t)g->npad * g->npad, 0.0);
  }
  launch_vxc_mx(s, nc, (int)p->d_vxc.n, p->grids[0]->hmol, p->d_mols.p, p->d_vxc.p);
  for (int m = 0; m < n; ++m) {
    chx_grid* g = p->grids[m];
    launch_finalize_sym(s, g->nbasis, g->npad, g->d_Mint.p, g->d_ref2int.p, V[m]);
  }
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
      if (g->mx_wide) {  // the pack's Vxc launch runs 128-AO super-tiles: re-plan the grid's items
        g->mx_wide = 0;
        set_vxc_chunk(g, g->vxc_chunk);
      }
      p->grids.push_back(g);
    }
    p->type = grids[0]->type;
    CHX_HIP(hipSetDevice(grids[0]->device));
    CHX_HIP(hipStreamCreate(&p->stream));
    p->d_en.alloc(2 * (size_t)n);
    build_items(p);
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
    static const int cand[] = {16, 32, 64, 128};
    hipEvent_t a, b;
    CHX_HIP(hipEventCreate(&a));
    CHX_HIP(hipEventCreate(&b));
    int best = -1;
    float best_ms = 0.f;
    for (int k = 0; k < 4; ++k) {
      for (chx_grid* g : p->grids) set_vxc_chunk(g, cand[k]);
      build_items(p);
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
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    for (chx_grid* g : p->grids) set_vxc_chunk(g, cand[best]);
    build_items(p);
    if (chunk_out) *chunk_out = cand[best];
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
