// Grid construction and batch planning (host side).
//
// Restates Grid::Grid (src/Grid/GridGeneration.cpp:27-162,232-310) for the
// GPU: atom-centred Lebedev x radial points, Becke partition weights (FP64,
// on the device), then this build's own batching: spatially compact batches
// of 64 points (recursive bisection along the longest box edge) instead of
// the reference's 128-point halves of the generation order. Batch membership
// changes only the order of summation.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <unordered_map>

#include "internal.hpp"

namespace chx {
namespace {

struct LebOrbit { int type; double a, b, v; };
struct LebRule { int n, first, count; };
#include "lebedev_table.inc"

const char* kElemNames[] = {"", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
                            "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni",
                            "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};

// Octahedral orbit of a generator (Lebedev-Laikov types 1-6).
void expand_orbit(const LebOrbit& o, std::vector<std::array<double, 4>>& out) {
  double g[3];
  switch (o.type) {
    case 1: g[0] = 1; g[1] = 0; g[2] = 0; break;
    case 2: g[0] = 0; g[1] = g[2] = std::sqrt(0.5); break;
    case 3: g[0] = g[1] = g[2] = std::sqrt(1.0 / 3.0); break;
    case 4: g[0] = g[1] = o.a; g[2] = std::sqrt(1.0 - 2.0 * o.a * o.a); break;
    case 5: g[0] = o.a; g[1] = std::sqrt(1.0 - o.a * o.a); g[2] = 0; break;
    case 6: g[0] = o.a; g[1] = o.b; g[2] = std::sqrt(1.0 - o.a * o.a - o.b * o.b); break;
    default: throw Error(CHX_ERR_ARG, "bad Lebedev orbit");
  }
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::vector<std::array<double, 3>> base;
  for (const auto& p : perm) {
    std::array<double, 3> c = {g[p[0]], g[p[1]], g[p[2]]};
    if (std::find(base.begin(), base.end(), c) == base.end()) base.push_back(c);
  }
  for (const auto& c : base)
    for (int s = 0; s < 8; ++s) {
      const double sx = (s & 1) ? -1.0 : 1.0, sy = (s & 2) ? -1.0 : 1.0, sz = (s & 4) ? -1.0 : 1.0;
      if ((c[0] == 0 && sx < 0) || (c[1] == 0 && sy < 0) || (c[2] == 0 && sz < 0)) continue;
      out.push_back({sx * c[0], sy * c[1], sz * c[2], o.v});
    }
}

const std::vector<std::array<double, 4>>& lebedev(int n) {
  static std::vector<std::vector<std::array<double, 4>>> cache(6000);
  if (n <= 0 || n >= (int)cache.size()) throw Error(CHX_ERR_FILE, "Lebedev rule not available");
  auto& r = cache[n];
  if (!r.empty()) return r;
  for (const LebRule& lr : kLebRules) {
    if (lr.n != n) continue;
    for (int k = 0; k < lr.count; ++k) expand_orbit(kLebOrbits[lr.first + k], r);
    if ((int)r.size() != n) throw Error(CHX_ERR_FILE, "Lebedev expansion size mismatch");
    return r;
  }
  throw Error(CHX_ERR_FILE, "Lebedev rule not available");
}

struct Radial { std::vector<double> r, w; std::vector<int> nang; };

// .grid file (Grids/README.md:1-11): "ngroups nshells", radial line ("em R" or
// "de2 alpha xmin xmax"), then "npoints nshells" per group from the nucleus out.
Radial read_grid_file(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw Error(CHX_ERR_FILE, "Missing element grid file in folder!");
  std::string line;
  int ngroups = 0, ntot = 0;
  if (!std::getline(f, line)) throw Error(CHX_ERR_FILE, "Corrupted grid file");
  { std::stringstream ss(line); ss >> ngroups >> ntot; }
  if (ngroups <= 0 || ntot <= 1) throw Error(CHX_ERR_FILE, "Corrupted grid file");
  if (!std::getline(f, line)) throw Error(CHX_ERR_FILE, "Corrupted grid file");
  std::stringstream ss(line);
  std::string formula;
  ss >> formula;
  Radial R;
  R.r.resize(ntot);
  R.w.resize(ntot);
  if (formula == "de2") {  // GridGeneration.cpp:90-100
    double alpha, x0, x1;
    ss >> alpha >> x0 >> x1;
    const double h = (x1 - x0) / (double)(ntot - 1);
    for (int i = 0; i < ntot; ++i) {
      const double xi = h * i + x0;
      R.r[i] = std::exp(alpha * xi - std::exp(-xi));
      R.w[i] = std::exp(3 * alpha * xi - 3 * std::exp(-xi)) * (alpha + std::exp(-xi)) * h;
    }
  } else if (formula == "em") {  // GridGeneration.cpp:103-113
    double Rm;
    ss >> Rm;
    for (int i = 0; i < ntot; ++i) {
      const double ii = i + 1;
      R.r[i] = Rm * std::pow(ii / (ntot + 1 - ii), 2);
      R.w[i] = 2 * std::pow(Rm, 3) * (ntot + 1) * std::pow(ii, 5) / std::pow(ntot + 1 - ii, 7);
    }
  } else {
    throw Error(CHX_ERR_FILE, "Unrecognized radial formula!");
  }
  for (int k = 0; k < ngroups; ++k) {
    if (!std::getline(f, line)) throw Error(CHX_ERR_FILE, "Corrupted grid file");
    std::stringstream gs(line);
    int npts = 0, nsh = 0;
    gs >> npts >> nsh;
    for (int s = 0; s < nsh; ++s) R.nang.push_back(npts);
  }
  if ((int)R.nang.size() != ntot) throw Error(CHX_ERR_FILE, "Corrupted grid file");
  return R;
}

// Nuclei with Z > 2 in a uniform hash of 1-bohr cells, to find the nucleus
// (if any) within r <= 0.5 bohr of a point by looking at 27 cells.
struct HeavyHash {
  std::unordered_map<uint64_t, std::vector<int>> cells;
  const std::vector<double>* xyz = nullptr;
  static uint64_t key(long ix, long iy, long iz) {
    return ((uint64_t)(ix & 0x1fffff) << 42) | ((uint64_t)(iy & 0x1fffff) << 21) | (uint64_t)(iz & 0x1fffff);
  }
  HeavyHash(const chx_mol& m) : xyz(&m.xyz) {
    for (int a = 0; a < m.natoms; ++a)
      if (m.Z[a] > 2)
        cells[key((long)std::floor(m.xyz[3 * a]), (long)std::floor(m.xyz[3 * a + 1]), (long)std::floor(m.xyz[3 * a + 2]))]
            .push_back(a);
  }
  // a heavy atom within sqrt(r2) of (x, y, z), sqrt(r2) <= 1; -1 if none
  int near(double x, double y, double z, double r2) const {
    if (cells.empty()) return -1;
    const long ix = (long)std::floor(x), iy = (long)std::floor(y), iz = (long)std::floor(z);
    for (long dx = -1; dx <= 1; ++dx)
      for (long dy = -1; dy <= 1; ++dy)
        for (long dz = -1; dz <= 1; ++dz) {
          auto it = cells.find(key(ix + dx, iy + dy, iz + dz));
          if (it == cells.end()) continue;
          for (int a : it->second) {
            const double ex = x - (*xyz)[3 * a], ey = y - (*xyz)[3 * a + 1], ez = z - (*xyz)[3 * a + 2];
            if (ex * ex + ey * ey + ez * ez < r2) return a;
          }
        }
    return -1;
  }
};

// Recursive bisection into batches of exactly kBatch points (one remainder).
void bisect(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& z,
            std::vector<long>& idx, long lo, long hi, std::vector<std::pair<long, long>>& out) {
  const long n = hi - lo;
  if (n <= kBatch) {
    out.push_back({lo, hi});
    return;
  }
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (long i = lo; i < hi; ++i) {
    const double c[3] = {x[idx[i]], y[idx[i]], z[idx[i]]};
    for (int d = 0; d < 3; ++d) { mn[d] = std::min(mn[d], c[d]); mx[d] = std::max(mx[d], c[d]); }
  }
  int ax = 0;
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
  const std::vector<double>& key = ax == 0 ? x : (ax == 1 ? y : z);
  const long nb = (n + kBatch - 1) / kBatch;
  const long left = kBatch * (nb / 2);
  std::nth_element(idx.begin() + lo, idx.begin() + lo + left, idx.begin() + hi,
                   [&](long a, long b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
  bisect(x, y, z, idx, lo, lo + left, out);
  bisect(x, y, z, idx, lo + left, hi, out);
}

uint64_t morton3(double fx, double fy, double fz) {
  auto spread = [](uint64_t v) {
    uint64_t r = 0;
    for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
    return r;
  };
  auto q = [](double f) {
    f = std::min(std::max(f, 0.0), 1.0);
    return (uint64_t)(f * 2097151.0);
  };
  return spread(q(fx)) | (spread(q(fy)) << 1) | (spread(q(fz)) << 2);
}

}  // namespace

const std::vector<std::array<double, 4>>& lebedev_rule(int n) { return lebedev(n); }

void ensure_stream(chx_grid* g) {
  if (g->stream) return;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(CHX_ERR_DEVICE, "no HIP device available");
  CHX_HIP(hipGetDevice(&g->device));
  CHX_HIP(hipStreamCreate(&g->stream));
}

// Build points, weights, batches and the internal AO layout.
void grid_build(chx_grid* g, const std::string& grid_name, const std::string& data_root) {
  const chx_mol& m = g->mol;
  const int nat = m.natoms;
  if (nat <= 0) throw Error(CHX_ERR_ARG, "molecule has no atoms");
  if (m.nbasis <= 0) throw Error(CHX_ERR_ARG, "molecule has no basis");
  // ---- atom-centred points in generation order (atom -> shell -> angular point)
  std::vector<double> px, py, pz, pw;
  std::vector<int> owner;
  for (int a = 0; a < nat; ++a) {
    const int Z = m.Z[a];
    if (Z < 1 || Z > 36) throw Error(CHX_ERR_ARG, "unsupported element");
    const Radial R = read_grid_file(data_root + "/Grids/" + grid_name + "/" + kElemNames[Z] + ".grid");
    const double* c = &m.xyz[3 * a];
    for (size_t s = 0; s < R.r.size(); ++s) {
      const auto& L = lebedev(R.nang[s]);
      for (const auto& l : L) {
        px.push_back(l[0] * R.r[s] + c[0]);
        py.push_back(l[1] * R.r[s] + c[1]);
        pz.push_back(l[2] * R.r[s] + c[2]);
        pw.push_back(R.w[s] * 4 * M_PI * l[3]);
        owner.push_back(a);
      }
    }
  }
  const long G = (long)px.size();
  // ---- Becke weights on the device (FP64); the data files are all read by now
  ensure_stream(g);
  {
    std::vector<double> inv(nat * (size_t)nat, 0.0);
    for (int b = 0; b < nat; ++b)
      for (int c = 0; c < nat; ++c) {
        if (b == c) continue;
        const double dx = m.xyz[3 * b] - m.xyz[3 * c], dy = m.xyz[3 * b + 1] - m.xyz[3 * c + 1],
                     dz = m.xyz[3 * b + 2] - m.xyz[3 * c + 2];
        inv[(size_t)b * nat + c] = std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    DevBuf<double> dx, dy, dz, dw, dxyz, dinv;
    DevBuf<int> down;
    dx.upload(px, g->stream); dy.upload(py, g->stream); dz.upload(pz, g->stream); dw.upload(pw, g->stream);
    down.upload(owner, g->stream);
    dxyz.upload(m.xyz, g->stream);
    dinv.upload(inv, g->stream);
    launch_becke(g->stream, G, dx.p, dy.p, dz.p, dw.p, down.p, nat, dxyz.p, dinv.p);
    CHX_HIP(hipMemcpyAsync(pw.data(), dw.p, sizeof(double) * G, hipMemcpyDeviceToHost, g->stream));
    CHX_HIP(hipStreamSynchronize(g->stream));
  }
  // ---- drop exactly-zero weights, batch the rest spatially
  std::vector<long> keep;
  keep.reserve(G);
  for (long i = 0; i < G; ++i)
    if (pw[i] != 0.0) keep.push_back(i);
  // The points within kCoreRadius of a nucleus with Z > 2 (atomic cores, where
  // the density is largest) are batched apart, per atom, so that the mixed
  // form's three-term core treatment covers few batches; the rest is bisected
  // as one set. Only the batch membership depends on this.
  std::vector<std::pair<long, long>> parts;
  {
    const double rc2 = kCoreRadius * kCoreRadius;
    const HeavyHash hh(m);
    std::vector<std::vector<long>> core(nat);
    std::vector<long> rest;
    rest.reserve(keep.size());
    for (long i : keep) {  // any nucleus with Z > 2, not only the point's own atom's
      const int a = hh.near(px[i], py[i], pz[i], rc2);
      if (a >= 0) core[a].push_back(i);
      else rest.push_back(i);
    }
    keep.clear();
    std::vector<long> seg = {0};
    for (int a = 0; a < nat; ++a) {
      keep.insert(keep.end(), core[a].begin(), core[a].end());
      seg.push_back((long)keep.size());
    }
    keep.insert(keep.end(), rest.begin(), rest.end());
    seg.push_back((long)keep.size());
    for (size_t k = 0; k + 1 < seg.size(); ++k)
      if (seg[k + 1] > seg[k]) bisect(px, py, pz, keep, seg[k], seg[k + 1], parts);
  }
  g->npoints = (long)keep.size();
  g->nbatches = (int)parts.size();
  const size_t nslots = (size_t)g->nbatches * kBatch;
  g->hx.assign(nslots, 0); g->hy.assign(nslots, 0); g->hz.assign(nslots, 0); g->hw.assign(nslots, 0);
  g->batch_n.assign(g->nbatches, 0);
  for (int b = 0; b < g->nbatches; ++b) {
    const long lo = parts[b].first, n = parts[b].second - parts[b].first;
    g->batch_n[b] = (int)n;
    for (int k = 0; k < kBatch; ++k) {
      const long src = keep[lo + (k < n ? k : 0)];
      const size_t s = (size_t)b * kBatch + k;
      g->hx[s] = px[src]; g->hy[s] = py[src]; g->hz[s] = pz[src];
      g->hw[s] = k < n ? pw[src] : 0.0;
    }
  }
  // ---- internal AO order: atoms along a Morton curve, shells in file order
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int a = 0; a < nat; ++a)
    for (int d = 0; d < 3; ++d) { mn[d] = std::min(mn[d], m.xyz[3 * a + d]); mx[d] = std::max(mx[d], m.xyz[3 * a + d]); }
  const double span = std::max({mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2], 1e-9});
  std::vector<int> aorder(nat);
  std::iota(aorder.begin(), aorder.end(), 0);
  std::vector<uint64_t> key(nat);
  for (int a = 0; a < nat; ++a)
    key[a] = morton3((m.xyz[3 * a] - mn[0]) / span, (m.xyz[3 * a + 1] - mn[1]) / span, (m.xyz[3 * a + 2] - mn[2]) / span);
  std::stable_sort(aorder.begin(), aorder.end(), [&](int p, int q) { return key[p] < key[q]; });
  g->nbasis = m.nbasis;
  g->npad = ((m.nbasis + kPadAO - 1) / kPadAO) * kPadAO;
  g->ntiles = g->npad / kTile;
  g->int2ref.assign(g->npad, -1);
  g->ref2int.assign(g->nbasis, -1);
  g->sh_xyz.clear(); g->sh_type.clear(); g->sh_nprim.clear(); g->sh_poff.clear(); g->sh_atom.clear();
  g->p_exp.clear(); g->p_coef.clear();
  std::vector<int> sh_int_ao;  // internal AO offset per internal shell
  int ao = 0;
  for (int a : aorder)
    for (const ShellH& s : m.shells) {
      if (s.center != a) continue;
      const int n = shell_size(s.type);
      for (int k = 0; k < n; ++k) {
        g->int2ref[ao + k] = s.ao_off + k;
        g->ref2int[s.ao_off + k] = ao + k;
      }
      g->sh_xyz.push_back(m.xyz[3 * a]); g->sh_xyz.push_back(m.xyz[3 * a + 1]); g->sh_xyz.push_back(m.xyz[3 * a + 2]);
      g->sh_type.push_back(s.type);
      g->sh_atom.push_back(a);
      g->sh_nprim.push_back(s.nprim);
      g->sh_poff.push_back((int)g->p_exp.size());
      for (int p = 0; p < s.nprim; ++p) {
        g->p_exp.push_back(m.exps[s.prim_off + p]);
        g->p_coef.push_back(m.coefs[s.prim_off + p]);
      }
      sh_int_ao.push_back(ao);
      ao += n;
    }
  // ---- tile segments
  g->tile_seg_off.assign(g->ntiles + 1, 0);
  g->seg.clear();
  for (int t = 0; t < g->ntiles; ++t) {
    g->tile_seg_off[t] = (int)g->seg.size();
    const int lo = t * kTile, hi = lo + kTile;
    for (size_t s = 0; s < sh_int_ao.size(); ++s) {
      const int a0 = sh_int_ao[s], a1 = a0 + shell_size(g->sh_type[s]);
      const int b0 = std::max(a0, lo), b1 = std::min(a1, hi);
      if (b0 >= b1) continue;
      g->seg.push_back(make_int4((int)s, b0 - a0, b1 - b0, b0 - lo));
    }
  }
  g->tile_seg_off[g->ntiles] = (int)g->seg.size();
  // ---- device copies
  hipStream_t st = g->stream;
  g->d_sh_xyz.upload(g->sh_xyz, st); g->d_sh_type.upload(g->sh_type, st); g->d_sh_nprim.upload(g->sh_nprim, st);
  g->d_sh_atom.upload(g->sh_atom, st);
  g->d_sh_poff.upload(g->sh_poff, st); g->d_p_exp.upload(g->p_exp, st); g->d_p_coef.upload(g->p_coef, st);
  std::vector<SegRec> rec(g->seg.size());  // lives until the synchronisation below
  g->max_nprim = 0;
  for (size_t k = 0; k < g->seg.size(); ++k) {
    const int4 sg = g->seg[k];
    const int sh = sg.x, np = g->sh_nprim[sh], po = g->sh_poff[sh];
    g->max_nprim = std::max(g->max_nprim, np);
    SegRec& r = rec[k];
    std::memset(&r, 0, sizeof(r));
    r.ax = g->sh_xyz[3 * sh]; r.ay = g->sh_xyz[3 * sh + 1]; r.az = g->sh_xyz[3 * sh + 2];
    r.meta = (g->sh_type[sh] & 0xff) | (std::min(np, kSegPrim) << 8) | (sg.y << 16) | (sg.z << 24);
    r.col = sg.w;
    for (int p = 0; p < np && p < kSegPrim; ++p) {
      r.e[p] = (float)g->p_exp[po + p];
      r.c[p] = (float)g->p_coef[po + p];
    }
  }
  g->d_segrec.upload(rec, st);
  g->d_tile_seg_off.upload(g->tile_seg_off, st); g->d_seg.upload(g->seg, st);
  g->d_int2ref.upload(g->int2ref, st); g->d_ref2int.upload(g->ref2int, st);
  g->d_x.upload(g->hx, st); g->d_y.upload(g->hy, st); g->d_z.upload(g->hz, st); g->d_w.upload(g->hw, st);
  g->d_bn.upload(g->batch_n, st);
  CHX_HIP(hipStreamSynchronize(st));
  g->ao_ready = false;
}

// Vxc work items of the mixed form for large bases: the batches are cut into
// chunks of `chunk` consecutive batches (the batch order is spatial, so a
// chunk covers a compact region), and the union of the alive tiles of a
// chunk's batches is cut into groups of 8 tiles in tile order. An item is an
// pair of groups I <= J of one chunk: the 256 x 256 blocks of M at those
// tiles' rows and columns, (I, J) and for I < J (J, I), accumulated over the
// chunk's batches. Tiles of a
// group that are not alive in a batch read the all-zero slot in the table,
// and the kernel skips their loads and products, so no MFMA multiplies a
// tile-pair block that the screening dropped. Items are chunk-major: the
// items that run together on one XCD share the chunk's operand slots in its
// L2.
static void build_union_plan(chx_grid* g, int chunk) {
  const int nb = g->nbatches, nt = g->ntiles, G = 2 * kCacheSuper;
  const int zero = (int)g->bt_list.size();
  g->cwork.clear();
  g->cpair_batches.clear();
  g->cslot.clear();
  g->vchunk.clear();
  g->vtab.clear();
  g->vtile.clear();
  std::vector<char> in_u(nt);
  std::vector<int> U;
  double useful = 0, dense = 0;
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int b1 = std::min(nb, b0 + chunk);
    std::fill(in_u.begin(), in_u.end(), 0);
    for (int b = b0; b < b1; ++b)
      for (int it = g->bt_off[b]; it < g->bt_off[b + 1]; ++it) in_u[g->bt_list[it]] = 1;
    U.clear();
    for (int t = 0; t < nt; ++t)
      if (in_u[t]) U.push_back(t);
    if (U.empty()) continue;
    const int nu = (int)U.size(), ng = (nu + G - 1) / G, npos = ng * G;
    const int ci = (int)g->vchunk.size();
    g->vchunk.push_back(make_int4((int)g->vtab.size(), b1 - b0, npos, (int)g->vtile.size()));
    for (int u = 0; u < npos; ++u) g->vtile.push_back(u < nu ? U[u] : -1);
    for (int b = b0; b < b1; ++b) {
      for (int u = 0; u < npos; ++u) {
        const int si = u < nu ? g->slot_of[(size_t)b * nt + U[u]] : -1;
        g->vtab.push_back(si >= 0 ? si : zero);
      }
      const double n = g->bt_off[b + 1] - g->bt_off[b];
      useful += n * n;
    }
    dense += (double)(b1 - b0) * nu * nu;
    for (int I = 0; I < ng; ++I)  // an item covers the blocks (I, J) and (J, I)
      for (int J = I; J < ng; ++J) g->cwork.push_back(make_int4(I, J, ci, 0));
  }
  g->vxc_exec_frac = dense > 0 ? useful / dense : 1.0;
}

// Vxc work items of the cached / mixed paths: (super-tile I, super-tile J,
// range of pair_batches) over 128-AO super-tiles, a chunk of batches where
// both are alive. Chunk-major order: blocks that run at the same time work on
// the same batches of different (I, J) pairs, so their operand tiles meet in
// L2. Fewer, longer items mean fewer FP64 atomics into M. Large bases in the
// mixed form take the chunk tile-group plan above.
void build_cwork(chx_grid* g) {
  const int nb = g->nbatches, nt = g->ntiles;
  const bool mixed = g->precision == CHX_PREC_MIXED;
  size_t chunk = g->vxc_chunk > 0 ? (size_t)g->vxc_chunk : (mixed ? 32 : kCacheChunk);
  if (mixed) chunk = std::min<size_t>(chunk, kMxMaxChunk);  // the kernel stages one item's slot table
  if (mixed && g->mx_wide) {
    build_union_plan(g, (int)chunk);
    return;
  }
  g->vchunk.clear();
  g->vtab.clear();
  g->vtile.clear();
  const int sup = kCacheSuper;  // tiles per super-tile
  const int nsc = g->npad / (sup * kTile);
  std::vector<std::vector<int>> cp((size_t)nsc * nsc);
  std::vector<int> alive_sup;
  for (int b = 0; b < nb; ++b) {
    alive_sup.clear();
    for (int sI = 0; sI < nsc; ++sI) {
      bool any = false;
      for (int h = 0; h < sup; ++h) any |= g->alive[(size_t)b * nt + sI * sup + h] != 0;
      if (any) alive_sup.push_back(sI);
    }
    for (int I : alive_sup)
      for (int J : alive_sup) cp[(size_t)I * nsc + J].push_back(b);
  }
  g->cwork.clear();
  g->cpair_batches.clear();
  size_t maxlen = 0;
  for (const auto& L : cp) maxlen = std::max(maxlen, L.size());
  for (size_t k = 0; k < maxlen; k += chunk)
    for (int I = 0; I < nsc; ++I)
      for (int J = 0; J < nsc; ++J) {
        const auto& L = cp[(size_t)I * nsc + J];
        if (k >= L.size()) continue;
        const int begin = (int)g->cpair_batches.size();
        const size_t e = std::min(L.size(), k + chunk);
        g->cpair_batches.insert(g->cpair_batches.end(), L.begin() + k, L.begin() + e);
        g->cwork.push_back(make_int4(I, J, begin, (int)g->cpair_batches.size()));
      }
  g->cslot.clear();
  if (mixed) {  // per batch of an item: the slots of I's and J's tiles, the all-zero slot if not alive
    const int zero = (int)g->bt_list.size();
    g->cslot.resize(g->cpair_batches.size() * 2 * sup);
    for (const int4& w : g->cwork)
      for (int k = w.z; k < w.w; ++k) {
        const int b = g->cpair_batches[k];
        for (int t = 0; t < sup; ++t) {
          const int si = g->slot_of[(size_t)b * nt + w.x * sup + t];
          const int sj = g->slot_of[(size_t)b * nt + w.y * sup + t];
          g->cslot[(size_t)k * 2 * sup + t] = si >= 0 ? si : zero;
          g->cslot[(size_t)k * 2 * sup + sup + t] = sj >= 0 ? sj : zero;
        }
      }
  }
}

// Uploads the mixed form's Vxc plan and points the descriptor at it.
static void upload_mx_plan(chx_grid* g, hipStream_t st) {
  // the plan the grid does not use stays a one-element placeholder (the copies
  // are asynchronous, so their sources outlive the call)
  static const std::vector<int> one{0};
  static const std::vector<int4> one4{make_int4(0, 0, 0, 0)};
  g->d_cslot.upload(g->cslot.empty() ? one : g->cslot, st);
  g->d_vchunk.upload(g->vchunk.empty() ? one4 : g->vchunk, st);
  g->d_vtab.upload(g->vtab.empty() ? one : g->vtab, st);
  g->d_vtile.upload(g->vtile.empty() ? one : g->vtile, st);
  g->hmol.cwork = g->d_cwork.p;
  g->hmol.pair_batches = g->d_cpair_batches.p;
  g->hmol.cslot = g->d_cslot.p;
  g->hmol.vchunk = g->d_vchunk.p;
  g->hmol.vtab = g->d_vtab.p;
  g->hmol.vtile = g->d_vtile.p;
}

// Rebuilds the Vxc work items for another chunk length (autotuning) without
// touching the AO tables.
void set_vxc_chunk(chx_grid* g, int chunk) {
  g->vxc_chunk = chunk;
  if (!g->ao_ready || !g->cached) return;
  ++g->plan_gen;
  build_cwork(g);
  hipStream_t st = g->stream;
  g->d_cwork.upload(g->cwork, st);
  g->d_cpair_batches.upload(g->cpair_batches, st);
  if (g->precision == CHX_PREC_MIXED) upload_mx_plan(g, st);
  CHX_HIP(hipStreamSynchronize(st));
}

// Vxc block shape of the mixed form (-1: 256-AO union groups from npad 1024
// on, 128-AO super-tiles below; 0, 1: either), rebuilding the work items.
void set_vxc_shape(chx_grid* g, int shape) {
  g->vxc_shape = shape;
  if (!g->ao_ready || g->precision != CHX_PREC_MIXED) return;
  g->mx_wide = (shape == 1 || (shape < 0 && g->npad >= 1024)) ? 1 : 0;
  set_vxc_chunk(g, g->vxc_chunk);
}

// Screening and work plans (Grid::getAO counterpart).
void grid_plan(chx_grid* g) {
  hipStream_t st = g->stream;
  const int nb = g->nbatches, nt = g->ntiles;
  g->d_alive.alloc((size_t)std::max(1, nb * nt));
  launch_screen(st, nb, nt, g->d_x.p, g->d_y.p, g->d_z.p, g->d_bn.p, g->view(), g->screen_thr, g->d_alive.p);
  g->alive.assign((size_t)nb * nt, 0);
  if (nb) CHX_HIP(hipMemcpyAsync(g->alive.data(), g->d_alive.p, (size_t)nb * nt, hipMemcpyDeviceToHost, st));
  CHX_HIP(hipStreamSynchronize(st));
  // real AO count per tile
  std::vector<int> tile_nao(nt, 0);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < kTile; ++k) tile_nao[t] += g->int2ref[t * kTile + k] >= 0 ? 1 : 0;
  g->bt_off.assign(nb + 1, 0);
  g->batch_work.assign(nb, 0.0);
  g->bt_list.clear();
  g->grp_batch.clear(); g->grp_tiles.clear(); g->grp_ntiles.clear();
  g->grp_off.assign(nb + 1, 0);
  double flops = 0, tiles = 0;
  for (int b = 0; b < nb; ++b) {
    g->bt_off[b] = (int)g->bt_list.size();
    g->grp_off[b] = (int)g->grp_batch.size();
    int nao = 0;
    std::vector<int> lst;
    for (int t = 0; t < nt; ++t)
      if (g->alive[(size_t)b * nt + t]) { lst.push_back(t); nao += tile_nao[t]; }
    g->bt_list.insert(g->bt_list.end(), lst.begin(), lst.end());
    for (size_t k = 0; k < lst.size(); k += kGroup) {
      g->grp_batch.push_back(b);
      const int cnt = (int)std::min<size_t>(kGroup, lst.size() - k);
      g->grp_ntiles.push_back(cnt);
      for (int q = 0; q < kGroup; ++q) g->grp_tiles.push_back(q < cnt ? lst[k + q] : 0);
    }
    tiles += (double)lst.size();
    g->batch_work[b] = (double)g->batch_n[b] * (double)nao * (double)nao;
    flops += g->batch_work[b];
  }
  g->bt_off[nb] = (int)g->bt_list.size();
  g->grp_off[nb] = (int)g->grp_batch.size();
  g->mean_tiles = nb ? tiles / nb : 0;
  // k contractions per build (GGA: Phi D and Phi^T B; meta-GGA adds three of
  // each for tau), 2 flops per multiply-add
  g->flops_build = (g->type >= 2 ? 8.0 : 2.0) * 2.0 * flops;
  // Vxc work items over super-tile pairs
  const int nsup = g->npad / (kSuper * kTile);
  std::vector<std::vector<int>> pairs((size_t)nsup * nsup);
  std::vector<int> sup;
  for (int b = 0; b < nb; ++b) {
    sup.clear();
    for (int s = 0; s < nsup; ++s) {
      bool any = false;
      for (int h = 0; h < kSuper; ++h) any |= g->alive[(size_t)b * nt + s * kSuper + h] != 0;
      if (any) sup.push_back(s);
    }
    for (size_t i = 0; i < sup.size(); ++i)
      for (size_t j = i; j < sup.size(); ++j) pairs[(size_t)sup[i] * nsup + sup[j]].push_back(b);
  }
  g->vwork.clear();
  g->pair_batches.clear();
  for (int I = 0; I < nsup; ++I)
    for (int J = I; J < nsup; ++J) {
      const auto& L = pairs[(size_t)I * nsup + J];
      for (size_t k = 0; k < L.size(); k += kVxcChunk) {
        const int begin = (int)g->pair_batches.size();
        const size_t e = std::min(L.size(), k + kVxcChunk);
        g->pair_batches.insert(g->pair_batches.end(), L.begin() + k, L.begin() + e);
        g->vwork.push_back(make_int4(I, J, begin, (int)g->pair_batches.size()));
      }
    }
  // Cached-AO plan: slots are the entries of bt_list.
  const size_t nslot = g->bt_list.size();
  g->slot_batch.assign(nslot, 0);
  g->slot_tile.assign(nslot, 0);
  g->slot_of.assign((size_t)nb * nt, -1);
  for (int b = 0; b < nb; ++b)
    for (int it = g->bt_off[b]; it < g->bt_off[b + 1]; ++it) {
      g->slot_batch[it] = b;
      g->slot_tile[it] = g->bt_list[it];
      g->slot_of[(size_t)b * nt + g->bt_list[it]] = it;
    }
  g->tile_nao = tile_nao;
  g->grp_slots.assign(g->grp_batch.size() * kGroup, 0);
  for (size_t k = 0; k < g->grp_batch.size(); ++k)
    for (int q = 0; q < g->grp_ntiles[k]; ++q)
      g->grp_slots[k * kGroup + q] = g->slot_of[(size_t)g->grp_batch[k] * nt + g->grp_tiles[k * kGroup + q]];
  // The mixed form always runs from the split operand cache (nT components
  // in two layouts per slot, 2 bytes per element) plus B^T of nT components.
  const bool mixed = g->precision == CHX_PREC_MIXED;
  const int nT = g->type >= 2 ? 4 : 1;
  const double cache_bytes =
      mixed ? (double)nslot * 3.0 * nT * kMxComp * sizeof(unsigned short)
            : (double)nslot * (4.0 + 1.0) * kBatch * kTile * sizeof(double);  // phi, grad phi and B
  if (mixed && g->ntiles > kMxMaxTiles)
    throw Error(CHX_ERR_UNSUPPORTED, "the mixed-precision form takes at most " + std::to_string(kMxMaxTiles * kTile) +
                                         " basis functions");
  if (mixed && g->max_nprim > kSegPrim)
    throw Error(CHX_ERR_UNSUPPORTED, "the mixed-precision form takes at most " + std::to_string(kSegPrim) +
                                         " primitives per shell");
  if (mixed) {  // a previous FP64 plan's tables are not needed any more
    g->d_cache.release();
    g->d_bmat.release();
  }
  size_t free_b = 0, total_b = 0;
  CHX_HIP(hipMemGetInfo(&free_b, &total_b));
  if (mixed && cache_bytes > 0.9 * (double)free_b)
    throw Error(CHX_ERR_DEVICE, "the mixed-precision form needs " + std::to_string(cache_bytes / 1e9) +
                                    " GB of device memory for its AO cache");
  g->cached = mixed || g->cache_mode == 1 || (g->cache_mode == -1 && cache_bytes < 0.8 * (double)free_b);
  // large bases: Vxc items over tile groups of batch chunks, 256 x 256 blocks
  g->mx_wide = mixed && (g->vxc_shape == 1 || (g->vxc_shape < 0 && g->npad >= 1024)) ? 1 : 0;
  if (g->cached) {
    build_cwork(g);
    g->d_slot_batch.upload(g->slot_batch, st); g->d_slot_tile.upload(g->slot_tile, st);
    g->d_slot_of.upload(g->slot_of, st); g->d_tile_nao.upload(g->tile_nao, st);
    g->d_grp_slots.upload(g->grp_slots, st);
    g->d_cwork.upload(g->cwork, st); g->d_cpair_batches.upload(g->cpair_batches, st);
    g->d_Mint.alloc((size_t)g->npad * g->npad);
    if (mixed) {
      g->d_cache.release();
      g->d_bmat.release();
      g->mx_nT = nT;
      g->mx_sstride = 2 * nT * kMxComp;  // AOg then AOt, nT components each
      // one extra all-zero slot at index nslot stands in for tiles that are not alive
      g->d_cmx.alloc((nslot + 1) * (size_t)g->mx_sstride);
      CHX_HIP(hipMemsetAsync(g->d_cmx.p + nslot * (size_t)g->mx_sstride, 0,
                             sizeof(unsigned short) * g->mx_sstride, st));
      g->bmx_comps = nT;
      // B^T of every slot, plus the all-zero slot that make_b rewrites for the
      // component count of each build's potential
      g->d_bmx.alloc((nslot + 1) * (size_t)nT * kMxComp);
      g->d_Dsplit.alloc(2 * (size_t)g->npad * g->npad);
      // core batches: any real point within core_radius of a nucleus with Z > 2
      std::vector<int> core_off(nb, -1), slot_core(nslot, -1);
      g->core_grps.clear();
      g->main_grps.clear();
      g->ncore_batches = 0;
      g->core_flops = 0.0;
      int ncore_slots = 0;
      if (g->core_radius > 0) {
        const double r2c = g->core_radius * g->core_radius;  // <= 1 bohr: the hash's reach
        const HeavyHash hh(g->mol);
        for (int b = 0; b < nb; ++b) {
          bool core = false;
          for (int q = 0; q < g->batch_n[b] && !core; ++q) {
            const size_t p = (size_t)b * kBatch + q;
            core = hh.near(g->hx[p], g->hy[p], g->hz[p], r2c) >= 0;
          }
          if (!core) {
            for (int k = g->grp_off[b]; k < g->grp_off[b + 1]; ++k) g->main_grps.push_back(k);
            continue;
          }
          ++g->ncore_batches;
          g->core_flops += 2.0 * g->batch_work[b];
          core_off[b] = ncore_slots;
          for (int it = g->bt_off[b]; it < g->bt_off[b + 1]; ++it) slot_core[it] = ncore_slots++;
          for (int k = g->grp_off[b]; k < g->grp_off[b + 1]; ++k) g->core_grps.push_back(k);
        }
      }
      static const std::vector<int> neg1{-1}, zero1{0};  // placeholders outlive the asynchronous copies
      g->d_core_off.upload(core_off.empty() ? neg1 : core_off, st);
      g->d_slot_core.upload(slot_core.empty() ? neg1 : slot_core, st);
      g->d_core_grps.upload(g->core_grps.empty() ? zero1 : g->core_grps, st);
      if (g->core_grps.empty()) g->main_grps.clear();  // no core batches: the launch takes every group
      g->d_main_grps.upload(g->main_grps.empty() ? zero1 : g->main_grps, st);
      g->d_cmx3.alloc(std::max<size_t>(1, (size_t)ncore_slots * nT * kBatch * kTile));
      launch_ao_cache_mx(st, (int)nslot, g->d_slot_batch.p, g->d_slot_tile.p, g->d_tile_nao.p, g->d_x.p, g->d_y.p,
                         g->d_z.p, g->view(), nT, g->mx_sstride, g->d_cmx.p, ncore_slots ? g->d_slot_core.p : nullptr,
                         g->d_cmx3.p);
      CHX_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
    } else {
      g->d_cmx.release();
      g->d_bmx.release();
      g->d_Dsplit.release();
      g->d_Dres.release();
      g->d_cpart.release();
      g->mx_nT = g->mx_sstride = g->bmx_comps = 0;
      g->d_cache.alloc(std::max<size_t>(1, nslot * (size_t)4 * kBatch * kTile));
      g->d_bmat.alloc(std::max<size_t>(1, nslot * (size_t)kBatch * kTile));
      launch_ao_cache(st, (int)nslot, g->d_slot_batch.p, g->d_slot_tile.p, g->d_tile_nao.p, g->d_x.p, g->d_y.p,
                      g->d_z.p, g->view(), g->d_cache.p);
    }
  } else {
    g->d_cache.release();
    g->d_bmat.release();
    g->d_Mint.release();
    g->d_cmx.release();
    g->d_bmx.release();
    g->d_Dsplit.release();
    g->d_Dres.release();
    g->d_cpart.release();
  }
  g->d_bt_off.upload(g->bt_off, st); g->d_bt_list.upload(g->bt_list, st);
  g->d_grp_batch.upload(g->grp_batch, st); g->d_grp_tiles.upload(g->grp_tiles, st);
  g->d_grp_ntiles.upload(g->grp_ntiles, st); g->d_grp_off.upload(g->grp_off, st);
  g->d_vwork.upload(g->vwork, st); g->d_pair_batches.upload(g->pair_batches, st);
  // per-iteration buffers
  const size_t nslots = (size_t)nb * kBatch;
  g->d_Dint.alloc((size_t)g->npad * g->npad);
  g->d_Vint.alloc((size_t)g->npad * g->npad);
  g->d_Din.alloc((size_t)g->nbasis * g->nbasis);
  g->d_Vout.alloc((size_t)g->nbasis * g->nbasis);
  g->d_part.alloc(std::max<size_t>(1, g->grp_batch.size() * 4 * kBatch));
  g->d_rho.alloc(std::max<size_t>(1, nslots));
  g->d_coef.alloc(std::max<size_t>(kCoef, kCoef * nslots));
  g->d_part_tau.alloc(g->type >= 2 ? std::max<size_t>(1, g->grp_batch.size() * 3 * kBatch) : 1);
  g->en_blocks = (int)std::max<size_t>(1, (nslots + 255) / 256);
  g->d_en_part.alloc(2 * (size_t)g->en_blocks);
  g->d_en.alloc(3);
  launch_fill(st, g->d_en.p, 3, 0.0);
  if (mixed) {  // split remainder of D per spin, and the partial sums of V . remainder
    g->d_Dres.alloc(2 * (size_t)g->nbasis * g->nbasis);
    g->d_cpart.alloc(2 * (size_t)corr_blocks(g->nbasis));
  }
  if (mixed) {  // descriptor of this grid for the mixed kernels
    MolDev& m = g->hmol;
    m = MolDev{};
    m.grp_batch = g->d_grp_batch.p; m.grp_tiles = g->d_grp_tiles.p; m.grp_ntiles = g->d_grp_ntiles.p;
    m.bt_off = g->d_bt_off.p; m.bt_list = g->d_bt_list.p;
    m.bx = g->d_x.p; m.by = g->d_y.p; m.bz = g->d_z.p;
    m.bv = g->view();
    m.cache = g->d_cmx.p; m.sstride = g->mx_sstride; m.toff = g->mx_nT * kMxComp;
    m.Ds = g->d_Dsplit.p; m.npad = g->npad; m.ntiles = g->ntiles; m.nslots = (int)g->bt_list.size();
    m.part = g->d_part.p; m.part_tau = g->d_part_tau.p;
    m.slot_batch = g->d_slot_batch.p; m.slot_tile = g->d_slot_tile.p;
    m.coef = g->d_coef.p; m.bt = g->d_bmx.p;
    m.slot_of = g->d_slot_of.p;
    m.Mint = g->d_Mint.p;
    m.cache3 = g->d_cmx3.p;
    m.core_off = g->d_core_off.p;
    m.core_grps = g->d_core_grps.p;
    m.main_grps = g->core_grps.empty() ? nullptr : g->d_main_grps.p;
    upload_mx_plan(g, st);
  }
  CHX_HIP(hipStreamSynchronize(st));
  g->ao_ready = true;
  ++g->plan_gen;
  g->have_density = g->have_eval = g->have_v = g->have_skeleton = false;
  g->corr_pending = g->corr_done = g->vout_ready = false;
  g->slot_grp.clear();
  g->slot_q.clear();
}

}  // namespace chx

int chx_lebedev_rule_impl(int n, double* xyzw) {
  const auto& r = chx::lebedev_rule(n);
  for (size_t i = 0; i < r.size(); ++i)
    for (int k = 0; k < 4; ++k) xyzw[4 * i + k] = r[i][k];
  return (int)r.size();
}

chx::BasisView chx_grid::view() const {
  chx::BasisView v;
  v.sh_xyz = d_sh_xyz.p;
  v.sh_atom = d_sh_atom.p;
  v.sh_type = d_sh_type.p;
  v.sh_nprim = d_sh_nprim.p;
  v.sh_poff = d_sh_poff.p;
  v.p_exp = d_p_exp.p;
  v.p_coef = d_p_coef.p;
  v.tile_seg_off = d_tile_seg_off.p;
  v.seg = d_seg.p;
  v.segrec = d_segrec.p;
  return v;
}
