// Host side of the density-fitted Coulomb engine (chx_jfit_*): primitive
// pairs and auxiliary primitives laid out for the kernels in coulomb.hip,
// the Boys table, and the C entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "coulomb.hpp"

namespace chx {
std::string& last_error();
void launch_fill(hipStream_t s, double* p, size_t n, double v);
void launch_jherm(hipStream_t s, const JDev& J, const double* D);
void launch_jgamma(hipStream_t s, const JDev& J, const int* fn_slot0, const int* fn_nprim, const int* fn_l,
                   const int* fn_m, double* gamma);
void launch_jcoulomb(hipStream_t s, const JDev& J, const int* slot_l, const double* cfit, double* gco, double* Jm);
void launch_joneint(hipStream_t s, const JDev& J, const double* bexp, double* S, double* T, double* V);
void launch_jmetric(hipStream_t s, const JDev& J, double* Vm);
}  // namespace chx

using namespace chx;

struct chx_jfit {
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, naux = 0, natoms = 0;
  long npairs = 0, npairs_raw = 0;
  double pair_thr = 1e-15;
  JDev J{};
  DevBuf<double> px, py, pz, p, K, pax, pay, paz, pbx, pby, pbz, bexp, h;
  DevBuf<int> ta, tb, aoa, aob;
  DevBuf<unsigned char> diag;
  DevBuf<double> ae, ax, ay, az, acoef, auxS, orbT, boys, nuc;
  DevBuf<int> afn, slot_l, fn_slot0, fn_nprim, fn_l, fn_m;
  DevBuf<int4> chunks;
  DevBuf<double> gpart, gco, dmat, jmat, gamma, cfit;
  ~chx_jfit() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <class F>
int jguard(F&& f) {
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

int lval(int type) { return type == 1 ? 1 : std::abs(type); }

// F_n(T) by its series F_n(T) = exp(-T) sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)),
// in long double (every term is positive, so the sum carries no cancellation).
long double boys_series(int n, long double T) {
  long double term = 1.0L / (2 * n + 1), sum = term;
  for (int k = 1; k < 2000; ++k) {
    term *= 2 * T / (2 * n + 2 * k + 1);
    sum += term;
    if (term < 1e-22L * sum) break;
  }
  return std::exp(-T) * sum;
}

// Real solid harmonics of degree l as Cartesian polynomials: the m = 0, cos and
// sin families, rows [m][cart] with cart in the jcidx order. Any basis of the
// degree-l harmonic polynomials spans the same fitting space; these are the
// standard ones (Helgaker, Jorgensen, Olsen, eq. 6.4.47-6.4.48, up to
// normalisation).
void solid_harmonics(int l, double* S /* 9 x 15 */) {
  std::fill(S, S + 9 * kJCart, 0.0);
  auto binom = [](int a, int b) {
    if (b < 0 || b > a) return 0.0;
    double r = 1;
    for (int i = 1; i <= b; ++i) r = r * (a - b + i) / i;
    return r;
  };
  int row = 0;
  for (int m = -l; m <= l; ++m, ++row) {
    const int am = std::abs(m);
    const double vm = m < 0 ? 0.5 : 0.0;
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u) {
        // v runs over vm, vm + 1, ... up to floor(|m|/2 - vm) + vm
        const int vmax2 = (int)std::floor(am / 2.0 - vm) ;
        for (int vi = 0; vi <= vmax2; ++vi) {
          const double v = vi + vm;
          const double sign = ((t + (int)std::lround(v - vm)) & 1) ? -1.0 : 1.0;
          const double c = sign * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) *
                           binom(am, (int)std::lround(2 * v));
          const int ex = 2 * t + am - (int)std::lround(2 * (u + v)), ey = (int)std::lround(2 * (u + v)),
                    ez = l - 2 * t - am;
          if (ex < 0 || ey < 0 || ez < 0) continue;
          S[row * kJCart + jcidx(ex, ey, ez)] += c;
        }
      }
  }
}

}  // namespace

extern "C" {

int chx_jfit_create(const chx_mol* m, const char* data_root, const char* aux_basis, chx_jfit** out) {
  return jguard([&] {
    if (!m || !data_root || !aux_basis || !out) throw Error(CHX_ERR_ARG, "bad arguments");
    if (m->nbasis == 0) throw Error(CHX_ERR_STATE, "the molecule has no basis set");
    for (const ShellH& sh : m->shells)
      if (lval(sh.type) > kJMaxLOrb)
        throw Error(CHX_ERR_UNSUPPORTED, "the Coulomb engine takes orbital shells up to d");
    std::unique_ptr<chx_jfit> j(new chx_jfit);
    CHX_HIP(hipGetDevice(&j->device));
    CHX_HIP(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking));
    hipStream_t st = j->stream;
    j->n = m->nbasis;
    j->natoms = m->natoms;

    // ---- orbital primitive pairs (shell a >= shell b, la >= lb after a swap)
    struct Pr {
      int L, la;
      double px, py, pz, p, K, pa[3], pb[3], b;
      int ta, tb, aoa, aob;
      unsigned char diag;
    };
    std::vector<Pr> prs;
    const int ns = (int)m->shells.size();
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b <= a; ++b) {
        const ShellH* A = &m->shells[a];
        const ShellH* B = &m->shells[b];
        if (lval(A->type) < lval(B->type)) std::swap(A, B);
        const double* ca = &m->xyz[3 * A->center];
        const double* cb = &m->xyz[3 * B->center];
        const double r2 = (ca[0] - cb[0]) * (ca[0] - cb[0]) + (ca[1] - cb[1]) * (ca[1] - cb[1]) +
                          (ca[2] - cb[2]) * (ca[2] - cb[2]);
        const int la = lval(A->type), lb = lval(B->type);
        for (int pa = 0; pa < A->nprim; ++pa)
          for (int pb = 0; pb < B->nprim; ++pb) {
            ++j->npairs_raw;
            const double ea = m->exps[A->prim_off + pa], eb = m->exps[B->prim_off + pb];
            const double p = ea + eb, mu = ea * eb / p;
            const double K = m->coefs[A->prim_off + pa] * m->coefs[B->prim_off + pb] * std::exp(-mu * r2);
            Pr q{};
            for (int d = 0; d < 3; ++d) {
              const double P = (ea * ca[d] + eb * cb[d]) / p;
              q.pa[d] = P - ca[d];
              q.pb[d] = P - cb[d];
            }
            // screening: the pair's overlap-like size |K| (pi/p)^3/2 with a
            // polynomial allowance for the angular factors
            const double ra = std::sqrt(q.pa[0] * q.pa[0] + q.pa[1] * q.pa[1] + q.pa[2] * q.pa[2]);
            const double rb = std::sqrt(q.pb[0] * q.pb[0] + q.pb[1] * q.pb[1] + q.pb[2] * q.pb[2]);
            const double size = std::abs(K) * std::pow(M_PI / p, 1.5) * std::pow(1.0 + ra + 1.0 / std::sqrt(p), la) *
                                std::pow(1.0 + rb + 1.0 / std::sqrt(p), lb);
            if (size < j->pair_thr) continue;
            q.L = la + lb;
            q.la = la;
            q.px = q.pa[0] + ca[0];
            q.py = q.pa[1] + ca[1];
            q.pz = q.pa[2] + ca[2];
            q.p = p;
            q.K = K;
            q.b = eb;
            q.ta = A->type;
            q.tb = B->type;
            q.aoa = A->ao_off;
            q.aob = B->ao_off;
            q.diag = a == b ? 1 : 0;
            prs.push_back(q);
          }
      }
    std::stable_sort(prs.begin(), prs.end(), [](const Pr& x, const Pr& y) { return x.L != y.L ? x.L < y.L : x.la > y.la; });
    const long np = (long)prs.size();
    j->npairs = np;
    if (np >= (long)INT32_MAX) throw Error(CHX_ERR_UNSUPPORTED, "too many primitive pairs");
    {
      std::vector<double> vpx(np), vpy(np), vpz(np), vp(np), vK(np), va[3], vb[3], vbe(np);
      for (int d = 0; d < 3; ++d) {
        va[d].resize(np);
        vb[d].resize(np);
      }
      std::vector<int> vta(np), vtb(np), vaa(np), vab(np);
      std::vector<unsigned char> vdg(np);
      for (long k = 0; k < np; ++k) {
        const Pr& q = prs[k];
        vpx[k] = q.px; vpy[k] = q.py; vpz[k] = q.pz; vp[k] = q.p; vK[k] = q.K; vbe[k] = q.b;
        for (int d = 0; d < 3; ++d) {
          va[d][k] = q.pa[d];
          vb[d][k] = q.pb[d];
        }
        vta[k] = q.ta; vtb[k] = q.tb; vaa[k] = q.aoa; vab[k] = q.aob; vdg[k] = q.diag;
      }
      j->px.upload(vpx, st); j->py.upload(vpy, st); j->pz.upload(vpz, st); j->p.upload(vp, st); j->K.upload(vK, st);
      j->pax.upload(va[0], st); j->pay.upload(va[1], st); j->paz.upload(va[2], st);
      j->pbx.upload(vb[0], st); j->pby.upload(vb[1], st); j->pbz.upload(vb[2], st);
      j->bexp.upload(vbe, st);
      j->ta.upload(vta, st); j->tb.upload(vtb, st); j->aoa.upload(vaa, st); j->aob.upload(vab, st);
      j->diag.upload(vdg, st);
      CHX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    }
    JDev& J = j->J;
    J.px = j->px.p; J.py = j->py.p; J.pz = j->pz.p; J.p = j->p.p; J.K = j->K.p;
    J.pax = j->pax.p; J.pay = j->pay.p; J.paz = j->paz.p; J.pbx = j->pbx.p; J.pby = j->pby.p; J.pbz = j->pbz.p;
    J.ta = j->ta.p; J.tb = j->tb.p; J.aoa = j->aoa.p; J.aob = j->aob.p; J.diag = j->diag.p;
    J.n = j->n;
    {
      int c = 0;
      for (int L = 0; L <= 2 * kJMaxLOrb; ++L) {
        J.cls[L] = c;
        while (c < np && prs[c].L == L) ++c;
      }
      J.cls[2 * kJMaxLOrb + 1] = (int)np;
      long off = 0;
      for (int L = 0; L <= 2 * kJMaxLOrb; ++L) {
        J.hoff[L] = off;
        off += (long)(J.cls[L + 1] - J.cls[L]) * jnherm(L);
      }
      j->h.alloc(std::max<long>(1, off));
      J.h = j->h.p;
      // gamma-pass rows: chunks of pairs of one class, about kJChunkRows in all
      std::vector<int4> ch;
      const long len = std::max<long>(64, (np + kJChunkRows - 1) / kJChunkRows);
      for (int L = 0; L <= 2 * kJMaxLOrb; ++L)
        for (long k = J.cls[L]; k < J.cls[L + 1]; k += len)
          ch.push_back(make_int4((int)k, (int)std::min<long>(J.cls[L + 1], k + len), L, 0));
      if (ch.empty()) ch.push_back(make_int4(0, 0, 0, 0));
      j->chunks.upload(ch, st);
      CHX_HIP(hipStreamSynchronize(st));
      J.chunks = j->chunks.p;
      J.nrows = np > 0 ? (int)ch.size() : 0;
    }

    // ---- auxiliary primitives, sorted by l, each class padded to 64 slots
    const std::vector<AuxShell> aux = aux_basis_shells(m, data_root, aux_basis);
    std::vector<int> shell_fn(aux.size());
    int nf = 0;
    for (size_t s = 0; s < aux.size(); ++s) {
      shell_fn[s] = nf;
      nf += 2 * aux[s].l + 1;
    }
    j->naux = nf;
    J.naux = nf;
    std::vector<double> vae, vax, vay, vaz, vco;
    std::vector<int> vfn, vsl, fslot0(nf), fnp(nf), fl(nf), fm(nf);
    for (int l = 0; l <= kJMaxLAux; ++l) {
      J.aoff[l] = (int)vae.size();
      int cnt = 0;
      for (size_t s = 0; s < aux.size(); ++s) {
        if (aux[s].l != l) continue;
        const int first = (int)vae.size();
        for (size_t q = 0; q < aux[s].e.size(); ++q) {
          const double* c = &m->xyz[3 * aux[s].center];
          const double e = aux[s].e[q];
          vae.push_back(e);
          vax.push_back(c[0]); vay.push_back(c[1]); vaz.push_back(c[2]);
          vco.push_back(((l & 1) ? -1.0 : 1.0) * aux[s].c[q] * std::pow(2.0 * e, -l));
          vfn.push_back(shell_fn[s]);
          vsl.push_back(l);
          ++cnt;
        }
        for (int mm = 0; mm < 2 * l + 1; ++mm) {
          fslot0[shell_fn[s] + mm] = first;
          fnp[shell_fn[s] + mm] = (int)aux[s].e.size();
          fl[shell_fn[s] + mm] = l;
          fm[shell_fn[s] + mm] = mm;
        }
      }
      J.anum[l] = cnt;
      while (vae.size() % 64) {  // padding: harmless primitive, never summed
        vae.push_back(1.0); vax.push_back(0.0); vay.push_back(0.0); vaz.push_back(0.0);
        vco.push_back(0.0); vfn.push_back(0); vsl.push_back(-1);
      }
    }
    J.aoff[kJMaxLAux + 1] = (int)vae.size();
    J.nslots = (int)vae.size();
    j->ae.upload(vae, st); j->ax.upload(vax, st); j->ay.upload(vay, st); j->az.upload(vaz, st);
    j->acoef.upload(vco, st); j->afn.upload(vfn, st); j->slot_l.upload(vsl, st);
    j->fn_slot0.upload(fslot0, st); j->fn_nprim.upload(fnp, st); j->fn_l.upload(fl, st); j->fn_m.upload(fm, st);
    J.ae = j->ae.p; J.ax = j->ax.p; J.ay = j->ay.p; J.az = j->az.p; J.acoef = j->acoef.p; J.afn = j->afn.p;

    // ---- tables: harmonic polynomials, orbital component polynomials, Boys function
    std::vector<double> S((kJMaxLAux + 1) * 9 * kJCart);
    for (int l = 0; l <= kJMaxLAux; ++l) solid_harmonics(l, &S[l * 9 * kJCart]);
    j->auxS.upload(S, st);
    std::vector<double> T(4 * 30, 0.0);
    const int types[4] = {0, 1, -1, -2};
    for (int ti = 0; ti < 4; ++ti) {
      const auto comp = component_monomials(types[ti]);
      for (size_t c = 0; c < comp.size(); ++c)
        for (const MonoTerm& t : comp[c]) T[ti * 30 + c * 6 + jcidx(t.i, t.j, t.k)] += t.c;
    }
    j->orbT.upload(T, st);
    std::vector<double> B((size_t)kBoysPts * kBoysOrders);
    for (int k = 0; k < kBoysPts; ++k)
      for (int n = 0; n < kBoysOrders; ++n) B[(size_t)k * kBoysOrders + n] = (double)boys_series(n, k * (long double)kBoysStep);
    j->boys.upload(B, st);
    std::vector<double> nuc(4 * m->natoms);
    for (int a = 0; a < m->natoms; ++a) {
      for (int d = 0; d < 3; ++d) nuc[4 * a + d] = m->xyz[3 * a + d];
      nuc[4 * a + 3] = m->Z[a];
    }
    j->nuc.upload(nuc, st);
    J.auxS = j->auxS.p; J.orbT = j->orbT.p; J.boys = j->boys.p; J.nuc = j->nuc.p; J.natoms = m->natoms;

    // ---- per-call buffers
    j->gpart.alloc(std::max<size_t>(1, (size_t)J.nrows * J.nslots * kJCart));
    J.gpart = j->gpart.p;
    j->gco.alloc((size_t)J.nslots * kJCart);
    J.gco = j->gco.p;
    j->dmat.alloc((size_t)j->n * j->n);
    j->jmat.alloc((size_t)j->n * j->n);
    j->gamma.alloc(std::max(1, nf));
    j->cfit.alloc(std::max(1, nf));
    CHX_HIP(hipStreamSynchronize(st));
    *out = j.release();
  });
}

int chx_jfit_info(const chx_jfit* j, int* naux, long* npairs, long* npairs_unscreened, int* naux_prims) {
  return jguard([&] {
    if (!j) throw Error(CHX_ERR_ARG, "bad arguments");
    if (naux) *naux = j->naux;
    if (npairs) *npairs = j->npairs;
    if (npairs_unscreened) *npairs_unscreened = j->npairs_raw;
    if (naux_prims) {
      int c = 0;
      for (int l = 0; l <= kJMaxLAux; ++l) c += j->J.anum[l];
      *naux_prims = c;
    }
  });
}

int chx_jfit_int1e(chx_jfit* j, double* S, double* T, double* V) {
  return jguard([&] {
    if (!j || !S || !T || !V) throw Error(CHX_ERR_ARG, "bad arguments");
    CHX_HIP(hipSetDevice(j->device));
    const size_t nn = (size_t)j->n * j->n;
    DevBuf<double> dS, dT, dV;
    dS.alloc(nn); dT.alloc(nn); dV.alloc(nn);
    launch_fill(j->stream, dS.p, nn, 0.0);
    launch_fill(j->stream, dT.p, nn, 0.0);
    launch_fill(j->stream, dV.p, nn, 0.0);
    launch_joneint(j->stream, j->J, j->bexp.p, dS.p, dT.p, dV.p);
    CHX_HIP(hipMemcpyAsync(S, dS.p, nn * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    CHX_HIP(hipMemcpyAsync(T, dT.p, nn * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    CHX_HIP(hipMemcpyAsync(V, dV.p, nn * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    CHX_HIP(hipStreamSynchronize(j->stream));
  });
}

int chx_jfit_metric(chx_jfit* j, double* Vpq) {
  return jguard([&] {
    if (!j || !Vpq) throw Error(CHX_ERR_ARG, "bad arguments");
    CHX_HIP(hipSetDevice(j->device));
    const size_t na2 = (size_t)j->naux * j->naux;
    DevBuf<double> dV;
    dV.alloc(na2);
    launch_fill(j->stream, dV.p, na2, 0.0);
    launch_jmetric(j->stream, j->J, dV.p);
    CHX_HIP(hipMemcpyAsync(Vpq, dV.p, na2 * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    CHX_HIP(hipStreamSynchronize(j->stream));
  });
}

int chx_jfit_gamma(chx_jfit* j, const double* D, double* gamma) {
  return jguard([&] {
    if (!j || !D || !gamma) throw Error(CHX_ERR_ARG, "bad arguments");
    CHX_HIP(hipSetDevice(j->device));
    const size_t nn = (size_t)j->n * j->n;
    CHX_HIP(hipMemcpyAsync(j->dmat.p, D, nn * sizeof(double), hipMemcpyHostToDevice, j->stream));
    launch_jherm(j->stream, j->J, j->dmat.p);
    launch_jgamma(j->stream, j->J, j->fn_slot0.p, j->fn_nprim.p, j->fn_l.p, j->fn_m.p, j->gamma.p);
    CHX_HIP(hipMemcpyAsync(gamma, j->gamma.p, j->naux * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    CHX_HIP(hipStreamSynchronize(j->stream));
  });
}

int chx_jfit_coulomb(chx_jfit* j, const double* cfit, double* Jm) {
  return jguard([&] {
    if (!j || !cfit || !Jm) throw Error(CHX_ERR_ARG, "bad arguments");
    CHX_HIP(hipSetDevice(j->device));
    const size_t nn = (size_t)j->n * j->n;
    CHX_HIP(hipMemcpyAsync(j->cfit.p, cfit, j->naux * sizeof(double), hipMemcpyHostToDevice, j->stream));
    launch_fill(j->stream, j->jmat.p, nn, 0.0);
    launch_jcoulomb(j->stream, j->J, j->slot_l.p, j->cfit.p, j->gco.p, j->jmat.p);
    CHX_HIP(hipMemcpyAsync(Jm, j->jmat.p, nn * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    CHX_HIP(hipStreamSynchronize(j->stream));
  });
}

void chx_jfit_destroy(chx_jfit* j) { delete j; }

}  // extern "C"
