// C++ drop-in test: the reference-shaped API of include/mpstools_amd.hpp
// (contractTwosite, contractSingle, addContractionLayer, TSLanzcosEngine) on
// the GPU against the CPU oracle on the same seeded XXZ chain.  Prints one
// PASS/FAIL line per check; exit status 0 only if all pass.
#define MPSTOOLS_AMD_THROW
#include "../../include/mpstools_amd.hpp"
#include <complex>

#include "../../oracle/oracle.hpp"

#include <cstdio>
#include <random>

using namespace mps;
typedef double Real;

static int g_fail = 0;
static void report(const char* name, double err, double tol) {
  const bool ok = err <= tol;
  std::printf("%s %s err=%.3e tol=%.1e\n", ok ? "PASS" : "FAIL", name, err, tol);
  if (!ok) ++g_fail;
}

// HeisenbergMPO bulk / boundary tensors (src/Hamiltonians.hpp:86-133), J=Jz=1, B=0.
static MPOMat<1, Real> heis(int site, int N) {
  MPOMat<1, Real> M;
  auto put = [&](uint wl, uint wr, uint out, uint in, Real v) { M[ukey<2>(wl, wr)][OpKeyType<1>(out, in)] = v; };
  auto Sz = [&](uint wl, uint wr, Real f) { put(wl, wr, 0, 0, -0.5 * f); put(wl, wr, 1, 1, 0.5 * f); };
  auto Id = [&](uint wl, uint wr) { put(wl, wr, 0, 0, 1); put(wl, wr, 1, 1, 1); };
  if (site == 0) {
    put(0, 1, 0, 1, 0.5); put(0, 2, 1, 0, 0.5); Sz(0, 3, 1.0); Id(0, 4);
    M.Dl = 1; M.Dr = 5;
  } else if (site == N - 1) {
    Id(0, 0); put(1, 0, 1, 0, 1); put(2, 0, 0, 1, 1); Sz(3, 0, 1.0);
    M.Dl = 5; M.Dr = 1;
  } else {
    Id(0, 0); put(1, 0, 1, 0, 1); put(2, 0, 0, 1, 1); Sz(3, 0, 1.0);
    put(4, 1, 0, 1, 0.5); put(4, 2, 1, 0, 0.5); Sz(4, 3, 1.0); Id(4, 4);
    M.Dl = 5; M.Dr = 5;
  }
  return M;
}

static oracle::MPOMat<double> to_oracle(const MPOMat<1, Real>& W) {
  oracle::MPOMat<double> o;
  o.Dl = W.Dl;
  o.Dr = W.Dr;
  for (auto& g : W)
    for (auto& e : g.second) o.m[oracle::UKey{g.first[0], g.first[1]}][oracle::OpKey{e.first.in(), e.first.out()}] = e.second;
  return o;
}
template <int Ns>
static oracle::BlockMatrix<Ns, double> to_oracle(const BlockMatrix<Ns, AbKeyType, Real>& B) {
  oracle::BlockMatrix<Ns, double> o;
  for (auto& kv : B) {
    uint32_t p[Ns];
    for (int s = 0; s < Ns; ++s) p[s] = kv.first(s);
    oracle::Mat<double> m((int64_t)kv.second.size1(), (int64_t)kv.second.size2());
    std::memcpy(m.a.data(), kv.second.data(), sizeof(double) * m.a.size());
    o[oracle::mk<Ns>(kv.first[0].getQ(), kv.first[1].getQ(), p)] = m;
  }
  return o;
}
template <int Ns>
static double rel(const BlockMatrix<Ns, AbKeyType, Real>& a, const oracle::BlockMatrix<Ns, double>& b) {
  if (a.size() != b.size()) return 1e300;
  double num = 0, den = 0;
  for (auto& kv : b) {
    uint32_t p[Ns];
    for (int s = 0; s < Ns; ++s) p[s] = kv.first.p[s];
    std::vector<uint> pv(p, p + Ns);
    TensorKeyType<2, Ns, AbKeyType> k;
    k[0] = AbKeyType(kv.first.q[0], 1);
    k[1] = AbKeyType(kv.first.q[1], 1);
    for (int s = 0; s < Ns; ++s) k(s) = p[s];
    auto it = a.find(k);
    if (it == a.end()) return 1e300;
    for (size_t i = 0; i < kv.second.a.size(); ++i) {
      const double d = it->second.data()[i] - kv.second.a[i];
      num += d * d;
      den += kv.second.a[i] * kv.second.a[i];
    }
  }
  return std::sqrt(num / std::max(den, 1e-300));
}

// complex128 counterparts: the same blocks with seeded imaginary parts, the
// oracle's complex template as the reference (its complex path is pinned
// against real contractions by tests/test_complex.py)
typedef std::complex<double> Cplx;
static MPOMat<1, Cplx> cplx_mpo(const MPOMat<1, Real>& W) {
  MPOMat<1, Cplx> o;
  o.Dl = W.Dl;
  o.Dr = W.Dr;
  for (auto& g : W)
    for (auto& e : g.second) o[g.first][e.first] = Cplx(e.second, 0.0);
  return o;
}
static oracle::MPOMat<Cplx> to_oracle_c(const MPOMat<1, Real>& W) {
  oracle::MPOMat<Cplx> o;
  o.Dl = W.Dl;
  o.Dr = W.Dr;
  for (auto& g : W)
    for (auto& e : g.second) o.m[oracle::UKey{g.first[0], g.first[1]}][oracle::OpKey{e.first.in(), e.first.out()}] = e.second;
  return o;
}
template <int Ns>
static BlockMatrix<Ns, AbKeyType, Cplx> with_imag(const BlockMatrix<Ns, AbKeyType, Real>& B, std::mt19937_64& rng) {
  std::normal_distribution<double> nd(0.0, 0.5);
  BlockMatrix<Ns, AbKeyType, Cplx> o;
  for (auto& kv : B) {
    matrix<Cplx> m(kv.second.size1(), kv.second.size2());
    for (size_t i = 0; i < m.size1() * m.size2(); ++i)
      m.data()[i] = Cplx(kv.second.data()[i], nd(rng) * std::fabs(kv.second.data()[i]));
    o[kv.first] = m;
  }
  return o;
}
template <int Ns>
static oracle::BlockMatrix<Ns, Cplx> to_oracle_c(const BlockMatrix<Ns, AbKeyType, Cplx>& B) {
  oracle::BlockMatrix<Ns, Cplx> o;
  for (auto& kv : B) {
    uint32_t p[Ns];
    for (int s = 0; s < Ns; ++s) p[s] = kv.first(s);
    oracle::Mat<Cplx> m((int64_t)kv.second.size1(), (int64_t)kv.second.size2());
    for (size_t i = 0; i < m.a.size(); ++i) m.a[i] = kv.second.data()[i];
    o[oracle::mk<Ns>(kv.first[0].getQ(), kv.first[1].getQ(), p)] = m;
  }
  return o;
}
template <int Ns>
static double rel_c(const BlockMatrix<Ns, AbKeyType, Cplx>& a, const oracle::BlockMatrix<Ns, Cplx>& b) {
  if (a.size() != b.size()) return 1e300;
  double num = 0, den = 0;
  for (auto& kv : b) {
    TensorKeyType<2, Ns, AbKeyType> k;
    k[0] = AbKeyType(kv.first.q[0], 1);
    k[1] = AbKeyType(kv.first.q[1], 1);
    for (int s = 0; s < Ns; ++s) k(s) = kv.first.p[s];
    auto it = a.find(k);
    if (it == a.end()) return 1e300;
    for (size_t i = 0; i < kv.second.a.size(); ++i) {
      num += std::norm(it->second.data()[i] - kv.second.a[i]);
      den += std::norm(kv.second.a[i]);
    }
  }
  return std::sqrt(num / std::max(den, 1e-300));
}

int main() {
  const int N = 12, Nup = 6, D = 16;
  std::mt19937_64 rng(20260001);
  std::normal_distribution<double> nd(0.0, 1.0);
  i_to_key<AbKeyType> itk;
  itk[0] = AbKeyType(0, 1);
  itk[1] = AbKeyType(1, 1);
  oracle::ItoKey oitk = {0, 1};
  // bond charges / dims: min of left and right counts, capped at D
  auto binom = [](int n, int k) { double r = 1; if (k < 0 || k > n) return 0.0; for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i; return r; };
  std::vector<std::map<unsigned long, int>> dims((size_t)N + 1);
  for (int l = 0; l <= N; ++l)
    for (int q = 0; q <= l && q <= Nup; ++q) {
      const double m = std::min(binom(l, q), binom(N - l, Nup - q));
      if (m > 0) dims[(size_t)l][(unsigned long)q] = (int)std::min<double>(m, D);
    }
  std::vector<BlockMatrix<1, AbKeyType, Real>> mps((size_t)N);
  for (int s = 0; s < N; ++s)
    for (auto& l : dims[(size_t)s])
      for (uint p = 0; p < 2; ++p) {
        auto r = dims[(size_t)s + 1].find(l.first + p);
        if (r == dims[(size_t)s + 1].end()) continue;
        matrix<Real> m((size_t)l.second, (size_t)r->second);
        for (size_t i = 0; i < m.size1() * m.size2(); ++i) m.data()[i] = nd(rng) / std::sqrt((double)D);
        mps[(size_t)s][TensorKeyType<2, 1, AbKeyType>(AbKeyType(l.first, 1), AbKeyType(r->first, 1), p)] = m;
      }
  const int c = N / 2 - 1;
  // environments on the GPU (drop-in) and on the CPU (oracle), layer by layer
  BlockMatrix<1, AbKeyType, Real> L;
  {
    matrix<Real> one(1, 1);
    one(0, 0) = 1;
    L[TensorKeyType<2, 1, AbKeyType>(AbKeyType(0, 1), AbKeyType(0, 1), 0u)] = one;
  }
  oracle::BlockMatrix<1, double> oL = oracle::getLMin<double>();
  double worst = 0;
  for (int s = 0; s < c; ++s) {
    L = addContractionLayer(L, mps[(size_t)s], heis(s, N), itk, 1);
    oL = oracle::addContractionLayer(oL, to_oracle<1>(mps[(size_t)s]), to_oracle(heis(s, N)), oitk, 1, 1);
    worst = std::max(worst, rel<1>(L, oL));
  }
  report("addContractionLayer left", worst, 1e-12);
  oracle::BlockMatrix<1, double> oR = oracle::getRMax(to_oracle<1>(mps[(size_t)N - 1]));
  BlockMatrix<1, AbKeyType, Real> R;
  for (auto& kv : oR) {
    matrix<Real> m((size_t)kv.second.r, (size_t)kv.second.c);
    std::memcpy(m.data(), kv.second.a.data(), sizeof(double) * kv.second.a.size());
    R[TensorKeyType<2, 1, AbKeyType>(AbKeyType(kv.first.q[0], 1), AbKeyType(kv.first.q[1], 1), 0u)] = m;
  }
  worst = 0;
  for (int s = N - 1; s > c + 1; --s) {
    R = addContractionLayer(R, mps[(size_t)s], heis(s, N), itk, -1);
    oR = oracle::addContractionLayer(oR, to_oracle<1>(mps[(size_t)s]), to_oracle(heis(s, N)), oitk, -1, 1);
    worst = std::max(worst, rel<1>(R, oR));
  }
  report("addContractionLayer right", worst, 1e-12);
  // mixed variant
  {
    auto Lp = addContractionLayer(L, mps[(size_t)c], heis(c, N), mps[(size_t)c], itk, 1);
    auto oLp = oracle::addContractionLayer(oL, to_oracle<1>(mps[(size_t)c]), to_oracle(heis(c, N)), oitk, 1, 1);
    report("addContractionLayer Ap", rel<1>(Lp, oLp), 1e-12);
  }
  // two-site psi and H_eff psi
  BlockMatrix<2, AbKeyType, Real> psi;
  for (auto& a : mps[(size_t)c])
    for (auto& b : mps[(size_t)c + 1])
      if (a.first[1] == b.first[0]) {
        matrix<Real> m(a.second.size1(), b.second.size2());
        for (size_t j = 0; j < m.size2(); ++j)
          for (size_t k = 0; k < a.second.size2(); ++k)
            for (size_t i = 0; i < m.size1(); ++i) m(i, j) += a.second(i, k) * b.second(k, j);
        psi[TensorKeyType<2, 2, AbKeyType>(a.first[0], b.first[1], std::vector<uint>{a.first(0), b.first(0)})] = m;
      }
  std::vector<i_to_key<AbKeyType>> itks = {itk, itk};
  BlockMatrix<2, AbKeyType, Real> y;
  contractTwosite(y, L, R, psi, heis(c, N), heis(c + 1, N), itks);
  oracle::BlockMatrix<2, double> oy;
  oracle::contractTwosite(oy, oL, oR, to_oracle<2>(psi), to_oracle(heis(c, N)), to_oracle(heis(c + 1, N)), oitk, oitk, 1);
  report("contractTwosite", rel<2>(y, oy), 1e-12);
  // plan cache: a reference-style Lanczos loop (src/lanzcos.hpp:197) calls
  // contractTwosite on fixed BL, BR with a new vector each step; the plan and
  // the device environments are built once, and a changed R is re-uploaded
  {
    const auto s0 = plan_cache_stats();
    BlockMatrix<2, AbKeyType, Real> x = psi, yy;
    double worst_c = 0;
    for (int it = 0; it < 5; ++it) {
      for (auto& kv : x)
        for (size_t i = 0; i < kv.second.size1() * kv.second.size2(); ++i) kv.second.data()[i] *= -1.25;
      contractTwosite(yy, L, R, x, heis(c, N), heis(c + 1, N), itks);
      oracle::BlockMatrix<2, double> oyy;
      oracle::contractTwosite(oyy, oL, oR, to_oracle<2>(x), to_oracle(heis(c, N)), to_oracle(heis(c + 1, N)), oitk,
                              oitk, 1);
      worst_c = std::max(worst_c, rel<2>(yy, oyy));
    }
    const auto s1 = plan_cache_stats();
    report("contractTwosite cached loop", worst_c, 1e-12);
    report("plan cache: no plan built in the loop", (double)(s1.builds - s0.builds), 0.0);
    report("plan cache: every loop call served by the cache", s1.hits - s0.hits >= 5 ? 0.0 : 1.0, 0.0);
    report("plan cache: environments not re-uploaded", (double)(s1.env_uploads - s0.env_uploads), 0.0);
    auto R2 = R;
    R2.begin()->second.data()[0] += 0.5;
    contractTwosite(yy, L, R2, x, heis(c, N), heis(c + 1, N), itks);
    oracle::BlockMatrix<2, double> oyy;
    auto oR2 = to_oracle<1>(R2);
    oracle::contractTwosite(oyy, oL, oR2, to_oracle<2>(x), to_oracle(heis(c, N)), to_oracle(heis(c + 1, N)), oitk, oitk,
                            1);
    const auto s2 = plan_cache_stats();
    report("plan cache: changed R uploaded once", (double)(s2.env_uploads - s1.env_uploads - 1), 0.0);
    report("plan cache: changed R result", rel<2>(yy, oyy), 1e-12);
  }
  // one-site
  {
    auto Rc1 = addContractionLayer(R, mps[(size_t)c + 1], heis(c + 1, N), itk, -1);
    auto oRc1 = oracle::addContractionLayer(oR, to_oracle<1>(mps[(size_t)c + 1]), to_oracle(heis(c + 1, N)), oitk, -1, 1);
    BlockMatrix<1, AbKeyType, Real> y1;
    contractSingle(y1, L, mps[(size_t)c], heis(c, N), Rc1, itk);
    oracle::BlockMatrix<1, double> oy1;
    oracle::contractSingle(oy1, oL, to_oracle<1>(mps[(size_t)c]), to_oracle(heis(c, N)), oRc1, oitk, 1);
    report("contractSingle", rel<1>(y1, oy1), 1e-12);
  }
  // Lanczos engine
  {
    TSLanzcosEngine<AbKeyType, Real> lan(L, R, mps[(size_t)c], mps[(size_t)c + 1], heis(c, N), heis(c + 1, N), itk, itk, 0);
    const int it = lan.simulate(200, 1e-11, true);
    auto mv = [&](oracle::BlockMatrix<2, double>& out, const oracle::BlockMatrix<2, double>& in) {
      oracle::contractTwosite(out, oL, oR, in, to_oracle(heis(c, N)), to_oracle(heis(c + 1, N)), oitk, oitk, 1);
    };
    auto r = oracle::lanczos<double>(oracle::twosite(to_oracle<1>(mps[(size_t)c]), to_oracle<1>(mps[(size_t)c + 1])), mv, 200, 1e-11, true);
    report("TSLanzcosEngine E0", std::fabs(lan.E(0) - r.es[0]), 1e-9);
    std::printf("INFO lanczos iterations gpu=%d oracle=%d\n", it, r.it);
    auto gs = lan.getState(0);
    auto ogs = oracle::getState(r, 0);
    double ov = 0;
    for (auto& kv : ogs) {
      TensorKeyType<2, 2, AbKeyType> k(AbKeyType(kv.first.q[0], 1), AbKeyType(kv.first.q[1], 1),
                                       std::vector<uint>{kv.first.p[0], kv.first.p[1]});
      auto f = gs.find(k);
      if (f == gs.end()) continue;
      for (size_t i = 0; i < kv.second.a.size(); ++i) ov += f->second.data()[i] * kv.second.a[i];
    }
    report("TSLanzcosEngine state overlap", std::fabs(std::fabs(ov) - 1.0), 1e-6);
    // the projected run against that ground state (src/lanzcos.hpp:281-414)
    TSLanzcosEngine<AbKeyType, Real> lp(L, R, mps[(size_t)c], mps[(size_t)c + 1], heis(c, N), heis(c + 1, N), itk, itk, 0);
    lp.simulate(200, 1e-11, true, gs, 0.0);
    const auto ogs_dev = to_oracle<2>(gs);
    auto rp = oracle::lanczos<double>(oracle::twosite(to_oracle<1>(mps[(size_t)c]), to_oracle<1>(mps[(size_t)c + 1])), mv,
                                      200, 1e-11, true, &ogs_dev);
    report("TSLanzcosEngine projected E0", std::fabs(lp.E(0) - rp.es[0]), 1e-9);
    report("TSLanzcosEngine projected above E0", lp.E(0) > lan.E(0) + 1e-8 ? 0.0 : 1.0, 0.0);
    auto xs = lp.getState(0);
    double ovp = 0;
    for (auto& kv : xs) {
      auto f = gs.find(kv.first);
      if (f == gs.end()) continue;
      for (size_t i = 0; i < kv.second.size1() * kv.second.size2(); ++i) ovp += f->second.data()[i] * kv.second.data()[i];
    }
    report("TSLanzcosEngine projected state orthogonal", std::fabs(ovp), 1e-9);
  }
  // complex128 blocks (planar on the device) against the oracle's complex template
  {
    std::mt19937_64 crng(20260002);
    auto Lc = with_imag<1>(L, crng), Rc = with_imag<1>(R, crng);
    auto psic = with_imag<2>(psi, crng);
    auto oLc = to_oracle_c<1>(Lc), oRc = to_oracle_c<1>(Rc);
    std::vector<i_to_key<AbKeyType>> itks2 = {itk, itk};
    BlockMatrix<2, AbKeyType, Cplx> yc;
    contractTwosite(yc, Lc, Rc, psic, cplx_mpo(heis(c, N)), cplx_mpo(heis(c + 1, N)), itks2);
    oracle::BlockMatrix<2, Cplx> oyc;
    oracle::contractTwosite(oyc, oLc, oRc, to_oracle_c<2>(psic), to_oracle_c(heis(c, N)), to_oracle_c(heis(c + 1, N)),
                            oitk, oitk, 1);
    report("contractTwosite complex128", rel_c<2>(yc, oyc), 1e-12);
    auto Ac = with_imag<1>(mps[(size_t)c], crng), Apc = with_imag<1>(mps[(size_t)c], crng);
    auto Lpc = addContractionLayer(Lc, Ac, cplx_mpo(heis(c, N)), itk, 1);
    auto oLpc = oracle::addContractionLayer(oLc, to_oracle_c<1>(Ac), to_oracle_c(heis(c, N)), oitk, 1, 1);
    report("addContractionLayer complex128", rel_c<1>(Lpc, oLpc), 1e-12);
    auto Lmc = addContractionLayer(Lc, Ac, cplx_mpo(heis(c, N)), Apc, itk, 1);
    const auto oApc = to_oracle_c<1>(Apc);
    auto oLmc = oracle::addContractionLayer(oLc, to_oracle_c<1>(Ac), to_oracle_c(heis(c, N)), oitk, 1, 1, &oApc);
    report("addContractionLayer Ap complex128", rel_c<1>(Lmc, oLmc), 1e-12);
    auto R1c = with_imag<1>(addContractionLayer(R, mps[(size_t)c + 1], heis(c + 1, N), itk, -1), crng);
    BlockMatrix<1, AbKeyType, Cplx> y1c;
    contractSingle(y1c, Lc, Ac, cplx_mpo(heis(c, N)), R1c, itk);
    oracle::BlockMatrix<1, Cplx> oy1c;
    oracle::contractSingle(oy1c, oLc, to_oracle_c<1>(Ac), to_oracle_c(heis(c, N)), to_oracle_c<1>(R1c), oitk, 1);
    report("contractSingle complex128", rel_c<1>(y1c, oy1c), 1e-12);
    // Lanczos from a complex start under the real (Hermitian) H_eff: the real ground state energy
    BlockMatrix<1, AbKeyType, Cplx> Lr, Rr;
    for (auto& kv : L) { matrix<Cplx> m(kv.second.size1(), kv.second.size2()); for (size_t i = 0; i < m.size1() * m.size2(); ++i) m.data()[i] = kv.second.data()[i]; Lr[kv.first] = m; }
    for (auto& kv : R) { matrix<Cplx> m(kv.second.size1(), kv.second.size2()); for (size_t i = 0; i < m.size1() * m.size2(); ++i) m.data()[i] = kv.second.data()[i]; Rr[kv.first] = m; }
    TSLanzcosEngine<AbKeyType, Cplx> lanc(Lr, Rr, with_imag<1>(mps[(size_t)c], crng), with_imag<1>(mps[(size_t)c + 1], crng),
                                          cplx_mpo(heis(c, N)), cplx_mpo(heis(c + 1, N)), itk, itk, 0);
    lanc.simulate(300, 1e-11, true);
    TSLanzcosEngine<AbKeyType, Real> lanr(L, R, mps[(size_t)c], mps[(size_t)c + 1], heis(c, N), heis(c + 1, N), itk, itk, 0);
    lanr.simulate(300, 1e-11, true);
    report("TSLanzcosEngine complex128 E0", std::fabs(lanc.E(0).real() - lanr.E(0)), 1e-9);
    auto gsc = lanc.getState(0);
    auto gsr = lanr.getState(0);
    Cplx ovc(0.0, 0.0);
    for (auto& kv : gsr) {
      auto f = gsc.find(kv.first);
      if (f == gsc.end()) continue;
      for (size_t i = 0; i < kv.second.size1() * kv.second.size2(); ++i) ovc += kv.second.data()[i] * f->second.data()[i];
    }
    report("TSLanzcosEngine complex128 state overlap", std::fabs(std::abs(ovc) - 1.0), 1e-6);
  }
  // error path: empty A
  try {
    BlockMatrix<1, AbKeyType, Real> empty;
    addContractionLayer(L, empty, heis(c, N), itk, 1);
    report("empty A raises", 1.0, 0.0);
  } catch (const std::runtime_error& e) {
    report("empty A raises", std::string(e.what()).find("A is empty") != std::string::npos ? 0.0 : 1.0, 0.0);
  }
  std::printf("%s\n", g_fail ? "FAILED" : "ALL PASS");
  return g_fail ? 1 : 0;
}
