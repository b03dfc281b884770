// check_spo / check_wfc / check of the distance tables for the reference-shaped C++ classes of
// include/qmcx_dropin.hpp, in the manner of the reference's src/Drivers/check_spo.cpp:205-347 and
// src/Drivers/check_wfc.cpp:190-485: the same particle configuration is evaluated by the drop-in
// (through libqmcx.so on the GPU) and by a reference implementation, and the errors are printed per
// quantity.  The reference implementation here is the FP64 oracle (oracle/qmc_oracle.cpp, linked as
// liboracle.so).  The system is a NiO-8-like cubic cell built here from a seeded generator (4 "Ni",
// 4 "O", 96 electrons, 48 orbitals on a 24^3 periodic B-spline table, J1 per species, J2 per spin
// pair, and in the last cases a three-body eeI Jastrow per ion species and spin pair, the components
// check_wfc takes with -f Det, J1, J2, J3).  Exit status 0 when every error is within its stated
// tolerance.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <vector>

#include "qmcx_dropin.hpp"

extern "C" {
void* orc_system_create(const double* lattice9, int nions, const double* ion_pos, const int* ion_species, int nspecies,
                        int nup, int ndn);
void orc_system_destroy(void* h);
void orc_system_set_j1(void* h, int species, int nparams, double rcut, double cusp, const double* params);
void orc_system_set_j2(void* h, int pair, int nparams, double rcut, double cusp, const double* params);
int orc_system_set_j3(void* h, int species, int pair, int neI, int nee, double rcut, const double* params,
                      int nparams);
void orc_system_set_spline(void* h, const int* ng, int norb, int npad, const float* coefs, const double* spo_lattice9,
                           int prec);
void orc_spo_evaluate(void* h, int npts, const double* r, double* psi, double* dpsi, double* d2psi);
void orc_distance_row(const double* R9, const double* pos, int n, const double* targets, int flip_ind, double* r,
                      double* dr);
double orc_run_vmc(void* h, int nw, int walker0, double* R, int nsweeps, uint64_t seed, double tau, int mode, int delay,
                   double* log_init, double* log_out, double* trace, double* G, double* L, long long* naccept,
                   int nthreads, double* sens);
void orc_move_randoms(uint64_t seed, uint64_t walker, uint64_t sweep, uint64_t iel, double* normal3, double* uacc);
}

namespace {

// splitmix64 stream: U(0, 1)
struct Rng {
  uint64_t s;
  double uniform() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

// the members of the reference's ParticleSet that the drop-in classes read
struct Pos {
  double v[3] = {0, 0, 0};
  double& operator[](int d) { return v[d]; }
  const double& operator[](int d) const { return v[d]; }
};
struct Particles {
  std::vector<Pos> R, G;
  std::vector<double> L;
  int active = -1;
  Pos activePos;
  int getTotalNum() const { return (int)R.size(); }
  const Pos& activeR(int iat) const { return iat == active ? activePos : R[iat]; }
  void makeMove(int iat, const double* d) {
    active = iat;
    for (int k = 0; k < 3; ++k) activePos[k] = R[iat][k] + d[k];
  }
  void acceptMove(int iat) {
    R[iat] = activePos;
    active = -1;
  }
  void rejectMove(int) { active = -1; }
};

int failures = 0;
void report(const char* what, double err, double tol) {
  const bool ok = err <= tol;
  std::printf("%-52s error %.3e (tolerance %.1e)%s\n", what, err, tol, ok ? "" : "  FAILED");
  if (!ok) ++failures;
}

constexpr double kA = 7.9;
// three-body (eeI) Jastrow: the reference's functor shape (3, 3) with its 26 free parameters, cutoff the
// cell's Wigner-Seitz radius (ThreeBodyJastrow.h:196-218, PolynomialFunctor3D.h:65-260)
constexpr int kJ3NeI = 3, kJ3Nee = 3, kJ3Params = 26;
constexpr double kJ3Rcut = 0.5 * kA;
constexpr int kNup = 48, kN = 96, kNorb = 48, kNpad = 48, kG = 24;

struct SystemData {
  double lattice[9] = {kA, 0, 0, 0, kA, 0, 0, 0, kA};
  std::vector<double> ions;
  std::vector<int> species;
  std::vector<float> coefs;
  std::vector<double> j1[2], j2[2];
  std::vector<double> j3[2][2];   // three-body functor parameters per (ion species, spin pair)
};

SystemData make_data(Rng& rng) {
  SystemData d;
  const double fcc[4][3] = {{0, 0, 0}, {0, .5, .5}, {.5, 0, .5}, {.5, .5, 0}};
  for (int sp = 0; sp < 2; ++sp)
    for (int i = 0; i < 4; ++i) {
      for (int k = 0; k < 3; ++k) d.ions.push_back(kA * std::fmod(fcc[i][k] + 0.5 * sp, 1.0));
      d.species.push_back(sp);
    }
  // periodic table [g+3]^3 x npad: ghosts repeat the first three planes of each axis
  const int gp = kG + 3;
  std::vector<float> core((size_t)kG * kG * kG * kNorb);
  for (auto& c : core) c = (float)(2.0 * rng.uniform() - 1.0);
  d.coefs.assign((size_t)gp * gp * gp * kNpad, 0.f);
  for (int i = 0; i < gp; ++i)
    for (int j = 0; j < gp; ++j)
      for (int k = 0; k < gp; ++k)
        for (int o = 0; o < kNorb; ++o)
          d.coefs[(((size_t)i * gp + j) * gp + k) * kNpad + o] =
              core[(((size_t)(i % kG) * kG + j % kG) * kG + k % kG) * kNorb + o];
  for (int s = 0; s < 2; ++s) {
    for (int p = 0; p < 8; ++p) d.j1[s].push_back(p == 7 ? 0.0 : 0.5 * rng.uniform());
    for (int p = 0; p < 8; ++p) d.j2[s].push_back(p == 7 ? 0.0 : 0.5 * rng.uniform());
  }
  for (int s = 0; s < 2; ++s)
    for (int pr = 0; pr < 2; ++pr)
      for (int p = 0; p < kJ3Params; ++p) d.j3[s][pr].push_back(0.04 * (rng.uniform() - 0.5));
  return d;
}

const double kRcut = 3.9, kCusp2[2] = {-0.25, -0.5};

void setup_qmcx(qmcx_system_t& sys, const SystemData& d, bool jastrows, bool j3 = false) {
  qmcx::check(qmcx_system_create(&sys, d.lattice, 8, d.ions.data(), d.species.data(), 2, kNup, kNup),
              "qmcx_system_create");
  const int ng[3] = {kG, kG, kG};
  qmcx::check(qmcx_system_set_spline(sys, ng, kNorb, kNpad, d.coefs.data(), d.lattice, 3), "qmcx_system_set_spline");
  if (jastrows)
    for (int s = 0; s < 2; ++s) {
      qmcx::check(qmcx_system_set_j1(sys, s, 8, kRcut, 0.0, d.j1[s].data()), "qmcx_system_set_j1");
      qmcx::check(qmcx_system_set_j2(sys, s, 8, kRcut, kCusp2[s], d.j2[s].data()), "qmcx_system_set_j2");
    }
  if (j3)
    for (int s = 0; s < 2; ++s)
      for (int pr = 0; pr < 2; ++pr)
        qmcx::check(qmcx_system_set_j3(sys, s, pr, kJ3NeI, kJ3Nee, kJ3Rcut, d.j3[s][pr].data(), kJ3Params),
                    "qmcx_system_set_j3");
}

void* setup_oracle(const SystemData& d, bool jastrows, bool j3 = false) {
  void* h = orc_system_create(d.lattice, 8, d.ions.data(), d.species.data(), 2, kNup, kNup);
  const int ng[3] = {kG, kG, kG};
  orc_system_set_spline(h, ng, kNorb, kNpad, d.coefs.data(), d.lattice, 8);
  if (jastrows)
    for (int s = 0; s < 2; ++s) {
      orc_system_set_j1(h, s, 8, kRcut, 0.0, d.j1[s].data());
      orc_system_set_j2(h, s, 8, kRcut, kCusp2[s], d.j2[s].data());
    }
  if (j3)
    for (int s = 0; s < 2; ++s)
      for (int pr = 0; pr < 2; ++pr)
        if (orc_system_set_j3(h, s, pr, kJ3NeI, kJ3Nee, kJ3Rcut, d.j3[s][pr].data(), kJ3Params) < 0)
          throw std::runtime_error("orc_system_set_j3");
  return h;
}

Particles random_particles(Rng& rng) {
  Particles P;
  P.R.resize(kN);
  P.G.resize(kN);
  P.L.assign(kN, 0.0);
  for (auto& r : P.R)
    for (int k = 0; k < 3; ++k) r[k] = kA * rng.uniform();
  return P;
}

// ---- check_spo: SPOSetHIP against the oracle's SPO evaluation
template <typename T>
void check_spo(qmcx_system_t sys, void* orc, Rng& rng, double tol) {
  const char* tag = sizeof(T) == 8 ? "FP64" : "FP32";
  qmcx::SPOSetHIP<T> spo(sys, kNorb);
  const int np = 16;
  double ev = 0, eg = 0, el = 0, evv = 0;
  std::vector<T> psi(kNorb), d2psi(kNorb), psiv(kNorb);
  std::vector<std::array<T, 3>> dpsi(kNorb);
  std::vector<double> rpsi(kNorb), rd(3 * kNorb), rl(kNorb);
  for (int p = 0; p < np; ++p) {
    Particles P = random_particles(rng);
    const double d[3] = {kA * (rng.uniform() - 0.3), kA * (rng.uniform() - 0.3), kA * (rng.uniform() - 0.3)};
    P.makeMove(5, d);   // includes points outside the cell
    spo.evaluate(P, 5, psi, dpsi, d2psi);
    spo.evaluate(P, 5, psiv);
    orc_spo_evaluate(orc, 1, P.activeR(5).v, rpsi.data(), rd.data(), rl.data());
    for (int j = 0; j < kNorb; ++j) {
      ev = std::max(ev, std::fabs(psi[j] - rpsi[j]) / (1 + std::fabs(rpsi[j])));
      evv = std::max(evv, std::fabs(psiv[j] - rpsi[j]) / (1 + std::fabs(rpsi[j])));
      for (int k = 0; k < 3; ++k) eg = std::max(eg, std::fabs(dpsi[j][k] - rd[3 * j + k]) / (1 + std::fabs(rd[3 * j + k])));
      el = std::max(el, std::fabs(d2psi[j] - rl[j]) / (1 + std::fabs(rl[j])));
    }
  }
  char buf[96];
  std::snprintf(buf, sizeof buf, "check_spo %s evaluate_v: value", tag);
  report(buf, evv, tol);
  std::snprintf(buf, sizeof buf, "check_spo %s evaluate_vgl: value", tag);
  report(buf, ev, tol);
  std::snprintf(buf, sizeof buf, "check_spo %s evaluate_vgl: gradient", tag);
  report(buf, eg, tol * 10);
  std::snprintf(buf, sizeof buf, "check_spo %s evaluate_vgl: d2psi (h_xx)", tag);
  report(buf, el, tol * 100);
  // multi_evaluate over four particle sets
  std::vector<Particles> Ps;
  for (int w = 0; w < 4; ++w) Ps.push_back(random_particles(rng));
  std::vector<Particles*> pl;
  std::vector<qmcx::SPOSetHIP<T>*> sl;
  std::vector<std::vector<T>> v(4, std::vector<T>(kNorb)), l(4, std::vector<T>(kNorb));
  std::vector<std::vector<std::array<T, 3>>> g(4, std::vector<std::array<T, 3>>(kNorb));
  std::vector<std::vector<T>*> vl, ll;
  std::vector<std::vector<std::array<T, 3>>*> gl;
  for (int w = 0; w < 4; ++w) {
    pl.push_back(&Ps[w]);
    sl.push_back(&spo);
    vl.push_back(&v[w]);
    ll.push_back(&l[w]);
    gl.push_back(&g[w]);
  }
  qmcx::SPOSetHIP<T>::multi_evaluate(sl, pl, 17, vl, gl, ll);
  double em = 0;
  for (int w = 0; w < 4; ++w) {
    orc_spo_evaluate(orc, 1, Ps[w].R[17].v, rpsi.data(), rd.data(), rl.data());
    for (int j = 0; j < kNorb; ++j) em = std::max(em, std::fabs(v[w][j] - rpsi[j]) / (1 + std::fabs(rpsi[j])));
  }
  std::snprintf(buf, sizeof buf, "check_spo %s multi_evaluate: value", tag);
  report(buf, em, tol);
}

// ---- distance tables: DistanceTableHIP (AA and BA) against the oracle's rows
template <typename T>
void check_distance_tables(qmcx_system_t sys, const SystemData& d, Rng& rng, double tol) {
  const char* tag = sizeof(T) == 8 ? "FP64" : "FP32";
  Particles P = random_particles(rng);
  qmcx::DistanceTableHIP<T> aa(sys, kN);
  qmcx::DistanceTableHIP<T> ba(sys, kN, 8, d.ions.data());
  aa.evaluate(P);
  ba.evaluate(P);
  std::vector<double> targets(3 * kN), r(kN), dr(3 * kN);
  for (int i = 0; i < kN; ++i)
    for (int k = 0; k < 3; ++k) targets[3 * i + k] = P.R[i][k];
  double eaa = 0, eba = 0, emove = 0;
  for (int i = 0; i < kN; ++i) {
    orc_distance_row(d.lattice, P.R[i].v, kN, targets.data(), i, r.data(), dr.data());
    for (int j = 0; j < kN; ++j) {
      if (j != i) eaa = std::max(eaa, std::fabs(aa.Distances[(size_t)i * kN + j] - r[j]));
      for (int k = 0; k < 3; ++k)
        eaa = std::max(eaa, std::fabs(aa.Displacements[((size_t)i * 3 + k) * kN + j] - dr[3 * j + k]));
    }
    orc_distance_row(d.lattice, P.R[i].v, 8, d.ions.data(), 0, r.data(), dr.data());
    for (int j = 0; j < 8; ++j) {
      eba = std::max(eba, std::fabs(ba.Distances[(size_t)i * 8 + j] - r[j]));
      for (int k = 0; k < 3; ++k)
        eba = std::max(eba, std::fabs(ba.Displacements[((size_t)i * 3 + k) * 8 + j] - dr[3 * j + k]));
    }
  }
  // a proposal of particle 3 (move), then update: row 3 becomes the proposal's
  const double disp[3] = {0.7, -1.9, 0.4};
  P.makeMove(3, disp);
  aa.move(P, P.activeR(3), 3);
  orc_distance_row(d.lattice, P.activeR(3).v, kN, targets.data(), 3, r.data(), dr.data());
  for (int j = 0; j < kN; ++j) emove = std::max(emove, std::fabs(aa.Temp_r[j] - r[j]));
  aa.update(3);
  for (int j = 0; j < kN; ++j) emove = std::max(emove, std::fabs(aa.Distances[(size_t)3 * kN + j] - r[j]));
  char buf[96];
  std::snprintf(buf, sizeof buf, "distance table AA %s: Distances, Displacements", tag);
  report(buf, eaa, tol);
  std::snprintf(buf, sizeof buf, "distance table BA %s: Distances, Displacements", tag);
  report(buf, eba, tol);
  std::snprintf(buf, sizeof buf, "distance table AA %s: move, update", tag);
  report(buf, emove, tol);
}

// ---- check_wfc: WaveFunctionBatchHIP's flex_* loop against the oracle's sweep (same walkers,
// same displacements and accept decisions), then WaveFunctionHIP for one walker
void check_wfc(qmcx_system_t sys, void* orc, const char* label, int prec, int delay, double tol_r, double tol_log) {
  const int W = 3;
  const uint64_t seed = 7;
  const double tau = 0.5;
  Rng rng{1234};
  std::vector<Particles> Ps;
  std::vector<double> R0((size_t)W * kN * 3);
  for (int w = 0; w < W; ++w) {
    Ps.push_back(random_particles(rng));
    for (int i = 0; i < kN; ++i)
      for (int k = 0; k < 3; ++k) R0[((size_t)w * kN + i) * 3 + k] = Ps[w].R[i][k];
  }
  std::vector<double> Rref = R0, log_init(W), log_out(W), trace((size_t)W * kN * 9), G((size_t)W * kN * 3),
                      L((size_t)W * kN);
  std::vector<long long> nacc(W);
  orc_run_vmc(orc, W, 0, Rref.data(), 1, seed, tau, 1, 1, log_init.data(), log_out.data(), trace.data(), G.data(),
              L.data(), nacc.data(), 0, nullptr);

  qmcx::WaveFunctionBatchHIP wf(sys, W, kN, delay, prec);
  std::vector<Particles*> pl;
  for (auto& P : Ps) pl.push_back(&P);
  wf.flex_evaluateLog(pl);
  double elog0 = 0, er = 0, egn = 0, egw = 0;
  int flips = 0;
  for (int w = 0; w < W; ++w) elog0 = std::max(elog0, std::fabs(wf.getLogValue(w) - log_init[w]) / std::fabs(log_init[w]));
  std::vector<std::array<double, 3>> gnow(W), gnew(W);
  std::vector<double> ratios(W);
  for (int iel = 0; iel < kN; ++iel) {
    wf.flex_evalGrad(pl, iel, gnow);
    std::vector<bool> acc(W);
    for (int w = 0; w < W; ++w) {
      double n3[3], u = 0;
      orc_move_randoms(seed, (uint64_t)w, 0, (uint64_t)iel, n3, &u);
      const double d[3] = {std::sqrt(tau) * n3[0], std::sqrt(tau) * n3[1], std::sqrt(tau) * n3[2]};
      Ps[w].makeMove(iel, d);
    }
    wf.flex_ratioGrad(pl, iel, ratios, gnew);
    for (int w = 0; w < W; ++w) {
      const double* t = trace.data() + ((size_t)w * kN + iel) * 9;
      er = std::max(er, std::fabs(ratios[w] - t[0]) / std::max(std::fabs(t[0]), 1.0));
      for (int k = 0; k < 3; ++k) {
        egn = std::max(egn, std::fabs(gnow[w][k] - t[3 + k]) / (1 + std::fabs(t[3 + k])));
        egw = std::max(egw, std::fabs(gnew[w][k] - t[6 + k]) / (1 + std::fabs(t[6 + k])));
      }
      double n3[3], u = 0;
      orc_move_randoms(seed, (uint64_t)w, 0, (uint64_t)iel, n3, &u);
      acc[w] = u < ratios[w] * ratios[w];   // Metropolis on the drop-in's own ratio
      if (acc[w] != (t[2] != 0.0)) ++flips;
      acc[w] = t[2] != 0.0;                 // continue on the oracle's chain
    }
    wf.flex_acceptrestoreMove(pl, acc, iel);
    for (int w = 0; w < W; ++w) {
      if (acc[w]) Ps[w].acceptMove(iel);
      else Ps[w].rejectMove(iel);
    }
  }
  wf.flex_completeUpdates();
  wf.flex_evaluateGL(pl);
  double elog = 0, eG = 0, eL = 0, eR = 0;
  for (int w = 0; w < W; ++w) {
    elog = std::max(elog, std::fabs(wf.getLogValue(w) - log_out[w]) / std::fabs(log_out[w]));
    for (int i = 0; i < kN; ++i) {
      for (int k = 0; k < 3; ++k) {
        eG = std::max(eG, std::fabs(Ps[w].G[i][k] - G[((size_t)w * kN + i) * 3 + k]) /
                              (1 + std::fabs(G[((size_t)w * kN + i) * 3 + k])));
        eR = std::max(eR, std::fabs(Ps[w].R[i][k] - Rref[((size_t)w * kN + i) * 3 + k]));
      }
      eL = std::max(eL, std::fabs(Ps[w].L[i] - L[(size_t)w * kN + i]) / (1 + std::fabs(L[(size_t)w * kN + i])));
    }
  }
  char buf[96];
  auto line = [&](const char* what, double e, double tol) {
    std::snprintf(buf, sizeof buf, "check_wfc %s %s", label, what);
    report(buf, e, tol);
  };
  line("flex_evaluateLog: V", elog0, tol_log);
  line("flex_evalGrad: G", egn, 100 * tol_r);
  line("flex_ratioGrad: ratio", er, tol_r);
  line("flex_ratioGrad: G", egw, 100 * tol_r);
  line("accept decisions differing", (double)flips, 0.0);
  line("positions after the sweep", eR, 1e-12);
  line("flex_evaluateGL: V", elog, tol_log);
  line("flex_evaluateGL: G", eG, 100 * tol_r);
  line("flex_evaluateGL: L", eL, 1000 * tol_r);

  // one walker through the single-walker calls: walker 0's first 12 moves
  qmcx::WaveFunctionHIP one(sys, kN, delay, prec);
  Particles P1 = random_particles(rng);
  for (int i = 0; i < kN; ++i)
    for (int k = 0; k < 3; ++k) P1.R[i][k] = R0[(size_t)i * 3 + k];
  one.evaluateLog(P1);
  double e1 = std::fabs(one.getLogValue() - log_init[0]) / std::fabs(log_init[0]);
  for (int iel = 0; iel < 12; ++iel) {
    const auto g0 = one.evalGrad(P1, iel);
    double n3[3], u = 0;
    orc_move_randoms(seed, 0, 0, (uint64_t)iel, n3, &u);
    const double d[3] = {std::sqrt(tau) * n3[0], std::sqrt(tau) * n3[1], std::sqrt(tau) * n3[2]};
    P1.makeMove(iel, d);
    std::array<double, 3> g1{};
    const double r = one.ratioGrad(P1, iel, g1);
    const double* t = trace.data() + (size_t)iel * 9;
    e1 = std::max(e1, std::fabs(r - t[0]) / std::max(std::fabs(t[0]), 1.0));
    for (int k = 0; k < 3; ++k) e1 = std::max(e1, std::fabs(g0[k] - t[3 + k]) / (100 * (1 + std::fabs(t[3 + k]))));
    if (t[2] != 0.0) {
      one.acceptMove(P1, iel);
      P1.acceptMove(iel);
    } else {
      one.restore(iel);
      P1.rejectMove(iel);
    }
  }
  one.completeUpdates();
  line("single walker: evaluateLog, evalGrad, ratioGrad", e1, tol_r);
}

}  // namespace

int main() {
  Rng rng{2024};
  const SystemData d = make_data(rng);
  qmcx_system_t sys = nullptr, sys_det = nullptr, sys_j3 = nullptr;
  setup_qmcx(sys, d, true);
  setup_qmcx(sys_det, d, false);
  setup_qmcx(sys_j3, d, true, true);
  void* orc = setup_oracle(d, true);
  void* orc_det = setup_oracle(d, false);
  void* orc_j3 = setup_oracle(d, true, true);
  try {
    check_spo<double>(sys, orc, rng, 1e-12);
    check_spo<float>(sys, orc, rng, 2e-5);
    check_distance_tables<double>(sys, d, rng, 1e-12);
    check_distance_tables<float>(sys, d, rng, 2e-5);
    check_wfc(sys_det, orc_det, "Det FP64 delay 1", QMCX_FP64, 1, 1e-9, 1e-11);
    check_wfc(sys, orc, "Det+J1+J2 FP64 delay 1", QMCX_FP64, 1, 1e-9, 1e-11);
    check_wfc(sys, orc, "Det+J1+J2 FP64 delay 4", QMCX_FP64, 4, 1e-9, 1e-11);
    check_wfc(sys, orc, "Det+J1+J2 FP32 delay 8", QMCX_FP32, 8, 2e-3, 1e-6);
    check_wfc(sys_j3, orc_j3, "Det+J1+J2+J3 FP64 delay 1", QMCX_FP64, 1, 1e-9, 1e-11);
    check_wfc(sys_j3, orc_j3, "Det+J1+J2+J3 FP64 delay 4", QMCX_FP64, 4, 1e-9, 1e-11);
    check_wfc(sys_j3, orc_j3, "Det+J1+J2+J3 FP32 delay 8", QMCX_FP32, 8, 2e-3, 1e-6);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    ++failures;
  }
  orc_system_destroy(orc);
  orc_system_destroy(orc_det);
  orc_system_destroy(orc_j3);
  qmcx_system_destroy(sys);
  qmcx_system_destroy(sys_det);
  qmcx_system_destroy(sys_j3);
  if (failures) {
    std::printf("%d checks FAILED\n", failures);
    return 1;
  }
  std::printf("All checks passed\n");
  return 0;
}
