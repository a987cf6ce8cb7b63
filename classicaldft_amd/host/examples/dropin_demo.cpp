// The hot path driven through the reference-named C++ API (config 1 and a
// short FIRE / DDFT run), printing machine-readable lines for the tests:
//   F_hs <value>            FMT::calculateFreeEnergyAndDerivatives on config 1
//   dF_sum <value>          sum of the FMT dF
//   mu <value>              DFT::mu_times_beta(0.5) with LJ mean field, kT = 1
//   fire <F> <fmax_raw>     after 20 fireMinimizer2 steps
//   ddft <F> <mass>         after 3 DDFT steps
//   mu_twf_bh_gauss_f <mu> <hsd> <vdw>  tWF (BH split) on a Gauss_F kernel, kT = 1.2
//   avdw_whdf_le <a>        lattice vdW parameter of WHDF on the LE kernel, kT = 0.9
//   hess <sum Hv> <v.Hv>    DFT::matrix_dot_v1 of a seeded direction (16^3, ideal + WB I + LJ/QF)
//   hess_fmt <sum> <v.Hv>   FMT::add_second_derivative of the same direction
//   ao <F> <sum dF> <muex>  FMT_AO_Species (Rp = 0.3, reservoir 0.4) with White Bear I on 16^3, BulkMuex(0.5)
//   fixed_mass_eval <F> <mu> <sum dF>   Species::setFixedMass(0.95 N), one evaluation (16^3, WB I + LJ/QF)
//   fixed_mass_fire <F> <mu> <N>        then 10 fireMinimizer2 steps
//   fixed_bg <F> <sum dF> <F onlyFex>   Species::setFixedBackground on another 16^3 system
//   fire_dir <F>            5 fireMinimizer2 steps with Minimizer::set_fixed_direction (ideal + WB I, mu = 0.4)
//   scaled_mf <a> <w0> <mu>  Interaction_Base::scale_interaction(0.5) on LJ/QF: a_vdw, getW(0), mu_times_beta(0.5)
//   log_det <ld> <var>       Log_Det(dft, 20, 0.05, 5e-4).calculate_log_det(7, 2, false) (16^3, ideal + WB I)
//   log_det_ddft <ld> <var>  Log_Det(ddft, 12, 40, 5e-3).calculate_log_det(7, 2, true): the DDFT dynamical matrix
//   coex <xs1> <xs2> <x1> <x2> <x> <xc> <Tc>  findSpinodal / findCoex / find_density_from_mu(mu(0.3)) /
//                           getCriticalPoint (16^3, ideal + WB I + LJ/QF at kT = 0.8)
//   ddft_dot <sum> <v.Mv>   DDFT::matrix_dot_v1 of the seeded direction (the dynamical matrix g . H)
//   eigen <min> <max> <n1> <n2>  Eigenvalues: smallest and (change_sign) largest Hessian eigenvalue (8^3, ideal + WB I)
//   open_ddft <F> <rho(0,0,0)> <rho(8,8,8)>  2 DDFT_IF_Open steps (16^3, WB I + LJ/QF, background 0.5)
//   points <a1> <a2> <cached> <F>  points-file Interaction (60 Fibonacci directions written to sphere_points.dat):
//                           a_vdw of a first object (kernel generated, cache written) and of a second one (cache
//                           read: 1), F of a 16^3 DFT with the cached kernel
//   mix <F> <sum dF_a> <sum dF_b>  FMT over two FMT_Species (d = 1 and 0.6, White Bear I, 16 x 16 x 24, dx = 0.1)
//   mix_dft <F> <F_id> <sum dF_a> <sum dF_b>  a DFT of two FMT_Species (mu 0.5 / 0.2, a field on the first)
//   mix_fire <F> <fmax_raw> <N_a> <N_b>      after 20 fireMinimizer2 steps over both species
//   mix_mass <F> <mu_b> <N_b>                the second species with a fixed mass (0.42 V); then
//   mix_mass_fire <F> <mu_b> <N_b>           after 10 fireMinimizer2 steps
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "classicaldft/classicaldft.h"

static double splitmix_u(uint64_t seed, uint64_t k) {  // (k+1)-th draw of splitmix64(seed)
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int main() {
  // config 1: 32^3, dx = 0.1, hsd = 1, White Bear I, rho = 0.6 (1 + 0.05 U[-1,1])
  double L[3] = {3.2, 3.2, 3.2};
  Density density(0.1, L);
  for (long i = 0; i < density.Ntot(); ++i) density.set(i, 0.6 * (1.0 + 0.05 * (2.0 * splitmix_u(0x5EED0001ULL, i) - 1.0)));
  FMT_Species species(density, 1.0, 0.0);
  WhiteBearI fmt;
  std::vector<Species*> all{&species};
  species.zeroForce();
  const double F = fmt.calculateFreeEnergyAndDerivatives(all);
  std::printf("F_hs %.17g\n", F);
  std::printf("dF_sum %.17g\n", species.getDF().accu());

  // 16^3 DFT with ideal gas + WB I + LJ/QF mean field at kT = 1
  double L2[3] = {1.6, 1.6, 1.6};
  Density d2(0.1, L2);
  for (long i = 0; i < d2.Ntot(); ++i) d2.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(11, i) - 1.0)));
  FMT_Species s2(d2, 1.0, 0.5);
  LJ lj(1.0, 1.0, 2.5);
  Interaction_Interpolation_QF qf(&s2, &s2, &lj, 1.0);
  WhiteBearI fmt2;
  DFT dft(&s2);
  dft.addHardCoreContribution(&fmt2);
  dft.addInteraction(&qf);
  std::printf("mu %.17g\n", dft.mu_times_beta(0.5));
  fireMinimizer2 fire(&dft);
  for (int i = 0; i < 20; ++i) fire.step();
  std::printf("fire %.17g %.17g\n", fire.getF(), fire.get_raw_force_max());

  Density d3(0.1, L2);
  for (long i = 0; i < d3.Ntot(); ++i) d3.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(21, i) - 1.0)));
  FMT_Species s3(d3, 1.0, 0.0);
  Interaction_Interpolation_QF qf3(&s3, &s3, &lj, 1.0);
  WhiteBearI fmt3;
  DFT dft3(&s3);
  dft3.addHardCoreContribution(&fmt3);
  dft3.addInteraction(&qf3);
  DDFT ddft(&dft3);
  for (int i = 0; i < 3; ++i) ddft.step();
  std::printf("ddft %.17g %.17g\n", ddft.getF(), ddft.get_mass());
  std::printf("hsd_lj_kT0.8 %.17g\n", lj.getHSD(0.8));

  // other potentials and kernel rules: tWF with the BH split on a 3-point Gauss_F kernel; WHDF on LE
  Density d5(0.1, L2);
  d5.set(0.5);
  FMT_Species s5(d5, 1.0, 0.0);
  tWF twf(1.0, 1.0, 2.0, 50.0);
  twf.setBH();
  Interaction_Gauss_F gf(&s5, &s5, &twf, 1.2, 3);
  WhiteBearI fmt5;
  DFT dft5(&s5);
  dft5.addHardCoreContribution(&fmt5);
  dft5.addInteraction(&gf);
  std::printf("mu_twf_bh_gauss_f %.17g %.17g %.17g\n", dft5.mu_times_beta(0.5), twf.getHSD(1.2),
              twf.getVDW_Parameter(1.2));
  WHDF whdf(1.0, 1.0, 2.0);
  Interaction_Interpolation_LE le(&s5, &s5, &whdf, 0.9);
  std::printf("avdw_whdf_le %.17g\n", le.getVDWParameter());

  // Hessian-vector products (DFT::matrix_dot_v1 and FMT::add_second_derivative) on a 16^3 system
  Density d4(0.1, L2);
  for (long i = 0; i < d4.Ntot(); ++i) d4.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(41, i) - 1.0)));
  FMT_Species s4(d4, 1.0, 0.0);
  Interaction_Interpolation_QF qf4(&s4, &s4, &lj, 1.0);
  WhiteBearI fmt4;
  DFT dft4(&s4);
  dft4.addHardCoreContribution(&fmt4);
  dft4.addInteraction(&qf4);
  DFT_Vec v(d4.Ntot()), Hv;
  for (long i = 0; i < d4.Ntot(); ++i) v.set(i, 2.0 * splitmix_u(43, i) - 1.0);
  dft4.matrix_dot_v1(v, Hv);
  std::printf("hess %.17g %.17g\n", Hv.accu(), Hv.dotWith(v));
  {
    FMT_Species s13(d4, 1.0, 0.0);
    WhiteBearI fmt13;
    DFT dft13(&s13);
    dft13.addHardCoreContribution(&fmt13);
    Log_Det ldet(dft13, 20, 0.05, 5e-4);
    double var = 0.0;
    const double ldv = ldet.calculate_log_det(7, 2, false, var);
    std::printf("log_det %.17g %.17g\n", ldv, var);
    DDFT ddft13(&dft13);
    Log_Det ldd(ddft13, 12, 40.0, 5e-3);
    const double ldd_v = ldd.calculate_log_det(7, 2, true, var);
    std::printf("log_det_ddft %.17g %.17g\n", ldd_v, var);
    DFT_Vec gv;
    ddft13.matrix_dot_v1(v, gv);
    std::printf("ddft_dot %.17g %.17g\n", gv.accu(), gv.dotWith(v));
  }
  {
    double L5[3] = {1.6, 1.6, 1.6};
    Density d5(0.2, L5);
    for (long i = 0; i < d5.Ntot(); ++i) d5.set(i, 0.4 * (1.0 + 0.1 * (2.0 * splitmix_u(47, i) - 1.0)));
    FMT_Species s14(d5, 1.0, 0.0);
    WhiteBearI fmt14;
    DFT dft14(&s14);
    dft14.addHardCoreContribution(&fmt14);
    Eigenvalues lo(dft14), hi(dft14);
    lo.set_tolerence(1e-13);
    hi.set_tolerence(1e-13);
    hi.set_change_sign(true);
    lo.calculate_eigenvector();
    hi.calculate_eigenvector();
    std::printf("eigen %.17g %.17g %d %d\n", lo.get_eigenvalue(), hi.get_eigenvalue(), lo.get_num_eval(),
                hi.get_num_eval());
  }
  {
    double L6[3] = {1.6, 1.6, 1.6};
    Density d6(0.1, L6);
    FMT_Species s15(d6, 1.0, 0.0);
    Interaction_Interpolation_QF qf15(&s15, &s15, &lj, 0.8);
    WhiteBearI fmt15;
    DFT dft15(&s15);
    dft15.addHardCoreContribution(&fmt15);
    dft15.addInteraction(&qf15);
    double xs1 = 0, xs2 = 0, x1 = 0, x2 = 0, xc = 0, Tc = 1.0;
    dft15.findSpinodal(1.2, 0.01, xs1, xs2, 1e-10);
    dft15.findCoex(1.2, 0.01, x1, x2, 1e-10);
    const double xmu = dft15.find_density_from_mu(dft15.mu_times_beta(0.3), 0.2, 0.5, 1e-13);
    dft15.getCriticalPoint(lj, xc, Tc);
    std::printf("coex %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", xs1, xs2, x1, x2, xmu, xc, Tc);
  }
  std::vector<DFT_Vec> vv(1), d2F(1);
  vv[0].set(v);
  d2F[0].zeros(d4.Ntot());
  std::vector<Species*> sp{&s4};
  fmt4.add_second_derivative(vv, d2F, sp);
  std::printf("hess_fmt %.17g %.17g\n", d2F[0].accu(), d2F[0].dotWith(v));

  // Asakura-Oosawa polymer of radius 0.3 at reservoir density 0.4
  Density d6(0.1, L2);
  for (long i = 0; i < d6.Ntot(); ++i) d6.set(i, 0.45 * (1.0 + 0.1 * (2.0 * splitmix_u(51, i) - 1.0)));
  FMT_AO_Species s6(d6, 1.0, 0.3, 0.4, 0.0);
  WhiteBearI fmt6;
  std::vector<Species*> a6{&s6};
  s6.zeroForce();
  const double F6 = fmt6.calculateFreeEnergyAndDerivatives(a6);
  std::printf("ao %.17g %.17g %.17g\n", F6, s6.getDF().accu(), fmt6.BulkMuex({0.5}, a6, 0));

  // species constraints: fixed mass (evaluation and FIRE), fixed background (evaluation, onlyFex)
  Density d7(0.1, L2);
  for (long i = 0; i < d7.Ntot(); ++i) d7.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(61, i) - 1.0)));
  FMT_Species s7(d7, 1.0, 0.0);
  Interaction_Interpolation_QF qf7(&s7, &s7, &lj, 1.0);
  WhiteBearI fmt7;
  DFT dft7(&s7);
  dft7.addHardCoreContribution(&fmt7);
  dft7.addInteraction(&qf7);
  s7.setFixedMass(0.95 * d7.getNumberAtoms());
  const double F7 = dft7.calculateFreeEnergyAndDerivatives(false);
  std::printf("fixed_mass_eval %.17g %.17g %.17g\n", F7, s7.getChemPotential(), s7.getDF().accu());
  fireMinimizer2 fire7(&dft7);
  fire7.set_snapshot_interval(0);
  for (int i = 0; i < 10; ++i) fire7.step();
  std::printf("fixed_mass_fire %.17g %.17g %.17g\n", fire7.getF(), s7.getChemPotential(), d7.getNumberAtoms());

  Density d8(0.1, L2);
  for (long i = 0; i < d8.Ntot(); ++i) d8.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(71, i) - 1.0)));
  FMT_Species s8(d8, 1.0, 0.3);
  Interaction_Interpolation_QF qf8(&s8, &s8, &lj, 1.0);
  WhiteBearI fmt8;
  DFT dft8(&s8);
  dft8.addHardCoreContribution(&fmt8);
  dft8.addInteraction(&qf8);
  s8.setFixedBackground(true);
  const double F8 = dft8.calculateFreeEnergyAndDerivatives(false);
  const double dF8 = s8.getDF().accu();
  const double F8x = dft8.calculateFreeEnergyAndDerivatives(true);
  std::printf("fixed_bg %.17g %.17g %.17g\n", F8, dF8, F8x);

  Density d9(0.1, L2);
  for (long i = 0; i < d9.Ntot(); ++i) d9.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(81, i) - 1.0)));
  FMT_Species s9(d9, 1.0, 0.4);
  WhiteBearI fmt9;
  DFT dft9(&s9);
  dft9.addHardCoreContribution(&fmt9);
  DFT_Vec e(d9.Ntot());
  double en = 0.0;
  for (long i = 0; i < d9.Ntot(); ++i) {
    e.set(i, 2.0 * splitmix_u(83, i) - 1.0);
    en += e.get(i) * e.get(i);
  }
  for (long i = 0; i < d9.Ntot(); ++i) e.set(i, e.get(i) / std::sqrt(en));
  fireMinimizer2 fire9(&dft9);
  fire9.set_snapshot_interval(0);
  fire9.set_fixed_direction(e);
  for (int i = 0; i < 5; ++i) fire9.step();
  std::printf("fire_dir %.17g\n", fire9.getF());

  {
    FMT_Species s12(d2, 1.0, 0.0);
    Interaction_Interpolation_QF qf12(&s12, &s12, &lj, 1.0);
    qf12.scale_interaction(0.5);
    WhiteBearI fmt12;
    DFT dft12(&s12);
    dft12.addHardCoreContribution(&fmt12);
    dft12.addInteraction(&qf12);
    std::printf("scaled_mf %.17g %.17g %.17g\n", qf12.getVDWParameter(), qf12.getW(0), dft12.mu_times_beta(0.5));
  }

  Density d10(0.1, L2);
  for (long i = 0; i < d10.Ntot(); ++i) d10.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(91, i) - 1.0)));
  FMT_Species s10(d10, 1.0, 0.0);
  Interaction_Interpolation_QF qf10(&s10, &s10, &lj, 1.0);
  WhiteBearI fmt10;
  DFT dft10(&s10);
  dft10.addHardCoreContribution(&fmt10);
  dft10.addInteraction(&qf10);
  DDFT_IF_Open open10(&dft10, 0.5);
  open10.step();
  open10.step();
  std::printf("open_ddft %.17g %.17g %.17g\n", open10.getF(), d10.get(0), d10.get((8 * 16 + 8) * 16 + 8));

  {
    std::FILE* pf = std::fopen("sphere_points.dat", "w");
    for (int i = 0; i < 60; ++i) {  // Fibonacci sphere
      const double z = 1 - 2 * (i + 0.5) / 60, phi = M_PI * (1 + std::sqrt(5.0)) * (i + 0.5), r = std::sqrt(1 - z * z);
      std::fprintf(pf, "%.17g %.17g %.17g\n", r * std::cos(phi), r * std::sin(phi), z);
    }
    std::fclose(pf);
  }
  Density d11(0.1, L2);
  for (long i = 0; i < d11.Ntot(); ++i) d11.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(101, i) - 1.0)));
  FMT_Species s11(d11, 1.0, 0.0, 7);
  Interaction pts1(&s11, &s11, &lj, 1.0, "sphere_points.dat");
  const double a1 = pts1.getVDWParameter();
  Interaction pts2(&s11, &s11, &lj, 1.0, "sphere_points.dat");
  const double a2 = pts2.getVDWParameter();
  WhiteBearI fmt11;
  DFT dft11(&s11);
  dft11.addHardCoreContribution(&fmt11);
  dft11.addInteraction(&pts2);
  const double F11 = dft11.calculateFreeEnergyAndDerivatives(false);
  std::printf("points %.17g %.17g %d %.17g\n", a1, a2, pts2.weights_from_cache() ? 1 : 0, F11);

  // a hard-sphere mixture: FMT::calculateFreeEnergyAndDerivatives over two FMT_Species of diameters 1 and 0.6
  double Lm[3] = {1.6, 1.6, 2.4};
  Density dma(0.1, Lm), dmb(0.1, Lm);
  for (long i = 0; i < dma.Ntot(); ++i) {
    dma.set(i, 0.3 * (1.0 + 0.1 * (2.0 * splitmix_u(61, i) - 1.0)));
    dmb.set(i, 0.5 * (1.0 + 0.1 * (2.0 * splitmix_u(62, i) - 1.0)));
  }
  FMT_Species sma(dma, 1.0, 0.0), smb(dmb, 0.6, 0.0);
  WhiteBearI fmtm;
  std::vector<Species*> both{&sma, &smb};
  sma.zeroForce();
  smb.zeroForce();
  const double Fm = fmtm.calculateFreeEnergyAndDerivatives(both);
  std::printf("mix %.17g %.17g %.17g\n", Fm, sma.getDF().accu(), smb.getDF().accu());

  // the same two species in a DFT (mu 0.5 and 0.2, a field on the first): one evaluation, then 20 fireMinimizer2
  // steps over both species
  Density dda(0.1, Lm), ddb(0.1, Lm);
  DFT_Vec field(dda.Ntot());
  for (long i = 0; i < dda.Ntot(); ++i) {
    dda.set(i, 0.3 * (1.0 + 0.1 * (2.0 * splitmix_u(63, i) - 1.0)));
    ddb.set(i, 0.4 * (1.0 + 0.1 * (2.0 * splitmix_u(64, i) - 1.0)));
    field.set(i, 0.2 * (2.0 * splitmix_u(65, i) - 1.0));
  }
  dda.set_external_field(field);
  FMT_Species sda(dda, 1.0, 0.5), sdb(ddb, 0.6, 0.2);
  WhiteBearI fmtd;
  DFT dftm(&sda);
  dftm.addSpecies(&sdb);
  dftm.addHardCoreContribution(&fmtd);
  const double Fd = dftm.calculateFreeEnergyAndDerivatives(false);
  std::printf("mix_dft %.17g %.17g %.17g %.17g\n", Fd, dftm.get_f_id(), sda.getDF().accu(), sdb.getDF().accu());
  fireMinimizer2 firem(&dftm);
  firem.set_snapshot_interval(0);
  for (int i = 0; i < 20; ++i) firem.step();
  std::printf("mix_fire %.17g %.17g %.17g %.17g\n", firem.getF(), firem.get_raw_force_max(), dda.getNumberAtoms(),
              ddb.getNumberAtoms());

  // the second species with a fixed mass (Species::setFixedMass): one evaluation and 10 fireMinimizer2 steps
  Density dfa(0.1, Lm), dfb(0.1, Lm);
  for (long i = 0; i < dfa.Ntot(); ++i) {
    dfa.set(i, 0.3 * (1.0 + 0.1 * (2.0 * splitmix_u(66, i) - 1.0)));
    dfb.set(i, 0.4 * (1.0 + 0.1 * (2.0 * splitmix_u(67, i) - 1.0)));
  }
  FMT_Species sfa(dfa, 1.0, 0.3), sfb(dfb, 0.6, 0.5);
  sfb.setFixedMass(0.42 * dfb.getVolume());
  WhiteBearI fmtf;
  DFT dftf(&sfa);
  dftf.addSpecies(&sfb);
  dftf.addHardCoreContribution(&fmtf);
  const double Ff = dftf.calculateFreeEnergyAndDerivatives(false);
  std::printf("mix_mass %.17g %.17g %.17g\n", Ff, sfb.getChemPotential(), dfb.getNumberAtoms());
  fireMinimizer2 firef(&dftf);
  firef.set_snapshot_interval(0);
  for (int i = 0; i < 10; ++i) firef.step();
  std::printf("mix_mass_fire %.17g %.17g %.17g\n", firef.getF(), sfb.getChemPotential(), dfb.getNumberAtoms());
  return 0;
}
