// Host sanitizer driver for the reference-named C++ API (classicaldft_amd/host/classicaldft.cpp).
//
// tests/test_sanitizers.py builds it together with classicaldft.cpp under AddressSanitizer and
// UndefinedBehaviorSanitizer (linked against the uninstrumented libcdft.so) and runs the host-side parts that need
// no GPU: DFT_Vec arithmetic, Lattice indexing with periodic wrap, the density file format (round trip and the
// errors for a missing file and a wrong size), species construction and its lattice check, and the potentials'
// hard-sphere diameters and van der Waals parameters.  Prints "sanitize host: ok" at the end.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "classicaldft/classicaldft.h"

namespace {
int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("sanitize host: FAILED %s\n", what);
    ++failures;
  }
}
}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  DFT_Vec a(10), b(10);
  for (long i = 0; i < 10; ++i) {
    a.set(i, 0.5 * i);
    b.set(i, 1.0);
  }
  a.IncrementBy_Scaled_Vector(b, 2.0);
  a.MultBy(0.5);
  expect(std::fabs(a.dotWith(b) - (0.25 * 45 + 10)) < 1e-12, "DFT_Vec arithmetic");
  expect(std::fabs(a.inf_norm() - (2.25 + 1)) < 1e-12, "DFT_Vec inf_norm");

  double L[3] = {1.2, 1.0, 1.4};
  Density d(0.1, L);
  expect(d.Ntot() == 12L * 10 * 14, "lattice size");
  expect(d.get_PBC_Pos(-1, 10, 15) == d.pos(11, 0, 1), "periodic wrap");
  for (long i = 0; i < d.Ntot(); ++i) d.set(i, 0.3 + 1e-3 * (i % 17));
  const std::string f = dir + "/sanitize_density.dat";
  d.writeDensity(f);
  Density e(0.1, L);
  e.readDensity(f.c_str());
  double diff = 0.0;
  for (long i = 0; i < d.Ntot(); ++i) diff = std::fmax(diff, std::fabs(d.get(i) - e.get(i)));
  expect(diff == 0.0, "density round trip");
  bool threw = false;
  try {
    e.readDensity((dir + "/no_such_file.dat").c_str());
  } catch (const std::runtime_error&) {
    threw = true;
  }
  expect(threw, "missing density file");
  double L2[3] = {1.2, 1.0, 1.2};
  Density g(0.1, L2);
  threw = false;
  try {
    g.readDensity(f.c_str());
  } catch (const std::runtime_error&) {
    threw = true;
  }
  expect(threw, "density file of another size");

  FMT_Species s(d, 1.0, -1.0);
  expect(s.getHSD() == 1.0, "species hsd");
  threw = false;
  try {
    FMT_Species bad(d, 0.05, 0.0);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  expect(threw, "hsd below the lattice spacing");

  LJ lj(1.0, 1.0, 2.5);
  tWF twf(1.0, 1.0, 2.5, 50.0);
  WHDF whdf(1.0, 1.0, 2.0);
  for (const Potential1* p : {static_cast<const Potential1*>(&lj), static_cast<const Potential1*>(&twf),
                              static_cast<const Potential1*>(&whdf)}) {
    const double h = p->getHSD(0.8), v = p->getVDW_Parameter(0.8);
    expect(h > 0.8 && h < 1.2 && std::isfinite(v), "potential hsd / vdw");
  }
  if (failures) return 1;
  std::printf("sanitize host: ok\n");
  return 0;
}
