// Density file round trip through the reference-named API (no GPU work):
//   density_io <in> <out> <dx> <Lx> <Ly> <Lz>
// reads <in> with Density::readDensity, writes it with Density::writeDensity and
// prints "N <number of values> sum <sum of the values read>".
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "classicaldft/classicaldft.h"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: density_io <in> <out> <dx> <Lx> <Ly> <Lz>\n");
    return 2;
  }
  try {
    double L[3] = {std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6])};
    Density d(std::atof(argv[3]), L);
    d.readDensity(argv[1]);
    d.writeDensity(argv[2]);
    double s = 0.0;
    for (long i = 0; i < d.Ntot(); ++i) s += d.get(i);
    std::printf("N %ld sum %.17g\n", d.Ntot(), s);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
