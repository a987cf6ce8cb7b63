// Host generator of the FMT cell weights by the reference's own double-precision expressions.
#pragma once

namespace cdft {

// 11 arrays of Nx*Ny*Nz doubles into w (include/Species.h:351-360 order), for spheres of diameter hsd.
// Throws std::runtime_error if hsd is less than the smallest lattice spacing (src/FMT_Species.cpp:321-322).
void reference_weights(int Nx, int Ny, int Nz, double dx, double dy, double dz, double hsd, double* w);

}  // namespace cdft
