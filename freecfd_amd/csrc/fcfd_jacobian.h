// Launchers of the FD-Jacobian kernels (fcfd_jacobian.hip).
#pragma once
#include "fcfd_device.h"

namespace fcfd {
// JL, JR: [25][F] per face (entry (var, eq) at (5*var+eq)*F + f); D: [25][ld]
// diagonal blocks (row-major 5x5 at k = 5*row+col); dt may be null (no time terms).
void launch_fd_jacobian(const DevGrid& g, const Params& P, const double* W, const double* G, const double* PHI,
                        const double* UPD, const double* dt, double* JL, double* JR, double* D, int* status, hipStream_t s);
}
