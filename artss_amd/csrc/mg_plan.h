// Device-resident description of the coarse end of a V-cycle.
//
// The levels from `first` down to the coarsest are small (a few thousand
// cells), where a kernel launch costs more than the work it carries.  One
// workgroup runs that whole part of VCycleMG::VCycleMultigrid
// (VCycleMG.cpp:215-288) in a single launch, with workgroup barriers where
// the reference finishes one operator before starting the next, and the
// coarse solve's residual test (VCycleMG.cpp:725-798) on the device instead
// of a host round trip per iteration.
#pragma once

#include "device.h"

namespace artss {

constexpr int kMaxTailLevels = 12;
constexpr int kMaxTailWaves = 16;
constexpr int kTailThreads = 1024;

struct TailLevel {
    LevelDesc L;
    DomainBC bc;                      // pressure ghost fill of this level
    const FaceEntry *faces;           // obstacle face writes, grouped in waves
    const FaceParam *params;
    int n_waves;
    long long wave_start[kMaxTailWaves + 1];
    double ax, ay, az, beta;          // smoother coefficients
    double *err1, *tmp, *res0, *res1, *err0;
    // Offsets (in doubles) into the kernel's dynamic LDS when the level's
    // arrays live there, -1 when they stay in HBM.  Order: err1, tmp, res0,
    // res1, err0.  At the first level res1 holds a copy of the right-hand side.
    int lds_off[5];
};

struct TailPlan {
    int first;        // finest level run by the kernel
    int last;         // n_level: the coarsest level
    int n_relax;
    int cgs;          // smoother: 0 weighted Jacobi, 1 colored Gauss-Seidel
    double w;         // smoother weight
    double dsign;
    int max_iter;     // coarse solve: iteration cap
    double tol;       // coarse solve: residual target
    int solve_smooths; // the coarse solve degenerates to a smoothing (Nx <= 4 and Ny <= 4)
    int lds_doubles;   // dynamic LDS size (doubles) for level arrays; 0: all in HBM
    double *io_err1;   // HBM error of the first level, written back when it lives in LDS
    TailLevel lv[kMaxTailLevels];  // index: level - first
};

}  // namespace artss
