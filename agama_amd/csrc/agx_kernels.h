// Kernel argument blocks shared by the launchers (agx_capi.hip) and the kernels.
#pragma once
#include "agx_tables.h"

struct AgxOrbitArgs {
    long n;
    const double* ic;        // n x 6 (x, y, z, vx, vy, vz), rotating-frame input as in the reference
    const double* tstart;    // n or null (0)
    const double* ttotal;    // n
    const int* trajsize;     // n (>= 1)
    const long* trajOffset;  // n: first output row of each orbit
    const int* order;        // n or null: processing order (a permutation of 0..n-1)
    double accuracy, Omega;
    long maxNumSteps;
    double* outTime;         // rows, or null
    double* outTraj;         // rows x 6, or null
    double* outFinal;        // n x 6, or null
    long* outNsteps;         // n, or null
    long* outNattempts;      // n, or null
    long* outNevals;         // n, or null
    double* outMaxde;        // n, or null
    int* outStatus;          // n, or null
    unsigned int* counter;   // work-queue head, zero on entry
};

enum AgxPrecision { AGX_FP64 = 0, AGX_FP32_TABLES = 1 };

// how FP64 node records are fetched: per-lane 8-byte loads, per-lane 16-byte loads, or the
// wavefront-cooperative gather through LDS (agx_device.h)
enum AgxFetch { AGX_FETCH_SCALAR = 0, AGX_FETCH_VEC2 = 1, AGX_FETCH_COOP = 2 };
