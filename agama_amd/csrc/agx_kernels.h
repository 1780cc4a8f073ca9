// Kernel argument blocks shared by the launchers (agx_capi.hip) and the kernels.
#pragma once
#include "agx_tables.h"

struct AgxOrbitArgs {
    long n;
    const double* ic;        // n x 6 (x, y, z, vx, vy, vz), rotating-frame input as in the reference
    const double* tstart;    // n or null (0)
    const double* ttotal;    // n
    const int* trajsize;     // n: >= 1 rows at regular intervals; -c: every step, c rows reserved;
                             // null: 1 (no rows are written without outTime/outTraj)
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
    // overflow pool of every-step recording (trajsize < 0): rows beyond an orbit's block go to
    // chunks of poolChunk rows, numbered from poolCounter (zero on entry) and placed at row
    // poolBase + chunk * poolChunk of the output buffers; chunkDir[oi * maxChunks + k] holds the
    // chunk of the orbit's k-th overflow chunk (-1 on entry).  poolCounter null: no pool.
    int* chunkDir = nullptr;
    unsigned int* poolCounter = nullptr;
    long poolBase = 0;
    int poolChunk = 0, maxChunks = 0, poolChunks = 0;
    // target plug-ins (agx_target.h), device memory, or null
    const struct AgxTargetList* targets = nullptr;
    // group mode for the longest orbits: the wavefronts whose index is a multiple of groupStride
    // (0: none) integrate the orbits order[0 .. groupOrbits) with AGX_GROUP lanes per orbit,
    // taking them from groupCounter (zero on entry); the other wavefronts take order[groupOrbits
    // ..) from counter
    long groupOrbits = 0;
    int groupStride = 0;
    unsigned int* groupCounter = nullptr;
    // the time each orbit reached (n, indexed by orbit), or null: the work-order probe's output
    double* outTend = nullptr;
};

// Placement of group mode's wavefronts (AgxOrbitArgs::groupStride): one wavefront in `stride`,
// or with AGX_GROUP_BLOCKS every wavefront of the workgroups b with (b / 8) % stride == 0 (whole
// workgroups, eight consecutive ones at a time, which the eight XCDs take in turn), so that the
// group wavefronts share their CUs' memory pipelines with each other only.
#ifndef AGX_GROUP_BLOCKS
#define AGX_GROUP_BLOCKS 0
#endif
__host__ __device__ inline bool agxIsGroupWave(int block, int waveInBlock, int wavesPerBlock, int stride)
{
    if(stride <= 0) return false;
    if(AGX_GROUP_BLOCKS) return (block / 8) % stride == 0;
    return (block * wavesPerBlock + waveInBlock) % stride == 0;
}

enum AgxPrecision { AGX_FP64 = 0, AGX_FP32_TABLES = 1 };

// how the FP64 kernels read the 2-D splines: node records with per-lane 8-byte loads, per-lane
// 16-byte loads or the wavefront-cooperative gather through LDS, or the per-cell polynomial
// tables (agx_device.h)
enum AgxFetch { AGX_FETCH_SCALAR = 0, AGX_FETCH_VEC2 = 1, AGX_FETCH_COOP = 2, AGX_FETCH_CELL = 3 };

// one recorded series - an orbit's trajectory or one of its deviation vectors - behind the
// quintic-spline interpolators of Orbit objects (agx_orbit_interp.hip)
struct AgxSeries {
    const double* t;         // timestamps of the rows
    const double* v;         // rows x 6: position and velocity part, rotating frame
    long rows;               // rows in the block; recording stops at the first NaN timestamp
    long node;               // first node of the series in the node array
};

// spline node: time, position (3), first derivative (3), second derivative (3) in the
// inertial frame
enum : int { AGX_NODE_DOUBLES = 10 };
