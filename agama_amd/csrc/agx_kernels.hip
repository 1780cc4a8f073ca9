// HIP kernels for gfx950: batched Multipole evaluation and the persistent DOP853 orbit
// integrator.
//
// Orbit kernel design (one orbit per lane, FP64 end to end):
//   - persistent: each lane pulls orbits from a global work queue (atomic counter) in the
//     order given by the host (longest expected integrations first) until the queue is empty;
//   - every loop iteration performs exactly ONE right-hand-side evaluation per active lane.
//     A per-lane phase (3 initialisation evaluations, 11 Runge-Kutta stages, 1 FSAL stage)
//     decides where the evaluation point comes from and where the result goes.  The
//     Multipole evaluation, ~90% of the work, is therefore a single code path executed by
//     the whole wavefront, even when lanes are at different points of their integrations
//     (a lane that starts a new orbit is only out of phase in the cheap stage sums);
//   - stage vectors live in registers with the reference's storage reuse (k11..k13 and the
//     b-sum reuse the registers of k2..k5, src/math_ode.cpp:241-258);
//   - dense output is formed lazily, only for steps that contain a trajectory sample.
// The integrator restates src/math_ode.cpp:35-423 (initTimeStep, DOP853 init/doStep/getSol),
// src/orbit.cpp:15-133, 182-208 (trajectory sampling, run loop, rotating-frame RHS).
#include <hip/hip_runtime.h>
#include <type_traits>
#include "agx_device.h"
#include "agx_kernels.h"
#include "agx_dop853.h"

using namespace agxd;

namespace {

// ---------------- batched potential / gradient ----------------
// minimum wavefronts per SIMD the evaluation kernel is compiled for
#ifndef AGX_EVAL_WAVES
#define AGX_EVAL_WAVES 2
#endif
// The table header lives in device memory (uniform pointer -> scalar loads); passing it by
// value would make the compiler copy its arrays into per-lane scratch at every launch.
// MODE: AgxFetch for FP64 tables, or -1 for the FP32 tables.  Every lane of the grid takes
// part in the evaluation (out-of-range lanes evaluate their block's first point and discard
// it) so that the cooperative fetcher sees wave-uniform control flow.
template<int MODE, bool NFW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AGX_EVAL_WAVES))) void evalKernel(const AgxMultipoleDev* __restrict__ Pp,
    long n, const double* __restrict__ pos, double* __restrict__ phi, double* __restrict__ grad)
{
    __shared__ vd2 stage[MODE == AGX_FETCH_COOP ? 4 * FetchCoop::LDS_DOUBLE2_PER_WAVE : 1];
    const AgxMultipoleDev& P = *Pp;
    // short-lived blocks: the grids are read from global memory (staging them costs more)
    const double AGX_GLOBAL* gR = gmem(P.gridR);
    const double AGX_GLOBAL* gT = gmem(P.gridT);
    const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool inRange = i0 < n;
    const long i = inRange ? i0 : (long)blockIdx.x * blockDim.x;
    const double x = pos[i * 3], y = pos[i * 3 + 1], z = pos[i * 3 + 2];
    double p, gx, gy, gz;
    if constexpr(MODE == AGX_FETCH_COOP) {
        FetchCoop f{ gmem(reinterpret_cast<const vd2*>(P.nodes)),
            (vd2 AGX_LDS*)(stage + (threadIdx.x >> 6) * FetchCoop::LDS_DOUBLE2_PER_WAVE),
            (int)(threadIdx.x & 63) };
        multipoleCar<decltype(f), const double AGX_GLOBAL*, NFW>(P, f, x, y, z, p, gx, gy, gz, gR, gT);
    } else if constexpr(MODE == AGX_FETCH_CELL) {
        FetchCell f(P.cells, P.cellBytes);
        multipoleCar<decltype(f), const double AGX_GLOBAL*, NFW>(P, f, x, y, z, p, gx, gy, gz, gR, gT);

    } else if constexpr(MODE == AGX_FETCH_VEC2) {
        FetchVec2 f{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
        multipoleCar<decltype(f), const double AGX_GLOBAL*, NFW>(P, f, x, y, z, p, gx, gy, gz, gR, gT);
    } else if constexpr(MODE == AGX_FETCH_SCALAR) {
        FetchScalar f{ gmem(P.nodes) };
        multipoleCar<decltype(f), const double AGX_GLOBAL*, NFW>(P, f, x, y, z, p, gx, gy, gz, gR, gT);
    } else {
        FetchFloat f{ gmem(P.nodesF) };
        multipoleCar<decltype(f), const double AGX_GLOBAL*, NFW>(P, f, x, y, z, p, gx, gy, gz, gR, gT);
    }
    if(!inRange) return;
    if(phi) phi[i] = p;
    if(grad) { grad[i * 3] = gx; grad[i * 3 + 1] = gy; grad[i * 3 + 2] = gz; }
}

// Threads per workgroup of the orbit kernel (fixed: it sizes the LDS block below).
constexpr int OBS = 256;
#ifndef AGX_K_IN_LDS
#define AGX_K_IN_LDS 0
#endif
#ifndef AGX_ALIGN_PHASES
#define AGX_ALIGN_PHASES 1
#endif

// whether the orbit kernel requests the next harmonic's records before evaluating the current
// one (more registers in flight, one L2 round trip fewer on the critical path per harmonic)
#ifndef AGX_ORBIT_PREFETCH
#define AGX_ORBIT_PREFETCH 1
#endif

// the ln r and tau grids of the orbit kernel, staged in LDS (at most this many values together)
constexpr int AGX_LDS_GRID_ORBIT = 512;
// LDS for the per-cell path's power-law arrays (10 (legLmax+1)^2 + legLmax + 1 doubles: legLmax <= 16)
constexpr int AGX_LDS_PL_ORBIT = 2944;
// LDS for the tau bucket table of the per-cell path (the host builds at most 4096 buckets)
constexpr int AGX_LDS_TAULUT_ORBIT = 4096;
typedef const double AGX_LDS* LdsGrid;

// one right-hand-side potential gradient in the inertial frame (orbit.cpp:182-208)
template<int MODE, bool NFW>
__device__ __forceinline__ void rhsEval(const AgxMultipoleDev& P, vd2* stage, double x, double y,
    double z, double& pot, double& gx, double& gy, double& gz, LdsGrid gR, LdsGrid gT, const PLTabs<LdsGrid>& plt)
{
    if constexpr(MODE == AGX_FETCH_COOP) {
        FetchCoop f{ gmem(reinterpret_cast<const vd2*>(P.nodes)),
            (vd2 AGX_LDS*)(stage + (threadIdx.x >> 6) * FetchCoop::LDS_DOUBLE2_PER_WAVE),
            (int)(threadIdx.x & 63) };
        multipoleCar<decltype(f), LdsGrid, NFW, AGX_ORBIT_PREFETCH != 0>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
    } else if constexpr(MODE == AGX_FETCH_CELL) {
        FetchCell f(P.cells, P.cellBytes);
        if(P.interp1d) multipoleCar<decltype(f), LdsGrid, NFW>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
        else multipoleCarCell<decltype(f), LdsGrid, NFW>(P, f, x, y, z, pot, gx, gy, gz, gR, gT, plt);

    } else if constexpr(MODE == AGX_FETCH_VEC2) {
        FetchVec2 f{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
        multipoleCar<decltype(f), LdsGrid, NFW, AGX_ORBIT_PREFETCH != 0>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
    } else if constexpr(MODE == AGX_FETCH_SCALAR) {
        FetchScalar f{ gmem(P.nodes) };
        multipoleCar<decltype(f), LdsGrid, NFW, AGX_ORBIT_PREFETCH != 0>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
    } else {
        FetchFloat f{ gmem(P.nodesF) };
        multipoleCar<decltype(f), LdsGrid, NFW, AGX_ORBIT_PREFETCH != 0>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
    }
}

// WAVES: wavefronts per SIMD the kernel is compiled for (register budget 512 / WAVES); both
// variants exist so that the autotuner can choose at run time.
// TGT: the variant that also feeds target plug-ins (agx_target.h)
// GRP: the variant with group mode (A.groupStride): some wavefronts integrate the longest orbits
// with AGX_GROUP lanes per orbit, each lane evaluating a share of the harmonics
// The arguments are read through a pointer to device memory (written by storeArgsKernel just
// before): passed by value, the whole block would be loaded into scalar registers at entry and
// kept live for the whole kernel, which spilled scalar registers into vector-register lanes.
template<int MODE, int WAVES, bool NFW, bool TGT = false, bool GRP = false>
__global__ __launch_bounds__(OBS) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
void orbitKernel(const AgxMultipoleDev* __restrict__ Pp, const AgxOrbitArgs* __restrict__ Ap)
{
    const AgxOrbitArgs& A = *Ap;
    // at one wavefront per SIMD the block has the CU's LDS to itself: the stage vectors k1..k5
    // and x live there too (the cooperative fetcher needs its own LDS staging area instead)
    typedef typename std::conditional<TGT, LaneTgtT<OBS>,
        LaneT<OBS, AGX_K_IN_LDS && WAVES == 1 && MODE != AGX_FETCH_COOP> >::type Lane;
    __shared__ double lds[Lane::LDS_FIELDS * OBS];
    __shared__ vd2 stage[MODE == AGX_FETCH_COOP ? (OBS / 64) * FetchCoop::LDS_DOUBLE2_PER_WAVE : 1];
    __shared__ double gridLds[AGX_LDS_GRID_ORBIT];
    // the per-cell path's power-law extrapolation arrays (the launcher checks that they fit)
    constexpr bool CELLS = MODE == AGX_FETCH_CELL;
    __shared__ double plLds[CELLS ? AGX_LDS_PL_ORBIT : 1];
    __shared__ int lutLds[CELLS ? AGX_LDS_TAULUT_ORBIT : 1];
    const AgxMultipoleDev& P = *Pp;
    // the grids in LDS (the launcher refuses larger ones): binary searches read LDS directly
    for(int i = threadIdx.x; i < P.nr + P.nt; i += blockDim.x)
        gridLds[i] = i < P.nr ? P.gridR[i] : P.gridT[i - P.nr];
    PLTabs<LdsGrid> plt{};
    if constexpr(CELLS) { if(!P.interp1d) plt = plStage(P, (double AGX_LDS*)plLds, (int AGX_LDS*)lutLds); }
    __syncthreads();
    const LdsGrid gR = (LdsGrid)gridLds, gT = gR + P.nr;
    Lane L;
    L.attach(lds + threadIdx.x);
    // wave-uniform: whether this wavefront works in group mode
    const bool group = GRP && agxIsGroupWave((int)blockIdx.x, (int)(threadIdx.x >> 6), OBS / 64, A.groupStride);
    fetchOrbit(A, L, group);
    const double accRel = A.accuracy;
    unsigned int it = 0;
    while(true) {
        const bool active = L.phase != PH_IDLE;
        if(!__any(active)) break;
        // Stage sequences are aligned to the wavefront's iteration counter: an attempt's first
        // stage starts when it % 12 == 0 (a lane that has just initialised waits), so all lanes
        // take the same branch of the phase machine below instead of several in turn.
        const bool run = active && !(L.phase == PH_S2 && AGX_ALIGN_PHASES && it % 12 != 0);
        it++;
        // ---- one right-hand-side evaluation (OrbitIntegrator<Car>::eval, orbit.cpp:182-208),
        // executed by every lane of the wavefront (idle lanes evaluate a placeholder point)
        double ca = 1, sa = 0;
        if(A.Omega != 0 && active) sincosA(A.Omega * (L.tb + L.toff), sa, ca);
        double pot, gx, gy, gz;
        if constexpr(GRP) {
            if(group) {
                if constexpr(MODE == AGX_FETCH_CELL) {
                    FetchCell fg(P.cells, P.cellBytes);
                    multipoleCarCellGroup<AGX_GROUP>(P, fg, L.y[0] * ca + L.y[1] * sa, L.y[1] * ca - L.y[0] * sa,
                        L.y[2], pot, gx, gy, gz, gR, gT, plt, (int)(threadIdx.x & (AGX_GROUP - 1)));
                } else {
                    FetchVec2 fg{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
                    multipoleCarGroup<AGX_GROUP>(P, fg, L.y[0] * ca + L.y[1] * sa, L.y[1] * ca - L.y[0] * sa,
                        L.y[2], pot, gx, gy, gz, gR, gT, (int)(threadIdx.x & (AGX_GROUP - 1)));
                }
            } else
                rhsEval<MODE, NFW>(P, stage, L.y[0] * ca + L.y[1] * sa, L.y[1] * ca - L.y[0] * sa, L.y[2],
                    pot, gx, gy, gz, gR, gT, plt);
        } else
            rhsEval<MODE, NFW>(P, stage, L.y[0] * ca + L.y[1] * sa, L.y[1] * ca - L.y[0] * sa, L.y[2],
                pot, gx, gy, gz, gR, gT, plt);
        if(!run) continue;
        double f[6];
        f[0] = L.y[3]; f[1] = L.y[4]; f[2] = L.y[5];
        f[3] = -gx * ca + gy * sa;
        f[4] = -gy * ca - gx * sa;
        f[5] = -gz;
        advancePhase(A, L, f, pot, accRel);
        if(L.phase == PH_IDLE) fetchOrbit(A, L, group);
    }
}

}  // namespace

// ---------------- launchers (called from agx_capi.hip) ----------------
// P: host copy of the header, dP: the same header in device memory
hipError_t agxLaunchEval(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, int precision, int fetch,
    long n, const double* pos, double* phi, double* grad, hipStream_t stream)
{
    (void)P;
    if(n <= 0) return hipSuccess;
    const int block = 256;
    const dim3 grid((unsigned)((n + block - 1) / block));
    if(P.nfwMass != 0) {   // composite with an NFW halo: per-cell tables, 16-byte per-lane fetch or FP32 tables
        if(precision == AGX_FP32_TABLES)
            hipLaunchKernelGGL((evalKernel<-1, true>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
        else if(fetch == AGX_FETCH_CELL && !P.interp1d)
            hipLaunchKernelGGL((evalKernel<AGX_FETCH_CELL, true>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
        else
            hipLaunchKernelGGL((evalKernel<AGX_FETCH_VEC2, true>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
    } else if(precision == AGX_FP32_TABLES)
        hipLaunchKernelGGL((evalKernel<-1, false>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
    else if(fetch == AGX_FETCH_CELL && !P.interp1d)
        hipLaunchKernelGGL((evalKernel<AGX_FETCH_CELL, false>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
    else if(fetch == AGX_FETCH_COOP)
        hipLaunchKernelGGL((evalKernel<AGX_FETCH_COOP, false>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
    else if(fetch == AGX_FETCH_VEC2)
        hipLaunchKernelGGL((evalKernel<AGX_FETCH_VEC2, false>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
    else
        hipLaunchKernelGGL((evalKernel<AGX_FETCH_SCALAR, false>), grid, dim3(block), 0, stream, dP, n, pos, phi, grad);
    return hipGetLastError();
}

int agxOrbitKernelBlockSize() { return OBS; }

namespace {
typedef void (*OrbitKernelFn)(const AgxMultipoleDev*, const AgxOrbitArgs*);
OrbitKernelFn orbitKernelFor(int precision, int fetch, int waves, bool nfw, bool targets = false, bool group = false)
{
    if(group)     // group mode: FP64 (node records or per-cell tables), 1 wave/SIMD, no NFW, no targets
        return fetch == AGX_FETCH_CELL ? orbitKernel<AGX_FETCH_CELL, 1, false, false, true>
                                       : orbitKernel<AGX_FETCH_VEC2, 1, false, false, true>;
    if(targets)   // target plug-ins: FP64, 16-byte per-lane fetch, 1 wave/SIMD
        return nfw ? orbitKernel<AGX_FETCH_VEC2, 1, true, true> : orbitKernel<AGX_FETCH_VEC2, 1, false, true>;
    if(nfw)   // composite with an NFW halo: per-cell tables or 16-byte per-lane fetch (or FP32 tables), 1 wave/SIMD
        return precision == AGX_FP32_TABLES ? orbitKernel<-1, 1, true> :
               fetch == AGX_FETCH_CELL ? orbitKernel<AGX_FETCH_CELL, 1, true> : orbitKernel<AGX_FETCH_VEC2, 1, true>;
    const bool two = waves == 2;
    if(precision == AGX_FP32_TABLES) return two ? orbitKernel<-1, 2, false> : orbitKernel<-1, 1, false>;
    if(fetch == AGX_FETCH_CELL) return two ? orbitKernel<AGX_FETCH_CELL, 2, false> : orbitKernel<AGX_FETCH_CELL, 1, false>;
    if(fetch == AGX_FETCH_COOP) return two ? orbitKernel<AGX_FETCH_COOP, 2, false> : orbitKernel<AGX_FETCH_COOP, 1, false>;
    if(fetch == AGX_FETCH_VEC2) return two ? orbitKernel<AGX_FETCH_VEC2, 2, false> : orbitKernel<AGX_FETCH_VEC2, 1, false>;
    return two ? orbitKernel<AGX_FETCH_SCALAR, 2, false> : orbitKernel<AGX_FETCH_SCALAR, 1, false>;
}

__global__ void storeArgsKernel(AgxOrbitArgs A, AgxOrbitArgs* dst)
{
    if(threadIdx.x == 0) *dst = A;
}

__global__ void fillU32Kernel(unsigned int* p, long n, unsigned int v)
{
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
}  // namespace

int agxOrbitKernelBlocksPerCU(int precision, int fetch, int waves, bool nfw, bool targets)
{
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitKernelFor(precision, fetch, waves, nfw,
        targets), OBS, 0);
    return e == hipSuccess ? nb : 1;
}

hipError_t agxLaunchOrbit(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, int precision, int fetch,
    int waves, const AgxOrbitArgs& A, int numBlocks, int blockSize, hipStream_t stream, AgxOrbitArgs* dArgs)
{
    (void)P;
    if(A.n <= 0) return hipSuccess;
    if(blockSize != OBS) return hipErrorInvalidValue;
    if(A.targets && precision == AGX_FP32_TABLES) return hipErrorInvalidValue;
    const bool group = A.groupStride > 0;
    if(group && (precision != AGX_FP64 || (fetch != AGX_FETCH_VEC2 && fetch != AGX_FETCH_CELL) || waves != 1 ||
                 P.nfwMass != 0 || P.interp1d || A.targets ||
                 A.poolCounter || !A.groupCounter || A.groupOrbits > A.n))
        return hipErrorInvalidValue;
    if(!dArgs || P.nr + P.nt > AGX_LDS_GRID_ORBIT) return hipErrorInvalidValue;
    if(fetch == AGX_FETCH_CELL && !P.interp1d &&
       ((6 + AGX_LEG_STRIDE) * (P.legLmax + 1) * (P.legLmax + 1) + P.legLmax + 1 > AGX_LDS_PL_ORBIT ||
        P.nTauLut > AGX_LDS_TAULUT_ORBIT))
        fetch = AGX_FETCH_VEC2;   // power-law arrays or bucket table too large for the kernel's LDS copy
    hipLaunchKernelGGL(storeArgsKernel, dim3(1), dim3(64), 0, stream, A, dArgs);
    hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL(orbitKernelFor(precision, fetch, waves, P.nfwMass != 0, A.targets != nullptr, group),
        dim3(numBlocks), dim3(OBS), 0, stream, dP, static_cast<const AgxOrbitArgs*>(dArgs));
    return hipGetLastError();
}

// a fill kernel instead of hipMemsetAsync: graph-capturable and safe to replay
hipError_t agxLaunchFillU32(unsigned int* p, long n, unsigned int v, hipStream_t stream)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fillU32Kernel, dim3(1), dim3(64), 0, stream, p, n, v);
    return hipGetLastError();
}
