// Work order for the persistent orbit kernel: longest expected integration first.
//
// A lane integrates one orbit at a time, so the batch cannot finish before its longest orbit
// does; starting the most expensive orbits first (longest-processing-time-first list
// scheduling) keeps that critical path inside the bulk of the work instead of at its end.
// The expected cost of an orbit is the number of dynamical times it spans,
//   cost = |T| / t_dyn,   t_dyn = sqrt(r / |grad Phi(x0)|)   (circular-orbit period / 2 pi),
// evaluated with the same Multipole kernel, then sorted descending with hipCUB radix sort.
// That estimate misses orbits whose step rate changes along the orbit (eccentric orbits that
// plunge towards the centre take many times the steps their starting radius suggests), so large
// batches are ranked again from a probe: every orbit is integrated for a few steps and its cost
// is extrapolated from the time those steps covered.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include "agx_device.h"
#include "agx_kernels.h"

using namespace agxd;

namespace {

__global__ __launch_bounds__(256) void costKernel(const AgxMultipoleDev* __restrict__ Pp, long n, const double* __restrict__ ic, const double* __restrict__ ttotal,
    float* __restrict__ key, int* __restrict__ idx)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const double x = ic[i * 6], y = ic[i * 6 + 1], z = ic[i * 6 + 2];
    double p, gx, gy, gz;
    FetchVec2 f{ gmem(reinterpret_cast<const vd2*>(Pp->nodes)) };
    multipoleCar(*Pp, f, x, y, z, p, gx, gy, gz);
    const double r = sqrt(x * x + y * y + z * z), g = sqrt(gx * gx + gy * gy + gz * gz);
    double c = fabs(ttotal[i]) * sqrt(g / fmax(r, 1e-300));
    if(!(c >= 0 && c < 3e38)) c = 3e38;   // NaN / inf inputs go first and fail fast
    key[i] = (float)c;
    idx[i] = (int)i;
}

// the probe's estimate of an orbit's cost: the steps it would take at the rate of its first
// probeSteps steps, probeSteps * |T| / |t_reached - t_start| (an orbit that the probe completed
// costs at most probeSteps; one that made no progress goes first and fails fast)
__global__ __launch_bounds__(256) void probeKeyKernel(long n, const double* __restrict__ tstart,
    const double* __restrict__ ttotal, const double* __restrict__ treached, int probeSteps,
    float* __restrict__ key, int* __restrict__ idx)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const double T = fabs(ttotal[i]), dt = fabs(treached[i] - (tstart ? tstart[i] : 0.0));
    double c = dt >= T ? (double)probeSteps : probeSteps * (T / dt);
    if(!(c >= 0 && c < 3e38)) c = 3e38;
    key[i] = (float)c;
    idx[i] = (int)i;
}

// the group wavefronts' share of a probed order, dealt round-robin: entry 8k + j of the first
// 8 * groupWaves entries takes rank j * groupWaves + k.  The groups of a wavefront take eight
// consecutive entries at the start, so each group wavefront receives one of the groupWaves
// longest orbits and seven shorter ones, instead of one wavefront receiving the eight longest
// (it would run all eight groups, in turn where their branches differ, for the longest orbit's
// whole duration).  One block; 8 * groupWaves <= AGX_DEAL_MAX.
constexpr int AGX_DEAL_MAX = 8192;
__global__ __launch_bounds__(1024) void dealGroupsKernel(int* __restrict__ order, int groupWaves)
{
    __shared__ int s[AGX_DEAL_MAX];
    const int m = 8 * groupWaves;
    for(int i = threadIdx.x; i < m; i += blockDim.x) s[i] = order[i];
    __syncthreads();
    for(int i = threadIdx.x; i < m; i += blockDim.x) order[i] = s[(i & 7) * groupWaves + (i >> 3)];
}

}  // namespace

// deal the first groupOrbits entries of `order` over the group wavefronts (dealGroupsKernel);
// groupOrbits: a multiple of 8, at most 8192
hipError_t agxLaunchDealGroups(int* order, long groupOrbits, hipStream_t stream)
{
    if(groupOrbits <= 0 || groupOrbits % 8 != 0 || groupOrbits > AGX_DEAL_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(dealGroupsKernel, dim3(1), dim3(1024), 0, stream, order, (int)(groupOrbits / 8));
    return hipGetLastError();
}

// bytes of workspace needed by agxLaunchSchedule for n orbits
size_t agxScheduleWorkspace(long n)
{
    size_t cub = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, cub, (const float*)nullptr, (float*)nullptr,
        (const int*)nullptr, (int*)nullptr, (int)n);
    const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    return 3 * a + cub + 256;
}

// writes the processing order into `order` (n ints); `work` holds agxScheduleWorkspace(n) bytes
hipError_t agxLaunchSchedule(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, long n,
    const double* ic, const double* ttotal, int* order, void* work, hipStream_t stream)
{
    if(n <= 0) return hipSuccess;
    const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    char* w = static_cast<char*>(work);
    float* keyIn = reinterpret_cast<float*>(w);
    float* keyOut = reinterpret_cast<float*>(w + a);
    int* idxIn = reinterpret_cast<int*>(w + 2 * a);
    void* cubTemp = w + 3 * a;
    size_t cubBytes = agxScheduleWorkspace(n) - 3 * a - 256;
    const int block = 256;
    hipLaunchKernelGGL(costKernel, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, stream,
        dP, n, ic, ttotal, keyIn, idxIn);
    hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairsDescending(cubTemp, cubBytes, keyIn, keyOut, idxIn, order,
        (int)n, 0, 32, stream);
}

// Work order from a probe: after agxLaunchSchedule, the caller integrates every orbit for at most
// probeSteps steps (in that order), writing the time each reached into agxScheduleProbeTimes(work,
// n); this ranks the orbits by the steps they would take at that rate (longest first) and returns
// the order (n ints inside the workspace; the first order is used as scratch)
double* agxScheduleProbeTimes(void* work, long n)
{
    const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    return reinterpret_cast<double*>(static_cast<char*>(work) + a);   // keyOut and idxIn (8n bytes)
}

hipError_t agxLaunchScheduleFromProbe(long n, const double* tstart, const double* ttotal, int probeSteps,
    int* order, void* work, int** finalOrder, hipStream_t stream)
{
    const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
    char* w = static_cast<char*>(work);
    float* keyIn = reinterpret_cast<float*>(w);
    float* keyOut = reinterpret_cast<float*>(w + a);
    int* idxOut = reinterpret_cast<int*>(w + 2 * a);
    void* cubTemp = w + 3 * a;
    size_t cubBytes = agxScheduleWorkspace(n) - 3 * a - 256;
    *finalOrder = idxOut;
    if(n <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(probeKeyKernel, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, stream,
        n, tstart, ttotal, agxScheduleProbeTimes(work, n), probeSteps, keyIn, order);
    hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairsDescending(cubTemp, cubBytes, keyIn, keyOut, order, idxOut,
        (int)n, 0, 32, stream);
}
