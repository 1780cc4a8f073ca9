// Work order for the persistent orbit kernel: longest expected integration first.
//
// A lane integrates one orbit at a time, so the batch cannot finish before its longest orbit
// does; starting the most expensive orbits first (longest-processing-time-first list
// scheduling) keeps that critical path inside the bulk of the work instead of at its end.
// The expected cost of an orbit is the number of dynamical times it spans,
//   cost = |T| / t_dyn,   t_dyn = sqrt(r / |grad Phi(x0)|)   (circular-orbit period / 2 pi),
// evaluated with the same Multipole kernel, then sorted descending with hipCUB radix sort.
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

}  // namespace

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
