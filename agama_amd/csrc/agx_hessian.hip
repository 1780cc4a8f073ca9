// Batched potential, gradient and Hessian of the Multipole (+ optional NFW) potential: the
// evaluation behind Potential.eval(..., der=True) and agama_evalPotential's deriv2
// (src/interface_python.cpp:3027-3207, src/interface_c.cpp:68-98), one point per lane.
#include <hip/hip_runtime.h>
#include "agx_hessian.h"
#include "agx_kernels.h"

using namespace agxd;

namespace {

__global__ __launch_bounds__(256) void evalHessKernel(const AgxMultipoleDev* __restrict__ Pp, long n,
    const double* __restrict__ pos, double* __restrict__ phi, double* __restrict__ grad,
    double* __restrict__ hess)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const AgxMultipoleDev& P = *Pp;
    const double x = pos[i * 3], y = pos[i * 3 + 1], z = pos[i * 3 + 2];
    FetchVec2 f{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
    double p, g[3], h[6];
    multipoleCarHess(P, f, x, y, z, p, g, h);
    if(phi) phi[i] = p;
    if(grad) for(int c = 0; c < 3; c++) grad[i * 3 + c] = g[c];
    if(hess) for(int c = 0; c < 6; c++) hess[i * 6 + c] = h[c];
}

}  // namespace

hipError_t agxLaunchEvalHess(const AgxMultipoleDev* dP, long n, const double* pos, double* phi,
    double* grad, double* hess, hipStream_t stream)
{
    if(n <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(evalHessKernel, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, stream,
        dP, n, pos, phi, grad, hess);
    return hipGetLastError();
}
