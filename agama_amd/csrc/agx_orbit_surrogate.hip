// Persistent DOP853 orbit kernel with the MLP surrogate force (gfx950).
//
// Same integrator and work queue as the FP64 orbit kernel (agx_dop853.h), one orbit per lane,
// one wavefront per workgroup.  Every loop iteration evaluates the surrogate for the whole
// wavefront: the 64 lanes' points go through the fused MFMA network in two passes of 32
// (agx_mlp.h), the l <= 2 FP64 base is evaluated per lane, and the force is assembled in FP64.
// The FP64 Multipole stays the judge: it gives the initial energy and the energy after every
// accepted step, so the recorded |E/E0 - 1| is the true energy error of the surrogate orbit.
// The three evaluations that start an orbit (initial energy, first-step estimate) use the FP64
// Multipole, so the first step size is free of the network's rounding noise.
// An orbit falls back - it is handed over to the FP64 kernel and integrated again from its
// initial conditions - as soon as that error exceeds the threshold, an evaluation point
// leaves the training radii, or an accepted step is 64 times shorter than the first one (the
// bf16 rounding in the network makes its force rough on small scales, and the error control
// can answer that with ever shorter steps).
//
// The stage sequence of every lane is aligned to the wavefront's iteration counter (the first
// stage of an attempt starts when it % 12 == 0; a lane that has just initialised waits), so
// the lanes reach the step's last evaluation - the one that needs the FP64 potential - in the
// same iteration, and the FP64 evaluation runs once per 12 iterations rather than in most.
#include <hip/hip_runtime.h>
#include "agx_device.h"
#include "agx_dop853.h"
#include "agx_mlp.h"
#include "agx_surrogate.h"

using namespace agxd;

namespace {

constexpr int SBS = 64;   // one wavefront per workgroup

__global__ __launch_bounds__(SBS) void orbitSurrogateKernel(const AgxMultipoleDev* __restrict__ Pp,
    const AgxMultipoleDev* __restrict__ Bp, const AgxMlpDev* __restrict__ Mp, AgxOrbitArgs A,
    AgxSurrogateOrbitArgs SA)
{
    __shared__ double lds[F_COUNT * SBS];
    __shared__ agx_bf16x8 sbuf[AGX_MLP_LDS_VEC];
    const AgxMultipoleDev& P = *Pp;
    const AgxMultipoleDev& B = *Bp;
    const AgxMlpDev& M = *Mp;
    const int lane = threadIdx.x & 63;
    LaneT<SBS> L;
    L.attach(lds + threadIdx.x);
    fetchOrbit(A, L);
    const double accRel = A.accuracy, rmin2 = SA.rmin * SA.rmin, rmax2 = SA.rmax * SA.rmax;
    const double kOut = M.outScale;
    unsigned int it = 0;
    double hRef = 0;   // first step of the lane's current orbit
    while(true) {
        const bool active = L.phase != PH_IDLE;
        if(!__any(active)) break;
        // a pilot launch ends as soon as its decision is known (enough orbits have fallen back);
        // the count is read with a device-scope atomic load, every 12 iterations
        if(SA.abortAt && it % 12 == 0 &&
           __any(__hip_atomic_load(SA.nfallback, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= SA.abortAt))
            break;
        const bool run = active && !(L.phase == PH_S2 && it % 12 != 0);
        it++;
        if(!__any(run)) continue;
        double ca = 1, sa = 0;
        if(A.Omega != 0 && active) sincosA(A.Omega * (L.tb + L.toff), sa, ca);
        const double x = L.y[0] * ca + L.y[1] * sa, y = L.y[1] * ca - L.y[0] * sa, z = L.y[2];
        // ---- surrogate: the network for the wavefront's points in two passes of 32
        float feat[4], out = 0.f, df[4] = { 0.f, 0.f, 0.f, 0.f };
        mlpFeatures(M, x, y, z, feat);
#pragma unroll 1
        for(int pass = 0; pass < 2; pass++) {
            if(!__any(run && (lane >> 5) == pass)) continue;
            const int src = pass * 32 + (lane & 31);
            float fp[4], o, d[4];
            for(int c = 0; c < 4; c++) fp[c] = __shfl(feat[c], src);
            mlpWave(M, fp, (agx_lds_bf16x8*)sbuf, o, d);
            if((lane >> 5) == pass) {
                out = o;
                for(int c = 0; c < 4; c++) df[c] = d[c];
            }
        }
        double pb, gbx, gby, gbz, ox, oy, oz;
        {
            FetchVec2 fb{ gmem(reinterpret_cast<const vd2*>(B.nodes)) };
            multipoleCar(B, fb, x, y, z, pb, gbx, gby, gbz);
        }
        mlpChain(M, x, y, z, df, ox, oy, oz);
        const double fac = 1 + kOut * (double)out, pk = pb * kOut;
        const double gx = gbx * fac + pk * ox, gy = gby * fac + pk * oy, gz = gbz * fac + pk * oz;
        // ---- the FP64 Multipole: initialisation, and judge of the energy of accepted steps
        double pot = pb * fac, fx = gx, fy = gy, fz = gz;
        const bool init = L.phase == PH_INIT_A || L.phase == PH_INIT_B || L.phase == PH_INIT_C;
        if(run && (init || L.phase == PH_FSAL)) {
            FetchVec2 fp{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
            double jx, jy, jz;
            multipoleCar(P, fp, x, y, z, pot, jx, jy, jz);
            if(init) { fx = jx; fy = jy; fz = jz; }
        }
        if(!run) continue;
        const double r2 = x * x + y * y + z * z;
        bool fallback = !(r2 >= rmin2 && r2 <= rmax2);
        double f[6];
        f[0] = L.y[3]; f[1] = L.y[4]; f[2] = L.y[5];
        f[3] = -fx * ca + fy * sa;
        f[4] = -fy * ca - fx * sa;
        f[5] = -fz;
        const int phaseBefore = L.phase, stepsBefore = L.nsteps;
        advancePhase(A, L, f, pot, accRel);
        if(phaseBefore == PH_INIT_C) hRef = fabs(L.h);
        // maxde is updated at accepted steps (also by the one that ends the orbit)
        const double mde = L.f(F_MAXDE);
        if(!(mde <= SA.maxRelDE)) fallback = true;
        if(L.phase != PH_IDLE && L.nsteps != stepsBefore && fabs(L.f(F_LASTH)) * 64 < hRef) fallback = true;
        if(fallback) {
            const unsigned int slot = atomicAdd(SA.nfallback, 1u);
            SA.fallback[slot] = L.oi;
            L.phase = PH_IDLE;
        }
        if(L.phase == PH_IDLE) fetchOrbit(A, L);
    }
}

__global__ void markFallbackKernel(const int* __restrict__ list, long n, int* __restrict__ status)
{
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        status[list[i]] |= 8;
}

}  // namespace

int agxOrbitSurrogateBlocksPerCU()
{
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitSurrogateKernel, SBS, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

hipError_t agxLaunchOrbitSurrogate(const AgxMultipoleDev* dP, const AgxMultipoleDev* dBase, const AgxMlpDev* dM,
    const AgxOrbitArgs& A, const AgxSurrogateOrbitArgs& SA, int numBlocks, hipStream_t stream)
{
    if(A.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(orbitSurrogateKernel, dim3(numBlocks), dim3(SBS), 0, stream, dP, dBase, dM, A, SA);
    return hipGetLastError();
}

hipError_t agxLaunchMarkFallback(const int* list, long n, int* status, hipStream_t stream)
{
    if(n <= 0 || !status) return hipSuccess;
    const long g = (n + 255) / 256;
    hipLaunchKernelGGL(markFallbackKernel, dim3((unsigned)(g < 4096 ? g : 4096)), dim3(256), 0, stream, list, n, status);
    return hipGetLastError();
}
