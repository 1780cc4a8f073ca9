// Multipole construction from a density on the device: the bulk of the Poisson solver of
// computePotentialCoefsFromSource (src/potential_multipole.cpp:1013-1114, 238-298) - the
// density at every (quadrature radius, angular node), its forward spherical-harmonic transform
// (src/math_sphharm.cpp:265-384) with the near-zero cut of eliminateNearZeros, and the radial
// integrals Pint/Pext - while the host keeps the grids, the transform matrix and the final
// restriction (host_multipole.cpp: poissonSetup, poissonFinish).
//
// Kernels: rhoKernel, one lane per (radius, angle) point; shKernel, one workgroup per
// quadrature radius (its density row in LDS, one lane per coefficient: the transform is a
// [nint x nang] x [nang x ncoef] product, then the row's largest magnitude sets the cut);
// poissonSegmentKernel, one lane per (term, radial segment) for the Gauss-Legendre integrals;
// poissonScanKernel, one lane per term for the two short recurrences along the grid.
#include <hip/hip_runtime.h>
#include <cmath>
#include <vector>
#include "agx_device.h"
#include "host_multipole.h"

using namespace agxd;

namespace {

constexpr int MAX_COMPONENTS = 8;

struct DensityArgs {
    int n;
    agx::DensityComp comp[MAX_COMPONENTS];
};

// power with exact special cases (math_core.cpp:135-148)
__device__ __forceinline__ double powE(double x, double n)
{
    if(n == 0.0) return 1;
    if(n == 1.0) return x;
    if(n == -1.0) return 1 / x;
    if(n == 2.0) return x * x;
    if(n == -2.0) return 1 / (x * x);
    if(n == 0.5) return sqrt(x);
    if(n == -0.5) return 1 / sqrt(x);
    if(n == 3.0) return x * x * x;
    if(n == -3.0) return 1 / (x * x * x);
    return pow(x, n);
}

// integer power by squaring (math_core.cpp:120-133)
__device__ __forceinline__ double ipowE(double x, int n)
{
    if(n < 0) { n = -n; x = 1 / x; }
    double r = 1;
    do { if(n % 2) r *= x; n >>= 1; x *= x; } while(n);
    return r;
}

// the density models of host_multipole.cpp (Dehnen src/potential_dehnen.cpp:37-42, Ferrers
// src/potential_ferrers.cpp:17-31, exponential disk, NFW, Plummer), summed in order
__device__ double densityAt(const DensityArgs& D, double R, double z, double phi)
{
    double sum = 0;
    for(int i = 0; i < D.n; i++) {
        const agx::DensityComp& c = D.comp[i];
        double v = 0;
        switch(c.type) {
        case agx::DENS_DEHNEN: {
            double s, co;
            sincosA(phi, s, co);
            const double x = R * co, y = R * s, M = c.p[0], a = c.p[1], g = c.p[2], p = c.p[3], q = c.p[4];
            const double m = sqrt(sqd(x) + sqd(y / p) + sqd(z / q));
            v = M * a * (3 - g) / (4 * M_PI * p * q) * powE(m, -g) * powE(a + m, g - 4);
            break;
        }
        case agx::DENS_FERRERS: {
            double s, co;
            sincosA(phi, s, co);
            const double m2 = sqd(R * co / c.p[0]) + sqd(R * s / c.p[1]) + sqd(z / c.p[2]);
            v = m2 > 1 ? 0 : c.p[3] * sqd(1 - m2);
            break;
        }
        case agx::DENS_DISK:
            v = c.p[0] * exp(-R / c.p[1]) * (0.5 / c.p[2]) * exp(-fabs(z) / c.p[2]);
            break;
        case agx::DENS_NFW: {
            const double r = sqrt(R * R + z * z);
            v = c.p[0] / (4 * M_PI) / (r * sqd(c.p[1] + r));
            break;
        }
        case agx::DENS_PLUMMER: {
            const double a = c.p[1];
            v = 3 * c.p[0] / (4 * M_PI * a * a * a) * pow(1 + (R * R + z * z) / (a * a), -2.5);
            break;
        }
        default: v = NAN;
        }
        sum = i == 0 ? v : sum + v;
    }
    return sum;
}

__global__ __launch_bounds__(256) void rhoKernel(DensityArgs D, long npts, int nang, const double* __restrict__ rint,
    const double* __restrict__ cth, const double* __restrict__ phi, double* __restrict__ vals)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= npts) return;
    const long ir = i / nang;
    const int ia = (int)(i - ir * nang);
    const double rad = rint[ir], ct = cth[ia];
    vals[i] = densityAt(D, rad * sqrt(1 - ct * ct), rad * ct, phi[ia]);
}

constexpr int SH_BLOCK = 256;

__global__ __launch_bounds__(SH_BLOCK) void shKernel(int nint, int nang, int ncoef, const double* __restrict__ vals,
    const double* __restrict__ B, double* __restrict__ ic)
{
    extern __shared__ double lds[];            // nang values, then SH_BLOCK partial maxima
    double* row = lds;
    double* red = lds + nang;
    const int ir = blockIdx.x;
    for(int ia = threadIdx.x; ia < nang; ia += SH_BLOCK) row[ia] = vals[(long)ir * nang + ia];
    __syncthreads();
    double mag = 0;
    for(int c = threadIdx.x; c < ncoef; c += SH_BLOCK) {
        double acc = 0;
        for(int ia = 0; ia < nang; ia++) acc += row[ia] * B[(long)ia * ncoef + c];
        ic[(long)c * nint + ir] = acc;
        mag = fmax(mag, fabs(acc));
    }
    red[threadIdx.x] = mag;
    __syncthreads();
    for(int w = SH_BLOCK / 2; w > 0; w >>= 1) {
        if(threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const double cut = red[0] * 1e-10;
    for(int c = threadIdx.x; c < ncoef; c += SH_BLOCK) {
        double& v = ic[(long)c * nint + ir];
        if(fabs(v) <= cut) v = 0;
    }
}

// the radial integrals of term c over segment k, in parallel over (c, k): the inner one with
// its propagation factor (R[k-1]/R[k])^(l+1) and the outer one with (R[k]/R[k+1])^l
// (computePotentialCoefsFromSource, potential_multipole.cpp:1013-1114); work holds four
// [ncoef][nr] arrays: inner, inner factor, outer, outer factor
__global__ __launch_bounds__(256) void poissonSegmentKernel(int ncoef, int nr, int glo, int nint,
    const int* __restrict__ active, const double* __restrict__ ic, const double* __restrict__ R,
    const double* __restrict__ rint, const double* __restrict__ gln, const double* __restrict__ glw,
    double* __restrict__ work)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= (long)ncoef * nr) return;
    const int c = (int)(i / nr), k = (int)(i - (long)c * nr);
    if(!active[c]) return;
    const int l = (int)sqrt((double)c);
    const double* q = ic + (long)c * nint;
    double* inner = work + (long)c * nr;
    double* innerF = inner + (long)ncoef * nr;
    double* outer = innerF + (long)ncoef * nr;
    double* outerF = outer + (long)ncoef * nr;
    double integr = 0;
    for(int s = 0; s < glo; s++) integr += q[k * glo + s] * glw[s] * ipowE(rint[k * glo + s] / R[k], l + 2);
    inner[k] = integr * (R[k] - (k > 0 ? R[k - 1] : 0)) * R[k];
    innerF[k] = k > 0 ? ipowE(R[k - 1] / R[k], l + 1) : 0;
    integr = 0;
    if(k == nr - 1) {
        for(int s = 0; s < glo; s++)
            integr += q[nr * glo + s] * glw[s] * ipowE(rint[nr * glo + s] / R[nr - 1], 1 - l) / sqd(gln[s]);
        outer[k] = integr * R[nr - 1] * R[nr - 1];
        outerF[k] = 0;
    } else {
        for(int s = 0; s < glo; s++) integr += q[(k + 1) * glo + s] * glw[s] * ipowE(rint[(k + 1) * glo + s] / R[k], 1 - l);
        outer[k] = integr * (R[k + 1] - R[k]) * R[k];
        outerF[k] = ipowE(R[k] / R[k + 1], l);
    }
}

// the two recurrences over the grid and the coefficients, one lane per term
__global__ __launch_bounds__(64) void poissonScanKernel(int ncoef, int nr, const int* __restrict__ active,
    const double* __restrict__ R, const double* __restrict__ work, double* __restrict__ Phi,
    double* __restrict__ dPhi)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= ncoef) return;
    double* ph = Phi + (long)c * nr;
    double* dph = dPhi + (long)c * nr;
    if(!active[c]) {
        for(int k = 0; k < nr; k++) ph[k] = dph[k] = 0;
        return;
    }
    const int l = (int)sqrt((double)c);
    const double* inner = work + (long)c * nr;
    const double* innerF = inner + (long)ncoef * nr;
    const double* outer = innerF + (long)ncoef * nr;
    const double* outerF = outer + (long)ncoef * nr;
    const double mul = -4 * M_PI / (2 * l + 1);
    // Pint runs outward; Pext inward (Pint values kept in ph until Pext is known)
    double Pint = 0;
    for(int k = 0; k < nr; k++) {
        Pint = inner[k] + (k > 0 ? Pint * innerF[k] : 0);
        ph[k] = Pint;
    }
    double Pext = outer[nr - 1];
    for(int k = nr - 1; k >= 0; k--) {
        if(k < nr - 1) Pext = outer[k] + Pext * outerF[k];
        const double Pi = ph[k];
        ph[k] = mul * (Pi + Pext);
        dph[k] = mul * (-(l + 1) * Pi + l * Pext) / R[k];
    }
}

// one device allocation carved into the solver's arrays (256-byte aligned)
struct Arena {
    char* base = nullptr;
    size_t used = 0, cap = 0;
    explicit Arena(size_t bytes) : cap(bytes) { if(hipMalloc((void**)&base, bytes) != hipSuccess) base = nullptr; }
    ~Arena() { if(base) (void)hipFree(base); }
    template<typename T> T* take(size_t count)
    {
        T* p = reinterpret_cast<T*>(base + used);
        used += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

}  // namespace

// the middle part of the Poisson solver on the device (see poissonStagesHost for the same on
// the host); Phi and dPhi receive ncoef x gridSizeR values
hipError_t agxPoissonDevice(const agx::PoissonSetup& S, const agx::Density& dens, std::vector<double>& Phi,
    std::vector<double>& dPhi)
{
    std::vector<agx::DensityComp> comps;
    dens.describe(comps);
    if(comps.empty() || comps.size() > (size_t)MAX_COMPONENTS) return hipErrorInvalidValue;
    DensityArgs D;
    D.n = (int)comps.size();
    for(int i = 0; i < D.n; i++) D.comp[i] = comps[i];
    const int nint = S.nint, nang = S.nang, nc = S.ncoef, nr = S.gridSizeR;
    const long npts = (long)nint * nang;
    auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    const size_t total = al(nint * 8) + 2 * al(nang * 8) + al(npts * 8) + al(S.B.size() * 8) + al((size_t)nc * nint * 8) +
        al(nr * 8) + 2 * al(S.glo * 8) + al(nc * 4) + al((size_t)nc * 4 * nr * 8) + 2 * al((size_t)nc * nr * 8);
    Arena A(total);
    if(!A.base) return hipErrorOutOfMemory;
    double* dr = A.take<double>(nint);
    double* dct = A.take<double>(nang);
    double* dph = A.take<double>(nang);
    double* dvals = A.take<double>(npts);
    double* dB = A.take<double>(S.B.size());
    double* dic = A.take<double>((size_t)nc * nint);
    double* dR = A.take<double>(nr);
    double* dgn = A.take<double>(S.glo);
    double* dgw = A.take<double>(S.glo);
    int* dact = A.take<int>(nc);
    double* dwork = A.take<double>((size_t)nc * 4 * nr);
    double* dP = A.take<double>((size_t)nc * nr);
    double* ddP = A.take<double>((size_t)nc * nr);
    hipError_t e;
#define AGX_TRY(x) do { e = (x); if(e != hipSuccess) return e; } while(0)
    AGX_TRY(hipMemcpy(dr, S.rint.data(), nint * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dct, S.cth.data(), nang * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dph, S.phi.data(), nang * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dB, S.B.data(), S.B.size() * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dR, S.R.data(), nr * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dgn, S.gln.data(), S.glo * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dgw, S.glw.data(), S.glo * 8, hipMemcpyHostToDevice));
    AGX_TRY(hipMemcpy(dact, S.active.data(), nc * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(rhoKernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, nullptr, D, npts, nang,
        dr, dct, dph, dvals);
    AGX_TRY(hipGetLastError());
    const size_t lds = (size_t)(nang + SH_BLOCK) * 8;
    if(lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(shKernel, dim3(nint), dim3(SH_BLOCK), lds, nullptr, nint, nang, nc, dvals, dB, dic);
    AGX_TRY(hipGetLastError());
    const long nseg = (long)nc * nr;
    hipLaunchKernelGGL(poissonSegmentKernel, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, nullptr, nc, nr,
        S.glo, nint, dact, dic, dR, dr, dgn, dgw, dwork);
    AGX_TRY(hipGetLastError());
    hipLaunchKernelGGL(poissonScanKernel, dim3((nc + 63) / 64), dim3(64), 0, nullptr, nc, nr, dact, dR, dwork, dP, ddP);
    AGX_TRY(hipGetLastError());
    Phi.resize((size_t)nc * nr);
    dPhi.resize((size_t)nc * nr);
    AGX_TRY(hipMemcpy(Phi.data(), dP, (size_t)nc * nr * 8, hipMemcpyDeviceToHost));
    AGX_TRY(hipMemcpy(dPhi.data(), ddP, (size_t)nc * nr * 8, hipMemcpyDeviceToHost));
#undef AGX_TRY
    return hipSuccess;
}
