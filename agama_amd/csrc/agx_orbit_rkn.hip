// Persistent orbit kernel with the 8th-order Runge-Kutta-Nystrom integrator DPRKN8
// (Dormand-Prince RKN6(8) with a PI step controller and a quintic Taylor dense output):
// restates src/math_ode.cpp:426-701 (OdeStepperDPRKN8) driven by src/orbit.cpp:52-84, 217-250
// (run loop, OrbitIntegrator<Car>::eval2).
//
// Same design as the DOP853 kernel (agx_kernels.hip): one orbit per lane, a persistent work
// queue in longest-first order, one right-hand-side evaluation per loop iteration chosen by a
// per-lane phase (3 initialisation evaluations, 8 stages, 1 final evaluation = 9 per attempt
// instead of DOP853's 12), stage sequences aligned across the wavefront.  Only positions enter
// the stages, so the stage vectors are 3-vectors; k2..k9 live in LDS.
#include <hip/hip_runtime.h>
#include "agx_device.h"
#include "agx_dop853.h"
#include "agx_kernels.h"

using namespace agxd;

namespace agxd {

// DPRKN8 coefficients (math_ode.cpp:453-518)
namespace rkn {
constexpr double
    c1 = 1.0 / 20, c2 = 1.0 / 10, c3 = 3.0 / 10, c4 = 1.0 / 2, c5 = 7.0 / 10, c6 = 9.0 / 10, c7 = 1.0, c8 = 1.0,
    a21 = 1.0 / 800, a31 = 1.0 / 600, a32 = 1.0 / 300, a41 = 9.0 / 200, a42 = -9.0 / 100, a43 = 9.0 / 100,
    a51 = -66701.0 / 197352, a52 = 28325.0 / 32892, a53 = -2665.0 / 5482, a54 = 2170.0 / 24669,
    a61 = 227015747.0 / 304251000, a62 = -54897451.0 / 30425100, a63 = 12942349.0 / 10141700,
    a64 = -9499.0 / 304251, a65 = 539.0 / 9250,
    a71 = -1131891597.0 / 901789000, a72 = 41964921.0 / 12882700, a73 = -6663147.0 / 3220675,
    a74 = 270954.0 / 644135, a75 = -108.0 / 5875, a76 = 114.0 / 1645,
    a81 = 13836959.0 / 3667458, a82 = -17731450.0 / 1833729, a83 = 1063919505.0 / 156478208,
    a84 = -33213845.0 / 39119552, a85 = 13335.0 / 28544, a86 = -705.0 / 14272, a87 = 1645.0 / 57088,
    a91 = 223.0 / 7938, a93 = 1175.0 / 8064, a94 = 925.0 / 6048, a95 = 41.0 / 448, a96 = 925.0 / 14112,
    a97 = 1175.0 / 72576,
    b1 = 223.0 / 7938, b3 = 1175.0 / 8064, b4 = 925.0 / 6048, b5 = 41.0 / 448, b6 = 925.0 / 14112,
    b7 = 1175.0 / 72576,
    bp1 = 223.0 / 7938, bp3 = 5875.0 / 36288, bp4 = 4625.0 / 21168, bp5 = 41.0 / 224,
    bp6 = 4625.0 / 21168, bp7 = 5875.0 / 36288, bp8 = 223.0 / 7938,
    bt1 = 223.0 / 7938 - 7987313.0 / 109941300, bt3 = 1175.0 / 8064 - 1610737.0 / 44674560,
    bt4 = 925.0 / 6048 - 10023263.0 / 33505920, bt5 = 41.0 / 448 + 497221.0 / 12409600,
    bt6 = 925.0 / 14112 - 10023263.0 / 78180480, bt7 = 1175.0 / 72576 - 1610737.0 / 402071040,
    bpt1 = 223.0 / 7938 - 7987313.0 / 109941300, bpt3 = 5875.0 / 36288 - 1610737.0 / 40207104,
    bpt4 = 4625.0 / 21168 - 10023263.0 / 23454144, bpt5 = 41.0 / 224 + 497221.0 / 6204800,
    bpt6 = 4625.0 / 21168 - 10023263.0 / 23454144, bpt7 = 5875.0 / 36288 - 1610737.0 / 40207104,
    bpt8 = 223.0 / 7938 + 4251941.0 / 54970650, bpt9 = -3.0 / 20,
    gam = 0.9, qmax = 10.0, qmin = 0.2, beta1 = 7. / 80, beta2 = 4. / 80;
}

enum : int { PR_IDLE = PH_IDLE, PR_INIT_A = PH_INIT_A, PR_INIT_B = PH_INIT_B, PR_INIT_C = PH_INIT_C,
    PR_S2 = 20, PR_S3, PR_S4, PR_S5, PR_S6, PR_S7, PR_S8, PR_S9, PR_FIN };

// LDS fields: the stage accelerations k2..k9 (3 each) in the slots the DOP853 layout uses for
// k6..k10, the shared bookkeeping fields at the same indices, and the controller's q_old
enum : int { R_K2 = 0, R_K3 = 3, R_K4 = 6, R_K5 = 9, R_K6 = 12, R_K7 = 15, R_K8 = 18, R_K9 = 21,
    R_QOLD = 24, R_COUNT = F_COUNT };
static_assert(R_QOLD + 1 <= F_H1, "stage slots overlap the bookkeeping fields");

template<int STRIDE>
struct LaneRKN {
    static constexpr int LDS_FIELDS = R_COUNT;
    static constexpr int EVALS_PER_ATTEMPT = 9;
    double* sh;
    int oi, phase, nbad, S, nsteps, natt, actual, niter;
    double x[6];               // position and velocity at the start of the step (after: at its end)
    double k1[3];              // acceleration at the start of the step
    double y[6];               // evaluation point (positions; at the final stage also the velocity)
    double toff, tb, tend, h, hnext, preverr, rem;
    __device__ __forceinline__ void attach(double* col) { sh = col; }
    __device__ __forceinline__ double& f(int field) const { return sh[field * STRIDE]; }
    __device__ __forceinline__ double& k(int slot, int i) const { return sh[(slot + i) * STRIDE]; }
};

// Dense output of component c at offset `off` inside the accepted step: the quintic Taylor
// series around the step's end (math_ode.cpp:656-701), formed on demand from the start values
// (L.x, L.k1), the end values (L.y, kn) and the step length.
template<int STRIDE>
__device__ __forceinline__ double denseOutRKN(const LaneRKN<STRIDE>& L, const double* kn, int c, double off)
{
    const double h = L.h, dt = off - h;
    if(dt == 0) return L.y[c];
    const int i = c % 3;
    const double xo = L.x[i], vo = L.x[3 + i], ko = L.k1[i], xn = L.y[i], vn = L.y[3 + i], kk = kn[i];
    const double jn = ((-60 * (xo - xn) / h + -24 * vo - 36 * vn) / h + -3 * ko + 9 * kk) / h;
    const double s4 = ((-360 * (xo - xn) / h + -168 * vo - 192 * vn) / h + -24 * ko + 36 * kk) / (h * h);
    const double s5 = ((-720 * (xo - xn) / h + -360 * vo - 360 * vn) / h + -60 * ko + 60 * kk) / (h * h * h);
    if(c < 3)
        return xn + dt * (vn + dt * (1. / 2) * (kk + dt * (1. / 3) * (jn + dt * (1. / 4) *
            (s4 + dt * (1. / 5) * s5))));
    return vn + dt * (kk + dt * (1. / 2) * (jn + dt * (1. / 3) * (s4 + dt * (1. / 4) * s5)));
}

// RuntimeTrajectory::processTimestep (orbit.cpp:15-47) for an accepted RKN step [tb, tb+h]
template<int STRIDE>
__device__ __forceinline__ void sampleStepRKN(const AgxOrbitArgs& A, LaneRKN<STRIDE>& L, const double* kn)
{
    if(L.S < 0) {   // sampling interval 0: getSol(0) of the first step, then every step's end
        const long cap = rowCapacity(L.S), base = trajOffsetOf(A, L.oi), row = (long)L.nsteps + 1;
        double d[6], o[6];
        if(L.nsteps == 0) {
            for(int c = 0; c < 6; c++) d[c] = denseOutRKN(L, kn, c, 0.0);
            toRotating(A.Omega, A.Omega * L.tb, d, o);
            storeRow(A, base, L.tb, o);
        }
        const long r = row < cap ? base + row : overflowRowCall(A.chunkDir, A.poolCounter, A.poolBase,
            A.poolChunk, A.maxChunks, A.poolChunks, L.oi, row - cap);
        if(r >= 0) {
            toRotating(A.Omega, A.Omega * (L.h + L.tb), L.y, o);
            storeRow(A, r, L.tb + L.h, o);
        }
        L.actual = 1;
        return;
    }
    if(L.S < 2) return;
    const double tbegin = L.tb, timestep = L.h;
    double t0 = L.f(F_T0);
    if(t0 != t0) { t0 = tbegin; L.f(F_T0) = t0; }
    const double dtS = L.f(F_DTS);
    const double sign = timestep >= 0 ? +1 : -1, tendS = tbegin + timestep;
    const double dtr = ROUNDOFF * fmax(fmax(fabs(tendS), fabs(tbegin)), fabs(t0));
    long iout = (long)ceil(sign * (tbegin - t0) / dtS);
    const long iend = (long)((sign * (tendS - t0) + dtr) / dtS);
    L.actual = (int)min(iend + 1, (long)L.S + 1);
    if(iout > iend) return;
    const long base = trajOffsetOf(A, L.oi);
    for(; iout <= iend; iout++) {
        const double timeout = sign * dtS * iout + t0;
        const double off = fmin(fmax(sign * (timeout - tbegin), 0.0), sign * timestep) * sign;
        double d[6], o[6];
        for(int c = 0; c < 6; c++) d[c] = denseOutRKN(L, kn, c, off);
        toRotating(A.Omega, A.Omega * (off + tbegin), d, o);
        if(iout >= 0 && iout < L.S) storeRow(A, base + iout, timeout, o);
    }
}

// RuntimeFncTarget::processTimestep (galaxymodel_target.h:141-154) for an accepted RKN step:
// ten equally spaced dense-output points (rotating frame), each weighted by a tenth of the step
template<int STRIDE>
__device__ __forceinline__ void targetStepRKN(const AgxOrbitArgs& A, const LaneRKN<STRIDE>& L, const double* kn)
{
    const AgxTargetList& TL = *A.targets;
    const int NUM_SAMPLES_PER_STEP = 10;
    const double timestep = L.h, substep = timestep / NUM_SAMPLES_PER_STEP;
    for(int s = 0; s < NUM_SAMPLES_PER_STEP; s++) {
        const double off = substep * (s + 0.5);
        double d[6], p[6];
        for(int c = 0; c < 3; c++) d[c] = denseOutRKN(L, kn, c, off);
        if(TL.needVelocity)
            for(int c = 3; c < 6; c++) d[c] = denseOutRKN(L, kn, c, off);
        else
            d[3] = d[4] = d[5] = 0;
        toRotating(A.Omega, A.Omega * (off + L.tb), d, p);
        for(int t = 0; t < TL.count; t++) {
            const AgxTargetDev& T = TL.t[t];
            targetAdd(T, T.cube + (long)L.oi * (T.numValues + 1), p, substep);
        }
    }
    for(int t = 0; t < TL.count; t++) {
        const AgxTargetDev& T = TL.t[t];
        T.cube[(long)L.oi * (T.numValues + 1) + T.numValues] += timestep;
    }
}

template<int STRIDE>
__device__ __forceinline__ bool beginAttemptRKN(const AgxOrbitArgs& A, LaneRKN<STRIDE>& L)
{
    if(L.h == 0 || !(L.h > -INFINITY && L.h < INFINITY)) { finishOrbit(A, L, 1); return false; }
    L.natt++;
    const double h = L.h;
    for(int i = 0; i < 3; i++) L.y[i] = L.x[i] + h * (rkn::c1 * L.x[3 + i] + h * (rkn::a21 * L.k1[i]));
    L.toff = rkn::c1 * h;
    L.phase = PR_S2;
    return true;
}

template<int STRIDE>
__device__ __forceinline__ void beginStepRKN(const AgxOrbitArgs& A, LaneRKN<STRIDE>& L)
{
    L.rem = L.tend - L.tb;
    if(L.rem == 0) { finishOrbit(A, L, 1); return; }
    const double sign = L.rem >= 0 ? +1 : -1;
    L.h = fmin(L.rem * sign, L.hnext) * sign;
    L.nbad = 0;
    L.niter = 0;
    L.preverr = INFINITY;
    beginAttemptRKN(A, L);
}

// one evaluation consumed: acceleration a (3) at the lane's evaluation point, potential pot;
// accRel is the stepper's relative tolerance 10 accuracy^0.9 (math_ode.cpp:428-431)
// TGT: also feed the target plug-ins of A.targets at every accepted step
template<int STRIDE, bool TGT = false>
__device__ __forceinline__ void advanceRKN(const AgxOrbitArgs& A, LaneRKN<STRIDE>& L, const double* a,
    double pot, double accRel)
{
    const double h = L.h;
    switch(L.phase) {
    case PR_INIT_A: {   // DPRKN8::init + initTimeStep on the first-order form w = (x, v)
        for(int i = 0; i < 3; i++) L.k1[i] = a[i];
        const double Ekin = 0.5 * (L.x[3] * L.x[3] + L.x[4] * L.x[4] + L.x[5] * L.x[5]);
        L.f(F_ACC) = fmin(1., fabs(pot + Ekin) / fmax(fabs(pot), Ekin));
        L.f(F_E0) = energyOf(L.x, pot, A.Omega);
        double nx = 0, nd = 0;
        for(int i = 0; i < 6; i++) nx += L.x[i] * L.x[i];
        for(int i = 0; i < 3; i++) nd += L.x[3 + i] * L.x[3 + i];
        for(int i = 0; i < 3; i++) nd += a[i] * a[i];
        const double h1 = fmax(1e-12, fmin(1e12, sqrt(nx / nd) * 0.01));
        L.f(F_H1) = h1;
        for(int i = 0; i < 3; i++) L.y[i] = L.x[i] + h1 * L.x[3 + i];   // Euler predictor (positions)
        L.toff = h1;
        L.phase = PR_INIT_B;
        break;
    }
    case PR_INIT_B: {   // k2 = f(w + h1 k1) = (v + h1 a, a2); Heun corrector point
        const double h1 = L.f(F_H1);
        for(int i = 0; i < 3; i++) {
            L.k(R_K2, i) = a[i];
            const double v2 = L.x[3 + i] + h1 * L.k1[i];
            L.y[i] = L.x[i] + 0.5 * h1 * (L.x[3 + i] + v2);
            L.k(R_K3, i) = L.x[3 + i] + 0.5 * h1 * (L.k1[i] + a[i]);   // corrector velocity
        }
        L.toff = h1;
        L.phase = PR_INIT_C;
        break;
    }
    case PR_INIT_C: {   // k3 = (v_c, a3): initTimeStep's step-size estimate over the 6 components
        const double h1 = L.f(F_H1);
        double h2 = INFINITY;
        for(int i = 0; i < 6; i++) {
            const int j = i % 3;
            double xi, xti, k1i, k2i, k3i;
            if(i < 3) {
                xi = L.x[i]; xti = L.y[i];
                k1i = L.x[3 + i]; k2i = L.x[3 + i] + h1 * L.k1[i]; k3i = L.k(R_K3, i);
            } else {
                xi = L.x[i]; xti = L.k(R_K3, j);
                k1i = L.k1[j]; k2i = L.k(R_K2, j); k3i = a[j];
            }
            const double d0 = fmax(0., fmax(fabs(xi), fabs(xti)));
            const double d1 = fmax(fabs(k1i), fabs(k3i));
            const double d2 = fabs((3 * k2i - k1i - 2 * k3i) / h1);
            const double d3 = fabs(6 * (k3i - k2i) / (h1 * h1));
            const double hh = (d0 * d2 + d1 * d1) / (d1 * d3 + d2 * d2);
            h2 = fmin(h2, hh);
        }
        L.hnext = fmax(1e-12, fmin(1e12, pow(accRel, 1. / 8) * sqrt(h2))) * L.f(F_ACC);
        L.f(R_QOLD) = 0.0001;
        if(A.ttotal[L.oi] == 0 || !(L.tend > -INFINITY && L.tend < INFINITY)) {
            L.f(F_LASTTB) = L.tb;
            L.f(F_LASTH) = 0;
            if(L.S >= 2 || L.S < 0) {
                double o[6];
                toRotating(A.Omega, A.Omega * L.tb, L.x, o);
                const long base = trajOffsetOf(A, L.oi);
                storeRow(A, base, L.tb, o);
                storeRow(A, base + 1, L.tb, o);
                L.actual = 2;
            } else
                L.actual = 1;
            finishOrbit(A, L, 0);
        } else
            beginStepRKN(A, L);
        break;
    }
#define RKN_STAGE(PH, SLOT, NEXT, CN, EXPR)                                                      \
    case PH:                                                                                      \
        for(int i = 0; i < 3; i++) {                                                              \
            L.k(SLOT, i) = a[i];                                                                  \
            L.y[i] = L.x[i] + h * (CN * L.x[3 + i] + h * (EXPR));                                 \
        }                                                                                         \
        L.toff = CN * h; L.phase = NEXT; break;
    RKN_STAGE(PR_S2, R_K2, PR_S3, rkn::c2, rkn::a31 * L.k1[i] + rkn::a32 * a[i])
    RKN_STAGE(PR_S3, R_K3, PR_S4, rkn::c3, rkn::a41 * L.k1[i] + rkn::a42 * L.k(R_K2, i) + rkn::a43 * a[i])
    RKN_STAGE(PR_S4, R_K4, PR_S5, rkn::c4, rkn::a51 * L.k1[i] + rkn::a52 * L.k(R_K2, i) + rkn::a53 * L.k(R_K3, i) + rkn::a54 * a[i])
    RKN_STAGE(PR_S5, R_K5, PR_S6, rkn::c5, rkn::a61 * L.k1[i] + rkn::a62 * L.k(R_K2, i) + rkn::a63 * L.k(R_K3, i) +
        rkn::a64 * L.k(R_K4, i) + rkn::a65 * a[i])
    RKN_STAGE(PR_S6, R_K6, PR_S7, rkn::c6, rkn::a71 * L.k1[i] + rkn::a72 * L.k(R_K2, i) + rkn::a73 * L.k(R_K3, i) +
        rkn::a74 * L.k(R_K4, i) + rkn::a75 * L.k(R_K5, i) + rkn::a76 * a[i])
    RKN_STAGE(PR_S7, R_K7, PR_S8, rkn::c7, rkn::a81 * L.k1[i] + rkn::a82 * L.k(R_K2, i) + rkn::a83 * L.k(R_K3, i) +
        rkn::a84 * L.k(R_K4, i) + rkn::a85 * L.k(R_K5, i) + rkn::a86 * L.k(R_K6, i) + rkn::a87 * a[i])
    RKN_STAGE(PR_S8, R_K8, PR_S9, rkn::c8, rkn::a91 * L.k1[i] + rkn::a93 * L.k(R_K3, i) + rkn::a94 * L.k(R_K4, i) +
        rkn::a95 * L.k(R_K5, i) + rkn::a96 * L.k(R_K6, i) + rkn::a97 * L.k(R_K7, i))
#undef RKN_STAGE
    case PR_S9:   // k9; the end point (x_n, v_n) becomes the evaluation point
        for(int i = 0; i < 3; i++) {
            L.k(R_K9, i) = a[i];
            L.y[i] = L.x[i] + h * (L.x[3 + i] + h * (rkn::b1 * L.k1[i] + rkn::b3 * L.k(R_K3, i) + rkn::b4 * L.k(R_K4, i) +
                rkn::b5 * L.k(R_K5, i) + rkn::b6 * L.k(R_K6, i) + rkn::b7 * L.k(R_K7, i)));
            L.y[3 + i] = L.x[3 + i] + h * (rkn::bp1 * L.k1[i] + rkn::bp3 * L.k(R_K3, i) + rkn::bp4 * L.k(R_K4, i) +
                rkn::bp5 * L.k(R_K5, i) + rkn::bp6 * L.k(R_K6, i) + rkn::bp7 * L.k(R_K7, i) + rkn::bp8 * L.k(R_K8, i));
        }
        L.toff = h; L.phase = PR_FIN; break;
    case PR_FIN: {   // k_n at (x_n, v_n): error estimate and PI step control (math_ode.cpp:597-654)
        const double Ekin = 0.5 * (L.y[3] * L.y[3] + L.y[4] * L.y[4] + L.y[5] * L.y[5]);
        const double accFac = fmin(1., fabs(pot + Ekin) / fmax(fabs(pot), Ekin));
        const double acc = accRel;
        double err = 0.0;
        for(int i = 0; i < 3; i++) {
            double denom = (acc * fmax(fabs(L.x[i]), fabs(L.y[i]))) * accFac;
            if(denom != 0) {
                const double xt = (h * h) * (rkn::bt1 * L.k1[i] + rkn::bt3 * L.k(R_K3, i) + rkn::bt4 * L.k(R_K4, i) +
                    rkn::bt5 * L.k(R_K5, i) + rkn::bt6 * L.k(R_K6, i) + rkn::bt7 * L.k(R_K7, i));
                err += (xt / denom) * (xt / denom);
            }
            denom = (acc * fmax(fabs(L.x[3 + i]), fabs(L.y[3 + i]))) * accFac;
            if(denom != 0) {
                const double vt = h * (rkn::bpt1 * L.k1[i] + rkn::bpt3 * L.k(R_K3, i) + rkn::bpt4 * L.k(R_K4, i) +
                    rkn::bpt5 * L.k(R_K5, i) + rkn::bpt6 * L.k(R_K6, i) + rkn::bpt7 * L.k(R_K7, i) + rkn::bpt8 * L.k(R_K8, i) +
                    rkn::bpt9 * L.k(R_K9, i));
                err += (vt / denom) * (vt / denom);
            }
        }
        err = sqrt(err / 6);
        double q1 = pow(err, rkn::beta1);
        if(err <= 1 || ++L.niter >= 12) {   // accepted
            const double q = acc * accFac == INFINITY ? 0 :
                fmin(1 / rkn::qmin, fmax(1 / rkn::qmax, q1 / pow(L.f(R_QOLD), rkn::beta2) / rkn::gam));
            L.hnext = fabs(h) / q;
            L.f(R_QOLD) = err;
            // energy (orbit.cpp:65-82), sampling, advance
            const double de = fabs(energyOf(L.y, pot, A.Omega) / L.f(F_E0) - 1);
            const double mde = L.f(F_MAXDE);
            if(de > mde || de != de) L.f(F_MAXDE) = de;
            sampleStepRKN(A, L, a);
            if constexpr(TGT) targetStepRKN(A, L, a);
            L.f(F_LASTTB) = L.tb;
            L.f(F_LASTH) = L.h;
            if(L.S == 1) L.actual = 1;
            for(int i = 0; i < 6; i++) L.x[i] = L.y[i];
            for(int i = 0; i < 3; i++) L.k1[i] = a[i];
            L.tb = L.h == L.rem ? L.tend : L.tb + L.h;
            L.nsteps++;
            if(L.nsteps >= A.maxNumSteps) { finishOrbit(A, L, 2); break; }
            if(L.tb == L.tend) { finishOrbit(A, L, 0); break; }
            beginStepRKN(A, L);
        } else {
            if(err > 0.5 * L.preverr) {
                if(++L.nbad >= 2) q1 = 1 / rkn::qmin;
            } else
                L.nbad = 0;
            L.preverr = err;
            L.h = h / fmin(1 / rkn::qmin, q1 / rkn::gam);
            beginAttemptRKN(A, L);
        }
        break;
    }
    default: break;
    }
}

}  // namespace agxd

namespace {

constexpr int RBS = 256;
// LDS of the per-cell variant: the power-law arrays (10 (legLmax+1)^2 + legLmax + 1 doubles,
// legLmax <= 16) and the tau bucket table (at most 4096 ints); the grids must fit AGX_LDS_GRID
constexpr int RKN_LDS_PL = 2944;
constexpr int RKN_LDS_TAULUT = 4096;
typedef const double AGX_LDS* RknLdsGrid;

template<int MODE, bool NFW>
__device__ __forceinline__ void rhsRKN(const AgxMultipoleDev& P, double x, double y, double z, double& pot,
    double& gx, double& gy, double& gz, const double* gR, const double* gT, RknLdsGrid lR, RknLdsGrid lT,
    const PLTabs<RknLdsGrid>& plt)
{
    if constexpr(MODE == AGX_FETCH_CELL) {   // per-cell tables, grids and power-law arrays in LDS
        FetchCell f(P.cells, P.cellBytes);
        if(P.interp1d) multipoleCar<FetchCell, RknLdsGrid, NFW>(P, f, x, y, z, pot, gx, gy, gz, lR, lT);
        else multipoleCarCell<FetchCell, RknLdsGrid, NFW>(P, f, x, y, z, pot, gx, gy, gz, lR, lT, plt);
    } else if constexpr(MODE == AGX_FETCH_VEC2) {
        FetchVec2 f{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
        multipoleCar<FetchVec2, const double*, NFW, true>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
    } else {
        FetchFloat f{ gmem(P.nodesF) };
        multipoleCar<FetchFloat, const double*, NFW, true>(P, f, x, y, z, pot, gx, gy, gz, gR, gT);
    }
}

// MODE: AGX_FETCH_CELL (per-cell FP64 tables), AGX_FETCH_VEC2 (FP64 node records) or -1 (FP32
// form); TGT: the variant feeding target plug-ins; GRP: the variant with group mode (per-cell
// tables, no NFW halo): the group wavefronts (agxIsGroupWave) integrate the longest orbits with
// AGX_GROUP lanes per orbit, every lane of a group holding the same state and evaluating a share
// of the harmonics (multipoleCarCellGroup), as in the DOP853 kernel
template<int MODE, bool NFW, bool TGT = false, bool GRP = false>
__global__ __launch_bounds__(RBS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void orbitKernelRKN(const AgxMultipoleDev* __restrict__ Pp, AgxOrbitArgs A)
{
    typedef LaneRKN<RBS> Lane;
    __shared__ double lds[Lane::LDS_FIELDS * RBS];
    __shared__ double gridLds[AGX_LDS_GRID];
    constexpr bool CELLS = MODE == AGX_FETCH_CELL;
    __shared__ double plLds[CELLS ? RKN_LDS_PL : 1];
    __shared__ int lutLds[CELLS ? RKN_LDS_TAULUT : 1];
    const AgxMultipoleDev& P = *Pp;
    const double *gR, *gT;
    PLTabs<RknLdsGrid> plt{};
    if constexpr(CELLS) { if(!P.interp1d) plt = plStage(P, (double AGX_LDS*)plLds, (int AGX_LDS*)lutLds); }
    stageGrids(P, gridLds, gR, gT);   // (synchronises the block)
    const RknLdsGrid lR = (RknLdsGrid)gridLds, lT = lR + P.nr;   // the per-cell variant: grids in LDS
    Lane L;
    L.attach(lds + threadIdx.x);
    // wave-uniform: whether this wavefront works in group mode
    const bool group = GRP && agxIsGroupWave((int)blockIdx.x, (int)(threadIdx.x >> 6), RBS / 64, A.groupStride);
    fetchOrbit(A, L, group);
    const double accRel = 10 * pow(A.accuracy, 0.9);
    unsigned int it = 0;
    while(true) {
        const bool active = L.phase != PH_IDLE;
        if(!__any(active)) break;
        // stage sequences aligned to the wavefront's iteration counter (9 evaluations per attempt)
        const bool run = active && !(L.phase == PR_S2 && it % 9 != 0);
        it++;
        double ca = 1, sa = 0;
        if(A.Omega != 0 && active) sincosA(A.Omega * (L.tb + L.toff), sa, ca);
        double pot, gx, gy, gz;
        const double ex = L.y[0] * ca + L.y[1] * sa, ey = L.y[1] * ca - L.y[0] * sa, ez = L.y[2];
        if constexpr(GRP) {
            if(group) {
                FetchCell fg(P.cells, P.cellBytes);
                multipoleCarCellGroup<AGX_GROUP>(P, fg, ex, ey, ez, pot, gx, gy, gz, lR, lT, plt,
                    (int)(threadIdx.x & (AGX_GROUP - 1)));
            } else
                rhsRKN<MODE, NFW>(P, ex, ey, ez, pot, gx, gy, gz, gR, gT, lR, lT, plt);
        } else
            rhsRKN<MODE, NFW>(P, ex, ey, ez, pot, gx, gy, gz, gR, gT, lR, lT, plt);
        if(!run) continue;
        double a[3];
        a[0] = -gx * ca + gy * sa;
        a[1] = -gy * ca - gx * sa;
        a[2] = -gz;
        advanceRKN<RBS, TGT>(A, L, a, pot, accRel);
        if(L.phase == PH_IDLE) fetchOrbit(A, L, group);
    }
}

}  // namespace

int agxOrbitKernelRKNBlocksPerCU(int precision, bool nfw)
{
    int nb = 0;
    hipError_t e;
    if(precision == AGX_FP32_TABLES)
        e = nfw ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitKernelRKN<-1, true>, RBS, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitKernelRKN<-1, false>, RBS, 0);
    else
        e = nfw ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitKernelRKN<AGX_FETCH_CELL, true>, RBS, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitKernelRKN<AGX_FETCH_CELL, false>, RBS, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

// whether the per-cell variant can take this expansion: its grids, power-law arrays and bucket
// table fit the kernel's LDS
bool agxOrbitRKNCellsFit(const AgxMultipoleDev& P)
{
    return P.nr + P.nt <= AGX_LDS_GRID &&
        (P.interp1d || (P.cells && P.nTauLut <= RKN_LDS_TAULUT &&
                        (6 + AGX_LEG_STRIDE) * (P.legLmax + 1) * (P.legLmax + 1) + P.legLmax + 1 <= RKN_LDS_PL));
}

// fetch: the orbit fetch setting (agx_set_fetch); the per-cell variant when it is AGX_FETCH_CELL
// and the expansion fits its LDS, else the node records.  Group mode (A.groupStride > 0) needs
// the per-cell variant, FP64, no NFW halo and no targets.
hipError_t agxLaunchOrbitRKN(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, int precision, int fetch,
    const AgxOrbitArgs& A, int numBlocks, hipStream_t stream)
{
    if(A.n <= 0) return hipSuccess;
    const bool nfw = P.nfwMass != 0;
    const dim3 g(numBlocks), b(RBS);
    const bool cells = fetch == AGX_FETCH_CELL && agxOrbitRKNCellsFit(P);
    if(A.groupStride > 0) {
        if(!cells || precision == AGX_FP32_TABLES || nfw || P.interp1d || A.targets || A.poolCounter ||
           !A.groupCounter || A.groupOrbits > A.n)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_CELL, false, false, true>), g, b, 0, stream, dP, A);
        return hipGetLastError();
    }
    if(A.targets) {   // target plug-ins: FP64 tables only
        if(precision == AGX_FP32_TABLES) return hipErrorInvalidValue;
        if(nfw) hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_VEC2, true, true>), g, b, 0, stream, dP, A);
        else hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_VEC2, false, true>), g, b, 0, stream, dP, A);
        return hipGetLastError();
    }
    if(precision == AGX_FP32_TABLES) {
        if(nfw) hipLaunchKernelGGL((orbitKernelRKN<-1, true>), g, b, 0, stream, dP, A);
        else hipLaunchKernelGGL((orbitKernelRKN<-1, false>), g, b, 0, stream, dP, A);
    } else if(cells) {
        if(nfw) hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_CELL, true>), g, b, 0, stream, dP, A);
        else hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_CELL, false>), g, b, 0, stream, dP, A);
    } else {
        if(nfw) hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_VEC2, true>), g, b, 0, stream, dP, A);
        else hipLaunchKernelGGL((orbitKernelRKN<AGX_FETCH_VEC2, false>), g, b, 0, stream, dP, A);
    }
    return hipGetLastError();
}
