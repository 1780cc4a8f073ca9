// Persistent DOP853 orbit kernel with the variational equation: the six deviation vectors of
// agama.orbit(der=True) and the deviation-vector growth behind lyapunov=True, restating
// RuntimeVariational (src/orbit_variational.cpp:63-404, inertial-frame form) with the
// Gauss-Legendre stepper Ode2StepperGL4<3> (src/math_ode.cpp:1017-1250), and the orbital
// period T_circ(E) that sets the Lyapunov exponent's unit (src/potential_utils.cpp:264-289,
// 608-619; src/math_core.cpp:866-1040; src/potential_base.cpp:257-283, 529-575).
//
// The orbit part is the DOP853 phase machine of agx_dop853.h.  After each accepted step a lane
// takes four more evaluations - the Hessian at the step's dense-output positions for the
// collocation times h*g_k (evalMat, orbit_variational.cpp:85-119) - then advances its six
// deviation vectors by one GL4 step, samples them like the trajectory and renormalises them
// when they have grown by e^10.  An iteration in which any lane of the wavefront needs a
// Hessian evaluates the Hessian path for the whole wavefront (it also returns the potential
// and gradient the other lanes need), so control flow stays uniform.
// Per-lane variational state (160 doubles: GL4 state of up to six vectors, the four A matrices,
// the four positions and bookkeeping) lives in global memory, field-major across the grid's
// lanes so that a wavefront's accesses are contiguous.
// NV = 6: the deviation vectors start as the identity and their weighted sum is the fiducial
// vector; NV = 1 (lyapunov=True without der=True): one vector with the fixed initial direction.
// The log-magnitude series of the fiducial vector is coarse-grained on the fly into the bins
// of calcLyapunovExponent (orbit_variational.cpp:372-404); the fit runs on the host.
#include <hip/hip_runtime.h>
#include "agx_device.h"
#include "agx_hessian.h"
#include "agx_dop853.h"
#include "agx_kernels.h"

namespace agxd {

enum : int { V_GL = 0, V_AM = 108, V_PX = 132, V_T0 = 144, V_ADDLOG = 145, V_DEVSIZE = 146, V_NEWSTEP = 147,
    V_PREVDT = 148, V_PERIOD = 149, V_BSTART = 150, V_BOFF = 151, V_BSUM = 152, V_BSIZE = 153,
    V_PREVT = 154, V_PREVV = 155, V_NBINS = 156, V_NPTS = 157, V_MINBIN = 158, V_COUNT = 160 };
enum : int { PV_0 = 40, PV_1, PV_2, PV_3 };

// Gauss-Legendre nodes on [0, 1] and the derived weights (math_ode.cpp:1081-1131)
namespace gl4 {
constexpr double g0 = 0.069431844202973713, g1 = 0.33000947820757187, g2 = 1 - g1, g3 = 1 - g0,
    su0 = 0.5 * g0 * g0, sv0 = su0 * g0 * -2 / 3, sw0 = su0 * g0 * (g1 * 2 / 3 - g0 / 6),
    sz0 = su0 * g0 * (g0 * (g1 + g2) / 6 - g0 * g0 / 15 - g1 * g2 * 2 / 3),
    su1 = 0.5 * g1 * g1, sv1 = su1 * (g1 / 3 - g0), sw1 = su1 * g1 * (g0 * 2 / 3 - g1 / 6),
    sz1 = su1 * g1 * (g1 * (g0 + g2) / 6 - g1 * g1 / 15 - g0 * g2 * 2 / 3),
    su2 = 0.5 * g2 * g2, sv2 = su2 * (g2 / 3 - g0), sw2 = su2 * (g2 / 3 * (g2 / 2 - g0 - g1) + g0 * g1),
    sz2 = su2 * g2 * (g2 * (g0 + g1) / 6 - g2 * g2 / 15 - g0 * g1 * 2 / 3),
    su3 = 0.5 * g3 * g3, sv3 = su3 * (g3 / 3 - g0), sw3 = su3 * (g3 / 3 * (g3 / 2 - g0 - g1) + g0 * g1),
    sz3 = su3 * (g3 * (g3 * (g3 * 0.6 - g0 - g1 - g2) / 6 + (g0 * g1 + g1 * g2 + g0 * g2) / 3) - g0 * g1 * g2),
    du0 = g0, dv0 = g0 * -g0 / 2, dw0 = g0 * g0 * (g1 / 2 - g0 / 6),
    dz0 = g0 * g0 * (g0 * (-g0 / 12 + g1 / 6 + g2 / 6) - g1 * g2 / 2),
    du1 = g1, dv1 = g1 * (g1 / 2 - g0), dw1 = g1 * g1 * (g0 / 2 - g1 / 6),
    dz1 = g1 * g1 * (g1 * (-g1 / 12 + g0 / 6 + g2 / 6) - g0 * g2 / 2),
    du2 = g2, dv2 = g2 * (g2 / 2 - g0), dw2 = g2 * (g2 * (g2 / 3 - g1 / 2 - g0 / 2) + g0 * g1),
    dz2 = g2 * g2 * (g2 * (-g2 / 12 + g0 / 6 + g1 / 6) - g0 * g1 / 2),
    du3 = g3, dv3 = g3 * (g3 / 2 - g0), dw3 = g3 * (g3 * (g3 / 3 - g1 / 2 - g0 / 2) + g0 * g1),
    dz3 = g3 * (g3 * (g0 * g1 / 2 + g0 * g2 / 2 + g1 * g2 / 2 - g3 * (g0 / 3 + g1 / 3 + g2 / 3 - g3 / 4)) - g0 * g1 * g2),
    mvu = 1 / (g1 - g0), mwu = 1 / (g2 - g0) / (g2 - g1), mwv = 1 / (g2 - g1),
    mzu = 1 / (g3 - g0) / (g3 - g1) / (g3 - g2), mzv = 1 / (g3 - g1) / (g3 - g2), mzw = 1 / (g3 - g2),
    mpu = 1. / 2, mpv = 1. / 2 * (1 - g0), mpw = 1. / 2 * (1 - g0) * (1 - g1),
    mpz = 1. / 2 * (1 - g0) * (1 - g1) * (1 - g2),
    mqv = 1. / 6, mqw = 1. / 6 * (2 - g0 - g1),
    mqz = 1. / 6 * ((1 - g0) * (1 - g1) + (1 - g1) * (1 - g2) + (1 - g2) * (1 - g0)),
    mrw = 1. / 12, mrz = 1. / 12 * (3 - g0 - g1 - g2), msz = 1. / 20;
constexpr double G[4] = { g0, g1, g2, g3 };
}

// initial direction of the single deviation vector (orbit_variational.cpp:45-47); the fiducial
// vector of the six-vector form is their combination with these weights
__device__ __forceinline__ double devVecInit(int i)
{
    // sqrt(7/90), sqrt(11/90), sqrt(13/90), sqrt(17/90), sqrt(19/90), sqrt(23/90)
    return i == 0 ? 0.27888667551135854 : i == 1 ? 0.3496029493900505 : i == 2 ? 0.38005847503304596 :
           i == 3 ? 0.4346134936801766 : i == 4 ? 0.45946829173634074 : 0.5055250296034367;
}

// outputs of the Lyapunov bookkeeping (null pointers: not requested)
struct AgxLyapOut {
    const double* period;   // [n] orbital period T_circ(E) of each orbit's starting point
    const double* maxTime;  // [n] duration used for the bin width (the requested |T|)
    double* bins;           // [n][2][maxBins]: bin times and ln w
    int* nbins;             // [n]
    double* lastTime;       // [n] time of the last recorded point since the start (NaN: none)
    int maxBins;
};

template<int STRIDE, int NV>
struct LaneVarT : public LaneT<STRIDE, false> {
    typedef LaneT<STRIDE, false> Base;
    static constexpr int NVEC = NV;
    double* vg;      // this lane's variational state (field-major, stride vstride)
    long vstride;
    double* dev;     // deviation-vector output (n rows * 36 doubles), or null
    AgxLyapOut ly;
    __device__ __forceinline__ double& v(int field) const { return vg[(long)field * vstride]; }
};

// GL4 getSol (math_ode.cpp:1040-1065) for component i (0..35) at offset timeOffset of the last step
template<int S, int NV>
__device__ __forceinline__ double gl4Sol(const LaneVarT<S, NV>& L, double timeOffset, int i)
{
    const double T = timeOffset - L.v(V_PREVDT);
    const int vec = i / 6, c = i - vec * 6;
    const int base = V_GL + vec * 18;
    if(T == 0) return L.v(base + c);
    const int d = c % 3;
    const double x = L.v(base + d), xd = L.v(base + 3 + d), p = L.v(base + 6 + d), q = L.v(base + 9 + d),
        r = L.v(base + 12 + d), s = L.v(base + 15 + d);
    if(c == d) return x + T * (xd + T * (p + T * (q + T * (r + T * s))));
    return xd + T * (2 * p + T * (3 * q + T * 4 * (r + T * 5 * s)));
}

__device__ __forceinline__ double logMagnitude6(const double* x)
{
    double norm = 0.;
    for(int i = 0; i < 6; i++) norm += x[i] * x[i];
    return 0.5 * log(norm);
}

// storeSolution (orbit_variational.cpp:201-254): write the six vectors at row ind (ind < 0: none)
// in the rotating frame, return the log-magnitude of the fiducial vector
template<int S, int NV>
__device__ __forceinline__ double storeSolution(const AgxOrbitArgs& A, LaneVarT<S, NV>& L, double tbegin,
    double timeOffset, long ind)
{
    const double mul = exp(L.v(V_ADDLOG));
    double fid[6] = { 0, 0, 0, 0, 0, 0 };
    double ca = 1, sa = 0;
    if(A.Omega != 0) sincosA(A.Omega * (tbegin + timeOffset), sa, ca);
    const bool write = NV == 6 && L.dev && ind >= 0;
    // the row's place: the orbit's block ([vector][cap][6]) or, beyond it, the chunk of the
    // overflow pool that the trajectory row of the same index took ([vector][chunk][6])
    double* rowBase = nullptr;
    long vstride = 0;
    if(write) {
        L.v(V_DEVSIZE) = (double)(ind + 1);
        const long cap = rowCapacity(L.S);
        if(ind < cap) {
            rowBase = L.dev + (trajOffsetOf(A, L.oi) * 36 + ind * 6);
            vstride = cap * 6;
        } else if(L.S < 0 && A.poolCounter) {
            const long k = (ind - cap) / A.poolChunk, r = ind - cap - k * A.poolChunk;
            const int c = k < A.maxChunks ? A.chunkDir[(long)L.oi * A.maxChunks + k] : -1;
            if(c >= 0) {
                rowBase = L.dev + ((A.poolBase + (long)c * A.poolChunk) * 36 + r * 6);
                vstride = (long)A.poolChunk * 6;
            }
        }
    }
    for(int vec = 0; vec < NV; vec++) {
        double q[6];
        for(int c = 0; c < 6; c++) q[c] = gl4Sol(L, timeOffset, vec * 6 + c);
        if(NV == 1) return logMagnitude6(q);
        if(rowBase) {
            double* o = rowBase + vec * vstride;
            o[0] = (q[0] * ca + q[1] * sa) * mul;
            o[1] = (q[1] * ca - q[0] * sa) * mul;
            o[2] = (q[2] * mul);
            o[3] = (q[3] * ca + q[4] * sa) * mul;
            o[4] = (q[4] * ca - q[3] * sa) * mul;
            o[5] = (q[5] * mul);
        }
        const double w = devVecInit(vec);
        for(int c = 0; c < 6; c++) fid[c] += q[c] * w;
    }
    return logMagnitude6(fid);
}

// six deviation vectors = the identity rotated into the inertial frame at tbegin (:136-155), or
// the single vector along the fixed initial direction (:157-160)
template<int S, int NV>
__device__ __forceinline__ void varInit(const AgxOrbitArgs& A, LaneVarT<S, NV>& L, double tbegin)
{
    L.v(V_NEWSTEP) = 1;
    L.v(V_PREVDT) = 0;
    L.v(V_T0) = tbegin;
    if(NV == 1) {
        for(int i = 0; i < 18; i++) L.v(V_GL + i) = i < 6 ? devVecInit(i) : 0;
        return;
    }
    double ca, sa;
    sincosA(A.Omega * tbegin, sa, ca);
    for(int i = 0; i < 108; i++) L.v(V_GL + i) = 0;
    auto set = [&](int flat, double val) { const int vec = flat / 6, c = flat - vec * 6; L.v(V_GL + vec * 18 + c) = val; };
    set(0, ca); set(7, ca); set(21, ca); set(28, ca);
    set(1, sa); set(22, sa);
    set(6, -sa); set(27, -sa);
    set(14, 1); set(35, 1);
    storeSolution(A, L, tbegin, 0, 0);
}

// Ode2StepperGL4::doStep for the six vectors with the stored A matrices (B = 0)
template<int S, int NV>
__device__ __forceinline__ void gl4Step(LaneVarT<S, NV>& L, double dt)
{
    using namespace gl4;
    const double dt2 = dt * dt, dt3 = dt2 * dt, idt = 1. / dt, idt2 = idt * idt, idt3 = idt2 * idt;
    double a[4][9];
    for(int k = 0; k < 4; k++) {
        const int f = V_AM + k * 6;
        a[k][0] = L.v(f); a[k][1] = L.v(f + 1); a[k][2] = L.v(f + 2);
        a[k][3] = a[k][1]; a[k][4] = L.v(f + 3); a[k][5] = L.v(f + 4);
        a[k][6] = a[k][2]; a[k][7] = a[k][5]; a[k][8] = L.v(f + 5);
    }
    const int niter = L.v(V_NEWSTEP) != 0 ? 8 : 4;
    for(int vec = 0; vec < NV; vec++) {
        const int b = V_GL + vec * 18;
        double x[3], xd[3], p[3], q[3], r[3], s[3];
        for(int d = 0; d < 3; d++) {
            x[d] = L.v(b + d); xd[d] = L.v(b + 3 + d); p[d] = L.v(b + 6 + d); q[d] = L.v(b + 9 + d);
            r[d] = L.v(b + 12 + d); s[d] = L.v(b + 15 + d);
        }
        double x0[3], x1[3], x2[3], x3[3], u[3], v[3], w[3], z[3], xh[3];
        for(int d = 0; d < 3; d++) {
            x0[d] = x[d] + xd[d] * dt * g0;
            x1[d] = x[d] + xd[d] * dt * g1;
            x2[d] = x[d] + xd[d] * dt * g2;
            x3[d] = x[d] + xd[d] * dt * g3;
            const double Q = dt * q[d], R = dt2 * r[d], Sv = dt3 * s[d];
            u[d] = 2 * p[d] + 6 * g0 * Q + 12 * g0 * g0 * R + 20 * g0 * g0 * g0 * Sv;
            v[d] = 6 * Q + 12 * (g0 + g1) * R + 20 * (g0 * g0 + g0 * g1 + g1 * g1) * Sv;
            w[d] = 12 * R + 20 * (g0 + g1 + g2) * Sv;
            z[d] = 20 * Sv;
        }
        auto rhs = [&](int k, int d) {
            double acc = 0;
            for(int j = 0; j < 3; j++) acc += a[k][d * 3 + j] * xh[j];
            return acc;
        };
        for(int it = 0; it < niter; it++) {
            for(int d = 0; d < 3; d++) xh[d] = x0[d] + dt2 * (su0 * u[d] + sv0 * v[d] + sw0 * w[d] + sz0 * z[d]);
            for(int d = 0; d < 3; d++) u[d] = rhs(0, d);
            for(int d = 0; d < 3; d++) xh[d] = x1[d] + dt2 * (su1 * u[d] + sv1 * v[d] + sw1 * w[d] + sz1 * z[d]);
            for(int d = 0; d < 3; d++) v[d] = (rhs(1, d) - u[d]) * mvu;
            for(int d = 0; d < 3; d++) xh[d] = x2[d] + dt2 * (su2 * u[d] + sv2 * v[d] + sw2 * w[d] + sz2 * z[d]);
            for(int d = 0; d < 3; d++) w[d] = (rhs(2, d) - u[d]) * mwu - v[d] * mwv;
            for(int d = 0; d < 3; d++) xh[d] = x3[d] + dt2 * (su3 * u[d] + sv3 * v[d] + sw3 * w[d] + sz3 * z[d]);
            for(int d = 0; d < 3; d++) z[d] = (rhs(3, d) - u[d]) * mzu - v[d] * mzv - w[d] * mzw;
        }
        for(int d = 0; d < 3; d++) {
            const double P = mpz * z[d] + mpw * w[d] + mpv * v[d] + mpu * u[d],
                Q = mqz * z[d] + mqw * w[d] + mqv * v[d], R = mrz * z[d] + mrw * w[d], Sv = msz * z[d];
            L.v(b + 6 + d) = P;
            L.v(b + 9 + d) = Q * idt;
            L.v(b + 12 + d) = R * idt2;
            L.v(b + 15 + d) = Sv * idt3;
            L.v(b + d) = x[d] + ((P - 2 * Q + 3 * R - 4 * Sv) * dt2 + xd[d] * dt);
            L.v(b + 3 + d) = xd[d] + (2 * P - 3 * Q + 4 * R - 5 * Sv) * dt;
        }
    }
    L.v(V_PREVDT) = dt;
    L.v(V_NEWSTEP) = 0;
}

// renormalize (orbit_variational.cpp:256-308, rescaling without orthogonalisation)
template<int S, int NV>
__device__ __forceinline__ void varRenormalize(LaneVarT<S, NV>& L, double timeStep)
{
    double lg = -INFINITY;
    for(int vec = 0; vec < NV; vec++) {
        double q[6];
        for(int c = 0; c < 6; c++) q[c] = gl4Sol(L, timeStep, vec * 6 + c);
        lg = fmax(lg, logMagnitude6(q));
    }
    const double mult = exp(-lg);
    L.v(V_ADDLOG) = L.v(V_ADDLOG) + lg;
    double vals[6 * NV];
    for(int i = 0; i < 6 * NV; i++) vals[i] = gl4Sol(L, timeStep, i) * mult;
    for(int i = 0; i < 18 * NV; i++) L.v(V_GL + i) = 0;
    for(int vec = 0; vec < NV; vec++)
        for(int c = 0; c < 6; c++) L.v(V_GL + vec * 18 + c) = vals[vec * 6 + c];
    L.v(V_NEWSTEP) = 1;
}

// calcLyapunovExponent's coarse-graining (orbit_variational.cpp:380-404), fed one point
// (time since start, ln|w|) at a time: each new point closes the pair (previous, new)
template<int S, int NV>
__device__ __forceinline__ void binPoint(LaneVarT<S, NV>& L, double tp, double vp)
{
    const double np = L.v(V_NPTS);
    L.v(V_NPTS) = np + 1;
    if(np > 0) {
        const double ti = L.v(V_PREVT), vi = L.v(V_PREVV);
        double bstart = L.v(V_BSTART);
        if(bstart != bstart) {
            bstart = ti;
            L.v(V_BSTART) = ti;
            L.v(V_BOFF) = vi;
            L.v(V_BSUM) = 0;
            L.v(V_BSIZE) = 0;
        }
        const double sum = L.v(V_BSUM) + (tp - ti) * exp(0.5 * (vi + vp) - L.v(V_BOFF));
        const double size = L.v(V_BSIZE) + 1;
        L.v(V_BSUM) = sum;
        L.v(V_BSIZE) = size;
        if(size >= 5 && tp >= bstart + L.v(V_MINBIN)) {
            const int nb = (int)L.v(V_NBINS);
            if(nb < L.ly.maxBins) {
                double* b = L.ly.bins + (long)L.oi * 2 * L.ly.maxBins;
                b[nb] = 0.5 * (bstart + tp);
                b[L.ly.maxBins + nb] = log(sum / (tp - bstart)) + L.v(V_BOFF);
            }
            L.v(V_NBINS) = nb + 1;
            L.v(V_BSTART) = NAN;
        }
    }
    L.v(V_PREVT) = tp;
    L.v(V_PREVV) = vp;
}

// the series ended: its last pair closes the open bin when it holds at least 5 points
template<int S, int NV>
__device__ __forceinline__ void binFinish(LaneVarT<S, NV>& L)
{
    const double bstart = L.v(V_BSTART);
    if(bstart == bstart && L.v(V_BSIZE) >= 5) {
        const int nb = (int)L.v(V_NBINS);
        const double tp = L.v(V_PREVT);
        if(nb < L.ly.maxBins) {
            double* b = L.ly.bins + (long)L.oi * 2 * L.ly.maxBins;
            b[nb] = 0.5 * (bstart + tp);
            b[L.ly.maxBins + nb] = log(L.v(V_BSUM) / (tp - bstart)) + L.v(V_BOFF);
        }
        L.v(V_NBINS) = nb + 1;
    }
    L.ly.nbins[L.oi] = (int)L.v(V_NBINS);
    L.ly.lastTime[L.oi] = L.v(V_NPTS) > 0 ? L.v(V_PREVT) : NAN;
}

// the generic bookkeeping plus the deviation-vector rows that were not recorded (NaN), and the
// zero-length case where the start is recorded (orbit.cpp:52-57 with RuntimeVariational)
template<int S, int NV>
__device__ __forceinline__ void finishOrbit(const AgxOrbitArgs& A, LaneVarT<S, NV>& L, int status)
{
    const bool zeroLength = status == 0 && L.v(V_T0) != L.v(V_T0) &&
        (A.ttotal[L.oi] == 0 || !(L.tend > -INFINITY && L.tend < INFINITY));
    if(zeroLength) {
        varInit(A, L, L.tb);
        // doStep(0) leaves x and x' unchanged; sampling interval 0 records the start again
        if(NV == 6 && L.dev && (L.S >= 2 || L.S < 0)) storeSolution(A, L, L.tb, 0, (long)L.v(V_DEVSIZE));
        if(L.ly.bins) binPoint(L, fabs(L.tb - L.v(V_T0)), storeSolution(A, L, L.tb, 0, -1));
    }
    if(L.ly.bins) binFinish(L);
    if(NV == 6 && L.dev) {
        const long done = L.v(V_T0) == L.v(V_T0) ? (long)L.v(V_DEVSIZE) : 0;
        const long base = trajOffsetOf(A, L.oi) * 36;
        for(int vec = 0; vec < 6; vec++)
            for(long r = done; r < L.S; r++)
                for(int c = 0; c < 6; c++) L.dev[base + ((long)vec * L.S + r) * 6 + c] = NAN;
    }
    finishOrbit<typename LaneVarT<S, NV>::Base>(A, L, status);
}

template<int S, int NV>
__device__ __forceinline__ void fetchOrbit(const AgxOrbitArgs& A, LaneVarT<S, NV>& L)
{
    fetchOrbit<typename LaneVarT<S, NV>::Base>(A, L);
    if(L.phase != PH_IDLE) {
        L.v(V_T0) = NAN;
        L.v(V_ADDLOG) = 0;
        L.v(V_DEVSIZE) = 0;
        if(L.ly.bins) {
            const double period = L.ly.period[L.oi];
            L.v(V_PERIOD) = period;
            L.v(V_MINBIN) = fmax(L.ly.maxTime[L.oi] * 0.01, period);
            L.v(V_BSTART) = NAN;
            L.v(V_NBINS) = 0;
            L.v(V_NPTS) = 0;
        }
    }
}

// accepted orbit step: energy, trajectory sampling and the step's dense-output positions at the
// collocation times; the orbit state advances, its clock after the variational step
template<int S, int NV>
__device__ __forceinline__ void acceptStep(const AgxOrbitArgs& A, LaneVarT<S, NV>& L, double pot)
{
    const double de = fabs(energyOf(L.y, pot, A.Omega) / L.f(F_E0) - 1);
    const double mde = L.f(F_MAXDE);
    if(de > mde || de != de) L.f(F_MAXDE) = de;
    sampleStep(A, L);
    L.f(F_LASTTB) = L.tb;
    L.f(F_LASTH) = L.h;
    if(L.S == 1) L.actual = 1;
    for(int k = 0; k < 4; k++)
        for(int c = 0; c < 3; c++) L.v(V_PX + k * 3 + c) = denseOut(L, c, L.h * gl4::G[k]);
    for(int i = 0; i < 6; i++) { L.x[i] = L.y[i]; L.k1[i] = L.k4[i]; }
    for(int c = 0; c < 3; c++) L.y[c] = L.v(V_PX + c);
    L.toff = L.h * gl4::g0;
    L.phase = PV_0;
}

// one Hessian (Car layout) at the k-th collocation position, rotated into the inertial frame
// with the same angle as the evaluation point; after the fourth: the variational step
template<int S, int NV>
__device__ __forceinline__ void advanceVar(const AgxOrbitArgs& A, LaneVarT<S, NV>& L, const double (&hs)[6],
    double ca, double sa)
{
    const int k = L.phase - PV_0;
    const double dx2 = hs[0], dy2 = hs[1], dz2 = hs[2], dxdy = hs[3], dydz = hs[4], dxdz = hs[5];
    const int f = V_AM + k * 6;
    L.v(f)     = -dx2 * (ca * ca) - dy2 * (sa * sa) + 2 * dxdy * (ca * sa);
    L.v(f + 1) = -dxdy * (ca * ca - sa * sa) - (dx2 - dy2) * (ca * sa);
    L.v(f + 2) = -dxdz * ca + dydz * sa;
    L.v(f + 3) = -dy2 * (ca * ca) - dx2 * (sa * sa) - 2 * dxdy * (ca * sa);
    L.v(f + 4) = -dydz * ca - dxdz * sa;
    L.v(f + 5) = -dz2;
    if(k < 3) {
        for(int c = 0; c < 3; c++) L.y[c] = L.v(V_PX + (k + 1) * 3 + c);
        L.toff = L.h * gl4::G[k + 1];
        L.phase++;
        return;
    }
    // RuntimeVariational::processTimestep (orbit_variational.cpp:121-199)
    const double tbegin = L.tb, timestep = L.h, tend = tbegin + timestep;
    if(L.v(V_T0) != L.v(V_T0)) varInit(A, L, tbegin);
    gl4Step(L, timestep);
    const double t0 = L.v(V_T0);
    if(NV == 6 && L.dev && L.S > 1) {
        const double dtS = L.f(F_DTS);
        const double sign = timestep >= 0 ? +1 : -1;
        const double dtr = ROUNDOFF * fmax(fmax(fabs(tend), fabs(tbegin)), fabs(t0));
        long iout = (long)ceil(sign * (tbegin - t0) / dtS);
        const long iend = (long)((sign * (tend - t0) + dtr) / dtS);
        iout = iout < 1 ? 1 : iout;
        for(; iout <= iend; iout++) {
            const double xo = (sign * iout * dtS - (tbegin - t0)) * sign, hi = sign * timestep;
            storeSolution(A, L, tbegin, (xo > hi ? hi : xo > 0.0 ? xo : 0.0) * sign, iout);
        }
    }
    // one row (trajsize 1), the next row (sampling interval 0, trajsize < 0 here) or none
    const long ind = !(NV == 6 && L.dev) ? -1 : L.S == 1 ? 0 : L.S < 0 ? (long)L.v(V_DEVSIZE) : -1;
    const double logDevVec = storeSolution(A, L, tbegin, timestep, ind);
    if(L.ly.bins) binPoint(L, fabs(tend - t0), logDevVec + L.v(V_ADDLOG));
    if(logDevVec > 10) varRenormalize(L, timestep);
    // the orbit's clock and step count (orbit.cpp:79-82)
    L.tb = L.h == L.rem ? L.tend : L.tb + L.h;
    L.nsteps++;
    if(L.nsteps >= A.maxNumSteps) { finishOrbit(A, L, 2); return; }
    if(L.tb == L.tend) { finishOrbit(A, L, 0); return; }
    beginStep(A, L);
}

// ---- orbital period of each orbit's starting point (the Lyapunov exponent's unit) ----

// Axisymmetrized<BasePotential>::evalCyl at (R, z): potential, dPhi/dR and d2Phi/dR2 averaged
// over 2*8+1 azimuths, 9 with doubled weights for y-reflection symmetric potentials
// (potential_base.cpp:257-283, 529-575)
template<class F>
__device__ __forceinline__ void axisymmetrizedEval(const AgxMultipoleDev& P, const F& fetch, double R,
    double z, double& Phi, double& dR, double& dR2)
{
    double p, g[3], h[6];
    constexpr int ST_YREFL = 2, ST_ZROT = 16;
    if((P.sym & ST_ZROT) == ST_ZROT || (R == 0 && z == 0) || R == INFINITY || fabs(z) == INFINITY) {
        multipoleCylHess(P, fetch, R, z, 0., p, g, h);
        Phi = p; dR = g[0]; dR2 = h[0];
        return;
    }
    const int nphi = 8;
    const bool ysym = (P.sym & ST_YREFL) == ST_YREFL;
    const int npoints = ysym ? nphi + 1 : 2 * nphi + 1;
    double a0 = 0, a1 = 0, a3 = 0;
    for(int i = 0; i < npoints; i++) {
        multipoleCylHess(P, fetch, R, z, M_PI / (nphi + 0.5) * i, p, g, h);
        const double w = ysym && i > 0 ? 2 : 1;
        a0 += p * w;
        a1 += g[0] * w;
        a3 += h[0] * w;
    }
    Phi = a0 * (1. / (2 * nphi + 1));
    dR = a1 * (1. / (2 * nphi + 1));
    dR2 = a3 * (1. / (2 * nphi + 1));
}

__device__ __forceinline__ double hermiteInterpD(double x, double x1, double x2, double y1, double y2,
    double dy1, double dy2)
{
    const double t = (x - x1) / (x2 - x1);
    return (1 - t) * (1 - t) * ((1 + 2 * t) * y1 + t * dy1 * (x2 - x1))
         + t * t * ((3 - 2 * t) * y2 + (t - 1) * dy2 * (x2 - x1));
}

// math_core.cpp:866-888
__device__ __forceinline__ double hermiteInterpMonotoneD(double x, double x1, double x2, double f1, double f2,
    double dfdx1, double dfdx2)
{
    if(!(fabs(dfdx1 + dfdx2) < INFINITY) || (dfdx1 >= 0 && dfdx2 < 0) || (dfdx2 >= 0 && dfdx1 < 0)) return NAN;
    const double dx = x2 - x1, ai = 1 / (-6 * (f2 - f1) / dx + 3 * (dfdx1 + dfdx2)),
        ba = -1 + (dfdx2 - dfdx1) * ai, ca = dfdx1 * ai, D = ba * ba - 4 * ca;
    if(!(fabs(ba + D) < INFINITY)) return NAN;
    if(D >= 0) {
        const double sqD = sqrt(D), t1 = 0.5 * (-ba - sqD), t2 = 0.5 * (-ba + sqD);
        if((t1 >= 0 && t1 <= 1) || (t2 >= 0 && t2 <= 1)) return NAN;
    }
    return hermiteInterpD(x, x1, x2, f1, f2, dfdx1, dfdx2);
}

// RcircRootFinder (potential_utils.cpp:264-289) of u = ln R seen through ScalingInf
// (u = 1/(1-s) - 1/s, math_core.cpp:302-306) as ScaledFnc (math_core.h:212-227) does
template<class F>
__device__ __forceinline__ void rcircScaled(const AgxMultipoleDev& P, const F& fetch, double E, double s,
    double& val, double& der)
{
    const double duds = 1 / ((1 - s) * (1 - s)) + 1 / (s * s);
    const double R = exp(1 / (1 - s) - 1 / s);
    double Phi, dR, dR2;
    axisymmetrizedEval(P, fetch, R, 0., Phi, dR, dR2);
    const double v = 0.5 * R * dR;
    val = Phi - E + (fabs(v) < INFINITY ? v : 0);
    if(!(fabs(val) < INFINITY)) {
        if(E == -INFINITY) val = Phi == E ? 0 : +1.0;
        else if(Phi == -INFINITY) val = -1.0;
        else if(R >= 1e100) val = +1.0;
    }
    der = ((1.5 * dR + 0.5 * R * dR2) * R) * duds;
}

// findRoot (math_core.cpp:895-1040) on s in [0, 1], relative tolerance 1e-10, and
// T_circ = 2 pi R / v_circ(R) (potential_utils.h:83-87, potential_utils.cpp:608-619)
template<class F>
__device__ double tcircOf(const AgxMultipoleDev& P, const F& fetch, double E)
{
    const double reltoler = 1e-10, DELTA_HERMITE = 1e-12;
    const int MAXITER = 64;
    double a = 0, b = 1, fa, fb, fdera, fderb;
    rcircScaled(P, fetch, E, a, fa, fdera);
    rcircScaled(P, fetch, E, b, fb, fderb);
    double sroot;
    if((fa < 0.0 && fb < 0.0) || (fa > 0.0 && fb > 0.0) || !(fabs(fa + fb) < INFINITY)) sroot = NAN;
    else if(fa == 0) sroot = fb == 0 ? 0.5 * (a + b) : a;
    else if(fb == 0) sroot = b;
    else {
        double c = a, fc = fa, fderc = fdera, d = b - c, e = b - c;
        if(fabs(fc) < fabs(fb)) {
            a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; fdera = fderb; fderb = fderc; fderc = fdera;
        }
        int numIter = 0;
        bool converged = false, done = false;
        const double abstoler = fabs(0. - 1.) * reltoler;
        while(!converged) {
            const double tol = 0.5 * 2.2204460492503131e-16 * fabs(b), cminusb = c - b, m = 0.5 * cminusb;
            if(fb == 0 || fabs(m) <= tol) { done = true; break; }
            if(fabs(e) < tol || fabs(fa) <= fabs(fb)) { d = m; e = m; }
            else {
                double dd = NAN;
                if(fderb * fderc > 0) dd = hermiteInterpMonotoneD(0, fb, fc, 0, cminusb, 1 / fderb, 1 / fderc);
                if(fabs(dd) < INFINITY && fmin(fabs(dd), fabs(cminusb - dd)) > fabs(cminusb) * DELTA_HERMITE)
                    d = dd;
                else {
                    double p, q, r, sr = fb / fa;
                    if(a == c) { p = 2 * m * sr; q = 1 - sr; }
                    else {
                        q = fa / fc; r = fb / fc;
                        p = sr * (2 * m * q * (q - r) - (b - a) * (r - 1));
                        q = (q - 1) * (r - 1) * (sr - 1);
                    }
                    if(p > 0) q = -q; else p = -p;
                    if(2 * p < fmin(3 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
                    else { d = m; e = m; }
                }
            }
            a = b; fa = fb; fdera = fderb;
            if(fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
            rcircScaled(P, fetch, E, b, fb, fderb);
            if(!(fabs(fb) < INFINITY)) { b = NAN; done = true; break; }
            if((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { c = a; fc = fa; fderc = fdera; d = b - c; e = b - c; }
            if(fabs(fc) < fabs(fb)) {
                a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; fdera = fderb; fderb = fderc; fderc = fdera;
            }
            numIter++;
            if(fabs(b - c) <= abstoler) {
                converged = true;
                const double offset = fb * (b - c) / (fc - fb);
                if((offset > 0 && offset < c - b) || (offset < 0 && offset > c - b)) b += offset;
            } else {
                const double offset = -fb / fderb;
                const bool bracketed = (offset > 0 && offset < c - b) || (offset < 0 && offset > c - b);
                if(bracketed && fabs(offset) < abstoler) { converged = true; b += offset; }
            }
            if(numIter >= MAXITER) converged = true;
        }
        (void)done;
        sroot = b;
    }
    const double R = exp(1 / (1 - sroot) - 1 / sroot);
    double vc;
    double Phi, dR, dR2;
    if(R == 0) {
        axisymmetrizedEval(P, fetch, 0., 0., Phi, dR, dR2);
        vc = fabs(Phi) < INFINITY ? 0 : INFINITY;
    } else {
        axisymmetrizedEval(P, fetch, R, 0., Phi, dR, dR2);
        vc = sqrt(R * dR);
    }
    return R / vc * 2 * M_PI;
}

}  // namespace agxd

using namespace agxd;

namespace {

constexpr int VBS = 256;

// period[i] = T_circ(E) at orbit i's starting point as the orbit integrator reports it: the
// initial condition taken into the inertial frame at tstart and back (orbit.cpp:104-133), its
// energy Phi + v^2/2 (orbit_variational.cpp:131-135); non-finite values become 0
__global__ __launch_bounds__(256) void tcircKernel(const AgxMultipoleDev* __restrict__ Pp, long n,
    const double* __restrict__ ic, const double* __restrict__ tstart, double Omega, double* __restrict__ period)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const AgxMultipoleDev& P = *Pp;
    double x[6];
    for(int c = 0; c < 6; c++) x[c] = ic[i * 6 + c];
    if(Omega != 0) {
        const double tb = tstart ? tstart[i] : 0.0;
        double ca = 1, sa = 0;
        sincosA(Omega * tb, sa, ca);
        double w[6];
        w[0] = x[0] * ca - x[1] * sa; w[1] = x[1] * ca + x[0] * sa; w[2] = x[2];
        w[3] = x[3] * ca - x[4] * sa; w[4] = x[4] * ca + x[3] * sa; w[5] = x[5];
        sincosA(Omega * (0 + tb), sa, ca);
        x[0] = w[0] * ca + w[1] * sa; x[1] = w[1] * ca - w[0] * sa; x[2] = w[2];
        x[3] = w[3] * ca + w[4] * sa; x[4] = w[4] * ca - w[3] * sa; x[5] = w[5];
    }
    FetchVec2 fv{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
    double pot, gx, gy, gz;
    multipoleCar<FetchVec2, const double AGX_GLOBAL*, true>(P, fv, x[0], x[1], x[2], pot, gx, gy, gz);
    const double E = pot + 0.5 * (x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
    const double T = tcircOf(P, fv, E);
    period[i] = fabs(T) < INFINITY ? T : 0;
}

// varState: V_COUNT * (gridDim.x * VBS) doubles; dev: deviation-vector output (NV = 6);
// ly: Lyapunov bookkeeping (ly.bins null: none)
template<bool NFW, int NV>
__global__ __launch_bounds__(VBS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void orbitVarKernel(const AgxMultipoleDev* __restrict__ Pp, AgxOrbitArgs A, double* varState, double* dev,
    AgxLyapOut ly)
{
    typedef LaneVarT<VBS, NV> Lane;
    __shared__ double lds[Lane::LDS_FIELDS * VBS];
    __shared__ double gridLds[AGX_LDS_GRID];
    const AgxMultipoleDev& P = *Pp;
    const double *gR, *gT;
    stageGrids(P, gridLds, gR, gT);
    Lane L;
    L.attach(lds + threadIdx.x);
    L.vstride = (long)gridDim.x * VBS;
    L.vg = varState + (long)blockIdx.x * VBS + threadIdx.x;
    L.dev = dev;
    L.ly = ly;
    fetchOrbit(A, L);
    const double accRel = A.accuracy;
    unsigned int it = 0;
    while(true) {
        const bool active = L.phase != PH_IDLE;
        if(!__any(active)) break;
        const bool run = active && !(L.phase == PH_S2 && it % 12 != 0);
        it++;
        double ca = 1, sa = 0;
        if(A.Omega != 0 && active) sincosA(A.Omega * (L.tb + L.toff), sa, ca);
        const double ex = L.y[0] * ca + L.y[1] * sa, ey = L.y[1] * ca - L.y[0] * sa, ez = L.y[2];
        const bool needH = active && L.phase >= PV_0;
        double pot, g[3], hs[6];
        FetchVec2 fv{ gmem(reinterpret_cast<const vd2*>(P.nodes)) };
        if(__any(needH))
            multipoleCarHess(P, fv, ex, ey, ez, pot, g, hs);
        else
            multipoleCar<FetchVec2, const double*, NFW, true>(P, fv, ex, ey, ez, pot, g[0], g[1], g[2], gR, gT);
        if(!run) continue;
        if(L.phase >= PV_0)
            advanceVar(A, L, hs, ca, sa);
        else {
            double f[6];
            f[0] = L.y[3]; f[1] = L.y[4]; f[2] = L.y[5];
            f[3] = -g[0] * ca + g[1] * sa;
            f[4] = -g[1] * ca - g[0] * sa;
            f[5] = -g[2];
            advancePhase(A, L, f, pot, accRel);
        }
        if(L.phase == PH_IDLE) fetchOrbit(A, L);
    }
}

}  // namespace

int agxOrbitVarKernelBlocksPerCU(bool nfw)
{
    int nb = 0;
    hipError_t e = nfw ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitVarKernel<true, 6>, VBS, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, orbitVarKernel<false, 6>, VBS, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

size_t agxOrbitVarStateDoubles(int numBlocks) { return (size_t)V_COUNT * numBlocks * VBS; }

hipError_t agxLaunchTcirc(const AgxMultipoleDev* dP, long n, const double* ic, const double* tstart,
    double Omega, double* period, hipStream_t stream)
{
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tcircKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, dP, n, ic, tstart,
        Omega, period);
    return hipGetLastError();
}

// dev non-null: six deviation vectors (der=True), else one vector; bins non-null: Lyapunov
// bookkeeping with period[n], maxTime[n] in and bins[n][2][maxBins], nbins[n], lastTime[n] out
hipError_t agxLaunchOrbitVar(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, const AgxOrbitArgs& A,
    int numBlocks, double* varState, double* dev, const double* period, const double* maxTime, double* bins,
    int* nbins, double* lastTime, int maxBins, hipStream_t stream)
{
    if(A.n <= 0) return hipSuccess;
    AgxLyapOut ly{ period, maxTime, bins, nbins, lastTime, maxBins };
    const dim3 g(numBlocks), b(VBS);
    if(dev) {
        if(P.nfwMass != 0) hipLaunchKernelGGL((orbitVarKernel<true, 6>), g, b, 0, stream, dP, A, varState, dev, ly);
        else hipLaunchKernelGGL((orbitVarKernel<false, 6>), g, b, 0, stream, dP, A, varState, dev, ly);
    } else {
        if(P.nfwMass != 0) hipLaunchKernelGGL((orbitVarKernel<true, 1>), g, b, 0, stream, dP, A, varState, dev, ly);
        else hipLaunchKernelGGL((orbitVarKernel<false, 1>), g, b, 0, stream, dP, A, varState, dev, ly);
    }
    return hipGetLastError();
}
