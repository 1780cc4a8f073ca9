// The DOP853 integrator of the orbit kernels as device functions on a per-lane state: the
// coefficients, the stage/step phase machine, dense output, trajectory sampling and the
// work-queue bookkeeping.  Restates src/math_ode.cpp:35-423 (initTimeStep, DOP853
// init/doStep/getSol) and src/orbit.cpp:15-133, 182-208 (trajectory sampling, run loop,
// rotating frame).  The lane state type is a template parameter so that kernels with
// different block sizes (the LDS stride) share the code.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "agx_device.h"
#include "agx_kernels.h"
#include "agx_target.h"

namespace agxd {

// ---------------- DOP853 coefficients (src/math_ode.cpp:111-236) ----------------
constexpr double
    c2 = 0.05260015195876773187856, c3 = 0.07890022793815159781784, c4 = 0.11835034190722739672676,
    c5 = 0.28164965809277260327324, c6 = 0.33333333333333333333333, c7 = 0.25000000000000000000000,
    c8 = 0.30769230769230769230769, c9 = 0.65128205128205128205128, c10 = 0.60000000000000000000000,
    c11 = 0.85714285714285714285714,
    a21 = 0.05260015195876773187856,
    a31 = 0.01972505698453789945446, a32 = 0.05917517095361369836338,
    a41 = 0.02958758547680684918169, a43 = 0.08876275643042054754507,
    a51 = 0.24136513415926668550237, a53 = -0.88454947932828608534486, a54 = 0.92483400326179200311574,
    a61 = 0.03703703703703703703704, a64 = 0.17082860872947387127960, a65 = 0.12546768756682242501669,
    a71 = 0.03710937500000000000000, a74 = 0.17025221101954403931498, a75 = 0.06021653898045596068502,
    a76 = -0.01757812500000000000000,
    a81 = 0.03709200011850479271088, a84 = 0.17038392571223999381021, a85 = 0.10726203044637328465181,
    a86 = -0.01531943774862440175279, a87 = 0.00827378916381402288758,
    a91 = 0.62411095871607571711443, a94 = -3.36089262944694129406857, a95 = -0.86821934684172600681819,
    a96 = 27.5920996994467083049416, a97 = 20.1540675504778934086187, a98 = -43.4898841810699588477366,
    a101 = 0.47766253643826436589043, a104 = -2.48811461997166764192642, a105 = -0.59029082683684299637145,
    a106 = 21.2300514481811942347289, a107 = 15.2792336328824235832597, a108 = -33.2882109689848629194453,
    a109 = -0.02033120170850862613582,
    a111 = -0.93714243008598732571704, a114 = 5.18637242884406370830024, a115 = 1.09143734899672957818500,
    a116 = -8.14978701074692612513997, a117 = -18.5200656599969598641566, a118 = 22.7394870993505042818970,
    a119 = 2.49360555267965238987089, a1110 = -3.04676447189821950038237,
    a121 = 2.27331014751653820792360, a124 = -10.5344954667372501984067, a125 = -2.00087205822486249909676,
    a126 = -17.9589318631187989172766, a127 = 27.9488845294199600508500, a128 = -2.85899827713502369474066,
    a129 = -8.87285693353062954433549, a1210 = 12.3605671757943030647266, a1211 = 0.64339274601576353035597,
    b1 = 0.05429373411656876223805, b6 = 4.45031289275240888144114, b7 = 1.89151789931450038304282,
    b8 = -5.80120396001058478146721, b9 = 0.31116436695781989440892, b10 = -0.15216094966251607855618,
    b11 = 0.20136540080403034837478, b12 = 0.04471061572777259051769,
    bhh1 = 0.24409448818897637795276, bhh2 = 0.73384668828161185734136, bhh3 = 0.02205882352941176470588,
    er1 = 0.01312004499419488073250, er6 = -1.22515644637620444072057, er7 = -0.49575894965725019152141,
    er8 = 1.66437718245498653696153, er9 = -0.35032884874997368168865, er10 = 0.33417911871301747902973,
    er11 = 0.08192320648511571246571, er12 = -0.02235530786388629525884,
    d41 = -5.40685903845352664250302, d46 = 367.268892700041893590281, d47 = 154.609958204083905482676,
    d48 = -505.920283865412564024766, d49 = 15.5975154819608130688200, d410 = -26.1936204184402805956691,
    d411 = -0.74003512364122230844721, d412 = 1.11776539319431476294221, d413 = -0.33333333333333333333333,
    d51 = 6.51987095363079615048119, d56 = -1066.34956011730205278592, d57 = -351.864047514639508625601,
    d58 = 1363.51955696662884408368, d59 = -112.727669432657582669864, d510 = 159.796191868560289612921,
    d511 = -2.13865100308788816220259, d512 = -3.75569172113289760348584, d513 = 7.00000000000000000000000,
    d61 = 10.4698004763293477204238, d66 = -1380.01473607038123167155, d67 = -531.219827862514074379012,
    d68 = 1866.98964341870892451324, d69 = -53.3302605020547902574560, d610 = 82.4147560258671369782481,
    d611 = 7.38443654502992069572676, d612 = 0.41729908012587751149843, d613 = -3.11111111111111111111111,
    d71 = -16.6338582677165354330709, d76 = 4516.16568914956011730205, d77 = 1393.85185384057776465219,
    d78 = -5687.52042419481539670071, d79 = 473.965563750151263163661, d710 = -661.810776942355889724311,
    d711 = -18.0180473354013232598119,
    fdec = 0.333, finc = 6.0, safe = 0.9, ROUNDOFF = 10 * 2.2204460492503131e-16;

enum : int { PH_IDLE = -1, PH_INIT_A = 0, PH_INIT_B, PH_INIT_C, PH_S2, PH_S3, PH_S4, PH_S5, PH_S6,
    PH_S7, PH_S8, PH_S9, PH_S10, PH_S11, PH_S12, PH_FSAL };

// Per-lane state kept in LDS, field-major with a stride of STRIDE doubles (the block size) so that a wavefront
// touches consecutive addresses: the stage vectors k6..k10 (used in the later stage sums,
// the error estimate and dense output) and the once-per-step bookkeeping.  Together with the
// register state this keeps the kernel within 256 VGPRs (two wavefronts per SIMD) while two
// workgroups share one CU's 160 KB of LDS.
enum : int { F_K6 = 0, F_K7 = 6, F_K8 = 12, F_K9 = 18, F_K10 = 24, F_H1 = 30, F_ACC = 31, F_E0 = 32,
    F_MAXDE = 33, F_T0 = 34, F_DTS = 35, F_LASTTB = 36, F_LASTH = 37, F_COUNT = 38,
    // with the stage vectors k1..k5 and the step's start x in LDS as well
    F_X = 38, F_K1 = 44, F_K2 = 50, F_K3 = 56, F_K4 = 62, F_K5 = 68, F_COUNT_K = 74 };

// a 6-vector stored in a lane's LDS column (stride = block size)
template<int STRIDE>
struct LdsVec {
    double* p;
    __device__ __forceinline__ double& operator[](int i) const { return p[i * STRIDE]; }
};

// KLDS: keep x and k1..k5 in LDS too (fewer registers live across the potential evaluation,
// for kernels whose occupancy is set by registers and that have LDS to spare)
template<int STRIDE, bool KLDS = false>
struct LaneT {
    typedef typename std::conditional<KLDS, LdsVec<STRIDE>, double[6]>::type Vec;
    static constexpr int LDS_FIELDS = KLDS ? F_COUNT_K : F_COUNT;
    static constexpr int EVALS_PER_ATTEMPT = 12;
    static constexpr bool HAS_TARGET = false;
    double* sh;                // this lane's LDS column
    int oi, phase, nbad, S, nsteps, natt, actual;
    Vec x, k1, k2, k3, k4, k5;
    double y[6];
    double toff, tb, tend, h, hnext, preverr, rem;
    __device__ __forceinline__ void attach(double* col)
    {
        sh = col;
        if constexpr(KLDS) {
            x.p = col + F_X * STRIDE; k1.p = col + F_K1 * STRIDE; k2.p = col + F_K2 * STRIDE;
            k3.p = col + F_K3 * STRIDE; k4.p = col + F_K4 * STRIDE; k5.p = col + F_K5 * STRIDE;
        }
    }
    __device__ __forceinline__ double& f(int field) const { return sh[field * STRIDE]; }
    __device__ __forceinline__ double& k6(int i) const { return sh[(F_K6 + i) * STRIDE]; }
    __device__ __forceinline__ double& k7(int i) const { return sh[(F_K7 + i) * STRIDE]; }
    __device__ __forceinline__ double& k8(int i) const { return sh[(F_K8 + i) * STRIDE]; }
    __device__ __forceinline__ double& k9(int i) const { return sh[(F_K9 + i) * STRIDE]; }
    __device__ __forceinline__ double& k10(int i) const { return sh[(F_K10 + i) * STRIDE]; }
};

// the lane of the orbit kernel that also feeds target plug-ins
template<int STRIDE>
struct LaneTgtT : LaneT<STRIDE, false> {
    static constexpr bool HAS_TARGET = true;
};

template<class V>
__device__ __forceinline__ double energyOf(const V& xv, double pot, double Omega)
{
    double E = pot + 0.5 * (xv[3] * xv[3] + xv[4] * xv[4] + xv[5] * xv[5]);
    if(Omega != 0) E -= Omega * (xv[0] * xv[4] - xv[1] * xv[3]);
    return E;
}

// dense output of component i at offset `off` inside the accepted step [x -> y] of length h
// (src/math_ode.cpp:376-423); k11, k12, k13 live in the k2, k3, k4 registers
template<class Lane>
__device__ __forceinline__ double denseOut(const Lane& L, int i, double off)
{
    if(off == L.h) return L.y[i];
    if(off == 0) return L.x[i];
    const double h = L.h, k6 = L.k6(i), k7 = L.k7(i), k8 = L.k8(i), k9 = L.k9(i), k10 = L.k10(i);
    const double r1 = L.x[i], xd = L.y[i] - L.x[i], r2 = xd, xc = h * L.k1[i] - xd, r3 = xc,
        r4 = xd - h * L.k4[i] - xc,
        r5 = h * (d41 * L.k1[i] + d46 * k6 + d47 * k7 + d48 * k8 + d49 * k9 +
             d410 * k10 + d411 * L.k2[i] + d412 * L.k3[i] + d413 * L.k4[i]),
        r6 = h * (d51 * L.k1[i] + d56 * k6 + d57 * k7 + d58 * k8 + d59 * k9 +
             d510 * k10 + d511 * L.k2[i] + d512 * L.k3[i] + d513 * L.k4[i]),
        r7 = h * (d61 * L.k1[i] + d66 * k6 + d67 * k7 + d68 * k8 + d69 * k9 +
             d610 * k10 + d611 * L.k2[i] + d612 * L.k3[i] + d613 * L.k4[i]),
        r8 = h * (d71 * L.k1[i] + d76 * k6 + d77 * k7 + d78 * k8 + d79 * k9 +
             d710 * k10 + d711 * L.k2[i] + 0 * L.k3[i] + 0 * L.k4[i]);
    const double p = off / h, q = 1.0 - p;
    return r1 + p * (r2 + q * (r3 + p * (r4 + q * (r5 + p * (r6 + q * (r7 + p * r8))))));
}

// rotate an inertial-frame state into the rotating frame at angle ang (orbit.cpp:119-133)
template<class V>
__device__ __forceinline__ void toRotating(double Omega, double ang, const V& d, double* o)
{
    if(Omega != 0) {
        double ca = 1, sa = 0;
        sincosA(ang, sa, ca);
        o[0] = d[0] * ca + d[1] * sa; o[1] = d[1] * ca - d[0] * sa; o[2] = d[2];
        o[3] = d[3] * ca + d[4] * sa; o[4] = d[4] * ca - d[3] * sa; o[5] = d[5];
    } else
        for(int c = 0; c < 6; c++) o[c] = d[c];
}

__device__ __forceinline__ void storeRow(const AgxOrbitArgs& A, long row, double t, const double* p)
{
    if(A.outTime) A.outTime[row] = t;
    if(A.outTraj) for(int c = 0; c < 6; c++) A.outTraj[row * 6 + c] = p[c];
}

__device__ __forceinline__ long trajOffsetOf(const AgxOrbitArgs& A, int oi)
{
    return A.trajOffset ? A.trajOffset[oi] : 0;
}

// row capacity of an orbit's trajectory block: trajsize S >= 1 rows, or -S rows when S < 0
// (sampling interval 0: every step is recorded)
__device__ __forceinline__ long rowCapacity(int S) { return S < 0 ? -(long)S : (long)S; }

// output row of an every-step orbit's row `row` counted from the end of its block: a row of the
// overflow pool, taking a new chunk when the row starts one (the lane writes its rows in order,
// so the chunk directory entry is set before it is read); -1 when there is no pool or it is
// exhausted (the host then integrates the orbit again with a block of its size)
__device__ __forceinline__ long overflowRow(int* chunkDir, unsigned int* poolCounter, long poolBase, int poolChunk,
    int maxChunks, int poolChunks, int oi, long row)
{
    if(!poolCounter) return -1;
    const long k = row / poolChunk, r = row - k * poolChunk;
    if(k >= maxChunks) return -1;
    int* dir = chunkDir + (long)oi * maxChunks + k;
    int c;
    if(r == 0) {
        const unsigned int q = atomicAdd(poolCounter, 1u);
        c = q < (unsigned int)poolChunks ? (int)q : -1;
        *dir = c;
    } else
        c = *dir;
    return c < 0 ? -1 : poolBase + (long)c * poolChunk + r;
}

__device__ __forceinline__ long overflowRow(const AgxOrbitArgs& A, int oi, long row)
{
    return overflowRow(A.chunkDir, A.poolCounter, A.poolBase, A.poolChunk, A.maxChunks, A.poolChunks, oi, row);
}

// the same as a call (fewer registers held across it in kernels where that matters)
__device__ __noinline__ long overflowRowCall(int* chunkDir, unsigned int* poolCounter, long poolBase,
    int poolChunk, int maxChunks, int poolChunks, int oi, long row)
{
    return overflowRow(chunkDir, poolCounter, poolBase, poolChunk, maxChunks, poolChunks, oi, row);
}

// RuntimeTrajectory::processTimestep (orbit.cpp:15-47) after an accepted step [tb, tb+h]:
// rows at regular intervals (trajsize S >= 2), or with sampling interval 0 (S < 0) the start of
// the first step and then the end of every step (orbit.cpp:38-45), in a block of -S rows
// continued in the overflow pool.  Both share one loop over the rows a step produces.
template<class Lane>
__device__ __forceinline__ void sampleStep(const AgxOrbitArgs& A, Lane& L)
{
    if(L.S == 0 || L.S == 1) return;   // a single final point is written when the orbit ends
    const bool every = L.S < 0;
    const double tbegin = L.tb, timestep = L.h;
    double t0 = 0, dtS = 0, sign = 1;
    long iout, iend;
    if(every) {
        iend = (long)L.nsteps + 1;
        iout = L.nsteps == 0 ? 0 : iend;
        L.actual = 1;
    } else {
        t0 = L.f(F_T0);
        if(t0 != t0) { t0 = tbegin; L.f(F_T0) = t0; }
        dtS = L.f(F_DTS);
        sign = timestep >= 0 ? +1 : -1;
        const double tendS = tbegin + timestep;
        const double dtr = ROUNDOFF * fmax(fmax(fabs(tendS), fabs(tbegin)), fabs(t0));
        iout = (long)ceil(sign * (tbegin - t0) / dtS);
        iend = (long)((sign * (tendS - t0) + dtr) / dtS);
        L.actual = (int)min(iend + 1, (long)L.S + 1);
    }
    if(iout > iend) return;
    const long base = trajOffsetOf(A, L.oi), lim = rowCapacity(L.S);
    for(; iout <= iend; iout++) {
        double timeout, off;
        if(every) {
            const bool last = iout == iend;
            timeout = last ? tbegin + timestep : tbegin;
            off = last ? timestep : 0.0;
        } else {
            timeout = sign * dtS * iout + t0;
            off = fmin(fmax(sign * (timeout - tbegin), 0.0), sign * timestep) * sign;
        }
        double d[6], o[6];
        for(int c = 0; c < 6; c++) d[c] = denseOut(L, c, off);
        toRotating(A.Omega, A.Omega * (off + tbegin), d, o);
        if(iout >= 0 && iout < lim) storeRow(A, base + iout, timeout, o);
        else if(every) {
            const long r = overflowRow(A, L.oi, iout - lim);
            if(r >= 0) storeRow(A, r, timeout, o);
        }
    }
}

// RuntimeFncTarget::processTimestep (galaxymodel_target.h:141-154): the step's duration, and
// ten equally spaced dense-output points (rotating frame) each weighted by a tenth of the step
template<class Lane>
__device__ __forceinline__ void targetStep(const AgxOrbitArgs& A, Lane& L)
{
    const AgxTargetList& TL = *A.targets;
    const int NUM_SAMPLES_PER_STEP = 10;
    const double timestep = L.h, substep = timestep / NUM_SAMPLES_PER_STEP;
    for(int s = 0; s < NUM_SAMPLES_PER_STEP; s++) {
        const double off = substep * (s + 0.5);
        double d[6], p[6];
        for(int c = 0; c < 3; c++) d[c] = denseOut(L, c, off);
        if(TL.needVelocity)
            for(int c = 3; c < 6; c++) d[c] = denseOut(L, c, off);
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

template<class Lane>
__device__ __forceinline__ void finishOrbit(const AgxOrbitArgs& A, Lane& L, int status)
{
    const int oi = L.oi;
    const long base = trajOffsetOf(A, oi);
    if(L.S == 1 || A.outFinal) {
        double fin[6];
        const double lastTb = L.f(F_LASTTB), lastH = L.f(F_LASTH);
        if(L.actual > 0)   // last recorded point: getSol(h) of the last processed step
            toRotating(A.Omega, A.Omega * (lastH + lastTb), L.x, fin);
        else
            for(int c = 0; c < 6; c++) fin[c] = NAN;
        if(L.S == 1) storeRow(A, base, L.actual > 0 ? lastTb + lastH : NAN, fin);
        if(A.outFinal) for(int c = 0; c < 6; c++) A.outFinal[(long)oi * 6 + c] = fin[c];
    }
    for(long r = L.actual > 0 ? L.actual : 0; r < L.S; r++) {
        const double nanrow[6] = { NAN, NAN, NAN, NAN, NAN, NAN };
        storeRow(A, base + r, NAN, nanrow);
    }
    if(A.outNsteps) A.outNsteps[oi] = L.nsteps;
    if(A.outNattempts) A.outNattempts[oi] = L.natt;
    if(A.outNevals) A.outNevals[oi] = 3 + Lane::EVALS_PER_ATTEMPT * (long)L.natt;   // initialisation + attempts
    if(A.outMaxde) A.outMaxde[oi] = L.f(F_MAXDE);
    if(A.outStatus) A.outStatus[oi] = status;
    if(A.outTend) A.outTend[oi] = L.tb;
    L.phase = PH_IDLE;
}

// group: this lane belongs to a group of AGX_GROUP lanes integrating one orbit together (all
// lanes of the group call this at the same time, their states being identical)
template<class Lane>
__device__ __forceinline__ void fetchOrbit(const AgxOrbitArgs& A, Lane& L, bool group = false)
{
    long q;
    if(group) {
        const int lane = threadIdx.x & 63, lead = lane & ~(AGX_GROUP - 1);
        unsigned int qq = 0;
        if(lane == lead) qq = atomicAdd(A.groupCounter, 1u);
        qq = __shfl(qq, lead, 64);
        q = (long)qq < A.groupOrbits ? (long)qq : -1;
    } else {
        q = (long)atomicAdd(A.counter, 1u) + A.groupOrbits;
        if(q >= A.n) q = -1;
    }
    if(q < 0) {
        L.oi = -1;
        L.phase = PH_IDLE;
        L.y[0] = 1; L.y[1] = L.y[2] = L.y[3] = L.y[4] = L.y[5] = 0;   // placeholder point
        return;
    }
    const int oi = A.order ? A.order[q] : (int)q;
    L.oi = oi;
    L.tb = A.tstart ? A.tstart[oi] : 0.0;
    const double T = A.ttotal[oi];
    L.tend = L.tb + T;
    L.S = A.trajsize ? A.trajsize[oi] : 1;
    L.f(F_DTS) = L.S > 1 ? fabs(T) / (L.S - 1) : INFINITY;
    L.f(F_T0) = NAN;
    L.f(F_MAXDE) = 0;
    L.f(F_LASTTB) = L.tb;
    L.f(F_LASTH) = 0;
    L.actual = 0;
    L.hnext = 0;
    L.h = 0;
    L.nsteps = L.natt = 0;
    double ic[6];
    for(int c = 0; c < 6; c++) ic[c] = A.ic[(long)oi * 6 + c];
    if(A.Omega != 0) {   // orbit.cpp:104-113
        double ca = 1, sa = 0;
        sincosA(A.Omega * L.tb, sa, ca);
        L.x[0] = ic[0] * ca - ic[1] * sa; L.x[1] = ic[1] * ca + ic[0] * sa; L.x[2] = ic[2];
        L.x[3] = ic[3] * ca - ic[4] * sa; L.x[4] = ic[4] * ca + ic[3] * sa; L.x[5] = ic[5];
    } else
        for(int c = 0; c < 6; c++) L.x[c] = ic[c];
    for(int c = 0; c < 6; c++) L.y[c] = L.x[c];
    L.toff = 0;
    L.phase = PH_INIT_A;
}

template<class Lane>
__device__ __forceinline__ bool beginAttempt(const AgxOrbitArgs& A, Lane& L)
{
    if(L.h == 0 || !(L.h > -INFINITY && L.h < INFINITY)) { finishOrbit(A, L, 1); return false; }
    L.natt++;
    for(int i = 0; i < 6; i++) L.y[i] = L.x[i] + L.h * a21 * L.k1[i];
    L.toff = c2 * L.h;
    L.phase = PH_S2;
    return true;
}

template<class Lane>
__device__ __forceinline__ void beginStep(const AgxOrbitArgs& A, Lane& L)
{
    L.rem = L.tend - L.tb;
    if(L.rem == 0) { finishOrbit(A, L, 1); return; }
    const double sign = L.rem >= 0 ? +1 : -1;
    L.h = fmin(L.rem * sign, L.hnext) * sign;
    L.nbad = 0;
    L.preverr = INFINITY;
    beginAttempt(A, L);
}

// step accepted (doStep returned h): energy, sampling, advance (orbit.cpp:65-82)
template<class Lane>
__device__ __forceinline__ void acceptStep(const AgxOrbitArgs& A, Lane& L, double pot)
{
    const double de = fabs(energyOf(L.y, pot, A.Omega) / L.f(F_E0) - 1);
    const double mde = L.f(F_MAXDE);
    if(de > mde || de != de) L.f(F_MAXDE) = de;
    sampleStep(A, L);
    if constexpr(Lane::HAS_TARGET) targetStep(A, L);
    L.f(F_LASTTB) = L.tb;
    L.f(F_LASTH) = L.h;
    if(L.S == 1) L.actual = 1;
    for(int i = 0; i < 6; i++) { L.x[i] = L.y[i]; L.k1[i] = L.k4[i]; }
    L.tb = L.h == L.rem ? L.tend : L.tb + L.h;
    L.nsteps++;
    if(L.nsteps >= A.maxNumSteps) { finishOrbit(A, L, 2); return; }
    if(L.tb == L.tend) { finishOrbit(A, L, 0); return; }
    beginStep(A, L);
}

// Consume one right-hand-side evaluation f (and the potential pot at the same point) according
// to the lane's phase and set up the next evaluation point: the DOP853 stage sequence, step
// control and bookkeeping of src/math_ode.cpp:35-373 and src/orbit.cpp:52-84.
template<class Lane>
__device__ __forceinline__ void advancePhase(const AgxOrbitArgs& A, Lane& L, const double* f, double pot,
    double accRel)
{
    const double h = L.h;
    switch(L.phase) {
    case PH_INIT_A: {   // DOP853::init + first evaluation of initTimeStep (math_ode.cpp:35-105)
        for(int i = 0; i < 6; i++) L.k1[i] = f[i];
        const double Ekin = 0.5 * (L.x[3] * L.x[3] + L.x[4] * L.x[4] + L.x[5] * L.x[5]);
        L.f(F_ACC) = fmin(1., fabs(pot + Ekin) / fmax(fabs(pot), Ekin));
        L.f(F_E0) = energyOf(L.x, pot, A.Omega);
        double nx = 0, nd = 0;
        for(int i = 0; i < 6; i++) { nx += L.x[i] * L.x[i]; nd += L.k1[i] * L.k1[i]; }
        const double h1 = fmax(1e-12, fmin(1e12, sqrt(nx / nd) * 0.01));
        L.f(F_H1) = h1;
        for(int i = 0; i < 6; i++) L.y[i] = L.x[i] + h1 * L.k1[i];
        L.toff = h1;
        L.phase = PH_INIT_B;
        break;
    }
    case PH_INIT_B: {
        const double h1 = L.f(F_H1);
        for(int i = 0; i < 6; i++) { L.k2[i] = f[i]; L.y[i] = L.x[i] + 0.5 * h1 * (L.k1[i] + L.k2[i]); }
        L.toff = h1;
        L.phase = PH_INIT_C;
        break;
    }
    case PH_INIT_C: {
        const double h1 = L.f(F_H1);
        for(int i = 0; i < 6; i++) L.k3[i] = f[i];
        double h2 = INFINITY;
        for(int i = 0; i < 6; i++) {
            const double d0 = fmax(0., fmax(fabs(L.x[i]), fabs(L.y[i])));
            const double d1 = fmax(fabs(L.k1[i]), fabs(L.k3[i]));
            const double d2 = fabs((3 * L.k2[i] - L.k1[i] - 2 * L.k3[i]) / h1);
            const double d3 = fabs(6 * (L.k3[i] - L.k2[i]) / (h1 * h1));
            const double hh = (d0 * d2 + d1 * d1) / (d1 * d3 + d2 * d2);
            h2 = fmin(h2, hh);
        }
        L.hnext = fmax(1e-12, fmin(1e12, pow(accRel, 1. / 8) * sqrt(h2))) * L.f(F_ACC);
        // run() (orbit.cpp:52-62): a zero-length or non-finite request records the start
        if(A.ttotal[L.oi] == 0 || !(L.tend > -INFINITY && L.tend < INFINITY)) {
            L.f(F_LASTTB) = L.tb;
            L.f(F_LASTH) = 0;
            if(L.S >= 2 || L.S < 0) {   // sampling interval 0: the start is recorded twice
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
            beginStep(A, L);
        break;
    }
    case PH_S2:
        for(int i = 0; i < 6; i++) { L.k2[i] = f[i]; L.y[i] = L.x[i] + h * (a31 * L.k1[i] + a32 * L.k2[i]); }
        L.toff = c3 * h; L.phase = PH_S3; break;
    case PH_S3:
        for(int i = 0; i < 6; i++) { L.k3[i] = f[i]; L.y[i] = L.x[i] + h * (a41 * L.k1[i] + a43 * L.k3[i]); }
        L.toff = c4 * h; L.phase = PH_S4; break;
    case PH_S4:
        for(int i = 0; i < 6; i++) {
            L.k4[i] = f[i];
            L.y[i] = L.x[i] + h * (a51 * L.k1[i] + a53 * L.k3[i] + a54 * L.k4[i]);
        }
        L.toff = c5 * h; L.phase = PH_S5; break;
    case PH_S5:
        for(int i = 0; i < 6; i++) {
            L.k5[i] = f[i];
            L.y[i] = L.x[i] + h * (a61 * L.k1[i] + a64 * L.k4[i] + a65 * L.k5[i]);
        }
        L.toff = c6 * h; L.phase = PH_S6; break;
    case PH_S6:
        for(int i = 0; i < 6; i++) {
            L.k6(i) = f[i];
            L.y[i] = L.x[i] + h * (a71 * L.k1[i] + a74 * L.k4[i] + a75 * L.k5[i] + a76 * f[i]);
        }
        L.toff = c7 * h; L.phase = PH_S7; break;
    case PH_S7:
        for(int i = 0; i < 6; i++) {
            L.k7(i) = f[i];
            L.y[i] = L.x[i] + h * (a81 * L.k1[i] + a84 * L.k4[i] + a85 * L.k5[i] + a86 * L.k6(i) +
                a87 * f[i]);
        }
        L.toff = c8 * h; L.phase = PH_S8; break;
    case PH_S8:
        for(int i = 0; i < 6; i++) {
            L.k8(i) = f[i];
            L.y[i] = L.x[i] + h * (a91 * L.k1[i] + a94 * L.k4[i] + a95 * L.k5[i] + a96 * L.k6(i) +
                a97 * L.k7(i) + a98 * f[i]);
        }
        L.toff = c9 * h; L.phase = PH_S9; break;
    case PH_S9:
        for(int i = 0; i < 6; i++) {
            L.k9(i) = f[i];
            L.y[i] = L.x[i] + h * (a101 * L.k1[i] + a104 * L.k4[i] + a105 * L.k5[i] + a106 * L.k6(i) +
                a107 * L.k7(i) + a108 * L.k8(i) + a109 * f[i]);
        }
        L.toff = c10 * h; L.phase = PH_S10; break;
    case PH_S10:
        for(int i = 0; i < 6; i++) {
            L.k10(i) = f[i];
            L.y[i] = L.x[i] + h * (a111 * L.k1[i] + a114 * L.k4[i] + a115 * L.k5[i] + a116 * L.k6(i) +
                a117 * L.k7(i) + a118 * L.k8(i) + a119 * L.k9(i) + a1110 * f[i]);
        }
        L.toff = c11 * h; L.phase = PH_S11; break;
    case PH_S11:   // k11 -> k2 registers
        for(int i = 0; i < 6; i++) {
            L.k2[i] = f[i];
            L.y[i] = L.x[i] + h * (a121 * L.k1[i] + a124 * L.k4[i] + a125 * L.k5[i] + a126 * L.k6(i) +
                a127 * L.k7(i) + a128 * L.k8(i) + a129 * L.k9(i) + a1210 * L.k10(i) + a1211 * f[i]);
        }
        L.toff = h; L.phase = PH_S12; break;
    case PH_S12:   // k12 -> k3 registers, b-sum -> k5 registers
        for(int i = 0; i < 6; i++) {
            L.k3[i] = f[i];
            L.k5[i] = b1 * L.k1[i] + b6 * L.k6(i) + b7 * L.k7(i) + b8 * L.k8(i) + b9 * L.k9(i) +
                b10 * L.k10(i) + b11 * L.k2[i] + b12 * f[i];
            L.y[i] = L.x[i] + h * L.k5[i];
        }
        L.toff = h; L.phase = PH_FSAL; break;
    case PH_FSAL: {   // k13 -> k4 registers; error control (math_ode.cpp:323-373)
        for(int i = 0; i < 6; i++) L.k4[i] = f[i];
        const double Ekin = 0.5 * (L.y[3] * L.y[3] + L.y[4] * L.y[4] + L.y[5] * L.y[5]);
        const double accFac = fmin(1., fabs(pot + Ekin) / fmax(fabs(pot), Ekin));
        double err5 = 0.0, err3 = 0.0;
        for(int i = 0; i < 6; i++) {
            const double sk = (0. + accRel * fmax(fabs(L.x[i]), fabs(L.y[i]))) * accFac;
            if(sk == 0) continue;
            const double e3 = (L.k5[i] - bhh1 * L.k1[i] - bhh2 * L.k9(i) - bhh3 * L.k3[i]) / sk;
            const double e5 = (er1 * L.k1[i] + er6 * L.k6(i) + er7 * L.k7(i) + er8 * L.k8(i) +
                er9 * L.k9(i) + er10 * L.k10(i) + er11 * L.k2[i] + er12 * L.k3[i]) / sk;
            err3 += e3 * e3;
            err5 += e5 * e5;
        }
        const double den = sqrt(6 * (err5 + 0.01 * err3));
        const double err = den == 0 ? 0 : err5 * fabs(h) / den;
        if(!(err > -INFINITY && err < INFINITY)) { finishOrbit(A, L, 1); break; }
        double fac = sqrt(sqrt(sqrt(err)));
        if(err <= 1) {
            L.hnext = fabs(h) * fmin(finc, safe / fac);
            acceptStep(A, L, pot);
        } else {
            bool forced = false;
            if(err > 0.5 * L.preverr) {
                L.nbad++;
                if(L.nbad >= 2) fac = 1 / fdec;
                if(L.nbad >= 4) { L.hnext = fabs(h); forced = true; }
            } else
                L.nbad = 0;
            if(forced)
                acceptStep(A, L, pot);
            else {
                L.preverr = err;
                L.h = h * fmax(fdec, safe / fac);
                beginAttempt(A, L);
            }
        }
        break;
    }
    default: break;
    }
}

}  // namespace agxd
