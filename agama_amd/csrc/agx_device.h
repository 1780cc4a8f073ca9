// Device functions for Multipole evaluation on gfx950 (FP64).
//
// Evaluation follows the reference routines (paths relative to its source tree):
//   sincosA / atan2A        src/math_core.cpp:171-248
//   binSearchD              src/math_core.cpp:250-280
//   quintic Hermite passes  src/math_spline.cpp:653-718 (evalQuinticSplines)
//   interior                src/potential_multipole.cpp:1969-2080 (MultipoleInterp2d::evalCyl)
//                           with src/math_spline.cpp:2535-2597 (QuinticSpline2d::evalDeriv)
//   power-law branches      src/potential_multipole.cpp:1600-1658, 722-848
//   Fourier synthesis       src/potential_multipole.cpp:666-717, src/math_sphharm.cpp:134-152
//   Car<->Cyl               src/coord.cpp:94-125, 607-614
// Design points of the device version:
//   - the cell search and the quintic weights depend only on the point, so they are computed
//     once per evaluation and shared by all azimuthal harmonics (the reference recomputes the
//     same values for each harmonic; the arithmetic is identical);
//   - harmonics are streamed in ascending m with the cos/sin(m phi) recurrence advanced in the
//     same loop, so no per-harmonic arrays live in registers (mmin = 0, every symmetry with
//     y-reflection); expansions with sine terms (mmin < 0) take a generic path with small
//     private arrays that keeps the reference's summation order.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "agx_tables.h"

namespace agxd {

// Table pointers are read from the header in device memory, so the compiler cannot infer their
// address space and would emit flat (generic) accesses; these qualifiers make global-memory
// and LDS accesses explicit.
#define AGX_GLOBAL __attribute__((address_space(1)))
#define AGX_LDS __attribute__((address_space(3)))
template<typename T>
__device__ __forceinline__ const T AGX_GLOBAL* gmem(const T* p) { return (const T AGX_GLOBAL*)p; }
// native 16- and 8-byte vectors (the HIP vector classes cannot be copied across address spaces)
typedef double vd2 __attribute__((ext_vector_type(2)));
typedef float vf2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double sqd(double x) { return x * x; }

__device__ __forceinline__ void sincosA(double x, double& s, double& c)
{
    double y = fabs(x);
    long quad = (long)(4 / M_PI * y);
    quad = (quad + 1) >> 1;
    y -= M_PI / 2 * quad;
    int q13 = (int)(quad & 1), q02 = 1 - q13;
    int signc = 1 - (int)(quad & 2), signs = signbit(x) ? -signc : signc;
    double y2 = y * y;
    double sy = y + y * (((((+1.5896230157654657e-10 * y2 - 2.5050747762857807e-8) * y2
        + 2.7557313621385725e-6) * y2 - 1.9841269829589539e-4) * y2 + 8.3333333333221186e-3) * y2
        - 0.1666666666666663073) * y2;
    double cy = 1.0 + ((((((-1.1358536521387682e-11 * y2 + 2.0875700841974730e-9) * y2
        - 2.7557314179296740e-7) * y2 + 2.4801587288851704e-5) * y2 - 0.0013888888888873056) * y2
        + 0.0416666666666665950) * y2 - 0.5) * y2;
    s = (q02 * sy + q13 * cy) * signs;
    c = (q02 * cy - q13 * sy) * signc;
}

__device__ __forceinline__ double atan1A(double x)
{
    double x2 = x * x,
    num = -403.11710541978266 + x2 * (-902.59989673314130 + x2 * (-707.35355100686270 +
          x2 * (-230.47304101738868 + x2 * (-28.595211396994470 + x2 * -0.9023747369388881)))),
    den = 1209.3513162593547 + x2 * (3433.4104799543807 + x2 * (3663.8135197609940 +
          x2 * (1821.3627056982116 + x2 * (423.04454079463840 + x2 * (39.917377889601520 + x2)))));
    return x * x2 * (num / den) + x;
}

__device__ __forceinline__ double atan2A(double y, double x)
{
    if(y == 0)
        return copysign(M_PI * (signbit(x) ? 1.0 : 0.0), y);
    bool negx = x < 0;
    double signy = y >= 0 ? 1 : -1, absy = y * signy, absx = negx ? -x : x;
    bool ybigger = absy > absx;
    double res = atan1A(ybigger ? -x / y : y / x);
    if(ybigger)   res += 0.5 * M_PI * signy;
    else if(negx) res += M_PI * signy;
    return res;
}

// arr: a global-memory pointer, or a generic one (e.g. to an LDS copy of the grid)
template<class GP>
__device__ __forceinline__ int binSearchD(double x, GP arr, int size)
{
    if(size < 1 || !(x >= arr[0])) return -1;
    if(x > arr[size - 1] || size < 2) return size - 1;
    int index = (int)((x - arr[0]) / (arr[size - 1] - arr[0]) * (size - 1));
    int indhi = size - 1;
    if(index == size - 1) return size - 2;
    if(x >= arr[index]) {
        if(x < arr[index + 1]) return index;
    } else {
        indhi = index;
        index = 0;
    }
    while(indhi > index + 1) {
        int i = (indhi + index) / 2;
        if(arr[i] > x) indhi = i;
        else index = i;
    }
    return index;
}

// weights of the quintic Hermite interpolant on [xl, xh] at x (value and first derivative),
// in the arithmetic type R of the evaluation (double, or float for the precision-split form)
template<typename R>
struct QuinticWT {
    bool atEnd;                 // x == xh: endpoint values are returned exactly
    R dx, dx2, h, P, Q, Rr, Pp, Qp, Rp;
    // the offsets x - xl and xh - xl are formed in double before rounding to R, so that the
    // float form's weights vary smoothly with x inside the cell
    __device__ __forceinline__ void init(double x, double xl, double xh)
    {
        atEnd = x == xh;
        dx = (R)(x - xl);
        dx2 = R(.5) * dx * dx;
        h = (R)(xh - xl);
        const double hiD = 1 / (xh - xl);
        const R hi = (R)hiD, h2 = h * h, t = (R)((x - xl) * hiD), t2 = t * t, t3 = t * t2,
            t1t = t * (R(1) - t), Px = R(30) * t1t * hi;
        P  = t3 * (R(10) - t * (R(15) - R(6) * t));
        Q  = t3 * (R(1) - t * R(0.5)) * h;
        Rr = t3 * (R(1) - t * (R(1.25) - R(0.5) * t)) * h2;
        Pp = Px * t1t;
        Qp = t2 * (R(3) - t * R(2));
        Rp = t2 * (R(3) - t * (R(5) - R(2.5) * t)) * h;
    }
    // value and derivative of one function given f, f', f'' at both ends
    __device__ __forceinline__ void eval2(R fl, R fh, R f1l, R f1h, R f2l, R f2h, R& f, R& df) const
    {
        if(atEnd) { f = fh; df = f1h; return; }
        const R fd = fh - fl - R(0.5) * h * (f1h + f1l),
            f1d = f1h - f1l - R(0.5) * h * (f2h + f2l), f2d = f2h - f2l;
        f  = fl + P * fd + dx * f1l + Q * f1d + dx2 * f2l + Rr * f2d;
        df = Pp * fd + f1l + Qp * f1d + dx * f2l + Rp * f2d;
    }
    __device__ __forceinline__ R eval1(R fl, R fh, R f1l, R f1h, R f2l, R f2h) const
    {
        if(atEnd) return fh;
        const R fd = fh - fl - R(0.5) * h * (f1h + f1l),
            f1d = f1h - f1l - R(0.5) * h * (f2h + f2l), f2d = f2h - f2l;
        return fl + P * fd + dx * f1l + Q * f1d + dx2 * f2l + Rr * f2d;
    }
};
typedef QuinticWT<double> QuinticW;

// ---- node fetchers: how a lane obtains the 9 values of one node record (f, f_x, f_y, f_xx,
// f_xy, f_yy, f_xxy, f_xyy, f_xxyy).  `rec` is the record's index in units of
// AGX_NODE_STRIDE doubles.  Every record sits at an 80-byte (16-byte aligned) boundary.

// one 8-byte load per value and lane
struct FetchScalar {
    typedef double real;
    const double AGX_GLOBAL* nodes;
    static constexpr bool uniform = false;
    __device__ __forceinline__ void record(int rec, double* v) const
    {
        const double AGX_GLOBAL* p = nodes + (size_t)rec * AGX_NODE_STRIDE;
        for(int q = 0; q < 9; q++) v[q] = p[q];
    }
};

// one 16-byte load per pair of values and lane
struct FetchVec2 {
    typedef double real;
    const vd2 AGX_GLOBAL* nodes;
    static constexpr bool uniform = false;
    __device__ __forceinline__ void record(int rec, double* v) const
    {
        const vd2 AGX_GLOBAL* p = nodes + (size_t)rec * (AGX_NODE_STRIDE / 2);
        for(int q = 0; q < 5; q++) {
            const vd2 t = p[q];
            v[2 * q] = t.x;
            if(2 * q + 1 < 9) v[2 * q + 1] = t.y;
        }
    }
};

// FP32 tables (precision-split form): the spline is evaluated in float arithmetic
struct FetchFloat {
    typedef float real;
    const float AGX_GLOBAL* nodes;
    static constexpr bool uniform = false;
    __device__ __forceinline__ void record(int rec, float* v) const
    {
        const vf2 AGX_GLOBAL* p2 = (const vf2 AGX_GLOBAL*)(nodes + (size_t)rec * AGX_NODE_STRIDE);
        for(int q = 0; q < 5; q++) {
            const vf2 t = p2[q];
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    }
};

// Cooperative gather of a whole wavefront: the 64 lanes' records (64 x 80 bytes each) are
// read with coalesced 16-byte loads - lane l fetches chunk k = l + 64 t of the wave's list -
// and redistributed through LDS, so that one load instruction touches about 13 records'
// cache lines instead of 64.  Divergent per-lane gathers are limited by the address unit at
// about one cache line per cycle per CU.  The four corner records of a harmonic are fetched
// as one batch (20 loads in flight).  All 64 lanes must call records4() together
// (wave-uniform control flow); `lds` is this wavefront's buffer of 4 * 64 * 5 vd2.
struct FetchCoop {
    typedef double real;
    const vd2 AGX_GLOBAL* nodes;
    vd2 AGX_LDS* lds;
    int lane;
    static constexpr bool uniform = true;
    static constexpr int LDS_DOUBLE2_PER_WAVE = 4 * 64 * 5;
    __device__ __forceinline__ void records4(int r0, int r1, int r2, int r3, double (&v)[4][10]) const
    {
        const int b0 = r0 * 5, b1 = r1 * 5, b2 = r2 * 5, b3 = r3 * 5;
        vd2 t0[5], t1[5], t2[5], t3[5];
#pragma unroll
        for(int t = 0; t < 5; t++) {
            const int k = lane + 64 * t, r = k / 5, c = k - 5 * r;
            t0[t] = nodes[__shfl(b0, r, 64) + c];
            t1[t] = nodes[__shfl(b1, r, 64) + c];
            t2[t] = nodes[__shfl(b2, r, 64) + c];
            t3[t] = nodes[__shfl(b3, r, 64) + c];
        }
#pragma unroll
        for(int t = 0; t < 5; t++) {
            const int k = lane + 64 * t;
            lds[k] = t0[t]; lds[320 + k] = t1[t]; lds[640 + k] = t2[t]; lds[960 + k] = t3[t];
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for(int j = 0; j < 4; j++)
#pragma unroll
            for(int q = 0; q < 5; q++) {
                const vd2 t = lds[j * 320 + lane * 5 + q];
                v[j][2 * q] = t.x;
                v[j][2 * q + 1] = t.y;
            }
        __builtin_amdgcn_wave_barrier();
    }
};

// per-lane fetchers: the four records are independent loads scheduled together
template<class F>
__device__ __forceinline__ void fetch4(const F& f, int r0, int r1, int r2, int r3,
    typename F::real (&v)[4][10])
{
    f.record(r0, v[0]); f.record(r1, v[1]); f.record(r2, v[2]); f.record(r3, v[3]);
}
__device__ __forceinline__ void fetch4(const FetchCoop& f, int r0, int r1, int r2, int r3,
    double (&v)[4][10])
{
    f.records4(r0, r1, r2, r3, v);
}

// 2-D quintic spline of one harmonic slot at a located cell: value and (ln r, tau) derivatives.
// recLL is the record of corner (xi, yi); (xi, yi+1) follows at +ns records and (xi+1, yi) at
// +nt*ns records.  The reference order of the nine node arrays (QuinticSpline2d::evalDeriv,
// math_spline.cpp:2572-2597) maps to the record as q = 0..8 above.
// value and (x, y) derivatives of one harmonic slot from its four corner records v
// (ll, lu, ul, uu): QuinticSpline2d::evalDeriv (src/math_spline.cpp:2535-2597)
template<typename R>
__device__ __forceinline__ void splineSlotFrom(const R (&v)[4][10], int offsetCorner,
    const QuinticWT<R>& wy, const QuinticWT<R>& wx, R& z, R& zx, R& zy)
{
    const R* ll = v[0]; const R* lu = v[1]; const R* ul = v[2]; const R* uu = v[3];
    const R off = offsetCorner == 0 ? ll[0] : offsetCorner == 1 ? lu[0] :
                  offsetCorner == 2 ? ul[0] : uu[0];
    // y-pass (tau) for the six functions {f, fx, fxx} at xlow/xupp, with f, f_y, f_yy ends
    R F0, F1, F2, F3, F4, F5, dF0, dF1, dF2, dF3, dF4, dF5;
    wy.eval2(ll[0] - off, lu[0] - off, ll[2], lu[2], ll[5], lu[5], F0, dF0);
    wy.eval2(ul[0] - off, uu[0] - off, ul[2], uu[2], ul[5], uu[5], F1, dF1);
    wy.eval2(ll[1], lu[1], ll[4], lu[4], ll[7], lu[7], F2, dF2);
    wy.eval2(ul[1], uu[1], ul[4], uu[4], ul[7], uu[7], F3, dF3);
    wy.eval2(ll[3], lu[3], ll[6], lu[6], ll[8], lu[8], F4, dF4);
    wy.eval2(ul[3], uu[3], ul[6], uu[6], ul[8], uu[8], F5, dF5);
    // x-pass (ln r)
    wx.eval2(F0, F1, F2, F3, F4, F5, z, zx);
    zy = wx.eval1(dF0, dF1, dF2, dF3, dF4, dF5);
    z += off;
}

// the four corner records of slot record recLL (cell corner (i, j)): (i,j), (i,j+1), (i+1,j), (i+1,j+1)
template<class F, typename R = typename F::real>
__device__ __forceinline__ void fetchSlot(const F& fetch, int recLL, int ns, int nt, R (&v)[4][10])
{
    fetch4(fetch, recLL, recLL + ns, recLL + nt * ns, recLL + nt * ns + ns, v);
}

template<class F, typename R = typename F::real>
__device__ __forceinline__ void splineSlot(const F& fetch, int recLL, int ns, int nt, int offsetCorner,
    const QuinticWT<R>& wy, const QuinticWT<R>& wx, R& z, R& zx, R& zy)
{
    R v[4][10];
    fetchSlot(fetch, recLL, ns, nt, v);
    splineSlotFrom(v, offsetCorner, wy, wx, z, zx, zy);
}

// cos/sin(k phi) for k = 1..m by the recurrence of trigMultiAngle (math_sphharm.cpp:134-152),
// started from (alpha, beta) = (2 sin^2(phi/2), sin(phi)); `c, s` hold the values for k and
// are advanced in place up to m
template<typename R>
struct TrigRecT {
    R alpha, beta, c, s;
    int k;
    __device__ __forceinline__ void init(double phi)
    {
        double sphi, cphi, a, b;
        sincosA(phi, sphi, cphi);
        sincosA(phi / 2, a, b);
        a *= a * 2;
        alpha = (R)a;
        beta = (R)sphi;
        c = 1; s = 0; k = 0;
    }
    __device__ __forceinline__ void restart() { c = 1; s = 0; k = 0; }
    __device__ __forceinline__ void advanceTo(int m)
    {
        while(k < m) {
            const R cn = c - (alpha * c + beta * s);
            const R sn = s - (alpha * s - beta * c);
            c = cn; s = sn; k++;
        }
    }
};

// interior evaluation in cylindrical coordinates: potential and (R, z, phi) gradient.
// Control flow is uniform across lanes (all harmonics are always evaluated; points outside
// the grid read cell (0,0) and return NaN, as the reference's spline does), which the
// cooperative fetcher requires.
// PF: request the next harmonic's corner records before evaluating the current one (software
// pipelining: more registers, one L2 round trip fewer on the critical path per harmonic)
template<class F, class GP, bool PF = false>
__device__ __forceinline__ void interiorEval(const AgxMultipoleDev& P, const F& fetch,
    double R, double z, double phi, double& pot, double& gR, double& gz, double& gphi,
    GP gridR, GP gridT)
{
    const double r = sqrt(R * R + z * z), logr = log(r), rplusRinv = 1. / (r + fabs(R)),
        tau = R == 0 ? (z > 0 ? 1. : z < 0 ? -1. : 0.) : z * rplusRinv;
    const int xi = binSearchD(logr, gridR, P.nr), yi = binSearchD(tau, gridT, P.nt);
    const bool inGrid = xi >= 0 && xi < P.nr - 1 && yi >= 0 && yi < P.nt - 1;
    const int xc = inGrid ? xi : 0, yc = inGrid ? yi : 0;
    typedef typename F::real Rt;
    const double xlow = gridR[xc], xupp = gridR[xc + 1], ylow = gridT[yc], yupp = gridT[yc + 1];
    const int offsetCorner = logr == xupp ? (tau == yupp ? 3 : 2) : (tau == yupp ? 1 : 0);
    QuinticWT<Rt> wy, wx;
    wy.init(tau, ylow, yupp);
    wx.init(logr, xlow, xupp);
    const int ns = P.ns, nt = P.nt, rec0 = (xc * nt + yc) * ns;

    // m = 0 first (it scales the other harmonics), then the others in ascending m as in the
    // reference's Fourier synthesis: negative m (sine terms) restart the trig recurrence,
    // positive m advance one running recurrence
    const int zs = P.zeroSlot;
    Rt P0, d0r, d0t;
    Rt va[4][10], vb[4][10];
    const int s0 = zs == 0 ? 1 : 0;
    if constexpr(PF) {
        fetchSlot(fetch, rec0 + zs, ns, nt, va);
        if(s0 < ns) fetchSlot(fetch, rec0 + s0, ns, nt, vb);
        splineSlotFrom(va, offsetCorner, wy, wx, P0, d0r, d0t);
    } else
        splineSlot(fetch, rec0 + zs, ns, nt, offsetCorner, wy, wx, P0, d0r, d0t);
    if(P.logScaling) {
        const Rt expX = exp(P0), val = Rt(1) / ((Rt)P.invPhi0 - expX), dPhidX = val * val * expX;
        P0 = val;
        d0r *= dPhidX;
        d0t *= dPhidX;
    }
    Rt trPot = P0, trDr = d0r, trDth = d0t, trDphi = 0;
    if(ns > 1) {
        TrigRecT<Rt> pos, neg;
        pos.init(phi);
        neg.alpha = pos.alpha; neg.beta = pos.beta;
        // harmonic slot s (not the m = 0 slot) from its corner records v
        auto addSlot = [&](int s, const Rt (&v)[4][10]) {
            const int m = P.slotM[s];
            Rt trig, dtrig;
            if(m > 0) {
                pos.advanceTo(m);
                trig = pos.c; dtrig = Rt(-m) * pos.s;
            } else {
                neg.restart();
                neg.advanceTo(-m);
                trig = neg.s; dtrig = Rt(-m) * neg.c;
            }
            Rt Pm, dmr, dmt;
            splineSlotFrom(v, offsetCorner, wy, wx, Pm, dmr, dmt);
            if(P.logScaling) {
                dmr = dmr * P0 + Pm * d0r;
                dmt = dmt * P0 + Pm * d0t;
                Pm *= P0;
            }
            trPot  += Pm * trig;
            trDr   += dmr * trig;
            trDth  += dmt * trig;
            trDphi += Pm * dtrig;
        };
        if constexpr(PF) {
        // vb holds slot s on entry; the two buffers alternate
        for(int s = s0; s < ns; ) {
            int s2 = s + 1;
            if(s2 == zs) s2++;
            if(s2 < ns) fetchSlot(fetch, rec0 + s2, ns, nt, va);
            addSlot(s, vb);
            if(s2 >= ns) break;
            int s3 = s2 + 1;
            if(s3 == zs) s3++;
            if(s3 < ns) fetchSlot(fetch, rec0 + s3, ns, nt, vb);
            addSlot(s2, va);
            s = s3;
        }
        } else {
        for(int s = 0; s < ns; s++) {
            if(s == zs) continue;
            fetchSlot(fetch, rec0 + s, ns, nt, va);
            addSlot(s, va);
        }
        }
    }
    const double rinv = 1 / r, r2inv = rinv * rinv;
    const double drdR = R * r2inv, drdz = z * r2inv, dthdR = -tau * rinv, dthdz = rinv - rplusRinv;
    pot = inGrid ? (double)trPot : NAN;
    gR = inGrid ? (double)trDr * drdR + (double)trDth * dthdR : NAN;
    gz = inGrid ? (double)trDr * drdz + (double)trDth * dthdz : NAN;
    gphi = inGrid ? (double)trDphi : NAN;
}

// ---- per-cell polynomial evaluation (FP64 kernels) ----
// The tensor-product quintic Hermite interpolant of one cell and harmonic is a bi-quintic
// polynomial; its 36 coefficients in the cell's local coordinates (u, v) = (ln r - lnr_i,
// tau - tau_j) are computed on the host from the corner records (buildCellPolys), so a lane
// evaluates p, dp/du, dp/dv by Horner's rule: 6 rows in v (value and derivative, 9 FMAs each),
// then 14 FMAs in u.  The same function as the two-pass Hermite form of
// QuinticSpline2d::evalDeriv (src/math_spline.cpp:2535-2597), in about half the arithmetic and
// without the per-point Hermite weights; all harmonics of a cell are one contiguous run read
// with buffer loads at a single 32-bit offset.
constexpr int AGX_CELL_SLOT_BYTES = AGX_CELL_COEFS * 8;
constexpr int AGX_CELL_CHUNKS = AGX_CELL_SLOT_BYTES / 16;   // 16-byte loads per slot

// Horner evaluation of one slot from its 36 coefficients (row a: c[6a .. 6a+5]).  Each
// coefficient sits in a register of its own (8-byte loads): with halves of 16-byte loads the
// compiler formed most multiply-adds as a copy of the addend plus v_fmac_f64, since the tuple's
// other half was still live.
__device__ __forceinline__ void cellHorner(const double (&c)[AGX_CELL_COEFS], double u, double v,
    double& p, double& pu, double& pv)
{
    double P = 0, Pu = 0, Pv = 0;
#pragma unroll
    for(int a = 5; a >= 0; a--) {
        const double* r = c + 6 * a;
        // row a: A = sum_b c[a][b] v^b and B = dA/dv
        double A = r[5] * v + r[4], B = r[5];
        B = B * v + A; A = A * v + r[3];
        B = B * v + A; A = A * v + r[2];
        B = B * v + A; A = A * v + r[1];
        B = B * v + A; A = A * v + r[0];
        if(a == 5) { P = A; Pv = B; }
        else { Pu = Pu * u + P; P = P * u + A; Pv = Pv * u + B; }
    }
    p = P; pu = Pu; pv = Pv;
}

// every lane reads its own cell
struct FetchCell {
    typedef double real;
    static constexpr bool uniform = false;
    __amdgpu_buffer_rsrc_t rsrc;
    __device__ __forceinline__ FetchCell(const double* cells, unsigned int bytes)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc((void*)cells, 0, (int)bytes, 0x00020000)) {}
    // 8 bytes at byte offset off
    __device__ __forceinline__ double ld(int off) const
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0));
    }
    // the coefficients of slot s of the cell at byte offset cellOff
    __device__ __forceinline__ void load(int cellOff, int s, double (&c)[AGX_CELL_COEFS]) const
    {
        const int off = cellOff + s * AGX_CELL_SLOT_BYTES;
#pragma unroll
        for(int t = 0; t < AGX_CELL_COEFS; t++) c[t] = ld(off + 8 * t);
    }
    // value and (u, v) derivatives of slot s of the cell at byte offset cellOff
    __device__ __forceinline__ void slot(int cellOff, int s, double u, double v, double& p, double& pu,
        double& pv) const
    {
        double c[AGX_CELL_COEFS];
        load(cellOff, s, c);
        cellHorner(c, u, v, p, pu, pv);
    }
};

template<class F> struct IsCellFetch : std::false_type {};
template<> struct IsCellFetch<FetchCell> : std::true_type {};

// interiorEval on the per-cell polynomial tables (same control flow and Fourier synthesis).
// With the per-lane fetcher the next harmonic's coefficients are requested before the current
// one is evaluated (two register buffers), so the memory round trips of the harmonics overlap.
template<class F, class GP>
__device__ __forceinline__ void interiorEvalCell(const AgxMultipoleDev& P, const F& fetch,
    double R, double z, double phi, double& pot, double& gR, double& gz, double& gphi,
    GP gridR, GP gridT)
{
    const double r = sqrt(R * R + z * z), logr = log(r), rplusRinv = 1. / (r + fabs(R)),
        tau = R == 0 ? (z > 0 ? 1. : z < 0 ? -1. : 0.) : z * rplusRinv;
    const int xi = binSearchD(logr, gridR, P.nr), yi = binSearchD(tau, gridT, P.nt);
    const bool inGrid = xi >= 0 && xi < P.nr - 1 && yi >= 0 && yi < P.nt - 1;
    const int xc = inGrid ? xi : 0, yc = inGrid ? yi : 0;
    const double u = logr - gridR[xc], v = tau - gridT[yc];
    const int ns = P.ns;
    const int off0 = (xc * (P.nt - 1) + yc) * ns * AGX_CELL_SLOT_BYTES;
    const int zs = P.zeroSlot;
    constexpr bool PF = true;
    double ca[AGX_CELL_COEFS], cb[AGX_CELL_COEFS];
    const int s0 = zs == 0 ? 1 : 0;   // the first slot after the m = 0 one
    double P0, d0r, d0t;
    if constexpr(PF) {
        fetch.load(off0, zs, ca);
        if(s0 < ns) fetch.load(off0, s0, cb);
        cellHorner(ca, u, v, P0, d0r, d0t);
    } else
        fetch.slot(off0, zs, u, v, P0, d0r, d0t);
    if(P.logScaling) {
        const double expX = exp(P0), val = 1. / (P.invPhi0 - expX), dPhidX = val * val * expX;
        P0 = val;
        d0r *= dPhidX;
        d0t *= dPhidX;
    }
    double trPot = P0, trDr = d0r, trDth = d0t, trDphi = 0;
    if(ns > 1) {
        TrigRecT<double> pos, neg;
        pos.init(phi);
        neg.alpha = pos.alpha; neg.beta = pos.beta;
        // harmonic slot s (not the m = 0 slot) from its value and derivatives
        auto addSlot = [&](int s, double Pm, double dmr, double dmt) {
            const int m = P.slotM[s];
            double trig, dtrig;
            if(m > 0) {
                pos.advanceTo(m);
                trig = pos.c; dtrig = double(-m) * pos.s;
            } else {
                neg.restart();
                neg.advanceTo(-m);
                trig = neg.s; dtrig = double(-m) * neg.c;
            }
            if(P.logScaling) {
                dmr = dmr * P0 + Pm * d0r;
                dmt = dmt * P0 + Pm * d0t;
                Pm *= P0;
            }
            trPot  += Pm * trig;
            trDr   += dmr * trig;
            trDth  += dmt * trig;
            trDphi += Pm * dtrig;
        };
        if constexpr(PF) {
            // cb holds slot s on entry; the two buffers alternate
            for(int s = s0; s < ns; ) {
                int s2 = s + 1;
                if(s2 == zs) s2++;
                if(s2 < ns) fetch.load(off0, s2, ca);
                double Pm, dmr, dmt;
                cellHorner(cb, u, v, Pm, dmr, dmt);
                addSlot(s, Pm, dmr, dmt);
                if(s2 >= ns) break;
                int s3 = s2 + 1;
                if(s3 == zs) s3++;
                if(s3 < ns) fetch.load(off0, s3, cb);
                cellHorner(ca, u, v, Pm, dmr, dmt);
                addSlot(s2, Pm, dmr, dmt);
                s = s3;
            }
        } else {
            for(int s = 0; s < ns; s++) {
                if(s == zs) continue;
                double Pm, dmr, dmt;
                fetch.slot(off0, s, u, v, Pm, dmr, dmt);
                addSlot(s, Pm, dmr, dmt);
            }
        }
    }
    const double rinv = 1 / r, r2inv = rinv * rinv;
    const double drdR = R * r2inv, drdz = z * r2inv, dthdR = -tau * rinv, dthdz = rinv - rplusRinv;
    pot = inGrid ? trPot : NAN;
    gR = inGrid ? trDr * drdR + trDth * dthdR : NAN;
    gz = inGrid ? trDr * drdz + trDth * dthdz : NAN;
    gphi = inGrid ? trDphi : NAN;
}

// ---- group evaluation: G lanes (an aligned group of the wavefront) share one point ----
// Used for the longest orbits of a batch, which would otherwise set the batch's duration on a
// single lane: lane g of the group evaluates the harmonic slots s = g, g + G, ... and the
// partial Fourier sums are added across the group.  Every lane of the group ends with the same
// bits (each butterfly step adds the partner's partial sum, and addition commutes).  All lanes of
// the wavefront must execute it together.
constexpr int AGX_GROUP = 8;

// a double moved between lanes by a DPP control (two 32-bit halves); every source lane the
// control names must be active
template<int CTRL>
__device__ __forceinline__ double dppMove(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
template<int G>
__device__ __forceinline__ double groupSum(double v)
{
    if constexpr(G == 8) {
        // the same butterfly as below with DPP moves instead of LDS permutes: quad_perm
        // [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1] (lane ^ 2), then row_half_mirror (lane i of
        // the 8 takes lane 7 - i, which lies in the other quad, holding the other quad's sum)
        v += dppMove<0xB1>(v);
        v += dppMove<0x4E>(v);
        v += dppMove<0x141>(v);
        return v;
    } else {
#pragma unroll
        for(int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o, 64);
        return v;
    }
}

// interiorEval with the harmonics split over the group; g = lane index within the group.
// The combination differs from the serial Fourier sum only in the order of additions.
template<int G, class F, class GP>
__device__ __forceinline__ void interiorEvalGroup(const AgxMultipoleDev& P, const F& fetch,
    double R, double z, double phi, double& pot, double& gR, double& gz, double& gphi,
    GP gridR, GP gridT, int g)
{
    const double r = sqrt(R * R + z * z), logr = log(r), rplusRinv = 1. / (r + fabs(R)),
        tau = R == 0 ? (z > 0 ? 1. : z < 0 ? -1. : 0.) : z * rplusRinv;
    const int xi = binSearchD(logr, gridR, P.nr), yi = binSearchD(tau, gridT, P.nt);
    const bool inGrid = xi >= 0 && xi < P.nr - 1 && yi >= 0 && yi < P.nt - 1;
    const int xc = inGrid ? xi : 0, yc = inGrid ? yi : 0;
    const double xlow = gridR[xc], xupp = gridR[xc + 1], ylow = gridT[yc], yupp = gridT[yc + 1];
    const int offsetCorner = logr == xupp ? (tau == yupp ? 3 : 2) : (tau == yupp ? 1 : 0);
    QuinticWT<double> wy, wx;
    wy.init(tau, ylow, yupp);
    wx.init(logr, xlow, xupp);
    const int ns = P.ns, nt = P.nt, rec0 = (xc * nt + yc) * ns, zs = P.zeroSlot;
    // the m = 0 value (log-unscaled) and the partial sums over this lane's other harmonics
    double P0 = 0, d0r = 0, d0t = 0, Am = 0, Ar = 0, At = 0, Ap = 0;
    if(g < ns) {
        TrigRecT<double> tr;
        tr.init(phi);
        for(int s = g; s < ns; s += G) {
            double Pm, dmr, dmt;
            splineSlot(fetch, rec0 + s, ns, nt, offsetCorner, wy, wx, Pm, dmr, dmt);
            if(s == zs) {
                if(P.logScaling) {
                    const double expX = exp(Pm), val = 1. / (P.invPhi0 - expX), dPhidX = val * val * expX;
                    Pm = val;
                    dmr *= dPhidX;
                    dmt *= dPhidX;
                }
                P0 = Pm; d0r = dmr; d0t = dmt;
            } else {
                const int m = P.slotM[s];
                tr.restart();
                tr.advanceTo(m > 0 ? m : -m);
                const double trig = m > 0 ? tr.c : tr.s, dtrig = double(-m) * (m > 0 ? tr.s : tr.c);
                Am += Pm * trig;
                Ar += dmr * trig;
                At += dmt * trig;
                Ap += Pm * dtrig;
            }
        }
    }
    P0 = groupSum<G>(P0); d0r = groupSum<G>(d0r); d0t = groupSum<G>(d0t);
    Am = groupSum<G>(Am); Ar = groupSum<G>(Ar); At = groupSum<G>(At); Ap = groupSum<G>(Ap);
    double trPot, trDr, trDth, trDphi;
    if(P.logScaling) {   // the other harmonics are stored as ratios to the m = 0 term
        trPot = P0 + P0 * Am;
        trDr = d0r + (P0 * Ar + d0r * Am);
        trDth = d0t + (P0 * At + d0t * Am);
        trDphi = P0 * Ap;
    } else {
        trPot = P0 + Am; trDr = d0r + Ar; trDth = d0t + At; trDphi = Ap;
    }
    const double rinv = 1 / r, r2inv = rinv * rinv;
    const double drdR = R * r2inv, drdz = z * r2inv, dthdR = -tau * rinv, dthdz = rinv - rplusRinv;
    pot = inGrid ? trPot : NAN;
    gR = inGrid ? trDr * drdR + trDth * dthdR : NAN;
    gz = inGrid ? trDr * drdz + trDth * dthdz : NAN;
    gphi = inGrid ? trDphi : NAN;
}

// sphHarmTransformInverseDeriv[2] (potential_multipole.cpp:722-821) followed by
// transformDerivsSphToCyl (:824-848): potential and (R, z, phi) gradient from the (l, m) terms
// of the index set I (fields lmax, mmax, mmin, step, lmin[m + mmax]), where
// term(l, m, C_lm, dC_lm / d ln r) supplies each term once; lmax >= 1
template<class Ind, class Term>
__device__ __forceinline__ void sphHarmSynthesis(const Ind& I, double R, double z, double phi, const Term& term,
    double& pot, double& gR, double& gz, double& gphi)
{
    const double tau = z == 0 ? 0 : z / (sqrt(R * R + z * z) + R);
    double vr, gr, gth, gph = 0;
    if(I.lmax == 2 && I.mmin == 0 && I.step == 2) {
        // sphHarmTransformInverseDeriv2 (potential_multipole.cpp:722-772)
        const double C2 = sqrt(1.25), D2 = sqrt(3.75);
        const double ct = 2 * tau / (1 + tau * tau), st = (1 - tau * tau) / (1 + tau * tau),
            cc = ct * ct, cs = ct * st, ss = st * st, Y20 = (3 * C2 * cc - C2), dY20 = -6 * C2 * cs;
        double P00, d00, P20, d20;
        term(0, 0, P00, d00);
        term(2, 0, P20, d20);
        vr = Y20 * P20 + P00;
        gr = Y20 * d20 + d00;
        gth = dY20 * P20;
        if(I.mmax == 2) {
            double sp, cp, P22, d22;
            term(2, 2, P22, d22);
            const double Y22 = D2 * ss, dY22 = 2 * D2 * cs;
            sincosA(2 * phi, sp, cp);
            vr  += Y22 * P22 * cp;
            gr  += Y22 * d22 * cp;
            gth += dY22 * P22 * cp;
            gph += Y22 * P22 * -2 * sp;
        }
    } else {
        // sphHarmTransformInverseDeriv (potential_multipole.cpp:778-821): per harmonic m the
        // Legendre recurrence of sphHarmArray (math_sphharm.cpp:63-132) streamed in l, then the
        // Fourier synthesis in the reference's order (m = 0 first, then ascending m)
        const double ct = 2 * tau / (1 + tau * tau), st = (1 - tau * tau) / (1 + tau * tau);
        TrigRecT<double> pos, neg;
        pos.init(phi);
        neg.alpha = pos.alpha; neg.beta = pos.beta;
        vr = gr = gth = 0;
        for(int pass = 0; pass < 2; pass++)
        for(int m = (pass == 0 ? 0 : I.mmin); m <= (pass == 0 ? 0 : I.mmax); m++) {
            if(pass == 1 && m == 0) continue;
            const int lmin = I.lmin[m + I.mmax];
            if(lmin > I.lmax) continue;
            const double mul = m == 0 ? 2 * 1.77245385090551602729816748334 :
                2 * 1.77245385090551602729816748334 * M_SQRT2;
            double Cv = 0, Cr = 0, Ct = 0;
            const int am = m < 0 ? -m : m;
            // legendrePmm
            double fact2m = 1, dfact = 1;
            for(int i = 2; i <= 2 * am; i++) fact2m *= i;
            for(int i = 2 * am - 1; i > 1; i -= 2) dfact *= i;
            double prefact = 0.5 / 1.77245385090551602729816748334 * sqrt((2 * am + 1) / fact2m);
            double Pl, dPl;
            if(am == 0) { Pl = prefact; dPl = 0; }
            else if(am == 1) { Pl = -st * prefact; dPl = -ct * prefact; }
            else {
                const double coef = prefact * dfact * (am % 2 == 1 ? -1 : 1);
                double sinm2 = 1, xb = st;
                for(int n = am - 2; n; n >>= 1) { if(n % 2) sinm2 *= xb; xb *= xb; }
                dPl = am * coef * sinm2 * st * ct;
                Pl = coef * sinm2 * (st * st);
            }
            double Plm1 = Pl / prefact, Plm = ct * (2 * am + 1) * Plm1, Plm2 = 0;
            for(int l = am; l <= I.lmax; l++) {
                if(l > am) {
                    if(l > am + 1) Plm = (ct * (2 * l - 1) * Plm1 - (l + am - 1) * Plm2) / (l - am);
                    prefact *= sqrt((2 * l + 1.) / (2 * l - 1.) * (l - am) / (l + am));
                    Pl = Plm * prefact;
                    double d = 0;
                    if(st >= 1e-8 || (am > 2 && st > 0)) d = (l * ct * Plm - (l + am) * Plm1) / st;
                    else if(am == 0) d = -l * (l + 1) / 2 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else if(am == 1) d = -l * (l + 1) / 2 * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else if(am == 2) d = l * (l + 1) * (l + 2) * (l - 1) / 4 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    dPl = prefact * d;
                    Plm2 = Plm1;
                    Plm1 = Plm;
                }
                if(l >= lmin && (l - lmin) % I.step == 0) {
                    double Pv, dPv;
                    term(l, m, Pv, dPv);
                    Cv += Pl * Pv * mul;
                    Cr += Pl * dPv * mul;
                    Ct += dPl * Pv * mul;
                }
            }
            if(m == 0) { vr = Cv; gr = Cr; gth = Ct; continue; }
            double trig, dtrig;
            if(m > 0) { pos.advanceTo(m); trig = pos.c; dtrig = -m * pos.s; }
            else { neg.restart(); neg.advanceTo(-m); trig = neg.s; dtrig = -m * neg.c; }
            vr  += Cv * trig;
            gr  += Cr * trig;
            gth += Ct * trig;
            gph += Cv * dtrig;
        }
    }
    pot = vr;
    // transformDerivsSphToCyl
    const double r2inv = 1 / (R * R + z * z), drdR = R * r2inv, drdz = z * r2inv;
    gR = gr * drdR + gth * drdz;
    gz = gr * drdz + gth * -drdR;
    gphi = gph;
}


// PowerLawMultipole::evalCyl: extrapolation inside rmin / outside rmax
__device__ __forceinline__ void powerLawEval(const AgxPowerLawDev& L, double R, double z, double phi,
    double& pot, double& gR, double& gz, double& gphi)
{
    const double rsq = R * R + z * z, dlogr = log(rsq / L.r0sq) * 0.5;
    const int lmax = (L.inner && rsq < L.r0sq * 1e-16) || (!L.inner && rsq > L.r0sq * 1e16) ? 0 : L.lmax;
    // one (l,m) term: Phi_lm and d Phi_lm / d ln r
    auto term = [&](int l, int m, double& Pv, double& dPv) {
        const int c = l * (l + 1) + m;
        const double s = gmem(L.S)[c], u = gmem(L.U)[c], w = gmem(L.W)[c], v = L.inner ? l : -l - 1;
        const double rv  = v != 0 ? exp(dlogr * v) : 1;
        const double rs  = s != v ? (s != 0 ? exp(dlogr * s) : 1) : rv;
        const double urs = u * rs * (s != v || u == 0 ? 1 : dlogr);
        const double wrv = w * rv;
        const double qr2 = v == 0 ? L.Q * rsq / L.r0sq : 0;
        Pv  = urs + wrv + qr2;
        dPv = urs * s + wrv * v + (s != v ? 0 : u * rs) + qr2 * 2;
    };
    if(lmax == 0) {
        double P0, dP0;
        term(0, 0, P0, dP0);
        pot = P0;
        const double rsqinv = rsq > 0 ? 1 / rsq : 0, Rr2 = rsq < INFINITY ? R * rsqinv : 0,
            zr2 = rsq < INFINITY ? z * rsqinv : 0;
        gR = dP0 * Rr2; gz = dP0 * zr2; gphi = 0;
        return;
    }
    sphHarmSynthesis(L, R, z, phi, term, pot, gR, gz, gphi);
}

// MultipoleInterp1d::evalCyl (potential_multipole.cpp:1753-1815), the interior of expansions
// with lmax <= 2: a quintic spline in ln r per (l, m) term (value and ln r derivative,
// QuinticSpline::evalDeriv, math_spline.cpp:1405-1429), the l=0 term log-unscaled and the others
// multiplied by it, then the spherical-harmonic synthesis.  Lanes outside the grid evaluate the
// end cell (their result is replaced by the extrapolation).
template<class GP>
__device__ __forceinline__ void interior1dEval(const AgxMultipoleDev& P, double R, double z, double phi,
    double& pot, double& gR, double& gz, double& gphi, GP gridR)
{
    const double r = sqrt(R * R + z * z), logr = log(r);
    int k = binSearchD(logr, gridR, P.nr);
    k = k < 0 ? 0 : k > P.nr - 2 ? P.nr - 2 : k;
    QuinticWT<double> w;
    w.init(logr, gridR[k], gridR[k + 1]);
    const vd2 AGX_GLOBAL* nodes = gmem(reinterpret_cast<const vd2*>(P.nodes1d));
    const int nc = P.nc1d;
    auto spline = [&](int c, double& f, double& df) {
        const int lo = (k * nc + c) * 2, hi = ((k + 1) * nc + c) * 2;
        const vd2 a0 = nodes[lo], a1 = nodes[lo + 1], b0 = nodes[hi], b1 = nodes[hi + 1];
        w.eval2(a0.x, b0.x, a0.y, b0.y, a1.x, b1.x, f, df);
    };
    double P0, dP0;
    spline(0, P0, dP0);
    if(P.logScaling) {
        const double expX = exp(P0), Phi = 1 / (P.invPhi0 - expX), dPhidX = Phi * Phi * expX;
        P0 = Phi;
        dP0 *= dPhidX;
    }
    if(P.ind1d.lmax == 0) {   // spherical: d/d ln r only, then transformDerivsSphToCyl
        const double r2inv = 1 / (R * R + z * z);
        pot = P0;
        gR = dP0 * (R * r2inv);
        gz = dP0 * (z * r2inv);
        gphi = 0;
        return;
    }
    auto term = [&](int l, int m, double& Pv, double& dPv) {
        const int c = l * (l + 1) + m;
        if(c == 0) { Pv = P0; dPv = dP0; return; }
        spline(c, Pv, dPv);
        if(P.logScaling) {
            dPv = dPv * P0 + Pv * dP0;
            Pv *= P0;
        }
    };
    sphHarmSynthesis(P.ind1d, R, z, phi, term, pot, gR, gz, gphi);
}

// Multipole::evalCyl dispatch (potential_multipole.cpp:1538-1548).  The interior spline is
// evaluated by every lane (uniform control flow for the cooperative fetcher); lanes in the
// power-law regions then overwrite the result.
// gridR/gridT: the ln r and tau grids, as global-memory pointers or generic ones (an LDS
// copy); null: the header's arrays
template<class F, class GP = const double AGX_GLOBAL*, bool PF = false>
__device__ __forceinline__ void multipoleCyl(const AgxMultipoleDev& P, const F& fetch,
    double R, double z, double phi, double& pot, double& gR, double& gz, double& gphi,
    GP gridR = nullptr, GP gridT = nullptr)
{
    const double rsq = R * R + z * z;
    if(P.interp1d)
        interior1dEval(P, R, z, phi, pot, gR, gz, gphi, gridR ? gridR : GP(gmem(P.gridR)));
    else if constexpr(IsCellFetch<F>::value)
        interiorEvalCell(P, fetch, R, z, phi, pot, gR, gz, gphi, gridR ? gridR : GP(gmem(P.gridR)),
            gridT ? gridT : GP(gmem(P.gridT)));
    else
        interiorEval<F, GP, PF>(P, fetch, R, z, phi, pot, gR, gz, gphi, gridR ? gridR : GP(gmem(P.gridR)),
            gridT ? gridT : GP(gmem(P.gridT)));
    if(rsq < P.rsqInner)
        powerLawEval(P.inner, R, z, phi, pot, gR, gz, gphi);
    else if(rsq > P.rsqOuter)
        powerLawEval(P.outer, R, z, phi, pot, gR, gz, gphi);
}

// NFW potential and radial derivative (src/potential_analytic.cpp:55-75, with its rational
// approximations of ln(1+x)/x and its derivative at small r/a, which set the float values)
// Grids in LDS: a block copies the ln r and tau grids (at most AGX_LDS_GRID doubles together)
// so that the binary searches of every evaluation read LDS instead of making dependent L2
// round trips.  Returns the pointers to use (the header's global arrays if they do not fit).
constexpr int AGX_LDS_GRID = 256;
__device__ __forceinline__ void stageGrids(const AgxMultipoleDev& P, double* lds, const double*& gR,
    const double*& gT)
{
    const int nr = P.nr, nt = P.nt;
    if(nr + nt > AGX_LDS_GRID) { gR = P.gridR; gT = P.gridT; return; }
    for(int i = threadIdx.x; i < nr + nt; i += blockDim.x) lds[i] = i < nr ? P.gridR[i] : P.gridT[i - nr];
    __syncthreads();
    gR = lds;
    gT = lds + nr;
}

__device__ __forceinline__ void nfwDeriv(double mass, double a, double r, double& pot, double& der)
{
    const double q = r / a;
    const double lnr = r == INFINITY ? 0. : q > 0.016 ? log(1 + q) / r :
        (1 + q * (1 + 11. / 60 * q)) / (1 + q * (1.5 + q * (0.6 + q * 0.05))) / a;
    pot = -mass * lnr;
    der = mass * (q > 0.013 ? (lnr - 1 / (r + a)) / r :
        (0.5 + 17. / 96 * q) / (1 + q * (27. / 16 + q * (0.75 + 11. / 160 * q))) / (a * a));
}

// ---- the FP64 kernels' evaluation on the per-cell tables ----
// reciprocal within an ulp of 1/x (x finite, non-zero and normal): the hardware estimate refined
// by two Newton steps (fewer instructions than the correctly rounded division)
__device__ __forceinline__ double rcpD(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// cell of the (near-uniform) ln r grid holding x: i with arr[i] <= x < arr[i+1] (n-2 at the last
// node), -1 outside [arr[0], arr[n-1]]; the same result as binSearchD, from a first guess
// (x - arr[0]) * scale corrected by one node, or the full search when the grid is not uniform
// enough for that
template<class GP>
__device__ __forceinline__ int cellIndexR(double x, GP arr, int n, double scale)
{
    const double a0 = arr[0];
    if(!(x >= a0) || x > arr[n - 1]) return -1;
    int i = (int)((x - a0) * scale);
    i = i < 0 ? 0 : i > n - 2 ? n - 2 : i;
    if(x < arr[i]) i--;
    else if(i < n - 2 && x >= arr[i + 1]) i++;
    if(i < 0 || x < arr[i] || (i < n - 2 && x >= arr[i + 1])) i = binSearchD(x, arr, n);
    return i;
}

// cell of the tau grid holding t in [-1, 1], from the bucket table (each bucket narrower than a
// cell, so at most one node lies inside it)
template<class GP, class IP>
__device__ __forceinline__ int cellIndexT(double t, GP arr, int n, IP lut, int nlut)
{
    if(!(t >= arr[0]) || t > arr[n - 1]) return -1;
    int b = (int)((t + 1) * (0.5 * nlut));
    b = b < 0 ? 0 : b > nlut - 1 ? nlut - 1 : b;
    int i = lut[b];
    while(i < n - 2 && t >= arr[i + 1]) i++;
    return i;
}

// PowerLawMultipole::evalCyl (src/potential_multipole.cpp:1600-1658) with the spherical-harmonic
// synthesis of sphHarmTransformInverseDeriv (:778-821) for the FP64 kernels' per-cell path: the
// same terms and sums as powerLawEval, with the Legendre normalisations and recurrence factors
// taken from a table (no square roots or divisions in the recurrence), (r/r0)^v for the integer
// powers v formed by multiplication, cos/sin(theta) = z/r, R/r and cos/sin(phi) given.  Agrees
// with powerLawEval to rounding.
// the arrays that powerLawEvalFast reads, in global memory or staged in LDS (plStage): S, U, W of
// the inner [0] and outer [1] extrapolation for c < (legLmax+1)^2, the Legendre table and legStart;
// and the tau bucket table of the cell search (lut), in the same address space
template<class GP>
struct PLTabs {
    typedef typename std::conditional<std::is_same<GP, const double AGX_LDS*>::value,
        const int AGX_LDS*, const int AGX_GLOBAL*>::type IP;
    GP S[2], U[2], W[2], leg, legStart;
    IP lut;
};
__device__ __forceinline__ PLTabs<const double AGX_GLOBAL*> plTabsGlobal(const AgxMultipoleDev& P)
{
    PLTabs<const double AGX_GLOBAL*> t;
    t.S[0] = gmem(P.inner.S); t.U[0] = gmem(P.inner.U); t.W[0] = gmem(P.inner.W);
    t.S[1] = gmem(P.outer.S); t.U[1] = gmem(P.outer.U); t.W[1] = gmem(P.outer.W);
    t.leg = gmem(P.legTab); t.legStart = gmem(P.legStart);
    t.lut = gmem(P.tauLut);
    return t;
}
// doubles of the LDS copy made by plStage
__device__ __forceinline__ int plStageDoubles(const AgxMultipoleDev& P)
{
    const int L1 = P.legLmax + 1;
    return (6 + AGX_LEG_STRIDE) * L1 * L1 + L1;
}
// copy the power-law arrays into `lds` (plStageDoubles(P) doubles) and the tau bucket table into
// `lut` (P.nTauLut ints) with the whole block; the caller synchronises the block afterwards
__device__ __forceinline__ PLTabs<const double AGX_LDS*> plStage(const AgxMultipoleDev& P, double AGX_LDS* lds,
    int AGX_LDS* lut)
{
    for(int i = threadIdx.x; i < P.nTauLut; i += blockDim.x) lut[i] = P.tauLut[i];
    const int L1 = P.legLmax + 1, nc = L1 * L1;
    const double* src[8] = { P.inner.S, P.inner.U, P.inner.W, P.outer.S, P.outer.U, P.outer.W, P.legTab, P.legStart };
    const int len[8] = { nc, nc, nc, nc, nc, nc, AGX_LEG_STRIDE * nc, L1 };
    PLTabs<const double AGX_LDS*> t;
    const double AGX_LDS* dst[8];
    int off = 0;
    for(int a = 0; a < 8; a++) {
        dst[a] = lds + off;
        for(int i = threadIdx.x; i < len[a]; i += blockDim.x) lds[off + i] = src[a][i];
        off += len[a];
    }
    t.S[0] = dst[0]; t.U[0] = dst[1]; t.W[0] = dst[2]; t.S[1] = dst[3]; t.U[1] = dst[4]; t.W[1] = dst[5];
    t.leg = dst[6]; t.legStart = dst[7];
    t.lut = lut;
    return t;
}

// powerLawEvalFast for expansions with legLmax <= LM, with every loop over (l, m) unrolled: the
// Legendre recurrence of each |m| is shared by m and -m (the same values), the table and
// coefficient offsets are constants, the recurrences of different |m| are independent chains,
// and exp(dlogr * S) is formed once: for extrapolations whose every S equals v or S[0]
// (AgxPowerLawDev::uniformS; the outer one of a finite-mass model).  The harmonics are added
// to the Fourier sum in the order 0, +-1, +-2, ... instead of 0, mmin, ..., mmax (rounding only).
template<int LM, class GP>
__device__ __forceinline__ void powerLawEvalUnrolled(const AgxMultipoleDev& P, const AgxPowerLawDev& L,
    const PLTabs<GP>& T, double R, double z, double r, double cphi, double sphi, double& pot, double& gR,
    double& gz, double& gphi)
{
    const double rsq = r * r;
    const double dlogr = log(rsq / L.r0sq) * 0.5, qq = exp(dlogr), qstep = L.inner ? qq : rcpD(qq);
    const double rinv = rcpD(r), ct = z * rinv, st = R * rinv, stinv = st > 0 ? rcpD(st) : 0.;
    const int side = L.inner ? 0 : 1;
    const GP S = T.S[side], U = T.U[side], W = T.W[side];
    constexpr int L1 = LM + 1;
    const int Lmax = L.lmax, mmin = L.mmin, mmax = L.mmax, stepMask = L.step - 1;
    const int stride = P.legLmax + 1;
    // (r/r0)^v for every l: v = l inside, -l-1 outside
    double qv[L1];
    qv[0] = L.inner ? 1. : qstep;
#pragma unroll
    for(int l = 1; l < L1; l++) qv[l] = qv[l - 1] * qstep;
    const double alpha = cphi >= 0 ? sphi * sphi * rcpD(1 + cphi) : 1 - cphi;
    double cm = 1, sm = 0;          // cos, sin(am phi) by the recurrence of trigMultiAngle
    // every term of the index set has S equal to v or to S[0] (the caller checks): one exp()
    const double S0 = S[0], eS0 = S0 != 0 ? exp(dlogr * S0) : 1;
    double vr = 0, gr = 0, gth = 0, gph = 0;
#pragma unroll
    for(int am = 0; am < L1; am++) {
        if(am > 0) {
            const double cn = cm - (alpha * cm + sphi * sm), sn = sm - (alpha * sm - sphi * cm);
            cm = cn; sm = sn;
        }
        if(am > Lmax || (am > mmax && -am < mmin)) continue;
        const bool usePos = am <= mmax && L.lmin[am + mmax] <= Lmax;
        const bool useNeg = am > 0 && -am >= mmin && L.lmin[-am + mmax] <= Lmax;
        if(!usePos && !useNeg) continue;
        const int lminP = usePos ? L.lmin[am + mmax] : 1 << 20, lminN = useNeg ? L.lmin[-am + mmax] : 1 << 20;
        const GP tab = T.leg + am * stride * AGX_LEG_STRIDE;
        double Plm1, dP0;
        if(am == 0) { Plm1 = 1; dP0 = 0; }
        else if(am == 1) { Plm1 = -st; dP0 = -ct; }
        else {
            double sinm2 = 1, xb = st;
#pragma unroll
            for(int k = am - 2; k; k >>= 1) { if(k % 2) sinm2 *= xb; xb *= xb; }
            const double coef = T.legStart[am];
            dP0 = am * coef * sinm2 * st * ct;
            Plm1 = coef * sinm2 * (st * st);
        }
        double Plm = ct * (2 * am + 1) * Plm1, Plm2 = 0;
        double CvP = 0, CrP = 0, CtP = 0, CvN = 0, CrN = 0, CtN = 0;
        const bool pole = !(st >= 1e-8 || (am > 2 && st > 0));
#pragma unroll
        for(int l = am; l < L1; l++) {
            if(l > Lmax) break;
            const GP e = tab + l * AGX_LEG_STRIDE;
            const double nrm = e[0];
            double Pl, dPl;
            if(l == am) { Pl = Plm1 * nrm; dPl = dP0 * nrm; }
            else {
                if(l > am + 1) Plm = ct * e[1] * Plm1 - e[2] * Plm2;
                Pl = Plm * nrm;
                double d = (l * ct * Plm - (l + am) * Plm1) * stinv;
                if(pole) {   // the reference's limits at the poles
                    if(am == 0) d = -l * (l + 1) / 2 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else if(am == 1) d = -l * (l + 1) / 2 * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else if(am == 2) d = l * (l + 1) * (l + 2) * (l - 1) / 4 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else d = 0;
                }
                dPl = nrm * d;
                Plm2 = Plm1;
                Plm1 = Plm;
            }
            const double v = L.inner ? l : -l - 1, rv = qv[l];
            auto term = [&](int m, double& Cv, double& Cr, double& Ct) {
                const int c = l * (l + 1) + m;
                const double sc = S[c], u = U[c], w = W[c];
                const double rs = sc != v ? eS0 : rv;
                const double urs = u * rs * (sc != v || u == 0 ? 1 : dlogr);
                const double wrv = w * rv;
                const double qr2 = v == 0 ? L.Q * rsq / L.r0sq : 0;
                const double Pv = urs + wrv + qr2, dPv = urs * sc + wrv * v + (sc != v ? 0 : u * rs) + qr2 * 2;
                Cv += Pl * Pv;
                Cr += Pl * dPv;
                Ct += dPl * Pv;
            };
            if(l >= lminP && ((l - lminP) & stepMask) == 0) term(am, CvP, CrP, CtP);
            if(l >= lminN && ((l - lminN) & stepMask) == 0) term(-am, CvN, CrN, CtN);
        }
        if(am == 0) { vr = CvP; gr = CrP; gth = CtP; continue; }
        if(useNeg) {   // sine terms: sin(am phi), d/dphi = am cos(am phi)
            vr  += CvN * sm;
            gr  += CrN * sm;
            gth += CtN * sm;
            gph += CvN * (am * cm);
        }
        if(usePos) {
            vr  += CvP * cm;
            gr  += CrP * cm;
            gth += CtP * cm;
            gph += CvP * (-am * sm);
        }
    }
    pot = vr;
    // transformDerivsSphToCyl
    const double r2inv = rinv * rinv, drdR = R * r2inv, drdz = z * r2inv;
    gR = gr * drdR + gth * drdz;
    gz = gr * drdz + gth * -drdR;
    gphi = gph;
}

template<class GP>
__device__ __forceinline__ void powerLawEvalFast(const AgxMultipoleDev& P, const AgxPowerLawDev& L,
    const PLTabs<GP>& T, double R, double z, double r, double cphi, double sphi, double& pot, double& gR,
    double& gz, double& gphi)
{
    const double rsq = r * r;
    const bool mono = (L.inner && rsq < L.r0sq * 1e-16) || (!L.inner && rsq > L.r0sq * 1e16) || L.lmax == 0;
    if(mono || (L.lmax == 2 && L.mmin == 0 && L.step == 2)) {   // the cheap branches as they are
        const double phi = R == 0 ? 0. : atan2A(R * sphi, R * cphi);
        powerLawEval(L, R, z, phi, pot, gR, gz, gphi);
        return;
    }
    if(P.legLmax <= 8 && L.uniformS) {
        powerLawEvalUnrolled<8>(P, L, T, R, z, r, cphi, sphi, pot, gR, gz, gphi);
        return;
    }
    const double dlogr = log(rsq / L.r0sq) * 0.5, q = exp(dlogr), qstep = L.inner ? q : rcpD(q);
    const double rinv = rcpD(r), ct = z * rinv, st = R * rinv, stinv = st > 0 ? rcpD(st) : 0.;
    const int side = L.inner ? 0 : 1;
    const GP S = T.S[side], U = T.U[side], W = T.W[side];
    const int L1 = P.legLmax + 1;
    TrigRecT<double> pos, neg;
    pos.alpha = cphi >= 0 ? sphi * sphi * rcpD(1 + cphi) : 1 - cphi;
    pos.beta = sphi;
    pos.restart();
    neg.alpha = pos.alpha; neg.beta = pos.beta;
    double vr = 0, gr = 0, gth = 0, gph = 0;
    for(int pass = 0; pass < 2; pass++)
    for(int m = (pass == 0 ? 0 : L.mmin); m <= (pass == 0 ? 0 : L.mmax); m++) {
        if(pass == 1 && m == 0) continue;
        const int lmin = L.lmin[m + L.mmax];
        if(lmin > L.lmax) continue;
        const int am = m < 0 ? -m : m;
        const GP tab = T.leg + am * L1 * AGX_LEG_STRIDE;
        // P_am^am and its theta derivative before normalisation (legendrePmm)
        double Plm1, dP0;
        if(am == 0) { Plm1 = 1; dP0 = 0; }
        else if(am == 1) { Plm1 = -st; dP0 = -ct; }
        else {
            double sinm2 = 1, xb = st;
            for(int k = am - 2; k; k >>= 1) { if(k % 2) sinm2 *= xb; xb *= xb; }
            const double coef = T.legStart[am];
            dP0 = am * coef * sinm2 * st * ct;
            Plm1 = coef * sinm2 * (st * st);
        }
        double Plm = ct * (2 * am + 1) * Plm1, Plm2 = 0;
        const double dam = am;
        // (r/r0)^v at l = am: v = l inside, -l-1 outside
        double qv = L.inner ? 1. : qstep;
        for(int k = 0; k < am; k++) qv *= qstep;
        double Cv = 0, Cr = 0, Ct = 0;
        // the table entries and coefficients of every l are read unconditionally (all indices
        // are in range), so that the loads of the next l can be issued early
#pragma unroll 2
        for(int l = am; l <= L.lmax; l++, qv *= qstep) {
            const GP e = tab + l * AGX_LEG_STRIDE;
            const double nrm = e[0], A = e[1], B = e[2], dl = e[3];
            const int c = l * (l + 1) + m;
            const double sc = S[c], u = U[c], w = W[c];
            double Pl, dPl;
            if(l == am) { Pl = Plm1 * nrm; dPl = dP0 * nrm; }
            else {
                if(l > am + 1) Plm = ct * A * Plm1 - B * Plm2;
                Pl = Plm * nrm;
                double d = (dl * ct * Plm - (dl + dam) * Plm1) * stinv;
                if(!(st >= 1e-8 || (am > 2 && st > 0))) {   // at the poles (the reference's limits)
                    if(am == 0) d = -l * (l + 1) / 2 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else if(am == 1) d = -l * (l + 1) / 2 * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else if(am == 2) d = l * (l + 1) * (l + 2) * (l - 1) / 4 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else d = 0;
                }
                dPl = nrm * d;
                Plm2 = Plm1;
                Plm1 = Plm;
            }
            if(l >= lmin && ((l - lmin) & (L.step - 1)) == 0) {   // step is 1 or 2
                const double v = L.inner ? dl : -dl - 1;
                const double rv = qv;
                const double rs = sc != v ? (sc != 0 ? exp(dlogr * sc) : 1) : rv;
                const double urs = u * rs * (sc != v || u == 0 ? 1 : dlogr);
                const double wrv = w * rv;
                const double qr2 = v == 0 ? L.Q * rsq / L.r0sq : 0;
                const double Pv = urs + wrv + qr2, dPv = urs * sc + wrv * v + (sc != v ? 0 : u * rs) + qr2 * 2;
                Cv += Pl * Pv;
                Cr += Pl * dPv;
                Ct += dPl * Pv;
            }
        }
        if(m == 0) { vr = Cv; gr = Cr; gth = Ct; continue; }
        double trig, dtrig;
        if(m > 0) { pos.advanceTo(m); trig = pos.c; dtrig = -m * pos.s; }
        else { neg.restart(); neg.advanceTo(-m); trig = neg.s; dtrig = -m * neg.c; }
        vr  += Cv * trig;
        gr  += Cr * trig;
        gth += Ct * trig;
        gph += Cv * dtrig;
    }
    pot = vr;
    // transformDerivsSphToCyl
    const double r2inv = rinv * rinv, drdR = R * r2inv, drdz = z * r2inv;
    gR = gr * drdR + gth * drdz;
    gz = gr * drdz + gth * -drdR;
    gphi = gph;
}

// powerLawEvalFast for a group of G lanes sharing one point (g = lane index within the group):
// lane g forms the Legendre chains of |m| = g, g + G, ... (each shared by m and -m, as in
// powerLawEvalUnrolled) and its part of the Fourier sum; the parts are added across the group.
// The chains have the same operations as in powerLawEvalUnrolled, so that one lane's terms carry
// the same bits; only the order of the final additions differs (rounding).  The whole group
// must take the call together (its lanes hold the same point, so they take the same branch).
template<int G, class GP>
__device__ __forceinline__ void powerLawEvalGroup(const AgxMultipoleDev& P, const AgxPowerLawDev& L,
    const PLTabs<GP>& T, double R, double z, double r, double cphi, double sphi, double& pot, double& gR,
    double& gz, double& gphi, int g)
{
    const double rsq = r * r;
    const bool mono = (L.inner && rsq < L.r0sq * 1e-16) || (!L.inner && rsq > L.r0sq * 1e16) || L.lmax == 0;
    if(mono || (L.lmax == 2 && L.mmin == 0 && L.step == 2)) {   // the cheap branches, on every lane
        const double phi = R == 0 ? 0. : atan2A(R * sphi, R * cphi);
        powerLawEval(L, R, z, phi, pot, gR, gz, gphi);
        return;
    }
    const double dlogr = log(rsq / L.r0sq) * 0.5, qq = exp(dlogr), qstep = L.inner ? qq : rcpD(qq);
    const double rinv = rcpD(r), ct = z * rinv, st = R * rinv, stinv = st > 0 ? rcpD(st) : 0.;
    const int side = L.inner ? 0 : 1;
    const GP S = T.S[side], U = T.U[side], W = T.W[side];
    const int Lmax = L.lmax, mmin = L.mmin, mmax = L.mmax, stepMask = L.step - 1;
    const int stride = P.legLmax + 1;
    const bool uni = L.uniformS != 0;
    const double S0 = S[0], eS0 = uni && S0 != 0 ? exp(dlogr * S0) : 1;
    const double alpha = cphi >= 0 ? sphi * sphi * rcpD(1 + cphi) : 1 - cphi;
    double vr = 0, gr = 0, gth = 0, gph = 0;
    for(int am = g; am <= Lmax; am += G) {
        if(am > mmax && -am < mmin) continue;
        const bool usePos = am <= mmax && L.lmin[am + mmax] <= Lmax;
        const bool useNeg = am > 0 && -am >= mmin && L.lmin[-am + mmax] <= Lmax;
        if(!usePos && !useNeg) continue;
        const int lminP = usePos ? L.lmin[am + mmax] : 1 << 20, lminN = useNeg ? L.lmin[-am + mmax] : 1 << 20;
        // cos, sin(am phi) by the recurrence of trigMultiAngle, and (r/r0)^v at l = am
        double cm = 1, sm = 0, rv = L.inner ? 1. : qstep;
        for(int k = 0; k < am; k++) {
            const double cn = cm - (alpha * cm + sphi * sm), sn = sm - (alpha * sm - sphi * cm);
            cm = cn; sm = sn;
            rv *= qstep;
        }
        const GP tab = T.leg + am * stride * AGX_LEG_STRIDE;
        double Plm1, dP0;
        if(am == 0) { Plm1 = 1; dP0 = 0; }
        else if(am == 1) { Plm1 = -st; dP0 = -ct; }
        else {
            double sinm2 = 1, xb = st;
            for(int k = am - 2; k; k >>= 1) { if(k % 2) sinm2 *= xb; xb *= xb; }
            const double coef = T.legStart[am];
            dP0 = am * coef * sinm2 * st * ct;
            Plm1 = coef * sinm2 * (st * st);
        }
        double Plm = ct * (2 * am + 1) * Plm1, Plm2 = 0;
        double CvP = 0, CrP = 0, CtP = 0, CvN = 0, CrN = 0, CtN = 0;
        const bool pole = !(st >= 1e-8 || (am > 2 && st > 0));
        for(int l = am; l <= Lmax; l++, rv *= qstep) {
            const GP e = tab + l * AGX_LEG_STRIDE;
            const double nrm = e[0];
            double Pl, dPl;
            if(l == am) { Pl = Plm1 * nrm; dPl = dP0 * nrm; }
            else {
                if(l > am + 1) Plm = ct * e[1] * Plm1 - e[2] * Plm2;
                Pl = Plm * nrm;
                double d = (l * ct * Plm - (l + am) * Plm1) * stinv;
                if(pole) {   // the reference's limits at the poles
                    if(am == 0) d = -l * (l + 1) / 2 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else if(am == 1) d = -l * (l + 1) / 2 * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else if(am == 2) d = l * (l + 1) * (l + 2) * (l - 1) / 4 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else d = 0;
                }
                dPl = nrm * d;
                Plm2 = Plm1;
                Plm1 = Plm;
            }
            const double v = L.inner ? l : -l - 1;
            auto term = [&](int m, double& Cv, double& Cr, double& Ct) {
                const int c = l * (l + 1) + m;
                const double sc = S[c], u = U[c], w = W[c];
                const double rs = sc != v ? (uni ? eS0 : sc != 0 ? exp(dlogr * sc) : 1) : rv;
                const double urs = u * rs * (sc != v || u == 0 ? 1 : dlogr);
                const double wrv = w * rv;
                const double qr2 = v == 0 ? L.Q * rsq / L.r0sq : 0;
                const double Pv = urs + wrv + qr2, dPv = urs * sc + wrv * v + (sc != v ? 0 : u * rs) + qr2 * 2;
                Cv += Pl * Pv;
                Cr += Pl * dPv;
                Ct += dPl * Pv;
            };
            if(l >= lminP && ((l - lminP) & stepMask) == 0) term(am, CvP, CrP, CtP);
            if(l >= lminN && ((l - lminN) & stepMask) == 0) term(-am, CvN, CrN, CtN);
        }
        if(am == 0) { vr += CvP; gr += CrP; gth += CtP; continue; }
        if(useNeg) {   // sine terms: sin(am phi), d/dphi = am cos(am phi)
            vr  += CvN * sm;
            gr  += CrN * sm;
            gth += CtN * sm;
            gph += CvN * (am * cm);
        }
        if(usePos) {
            vr  += CvP * cm;
            gr  += CrP * cm;
            gth += CtP * cm;
            gph += CvP * (-am * sm);
        }
    }
    vr = groupSum<G>(vr); gr = groupSum<G>(gr); gth = groupSum<G>(gth); gph = groupSum<G>(gph);
    pot = vr;
    // transformDerivsSphToCyl
    const double r2inv = rinv * rinv, drdR = R * r2inv, drdz = z * r2inv;
    gR = gr * drdR + gth * drdz;
    gz = gr * drdz + gth * -drdR;
    gphi = gph;
}

// multipoleCar on the per-cell tables (fetch: FetchCell).  The same function as
// Multipole::evalCyl behind BasePotentialCyl::evalCar (src/potential_multipole.cpp:1538-1548,
// 1969-2080), arranged for fewer instructions: cos and sin of the azimuth come from x/R and
// y/R instead of sincos(atan2(y, x)) (the angle itself is formed only for the power-law
// branches), the trigonometric recurrence starts from them (2 sin^2(phi/2) = sin^2 / (1 + cos)),
// reciprocals replace divisions, and the cell search uses the uniform ln r grid and the tau
// bucket table.  Results agree with the reference's sequence to rounding.
template<class F, class GP, bool NFW>
__device__ __forceinline__ void multipoleCarCell(const AgxMultipoleDev& P, const F& fetch,
    double x, double y, double z, double& pot, double& gx, double& gy, double& gz, GP gridR, GP gridT,
    const PLTabs<GP>& plt)
{
    const double R2 = x * x + y * y, R = sqrt(R2), rsq = R2 + z * z, r = sqrt(rsq);
    const bool axis = R == 0;
    const double Rinv = axis ? 0. : rcpD(R);
    const double cphi = axis ? 1. : x * Rinv, sphi = axis ? 0. : y * Rinv;
    const double logr = log(r), rplusRinv = rcpD(r + R),
        tau = axis ? (z > 0 ? 1. : z < 0 ? -1. : 0.) : z * rplusRinv;
    const int xi = cellIndexR(logr, gridR, P.nr, P.lnrScale),
        yi = cellIndexT(tau, gridT, P.nt, plt.lut, P.nTauLut);
    const bool inGrid = xi >= 0 && yi >= 0;
    const int xc = inGrid ? xi : 0, yc = inGrid ? yi : 0;
    const double u = logr - gridR[xc], v = tau - gridT[yc];
    const int ns = P.ns, zs = P.zeroSlot;
    const int off0 = (xc * (P.nt - 1) + yc) * ns * AGX_CELL_SLOT_BYTES;
    constexpr bool PF = true;
    double ca[AGX_CELL_COEFS], cb[AGX_CELL_COEFS];
    const int s0 = zs == 0 ? 1 : 0;   // the first slot after the m = 0 one
    double P0, d0r, d0t;
    if constexpr(PF) {
        fetch.load(off0, zs, ca);
        if(s0 < ns) fetch.load(off0, s0, cb);
        cellHorner(ca, u, v, P0, d0r, d0t);
    } else
        fetch.slot(off0, zs, u, v, P0, d0r, d0t);
    if(P.logScaling) {
        const double expX = exp(P0), val = rcpD(P.invPhi0 - expX), dPhidX = val * val * expX;
        P0 = val;
        d0r *= dPhidX;
        d0t *= dPhidX;
    }
    double trPot = P0, trDr = d0r, trDth = d0t, trDphi = 0;
    if(ns > 1) {
        // cos/sin(k phi) by the recurrence of trigMultiAngle (math_sphharm.cpp:134-152)
        TrigRecT<double> pos, neg;
        pos.alpha = cphi >= 0 ? sphi * sphi * rcpD(1 + cphi) : 1 - cphi;
        pos.beta = sphi;
        pos.restart();
        neg.alpha = pos.alpha; neg.beta = pos.beta;
        auto addSlot = [&](int s, double Pm, double dmr, double dmt) {
            const int m = P.slotM[s];
            double trig, dtrig;
            if(m > 0) {
                pos.advanceTo(m);
                trig = pos.c; dtrig = double(-m) * pos.s;
            } else {
                neg.restart();
                neg.advanceTo(-m);
                trig = neg.s; dtrig = double(-m) * neg.c;
            }
            if(P.logScaling) {
                dmr = dmr * P0 + Pm * d0r;
                dmt = dmt * P0 + Pm * d0t;
                Pm *= P0;
            }
            trPot  += Pm * trig;
            trDr   += dmr * trig;
            trDth  += dmt * trig;
            trDphi += Pm * dtrig;
        };
        if constexpr(PF) {
            // cb holds slot s on entry; the two buffers alternate
            for(int s = s0; s < ns; ) {
                int s2 = s + 1;
                if(s2 == zs) s2++;
                if(s2 < ns) fetch.load(off0, s2, ca);
                double Pm, dmr, dmt;
                cellHorner(cb, u, v, Pm, dmr, dmt);
                addSlot(s, Pm, dmr, dmt);
                if(s2 >= ns) break;
                int s3 = s2 + 1;
                if(s3 == zs) s3++;
                if(s3 < ns) fetch.load(off0, s3, cb);
                cellHorner(ca, u, v, Pm, dmr, dmt);
                addSlot(s2, Pm, dmr, dmt);
                s = s3;
            }
        } else {
            for(int s = 0; s < ns; s++) {
                if(s == zs) continue;
                double Pm, dmr, dmt;
                fetch.slot(off0, s, u, v, Pm, dmr, dmt);
                addSlot(s, Pm, dmr, dmt);
            }
        }
    }
    const double rinv = rcpD(r), r2inv = rinv * rinv;
    const double drdR = R * r2inv, drdz = z * r2inv, dthdR = -tau * rinv, dthdz = rinv - rplusRinv;
    double gR = inGrid ? trDr * drdR + trDth * dthdR : NAN;
    double gzc = inGrid ? trDr * drdz + trDth * dthdz : NAN;
    double gphi = inGrid ? trDphi : NAN;
    pot = inGrid ? trPot : NAN;
    // power-law extrapolation (lanes beyond the grid)
    if(rsq < P.rsqInner)
        powerLawEvalFast(P, P.inner, plt, R, z, r, cphi, sphi, pot, gR, gzc, gphi);
    else if(rsq > P.rsqOuter)
        powerLawEvalFast(P, P.outer, plt, R, z, r, cphi, sphi, pot, gR, gzc, gphi);
    // BasePotentialCyl::evalCar: gradient back to Cartesian (on the axis the reference takes
    // dR/dx = dR/dy = dphi/dx = dphi/dy = 1)
    const double dRdx = axis ? 1. : cphi, dRdy = axis ? 1. : sphi,
        dphidx = axis ? 1. : -sphi * Rinv, dphidy = axis ? 1. : cphi * Rinv;
    gx = gR * dRdx + gphi * dphidx;
    gy = gR * dRdy + gphi * dphidy;
    gz = gzc;
    if(NFW && P.nfwMass != 0) {
        double pn, dn;
        nfwDeriv(P.nfwMass, P.nfwScale, r, pn, dn);
        const double xr = r == 0 ? 0. : x * rinv, yr = r == 0 ? 0. : y * rinv, zr = r == 0 ? 0. : z * rinv;
        pot = pot + pn;
        gx = xr * dn + gx;
        gy = yr * dn + gy;
        gz = zr * dn + gz;
    }
}

// multipoleCarCell for an aligned group of G lanes sharing one point (the longest orbits of a
// batch): lane g evaluates the harmonic slots s = g, g + G, ... and the partial sums (the m = 0
// value and derivatives, the m != 0 Fourier sums) are added across the group by a butterfly,
// which leaves the same bits on every lane of the group.  Only the order of the additions
// differs from the one-lane form.  All lanes of the wavefront must execute it together.
template<int G, class GP>
__device__ __forceinline__ void multipoleCarCellGroup(const AgxMultipoleDev& P, const FetchCell& fetch,
    double x, double y, double z, double& pot, double& gx, double& gy, double& gz, GP gridR, GP gridT,
    const PLTabs<GP>& plt, int g)
{
    const double R2 = x * x + y * y, R = sqrt(R2), rsq = R2 + z * z, r = sqrt(rsq);
    const bool axis = R == 0;
    const double Rinv = axis ? 0. : rcpD(R);
    const double cphi = axis ? 1. : x * Rinv, sphi = axis ? 0. : y * Rinv;
    const double logr = log(r), rplusRinv = rcpD(r + R),
        tau = axis ? (z > 0 ? 1. : z < 0 ? -1. : 0.) : z * rplusRinv;
    const int xi = cellIndexR(logr, gridR, P.nr, P.lnrScale),
        yi = cellIndexT(tau, gridT, P.nt, plt.lut, P.nTauLut);
    const bool inGrid = xi >= 0 && yi >= 0;
    const int xc = inGrid ? xi : 0, yc = inGrid ? yi : 0;
    const double u = logr - gridR[xc], v = tau - gridT[yc];
    const int ns = P.ns, zs = P.zeroSlot;
    const int off0 = (xc * (P.nt - 1) + yc) * ns * AGX_CELL_SLOT_BYTES;
    double P0 = 0, d0r = 0, d0t = 0, Am = 0, Ar = 0, At = 0, Ap = 0;
    if(g < ns) {
        TrigRecT<double> tr;
        tr.alpha = cphi >= 0 ? sphi * sphi * rcpD(1 + cphi) : 1 - cphi;
        tr.beta = sphi;
        for(int s = g; s < ns; s += G) {
            double Pm, dmr, dmt;
            fetch.slot(off0, s, u, v, Pm, dmr, dmt);
            if(s == zs) {
                if(P.logScaling) {
                    const double expX = exp(Pm), val = rcpD(P.invPhi0 - expX), dPhidX = val * val * expX;
                    Pm = val;
                    dmr *= dPhidX;
                    dmt *= dPhidX;
                }
                P0 = Pm; d0r = dmr; d0t = dmt;
            } else {
                const int m = P.slotM[s];
                tr.restart();
                tr.advanceTo(m > 0 ? m : -m);
                const double trig = m > 0 ? tr.c : tr.s, dtrig = double(-m) * (m > 0 ? tr.s : tr.c);
                Am += Pm * trig;
                Ar += dmr * trig;
                At += dmt * trig;
                Ap += Pm * dtrig;
            }
        }
    }
    P0 = groupSum<G>(P0); d0r = groupSum<G>(d0r); d0t = groupSum<G>(d0t);
    Am = groupSum<G>(Am); Ar = groupSum<G>(Ar); At = groupSum<G>(At); Ap = groupSum<G>(Ap);
    double trPot, trDr, trDth, trDphi;
    if(P.logScaling) {   // the other harmonics are stored as ratios to the m = 0 term
        trPot = P0 + P0 * Am;
        trDr = d0r + (P0 * Ar + d0r * Am);
        trDth = d0t + (P0 * At + d0t * Am);
        trDphi = P0 * Ap;
    } else {
        trPot = P0 + Am; trDr = d0r + Ar; trDth = d0t + At; trDphi = Ap;
    }
    const double rinv = rcpD(r), r2inv = rinv * rinv;
    const double drdR = R * r2inv, drdz = z * r2inv, dthdR = -tau * rinv, dthdz = rinv - rplusRinv;
    double gR = inGrid ? trDr * drdR + trDth * dthdR : NAN;
    double gzc = inGrid ? trDr * drdz + trDth * dthdz : NAN;
    double gphi = inGrid ? trDphi : NAN;
    pot = inGrid ? trPot : NAN;
    if(rsq < P.rsqInner)
        powerLawEvalGroup<G>(P, P.inner, plt, R, z, r, cphi, sphi, pot, gR, gzc, gphi, g);
    else if(rsq > P.rsqOuter)
        powerLawEvalGroup<G>(P, P.outer, plt, R, z, r, cphi, sphi, pot, gR, gzc, gphi, g);
    const double dRdx = axis ? 1. : cphi, dRdy = axis ? 1. : sphi,
        dphidx = axis ? 1. : -sphi * Rinv, dphidy = axis ? 1. : cphi * Rinv;
    gx = gR * dRdx + gphi * dphidx;
    gy = gR * dRdy + gphi * dphidy;
    gz = gzc;
}

// BasePotentialCyl::evalCar: Car -> Cyl, evaluate, gradient back to Car (+dPhi/dx)
// NFW: whether the optional analytic halo is handled (kernels for potentials without one are
// instantiated with false: keeping x, y, z live to the end costs a wavefront of occupancy)
template<class F, class GP = const double AGX_GLOBAL*, bool NFW = true, bool PF = false>
__device__ __forceinline__ void multipoleCar(const AgxMultipoleDev& P, const F& fetch,
    double x, double y, double z, double& pot, double& gx, double& gy, double& gz,
    GP gridR = nullptr, GP gridT = nullptr)
{
    // (kernels with grids in LDS call multipoleCarCell themselves, with the power-law arrays in LDS)
    if constexpr(IsCellFetch<F>::value && std::is_same<GP, const double AGX_GLOBAL*>::value) {
        if(!P.interp1d) {
            multipoleCarCell<F, GP, NFW>(P, fetch, x, y, z, pot, gx, gy, gz, gridR ? gridR : GP(gmem(P.gridR)),
                gridT ? gridT : GP(gmem(P.gridT)), plTabsGlobal(P));
            return;
        }
    }
    const double R = sqrt(x * x + y * y);
    double dRdx, dRdy, dphidx, dphidy, phi;
    if(R == 0) {
        dRdx = dRdy = dphidx = dphidy = 1.;
        phi = 0;
    } else {
        const double cosphi = x / R, sinphi = y / R;
        dRdx = cosphi; dRdy = sinphi; dphidx = -sinphi / R; dphidy = cosphi / R;
        phi = atan2A(y, x);
    }
    double gR, gzc, gphi;
    multipoleCyl<F, GP, PF>(P, fetch, R, z, phi, pot, gR, gzc, gphi, gridR, gridT);
    gx = gR * dRdx + gphi * dphidx;
    gy = gR * dRdy + gphi * dphidy;
    gz = gzc;
    if(NFW && P.nfwMass != 0) {
        // the spherical component is evaluated in Cartesian coordinates and added first,
        // the cylindrical one's converted gradient after it (potential_composite.cpp:134-151)
        const double r = sqrt(x * x + y * y + z * z);
        double pn, dn;
        nfwDeriv(P.nfwMass, P.nfwScale, r, pn, dn);
        const double xr = r == 0 ? 0. : x / r, yr = r == 0 ? 0. : y / r, zr = r == 0 ? 0. : z / r;
        pot = pot + pn;
        gx = xr * dn + gx;
        gy = yr * dn + gy;
        gz = zr * dn + gz;
    }
}

// multipoleCar (without the NFW halo) for a group of G lanes sharing one point
template<int G, class F, class GP>
__device__ __forceinline__ void multipoleCarGroup(const AgxMultipoleDev& P, const F& fetch,
    double x, double y, double z, double& pot, double& gx, double& gy, double& gz, GP gridR, GP gridT, int g)
{
    const double R = sqrt(x * x + y * y);
    double dRdx, dRdy, dphidx, dphidy, phi;
    if(R == 0) {
        dRdx = dRdy = dphidx = dphidy = 1.;
        phi = 0;
    } else {
        const double cosphi = x / R, sinphi = y / R;
        dRdx = cosphi; dRdy = sinphi; dphidx = -sinphi / R; dphidy = cosphi / R;
        phi = atan2A(y, x);
    }
    double gR, gzc, gphi;
    const double rsq = R * R + z * z;
    interiorEvalGroup<G>(P, fetch, R, z, phi, pot, gR, gzc, gphi, gridR, gridT, g);
    if(rsq < P.rsqInner)
        powerLawEval(P.inner, R, z, phi, pot, gR, gzc, gphi);
    else if(rsq > P.rsqOuter)
        powerLawEval(P.outer, R, z, phi, pot, gR, gzc, gphi);
    gx = gR * dRdx + gphi * dphidx;
    gy = gR * dRdy + gphi * dphidy;
    gz = gzc;
}

}  // namespace agxd
