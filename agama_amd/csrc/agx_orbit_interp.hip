// Interpolated orbits: the quintic splines behind agama.Orbit objects, built on the device from
// the rows recorded by the orbit kernels, and their batched evaluation.  Restates
// createOrbitObject (src/interface_python.cpp:6793-6880: rows taken into the inertial frame,
// duplicate timestamps skipped, time order restored for backward integration), the quintic
// spline from values and first derivatives (src/math_spline.cpp:455-486 with the tridiagonal
// solve of src/math_linalg.cpp:280-298), its evaluation (math_spline.cpp:653-718, 1405-1429)
// and Orbit.__call__ (interface_python.cpp:6643-6714).
//
// Build: one lane per series walks its rows twice (nodes, then the tridiagonal sweep of the
// three coordinates, which share the matrix: one pivot row per node).  Evaluation: one lane
// per (series, time) query with a binary search over the series' node times.
#include <hip/hip_runtime.h>
#include <cmath>
#include "agx_device.h"
#include "agx_kernels.h"

using namespace agxd;

namespace {

constexpr int ND = AGX_NODE_DOUBLES;

__global__ __launch_bounds__(256) void buildSplinesKernel(const AgxSeries* __restrict__ ser, long ns,
    double Omega, double* __restrict__ nodes, double* __restrict__ piv, int* __restrict__ count,
    int* __restrict__ reversedOut)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(s >= ns) return;
    const AgxSeries S = ser[s];
    long rows = 0;   // recorded rows: those before the first NaN timestamp
    while(rows < S.rows && S.t[rows] == S.t[rows]) rows++;
    const bool reversed = rows > 1 && S.t[1] < S.t[0];
    double* nd = nodes + S.node * ND;
    double* U = piv + S.node;
    long k = 0;
    for(long idx = 0; idx < rows; idx++) {
        const long r = reversed ? rows - 1 - idx : idx;
        const double ti = S.t[r];
        if(k > 0 && ti == nd[(k - 1) * ND]) continue;   // a repeated timestamp
        const double* p = S.v + r * 6;
        double ca = 1, sa = 0;
        sincosA(Omega * ti, sa, ca);
        double* o = nd + k * ND;
        o[0] = ti;
        o[1] = p[0] * ca - p[1] * sa;
        o[2] = p[1] * ca + p[0] * sa;
        o[3] = p[2];
        o[4] = p[3] * ca - p[4] * sa;
        o[5] = p[4] * ca + p[3] * sa;
        o[6] = p[5];
        k++;
    }
    count[s] = (int)k;
    reversedOut[s] = reversed ? 1 : 0;
    if(k < 2) {
        for(long i = 0; i < k; i++) nd[i * ND + 7] = nd[i * ND + 8] = nd[i * ND + 9] = 0;
        return;
    }
    // second derivatives: continuity of the third derivative at interior nodes, vanishing
    // fourth derivative at the two ends; forward elimination, then back-substitution
    for(long i = 0; i < k; i++) {
        const double* c = nd + i * ND;
        double di = 0, lo = 0, up = 0, rhs[3] = { 0, 0, 0 };
        if(i > 0) {
            const double* b = c - ND;
            const double hi = 1 / (c[0] - b[0]);
            di += hi * 3;
            lo = -hi;
            for(int d = 0; d < 3; d++)
                rhs[d] -= (20 * (c[1 + d] - b[1 + d]) * hi - 12 * c[4 + d] - 8 * b[4 + d]) * hi * hi;
        }
        if(i < k - 1) {
            const double* a = c + ND;
            const double hi = 1 / (a[0] - c[0]);
            di += hi * 3;
            up = -hi;
            for(int d = 0; d < 3; d++)
                rhs[d] += (20 * (a[1 + d] - c[1 + d]) * hi - 12 * c[4 + d] - 8 * a[4 + d]) * hi * hi;
        }
        if(i == 0) {
            const double* a = c + ND;
            const double hi = 1 / (a[0] - c[0]);
            up = -2 * hi;
            for(int d = 0; d < 3; d++)
                rhs[d] = (30 * (a[1 + d] - c[1 + d]) * hi - 14 * a[4 + d] - 16 * c[4 + d]) * hi * hi;
        }
        if(i == k - 1) {
            const double* b = c - ND;
            const double hi = 1 / (c[0] - b[0]);
            lo = -2 * hi;
            for(int d = 0; d < 3; d++)
                rhs[d] = (-30 * (c[1 + d] - b[1 + d]) * hi + 14 * b[4 + d] + 16 * c[4 + d]) * hi * hi;
        }
        const double ip = 1. / (di - (i > 0 ? lo * U[i - 1] : 0));
        U[i] = up * ip;
        double* o = nd + i * ND;
        for(int d = 0; d < 3; d++)
            o[7 + d] = (rhs[d] - (i > 0 ? lo * o[7 + d - ND] : 0)) * ip;
    }
    for(long i = k - 2; i >= 0; i--)
        for(int d = 0; d < 3; d++) nd[i * ND + 7 + d] -= U[i] * nd[(i + 1) * ND + 7 + d];
}

// value, first and second derivative of the three coordinates at x in [node a, node a+1]
__device__ __forceinline__ void quinticSegment(const double* a, double x, double* f, double* df, double* d2f)
{
    const double* b = a + ND;
    const double xl = a[0], xh = b[0];
    if(x == xh) {   // the right end exactly (no rounding)
        for(int d = 0; d < 3; d++) { f[d] = b[1 + d]; df[d] = b[4 + d]; d2f[d] = b[7 + d]; }
        return;
    }
    const double dx = x - xl, h = xh - xl, hi = 1 / h, t = dx * hi, t2 = t * t, t3 = t * t2,
        u = t * (1 - t), w = 30 * u * hi;
    // Hermite basis of the differences beyond the quadratic Taylor part and its derivatives
    const double P = t3 * (10 - t * (15 - 6 * t)), Q = t3 * (1 - t * 0.5) * h,
        R = t3 * (1 - t * (1.25 - 0.5 * t)) * h * h,
        P1 = w * u, Q1 = t2 * (3 - t * 2), R1 = t2 * (3 - t * (5 - 2.5 * t)) * h,
        P2 = w * hi * (2 - 4 * t), Q2 = w * 0.2, R2 = t * (6 - t * (15 - 10 * t));
    for(int d = 0; d < 3; d++) {
        const double fl = a[1 + d], fh = b[1 + d], gl = a[4 + d], gh = b[4 + d], cl = a[7 + d], ch = b[7 + d];
        const double e0 = fh - fl - 0.5 * h * (gh + gl), e1 = gh - gl - 0.5 * h * (ch + cl), e2 = ch - cl;
        f[d] = fl + P * e0 + dx * gl + Q * e1 + 0.5 * dx * dx * cl + R * e2;
        df[d] = P1 * e0 + gl + Q1 * e1 + dx * cl + R1 * e2;
        d2f[d] = P2 * e0 + Q2 * e1 + cl + R2 * e2;
    }
}

// mode 0: Orbit.__call__ - position and velocity in the rotating frame, NaN outside the
// recorded time span; mode 1: the three inertial-frame splines with their derivatives
// (x, y, z, x', y', z', x'', y'', z''), linear extrapolation outside the span
__global__ __launch_bounds__(256) void evalSplinesKernel(const AgxSeries* __restrict__ ser,
    const int* __restrict__ count, const double* __restrict__ nodes, double Omega, long m,
    const long* __restrict__ which, const double* __restrict__ tq, int mode, double* __restrict__ out)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= m) return;
    const long s = which[q];
    const double x = tq[q];
    const int n = count[s];
    const double* nd = nodes + ser[s].node * ND;
    const int width = mode == 0 ? 6 : 9;
    double* o = out + q * width;
    const bool inside = n > 0 && x >= nd[0] && x <= nd[(long)(n - 1) * ND];
    if(x != x || n == 0 || (mode == 0 && !inside)) {
        for(int c = 0; c < width; c++) o[c] = NAN;
        return;
    }
    double f[3], df[3], d2f[3];
    if(n >= 2 && inside) {
        int lo = 0, hi = n - 1;   // the interval [lo, lo+1] holding x (the last one at the right end)
        while(hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if(nd[(long)mid * ND] > x) hi = mid; else lo = mid;
        }
        quinticSegment(nd + (long)lo * ND, x, f, df, d2f);
    } else {   // beyond an end, or a single node: the end value continued with its slope
        const double* e = x < nd[0] ? nd : nd + (long)(n - 1) * ND;
        for(int d = 0; d < 3; d++) {
            f[d] = e[1 + d] + (e[4 + d] == 0 ? 0 : e[4 + d] * (x - e[0]));
            df[d] = e[4 + d];
            d2f[d] = 0;
        }
    }
    if(mode == 1) {
        for(int d = 0; d < 3; d++) { o[d] = f[d]; o[3 + d] = df[d]; o[6 + d] = d2f[d]; }
        return;
    }
    double ca = 1, sa = 0;
    if(Omega != 0) sincosA(Omega * x, sa, ca);
    o[0] = f[0] * ca + f[1] * sa;
    o[1] = f[1] * ca - f[0] * sa;
    o[2] = f[2];
    o[3] = df[0] * ca + df[1] * sa;
    o[4] = df[1] * ca - df[0] * sa;
    o[5] = df[2];
}

// copy row segments: segment s moves rows[s] rows of `width` doubles from row srcRow[s] of src
// to row dstRow[s] of dst; one workgroup per segment at a time (coalesced within a segment)
__global__ __launch_bounds__(256) void gatherRowsKernel(long nseg, const long* __restrict__ srcRow,
    const long* __restrict__ dstRow, const long* __restrict__ rows, int width, const double* __restrict__ src,
    double* __restrict__ dst)
{
    for(long s = blockIdx.x; s < nseg; s += gridDim.x) {
        const long len = rows[s] * width;
        const double* a = src + srcRow[s] * width;
        double* b = dst + dstRow[s] * width;
        for(long k = threadIdx.x; k < len; k += blockDim.x) b[k] = a[k];
    }
}

}  // namespace

hipError_t agxLaunchGatherRows(long nseg, const long* srcRow, const long* dstRow, const long* rows, int width,
    const double* src, double* dst, hipStream_t stream)
{
    if(nseg <= 0) return hipSuccess;
    const long blocks = nseg < 65536 ? nseg : 65536;
    hipLaunchKernelGGL(gatherRowsKernel, dim3((unsigned)blocks), dim3(256), 0, stream, nseg, srcRow, dstRow, rows,
        width, src, dst);
    return hipGetLastError();
}

hipError_t agxLaunchBuildSplines(const AgxSeries* dSeries, long ns, double Omega, double* nodes,
    double* pivots, int* count, int* reversed, hipStream_t stream)
{
    if(ns <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(buildSplinesKernel, dim3((unsigned)((ns + block - 1) / block)), dim3(block), 0, stream,
        dSeries, ns, Omega, nodes, pivots, count, reversed);
    return hipGetLastError();
}

hipError_t agxLaunchEvalSplines(const AgxSeries* dSeries, const int* count, const double* nodes, double Omega,
    long m, const long* which, const double* t, int mode, double* out, hipStream_t stream)
{
    if(m <= 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(evalSplinesKernel, dim3((unsigned)((m + block - 1) / block)), dim3(block), 0, stream,
        dSeries, count, nodes, Omega, m, which, t, mode, out);
    return hipGetLastError();
}
