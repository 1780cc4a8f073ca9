// Target plug-ins collected during orbit integration on the device: RuntimeFncTarget
// (src/galaxymodel_target.h:95-155: ten dense-output samples per accepted step, each weighted
// by its share of the step) with the density discretisations TargetDensitySphHarm
// (src/galaxymodel_densitygrid.cpp:356-411: linear B-splines in r times even spherical
// harmonics), TargetDensityCylindrical<0/1> (:486-545, 92-118: top-hat cells or bilinear
// B-splines in (R, |z|) times even azimuthal Fourier terms), TargetDensityClassic<0/1> (:146-239,
// 59-90: spheroidal shells, three panes of strips each, top-hat cells or trilinear nodes) and the
// kinematic
// TargetKinemShell<N> (src/galaxymodel_losvd.cpp:568-576: v_r^2 and v_t^2 on B-splines of degree
// N in r, src/math_spline.cpp:35-82, 1877-1886).  One lane accumulates its orbit's datacube in FP64 in global memory (its own
// row: no atomics); the host normalises by the integration time at the end.
#pragma once
#include <hip/hip_runtime.h>
#include "agx_device.h"

enum : int { AGX_TARGET_DENSITY_SPHHARM = 1, AGX_TARGET_DENSITY_CYL_TOPHAT = 2, AGX_TARGET_DENSITY_CYL_LINEAR = 3,
    AGX_TARGET_KINEM_SHELL = 4, AGX_TARGET_DENSITY_CLASSIC_TOPHAT = 5, AGX_TARGET_DENSITY_CLASSIC_LINEAR = 6 };
enum : int { AGX_MAX_TARGETS = 4 };

struct AgxTargetDev {
    int type;
    int lmax, mmax;          // even orders of the angular expansion (KinemShell: lmax = B-spline degree;
                             // DensityClassic: mmax = strips per pane)
    int nr;                  // radial grid nodes (the origin implied)
    int nz;                  // vertical grid nodes (cylindrical targets; z = 0 implied)
    int numValues;           // datacube length; the row holds one more value: the accumulated time
    const double* gridr;     // nr nodes, increasing, > 0 (spherical radius, or cylindrical R)
    const double* gridz;     // nz nodes, increasing, > 0 (cylindrical targets)
    double* cube;            // n x (numValues + 1) per orbit
    double axis[3];          // DensityClassic: axes of the spheroidal shells, product 1
};

struct AgxTargetList {
    int count = 0;
    int needVelocity = 0;    // a kinematic target is present: samples carry velocities
    AgxTargetDev t[AGX_MAX_TARGETS];
};

namespace agxd {

// TargetDensitySphHarm::addPoint (galaxymodel_densitygrid.cpp:371-411): the point's share of
// the two radial nodes around r of every even (l, m) term
__device__ __noinline__ void sphHarmTargetAdd(const AgxTargetDev& T, double* values, double x, double y, double z,
    double mult)
{
    const double R = sqrt(x * x + y * y), phi = atan2A(y, x);
    const double r = sqrt(R * R + z * z), tau = z / (r + R);
    if(r == 0) { values[0] += mult; return; }
    const int nr = T.nr;
    const int indr = binSearchD(r, T.gridr, nr) + 1;
    if(indr >= nr || mult == 0) return;
    const double offr = indr == 0 ? r / T.gridr[0] : (r - T.gridr[indr - 1]) / (T.gridr[indr] - T.gridr[indr - 1]);
    // cos(k phi) by the recurrence of trigMultiAngle (math_sphharm.cpp:134-152), two steps per even m
    double sphi, cphi, sh, ch;
    sincosA(phi, sphi, cphi);
    sincosA(phi / 2, sh, ch);
    const double alpha = 2 * sh * sh, beta = sphi;
    double c1 = 1, s1 = 0;   // cos, sin of k*phi, k = 0
    const double ct = 2 * tau / (1 + tau * tau), st = (1 - tau * tau) / (1 + tau * tau);
    int offset = indr + 1;
    for(int m = 0; m <= T.mmax; m += 2) {
        if(m > 0)
            for(int k = 0; k < 2; k++) {
                const double cn = c1 - (alpha * c1 + beta * s1), sn = s1 - (alpha * s1 - beta * c1);
                c1 = cn; s1 = sn;
            }
        const double angular = m == 0 ? 1. : M_SQRT2 * c1;
        // normalised associated Legendre functions P_l^m(theta), l = m..lmax (sphHarmArray,
        // math_sphharm.cpp:22-132, values only)
        double fact2m = 1, dfact = 1;
        for(int i = 2; i <= 2 * m; i++) fact2m *= i;
        for(int i = 2 * m - 1; i > 1; i -= 2) dfact *= i;
        double prefact = 0.5 / 1.77245385090551602729816748334 * sqrt((2 * m + 1) / fact2m);
        double Pl;
        if(T.lmax == 0) Pl = prefact;
        else if(m == 0) Pl = prefact;
        else {
            const double coef = prefact * dfact;   // (-1)^m = 1 for even m
            double sinm2 = 1, xb = st;
            for(int n = m - 2; n; n >>= 1) { if(n % 2) sinm2 *= xb; xb *= xb; }
            Pl = coef * sinm2 * (st * st);
        }
        double Plm1 = Pl / prefact, Plm = ct * (2 * m + 1) * Plm1, Plm2 = 0;
        for(int l = m; l <= T.lmax; l++) {
            if(l > m) {
                if(l > m + 1) Plm = (ct * (2 * l - 1) * Plm1 - (l + m - 1) * Plm2) / (l - m);
                prefact *= sqrt((2 * l + 1.) / (2 * l - 1.) * (l - m) / (l + m));
                Pl = Plm * prefact;
                Plm2 = Plm1;
                Plm1 = Plm;
            }
            if((l - m) % 2) continue;
            const double val = mult * Pl * 2 * 1.77245385090551602729816748334 * angular;
            values[offset] += val * offr;
            if(indr > 0 || l == 0) values[offset - 1] += val * (1 - offr);
            offset += nr;
        }
    }
}

// TargetDensityCylindrical<N>::addPoint (galaxymodel_densitygrid.cpp:505-545) with the
// corner indexing of getCornerIndicesCylindrical (:92-118); N = 0 top-hat cells, 1 bilinear
__device__ __noinline__ void cylTargetAdd(const AgxTargetDev& T, double* values, double x, double y, double z,
    double mult)
{
    const int N = T.type == AGX_TARGET_DENSITY_CYL_LINEAR ? 1 : 0;
    const double zz = fabs(z), R = sqrt(x * x + y * y), phi = atan2A(y, x);
    const int nR = T.nr, nz = T.nz;
    const int indR = binSearchD(R, T.gridr, nR) + 1, indz = binSearchD(zz, T.gridz, nz) + 1;
    if(indR >= nR || indz >= nz || mult == 0) return;
    const double prevR = indR > 0 ? T.gridr[indR - 1] : 0., prevz = indz > 0 ? T.gridz[indz - 1] : 0.;
    const double offR = (R - prevR) / (T.gridr[indR] - prevR), offz = (zz - prevz) / (T.gridz[indz] - prevz);
    // cos(k phi) by the recurrence of trigMultiAngle (math_sphharm.cpp:134-152), two steps per even m
    double sphi, cphi, sh, ch;
    sincosA(phi, sphi, cphi);
    sincosA(phi / 2, sh, ch);
    const double alpha = 2 * sh * sh, beta = sphi;
    double c1 = 1, s1 = 0;
    for(int m = 0; m <= T.mmax; m += 2) {
        if(m > 0)
            for(int k = 0; k < 2; k++) {
                const double cn = c1 - (alpha * c1 + beta * s1), sn = s1 - (alpha * s1 - beta * c1);
                c1 = cn; s1 = sn;
            }
        const double val = mult * (m == 0 ? 1. : 2 * c1);
        if(N == 0) {
            values[indR + (indz + m / 2 * nz) * nR] += val;
            continue;
        }
        int offsetm = 0, gR = nR, iR = indR;
        if(m > 0) {   // no basis functions at R = 0 for m > 0
            offsetm = (1 + m / 2 * nR) * (nz + 1);
            gR -= 1;
            iR -= 1;
        }
        const int ll = offsetm + iR + indz * (gR + 1), ul = ll + 1, lu = ll + gR + 1, uu = lu + 1;
        values[ul] += val * offR * (1 - offz);
        values[uu] += val * offR * offz;
        if(m == 0 || indR > 0) {
            values[ll] += val * (1 - offR) * (1 - offz);
            values[lu] += val * (1 - offR) * offz;
        }
    }
}

// TargetDensityClassic<N>::addPoint (galaxymodel_densitygrid.cpp:169-239) with the cell and
// corner numbering of getCornerIndicesClassic (:59-90).  The reference throws for a point whose
// pane cannot be determined (a NaN coordinate); such a sample adds nothing here.
__device__ __noinline__ void classicTargetAdd(const AgxTargetDev& T, double* values, double px, double py, double pz,
    double mult)
{
    const int N = T.type == AGX_TARGET_DENSITY_CLASSIC_LINEAR ? 1 : 0, S = T.mmax, numShells = T.nr;
    const int valuesPerShell = 3 * S * (S + N) + N;
    const double X = fabs(px / T.axis[0]), Y = fabs(py) / T.axis[1], Z = fabs(pz) / T.axis[2];
    const double r = sqrt(X * X + Y * Y + Z * Z);
    const int indShell = binSearchD(r, T.gridr, numShells) + 1;
    if(indShell >= numShells || mult == 0) return;
    int pane;
    double c0, c1, c2;
    if(X >= Y && X > Z)      { pane = 0; c0 = X; c1 = Y; c2 = Z; }
    else if(Y >= Z && Y > X) { pane = 1; c0 = Y; c1 = Z; c2 = X; }
    else if(Z >= X && Z >= Y){ pane = 2; c0 = Z; c1 = X; c2 = Y; }
    else return;
    double ratio1 = fmin(atan(c1 / c0) * (4 / M_PI), 1.) * S;
    double ratio2 = fmin(atan(c2 / c0) * (4 / M_PI), 1.) * S;
    const int ind1 = min((int)ratio1, S - 1), ind2 = min((int)ratio2, S - 1);
    if(N == 0) {
        values[ind1 + (ind2 + pane * S) * S + indShell * valuesPerShell] += mult;
        return;
    }
    const int indll = ind1 + (ind2 + pane * S) * (S + 1), indul = indll + 1;
    int indlu, induu;
    if(ind2 < S - 1) {
        indlu = indll + S + 1;
        induu = indlu + 1;
    } else {                       // the upper row lies on the adjacent pane
        const int adjpane = (pane + 2) % 3;
        indlu = (ind1 + adjpane * S + 1) * (S + 1) - 1;
        induu = ind1 < S - 1 ? indlu + S + 1 : 3 * S * (S + 1);   // the node where three panes meet
    }
    ratio1 -= ind1;
    ratio2 -= ind2;
    double val = indShell == 0 ? mult * r / T.gridr[0]
        : mult * (r - T.gridr[indShell - 1]) / (T.gridr[indShell] - T.gridr[indShell - 1]);
    double* off = values + indShell * valuesPerShell + 1;
    off[indll] += (1 - ratio1) * (1 - ratio2) * val;
    off[indul] += ratio1 * (1 - ratio2) * val;
    off[indlu] += (1 - ratio1) * ratio2 * val;
    off[induu] += ratio1 * ratio2 * val;
    if(indShell == 0) {
        values[0] += mult - val;
    } else {
        off -= valuesPerShell;
        val = mult - val;
        off[indll] += (1 - ratio1) * (1 - ratio2) * val;
        off[indul] += ratio1 * (1 - ratio2) * val;
        off[indlu] += (1 - ratio1) * ratio2 * val;
        off[induu] += ratio1 * ratio2 * val;
    }
}

// linInt of math_spline.cpp:35-45: node positions clamped to the grid
__device__ __forceinline__ double linIntD(double x, const double* g, int size, int i1, int i2)
{
    const double x1 = g[i1 < 0 ? 0 : i1 >= size ? size - 1 : i1], x2 = g[i2 < 0 ? 0 : i2 >= size ? size - 1 : i2];
    return x1 == x2 ? (x == x1 ? 1 : 0) : (x - x1) / (x2 - x1);
}

// BsplineInterpolator1d<N>::addPoint (math_spline.cpp:1877-1886) with de Boor's recurrence
// (bsplineValues, :67-82)
__device__ void bsplineAddD(int N, const double* grid, int size, double x, double mult, double* values)
{
    const int ind = binSearchD(x, grid, size);
    if(ind < 0 || ind >= size - 1) return;
    double B[4] = { 0, 0, 0, 0 };
    B[N] = 1;
    for(int l = 1; l <= N; l++) {
        double Bip1 = 0;
        for(int i = N, j = ind; j >= ind - l; i--, j--) {
            const double Bi = B[i] * linIntD(x, grid, size, j, j + l) + Bip1 * linIntD(x, grid, size, j + l + 1, j + 1);
            Bip1 = B[i];
            B[i] = Bi;
        }
    }
    for(int i = 0; i <= N; i++) values[i + ind] += mult * B[i];
}

// TargetKinemShell<N>::addPoint (galaxymodel_losvd.cpp:568-576)
__device__ __noinline__ void kinemShellTargetAdd(const AgxTargetDev& T, double* values, const double* p, double mult)
{
    const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2], r = sqrt(r2);
    const double vr = p[0] * p[3] + p[1] * p[4] + p[2] * p[5];
    const double vr2 = vr * vr / r2, vt2 = p[3] * p[3] + p[4] * p[4] + p[5] * p[5] - vr2;
    const int nb = T.nr + T.lmax - 1;
    bsplineAddD(T.lmax, T.gridr, T.nr, r, mult * vr2, values);
    bsplineAddD(T.lmax, T.gridr, T.nr, r, mult * vt2, values + nb);
}

// p: position, and velocity when the list needs it, in the rotating frame
__device__ __forceinline__ void targetAdd(const AgxTargetDev& T, double* row, const double* p, double mult)
{
    if(T.type == AGX_TARGET_DENSITY_SPHHARM) sphHarmTargetAdd(T, row, p[0], p[1], p[2], mult);
    else if(T.type == AGX_TARGET_KINEM_SHELL) kinemShellTargetAdd(T, row, p, mult);
    else if(T.type >= AGX_TARGET_DENSITY_CLASSIC_TOPHAT) classicTargetAdd(T, row, p[0], p[1], p[2], mult);
    else cylTargetAdd(T, row, p[0], p[1], p[2], mult);
}

}  // namespace agxd
