// Second derivatives of the Multipole potential on the device (FP64): BasePotential::eval with
// deriv2 for the Multipole path, restating
//   QuinticSpline2d::evalDeriv z_xx, z_xy, z_yy        src/math_spline.cpp:2535-2597
//   evalQuinticSplines (second-derivative weights)     src/math_spline.cpp:653-718
//   MultipoleInterp2d::evalCyl with hess               src/potential_multipole.cpp:1969-2080
//   PowerLawMultipole::evalCyl with hess               src/potential_multipole.cpp:1600-1658
//   sphHarmTransformInverseDeriv[2], fourierTransformAzimuth, transformDerivsSphToCyl
//                                                      src/potential_multipole.cpp:666-848
//   sphHarmArray / legendrePmm second derivatives      src/math_sphharm.cpp:22-132
//   toPosDeriv<Car,Cyl>, toHess<Cyl,Car>, toHess<Sph,Cyl>, evalAndConvertSph<Car>
//                                                      src/coord.cpp:94-125, 700-720, 779-799, 923-956
//   NFW::evalDeriv deriv2                              src/potential_analytic.cpp:55-75
//   Composite (evalComponents<Cyl,Sph>)                src/potential_composite.cpp:51-163
// The value and gradient parts repeat the arithmetic of agx_device.h; this path is a separate
// set of functions so that the orbit kernels' register budget is not affected.
// Hessian layouts: Cyl {dR2, dz2, dphi2, dRdz, dzdphi, dRdphi}; Car {dx2, dy2, dz2, dxdy, dydz, dxdz}.
#pragma once
#include "agx_device.h"

namespace agxd {

// quintic Hermite weights with the second derivative (evalQuinticSplines, d2f output)
struct QuinticW2 {
    bool atEnd;
    double dx, dx2, h, P, Q, Rr, Pp, Qp, Rp, Ppp, Qpp, Rpp;
    __device__ __forceinline__ void init(double x, double xl, double xh)
    {
        atEnd = x == xh;
        dx = x - xl;
        dx2 = .5 * dx * dx;
        h = xh - xl;
        const double hi = 1 / h, h2 = h * h, t = dx * hi, t2 = t * t, t3 = t * t2, t1t = t * (1 - t),
            Px = 30 * t1t * hi;
        P  = t3 * (10 - t * (15 - 6 * t));
        Q  = t3 * (1 - t * 0.5) * h;
        Rr = t3 * (1 - t * (1.25 - 0.5 * t)) * h2;
        Pp = Px * t1t;
        Qp = t2 * (3 - t * 2);
        Rp = t2 * (3 - t * (5 - 2.5 * t)) * h;
        Ppp = Px * hi * (2 - 4 * t);
        Qpp = Px * 0.2;
        Rpp = t * (6 - t * (15 - 10 * t));
    }
    // value, first and second derivative of one function given f, f', f'' at both ends
    __device__ __forceinline__ void eval3(double fl, double fh, double f1l, double f1h, double f2l,
        double f2h, double& f, double& df, double& d2f) const
    {
        if(atEnd) { f = fh; df = f1h; d2f = f2h; return; }
        const double fd = fh - fl - 0.5 * h * (f1h + f1l), f1d = f1h - f1l - 0.5 * h * (f2h + f2l),
            f2d = f2h - f2l;
        f   = fl + P * fd + dx * f1l + Q * f1d + dx2 * f2l + Rr * f2d;
        df  = Pp * fd + f1l + Qp * f1d + dx * f2l + Rp * f2d;
        d2f = Ppp * fd + Qpp * f1d + f2l + Rpp * f2d;
    }
};

// value, first and second (ln r, tau) derivatives of one harmonic slot from its four corner
// records v (ll, lu, ul, uu), in the record order f, f_x, f_y, f_xx, f_xy, f_yy, f_xxy, f_xyy, f_xxyy
__device__ __forceinline__ void splineSlotHess(const double (&v)[4][10], int offsetCorner,
    const QuinticW2& wy, const QuinticW2& wx, double& z, double& zx, double& zy, double& zxx,
    double& zxy, double& zyy)
{
    const double* ll = v[0]; const double* lu = v[1]; const double* ul = v[2]; const double* uu = v[3];
    const double off = offsetCorner == 0 ? ll[0] : offsetCorner == 1 ? lu[0] : offsetCorner == 2 ? ul[0] : uu[0];
    // y-pass: the six intermediate functions {f(xl), f(xu), fx(xl), fx(xu), fxx(xl), fxx(xu)}
    // with their first (dF) and second (d2F) y-derivatives
    double F[6], dF[6], d2F[6];
    wy.eval3(ll[0] - off, lu[0] - off, ll[2], lu[2], ll[5], lu[5], F[0], dF[0], d2F[0]);
    wy.eval3(ul[0] - off, uu[0] - off, ul[2], uu[2], ul[5], uu[5], F[1], dF[1], d2F[1]);
    wy.eval3(ll[1], lu[1], ll[4], lu[4], ll[7], lu[7], F[2], dF[2], d2F[2]);
    wy.eval3(ul[1], uu[1], ul[4], uu[4], ul[7], uu[7], F[3], dF[3], d2F[3]);
    wy.eval3(ll[3], lu[3], ll[6], lu[6], ll[8], lu[8], F[4], dF[4], d2F[4]);
    wy.eval3(ul[3], uu[3], ul[6], uu[6], ul[8], uu[8], F[5], dF[5], d2F[5]);
    // x-pass
    wx.eval3(F[0], F[1], F[2], F[3], F[4], F[5], z, zx, zxx);
    double unused;
    wx.eval3(dF[0], dF[1], dF[2], dF[3], dF[4], dF[5], zy, zxy, unused);
    double u1, u2;
    wx.eval3(d2F[0], d2F[1], d2F[2], d2F[3], d2F[4], d2F[5], zyy, u1, u2);
    z += off;
}

// toHess<Sph,Cyl> (coord.cpp:779-799) for source derivatives g = {dr, dtheta, dphi},
// h = {dr2, dtheta2, dphi2, drdtheta, drdphi, dthetadphi}; out: Cyl layout
__device__ __forceinline__ void sphToCylHess(const double (&g)[3], const double (&h)[6], double drdR,
    double drdz, double dthdR, double dthdz, double d2rdR2, double d2rdRdz, double d2rdz2, double d2thdR2,
    double d2thdRdz, double d2thdz2, double (&o)[6])
{
    const double dr2 = h[0], dth2 = h[1], dph2 = h[2], drdth = h[3], drdph = h[4], dthdph = h[5];
    o[0] = (dr2 * drdR + drdth * dthdR) * drdR + (drdth * drdR + dth2 * dthdR) * dthdR +
        g[0] * d2rdR2 + g[1] * d2thdR2;
    o[3] = (dr2 * drdz + drdth * dthdz) * drdR + (drdth * drdz + dth2 * dthdz) * dthdR +
        g[0] * d2rdRdz + g[1] * d2thdRdz;
    o[1] = (dr2 * drdz + drdth * dthdz) * drdz + (drdth * drdz + dth2 * dthdz) * dthdz +
        g[0] * d2rdz2 + g[1] * d2thdz2;
    o[5] = drdph * drdR + dthdph * dthdR;
    o[4] = drdph * drdz + dthdph * dthdz;
    o[2] = dph2;
}

// interior (spline) region: potential, Cyl gradient {dR, dz, dphi} and Cyl Hessian
template<class F>
__device__ __forceinline__ void interiorEvalHess(const AgxMultipoleDev& P, const F& fetch, double R,
    double z, double phi, double& pot, double (&gc)[3], double (&hc)[6])
{
    const double AGX_GLOBAL* gridR = gmem(P.gridR);
    const double AGX_GLOBAL* gridT = gmem(P.gridT);
    const double r = sqrt(R * R + z * z), logr = log(r), rplusRinv = 1. / (r + fabs(R)),
        tau = R == 0 ? (z > 0 ? 1. : z < 0 ? -1. : 0.) : z * rplusRinv;
    const int xi = binSearchD(logr, gridR, P.nr), yi = binSearchD(tau, gridT, P.nt);
    if(!(xi >= 0 && xi < P.nr - 1 && yi >= 0 && yi < P.nt - 1)) {
        pot = NAN;
        for(int i = 0; i < 3; i++) gc[i] = NAN;
        for(int i = 0; i < 6; i++) hc[i] = NAN;
        return;
    }
    const double xlow = gridR[xi], xupp = gridR[xi + 1], ylow = gridT[yi], yupp = gridT[yi + 1];
    const int offsetCorner = logr == xupp ? (tau == yupp ? 3 : 2) : (tau == yupp ? 1 : 0);
    QuinticW2 wy, wx;
    wy.init(tau, ylow, yupp);
    wx.init(logr, xlow, xupp);
    const int ns = P.ns, nt = P.nt, rec0 = (xi * nt + yi) * ns, zs = P.zeroSlot;
    double v[4][10];
    double P0, d0r, d0t, d0rr, d0rt, d0tt;
    fetchSlot(fetch, rec0 + zs, ns, nt, v);
    splineSlotHess(v, offsetCorner, wy, wx, P0, d0r, d0t, d0rr, d0rt, d0tt);
    if(P.logScaling) {
        const double expX = exp(P0), val = 1 / (P.invPhi0 - expX), dPhidX = val * val * expX,
            d2PhidX2 = dPhidX * val * (P.invPhi0 + expX);
        d0rr = dPhidX * d0rr + d2PhidX2 * d0r * d0r;
        d0tt = dPhidX * d0tt + d2PhidX2 * d0t * d0t;
        d0rt = dPhidX * d0rt + d2PhidX2 * d0r * d0t;
        d0r *= dPhidX;
        d0t *= dPhidX;
        P0 = val;
    }
    // Fourier synthesis (fourierTransformAzimuth with hess) in the reference's order
    double tp = P0, g[3] = { d0r, d0t, 0 }, h[6] = { d0rr, d0tt, 0, d0rt, 0, 0 };
    TrigRecT<double> pos, neg;
    pos.init(phi);
    neg.alpha = pos.alpha; neg.beta = pos.beta;
    for(int s = 0; s < ns; s++) {
        if(s == zs) continue;
        const int m = P.slotM[s];
        double trig, dtrig;
        if(m > 0) { pos.advanceTo(m); trig = pos.c; dtrig = -m * pos.s; }
        else { neg.restart(); neg.advanceTo(-m); trig = neg.s; dtrig = -m * neg.c; }
        const double d2trig = -m * m * trig;
        double Pm, dmr, dmt, dmrr, dmrt, dmtt;
        fetchSlot(fetch, rec0 + s, ns, nt, v);
        splineSlotHess(v, offsetCorner, wy, wx, Pm, dmr, dmt, dmrr, dmrt, dmtt);
        if(P.logScaling) {
            dmrr = dmrr * P0 + Pm * d0rr + 2 * dmr * d0r;
            dmtt = dmtt * P0 + Pm * d0tt + 2 * dmt * d0t;
            dmrt = dmrt * P0 + Pm * d0rt + dmr * d0t + dmt * d0r;
            dmr = dmr * P0 + Pm * d0r;
            dmt = dmt * P0 + Pm * d0t;
            Pm *= P0;
        }
        tp   += Pm * trig;
        g[0] += dmr * trig;
        g[1] += dmt * trig;
        g[2] += Pm * dtrig;
        h[0] += dmrr * trig;
        h[3] += dmrt * trig;
        h[1] += dmtt * trig;
        h[4] += dmr * dtrig;
        h[5] += dmt * dtrig;
        h[2] += Pm * d2trig;
    }
    pot = tp;
    const double rinv = 1 / r, r2inv = rinv * rinv;
    const double drdR = R * r2inv, drdz = z * r2inv, dthdR = -tau * rinv, dthdz = rinv - rplusRinv;
    gc[0] = g[0] * drdR + g[1] * dthdR;
    gc[1] = g[0] * drdz + g[1] * dthdz;
    gc[2] = g[2];
    const double d2rdR2 = drdz * drdz - drdR * drdR, d2rdRdz = -2 * drdR * drdz, d2rdz2 = -d2rdR2,
        d2thdR2 = z * r2inv * rinv, d2thdRdz = dthdR * dthdR - drdR * drdR * r * rplusRinv,
        d2thdz2 = -d2thdR2 - dthdR * rplusRinv;
    sphToCylHess(g, h, drdR, drdz, dthdR, dthdz, d2rdR2, d2rdRdz, d2rdz2, d2thdR2, d2thdRdz, d2thdz2, hc);
}

// PowerLawMultipole::evalCyl with the Hessian
// transformDerivsSphToCyl with the Hessian (potential_multipole.cpp:824-848): (ln r, theta, phi)
// gradient g and Hessian h {dr2, dth2, dph2, drdth, drdph, dthdph} -> Cyl layout
__device__ __forceinline__ void sphToCylDerivs(double R, double z, const double (&g)[3], const double (&h)[6],
    double (&gc)[3], double (&hc)[6])
{
    // transformDerivsSphToCyl with the Hessian
    const double r2inv = 1 / (R * R + z * z), drdR = R * r2inv, drdz = z * r2inv, dthdR = drdz, dthdz = -drdR;
    gc[0] = g[0] * drdR + g[1] * dthdR;
    gc[1] = g[0] * drdz + g[1] * dthdz;
    gc[2] = g[2];
    const double d2rdR2 = drdz * drdz - drdR * drdR, d2rdRdz = -2 * drdR * drdz, d2rdz2 = -d2rdR2;
    sphToCylHess(g, h, drdR, drdz, dthdR, dthdz, d2rdR2, d2rdRdz, d2rdz2, d2rdRdz, -d2rdR2, -d2rdRdz, hc);
}


// sphHarmTransformInverseDeriv[2] with the Hessian (potential_multipole.cpp:722-821) and
// transformDerivsSphToCyl (:824-848) for the (l, m) terms of the index set I; term(l, m, C, dC, d2C)
// gives each term and its first two ln r derivatives; lmax >= 1
template<class Ind, class Term>
__device__ __forceinline__ void sphHarmSynthesisHess(const Ind& I, double R, double z, double phi, const Term& term,
    double& pot, double (&gc)[3], double (&hc)[6])
{
    const double tau = z == 0 ? 0 : z / (sqrt(R * R + z * z) + R);
    double vr, g[3] = { 0, 0, 0 }, h[6] = { 0, 0, 0, 0, 0, 0 };   // Sph: {dr2, dth2, dph2, drdth, drdph, dthdph}
    if(I.lmax == 2 && I.mmin == 0 && I.step == 2) {
        const double C2 = sqrt(1.25), D2 = sqrt(3.75);
        const double ct = 2 * tau / (1 + tau * tau), st = (1 - tau * tau) / (1 + tau * tau),
            cc = ct * ct, cs = ct * st, ss = st * st, Y20 = (3 * C2 * cc - C2), dY20 = -6 * C2 * cs,
            d2Y20 = -12 * C2 * cc + 6 * C2;
        double P00, d00, dd00, P20, d20, dd20;
        term(0, 0, P00, d00, dd00);
        term(2, 0, P20, d20, dd20);
        vr = Y20 * P20 + P00;
        g[0] = Y20 * d20 + d00;
        g[1] = dY20 * P20;
        h[0] = Y20 * dd20 + dd00;
        h[3] = dY20 * d20;
        h[1] = d2Y20 * P20;
        if(I.mmax == 2) {
            double sp, cp, P22, d22, dd22;
            term(2, 2, P22, d22, dd22);
            const double Y22 = D2 * ss, dY22 = 2 * D2 * cs, d2Y22 = 4 * D2 * cc - 2 * D2;
            sincosA(2 * phi, sp, cp);
            vr   += Y22 * P22 * cp;
            g[0] += Y22 * d22 * cp;
            g[1] += dY22 * P22 * cp;
            g[2] += Y22 * P22 * -2 * sp;
            h[0] += Y22 * dd22 * cp;
            h[3] += dY22 * d22 * cp;
            h[1] += d2Y22 * P22 * cp;
            h[4] += Y22 * d22 * -2 * sp;
            h[5] += dY22 * P22 * -2 * sp;
            h[2] += Y22 * P22 * -4 * cp;
        }
    } else {
        const double ct = 2 * tau / (1 + tau * tau), st = (1 - tau * tau) / (1 + tau * tau);
        const double EPS = 1e-8;
        TrigRecT<double> pos, neg;
        pos.init(phi);
        neg.alpha = pos.alpha; neg.beta = pos.beta;
        vr = 0;
        for(int pass = 0; pass < 2; pass++)
        for(int m = (pass == 0 ? 0 : I.mmin); m <= (pass == 0 ? 0 : I.mmax); m++) {
            if(pass == 1 && m == 0) continue;
            const int lmin = I.lmin[m + I.mmax];
            if(lmin > I.lmax) continue;
            const double mul = m == 0 ? 2 * 1.77245385090551602729816748334 :
                2 * 1.77245385090551602729816748334 * M_SQRT2;
            double Cv = 0, Cr = 0, Ct = 0, Crr = 0, Crt = 0, Ctt = 0;
            const int am = m < 0 ? -m : m;
            // legendrePmm with the second derivative
            double fact2m = 1, dfact = 1;
            for(int i = 2; i <= 2 * am; i++) fact2m *= i;
            for(int i = 2 * am - 1; i > 1; i -= 2) dfact *= i;
            double prefact = 0.5 / 1.77245385090551602729816748334 * sqrt((2 * am + 1) / fact2m);
            double Pl, dPl, d2Pl;
            if(am == 0) { Pl = prefact; dPl = 0; d2Pl = 0; }
            else if(am == 1) { Pl = -st * prefact; dPl = -ct * prefact; d2Pl = st * prefact; }
            else {
                const double coef = prefact * dfact * (am % 2 == 1 ? -1 : 1);
                double sinm2 = 1, xb = st;
                for(int n = am - 2; n; n >>= 1) { if(n % 2) sinm2 *= xb; xb *= xb; }
                dPl = am * coef * sinm2 * st * ct;
                d2Pl = am * coef * sinm2 * (am * (ct * ct) - 1);
                Pl = coef * sinm2 * (st * st);
            }
            double Plm1 = Pl / prefact, Plm = ct * (2 * am + 1) * Plm1, Plm2 = 0;
            double d2Plm1 = st, d2Plm2 = 0, d2Plm = 12 * ct * st;   // m = 1 pole recurrence
            for(int l = am; l <= I.lmax; l++) {
                if(l > am) {
                    if(l > am + 1) Plm = (ct * (2 * l - 1) * Plm1 - (l + am - 1) * Plm2) / (l - am);
                    prefact *= sqrt((2 * l + 1.) / (2 * l - 1.) * (l - am) / (l + am));
                    Pl = Plm * prefact;
                    double d = 0;
                    if(st >= EPS || (am > 2 && st > 0)) d = (l * ct * Plm - (l + am) * Plm1) / st;
                    else if(am == 0) d = -l * (l + 1) / 2 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    else if(am == 1) d = -l * (l + 1) / 2 * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else if(am == 2) d = l * (l + 1) * (l + 2) * (l - 1) / 4 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
                    dPl = prefact * d;
                    if(st >= EPS || (am > 2 && st > 0))
                        d2Pl = ct * dPl / (-st) - (l * (l + 1) - (am / st) * (am / st)) * Pl;
                    else if(am == 0)
                        d2Pl = -l * (l + 1) / 2 * prefact * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else if(am == 1) {
                        if(l > am + 1) {
                            const double twodPlm1 = -l * (l - 1) * (ct > 0 || l % 2 == 1 ? 1 : -1);
                            d2Plm = ((2 * l - 1) * (ct * (d2Plm1 - Plm1) - st * twodPlm1) - l * d2Plm2) / (l - 1);
                        }
                        d2Pl = prefact * d2Plm;
                        d2Plm2 = d2Plm1;
                        d2Plm1 = d2Plm;
                    }
                    else if(am == 2)
                        d2Pl = l * (l + 1) * (l + 2) * (l - 1) / 4 * prefact * (ct > 0 || l % 2 == 0 ? 1 : -1);
                    else
                        d2Pl = 0;
                    Plm2 = Plm1;
                    Plm1 = Plm;
                }
                if(l >= lmin && (l - lmin) % I.step == 0) {
                    double Pv, dPv, d2Pv;
                    term(l, m, Pv, dPv, d2Pv);
                    Cv  += Pl * Pv * mul;
                    Cr  += Pl * dPv * mul;
                    Ct  += dPl * Pv * mul;
                    Crr += Pl * d2Pv * mul;
                    Crt += dPl * dPv * mul;
                    Ctt += d2Pl * Pv * mul;
                }
            }
            if(m == 0) {
                vr = Cv; g[0] = Cr; g[1] = Ct; g[2] = 0;
                h[0] = Crr; h[3] = Crt; h[1] = Ctt; h[4] = h[5] = h[2] = 0;
                continue;
            }
            double trig, dtrig;
            if(m > 0) { pos.advanceTo(m); trig = pos.c; dtrig = -m * pos.s; }
            else { neg.restart(); neg.advanceTo(-m); trig = neg.s; dtrig = -m * neg.c; }
            const double d2trig = -m * m * trig;
            vr   += Cv * trig;
            g[0] += Cr * trig;
            g[1] += Ct * trig;
            g[2] += Cv * dtrig;
            h[0] += Crr * trig;
            h[3] += Crt * trig;
            h[1] += Ctt * trig;
            h[4] += Cr * dtrig;
            h[5] += Ct * dtrig;
            h[2] += Cv * d2trig;
        }
    }
    pot = vr;
    sphToCylDerivs(R, z, g, h, gc, hc);
}

__device__ __forceinline__ void powerLawEvalHess(const AgxPowerLawDev& L, double R, double z, double phi,
    double& pot, double (&gc)[3], double (&hc)[6])
{
    const double rsq = R * R + z * z, dlogr = log(rsq / L.r0sq) * 0.5;
    const int lmax = (L.inner && rsq < L.r0sq * 1e-16) || (!L.inner && rsq > L.r0sq * 1e16) ? 0 : L.lmax;
    auto term = [&](int l, int m, double& Pv, double& dPv, double& d2Pv) {
        const int c = l * (l + 1) + m;
        const double s = gmem(L.S)[c], u = gmem(L.U)[c], w = gmem(L.W)[c], v = L.inner ? l : -l - 1;
        const double rv  = v != 0 ? exp(dlogr * v) : 1;
        const double rs  = s != v ? (s != 0 ? exp(dlogr * s) : 1) : rv;
        const double urs = u * rs * (s != v || u == 0 ? 1 : dlogr);
        const double wrv = w * rv;
        const double qr2 = v == 0 ? L.Q * rsq / L.r0sq : 0;
        Pv   = urs + wrv + qr2;
        dPv  = urs * s + wrv * v + (s != v ? 0 : u * rs) + qr2 * 2;
        d2Pv = urs * s * s + wrv * v * v + (s != v ? 0 : 2 * s * u * rs) + qr2 * 4;
    };
    if(lmax == 0) {
        double P0, dP0, d2P0;
        term(0, 0, P0, dP0, d2P0);
        pot = P0;
        const double rsqinv = rsq > 0 ? 1 / rsq : 0, Rr2 = rsq < INFINITY ? R * rsqinv : 0,
            zr2 = rsq < INFINITY ? z * rsqinv : 0;
        gc[0] = dP0 * Rr2; gc[1] = dP0 * zr2; gc[2] = 0;
        const double d2 = d2P0 - 2 * dP0;
        hc[0] = d2 * (Rr2 * Rr2) + dP0 * rsqinv;
        hc[1] = d2 * (zr2 * zr2) + dP0 * rsqinv;
        hc[3] = d2 * Rr2 * zr2;
        hc[2] = hc[4] = hc[5] = 0;
        return;
    }
    sphHarmSynthesisHess(L, R, z, phi, term, pot, gc, hc);
}

// MultipoleInterp1d::evalCyl with the Hessian (potential_multipole.cpp:1753-1815): 1-d quintic
// splines in ln r with second derivatives, the l=0 term log-unscaled, the others multiplied by it
__device__ __forceinline__ void interior1dEvalHess(const AgxMultipoleDev& P, double R, double z, double phi,
    double& pot, double (&gc)[3], double (&hc)[6])
{
    const double AGX_GLOBAL* gridR = gmem(P.gridR);
    const double r = sqrt(R * R + z * z), logr = log(r);
    int k = binSearchD(logr, gridR, P.nr);
    k = k < 0 ? 0 : k > P.nr - 2 ? P.nr - 2 : k;
    QuinticW2 w;
    w.init(logr, gridR[k], gridR[k + 1]);
    const vd2 AGX_GLOBAL* nodes = gmem(reinterpret_cast<const vd2*>(P.nodes1d));
    const int nc = P.nc1d;
    auto spline = [&](int c, double& f, double& df, double& d2f) {
        const int lo = (k * nc + c) * 2, hi = ((k + 1) * nc + c) * 2;
        const vd2 a0 = nodes[lo], a1 = nodes[lo + 1], b0 = nodes[hi], b1 = nodes[hi + 1];
        w.eval3(a0.x, b0.x, a0.y, b0.y, a1.x, b1.x, f, df, d2f);
    };
    double P0, dP0, d2P0;
    spline(0, P0, dP0, d2P0);
    if(P.logScaling) {
        const double expX = exp(P0), Phi = 1 / (P.invPhi0 - expX), dPhidX = Phi * Phi * expX;
        P0 = Phi;
        d2P0 = dPhidX * (d2P0 + dP0 * dP0 * Phi * (P.invPhi0 + expX));
        dP0 *= dPhidX;
    }
    if(P.ind1d.lmax == 0) {
        const double g[3] = { dP0, 0, 0 }, h[6] = { d2P0, 0, 0, 0, 0, 0 };
        pot = P0;
        sphToCylDerivs(R, z, g, h, gc, hc);
        return;
    }
    auto term = [&](int l, int m, double& Pv, double& dPv, double& d2Pv) {
        const int c = l * (l + 1) + m;
        if(c == 0) { Pv = P0; dPv = dP0; d2Pv = d2P0; return; }
        spline(c, Pv, dPv, d2Pv);
        if(P.logScaling) {
            d2Pv = d2Pv * P0 + 2 * dPv * dP0 + Pv * d2P0;
            dPv = dPv * P0 + Pv * dP0;
            Pv *= P0;
        }
    };
    sphHarmSynthesisHess(P.ind1d, R, z, phi, term, pot, gc, hc);
}

// NFW: potential, radial first and second derivatives
__device__ __forceinline__ void nfwDeriv2(double mass, double a, double r, double& pot, double& der, double& der2)
{
    nfwDeriv(mass, a, r, pot, der);
    const double q = r / a;
    const double lnr = r == INFINITY ? 0. : q > 0.016 ? log(1 + q) / r :
        (1 + q * (1 + 11. / 60 * q)) / (1 + q * (1.5 + q * (0.6 + q * 0.05))) / a;
    der2 = -mass * (q > 0.010 ? (2 * lnr - (2 * a + 3 * r) / ((a + r) * (a + r))) / (r * r) :
        1 / (1.5 + q * (27. / 8 + q * (351. / 160 + 183. / 640 * q))) / (a * a * a));
}

// BasePotential::eval(PosCyl) with gradient {dR, dz, dphi} and Hessian (Cyl layout) for the
// Multipole (+ optional NFW, evalAndConvertSph<Cyl>, coord.cpp:958-990): the components are
// summed in order, the Multipole first (Composite::evalCyl, potential_composite.cpp:315-320)
template<class F>
__device__ __forceinline__ void multipoleCylHess(const AgxMultipoleDev& P, const F& fetch, double R,
    double z, double phi, double& pot, double (&gc)[3], double (&hc)[6])
{
    const double rsq = R * R + z * z;
    if(rsq < P.rsqInner) powerLawEvalHess(P.inner, R, z, phi, pot, gc, hc);
    else if(rsq > P.rsqOuter) powerLawEvalHess(P.outer, R, z, phi, pot, gc, hc);
    else if(P.interp1d) interior1dEvalHess(P, R, z, phi, pot, gc, hc);
    else interiorEvalHess(P, fetch, R, z, phi, pot, gc, hc);
    if(P.nfwMass != 0) {
        const double r = sqrt(R * R + z * z);
        double pn, dn, d2n;
        nfwDeriv2(P.nfwMass, P.nfwScale, r, pn, dn, d2n);
        const double Rr = r == 0 ? 0. : R / r, zr = r == 0 ? 0. : z / r;
        double dor = dn / r, dd = d2n - dor;
        if(r == 0) { dd = 0; if(dn == 0) dor = d2n; }
        pot = pot + pn;
        gc[0] = gc[0] + Rr * dn;
        gc[1] = gc[1] + zr * dn;
        hc[0] = hc[0] + (Rr * Rr * dd + dor);
        hc[1] = hc[1] + (zr * zr * dd + dor);
        hc[3] = hc[3] + Rr * zr * dd;
    }
}

// BasePotential::eval(PosCar) with deriv2 for the Multipole (+ optional NFW) potential:
// gradient g = +dPhi/dx and Hessian h in the Car layout
template<class F>
__device__ __forceinline__ void multipoleCarHess(const AgxMultipoleDev& P, const F& fetch, double x,
    double y, double z, double& pot, double (&g)[3], double (&h)[6])
{
    // toPosDeriv<Car,Cyl> with second derivatives
    const double R2 = x * x + y * y, R = sqrt(R2);
    double dRdx, dRdy, dphidx, dphidy, d2Rdx2, d2Rdy2, d2Rdxdy, d2phidx2, d2phidy2, d2phidxdy, phi;
    if(R == 0) {
        dRdx = dRdy = dphidx = dphidy = 1.;
        d2Rdx2 = d2Rdy2 = d2Rdxdy = d2phidx2 = d2phidy2 = d2phidxdy = 1.;
        phi = 0;
    } else {
        const double cosphi = x / R, sinphi = y / R;
        dRdx = cosphi; dRdy = sinphi; dphidx = -sinphi / R; dphidy = cosphi / R;
        d2Rdx2 = sinphi * sinphi / R;
        d2Rdy2 = cosphi * cosphi / R;
        d2Rdxdy = -sinphi * cosphi / R;
        d2phidx2 = 2 * sinphi * cosphi / R2;
        d2phidy2 = -d2phidx2;
        d2phidxdy = (sinphi * sinphi - cosphi * cosphi) / R2;
        phi = atan2A(y, x);
    }
    const double rsq = R * R + z * z;
    double gc[3], hc[6];
    if(rsq < P.rsqInner) powerLawEvalHess(P.inner, R, z, phi, pot, gc, hc);
    else if(rsq > P.rsqOuter) powerLawEvalHess(P.outer, R, z, phi, pot, gc, hc);
    else if(P.interp1d) interior1dEvalHess(P, R, z, phi, pot, gc, hc);
    else interiorEvalHess(P, fetch, R, z, phi, pot, gc, hc);
    // toGrad<Cyl,Car>, toHess<Cyl,Car>
    const double dR2 = hc[0], dz2 = hc[1], dphi2 = hc[2], dRdz = hc[3], dzdphi = hc[4], dRdphi = hc[5];
    double gx = gc[0] * dRdx + gc[2] * dphidx, gy = gc[0] * dRdy + gc[2] * dphidy, gz = gc[1];
    double hxx = (dR2 * dRdx + dRdphi * dphidx) * dRdx + (dRdphi * dRdx + dphi2 * dphidx) * dphidx +
        gc[0] * d2Rdx2 + gc[2] * d2phidx2;
    double hxy = (dR2 * dRdy + dRdphi * dphidy) * dRdx + (dRdphi * dRdy + dphi2 * dphidy) * dphidx +
        gc[0] * d2Rdxdy + gc[2] * d2phidxdy;
    double hyy = (dR2 * dRdy + dRdphi * dphidy) * dRdy + (dRdphi * dRdy + dphi2 * dphidy) * dphidy +
        gc[0] * d2Rdy2 + gc[2] * d2phidy2;
    double hxz = dRdz * dRdx + dzdphi * dphidx, hyz = dRdz * dRdy + dzdphi * dphidy, hzz = dz2;
    if(P.nfwMass != 0) {
        // the spherical component is added first, the cylindrical one's converted derivatives
        // after it (potential_composite.cpp:134-163); evalAndConvertSph<Car> for the halo
        const double r = sqrt(x * x + y * y + z * z);
        double pn, dn, d2n;
        nfwDeriv2(P.nfwMass, P.nfwScale, r, pn, dn, d2n);
        const double xr = r == 0 ? 0. : x / r, yr = r == 0 ? 0. : y / r, zr = r == 0 ? 0. : z / r;
        double dor = dn / r, dd = d2n - dor;
        if(r == 0) { dd = 0; if(dn == 0) dor = d2n; }
        pot = pot + pn;
        gx = xr * dn + gx;
        gy = yr * dn + gy;
        gz = zr * dn + gz;
        hxx = (xr * xr * dd + dor) + hxx;
        hyy = (yr * yr * dd + dor) + hyy;
        hzz = (zr * zr * dd + dor) + hzz;
        hxy = xr * yr * dd + hxy;
        hyz = yr * zr * dd + hyz;
        hxz = xr * zr * dd + hxz;
    }
    g[0] = gx; g[1] = gy; g[2] = gz;
    h[0] = hxx; h[1] = hyy; h[2] = hzz; h[3] = hxy; h[4] = hyz; h[5] = hxz;
}

}  // namespace agxd
