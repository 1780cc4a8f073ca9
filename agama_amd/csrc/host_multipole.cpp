// Host-side construction of Multipole potentials (see host_multipole.h).
//
// The numerical steps follow the reference construction chain so that the tables agree with
// it to rounding:
//   multipoleFromDensity   src/potential_multipole.cpp:1400-1441 (createMultipole),
//                          :1013-1114 (computePotentialCoefsFromSource), :238-298
//                          (computeSphHarmCoefs), :150-185 (restrictSphHarmCoefs),
//                          src/math_sphharm.cpp:265-384 (Fourier / SH forward transforms)
//   readCoefFile/write     src/potential_factory.cpp:621-641, 725-747, 834-852, 909-929
//   buildTables            src/potential_multipole.cpp:1854-1967 (MultipoleInterp2d ctor),
//                          :1828-1852 (createGridInTau), :553-654 (extrapolation coefs),
//                          src/math_spline.cpp:455-486, 2478-2533 (quintic spline setup)
// Gauss-Legendre nodes are computed (reference's own Halley iteration, math_core.cpp:1237)
// for every order rather than read from 16-digit tables.
#include "host_multipole.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace agx {
namespace {

const double SQRT_DBL_EPS  = 1.4901161193847656e-08;
const double ROOT3_DBL_EPS = 6.0554544523933429e-06;
const double SQRTPI = 1.77245385090551602729816748334;

template<typename T> inline T sq(T x) { return x*x; }
inline bool finite(double x) { return x > -INFINITY && x < INFINITY; }
inline bool symHas(int s, int f) { return (s & f) == f; }

std::string lowerStr(std::string s) { for(auto& ch : s) ch = (char)tolower(ch); return s; }
std::string trimStr(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// integer power by squaring (math_core.cpp:120-133)
double ipow(double x, int n)
{
    if(n < 0) { n = -n; x = 1/x; }
    double r = 1;
    do { if(n % 2) r *= x; n >>= 1; x *= x; } while(n);
    return r;
}

// power with a real exponent and exact special cases (math_core.cpp:135-148)
double powA(double x, double n)
{
    if(n == 0.0) return 1;
    if(n == 1.0) return x;
    if(n ==-1.0) return 1/x;
    if(n == 2.0) return x*x;
    if(n ==-2.0) return 1/(x*x);
    if(n == 0.5) return sqrt(x);
    if(n ==-0.5) return 1/sqrt(x);
    if(n == 3.0) return x*x*x;
    if(n ==-3.0) return 1/(x*x*x);
    return std::pow(x, n);
}

// sin/cos with multiples of pi/2 mapped exactly (math_core.cpp:171-206)
void sincosA(double x, double& s, double& c)
{
    double y = fabs(x);
    long quad = long(4/M_PI * y);
    quad = (quad+1) >> 1;
    y -= M_PI/2 * quad;
    int q13 = quad & 1, q02 = 1 - q13;
    int signc = 1 - (quad & 2), signs = std::signbit(x) ? -signc : signc;
    double y2 = y*y;
    double sy = y + y * (((((+1.5896230157654657e-10 * y2 - 2.5050747762857807e-8) * y2
        + 2.7557313621385725e-6) * y2 - 1.9841269829589539e-4) * y2 + 8.3333333333221186e-3) * y2
        - 0.1666666666666663073) * y2;
    double cy = 1.0 + ((((((-1.1358536521387682e-11 * y2 + 2.0875700841974730e-9) * y2
        - 2.7557314179296740e-7) * y2 + 2.4801587288851704e-5) * y2 - 0.0013888888888873056) * y2
        + 0.0416666666666665950) * y2 - 0.5) * y2;
    s = (q02 * sy + q13 * cy) * signs;
    c = (q02 * cy - q13 * sy) * signc;
}

// n-point Gauss-Legendre rule on [x1,x2] (math_core.cpp:1237-1283, on-the-fly branch)
void glRule(double x1, double x2, int N, double* nodes, double* weights)
{
    if(N == 1) { nodes[0] = 0.5*(x1+x2); weights[0] = x2-x1; return; }
    for(int i = 0; i <= N/2 && i <= N-1-i; i++) {
        double x = (N == i*2+1) ? 0 : (1 - 0.125 * (N-1) / (N*N*N)) * cos(M_PI * (i+0.75) / (N+0.5)), w;
        while(1) {
            double Pm2, Pm1 = x, P = 1.5 * x*x - 0.5;
            for(int l = 3; l <= N; l++) {
                Pm2 = Pm1; Pm1 = P;
                P = x * Pm1 + (1-1./l) * (x * Pm1 - Pm2);
            }
            double der = (Pm1 - x*P) * N / (1-x*x), der2 = (2*x*der - N*(N+1)*P) / (1-x*x),
                dx = -P / (der - 0.5 * P * der2 / der);
            x += dx;
            if(fabs(dx) <= 1e-10) { w = 2 / ((1-x*x) * sq(der + dx*der2)); break; }
        }
        nodes[i]     = 0.5 * (x1 + x2 - x * (x2-x1));
        nodes[N-1-i] = 0.5 * (x1 + x2 + x * (x2-x1));
        weights[i] = weights[N-1-i] = 0.5 * w * (x2-x1);
    }
}

// ---- spherical-harmonic index set (math_sphharm.h:116-157, math_sphharm.cpp:156-230) ----
struct Ind {
    int lmax = 0, mmax = 0, step = 1, sym = ST_SPHERICAL;
    std::vector<int> lminArr;
    Ind() {}
    Ind(int l, int m, int s) : lmax(l), mmax(m), step(symHas(s, ST_ZREFL) || symHas(s, ST_REFL) ? 2 : 1), sym(s)
    {
        if(lmax < 0 || mmax < 0 || mmax > lmax)
            throw std::invalid_argument("SphHarmIndices: incorrect indexing scheme requested");
        if(mmax == 0) sym |= ST_ZROT | ST_XREFL | ST_YREFL;
        if(lmax == 0) sym |= ST_ROT | ST_ZREFL | ST_REFL;
        lminArr.resize(2*mmax+1);
        for(int mm = -mmax; mm <= mmax; mm++) {
            int lm = abs(mm);
            if(symHas(sym, ST_REFL) && mm % 2 != 0) lm = abs(mm) + 1;
            if((symHas(sym, ST_YREFL) && mm < 0) || (symHas(sym, ST_XREFL) && ((mm < 0) ^ (mm % 2 != 0)))
                || (symHas(sym, ST_BISYM) && mm % 2 != 0))
                lm = lmax + 1;
            lminArr[mm + mmax] = lm;
        }
    }
    unsigned size() const { return (lmax+1)*(lmax+1); }
    static unsigned index(int l, int m) { return l*(l+1)+m; }
    static int indexL(unsigned c) { return (int)sqrt((double)c); }
    static int indexM(unsigned c) { int l = indexL(c); return (int)c - l*(l+1); }
    int lmin(int m) const { return m >= -mmax && m <= mmax ? lminArr[m+mmax] : lmax+1; }
    int mmin() const { return symHas(sym, ST_YREFL) ? 0 : -mmax; }
    int thetasize() const { return symHas(sym, ST_ZREFL) ? lmax/2+1 : lmax+1; }
};

Ind indicesFromCoefs(const std::vector<double>& C)
{
    int lmax = (int)sqrt((double)C.size()) - 1;
    if(lmax < 0 || (int)C.size() != sq(lmax+1))
        throw std::invalid_argument("invalid size of the coefficients array");
    int sym = ST_SPHERICAL, mmax = 0;
    for(unsigned c = 0; c < C.size(); c++) {
        if(!finite(C[c])) throw std::domain_error("coefficient not finite");
        if(C[c] != 0) {
            int l = Ind::indexL(c), m = Ind::indexM(c);
            if(l % 2 == 1) sym &= ~ST_REFL;
            if(m < 0) sym &= ~ST_YREFL;
            if((l+m) % 2 == 1) sym &= ~ST_ZREFL;
            if((m < 0) ^ (m % 2 != 0)) sym &= ~ST_XREFL;
            if(m != 0) { sym &= ~ST_ZROT; mmax = std::max(mmax, abs(m)); }
            if(l > 0) sym &= ~ST_ROT;
        }
    }
    return Ind(lmax, mmax, sym);
}

Ind indicesFromCoefs2(const std::vector<std::vector<double> >& A, const std::vector<std::vector<double> >& B)
{
    std::vector<double> D(A.size());
    for(size_t c = 0; c < A.size(); c++)
        for(size_t k = 0; k < A[c].size(); k++)
            D[c] += fabs(A[c][k]) + fabs(B[c][k]);
    return indicesFromCoefs(D);
}

// normalised associated Legendre functions and theta-derivatives (math_sphharm.cpp:22-132)
void legendreArray(int lmax, int m, double tau, double* P, double* dP)
{
    if(lmax == 0) { P[0] = 0.5/SQRTPI; if(dP) dP[0] = 0; return; }
    const double ct = 2*tau / (1 + tau*tau), st = (1 - tau*tau) / (1 + tau*tau);
    double fact2m = 1, dfact = 1;
    for(int i = 2; i <= 2*m; i++) fact2m *= i;
    for(int i = 2*m-1; i > 1; i -= 2) dfact *= i;
    double prefact = 0.5/SQRTPI * sqrt((2*m+1) / fact2m);
    if(m == 0) { P[0] = prefact; if(dP) dP[0] = 0; }
    else if(m == 1) { P[0] = -st * prefact; if(dP) dP[0] = -ct * prefact; }
    else {
        double coef = prefact * dfact * (m % 2 == 1 ? -1 : 1), sinm2 = ipow(st, m-2);
        if(dP) dP[0] = m * coef * sinm2 * st * ct;
        P[0] = coef * sinm2 * sq(st);
    }
    if(lmax == m) return;
    double Plm1 = P[0] / prefact, Plm = ct * (2*m+1) * Plm1, Plm2 = 0;
    for(int l = m+1; l <= lmax; l++) {
        int i = l - m;
        if(l > m+1) Plm = (ct * (2*l-1) * Plm1 - (l+m-1) * Plm2) / (l-m);
        prefact *= sqrt((2*l+1.) / (2*l-1.) * (l-m) / (l+m));
        P[i] = Plm * prefact;
        if(dP) {
            double d = 0;
            if(st >= 1e-8 || (m > 2 && st > 0)) d = (l * ct * Plm - (l+m) * Plm1) / st;
            else if(m == 0) d = -l*(l+1)/2 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
            else if(m == 1) d = -l*(l+1)/2 * (ct > 0 || l % 2 == 0 ? 1 : -1);
            else if(m == 2) d = l*(l+1)*(l+2)*(l-1)/4 * st * (ct > 0 || l % 2 == 1 ? 1 : -1);
            dP[i] = prefact * d;
        }
        Plm2 = Plm1; Plm1 = Plm;
    }
}

// cos(k phi), sin(k phi), k=1..m by the stable recurrence (math_sphharm.cpp:134-152)
void trigMulti(double phi, int m, bool needSine, double* out)
{
    if(m < 1) return;
    double alpha, beta, sinphi, cosphi, s1 = 0, c1 = 1;
    sincosA(phi, sinphi, cosphi);
    sincosA(phi/2, alpha, beta);
    alpha *= alpha*2;
    beta = sinphi;
    for(int k = 0; k < m; k++) {
        cosphi = c1 - (alpha * c1 + beta * s1);
        sinphi = s1 - (alpha * s1 - beta * c1);
        out[k] = cosphi;
        if(needSine) out[k+m] = sinphi;
        c1 = cosphi; s1 = sinphi;
    }
}

// ---- forward spherical-harmonic transform on the GL x uniform-phi grid ----
struct ForwardSH {
    Ind ind;
    bool useSine;
    int nfnc;
    std::vector<double> trigFnc, costh, legFnc;
    ForwardSH(const Ind& I) : ind(I), useSine(I.mmin() < 0)
    {
        const int mmax = ind.mmax, nphi = mmax+1;
        nfnc = useSine ? 2*mmax+1 : mmax+1;
        trigFnc.assign(nphi * nfnc, 0.);
        const double weight = M_PI / (mmax+0.5);
        for(int k = 0; k < nphi; k++) {
            trigFnc[k*nfnc] = 1.;
            if(mmax > 0) trigMulti(phiNode(k), mmax, useSine, &trigFnc[k*nfnc+1]);
            for(int m = 0; m < nfnc; m++)
                trigFnc[k*nfnc+m] *= weight * (useSine || k == 0 ? 1 : 2);
        }
        const int ngrid = ind.lmax+1, ntheta = ind.lmax/2+1, nleg = (ind.lmax+1)*(ind.mmax+1);
        legFnc.assign(ntheta * nleg, 0.);
        costh.resize(ngrid);
        std::vector<double> nodes(ngrid), weights(ngrid);
        glRule(-1, 1, ngrid, &nodes[0], &weights[0]);
        for(int j = 0; j < ngrid; j++) {
            costh[j] = nodes[ngrid-1-j];
            if(j >= ntheta) continue;
            for(int m = 0; m <= ind.mmax; m++) {
                double tau = costh[j] / (sqrt(1 - sq(costh[j])) + 1);
                legendreArray(ind.lmax, m, tau, &legFnc[indLeg(j, m, m)], NULL);
                for(int l = m; l <= ind.lmax; l++)
                    legFnc[indLeg(j, l, m)] *= weights[j] * (0.5/SQRTPI) * (m > 0 ? M_SQRT2 : 1);
            }
        }
    }
    double phiNode(int k) const { return k * M_PI / (ind.mmax + 0.5); }
    unsigned indLeg(int j, int l, int m) const
    {
        int ntheta = ind.lmax/2+1, nleg = (ind.lmax+1)*(ind.mmax+1), am = abs(m);
        int aj = j < ntheta ? j : ind.lmax - j;
        return aj * nleg + am * (ind.lmax+1) + l;
    }
    int size() const { return ind.thetasize() * nfnc; }
    double costheta(int i) const { return costh[i / nfnc]; }
    double phi(int i) const { return phiNode(i % nfnc); }
    void fourier(const double* values, double* coefs, int stride) const
    {
        const int mmax = ind.mmax;
        for(int mm = 0; mm < nfnc; mm++) {
            coefs[mm] = 0;
            int m = useSine ? mm - mmax : mm;
            for(int k = 0; k < nfnc; k++) {
                int iphi = k <= mmax ? k : 2*mmax+1-k;
                int ifn = m >= 0 ? m : mmax - m;
                double fnc = trigFnc[iphi*nfnc + ifn];
                if(m < 0 && k > mmax) fnc *= -1;
                coefs[mm] += fnc * values[k*stride];
            }
        }
    }
    void transform(const double* values, double* coefs, int stride) const
    {
        if(ind.size() == 1) { coefs[0] = values[0]; return; }
        std::fill(coefs, coefs + ind.size(), 0.);
        const int mmin = ind.mmin(), ngrid = ind.lmax+1, nfour = ind.mmax - mmin + 1,
            nsamp = ind.thetasize();
        std::vector<double> valm(nsamp * nfour);
        for(int j = 0; j < nsamp; j++)
            fourier(&values[j * nfnc * stride], &valm[j * nfour], stride);
        for(int m = mmin; m <= ind.mmax; m++)
            for(int l = ind.lmin(m); l <= ind.lmax; l += ind.step) {
                unsigned c = Ind::index(l, m);
                for(int j = 0; j < ngrid; j++) {
                    int js = j < nsamp ? j : ngrid-1-j;
                    double P = legFnc[indLeg(j, l, m)];
                    if((l+m) % 2 == 1 && j > ind.lmax/2) P *= -1;
                    coefs[c] += P * valm[js * nfour + m - mmin];
                }
            }
    }
};

void eliminateNearZeros(std::vector<double>& v, double thr)
{
    double mag = 0;
    for(double x : v) mag = fmax(mag, fabs(x));
    mag *= thr;
    for(double& x : v) if(fabs(x) <= mag) x = 0;
}

// ---- density models ----
struct DehnenDensity : Density {        // src/potential_dehnen.cpp:37-42
    double M, a, gamma, p, q;
    double rhoCyl(double R, double z, double phi) const override
    {
        double s, c; sincosA(phi, s, c);
        double x = R*c, y = R*s;
        double m = sqrt(sq(x) + sq(y/p) + sq(z/q));
        return M * a * (3-gamma) / (4*M_PI*p*q) * powA(m, -gamma) * powA(a+m, gamma-4);
    }
    int symmetry() const override
    { return p == 1 && q == 1 ? ST_SPHERICAL : p == 1 ? ST_AXISYM : ST_TRIAXIAL; }
    void describe(std::vector<DensityComp>& out) const override
    { DensityComp d; d.type = DENS_DEHNEN; d.p[0] = M; d.p[1] = a; d.p[2] = gamma; d.p[3] = p; d.p[4] = q; out.push_back(d); }
};
struct FerrersDensity : Density {       // src/potential_ferrers.cpp:17-31
    double a, b, c, rho0;
    double rhoCyl(double R, double z, double phi) const override
    {
        double s, co; sincosA(phi, s, co);
        double m2 = sq(R*co/a) + sq(R*s/b) + sq(z/c);
        return m2 > 1 ? 0 : rho0 * sq(1-m2);
    }
    int symmetry() const override { return ST_TRIAXIAL; }
    void describe(std::vector<DensityComp>& out) const override
    { DensityComp d; d.type = DENS_FERRERS; d.p[0] = a; d.p[1] = b; d.p[2] = c; d.p[3] = rho0; out.push_back(d); }
};
struct DiskExpDensity : Density {       // src/potential_disk.h:97-139 (exponential in R and |z|)
    double Sigma0, Rd, h;
    double rhoCyl(double R, double z, double) const override
    { return Sigma0 * exp(-R/Rd) * (0.5/h) * exp(-fabs(z)/h); }
    int symmetry() const override { return ST_AXISYM; }
    void describe(std::vector<DensityComp>& out) const override
    { DensityComp d; d.type = DENS_DISK; d.p[0] = Sigma0; d.p[1] = Rd; d.p[2] = h; out.push_back(d); }
};
struct NFWDensity : Density {           // Laplacian of src/potential_analytic.cpp:55-75 / (4 pi)
    double M, rs;
    double rhoCyl(double R, double z, double) const override
    {
        double r = sqrt(R*R + z*z);
        return M / (4*M_PI) / (r * sq(rs + r));
    }
    int symmetry() const override { return ST_SPHERICAL; }
    void describe(std::vector<DensityComp>& out) const override
    { DensityComp d; d.type = DENS_NFW; d.p[0] = M; d.p[1] = rs; out.push_back(d); }
};
struct PlummerDensity : Density {
    double M, a;
    double rhoCyl(double R, double z, double) const override
    { return 3*M / (4*M_PI*a*a*a) * pow(1 + (R*R + z*z) / (a*a), -2.5); }
    int symmetry() const override { return ST_SPHERICAL; }
    void describe(std::vector<DensityComp>& out) const override
    { DensityComp d; d.type = DENS_PLUMMER; d.p[0] = M; d.p[1] = a; out.push_back(d); }
};
struct SumDensity : Density {
    std::vector<std::unique_ptr<Density> > parts;
    double rhoCyl(double R, double z, double phi) const override
    {
        double s = 0;
        for(auto& p : parts) s += p->rhoCyl(R, z, phi);
        return s;
    }
    int symmetry() const override
    {
        int s = ST_SPHERICAL;
        for(auto& p : parts) s &= p->symmetry();
        return s;
    }
    void describe(std::vector<DensityComp>& out) const override { for(auto& p : parts) p->describe(out); }
};

std::unique_ptr<Density> makeOne(const std::string& type, const Params& p, const std::string& pre)
{
    auto num = [&](const char* k, double d) { return p.num(pre + k, d); };
    std::string t = lowerStr(type);
    if(t == "dehnen" || t == "hernquist") {
        auto d = std::unique_ptr<DehnenDensity>(new DehnenDensity);
        d->M = num("mass", 1); d->a = num("scaleradius", 1);
        d->gamma = t == "hernquist" ? 1 : num("gamma", 1);
        d->p = num("axisratioy", 1); d->q = num("axisratioz", 1);
        if(!(d->a > 0) || d->gamma < 0 || d->gamma > 2) throw std::invalid_argument("Dehnen: invalid parameters");
        return std::move(d);
    }
    if(t == "ferrers") {
        auto d = std::unique_ptr<FerrersDensity>(new FerrersDensity);
        double M = num("mass", 1), R = num("scaleradius", 1), pp = num("axisratioy", 1), qq = num("axisratioz", 1);
        if(!(R > 0) || !(1 > pp && pp > qq && qq > 0))
            throw std::invalid_argument("Ferrers: axis ratios must satisfy 0 < q < p < 1");
        d->a = R; d->b = R*pp; d->c = R*qq; d->rho0 = M*105./(32*M_PI*d->a*d->b*d->c);
        return std::move(d);
    }
    if(t == "disk") {
        auto d = std::unique_ptr<DiskExpDensity>(new DiskExpDensity);
        d->Rd = num("scaleradius", 1); d->h = num("scaleheight", 0.1);
        if(!(d->h > 0)) throw std::invalid_argument("Disk: only exponential vertical profiles (scaleHeight>0) are supported");
        d->Sigma0 = p.has(pre + "surfacedensity") ? num("surfacedensity", 0) :
            num("mass", 1) / (2*M_PI*sq(d->Rd));
        return std::move(d);
    }
    if(t == "nfw") {
        auto d = std::unique_ptr<NFWDensity>(new NFWDensity);
        d->M = num("mass", 1); d->rs = num("scaleradius", 1);
        return std::move(d);
    }
    if(t == "plummer") {
        auto d = std::unique_ptr<PlummerDensity>(new PlummerDensity);
        d->M = num("mass", 1); d->a = num("scaleradius", 1);
        return std::move(d);
    }
    throw std::invalid_argument("unknown density type: " + type);
}

// 1-D quintic spline: second derivatives from values and first derivatives with zero 4th
// derivative at both ends (math_spline.cpp:455-486), tridiagonal solve (math_linalg.cpp:289-298)
std::vector<double> quinticSecondDerivs(const std::vector<double>& x, const std::vector<double>& f,
    const std::vector<double>& d)
{
    const size_t n = x.size();
    std::vector<double> lo(n, 0.), di(n, 0.), up(n, 0.), r(n, 0.);
    for(size_t i = 0; i < n; i++) {
        if(i > 0) {
            double hi = 1 / (x[i] - x[i-1]);
            di[i] += hi * 3;
            up[i-1] = -hi;
            r[i] -= (20 * (f[i] - f[i-1]) * hi - 12 * d[i] - 8 * d[i-1]) * hi * hi;
        }
        if(i < n-1) {
            double hi = 1 / (x[i+1] - x[i]);
            di[i] += hi * 3;
            lo[i+1] = -hi;
            r[i] += (20 * (f[i+1] - f[i]) * hi - 12 * d[i] - 8 * d[i+1]) * hi * hi;
        }
    }
    double hi = 1 / (x[1] - x[0]);
    up[0] = -2*hi;
    r[0] = (30 * (f[1] - f[0]) * hi - 14 * d[1] - 16 * d[0]) * hi * hi;
    hi = 1 / (x[n-1] - x[n-2]);
    lo[n-1] = -2*hi;
    r[n-1] = (-30 * (f[n-1] - f[n-2]) * hi + 14 * d[n-2] + 16 * d[n-1]) * hi * hi;
    std::vector<double> U(n);
    for(size_t i = 0; i < n; i++) {
        double ipiv = 1. / (di[i] - (i > 0 ? lo[i] * U[i-1] : 0));
        U[i] = up[i] * ipiv;
        r[i] = (r[i] - (i > 0 ? lo[i] * r[i-1] : 0)) * ipiv;
    }
    for(long i = (long)n-2; i >= 0; i--)
        r[i] -= U[i] * r[i+1];
    return r;
}

// Brent root finder, derivative-free branch (math_core.cpp:895-1040)
template<typename F>
double brentRoot(const F& fnc, double xlower, double xupper, double reltoler)
{
    double a = xlower, b = xupper, fa = fnc(a), fb = fnc(b);
    if((fa < 0 && fb < 0) || (fa > 0 && fb > 0) || !finite(fa+fb)) return NAN;
    if(fa == 0) return fb == 0 ? 0.5*(a+b) : a;
    if(fb == 0) return b;
    double c = a, fc = fa, d = b - c, e = b - c;
    if(fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    int it = 0;
    bool conv = false;
    const double abstol = fabs(xlower - xupper) * reltoler;
    do {
        double tol = 0.5 * DBL_EPSILON * fabs(b), cmb = c - b, m = 0.5 * cmb;
        if(fb == 0 || fabs(m) <= tol) return b;
        if(fabs(e) < tol || fabs(fa) <= fabs(fb)) { d = m; e = m; }
        else {
            double p, q, r, s = fb / fa;
            if(a == c) { p = 2 * m * s; q = 1 - s; }
            else {
                q = fa / fc; r = fb / fc;
                p = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
                q = (q - 1) * (r - 1) * (s - 1);
            }
            if(p > 0) q = -q; else p = -p;
            if(2 * p < fmin(3 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
            else { d = m; e = m; }
        }
        a = b; fa = fb;
        if(fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
        fb = fnc(b);
        if(!finite(fb)) return NAN;
        if((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { c = a; fc = fa; d = b - c; e = b - c; }
        if(fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        it++;
        if(fabs(b - c) <= abstol) {
            conv = true;
            double off = fb * (b - c) / (fc - fb);
            if((off > 0 && off < c - b) || (off < 0 && off > c - b)) b += off;
        }
        if(it >= 64) conv = true;
    } while(!conv);
    return b;
}

// Lambert W on [-1/e,0): Halley iteration from series / asymptotic starting points
// (Corless et al. 1996).  The reference calls GSL's gsl_sf_lambert_W0 / Wm1.
double lambertW(double x, bool wm1)
{
    if(!(x >= -0.36787944117144232159553) || (wm1 && x >= 0)) return NAN;
    if(x == 0) return wm1 ? -INFINITY : 0;
    double w;
    if(x < -0.25) {
        double p = sqrt(fmax(0., 2 * (2.718281828459045235360287 * x + 1)));
        if(p == 0) return -1;
        if(wm1) p = -p;
        w = -1 + p * (1 + p * (-1./3 + p * (11./72 + p * (-43./540))));
    } else if(!wm1) w = x * (1 - x * (1 - 1.5 * x));
    else { double L1 = log(-x), L2 = log(-L1); w = L1 - L2 + L2 / L1; }
    for(int it = 0; it < 50; it++) {
        double ew = exp(w), f = w * ew - x, wp1 = w + 1;
        if(wp1 == 0) break;
        double dw = f / (ew * wp1 - (w + 2) * f / (2 * wp1));
        if(!finite(dw)) break;
        w -= dw;
        if(fabs(dw) <= 4 * DBL_EPSILON * fabs(w)) break;
    }
    return w;
}

// power-law extrapolation coefficients (potential_multipole.cpp:526-614)
void extrapolationCoefs(double Phi1, double Phi2, double dPhi1, double dPhi2, double r1, double r2,
    int v, double s0, double& s, double& U, double& W, double& Q)
{
    if(v == 0) {
        const double coef = (Phi2 - Phi1) / r1, ratio = r2 / r1;
        auto finder = [&](double ss) {
            double rsm1 = pow(ratio, ss-1);
            return coef * (rsm1 - ratio) + (dPhi2 - dPhi1 * rsm1) * 0.5 * (sq(ratio) - 1) +
                (dPhi1 * ratio - dPhi2) * (ratio * rsm1 - 1) / ss;
        };
        s = brentRoot(finder, 2 + 3*SQRT_DBL_EPS, 8.0, SQRT_DBL_EPS);
        double rsm1 = pow(r2 / r1, s-1);
        U = (dPhi2 * r1 - dPhi1 * r2) / (rsm1 - r2 / r1) / s;
        Q = (dPhi1 * rsm1 - dPhi2) * r1 / (rsm1 - r2 / r1) * 0.5;
        double rho0 = 6*Q, rho1 = 6*Q + s * (s+1) * U;
        if(s == s && rho0 * rho1 > 0) { W = Phi1 - U - Q; return; }
        Q = 0;
    }
    double lnr = log(r2/r1);
    double num1 = r1*dPhi1, num2 = v*Phi1, den1 = Phi1, den2 = Phi2 * exp(-v*lnr);
    double A = lnr * (num1 - num2) / (den1 - den2);
    const double SAFE = 100*DBL_EPSILON;
    bool roundoff = fabs(num1-num2) < std::max(fabs(num1), fabs(num2)) * SAFE ||
        fabs(den1-den2) < std::max(fabs(den1), fabs(den2)) * SAFE;
    if(!finite(A) || A >= 0 || roundoff) { U = 0; s = 0; W = Phi1; return; }
    s = fabs(A+1) < SQRT_DBL_EPS ? v : v + (A - lambertW(A * exp(A), A > -1)) / lnr;
    if(v >= 0 && !finite(s)) s = 2;
    if(v < 0 && (!finite(s) || s >= 0)) s = -2;
    if(v > 0) s = fmax(s, s0);
    if(v < -1) s = fmin(s, s0);
    if(s != v) { U = (r1*dPhi1 - v*Phi1) / (s-v); W = (r1*dPhi1 - s*Phi1) / (v-s); }
    else { U = r1*dPhi1 - v*Phi1; W = Phi1; }
}

HostPowerLaw powerLaw(const Coefs& C, bool inner)
{
    const unsigned nc = C.Phi.size();
    HostPowerLaw P;
    P.S.assign(nc, 0.); P.U.assign(nc, 0.); P.W.assign(nc, 0.);
    double Q = 0;
    unsigned i1 = inner ? 0 : C.Phi[0].size()-1, i2 = inner ? 1 : C.Phi[0].size()-2;
    for(unsigned c = 0; c < nc; c++) {
        int l = Ind::indexL(c);
        extrapolationCoefs(C.Phi[c][i1], C.Phi[c][i2], C.dPhi[c][i1], C.dPhi[c][i2],
            C.radii[i1], C.radii[i2], inner ? l : -l-1, P.S[0], P.S[c], P.U[c], P.W[c], Q);
    }
    Ind ind = indicesFromCoefs(P.U);   // potential_multipole.cpp:1589
    P.Q = Q;
    P.r0sq = C.radii[i1] * C.radii[i1];
    P.inner = inner;
    P.lmax = ind.lmax; P.mmax = ind.mmax; P.mmin = ind.mmin(); P.step = ind.step;
    P.lmin.resize(2*ind.mmax+1);
    for(int m = -ind.mmax; m <= ind.mmax; m++) P.lmin[m + ind.mmax] = ind.lmin(m);
    return P;
}

std::vector<double> gridInTau(int lmax)
{
    int nGL = lmax+1;
    std::vector<double> tau(nGL+2), dummy(nGL);
    glRule(-1, 1, nGL, &tau[1], &dummy[0]);
    for(int i = 1; i <= nGL; i++) tau[i] /= 1 + sqrt(1 - sq(tau[i]));
    tau.back() = 1;
    tau.front() = -1;
    const int over = 3, nt = (nGL+1) * over + 1;
    std::vector<double> g(nt);
    for(int i = 0; i <= nGL; i++)
        for(int k = 0; k < over; k++)
            g[i*over + k] = (tau[i] * (over-k) + tau[i+1] * k) / over;
    g.back() = 1;
    return g;
}

}  // namespace

void sincosAgama(double x, double& s, double& c) { sincosA(x, s, c); }

// ---------------------------------------------------------------------------------------
Params Params::parse(const std::string& text)
{
    Params p;
    std::string cur;
    std::vector<std::string> items;
    for(char ch : text) {
        if(ch == ',' || ch == '\n' || ch == ';') { items.push_back(cur); cur.clear(); }
        else cur += ch;
    }
    items.push_back(cur);
    for(auto& it : items) {
        std::string s = trimStr(it);
        if(s.empty() || s[0] == '#' || s[0] == '[') continue;
        size_t eq = s.find('=');
        if(eq == std::string::npos) throw std::invalid_argument("expected key=value, got '" + s + "'");
        p.kv[lowerStr(trimStr(s.substr(0, eq)))] = trimStr(s.substr(eq+1));
    }
    return p;
}
bool Params::has(const std::string& k) const { return kv.count(lowerStr(k)) > 0; }
std::string Params::str(const std::string& k, const std::string& d) const
{ auto it = kv.find(lowerStr(k)); return it == kv.end() ? d : it->second; }
double Params::num(const std::string& k, double d) const
{
    auto it = kv.find(lowerStr(k));
    if(it == kv.end()) return d;
    char* end = NULL;
    double v = strtod(it->second.c_str(), &end);
    if(end == it->second.c_str() || *end != 0) throw std::invalid_argument("parameter " + k + " is not a number");
    return v;
}
void Params::checkOnly(const std::vector<std::string>& allowed) const
{
    for(auto& it : kv) {
        bool ok = false;
        for(auto& a : allowed) {
            std::string al = lowerStr(a);
            if(it.first == al || (al.back() == '*' && it.first.compare(0, al.size()-1, al, 0, al.size()-1) == 0))
                ok = true;
        }
        if(!ok) throw std::invalid_argument("unknown parameter '" + it.first + "'");
    }
}

int symmetryFromName(const std::string& name)   // src/potential_factory.cpp:216-237
{
    if(name.empty()) return ST_UNKNOWN;
    switch(tolower(name[0])) {
        case 's': return ST_SPHERICAL;
        case 'a': return ST_AXISYM;
        case 't': return ST_TRIAXIAL;
        case 'b': return ST_BISYM;
        case 'r': return ST_REFL;
        case 'n': return ST_NONE;
    }
    char* end = NULL;
    long v = strtol(name.c_str(), &end, 10);
    if(end == name.c_str() || *end != 0 || v < 0 || v >= ST_UNKNOWN)
        throw std::runtime_error("Invalid symmetry type: " + name);
    return (int)v;
}

std::string symmetryName(int sym)                // src/potential_factory.cpp:240-252
{
    switch(sym) {
        case ST_UNKNOWN: return "Unknown";
        case ST_NONE: return "None";
        case ST_REFL: return "Reflection";
        case ST_BISYM: return "Bisymmetric";
        case ST_TRIAXIAL: return "Triaxial";
        case ST_AXISYM: return "Axisymmetric";
        case ST_SPHERICAL: return "Spherical";
        default: return std::to_string(sym);
    }
}

int Coefs::lmax() const { return (int)sqrt((double)Phi.size()) - 1; }

std::unique_ptr<Density> makeDensity(const Params& p)
{
    std::string type = p.str("density");
    if(type.find('+') == std::string::npos)
        return makeOne(type, p, "");
    // components "a+b+c": parameters of component i are given with the prefix "i."
    // (e.g. "density=Disk+Dehnen, 0.mass=0.05, 0.scaleRadius=3, 1.mass=0.01, ...")
    auto sum = std::unique_ptr<SumDensity>(new SumDensity);
    std::string cur;
    int idx = 0;
    for(size_t i = 0; i <= type.size(); i++) {
        if(i == type.size() || type[i] == '+') {
            sum->parts.push_back(makeOne(trimStr(cur), p, std::to_string(idx++) + "."));
            cur.clear();
        } else cur += type[i];
    }
    return std::move(sum);
}

Coefs restrictCoefs(Coefs&& C, int lmax, int mmax, int symExp, int gridSizeR);

Coefs multipoleFromDensity(const Density& dens, int lmax, int mmax, int symExp,
    int gridSizeR, double rmin, double rmax)
{
    if(gridSizeR < 2 || !(rmin > 0) || !(rmax > rmin))
        throw std::invalid_argument("Multipole: invalid grid parameters (rmin, rmax must be given, 0<rmin<rmax)");
    if(lmax < 0 || mmax < 0 || mmax > lmax)
        throw std::invalid_argument("Multipole: invalid choice of expansion order");
    int symSrc = dens.symmetry();
    symExp = symExp == ST_UNKNOWN ? symSrc : (symSrc | symExp);
    // createExpGrid (math_spline.cpp:4223-4234)
    std::vector<double> R(gridSizeR);
    {
        double lmin = log(rmin), lmx = log(rmax);
        R.front() = rmin; R.back() = rmax;
        for(int k = 1; k < gridSizeR-1; k++) R[k] = exp(lmin + k*(lmx-lmin)/(gridSizeR-1));
    }
    int lmaxSrc = std::max(lmax + 6, 12), mmaxSrc = std::max(mmax + 6, 12);
    if(symHas(symSrc, ST_SPHERICAL)) lmaxSrc = 0;
    if(symHas(symSrc, ST_ZROT)) mmaxSrc = 0;
    if(symHas(symExp, ST_SPHERICAL)) lmax = 0;
    if(symHas(symExp, ST_ZROT)) mmax = 0;
    Ind ind(lmaxSrc, mmaxSrc, symSrc);

    // computePotentialCoefsFromSource (potential_multipole.cpp:1013-1114)
    const int GLO = 10;
    std::vector<double> gln(GLO), glw(GLO);
    glRule(0, 1, GLO, &gln[0], &glw[0]);
    const size_t nint = (size_t)(gridSizeR+1) * GLO;
    std::vector<double> rint(nint);
    for(int k = 0; k <= gridSizeR; k++) {
        double rkm1 = k > 0 ? R[k-1] : 0, dR = k < gridSizeR ? R[k] - rkm1 : R.back();
        for(int s = 0; s < GLO; s++)
            rint[k*GLO + s] = k < gridSizeR ? rkm1 + gln[s] * dR : R.back() / gln[s];
    }
    // computeSphHarmCoefs for a density (potential_multipole.cpp:238-298)
    ForwardSH tr(ind);
    const int nang = tr.size();
    std::vector<double> values(nint * nang);
    for(size_t ir = 0; ir < nint; ir++)
        for(int ia = 0; ia < nang; ia++) {
            double rad = rint[ir], ct = tr.costheta(ia);
            values[ir*nang + ia] = dens.rhoCyl(rad * sqrt(1 - sq(ct)), rad * ct, tr.phi(ia));
        }
    std::vector<std::vector<double> > ic(ind.size(), std::vector<double>(nint));
    std::vector<double> sh(ind.size());
    for(size_t ir = 0; ir < nint; ir++) {
        tr.transform(&values[ir*nang], &sh[0], 1);
        eliminateNearZeros(sh, 1e-10);
        for(unsigned c = 0; c < ind.size(); c++) ic[c][ir] = sh[c];
    }
    Coefs C;
    C.radii = R;
    C.Phi.assign(ind.size(), std::vector<double>(gridSizeR, 0.));
    C.dPhi.assign(ind.size(), std::vector<double>(gridSizeR, 0.));
    std::vector<double> Pint(gridSizeR), Pext(gridSizeR);
    for(int m = ind.mmin(); m <= ind.mmax; m++)
        for(int l = ind.lmin(m); l <= ind.lmax; l += ind.step) {
            unsigned c = Ind::index(l, m);
            for(int k = 0; k < gridSizeR; k++) {
                double integr = 0;
                for(int s = 0; s < GLO; s++)
                    integr += ic[c][k*GLO + s] * glw[s] * ipow(rint[k*GLO + s] / R[k], l+2);
                Pint[k] = integr * (R[k] - (k > 0 ? R[k-1] : 0)) * R[k] +
                    (k > 0 ? Pint[k-1] * ipow(R[k-1] / R[k], l+1) : 0);
            }
            {
                double integr = 0;
                for(int s = 0; s < GLO; s++)
                    integr += ic[c][gridSizeR*GLO + s] * glw[s] *
                        ipow(rint[gridSizeR*GLO + s] / R[gridSizeR-1], 1-l) / sq(gln[s]);
                Pext[gridSizeR-1] = integr * R[gridSizeR-1] * R[gridSizeR-1];
            }
            for(int k = gridSizeR-2; k >= 0; k--) {
                double integr = 0;
                for(int s = 0; s < GLO; s++)
                    integr += ic[c][(k+1)*GLO + s] * glw[s] * ipow(rint[(k+1)*GLO + s] / R[k], 1-l);
                Pext[k] = integr * (R[k+1] - R[k]) * R[k] + Pext[k+1] * ipow(R[k] / R[k+1], l);
            }
            double mul = -4*M_PI / (2*l+1);
            for(int k = 0; k < gridSizeR; k++) {
                C.Phi[c][k] = mul * (Pint[k] + Pext[k]);
                C.dPhi[c][k] = mul * (-(l+1)*Pint[k] + l*Pext[k]) / R[k];
            }
        }
    return restrictCoefs(std::move(C), lmax, mmax, symExp, gridSizeR);
}

// restrictSphHarmCoefs<2> (potential_multipole.cpp:150-185): terms beyond (lmax, mmax) or not
// allowed by the expansion's symmetry set to zero, trailing empty orders dropped
Coefs restrictCoefs(Coefs&& C, int lmax, int mmax, int symExp, int gridSizeR)
{
    Ind ir(lmax, mmax, symExp);
    std::vector<bool> allowed(ir.size(), false);
    for(int m = ir.mmin(); m <= mmax; m++)
        for(int l = ir.lmin(m); l <= lmax; l += ir.step) allowed[Ind::index(l, m)] = true;
    int lmaxActual = 0;
    for(unsigned c = 0; c < C.Phi.size(); c++) {
        int l = Ind::indexL(c), m = Ind::indexM(c);
        bool zero = true;
        if(l > lmax || abs(m) > mmax || !allowed[c]) {
            C.Phi[c].assign(gridSizeR, 0.);
            C.dPhi[c].assign(gridSizeR, 0.);
        } else {
            for(int k = 0; k < gridSizeR; k++) zero &= C.Phi[c][k] == 0 && C.dPhi[c][k] == 0;
        }
        if(!zero) lmaxActual = std::max(lmaxActual, l);
    }
    if((int)C.Phi.size() > sq(lmaxActual+1)) {
        C.Phi.resize(sq(lmaxActual+1));
        C.dPhi.resize(sq(lmaxActual+1));
    }
    return C;
}

PoissonSetup poissonSetup(const Density& dens, int lmax, int mmax, int symExp,
    int gridSizeR, double rmin, double rmax)
{
    // the same grids, orders and quadrature as multipoleFromDensity
    if(gridSizeR < 2 || !(rmin > 0) || !(rmax > rmin))
        throw std::invalid_argument("Multipole: invalid grid parameters (rmin, rmax must be given, 0<rmin<rmax)");
    if(lmax < 0 || mmax < 0 || mmax > lmax)
        throw std::invalid_argument("Multipole: invalid choice of expansion order");
    PoissonSetup S;
    const int symSrc = dens.symmetry();
    symExp = symExp == ST_UNKNOWN ? symSrc : (symSrc | symExp);
    S.R.resize(gridSizeR);
    {
        double lmin = log(rmin), lmx = log(rmax);
        S.R.front() = rmin; S.R.back() = rmax;
        for(int k = 1; k < gridSizeR-1; k++) S.R[k] = exp(lmin + k*(lmx-lmin)/(gridSizeR-1));
    }
    int lmaxSrc = std::max(lmax + 6, 12), mmaxSrc = std::max(mmax + 6, 12);
    if(symHas(symSrc, ST_SPHERICAL)) lmaxSrc = 0;
    if(symHas(symSrc, ST_ZROT)) mmaxSrc = 0;
    if(symHas(symExp, ST_SPHERICAL)) lmax = 0;
    if(symHas(symExp, ST_ZROT)) mmax = 0;
    S.gridSizeR = gridSizeR; S.lmax = lmax; S.mmax = mmax; S.symExp = symExp;
    Ind ind(lmaxSrc, mmaxSrc, symSrc);
    S.glo = 10;
    S.gln.resize(S.glo); S.glw.resize(S.glo);
    glRule(0, 1, S.glo, &S.gln[0], &S.glw[0]);
    S.nint = (gridSizeR+1) * S.glo;
    S.rint.resize(S.nint);
    for(int k = 0; k <= gridSizeR; k++) {
        double rkm1 = k > 0 ? S.R[k-1] : 0, dR = k < gridSizeR ? S.R[k] - rkm1 : S.R.back();
        for(int s = 0; s < S.glo; s++)
            S.rint[k*S.glo + s] = k < gridSizeR ? rkm1 + S.gln[s] * dR : S.R.back() / S.gln[s];
    }
    ForwardSH tr(ind);
    S.nang = tr.size();
    S.ncoef = (int)ind.size();
    S.cth.resize(S.nang); S.phi.resize(S.nang);
    for(int ia = 0; ia < S.nang; ia++) { S.cth[ia] = tr.costheta(ia); S.phi[ia] = tr.phi(ia); }
    // the transform as a matrix: coefs[c] = sum_{j,k} P(j; l, m) F(m, k) values[j][k], with the
    // Fourier weights F of ForwardSH::fourier and the Legendre weights of ForwardSH::transform
    S.B.assign((size_t)S.nang * S.ncoef, 0.);
    S.active.assign(S.ncoef, 0);
    for(int m = ind.mmin(); m <= ind.mmax; m++)
        for(int l = ind.lmin(m); l <= ind.lmax; l += ind.step) S.active[Ind::index(l, m)] = 1;
    if(ind.size() == 1) {
        S.B[0] = 1;
        return S;
    }
    const int nfnc = tr.nfnc, mmx = ind.mmax, ngrid = ind.lmax+1, nsamp = ind.thetasize();
    std::vector<double> F(nfnc);
    for(int m = ind.mmin(); m <= mmx; m++) {
        for(int k = 0; k < nfnc; k++) {
            const int iphi = k <= mmx ? k : 2*mmx+1-k, ifn = m >= 0 ? m : mmx - m;
            F[k] = tr.trigFnc[iphi*nfnc + ifn] * (m < 0 && k > mmx ? -1 : 1);
        }
        for(int l = ind.lmin(m); l <= ind.lmax; l += ind.step) {
            const unsigned c = Ind::index(l, m);
            for(int j = 0; j < ngrid; j++) {
                const int js = j < nsamp ? j : ngrid-1-j;
                double P = tr.legFnc[tr.indLeg(j, l, m)];
                if((l+m) % 2 == 1 && j > ind.lmax/2) P *= -1;
                for(int k = 0; k < nfnc; k++) S.B[(size_t)(js*nfnc + k) * S.ncoef + c] += P * F[k];
            }
        }
    }
    return S;
}

void poissonStagesHost(const PoissonSetup& S, const Density& dens, std::vector<double>& Phi,
    std::vector<double>& dPhi)
{
    const int nint = S.nint, nang = S.nang, nc = S.ncoef, nr = S.gridSizeR, G = S.glo;
    std::vector<double> ic((size_t)nc * nint), vals(nang), sh(nc);
    for(int ir = 0; ir < nint; ir++) {
        const double rad = S.rint[ir];
        for(int ia = 0; ia < nang; ia++)
            vals[ia] = dens.rhoCyl(rad * sqrt(1 - sq(S.cth[ia])), rad * S.cth[ia], S.phi[ia]);
        double mag = 0;
        for(int c = 0; c < nc; c++) {
            double acc = 0;
            for(int ia = 0; ia < nang; ia++) acc += vals[ia] * S.B[(size_t)ia * nc + c];
            sh[c] = acc;
            mag = fmax(mag, fabs(acc));
        }
        for(int c = 0; c < nc; c++) ic[(size_t)c * nint + ir] = fabs(sh[c]) <= mag * 1e-10 ? 0 : sh[c];
    }
    Phi.assign((size_t)nc * nr, 0.);
    dPhi.assign((size_t)nc * nr, 0.);
    std::vector<double> Pint(nr), Pext(nr);
    for(int c = 0; c < nc; c++) {
        if(!S.active[c]) continue;
        const int l = Ind::indexL(c);
        const double* q = &ic[(size_t)c * nint];
        for(int k = 0; k < nr; k++) {
            double integr = 0;
            for(int s = 0; s < G; s++) integr += q[k*G + s] * S.glw[s] * ipow(S.rint[k*G + s] / S.R[k], l+2);
            Pint[k] = integr * (S.R[k] - (k > 0 ? S.R[k-1] : 0)) * S.R[k] +
                (k > 0 ? Pint[k-1] * ipow(S.R[k-1] / S.R[k], l+1) : 0);
        }
        {
            double integr = 0;
            for(int s = 0; s < G; s++)
                integr += q[nr*G + s] * S.glw[s] * ipow(S.rint[nr*G + s] / S.R[nr-1], 1-l) / sq(S.gln[s]);
            Pext[nr-1] = integr * S.R[nr-1] * S.R[nr-1];
        }
        for(int k = nr-2; k >= 0; k--) {
            double integr = 0;
            for(int s = 0; s < G; s++) integr += q[(k+1)*G + s] * S.glw[s] * ipow(S.rint[(k+1)*G + s] / S.R[k], 1-l);
            Pext[k] = integr * (S.R[k+1] - S.R[k]) * S.R[k] + Pext[k+1] * ipow(S.R[k] / S.R[k+1], l);
        }
        const double mul = -4*M_PI / (2*l+1);
        for(int k = 0; k < nr; k++) {
            Phi[(size_t)c * nr + k] = mul * (Pint[k] + Pext[k]);
            dPhi[(size_t)c * nr + k] = mul * (-(l+1)*Pint[k] + l*Pext[k]) / S.R[k];
        }
    }
}

Coefs poissonFinish(const PoissonSetup& S, const std::vector<double>& Phi, const std::vector<double>& dPhi)
{
    Coefs C;
    C.radii = S.R;
    C.Phi.assign(S.ncoef, std::vector<double>(S.gridSizeR, 0.));
    C.dPhi.assign(S.ncoef, std::vector<double>(S.gridSizeR, 0.));
    for(int c = 0; c < S.ncoef; c++)
        for(int k = 0; k < S.gridSizeR; k++) {
            C.Phi[c][k] = Phi[(size_t)c * S.gridSizeR + k];
            C.dPhi[c][k] = dPhi[(size_t)c * S.gridSizeR + k];
        }
    return restrictCoefs(std::move(C), S.lmax, S.mmax, S.symExp, S.gridSizeR);
}

Coefs readCoefFile(const std::string& path)
{
    std::ifstream in(path);
    if(!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    int nr = -1, lmax = -1;
    bool inCoefs = false;
    std::vector<std::string> lines;
    while(std::getline(in, line)) {
        std::string t = trimStr(line);
        if(!inCoefs) {
            if(t == "Coefficients") { inCoefs = true; continue; }
            size_t eq = t.find('=');
            if(eq == std::string::npos) continue;
            std::string k = lowerStr(trimStr(t.substr(0, eq))), v = trimStr(t.substr(eq+1));
            if(k == "gridsizer") nr = atoi(v.c_str());
            else if(k == "lmax") lmax = atoi(v.c_str());
            else if(k == "type" && lowerStr(v) != "multipole")
                throw std::runtime_error("coefficient file is not of type Multipole");
        } else lines.push_back(t);
    }
    bool ok = nr >= 2 && lmax >= 0 && (int)lines.size() >= 2*nr+5 && lines[0] == "#Phi" &&
        lines[nr+3] == "#dPhi/dr" && lines[1].size() > 1 && lines[1][0] == '#' &&
        lines[nr+4].size() > 1 && lines[nr+4][0] == '#';
    if(!ok) throw std::runtime_error("Error loading Multipole potential from " + path);
    const unsigned nc = sq(lmax+1);
    Coefs C;
    C.radii.resize(nr);
    C.Phi.assign(nc, std::vector<double>(nr));
    C.dPhi.assign(nc, std::vector<double>(nr));
    for(int blk = 0; blk < 2; blk++) {
        int start = blk == 0 ? 2 : nr+5;
        for(int n = 0; n < nr; n++) {
            std::istringstream is(lines[start+n]);
            std::vector<double> f;
            std::string tok;
            while(is >> tok) f.push_back(strtod(tok.c_str(), NULL));
            if(f.size() != nc+1) throw std::runtime_error("Error loading Multipole potential: bad row");
            C.radii[n] = f[0];
            for(unsigned c = 0; c < nc; c++) (blk == 0 ? C.Phi : C.dPhi)[c][n] = f[c+1];
        }
    }
    return C;
}

void writeCoefFile(const std::string& path, const Coefs& C)
{
    Ind ind = indicesFromCoefs2(C.Phi, C.dPhi);
    std::ofstream out(path);
    if(!out) throw std::runtime_error("cannot write " + path);
    char buf[64];
    out << "[Potential]\ntype=Multipole\n";
    out << "gridSizeR=" << C.radii.size() << "\n";
    out << "lmax=" << C.lmax() << "\n";
    out << "symmetry=" << symmetryName(ind.sym) << "\n";
    out << "Coefficients\n";
    for(int blk = 0; blk < 2; blk++) {
        out << (blk == 0 ? "#Phi\n" : "\n#dPhi/dr\n") << "#radius";
        for(int l = 0; l <= C.lmax(); l++)
            for(int m = -l; m <= l; m++) out << "\tl=" << l << ",m=" << m;
        out << "\n";
        for(size_t n = 0; n < C.radii.size(); n++) {
            snprintf(buf, sizeof(buf), "%.17g", C.radii[n]);
            out << buf;
            for(size_t c = 0; c < C.Phi.size(); c++) {
                double v = (blk == 0 ? C.Phi : C.dPhi)[c][n];
                if(v == 0) out << "\t0";
                else { snprintf(buf, sizeof(buf), "\t%.17g", v); out << buf; }
            }
            out << "\n";
        }
    }
}

Coefs truncateCoefs(const Coefs& C, int lmax)
{
    Coefs t = C;
    const int L = C.lmax();
    for(int l = lmax + 1; l <= L; l++)
        for(int m = -l; m <= l; m++) {
            std::fill(t.Phi[l * (l + 1) + m].begin(), t.Phi[l * (l + 1) + m].end(), 0.);
            std::fill(t.dPhi[l * (l + 1) + m].begin(), t.dPhi[l * (l + 1) + m].end(), 0.);
        }
    return t;
}

// 1-D quintic Hermite interpolation on [0, h] as a polynomial in u: a[0..5] from
// d = (f(0), f(h), f'(0), f'(h), f''(0), f''(h)).  Expanding the weights of the Hermite form
// (P = 10t^3 - 15t^4 + 6t^5, Q/h = t^3 - t^4/2, R/h^2 = t^3 - 1.25t^4 + t^5/2 with t = u/h, applied
// to fd = f(h) - f(0) - h/2 (f'(0) + f'(h)), f1d = f'(h) - f'(0) - h/2 (f''(0) + f''(h)) and
// f2d = f''(h) - f''(0), besides f(0) + f'(0) u + f''(0) u^2/2) gives the coefficients below.
// The differences are formed first, as in the Hermite form: values of one magnitude that vary
// little across a cell cancel exactly instead of after scaling by 1/h^3.
static void hermiteToPoly(long double h, const long double d[6], long double a[6])
{
    const long double fd = (d[1] - d[0]) - h / 2 * (d[2] + d[3]),
        f1d = (d[3] - d[2]) - h / 2 * (d[4] + d[5]), f2d = d[5] - d[4];
    const long double h2 = h * h, h3 = h2 * h;
    a[0] = d[0];
    a[1] = d[2];
    a[2] = d[4] / 2;
    a[3] = (10 * fd + h * f1d + h2 * f2d) / h3;
    a[4] = (-15 * fd - 0.5L * h * f1d - 1.25L * h2 * f2d) / (h3 * h);
    a[5] = (6 * fd + 0.5L * h2 * f2d) / (h3 * h2);
}

std::vector<double> buildCellPolys(const HostTables& T)
{
    if(T.interp1d || T.nr < 2 || T.nt < 2) return {};
    const int nr = T.nr, nt = T.nt, ns = T.ns;
    std::vector<double> cells((size_t)(nr - 1) * (nt - 1) * ns * AGX_CELL_COEFS);
    // record index q of the derivative d^px/dx^px d^py/dy^py (agx_tables.h)
    static const int Q[3][3] = { { 0, 2, 5 }, { 1, 4, 7 }, { 3, 6, 8 } };
    for(int i = 0; i < nr - 1; i++) {
        const long double hx = (long double)T.gridR[i + 1] - T.gridR[i];
        for(int j = 0; j < nt - 1; j++) {
            const long double hy = (long double)T.gridT[j + 1] - T.gridT[j];
            for(int s = 0; s < ns; s++) {
                // the tau pass for each x-datum alpha = 2 px + ex (f, f_x, f_xx at x_i, x_i+1),
                // as QuinticSpline2d's first pass; then the ln r pass on each tau coefficient
                long double Y[6][6], d[6], a[6];
                for(int al = 0; al < 6; al++) {
                    const int px = al / 2, ex = al % 2;
                    for(int be = 0; be < 6; be++) {
                        const int py = be / 2, ey = be % 2;
                        const size_t node = (size_t)(i + ex) * nt + (j + ey);
                        d[be] = T.nodes[(node * ns + s) * AGX_NODE_STRIDE + Q[px][py]];
                    }
                    hermiteToPoly(hy, d, Y[al]);
                }
                double* dst = &cells[(((size_t)i * (nt - 1) + j) * ns + s) * AGX_CELL_COEFS];
                for(int b = 0; b < 6; b++) {
                    for(int al = 0; al < 6; al++) d[al] = Y[al][b];
                    hermiteToPoly(hx, d, a);
                    for(int k = 0; k < 6; k++) dst[k * 6 + b] = (double)a[k];
                }
            }
        }
    }
    return cells;
}

std::vector<double> buildLegendreTable(int lmax)
{
    const int L1 = lmax + 1;
    std::vector<double> t((size_t)L1 * L1 * AGX_LEG_STRIDE + L1, 0.);
    const long double pi = 3.14159265358979323846264338327950288L;
    for(int m = 0; m <= lmax; m++) {
        // sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), by its closed form at l = m and the ratio in l
        long double fact2m = 1, dfact = 1;
        for(int i = 2; i <= 2 * m; i++) fact2m *= i;
        for(int i = 2 * m - 1; i > 1; i -= 2) dfact *= i;
        long double pf = 0.5L / sqrtl(pi) * sqrtl((2 * m + 1) / fact2m);
        const long double mul = m == 0 ? 2 * sqrtl(pi) : 2 * sqrtl(pi) * sqrtl(2.0L);
        for(int l = m; l <= lmax; l++) {
            if(l > m) pf *= sqrtl((2 * l + 1.0L) / (2 * l - 1.0L) * (l - m) / (l + m));
            double* e = &t[((size_t)m * L1 + l) * AGX_LEG_STRIDE];
            e[0] = (double)(pf * mul);
            e[1] = l > m ? (double)((2 * l - 1.0L) / (l - m)) : 0.;
            e[2] = l > m ? (double)((l + m - 1.0L) / (l - m)) : 0.;
            e[3] = l;
        }
        t[(size_t)L1 * L1 * AGX_LEG_STRIDE + m] = (double)(dfact * (m % 2 ? -1 : 1));
    }
    return t;
}

std::vector<int> buildTauLookup(const HostTables& T)
{
    if(T.interp1d || T.nt < 2) return {};
    double minw = INFINITY;
    for(int j = 0; j + 1 < T.nt; j++) minw = std::min(minw, T.gridT[j + 1] - T.gridT[j]);
    int nb = 64;
    while(nb < 4096 && 2.0 / nb >= minw) nb *= 2;
    std::vector<int> lut(nb);
    for(int b = 0; b < nb; b++) {
        const double t = -1 + 2.0 * b / nb;
        int j = 0;
        while(j < T.nt - 2 && t >= T.gridT[j + 1]) j++;
        lut[b] = j;
    }
    return lut;
}

// MultipoleInterp1d ctor (potential_multipole.cpp:1691-1738): the extrapolation's 1/Phi_00(0)
// from the two innermost nodes, log-scaling when every Phi_00(r_k) < 0, and a quintic spline in
// ln r (values and ln r derivatives at the nodes, second derivatives by constructQuinticSpline,
// math_spline.cpp:1376-1389) for each (l,m) term of the index set
static HostTables buildTables1d(const Coefs& C, const Ind& ind)
{
    const int nr = (int)C.radii.size();
    HostTables T;
    T.interp1d = 1;
    T.nr = nr; T.nt = 0; T.ns = 0;
    T.lmax = ind.lmax; T.mmax = ind.mmax; T.mmin = ind.mmin(); T.step = ind.step; T.sym = ind.sym;
    double s, U, W, Q;
    extrapolationCoefs(C.Phi[0][0], C.Phi[0][1], C.dPhi[0][0], C.dPhi[0][1], C.radii[0], C.radii[1],
        0, NAN, s, U, W, Q);
    T.invPhi0 = s > 0 ? 1./W : 0;
    bool logScaling = true;
    T.gridR.resize(nr);
    for(int k = 0; k < nr; k++) {
        T.gridR[k] = log(C.radii[k]);
        logScaling &= C.Phi[0][k] < 0;
        if(C.Phi[0][k] * T.invPhi0 >= 1) T.invPhi0 = 0;
    }
    T.logScaling = logScaling;
    T.nc1d = (ind.lmax + 1) * (ind.lmax + 1);
    T.nodes1d.assign((size_t)nr * T.nc1d * 4, 0.);
    T.term1d.assign(T.nc1d, 0);
    std::vector<double> f(nr), df(nr);
    for(int m = ind.mmin(); m <= ind.mmax; m++)
        for(int l = ind.lmin(m); l <= ind.lmax; l += ind.step) {
            const unsigned c = Ind::index(l, m);
            T.term1d[c] = 1;
            for(int k = 0; k < nr; k++) {
                if(logScaling) {
                    if(c == 0) {
                        f[k]  = log(T.invPhi0 - 1 / C.Phi[c][k]);
                        df[k] = C.radii[k] * C.dPhi[c][k] / (C.Phi[c][k] * (T.invPhi0 * C.Phi[c][k] - 1));
                    } else {
                        f[k]  = C.Phi[c][k] / C.Phi[0][k];
                        df[k] = (C.dPhi[c][k] - f[k] * C.dPhi[0][k]) * C.radii[k] / C.Phi[0][k];
                    }
                } else {
                    f[k] = C.Phi[c][k];
                    df[k] = C.radii[k] * C.dPhi[c][k];
                }
                if(!finite(f[k]) || !finite(df[k]))
                    throw std::invalid_argument("QuinticSpline: non-finite input");
            }
            const std::vector<double> d2 = quinticSecondDerivs(T.gridR, f, df);
            for(int k = 0; k < nr; k++) {
                double* dst = &T.nodes1d[((size_t)k * T.nc1d + c) * 4];
                dst[0] = f[k]; dst[1] = df[k]; dst[2] = d2[k];
            }
        }
    T.lmax1d = ind.lmax; T.mmax1d = ind.mmax; T.mmin1d = ind.mmin(); T.step1d = ind.step;
    T.lmin1d.resize(2 * ind.mmax + 1);
    for(int m = -ind.mmax; m <= ind.mmax; m++) T.lmin1d[m + ind.mmax] = ind.lmin(m);
    T.rsqInner = sq(C.radii.front() * (1 + ROOT3_DBL_EPS));
    T.rsqOuter = sq(C.radii.back() * (1 - ROOT3_DBL_EPS));
    T.inner = powerLaw(C, true);
    T.outer = powerLaw(C, false);
    return T;
}

HostTables buildTables(const Coefs& C, bool allowLowOrder)
{
    const int nr = (int)C.radii.size();
    Ind ind = indicesFromCoefs2(C.Phi, C.dPhi);
    bool correct = nr >= 2 && C.Phi.size() == ind.size() && C.dPhi.size() == ind.size();
    for(unsigned c = 0; correct && c < ind.size(); c++)
        correct &= (int)C.Phi[c].size() == nr && (int)C.dPhi[c].size() == nr;
    for(int k = 1; correct && k < nr; k++) correct &= C.radii[k] > C.radii[k-1];
    if(!correct) throw std::invalid_argument("Multipole: invalid radial grid");
    if(ind.mmax > AGX_MAX_M) throw std::invalid_argument("Multipole: mmax > 16 is not supported");
    if(ind.lmax <= 2 && !allowLowOrder) return buildTables1d(C, ind);

    HostTables T;
    T.nr = nr; T.lmax = ind.lmax; T.mmax = ind.mmax; T.mmin = ind.mmin(); T.step = ind.step; T.sym = ind.sym;
    // MultipoleInterp2d ctor (potential_multipole.cpp:1854-1967)
    double s, U, W, Q;
    extrapolationCoefs(C.Phi[0][0], C.Phi[0][1], C.dPhi[0][0], C.dPhi[0][1], C.radii[0], C.radii[1],
        0, NAN, s, U, W, Q);
    T.invPhi0 = s > 0 ? 1./W : 0;
    T.gridR.resize(nr);
    for(int k = 0; k < nr; k++) T.gridR[k] = log(C.radii[k]);
    T.gridT = gridInTau(ind.lmax);
    const int nt = T.nt = (int)T.gridT.size();
    const size_t n2 = (size_t)nr * nt;
    // slots in the order of the evaluation loop (m ascending from mmin); construction order
    // processes m = 0 first, as the reference does
    for(int m = ind.mmin(); m <= ind.mmax; m++)
        if(ind.lmin(m) <= ind.lmax) T.slotM.push_back(m);
    T.ns = (int)T.slotM.size();
    for(int si = 0; si < T.ns; si++) if(T.slotM[si] == 0) T.zeroSlot = si;
    T.nodes.assign(n2 * T.ns * AGX_NODE_STRIDE, 0.);
    std::vector<double> Pv(n2), PdR(n2), PdT(n2), PdRdT(n2), P0v, P0dR, P0dT, P0dRdT;
    std::vector<double> Plm(ind.lmax+1), dPlm(ind.lmax+1);
    bool logScaling = true;
    for(int mm = 0; mm <= ind.mmax - ind.mmin(); mm++) {
        int m = mm <= ind.mmax ? mm : ind.mmax - mm;
        int lmin = ind.lmin(m);
        if(lmin > ind.lmax) continue;
        int am = abs(m);
        double mul = m == 0 ? 2*SQRTPI : 2*SQRTPI*M_SQRT2;
        for(int j = 0; j < nt; j++) {
            legendreArray(ind.lmax, am, T.gridT[j], &Plm[0], &dPlm[0]);
            for(int k = 0; k < nr; k++) {
                double val = 0, dR = 0, dT = 0, dRdT = 0;
                for(int l = lmin; l <= ind.lmax; l += ind.step) {
                    unsigned c = Ind::index(l, m);
                    val  +=  C.Phi[c][k] *  Plm[l-am];
                    dR   += C.dPhi[c][k] *  Plm[l-am];
                    dT   +=  C.Phi[c][k] * dPlm[l-am];
                    dRdT += C.dPhi[c][k] * dPlm[l-am];
                }
                if(m == 0) {
                    logScaling &= val < 0;
                    if(T.invPhi0 <= 1 / (mul * val)) T.invPhi0 = 0;
                }
                size_t i = (size_t)k * nt + j;
                Pv[i]    = mul * val;
                PdR[i]   = mul * dR * C.radii[k];
                PdT[i]   = mul * dT * -2 / (sq(T.gridT[j]) + 1);
                PdRdT[i] = mul * dRdT * C.radii[k] * -2 / (sq(T.gridT[j]) + 1);
            }
        }
        if(m == 0) {
            P0v = Pv; P0dR = PdR; P0dT = PdT; P0dRdT = PdRdT;
            if(logScaling)
                for(size_t i = 0; i < n2; i++) {
                    double dsc = 1 / (P0v[i] * (T.invPhi0 * P0v[i] - 1));
                    PdR[i] *= dsc;
                    PdT[i] *= dsc;
                    PdRdT[i] = dsc * PdRdT[i] + (1 - 2 * T.invPhi0 * P0v[i]) * PdR[i] * PdT[i];
                    Pv[i] = log(T.invPhi0 - 1 / P0v[i]);
                }
        } else if(logScaling) {
            for(size_t i = 0; i < n2; i++) {
                double rel = Pv[i] / P0v[i];
                Pv[i] = rel;
                PdR[i] = (PdR[i] - rel * P0dR[i]) / P0v[i];
                PdT[i] = (PdT[i] - rel * P0dT[i]) / P0v[i];
                PdRdT[i] = (PdRdT[i] - rel * P0dRdT[i] - PdR[i] * P0dT[i] - PdT[i] * P0dR[i]) / P0v[i];
            }
        }
        for(size_t i = 0; i < n2; i++)
            if(!finite(Pv[i]) || !finite(PdR[i]) || !finite(PdT[i]) || !finite(PdRdT[i]))
                throw std::invalid_argument("QuinticSpline2d: non-finite input");
        // QuinticSpline2d::setupWithMixedDeriv (math_spline.cpp:2478-2533)
        std::vector<double> fxx(n2), fyy(n2), fxxy(n2), fxyy(n2), fxxyy(n2);
        {
            std::vector<double> t(nr), tx(nr), ty(nr), txy(nr);
            for(int j = 0; j < nt; j++) {
                for(int i = 0; i < nr; i++) {
                    t[i] = Pv[i*nt+j]; tx[i] = PdR[i*nt+j]; ty[i] = PdT[i*nt+j]; txy[i] = PdRdT[i*nt+j];
                }
                std::vector<double> a = quinticSecondDerivs(T.gridR, t, tx), b = quinticSecondDerivs(T.gridR, ty, txy);
                for(int i = 0; i < nr; i++) { fxx[i*nt+j] = a[i]; fxxy[i*nt+j] = b[i]; }
            }
            std::vector<double> u(nt), ux(nt), uy(nt), uxx(nt), uxy(nt), uxxy(nt);
            for(int i = 0; i < nr; i++) {
                for(int j = 0; j < nt; j++) {
                    u[j] = Pv[i*nt+j]; ux[j] = PdR[i*nt+j]; uy[j] = PdT[i*nt+j];
                    uxx[j] = fxx[i*nt+j]; uxy[j] = PdRdT[i*nt+j]; uxxy[j] = fxxy[i*nt+j];
                }
                std::vector<double> a = quinticSecondDerivs(T.gridT, u, uy),
                    b = quinticSecondDerivs(T.gridT, ux, uxy), c = quinticSecondDerivs(T.gridT, uxx, uxxy);
                for(int j = 0; j < nt; j++) { fyy[i*nt+j] = a[j]; fxyy[i*nt+j] = b[j]; fxxyy[i*nt+j] = c[j]; }
            }
        }
        int slot = (int)(std::find(T.slotM.begin(), T.slotM.end(), m) - T.slotM.begin());
        for(size_t i = 0; i < n2; i++) {
            double* dst = &T.nodes[(i * T.ns + slot) * AGX_NODE_STRIDE];
            dst[0] = Pv[i]; dst[1] = PdR[i]; dst[2] = PdT[i]; dst[3] = fxx[i]; dst[4] = PdRdT[i];
            dst[5] = fyy[i]; dst[6] = fxxy[i]; dst[7] = fxyy[i]; dst[8] = fxxyy[i];
        }
    }
    T.logScaling = logScaling;
    T.rsqInner = sq(C.radii.front() * (1 + ROOT3_DBL_EPS));
    T.rsqOuter = sq(C.radii.back() * (1 - ROOT3_DBL_EPS));
    T.inner = powerLaw(C, true);
    T.outer = powerLaw(C, false);
    return T;
}

}  // namespace agx
