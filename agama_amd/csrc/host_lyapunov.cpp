// Lyapunov exponent fit on the host (see host_lyapunov.h).  The reference fits with GSL's
// lmsder (src/math_fit.cpp:258-315), a MINPACK Levenberg-Marquardt, and stops as soon as one
// iteration changes every parameter by at most 1e-6 relative.  GSL is not part of this build:
// the solver below is a Levenberg-Marquardt of its own (Marquardt's scaled damping with
// Nielsen's update of the damping factor) that iterates to the least-squares minimum (every
// parameter settled to 1e-12 relative, or no further decrease of the sum of squares), so that
// its result does not depend on the size of its first steps.  Non-finite residuals are
// replaced by 1e10 as the reference's wrapper does (math_fit.cpp:94-117).
#include "host_lyapunov.h"

#include <cmath>
#include <vector>

namespace agx {

namespace {

inline double sq(double x) { return x * x; }

// F(x) = ln[(exp(x)-1)/x] and its derivative without overflow (orbit_variational.cpp:312-324)
inline double logexpm1overx(double x)
{
    return fabs(x) < 1e-2 ? x * (0.5 + x * (1./24 - (1./2880) * x*x)) :
        x < 33 ? log(expm1(x) / x) : x - log(x);
}
inline double derlogexpm1overx(double x)
{
    return fabs(x) < 1e-2 ? 0.5 + x * (1./12 - (1./720) * x*x) : -1/x - 1/expm1(-x);
}

struct FitData { const double* t; const double* lnw; int n; };

// the fitting function of orbit_variational.cpp:326-356: residuals A + ln t + [t >= tch]
// (F(lambda t) - F(lambda tch)) - ln w and their derivatives in (lambda, A, tch)
void lyapResiduals(const double* p, double* r, double* jac, void* ctx)
{
    const FitData& d = *static_cast<const FitData*>(ctx);
    const double lambda = p[0], A = p[1], tch = p[2];
    for(int k = 0; k < d.n; k++) {
        const double t = d.t[k];
        if(r)
            r[k] = A + log(t) + (t < tch ? 0 : logexpm1overx(lambda * t) - logexpm1overx(lambda * tch)) - d.lnw[k];
        if(jac) {
            jac[k*3+0] = t < tch ? 0 : (t - tch) * derlogexpm1overx(lambda * t);
            jac[k*3+1] = 1;
            jac[k*3+2] = t < tch ? 0 : -lambda * derlogexpm1overx(lambda * tch);
        }
    }
}

// solve the 3x3 symmetric positive system M x = b (Cholesky); false if not positive definite
bool solve3(const double M[9], const double b[3], double x[3])
{
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for(int i = 0; i < 3; i++)
        for(int j = 0; j <= i; j++) {
            double s = M[i*3+j];
            for(int k = 0; k < j; k++) s -= L[i*3+k] * L[j*3+k];
            if(i == j) {
                if(!(s > 0)) return false;
                L[i*3+i] = sqrt(s);
            } else
                L[i*3+j] = s / L[j*3+j];
        }
    double y[3];
    for(int i = 0; i < 3; i++) {
        double s = b[i];
        for(int k = 0; k < i; k++) s -= L[i*3+k] * y[k];
        y[i] = s / L[i*3+i];
    }
    for(int i = 2; i >= 0; i--) {
        double s = y[i];
        for(int k = i+1; k < 3; k++) s -= L[k*3+i] * x[k];
        x[i] = s / L[i*3+i];
    }
    return true;
}

}  // namespace

int levenbergMarquardt3(agx_resid_fn f, void* ctx, int m, double p[3], double relToler, int maxCalls)
{
    std::vector<double> r(m), rn(m), J(m * 3);
    int calls = 0;
    auto evalR = [&](const double* x, double* out) {
        f(x, out, nullptr, ctx);
        calls++;
        bool ok = true;
        for(int k = 0; k < m; k++) ok &= std::isfinite(out[k]);
        if(!ok) for(int k = 0; k < m; k++) out[k] = 1e10;
        double s = 0;
        for(int k = 0; k < m; k++) s += sq(out[k]);
        return s;
    };
    auto evalJ = [&](const double* x) { f(x, nullptr, J.data(), ctx); calls++; };
    double cost = evalR(p, r.data());
    evalJ(p);
    double A[9], g[3], mu = -1, nu = 2;
    bool converged = false;
    while(!converged && calls < maxCalls) {
        for(int i = 0; i < 3; i++) {
            g[i] = 0;
            for(int j = 0; j < 3; j++) A[i*3+j] = 0;
        }
        for(int k = 0; k < m; k++)
            for(int i = 0; i < 3; i++) {
                g[i] += J[k*3+i] * r[k];
                for(int j = 0; j < 3; j++) A[i*3+j] += J[k*3+i] * J[k*3+j];
            }
        if(mu < 0) mu = 1e-3;   // relative to the diagonal of J^T J (Marquardt's scaling)
        // one iteration: damp until the step lowers the cost
        bool accepted = false;
        double pn[3];
        while(!accepted && calls < maxCalls) {
            double M[9], h[3], mg[3] = { -g[0], -g[1], -g[2] };
            for(int i = 0; i < 9; i++) M[i] = A[i];
            for(int i = 0; i < 3; i++) M[i*4] += mu * fmax(A[i*4], 1e-300);
            if(!solve3(M, mg, h)) {
                mu *= nu;
                nu *= 2;
                if(!(mu < 1e300)) return -calls;   // singular or non-finite normal matrix
                continue;
            }
            for(int i = 0; i < 3; i++) pn[i] = p[i] + h[i];
            const double costn = evalR(pn, rn.data());
            double pred = 0;   // predicted reduction h^T (mu D h - g)
            for(int i = 0; i < 3; i++) pred += h[i] * (mu * fmax(A[i*4], 1e-300) * h[i] - g[i]);
            const double rho = pred > 0 ? (cost - costn) / pred : -1;
            if(!(rho > 0 && costn < cost) && !(cost - costn < 0) && costn == cost) {
                // the step changes nothing: a stationary point
                converged = true;
                break;
            }
            if(rho > 0 && costn < cost) {
                accepted = true;
                mu *= fmax(1./3, 1 - pow(2 * rho - 1, 3));
                nu = 2;
                converged = !(cost - costn > 1e-15 * cost);   // no measurable decrease left
                bool settled = true;
                for(int i = 0; i < 3; i++) settled &= fabs(pn[i] - p[i]) <= relToler * fabs(pn[i]);
                converged |= settled;
                for(int i = 0; i < 3; i++) p[i] = pn[i];
                r.swap(rn);
                cost = costn;
                if(!converged && calls < maxCalls) evalJ(p);
            } else {
                mu *= nu;
                nu *= 2;
                if(!(mu < 1e300)) return -calls;   // no descent left: stationary point
            }
        }
    }
    return converged ? calls : -calls;
}

void lyapunovFromBins(const double* t, const double* lnw, int nb, double maxTime, double period,
    double out[2])
{
    out[0] = out[1] = NAN;
    if(!(maxTime == maxTime) || nb < 5) return;   // nothing recorded, or the orbit is too short
    out[0] = 0;
    out[1] = INFINITY;
    // gsl_fit_linear: slope and intercept from running means (math_fit.cpp:210-224)
    double mx = 0, my = 0, dx2 = 0, dxdy = 0;
    for(int i = 0; i < nb; i++) {
        mx += (t[i] - mx) / (i + 1.0);
        my += (lnw[i] - my) / (i + 1.0);
    }
    for(int i = 0; i < nb; i++) {
        const double dx = t[i] - mx, dy = lnw[i] - my;
        dx2 += (dx * dx - dx2) / (i + 1.0);
        dxdy += (dx * dy - dxdy) / (i + 1.0);
    }
    double params[3];
    params[0] = dxdy / dx2;
    params[1] = my - mx * params[0];
    params[2] = maxTime * 0.5;
    FitData d{ t, lnw, nb };
    levenbergMarquardt3(lyapResiduals, &d, nb, params, 1e-12, 1000);
    const double lambda = params[0];
    double tch = params[2];
    if(lambda < 0 || tch >= t[nb - 2])
        return;
    if(tch < t[0])
        tch = 0;
    const double delta_lnw = logexpm1overx(lambda * t[nb - 1]) - logexpm1overx(lambda * tch);
    if(delta_lnw < M_LN2)
        return;
    out[0] = lambda * period;
    out[1] = tch / period;
}

}  // namespace agx
