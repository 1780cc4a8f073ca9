// Lyapunov exponent from the coarse-grained growth of the deviation vector: the fitting part of
// calcLyapunovExponent (src/orbit_variational.cpp:406-448) on the bins accumulated by the
// variational orbit kernel.
#pragma once

namespace agx {

// t[nb], lnw[nb]: bin times and log of the mean deviation-vector magnitude; maxTime: time of the
// last recorded point since the start (NaN: nothing was recorded); period: T_circ of the orbit.
// out = {lambda * period, t_chaos / period}: NaN NaN without data or with fewer than 5 bins,
// 0 and infinity when no significant exponential growth is found.
void lyapunovFromBins(const double* t, const double* lnw, int nb, double maxTime, double period,
    double out[2]);

// the least-squares Levenberg-Marquardt solver used by the fit: minimises sum_k r_k(p)^2 over
// 3 parameters, stopping when every parameter changes by at most relToler * |p| in an
// iteration, the sum of squares stops decreasing, or after maxCalls residual and Jacobian
// evaluations; returns the number of evaluations (negative: not converged)
typedef void (*agx_resid_fn)(const double* p, double* r, double* jac, void* ctx);
int levenbergMarquardt3(agx_resid_fn f, void* ctx, int m, double p[3], double relToler, int maxCalls);

}  // namespace agx
