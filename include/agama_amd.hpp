/*  agama_amd C++ layer: the reference's C++ surface for the Multipole + orbit path, over the
    C ABI of include/agama_amd.h (header only; link with libagama_amd.so).

    Mirrors, for a host program written against Agama's C++ classes:
      - potential::BasePotential::eval(const PosCar&, double*, GradCar*, HessCar*, double)
        (src/potential_base.h:153-156): MultipolePotential::eval, with NULL outputs skipped, and a
        batched evalMany (no counterpart; one device launch for many points);
      - shared ownership of immutable potentials (src/potential_base.h:30-46, PtrPotential):
        MultipolePotential::create returns std::shared_ptr<const MultipolePotential>;
      - orbit::OrbitIntParams (src/orbit.h:116-129) and the integration of a batch of orbits, one
        OrbitIntegrator<Car> per orbit in the reference (src/orbit.h:165-277,
        src/interface_python.cpp:8099-8127): integrateOrbits (final states) and
        integrateTrajectories (orbit::Trajectory samples at trajsize regular times);
      - errors as exceptions (std::runtime_error carrying agx_last_error()).
    The value types have the field names of the reference's coord::PosCar, GradCar, HessCar and
    PosVelCar (src/coord.h:190-196, 353-375, 457-498), so that code holding the reference's types
    copies them member for member (INTEGRATION.md shows the adapter class). */
#ifndef AGAMA_AMD_HPP
#define AGAMA_AMD_HPP
#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include "agama_amd.h"

namespace agama_amd {

struct PosCar { double x, y, z; };
struct GradCar { double dx, dy, dz; };
struct HessCar { double dx2, dy2, dz2, dxdy, dydz, dxdz; };
struct PosVelCar { double x, y, z, vx, vy, vz; };
/// samples of one orbit: (phase-space point, time), as orbit::Trajectory (src/orbit.h:28)
typedef std::vector<std::pair<PosVelCar, double> > Trajectory;

/// orbit::OrbitIntParams (src/orbit.h:116-129); HERMITE is not provided by this library
struct OrbitIntParams {
    enum Method { DOP853, DPRKN8 };
    double accuracy;
    size_t maxNumSteps;
    Method method;
    OrbitIntParams(double _accuracy = 1e-8, size_t _maxNumSteps = 100000000, Method _method = DOP853) :
        accuracy(_accuracy), maxNumSteps(_maxNumSteps), method(_method) {}
};

/// per-orbit counters of a batch (accepted steps, step attempts, max |E/E0 - 1|, status:
/// 0 done, 1 stepper failure, 2 maxNumSteps reached)
struct OrbitBatchInfo {
    std::vector<long> nsteps, nattempts;
    std::vector<double> maxRelEnergyError;
    std::vector<int> status;
};

inline void check(int rc, const char* what)
{
    if(rc != AGX_OK) throw std::runtime_error(std::string(what) + ": " + agx_last_error());
}

/// Multipole potential (optionally with one analytic NFW halo) evaluated on the GPU
class MultipolePotential {
public:
    /// params: Agama's key=value form ("type=Multipole, file=coefs.ini" or a density), see
    /// agx_potential_create; throws std::runtime_error on failure
    static std::shared_ptr<const MultipolePotential> create(const std::string& params)
    {
        agx_potential* h = agx_potential_create(params.c_str());
        if(!h) throw std::runtime_error("agama_amd: " + std::string(agx_last_error()));
        return std::shared_ptr<const MultipolePotential>(new MultipolePotential(h));
    }
    ~MultipolePotential() { agx_potential_destroy(h_); }
    MultipolePotential(const MultipolePotential&) = delete;
    MultipolePotential& operator=(const MultipolePotential&) = delete;

    /// BasePotential::eval for Cartesian coordinates: potential, gradient (+dPhi/dx) and
    /// Hessian, each skipped when NULL
    void eval(const PosCar& pos, double* potential = NULL, GradCar* deriv = NULL, HessCar* deriv2 = NULL,
        double time = 0) const
    {
        const double p[3] = { pos.x, pos.y, pos.z };
        double g[3], h[6];
        const double val = agx_eval_potential(h_, p, time, g, deriv2 ? h : NULL);
        if(std::isnan(val) && agx_last_error()[0]) throw std::runtime_error("agama_amd eval: " + std::string(agx_last_error()));
        if(potential) *potential = val;
        if(deriv) { deriv->dx = g[0]; deriv->dy = g[1]; deriv->dz = g[2]; }
        if(deriv2) {
            deriv2->dx2 = h[0]; deriv2->dy2 = h[1]; deriv2->dz2 = h[2];
            deriv2->dxdy = h[3]; deriv2->dydz = h[4]; deriv2->dxdz = h[5];
        }
    }

    /// the same for n points in one device launch (arrays of n values; NULL outputs skipped)
    void evalMany(size_t n, const PosCar* pos, double* potential, GradCar* deriv, HessCar* deriv2 = NULL,
        double time = 0) const
    {
        static_assert(sizeof(PosCar) == 3 * sizeof(double) && sizeof(GradCar) == 3 * sizeof(double) &&
            sizeof(HessCar) == 6 * sizeof(double), "packed value types");
        const double* p = reinterpret_cast<const double*>(pos);
        if(deriv2)
            check(agx_eval_hess_many(h_, (long)n, p, time, potential, reinterpret_cast<double*>(deriv),
                reinterpret_cast<double*>(deriv2), AGX_PRECISION_FP64), "agama_amd evalMany");
        else
            check(agx_eval_many(h_, (long)n, p, time, potential, reinterpret_cast<double*>(deriv),
                AGX_PRECISION_FP64), "agama_amd evalMany");
    }

    const agx_potential* handle() const { return h_; }

private:
    explicit MultipolePotential(agx_potential* h) : h_(h) {}
    agx_potential* h_;
};

namespace detail {
inline int methodFlag(const OrbitIntParams& p) { return p.method == OrbitIntParams::DPRKN8 ? AGX_METHOD_DPRKN8 : 0; }
inline void fillInfo(OrbitBatchInfo* info, size_t n, std::vector<long>& ns, std::vector<long>& na,
    std::vector<double>& de, std::vector<int>& st)
{
    if(!info) return;
    info->nsteps.swap(ns); info->nattempts.swap(na); info->maxRelEnergyError.swap(de); info->status.swap(st);
    (void)n;
}
}  // namespace detail

/// Integrate orbits from ic[i] at time tstart (scalar) for time T[i] each, in a frame rotating
/// at Omega; returns each orbit's last point (OrbitIntegrator<Car>::run followed by the final
/// sample of RuntimeTrajectory with trajsize=1).  One persistent device launch for the batch.
inline std::vector<PosVelCar> integrateOrbits(const MultipolePotential& pot, const std::vector<PosVelCar>& ic,
    const std::vector<double>& T, double Omega = 0, const OrbitIntParams& params = OrbitIntParams(),
    double tstart = 0, OrbitBatchInfo* info = NULL)
{
    const size_t n = ic.size();
    if(T.size() != n) throw std::invalid_argument("agama_amd integrateOrbits: ic and T differ in length");
    std::vector<double> ts(n, tstart);
    std::vector<int> sz(n, 1);
    std::vector<PosVelCar> fin(n);
    std::vector<double> tout(n), de(n);
    std::vector<long> ns(n), na(n);
    std::vector<int> st(n);
    check(agx_orbit_batch(pot.handle(), (long)n, reinterpret_cast<const double*>(ic.data()), ts.data(), T.data(),
        sz.data(), params.accuracy, Omega, (long)params.maxNumSteps, tout.data(), NULL,
        reinterpret_cast<double*>(fin.data()), ns.data(), na.data(), NULL, de.data(), st.data(),
        detail::methodFlag(params)), "agama_amd integrateOrbits");
    detail::fillInfo(info, n, ns, na, de, st);
    return fin;
}

/// the same with `trajsize` samples per orbit at regular intervals T[i]/(trajsize-1) (the
/// first at tstart, the last at tstart + T[i]); orbits that stop early have fewer samples
inline std::vector<Trajectory> integrateTrajectories(const MultipolePotential& pot, const std::vector<PosVelCar>& ic,
    const std::vector<double>& T, int trajsize, double Omega = 0, const OrbitIntParams& params = OrbitIntParams(),
    double tstart = 0, OrbitBatchInfo* info = NULL)
{
    const size_t n = ic.size();
    if(T.size() != n) throw std::invalid_argument("agama_amd integrateTrajectories: ic and T differ in length");
    if(trajsize < 1) throw std::invalid_argument("agama_amd integrateTrajectories: trajsize must be >= 1");
    std::vector<double> ts(n, tstart), de(n);
    std::vector<int> sz(n, trajsize), st(n);
    std::vector<double> tout(n * trajsize);
    std::vector<PosVelCar> xout(n * trajsize);
    std::vector<long> ns(n), na(n);
    check(agx_orbit_batch(pot.handle(), (long)n, reinterpret_cast<const double*>(ic.data()), ts.data(), T.data(),
        sz.data(), params.accuracy, Omega, (long)params.maxNumSteps, tout.data(),
        reinterpret_cast<double*>(xout.data()), NULL, ns.data(), na.data(), NULL, de.data(), st.data(),
        detail::methodFlag(params)), "agama_amd integrateTrajectories");
    std::vector<Trajectory> out(n);
    for(size_t i = 0; i < n; i++)
        for(int k = 0; k < trajsize; k++) {
            const size_t r = i * trajsize + k;
            if(std::isnan(tout[r])) break;   // NaN padding after an early stop
            out[i].push_back(std::make_pair(xout[r], tout[r]));
        }
    detail::fillInfo(info, n, ns, na, de, st);
    return out;
}

}  // namespace agama_amd
#endif
