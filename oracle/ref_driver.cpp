/*  Test-only driver around the reference's own DOP853 (math_ode.cpp, compiled from the
    reference tree by ref_build.sh into oracle/_ref).  It exposes the same entry points as
    orc_ode_dop853 / orc_ode_dop853_fixed in agama_oracle.h so that tests can compare the
    restated stepper with the original one bit for bit on identical right-hand sides.
    The right-hand side is a C callback; nothing here restates the stepper. */
#include "math_ode.h"
#include <cmath>

typedef void (*ref_rhs_fn)(double t, const double* x, double* dxdt, void* ctx);

namespace {
class CallbackOde : public math::IOdeSystem {
public:
    int n; ref_rhs_fn fn; void* ctx; double tb;
    CallbackOde(int _n, ref_rhs_fn _fn, void* _ctx) : n(_n), fn(_fn), ctx(_ctx), tb(0) {}
    virtual void eval(const double t, const double x[], double dxdt[], double* af) const
    { (void)af; fn(tb + t, x, dxdt, ctx); }
    virtual unsigned int size() const { return n; }
};

class CallbackOde2 : public math::IOdeSystem2ndOrder {
public:
    int n; ref_rhs_fn fn; void* ctx; double tb;
    CallbackOde2(int _n, ref_rhs_fn _fn, void* _ctx) : n(_n), fn(_fn), ctx(_ctx), tb(0) {}
    virtual void eval2(const double t, const double xv[], double d2xdt2[], double* d3, double* af) const
    { (void)d3; (void)af; fn(tb + t, xv, d2xdt2, ctx); }
    virtual unsigned int size() const { return n; }
};

typedef void (*ref_mat_fn)(double t, double* a, double* b, void* ctx);

class CallbackLinear : public math::IOdeSystem2ndOrderLinear {
public:
    ref_mat_fn fn; void* ctx; double tb;
    CallbackLinear(ref_mat_fn _fn, void* _ctx) : fn(_fn), ctx(_ctx), tb(0) {}
    virtual void evalMat(const double t, double a[], double b[]) const { fn(tb + t, a, b, ctx); }
    virtual unsigned int size() const { return 6; }
};
}

extern "C" {

/* the reference's Ode2StepperGL4<3> on x'' = A(t) x + B(t) x' for nvec vectors (x0: nvec*6
   values, each {x, x'}); steps of the given lengths; records the end state of every step and
   the dense output at the fractions frac[0..nfrac) of every step */
int ref_ode_gl4(int nvec, ref_mat_fn mat, void* ctx, const double* x0, int nsteps, const double* steps,
    double* out_states, int nfrac, const double* frac, double* out_dense)
{
    CallbackLinear sys(mat, ctx);
    math::Ode2StepperGL4<3> st(sys, nvec);
    st.init(x0);
    double tb = 0;
    for(int s=0; s<nsteps; s++) {
        st.doStep(steps[s]);
        for(int i=0; i<nvec*6; i++) out_states[(long)s*nvec*6+i] = st.getSol(steps[s], i);
        for(int k=0; k<nfrac; k++)
            for(int i=0; i<nvec*6; i++)
                out_dense[((long)s*nfrac+k)*nvec*6+i] = st.getSol(steps[s] * frac[k], i);
        tb += steps[s];
        sys.tb = tb;
    }
    return 0;
}

/* adaptive integration over [0, ttotal]; dense output at sorted tsample[]; end state */
int ref_ode_dop853(int ndim, ref_rhs_fn rhs, void* ctx, const double* x0, double ttotal,
    double accuracy, double* out_xend, int nsample, const double* tsample, double* out_sample,
    long* out_nsteps)
{
    CallbackOde sys(ndim, rhs, ctx);
    math::OdeStepperDOP853 st(sys, accuracy);
    st.init(x0);
    double tb = 0, hlast = 0;
    long ns = 0;
    int is = 0;
    while(tb != ttotal) {
        double rem = ttotal - tb;
        double h = st.doStep(rem);
        if(!(h > 0)) return 1;
        while(is < nsample && tsample[is] <= tb + h) {
            for(int i=0; i<ndim; i++) out_sample[is*ndim+i] = st.getSol(tsample[is] - tb, i);
            is++;
        }
        tb = h == rem ? ttotal : tb + h;
        sys.tb = tb;
        hlast = h;
        ns++;
    }
    for(int i=0; i<ndim; i++) out_xend[i] = st.getSol(hlast, i);
    if(out_nsteps) *out_nsteps = ns;
    return 0;
}

/* the reference's DPRKN8 on x'' = acc(t, x), ndim = 2N; as ref_ode_dop853 */
int ref_ode_dprkn8(int ndim, ref_rhs_fn acc, void* ctx, const double* x0, double ttotal,
    double accuracy, double* out_xend, int nsample, const double* tsample, double* out_sample,
    long* out_nsteps)
{
    CallbackOde2 sys(ndim, acc, ctx);
    math::OdeStepperDPRKN8 st(sys, accuracy);
    st.init(x0);
    double tb = 0, hlast = 0;
    long ns = 0;
    int is = 0;
    while(tb != ttotal) {
        double rem = ttotal - tb;
        double h = st.doStep(rem);
        if(!(h > 0)) return 1;
        while(is < nsample && tsample[is] <= tb + h) {
            for(int i=0; i<ndim; i++) out_sample[is*ndim+i] = st.getSol(tsample[is] - tb, i);
            is++;
        }
        tb = h == rem ? ttotal : tb + h;
        sys.tb = tb;
        hlast = h;
        ns++;
    }
    for(int i=0; i<ndim; i++) out_xend[i] = st.getSol(hlast, i);
    if(out_nsteps) *out_nsteps = ns;
    return 0;
}

/* nsteps fixed steps of length h (accuracy = infinity accepts every step); records the
   state at the end of each step and the dense output at its midpoint */
int ref_ode_dop853_fixed(int ndim, ref_rhs_fn rhs, void* ctx, const double* x0, double h,
    int nsteps, double* out_states, double* out_mid)
{
    CallbackOde sys(ndim, rhs, ctx);
    math::OdeStepperDOP853 st(sys, INFINITY);
    st.init(x0);
    for(int s=0; s<nsteps; s++) {
        double hh = st.doStep(h);
        if(!(hh > 0)) return 1;
        for(int i=0; i<ndim; i++) {
            out_states[s*ndim+i] = st.getSol(hh, i);
            if(out_mid) out_mid[s*ndim+i] = st.getSol(0.5*hh, i);
        }
        sys.tb += hh;
    }
    return 0;
}

}
