/*  Operation count of the oracle (test infrastructure, like the rest of oracle/): the oracle's
    source compiled with every `double` replaced by `Counted`, a scalar that holds a double and
    counts the floating-point operations performed on it.  Nothing in the agama_amd product
    links or calls it; oracle/flopcount.py runs it on samples of the benchmark workloads and
    prints the per-unit counts that bench.py keeps as constants.

    What counts as one operation (each on a double, as in the reference's own arithmetic):
    + - * / (binary), sqrt, and every libm function (exp, log, pow, cbrt, atan, ...).
    Negation, fabs, comparisons, min/max and conversions count nothing.

    The counters are per thread (the oracle's orbit loops run on OpenMP threads);
    flc_ops() adds them up over the threads of the last team size used.
*/
#include <functional>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <algorithm>
#include <memory>
#include <atomic>
#include <istream>
#include <type_traits>
#include <omp.h>

namespace flc {
thread_local long long ops = 0;
}

struct Counted {
    double v;
    Counted() = default;
    constexpr Counted(double x) : v(x) {}
    // conversions are explicit (so that `c ? counted : 0.` and the like resolve to Counted)
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator bool() const { return v != 0; }
    explicit constexpr operator int() const { return (int)v; }
    explicit constexpr operator long() const { return (long)v; }
    explicit constexpr operator long long() const { return (long long)v; }
    explicit constexpr operator unsigned() const { return (unsigned)v; }
    explicit constexpr operator unsigned long() const { return (unsigned long)v; }
    explicit constexpr operator float() const { return (float)v; }
    Counted& operator+=(Counted b) { ++flc::ops; v += b.v; return *this; }
    Counted& operator-=(Counted b) { ++flc::ops; v -= b.v; return *this; }
    Counted& operator*=(Counted b) { ++flc::ops; v *= b.v; return *this; }
    Counted& operator/=(Counted b) { ++flc::ops; v /= b.v; return *this; }
};
static_assert(sizeof(Counted) == sizeof(double) && std::is_trivially_copyable<Counted>::value,
    "Counted must be layout-compatible with double (the C API passes double arrays)");

inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted operator+(Counted a) { return a; }
#define FLC_BINOP(OP)                                                                           \
    inline Counted operator OP(Counted a, Counted b) { ++flc::ops; return Counted(a.v OP b.v); } \
    template<class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>     \
    inline Counted operator OP(Counted a, T b) { ++flc::ops; return Counted(a.v OP (double)b); } \
    template<class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>     \
    inline Counted operator OP(T a, Counted b) { ++flc::ops; return Counted((double)a OP b.v); }
FLC_BINOP(+)
FLC_BINOP(-)
FLC_BINOP(*)
FLC_BINOP(/)
#undef FLC_BINOP

#define FLC_CMP(OP)                                                                          \
    inline bool operator OP(Counted a, Counted b) { return a.v OP b.v; }                     \
    template<class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>  \
    inline bool operator OP(Counted a, T b) { return a.v OP (double)b; }                     \
    template<class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>  \
    inline bool operator OP(T a, Counted b) { return (double)a OP b.v; }
FLC_CMP(==)
FLC_CMP(!=)
FLC_CMP(<)
FLC_CMP(>)
FLC_CMP(<=)
FLC_CMP(>=)
#undef FLC_CMP

#define FLC_FN1(F)                                                                  \
    inline Counted F(Counted a) { ++flc::ops; return Counted(std::F(a.v)); }
FLC_FN1(sqrt) FLC_FN1(exp) FLC_FN1(log) FLC_FN1(cbrt) FLC_FN1(atan) FLC_FN1(sin) FLC_FN1(cos)
FLC_FN1(tan) FLC_FN1(asin) FLC_FN1(acos) FLC_FN1(expm1) FLC_FN1(log1p) FLC_FN1(log10) FLC_FN1(sinh)
FLC_FN1(cosh) FLC_FN1(tanh)
#undef FLC_FN1
inline Counted pow(Counted a, Counted b) { ++flc::ops; return Counted(std::pow(a.v, b.v)); }
inline Counted pow(Counted a, double b) { ++flc::ops; return Counted(std::pow(a.v, b)); }
inline Counted pow(Counted a, int b) { ++flc::ops; return Counted(std::pow(a.v, b)); }
inline Counted pow(double a, Counted b) { ++flc::ops; return Counted(std::pow(a, b.v)); }
inline Counted atan2(Counted a, Counted b) { ++flc::ops; return Counted(std::atan2(a.v, b.v)); }
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted fmax(Counted a, Counted b) { return Counted(std::fmax(a.v, b.v)); }
inline Counted fmax(double a, Counted b) { return Counted(std::fmax(a, b.v)); }
inline Counted fmax(Counted a, double b) { return Counted(std::fmax(a.v, b)); }
inline Counted fmin(Counted a, Counted b) { return Counted(std::fmin(a.v, b.v)); }
inline Counted fmin(double a, Counted b) { return Counted(std::fmin(a, b.v)); }
inline Counted fmin(Counted a, double b) { return Counted(std::fmin(a.v, b)); }
inline Counted floor(Counted a) { return Counted(std::floor(a.v)); }
inline Counted ceil(Counted a) { return Counted(std::ceil(a.v)); }
inline std::istream& operator>>(std::istream& s, Counted& a) { return s >> a.v; }
inline std::ostream& operator<<(std::ostream& s, Counted a) { return s << a.v; }
namespace std {
inline Counted max(Counted a, double b) { return a.v < b ? Counted(b) : a; }
inline Counted max(double a, Counted b) { return a < b.v ? b : Counted(a); }
inline Counted min(Counted a, double b) { return b < a.v ? Counted(b) : a; }
inline Counted min(double a, Counted b) { return b.v < a ? b : Counted(a); }
template<class T, class = typename std::enable_if<std::is_integral<T>::value>::type>
inline T min(Counted a, T b) { return (T)a.v < b ? (T)a.v : b; }
template<class T, class = typename std::enable_if<std::is_integral<T>::value>::type>
inline T max(Counted a, T b) { return (T)a.v > b ? (T)a.v : b; }
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline bool signbit(Counted a) { return std::signbit(a.v); }
inline Counted copysign(Counted a, Counted b) { return Counted(std::copysign(a.v, b.v)); }
inline Counted copysign(double a, Counted b) { return Counted(std::copysign(a, b.v)); }
inline Counted copysign(Counted a, double b) { return Counted(std::copysign(a.v, b)); }
inline bool isnan(Counted a) { return std::isnan(a.v); }
inline bool isfinite(Counted a) { return std::isfinite(a.v); }
inline bool isinf(Counted a) { return std::isinf(a.v); }
}

#define double Counted
#include "agama_oracle.cpp"
#undef double

extern "C" {
// operations counted since flc_reset, over the threads of a team of `nthreads`
long long flc_ops(int nthreads)
{
    std::atomic<long long> total{0};
#pragma omp parallel num_threads(nthreads)
    total += flc::ops;
    return total.load();
}
void flc_reset(int nthreads)
{
#pragma omp parallel num_threads(nthreads)
    flc::ops = 0;
}
// operations counted for a fixed set of expressions (the counting rules): 9 in c (2 products, a
// quotient, 2 sums, a difference, sqrt, exp, log), 4 compound assignments, 4 in e (with integer
// and double operands), 2 in f (a product and a sum: negation and fabs count nothing), none for
// fmax, floor, comparisons or conversions -> 19; `out` receives the values
long long flc_selftest(double x, double* out)
{
    const long long before = flc::ops;
    Counted a = x, b = 2;
    Counted c = a * b + a / b - sqrt(a) + exp(b) * log(a);
    Counted d = a;
    d += b; d -= 1; d *= b; d /= 2;
    Counted e = 2 * a + a * 3.0 - 1;
    Counted f = -fabs(-a) + Counted(0.0) * 0;   // the product counts, not the sign or fabs
    const bool cmp = a < b && (int)a >= 0 && fmax(a, b) >= floor(b);
    out[0] = (double)c; out[1] = (double)d; out[2] = (double)e; out[3] = (double)f; out[4] = cmp;
    return flc::ops - before;
}
}
