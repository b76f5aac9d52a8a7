// Forward-mode dual numbers: a value and N directional derivatives.
//
// Every operation computes its value part with the same double operation
// the plain code performs (so a function evaluated on duals returns the
// plain function's value bit for bit) and carries the derivatives by the
// chain rule.  Comparisons look at the value only, so branches follow the
// plain code's path; the derivative is that of the branch taken.
#pragma once
#include <cmath>
#include <type_traits>

#if defined(__HIPCC__)
#define NSA_HD __host__ __device__
#else
#define NSA_HD
#endif

namespace nsa {

template <int N>
struct Dual {
    double v;
    double d[N];
    NSA_HD Dual() : v(0.0) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    NSA_HD Dual(double a) : v(a) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    NSA_HD Dual& operator+=(const Dual& b) { return *this = *this + b; }
    NSA_HD Dual& operator-=(const Dual& b) { return *this = *this - b; }
    NSA_HD Dual& operator*=(const Dual& b) { return *this = *this * b; }
    NSA_HD Dual& operator/=(const Dual& b) { return *this = *this / b; }

    friend NSA_HD Dual operator-(const Dual& a) {
        Dual r;
        r.v = -a.v;
        for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
        return r;
    }
    friend NSA_HD Dual operator+(const Dual& a) { return a; }
    friend NSA_HD Dual operator+(const Dual& a, const Dual& b) {
        Dual r;
        r.v = a.v + b.v;
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
        return r;
    }
    friend NSA_HD Dual operator-(const Dual& a, const Dual& b) {
        Dual r;
        r.v = a.v - b.v;
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
        return r;
    }
    friend NSA_HD Dual operator*(const Dual& a, const Dual& b) {
        Dual r;
        r.v = a.v * b.v;
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
        return r;
    }
    friend NSA_HD Dual operator/(const Dual& a, const Dual& b) {
        Dual r;
        r.v = a.v / b.v;
        for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
        return r;
    }
    friend NSA_HD bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    friend NSA_HD bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    friend NSA_HD bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    friend NSA_HD bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    friend NSA_HD bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
    friend NSA_HD bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

    friend NSA_HD Dual exp(const Dual& a) {
        Dual r;
        r.v = ::exp(a.v);
        for (int k = 0; k < N; ++k) r.d[k] = r.v * a.d[k];
        return r;
    }
    friend NSA_HD Dual log(const Dual& a) {
        Dual r;
        r.v = ::log(a.v);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / a.v;
        return r;
    }
    friend NSA_HD Dual sqrt(const Dual& a) {
        Dual r;
        r.v = ::sqrt(a.v);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
        return r;
    }
    friend NSA_HD Dual fabs(const Dual& a) { return (a.v < 0.0) ? -a : a; }
    friend NSA_HD Dual atan(const Dual& a) {
        Dual r;
        r.v = ::atan(a.v);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (1.0 + a.v * a.v);
        return r;
    }
    // a^b = exp(b log a) for a > 0; the value is pow's
    friend NSA_HD Dual pow(const Dual& a, const Dual& b) {
        Dual r;
        r.v = ::pow(a.v, b.v);
        const double la = ::log(a.v);
        for (int k = 0; k < N; ++k) r.d[k] = r.v * (b.d[k] * la + b.v * a.d[k] / a.v);
        return r;
    }
};

// Value and derivative access that also accepts plain doubles.
NSA_HD inline double value_of(double a) { return a; }
template <int N>
NSA_HD inline double value_of(const Dual<N>& a) { return a.v; }

template <class T>
struct is_dual : std::false_type {};
template <int N>
struct is_dual<Dual<N>> : std::true_type {};

}  // namespace nsa
