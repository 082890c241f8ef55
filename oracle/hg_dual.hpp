// hg_dual.hpp — forward-mode dual numbers for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  Dual<N> carries a value and N directional
// derivatives; instantiating the templated passes of hg_oracle.hpp with it
// gives the exact derivatives of the FP64 model along N seeded directions.
// Comparisons use the value only (orc::val), so every branch of the model
// (clamps, max/min, the talus test, the erosion cases) follows the primal run
// and its derivative is the derivative of the branch taken (a subgradient at
// the switching points).
#pragma once

#include <cmath>

#include "hg_oracle.hpp"

namespace orc {

template <int N> struct Dual {
    double v = 0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}                    // NOLINT: implicit from constants
    Dual(float x) : v(double(x)) {}             // NOLINT
    Dual(int x) : v(double(x)) {}               // NOLINT
    static Dual seed(double x, int k) { Dual r(x); r.d[k] = 1.0; return r; }
    Dual& operator+=(const Dual& o) { v += o.v; for (int i = 0; i < N; ++i) d[i] += o.d[i]; return *this; }
    Dual& operator-=(const Dual& o) { v -= o.v; for (int i = 0; i < N; ++i) d[i] -= o.d[i]; return *this; }
    Dual& operator*=(const Dual& o) { *this = *this * o; return *this; }
    Dual& operator/=(const Dual& o) { *this = *this / o; return *this; }
    friend Dual operator+(Dual a, const Dual& b) { a += b; return a; }
    friend Dual operator-(Dual a, const Dual& b) { a -= b; return a; }
    friend Dual operator-(const Dual& a) { Dual r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
    friend Dual operator*(const Dual& a, const Dual& b) {
        Dual r; r.v = a.v * b.v;
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
        return r;
    }
    friend Dual operator/(const Dual& a, const Dual& b) {
        Dual r; r.v = a.v / b.v;
        const double inv = 1.0 / b.v;
        for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
        return r;
    }
};

template <int N> inline double val(const Dual<N>& x) { return x.v; }

template <int N> inline Dual<N> sqrt(const Dual<N>& x) {
    Dual<N> r; r.v = std::sqrt(x.v);
    const double g = r.v > 0 ? 0.5 / r.v : 0.0;
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * g;
    return r;
}
template <int N> inline Dual<N> atan(const Dual<N>& x) {
    Dual<N> r; r.v = std::atan(x.v);
    const double g = 1.0 / (1.0 + x.v * x.v);
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * g;
    return r;
}
template <int N> inline Dual<N> floor(const Dual<N>& x) { return Dual<N>(std::floor(x.v)); }
template <int N> inline Dual<N> fabs(const Dual<N>& x) { return x.v < 0 ? -x : x; }

// Shadow<K>: a value computed along with K perturbed copies of it.  Every
// comparison of the model reads the base value (orc::val), so the copies
// follow the branches of the unperturbed run: a central difference of two
// copies is the difference quotient of the rollout with its branches frozen,
// whose limit is the derivative of the branch taken, the gradient that the
// dual numbers (and the adjoint) compute.
template <int K> struct Shadow {
    double v = 0;
    double c[K] = {};
    Shadow() = default;
    Shadow(double x) : v(x) { for (int i = 0; i < K; ++i) c[i] = x; }   // NOLINT
    Shadow(float x) : Shadow(double(x)) {}                                // NOLINT
    Shadow(int x) : Shadow(double(x)) {}                                  // NOLINT
    Shadow& operator+=(const Shadow& o) { v += o.v; for (int i = 0; i < K; ++i) c[i] += o.c[i]; return *this; }
    Shadow& operator-=(const Shadow& o) { v -= o.v; for (int i = 0; i < K; ++i) c[i] -= o.c[i]; return *this; }
    Shadow& operator*=(const Shadow& o) { v *= o.v; for (int i = 0; i < K; ++i) c[i] *= o.c[i]; return *this; }
    Shadow& operator/=(const Shadow& o) { v /= o.v; for (int i = 0; i < K; ++i) c[i] /= o.c[i]; return *this; }
    friend Shadow operator+(Shadow a, const Shadow& b) { a += b; return a; }
    friend Shadow operator-(Shadow a, const Shadow& b) { a -= b; return a; }
    friend Shadow operator*(Shadow a, const Shadow& b) { a *= b; return a; }
    friend Shadow operator/(Shadow a, const Shadow& b) { a /= b; return a; }
    friend Shadow operator-(const Shadow& a) { Shadow r; r.v = -a.v; for (int i = 0; i < K; ++i) r.c[i] = -a.c[i]; return r; }
};
template <int K> inline double val(const Shadow<K>& x) { return x.v; }
template <int K> inline Shadow<K> sqrt(const Shadow<K>& x) {
    Shadow<K> r; r.v = std::sqrt(x.v); for (int i = 0; i < K; ++i) r.c[i] = std::sqrt(x.c[i] > 0 ? x.c[i] : 0.0); return r;
}
template <int K> inline Shadow<K> atan(const Shadow<K>& x) {
    Shadow<K> r; r.v = std::atan(x.v); for (int i = 0; i < K; ++i) r.c[i] = std::atan(x.c[i]); return r;
}
// floor and |.| are branches too: the copies take the base run's.
template <int K> inline Shadow<K> floor(const Shadow<K>& x) { return Shadow<K>(std::floor(x.v)); }
template <int K> inline Shadow<K> fabs(const Shadow<K>& x) { return x.v < 0 ? -x : x; }

}  // namespace orc
