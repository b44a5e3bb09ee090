// Host build of the generated pointwise BSSN equations and constraints
// (Hamiltonian, momentum, psi4), for checking the
// generated code itself on the CPU against the reference's golden vectors.
// Not part of the device path. Input layout per point: 163 doubles
// ([0,24) fields, [24,96) first derivatives 3*v+d, [96,162) second
// derivatives 6*m+c of alpha, chi, beta0..2, gt0..5, [162] eta).
#include <cmath>
#include <cstdint>

#include "bssn_rhs_generated.h"
#include "bssn_rhs_gauge_generated.h"
#include "bssn_ham_generated.h"

namespace {
// Coordinates for the callers without any: the psi4 expressions are still
// evaluated (sqrt keeps them alive on the host) and must read valid memory.
constexpr double kNoCoords[3] = {1.0, 2.0, 3.0};
template <typename T>
struct In163 {
    const double *p;
    template <int V>
    T fld() const { return (T)p[V]; }
    template <int V, int D>
    T d1() const { return (T)p[24 + 3 * V + D]; }
    template <int M, int C>
    T d2() const { return (T)p[96 + 6 * M + C]; }
    const double *xyz = kNoCoords;  // coordinates, read by psi4 only
    template <int D>
    T crd() const { return (T)xyz[D]; }
};
template <typename T>
struct Out24 {
    double *o;
    template <int V>
    void put(T x) { o[V] = (double)x; }
};
template <typename T>
void run(const double *in, int64_t n, const double *lam, const double *lamf, double *out) {
    T l[4] = {(T)lam[0], (T)lam[1], (T)lam[2], (T)lam[3]};
    T lf[2] = {(T)lamf[0], (T)lamf[1]};
    for (int64_t q = 0; q < n; q++) {
        In163<T> a{in + 163 * q};
        Out24<T> b{out + 24 * q};
        dgr::bssn_rhs_point<T>(a, (T)in[163 * q + 162], l, lf, b);
    }
}
template <typename T>
struct Out1 {  // keeps the Hamiltonian constraint (output 0) only
    double *o;
    template <int V>
    void put(T x) {
        if (V == 0) *o = (double)x;
    }
};
template <typename T>
struct Out6 {  // ham, mom0, mom1, mom2, psi4 real, psi4 imaginary
    double *o;
    template <int V>
    void put(T x) { o[V] = (double)x; }
};
}  // namespace

extern "C" void dgr_host_ham_points(const double *in, int64_t n, double *out) {
    for (int64_t q = 0; q < n; q++) {
        In163<double> a{in + 163 * q};
        Out1<double> b{out + q};
        dgr::bssn_ham_point<double>(a, b);
    }
}

// Six outputs per point (ham, mom0..2, psi4 real and imaginary) at the
// coordinates xyz (three per point).
extern "C" void dgr_host_constraint_points(const double *in, const double *xyz, int64_t n, double *out) {
    for (int64_t q = 0; q < n; q++) {
        In163<double> a{in + 163 * q, xyz + 3 * q};
        Out6<double> b{out + 6 * q};
        dgr::bssn_ham_point<double>(a, b);
    }
}

extern "C" void dgr_host_bssn_points(const double *in, int64_t n, const double *lam, const double *lamf,
                                     double *out) {
    run<double>(in, n, lam, lamf, out);
}

extern "C" void dgr_host_bssn_points_f32(const double *in, int64_t n, const double *lam, const double *lamf,
                                         double *out) {
    run<float>(in, n, lam, lamf, out);
}

// The generated equations with the gauge variants (Rochester gauge when
// rochester != 0, the slow-start lapse at time t when ssl != 0, formed here as
// the device kernels form it: 0.6 sqrt(chi) (alpha - sqrt(chi)) exp(-t^2/800)).
extern "C" void dgr_host_bssn_points_gauge(const double *in, int64_t n, const double *lam, const double *lamf,
                                           const double *xi, int rochester, int ssl, double t, double *out) {
    const double e = std::exp(-1.0 / 800.0 * (t * t));
    for (int64_t q = 0; q < n; q++) {
        const double *p = in + 163 * q;
        In163<double> a{p};
        Out24<double> b{out + 24 * q};
        double sslt = 0.0;
        if (ssl) {
            const double w = std::sqrt(p[1]);
            sslt = 0.6 * w * (p[0] - w) * e;
        }
        dgr::bssn_rhs_point_gauge<double>(a, p[162], lam, lamf, xi, rochester, sslt, b);
    }
}
