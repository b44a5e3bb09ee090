// Finite-difference stencils of the BSSN path, evaluated on a tile staged in
// LDS. `s` points at the sample, `st` is the stride between neighbours along
// the axis, `a` the sample's index along the axis inside its padded block of
// extent `n`, `lo`/`hi` whether the block's low/high face on that axis is a
// domain boundary. Each expression keeps the term order of the reference
// stencil it restates, and floating-point contraction is off, so the sums
// round exactly as the reference's do (the right-hand sides of near-puncture
// data are small differences of these sums, so FMA rounding would show).
#pragma once

namespace dgr {

template <typename T>
struct Axis1 {  // deriv644 first derivative constants (derivs.cpp:2621-2623)
    T by12, by60;
    __device__ explicit Axis1(T h) {
#pragma clang fp contract(off)
        const T ih = T(1.0) / h;
        by12 = ih / T(12.0);
        by60 = ih / T(60.0);
    }
};

template <typename T>
struct Axis2 {  // deriv644 second derivative constants (derivs.cpp:2950-2952)
    T by12, by180;
    __device__ explicit Axis2(T h) {
#pragma clang fp contract(off)
        const T ih2 = T(1.0) / (h * h);
        by180 = ih2 / T(180.0);
        by12 = ih2 / T(12.0);
    }
};

template <typename T>
struct AxisKO {  // ko_deriv42 constants (derivs.cpp:966-972)
    T pre, smr3, smr2, smr1;
    __device__ explicit AxisKO(T h) {
#pragma clang fp contract(off)
        pre = T(-1.0) / T(64.0) / h;
        smr3 = T(59.0) / T(48.0) * T(64) * h;
        smr2 = T(43.0) / T(48.0) * T(64) * h;
        smr1 = T(49.0) / T(48.0) * T(64) * h;
    }
};

// deriv644_{x,y,z}: interior derivs.cpp:2646-2649, closures :2654-2697.
template <typename T>
__device__ __forceinline__ T d1(const T *s, int st, int a, int n, bool lo, bool hi,
                                const Axis1<T> &c) {
#pragma clang fp contract(off)
#define U(o) s[(o) * st]
    const int ie = n - 3;
    if (lo && a < 6) {
        if (a == 3) return (T(-25.0) * U(0) + T(48.0) * U(1) - T(36.0) * U(2) + T(16.0) * U(3) - T(3.0) * U(4)) * c.by12;
        if (a == 4) return (T(-3.0) * U(-1) - T(10.0) * U(0) + T(18.0) * U(1) - T(6.0) * U(2) + U(3)) * c.by12;
        return (+U(-2) - T(8.0) * U(-1) + T(8.0) * U(1) - U(2)) * c.by12;
    }
    if (hi && a >= ie - 3) {
        if (a == ie - 3) return (+U(-2) - T(8.0) * U(-1) + T(8.0) * U(1) - U(2)) * c.by12;
        if (a == ie - 2) return (-U(-3) + T(6.0) * U(-2) - T(18.0) * U(-1) + T(10.0) * U(0) + T(3.0) * U(1)) * c.by12;
        return (T(3.0) * U(-4) - T(16.0) * U(-3) + T(36.0) * U(-2) - T(48.0) * U(-1) + T(25.0) * U(0)) * c.by12;
    }
    return (-U(-3) + T(9.0) * U(-2) - T(45.0) * U(-1) + T(45.0) * U(1) - T(9.0) * U(2) + U(3)) * c.by60;
#undef U
}

// deriv644_{xx,yy,zz}: interior derivs.cpp:2975-2979, closures :2984-3047.
template <typename T>
__device__ __forceinline__ T d2(const T *s, int st, int a, int n, bool lo, bool hi,
                                const Axis2<T> &c) {
#pragma clang fp contract(off)
#define U(o) s[(o) * st]
    const int ie = n - 3;
    if (lo && a < 6) {
        if (a == 3) return (T(45.0) * U(0) - T(154.0) * U(1) + T(214.0) * U(2) - T(156.0) * U(3) + T(61.0) * U(4) - T(10.0) * U(5)) * c.by12;
        if (a == 4) return (T(10.0) * U(-1) - T(15.0) * U(0) - T(4.0) * U(1) + T(14.0) * U(2) - T(6.0) * U(3) + U(4)) * c.by12;
        return (-U(-2) + T(16.0) * U(-1) - T(30.0) * U(0) + T(16.0) * U(1) - U(2)) * c.by12;
    }
    if (hi && a >= ie - 3) {
        if (a == ie - 3) return (-U(-2) + T(16.0) * U(-1) - T(30.0) * U(0) + T(16.0) * U(1) - U(2)) * c.by12;
        if (a == ie - 2) return (U(-4) - T(6.0) * U(-3) + T(14.0) * U(-2) - T(4.0) * U(-1) - T(15.0) * U(0) + T(10.0) * U(1)) * c.by12;
        return (T(-10.0) * U(-5) + T(61.0) * U(-4) - T(156.0) * U(-3) + T(214.0) * U(-2) - T(154.0) * U(-1) + T(45.0) * U(0)) * c.by12;
    }
    return (T(2.0) * U(-3) - T(27.0) * U(-2) + T(270.0) * U(-1) - T(490.0) * U(0) + T(270.0) * U(1) - T(27.0) * U(2) + T(2.0) * U(3)) * c.by180;
#undef U
}

// ko_deriv42_{x,y,z}: interior derivs.cpp:988-993, closures :996-1056.
template <typename T>
__device__ __forceinline__ T ko(const T *s, int st, int a, int n, bool lo, bool hi,
                                const AxisKO<T> &c) {
#pragma clang fp contract(off)
#define U(o) s[(o) * st]
    const int ie = n - 3;
    if (lo && a < 6) {
        if (a == 3) return (U(3) - T(3.0) * U(2) + T(3.0) * U(1) - U(0)) / c.smr3;
        if (a == 4) return (U(3) - T(6.0) * U(2) + T(12.0) * U(1) - T(10.0) * U(0) + T(3.0) * U(-1)) / c.smr2;
        return (U(3) - T(6.0) * U(2) + T(15.0) * U(1) - T(19.0) * U(0) + T(12.0) * U(-1) - T(3.0) * U(-2)) / c.smr1;
    }
    if (hi && a >= ie - 3) {
        if (a == ie - 3) return (U(-3) - T(6.0) * U(-2) + T(15.0) * U(-1) - T(19.0) * U(0) + T(12.0) * U(1) - T(3.0) * U(2)) / c.smr1;
        if (a == ie - 2) return (U(-3) - T(6.0) * U(-2) + T(12.0) * U(-1) - T(10.0) * U(0) + T(3.0) * U(1)) / c.smr2;
        return (U(-3) - T(3.0) * U(-2) + T(3.0) * U(-1) - U(0)) / c.smr3;
    }
    return c.pre * (-U(-3) + T(6.0) * U(-2) - T(15.0) * U(-1) + T(20.0) * U(0) - T(15.0) * U(1) + T(6.0) * U(2) - U(3));
#undef U
}

// Stencil constants of one axis for a whole block, computed once with the
// same operations as Axis1/Axis2/AxisKO (so the values are identical).
template <typename T>
struct AxisConstsT {
    T d1_12, d1_60, d2_12, d2_180, ko_pre, ko_s3, ko_s2, ko_s1;
};
using AxisConsts = AxisConstsT<double>;

__host__ __device__ inline AxisConsts axis_consts(double h) {
#pragma clang fp contract(off)
    AxisConsts k;
    const double ih = 1.0 / h;
    k.d1_12 = ih / 12.0;
    k.d1_60 = ih / 60.0;
    const double ih2 = 1.0 / (h * h);
    k.d2_180 = ih2 / 180.0;
    k.d2_12 = ih2 / 12.0;
    k.ko_pre = -1.0 / 64.0 / h;
    k.ko_s3 = 59.0 / 48.0 * 64 * h;
    k.ko_s2 = 43.0 / 48.0 * 64 * h;
    k.ko_s1 = 49.0 / 48.0 * 64 * h;
    return k;
}

// The same stencils with a compile-time stride ST, split into the interior
// formula and the face closure: *_edge returns the closure value where the
// point is one of the three nearest a flagged face (lo: a < 6, hi: a >= ie-3
// with ie = n-3) and `r` otherwise, so a caller whose tile touches no flagged
// face on this axis skips it altogether. Expressions as in d1/d2/ko above.
#define U(o) s[(o) * ST]
template <int ST, typename T>
__device__ __forceinline__ T d1_in(const T *s, T by60) {
#pragma clang fp contract(off)
    return (-U(-3) + T(9.0) * U(-2) - T(45.0) * U(-1) + T(45.0) * U(1) - T(9.0) * U(2) + U(3)) * by60;
}

template <int ST, typename T>
__device__ __forceinline__ T d1_edge(const T *s, T r, int a, int ie, bool lo, bool hi, T by12) {
#pragma clang fp contract(off)
    if (lo && a < 6) {
        if (a == 3) return (-T(25.0) * U(0) + T(48.0) * U(1) - T(36.0) * U(2) + T(16.0) * U(3) - T(3.0) * U(4)) * by12;
        if (a == 4) return (-T(3.0) * U(-1) - T(10.0) * U(0) + T(18.0) * U(1) - T(6.0) * U(2) + U(3)) * by12;
        return (+U(-2) - T(8.0) * U(-1) + T(8.0) * U(1) - U(2)) * by12;
    }
    if (hi && a >= ie - 3) {
        if (a == ie - 3) return (+U(-2) - T(8.0) * U(-1) + T(8.0) * U(1) - U(2)) * by12;
        if (a == ie - 2) return (-U(-3) + T(6.0) * U(-2) - T(18.0) * U(-1) + T(10.0) * U(0) + T(3.0) * U(1)) * by12;
        return (T(3.0) * U(-4) - T(16.0) * U(-3) + T(36.0) * U(-2) - T(48.0) * U(-1) + T(25.0) * U(0)) * by12;
    }
    return r;
}

template <int ST, typename T>
__device__ __forceinline__ T d2_in(const T *s, T by180) {
#pragma clang fp contract(off)
    return (T(2.0) * U(-3) - T(27.0) * U(-2) + T(270.0) * U(-1) - T(490.0) * U(0) + T(270.0) * U(1) - T(27.0) * U(2) + T(2.0) * U(3)) * by180;
}

template <int ST, typename T>
__device__ __forceinline__ T d2_edge(const T *s, T r, int a, int ie, bool lo, bool hi, T by12) {
#pragma clang fp contract(off)
    if (lo && a < 6) {
        if (a == 3) return (T(45.0) * U(0) - T(154.0) * U(1) + T(214.0) * U(2) - T(156.0) * U(3) + T(61.0) * U(4) - T(10.0) * U(5)) * by12;
        if (a == 4) return (T(10.0) * U(-1) - T(15.0) * U(0) - T(4.0) * U(1) + T(14.0) * U(2) - T(6.0) * U(3) + U(4)) * by12;
        return (-U(-2) + T(16.0) * U(-1) - T(30.0) * U(0) + T(16.0) * U(1) - U(2)) * by12;
    }
    if (hi && a >= ie - 3) {
        if (a == ie - 3) return (-U(-2) + T(16.0) * U(-1) - T(30.0) * U(0) + T(16.0) * U(1) - U(2)) * by12;
        if (a == ie - 2) return (U(-4) - T(6.0) * U(-3) + T(14.0) * U(-2) - T(4.0) * U(-1) - T(15.0) * U(0) + T(10.0) * U(1)) * by12;
        return (-T(10.0) * U(-5) + T(61.0) * U(-4) - T(156.0) * U(-3) + T(214.0) * U(-2) - T(154.0) * U(-1) + T(45.0) * U(0)) * by12;
    }
    return r;
}

template <int ST, typename T>
__device__ __forceinline__ T ko_in(const T *s, T pre) {
#pragma clang fp contract(off)
    return pre * (-U(-3) + T(6.0) * U(-2) - T(15.0) * U(-1) + T(20.0) * U(0) - T(15.0) * U(1) + T(6.0) * U(2) - U(3));
}

template <int ST, typename T, typename K>
__device__ __forceinline__ T ko_edge(const T *s, T r, int a, int ie, bool lo, bool hi, const K &k) {
#pragma clang fp contract(off)
    if (lo && a < 6) {
        if (a == 3) return (U(3) - T(3.0) * U(2) + T(3.0) * U(1) - U(0)) / T(k.ko_s3);
        if (a == 4) return (U(3) - T(6.0) * U(2) + T(12.0) * U(1) - T(10.0) * U(0) + T(3.0) * U(-1)) / T(k.ko_s2);
        return (U(3) - T(6.0) * U(2) + T(15.0) * U(1) - T(19.0) * U(0) + T(12.0) * U(-1) - T(3.0) * U(-2)) / T(k.ko_s1);
    }
    if (hi && a >= ie - 3) {
        if (a == ie - 3) return (U(-3) - T(6.0) * U(-2) + T(15.0) * U(-1) - T(19.0) * U(0) + T(12.0) * U(1) - T(3.0) * U(2)) / T(k.ko_s1);
        if (a == ie - 2) return (U(-3) - T(6.0) * U(-2) + T(12.0) * U(-1) - T(10.0) * U(0) + T(3.0) * U(1)) / T(k.ko_s2);
        return (U(-3) - T(3.0) * U(-2) + T(3.0) * U(-1) - U(0)) / T(k.ko_s3);
    }
    return r;
}
#undef U

}  // namespace dgr
