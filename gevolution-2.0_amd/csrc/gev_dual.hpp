// gev_dual.hpp -- forward-mode dual numbers for the per-particle Jacobians of
// the reverse-mode kernels.  Dual<R, K> carries a value and its derivatives in
// K directions; the functor physics of gev_pcommon.hpp, written over a generic
// real type, then yields the exact derivative of the arithmetic the forward
// kernels execute.  Field values read from memory stay plain R constants.
#pragma once
#include <type_traits>

#include "gev_pcommon.hpp"

namespace gev {

template <typename R, int K> struct Dual {
    R v;
    R d[K];
    __device__ __forceinline__ Dual() {}
    template <typename S, typename = typename std::enable_if<std::is_arithmetic<S>::value>::type>
    __device__ __forceinline__ Dual(S x) : v((R)x)
    {
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = (R)0;
    }
    // the value x seeded as direction j
    __device__ __forceinline__ static Dual seed(R x, int j)
    {
        Dual r((R)0);
        r.v = x;
        r.d[j] = (R)1;
        return r;
    }
    __device__ __forceinline__ Dual &operator+=(const Dual &b)
    {
        v += b.v;
#pragma unroll
        for (int k = 0; k < K; k++) d[k] += b.d[k];
        return *this;
    }
    __device__ __forceinline__ Dual &operator-=(const Dual &b)
    {
        v -= b.v;
#pragma unroll
        for (int k = 0; k < K; k++) d[k] -= b.d[k];
        return *this;
    }
    __device__ __forceinline__ Dual &operator*=(const Dual &b)
    {
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = d[k] * b.v + v * b.d[k];
        v *= b.v;
        return *this;
    }
};

// the real type under a dual number (the type itself for a plain real)
template <typename R> struct DualBase { using type = R; };
template <typename R, int K> struct DualBase<Dual<R, K>> { using type = R; };

template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator+(Dual<R, K> a, const Dual<R, K> &b) { return a += b; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator-(Dual<R, K> a, const Dual<R, K> &b) { return a -= b; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator*(Dual<R, K> a, const Dual<R, K> &b) { return a *= b; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator-(const Dual<R, K> &a)
{
    Dual<R, K> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = -a.d[k];
    return r;
}
// with a plain constant on either side
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator*(const Dual<R, K> &a, R s)
{
    Dual<R, K> r;
    r.v = a.v * s;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * s;
    return r;
}
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator*(R s, const Dual<R, K> &a) { return a * s; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator+(Dual<R, K> a, R s) { a.v += s; return a; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator+(R s, Dual<R, K> a) { a.v += s; return a; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator-(Dual<R, K> a, R s) { a.v -= s; return a; }
template <typename R, int K> __device__ __forceinline__ Dual<R, K> operator-(R s, const Dual<R, K> &a)
{
    Dual<R, K> r = -a;
    r.v += s;
    return r;
}

// the reciprocal and square roots of the forward kernels (gev_pcommon.hpp),
// with their exact derivatives
template <typename R, int K> __device__ __forceinline__ Dual<R, K> frcp(const Dual<R, K> &a)
{
    const R r = frcp(a.v), dr = -r * r;
    Dual<R, K> o;
    o.v = r;
#pragma unroll
    for (int k = 0; k < K; k++) o.d[k] = a.d[k] * dr;
    return o;
}
template <typename R, int K> __device__ __forceinline__ Dual<R, K> fsqrt(const Dual<R, K> &a)
{
    const R s = fsqrt(a.v), ds = (R)0.5 * frcp(s);
    Dual<R, K> o;
    o.v = s;
#pragma unroll
    for (int k = 0; k < K; k++) o.d[k] = a.d[k] * ds;
    return o;
}
template <typename R, int K> __device__ __forceinline__ Dual<R, K> frsqrt(const Dual<R, K> &a)
{
    const R s = frsqrt(a.v), ds = (R)-0.5 * s * s * s;
    Dual<R, K> o;
    o.v = s;
#pragma unroll
    for (int k = 0; k < K; k++) o.d[k] = a.d[k] * ds;
    return o;
}

}  // namespace gev
