// Real and complex double arithmetic behind one interface, so the dense
// solver kernels are written once for zhegv-like and dsygv-like batches.
#pragma once
#include <hip/hip_runtime.h>

namespace olcao {

template <bool CPLX>
struct Num;
template <>
struct Num<true> {
    struct T {
        double x, y;
    };
    static __device__ T zero() { return {0.0, 0.0}; }
    static __device__ T mk(double r, double i) { return {r, i}; }
    static __device__ T add(T a, T b) { return {a.x + b.x, a.y + b.y}; }
    static __device__ T sub(T a, T b) { return {a.x - b.x, a.y - b.y}; }
    static __device__ T mul(T a, T b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
    static __device__ T cmul(T a, T b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }   // conj(a) b
    static __device__ T scale(T a, double s) { return {a.x * s, a.y * s}; }
    static __device__ T conj(T a) { return {a.x, -a.y}; }
    static __device__ T sel(bool c, T a, T b) { return {c ? a.x : b.x, c ? a.y : b.y}; }
    static __device__ double re(T a) { return a.x; }
    static __device__ double abs2(T a) { return a.x * a.x + a.y * a.y; }
    static __device__ T load(const double *p, long long i) { return {p[2 * i], p[2 * i + 1]}; }
    static __device__ void store(double *p, long long i, T v)
    {
        p[2 * i] = v.x;
        p[2 * i + 1] = v.y;
    }
};
template <>
struct Num<false> {
    typedef double T;
    static __device__ T zero() { return 0.0; }
    static __device__ T mk(double r, double) { return r; }
    static __device__ T add(T a, T b) { return a + b; }
    static __device__ T sub(T a, T b) { return a - b; }
    static __device__ T mul(T a, T b) { return a * b; }
    static __device__ T cmul(T a, T b) { return a * b; }
    static __device__ T scale(T a, double s) { return a * s; }
    static __device__ T conj(T a) { return a; }
    static __device__ T sel(bool c, T a, T b) { return c ? a : b; }
    static __device__ double re(T a) { return a; }
    static __device__ double abs2(T a) { return a * a; }
    static __device__ T load(const double *p, long long i) { return p[i]; }
    static __device__ void store(double *p, long long i, T v) { p[i] = v; }
};

}  // namespace olcao
