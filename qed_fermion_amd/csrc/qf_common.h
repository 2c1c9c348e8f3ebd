// qf_common.h -- shared types for the MI355X (gfx950) HMC fermion kernels.
//
// Complex numbers are float2 / double2 (re, im) so a site is one 8- or
// 16-byte vector load.  Lattice layouts follow the reference's permuted
// views (qed_fermion/hmc_sampler_batch.py:2166, 764):
//   phi  [chain][tau][y][x][mu]   real (float or double), mu = 0: x-link
//   vec  [rhs][tau][y][x]         complex, site index tau*Vs + y*L + x
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qf {

template <typename R> struct CT;
template <> struct CT<float> { using T = float2; using T2 = float4; };
template <> struct CT<double> { using T = double2; using T2 = double4; };
template <typename R> using cplx = typename CT<R>::T;

template <typename R> __device__ __forceinline__ cplx<R> mk(R re, R im) { cplx<R> c; c.x = re; c.y = im; return c; }
template <typename R> __device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ cplx<R> cscale(R s, cplx<R> a) { return mk<R>(s * a.x, s * a.y); }
// a * b
template <typename R> __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename R> __device__ __forceinline__ cplx<R> cjmul(cplx<R> a, cplx<R> b) {
    return mk<R>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// c*v + h*w  (one pair update of a D_g factor)
template <typename R> __device__ __forceinline__ cplx<R> pair_upd(R c, cplx<R> v, cplx<R> h, cplx<R> w) {
    return mk<R>(c * v.x + (h.x * w.x - h.y * w.y), c * v.y + (h.x * w.y + h.y * w.x));
}
// c*v + conj(h)*w
template <typename R> __device__ __forceinline__ cplx<R> pair_updc(R c, cplx<R> v, cplx<R> h, cplx<R> w) {
    return mk<R>(c * v.x + (h.x * w.x + h.y * w.y), c * v.y + (h.x * w.y - h.y * w.x));
}

template <typename R> __device__ __forceinline__ void qsincos(R a, R* s, R* c);
template <> __device__ __forceinline__ void qsincos<float>(float a, float* s, float* c) { sincosf(a, s, c); }
template <> __device__ __forceinline__ void qsincos<double>(double a, double* s, double* c) { sincos(a, s, c); }

// Trotter factor constants: a_1 = t*dtau (group 1), a_{2,3,4} = t*dtau/2
// (hmc_sampler_batch.py:1508-1550).  sh1/ch1 for group 1, sh2/ch2 otherwise.
template <typename R> struct Hop {
    R ch1, sh1, ch2, sh2;
};

// Per-RHS scalar state of the batched PCG (parity double-buffered).
struct CgState {
    double rz;     // Re(r^dag z) of the last completed iteration
    double act;    // 1.0 while the chain is still iterating
    double err;    // ||r|| / ||b|| after the last completed iteration
    double iters;  // reference iteration counter
};

}  // namespace qf
