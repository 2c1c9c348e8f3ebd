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

// FP32: hardware sin/cos (v_sin_f32/v_cos_f32 after a multiply by 1/2pi),
// absolute error ~1e-6 for |phi| <= pi -- the FP32 operator is only ever
// used as the inner operator of a solve or at the reference's complex64
// accuracy.  FP64: accurate sincos.
template <typename R> __device__ __forceinline__ void qsincos(R a, R* s, R* c);
template <> __device__ __forceinline__ void qsincos<float>(float a, float* s, float* c) { __sincosf(a, s, c); }
template <> __device__ __forceinline__ void qsincos<double>(double a, double* s, double* c) { sincos(a, s, c); }

// Work-group barrier for LDS exchanges only: waits for this wave's LDS
// operations, then s_barrier.  Unlike __syncthreads() it does not wait for
// outstanding global loads, so loads issued for the next slice stay in flight
// across the exchange barriers of the current one.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The MI355X places the work-groups of a launch on its 8 XCDs round-robin in dispatch order
// (w = 0, 8, 16, ... on the first).  xcd_order(w, G) renumbers them so that each XCD takes one
// contiguous range of logical work items, in order: neighbouring items then run on one XCD at
// about the same time and share its L2 (a bijection of [0, G) for any G).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_order(int w, int G) {
    const int q = G / kXcds, r = G - q * kXcds;
    const int j = w % kXcds, k = w / kXcds;
    return j * q + min(j, r) + k;
}

// v + h*w (normalised pair update: 4 FMAs)
template <typename R> __device__ __forceinline__ cplx<R> pair_add(cplx<R> v, cplx<R> h, cplx<R> w) {
    return mk<R>(fma(-h.y, w.y, fma(h.x, w.x, v.x)), fma(h.y, w.x, fma(h.x, w.y, v.y)));
}
// v + conj(h)*w
template <typename R> __device__ __forceinline__ cplx<R> pair_addc(cplx<R> v, cplx<R> h, cplx<R> w) {
    return mk<R>(fma(h.y, w.y, fma(h.x, w.x, v.x)), fma(-h.y, w.x, fma(h.x, w.y, v.y)));
}

// Raw buffer loads (buffer_load with a per-lane VGPR byte offset and a
// uniform SGPR byte offset): a whole patch row or slice is addressed by adding
// uniform offsets on the scalar unit, so the VALU does no address arithmetic.
// The resource covers [p, p + bytes); every access here is in range.
// Stores are NOT done this way: a buffer store of more than 64 bits whose
// soffset is an SGPR is treated by the compiler as free of the store-data
// hazard, and on gfx950 the next VALU write of its data VGPRs corrupted the
// stored value (lanes 12-15 of each 16-lane group).  Stores use global
// instructions (uniform 64-bit base + per-lane 32-bit offset), which the
// compiler's hazard model covers.
typedef unsigned int qf_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int qf_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// one complex value
template <typename R> __device__ __forceinline__ cplx<R> buf_ld1(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so);
template <> __device__ __forceinline__ float2 buf_ld1<float>(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so) {
    const qf_u32x2 q = __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, 0);
    return make_float2(__uint_as_float(q.x), __uint_as_float(q.y));
}
template <> __device__ __forceinline__ double2 buf_ld1<double>(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so) {
    const qf_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, vo, so, 0);
    return make_double2(__hiloint2double(q.y, q.x), __hiloint2double(q.w, q.z));
}
// two adjacent complex values
template <typename R> __device__ __forceinline__ void buf_ld2(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so, cplx<R>& a, cplx<R>& b);
template <> __device__ __forceinline__ void buf_ld2<float>(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so, float2& a, float2& b) {
    const qf_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, vo, so, 0);
    a = make_float2(__uint_as_float(q.x), __uint_as_float(q.y));
    b = make_float2(__uint_as_float(q.z), __uint_as_float(q.w));
}
template <> __device__ __forceinline__ void buf_ld2<double>(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so, double2& a, double2& b) {
    a = buf_ld1<double>(rs, vo, so);
    b = buf_ld1<double>(rs, vo + 16, so);
}
// Trotter factor constants: a_1 = t*dtau (group 1), a_{2,3,4} = t*dtau/2
// (hmc_sampler_batch.py:1508-1550).  ch/sh: cosh/sinh of a_1 (suffix 1) and of
// a_{2,3,4} (suffix 2).  Normalised form: D_g = cosh(a_g) (I + tanh(a_g) K_g),
// so B_tau = cn * prod_g (I + th_g K_g) with cn = ch1 * ch2^6.
template <typename R> struct Hop {
    R ch1, sh1, ch2, sh2;
    R th1, th2, cn;
    double dtau_, t_;   // the parameters they were built from
};

// One lattice of a ragged batch: independent lattices of different L solved
// in the same launches, one right-hand side (chain) each.  Offsets are in
// elements of the concatenated arrays.  The plan is a RaggedHead followed by
// nprob Prob entries; the kernels read it from device memory and find their
// lattice from the work-group index (wg0 / bwg0: first work-group of a
// lattice in the operator / update launch).
struct Prob {
    int L, Lt;        // lattice
    int T, nA;        // operator launch: tau slices per work-group, runs
    int wg0;          // first operator work-group
    int nsx, nB;      // update launch: site tiles per slice row, tiles in total
    int bwg0;         // first update work-group
    long long voff;   // vectors (complex)
    long long poff;   // phi (reals)
    long long hoff;   // hop table (complex)
    long long toff;   // preconditioner table (13 * Lt reals)
    int pa0, pb0;     // operator / update partial sums
    int nband, H, hal, nown;   // bands of a lattice too wide for one work-group per slice (nband 1: none)
};
struct RaggedHead {
    int nprob, by, nthreads;    // patch height and work-group size of the operator launch
    int nwg_op, nwg_upd, tb;    // launch sizes; update tile slices
    int nA_tot, nB_tot;         // partial sums
    long long nsites, ntab;     // total sites; preconditioner table length
};

// Per-RHS scalar state of the batched PCG (parity double-buffered).
struct CgState {
    double rz;     // Re(r^dag z) of the last completed iteration
    double act;    // 1.0 while the chain is still iterating
    double err;    // ||r|| / ||b|| after the last completed iteration
    double iters;  // reference iteration counter
    double stat;   // 0, or the failure that stopped the chain (kSolveNonfinite, kSolveBreakdown)
};

// Per-chain solve status (the QF_SOLVE_* codes of the C ABI).
constexpr int kSolveConverged = 0;   // relative residual <= rtol
constexpr int kSolveMaxIter = 1;     // max_iter iterations without reaching rtol
constexpr int kSolveNonfinite = 2;   // NaN or Inf in the residual, Re(r^dag z) or Re(p^dag A p)
constexpr int kSolveBreakdown = 3;   // Re(p^dag A p) <= 0: A not positive definite on p

}  // namespace qf
