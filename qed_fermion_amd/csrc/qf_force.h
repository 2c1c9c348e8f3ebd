// qf_force.h -- adjoint fermion-force kernel.
//
// S_f = psi^dag (M^dag M)^{-1} psi, xi = (M^dag M)^{-1} psi.  The force
// Ft = -dS_f/dphi = 2 Re[(M xi)^dag (dM/dphi) xi] (force_f_sparse,
// hmc_sampler_batch.py:2230-2330).  dM/dphi_tau lives in the block
// (tau+1, tau) with sign sD(tau), so per slice
//   Ft[tau, link] = 2 Re[a_tau^dag (dB_tau/dphi) xi_tau],
//   a_tau = sD(tau) xi_{tau+1} + B_tau xi_tau.
// B_tau = F_0 F_1 ... F_6 (F_k = D_{seq(k)}, palindromic), so
//   a^dag dB xi = sum_k Q_k(a)^dag dF_k Q_{6-k}(xi),   Q_m = F_{m-1}..F_0 applied.
// Q_k(a) is built forward (A <- F_k A) and Q_{6-k}(xi) backward from
// Q_7(xi) = B xi with inverse factors (W <- F_k^{-1} W), so nothing but the
// running pair (A, W) is kept: one LDS exchange of both per factor.
// For bond (i,j) with h = sinh(a_g) e^{i phi}:
//   u^dag dD w = conj(u_i) (i h) w_j + conj(u_j) (-i conj h) w_i;
// the thread that owns the first site i writes the link.
#pragma once
#include "qf_operator.h"

namespace qf {

template <typename R>
__device__ __forceinline__ R force_term(cplx<R> ai, cplx<R> wi, cplx<R> aj, cplx<R> wj, cplx<R> h) {
    const cplx<R> c1 = mk<R>(-h.y, h.x);    // i h
    const cplx<R> c2 = mk<R>(-h.y, -h.x);   // -i conj(h)
    const cplx<R> t1 = cmul<R>(c1, wj), t2 = cmul<R>(c2, wi);
    return R(2) * ((ai.x * t1.x + ai.y * t1.y) + (aj.x * t2.x + aj.y * t2.y));
}

template <typename R, int BY, int G>
__device__ __forceinline__ void force_factor(cplx<R> (&Av)[2 * BY], cplx<R> (&Wv)[2 * BY], R (&F)[2 * BY][2], const Coef<R, BY>& cf,
                                             const Hop<R>& hp, cplx<R>* lds, const Geo& g, int& phase) {
    cplx<R> v2[2][2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) { v2[0][k] = Av[k]; v2[1][k] = Wv[k]; }
    cplx<R> pr[2][BY];
    exchange<R, BY, G, 2>(v2, pr, lds, g, phase);
    // W <- F^{-1} W on own sites (old W kept in v2[1])
    factor_update<R, BY, G, -1>(Wv, pr[1], cf, hp);
    if constexpr (G == 1 || G == 3) {
        const R c = (G == 1) ? hp.ch1 : hp.ch2;
        constexpr int rc = (G == 1) ? 1 : 0;
#pragma unroll
        for (int r = 0, j = 0; r < BY; ++r) {
            if ((r & 1) != rc) {
                F[2 * r][0] += force_term<R>(Av[2 * r], Wv[2 * r], Av[2 * r + 1], Wv[2 * r + 1], cf.ex[r][1]);
            } else {
                const cplx<R> h = cf.ex[r][2];
                const cplx<R> wold_i = v2[1][2 * r + 1], wold_j = pr[1][2 * j + 1];
                // partner's new W: c w_j - conj(h) w_i
                const cplx<R> wj = mk<R>(c * wold_j.x - (h.x * wold_i.x + h.y * wold_i.y), c * wold_j.y - (h.x * wold_i.y - h.y * wold_i.x));
                F[2 * r + 1][0] += force_term<R>(Av[2 * r + 1], Wv[2 * r + 1], pr[0][2 * j + 1], wj, h);
                ++j;
            }
        }
    } else {
        const R c = hp.ch2;
        constexpr int cc = (G == 2) ? 1 : 0;
        constexpr int ci = 1 - cc;
#pragma unroll
        for (int r = 0; r < BY; r += 2)
            F[2 * r + ci][1] += force_term<R>(Av[2 * r + ci], Wv[2 * r + ci], Av[2 * (r + 1) + ci], Wv[2 * (r + 1) + ci], cf.ey[ci][r + 1]);
#pragma unroll
        for (int r = 1; r + 1 < BY; r += 2)
            F[2 * r + cc][1] += force_term<R>(Av[2 * r + cc], Wv[2 * r + cc], Av[2 * (r + 1) + cc], Wv[2 * (r + 1) + cc], cf.ey[cc][r + 1]);
        const cplx<R> h = cf.ey[cc][BY];
        const cplx<R> wold_i = v2[1][2 * (BY - 1) + cc], wold_j = pr[1][1];
        const cplx<R> wj = mk<R>(c * wold_j.x - (h.x * wold_i.x + h.y * wold_i.y), c * wold_j.y - (h.x * wold_i.y - h.y * wold_i.x));
        F[2 * (BY - 1) + cc][1] += force_term<R>(Av[2 * (BY - 1) + cc], Wv[2 * (BY - 1) + cc], pr[0][1], wj, h);
    }
    // A <- F A
    factor_update<R, BY, G, 1>(Av, pr[0], cf, hp);
}

template <typename R>
struct ForceArgs {
    const R* phi;
    long long phi_cs;
    const cplx<R>* xi;
    long long vs;
    R* Ft;               // [rho][tau][y][x][mu]
    long long f_cs;
    int L, Lt, T;
    Hop<R> hp;
    int nband, H, hal, nown;   // bands (blockIdx.z), as in MdmArgs
};

template <typename R, int BY, int NT>
__global__ void __launch_bounds__(NT) force_kernel(ForceArgs<R> A) {
    extern __shared__ __align__(16) unsigned char qf_smem[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(qf_smem);
    const int rho = blockIdx.y;
    const int L = A.L, Lt = A.Lt, Vs = L * L;
    const int t0 = blockIdx.x * A.T;
    const int t1 = min(t0 + A.T, Lt);
    const Geo g = make_geo<BY>(L, threadIdx.x, band_of(L, A.nband, A.H, A.hal, A.nown, blockIdx.z));
    const R* phi = A.phi + rho * A.phi_cs;
    const cplx<R>* xi = A.xi + rho * A.vs;
    R* Ft = A.Ft + rho * A.f_cs;
    int phase = 0;
    Coef<R, BY> cf;
    cplx<R> X[2 * BY], Xn[2 * BY], Av[2 * BY], Wv[2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) X[k] = Xn[k] = Av[k] = Wv[k] = mk<R>(0, 0);
    // every lane loads (padding lanes mirror patch (0,0) and write identical
    // exchange values); only valid lanes store
    load_patch<R, BY>(X, xi + (long long)t0 * Vs, g);
    for (int tau = t0; tau < t1; ++tau) {
        const int tn = (tau + 1 == Lt) ? 0 : tau + 1;
        load_coef<R, BY, false>(cf, phi + (long long)tau * 2 * Vs, g, A.hp);
        load_patch<R, BY>(Xn, xi + (long long)tn * Vs, g);
#pragma unroll
        for (int k = 0; k < 2 * BY; ++k) Wv[k] = X[k];
        apply_B<R, BY, false>(Wv, cf, A.hp, lds, g, phase);   // W = Q_7(xi) = B xi
        axpy_patch<R, BY>(Av, Wv, (R)sgnD(tau, Lt), Xn);   // a = B xi + sD xi_{tau+1}
        R F[2 * BY][2];
#pragma unroll
        for (int k = 0; k < 2 * BY; ++k) F[k][0] = F[k][1] = R(0);
        force_factor<R, BY, 4>(Av, Wv, F, cf, A.hp, lds, g, phase);
        force_factor<R, BY, 3>(Av, Wv, F, cf, A.hp, lds, g, phase);
        force_factor<R, BY, 2>(Av, Wv, F, cf, A.hp, lds, g, phase);
        force_factor<R, BY, 1>(Av, Wv, F, cf, A.hp, lds, g, phase);
        force_factor<R, BY, 2>(Av, Wv, F, cf, A.hp, lds, g, phase);
        force_factor<R, BY, 3>(Av, Wv, F, cf, A.hp, lds, g, phase);
        force_factor<R, BY, 4>(Av, Wv, F, cf, A.hp, lds, g, phase);
        if (g.own) {
            cplx<R>* fo = reinterpret_cast<cplx<R>*>(Ft + (long long)tau * 2 * Vs);
#pragma unroll
            for (int r = 0; r < BY; ++r) {
                fo[(g.gy0 + r) * L + 2 * g.a] = mk<R>(F[2 * r][0], F[2 * r][1]);
                fo[(g.gy0 + r) * L + 2 * g.a + 1] = mk<R>(F[2 * r + 1][0], F[2 * r + 1][1]);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * BY; ++k) X[k] = Xn[k];
    }
}

}  // namespace qf
