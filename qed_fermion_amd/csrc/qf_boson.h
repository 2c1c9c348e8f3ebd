// qf_boson.h -- boson sector of the trajectory (FP64), on the fermion-side
// layout phi[rho][tau][y][x][mu].
//
// S_tau  = cJ sum (1 - cos(d)) (compact) or cJ sum d^2/2, d = phi_{tau+1} - phi_tau
//          (action_boson_tau, hmc_sampler_batch.py:1363-1372)
// S_plaq = cK sum cos(curl), curl = phi_x(r) + phi_y(r+x) - phi_x(r+y) - phi_y(r)
//          (action_boson_plaq_matfree, 1412-1436)
// forces are -dS/dphi (force_b_tau_cmp/noncmp 1338-1360, force_b_plaq_matfree 1438-1459).
// The inner leapfrog of leapfrog_cmp (2332-2347) is a pair of kernels per
// sub-step: kick+drift (elementwise) and force+kick (nearest neighbours).
#pragma once
#include "qf_reduce.h"

namespace qf {

struct BosonArgs {
    double* x;
    double* p;
    double* Fp;
    double* Ft;
    const double* dt;        // per chain
    const double* tau_mask;  // [Lt] or nullptr (all ones)
    long long cs;            // reals per chain
    int L, Lt, nrhs;
    double cJ, cK, m;
    int compact;
    int nsub;                // leapfrog sub-steps (5 in the reference)
};

__device__ __forceinline__ double curl_at(const double* ps, int L, int x, int y) {
    const int xp = (x + 1 == L) ? 0 : x + 1, yp = (y + 1 == L) ? 0 : y + 1;
    return ps[2 * (y * L + x)] + ps[2 * (y * L + xp) + 1] - ps[2 * (yp * L + x)] - ps[2 * (y * L + x) + 1];
}

// Forces at every link; also p += scale * (Fp + Ft) * mask when kick != 0.
__global__ void boson_force_kernel(BosonArgs A, int kick) {
    const int L = A.L, Lt = A.Lt, Vs = L * L;
    const long long nsite = (long long)A.nrhs * Lt * Vs;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nsite; idx += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(idx % Vs);
        const long long q = idx / Vs;
        const int tau = (int)(q % Lt);
        const int rho = (int)(q / Lt);
        const int x = s % L, y = s / L;
        const double* xr = A.x + rho * A.cs;
        const double* ps = xr + (long long)tau * 2 * Vs;
        const double* pn = xr + (long long)((tau + 1 == Lt) ? 0 : tau + 1) * 2 * Vs;
        const double* pp = xr + (long long)(tau == 0 ? Lt - 1 : tau - 1) * 2 * Vs;
        const int xm = x == 0 ? L - 1 : x - 1, ym = y == 0 ? L - 1 : y - 1;
        const double sc = sin(curl_at(ps, L, x, y));
        const double fpx = A.cK * (sc - sin(curl_at(ps, L, x, ym)));
        const double fpy = A.cK * (-sc + sin(curl_at(ps, L, xm, y)));
        double ft[2];
#pragma unroll
        for (int mu = 0; mu < 2; ++mu) {
            const double c = ps[2 * s + mu];
            const double d1 = pn[2 * s + mu] - c, d2 = c - pp[2 * s + mu];
            ft[mu] = A.compact ? -A.cJ * (-sin(d1) + sin(d2)) : -A.cJ * (-d1 + d2);
        }
        const long long o = rho * A.cs + (long long)tau * 2 * Vs + 2 * s;
        A.Fp[o] = fpx; A.Fp[o + 1] = fpy;
        A.Ft[o] = ft[0]; A.Ft[o + 1] = ft[1];
        if (kick) {
            const double h = A.dt[rho] / 2 / A.nsub * (A.tau_mask ? A.tau_mask[tau] : 1.0);
            A.p[o] = A.p[o] + (fpx + ft[0]) * h;
            A.p[o + 1] = A.p[o + 1] + (fpy + ft[1]) * h;
        }
    }
}

// p += (Fp + Ft) dt/2/M mask; x += p/m dt/M mask
__global__ void boson_kick_drift_kernel(BosonArgs A) {
    const long long n = (long long)A.nrhs * A.cs;
    const int per = 2 * A.L * A.L;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int rho = (int)(i / A.cs);
        const int tau = (int)((i % A.cs) / per);
        const double mk_ = A.tau_mask ? A.tau_mask[tau] : 1.0;
        const double dt = A.dt[rho];
        double p = A.p[i] + (A.Fp[i] + A.Ft[i]) * (dt / 2 / A.nsub * mk_);
        A.p[i] = p;
        A.x[i] = A.x[i] + p / A.m * (dt / A.nsub * mk_);
    }
}

// p += scale * dt[rho] * F * mask  (fermion kicks, hmc_sampler_batch.py:2505, 2530)
__global__ void momentum_kick_kernel(double* p, const double* F, const double* dt, double scale, const double* tau_mask, long long cs,
                                     int per, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int rho = (int)(i / cs);
        const int tau = (int)((i % cs) / per);
        const double mk_ = tau_mask ? tau_mask[tau] : 1.0;
        p[i] = p[i] + dt[rho] * scale * F[i] * mk_;
    }
}

// Per-chain partial sums of (S_tau, S_plaq, sum p^2 / 2m, Re psi^dag xi);
// grid (nblk, nrhs), partials[rho][4][nblk].
template <typename RV>
__global__ void __launch_bounds__(256) hamiltonian_partial_kernel(const double* X, const double* P, const cplx<RV>* psi, const cplx<RV>* xi,
                                                                  long long cs, long long vs, int L, int Lt, double cJ, double cK,
                                                                  int compact, double m, double* partials) {
    __shared__ double red[64];
    const int rho = blockIdx.y;
    const int Vs = L * L;
    const long long nsite = (long long)Lt * Vs;
    const double* xr = X + rho * cs;
    const double* pr = P ? P + rho * cs : nullptr;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nsite; idx += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(idx % Vs), tau = (int)(idx / Vs);
        const int x = s % L, y = s / L;
        const double* ps = xr + (long long)tau * 2 * Vs;
        const double* pn = xr + (long long)((tau + 1 == Lt) ? 0 : tau + 1) * 2 * Vs;
#pragma unroll
        for (int mu = 0; mu < 2; ++mu) {
            const double d = pn[2 * s + mu] - ps[2 * s + mu];
            v[0] += compact ? 1.0 - cos(d) : 0.5 * d * d;
            if (pr) {
                const double q = pr[(long long)tau * 2 * Vs + 2 * s + mu];
                v[2] += q * q;
            }
        }
        v[1] += cos(curl_at(ps, L, x, y));
        if (psi) {
            const cplx<RV> a = psi[rho * vs + idx], b = xi[rho * vs + idx];
            v[3] += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
        }
    }
    block_sum<4>(v, red);
    if (threadIdx.x == 0) {
        double* o = partials + (long long)rho * 4 * gridDim.x;
        o[0 * gridDim.x + blockIdx.x] = cJ * v[0];
        o[1 * gridDim.x + blockIdx.x] = cK * v[1];
        o[2 * gridDim.x + blockIdx.x] = v[2] / (2.0 * m);
        o[3 * gridDim.x + blockIdx.x] = v[3];
    }
}

// H[rho] = sum of the four parts; parts[rho][4] optional.
__global__ void hamiltonian_finalize_kernel(const double* partials, int nblk, double* H, double* parts) {
    __shared__ double red[64];
    const int rho = blockIdx.x;
    const double* o = partials + (long long)rho * 4 * nblk;
    double s[4];
    const double* src[4] = {o, o + nblk, o + 2 * nblk, o + 3 * nblk};
    sum_partials<4>(src, nblk, s, red);
    if (threadIdx.x == 0) {
        H[rho] = s[0] + s[1] + s[2] + s[3];
        if (parts)
            for (int k = 0; k < 4; ++k) parts[rho * 4 + k] = s[k];
    }
}

}  // namespace qf
