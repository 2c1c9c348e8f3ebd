// qf_obs.h -- contractions of the bond-bond stochastic estimators
// (qed_fermion/stochastic_estimator.py:1393-1654, T1 / T21 / T2 / T4 / Tvv).
//
// Each estimator is a correlation over Delta = (dy, dx) of two fields a, b
// built from a pair of noise vectors (s < s'), averaged over pairs and tau
// slices:  C(Delta) = (1/V) sum_r a(r - Delta) b(r).  The caller takes 2-D
// FFTs of the fields this file writes (obs_fields_kernel); the k-space
// products Ahat(k) Bhat(-k) are then summed over every slice of every pair in
// FP64 (obs_kspace_kernel), so one inverse transform of a [Ly, Lx] array per
// estimator finishes the job instead of one per pair batch.
//
// Layouts: eta and G_eta = M^{-1} eta are [nrv][tau][y][x] complex; phix is
// [tau][y][x] real, the x-link angle of the chain.  x-neighbours wrap.
#pragma once
#include "qf_common.h"

namespace qf {

enum { OBS_T1 = 0, OBS_T21 = 1, OBS_T2 = 2, OBS_T4 = 3, OBS_V = 4 };

template <typename R> struct ObsArgs {
    const cplx<R>* eta;   // noise vectors (conjugated here: e = eta*)
    const cplx<R>* G;     // G_eta
    const R* phix;        // x-link angles of the chain
    const int* pairs;     // [2 * npairs] (s, s'), pair estimators only
    int p0, np;           // pairs (OBS_V: noise vectors) [p0, p0 + np) of this launch
    int nrv, L;
    long long N;          // sites per vector
    cplx<R>* A;           // [np][N]
    cplx<R>* B;           // [np][N], pair estimators only
};

template <typename R> __device__ __forceinline__ cplx<R> cconj(cplx<R> a) { return mk<R>(a.x, -a.y); }

// e^{i p}, accurate (the estimators are read at the precision of their inputs)
template <typename R> __device__ __forceinline__ cplx<R> expi(R p);
template <> __device__ __forceinline__ float2 expi<float>(float p) {
    float s, c;
    sincosf(p, &s, &c);
    return make_float2(c, s);
}
template <> __device__ __forceinline__ double2 expi<double>(double p) {
    double s, c;
    sincos(p, &s, &c);
    return make_double2(c, s);
}

// Fields of one estimator for each (pair, site) of the launch.  With
// e = eta*, x+ the site r + x, U = e^{i phix(r)}, Ub = conj(U):
//   T1  a = e_s(r) G_s'(r+) U                          b = G_s(r+) e_s'(r) U            (1419-1425)
//   T21 a = (e_s(r+) G_s'(r) + e_s'(r+) G_s(r)) Ub/2    b = (G_s(r+) e_s'(r) + G_s'(r+) e_s(r)) U/2   (1473-1487)
//   T2  a = (e_s(r) G_s'(r+) + e_s'(r) G_s(r+)) U/2     b = (G_s(r) e_s'(r+) + G_s'(r) e_s(r+)) Ub/2  (1530-1544)
//   T4  a = e_s(r+) G_s'(r) Ub                          b = G_s(r) e_s'(r+) Ub           (1587-1593)
//   V   a = G_s(r+) e_s(r) U + G_s(r) e_s(r+) Ub  (one noise vector s)                   (1631-1636)
template <typename R, int OBS>
__global__ void __launch_bounds__(256) obs_fields_kernel(ObsArgs<R> a) {
    const long long total = (long long)a.np * a.N;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int j = (int)(i / a.N);
        const long long site = i - (long long)j * a.N;
        const int x = (int)(site % a.L);
        const long long xp = site - x + (x + 1 == a.L ? 0 : x + 1);
        int s, t;
        if constexpr (OBS == OBS_V) {
            s = a.p0 + j;
            t = s;
        } else {
            s = a.pairs[2 * (long long)(a.p0 + j)];
            t = a.pairs[2 * (long long)(a.p0 + j) + 1];
        }
        if (s < 0 || s >= a.nrv || t < 0 || t >= a.nrv) {   // malformed pair list: zero fields, no access
            a.A[i] = mk<R>(0, 0);
            if constexpr (OBS != OBS_V) a.B[i] = mk<R>(0, 0);
            continue;
        }
        const cplx<R>* es = a.eta + (long long)s * a.N;
        const cplx<R>* gs = a.G + (long long)s * a.N;
        const cplx<R> U = expi<R>(a.phix[site]);
        const cplx<R> Ub = cconj<R>(U);
        const cplx<R> e0 = cconj<R>(es[site]), e1 = cconj<R>(es[xp]);
        const cplx<R> g0 = gs[site], g1 = gs[xp];
        if constexpr (OBS == OBS_V) {
            a.A[i] = cadd<R>(cmul<R>(cmul<R>(g1, e0), U), cmul<R>(cmul<R>(g0, e1), Ub));
        } else {
            const cplx<R>* et = a.eta + (long long)t * a.N;
            const cplx<R>* gt = a.G + (long long)t * a.N;
            const cplx<R> f0 = cconj<R>(et[site]), f1 = cconj<R>(et[xp]);
            const cplx<R> h0 = gt[site], h1 = gt[xp];
            cplx<R> va, vb;
            if constexpr (OBS == OBS_T1) {
                va = cmul<R>(cmul<R>(e0, h1), U);
                vb = cmul<R>(cmul<R>(g1, f0), U);
            } else if constexpr (OBS == OBS_T21) {
                va = cscale<R>((R)0.5, cmul<R>(cadd<R>(cmul<R>(e1, h0), cmul<R>(f1, g0)), Ub));
                vb = cscale<R>((R)0.5, cmul<R>(cadd<R>(cmul<R>(g1, f0), cmul<R>(h1, e0)), U));
            } else if constexpr (OBS == OBS_T2) {
                va = cscale<R>((R)0.5, cmul<R>(cadd<R>(cmul<R>(e0, h1), cmul<R>(f0, g1)), U));
                vb = cscale<R>((R)0.5, cmul<R>(cadd<R>(cmul<R>(g0, f1), cmul<R>(h0, e1)), Ub));
            } else {
                va = cmul<R>(cmul<R>(e1, h0), Ub);
                vb = cmul<R>(cmul<R>(g0, f1), Ub);
            }
            a.A[i] = va;
            a.B[i] = vb;
        }
    }
}

// part[w][k] (+)= sum over slices sl = w, w + nwg, ... of Ahat[sl][k] * Bhat[sl][-k]
// (FP64), -k = ((L - ky) % L, (L - kx) % L); with lin, part[nwg + w][k] (+)=
// sum Ahat[sl][k].  One work-group per w; acc = 0 overwrites.
template <typename R>
__global__ void __launch_bounds__(256) obs_kspace_kernel(const cplx<R>* Ah, const cplx<R>* Bh, long long nsl, int L, int lin, int acc,
                                                         double2* part, int nwg) {
    const int V = L * L;
    const int w = blockIdx.x;
    double2* out = part + (long long)w * V;
    double2* outl = part + (long long)(nwg + w) * V;
    for (int k = threadIdx.x; k < V; k += blockDim.x) {
        const int ky = k / L, kx = k - ky * L;
        const int nk = (ky ? L - ky : 0) * L + (kx ? L - kx : 0);
        double pr = 0, pi = 0, lr = 0, li = 0;
        for (long long sl = w; sl < nsl; sl += nwg) {
            const cplx<R> u = Ah[sl * V + k];
            const cplx<R> v = Bh[sl * V + nk];
            pr += (double)u.x * v.x - (double)u.y * v.y;
            pi += (double)u.x * v.y + (double)u.y * v.x;
            lr += u.x;
            li += u.y;
        }
        if (acc) {
            const double2 o = out[k];
            pr += o.x;
            pi += o.y;
        }
        out[k] = make_double2(pr, pi);
        if (lin) {
            if (acc) {
                const double2 o = outl[k];
                lr += o.x;
                li += o.y;
            }
            outl[k] = make_double2(lr, li);
        }
    }
}

// out[k] = sum_w part[w * n + k], w in order
__global__ void __launch_bounds__(256) obs_reduce_kernel(const double2* part, int nw, int n, double2* out) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double r = 0, i = 0;
        for (int w = 0; w < nw; ++w) {
            const double2 p = part[(long long)w * n + k];
            r += p.x;
            i += p.y;
        }
        out[k] = make_double2(r, i);
    }
}

// Same-noise-vector sums that the estimators add at one Delta each
// (part[block][3], FP64), over every noise vector and site:
//   0: G(r+) e(r-) U(r) U(r-)          T1 at (0, L-1)   (1434-1440)
//   1: G(r) e(r)                        T21, T2 at (0, 0) (1496-1500, 1553-1557)
//   2: G(r) e(r++) Ub(r) Ub(r+)         T4 at (0, 1)     (1601-1608)
template <typename R>
__global__ void __launch_bounds__(256) obs_diag_kernel(const cplx<R>* eta, const cplx<R>* G, const R* phix, int nrv, int L,
                                                       long long N, double2* part) {
    __shared__ double red[6][256];
    const long long total = (long long)nrv * N;
    const long long stride = (long long)gridDim.x * blockDim.x;
    double s0r = 0, s0i = 0, s1r = 0, s1i = 0, s2r = 0, s2i = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const long long sv = i / N;
        const long long site = i - sv * N;
        const int x = (int)(site % L);
        const long long row = site - x;
        const long long xp = row + (x + 1) % L, xm = row + (x + L - 1) % L, xpp = row + (x + 2) % L;
        const cplx<R>* e = eta + sv * N;
        const cplx<R>* g = G + sv * N;
        const cplx<R> U0 = expi<R>(phix[site]), Um = expi<R>(phix[xm]), Up = expi<R>(phix[xp]);
        const cplx<R> d0 = cmul<R>(cmul<R>(g[xp], cconj<R>(e[xm])), cmul<R>(U0, Um));
        const cplx<R> d1 = cmul<R>(g[site], cconj<R>(e[site]));
        const cplx<R> d2 = cmul<R>(cmul<R>(g[site], cconj<R>(e[xpp])), cconj<R>(cmul<R>(U0, Up)));
        s0r += d0.x; s0i += d0.y;
        s1r += d1.x; s1i += d1.y;
        s2r += d2.x; s2i += d2.y;
    }
    const int t = threadIdx.x;
    red[0][t] = s0r; red[1][t] = s0i; red[2][t] = s1r; red[3][t] = s1i; red[4][t] = s2r; red[5][t] = s2i;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int c = 0; c < 6; ++c) red[c][t] += red[c][t + h];
        __syncthreads();
    }
    if (t < 3) part[(long long)blockIdx.x * 3 + t] = make_double2(red[2 * t][0], red[2 * t + 1][0]);
}

}  // namespace qf
