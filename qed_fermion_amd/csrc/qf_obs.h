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
    const cplx<R>* U;     // e^{i phix}: x-link phases of the chain
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

// Fields of one estimator at one site of noise vectors s, t (t = s for OBS_V).
// With e = eta*, x+ the site r + x, U = e^{i phix(r)}, Ub = conj(U):
//   T1  a = e_s(r) G_t(r+) U                         b = G_s(r+) e_t(r) U            (1419-1425)
//   T21 a = (e_s(r+) G_t(r) + e_t(r+) G_s(r)) Ub/2    b = (G_s(r+) e_t(r) + G_t(r+) e_s(r)) U/2   (1473-1487)
//   T2  a = (e_s(r) G_t(r+) + e_t(r) G_s(r+)) U/2     b = (G_s(r) e_t(r+) + G_t(r) e_s(r+)) Ub/2  (1530-1544)
//   T4  a = e_s(r+) G_t(r) Ub                         b = G_s(r) e_t(r+) Ub           (1587-1593)
//   V   a = G_s(r+) e_s(r) U + G_s(r) e_s(r+) Ub                                     (1631-1636)
// site indexes a vector (tau * V + y * L + x), xp is its x-neighbour; Uf = e^{i phix}.
template <typename R, int OBS>
__device__ __forceinline__ void obs_site(const cplx<R>* eta, const cplx<R>* G, const cplx<R>* Uf, long long N, int s, int t, long long site,
                                         long long xp, cplx<R>& va, cplx<R>& vb) {
    const cplx<R>* es = eta + (long long)s * N;
    const cplx<R>* gs = G + (long long)s * N;
    const cplx<R> U = Uf[site];
    const cplx<R> Ub = cconj<R>(U);
    const cplx<R> e0 = cconj<R>(es[site]), e1 = cconj<R>(es[xp]);
    const cplx<R> g0 = gs[site], g1 = gs[xp];
    if constexpr (OBS == OBS_V) {
        va = cadd<R>(cmul<R>(cmul<R>(g1, e0), U), cmul<R>(cmul<R>(g0, e1), Ub));
        vb = va;
    } else {
        const cplx<R>* et = eta + (long long)t * N;
        const cplx<R>* gt = G + (long long)t * N;
        const cplx<R> f0 = cconj<R>(et[site]), f1 = cconj<R>(et[xp]);
        const cplx<R> h0 = gt[site], h1 = gt[xp];
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
    }
}

// U[i] = e^{i phix[i]} (accurate sincos), once per estimator call.
template <typename R>
__global__ void __launch_bounds__(256) obs_phase_kernel(const R* phix, cplx<R>* U, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) U[i] = expi<R>(phix[i]);
}

// Fields of every (pair, site) of the launch into A, B (rocFFT path).
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
        cplx<R> va, vb;
        obs_site<R, OBS>(a.eta, a.G, a.U, a.N, s, t, site, xp, va, vb);
        a.A[i] = va;
        if constexpr (OBS != OBS_V) a.B[i] = vb;
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

// out[k] = sum_w part[w * n + k] in a fixed order: a block of 256 threads takes
// 16 consecutive k, and 16 lanes per k each sum every 16th row (w = lane, lane
// + 16, ...) before the 16 partial sums are added in lane order.
__global__ void __launch_bounds__(256) obs_reduce_kernel(const double2* part, int nw, int n, double2* out) {
    __shared__ double2 red[16][17];
    const int kk = threadIdx.x & 15, lane = threadIdx.x >> 4;
    const int k = blockIdx.x * 16 + kk;
    double r = 0, i = 0;
    if (k < n) {
        for (int w = lane; w < nw; w += 16) {
            const double2 p = part[(long long)w * n + k];
            r += p.x;
            i += p.y;
        }
    }
    red[lane][kk] = make_double2(r, i);
    __syncthreads();
    if (lane == 0 && k < n) {
        double sr = 0, si = 0;
        for (int j = 0; j < 16; ++j) {
            sr += red[j][kk].x;
            si += red[j][kk].y;
        }
        out[k] = make_double2(sr, si);
    }
}

// Same-noise-vector sums that the estimators add at one Delta each
// (part[block][3], FP64), over every noise vector and site:
//   0: G(r+) e(r-) U(r) U(r-)          T1 at (0, L-1)   (1434-1440)
//   1: G(r) e(r)                        T21, T2 at (0, 0) (1496-1500, 1553-1557)
//   2: G(r) e(r++) Ub(r) Ub(r+)         T4 at (0, 1)     (1601-1608)
template <typename R>
__global__ void __launch_bounds__(256) obs_diag_kernel(const cplx<R>* eta, const cplx<R>* G, const cplx<R>* Uf, int nrv, int L,
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
        const cplx<R> U0 = Uf[site], Um = Uf[xm], Up = Uf[xp];
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

// ---------------------------------------------------------------- fused path
// One work-group takes items (tau slice, pair) in turn: it builds a and b of
// the slice in LDS, transforms both in place with a mixed-radix Stockham
// autosort FFT (the per-stage formulation of Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC 2008:
// stage of radix P with Ns = product of the earlier radices reads
// x_r = src[j + r L/P] e^{-2 pi i r (j mod Ns)/(Ns P)}, takes their P-point DFT
// and writes it to dst[(j - j mod Ns) P + j mod Ns + q Ns]), rows then
// columns, and adds Ahat(k) Bhat(-k) to FP64 registers.  Only the inputs (from
// L2) and one partial per work-group touch global memory.

constexpr int kObsMaxFac = 8;     // radices of one transform (L <= 78 needs at most 4)
constexpr int kObsThreads = 512;  // threads of a fused work-group (up to 256 VGPRs per lane, no spills)

template <typename R> struct ObsFusedArgs {
    const cplx<R>* eta;
    const cplx<R>* G;
    const cplx<R>* U;           // e^{i phix}
    const int* pairs;
    int count;                  // pairs (OBS_V: noise vectors)
    int nrv, L, Lt;
    long long N;
    int nfac;
    unsigned long long facs;    // radices in stage order, 8 bits each (stage f: bits 8f..8f+7), product L
    double2* part;              // [gridDim.x][V] (+ [gridDim.x][V] sums of Ahat for OBS_V)
};

template <typename R>
__device__ __forceinline__ cplx<R> tw_mul(cplx<R> u, const cplx<R>* tw, int e) {
    return e ? cmul<R>(u, tw[e]) : u;
}

// floor(a / d) for 0 <= a < 2^16 and 1 <= d <= 2^8 from rd = 1.0f / d: the
// fractional part of (a + 0.5) / d is at least 0.5 / d > 1e-3, far above the
// float rounding of the product, so the truncation is exact.  Replaces the
// ~20-instruction integer division in the FFT's index arithmetic.
__device__ __forceinline__ int fdiv(int a, float rd) { return (int)(((float)a + 0.5f) * rd); }

// First element of line l: rows (col = false) of the [n][L][L] arrays are
// lines at l*L, stride 1; columns are line l = array*L + x at
// array*L*L + x, stride L.
__device__ __forceinline__ int line_base(int l, int L, float rL, bool col) {
    if (!col) return l * L;
    const int arr = fdiv(l, rL);
    return arr * L * L + (l - arr * L);
}

// One Stockham stage of compile-time radix P over nlines lines of length L;
// tw[m] = e^{-2 pi i m / L}.
template <typename R, int P>
__device__ __forceinline__ void stockham_stage(const cplx<R>* src, cplx<R>* dst, const cplx<R>* tw, int L, int Ns, int nlines, bool col) {
    const int nb = L / P, total = nlines * nb, tstep = L / (Ns * P), wstep = L / P, es = col ? L : 1;
    const float rnb = 1.0f / nb, rnl = 1.0f / nlines, rNs = 1.0f / Ns, rL = 1.0f / L;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        // rows: neighbouring threads take neighbouring butterflies of a row; columns: the same
        // butterfly of neighbouring columns, so LDS accesses stay at unit stride either way
        int line, j;
        if (col) {
            j = fdiv(b, rnl);
            line = b - j * nlines;
        } else {
            line = fdiv(b, rnb);
            j = b - line * nb;
        }
        const int k = j - fdiv(j, rNs) * Ns;
        const int sb = line_base(line, L, rL, col);
        cplx<R> v[P];
        // twiddles w^r, w = tw[k * tstep] (k * tstep < L), by repeated multiplication: one LDS
        // read per butterfly instead of P - 1 scattered ones
        v[0] = src[sb + j * es];
        if (k == 0) {
#pragma unroll
            for (int r = 1; r < P; ++r) v[r] = src[sb + (j + r * nb) * es];
        } else {
            const cplx<R> w1 = tw[k * tstep];
            cplx<R> w = w1;
#pragma unroll
            for (int r = 1; r < P; ++r) {
                v[r] = cmul<R>(src[sb + (j + r * nb) * es], w);
                if (r + 1 < P) w = cmul<R>(w, w1);
            }
        }
        const int d0 = sb + ((j - k) * P + k) * es, dq = Ns * es;
        if constexpr (P == 2) {
            dst[d0] = cadd<R>(v[0], v[1]);
            dst[d0 + dq] = csub<R>(v[0], v[1]);
        } else if constexpr (P == 3) {
            const R h = (R)0.86602540378443864676;    // sin(2 pi / 3)
            const cplx<R> sm = cadd<R>(v[1], v[2]), df = csub<R>(v[1], v[2]);
            const cplx<R> m = mk<R>(v[0].x - (R)0.5 * sm.x, v[0].y - (R)0.5 * sm.y);
            dst[d0] = cadd<R>(v[0], sm);
            dst[d0 + dq] = mk<R>(m.x + h * df.y, m.y - h * df.x);           // m - i h df
            dst[d0 + 2 * dq] = mk<R>(m.x - h * df.y, m.y + h * df.x);       // m + i h df
        } else if constexpr (P == 4) {
            const cplx<R> a = cadd<R>(v[0], v[2]), bb = csub<R>(v[0], v[2]);
            const cplx<R> c = cadd<R>(v[1], v[3]), d = csub<R>(v[1], v[3]);
            dst[d0] = cadd<R>(a, c);
            dst[d0 + dq] = mk<R>(bb.x + d.y, bb.y - d.x);                    // b - i d
            dst[d0 + 2 * dq] = csub<R>(a, c);
            dst[d0 + 3 * dq] = mk<R>(bb.x - d.y, bb.y + d.x);                // b + i d
        } else {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                cplx<R> y = v[0];
#pragma unroll
                for (int r = 1; r < P; ++r) y = cadd<R>(y, cmul<R>(v[r], tw[((q * r) % P) * wstep]));
                dst[d0 + q * dq] = y;
            }
        }
    }
}

// The same for any radix P (larger primes): inputs are re-read from LDS per output.
template <typename R>
__device__ __forceinline__ void stockham_stage_any(const cplx<R>* src, cplx<R>* dst, const cplx<R>* tw, int L, int P, int Ns, int nlines,
                                                   bool col) {
    const int nb = L / P, total = nlines * nb, tstep = L / (Ns * P), wstep = L / P, es = col ? L : 1;
    const float rnb = 1.0f / nb, rnl = 1.0f / nlines, rNs = 1.0f / Ns, rL = 1.0f / L;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        int line, j;
        if (col) {
            j = fdiv(b, rnl);
            line = b - j * nlines;
        } else {
            line = fdiv(b, rnb);
            j = b - line * nb;
        }
        const int k = j - fdiv(j, rNs) * Ns;
        const int sb = line_base(line, L, rL, col);
        const int d0 = sb + ((j - k) * P + k) * es;
        for (int q = 0; q < P; ++q) {
            cplx<R> y = mk<R>(0, 0);
            int qr = 0;                                   // (q * r) mod P
            for (int r = 0; r < P; ++r) {
                const cplx<R> u = tw_mul<R>(src[sb + (j + r * nb) * es], tw, r * k * tstep);
                y = cadd<R>(y, tw_mul<R>(u, tw, qr * wstep));
                qr += q;
                if (qr >= P) qr -= P;
            }
            dst[d0 + q * Ns * es] = y;
        }
    }
}

// Forward 2-D transforms of the n arrays [n][L][L] at X (LDS) in place, T
// ([n][L][L]) as the ping-pong buffer: rows of all arrays, then columns;
// 2 * nfac stages leave the result in X.
// SMALL: every radix is 2, 3 or 4 (L = 2^a 3^b), so the other butterflies are not compiled
// in and do not hold registers.
template <typename R, bool SMALL>
__device__ void stockham_2d(cplx<R>* X, cplx<R>* T, const cplx<R>* tw, int L, int n, unsigned long long facs, int nfac) {
    cplx<R>* src = X;
    cplx<R>* dst = T;
    for (int pass = 0; pass < 2; ++pass) {
        const bool col = pass != 0;
        int Ns = 1;
        for (int f = 0; f < nfac; ++f) {
            const int P = (int)((facs >> (8 * f)) & 0xFF);
            if constexpr (SMALL) {
                if (P == 4) stockham_stage<R, 4>(src, dst, tw, L, Ns, n * L, col);
                else if (P == 3) stockham_stage<R, 3>(src, dst, tw, L, Ns, n * L, col);
                else stockham_stage<R, 2>(src, dst, tw, L, Ns, n * L, col);
            } else {
                switch (P) {
                    case 2: stockham_stage<R, 2>(src, dst, tw, L, Ns, n * L, col); break;
                    case 3: stockham_stage<R, 3>(src, dst, tw, L, Ns, n * L, col); break;
                    case 4: stockham_stage<R, 4>(src, dst, tw, L, Ns, n * L, col); break;
                    case 5: stockham_stage<R, 5>(src, dst, tw, L, Ns, n * L, col); break;
                    case 7: stockham_stage<R, 7>(src, dst, tw, L, Ns, n * L, col); break;
                    default: stockham_stage_any<R>(src, dst, tw, L, P, Ns, n * L, col); break;
                }
            }
            __syncthreads();
            Ns *= P;
            cplx<R>* t = src;
            src = dst;
            dst = t;
        }
    }
}

// part[w][k] = sum over the work-group's items of Ahat(k) Bhat(-k); for OBS_V
// (Bhat = Ahat) also part[gridDim.x + w][k] = sum Ahat(k).  NT threads,
// MAXK * NT >= L*L.  DUAL: A and B are transformed together as one batch of
// 2L lines (LDS [A | B] + [T | T]); otherwise one after the other through T.
template <typename R, int OBS, int NT, int MAXK, bool DUAL, bool SMALL>
__global__ void __launch_bounds__(NT) obs_fused_kernel(ObsFusedArgs<R> a) {
    extern __shared__ __align__(16) unsigned char obs_lds[];
    const int L = a.L, V = L * L;
    const float rL = 1.0f / L;
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(obs_lds);
    cplx<R>* A = tw + L;
    cplx<R>* B = A + V;
    cplx<R>* T = (OBS == OBS_V) ? B : B + V;    // OBS_V needs no B
    for (int j = threadIdx.x; j < L; j += NT) {
        double sn, cs;
        sincos(-6.283185307179586476925 * j / L, &sn, &cs);
        tw[j] = mk<R>((R)cs, (R)sn);
    }
    double pr[MAXK], pi[MAXK], lr[MAXK], li[MAXK];
#pragma unroll
    for (int m = 0; m < MAXK; ++m) pr[m] = pi[m] = lr[m] = li[m] = 0.0;
    __syncthreads();
    const long long items = (long long)a.count * a.Lt;
    for (long long it = blockIdx.x; it < items; it += gridDim.x) {
        const int tau = (int)(it / a.count);
        const int pidx = (int)(it - (long long)tau * a.count);
        int s, t;
        if constexpr (OBS == OBS_V) {
            s = t = pidx;
        } else {
            s = a.pairs[2 * (long long)pidx];
            t = a.pairs[2 * (long long)pidx + 1];
        }
        if (s < 0 || s >= a.nrv || t < 0 || t >= a.nrv) continue;   // malformed pair: skipped (uniform in the work-group)
        const long long off = (long long)tau * V;
#pragma unroll
        for (int m = 0; m < MAXK; ++m) {
            const int q = threadIdx.x + m * NT;
            if (q < V) {
                const int x = q - fdiv(q, rL) * L;
                cplx<R> va, vb;
                obs_site<R, OBS>(a.eta, a.G, a.U, a.N, s, t, off + q, off + q - x + (x + 1 == L ? 0 : x + 1), va, vb);
                A[q] = va;
                if constexpr (OBS != OBS_V) B[q] = vb;
            }
        }
        __syncthreads();
        if constexpr (OBS == OBS_V) {
            stockham_2d<R, SMALL>(A, T, tw, L, 1, a.facs, a.nfac);
        } else if constexpr (DUAL) {
            stockham_2d<R, SMALL>(A, T, tw, L, 2, a.facs, a.nfac);
        } else {
            stockham_2d<R, SMALL>(A, T, tw, L, 1, a.facs, a.nfac);
            stockham_2d<R, SMALL>(B, T, tw, L, 1, a.facs, a.nfac);
        }
        const cplx<R>* Bh = (OBS == OBS_V) ? A : B;
#pragma unroll
        for (int m = 0; m < MAXK; ++m) {
            const int k = threadIdx.x + m * NT;
            if (k < V) {
                const int ky = fdiv(k, rL), kx = k - ky * L;
                const cplx<R> u = A[k];
                const cplx<R> v = Bh[(ky ? L - ky : 0) * L + (kx ? L - kx : 0)];
                pr[m] += (double)u.x * v.x - (double)u.y * v.y;
                pi[m] += (double)u.x * v.y + (double)u.y * v.x;
                if constexpr (OBS == OBS_V) {
                    lr[m] += u.x;
                    li[m] += u.y;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < MAXK; ++m) {
        const int k = threadIdx.x + m * NT;
        if (k < V) {
            a.part[(long long)blockIdx.x * V + k] = make_double2(pr[m], pi[m]);
            if constexpr (OBS == OBS_V) a.part[(long long)(gridDim.x + blockIdx.x) * V + k] = make_double2(lr[m], li[m]);
        }
    }
}

// LDS bytes of obs_fused_kernel: twiddles, A, B (pair estimators) and the
// ping-pong buffer (2 arrays when DUAL).
template <typename R> inline size_t obs_fused_lds(int L, int obs, bool dual) {
    const size_t V = (size_t)L * L;
    const size_t arrays = obs == OBS_V ? 2 : (dual ? 4 : 3);
    return sizeof(cplx<R>) * ((size_t)L + arrays * V);
}

}  // namespace qf
