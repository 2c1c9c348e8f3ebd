// qf_patch.h -- register-blocked application of the checkerboard transfer
// matrix B_tau = D4 D3 D2 D1 D2 D3 D4 on one tau slice.
//
// Reference semantics: hmc_sampler_batch.py:937-969 (bond groups) and
// 1552-1611 (factor order, cosh/sinh weights).  Group g holds the bonds
// (i, i+mu) whose first site i has parity p = (x+y) mod 2:
//   g1: mu=x, p=0   g2: mu=y, p=0   g3: mu=x, p=1   g4: mu=y, p=1
// and D_g = cosh(a_g) I + sinh(a_g) K_g with K_g[i,j] = e^{i phi_mu(i)},
// K_g[j,i] = e^{-i phi_mu(i)}; a_1 = t dtau, a_{2,3,4} = t dtau / 2.
//
// MI355X mapping: one work-group owns a whole L x L slice; every thread owns a
// 2 x BY patch of sites (x = 2a, 2a+1; y = b*BY .. b*BY+BY-1) held in VGPRs.
// Because patches start on even coordinates, half of the pairs of every
// factor are inside a patch and are updated in registers; only the pairs that
// cross a patch edge go through LDS (one ds_write/ds_read per crossing site,
// one barrier per factor, two LDS buffers alternating so a single barrier
// orders both the RAW and the WAR hazard).
#pragma once
#include "qf_common.h"

namespace qf {

// Factor order of B_tau, applied right to left (first entry applied first).
__host__ __device__ constexpr int seq_group(int k) { return (k == 3) ? 1 : (k == 2 || k == 4) ? 2 : (k == 1 || k == 5) ? 3 : 4; }

struct Geo {
    int L;       // lattice extent (even)
    int a, b;    // patch coordinates: a in [0, L/2), b in [0, L/BY)
    int nb;      // L / BY
    int y0;      // b * BY
    bool valid;  // false for padding lanes of the last wave
};

template <int BY>
__device__ __forceinline__ Geo make_geo(int L, int tid) {
    Geo g;
    g.L = L;
    const int na = L >> 1;
    g.nb = L / BY;
    g.valid = tid < na * g.nb;
    const int t = g.valid ? tid : 0;
    g.a = t % na;
    g.b = t / na;
    g.y0 = g.b * BY;
    return g;
}

// Hopping coefficients of one slice for one patch, pre-scaled by sinh(a_g).
//   ex[r][0] = sh * e^{i phi_x(2a-1, y)}   (bond from the left neighbour)
//   ex[r][1] = sh * e^{i phi_x(2a,   y)}   (bond inside the patch)
//   ex[r][2] = sh * e^{i phi_x(2a+1, y)}   (bond to the right neighbour)
//   ey[c][0] = sh2 * e^{i phi_y(x_c, y0-1)},  ey[c][r+1] = sh2 * e^{i phi_y(x_c, y0+r)}
// Row r (y = y0 + r, y0 even): the inside bond belongs to g1 when r is even
// and to g3 when r is odd; the edge bonds belong to the other x-group.
template <typename R, int BY>
struct Coef {
    cplx<R> ex[BY][3];
    cplx<R> ey[2][BY + 1];
};

template <typename R>
__device__ __forceinline__ cplx<R> phase(R sh, R ph) {
    R s, c;
    qsincos<R>(ph, &s, &c);
    return mk<R>(sh * c, sh * s);
}

// phi_s: slice base, layout [y][x][mu] (R reals).
template <typename R, int BY>
__device__ __forceinline__ void load_coef(Coef<R, BY>& cf, const R* __restrict__ phi_s, const Geo& g, const Hop<R>& hp) {
    const int L = g.L;
    const cplx<R>* ps = reinterpret_cast<const cplx<R>*>(phi_s);  // one (phi_x, phi_y) pair per site
    const int x0 = 2 * g.a;
    const int xm = (x0 + L - 1) % L;
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const int y = g.y0 + r;
        const cplx<R> p0 = ps[y * L + x0];
        const cplx<R> p1 = ps[y * L + x0 + 1];
        const R pxm = phi_s[2 * (y * L + xm)];
        const R sh_in = (r & 1) ? hp.sh2 : hp.sh1;   // g1 (r even) or g3 (r odd)
        const R sh_ed = (r & 1) ? hp.sh1 : hp.sh2;
        cf.ex[r][0] = phase<R>(sh_ed, pxm);
        cf.ex[r][1] = phase<R>(sh_in, p0.x);
        cf.ex[r][2] = phase<R>(sh_ed, p1.x);
        cf.ey[0][r + 1] = phase<R>(hp.sh2, p0.y);
        cf.ey[1][r + 1] = phase<R>(hp.sh2, p1.y);
    }
    const int ym = (g.y0 + L - 1) % L;
    const cplx<R> q0 = ps[ym * L + x0];
    const cplx<R> q1 = ps[ym * L + x0 + 1];
    cf.ey[0][0] = phase<R>(hp.sh2, q0.y);
    cf.ey[1][0] = phase<R>(hp.sh2, q1.y);
}

// LDS exchange buffer: two alternating halves, each NV * (L/2)*L complex
// (NV vectors exchanged together).  Callers size the dynamic LDS with
// lds_bytes<R>(L, NV).
__host__ __device__ constexpr int xstride(int L) { return (L >> 1) * L; }
template <typename R>
inline size_t lds_bytes(int L, int NV) { return (size_t)2 * NV * xstride(L) * sizeof(cplx<R>); }

// Partner values of the sites that cross a patch edge for group G.  Writes the
// patch's edge values, barriers, reads the neighbours' edge values.  pr[] gets
// the partner value for the edge sites (index: x-groups -> 2*j + {0: left
// partner of the cx=0 site, 1: right partner of the cx=1 site} for the j-th
// edge row; y-groups -> {0: partner below r=0, 1: partner above r=BY-1}).
template <typename R, int BY, int G, int NV>
__device__ __forceinline__ void exchange(cplx<R> (*v)[2 * BY], cplx<R> (*pr)[BY], cplx<R>* lds, const Geo& g, int& phase) {
    const int L = g.L;
    const int half = xstride(L);                 // per-vector sub-buffer
    cplx<R>* buf = lds + (phase & 1) * NV * half;
    if constexpr (G == 1 || G == 3) {
        constexpr int rc = (G == 1) ? 1 : 0;   // edge rows of group G have r % 2 == rc
        const int x0 = 2 * g.a;
        if (g.valid) {
#pragma unroll
            for (int n = 0; n < NV; ++n)
#pragma unroll
                for (int r = rc; r < BY; r += 2) {
                    cplx<R>* row = buf + n * half + ((g.y0 + r) >> 1) * L;
                    row[x0] = v[n][2 * r];
                    row[x0 + 1] = v[n][2 * r + 1];
                }
        }
        __syncthreads();
        if (g.valid) {
            const int xl = (x0 + L - 1) % L, xr = (x0 + 2) % L;
#pragma unroll
            for (int n = 0; n < NV; ++n)
#pragma unroll
                for (int r = rc, j = 0; r < BY; r += 2, ++j) {
                    const cplx<R>* row = buf + n * half + ((g.y0 + r) >> 1) * L;
                    pr[n][2 * j] = row[xl];
                    pr[n][2 * j + 1] = row[xr];
                }
        }
    } else {
        constexpr int cc = (G == 2) ? 1 : 0;   // edge column of group G
        const int na = L >> 1;
        if (g.valid) {
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                cplx<R>* s = buf + n * half + 2 * (g.b * na + g.a);
                s[0] = v[n][cc];
                s[1] = v[n][2 * (BY - 1) + cc];
            }
        }
        __syncthreads();
        if (g.valid) {
            const int bm = (g.b + g.nb - 1) % g.nb, bp = (g.b + 1) % g.nb;
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                pr[n][0] = buf[n * half + 2 * (bm * na + g.a) + 1];
                pr[n][1] = buf[n * half + 2 * (bp * na + g.a) + 0];
            }
        }
    }
    ++phase;
}

// v <- D_G v (SGN = +1) or D_G^{-1} v (SGN = -1; D^{-1} = cosh - sinh K since
// K_g^2 = 1 on a perfect matching), given the edge partner values pr.
template <typename R, int BY, int G, int SGN>
__device__ __forceinline__ void factor_update(cplx<R> (&v)[2 * BY], const cplx<R> (&pr)[BY], const Coef<R, BY>& cf, const Hop<R>& hp) {
    if constexpr (G == 1 || G == 3) {
        const R c = (G == 1) ? hp.ch1 : hp.ch2;
        constexpr int rc = (G == 1) ? 1 : 0;
#pragma unroll
        for (int r = 0, j = 0; r < BY; ++r) {
            if ((r & 1) != rc) {   // bond inside the patch, first site cx=0
                const cplx<R> h = SGN > 0 ? cf.ex[r][1] : mk<R>(-cf.ex[r][1].x, -cf.ex[r][1].y);
                const cplx<R> v0 = v[2 * r], v1 = v[2 * r + 1];
                v[2 * r] = pair_upd<R>(c, v0, h, v1);
                v[2 * r + 1] = pair_updc<R>(c, v1, h, v0);
            } else {               // both sites on edge bonds
                const cplx<R> hl = SGN > 0 ? cf.ex[r][0] : mk<R>(-cf.ex[r][0].x, -cf.ex[r][0].y);
                const cplx<R> hr = SGN > 0 ? cf.ex[r][2] : mk<R>(-cf.ex[r][2].x, -cf.ex[r][2].y);
                v[2 * r] = pair_updc<R>(c, v[2 * r], hl, pr[2 * j]);
                v[2 * r + 1] = pair_upd<R>(c, v[2 * r + 1], hr, pr[2 * j + 1]);
                ++j;
            }
        }
    } else {
        const R c = hp.ch2;
        constexpr int cc = (G == 2) ? 1 : 0;
        constexpr int ci = 1 - cc;
#pragma unroll
        for (int r = 0; r < BY; r += 2) {   // inside column: pairs (r, r+1), r even
            const cplx<R> h = SGN > 0 ? cf.ey[ci][r + 1] : mk<R>(-cf.ey[ci][r + 1].x, -cf.ey[ci][r + 1].y);
            const cplx<R> v0 = v[2 * r + ci], v1 = v[2 * (r + 1) + ci];
            v[2 * r + ci] = pair_upd<R>(c, v0, h, v1);
            v[2 * (r + 1) + ci] = pair_updc<R>(c, v1, h, v0);
        }
#pragma unroll
        for (int r = 1; r + 1 < BY; r += 2) {   // edge column: pairs (r, r+1), r odd
            const cplx<R> h = SGN > 0 ? cf.ey[cc][r + 1] : mk<R>(-cf.ey[cc][r + 1].x, -cf.ey[cc][r + 1].y);
            const cplx<R> v0 = v[2 * r + cc], v1 = v[2 * (r + 1) + cc];
            v[2 * r + cc] = pair_upd<R>(c, v0, h, v1);
            v[2 * (r + 1) + cc] = pair_updc<R>(c, v1, h, v0);
        }
        const cplx<R> hb = SGN > 0 ? cf.ey[cc][0] : mk<R>(-cf.ey[cc][0].x, -cf.ey[cc][0].y);
        const cplx<R> ht = SGN > 0 ? cf.ey[cc][BY] : mk<R>(-cf.ey[cc][BY].x, -cf.ey[cc][BY].y);
        v[cc] = pair_updc<R>(c, v[cc], hb, pr[0]);
        v[2 * (BY - 1) + cc] = pair_upd<R>(c, v[2 * (BY - 1) + cc], ht, pr[1]);
    }
}

template <typename R, int BY, int G>
__device__ __forceinline__ void factor(cplx<R> (&v)[2 * BY], const Coef<R, BY>& cf, const Hop<R>& hp, cplx<R>* lds, const Geo& g,
                                       int& phase) {
    cplx<R> pr[1][BY];
    exchange<R, BY, G, 1>(reinterpret_cast<cplx<R>(*)[2 * BY]>(&v), pr, lds, g, phase);
    factor_update<R, BY, G, 1>(v, pr[0], cf, hp);
}

// v <- B_tau v
template <typename R, int BY>
__device__ __forceinline__ void apply_B(cplx<R> (&v)[2 * BY], const Coef<R, BY>& cf, const Hop<R>& hp, cplx<R>* lds, const Geo& g,
                                        int& phase) {
    factor<R, BY, 4>(v, cf, hp, lds, g, phase);
    factor<R, BY, 3>(v, cf, hp, lds, g, phase);
    factor<R, BY, 2>(v, cf, hp, lds, g, phase);
    factor<R, BY, 1>(v, cf, hp, lds, g, phase);
    factor<R, BY, 2>(v, cf, hp, lds, g, phase);
    factor<R, BY, 3>(v, cf, hp, lds, g, phase);
    factor<R, BY, 4>(v, cf, hp, lds, g, phase);
}

// v <- (first nfac factors of B_tau) v, nfac in [0, 7] (uniform over the block).
template <typename R, int BY>
__device__ __forceinline__ void apply_B_partial(cplx<R> (&v)[2 * BY], int nfac, const Coef<R, BY>& cf, const Hop<R>& hp, cplx<R>* lds,
                                                const Geo& g, int& phase) {
    for (int k = 0; k < nfac; ++k) {
        switch (seq_group(k)) {
            case 1: factor<R, BY, 1>(v, cf, hp, lds, g, phase); break;
            case 2: factor<R, BY, 2>(v, cf, hp, lds, g, phase); break;
            case 3: factor<R, BY, 3>(v, cf, hp, lds, g, phase); break;
            default: factor<R, BY, 4>(v, cf, hp, lds, g, phase); break;
        }
    }
}

// Patch load / store: row r holds two adjacent sites -> one 16 B (c64) or
// 2 x 16 B (c128) access per row, adjacent threads adjacent in memory.
template <typename R, int BY>
__device__ __forceinline__ void load_patch(cplx<R> (&v)[2 * BY], const cplx<R>* __restrict__ s, const Geo& g) {
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const cplx<R>* p = s + (g.y0 + r) * g.L + 2 * g.a;
        v[2 * r] = p[0];
        v[2 * r + 1] = p[1];
    }
}

template <typename R, int BY>
__device__ __forceinline__ void store_patch(cplx<R>* __restrict__ s, const cplx<R> (&v)[2 * BY], const Geo& g) {
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        cplx<R>* p = s + (g.y0 + r) * g.L + 2 * g.a;
        p[0] = v[2 * r];
        p[1] = v[2 * r + 1];
    }
}

}  // namespace qf
