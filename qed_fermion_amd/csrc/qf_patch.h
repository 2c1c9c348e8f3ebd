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

// The work-group's region is H rows of the lattice (all L, or a band), L sites
// wide, periodic in x; its rows wrap within the region, so for a band the
// partners of its edge rows are wrong and the values they reach are only kept
// where the band's own rows lie far enough inside (the halo, see RowBand).
struct Geo {
    int L;       // lattice extent (even): row length and x period
    int H;       // rows of the region (L, or a band with its halo)
    int a, b;    // patch coordinates: a in [0, L/2), b in [0, H/BY)
    int nb;      // H / BY
    int y0;      // b * BY, row in the region
    int gy0;     // the lattice row of y0
    bool valid;  // false for padding lanes of the last wave
    bool own;    // valid and in the region's own rows: stores and partial sums
};

// A band: region rows [row0, row0 + H) of the lattice (mod L), of which rows
// [hal, hal + nown) are its own output (the rest is halo, recomputed by the
// neighbouring bands).  The whole lattice is row0 = 0, H = L, hal = 0, nown = L.
struct RowBand {
    int row0, H, hal, nown;
};

template <int BY>
__device__ __forceinline__ Geo make_geo(int L, int tid, const RowBand& bd) {
    Geo g;
    g.L = L;
    g.H = bd.H;
    const int na = L >> 1;
    g.nb = bd.H / BY;
    g.valid = tid < na * g.nb;
    // padding lanes mirror patch (0,0): they load and exchange the same values
    // as thread 0 (identical LDS writes are benign) and never store
    const int t = g.valid ? tid : 0;
    g.a = t % na;
    g.b = t / na;
    g.y0 = g.b * BY;
    const int gy = bd.row0 + g.y0;
    g.gy0 = gy >= L ? gy - L : gy;
    g.own = g.valid && g.y0 >= bd.hal && g.y0 < bd.hal + bd.nown;
    return g;
}

template <int BY>
__device__ __forceinline__ Geo make_geo(int L, int tid) {
    return make_geo<BY>(L, tid, RowBand{0, L, 0, L});
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

// phi_s: slice base, layout [y][x][mu] (R reals).  NORM: coefficients scaled
// by tanh(a_g) for the normalised factors I + th K; otherwise by sinh(a_g).
template <typename R, int BY, bool NORM = true>
__device__ __forceinline__ void load_coef(Coef<R, BY>& cf, const R* __restrict__ phi_s, const Geo& g, const Hop<R>& hp) {
    const int L = g.L;
    const cplx<R>* ps = reinterpret_cast<const cplx<R>*>(phi_s);  // one (phi_x, phi_y) pair per site
    const int x0 = 2 * g.a;
    const int xm = x0 == 0 ? L - 1 : x0 - 1;
    const R s1 = NORM ? hp.th1 : hp.sh1, s2 = NORM ? hp.th2 : hp.sh2;
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const int y = g.gy0 + r;
        const cplx<R> p0 = ps[y * L + x0];
        const cplx<R> p1 = ps[y * L + x0 + 1];
        const R pxm = phi_s[2 * (y * L + xm)];
        const R sh_in = (r & 1) ? s2 : s1;   // g1 (r even) or g3 (r odd)
        const R sh_ed = (r & 1) ? s1 : s2;
        cf.ex[r][0] = phase<R>(sh_ed, pxm);
        cf.ex[r][1] = phase<R>(sh_in, p0.x);
        cf.ex[r][2] = phase<R>(sh_ed, p1.x);
        cf.ey[0][r + 1] = phase<R>(s2, p0.y);
        cf.ey[1][r + 1] = phase<R>(s2, p1.y);
    }
    const int ym = g.gy0 == 0 ? L - 1 : g.gy0 - 1;
    const cplx<R> q0 = ps[ym * L + x0];
    const cplx<R> q1 = ps[ym * L + x0 + 1];
    cf.ey[0][0] = phase<R>(s2, q0.y);
    cf.ey[1][0] = phase<R>(s2, q1.y);
}

// Two adjacent sites (x even) as one 16-byte (complex64) or two 16-byte
// (complex128) accesses.
template <typename R> __device__ __forceinline__ void ld_pair(const cplx<R>* p, cplx<R>& a, cplx<R>& b);
template <> __device__ __forceinline__ void ld_pair<float>(const float2* p, float2& a, float2& b) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    a = make_float2(q.x, q.y);
    b = make_float2(q.z, q.w);
}
template <> __device__ __forceinline__ void ld_pair<double>(const double2* p, double2& a, double2& b) {
    a = p[0];
    b = p[1];
}
template <typename R> __device__ __forceinline__ void st_pair(cplx<R>* p, cplx<R> a, cplx<R> b);
template <> __device__ __forceinline__ void st_pair<float>(float2* p, float2 a, float2 b) {
    *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
}
template <> __device__ __forceinline__ void st_pair<double>(double2* p, double2 a, double2 b) {
    p[0] = a;
    p[1] = b;
}

// Coefficients from a precomputed hop table hop[tau][y][x][mu] = th_g e^{i phi_mu}
// (built once per field by hop_table_kernel): no sincos in the operator.
template <typename R, int BY>
__device__ __forceinline__ void load_coef_tab(Coef<R, BY>& cf, const cplx<R>* __restrict__ hop_s, const Geo& g) {
    const int L = g.L;
    const int x0 = 2 * g.a;
    const int xm = x0 == 0 ? L - 1 : x0 - 1;
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const int y = g.gy0 + r;
        cplx<R> h0x, h0y, h1x, h1y;
        ld_pair<R>(hop_s + 2 * (y * L + x0), h0x, h0y);
        ld_pair<R>(hop_s + 2 * (y * L + x0 + 1), h1x, h1y);
        cf.ex[r][0] = hop_s[2 * (y * L + xm)];
        cf.ex[r][1] = h0x;
        cf.ex[r][2] = h1x;
        cf.ey[0][r + 1] = h0y;
        cf.ey[1][r + 1] = h1y;
    }
    const int ym = g.gy0 == 0 ? L - 1 : g.gy0 - 1;
    cf.ey[0][0] = hop_s[2 * (ym * L + x0) + 1];
    cf.ey[1][0] = hop_s[2 * (ym * L + x0 + 1) + 1];
}

// Per-lane byte offsets inside one slice for buffer addressing: the patch
// origin in a vector slice, and the hop-table entries a patch needs.  Row and
// slice offsets are uniform and go in the scalar offset.
struct PatchOff {
    unsigned v;    // vector: site (2a, y0)
    unsigned h;    // hop: sites (2a, y0) and (2a+1, y0), both mu
    unsigned hm;   // hop: x-entry of (xm, y0), xm = 2a-1 mod L
    unsigned hy;   // hop: y-entry of (2a, ym), ym = y0-1 mod L
};

template <typename R>
__device__ __forceinline__ PatchOff patch_off(const Geo& g) {
    constexpr unsigned S = sizeof(cplx<R>);
    const int L = g.L, x0 = 2 * g.a;
    const int xm = x0 == 0 ? L - 1 : x0 - 1;
    const int ym = g.gy0 == 0 ? L - 1 : g.gy0 - 1;
    PatchOff o;
    o.v = (unsigned)(g.gy0 * L + x0) * S;
    o.h = (unsigned)(g.gy0 * L + x0) * 2 * S;
    o.hm = (unsigned)(g.gy0 * L + xm) * 2 * S;
    o.hy = (unsigned)(ym * L + x0) * 2 * S + S;
    return o;
}

// Patch of one vector slice; so = byte offset of the slice.
template <typename R, int BY>
__device__ __forceinline__ void load_patch_buf(cplx<R> (&v)[2 * BY], __amdgpu_buffer_rsrc_t rs, const PatchOff& o, unsigned so, int L) {
    constexpr unsigned S = sizeof(cplx<R>);
#pragma unroll
    for (int r = 0; r < BY; ++r) buf_ld2<R>(rs, o.v, so + (unsigned)(r * L) * S, v[2 * r], v[2 * r + 1]);
}

// Store of the patch into one vector slice (slice: uniform base pointer).
template <typename R, int BY>
__device__ __forceinline__ void store_patch_g(const cplx<R> (&v)[2 * BY], cplx<R>* slice, const PatchOff& o, int L) {
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        cplx<R>* row = slice + r * L;   // uniform
        st_pair<R>(reinterpret_cast<cplx<R>*>(reinterpret_cast<char*>(row) + o.v), v[2 * r], v[2 * r + 1]);
    }
}

// Coefficients of one slice from the hop table; so = byte offset of the slice
// (tau * 2 * Vs * sizeof(cplx)).
template <typename R, int BY>
__device__ __forceinline__ void load_coef_buf(Coef<R, BY>& cf, __amdgpu_buffer_rsrc_t rs, const PatchOff& o, unsigned so, int L) {
    constexpr unsigned S = sizeof(cplx<R>);
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const unsigned row = so + (unsigned)(r * L) * 2 * S;
        cplx<R> h0y, h1y;
        buf_ld2<R>(rs, o.h, row, cf.ex[r][1], h0y);
        buf_ld2<R>(rs, o.h + 2 * S, row, cf.ex[r][2], h1y);
        cf.ex[r][0] = buf_ld1<R>(rs, o.hm, row);
        cf.ey[0][r + 1] = h0y;
        cf.ey[1][r + 1] = h1y;
    }
    cf.ey[0][0] = buf_ld1<R>(rs, o.hy, so);
    cf.ey[1][0] = buf_ld1<R>(rs, o.hy + 2 * S, so);
}

// LDS exchange buffer: two alternating halves, each NV * (L/2)*L complex
// (NV vectors exchanged together).  Callers size the dynamic LDS with
// lds_bytes<R>(L, NV).  Edge values are kept in four planes indexed by the
// patch column a, so a wave's writes and reads are contiguous (bank-conflict
// free); the NV vectors' values of one site share one slot of NV complex, so
// one 16-byte (NV = 2, complex64) LDS access moves both:
//   x-groups: EL[row][a] (cx=0 site), ER[row][a] (cx=1 site), row = y >> 1
//   y-groups: BOT[b][a] (r=0 site), TOP[b][a] (r=BY-1 site)
__host__ __device__ constexpr int xstride(int L, int H) { return (L >> 1) * H; }
template <typename R>
inline size_t lds_bytes(int L, int NV, int H = 0) { return (size_t)2 * NV * xstride(L, H > 0 ? H : L) * sizeof(cplx<R>); }

// Write / read one slot of NV complex values (vector n's value at slot[n]).
template <typename R, int NV>
struct SlotIO {
    template <typename F>
    __device__ __forceinline__ static void put(cplx<R>* slot, F&& val) {
#pragma unroll
        for (int n = 0; n < NV; ++n) slot[n] = val(n);
    }
    template <typename F>
    __device__ __forceinline__ static void get(const cplx<R>* slot, F&& sink) {
#pragma unroll
        for (int n = 0; n < NV; ++n) sink(n, slot[n]);
    }
};
template <>
struct SlotIO<float, 2> {
    template <typename F>
    __device__ __forceinline__ static void put(float2* slot, F&& val) {
        const float2 a = val(0), b = val(1);
        *reinterpret_cast<float4*>(slot) = make_float4(a.x, a.y, b.x, b.y);
    }
    template <typename F>
    __device__ __forceinline__ static void get(const float2* slot, F&& sink) {
        const float4 q = *reinterpret_cast<const float4*>(slot);
        sink(0, make_float2(q.x, q.y));
        sink(1, make_float2(q.z, q.w));
    }
};

// Partner values of the sites that cross a patch edge for group G.  Writes the
// patch's edge values, barriers, reads the neighbours' edge values.  pr[] gets
// the partner value for the edge sites (index: x-groups -> 2*j + {0: left
// partner of the cx=0 site, 1: right partner of the cx=1 site} for the j-th
// edge row; y-groups -> {0: partner below r=0, 1: partner above r=BY-1}).
template <typename R, int BY, int G, int NV>
__device__ __forceinline__ void exchange(cplx<R> (*v)[2 * BY], cplx<R> (*pr)[BY], cplx<R>* lds, const Geo& g, int& phase) {
    const int L = g.L;
    const int na = L >> 1;
    const int plane = xstride(L, g.H) >> 1;      // na * H/2 slots
    cplx<R>* buf = lds + (phase & 1) * NV * xstride(L, g.H);
    using IO = SlotIO<R, NV>;
    if constexpr (G == 1 || G == 3) {
        constexpr int rc = (G == 1) ? 1 : 0;   // edge rows of group G have r % 2 == rc
#pragma unroll
        for (int r = rc; r < BY; r += 2) {
            const int o = ((g.y0 + r) >> 1) * na + g.a;
            IO::put(buf + o * NV, [&](int n) { return v[n][2 * r]; });
            IO::put(buf + (plane + o) * NV, [&](int n) { return v[n][2 * r + 1]; });
        }
        lds_barrier();
        const int am = g.a == 0 ? na - 1 : g.a - 1, ap = g.a + 1 == na ? 0 : g.a + 1;
#pragma unroll
        for (int r = rc, j = 0; r < BY; r += 2, ++j) {
            const int rowo = ((g.y0 + r) >> 1) * na;
            IO::get(buf + (plane + rowo + am) * NV, [&](int n, cplx<R> c) { pr[n][2 * j] = c; });
            IO::get(buf + (rowo + ap) * NV, [&](int n, cplx<R> c) { pr[n][2 * j + 1] = c; });
        }
    } else {
        constexpr int cc = (G == 2) ? 1 : 0;   // edge column of group G
        const int yplane = g.nb * na;
        const int o = g.b * na + g.a;
        IO::put(buf + o * NV, [&](int n) { return v[n][cc]; });
        IO::put(buf + (yplane + o) * NV, [&](int n) { return v[n][2 * (BY - 1) + cc]; });
        lds_barrier();
        const int bm = g.b == 0 ? g.nb - 1 : g.b - 1, bp = g.b + 1 == g.nb ? 0 : g.b + 1;
        IO::get(buf + (yplane + bm * na + g.a) * NV, [&](int n, cplx<R> c) { pr[n][0] = c; });
        IO::get(buf + (bp * na + g.a) * NV, [&](int n, cplx<R> c) { pr[n][1] = c; });
    }
    ++phase;
}

// v <- D_G v given the edge partner values pr.  NORM: normalised factor
// (I + th K_G), coefficients scaled by tanh.  Otherwise the plain factor
// cosh + sinh K (SGN = +1) or its inverse cosh - sinh K (SGN = -1; K_g^2 = 1
// on a perfect matching), coefficients scaled by sinh.
template <typename R, int BY, int G, int SGN, bool NORM = false>
__device__ __forceinline__ void factor_update(cplx<R> (&v)[2 * BY], const cplx<R> (&pr)[BY], const Coef<R, BY>& cf, const Hop<R>& hp) {
    static_assert(!NORM || SGN > 0, "normalised factors are applied forward only");
    auto upd = [&](cplx<R> a, cplx<R> h, cplx<R> w) -> cplx<R> {
        if constexpr (NORM) return pair_add<R>(a, h, w);
        const R c = (G == 1) ? hp.ch1 : hp.ch2;
        const cplx<R> hh = SGN > 0 ? h : mk<R>(-h.x, -h.y);
        return pair_upd<R>(c, a, hh, w);
    };
    auto updc = [&](cplx<R> a, cplx<R> h, cplx<R> w) -> cplx<R> {
        if constexpr (NORM) return pair_addc<R>(a, h, w);
        const R c = (G == 1) ? hp.ch1 : hp.ch2;
        const cplx<R> hh = SGN > 0 ? h : mk<R>(-h.x, -h.y);
        return pair_updc<R>(c, a, hh, w);
    };
    if constexpr (G == 1 || G == 3) {
        constexpr int rc = (G == 1) ? 1 : 0;
#pragma unroll
        for (int r = 0, j = 0; r < BY; ++r) {
            if ((r & 1) != rc) {   // bond inside the patch, first site cx=0
                const cplx<R> v0 = v[2 * r], v1 = v[2 * r + 1];
                v[2 * r] = upd(v0, cf.ex[r][1], v1);
                v[2 * r + 1] = updc(v1, cf.ex[r][1], v0);
            } else {               // both sites on edge bonds
                v[2 * r] = updc(v[2 * r], cf.ex[r][0], pr[2 * j]);
                v[2 * r + 1] = upd(v[2 * r + 1], cf.ex[r][2], pr[2 * j + 1]);
                ++j;
            }
        }
    } else {
        constexpr int cc = (G == 2) ? 1 : 0;
        constexpr int ci = 1 - cc;
#pragma unroll
        for (int r = 0; r < BY; r += 2) {   // inside column: pairs (r, r+1), r even
            const cplx<R> v0 = v[2 * r + ci], v1 = v[2 * (r + 1) + ci];
            v[2 * r + ci] = upd(v0, cf.ey[ci][r + 1], v1);
            v[2 * (r + 1) + ci] = updc(v1, cf.ey[ci][r + 1], v0);
        }
#pragma unroll
        for (int r = 1; r + 1 < BY; r += 2) {   // edge column: pairs (r, r+1), r odd
            const cplx<R> v0 = v[2 * r + cc], v1 = v[2 * (r + 1) + cc];
            v[2 * r + cc] = upd(v0, cf.ey[cc][r + 1], v1);
            v[2 * (r + 1) + cc] = updc(v1, cf.ey[cc][r + 1], v0);
        }
        v[cc] = updc(v[cc], cf.ey[cc][0], pr[0]);
        v[2 * (BY - 1) + cc] = upd(v[2 * (BY - 1) + cc], cf.ey[cc][BY], pr[1]);
    }
}

template <typename R, int BY, int G, bool NORM = true>
__device__ __forceinline__ void factor(cplx<R> (&v)[2 * BY], const Coef<R, BY>& cf, const Hop<R>& hp, cplx<R>* lds, const Geo& g,
                                       int& phase) {
    cplx<R> pr[1][BY];
    exchange<R, BY, G, 1>(reinterpret_cast<cplx<R>(*)[2 * BY]>(&v), pr, lds, g, phase);
    factor_update<R, BY, G, 1, NORM>(v, pr[0], cf, hp);
}

// Two independent factor applications sharing one exchange barrier:
// v1 <- D_G(c1) v1 and v2 <- D_G(c2) v2 (different slices' coefficients).
template <typename R, int BY, int G>
__device__ __forceinline__ void factor2(cplx<R> (&v1)[2 * BY], const Coef<R, BY>& c1, cplx<R> (&v2)[2 * BY], const Coef<R, BY>& c2,
                                        const Hop<R>& hp, cplx<R>* lds, const Geo& g, int& phase) {
    cplx<R> vv[2][2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) { vv[0][k] = v1[k]; vv[1][k] = v2[k]; }
    cplx<R> pr[2][BY];
    exchange<R, BY, G, 2>(vv, pr, lds, g, phase);
    factor_update<R, BY, G, 1, true>(v1, pr[0], c1, hp);
    factor_update<R, BY, G, 1, true>(v2, pr[1], c2, hp);
}

// v1 <- B(c1) v1 / cn and v2 <- B(c2) v2 / cn (normalised factors) with 7
// barriers instead of 14.
template <typename R, int BY>
__device__ __forceinline__ void apply_B2(cplx<R> (&v1)[2 * BY], const Coef<R, BY>& c1, cplx<R> (&v2)[2 * BY], const Coef<R, BY>& c2,
                                         const Hop<R>& hp, cplx<R>* lds, const Geo& g, int& phase) {
    factor2<R, BY, 4>(v1, c1, v2, c2, hp, lds, g, phase);
    factor2<R, BY, 3>(v1, c1, v2, c2, hp, lds, g, phase);
    factor2<R, BY, 2>(v1, c1, v2, c2, hp, lds, g, phase);
    factor2<R, BY, 1>(v1, c1, v2, c2, hp, lds, g, phase);
    factor2<R, BY, 2>(v1, c1, v2, c2, hp, lds, g, phase);
    factor2<R, BY, 3>(v1, c1, v2, c2, hp, lds, g, phase);
    factor2<R, BY, 4>(v1, c1, v2, c2, hp, lds, g, phase);
}

// v <- B_tau v (NORM = false) or B_tau v / cn (NORM = true, the default)
template <typename R, int BY, bool NORM = true>
__device__ __forceinline__ void apply_B(cplx<R> (&v)[2 * BY], const Coef<R, BY>& cf, const Hop<R>& hp, cplx<R>* lds, const Geo& g,
                                        int& phase) {
    factor<R, BY, 4, NORM>(v, cf, hp, lds, g, phase);
    factor<R, BY, 3, NORM>(v, cf, hp, lds, g, phase);
    factor<R, BY, 2, NORM>(v, cf, hp, lds, g, phase);
    factor<R, BY, 1, NORM>(v, cf, hp, lds, g, phase);
    factor<R, BY, 2, NORM>(v, cf, hp, lds, g, phase);
    factor<R, BY, 3, NORM>(v, cf, hp, lds, g, phase);
    factor<R, BY, 4, NORM>(v, cf, hp, lds, g, phase);
}

// v <- (first nfac normalised factors of B_tau) v, nfac in [0, 7] (uniform
// over the block); the caller multiplies by the product of their cosh.
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

// ---------------------------------------------------------------------------
// Compact form for the operator kernels: every bond coefficient is held once,
// by the patch that owns the bond's first site.  For an edge bond (i, j) the
// owner of i sends conj(h) u_i instead of u_i, so the partner only adds what
// it receives (u_j += conj(h) u_i) and needs no copy of h; the owner reads the
// raw u_j and applies h itself.  In an edge row of an x-group the cx=0 site is
// the second site of the bond from the left neighbour and the cx=1 site the
// first site of the bond to the right; in the edge column of a y-group the
// r=0 site is the second site of the bond from below and r=BY-1 the first
// site of the bond upwards.
//   ex[r][0] = th e^{i phi_x(2a,   y0+r)}   (bond inside the patch)
//   ex[r][1] = th e^{i phi_x(2a+1, y0+r)}   (bond to the right neighbour)
//   ey[c][r] = th e^{i phi_y(2a+c, y0+r)}   (bond upwards from row r)
template <typename R, int BY>
struct CoefC {
    cplx<R> ex[BY][2];
    cplx<R> ey[2][BY];
};

template <typename R, int BY>
__device__ __forceinline__ void load_coefc_buf(CoefC<R, BY>& cf, __amdgpu_buffer_rsrc_t rs, const PatchOff& o, unsigned so, int L) {
    constexpr unsigned S = sizeof(cplx<R>);
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const unsigned row = so + (unsigned)(r * L) * 2 * S;
        buf_ld2<R>(rs, o.h, row, cf.ex[r][0], cf.ey[0][r]);
        buf_ld2<R>(rs, o.h + 2 * S, row, cf.ex[r][1], cf.ey[1][r]);
    }
}

// From phi (sincos), scaled by tanh(a_g) for the normalised factors.
template <typename R, int BY>
__device__ __forceinline__ void load_coefc(CoefC<R, BY>& cf, const R* __restrict__ phi_s, const Geo& g, const Hop<R>& hp) {
    const int L = g.L;
    const cplx<R>* ps = reinterpret_cast<const cplx<R>*>(phi_s);
    const int x0 = 2 * g.a;
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const int y = g.gy0 + r;
        const cplx<R> p0 = ps[y * L + x0];
        const cplx<R> p1 = ps[y * L + x0 + 1];
        const R t_in = (r & 1) ? hp.th2 : hp.th1;   // g1 (r even) or g3 (r odd)
        const R t_ed = (r & 1) ? hp.th1 : hp.th2;
        cf.ex[r][0] = phase<R>(t_in, p0.x);
        cf.ex[r][1] = phase<R>(t_ed, p1.x);
        cf.ey[0][r] = phase<R>(hp.th2, p0.y);
        cf.ey[1][r] = phase<R>(hp.th2, p1.y);
    }
}

// The same coefficients from the raw angles of the patch, q[2r] = (phi_x, phi_y) of site
// (2a, y0+r) and q[2r+1] those of (2a+1, y0+r): the layout load_patch_buf gives when it
// reads a phi slice, so the angles can be loaded one step before the sincos is taken.
template <typename R, int BY>
__device__ __forceinline__ void coefc_from_angles(CoefC<R, BY>& cf, const cplx<R> (&q)[2 * BY], const Hop<R>& hp) {
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const R t_in = (r & 1) ? hp.th2 : hp.th1;
        const R t_ed = (r & 1) ? hp.th1 : hp.th2;
        cf.ex[r][0] = phase<R>(t_in, q[2 * r].x);
        cf.ex[r][1] = phase<R>(t_ed, q[2 * r + 1].x);
        cf.ey[0][r] = phase<R>(hp.th2, q[2 * r].y);
        cf.ey[1][r] = phase<R>(hp.th2, q[2 * r + 1].y);
    }
}

// conj(h) * u
template <typename R> __device__ __forceinline__ cplx<R> cjmul_h(cplx<R> h, cplx<R> u) {
    return mk<R>(fma(h.x, u.x, h.y * u.y), fma(h.x, u.y, -h.y * u.x));
}

// Lane-indexed exchange layout of the compact factors: a phase buffer holds
// NPL = max(BY, 2) planes of nl slots (nl = lanes of the work-group), slot =
// patch index (= threadIdx.x for valid lanes), NV complex per slot.  A
// wave's writes are then one contiguous run and its reads a run shifted by
// one patch (or one patch row), so LDS accesses are free of bank conflicts.
//   x-group, j-th edge row: plane 2j = EL (cx=0 site, raw),
//                           plane 2j+1 = ER (cx=1 site, premultiplied)
//   y-group: plane 0 = BOT (r=0 site, raw), plane 1 = TOP (r=BY-1, premultiplied)
// Both phase buffers are sized for NV = 2 (lds_bytes_c), 2*NPL*nl slots apart.
__host__ __device__ constexpr int npl_of(int BY) { return BY > 2 ? BY : 2; }
template <typename R>
inline size_t lds_bytes_c(int BY, int nlanes) { return (size_t)2 * npl_of(BY) * nlanes * 2 * sizeof(cplx<R>); }

struct XAddr {
    int me;            // own slot
    int l, r, d, u;    // slots of the left / right / lower / upper neighbour patches
    int nl;            // slots per plane
};

// base: first slot of the lane's slice (a work-group holding several slices side by side
// gives each its own range of slots)
template <int BY>
__device__ __forceinline__ XAddr make_xaddr(const Geo& g, int base = 0) {
    const int na = g.L >> 1;
    XAddr x;
    x.me = threadIdx.x;
    x.nl = blockDim.x;
    const int am = g.a == 0 ? na - 1 : g.a - 1, ap = g.a + 1 == na ? 0 : g.a + 1;
    const int bm = g.b == 0 ? g.nb - 1 : g.b - 1, bp = g.b + 1 == g.nb ? 0 : g.b + 1;
    x.l = base + g.b * na + am;
    x.r = base + g.b * na + ap;
    x.d = base + bm * na + g.a;
    x.u = base + bp * na + g.a;
    return x;
}

// One normalised factor (I + th K_G) on NV vectors with coefficient sets
// cf[n], one exchange barrier for all of them.  Inside pairs are updated
// between the edge writes and the barrier.
template <typename R, int BY, int G, int NV>
__device__ __forceinline__ void factor_c(cplx<R> (*v)[2 * BY], const CoefC<R, BY>* const* cf, cplx<R>* lds, const XAddr& xa, int& phase) {
    static_assert(NV == 1 || NV == 2, "one or two vectors per exchange");
    constexpr int NPL = npl_of(BY);
    cplx<R>* buf = lds + (phase & 1) * (NPL * 2) * xa.nl;
    auto at = [&](int plane, int slot) { return buf + (plane * xa.nl + slot) * NV; };
    using IO = SlotIO<R, NV>;
    if constexpr (G == 1 || G == 3) {
        constexpr int rc = (G == 1) ? 1 : 0;   // edge rows have r % 2 == rc
#pragma unroll
        for (int r = rc, j = 0; r < BY; r += 2, ++j) {
            IO::put(at(2 * j, xa.me), [&](int n) { return v[n][2 * r]; });
            IO::put(at(2 * j + 1, xa.me), [&](int n) { return cjmul_h<R>(cf[n]->ex[r][1], v[n][2 * r + 1]); });
        }
#pragma unroll
        for (int n = 0; n < NV; ++n)
#pragma unroll
            for (int r = 1 - rc; r < BY; r += 2) {
                const cplx<R> v0 = v[n][2 * r], v1 = v[n][2 * r + 1], h = cf[n]->ex[r][0];
                v[n][2 * r] = pair_add<R>(v0, h, v1);
                v[n][2 * r + 1] = pair_addc<R>(v1, h, v0);
            }
        lds_barrier();
#pragma unroll
        for (int r = rc, j = 0; r < BY; r += 2, ++j) {
            IO::get(at(2 * j + 1, xa.l), [&](int n, cplx<R> p) { v[n][2 * r] = cadd<R>(v[n][2 * r], p); });
            IO::get(at(2 * j, xa.r), [&](int n, cplx<R> w) { v[n][2 * r + 1] = pair_add<R>(v[n][2 * r + 1], cf[n]->ex[r][1], w); });
        }
    } else {
        constexpr int cc = (G == 2) ? 1 : 0;   // edge column
        constexpr int ci = 1 - cc;             // inside column
        IO::put(at(0, xa.me), [&](int n) { return v[n][cc]; });
        IO::put(at(1, xa.me), [&](int n) { return cjmul_h<R>(cf[n]->ey[cc][BY - 1], v[n][2 * (BY - 1) + cc]); });
#pragma unroll
        for (int n = 0; n < NV; ++n) {
#pragma unroll
            for (int r = 0; r < BY; r += 2) {   // inside column: pairs (r, r+1), r even
                const cplx<R> v0 = v[n][2 * r + ci], v1 = v[n][2 * (r + 1) + ci], h = cf[n]->ey[ci][r];
                v[n][2 * r + ci] = pair_add<R>(v0, h, v1);
                v[n][2 * (r + 1) + ci] = pair_addc<R>(v1, h, v0);
            }
#pragma unroll
            for (int r = 1; r + 1 < BY; r += 2) {   // edge column: pairs (r, r+1), r odd
                const cplx<R> v0 = v[n][2 * r + cc], v1 = v[n][2 * (r + 1) + cc], h = cf[n]->ey[cc][r];
                v[n][2 * r + cc] = pair_add<R>(v0, h, v1);
                v[n][2 * (r + 1) + cc] = pair_addc<R>(v1, h, v0);
            }
        }
        lds_barrier();
        IO::get(at(1, xa.d), [&](int n, cplx<R> p) { v[n][cc] = cadd<R>(v[n][cc], p); });
        IO::get(at(0, xa.u), [&](int n, cplx<R> w) {
            v[n][2 * (BY - 1) + cc] = pair_add<R>(v[n][2 * (BY - 1) + cc], cf[n]->ey[cc][BY - 1], w);
        });
    }
    ++phase;
}

// v1 <- B(c1) v1 / cn and v2 <- B(c2) v2 / cn with 7 barriers.
template <typename R, int BY>
__device__ __forceinline__ void apply_B2c(cplx<R> (&v1)[2 * BY], const CoefC<R, BY>& c1, cplx<R> (&v2)[2 * BY], const CoefC<R, BY>& c2,
                                          cplx<R>* lds, const XAddr& g, int& phase) {
    cplx<R> vv[2][2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) { vv[0][k] = v1[k]; vv[1][k] = v2[k]; }
    const CoefC<R, BY>* cf[2] = {&c1, &c2};
    factor_c<R, BY, 4, 2>(vv, cf, lds, g, phase);
    factor_c<R, BY, 3, 2>(vv, cf, lds, g, phase);
    factor_c<R, BY, 2, 2>(vv, cf, lds, g, phase);
    factor_c<R, BY, 1, 2>(vv, cf, lds, g, phase);
    factor_c<R, BY, 2, 2>(vv, cf, lds, g, phase);
    factor_c<R, BY, 3, 2>(vv, cf, lds, g, phase);
    factor_c<R, BY, 4, 2>(vv, cf, lds, g, phase);
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) { v1[k] = vv[0][k]; v2[k] = vv[1][k]; }
}

// v <- B(c) v / cn (one vector)
template <typename R, int BY>
__device__ __forceinline__ void apply_Bc(cplx<R> (&v)[2 * BY], const CoefC<R, BY>& c, cplx<R>* lds, const XAddr& g, int& phase) {
    cplx<R> vv[1][2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) vv[0][k] = v[k];
    const CoefC<R, BY>* cf[1] = {&c};
    factor_c<R, BY, 4, 1>(vv, cf, lds, g, phase);
    factor_c<R, BY, 3, 1>(vv, cf, lds, g, phase);
    factor_c<R, BY, 2, 1>(vv, cf, lds, g, phase);
    factor_c<R, BY, 1, 1>(vv, cf, lds, g, phase);
    factor_c<R, BY, 2, 1>(vv, cf, lds, g, phase);
    factor_c<R, BY, 3, 1>(vv, cf, lds, g, phase);
    factor_c<R, BY, 4, 1>(vv, cf, lds, g, phase);
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) v[k] = vv[0][k];
}

// Patch load / store: row r holds two adjacent sites -> one 16 B (c64) or
// 2 x 16 B (c128) access per row, adjacent threads adjacent in memory.
template <typename R, int BY>
__device__ __forceinline__ void load_patch(cplx<R> (&v)[2 * BY], const cplx<R>* __restrict__ s, const Geo& g) {
    const cplx<R>* p = s + g.gy0 * g.L + 2 * g.a;
#pragma unroll
    for (int r = 0; r < BY; ++r) ld_pair<R>(p + r * g.L, v[2 * r], v[2 * r + 1]);
}

template <typename R, int BY>
__device__ __forceinline__ void store_patch(cplx<R>* __restrict__ s, const cplx<R> (&v)[2 * BY], const Geo& g) {
    cplx<R>* p = s + g.gy0 * g.L + 2 * g.a;
#pragma unroll
    for (int r = 0; r < BY; ++r) st_pair<R>(p + r * g.L, v[2 * r], v[2 * r + 1]);
}

// Hop table of one field: hop[tau][y][x][mu] = th_g(x,y,mu) e^{i phi_mu(x,y)},
// th = tanh(t dtau) for the x-bonds of even sites (group 1), tanh(t dtau/2)
// otherwise.  Accurate sincos: built once per field, read by every iteration.
// The body covers entries i0, i0 + stride, ... (a grid-stride loop).
template <typename R>
__device__ __forceinline__ void hop_table_body(const R* __restrict__ phi, long long phi_cs, cplx<R>* __restrict__ hop, int nchain, int L,
                                               int Lt, double th1, double th2, long long i0, long long stride) {
    const int Vs = L * L;
    const long long n = (long long)nchain * Lt * Vs;
    for (long long i = i0; i < n; i += stride) {
        const int s = (int)(i % Vs);
        const long long q = i / Vs;            // chain * Lt + tau
        const int chain = (int)(q / Lt), tau = (int)(q % Lt);
        const int x = s % L, y = s / L;
        const R* ph = phi + chain * phi_cs + ((long long)tau * Vs + s) * 2;
        cplx<R>* h = hop + (q * Vs + s) * 2;
        R sx, cx, sy, cy;
        if constexpr (sizeof(R) == 4) {
            sincosf(ph[0], &sx, &cx);
            sincosf(ph[1], &sy, &cy);
        } else {
            sincos(ph[0], &sx, &cx);
            sincos(ph[1], &sy, &cy);
        }
        const R tx = ((x + y) & 1) ? (R)th2 : (R)th1;
        h[0] = mk<R>(tx * cx, tx * sx);
        h[1] = mk<R>((R)th2 * cy, (R)th2 * sy);
    }
}

template <typename R>
__global__ void hop_table_kernel(const R* __restrict__ phi, long long phi_cs, cplx<R>* __restrict__ hop, int nchain, int L, int Lt,
                                 double th1, double th2) {
    hop_table_body<R>(phi, phi_cs, hop, nchain, L, Lt, th1, th2, blockIdx.x * (long long)blockDim.x + threadIdx.x,
                      (long long)gridDim.x * blockDim.x);
}

// Hop tables of a ragged batch (every lattice's table in one launch; site i of
// the concatenated vectors belongs to the last lattice with voff <= i).
template <typename R>
__device__ __forceinline__ void hop_table_ragged_body(const R* __restrict__ phi, cplx<R>* __restrict__ hop, const Prob* __restrict__ prob,
                                                      int nprob, long long nsites, double th1, double th2, long long i0, long long stride) {
    for (long long i = i0; i < nsites; i += stride) {
        int k = 0;
        for (int j = 1; j < nprob; ++j)
            if (prob[j].voff <= i) k = j;
        const Prob& P = prob[k];
        const long long loc = i - P.voff;
        const int Vs = P.L * P.L;
        const int s = (int)(loc % Vs);
        const int x = s % P.L, y = s / P.L;
        const R* ph = phi + P.poff + 2 * loc;
        cplx<R>* h = hop + P.hoff + 2 * loc;
        R sx, cx, sy, cy;
        if constexpr (sizeof(R) == 4) {
            sincosf(ph[0], &sx, &cx);
            sincosf(ph[1], &sy, &cy);
        } else {
            sincos(ph[0], &sx, &cx);
            sincos(ph[1], &sy, &cy);
        }
        const R tx = ((x + y) & 1) ? (R)th2 : (R)th1;
        h[0] = mk<R>(tx * cx, tx * sx);
        h[1] = mk<R>((R)th2 * cy, (R)th2 * sy);
    }
}

template <typename R>
__global__ void hop_table_ragged_kernel(const R* __restrict__ phi, cplx<R>* __restrict__ hop, const Prob* __restrict__ prob, int nprob,
                                        long long nsites, double th1, double th2) {
    hop_table_ragged_body<R>(phi, hop, prob, nprob, nsites, th1, th2, blockIdx.x * (long long)blockDim.x + threadIdx.x,
                             (long long)gridDim.x * blockDim.x);
}

}  // namespace qf
