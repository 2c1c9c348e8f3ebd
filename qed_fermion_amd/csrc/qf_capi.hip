// qf_capi.hip -- C-ABI of the MI355X HMC fermion library (libqf_hip.so).
//
// Every entry point takes raw device pointers, sizes and a hipStream_t passed
// as void*, launches on that stream only, never synchronises the host and
// never allocates, so any sequence of calls can be captured into a hipGraph by
// the caller.  Errors are returned as status codes; qf_last_error() gives the
// message of the calling thread's last failure.  Declarations and the
// reference interface each one replaces: include/qed_fermion_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/qed_fermion_amd.h"
#include "qf_boson.h"
#include "qf_cg.h"
#include "qf_force.h"
#include "qf_obs.h"
#include "qf_operator.h"
#include "qf_persist.h"

using namespace qf;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define QF_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail(QF_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define QF_LAUNCHED()                                                                          \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess) return fail(QF_ERR_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- tiling
struct Tuning {
    int by = 0, t = 0, tb = 0;   // 0 = automatic
    int ns = 0;                  // right-hand sides per work-group of the plain M^dag M from phi (0: one)
};
Tuning g_tune[2];                // per precision: applies to every shape
// Per-shape entries (set by the autotuner); nrhs = 0 matches any batch.
struct ShapeTuning {
    int prec, L, Lt, nrhs;
    Tuning tu;
};
std::vector<ShapeTuning> g_shape_tune;
std::mutex g_tune_mu;
std::atomic<int> g_paired_x{1};   // qf_set_paired_x_updates
std::atomic<long long> g_pers_timeout_us{10000000};   // qf_set_persistent_timeout_us

// Work-groups of up to 512 threads get 256 VGPRs per lane; only the 2x2
// patch may use 1024-thread work-groups (L up to 64).
inline int launch_bound(int by) { return by == 2 ? 1024 : 512; }
inline int nthreads(int L, int by) { return (L / 2) * (L / by); }
inline bool by_ok(int L, int by) { return by > 0 && L % by == 0 && by % 2 == 0 && nthreads(L, by) <= launch_bound(by); }

struct Tiling {
    int by, T, nA;     // operator kernels
    int TB, nsx, nB;   // update kernel: tile of TB slices x cgu_tx(TB) sites
    int ns;            // right-hand sides per work-group of the plain M^dag M from phi
};

// Update-kernel tiles: 512 threads over TB in {16, 32, 64, 128} tau slices of
// cgu_tx(TB) sites (32 sites for TB = 16, else 16 sites: 128-byte rows in
// complex64); TB >= 64 gives each thread TB / 32 slices.
inline bool tb_ok(int tb) { return tb == 16 || tb == 32 || tb == 64 || tb == 128; }
inline int cgu_tx(int tb) { return tb == 16 ? 32 : 16; }

// Tuning that applies to (prec, L, Lt, nrhs): exact shape, then any batch
// size, then the per-precision setting.
Tuning find_tuning(int prec, int L, int Lt, int nrhs) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const ShapeTuning* any = nullptr;
    for (const auto& e : g_shape_tune) {
        if (e.prec != prec || e.L != L || e.Lt != Lt) continue;
        if (e.nrhs == nrhs) return e.tu;
        if (e.nrhs == 0) any = &e;
    }
    return any ? any->tu : g_tune[prec];
}

// Default patch height: complex64 kernels fit the 2x4 patch in 256 VGPRs
// without spilling (more independent work per exchange barrier); complex128
// kernels only fit the 2x2 patch.  Patches that need the 1024-thread
// work-group come last.
int pick_by_default(int L, int prec) {
    const int order32[3] = {4, 2, 6}, order64[3] = {2, 4, 6};
    const int* order = prec == QF_PREC_F64 ? order64 : order32;
    for (int k = 0; k < 3; ++k)
        if (by_ok(L, order[k]) && nthreads(L, order[k]) <= 512) return order[k];
    for (int k = 0; k < 3; ++k)
        if (by_ok(L, order[k])) return order[k];
    return 0;
}

int pick_by(int L, int prec, int want = 0) {
    if (want && by_ok(L, want)) return want;
    return pick_by_default(L, prec);
}

// Lattices too wide for a whole slice per work-group (no patch shape fits) are
// split into bands of rows: 2x2 patches, up to 1024 threads, each work-group
// computing its band plus `hal` halo rows on both sides (4 rows for one B
// application, 8 for M^dag M and the force, whose factor sweeps reach twice
// as far), with nband bands of nown rows each (the last one shorter).
struct BandCfg {
    int nband = 1, H = 0, hal = 0, nown = 0, nthr = 0;
};
inline bool square_ok(int L) { return pick_by_default(L, 0) != 0; }
// A launch with 2x2 patches on a lattice whose whole slice of 2x2 patches exceeds a
// work-group runs in bands (chosen by the tiling when no patch shape fits, or
// requested as patch height 2 through the tuning).
inline bool band_mode(int L, int by) { return by == 2 && !by_ok(L, 2); }
constexpr int kHalB = 4, kHalMdm = 8;

BandCfg band_cfg(int L, int hal) {
    BandCfg c;
    const int half = L / 2;
    const int rows_max = 2 * (1024 / half);          // band rows a 1024-thread work-group holds
    const int own_max = rows_max - 2 * hal;
    if (own_max < 2) return c;                        // nband = 1, nthr = 0: does not fit
    c.nband = (L + own_max - 1) / own_max;
    c.nown = 2 * ((L + 2 * c.nband - 1) / (2 * c.nband));
    c.nband = (L + c.nown - 1) / c.nown;
    c.hal = hal;
    c.H = c.nown + 2 * hal;
    c.nthr = half * (c.H / 2);
    return c;
}

template <typename A>
void set_band(A& a, const BandCfg& c) {
    a.nband = c.nband; a.H = c.H; a.hal = c.hal; a.nown = c.nown;
}

// Right-hand sides per work-group of the plain M^dag M from phi (whole slices): up to `want`,
// as many slices as fit one work-group (1024 threads with 2 x 2 patches, else 512), and the
// group's vectors within the 32-bit offsets of its loads.
int rhs_per_group(int L, int Lt, int by, int prec, int nrhs, int want) {
    if (want <= 1 || nrhs <= 1 || by <= 0 || !by_ok(L, by)) return 1;
    const int nts = nthreads(L, by);
    const long long vbytes = (long long)L * L * Lt * (prec == QF_PREC_F64 ? 16 : 8);
    int ns = std::min(std::min(want, nrhs), (by == 2 ? 1024 : 512) / nts);
    while (ns > 1 && ns * vbytes >= (1LL << 31)) --ns;
    return std::max(ns, 1);
}

Tiling tiling(int prec, int nrhs, int L, int Lt) {
    Tiling t;
    const Tuning tu = find_tuning(prec, L, Lt, nrhs);
    const bool band = !square_ok(L) || band_mode(L, tu.by);
    t.by = band ? 2 : pick_by(L, prec, tu.by);
    const long long units = (long long)nrhs * Lt;
    // tau slices per operator work-group (untuned), where the autotuner lands: about 512
    // work-groups over the batch (two per CU; 8 chains at L=36, Ltau=360: 6 slices, 32 chains:
    // 23), 256 for bands (one work-group of 9 or more waves per CU; one chain at L=66: 6), at
    // most 24 slices.  Small batches get runs down to one slice: their launches are
    // latency-bound, and the extra B application of each run's prologue costs less than
    // marching (one FP64 chain at L=24, Ltau=240: 1 slice, 35% faster per trajectory than 4).
    const long long nbands = band ? band_cfg(L, kHalMdm).nband : 1;
    const long long target = band ? 256 : 512;
    t.T = tu.t ? tu.t : (int)std::max(1LL, std::min(24LL, (units * nbands + target - 1) / target));
    t.T = std::min(t.T, Lt);
    t.nA = (Lt + t.T - 1) / t.T;
    if (band) t.nA *= band_cfg(L, kHalMdm).nband;   // one partial sum per (run, band)
    const int Vs = L * L;
    // update kernel tile: TB tau slices x cgu_tx(TB) sites; the 12-slice
    // preconditioner halo costs 12 / TB extra reads of r and M^dag M p
    t.TB = tu.tb ? tu.tb : (Lt >= 128 ? 128 : Lt >= 64 ? 64 : Lt > 16 ? 32 : 16);
    t.nsx = (Vs + cgu_tx(t.TB) - 1) / cgu_tx(t.TB);
    t.nB = t.nsx * ((Lt + t.TB - 1) / t.TB);
    t.ns = band ? 1 : rhs_per_group(L, Lt, t.by, prec, nrhs, tu.ns);
    (void)units;
    return t;
}

template <typename R>
Hop<R> make_hop(double dtau, double t) {
    Hop<R> h;
    h.ch1 = (R)std::cosh(t * dtau);
    h.sh1 = (R)std::sinh(t * dtau);
    h.ch2 = (R)std::cosh(0.5 * t * dtau);
    h.sh2 = (R)std::sinh(0.5 * t * dtau);
    h.th1 = (R)std::tanh(t * dtau);
    h.th2 = (R)std::tanh(0.5 * t * dtau);
    h.cn = (R)(std::cosh(t * dtau) * std::pow(std::cosh(0.5 * t * dtau), 6));
    h.dtau_ = dtau;
    h.t_ = t;
    return h;
}

// Product of cosh(a_g) over the first nfac factors of D4 D3 D2 D1 D2 D3 D4.
double cosh_prefix(int nfac, double dtau, double t) {
    double c = 1.0;
    for (int k = 0; k < nfac; ++k) c *= (k == 3) ? std::cosh(t * dtau) : std::cosh(0.5 * t * dtau);
    return c;
}

int check_lattice(int L, int Lt, int nrhs) {
    if (L < 2 || (L & 1)) return fail(QF_ERR_ODD_L, "L=%d: the checkerboard bond groups need an even L >= 2", L);
    if (Lt < 1) return fail(QF_ERR_INVALID, "Ltau=%d must be >= 1", Lt);
    if (nrhs < 1 || nrhs > 65535) return fail(QF_ERR_INVALID, "nrhs=%d out of range [1, 65535]", nrhs);
    if (!square_ok(L) && band_cfg(L, kHalMdm).nthr == 0)
        return fail(QF_ERR_UNSUPPORTED, "L=%d: no band of 2x2 patches with its halo fits a 1024-thread work-group (L <= 226)", L);
    return QF_OK;
}

// Raise a kernel's dynamic-LDS limit above 64 KB once (not a stream operation,
// so it is also safe while the caller captures a graph).
template <typename KernelT>
int set_lds(KernelT kern, size_t bytes) {
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> done;
    if (bytes <= 65536) return QF_OK;
    const void* key = reinterpret_cast<const void*>(kern);
    std::lock_guard<std::mutex> lk(mu);
    for (auto& d : done)
        if (d.first == key && d.second >= bytes) return QF_OK;
    hipError_t e = hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return fail(QF_ERR_HIP, "LDS of %zu bytes refused: %s", bytes, hipGetErrorString(e));
    done.emplace_back(key, bytes);
    return QF_OK;
}

inline dim3 blk_of(int L, int by) { return dim3((unsigned)(((nthreads(L, by) + 63) / 64) * 64)); }

// ---------------------------------------------------------------- launchers
// Pipelined (two vectors per exchange) where the registers allow it:
// complex64 with 2 x 2 and 2 x 4 patches; complex128 and 2 x 6 run one vector
// per exchange.
template <typename R, int BY>
constexpr bool mdm_pipe() { return sizeof(R) == 4 && BY <= 4; }

template <typename R, int BY, int MODE, bool HOPT>
int launch_mdm_k(const MdmArgs<R>& a0, int nrhs, hipStream_t s) {
    MdmArgs<R> a = a0;
    a.nrhs = nrhs;
    constexpr bool PIPE = mdm_pipe<R, BY>();
    dim3 blk = blk_of(a.L, BY);
    a.nband = 1;
    if (band_mode(a.L, BY)) {   // bands of 2x2 patches
        const BandCfg c = band_cfg(a.L, kHalMdm);
        set_band(a, c);
        blk = dim3((unsigned)(((c.nthr + 63) / 64) * 64));
    }
    // several right-hand sides per work-group (plain M^dag M from phi, whole slices)
    int ns = 1;
    if constexpr (MODE == MDM_PLAIN && PIPE && !HOPT) {
        if (a.nsub > 1 && a.nband == 1 && a.prob == nullptr && a.dot_part == nullptr) {   // partials are per (rhs, run)
            const int nts = nthreads(a.L, BY);
            const long long vbytes = (long long)a.vs * (long long)sizeof(cplx<R>);
            ns = std::min(std::min(a.nsub, nrhs), (BY == 2 ? 1024 : 512) / nts);
            while (ns > 1 && (ns * vbytes >= (1LL << 31) || ns * a.phi_cs * (long long)sizeof(R) >= (1LL << 31) ||
                              (long long)a.L * a.L * a.Lt * 2 * (long long)sizeof(R) >= (1LL << 31))) --ns;
            if (ns > 1) blk = dim3((unsigned)(((ns * nts + 63) / 64) * 64));
        }
    }
    a.nsub = ns;
    const dim3 grid((unsigned)(((a.Lt + a.T - 1) / a.T) * a.nband * ((nrhs + ns - 1) / ns)));
    const size_t lds = lds_bytes_c<R>(BY, (int)blk.x);
    if (blk.x <= 512) {
        auto k = mdagm_kernel<R, BY, MODE, 512, HOPT, PIPE>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    } else if constexpr (BY == 2) {
        auto k = mdagm_kernel<R, BY, MODE, 1024, HOPT, PIPE>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    }
    QF_LAUNCHED();
    return QF_OK;
}

// CG step A: complex64 computes its coefficients from phi (hardware sincos:
// measured about 5% faster per iteration than reading the 16-byte-per-site hop
// table), except with 2 x 6 patches, whose from-phi form spills 66 VGPRs (22
// from the table: the L=66 trajectory 5.6% faster); complex128 reads the solve's
// hop table.  The FP64 residual of the reliable updates reads phi; plain
// applications take either.
template <typename R> constexpr bool cg_uses_hop(int by) { return sizeof(R) == 8 || by == 6; }

template <typename R, int BY, int MODE>
int launch_mdm_by(const MdmArgs<R>& a, int nrhs, hipStream_t s) {
    if constexpr (MODE == MDM_CG) {
        if constexpr (sizeof(R) == 4 && BY != 6) return launch_mdm_k<R, BY, MODE, false>(a, nrhs, s);
        if (!a.hop) return fail(QF_ERR_INVALID, "CG step without a hop table");
        return launch_mdm_k<R, BY, MODE, true>(a, nrhs, s);
    } else if constexpr (MODE == MDM_RESID) {
        return launch_mdm_k<R, BY, MODE, false>(a, nrhs, s);
    } else {
        return a.hop ? launch_mdm_k<R, BY, MODE, true>(a, nrhs, s) : launch_mdm_k<R, BY, MODE, false>(a, nrhs, s);
    }
}

template <typename R, int MODE>
int launch_mdm(const MdmArgs<R>& a, int by, int nrhs, hipStream_t s) {
    switch (by) {
        case 2: return launch_mdm_by<R, 2, MODE>(a, nrhs, s);
        case 4: return launch_mdm_by<R, 4, MODE>(a, nrhs, s);
        case 6: return launch_mdm_by<R, 6, MODE>(a, nrhs, s);
    }
    return fail(QF_ERR_UNSUPPORTED, "patch height %d", by);
}

template <typename R, int BY, int OP>
int launch_op_by(const OpArgs<R>& a0, int nrhs, hipStream_t s) {
    OpArgs<R> a = a0;
    dim3 blk = blk_of(a.L, BY);
    BandCfg c;
    c.H = a.L;
    if (band_mode(a.L, BY)) {
        c = band_cfg(a.L, kHalB);
        blk = dim3((unsigned)(((c.nthr + 63) / 64) * 64));
    }
    set_band(a, c);
    const size_t lds = lds_bytes<R>(a.L, 1, c.H);
    const dim3 grid((a.Lt + a.T - 1) / a.T, nrhs, c.nband);
    if (blk.x <= 512) {
        auto k = op_kernel<R, BY, OP, 512>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    } else if constexpr (BY == 2) {
        auto k = op_kernel<R, BY, OP, 1024>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    }
    QF_LAUNCHED();
    return QF_OK;
}

template <typename R, int OP>
int launch_op(const OpArgs<R>& a, int by, int nrhs, hipStream_t s) {
    switch (by) {
        case 2: return launch_op_by<R, 2, OP>(a, nrhs, s);
        case 4: return launch_op_by<R, 4, OP>(a, nrhs, s);
        case 6: return launch_op_by<R, 6, OP>(a, nrhs, s);
    }
    return fail(QF_ERR_UNSUPPORTED, "patch height %d", by);
}

template <typename R, int BY>
int launch_force_by(const ForceArgs<R>& a0, int nrhs, hipStream_t s) {
    ForceArgs<R> a = a0;
    dim3 blk = blk_of(a.L, BY);
    BandCfg c;
    c.H = a.L;
    if (band_mode(a.L, BY)) {
        c = band_cfg(a.L, kHalMdm);
        blk = dim3((unsigned)(((c.nthr + 63) / 64) * 64));
    }
    set_band(a, c);
    const size_t lds = lds_bytes<R>(a.L, 2, c.H);
    const dim3 grid((a.Lt + a.T - 1) / a.T, nrhs, c.nband);
    if (blk.x <= 512) {
        auto k = force_kernel<R, BY, 512>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    } else if constexpr (BY == 2) {
        auto k = force_kernel<R, BY, 1024>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    }
    QF_LAUNCHED();
    return QF_OK;
}

template <typename R, int MODE, bool PRE>
int launch_cgu_p(const CguArgs<R>& a0, const Tiling& t, int nrhs, hipStream_t s) {
    const dim3 grid(t.nsx, (a0.Lt + t.TB - 1) / t.TB, nrhs), blk(512);
    CguArgs<R> a = a0;
    a.xcd = ((long long)a.Vs * (long long)sizeof(cplx<R>)) % 128 != 0;   // slices that end inside a line
    switch (t.TB) {
        case 16: hipLaunchKernelGGL((cg_update_kernel<R, MODE, PRE, 32, 16, 1>), grid, blk, 0, s, a); break;
        case 32: hipLaunchKernelGGL((cg_update_kernel<R, MODE, PRE, 16, 32, 1>), grid, blk, 0, s, a); break;
        case 64: hipLaunchKernelGGL((cg_update_kernel<R, MODE, PRE, 16, 64, 2>), grid, blk, 0, s, a); break;
        case 128: hipLaunchKernelGGL((cg_update_kernel<R, MODE, PRE, 16, 128, 4>), grid, blk, 0, s, a); break;
        default: return fail(QF_ERR_UNSUPPORTED, "update tile of %d slices", t.TB);
    }
    QF_LAUNCHED();
    return QF_OK;
}

template <typename R, int MODE>
int launch_cgu(const CguArgs<R>& a, const Tiling& t, int nrhs, hipStream_t s) {
    return a.ptab ? launch_cgu_p<R, MODE, true>(a, t, nrhs, s) : launch_cgu_p<R, MODE, false>(a, t, nrhs, s);
}

// ------------------------------------------------------------ ragged batch
// Every lattice of the batch runs with 2 x 2 patches (valid for every even L)
// in one work-group size (the largest lattice's); smaller lattices leave the
// surplus lanes as padding.
template <typename R, int MODE, bool HOPT>
int launch_mdm_ragged(const MdmArgs<R>& a, const RaggedHead& h, hipStream_t s) {
    const dim3 grid((unsigned)h.nwg_op), blk((unsigned)h.nthreads);
    const size_t lds = lds_bytes_c<R>(2, h.nthreads);
    if (h.nthreads <= 512) {
        auto k = mdagm_kernel<R, 2, MODE, 512, HOPT, mdm_pipe<R, 2>()>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    } else {
        auto k = mdagm_kernel<R, 2, MODE, 1024, HOPT, mdm_pipe<R, 2>()>;
        int rc = set_lds(k, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k, grid, blk, lds, s, a);
    }
    QF_LAUNCHED();
    return QF_OK;
}

template <typename R, int MODE>
int launch_cgu_ragged(const CguArgs<R>& a, const RaggedHead& h, hipStream_t s) {
    const dim3 grid((unsigned)h.nwg_upd), blk(512);
    if (a.ptab)
        hipLaunchKernelGGL((cg_update_kernel<R, MODE, true, 16, 64, 2>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((cg_update_kernel<R, MODE, false, 16, 64, 2>), grid, blk, 0, s, a);
    QF_LAUNCHED();
    return QF_OK;
}

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename R>
int launch_hop_table(const R* phi, long long phi_cs, cplx<R>* hop, int nchain, int L, int Lt, double dtau, double t, hipStream_t s) {
    const long long n = (long long)nchain * L * L * Lt;
    hipLaunchKernelGGL((hop_table_kernel<R>), dim3(std::min<long long>((n + 255) / 256, 16384)), dim3(256), 0, s, phi, phi_cs, hop, nchain,
                       L, Lt, std::tanh(t * dtau), std::tanh(0.5 * t * dtau));
    QF_LAUNCHED();
    return QF_OK;
}

// Workspace carve-up of the PCG solvers (identical in size query and solve).
struct CgWork {
    void* R[2];
    void* P[2];
    void* Z;
    void* OP;
    void* XLO;   // mixed only (complex64)
    void* BLO;   // mixed only (complex64)
    void* HOP;   // hop table of the solve's field(s), [chain][tau][y][x][mu] complex
    double* rzp;
    double* rrp;
    double* pap;
    CgState* st;
    double* normb;
    double* xal;    // alpha of a held iteration, [nrhs]
    double* rp64;   // mixed only
    double* ru_rr;  // mixed only: [2][nrhs]
    double* ru_it;  // mixed only: [2][nrhs]
    int* ru_flag;   // mixed only: [2][nrhs]
    int* brk;       // [nrhs] breakdown flag of the last update step
    size_t bytes;
};

CgWork carve(void* base, int prec, int mixed, int nrhs, int L, int Lt, const Tiling& t, const Tiling& t64) {
    CgWork w;
    const size_t N = (size_t)L * L * Lt;
    const size_t vb = al256((size_t)nrhs * N * (prec == QF_PREC_F64 ? 16 : 8));
    char* p = static_cast<char*>(base);
    size_t off = 0;
    auto take = [&](size_t n) { void* r = p ? p + off : nullptr; off += al256(n); return r; };
    w.R[0] = take(vb); w.R[1] = take(vb); w.P[0] = take(vb); w.P[1] = take(vb); w.Z = take(vb); w.OP = take(vb);
    w.XLO = mixed ? take(vb) : nullptr;
    w.BLO = mixed ? take(vb) : nullptr;
    w.HOP = take(2 * vb);
    w.rzp = (double*)take(sizeof(double) * 2 * nrhs * t.nB);
    w.rrp = (double*)take(sizeof(double) * 2 * nrhs * t.nB);
    w.pap = (double*)take(sizeof(double) * nrhs * t.nA);
    w.st = (CgState*)take(sizeof(CgState) * 2 * nrhs);
    w.normb = (double*)take(sizeof(double) * nrhs);
    w.xal = (double*)take(sizeof(double) * nrhs);
    w.rp64 = mixed ? (double*)take(sizeof(double) * nrhs * t64.nA) : nullptr;
    w.ru_rr = mixed ? (double*)take(sizeof(double) * 2 * nrhs) : nullptr;
    w.ru_it = mixed ? (double*)take(sizeof(double) * 2 * nrhs) : nullptr;
    w.ru_flag = mixed ? (int*)take(sizeof(int) * 2 * nrhs) : nullptr;
    w.brk = (int*)take(sizeof(int) * nrhs);
    w.bytes = off;
    return w;
}

template <typename R>
int pcg_iterate(const R* phi, long long phi_cs, const cplx<R>* b_in, cplx<R>* x, int nrhs, int L, int Lt, const Hop<R>& hp,
                const R* ptab, double rtol, int max_iter, const CgWork& w, const Tiling& t, hipStream_t s,
                // mixed-precision reliable updates (R = float only)
                int ru_period, const double* phi64, long long phi64_cs, const double2* b_hi, double2* x_hi, const Hop<double>* hp64,
                const Tiling* t64, double* err_out, double* iters_out, int* status_out) {
    const int Vs = L * L;
    const long long N = (long long)Vs * Lt;
    CguArgs<R> u{};
    u.vs = N; u.Vs = Vs; u.Lt = Lt; u.TB = t.TB; u.ptab = ptab; u.nA = t.nA;
    MdmArgs<R> m{};
    m.phi = phi; m.phi_cs = phi_cs; m.vs = N; m.L = L; m.Lt = Lt; m.T = t.T; m.hp = hp;
    m.nB = t.nB; m.normb = w.normb; m.rtol = rtol; m.dot_part = w.pap; m.z = (const cplx<R>*)w.Z; m.y = (cplx<R>*)w.OP;
    const long long ps = (long long)nrhs * t.nB;   // partial slot stride
    // complex128 (and complex64 with 2 x 6 patches): hop table of the field(s),
    // read by the iterations instead of recomputing sincos; other complex64
    // iterations take phi directly
    if (cg_uses_hop<R>(t.by)) {
        const int nchain = phi_cs ? nrhs : 1;
        int rc0 = launch_hop_table<R>(phi, phi_cs, (cplx<R>*)w.HOP, nchain, L, Lt, hp.dtau_, hp.t_, s);
        if (rc0) return rc0;
        m.hop = (const cplx<R>*)w.HOP;
        m.hop_cs = phi_cs ? 2 * N : 0;
    }

    // init: x = 0, r = b, z = P b, partials into slot 1, state slot 1
    u.rold = b_in; u.x = x; u.rnew = (cplx<R>*)w.R[1]; u.z = (cplx<R>*)w.Z;
    u.st_init = w.st + nrhs; u.rz_part = w.rzp + ps; u.rr_part = w.rrp + ps; u.brk = w.brk;
    m.brk = w.brk;
    int rc = launch_cgu<R, CGU_INIT>(u, t, nrhs, s);
    if (rc) return rc;

    // Mixed precision: reliable-update checks every kRuCheck iterations (see
    // ru_decide_flush_kernel); the check after iteration i uses parity c & 1.
    constexpr int kRuCheck = 10;
    const bool mixed = ru_period > 0;
    const long long n_el = (long long)nrhs * N;
    const int ru_blocks = (int)std::max(1LL, std::min((N + 1023) / 1024, 512LL));
    int check = 0;
    auto reliable_update = [&](int iters_done, int slot, int force, int cur) -> int {
        RuArgs a{};
        const int c = check & 1, pc = c ^ 1;
        a.xhi = x_hi; a.xlo = (float2*)x; a.N = N;
        a.rr_part = w.rrp + slot * ps; a.nB = t.nB; a.rp64 = w.rp64; a.nA64 = t64->nA;
        a.rr_prev = w.ru_rr + pc * nrhs; a.rr_out = w.ru_rr + c * nrhs;
        a.it_prev = w.ru_it + pc * nrhs; a.it_out = w.ru_it + c * nrhs;
        a.flag_prev = w.ru_flag + pc * nrhs; a.flag_out = w.ru_flag + c * nrhs;
        a.st = w.st + cur * nrhs; a.iter = iters_done; a.ru_max = ru_period; a.delta2 = 1e-4; a.force = force;
        a.normb = (const double*)w.normb; a.rtol2 = rtol * rtol;
        hipLaunchKernelGGL(ru_decide_flush_kernel, dim3(ru_blocks, nrhs), dim3(256), 0, s, a);
        QF_LAUNCHED();
        MdmArgs<double> r{};
        r.phi = phi64; r.phi_cs = phi64_cs; r.x = x_hi; r.vs = N; r.L = L; r.Lt = Lt; r.T = t64->T; r.hp = *hp64;
        r.b = b_hi; r.r_lo = (float2*)w.R[cur]; r.dot_part = w.rp64; r.ru_flag = w.ru_flag + c * nrhs;
        int rc2 = launch_mdm<double, MDM_RESID>(r, t64->by, nrhs, s);
        if (rc2) return rc2;
        ++check;
        return QF_OK;
    };
    (void)n_el;

    // x updates go in pairs (xmode of cg_update_kernel): an iteration holds its
    // step unless x is read after it (a reliable-update check, the end), and the
    // iteration after a held one applies both steps.
    const bool paired = g_paired_x.load() != 0;
    bool held = false;
    for (int i = 0; i < max_iter; ++i) {
        const int cur = i & 1, prv = cur ^ 1;
        const bool ru_due = mixed && (i + 1) % kRuCheck == 0 && i + 1 < max_iter;
        const int xmode = !paired ? 0 : held ? 2 : (ru_due || i + 1 == max_iter ? 0 : 1);
        held = xmode == 1;
        m.iter = i;
        // mixed precision: a chain may only stop on a true (FP64) residual,
        // i.e. right after a reliable update (stop_flag of the last check);
        // the FP32 iterated residual can drift below the true one.
        m.may_stop = 1;
        m.stop_flag = nullptr;
        if (mixed && i > 0) {
            if (i % kRuCheck == 0) m.stop_flag = w.ru_flag + ((check - 1) & 1) * nrhs;
            else m.may_stop = 0;
        }
        m.rz_part = w.rzp + prv * ps; m.rr_part = w.rrp + prv * ps;
        m.st_prev = w.st + prv * nrhs; m.st_out = w.st + cur * nrhs;
        m.pold = (const cplx<R>*)w.P[prv]; m.pnew = (cplx<R>*)w.P[cur];
        rc = launch_mdm<R, MDM_CG>(m, t.by, nrhs, s);
        if (rc) return rc;
        if (mixed && i == 0) {
            hipLaunchKernelGGL(ru_init_kernel, dim3((nrhs + 63) / 64), dim3(64), 0, s, (const double*)w.normb, w.ru_rr + nrhs,
                               w.ru_it + nrhs, w.ru_flag + nrhs, nrhs);
            QF_LAUNCHED();
        }
        u.rold = (const cplx<R>*)w.R[prv]; u.op = (const cplx<R>*)w.OP; u.p = (const cplx<R>*)w.P[cur];
        u.x = x; u.rnew = (cplx<R>*)w.R[cur]; u.z = (cplx<R>*)w.Z;
        u.pap = w.pap; u.st = w.st + cur * nrhs;
        u.rz_part = w.rzp + cur * ps; u.rr_part = w.rrp + cur * ps;
        u.ru_flag = nullptr;
        u.xmode = xmode; u.pprev = (const cplx<R>*)w.P[prv]; u.xalpha = w.xal; u.st_held = w.st + prv * nrhs;
        rc = launch_cgu<R, CGU_ITER>(u, t, nrhs, s);
        if (rc) return rc;
        if (ru_due) {
            // reliable update where needed: x_hi += x_lo; r = b - A64 x_hi (FP64); z = P r
            rc = reliable_update(i + 1, cur, 0, cur);
            if (rc) return rc;
            CguArgs<R> f = u;
            f.rold = (const cplx<R>*)w.R[cur];
            f.ru_flag = w.ru_flag + ((check - 1) & 1) * nrhs;
            rc = launch_cgu<R, CGU_REFRESH>(f, t, nrhs, s);
            if (rc) return rc;
        }
    }
    const int last = (max_iter - 1) & 1;
    if (mixed) {
        rc = reliable_update(max_iter, last, 1, last);
        if (rc) return rc;
        hipLaunchKernelGGL(cg_finalize_kernel, dim3(nrhs), dim3(64), 0, s, (const double*)w.rp64, t64->nA,
                           (const CgState*)(w.st + last * nrhs), (const double*)w.normb, rtol, 1, err_out, iters_out,
                           status_out, (const Prob*)nullptr);
    } else {
        hipLaunchKernelGGL(cg_finalize_kernel, dim3(nrhs), dim3(64), 0, s, (const double*)(w.rrp + last * ps), t.nB,
                           (const CgState*)(w.st + last * nrhs), (const double*)w.normb, rtol, 0, err_out, iters_out,
                           status_out, (const Prob*)nullptr);
    }
    QF_LAUNCHED();
    return QF_OK;
}

// ------------------------------------------------------------ persistent solve
// Tiling of the persistent launch: the operator phase sets the work-group
// (blk_of for the patch height of the solve's tiling, at most 512 threads);
// the update phase uses kPersTX x (blk * kPersSPT / kPersTX) tiles.
struct PersPlan {
    Tiling t;
    int blk, ntile;
};

PersPlan pers_plan(int prec, int nrhs, int L, int Lt) {
    PersPlan p;
    p.t = tiling(prec, nrhs, L, Lt);
    p.blk = (int)blk_of(L, p.t.by).x;
    p.t.TB = p.blk * kPersSPT / kPersTX;
    p.t.nsx = (L * L + kPersTX - 1) / kPersTX;
    p.ntile = (Lt + p.t.TB - 1) / p.t.TB;
    p.t.nB = p.t.nsx * p.ntile;
    return p;
}

// Barrier counters after the solver workspace (zero before the first use; every solve
// leaves them zero).
constexpr size_t kPersBarBytes = 256;

template <typename R, int BY>
int launch_pers(PcgPersArgs<R>& a, int blk, hipStream_t s) {
    constexpr bool PIPE = mdm_pipe<R, BY>();
    constexpr bool HOPT = cg_uses_hop<R>(BY);
    auto k = pcg_persistent_kernel<R, BY, HOPT, PIPE>;
    if (kPersTX * a.u.TB / kPersSPT > blk) return fail(QF_ERR_INVALID, "persistent solve: update tile larger than the work-group");
    const size_t lds = std::max<size_t>(lds_bytes_c<R>(BY, blk), cgu_stage_bytes<R>(a.u.TB, kPersTX));
    int rc = set_lds(k, lds);
    if (rc) return rc;
    int dev = 0, ncu = 0, per = 0, khz = 0;
    QF_HIP(hipGetDevice(&dev));
    QF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    QF_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    QF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, blk, lds));
    if (per < 1) return fail(QF_ERR_UNSUPPORTED, "persistent solve: no work-group of %d threads fits a CU", blk);
    const long long items = a.prob ? std::max(a.nAi, a.nBi) : std::max((long long)a.nrhs * a.nA, (long long)a.nrhs * a.nsx * a.ntile);
    const int G = (int)std::max(1LL, std::min<long long>((long long)per * ncu, items));
    a.timeout_ticks = g_pers_timeout_us.load() * (long long)khz / 1000LL;   // wall-clock ticks per wait
    void* args[] = {&a};
    QF_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k), dim3(G), dim3(blk), args, (unsigned)lds, s));
    return QF_OK;
}

template <typename R>
int pcg_persistent(const R* phi, long long phi_cs, const cplx<R>* b, cplx<R>* x, int nrhs, int L, int Lt, const Hop<R>& hp, const R* ptab,
                   double rtol, int max_iter, const CgWork& w, void* bar, const PersPlan& pl, hipStream_t s, double* err_out,
                   double* iters_out, int* status_out) {
    const Tiling& t = pl.t;
    const long long N = (long long)L * L * Lt;
    PcgPersArgs<R> a{};
    a.u.vs = N; a.u.Vs = L * L; a.u.Lt = Lt; a.u.TB = t.TB; a.u.ptab = ptab; a.u.nA = t.nA;
    a.m.phi = phi; a.m.phi_cs = phi_cs; a.m.vs = N; a.m.L = L; a.m.Lt = Lt; a.m.T = t.T; a.m.hp = hp; a.m.nrhs = nrhs;
    a.m.nB = t.nB; a.m.normb = w.normb; a.m.rtol = rtol; a.m.dot_part = w.pap; a.m.z = (const cplx<R>*)w.Z; a.m.y = (cplx<R>*)w.OP;
    a.nrhs = nrhs; a.nA = t.nA; a.nsx = t.nsx; a.ntile = pl.ntile; a.max_iter = max_iter; a.paired = g_paired_x.load() != 0;
    a.b = b; a.x = x;
    a.R2[0] = (cplx<R>*)w.R[0]; a.R2[1] = (cplx<R>*)w.R[1]; a.P2[0] = (cplx<R>*)w.P[0]; a.P2[1] = (cplx<R>*)w.P[1];
    a.Z = (cplx<R>*)w.Z; a.OP = (cplx<R>*)w.OP;
    a.rzp = w.rzp; a.rrp = w.rrp; a.pap = w.pap; a.ps = (long long)nrhs * t.nB;
    a.st = w.st; a.normb = w.normb; a.xal = w.xal; a.brk = w.brk;
    if (cg_uses_hop<R>(t.by)) {
        a.phi = phi; a.phi_cs = phi_cs; a.hop = (cplx<R>*)w.HOP; a.nchain = phi_cs ? nrhs : 1;
        a.th1 = std::tanh(hp.t_ * hp.dtau_); a.th2 = std::tanh(0.5 * hp.t_ * hp.dtau_);
        a.m.hop = (const cplx<R>*)w.HOP;
        a.m.hop_cs = phi_cs ? 2 * N : 0;
    }
    a.rtol = rtol; a.err_out = err_out; a.iters_out = iters_out; a.status_out = status_out;
    unsigned* c = static_cast<unsigned*>(bar);
    a.gb.count = c; a.gb.done = c + 16; a.gb.abort = c + 32;
    switch (t.by) {
        case 2: return launch_pers<R, 2>(a, pl.blk, s);
        case 4: return launch_pers<R, 4>(a, pl.blk, s);
        case 6: return launch_pers<R, 6>(a, pl.blk, s);
    }
    return fail(QF_ERR_UNSUPPORTED, "patch height %d", t.by);
}

// ------------------------------------------------------------ preconditioner
// Float32 roundings of the manual band constants of
// preconditioners_orig/precon_manual.py:37-167 (the reference stores them as
// complex64).  edge/tail/middle per rank 0..6, corners per rank 1..5.
const float kEdge0[] = {0.91698223f, 1.02107012f, 1.1740706f, 1.24646831f, 1.28092396f, 1.29741096f, 1.30534041f, 1.30917168f, 1.3110323f, 1.31193936f};
const float kTail0[] = {1.31167626f, 1.31074941f, 1.30885756f, 1.30498219f, 1.29699361f, 1.28042281f, 1.24588311f, 1.17343438f, 1.02052784f};
const float kEdge1[] = {0.475227f, 0.69804072f, 0.80315554f, 0.85303694f, 0.87684089f, 0.88826168f, 0.89376724f, 0.89643389f, 0.89773124f, 0.89836514f, 0.89867532f, 0.89882821f};
const float kTail1[] = {0.89785033f, 0.89652652f, 0.89382064f, 0.88826007f, 0.87676316f, 0.85285789f, 0.80287397f, 0.69771057f};
const float kEdge2[] = {0.32525402f, 0.47824016f, 0.550632f, 0.58508581f, 0.60157132f, 0.60949975f, 0.61333305f, 0.61519265f};
const float kTail2[] = {0.61507541f, 0.61318511f, 0.60930747f, 0.60131705f, 0.58475631f, 0.5502165f, 0.47778079f};
const float kEdge3[] = {0.22288805f, 0.32804558f, 0.3779504f, 0.40176818f, 0.41319385f, 0.4187026f, 0.42137364f, 0.42267111f};
const float kTail3[] = {0.42132217f, 0.41861743f, 0.41305307f, 0.40155935f, 0.37766892f, 0.3276957f};
const float kEdge4[] = {0.15300766f, 0.22541766f, 0.25988016f, 0.27637246f, 0.28430402f, 0.28813654f, 0.29000109f, 0.29090756f, 0.291352f};
const float kTail4[] = {0.29132634f, 0.29086968f, 0.28994057f, 0.28805143f, 0.28417516f, 0.27618375f, 0.25962788f, 0.22511749f};
const float kEdge5[] = {0.10518534f, 0.15511395f, 0.17894313f, 0.190377f, 0.19588688f, 0.1985613f, 0.19986567f, 0.20050026f, 0.20081173f, 0.20096488f, 0.20104057f};
const float kTail5[] = {0.20103745f, 0.20095888f, 0.20079921f, 0.20047785f, 0.19982809f, 0.19850214f, 0.19579838f, 0.19023941f, 0.17874525f, 0.15486649f};
const float kEdge6[] = {0.10691644f, 0.12341493f, 0.13135427f, 0.13519259f, 0.13706173f, 0.13796897f, 0.13841458f};
const float kTail6[] = {0.13839549f, 0.13793865f, 0.13701266f, 0.13511762f, 0.13124515f, 0.12326149f, 0.10670882f};
const float kMid[] = {1.312381281f, 0.8989771f, 0.616974f, 0.42391914f, 0.29178312f, 0.20111492f, 0.13884714f};
const float kCorner[] = {-0.47505581f, -0.32495788f, -0.22264314f, -0.15278356f, -0.10501137f};

struct Band {
    const float* e;
    int ne;
    const float* t;
    int nt;
};
const Band kBand[7] = {{kEdge0, 10, kTail0, 9},  {kEdge1, 12, kTail1, 8}, {kEdge2, 8, kTail2, 7}, {kEdge3, 8, kTail3, 6},
                       {kEdge4, 9, kTail4, 8},   {kEdge5, 11, kTail5, 10}, {kEdge6, 7, kTail6, 7}};

}  // namespace

// =================================================================== C ABI
namespace {

// Radices of a length-L transform for the fused estimator kernel: 4s, then 2,
// 3, 5, 7 and any larger prime factors; false if more than kObsMaxFac.
bool obs_factor(int L, unsigned long long& facs, int& nfac) {
    facs = 0;
    nfac = 0;
    int n = L;
    auto push = [&](int p) {
        if (nfac >= kObsMaxFac) return false;
        facs |= (unsigned long long)p << (8 * nfac++);
        return true;
    };
    while (n % 4 == 0) { if (!push(4)) return false; n /= 4; }
    for (int p = 2; p <= n; ++p)
        while (n % p == 0) { if (!push(p)) return false; n /= p; }
    return true;
}


// Launch shape of the fused estimator kernel for (R, estimator, L):
// kObsThreads threads, MAXK sites per thread, DUAL when [A | B] and two ping-pong
// arrays fit in LDS.  Calls f(kernel, lds); QF_ERR_UNSUPPORTED when nothing fits.
template <typename R, int O, typename F> int obs_fused_pick(int L, F&& f) {
    const int V = L * L;
    constexpr size_t kLds = 160 * 1024;
    const bool dual = O != OBS_V && obs_fused_lds<R>(L, O, true) <= kLds;
    const size_t lds = obs_fused_lds<R>(L, O, dual);
    if (lds > kLds || V > 6 * 1024) return fail(QF_ERR_UNSUPPORTED, "fused estimator kernel: %zu bytes of LDS at L = %d", lds, L);
    int m = L;                      // only radices 2, 3, 4 (L = 2^a 3^b)?
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    const bool small = m == 1;
    auto by_k = [&](auto dual_tag, auto small_tag) -> int {
        constexpr bool D = decltype(dual_tag)::value, SM = decltype(small_tag)::value;
        if (V <= 1024) return f(obs_fused_kernel<R, O, kObsThreads, 2, D, SM>, lds);
        if (V <= 2048) return f(obs_fused_kernel<R, O, kObsThreads, 4, D, SM>, lds);
        if (V <= 4096) return f(obs_fused_kernel<R, O, kObsThreads, 8, D, SM>, lds);
        return f(obs_fused_kernel<R, O, kObsThreads, 12, D, SM>, lds);
    };
    auto by_small = [&](auto dual_tag) -> int {
        return small ? by_k(dual_tag, std::true_type{}) : by_k(dual_tag, std::false_type{});
    };
    if constexpr (O == OBS_V) return by_small(std::false_type{});
    else return dual ? by_small(std::true_type{}) : by_small(std::false_type{});
}

template <typename R, typename F> int obs_fused_dispatch(int obs, int L, F&& f) {
    switch (obs) {
        case QF_OBS_T1: return obs_fused_pick<R, OBS_T1>(L, f);
        case QF_OBS_T21: return obs_fused_pick<R, OBS_T21>(L, f);
        case QF_OBS_T2: return obs_fused_pick<R, OBS_T2>(L, f);
        case QF_OBS_T4: return obs_fused_pick<R, OBS_T4>(L, f);
        default: return obs_fused_pick<R, OBS_V>(L, f);
    }
}

}  // namespace

// Workspace of qf_pcg_solve_noalloc: a file-scope device array, so that callers which
// cannot allocate (code under graph capture, serving loops, harnesses that forbid
// hipMalloc) still get a solver.  Every GPU the library runs on holds its own copy of
// these kQfScratchBytes until the process ends.  In host code `scratch` only names the
// symbol: its device address, per device, comes from hipGetSymbolAddress and is the only
// thing passed to kernels.
constexpr size_t kQfScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kQfScratchBytes];

namespace {
constexpr int kQfMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kQfMaxDevices] = {};

// Device address of the current device's scratch array (looked up once per device).
int scratch_address(void** out) {
    int dev = -1;
    QF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kQfMaxDevices) return fail(QF_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kQfMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        QF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) return fail(QF_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return QF_OK;
}
}  // namespace

extern "C" {

int qf_version(void) { return QF_ABI_VERSION; }

const char* qf_last_error(void) { return g_err.c_str(); }

int qf_set_tuning(int prec, int by, int t, int tb) {
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    if (by != 0 && (by % 2 || by < 2 || by > 6)) return fail(QF_ERR_INVALID, "patch height %d not in {0,2,4,6}", by);
    if (t < 0 || (tb != 0 && !tb_ok(tb))) return fail(QF_ERR_INVALID, "bad tile size (t=%d, tb=%d not in {0,16,32,64,128})", t, tb);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[prec].by = by;
    g_tune[prec].t = t;
    g_tune[prec].tb = tb;
    return QF_OK;
}

int qf_set_tuning_shape(int prec, int L, int Ltau, int nrhs, int by, int t, int tb) {
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    if (by != 0 && !by_ok(L, by) && !(by == 2 && band_cfg(L, kHalMdm).nthr > 0)) return fail(QF_ERR_INVALID, "patch height %d does not fit L=%d", by, L);
    if (t < 0 || (tb != 0 && !tb_ok(tb)) || nrhs < 0 || L < 2 || Ltau < 1) return fail(QF_ERR_INVALID, "bad shape or tile size");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& e : g_shape_tune)
        if (e.prec == prec && e.L == L && e.Lt == Ltau && e.nrhs == nrhs) {
            e.tu.by = by; e.tu.t = t; e.tu.tb = tb;
            return QF_OK;
        }
    ShapeTuning e;
    e.prec = prec; e.L = L; e.Lt = Ltau; e.nrhs = nrhs;
    e.tu.by = by; e.tu.t = t; e.tu.tb = tb;
    g_shape_tune.push_back(e);
    return QF_OK;
}

int qf_set_tuning_ns(int prec, int L, int Ltau, int nrhs, int ns) {
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    if (ns < 0 || ns > 64 || nrhs < 0 || L < 2 || Ltau < 1) return fail(QF_ERR_INVALID, "bad shape or ns=%d", ns);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& e : g_shape_tune)
        if (e.prec == prec && e.L == L && e.Lt == Ltau && e.nrhs == nrhs) {
            e.tu.ns = ns;
            return QF_OK;
        }
    ShapeTuning e;
    e.prec = prec; e.L = L; e.Lt = Ltau; e.nrhs = nrhs;
    e.tu.ns = ns;
    g_shape_tune.push_back(e);
    return QF_OK;
}

int qf_clear_tuning(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_shape_tune.clear();
    g_tune[0] = Tuning();
    g_tune[1] = Tuning();
    return QF_OK;
}

int qf_set_paired_x_updates(int on) { return g_paired_x.exchange(on ? 1 : 0); }

long long qf_set_persistent_timeout_us(long long us) { return g_pers_timeout_us.exchange(us < 0 ? 0 : us); }

int qf_get_tiling(int prec, int nrhs, int L, int Ltau, int* out4) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    Tiling t = tiling(prec, nrhs, L, Ltau);
    out4[0] = t.by; out4[1] = t.T; out4[2] = t.TB; out4[3] = t.nA;
    return QF_OK;
}

int qf_get_rhs_per_group(int prec, int nrhs, int L, int Ltau, int* ns) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!ns) return fail(QF_ERR_INVALID, "null pointer");
    *ns = tiling(prec, nrhs, L, Ltau).ns;
    return QF_OK;
}

int qf_precon_table(int Ltau, double* table) {
    if (!table) return fail(QF_ERR_INVALID, "null table");
    if (Ltau < 26) return fail(QF_ERR_LTAU, "the manual preconditioner needs Ltau >= 26 (got %d)", Ltau);
    std::memset(table, 0, sizeof(double) * 13 * (size_t)Ltau);
    for (int tau = 0; tau < Ltau; ++tau) {
        const Band& b = kBand[0];
        float v = tau < b.ne ? b.e[tau] : tau >= Ltau - b.nt ? b.t[tau - (Ltau - b.nt)] : kMid[0];
        table[tau * 13 + 6] = v;
    }
    for (int r = 1; r <= 6; ++r) {
        const Band& b = kBand[r];
        const int n = Ltau - r - (r == 6 ? 1 : 0);
        const int row0 = r == 6 ? 1 : 0;
        for (int k = 0; k < n; ++k) {
            const float v = k < b.ne ? b.e[k] : k >= n - b.nt ? b.t[k - (n - b.nt)] : kMid[r];
            table[(row0 + k) * 13 + 6 + r] += v;
            table[(row0 + k + r) * 13 + 6 - r] += v;
        }
    }
    for (int r = 1; r <= 5; ++r) {
        table[0 * 13 + 6 - r] += kCorner[r - 1];
        table[(Ltau - r) * 13 + 6 + r] += kCorner[r - 1];
    }
    return QF_OK;
}

int qf_precon_table_from_csr(const long long* crow, const long long* col, const double* val_re, const double* val_im, long long nrows,
                             int Vs, int Ltau, double* table) {
    if (!crow || !col || !val_re || !table) return fail(QF_ERR_INVALID, "null argument");
    if (Vs < 1 || Ltau < 1 || nrows != (long long)Vs * Ltau) return fail(QF_ERR_INVALID, "CSR has %lld rows, expected Vs*Ltau", nrows);
    std::vector<char> seen((size_t)13 * Ltau, 0);
    std::memset(table, 0, sizeof(double) * 13 * (size_t)Ltau);
    // pass 1: site 0 of every slice defines the tau stencil
    for (int tau = 0; tau < Ltau; ++tau) {
        const long long row = (long long)tau * Vs;
        for (long long k = crow[row]; k < crow[row + 1]; ++k) {
            const long long c = col[k];
            if (c < 0 || c >= nrows) return fail(QF_ERR_PRECON, "column %lld out of range", c);
            if (c % Vs != 0) return fail(QF_ERR_PRECON, "row %lld couples different sites: not P_tau (x) I", row);
            int o = (int)(c / Vs) - tau;
            if (o > Ltau / 2) o -= Ltau;
            if (o < -Ltau / 2) o += Ltau;
            if (o < -6 || o > 6) return fail(QF_ERR_PRECON, "tau offset %d outside the 13-band", o);
            if (val_im && val_im[k] != 0.0) return fail(QF_ERR_PRECON, "imaginary preconditioner entry");
            table[tau * 13 + 6 + o] += val_re[k];
            seen[tau * 13 + 6 + o] = 1;
        }
    }
    // pass 2: every other site repeats the same stencil
    for (long long row = 0; row < nrows; ++row) {
        const int tau = (int)(row / Vs), s = (int)(row % Vs);
        double acc[13] = {0};
        for (long long k = crow[row]; k < crow[row + 1]; ++k) {
            const long long c = col[k];
            if (c < 0 || c >= nrows || c % Vs != s) return fail(QF_ERR_PRECON, "row %lld: entry off the site diagonal", row);
            int o = (int)(c / Vs) - tau;
            if (o > Ltau / 2) o -= Ltau;
            if (o < -Ltau / 2) o += Ltau;
            if (o < -6 || o > 6) return fail(QF_ERR_PRECON, "row %lld: tau offset %d outside the band", row, o);
            acc[o + 6] += val_re[k];
        }
        for (int o = 0; o < 13; ++o)
            if (acc[o] != table[tau * 13 + o]) return fail(QF_ERR_PRECON, "row %lld: stencil differs from site 0 of its slice", row);
    }
    return QF_OK;
}

int qf_apply(int op, int prec, const void* phi, long long phi_cs, const void* x, void* y, int nrhs, int L, int Ltau, double dtau, double t,
             void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!phi || !x || !y) return fail(QF_ERR_INVALID, "null pointer");
    if (x == y) return fail(QF_ERR_INVALID, "x and y must not alias");
    hipStream_t s = (hipStream_t)stream;
    const Tiling tl = tiling(prec, nrhs, L, Ltau);
    const long long N = (long long)L * L * Ltau;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        if (op == QF_OP_MDAGM) {
            MdmArgs<R> a{};
            a.phi = (const R*)phi; a.phi_cs = phi_cs; a.x = (const cplx<R>*)x; a.y = (cplx<R>*)y; a.vs = N;
            a.L = L; a.Lt = Ltau; a.T = tl.T; a.hp = make_hop<R>(dtau, t); a.nsub = tl.ns;
            return launch_mdm<R, MDM_PLAIN>(a, tl.by, nrhs, s);
        }
        OpArgs<R> a{};
        a.phi = (const R*)phi; a.phi_cs = phi_cs; a.x = (const cplx<R>*)x; a.y = (cplx<R>*)y; a.vs = N;
        a.L = L; a.Lt = Ltau; a.T = tl.T; a.hp = make_hop<R>(dtau, t); a.nfac = 7; a.cpart = a.hp.cn;
        if (op >= QF_OP_BPART0 && op <= QF_OP_BPART0 + 7) {
            a.nfac = op - QF_OP_BPART0;
            a.cpart = (R)cosh_prefix(a.nfac, dtau, t);
            return launch_op<R, OP_B>(a, tl.by, nrhs, s);
        }
        if (op == QF_OP_M) return launch_op<R, OP_M>(a, tl.by, nrhs, s);
        if (op == QF_OP_MDAG) return launch_op<R, OP_MDAG>(a, tl.by, nrhs, s);
        if (op == QF_OP_B) return launch_op<R, OP_B>(a, tl.by, nrhs, s);
        return fail(QF_ERR_INVALID, "unknown operator %d", op);
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

int qf_hop_table(int prec, const void* phi, long long phi_cs, int nchain, int L, int Ltau, double dtau, double t, void* hop, void* stream) {
    int rc = check_lattice(L, Ltau, nchain);
    if (rc) return rc;
    if (!phi || !hop) return fail(QF_ERR_INVALID, "null pointer");
    if (prec == QF_PREC_F32) return launch_hop_table<float>((const float*)phi, phi_cs, (float2*)hop, nchain, L, Ltau, dtau, t, (hipStream_t)stream);
    if (prec == QF_PREC_F64)
        return launch_hop_table<double>((const double*)phi, phi_cs, (double2*)hop, nchain, L, Ltau, dtau, t, (hipStream_t)stream);
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

int qf_apply_hop(int op, int prec, const void* hop, long long hop_cs, const void* x, void* y, int nrhs, int L, int Ltau, double dtau, double t,
                 void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!hop || !x || !y) return fail(QF_ERR_INVALID, "null pointer");
    if (x == y) return fail(QF_ERR_INVALID, "x and y must not alias");
    hipStream_t s = (hipStream_t)stream;
    const Tiling tl = tiling(prec, nrhs, L, Ltau);
    const long long N = (long long)L * L * Ltau;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        if (op == QF_OP_MDAGM) {
            MdmArgs<R> a{};
            a.hop = (const cplx<R>*)hop; a.hop_cs = hop_cs; a.x = (const cplx<R>*)x; a.y = (cplx<R>*)y; a.vs = N;
            a.L = L; a.Lt = Ltau; a.T = tl.T; a.hp = make_hop<R>(dtau, t);
            return launch_mdm<R, MDM_PLAIN>(a, tl.by, nrhs, s);
        }
        OpArgs<R> a{};
        a.hop = (const cplx<R>*)hop; a.hop_cs = hop_cs; a.x = (const cplx<R>*)x; a.y = (cplx<R>*)y; a.vs = N;
        a.L = L; a.Lt = Ltau; a.T = tl.T; a.hp = make_hop<R>(dtau, t); a.nfac = 7; a.cpart = a.hp.cn;
        if (op == QF_OP_M) return launch_op<R, OP_M>(a, tl.by, nrhs, s);
        if (op == QF_OP_MDAG) return launch_op<R, OP_MDAG>(a, tl.by, nrhs, s);
        if (op == QF_OP_B) return launch_op<R, OP_B>(a, tl.by, nrhs, s);
        return fail(QF_ERR_INVALID, "unknown operator %d", op);
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

int qf_mhm_vec(int prec, const void* phi, const void* x, void* y, int bs, int L, int Ltau, double dtau, void* stream) {
    return qf_apply(QF_OP_MDAGM, prec, phi, (long long)2 * L * L * Ltau, x, y, bs, L, Ltau, dtau, 1.0, stream);
}

int qf_mh_vec(int prec, const void* phi, const void* x, void* y, int bs, int L, int Ltau, double dtau, void* stream) {
    return qf_apply(QF_OP_MDAG, prec, phi, (long long)2 * L * L * Ltau, x, y, bs, L, Ltau, dtau, 1.0, stream);
}

int qf_precon_vec(int prec, const void* ptab, const void* r, void* z, int nrhs, int Vs, int Ltau, void* stream) {
    if (!ptab || !r || !z) return fail(QF_ERR_INVALID, "null pointer");
    if (r == z) return fail(QF_ERR_INVALID, "r and z must not alias");
    if (Vs < 1 || Ltau < 1 || nrhs < 1 || nrhs > 65535) return fail(QF_ERR_INVALID, "bad sizes");
    hipStream_t s = (hipStream_t)stream;
    Tiling t{};
    t.TB = Ltau >= 128 ? 128 : Ltau >= 64 ? 64 : Ltau > 16 ? 32 : 16;
    t.nsx = (Vs + cgu_tx(t.TB) - 1) / cgu_tx(t.TB);
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        CguArgs<R> a{};
        a.rold = (const cplx<R>*)r; a.z = (cplx<R>*)z; a.vs = (long long)Vs * Ltau; a.Vs = Vs; a.Lt = Ltau; a.TB = t.TB;
        a.ptab = (const R*)ptab;
        return launch_cgu<R, CGU_PRECON>(a, t, nrhs, s);
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

size_t qf_pcg_workspace_bytes(int prec, int mixed, int nrhs, int L, int Ltau) {
    if (check_lattice(L, Ltau, nrhs)) return 0;
    const Tiling t = tiling(mixed ? QF_PREC_F32 : prec, nrhs, L, Ltau);
    const Tiling t64 = tiling(QF_PREC_F64, nrhs, L, Ltau);
    return carve(nullptr, mixed ? QF_PREC_F32 : prec, mixed, nrhs, L, Ltau, t, t64).bytes;
}

int qf_pcg_solve(int prec, const void* phi, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau, double dtau, double t,
                 const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes, double* err_out, double* iters_out,
                 int* status_out, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!phi || !b || !x || !work || !err_out || !iters_out) return fail(QF_ERR_INVALID, "null pointer");
    if (max_iter < 1) return fail(QF_ERR_INVALID, "max_iter=%d must be >= 1", max_iter);
    const Tiling tl = tiling(prec, nrhs, L, Ltau);
    const Tiling t64 = tiling(QF_PREC_F64, nrhs, L, Ltau);
    const CgWork w = carve(work, prec, 0, nrhs, L, Ltau, tl, t64);
    if (work_bytes < w.bytes) return fail(QF_ERR_INVALID, "workspace of %zu bytes, need %zu", work_bytes, w.bytes);
    hipStream_t s = (hipStream_t)stream;
    if (prec == QF_PREC_F32)
        return pcg_iterate<float>((const float*)phi, phi_cs, (const float2*)b, (float2*)x, nrhs, L, Ltau, make_hop<float>(dtau, t),
                                  (const float*)ptab, rtol, max_iter, w, tl, s, 0, nullptr, 0, nullptr, nullptr, nullptr, nullptr, err_out,
                                  iters_out, status_out);
    if (prec == QF_PREC_F64)
        return pcg_iterate<double>((const double*)phi, phi_cs, (const double2*)b, (double2*)x, nrhs, L, Ltau, make_hop<double>(dtau, t),
                                   (const double*)ptab, rtol, max_iter, w, tl, s, 0, nullptr, 0, nullptr, nullptr, nullptr, nullptr,
                                   err_out, iters_out, status_out);
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

size_t qf_pcg_noalloc_max_bytes(void) { return kQfScratchBytes; }

int qf_pcg_solve_noalloc(int prec, const void* phi, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau, double dtau,
                         double t, const void* ptab, double rtol, int max_iter, double* err_out, double* iters_out, int* status_out,
                         void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    const size_t need = qf_pcg_workspace_bytes(prec, 0, nrhs, L, Ltau);
    if (need == 0 || need > kQfScratchBytes)
        return fail(QF_ERR_INVALID, "qf_pcg_solve_noalloc: this solve needs %zu bytes of workspace, more than the %zu bytes of the "
                                    "library's static scratch (use qf_pcg_solve with a caller workspace)", need, kQfScratchBytes);
    void* work = nullptr;
    rc = scratch_address(&work);
    if (rc) return rc;
    // the solve writes every part of the workspace before reading it, so the previous call's
    // contents need no clearing
    return qf_pcg_solve(prec, phi, phi_cs, b, x, nrhs, L, Ltau, dtau, t, ptab, rtol, max_iter, work, kQfScratchBytes, err_out,
                        iters_out, status_out, stream);
}

size_t qf_pcg_persistent_workspace_bytes(int prec, int nrhs, int L, int Ltau) {
    if (check_lattice(L, Ltau, nrhs)) return 0;
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) {
        fail(QF_ERR_INVALID, "prec %d", prec);
        return 0;
    }
    const PersPlan pl = pers_plan(prec, nrhs, L, Ltau);
    if (pl.blk > 512 || !square_ok(L)) {
        fail(QF_ERR_UNSUPPORTED, "L=%d: the persistent solve takes work-groups of at most 512 threads", L);
        return 0;
    }
    const Tiling t64 = tiling(QF_PREC_F64, nrhs, L, Ltau);
    return carve(nullptr, prec, 0, nrhs, L, Ltau, pl.t, t64).bytes + kPersBarBytes;
}

int qf_pcg_solve_persistent(int prec, const void* phi, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau, double dtau,
                            double t, const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes, double* err_out,
                            double* iters_out, int* status_out, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!phi || !b || !x || !work || !err_out || !iters_out) return fail(QF_ERR_INVALID, "null pointer");
    if (!ptab) return fail(QF_ERR_INVALID, "the persistent solve needs the preconditioner table");
    if (max_iter < 1 || max_iter > 1000000) return fail(QF_ERR_INVALID, "max_iter=%d out of range [1, 1000000]", max_iter);
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    const PersPlan pl = pers_plan(prec, nrhs, L, Ltau);
    if (pl.blk > 512 || !square_ok(L)) return fail(QF_ERR_UNSUPPORTED, "L=%d: the persistent solve takes work-groups of at most 512 threads", L);
    const Tiling t64 = tiling(QF_PREC_F64, nrhs, L, Ltau);
    const CgWork w = carve(work, prec, 0, nrhs, L, Ltau, pl.t, t64);
    if (work_bytes < w.bytes + kPersBarBytes)
        return fail(QF_ERR_INVALID, "workspace of %zu bytes, need %zu", work_bytes, w.bytes + kPersBarBytes);
    void* bar = static_cast<char*>(work) + w.bytes;
    hipStream_t s = (hipStream_t)stream;
    if (prec == QF_PREC_F32)
        return pcg_persistent<float>((const float*)phi, phi_cs, (const float2*)b, (float2*)x, nrhs, L, Ltau, make_hop<float>(dtau, t),
                                     (const float*)ptab, rtol, max_iter, w, bar, pl, s, err_out, iters_out, status_out);
    return pcg_persistent<double>((const double*)phi, phi_cs, (const double2*)b, (double2*)x, nrhs, L, Ltau, make_hop<double>(dtau, t),
                                  (const double*)ptab, rtol, max_iter, w, bar, pl, s, err_out, iters_out, status_out);
}

int qf_pcg_mixed_solve(const double* phi64, const float* phi32, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau,
                       double dtau, double t, const float* ptab32, double rtol, int max_iter, int ru_period, void* work,
                       size_t work_bytes, double* err_out, double* iters_out, int* status_out, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!phi64 || !phi32 || !b || !x || !work || !err_out || !iters_out) return fail(QF_ERR_INVALID, "null pointer");
    if (max_iter < 1) return fail(QF_ERR_INVALID, "max_iter=%d must be >= 1", max_iter);
    if (ru_period < 1) return fail(QF_ERR_INVALID, "reliable-update period %d must be >= 1", ru_period);
    const Tiling tl = tiling(QF_PREC_F32, nrhs, L, Ltau);
    const Tiling t64 = tiling(QF_PREC_F64, nrhs, L, Ltau);
    const CgWork w = carve(work, QF_PREC_F32, 1, nrhs, L, Ltau, tl, t64);
    if (work_bytes < w.bytes) return fail(QF_ERR_INVALID, "workspace of %zu bytes, need %zu", work_bytes, w.bytes);
    hipStream_t s = (hipStream_t)stream;
    const long long n = (long long)nrhs * L * L * Ltau;
    const dim3 g1(std::min<long long>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(demote_kernel, g1, dim3(256), 0, s, (const double2*)b, (float2*)w.BLO, n);
    QF_LAUNCHED();
    hipLaunchKernelGGL(zero_kernel, g1, dim3(256), 0, s, (double*)x, 2 * n);
    QF_LAUNCHED();
    const Hop<double> hp64 = make_hop<double>(dtau, t);
    return pcg_iterate<float>(phi32, phi_cs, (const float2*)w.BLO, (float2*)w.XLO, nrhs, L, Ltau, make_hop<float>(dtau, t), ptab32, rtol,
                              max_iter, w, tl, s, ru_period, phi64, phi_cs, (const double2*)b, (double2*)x, &hp64, &t64, err_out,
                              iters_out, status_out);
}

// ---------------------------------------------------------------- ragged batch
size_t qf_ragged_plan_bytes(int nprob) {
    if (nprob < 1) return 0;
    return sizeof(RaggedHead) + (size_t)nprob * sizeof(Prob);
}

int qf_ragged_plan(int nprob, const int* Ls, const int* Lts, void* plan, size_t plan_bytes) {
    if (nprob < 1 || nprob > 4096 || !Ls || !Lts || !plan) return fail(QF_ERR_INVALID, "bad ragged plan arguments");
    if (plan_bytes < qf_ragged_plan_bytes(nprob)) return fail(QF_ERR_INVALID, "plan buffer of %zu bytes, need %zu", plan_bytes,
                                                               qf_ragged_plan_bytes(nprob));
    RaggedHead* h = static_cast<RaggedHead*>(plan);
    Prob* P = reinterpret_cast<Prob*>(h + 1);
    std::memset(plan, 0, qf_ragged_plan_bytes(nprob));
    h->nprob = nprob; h->by = 2; h->tb = 64;
    int maxthr = 0;
    long long voff = 0, toff = 0, wg = 0, bwg = 0, pa = 0, pb = 0;
    for (int k = 0; k < nprob; ++k) {
        const int L = Ls[k], Lt = Lts[k];
        if (L < 2 || (L & 1)) return fail(QF_ERR_ODD_L, "lattice %d: L=%d must be even and >= 2", k, L);
        if (Lt < 1) return fail(QF_ERR_INVALID, "lattice %d: Ltau=%d", k, Lt);
        Prob& q = P[k];
        q.L = L; q.Lt = Lt;
        q.T = std::min(Lt, 4);
        q.nA = (Lt + q.T - 1) / q.T;
        q.nband = 1; q.H = L; q.hal = 0; q.nown = L;
        int thr = nthreads(L, 2);
        if (!by_ok(L, 2)) {   // too wide for one work-group per slice: bands of 2x2 patches
            const BandCfg c = band_cfg(L, kHalMdm);
            if (c.nthr == 0) return fail(QF_ERR_UNSUPPORTED, "lattice %d: L=%d has no band of 2x2 patches that fits (L <= 226)", k, L);
            q.nband = c.nband; q.H = c.H; q.hal = c.hal; q.nown = c.nown;
            q.nA *= c.nband;
            thr = c.nthr;
        }
        q.wg0 = (int)wg;
        q.nsx = (L * L + cgu_tx(h->tb) - 1) / cgu_tx(h->tb);
        q.nB = q.nsx * ((Lt + h->tb - 1) / h->tb);
        q.bwg0 = (int)bwg;
        q.voff = voff; q.poff = 2 * voff; q.hoff = 2 * voff; q.toff = toff;
        q.pa0 = (int)pa; q.pb0 = (int)pb;
        maxthr = std::max(maxthr, thr);
        wg += q.nA; bwg += q.nB; pa += q.nA; pb += q.nB;
        voff += (long long)L * L * Lt; toff += 13LL * Lt;
        if (wg > (1LL << 30) || bwg > (1LL << 30)) return fail(QF_ERR_INVALID, "ragged batch too large");
    }
    h->nthreads = ((maxthr + 63) / 64) * 64;
    h->nwg_op = (int)wg; h->nwg_upd = (int)bwg;
    h->nA_tot = (int)pa; h->nB_tot = (int)pb;
    h->nsites = voff; h->ntab = toff;
    return QF_OK;
}

static int ragged_head(const void* plan_host, const void* plan_dev, const RaggedHead** h, const Prob** pd) {
    if (!plan_host || !plan_dev) return fail(QF_ERR_INVALID, "null plan");
    *h = static_cast<const RaggedHead*>(plan_host);
    if ((*h)->nprob < 1 || (*h)->by != 2 || (*h)->nthreads < 64 || (*h)->nthreads > 1024)
        return fail(QF_ERR_INVALID, "not a ragged plan (build it with qf_ragged_plan)");
    *pd = reinterpret_cast<const Prob*>(static_cast<const RaggedHead*>(plan_dev) + 1);
    return QF_OK;
}

int qf_apply_ragged(int op, int prec, const void* plan_host, const void* plan_dev, const void* phi, const void* x, void* y, double dtau,
                    double t, void* stream) {
    const RaggedHead* h;
    const Prob* pd;
    int rc = ragged_head(plan_host, plan_dev, &h, &pd);
    if (rc) return rc;
    if (op != QF_OP_MDAGM) return fail(QF_ERR_UNSUPPORTED, "ragged batches apply M^dag M only (op %d)", op);
    if (!phi || !x || !y) return fail(QF_ERR_INVALID, "null pointer");
    if (x == y) return fail(QF_ERR_INVALID, "x and y must not alias");
    hipStream_t s = (hipStream_t)stream;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        MdmArgs<R> a{};
        a.phi = (const R*)phi; a.x = (const cplx<R>*)x; a.y = (cplx<R>*)y; a.hp = make_hop<R>(dtau, t);
        a.prob = pd; a.nprob = h->nprob; a.nrhs = 1;
        return launch_mdm_ragged<R, MDM_PLAIN, false>(a, *h, s);
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

}  // extern "C"

namespace {
// Workspace of a ragged PCG solve (mixed: complex64 iterations with FP64 reliable updates).
struct RgWork {
    void* R[2];
    void* P[2];
    void* Z;
    void* OP;
    void* HOP;       // complex128 only
    void* XLO;       // mixed only: the complex64 iterate
    void* BLO;       // mixed only: b in complex64
    double* rzp;
    double* rrp;
    double* pap;
    double* rp64;    // mixed only: FP64 residual partials (operator layout)
    CgState* st;
    double* normb;
    int* brk;
    double* ru_rr;   // mixed only: [2][nprob]
    double* ru_it;
    int* ru_flag;
    size_t bytes;
};

RgWork carve_ragged(void* base, int prec, bool mixed, const RaggedHead* h) {
    RgWork w{};
    const size_t vb = al256((size_t)h->nsites * (prec == QF_PREC_F64 ? 16 : 8));
    const int np = h->nprob;
    char* p = static_cast<char*>(base);
    size_t off = 0;
    auto take = [&](size_t n) { void* r = p ? p + off : nullptr; off += al256(n); return r; };
    w.R[0] = take(vb); w.R[1] = take(vb); w.P[0] = take(vb); w.P[1] = take(vb); w.Z = take(vb); w.OP = take(vb);
    w.HOP = prec == QF_PREC_F64 ? take(2 * vb) : nullptr;
    w.XLO = mixed ? take(vb) : nullptr;
    w.BLO = mixed ? take(vb) : nullptr;
    w.rzp = (double*)take(sizeof(double) * 2 * h->nB_tot);
    w.rrp = (double*)take(sizeof(double) * 2 * h->nB_tot);
    w.pap = (double*)take(sizeof(double) * h->nA_tot);
    w.rp64 = mixed ? (double*)take(sizeof(double) * h->nA_tot) : nullptr;
    w.st = (CgState*)take(sizeof(CgState) * 2 * np);
    w.normb = (double*)take(sizeof(double) * np);
    w.brk = (int*)take(sizeof(int) * np);
    w.ru_rr = mixed ? (double*)take(sizeof(double) * 2 * np) : nullptr;
    w.ru_it = mixed ? (double*)take(sizeof(double) * 2 * np) : nullptr;
    w.ru_flag = mixed ? (int*)take(sizeof(int) * 2 * np) : nullptr;
    w.bytes = off;
    return w;
}

// PCG over a ragged batch: one operator launch and one update launch per iteration for
// every lattice (the recurrence of pcg_iterate, one chain per lattice, x updated every
// iteration).  ru_period > 0 (complex64): the mixed-precision solver of pcg_iterate, with
// x the complex64 iterate, reliable updates checked every 10 iterations into x_hi (FP64
// residual from phi64 and b_hi), and chains stopping only on a true residual.
template <typename R>
int pcg_ragged(const RaggedHead* h, const Prob* pd, const R* phi, const cplx<R>* b, cplx<R>* x, double dtau, double t, const R* ptab,
               double rtol, int max_iter, const RgWork& w, double* err_out, double* iters_out, int* status_out, hipStream_t s,
               int ru_period = 0, const double* phi64 = nullptr, const double2* b_hi = nullptr, double2* x_hi = nullptr) {
    const Hop<R> hp = make_hop<R>(dtau, t);
    const int np = h->nprob;
    const long long ps = h->nB_tot;
    const bool mixed = ru_period > 0;
    if constexpr (sizeof(R) == 8) {
        hipLaunchKernelGGL((hop_table_ragged_kernel<R>), dim3((unsigned)std::min<long long>((h->nsites + 255) / 256, 16384)),
                           dim3(256), 0, s, phi, (cplx<R>*)w.HOP, pd, np, h->nsites, std::tanh(t * dtau), std::tanh(0.5 * t * dtau));
        QF_LAUNCHED();
    }
    CguArgs<R> u{};
    u.ptab = ptab; u.prob = pd; u.nprob = np;
    u.rold = b; u.x = x; u.rnew = (cplx<R>*)w.R[1]; u.z = (cplx<R>*)w.Z;
    u.st_init = w.st + np; u.rz_part = w.rzp + ps; u.rr_part = w.rrp + ps; u.brk = w.brk;
    int rc = launch_cgu_ragged<R, CGU_INIT>(u, *h, s);
    if (rc) return rc;
    MdmArgs<R> m{};
    m.phi = phi; m.hop = (const cplx<R>*)w.HOP; m.hp = hp; m.prob = pd; m.nprob = np; m.nrhs = 1;
    m.normb = w.normb; m.rtol = rtol; m.dot_part = w.pap; m.z = (const cplx<R>*)w.Z; m.y = (cplx<R>*)w.OP; m.may_stop = 1;
    m.brk = w.brk;

    // mixed precision: the reliable-update check after iteration i uses parity check & 1
    constexpr int kRuCheck = 10;
    const Prob* ph = reinterpret_cast<const Prob*>(h + 1);   // host copy of the plan
    long long nmax = 1;
    for (int k = 0; k < np; ++k) nmax = std::max(nmax, (long long)ph[k].L * ph[k].L * ph[k].Lt);
    const int ru_blocks = (int)std::max(1LL, std::min((nmax + 1023) / 1024, 512LL));
    const Hop<double> hp64 = make_hop<double>(dtau, t);
    int check = 0;
    auto reliable_update = [&](int iters_done, int slot, int force, int cur) -> int {
        RuArgs a{};
        const int c = check & 1, pc = c ^ 1;
        a.xhi = x_hi; a.xlo = (float2*)x; a.prob = pd;
        a.rr_part = w.rrp + slot * ps; a.rp64 = w.rp64;
        a.rr_prev = w.ru_rr + pc * np; a.rr_out = w.ru_rr + c * np;
        a.it_prev = w.ru_it + pc * np; a.it_out = w.ru_it + c * np;
        a.flag_prev = w.ru_flag + pc * np; a.flag_out = w.ru_flag + c * np;
        a.st = w.st + cur * np; a.iter = iters_done; a.ru_max = ru_period; a.delta2 = 1e-4; a.force = force;
        a.normb = (const double*)w.normb; a.rtol2 = rtol * rtol;
        hipLaunchKernelGGL(ru_decide_flush_kernel, dim3(ru_blocks, np), dim3(256), 0, s, a);
        QF_LAUNCHED();
        MdmArgs<double> r{};
        r.phi = phi64; r.x = x_hi; r.hp = hp64; r.prob = pd; r.nprob = np; r.nrhs = 1;
        r.b = b_hi; r.r_lo = (float2*)w.R[cur]; r.dot_part = w.rp64; r.ru_flag = w.ru_flag + c * np;
        int rc2 = launch_mdm_ragged<double, MDM_RESID, false>(r, *h, s);
        if (rc2) return rc2;
        ++check;
        return QF_OK;
    };

    for (int i = 0; i < max_iter; ++i) {
        const int cur = i & 1, prv = cur ^ 1;
        m.iter = i;
        m.may_stop = 1;
        m.stop_flag = nullptr;
        if (mixed && i > 0) {
            if (i % kRuCheck == 0) m.stop_flag = w.ru_flag + ((check - 1) & 1) * np;
            else m.may_stop = 0;
        }
        m.rz_part = w.rzp + prv * ps; m.rr_part = w.rrp + prv * ps;
        m.st_prev = w.st + prv * np; m.st_out = w.st + cur * np;
        m.pold = (const cplx<R>*)w.P[prv]; m.pnew = (cplx<R>*)w.P[cur];
        rc = sizeof(R) == 8 ? launch_mdm_ragged<R, MDM_CG, true>(m, *h, s) : launch_mdm_ragged<R, MDM_CG, false>(m, *h, s);
        if (rc) return rc;
        if (mixed && i == 0) {
            hipLaunchKernelGGL(ru_init_kernel, dim3((np + 63) / 64), dim3(64), 0, s, (const double*)w.normb, w.ru_rr + np, w.ru_it + np,
                               w.ru_flag + np, np);
            QF_LAUNCHED();
        }
        u.rold = (const cplx<R>*)w.R[prv]; u.op = (const cplx<R>*)w.OP; u.p = (const cplx<R>*)w.P[cur];
        u.x = x; u.rnew = (cplx<R>*)w.R[cur]; u.z = (cplx<R>*)w.Z;
        u.pap = w.pap; u.st = w.st + cur * np;
        u.rz_part = w.rzp + cur * ps; u.rr_part = w.rrp + cur * ps;
        u.ru_flag = nullptr;
        rc = launch_cgu_ragged<R, CGU_ITER>(u, *h, s);
        if (rc) return rc;
        if (mixed && (i + 1) % kRuCheck == 0 && i + 1 < max_iter) {
            // reliable update where needed: x_hi += x_lo; r = b - A64 x_hi (FP64); z = P r
            rc = reliable_update(i + 1, cur, 0, cur);
            if (rc) return rc;
            CguArgs<R> f = u;
            f.rold = (const cplx<R>*)w.R[cur];
            f.ru_flag = w.ru_flag + ((check - 1) & 1) * np;
            rc = launch_cgu_ragged<R, CGU_REFRESH>(f, *h, s);
            if (rc) return rc;
        }
    }
    const int last = (max_iter - 1) & 1;
    if (mixed) {
        rc = reliable_update(max_iter, last, 1, last);
        if (rc) return rc;
        hipLaunchKernelGGL(cg_finalize_kernel, dim3(np), dim3(64), 0, s, (const double*)w.rp64, 0, (const CgState*)(w.st + last * np),
                           (const double*)w.normb, rtol, 1, err_out, iters_out, status_out, pd);
    } else {
        hipLaunchKernelGGL(cg_finalize_kernel, dim3(np), dim3(64), 0, s, (const double*)(w.rrp + last * ps), 0,
                           (const CgState*)(w.st + last * np), (const double*)w.normb, rtol, 0, err_out, iters_out, status_out, pd);
    }
    QF_LAUNCHED();
    return QF_OK;
}

// The ragged PCG as one persistent launch: the plan's operator and update work items,
// work-groups of at least 256 threads (the 16 x 64 update tiles of the plan, four slices per
// thread), x updated every iteration as in pcg_ragged.
template <typename R>
int pcg_ragged_persistent(const RaggedHead* h, const Prob* pd, const R* phi, const cplx<R>* b, cplx<R>* x, double dtau, double t,
                          const R* ptab, double rtol, int max_iter, const RgWork& w, void* bar, double* err_out, double* iters_out,
                          int* status_out, hipStream_t s) {
    const int np = h->nprob;
    const int blk = std::max(h->nthreads, kPersTX * h->tb / kPersSPT);
    if (blk > 512) return fail(QF_ERR_UNSUPPORTED, "persistent ragged solve: work-groups of %d threads (at most 512: L up to 44)", blk);
    PcgPersArgs<R> a{};
    a.u.ptab = ptab; a.u.prob = pd; a.u.nprob = np; a.u.TB = h->tb;
    a.m.phi = phi; a.m.hop = (const cplx<R>*)w.HOP; a.m.hp = make_hop<R>(dtau, t); a.m.prob = pd; a.m.nprob = np; a.m.nrhs = 1;
    a.m.normb = w.normb; a.m.rtol = rtol; a.m.dot_part = w.pap; a.m.z = (const cplx<R>*)w.Z; a.m.y = (cplx<R>*)w.OP;
    a.nrhs = np; a.max_iter = max_iter; a.paired = 0;
    a.b = b; a.x = x;
    a.R2[0] = (cplx<R>*)w.R[0]; a.R2[1] = (cplx<R>*)w.R[1]; a.P2[0] = (cplx<R>*)w.P[0]; a.P2[1] = (cplx<R>*)w.P[1];
    a.Z = (cplx<R>*)w.Z; a.OP = (cplx<R>*)w.OP;
    a.rzp = w.rzp; a.rrp = w.rrp; a.pap = w.pap; a.ps = h->nB_tot;
    a.st = w.st; a.normb = w.normb; a.brk = w.brk;
    if constexpr (sizeof(R) == 8) {
        a.phi = phi; a.hop = (cplx<R>*)w.HOP;
        a.th1 = std::tanh(t * dtau); a.th2 = std::tanh(0.5 * t * dtau);
    }
    a.rtol = rtol; a.err_out = err_out; a.iters_out = iters_out; a.status_out = status_out;
    unsigned* c = static_cast<unsigned*>(bar);
    a.gb.count = c; a.gb.done = c + 16; a.gb.abort = c + 32;
    a.prob = pd; a.nAi = h->nwg_op; a.nBi = h->nwg_upd; a.nsites = h->nsites;
    return launch_pers<R, 2>(a, blk, s);
}
}  // namespace

extern "C" {

size_t qf_pcg_ragged_workspace_bytes(int prec, const void* plan_host) {
    if (!plan_host || (prec != QF_PREC_F32 && prec != QF_PREC_F64)) return 0;
    const RaggedHead* h = static_cast<const RaggedHead*>(plan_host);
    if (h->nprob < 1) return 0;
    return carve_ragged(nullptr, prec, false, h).bytes;
}

size_t qf_pcg_mixed_ragged_workspace_bytes(const void* plan_host) {
    if (!plan_host) return 0;
    const RaggedHead* h = static_cast<const RaggedHead*>(plan_host);
    if (h->nprob < 1) return 0;
    return carve_ragged(nullptr, QF_PREC_F32, true, h).bytes;
}

int qf_pcg_solve_ragged(int prec, const void* plan_host, const void* plan_dev, const void* phi, const void* b, void* x, double dtau,
                        double t, const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes, double* err_out,
                        double* iters_out, int* status_out, void* stream) {
    const RaggedHead* h;
    const Prob* pd;
    int rc = ragged_head(plan_host, plan_dev, &h, &pd);
    if (rc) return rc;
    if (!phi || !b || !x || !work || !err_out || !iters_out) return fail(QF_ERR_INVALID, "null pointer");
    if (max_iter < 1) return fail(QF_ERR_INVALID, "max_iter=%d must be >= 1", max_iter);
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    const RgWork w = carve_ragged(work, prec, false, h);
    if (work_bytes < w.bytes) return fail(QF_ERR_INVALID, "workspace of %zu bytes, need %zu", work_bytes, w.bytes);
    hipStream_t s = (hipStream_t)stream;
    if (prec == QF_PREC_F32)
        return pcg_ragged<float>(h, pd, (const float*)phi, (const float2*)b, (float2*)x, dtau, t, (const float*)ptab, rtol, max_iter, w,
                                 err_out, iters_out, status_out, s);
    return pcg_ragged<double>(h, pd, (const double*)phi, (const double2*)b, (double2*)x, dtau, t, (const double*)ptab, rtol, max_iter, w,
                              err_out, iters_out, status_out, s);
}

size_t qf_pcg_ragged_persistent_workspace_bytes(int prec, const void* plan_host) {
    const size_t n = qf_pcg_ragged_workspace_bytes(prec, plan_host);
    return n ? n + kPersBarBytes : 0;
}

int qf_pcg_solve_ragged_persistent(int prec, const void* plan_host, const void* plan_dev, const void* phi, const void* b, void* x,
                                   double dtau, double t, const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes,
                                   double* err_out, double* iters_out, int* status_out, void* stream) {
    const RaggedHead* h;
    const Prob* pd;
    int rc = ragged_head(plan_host, plan_dev, &h, &pd);
    if (rc) return rc;
    if (!phi || !b || !x || !work || !err_out || !iters_out) return fail(QF_ERR_INVALID, "null pointer");
    if (!ptab) return fail(QF_ERR_INVALID, "the persistent solve needs the preconditioner tables");
    if (max_iter < 1 || max_iter > 1000000) return fail(QF_ERR_INVALID, "max_iter=%d out of range [1, 1000000]", max_iter);
    if (prec != QF_PREC_F32 && prec != QF_PREC_F64) return fail(QF_ERR_INVALID, "prec %d", prec);
    const RgWork w = carve_ragged(work, prec, false, h);
    if (work_bytes < w.bytes + kPersBarBytes)
        return fail(QF_ERR_INVALID, "workspace of %zu bytes, need %zu", work_bytes, w.bytes + kPersBarBytes);
    void* bar = static_cast<char*>(work) + w.bytes;
    hipStream_t s = (hipStream_t)stream;
    if (prec == QF_PREC_F32)
        return pcg_ragged_persistent<float>(h, pd, (const float*)phi, (const float2*)b, (float2*)x, dtau, t, (const float*)ptab, rtol,
                                            max_iter, w, bar, err_out, iters_out, status_out, s);
    return pcg_ragged_persistent<double>(h, pd, (const double*)phi, (const double2*)b, (double2*)x, dtau, t, (const double*)ptab, rtol,
                                         max_iter, w, bar, err_out, iters_out, status_out, s);
}

int qf_pcg_mixed_solve_ragged(const void* plan_host, const void* plan_dev, const double* phi64, const float* phi32, const void* b, void* x,
                              double dtau, double t, const float* ptab32, double rtol, int max_iter, int ru_period, void* work,
                              size_t work_bytes, double* err_out, double* iters_out, int* status_out, void* stream) {
    const RaggedHead* h;
    const Prob* pd;
    int rc = ragged_head(plan_host, plan_dev, &h, &pd);
    if (rc) return rc;
    if (!phi64 || !phi32 || !b || !x || !work || !err_out || !iters_out) return fail(QF_ERR_INVALID, "null pointer");
    if (max_iter < 1) return fail(QF_ERR_INVALID, "max_iter=%d must be >= 1", max_iter);
    if (ru_period < 1) return fail(QF_ERR_INVALID, "reliable-update period %d must be >= 1", ru_period);
    const RgWork w = carve_ragged(work, QF_PREC_F32, true, h);
    if (work_bytes < w.bytes) return fail(QF_ERR_INVALID, "workspace of %zu bytes, need %zu", work_bytes, w.bytes);
    hipStream_t s = (hipStream_t)stream;
    const long long n = h->nsites;
    const dim3 g1(std::min<long long>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(demote_kernel, g1, dim3(256), 0, s, (const double2*)b, (float2*)w.BLO, n);
    QF_LAUNCHED();
    hipLaunchKernelGGL(zero_kernel, g1, dim3(256), 0, s, (double*)x, 2 * n);
    QF_LAUNCHED();
    return pcg_ragged<float>(h, pd, phi32, (const float2*)w.BLO, (float2*)w.XLO, dtau, t, ptab32, rtol, max_iter, w, err_out, iters_out,
                             status_out, s, ru_period, phi64, (const double2*)b, (double2*)x);
}

int qf_fermion_force(int prec, const void* phi, long long phi_cs, const void* xi, void* Ft, int nrhs, int L, int Ltau, double dtau,
                     double t, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!phi || !xi || !Ft) return fail(QF_ERR_INVALID, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    const Tiling tl = tiling(prec, nrhs, L, Ltau);
    const long long N = (long long)L * L * Ltau;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        ForceArgs<R> a{};
        a.phi = (const R*)phi; a.phi_cs = phi_cs; a.xi = (const cplx<R>*)xi; a.vs = N; a.Ft = (R*)Ft; a.f_cs = 2 * N;
        a.L = L; a.Lt = Ltau; a.T = tl.T; a.hp = make_hop<R>(dtau, t);
        switch (tl.by) {
            case 2: return launch_force_by<R, 2>(a, nrhs, s);
            case 4: return launch_force_by<R, 4>(a, nrhs, s);
            case 6: return launch_force_by<R, 6>(a, nrhs, s);
        }
        return fail(QF_ERR_UNSUPPORTED, "patch height %d", tl.by);
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

int qf_boson_force(const double* x, double* Fp, double* Ft, int nrhs, int L, int Ltau, double cJ, double cK, int compact, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!x || !Fp || !Ft) return fail(QF_ERR_INVALID, "null pointer");
    BosonArgs a{};
    a.x = (double*)x; a.Fp = Fp; a.Ft = Ft; a.cs = 2LL * L * L * Ltau; a.L = L; a.Lt = Ltau; a.nrhs = nrhs;
    a.cJ = cJ; a.cK = cK; a.compact = compact; a.nsub = 1; a.m = 1.0;
    const long long ns = (long long)nrhs * L * L * Ltau;
    hipLaunchKernelGGL(boson_force_kernel, dim3(std::min<long long>((ns + 255) / 256, 16384)), dim3(256), 0, (hipStream_t)stream, a, 0);
    QF_LAUNCHED();
    return QF_OK;
}

int qf_leapfrog_boson(double* x, double* p, double* Fp, double* Ft, const double* dt, const double* tau_mask, double m, int nsub, int nrhs,
                      int L, int Ltau, double cJ, double cK, int compact, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!x || !p || !Fp || !Ft || !dt) return fail(QF_ERR_INVALID, "null pointer");
    if (nsub < 1) return fail(QF_ERR_INVALID, "nsub=%d", nsub);
    BosonArgs a{};
    a.x = x; a.p = p; a.Fp = Fp; a.Ft = Ft; a.dt = dt; a.tau_mask = tau_mask; a.cs = 2LL * L * L * Ltau; a.L = L; a.Lt = Ltau;
    a.nrhs = nrhs; a.cJ = cJ; a.cK = cK; a.m = m; a.compact = compact; a.nsub = nsub;
    hipStream_t s = (hipStream_t)stream;
    const long long ns = (long long)nrhs * L * L * Ltau, nr = 2 * ns;
    for (int k = 0; k < nsub; ++k) {
        hipLaunchKernelGGL(boson_kick_drift_kernel, dim3(std::min<long long>((nr + 255) / 256, 16384)), dim3(256), 0, s, a);
        QF_LAUNCHED();
        hipLaunchKernelGGL(boson_force_kernel, dim3(std::min<long long>((ns + 255) / 256, 16384)), dim3(256), 0, s, a, 1);
        QF_LAUNCHED();
    }
    return QF_OK;
}

int qf_momentum_kick(double* p, const double* F, const double* dt, double scale, const double* tau_mask, int nrhs, int L, int Ltau,
                     void* stream) {
    if (!p || !F || !dt) return fail(QF_ERR_INVALID, "null pointer");
    if (nrhs < 1 || L < 1 || Ltau < 1) return fail(QF_ERR_INVALID, "bad sizes");
    const long long cs = 2LL * L * L * Ltau, n = nrhs * cs;
    hipLaunchKernelGGL(momentum_kick_kernel, dim3(std::min<long long>((n + 255) / 256, 16384)), dim3(256), 0, (hipStream_t)stream, p, F, dt,
                       scale, tau_mask, cs, 2 * L * L, n);
    QF_LAUNCHED();
    return QF_OK;
}

size_t qf_hamiltonian_workspace_bytes(int nrhs) { return sizeof(double) * 4 * 64 * (size_t)std::max(nrhs, 1); }

int qf_hamiltonian(int prec, const double* x, const double* p, const void* psi, const void* xi, double* H, double* parts, int nrhs, int L,
                   int Ltau, double cJ, double cK, int compact, double m, void* work, size_t work_bytes, void* stream) {
    int rc = check_lattice(L, Ltau, nrhs);
    if (rc) return rc;
    if (!x || !H || !work) return fail(QF_ERR_INVALID, "null pointer");
    if ((psi == nullptr) != (xi == nullptr)) return fail(QF_ERR_INVALID, "psi and xi must both be given or both be null");
    if (work_bytes < qf_hamiltonian_workspace_bytes(nrhs)) return fail(QF_ERR_INVALID, "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const int nblk = 64;
    const long long cs = 2LL * L * L * Ltau, vs = (long long)L * L * Ltau;
    double* part = (double*)work;
    if (prec == QF_PREC_F64)
        hipLaunchKernelGGL((hamiltonian_partial_kernel<double>), dim3(nblk, nrhs), dim3(256), 0, s, x, p, (const double2*)psi,
                           (const double2*)xi, cs, vs, L, Ltau, cJ, cK, compact, m, part);
    else if (prec == QF_PREC_F32)
        hipLaunchKernelGGL((hamiltonian_partial_kernel<float>), dim3(nblk, nrhs), dim3(256), 0, s, x, p, (const float2*)psi,
                           (const float2*)xi, cs, vs, L, Ltau, cJ, cK, compact, m, part);
    else
        return fail(QF_ERR_INVALID, "prec %d", prec);
    QF_LAUNCHED();
    hipLaunchKernelGGL(hamiltonian_finalize_kernel, dim3(nrhs), dim3(64), 0, s, (const double*)part, nblk, H, parts);
    QF_LAUNCHED();
    return QF_OK;
}

// ---------------------------------------------------------------- bond-bond estimators
int qf_obs_phase(int prec, const void* phix, void* U, long long n, void* stream) {
    if (!phix || !U) return fail(QF_ERR_INVALID, "null pointer");
    if (n < 1) return fail(QF_ERR_INVALID, "bad sizes");
    const dim3 grid((unsigned)std::min<long long>((n + 255) / 256, 16384)), blk(256);
    if (prec == QF_PREC_F32)
        hipLaunchKernelGGL((obs_phase_kernel<float>), grid, blk, 0, (hipStream_t)stream, (const float*)phix, (float2*)U, n);
    else if (prec == QF_PREC_F64)
        hipLaunchKernelGGL((obs_phase_kernel<double>), grid, blk, 0, (hipStream_t)stream, (const double*)phix, (double2*)U, n);
    else
        return fail(QF_ERR_INVALID, "prec %d", prec);
    QF_LAUNCHED();
    return QF_OK;
}

int qf_obs_fields(int prec, int obs, const void* eta, const void* G, const void* U, const int* pairs, int p0, int np, int nrv, int L,
                  int Ltau, void* A, void* B, void* stream) {
    if (!eta || !G || !U || !A) return fail(QF_ERR_INVALID, "null pointer");
    if (obs < QF_OBS_T1 || obs > QF_OBS_V) return fail(QF_ERR_INVALID, "estimator %d", obs);
    if (obs != QF_OBS_V && (!pairs || !B)) return fail(QF_ERR_INVALID, "pair estimators need pairs and B");
    if (L < 2 || Ltau < 1 || nrv < 1 || np < 1 || p0 < 0) return fail(QF_ERR_INVALID, "bad sizes");
    if (obs == QF_OBS_V && p0 + np > nrv) return fail(QF_ERR_INVALID, "noise vectors [%d, %d) beyond nrv %d", p0, p0 + np, nrv);
    const long long N = (long long)L * L * Ltau, total = (long long)np * N;
    const dim3 grid((unsigned)std::min<long long>((total + 255) / 256, 32768)), blk(256);
    hipStream_t s = (hipStream_t)stream;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        ObsArgs<R> a{};
        a.eta = (const cplx<R>*)eta; a.G = (const cplx<R>*)G; a.U = (const cplx<R>*)U; a.pairs = pairs;
        a.p0 = p0; a.np = np; a.nrv = nrv; a.L = L; a.N = N; a.A = (cplx<R>*)A; a.B = (cplx<R>*)B;
        switch (obs) {
            case QF_OBS_T1: hipLaunchKernelGGL((obs_fields_kernel<R, OBS_T1>), grid, blk, 0, s, a); break;
            case QF_OBS_T21: hipLaunchKernelGGL((obs_fields_kernel<R, OBS_T21>), grid, blk, 0, s, a); break;
            case QF_OBS_T2: hipLaunchKernelGGL((obs_fields_kernel<R, OBS_T2>), grid, blk, 0, s, a); break;
            case QF_OBS_T4: hipLaunchKernelGGL((obs_fields_kernel<R, OBS_T4>), grid, blk, 0, s, a); break;
            default: hipLaunchKernelGGL((obs_fields_kernel<R, OBS_V>), grid, blk, 0, s, a); break;
        }
        QF_LAUNCHED();
        return QF_OK;
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

int qf_obs_kspace(int prec, const void* A_hat, const void* B_hat, long long nslices, int L, int lin, int accumulate, void* part, int nwg,
                  void* stream) {
    if (!A_hat || !B_hat || !part) return fail(QF_ERR_INVALID, "null pointer");
    if (L < 2 || nslices < 1 || nwg < 1 || nwg > 65535) return fail(QF_ERR_INVALID, "bad sizes");
    hipStream_t s = (hipStream_t)stream;
    if (prec == QF_PREC_F32)
        hipLaunchKernelGGL((obs_kspace_kernel<float>), dim3(nwg), dim3(256), 0, s, (const float2*)A_hat, (const float2*)B_hat, nslices, L,
                           lin, accumulate, (double2*)part, nwg);
    else if (prec == QF_PREC_F64)
        hipLaunchKernelGGL((obs_kspace_kernel<double>), dim3(nwg), dim3(256), 0, s, (const double2*)A_hat, (const double2*)B_hat, nslices,
                           L, lin, accumulate, (double2*)part, nwg);
    else
        return fail(QF_ERR_INVALID, "prec %d", prec);
    QF_LAUNCHED();
    return QF_OK;
}

int qf_obs_reduce(const void* part, int nw, int n, void* out, void* stream) {
    if (!part || !out) return fail(QF_ERR_INVALID, "null pointer");
    if (nw < 1 || n < 1) return fail(QF_ERR_INVALID, "bad sizes");
    hipLaunchKernelGGL(obs_reduce_kernel, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, (hipStream_t)stream, (const double2*)part, nw, n,
                       (double2*)out);
    QF_LAUNCHED();
    return QF_OK;
}

int qf_obs_diag(int prec, const void* eta, const void* G, const void* U, int nrv, int L, int Ltau, void* part, int nblk, void* stream) {
    if (!eta || !G || !U || !part) return fail(QF_ERR_INVALID, "null pointer");
    if (L < 2 || Ltau < 1 || nrv < 1 || nblk < 1 || nblk > 65535) return fail(QF_ERR_INVALID, "bad sizes");
    const long long N = (long long)L * L * Ltau;
    hipStream_t s = (hipStream_t)stream;
    if (prec == QF_PREC_F32)
        hipLaunchKernelGGL((obs_diag_kernel<float>), dim3(nblk), dim3(256), 0, s, (const float2*)eta, (const float2*)G, (const float2*)U, nrv,
                           L, N, (double2*)part);
    else if (prec == QF_PREC_F64)
        hipLaunchKernelGGL((obs_diag_kernel<double>), dim3(nblk), dim3(256), 0, s, (const double2*)eta, (const double2*)G,
                           (const double2*)U, nrv, L, N, (double2*)part);
    else
        return fail(QF_ERR_INVALID, "prec %d", prec);
    QF_LAUNCHED();
    return QF_OK;
}

int qf_obs_fused_grid(int prec, int L) {
    if (L < 2 || L > 78) return 0;
    unsigned long long fc;
    int n;
    if (!obs_factor(L, fc, n)) return 0;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) return 0;
    // resident work-groups per CU of the least-occupying pair-estimator kernel at this L (Tvv's
    // kernel, a twentieth of the work at Nrv = 40, may then run its work-groups in two rounds)
    int best = 1 << 30;
    auto occ = [&](auto tag) {
        using R = decltype(tag);
        for (int obs = QF_OBS_T1; obs <= QF_OBS_T4; ++obs) {
            const int rc = obs_fused_dispatch<R>(obs, L, [&](auto kernel, size_t lds) -> int {
                int b = 0;
                if (set_lds(kernel, lds) != QF_OK || hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, kObsThreads, lds) != hipSuccess)
                    b = 0;
                best = std::min(best, b);
                return QF_OK;
            });
            if (rc) best = 0;
        }
    };
    if (prec == QF_PREC_F32) occ(float{});
    else if (prec == QF_PREC_F64) occ(double{});
    else return 0;
    if (best <= 0 || best == (1 << 30)) return 0;
    return std::min(best * ncu, 65535);
}

int qf_obs_fused(int prec, int obs, const void* eta, const void* G, const void* U, const int* pairs, int count, int nrv, int L, int Ltau,
                 void* part, int nwg, void* stream) {
    if (!eta || !G || !U || !part) return fail(QF_ERR_INVALID, "null pointer");
    if (obs < QF_OBS_T1 || obs > QF_OBS_V) return fail(QF_ERR_INVALID, "estimator %d", obs);
    if (obs != QF_OBS_V && !pairs) return fail(QF_ERR_INVALID, "pair estimators need pairs");
    if (L < 2 || Ltau < 1 || nrv < 1 || count < 1 || nwg < 1 || nwg > 65535) return fail(QF_ERR_INVALID, "bad sizes");
    if (obs == QF_OBS_V && count > nrv) return fail(QF_ERR_INVALID, "%d noise vectors beyond nrv %d", count, nrv);
    if (L > 78) return fail(QF_ERR_UNSUPPORTED, "fused estimator kernel: L = %d > 78", L);
    unsigned long long facs;
    int nfac;
    if (!obs_factor(L, facs, nfac)) return fail(QF_ERR_UNSUPPORTED, "fused estimator kernel: L = %d has too many factors", L);
    hipStream_t s = (hipStream_t)stream;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        ObsFusedArgs<R> a{};
        a.eta = (const cplx<R>*)eta; a.G = (const cplx<R>*)G; a.U = (const cplx<R>*)U; a.pairs = pairs;
        a.count = count; a.nrv = nrv; a.L = L; a.Lt = Ltau; a.N = (long long)L * L * Ltau; a.nfac = nfac; a.facs = facs;
        a.part = (double2*)part;
        return obs_fused_dispatch<R>(obs, L, [&](auto kernel, size_t lds) -> int {
            int rc = set_lds(kernel, lds);
            if (rc) return rc;
            hipLaunchKernelGGL(kernel, dim3(nwg), dim3(kObsThreads), lds, s, a);
            QF_LAUNCHED();
            return QF_OK;
        });
    };
    if (prec == QF_PREC_F32) return run(float{});
    if (prec == QF_PREC_F64) return run(double{});
    return fail(QF_ERR_INVALID, "prec %d", prec);
}

}  // extern "C"


