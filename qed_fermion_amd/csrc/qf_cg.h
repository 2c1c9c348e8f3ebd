// qf_cg.h -- fused PCG update step and preconditioner application.
//
// The preconditioner of the reference (preconditioners_orig/precon_manual.py:
// 15-175, loaded at hmc_sampler_batch.py:461-547) is P = P_tau (x) I_Vs with
// P_tau a symmetric 13-band matrix plus antiperiodic corners.  Here it is a
// cyclic 13-point tau stencil: z_tau = sum_{o=-6..6} c[tau][o] r_{(tau+o) mod Nt}
// (the corners (0, Nt-r) and (Nt-r, 0) are the cyclic offsets -r / +r of rows
// 0 and Nt-r, which the open band never uses).
//
// A block owns a tile of sites x tau slices; the new residual of the tile and
// of its 12-slice halo is staged in LDS, so every r value is read from memory
// (1 + 12/TT) times instead of 13 times.  The tile's rows of the band table go
// through LDS too, loaded together with the vectors, so no global load is
// waited for after the barrier.
//
// Update step B_i of preconditioned_cg_fast_test (hmc_sampler_batch.py:797-844):
//   alpha = rz_old / Re(p^dag Op);  x += alpha p;  r -= alpha Op;  z = P r;
//   partial sums of Re(r^dag z) and |r|^2 per block (FP64).
// r is double buffered (rold -> rnew) because z needs the new r of the
// neighbouring slices, which it recomputes on the fly from rold and Op.
//
// x is only needed at the end of a solve and at a reliable update, so the
// iterations may hold one step back: a held iteration (xmode 1) keeps its
// alpha and leaves x untouched; the next iteration (xmode 2) applies
// x += alpha' p' + alpha p with p' still in the other p buffer.  In FP32 or
// FP64 the stored x equals the step-by-step x bit for bit (the same two
// rounded multiply-adds, the first no longer stored in between).  A chain that
// stops right after a held iteration applies alpha' p' alone.
#pragma once
#include "qf_reduce.h"

namespace qf {

enum { CGU_INIT = 0, CGU_ITER = 1, CGU_REFRESH = 2, CGU_PRECON = 3 };

template <typename R>
struct CguArgs {
    const cplx<R>* rold;   // INIT: b; ITER: r_{i-1}; REFRESH/PRECON: r
    const cplx<R>* op;     // ITER: M^dag M p
    const cplx<R>* p;      // ITER
    cplx<R>* x;            // ITER: x += alpha p; INIT: x = 0
    cplx<R>* rnew;         // INIT/ITER: new r
    cplx<R>* z;            // z = P r
    long long vs;
    int Vs, Lt, TB;
    const R* ptab;         // [Lt][13] or nullptr (no preconditioner)
    const double* pap;     // ITER: Re(p^dag Op) partials [nrhs][nA]
    int nA;
    const CgState* st;     // ITER: state written by step A of this iteration
    CgState* st_init;      // INIT: state slot read by step A_0
    double* rz_part;       // [nrhs][nblk]
    double* rr_part;
    const int* ru_flag;    // REFRESH: skip chains whose reliable-update flag is 0 (nullptr: never skip)
    int xmode;             // ITER: 0 x += alpha p; 1 hold (alpha into xalpha, x untouched);
                           //       2 x += alpha' p' + alpha p (alpha', p' of the held iteration)
    const cplx<R>* pprev;  // ITER, xmode 2: p' (search direction of the held iteration)
    double* xalpha;        // ITER: [nrhs] alpha of the held iteration (xmode 1 writes, 2 reads)
    const CgState* st_held;  // ITER, xmode 2: state written by step A of the held iteration
    int* brk;              // [nrhs] breakdown flag of the last ITER step (INIT clears it; optional)
    const Prob* prob;      // ragged batch (nullptr: uniform; grid (nsx, tiles, nrhs))
    int nprob;
    int xcd;               // uniform grid: 1 runs the tiles in XCD order (cg_update_kernel)
};

// A block of TX * TT / SPT threads owns a tile of TX sites x TT slices; each
// thread owns SPT slices (rows sr, sr + RT, ...) of one site.  The new
// residual of the tile's TT + 12 slices (the 13-point tau window) is staged in
// LDS, so no thread marches in tau and all loads of a block are in flight at
// once.  Slices outside [0, Lt) of the window wrap cyclically (P is cyclic).
// Partial sums: one (rz, rr) pair per block, blk = bx + nsx * by.
// TT = 0: the tile height is A.TB, set at run time for a persistent launch
// whose work-group size is set by the operator phase (TX * A.TB / SPT threads
// work, at most blockDim.x; the rest only join the barriers); the staged
// residual and band rows then live in the caller's LDS at stage
// (cgu_stage_bytes).  The body takes the tile
// coordinates (bx, by) of rhs bz in a grid of gx x gy tiles per rhs (lin:
// linear tile index of a ragged batch); consecutive calls in one work-group
// must be separated by a work-group barrier.
template <typename R>
__host__ __device__ constexpr int cgu_cst() { return (13 * (int)sizeof(R) + 15) / 16 * 16 / (int)sizeof(R); }
template <typename R>
__host__ __device__ constexpr unsigned cgu_ring_bytes(int tt, int tx) { return (unsigned)((tt + 12) * tx) * (unsigned)sizeof(cplx<R>); }
template <typename R>
__host__ __device__ constexpr unsigned cgu_stage_bytes(int tt, int tx) {
    return (cgu_ring_bytes<R>(tt, tx) + 15u) / 16u * 16u + (unsigned)(tt * cgu_cst<R>()) * (unsigned)sizeof(R);
}

template <typename R, int MODE, bool PRE, int TX, int TT, int SPT>
__device__ __forceinline__ void cgu_body(const CguArgs<R>& A, int lin, int bx_, int by_, int bz_, int gx, int gy,
                                         unsigned char* stage = nullptr) {
    constexpr int H = PRE ? 12 : 0;   // halo slices
    const int tt = TT > 0 ? TT : A.TB;
    const int ROWS = tt + H;
    const int RT = tt / SPT;      // thread rows
    // loads per thread of the tile and its halo, ceil((TT + H) / RT); a runtime tile needs at
    // most that of the smallest work-group (64 threads, RT = 64 / TX)
    constexpr int NLD = TT > 0 ? (TT + H + TT / SPT - 1) / (TT / SPT) : SPT + (H * TX + 63) / 64;
    static_assert(TT % SPT == 0, "slices per thread must divide the tile");
    // preconditioner rows of the tile's TT slices, padded to 16-byte multiples
    constexpr int CST = cgu_cst<R>();
    __shared__ cplx<R> ring_s[PRE && TT > 0 ? (TT + H) * TX : 1];
    __shared__ __align__(16) R crow_s[PRE && TT > 0 ? TT * CST : 1];
    cplx<R>* const ring = TT > 0 ? ring_s : reinterpret_cast<cplx<R>*>(stage);
    R* const crow = TT > 0 ? crow_s : reinterpret_cast<R*>(stage + (cgu_ring_bytes<R>(tt, TX) + 15u) / 16u * 16u);
    __shared__ double red[64];
    const int NTHR = TX * tt / SPT;
    const bool live = TT > 0 || (int)threadIdx.x < NTHR;   // runtime tiles may leave threads idle
    constexpr int NCE = PRE ? (13 * SPT + TX - 1) / TX : 1;   // coefficients per thread
    int rho, Vs, Lt, bx, by, blk, nA;
    long long vo, pout, pin;
    const R* ptab = A.ptab;
    if (A.prob != nullptr) {   // ragged batch: one linear grid over every lattice's tiles
        int k = 0;
        for (int j = 1; j < A.nprob; ++j)
            if (A.prob[j].bwg0 <= lin) k = j;
        const Prob& P = A.prob[k];
        rho = k; Vs = P.L * P.L; Lt = P.Lt; blk = lin - P.bwg0;
        bx = blk % P.nsx; by = blk / P.nsx;
        vo = P.voff; pout = P.pb0 + blk; pin = P.pa0; nA = P.nA;
        if (ptab) ptab += P.toff;
    } else {
        rho = bz_; Vs = A.Vs; Lt = A.Lt; bx = bx_; by = by_;
        blk = bx + gx * by;
        vo = rho * A.vs; pout = (long long)rho * gx * gy + blk; pin = (long long)rho * A.nA; nA = A.nA;
    }
    const int sx = threadIdx.x % TX, sr = threadIdx.x / TX;
    const int s = bx * TX + sx;
    const int t0 = by * tt;
    const bool col = s < Vs;
    const int sc = col ? s : 0;   // loads of padding columns read site 0 (never stored)

    if constexpr (MODE == CGU_ITER) {
        if (A.st[rho].act == 0.0) {   // converged chain: x and r stay frozen (uniform)
            // stopped right after a held iteration (ah = 0: the held step broke down, x stays);
            // xalpha and st_held are only read in paired mode
            const R ah = (A.xmode == 2 && A.st_held[rho].act != 0.0) ? (R)A.xalpha[rho] : R(0);
            if (ah != R(0)) {
#pragma unroll
                for (int j = 0; j < SPT; ++j) {
                    const int tau = t0 + sr + j * RT;
                    if (live && col && tau < Lt) {
                        const long long io = vo + (long long)tau * Vs + s;
                        const cplx<R> xv = A.x[io], pv = A.pprev[io];
                        A.x[io] = mk<R>(xv.x + ah * pv.x, xv.y + ah * pv.y);
                    }
                }
            }
            return;
        }
    }
    if constexpr (MODE == CGU_REFRESH) {
        if (A.ru_flag != nullptr && A.ru_flag[rho] == 0) return;
    }
    if constexpr (MODE == CGU_INIT) {
        if (blk == 0 && threadIdx.x == 0) {
            CgState o; o.rz = 1.0; o.act = 1.0; o.err = 1.0; o.iters = 0.0; o.stat = 0.0;
            A.st_init[rho] = o;
            if (A.brk != nullptr) A.brk[rho] = 0;
        }
    }

    // ---- loads (none depends on the step size); the partials of Re(p^dag Op) first
    Partials<1, 2> pf{};
    if constexpr (MODE == CGU_ITER) {
        const double* src[1] = {A.pap + pin};
        pf = load_partials<1, 2>(src, nA);
    }
    const cplx<R>* rold = A.rold + vo;
    const cplx<R>* op = A.op + vo;
    auto slice_of = [&](int h) { return ((t0 - H / 2 + h) % Lt + Lt) % Lt; };
    cplx<R> rw[PRE ? NLD : SPT], ow[PRE ? NLD : SPT];
    if constexpr (PRE) {
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            const int h = sr + k * RT;
            const long long i = (long long)slice_of(h < ROWS ? h : 0) * Vs + sc;
            rw[k] = mk<R>(0, 0); ow[k] = mk<R>(0, 0);
            if (h < ROWS) {
                rw[k] = rold[i];
                if constexpr (MODE == CGU_ITER) ow[k] = op[i];
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const long long i = (long long)min(t0 + sr + j * RT, Lt - 1) * Vs + sc;
            rw[j] = rold[i];
            ow[j] = mk<R>(0, 0);
            if constexpr (MODE == CGU_ITER) ow[j] = op[i];
        }
    }
    // the stencil's coefficients go through LDS with the ring, so no global load
    // waits behind the barrier; they are loaded here with the vectors
    R cv[NCE];
    if constexpr (PRE) {
#pragma unroll
        for (int q = 0; q < NCE; ++q) {
            const int e = threadIdx.x + q * NTHR;
            const int row = e / 13, k = e - row * 13;
            cv[q] = e < tt * 13 ? ptab[min(t0 + row, Lt - 1) * 13 + k] : R(0);
        }
    }
    cplx<R> pp[SPT], xx[SPT], ph[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        pp[j] = xx[j] = ph[j] = mk<R>(0, 0);
        if constexpr (MODE == CGU_ITER) {
            const long long io = vo + (long long)min(t0 + sr + j * RT, Lt - 1) * Vs + sc;
            if (A.xmode != 1) {
                pp[j] = A.p[io];
                xx[j] = A.x[io];
            }
            if (A.xmode == 2) ph[j] = A.pprev[io];
        }
    }

    R alpha = R(0);
    bool brk = false;   // uniform over the chain: every block reduces the same partials in the same order
    if constexpr (MODE == CGU_ITER) {
        double sums[1];
        const double* src[1] = {A.pap + pin};
        sum_partials_pf<1, 2>(pf, src, nA, sums, red);
        const double a64 = A.st[rho].rz / sums[0];
        // breakdown: Re(p^dag A p) <= 0 or a non-finite step.  The step is skipped (x and r
        // stay, so no NaN reaches them) and step A of the next iteration stops the chain.
        const bool nonfinite = !isfinite(sums[0]) || (sums[0] > 0.0 && !isfinite(a64));
        brk = nonfinite || !(sums[0] > 0.0);
        if (brk && blk == 0 && threadIdx.x == 0 && A.brk != nullptr) A.brk[rho] = nonfinite ? kSolveNonfinite : kSolveBreakdown;
        alpha = brk ? R(0) : (R)a64;
        if (A.xmode == 1 && blk == 0 && threadIdx.x == 0) A.xalpha[rho] = (double)alpha;
        if (A.xmode == 2) {   // x_{i} = x_{i-1} + alpha' p', kept in registers
            const R ah = (R)A.xalpha[rho];
            if (ah != R(0)) {
#pragma unroll
                for (int j = 0; j < SPT; ++j) xx[j] = mk<R>(xx[j].x + ah * ph[j].x, xx[j].y + ah * ph[j].y);
            }
        }
        if (!brk) {
#pragma unroll
            for (int k = 0; k < (PRE ? NLD : SPT); ++k) rw[k] = mk<R>(rw[k].x - alpha * ow[k].x, rw[k].y - alpha * ow[k].y);
        }
    }
    if constexpr (PRE) {
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            const int h = sr + k * RT;
            if (live && h < ROWS) ring[h * TX + sx] = rw[k];
        }
#pragma unroll
        for (int q = 0; q < NCE; ++q) {
            const int e = threadIdx.x + q * NTHR;
            const int row = e / 13;
            if (live && e < tt * 13) crow[row * CST + (e - row * 13)] = cv[q];
        }
        __syncthreads();
    }

    double rr = 0.0, rz = 0.0;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int st = sr + j * RT;
        const int tau = t0 + st;
        cplx<R> r, zz;
        if constexpr (PRE) {
            r = ring[(st + H / 2) * TX + sx];
            const R* c = crow + st * CST;
            zz = mk<R>(0, 0);
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                const cplx<R> w = ring[(st + k) * TX + sx];
                zz = mk<R>(zz.x + c[k] * w.x, zz.y + c[k] * w.y);
            }
        } else {
            r = rw[j];
            zz = r;
        }
        if (live && col && tau < Lt) {
            const long long io = vo + (long long)tau * Vs + s;
            if constexpr (MODE == CGU_ITER) {
                if (A.xmode != 1) A.x[io] = brk ? xx[j] : mk<R>(xx[j].x + alpha * pp[j].x, xx[j].y + alpha * pp[j].y);
            }
            if constexpr (MODE == CGU_INIT) A.x[io] = mk<R>(0, 0);
            if constexpr (MODE == CGU_INIT || MODE == CGU_ITER) A.rnew[io] = r;
            A.z[io] = zz;
            rr += (double)r.x * r.x + (double)r.y * r.y;
            rz += (double)r.x * zz.x + (double)r.y * zz.y;
        }
    }
    if constexpr (MODE != CGU_PRECON) {
        double v[2] = {rz, rr};
        block_sum<2>(v, red);
        if (threadIdx.x == 0) {
            A.rz_part[pout] = v[0];
            A.rr_part[pout] = v[1];
        }
    }
}

// A tile's rows of 16 sites are whole 128-byte lines only where a slice is a whole number of
// lines; elsewhere (complex64 at L = 66: 4356 sites, 34848 B) each row shares a line with the
// next column's row, and with A.xcd the tiles run in XCD order (xcd_order) so that one XCD runs
// both columns and fetches the line once (the launcher sets it for such slices only: on aligned
// slices the order measured a few percent slower).
template <typename R, int MODE, bool PRE, int TX, int TT, int SPT>
__global__ void __launch_bounds__(TX * TT / SPT) cg_update_kernel(CguArgs<R> A) {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int G = gx * gy * (int)gridDim.z;
    const int w = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
    const int l = A.xcd ? xcd_order(w, G) : w;
    const int bx = l % gx, rest = l / gx;
    cgu_body<R, MODE, PRE, TX, TT, SPT>(A, l, bx, rest % gy, rest / gy, gx, gy);
}

// Scalars after the last iteration (the reference returns the error of the
// last iteration and, without graphs, the iteration counter, 828-846).
// One block per RHS.  If true_res, rr partials hold |b - A x|^2 in FP64 and the
// error is reported from them even for converged chains.  status_out (optional)
// gets the chain's kSolve* code.
__device__ __forceinline__ void cg_finalize_body(int rho, const double* rr_part, int n, const CgState* st_prev, const double* normb,
                                                 double rtol, int true_res, double* err_out, double* iters_out, int* status_out,
                                                 const Prob* prob) {
    __shared__ double red[64];
    double sums[1];
    long long off = (long long)rho * n;
    if (prob != nullptr) {   // ragged batch: this lattice's update (or, true_res, operator) partials
        off = true_res ? prob[rho].pa0 : prob[rho].pb0;
        n = true_res ? prob[rho].nA : prob[rho].nB;
    }
    const double* src[1] = {rr_part + off};
    sum_partials<1>(src, n, sums, red);
    if (threadIdx.x == 0) {
        const CgState S = st_prev[rho];
        const double nb = normb[rho];
        double err, it = S.iters;
        if (true_res) {
            err = nb > 0.0 ? sqrt(sums[0]) / nb : 0.0;
            if (S.act != 0.0 && err > rtol) it += 1.0;
        } else if (S.act == 0.0) {
            err = S.err;
        } else {
            err = nb > 0.0 ? sqrt(sums[0]) / nb : 0.0;
            if (err > rtol) it += 1.0;
        }
        err_out[rho] = err;
        iters_out[rho] = it;
        if (status_out != nullptr) {
            int code = (int)S.stat;
            if (code == 0) code = !isfinite(err) ? kSolveNonfinite : err <= rtol ? kSolveConverged : kSolveMaxIter;
            status_out[rho] = code;
        }
    }
}

__global__ void cg_finalize_kernel(const double* rr_part, int n, const CgState* st_prev, const double* normb, double rtol,
                                   int true_res, double* err_out, double* iters_out, int* status_out, const Prob* prob) {
    cg_finalize_body((int)blockIdx.x, rr_part, n, st_prev, normb, rtol, true_res, err_out, iters_out, status_out, prob);
}

// Adaptive reliable update of the mixed-precision solver (checked every few
// iterations, entirely on the device so the solve stays graph-capturable):
// a chain updates when its iterated residual has dropped by delta since the
// last true residual, or has reached rtol (a chain may only stop on a true
// residual, so this check ends the solve without iterating on to the next
// delta drop), or ru_max iterations have passed, or at the end (force).
// The update flushes x_hi += x_lo, x_lo = 0 here; the FP64 residual kernel and
// the refresh kernel that follow read the per-chain flag and exit if it is 0.
// State is parity double-buffered per check (prev -> out) because every block
// reads it while block 0 of each chain writes the next value.
struct RuArgs {
    double2* xhi;
    float2* xlo;
    long long N;               // elements per chain
    const double* rr_part;     // iterated |r|^2 partials of the last update step, [nrhs][nB]
    int nB;
    const double* rp64;        // |b - A x_hi|^2 partials of the last FP64 residual, [nrhs][nA64]
    int nA64;
    const double* rr_prev;     // |r_true|^2 at the last reliable update (prev parity)
    double* rr_out;
    const double* it_prev;     // iteration of the last reliable update
    double* it_out;
    const int* flag_prev;      // 1 if the previous check performed an update
    int* flag_out;
    const CgState* st;         // active mask
    int iter;                  // iterations completed
    int ru_max;
    double delta2;
    const double* normb;       // |b| per chain
    double rtol2;              // rtol^2 (0: no convergence trigger)
    int force;
    const Prob* prob;          // ragged batch: per-lattice vector offsets and partial-sum ranges (nullptr: uniform)
};

__global__ void __launch_bounds__(256) ru_decide_flush_kernel(RuArgs A) {
    __shared__ double red[64];
    const int rho = blockIdx.y;
    const bool prev_updated = A.flag_prev[rho] != 0;
    const Prob* P = A.prob != nullptr ? A.prob + rho : nullptr;
    double s0[1], s1[1];
    const double* src0[1] = {A.rr_part + (P ? (long long)P->pb0 : (long long)rho * A.nB)};
    const double* src1[1] = {A.rp64 + (P ? (long long)P->pa0 : (long long)rho * A.nA64)};
    sum_partials<1>(src0, P ? P->nB : A.nB, s0, red);
    sum_partials<1>(src1, prev_updated ? (P ? P->nA : A.nA64) : 0, s1, red);   // true residual of the last update
    const double rr_now = s0[0];
    const double rr_last = prev_updated ? s1[0] : A.rr_prev[rho];
    const double since = (double)A.iter - A.it_prev[rho];
    const bool active = A.st[rho].act != 0.0;
    const double nb = A.normb[rho];
    const bool at_rtol = rr_now <= A.rtol2 * nb * nb;
    const bool doit = A.force || (active && (since >= A.ru_max || rr_now < A.delta2 * rr_last || at_rtol));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.flag_out[rho] = doit ? 1 : 0;
        A.rr_out[rho] = rr_last;
        A.it_out[rho] = doit ? (double)A.iter : A.it_prev[rho];
    }
    if (!doit) return;
    const long long n = P ? (long long)P->L * P->L * P->Lt : A.N;
    double2* xh = A.xhi + (P ? P->voff : rho * A.N);
    float2* xl = A.xlo + (P ? P->voff : rho * A.N);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float2 l = xl[i];
        double2 h = xh[i];
        h.x += (double)l.x;
        h.y += (double)l.y;
        xh[i] = h;
        xl[i] = make_float2(0.f, 0.f);
    }
}

// Reliable-update state at the start: |r_true|^2 = |b|^2 at iteration 0.
__global__ void ru_init_kernel(const double* normb, double* rr, double* it, int* flag, int nrhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrhs) {
        rr[i] = normb[i] * normb[i];
        it[i] = 0.0;
        flag[i] = 0;
    }
}

// x_hi += x_lo; x_lo = 0 (reliable update of the mixed-precision solver).
__global__ void flush_x_kernel(double2* __restrict__ xhi, float2* __restrict__ xlo, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float2 l = xlo[i];
        double2 h = xhi[i];
        h.x += (double)l.x;
        h.y += (double)l.y;
        xhi[i] = h;
        xlo[i] = make_float2(0.f, 0.f);
    }
}

__global__ void demote_kernel(const double2* __restrict__ src, float2* __restrict__ dst, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double2 v = src[i];
        dst[i] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ void zero_kernel(double* __restrict__ p, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace qf
