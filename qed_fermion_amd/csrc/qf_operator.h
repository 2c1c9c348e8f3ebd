// qf_operator.h -- tau-marching fermion operator kernels.
//
// M (get_M_sparse, hmc_sampler_batch.py:1884-1906):
//   (Mx)_tau = x_tau + sM(tau) B_{tau-1} x_{tau-1},  sM = +1 at tau = 0 (antiperiodic), else -1
// M^dag (B_tau Hermitian):
//   (M^dag y)_tau = y_tau + sD(tau) B_tau y_{tau+1},  sD = +1 at tau = Nt-1, else -1
// M^dag M fused: a work-group owns a run of T slices [t0, t1) of one RHS and
// marches in tau keeping x_tau and y_tau = (Mx)_tau in registers, so each
// output slice costs two B_tau applications that share one set of phases:
//   u = B_tau x_tau;  y_{tau+1} = x_{tau+1} + sM u;  w = B_tau y_{tau+1};
//   (M^dag M x)_tau = y_tau + sD w
// plus one B application of prologue per run.  HBM traffic per output slice:
// read x (+ one halo slice per run), read phi_tau, write y.
#pragma once
#include "qf_patch.h"
#include "qf_reduce.h"

namespace qf {

__device__ __forceinline__ int sgnM(int tau) { return tau == 0 ? 1 : -1; }
__device__ __forceinline__ int sgnD(int tau, int Lt) { return tau == Lt - 1 ? 1 : -1; }

template <typename R, int BY>
__device__ __forceinline__ void axpy_patch(cplx<R> (&out)[2 * BY], const cplx<R> (&a)[2 * BY], R s, const cplx<R> (&u)[2 * BY]) {
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) out[k] = mk<R>(a[k].x + s * u[k].x, a[k].y + s * u[k].y);
}

template <typename R, int BY>
__device__ __forceinline__ double redot_patch(const cplx<R> (&a)[2 * BY], const cplx<R> (&b)[2 * BY]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) s += (double)a[k].x * (double)b[k].x + (double)a[k].y * (double)b[k].y;
    return s;
}

// Arguments of the M^dag M family.  Pointers are per-launch bases; per-RHS
// offsets are rho * vs (vectors) and rho * phi_cs (phi; 0 = shared phi).
template <typename R>
struct MdmArgs {
    const R* phi;
    long long phi_cs;
    const cplx<R>* hop;     // optional hop table (replaces phi): [chain][tau][y][x][mu]
    long long hop_cs;       // complex per chain (0 = shared)
    const cplx<R>* x;
    cplx<R>* y;
    long long vs;
    int L, Lt, T;
    Hop<R> hp;
    double* dot_part;       // [nrhs][nblk]: ||(Mx)_tau||^2 = Re(x^dag M^dag M x) over the run (optional)
    // MODE 1 (CG step A): x_tau := z_tau + beta p_tau, written to pnew
    const cplx<R>* z;
    const cplx<R>* pold;
    cplx<R>* pnew;
    const double* rz_part;  // partials of the previous update step, [nrhs][nB]
    const double* rr_part;
    int nB;
    const CgState* st_prev;
    CgState* st_out;
    double* normb;
    int iter;
    double rtol;
    int may_stop;           // 0: the partials are an iterated (not true) residual; do not deactivate
    const int* stop_flag;   // if set: per-chain may_stop (1 right after a reliable update)
    const int* ru_flag;     // MODE 2: skip chains whose reliable-update flag is 0
    // MODE 2 (true residual in FP64): r = b - M^dag M x written as complex64
    const cplx<R>* b;
    float2* r_lo;
};

enum { MDM_PLAIN = 0, MDM_CG = 1, MDM_RESID = 2 };

template <typename R, int BY, int MODE, int NT>
__global__ void __launch_bounds__(NT) mdagm_kernel(MdmArgs<R> A) {
    extern __shared__ __align__(16) unsigned char qf_smem[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(qf_smem);
    __shared__ double red[64];
    const int rho = blockIdx.y;
    const int L = A.L, Lt = A.Lt, Vs = L * L;
    const int t0 = blockIdx.x * A.T;
    const int t1 = min(t0 + A.T, Lt);
    const Geo g = make_geo<BY>(L, threadIdx.x);
    const R* phi = A.phi + rho * A.phi_cs;
    const long long vo = rho * A.vs;
    R beta = R(0);
    if constexpr (MODE == MDM_RESID) {
        if (A.ru_flag != nullptr && A.ru_flag[rho] == 0) return;   // uniform over the block
    }

    if constexpr (MODE == MDM_CG) {
        // Scalars of the previous iteration, reduced in a fixed order by
        // every block (hmc_sampler_batch.py:806-844: error, active mask,
        // beta = rz_new / rz_old).
        double sums[2];
        const double* src[2] = {A.rz_part + (long long)rho * A.nB, A.rr_part + (long long)rho * A.nB};
        sum_partials<2>(src, A.nB, sums, red);
        const CgState S = A.st_prev[rho];
        const double nb = (A.iter == 0) ? sqrt(sums[1]) : A.normb[rho];
        double act, err, rzv, it = S.iters, bt = 0.0;
        if (S.act == 0.0) {
            act = 0.0; err = S.err; rzv = S.rz;
        } else {
            err = nb > 0.0 ? sqrt(sums[1]) / nb : 0.0;
            const int may = A.stop_flag ? A.stop_flag[rho] : A.may_stop;
            act = (err > A.rtol || !may) ? 1.0 : 0.0;
            rzv = sums[0];
            if (A.iter > 0) {
                bt = sums[0] / S.rz;
                if (act != 0.0) it += 1.0;
            }
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            CgState o; o.rz = rzv; o.act = act; o.err = err; o.iters = it;
            A.st_out[rho] = o;
            if (A.iter == 0) A.normb[rho] = nb;
        }
        if (act == 0.0) return;   // uniform over the block
        beta = (R)bt;
    }

    const cplx<R>* x = A.x + vo;
    const cplx<R>* hop = A.hop ? A.hop + rho * A.hop_cs : nullptr;
    auto coef = [&](Coef<R, BY>& cf, int tau) {
        if (hop)
            load_coef_tab<R, BY>(cf, hop + (long long)tau * 2 * Vs, g);
        else
            load_coef<R, BY>(cf, phi + (long long)tau * 2 * Vs, g, A.hp);
    };
    auto ldx = [&](int tau, cplx<R> (&v)[2 * BY]) {
        if constexpr (MODE == MDM_CG) {
            const long long so = vo + (long long)tau * Vs;
            cplx<R> zz[2 * BY];
            load_patch<R, BY>(zz, A.z + so, g);
            if (A.iter == 0) {
#pragma unroll
                for (int k = 0; k < 2 * BY; ++k) v[k] = zz[k];
            } else {
                cplx<R> pp[2 * BY];
                load_patch<R, BY>(pp, A.pold + so, g);
                axpy_patch<R, BY>(v, zz, beta, pp);
            }
        } else {
            load_patch<R, BY>(v, x + (long long)tau * Vs, g);
        }
    };

    // tau pipeline: the second B of slice tau (W = B_tau y_{tau+1}) and the
    // first B of slice tau+1 (U = B_{tau+1} x_{tau+1}) are independent, so they
    // run together and share every exchange barrier (7 per slice, not 14).
    // The dot product accumulated is ||y||^2 = Re(x^dag M^dag M x) summed over
    // the run's slices (y = Mx is in registers; no x copy needed).
    int phase = 0;
    Coef<R, BY> cfc, cfn;
    cplx<R> X[2 * BY], Y[2 * BY], Yn[2 * BY], U[2 * BY], W[2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) X[k] = Y[k] = Yn[k] = U[k] = W[k] = mk<R>(0, 0);
    double dsum = 0.0;

    // prologue: U = B_{t0-1} x_{t0-1}, W = B_{t0} x_{t0};  y_{t0}, y_{t0+1}
    const int tm = (t0 + Lt - 1) % Lt;
    const int t0n = (t0 + 1 == Lt) ? 0 : t0 + 1;
    coef(cfn, tm);
    coef(cfc, t0);
    ldx(tm, U);
    ldx(t0, X);
    if constexpr (MODE == MDM_CG) {
        if (g.valid) store_patch<R, BY>(A.pnew + vo + (long long)t0 * Vs, X, g);
    }
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) W[k] = X[k];
    apply_B2<R, BY>(U, cfn, W, cfc, A.hp, lds, g, phase);
    const R cn = A.hp.cn;   // apply_B returns B v / cn
    axpy_patch<R, BY>(Y, X, sgnM(t0) * cn, U);
    ldx(t0n, X);
    axpy_patch<R, BY>(Yn, X, sgnM(t0n) * cn, W);

    for (int tau = t0; tau < t1; ++tau) {
        const int tn = (tau + 1 == Lt) ? 0 : tau + 1;
        const bool more = tau + 1 < t1;
#pragma unroll
        for (int k = 0; k < 2 * BY; ++k) W[k] = Yn[k];
        if (more) {
            coef(cfn, tn);
#pragma unroll
            for (int k = 0; k < 2 * BY; ++k) U[k] = X[k];
            apply_B2<R, BY>(W, cfc, U, cfn, A.hp, lds, g, phase);   // W = B_tau y_{tau+1}, U = B_{tau+1} x_{tau+1}
        } else {
            apply_B<R, BY>(W, cfc, A.hp, lds, g, phase);
        }
        cplx<R> Z[2 * BY];
        axpy_patch<R, BY>(Z, Y, sgnD(tau, Lt) * cn, W);
        if (g.valid) {
            const long long so = vo + (long long)tau * Vs;
            if constexpr (MODE == MDM_RESID) {
                cplx<R> bb[2 * BY];
                load_patch<R, BY>(bb, A.b + so, g);
#pragma unroll
                for (int k = 0; k < 2 * BY; ++k) {
                    const R rx = bb[k].x - Z[k].x, ry = bb[k].y - Z[k].y;
                    dsum += (double)rx * rx + (double)ry * ry;
                    Z[k] = mk<R>(rx, ry);
                }
                float2* rp = A.r_lo + so + (g.y0) * L + 2 * g.a;
#pragma unroll
                for (int r = 0; r < BY; ++r) {
                    rp[r * L] = make_float2((float)Z[2 * r].x, (float)Z[2 * r].y);
                    rp[r * L + 1] = make_float2((float)Z[2 * r + 1].x, (float)Z[2 * r + 1].y);
                }
            } else {
                store_patch<R, BY>(A.y + so, Z, g);
                dsum += redot_patch<R, BY>(Y, Y);
            }
        }
        if (more) {
            const int tnn = (tn + 1 == Lt) ? 0 : tn + 1;
#pragma unroll
            for (int k = 0; k < 2 * BY; ++k) Y[k] = Yn[k];
            if constexpr (MODE == MDM_CG) {
                if (g.valid) store_patch<R, BY>(A.pnew + vo + (long long)tn * Vs, X, g);
            }
            ldx(tnn, X);
            axpy_patch<R, BY>(Yn, X, sgnM(tnn) * cn, U);        // y_{tau+2}
            cfc = cfn;
        }
    }

    if (A.dot_part != nullptr) {
        double v[1] = {dsum};
        block_sum<1>(v, red);
        if (threadIdx.x == 0) A.dot_part[(long long)rho * gridDim.x + blockIdx.x] = v[0];
    }
}

// Single-sided operators: OP_M (y = Mx), OP_MDAG (y = M^dag x), OP_B
// (y_tau = B_tau x_tau for every slice; the per-slice product of
// _C.b_vec_per_tau, hmc_sampler_batch.py:2181).
enum { OP_M = 1, OP_MDAG = 2, OP_B = 3 };

template <typename R>
struct OpArgs {
    const R* phi;
    long long phi_cs;
    const cplx<R>* hop;
    long long hop_cs;
    const cplx<R>* x;
    cplx<R>* y;
    long long vs;
    int L, Lt, T;
    Hop<R> hp;
    int nfac;   // OP_B: number of leading factors of B_tau applied (7 = all)
    R cpart;    // OP_B: product of cosh(a_g) over those factors (cn for all 7)
};

template <typename R, int BY, int OP, int NT>
__global__ void __launch_bounds__(NT) op_kernel(OpArgs<R> A) {
    extern __shared__ __align__(16) unsigned char qf_smem[];
    cplx<R>* lds = reinterpret_cast<cplx<R>*>(qf_smem);
    const int rho = blockIdx.y;
    const int L = A.L, Lt = A.Lt, Vs = L * L;
    const int t0 = blockIdx.x * A.T;
    const int t1 = min(t0 + A.T, Lt);
    const Geo g = make_geo<BY>(L, threadIdx.x);
    const R* phi = A.phi + rho * A.phi_cs;
    const cplx<R>* x = A.x + rho * A.vs;
    cplx<R>* y = A.y + rho * A.vs;
    const cplx<R>* hop = A.hop ? A.hop + rho * A.hop_cs : nullptr;
    auto coef = [&](Coef<R, BY>& c, int tau) {
        if (hop)
            load_coef_tab<R, BY>(c, hop + (long long)tau * 2 * Vs, g);
        else
            load_coef<R, BY>(c, phi + (long long)tau * 2 * Vs, g, A.hp);
    };
    int phase = 0;
    Coef<R, BY> cf;
    cplx<R> X[2 * BY], U[2 * BY];
#pragma unroll
    for (int k = 0; k < 2 * BY; ++k) X[k] = U[k] = mk<R>(0, 0);

    if constexpr (OP == OP_M) {
        int tp = (t0 + Lt - 1) % Lt;
        // every lane loads (padding lanes mirror patch (0,0)); only valid lanes store
        load_patch<R, BY>(U, x + (long long)tp * Vs, g);
        for (int tau = t0; tau < t1; ++tau) {
            coef(cf, tp);
            apply_B<R, BY>(U, cf, A.hp, lds, g, phase);        // U = B_{tau-1} x_{tau-1}
            load_patch<R, BY>(X, x + (long long)tau * Vs, g);
            if (g.valid) {
                cplx<R> Z[2 * BY];
                axpy_patch<R, BY>(Z, X, sgnM(tau) * A.hp.cn, U);
                store_patch<R, BY>(y + (long long)tau * Vs, Z, g);
            }
#pragma unroll
            for (int k = 0; k < 2 * BY; ++k) U[k] = X[k];
            tp = tau;
        }
    } else if constexpr (OP == OP_MDAG) {
        load_patch<R, BY>(X, x + (long long)t0 * Vs, g);
        for (int tau = t0; tau < t1; ++tau) {
            const int tn = (tau + 1 == Lt) ? 0 : tau + 1;
            coef(cf, tau);
            load_patch<R, BY>(U, x + (long long)tn * Vs, g);
            cplx<R> Xn[2 * BY];
#pragma unroll
            for (int k = 0; k < 2 * BY; ++k) Xn[k] = U[k];
            apply_B<R, BY>(U, cf, A.hp, lds, g, phase);        // U = B_tau x_{tau+1}
            if (g.valid) {
                cplx<R> Z[2 * BY];
                axpy_patch<R, BY>(Z, X, sgnD(tau, Lt) * A.hp.cn, U);
                store_patch<R, BY>(y + (long long)tau * Vs, Z, g);
            }
#pragma unroll
            for (int k = 0; k < 2 * BY; ++k) X[k] = Xn[k];
        }
    } else {
        for (int tau = t0; tau < t1; ++tau) {
            coef(cf, tau);
            load_patch<R, BY>(U, x + (long long)tau * Vs, g);
            if (A.nfac == 7)
                apply_B<R, BY>(U, cf, A.hp, lds, g, phase);
            else
                apply_B_partial<R, BY>(U, A.nfac, cf, A.hp, lds, g, phase);
#pragma unroll
            for (int k = 0; k < 2 * BY; ++k) U[k] = mk<R>(A.cpart * U[k].x, A.cpart * U[k].y);
            if (g.valid) store_patch<R, BY>(y + (long long)tau * Vs, U, g);
        }
    }
}

}  // namespace qf
