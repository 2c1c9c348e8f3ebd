// qf_persist.h -- the batched PCG of pcg_iterate as ONE persistent launch.
//
// Same recurrence and kernels' bodies as the two-launches-per-iteration solve
// (preconditioned_cg_fast_test, hmc_sampler_batch.py:751-846): step A
// (p = z + beta p, Op = M^dag M p, Re(p^dag Op) partials; mdagm_body MDM_CG)
// and step B (alpha, r, x, z = P r, partials; cgu_body CGU_ITER), each over
// all right-hand sides, with a grid barrier between them.  The work-groups of
// the launch (all resident: cooperative launch) take the work items of a phase
// in turn.  The solve leaves the iteration loop as soon as no chain was active
// in an iteration's step A (the reference's fixed max_iter stays the upper
// bound; frozen chains do nothing either way, so the outputs equal those of
// max_iter iterations), then writes err / iters / status as cg_finalize_kernel.
//
// Grid barrier: agent-scope release, one atomic arrival per work-group on a
// counter that only grows within the launch, a bounded spin, agent-scope
// acquire.  A spin that exceeds the time limit raises an abort flag; every
// work-group then leaves, and the last one out marks every chain
// kSolveAborted.  The last work-group out also zeroes the counters, so the
// next launch on the workspace starts from zero (the caller zeroes them once,
// before the first use).
#pragma once
#include "qf_cg.h"
#include "qf_operator.h"

namespace qf {

constexpr int kSolveAborted = 4;   // a grid barrier of the persistent solve timed out

struct GridBar {
    unsigned* count;   // arrivals of this launch
    unsigned* done;    // work-groups that have left
    unsigned* abort;   // nonzero after a timed-out wait
};

__device__ __forceinline__ unsigned ld_agent(const unsigned* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_agent(unsigned* p, unsigned v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Every work-group of the launch must call it the same number of times.  Returns false
// (in every thread) if the wait timed out here or elsewhere.
__device__ __forceinline__ bool grid_barrier(const GridBar& gb, unsigned& epoch, long long timeout_ticks) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned target = (epoch + 1) * gridDim.x;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(gb.count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        const long long t0 = wall_clock64();
        while (ld_agent(gb.count) < target) {
            if (ld_agent(gb.abort) != 0) { ok = 0; break; }
            if (wall_clock64() - t0 > timeout_ticks) { st_agent(gb.abort, 1u); ok = 0; break; }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    ++epoch;
    return s_ok != 0;
}

template <typename R>
struct PcgPersArgs {
    MdmArgs<R> m;        // step A (the per-iteration fields are set by the kernel)
    CguArgs<R> u;        // step B / start (likewise)
    int nrhs, nA;        // right-hand sides; operator runs per rhs
    int nsx, ntile;      // update tiles per rhs: site columns x slice tiles
    int max_iter, paired;
    const cplx<R>* b;
    cplx<R>* x;
    cplx<R>* R2[2];
    cplx<R>* P2[2];
    cplx<R>* Z;
    cplx<R>* OP;
    double* rzp;         // partial sums, two parity slots of ps each
    double* rrp;
    double* pap;
    long long ps;
    CgState* st;         // two parity slots of nrhs
    double* normb;
    double* xal;
    int* brk;
    // hop table built in the first phase (HOPT)
    const R* phi;
    long long phi_cs;
    cplx<R>* hop;
    int nchain;
    double th1, th2;     // tanh(t dtau), tanh(t dtau / 2)
    double rtol;
    double* err_out;
    double* iters_out;
    int* status_out;
    GridBar gb;
    long long timeout_ticks;
    // ragged batch (prob != nullptr): independent lattices, one chain each, with the
    // operator and update work items of the plan (nrhs = lattices)
    const Prob* prob;
    int nAi, nBi;
    long long nsites;
};

// Update tiles of the persistent solve: 16 sites x P.u.TB slices, four per thread (uniform
// lattices: TB = blockDim / 4; a ragged plan's TB = 64 with 256 working threads).
constexpr int kPersTX = 16, kPersSPT = 4;

template <typename R, int BY, bool HOPT, bool PIPE>
__global__ void __launch_bounds__(512) pcg_persistent_kernel(PcgPersArgs<R> P) {
    // the update phase stages its tile in the dynamic LDS that the operator phase uses for its
    // exchanges (the launcher sizes it for both)
    extern __shared__ __align__(16) unsigned char qf_smem[];
    const int G = gridDim.x;
    const int nrhs = P.nrhs;
    const bool rag = P.prob != nullptr;
    const int nBt = P.nsx * P.ntile;
    const int nAi = rag ? P.nAi : nrhs * P.nA, nBi = rag ? P.nBi : nrhs * nBt;
    unsigned epoch = 0;
    bool ok = true;
    int last = 1;   // parity slot of the last completed iteration (1: the start)

    if constexpr (HOPT) {
        const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)G * blockDim.x;
        if (rag)
            hop_table_ragged_body<R>(P.phi, P.hop, P.prob, nrhs, P.nsites, P.th1, P.th2, i0, stride);
        else
            hop_table_body<R>(P.phi, P.phi_cs, P.hop, P.nchain, P.m.L, P.m.Lt, P.th1, P.th2, i0, stride);
    }
    {   // start: x = 0, r = b, z = P b, partials into slot 1, state slot 1
        CguArgs<R> u = P.u;
        u.rold = P.b; u.x = P.x; u.rnew = P.R2[1]; u.z = P.Z;
        u.st_init = P.st + nrhs; u.rz_part = P.rzp + P.ps; u.rr_part = P.rrp + P.ps; u.brk = P.brk;
        for (int it = blockIdx.x; it < nBi; it += G) {
            const int rho = rag ? 0 : it / nBt, k = rag ? 0 : it - rho * nBt;   // a ragged tile is found from it
            cgu_body<R, CGU_INIT, true, kPersTX, 0, kPersSPT>(u, it, rag ? 0 : k % P.nsx, rag ? 0 : k / P.nsx, rho, P.nsx, P.ntile, qf_smem);
            __syncthreads();
        }
    }
    ok = grid_barrier(P.gb, epoch, P.timeout_ticks);
    bool held = false;
    for (int i = 0; ok && i < P.max_iter; ++i) {
        const int cur = i & 1, prv = cur ^ 1;
        const int xmode = !P.paired ? 0 : held ? 2 : (i + 1 == P.max_iter ? 0 : 1);
        held = xmode == 1;
        MdmArgs<R> m = P.m;
        m.iter = i; m.may_stop = 1; m.stop_flag = nullptr; m.brk = P.brk;
        m.rz_part = P.rzp + prv * P.ps; m.rr_part = P.rrp + prv * P.ps;
        m.st_prev = P.st + prv * nrhs; m.st_out = P.st + cur * nrhs;
        m.pold = P.P2[prv]; m.pnew = P.P2[cur];
        for (int it = blockIdx.x; it < nAi; it += G) {
            mdagm_body<R, BY, MDM_CG, HOPT, PIPE>(m, it, nAi);
            __syncthreads();
        }
        if (!(ok = grid_barrier(P.gb, epoch, P.timeout_ticks))) break;
        CguArgs<R> u = P.u;
        u.rold = P.R2[prv]; u.op = P.OP; u.p = P.P2[cur]; u.x = P.x; u.rnew = P.R2[cur]; u.z = P.Z;
        u.pap = P.pap; u.st = P.st + cur * nrhs; u.rz_part = P.rzp + cur * P.ps; u.rr_part = P.rrp + cur * P.ps;
        u.ru_flag = nullptr; u.xmode = xmode; u.pprev = P.P2[prv]; u.xalpha = P.xal; u.st_held = P.st + prv * nrhs; u.brk = P.brk;
        for (int it = blockIdx.x; it < nBi; it += G) {
            const int rho = rag ? 0 : it / nBt, k = rag ? 0 : it - rho * nBt;   // a ragged tile is found from it
            cgu_body<R, CGU_ITER, true, kPersTX, 0, kPersSPT>(u, it, rag ? 0 : k % P.nsx, rag ? 0 : k / P.nsx, rho, P.nsx, P.ntile, qf_smem);
            __syncthreads();
        }
        if (!(ok = grid_barrier(P.gb, epoch, P.timeout_ticks))) break;
        last = cur;
        // leave once no chain was active in this iteration's step A: step B then only applied a
        // held step of the previous iteration, and every later iteration would do nothing
        int any = 0;
        for (int r = threadIdx.x; r < nrhs; r += blockDim.x) any |= P.st[cur * nrhs + r].act != 0.0;
        if (!__syncthreads_or(any)) break;
    }
    if (ok) {
        for (int rho = blockIdx.x; rho < nrhs; rho += G) {
            cg_finalize_body(rho, P.rrp + last * P.ps, nBt, P.st + last * nrhs, P.normb, P.rtol, 0, P.err_out, P.iters_out, P.status_out,
                             P.prob);
            __syncthreads();
        }
    }
    // leave; the last work-group out resets the barrier for the next launch, after
    // marking every chain if a wait timed out
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        s_last = __hip_atomic_fetch_add(P.gb.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == (unsigned)G;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (s_last) {
        const bool aborted = ld_agent(P.gb.abort) != 0;
        if (aborted) {
            for (int r = threadIdx.x; r < nrhs; r += blockDim.x) {
                if (P.status_out != nullptr) P.status_out[r] = kSolveAborted;
                P.err_out[r] = __builtin_nan("");
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            st_agent(P.gb.count, 0u);
            st_agent(P.gb.abort, 0u);
            st_agent(P.gb.done, 0u);
        }
    }
}

}  // namespace qf
