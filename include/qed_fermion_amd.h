/*
 * qed_fermion_amd.h -- C ABI of libqf_hip.so, the MI355X (gfx950) HMC fermion
 * hot path of qed_fermion.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers owned by the caller (the Python
 *    layer passes torch tensors' data_ptr()); the library allocates nothing and
 *    never synchronises the host, so every call can be captured into a hipGraph.
 *  - stream is a hipStream_t passed as void* (NULL = default stream).
 *  - prec: QF_PREC_F32 = complex64 vectors + float32 phi (the reference's
 *    dtypes, hmc_sampler_batch.py:118-123); QF_PREC_F64 = complex128 + float64.
 *  - Layouts are the reference's permuted views (hmc_sampler_batch.py:2166):
 *      phi  [chain][Ltau][L][L][2]  (tau, y, x, mu), chain stride phi_cs reals
 *           (0 = one phi shared by all right-hand sides)
 *      vec  [rhs][Ltau][L][L]       complex, rhs stride Ltau*L*L
 *      Ft   same layout as phi
 *  - Return value: QF_OK or an error code; qf_last_error() has the message
 *    (thread-local).  No exceptions cross the ABI.
 *
 * File:line citations are relative to the reference repository
 * (KexinFeng/qed_fermion).
 */
#ifndef QED_FERMION_AMD_H
#define QED_FERMION_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QF_ABI_VERSION 2

enum {
    QF_OK = 0,
    QF_ERR_INVALID = 1,     /* bad argument, size or null pointer */
    QF_ERR_ODD_L = 2,       /* L odd: checkerboard groups (hmc_sampler_batch.py:937-969) need even L */
    QF_ERR_LTAU = 3,        /* manual preconditioner needs Ltau >= 26 (precon_manual.py:50-167) */
    QF_ERR_HIP = 4,         /* HIP runtime / launch failure */
    QF_ERR_UNSUPPORTED = 5, /* lattice shape with no valid work-group tiling */
    QF_ERR_PRECON = 6       /* CSR preconditioner is not P_tau (x) I_Vs with a 13-band tau stencil */
};

enum { QF_PREC_F32 = 0, QF_PREC_F64 = 1 };

/* Per-chain outcome of a PCG solve, written on the device to the solvers'
 * status_out[nrhs] (int).  A chain stops at the first failure with x left at
 * its last finite iterate.  The reference's preconditioned_cg_fast_test
 * (hmc_sampler_batch.py:751-846) has no such status: a NaN or a breakdown
 * there propagates into x, the force and H. */
enum {
    QF_SOLVE_CONVERGED = 0, /* relative residual <= rtol */
    QF_SOLVE_MAX_ITER = 1,  /* max_iter iterations without reaching rtol (the reference's fixed budget) */
    QF_SOLVE_NONFINITE = 2, /* NaN or Inf in the residual, Re(r^dag z) or Re(p^dag A p): a NaN in phi or b */
    QF_SOLVE_BREAKDOWN = 3, /* Re(p^dag A p) <= 0: the operator is not positive definite on p */
    QF_SOLVE_ABORTED = 4    /* qf_pcg_solve_persistent only: a grid barrier timed out (err_out is NaN) */
};

enum { QF_OP_M = 1, QF_OP_MDAG = 2, QF_OP_B = 3, QF_OP_MDAGM = 4, QF_OP_BPART0 = 16 };
/* QF_OP_BPART0 + m (m = 0..7): y_tau = Q_m x_tau, the first m factors of the
 * B_tau sequence D4, D3, D2, D1, D2, D3, D4 -- the intermediates that
 * _C.b_vec_per_tau(..., interm=True) returns (hmc_sampler_batch.py:2183-2189). */

int qf_version(void);
const char* qf_last_error(void);

/* Work-group tiling: patch height by in {2,4,6} (0 = auto), tau slices per
 * operator work-group t (0 = auto), tau slices of a PCG update tile tb in
 * {16, 32, 64, 128} (0 = auto; 512-thread tiles of 32 sites for tb = 16, 16
 * sites otherwise).  Set by the autotuner (qed_fermion_amd/autotune.py); no
 * counterpart in the reference, whose kernels take BLOCK_SIZE=(4,8)
 * (hmc_sampler_batch.py:111). */
int qf_set_tuning(int prec, int by, int t, int tb);
/* The same for one lattice shape (nrhs = 0: any batch size); looked up
 * before the per-precision setting.  Written by the autotuner. */
int qf_set_tuning_shape(int prec, int L, int Ltau, int nrhs, int by, int t, int tb);
/* Forget every tuning entry (back to the built-in defaults). */
int qf_clear_tuning(void);
/* out4 = {patch height, tau slices per operator work-group, tau slices per
 * update work-group, operator work-groups per right-hand side}. */
int qf_get_tiling(int prec, int nrhs, int L, int Ltau, int* out4);
/* Right-hand sides per work-group of the plain M^dag M from phi (qf_apply
 * QF_OP_MDAGM) for one lattice shape (nrhs = 0: any batch size): ns > 1 runs
 * the same tau run of ns consecutive right-hand sides side by side in one
 * work-group, to fill whole wavefronts where one slice leaves a wavefront mostly
 * empty (L=36: 324 lanes per slice, 972 for three); 0 or 1 = one.  Whole-slice
 * tilings only; capped by the 1024-thread work-group (512 above 2x2 patches).
 * Set by the autotuner; no counterpart in the reference. */
int qf_set_tuning_ns(int prec, int L, int Ltau, int nrhs, int ns);
/* The right-hand sides per work-group qf_apply's M^dag M from phi would use. */
int qf_get_rhs_per_group(int prec, int nrhs, int L, int Ltau, int* ns);
/* PCG x updates in pairs (on = 1, the default): an iteration that is not
 * followed by a read of x holds x += alpha p and the next iteration applies
 * both steps, so x and p are read every other iteration.  on = 0 updates x in
 * every iteration as hmc_sampler_batch.py:802 does; the results are the same.
 * Returns the previous setting.  No counterpart in the reference. */
int qf_set_paired_x_updates(int on);
/* y = op(phi) x for nrhs right-hand sides.  op = QF_OP_M: get_M_sparse's M
 * (hmc_sampler_batch.py:1884-1906); QF_OP_MDAG: M^dag (_C.mh_vec, :2406);
 * QF_OP_MDAGM: M^dag M (_C.mhm_vec, :772, :799); QF_OP_B: y_tau = B_tau x_tau
 * per slice (_C.b_vec_per_tau with interm=False, :2181). */
int qf_apply(int op, int prec, const void* phi, long long phi_cs, const void* x, void* y, int nrhs, int L, int Ltau, double dtau,
             double t, void* stream);

/* Hop table of nchain fields: hop[chain][tau][y][x][mu] = tanh(a_g) e^{i phi_mu} (complex,
 * 2*Ltau*L*L per chain), the per-link factor coefficients of get_M's dB matrices
 * (hmc_sampler_batch.py:1560-1598) normalised by cosh(a_g).  qf_apply_hop applies the
 * operators from a table instead of phi (no sincos per application; hop_cs = 0 shares
 * one table between all right-hand sides).  No separate counterpart in the reference,
 * whose kernels take the boson field directly. */
int qf_hop_table(int prec, const void* phi, long long phi_cs, int nchain, int L, int Ltau, double dtau, double t, void* hop, void* stream);
int qf_apply_hop(int op, int prec, const void* hop, long long hop_cs, const void* x, void* y, int nrhs, int L, int Ltau, double dtau, double t,
                 void* stream);

/* _C.mhm_vec(boson, x, Lx, dtau, *BLOCK_SIZE) (hmc_sampler_batch.py:772, 799): batched y = M^dag M x, t = 1. */
int qf_mhm_vec(int prec, const void* phi, const void* x, void* y, int bs, int L, int Ltau, double dtau, void* stream);

/* _C.mh_vec(boson, R, Lx, dtau, *BLOCK_SIZE) (hmc_sampler_batch.py:2406; stochastic_estimator.py:388, 430): y = M^dag x. */
int qf_mh_vec(int prec, const void* phi, const void* x, void* y, int bs, int L, int Ltau, double dtau, void* stream);

/* _C.precon_vec(r, precon_csr, Lx) (hmc_sampler_batch.py:773, 838): z = P r with P given as its
 * cyclic tau-band table ptab[Ltau][13] (float for F32, double for F64) instead of a CSR. */
int qf_precon_vec(int prec, const void* ptab, const void* r, void* z, int nrhs, int Vs, int Ltau, void* stream);

/* HOST functions.  get_precon_man(Lx, Ltau) (precon_manual.py:15-175) as a [Ltau][13] table:
 * table[tau*13 + o + 6] multiplies r_{(tau+o) mod Ltau}. */
int qf_precon_table(int Ltau, double* table);
/* Verify that a CSR matrix (reset_precon's precon_csr, hmc_sampler_batch.py:547) is
 * P_tau (x) I_Vs with a 13-point cyclic tau band and extract its table.  val_im may be NULL. */
int qf_precon_table_from_csr(const long long* crow, const long long* col, const double* val_re, const double* val_im, long long nrows,
                             int Vs, int Ltau, double* table);

/* Batched PCG for M^dag M x = b, the recurrence of preconditioned_cg_fast_test
 * (hmc_sampler_batch.py:751-846): x0 = 0, fixed max_iter with the per-chain
 * active mask (converged chains freeze; their kernels exit early).  ptab NULL
 * = no preconditioner.  Outputs on the device: err_out[nrhs] (last relative
 * residual), iters_out[nrhs] (the reference's iteration counter) and, when
 * status_out is not NULL, status_out[nrhs] (a QF_SOLVE_* code per chain).
 * Two kernels per iteration; all scalars stay on the device.  work: at least
 * qf_pcg_workspace_bytes(prec, 0, nrhs, L, Ltau) bytes of device memory. */
size_t qf_pcg_workspace_bytes(int prec, int mixed, int nrhs, int L, int Ltau);
int qf_pcg_solve(int prec, const void* phi, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau, double dtau, double t,
                 const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes, double* err_out, double* iters_out,
                 int* status_out, void* stream);

/* qf_pcg_solve with its workspace in the library's own static device memory instead of a
 * caller buffer: no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * the call itself allocates nothing.  The scratch is qf_pcg_noalloc_max_bytes() = 64 MiB per
 * GPU, held by every GPU the library runs on until the process ends; a solve whose
 * qf_pcg_workspace_bytes(prec, 0, nrhs, L, Ltau) exceeds it is rejected with QF_ERR_INVALID
 * (one complex128 chain at L=36, Ltau=360 fits).  All calls on one GPU share that GPU's
 * scratch: calls that overlap on one GPU (other streams, other threads) race. */
size_t qf_pcg_noalloc_max_bytes(void);
int qf_pcg_solve_noalloc(int prec, const void* phi, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau, double dtau,
                         double t, const void* ptab, double rtol, int max_iter, double* err_out, double* iters_out, int* status_out,
                         void* stream);

/* qf_pcg_solve as ONE persistent cooperative launch: the same recurrence and
 * the same per-iteration work (the operator and update kernels' bodies), with
 * the work-groups taking each phase's items in turn and a grid barrier
 * between the phases, so the whole solve is one dispatch for any number of
 * right-hand sides (the multi-RHS solve of stochastic_estimator.py:427-433).
 * The launch leaves as soon as no chain is active (the reference's fixed
 * max_iter is the upper bound; frozen chains do nothing, so err_out,
 * iters_out, status_out and x equal those of qf_pcg_solve up to the order of
 * the partial sums).  ptab must not be NULL.  Every barrier wait is bounded
 * (10 s): a timed-out solve marks every chain QF_SOLVE_ABORTED.  work: at
 * least n = qf_pcg_persistent_workspace_bytes(prec, nrhs, L, Ltau) bytes; the
 * 256 bytes at offset n - 256 (barrier counters) must be zero before the
 * first call on that workspace (every call leaves them zero).  Calls sharing a workspace
 * must not overlap.  L up to 64 (work-groups of at most 512 threads); returns
 * QF_ERR_UNSUPPORTED otherwise.  Not for stream capture. */
size_t qf_pcg_persistent_workspace_bytes(int prec, int nrhs, int L, int Ltau);
/* Time limit of each grid-barrier wait of qf_pcg_solve_persistent, in microseconds
 * (default 10 s); returns the previous value.  A limit of 0 makes every wait time out at
 * once, which exercises the abort path (every work-group leaves, every chain is marked
 * QF_SOLVE_ABORTED) without a fault.  No counterpart in the reference. */
long long qf_set_persistent_timeout_us(long long us);
int qf_pcg_solve_persistent(int prec, const void* phi, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau, double dtau,
                            double t, const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes, double* err_out,
                            double* iters_out, int* status_out, void* stream);

/* Mixed-precision PCG (same recurrence and outputs): complex64 vectors and
 * operator with FP64 dot products, and every ru_period iterations a reliable
 * update x_hi += x_lo; r = b - (M^dag M)_FP64 x_hi.  b and x are complex128;
 * phi64 is the float64 master field and phi32 its float32 copy.  err_out is the
 * FP64 true residual of the returned x.  Replaces the complex64 CG of
 * preconditioned_cg_fast_test (hmc_sampler_batch.py:751-846). */
int qf_pcg_mixed_solve(const double* phi64, const float* phi32, long long phi_cs, const void* b, void* x, int nrhs, int L, int Ltau,
                       double dtau, double t, const float* ptab32, double rtol, int max_iter, int ru_period, void* work,
                       size_t work_bytes, double* err_out, double* iters_out, int* status_out, void* stream);

/* Ragged batch: nprob independent lattices of different L (one chain each)
 * in the same launches (for example L in {12, 18, 24, 30} with Ltau = 10 L);
 * no counterpart in the reference, whose batches share one L.  Data are concatenated per lattice
 * (vectors: Ltau*L*L complex each; phi: 2*Ltau*L*L reals; preconditioner
 * tables: 13*Ltau each).  qf_ragged_plan fills a host plan of
 * qf_ragged_plan_bytes(nprob) bytes; the caller keeps an identical copy in
 * device memory and passes both (the host copy sizes the launches, the device
 * copy is read by the kernels).  Every lattice runs with 2 x 2 patches. */
size_t qf_ragged_plan_bytes(int nprob);
int qf_ragged_plan(int nprob, const int* L, const int* Ltau, void* plan, size_t plan_bytes);
/* y = M^dag M x for every lattice of the batch in one launch (op must be QF_OP_MDAGM). */
int qf_apply_ragged(int op, int prec, const void* plan_host, const void* plan_dev, const void* phi, const void* x, void* y, double dtau,
                    double t, void* stream);
/* Batched PCG over a ragged batch (the recurrence of qf_pcg_solve; err_out and
 * iters_out per lattice); every iteration is one operator and one update launch
 * for the whole batch. */
size_t qf_pcg_ragged_workspace_bytes(int prec, const void* plan_host);
int qf_pcg_solve_ragged(int prec, const void* plan_host, const void* plan_dev, const void* phi, const void* b, void* x, double dtau,
                        double t, const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes, double* err_out,
                        double* iters_out, int* status_out, void* stream);

/* qf_pcg_solve_ragged as ONE persistent cooperative launch (the persistent
 * solve of qf_pcg_solve_persistent over the plan's work items): same outputs,
 * one dispatch, leaving as soon as every lattice has stopped.  ptab (the
 * concatenated tables) must not be NULL; the plan's largest lattice may need
 * at most 512 threads (L up to 44).  work: at least
 * n = qf_pcg_ragged_persistent_workspace_bytes(prec, plan_host) bytes; the 256
 * bytes at offset n - 256 (barrier counters) must be zero before the first
 * call on that workspace (every call leaves them zero).  Not for stream
 * capture. */
size_t qf_pcg_ragged_persistent_workspace_bytes(int prec, const void* plan_host);
int qf_pcg_solve_ragged_persistent(int prec, const void* plan_host, const void* plan_dev, const void* phi, const void* b, void* x,
                                   double dtau, double t, const void* ptab, double rtol, int max_iter, void* work, size_t work_bytes,
                                   double* err_out, double* iters_out, int* status_out, void* stream);

/* The mixed-precision solver of qf_pcg_mixed_solve over a ragged batch: complex64
 * iterations with FP64 dot products and adaptive FP64 reliable updates (at most every
 * ru_period iterations), phi64 / phi32 the concatenated fields, b and x complex128, ptab32
 * the concatenated float band tables.  A lattice stops only on an FP64 true residual, and
 * err_out is that residual.  No counterpart in the reference, whose batches share one L. */
size_t qf_pcg_mixed_ragged_workspace_bytes(const void* plan_host);
int qf_pcg_mixed_solve_ragged(const void* plan_host, const void* plan_dev, const double* phi64, const float* phi32, const void* b, void* x,
                              double dtau, double t, const float* ptab32, double rtol, int max_iter, int ru_period, void* work,
                              size_t work_bytes, double* err_out, double* iters_out, int* status_out, void* stream);

/* Fermion force Ft = -dS_f/dphi given xi = (M^dag M)^{-1} psi: the second half of
 * force_f_fast (hmc_sampler_batch.py:2174-2227, per-tau _C.b_vec_per_tau calls and
 * df_dot gathers 1969-2006) fused into one launch.  Ft has the phi layout. */
int qf_fermion_force(int prec, const void* phi, long long phi_cs, const void* xi, void* Ft, int nrhs, int L, int Ltau, double dtau,
                     double t, void* stream);

/* Boson sector (FP64, phi layout).  force_b_plaq_matfree + force_b_tau_cmp/noncmp
 * (hmc_sampler_batch.py:1338-1360, 1438-1459). */
int qf_boson_force(const double* x, double* Fp, double* Ft, int nrhs, int L, int Ltau, double cJ, double cK, int compact, void* stream);
/* leapfrog_cmp (hmc_sampler_batch.py:2332-2347): nsub boson sub-steps of dt/nsub per chain;
 * Fp/Ft hold the forces at x on entry and are updated.  tau_mask may be NULL. */
int qf_leapfrog_boson(double* x, double* p, double* Fp, double* Ft, const double* dt, const double* tau_mask, double m, int nsub, int nrhs,
                      int L, int Ltau, double cJ, double cK, int compact, void* stream);
/* p += scale * dt[chain] * F * tau_mask (fermion kicks, hmc_sampler_batch.py:2505, 2530). */
int qf_momentum_kick(double* p, const double* F, const double* dt, double scale, const double* tau_mask, int nrhs, int L, int Ltau,
                     void* stream);
/* H = S_tau + S_plaq + sum p^2/2m + Re(psi^dag xi) per chain in FP64
 * (hmc_sampler_batch.py:2417-2424, 2611-2618).  p, psi, xi may be NULL (terms
 * skipped; psi and xi together).  parts[nrhs][4] optional. */
size_t qf_hamiltonian_workspace_bytes(int nrhs);
int qf_hamiltonian(int prec, const double* x, const double* p, const void* psi, const void* xi, double* H, double* parts, int nrhs, int L,
                   int Ltau, double cJ, double cK, int compact, double m, void* work, size_t work_bytes, void* stream);

/* Bond-bond stochastic estimators (stochastic_estimator.py:1393-1654, T1 / T21 /
 * T2 / T4 pair estimators and Tvv; combined by get_bond_bond_per_b2, 2594-2641).
 * Each is a correlation C(Delta) = (1/V) sum_r a(r - Delta) b(r) of fields a, b
 * built from two noise vectors, averaged over pairs s < s' and tau slices.
 * The reference computes it per pair batch with three FFTs (1428-1430); here
 * the caller 2-D-transforms the fields of a chunk of pairs (any FFT library,
 * e.g. rocFFT), qf_obs_kspace sums Ahat(k) Bhat(-k) over every slice in FP64
 * into per-work-group partials, and one transform of the reduced [L, L]
 * array gives C.  eta and G (= M^{-1} eta) are [nrv][Ltau][L][L] complex
 * (complex64 for QF_PREC_F32, complex128 for QF_PREC_F64); phix [Ltau][L][L]
 * is the chain's x-link angle in the matching real type and U = e^{i phix}
 * its phase field (qf_obs_phase), complex of the same precision. */
#define QF_OBS_T1 0
#define QF_OBS_T21 1
#define QF_OBS_T2 2
#define QF_OBS_T4 3
#define QF_OBS_V 4
/* U[i] = exp(i phix[i]) for i < n (accurate sincos). */
int qf_obs_phase(int prec, const void* phix, void* U, long long n, void* stream);
/* Fields a -> A[j], b -> B[j] for pairs j = p0 .. p0+np-1 of pairs[2*j], pairs[2*j+1]
 * (s < s', torch.combinations order; entries outside [0, nrv) give zero fields).
 * QF_OBS_V writes v_s of Tvv for noise vectors p0 .. p0+np-1 into A (pairs, B unused).
 * A, B: [np][Ltau*L*L]. */
int qf_obs_fields(int prec, int obs, const void* eta, const void* G, const void* U, const int* pairs, int p0, int np, int nrv, int L,
                  int Ltau, void* A, void* B, void* stream);
/* part[w][k] (+)= sum over slices sl = w, w+nwg, ... < nslices of A_hat[sl][k] * B_hat[sl][-k]
 * (double2, [nwg][L*L]); with lin also part[nwg + w][k] (+)= sum A_hat[sl][k].
 * accumulate = 0 overwrites, so the first chunk needs no cleared buffer. */
int qf_obs_kspace(int prec, const void* A_hat, const void* B_hat, long long nslices, int L, int lin, int accumulate, void* part, int nwg,
                  void* stream);
/* out[k] = sum_{w < nw} part[w*n + k] (double2), in order. */
int qf_obs_reduce(const void* part, int nw, int n, void* out, void* stream);
/* Same-noise-vector sums the estimators add at one Delta each, over all nrv
 * vectors and sites, as part[block][3] double2 (nblk blocks; reduce with
 * qf_obs_reduce(part, nblk, 3, ...)): 0: G(r+x) e(r-x) U(r) U(r-x) (T1),
 * 1: G(r) e(r) (T21, T2), 2: G(r) e(r+2x) conj(U(r) U(r+x)) (T4); e = conj(eta),
 * U = exp(i phix) from qf_obs_phase. */
int qf_obs_diag(int prec, const void* eta, const void* G, const void* U, int nrv, int L, int Ltau, void* part, int nblk, void* stream);
/* The whole pair sum of one estimator in one launch, no FFT library: each
 * work-group builds a and b of one (tau slice, pair) in LDS, transforms both
 * with its own mixed-radix FFT and adds Ahat(k) Bhat(-k) in FP64 registers.
 * part: [nwg][L*L] double2 (QF_OBS_V: [2*nwg][L*L], the second half the sums
 * of Ahat), reduced with qf_obs_reduce.  count: pairs (QF_OBS_V: noise
 * vectors, <= nrv).  Returns QF_ERR_UNSUPPORTED where the slices do not fit
 * in LDS (complex128 with L > 58, any L > 78). */
int qf_obs_fused_grid(int prec, int L);   /* work-groups that fill the device with the pair estimators; 0 = unsupported */
int qf_obs_fused(int prec, int obs, const void* eta, const void* G, const void* U, const int* pairs, int count, int nrv, int L, int Ltau,
                 void* part, int nwg, void* stream);

#ifdef __cplusplus
}
#endif

#endif
