// hf_kernels.hpp -- parameter blocks and host launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hf_device.hpp"

namespace hf {

struct AsmParams {
    double inverse_time;       // 1/time_step (mod_setup.f90:45)
    double kb[6];              // kappaBound x1, Nx, y1, Ny, z1, Nz
    double tb[6];              // T_Bath x1, x2, y1, y2, z1, z2
    double cg_pd[6];           // CG face power densities (mod_boundary.f90:43-66)
    double br;
    int32_t isteady, icattaneo;
    int32_t periodic[3];       // periodic axes (no boundary terms; length >= 3)
    int32_t cg[6];
};

struct HeatParams {
    double qmode[7];           // mode power of this step, index = iheater (mod_heating.f90:59-117)
    double q5b, q5c;           // mode 5: + (q5b * tau) * q5c
    double dt2;                // time_step**2 (mod_heating.f90:56)
    double tb4;                // T_Bath**4
    double heated_volume;
    int32_t any_heater;
};

// Heater source of one cell and step (src/heatflow/mod_heating.f90:59-123): qmode[m] is the
// mode-m power of this step computed on the host; mode 5 adds (q5b * tau) * q5c with
// tau = tau_s * dt^2 (:56, :104-105); the radiative term -em L1 L2 sigma (T^4 - T_Bath^4) on
// every cell (:121-123).
__device__ __forceinline__ double heater_q(const HeatParams &hp, int8_t mode, double tau_s,
                                           double ells, double tp) {
    double q;
    switch (mode) {
        case 0: q = 0.0; break;
        case 5: q = hp.qmode[5] + hp.q5b * (tau_s * hp.dt2) * hp.q5c; break;
        default: q = (mode >= 1 && mode <= 6) ? hp.qmode[mode] : 0.0; break;
    }
    return q - ells * (pow(tp, 4.0) - hp.tb4);
}

// Mode of S (compile-time in the step kernels): kSteady (isteady /= 0), kFourier (transient,
// icattaneo /= 1), kCattaneo (transient, icattaneo == 1)
enum : int { kSteady = 0, kFourier = 1, kCattaneo = 2 };
__host__ __device__ constexpr int s_mode(int32_t isteady, int32_t icattaneo) {
    return isteady != 0 ? kSteady : (icattaneo == 1 ? kCattaneo : kFourier);
}

// rc_mask: the index mask of the rho c_p and tau' rho c_p reads, ~0 (per cell) or 0 (uniform
// fields: every cell reads element 0)
struct RhsParams {
    double inverse_time;
    double tol;
    int64_t itmax;
    int32_t isteady, icattaneo;
    uint32_t rc_mask = ~0u;
};

struct IrParams {
    int64_t itmax;             // refinement sweep cap
    int32_t from_state;
};

// the fused mixed-precision step (k_ir_step_z)
struct StepParams {
    double inverse_time;
    int32_t isteady, icattaneo;
    int32_t tiny_guess;        // form the next residual for this outcome of its TINY test
    int32_t store_res;         // store this step's residual (needed only for a further sweep)
    uint32_t rc_mask = ~0u;    // as RhsParams::rc_mask
};

// adjoint step (hf_adjoint.hip): parameters and the derivatives of hmatrixfunc's face
// conductance G = 2 k_I k_J / ((k_I + k_J) L^2) [x BR if k_I != k_J] with respect to k_I, for
// an interior face and a bath face (src/heatflow/mod_hmatrix.f90:288-314, 357-385)
struct AdjParams {
    double inverse_time;
    double kb[6], tb[6], br;
    int32_t isteady, icattaneo;
    int32_t bface[6];   // face is a domain boundary that enters the diagonal (non-periodic)
    int32_t cg[6];
    // index masks of the rho c_p / tau' rho c_p and volume reads: 0 where the field is uniform
    // (every cell reads element 0), else ~0
    uint32_t rc_mask, vol_mask;
    // radiative source (src/heatflow/mod_heating.f90:121-123, 164-167): with heaters present,
    // Qdens_I holds -em_I L1 L2 sigma (T_I^4 - T_bath^4) / V_heated, so dS^n_I/dT^n_I gains
    // rad_scale * ells_I * (T^n_I)^3 with rad_scale = 4 / V_heated and ells = em L1 L2 sigma;
    // rad_scale = 0 when no cell radiates
    double rad_scale;
    const double *ells;   // em L1 L2 sigma per cell
    const double *tn;     // T^n, the state the step started from
};

__device__ __forceinline__ double dG_int(double ki, double kj, double L, double br) {
    double s = ki + kj;
    double v = 2.0 * kj * kj / (s * s * L * L);
    return ki != kj ? v * br : v;
}

__device__ __forceinline__ double dG_bnd(double kb, double k, double L, double br) {
    double s = kb + k;
    double v = 2.0 * kb * kb / (s * s * L * L);
    return k != kb ? v * br : v;
}

// z-marching adjoint step on a grid without periodic axes (the same per-cell sums as the flat
// k_adj_step of hf_adjoint.hip, in the same order)
void launch_adj_step_z(const Dims &g, const AdjParams &q, const double *T, const double *lam,
                       const double *kappa, const double *lx, const double *ly, const double *lz,
                       const double *rhoc, const double *cvec, const double *vol, double *g_lo,
                       double *g_new, double *gk, double *gp, hipStream_t st);

void launch_assemble(const Dims &g, const AsmParams &ap, const double *kappa, const double *lx,
                     const double *ly, const double *lz, const double *rhoc,
                     const double *tau_s, double *d, double *cx, double *cy, double *cz,
                     double *bvec, Scal *sc, hipStream_t st);
void launch_scale(const Dims &g, const double *d, const double *cx, const double *cy,
                  const double *cz, double *sv, double *hx, double *hy, double *hz, float *fx,
                  float *fy, float *fz, hipStream_t st);
void launch_coo(const Dims &g, const double *d, const double *cx, const double *cy,
                const double *cz, int64_t *cnt, int64_t *off, int64_t *bsum, int64_t *total,
                double *val, int64_t *irow, int64_t *jcol, int64_t len, int64_t *count_host,
                hipStream_t st);
void launch_qdens(const Dims &g, const HeatParams &hp, const int8_t *iheater, const double *ells,
                  const double *tau_s, const double *tp, const double *qextra, double *qd,
                  Scal *sc, hipStream_t st);
void launch_rhs_res_state64(const Dims &g, const RhsParams &rp, const double *tp,
                            const double *tpp, const double *rhoc, const double *cvec,
                            const double *qd, const double *bvec, const double *d,
                            const double *cx, const double *cy, const double *cz, double *rhat,
                            double *sout, Scal *sc, double *partials, hipStream_t st);
void launch_rhs_res_state32(const Dims &g, const RhsParams &rp, const double *tp,
                            const double *tpp, const double *rhoc, const double *cvec,
                            const double *qd, const double *bvec, const double *d,
                            const double *cx, const double *cy, const double *cz, float *rhat,
                            double *sout, double *xout, Scal *sc, double *partials,
                            hipStream_t st);
void launch_rhs_res_solve64(const Dims &g, const RhsParams &rp, const double *x,
                            const double *b, const double *d, const double *cx,
                            const double *cy, const double *cz, double *rhat, Scal *sc,
                            double *partials, hipStream_t st);
void launch_rhs_res_solve32(const Dims &g, const RhsParams &rp, const double *x,
                            const double *b, const double *d, const double *cx,
                            const double *cy, const double *cz, float *rhat, Scal *sc,
                            double *partials, hipStream_t st);
// first: the solve's first iteration (p = r; reductions only, nothing stored).  Iteration
// k >= 1 also adds the previous iteration's term to the solution: y = alpha p_old (yfirst,
// k == 1) or y += alpha p_old, alpha being the scalars' (that iteration's) alpha.
template <typename T>
void launch_cg_a(const Dims &g, bool first, const T *r, const T *pold, T *pnew, T *q, T *y,
                 bool yfirst, const T *hx, const T *hy, const T *hz, Scal *sc, double *partials,
                 hipStream_t st);
// the rest of the first iteration: rw = r - alpha A^ r (r unchanged)
template <typename T>
void launch_cg_b1(const Dims &g, const T *r, const T *hx, const T *hy, const T *hz, T *rw,
                  Scal *sc, double *partials, hipStream_t st);
// r -= alpha q
template <typename T>
void launch_cg_b(const Dims &g, T *r, const T *q, Scal *sc, double *partials, hipStream_t st);
// after a solve: y += alpha p_K for its last iteration K >= 1 (pa, pb = p_k for odd / even k)
template <typename T>
void launch_cg_tail(const Dims &g, T *y, const T *pa, const T *pb, const Scal *sc,
                    hipStream_t st);
// The consumers of a solve's y (CgY: y after k_cg_tail, or alpha r0 when the solve ended in
// its first iteration) and of that solve's scalars.
// linbcg's stop norms for hf_solve (k_wsum2 + k_fin_wstop): mode 0 records ||b||^2 of v and
// switches the scaled-residual test off; mode 1 / 2 (itol 1 / 2) the weighted norm^2 of the
// scaled residual v after iteration expect - 1, ending the solve when it meets tol
void launch_wstop(const Dims &g, const double *v, const double *s, int32_t mode, int64_t expect,
                  double tol, Scal *sc, double *partials, hipStream_t st);
void launch_finalize(const Dims &g, bool from_state, const double *tp, double *tpp_io,
                     const double *sv, const double *y, const double *r0, Scal *sc,
                     hipStream_t st);
void launch_ir_res(const Dims &g, const IrParams &ip, const double *x, const float *y,
                   const float *r0, const Scal *inner, const double *sv, const double *S,
                   const double *d, const double *cx, const double *cy, const double *cz,
                   const double *tp, double *xout, float *rhat, Scal *sc, double *partials,
                   hipStream_t st);
// Fused mixed-precision step: x' = base + s y (base = the state's x0, or xb), this step's
// FP64 residual at x' (FP32 -> rres if sp.store_res; its norm always), x' -> xout, and the next
// step's scaled r^0 from (x', T^n), with the TINY shift if sp.tiny_guess (FP32 -> rnext);
// 4 partials.  The pass is a function of its inputs: run again, it writes the same values.
// qb = Qdens + B (per-step constant in this path)
void launch_ir_step(const Dims &g, const StepParams &sp, bool from_state, const double *tp,
                    const double *tpp, const double *xb, const float *y, const float *r0,
                    const Scal *inner, const double *rhoc, const double *cvec,
                    const double *qb, const double *d, const double *cx, const double *cy,
                    const double *cz, double *xout, float *rres, float *rnext, Scal *sc,
                    double *partials, hipStream_t st);
// Fused FP64 step: k_finalize of this step (x -> xout) and k_rhs_res_z of the next from
// (x, T^n), with the next step's TINY shift if `guess` (r^0 -> rnext); bit-identical to the two
// kernels run apart.  2 partials (as k_rhs_res_z); NaN flags of the next S go to flags_next.
void launch_step64(const Dims &g, const RhsParams &rp, bool guess, const double *tp,
                   const double *tpp, const double *y, const double *r0, const double *rhoc,
                   const double *cvec, const double *qd, const double *bvec, const double *d,
                   const double *cx, const double *cy, const double *cz, double *xout,
                   double *rnext, Scal *sc, double *partials, hipStream_t st);
// one-block finishing kernels: sum the per-block partials in block order, update the scalars
// allow_short: the solve may end after its first iteration (k_fin_a)
void launch_fin_start(const Dims &g, const double *partials, Scal *sc, double tol, int64_t itmax,
                      bool allow_short, hipStream_t st);
void launch_fin_a(const Dims &g, const double *partials, Scal *sc, hipStream_t st);
void launch_fin_b(const Dims &g, const double *partials, Scal *sc, hipStream_t st);
void launch_fin_b1(const Dims &g, const double *partials, Scal *sc, hipStream_t st);  // after k_cg_b1
void launch_fin_ir(const Dims &g, const double *partials, Scal *sc, int64_t sweeps_max,
                   hipStream_t st);
// tiny_guess: the TINY case the step's rnext was formed for; reused: the sweep took the last
// inner solve's alpha (launch_carry_start), passes only within the inner target, and updates
// inner->alpha to this step's own step length
void launch_fin_step(const Dims &g, const double *partials, Scal *sc, Scal *inner,
                     int64_t sweeps_max, bool tiny_guess, bool reused, bool freeze,
                     hipStream_t st);
// two reusing steps of the mixed path in one pass, and its finishing kernel (hf_stencil.hip,
// hf_kernels.hip)
void launch_ir_step2(const Dims &g, const StepParams &sp, const double *tp, const double *tpp,
                     const float *r0, const Scal *inner, const double *rhoc, const double *cvec,
                     const double *qb, const double *d, const double *cx, const double *cy,
                     const double *cz, double *x1out, double *x2out, float *rnext, Scal *sc,
                     double *partials, hipStream_t st);
void launch_fin_step2(const Dims &g, const double *partials, Scal *sc, Scal *inner,
                      bool tiny_guess, bool freeze, bool solved, bool reject0, double tol,
                      double inner_tol, double safety, hipStream_t st);
// after launch_step64: the next step's norms for the TINY case `tiny_guess`
void launch_fin_next(const Dims &g, const double *partials, Scal *sc, bool tiny_guess,
                     hipStream_t st);
// start of a carried step (k_step_begin; the scalars from the norms the last fused pass left,
// as launch_fin_start for a residual from the state) and, for a mixed-precision step (inner not
// null), the first inner solve's scalars (as launch_inner_begin(sc, inner, ..., outer_tol =
// tol)), in one launch.  reuse (mixed): no first inner solve; the first sweep takes
// y = alpha r^0 with the last solve's alpha and must meet the inner target
// copy of `bytes` (a multiple of 8) device bytes into pinned host memory, then *seq_host = seq
void launch_publish(const void *src, void *dst, size_t bytes, int64_t *seq_host, int64_t seq,
                    hipStream_t st);
void launch_carry_start(Scal *sc, Scal *inner, double tol, int64_t itmax, bool allow_short,
                        double inner_tol, int64_t inner_itmax, double safety, bool reuse,
                        hipStream_t st);
void launch_inner_begin(Scal *outer, Scal *inner, double inner_tol, double outer_tol,
                        int64_t inner_itmax, double safety, bool allow_short, hipStream_t st);
void launch_step_begin(Scal *sc, hipStream_t st);
void launch_tiny_from_state(const Dims &g, const double *tp, const double *tpp, Scal *sc,
                            hipStream_t st);
void launch_sum2(const Dims &g, const double *a, const double *b, double *out, hipStream_t st);
// out = sum_{j < m} w[j] v[j] (v[0] the newest of the last m solutions)
constexpr int kExtrapMax = 8;
struct ExtrapArgs {
    const double *v[kExtrapMax];
    double w[kExtrapMax];
    int32_t m;
};
void launch_extrap(const Dims &g, const ExtrapArgs &a, double *out, hipStream_t st);

// ---- FP64 time steps of a small grid in one persistent workgroup (hf_persist.hip) ----------

constexpr int kSmallThreads = 1024;             // the one workgroup
constexpr uint32_t kSmallMaxCells = 8192;       // grids up to this many cells take the path
constexpr int32_t kSmallMaxSteps = 256;         // steps per launch (per-step heat parameters)

// what a run of k_run_small leaves (written by thread 0 at its end)
struct SmallOut {
    int64_t steps;          // steps completed (a step stopped by a NaN guard is not counted)
    int64_t iters;          // CG iterations over them
    int64_t last_iter;      // of the last completed step
    int64_t first_stops;    // solves that ended on their first iteration
    double last_err;        // ||r^|| / ||b^|| of the last completed step
    int32_t flags;          // NaN bits of the step that stopped the run, TINY_NEXT of the last
    int32_t rot;            // (T^n, T^{n-1}, spare) = t[rot], t[(rot + 1) % 3], t[(rot + 2) % 3]
};

struct SmallRun {
    Dims g;
    double inverse_time, tol;
    int64_t itmax;
    int32_t isteady, icattaneo, allow_short, nsteps;
    int32_t dyn;            // Qdens from T^n every step (hp[step]): time-dependent heating
    uint32_t rc_mask;       // as RhsParams::rc_mask
    double *t[3];           // T^n, T^{n-1}, spare
    const double *rhoc, *cvec, *bvec, *d, *cx, *cy, *cz, *hx, *hy, *hz, *sv;
    double *qd;             // Qdens (read; written per step when dyn)
    const HeatParams *hp;   // per step (dyn)
    const int8_t *iheater;
    const double *ells, *tau_s, *qextra;
    double *r, *p0, *p1, *q, *y;   // CG vectors
    Scal *sc;               // flags in (TINY_NEXT of the last step) and out
    SmallOut *out;
};
void launch_run_small(const SmallRun &a, hipStream_t st);

}  // namespace hf
