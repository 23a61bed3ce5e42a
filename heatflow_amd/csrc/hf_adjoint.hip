// hf_adjoint.hip -- reverse-time adjoint of the HeatFlow time step (no reference counterpart).
//
// Forward step n (src/heatflow/mod_evolve.f90:112-170):  H T^{n+1} = S^n,
//   S^n = -a T^n - Qdens - B + c (T^{n-1} - 2 T^n),  a = rho c_p / dt,  c = tau' rho c_p
// (c = 0 for Fourier, a = c = 0 for steady state).  H is symmetric (mod_setup.f90:138-145), so
// the adjoint solve is the forward solver on the same operator:
//   lambda^{n+1} = H^{-1} gbar^{n+1},
//   gbar^n     += -(a + 2c) lambda^{n+1},   gbar^{n-1} += c lambda^{n+1},
//   gbar^n_I   += 4 em_I L1 L2 sigma (T^n_I)^3 / V_heated lambda^{n+1}_I   (radiative source:
//                 with heaters present Qdens_I = (Q_I - em_I L1 L2 sigma (T^n_I^4 - T_bath^4))
//                 / V_heated, src/heatflow/mod_heating.f90:121-123, 164-167),
//   dL/dkappa_I += sum_faces dG/dkappa_I (lambda_I - lambda_J)(T_I - T_J)
//                + sum_bfaces dGb/dkappa_I lambda_I (T_I - [T_bath surviving in B_I]),
//   dL/dP_I     += -lambda_I / V_I          (P_I enters as Qdens_I += P_I / V_I),
// with T = T^{n+1}.  G is the harmonic-mean face conductance of hmatrixfunc
// (src/heatflow/mod_hmatrix.f90:288-314, 357-385) for per-axis uniform lengths:
//   G = 2 k_I k_J / ((k_I + k_J) L^2) [x BR if k_I != k_J],  dG/dk_I = 2 k_J^2 / ((k_I+k_J)^2 L^2).
// The BR switch is treated as locally constant (it is discontinuous at k_I == k_J).
//
// Memory: (T^n, T^{n-1}) checkpoints every K steps plus the K states of the segment being
// reversed, recomputed from its checkpoint; for K = 1 every state once (nsteps + 1 of them) and
// nothing recomputed.  K = 0 takes K = 1 when that fits in the device memory free at the call,
// else K = ceil(sqrt(nsteps)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "hf_ctx.hpp"

namespace hf {

struct AdjointState {
    size_t N = 0;
    size_t ckpt_cap = 0;      // doubles allocated for ckpt
    int64_t k_cap = 0;
    // nseg x (T^n, T^{n-1}), T^n at ckpt + 2 N s; for K = 1, T^0 .. T^nsteps at ckpt + N n
    double *ckpt = nullptr;
    double *seg = nullptr;    // max(K - 1, 1) x N: the recomputed states of one segment
    double *g_hi = nullptr, *g_lo = nullptr, *g_new = nullptr;
    // the last adjoint solutions, newest first (hist[0] is lambda^{n+1}), and the next guess
    double *hist[kExtrapMax] = {};
    double *lam_next = nullptr;
    double *gk = nullptr, *gp = nullptr;
    double *probe_val = nullptr;  // device, nprobe
    int64_t *probe_idx = nullptr;
    double *probe_w = nullptr, *probe_t = nullptr;
    int64_t nprobe_cap = 0;
    double *fin = nullptr;    // end state of the forward run (T^N, T^{N-1})
    double *seed = nullptr;   // hf_grad's dL/dT^N
    int64_t itime0 = 1, nsteps = 0, K = 1;
    bool valid = false;
};

void adjoint_free(AdjointState *a) {
    if (!a) return;
    hipFree(a->ckpt); hipFree(a->seg);
    hipFree(a->g_hi); hipFree(a->g_lo); hipFree(a->g_new);
    for (double *h : a->hist) hipFree(h);
    hipFree(a->lam_next);
    hipFree(a->gk); hipFree(a->gp);
    hipFree(a->probe_val); hipFree(a->probe_idx); hipFree(a->probe_w); hipFree(a->probe_t);
    hipFree(a->fin); hipFree(a->seed);
    delete a;
}

void adjoint_invalidate(AdjointState *a) {
    if (a) a->valid = false;
}


// One reverse step: gradient contributions of step n (T = T^{n+1}, lam = lambda^{n+1}) and the
// adjoint history update.  g_lo (partial gbar^n) += -(a+2c) lam; g_new = c lam (gbar^{n-1}).
__global__ void __launch_bounds__(kBlock)
k_adj_step(Dims g, AdjParams q, const double *__restrict__ T, const double *__restrict__ lam,
           const double *__restrict__ kappa, const double *__restrict__ lx,
           const double *__restrict__ ly, const double *__restrict__ lz,
           const double *__restrict__ rhoc, const double *__restrict__ cvec,
           const double *__restrict__ vol, double *__restrict__ g_lo,
           double *__restrict__ g_new, double *__restrict__ gk, double *__restrict__ gp) {
    // cell lengths are uniform per axis (checked when the adjoint starts): read once
    const double Ls[3] = {lx[0], ly[0], lz[0]};
    HF_FOR_CELLS(g, I) {
        Nbr n = neighbours(g, I);
        const double li = lam[I], ti = T[I], ki = kappa[I];
        double acc = 0.0;
        const bool has[6] = {n.hxm, n.hxp, n.hym, n.hyp, n.hzm, n.hzp};
        const uint32_t nb[6] = {n.xm, n.xp, n.ym, n.yp, n.zm, n.zp};
        const uint32_t pos[3] = {n.x, n.y, n.z};
        const uint32_t len[3] = {g.nx, g.ny, g.nz};
        // which boundary faces still contribute to B_I after CG assignments (mod_boundary.f90:72-150)
        int survive = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int ax = f >> 1;
            const bool at_face = pos[ax] == ((f & 1) ? len[ax] - 1 : 0u);
            if (!q.bface[f] || !at_face) continue;
            if (q.cg[f]) survive = 0;
            else survive |= (1 << f);
        }
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int ax = f >> 1;
            if (has[f]) {
                const uint32_t J = nb[f];
                acc += dG_int(ki, kappa[J], Ls[ax], q.br) * (li - lam[J]) * (ti - T[J]);
            } else {
                const bool at_face = pos[ax] == ((f & 1) ? len[ax] - 1 : 0u);
                if (q.bface[f] && at_face) {
                    const double tb = (survive >> f) & 1 ? q.tb[f] : 0.0;
                    acc += dG_bnd(q.kb[f], ki, Ls[ax], q.br) * li * (ti - tb);
                }
            }
        }
        gk[I] += acc;
        gp[I] += -li / vol[I & q.vol_mask];
        double a = 0.0, cc = 0.0;
        if (q.isteady == 0) {
            a = q.inverse_time * rhoc[I & q.rc_mask];
            cc = q.icattaneo == 1 ? cvec[I & q.rc_mask] : 0.0;
        }
        double gl = g_lo[I] + -(a + 2.0 * cc) * li;
        if (q.rad_scale != 0.0) {
            const double tn = q.tn[I];
            gl = gl + q.rad_scale * q.ells[I] * (tn * tn * tn) * li;
        }
        g_lo[I] = gl;
        g_new[I] = cc * li;
    }
}

// loss and the seed gbar^{N} at the probes; one thread, fixed order (duplicates accumulate)
__global__ void k_adj_seed(int64_t np, const int64_t *__restrict__ idx,
                           const double *__restrict__ w, const double *__restrict__ t,
                           const double *__restrict__ T, double *__restrict__ g,
                           double *__restrict__ out_loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double L = 0.0;
    for (int64_t k = 0; k < np; ++k) {
        const double diff = T[idx[k]] - t[k];
        L += w[k] * diff * diff;
        g[idx[k]] += 2.0 * w[k] * diff;
    }
    *out_loss = L;
}

// Points of the adjoint guess's extrapolation.  At 384^3 (500 steps, tol 1e-13) the gradient
// took 5.33 s with 2 points (linear), 4.80 with 3, 4.48 with 4, 4.21 with 5, 4.15 with 6 and
// 4.32 with 8 (FP32-IR; FP64 6.35, 5.13 with 4 and 4.43 with 6).
constexpr int kAdjExtrapPoints = 6;
static_assert(kAdjExtrapPoints <= kExtrapMax, "extrapolation history");

static int ensure(hf_ctx *c, double **p, size_t n) {
    if (*p) return HF_OK;
    if (hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return ctx_set_err(c, HF_ERR_ALLOC, "hf_grad: device allocation failed");
    }
    return HF_OK;
}

#define ATRY(x)                        \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != HF_OK) return rc_;  \
    } while (0)
#define AHIP(c, call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ctx_set_err(c, HF_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// restore (T^n, T^{n-1}) into the context state and recompute the TINY flag
static int restore_state(hf_ctx *c, const double *tp, const double *tpp) {
    const size_t B8 = (size_t)c->g.N * 8;
    reset_step_heuristics(c);
    AHIP(c, hipMemcpyAsync(c->T[0], tp, B8, hipMemcpyDeviceToDevice, c->st));
    AHIP(c, hipMemcpyAsync(c->T[1], tpp, B8, hipMemcpyDeviceToDevice, c->st));
    launch_step_begin(c->sc, c->st);
    launch_tiny_from_state(c->g, c->T[0], c->T[1], c->sc, c->st);
    AHIP(c, hipGetLastError());
    return HF_OK;
}

static int validate(hf_ctx *c) {
    if (!c->assembled || !c->has_state)
        return ctx_set_err(c, HF_ERR_STATE, "adjoint: needs hf_assemble and hf_set_state first");
    // per-axis uniform lengths (the reference's system.in gives Lx/nx etc., mod_inputs.f90:619-622),
    // found by k_uniform_check at upload
    if (!c->uniform_len)
        return ctx_set_err(c, HF_ERR_ARG, "adjoint: cell lengths must be uniform per axis");
    return HF_OK;
}

// Forward run of nsteps from the current state, keeping (T^n, T^{n-1}) every ckpt_every steps.
// The checkpoint spacing for ckpt_every = 0: every state (K = 1) when nsteps + 1 states fit in
// the device memory free now, counting what this context's checkpoints already hold and
// leaving max(2 GiB, 5% of the device) to spare; else ceil(sqrt(nsteps)).
static int64_t auto_spacing(hf_ctx *c, int64_t nsteps) {
    const int64_t ks = std::max<int64_t>(1, (int64_t)std::ceil(std::sqrt((double)nsteps)));
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        (void)hipGetLastError();
        return ks;
    }
    const AdjointState *A = c->adj;
    const size_t held = A ? ((A->ckpt ? A->ckpt_cap : 0) + (A->seg ? A->k_cap * A->N : 0)) * 8 : 0;
    const size_t need = (size_t)(nsteps + 1) * c->g.N * 8;
    const size_t spare = std::max<size_t>((size_t)2 << 30, totalb / 20);
    return need + spare <= freeb + held ? 1 : ks;
}

int adjoint_forward(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t ckpt_every) {
    ATRY(validate(c));
    if (nsteps < 1 || ckpt_every < 0)
        return ctx_set_err(c, HF_ERR_ARG, "hf_forward_checkpointed: bad step counts");
    const size_t N = c->g.N, B8 = N * 8;
    int64_t K = ckpt_every > 0 ? ckpt_every : auto_spacing(c, nsteps);
    if (!c->adj) c->adj = new AdjointState();
    AdjointState &A = *c->adj;
    // the checkpoints are rewritten below: valid again only once the run has completed
    A.valid = false;
    // the checkpoint buffer for spacing K; spacing 0 that chose every state falls back to
    // ceil(sqrt(nsteps)) if the allocation fails after all (memory taken since the estimate)
    for (;;) {
        const size_t need =
            K == 1 ? (size_t)(nsteps + 1) * N : (size_t)2 * N * ((nsteps + K - 1) / K);
        // a segment of K steps is recomputed up to its second-to-last state (the last one is
        // the next checkpoint's): K - 1 states
        const int64_t kseg = std::max<int64_t>(K - 1, 1);
        if (A.N != N || A.ckpt_cap < need || A.k_cap < kseg) {
            hipFree(A.ckpt); hipFree(A.seg);
            A.ckpt = A.seg = nullptr;
            A.ckpt_cap = need;
            A.k_cap = std::max(kseg, A.N == N ? A.k_cap : 0);
        }
        // (debug bit 10, a test aid: the every-state allocation of spacing 0 fails)
        const bool auto_every = ckpt_every == 0 && K == 1;
        const int rc = auto_every && (c->debug & 1024)
                           ? ctx_set_err(c, HF_ERR_ALLOC, "hf_grad: device allocation failed")
                           : ensure(c, &A.ckpt, A.ckpt_cap);
        const int64_t ks = std::max<int64_t>(1, (int64_t)std::ceil(std::sqrt((double)nsteps)));
        if (rc != HF_OK && auto_every && ks > 1) {
            A.ckpt_cap = 0;
            K = ks;
            continue;
        }
        ATRY(rc);
        break;
    }
    const int64_t nseg = (nsteps + K - 1) / K;
    const bool every = K == 1;   // every state once, nothing recomputed
    if (A.N != N) {
        hipFree(A.g_hi); hipFree(A.g_lo); hipFree(A.g_new);
        for (double *&h : A.hist) {
            hipFree(h);
            h = nullptr;
        }
        hipFree(A.lam_next);
        hipFree(A.gk); hipFree(A.gp); hipFree(A.fin);
        A.g_hi = A.g_lo = A.g_new = A.lam_next = A.gk = A.gp = A.fin = nullptr;
        A.N = N;
    }
    ATRY(ensure(c, &A.seg, (size_t)N * A.k_cap));
    ATRY(ensure(c, &A.fin, (size_t)2 * N));
    for (int64_t s = 0; s < nseg; ++s) {
        if (every) {
            AHIP(c, hipMemcpyAsync(A.ckpt + N * s, c->T[0], B8, hipMemcpyDeviceToDevice, c->st));
        } else {
            AHIP(c, hipMemcpyAsync(A.ckpt + (size_t)2 * N * s, c->T[0], B8,
                                   hipMemcpyDeviceToDevice, c->st));
            AHIP(c, hipMemcpyAsync(A.ckpt + (size_t)2 * N * s + N, c->T[1], B8,
                                   hipMemcpyDeviceToDevice, c->st));
        }
        const int64_t e = std::min(nsteps, (s + 1) * K);
        for (int64_t n = s * K; n < e; ++n) ATRY(ctx_step(c, itime0 + n, n + 1 == nsteps));
    }
    if (every)
        AHIP(c, hipMemcpyAsync(A.ckpt + N * nseg, c->T[0], B8, hipMemcpyDeviceToDevice, c->st));
    AHIP(c, hipMemcpyAsync(A.fin, c->T[0], B8, hipMemcpyDeviceToDevice, c->st));
    AHIP(c, hipMemcpyAsync(A.fin + N, c->T[1], B8, hipMemcpyDeviceToDevice, c->st));
    AHIP(c, hipStreamSynchronize(c->st));
    A.itime0 = itime0;
    A.nsteps = nsteps;
    A.K = K;
    A.valid = true;
    return HF_OK;
}

// Reverse sweep of the last adjoint_forward, seeded with gbar^{N} = dL/dT^{N} (length N, host
// or device).  Leaves the context at the end state of the forward run.
int adjoint_backward(hf_ctx *c, const double *seed, double *dl_dkappa, double *dl_dpower) {
    if (!c->adj || !c->adj->valid)
        return ctx_set_err(c, HF_ERR_STATE, "hf_adjoint: no checkpointed forward run");
    if (!seed) return ctx_set_err(c, HF_ERR_ARG, "hf_adjoint: null seed");
    AdjointState &A = *c->adj;
    const hf_params &p = c->p;
    const size_t N = c->g.N, B8 = N * 8;
    ATRY(ensure(c, &A.g_hi, N)); ATRY(ensure(c, &A.g_lo, N)); ATRY(ensure(c, &A.g_new, N));
    ATRY(ensure(c, &A.gk, N)); ATRY(ensure(c, &A.gp, N)); ATRY(ensure(c, &A.lam_next, N));
    // history buffers for the extrapolated guesses: debug bit 5 keeps two (linear guess)
    const int hmax = (c->debug & 32) ? 2 : kAdjExtrapPoints;
    for (int j = 0; j < hmax; ++j) ATRY(ensure(c, &A.hist[j], N));
    AHIP(c, hipMemcpyAsync(A.g_hi, seed, B8, hipMemcpyDefault, c->st));
    AHIP(c, hipMemsetAsync(A.g_lo, 0, B8, c->st));
    AHIP(c, hipMemsetAsync(A.hist[0], 0, B8, c->st));
    AHIP(c, hipMemsetAsync(A.gk, 0, B8, c->st));
    AHIP(c, hipMemsetAsync(A.gp, 0, B8, c->st));

    AdjParams q{};
    q.inverse_time = 1.0 / p.time_step;
    for (int f = 0; f < 6; ++f) {
        q.kb[f] = p.kappa_bound[f];
        q.tb[f] = p.t_bath[f];
        q.cg[f] = p.cg[f] != 0;
        q.bface[f] = !p.periodic[f >> 1];
    }
    q.br = p.br;
    q.isteady = p.isteady;
    q.icattaneo = p.icattaneo;
    q.rc_mask = c->uniform_rc && (c->debug & 8) == 0 ? 0u : ~0u;
    q.vol_mask = c->uniform_vol && (c->debug & 8) == 0 ? 0u : ~0u;
    // the radiative term enters S only with heaters present (mod_evolve.f90:79-80)
    q.rad_scale = (c->any_heater && c->any_em && c->heated_volume > 0.0)
                      ? 4.0 / c->heated_volume : 0.0;
    q.ells = c->ells;

    const int64_t K = A.K, nsteps = A.nsteps, itime0 = A.itime0;
    const int64_t nseg = (nsteps + K - 1) / K;
    const double atol = c->adj_tol > 0.0 ? c->adj_tol : p.tol;   // hf_set_adjoint_tol
    // the z-marching kernel for grids without periodic axes; debug bit 4 keeps the flat one
    const bool zmarch_adj =
        !(p.periodic[0] || p.periodic[1] || p.periodic[2]) && (c->debug & 16) == 0;
    // adjoint solves done; hist[0 .. min(solved, hmax) - 1] hold the last solutions, newest first
    int64_t solved = 0;
    for (int64_t s = nseg - 1; s >= 0; --s) {
        const int64_t b = s * K, e = std::min(nsteps, (s + 1) * K);
        // T^e, the state after the segment's last step, is the next checkpoint's (or the final
        // state): the recomputation stops one step short (none at all for ckpt_every = 1,
        // which keeps every state)
        const double *Te = K == 1                 ? A.ckpt + N * (s + 1)
                           : s + 1 < nseg         ? A.ckpt + (size_t)2 * N * (s + 1)
                                                  : A.fin;
        if (e - 1 > b)
            ATRY(restore_state(c, A.ckpt + (size_t)2 * N * s, A.ckpt + (size_t)2 * N * s + N));
        for (int64_t n = b; n < e - 1; ++n) {
            ATRY(ctx_step(c, itime0 + n, false));
            AHIP(c, hipMemcpyAsync(A.seg + (size_t)N * (n - b), c->T[0], B8,
                                   hipMemcpyDeviceToDevice, c->st));
        }
        for (int64_t n = e - 1; n >= b; --n) {
            int64_t it = 0;
            // lambda^{n+1} = H^{-1} gbar^{n+1}.  Initial guess: 0 for the first solve, then
            // the polynomial through the last m = min(solved, hmax) solutions at equal steps,
            // extrapolated one step: weights (-1)^(j) C(m, j+1) for the j-th newest (m = 2:
            // 2 lambda^{n+2} - lambda^{n+3}, the adjoint counterpart of the forward guess
            // 2 T^n - T^{n-1}; m = 3: 3, -3, 1; ...).
            if (solved == 0) {
                ATRY(ctx_solve_dev(c, A.g_hi, A.hist[0], atol, p.itmax, &it));
            } else {
                const int m = (int)std::min<int64_t>(solved, hmax);
                ExtrapArgs ea{};
                ea.m = m;
                double binom = 1.0;   // C(m, j+1), built up by the multiplicative formula
                for (int j = 0; j < m; ++j) {
                    binom = binom * (double)(m - j) / (double)(j + 1);
                    ea.v[j] = A.hist[j];
                    ea.w[j] = (j % 2 == 0) ? binom : -binom;
                }
                launch_extrap(c->g, ea, A.lam_next, c->st);
                ATRY(ctx_solve_dev(c, A.g_hi, A.lam_next, atol, p.itmax, &it));
                // rotate: the oldest buffer in use takes the next guess
                double *oldest = A.hist[hmax - 1];
                for (int j = hmax - 1; j > 0; --j) A.hist[j] = A.hist[j - 1];
                A.hist[0] = A.lam_next;
                A.lam_next = oldest;
            }
            solved += 1;
            c->stats.adjoint_solves += 1;
            c->stats.adjoint_iters += it;
            const double *Tn1 = n == e - 1 ? Te : A.seg + (size_t)N * (n - b);
            // T^n, the state step n started from (the radiative term's derivative)
            q.tn = K == 1   ? A.ckpt + N * s
                   : n == b ? A.ckpt + (size_t)2 * N * s
                            : A.seg + (size_t)N * (n - 1 - b);
            if (zmarch_adj)
                launch_adj_step_z(c->g, q, Tn1, A.hist[0], c->kappa, c->lx, c->ly, c->lz, c->rhoc,
                                  c->cvec, c->vol, A.g_lo, A.g_new, A.gk, A.gp, c->st);
            else
                hipLaunchKernelGGL(k_adj_step, dim3(c->g.nblk), dim3(kBlock), 0, c->st, c->g, q,
                                   Tn1, A.hist[0], c->kappa, c->lx, c->ly, c->lz, c->rhoc, c->cvec,
                                   c->vol, A.g_lo, A.g_new, A.gk, A.gp);
            AHIP(c, hipGetLastError());
            // shift: gbar^n (complete) becomes the next right-hand side
            double *old_hi = A.g_hi;
            A.g_hi = A.g_lo;
            A.g_lo = A.g_new;
            A.g_new = old_hi;
        }
    }
    if (dl_dkappa) AHIP(c, hipMemcpyAsync(dl_dkappa, A.gk, B8, hipMemcpyDefault, c->st));
    if (dl_dpower) AHIP(c, hipMemcpyAsync(dl_dpower, A.gp, B8, hipMemcpyDefault, c->st));
    ATRY(restore_state(c, A.fin, A.fin + N));
    AHIP(c, hipStreamSynchronize(c->st));
    return HF_OK;
}

int adjoint_grad(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t ckpt_every, int64_t nprobe,
                 const int64_t *probe_idx, const double *target, const double *weight,
                 double *loss, double *dl_dkappa, double *dl_dpower) {
    if (nprobe < 1 || !probe_idx || !target)
        return ctx_set_err(c, HF_ERR_ARG, "hf_grad: bad probe arguments");
    const size_t N = c->g.N, B8 = N * 8;
    for (int64_t k = 0; k < nprobe; ++k)
        if (probe_idx[k] < 0 || probe_idx[k] >= (int64_t)N)
            return ctx_set_err(c, HF_ERR_ARG, "hf_grad: probe index out of range");
    ATRY(adjoint_forward(c, itime0, nsteps, ckpt_every));
    AdjointState &A = *c->adj;
    if (A.nprobe_cap < nprobe) {
        hipFree(A.probe_val); hipFree(A.probe_idx); hipFree(A.probe_w); hipFree(A.probe_t);
        A.probe_val = A.probe_w = A.probe_t = nullptr;
        A.probe_idx = nullptr;
        ATRY(ensure(c, &A.probe_val, 1));
        ATRY(ensure(c, &A.probe_w, nprobe));
        ATRY(ensure(c, &A.probe_t, nprobe));
        if (hipMalloc(&A.probe_idx, nprobe * sizeof(int64_t)) != hipSuccess)
            return ctx_set_err(c, HF_ERR_ALLOC, "hf_grad: probe allocation failed");
        A.nprobe_cap = nprobe;
    }
    ATRY(ensure(c, &A.seed, N));
    std::vector<double> w(nprobe, 1.0);
    if (weight) std::copy(weight, weight + nprobe, w.begin());
    AHIP(c, hipMemcpyAsync(A.probe_idx, probe_idx, nprobe * 8, hipMemcpyDefault, c->st));
    AHIP(c, hipMemcpyAsync(A.probe_t, target, nprobe * 8, hipMemcpyDefault, c->st));
    AHIP(c, hipMemcpyAsync(A.probe_w, w.data(), nprobe * 8, hipMemcpyHostToDevice, c->st));
    AHIP(c, hipMemsetAsync(A.seed, 0, B8, c->st));
    hipLaunchKernelGGL(k_adj_seed, dim3(1), dim3(64), 0, c->st, nprobe, A.probe_idx, A.probe_w,
                       A.probe_t, c->T[0], A.seed, A.probe_val);
    AHIP(c, hipGetLastError());
    double L = 0.0;
    AHIP(c, hipMemcpyAsync(&L, A.probe_val, 8, hipMemcpyDeviceToHost, c->st));
    AHIP(c, hipStreamSynchronize(c->st));
    ATRY(adjoint_backward(c, A.seed, dl_dkappa, dl_dpower));
    if (loss) *loss = L;
    return HF_OK;
}

}  // namespace hf
