/*
 * heatflow_hip.h -- C-ABI of the MI355X (gfx950) HeatFlow time-step library.
 *
 * The library replaces the per-timestep hot path of HeatFlow's Fortran program:
 *   - H-matrix assembly        setup::sparse_Hmatrix  (src/heatflow/mod_setup.f90:89-152)
 *                              hmatrixmod::hmatrixfunc (src/heatflow/mod_hmatrix.f90:62-248)
 *   - right-hand side          boundary_vector::boundary (src/heatflow/mod_boundary.f90:24-152),
 *                              Heating::heater (src/heatflow/mod_heating.f90:27-169),
 *                              cattaneo::S_catS (src/heatflow/mod_cattaneo.f90:23-42),
 *                              S build and x0 (src/heatflow/mod_evolve.f90:112-141)
 *   - linear solve             solver::linbcg (src/heatflow/mod_Sparse.f90:372-434)
 *   - the time step itself     evolution::simulate (src/heatflow/mod_evolve.f90:45-177)
 *
 * Plain C types only.  Every array pointer may be a host pointer or a device pointer
 * (the library copies with hipMemcpyDefault).  Arrays over cells have N = nx*ny*nz entries
 * in the reference's linear order: x fastest, then y, then z
 * (src/heatflow/mod_hmatrix.f90:72-74).  All entry points return 0 on success or an HF_ERR_*
 * code; hf_last_error() gives a message.  One context = one HIP stream on the current device;
 * calls on one context must not overlap.
 */
#ifndef HEATFLOW_HIP_H
#define HEATFLOW_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes.  HF_ERR_NAN_* map one-to-one onto the six fatal NaN guards of simulate()
 * (src/heatflow/mod_evolve.f90:70-73, 82-85, 86-89, 102-105, 124-127, 153-158); a Fortran
 * caller prints the reference message and stops with status 1. */
enum {
    HF_OK = 0,
    HF_ERR_NAN_B = 1,
    HF_ERR_NAN_Q = 2,
    HF_ERR_NAN_QDENS = 3,
    HF_ERR_NAN_SCAT = 4,
    HF_ERR_NAN_S = 5,
    HF_ERR_NAN_X = 6,
    HF_ERR_ARG = 10,        /* invalid argument or unsupported geometry               */
    HF_ERR_ALLOC = 11,      /* device allocation failed                               */
    HF_ERR_HIP = 12,        /* HIP runtime error                                      */
    HF_ERR_STATE = 13,      /* call order: e.g. hf_step before hf_assemble            */
    HF_ERR_NO_DEVICE = 14,  /* no HIP device                                          */
    HF_ERR_UNSTABLE = 15    /* hf_check_stability: the stability condition fails
                               (setup::stability prints and stops, mod_setup.f90:175-197) */
};

/* Solver precision modes */
enum {
    HF_PREC_FP64 = 0,       /* FP64 throughout (the drop-in)                          */
    HF_PREC_FP32_IR = 1     /* FP32 inner CG (FP64 dot accumulation) inside FP64
                               iterative refinement; history, RHS and residual in FP64 */
};

/* Scalar run parameters: the members of module `inputs` that the path reads
 * (src/heatflow/mod_inputs.f90:80-107, keys parsed at :283-327), plus solver controls.
 * The face order of the 6-vectors is x1, xN, y1, yN, z1, zN. */
typedef struct hf_params {
    int64_t nx, ny, nz;
    double time_step;          /* time_step                                             */
    int32_t icattaneo;         /* icattaneo (0 Fourier, 1 Cattaneo)                     */
    int32_t isteady;           /* isteady                                               */
    int32_t periodic[3];       /* Periodicx/y/z                                         */
    int32_t cg[6];             /* CG_x_m, CG_x_p, CG_y_m, CG_y_p, CG_z_m, CG_z_p        */
    double kappa_bound[6];     /* kappaBoundx1, kappaBoundNx, ..y1, ..Ny, ..z1, ..Nz    */
    double t_bath[6];          /* T_Bathx1, T_Bathx2, T_Bathy1, T_Bathy2, T_Bathz1, z2 */
    double br;                 /* BR                                                    */
    double t_bath_cg;          /* T_BathCG                                              */
    double power_in;           /* power_in                                              */
    double freq;               /* freq                                                  */
    double t_bath_rad;         /* T_Bath (radiative term, mod_heating.f90:121-123)      */
    int64_t heated_steps;      /* heattime                                              */
    /* solver controls (the reference hard-codes itol=1, tol=1e-32, itmax=50000 at
     * src/heatflow/mod_evolve.f90:142-144) */
    double tol;                /* relative residual tolerance of the (outer) solve      */
    int64_t itmax;             /* iteration cap of the (outer) solve; in mixed mode the
                                  refinement sweeps are capped at min(itmax, 64) and stop
                                  when a sweep no longer halves the FP64 residual        */
    int32_t precision;         /* HF_PREC_*                                             */
    int32_t inner_itmax;       /* FP32 inner CG iteration cap per refinement sweep      */
    double inner_tol;          /* FP32 inner CG relative tolerance per sweep            */
} hf_params;

typedef struct hf_ctx hf_ctx;

/* Fills *p with the reference defaults (src/heatflow/mod_inputs.f90:213-270) and the
 * reference's hard-coded solver controls. */
void hf_params_default(hf_params *p);

/* Creates a context for one grid; allocates device memory for N cells.
 * Replaces the allocations of set_global_variables (src/heatflow/mod_setup.f90:38-45).
 * Geometry the reference itself mishandles is rejected with HF_ERR_ARG: nx == 1 with N > 1
 * (duplicate COO offsets overflow `len`, mod_setup.f90:96-112) and periodic axes shorter than 3
 * (length 1 zeroes the diagonal in hmatrixfunc's wrap test, mod_hmatrix.f90:213-227; length 2
 * duplicates a COO offset). */
int hf_create(hf_ctx **ctx, const hf_params *p);

/* Uploads the per-cell material fields of grid(:,:,:) (src/heatflow/mod_constructions.f90:25-31)
 * as structure-of-arrays, after set_global_variables has filled them
 * (src/heatflow/mod_setup.f90:53-68): tau_scaled is grid%tau = tau/dt^2.  len_x/y/z are
 * grid%Length(1:3) (mod_inputs.f90:619-622), volume is grid%volume.  iheater may be NULL
 * (no heaters), em may be NULL (zero).  qdens_extra (may be NULL) is a superset input with no
 * reference counterpart: a per-cell volumetric power density in W/m^3 added to Qdens. */
int hf_upload_cells(hf_ctx *ctx, const double *kappa, const double *rho,
                    const double *heat_capacity, const double *tau_scaled, const double *em,
                    const double *len_x, const double *len_y, const double *len_z,
                    const double *volume, const int64_t *iheater, const double *qdens_extra);

/* One material of material()'s table (src/heatflow/mod_material.f90:54-267): the fields the
 * time step uses.  File materials (index 1..99) are what the caller read from mat.in
 * (mod_inputs.f90:655-720); the built-in ones (index > 99) are the library's. */
typedef struct hf_material {
    int64_t index;
    double kappa, rho, heat_capacity;
    double tau;                /* physical relaxation time (s)                          */
    double em;                 /* emissivity                                            */
} hf_material;

/* The library's copy of built-in material `index` (> 99; mod_material.f90:116-265): *m as
 * material() sets it (its single-precision literals, widened), and whether material() assigns
 * tau and em for it (*sets_tau, *sets_em; otherwise the value of the previous cell in the
 * setup loop stays).  HF_ERR_ARG for an unknown index.  Needs no device. */
int hf_builtin_material(int64_t index, hf_material *m, int32_t *sets_tau, int32_t *sets_em);

/* Setup on the device from material ids: the per-cell fields that set_global_variables
 * derives with material() (src/heatflow/mod_setup.f90:34-68) -- kappa, rho, heat_capacity,
 * grid%tau = tau / dt^2, em -- then as hf_upload_cells.  imaterial_type (N ids) and iheater
 * (N flags, may be NULL) are grid%imaterial_type and grid%iheater of read_system, host or
 * device memory; len_x/y/z are its uniform cell lengths Lx/nx, Ly/ny, Lz/nz and the volume
 * their product (mod_inputs.f90:619-622).  Where a built-in material does not assign tau or em,
 * a cell takes the value of the last earlier cell (loop order) whose material did, 0 before
 * the first (the reference leaves that local uninitialised).  An unknown id is HF_ERR_ARG
 * (material() prints and exits, mod_material.f90:105-107, 262-264). */
int hf_upload_materials(hf_ctx *ctx, const int64_t *imaterial_type, const int64_t *iheater,
                        int32_t nfile, const hf_material *file_materials, double len_x,
                        double len_y, double len_z);

/* setup::stability over the uploaded cells (src/heatflow/mod_setup.f90:158-199, run when
 * _Check_Stability): dt kappa/(rho c_p) (1/L1^2 + 1/L2^2 + 1/L3^2) must not exceed 1/12 (the
 * reference's single-precision constant), nor the same with kappaBoundx1 on cells of the x1,
 * y1 or z1 faces.  HF_OK, or HF_ERR_UNSTABLE with *cell the first failing cell in loop order
 * (-1 if none) and *value its failing value (either may be NULL). */
int hf_check_stability(hf_ctx *ctx, int64_t *cell, double *value);

/* Assembles H on the device as a 7-point stencil (diagonal + 3 forward couplings) and the
 * boundary vector B.  Replaces setup::sparse_Hmatrix (src/heatflow/mod_setup.f90:89-152).
 * Values follow hmatrixfunc's operand order (src/heatflow/mod_hmatrix.f90:62-391). */
int hf_assemble(hf_ctx *ctx);

/* Exports the assembled H in the reference's COO layout, globe_data::ra
 * (src/heatflow/mod_global.f90:14, type sprs2_dp src/heatflow/mod_SPtype.f90:56-61):
 * 1-based (val, irow, jcol) in sparse_Hmatrix's emission order, `len` from the reference
 * formula (mod_setup.f90:96-99), entries past `count` hold the sentinel (0, -2, -1)
 * (mod_setup.f90:104-106).  With val/irow/jcol NULL only n/len/count are returned.
 * Arrays must hold `len` entries. */
int hf_export_coo(hf_ctx *ctx, int64_t *n, int64_t *len, int64_t *count, double *val,
                  int64_t *irow, int64_t *jcol);

/* hf_export_coo for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference.
 * Same outputs as hf_export_coo, but val, irow and jcol must be device memory (`len` entries
 * each; all three, or all NULL for n/len/count only) and the call allocates nothing: its scan
 * scratch is a fixed 2 MiB device array that each GPU holds for the life of the process,
 * which covers grids of up to 2^30 cells (HF_ERR_ARG beyond).  Calls that overlap in time on
 * one GPU share that array and race. */
int hf_export_coo_noalloc(hf_ctx *ctx, int64_t *n, int64_t *len, int64_t *count, double *val,
                          int64_t *irow, int64_t *jcol);

/* Sets / reads globe_data::Temp_p and Temp_pp (src/heatflow/mod_global.f90:10), i.e. T^n and
 * T^{n-1}.  hf_set_state replaces initial_evolve's assignment
 * (src/heatflow/mod_init_evolve.f90:34-51). */
int hf_set_state(hf_ctx *ctx, const double *temp_p, const double *temp_pp);
int hf_get_state(hf_ctx *ctx, double *temp_p, double *temp_pp);

/* One time step: evolution::simulate(itime) (src/heatflow/mod_evolve.f90:45-177).  Builds S,
 * the initial guess, solves H x = S to `tol`, rotates the history.  *iter gets the solver's
 * iteration count (outer FP64 + inner FP32 iterations in mixed mode) and *err its final
 * relative residual (both may be NULL).  NaN guards return HF_ERR_NAN_*. */
int hf_step(hf_ctx *ctx, int64_t itime, int64_t *iter, double *err);

/* nsteps consecutive time steps itime0 .. itime0+nsteps-1 without host copies of the state
 * (the reference driver loop, src/heatflow.f90:67-90).  *iters_total may be NULL. */
int hf_run(hf_ctx *ctx, int64_t itime0, int64_t nsteps, int64_t *iters_total);

/* linbcg-compatible solve of H x = b on the assembled H (src/heatflow/mod_Sparse.f90:372-434):
 * x is the initial guess on entry and the solution on exit.  The stop rule is linbcg's,
 * checked after every iteration on the recursively updated residual r: itol = 1 stops when
 * ||r|| / ||b|| <= tol, itol = 2 when ||r / diag H|| / ||b|| <= tol (:395, 418-425); other itol
 * values are HF_ERR_ARG.  The iterates are Jacobi-preconditioned CG's (linbcg's BiCG on the
 * symmetric H).  *iter and *err as in the reference (unused by its caller); an FP64 context. */
int hf_solve(hf_ctx *ctx, const double *b, double *x, int32_t itol, double tol, int64_t itmax,
             int64_t *iter, double *err);

/* Reads intermediate vectors of the last hf_step for parity checks:
 * which = 0: B (boundary vector, mod_boundary.f90:24-152)
 *         1: Qdens of the last step (mod_heating.f90:164-167)
 *         2: S of the last step (mod_evolve.f90:112-119)
 *         3: diagonal of H, 4/5/6: forward couplings in x/y/z (0 where absent)
 * and the uploaded per-cell fields: 7 kappa, 8 lin_rhoc = rho c_p, 9 grid%tau = tau/dt^2,
 * 10 em L1 L2 sigma (the radiative coefficient), 11 volume. */
int hf_get_vector(hf_ctx *ctx, int32_t which, double *out);

/* Debug switches (no reference counterpart): bit 0 keeps S of every step for hf_get_vector(2)
 * (costs one extra N-vector write per step; both precision modes then run the unfused
 * step kernels); bit 1 turns off the first-iteration stop of CG (a solve whose first
 * iteration already meets the tolerance ends there, with the next residual norm predicted
 * from p.q and q.q), and with it the reuse of its step length; bit 2 turns off that reuse
 * alone (mixed mode: a step after one whose inner solve ended on its first iteration takes
 * y = alpha r^0 with that solve's alpha, and solves only if the result misses the inner
 * target); bit 3 keeps per-cell reads of rho c_p and tau' rho c_p where both are uniform,
 * and of the cell volume in the adjoint (found at upload; the kernels then read one value for
 * every cell); bit 4 runs the adjoint step with the flat per-cell kernel instead of the
 * z-marching one (grids with a periodic axis always use the flat kernel); bit 5 starts every
 * adjoint solve from the linear extrapolation 2 lambda^{n+2} - lambda^{n+3} instead of the
 * polynomial through the last six solutions; bit 6 runs the FP64 steps of a small grid with the
 * multi-kernel path instead of the persistent one (FP64 contexts of at most 8192 cells run
 * whole steps, up to 256 per launch, in one workgroup: no launch per CG iteration and one host
 * synchronisation per launch; its sums group differently, so the two paths agree to the
 * tolerance, not bit for bit); bit 7 turns off the two-step pass of mixed mode (where a step
 * would reuse the inner step length, or its own inner solve ends on its first iteration, one
 * z-march does it and the next step, reading the operator, Qdens + B and the state once; a step
 * whose test misses is redone alone); bit 8 keeps the reused step length fixed instead of
 * updating it after each reusing step (then the two-step pass and single steps give the same
 * bits); bit 9 fails the first step's test in every two-step pass, so that each pair falls back
 * (a test aid); bit 10 makes the every-state checkpoint allocation of spacing 0 fail, so that it
 * falls back to ceil(sqrt(nsteps)) checkpoints (a test aid). */
int hf_set_debug(hf_ctx *ctx, int32_t flags);

/* Timing / introspection for the benchmark: achieved counters since the last reset. */
typedef struct hf_stats {
    int64_t steps;            /* time steps taken                                      */
    int64_t outer_iters;      /* FP64 CG iterations (or refinement sweeps in mixed mode) */
    int64_t inner_iters;      /* FP32 inner CG iterations (mixed mode)                 */
    double last_err;          /* final relative residual of the last solve             */
    int64_t first_iter_stops; /* CG solves that ended on their first iteration          */
    int64_t tiny_misses;      /* mixed-mode steps whose successor's TINY test was guessed
                                 wrong (that successor started from the state)          */
    int64_t residual_reruns;  /* mixed-mode passes run again to store the residual a
                                 further refinement sweep needed                        */
    int64_t alpha_reuses;     /* mixed-mode steps whose first sweep reused the last inner
                                 solve's step length instead of solving                 */
    int64_t reuse_misses;     /* of those, steps whose residual missed the inner target
                                 (ordinary sweeps followed)                             */
    int64_t adjoint_solves;   /* adjoint solves (hf_adjoint, hf_grad)                   */
    int64_t adjoint_iters;    /* their CG iterations (FP32 inner ones in mixed mode)    */
} hf_stats;
int hf_get_stats(hf_ctx *ctx, hf_stats *st);
int hf_reset_stats(hf_ctx *ctx);

/* Kernel timing (no reference counterpart; for the benchmark).  While enabled, each kernel
 * launch of a time step is bracketed by a pair of HIP events on the context's stream.
 * hf_kernel_times fills, per kernel id (HF_K_*), the summed device time in ms, the number of
 * launches that did work (CG launches queued past convergence exit at once and are not
 * counted) and the algorithmic bytes those launches move (compulsory reads + writes of the
 * arrays each kernel touches, every array element counted once).  Reading resets the pool. */
enum {
    HF_K_QDENS = 0, HF_K_RHS64, HF_K_RHS32, HF_K_CG_A64, HF_K_CG_B64, HF_K_CG_A32, HF_K_CG_B32,
    HF_K_FINAL, HF_K_IR_RES, HF_K_IR_STEP, HF_K_STEP64, HF_K_PERSIST, HF_K_IR_STEP2, HF_K_COUNT
};
int hf_set_timing(hf_ctx *ctx, int32_t on);
int hf_kernel_times(hf_ctx *ctx, double *ms, int64_t *launches, double *bytes);

/* Changes the solver controls of an existing context (tol, itmax, precision, inner_*). */
int hf_set_solver(hf_ctx *ctx, double tol, int64_t itmax, int32_t precision, double inner_tol,
                  int32_t inner_itmax);

/* ---- adjoint (no reference counterpart: reverse-time gradients) ---------------------------
 * hf_forward_checkpointed runs nsteps time steps (itime0 .. itime0+nsteps-1) from the current
 * state, keeping (T^n, T^{n-1}) on the device every `ckpt_every` steps; the context ends at the
 * final state.  ckpt_every = 1 keeps every state once instead (nsteps + 1 states of N doubles)
 * and nothing is recomputed; ckpt_every = 0 takes 1 when those states fit in the device memory
 * free at the call (leaving max(2 GiB, 5% of the device) spare), else ceil(sqrt(nsteps)).
 * hf_adjoint then runs the reverse sweep (each segment recomputed from its
 * checkpoint) seeded with dL/dT^{final} (length N), and returns the gradients with respect to
 * the per-cell conductivity kappa (dl_dkappa, length N) and to a per-cell power P_I that enters
 * as Qdens_I += P_I / volume_I (dl_dpower, length N).  The adjoint solves use the context's
 * precision mode (FP64, or FP32 inner CG inside FP64 refinement) and tolerance (or the one
 * hf_set_adjoint_tol gives).  The temperature-dependent radiative source of heated systems
 * (src/heatflow/mod_heating.f90:121-123) enters through its derivative
 * 4 em L1 L2 sigma (T^n)^3 / V_heated; heater modes 2-6 depend on time only.  Requires
 * per-axis uniform cell lengths (HF_ERR_ARG otherwise); the BR switch (x BR where neighbouring
 * kappas differ) is treated as locally constant.  hf_upload_cells, hf_assemble, hf_set_fields
 * and hf_set_solver discard the checkpoints (hf_adjoint then returns HF_ERR_STATE).  The context
 * is left at the end state of the forward run. */
int hf_forward_checkpointed(hf_ctx *ctx, int64_t itime0, int64_t nsteps, int64_t ckpt_every);
int hf_adjoint(hf_ctx *ctx, const double *dl_dt_final, double *dl_dkappa, double *dl_dpower);

/* Scaled-residual tolerance of the adjoint solves, apart from the forward's (0: the forward's,
 * the default).  Their right-hand side carries no offset like the forward's 300 K, so the same
 * relative tolerance asks more of them; the gradient's error is set by the forward solves down
 * to a much looser adjoint tolerance (at 384^3, forward 1e-13: unchanged up to 1e-10). */
int hf_set_adjoint_tol(hf_ctx *ctx, double tol);

/* Convenience: forward + reverse for L = sum_k w_k (T^{final}[probe_k] - target_k)^2
 * (weight may be NULL for w = 1); *loss gets L. */
int hf_grad(hf_ctx *ctx, int64_t itime0, int64_t nsteps, int64_t ckpt_every, int64_t nprobe,
            const int64_t *probe_idx, const double *target, const double *weight,
            double *loss, double *dl_dkappa, double *dl_dpower);

/* Replaces the per-cell conductivity and/or the superset per-cell power density (either may be
 * NULL) of an uploaded context; call hf_assemble afterwards.  Used by the autograd wrapper. */
int hf_set_fields(hf_ctx *ctx, const double *kappa, const double *qdens_extra);

/* Last error message of this context (or of the last failed hf_create when ctx is NULL). */
const char *hf_last_error(const hf_ctx *ctx);

/* Releases the context and its device memory. */
void hf_destroy(hf_ctx *ctx);

/* Library version string. */
const char *hf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HEATFLOW_HIP_H */
