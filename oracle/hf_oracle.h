/*
 * hf_oracle.h -- CPU restatement of HeatFlow's per-timestep path (TEST INFRASTRUCTURE).
 *
 * This is the oracle: a plain-C, serial, literal restatement of the reference
 * Fortran algorithm, used only by tests/, the smoke check and bench.py's
 * cpu_baseline leg, and only as the checker.  It is never linked into or called by
 * the product library (heatflow_amd/csrc).
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * HeatFlow repository root).  Indices are 1-based in the COO export, exactly as the
 * reference's sprs2_dp (src/heatflow/mod_SPtype.f90:56-61); cell arrays are 0-based,
 * x fastest, linear index I-1 = (ix-1) + nx*(iy-1) + nx*ny*(iz-1).
 */
#ifndef HF_ORACLE_H
#define HF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar inputs: the subset of module `inputs` (src/heatflow/mod_inputs.f90:80-107)
 * and `globe_data` (src/heatflow/mod_global.f90:5-17) that the path reads. */
typedef struct hfo_system {
    int64_t nx, ny, nz;
    double time_step;            /* inputs::time_step                         */
    int64_t icattaneo, isteady;  /* inputs::icattaneo, isteady                */
    int32_t periodic[3];         /* Periodicx/y/z                             */
    double kappa_bound[6];       /* kappaBoundx1, NX, y1, Ny, z1, Nz          */
    double t_bath[6];            /* T_Bathx1, x2, y1, y2, z1, z2              */
    double br;                   /* BR                                        */
    int32_t cg[6];               /* CG_x_m, CG_x_p, CG_y_m, CG_y_p, CG_z_m, CG_z_p */
    double t_bath_cg;            /* T_BathCG                                  */
    double power_in, freq;       /* power_in, freq                            */
    int64_t heated_steps;        /* heattime                                  */
    double t_bath_rad;           /* T_Bath (radiative term)                   */
    /* per-cell arrays (length N), the fields of heatblock
     * (src/heatflow/mod_constructions.f90:25-31) after set_global_variables
     * (src/heatflow/mod_setup.f90:53-68): tau is already scaled by 1/dt^2. */
    const double *kappa, *rho, *heat_capacity, *tau, *em;
    const double *len_x, *len_y, *len_z, *volume;
    const int64_t *iheater;
    /* Superset input (no reference counterpart): extra volumetric power density
     * added to Qdens; NULL for the plain reference path. */
    const double *qdens_extra;
} hfo_system;

/* solver controls, src/heatflow/mod_evolve.f90:142-144 */
typedef struct hfo_solver {
    int64_t itol;     /* 1: ||r||/||b||, 2: ||z||/||b||  */
    double tol;       /* reference: 1e-32                */
    int64_t itmax;    /* reference: 50000                */
} hfo_solver;

/* Error codes of hfo_simulate: the six NaN guards of simulate()
 * (src/heatflow/mod_evolve.f90:70-73, 82-89, 102-105, 124-127, 153-158). */
enum {
    HFO_OK = 0,
    HFO_NAN_B = 1, HFO_NAN_Q = 2, HFO_NAN_QDENS = 3, HFO_NAN_SCAT = 4,
    HFO_NAN_S = 5, HFO_NAN_X = 6
};

/* hmatrixmod::hmatrixfunc(i, j) (src/heatflow/mod_hmatrix.f90:62-248); i, j 1-based. */
double hfo_hmatrixfunc(const hfo_system *s, int64_t i, int64_t j);

/* setup::sparse_Hmatrix (src/heatflow/mod_setup.f90:89-152).  Returns len (the
 * reference's formula, :96-99); *count gets the number of filled entries.  val/irow/
 * jcol must hold hfo_coo_len(s) entries; the tail keeps the sentinels (0,-2,-1). */
int64_t hfo_coo_len(const hfo_system *s);
int64_t hfo_sparse_hmatrix(const hfo_system *s, double *val, int64_t *irow, int64_t *jcol,
                           int64_t *count);

/* boundary_vector::boundary(B) (src/heatflow/mod_boundary.f90:24-152); B zeroed first as
 * simulate() does (mod_evolve.f90:55). */
void hfo_boundary(const hfo_system *s, double *B);

/* Heating::heater(itime, Q, Qdens) (src/heatflow/mod_heating.f90:27-169).
 * Returns heated_volume. */
double hfo_heater(const hfo_system *s, int64_t itime, const double *temp_p, double *Q,
                  double *Qdens);

/* cattaneo::S_catS (src/heatflow/mod_cattaneo.f90:23-42) */
void hfo_scat(const hfo_system *s, const double *temp_p, const double *temp_pp, double *scat);

/* Order-sensitive hash of the COO hfo_sparse_hmatrix would write (positions 0..len-1, the
 * sentinel tail included), computed without storing it; and the hash of one entry (the sum
 * over the entries, mod 2^64, is the COO's hash). */
int64_t hfo_coo_hash(const hfo_system *s, uint64_t *hash, int64_t *count_out);
uint64_t hfo_coo_entry_hash(int64_t k, int64_t i, int64_t j, double v);

/* solver::linbcg (src/heatflow/mod_Sparse.f90:372-434) over a COO matrix in the
 * reference layout (SRSax/SRStx/SRSdiag, :293-366).  x is in/out. */
void hfo_linbcg(int64_t n, int64_t len, const double *val, const int64_t *irow,
                const int64_t *jcol, const double *b, double *x, int64_t itol, double tol,
                int64_t itmax, int64_t *iter, double *err);

/* setup::stability (src/heatflow/mod_setup.f90:158-199) over the cells in loop order, as
 * set_global_variables calls it per cell (:65): returns the first cell index (0-based) whose
 * test fails, or -1; *value gets that cell's failing var_stability.  *boundary_test is 1 when
 * the boundary test (kappaBoundx1 on the x1, y1, z1 faces, :182-194) failed, else 0. */
int64_t hfo_stability(const hfo_system *s, double *value, int32_t *boundary_test);

/* evolution::simulate(itime) (src/heatflow/mod_evolve.f90:45-177) on an assembled COO
 * H; updates temp_p / temp_pp in place (history rotation :169-170).  Returns an HFO_*
 * code.  If iters/err are non-NULL they receive linbcg's iteration count and error. */
int hfo_simulate(const hfo_system *s, const hfo_solver *sv, int64_t len, const double *val,
                 const int64_t *irow, const int64_t *jcol, int64_t itime, double *temp_p,
                 double *temp_pp, int64_t *iters, double *err);

/* Convenience: the full reference run loop (src/heatflow.f90:67-90 with
 * initial_evolve src/heatflow/mod_init_evolve.f90:46-48): assemble once, set
 * T_p = T_pp = t_system unless temp_p/temp_pp already hold a state (init_state=0),
 * then run ntime steps.  Returns HFO_* and the total iteration count. */
int hfo_run(const hfo_system *s, const hfo_solver *sv, int64_t ntime, double t_system,
            int init_state, double *temp_p, double *temp_pp, int64_t *total_iters);

#ifdef __cplusplus
}
#endif
#endif
