/*
 * hf_oracle.c -- serial CPU restatement of HeatFlow's per-timestep path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle that tests/, the smoke check and the CPU
 * baseline leg of bench.py compare against.  The product (heatflow_amd) never links
 * or calls it.
 *
 * The arithmetic follows the reference's operand order statement by statement
 * (Fortran evaluates a*b*c as (a*b)*c and a+b+c as (a+b)+c); it is compiled without
 * FMA contraction (-ffp-contract=off) so that it rounds like the reference build.
 * Each function cites the reference file:line it restates.
 */
#include "hf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1.0e-12                   /* src/heatflow/mod_constants.f90:17 */
#define STEFBOLTZ 5.670373e-8          /* src/heatflow/mod_constants.f90:16 */
#define PI_REF 3.141592653589793       /* src/heatflow/mod_constants.f90:12 */

/* 1-based (x,y,z) -> 0-based linear cell index */
static inline int64_t cell(const hfo_system *s, int64_t x, int64_t y, int64_t z) {
    return (x - 1) + s->nx * ((y - 1) + s->ny * (z - 1));
}

/* Fortran MOD for positive operands; altmod: src/heatflow/mod_hmatrix.f90:330-338 */
static inline int64_t altmod(int64_t a, int64_t b) {
    int64_t c = a % b;
    if (c == 0) c = b;
    return c;
}

/* calculate_alpha, src/heatflow/mod_hmatrix.f90:254-267 */
static double calc_alpha(const hfo_system *s, int64_t x, int64_t y, int64_t z, int64_t i) {
    double alpha = 0.0;
    int64_t c = cell(s, x, y, z);
    double tau = s->tau[c];
    double inverse_time = 1.0 / s->time_step;           /* mod_setup.f90:45 */
    double lin_rhoc = s->rho[i - 1] * s->heat_capacity[i - 1]; /* mod_setup.f90:64 */
    if (s->isteady == 0) {
        if (s->icattaneo == 0) tau = 0.0;
        alpha = (tau * lin_rhoc) + (inverse_time * lin_rhoc);
    } else {
        alpha = 0.0;
    }
    return alpha;
}

/* boundry_diag_term, src/heatflow/mod_hmatrix.f90:347-391 */
static double boundary_diag_term(const hfo_system *s, int64_t xb, int64_t yb, int64_t zb,
                                 int64_t x, int64_t y, int64_t z, double kappa_ab) {
    int64_t c = cell(s, x, y, z);
    double kappa = s->kappa[c];
    const double *kb = s->kappa_bound;
    if (xb != x) {
        double L = s->len_x[c];
        if (xb < 1) {
            kappa_ab = ((2.0 * kb[0] * kappa) / (kb[0] + kappa)) / (L * L);
            if (kappa != kb[0]) kappa_ab = kappa_ab * s->br;
        } else if (xb > s->nx) {
            kappa_ab = ((2.0 * kb[1] * kappa) / (kb[1] + kappa)) / (L * L);
            if (kappa != kb[1]) kappa_ab = kappa_ab * s->br;
        }
    } else if (yb != y) {
        double L = s->len_y[c];
        if (yb < 1) {
            kappa_ab = ((2.0 * kb[2] * kappa) / (kb[2] + kappa)) / (L * L);
            if (kappa != kb[2]) kappa_ab = kappa_ab * s->br;
        } else if (yb > s->ny) {
            kappa_ab = ((2.0 * kb[3] * kappa) / (kb[3] + kappa)) / (L * L);
            if (kappa != kb[3]) kappa_ab = kappa_ab * s->br;
        }
    } else if (zb != z) {
        double L = s->len_z[c];
        if (zb < 1) {
            kappa_ab = ((2.0 * kb[4] * kappa) / (kb[4] + kappa)) / (L * L);
            if (kappa != kb[4]) kappa_ab = kappa_ab * s->br;
        } else if (zb > s->nz) {
            kappa_ab = ((2.0 * kb[5] * kappa) / (kb[5] + kappa)) / (L * L);
            if (kappa != kb[5]) kappa_ab = kappa_ab * s->br;
        }
    }
    return kappa_ab;
}

/* calculate_conductivity, src/heatflow/mod_hmatrix.f90:273-321 */
static double calc_conductivity(const hfo_system *s, int64_t xi, int64_t yi, int64_t zi,
                                int64_t xo, int64_t yo, int64_t zo) {
    double kappa_ab = 0.0;
    if (xi >= 1 && xi <= s->nx && yi >= 1 && yi <= s->ny && zi >= 1 && zi <= s->nz) {
        int64_t ci = cell(s, xi, yi, zi), co = cell(s, xo, yo, zo);
        double kin = s->kappa[ci], kout = s->kappa[co];
        const double *Lin, *Lout;
        int axis_set = 1;
        if (xi != xo) {
            Lin = s->len_x; Lout = s->len_x;
        } else if (yi != yo) {
            Lin = s->len_y; Lout = s->len_y;
        } else if (zi != zo) {
            Lin = s->len_z; Lout = s->len_z;
        } else {
            axis_set = 0; Lin = Lout = NULL;
        }
        if (axis_set) {
            double li = Lin[ci], lo = Lout[co];
            kappa_ab = (li + lo) * kin * kout / (li * kout + lo * kin);
            kappa_ab = kappa_ab / (lo * lo);
        }
        if (kin != kout) kappa_ab = kappa_ab * s->br;
        return kappa_ab;
    }
    return boundary_diag_term(s, xi, yi, zi, xo, yo, zo, kappa_ab);
}

/* hmatrixfunc, src/heatflow/mod_hmatrix.f90:62-248 (the if-chain order is kept:
 * later matches override earlier ones, as in the reference) */
double hfo_hmatrixfunc(const hfo_system *s, int64_t i, int64_t j) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    int64_t x = altmod(i, nx);
    int64_t y = ((i - altmod(i, nx)) / nx) % ny + 1;
    int64_t z = (i - altmod(i, nx * ny)) / (nx * ny) + 1;
    double alpha = calc_alpha(s, x, y, z, i);
    int64_t xp = x + 1, xm = x - 1, yp = y + 1, ym = y - 1, zp = z + 1, zm = z - 1;
    if (s->periodic[0]) { if (xp > nx) xp = 1; if (xm < 1) xm = nx; }
    if (s->periodic[1]) { if (yp > ny) yp = 1; if (ym < 1) ym = ny; }
    if (s->periodic[2]) { if (zp > nz) zp = 1; if (zm < 1) zm = nz; }

    double A = calc_conductivity(s, xm, y, z, x, y, z);
    double B = calc_conductivity(s, xp, y, z, x, y, z);
    double D = calc_conductivity(s, x, ym, z, x, y, z);
    double E = calc_conductivity(s, x, yp, z, x, y, z);
    double F = calc_conductivity(s, x, y, zm, x, y, z);
    double G = calc_conductivity(s, x, y, zp, x, y, z);

    int64_t d = i - j;
    double H = 0.0;
    if (d == 0) H = -(A + B + D + E + F + G) - alpha;
    if (d == 1) H = (x == 1) ? 0.0 : A;
    if (d == -1) H = (x == nx) ? 0.0 : B;
    if (d == nx) H = (y == 1) ? 0.0 : D;
    if (d == -nx) H = (y == ny) ? 0.0 : E;
    if (d == nx * ny) H = (z == 1) ? 0.0 : F;
    if (d == -(nx * ny)) H = (z == nz) ? 0.0 : G;
    if (s->periodic[0]) {
        if (d == -(nx - 1) && x == 1) H = A;
        if (d == (nx - 1) && x == nx) H = B;
    }
    if (s->periodic[1]) {
        if (d == -(ny - 1) * nx && y == 1) H = D;
        if (d == (ny - 1) * nx && y == ny) H = E;
    }
    if (s->periodic[2]) {
        if (d == -(nz - 1) * nx * ny && z == 1) H = F;
        if (d == (nz - 1) * nx * ny && z == nz) H = G;
    }
    return H;
}

/* len formula, src/heatflow/mod_setup.f90:96-99 */
int64_t hfo_coo_len(const hfo_system *s) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    int64_t len = 7 * nx * ny * nz - 2 * (nx * ny + ny * nz + nz * nx);
    if (s->periodic[0]) len += 2 * ny * nz;
    if (s->periodic[1]) len += 2 * nz * nx;
    if (s->periodic[2]) len += 2 * nx * ny;
    return len;
}

/* Where sparse_Hmatrix's entries go: the COO arrays, or (arrays NULL) an order-sensitive hash
 * of (position, irow, jcol, value bits) summed over the entries, sentinels included.  The mix is
 * the splitmix64 output function (S. Vigna, public domain). */
typedef struct {
    double *val;
    int64_t *irow, *jcol;
    uint64_t hash;
} coo_sink;

static uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t hfo_coo_entry_hash(int64_t k, int64_t i, int64_t j, double v) {
    uint64_t vb;
    memcpy(&vb, &v, sizeof vb);
    return mix64(mix64(mix64(mix64((uint64_t)k) + (uint64_t)i) + (uint64_t)j) + vb);
}

static void put(coo_sink *o, int64_t k, double v, int64_t i, int64_t j) {
    if (o->val) { o->val[k] = v; o->irow[k] = i; o->jcol[k] = j; }
    else o->hash += hfo_coo_entry_hash(k, i, j, v);
}

static int64_t emit_coo(const hfo_system *s, coo_sink *o, int64_t *count_out);

/* sparse_Hmatrix, src/heatflow/mod_setup.f90:89-152 */
int64_t hfo_sparse_hmatrix(const hfo_system *s, double *val, int64_t *irow, int64_t *jcol,
                           int64_t *count_out) {
    coo_sink o = {val, irow, jcol, 0};
    return emit_coo(s, &o, count_out);
}

/* The hash of the COO that hfo_sparse_hmatrix would write (entries past `count` are the
 * sentinels (0, -2, -1)), without storing it: for grids whose COO does not fit in memory. */
int64_t hfo_coo_hash(const hfo_system *s, uint64_t *hash, int64_t *count_out) {
    coo_sink o = {NULL, NULL, NULL, 0};
    int64_t len = emit_coo(s, &o, count_out);
    *hash = o.hash;
    return len;
}

static int64_t emit_coo(const hfo_system *s, coo_sink *o, int64_t *count_out) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz, NA = nx * ny * nz;
    int64_t len = hfo_coo_len(s);
    if (o->val)                                         /* :104-106 */
        for (int64_t k = 0; k < len; ++k) { o->val[k] = 0.0; o->irow[k] = -2; o->jcol[k] = -1; }
    int64_t addit[6], na = 0;                          /* :107-112 */
    addit[na++] = 1;
    if (s->periodic[0]) addit[na++] = nx - 1;
    if (ny > 1) addit[na++] = nx;
    if (s->periodic[1] && ny > 1) addit[na++] = (ny - 1) * nx;
    if (nz > 1) addit[na++] = nx * ny;
    if (s->periodic[2] && nz > 1) addit[na++] = (nz - 1) * ny * nx;

    int64_t count = 0;
    for (int64_t j = 1; j <= NA; ++j) {                /* parent_loop :121-148 */
        int64_t i = j;
        double H0 = hfo_hmatrixfunc(s, i, j);
        /* the reference writes past `len` only if the formula under-counts; guard it */
        if (count < len) put(o, count, H0, i, j);
        count++;
        for (int64_t k = 0; k < na; ++k) {
            i = j + addit[k];
            if (i > NA) break;                          /* cycle parent_loop */
            H0 = hfo_hmatrixfunc(s, i, j);
            if (fabs(H0) < TINY) continue;              /* cycle neighbour_loop */
            if (count < len) put(o, count, H0, i, j);
            count++;
            if (count < len) put(o, count, H0, j, i);
            count++;
        }
    }
    if (!o->val)
        for (int64_t k = count; k < len; ++k) put(o, k, 0.0, -2, -1);
    *count_out = count;
    return len;
}

/* boundary, src/heatflow/mod_boundary.f90:24-152 */
void hfo_boundary(const hfo_system *s, double *B) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    const double *kb = s->kappa_bound, *tb = s->t_bath;
    double pd[6] = {0, 0, 0, 0, 0, 0};                 /* :43-66 power densities */
    double vol111 = s->volume[0];
    if (s->cg[5]) pd[5] = s->t_bath_cg / (s->volume[cell(s, 1, 1, nz)] * (double)nx * (double)ny);
    if (s->cg[4]) pd[4] = s->t_bath_cg / (vol111 * (double)nx * (double)ny);
    if (s->cg[2]) pd[2] = s->t_bath_cg / (vol111 * (double)nx * (double)nz);
    if (s->cg[3]) pd[3] = s->t_bath_cg / (s->volume[cell(s, 1, ny, 1)] * (double)nx * (double)nz);
    if (s->cg[0]) pd[0] = s->t_bath_cg / (vol111 * (double)ny * (double)nz);
    if (s->cg[1]) pd[1] = s->t_bath_cg / (s->volume[cell(s, nx, 1, 1)] * (double)ny * (double)nz);

    int64_t I = 0;
    for (int64_t iz = 1; iz <= nz; ++iz)
        for (int64_t iy = 1; iy <= ny; ++iy)
            for (int64_t ix = 1; ix <= nx; ++ix, ++I) {
                double kappa = s->kappa[I];
                B[I] = 0.0;                              /* simulate: B = 0 (:55) */
                for (int ax = 0; ax < 3; ++ax) {
                    if (s->periodic[ax]) continue;
                    int64_t pos = ax == 0 ? ix : (ax == 1 ? iy : iz);
                    int64_t n = ax == 0 ? nx : (ax == 1 ? ny : nz);
                    const double *L = ax == 0 ? s->len_x : (ax == 1 ? s->len_y : s->len_z);
                    for (int side = 0; side < 2; ++side) {
                        int f = 2 * ax + side;
                        if (pos != (side == 0 ? 1 : n)) continue;
                        if (s->cg[f]) {
                            B[I] = pd[f];
                        } else {
                            double kh = (2.0 * kappa * kb[f] / (kappa + kb[f])) / (L[I] * L[I]);
                            if (kappa != kb[f]) kh = kh * s->br;
                            B[I] = B[I] + (kh) * tb[f];
                        }
                    }
                }
            }
}

/* heater, src/heatflow/mod_heating.f90:27-169 */
double hfo_heater(const hfo_system *s, int64_t itime, const double *temp_p, double *Q,
                  double *Qdens) {
    const int64_t N = s->nx * s->ny * s->nz;
    double POWER = s->power_in;
    double time = s->time_step * (double)itime;
    double time_pulse = (double)s->heated_steps * s->time_step;
    double heated_volume = 0.0;
    double tb4 = pow(s->t_bath_rad, 4.0);
    for (int64_t IA = 0; IA < N; ++IA) {
        double q = 0.0;
        double tau = s->tau[IA] * pow(s->time_step, 2.0);   /* :56 */
        switch (s->iheater[IA]) {
            case 0: q = 0.0; break;
            case 1: q = POWER; break;
            case 2: q = (time <= time_pulse) ? POWER : 0.0; break;
            case 3: {
                double x = time * 2.0 * PI_REF * s->freq;
                double x2 = s->time_step * 2.0 * PI_REF * s->freq;
                q = POWER * 0.5 * ((x2) - sin(x + x2) * cos(x + x2) + sin(x) * cos(x)) / x2;
                break;
            }
            case 4: q = POWER * pow(sin(time * 2.0 * PI_REF * s->freq), 2.0); break;
            case 5:
                q = POWER * pow(sin(time * 2.0 * PI_REF * s->freq), 2.0)
                    + POWER * 2.0 * PI_REF * s->freq * tau
                          * sin(2.0 * time * 2.0 * PI_REF * s->freq);
                break;
            case 6: q = (itime == 1) ? POWER : 0.0; break;
            default: q = 0.0; break;   /* select case with no match leaves Q(IA)=0 */
        }
        q = q - s->em[IA] * s->len_x[IA] * s->len_y[IA] * STEFBOLTZ
                    * (pow(temp_p[IA], 4.0) - tb4);     /* :121-123 */
        Q[IA] = q;
        if (s->iheater[IA] > 0) heated_volume = heated_volume + s->volume[IA];
    }
    for (int64_t IA = 0; IA < N; ++IA) Qdens[IA] = 0.0;
    if (heated_volume > 0.0)                            /* :164-167 */
        for (int64_t IA = 0; IA < N; ++IA) Qdens[IA] = Q[IA] / heated_volume;
    return heated_volume;
}

/* S_catS, src/heatflow/mod_cattaneo.f90:23-42 */
void hfo_scat(const hfo_system *s, const double *temp_p, const double *temp_pp, double *scat) {
    const int64_t N = s->nx * s->ny * s->nz;
    for (int64_t I = 0; I < N; ++I) {
        double lin_rhoc = s->rho[I] * s->heat_capacity[I];
        scat[I] = (temp_pp[I] - 2.0 * temp_p[I]) * (s->tau[I] * lin_rhoc);
    }
}

/* SRSax, src/heatflow/mod_Sparse.f90:293-315 (serial order) */
static void srsax(int64_t n, int64_t len, const double *val, const int64_t *irow,
                  const int64_t *jcol, const double *x, double *b) {
    for (int64_t i = 0; i < n; ++i) b[i] = 0.0;
    for (int64_t j = 0; j < len; ++j) {
        int64_t i = irow[j];
        if (i > 0 && i <= n) b[i - 1] = b[i - 1] + val[j] * x[jcol[j] - 1];
    }
}

/* SRStx, src/heatflow/mod_Sparse.f90:317-337 (serial order) */
static void srstx(int64_t n, int64_t len, const double *val, const int64_t *irow,
                  const int64_t *jcol, const double *x, double *b) {
    for (int64_t i = 0; i < n; ++i) b[i] = 0.0;
    for (int64_t i = 0; i < len; ++i) {
        int64_t j = jcol[i];
        if (j > 0 && j <= n) b[j - 1] = b[j - 1] + val[i] * x[irow[i] - 1];
    }
}

/* asolve -> SRSdiag, src/heatflow/mod_Sparse.f90:452-466, 350-366: the diagonal is
 * re-extracted by a scan over all entries on every call, as the reference does. */
static void asolve(int64_t n, int64_t len, const double *val, const int64_t *irow,
                   const int64_t *jcol, const double *b, double *x) {
    int64_t k = 0;
    for (int64_t c = 0; c < len; ++c)
        if (irow[c] == jcol[c]) { if (k < n) x[k] = val[c]; k++; }
    for (int64_t i = 0; i < n; ++i) x[i] = b[i] / x[i];
}

static double dot(int64_t n, const double *a, const double *b) {
    double s = 0.0;   /* dot_product: sequential sum */
    for (int64_t i = 0; i < n; ++i) s = s + a[i] * b[i];
    return s;
}

/* linbcg, src/heatflow/mod_Sparse.f90:372-434 */
void hfo_linbcg(int64_t n, int64_t len, const double *val, const int64_t *irow,
                const int64_t *jcol, const double *b, double *x, int64_t itol, double tol,
                int64_t itmax, int64_t *iter_out, double *err_out) {
    double *p = malloc(sizeof(double) * n), *pp = malloc(sizeof(double) * n);
    double *r = malloc(sizeof(double) * n), *rr = malloc(sizeof(double) * n);
    double *z = malloc(sizeof(double) * n), *zz = malloc(sizeof(double) * n);
    double ak, akden, bk, bkden, bknum, bnrm, err = 0.0;
    int64_t iter;
    srsax(n, len, val, irow, jcol, x, r);                  /* :390 */
    for (int64_t i = 0; i < n; ++i) { r[i] = b[i] - r[i]; rr[i] = r[i]; }
    bnrm = sqrt(dot(n, b, b));                              /* :397 */
    asolve(n, len, val, irow, jcol, r, z);
    asolve(n, len, val, irow, jcol, rr, zz);
    bknum = dot(n, z, rr);
    for (int64_t i = 0; i < n; ++i) { p[i] = z[i]; pp[i] = zz[i]; }
    for (iter = 1; iter <= itmax; ++iter) {                 /* solve_loop :404-433 */
        srsax(n, len, val, irow, jcol, p, z);
        bkden = bknum;
        akden = dot(n, z, pp);
        ak = bknum / akden;
        srstx(n, len, val, irow, jcol, pp, zz);
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] + ak * p[i];
            r[i] = r[i] - ak * z[i];
            rr[i] = rr[i] - ak * zz[i];
        }
        asolve(n, len, val, irow, jcol, r, z);
        if (itol == 1) err = sqrt(dot(n, r, r)) / bnrm;
        else if (itol == 2) err = sqrt(dot(n, z, z)) / bnrm;
        if (err <= tol) break;
        asolve(n, len, val, irow, jcol, rr, zz);
        bknum = dot(n, z, rr);
        bk = bknum / bkden;
        for (int64_t i = 0; i < n; ++i) { p[i] = bk * p[i] + z[i]; pp[i] = bk * pp[i] + zz[i]; }
    }
    /* Fortran DO leaves iter = itmax+1 when the loop runs out */
    if (iter_out) *iter_out = iter;
    if (err_out) *err_out = err;
    free(p); free(pp); free(r); free(rr); free(z); free(zz);
}

static int any_nan(int64_t n, const double *v) {
    for (int64_t i = 0; i < n; ++i) if (isnan(v[i])) return 1;
    return 0;
}

/* simulate, src/heatflow/mod_evolve.f90:45-177 */
int hfo_simulate(const hfo_system *s, const hfo_solver *sv, int64_t len, const double *val,
                 const int64_t *irow, const int64_t *jcol, int64_t itime, double *temp_p,
                 double *temp_pp, int64_t *iters, double *err_out) {
    const int64_t N = s->nx * s->ny * s->nz;
    int rc = HFO_OK;
    double *S = calloc(N, sizeof(double)), *x = calloc(N, sizeof(double));
    double *Q = calloc(N, sizeof(double)), *Qdens = calloc(N, sizeof(double));
    double *SCAT = calloc(N, sizeof(double)), *B = calloc(N, sizeof(double));
    double inverse_time = 1.0 / s->time_step;

    hfo_boundary(s, B);                                   /* :66 */
    if (any_nan(N, B)) { rc = HFO_NAN_B; goto done; }
    int any_heater = 0;
    for (int64_t I = 0; I < N; ++I) if (s->iheater[I] > 0) { any_heater = 1; break; }
    if (any_heater) {                                     /* :79-90 */
        hfo_heater(s, itime, temp_p, Q, Qdens);
        if (any_nan(N, Q)) { rc = HFO_NAN_Q; goto done; }
        if (any_nan(N, Qdens)) { rc = HFO_NAN_QDENS; goto done; }
    }
    if (s->qdens_extra)        /* superset input: per-cell power density (no reference counterpart) */
        for (int64_t I = 0; I < N; ++I) Qdens[I] = Qdens[I] + s->qdens_extra[I];
    if (s->icattaneo == 1) {                              /* :98-106 */
        hfo_scat(s, temp_p, temp_pp, SCAT);
        if (any_nan(N, SCAT)) { rc = HFO_NAN_SCAT; goto done; }
    }
    for (int64_t I = 0; I < N; ++I) {                     /* :112-119 */
        if (s->isteady == 0) {
            double lin_rhoc = s->rho[I] * s->heat_capacity[I];
            S[I] = -(inverse_time * temp_p[I] * lin_rhoc) - Qdens[I] - B[I];
            if (s->icattaneo == 1) S[I] = S[I] + SCAT[I];
        } else {
            S[I] = -Qdens[I] - B[I];
        }
    }
    if (any_nan(N, S)) { rc = HFO_NAN_S; goto done; }
    int small = 0;                                        /* :140-141 */
    for (int64_t I = 0; I < N; ++I) {
        x[I] = temp_p[I] + (temp_p[I] - temp_pp[I]);
        if (x[I] - temp_p[I] < TINY) small = 1;
    }
    if (small) for (int64_t I = 0; I < N; ++I) x[I] = x[I] + TINY;
    int64_t iter = 0;
    double err = 0.0;
    hfo_linbcg(N, len, val, irow, jcol, S, x, sv->itol, sv->tol, sv->itmax, &iter, &err);
    if (iters) *iters = iter;
    if (err_out) *err_out = err;
    if (any_nan(N, x)) { rc = HFO_NAN_X; goto done; }
    for (int64_t I = 0; I < N; ++I) { temp_pp[I] = temp_p[I]; temp_p[I] = x[I]; }  /* :169-170 */
done:
    free(S); free(x); free(Q); free(Qdens); free(SCAT); free(B);
    return rc;
}

int hfo_run(const hfo_system *s, const hfo_solver *sv, int64_t ntime, double t_system,
            int init_state, double *temp_p, double *temp_pp, int64_t *total_iters) {
    const int64_t N = s->nx * s->ny * s->nz;
    int64_t len = hfo_coo_len(s), count = 0;
    double *val = malloc(sizeof(double) * len);
    int64_t *irow = malloc(sizeof(int64_t) * len), *jcol = malloc(sizeof(int64_t) * len);
    hfo_sparse_hmatrix(s, val, irow, jcol, &count);
    if (init_state)
        for (int64_t I = 0; I < N; ++I) { temp_p[I] = t_system; temp_pp[I] = t_system; }
    int rc = HFO_OK;
    int64_t tot = 0;
    for (int64_t it = 1; it <= ntime && rc == HFO_OK; ++it) {
        int64_t iters = 0;
        rc = hfo_simulate(s, sv, len, val, irow, jcol, it, temp_p, temp_pp, &iters, NULL);
        tot += iters;
    }
    if (total_iters) *total_iters = tot;
    free(val); free(irow); free(jcol);
    return rc;
}

/* setup::stability, src/heatflow/mod_setup.f90:158-199.  The threshold 1.0/12.0 is a
 * default-real (single precision) expression in the reference. */
int64_t hfo_stability(const hfo_system *s, double *value, int32_t *boundary_test) {
    const float thresh = 1.0f / 12.0f;
    int64_t I = 0;
    for (int64_t iz = 1; iz <= s->nz; ++iz)
        for (int64_t iy = 1; iy <= s->ny; ++iy)
            for (int64_t ix = 1; ix <= s->nx; ++ix, ++I) {
                const double l1 = s->len_x[I], l2 = s->len_y[I], l3 = s->len_z[I];
                const double rhoc = s->rho[I] * s->heat_capacity[I];
                double alpha = s->kappa[I] / rhoc;                                  /* :170 */
                double var = s->time_step * alpha *
                             (1 / (l1 * l1) + 1 / (l2 * l2) + 1 / (l3 * l3));        /* :171-173 */
                if (var > (double)thresh) {                                         /* :176 */
                    if (value) *value = var;
                    if (boundary_test) *boundary_test = 0;
                    return I;
                }
                if (ix == 1 || iy == 1 || iz == 1) {                                /* :187 */
                    alpha = s->kappa_bound[0] / rhoc;
                    var = s->time_step * alpha * (1 / (l1 * l1) + 1 / (l2 * l2) + 1 / (l3 * l3));
                    if (var > (double)thresh) {
                        if (value) *value = var;
                        if (boundary_test) *boundary_test = 1;
                        return I;
                    }
                }
            }
    return -1;
}
