// hf_kernels.hip -- gfx950 kernels of the HeatFlow time-step path.
//
// Kernels (one thread per cell per loop trip, see hf_device.hpp for the launch layout):
//   k_assemble     H as a 7-point stencil + boundary vector B        (mod_hmatrix.f90, mod_boundary.f90)
//   k_scale        Jacobi scaling s = 1/sqrt(-diag), scaled couplings, FP32 copies
//   k_coo_*        export in the reference COO order                  (mod_setup.f90:89-152)
//   k_qdens        per-step heater power density                      (mod_heating.f90:27-169)
//   k_cg_b         r -= alpha q, r.r                                   (mod_Sparse.f90:412-420)
//   k_cg_tail      y += alpha p of the last iteration
//   k_finalize     x = x0 + s*y, NaN guard, TINY flag of the next step (mod_evolve.f90:140-170)
//   k_fin_step     refinement test and next-step norms after the fused mixed step
//   k_fin_next     next-step norms after the fused FP64 step
//   k_extrap       polynomial extrapolation of the last solutions: the adjoint solve's guess
//   k_fin_*        one-block finishing kernels of the reductions (scalars, convergence)
// The stencil kernels (fused RHS/residual, CG SpMV, refinement residual) are in hf_stencil.hip.
//
// All HIP code here is compiled with -ffp-contract=off so every product and sum rounds as
// written; the assembly therefore reproduces hmatrixfunc's values bit for bit.
#include "hf_device.hpp"
#include "hf_kernels.hpp"

namespace hf {

// ---- assembly ------------------------------------------------------------------------------

// calculate_conductivity, interior branch (src/heatflow/mod_hmatrix.f90:288-315): `in` is the
// neighbour whose coordinate differs, `out` the cell the value is computed for.
__device__ __forceinline__ double cond_int(double kin, double kout, double lin, double lout,
                                           double br) {
    double kab = (lin + lout) * kin * kout / (lin * kout + lout * kin);
    kab = kab / (lout * lout);
    if (kin != kout) kab = kab * br;
    return kab;
}

// boundry_diag_term (src/heatflow/mod_hmatrix.f90:357-385)
__device__ __forceinline__ double cond_bnd(double kb, double kappa, double L, double br) {
    double kab = ((2.0 * kb * kappa) / (kb + kappa)) / (L * L);
    if (kappa != kb) kab = kab * br;
    return kab;
}

// Per-face value of the diagonal conductance of cell I towards one side of one axis:
// neighbour J (if any), the boundary term otherwise.
__device__ __forceinline__ double face_cond(bool has_nbr, uint32_t I, uint32_t J,
                                            const double *__restrict__ kappa,
                                            const double *__restrict__ L, double kb, double br) {
    if (has_nbr) return cond_int(kappa[J], kappa[I], L[J], L[I], br);
    return cond_bnd(kb, kappa[I], L[I], br);
}

__global__ void __launch_bounds__(kBlock)
k_assemble(Dims g, AsmParams ap, const double *__restrict__ kappa, const double *__restrict__ lx,
           const double *__restrict__ ly, const double *__restrict__ lz,
           const double *__restrict__ rhoc, const double *__restrict__ tau_s,
           double *__restrict__ d, double *__restrict__ cx, double *__restrict__ cy,
           double *__restrict__ cz, double *__restrict__ bvec, Scal *sc) {
    int32_t fl = 0;
    HF_FOR_CELLS(g, I) {
        Nbr n = neighbours(g, I);
        // calculate_alpha (mod_hmatrix.f90:254-267)
        double alpha = 0.0;
        if (ap.isteady == 0) {
            double tau = ap.icattaneo == 0 ? 0.0 : tau_s[I];
            alpha = (tau * rhoc[I]) + (ap.inverse_time * rhoc[I]);
        }
        // A..G: in = neighbour, out = I (mod_hmatrix.f90:133-138)
        double A = face_cond(n.hxm, I, n.xm, kappa, lx, ap.kb[0], ap.br);
        double B = face_cond(n.hxp, I, n.xp, kappa, lx, ap.kb[1], ap.br);
        double D = face_cond(n.hym, I, n.ym, kappa, ly, ap.kb[2], ap.br);
        double E = face_cond(n.hyp, I, n.yp, kappa, ly, ap.kb[3], ap.br);
        double F = face_cond(n.hzm, I, n.zm, kappa, lz, ap.kb[4], ap.br);
        double G = face_cond(n.hzp, I, n.zp, kappa, lz, ap.kb[5], ap.br);
        d[I] = -(A + B + D + E + F + G) - alpha;  // :144
        // forward couplings: value of hmatrixfunc(i, j) for the pair, in = lower linear index
        double c;
        c = 0.0;
        if (n.hxp) {
            uint32_t lo = n.xp > I ? I : n.xp, hi = n.xp > I ? n.xp : I;
            c = cond_int(kappa[lo], kappa[hi], lx[lo], lx[hi], ap.br);
            if (fabs(c) < kTiny) c = 0.0;  // dropped from the COO (mod_setup.f90:137)
        }
        cx[I] = c;
        c = 0.0;
        if (n.hyp) {
            uint32_t lo = n.yp > I ? I : n.yp, hi = n.yp > I ? n.yp : I;
            c = cond_int(kappa[lo], kappa[hi], ly[lo], ly[hi], ap.br);
            if (fabs(c) < kTiny) c = 0.0;
        }
        cy[I] = c;
        c = 0.0;
        if (n.hzp) {
            uint32_t lo = n.zp > I ? I : n.zp, hi = n.zp > I ? n.zp : I;
            c = cond_int(kappa[lo], kappa[hi], lz[lo], lz[hi], ap.br);
            if (fabs(c) < kTiny) c = 0.0;
        }
        cz[I] = c;

        // boundary vector (src/heatflow/mod_boundary.f90:72-150)
        double kap = kappa[I];
        double b = 0.0;
        const uint32_t pos[3] = {n.x, n.y, n.z};
        const uint32_t len[3] = {g.nx, g.ny, g.nz};
        const double Ls[3] = {lx[I], ly[I], lz[I]};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (ap.periodic[ax]) continue;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int f = 2 * ax + side;
                if (pos[ax] != (side == 0 ? 0u : len[ax] - 1)) continue;
                if (ap.cg[f]) {
                    b = ap.cg_pd[f];
                } else {
                    double kh = (2.0 * kap * ap.kb[f] / (kap + ap.kb[f])) / (Ls[ax] * Ls[ax]);
                    if (kap != ap.kb[f]) kh = kh * ap.br;
                    b = b + (kh) * ap.tb[f];
                }
            }
        }
        bvec[I] = b;
        if (isnan(b)) fl |= FLAG_NAN_B;
    }
    block_flags(fl, &sc->flags);
}

// s = 1/sqrt(-diag) and the couplings of A^ = S(-H)S = I - C^, C^_IJ = s_I c_IJ s_J (FP64 and
// FP32 copies).  The scaled coupling of a pair is stored once, at the same index as c.
__global__ void __launch_bounds__(kBlock)
k_scale(Dims g, const double *__restrict__ d, const double *__restrict__ cx,
        const double *__restrict__ cy, const double *__restrict__ cz, double *__restrict__ sv,
        double *__restrict__ hx, double *__restrict__ hy, double *__restrict__ hz,
        float *__restrict__ fx, float *__restrict__ fy, float *__restrict__ fz) {
    HF_FOR_CELLS(g, I) {
        Nbr n = neighbours(g, I);
        double si = 1.0 / sqrt(-d[I]);
        sv[I] = si;
        double vx = n.hxp ? si * cx[I] * (1.0 / sqrt(-d[n.xp])) : 0.0;
        double vy = n.hyp ? si * cy[I] * (1.0 / sqrt(-d[n.yp])) : 0.0;
        double vz = n.hzp ? si * cz[I] * (1.0 / sqrt(-d[n.zp])) : 0.0;
        if (hx) { hx[I] = vx; hy[I] = vy; hz[I] = vz; }
        if (fx) { fx[I] = (float)vx; fy[I] = (float)vy; fz[I] = (float)vz; }
    }
}

// ---- COO export (reference order, mod_setup.f90:121-148) ------------------------------------

__global__ void __launch_bounds__(kBlock)
k_coo_count(Dims g, const double *__restrict__ cx, const double *__restrict__ cy,
            const double *__restrict__ cz, int64_t *__restrict__ cnt) {
    HF_FOR_CELLS(g, j) {
        Nbr n = neighbours(g, j);
        uint32_t rows[6];
        double vals[6];
        int k = coo_slots(g, n, j, cx, cy, cz, rows, vals);
        cnt[j] = 1 + 2 * k;
    }
}

__global__ void __launch_bounds__(kBlock)
k_coo_fill(Dims g, const double *__restrict__ d, const double *__restrict__ cx,
           const double *__restrict__ cy, const double *__restrict__ cz,
           const int64_t *__restrict__ off, double *__restrict__ val, int64_t *__restrict__ irow,
           int64_t *__restrict__ jcol) {
    HF_FOR_CELLS(g, j) {
        Nbr n = neighbours(g, j);
        uint32_t rows[6];
        double vals[6];
        int k = coo_slots(g, n, j, cx, cy, cz, rows, vals);
        int64_t o = off[j];
        val[o] = d[j]; irow[o] = (int64_t)j + 1; jcol[o] = (int64_t)j + 1; ++o;
        for (int s = 0; s < k; ++s) {
            val[o] = vals[s]; irow[o] = (int64_t)rows[s] + 1; jcol[o] = (int64_t)j + 1; ++o;
            val[o] = vals[s]; irow[o] = (int64_t)j + 1; jcol[o] = (int64_t)rows[s] + 1; ++o;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_coo_tail(int64_t count, int64_t len, double *__restrict__ val, int64_t *__restrict__ irow,
           int64_t *__restrict__ jcol) {
    for (int64_t k = count + (int64_t)blockIdx.x * kBlock + threadIdx.x; k < len;
         k += (int64_t)gridDim.x * kBlock) {
        val[k] = 0.0; irow[k] = -2; jcol[k] = -1;
    }
}

// exclusive scan of int64 counts in three passes (block sums, one-block scan, apply)
__global__ void __launch_bounds__(kBlock)
k_scan_blocksum(const int64_t *__restrict__ in, int64_t n, int64_t *__restrict__ bsum) {
    __shared__ int64_t lds[kBlock];
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    lds[threadIdx.x] = i < n ? in[i] : 0;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = lds[0];
}

__global__ void __launch_bounds__(1024)
k_scan_single(int64_t *__restrict__ bsum, int64_t nb, int64_t *__restrict__ total) {
    __shared__ int64_t lds[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = threadIdx.x * per;
    int64_t s = 0;
    for (int64_t k = 0; k < per && b0 + k < nb; ++k) s += bsum[b0 + k];
    lds[threadIdx.x] = s;
    __syncthreads();
    // inclusive Hillis-Steele scan over the 1024 thread sums
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = threadIdx.x >= (unsigned)off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = threadIdx.x == 0 ? 0 : lds[threadIdx.x - 1];
    for (int64_t k = 0; k < per && b0 + k < nb; ++k) {
        int64_t v = bsum[b0 + k];
        bsum[b0 + k] = run;
        run += v;
    }
    if (threadIdx.x == 1023) *total = lds[1023];
}

__global__ void __launch_bounds__(kBlock)
k_scan_apply(const int64_t *__restrict__ in, int64_t n, const int64_t *__restrict__ bsum,
             int64_t *__restrict__ out) {
    __shared__ int64_t lds[kBlock];
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int64_t v = i < n ? in[i] : 0;
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int64_t t = threadIdx.x >= (unsigned)off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = bsum[blockIdx.x] + lds[threadIdx.x] - v;
}

// ---- right-hand side -----------------------------------------------------------------------

// Heater power density Qdens of one step (src/heatflow/mod_heating.f90:59-167) plus the
// superset per-cell density (heater_q: hf_kernels.hpp).
__global__ void __launch_bounds__(kBlock)
k_qdens(Dims g, HeatParams hp, const int8_t *__restrict__ iheater,
        const double *__restrict__ ells, const double *__restrict__ tau_s,
        const double *__restrict__ tp, const double *__restrict__ qextra,
        double *__restrict__ qd, Scal *sc) {
    int32_t fl = 0;
    HF_FOR_CELLS(g, I) {
        double q = 0.0, qdv = 0.0;
        if (hp.any_heater) {
            q = heater_q(hp, iheater[I], tau_s[I], ells[I], tp[I]);
            qdv = hp.heated_volume > 0.0 ? q / hp.heated_volume : 0.0;
            if (isnan(q)) fl |= FLAG_NAN_Q;
            if (isnan(qdv)) fl |= FLAG_NAN_QDENS;
        }
        if (qextra) qdv = qdv + qextra[I];
        qd[I] = qdv;
    }
    block_flags(fl, &sc->flags);
}

// Start of a solve: ||b^||^2, r.r, thresholds (after k_rhs_res).
__global__ void __launch_bounds__(kBlock)
k_fin_start(uint32_t nb, const double *partials, Scal *sc, double tol, int64_t itmax,
            int32_t allow_short) {
    double v[2];
    sum_partials<2>(partials, nb, v);
    if (threadIdx.x != 0) return;
    sc->bnorm2 = v[0];
    sc->rr = v[1];
    sc->rr_start = v[1];
    sc->outer_rr = v[1];
    sc->thresh2 = tol * tol * v[0];
    sc->iter = 0;
    sc->itmax = itmax;
    sc->first = 1;
    sc->pq = 0.0;
    sc->alpha = 0.0;
    sc->beta = 0.0;
    sc->shortcut = 0;
    sc->allow_short = allow_short;
    sc->yset = 0;
    sc->ypend = 0;
    sc->kpend = 0;
    // An exactly zero residual needs no iteration (the reference would divide 0/0).
    sc->done = (v[1] == 0.0) || isnan(v[1]) || itmax <= 0;
    if (isnan(v[1])) sc->flags |= FLAG_BREAKDOWN;
}

// ---- CG on A^ = I - C^ (unit diagonal; the Jacobi-preconditioned CG of linbcg in scaled form)

// alpha = r.r / p.q (after k_cg_a).  On the first iteration p = r0, so the next residual's
// norm is known without forming it: |r0 - alpha q|^2 = rr - 2 alpha p.q + alpha^2 q.q
// = alpha^2 q.q - rr.  If that already meets the stop test the solve ends here with
// y = alpha p0 (`shortcut`; the term stays pending for the consumers), and k_cg_b is skipped.
// After iteration k >= 1's k_cg_a, y holds the terms of iterations 0 .. k-1.
__global__ void __launch_bounds__(kBlock)
k_fin_a(uint32_t nb, const double *partials, Scal *sc) {
    if (sc->done) return;
    double v[2];
    sum_partials<2>(partials, nb, v);
    if (threadIdx.x != 0) return;
    if (!sc->first) {
        sc->yset = 1;
        sc->ypend = 0;
    }
    sc->pq = v[0];
    sc->alpha = sc->rr / v[0];
    if (!(v[0] > 0.0)) {  // A^ is SPD: p.q <= 0 or NaN is a breakdown
        sc->done = 1;
        sc->alpha = 0.0;
        sc->flags |= FLAG_BREAKDOWN;
        return;
    }
    if (sc->first && sc->allow_short) {
        const double a = sc->alpha;
        const double rr1 = a * a * v[1] - sc->rr;
        if (rr1 <= sc->thresh2 && sc->itmax >= 1) {
            sc->rr = rr1 > 0.0 ? rr1 : 0.0;
            sc->iter = 1;
            sc->shortcut = 1;
            sc->ypend = 1;
            sc->kpend = 0;
            sc->done = 1;
        }
    }
}

// r -= alpha q, partial r.r, with 16-byte accesses: block b (XCD-mapped) takes a contiguous
// chunk of the 16-byte vectors; the cells past the last full vector go to logical block 0.
// (y += alpha p is left to the next iteration's k_cg_a, which reads p anyway.)
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_b(Dims g, T *__restrict__ r, const T *__restrict__ q, Scal *sc, double *partials) {
    constexpr int W = 16 / sizeof(T);
    using V = T __attribute__((ext_vector_type(W)));
    if (sc->done) return;
    const T alpha = (T)sc->alpha;
    double acc[1] = {0.0};
    const uint32_t nv = g.N / W;
    const uint32_t b = logical_block(g);
    const uint32_t per = (nv + g.nblk - 1) / g.nblk;
    const uint64_t v0 = (uint64_t)b * per;
    const uint32_t v1 = (uint32_t)(v0 + per < nv ? v0 + per : nv);
    for (uint32_t v = (uint32_t)(v0 < nv ? v0 : nv) + threadIdx.x; v < v1; v += kBlock) {
        const V qv = __builtin_nontemporal_load((const V *)q + v);
        V rv = __builtin_nontemporal_load((const V *)r + v);
#pragma unroll
        for (int k = 0; k < W; ++k) {
            rv[k] = rv[k] - alpha * qv[k];
            acc[0] += (double)rv[k] * (double)rv[k];
        }
        __builtin_nontemporal_store(rv, (V *)r + v);
    }
    if (b == 0) {
        for (uint32_t I = nv * W + threadIdx.x; I < g.N; I += kBlock) {
            const T ri = r[I] - alpha * q[I];
            r[I] = ri;
            acc[0] += (double)ri * (double)ri;
        }
    }
    block_partials<1>(acc, partials);
}

// Norms of linbcg's own stop rule (src/heatflow/mod_Sparse.f90:395, 418-425) for hf_solve.
// The CG here runs on the scaled system, whose residual is r^ = -s r with s = 1/sqrt(-diag H),
// so r = -r^ / s and the Jacobi-preconditioned residual z = r / diag H = s r^.  Sum of w_I^2 v_I^2
// over the cells: mode 0 w = 1 (||b||^2), mode 1 w = 1/s (||r||^2 from r^, itol 1), mode 2
// w = s (||z||^2, itol 2).  expect >= 0: only when iteration expect - 1 has just completed
// (sc->iter == expect), else the launch returns at once (its finishing kernel skips as well).
__global__ void __launch_bounds__(kBlock)
k_wsum2(Dims g, const double *__restrict__ v, const double *__restrict__ s, int32_t mode,
        int64_t expect, const Scal *sc, double *partials) {
    if (expect >= 0 && sc->iter != expect) return;
    double acc[1] = {0.0};
    HF_FOR_CELLS(g, I) {
        double t = v[I];
        if (mode == 1) t = t / s[I];
        else if (mode == 2) t = t * s[I];
        acc[0] += t * t;
    }
    block_partials<1>(acc, partials);
}

// mode 0: ||b||^2 -> wb2, and wrr = 0 (a solve that needs no iteration reports err 0).
// Else, when iteration expect - 1 has just completed: wrr = its weighted norm^2, and the solve
// is done if wrr <= tol^2 ||b||^2 (linbcg: err <= tol).
__global__ void __launch_bounds__(kBlock)
k_fin_wstop(uint32_t nb, const double *partials, Scal *sc, double tol, int32_t mode,
            int64_t expect) {
    if (mode != 0 && sc->iter != expect) return;
    double v[1];
    sum_partials<1>(partials, nb, v);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        sc->wb2 = v[0];
        sc->wrr = 0.0;
        // the solve's own test on the scaled residual is switched off (thresh2 < 0 never
        // holds; the first-iteration prediction cannot meet it either)
        sc->thresh2 = -1.0;
        return;
    }
    sc->wrr = v[0];
    if (v[0] <= tol * tol * sc->wb2) sc->done = 1;
}

// After a solve: y += alpha p_K, the term of its last iteration K >= 1 (p_K = pa for odd K, pb
// for even); nothing when K = 0 or no term is pending (CgY).  16-byte accesses as k_cg_b.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_tail(Dims g, T *__restrict__ y, const T *__restrict__ pa, const T *__restrict__ pb,
          const Scal *sc) {
    constexpr int W = 16 / sizeof(T);
    using V = T __attribute__((ext_vector_type(W)));
    if (!(sc->ypend && sc->yset)) return;
    const T *__restrict__ p = ((sc->kpend - 1) & 1) ? pb : pa;
    const T alpha = (T)sc->alpha;
    const uint32_t nv = g.N / W;
    const uint32_t b = logical_block(g);
    const uint32_t per = (nv + g.nblk - 1) / g.nblk;
    const uint64_t v0 = (uint64_t)b * per;
    const uint32_t v1 = (uint32_t)(v0 + per < nv ? v0 + per : nv);
    for (uint32_t v = (uint32_t)(v0 < nv ? v0 : nv) + threadIdx.x; v < v1; v += kBlock) {
        const V pv = __builtin_nontemporal_load((const V *)p + v);
        V yv = __builtin_nontemporal_load((const V *)y + v);
#pragma unroll
        for (int k = 0; k < W; ++k) yv[k] = yv[k] + alpha * pv[k];
        __builtin_nontemporal_store(yv, (V *)y + v);
    }
    if (b == 0)
        for (uint32_t I = nv * W + threadIdx.x; I < g.N; I += kBlock) y[I] = y[I] + alpha * p[I];
}

// beta, convergence test on ||r^|| / ||b^|| (after k_cg_b or k_cg_b1_z); the term of this
// iteration k, alpha p_k, is pending in y
__global__ void __launch_bounds__(kBlock)
k_fin_b(uint32_t nb, const double *partials, Scal *sc) {
    if (sc->done) return;
    double v[1];
    sum_partials<1>(partials, nb, v);
    if (threadIdx.x != 0) return;
    const double rr = v[0];
    sc->kpend = sc->iter;
    sc->ypend = 1;
    sc->iter += 1;
    sc->beta = rr / sc->rr;
    sc->rr = rr;
    sc->first = 0;
    if (rr <= sc->thresh2 || sc->iter >= sc->itmax || isnan(rr)) sc->done = 1;
    if (isnan(rr)) sc->flags |= FLAG_BREAKDOWN;
}

// x = x0 + s*y written over T^{n-1} (in place per cell), NaN guard of x (mod_evolve.f90:153),
// and the TINY test of the next step's initial guess (:141).  When FromState is false the
// explicit initial guess x is updated instead (linbcg-compatible solve).
template <bool FromState>
__global__ void __launch_bounds__(kBlock)
k_finalize(Dims g, const double *__restrict__ tp, double *__restrict__ tpp_io,
           const double *__restrict__ sv, const double *__restrict__ y,
           const double *__restrict__ r0, Scal *sc) {
    const bool tiny = sc->tiny != 0;
    const CgY<double> cy = cg_y(sc, y, r0);
    const bool have_y = cy.any;
    int32_t fl = 0;
    HF_FOR_CELLS(g, I) {
        double x0, tpi = 0.0;
        if (FromState) {
            tpi = ldnt(&tp[I]);
            x0 = tpi + (tpi - ldnt(&tpp_io[I]));
            if (tiny) x0 = x0 + kTiny;
        } else {
            x0 = ldnt(&tpp_io[I]);
        }
        const double x = have_y ? x0 + ldnt(&sv[I]) * cy.at(I) : x0;
        stnt(&tpp_io[I], x);
        if (isnan(x)) fl |= FLAG_NAN_X;
        if (FromState && (x + (x - tpi)) - x < kTiny) fl |= FLAG_TINY_NEXT;
    }
    block_flags(fl, &sc->flags);
}

// outer residual of a refinement sweep (after k_ir_res)
__global__ void __launch_bounds__(kBlock)
k_fin_ir(uint32_t nb, const double *partials, Scal *sc, int64_t sweeps_max) {
    double v[1];
    sum_partials<1>(partials, nb, v);
    if (threadIdx.x != 0) return;
    // refinement stops at the tolerance, at the sweep cap, or when a sweep no longer halves
    // the FP64 residual (it has reached the FP64 round-off floor of H x)
    const double prev = sc->outer_rr;
    sc->outer_rr = v[0];
    sc->iter += 1;   // refinement sweeps
    const bool stalled = sc->iter > 1 && v[0] > 0.25 * prev;
    sc->done = (v[0] <= sc->thresh2) || sc->iter >= sweeps_max || stalled || isnan(v[0]);
    if (isnan(v[0])) sc->flags |= FLAG_BREAKDOWN;
}

// Starts an inner FP32 solve from the outer scalars: thresholds relative to this sweep's
// residual, floored by the outer target (no point solving further than the outer test needs).
// Run by one thread.
static __device__ double inner_target(const Scal *outer, double inner_tol, double outer_tol,
                                      double safety) {
    const double t_rel = inner_tol * inner_tol * outer->outer_rr;
    const double t_abs = safety * safety * outer_tol * outer_tol * outer->bnorm2;
    return t_rel > t_abs ? t_rel : t_abs;
}
static __device__ void inner_begin(Scal *outer, Scal *inner, double inner_tol, double outer_tol,
                                   int64_t inner_itmax, double safety, int32_t allow_short) {
    const double rr = outer->outer_rr;
    inner->rr = rr;
    inner->rr_start = rr;
    inner->bnorm2 = outer->bnorm2;
    inner->thresh2 = inner_target(outer, inner_tol, outer_tol, safety);
    inner->iter = 0;
    inner->itmax = inner_itmax;
    inner->first = 1;
    inner->beta = 0.0;
    inner->alpha = 0.0;
    inner->shortcut = 0;
    inner->yset = 0;
    inner->ypend = 0;
    inner->kpend = 0;
    inner->allow_short = allow_short;
    inner->done = (rr == 0.0) || isnan(rr) || inner_itmax <= 0;
    // the refinement pass after this solve recomputes the next step's TINY test and S: only
    // the last sweep of a step may set them
    outer->flags &= ~FLAG_TINY_NEXT;
    outer->flags_next = 0;
}
__global__ void k_inner_begin(Scal *outer, Scal *inner, double inner_tol, double outer_tol,
                              int64_t inner_itmax, double safety, int32_t allow_short) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    inner_begin(outer, inner, inner_tol, outer_tol, inner_itmax, safety, allow_short);
}

// Clears the per-step flag bits, keeping TINY_NEXT -> tiny for this step; NaN bits of this
// step's S found ahead by the previous fused step become current.
static __device__ void step_begin(Scal *sc) {
    sc->tiny = (sc->flags & FLAG_TINY_NEXT) ? 1 : 0;
    sc->flags = (sc->flags & FLAG_NAN_B) | sc->flags_next;
    sc->flags_next = 0;
}
__global__ void k_step_begin(Scal *sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    step_begin(sc);
}

// Start of a carried mixed-precision step: the previous fused pass left this step's ||b^||^2
// and its ||r^0||^2 for the TINY case it guessed (as k_fin_start for a residual from the
// state).  The host carries only a right guess; the other case's norm is NaN.
static __device__ void carry_begin(Scal *sc, double tol, int64_t itmax, int32_t allow_short) {
    const double rr = sc->next_rr[sc->tiny ? 1 : 0];
    sc->bnorm2 = sc->next_bnorm2;
    sc->rr = rr;
    sc->rr_start = rr;
    sc->outer_rr = rr;
    sc->thresh2 = tol * tol * sc->next_bnorm2;
    sc->iter = 0;
    sc->itmax = itmax;
    sc->first = 1;
    sc->pq = 0.0;
    sc->alpha = 0.0;
    sc->beta = 0.0;
    sc->shortcut = 0;
    sc->yset = 0;
    sc->ypend = 0;
    sc->kpend = 0;
    sc->allow_short = allow_short;
    sc->done = (rr == 0.0) || isnan(rr) || itmax <= 0;
    if (isnan(rr)) sc->flags |= FLAG_BREAKDOWN;
}

// Start of a carried mixed-precision step and of its first inner solve in one launch:
// step_begin, carry_begin, inner_begin.
// With `reuse`, the first sweep takes y = alpha r^0 with the step length alpha of the last
// inner solve (kept in inner->alpha) instead of solving: the inner scalars describe a solve
// that ended on its first iteration, and the sweep's FP64 residual must meet half the target
// (in norm) the solve would have had (reuse_thresh2, k_fin_step), so that a kept sweep is no
// less accurate than a solve's.
__global__ void k_carry_start(Scal *sc, Scal *inner, double tol, int64_t itmax,
                              int32_t allow_short, double inner_tol, int64_t inner_itmax,
                              double safety, int32_t reuse) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    step_begin(sc);
    carry_begin(sc, tol, itmax, allow_short);
    if (!inner) return;
    if (reuse) {
        sc->reuse_thresh2 = 0.25 * inner_target(sc, inner_tol, tol, safety);
        inner->iter = 0;
        inner->shortcut = 1;
        inner->yset = 0;
        inner->ypend = 1;
        inner->kpend = 0;
        inner->done = 1;
        // as inner_begin: only the last sweep of a step sets the next step's flags
        sc->flags &= ~FLAG_TINY_NEXT;
        sc->flags_next = 0;
    } else {
        inner_begin(sc, inner, inner_tol, tol, inner_itmax, safety, allow_short);
    }
}

// After the fused mixed step: FP64 residual of this step (refinement test as k_fin_ir) and
// the next step's norms.  partials: ||r^||^2, ||b^_next||^2, ||r^0_next||^2 (TINY case
// `guess`), r^0 . r'.  A sweep that reused alpha (y = alpha r^0, r' = r^0 - alpha A^ r^0)
// leaves the step length a solve would have taken, alpha r^0.r^0 / (r^0.r^0 - r^0.r'), in
// inner->alpha for the next reuse, so the reused step length is never more than one step old.
__global__ void __launch_bounds__(kBlock)
k_fin_step(uint32_t nb, const double *partials, Scal *sc, Scal *inner, int64_t sweeps_max,
           int32_t guess, int32_t reused, int32_t freeze) {
    double v[4];
    sum_partials<4>(partials, nb, v);
    if (threadIdx.x != 0) return;
    if (reused && !freeze) {
        const double r0r0 = sc->rr_start, den = r0r0 - v[3];
        if (den > 0.0 && isfinite(den)) inner->alpha = inner->alpha * (r0r0 / den);
    }
    const double prev = sc->outer_rr;
    // a sweep that reused alpha passes only within the inner target (reuse_thresh2), and within
    // the outer tolerance like any sweep
    const double thr = reused ? fmin(sc->reuse_thresh2, sc->thresh2) : sc->thresh2;
    sc->outer_rr = v[0];
    sc->iter += 1;
    const bool stalled = sc->iter > 1 && v[0] > 0.25 * prev;
    sc->done = (v[0] <= thr) || sc->iter >= sweeps_max || stalled || isnan(v[0]);
    if (isnan(v[0])) sc->flags |= FLAG_BREAKDOWN;
    sc->next_bnorm2 = v[1];
    sc->next_rr[guess ? 1 : 0] = v[2];
    sc->next_rr[guess ? 0 : 1] = __builtin_nan("");
}

// After a two-step pass (k_ir_step2_z): the sums of both steps, their tests and the outcome in
// sc->pair_ok[0]: 2 both steps stand, 1 only step n (step n+1's test failed, its TINY outcome
// differed from the guess, or a NaN guard of step n+1 tripped), 0 neither (step n's test
// failed or a NaN guard of step n tripped: the host runs step n again, alone, or, when step n
// took alpha from its own solve (`solved`), goes on with its single-step pass).  Step n meets
// the test of a sweep that reused alpha (k_fin_step: half the inner target and the outer
// tolerance), or the outer tolerance alone when `solved`; step n+1 the reuse test with its
// targets formed from the pass's own sums, as k_carry_start forms them from the carried ones.
// The step length of the last step that stands and reused alpha is left in inner->alpha for
// the next reuse (unless `freeze`).  pair_ok[1]: the TINY outcome of the step after the pair
// (carry the pass's residual if it equals the guess).  `reject0` (a test aid) fails step n's
// test, so that every pair falls back.
__global__ void __launch_bounds__(kBlock)
k_fin_step2(uint32_t nb, const double *partials, Scal *sc, Scal *inner, int32_t guess,
            int32_t freeze, int32_t solved, int32_t reject0, double tol, double inner_tol,
            double safety) {
    double v[8];
    sum_partials<8>(partials, nb, v);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 8; ++k) sc->pair_v[k] = v[k];
    const int32_t nan0 = (sc->flags & FLAG_NAN_X) | (sc->flags_next & (FLAG_NAN_SCAT | FLAG_NAN_S));
    const int32_t nan1 = sc->flags2 & (FLAG_NAN_X2 | FLAG_NAN_SCAT2 | FLAG_NAN_S2);
    const bool t1 = (sc->flags & FLAG_TINY_NEXT) != 0;
    const bool t2 = (sc->flags2 & FLAG_TINY_NEXT2) != 0;
    // step n: ||r'||^2 = v0, ||b^_{n+1}||^2 = v1, ||r^0_{n+1}||^2 = v2, r^0_n . r' = v3; a
    // step whose alpha came from its own solve meets the outer tolerance only (as k_fin_step)
    const bool ok0 = !reject0 && !nan0 &&
                     v[0] <= (solved ? sc->thresh2 : fmin(sc->reuse_thresh2, sc->thresh2));
    // step n+1: v4 .. v7 likewise, its targets from v1, v2
    const double thresh2_1 = tol * tol * v[1];
    const double t_rel = inner_tol * inner_tol * v[2];
    const double t_abs = safety * safety * tol * tol * v[1];
    const double reuse2_1 = 0.25 * (t_rel > t_abs ? t_rel : t_abs);
    const bool ok1 = !nan1 && (t1 == (guess != 0)) && v[4] <= fmin(reuse2_1, thresh2_1);
    const int32_t outcome = ok0 ? (ok1 ? 2 : 1) : 0;
    sc->pair_ok[0] = outcome;
    sc->pair_ok[1] = t2 ? 1 : 0;
    if (!freeze) {
        double r0r0 = 0.0, den = 0.0;
        if (outcome == 2) {
            r0r0 = v[2];
            den = r0r0 - v[7];
        } else if (outcome == 1 && !solved) {
            r0r0 = sc->rr_start;
            den = r0r0 - v[3];
        }
        if (outcome > 0 && den > 0.0 && isfinite(den)) inner->alpha = inner->alpha * (r0r0 / den);
    }
    if (outcome == 2) {
        sc->bnorm2 = v[1];
        sc->outer_rr = v[4];
        sc->next_bnorm2 = v[5];
        sc->next_rr[guess ? 1 : 0] = v[6];
        sc->next_rr[guess ? 0 : 1] = __builtin_nan("");
        // the TINY outcome of the step after the pair, for its step_begin
        sc->flags = (sc->flags & ~FLAG_TINY_NEXT) | (t2 ? FLAG_TINY_NEXT : 0);
        sc->flags_next = 0;
    } else if (outcome == 1) {
        sc->outer_rr = v[0];
        sc->flags_next = 0;
    } else if (solved) {
        // nothing stands: step n goes on from its solve with the single-step pass, which sets
        // the flags again
        sc->flags &= ~(FLAG_NAN_X | FLAG_TINY_NEXT);
        sc->flags_next = 0;
    } else {
        // nothing stands: step n runs again from the state with its own TINY outcome
        sc->flags = (sc->flags & FLAG_NAN_B) | (sc->tiny ? FLAG_TINY_NEXT : 0);
        sc->flags_next = 0;
    }
    sc->flags2 = 0;
}

// After the fused FP64 step: the next step's norms ||b^_next||^2, ||r^0_next||^2 (TINY case
// `guess`; the other case's is NaN, never carried)
__global__ void __launch_bounds__(kBlock)
k_fin_next(uint32_t nb, const double *partials, Scal *sc, int32_t guess) {
    double v[2];
    sum_partials<2>(partials, nb, v);
    if (threadIdx.x != 0) return;
    sc->next_bnorm2 = v[0];
    sc->next_rr[guess ? 1 : 0] = v[1];
    sc->next_rr[guess ? 0 : 1] = __builtin_nan("");
}

// out = a + b (the fused mixed step's per-step constant part of S: Qdens + B)
__global__ void __launch_bounds__(kBlock)
k_sum2(Dims g, const double *__restrict__ a, const double *__restrict__ b,
       double *__restrict__ out) {
    HF_FOR_CELLS(g, I) { stnt(&out[I], ldnt(&a[I]) + ldnt(&b[I])); }
}

// out = sum_j w_j v_j over the last m solutions, newest first: the polynomial extrapolation
// through them (the adjoint solve's initial guess; weights from the host).  16-byte accesses
// as k_cg_b: block b (XCD-mapped) takes a contiguous chunk of the cell pairs; an odd last cell
// goes to logical block 0.
__global__ void __launch_bounds__(kBlock)
k_extrap(Dims g, ExtrapArgs a, double *__restrict__ out) {
    using V = double __attribute__((ext_vector_type(2)));
    const uint32_t nv = g.N / 2;
    const uint32_t b = logical_block(g);
    const uint32_t per = (nv + g.nblk - 1) / g.nblk;
    const uint64_t v0 = (uint64_t)b * per;
    const uint32_t v1 = (uint32_t)(v0 + per < nv ? v0 + per : nv);
    for (uint32_t i = (uint32_t)(v0 < nv ? v0 : nv) + threadIdx.x; i < v1; i += kBlock) {
        V v = a.w[0] * __builtin_nontemporal_load((const V *)a.v[0] + i);
#pragma unroll
        for (int j = 1; j < kExtrapMax; ++j)
            if (j < a.m) v = v + a.w[j] * __builtin_nontemporal_load((const V *)a.v[j] + i);
        __builtin_nontemporal_store(v, (V *)out + i);
    }
    if (b == 0) {
        for (uint32_t I = nv * 2 + threadIdx.x; I < g.N; I += kBlock) {
            double v = a.w[0] * a.v[0][I];
            for (int j = 1; j < a.m; ++j) v = v + a.w[j] * a.v[j][I];
            out[I] = v;
        }
    }
}

// TINY rule from an arbitrary state (first step after hf_set_state)
__global__ void __launch_bounds__(kBlock)
k_tiny_from_state(Dims g, const double *__restrict__ tp, const double *__restrict__ tpp,
                  Scal *sc) {
    int32_t fl = 0;
    HF_FOR_CELLS(g, I) {
        double x = tp[I] + (tp[I] - tpp[I]);
        if (x - tp[I] < kTiny) fl |= FLAG_TINY_NEXT;
    }
    block_flags(fl, &sc->flags);
}

// ---- host-side launchers -----------------------------------------------------------------

#define HF_LAUNCH(kern, g, st, ...) \
    hipLaunchKernelGGL(kern, dim3((g).nblk), dim3(kBlock), 0, st, __VA_ARGS__)

void launch_assemble(const Dims &g, const AsmParams &ap, const double *kappa, const double *lx,
                     const double *ly, const double *lz, const double *rhoc,
                     const double *tau_s, double *d, double *cx, double *cy, double *cz,
                     double *bvec, Scal *sc, hipStream_t st) {
    HF_LAUNCH(k_assemble, g, st, g, ap, kappa, lx, ly, lz, rhoc, tau_s, d, cx, cy, cz, bvec, sc);
}

void launch_scale(const Dims &g, const double *d, const double *cx, const double *cy,
                  const double *cz, double *sv, double *hx, double *hy, double *hz, float *fx,
                  float *fy, float *fz, hipStream_t st) {
    HF_LAUNCH(k_scale, g, st, g, d, cx, cy, cz, sv, hx, hy, hz, fx, fy, fz);
}

void launch_coo(const Dims &g, const double *d, const double *cx, const double *cy,
                const double *cz, int64_t *cnt, int64_t *off, int64_t *bsum, int64_t *total,
                double *val, int64_t *irow, int64_t *jcol, int64_t len, int64_t *count_host,
                hipStream_t st) {
    HF_LAUNCH(k_coo_count, g, st, g, cx, cy, cz, cnt);
    const int64_t n = g.N;
    const int64_t nb = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_scan_blocksum, dim3((unsigned)nb), dim3(kBlock), 0, st, cnt, n, bsum);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kBlock), 0, st, cnt, n, bsum, off);
    (void)hipMemcpyAsync(count_host, total, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    if (val) {
        HF_LAUNCH(k_coo_fill, g, st, g, d, cx, cy, cz, off, val, irow, jcol);
        if (*count_host < len) {
            int64_t rem = len - *count_host;
            unsigned nbt = (unsigned)((rem + kBlock - 1) / kBlock);
            if (nbt > 65536u) nbt = 65536u;
            hipLaunchKernelGGL(k_coo_tail, dim3(nbt), dim3(kBlock), 0, st, *count_host, len, val,
                               irow, jcol);
        }
    }
}

void launch_qdens(const Dims &g, const HeatParams &hp, const int8_t *iheater, const double *ells,
                  const double *tau_s, const double *tp, const double *qextra, double *qd,
                  Scal *sc, hipStream_t st) {
    HF_LAUNCH(k_qdens, g, st, g, hp, iheater, ells, tau_s, tp, qextra, qd, sc);
}

#define HF_FIN(kern, nb, st, ...) \
    hipLaunchKernelGGL(kern, dim3(1), dim3(kBlock), 0, st, nb, __VA_ARGS__)

template <typename T>
void launch_cg_b(const Dims &g, T *r, const T *q, Scal *sc, double *partials, hipStream_t st) {
    HF_LAUNCH(k_cg_b<T>, g, st, g, r, q, sc, partials);
}
template void launch_cg_b<double>(const Dims &, double *, const double *, Scal *, double *,
                                  hipStream_t);
template void launch_cg_b<float>(const Dims &, float *, const float *, Scal *, double *,
                                 hipStream_t);

template <typename T>
void launch_cg_tail(const Dims &g, T *y, const T *pa, const T *pb, const Scal *sc,
                    hipStream_t st) {
    HF_LAUNCH(k_cg_tail<T>, g, st, g, y, pa, pb, sc);
}
template void launch_cg_tail<double>(const Dims &, double *, const double *, const double *,
                                     const Scal *, hipStream_t);
template void launch_cg_tail<float>(const Dims &, float *, const float *, const float *,
                                    const Scal *, hipStream_t);

void launch_wstop(const Dims &g, const double *v, const double *s, int32_t mode, int64_t expect,
                  double tol, Scal *sc, double *partials, hipStream_t st) {
    HF_LAUNCH(k_wsum2, g, st, g, v, s, mode, expect, sc, partials);
    hipLaunchKernelGGL(k_fin_wstop, dim3(1), dim3(kBlock), 0, st, g.nblk, partials, sc, tol, mode,
                       expect);
}

void launch_finalize(const Dims &g, bool from_state, const double *tp, double *tpp_io,
                     const double *sv, const double *y, const double *r0, Scal *sc,
                     hipStream_t st) {
    if (from_state)
        HF_LAUNCH(k_finalize<true>, g, st, g, tp, tpp_io, sv, y, r0, sc);
    else
        HF_LAUNCH(k_finalize<false>, g, st, g, tp, tpp_io, sv, y, r0, sc);
}

// the stencil kernels run g.zblk blocks of the Dims they were launched with (their z-chunking),
// the flat ones g.nblk; a finishing kernel reads that many partials
void launch_fin_start(const Dims &g, const double *partials, Scal *sc, double tol, int64_t itmax,
                      bool allow_short, hipStream_t st) {
    // after k_rhs_res_z, which runs the pipelined kernels' z-chunking
    HF_FIN(k_fin_start, g.zlay[kZPipe].zblk, st, partials, sc, tol, itmax, (int32_t)allow_short);
}
void launch_fin_a(const Dims &g, const double *partials, Scal *sc, hipStream_t st) {
    HF_FIN(k_fin_a, g.zblk, st, partials, sc);
}
void launch_fin_b(const Dims &g, const double *partials, Scal *sc, hipStream_t st) {
    HF_FIN(k_fin_b, g.nblk, st, partials, sc);
}
void launch_fin_b1(const Dims &g, const double *partials, Scal *sc, hipStream_t st) {
    HF_FIN(k_fin_b, g.zblk, st, partials, sc);
}
void launch_fin_ir(const Dims &g, const double *partials, Scal *sc, int64_t sweeps_max,
                   hipStream_t st) {
    HF_FIN(k_fin_ir, g.zblk, st, partials, sc, sweeps_max);
}
void launch_fin_step(const Dims &g, const double *partials, Scal *sc, Scal *inner,
                     int64_t sweeps_max, bool tiny_guess, bool reused, bool freeze,
                     hipStream_t st) {
    // after k_ir_step_z (pipelined z-chunking)
    HF_FIN(k_fin_step, g.zlay[kZPipe].zblk, st, partials, sc, inner, sweeps_max,
           (int32_t)tiny_guess, (int32_t)reused, (int32_t)freeze);
}
void launch_fin_step2(const Dims &g, const double *partials, Scal *sc, Scal *inner,
                      bool tiny_guess, bool freeze, bool solved, bool reject0, double tol,
                      double inner_tol, double safety, hipStream_t st) {
    // after k_ir_step2_z (its own tiles)
    HF_FIN(k_fin_step2, g.p2blk, st, partials, sc, inner, (int32_t)tiny_guess,
           (int32_t)freeze, (int32_t)solved, (int32_t)reject0, tol, inner_tol, safety);
}
void launch_fin_next(const Dims &g, const double *partials, Scal *sc, bool tiny_guess,
                     hipStream_t st) {
    // after k_step64_z (pipelined z-chunking)
    HF_FIN(k_fin_next, g.zlay[kZPipe].zblk, st, partials, sc, (int32_t)tiny_guess);
}
// Copies `words` 8-byte words from src (device) to dst (pinned host memory), then, after a
// system-scope fence, writes `seq` to *seq_host: the host sees the copy complete once it sees
// the sequence number (one thread: a few hundred bytes).
__global__ void k_publish(const int64_t *src, int64_t *dst, int32_t words, int64_t *seq_host,
                          int64_t seq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int32_t k = 0; k < words; ++k) dst[k] = src[k];
    __threadfence_system();
    *(volatile int64_t *)seq_host = seq;
}
void launch_publish(const void *src, void *dst, size_t bytes, int64_t *seq_host, int64_t seq,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, (const int64_t *)src,
                       (int64_t *)dst, (int32_t)(bytes / 8), seq_host, seq);
}

void launch_carry_start(Scal *sc, Scal *inner, double tol, int64_t itmax, bool allow_short,
                        double inner_tol, int64_t inner_itmax, double safety, bool reuse,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_carry_start, dim3(1), dim3(64), 0, st, sc, inner, tol, itmax,
                       (int32_t)allow_short, inner_tol, inner_itmax, safety, (int32_t)reuse);
}

void launch_inner_begin(Scal *outer, Scal *inner, double inner_tol, double outer_tol,
                        int64_t inner_itmax, double safety, bool allow_short, hipStream_t st) {
    hipLaunchKernelGGL(k_inner_begin, dim3(1), dim3(64), 0, st, outer, inner, inner_tol,
                       outer_tol, inner_itmax, safety, (int32_t)allow_short);
}

void launch_step_begin(Scal *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, st, sc);
}

void launch_sum2(const Dims &g, const double *a, const double *b, double *out, hipStream_t st) {
    HF_LAUNCH(k_sum2, g, st, g, a, b, out);
}
void launch_extrap(const Dims &g, const ExtrapArgs &a, double *out, hipStream_t st) {
    HF_LAUNCH(k_extrap, g, st, g, a, out);
}
void launch_tiny_from_state(const Dims &g, const double *tp, const double *tpp, Scal *sc,
                            hipStream_t st) {
    HF_LAUNCH(k_tiny_from_state, g, st, g, tp, tpp, sc);
}

}  // namespace hf
