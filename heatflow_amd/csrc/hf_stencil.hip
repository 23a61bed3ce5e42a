// hf_stencil.hip -- the 7-point stencil kernels of the time step, z-marching.
//
// A block is a tile_x x tile_y (x by y) tile of columns; each thread walks its column through a chunk
// of z-planes and keeps the stencil operand of planes z-1, z, z+1 -- and the z coupling of
// plane z-1 -- in registers, so the z neighbours are never re-read; the x and y neighbours of
// plane z come from the same tile rows (L1) or the adjacent tile (same XCD, L2).  Tiles are
// numbered x fastest, then y, then z-chunk, and hardware block h runs logical tile
// (h % 8) * (zblk / 8) + h / 8: each XCD owns a contiguous range of tiles, i.e. whole z-chunks.
//
// Sums run in the ascending-column order of the reference's serial SpMV (z-, y-, x-, diag,
// x+, y+, z+; periodic wrap terms last), the same as the per-cell formulation.
//
//   k_rhs_res_z  S, x0 = 2 T^n - T^{n-1} (+TINY), r^ = -s (S - H x0), norms  (mod_evolve.f90:112-141)
//   k_cg_a_z     p = r + beta p, q = (I - C^) p, p.q                        (mod_Sparse.f90:404-411)
//   k_ir_res_z   x' = x + s y, r^ = -s (S - H x'), ||r^||^2  (mixed-precision refinement)
#include "hf_device.hpp"
#include "hf_kernels.hpp"

namespace hf {

struct ZCol {
    uint32_t col, x, y, z0, z1;
    bool active;
    // x / y neighbours: existence and index offsets (constant along the column)
    bool hxm, hxp, hym, hyp;
    int32_t oxm, oxp, oym, oyp;   // signed offsets from I
    bool wxm, wxp, wym, wyp;      // the neighbour is a periodic wrap (larger/smaller index)
};

__device__ __forceinline__ ZCol zcolumn(const Dims &g) {
    ZCol c;
    const uint32_t h = blockIdx.x;
    const uint32_t lt = (h & 7u) * (g.zblk >> 3) + (h >> 3);
    const uint32_t per_plane = g.ztx * g.zty;
    const uint32_t zc = lt / per_plane;
    const uint32_t rem = lt - zc * per_plane;
    const uint32_t by = rem / g.ztx, bx = rem - by * g.ztx;
    c.x = bx * g.tile_x + (threadIdx.x & (g.tile_x - 1));
    c.y = by * g.tile_y + (threadIdx.x >> g.tile_xs);
    c.z0 = zc * g.zchunk;
    c.z1 = c.z0 + g.zchunk < g.nz ? c.z0 + g.zchunk : g.nz;
    c.active = lt < g.ztiles && c.x < g.nx && c.y < g.ny && c.z0 < g.nz;
    c.col = c.x + g.nx * c.y;
    c.hxm = c.x > 0 || g.per[0];
    c.hxp = c.x + 1 < g.nx || g.per[0];
    c.hym = c.y > 0 || g.per[1];
    c.hyp = c.y + 1 < g.ny || g.per[1];
    c.wxm = c.x == 0;
    c.wxp = c.x + 1 == g.nx;
    c.wym = c.y == 0;
    c.wyp = c.y + 1 == g.ny;
    c.oxm = c.wxm ? (int32_t)(g.nx - 1) : -1;
    c.oxp = c.wxp ? -(int32_t)(g.nx - 1) : 1;
    c.oym = c.wym ? (int32_t)((g.ny - 1) * g.nx) : -(int32_t)g.nx;
    c.oyp = c.wyp ? -(int32_t)((g.ny - 1) * g.nx) : (int32_t)g.nx;
    return c;
}

// Sum of the seven stencil terms of H x in the reference's serial order; xm_z/xp_z and cz_m are
// the z neighbours' operands and the z- coupling carried in registers.
template <class XF>
__device__ __forceinline__ double hx_sum(const Dims &g, const ZCol &c, uint32_t I, uint32_t z,
                                         double di, const double *__restrict__ cx,
                                         const double *__restrict__ cy, double czc, double xi,
                                         double xz_m, double xz_p, double cz_m, XF X) {
    const bool hzm = z > 0 || g.per[2], hzp = z + 1 < g.nz || g.per[2];
    const bool wzm = z == 0, wzp = z + 1 == g.nz;
    double h = 0.0;
    if (hzm && !wzm) h = h + cz_m * xz_m;
    if (c.hym && !c.wym) h = h + cy[I + c.oym] * X(I + c.oym);
    if (c.hxm && !c.wxm) h = h + cx[I + c.oxm] * X(I + c.oxm);
    h = h + di * xi;
    if (c.hxp && !c.wxp) h = h + cx[I] * X(I + c.oxp);
    if (c.hyp && !c.wyp) h = h + cy[I] * X(I + c.oyp);
    if (hzp && !wzp) h = h + czc * xz_p;
    // periodic wrap couplings
    if (c.hxm && c.wxm) h = h + cx[I + c.oxm] * X(I + c.oxm);
    if (c.hxp && c.wxp) h = h + cx[I] * X(I + c.oxp);
    if (c.hym && c.wym) h = h + cy[I + c.oym] * X(I + c.oym);
    if (c.hyp && c.wyp) h = h + cy[I] * X(I + c.oyp);
    if (hzm && wzm) h = h + cz_m * xz_m;
    if (hzp && wzp) h = h + czc * xz_p;
    return h;
}

// index of the z-1 / z+1 neighbour of (col, z) (periodic wrap; only used when it exists)
__device__ __forceinline__ uint32_t zm_index(const Dims &g, uint32_t col, uint32_t z) {
    return z > 0 ? col + g.nxny * (z - 1) : col + g.nxny * (g.nz - 1);
}
__device__ __forceinline__ uint32_t zp_index(const Dims &g, uint32_t col, uint32_t z) {
    return z + 1 < g.nz ? col + g.nxny * (z + 1) : col;
}

template <typename OutT, bool FromState, bool StoreS, bool StoreX>
__global__ void __launch_bounds__(1024)
k_rhs_res_z(Dims g, RhsParams rp, const double *__restrict__ tp, const double *__restrict__ tpp,
            const double *__restrict__ xin, const double *__restrict__ bin,
            const double *__restrict__ rhoc, const double *__restrict__ cvec,
            const double *__restrict__ qd, const double *__restrict__ bvec,
            const double *__restrict__ d, const double *__restrict__ cx,
            const double *__restrict__ cy, const double *__restrict__ cz,
            OutT *__restrict__ rhat, double *__restrict__ sout, double *__restrict__ xout,
            Scal *sc, double *partials) {
    const bool tiny = sc->tiny != 0;
    auto X = [&](uint32_t J) -> double {
        if (FromState) {
            double x0 = tp[J] + (tp[J] - tpp[J]);  // mod_evolve.f90:140
            if (tiny) x0 = x0 + kTiny;             // :141
            return x0;
        } else {
            return xin[J];
        }
    };
    double acc[2] = {0.0, 0.0};
    int32_t fl = 0;
    const ZCol c = zcolumn(g);
    if (c.active) {
        const bool per_z = g.per[2];
        const bool h0 = c.z0 > 0 || per_z;
        double xz_m = 0.0, cz_m = 0.0;
        if (h0) {
            const uint32_t J = zm_index(g, c.col, c.z0);
            xz_m = X(J);
            cz_m = cz[J];
        }
        uint32_t I = c.col + g.nxny * c.z0;
        double xi = X(I);
        for (uint32_t z = c.z0; z < c.z1; ++z, I += g.nxny) {
            const bool hzp = z + 1 < g.nz || per_z;
            const double xz_p = hzp ? X(zp_index(g, c.col, z)) : 0.0;
            double S;
            if (FromState) {
                // mod_evolve.f90:112-119, S_catS mod_cattaneo.f90:35-36
                if (rp.isteady == 0) {
                    S = -(rp.inverse_time * tp[I] * ldnt(&rhoc[I])) - ldnt(&qd[I]) - ldnt(&bvec[I]);
                    if (rp.icattaneo == 1) {
                        double scat = (tpp[I] - 2.0 * tp[I]) * ldnt(&cvec[I]);
                        if (isnan(scat)) fl |= FLAG_NAN_SCAT;
                        S = S + scat;
                    }
                } else {
                    S = -ldnt(&qd[I]) - ldnt(&bvec[I]);
                }
                if (isnan(S)) fl |= FLAG_NAN_S;
            } else {
                S = ldnt(&bin[I]);
            }
            const double di = ldnt(&d[I]);
            const double czc = ldnt(&cz[I]);
            const double hxv = hx_sum(g, c, I, z, di, cx, cy, czc, xi, xz_m, xz_p, cz_m, X);
            const double r = S - hxv;
            const double s = 1.0 / sqrt(-di);
            const double rh = -s * r;
            stnt(&rhat[I], (OutT)rh);
            if (StoreS) stnt(&sout[I], S);
            if (StoreX) stnt(&xout[I], xi);
            const double bh = s * S;
            const double rq = (double)(OutT)rh;   // r.r of the stored (possibly FP32) residual
            acc[0] += bh * bh;
            acc[1] += rq * rq;
            xz_m = xi;
            xi = xz_p;
            cz_m = czc;
        }
    }
    block_flags(fl, &sc->flags);
    block_partials<2>(acc, partials);
}

// p = r + beta * p_old (p = r on the first iteration), q = A^ p, partial p.q.
template <typename T>
__global__ void __launch_bounds__(1024)
k_cg_a_z(Dims g, const T *__restrict__ r, const T *__restrict__ pold, T *__restrict__ pnew,
         T *__restrict__ q, const T *__restrict__ hx, const T *__restrict__ hy,
         const T *__restrict__ hz, Scal *sc, double *partials) {
    if (sc->done) return;
    const bool first = sc->first != 0;
    const T beta = (T)sc->beta;
    auto P = [&](uint32_t J) -> T { return first ? r[J] : (T)(r[J] + beta * pold[J]); };
    double acc[1] = {0.0};
    const ZCol c = zcolumn(g);
    if (c.active) {
        const bool per_z = g.per[2];
        T pz_m = (T)0, hz_m = (T)0;
        if (c.z0 > 0 || per_z) {
            const uint32_t J = zm_index(g, c.col, c.z0);
            pz_m = P(J);
            hz_m = hz[J];
        }
        uint32_t I = c.col + g.nxny * c.z0;
        T pi = P(I);
        for (uint32_t z = c.z0; z < c.z1; ++z, I += g.nxny) {
            const bool hzm = z > 0 || per_z, hzp = z + 1 < g.nz || per_z;
            const T pz_p = hzp ? P(zp_index(g, c.col, z)) : (T)0;
            const T hzc = ldnt(&hz[I]);
            T s = (T)0;
            if (hzm) s = s + hz_m * pz_m;
            if (c.hym) s = s + hy[I + c.oym] * P(I + c.oym);
            if (c.hxm) s = s + hx[I + c.oxm] * P(I + c.oxm);
            if (c.hxp) s = s + hx[I] * P(I + c.oxp);
            if (c.hyp) s = s + hy[I] * P(I + c.oyp);
            if (hzp) s = s + hzc * pz_p;
            const T qi = pi - s;
            stnt(&pnew[I], pi);
            stnt(&q[I], qi);
            acc[0] += (double)pi * (double)qi;
            pz_m = pi;
            pi = pz_p;
            hz_m = hzc;
        }
    }
    block_partials<1>(acc, partials);
}

// Mixed-precision refinement: x' = x + s y (y FP32) formed on the fly, r^ = -s (S - H x')
// stored as FP32 for the next inner solve, ||r^||^2, NaN and next-step TINY tests on x'.
__global__ void __launch_bounds__(1024)
k_ir_res_z(Dims g, IrParams ip, const double *__restrict__ x, const float *__restrict__ y,
           const double *__restrict__ sv, const double *__restrict__ S,
           const double *__restrict__ d, const double *__restrict__ cx,
           const double *__restrict__ cy, const double *__restrict__ cz,
           const double *__restrict__ tp, double *__restrict__ xout, float *__restrict__ rhat,
           Scal *sc, double *partials) {
    auto X = [&](uint32_t J) -> double { return x[J] + sv[J] * (double)y[J]; };
    double acc[1] = {0.0};
    int32_t fl = 0;
    const ZCol c = zcolumn(g);
    if (c.active) {
        const bool per_z = g.per[2];
        double xz_m = 0.0, cz_m = 0.0;
        if (c.z0 > 0 || per_z) {
            const uint32_t J = zm_index(g, c.col, c.z0);
            xz_m = X(J);
            cz_m = cz[J];
        }
        uint32_t I = c.col + g.nxny * c.z0;
        double xi = X(I);
        for (uint32_t z = c.z0; z < c.z1; ++z, I += g.nxny) {
            const bool hzp = z + 1 < g.nz || per_z;
            const double xz_p = hzp ? X(zp_index(g, c.col, z)) : 0.0;
            const double czc = ldnt(&cz[I]);
            const double hxv = hx_sum(g, c, I, z, ldnt(&d[I]), cx, cy, czc, xi, xz_m, xz_p, cz_m, X);
            const double rh = -sv[I] * (ldnt(&S[I]) - hxv);
            const float rf = (float)rh;
            stnt(&rhat[I], rf);
            stnt(&xout[I], xi);
            acc[0] += (double)rf * (double)rf;
            if (isnan(xi)) fl |= FLAG_NAN_X;
            if (ip.from_state && (xi + (xi - ldnt(&tp[I]))) - xi < kTiny) fl |= FLAG_TINY_NEXT;
            xz_m = xi;
            xi = xz_p;
            cz_m = czc;
        }
    }
    block_flags(fl, &sc->flags);
    block_partials<1>(acc, partials);
}

// ---- launchers -----------------------------------------------------------------------------

#define HF_ZLAUNCH(kern, g, st, ...) \
    hipLaunchKernelGGL(kern, dim3((g).zblk), dim3((g).tile_x * (g).tile_y), 0, st, __VA_ARGS__)

void launch_rhs_res_state64(const Dims &g, const RhsParams &rp, const double *tp,
                            const double *tpp, const double *rhoc, const double *cvec,
                            const double *qd, const double *bvec, const double *d,
                            const double *cx, const double *cy, const double *cz, double *rhat,
                            double *sout, Scal *sc, double *partials, hipStream_t st) {
    if (sout)
        HF_ZLAUNCH((k_rhs_res_z<double, true, true, false>), g, st, g, rp, tp, tpp, nullptr,
                   nullptr, rhoc, cvec, qd, bvec, d, cx, cy, cz, rhat, sout, nullptr, sc, partials);
    else
        HF_ZLAUNCH((k_rhs_res_z<double, true, false, false>), g, st, g, rp, tp, tpp, nullptr,
                   nullptr, rhoc, cvec, qd, bvec, d, cx, cy, cz, rhat, nullptr, nullptr, sc,
                   partials);
}

void launch_rhs_res_state32(const Dims &g, const RhsParams &rp, const double *tp,
                            const double *tpp, const double *rhoc, const double *cvec,
                            const double *qd, const double *bvec, const double *d,
                            const double *cx, const double *cy, const double *cz, float *rhat,
                            double *sout, double *xout, Scal *sc, double *partials,
                            hipStream_t st) {
    HF_ZLAUNCH((k_rhs_res_z<float, true, true, true>), g, st, g, rp, tp, tpp, nullptr, nullptr,
               rhoc, cvec, qd, bvec, d, cx, cy, cz, rhat, sout, xout, sc, partials);
}

void launch_rhs_res_solve64(const Dims &g, const RhsParams &rp, const double *x,
                            const double *b, const double *d, const double *cx,
                            const double *cy, const double *cz, double *rhat, Scal *sc,
                            double *partials, hipStream_t st) {
    HF_ZLAUNCH((k_rhs_res_z<double, false, false, false>), g, st, g, rp, nullptr, nullptr, x, b,
               nullptr, nullptr, nullptr, nullptr, d, cx, cy, cz, rhat, nullptr, nullptr, sc,
               partials);
}

void launch_rhs_res_solve32(const Dims &g, const RhsParams &rp, const double *x,
                            const double *b, const double *d, const double *cx,
                            const double *cy, const double *cz, float *rhat, Scal *sc,
                            double *partials, hipStream_t st) {
    HF_ZLAUNCH((k_rhs_res_z<float, false, false, false>), g, st, g, rp, nullptr, nullptr, x, b,
               nullptr, nullptr, nullptr, nullptr, d, cx, cy, cz, rhat, nullptr, nullptr, sc,
               partials);
}

template <typename T>
void launch_cg_a(const Dims &g, const T *r, const T *pold, T *pnew, T *q, const T *hx,
                 const T *hy, const T *hz, Scal *sc, double *partials, hipStream_t st) {
    HF_ZLAUNCH(k_cg_a_z<T>, g, st, g, r, pold, pnew, q, hx, hy, hz, sc, partials);
}
template void launch_cg_a<double>(const Dims &, const double *, const double *, double *,
                                  double *, const double *, const double *, const double *,
                                  Scal *, double *, hipStream_t);
template void launch_cg_a<float>(const Dims &, const float *, const float *, float *, float *,
                                 const float *, const float *, const float *, Scal *, double *,
                                 hipStream_t);

void launch_ir_res(const Dims &g, const IrParams &ip, const double *x, const float *y,
                   const double *sv, const double *S, const double *d, const double *cx,
                   const double *cy, const double *cz, const double *tp, double *xout,
                   float *rhat, Scal *sc, double *partials, hipStream_t st) {
    HF_ZLAUNCH(k_ir_res_z, g, st, g, ip, x, y, sv, S, d, cx, cy, cz, tp, xout, rhat, sc, partials);
}

}  // namespace hf
