// hf_persist.hip -- FP64 time steps of a small grid in one persistent workgroup.
//
// On a small grid (kSmallMaxCells cells) a step of the multi-kernel path is bound by launch and
// host round-trip latency, not by memory: a CG iteration is four launches and every solve ends
// in a host read of the solver scalars.  Here one workgroup of kSmallThreads threads runs whole
// time steps -- Qdens of time-dependent heating, S and the residual, the CG iterations with
// their reductions, x, the NaN guards and the history rotation -- for up to kSmallMaxSteps steps
// per launch, with workgroup barriers in place of kernel boundaries and the convergence
// decisions taken on the device.  Every cell's arrays stay in the CU's L1 / the XCD's L2.
//
// The arithmetic is that of the multi-kernel FP64 step (evolution::simulate,
// src/heatflow/mod_evolve.f90:45-177, as CG on the Jacobi-scaled operator): the same per-cell
// formulas and stencil sum orders, the same stop rules (||r^||^2 <= tol^2 ||b^||^2, itmax, the
// first-iteration stop from the predicted norm, breakdown), the same TINY rule; only the
// grouping of the dot-product sums differs (one workgroup instead of per-block partials).
#include <hip/hip_runtime.h>

#include "hf_kernels.hpp"

namespace hf {

// S - modes as k_rhs_res_z: 0 steady, 1 Fourier, 2 Cattaneo
template <int Mode, bool Dyn>
__global__ void __launch_bounds__(kSmallThreads) k_run_small(SmallRun a) {
    __shared__ double red[3 * kMaxWaves];
    __shared__ int32_t sfl;
    const Dims &g = a.g;
    const uint32_t N = g.N;
    const uint32_t tid = threadIdx.x;
    const bool per_x = g.per[0], per_y = g.per[1], per_z = g.per[2];
    int cur = 0;   // t[cur] = T^n, t[(cur + 1) % 3] = T^{n-1}, t[(cur + 2) % 3] = spare
    const int32_t f_in = a.sc->flags;
    bool tiny = (f_in & FLAG_TINY_NEXT) != 0;
    int64_t steps = 0, iters = 0, last_iter = 0, first_stops = 0;
    double last_err = 0.0;
    int32_t stop_flags = 0;
    for (int32_t step = 0; step < a.nsteps; ++step) {
        const double *__restrict__ tp = a.t[cur];
        const double *__restrict__ tpp = a.t[(cur + 1) % 3];
        double *__restrict__ xo = a.t[(cur + 2) % 3];
        __syncthreads();   // every thread is past its reads of sfl (and of the last step's x)
        if (tid == 0) sfl = 0;
        __syncthreads();
        int32_t fl = 0;
        // ---- Qdens of this step from T^n (k_qdens; mod_heating.f90:27-169)
        if (Dyn) {
            const HeatParams &hp = a.hp[step];
            for (uint32_t I = tid; I < N; I += kSmallThreads) {
                double q = 0.0, qdv = 0.0;
                if (hp.any_heater) {
                    q = heater_q(hp, a.iheater[I], a.tau_s[I], a.ells[I], tp[I]);
                    qdv = hp.heated_volume > 0.0 ? q / hp.heated_volume : 0.0;
                    if (isnan(q)) fl |= FLAG_NAN_Q;
                    if (isnan(qdv)) fl |= FLAG_NAN_QDENS;
                }
                if (a.qextra) qdv = qdv + a.qextra[I];
                a.qd[I] = qdv;
            }
            __syncthreads();
        }
        // ---- S, x0, r^ = -s (S - H x0), ||b^||^2, ||r^||^2 (k_rhs_res_z; mod_evolve.f90:112-141)
        auto X = [&](uint32_t J) -> double {
            double x0 = tp[J] + (tp[J] - tpp[J]);
            if (tiny) x0 = x0 + kTiny;
            return x0;
        };
        double acc[2] = {0.0, 0.0};
        for (uint32_t I = tid; I < N; I += kSmallThreads) {
            const Nbr n = neighbours(g, I);
            double S;
            const double tpi = tp[I];
            if (Mode != kSteady) {
                S = -(a.inverse_time * tpi * a.rhoc[I & a.rc_mask]) - a.qd[I] - a.bvec[I];
                if (Mode == kCattaneo) {
                    const double scat = (tpp[I] - 2.0 * tpi) * a.cvec[I & a.rc_mask];
                    if (isnan(scat)) fl |= FLAG_NAN_SCAT;
                    S = S + scat;
                }
            } else {
                S = -a.qd[I] - a.bvec[I];
            }
            if (isnan(S)) fl |= FLAG_NAN_S;
            // H x0 in the reference's serial order: z-, y-, x-, diag, x+, y+, z+, wrap terms last
            const bool xm_in = n.x > 0, xp_in = n.x + 1 < g.nx;
            const bool ym_in = n.y > 0, yp_in = n.y + 1 < g.ny;
            const bool zm_in = n.z > 0, zp_in = n.z + 1 < g.nz;
            const double czm = n.hzm ? a.cz[n.zm] : 0.0;
            const double cym = n.hym ? a.cy[n.ym] : 0.0;
            const double cxm = n.hxm ? a.cx[n.xm] : 0.0;
            const double xi = X(I);
            double h = 0.0;
            h = h + czm * (zm_in ? X(n.zm) : 0.0);
            h = h + cym * (ym_in ? X(n.ym) : 0.0);
            h = h + cxm * (xm_in ? X(n.xm) : 0.0);
            h = h + a.d[I] * xi;
            h = h + a.cx[I] * (xp_in ? X(n.xp) : 0.0);
            h = h + a.cy[I] * (yp_in ? X(n.yp) : 0.0);
            h = h + a.cz[I] * (zp_in ? X(n.zp) : 0.0);
            if (per_x && !xm_in) h = h + cxm * X(n.xm);
            if (per_x && !xp_in) h = h + a.cx[I] * X(n.xp);
            if (per_y && !ym_in) h = h + cym * X(n.ym);
            if (per_y && !yp_in) h = h + a.cy[I] * X(n.yp);
            if (per_z && !zm_in) h = h + czm * X(n.zm);
            if (per_z && !zp_in) h = h + a.cz[I] * X(n.zp);
            const double r = S - h;
            const double s = 1.0 / sqrt(-a.d[I]);
            const double rh = -s * r;
            a.r[I] = rh;
            const double bh = s * S;
            acc[0] += bh * bh;
            acc[1] += rh * rh;
        }
        if (fl) atomicOr(&sfl, fl);
        block_sum<2>(acc, red);   // its barriers also publish r^ and sfl
        const int32_t f_rhs = sfl;
        if (f_rhs & (FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT | FLAG_NAN_S)) {
            stop_flags = f_rhs;
            break;
        }
        // ---- CG on A^ y = r^ (linbcg, mod_Sparse.f90:372-434, as scaled CG)
        const double bnorm2 = acc[0];
        double rr = acc[1];
        const double thresh2 = a.tol * a.tol * bnorm2;
        bool done = (rr == 0.0) || isnan(rr) || a.itmax <= 0;
        if (isnan(rr)) stop_flags |= FLAG_BREAKDOWN;
        int64_t it = 0, k = 0, yterms = 0;
        double alpha = 0.0, beta = 0.0;
        bool pending = false, shortcut = false;   // the term alpha p_k is still to be added
        double *pold = a.p1, *pnew = a.p0;
        const double *ppend = nullptr;
        while (!done) {
            // p_k = r + beta p_{k-1} (p_0 = r^), q = A^ p_k with p_k of the neighbours formed
            // from r and p_{k-1}; y gains the previous iteration's term alpha_{k-1} p_{k-1}
            auto P = [&](uint32_t J) -> double {
                return k == 0 ? a.r[J] : a.r[J] + beta * pold[J];
            };
            double acc2[2] = {0.0, 0.0};
            for (uint32_t I = tid; I < N; I += kSmallThreads) {
                const Nbr n = neighbours(g, I);
                const double pi = P(I);
                double sm = 0.0;
                sm = sm + (n.hzm ? a.hz[n.zm] : 0.0) * (n.hzm ? P(n.zm) : 0.0);
                sm = sm + (n.hym ? a.hy[n.ym] : a.hy[I]) * (n.hym ? P(n.ym) : 0.0);
                sm = sm + (n.hxm ? a.hx[n.xm] : a.hx[I]) * (n.hxm ? P(n.xm) : 0.0);
                sm = sm + a.hx[I] * (n.hxp ? P(n.xp) : 0.0);
                sm = sm + a.hy[I] * (n.hyp ? P(n.yp) : 0.0);
                sm = sm + a.hz[I] * (n.hzp ? P(n.zp) : 0.0);
                const double qv = pi - sm;
                pnew[I] = pi;
                a.q[I] = qv;
                if (k >= 1) a.y[I] = yterms == 0 ? alpha * pold[I] : a.y[I] + alpha * pold[I];
                acc2[0] += pi * qv;
                acc2[1] += qv * qv;
            }
            block_sum<2>(acc2, red);
            if (k >= 1) {
                yterms += 1;
                pending = false;
            }
            const double pq = acc2[0];
            alpha = rr / pq;
            if (!(pq > 0.0)) {   // A^ is SPD: p.q <= 0 or NaN is a breakdown
                alpha = 0.0;
                stop_flags |= FLAG_BREAKDOWN;
                break;
            }
            if (k == 0 && a.allow_short && a.itmax >= 1) {
                // |r0 - alpha q|^2 = alpha^2 q.q - r.r when p = r0
                const double rr1 = alpha * alpha * acc2[1] - rr;
                if (rr1 <= thresh2) {
                    rr = rr1 > 0.0 ? rr1 : 0.0;
                    it = 1;
                    shortcut = true;
                    pending = true;
                    ppend = pnew;
                    break;
                }
            }
            // r -= alpha q, ||r||^2
            double acc1[1] = {0.0};
            for (uint32_t I = tid; I < N; I += kSmallThreads) {
                const double ri = a.r[I] - alpha * a.q[I];
                a.r[I] = ri;
                acc1[0] += ri * ri;
            }
            block_sum<1>(acc1, red);
            const double rrn = acc1[0];
            beta = rrn / rr;
            rr = rrn;
            it += 1;
            pending = true;
            ppend = pnew;
            if (rr <= thresh2 || it >= a.itmax || isnan(rr)) done = true;
            if (isnan(rr)) stop_flags |= FLAG_BREAKDOWN;
            double *t = pold;
            pold = pnew;
            pnew = t;
            k += 1;
        }
        // ---- x = x0 + s y over the spare buffer, NaN guard of x, TINY test of the next x0
        // (k_finalize; mod_evolve.f90:140-158)
        const bool have_y = yterms > 0 || pending;
        bool tiny_next = false;
        fl = 0;
        for (uint32_t I = tid; I < N; I += kSmallThreads) {
            const double tpi = tp[I];
            double x0 = tpi + (tpi - tpp[I]);
            if (tiny) x0 = x0 + kTiny;
            double x = x0;
            if (have_y) {
                double yv;
                if (pending)
                    yv = yterms == 0 ? alpha * ppend[I] : a.y[I] + alpha * ppend[I];
                else
                    yv = a.y[I];
                x = x0 + a.sv[I] * yv;
            }
            xo[I] = x;
            if (isnan(x)) fl |= FLAG_NAN_X;
            if ((x + (x - tpi)) - x < kTiny) fl |= FLAG_TINY_NEXT;
        }
        __syncthreads();   // every thread is past its read of f_rhs
        if (tid == 0) sfl = 0;
        __syncthreads();
        if (fl) atomicOr(&sfl, fl);
        __syncthreads();
        const int32_t f_x = sfl;
        tiny_next = (f_x & FLAG_TINY_NEXT) != 0;
        // rotation: T^{n-1} <- T^n, T^n <- x (mod_evolve.f90:169-170)
        cur = (cur + 2) % 3;
        tiny = tiny_next;
        steps += 1;
        iters += it;
        last_iter = it;
        if (shortcut) first_stops += 1;
        last_err = bnorm2 > 0.0 ? sqrt(rr / bnorm2) : 0.0;
        if (f_x & FLAG_NAN_X) {
            stop_flags |= FLAG_NAN_X;
            break;
        }
    }
    if (tid == 0) {
        const int32_t nan_bits = stop_flags & (FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT |
                                               FLAG_NAN_S | FLAG_NAN_X | FLAG_BREAKDOWN);
        const int32_t fo = nan_bits | (tiny ? FLAG_TINY_NEXT : 0);
        a.out->steps = steps;
        a.out->iters = iters;
        a.out->last_iter = last_iter;
        a.out->first_stops = first_stops;
        a.out->last_err = last_err;
        a.out->flags = fo;
        a.out->rot = cur;
        // the solver scalars as the next multi-kernel step reads them (step_begin)
        a.sc->flags = (f_in & FLAG_NAN_B) | fo;
        a.sc->flags_next = 0;
    }
}

void launch_run_small(const SmallRun &a, hipStream_t st) {
    const int mode = s_mode(a.isteady, a.icattaneo);
#define HF_RUN_SMALL(M)                                                                       \
    do {                                                                                      \
        if (a.dyn)                                                                            \
            hipLaunchKernelGGL((k_run_small<M, true>), dim3(1), dim3(kSmallThreads), 0, st, a); \
        else                                                                                  \
            hipLaunchKernelGGL((k_run_small<M, false>), dim3(1), dim3(kSmallThreads), 0, st, a); \
    } while (0)
    if (mode == kSteady) HF_RUN_SMALL(kSteady);
    else if (mode == kFourier) HF_RUN_SMALL(kFourier);
    else HF_RUN_SMALL(kCattaneo);
#undef HF_RUN_SMALL
}

}  // namespace hf
