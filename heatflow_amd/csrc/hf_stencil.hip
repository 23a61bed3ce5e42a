// hf_stencil.hip -- the 7-point stencil kernels of the time step: z-marching with the operand
// plane staged in LDS.
//
// A block of kZTX x kZTY threads owns a tile of columns and walks it through a chunk of
// z-planes.  Each thread keeps the stencil operand of planes z-1, z, z+1 (and the z coupling of
// plane z-1) of its own column in registers, so z neighbours are never re-read.  The operand of
// plane z, with a one-cell halo ring around the tile, sits in LDS: every cell of the tile and
// ring is loaded once per plane and the four in-plane neighbours are read from LDS.  Two LDS
// buffers alternate, so plane z+1 is staged while plane z is computed, with one barrier per
// plane.  Periodic wrap neighbours (boundary columns of a periodic axis) are read from memory.
// The kernels with many inputs per cell (k_rhs_res_z, k_ir_step_z) are software-pipelined: the
// next plane's loads are issued before the barrier of the current one (zmarch_pipe).
// Tiles are numbered x fastest, then y, then z-chunk; hardware block h runs logical tile
// (h % 8) * (T / 8) + h / 8 (T: the kernel's tile count rounded up to a multiple of 8), so
// each XCD owns a contiguous range of tiles (whole z-chunks) and the ring rows of neighbouring
// tiles come from the same L2.
//
// Sums run in the ascending-column order of the reference's serial SpMV (z-, y-, x-, diag,
// x+, y+, z+; periodic wrap terms last), the same as the per-cell formulation.
//
//   k_rhs_res_z  S, x0 = 2 T^n - T^{n-1} (+TINY), r^ = -s (S - H x0), norms  (mod_evolve.f90:112-141)
//   k_cg_a_z     p = r + beta p, q = (I - C^) p, p.q                        (mod_Sparse.f90:404-411)
//   k_cg_b1_z    the first iteration's y, r updates with q recomputed
//   k_ir_res_z   x' = x + s y, r^ = -s (S - H x'), ||r^||^2  (mixed-precision refinement)
//   k_ir_step_z  the refinement residual fused with the next step's right-hand side
//   k_step64_z   k_finalize fused with the next step's k_rhs_res_z (FP64 path)
#include <type_traits>

#include "hf_device.hpp"
#include "hf_kernels.hpp"

namespace hf {

struct ZCol {
    uint32_t col, x, y, z0, z1;
    uint32_t bx, by, lx, ly;      // tile coordinates, thread position in the tile
    bool active;                  // this thread's column is inside the grid
    bool zvalid;                  // this block has planes to walk
    // x / y neighbours: existence and index offsets (constant along the column)
    bool hxm, hxp, hym, hyp;
    int32_t oxm, oxp, oym, oyp;   // signed offsets from I
    bool wxm, wxp, wym, wyp;      // the neighbour is a periodic wrap (larger/smaller index)
};

// Logical tile of hardware block h: the kernel's own `ztiles` tiles (rounded up to a multiple
// of 8) are spread over the 8 XCDs in contiguous ranges; blocks past them (the launch grid is
// the largest tile count of the chunking class) get no tile.
__device__ __forceinline__ uint32_t logical_tile(uint32_t ztiles) {
    const uint32_t h = blockIdx.x;
    const uint32_t tr = (ztiles + 7u) & ~7u;
    return h < tr ? (h & 7u) * (tr >> 3) + (h >> 3) : ztiles;
}

__device__ __forceinline__ ZCol zcolumn(const Dims &g) {
    ZCol c;
    const uint32_t lt = logical_tile(g.ztiles);
    const uint32_t per_plane = g.ztx * g.zty;
    const uint32_t zc = lt / per_plane;
    const uint32_t rem = lt - zc * per_plane;
    c.by = rem / g.ztx;
    c.bx = rem - c.by * g.ztx;
    c.lx = threadIdx.x % kZTX;
    c.ly = threadIdx.x / kZTX;
    c.x = c.bx * kZTX + c.lx;
    c.y = c.by * kZTY + c.ly;
    c.z0 = zc * g.zchunk;
    c.z1 = c.z0 + g.zchunk < g.nz ? c.z0 + g.zchunk : g.nz;
    c.zvalid = lt < g.ztiles && c.z0 < g.nz;
    c.active = c.zvalid && c.x < g.nx && c.y < g.ny;
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

// LDS planes of the operand: two buffers of (kZTY + 2) x (kZTX + 2) values (tile + halo ring).
template <typename V>
struct PlaneBuf {
    V v[2][kZTY + 2][kZTX + 2];
};

// In-plane neighbours of a cell, read from the staged LDS plane when used (c points at the
// cell's own slot); meaningful where the neighbour exists and is not a periodic wrap.
template <typename V>
struct InPlane {
    const V *c;
    __device__ __forceinline__ V xm() const { return c[-1]; }
    __device__ __forceinline__ V xp() const { return c[1]; }
    __device__ __forceinline__ V ym() const { return c[-(int)(kZTX + 2)]; }
    __device__ __forceinline__ V yp() const { return c[(int)(kZTX + 2)]; }
};

// A pair of stencil operands staged together (the fused step: x' and the next x0).
struct D2 {
    double a, b;
    __device__ D2() = default;
    __device__ constexpr D2(double v) : a(v), b(v) {}
    __device__ constexpr D2(double x, double y) : a(x), b(y) {}
};

// 1/sqrt(v) within an ulp: the hardware estimate y (v_rsq_f64; on gfx950 its relative error
// measured at most 5.2e-8 over 2^20 values from 1e-300 to 1e301) refined by one third-order
// Newton (Householder) step, e = 1 - v y^2, y' = y + y e (1/2 + 3/8 e), which leaves an error
// of order e^3 before the last roundings: at most 0.99 ulp and 86% correctly rounded over the
// same values, against 1.48 ulp and 73% for 1.0 / sqrt(v).  Where e is NaN (v zero, infinite,
// negative or NaN) the estimate stands: inf, 0 or NaN, as 1/sqrt gives.  Nine instructions,
// against about thirty for a rounded sqrt and a divide.
__device__ __forceinline__ double inv_sqrt(double v) {
    const double y = __builtin_amdgcn_rsq(v);
    const double e = __builtin_fma(-(v * y), y, 1.0);
    const double r = __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
    return e == e ? r : y;
}

// One component of the in-plane neighbours of a staged pair.
template <int K>
struct InPlaneOf {
    const D2 *c;
    __device__ __forceinline__ static double pick(const D2 &v) { return K == 0 ? v.a : v.b; }
    __device__ __forceinline__ double xm() const { return pick(c[-1]); }
    __device__ __forceinline__ double xp() const { return pick(c[1]); }
    __device__ __forceinline__ double ym() const { return pick(c[-(int)(kZTX + 2)]); }
    __device__ __forceinline__ double yp() const { return pick(c[(int)(kZTX + 2)]); }
};

// The halo-ring cell a thread stages (threads < kZRing; corners are not needed): its slot in
// the LDS plane and its column, or kNoRing.  Ring cells outside the grid are not loaded: a
// wrap neighbour is read from memory by its user.
struct Ring {
    int slot;
    uint32_t col;
};
constexpr uint32_t kNoRing = 0xffffffffu;

constexpr int kZRing = 2 * (int)kZTX + 2 * (int)kZTY;
static_assert(kZRing <= (int)(kZTX * kZTY), "ring loader needs one thread per ring cell");

__device__ __forceinline__ Ring ring_cell(const Dims &g, const ZCol &c) {
    Ring r{0, kNoRing};
    const int t = (int)threadIdx.x;
    const int TX = (int)kZTX, TY = (int)kZTY;
    if (t >= kZRing || !c.zvalid) return r;
    int lx, ly;
    if (t < TX) {
        lx = t + 1; ly = 0;
    } else if (t < 2 * TX) {
        lx = t - TX + 1; ly = TY + 1;
    } else if (t < 2 * TX + TY) {
        lx = 0; ly = t - 2 * TX + 1;
    } else {
        lx = TX + 1; ly = t - 2 * TX - TY + 1;
    }
    const int64_t gx = (int64_t)c.bx * TX + lx - 1, gy = (int64_t)c.by * TY + ly - 1;
    r.slot = ly * (TX + 2) + lx;
    if (gx >= 0 && gx < (int64_t)g.nx && gy >= 0 && gy < (int64_t)g.ny)
        r.col = (uint32_t)gx + g.nx * (uint32_t)gy;
    return r;
}

// index of the z-1 / z+1 neighbour of (col, z) (periodic wrap; only used when it exists)
__device__ __forceinline__ uint32_t zm_index(const Dims &g, uint32_t col, uint32_t z) {
    return z > 0 ? col + g.nxny * (z - 1) : col + g.nxny * (g.nz - 1);
}
__device__ __forceinline__ uint32_t zp_index(const Dims &g, uint32_t col, uint32_t z) {
    return z + 1 < g.nz ? col + g.nxny * (z + 1) : col;
}

// Walks this block's z-chunk.  X(J) is the stencil operand at cell J (it may combine several
// arrays); cell(I, z, xi, xz_m, xz_p, nb) runs for every active cell with its own operand, the
// z neighbours' operands and the in-plane neighbours from LDS.  Every thread of the block runs
// the same number of trips (one barrier each); inactive threads only stage.
template <typename V, class XF, class CF>
__device__ __forceinline__ void zmarch(const Dims &g, const ZCol &c, PlaneBuf<V> &sp, XF X,
                                       CF cell) {
    const Ring rg = ring_cell(g, c);
    const bool per_z = g.per[2];
    V xz_m = (V)0, xi = (V)0;
    if (c.active) {
        if (c.z0 > 0 || per_z) xz_m = X(zm_index(g, c.col, c.z0));
        xi = X(c.col + g.nxny * c.z0);
    }
    if (c.zvalid) {
        sp.v[0][c.ly + 1][c.lx + 1] = xi;
        if (rg.col != kNoRing) (&sp.v[0][0][0])[rg.slot] = X(rg.col + g.nxny * c.z0);
    }
    int buf = 0;
    uint32_t I = c.col + g.nxny * c.z0;
    for (uint32_t z = c.z0; z < c.z1; ++z, I += g.nxny) {
        const bool more = z + 1 < c.z1;
        V xz_p = (V)0, rnext = (V)0;
        if (c.active && (z + 1 < g.nz || per_z)) xz_p = X(zp_index(g, c.col, z));
        if (more && rg.col != kNoRing) rnext = X(rg.col + g.nxny * (z + 1));
        __syncthreads();   // plane z is staged in sp.v[buf]
        if (c.active) {
            const InPlane<V> nb{&sp.v[buf][c.ly + 1][c.lx + 1]};
            cell(I, z, xi, xz_m, xz_p, nb);
        }
        if (more) {
            // sp.v[buf ^ 1] was last read before this trip's barrier
            sp.v[buf ^ 1][c.ly + 1][c.lx + 1] = xz_p;
            if (rg.col != kNoRing) (&sp.v[buf ^ 1][0][0])[rg.slot] = rnext;
        }
        buf ^= 1;
        xz_m = xi;
        xi = xz_p;
    }
}

// zmarch that also hands each cell a value of its own (own(J): loads of the same addresses as
// X's, which the compiler shares), carried in registers from the load one plane ahead.
template <typename V, typename O, class XF, class OF, class CF>
__device__ __forceinline__ void zmarch_own(const Dims &g, const ZCol &c, PlaneBuf<V> &sp, XF X,
                                           OF own, CF cell) {
    const Ring rg = ring_cell(g, c);
    const bool per_z = g.per[2];
    V xz_m = (V)0, xi = (V)0;
    O oi{};
    if (c.active) {
        if (c.z0 > 0 || per_z) xz_m = X(zm_index(g, c.col, c.z0));
        const uint32_t J0 = c.col + g.nxny * c.z0;
        xi = X(J0);
        oi = own(J0);
    }
    if (c.zvalid) {
        sp.v[0][c.ly + 1][c.lx + 1] = xi;
        if (rg.col != kNoRing) (&sp.v[0][0][0])[rg.slot] = X(rg.col + g.nxny * c.z0);
    }
    int buf = 0;
    uint32_t I = c.col + g.nxny * c.z0;
    for (uint32_t z = c.z0; z < c.z1; ++z, I += g.nxny) {
        const bool more = z + 1 < c.z1;
        V xz_p = (V)0, rnext = (V)0;
        O op{};
        if (c.active && (z + 1 < g.nz || per_z)) {
            const uint32_t J = zp_index(g, c.col, z);
            xz_p = X(J);
            op = own(J);
        }
        if (more && rg.col != kNoRing) rnext = X(rg.col + g.nxny * (z + 1));
        __syncthreads();   // plane z is staged in sp.v[buf]
        if (c.active) {
            const InPlane<V> nb{&sp.v[buf][c.ly + 1][c.lx + 1]};
            cell(I, z, xi, xz_m, xz_p, nb, oi);
        }
        if (more) {
            sp.v[buf ^ 1][c.ly + 1][c.lx + 1] = xz_p;
            if (rg.col != kNoRing) (&sp.v[buf ^ 1][0][0])[rg.slot] = rnext;
        }
        buf ^= 1;
        xz_m = xi;
        xi = xz_p;
        oi = op;
    }
}

// zmarch, software-pipelined.  The operand is formed from a raw record R = load(J) by form(R);
// the cell's other inputs are a record C = cload(I) (I a cell of the thread's column, or 0 for
// an inactive thread).  While plane z is computed, the loads of plane z+1's cell record and
// halo ring and of plane z+2's own operand record are in flight, so a trip waits only on loads
// issued one trip earlier.  cell(I, z, xi, xz_m, xz_p, nb, cur, ri) runs for every active
// cell; ri is its own operand record of plane z.
// The next trips' loads are issued on every trip by every thread, without branches, so that
// the compiler's wait counts know how many are in flight; where there is nothing to load they
// read an address the thread read recently (cache hits): its own cell, or cell 0.  Two trips
// per loop iteration alternate the register sets that carry loads between trips (a copy at
// the end of a trip would wait for the loads just issued; the compiler does not unroll a loop
// with a barrier by itself).  Wrap = false: no axis is periodic.
// The own cell's operand is formed by fown(R, E &), which may also hand the cell values it
// computed on the way (E: the Jacobi scale, say), so the cell need not compute them again;
// cell(..., ri, ei) gets them with its raw record.
template <typename V, typename R, typename C, typename E, bool Wrap, class XL, class XF,
          class XO, class CL, class CF>
__device__ __forceinline__ void zmarch_pipe_own(const Dims &g, const ZCol &c, PlaneBuf<V> &sp,
                                                XL load, XF form, XO fown, CL cload, CF cell) {
    const Ring rg = ring_cell(g, c);
    const bool per_z = Wrap && g.per[2];
    V xz_m = (V)0, xi = (V)0;
    R rA{}, rB{}, ri{};
    C cA{}, cB{};
    E ei{};
    if (c.active) {
        if (c.z0 > 0 || per_z) xz_m = form(load(zm_index(g, c.col, c.z0)));
        ri = load(c.col + g.nxny * c.z0);
        xi = fown(ri, ei);
        if (c.z0 + 1 < g.nz || per_z) rA = load(zp_index(g, c.col, c.z0));
        cA = cload(c.col + g.nxny * c.z0);
    }
    if (c.zvalid) {
        sp.v[0][c.ly + 1][c.lx + 1] = xi;
        if (rg.col != kNoRing) (&sp.v[0][0][0])[rg.slot] = form(load(rg.col + g.nxny * c.z0));
    }
    int buf = 0;
    auto trip = [&](uint32_t z, uint32_t I, const R &ri, const R &r1, const C &cur, R &r2,
                    C &cn) {
        const bool more = z + 1 < c.z1;
        const bool has_p = c.active && (z + 1 < g.nz || per_z);
        const bool has_p2 = c.active && more && (z + 2 < g.nz || per_z);
        const bool has_r = more && rg.col != kNoRing;
        const uint32_t own = c.active ? I : 0u;
        r2 = load(has_p2 ? zp_index(g, c.col, z + 1) : own);
        // (opaque to the compiler, which would otherwise load under a branch on the ring test)
        uint32_t jr = rg.col != kNoRing ? rg.col + g.nxny * (more ? z + 1 : z) : 0u;
        asm volatile("" : "+v"(jr));
        const R rr = load(jr);
        cn = cload(c.active && more ? I + g.nxny : own);
        __syncthreads();   // plane z is staged in sp.v[buf]
        V xz_p = (V)0;
        E en{};
        if (c.active) {
            xz_p = has_p ? fown(r1, en) : (V)0;
            const InPlane<V> nb{&sp.v[buf][c.ly + 1][c.lx + 1]};
            cell(I, z, xi, xz_m, xz_p, nb, cur, ri, ei);
        }
        if (more) {
            // sp.v[buf ^ 1] was last read before this trip's barrier
            sp.v[buf ^ 1][c.ly + 1][c.lx + 1] = xz_p;
            if (has_r) (&sp.v[buf ^ 1][0][0])[rg.slot] = form(rr);
        }
        buf ^= 1;
        xz_m = xi;
        xi = xz_p;
        ei = en;
    };
    // Six trips per iteration: the own operand records of planes z, z+1, z+2 rotate through
    // three register sets and the cell records of planes z, z+1 through two, so every loaded
    // record keeps its registers from its load to its last use (a record that changed
    // registers would be copied at the loop's back edge, waiting for the loads just issued).
    const uint32_t P = g.nxny;
    uint32_t I = c.col + P * c.z0;
    for (uint32_t z = c.z0; z < c.z1; z += 6, I += 6 * P) {
        trip(z, I, ri, rA, cA, rB, cB);
        if (z + 1 >= c.z1) break;
        trip(z + 1, I + P, rA, rB, cB, ri, cA);
        if (z + 2 >= c.z1) break;
        trip(z + 2, I + 2 * P, rB, ri, cA, rA, cB);
        if (z + 3 >= c.z1) break;
        trip(z + 3, I + 3 * P, ri, rA, cB, rB, cA);
        if (z + 4 >= c.z1) break;
        trip(z + 4, I + 4 * P, rA, rB, cA, ri, cB);
        if (z + 5 >= c.z1) break;
        trip(z + 5, I + 5 * P, rB, ri, cB, rA, cA);
    }
}

struct NoOwn {};

template <typename V, typename R, typename C, bool Wrap, class XL, class XF, class CL, class CF>
__device__ __forceinline__ void zmarch_pipe(const Dims &g, const ZCol &c, PlaneBuf<V> &sp,
                                            XL load, XF form, CL cload, CF cell) {
    zmarch_pipe_own<V, R, C, NoOwn, Wrap>(
        g, c, sp, load, form, [&](const R &r, NoOwn &) -> V { return form(r); }, cload,
        [&](uint32_t I, uint32_t z, V xi, V xz_m, V xz_p, const InPlane<V> &nb, const C &cur,
            const R &ri, const NoOwn &) { cell(I, z, xi, xz_m, xz_p, nb, cur, ri); });
}

// Offsets from a cell to its in-grid x- / y- neighbours' couplings (stored at the neighbour;
// periodic wrap terms are added apart): 0 where the neighbour is absent or the thread
// inactive.  Opaque to the compiler, which would otherwise turn a load at I + offset into a
// load under a branch reusing the one at I (and a branch stops loads being issued ahead).
struct NbOff {
    int32_t xm, ym;
};
__device__ __forceinline__ NbOff nb_offsets_in(const Dims &g, const ZCol &c) {
    int32_t xm = c.active && c.x > 0 ? -1 : 0;
    int32_t ym = c.active && c.y > 0 ? -(int32_t)g.nx : 0;
    asm volatile("" : "+v"(xm), "+v"(ym));
    return NbOff{xm, ym};
}

// the couplings of H a cell needs: diag, its own x+ / y+ / z+ ones and those of its in-grid
// x- / y- neighbours (read at the cell itself where that neighbour is absent)
struct HCell {
    double d, cx, cy, cz, cxm, cym;
};
__device__ __forceinline__ HCell hcell_load(uint32_t I, const NbOff &o, const double *d,
                                            const double *cx, const double *cy,
                                            const double *cz) {
    return HCell{ldnt(&d[I]), cx[I], cy[I], ldnt(&cz[I]), cx[I + o.xm], cy[I + o.ym]};
}

// hx_sum with the couplings preloaded (same terms, same order; wrap terms read from memory)
template <bool Wrap, class NB, class XF>
__device__ __forceinline__ double hx_sum_v(const Dims &g, const ZCol &c, uint32_t I, uint32_t z,
                                           const HCell &hc, const double *__restrict__ cx,
                                           const double *__restrict__ cy, double xi, double xz_m,
                                           double xz_p, double cz_m, const NB &nb, XF X) {
    const bool ym_in = c.y > 0, yp_in = c.y + 1 < g.ny;
    const bool xm_in = c.x > 0, xp_in = c.x + 1 < g.nx;
    double h = 0.0;
    h = h + cz_m * (z > 0 ? xz_m : 0.0);
    h = h + hc.cym * (ym_in ? nb.ym() : 0.0);
    h = h + hc.cxm * (xm_in ? nb.xm() : 0.0);
    h = h + hc.d * xi;
    h = h + hc.cx * (xp_in ? nb.xp() : 0.0);
    h = h + hc.cy * (yp_in ? nb.yp() : 0.0);
    h = h + hc.cz * (z + 1 < g.nz ? xz_p : 0.0);
    if (!Wrap) return h;
    const bool hzm = z > 0 || g.per[2], hzp = z + 1 < g.nz || g.per[2];
    if (c.hxm && c.wxm) h = h + cx[I + c.oxm] * X(I + c.oxm);
    if (c.hxp && c.wxp) h = h + hc.cx * X(I + c.oxp);
    if (c.hym && c.wym) h = h + cy[I + c.oym] * X(I + c.oym);
    if (c.hyp && c.wyp) h = h + hc.cy * X(I + c.oyp);
    if (hzm && z == 0) h = h + cz_m * xz_m;
    if (hzp && z + 1 == g.nz) h = h + hc.cz * xz_p;
    return h;
}

// Sum of the seven stencil terms of H x in the reference's serial order: xz_m / xz_p and cz_m
// are the z neighbours' operands and the z- coupling carried in registers, nb the in-plane
// neighbours from LDS; periodic wrap terms (last) read the operand X from memory.  Wrap =
// false: no axis is periodic (the wrap terms are left out).
template <bool Wrap = true, class NB, class XF>
__device__ __forceinline__ double hx_sum(const Dims &g, const ZCol &c, uint32_t I, uint32_t z,
                                         double di, const double *__restrict__ cx,
                                         const double *__restrict__ cy, double czc, double xi,
                                         double xz_m, double xz_p, double cz_m, const NB &nb,
                                         XF X) {
    const bool hzm = z > 0 || g.per[2], hzp = z + 1 < g.nz || g.per[2];
    const bool wzm = z == 0, wzp = z + 1 == g.nz;
    // Terms of neighbours inside the grid, without branches: an absent neighbour contributes
    // coefficient * 0 = +-0, which leaves the sum unchanged, and its coefficient is read at I
    // (a valid address).  Wrap terms, last, keep their branches (periodic axes only).
    const bool ym_in = c.y > 0, yp_in = c.y + 1 < g.ny;
    const bool xm_in = c.x > 0, xp_in = c.x + 1 < g.nx;
    double h = 0.0;
    h = h + cz_m * (z > 0 ? xz_m : 0.0);
    h = h + cy[ym_in ? I - g.nx : I] * (ym_in ? nb.ym() : 0.0);
    h = h + cx[xm_in ? I - 1 : I] * (xm_in ? nb.xm() : 0.0);
    h = h + di * xi;
    h = h + cx[I] * (xp_in ? nb.xp() : 0.0);
    h = h + cy[I] * (yp_in ? nb.yp() : 0.0);
    h = h + czc * (z + 1 < g.nz ? xz_p : 0.0);
    if (!Wrap) return h;
    // periodic wrap couplings
    if (c.hxm && c.wxm) h = h + cx[I + c.oxm] * X(I + c.oxm);
    if (c.hxp && c.wxp) h = h + cx[I] * X(I + c.oxp);
    if (c.hym && c.wym) h = h + cy[I + c.oym] * X(I + c.oym);
    if (c.hyp && c.wyp) h = h + cy[I] * X(I + c.oyp);
    if (hzm && wzm) h = h + cz_m * xz_m;
    if (hzp && wzp) h = h + czc * xz_p;
    return h;
}

// Mode of S: kSteady, kFourier, kCattaneo (hf_kernels.hpp).  Kernels built for a solve (not
// FromState) use kSteady.

// operand inputs of k_rhs_res_z: T^n, T^{n-1} (state) or x (solve)
struct RhsRaw {
    double a, b;
};
// the cell's other inputs: S inputs (state: rho c_p, tau' rho c_p, Qdens, B; solve: b in qd)
// and the couplings
struct RhsCell {
    double rc, cv, qd, bv;
    HCell h;
};

// (software-pipelined z-march; Wrap = some axis is periodic)
template <typename OutT, bool FromState, bool StoreS, bool StoreX, int Mode, bool Wrap>
__global__ void __launch_bounds__(kZTX * kZTY)
k_rhs_res_z(Dims g, RhsParams rp, const double *__restrict__ tp, const double *__restrict__ tpp,
            const double *__restrict__ xin, const double *__restrict__ bin,
            const double *__restrict__ rhoc, const double *__restrict__ cvec,
            const double *__restrict__ qd, const double *__restrict__ bvec,
            const double *__restrict__ d, const double *__restrict__ cx,
            const double *__restrict__ cy, const double *__restrict__ cz,
            OutT *__restrict__ rhat, double *__restrict__ sout, double *__restrict__ xout,
            Scal *sc, double *partials) {
    __shared__ PlaneBuf<double> sp;
    const bool tiny = sc->tiny != 0;
    auto load = [&](uint32_t J) -> RhsRaw {
        return FromState ? RhsRaw{tp[J], tpp[J]} : RhsRaw{xin[J], 0.0};
    };
    auto form = [&](const RhsRaw &v) -> double {
        if (FromState) {
            double x0 = v.a + (v.a - v.b);   // mod_evolve.f90:140
            if (tiny) x0 = x0 + kTiny;       // :141
            return x0;
        } else {
            return v.a;
        }
    };
    auto X = [&](uint32_t J) -> double { return form(load(J)); };
    double acc[2] = {0.0, 0.0};
    int32_t fl = 0;
    const ZCol c = zcolumn(g);
    const NbOff o = nb_offsets_in(g, c);
    auto cload = [&](uint32_t I) -> RhsCell {
        RhsCell v;
        v.rc = (FromState && Mode != kSteady) ? ldnt(&rhoc[I & rp.rc_mask]) : 0.0;
        v.cv = (FromState && Mode == kCattaneo) ? ldnt(&cvec[I & rp.rc_mask]) : 0.0;
        v.qd = FromState ? ldnt(&qd[I]) : ldnt(&bin[I]);
        v.bv = FromState ? ldnt(&bvec[I]) : 0.0;
        v.h = hcell_load(I, o, d, cx, cy, cz);
        return v;
    };
    double cz_m = 0.0;
    if (c.active && (c.z0 > 0 || (Wrap && g.per[2]))) cz_m = cz[zm_index(g, c.col, c.z0)];
    zmarch_pipe<double, RhsRaw, RhsCell, Wrap>(
        g, c, sp, load, form, cload,
        [&](uint32_t I, uint32_t z, double xi, double xz_m, double xz_p,
            const InPlane<double> &nb, const RhsCell &cv, const RhsRaw &ri) {
            double S;
            if (FromState) {
                // mod_evolve.f90:112-119, S_catS mod_cattaneo.f90:35-36
                if (Mode != kSteady) {
                    S = -(rp.inverse_time * ri.a * cv.rc) - cv.qd - cv.bv;
                    if (Mode == kCattaneo) {
                        const double scat = (ri.b - 2.0 * ri.a) * cv.cv;
                        if (isnan(scat)) fl |= FLAG_NAN_SCAT;
                        S = S + scat;
                    }
                } else {
                    S = -cv.qd - cv.bv;
                }
                if (isnan(S)) fl |= FLAG_NAN_S;
            } else {
                S = cv.qd;
            }
            const double hxv = hx_sum_v<Wrap>(g, c, I, z, cv.h, cx, cy, xi, xz_m, xz_p, cz_m, nb,
                                              X);
            const double r = S - hxv;
            const double s = 1.0 / sqrt(-cv.h.d);
            const double rh = -s * r;
            stnt(&rhat[I], (OutT)rh);
            if (StoreS) stnt(&sout[I], S);
            if (StoreX) stnt(&xout[I], xi);
            const double bh = s * S;
            const double rq = (double)(OutT)rh;   // r.r of the stored (possibly FP32) residual
            acc[0] += bh * bh;
            acc[1] += rq * rq;
            cz_m = cv.h.cz;
        });
    block_flags(fl, &sc->flags);
    block_partials<2>(acc, partials);
}

// q = A^ p = p - C^ p with the z coupling of plane z-1 carried in hz_m (order z-, y-, x-, x+,
// y+, z+; periodic wrap neighbours read through P from memory)
template <typename T, class PF>
__device__ __forceinline__ T ahat(const Dims &g, const ZCol &c, uint32_t I, uint32_t z, T pi,
                                  T pz_m, T pz_p, T hz_m, T hzc, const T *__restrict__ hx,
                                  const T *__restrict__ hy, const InPlane<T> &nb, PF P) {
    const bool per_z = g.per[2];
    const bool hzm = z > 0 || per_z, hzp = z + 1 < g.nz || per_z;
    // absent neighbours contribute coefficient * 0 (coefficient read at I); wrap neighbours are
    // read from memory
    T vym = (T)0, vxm = (T)0, vxp = (T)0, vyp = (T)0;
    if (c.hym) vym = c.wym ? P(I + c.oym) : nb.ym();
    if (c.hxm) vxm = c.wxm ? P(I + c.oxm) : nb.xm();
    if (c.hxp) vxp = c.wxp ? P(I + c.oxp) : nb.xp();
    if (c.hyp) vyp = c.wyp ? P(I + c.oyp) : nb.yp();
    T s = (T)0;
    s = s + hz_m * (hzm ? pz_m : (T)0);
    s = s + hy[c.hym ? I + c.oym : I] * vym;
    s = s + hx[c.hxm ? I + c.oxm : I] * vxm;
    s = s + hx[I] * vxp;
    s = s + hy[I] * vyp;
    s = s + hzc * (hzp ? pz_p : (T)0);
    return pi - s;
}

// CG iteration k >= 1: p = r + beta p_old, q = A^ p, partials p.q and q.q.
// First iteration (First): p = r, and only the partials: p and q are not stored, because the
// solve either stops here (k_fin_a's first-iteration test; y = alpha r is formed by its
// consumer) or k_cg_b1_z recomputes q.
// Iteration k >= 1 also adds iteration k-1's term to the solution: y = alpha p_old (yfirst,
// k == 1) or y += alpha p_old, with alpha still that iteration's.
template <typename T, bool First>
__global__ void __launch_bounds__(kZTX * kZTY)
k_cg_a_z(Dims g, const T *__restrict__ r, const T *__restrict__ pold, T *__restrict__ pnew,
         T *__restrict__ q, T *__restrict__ y, int32_t yfirst, const T *__restrict__ hx,
         const T *__restrict__ hy, const T *__restrict__ hz, Scal *sc, double *partials) {
    __shared__ PlaneBuf<T> sp;
    if (sc->done) return;
    const T beta = (T)sc->beta;
    const T alpha = (T)sc->alpha;
    auto P = [&](uint32_t J) -> T { return First ? r[J] : (T)(r[J] + beta * pold[J]); };
    auto own = [&](uint32_t J) -> T { return First ? (T)0 : pold[J]; };
    double acc[2] = {0.0, 0.0};
    const ZCol c = zcolumn(g);
    T hz_m = (T)0;
    if (c.active && (c.z0 > 0 || g.per[2])) hz_m = hz[zm_index(g, c.col, c.z0)];
    zmarch_own<T, T>(g, c, sp, P, own, [&](uint32_t I, uint32_t z, T pi, T pz_m, T pz_p,
                                           const InPlane<T> &nb, T po) {
        const T hzc = ldnt(&hz[I]);
        const T qi = ahat(g, c, I, z, pi, pz_m, pz_p, hz_m, hzc, hx, hy, nb, P);
        if (!First) {
            stnt(&pnew[I], pi);
            stnt(&q[I], qi);
            stnt(&y[I], yfirst ? (T)(alpha * po) : (T)(ldnt(&y[I]) + alpha * po));
        }
        acc[0] += (double)pi * (double)qi;
        acc[1] += (double)qi * (double)qi;
        hz_m = hzc;
    });
    block_partials<2>(acc, partials);
}

// Second half of the first CG iteration when the solve goes on: q = A^ r recomputed,
// rw = r - alpha q (into a separate vector: r is the stencil operand), partial rw.rw.  The
// right-hand side r is left unchanged; it is p_old of iteration 1 (whose k_cg_a forms
// y = alpha r).
template <typename T>
__global__ void __launch_bounds__(kZTX * kZTY)
k_cg_b1_z(Dims g, const T *__restrict__ r, const T *__restrict__ hx, const T *__restrict__ hy,
          const T *__restrict__ hz, T *__restrict__ rw, Scal *sc, double *partials) {
    __shared__ PlaneBuf<T> sp;
    if (sc->done) return;
    const T alpha = (T)sc->alpha;
    auto P = [&](uint32_t J) -> T { return r[J]; };
    double acc[1] = {0.0};
    const ZCol c = zcolumn(g);
    T hz_m = (T)0;
    if (c.active && (c.z0 > 0 || g.per[2])) hz_m = hz[zm_index(g, c.col, c.z0)];
    zmarch<T>(g, c, sp, P, [&](uint32_t I, uint32_t z, T pi, T pz_m, T pz_p,
                               const InPlane<T> &nb) {
        const T hzc = ldnt(&hz[I]);
        const T qi = ahat(g, c, I, z, pi, pz_m, pz_p, hz_m, hzc, hx, hy, nb, P);
        const T ri = pi - alpha * qi;
        stnt(&rw[I], ri);
        acc[0] += (double)ri * (double)ri;
        hz_m = hzc;
    });
    block_partials<1>(acc, partials);
}

// Mixed-precision refinement: x' = x + s y (y FP32) formed on the fly, r^ = -s (S - H x')
// stored as FP32 for the next inner solve, ||r^||^2, NaN and next-step TINY tests on x'.
__global__ void __launch_bounds__(kZTX * kZTY)
k_ir_res_z(Dims g, IrParams ip, const double *__restrict__ x, const float *__restrict__ y,
           const float *__restrict__ r0, const Scal *inner, const double *__restrict__ sv,
           const double *__restrict__ S, const double *__restrict__ d,
           const double *__restrict__ cx, const double *__restrict__ cy,
           const double *__restrict__ cz, const double *__restrict__ tp,
           double *__restrict__ xout, float *__restrict__ rhat, Scal *sc, double *partials) {
    __shared__ PlaneBuf<double> sp;
    // the inner solution (CgY; alpha r0 formed in FP32 as k_cg_b1_z's terms are)
    const CgY<float> ysol = cg_y(inner, y, r0);
    auto X = [&](uint32_t J) -> double {
        return ysol.any ? x[J] + sv[J] * (double)ysol.at(J) : x[J];
    };
    double acc[1] = {0.0};
    int32_t fl = 0;
    const ZCol c = zcolumn(g);
    double cz_m = 0.0;
    if (c.active && (c.z0 > 0 || g.per[2])) cz_m = cz[zm_index(g, c.col, c.z0)];
    zmarch<double>(g, c, sp, X, [&](uint32_t I, uint32_t z, double xi, double xz_m, double xz_p,
                                    const InPlane<double> &nb) {
        const double czc = ldnt(&cz[I]);
        const double hxv =
            hx_sum(g, c, I, z, ldnt(&d[I]), cx, cy, czc, xi, xz_m, xz_p, cz_m, nb, X);
        const double rh = -sv[I] * (ldnt(&S[I]) - hxv);
        const float rf = (float)rh;
        stnt(&rhat[I], rf);
        stnt(&xout[I], xi);
        acc[0] += (double)rf * (double)rf;
        if (isnan(xi)) fl |= FLAG_NAN_X;
        if (ip.from_state && (xi + (xi - ldnt(&tp[I]))) - xi < kTiny) fl |= FLAG_TINY_NEXT;
        cz_m = czc;
    });
    block_flags(fl, &sc->flags);
    block_partials<1>(acc, partials);
}

// operand inputs of the fused step: T^n, T^{n-1} (or the refined x), the diagonal (s is
// formed from it, reading s as well would cost 8 bytes a cell) and the inner solution (y, or
// r0 when scaled by alpha: CgY).  The cell's own record of plane z also gives it T^n, T^{n-1}
// and the diagonal.  s = inv_sqrt(-d) may differ from k_scale's 1/sqrt(-d) in the last bit:
// it scales the update x0 + s y and the residual the pass forms, both of which the refinement
// measures in FP64 (the two-step pass forms it the same way, so both give the same bits).
struct StepRaw {
    double tp, b, d;   // d: the diagonal
    float y;
};

// the other per-cell inputs of the fused step: diag, z+ coupling, Qdens + B, rho c_p,
// tau' rho c_p, the x+ / y+ couplings and those of the x- / y- neighbours (stored at the
// neighbour; read at the cell itself where that neighbour is absent), T^{n-1} (refined pass)
struct StepCell {
    double cz, qb, rc, cv, cx, cy, cxm, cym, tpp;
};

// Fused step of the mixed-precision path, one pass over the grid after an inner solve:
//   x' = base + s y               base = x0 of this step from the state (2 T^n - T^{n-1},
//                                 +TINY, mod_evolve.f90:140-141) or a refined xb
//   r^ = -s (S_n - H x')          this step's FP64 refinement residual -> rres (FP32)
//   x1 = x' + (x' - T^n)          the next step's initial guess (no TINY yet)
//   r^0 = -s (S_{n+1} - H x1)     the next step's residual, S_{n+1} from (x', T^n) -> rnext
//                                 (S uses qb = Qdens + B, constant over the run), or with
//                                 the TINY shift x1 + TINY: r^0 + TINY s (H 1)
// The TINY test of the next step (any x1 - x' < TINY) is an OR over the grid, known only after
// the pass: rnext is formed for the guessed outcome (StepParams::tiny_guess), and the flag the
// pass raises tells the host whether the guess held (else the next step starts afresh).
// rres is read only by a further refinement sweep, also known only after the pass: it is
// stored when the host expects one (StepParams::store_res), else the pass is run again.  x'
// and x1 are staged together in LDS (two stencil operands, one halo ring).
//
// The z-march is software-pipelined: while plane z is computed, the loads of plane z+1's cell
// inputs and halo ring and of plane z+2's own operand inputs are in flight, so a trip waits on
// loads issued one trip earlier.  That takes ~128 VGPRs: 4 waves per SIMD (2 blocks of 8 waves
// per CU), which the staged pair of operands allowed anyway; at 512^3 it runs ~8% faster than
// the same pass with each trip's loads issued after its barrier.  Transient is isteady == 0,
// Cattaneo icattaneo == 1, Wrap some axis periodic (compile-time, so the trip is straight-line
// code the scheduler can order freely).
template <bool FromState, bool Transient, bool Cattaneo, bool Wrap, bool UniformRc>
__global__ void __launch_bounds__(kZTX * kZTY) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_ir_step_z(Dims g, StepParams sp_, const double *__restrict__ tp,
            const double *__restrict__ tpp, const double *__restrict__ xb,
            const float *__restrict__ y, const float *__restrict__ r0, const Scal *inner,
            const double *__restrict__ rhoc, const double *__restrict__ cvec,
            const double *__restrict__ qb,
            const double *__restrict__ d, const double *__restrict__ cx,
            const double *__restrict__ cy, const double *__restrict__ cz,
            double *__restrict__ xout, float *__restrict__ rres, float *__restrict__ rnext,
            Scal *sc, double *partials) {
    __shared__ PlaneBuf<D2> sp;
    const bool tiny = sc->tiny != 0;
    // rho c_p and tau' rho c_p of every cell when both are uniform (UniformRc): one scalar
    // read instead of a load per cell, and two fewer doubles carried per cell record
    const double rc_u = UniformRc ? rhoc[0] : 0.0;
    const double cv_u = UniformRc && Cattaneo ? cvec[0] : 0.0;
    // the inner solution (CgY), formed in FP32 as k_cg_b1_z forms its terms; 0 when the solve
    // needed no iteration
    const CgY<float> ysol = cg_y(inner, y, r0);
    const float ysc = ysol.any ? ysol.scale : 0.0f;
    const ZCol c = zcolumn(g);
    const Ring rg = ring_cell(g, c);
    const bool per_z = Wrap && g.per[2];
    const bool xm_in = c.x > 0, xp_in = c.x + 1 < g.nx;
    const bool ym_in = c.y > 0, yp_in = c.y + 1 < g.ny;
    auto load = [&](uint32_t J) -> StepRaw {
        return StepRaw{tp[J], FromState ? tpp[J] : xb[J], d[J], ysol.v[J]};
    };
    auto form = [&](const StepRaw &r) -> D2 {
        double base;
        if (FromState) {
            base = r.tp + (r.tp - r.b);
            if (tiny) base = base + kTiny;
        } else {
            base = r.b;
        }
        const double xn = __builtin_fma(inv_sqrt(-r.d), (double)(ysc * r.y), base);
        return D2(xn, xn + (xn - r.tp));
    };
    // form for the thread's own cells, also handing back s (same operations)
    auto form_s = [&](const StepRaw &r, double &si) -> D2 {
        double base;
        if (FromState) {
            base = r.tp + (r.tp - r.b);
            if (tiny) base = base + kTiny;
        } else {
            base = r.b;
        }
        si = inv_sqrt(-r.d);
        const double xn = __builtin_fma(si, (double)(ysc * r.y), base);
        return D2(xn, xn + (xn - r.tp));
    };
    auto X = [&](uint32_t J) -> D2 { return form(load(J)); };
    const NbOff o = nb_offsets_in(g, c);
    // I must be a cell of this thread's column (active threads) or 0
    auto cload = [&](uint32_t I) -> StepCell {
        StepCell v;
        v.cz = ldnt(&cz[I]);
        v.qb = ldnt(&qb[I]);
        v.rc = Transient && !UniformRc ? ldnt(&rhoc[I & sp_.rc_mask]) : 0.0;
        v.cv = Transient && Cattaneo && !UniformRc ? ldnt(&cvec[I & sp_.rc_mask]) : 0.0;
        v.cx = cx[I];
        v.cy = cy[I];
        v.cxm = cx[I + o.xm];
        v.cym = cy[I + o.ym];
        v.tpp = FromState ? 0.0 : ldnt(&tpp[I]);
        return v;
    };
    // H v of the cell in the reference's serial order (as hx_sum), K picks x' or x1; each term
    // is one fused multiply-add (the mixed step measures its residual in FP64: the rounding of
    // the products is not the oracle's, unlike the FP64 path's sums)
    auto hsum = [&](uint32_t I, uint32_t z, const StepCell &cv, double di, double cz_m,
                    const D2 &xi, const D2 &xz_m, const D2 &xz_p, const D2 *nbc,
                    auto pick) -> double {
        double h = 0.0;
        h = __builtin_fma(cz_m, z > 0 ? pick(xz_m) : 0.0, h);
        h = __builtin_fma(cv.cym, ym_in ? pick(nbc[-(int)(kZTX + 2)]) : 0.0, h);
        h = __builtin_fma(cv.cxm, xm_in ? pick(nbc[-1]) : 0.0, h);
        h = __builtin_fma(di, pick(xi), h);
        h = __builtin_fma(cv.cx, xp_in ? pick(nbc[1]) : 0.0, h);
        h = __builtin_fma(cv.cy, yp_in ? pick(nbc[(int)(kZTX + 2)]) : 0.0, h);
        h = __builtin_fma(cv.cz, z + 1 < g.nz ? pick(xz_p) : 0.0, h);
        if (Wrap) {
            const bool hzm = z > 0 || g.per[2], hzp = z + 1 < g.nz || g.per[2];
            if (c.hxm && c.wxm) h = __builtin_fma(cx[I + c.oxm], pick(X(I + c.oxm)), h);
            if (c.hxp && c.wxp) h = __builtin_fma(cv.cx, pick(X(I + c.oxp)), h);
            if (c.hym && c.wym) h = __builtin_fma(cy[I + c.oym], pick(X(I + c.oym)), h);
            if (c.hyp && c.wyp) h = __builtin_fma(cv.cy, pick(X(I + c.oyp)), h);
            if (hzm && z == 0) h = __builtin_fma(cz_m, pick(xz_m), h);
            if (hzp && z + 1 == g.nz) h = __builtin_fma(cv.cz, pick(xz_p), h);
        }
        return h;
    };
    auto pick_a = [](const D2 &v) { return v.a; };
    auto pick_b = [](const D2 &v) { return v.b; };
    const bool guess = sp_.tiny_guess != 0;
    // This thread's sums over its cells (|r^|^2, |b^|^2, |r^0|^2, r^0 . r'), each kept in its
    // own LDS slot and added to in cell order, rather than in registers: the pipelined trip
    // uses nearly all 128 VGPRs that 4 waves per SIMD allow
    __shared__ double part[4][kZTX * kZTY];
#pragma unroll
    for (int k = 0; k < 4; ++k) part[k][threadIdx.x] = 0.0;
    auto add_part = [&](int k, double v) {
        __hip_atomic_fetch_add(&part[k][threadIdx.x], v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    // flags of this step (NAN_X, TINY_NEXT) and of the next step's S (NAN_SCAT, NAN_S): distinct
    // bits in one word
    int32_t fl = 0;
    double cz_m = 0.0;
    D2 xz_m = 0.0, xi = 0.0;
    double s_i = 0.0;     // s of the own cell of plane z, formed with xi
    StepRaw ri{}, r1{};   // own operand inputs of planes z and z+1
    StepCell cur{};       // cell inputs of plane z
    if (c.active) {
        if (c.z0 > 0 || per_z) {
            const uint32_t J = zm_index(g, c.col, c.z0);
            xz_m = X(J);
            cz_m = cz[J];
        }
        ri = load(c.col + g.nxny * c.z0);
        xi = form_s(ri, s_i);
        if (c.z0 + 1 < g.nz || per_z) r1 = load(zp_index(g, c.col, c.z0));
        cur = cload(c.col + g.nxny * c.z0);
    }
    if (c.zvalid) {
        sp.v[0][c.ly + 1][c.lx + 1] = xi;
        if (rg.col != kNoRing) (&sp.v[0][0][0])[rg.slot] = X(rg.col + g.nxny * c.z0);
    }
    int buf = 0;
    // One trip: plane z (cell I of the column) with its own operand inputs of planes z (ri)
    // and z+1 (r1) and its cell inputs (cur), loaded one and two trips earlier; loads r2 (z+2),
    // cn (z+1) and, on a ring thread, the ring cell's inputs of plane z+1 (formed into the LDS
    // plane at the end of the trip).
    auto trip = [&](uint32_t z, uint32_t I, const StepRaw &ri, const StepRaw &r1,
                    const StepCell &cur, StepRaw &r2, StepCell &cn) {
        const bool more = z + 1 < c.z1;
        const bool has_p = c.active && (z + 1 < g.nz || per_z);
        const bool has_p2 = c.active && more && (z + 2 < g.nz || per_z);
        const bool has_r = more && rg.col != kNoRing;
        // The next trips' loads are issued on every trip by every thread, without branches, so
        // that the compiler's wait counts know how many are in flight and this trip waits only
        // for the previous trip's.  Where there is nothing to load they read an address this
        // thread read recently (cache hits): its own cell, or cell 0 (inactive threads).
        const uint32_t own = c.active ? I : 0u;
        r2 = load(has_p2 ? zp_index(g, c.col, z + 1) : own);
        // (the address is opaque to the compiler, which would otherwise load under a branch on
        // the ring test and zero the record elsewhere: that write waits for the record's
        // previous load)
        uint32_t jr = rg.col != kNoRing ? rg.col + g.nxny * (more ? z + 1 : z) : 0u;
        asm volatile("" : "+v"(jr));
        const StepRaw rr = load(jr);
        cn = cload(c.active && more ? I + g.nxny : own);
        __syncthreads();   // plane z is staged in sp.v[buf]
        D2 xz_p = 0.0;
        double s_p = 0.0;
        if (c.active) {
            xz_p = has_p ? form_s(r1, s_p) : D2(0.0);
            const D2 *nbc = &sp.v[buf][c.ly + 1][c.lx + 1];
            const double tpi = ri.tp;
            const double tppi = FromState ? ri.b : cur.tpp;
            // S of this step (mod_evolve.f90:112-119, S_catS mod_cattaneo.f90:35-36) and of
            // the next, with Qdens + B summed once per run (qb)
            double Sn, Sn1;
            if (Transient) {
                const double rci = UniformRc ? rc_u : cur.rc;
                const double cvi = UniformRc ? cv_u : cur.cv;
                Sn = __builtin_fma(-(sp_.inverse_time * tpi), rci, -cur.qb);
                Sn1 = __builtin_fma(-(sp_.inverse_time * xi.a), rci, -cur.qb);
                if (Cattaneo) {
                    // (a product and a sum, as the next step's scat below and in the
                    // two-step pass)
                    Sn = Sn + (tppi - 2.0 * tpi) * cvi;
                    const double scat = (tpi - 2.0 * xi.a) * cvi;
                    if (isnan(scat)) fl |= FLAG_NAN_SCAT;
                    Sn1 = Sn1 + scat;
                }
            } else {
                Sn = -cur.qb;
                Sn1 = Sn;
            }
            if (isnan(Sn1)) fl |= FLAG_NAN_S;
            const double h1 = hsum(I, z, cur, ri.d, cz_m, xi, xz_m, xz_p, nbc, pick_a);
            const double h2 = hsum(I, z, cur, ri.d, cz_m, xi, xz_m, xz_p, nbc, pick_b);
            const double s = s_i;
            const float rf = (float)(-s * (Sn - h1));
            double ra = -s * (Sn1 - h2);
            if (guess) {
                // H 1: the row sum over the same terms (absent couplings are stored as 0)
                double h1s = ri.d;
                h1s = h1s + cz_m;
                // (an absent neighbour's term is selected away: h1s starts at the nonzero
                // diagonal, so adding 0 or selecting it out gives the same sum)
                h1s = h1s + (c.hym ? (Wrap ? cy[I + c.oym] : cur.cym) : 0.0);
                h1s = h1s + (c.hxm ? (Wrap ? cx[I + c.oxm] : cur.cxm) : 0.0);
                h1s = h1s + cur.cx;
                h1s = h1s + cur.cy;
                h1s = h1s + cur.cz;
                ra = __builtin_fma(kTiny * s, h1s, ra);
            }
            const float fn = (float)ra;
            stnt(&xout[I], xi.a);
            if (sp_.store_res) stnt(&rres[I], rf);
            stnt(&rnext[I], fn);
            const double bh = s * Sn1;
            add_part(0, (double)rf * (double)rf);
            add_part(1, bh * bh);
            add_part(2, (double)fn * (double)fn);
            // r^0 . r' when y = alpha r^0 (the sweep reused alpha): gives r^0 . A^ r^0
            add_part(3, (double)ri.y * (double)rf);
            if (isnan(xi.a)) fl |= FLAG_NAN_X;
            if (xi.b - xi.a < kTiny) fl |= FLAG_TINY_NEXT;   // mod_evolve.f90:141, next step
            cz_m = cur.cz;
        }
        if (more) {
            // sp.v[buf ^ 1] was last read before this trip's barrier
            sp.v[buf ^ 1][c.ly + 1][c.lx + 1] = xz_p;
            if (has_r) (&sp.v[buf ^ 1][0][0])[rg.slot] = form(rr);
        }
        buf ^= 1;
        xz_m = xi;
        xi = xz_p;
        s_i = s_p;
    };
    // Six trips per iteration: the own operand records of planes z, z+1, z+2 rotate through
    // three register sets and the cell records of planes z, z+1 through two, so every loaded
    // record keeps its registers from its load to its last use.  A record that changed
    // registers between trips would be copied at the loop's back edge, and the copy waits for
    // the loads just issued.  (The compiler does not unroll a loop with a barrier by itself.)
    StepRaw r2{};
    StepCell cn{};
    const uint32_t P = g.nxny;
    uint32_t I = c.col + P * c.z0;
    for (uint32_t z = c.z0; z < c.z1; z += 6, I += 6 * P) {
        trip(z, I, ri, r1, cur, r2, cn);
        if (z + 1 >= c.z1) break;
        trip(z + 1, I + P, r1, r2, cn, ri, cur);
        if (z + 2 >= c.z1) break;
        trip(z + 2, I + 2 * P, r2, ri, cur, r1, cn);
        if (z + 3 >= c.z1) break;
        trip(z + 3, I + 3 * P, ri, r1, cn, r2, cur);
        if (z + 4 >= c.z1) break;
        trip(z + 4, I + 4 * P, r1, r2, cur, ri, cn);
        if (z + 5 >= c.z1) break;
        trip(z + 5, I + 5 * P, r2, ri, cn, r1, cur);
    }
    block_flags(fl & (FLAG_NAN_X | FLAG_TINY_NEXT), &sc->flags);
    block_flags(fl & (FLAG_NAN_SCAT | FLAG_NAN_S), &sc->flags_next);
    double acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = part[k][threadIdx.x];
    block_partials<4>(acc, partials);
}

// operand inputs of the fused FP64 step: T^n, T^{n-1}, the diagonal (s is formed from it as
// k_scale forms it) and the solution (y, or r0 scaled by alpha: CgY)
struct Step64Raw {
    double tp, tpp, d, y;
};

// what forming a cell's own operand yields on the way: x^{n+1} and s = 1/sqrt(-d)
struct Own64 {
    double x, s;
};

// Fused step of the FP64 path, one pass over the grid after the solve: k_finalize of this step
// followed by k_rhs_res_z of the next, with the same operations in the same order, so results
// are bit-identical to the two kernels run apart:
//   x = x0 + s y                  x0 = 2 T^n - T^{n-1} (+TINY) (mod_evolve.f90:140-141) -> xout
//   x1 = x + (x - T^n) (+TINY)    the next step's initial guess, staged in LDS
//   r^0 = -s (S_{n+1} - H x1)     S_{n+1} from (x, T^n) (mod_evolve.f90:112-119) -> rnext
// The next step's TINY case (any x1 - x < TINY) is an OR over the grid, known only after the
// pass: x1 is formed for the guessed outcome `guess`, and the flag the pass raises tells the
// host whether the guess held (else the next step starts from k_rhs_res_z).  NaN flags of
// S_{n+1} go to flags_next, which the next step's start makes current.
template <int Mode, bool Wrap>
__global__ void __launch_bounds__(kZTX * kZTY)
k_step64_z(Dims g, RhsParams rp, int32_t guess, const double *__restrict__ tp,
           const double *__restrict__ tpp, const double *__restrict__ y,
           const double *__restrict__ r0, const double *__restrict__ rhoc,
           const double *__restrict__ cvec, const double *__restrict__ qd,
           const double *__restrict__ bvec, const double *__restrict__ d,
           const double *__restrict__ cx, const double *__restrict__ cy,
           const double *__restrict__ cz, double *__restrict__ xout,
           double *__restrict__ rnext, Scal *sc, double *partials) {
    __shared__ PlaneBuf<double> sp;
    const bool tiny = sc->tiny != 0;
    const CgY<double> ysol = cg_y(sc, y, r0);
    auto load = [&](uint32_t J) -> Step64Raw { return Step64Raw{tp[J], tpp[J], d[J], ysol.v[J]}; };
    // x^{n+1} of a cell as k_finalize forms it
    auto xnew = [&](const Step64Raw &v) -> double {
        double x0 = v.tp + (v.tp - v.tpp);
        if (tiny) x0 = x0 + kTiny;
        return ysol.any ? x0 + (1.0 / sqrt(-v.d)) * (ysol.scale * v.y) : x0;
    };
    // the next step's x0 as k_rhs_res_z forms it from the state (x, T^n)
    auto form = [&](const Step64Raw &v) -> double {
        const double x = xnew(v);
        double x1 = x + (x - v.tp);
        if (guess) x1 = x1 + kTiny;
        return x1;
    };
    // the same for the thread's own cells, handing the cell its x and s (same operations)
    auto fown = [&](const Step64Raw &v, Own64 &e) -> double {
        e.s = 1.0 / sqrt(-v.d);
        double x0 = v.tp + (v.tp - v.tpp);
        if (tiny) x0 = x0 + kTiny;
        e.x = ysol.any ? x0 + e.s * (ysol.scale * v.y) : x0;
        double x1 = e.x + (e.x - v.tp);
        if (guess) x1 = x1 + kTiny;
        return x1;
    };
    auto X = [&](uint32_t J) -> double { return form(load(J)); };
    double acc[2] = {0.0, 0.0};
    int32_t fl = 0, fl_next = 0;
    const ZCol c = zcolumn(g);
    const NbOff o = nb_offsets_in(g, c);
    // the diagonal comes with the cell's own operand record (ri.d), not with these
    auto cload = [&](uint32_t I) -> RhsCell {
        RhsCell v;
        v.rc = Mode != kSteady ? ldnt(&rhoc[I & rp.rc_mask]) : 0.0;
        v.cv = Mode == kCattaneo ? ldnt(&cvec[I & rp.rc_mask]) : 0.0;
        v.qd = ldnt(&qd[I]);
        v.bv = ldnt(&bvec[I]);
        v.h = HCell{0.0, cx[I], cy[I], ldnt(&cz[I]), cx[I + o.xm], cy[I + o.ym]};
        return v;
    };
    double cz_m = 0.0;
    if (c.active && (c.z0 > 0 || (Wrap && g.per[2]))) cz_m = cz[zm_index(g, c.col, c.z0)];
    zmarch_pipe_own<double, Step64Raw, RhsCell, Own64, Wrap>(
        g, c, sp, load, form, fown, cload,
        [&](uint32_t I, uint32_t z, double xi, double xz_m, double xz_p,
            const InPlane<double> &nb, const RhsCell &cv, const Step64Raw &ri, const Own64 &e) {
            const double x = e.x;
            stnt(&xout[I], x);
            if (isnan(x)) fl |= FLAG_NAN_X;                                  // mod_evolve.f90:153
            if ((x + (x - ri.tp)) - x < kTiny) fl |= FLAG_TINY_NEXT;          // :141, next step
            double S;
            if (Mode != kSteady) {
                S = -(rp.inverse_time * x * cv.rc) - cv.qd - cv.bv;
                if (Mode == kCattaneo) {
                    const double scat = (ri.tp - 2.0 * x) * cv.cv;
                    if (isnan(scat)) fl_next |= FLAG_NAN_SCAT;
                    S = S + scat;
                }
            } else {
                S = -cv.qd - cv.bv;
            }
            if (isnan(S)) fl_next |= FLAG_NAN_S;
            HCell h = cv.h;
            h.d = ri.d;
            const double hxv = hx_sum_v<Wrap>(g, c, I, z, h, cx, cy, xi, xz_m, xz_p, cz_m, nb, X);
            const double r = S - hxv;
            const double s = e.s;
            const double rh = -s * r;
            stnt(&rnext[I], rh);
            const double bh = s * S;
            acc[0] += bh * bh;
            acc[1] += rh * rh;
            cz_m = cv.h.cz;
        });
    block_flags(fl, &sc->flags);
    block_flags(fl_next, &sc->flags_next);
    block_partials<2>(acc, partials);
}

// ---- CG kernels with W x-neighbouring cells per thread (nx a multiple of W) ----------------
//
// FP32 arrays moved one float per lane per instruction run well below the FP64 kernels'
// bandwidth.  With W cells per thread every lane moves W values per access: 16 bytes for FP64
// pairs (W = 2) and FP32 quads (W = 4), 8 for FP32 pairs.  A block of kZTX x kZTY threads owns
// W kZTX x kZTY columns (tiles numbered as above, ztxW per row, chunks of zchunkW planes).  The
// LDS plane holds W-tuples: the x neighbours of a tuple are the last cell of the tuple on its
// left and the first of the one on its right.  Sums per cell are the scalar kernels' (same
// terms, same order).

template <typename T, int W>
using vecw = T __attribute__((ext_vector_type(W)));

// the W-cell tiling of a Dims
template <int W>
struct TileW {
    static __device__ __forceinline__ uint32_t ztx(const Dims &g) {
        return W == 2 ? g.ztx2 : g.ztx4;
    }
    static __device__ __forceinline__ uint32_t zchunk(const Dims &g) {
        return W == 2 ? g.zchunk2 : g.zchunk4;
    }
    static __device__ __forceinline__ uint32_t ztiles(const Dims &g) {
        return W == 2 ? g.ztiles2 : g.ztiles4;
    }
};

struct ZColW {
    uint32_t col, x, y, z0, z1;   // x: the first cell of the tuple
    uint32_t bx, by, lx, ly;
    bool active, zvalid;
    bool hym, hyp, wym, wyp;
    int32_t oym, oyp;
    bool hxm_a, wxm_a;            // the first cell's x- neighbour exists / is the wrap (x == 0)
    bool hxp_b, wxp_b;            // the last cell's x+ neighbour exists / is the wrap
    int32_t oxm_a, oxp_b;         // offsets from the first cell's index of those neighbours
};

template <int W>
__device__ __forceinline__ ZColW zcolumnw(const Dims &g) {
    ZColW c;
    const uint32_t lt = logical_tile(TileW<W>::ztiles(g));
    const uint32_t tx = TileW<W>::ztx(g);
    const uint32_t per_plane = tx * g.zty;
    const uint32_t zc = lt / per_plane;
    const uint32_t rem = lt - zc * per_plane;
    c.by = rem / tx;
    c.bx = rem - c.by * tx;
    c.lx = threadIdx.x % kZTX;
    c.ly = threadIdx.x / kZTX;
    c.x = c.bx * (W * kZTX) + W * c.lx;
    c.y = c.by * kZTY + c.ly;
    c.z0 = zc * TileW<W>::zchunk(g);
    c.z1 = c.z0 + TileW<W>::zchunk(g) < g.nz ? c.z0 + TileW<W>::zchunk(g) : g.nz;
    c.zvalid = lt < TileW<W>::ztiles(g) && c.z0 < g.nz;
    c.active = c.zvalid && c.x < g.nx && c.y < g.ny;   // nx % W == 0: x + W - 1 < nx as well
    c.col = c.x + g.nx * c.y;
    c.hym = c.y > 0 || g.per[1];
    c.hyp = c.y + 1 < g.ny || g.per[1];
    c.wym = c.y == 0;
    c.wyp = c.y + 1 == g.ny;
    c.oym = c.wym ? (int32_t)((g.ny - 1) * g.nx) : -(int32_t)g.nx;
    c.oyp = c.wyp ? -(int32_t)((g.ny - 1) * g.nx) : (int32_t)g.nx;
    c.wxm_a = c.x == 0;
    c.hxm_a = c.x > 0 || g.per[0];
    c.oxm_a = c.wxm_a ? (int32_t)(g.nx - 1) : -1;
    c.wxp_b = c.x + W == g.nx;
    c.hxp_b = c.x + W < g.nx || g.per[0];
    c.oxp_b = c.wxp_b ? -(int32_t)(g.nx - W) : W;
    return c;
}

template <int W>
__device__ __forceinline__ Ring ring_cellw(const Dims &g, const ZColW &c) {
    Ring r{0, kNoRing};
    const int t = (int)threadIdx.x;
    const int TX = (int)kZTX, TY = (int)kZTY;
    if (t >= kZRing || !c.zvalid) return r;
    int lx, ly;
    if (t < TX) {
        lx = t + 1; ly = 0;
    } else if (t < 2 * TX) {
        lx = t - TX + 1; ly = TY + 1;
    } else if (t < 2 * TX + TY) {
        lx = 0; ly = t - 2 * TX + 1;
    } else {
        lx = TX + 1; ly = t - 2 * TX - TY + 1;
    }
    const int64_t px = (int64_t)c.bx * TX + lx - 1, gy = (int64_t)c.by * TY + ly - 1;
    r.slot = ly * (TX + 2) + lx;
    if (px >= 0 && W * px < (int64_t)g.nx && gy >= 0 && gy < (int64_t)g.ny)
        r.col = (uint32_t)(W * px) + g.nx * (uint32_t)gy;
    return r;
}

// zmarch for W-tuples: X(J) gives the operand tuple at the first cell J; own(J) a tuple of the
// cells' own (loads shared with X's, carried in registers as in zmarch_own).
template <typename T, int W, class XF, class OF, class CF>
__device__ __forceinline__ void zmarchw(const Dims &g, const ZColW &c,
                                        PlaneBuf<vecw<T, W>> &sp, XF X, OF own, CF cell) {
    using TW = vecw<T, W>;
    const Ring rg = ring_cellw<W>(g, c);
    const bool per_z = g.per[2];
    TW xz_m = (TW)(T)0, xi = (TW)(T)0, oi = (TW)(T)0;
    if (c.active) {
        if (c.z0 > 0 || per_z) xz_m = X(zm_index(g, c.col, c.z0));
        const uint32_t J0 = c.col + g.nxny * c.z0;
        xi = X(J0);
        oi = own(J0);
    }
    if (c.zvalid) {
        sp.v[0][c.ly + 1][c.lx + 1] = xi;
        if (rg.col != kNoRing) (&sp.v[0][0][0])[rg.slot] = X(rg.col + g.nxny * c.z0);
    }
    int buf = 0;
    uint32_t I = c.col + g.nxny * c.z0;
    for (uint32_t z = c.z0; z < c.z1; ++z, I += g.nxny) {
        const bool more = z + 1 < c.z1;
        TW xz_p = (TW)(T)0, rnext = (TW)(T)0, op = (TW)(T)0;
        if (c.active && (z + 1 < g.nz || per_z)) {
            const uint32_t J = zp_index(g, c.col, z);
            xz_p = X(J);
            op = own(J);
        }
        if (more && rg.col != kNoRing) rnext = X(rg.col + g.nxny * (z + 1));
        __syncthreads();   // plane z is staged in sp.v[buf]
        if (c.active) {
            const InPlane<TW> nb{&sp.v[buf][c.ly + 1][c.lx + 1]};
            cell(I, z, xi, xz_m, xz_p, nb, oi);
        }
        if (more) {
            sp.v[buf ^ 1][c.ly + 1][c.lx + 1] = xz_p;
            if (rg.col != kNoRing) (&sp.v[buf ^ 1][0][0])[rg.slot] = rnext;
        }
        buf ^= 1;
        xz_m = xi;
        xi = xz_p;
        oi = op;
    }
}

template <typename T, int W>
__device__ __forceinline__ vecw<T, W> ldw(const T *p) { return *(const vecw<T, W> *)p; }
template <typename T, int W>
__device__ __forceinline__ vecw<T, W> ldntw(const T *p) {
    return __builtin_nontemporal_load((const vecw<T, W> *)p);
}
template <typename T, int W>
__device__ __forceinline__ void stntw(T *p, vecw<T, W> v) {
    __builtin_nontemporal_store(v, (vecw<T, W> *)p);
}

// q = A^ p for the tuple p, the scalar ahat's terms and order per cell; P1 reads the operand of
// one (wrap) cell from memory, PW the tuple at a first cell.
template <typename T, int W, class P1F, class PWF>
__device__ __forceinline__ vecw<T, W> ahatw(const Dims &g, const ZColW &c, uint32_t I, uint32_t z,
                                            vecw<T, W> pi, vecw<T, W> pz_m, vecw<T, W> pz_p,
                                            vecw<T, W> hz_m, vecw<T, W> hzc,
                                            const T *__restrict__ hx, const T *__restrict__ hy,
                                            const InPlane<vecw<T, W>> &nb, P1F P1, PWF PW) {
    using TW = vecw<T, W>;
    const bool per_z = g.per[2];
    const bool hzm = z > 0 || per_z, hzp = z + 1 < g.nz || per_z;
    const TW hxw = ldw<T, W>(&hx[I]);     // coupling of each cell to its x+ neighbour
    // absent neighbours contribute coefficient * 0 (coefficients read at I); wrap neighbours
    // are read from memory
    TW vym = (TW)(T)0, vyp = (TW)(T)0;
    T vxm = (T)0, vxp = (T)0;
    if (c.hym) vym = c.wym ? PW(I + c.oym) : nb.ym();
    if (c.hyp) vyp = c.wyp ? PW(I + c.oyp) : nb.yp();
    if (c.hxm_a) vxm = c.wxm_a ? P1(I + c.oxm_a) : nb.xm()[W - 1];
    if (c.hxp_b) vxp = c.wxp_b ? P1(I + c.oxp_b) : nb.xp()[0];
    const TW zm = hzm ? pz_m : (TW)(T)0, zp = hzp ? pz_p : (TW)(T)0;
    const TW cym = ldw<T, W>(&hy[c.hym ? I + c.oym : I]), cyp = ldw<T, W>(&hy[I]);
    const T cxm = hx[c.hxm_a ? I + c.oxm_a : I];
    TW q;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        T s = (T)0;
        s = s + hz_m[k] * zm[k];
        s = s + cym[k] * vym[k];
        // x-: the first cell's left neighbour, else the previous cell (coupling stored there);
        // x+: the next cell, or the right neighbour of the last one
        s = s + (k == 0 ? cxm : hxw[k > 0 ? k - 1 : 0]) * (k == 0 ? vxm : pi[k > 0 ? k - 1 : 0]);
        s = s + hxw[k] * (k == W - 1 ? vxp : pi[k < W - 1 ? k + 1 : 0]);
        s = s + cyp[k] * vyp[k];
        s = s + hzc[k] * zp[k];
        q[k] = pi[k] - s;
    }
    return q;
}

template <typename T, int W, bool First>
__global__ void __launch_bounds__(kZTX * kZTY)
k_cg_a_zw(Dims g, const T *__restrict__ r, const T *__restrict__ pold, T *__restrict__ pnew,
          T *__restrict__ q, T *__restrict__ y, int32_t yfirst, const T *__restrict__ hx,
          const T *__restrict__ hy, const T *__restrict__ hz, Scal *sc, double *partials) {
    using TW = vecw<T, W>;
    __shared__ PlaneBuf<TW> sp;
    if (sc->done) return;
    const T beta = (T)sc->beta;
    const T alpha = (T)sc->alpha;
    auto P1 = [&](uint32_t J) -> T { return First ? r[J] : (T)(r[J] + beta * pold[J]); };
    auto PW = [&](uint32_t J) -> TW {
        if (First) return ldw<T, W>(&r[J]);
        const TW a = ldw<T, W>(&r[J]), b = ldw<T, W>(&pold[J]);
        TW v;
#pragma unroll
        for (int k = 0; k < W; ++k) v[k] = a[k] + beta * b[k];
        return v;
    };
    auto own = [&](uint32_t J) -> TW { return First ? (TW)(T)0 : ldw<T, W>(&pold[J]); };
    double acc[2] = {0.0, 0.0};
    const ZColW c = zcolumnw<W>(g);
    TW hz_m = (TW)(T)0;
    if (c.active && (c.z0 > 0 || g.per[2])) hz_m = ldw<T, W>(&hz[zm_index(g, c.col, c.z0)]);
    zmarchw<T, W>(g, c, sp, PW, own, [&](uint32_t I, uint32_t z, TW pi, TW pz_m, TW pz_p,
                                         const InPlane<TW> &nb, TW po) {
        const TW hzc = ldntw<T, W>(&hz[I]);
        const TW qi = ahatw<T, W>(g, c, I, z, pi, pz_m, pz_p, hz_m, hzc, hx, hy, nb, P1, PW);
        if (!First) {
            stntw<T, W>(&pnew[I], pi);
            stntw<T, W>(&q[I], qi);
            // iteration k-1's term of the solution (alpha p_old), as k_cg_a_z
            TW yi = yfirst ? (TW)(T)0 : ldntw<T, W>(&y[I]);
#pragma unroll
            for (int k = 0; k < W; ++k) yi[k] = yfirst ? alpha * po[k] : yi[k] + alpha * po[k];
            stntw<T, W>(&y[I], yi);
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            acc[0] += (double)pi[k] * (double)qi[k];
            acc[1] += (double)qi[k] * (double)qi[k];
        }
        hz_m = hzc;
    });
    block_partials<2>(acc, partials);
}

template <typename T, int W>
__global__ void __launch_bounds__(kZTX * kZTY)
k_cg_b1_zw(Dims g, const T *__restrict__ r, const T *__restrict__ hx, const T *__restrict__ hy,
           const T *__restrict__ hz, T *__restrict__ rw, Scal *sc, double *partials) {
    using TW = vecw<T, W>;
    __shared__ PlaneBuf<TW> sp;
    if (sc->done) return;
    const T alpha = (T)sc->alpha;
    auto P1 = [&](uint32_t J) -> T { return r[J]; };
    auto PW = [&](uint32_t J) -> TW { return ldw<T, W>(&r[J]); };
    auto own = [](uint32_t) -> TW { return (TW)(T)0; };
    double acc[1] = {0.0};
    const ZColW c = zcolumnw<W>(g);
    TW hz_m = (TW)(T)0;
    if (c.active && (c.z0 > 0 || g.per[2])) hz_m = ldw<T, W>(&hz[zm_index(g, c.col, c.z0)]);
    zmarchw<T, W>(g, c, sp, PW, own, [&](uint32_t I, uint32_t z, TW pi, TW pz_m, TW pz_p,
                                         const InPlane<TW> &nb, TW) {
        const TW hzc = ldntw<T, W>(&hz[I]);
        const TW qi = ahatw<T, W>(g, c, I, z, pi, pz_m, pz_p, hz_m, hzc, hx, hy, nb, P1, PW);
        TW ri;
#pragma unroll
        for (int k = 0; k < W; ++k) ri[k] = pi[k] - alpha * qi[k];
        stntw<T, W>(&rw[I], ri);
#pragma unroll
        for (int k = 0; k < W; ++k) acc[0] += (double)ri[k] * (double)ri[k];
        hz_m = hzc;
    });
    block_partials<1>(acc, partials);
}

// ---- launchers -----------------------------------------------------------------------------

// ---- two reusing steps in one z-march (k_ir_step2_z) ---------------------------------------
//
// In the carried regime where the inner step length is reused (or a step's own inner solve ends
// on its first iteration, and the next step reuses that alpha), a mixed step is one stencil
// pass of per-cell arithmetic: x' = x0 + s (alpha r^0), its FP64 residual, and the next step's
// r^0.
// Two such steps are two passes over the same operator, Qdens + B and state.  k_ir_step2_z does
// both in one z-march, so those arrays are read once for two steps (80 instead of 144
// algorithmic bytes a cell):
//   step n     (as k_ir_step_z): x'_n, rho_n = -s (S_n - H x'_n), r^0_{n+1} from x1_n
//   step n+1   x'_{n+1} = x1_n (+TINY) + s (alpha r^0_{n+1}), rho_{n+1}, r^0_{n+2} from x1_{n+1}
// with the same alpha for both steps and the TINY outcome of both next steps guessed as the
// last one seen.  Every per-cell operation, its order and its rounding (r^0 is rounded to FP32
// as the single-step pass stores it) are those of two k_ir_step_z passes, so a pair whose tests
// pass gives the bits of two single steps that reuse the same alpha.  The residual sums run over
// the same cells grouped by this kernel's tiles: they agree with the single passes' to rounding,
// as does the step length formed from them for the next pair.
//
// Step n+1 at a cell needs r^0_{n+1} of its neighbours, which needs x1_n two cells away: the
// (x'_n, x1_n) plane ("a") is formed on the tile and a two-cell ring, r^0_{n+1} and
// (x'_{n+1}, x1_{n+1}) ("b") on the tile and a one-cell ring, and step n+1 runs one plane
// behind step n.  Trip q of the march forms a(q); then (barrier) step n on plane q-1 for the
// tile and the one-cell ring (its residual test on the tile only) and step n+1's z+ term and
// residual on plane q-2; then (barrier) step n+1's other terms on plane q-1.  Three a-planes
// rotate in LDS beside one b-plane (the own cell's b of the plane below stays in registers) and
// two planes of the b-region's (s, d, T^n).  Each thread takes its own tile cell as its first
// position in every stage, so what one stage hands the next for that cell stays in its
// registers; the a-ring's cells are second positions of the first waves, the b-ring's of the
// waves after them.  Non-periodic grids only.
//
// Outputs: x'_n -> x1out, x'_{n+1} -> x2out, r^0_{n+2} (FP32) -> rnext; partials 0-3 of step n,
// 4-7 of step n+1, each ||rho||^2, ||b^_next||^2, ||r^0_next||^2, r^0 . rho; flags: TINY_NEXT
// of step n in Scal::flags and a NaN anywhere in step n (x'_n, scat, S_{n+1}: tested on
// S_{n+1}, which any of them makes NaN) as NAN_S in flags_next; those of step n+1 (TINY_NEXT2,
// NAN_S2) in flags2.  The pair needs only whether each step met a NaN: it then falls back to
// single steps, whose flags name the quantity.
constexpr int kAW = (int)kP2X + 4, kAH = (int)kP2Y + 4, kAN = kAW * kAH;
constexpr int kBW = (int)kP2X + 2, kBH = (int)kP2Y + 2, kBN = kBW * kBH;
constexpr int kTT = (int)(kP2X * kP2Y);
static_assert(kAN - kTT <= kTT && kBN - kTT <= kTT, "at most two positions a thread");

// A plane of (x', x1) pairs is held as two arrays, x' then x1, a multiple of 64 doubles apart:
// a lane reading one half of its neighbours' pairs reads consecutive doubles (no bank
// conflicts), and one ds_read2st64_b64 reads both halves of a pair.
constexpr int kANp = (kAN + 63) / 64 * 64, kBNp = (kBN + 63) / 64 * 64;
struct Step2Lds {
    double a[3][2][kANp];
    double b[2][kBNp];
    // the Jacobi scale, diagonal and T^n of plane q's b-region cells, formed by the operand
    // stage on trip q and read by step n on trip q + 1 (two planes rotate)
    double sdt[2][3][kBNp];
};
static_assert(sizeof(Step2Lds) <= 160 * 1024, "the two-step pass's LDS");

// the k-th position (k >= 0) of a region of W x H around the tile (ring width R): k < T is tile
// cell (k % TX, k / TX) at (k % TX + R, k / TX + R); then the ring, row by row for its full
// rows (the R top and R bottom rows), then the R left and R right columns of the tile rows
template <int R>
__device__ __forceinline__ void region_pos(int k, int &px, int &py) {
    constexpr int TX = (int)kP2X, TY = (int)kP2Y, W = TX + 2 * R;
    if (k < kTT) {
        px = k % TX + R;
        py = k / TX + R;
        return;
    }
    k -= kTT;
    if (k < 2 * R * W) {
        const int row = k / W;
        px = k - row * W;
        py = row < R ? row : TY + row;   // rows 0..R-1 and TY+R..TY+2R-1
        return;
    }
    k -= 2 * R * W;
    const int col = k / TY;              // columns 0..R-1, then TX+R..TX+2R-1
    py = k - col * TY + R;
    px = col < R ? col : TX + col;
}

// Element at a 32-bit byte offset from p (the address form with the base in scalar registers
// and one 32-bit offset a lane, shared by the arrays of one element size); the caller
// guarantees the offset fits (grids of fewer than 2^29 cells).
template <typename T>
__device__ __forceinline__ T ld_at(const T *p, uint32_t byte_off) {
    return *(const T *)((const char *)p + byte_off);
}
template <typename T>
__device__ __forceinline__ void st_at(T *p, uint32_t byte_off, T v) {
    __builtin_nontemporal_store(v, (T *)((char *)p + byte_off));
}

// ---- cross-lane sums (gfx950 DPP and permlane swaps) ----------------------------------------

// A 64-bit value of another lane under DPP control Ctrl (both halves moved alike; all lanes of
// the wave active).
template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, Ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), Ctrl, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32));
}
__device__ __forceinline__ double f64_of(uint32_t lo, uint32_t hi) {
    return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
}
// a's upper 32 lanes <-> b's lower 32 lanes
__device__ __forceinline__ void swap32_f64(double &a, double &b) {
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
    const auto hi =
        __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
    a = f64_of(lo[0], hi[0]);
    b = f64_of(lo[1], hi[1]);
}
// a's odd 16-lane rows <-> b's even rows
__device__ __forceinline__ void swap16_f64(double &a, double &b) {
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
    const auto hi =
        __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
    a = f64_of(lo[0], hi[0]);
    b = f64_of(lo[1], hi[1]);
}
// Eight values a lane summed over the wave's 64 lanes: lane L gets the sum of component L >> 3.
// Each exchange halves the components a lane carries (32-lane swap, 16-lane swap, a rotation by
// 8 within rows), then three more sum the 8 lanes of a group; the order is fixed, so the sums
// are deterministic.  All lanes must be active.
__device__ __forceinline__ double wave_sum8(const double (&v)[8]) {
    double w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double a = v[k], b = v[k + 4];
        swap32_f64(a, b);
        w[k] = a + b;   // lanes 0-31: component k; 32-63: component k + 4
    }
    double u[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double a = w[k], b = w[k + 2];
        swap16_f64(a, b);
        u[k] = a + b;   // rows 0-3 (16 lanes each): components (0, 1), (2, 3), (4, 5), (6, 7)
    }
    const bool hi8 = (threadIdx.x & 8) != 0;
    const double keep = hi8 ? u[1] : u[0], send = hi8 ? u[0] : u[1];
    double x = keep + dpp_f64<0x128>(send);   // row_ror:8 -> component L >> 3
    x = x + dpp_f64<0x141>(x);                // row_half_mirror
    x = x + dpp_f64<0x4E>(x);                 // quad_perm [2, 3, 0, 1]
    x = x + dpp_f64<0xB1>(x);                 // quad_perm [1, 0, 3, 2]
    return x;
}

template <bool Cattaneo, bool UniformRc>
__global__ void __launch_bounds__(kP2X * kP2Y) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_ir_step2_z(Dims g, StepParams sp_, const double *__restrict__ tp,
             const double *__restrict__ tpp, const float *__restrict__ r0, const Scal *inner,
             const double *__restrict__ rhoc, const double *__restrict__ cvec,
             const double *__restrict__ qb, const double *__restrict__ d,
             const double *__restrict__ cx, const double *__restrict__ cy,
             const double *__restrict__ cz, double *__restrict__ x1out,
             double *__restrict__ x2out, float *__restrict__ rnext, Scal *sc,
             double *partials) {
    __shared__ Step2Lds L;
    const bool tiny = sc->tiny != 0;           // step n's TINY outcome (known)
    const bool guess = sp_.tiny_guess != 0;    // steps n+1 and n+2, guessed
    const float ysc = (float)inner->alpha;     // the reused step length (y = alpha r^0)
    const double rc_u = UniformRc ? rhoc[0] : 0.0;
    const double cv_u = UniformRc && Cattaneo ? cvec[0] : 0.0;
    const double it = sp_.inverse_time;
    // this block's tile (XCD-aware as the other z-marching kernels) and z-chunk
    const uint32_t lt = logical_tile(g.p2tiles);
    const uint32_t per_plane = g.p2x * g.p2y;
    const uint32_t zc = lt / per_plane, rem = lt - zc * per_plane;
    const uint32_t tby = rem / g.p2x, tbx = rem - tby * g.p2x;
    const uint32_t zchunk = g.p2zc;
    const uint32_t z0 = zc * zchunk, z1 = z0 + zchunk < g.nz ? z0 + zchunk : g.nz;
    const bool zvalid = lt < g.p2tiles && z0 < g.nz;
    const int t = (int)threadIdx.x;
    const uint32_t lx = (uint32_t)t % kP2X, ly = (uint32_t)t / kP2X;
    const uint32_t ox = tbx * kP2X + lx, oy = tby * kP2Y + ly;
    const bool active = zvalid && ox < g.nx && oy < g.ny;
    const uint32_t ocol = ox + g.nx * oy;
    const int64_t gx0 = (int64_t)tbx * kP2X, gy0 = (int64_t)tby * kP2Y;
    // a second a-position (on the ring) for the first threads; a second b-position for the
    // threads of the waves after theirs, so that no wave runs two ring stages
    constexpr int kBOff = (kAN - kTT + 63) / 64 * 64;
    static_assert(kBOff + (kBN - kTT) <= kTT, "the b-ring's threads");
    const bool two_a = t < kAN - kTT;
    const bool two_b = t >= kBOff && t < kBOff + (kBN - kTT);
    const uint32_t P = g.nxny;
    // grid column of a region position (R: ring width), or kNoCol outside the grid
    constexpr uint32_t kNoCol = 0xffffffffu;
    auto col_of = [&](int px, int py, int R) -> uint32_t {
        const int64_t x = gx0 + px - R, y = gy0 + py - R;
        if (!zvalid || x < 0 || y < 0 || x >= (int64_t)g.nx || y >= (int64_t)g.ny) return kNoCol;
        return (uint32_t)x + g.nx * (uint32_t)y;
    };
    // the thread's a-positions (k = 0: its tile cell) and b-positions: region coordinates
    // (x | y << 16) and grid columns
    uint32_t apos[2], bpos[2], acol[2], bcol[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        int px, py;
        region_pos<2>(k == 0 || two_a ? t + k * kTT : t, px, py);
        apos[k] = (uint32_t)px | (uint32_t)py << 16;
        acol[k] = k == 0 || two_a ? col_of(px, py, 2) : kNoCol;
        region_pos<1>(k == 1 && two_b ? kTT + (t - kBOff) : t, px, py);
        bpos[k] = (uint32_t)px | (uint32_t)py << 16;
        bcol[k] = k == 0 || two_b ? col_of(px, py, 1) : kNoCol;
    }
    auto pos_x = [](uint32_t v) { return (int)(v & 0xffffu); };
    auto pos_y = [](uint32_t v) { return (int)(v >> 16); };
    auto mod3 = [](int v) { return ((v % 3) + 3) % 3; };
    // the residual sums: each trip's eight per-cell terms are summed over the wave at once
    // (wave_sum8), and lane L accumulates component L >> 3
    double acc = 0.0;
    int32_t fl = 0, fl2 = 0;

    // Loads.  Every thread issues the same loads on every trip, without branches, so that the
    // compiler's wait counts know how many are in flight; a position with nothing to load
    // reads cell 0.  A trip's loads (a(q), then step n's) are issued on the trip before, ahead
    // of its last barrier.
    struct ALd {
        double tp, tpp, d;
        float r;
    };
    struct BLd {
        double d, cx, cy, cz, cxm, cym, czm, qb, tp, rc, cv;
    };
    auto a_load = [&](int k, int q) {
        const bool ok = acol[k] != kNoCol && q >= 0 && q < (int)g.nz;
        uint32_t J = ok ? acol[k] + P * (uint32_t)q : 0u;
        asm volatile("" : "+v"(J));
        const uint32_t o8 = J * 8u;
        ALd v;
        v.tp = ld_at(tp, o8);
        v.tpp = ld_at(tpp, o8);
        v.d = ld_at(d, o8);
        v.r = ld_at(r0, J * 4u);
        return v;
    };
    // step n's operator row, Qdens + B and T^n at b-position k of plane p (the tile cell takes
    // d and T^n from its a-loads of the trip before)
    auto b_load = [&](int k, int p) {
        const bool ok = bcol[k] != kNoCol && p >= 0 && p < (int)g.nz;
        const uint32_t x = (uint32_t)(gx0 + pos_x(bpos[k]) - 1);
        const uint32_t y = (uint32_t)(gy0 + pos_y(bpos[k]) - 1);
        const uint32_t J = ok ? bcol[k] + P * (uint32_t)p : 0u;
        uint32_t o8 = J * 8u;
        uint32_t oxm = ok && x > 0 ? o8 - 8u : o8;
        uint32_t oym = ok && y > 0 ? o8 - 8u * g.nx : o8;
        uint32_t ozm = ok && p > 0 ? o8 - 8u * P : o8;
        asm volatile("" : "+v"(o8), "+v"(oxm), "+v"(oym), "+v"(ozm));
        BLd v;
        v.cx = ld_at(cx, o8);
        v.cy = ld_at(cy, o8);
        v.cz = ld_at(cz, o8);
        v.cxm = ld_at(cx, oxm);
        v.cym = ld_at(cy, oym);
        v.czm = ld_at(cz, ozm);
        v.qb = ld_at(qb, o8);
        if (!UniformRc) {
            v.rc = rhoc[J & sp_.rc_mask];
            v.cv = Cattaneo ? cvec[J & sp_.rc_mask] : 0.0;
        }
        return v;
    };

    // own-cell values handed from stage to stage
    double own_tpp = 0.0;   // a(q) -> step n on plane q
    float own_r = 0.0f;
    D2 b_prev = 0.0;   // own b of the plane before the one step n ran on last
    // step n+1 at the own cell of plane p, formed on the trip step n ran on p, with the z+ term
    // (b of plane p + 1) added on the next trip: H x'_{n+1} and H x1_{n+1} without it, the
    // tiny-guess row sum without it, the z+ coupling, s, S_{n+1}, S_{n+2}, r^0_{n+1}
    double w_ha = 0.0, w_hb = 0.0, w_h1 = 0.0, w_cz = 0.0, w_s = 0.0, w_S1 = 0.0, w_S2 = 0.0;
    float w_fn = 0.0f;
    const int zq0 = (int)z0 - 2, zq1 = (int)z1 + 1;
    ALd al0 = a_load(0, zq0), al1 = a_load(1, zq0);
    BLd bl0 = b_load(0, zq0 - 1);
    for (int q = zq0; q <= zq1; ++q) {
        const int p = q - 1;
        const BLd bl1 = b_load(1, p);   // the ring position's
        // ---- a(q): x'_n and x1_n on the tile and a two-cell ring
        double *A = L.a[mod3(q)][0];
        const bool zin = q >= 0 && q < (int)g.nz;
        double(*SD)[kBNp] = L.sdt[q & 1];
        auto stage_a = [&](const ALd &v, int k) {
            D2 o = 0.0;
            const int px = pos_x(apos[k]), py = pos_y(apos[k]);
            if (zin && acol[k] != kNoCol) {
                const double sa = inv_sqrt(-v.d);
                double base = v.tp + (v.tp - v.tpp);   // mod_evolve.f90:140
                if (tiny) base = base + kTiny;         // :141
                const double xn = __builtin_fma(sa, (double)(ysc * v.r), base);
                o = D2(xn, xn + (xn - v.tp));
                // inside the b-region (all but the outer ring): for step n on the next trip
                if (px >= 1 && px <= kBW && py >= 1 && py <= kBH) {
                    const int bi = (py - 1) * kBW + (px - 1);
                    SD[0][bi] = sa;
                    SD[1][bi] = v.d;
                    SD[2][bi] = v.tp;
                }
            }
            A[py * kAW + px] = o.a;
            A[kANp + py * kAW + px] = o.b;
        };
        stage_a(al0, 0);
        if (two_a) stage_a(al1, 1);
        __syncthreads();   // a(q) staged
        // ---- step n on plane p = q - 1 (ring positions; its test on the tile only)
        const bool step_n = p >= (int)z0 - 1 && p <= (int)z1 && p >= 0 && p < (int)g.nz;
        // the tile's own planes: its residual test, sums and flags (the planes z0 - 1 and z1
        // are the neighbouring chunks')
        const bool own_plane = p >= (int)z0 && p < (int)z1;
        const double *Am = L.a[mod3(p - 1)][0], *A0 = L.a[mod3(p)][0], *Ap = L.a[mod3(p + 1)][0];
        // the pair at position i of an a-plane
        auto ga = [](const double *P, int i) { return D2(P[i], P[kANp + i]); };
        // this trip's terms of the residual sums: step n on plane p (0-3), step n+1 on p - 1
        double tv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        // the own cell's operator row, S_{n+1}, x'_n, r^0_{n+1} and b: step n+1 on plane p
        D2 o_b = 0.0;
        double o_d = 0.0, o_cxi = 0.0, o_cyi = 0.0, o_czi = 0.0, o_cxm = 0.0, o_cym = 0.0;
        double o_czm = 0.0, o_s = 0.0, o_S1 = 0.0, o_S2 = 0.0, o_h1 = 0.0, o_xn = 0.0;
        float o_fn = 0.0f;
        // position k of this thread (k = 0: its tile cell)
        auto stage_b = [&](auto kc, const BLd &ld) {
            constexpr int k = decltype(kc)::value;
            const int px = pos_x(bpos[k]), py = pos_y(bpos[k]);
            D2 v = 0.0;
            if (step_n && bcol[k] != kNoCol) {
                const uint32_t x = (uint32_t)(gx0 + px - 1), y = (uint32_t)(gy0 + py - 1);
                const bool xm_in = x > 0, ym_in = y > 0, zm_in = p > 0;
                const int ai = (py + 1) * kAW + (px + 1);   // this position in an a-plane
                const D2 xi = ga(A0, ai);
                const int sb = py * kBW + px;   // this position in the b-region
                const double(*SDp)[kBNp] = L.sdt[p & 1];
                const double di = SDp[1][sb];
                const double cxi = ld.cx, cyi = ld.cy, czi = ld.cz;
                const double cxm = ld.cxm, cym = ld.cym;
                const double czm = zm_in ? ld.czm : 0.0;
                const double qbi = ld.qb;
                const double tpi = SDp[2][sb];
                const double rci = UniformRc ? rc_u : ld.rc;
                const double cvi = Cattaneo ? (UniformRc ? cv_u : ld.cv) : 0.0;
                // H x (pick: x' or x1) in the reference's serial order; an a-plane holds 0 at
                // the positions outside the grid, so an absent neighbour adds coefficient x 0
                auto hsum = [&](auto pick) {
                    double h = 0.0;
                    h = __builtin_fma(czm, pick(ga(Am, ai)), h);
                    h = __builtin_fma(cym, pick(ga(A0, ai - kAW)), h);
                    h = __builtin_fma(cxm, pick(ga(A0, ai - 1)), h);
                    h = __builtin_fma(di, pick(xi), h);
                    h = __builtin_fma(cxi, pick(ga(A0, ai + 1)), h);
                    h = __builtin_fma(cyi, pick(ga(A0, ai + kAW)), h);
                    h = __builtin_fma(czi, pick(ga(Ap, ai)), h);
                    return h;
                };
                const double s = SDp[0][sb];
                // S_{n+1} from (x'_n, T^n) (mod_evolve.f90:112-119, S_catS)
                double S1 = __builtin_fma(-(it * xi.a), rci, -qbi);
                if (Cattaneo) {
                    const double scat = (tpi - 2.0 * xi.a) * cvi;
                    S1 = S1 + scat;
                }
                double ra = -s * (S1 - hsum([](const D2 &w) { return w.b; }));
                if (guess) {
                    double h1s = di;
                    h1s = h1s + czm;
                    h1s = h1s + (ym_in ? cym : 0.0);
                    h1s = h1s + (xm_in ? cxm : 0.0);
                    h1s = h1s + cxi;
                    h1s = h1s + cyi;
                    h1s = h1s + czi;
                    ra = __builtin_fma(kTiny * s, h1s, ra);
                }
                const float fn = (float)ra;
                // step n+1's x'_{n+1} = x1_n (+TINY) + s (alpha r^0_{n+1}), x1_{n+1}
                double base = xi.a + (xi.a - tpi);
                if (guess) base = base + kTiny;
                const double xn1 = __builtin_fma(s, (double)(ysc * fn), base);
                v = D2(xn1, xn1 + (xn1 - xi.a));
                if (k == 0 && own_plane) {
                    // the tile cell: step n's residual test and sums
                    double Sn = __builtin_fma(-(it * tpi), rci, -qbi);
                    if (Cattaneo) Sn = Sn + (own_tpp - 2.0 * tpi) * cvi;
                    // a NaN in x'_n or in its scat makes S_{n+1} NaN: one test flags step n
                    // (the pair then falls back to single steps, which tell them apart)
                    if (isnan(S1)) fl |= FLAG_NAN_S;
                    const float rf = (float)(-s * (Sn - hsum([](const D2 &w) { return w.a; })));
                    const double bh = s * S1;
                    tv[0] = (double)rf * (double)rf;
                    tv[1] = bh * bh;
                    tv[2] = (double)fn * (double)fn;
                    tv[3] = (double)own_r * (double)rf;
                    if (xi.b - xi.a < kTiny) fl |= FLAG_TINY_NEXT;
                    // step n+1's S_{n+2} from (x'_{n+1}, x'_n), and its tiny-guess row sum but
                    // for the z+ coupling
                    double S2 = __builtin_fma(-(it * v.a), rci, -qbi);
                    if (Cattaneo) {
                        const double scat = (xi.a - 2.0 * v.a) * cvi;
                        S2 = S2 + scat;
                    }
                    if (isnan(S2)) fl2 |= FLAG_NAN_S2;   // x'_{n+1}, its scat or S_{n+2}
                    o_S2 = S2;
                    if (guess) {
                        double h1s = di;
                        h1s = h1s + czm;
                        h1s = h1s + (ym_in ? cym : 0.0);
                        h1s = h1s + (xm_in ? cxm : 0.0);
                        h1s = h1s + cxi;
                        h1s = h1s + cyi;
                        o_h1 = h1s;
                    }
                }
                if (k == 0) {
                    o_b = v;
                    o_d = di;
                    o_cxi = cxi;
                    o_cyi = cyi;
                    o_czi = czi;
                    o_cxm = cxm;
                    o_cym = cym;
                    o_czm = czm;
                    o_s = s;
                    o_S1 = S1;
                    o_xn = xi.a;
                    o_fn = fn;
                }
            }
            L.b[0][py * kBW + px] = v.a;
            L.b[1][py * kBW + px] = v.b;
        };
        if (two_b) stage_b(std::integral_constant<int, 1>{}, bl1);
        stage_b(std::integral_constant<int, 0>{}, bl0);
        // ---- step n+1 on plane p2 = q - 2 (tile cells): add the z+ term, b(q - 1) of the own
        // cell, to what the previous trip formed, and finish
        const int p2 = q - 2;
        if (active && p2 >= (int)z0 && p2 < (int)z1) {
            // (o_b is 0 when plane p2 + 1 is outside the grid)
            const double ha = __builtin_fma(w_cz, o_b.a, w_ha);
            const double hb = __builtin_fma(w_cz, o_b.b, w_hb);
            const float rf = (float)(-w_s * (w_S1 - ha));
            double ra = -w_s * (w_S2 - hb);
            if (guess) ra = __builtin_fma(kTiny * w_s, w_h1 + w_cz, ra);
            const float fn = (float)ra;
            st_at(rnext, (ocol + P * (uint32_t)p2) * 4u, fn);
            const double bh = w_s * w_S2;
            tv[4] = (double)rf * (double)rf;
            tv[5] = bh * bh;
            tv[6] = (double)fn * (double)fn;
            tv[7] = (double)w_fn * (double)rf;
        }
        acc += wave_sum8(tv);
        own_tpp = al0.tpp;
        own_r = al0.r;
        // the next trip's loads: a(q + 1), then step n on plane q
        al0 = a_load(0, q + 1);
        al1 = a_load(1, q + 1);
        bl0 = b_load(0, q);
        __syncthreads();   // b(p) staged; a(q - 2) no longer read
        // ---- step n+1 on plane p (tile cells) but for its z+ term: b(p) in LDS, the own b of
        // plane p - 1 in a register
        if (active && own_plane) {
            const int bi = ((int)ly + 1) * kBW + ((int)lx + 1);
            const D2 xi = o_b;
            auto gb = [&](int i) { return D2(L.b[0][i], L.b[1][i]); };
            // (the b-plane holds 0 at the positions outside the grid, and b_prev is 0 on plane 0)
            auto hsum6 = [&](auto pick) {
                double h = 0.0;
                h = __builtin_fma(o_czm, pick(b_prev), h);
                h = __builtin_fma(o_cym, pick(gb(bi - kBW)), h);
                h = __builtin_fma(o_cxm, pick(gb(bi - 1)), h);
                h = __builtin_fma(o_d, pick(xi), h);
                h = __builtin_fma(o_cxi, pick(gb(bi + 1)), h);
                h = __builtin_fma(o_cyi, pick(gb(bi + kBW)), h);
                return h;
            };
            w_ha = hsum6([](const D2 &w) { return w.a; });
            w_hb = hsum6([](const D2 &w) { return w.b; });
            w_h1 = o_h1;
            w_cz = o_czi;
            w_s = o_s;
            w_S1 = o_S1;
            w_S2 = o_S2;
            w_fn = o_fn;
            const uint32_t o8 = (ocol + P * (uint32_t)p) * 8u;
            st_at(x1out, o8, o_xn);
            st_at(x2out, o8, xi.a);
            if (xi.b - xi.a < kTiny) fl2 |= FLAG_TINY_NEXT2;
        }
        b_prev = o_b;
    }
    block_flags(fl & (FLAG_NAN_X | FLAG_TINY_NEXT), &sc->flags);
    block_flags(fl & (FLAG_NAN_SCAT | FLAG_NAN_S), &sc->flags_next);
    block_flags(fl2, &sc->flags2);
    // partials: components 0-3 (step n), then 4-7 (step n+1); the waves' sums in wave order,
    // staged in the a-planes (no longer read)
    static_assert(kTT / 64 * 8 <= 2 * kANp, "wave sums fit an a-plane");
    double *red = L.a[0][0];
    const int lane = t & 63, wv = t >> 6;
    __syncthreads();
    if ((lane & 7) == 0) red[(lane >> 3) * (kTT / 64) + wv] = acc;
    __syncthreads();
    if (t < 8) {
        double sum = 0.0;
        for (int w = 0; w < kTT / 64; ++w) sum += red[t * (kTT / 64) + w];
        partials[(size_t)t * gridDim.x + blockIdx.x] = sum;
    }
}

void launch_ir_step2(const Dims &g0, const StepParams &sp, const double *tp, const double *tpp,
                     const float *r0, const Scal *inner, const double *rhoc, const double *cvec,
                     const double *qb, const double *d, const double *cx, const double *cy,
                     const double *cz, double *x1out, double *x2out, float *rnext, Scal *sc,
                     double *partials, hipStream_t st) {
    const Dims g = zdims(g0, kZPipe);
    const bool uni = sp.rc_mask == 0;
    if (sp.icattaneo == 1) {
        if (uni)
            hipLaunchKernelGGL((k_ir_step2_z<true, true>), dim3(g.p2blk), dim3(kP2X * kP2Y), 0, st,
                               g, sp, tp, tpp, r0, inner, rhoc, cvec, qb, d, cx, cy, cz, x1out,
                               x2out, rnext, sc, partials);
        else
            hipLaunchKernelGGL((k_ir_step2_z<true, false>), dim3(g.p2blk), dim3(kP2X * kP2Y), 0,
                               st, g, sp, tp, tpp, r0, inner, rhoc, cvec, qb, d, cx, cy, cz,
                               x1out, x2out, rnext, sc, partials);
    } else {
        if (uni)
            hipLaunchKernelGGL((k_ir_step2_z<false, true>), dim3(g.p2blk), dim3(kP2X * kP2Y), 0,
                               st, g, sp, tp, tpp, r0, inner, rhoc, cvec, qb, d, cx, cy, cz,
                               x1out, x2out, rnext, sc, partials);
        else
            hipLaunchKernelGGL((k_ir_step2_z<false, false>), dim3(g.p2blk), dim3(kP2X * kP2Y), 0,
                               st, g, sp, tp, tpp, r0, inner, rhoc, cvec, qb, d, cx, cy, cz,
                               x1out, x2out, rnext, sc, partials);
    }
}

#define HF_ZLAUNCH(kern, g, st, ...) \
    hipLaunchKernelGGL(kern, dim3((g).zblk), dim3(kZTX * kZTY), 0, st, __VA_ARGS__)

// k_rhs_res_z for the run-time S mode and periodicity, with the pipelined kernels' z-chunking
// (launch_fin_start reads its partials)
// ---- adjoint step ----------------------------------------------------------------------------

// the adjoint step's stencil operand of a cell: lambda^{n+1}, T^{n+1}, kappa
struct AdjV {
    double l, t, k;
    __device__ AdjV() = default;
    __device__ constexpr AdjV(double v) : l(v), t(v), k(v) {}
    __device__ constexpr AdjV(double a, double b, double c) : l(a), t(b), k(c) {}
};

// k_adj_step (hf_adjoint.hip) as a z-march: each cell's face terms over x-, x+, y-, y+, z-, z+
// (the same terms in the same order), with the neighbours' (lambda, T, kappa) from the staged
// LDS plane and registers instead of global memory.  No periodic axis (those grids run the
// flat kernel).
__global__ void __launch_bounds__(kZTX * kZTY)
k_adj_step_z(Dims g, AdjParams q, const double *__restrict__ T, const double *__restrict__ lam,
             const double *__restrict__ kappa, const double *__restrict__ lx,
             const double *__restrict__ ly, const double *__restrict__ lz,
             const double *__restrict__ rhoc, const double *__restrict__ cvec,
             const double *__restrict__ vol, double *__restrict__ g_lo,
             double *__restrict__ g_new, double *__restrict__ gk, double *__restrict__ gp) {
    __shared__ PlaneBuf<AdjV> sp;
    // cell lengths are uniform per axis (checked when the adjoint starts): read once
    const double Ls[3] = {lx[0], ly[0], lz[0]};
    const ZCol c = zcolumn(g);
    auto X = [&](uint32_t J) -> AdjV { return AdjV(lam[J], T[J], kappa[J]); };
    zmarch<AdjV>(g, c, sp, X, [&](uint32_t I, uint32_t z, AdjV xi, AdjV xz_m, AdjV xz_p,
                                  const InPlane<AdjV> &nb) {
        const double li = xi.l, ti = xi.t, ki = xi.k;
        const uint32_t pos[3] = {c.x, c.y, z};
        const uint32_t len[3] = {g.nx, g.ny, g.nz};
        const bool has[6] = {c.x > 0, c.x + 1 < g.nx, c.y > 0, c.y + 1 < g.ny, z > 0,
                             z + 1 < g.nz};
        // which boundary faces still contribute to B_I after CG assignments
        // (mod_boundary.f90:72-150)
        int survive = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int ax = f >> 1;
            const bool at_face = pos[ax] == ((f & 1) ? len[ax] - 1 : 0u);
            if (!q.bface[f] || !at_face) continue;
            if (q.cg[f]) survive = 0;
            else survive |= (1 << f);
        }
        double acc = 0.0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int ax = f >> 1;
            if (has[f]) {
                const AdjV v = f == 0 ? nb.xm() : f == 1 ? nb.xp() : f == 2 ? nb.ym()
                             : f == 3 ? nb.yp() : f == 4 ? xz_m : xz_p;
                acc += dG_int(ki, v.k, Ls[ax], q.br) * (li - v.l) * (ti - v.t);
            } else if (q.bface[f]) {
                const double tb = (survive >> f) & 1 ? q.tb[f] : 0.0;
                acc += dG_bnd(q.kb[f], ki, Ls[ax], q.br) * li * (ti - tb);
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
    });
}

void launch_adj_step_z(const Dims &g0, const AdjParams &q, const double *T, const double *lam,
                       const double *kappa, const double *lx, const double *ly, const double *lz,
                       const double *rhoc, const double *cvec, const double *vol, double *g_lo,
                       double *g_new, double *gk, double *gp, hipStream_t st) {
    const Dims g = zdims(g0, kZDefault);
    HF_ZLAUNCH(k_adj_step_z, g, st, g, q, T, lam, kappa, lx, ly, lz, rhoc, cvec, vol, g_lo, g_new,
               gk, gp);
}

template <typename OutT, bool FS, bool SS, bool SX>
static void launch_rhs(const Dims &g0, const RhsParams &rp, const double *tp, const double *tpp,
                       const double *xin, const double *bin, const double *rhoc,
                       const double *cvec, const double *qd, const double *bvec, const double *d,
                       const double *cx, const double *cy, const double *cz, OutT *rhat,
                       double *sout, double *xout, Scal *sc, double *partials, hipStream_t st) {
    const Dims g = zdims(g0, kZPipe);
#define HF_RHS(M, W)                                                                        \
    HF_ZLAUNCH((k_rhs_res_z<OutT, FS, SS, SX, M, W>), g, st, g, rp, tp, tpp, xin, bin, rhoc, \
               cvec, qd, bvec, d, cx, cy, cz, rhat, sout, xout, sc, partials)
    const int mode = FS ? s_mode(rp.isteady, rp.icattaneo) : kSteady;
    const bool wrap = g.per[0] || g.per[1] || g.per[2];
    if (mode == kCattaneo) {
        if (wrap) HF_RHS(kCattaneo, true); else HF_RHS(kCattaneo, false);
    } else if (mode == kFourier) {
        if (wrap) HF_RHS(kFourier, true); else HF_RHS(kFourier, false);
    } else {
        if (wrap) HF_RHS(kSteady, true); else HF_RHS(kSteady, false);
    }
#undef HF_RHS
}

void launch_rhs_res_state64(const Dims &g, const RhsParams &rp, const double *tp,
                            const double *tpp, const double *rhoc, const double *cvec,
                            const double *qd, const double *bvec, const double *d,
                            const double *cx, const double *cy, const double *cz, double *rhat,
                            double *sout, Scal *sc, double *partials, hipStream_t st) {
    if (sout)
        launch_rhs<double, true, true, false>(g, rp, tp, tpp, nullptr, nullptr, rhoc, cvec, qd,
                                              bvec, d, cx, cy, cz, rhat, sout, nullptr, sc,
                                              partials, st);
    else
        launch_rhs<double, true, false, false>(g, rp, tp, tpp, nullptr, nullptr, rhoc, cvec, qd,
                                               bvec, d, cx, cy, cz, rhat, nullptr, nullptr, sc,
                                               partials, st);
}

void launch_rhs_res_state32(const Dims &g, const RhsParams &rp, const double *tp,
                            const double *tpp, const double *rhoc, const double *cvec,
                            const double *qd, const double *bvec, const double *d,
                            const double *cx, const double *cy, const double *cz, float *rhat,
                            double *sout, double *xout, Scal *sc, double *partials,
                            hipStream_t st) {
    if (sout && xout)
        launch_rhs<float, true, true, true>(g, rp, tp, tpp, nullptr, nullptr, rhoc, cvec, qd,
                                            bvec, d, cx, cy, cz, rhat, sout, xout, sc, partials,
                                            st);
    else
        launch_rhs<float, true, false, false>(g, rp, tp, tpp, nullptr, nullptr, rhoc, cvec, qd,
                                              bvec, d, cx, cy, cz, rhat, nullptr, nullptr, sc,
                                              partials, st);
}

void launch_rhs_res_solve64(const Dims &g, const RhsParams &rp, const double *x,
                            const double *b, const double *d, const double *cx,
                            const double *cy, const double *cz, double *rhat, Scal *sc,
                            double *partials, hipStream_t st) {
    launch_rhs<double, false, false, false>(g, rp, nullptr, nullptr, x, b, nullptr, nullptr,
                                            nullptr, nullptr, d, cx, cy, cz, rhat, nullptr,
                                            nullptr, sc, partials, st);
}

void launch_rhs_res_solve32(const Dims &g, const RhsParams &rp, const double *x,
                            const double *b, const double *d, const double *cx,
                            const double *cy, const double *cz, float *rhat, Scal *sc,
                            double *partials, hipStream_t st) {
    launch_rhs<float, false, false, false>(g, rp, nullptr, nullptr, x, b, nullptr, nullptr,
                                           nullptr, nullptr, d, cx, cy, cz, rhat, nullptr,
                                           nullptr, sc, partials, st);
}

template <typename T>
void launch_cg_a(const Dims &g, bool first, const T *r, const T *pold, T *pnew, T *q, T *y,
                 bool yfirst, const T *hx, const T *hy, const T *hz, Scal *sc, double *partials,
                 hipStream_t st) {
    const int32_t yf = yfirst ? 1 : 0;
#define HF_CGA(...)                                                                          \
    HF_ZLAUNCH((__VA_ARGS__), g, st, g, r, pold, pnew, q, y, yf, hx, hy, hz, sc, partials)
    if constexpr (std::is_same<T, float>::value) {
        if (g.quads) {
            if (first) HF_CGA(k_cg_a_zw<T, 4, true>);
            else HF_CGA(k_cg_a_zw<T, 4, false>);
            return;
        }
    }
    if (g.pairs) {
        if (first) HF_CGA(k_cg_a_zw<T, 2, true>);
        else HF_CGA(k_cg_a_zw<T, 2, false>);
        return;
    }
    if (first) HF_CGA(k_cg_a_z<T, true>);
    else HF_CGA(k_cg_a_z<T, false>);
#undef HF_CGA
}
template void launch_cg_a<double>(const Dims &, bool, const double *, const double *, double *,
                                  double *, double *, bool, const double *, const double *,
                                  const double *, Scal *, double *, hipStream_t);
template void launch_cg_a<float>(const Dims &, bool, const float *, const float *, float *,
                                 float *, float *, bool, const float *, const float *,
                                 const float *, Scal *, double *, hipStream_t);

template <typename T>
void launch_cg_b1(const Dims &g, const T *r, const T *hx, const T *hy, const T *hz, T *rw,
                  Scal *sc, double *partials, hipStream_t st) {
    if constexpr (std::is_same<T, float>::value) {
        if (g.quads) {
            HF_ZLAUNCH((k_cg_b1_zw<T, 4>), g, st, g, r, hx, hy, hz, rw, sc, partials);
            return;
        }
    }
    if (g.pairs) {
        HF_ZLAUNCH((k_cg_b1_zw<T, 2>), g, st, g, r, hx, hy, hz, rw, sc, partials);
        return;
    }
    HF_ZLAUNCH(k_cg_b1_z<T>, g, st, g, r, hx, hy, hz, rw, sc, partials);
}
template void launch_cg_b1<double>(const Dims &, const double *, const double *, const double *,
                                   const double *, double *, Scal *, double *,
                                   hipStream_t);
template void launch_cg_b1<float>(const Dims &, const float *, const float *, const float *,
                                  const float *, float *, Scal *, double *, hipStream_t);

void launch_ir_res(const Dims &g, const IrParams &ip, const double *x, const float *y,
                   const float *r0, const Scal *inner, const double *sv, const double *S,
                   const double *d, const double *cx, const double *cy, const double *cz,
                   const double *tp, double *xout, float *rhat, Scal *sc, double *partials,
                   hipStream_t st) {
    HF_ZLAUNCH(k_ir_res_z, g, st, g, ip, x, y, r0, inner, sv, S, d, cx, cy, cz, tp, xout, rhat,
               sc, partials);
}

// (pipelined z-chunking; launch_fin_next reads its partials)
void launch_step64(const Dims &g0, const RhsParams &rp, bool guess, const double *tp,
                   const double *tpp, const double *y, const double *r0, const double *rhoc,
                   const double *cvec, const double *qd, const double *bvec, const double *d,
                   const double *cx, const double *cy, const double *cz, double *xout,
                   double *rnext, Scal *sc, double *partials, hipStream_t st) {
    const Dims g = zdims(g0, kZPipe);
    const int32_t gs = guess ? 1 : 0;
#define HF_STEP64(M, W)                                                                       \
    HF_ZLAUNCH((k_step64_z<M, W>), g, st, g, rp, gs, tp, tpp, y, r0, rhoc, cvec, qd, bvec, d, \
               cx, cy, cz, xout, rnext, sc, partials)
    const int mode = s_mode(rp.isteady, rp.icattaneo);
    const bool wrap = g.per[0] || g.per[1] || g.per[2];
    if (mode == kCattaneo) {
        if (wrap) HF_STEP64(kCattaneo, true); else HF_STEP64(kCattaneo, false);
    } else if (mode == kFourier) {
        if (wrap) HF_STEP64(kFourier, true); else HF_STEP64(kFourier, false);
    } else {
        if (wrap) HF_STEP64(kSteady, true); else HF_STEP64(kSteady, false);
    }
#undef HF_STEP64
}

// (pipelined z-chunking; launch_fin_step reads its partials)
void launch_ir_step(const Dims &g0, const StepParams &sp, bool from_state, const double *tp,
                    const double *tpp, const double *xb, const float *y, const float *r0,
                    const Scal *inner, const double *rhoc, const double *cvec,
                    const double *qb, const double *d, const double *cx, const double *cy,
                    const double *cz, double *xout, float *rres, float *rnext, Scal *sc,
                    double *partials, hipStream_t st) {
    const Dims g = zdims(g0, kZPipe);
#define HF_IR_STEP_W(F, T, C, U)                                                               \
    do {                                                                                       \
        if (wrap)                                                                              \
            HF_ZLAUNCH((k_ir_step_z<F, T, C, true, U>), g, st, g, sp, tp, tpp, xb, y, r0, inner, \
                       rhoc, cvec, qb, d, cx, cy, cz, xout, rres, rnext, sc, partials);        \
        else                                                                                   \
            HF_ZLAUNCH((k_ir_step_z<F, T, C, false, U>), g, st, g, sp, tp, tpp, xb, y, r0,      \
                       inner, rhoc, cvec, qb, d, cx, cy, cz, xout, rres, rnext, sc, partials);  \
    } while (0)
#define HF_IR_STEP(F, T, C)                                                                   \
    do {                                                                                      \
        if (T && sp.rc_mask == 0) HF_IR_STEP_W(F, T, C, true);                                \
        else HF_IR_STEP_W(F, T, C, false);                                                    \
    } while (0)
    const bool tr = sp.isteady == 0, cat = sp.icattaneo == 1;
    const bool wrap = g.per[0] || g.per[1] || g.per[2];
    if (from_state) {
        if (!tr) HF_IR_STEP_W(true, false, false, false);
        else if (cat) HF_IR_STEP(true, true, true);
        else HF_IR_STEP(true, true, false);
    } else {
        if (!tr) HF_IR_STEP_W(false, false, false, false);
        else if (cat) HF_IR_STEP(false, true, true);
        else HF_IR_STEP(false, true, false);
    }
#undef HF_IR_STEP
#undef HF_IR_STEP_W
}

}  // namespace hf
