// hf_device.hpp -- shared device-side definitions of the HeatFlow gfx950 library.
//
// Layout: one HIP thread handles one cell per loop trip; every array over cells is SoA in the
// reference's linear order (x fastest, then y, then z; src/heatflow/mod_hmatrix.f90:72-74).
// A launch uses a fixed grid of `nblk` 256-thread blocks (a multiple of 8); hardware block h
// is mapped to the logical chunk (h % 8) * (nblk / 8) + h / 8, so each of the 8 XCDs of an
// MI355X sweeps one contiguous 1/8 of the domain and z-neighbour planes stay in that XCD's L2.
// All reductions are deterministic: wave-level butterfly, then a fixed-order LDS sum, one
// partial per block, and a one-block kernel sums the partials in block order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hf {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxWaves = 16;            // largest block: 1024 threads (z-marching kernels)
// z-marching stencil tile: kZTX x kZTY threads, one column each
constexpr uint32_t kZTX = 128, kZTY = 4;
constexpr double kTiny = 1.0e-12;        // constants::TINY, src/heatflow/mod_constants.f90:17
constexpr double kStefBoltz = 5.670373e-8;  // src/heatflow/mod_constants.f90:16

// floor(n / d) for n < 2^31 by a multiply-shift: m = ceil(2^(32+l) / d), 2^(l-1) < d <= 2^l.
struct FastDiv {
    uint64_t m;
    uint32_t d, s;
    void init(uint32_t dd) {
        uint32_t l = 0;
        while ((1ull << l) < dd) ++l;
        d = dd;
        s = 32 + l;
        m = ((1ull << s) + dd - 1) / dd;
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        return (uint32_t)(((uint64_t)n * m) >> s);
    }
};

struct Dims {
    uint32_t nx, ny, nz, N, nxny;
    uint32_t ntiles;     // tiles of kBlock consecutive cells
    uint32_t tpb;        // tiles per block (chunked layout)
    uint32_t nblk;       // launch grid (multiple of 8)
    int32_t layout;      // LAYOUT_* bits
    // z-marching stencil kernels: blocks of kZTX x kZTY threads, ztx x zty tiles per plane,
    // z-chunks of `zchunk` planes, ztiles = ztx * zty * chunks.  The FP32 kernels take two
    // x-neighbouring cells per thread when nx is even (`pairs`): tiles of 2 kZTX cells, ztx2 per
    // row, chunks of zchunk2 planes, ztiles2 tiles.  Every stencil launch runs zblk blocks (the
    // larger tile count rounded up to a multiple of 8; blocks past a kernel's tiles only write
    // zero partials), so the finishing kernels see one partial count.
    uint32_t ztx, zty, zchunk, ztiles, zblk;
    uint32_t ztx2, zchunk2, ztiles2;
    int32_t pairs;
    // four x-neighbouring cells per thread (FP32 CG kernels, nx a multiple of 4: `quads`)
    uint32_t ztx4, zchunk4, ztiles4;
    int32_t quads;
    int32_t per[3];      // periodic axes (length >= 3)
    FastDiv dnx, dny;
    // z-chunkings for ~4096 (the active one above), ~2048 and ~512 blocks at 512^3; zdims()
    // selects one for a kernel and its finishing kernel
    struct ZLayout {
        uint32_t zchunk, ztiles, zchunk2, ztiles2, zchunk4, ztiles4, zblk;
    } zlay[3];
    // the two-step pass (k_ir_step2_z): tiles of kP2X x kP2Y columns, p2x x p2y a plane, z-chunks
    // of p2zc planes; p2tiles tiles, p2blk blocks (a multiple of 8)
    uint32_t p2x, p2y, p2zc, p2tiles, p2blk;
};
// 128 x 8 (1024 threads, one block a CU by its LDS): of 16x16, 32x8, 32x16, 64x8, 64x16,
// 128x8 and 256x4 tiles timed on the 512^3 workload, the fastest -- the wide rows keep the
// ring's column cells (one cache line each) few
constexpr uint32_t kP2X = 128, kP2Y = 8;

// z-chunkings (Dims::zlay): kZDefault; kZPipe for the software-pipelined kernels
// (k_rhs_res_z, k_ir_step_z), which favour fewer, longer chunks; kZColumn, whole columns (512
// blocks at 512^3, all resident at once) for the CG kernels with 16-byte accesses (FP64 pairs,
// FP32 quads).
enum : int { kZDefault = 0, kZPipe = 1, kZColumn = 2 };
inline Dims zdims(const Dims &g, int cls) {
    Dims d = g;
    const Dims::ZLayout &l = g.zlay[cls];
    d.zchunk = l.zchunk;
    d.ztiles = l.ztiles;
    d.zchunk2 = l.zchunk2;
    d.ztiles2 = l.ztiles2;
    d.zchunk4 = l.zchunk4;
    d.ztiles4 = l.ztiles4;
    d.zblk = l.zblk;
    return d;
}

// Tile-to-block layouts.  CHUNKED: block b owns tiles [b*tpb, (b+1)*tpb); otherwise block b
// takes tiles b, b+nblk, b+2 nblk, ... (all blocks sweep the grid together).  XCD: hardware
// block h acts as logical block (h % 8) * (nblk / 8) + h / 8, so the blocks one XCD runs are
// neighbours in the logical order (shared L2 for their halos).
enum : int32_t { LAYOUT_CHUNKED = 1, LAYOUT_XCD = 2 };

// Device-resident solver scalars.  Written only by the one-block finishing kernels.
struct Scal {
    double rr;        // current r.r of the (scaled) CG residual
    double rr_start;  // r.r at the start of this solve
    double bnorm2;    // ||b_hat||^2 of the outer system
    double thresh2;   // stop when rr <= thresh2
    double pq, alpha, beta;
    double outer_rr;  // mixed mode: FP64 residual^2 after the last refinement sweep
    double reuse_thresh2;  // mixed mode: the inner target a step that reuses alpha must meet
    int64_t iter;     // iterations of this solve
    int64_t itmax;
    int32_t done;
    int32_t first;    // next CG iteration is the first one (p = r, y = 0)
    int32_t flags;    // HF_FLAG_* bits (sticky until the host clears them)
    int32_t tiny;     // x0 += TINY for the next step (mod_evolve.f90:141)
    int32_t shortcut; // the solve stopped after its first iteration: y = alpha p (p = r0)
    int32_t flags_next;  // NaN bits of the next step's S, computed ahead by the fused step
    int32_t allow_short; // k_fin_a may end a solve after its first iteration
    int32_t varying;     // set at upload: bit 0 rho c_p or tau' rho c_p, bit 1 the volume
                         // differs between cells
    // next step's ||b^||^2 and ||r^0||^2 without / with the TINY shift (fused mixed step)
    double next_bnorm2, next_rr[2];
    // the solution y (CgY below): y holds at least one term (yset); the term alpha p_kpend
    // (alpha = this Scal's alpha) is still to be added (ypend)
    int32_t yset, ypend;
    int64_t kpend;
    // linbcg's own stop rule (hf_solve): ||b||^2 unscaled and the last iteration's weighted
    // residual norm^2 (itol 1: ||r||^2, itol 2: ||r / diag H||^2)
    double wb2, wrr;
    // two reusing steps in one pass (k_ir_step2_z, k_fin_step2): flags of the pass, the sums
    // (per step: ||r'||^2, r^0 . r', ||b^_next||^2, ||r^0_next||^2), and whether each step met
    // its tests (pair_ok[0]: step n; pair_ok[1]: step n+1, given step n)
    int32_t flags2;
    int32_t pair_ok[2];
    double pair_v[8];
};

// The solution y of a CG solve on A^ y = r0, as its consumers read it.  The SpMV pass of
// iteration k >= 1 adds the previous iteration's term alpha_{k-1} p_{k-1} to y, and k_cg_tail,
// queued after a solve, adds the term of its last iteration K >= 1.  A solve that ended in its
// first iteration (K = 0) leaves y = alpha r0 unformed (its consumer reads r0 with that scale);
// one that needed no iteration leaves y = 0.  The values are those an eager update forms.
template <typename T>
struct CgY {
    const T *v;      // y, or r0 (K = 0)
    T scale;         // 1, or alpha (K = 0)
    bool any;        // y is not 0
    __device__ __forceinline__ T at(uint32_t J) const { return scale * v[J]; }
};
template <typename T>
__device__ __forceinline__ CgY<T> cg_y(const Scal *sc, const T *y, const T *r0) {
    CgY<T> c;
    const bool k0 = sc->ypend && !sc->yset;   // only p_0 = r0's term
    c.v = k0 ? r0 : y;
    c.scale = k0 ? (T)sc->alpha : (T)1;
    c.any = sc->ypend || sc->yset;
    return c;
}

enum : int32_t {
    FLAG_NAN_B = 1, FLAG_NAN_Q = 2, FLAG_NAN_QDENS = 4, FLAG_NAN_SCAT = 8, FLAG_NAN_S = 16,
    FLAG_NAN_X = 32, FLAG_TINY_NEXT = 64, FLAG_BREAKDOWN = 128,
    // the second step of a two-step pass (k_ir_step2_z; Scal::flags2): its x, the TINY test of
    // the step after it, the NaN guards of that step's S
    FLAG_NAN_X2 = 256, FLAG_TINY_NEXT2 = 512, FLAG_NAN_SCAT2 = 1024, FLAG_NAN_S2 = 2048
};

__device__ __forceinline__ uint32_t logical_block(const Dims &g) {
    const uint32_t h = blockIdx.x;
    return (g.layout & LAYOUT_XCD) ? (h & 7u) * (g.nblk >> 3) + (h >> 3) : h;
}

// Tiles of this block: t = first; t < end; t += step; cell I = t * kBlock + threadIdx.x.
struct Tiles {
    uint32_t first, end, step;
};

__device__ __forceinline__ Tiles block_tiles(const Dims &g) {
    const uint32_t b = logical_block(g);
    Tiles t;
    if (g.layout & LAYOUT_CHUNKED) {
        const uint64_t f = (uint64_t)b * g.tpb, e = f + g.tpb;
        t.first = (uint32_t)(f < g.ntiles ? f : g.ntiles);
        t.end = (uint32_t)(e < g.ntiles ? e : g.ntiles);
        t.step = 1;
    } else {
        t.first = b;
        t.end = g.ntiles;
        t.step = g.nblk;
    }
    return t;
}

// Loop over the cells of this block (one cell per thread per trip); I < N inside.
#define HF_FOR_CELLS(g, I)                                                          \
    for (Tiles _tl = block_tiles(g); _tl.first < _tl.end; _tl.first += _tl.step)  \
        for (uint32_t I = _tl.first * kBlock + threadIdx.x; I < (g).N; I = (g).N)

struct Nbr {
    uint32_t xm, xp, ym, yp, zm, zp;
    bool hxm, hxp, hym, hyp, hzm, hzp;
    uint32_t x, y, z;
};

// Neighbour indices of cell I; h** = false where the face is a domain boundary (or a
// length-1 periodic axis).  Periodic axes wrap (hmatrixfunc's xp/xm logic, :78-92).
__device__ __forceinline__ Nbr neighbours(const Dims &g, uint32_t I) {
    Nbr n;
    uint32_t t = g.dnx.div(I);
    n.x = I - t * g.nx;
    n.z = g.dny.div(t);
    n.y = t - n.z * g.ny;
    // x
    n.hxm = n.x > 0 || g.per[0];
    n.xm = n.x > 0 ? I - 1 : I + (g.nx - 1);
    n.hxp = n.x + 1 < g.nx || g.per[0];
    n.xp = n.x + 1 < g.nx ? I + 1 : I - (g.nx - 1);
    // y
    n.hym = n.y > 0 || g.per[1];
    n.ym = n.y > 0 ? I - g.nx : I + (g.ny - 1) * g.nx;
    n.hyp = n.y + 1 < g.ny || g.per[1];
    n.yp = n.y + 1 < g.ny ? I + g.nx : I - (g.ny - 1) * g.nx;
    // z
    n.hzm = n.z > 0 || g.per[2];
    n.zm = n.z > 0 ? I - g.nxny : I + (g.nz - 1) * g.nxny;
    n.hzp = n.z + 1 < g.nz || g.per[2];
    n.zp = n.z + 1 < g.nz ? I + g.nxny : I - (g.nz - 1) * g.nxny;
    return n;
}

// Streaming accesses: arrays a kernel touches only at its own cell are loaded and stored with
// the non-temporal hint, so they do not displace the stencil operands (whose neighbours are
// re-read) from L2.
template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void stnt(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

// ---- deterministic reductions ------------------------------------------------------------

template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    }
}

// Block sum over blockDim.x threads (<= 1024); the result is valid in every thread.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *lds /* K*kMaxWaves */) {
    wave_sum<K>(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * kMaxWaves + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += lds[k * kMaxWaves + i];
        v[k] = s;
    }
}

// Per-block partial sums: every block writes its K block sums to partials[k * gridDim.x + b]
// with plain stores.  A one-block finishing kernel launched next on the same stream (kernel
// boundaries make the stores visible across the XCDs' L2s) sums them in block order and updates
// the solver scalars -- no device-scope fences or contended atomics inside the wide kernel.
template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], double *partials) {
    __shared__ double lds[K * kMaxWaves];
    block_sum<K>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = v[k];
    }
}

// Fixed-order sum of nb partials per component, for a one-block kernel of kBlock threads.
template <int K>
__device__ __forceinline__ void sum_partials(const double *partials, uint32_t nb,
                                             double (&out)[K]) {
    __shared__ double lds[K * kMaxWaves];
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += partials[(size_t)k * nb + i];
    }
    block_sum<K>(acc, lds);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = acc[k];
}

// OR-reduce an int flag word of the block into *dst with one atomic per block (order-free).
__device__ __forceinline__ void block_flags(int32_t f, int32_t *dst) {
    __shared__ int32_t sf;
    if (threadIdx.x == 0) sf = 0;
    __syncthreads();
    if (f) atomicOr(&sf, f);
    __syncthreads();
    if (threadIdx.x == 0 && sf) atomicOr(dst, sf);
}

// ---- COO emission (sparse_Hmatrix, src/heatflow/mod_setup.f90:89-152) ------------------------

// a coupling sparse_Hmatrix emits: |H0| >= TINY (mod_setup.f90:130)
__device__ __forceinline__ bool emitted(double c) { return !(fabs(c) < kTiny); }

// emitted couplings of column j in addit order: x+1, x wrap, y+nx, y wrap, z+nxny, z wrap
__device__ __forceinline__ int coo_slots(const Dims &g, const Nbr &n, uint32_t j,
                                         const double *cx, const double *cy, const double *cz,
                                         uint32_t *rows, double *vals) {
    int k = 0;
    if (n.x + 1 < g.nx && emitted(cx[j])) { rows[k] = j + 1; vals[k] = cx[j]; ++k; }
    if (g.per[0] && n.x == 0) {
        uint32_t i = j + g.nx - 1;
        if (emitted(cx[i])) { rows[k] = i; vals[k] = cx[i]; ++k; }
    }
    if (g.ny > 1 && n.y + 1 < g.ny && emitted(cy[j])) { rows[k] = j + g.nx; vals[k] = cy[j]; ++k; }
    if (g.per[1] && n.y == 0) {
        uint32_t i = j + (g.ny - 1) * g.nx;
        if (emitted(cy[i])) { rows[k] = i; vals[k] = cy[i]; ++k; }
    }
    if (g.nz > 1 && n.z + 1 < g.nz && emitted(cz[j])) { rows[k] = j + g.nxny; vals[k] = cz[j]; ++k; }
    if (g.per[2] && n.z == 0) {
        uint32_t i = j + (g.nz - 1) * g.nxny;
        if (emitted(cz[i])) { rows[k] = i; vals[k] = cz[i]; ++k; }
    }
    return k;
}

}  // namespace hf
