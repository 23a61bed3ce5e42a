// hf_api.hip -- host side of the C-ABI declared in include/heatflow_hip.h.
//
// One hf_ctx owns one HIP stream and all device buffers of one grid.  The time step mirrors
// evolution::simulate (src/heatflow/mod_evolve.f90:45-177): heater (when time-dependent) ->
// fused S / x0 / residual -> CG on the Jacobi-scaled operator with device-resident scalars ->
// x = x0 + s*y -> history rotation by pointer swap.  The host synchronises once per batch of
// CG iterations (the batch length follows the previous step's iteration count) to read the
// convergence flag; no per-iteration host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/heatflow_hip.h"
#include "hf_device.hpp"
#include "hf_kernels.hpp"
#include "hf_ctx.hpp"

using namespace hf;

static thread_local std::string g_create_error;

// ---------------------------------------------------------------------------------------------

static int set_err(hf_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg; else g_create_error = msg;
    return code;
}

#define HIPCHK(c, call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err(c, HF_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static int dalloc(hf_ctx *c, T **ptr, size_t n) {
    if (*ptr) return HF_OK;
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(c, HF_ERR_ALLOC, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                            " bytes failed: " + hipGetErrorString(e));
    }
    c->allocs.push_back(v);
    *ptr = (T *)v;
    return HF_OK;
}

#define TRY(x)                       \
    do {                             \
        int rc_ = (x);               \
        if (rc_ != HF_OK) return rc_; \
    } while (0)

static int check_launch(hf_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, HF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return HF_OK;
}

// CG solves may stop after their first iteration unless debug bit 1 turns that off
static bool allow_short(const hf_ctx *c) { return (c->debug & 2) == 0; }
// index mask of the rho c_p / tau' rho c_p reads: 0 when both are uniform (every cell reads
// element 0; debug bit 3 keeps the per-cell reads)
static uint32_t rc_mask(const hf_ctx *c) {
    return c->uniform_rc && (c->debug & 8) == 0 ? 0u : ~0u;
}
// bytes per cell of the rho c_p and tau' rho c_p reads of a step kernel
static double rc_bytes(const hf_ctx *c) { return rc_mask(c) ? 16.0 : 0.0; }

// ---- kernel timing -----------------------------------------------------------------------

// Opens a timed launch: records the start event on the context stream; returns the slot or -1.
static long tbeg(hf_ctx *c, int id, double bytes) {
    if (!c->timing) return -1;
    if (c->tused == c->tpool.size()) {
        hf_ctx::Timed t{};
        if (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        c->tpool.push_back(t);
    }
    hf_ctx::Timed &t = c->tpool[c->tused];
    t.id = id;
    t.bytes = bytes;
    t.valid = true;
    (void)hipEventRecord(t.a, c->st);
    return (long)c->tused++;
}

static void tend(hf_ctx *c, long slot) {
    if (slot >= 0) (void)hipEventRecord(c->tpool[slot].b, c->st);
}

// cus: compute units of the device (resident blocks per launch round scale with it)
static void make_dims(Dims &g, const hf_params &p, int cus) {
    g.nx = (uint32_t)p.nx; g.ny = (uint32_t)p.ny; g.nz = (uint32_t)p.nz;
    g.nxny = g.nx * g.ny;
    g.N = g.nxny * g.nz;
    g.per[0] = p.periodic[0] && p.nx >= 3;
    g.per[1] = p.periodic[1] && p.ny >= 3;
    g.per[2] = p.periodic[2] && p.nz >= 3;
    g.dnx.init(g.nx);
    g.dny.init(g.ny);
    // launch layout: tiles of kBlock cells over a grid of at most `cap` blocks (multiple of 8);
    // HF_NBLK / HF_LAYOUT override the defaults (tuning sweeps)
    uint32_t cap = 16384;
    if (const char *e = getenv("HF_NBLK")) cap = std::max(8u, (uint32_t)atoi(e) / 8u * 8u);
    g.layout = LAYOUT_CHUNKED | LAYOUT_XCD;
    if (const char *e = getenv("HF_LAYOUT")) g.layout = atoi(e);
    g.ntiles = (uint32_t)(((uint64_t)g.N + kBlock - 1) / kBlock);
    uint64_t nb = std::min<uint64_t>(g.ntiles, cap);
    g.tpb = (uint32_t)((g.ntiles + nb - 1) / nb);
    nb = (g.ntiles + g.tpb - 1) / g.tpb;
    g.nblk = (uint32_t)((nb + 7) / 8 * 8);
    // z-marching tiles of kZTX x kZTY columns: enough z-chunks for ~`zcap` blocks, one chunking
    // per kernel class (measured at 512^3, 1 to 12% apart; HF_ZBLK, HF_ZBLK_PIPE, HF_ZBLK_COL
    // override).  Fewer, longer chunks cost fewer prologue planes; the CG kernels with 8-byte
    // accesses (FP32 pairs) ran best at ~4096 blocks, the pipelined kernels at ~2048 and the
    // 16-byte ones (FP64 pairs, FP32 quads) with whole columns (one round of resident blocks).
    const uint32_t tx = kZTX, ty = kZTY;
    g.zty = (g.ny + ty - 1) / ty;
    g.ztx = (g.nx + tx - 1) / tx;
    g.ztx2 = (g.nx + 2 * tx - 1) / (2 * tx);
    g.ztx4 = (g.nx + 4 * tx - 1) / (4 * tx);
    g.pairs = (g.nx % 2 == 0) && getenv("HF_NO_PAIRS") == nullptr;
    g.quads = g.pairs && (g.nx % 4 == 0) && getenv("HF_NO_QUADS") == nullptr;
    // Chunk length and tile count for about zcap blocks, tiles of `ztx` x zty per plane.  A
    // launch runs in rounds of `slots` resident blocks, so its time goes about as rounds x
    // (planes per chunk + 2 prologue planes).  The zcap chunking is kept unless a chunk count
    // within a factor 2 of it costs 5% less by that measure: a tile count just past a multiple
    // of the slots leaves a round nearly empty (at 384^3 whole columns gave 576 blocks for 512
    // slots).
    auto chunking = [&](uint32_t zcap, uint64_t slots, uint32_t ztx, uint32_t *zchunk,
                        uint32_t *ztiles) {
        const uint64_t per_plane = (uint64_t)ztx * g.zty;
        auto chunks = [&](uint64_t n) {   // chunk count actually needed for ~n chunks
            const uint64_t zc = (g.nz + n - 1) / n;
            return (g.nz + zc - 1) / zc;
        };
        auto cost = [&](uint64_t nn) {
            const uint64_t zc = (g.nz + nn - 1) / nn;
            return ((per_plane * nn + slots - 1) / slots) * (zc + 2);
        };
        const uint64_t n0 = chunks(std::min<uint64_t>(
            g.nz, std::max<uint64_t>(1, (zcap + per_plane - 1) / per_plane)));
        const uint64_t hi = std::max<uint64_t>(1, std::min<uint64_t>(g.nz, 2 * n0));
        const uint64_t lo = std::min<uint64_t>(hi, std::max<uint64_t>(1, n0 / 2));
        uint64_t best = n0, best_cost = cost(n0);
        for (uint64_t n = lo; n <= hi; ++n) {
            const uint64_t nn = chunks(n);
            if (cost(nn) * 20 < best_cost * 19) {
                best = nn;
                best_cost = cost(nn);
            }
        }
        *zchunk = (uint32_t)((g.nz + best - 1) / best);
        *ztiles = (uint32_t)(per_plane * best);
    };
    auto env_cap = [](const char *name, uint32_t dflt) {
        const char *e = getenv(name);
        return e ? std::max(8u, (uint32_t)atoi(e)) : dflt;
    };
    const uint32_t caps[3] = {env_cap("HF_ZBLK", 4096), env_cap("HF_ZBLK_PIPE", 2048),
                              env_cap("HF_ZBLK_COL", 512)};
    // resident 512-thread blocks per CU: ~4 for the default class, 2 for the pipelined kernels
    // (128 VGPRs) and the 16-byte CG kernels (~90 VGPRs)
    const uint64_t slots[3] = {4ull * (uint64_t)cus, 2ull * (uint64_t)cus, 2ull * (uint64_t)cus};
    for (int k = 0; k < 3; ++k) {
        Dims::ZLayout &l = g.zlay[k];
        chunking(caps[k], slots[k], g.ztx, &l.zchunk, &l.ztiles);
        chunking(caps[k], slots[k], g.ztx2, &l.zchunk2, &l.ztiles2);
        chunking(caps[k], slots[k], g.ztx4, &l.zchunk4, &l.ztiles4);
        uint32_t zt = l.ztiles;
        if (g.pairs) zt = std::max(zt, l.ztiles2);
        if (g.quads) zt = std::max(zt, l.ztiles4);
        l.zblk = (zt + 7) / 8 * 8;
    }
    const Dims::ZLayout &l0 = g.zlay[kZDefault];
    g.zchunk = l0.zchunk;
    g.ztiles = l0.ztiles;
    g.zchunk2 = l0.zchunk2;
    g.ztiles2 = l0.ztiles2;
    g.zchunk4 = l0.zchunk4;
    g.ztiles4 = l0.ztiles4;
    g.zblk = l0.zblk;
    {
        // the two-step pass: one block a CU (its LDS), and a z-march of zc planes takes zc + 4
        // trips (the pipeline's fill and drain), so the chunk count c minimises
        // ceil(tiles(c) / CUs) * (ceil(nz / c) + 4): whole columns when a plane has about as many
        // tiles as there are CUs (512^3 on 256 CUs: 256 columns, one round)
        g.p2x = (g.nx + kP2X - 1) / kP2X;
        g.p2y = (g.ny + kP2Y - 1) / kP2Y;
        const uint64_t per_plane = (uint64_t)g.p2x * g.p2y;
        uint32_t best_c = 1;
        uint64_t best = ~0ull;
        for (uint32_t c = 1; c <= std::max(1u, g.nz / 16u); ++c) {
            const uint32_t zc = (g.nz + c - 1) / c;
            if ((g.nz + zc - 1) / zc != c) continue;   // c chunks of zc planes, none empty
            const uint64_t rounds = (per_plane * c + cus - 1) / cus;
            const uint64_t cost = rounds * (zc + 4);
            if (cost < best) {
                best = cost;
                best_c = c;
            }
        }
        if (const char *e = getenv("HF_P2_CHUNKS")) best_c = std::max(1, atoi(e));
        g.p2zc = (g.nz + best_c - 1) / best_c;
        g.p2tiles = (uint32_t)(per_plane * ((g.nz + g.p2zc - 1) / g.p2zc));
        g.p2blk = (g.p2tiles + 7) / 8 * 8;
    }
}

// the largest z-march grid over the chunkings (partial sums per block)
static uint32_t max_zblk(const Dims &g) {
    return std::max(std::max(g.zlay[0].zblk, std::max(g.zlay[1].zblk, g.zlay[2].zblk)), g.p2blk);
}

// the z-chunking of the CG kernels in precision T: whole columns for the 16-byte-access
// kernels (FP64 pairs, FP32 quads)
template <typename T>
static Dims cg_dims(const Dims &g) {
    const bool wide = std::is_same<T, double>::value ? g.pairs : g.quads;
    return wide ? zdims(g, kZColumn) : g;
}

// ---------------------------------------------------------------------------------------------

extern "C" void hf_params_default(hf_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 1;
    p->time_step = 1.0;      // mod_inputs.f90:226
    p->icattaneo = 1;        // :229
    p->isteady = 0;
    p->br = 1.0;             // :242
    p->freq = 1.0;           // :227
    p->tol = 1.0e-32;        // mod_evolve.f90:143
    p->itmax = 50000;        // :144
    p->precision = HF_PREC_FP64;
    p->inner_tol = 1.0e-6;
    p->inner_itmax = 200;
}

extern "C" const char *hf_version(void) { return "heatflow_hip 0.1 (gfx950)"; }

extern "C" const char *hf_last_error(const hf_ctx *c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

static bool allocs_base(hf_ctx *c, int *rc) {
    const size_t N = c->g.N;
    int r = HF_OK;
    if ((r = dalloc(c, &c->kappa, N)) || (r = dalloc(c, &c->rhoc, N)) ||
        (r = dalloc(c, &c->tau_s, N)) || (r = dalloc(c, &c->cvec, N)) ||
        (r = dalloc(c, &c->ells, N)) || (r = dalloc(c, &c->lx, N)) || (r = dalloc(c, &c->ly, N)) ||
        (r = dalloc(c, &c->lz, N)) || (r = dalloc(c, &c->vol, N)) ||
        (r = dalloc(c, &c->iheater, N)) || (r = dalloc(c, &c->d, N)) ||
        (r = dalloc(c, &c->cx, N)) || (r = dalloc(c, &c->cy, N)) || (r = dalloc(c, &c->cz, N)) ||
        (r = dalloc(c, &c->bvec, N)) || (r = dalloc(c, &c->sv, N)) || (r = dalloc(c, &c->qd, N)) ||
        (r = dalloc(c, &c->T[0], N)) || (r = dalloc(c, &c->T[1], N)) ||
        (r = dalloc(c, &c->sc, 2)) ||
        (r = dalloc(c, &c->partials, (size_t)8 * std::max(c->g.nblk, max_zblk(c->g))))) {
        *rc = r;
        return false;
    }
    c->sci = c->sc + 1;   // outer and inner solver scalars side by side: one read-back
    return true;
}

static int alloc_mode(hf_ctx *c) {
    const size_t N = c->g.N;
    if (c->p.precision == HF_PREC_FP32_IR) {
        TRY(dalloc(c, &c->fx, N)); TRY(dalloc(c, &c->fy, N)); TRY(dalloc(c, &c->fz, N));
        TRY(dalloc(c, &c->r32, N)); TRY(dalloc(c, &c->p32[0], N)); TRY(dalloc(c, &c->p32[1], N));
        TRY(dalloc(c, &c->q32, N)); TRY(dalloc(c, &c->y32, N));
        TRY(dalloc(c, &c->S, N)); TRY(dalloc(c, &c->T[2], N));
        TRY(dalloc(c, &c->rn[0], N)); TRY(dalloc(c, &c->rn[1], N));
        TRY(dalloc(c, &c->rw32, N)); TRY(dalloc(c, &c->qb, N));
        c->qb_valid = false;
    } else {
        TRY(dalloc(c, &c->hx, N)); TRY(dalloc(c, &c->hy, N)); TRY(dalloc(c, &c->hz, N));
        TRY(dalloc(c, &c->r64, N)); TRY(dalloc(c, &c->p64[0], N)); TRY(dalloc(c, &c->p64[1], N));
        TRY(dalloc(c, &c->rw64, N));
        TRY(dalloc(c, &c->q64, N)); TRY(dalloc(c, &c->y64, N));
        TRY(dalloc(c, &c->r64b, N)); TRY(dalloc(c, &c->T[2], N));
    }
    return HF_OK;
}

extern "C" int hf_create(hf_ctx **out, const hf_params *p) {
    if (!out || !p) return set_err(nullptr, HF_ERR_ARG, "hf_create: null argument");
    *out = nullptr;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1)
        return set_err(nullptr, HF_ERR_ARG, "hf_create: nx, ny, nz must be >= 1");
    const long double n = (long double)p->nx * p->ny * p->nz;
    if (n > 2147483647.0L) return set_err(nullptr, HF_ERR_ARG, "hf_create: more than 2^31-1 cells");
    if (p->nx == 1 && n > 1)
        return set_err(nullptr, HF_ERR_ARG,
                       "hf_create: nx == 1 with ny*nz > 1 is not supported (the reference's "
                       "sparse_Hmatrix emits duplicate offsets and overruns its COO length)");
    for (int a = 0; a < 3; ++a) {
        int64_t len = a == 0 ? p->nx : (a == 1 ? p->ny : p->nz);
        if (p->periodic[a] && len < 3)
            return set_err(nullptr, HF_ERR_ARG,
                           "hf_create: periodic axes need length >= 3 (length 1 zeroes the "
                           "diagonal in hmatrixfunc's wrap test, length 2 duplicates COO offsets)");
    }
    if (!(p->time_step > 0.0) && p->isteady == 0)
        return set_err(nullptr, HF_ERR_ARG, "hf_create: time_step must be > 0");
    if (p->precision != HF_PREC_FP64 && p->precision != HF_PREC_FP32_IR)
        return set_err(nullptr, HF_ERR_ARG, "hf_create: unknown precision mode");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_err(nullptr, HF_ERR_NO_DEVICE, "hf_create: no HIP device");
    }
    hf_ctx *c = new hf_ctx();
    c->p = *p;
    hipGetDevice(&c->dev);
    if (const char *e = getenv("HF_SMALL_MAX_CELLS")) c->small_max = (uint32_t)atol(e);
    {
        // the static scratch of hf_export_coo_noalloc on this device (once per device)
        const hipError_t e = hf::noalloc_setup(c->dev);
        if (e != hipSuccess) {
            delete c;
            return set_err(nullptr, HF_ERR_HIP,
                           std::string("hf_create: hipGetSymbolAddress(scratch): ") +
                               hipGetErrorString(e));
        }
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->dev) != hipSuccess ||
        cus < 1) {
        (void)hipGetLastError();
        cus = 256;
    }
    make_dims(c->g, c->p, cus);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        std::string m = "hf_create: hipStreamCreate failed";
        delete c;
        return set_err(nullptr, HF_ERR_HIP, m);
    }
    int rc = HF_OK;
    if (!allocs_base(c, &rc) || (rc = alloc_mode(c)) != HF_OK ||
        hipHostMalloc((void **)&c->h_sc, 2 * sizeof(Scal) + 64, hipHostMallocCoherent) !=
            hipSuccess) {
        std::string m = c->err.empty() ? "hf_create: allocation failed" : c->err;
        hf_destroy(c);
        return set_err(nullptr, rc ? rc : HF_ERR_ALLOC, m);
    }
    c->h_sci = c->h_sc + 1;
    c->h_seq = (int64_t *)(c->h_sc + 2);
    *(volatile int64_t *)c->h_seq = 0;
    hipMemsetAsync(c->sc, 0, 2 * sizeof(Scal), c->st);
    if (hipStreamSynchronize(c->st) != hipSuccess) {
        hf_destroy(c);
        return set_err(nullptr, HF_ERR_HIP, "hf_create: stream sync failed");
    }
    // static assembly parameters
    AsmParams &ap = c->ap;
    ap.inverse_time = 1.0 / p->time_step;
    for (int f = 0; f < 6; ++f) {
        ap.kb[f] = p->kappa_bound[f];
        ap.tb[f] = p->t_bath[f];
        ap.cg[f] = p->cg[f] != 0;
        ap.cg_pd[f] = 0.0;
    }
    ap.br = p->br;
    ap.isteady = p->isteady;
    ap.icattaneo = p->icattaneo;
    for (int a = 0; a < 3; ++a) ap.periodic[a] = p->periodic[a] != 0;
    *out = c;
    return HF_OK;
}

extern "C" void hf_destroy(hf_ctx *c) {
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    hf::adjoint_free(c->adj);
    for (auto &t : c->tpool) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    for (void *v : c->allocs) hipFree(v);
    if (c->h_sc) hipHostFree(c->h_sc);   // h_sci is h_sc + 1
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

// derived per-cell fields: lin_rhoc = rho*c_p (mod_setup.f90:64), tau*lin_rhoc (S_catS),
// em*L1*L2*StefBoltz (mod_heating.f90:121-122)
__global__ void k_derived(uint32_t N, const double *__restrict__ rho,
                          const double *__restrict__ hc, const double *__restrict__ tau_s,
                          const double *__restrict__ em, const double *__restrict__ lx,
                          const double *__restrict__ ly, double *__restrict__ rhoc,
                          double *__restrict__ cvec, double *__restrict__ ells) {
    uint32_t I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double rc = rho[I] * hc[I];
    rhoc[I] = rc;
    cvec[I] = tau_s[I] * rc;
    ells[I] = em ? em[I] * lx[I] * ly[I] * kStefBoltz : 0.0;
}

// Sets bit 0 of sc->varying when rho c_p or tau' rho c_p differs from its value at cell 0
// somewhere, bit 1 when the volume does, bit 2 when a cell length along some axis does
// (sc->varying cleared before).  Uniform fields let the step and adjoint kernels read element 0
// only; the adjoint needs uniform lengths per axis.
__global__ void __launch_bounds__(kBlock)
k_uniform_check(uint32_t N, const double *__restrict__ rhoc, const double *__restrict__ cvec,
                const double *__restrict__ vol, const double *__restrict__ lx,
                const double *__restrict__ ly, const double *__restrict__ lz, Scal *sc) {
    const double r0 = rhoc[0], c0 = cvec[0], v0 = vol[0], x0 = lx[0], y0 = ly[0], z0 = lz[0];
    int32_t diff = 0;
    for (uint32_t I = blockIdx.x * kBlock + threadIdx.x; I < N; I += gridDim.x * kBlock) {
        if (!(ldnt(&rhoc[I]) == r0) || !(ldnt(&cvec[I]) == c0)) diff |= 1;
        if (!(ldnt(&vol[I]) == v0)) diff |= 2;
        if (!(ldnt(&lx[I]) == x0) || !(ldnt(&ly[I]) == y0) || !(ldnt(&lz[I]) == z0)) diff |= 4;
    }
    block_flags(diff, &sc->varying);
}

extern "C" int hf_upload_cells(hf_ctx *c, const double *kappa, const double *rho,
                               const double *heat_capacity, const double *tau_scaled,
                               const double *em, const double *len_x, const double *len_y,
                               const double *len_z, const double *volume,
                               const int64_t *iheater, const double *qdens_extra) {
    if (!c) return HF_ERR_ARG;
    c->carry = false;   // device state or scratch changes: no carried residual
    hf::adjoint_invalidate(c->adj);   // new fields: the checkpoints no longer belong to them
    if (!kappa || !rho || !heat_capacity || !tau_scaled || !len_x || !len_y || !len_z || !volume)
        return set_err(c, HF_ERR_ARG, "hf_upload_cells: required array is NULL");
    const size_t N = c->g.N, B8 = N * sizeof(double);
    double *rho_d = nullptr, *hc_d = nullptr, *em_d = nullptr;
    // rho and heat_capacity are staged in the (not yet used) CG work buffers
    rho_d = c->p.precision == HF_PREC_FP64 ? c->r64 : c->S;
    hc_d = c->p.precision == HF_PREC_FP64 ? c->q64 : c->T[2];
    HIPCHK(c, hipMemcpyAsync(c->kappa, kappa, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(rho_d, rho, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(hc_d, heat_capacity, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(c->tau_s, tau_scaled, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(c->lx, len_x, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(c->ly, len_y, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(c->lz, len_z, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(c->vol, volume, B8, hipMemcpyDefault, c->st));
    if (em) {
        em_d = c->qd;  // staged; qd is rebuilt at assembly
        HIPCHK(c, hipMemcpyAsync(em_d, em, B8, hipMemcpyDefault, c->st));
    }
    hipLaunchKernelGGL(k_derived, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->st,
                       (uint32_t)N, rho_d, hc_d, c->tau_s, em_d, c->lx, c->ly, c->rhoc, c->cvec,
                       c->ells);
    TRY(check_launch(c, "k_derived"));
    {
        int32_t varying = 0;
        HIPCHK(c, hipMemsetAsync(&c->sc->varying, 0, sizeof(int32_t), c->st));
        const uint32_t nb = (uint32_t)std::min<size_t>((N + kBlock - 1) / kBlock, 8192);
        hipLaunchKernelGGL(k_uniform_check, dim3(nb), dim3(kBlock), 0, c->st, (uint32_t)N,
                           c->rhoc, c->cvec, c->vol, c->lx, c->ly, c->lz, c->sc);
        TRY(check_launch(c, "k_uniform_check"));
        HIPCHK(c, hipMemcpyAsync(&varying, &c->sc->varying, sizeof(int32_t),
                                 hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->uniform_rc = (varying & 1) == 0;
        c->uniform_vol = (varying & 2) == 0;
        c->uniform_len = (varying & 4) == 0;
    }
    if (qdens_extra) {
        TRY(dalloc(c, &c->qextra, N));
        HIPCHK(c, hipMemcpyAsync(c->qextra, qdens_extra, B8, hipMemcpyDefault, c->st));
    }
    // heater bookkeeping on the host: heated_volume is the sequential sum of grid%volume
    // over iheater > 0 in loop order (mod_heating.f90:137-141)
    std::vector<int64_t> ih(iheater ? N : 0);
    std::vector<double> vh, emh;
    if (iheater) HIPCHK(c, hipMemcpy(ih.data(), iheater, N * 8, hipMemcpyDefault));
    c->any_heater = false;
    c->dyn_heat = false;
    c->any_em = false;
    c->heated_volume = 0.0;
    std::vector<int8_t> ih8(N, 0);
    if (iheater) {
        vh.resize(N);
        HIPCHK(c, hipMemcpy(vh.data(), volume, B8, hipMemcpyDefault));
        bool other_modes = false;
        for (size_t i = 0; i < N; ++i) {
            int64_t m = ih[i];
            ih8[i] = m <= 0 ? 0 : (m <= 6 ? (int8_t)m : 7);
            if (m > 0) {
                c->any_heater = true;
                c->heated_volume = c->heated_volume + vh[i];
                if (m != 1 && m <= 6) other_modes = true;
            }
        }
        bool any_em = false;
        if (em && c->any_heater) {
            emh.resize(N);
            HIPCHK(c, hipMemcpy(emh.data(), em, B8, hipMemcpyDefault));
            for (size_t i = 0; i < N; ++i)
                if (emh[i] != 0.0) { any_em = true; break; }
        }
        c->dyn_heat = c->any_heater && (other_modes || any_em);
        c->any_em = any_em;
    }
    HIPCHK(c, hipMemcpyAsync(c->iheater, ih8.data(), N, hipMemcpyHostToDevice, c->st));
    // CG_dir power densities use grid(1,1,1), (1,1,nz), (1,ny,1), (nx,1,1) volumes
    // (mod_boundary.f90:225-248)
    {
        const hf_params &p = c->p;
        auto vol_at = [&](int64_t x, int64_t y, int64_t z) -> double {
            double v = 0.0;
            size_t idx = (size_t)(x + p.nx * (y + p.ny * z));
            hipMemcpy(&v, volume + idx, sizeof(double), hipMemcpyDefault);
            return v;
        };
        const double nx = (double)p.nx, ny = (double)p.ny, nz = (double)p.nz;
        if (p.cg[5]) c->ap.cg_pd[5] = p.t_bath_cg / (vol_at(0, 0, p.nz - 1) * nx * ny);
        if (p.cg[4]) c->ap.cg_pd[4] = p.t_bath_cg / (vol_at(0, 0, 0) * nx * ny);
        if (p.cg[2]) c->ap.cg_pd[2] = p.t_bath_cg / (vol_at(0, 0, 0) * nx * nz);
        if (p.cg[3]) c->ap.cg_pd[3] = p.t_bath_cg / (vol_at(0, p.ny - 1, 0) * nx * nz);
        if (p.cg[0]) c->ap.cg_pd[0] = p.t_bath_cg / (vol_at(0, 0, 0) * ny * nz);
        if (p.cg[1]) c->ap.cg_pd[1] = p.t_bath_cg / (vol_at(p.nx - 1, 0, 0) * ny * nz);
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->uploaded = true;
    c->assembled = false;
    return HF_OK;
}

static HeatParams heat_params(const hf_ctx *c, int64_t itime) {
    // mode powers of this step (src/heatflow/mod_heating.f90:39-117), host libm
    const double PI = 3.141592653589793;
    const hf_params &p = c->p;
    HeatParams hp{};
    const double POWER = p.power_in;
    const double time = p.time_step * (double)itime;
    const double time_pulse = (double)p.heated_steps * p.time_step;
    hp.qmode[0] = 0.0;
    hp.qmode[1] = POWER;
    hp.qmode[2] = (time <= time_pulse) ? POWER : 0.0;
    {
        double x = time * 2.0 * PI * p.freq;
        double x2 = p.time_step * 2.0 * PI * p.freq;
        hp.qmode[3] = POWER * 0.5 * ((x2) - sin(x + x2) * cos(x + x2) + sin(x) * cos(x)) / x2;
    }
    hp.qmode[4] = POWER * pow(sin(time * 2.0 * PI * p.freq), 2.0);
    hp.qmode[5] = POWER * pow(sin(time * 2.0 * PI * p.freq), 2.0);
    hp.q5b = POWER * 2.0 * PI * p.freq;
    hp.q5c = sin(2.0 * time * 2.0 * PI * p.freq);
    hp.qmode[6] = (itime == 1) ? POWER : 0.0;
    hp.dt2 = pow(p.time_step, 2.0);
    hp.tb4 = pow(p.t_bath_rad, 4.0);
    hp.heated_volume = c->heated_volume;
    hp.any_heater = c->any_heater ? 1 : 0;
    return hp;
}

// reads back `count` consecutive scalar blocks
// Brings `count` solver-scalar records to the host once the stream's work before them is done.
// The device copies them into the pinned mirror and then raises a sequence number (k_publish),
// which the host watches for up to 2 ms before it falls back to a blocking synchronisation of
// the stream (which also reports a failed kernel): watching the word saves the blocking wait's
// wake-up latency, tens of microseconds a step.
static_assert(sizeof(Scal) % 8 == 0, "k_publish copies 8-byte words");
static int sync_scal(hf_ctx *c, Scal *dev, Scal *host, int count = 1) {
    const int64_t want = ++c->seq;
    launch_publish(dev, host, count * sizeof(Scal), c->h_seq, want, c->st);
    TRY(check_launch(c, "k_publish"));
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t k = 1;; ++k) {
        if (*(volatile int64_t *)c->h_seq == want) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return HF_OK;
        }
        if ((k & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
            break;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::atomic_thread_fence(std::memory_order_acquire);
    return HF_OK;
}

extern "C" int hf_assemble(hf_ctx *c) {
    if (!c) return HF_ERR_ARG;
    c->carry = false;   // device state or scratch changes: no carried residual
    hf::adjoint_invalidate(c->adj);   // a new operator: the checkpoints no longer belong to it
    if (!c->uploaded) return set_err(c, HF_ERR_STATE, "hf_assemble: cells not uploaded");
    launch_assemble(c->g, c->ap, c->kappa, c->lx, c->ly, c->lz, c->rhoc, c->tau_s, c->d, c->cx,
                    c->cy, c->cz, c->bvec, c->sc, c->st);
    c->qb_valid = false;
    TRY(check_launch(c, "k_assemble"));
    launch_scale(c->g, c->d, c->cx, c->cy, c->cz, c->sv, c->hx, c->hy, c->hz, c->fx, c->fy, c->fz,
                 c->st);
    TRY(check_launch(c, "k_scale"));
    if (!c->dyn_heat) {
        // time-invariant Qdens (constant heaters, no radiative term): built once
        HeatParams hp = heat_params(c, 1);
        hp.tb4 = 0.0;
        launch_qdens(c->g, hp, c->iheater, c->ells, c->tau_s, c->kappa /* unused: ells == 0 */,
                     c->qextra, c->qd, c->sc, c->st);
        c->qb_valid = false;
        TRY(check_launch(c, "k_qdens"));
    }
    TRY(sync_scal(c, c->sc, c->h_sc));
    c->setup_flags = c->h_sc->flags & (FLAG_NAN_B | FLAG_NAN_Q | FLAG_NAN_QDENS);
    c->assembled = true;
    return HF_OK;
}

extern "C" int hf_export_coo(hf_ctx *c, int64_t *n, int64_t *len, int64_t *count, double *val,
                             int64_t *irow, int64_t *jcol) {
    if (!c) return HF_ERR_ARG;
    c->carry = false;   // device state or scratch changes: no carried residual
    if (!c->assembled) return set_err(c, HF_ERR_STATE, "hf_export_coo: not assembled");
    const hf_params &p = c->p;
    int64_t L = 7 * p.nx * p.ny * p.nz - 2 * (p.nx * p.ny + p.ny * p.nz + p.nz * p.nx);
    if (p.periodic[0]) L += 2 * p.ny * p.nz;
    if (p.periodic[1]) L += 2 * p.nz * p.nx;
    if (p.periodic[2]) L += 2 * p.nx * p.ny;
    const size_t N = c->g.N;
    const size_t nb = (N + kBlock - 1) / kBlock;
    int64_t *cnt = nullptr, *off = nullptr, *bsum = nullptr, *total = nullptr;
    double *dval = nullptr;
    int64_t *dirow = nullptr, *djcol = nullptr;
    auto cleanup = [&]() {
        hipFree(cnt); hipFree(off); hipFree(bsum); hipFree(total);
        hipFree(dval); hipFree(dirow); hipFree(djcol);
    };
    if (hipMalloc(&cnt, N * 8) != hipSuccess || hipMalloc(&off, N * 8) != hipSuccess ||
        hipMalloc(&bsum, nb * 8) != hipSuccess || hipMalloc(&total, 8) != hipSuccess) {
        cleanup();
        return set_err(c, HF_ERR_ALLOC, "hf_export_coo: scratch allocation failed");
    }
    const bool want = val && irow && jcol;
    const int64_t cap = std::max<int64_t>(L, (int64_t)(7 * N));
    if (want && (hipMalloc(&dval, cap * 8) != hipSuccess || hipMalloc(&dirow, cap * 8) != hipSuccess ||
                 hipMalloc(&djcol, cap * 8) != hipSuccess)) {
        cleanup();
        return set_err(c, HF_ERR_ALLOC, "hf_export_coo: output allocation failed");
    }
    int64_t cnt_host = 0;
    launch_coo(c->g, c->d, c->cx, c->cy, c->cz, cnt, off, bsum, total, want ? dval : nullptr,
               dirow, djcol, L, &cnt_host, c->st);
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && want) {
        // the reference never writes past len; an over-full export keeps the first len entries
        int64_t ncopy = L;
        e = hipMemcpy(val, dval, ncopy * 8, hipMemcpyDefault);
        if (e == hipSuccess) e = hipMemcpy(irow, dirow, ncopy * 8, hipMemcpyDefault);
        if (e == hipSuccess) e = hipMemcpy(jcol, djcol, ncopy * 8, hipMemcpyDefault);
    }
    cleanup();
    if (e != hipSuccess) return set_err(c, HF_ERR_HIP, std::string("hf_export_coo: ") + hipGetErrorString(e));
    if (n) *n = (int64_t)N;
    if (len) *len = L;
    if (count) *count = cnt_host;
    return HF_OK;
}

namespace hf {
void reset_step_heuristics(hf_ctx *c) {
    const hf_ctx fresh{};
    c->carry = false;
    c->tiny_guess = fresh.tiny_guess;
    c->last_sweeps = fresh.last_sweeps;
    c->reuse_ok = fresh.reuse_ok;
    c->reuse_pause = fresh.reuse_pause;
    c->reuse_miss_run = fresh.reuse_miss_run;
    c->pred_iters = fresh.pred_iters;
    c->pred_inner = fresh.pred_inner;
}
}  // namespace hf

extern "C" int hf_set_state(hf_ctx *c, const double *tp, const double *tpp) {
    if (!c || !tp || !tpp) return set_err(c, HF_ERR_ARG, "hf_set_state: null argument");
    // a new state: no carried residual, and the heuristics start as in a new context
    hf::reset_step_heuristics(c);
    c->pred_adj = hf_ctx{}.pred_adj;
    const size_t B8 = (size_t)c->g.N * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(c->T[0], tp, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(c->T[1], tpp, B8, hipMemcpyDefault, c->st));
    launch_step_begin(c->sc, c->st);   // clears stale per-step flags
    launch_tiny_from_state(c->g, c->T[0], c->T[1], c->sc, c->st);
    TRY(check_launch(c, "k_tiny_from_state"));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->has_state = true;
    return HF_OK;
}

extern "C" int hf_get_state(hf_ctx *c, double *tp, double *tpp) {
    if (!c) return HF_ERR_ARG;
    if (!c->has_state) return set_err(c, HF_ERR_STATE, "hf_get_state: no state");
    const size_t B8 = (size_t)c->g.N * sizeof(double);
    if (tp) HIPCHK(c, hipMemcpyAsync(tp, c->T[0], B8, hipMemcpyDefault, c->st));
    if (tpp) HIPCHK(c, hipMemcpyAsync(tpp, c->T[1], B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HF_OK;
}

static int flags_to_rc(hf_ctx *c, int32_t fl) {
    // the reference's guard order: B, Q, Qdens, S_CAT, S, x (mod_evolve.f90:70-158)
    if (fl & FLAG_NAN_B) return set_err(c, HF_ERR_NAN_B, "fatal error: NAN in B vector");
    if (fl & FLAG_NAN_Q) return set_err(c, HF_ERR_NAN_Q, "fatal error: NAN in Q vector");
    if (fl & FLAG_NAN_QDENS) return set_err(c, HF_ERR_NAN_QDENS, "fatal error: NAN in Qdens vector");
    if (fl & FLAG_NAN_SCAT) return set_err(c, HF_ERR_NAN_SCAT, "fatal error: NAN in S_CAT vector");
    if (fl & FLAG_NAN_S) return set_err(c, HF_ERR_NAN_S, "fatal error: NAN in S vector");
    if (fl & FLAG_NAN_X) return set_err(c, HF_ERR_NAN_X, "fatal error: NAN in x tempurature vector");
    return HF_OK;
}

// Timing slots of CG launches: iteration k of the solve, and whether it is the k_cg_b launch.
// Launches queued past convergence return at once and their slots are dropped; so is the
// first k_cg_b of a solve that stopped on its first iteration (Scal::shortcut).
struct SlotRec {
    long slot;
    int64_t k;
    bool is_b;
};

static void drop_idle_slots(hf_ctx *c, const std::vector<SlotRec> &slots, int64_t iters,
                            bool shortcut) {
    for (const SlotRec &sl : slots)
        if (sl.k >= iters || (sl.is_b && shortcut && sl.k == 0)) c->tpool[sl.slot].valid = false;
}

// Vectors of one CG solve on A^ y = rin.
template <typename T>
struct CgVecs {
    const T *rin;              // right-hand side, left unchanged (p_old of iteration 1)
    T *rw, *p[2], *q, *y;      // residual from iteration 1 on, directions, A^ p, solution
    const T *hx, *hy, *hz;     // scaled couplings
};

// linbcg's own stop rule for an FP64 solve (hf_solve): mode 1 (itol 1) stops on ||r|| / ||b||,
// mode 2 (itol 2) on ||r / diag H|| / ||b|| (src/heatflow/mod_Sparse.f90:418-425), both checked
// after every iteration in place of the scaled-residual test
struct WStop {
    int32_t mode;
    double tol;
};

// Queues CG iterations k0 .. k0+count-1 (launches past convergence return at once).
// Iteration 0: k_cg_a (first: p = rin, only p.q and q.q), k_fin_a (which may end the solve:
// y = alpha rin is then formed by the consumer), k_cg_b1 (q recomputed; y, rw), k_fin_b.
// Iteration k >= 1: k_cg_a (p = rw + beta p_old, p_old = rin for k = 1), k_cg_b on rw.
// With a WStop, each iteration ends with the weighted norm of its residual (rw) and the test.
template <typename T>
static void queue_cg(hf_ctx *c, Scal *sc, const CgVecs<T> &v, int64_t k0, int64_t count,
                     std::vector<SlotRec> *slots, int ida, int idb, const WStop *ws = nullptr) {
    const double N = (double)c->g.N, w = (double)sizeof(T);
    const Dims gc = cg_dims<T>(c->g);   // z-chunking of the stencil CG kernels
    for (int64_t k = k0; k < k0 + count; ++k) {
        long sa, sb;
        if (k == 0) {
            // k_cg_a (first): read r, 3 couplings
            sa = tbeg(c, ida, N * w * 4);
            launch_cg_a<T>(gc, true, v.rin, nullptr, nullptr, nullptr, nullptr, false, v.hx, v.hy,
                           v.hz, sc, c->partials, c->st);
            tend(c, sa);
            launch_fin_a(gc, c->partials, sc, c->st);
            // k_cg_b1: read r, 3 couplings; write rw
            sb = tbeg(c, idb, N * w * 5);
            launch_cg_b1<T>(gc, v.rin, v.hx, v.hy, v.hz, v.rw, sc, c->partials, c->st);
            tend(c, sb);
            launch_fin_b1(gc, c->partials, sc, c->st);
        } else {
            // p_k lives in p[(k - 1) & 1]; p_0 is the right-hand side
            const T *pold = k == 1 ? v.rin : v.p[k & 1];
            T *pnew = v.p[(k - 1) & 1];
            // k_cg_a: read r, p_old, (y), 3 couplings; write p, q, y (iteration k-1's term)
            sa = tbeg(c, ida, N * w * (k == 1 ? 8 : 9));
            launch_cg_a<T>(gc, false, v.rw, pold, pnew, v.q, v.y, k == 1, v.hx, v.hy, v.hz, sc,
                           c->partials, c->st);
            tend(c, sa);
            launch_fin_a(gc, c->partials, sc, c->st);
            // k_cg_b: read r, q; write r
            sb = tbeg(c, idb, N * w * 3);
            launch_cg_b<T>(c->g, v.rw, v.q, sc, c->partials, c->st);
            tend(c, sb);
            launch_fin_b(c->g, c->partials, sc, c->st);
        }
        if constexpr (std::is_same<T, double>::value) {
            if (ws)
                launch_wstop(c->g, v.rw, c->sv, ws->mode, k + 1, ws->tol, sc, c->partials,
                             c->st);
        }
        if (slots && sa >= 0) slots->push_back({sa, k, false});
        if (slots && sb >= 0) slots->push_back({sb, k, true});
    }
}

// CG on A^ y = rin until sc->done, in batches of `pred` iterations per host synchronisation.
template <typename T>
static int run_cg(hf_ctx *c, Scal *sc, Scal *h, const CgVecs<T> &v, int64_t *pred,
                  int64_t *iters, int ida, int idb, const WStop *ws = nullptr) {
    int64_t batch = std::max<int64_t>(1, *pred);
    int64_t queued = 0;
    std::vector<SlotRec> slots;
    for (;;) {
        queue_cg<T>(c, sc, v, queued, batch, &slots, ida, idb, ws);
        queued += batch;
        TRY(check_launch(c, "k_cg"));
        TRY(sync_scal(c, sc, h));
        if (h->done) break;
        batch = std::max<int64_t>(2, batch);
    }
    // the last iteration's term of y (iterations >= 1 leave it to their successor)
    if (h->ypend && h->yset) {
        launch_cg_tail<T>(c->g, v.y, v.p[0], v.p[1], sc, c->st);
        TRY(check_launch(c, "k_cg_tail"));
    }
    drop_idle_slots(c, slots, h->iter, h->shortcut != 0);
    if (h->shortcut) c->stats.first_iter_stops += 1;
    *iters = h->iter;
    // next batch: what this solve needed, +1 unless it stopped on its first iteration (a batch
    // cut short only costs one more host synchronisation)
    *pred = std::max<int64_t>(1, h->iter + (h->shortcut ? 0 : 1));
    return HF_OK;
}

static CgVecs<double> cg64(hf_ctx *c, const double *rin) {
    return CgVecs<double>{rin, c->rw64, {c->p64[0], c->p64[1]}, c->q64, c->y64, c->hx, c->hy, c->hz};
}
static CgVecs<float> cg32(hf_ctx *c, const float *rin) {
    return CgVecs<float>{rin, c->rw32, {c->p32[0], c->p32[1]}, c->q32, c->y32, c->fx, c->fy, c->fz};
}

// FP64 step with k_finalize fused into the next step's right-hand side (k_step64_z): per step
// one solve and one fused pass; the first step after a change of state starts from k_rhs_res.
// Results are bit-identical to step_fp64's unfused kernels.  Needs a per-step constant Qdens
// (no time-dependent heater modes or radiation: those recompute Qdens from T^n first).  The
// next step's residual depends on its TINY test, an OR over the grid that the pass itself
// decides: the pass forms it for the outcome of the last test (c->tiny_guess), and a wrong guess
// makes the next step start from k_rhs_res.
static int step_fp64_fused(hf_ctx *c, int64_t itime, int64_t *iter, double *err) {
    (void)itime;
    const hf_params &p = c->p;
    const double N = (double)c->g.N;
    RhsParams rp{1.0 / p.time_step, p.tol, p.itmax, p.isteady, p.icattaneo, rc_mask(c)};
    const double *rin;
    if (c->carry) {
        // step start and the scalars of the carried residual in one launch
        launch_carry_start(c->sc, nullptr, rp.tol, rp.itmax, allow_short(c), 0.0, 0, 0.0, false,
                           c->st);
        rin = c->carry_r64;
    } else {
        launch_step_begin(c->sc, c->st);
        // k_rhs_res: read Tp, Tpp, rho c_p, tau' rho c_p, Qdens, B, diag, 3 couplings; write r^
        long sr = tbeg(c, HF_K_RHS64, N * (8 * 9 + rc_bytes(c)));
        launch_rhs_res_state64(c->g, rp, c->T[0], c->T[1], c->rhoc, c->cvec, c->qd, c->bvec,
                               c->d, c->cx, c->cy, c->cz, c->r64, nullptr, c->sc, c->partials,
                               c->st);
        tend(c, sr);
        launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
        TRY(check_launch(c, "k_rhs_res"));
        rin = c->r64;
    }
    c->carry = false;
    int64_t it = 0;
    TRY(run_cg<double>(c, c->sc, c->h_sc, cg64(c, rin), &c->pred_iters, &it, HF_K_CG_A64,
                       HF_K_CG_B64));
    int32_t fl = c->h_sc->flags | c->setup_flags;
    if (fl & (FLAG_NAN_B | FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT | FLAG_NAN_S))
        return flags_to_rc(c, fl);
    const double e = c->h_sc->bnorm2 > 0 ? std::sqrt(c->h_sc->rr / c->h_sc->bnorm2) : 0.0;
    // the next step's right-hand side goes to the vector this solve did not read
    double *rn = rin == c->r64 ? c->r64b : c->r64;
    // k_step64: read Tp, Tpp, diag, y, rho c_p, tau' rho c_p, Qdens, B, 3 couplings; write x and
    // the next r^
    long sf = tbeg(c, HF_K_STEP64, N * (8 * 11 + rc_bytes(c)));
    launch_step64(c->g, rp, c->tiny_guess, c->T[0], c->T[1], c->y64, rin, c->rhoc, c->cvec,
                  c->qd, c->bvec, c->d, c->cx, c->cy, c->cz, c->T[2], rn, c->sc, c->partials,
                  c->st);
    tend(c, sf);
    launch_fin_next(c->g, c->partials, c->sc, c->tiny_guess, c->st);
    TRY(check_launch(c, "k_step64"));
    TRY(sync_scal(c, c->sc, c->h_sc));
    // rotation: Temp_pp <- Temp_p, Temp_p <- x (mod_evolve.f90:169-170); the old Temp_pp buffer
    // becomes the spare
    double *old_tpp = c->T[1];
    c->T[1] = c->T[0];
    c->T[0] = c->T[2];
    c->T[2] = old_tpp;
    // the next step starts from the residual the pass left if its TINY case was guessed
    const bool tiny_next = (c->h_sc->flags & FLAG_TINY_NEXT) != 0;
    c->carry = tiny_next == c->tiny_guess;
    if (!c->carry) c->stats.tiny_misses += 1;
    c->tiny_guess = tiny_next;
    c->carry_r64 = rn;
    c->stats.steps += 1;
    c->stats.outer_iters += it;
    c->stats.last_err = e;
    if (iter) *iter = it;
    if (err) *err = e;
    return HF_OK;
}

static int step_fp64(hf_ctx *c, int64_t itime, int64_t *iter, double *err) {
    if (!c->dyn_heat && !(c->debug & 1)) return step_fp64_fused(c, itime, iter, err);
    c->carry = false;
    const hf_params &p = c->p;
    launch_step_begin(c->sc, c->st);
    const double N = (double)c->g.N;
    if (c->dyn_heat) {
        HeatParams hp = heat_params(c, itime);
        long sq = tbeg(c, HF_K_QDENS, N * (1 + 8 * 3 + 8 + (c->qextra ? 8 : 0)));
        launch_qdens(c->g, hp, c->iheater, c->ells, c->tau_s, c->T[0], c->qextra, c->qd, c->sc,
                     c->st);
        c->qb_valid = false;
        tend(c, sq);
    }
    RhsParams rp{1.0 / p.time_step, p.tol, p.itmax, p.isteady, p.icattaneo, rc_mask(c)};
    // k_rhs_res: read Tp, Tpp, rho c_p, tau' rho c_p, Qdens, B, diag, 3 couplings; write r^
    long sr = tbeg(c, HF_K_RHS64, N * (8 * (9 + ((c->debug & 1) ? 1 : 0)) + rc_bytes(c)));
    launch_rhs_res_state64(c->g, rp, c->T[0], c->T[1], c->rhoc, c->cvec, c->qd, c->bvec, c->d,
                           c->cx, c->cy, c->cz, c->r64, (c->debug & 1) ? c->S : nullptr, c->sc,
                           c->partials, c->st);
    tend(c, sr);
    launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
    TRY(check_launch(c, "k_rhs_res"));
    int64_t it = 0;
    TRY(run_cg<double>(c, c->sc, c->h_sc, cg64(c, c->r64), &c->pred_iters, &it, HF_K_CG_A64,
                       HF_K_CG_B64));
    int32_t fl = c->h_sc->flags | c->setup_flags;
    if (fl & (FLAG_NAN_B | FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT | FLAG_NAN_S))
        return flags_to_rc(c, fl);
    const double e = c->h_sc->bnorm2 > 0 ? std::sqrt(c->h_sc->rr / c->h_sc->bnorm2) : 0.0;
    // k_finalize: read Tp, Tpp, s, y; write T^{n+1}
    long sf = tbeg(c, HF_K_FINAL, N * 8 * 5);
    launch_finalize(c->g, true, c->T[0], c->T[1], c->sv, c->y64, c->r64, c->sc, c->st);
    tend(c, sf);
    TRY(check_launch(c, "k_finalize"));
    std::swap(c->T[0], c->T[1]);  // Temp_pp <- Temp_p, Temp_p <- x (mod_evolve.f90:169-170)
    c->stats.steps += 1;
    c->stats.outer_iters += it;
    c->stats.last_err = e;
    if (iter) *iter = it;
    if (err) *err = e;
    return HF_OK;
}

// inner CG target floored at this fraction of the outer tolerance
constexpr double kInnerSafety = 0.25;
// steps without alpha reuse after a step whose reused alpha missed the inner target
constexpr int64_t kReusePause = 16;

// Queues the inner FP32 CG of one refinement sweep on A^ y = r (r unchanged); *batch_out is
// the number of iterations queued.  begin: set up the inner scalars (else already done).
static int inner_solve(hf_ctx *c, const float *r, int64_t *batch_out, std::vector<SlotRec> &slots,
                       bool begin = true) {
    const hf_params &p = c->p;
    if (begin)
        launch_inner_begin(c->sc, c->sci, p.inner_tol, p.tol, p.inner_itmax, kInnerSafety,
                           allow_short(c), c->st);
    const int64_t batch = std::max<int64_t>(1, c->pred_inner);
    queue_cg<float>(c, c->sci, cg32(c, r), 0, batch, &slots, HF_K_CG_A32, HF_K_CG_B32);
    // the last iteration's term of y, when the solve went past its first iteration (the
    // kernel returns at once otherwise)
    launch_cg_tail<float>(c->g, c->y32, c->p32[0], c->p32[1], c->sci, c->st);
    *batch_out = batch;
    return check_launch(c, "k_cg32");
}

// the first mixed-mode residual vector that is none of the given ones
static float *other_rvec(hf_ctx *c, const float *a, const float *b = nullptr) {
    float *all[3] = {c->r32, c->rn[0], c->rn[1]};
    for (float *v : all)
        if (v != a && v != b) return v;
    return nullptr;
}

// Mixed-precision step with the refinement residual fused into the next step's right-hand
// side (k_ir_step_z): per step one inner solve and one fused pass; the first step after a
// change of state starts from k_rhs_res.  Needs a per-step constant Qdens (no time-dependent
// heater modes or radiation: those recompute Qdens from T^n first).
// The next step's residual depends on its TINY test, an OR over the grid that the pass itself
// decides: the pass forms it for the outcome of the last test (c->tiny_guess), and a wrong
// guess makes the next step start from k_rhs_res.  The pass stores this step's residual only
// for a sweep that the last step followed with another (c->last_sweeps); when a sweep then
// needs one after all, the same pass runs again with the store.
// A carried step after one whose inner solve ended on its first iteration (y = alpha r^0) and
// needed no second sweep takes y = alpha r^0 with that alpha instead of solving: A^ is close to
// the identity here, so alpha changes little from step to step.  The sweep's FP64 residual
// must then meet half the inner solve's own target (in norm); if it does not, the step goes on
// with ordinary sweeps, and reuse pauses for kReusePause steps, twice as long after each
// further miss in a row.  A reusing sweep also yields its own step's alpha (k_fin_step), which
// the next step reuses: the lag is one step, so the small residual component along r^0 that a
// lagged alpha leaves does not build up over long runs of reuse.
// A carried step whose inner solve ends on its first iteration (y = alpha r^0, as a reused step
// length gives it) runs as the first step of a two-step pass with the next step, which reuses
// that alpha, when `pair_adv` is given (the caller has a next step to give): *pair_adv is then the
// number of steps taken (1 or 2), or 0 when the step ran alone.
static int pair_pass(hf_ctx *c, bool solved, int64_t *advanced, int64_t *iters);
static bool pair_grid_ok(const hf_ctx *c);
static int step_ir_fused(hf_ctx *c, int64_t itime, int64_t *iter, double *err,
                         int64_t *pair_adv = nullptr) {
    (void)itime;
    const hf_params &p = c->p;
    const double N = (double)c->g.N;
    RhsParams rp{1.0 / p.time_step, p.tol, p.itmax, p.isteady, p.icattaneo, rc_mask(c)};
    const float *rin;
    bool inner_started = false, reuse = false;
    if (pair_adv) *pair_adv = 0;
    const bool carried = c->carry;
    if (c->carry) {
        reuse = c->reuse_ok && c->reuse_pause == 0 && allow_short(c) && (c->debug & 4) == 0;
        // step, carried scalars and the first inner solve's scalars (or the reused alpha) in
        // one launch
        launch_carry_start(c->sc, c->sci, rp.tol, rp.itmax, allow_short(c), p.inner_tol,
                           p.inner_itmax, kInnerSafety, reuse, c->st);
        inner_started = true;
        rin = c->carry_r;
    } else {
        launch_step_begin(c->sc, c->st);
        // k_rhs_res: read 10 FP64 arrays; write r^ (FP32)
        long sr = tbeg(c, HF_K_RHS32, N * (8 * 8 + rc_bytes(c) + 4));
        launch_rhs_res_state32(c->g, rp, c->T[0], c->T[1], c->rhoc, c->cvec, c->qd, c->bvec,
                               c->d, c->cx, c->cy, c->cz, c->r32, nullptr, nullptr, c->sc,
                               c->partials, c->st);
        tend(c, sr);
        launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
        // Qdens + B, once per change of either (they change only through calls that end a
        // carried run)
        if (!c->qb_valid) {
            launch_sum2(c->g, c->qd, c->bvec, c->qb, c->st);
            c->qb_valid = true;
        }
        TRY(check_launch(c, "k_rhs_res"));
        rin = c->r32;
    }
    c->carry = false;
    StepParams sp{1.0 / p.time_step, p.isteady, p.icattaneo, c->tiny_guess ? 1 : 0, 0,
                  rc_mask(c)};
    const int64_t sweeps_max = std::min<int64_t>(p.itmax, 64);
    // refined x ping-pongs between the spare state buffer and S (unused by this path)
    double *xbuf[2] = {c->T[2], c->S};
    double *xa = nullptr;
    int nb = 0;
    int64_t inner_total = 0, sweeps = 0;
    bool short_stop = false;   // the last sweep's inner solve ended on its first iteration
    for (;;) {
        std::vector<SlotRec> slots;
        int64_t batch = 0;
        const bool reused = reuse && sweeps == 0;
        if (!reused) TRY(inner_solve(c, rin, &batch, slots, !inner_started));
        inner_started = false;
        if (pair_adv && carried && !reused && sweeps == 0 && c->reuse_pause == 0 &&
            pair_grid_ok(c)) {
            TRY(sync_scal(c, c->sci, c->h_sci, 1));   // did the solve end on its first iteration?
            if (c->h_sci->shortcut && c->h_sci->done) {
                int64_t adv = 0, it2 = 0;
                TRY(pair_pass(c, true, &adv, &it2));
                if (adv > 0) {
                    const int64_t in_it = c->h_sci->iter;
                    drop_idle_slots(c, slots, in_it, true);
                    c->stats.first_iter_stops += 1;
                    c->stats.inner_iters += in_it;
                    c->pred_inner = std::max<int64_t>(1, in_it + 1);
                    *pair_adv = adv;
                    if (iter) *iter = in_it + it2;
                    if (err) *err = c->stats.last_err;
                    return HF_OK;
                }
            }
        }
        double *xo = xbuf[nb];
        // the pass reads y (or the right-hand side, for y = alpha r) and writes two residual
        // vectors: neither may be the right-hand side
        float *rres = other_rvec(c, rin);
        float *rn = other_rvec(c, rin, rres);
        // k_ir_step: read T^n, T^{n-1} (or x), y or r (FP32), rho c_p, tau' rho c_p,
        // Qdens + B, diag, 3 couplings; write x', the next FP32 residual (and this step's)
        sp.store_res = sweeps + 1 < c->last_sweeps ? 1 : 0;
        auto pass = [&]() {
            long si = tbeg(c, HF_K_IR_STEP,
                           N * (8 * 7 + rc_bytes(c) + 4 + 8 + 4 + (sp.store_res ? 4 : 0)) +
                               (xa ? N * 8 : 0));
            launch_ir_step(c->g, sp, xa == nullptr, c->T[0], c->T[1], xa, c->y32, rin, c->sci,
                           c->rhoc, c->cvec, c->qb, c->d, c->cx, c->cy, c->cz, xo, rres, rn,
                           c->sc, c->partials, c->st);
            tend(c, si);
        };
        pass();
        launch_fin_step(c->g, c->partials, c->sc, c->sci, sweeps_max, c->tiny_guess, reused,
                        (c->debug & 256) != 0, c->st);
        TRY(check_launch(c, "k_ir_step"));
        TRY(sync_scal(c, c->sc, c->h_sc, 2));   // outer and inner scalars
        sweeps += 1;
        if (!reused) {
            const int64_t in_it = c->h_sci->iter;
            drop_idle_slots(c, slots, in_it, c->h_sci->shortcut != 0);
            short_stop = c->h_sci->shortcut != 0;
            if (short_stop) c->stats.first_iter_stops += 1;
            inner_total += in_it;
            // next batch: what this sweep needed + 1 (a batch cut short costs a refinement
            // sweep)
            c->pred_inner = c->h_sci->done ? std::max<int64_t>(1, in_it + 1) : batch * 2;
        }
        if (!c->h_sc->done && !sp.store_res) {
            // a further sweep needs this step's residual: the same pass again, with the store
            // (its other outputs, partials and flags come out the same; the scalars stay)
            sp.store_res = 1;
            pass();
            TRY(check_launch(c, "k_ir_step"));
            c->stats.residual_reruns += 1;
        }
        xa = xo;
        nb ^= 1;
        rin = rres;
        c->carry_r = rn;
        int32_t fl = c->h_sc->flags | c->setup_flags;
        if (fl & (FLAG_NAN_B | FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT | FLAG_NAN_S))
            return flags_to_rc(c, fl);
        if (c->h_sc->done) break;
    }
    // rotation: Temp_pp <- Temp_p, Temp_p <- x'; the two spares take the old Temp_pp buffer
    // and the unused refinement buffer
    double *old_tp = c->T[0], *old_tpp = c->T[1];
    double *other = xbuf[nb];   // the buffer the last sweep did not write
    c->T[0] = xa;
    c->T[1] = old_tp;
    c->T[2] = old_tpp;
    c->S = other;
    // the next step starts from the residual the last pass left if its TINY case was guessed
    const bool tiny_next = (c->h_sc->flags & FLAG_TINY_NEXT) != 0;
    c->carry = tiny_next == c->tiny_guess;
    if (!c->carry) c->stats.tiny_misses += 1;
    c->tiny_guess = tiny_next;
    c->last_sweeps = sweeps;
    // alpha reuse: after a step done in one sweep whose alpha was reused or came from a solve
    // that ended on its first iteration
    if (reuse) {
        c->stats.alpha_reuses += 1;
        if (sweeps > 1) {
            c->stats.reuse_misses += 1;
            c->reuse_pause = kReusePause << std::min<int64_t>(c->reuse_miss_run, 6);
            c->reuse_miss_run += 1;
        } else {
            c->reuse_miss_run = 0;
        }
    } else if (c->reuse_pause > 0) {
        c->reuse_pause -= 1;
    }
    c->reuse_ok = sweeps == 1 && (reuse || short_stop);
    const double e = c->h_sc->bnorm2 > 0 ? std::sqrt(c->h_sc->outer_rr / c->h_sc->bnorm2) : 0.0;
    c->stats.steps += 1;
    c->stats.outer_iters += sweeps;
    c->stats.inner_iters += inner_total;
    c->stats.last_err = e;
    if (iter) *iter = inner_total + sweeps;
    if (err) *err = e;
    return HF_OK;
}

// Whether the next two steps can run as one two-step pass (k_ir_step2_z): the fused mixed path
// is carrying a residual into a step that will reuse the inner step length after a one-sweep
// step, the grid has no periodic axis, the step is transient, and the grid has fewer than 2^29
// cells (the pass addresses its arrays with 32-bit byte offsets).  Debug bit 7 turns pairs off.
static bool pair_grid_ok(const hf_ctx *c) {
    const hf_params &p = c->p;
    return p.precision == HF_PREC_FP32_IR && !c->dyn_heat && (c->debug & (1 | 4 | 128)) == 0 &&
           allow_short(c) && c->last_sweeps == 1 && p.isteady == 0 && !c->g.per[0] &&
           !c->g.per[1] && !c->g.per[2] && c->qb_valid && (uint64_t)c->g.N < (1ull << 29);
}
static bool pair_ready(const hf_ctx *c) {
    return pair_grid_ok(c) && c->carry && c->reuse_ok && c->reuse_pause == 0;
}

// Two steps in one pass, the second reusing the first's inner step length.  The first reuses the
// last step's (`solved` false), or took its own from an inner solve that ended on its first
// iteration, already queued with the carried residual as its right-hand side (`solved`).
// *advanced: 2 when both stand, 1 when only the first does (the next step then starts from the
// state), 0 when neither does (the caller runs the step alone; with `solved`, the caller goes
// on with its single-step pass from the same solve, and nothing here has changed).  A step that
// fails its test after reusing alpha counts as a reuse that missed, as in the single-step path,
// and pauses reuse; a wrong TINY guess counts as a TINY miss.
static int pair_pass(hf_ctx *c, bool solved, int64_t *advanced, int64_t *iters) {
    const hf_params &p = c->p;
    const double N = (double)c->g.N;
    *advanced = 0;
    if (!solved)
        launch_carry_start(c->sc, c->sci, p.tol, p.itmax, allow_short(c), p.inner_tol,
                           p.inner_itmax, kInnerSafety, true, c->st);
    StepParams sp{1.0 / p.time_step, p.isteady, p.icattaneo, c->tiny_guess ? 1 : 0, 0,
                  rc_mask(c)};
    const float *rin = c->carry_r;
    float *rout = other_rvec(c, rin);
    double *x1 = c->T[2], *x2 = c->S;
    // k_ir_step2: read T^n, T^{n-1}, r^0 (FP32), Qdens + B, diag, 3 couplings (rho c_p, tau' rho
    // c_p per cell unless uniform); write x'_n, x'_{n+1}, r^0_{n+2} (FP32)
    long sl = tbeg(c, HF_K_IR_STEP2, N * (8 * 7 + 4 + rc_bytes(c) + 8 + 8 + 4));
    launch_ir_step2(c->g, sp, c->T[0], c->T[1], rin, c->sci, c->rhoc, c->cvec, c->qb, c->d,
                    c->cx, c->cy, c->cz, x1, x2, rout, c->sc, c->partials, c->st);
    tend(c, sl);
    launch_fin_step2(c->g, c->partials, c->sc, c->sci, c->tiny_guess, (c->debug & 256) != 0,
                     solved, (c->debug & 512) != 0, p.tol, p.inner_tol, kInnerSafety, c->st);
    TRY(check_launch(c, "k_ir_step2"));
    TRY(sync_scal(c, c->sc, c->h_sc, 2));
    const int32_t outcome = c->h_sc->pair_ok[0];
    if (solved && outcome == 0) return HF_OK;
    // steps that took y = alpha r^0 with a reused alpha, as the single-step path counts them:
    // step n unless `solved`, and step n+1 unless step n's TINY outcome differed from the guess
    // (step n+1 then starts from the state)
    const bool t1_guessed = ((c->h_sc->flags & FLAG_TINY_NEXT) != 0) == c->tiny_guess;
    c->stats.alpha_reuses +=
        (solved ? 0 : 1) + (outcome == 2 || (outcome == 1 && t1_guessed) ? 1 : 0);
    if (outcome == 0) {
        // step n runs again alone from the state (its test failed or a NaN guard tripped)
        c->carry = false;
        c->stats.reuse_misses += 1;
        c->reuse_pause = kReusePause << std::min<int64_t>(c->reuse_miss_run, 6);
        c->reuse_miss_run += 1;
        c->reuse_ok = false;
        return HF_OK;
    }
    double *old_tp = c->T[0], *old_tpp = c->T[1];
    if (outcome == 2) {
        // T^{n+2} = x'_{n+1}, T^{n+1} = x'_n; the old state buffers become the spares
        c->T[0] = x2;
        c->T[1] = x1;
        c->T[2] = old_tpp;
        c->S = old_tp;
        const bool t2 = c->h_sc->pair_ok[1] != 0;
        c->carry = t2 == c->tiny_guess;
        if (!c->carry) c->stats.tiny_misses += 1;
        c->tiny_guess = t2;
        c->carry_r = rout;
        c->reuse_miss_run = 0;
        c->reuse_ok = true;
    } else {
        // T^{n+1} = x'_n; step n+1 starts from the state
        c->T[0] = x1;
        c->T[1] = old_tp;
        c->T[2] = old_tpp;
        c->S = x2;
        const bool t1 = (c->h_sc->flags & FLAG_TINY_NEXT) != 0;
        if (t1 != c->tiny_guess) {
            c->stats.tiny_misses += 1;
            c->reuse_ok = true;
        } else {
            // step n+1 missed its test (or a NaN guard tripped: its own run reports it)
            c->stats.reuse_misses += 1;
            c->reuse_pause = kReusePause << std::min<int64_t>(c->reuse_miss_run, 6);
            c->reuse_miss_run += 1;
            c->reuse_ok = false;
        }
        c->carry = false;
        c->tiny_guess = t1;
    }
    c->last_sweeps = 1;
    const double e =
        c->h_sc->bnorm2 > 0 ? std::sqrt(c->h_sc->outer_rr / c->h_sc->bnorm2) : 0.0;
    c->stats.steps += outcome;
    c->stats.outer_iters += outcome;
    c->stats.last_err = e;
    *advanced = outcome;
    *iters = outcome;
    return HF_OK;
}

static int step_ir_pair(hf_ctx *c, int64_t *advanced, int64_t *iters) {
    return pair_pass(c, false, advanced, iters);
}

static int step_ir(hf_ctx *c, int64_t itime, int64_t *iter, double *err,
                  int64_t *pair_adv = nullptr) {
    if (!c->dyn_heat && !(c->debug & 1)) return step_ir_fused(c, itime, iter, err, pair_adv);
    c->carry = false;
    const hf_params &p = c->p;
    launch_step_begin(c->sc, c->st);
    const double N = (double)c->g.N;
    if (c->dyn_heat) {
        HeatParams hp = heat_params(c, itime);
        long sq = tbeg(c, HF_K_QDENS, N * (1 + 8 * 3 + 8 + (c->qextra ? 8 : 0)));
        launch_qdens(c->g, hp, c->iheater, c->ells, c->tau_s, c->T[0], c->qextra, c->qd, c->sc,
                     c->st);
        c->qb_valid = false;
        tend(c, sq);
    }
    RhsParams rp{1.0 / p.time_step, p.tol, p.itmax, p.isteady, p.icattaneo, rc_mask(c)};
    // x0 goes to T[2]; S is kept for the refinement residuals.
    // k_rhs_res: read 10 FP64 arrays; write r^ (FP32), S and x0 (FP64)
    long sr = tbeg(c, HF_K_RHS32, N * (8 * 8 + rc_bytes(c) + 4 + 8 + 8));
    launch_rhs_res_state32(c->g, rp, c->T[0], c->T[1], c->rhoc, c->cvec, c->qd, c->bvec, c->d,
                           c->cx, c->cy, c->cz, c->r32, c->S, c->T[2], c->sc, c->partials, c->st);
    tend(c, sr);
    launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
    TRY(check_launch(c, "k_rhs_res"));
    double *xa = c->T[2], *xb = c->T[1];  // T[1] (Temp_pp) is free once x0 and S exist
    // T[1] is still needed: the refined x of the last sweep must not overwrite Temp_pp before
    // the rotation, but Temp_pp is only read by k_rhs_res (done above), so it may host x.
    // refinement sweeps: each reduces the FP64 residual by the inner tolerance, so a handful
    // suffice; the cap guards tolerances below the FP64 round-off floor
    IrParams ip{std::min<int64_t>(p.itmax, 64), 1};
    int64_t inner_total = 0, sweeps = 0;
    float *rin = c->r32;
    for (;;) {
        std::vector<SlotRec> slots;
        int64_t batch = 0;
        TRY(inner_solve(c, rin, &batch, slots));
        float *rout = other_rvec(c, rin);   // the pass may read y as alpha * rin
        // k_ir_res: read x, y (FP32), s, S, diag, 3 couplings, Tp; write x, r^ (FP32)
        long si = tbeg(c, HF_K_IR_RES, N * (8 * 8 + 4 + 8 + 4));
        launch_ir_res(c->g, ip, xa, c->y32, rin, c->sci, c->sv, c->S, c->d, c->cx, c->cy, c->cz,
                      c->T[0], xb, rout, c->sc, c->partials, c->st);
        tend(c, si);
        launch_fin_ir(c->g, c->partials, c->sc, ip.itmax, c->st);
        TRY(check_launch(c, "k_ir_res"));
        std::swap(xa, xb);
        rin = rout;
        TRY(sync_scal(c, c->sc, c->h_sc, 2));   // outer and inner scalars
        const int64_t in_it = c->h_sci->iter;
        drop_idle_slots(c, slots, in_it, c->h_sci->shortcut != 0);
        if (c->h_sci->shortcut) c->stats.first_iter_stops += 1;
        inner_total += in_it;
        sweeps += 1;
        // next prediction: what this sweep needed (+1), or a longer batch if it was cut short
        // next batch: what this sweep needed + 1 (a batch cut short costs a refinement sweep)
        c->pred_inner = c->h_sci->done ? std::max<int64_t>(1, in_it + 1) : batch * 2;
        int32_t fl = c->h_sc->flags | c->setup_flags;
        if (fl & (FLAG_NAN_B | FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT | FLAG_NAN_S))
            return flags_to_rc(c, fl);
        if (c->h_sc->done) break;
    }
    // rotation: Temp_pp <- Temp_p, Temp_p <- x; the spare takes the old Temp_pp buffer
    double *old_tp = c->T[0], *old_tpp = c->T[1];
    double *spare = (xa == old_tpp) ? c->T[2] : old_tpp;
    c->T[0] = xa;
    c->T[1] = old_tp;
    c->T[2] = spare;
    const double e = c->h_sc->bnorm2 > 0 ? std::sqrt(c->h_sc->outer_rr / c->h_sc->bnorm2) : 0.0;
    c->stats.steps += 1;
    c->stats.outer_iters += sweeps;
    c->stats.inner_iters += inner_total;
    c->stats.last_err = e;
    if (iter) *iter = inner_total + sweeps;
    if (err) *err = e;
    return HF_OK;
}

// Steps of an FP64 context on a grid of at most kSmallMaxCells cells run in one persistent
// workgroup (k_run_small), unless debug bit 0 (keep S) or bit 6 asks for the multi-kernel path.
static bool small_path(const hf_ctx *c) {
    return c->p.precision == HF_PREC_FP64 && c->hx != nullptr && c->g.N <= c->small_max &&
           (c->debug & (1 | 64)) == 0;
}

// nsteps FP64 steps from itime0 on the persistent small-grid path, kSmallMaxSteps per launch
// and one host synchronisation per launch.  The scalars of the multi-kernel path are left as
// its next step expects them (TINY outcome in sc->flags, no carried residual).
static int run_small(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t *iters_total,
                     int64_t *last_iter, double *last_err) {
    if (!c->small_out) TRY(dalloc(c, &c->small_out, 1));
    if (c->dyn_heat && !c->small_hp) TRY(dalloc(c, &c->small_hp, kSmallMaxSteps));
    c->carry = false;
    const hf_params &p = c->p;
    const double N = (double)c->g.N;
    int64_t tot = 0, done = 0;
    while (done < nsteps) {
        const int32_t ns = (int32_t)std::min<int64_t>(nsteps - done, kSmallMaxSteps);
        if (c->dyn_heat) {
            c->small_hp_host.resize(ns);
            for (int32_t k = 0; k < ns; ++k) c->small_hp_host[k] = heat_params(c, itime0 + done + k);
            HIPCHK(c, hipMemcpyAsync(c->small_hp, c->small_hp_host.data(),
                                     ns * sizeof(HeatParams), hipMemcpyHostToDevice, c->st));
            c->qb_valid = false;
        }
        SmallRun a{};
        a.g = c->g;
        a.inverse_time = 1.0 / p.time_step;
        a.tol = p.tol;
        a.itmax = p.itmax;
        a.isteady = p.isteady;
        a.icattaneo = p.icattaneo;
        a.allow_short = allow_short(c) ? 1 : 0;
        a.nsteps = ns;
        a.dyn = c->dyn_heat ? 1 : 0;
        a.rc_mask = rc_mask(c);
        a.t[0] = c->T[0];
        a.t[1] = c->T[1];
        a.t[2] = c->T[2];
        a.rhoc = c->rhoc;
        a.cvec = c->cvec;
        a.bvec = c->bvec;
        a.d = c->d;
        a.cx = c->cx;
        a.cy = c->cy;
        a.cz = c->cz;
        a.hx = c->hx;
        a.hy = c->hy;
        a.hz = c->hz;
        a.sv = c->sv;
        a.qd = c->qd;
        a.hp = c->small_hp;
        a.iheater = c->iheater;
        a.ells = c->ells;
        a.tau_s = c->tau_s;
        a.qextra = c->qextra;
        a.r = c->r64;
        a.p0 = c->p64[0];
        a.p1 = c->p64[1];
        a.q = c->q64;
        a.y = c->y64;
        a.sc = c->sc;
        a.out = c->small_out;
        long slot = tbeg(c, HF_K_PERSIST, 0.0);
        launch_run_small(a, c->st);
        tend(c, slot);
        TRY(check_launch(c, "k_run_small"));
        SmallOut o{};
        HIPCHK(c, hipMemcpyAsync(&o, c->small_out, sizeof(SmallOut), hipMemcpyDeviceToHost,
                                 c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        double *t3[3] = {c->T[0], c->T[1], c->T[2]};
        c->T[0] = t3[o.rot % 3];
        c->T[1] = t3[(o.rot + 1) % 3];
        c->T[2] = t3[(o.rot + 2) % 3];
        // algorithmic bytes as the multi-kernel FP64 step counts them: right-hand side, 120 B a
        // CG iteration, x; plus Qdens when it is recomputed
        if (slot >= 0)
            c->tpool[slot].bytes = N * ((double)o.steps * (8 * 9 + rc_bytes(c) + 8 * 5 +
                                                           (c->dyn_heat ? 33.0 : 0.0)) +
                                        120.0 * (double)o.iters);
        c->tiny_guess = (o.flags & FLAG_TINY_NEXT) != 0;
        c->h_sc->flags = o.flags;
        c->stats.steps += o.steps;
        c->stats.outer_iters += o.iters;
        c->stats.first_iter_stops += o.first_stops;
        if (o.steps > 0) c->stats.last_err = o.last_err;
        tot += o.iters;
        done += o.steps;
        if (last_iter) *last_iter = o.last_iter;
        if (last_err) *last_err = o.last_err;
        const int32_t nan = o.flags & (FLAG_NAN_Q | FLAG_NAN_QDENS | FLAG_NAN_SCAT | FLAG_NAN_S |
                                       FLAG_NAN_X);
        if (nan) return flags_to_rc(c, nan);
        if (o.steps != ns)
            return set_err(c, HF_ERR_HIP, "k_run_small: the run ended early without a NaN flag");
    }
    if (iters_total) *iters_total = tot;
    return HF_OK;
}

static int step_common(hf_ctx *c, int64_t itime, int64_t *iter, double *err, bool final_sync,
                       int64_t *pair_adv = nullptr) {
    if (pair_adv) *pair_adv = 0;
    if (!c->assembled) return set_err(c, HF_ERR_STATE, "hf_step: not assembled");
    if (!c->has_state) return set_err(c, HF_ERR_STATE, "hf_step: no state (hf_set_state)");
    if (c->setup_flags) return flags_to_rc(c, c->setup_flags);
    if (small_path(c)) return run_small(c, itime, 1, nullptr, iter, err);
    int rc = c->p.precision == HF_PREC_FP32_IR ? step_ir(c, itime, iter, err, pair_adv)
                                               : step_fp64(c, itime, iter, err);
    if (rc != HF_OK) return rc;
    if (final_sync) {
        TRY(sync_scal(c, c->sc, c->h_sc));
        if (c->h_sc->flags & FLAG_NAN_X) return flags_to_rc(c, FLAG_NAN_X);
    }
    return HF_OK;
}

extern "C" int hf_step(hf_ctx *c, int64_t itime, int64_t *iter, double *err) {
    if (!c) return HF_ERR_ARG;
    return step_common(c, itime, iter, err, true);
}

extern "C" int hf_run(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t *iters_total) {
    if (!c) return HF_ERR_ARG;
    if (small_path(c) && nsteps > 0) {
        if (!c->assembled) return set_err(c, HF_ERR_STATE, "hf_run: not assembled");
        if (!c->has_state) return set_err(c, HF_ERR_STATE, "hf_run: no state (hf_set_state)");
        if (c->setup_flags) return flags_to_rc(c, c->setup_flags);
        return run_small(c, itime0, nsteps, iters_total, nullptr, nullptr);
    }
    int64_t tot = 0;
    for (int64_t k = 0; k < nsteps;) {
        int64_t it = 0;
        if (k + 1 < nsteps && c->assembled && c->has_state && !c->setup_flags &&
            pair_ready(c)) {
            int64_t adv = 0;
            int rc = step_ir_pair(c, &adv, &it);
            if (rc != HF_OK) return rc;
            if (adv > 0) {
                k += adv;
                tot += it;
                continue;
            }
        }
        // NaN in x of step k is seen at the next step's first synchronisation (the flag is
        // sticky within a step) or at the final check.  A step whose inner solve ends on its
        // first iteration may take the next step with it (pair_adv 2).
        int64_t pair_adv = 0;
        int rc = step_common(c, itime0 + k, &it, nullptr, k + 1 == nsteps,
                             k + 1 < nsteps ? &pair_adv : nullptr);
        if (rc != HF_OK) return rc;
        if (c->h_sc->flags & FLAG_NAN_X) return flags_to_rc(c, FLAG_NAN_X);
        tot += it;
        k += pair_adv > 0 ? pair_adv : 1;
    }
    if (iters_total) *iters_total = tot;
    return HF_OK;
}

extern "C" int hf_solve(hf_ctx *c, const double *b, double *x, int32_t itol, double tol,
                        int64_t itmax, int64_t *iter, double *err) {
    if (!c || !b || !x) return set_err(c, HF_ERR_ARG, "hf_solve: null argument");
    if (itol != 1 && itol != 2)
        return set_err(c, HF_ERR_ARG, "hf_solve: itol must be 1 or 2 (mod_Sparse.f90:418-423)");
    c->carry = false;   // device state or scratch changes: no carried residual
    if (!c->assembled) return set_err(c, HF_ERR_STATE, "hf_solve: not assembled");
    if (c->p.precision != HF_PREC_FP64 || !c->hx)
        return set_err(c, HF_ERR_STATE, "hf_solve: needs an FP64 context");
    const size_t N = c->g.N, B8 = N * sizeof(double);
    double *bd = nullptr, *xd = nullptr;
    TRY(dalloc(c, &c->S, N));
    bd = c->S;
    TRY(dalloc(c, &c->T[2], N));
    xd = c->T[2];
    HIPCHK(c, hipMemcpyAsync(bd, b, B8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipMemcpyAsync(xd, x, B8, hipMemcpyDefault, c->st));
    RhsParams rp{0.0, tol, itmax, 0, 0};
    launch_rhs_res_solve64(c->g, rp, xd, bd, c->d, c->cx, c->cy, c->cz, c->r64, c->sc,
                           c->partials, c->st);
    launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
    // linbcg's stop rule: ||b|| unscaled, then the itol norm of every iteration's residual
    const WStop ws{itol, tol};
    launch_wstop(c->g, bd, nullptr, 0, -1, tol, c->sc, c->partials, c->st);
    TRY(check_launch(c, "k_rhs_res(solve)"));
    int64_t it = 0, pred = std::max<int64_t>(4, c->pred_iters);
    TRY(run_cg<double>(c, c->sc, c->h_sc, cg64(c, c->r64),
                       &pred, &it, HF_K_CG_A64, HF_K_CG_B64, &ws));
    launch_finalize(c->g, false, nullptr, xd, c->sv, c->y64, c->r64, c->sc, c->st);
    TRY(check_launch(c, "k_finalize(solve)"));
    HIPCHK(c, hipMemcpyAsync(x, xd, B8, hipMemcpyDefault, c->st));
    TRY(sync_scal(c, c->sc, c->h_sc));
    if (iter) *iter = it;
    if (err) *err = c->h_sc->wb2 > 0 ? std::sqrt(c->h_sc->wrr / c->h_sc->wb2) : 0.0;
    return HF_OK;
}

extern "C" int hf_get_vector(hf_ctx *c, int32_t which, double *out) {
    if (!c || !out) return HF_ERR_ARG;
    const double *src = nullptr;
    switch (which) {
        case 0: src = c->bvec; break;
        case 1: src = c->qd; break;
        case 2: src = c->S; break;
        case 3: src = c->d; break;
        case 4: src = c->cx; break;
        case 5: src = c->cy; break;
        case 6: src = c->cz; break;
        case 7: src = c->kappa; break;
        case 8: src = c->rhoc; break;
        case 9: src = c->tau_s; break;
        case 10: src = c->ells; break;
        case 11: src = c->vol; break;
        default: return set_err(c, HF_ERR_ARG, "hf_get_vector: unknown vector id");
    }
    if (!src) return set_err(c, HF_ERR_STATE, "hf_get_vector: vector not available");
    if (which >= 7 ? !c->uploaded : !c->assembled)
        return set_err(c, HF_ERR_STATE, "hf_get_vector: not uploaded / assembled");
    HIPCHK(c, hipMemcpyAsync(out, src, (size_t)c->g.N * 8, hipMemcpyDefault, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HF_OK;
}

extern "C" int hf_set_debug(hf_ctx *c, int32_t flags) {
    if (!c) return HF_ERR_ARG;
    c->carry = false;   // device state or scratch changes: no carried residual
    c->debug = flags;
    if (flags & 1) TRY(dalloc(c, &c->S, c->g.N));
    return HF_OK;
}

extern "C" int hf_get_stats(hf_ctx *c, hf_stats *st) {
    if (!c || !st) return HF_ERR_ARG;
    *st = c->stats;
    return HF_OK;
}

extern "C" int hf_reset_stats(hf_ctx *c) {
    if (!c) return HF_ERR_ARG;
    c->stats = hf_stats{};
    return HF_OK;
}

extern "C" int hf_set_timing(hf_ctx *c, int32_t on) {
    if (!c) return HF_ERR_ARG;
    c->timing = on != 0;
    c->tused = 0;
    return HF_OK;
}

extern "C" int hf_kernel_times(hf_ctx *c, double *ms, int64_t *launches, double *bytes) {
    if (!c) return HF_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int k = 0; k < HF_K_COUNT; ++k) {
        if (ms) ms[k] = 0.0;
        if (launches) launches[k] = 0;
        if (bytes) bytes[k] = 0.0;
    }
    for (size_t i = 0; i < c->tused; ++i) {
        const hf_ctx::Timed &t = c->tpool[i];
        if (!t.valid || t.id < 0 || t.id >= HF_K_COUNT) continue;
        float e = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&e, t.a, t.b));
        if (ms) ms[t.id] += (double)e;
        if (launches) launches[t.id] += 1;
        if (bytes) bytes[t.id] += t.bytes;
    }
    c->tused = 0;
    return HF_OK;
}

extern "C" int hf_set_solver(hf_ctx *c, double tol, int64_t itmax, int32_t precision,
                             double inner_tol, int32_t inner_itmax) {
    if (!c) return HF_ERR_ARG;
    c->carry = false;   // device state or scratch changes: no carried residual
    if (precision != HF_PREC_FP64 && precision != HF_PREC_FP32_IR)
        return set_err(c, HF_ERR_ARG, "hf_set_solver: unknown precision");
    // the reverse sweep recomputes segments with the current solver: checkpoints of a forward
    // run under other controls would not match
    hf::adjoint_invalidate(c->adj);
    const int32_t old = c->p.precision;
    c->p.tol = tol;
    c->p.itmax = itmax;
    c->p.precision = precision;
    c->p.inner_tol = inner_tol;
    c->p.inner_itmax = inner_itmax;
    if (precision != old) {
        TRY(alloc_mode(c));
        if (c->assembled) {
            launch_scale(c->g, c->d, c->cx, c->cy, c->cz, c->sv, c->hx, c->hy, c->hz, c->fx,
                         c->fy, c->fz, c->st);
            TRY(check_launch(c, "k_scale"));
            HIPCHK(c, hipStreamSynchronize(c->st));
        }
    }
    return HF_OK;
}

// ---- internal entry points shared with the adjoint ------------------------------------------

namespace hf {
int ctx_step(hf_ctx *c, int64_t itime, bool final_sync) {
    return step_common(c, itime, nullptr, nullptr, final_sync);
}
int ctx_set_err(hf_ctx *c, int code, const std::string &msg) { return set_err(c, code, msg); }
int ctx_alloc(hf_ctx *c, double **ptr, size_t n) { return dalloc(c, ptr, n); }

int ctx_solve_dev(hf_ctx *c, const double *b, double *x, double tol, int64_t itmax,
                  int64_t *iters) {
    c->carry = false;   // the solve reuses the step scratch
    if (c->p.precision == HF_PREC_FP32_IR) {
        // FP32 inner CG inside FP64 refinement, x refined in place (ping-pong with T[2])
        RhsParams rp{0.0, tol, itmax, 0, 0};
        launch_rhs_res_solve32(c->g, rp, x, b, c->d, c->cx, c->cy, c->cz, c->r32, c->sc,
                               c->partials, c->st);
        launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
        TRY(check_launch(c, "k_rhs_res(solve32)"));
        IrParams ip{std::min<int64_t>(itmax, 64), 0};
        double *xa = x, *xb = c->T[2];
        float *rin = c->r32;
        int64_t inner_total = 0;
        for (;;) {
            launch_inner_begin(c->sc, c->sci, c->p.inner_tol, tol, c->p.inner_itmax, kInnerSafety,
                               allow_short(c), c->st);
            const int64_t batch = std::max<int64_t>(1, c->pred_adj);
            queue_cg<float>(c, c->sci, cg32(c, rin), 0, batch, nullptr, HF_K_CG_A32, HF_K_CG_B32);
            launch_cg_tail<float>(c->g, c->y32, c->p32[0], c->p32[1], c->sci, c->st);
            float *rout = other_rvec(c, rin);
            launch_ir_res(c->g, ip, xa, c->y32, rin, c->sci, c->sv, b, c->d, c->cx, c->cy, c->cz,
                          nullptr, xb, rout, c->sc, c->partials, c->st);
            launch_fin_ir(c->g, c->partials, c->sc, ip.itmax, c->st);
            TRY(check_launch(c, "k_ir_res(solve)"));
            std::swap(xa, xb);
            rin = rout;
            TRY(sync_scal(c, c->sc, c->h_sc, 2));   // outer and inner scalars
            inner_total += c->h_sci->iter;
            c->pred_adj = c->h_sci->done ? std::max<int64_t>(1, c->h_sci->iter + 1) : batch * 2;
            if (c->h_sc->done) break;
        }
        if (xa != x)
            HIPCHK(c, hipMemcpyAsync(x, xa, (size_t)c->g.N * 8, hipMemcpyDeviceToDevice, c->st));
        if (iters) *iters = inner_total;
        return HF_OK;
    }
    if (c->p.precision != HF_PREC_FP64 || !c->hx)
        return set_err(c, HF_ERR_STATE, "solve: no FP64 operator");
    RhsParams rp{0.0, tol, itmax, 0, 0};
    launch_rhs_res_solve64(c->g, rp, x, b, c->d, c->cx, c->cy, c->cz, c->r64, c->sc, c->partials,
                           c->st);
    launch_fin_start(c->g, c->partials, c->sc, rp.tol, rp.itmax, allow_short(c), c->st);
    TRY(check_launch(c, "k_rhs_res(solve)"));
    int64_t it = 0;
    TRY(run_cg<double>(c, c->sc, c->h_sc, cg64(c, c->r64),
                       &c->pred_adj, &it, HF_K_CG_A64, HF_K_CG_B64));
    launch_finalize(c->g, false, nullptr, x, c->sv, c->y64, c->r64, c->sc, c->st);
    TRY(check_launch(c, "k_finalize(solve)"));
    if (iters) *iters = it;
    return HF_OK;
}
}  // namespace hf

extern "C" int hf_forward_checkpointed(hf_ctx *c, int64_t itime0, int64_t nsteps,
                                       int64_t ckpt_every) {
    if (!c) return HF_ERR_ARG;
    return hf::adjoint_forward(c, itime0, nsteps, ckpt_every);
}

extern "C" int hf_adjoint(hf_ctx *c, const double *dl_dt_final, double *dl_dkappa,
                          double *dl_dpower) {
    if (!c) return HF_ERR_ARG;
    return hf::adjoint_backward(c, dl_dt_final, dl_dkappa, dl_dpower);
}

extern "C" int hf_set_fields(hf_ctx *c, const double *kappa, const double *qdens_extra) {
    if (!c) return HF_ERR_ARG;
    c->carry = false;   // device state or scratch changes: no carried residual
    if (!c->uploaded) return set_err(c, HF_ERR_STATE, "hf_set_fields: cells not uploaded");
    const size_t B8 = (size_t)c->g.N * 8;
    if (kappa) HIPCHK(c, hipMemcpyAsync(c->kappa, kappa, B8, hipMemcpyDefault, c->st));
    if (qdens_extra) {
        TRY(dalloc(c, &c->qextra, c->g.N));
        HIPCHK(c, hipMemcpyAsync(c->qextra, qdens_extra, B8, hipMemcpyDefault, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->assembled = false;
    hf::adjoint_invalidate(c->adj);
    return HF_OK;
}

extern "C" int hf_set_adjoint_tol(hf_ctx *c, double tol) {
    if (!c) return HF_ERR_ARG;
    if (!(tol >= 0.0)) return set_err(c, HF_ERR_ARG, "hf_set_adjoint_tol: tol must be >= 0");
    c->adj_tol = tol;
    return HF_OK;
}

extern "C" int hf_grad(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t ckpt_every,
                       int64_t nprobe, const int64_t *probe_idx, const double *target,
                       const double *weight, double *loss, double *dl_dkappa,
                       double *dl_dpower) {
    if (!c) return HF_ERR_ARG;
    return hf::adjoint_grad(c, itime0, nsteps, ckpt_every, nprobe, probe_idx, target, weight,
                            loss, dl_dkappa, dl_dpower);
}
