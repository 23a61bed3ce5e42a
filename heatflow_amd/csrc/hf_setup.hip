// hf_setup.hip -- setup on the device: the per-cell material fields from material ids
// (set_global_variables and material(), src/heatflow/mod_setup.f90:34-68,
// src/heatflow/mod_material.f90:54-267) and the stability check (mod_setup.f90:158-199).
//
// The material lookup runs per cell against a table of the file materials the caller read
// from mat.in and the built-in ones (index > 99) below.  material() assigns tau only for file
// materials and the built-ins 100, 320, 340, 350, and em only for file materials; elsewhere the
// value set_global_variables passes in stays, i.e. the value of the last earlier cell (in loop
// order) whose material assigned it.  That carry is an inclusive prefix maximum over cell
// indices ("the last cell at or before I that assigns tau"), computed per chunk of kSetupChunk
// cells, across chunks by one block, and within each chunk.  Cells before the first assigning
// cell get 0 (an uninitialised local in the reference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "hf_ctx.hpp"

namespace hf {

// Built-in materials of material() (mod_material.f90:116-265).  The literals there are default
// REAL (single precision) assigned to real(real12) variables, so the values are the
// single-precision roundings of the decimal constants.
struct BuiltinMaterial {
    int64_t index;
    float kappa, rho, heat_capacity, tau;
    int32_t sets_tau;
};
static const BuiltinMaterial kBuiltins[] = {
    {100, 1.0f, 1.0f, 1.0f, 0.0f, 1},          // air (:119-126)
    {140, 130.0f, 2328.0f, 4200.0f, 0.0f, 0},    // silicon (:138-144)
    {3001, 0.19f, 941.0f, 840.0f, 0.0f, 0},      // plastic (:148-153)
    {3002, 1.42f, 1700.0f, 1100.0f, 0.0f, 0},    // soil 1 (:156-161)
    {3003, 1.59f, 1800.0f, 1400.0f, 0.0f, 0},    // soil 2 (:164-169)
    {3004, 1.25f, 2300.0f, 2300.0f, 0.0f, 0},    // clay (:172-177)
    {9001, 0.6f, 1090.0f, 3421.0f, 0.0f, 0},     // muscle (:180-185)
    {9002, 0.209f, 1109.0f, 3391.0f, 0.0f, 0},   // epidermis (:189-194)
    {9003, 0.317f, 911.0f, 2348.0f, 0.0f, 0},    // fat (:198-203)
    {310, 130.0f, 2329.0f, 700.0f, 0.0f, 0},     // Si (:206-211)
    {320, 1.4f, 2400.0f, 720.0f, 1.0f, 1},       // SiO2 (:214-220)
    {330, 427.0f, 10524.0f, 236.0f, 0.0f, 0},    // silver (:223-228)
    {340, 0.2f, 1153.0f, 1500.0f, 1.0f, 1},      // SU8 (:231-237)
    {350, 10.2f, 7140.0f, 361.0f, 1.0f, 1},      // ITO (:240-246)
};
constexpr int kNumBuiltins = (int)(sizeof(kBuiltins) / sizeof(kBuiltins[0]));

// one entry of the device lookup table
struct MatEntry {
    int64_t index;
    double kappa, rho, heat_capacity, tau, em;
    int32_t sets_tau, sets_em;
};

constexpr uint32_t kSetupChunk = 4096;

// per cell: kappa, rho, c_p of its material; the table slot; this cell's index where its
// material assigns tau / em, else -1.  Unknown ids: *bad gets the smallest such cell.
__global__ void __launch_bounds__(kBlock)
k_mat_lookup(uint32_t N, const int64_t *__restrict__ ids, int32_t nmat,
             const MatEntry *__restrict__ tab, double *__restrict__ kappa,
             double *__restrict__ rho, double *__restrict__ hc, int32_t *__restrict__ slot,
             int32_t *__restrict__ src_tau, int32_t *__restrict__ src_em,
             unsigned long long *bad) {
    for (uint32_t I = blockIdx.x * kBlock + threadIdx.x; I < N; I += gridDim.x * kBlock) {
        const int64_t id = ids[I];
        int32_t k = -1;
        if (id > 0)
            for (int32_t m = 0; m < nmat; ++m)
                if (tab[m].index == id) {
                    k = m;
                    break;
                }
        if (k < 0) {
            atomicMin(bad, (unsigned long long)I);
            k = 0;
        }
        kappa[I] = tab[k].kappa;
        rho[I] = tab[k].rho;
        hc[I] = tab[k].heat_capacity;
        slot[I] = k;
        src_tau[I] = tab[k].sets_tau ? (int32_t)I : -1;
        src_em[I] = tab[k].sets_em ? (int32_t)I : -1;
    }
}

// chunk c: the largest source index in its cells (-1 if none), for tau and em
__global__ void __launch_bounds__(kBlock)
k_carry_chunk(uint32_t N, const int32_t *__restrict__ src_tau, const int32_t *__restrict__ src_em,
              int32_t *__restrict__ cmax) {
    __shared__ int32_t lt[kBlock], le[kBlock];
    const uint64_t c0 = (uint64_t)blockIdx.x * kSetupChunk;
    const uint64_t c1 = c0 + kSetupChunk < N ? c0 + kSetupChunk : N;
    int32_t mt = -1, me = -1;
    for (uint64_t i = c0 + threadIdx.x; i < c1; i += kBlock) {
        mt = max(mt, src_tau[i]);
        me = max(me, src_em[i]);
    }
    lt[threadIdx.x] = mt;
    le[threadIdx.x] = me;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            lt[threadIdx.x] = max(lt[threadIdx.x], lt[threadIdx.x + h]);
            le[threadIdx.x] = max(le[threadIdx.x], le[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cmax[2 * blockIdx.x] = lt[0];
        cmax[2 * blockIdx.x + 1] = le[0];
    }
}

// exclusive prefix maximum of the nb chunk maxima in place (one block of 1024 threads)
__global__ void __launch_bounds__(1024)
k_carry_scan(int32_t *__restrict__ cmax, int64_t nb) {
    __shared__ int32_t lt[1024], le[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int32_t mt = -1, me = -1;
    for (int64_t k = 0; k < per && b0 + k < nb; ++k) {
        mt = max(mt, cmax[2 * (b0 + k)]);
        me = max(me, cmax[2 * (b0 + k) + 1]);
    }
    lt[threadIdx.x] = mt;
    le[threadIdx.x] = me;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t vt = threadIdx.x >= (unsigned)off ? lt[threadIdx.x - off] : -1;
        const int32_t ve = threadIdx.x >= (unsigned)off ? le[threadIdx.x - off] : -1;
        __syncthreads();
        lt[threadIdx.x] = max(lt[threadIdx.x], vt);
        le[threadIdx.x] = max(le[threadIdx.x], ve);
        __syncthreads();
    }
    int32_t rt = threadIdx.x == 0 ? -1 : lt[threadIdx.x - 1];
    int32_t re = threadIdx.x == 0 ? -1 : le[threadIdx.x - 1];
    for (int64_t k = 0; k < per && b0 + k < nb; ++k) {
        const int32_t vt = cmax[2 * (b0 + k)], ve = cmax[2 * (b0 + k) + 1];
        cmax[2 * (b0 + k)] = rt;
        cmax[2 * (b0 + k) + 1] = re;
        rt = max(rt, vt);
        re = max(re, ve);
    }
}

// each cell's tau and em from the last cell at or before it whose material assigns them
// (rounds of kBlock cells, inclusive block scans carried from chunk to chunk); tau is stored
// scaled as set_global_variables stores grid%tau, tau * inverse_time * inverse_time
// (mod_setup.f90:45, 62)
__global__ void __launch_bounds__(kBlock)
k_carry_apply(uint32_t N, const int32_t *__restrict__ src_tau, const int32_t *__restrict__ src_em,
              const int32_t *__restrict__ cmax, const int32_t *__restrict__ slot,
              const MatEntry *__restrict__ tab, double inverse_time,
              double *__restrict__ tau_scaled, double *__restrict__ em) {
    __shared__ int32_t lt[kBlock], le[kBlock];
    const uint64_t c0 = (uint64_t)blockIdx.x * kSetupChunk;
    const uint64_t c1 = c0 + kSetupChunk < N ? c0 + kSetupChunk : N;
    int32_t ct = cmax[2 * blockIdx.x], ce = cmax[2 * blockIdx.x + 1];
    for (uint64_t r = c0; r < c1; r += kBlock) {
        const uint64_t i = r + threadIdx.x;
        lt[threadIdx.x] = i < c1 ? src_tau[i] : -1;
        le[threadIdx.x] = i < c1 ? src_em[i] : -1;
        __syncthreads();
        for (int off = 1; off < kBlock; off <<= 1) {
            const int32_t vt = threadIdx.x >= (unsigned)off ? lt[threadIdx.x - off] : -1;
            const int32_t ve = threadIdx.x >= (unsigned)off ? le[threadIdx.x - off] : -1;
            __syncthreads();
            lt[threadIdx.x] = max(lt[threadIdx.x], vt);
            le[threadIdx.x] = max(le[threadIdx.x], ve);
            __syncthreads();
        }
        const int32_t st = max(ct, lt[threadIdx.x]), se = max(ce, le[threadIdx.x]);
        if (i < c1) {
            const double t = st >= 0 ? tab[slot[st]].tau : 0.0;
            tau_scaled[i] = t * inverse_time * inverse_time;
            em[i] = se >= 0 ? tab[slot[se]].em : 0.0;
        }
        ct = max(ct, lt[kBlock - 1]);
        ce = max(ce, le[kBlock - 1]);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock)
k_fill3(uint32_t N, double lx, double ly, double lz, double *__restrict__ ax,
        double *__restrict__ ay, double *__restrict__ az, double *__restrict__ vol) {
    for (uint32_t I = blockIdx.x * kBlock + threadIdx.x; I < N; I += gridDim.x * kBlock) {
        ax[I] = lx;
        ay[I] = ly;
        az[I] = lz;
        vol[I] = lx * ly * lz;   // read_system, mod_inputs.f90:622
    }
}

// stability (mod_setup.f90:158-199): alpha = kappa / (rho c_p), var = dt alpha (1/L1^2 + 1/L2^2
// + 1/L3^2) against 1/12 (a single-precision constant there); cells on the x1, y1 or z1 face
// also with kappaBoundx1 in place of kappa.  *first gets the smallest failing cell index, and
// bit 0 / bit 1 of its flag word whether the cell or its boundary test failed.
__global__ void __launch_bounds__(kBlock)
k_stability(Dims g, double dt, double kb_x1, const double *__restrict__ kappa,
            const double *__restrict__ rhoc, const double *__restrict__ lx,
            const double *__restrict__ ly, const double *__restrict__ lz,
            unsigned long long *first) {
    const double thresh = (double)(1.0f / 12.0f);
    for (uint32_t I = blockIdx.x * kBlock + threadIdx.x; I < g.N; I += gridDim.x * kBlock) {
        const double sum = 1.0 / (lx[I] * lx[I]) + 1.0 / (ly[I] * ly[I]) + 1.0 / (lz[I] * lz[I]);
        const double alpha = kappa[I] / rhoc[I];
        bool bad = dt * alpha * sum > thresh;
        if (!bad) {
            const uint32_t t = g.dnx.div(I);
            const uint32_t x = I - t * g.nx, z = g.dny.div(t), y = t - z * g.ny;
            if (x == 0 || y == 0 || z == 0) bad = dt * (kb_x1 / rhoc[I]) * sum > thresh;
        }
        if (bad) atomicMin(first, (unsigned long long)I);
    }
}

static std::vector<MatEntry> lookup_table(int32_t nfile, const hf_material *file) {
    std::vector<MatEntry> t;
    for (int32_t i = 0; i < nfile; ++i) {
        const hf_material &m = file[i];
        t.push_back(MatEntry{m.index, m.kappa, m.rho, m.heat_capacity, m.tau, m.em, 1, 1});
    }
    for (const BuiltinMaterial &b : kBuiltins)
        t.push_back(MatEntry{b.index, (double)b.kappa, (double)b.rho, (double)b.heat_capacity,
                             (double)b.tau, 0.0, b.sets_tau, 0});
    return t;
}

}  // namespace hf

extern "C" int hf_builtin_material(int64_t index, hf_material *m, int32_t *sets_tau,
                                   int32_t *sets_em) {
    using namespace hf;
    for (const BuiltinMaterial &b : kBuiltins)
        if (b.index == index) {
            if (m) *m = hf_material{b.index, (double)b.kappa, (double)b.rho,
                                    (double)b.heat_capacity, (double)b.tau, 0.0};
            if (sets_tau) *sets_tau = b.sets_tau;
            if (sets_em) *sets_em = 0;
            return HF_OK;
        }
    return HF_ERR_ARG;
}

extern "C" int hf_upload_materials(hf_ctx *c, const int64_t *imaterial_type,
                                   const int64_t *iheater, int32_t nfile,
                                   const hf_material *file_materials, double len_x, double len_y,
                                   double len_z) {
    using namespace hf;
    if (!c) return HF_ERR_ARG;
    if (!imaterial_type || (nfile > 0 && !file_materials) || nfile < 0 || nfile > 99)
        return ctx_set_err(c, HF_ERR_ARG, "hf_upload_materials: bad material arguments");
    for (int32_t i = 0; i < nfile; ++i)
        if (file_materials[i].index < 1 || file_materials[i].index > 99)
            return ctx_set_err(c, HF_ERR_ARG,
                               "hf_upload_materials: file material indices are 1..99 "
                               "(mod_material.f90:86)");
    const uint32_t N = c->g.N;
    const size_t B8 = (size_t)N * 8;
    const std::vector<MatEntry> tab = lookup_table(nfile, file_materials);
    const int64_t nb = ((int64_t)N + kSetupChunk - 1) / kSetupChunk;
    // staging: 9 double and 3 int32 N-vectors, the id array, the table, the chunk carries
    std::vector<void *> bufs;
    auto get = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        bufs.push_back(p);
        return p;
    };
    auto release = [&]() {
        (void)hipStreamSynchronize(c->st);
        for (void *p : bufs) hipFree(p);
    };
    double *kap = (double *)get(B8), *rho = (double *)get(B8), *hc = (double *)get(B8),
           *taus = (double *)get(B8), *em = (double *)get(B8), *ax = (double *)get(B8),
           *ay = (double *)get(B8), *az = (double *)get(B8), *vol = (double *)get(B8);
    int32_t *slot = (int32_t *)get((size_t)N * 4), *st = (int32_t *)get((size_t)N * 4),
            *se = (int32_t *)get((size_t)N * 4), *cmax = (int32_t *)get((size_t)nb * 8);
    int64_t *ids = (int64_t *)get(B8);
    MatEntry *dtab = (MatEntry *)get(tab.size() * sizeof(MatEntry));
    unsigned long long *bad = (unsigned long long *)get(8);
    if (bufs.size() != 16 || std::find(bufs.begin(), bufs.end(), nullptr) != bufs.end()) {
        release();
        return ctx_set_err(c, HF_ERR_ALLOC, "hf_upload_materials: staging allocation failed");
    }
    const unsigned long long none = ~0ull;
    unsigned long long bad_h = none;
    hipError_t e = hipMemcpyAsync(ids, imaterial_type, B8, hipMemcpyDefault, c->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(MatEntry), hipMemcpyHostToDevice,
                           c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(bad, &none, 8, hipMemcpyHostToDevice, c->st);
    const unsigned nblk = (unsigned)std::min<int64_t>(((int64_t)N + kBlock - 1) / kBlock, 65536);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mat_lookup, dim3(nblk), dim3(kBlock), 0, c->st, N, ids,
                           (int32_t)tab.size(), dtab, kap, rho, hc, slot, st, se, bad);
        hipLaunchKernelGGL(k_carry_chunk, dim3((unsigned)nb), dim3(kBlock), 0, c->st, N, st, se,
                           cmax);
        hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(1024), 0, c->st, cmax, nb);
        hipLaunchKernelGGL(k_carry_apply, dim3((unsigned)nb), dim3(kBlock), 0, c->st, N, st, se,
                           cmax, slot, dtab, 1.0 / c->p.time_step, taus, em);
        hipLaunchKernelGGL(k_fill3, dim3(nblk), dim3(kBlock), 0, c->st, N, len_x, len_y, len_z,
                           ax, ay, az, vol);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad_h, bad, 8, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) {
        release();
        return ctx_set_err(c, HF_ERR_HIP, std::string("hf_upload_materials: ") + hipGetErrorString(e));
    }
    if (bad_h != none) {
        int64_t id = 0;
        (void)hipMemcpy(&id, ids + bad_h, 8, hipMemcpyDeviceToHost);
        release();
        return ctx_set_err(c, HF_ERR_ARG,
                           "hf_upload_materials: imaterial_type not recognized: " +
                               std::to_string(id) + " (cell " + std::to_string(bad_h) + ")");
    }
    const int rc = hf_upload_cells(c, kap, rho, hc, taus, em, ax, ay, az, vol, iheater, nullptr);
    release();
    return rc;
}

extern "C" int hf_check_stability(hf_ctx *c, int64_t *cell, double *value) {
    using namespace hf;
    if (!c) return HF_ERR_ARG;
    if (!c->uploaded) return ctx_set_err(c, HF_ERR_STATE, "hf_check_stability: cells not uploaded");
    unsigned long long *first = nullptr;
    if (hipMalloc(&first, 8) != hipSuccess) {
        (void)hipGetLastError();
        return ctx_set_err(c, HF_ERR_ALLOC, "hf_check_stability: allocation failed");
    }
    const unsigned long long none = ~0ull;
    unsigned long long h = none;
    const unsigned nblk = (unsigned)std::min<int64_t>(((int64_t)c->g.N + kBlock - 1) / kBlock, 65536);
    hipError_t e = hipMemcpyAsync(first, &none, 8, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_stability, dim3(nblk), dim3(kBlock), 0, c->st, c->g, c->p.time_step,
                           c->p.kappa_bound[0], c->kappa, c->rhoc, c->lx, c->ly, c->lz, first);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, first, 8, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    hipFree(first);
    if (e != hipSuccess)
        return ctx_set_err(c, HF_ERR_HIP, std::string("hf_check_stability: ") + hipGetErrorString(e));
    if (cell) *cell = h == none ? -1 : (int64_t)h;
    if (h == none) return HF_OK;
    // the failing cell's values of both tests, in the reference's order: the cell first
    double k = 0, rc = 0, ax = 0, ay = 0, az = 0;
    (void)hipMemcpy(&k, c->kappa + h, 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&rc, c->rhoc + h, 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&ax, c->lx + h, 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&ay, c->ly + h, 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&az, c->lz + h, 8, hipMemcpyDeviceToHost);
    const double sum = 1.0 / (ax * ax) + 1.0 / (ay * ay) + 1.0 / (az * az);
    const double dt = c->p.time_step;
    const double v_cell = dt * (k / rc) * sum, v_bnd = dt * (c->p.kappa_bound[0] / rc) * sum;
    const bool cell_fails = v_cell > (double)(1.0f / 12.0f);
    if (value) *value = cell_fails ? v_cell : v_bnd;
    return ctx_set_err(c, HF_ERR_UNSTABLE,
                       cell_fails ? "Stability condition not met (cell " + std::to_string(h) + ")"
                                  : "Stability condition at boundary not met (cell " +
                                        std::to_string(h) + ")");
}
