// hf_noalloc.hip -- the COO export for callers that cannot allocate device memory
// (hf_export_coo_noalloc; no reference counterpart).
//
// hf_export_coo (hf_api.hip) allocates its scan arrays and output staging on every call.  This
// variant writes straight into caller-owned device arrays, and the only scratch its kernels need
// -- one exclusive offset per chunk of kCooChunk cells and the total count -- lives in the
// file-scope device array `scratch` below, a fixed kScratchWords int64 words (2 MiB) that every
// GPU the library runs on holds until the process ends.  It covers grids of up to
// kMaxNoallocCells cells.
//
// Host code never passes `scratch` itself to a kernel or a copy (in host code the name is a
// host-side stand-in, not a device address): hf_create calls noalloc_setup, which takes the
// array's device address on the context's device with hipGetSymbolAddress once per device and
// keeps it in g_scratch[device]; the entry point passes that address to its kernels.
//
// Every call on one GPU shares that GPU's array: two calls that overlap in time on the same GPU
// (on different streams or host threads) race.  Each call rewrites every word it later reads
// (the chunk counts, then the offsets and the total), so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "hf_ctx.hpp"

namespace hf {

constexpr uint32_t kCooChunk = 4096;                 // cells per chunk: one block, 16 rounds
constexpr int64_t kScratchChunks = int64_t(1) << 18;  // chunk offsets the array holds
constexpr int64_t kScratchWords = kScratchChunks + 64;  // + the total (and padding)
constexpr int64_t kMaxNoallocCells = kScratchChunks * kCooChunk;   // 2^30 cells

static __device__ __align__(256) int64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
static int64_t *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Takes the device address of `scratch` on the current device (once per device).  Returns
// hipSuccess or the error of hipGetSymbolAddress.
hipError_t noalloc_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return hipSuccess;
    int64_t *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_scratch[dev] = ptr;
    return hipSuccess;
}

static int64_t *scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[dev];
}

// Entries sparse_Hmatrix emits for column j: the diagonal and two per emitted coupling.
__device__ __forceinline__ int64_t coo_count_of(const Dims &g, uint32_t j, const double *cx,
                                                const double *cy, const double *cz) {
    const Nbr n = neighbours(g, j);
    uint32_t rows[6];
    double vals[6];
    return 1 + 2 * coo_slots(g, n, j, cx, cy, cz, rows, vals);
}

// chunk b (cells [b C, (b+1) C)): its entry count -> cnt[b]
__global__ void __launch_bounds__(kBlock)
k_coo_chunk_count(Dims g, const double *__restrict__ cx, const double *__restrict__ cy,
                  const double *__restrict__ cz, int64_t *__restrict__ cnt) {
    __shared__ int64_t lds[kBlock];
    const uint64_t c0 = (uint64_t)blockIdx.x * kCooChunk;
    const uint64_t c1 = c0 + kCooChunk < g.N ? c0 + kCooChunk : g.N;
    int64_t s = 0;
    for (uint64_t j = c0 + threadIdx.x; j < c1; j += kBlock)
        s += coo_count_of(g, (uint32_t)j, cx, cy, cz);
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) lds[threadIdx.x] += lds[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[blockIdx.x] = lds[0];
}

// Exclusive scan of the nb chunk counts in place (one block of 1024 threads); the total goes
// to *total.
__global__ void __launch_bounds__(1024)
k_coo_chunk_scan(int64_t *__restrict__ cnt, int64_t nb, int64_t *__restrict__ total) {
    __shared__ int64_t lds[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int64_t s = 0;
    for (int64_t k = 0; k < per && b0 + k < nb; ++k) s += cnt[b0 + k];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = threadIdx.x >= (unsigned)off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = threadIdx.x == 0 ? 0 : lds[threadIdx.x - 1];
    for (int64_t k = 0; k < per && b0 + k < nb; ++k) {
        const int64_t v = cnt[b0 + k];
        cnt[b0 + k] = run;
        run += v;
    }
    if (threadIdx.x == 1023) *total = lds[1023];
}

// Chunk b writes its entries in sparse_Hmatrix's order from offset off[b]: rounds of kBlock
// consecutive cells, each cell's offset the chunk's running base plus a block-wide exclusive
// scan of the round's counts.  Entries at or past `len` are not written (the reference never
// writes past len).
__global__ void __launch_bounds__(kBlock)
k_coo_chunk_fill(Dims g, const double *__restrict__ d, const double *__restrict__ cx,
                 const double *__restrict__ cy, const double *__restrict__ cz,
                 const int64_t *__restrict__ off, int64_t len, double *__restrict__ val,
                 int64_t *__restrict__ irow, int64_t *__restrict__ jcol) {
    __shared__ int64_t lds[kBlock];
    const uint64_t c0 = (uint64_t)blockIdx.x * kCooChunk;
    const uint64_t c1 = c0 + kCooChunk < g.N ? c0 + kCooChunk : g.N;
    int64_t base = off[blockIdx.x];
    for (uint64_t r = c0; r < c1; r += kBlock) {
        const uint64_t jj = r + threadIdx.x;
        const bool mine = jj < c1;
        const uint32_t j = (uint32_t)jj;
        uint32_t rows[6];
        double vals[6];
        int k = 0;
        if (mine) k = coo_slots(g, neighbours(g, j), j, cx, cy, cz, rows, vals);
        const int64_t cnt = mine ? 1 + 2 * k : 0;
        // inclusive Hillis-Steele scan of the round's counts
        lds[threadIdx.x] = cnt;
        __syncthreads();
        for (int h = 1; h < kBlock; h <<= 1) {
            const int64_t v = threadIdx.x >= (unsigned)h ? lds[threadIdx.x - h] : 0;
            __syncthreads();
            lds[threadIdx.x] += v;
            __syncthreads();
        }
        const int64_t incl = lds[threadIdx.x];
        const int64_t round_total = lds[kBlock - 1];
        __syncthreads();   // lds is rewritten by the next round
        if (mine) {
            int64_t o = base + incl - cnt;
            auto put = [&](double v, int64_t i, int64_t jc) {
                if (o < len) {
                    val[o] = v;
                    irow[o] = i;
                    jcol[o] = jc;
                }
                ++o;
            };
            put(d[j], (int64_t)j + 1, (int64_t)j + 1);
            for (int s = 0; s < k; ++s) {
                put(vals[s], (int64_t)rows[s] + 1, (int64_t)j + 1);
                put(vals[s], (int64_t)j + 1, (int64_t)rows[s] + 1);
            }
        }
        base += round_total;
    }
}

// the sentinel (0, -2, -1) at positions [*count, len) (mod_setup.f90:104-106)
__global__ void __launch_bounds__(kBlock)
k_coo_sentinel(const int64_t *__restrict__ count, int64_t len, double *__restrict__ val,
               int64_t *__restrict__ irow, int64_t *__restrict__ jcol) {
    for (int64_t k = *count + (int64_t)blockIdx.x * kBlock + threadIdx.x; k < len;
         k += (int64_t)gridDim.x * kBlock) {
        val[k] = 0.0;
        irow[k] = -2;
        jcol[k] = -1;
    }
}

static bool on_device(const void *p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace hf

extern "C" int hf_export_coo_noalloc(hf_ctx *c, int64_t *n, int64_t *len, int64_t *count,
                                     double *val, int64_t *irow, int64_t *jcol) {
    using namespace hf;
    if (!c) return HF_ERR_ARG;
    if (!c->assembled) return ctx_set_err(c, HF_ERR_STATE, "hf_export_coo_noalloc: not assembled");
    const int64_t N = c->g.N;
    if (N > kMaxNoallocCells)
        return ctx_set_err(c, HF_ERR_ARG,
                           "hf_export_coo_noalloc: " + std::to_string(N) + " cells exceed the " +
                               std::to_string(kMaxNoallocCells) +
                               " its fixed scratch covers (use hf_export_coo)");
    const bool want = val || irow || jcol;
    if (want && !(val && irow && jcol))
        return ctx_set_err(c, HF_ERR_ARG, "hf_export_coo_noalloc: give all of val, irow, jcol or none");
    if (want && !(on_device(val) && on_device(irow) && on_device(jcol)))
        return ctx_set_err(c, HF_ERR_ARG,
                           "hf_export_coo_noalloc: val, irow and jcol must be device memory");
    int64_t *sp = scratch_of(c->dev);
    if (!sp) return ctx_set_err(c, HF_ERR_STATE, "hf_export_coo_noalloc: scratch not set up on this device");
    // the reference's len (mod_setup.f90:96-99)
    const hf_params &p = c->p;
    int64_t L = 7 * p.nx * p.ny * p.nz - 2 * (p.nx * p.ny + p.ny * p.nz + p.nz * p.nx);
    if (p.periodic[0]) L += 2 * p.ny * p.nz;
    if (p.periodic[1]) L += 2 * p.nz * p.nx;
    if (p.periodic[2]) L += 2 * p.nx * p.ny;
    const int64_t nb = (N + kCooChunk - 1) / kCooChunk;
    int64_t *chunk_off = sp, *total = sp + kScratchChunks;
    hipLaunchKernelGGL(k_coo_chunk_count, dim3((unsigned)nb), dim3(kBlock), 0, c->st, c->g, c->cx,
                       c->cy, c->cz, chunk_off);
    hipLaunchKernelGGL(k_coo_chunk_scan, dim3(1), dim3(1024), 0, c->st, chunk_off, nb, total);
    if (want) {
        hipLaunchKernelGGL(k_coo_chunk_fill, dim3((unsigned)nb), dim3(kBlock), 0, c->st, c->g,
                           c->d, c->cx, c->cy, c->cz, chunk_off, L, val, irow, jcol);
        hipLaunchKernelGGL(k_coo_sentinel, dim3(1024), dim3(kBlock), 0, c->st, total, L, val,
                           irow, jcol);
    }
    int64_t cnt_host = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt_host, total, sizeof(int64_t), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess)
        return ctx_set_err(c, HF_ERR_HIP, std::string("hf_export_coo_noalloc: ") + hipGetErrorString(e));
    if (n) *n = N;
    if (len) *len = L;
    if (count) *count = cnt_host;
    return HF_OK;
}
