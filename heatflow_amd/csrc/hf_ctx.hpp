// hf_ctx.hpp -- the opaque context behind hf_ctx* and the internal entry points shared by the
// forward (hf_api.hip) and adjoint (hf_adjoint.hip) translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/heatflow_hip.h"
#include "hf_device.hpp"
#include "hf_kernels.hpp"

namespace hf {
struct AdjointState;
void adjoint_free(AdjointState *a);
void adjoint_invalidate(AdjointState *a);
}  // namespace hf


using hf::Dims;
using hf::AsmParams;
using hf::Scal;

struct hf_ctx {
    hf_params p{};
    Dims g{};
    AsmParams ap{};
    hipStream_t st = nullptr;
    int dev = 0;
    std::string err;
    int debug = 0;

    // per-cell inputs and derived fields (device)
    double *kappa = nullptr, *rhoc = nullptr, *tau_s = nullptr, *cvec = nullptr, *ells = nullptr;
    double *lx = nullptr, *ly = nullptr, *lz = nullptr, *vol = nullptr, *qextra = nullptr;
    int8_t *iheater = nullptr;
    // assembled operator
    double *d = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr, *bvec = nullptr;
    double *sv = nullptr, *hx = nullptr, *hy = nullptr, *hz = nullptr;
    float *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double *qd = nullptr;        // Qdens (constant, or of the last step)
    // state: T[0] = Temp_p, T[1] = Temp_pp, T[2] = spare (the fused steps' x, the refined x)
    double *T[3] = {nullptr, nullptr, nullptr};
    double *S = nullptr;
    // FP64 CG (r64: right-hand side of a solve, rw64: its residual from iteration 1 on)
    double *r64 = nullptr, *rw64 = nullptr, *p64[2] = {nullptr, nullptr}, *q64 = nullptr,
           *y64 = nullptr;
    // the fused FP64 step writes the next step's right-hand side into the vector of r64, r64b
    // that the current solve does not read; carry_r64 is the one the next step starts from
    double *r64b = nullptr, *carry_r64 = nullptr;
    // FP32 inner CG
    float *r32 = nullptr, *rw32 = nullptr, *p32[2] = {nullptr, nullptr}, *q32 = nullptr,
          *y32 = nullptr;
    // mixed-mode residual vectors besides r32: a refinement pass writes its residuals into
    // vectors other than the right-hand side it read y from.  `carry` is set when the last
    // fused step left the next step's r^0 (and the scalars) valid for the current state;
    // `carry_r` is that r^0, formed for the TINY outcome `tiny_guess` (the last one seen)
    float *rn[2] = {nullptr, nullptr};
    double *qb = nullptr;        // Qdens + B for the fused step
    bool qb_valid = false;       // qb holds the current Qdens + B (cleared when either changes)
    bool carry = false;
    bool tiny_guess = true;
    int64_t last_sweeps = 1;     // refinement sweeps the last fused step needed
    // the next carried step may reuse the last inner solve's alpha (it ended on its first
    // iteration, or reused alpha itself, and the step needed one sweep); steps left before
    // reuse is tried again after a miss
    bool reuse_ok = false;
    int64_t reuse_pause = 0, reuse_miss_run = 0;
    float *carry_r = nullptr;
    // reductions / scalars
    Scal *sc = nullptr, *sci = nullptr;
    double *partials = nullptr;
    Scal *h_sc = nullptr;        // pinned mirror (coherent: k_publish writes it)
    Scal *h_sci = nullptr;
    int64_t *h_seq = nullptr;    // pinned: the sequence number of the last k_publish
    int64_t seq = 0;

    bool uploaded = false, assembled = false, has_state = false;
    bool any_heater = false, dyn_heat = false;
    bool any_em = false;         // some cell has a nonzero emissivity (with heaters: radiates)
    bool uniform_rc = false;     // rho c_p and tau' rho c_p are the same in every cell
    bool uniform_vol = false;    // the cell volume is the same in every cell
    bool uniform_len = false;    // each axis has one cell length (the adjoint's precondition)
    double heated_volume = 0.0;
    int32_t setup_flags = 0;
    // CG iterations queued before the first read-back of a solve (then the last solve's count
    // + 1, doubling while a batch falls short): launches queued past convergence return at once
    // but still cost a launch each
    int64_t pred_iters = 4, pred_inner = 4, pred_adj = 8;
    double adj_tol = 0.0;        // tolerance of the adjoint solves (0: p.tol)
    hf_stats stats{};
    // persistent small-grid steps (hf_persist.hip): run summary and per-step heat parameters
    uint32_t small_max = hf::kSmallMaxCells;   // HF_SMALL_MAX_CELLS overrides (tuning)
    hf::SmallOut *small_out = nullptr;
    hf::HeatParams *small_hp = nullptr;
    std::vector<hf::HeatParams> small_hp_host;
    std::vector<void *> allocs;
    hf::AdjointState *adj = nullptr;

    // kernel timing pool (hf_set_timing / hf_kernel_times)
    struct Timed {
        int id;
        hipEvent_t a, b;
        double bytes;
        bool valid;
    };
    bool timing = false;
    std::vector<Timed> tpool;
    size_t tused = 0;
};


namespace hf {
// Returns the step-to-step heuristics of both precision modes (carried residual, TINY guess,
// sweep and step-length-reuse history, CG batch predictions) to their values in a new context,
// so that steps from a state set by the caller run the same way whatever ran before (hf_api.hip)
void reset_step_heuristics(hf_ctx *c);
// one time step without the final host synchronisation (hf_api.hip)
int ctx_step(hf_ctx *c, int64_t itime, bool final_sync);
// FP64 solve of H x = b on device vectors; x holds the initial guess (hf_api.hip)
int ctx_solve_dev(hf_ctx *c, const double *b, double *x, double tol, int64_t itmax,
                  int64_t *iters);
int ctx_set_err(hf_ctx *c, int code, const std::string &msg);
int ctx_alloc(hf_ctx *c, double **ptr, size_t n);
// reverse-time gradient (hf_adjoint.hip)
int adjoint_forward(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t ckpt_every);
int adjoint_backward(hf_ctx *c, const double *seed, double *dl_dkappa, double *dl_dpower);
// takes the device address of the no-allocation export's static scratch array on the current
// device, once per device (hf_noalloc.hip)
hipError_t noalloc_setup(int dev);
int adjoint_grad(hf_ctx *c, int64_t itime0, int64_t nsteps, int64_t ckpt_every, int64_t nprobe,
                 const int64_t *probe_idx, const double *target, const double *weight,
                 double *loss, double *dl_dkappa, double *dl_dpower);
}  // namespace hf
