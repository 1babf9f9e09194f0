// pam_api.cpp -- the C ABI of libpam (include/pam.h): context, model and
// geometry set-up, and the forward / RTM / gradient time loops that drive the
// kernels of pam_kernels.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pam.h"
#include "pam_internal.h"

using namespace pam;

// the last failure of a context-free call on this thread (pam_last_error(NULL))
extern thread_local std::string g_free_err;

// The fields one set of shots is advanced with: the two wavefield buffers and
// the CPML memory variables ([max_shots][plane] each; psi twice, double-
// buffered across steps, psi_flip selecting the half that holds the previous
// step's psi).  The context holds the main set and, for checkpointed adjoint
// runs, a second set that recomputes source-field segments while the main set
// carries the adjoint field.
struct FieldSet {
    float *U0 = nullptr, *U1 = nullptr;
    float *psix = nullptr, *psiz = nullptr, *zetax = nullptr, *zetaz = nullptr;
    int psi_flip = 0;
};

struct pam_ctx {
    pam_grid g{};
    int phys = 0, abc = 0;
    pam_abc_params ap{};
    int nzp = 0, nxp = 0, P = 0;
    long long plane = 0;  // nzp * P
    hipStream_t st = nullptr;
    bool own_stream = false;
    int max_shots = 1;
    // the two lanes of pam_forward and group runs: streams created together
    // (so that they are served by different hardware queues; neither is the
    // caller's stream), forked from and joined back into st
    hipStream_t lane_st[2] = {nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};

    // model coefficient fields
    float *v2 = nullptr, *eps = nullptr, *del = nullptr, *cs = nullptr, *sn = nullptr;
    float *cerx = nullptr, *cerz = nullptr;
    float *cax = nullptr, *cbx = nullptr, *caz = nullptr, *cbz = nullptr;
    float *vp_pad = nullptr;
    float *th_pad = nullptr;  // tilt in radians, padded (TTI)
    bool model_set = false;
    // wavelet
    float *wav = nullptr;
    float wav_absmax = 0.f;
    bool wav_set = false;
    // geometry
    std::vector<int> hsz, hsx, hrz, hrx;
    int nshot = 0, nrec = 0, nslot = 0;
    int *d_rz = nullptr, *d_rx = nullptr, *d_rrow = nullptr, *d_rhead = nullptr, *d_rnext = nullptr;
    int *d_ssz = nullptr, *d_ssx = nullptr;  // [max_shots]
    bool geom_set = false;
    char errbuf[256] = {0};
    // set by k_nonfinite when a checked field holds NaN or Inf
    int *d_bad = nullptr;
    // wavefields and CPML memory: the main set and the recompute set
    FieldSet fs, aux;
    long long psi_half = 0;  // floats between the two halves of psi
    // per-step interface state (slot 0)
    float *cur = nullptr, *prv = nullptr;
    int next_k = -1;
    // per-step adjoint state (slot 0 of the second field set): the shot's data,
    // the adjoint fields, the next step (-2: none in progress, -1: the first
    // step may be any t), and the source field of the step in the kernels'
    // pitch (damped once more for the linearised adjoint)
    const float *adj_data = nullptr;
    float *adj_cur = nullptr, *adj_prv = nullptr;
    int adj_next_t = -2;
    float *adj_S = nullptr;
    long long adj_S_bytes = 0;
    // step-kernel variant chosen by the autotuner, per batch size
    std::map<int, int> tuned;
    std::map<int, double> tuned_ms;  // its time per shot-step
    std::map<int, int> tuned_lanes;  // variant per lane batch, timed on two lanes at once
    int auto_batch = 0;              // shots per launch chosen by the autotuner (0: not yet)
    int last_tile = -1;              // variant of the last whole run
    int last_batch = 0;              // shots per launch of the last whole run
    // options
    int batch = 0, tile = -1, graph = 1, storage = PAM_STORE_ALL_F32, ckpt_step = 0;
    int reset_adj_cpml = 1, normalize = 1, adj_range = 0, snapshots = 50, lanes = 2;
    long long group_bytes = 56LL << 20;  // wavefields of one sub-group of a group launch (0: no split)
    long long steps_run = 0;  // time steps launched by the current adjoint run
    // checkpoints of the main set ([segment][field][shot][plane]) and the
    // psi_flip at each
    float *ckpt = nullptr;
    long long ckpt_bytes = 0;
    std::vector<int> ckpt_flip;
    unsigned int *rvbits = nullptr;  // float16 slots: scale bits [slot][field][shot]
    long long rvbits_bytes = 0;
    unsigned int *smax = nullptr;  // per-snapshot max |S_k| (compressed storage)
    long long smax_bytes = 0;
    // the temporal-blocking variant's second field pair: U0 planes of every
    // shot slot, then U1 planes (slot offsets as in fs)
    float *tmp = nullptr;
    long long tmp_bytes = 0;
    void *save = nullptr;
    long long save_bytes = 0;
    float *save_scale = nullptr;
    float *acc_ilum = nullptr, *acc_img = nullptr;
    long long acc_bytes = 0;
    float *sbd = nullptr;  // boundary saving: D_t of the shots ([shot][nz][nx])
    long long sbd_bytes = 0;
    float *acc_multi = nullptr;  // per-shot eps / delta / theta partials
    long long acc_multi_bytes = 0;
    float *maxv = nullptr;
    float *ilum_tmp = nullptr;
    long long ilum_tmp_bytes = 0;
    double *d_acc = nullptr;
    // exact-adjoint scratch: saved forward states, adjoint work fields, gradient
    float *ex_save = nullptr, *ex_work = nullptr, *ex_gacc = nullptr;
    long long ex_save_bytes = 0, ex_work_bytes = 0, ex_gacc_bytes = 0;
    // timing
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<hipEvent_t> kev;
    double loop_ms = 0.0, kernel_ms = 0.0;
    long long launches = 0;
    int profile = 0;
    // instantiated step-loop graphs, keyed by what their launches bake in;
    // `epoch` changes whenever the model, wavelet, geometry or an option does
    std::map<std::string, hipGraphExec_t> graphs;
    unsigned epoch = 0;
    // graph-replayed forward runs gather into this fixed buffer ([max_shots][nt_data][nrec])
    // and copy it out, so that one graph serves every batch of a run
    float *stage = nullptr;
    long long stage_bytes = 0;
    std::string err;
};

static int fail(pam_ctx *c, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(pam_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HCK(call)                                                                                  \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return fail(ctx, PAM_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                        __LINE__);                                                                 \
    } while (0)
#define KCK(call)                                                                                  \
    do {                                                                                           \
        int _e = (call);                                                                           \
        if (_e != 0)                                                                               \
            return fail(ctx, PAM_E_HIP, "kernel launch %s failed: %s (%s:%d)", #call,              \
                        hipGetErrorString((hipError_t)_e), __FILE__, __LINE__);                    \
    } while (0)

static int dalloc(pam_ctx *ctx, void **p, size_t bytes) {
    if (*p) return PAM_OK;
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, PAM_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return PAM_OK;
}
#define DALLOC(p, bytes)                                                        \
    do {                                                                        \
        int _r = dalloc(ctx, (void **)&(p), (size_t)(bytes));                   \
        if (_r) return _r;                                                      \
    } while (0)

static void dfree(void *p) {
    if (p) (void)hipFree(p);
}

extern "C" const char *pam_version(void) { return "pam-mi355x 0.1 (gfx950)"; }
extern "C" int pam_tile_count(void) { return tile_count(); }

extern "C" int pam_create(const pam_grid *grid, int physics, int abc, const pam_abc_params *abcp, void *hip_stream,
                          int max_shots, pam_ctx **out) {
    if (!out) return PAM_E_ARG;
    *out = nullptr;
    pam_ctx *ctx = new pam_ctx();
    auto bad = [&](const char *msg) {
        ctx->err = msg;
        *out = ctx;  // returned so that pam_last_error can report; caller destroys
        return PAM_E_ARG;
    };
    if (!grid) return bad("grid is NULL");
    if (physics < PAM_ACOUSTIC || physics > PAM_TTI) return bad("physics must be 0 (acoustic), 1 (VTI) or 2 (TTI)");
    if (abc != PAM_CERJAN && abc != PAM_CPML) return bad("abc must be 0 (Cerjan) or 1 (CPML)");
    const pam_grid &g = *grid;
    if (g.nz < 1 || g.nx < 1) return bad("nz and nx must be positive");
    if (g.n_abc < 0) return bad("n_abc must be >= 0");
    if (abc == PAM_CPML && g.n_abc < 8) return bad("CPML needs n_abc >= 8");
    if (g.nz + 2 * g.n_abc < 9 || g.nx + 2 * g.n_abc < 9) return bad("padded grid must be at least 9x9");
    if (!(g.dx > 0.f) || !(g.dz > 0.f) || !(g.dt > 0.f)) return bad("dx, dz and dt must be positive");
    if (g.nt < 1 || g.itlag < 0 || g.itlag > g.nt) return bad("need nt >= 1 and 0 <= itlag <= nt");
    if (max_shots < 1) return bad("max_shots must be >= 1");
    {
        // the kernels address one shot's field (and one snapshot of it) with a
        // 32-bit byte offset from a per-shot base: a padded plane must stay
        // below 4 GiB
        const long long nzp = (long long)g.nz + 2LL * g.n_abc, nxp = (long long)g.nx + 2LL * g.n_abc;
        const long long plane_bytes = nzp * ((nxp + 63) / 64 * 64) * 4;
        if (plane_bytes >= (1LL << 32)) {
            std::snprintf(ctx->errbuf, sizeof ctx->errbuf,
                          "padded grid %lldx%lld: one field of %lld bytes reaches 4 GiB (the kernels' 32-bit "
                          "per-shot offsets); split the model",
                          nzp, nxp, plane_bytes);
            return bad(ctx->errbuf);
        }
    }
    ctx->g = g;
    ctx->phys = physics;
    ctx->abc = abc;
    if (abcp) ctx->ap = *abcp;
    ctx->nzp = g.nz + 2 * g.n_abc;
    ctx->nxp = g.nx + 2 * g.n_abc;
    ctx->P = (ctx->nxp + 63) / 64 * 64;
    ctx->plane = (long long)ctx->nzp * ctx->P;
    ctx->max_shots = max_shots;
    if (hip_stream) {
        ctx->st = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreate(&ctx->st) != hipSuccess) return bad("hipStreamCreate failed");
        ctx->own_stream = true;
    }
    *out = ctx;
    // the allocation-free entry points' scratch address on this device, looked
    // up here once so that their first call does no lookup (a failure is
    // reported by those calls themselves)
    {
        double *sp = nullptr;
        if (pam_noalloc_scratch(&sp) != PAM_OK) g_free_err.clear();
    }
    // wavefields and per-shot CPML memory
    const size_t fb = (size_t)ctx->plane * max_shots * sizeof(float);
    DALLOC(ctx->fs.U0, fb);
    DALLOC(ctx->fs.U1, fb);
    if (abc == PAM_CPML) {
        DALLOC(ctx->fs.psix, 2 * fb);
        DALLOC(ctx->fs.psiz, 2 * fb);
        ctx->psi_half = ctx->plane * max_shots;
        DALLOC(ctx->fs.zetax, fb);
        DALLOC(ctx->fs.zetaz, fb);
    }
    const size_t pb = (size_t)ctx->plane * sizeof(float);
    DALLOC(ctx->v2, pb);
    if (physics >= 1) {
        DALLOC(ctx->eps, pb);
        DALLOC(ctx->del, pb);
    }
    if (physics == 2) {
        DALLOC(ctx->cs, pb);
        DALLOC(ctx->sn, pb);
        DALLOC(ctx->th_pad, pb);
    }
    DALLOC(ctx->vp_pad, pb);
    DALLOC(ctx->cerx, (size_t)ctx->P * sizeof(float));
    DALLOC(ctx->cerz, (size_t)ctx->nzp * sizeof(float));
    if (abc == PAM_CPML) {
        DALLOC(ctx->cax, pb);
        DALLOC(ctx->cbx, pb);
        DALLOC(ctx->caz, pb);
        DALLOC(ctx->cbz, pb);
    }
    DALLOC(ctx->wav, (size_t)g.nt * sizeof(float));
    DALLOC(ctx->d_ssz, (size_t)max_shots * sizeof(int));
    DALLOC(ctx->d_ssx, (size_t)max_shots * sizeof(int));
    DALLOC(ctx->maxv, 64);
    // misfit: one float64 partial per block of the first reduction pass, then the sum
    DALLOC(ctx->d_acc, (size_t)(kMisfitParts + 1) * sizeof(double));
    DALLOC(ctx->d_bad, 64);
    HCK(hipMemsetAsync(ctx->fs.U0, 0, fb, ctx->st));
    HCK(hipMemsetAsync(ctx->fs.U1, 0, fb, ctx->st));
    // Cerjan factors: 1 in the interior, A[i] in the strips (cpp/Modeling.cu:111-120)
    {
        const int N = g.n_abc;
        std::vector<float> A(std::max(N, 1), 1.0f);
        for (int i = 0; i < N; ++i) {
            float fb_;
            if (ctx->ap.cerjan_factor > 0.f)
                fb_ = ctx->ap.cerjan_factor * (float)(N - i);
            else
                fb_ = (float)(N - i) / (1.4142f * (3.0f * (float)N));
            A[i] = expf(-fb_ * fb_);
        }
        std::vector<float> cx(ctx->P, 1.0f), cz(ctx->nzp, 1.0f);
        for (int i = 0; i < ctx->nxp; ++i) {
            float f = 1.0f;
            if (i < N) f *= A[i];
            if (i >= ctx->nxp - N) f *= A[ctx->nxp - 1 - i];
            cx[i] = f;
        }
        for (int i = 0; i < ctx->nzp; ++i) {
            float f = 1.0f;
            if (i < N) f *= A[i];
            if (i >= ctx->nzp - N) f *= A[ctx->nzp - 1 - i];
            cz[i] = f;
        }
        HCK(hipMemcpyAsync(ctx->cerx, cx.data(), cx.size() * sizeof(float), hipMemcpyHostToDevice, ctx->st));
        HCK(hipMemcpyAsync(ctx->cerz, cz.data(), cz.size() * sizeof(float), hipMemcpyHostToDevice, ctx->st));
        HCK(hipStreamSynchronize(ctx->st));
    }
    HCK(hipEventCreate(&ctx->e0));
    HCK(hipEventCreate(&ctx->e1));
    return PAM_OK;
}

extern "C" int pam_destroy(pam_ctx *ctx) {
    if (!ctx) return PAM_OK;
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    float *bufs[] = {ctx->v2,  ctx->eps,  ctx->del,   ctx->cs,    ctx->sn,   ctx->cerx, ctx->cerz,
                     ctx->cax, ctx->cbx,  ctx->caz,   ctx->cbz,   ctx->vp_pad, ctx->wav, ctx->fs.U0,
                     ctx->fs.U1,  ctx->fs.psix, ctx->fs.psiz,  ctx->fs.zetax, ctx->fs.zetaz, ctx->tmp, ctx->save_scale,
                     ctx->aux.U0, ctx->aux.U1, ctx->aux.psix, ctx->aux.psiz, ctx->aux.zetax, ctx->aux.zetaz, ctx->ckpt,
                     ctx->acc_ilum, ctx->maxv, ctx->ilum_tmp, ctx->th_pad, ctx->acc_multi,
                     ctx->sbd, ctx->stage, ctx->ex_save, ctx->ex_work, ctx->ex_gacc, ctx->adj_S};
    for (float *b : bufs) dfree(b);
    dfree(ctx->save);
    dfree(ctx->d_acc);
    dfree(ctx->d_bad);
    dfree(ctx->smax);
    dfree(ctx->rvbits);
    int *ibufs[] = {ctx->d_rz, ctx->d_rx, ctx->d_rrow, ctx->d_rhead, ctx->d_rnext, ctx->d_ssz, ctx->d_ssx};
    for (int *b : ibufs) dfree(b);
    for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto e : ctx->kev) (void)hipEventDestroy(e);
    if (ctx->e0) (void)hipEventDestroy(ctx->e0);
    if (ctx->e1) (void)hipEventDestroy(ctx->e1);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    for (int i = 0; i < 2; ++i) {
        if (ctx->join_ev[i]) (void)hipEventDestroy(ctx->join_ev[i]);
        if (ctx->lane_st[i]) (void)hipStreamDestroy(ctx->lane_st[i]);
    }
    if (ctx->own_stream && ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return PAM_OK;
}

// the last failure of a context-free call on this thread
thread_local std::string g_free_err;

extern "C" const char *pam_last_error(const pam_ctx *ctx) { return ctx ? ctx->err.c_str() : g_free_err.c_str(); }

extern "C" int pam_set_option(pam_ctx *ctx, int key, long long v) {
    if (!ctx) return PAM_E_ARG;
    ++ctx->epoch;
    switch (key) {
        case PAM_OPT_BATCH:
            if (v < 0 || v > ctx->max_shots) return fail(ctx, PAM_E_ARG, "batch must be in [0, max_shots]");
            ctx->auto_batch = 0;
            ctx->batch = (int)v;
            return PAM_OK;
        case PAM_OPT_TILE:
            if (v < -1 || v >= tile_count()) return fail(ctx, PAM_E_ARG, "tile index out of range");
            ctx->tile = (int)v;
            return PAM_OK;
        case PAM_OPT_GRAPH:
            ctx->graph = v ? 1 : 0;
            return PAM_OK;
        case PAM_OPT_STORAGE:
            if (v < PAM_STORE_ALL_F32 || v > PAM_STORE_RANDOM_BOUNDARY) return fail(ctx, PAM_E_ARG, "unknown storage");
            ctx->storage = (int)v;
            return PAM_OK;
        case PAM_OPT_CKPT_STEP:
            if (v < 0) return fail(ctx, PAM_E_ARG, "checkpoint step must be >= 0");
            ctx->ckpt_step = (int)v;
            return PAM_OK;
        case PAM_OPT_RESET_ADJ_CPML:
            ctx->reset_adj_cpml = v ? 1 : 0;
            return PAM_OK;
        case PAM_OPT_NORMALIZE:
            if (v < 0 || v > 2) return fail(ctx, PAM_E_ARG, "normalize must be 0, 1 or 2");
            ctx->normalize = (int)v;
            return PAM_OK;
        case PAM_OPT_ADJ_RANGE:
            if (v < 0 || v > 1) return fail(ctx, PAM_E_ARG, "adjoint range must be 0 or 1");
            ctx->adj_range = (int)v;
            return PAM_OK;
        case PAM_OPT_PROFILE:
            if (v < 0 || v > 2) return fail(ctx, PAM_E_ARG, "profile must be 0, 1 or 2");
            ctx->profile = (int)v;
            return PAM_OK;
        case PAM_OPT_SNAPSHOTS:
            if (v < 1 || v > 1000000) return fail(ctx, PAM_E_ARG, "snapshots must be in [1, 1e6]");
            ctx->snapshots = (int)v;
            return PAM_OK;
        case PAM_OPT_LANES:
            if (v < 1 || v > 2) return fail(ctx, PAM_E_ARG, "lanes must be 1 or 2");
            ctx->lanes = (int)v;
            return PAM_OK;
        case PAM_OPT_GROUP_BYTES:
            if (v < 0) return fail(ctx, PAM_E_ARG, "group bytes must be >= 0");
            ctx->group_bytes = v;
            return PAM_OK;
        default:
            return fail(ctx, PAM_E_ARG, "unknown option %d", key);
    }
}

extern "C" int pam_get_option(const pam_ctx *ctx, int key, long long *value) {
    if (!ctx || !value) return PAM_E_ARG;
    switch (key) {
        case PAM_OPT_BATCH: *value = ctx->batch > 0 ? ctx->batch : ctx->last_batch; return PAM_OK;
        case PAM_OPT_TILE: *value = ctx->tile >= 0 ? ctx->tile : ctx->last_tile; return PAM_OK;
        case PAM_OPT_GRAPH: *value = ctx->graph; return PAM_OK;
        case PAM_OPT_STORAGE: *value = ctx->storage; return PAM_OK;
        case PAM_OPT_CKPT_STEP: *value = ctx->ckpt_step; return PAM_OK;
        case PAM_OPT_RESET_ADJ_CPML: *value = ctx->reset_adj_cpml; return PAM_OK;
        case PAM_OPT_NORMALIZE: *value = ctx->normalize; return PAM_OK;
        case PAM_OPT_ADJ_RANGE: *value = ctx->adj_range; return PAM_OK;
        case PAM_OPT_PROFILE: *value = ctx->profile; return PAM_OK;
        case PAM_OPT_SNAPSHOTS: *value = ctx->snapshots; return PAM_OK;
        case PAM_OPT_LANES: *value = ctx->lanes; return PAM_OK;
        case PAM_OPT_GROUP_BYTES: *value = ctx->group_bytes; return PAM_OK;
        default: return PAM_E_ARG;
    }
}

extern "C" int pam_set_model(pam_ctx *ctx, const float *vp, const float *eps, const float *delta,
                             const float *theta, int layout) {
    if (!ctx) return PAM_E_ARG;
    ++ctx->epoch;
    if (!vp) return fail(ctx, PAM_E_ARG, "vp is NULL");
    if (ctx->phys >= 1 && (!eps || !delta)) return fail(ctx, PAM_E_ARG, "VTI/TTI need eps and delta");
    if (ctx->phys == 2 && !theta) return fail(ctx, PAM_E_ARG, "TTI needs theta");
    if (layout != PAM_INTERIOR && layout != PAM_PADDED) return fail(ctx, PAM_E_ARG, "layout must be 0 or 1");
    const int N = ctx->g.n_abc;
    auto put = [&](const float *src, float *dst) -> int {
        if (layout == PAM_INTERIOR) return launch_expand(src, ctx->g.nz, ctx->g.nx, N, dst, ctx->P, ctx->st);
        return launch_copy_padded(src, ctx->nzp, ctx->nxp, dst, ctx->P, ctx->st);
    };
    KCK(put(vp, ctx->vp_pad));
    float *th = nullptr;
    if (ctx->phys >= 1) {
        KCK(put(eps, ctx->eps));
        KCK(put(delta, ctx->del));
    }
    if (ctx->phys == 2) {
        KCK(put(theta, ctx->th_pad));
        th = ctx->th_pad;
    }
    KCK(launch_coeffs(ctx->phys, ctx->vp_pad, th, ctx->g.dt, ctx->nzp, ctx->nxp, ctx->P, ctx->v2, ctx->cs, ctx->sn,
                      ctx->st));
    if (ctx->abc == PAM_CPML)
        KCK(launch_cpml_coeffs(ctx->vp_pad, ctx->nzp, ctx->nxp, ctx->P, N, ctx->g.dx, ctx->g.dz, ctx->g.dt,
                               ctx->ap.cpml_d0, ctx->ap.cpml_f_pico, ctx->cax, ctx->cbx, ctx->caz, ctx->cbz, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    ctx->model_set = true;
    return PAM_OK;
}

extern "C" int pam_set_wavelet(pam_ctx *ctx, const float *src) {
    if (!ctx) return PAM_E_ARG;
    ++ctx->epoch;
    if (!src) return fail(ctx, PAM_E_ARG, "wavelet is NULL");
    HCK(hipMemcpyAsync(ctx->wav, src, (size_t)ctx->g.nt * sizeof(float), hipMemcpyDeviceToDevice, ctx->st));
    std::vector<float> h(ctx->g.nt);
    HCK(hipMemcpyAsync(h.data(), src, h.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    ctx->wav_absmax = 0.f;
    for (float v : h) ctx->wav_absmax = std::max(ctx->wav_absmax, std::fabs(v));
    ctx->wav_set = true;
    return PAM_OK;
}

extern "C" int pam_set_geometry(pam_ctx *ctx, const int *sz, const int *sx, int nshot, const int *rz, const int *rx,
                                int nrec) {
    if (!ctx) return PAM_E_ARG;
    ++ctx->epoch;
    if (nshot < 0 || nrec < 0) return fail(ctx, PAM_E_ARG, "negative counts");
    if ((nshot && (!sz || !sx)) || (nrec && (!rz || !rx))) return fail(ctx, PAM_E_ARG, "NULL geometry array");
    // sources and receivers must lie where the stencil updates the field: the
    // outer 4-cell rim of the padded grid is never updated
    for (int i = 0; i < nshot; ++i)
        if (sz[i] < 4 || sz[i] >= ctx->nzp - 4 || sx[i] < 4 || sx[i] >= ctx->nxp - 4)
            return fail(ctx, PAM_E_ARG, "source %d at (%d,%d) outside the stencil interior [4,%d)x[4,%d)", i, sz[i],
                        sx[i], ctx->nzp - 4, ctx->nxp - 4);
    for (int i = 0; i < nrec; ++i)
        if (rz[i] < 4 || rz[i] >= ctx->nzp - 4 || rx[i] < 4 || rx[i] >= ctx->nxp - 4)
            return fail(ctx, PAM_E_ARG, "receiver %d at (%d,%d) outside the stencil interior [4,%d)x[4,%d)", i,
                        rz[i], rx[i], ctx->nzp - 4, ctx->nxp - 4);
    ctx->hsz.assign(sz, sz + nshot);
    ctx->hsx.assign(sx, sx + nshot);
    ctx->hrz.assign(rz, rz + nrec);
    ctx->hrx.assign(rx, rx + nrec);
    ctx->nshot = nshot;
    ctx->nrec = nrec;
    // receiver tables: row -> slot, (slot, column) -> first receiver, receiver -> next at the same cell
    std::vector<int> rrow(ctx->nzp, -1);
    int nslot = 0;
    for (int r = 0; r < nrec; ++r)
        if (rrow[rz[r]] < 0) rrow[rz[r]] = nslot++;
    std::vector<int> head((size_t)std::max(nslot, 1) * ctx->P, -1), next(std::max(nrec, 1), -1);
    for (int r = nrec - 1; r >= 0; --r) {
        const size_t h = (size_t)rrow[rz[r]] * ctx->P + rx[r];
        next[r] = head[h];
        head[h] = r;
    }
    dfree(ctx->d_rz);
    dfree(ctx->d_rx);
    dfree(ctx->d_rrow);
    dfree(ctx->d_rhead);
    dfree(ctx->d_rnext);
    ctx->d_rz = ctx->d_rx = ctx->d_rrow = ctx->d_rhead = ctx->d_rnext = nullptr;
    DALLOC(ctx->d_rz, std::max(nrec, 1) * sizeof(int));
    DALLOC(ctx->d_rx, std::max(nrec, 1) * sizeof(int));
    DALLOC(ctx->d_rrow, rrow.size() * sizeof(int));
    DALLOC(ctx->d_rhead, head.size() * sizeof(int));
    DALLOC(ctx->d_rnext, next.size() * sizeof(int));
    if (nrec) {
        HCK(hipMemcpyAsync(ctx->d_rz, rz, nrec * sizeof(int), hipMemcpyHostToDevice, ctx->st));
        HCK(hipMemcpyAsync(ctx->d_rx, rx, nrec * sizeof(int), hipMemcpyHostToDevice, ctx->st));
    }
    HCK(hipMemcpyAsync(ctx->d_rrow, rrow.data(), rrow.size() * sizeof(int), hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(ctx->d_rhead, head.data(), head.size() * sizeof(int), hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(ctx->d_rnext, next.data(), next.size() * sizeof(int), hipMemcpyHostToDevice, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    ctx->nslot = nslot;
    ctx->geom_set = true;
    return PAM_OK;
}

// ------------------------------------------------------------------ helpers
// Blow-up detection (the reference only warns at set-up, cpp/Modeling.cu:321-326):
// a run clears the flag, checks the fields it ends with (NaN / Inf persist once
// an unstable step has produced them), and reports after its last
// synchronisation.
static int bad_clear(pam_ctx *ctx) {
    HCK(hipMemsetAsync(ctx->d_bad, 0, sizeof(int), ctx->st));
    return PAM_OK;
}
static int bad_check(pam_ctx *ctx, const float *x, long long n, hipStream_t st = nullptr) {
    KCK(launch_nonfinite(x, n, ctx->d_bad, st ? st : ctx->st));
    return PAM_OK;
}
static int bad_report(pam_ctx *ctx, const char *what) {
    int h = 0;
    HCK(hipMemcpyAsync(&h, ctx->d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    if (h)
        return fail(ctx, PAM_E_NONFINITE,
                    "%s: the wavefield holds NaN or Inf (an unstable run: dt above the stability limit of the "
                    "model, or a non-finite input)",
                    what);
    return PAM_OK;
}

static int ready(pam_ctx *ctx) {
    if (!ctx->model_set) return fail(ctx, PAM_E_STATE, "no model set (pam_set_model)");
    if (!ctx->wav_set) return fail(ctx, PAM_E_STATE, "no wavelet set (pam_set_wavelet)");
    if (!ctx->geom_set) return fail(ctx, PAM_E_STATE, "no geometry set (pam_set_geometry)");
    return PAM_OK;
}

static void base_step_args(pam_ctx *ctx, StepArgs &a, int nshots) {
    std::memset(&a, 0, sizeof(a));
    const pam_grid &g = ctx->g;
    a.nzp = ctx->nzp;
    a.nxp = ctx->nxp;
    a.P = ctx->P;
    a.N = g.n_abc;
    a.nz = g.nz;
    a.nx = g.nx;
    a.nshots = nshots;
    a.sstride = ctx->plane;
    a.idx2 = 1.0f / (g.dx * g.dx);
    a.idz2 = 1.0f / (g.dz * g.dz);
    a.idx = 1.0f / g.dx;
    a.idz = 1.0f / g.dz;
    a.idxdz = 1.0f / (g.dz * g.dx);
    a.idt2 = 1.0f / (g.dt * g.dt);
    a.dxdz = g.dx * g.dz;
    a.nt = g.nt;
    a.itlag = g.itlag;
    a.nrec = ctx->nrec;
    a.v2 = ctx->v2;
    a.eps = ctx->eps;
    a.del = ctx->del;
    a.cs = ctx->cs;
    a.sn = ctx->sn;
    a.cerx = ctx->cerx;
    a.cerz = ctx->cerz;
    a.psix = ctx->fs.psix;
    a.psiz = ctx->fs.psiz;
    a.zetax = ctx->fs.zetax;
    a.zetaz = ctx->fs.zetaz;
    a.cax = ctx->cax;
    a.cbx = ctx->cbx;
    a.caz = ctx->caz;
    a.cbz = ctx->cbz;
    a.ssz = ctx->d_ssz;
    a.ssx = ctx->d_ssx;
    a.wav = ctx->wav;
    a.rrow = ctx->d_rrow;
    a.rhead = ctx->d_rhead;
    a.rnext = ctx->d_rnext;
    a.flags = (ctx->abc == PAM_CERJAN) ? F_CERJAN : 0;
}

static void base_psi_args(pam_ctx *ctx, PsiArgs &p, int nshots, const float *cur) {
    p.nzp = ctx->nzp;
    p.nxp = ctx->nxp;
    p.P = ctx->P;
    p.N = ctx->g.n_abc;
    p.nshots = nshots;
    p.sstride = ctx->plane;
    p.idx = 1.0f / ctx->g.dx;
    p.idz = 1.0f / ctx->g.dz;
    p.cur = cur;
    p.psix = ctx->fs.psix;
    p.psiz = ctx->fs.psiz;
    p.cax = ctx->cax;
    p.cbx = ctx->cbx;
    p.caz = ctx->caz;
    p.cbz = ctx->cbz;
}

// (on stream st, or the context's when null)
static int zero_fields(pam_ctx *ctx, FieldSet &f, int nshots, bool fields, bool cpml, hipStream_t st = nullptr) {
    if (!st) st = ctx->st;
    const size_t fb = (size_t)ctx->plane * nshots * sizeof(float);
    if (fields) {
        HCK(hipMemsetAsync(f.U0, 0, fb, st));
        HCK(hipMemsetAsync(f.U1, 0, fb, st));
    }
    if (cpml && ctx->abc == PAM_CPML) {
        HCK(hipMemsetAsync(f.psix, 0, fb, st));
        HCK(hipMemsetAsync(f.psiz, 0, fb, st));
        HCK(hipMemsetAsync(f.psix + ctx->psi_half, 0, fb, st));
        HCK(hipMemsetAsync(f.psiz + ctx->psi_half, 0, fb, st));
        f.psi_flip = 0;
        HCK(hipMemsetAsync(f.zetax, 0, fb, st));
        HCK(hipMemsetAsync(f.zetaz, 0, fb, st));
    }
    return PAM_OK;
}
static int zero_shot_state(pam_ctx *ctx, int nshots, bool fields, bool cpml) {
    return zero_fields(ctx, ctx->fs, nshots, fields, cpml);
}

// source cells of shots shot0 .. shot0+nshots-1 into slots slot0.. of d_ssz / d_ssx
static int upload_sources(pam_ctx *ctx, int shot0, int nshots, int slot0 = 0, hipStream_t st = nullptr) {
    if (!st) st = ctx->st;
    for (int s = 0; s < nshots; ++s)
        if (shot0 + s < 0 || shot0 + s >= ctx->nshot)
            return fail(ctx, PAM_E_ARG, "shot %d out of range [0,%d)", shot0 + s, ctx->nshot);
    HCK(hipMemcpyAsync(ctx->d_ssz + slot0, ctx->hsz.data() + shot0, nshots * sizeof(int), hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(ctx->d_ssx + slot0, ctx->hsx.data() + shot0, nshots * sizeof(int), hipMemcpyHostToDevice, st));
    return PAM_OK;
}

// One forward time step of `nshots` shots on the field set f: psi pre-pass
// (CPML, unless the kernel variant updates psi itself) then the fused step,
// on stream st (the context's when null).
static int fwd_step_fs(pam_ctx *ctx, FieldSet &f, StepArgs &a, int tile, float *cur, float *prv, int k,
                       hipEvent_t *ev, hipStream_t st = nullptr) {
    if (!st) st = ctx->st;
    if (ctx->abc == PAM_CPML) {
        const long long in = f.psi_flip ? ctx->psi_half : 0, out = f.psi_flip ? 0 : ctx->psi_half;
        a.psix_in = f.psix + in;
        a.psiz_in = f.psiz + in;
        a.psix = f.psix + out;
        a.psiz = f.psiz + out;
        a.zetax = f.zetax;
        a.zetaz = f.zetaz;
        if (!tile_fuses_psi(tile)) {
            PsiArgs p;
            base_psi_args(ctx, p, a.nshots, cur);
            p.psix_in = a.psix_in;
            p.psiz_in = a.psiz_in;
            p.psix = a.psix;
            p.psiz = a.psiz;
            KCK(launch_psi(p, st));
        }
    }
    a.cur = cur;
    a.prv = prv;
    a.k = k;
    KCK(launch_step(ctx->phys, ctx->abc == PAM_CPML, tile, a, st, ev));  // ev: the kernel's own start / end
    ++ctx->steps_run;
    if (ctx->abc == PAM_CPML) f.psi_flip ^= 1;
    return PAM_OK;
}
static int fwd_step(pam_ctx *ctx, StepArgs &a, int tile, float *cur, float *prv, int k, hipEvent_t *ev) {
    return fwd_step_fs(ctx, ctx->fs, a, tile, cur, prv, k, ev);
}

// The temporal-blocking variant: nlaunch launches of tblock_steps steps from
// step k0 on the shots of f (a view of the context's field set), alternating
// between f's pair (cur, prv) and the matching slots of the second pair; with
// nlaunch even the state ends in (cur, prv) with cur holding the last level,
// as nlaunch * tblock_steps single steps (an even number) would leave it.
// (the second pair is allocated by ensure_tblock beforehand: never during a
// graph capture)
static int ensure_scratch(pam_ctx *ctx, void **p, long long *have, long long need);
static int ensure_tblock(pam_ctx *ctx) {
    return ensure_scratch(ctx, (void **)&ctx->tmp, &ctx->tmp_bytes, 2LL * ctx->max_shots * ctx->plane * 4);
}
static int tblock_launches(pam_ctx *ctx, const FieldSet &f, const StepArgs &a, hipStream_t st, float *cur,
                           float *prv, int k0, int nlaunch) {
    if (nlaunch & 1) return fail(ctx, PAM_E_STATE, "temporal blocking needs an even number of launches");
    if (!ctx->tmp || ctx->tmp_bytes < 2LL * ctx->max_shots * ctx->plane * 4)
        return fail(ctx, PAM_E_STATE, "temporal blocking without its second field pair");
    const long long slot = f.U0 - ctx->fs.U0;  // floats from the set's first shot slot
    if (slot < 0 || slot + (long long)a.nshots * ctx->plane > (long long)ctx->max_shots * ctx->plane)
        return fail(ctx, PAM_E_STATE, "temporal blocking outside the context's field set");
    float *x0 = ctx->tmp + slot, *x1 = ctx->tmp + (long long)ctx->max_shots * ctx->plane + slot;
    const int T = tblock_steps(ctx->phys);
    for (int j = 0; j < nlaunch; ++j) {
        const int k = k0 + j * T;
        if ((j & 1) == 0)
            KCK(launch_tblock(ctx->phys, a, cur, prv, x0, x1, k, T, st));
        else
            KCK(launch_tblock(ctx->phys, a, x0, x1, cur, prv, k, T, st));
        ctx->steps_run += T;
    }
    return PAM_OK;
}
static bool tblock_usable(const pam_ctx *ctx, int tile, int flags) {
    return tile_is_tblock(tile) && ctx->abc == PAM_CERJAN && (flags & ~F_FWD_MASK) == 0;
}

// The autotuner: times every step-kernel variant on the wavefields of `nb`
// shots (a few plain steps each, HIP events on the context's stream) and keeps
// the fastest for that batch size.  The fields it steps on are cleared before
// and after, so it must run before a shot is primed.
static constexpr int kDefaultTile = 5;
static constexpr int kTuneRounds = 3;

static int autotune(pam_ctx *ctx, int nb, int *best, double *best_per_shot_step) {
    *best = kDefaultTile;
    *best_per_shot_step = 1e300;
    StepArgs a;
    base_step_args(ctx, a, nb);
    int r = zero_shot_state(ctx, nb, true, true);
    if (r) return r;
    double best_ms = 1e300;  // per step
    for (int t = 0; t < tile_count(); ++t) {
        float *c = ctx->fs.U0, *p = ctx->fs.U1;
        const bool tb = tile_is_tblock(t);
        if (tb) {
            if (!tblock_usable(ctx, t, a.flags)) continue;
            r = ensure_tblock(ctx);
            if (r) return r;
        }
        // warm-up steps, then the best of kTuneRounds timed windows (variants
        // within a few percent of each other were picked by chance from one)
        const int nsteps = tb ? 2 * tblock_steps(ctx->phys) : 6;
        auto steps = [&](int k0, int n) -> int {
            if (tb) return tblock_launches(ctx, ctx->fs, a, ctx->st, c, p, k0, n / tblock_steps(ctx->phys));
            for (int i = 0; i < n; ++i) {
                int rr = fwd_step(ctx, a, t, c, p, k0 + i, nullptr);
                if (rr) return rr;
                std::swap(c, p);
            }
            return PAM_OK;
        };
        r = steps(0, tb ? nsteps : 2);
        if (r) return r;
        double t_ms = 1e300;
        for (int round = 0; round < kTuneRounds; ++round) {
            HCK(hipEventRecord(ctx->e0, ctx->st));
            r = steps(nsteps * (round + 1), nsteps);
            if (r) return r;
            HCK(hipEventRecord(ctx->e1, ctx->st));
            HCK(hipEventSynchronize(ctx->e1));
            float ms = 0.f;
            HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
            t_ms = std::min(t_ms, (double)ms / nsteps);
        }
        if (t_ms < best_ms) {
            best_ms = t_ms;
            *best = t;
        }
    }
    *best_per_shot_step = best_ms / nb;
    return zero_shot_state(ctx, nb, true, true);
}

// variant for a batch of nshots: the option if set, else the autotuner's
// choice (tuned on first use when `may_tune`, else the default)
static int choose_tile(pam_ctx *ctx, int nshots, bool may_tune) {
    if (ctx->tile >= 0) return ctx->tile;
    auto it = ctx->tuned.find(nshots);
    if (it != ctx->tuned.end()) return it->second;
    if (!may_tune) return kDefaultTile;
    int best = kDefaultTile;
    double per = 1e300;
    if (autotune(ctx, nshots, &best, &per) != PAM_OK) {
        (void)hipGetLastError();
        best = kDefaultTile;
    }
    ctx->tuned[nshots] = best;
    ctx->tuned_ms[nshots] = per;
    return best;
}

// Shots advanced per launch when PAM_OPT_BATCH is 0: the autotuner times the
// batch sizes max_shots, max_shots/2, ... down to 8 (each with its best
// kernel variant) and keeps the one with the shortest time per shot-step.
// Small batches keep the working set in the 256 MB MALL, large ones
// amortise launches and model reads; the best size depends on the grid.
static int auto_batch(pam_ctx *ctx) {
    if (ctx->auto_batch > 0) return ctx->auto_batch;
    int best = ctx->max_shots;
    double best_per = 1e300;
    for (int nb = ctx->max_shots; nb >= 1; nb /= 2) {
        (void)choose_tile(ctx, nb, true);
        const double per = ctx->tuned_ms.count(nb) ? ctx->tuned_ms[nb] : 1e300;
        if (per < best_per) {
            best_per = per;
            best = nb;
        }
        if (nb <= 8) break;
    }
    ctx->auto_batch = best;
    return best;
}

static int ensure_events(pam_ctx *ctx, size_t n) {
    while (ctx->kev.size() < n) {
        hipEvent_t e;
        HCK(hipEventCreate(&e));
        ctx->kev.push_back(e);
    }
    return PAM_OK;
}

// ------------------------------------------------------------ per-step API
extern "C" int pam_reset(pam_ctx *ctx, int shot) {
    if (!ctx) return PAM_E_ARG;
    int r = ready(ctx);
    if (r) return r;
    if (shot < 0 || shot >= ctx->nshot) return fail(ctx, PAM_E_ARG, "shot %d out of range", shot);
    (void)choose_tile(ctx, 1, true);
    r = zero_shot_state(ctx, 1, true, true);
    if (r) return r;
    r = upload_sources(ctx, shot, 1);
    if (r) return r;
    KCK(launch_inject_point(ctx->fs.U0, ctx->plane, ctx->d_ssz, ctx->d_ssx, ctx->P, ctx->wav, 0, 1, ctx->st));
    ctx->cur = ctx->fs.U0;
    ctx->prv = ctx->fs.U1;
    ctx->next_k = 0;
    return PAM_OK;
}

extern "C" int pam_forward_step(pam_ctx *ctx, int k, float *seis) {
    if (!ctx) return PAM_E_ARG;
    if (ctx->next_k < 0) return fail(ctx, PAM_E_STATE, "pam_reset must precede pam_forward_step");
    if (k != ctx->next_k) return fail(ctx, PAM_E_STATE, "steps must be issued in order: expected k=%d", ctx->next_k);
    if (k >= ctx->g.nt) return fail(ctx, PAM_E_ARG, "k=%d beyond nt=%d", k, ctx->g.nt);
    StepArgs a;
    base_step_args(ctx, a, 1);
    a.flags |= F_SRC;
    if (seis && k >= ctx->g.itlag) {
        a.flags |= F_GATHER;
        a.seis = seis;
        a.dstride = (long long)(ctx->g.nt - ctx->g.itlag) * ctx->nrec;
    }
    int r = fwd_step(ctx, a, choose_tile(ctx, 1, false), ctx->cur, ctx->prv, k, nullptr);
    if (r) return r;
    std::swap(ctx->cur, ctx->prv);
    ctx->next_k = k + 1;
    return PAM_OK;
}

extern "C" int pam_get_field(pam_ctx *ctx, int which, float *out) {
    if (!ctx || !out) return PAM_E_ARG;
    if (which < 0 || which > 3) return fail(ctx, PAM_E_ARG, "which must be 0 .. 3");
    if (which < 2 && ctx->next_k < 0) return fail(ctx, PAM_E_STATE, "no shot in progress");
    if (which >= 2 && ctx->adj_next_t < -1) return fail(ctx, PAM_E_STATE, "no adjoint shot in progress");
    const float *src = which == 0 ? ctx->cur : which == 1 ? ctx->prv : which == 2 ? ctx->adj_cur : ctx->adj_prv;
    KCK(launch_pack_padded(src, ctx->nzp, ctx->nxp, ctx->P, out, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return PAM_OK;
}

// ------------------------------------------------------------ per-step adjoint
static int ensure_scratch(pam_ctx *ctx, void **p, long long *have, long long need);
static int alloc_fieldset(pam_ctx *ctx, FieldSet &f);

extern "C" int pam_adjoint_reset(pam_ctx *ctx, int shot, const float *data) {
    if (!ctx) return PAM_E_ARG;
    int r = ready(ctx);
    if (r) return r;
    if (shot < 0 || shot >= ctx->nshot) return fail(ctx, PAM_E_ARG, "shot %d out of range", shot);
    if (!data) return fail(ctx, PAM_E_ARG, "data is NULL");
    r = alloc_fieldset(ctx, ctx->aux);
    if (r) return r;
    r = ensure_scratch(ctx, (void **)&ctx->adj_S, &ctx->adj_S_bytes, ctx->plane * 4);
    if (r) return r;
    // the adjoint field cleared; the CPML memory cleared, or with
    // PAM_OPT_RESET_ADJ_CPML 0 taken over from the forward per-step state as
    // the reference's driver leaves it (src/Migration.py:1000-1039)
    const bool keep = ctx->abc == PAM_CPML && !ctx->reset_adj_cpml;
    r = zero_fields(ctx, ctx->aux, 1, true, !keep);
    if (r) return r;
    if (keep) {
        if (ctx->next_k < 0) return fail(ctx, PAM_E_STATE, "PAM_OPT_RESET_ADJ_CPML 0 needs a forward per-step run");
        const size_t fb = (size_t)ctx->plane * 4;
        for (long long h : {0LL, ctx->psi_half}) {
            HCK(hipMemcpyAsync(ctx->aux.psix + h, ctx->fs.psix + h, fb, hipMemcpyDeviceToDevice, ctx->st));
            HCK(hipMemcpyAsync(ctx->aux.psiz + h, ctx->fs.psiz + h, fb, hipMemcpyDeviceToDevice, ctx->st));
        }
        HCK(hipMemcpyAsync(ctx->aux.zetax, ctx->fs.zetax, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(ctx->aux.zetaz, ctx->fs.zetaz, fb, hipMemcpyDeviceToDevice, ctx->st));
        ctx->aux.psi_flip = ctx->fs.psi_flip;
    }
    ctx->adj_data = data;
    ctx->adj_cur = ctx->aux.U0;
    ctx->adj_prv = ctx->aux.U1;
    ctx->adj_next_t = -1;
    HCK(hipStreamSynchronize(ctx->st));
    return PAM_OK;
}

extern "C" int pam_adjoint_step(pam_ctx *ctx, int t, const float *S_t, float *image, float *ilum) {
    if (!ctx) return PAM_E_ARG;
    if (ctx->adj_next_t < -1) return fail(ctx, PAM_E_STATE, "pam_adjoint_reset must precede pam_adjoint_step");
    const pam_grid &g = ctx->g;
    if (t < 0 || t >= g.nt) return fail(ctx, PAM_E_ARG, "t=%d outside [0, nt=%d)", t, g.nt);
    if (ctx->adj_next_t >= 0 && t != ctx->adj_next_t)
        return fail(ctx, PAM_E_STATE, "adjoint steps must be issued in decreasing order: expected t=%d", ctx->adj_next_t);
    const bool linear_adj = ctx->phys >= 1 && ctx->abc == PAM_CERJAN;
    if (!S_t && (image || ilum || linear_adj))
        return fail(ctx, PAM_E_ARG, "S_t is NULL (imaging and the linearised VTI/TTI adjoint read it)");
    const long long nd = (long long)(g.nt - g.itlag) * ctx->nrec;
    const long long nint = (long long)g.nz * g.nx;
    // the first step: data row t into the adjoint field; each step then adds row
    // t-1 to the field it writes (writer-side, as the whole-run adjoint pass does)
    if (ctx->adj_next_t < 0 && t >= g.itlag)
        KCK(launch_inject_rec(ctx->adj_cur, ctx->plane, ctx->P, ctx->d_rz, ctx->d_rx, ctx->nrec, ctx->adj_data, nd,
                              t - g.itlag, g.dx * g.dz, 1, ctx->st));
    if (S_t) {
        KCK(launch_copy_padded(S_t, ctx->nzp, ctx->nxp, ctx->adj_S, ctx->P, ctx->st));
        // the linearised adjoint reads the source field as the forward step
        // saves it: damped once more (interior cells are unchanged)
        if (linear_adj) KCK(launch_cerjan_scale(ctx->adj_S, ctx->nzp, ctx->nxp, ctx->P, ctx->cerx, ctx->cerz, ctx->st));
        if (ilum) KCK(launch_ilum_acc(ctx->adj_S, ctx->P, g.n_abc, g.nz, g.nx, ilum, ctx->st));
    }
    if (!linear_adj) {
        StepArgs b;
        base_step_args(ctx, b, 1);
        b.flags |= F_RECINJ | (image ? F_IMAGE : 0);
        b.data = ctx->adj_data;
        b.dstride = nd;
        b.save_kind = SV_F32_PAD;
        b.save = ctx->adj_S;
        b.save_k0 = t;
        b.save_kstride = b.save_sstride = ctx->plane;
        b.image = image;
        b.acc_sstride = nint;
        const int r = fwd_step_fs(ctx, ctx->aux, b, choose_tile(ctx, 1, false), ctx->adj_cur, ctx->adj_prv, t, nullptr);
        if (r) return r;
    } else {
        AdjArgs c;
        std::memset(&c, 0, sizeof(c));
        c.nzp = ctx->nzp;
        c.nxp = ctx->nxp;
        c.P = ctx->P;
        c.N = g.n_abc;
        c.nz = g.nz;
        c.nx = g.nx;
        c.nshots = 1;
        c.sstride = ctx->plane;
        c.idx2 = 1.0f / (g.dx * g.dx);
        c.idz2 = 1.0f / (g.dz * g.dz);
        c.idx = 1.0f / g.dx;
        c.idz = 1.0f / g.dz;
        c.idxdz = 1.0f / (g.dz * g.dx);
        c.idt2 = 1.0f / (g.dt * g.dt);
        c.dxdz = g.dx * g.dz;
        c.k = t;
        c.nt = g.nt;
        c.itlag = g.itlag;
        c.nrec = ctx->nrec;
        c.cur = ctx->adj_cur;
        c.prv = ctx->adj_prv;
        c.P_ = ctx->adj_S;
        c.psstride = c.pkstride = ctx->plane;
        c.v2 = ctx->v2;
        c.eps = ctx->eps;
        c.del = ctx->del;
        c.cs = ctx->cs;
        c.sn = ctx->sn;
        c.cerx = ctx->cerx;
        c.cerz = ctx->cerz;
        c.rrow = ctx->d_rrow;
        c.rhead = ctx->d_rhead;
        c.rnext = ctx->d_rnext;
        c.data = ctx->adj_data;
        c.dstride = nd;
        c.image = image;
        c.acc_sstride = nint;
        KCK(launch_adj_step(ctx->phys, c, ctx->st));
    }
    std::swap(ctx->adj_cur, ctx->adj_prv);
    ctx->adj_next_t = t - 1;
    return PAM_OK;
}

// ------------------------------------------------------------ whole runs
// (re)allocates a scratch buffer to hold at least `need` bytes
static int ensure_scratch(pam_ctx *ctx, void **p, long long *have, long long need) {
    if (*have >= need && *p) return PAM_OK;
    dfree(*p);
    *p = nullptr;
    *have = 0;
    int r = dalloc(ctx, p, (size_t)need);
    if (r) return r;
    *have = need;
    return PAM_OK;
}

static int batch_size(pam_ctx *ctx, int nshots) {
    const int b = ctx->batch > 0 ? ctx->batch : auto_batch(ctx);
    return std::max(1, std::min(b, std::min(nshots, ctx->max_shots)));
}

static int finish_timing(pam_ctx *ctx, long long nk) {
    HCK(hipEventSynchronize(ctx->e1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
    ctx->loop_ms += ms;
    if (nk > 0 && ctx->profile) {
        double acc = 0.0;
        for (long long i = 0; i < nk; ++i) {
            float m = 0.f;
            HCK(hipEventElapsedTime(&m, ctx->kev[2 * i], ctx->kev[2 * i + 1]));
            acc += m;
        }
        ctx->kernel_ms += acc;
        ctx->launches += nk;
    }
    return PAM_OK;
}

// PAM_OPT_PROFILE 2: the step kernel alone, kStepTimingReps launches back to
// back on the fields of the run's last batch (the forward instance: source
// injection and receiver gather into seis on, steps itlag .. ; no CPML
// pre-pass between them, psi as the run left it), bracketed by two HIP events
// on the context's stream: kernel_ms / kernel_launches of pam_last_timing.
// The fields and the first nb gathers of seis are overwritten.
static constexpr int kStepTimingReps = 200;
static int time_step_kernel(pam_ctx *ctx, int nb, int tile, float *seis, long long nd) {
    StepArgs a;
    base_step_args(ctx, a, nb);
    a.flags |= F_SRC | F_GATHER;
    a.seis = seis;
    a.dstride = nd;
    FieldSet &f = ctx->fs;
    if (ctx->abc == PAM_CPML) {
        const long long in = f.psi_flip ? ctx->psi_half : 0, out = f.psi_flip ? 0 : ctx->psi_half;
        a.psix_in = f.psix + in;
        a.psiz_in = f.psiz + in;
        a.psix = f.psix + out;
        a.psiz = f.psiz + out;
        a.zetax = f.zetax;
        a.zetaz = f.zetaz;
    }
    const int span = std::max(1, ctx->g.nt - ctx->g.itlag);
    float *c = f.U0, *p = f.U1;
    for (int i = -4; i < kStepTimingReps; ++i) {  // 4 untimed launches first
        if (i == 0) HCK(hipEventRecord(ctx->e0, ctx->st));
        a.cur = c;
        a.prv = p;
        a.k = ctx->g.itlag + ((i + 4) % span);
        KCK(launch_step(ctx->phys, ctx->abc == PAM_CPML, tile, a, ctx->st, nullptr));
        std::swap(c, p);
    }
    HCK(hipEventRecord(ctx->e1, ctx->st));
    HCK(hipEventSynchronize(ctx->e1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
    ctx->kernel_ms = ms;
    ctx->launches = kStepTimingReps;
    return PAM_OK;
}

// Runs `body` (host code that enqueues the launches of a step loop on the
// context's stream) through an instantiated hipGraph cached under `key`:
// captured on first use, replayed afterwards.  The launches must be
// capturable (kernels and device copies: no memset nodes, no events) and
// `key` must name everything they bake in that can change between calls;
// `replay_fix` applies the host-side effects of `body` when it is skipped.
// The graphs go to `cache` (the context's own, or a group's: a group's graphs
// bake in its device tables and are destroyed with it).
template <class Body, class Fix>
static int run_graph_in(pam_ctx *ctx, std::map<std::string, hipGraphExec_t> &cache, const std::string &key,
                        Body &&body, Fix &&replay_fix, hipStream_t st = nullptr) {
    if (!st) st = ctx->st;
    auto it = cache.find(key);
    if (it != cache.end()) {
        replay_fix();
        HCK(hipGraphLaunch(it->second, st));
        return PAM_OK;
    }
    if (cache.size() >= 16) {  // bound the cache: drop every entry once none can be running
        HCK(hipStreamSynchronize(ctx->st));
        for (hipStream_t l : ctx->lane_st)
            if (l) HCK(hipStreamSynchronize(l));
        for (auto &kv : cache) (void)hipGraphExecDestroy(kv.second);
        cache.clear();
    }
    HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int r = body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(ctx, PAM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(ctx, PAM_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    cache[key] = ex;
    HCK(hipGraphLaunch(ex, st));
    return PAM_OK;
}
template <class Body, class Fix>
static int run_graph(pam_ctx *ctx, const std::string &key, Body &&body, Fix &&replay_fix, hipStream_t st = nullptr) {
    return run_graph_in(ctx, ctx->graphs, key, body, replay_fix, st);
}

// forward loop over nt steps for shots already primed in U0 (cur) / U1 (prv);
// extra flags / args are applied to every step.  With PAM_OPT_PROFILE each
// step kernel is bracketed by HIP events; otherwise, with PAM_OPT_GRAPH, the
// loop is replayed as a captured graph.
static int run_forward_loop_fs(pam_ctx *ctx, FieldSet &f, hipStream_t st, StepArgs &a, int tile, bool timed);
static int run_forward_loop(pam_ctx *ctx, StepArgs &a, int tile, bool timed) {
    return run_forward_loop_fs(ctx, ctx->fs, ctx->st, a, tile, timed);
}
// the same on field set f and stream st
static int run_forward_loop_fs(pam_ctx *ctx, FieldSet &f, hipStream_t st, StepArgs &a, int tile, bool timed) {
    const int nt = ctx->g.nt;
    const bool events = timed && ctx->profile == 1;
    if (events) {
        int r = ensure_events(ctx, 2 * (size_t)nt);
        if (r) return r;
    }
    if (timed) HCK(hipEventRecord(ctx->e0, st));
    const int base_flags = a.flags;
    // the temporal-blocking variant: an even number of tblock_steps-step
    // launches after the first nt % (2 tblock_steps) steps taken singly (its
    // inner steps gather only from itlag on themselves)
    const bool tb = !events && tblock_usable(ctx, tile, base_flags);
    const int nsingle = tb ? nt % (2 * tblock_steps(ctx->phys)) : nt;
    if (tb) {
        int r = ensure_tblock(ctx);
        if (r) return r;
    }
    auto body = [&]() -> int {
        float *cur = f.U0, *prv = f.U1;
        for (int k = 0; k < nsingle; ++k) {
            a.flags = base_flags;
            if (!(k >= ctx->g.itlag)) a.flags &= ~F_GATHER;
            int r = fwd_step_fs(ctx, f, a, tile, cur, prv, k, events ? &ctx->kev[2 * k] : nullptr, st);
            if (r) return r;
            std::swap(cur, prv);
        }
        if (tb) {
            a.flags = base_flags;
            int r = tblock_launches(ctx, f, a, st, cur, prv, nsingle, (nt - nsingle) / tblock_steps(ctx->phys));
            if (r) return r;
        }
        return PAM_OK;
    };
    int r;
    if (ctx->graph && !events) {
        char key[256];
        std::snprintf(key, sizeof key, "fwd|%u|%d|%d|%d|%p|%p|%p|%p", ctx->epoch, a.nshots, tile, base_flags,
                      (void *)a.seis, (const void *)a.data, (void *)f.U0, (const void *)a.ssz);
        r = run_graph(
                ctx, key, body,
                [&]() {
                    if (ctx->abc == PAM_CPML) f.psi_flip ^= (nt & 1);
                },
                st);
    } else {
        r = body();
    }
    a.flags = base_flags;
    if (r) return r;
    if (timed) {
        HCK(hipEventRecord(ctx->e1, st));
        r = finish_timing(ctx, events ? nt : 0);
        if (r) return r;
        if (!events) {
            float ms = 0.f;
            HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
            ctx->kernel_ms += ms;
            ctx->launches += nt;
        }
    }
    return PAM_OK;
}

// The lane streams (default priority, created together on first use) and
// their fork from / join into the context's stream.
static int ensure_lanes(pam_ctx *ctx) {
    for (int i = 0; i < 2; ++i) {
        if (!ctx->lane_st[i]) HCK(hipStreamCreate(&ctx->lane_st[i]));
        if (!ctx->join_ev[i]) HCK(hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming));
    }
    if (!ctx->fork_ev) HCK(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    return PAM_OK;
}
static int lanes_fork(pam_ctx *ctx) {
    HCK(hipEventRecord(ctx->fork_ev, ctx->st));
    for (hipStream_t l : ctx->lane_st) HCK(hipStreamWaitEvent(l, ctx->fork_ev, 0));
    return PAM_OK;
}
// the context's stream waits for both lanes, on every path (r0: the result so
// far); if that cannot be arranged the lanes are drained
static int lanes_join(pam_ctx *ctx, int r0) {
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipEventRecord(ctx->join_ev[i], ctx->lane_st[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->st, ctx->join_ev[i], 0);
    }
    if (e != hipSuccess)
        for (hipStream_t l : ctx->lane_st) (void)hipStreamSynchronize(l);
    if (r0) return r0;
    if (e != hipSuccess) return fail(ctx, PAM_E_HIP, "joining the lanes: %s", hipGetErrorString(e));
    return PAM_OK;
}

// The autotuner for runs on two lanes: every variant timed with both lanes
// stepping Bl shots at once (a few plain steps each, lanes forked from and
// joined into the context's stream, HIP events on it), so that what one lane's
// CPML pre-pass and launch gaps leave to the other counts.  The fields of both
// lanes' slots are cleared before and after.
static int autotune_lanes(pam_ctx *ctx, int Bl, int *best) {
    *best = kDefaultTile;
    int r = ensure_lanes(ctx);
    if (r) return r;
    FieldSet lane1 = ctx->fs;
    const long long off = (long long)Bl * ctx->plane;
    lane1.U0 += off;
    lane1.U1 += off;
    if (ctx->abc == PAM_CPML) {
        lane1.psix += off;
        lane1.psiz += off;
        lane1.zetax += off;
        lane1.zetaz += off;
    }
    FieldSet *fsl[2] = {&ctx->fs, &lane1};
    StepArgs a[2];
    for (int l = 0; l < 2; ++l) base_step_args(ctx, a[l], Bl);
    r = zero_shot_state(ctx, 2 * Bl, true, true);
    if (r) return r;
    double best_ms = 1e300;  // per step
    for (int t = 0; t < tile_count(); ++t) {
        float *c[2] = {fsl[0]->U0, fsl[1]->U0}, *p[2] = {fsl[0]->U1, fsl[1]->U1};
        const bool tb = tile_is_tblock(t);
        if (tb && !tblock_usable(ctx, t, a[0].flags)) continue;
        if (tb) {
            r = ensure_tblock(ctx);
            if (r) return r;
        }
        // n single steps, or n / tblock_steps launches of the temporal-blocking variant
        auto steps = [&](int k0, int n) -> int {
            if (tb) {
                for (int l = 0; l < 2; ++l) {
                    int rr = tblock_launches(ctx, *fsl[l], a[l], ctx->lane_st[l], c[l], p[l], k0,
                                             n / tblock_steps(ctx->phys));
                    if (rr) return rr;
                }
                return PAM_OK;
            }
            for (int i = 0; i < n; ++i)
                for (int l = 0; l < 2; ++l) {
                    int rr = fwd_step_fs(ctx, *fsl[l], a[l], t, c[l], p[l], k0 + i, nullptr, ctx->lane_st[l]);
                    if (rr) return rr;
                    std::swap(c[l], p[l]);
                }
            return PAM_OK;
        };
        const int nwarm = tb ? 2 * tblock_steps(ctx->phys) : 2, ntimed = tb ? 2 * tblock_steps(ctx->phys) : 6;
        r = lanes_fork(ctx);
        if (r) return r;
        r = lanes_join(ctx, steps(0, nwarm));
        if (r) return r;
        double t_ms = 1e300;  // the best of kTuneRounds windows
        for (int round = 0; round < kTuneRounds; ++round) {
            HCK(hipEventRecord(ctx->e0, ctx->st));
            r = lanes_fork(ctx);
            if (r) return r;
            r = lanes_join(ctx, steps(nwarm + round * ntimed, ntimed));
            if (r) return r;
            HCK(hipEventRecord(ctx->e1, ctx->st));
            HCK(hipEventSynchronize(ctx->e1));
            float ms = 0.f;
            HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
            t_ms = std::min(t_ms, (double)ms / ntimed);
        }
        if (t_ms < best_ms) {
            best_ms = t_ms;
            *best = t;
        }
    }
    return zero_shot_state(ctx, 2 * Bl, true, true);
}
// variant for two lanes of Bl shots: the option if set, else tuned on two lanes
static int choose_tile_lanes(pam_ctx *ctx, int Bl) {
    if (ctx->tile >= 0) return ctx->tile;
    auto it = ctx->tuned_lanes.find(Bl);
    if (it != ctx->tuned_lanes.end()) return it->second;
    int best = kDefaultTile;
    if (autotune_lanes(ctx, Bl, &best) != PAM_OK) {
        (void)hipGetLastError();
        best = choose_tile(ctx, Bl, true);
    }
    ctx->tuned_lanes[Bl] = best;
    return best;
}

// pam_forward on two lanes of Bl shots per launch: batches alternate between
// the two lane streams, each lane with its own shot slots, source slots,
// gather staging and cached graph; the lanes start after the work already
// queued on the context's stream, which waits for both at the end.  One
// lane's psi pre-pass and launch gaps then overlap the other's step kernel.
// Every batch is computed exactly as on one lane.
static int forward_two_lanes(pam_ctx *ctx, int shot0, int nshots, float *seis, int Bl, long long nd) {
    int rl = ensure_lanes(ctx);
    if (rl) return rl;
    const int tile = choose_tile_lanes(ctx, Bl);
    ctx->last_tile = tile;
    ctx->last_batch = Bl;
    const bool staged = ctx->graph != 0;
    if (staged) {
        int r = ensure_scratch(ctx, (void **)&ctx->stage, &ctx->stage_bytes, 2LL * Bl * nd * 4);
        if (r) return r;
    }
    // lane 1's view of the field set: shot slots Bl .. 2 Bl - 1
    FieldSet lane1 = ctx->fs;
    const long long off = (long long)Bl * ctx->plane;
    lane1.U0 += off;
    lane1.U1 += off;
    if (ctx->abc == PAM_CPML) {
        lane1.psix += off;
        lane1.psiz += off;
        lane1.zetax += off;
        lane1.zetaz += off;
    }
    FieldSet *fsl[2] = {&ctx->fs, &lane1};
    hipStream_t stl[2] = {ctx->lane_st[0], ctx->lane_st[1]};
    HCK(hipEventRecord(ctx->e0, ctx->st));
    int nbatch = 0;
    auto batches = [&]() -> int {
        for (int b0 = 0; b0 < nshots; b0 += Bl, ++nbatch) {
            const int lane = nbatch & 1;
            const int nb = std::min(Bl, nshots - b0);
            FieldSet &f = *fsl[lane];
            hipStream_t st = stl[lane];
            int r = zero_fields(ctx, f, nb, true, true, st);
            if (r) return r;
            r = upload_sources(ctx, shot0 + b0, nb, lane * Bl, st);
            if (r) return r;
            float *dst = staged ? ctx->stage + (long long)lane * Bl * nd : seis + (long long)b0 * nd;
            HCK(hipMemsetAsync(dst, 0, (size_t)nb * nd * sizeof(float), st));
            KCK(launch_inject_point(f.U0, ctx->plane, ctx->d_ssz + lane * Bl, ctx->d_ssx + lane * Bl, ctx->P,
                                    ctx->wav, 0, nb, st));
            StepArgs a;
            base_step_args(ctx, a, nb);
            a.flags |= F_SRC | F_GATHER;
            a.seis = dst;
            a.dstride = nd;
            a.ssz = ctx->d_ssz + lane * Bl;
            a.ssx = ctx->d_ssx + lane * Bl;
            r = run_forward_loop_fs(ctx, f, st, a, tile, false);
            if (r) return r;
            r = bad_check(ctx, f.U0, (long long)nb * ctx->plane, st);
            if (r) return r;
            r = bad_check(ctx, f.U1, (long long)nb * ctx->plane, st);
            if (r) return r;
            if (staged)
                HCK(hipMemcpyAsync(seis + (long long)b0 * nd, dst, (size_t)nb * nd * sizeof(float),
                                   hipMemcpyDeviceToDevice, st));
        }
        return PAM_OK;
    };
    rl = bad_clear(ctx);
    if (rl) return rl;
    rl = lanes_fork(ctx);
    if (rl) return rl;
    rl = lanes_join(ctx, batches());  // the context's stream then orders after all lane work
    if (rl) return rl;
    HCK(hipEventRecord(ctx->e1, ctx->st));
    HCK(hipEventSynchronize(ctx->e1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
    ctx->loop_ms += ms;
    ctx->kernel_ms += ms;
    ctx->launches += (long long)nbatch * ctx->g.nt;
    return bad_report(ctx, "pam_forward");
}

extern "C" int pam_forward(pam_ctx *ctx, int shot0, int nshots, float *seis) {
    if (!ctx) return PAM_E_ARG;
    int r = ready(ctx);
    if (r) return r;
    if (nshots < 1) return fail(ctx, PAM_E_ARG, "nshots must be >= 1");
    if (!seis) return fail(ctx, PAM_E_ARG, "seis is NULL");
    const long long nd = (long long)(ctx->g.nt - ctx->g.itlag) * ctx->nrec;
    ctx->loop_ms = ctx->kernel_ms = 0.0;
    ctx->launches = 0;
    ctx->next_k = -1;
    int B = batch_size(ctx, nshots);
    if (ctx->lanes > 1 && nshots > 1 && ctx->max_shots > 1) {
        // two lanes: each half the batch (or PAM_OPT_BATCH), in shot slots
        // [0, Bl) and [Bl, 2 Bl) of the field set, on the context's stream and
        // the second one; batch i goes to lane i % 2.  Profiled runs make the
        // same launches on one lane.
        const int Bl = std::max(1, std::min(ctx->batch > 0 ? ctx->batch : (B + 1) / 2, ctx->max_shots / 2));
        if (!ctx->profile) return forward_two_lanes(ctx, shot0, nshots, seis, Bl, nd);
        B = Bl;
    }
    const int tile = choose_tile(ctx, B, true);
    ctx->last_tile = tile;
    ctx->last_batch = B;
    const bool staged = ctx->graph && !ctx->profile;
    if (staged) {
        r = ensure_scratch(ctx, (void **)&ctx->stage, &ctx->stage_bytes, (long long)B * nd * 4);
        if (r) return r;
    }
    r = bad_clear(ctx);
    if (r) return r;
    for (int b0 = 0; b0 < nshots; b0 += B) {
        const int nb = std::min(B, nshots - b0);
        r = zero_shot_state(ctx, nb, true, true);
        if (r) return r;
        r = upload_sources(ctx, shot0 + b0, nb);
        if (r) return r;
        float *dst = staged ? ctx->stage : seis + (long long)b0 * nd;
        HCK(hipMemsetAsync(dst, 0, (size_t)nb * nd * sizeof(float), ctx->st));
        KCK(launch_inject_point(ctx->fs.U0, ctx->plane, ctx->d_ssz, ctx->d_ssx, ctx->P, ctx->wav, 0, nb, ctx->st));
        StepArgs a;
        base_step_args(ctx, a, nb);
        a.flags |= F_SRC | F_GATHER;
        a.seis = dst;
        a.dstride = nd;
        r = run_forward_loop(ctx, a, tile, true);
        if (r) return r;
        r = bad_check(ctx, ctx->fs.U0, (long long)nb * ctx->plane);
        if (r) return r;
        r = bad_check(ctx, ctx->fs.U1, (long long)nb * ctx->plane);
        if (r) return r;
        if (staged)
            HCK(hipMemcpyAsync(seis + (long long)b0 * nd, ctx->stage, (size_t)nb * nd * sizeof(float),
                               hipMemcpyDeviceToDevice, ctx->st));
    }
    r = bad_report(ctx, "pam_forward");
    if (r == PAM_OK && ctx->profile == 2) r = time_step_kernel(ctx, std::min(B, nshots), tile, seis, nd);
    return r;
}

// ------------------------------------------------------------ preprocessing
static int free_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int free_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_free_err = buf;
    return code;
}
#define FCK(call)                                                                                          \
    do {                                                                                                   \
        hipError_t _e = (hipError_t)(call);                                                                \
        if (_e != hipSuccess) return free_fail(PAM_E_HIP, "%s: %s", #call, hipGetErrorString(_e));         \
    } while (0)

// stream-ordered scratch released when the call returns
struct FreeScratch {
    hipStream_t st;
    std::vector<void *> p;
    explicit FreeScratch(hipStream_t s) : st(s) {}
    void *get(size_t bytes) {
        void *q = nullptr;
        if (hipMallocAsync(&q, bytes ? bytes : 4, st) != hipSuccess) return nullptr;
        p.push_back(q);
        return q;
    }
    ~FreeScratch() {
        for (void *q : p) (void)hipFreeAsync(q, st);
        (void)hipStreamSynchronize(st);
    }
};

extern "C" int pam_smooth_model(const float *in, float *out, const unsigned char *mask, int nz, int nx, double sigma,
                                int parameter, int mode, void *hip_stream) {
    if (!in || !out || nz < 1 || nx < 1 || !(sigma > 0.0) || (mode != PAM_PREP_PY && mode != PAM_PREP_CPP))
        return free_fail(PAM_E_ARG, "pam_smooth_model: invalid arguments");
    hipStream_t st = (hipStream_t)hip_stream;
    const long long n = (long long)nz * nx;
    FreeScratch sc(st);
    float *s_old = (float *)sc.get(n * 4), *tmp = (float *)sc.get(n * 4);
    if (!s_old || !tmp) return free_fail(PAM_E_NOMEM, "pam_smooth_model: scratch");
    FCK(launch_slowness(in, s_old, n, parameter, st));
    if (mode == PAM_PREP_PY) {
        // the normalised 2-D Gaussian of src/utils.py:51-74
        int ks = (int)std::ceil(6.0 * sigma + 1.0);
        if (ks % 2 == 0) ++ks;
        const int h = ks / 2;
        if (nz <= 2 * h || nx <= 2 * h)
            return free_fail(PAM_E_ARG, "pam_smooth_model: a %dx%d kernel does not fit a %dx%d model", ks, ks, nz, nx);
        std::vector<double> val((size_t)ks * ks);
        double vsum = 0.0;
        for (int i = 0; i < ks; ++i)
            for (int j = 0; j < ks; ++j) {
                const double x = i - h, z = j - h;
                val[(size_t)i * ks + j] = (1.0 / (2.0 * M_PI * sigma * sigma)) * std::exp(-(x * x + z * z) / (2.0 * sigma * sigma));
                vsum += val[(size_t)i * ks + j];
            }
        std::vector<float> w((size_t)ks * ks);
        double wsum = 0.0;
        for (size_t i = 0; i < w.size(); ++i) {
            w[i] = (float)val[i] / (float)vsum;
            wsum += (double)w[i];
        }
        float *dw = (float *)sc.get(w.size() * 4);
        int *first = (int *)sc.get((size_t)nx * 4);
        if (!dw || !first) return free_fail(PAM_E_NOMEM, "pam_smooth_model: scratch");
        FCK(hipMemcpyAsync(dw, w.data(), w.size() * 4, hipMemcpyHostToDevice, st));
        FCK(launch_first_dry(mask, nz, nx, first, st));
        FCK(launch_smooth_conv2d(s_old, nz, nx, dw, ks, wsum, tmp, st));
        FCK(launch_smooth_finish_py(s_old, tmp, first, nz, nx, h, parameter, out, st));
        FCK(hipStreamSynchronize(st));  // the host weights are released
    } else {
        // the normalised 1-D Gaussian of cpp/Migration.cu:147-169, float32
        int ks = (int)std::ceil(6.0f * (float)sigma + 1.0f);
        if (ks % 2 == 0) ++ks;
        const int h = ks / 2;
        const float sg = (float)sigma;
        std::vector<float> w(ks);
        float total = 0.0f;
        for (int i = 0; i < ks; ++i) {
            const float x = (float)(i - h);
            const float num = -(x * x);
            const float den = 2.0f * sg * sg;
            w[i] = std::exp(num / den);
            total = total + w[i];
        }
        for (int i = 0; i < ks; ++i) w[i] = w[i] / total;
        float *dw = (float *)sc.get((size_t)ks * 4), *sm = (float *)sc.get(n * 4);
        if (!dw || !sm) return free_fail(PAM_E_NOMEM, "pam_smooth_model: scratch");
        FCK(hipMemcpyAsync(dw, w.data(), (size_t)ks * 4, hipMemcpyHostToDevice, st));
        FCK(launch_smooth_sep(s_old, nz, nx, dw, ks, 0, tmp, st));
        FCK(launch_smooth_sep(tmp, nz, nx, dw, ks, 1, sm, st));
        FCK(launch_smooth_finish_cpp(s_old, sm, mask, n, parameter, out, st));
        FCK(hipStreamSynchronize(st));
    }
    return PAM_OK;
}

extern "C" int pam_mute(float *seis, int nshots, int nt, int nrec, const double *rec_x, const double *rec_z,
                        const double *src_x, const double *src_z, double dt, double shift, double window, double v0,
                        int mode, void *hip_stream) {
    if (!seis || nshots < 1 || nt < 1 || nrec < 1 || !rec_x || !rec_z || !src_x || !src_z || !(v0 > 0.0) ||
        (mode != PAM_PREP_PY && mode != PAM_PREP_CPP))
        return free_fail(PAM_E_ARG, "pam_mute: invalid arguments");
    hipStream_t st = (hipStream_t)hip_stream;
    const size_t m = (size_t)nshots * nrec;
    std::vector<double> t64;
    std::vector<float> t32;
    if (mode == PAM_PREP_PY) {
        t64.resize(2 * m);
        for (int s = 0; s < nshots; ++s)
            for (int r = 0; r < nrec; ++r) {
                const double dz = rec_z[r] - src_z[s], dx = rec_x[r] - src_x[s];
                const double t1 = std::sqrt(dz * dz + dx * dx) / v0 + shift;
                t64[2 * ((size_t)s * nrec + r)] = t1;
                t64[2 * ((size_t)s * nrec + r) + 1] = t1 + window;
            }
    } else {
        t32.resize(2 * m);
        for (int s = 0; s < nshots; ++s)
            for (int r = 0; r < nrec; ++r) {
                const float dz = (float)rec_z[r] - (float)src_z[s], dx = (float)rec_x[r] - (float)src_x[s];
                const float dz2 = dz * dz, dx2 = dx * dx;
                const float dist = std::sqrt(dz2 + dx2);
                const float q = dist / (float)v0;
                const float t1 = q + (float)shift;
                t32[2 * ((size_t)s * nrec + r)] = t1;
                t32[2 * ((size_t)s * nrec + r) + 1] = t1 + (float)window;
            }
    }
    FreeScratch sc(st);
    const size_t bytes = mode == PAM_PREP_PY ? t64.size() * 8 : t32.size() * 4;
    void *d = sc.get(bytes);
    if (!d) return free_fail(PAM_E_NOMEM, "pam_mute: scratch");
    FCK(hipMemcpyAsync(d, mode == PAM_PREP_PY ? (const void *)t64.data() : (const void *)t32.data(), bytes,
                       hipMemcpyHostToDevice, st));
    FCK(launch_mute(seis, nshots, nt, nrec, d, dt, mode == PAM_PREP_CPP, st));
    FCK(hipStreamSynchronize(st));
    return PAM_OK;
}

// ------------------------------------------------------------ groups of grids
// Several contexts (models of different sizes, each with its own shots)
// advanced together: one z-march launch per time step covers every shot of
// every grid, blocks assigned to grids through a prefix table.
struct pam_group {
    std::vector<pam_ctx *> ctx;
    StepArgs *d_list = nullptr;
    int *d_boff = nullptr, *d_pboff = nullptr;  // step-kernel and psi-pass block prefix tables (2n: per sub-group)
    std::map<std::string, int> tuned;  // step-kernel variant per shot layout
    int last_tile = -1;
    // the group's step-loop graphs: they read d_list / d_boff / d_pboff, so they
    // live and die with the group (a new group over the same members must not
    // replay a graph that reads a destroyed group's tables)
    std::map<std::string, hipGraphExec_t> graphs;
    // the two streams the gradient workers run their members on (default
    // priority, created on first use) and the fork / join event
    hipStream_t wst[2] = {nullptr, nullptr};
    hipEvent_t wev = nullptr;
};

extern "C" int pam_group_create(pam_ctx *const *ctxs, int n, pam_group **out) {
    if (!ctxs || n < 1 || !out) return PAM_E_ARG;
    pam_ctx *c0 = ctxs[0];
    if (!c0) return PAM_E_ARG;
    for (int i = 0; i < n; ++i) {
        const pam_ctx *c = ctxs[i];
        if (!c) return fail(c0, PAM_E_ARG, "group member %d is NULL", i);
        if (c->phys != c0->phys || c->abc != c0->abc)
            return fail(c0, PAM_E_ARG, "group members must share physics and absorbing boundary (member %d)", i);
        if (c->g.nt != c0->g.nt) return fail(c0, PAM_E_ARG, "group members must share nt (member %d)", i);
    }
    pam_group *g = new pam_group();
    g->ctx.assign(ctxs, ctxs + n);
    if (hipMalloc((void **)&g->d_list, sizeof(StepArgs) * n) != hipSuccess ||
        hipMalloc((void **)&g->d_boff, sizeof(int) * 2 * n) != hipSuccess ||
        hipMalloc((void **)&g->d_pboff, sizeof(int) * 2 * n) != hipSuccess) {
        dfree(g->d_list);
        dfree(g->d_boff);
        dfree(g->d_pboff);
        delete g;
        return fail(c0, PAM_E_NOMEM, "group argument tables");
    }
    *out = g;
    return PAM_OK;
}

extern "C" int pam_group_destroy(pam_group *g) {
    if (!g) return PAM_OK;
    if (!g->ctx.empty()) {
        pam_ctx *c0 = g->ctx[0];
        if (c0->st) (void)hipStreamSynchronize(c0->st);
        for (hipStream_t l : c0->lane_st)
            if (l) (void)hipStreamSynchronize(l);
    }
    for (auto &kv : g->graphs) (void)hipGraphExecDestroy(kv.second);
    g->graphs.clear();
    for (hipStream_t &w : g->wst)
        if (w) {
            (void)hipStreamSynchronize(w);
            (void)hipStreamDestroy(w);
            w = nullptr;
        }
    if (g->wev) (void)hipEventDestroy(g->wev);
    dfree(g->d_list);
    dfree(g->d_boff);
    dfree(g->d_pboff);
    delete g;
    return PAM_OK;
}

extern "C" int pam_group_forward(pam_group *g, const int *nshots, float *const *seis, double *loop_ms) {
    if (!g || !nshots || !seis) return PAM_E_ARG;
    pam_ctx *c0 = g->ctx[0];
    pam_ctx *const ctx = c0;  // errors of the group go to member 0
    const int n = (int)g->ctx.size();
    const int nt = c0->g.nt;
    for (int i = 0; i < n; ++i) {
        pam_ctx *c = g->ctx[i];
        int r = ready(c);
        if (r) return fail(c0, r, "group member %d: %s", i, c->err.c_str());
        if (nshots[i] < 1 || nshots[i] > c->max_shots || nshots[i] > c->nshot)
            return fail(c0, PAM_E_ARG, "group member %d: nshots %d outside [1, min(max_shots, shots)]", i, nshots[i]);
        if (!seis[i]) return fail(c0, PAM_E_ARG, "group member %d: seis is NULL", i);
    }
    // sub-groups of consecutive members whose wavefields (two per shot) stay
    // within member 0's PAM_OPT_GROUP_BYTES, so that each launch's working set
    // stays in the MALL as a single grid's batch does (a member larger than
    // that forms a sub-group of its own); sub-group j holds members
    // [first[j], first[j+1]) and its block tables start at first[j] + j
    std::vector<int> first;
    long long acc = 0;
    for (int i = 0; i < n; ++i) {
        const long long bytes = 8LL * g->ctx[i]->plane * nshots[i];
        if (i == 0 || (c0->group_bytes > 0 && acc + bytes > c0->group_bytes)) {
            first.push_back(i);
            acc = 0;
        }
        acc += bytes;
    }
    const int nsub = (int)first.size();
    first.push_back(n);
    std::vector<StepArgs> list(n);
    std::vector<int> boff(n + nsub, 0), pboff(n + nsub, 0), nblk(nsub, 0), npblk(nsub, 0);
    // the variant (the 16-row tile or a z-march): member 0's PAM_OPT_TILE when
    // it names one of these, else timed
    std::string layout;
    for (int i = 0; i < n; ++i) layout += std::to_string(nshots[i]) + ",";
    int tile = (c0->tile >= 0 && tile_has_multi(c0->tile)) ? c0->tile : -1;
    auto build = [&](int t) -> int {
        const int ch = march_chunks(t);
        for (int j = 0; j < nsub; ++j) {
            const int o = first[j] + j;  // this sub-group's tables
            boff[o] = pboff[o] = 0;
            for (int i = first[j]; i < first[j + 1]; ++i) {
                pam_ctx *c = g->ctx[i];
                const int nb = nshots[i];
                const long long nd = (long long)(c->g.nt - c->g.itlag) * c->nrec;
                StepArgs &a = list[i];
                base_step_args(c, a, nb);
                a.flags |= F_SRC | F_GATHER;
                a.seis = seis[i];
                a.dstride = nd;
                a.cur = c->fs.U0;
                a.prv = c->fs.U1;
                if (c->abc == PAM_CPML) {
                    a.psix_in = c->fs.psix;
                    a.psiz_in = c->fs.psiz;
                    a.psix = c->fs.psix + c->psi_half;
                    a.psiz = c->fs.psiz + c->psi_half;
                }
                a.nxt = (a.nxp + 63) / 64;
                a.nzt = (t == 0) ? (a.nzp + 15) / 16 : (a.nzp - 8 + ch * 16 - 1) / (ch * 16);
                const int l = o + (i - first[j]);
                boff[l + 1] = boff[l] + a.nxt * a.nzt * nb;
                const bool psi_pass = t == 0 && c->abc == PAM_CPML;  // the tile reads psi from a pre-pass
                pboff[l + 1] = pboff[l] + (psi_pass ? psi_blocks_per_shot(c->nzp, c->nxp, c->g.n_abc) * nb : 0);
            }
            const int e = o + (first[j + 1] - first[j]);
            nblk[j] = boff[e];
            npblk[j] = pboff[e];
        }
        HCK(hipMemcpyAsync(g->d_list, list.data(), sizeof(StepArgs) * n, hipMemcpyHostToDevice, c0->st));
        HCK(hipMemcpyAsync(g->d_boff, boff.data(), sizeof(int) * (n + nsub), hipMemcpyHostToDevice, c0->st));
        HCK(hipMemcpyAsync(g->d_pboff, pboff.data(), sizeof(int) * (n + nsub), hipMemcpyHostToDevice, c0->st));
        HCK(hipStreamSynchronize(c0->st));  // the host tables are reused
        return PAM_OK;
    };
    // step k of sub-group j on stream st
    auto launch_sub = [&](int t, int j, int k, hipStream_t st) -> int {
        const int o = first[j] + j, f = first[j];
        KCK(launch_step_multi(c0->phys, c0->abc == PAM_CPML, t, g->d_list + f, g->d_boff + o, first[j + 1] - f,
                              nblk[j], k, g->d_pboff + o, npblk[j], st));
        return PAM_OK;
    };
    // prime every member's shots: cleared fields and memory, sources, S_0
    auto prime = [&]() -> int {
        for (int i = 0; i < n; ++i) {
            pam_ctx *c = g->ctx[i];
            int r = zero_shot_state(c, nshots[i], true, true);
            if (r) return r;
            r = upload_sources(c, 0, nshots[i]);
            if (r) return r;
            KCK(launch_inject_point(c->fs.U0, c->plane, c->d_ssz, c->d_ssx, c->P, c->wav, 0, nshots[i], c->st));
        }
        // the group launches on member 0's stream (and its second lane)
        for (int i = 0; i < n; ++i) HCK(hipStreamSynchronize(g->ctx[i]->st));
        return PAM_OK;
    };
    // sub-group j runs on lane j % 2 of member 0 (one lane: member 0's stream)
    const int lanes = (c0->lanes > 1 && nsub > 1) ? 2 : 1;
    if (lanes > 1) {
        const int rl = ensure_lanes(c0);
        if (rl) return rl;
    }
    auto lane_stream = [&](int j) { return lanes > 1 ? c0->lane_st[j & 1] : c0->st; };
    auto fork = [&]() -> int { return lanes > 1 ? lanes_fork(c0) : PAM_OK; };
    auto join = [&](int r0) -> int { return lanes > 1 ? lanes_join(c0, r0) : r0; };
    if (tile < 0) {
        auto it = g->tuned.find(layout);
        if (it != g->tuned.end()) {
            tile = it->second;
        } else {  // time 12 steps of every sub-group with each variant, on the lanes of the
                  // run, twice (the first pass also warms the kernels and caches): the faster
            double best = 1e300;
            tile = kDefaultTile;
            for (int t = 0; t < tile_count(); ++t) {
                if (!tile_has_multi(t)) continue;
                int r = build(t);
                if (r) return r;
                for (int pass = 0; pass < 2; ++pass) {
                    r = prime();
                    if (r) return r;
                    HCK(hipEventRecord(c0->e0, c0->st));
                    r = fork();
                    if (r) return r;
                    for (int j = 0; j < nsub && !r; ++j)
                        for (int k = 0; k < 12 && !r; ++k) r = launch_sub(t, j, k, lane_stream(j));
                    r = join(r);
                    if (r) return r;
                    HCK(hipEventRecord(c0->e1, c0->st));
                    HCK(hipEventSynchronize(c0->e1));
                    float ms = 0.f;
                    HCK(hipEventElapsedTime(&ms, c0->e0, c0->e1));
                    if (ms < best) {
                        best = ms;
                        tile = t;
                    }
                }
            }
            g->tuned[layout] = tile;
        }
    }
    g->last_tile = tile;
    int r = prime();
    if (r) return r;
    for (int i = 0; i < n; ++i) {
        pam_ctx *c = g->ctx[i];
        const long long nd = (long long)(c->g.nt - c->g.itlag) * c->nrec;
        HCK(hipMemsetAsync(seis[i], 0, (size_t)nshots[i] * nd * sizeof(float), c0->st));
    }
    r = build(tile);
    if (r) return r;
    // each sub-group runs its nt steps on its lane, one launch per step (two
    // with the psi pass)
    HCK(hipEventRecord(c0->e0, c0->st));
    r = fork();
    if (r) return r;
    auto subgroups = [&]() -> int {
        for (int j = 0; j < nsub; ++j) {
            hipStream_t st = lane_stream(j);
            auto body = [&]() -> int {
                for (int k = 0; k < nt; ++k) {
                    int rr = launch_sub(tile, j, k, st);
                    if (rr) return rr;
                }
                return PAM_OK;
            };
            int rr;
            if (c0->graph) {
                std::string key = "grp|" + std::to_string(tile) + "|" + std::to_string(j) + "|";
                for (int i = first[j]; i < first[j + 1]; ++i) {
                    char buf[96];
                    std::snprintf(buf, sizeof buf, "%p:%u:%d:%p;", (void *)g->ctx[i], g->ctx[i]->epoch, nshots[i],
                                  (void *)seis[i]);
                    key += buf;
                }
                rr = run_graph_in(c0, g->graphs, key, body, []() {}, st);
            } else {
                rr = body();
            }
            if (rr) return rr;
        }
        return PAM_OK;
    };
    r = join(subgroups());
    if (r) return r;
    HCK(hipEventRecord(c0->e1, c0->st));
    HCK(hipEventSynchronize(c0->e1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, c0->e0, c0->e1));
    if (loop_ms) *loop_ms = ms;
    return PAM_OK;
}

extern "C" int pam_group_get_tile(const pam_group *g) { return g ? g->last_tile : -1; }

static int adjoint_run(pam_ctx *ctx, int shot0, int nshots, const float *data, bool grad, float *image, float *ilum,
                       float *g_eps, float *g_del, float *g_th);

// The scratch of the whole-run passes (snapshots, checkpoints, revolve slots,
// accumulators, staging) released once the context's streams are idle; the
// step-loop graphs that bake their addresses in go with them.  The next run
// allocates again.
struct ScratchBuf {
    void **p;
    long long *bytes;
};
constexpr int kScratchBufs = 13;
static void run_scratch_bufs(pam_ctx *ctx, ScratchBuf *b) {
    const ScratchBuf l[kScratchBufs] = {
        {(void **)&ctx->save, &ctx->save_bytes},         {(void **)&ctx->ckpt, &ctx->ckpt_bytes},
        {(void **)&ctx->rvbits, &ctx->rvbits_bytes},     {(void **)&ctx->smax, &ctx->smax_bytes},
        {(void **)&ctx->tmp, &ctx->tmp_bytes},           {(void **)&ctx->acc_ilum, &ctx->acc_bytes},
        {(void **)&ctx->sbd, &ctx->sbd_bytes},           {(void **)&ctx->acc_multi, &ctx->acc_multi_bytes},
        {(void **)&ctx->ilum_tmp, &ctx->ilum_tmp_bytes}, {(void **)&ctx->stage, &ctx->stage_bytes},
        {(void **)&ctx->ex_save, &ctx->ex_save_bytes},   {(void **)&ctx->ex_work, &ctx->ex_work_bytes},
        {(void **)&ctx->ex_gacc, &ctx->ex_gacc_bytes}};
    for (int i = 0; i < kScratchBufs; ++i) b[i] = l[i];
}
static void drop_graphs(pam_ctx *ctx) {
    for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}
static void release_run_scratch(pam_ctx *ctx) {
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    for (hipStream_t l : ctx->lane_st)
        if (l) (void)hipStreamSynchronize(l);
    drop_graphs(ctx);
    ScratchBuf bufs[kScratchBufs];
    run_scratch_bufs(ctx, bufs);
    for (ScratchBuf &b : bufs) {
        dfree(*b.p);
        *b.p = nullptr;
        *b.bytes = 0;
    }
}
// The run scratch of `from` (idle) and of `to` trade places, so that `to`
// starts with the buffers `from` sized, without a free (hipFree synchronizes
// the whole device); the graphs of both, which may read either's buffers, are
// dropped.
static void swap_run_scratch(pam_ctx *from, pam_ctx *to) {
    drop_graphs(from);
    drop_graphs(to);
    ScratchBuf a[kScratchBufs], b[kScratchBufs];
    run_scratch_bufs(from, a);
    run_scratch_bufs(to, b);
    for (int i = 0; i < kScratchBufs; ++i) {
        std::swap(*a[i].p, *b[i].p);
        std::swap(*a[i].bytes, *b[i].bytes);
    }
}

// the regions of a whole state's fields (the two wavefields; psi_x, psi_z,
// zeta_x, zeta_z on their strips), in the order of the revolve slots
static void state_regions(const pam_ctx *ctx, FieldRegion *reg) {
    reg[0] = reg[1] = region_all(ctx->nzp, ctx->P);
    if (ctx->abc == PAM_CPML) {
        reg[2] = reg[4] = region_xstrips(ctx->nzp, ctx->nxp, ctx->P, ctx->g.n_abc);
        reg[3] = reg[5] = region_zstrips(ctx->nzp, ctx->P, ctx->g.n_abc);
    }
}

// FWI gradients of every member (pam_gradient's g_m, with each member's own
// storage, batch and lane options).  Two host workers take the members, the
// largest first, each on a stream of its own, so that two members' passes
// share the device (a small member alone leaves most of it idle); a worker
// hands its run scratch from one member to the next rather than freeing it
// (a free would synchronize the device and so the other worker), and every
// member's scratch is released at the end: the group needs about two members'
// scratch at a time, not the sum.  Each member's gradient is computed as when
// it runs alone.  g_norm2[i] = sum of g_m[i]^2 over the interior in float64
// (two passes in a fixed order).
extern "C" int pam_group_gradient(pam_group *g, const int *nshots, const float *const *residual, float *const *g_m,
                                  double *g_norm2, double *loop_ms) {
    if (!g || !nshots || !residual || !g_m) return PAM_E_ARG;
    pam_ctx *c0 = g->ctx[0];
    pam_ctx *const ctx = c0;  // errors of the group go to member 0
    const int n = (int)g->ctx.size();
    for (int i = 0; i < n; ++i) {
        pam_ctx *c = g->ctx[i];
        int r = ready(c);
        if (r) return fail(c0, r, "group member %d: %s", i, c->err.c_str());
        if (nshots[i] < 1 || nshots[i] > c->nshot)
            return fail(c0, PAM_E_ARG, "group member %d: nshots %d outside [1, shots]", i, nshots[i]);
        if (!residual[i] || !g_m[i]) return fail(c0, PAM_E_ARG, "group member %d: residual or g_m is NULL", i);
    }
    int dev = 0;
    HCK(hipGetDevice(&dev));
    const int nw = n > 1 ? 2 : 1;
    for (int w = 0; w < nw; ++w)
        if (!g->wst[w]) HCK(hipStreamCreate(&g->wst[w]));
    if (!g->wev) HCK(hipEventCreateWithFlags(&g->wev, hipEventDisableTiming));
    // the workers' streams start after the work already queued on every member's stream
    for (int i = 0; i < n; ++i) {
        HCK(hipEventRecord(g->wev, g->ctx[i]->st));
        for (int w = 0; w < nw; ++w) HCK(hipStreamWaitEvent(g->wst[w], g->wev, 0));
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (long long)g->ctx[x]->plane * nshots[x] > (long long)g->ctx[y]->plane * nshots[y];
    });
    std::atomic<int> next(0);
    std::atomic<bool> failed(false);
    std::vector<int> rc(n, PAM_OK);
    auto worker = [&](int w) {
        (void)hipSetDevice(dev);
        pam_ctx *held = nullptr;  // the member whose scratch this worker carries
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= n || failed.load()) break;
            const int i = order[k];
            pam_ctx *c = g->ctx[i];
            if (held) swap_run_scratch(held, c);
            hipStream_t own = c->st;
            c->st = g->wst[w];
            int r = adjoint_run(c, 0, nshots[i], residual[i], true, g_m[i], nullptr, nullptr, nullptr, nullptr);
            if (!r && g_norm2) {
                if (launch_sumsq(g_m[i], (long long)c->g.nz * c->g.nx, c->d_acc, c->d_acc + kMisfitParts, c->st))
                    r = fail(c, PAM_E_HIP, "gradient norm: %s", hipGetErrorString(hipGetLastError()));
                if (!r && hipMemcpyAsync(&g_norm2[i], c->d_acc + kMisfitParts, sizeof(double), hipMemcpyDeviceToHost,
                                         c->st) != hipSuccess)
                    r = fail(c, PAM_E_HIP, "gradient norm copy");
            }
            if (hipStreamSynchronize(c->st) != hipSuccess && !r) r = fail(c, PAM_E_HIP, "member stream");
            c->st = own;
            rc[i] = r;
            held = c;
            if (r) {
                failed.store(true);
                break;
            }
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    if (nw > 1) {
        std::thread other(worker, 1);
        worker(0);
        other.join();
    } else {
        worker(0);
    }
    const auto t1 = std::chrono::steady_clock::now();
    for (int w = 0; w < nw; ++w) {
        HCK(hipEventRecord(g->wev, g->wst[w]));
        HCK(hipStreamWaitEvent(c0->st, g->wev, 0));
    }
    for (int i = 0; i < n; ++i) release_run_scratch(g->ctx[i]);
    for (int i = 0; i < n; ++i)
        if (rc[i]) return fail(c0, rc[i], "group member %d: %s", i, g->ctx[i]->err.c_str());
    if (loop_ms) *loop_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return PAM_OK;
}

extern "C" int pam_last_timing(const pam_ctx *ctx, double *loop_ms, double *kernel_ms, long long *kernel_launches) {
    if (!ctx) return PAM_E_ARG;
    if (loop_ms) *loop_ms = ctx->loop_ms;
    if (kernel_ms) *kernel_ms = ctx->kernel_ms;
    if (kernel_launches) *kernel_launches = ctx->launches;
    return PAM_OK;
}

// ------------------------------------------------------------ adjoint passes

static int alloc_fieldset(pam_ctx *ctx, FieldSet &f) {
    const size_t fb = (size_t)ctx->plane * ctx->max_shots * sizeof(float);
    DALLOC(f.U0, fb);
    DALLOC(f.U1, fb);
    if (ctx->abc == PAM_CPML) {
        DALLOC(f.psix, 2 * fb);
        DALLOC(f.psiz, 2 * fb);
        DALLOC(f.zetax, fb);
        DALLOC(f.zetaz, fb);
        HCK(hipMemsetAsync(f.psix, 0, 2 * fb, ctx->st));
        HCK(hipMemsetAsync(f.psiz, 0, 2 * fb, ctx->st));
    }
    return PAM_OK;
}

// fields held by a checkpoint: the two wavefields, and for CPML the previous
// step's psi (x, z) and zeta (x, z)
static int ckpt_fields(const pam_ctx *ctx) { return ctx->abc == PAM_CPML ? 6 : 2; }

// copy (cur, prv, CPML state) of nb shots of the main set into checkpoint j
static int save_ckpt(pam_ctx *ctx, int j, int nb, const float *cur, const float *prv) {
    const size_t fb = (size_t)ctx->plane * nb * sizeof(float);
    const long long fl = ctx->plane * nb;  // floats per checkpointed field
    float *c = ctx->ckpt + (long long)j * ckpt_fields(ctx) * fl;
    HCK(hipMemcpyAsync(c, cur, fb, hipMemcpyDeviceToDevice, ctx->st));
    HCK(hipMemcpyAsync(c + fl, prv, fb, hipMemcpyDeviceToDevice, ctx->st));
    if (ctx->abc == PAM_CPML) {
        const long long in = ctx->fs.psi_flip ? ctx->psi_half : 0;
        HCK(hipMemcpyAsync(c + 2 * fl, ctx->fs.psix + in, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(c + 3 * fl, ctx->fs.psiz + in, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(c + 4 * fl, ctx->fs.zetax, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(c + 5 * fl, ctx->fs.zetaz, fb, hipMemcpyDeviceToDevice, ctx->st));
    }
    ctx->ckpt_flip[j] = ctx->fs.psi_flip;
    return PAM_OK;
}

// restore checkpoint j into the recompute set (U0 = cur, U1 = prv)
static int restore_ckpt(pam_ctx *ctx, int j, int nb) {
    const size_t fb = (size_t)ctx->plane * nb * sizeof(float);
    const long long fl = ctx->plane * nb;
    const float *c = ctx->ckpt + (long long)j * ckpt_fields(ctx) * fl;
    FieldSet &f = ctx->aux;
    HCK(hipMemcpyAsync(f.U0, c, fb, hipMemcpyDeviceToDevice, ctx->st));
    HCK(hipMemcpyAsync(f.U1, c + fl, fb, hipMemcpyDeviceToDevice, ctx->st));
    if (ctx->abc == PAM_CPML) {
        f.psi_flip = ctx->ckpt_flip[j];
        const long long in = f.psi_flip ? ctx->psi_half : 0;
        HCK(hipMemcpyAsync(f.psix + in, c + 2 * fl, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(f.psiz + in, c + 3 * fl, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(f.zetax, c + 4 * fl, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(f.zetaz, c + 5 * fl, fb, hipMemcpyDeviceToDevice, ctx->st));
    }
    return PAM_OK;
}

// Binomial checkpointing (Griewank's treeverse): beta(s, r) = C(s + r, s) is
// the longest chain that s slots reverse with no step advanced more than r
// times.
static double revolve_beta(int s, int r) {
    double b = 1.0;
    for (int i = 1; i <= r; ++i) b = b * (s + i) / i;
    return b;
}
// First checkpoint of an n-step range (n >= 2) with s >= 1 free slots: for
// the smallest r with beta(s, r) >= n, the offset m <= beta(s, r - 1) (the left
// part is then reversed with r - 1 more sweeps) with n - m <= beta(s - 1, r)
// (the right part with one slot fewer), which beta(s, r) = beta(s, r - 1) +
// beta(s - 1, r) guarantees.
static int revolve_split(int n, int s) {
    int r = 1;
    while (revolve_beta(s, r) < n) ++r;
    const double m = std::min(revolve_beta(s, r - 1), (double)(n - 1));
    return std::max(1, (int)m);
}

// ---- uniform checkpoints on two lanes
// The settings of one uniform-checkpoint run that every batch shares.
struct CkptRun {
    int shot0 = 0;
    const float *data = nullptr;
    long long nd = 0, nint = 0, per_step = 0;
    long long acc_st = 0;  // floats between shots' accumulators
    int acc_pad = 0;       // accumulators and snapshots padded
    std::function<int(const float *, int, float *, hipStream_t)> reduce;  // accumulators into an output
    bool grad = false, d2t = false;
    int tile = 0, C = 0, nseg = 0, t_lo = 0, t_hi = 0, nbuf = 0, skind = 0;
    float *image = nullptr, *ilum = nullptr;
    // save-all runs on the lanes: every step's snapshot (elem bytes per element)
    // instead of checkpoints; half: 16-bit snapshots scaled by a max found
    // step by step (smax0: the bound of the first step's scale)
    bool saveall = false, half = false;
    int elem = 4;
    float smax0 = 0.0f;
    std::vector<hipEvent_t> red_ev;   // reduction of batch i done
    std::vector<char> red_issued;     // batch i's reduction enqueued
};

// One batch of a uniform-checkpoint run on one lane, issued a time step at a
// time (ckpt_tick) so that one host thread can keep two lanes' queues filled:
// the forward pass with illumination saving the checkpoints, then segment by
// segment (the last first) the recompute storing S_k or D_k and the adjoint
// steps over it, then the reduction of the batch's accumulators, enqueued
// after the previous batch's so that the sums are those of one lane.
struct CkptLane {
    FieldSet F, A;  // the lane's views of the two field sets
    float *save = nullptr, *acc_ilum = nullptr, *acc_img = nullptr, *ckpt = nullptr;
    unsigned int *smax = nullptr;  // the lane's snapshot scales (16-bit save-all)
    std::vector<int> flip;  // psi buffer of each checkpoint
    hipStream_t st = nullptr;
    int slot0 = 0, Bl = 0;
    // the batch in flight
    int batch = -1, b0 = 0, nb = 0, phase = 0;
    int k = 0, j = 0, seg = 0, t = 0, lo = 0, hi = 0, k0 = 0, k1 = 0;
    StepArgs a{}, b{}, f{};
    float *cur = nullptr, *prv = nullptr, *fc = nullptr, *fp = nullptr;
};

static FieldSet fieldset_view(const pam_ctx *ctx, const FieldSet &f, int slot0) {
    FieldSet v = f;
    const long long off = (long long)slot0 * ctx->plane;
    v.U0 += off;
    v.U1 += off;
    if (ctx->abc == PAM_CPML) {
        v.psix += off;
        v.psiz += off;
        v.zetax += off;
        v.zetaz += off;
    }
    return v;
}

// save_ckpt / restore_ckpt on a lane
static int ckpt_lane_save(pam_ctx *ctx, CkptLane &L, int j) {
    const size_t fb = (size_t)ctx->plane * L.nb * sizeof(float);
    const long long fl = ctx->plane * L.Bl;  // a lane's checkpoint fields hold Bl shots
    float *c = L.ckpt + (long long)j * ckpt_fields(ctx) * fl;
    HCK(hipMemcpyAsync(c, L.cur, fb, hipMemcpyDeviceToDevice, L.st));
    HCK(hipMemcpyAsync(c + fl, L.prv, fb, hipMemcpyDeviceToDevice, L.st));
    if (ctx->abc == PAM_CPML) {
        const long long in = L.F.psi_flip ? ctx->psi_half : 0;
        HCK(hipMemcpyAsync(c + 2 * fl, L.F.psix + in, fb, hipMemcpyDeviceToDevice, L.st));
        HCK(hipMemcpyAsync(c + 3 * fl, L.F.psiz + in, fb, hipMemcpyDeviceToDevice, L.st));
        HCK(hipMemcpyAsync(c + 4 * fl, L.F.zetax, fb, hipMemcpyDeviceToDevice, L.st));
        HCK(hipMemcpyAsync(c + 5 * fl, L.F.zetaz, fb, hipMemcpyDeviceToDevice, L.st));
    }
    L.flip[j] = L.F.psi_flip;
    return PAM_OK;
}
static int ckpt_lane_restore(pam_ctx *ctx, CkptLane &L, int j) {
    const size_t fb = (size_t)ctx->plane * L.nb * sizeof(float);
    const long long fl = ctx->plane * L.Bl;
    const float *c = L.ckpt + (long long)j * ckpt_fields(ctx) * fl;
    HCK(hipMemcpyAsync(L.A.U0, c, fb, hipMemcpyDeviceToDevice, L.st));
    HCK(hipMemcpyAsync(L.A.U1, c + fl, fb, hipMemcpyDeviceToDevice, L.st));
    if (ctx->abc == PAM_CPML) {
        L.A.psi_flip = L.flip[j];
        const long long in = L.A.psi_flip ? ctx->psi_half : 0;
        HCK(hipMemcpyAsync(L.A.psix + in, c + 2 * fl, fb, hipMemcpyDeviceToDevice, L.st));
        HCK(hipMemcpyAsync(L.A.psiz + in, c + 3 * fl, fb, hipMemcpyDeviceToDevice, L.st));
        HCK(hipMemcpyAsync(L.A.zetax, c + 4 * fl, fb, hipMemcpyDeviceToDevice, L.st));
        HCK(hipMemcpyAsync(L.A.zetaz, c + 5 * fl, fb, hipMemcpyDeviceToDevice, L.st));
    }
    return PAM_OK;
}

// Issues the next piece of lane L's batch (one time step, or the batch's set-up
// or reduction); *done once the reduction is enqueued.  The reduction waits
// (returns without issuing) until the previous batch's is enqueued.
static int ckpt_tick(pam_ctx *ctx, CkptRun &R, CkptLane &L, bool *done) {
    const pam_grid &g = ctx->g;
    const int nt = g.nt;
    *done = false;
    switch (L.phase) {
        case 0: {  // forward pass set-up
            int r = zero_fields(ctx, L.F, L.nb, true, true, L.st);
            if (r) return r;
            r = upload_sources(ctx, R.shot0 + L.b0, L.nb, L.slot0, L.st);
            if (r) return r;
            HCK(hipMemsetAsync(L.acc_ilum, 0, (size_t)L.nb * R.acc_st * 4, L.st));
            HCK(hipMemsetAsync(L.acc_img, 0, (size_t)L.nb * R.acc_st * 4, L.st));
            KCK(launch_inject_point(L.F.U0, ctx->plane, ctx->d_ssz + L.slot0, ctx->d_ssx + L.slot0, ctx->P, ctx->wav,
                                    0, L.nb, L.st));
            base_step_args(ctx, L.a, L.nb);
            L.a.flags |= F_SRC | F_ILUM;
            if (R.saveall) {
                // every step's snapshot (S_k, or D_k for a gradient), 16-bit ones scaled
                // by the max the step before finds
                L.a.flags |= F_SAVE | (R.d2t ? F_D2T : 0) | (R.half ? F_SMAX : 0);
                if (R.half) {
                    HCK(hipMemsetAsync(L.smax, 0, (size_t)L.nb * nt * 4 * kSmaxSlots, L.st));
                    KCK(launch_smax0(L.smax, nt, ctx->d_ssz + L.slot0, ctx->d_ssx + L.slot0, g.n_abc, ctx->nzp,
                                     ctx->nxp, ctx->wav, L.nb, R.smax0, L.st));
                    L.a.smax = L.smax;
                    L.a.save_scale = reinterpret_cast<const float *>(L.smax);
                }
            }
            L.a.save_kind = R.skind;
            L.a.save = L.save;
            L.a.save_k0 = 0;
            L.a.save_kstride = R.per_step;
            L.a.save_sstride = R.per_step * R.nbuf;
            L.a.ilum = L.acc_ilum;
            L.a.image = L.acc_img;
            L.a.acc_sstride = R.acc_st;
            L.a.acc_pad = R.acc_pad;
            L.a.ssz = ctx->d_ssz + L.slot0;
            L.a.ssx = ctx->d_ssx + L.slot0;
            L.cur = L.F.U0;
            L.prv = L.F.U1;
            L.k = 0;
            L.j = 0;
            L.phase = 1;
            return PAM_OK;
        }
        case 1: {  // forward step k, checkpoint j first when it starts a segment
            if (!R.saveall && L.j < R.nseg && L.k == std::max(0, L.j * R.C - 1)) {
                int r = ckpt_lane_save(ctx, L, L.j);
                if (r) return r;
                ++L.j;
            }
            int r = fwd_step_fs(ctx, L.F, L.a, R.tile, L.cur, L.prv, L.k, nullptr, L.st);
            if (r) return r;
            std::swap(L.cur, L.prv);
            if (++L.k < nt) return PAM_OK;
            // backward pass set-up: the adjoint field reuses the lane's main set
            r = zero_fields(ctx, L.F, L.nb, true, ctx->reset_adj_cpml != 0, L.st);
            if (r) return r;
            if (R.t_hi >= g.itlag)
                KCK(launch_inject_rec(L.F.U0, ctx->plane, ctx->P, ctx->d_rz, ctx->d_rx, ctx->nrec,
                                      R.data + (long long)L.b0 * R.nd, R.nd, R.t_hi - g.itlag, g.dx * g.dz, L.nb,
                                      L.st));
            L.cur = L.F.U0;
            L.prv = L.F.U1;
            base_step_args(ctx, L.b, L.nb);
            L.b.flags |= F_RECINJ | (R.grad ? F_GRADM : F_IMAGE) | (R.d2t ? F_D2T : 0);
            L.b.data = R.data + (long long)L.b0 * R.nd;
            L.b.dstride = R.nd;
            L.b.save_kind = R.skind;
            L.b.save = L.save;
            L.b.save_kstride = R.per_step;
            L.b.save_sstride = R.per_step * R.nbuf;
            L.b.image = L.acc_img;
            L.b.acc_sstride = R.acc_st;
            L.b.acc_pad = R.acc_pad;
            L.b.ssz = ctx->d_ssz + L.slot0;
            L.b.ssx = ctx->d_ssx + L.slot0;
            if (R.saveall) {  // the snapshots of every step are in place: one backward sweep
                L.b.save_k0 = 0;
                L.b.save_scale = reinterpret_cast<const float *>(L.smax);
                L.seg = 0;
                L.lo = R.t_lo;
                L.hi = R.t_hi;
                L.t = R.t_hi;
                L.phase = 4;
                return PAM_OK;
            }
            L.seg = R.nseg - 1;
            L.phase = 2;
            return PAM_OK;
        }
        case 2: {  // segment start: restore its checkpoint
            if (L.seg < 0) {
                L.phase = 5;
                return PAM_OK;
            }
            L.lo = std::max(R.t_lo, L.seg * R.C);
            L.hi = std::min(R.t_hi, (L.seg + 1) * R.C - 1);
            if (L.lo > L.hi) {
                --L.seg;
                return PAM_OK;
            }
            L.k0 = std::max(0, L.seg * R.C - 1);
            L.k1 = std::min(nt - 1, L.hi + 1);
            int r = ckpt_lane_restore(ctx, L, L.seg);
            if (r) return r;
            L.f = L.a;
            L.f.flags = (L.a.flags & (F_CERJAN | F_SRC)) | F_SAVE | (R.d2t ? F_D2T : 0);
            L.f.save_k0 = L.k0;
            L.fc = L.A.U0;
            L.fp = L.A.U1;
            L.k = L.k0;
            L.phase = 3;
            return PAM_OK;
        }
        case 3: {  // recompute step k of the segment, storing S_k or D_k
            int r = fwd_step_fs(ctx, L.A, L.f, R.tile, L.fc, L.fp, L.k, nullptr, L.st);
            if (r) return r;
            std::swap(L.fc, L.fp);
            if (++L.k <= L.k1) return PAM_OK;
            L.b.save_k0 = L.k0;
            L.t = L.hi;
            L.phase = 4;
            return PAM_OK;
        }
        case 4: {  // adjoint step t
            int r = fwd_step_fs(ctx, L.F, L.b, R.tile, L.cur, L.prv, L.t, nullptr, L.st);
            if (r) return r;
            std::swap(L.cur, L.prv);
            if (--L.t < L.lo) {
                --L.seg;
                L.phase = 2;
            }
            return PAM_OK;
        }
        case 5: {  // the batch's sums, after the previous batch's
            if (L.batch > 0) {
                if (!R.red_issued[L.batch - 1]) return PAM_OK;
                HCK(hipStreamWaitEvent(L.st, R.red_ev[L.batch - 1], 0));
            }
            int r = R.reduce(L.acc_img, L.nb, R.image, L.st);
            if (r) return r;
            r = R.reduce(L.acc_ilum, L.nb, R.ilum, L.st);
            if (r) return r;
            HCK(hipEventRecord(R.red_ev[L.batch], L.st));
            R.red_issued[L.batch] = 1;
            *done = true;
            return PAM_OK;
        }
        default:
            return fail(ctx, PAM_E_STATE, "checkpoint lane in an unknown phase");
    }
}

// The batches of a uniform-checkpoint run (sizes `sizes`, in shot order) on two
// lanes of at most Bl shots each, interleaved a time step at a time; the lanes
// start after the work queued on the context's stream, which waits for both.
static int ckpt_two_lanes(pam_ctx *ctx, CkptRun &R, const std::vector<int> &sizes, int Bl) {
    int r = ensure_lanes(ctx);
    if (r) return r;
    const int nbat = (int)sizes.size();
    R.red_ev.assign(nbat, nullptr);
    R.red_issued.assign(nbat, 0);
    for (auto &e : R.red_ev) HCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    CkptLane L[2];
    for (int l = 0; l < 2; ++l) {
        L[l].slot0 = l * Bl;
        L[l].Bl = Bl;
        L[l].F = fieldset_view(ctx, ctx->fs, l * Bl);
        if (!R.saveall) {
            L[l].A = fieldset_view(ctx, ctx->aux, l * Bl);
            L[l].ckpt = ctx->ckpt + (long long)l * R.nseg * ckpt_fields(ctx) * ctx->plane * Bl;
        }
        // the lane's share of the snapshot buffer (R.elem bytes per element)
        L[l].save = reinterpret_cast<float *>(static_cast<char *>(ctx->save) +
                                              (long long)l * Bl * R.per_step * R.nbuf * R.elem);
        if (R.half) L[l].smax = ctx->smax + (long long)l * Bl * ctx->g.nt * kSmaxSlots;
        L[l].acc_ilum = ctx->acc_ilum + (long long)l * Bl * R.acc_st;
        L[l].acc_img = ctx->acc_img + (long long)l * Bl * R.acc_st;
        L[l].flip.assign(R.nseg, 0);
        L[l].st = ctx->lane_st[l];
    }
    r = lanes_fork(ctx);
    if (r) return r;
    auto body = [&]() -> int {
        int next = 0, b0 = 0;
        auto start = [&](CkptLane &X) {
            X.batch = next;
            X.b0 = b0;
            X.nb = sizes[next];
            X.phase = 0;
            b0 += sizes[next];
            ++next;
        };
        bool active[2] = {false, false};
        for (int l = 0; l < 2 && next < nbat; ++l) {
            start(L[l]);
            active[l] = true;
        }
        while (active[0] || active[1]) {
            for (int l = 0; l < 2; ++l) {
                if (!active[l]) continue;
                bool done = false;
                int rr = ckpt_tick(ctx, R, L[l], &done);
                if (rr) return rr;
                if (!done) continue;
                if (next < nbat)
                    start(L[l]);
                else
                    active[l] = false;
            }
        }
        return PAM_OK;
    };
    r = lanes_join(ctx, body());
    for (auto &e : R.red_ev)
        if (e) (void)hipEventDestroy(e);
    R.red_ev.clear();
    return r;
}

// RTM (imaging) or FWI gradient (F_GRADM) over shots.  The source field is
// kept by one of the strategies of enum pam_storage:
//  - every step in float32 (src/Migration.py:996-1037);
//  - every step in float16 / bfloat16 scaled by the snapshot's max |S_k| over
//    the interior, found by the step that produces the field (F_SMAX);
//  - uniform checkpoints of the whole state every C steps in HBM, each segment
//    recomputed into a (C+2)-step buffer on a second field set just before
//    the adjoint pass crosses it (src/Migration.py:1114-1263 keeps them in host
//    files); the recomputed fields are bit-identical to the first pass;
//  - boundary saving: the 4-cell rims inside the interior edge of every step
//    (float32, or float16 / bfloat16 scaled per step), the interior rebuilt
//    backwards in time on the second field set one step ahead of the adjoint
//    pass (src/Migration.py:1266-1410; the reference's time alignment is
//    re-derived: S_{k-1} = v^2 dt^2 L(S_k) + 2 S_k - S_{k+1} + wav[k+1] at the
//    source, rims written back, zero on the absorbing layer);
//  - binomial checkpointing: the whole state in a fixed number of slots
//    (float32 or float16), the forward pass recomputed on the second field set
//    in treeverse order; each adjoint step's S_t (and D_t) comes from one
//    recomputed forward step.
static int adjoint_run(pam_ctx *ctx, int shot0, int nshots, const float *data, bool grad, float *image, float *ilum,
                       float *g_eps, float *g_del, float *g_th) {
    int r = ready(ctx);
    if (r) return r;
    ctx->adj_next_t = -2;  // the second field set is reused
    if (nshots < 1) return fail(ctx, PAM_E_ARG, "nshots must be >= 1");
    if (!data || !image) return fail(ctx, PAM_E_ARG, "data and image must not be NULL");
    const pam_grid &g = ctx->g;
    const int nt = g.nt;
    if (nt < 3) return fail(ctx, PAM_E_ARG, "the adjoint pass needs nt >= 3");
    const long long nd = (long long)(nt - g.itlag) * ctx->nrec;
    const long long nint = (long long)g.nz * g.nx;
    const bool linear_adj = (ctx->phys >= 1 && ctx->abc == PAM_CERJAN);
    const int storage = ctx->storage;
    const bool ckpt = storage == PAM_STORE_CHECKPOINT;
    const bool half = storage == PAM_STORE_ALL_F16 || storage == PAM_STORE_ALL_BF16;
    const bool sb = storage >= PAM_STORE_BOUNDARY_F32 && storage <= PAM_STORE_BOUNDARY_BF16;
    const bool rv = storage == PAM_STORE_REVOLVE_F32 || storage == PAM_STORE_REVOLVE_F16;
    const bool rv_half = storage == PAM_STORE_REVOLVE_F16;
    const bool rbc = storage == PAM_STORE_RANDOM_BOUNDARY;
    const bool cpml = ctx->abc == PAM_CPML;
    const int sb_kind = sb ? storage - PAM_STORE_BOUNDARY_F32 : 0;  // 0 float32, 1 float16, 2 bfloat16
    if (sb && (g.nz < 9 || g.nx < 9 || g.n_abc < 4))
        return fail(ctx, PAM_E_ARG, "boundary saving needs nz, nx >= 9 and n_abc >= 4");
    if (half && linear_adj)
        return fail(ctx, PAM_E_UNSUPPORTED,
                    "float16 / bfloat16 snapshots feed the pressure-form adjoint only; the linearised VTI/TTI "
                    "adjoint (Cerjan) reads float32 padded snapshots: use PAM_STORE_ALL_F32 or PAM_STORE_CHECKPOINT");
    if (half && grad && (g_eps || g_del || g_th))
        return fail(ctx, PAM_E_UNSUPPORTED,
                    "the eps / delta / theta partials read float32 snapshots of S_t: use PAM_STORE_ALL_F32, "
                    "PAM_STORE_CHECKPOINT or a revolve / boundary strategy with them, not float16 / bfloat16 "
                    "save-all snapshots");
    // boundary saving: the adjoint step reads D_t from a one-step interior
    // buffer (gradient) or S_t from the rebuilt padded field (imaging)
    int skind = (sb || rbc) ? (grad ? SV_F32_INT : SV_F32_PAD)
                         : linear_adj ? SV_F32_PAD
                         : rv ? SV_F32_INT
                                      : (storage == PAM_STORE_ALL_F16
                                             ? SV_F16_INT
                                             : (storage == PAM_STORE_ALL_BF16 ? SV_BF16_INT : SV_F32_INT));
    const int elem = (half || sb_kind) ? 2 : 4;
    const bool multi = grad && (g_eps || g_del || g_th);
    // save-all (float32, float16 or bfloat16) or checkpoints with the
    // pressure-form adjoint and no partials: accumulators and snapshots are
    // sized as padded planes, and when the kernel variant's passes run the
    // 16-row tile family they are laid out padded (rows 256-byte aligned, four
    // cells per thread in one access; unaligned interior rows took twice as
    // long to update); other variants index them as interior arrays at the
    // same strides
    const bool pad_ok = !linear_adj && !multi && (storage == PAM_STORE_ALL_F32 || half || ckpt);
    const long long acc_st = pad_ok ? ctx->plane : nint;  // floats between shots' accumulators
    const long long per_step = rbc ? 1 : sb ? sb_width(g.nz, g.nx) : (skind == SV_F32_PAD || pad_ok) ? ctx->plane : nint;  // per step
    // default interval 2 ceil(sqrt(nt)): every step is recomputed once whatever
    // the interval, so a longer one only trades checkpoint copies for segment
    // buffer (2% faster than ceil(sqrt(nt)) on a 451x1801 VTI grid, nt 5099)
    const int C = ckpt ? (ctx->ckpt_step > 0 ? ctx->ckpt_step
                                             : std::min(nt, 2 * (int)std::ceil(std::sqrt((double)nt))))
                       : nt;
    // backward steps t in [t_lo, t_hi]: [1, nt-2] (Python reference) or [0, nt-1] (C++)
    if (grad && ctx->adj_range == 1)
        return fail(ctx, PAM_E_ARG, "the FWI gradient needs the adjoint range nt-2 .. 1 (PAM_OPT_ADJ_RANGE 0)");
    const int t_hi = ctx->adj_range ? nt - 1 : nt - 2, t_lo = ctx->adj_range ? 0 : 1;
    const int nseg = ckpt ? t_hi / C + 1 : 1;  // segments of backward steps
    const int nbuf = (rv || rbc) ? 1 : ckpt ? C + 2 : nt;  // snapshots held at once
    const int nslot = rv ? std::max(1, std::min(ctx->snapshots, t_hi - t_lo + 1)) : 0;  // revolve slots
    const int nmulti = (ctx->phys == 2) ? 3 : 2;
    // the gradient needs S only through S_{t+1} - 2 S_t + S_{t-1}: store that
    // difference (formed by the forward step in float32, the same arithmetic the
    // adjoint step would do) unless the partials or the linearised adjoint need S
    const bool d2t = grad && (sb || rv || rbc || (!multi && !linear_adj));
    // choose the batch so that the snapshots (and checkpoints) fit in free memory
    size_t freeb = 0, totalb = 0;
    HCK(hipMemGetInfo(&freeb, &totalb));
    int B = batch_size(ctx, nshots);
    if (ctx->batch <= 0) {
        // besides the two fields, every shot carries its accumulators and a
        // snapshot slot per step through these passes: 5/6 of a forward batch
        // of 8 or more keeps the working set in cache (13 shots per batch ran
        // 5% faster than 16 on a 451x1801 VTI grid), spread evenly over the
        // batches
        const int fb = auto_batch(ctx);
        B = std::max(1, std::min(fb >= 8 ? fb * 5 / 6 : fb, std::min(nshots, ctx->max_shots)));
        const int nbat = (nshots + B - 1) / B;
        B = (nshots + nbat - 1) / nbat;
    }
    // uniform checkpoints with the pressure-form adjoint and no partials:
    // batches of Bl shots alternate between two lanes, as in pam_forward (an
    // explicit batch is the per-lane one); the buffers hold both lanes' shots
    // (and save-all runs in float32, float16 or bfloat16 on the same lanes)
    const bool saveall = storage == PAM_STORE_ALL_F32 || half;
    bool two = (ckpt || saveall) && !multi && !linear_adj && ctx->lanes == 2 && nshots > 1 && ctx->max_shots >= 2;
    int Bl = 0;
    if (two) {
        Bl = ctx->batch > 0 ? ctx->batch : (B + 1) / 2;
        Bl = std::max(1, std::min(Bl, std::min(nshots, ctx->max_shots / 2)));
        B = 2 * Bl;
    }
    const long long per_shot_bytes = per_step * nbuf * elem + 2 * acc_st * 4 +
                                     (multi ? nmulti * nint * 4 : 0) + (half ? (long long)nt * 4 * kSmaxSlots : 0) +
                                     (ckpt ? (long long)nseg * ckpt_fields(ctx) * ctx->plane * 4 : 0) +
                                     (sb ? (long long)nt * 4 + nint * 4 : 0) +
                                     (rv ? (long long)nslot * ckpt_fields(ctx) * ctx->plane * (rv_half ? 2 : 4) +
                                               nint * 4 : 0) +
                                     (rbc ? nint * 4 : 0);
    const long long have = (long long)ctx->save_bytes + ctx->acc_bytes + ctx->tmp_bytes + ctx->acc_multi_bytes +
                           ctx->ckpt_bytes + ctx->smax_bytes + ctx->sbd_bytes + ctx->rvbits_bytes;
    long long budget = (long long)((freeb + have) * 0.85);
    if ((ckpt || sb || rv || rbc) && !ctx->aux.U0)
        budget -= (long long)ctx->plane * ctx->max_shots * 4 * (ctx->abc == PAM_CPML ? 8 : 2);
    while (B > 1 && (long long)B * per_shot_bytes > budget) --B;
    if ((long long)B * per_shot_bytes > budget)
        return fail(ctx, PAM_E_NOMEM, "one shot's snapshots (%lld bytes) exceed free device memory", per_shot_bytes);
    if (two) {
        Bl = std::min(Bl, B / 2);
        two = Bl >= 1;
        if (two) B = 2 * Bl;
    }
    r = ensure_scratch(ctx, &ctx->save, &ctx->save_bytes, (long long)B * per_step * nbuf * elem);
    if (r) return r;
    r = ensure_scratch(ctx, (void **)&ctx->acc_ilum, &ctx->acc_bytes, (long long)B * acc_st * 4 * 2);
    if (r) return r;
    ctx->acc_img = ctx->acc_ilum + (long long)B * acc_st;
    if (half || sb) {
        // save-all: kSmaxSlots words per shot and step; boundary saving: one
        r = ensure_scratch(ctx, (void **)&ctx->smax, &ctx->smax_bytes, (long long)B * nt * 4 * (half ? kSmaxSlots : 1));
        if (r) return r;
    }
    if (sb || rbc) {
        r = ensure_scratch(ctx, (void **)&ctx->sbd, &ctx->sbd_bytes, (long long)B * nint * 4);
        if (r) return r;
        r = alloc_fieldset(ctx, ctx->aux);
        if (r) return r;
    }
    if (rv) {
        r = ensure_scratch(ctx, (void **)&ctx->ckpt, &ctx->ckpt_bytes,
                           (long long)nslot * ckpt_fields(ctx) * ctx->plane * B * (rv_half ? 2 : 4));
        if (r) return r;
        if (grad) {
            r = ensure_scratch(ctx, (void **)&ctx->sbd, &ctx->sbd_bytes, (long long)B * nint * 4);
            if (r) return r;
        }
        if (rv_half) {
            r = ensure_scratch(ctx, (void **)&ctx->rvbits, &ctx->rvbits_bytes,
                               (long long)nslot * ckpt_fields(ctx) * B * 4);
            if (r) return r;
        }
        r = alloc_fieldset(ctx, ctx->aux);
        if (r) return r;
    }
    if (ckpt) {
        r = ensure_scratch(ctx, (void **)&ctx->ckpt, &ctx->ckpt_bytes,
                           (long long)nseg * ckpt_fields(ctx) * ctx->plane * B * 4);
        if (r) return r;
        r = alloc_fieldset(ctx, ctx->aux);
        if (r) return r;
        ctx->ckpt_flip.assign(nseg, 0);
    }
    float *out_multi[3] = {g_eps, g_del, g_th};
    if (multi) {
        r = ensure_scratch(ctx, (void **)&ctx->acc_multi, &ctx->acc_multi_bytes, (long long)B * nmulti * nint * 4);
        if (r) return r;
        for (int m = 0; m < nmulti; ++m)
            if (out_multi[m]) HCK(hipMemsetAsync(out_multi[m], 0, nint * sizeof(float), ctx->st));
    }
    HCK(hipMemsetAsync(image, 0, nint * sizeof(float), ctx->st));
    float *ilum_acc = ilum;
    if (!ilum_acc) {
        r = ensure_scratch(ctx, (void **)&ctx->ilum_tmp, &ctx->ilum_tmp_bytes, nint * 4);
        if (r) return r;
        ilum_acc = ctx->ilum_tmp;
    }
    HCK(hipMemsetAsync(ilum_acc, 0, nint * sizeof(float), ctx->st));
    ctx->loop_ms = ctx->kernel_ms = 0.0;
    ctx->launches = 0;
    ctx->next_k = -1;
    int tile0 = choose_tile(ctx, two ? Bl : B, true);
    ctx->steps_run = 0;
    if ((half || sb || rv || rbc) && tile_shape(tile0).V > 0) tile0 = kDefaultTile;  // k_step has no fused max / D_t
    const bool accp = pad_ok && tile_pads_acc(tile0);
    if (accp) skind = !half ? SV_F32_PAD : (storage == PAM_STORE_ALL_F16 ? SV_F16_PAD : SV_BF16_PAD);
    // the accumulators of nb shots into the outputs (padded or interior layout)
    auto reduce_acc = [&](const float *parts, int nb, float *out, hipStream_t st) -> int {
        if (accp)
            KCK(launch_reduce_shots_pad(parts, acc_st, nb, g.nz, g.nx, ctx->P, g.n_abc, out, 1, st));
        else
            KCK(launch_reduce_shots(parts, acc_st, nb, (int)nint, out, 1, st));
        return PAM_OK;
    };
    ctx->last_tile = tile0;
    ctx->last_batch = two ? Bl : B;
    const long long pint = (long long)g.n_abc * ctx->P + g.n_abc;  // interior origin in a padded field
    HCK(hipEventRecord(ctx->e0, ctx->st));
    if (two) {
        CkptRun R;
        R.shot0 = shot0;
        R.data = data;
        R.nd = nd;
        R.nint = nint;
        R.acc_st = acc_st;
        R.acc_pad = accp ? 1 : 0;
        R.reduce = reduce_acc;
        R.per_step = per_step;
        R.grad = grad;
        R.d2t = d2t;
        R.tile = tile0;
        R.C = C;
        R.nseg = nseg;
        R.t_lo = t_lo;
        R.t_hi = t_hi;
        R.nbuf = nbuf;
        R.skind = skind;
        R.image = image;
        R.ilum = ilum_acc;
        R.saveall = saveall;
        R.half = half;
        R.elem = elem;
        R.smax0 = grad ? 3.0f * ctx->wav_absmax : 0.0f;
        // an explicit batch: Bl, Bl, ..., the rest (the batches one lane would
        // run); the automatic one: as even as Bl allows
        std::vector<int> sizes;
        if (ctx->batch > 0) {
            for (int b0 = 0; b0 < nshots; b0 += Bl) sizes.push_back(std::min(Bl, nshots - b0));
        } else {
            // an even number of batches, so that neither lane ends alone
            int nbat = (nshots + Bl - 1) / Bl;
            if (nbat > 1 && (nbat & 1) && nbat < nshots) ++nbat;
            for (int i = 0; i < nbat; ++i) sizes.push_back(nshots / nbat + (i < nshots % nbat ? 1 : 0));
        }
        r = ckpt_two_lanes(ctx, R, sizes, Bl);
        if (r) return r;
    }
    for (int b0 = 0; !two && b0 < nshots; b0 += B) {
        const int nb = std::min(B, nshots - b0);
        // ---- forward pass: illumination, and snapshots or checkpoints
        r = zero_shot_state(ctx, nb, true, true);
        if (r) return r;
        r = upload_sources(ctx, shot0 + b0, nb);
        if (r) return r;
        if (!accp && skind == SV_F32_PAD && !sb)
            HCK(hipMemsetAsync(ctx->save, 0, (size_t)nb * per_step * nbuf * elem, ctx->st));
        HCK(hipMemsetAsync(ctx->acc_ilum, 0, (size_t)nb * acc_st * 4, ctx->st));
        HCK(hipMemsetAsync(ctx->acc_img, 0, (size_t)nb * acc_st * 4, ctx->st));
        if (multi) HCK(hipMemsetAsync(ctx->acc_multi, 0, (size_t)B * nmulti * nint * 4, ctx->st));
        KCK(launch_inject_point(ctx->fs.U0, ctx->plane, ctx->d_ssz, ctx->d_ssx, ctx->P, ctx->wav, 0, nb, ctx->st));
        if (half) {
            HCK(hipMemsetAsync(ctx->smax, 0, (size_t)nb * nt * 4 * kSmaxSlots, ctx->st));
            // scale of S_0 (imaging) or a bound of |D_0| (gradient): 3 max|wavelet| bounds
            // |S_1 - 2 S_0| under the CFL condition
            KCK(launch_smax0(ctx->smax, nt, ctx->d_ssz, ctx->d_ssx, g.n_abc, ctx->nzp, ctx->nxp, ctx->wav, nb,
                             grad ? 3.0f * ctx->wav_absmax : 0.0f, ctx->st));
        }
        StepArgs a;
        base_step_args(ctx, a, nb);
        // (D_t is formed here only when this pass saves it: with checkpoints the
        // recomputation forms it, so the pass runs the forward-only kernel)
        a.flags |= F_SRC | F_ILUM | ((ckpt || sb || rbc) ? 0 : F_SAVE) | (half ? F_SMAX : 0) |
                   ((d2t && !sb && !rbc && !ckpt) ? F_D2T : 0);
        a.save_kind = skind;
        a.save = ctx->save;
        a.save_k0 = 0;
        a.save_kstride = per_step;
        a.save_sstride = per_step * nbuf;
        a.save_scale = reinterpret_cast<const float *>(ctx->smax);
        a.smax = ctx->smax;
        a.ilum = ctx->acc_ilum;
        a.image = ctx->acc_img;
        a.acc_sstride = acc_st;
        a.acc_pad = accp ? 1 : 0;
        if (!rv) {
            float *cur = ctx->fs.U0, *prv = ctx->fs.U1;
            // boundary saving: rims of S_k into step k of the strip buffer
            auto sb_save = [&](const float *U, int k) -> int {
                KCK(launch_sb_save(U, ctx->plane, ctx->P, g.n_abc, g.nz, g.nx, nb, sb_kind, ctx->save,
                                   per_step * nt, k, reinterpret_cast<float *>(ctx->smax), nt, ctx->st));
                return PAM_OK;
            };
            if (sb) {
                r = sb_save(cur, 0);
                if (r) return r;
            }
            int j = 0;
            for (int k = 0; k < nt; ++k) {
                if (ckpt && j < nseg && k == std::max(0, j * C - 1)) {
                    r = save_ckpt(ctx, j, nb, cur, prv);
                    if (r) return r;
                    ++j;
                }
                if (rbc) {
                    // undamped through the random layer: no Cerjan factors, no CPML terms
                    StepArgs u = a;
                    u.flags &= ~F_CERJAN;
                    u.cur = cur;
                    u.prv = prv;
                    u.k = k;
                    KCK(launch_step(ctx->phys, false, tile0, u, ctx->st));
                    ++ctx->steps_run;
                } else {
                    r = fwd_step(ctx, a, tile0, cur, prv, k, nullptr);
                    if (r) return r;
                }
                if (sb && k + 1 < nt) {
                    r = sb_save(prv, k + 1);
                    if (r) return r;
                }
                std::swap(cur, prv);
            }
            // the random-boundary rebuild starts from S_nt and S_{nt-1} (whole grid)
            if (rbc) {
                HCK(hipMemcpyAsync(ctx->aux.U0, cur, (size_t)nb * ctx->plane * 4, hipMemcpyDeviceToDevice, ctx->st));
                HCK(hipMemcpyAsync(ctx->aux.U1, prv, (size_t)nb * ctx->plane * 4, hipMemcpyDeviceToDevice, ctx->st));
            }
            // the rebuild starts from S_nt and S_{nt-1} (interiors; zero on the absorbing layer)
            if (sb) {
                KCK(launch_copy_interior(cur, ctx->aux.U0, ctx->plane, ctx->nzp, ctx->P, g.n_abc, g.nz, g.nx, nb,
                                         ctx->st));
                KCK(launch_copy_interior(prv, ctx->aux.U1, ctx->plane, ctx->nzp, ctx->P, g.n_abc, g.nz, g.nx, nb,
                                         ctx->st));
            }
        }
        // ---- backward pass: the adjoint field reuses the main set
        r = zero_shot_state(ctx, nb, true, ctx->reset_adj_cpml != 0);
        if (r) return r;
        const int t0 = t_hi;
        if (t0 >= g.itlag)
            KCK(launch_inject_rec(ctx->fs.U0, ctx->plane, ctx->P, ctx->d_rz, ctx->d_rx, ctx->nrec,
                                  data + (long long)b0 * nd, nd, t0 - g.itlag, g.dx * g.dz, nb, ctx->st));
        float *cur = ctx->fs.U0, *prv = ctx->fs.U1;
        int k0 = 0;  // step held in snapshot slot 0
        // boundary saving: the rebuilt source field, S_{lev+1} in ra and S_lev in rb
        float *ra = ctx->aux.U0, *rb = ctx->aux.U1;
        int lev = nt - 1;
        // one reverse step: S_{lev-1} = v^2 dt^2 L(S_lev) + 2 S_lev - S_{lev+1} over
        // the padded grid (undamped, no CPML terms) into ra, then the rims of
        // step lev-1, the cleared outer band and wav[lev+1] at the source; with
        // with_d, D_lev = S_{lev-1} - 2 S_lev + S_{lev+1} into sbd on the way
        auto rev_step = [&](bool with_d) -> int {
            StepArgs f;
            base_step_args(ctx, f, nb);
            f.flags = with_d ? (F_D2T | F_SAVE | F_D2T_OP) : 0;
            f.save_kind = SV_F32_INT;
            f.save = ctx->sbd;
            f.save_k0 = lev;
            f.save_kstride = nint;
            f.save_sstride = nint;
            f.cur = rb;
            f.prv = ra;
            f.k = lev;
            KCK(launch_step(ctx->phys, false, tile0, f, ctx->st));
            if (rbc) {
                if (lev + 1 < nt)
                    KCK(launch_rbc_source(ra, ctx->plane, ctx->P, with_d ? ctx->sbd : nullptr, nint, g.n_abc, g.nz, g.nx,
                                          ctx->d_ssz, ctx->d_ssx, ctx->wav, lev + 1, nb, ctx->st));
            } else {
                KCK(launch_sb_restore(ra, ctx->plane, ctx->P, g.n_abc, g.nz, g.nx, nb, sb_kind, ctx->save,
                                      per_step * nt, reinterpret_cast<const float *>(ctx->smax), nt, lev - 1,
                                      with_d ? ctx->sbd : nullptr, nint, ctx->d_ssz, ctx->d_ssx, ctx->wav, lev + 1,
                                      ctx->st));
            }
            std::swap(ra, rb);
            --lev;
            return PAM_OK;
        };
        // eps / delta / theta partials of step t (src/Migration.py:1053-1064)
        auto partials = [&](int t, const float *lam) -> int {
            const bool rebuilt = sb || rbc;
            const float *S = rebuilt ? ra + pint
                                     : (const float *)ctx->save + (long long)(t - k0) * per_step +
                                           (skind == SV_F32_PAD ? pint : 0);
            const int sP = (rebuilt || skind == SV_F32_PAD) ? ctx->P : g.nx;
            const long long ss = rebuilt ? ctx->plane : per_step * nbuf;
            float *acc = ctx->acc_multi;
            return launch_grad_multi(ctx->phys, S, sP, ss, lam + pint, ctx->P, ctx->plane, nb, g.nz, g.nx,
                                     g.dx, g.dz, ctx->eps + pint, ctx->del + pint,
                                     ctx->th_pad ? ctx->th_pad + pint : nullptr, ctx->P, acc,
                                     acc + (long long)B * nint, nmulti == 3 ? acc + 2LL * B * nint : nullptr, ctx->st);
        };
        StepArgs b;
        base_step_args(ctx, b, nb);
        b.flags |= F_RECINJ | (grad ? F_GRADM : F_IMAGE) | (d2t ? F_D2T : 0);
        b.data = data + (long long)b0 * nd;
        b.dstride = nd;
        b.save_kind = skind;
        b.save = ctx->save;
        b.save_kstride = per_step;
        b.save_sstride = per_step * nbuf;
        if (sb || rbc) {
            b.save = grad ? (void *)ctx->sbd : nullptr;
            b.save_kstride = b.save_sstride = grad ? nint : ctx->plane;
        }
        if (rv && grad) {
            b.save = ctx->sbd;
            b.save_kstride = b.save_sstride = nint;
        }
        b.save_scale = reinterpret_cast<const float *>(ctx->smax);
        b.image = ctx->acc_img;
        b.acc_sstride = acc_st;
        b.acc_pad = accp ? 1 : 0;
        AdjArgs c;
        std::memset(&c, 0, sizeof(c));
        if (linear_adj) {
            c.nzp = ctx->nzp;
            c.nxp = ctx->nxp;
            c.P = ctx->P;
            c.N = g.n_abc;
            c.nz = g.nz;
            c.nx = g.nx;
            c.nshots = nb;
            c.sstride = ctx->plane;
            c.idx2 = 1.0f / (g.dx * g.dx);
            c.idz2 = 1.0f / (g.dz * g.dz);
            c.idx = 1.0f / g.dx;
            c.idz = 1.0f / g.dz;
            c.idxdz = 1.0f / (g.dz * g.dx);
            c.idt2 = 1.0f / (g.dt * g.dt);
            c.dxdz = g.dx * g.dz;
            c.nt = nt;
            c.itlag = g.itlag;
            c.nrec = ctx->nrec;
            c.v2 = ctx->v2;
            c.eps = ctx->eps;
            c.del = ctx->del;
            c.cs = ctx->cs;
            c.sn = ctx->sn;
            c.cerx = ctx->cerx;
            c.cerz = ctx->cerz;
            c.rrow = ctx->d_rrow;
            c.rhead = ctx->d_rhead;
            c.rnext = ctx->d_rnext;
            c.data = data + (long long)b0 * nd;
            c.dstride = nd;
            c.image = ctx->acc_img;
            c.acc_sstride = nint;
            c.psstride = per_step * nbuf;
            c.gradm = grad ? 1 : 0;
            c.pkstride = per_step;
        }
        // the adjoint step at t, with S_t / D_t in place (snapshot slot t - k0,
        // the rebuilt field, or the D buffer)
        auto adj_at = [&](int t) -> int {
            if (!linear_adj) {
                int rr = fwd_step(ctx, b, tile0, cur, prv, t, nullptr);
                if (rr) return rr;
            } else {
                c.cur = cur;
                c.prv = prv;
                c.k = t;
                c.P_ = (const float *)ctx->save + (long long)(t - k0) * per_step;
                if (sb || rbc) {
                    c.P_ = grad ? ra : rb;
                    c.psstride = ctx->plane;
                    c.D_ = grad ? ctx->sbd : nullptr;
                    c.dsstride = nint;
                }
                if (rv) {
                    c.D_ = grad ? ctx->sbd : nullptr;
                    c.dsstride = nint;
                }
                KCK(launch_adj_step(ctx->phys, c, ctx->st));
                ++ctx->steps_run;
            }
            if (multi) KCK(partials(t, cur));
            std::swap(cur, prv);
            return PAM_OK;
        };
        if (rv) {
            // ---- binomial checkpointing on the second field set (the working set)
            FieldSet &w = ctx->aux;
            const int nf = ckpt_fields(ctx);
            const long long fl = ctx->plane * nb;  // floats per field of a slot
            const int es = rv_half ? 2 : 4;
            float *wc = w.U0, *wp = w.U1;
            int ws = 0;         // step whose state the working set holds
            bool first = true;  // the first descent: illuminate every step once
            std::vector<int> slot_of(nt + 1, -1), flip_of(nslot, 0), pool;
            for (int i = nslot - 1; i >= 0; --i) pool.push_back(i);
            r = zero_fields(ctx, w, nb, true, true);
            if (r) return r;
            KCK(launch_inject_point(w.U0, ctx->plane, ctx->d_ssz, ctx->d_ssx, ctx->P, ctx->wav, 0, nb, ctx->st));
            StepArgs f = a;
            const int fbase = (a.flags & F_CERJAN) | F_SRC;
            f.save_kind = skind;
            f.save = ctx->save;
            f.save_kstride = f.save_sstride = per_step;
            f.dsave = ctx->sbd;
            f.dsave_sstride = nint;
            auto wstep = [&](int k, int extra) -> int {
                f.flags = fbase | extra | (first ? F_ILUM : 0);
                f.save_k0 = k;
                int rr = fwd_step_fs(ctx, w, f, tile0, wc, wp, k, nullptr);
                if (rr) return rr;
                std::swap(wc, wp);
                ws = k + 1;
                return PAM_OK;
            };
            // the state of the working set: both wavefields, and for CPML the
            // previous step's psi and zeta
            auto fields = [&](float **pf) {
                pf[0] = wc;
                pf[1] = wp;
                if (cpml) {
                    const long long in = w.psi_flip ? ctx->psi_half : 0;
                    pf[2] = w.psix + in;
                    pf[3] = w.psiz + in;
                    pf[4] = w.zetax;
                    pf[5] = w.zetaz;
                }
            };
            auto store = [&](int k) -> int {
                if (pool.empty()) return fail(ctx, PAM_E_STATE, "revolve: no free checkpoint slot");
                const int sl = pool.back();
                pool.pop_back();
                slot_of[k] = sl;
                flip_of[sl] = w.psi_flip;
                float *pf[6];
                fields(pf);
                if (rv_half) {
                    FieldRegion reg[6];
                    state_regions(ctx, reg);
                    KCK(launch_pack_half(pf, nf, ctx->plane, ctx->plane, nb,
                                         (char *)ctx->ckpt + (long long)sl * nf * fl * es, fl, ctx->plane,
                                         ctx->rvbits + (long long)sl * nf * B, B, ctx->st, reg));
                    return PAM_OK;
                }
                for (int i = 0; i < nf; ++i) {
                    char *dst = (char *)ctx->ckpt + ((long long)sl * nf + i) * fl * es;
                    HCK(hipMemcpyAsync(dst, pf[i], (size_t)fl * 4, hipMemcpyDeviceToDevice, ctx->st));
                }
                return PAM_OK;
            };
            auto load = [&](int k) -> int {
                const int sl = slot_of[k];
                if (sl < 0) return fail(ctx, PAM_E_STATE, "revolve: step %d has no checkpoint", k);
                wc = w.U0;
                wp = w.U1;
                w.psi_flip = flip_of[sl];
                float *pf[6];
                fields(pf);
                if (rv_half) {
                    FieldRegion reg[6];
                    state_regions(ctx, reg);
                    KCK(launch_unpack_half((const char *)ctx->ckpt + (long long)sl * nf * fl * es, fl, ctx->plane,
                                           ctx->plane, nb, nf, ctx->rvbits + (long long)sl * nf * B, B, pf, ctx->plane,
                                           ctx->st, reg));
                } else {
                    for (int i = 0; i < nf; ++i) {
                        const char *src = (const char *)ctx->ckpt + ((long long)sl * nf + i) * fl * es;
                        HCK(hipMemcpyAsync(pf[i], src, (size_t)fl * 4, hipMemcpyDeviceToDevice, ctx->st));
                    }
                }
                ws = k;
                return PAM_OK;
            };
            auto advance = [&](int from, int to) -> int {
                if (ws != from) {
                    int rr = load(from);
                    if (rr) return rr;
                }
                for (int k = from; k < to; ++k) {
                    int rr = wstep(k, 0);
                    if (rr) return rr;
                }
                return PAM_OK;
            };
            // the adjoint step t: one forward step from state t saves S_t (and D_t)
            auto leaf = [&](int t) -> int {
                int rr = advance(t, t);
                if (rr) return rr;
                rr = wstep(t, F_SAVE | (grad ? F_DSAVE : 0));
                if (rr) return rr;
                if (first) {  // the steps past the adjoint range, for the illumination
                    for (int k = t + 1; k < nt; ++k) {
                        rr = wstep(k, 0);
                        if (rr) return rr;
                    }
                    first = false;
                }
                k0 = t;
                b.save_k0 = t;
                return adj_at(t);
            };
            // adjoint steps hi-1 .. lo with a slot at lo and s more free slots
            std::function<int(int, int, int)> rev = [&](int lo, int hi, int s) -> int {
                const int n = hi - lo;
                int rr;
                if (n == 1) return leaf(lo);
                if (s == 0) {
                    for (int t = hi - 1; t >= lo; --t) {
                        rr = advance(lo, t);
                        if (rr) return rr;
                        rr = leaf(t);
                        if (rr) return rr;
                    }
                    return PAM_OK;
                }
                const int m = lo + revolve_split(n, s);
                rr = advance(lo, m);
                if (rr) return rr;
                rr = store(m);
                if (rr) return rr;
                rr = rev(m, hi, s - 1);
                if (rr) return rr;
                pool.push_back(slot_of[m]);
                slot_of[m] = -1;
                return rev(lo, m, s);
            };
            r = advance(0, t_lo);
            if (r) return r;
            r = store(t_lo);
            if (r) return r;
            r = rev(t_lo, t_hi + 1, nslot - 1);
            if (r) return r;
        } else {
            for (int seg = nseg - 1; seg >= 0; --seg) {
                int hi = t0, lo = t_lo;
                if (ckpt) {
                    lo = std::max(t_lo, seg * C);
                    hi = std::min(t0, (seg + 1) * C - 1);
                    if (lo > hi) continue;
                    // recompute the source field of steps k0 .. hi+1 from checkpoint seg
                    k0 = std::max(0, seg * C - 1);
                    const int k1 = std::min(nt - 1, hi + 1);
                    r = restore_ckpt(ctx, seg, nb);
                    if (r) return r;
                    StepArgs f = a;
                    f.flags = (a.flags & (F_CERJAN | F_SRC)) | F_SAVE | (d2t ? F_D2T : 0);
                    f.save_k0 = k0;
                    float *fc = ctx->aux.U0, *fp = ctx->aux.U1;
                    for (int k = k0; k <= k1; ++k) {
                        r = fwd_step_fs(ctx, ctx->aux, f, tile0, fc, fp, k, nullptr);
                        if (r) return r;
                        std::swap(fc, fp);
                    }
                }
                b.save_k0 = k0;
                for (int t = hi; t >= lo; --t) {
                    if (sb || rbc) {
                        // rebuild S_t (imaging: into rb) or S_{t-1} and D_t (gradient: S_t in ra)
                        while (lev > t) {
                            r = rev_step(false);
                            if (r) return r;
                        }
                        if (grad) {
                            r = rev_step(true);
                            if (r) return r;
                        } else {
                            b.save = rb;
                        }
                        b.save_k0 = t;
                    }
                    r = adj_at(t);
                    if (r) return r;
                }
            }
        }
        r = reduce_acc(ctx->acc_img, nb, image, ctx->st);
        if (r) return r;
        r = reduce_acc(ctx->acc_ilum, nb, ilum_acc, ctx->st);
        if (r) return r;
        if (multi)
            for (int m = 0; m < nmulti; ++m)
                if (out_multi[m])
                    KCK(launch_reduce_shots(ctx->acc_multi + (long long)m * B * nint, nint, nb, (int)nint,
                                            out_multi[m], 1, ctx->st));
    }
    // the image / gradient and illumination sum the source and adjoint fields
    // of every step: a blow-up of either leaves NaN or Inf in them (checked
    // before the normalisation, which a zero illumination would make 0 / 0)
    r = bad_clear(ctx);
    if (r) return r;
    r = bad_check(ctx, image, nint);
    if (r) return r;
    r = bad_check(ctx, ilum_acc, nint);
    if (r) return r;
    if (ctx->normalize) {
        KCK(launch_maxabs(ilum_acc, (int)nint, ctx->maxv, ctx->st));
        KCK(launch_normalize(image, ilum_acc, (int)nint, ctx->normalize, ctx->maxv, ctx->st));
        if (multi)
            for (int m = 0; m < nmulti; ++m)
                if (out_multi[m])
                    KCK(launch_normalize(out_multi[m], ilum_acc, (int)nint, ctx->normalize, ctx->maxv, ctx->st));
    }
    HCK(hipEventRecord(ctx->e1, ctx->st));
    HCK(hipEventSynchronize(ctx->e1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
    ctx->loop_ms = ms;
    ctx->kernel_ms = ms;
    ctx->launches = ctx->steps_run;
    return bad_report(ctx, grad ? "pam_gradient" : "pam_rtm");
}

extern "C" int pam_rtm(pam_ctx *ctx, int shot0, int nshots, const float *data, float *image, float *ilum) {
    if (!ctx) return PAM_E_ARG;
    return adjoint_run(ctx, shot0, nshots, data, false, image, ilum, nullptr, nullptr, nullptr);
}

extern "C" int pam_gradient(pam_ctx *ctx, int shot0, int nshots, const float *residual, float *g_m, float *g_eps,
                            float *g_delta, float *g_theta, float *ilum) {
    if (!ctx) return PAM_E_ARG;
    if ((g_eps || g_delta || g_theta) && ctx->phys == PAM_ACOUSTIC)
        return fail(ctx, PAM_E_ARG, "eps / delta / theta partials need VTI or TTI physics");
    if (g_theta && ctx->phys != PAM_TTI) return fail(ctx, PAM_E_ARG, "the theta partial needs TTI physics");
    return adjoint_run(ctx, shot0, nshots, residual, true, g_m, ilum, g_eps, g_delta, g_theta);
}

extern "C" int pam_misfit(pam_ctx *ctx, int nshots, const float *dobs, const float *dcal, float *residual,
                          double *misfit) {
    if (!ctx) return PAM_E_ARG;
    if (!dobs || !dcal || !misfit) return fail(ctx, PAM_E_ARG, "NULL argument");
    if (nshots < 1) return fail(ctx, PAM_E_ARG, "nshots must be >= 1");
    if (!ctx->geom_set) return fail(ctx, PAM_E_STATE, "no geometry set (pam_set_geometry)");
    const long long n = (long long)nshots * (ctx->g.nt - ctx->g.itlag) * ctx->nrec;
    if (n < 1) return fail(ctx, PAM_E_ARG, "no samples (nt_data * nrec is 0)");
    // two passes in a fixed order: the same inputs give the same bits
    KCK(launch_residual(dobs, dcal, residual, n, ctx->d_acc, ctx->d_acc + kMisfitParts, ctx->st));
    HCK(hipMemcpyAsync(misfit, ctx->d_acc + kMisfitParts, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return PAM_OK;
}

// ------------------------------------------------------------ exact adjoint
// The saved state a reverse step reads: C_k and, for CPML, the psi and zeta
// step k wrote (of step k - 1 only psi and zeta are read)
struct ExState {
    const float *C = nullptr, *psx = nullptr, *psz = nullptr, *zex = nullptr, *zez = nullptr;
};
// the state in a snapshot slot (fields blk floats apart)
static ExState ex_slot(const float *slot, long long blk) {
    ExState e;
    e.C = slot;
    e.psx = slot + blk;
    e.psz = slot + 2 * blk;
    e.zex = slot + 3 * blk;
    e.zez = slot + 4 * blk;
    return e;
}
// Copy the state of step k held by field set f (C_k = cur, and for CPML the
// psi and zeta the step just wrote) into a snapshot slot: fields `blk` floats
// apart, shots one plane apart.
static int exact_copy_state(pam_ctx *ctx, const FieldSet &f, const float *cur, float *slot, int nb, long long blk) {
    const size_t fb = (size_t)ctx->plane * nb * sizeof(float);
    HCK(hipMemcpyAsync(slot, cur, fb, hipMemcpyDeviceToDevice, ctx->st));
    if (ctx->abc == PAM_CPML) {
        const long long newest = f.psi_flip ? ctx->psi_half : 0;
        HCK(hipMemcpyAsync(slot + blk, f.psix + newest, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(slot + 2 * blk, f.psiz + newest, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(slot + 3 * blk, f.zetax, fb, hipMemcpyDeviceToDevice, ctx->st));
        HCK(hipMemcpyAsync(slot + 4 * blk, f.zetaz, fb, hipMemcpyDeviceToDevice, ctx->st));
    }
    return PAM_OK;
}

// Binomial (revolve) checkpointing of the exact reverse sweep: the forward
// steps run on the second field set, whose whole state (both wavefields and,
// for CPML, the previous step's psi and zeta) is kept at up to nslot steps in
// float32 or, with `half`, each field scaled by its max |value| and rounded to
// float16.  Reverse step t needs the state of step t (U_t with the psi and
// zeta step t writes) and the psi and zeta of step t - 1: the zeta of step
// t - 1 is copied into Sm before step t runs (step t overwrites it), and after
// it U_t, both psi halves (psi_t and psi_{t-1}, double-buffered) and zeta_t are
// read where they are (copying the ten fields per reverse step cost a quarter
// of the reverse sweep).  Steps are reversed in treeverse order from nt - 1
// down to 0.
template <class Rev>
static int exact_revolve(pam_ctx *ctx, StepArgs &a, int tile, int nb, int nslot, bool half, long long blk, float *Sm,
                         Rev &reverse_step) {
    const int nt = ctx->g.nt;
    const int nf = ckpt_fields(ctx);
    const long long fl = ctx->plane * nb;  // floats per field of a slot
    const int es = half ? 2 : 4;
    const bool cpml = ctx->abc == PAM_CPML;
    FieldSet &w = ctx->aux;
    int r = zero_fields(ctx, w, nb, true, true);
    if (r) return r;
    KCK(launch_inject_point(w.U0, ctx->plane, ctx->d_ssz, ctx->d_ssx, ctx->P, ctx->wav, 0, nb, ctx->st));
    float *wc = w.U0, *wp = w.U1;
    int ws = 0;  // step whose state the working set holds
    std::vector<int> slot_of(nt + 1, -1), flip_of(nslot, 0), pool;
    for (int i = nslot - 1; i >= 0; --i) pool.push_back(i);
    auto fields = [&](float **pf) {
        pf[0] = wc;
        pf[1] = wp;
        if (cpml) {
            const long long in = w.psi_flip ? ctx->psi_half : 0;
            pf[2] = w.psix + in;
            pf[3] = w.psiz + in;
            pf[4] = w.zetax;
            pf[5] = w.zetaz;
        }
    };
    auto store = [&](int k) -> int {
        if (pool.empty()) return fail(ctx, PAM_E_STATE, "revolve: no free checkpoint slot");
        const int sl = pool.back();
        pool.pop_back();
        slot_of[k] = sl;
        flip_of[sl] = w.psi_flip;
        float *pf[6];
        fields(pf);
        if (half) {
            FieldRegion reg[6];
            state_regions(ctx, reg);
            KCK(launch_pack_half(pf, nf, ctx->plane, ctx->plane, nb, (char *)ctx->ckpt + (long long)sl * nf * fl * es,
                                 fl, ctx->plane, ctx->rvbits + (long long)sl * nf * nb, nb, ctx->st, reg));
            return PAM_OK;
        }
        for (int i = 0; i < nf; ++i) {
            char *dst = (char *)ctx->ckpt + ((long long)sl * nf + i) * fl * es;
            HCK(hipMemcpyAsync(dst, pf[i], (size_t)fl * 4, hipMemcpyDeviceToDevice, ctx->st));
        }
        return PAM_OK;
    };
    auto load = [&](int k) -> int {
        const int sl = slot_of[k];
        if (sl < 0) return fail(ctx, PAM_E_STATE, "revolve: step %d has no checkpoint", k);
        wc = w.U0;
        wp = w.U1;
        w.psi_flip = flip_of[sl];
        float *pf[6];
        fields(pf);
        if (half) {
            FieldRegion reg[6];
            state_regions(ctx, reg);
            KCK(launch_unpack_half((const char *)ctx->ckpt + (long long)sl * nf * fl * es, fl, ctx->plane, ctx->plane, nb,
                                   nf, ctx->rvbits + (long long)sl * nf * nb, nb, pf, ctx->plane, ctx->st, reg));
        } else {
            for (int i = 0; i < nf; ++i) {
                const char *src = (const char *)ctx->ckpt + ((long long)sl * nf + i) * fl * es;
                HCK(hipMemcpyAsync(pf[i], src, (size_t)fl * 4, hipMemcpyDeviceToDevice, ctx->st));
            }
        }
        ws = k;
        return PAM_OK;
    };
    auto wstep = [&](int k) -> int {
        int rr = fwd_step_fs(ctx, w, a, tile, wc, wp, k, nullptr);
        if (rr) return rr;
        std::swap(wc, wp);
        ws = k + 1;
        return PAM_OK;
    };
    auto advance = [&](int from, int to) -> int {
        if (ws != from) {
            int rr = load(from);
            if (rr) return rr;
        }
        for (int k = from; k < to; ++k) {
            int rr = wstep(k);
            if (rr) return rr;
        }
        return PAM_OK;
    };
    auto leaf = [&](int t) -> int {
        int rr = advance(t, t);
        if (rr) return rr;
        if (t >= 1 && cpml) {  // zeta of step t - 1, on its strips (the adjoint step reads no other cell)
            const FieldRegion rg[2] = {region_xstrips(ctx->nzp, ctx->nxp, ctx->P, ctx->g.n_abc),
                                       region_zstrips(ctx->nzp, ctx->P, ctx->g.n_abc)};
            const float *zs[2] = {w.zetax, w.zetaz};
            float *zd[2] = {Sm + 3 * blk, Sm + 4 * blk};
            HCK((hipError_t)launch_copy_region(zs, zd, 2, ctx->plane, nb, rg, ctx->st));
        }
        float *ut = wc;
        rr = wstep(t);
        if (rr) return rr;
        ExState st, sm;
        st.C = ut;
        if (cpml) {
            const long long newest = w.psi_flip ? ctx->psi_half : 0, older = w.psi_flip ? 0 : ctx->psi_half;
            st.psx = w.psix + newest;
            st.psz = w.psiz + newest;
            st.zex = w.zetax;
            st.zez = w.zetaz;
            sm.psx = w.psix + older;
            sm.psz = w.psiz + older;
            sm.zex = Sm + 3 * blk;
            sm.zez = Sm + 4 * blk;
        }
        return reverse_step(t, st, t >= 1 ? &sm : nullptr);
    };
    std::function<int(int, int, int)> rev = [&](int lo, int hi, int s) -> int {
        const int n = hi - lo;
        int rr;
        if (n == 1) return leaf(lo);
        if (s == 0) {
            for (int t = hi - 1; t >= lo; --t) {
                rr = advance(lo, t);
                if (rr) return rr;
                rr = leaf(t);
                if (rr) return rr;
            }
            return PAM_OK;
        }
        const int m = lo + revolve_split(n, s);
        rr = advance(lo, m);
        if (rr) return rr;
        rr = store(m);
        if (rr) return rr;
        rr = rev(m, hi, s - 1);
        if (rr) return rr;
        pool.push_back(slot_of[m]);
        slot_of[m] = -1;
        return rev(lo, m, s);
    };
    r = store(0);
    if (r) return r;
    return rev(0, nt, nslot - 1);
}

// Backward of pam_forward (acoustic physics): a forward pass that keeps the
// state of every step in HBM (or uniform checkpoints of the whole state, the
// segments recomputed on the second field set, with PAM_STORE_CHECKPOINT),
// then the reverse sweep of pam_exact.hip, shots in batches.
extern "C" int pam_forward_adjoint(pam_ctx *ctx, int shot0, int nshots, const float *dseis, float *g_vp, int layout) {
    if (!ctx) return PAM_E_ARG;
    int r = ready(ctx);
    if (r) return r;
    if (ctx->phys != PAM_ACOUSTIC) return fail(ctx, PAM_E_UNSUPPORTED, "the exact adjoint covers acoustic physics");
    ctx->adj_next_t = -2;  // the second field set is reused
    if (nshots < 1) return fail(ctx, PAM_E_ARG, "nshots must be >= 1");
    if (!dseis || !g_vp) return fail(ctx, PAM_E_ARG, "dseis and g_vp must not be NULL");
    if (layout != PAM_INTERIOR && layout != PAM_PADDED) return fail(ctx, PAM_E_ARG, "layout must be 0 or 1");
    const int storage = ctx->storage;
    if (storage != PAM_STORE_ALL_F32 && storage != PAM_STORE_CHECKPOINT && storage != PAM_STORE_REVOLVE_F32 &&
        storage != PAM_STORE_REVOLVE_F16)
        return fail(ctx, PAM_E_UNSUPPORTED,
                    "the exact adjoint keeps the state of every step, uniform checkpoints or revolve slots: use "
                    "PAM_STORE_ALL_F32, PAM_STORE_CHECKPOINT or PAM_STORE_REVOLVE_F32 / _F16");
    const pam_grid &g = ctx->g;
    const int nt = g.nt;
    const long long nd = (long long)(nt - g.itlag) * ctx->nrec;
    const bool cpml = ctx->abc == PAM_CPML;
    const bool ckpt = storage == PAM_STORE_CHECKPOINT;
    const bool rv = storage == PAM_STORE_REVOLVE_F32 || storage == PAM_STORE_REVOLVE_F16;
    const bool rv_half = storage == PAM_STORE_REVOLVE_F16;
    const int nslot = rv ? std::max(1, std::min(ctx->snapshots, nt)) : 0;  // revolve slots
    const int nfc = ckpt_fields(ctx);                                          // fields of a checkpoint or slot
    const int nf = cpml ? 5 : 1;           // saved fields per step
    const int nwork = cpml ? 6 : 0;        // mu (x, z), nu (x, z) in and out
    const int nacc = cpml ? 5 : 1;         // dJ/dV, and dJ/da, dJ/db in x and z
    const int C = ckpt ? (ctx->ckpt_step > 0 ? ctx->ckpt_step : (int)std::ceil(std::sqrt((double)nt))) : nt;
    const int nseg = ckpt ? (nt - 1) / C + 1 : 1;
    const int nbuf = rv ? 2 : (ckpt ? C + 1 : nt);  // states k0 .. hi of a segment (revolve: t - 1 and t)
    const long long pl = ctx->plane;
    size_t freeb = 0, totalb = 0;
    HCK(hipMemGetInfo(&freeb, &totalb));
    int B = batch_size(ctx, nshots);
    const long long per_shot = pl * 4 * ((long long)nbuf * nf + nwork + nacc) +
                               (ckpt ? (long long)nseg * nfc * pl * 4 : 0) +
                               (rv ? (long long)nslot * nfc * (pl * (rv_half ? 2 : 4) + 4) : 0);
    const long long have = ctx->ex_save_bytes + ctx->ex_work_bytes + ctx->ckpt_bytes + ctx->rvbits_bytes;
    long long budget = (long long)((freeb + have) * 0.85);
    if ((ckpt || rv) && !ctx->aux.U0) budget -= pl * ctx->max_shots * 4 * (cpml ? 8 : 2);
    while (B > 1 && (long long)B * per_shot > budget) --B;
    if ((long long)B * per_shot > budget)
        return fail(ctx, PAM_E_NOMEM, "one shot's saved states (%lld bytes) exceed free device memory", per_shot);
    const long long blk = (long long)B * pl;  // floats per field block of a slot
    r = ensure_scratch(ctx, (void **)&ctx->ex_save, &ctx->ex_save_bytes, blk * nbuf * nf * 4);
    if (r) return r;
    r = ensure_scratch(ctx, (void **)&ctx->ex_work, &ctx->ex_work_bytes, blk * (nwork + nacc) * 4);
    if (r) return r;
    r = ensure_scratch(ctx, (void **)&ctx->ex_gacc, &ctx->ex_gacc_bytes, pl * 4);
    if (r) return r;
    if (ckpt) {
        r = ensure_scratch(ctx, (void **)&ctx->ckpt, &ctx->ckpt_bytes, (long long)nseg * nfc * pl * B * 4);
        if (r) return r;
        r = alloc_fieldset(ctx, ctx->aux);
        if (r) return r;
        ctx->ckpt_flip.assign(nseg, 0);
    }
    if (rv) {
        r = ensure_scratch(ctx, (void **)&ctx->ckpt, &ctx->ckpt_bytes,
                           (long long)nslot * nfc * pl * B * (rv_half ? 2 : 4));
        if (r) return r;
        if (rv_half) {
            r = ensure_scratch(ctx, (void **)&ctx->rvbits, &ctx->rvbits_bytes, (long long)nslot * nfc * B * 4);
            if (r) return r;
        }
        r = alloc_fieldset(ctx, ctx->aux);
        if (r) return r;
    }
    HCK(hipMemsetAsync(ctx->ex_gacc, 0, (size_t)pl * 4, ctx->st));
    ctx->loop_ms = ctx->kernel_ms = 0.0;
    ctx->launches = 0;
    ctx->next_k = -1;
    const int tile = choose_tile(ctx, B, true);
    ctx->last_tile = tile;
    ctx->last_batch = B;
    auto slot = [&](int j) { return ctx->ex_save + (long long)j * nf * blk; };
    HCK(hipEventRecord(ctx->e0, ctx->st));
    for (int b0 = 0; b0 < nshots; b0 += B) {
        const int nb = std::min(B, nshots - b0);
        // ---- forward pass: the state of every step, or checkpoints (revolve:
        // none, the reverse sweep runs the forward steps it needs)
        r = zero_shot_state(ctx, nb, true, true);
        if (r) return r;
        r = upload_sources(ctx, shot0 + b0, nb);
        if (r) return r;
        KCK(launch_inject_point(ctx->fs.U0, pl, ctx->d_ssz, ctx->d_ssx, ctx->P, ctx->wav, 0, nb, ctx->st));
        StepArgs a;
        base_step_args(ctx, a, nb);
        a.flags |= F_SRC;
        if (!rv) {
            float *cur = ctx->fs.U0, *prv = ctx->fs.U1;
            int j = 0;
            for (int k = 0; k < nt; ++k) {
                if (ckpt && j < nseg && k == std::max(0, j * C - 1)) {
                    r = save_ckpt(ctx, j, nb, cur, prv);
                    if (r) return r;
                    ++j;
                }
                r = fwd_step(ctx, a, tile, cur, prv, k, nullptr);
                if (r) return r;
                if (!ckpt) {
                    r = exact_copy_state(ctx, ctx->fs, cur, slot(k), nb, blk);
                    if (r) return r;
                }
                std::swap(cur, prv);
            }
        }
        // ---- reverse sweep
        r = zero_shot_state(ctx, nb, true, false);
        if (r) return r;
        HCK(hipMemsetAsync(ctx->ex_work, 0, (size_t)blk * (nwork + nacc) * 4, ctx->st));
        ExactArgs e;
        std::memset(&e, 0, sizeof(e));
        e.nzp = ctx->nzp;
        e.nxp = ctx->nxp;
        e.P = ctx->P;
        e.N = g.n_abc;
        e.nshots = nb;
        e.cpml = cpml ? 1 : 0;
        e.sstride = pl;
        e.fstride = pl;
        e.idx2 = 1.0f / (g.dx * g.dx);
        e.idz2 = 1.0f / (g.dz * g.dz);
        e.idx = 1.0f / g.dx;
        e.idz = 1.0f / g.dz;
        e.v2 = ctx->v2;
        e.cerx = ctx->cerx;
        e.cerz = ctx->cerz;
        e.cax = ctx->cax;
        e.cbx = ctx->cbx;
        e.caz = ctx->caz;
        e.cbz = ctx->cbz;
        float *X = ctx->fs.U0, *Y = ctx->fs.U1;
        float *w = ctx->ex_work;
        float *nxi = w + 2 * blk, *nzi = w + 3 * blk, *nxo = w + 4 * blk, *nzo = w + 5 * blk;
        if (cpml) {
            e.mux = w;
            e.muz = w + blk;
            e.gax = w + (nwork + 1) * blk;
            e.gbx = w + (nwork + 2) * blk;
            e.gaz = w + (nwork + 3) * blk;
            e.gbz = w + (nwork + 4) * blk;
        }
        e.gV = w + nwork * blk;
        const float *data = dseis + (long long)b0 * nd;
        // reverse step t from the saved state of step t (S) and of step t - 1 (Sm,
        // NULL at t = 0; only its psi and zeta are read)
        auto reverse_step = [&](int t, const ExState &S, const ExState *Sm) -> int {
            e.X = X;
            e.Y = Y;
            e.C = S.C;
            if (cpml) {
                e.psx = S.psx;
                e.psz = S.psz;
                e.zex = S.zex;
                e.zez = S.zez;
                e.psxm = Sm ? Sm->psx : nullptr;
                e.pszm = Sm ? Sm->psz : nullptr;
                e.zexm = Sm ? Sm->zex : nullptr;
                e.zezm = Sm ? Sm->zez : nullptr;
                e.nux_in = nxi;
                e.nuz_in = nzi;
                e.nux_out = nxo;
                e.nuz_out = nzo;
                KCK(launch_exact_mu(e, ctx->st));
            }
            KCK(launch_exact_lambda(e, ctx->st));
            std::swap(nxi, nxo);
            std::swap(nzi, nzo);
            if (t >= g.itlag)
                KCK(launch_exact_rec(Y, pl, ctx->P, ctx->d_rz, ctx->d_rx, ctx->nrec, data, nd, t - g.itlag,
                                     cpml ? nullptr : ctx->cerx, ctx->cerz, nb, ctx->st));
            std::swap(X, Y);
            return PAM_OK;
        };
        if (rv) {
            r = exact_revolve(ctx, a, tile, nb, nslot, rv_half, blk, slot(0), reverse_step);
            if (r) return r;
        }
        for (int seg = rv ? -1 : nseg - 1; seg >= 0; --seg) {
            const int lo = ckpt ? seg * C : 0;
            const int hi = ckpt ? std::min(nt - 1, (seg + 1) * C - 1) : nt - 1;
            const int k0 = ckpt ? std::max(0, lo - 1) : 0;
            if (ckpt) {
                r = restore_ckpt(ctx, seg, nb);
                if (r) return r;
                float *fc = ctx->aux.U0, *fp = ctx->aux.U1;
                for (int k = k0; k <= hi; ++k) {
                    r = fwd_step_fs(ctx, ctx->aux, a, tile, fc, fp, k, nullptr);
                    if (r) return r;
                    r = exact_copy_state(ctx, ctx->aux, fc, slot(k - k0), nb, blk);
                    if (r) return r;
                    std::swap(fc, fp);
                }
            }
            for (int t = hi; t >= lo; --t) {
                const bool has_m = t >= 1 && t - 1 >= k0;
                ExState sm;
                if (has_m) sm = ex_slot(slot(t - 1 - k0), blk);
                r = reverse_step(t, ex_slot(slot(t - k0), blk), has_m ? &sm : nullptr);
                if (r) return r;
            }
        }
        KCK(launch_exact_finish(e, ctx->vp_pad, g.dx, g.dz, g.dt, ctx->ap.cpml_d0, ctx->ap.cpml_f_pico, ctx->ex_gacc,
                                ctx->st));
    }
    if (layout == PAM_PADDED)
        KCK(launch_pack_padded(ctx->ex_gacc, ctx->nzp, ctx->nxp, ctx->P, g_vp, ctx->st));
    else
        KCK(launch_expand_T(ctx->ex_gacc, ctx->nzp, ctx->nxp, ctx->P, g.n_abc, g_vp, ctx->st));
    r = bad_clear(ctx);
    if (r) return r;
    r = bad_check(ctx, g_vp, layout == PAM_PADDED ? (long long)ctx->nzp * ctx->nxp : (long long)g.nz * g.nx);
    if (r) return r;
    HCK(hipEventRecord(ctx->e1, ctx->st));
    HCK(hipEventSynchronize(ctx->e1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
    ctx->loop_ms = ms;
    return bad_report(ctx, "pam_forward_adjoint");
}
