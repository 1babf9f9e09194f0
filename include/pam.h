/*
 * pam.h -- C ABI of libpam, the MI355X (gfx950) 2D finite-difference engine:
 * 8th-order acoustic / pseudo-acoustic VTI / TTI time stepping with Cerjan or
 * CPML absorbing boundaries, source injection and receiver gathering, RTM
 * cross-correlation imaging and the FWI adjoint-state gradient.
 *
 * Every entry point takes plain pointers and sizes.  Array arguments named
 * "device" are HIP device pointers owned by the caller (for example a torch
 * tensor's data_ptr()); the library owns its scratch (wavefields, CPML memory
 * variables, snapshot pools).  All work is enqueued on the context's stream;
 * entry points that return results to the host synchronise that stream.
 * Return value: PAM_OK (0) or a negative PAM_E_* code; pam_last_error() then
 * holds a message.  Nothing aborts the process.  One context must not be used
 * from two threads at once; separate contexts may.
 *
 * Grid conventions (the reference's): the interior model is [nz][nx] floats,
 * x fastest; the padded grid adds n_abc cells on every side,
 * nzp = nz + 2 n_abc, nxp = nx + 2 n_abc.  Source and receiver positions are
 * padded-grid cell indices (iz, ix), as the reference computes them
 * (cpp/Survey.cpp:150-158, src/survey.py:131-138).  Seismograms are
 * [nt_data][nrec] per shot, nt_data = nt - itlag; images and gradients are
 * interior [nz][nx].
 */
#ifndef PAM_H
#define PAM_H

#ifdef __cplusplus
extern "C" {
#endif

#define PAM_OK 0
#define PAM_E_ARG (-1)         /* invalid argument or size */
#define PAM_E_HIP (-2)         /* a HIP runtime call failed */
#define PAM_E_STATE (-3)       /* call out of order (e.g. no model set) */
#define PAM_E_NOMEM (-4)       /* device allocation failed */
#define PAM_E_UNSUPPORTED (-5) /* combination not provided */
#define PAM_E_NONFINITE (-6)   /* wavefield blew up (stability check) */

enum pam_physics { PAM_ACOUSTIC = 0, PAM_VTI = 1, PAM_TTI = 2 };
enum pam_abc { PAM_CERJAN = 0, PAM_CPML = 1 };
enum pam_layout { PAM_INTERIOR = 0, PAM_PADDED = 1 };

/* source-wavefield strategy of the adjoint passes (rtm / gradient) */
enum pam_storage {
    PAM_STORE_ALL_F32 = 0,  /* every step saved, float32 */
    PAM_STORE_ALL_F16 = 1,  /* every step saved, float16 with a per-snapshot scale */
    PAM_STORE_ALL_BF16 = 2, /* every step saved, bfloat16 with a per-snapshot scale (F16 / BF16:
                               not with the linearised VTI/TTI Cerjan adjoint nor with eps / delta /
                               theta partials, which read float32 snapshots: PAM_E_UNSUPPORTED) */
    PAM_STORE_CHECKPOINT = 3, /* uniform checkpoints, segments recomputed (exact) */
    /* boundary saving (src/Migration.py:1266-1410): the 4-cell rims just inside
       the interior edge of every step's source field are kept, and the backward
       pass rebuilds the source field on the interior by running the undamped
       step backwards in time (zero on the absorbing layer); rims in float32, or
       float16 / bfloat16 scaled by each step's max |rim value| */
    PAM_STORE_BOUNDARY_F32 = 4,
    PAM_STORE_BOUNDARY_F16 = 5,
    PAM_STORE_BOUNDARY_BF16 = 6,
    /* binomial (revolve) checkpointing with PAM_OPT_SNAPSHOTS slots of the whole
       state (wavefields and CPML memory): the forward pass is recomputed from
       the slots in the order of Griewank's treeverse, no step more than r times
       for the smallest r with C(s + r, s) >= the adjoint steps; slots in float32
       (the recomputed fields are bit-identical to the first pass) or float16
       scaled per field and shot.  No counterpart in the reference. */
    PAM_STORE_REVOLVE_F32 = 7,
    PAM_STORE_REVOLVE_F16 = 8,
    /* random boundary (src/Migration.py:1413-1561): the context's model is a
       random-boundary model (its absorbing layer randomised by the caller, e.g.
       prep.random_boundary); the forward pass runs undamped through it and the
       backward pass rebuilds the source field by reversing the undamped steps
       over the whole grid (no storage); the adjoint field keeps the absorbing
       boundary */
    PAM_STORE_RANDOM_BOUNDARY = 9
};

/* keys of pam_set_option */
enum pam_option {
    PAM_OPT_BATCH = 1,     /* shots advanced together in one launch; 0 (default): the
                              autotuner times max_shots, max_shots/2, ... 8 and keeps the
                              fastest per shot-step (small batches stay in the MALL) */
    PAM_OPT_TILE = 2,      /* step-kernel variant in [0, pam_tile_count()); -1 (default):
                              the autotuner times every variant on the first run of a
                              batch size and keeps the fastest.  Variant 12 advances
                              4 steps per launch (small grids, Cerjan boundary, forward
                              modelling); elsewhere it runs as variant 0 */
    PAM_OPT_GRAPH = 3,     /* 1 (default): a whole run's step loop is captured once into a
                              hipGraph (per batch of shots, cached while the model, geometry,
                              options and output buffers stay the same) and replayed; 0: the
                              steps are launched one by one */
    PAM_OPT_STORAGE = 4,   /* enum pam_storage for rtm / gradient (default ALL_F32) */
    PAM_OPT_CKPT_STEP = 5, /* checkpoint interval for PAM_STORE_CHECKPOINT (0: 2 ceil(sqrt(nt)) for
                              pam_rtm / pam_gradient, ceil(sqrt(nt)) for pam_forward_adjoint) */
    PAM_OPT_RESET_ADJ_CPML = 6, /* 1 (default): clear CPML memory before the
                                   backward pass; 0: keep the forward pass's
                                   values as src/Migration.py does */
    PAM_OPT_NORMALIZE = 7, /* image/gradient normalisation: 0 none, 1 ilum+1e-5 max
                              (src/Migration.py:1086, default), 2 ilum>1e-12
                              (cpp/Migration.cu:470-484) */
    PAM_OPT_ADJ_RANGE = 8, /* adjoint steps: 0 (default) t = nt-2 .. 1 as src/Migration.py:1039;
                              1 t = nt-1 .. 0 as cpp/Migration.cu:383-390 (imaging only: the
                              gradient needs S_{t+1} and S_{t-1}) */
    PAM_OPT_PROFILE = 9,   /* 1: HIP events around every step kernel of pam_forward (the
                              kernel time pam_last_timing reports; launches are then not
                              graph-replayed); 2: after the run, the step kernel alone
                              launched 200 times back to back on the last batch's fields
                              between two HIP events (no per-launch event overhead; the
                              fields and that batch's gathers are overwritten); 0
                              (default): no kernel timing */
    PAM_OPT_SNAPSHOTS = 10, /* checkpoint slots of PAM_STORE_REVOLVE_* (default 50, >= 1) */
    PAM_OPT_LANES = 11,     /* pam_forward: 2 (default) advances two batches at once on two
                               streams the context creates together (default priority),
                               forked from its stream and joined back into it, each lane
                               PAM_OPT_BATCH shots per launch, or half the autotuned batch
                               when that is 0, so that one lane's CPML pre-pass and launch
                               gaps overlap the other's step kernel; 1: one batch at a time
                               on the context's stream.  Runs with PAM_OPT_PROFILE use one
                               lane.  Group runs use member 0's lanes for their sub-groups.
                               pam_rtm / pam_gradient with PAM_STORE_CHECKPOINT (without
                               eps / delta / theta partials or the linearised VTI/TTI
                               Cerjan adjoint) use them too: PAM_OPT_BATCH shots per lane,
                               or half the one-lane batch; the results equal one lane's
                               with the same batch bit for bit. */
    PAM_OPT_GROUP_BYTES = 12  /* pam_group_forward (read from member 0): wavefield bytes
                                 (two fields per shot) of one sub-group of consecutive
                                 members advanced in one launch per step; default 56 MiB,
                                 so that a launch's working set stays in the 256 MB MALL;
                                 0: the whole group in one launch per step */
};

typedef struct pam_grid {
    int nz, nx;       /* interior model size */
    int n_abc;        /* absorbing width on every side */
    float dz, dx, dt; /* spacing [m], time step [s] */
    int nt;           /* total time steps = itlag + nt_data */
    int itlag;        /* first recorded step */
} pam_grid;

typedef struct pam_abc_params {
    float cerjan_factor; /* A[i] = exp(-(factor (N-i))^2); <= 0: the reference's
                            factor 1/(1.4142*3N) (cpp/Modeling.cu:111-120) */
    float cpml_d0;       /* -(M+1) ln R, src/Modeling2D.py:120-126 */
    float cpml_f_pico;   /* fcut/3 */
} pam_abc_params;

typedef struct pam_ctx pam_ctx;

const char *pam_version(void);

/* Replaces the constructors Modeling(Survey*) (cpp/Modeling.cuh:14) and
 * wavefield(parameters) (src/Modeling2D.py:23-25) plus initializeFields
 * (cpp/Modeling.cu:40-82): allocates the wavefields of up to max_shots shots
 * advanced together.  hip_stream may be NULL (a private stream is created). */
int pam_create(const pam_grid *grid, int physics, int abc, const pam_abc_params *abc_params,
               void *hip_stream, int max_shots, pam_ctx **out);
/* Replaces Modeling::freeMemory (cpp/Modeling.cu:13-38). */
int pam_destroy(pam_ctx *ctx);
/* the message of the context's last failure (NULL: of the calling thread's
 * last failed context-free call) */
const char *pam_last_error(const pam_ctx *ctx);
int pam_set_option(pam_ctx *ctx, int key, long long value);
/* Current value of an option; for PAM_OPT_TILE / PAM_OPT_BATCH left on auto,
 * the variant / batch the autotuner chose for the last whole run (-1 / 0
 * before the first). */
int pam_get_option(const pam_ctx *ctx, int key, long long *value);
/* number of kernel variants (tile shapes) the autotuner may choose from */
int pam_tile_count(void);

/* Replaces Modeling::setModel / expandModel (cpp/Modeling.cu:130-182,329-382)
 * and wavefield.ExpandModel (src/Modeling2D.py:38-54): device float arrays,
 * interior [nz][nx] (expanded by edge replication) or padded [nzp][nxp].
 * eps/delta are required for VTI and TTI, theta (radians) for TTI. */
int pam_set_model(pam_ctx *ctx, const float *vp, const float *eps, const float *delta,
                  const float *theta, int layout);
/* Replaces Modeling::createWavelet (cpp/Modeling.cu:84-100) /
 * wavefield.createSourceWavelet (src/Modeling2D.py:27-31): device [nt]. */
int pam_set_wavelet(pam_ctx *ctx, const float *src);
/* Replaces the geometry part of Survey (cpp/Survey.cpp:128-176): host arrays of
 * padded-grid indices, nshot sources and nrec receivers shared by all shots. */
int pam_set_geometry(pam_ctx *ctx, const int *sz, const int *sx, int nshot, const int *rz,
                     const int *rx, int nrec);

/* Per-step interface, mirroring Modeling::forward_step(k) (cpp/Modeling.cu:421-431)
 * and wavefield.forward_stepGPU(k) + store_seismogram (src/Modeling2D.py:292-300,
 * 363-386) for one shot held in slot 0.  pam_reset clears the slot's fields for
 * a shot; pam_forward_step advances one step (injection, update, absorbing
 * boundary, receiver gather into seis row k-itlag when k >= itlag and seis is
 * not NULL).  Steps must be issued in order k = 0, 1, ... */
int pam_reset(pam_ctx *ctx, int shot);
int pam_forward_step(pam_ctx *ctx, int k, float *seis /* device [nt_data][nrec] or NULL */);
/* Copies a padded [nzp][nxp] field of slot 0 to a device buffer.  which: 0 the
 * current field c_k (after the injection of step k, k the next forward step), 1
 * the previous field; 2 the adjoint field lambda_t of the next adjoint step t
 * (after the injection of data row t), 3 the adjoint field before it. */
int pam_get_field(pam_ctx *ctx, int which, float *out);

/* Per-step adjoint interface, mirroring Migration::backward_step(k, P)
 * (cpp/Migration.cu:238-251, cpp/Migration.cuh:56) and
 * migration.backward_step[GPU](k, P) (src/Migration.py:237-286) with the
 * imaging of the drivers' loops (cpp/Migration.cu:383-390: image += P lambda,
 * ilum += P^2 over the interior): one shot in slot 0 of the context's second
 * field set, so that it runs beside the per-step forward interface.
 * pam_adjoint_reset binds the shot's data (device [nt_data][nrec], the muted
 * record or the residual, read until the next reset) and clears the adjoint
 * field, and the CPML memory unless PAM_OPT_RESET_ADJ_CPML is 0 (then the
 * memory the per-step forward run left is taken over, as src/Migration.py
 * does).  pam_adjoint_step(t) runs one adjoint step: data row t injected at
 * the receivers (divided by dx dz) when t >= itlag, the adjoint update -- the
 * forward operator (acoustic, and VTI/TTI with CPML) or the linearised VTI/TTI
 * adjoint reading S_t (Cerjan) --, the absorbing boundary, and, with image /
 * ilum (device [nz][nx], may be NULL), image += S_t lambda_t and ilum += S_t^2
 * on the interior.  S_t is the source field of step t as pam_get_field(ctx, 0)
 * returns it before forward step t (device padded [nzp][nxp]); it may be NULL
 * when nothing reads it.  Steps are issued in decreasing order, t, t-1, ... 0,
 * from any first t.  A whole-run pam_rtm / pam_gradient / pam_forward_adjoint
 * ends the per-step adjoint (PAM_E_STATE until the next reset).  One shot
 * driven this way gives pam_rtm's image (PAM_OPT_NORMALIZE 0) bit for bit. */
int pam_adjoint_reset(pam_ctx *ctx, int shot, const float *data);
int pam_adjoint_step(pam_ctx *ctx, int t, const float *S_t, float *image, float *ilum);

/* Whole-run forward modeling, replacing Modeling::solveWaveEquation
 * (cpp/Modeling.cu:433-479) and wavefield.solveWaveEquationGPU
 * (src/Modeling2D.py:423-467): shots [shot0, shot0+nshots) -> device seis
 * [nshots][nt_data][nrec].  Shots are advanced PAM_OPT_BATCH at a time. */
int pam_forward(pam_ctx *ctx, int shot0, int nshots, float *seis);

/* Whole-run RTM, replacing Migration::solveReverseTimeMigrationOntheFly
 * (cpp/Migration.cu:359-397) and migration.solveBackwardWaveEquationOntheFlyGPU
 * with fwi=False (src/Migration.py:963-1111): data is the (muted) observed data
 * [nshots][nt_data][nrec] (device).  image and ilum (device [nz][nx]) receive the
 * sums over the shots; image is normalised per PAM_OPT_NORMALIZE. */
int pam_rtm(pam_ctx *ctx, int shot0, int nshots, const float *data, float *image, float *ilum);

/* FWI gradient with fwi=True semantics (src/Migration.py:1039-1092,
 * src/Inversion.py:55-86): residual = dobs - dcal [nshots][nt_data][nrec].
 * g_m receives the normalised gradient w.r.t. m = 1/vp^2; g_eps, g_delta,
 * g_theta (may be NULL) the multiparameter partials of src/CudaKernels.py:1174-1379.
 * ilum may be NULL. */
int pam_gradient(pam_ctx *ctx, int shot0, int nshots, const float *residual, float *g_m,
                 float *g_eps, float *g_delta, float *g_theta, float *ilum);

/* Misfit X = sum_shots 1/2 sum (dobs - dcal)^2 (src/Inversion.py:13-53) with a
 * float64 reduction; also writes residual = dobs - dcal when residual != NULL.
 * dobs/dcal/residual are device [nshots][nt_data][nrec]. */
int pam_misfit(pam_ctx *ctx, int nshots, const float *dobs, const float *dcal, float *residual,
               double *misfit);

/* pam_misfit without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It takes no
 * context: dobs / dcal / residual (may be NULL) are device
 * [nshots][nt_data][nrec], and the misfit sum_shots 1/2 sum (dobs - dcal)^2 goes
 * to the DEVICE double `misfit`; the work is only enqueued on hip_stream (NULL:
 * the null stream), on the calling thread's current device, and the call
 * returns without synchronising.  The result equals pam_misfit's bit for bit
 * (the same two fixed-order passes).  The per-block partial sums live in a
 * fixed 8 KiB array per GPU that the library holds until the process ends:
 * every call on one GPU shares it, so calls that overlap on one GPU (two
 * streams, two threads) race; order them.  At most
 * PAM_MISFIT_NOALLOC_MAX_SAMPLES samples per call (PAM_E_ARG beyond);
 * failures report through pam_last_error(NULL). */
#define PAM_MISFIT_NOALLOC_MAX_SAMPLES (1LL << 40)
int pam_misfit_noalloc(int nshots, int nt_data, int nrec, const float *dobs, const float *dcal, float *residual,
                       double *misfit, void *hip_stream);

/* Exact discrete adjoint of pam_forward (acoustic physics, Cerjan or CPML):
 * given dseis = dJ/dseis [nshots][nt_data][nrec] (device) for the seismograms
 * pam_forward produces for shots shot0 .. shot0+nshots-1, writes dJ/dvp into
 * g_vp (device [nz][nx] when layout = PAM_INTERIOR -- the transpose of the
 * edge-replicating expansion applied --, or [nzp][nxp] when PAM_PADDED).  It
 * is the transpose of the forward map's Jacobian, including the dependence of
 * the CPML profiles on vp: the backward a torch.autograd.Function needs.  The
 * forward states are kept per step in float32 (PAM_STORE_ALL_F32), as
 * uniform checkpoints recomputed per segment (PAM_STORE_CHECKPOINT, interval
 * PAM_OPT_CKPT_STEP or ceil(sqrt(nt))), or in PAM_OPT_SNAPSHOTS binomial
 * (revolve) slots of the whole state, float32 (PAM_STORE_REVOLVE_F32: the
 * same gradient bit for bit) or each field scaled by its max |value| and
 * stored in float16 (PAM_STORE_REVOLVE_F16: the adjoint of a forward pass
 * restarted from rounded states, so the gradient moves by about the float16
 * rounding).  No counterpart in the reference, whose gradient (pam_gradient)
 * is preconditioned and normalised. */
int pam_forward_adjoint(pam_ctx *ctx, int shot0, int nshots, const float *dseis, float *g_vp, int layout);

/* Elapsed GPU milliseconds of the last whole-run call's step loop and, with
 * PAM_OPT_PROFILE 1, the summed duration and the number of its step-kernel
 * launches (HIP events around each; with 2, of the back-to-back timing
 * launches after the run); with PAM_OPT_PROFILE 0 kernel_ms is the
 * loop time and kernel_launches the number of steps.  After pam_rtm /
 * pam_gradient, kernel_launches counts the time steps run (forward,
 * recomputed and adjoint) over all shots' batches. */
int pam_last_timing(const pam_ctx *ctx, double *loop_ms, double *kernel_ms, long long *kernel_launches);

/* Preprocessing of the migration / inversion drivers, on the GPU.  These take
 * no context: device arrays, sizes and a HIP stream (NULL: the null stream);
 * they return when the work is done.  pam_last_error(NULL) explains a
 * failure of the calling thread's last such call. */
enum pam_prep_mode {
    PAM_PREP_PY = 0, /* the Python driver's rule */
    PAM_PREP_CPP = 1 /* the C++ driver's rule */
};
/* Model smoothing before migration (in / out device [nz][nx], velocity or,
 * with parameter != 0, an anisotropy parameter smoothed as is; water_mask
 * device [nz][nx] bytes or NULL; sigma in cells).  PAM_PREP_PY replaces
 * smooth_model (src/utils.py:76-129): the slowness convolved with the
 * normalised ceil(6 sigma + 1)-point 2-D Gaussian (float64 accumulation) on the
 * cells a whole kernel fits, the border replicated from the nearest such cell,
 * each column restored above its first cell outside the mask (PAM_E_ARG when
 * the kernel does not fit in the model).  PAM_PREP_CPP replaces the C++
 * driver's smoothing (cpp/Migration.cu:147-236): separable float32 passes
 * along x then z with clamped indices, every masked cell restored. */
int pam_smooth_model(const float *in, float *out, const unsigned char *water_mask, int nz, int nx, double sigma,
                     int parameter, int mode, void *hip_stream);
/* Linear top mute of shot gathers, replacing Mute (src/utils.py:30-49,
 * float64; cpp/Migration.cu:86-115, float32): seis device
 * [nshots][nt][nrec], in place; host coordinates in metres, rec_x / rec_z
 * [nrec] and src_x / src_z [nshots]; zero before t1 = offset / v0 + shift, a
 * linear ramp to t2 = t1 + window, unchanged after. */
int pam_mute(float *seis, int nshots, int nt, int nrec, const double *rec_x, const double *rec_z, const double *src_x,
             const double *src_z, double dt, double shift, double window, double v0, int mode, void *hip_stream);

/* Groups of grids (several models of different sizes, each with its own
 * shots), advanced together in one kernel launch per time step: the batched,
 * uneven-grid counterpart of running the reference's forward driver
 * (cpp/Modeling.cu:433-479, src/Modeling2D.py:423-467) once per model.  The
 * members share physics, absorbing boundary and nt; each keeps its own grid,
 * model, wavelet and geometry.  A group call primes every member on its own
 * stream, runs the steps on member 0's stream and returns when they are done.
 * Errors are reported through pam_last_error(ctxs[0]). */
typedef struct pam_group pam_group;
int pam_group_create(pam_ctx *const *ctxs, int n, pam_group **out);
int pam_group_destroy(pam_group *grp);
/* Forward modeling of shots [0, nshots[i]) of every member i (nshots[i] <=
 * its max_shots) into seis[i] (device [nshots[i]][nt_data_i][nrec_i]); the
 * step loop is graph-replayed when member 0 has PAM_OPT_GRAPH set.  The
 * members are split into sub-groups of consecutive members within member 0's
 * PAM_OPT_GROUP_BYTES; each sub-group runs its steps with one launch per step
 * (two for variant 0 with CPML: the psi pass, then the step), and with
 * PAM_OPT_LANES 2 sub-groups alternate between member 0's two lane streams.  The variant -- 0 (the 16-row tile) or a z-march (4..6) -- is member
 * 0's PAM_OPT_TILE when that names one of these, else timed on the first call
 * for the shot layout.  loop_ms (may be NULL) receives the GPU time of the
 * steps. */
int pam_group_forward(pam_group *grp, const int *nshots, float *const *seis, double *loop_ms);
/* the step-kernel variant of the last pam_group_forward (-1 before) */
int pam_group_get_tile(const pam_group *grp);
/* FWI gradients of the members: member i's pam_gradient g_m for shots
 * 0 .. nshots[i]-1 and its residual[i] (device [nshots[i]][nt_data][nrec]),
 * into g_m[i] (device [nz][nx]), each with its own storage, batch and lane
 * options (float16-scaled snapshots: PAM_STORE_ALL_F16, or
 * PAM_STORE_REVOLVE_F16 slots).  Two host workers take the members (the
 * largest first), each on a stream of the group's own, so that two members
 * share the device; a worker passes its run scratch (snapshots, checkpoints,
 * slots) from one member to the next and all of it is released at the end, so
 * the group needs about two members' scratch at a time, not the sum.  Each
 * gradient is the one the member computes alone.  The workers start after the
 * work queued on the members' streams, and member 0's stream is ordered after
 * them.  g_norm2 (host [n], or NULL): sum of g_m[i]^2 over the interior in
 * float64, summed in a fixed order (the same inputs give the same bits);
 * loop_ms (or NULL): the wall time of the members' runs.  Errors report
 * through member 0.
 * Extends the reference's per-model FWI (src/Inversion.py:55-86, one model
 * per run) to a list of models of different sizes. */
int pam_group_gradient(pam_group *grp, const int *nshots, const float *const *residual, float *const *g_m,
                       double *g_norm2, double *loop_ms);

#ifdef __cplusplus
}
#endif
#endif /* PAM_H */
