// pam_internal.h -- kernel argument blocks and launch helpers shared by
// pam_kernels.hip (device code) and pam_api.cpp (the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pam {

// 1 / m for a 16-bit storage scale m = max |v|: 0 when m is 0 or below the
// smallest normal float (1 / m would overflow to Inf, and 0 * Inf is NaN); the
// values it scales are then stored as 0
__host__ __device__ inline float scale_inv(float m) { return (m >= 1.17549435e-38f) ? 1.0f / m : 0.0f; }

// flags of StepArgs::flags (wave-uniform switches inside k_step)
enum : int {
    F_CERJAN = 1,       // Cerjan damping (cerx/cerz)
    F_SRC = 2,          // writer-side injection of wav[k+1] at (ssz, ssx)
    F_GATHER = 4,       // receiver gather of the input field into seis row k-itlag
    F_RECINJ = 8,       // writer-side injection of data row (k-1)-itlag / dxdz at receivers
    F_SAVE = 16,        // store the (damped) input field into save slot k
    F_ILUM = 32,        // ilum += input^2 on the interior
    F_IMAGE = 64,       // image += S_k * input on the interior
    F_GRADM = 128,      // grad += input * (S_{k+1} - 2 S_k + S_{k-1}) / dt^2 on the interior
    F_ADJ_EXACT = 256,  // (unused bit)
    F_SMAX = 512,       // max |new field| over the interior -> smax[shot][k+1] (scale of snapshot k+1)
    F_D2T = 1024,       // snapshots hold D_k = S_{k+1} - 2 S_k + S_{k-1} (formed in float32 by step k),
                        // F_SMAX then takes max |D_k| (the lagged scale of D_{k+1}) and F_GRADM reads D_k
    F_D2T_OP = 2048,    // with F_D2T | F_SAVE: D_k is the operator term v^2 dt^2 L(c_k) of the update
                        // instead of out - 2 c + prev (equal in exact arithmetic; free of the
                        // cancellation between time levels when the levels carry rounding noise)
    F_DSAVE = 4096,     // D_k = S_{k+1} - 2 S_k + S_{k-1} of the interior into dsave (float32), besides
                        // any F_SAVE snapshot of the input field
};
#define MODE_HAS_CERJAN(f) (((f)&F_CERJAN) != 0)
// the flags a plain forward step uses: kernels instantiated for them alone
// (no snapshot, imaging or gradient paths) are launched when nothing else is set
constexpr int F_FWD_MASK = F_CERJAN | F_SRC | F_GATHER;
// the flag sets of the three passes of an imaging or gradient run with the
// source field recomputed from checkpoints or boundaries: the forward pass
// with illumination, the recompute that stores S_k or D_k, and the adjoint
// pass (of a gradient or of an image); the 16-row tile has an instance for
// each, and one for every flag
constexpr int F_PASS_ILUM = F_FWD_MASK | F_ILUM;
constexpr int F_PASS_SAVE = F_FWD_MASK | F_SAVE | F_D2T;
constexpr int F_PASS_GRAD = F_CERJAN | F_RECINJ | F_GRADM | F_D2T;
constexpr int F_PASS_IMAGE = F_CERJAN | F_RECINJ | F_IMAGE;
// the forward pass of a save-all run: illumination, the snapshot (S_k, or D_k
// for a gradient) and, for 16-bit snapshots, the max that scales the next one
constexpr int F_PASS_SAVEALL = F_FWD_MASK | F_ILUM | F_SAVE | F_SMAX | F_D2T;
constexpr int F_ALL = 0x7fffffff;

// F_SMAX spreads the max of each (shot, step) over kSmaxSlots words: block b
// updates word b % kSmaxSlots, so that a few dozen blocks rather than every
// block of a shot contend for one atomic (same-address atomics serialize: about
// 27 us per step for 1100 blocks a shot on a 451 x 1801 grid); the scale is the
// max over the words
constexpr int kSmaxSlots = 32;

// snapshot storage kinds
// (_INT: interior arrays [nz][nx]; _PAD: padded planes [nzp][P], rows 256-byte
// aligned, four cells per thread in one access)
enum : int { SV_NONE = 0, SV_F32_INT = 1, SV_F32_PAD = 2, SV_F16_INT = 3, SV_BF16_INT = 4, SV_F16_PAD = 5, SV_BF16_PAD = 6 };

struct StepArgs {
    int nzp, nxp, P, N;   // padded grid, row pitch (floats), absorbing width
    int nz, nx;           // interior
    int nxt, nzt, nshots; // x tiles, z tiles, shots in this launch
    long long sstride;    // floats between consecutive shots' fields
    float idx2, idz2, idx, idz, idxdz, idt2, dxdz;
    int k;                // step index of the input field
    int nt, itlag, nrec;
    int flags;

    const float *cur;     // input field c_k (shot 0 of the launch)
    float *prv;           // previous field in, new field out (in place)
    // model coefficient fields, padded [nzp][P], shared by the shots
    const float *v2;      // (vp*vp)*(dt*dt)
    const float *eps, *del, *cs, *sn;
    const float *cerx, *cerz;  // Cerjan factors per column [P] / per row [nzp]
    // CPML (full-grid fields, zero outside the strips); psi is double-buffered
    // across steps: psix_in/psiz_in hold psi of the previous step, psix/psiz
    // receive the new psi (written by k_psi, or by k_march itself)
    const float *psix_in, *psiz_in;
    float *psix, *psiz;
    float *zetax, *zetaz;
    const float *cax, *cbx, *caz, *cbz;
    // sources (writer injection)
    const int *ssz, *ssx;      // [nshots]
    const float *wav;          // [nt]
    // receivers
    const int *rrow;           // [nzp] -> slot or -1
    const int *rhead;          // [nslot][P] -> first receiver at the cell or -1
    const int *rnext;          // [nrec] -> next receiver at the same cell or -1
    float *seis;               // gather target [nshots][nt_data][nrec]
    const float *data;         // injection source [nshots][nt_data][nrec]
    long long dstride;         // floats between shots in seis/data
    // snapshots
    int save_kind;
    void *save;                // base of shot 0's snapshot buffer
    long long save_sstride;    // elements between shots
    long long save_kstride;    // elements between steps
    int save_k0;               // step held in save slot 0 (slot of step k is k - save_k0)
    const float *save_scale;   // [nshots][nt][kSmaxSlots] max |S_k| over the interior (compressed kinds)
    unsigned int *smax;        // F_SMAX target, the same array as float bits (atomicMax)
    const float *save_prev, *save_next;  // (unused; kept for ABI stability)
    float *dsave;              // F_DSAVE target [nshots][nz][nx], dsave_sstride apart
    long long dsave_sstride;
    // accumulators, per shot: interior [nshots][nz][nx], or with acc_pad padded
    // planes [nshots][nzp][P] (acc_sstride apart) whose float32 snapshots
    // (save_kind SV_F32_PAD) are padded too, read and written one aligned float4
    // per thread (16-row tile family only)
    float *ilum, *image;
    long long acc_sstride;
    int acc_pad;
};

// CPML psi pre-pass
struct PsiArgs {
    int nzp, nxp, P, N, nshots;
    long long sstride;
    float idx, idz;
    const float *cur;
    const float *psix_in, *psiz_in;  // psi of the previous step
    float *psix, *psiz;              // new psi
    const float *cax, *cbx, *caz, *cbz;
};

// linearised VTI/TTI adjoint (Cerjan)
struct AdjArgs {
    int nzp, nxp, P, N, nz, nx, nshots;
    long long sstride;
    float idx2, idz2, idx, idz, idxdz, idt2, dxdz;
    int k, nt, itlag, nrec, flags;
    const float *cur;      // lambda_t
    float *prv;            // lambda_{t+1} in, lambda_{t-1} out
    const float *P_;       // source field S_t, padded (damped), shot 0
    long long psstride;    // stride between shots in P_
    const float *v2, *eps, *del, *cs, *sn, *cerx, *cerz;
    const int *rrow, *rhead, *rnext;
    const float *data;
    long long dstride;
    float *image;          // per shot
    long long acc_sstride;
    int gradm;             // image += lambda (S_{t+1} - 2 S_t + S_{t-1}) / dt^2 instead of lambda S_t
    long long pkstride;    // elements between consecutive saved steps of P_
    const float *D_;       // with gradm: D_t = S_{t+1} - 2 S_t + S_{t-1} on the interior
                           // ([nshots][nz][nx], dsstride apart) instead of three levels of P_
    long long dsstride;
};

// V > 0: k_step with V cells per lane and TZ rows per wavefront; V == 0: k_tile
// with TZ rows per block; V == -1: k_march with TZ chunks; V == -2: k_tile_pipe
// with interior tiles in groups of TZ shots; V == -3: the same, its edge tiles
// updating the CPML psi themselves
struct TileShape {
    int V, TZ;
};
int tile_count();
TileShape tile_shape(int idx);
// true when the step-kernel variant updates the CPML psi itself (no k_psi pre-pass)
bool tile_fuses_psi(int idx);
// true for the variants whose passes run the 16-row tile family (k_tile,
// k_tile_pipe), which read and write padded accumulators and snapshots
bool tile_pads_acc(int idx);

// launchers (return hipError_t as int)
// ev: null, or two events that the launch stamps with the kernel's start and end
// CPML psi pass of a group of grids (block prefix table pboff, see k_psi_multi)
int launch_psi_multi(const StepArgs *list, const int *pboff, int ngrid, int npblocks, int k, hipStream_t st);
// blocks of one shot's psi pass (0 without absorbing strips)
int psi_blocks_per_shot(int nzp, int nxp, int N);
int launch_step(int phys, bool cpml, int tile, const StepArgs &a, hipStream_t st, const hipEvent_t *ev = nullptr);
int launch_psi(const PsiArgs &a, hipStream_t st);
// One time step k of a group of grids in one launch (variant 0, the 16-row
// tile, or a z-march variant): list[g] (device) holds grid g's arguments as of
// step 0, with nxt / nzt set for the variant; grid g owns step-kernel blocks
// [boff[g], boff[g+1]) of nblocks = boff[ngrid] and, for variant 0 with CPML,
// psi-pass blocks [pboff[g], pboff[g+1]) of npblocks = pboff[ngrid].
int launch_step_multi(int phys, bool cpml, int tile, const StepArgs *list, const int *boff, int ngrid, int nblocks,
                      int k, const int *pboff, int npblocks, hipStream_t st);
// the step-kernel variants a group launch offers
bool tile_has_multi(int tile);
// row chunks of 16 per block of a z-march variant (0 for the other families)
int march_chunks(int tile);
// steps advanced per launch by the temporal-blocking variant (k_tblock); 8 for
// acoustic measured slower on a 301 x 501 grid (2.49 against 2.30 us per step:
// its 132 KB of LDS leave one block per CU)
inline int tblock_steps(int) { return 4; }
// nsub steps k0 .. k0 + nsub - 1 (1 <= nsub <= tblock_steps(phys)) of a.nshots shots with
// the Cerjan boundary and forward flags, from (in_cur, in_prv) = (U_k0, U_k0-1)
// into (out_cur, out_prv) = (U_k0+nsub, U_k0+nsub-1): a separate buffer pair
int launch_tblock(int phys, const StepArgs &a, const float *in_cur, const float *in_prv, float *out_cur,
                  float *out_prv, int k0, int nsub, hipStream_t st);
// true for the temporal-blocking variant
bool tile_is_tblock(int idx);
// one step of the linearised VTI/TTI adjoint (products and update in one pass)
int launch_adj_step(int phys, const AdjArgs &a, hipStream_t st);

// small utility kernels
int launch_expand(const float *m, int nz, int nx, int N, float *out, int P, hipStream_t st);
int launch_copy_padded(const float *m, int nzp, int nxp, float *out, int P, hipStream_t st);
int launch_coeffs(int phys, const float *vp, const float *theta_in, float dt, int nzp, int nxp, int P, float *v2,
                  float *cs, float *sn, hipStream_t st);
int launch_cpml_coeffs(const float *vp, int nzp, int nxp, int P, int N, float dx, float dz, float dt, float d0,
                       float fpico, float *ax, float *bx, float *az, float *bz, hipStream_t st);
int launch_inject_point(float *U, long long sstride, const int *ssz, const int *ssx, int P, const float *wav,
                        int k, int nshots, hipStream_t st);
// max |S_0| over the interior of each shot (the injected source sample) into smax[s * nt]
// (bound > 0: smax[s * nt] = bound instead -- the scale of D_0)
int launch_smax0(unsigned int *smax, int nt, const int *ssz, const int *ssx, int N, int nzp, int nxp,
                 const float *wav, int nshots, float bound, hipStream_t st);
int launch_inject_rec(float *U, long long sstride, int P, const int *rz, const int *rx, int nrec,
                      const float *data, long long dstride, int row, float dxdz, int nshots, hipStream_t st);
int launch_pack_padded(const float *U, int nzp, int nxp, int P, float *out, hipStream_t st);
int launch_reduce_shots_pad(const float *parts, long long sstride, int nshots, int nz, int nx, int P, int N,
                            float *out, int accumulate, hipStream_t st);
int launch_reduce_shots(const float *parts, long long sstride, int nshots, int n, float *out, int accumulate,
                        hipStream_t st);
int launch_normalize(float *image, const float *ilum, int n, int mode, const float *maxv, hipStream_t st);
int launch_maxabs(const float *x, int n, float *out, hipStream_t st);
// Misfit in two deterministic passes: residual = dobs - dcal and per-block
// float64 partial sums of residual^2 / 2 into part[0 .. misfit_parts(n)), then
// one block sums the partials in a fixed order into *out.  The same n gives the
// same bits on every call; part needs kMisfitParts doubles.
constexpr int kMisfitParts = 1024;
int misfit_parts(long long n);
int launch_residual(const float *dobs, const float *dcal, float *res, long long n, double *part, double *out,
                    hipStream_t st);
// *out = sum of x[i]^2 over n floats in float64, two passes in a fixed order
// (per-block partials into part[misfit_parts(n)], then one block)
int launch_sumsq(const float *x, long long n, double *part, double *out, hipStream_t st);
// the device address of the allocation-free entry points' scratch array on the
// calling thread's current device (looked up once per device)
int pam_noalloc_scratch(double **out);
// U = (U * cerx[x]) * cerz[z] over a padded field of pitch P
int launch_cerjan_scale(float *U, int nzp, int nxp, int P, const float *cerx, const float *cerz, hipStream_t st);
// ilum [nz][nx] += S^2 on the interior of the padded field S (pitch P, absorbing width N)
int launch_ilum_acc(const float *S, int P, int N, int nz, int nx, float *ilum, hipStream_t st);
// *flag = 1 when any of the n floats of x is NaN or Inf (flag is not cleared)
int launch_nonfinite(const float *x, long long n, int *flag, hipStream_t st);
// multiparameter partials (src/CudaKernels.py:1174-1379) of nshots shots:
// S (saved source field at step t) and lam (adjoint field) are given at their
// interior origin with row pitches sP / lP and shot strides ss / ls; the
// partials accumulate into g* + shot * nz * nx ([nz][nx] per shot); the model
// fields eps / del / theta (radians) are given at their interior origin with
// row pitch mP
int launch_grad_multi(int phys, const float *S, int sP, long long ss, const float *lam, int lP, long long ls,
                      int nshots, int nz, int nx, float dx, float dz, const float *eps, const float *del,
                      const float *theta, int mP, float *geps, float *gdel, float *gth, hipStream_t st);

// Boundary saving: the 4-cell rims just inside the interior edge of a padded
// field, sb_width(nz, nx) elements per step in the order [top 4][nx]
// [bottom 4][nx] [left nz-8][4] [right nz-8][4]; kind 0 float32, 1 float16,
// 2 bfloat16 (16-bit kinds hold v / m, m = the step's max |v| in scale[shot][k]).
long long sb_width(int nz, int nx);
int launch_sb_save(const float *U, long long sstride, int P, int N, int nz, int nx, int nshots, int kind,
                   void *strips, long long strip_sstride, int k, float *scale, int nt, hipStream_t st);
// rims of step L into U, the 4-cell band outside the interior cleared, wav[wk]
// added at a source cell of the inner region (wk in [0, nt)); with D, the rim
// cells of D_{L+1} re-formed from the saved rims of steps L, L+1, L+2
int launch_sb_restore(float *U, long long sstride, int P, int N, int nz, int nx, int nshots, int kind,
                      const void *strips, long long strip_sstride, const float *scale, int nt, int L, float *D,
                      long long dsstride, const int *ssz, const int *ssx, const float *wav, int wk, hipStream_t st);
// random-boundary rebuild: wav[wk] at each shot's source cell of U and, when
// interior, of D (interior [nz][nx] per shot; NULL: none)
int launch_rbc_source(float *U, long long sstride, int P, float *D, long long dsstride, int N, int nz, int nx,
                      const int *ssz, const int *ssx, const float *wav, int wk, int nshots, hipStream_t st);
// dst = src on the interior, 0 on the absorbing layer (padded fields)
int launch_copy_interior(const float *src, float *dst, long long sstride, int nzp, int P, int N, int nz, int nx,
                         int nshots, hipStream_t st);

// float16 copies of nf (<= kMaxFields) fields x[f] of n floats per shot
// (shots sstride apart) into q + f qfstride (halfs; shots qsstride apart),
// each field and shot scaled by its max |x| (float bits into
// bits[f bfstride + shot]); and back
constexpr int kMaxFields = 6;
// The cells of a padded field that a pass touches: rows [r0[b], r1[b]) of up
// to two row bands times float4 groups [g0[c], g1[c]) of up to two column
// bands (a band with r1 <= r0 or g1 <= g0 is empty).  Off its region a field
// is zero (the CPML psi and zeta off their strips), so the max, the pack and
// the unpack skip those cells; the slot keeps the field's layout.
struct FieldRegion {
    int r0[2], r1[2], g0[2], g1[2];
    int p4;  // float4 per row of the field
};
// the whole padded plane of nzp rows of P floats
FieldRegion region_all(int nzp, int P);
// the x-strip (columns [4, N) and [nxp-N, nxp-4), rows [4, nzp-4)) or
// z-strip (rows [4, N) and [nzp-N, nzp-4), every column) cells, in float4 groups
FieldRegion region_xstrips(int nzp, int nxp, int P, int N);
FieldRegion region_zstrips(int nzp, int P, int N);
// y[f] = x[f] on reg[f] of each of nshots shots (shot stride sstride floats);
// whole fields where the pointers or the stride are not 16-byte multiples
int launch_copy_region(const float *const *x, float *const *y, int nf, long long sstride, int nshots,
                       const FieldRegion *reg, hipStream_t st);
// reg (nullptr: every field whole) gives each field's region; n = nzp P
int launch_pack_half(const float *const *x, int nf, long long sstride, long long n, int nshots, void *q,
                     long long qfstride, long long qsstride, unsigned int *bits, long long bfstride, hipStream_t st,
                     const FieldRegion *reg = nullptr);
int launch_unpack_half(const void *q, long long qfstride, long long qsstride, long long n, int nshots, int nf,
                       const unsigned int *bits, long long bfstride, float *const *x, long long sstride,
                       hipStream_t st, const FieldRegion *reg = nullptr);

// preprocessing (pam_prep.hip)
int launch_slowness(const float *f, float *s, long long n, int parameter, hipStream_t st);
int launch_first_dry(const unsigned char *mask, int nz, int nx, int *first, hipStream_t st);
int launch_smooth_conv2d(const float *s, int nz, int nx, const float *w, int ks, double total, float *conv,
                         hipStream_t st);
int launch_smooth_finish_py(const float *s_old, const float *conv, const int *first, int nz, int nx, int h,
                            int parameter, float *out, hipStream_t st);
int launch_smooth_sep(const float *in, int nz, int nx, const float *w, int ks, int dir, float *out, hipStream_t st);
int launch_smooth_finish_cpp(const float *s_old, const float *sm, const unsigned char *mask, long long n,
                             int parameter, float *out, hipStream_t st);
// t12: [nshots][nrec][2] (t1, t2), float64 (cpp = 0) or float32 (cpp = 1)
int launch_mute(float *seis, int nshots, int nt, int nrec, const void *t12, double dt, int cpp, hipStream_t st);

// Exact discrete adjoint of the acoustic forward map (pam_forward_adjoint).
// Fields are padded [nzp][P] per shot, shots `sstride` floats apart; the saved
// forward state of step k (C_k = current field after injection, and for CPML
// psi_k / zeta_k after the step) is `fstride` floats apart per shot.
struct ExactArgs {
    int nzp, nxp, P, N, nshots, cpml;
    long long sstride, fstride;
    float idx2, idz2, idx, idz;
    const float *v2, *cerx, *cerz, *cax, *cbx, *caz, *cbz;
    const float *X;                          // lambda_{k+1}
    float *Y;                                // lambda_{k+2} in, lambda_k out
    float *mux, *muz;                        // mu_{k+1} in, mu_k out (CPML)
    const float *nux_in, *nuz_in;            // nu_{k+1} (CPML)
    float *nux_out, *nuz_out;                // nu_k (CPML)
    const float *C, *psx, *psz, *zex, *zez;  // saved state of step k
    const float *psxm, *pszm, *zexm, *zezm;  // saved psi / zeta of step k-1 (NULL at k = 0)
    float *gV, *gax, *gbx, *gaz, *gbz;       // per-shot accumulators
};
int launch_exact_mu(const ExactArgs &a, hipStream_t st);
int launch_exact_lambda(const ExactArgs &a, hipStream_t st);
// Y[rz, rx] += w * data[row] with w the Cerjan factor of the cell (1 for CPML)
int launch_exact_rec(float *Y, long long sstride, int P, const int *rz, const int *rx, int nrec, const float *data,
                     long long dstride, int row, const float *cerx, const float *cerz, int nshots, hipStream_t st);
// gacc += sum over shots of dJ/dvp from the per-shot accumulators (V and, for
// CPML, the profile derivatives)
int launch_exact_finish(const ExactArgs &a, const float *vp, float dx, float dz, float dt, float d0, float fpico,
                        float *gacc, hipStream_t st);
// out [nz][nx] = transpose of the edge-replicating expansion of gpad
int launch_expand_T(const float *gpad, int nzp, int nxp, int P, int N, float *out, hipStream_t st);

}  // namespace pam
