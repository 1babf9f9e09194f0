// pam_kernels.hip -- gfx950 kernels of the 2D finite-difference engine.
//
// The hot kernel (k_step) advances one time step of one or many shots in one
// launch.  Each wavefront owns an x-tile of 64*V columns (V consecutive cells
// per lane, vector loads) and marches down TZ rows.  Rows of the current field
// enter a 9-deep register queue as they are loaded, so each cell of the
// current field is fetched once per tile (plus an 8-row z-halo); the
// x-neighbours of the entering row come from shifted vector loads that hit
// the cache.  First and second x-derivatives are formed when a row enters and
// are queued as well, which gives the TTI mixed derivative as d/dz(d/dx U)
// without a 2-D halo.  The rest of a step happens in the same pass: the
// update, Cerjan damping (the damping the reference applies to U_k after the
// update is re-applied when U_k is read back as U_{k-1}, which gives the same
// values without writing U_k twice), the CPML zeta update, the source
// injection of the next step (added to the value written), the receiver
// gather of this step, and for the adjoint the data injection and the
// imaging / gradient accumulation.
//
// Reference kernels restated (semantics, not code): cpp/Modeling.cu:481-749,
// src/CudaKernels.py:3-721 and :1174-1715, src/utils.py:380-942.
#include "pam_variants.cuh"

namespace pam {

// ---------------------------------------------------------------- CPML psi
// psi = a psi + b dU on the strips, from the current field (src/CudaKernels.py:92-163).
// Two kinds of thread, split at a block boundary (block-uniform role):
//  - x-strip cells, one per thread (rows [4, nzp-4), columns [4, N) and
//    [nxp-N, nxp-4)): the x-neighbours share the cell's cache lines;
//  - z-strip columns: a thread marches down a segment of up to kPsiSeg rows of
//    one z-strip at one column with a 9-row register queue of U, so each row
//    of U is loaded once per segment rather than once per neighbour.
constexpr int kPsiSeg = 12;

// thread t of shot s; x-strip threads first (nx_threads, a whole number of
// blocks), then nsegz segments of each z-strip column
__device__ __forceinline__ void psi_body(const PsiArgs &a, const long long t, const int s, const long long nx_threads,
                                         const int nsegz) {
    const long long so = (long long)s * a.sstride;
    const int wN = a.N - 4;
    // per-shot bases; lane offsets are 32-bit (at())
    const float *__restrict__ U = a.cur + so;
    const float *__restrict__ psix_in = a.psix_in + so;
    const float *__restrict__ psiz_in = a.psiz_in + so;
    float *__restrict__ psix = a.psix + so;
    float *__restrict__ psiz = a.psiz + so;
    const unsigned P = (unsigned)a.P;
    if (t < nx_threads) {
        const long long nxs = (long long)(a.nzp - 8) * 2 * wN;
        if (t >= nxs) return;
        const int row = 4 + (int)(t / (2 * wN));
        const int cc = (int)(t % (2 * wN));
        const int ix = (cc < wN) ? 4 + cc : a.nxp - a.N + (cc - wN);
        const unsigned i = (unsigned)row * P + (unsigned)ix;
        const float *c = at(U, i);
        const float d1 = KA1 * (c[1] - c[-1]) + KA2 * (c[2] - c[-2]) + KA3 * (c[3] - c[-3]) + KA4 * (c[4] - c[-4]);
        *at(psix, i) = *at(a.cax, i) * *at(psix_in, i) + *at(a.cbx, i) * (d1 * a.idx);
        return;
    }
    // z-strip column segment: (side, segment, column)
    const long long u = t - nx_threads;
    const int ncol = a.nxp - 8;
    const long long per_side = (long long)nsegz * ncol;
    if (u >= 2 * per_side) return;
    const int side = (int)(u / per_side);
    const long long v = u - side * per_side;
    const int seg = (int)(v / ncol);
    const int ix = 4 + (int)(v - (long long)seg * ncol);
    const int z0 = (side == 0 ? 4 : a.nzp - a.N) + seg * kPsiSeg;
    const int z1 = min(z0 + kPsiSeg, side == 0 ? a.N : a.nzp - 4);  // one past the last row
    float q[9];
#pragma unroll
    for (int h = 0; h < 8; ++h) q[h] = *at(U, (unsigned)(z0 - 4 + h) * P + (unsigned)ix);
#pragma unroll
    for (int r = 0; r < kPsiSeg; ++r) {
        const int z = z0 + r;
        if (z >= z1) break;  // thread-uniform bound of the segment
        const unsigned i = (unsigned)z * P + (unsigned)ix;
        q[8] = *at(U, i + 4 * P);
        const float d1 = KA1 * (q[5] - q[3]) + KA2 * (q[6] - q[2]) + KA3 * (q[7] - q[1]) + KA4 * (q[8] - q[0]);
        *at(psiz, i) = *at(a.caz, i) * *at(psiz_in, i) + *at(a.cbz, i) * (d1 * a.idz);
#pragma unroll
        for (int h = 0; h < 8; ++h) q[h] = q[h + 1];
    }
}

// threads of one shot's psi pass: x-strip cells (rounded up to whole blocks)
// and the z-strip column segments
static __host__ __device__ inline void psi_threads(int nzp, int nxp, int N, long long *nx_threads, int *nsegz,
                                                   long long *n) {
    const long long nxs = (long long)(nzp - 8) * 2 * (N - 4);
    *nx_threads = (nxs + 255) / 256 * 256;
    *nsegz = (N - 4 + kPsiSeg - 1) / kPsiSeg;
    *n = *nx_threads + 2LL * (*nsegz) * (nxp - 8);
}

// One launch for every shot.  Blocks go to the XCDs round-robin (block h to
// XCD h mod 8); block h = x + 8 (s + nshots q) runs segment block g = 8 q + x
// of shot s, so the segment blocks are spread over the XCDs as they come
// (the x-strip cell blocks and the heavier z-strip column blocks alike), while
// the shots of one segment block -- which read the same CPML coefficients --
// run back to back on one XCD and share its L2.  bps8 = bps rounded up to a
// multiple of 8; blocks past bps leave.
__global__ __launch_bounds__(256) void k_psi(const PsiArgs a, long long nx_threads, int nsegz, int bps) {
    const int h = blockIdx.x;
    const int q = (h >> 3) / a.nshots, s = (h >> 3) - q * a.nshots;
    const int g = q * 8 + (h & 7);
    if (g >= bps) return;  // block-uniform
    psi_body(a, (long long)g * 256 + threadIdx.x, s, nx_threads, nsegz);
}

// The psi pass of a group of grids in one launch: block b belongs to grid g
// with pboff[g] <= b < pboff[g + 1] (blocks of shot 0, then shot 1, ...);
// list[g] holds the grid's arguments as of step 0, its wavefields and psi
// halves trading places on odd steps as in k_tile_multi.
__global__ __launch_bounds__(256) void k_psi_multi(const StepArgs *__restrict__ list, const int *__restrict__ pboff,
                                                   int ngrid, int k) {
    const int b = blockIdx.x;
    if (b >= pboff[ngrid]) return;  // block-uniform
    int g = 0;
    while (g + 1 < ngrid && b >= pboff[g + 1]) ++g;
    const StepArgs &sa = list[g];
    PsiArgs p;
    p.nzp = sa.nzp;
    p.nxp = sa.nxp;
    p.P = sa.P;
    p.N = sa.N;
    p.nshots = sa.nshots;
    p.sstride = sa.sstride;
    p.idx = sa.idx;
    p.idz = sa.idz;
    p.cur = (k & 1) ? sa.prv : sa.cur;
    p.psix_in = (k & 1) ? sa.psix : sa.psix_in;
    p.psiz_in = (k & 1) ? sa.psiz : sa.psiz_in;
    p.psix = const_cast<float *>((k & 1) ? sa.psix_in : sa.psix);
    p.psiz = const_cast<float *>((k & 1) ? sa.psiz_in : sa.psiz);
    p.cax = sa.cax;
    p.cbx = sa.cbx;
    p.caz = sa.caz;
    p.cbz = sa.cbz;
    long long nx_threads, n;
    int nsegz;
    psi_threads(p.nzp, p.nxp, p.N, &nx_threads, &nsegz, &n);
    const int per_shot = (int)((n + 255) / 256);
    const int bb = b - pboff[g];
    const int s = bb / per_shot;
    psi_body(p, (long long)(bb - s * per_shot) * 256 + threadIdx.x, s, nx_threads, nsegz);
}

// ------------------------------------------- linearised VTI/TTI adjoint (Cerjan)
// The products of one cell (src/CudaKernels.py:1382-1472, :1534-1633):
// o[0..4] = A*lam, B*lam, Q*Cx*lam, Q*Cz*lam, H*lam, where Q, Cx, Cz and the
// Sd inside A, B come from the source field; Pf(dx, dz) is the source field at
// the cell's (x + dx, z + dz).
template <int PHYS, typename PF>
__device__ __forceinline__ void adj_products_cell(const PF &Pf, const float e, const float dl, const float c_,
                                                  const float sn_, const float lam, const AdjArgs &a, float o[5]) {
    const float pxx = (KC0 * Pf(0, 0) + KC1 * (Pf(1, 0) + Pf(-1, 0)) + KC2 * (Pf(2, 0) + Pf(-2, 0)) +
                       KC3 * (Pf(3, 0) + Pf(-3, 0)) + KC4 * (Pf(4, 0) + Pf(-4, 0))) *
                      a.idx2;
    const float pzz = (KC0 * Pf(0, 0) + KC1 * (Pf(0, 1) + Pf(0, -1)) + KC2 * (Pf(0, 2) + Pf(0, -2)) +
                       KC3 * (Pf(0, 3) + Pf(0, -3)) + KC4 * (Pf(0, 4) + Pf(0, -4))) *
                      a.idz2;
    const float px = (KA1 * (Pf(1, 0) - Pf(-1, 0)) + KA2 * (Pf(2, 0) - Pf(-2, 0)) + KA3 * (Pf(3, 0) - Pf(-3, 0)) +
                      KA4 * (Pf(4, 0) - Pf(-4, 0))) *
                     a.idx;
    const float pz = (KA1 * (Pf(0, 1) - Pf(0, -1)) + KA2 * (Pf(0, 2) - Pf(0, -2)) + KA3 * (Pf(0, 3) - Pf(0, -3)) +
                      KA4 * (Pf(0, 4) - Pf(0, -4))) *
                     a.idz;
    float X = px, Z = pz, c = 1.f, sn = 0.f;
    if (PHYS == 2) {
        c = c_;
        sn = sn_;
        X = px * c - pz * sn;
        Z = px * sn + pz * c;
    }
    const float X2 = X * X, Z2 = Z * Z, X4 = X2 * X2, Z4 = Z2 * Z2;
    const float num = -2.0f * (e - dl) * X2 * Z2;
    const float den = (1.0f + 2.0f * e) * X4 + Z4 + 2.0f * (1.0f + dl) * X2 * Z2;
    float Sd = 0.f, Cx = 0.f, Cz = 0.f;
    if (fabsf(den) > 1.0e-12f) {
        Sd = num / den;
        if (PHYS == 1) {
            const float fac = 4.0f * (e - dl) * ((1.0f + 2.0f * e) * X4 - Z4) / (den * den);
            Cx = fac * px * Z2;
            Cz = -fac * X2 * pz;
        } else {
            const float K = (1.0f + 2.0f * e) * X4 - Z4;
            const float fac = 4.0f * (e - dl) * X * Z * K / (den * den);
            Cx = fac * pz;
            Cz = -fac * px;
        }
    }
    const float Q = pxx + pzz;
    o[4] = 0.f;
    if (PHYS == 1) {
        o[0] = (1.0f + 2.0f * e + Sd) * lam;
        o[1] = (1.0f + Sd) * lam;
    } else {
        o[0] = ((1.0f + 2.0f * e) * (c * c) + (sn * sn) + Sd) * lam;
        o[1] = ((1.0f + 2.0f * e) * (sn * sn) + (c * c) + Sd) * lam;
        o[4] = (2.0f * e * (2.0f * sn * c)) * lam;
    }
    o[2] = Q * Cx * lam;
    o[3] = Q * Cz * lam;
}

// The adjoint operator at one cell: dxx(A lam) + dzz(B lam) [- dxz(H lam)] -
// dx(Q Cx lam) - dz(Q Cz lam); T(k, dx, dz) is product k at (x + dx, z + dz).
template <int PHYS, typename TF>
__device__ __forceinline__ float adj_operator(const TF &T, const AdjArgs &a) {
    const float dxxA = (KC0 * T(0, 0, 0) + KC1 * (T(0, 1, 0) + T(0, -1, 0)) + KC2 * (T(0, 2, 0) + T(0, -2, 0)) +
                        KC3 * (T(0, 3, 0) + T(0, -3, 0)) + KC4 * (T(0, 4, 0) + T(0, -4, 0))) *
                       a.idx2;
    const float dzzB = (KC0 * T(1, 0, 0) + KC1 * (T(1, 0, 1) + T(1, 0, -1)) + KC2 * (T(1, 0, 2) + T(1, 0, -2)) +
                        KC3 * (T(1, 0, 3) + T(1, 0, -3)) + KC4 * (T(1, 0, 4) + T(1, 0, -4))) *
                       a.idz2;
    const float dxQ = (KA1 * (T(2, 1, 0) - T(2, -1, 0)) + KA2 * (T(2, 2, 0) - T(2, -2, 0)) +
                       KA3 * (T(2, 3, 0) - T(2, -3, 0)) + KA4 * (T(2, 4, 0) - T(2, -4, 0))) *
                      a.idx;
    const float dzQ = (KA1 * (T(3, 0, 1) - T(3, 0, -1)) + KA2 * (T(3, 0, 2) - T(3, 0, -2)) +
                       KA3 * (T(3, 0, 3) - T(3, 0, -3)) + KA4 * (T(3, 0, 4) - T(3, 0, -4))) *
                      a.idz;
    if (PHYS == 2) {
        float acc = 0.f;
        const float aa[4] = {KA1, KA2, KA3, KA4};
#pragma unroll
        for (int ii = 1; ii <= 4; ++ii)
#pragma unroll
            for (int jj = 1; jj <= 4; ++jj)
                acc += aa[ii - 1] * aa[jj - 1] * (T(4, ii, jj) - T(4, ii, -jj) + T(4, -ii, -jj) - T(4, -ii, jj));
        return dxxA + dzzB - acc * a.idxdz - dxQ - dzQ;
    }
    return dxxA + dzzB - dxQ - dzQ;
}

// lambda_new = 2 lam - lam_prev + V op at padded cell (ix, iz) of shot s, with
// Cerjan damping, the data injection of the next adjoint step and the imaging
// with S_t (or the gradient with D_t)
__device__ __forceinline__ void adj_finish(const AdjArgs &a, const int s, const int ix, const int iz, const float op) {
    const long long P = a.P;
    const long long i = (long long)iz * P + ix;
    const long long so = (long long)s * a.sstride;
    const float lam = a.cur[so + i];
    const float fx = a.cerx[ix], fz = a.cerz[iz];
    const float pe = (a.prv[so + i] * fx) * fz;
    float out = ((2.0f * lam - pe + a.v2[i] * op) * fx) * fz;
    if (a.rrow) {
        const int slot = a.rrow[iz];
        const int tn = a.k - 1 - a.itlag;
        if (slot >= 0 && tn >= 0) {
            int rr = a.rhead[(long long)slot * P + ix];
            while (rr >= 0) {
                out += a.data[(long long)s * a.dstride + (long long)tn * a.nrec + rr] / a.dxdz;
                rr = a.rnext[rr];
            }
        }
    }
    a.prv[so + i] = out;
    if (a.image && iz >= a.N && iz < a.nzp - a.N && ix >= a.N && ix < a.nxp - a.N) {
        const long long ii = (long long)(iz - a.N) * a.nx + (ix - a.N);
        const float *Ps = a.P_ + (long long)s * a.psstride + i;
        if (a.gradm) {
            const float d2t = (a.D_ ? a.D_[(long long)s * a.dsstride + ii]
                                    : (Ps[a.pkstride] - 2.0f * Ps[0] + Ps[-a.pkstride])) *
                              a.idt2;
            a.image[(long long)s * a.acc_sstride + ii] += d2t * lam;
        } else {
            a.image[(long long)s * a.acc_sstride + ii] += Ps[0] * lam;
        }
    }
}

// One adjoint step of the linearised VTI/TTI operator, products and update in
// one pass: a block takes one 64 x 16 tile of one shot (tiles of a shot group
// adjacent, so the model is reused from L2; XCD-contiguous ranges), stages the
// source field with an 8-cell halo in LDS, forms the products over the tile
// with a 4-cell halo in LDS (zero outside the stencil interior, as the
// two-pass form wrote them), then updates its cells from them.  No product
// plane goes through HBM (the two-pass form wrote and re-read 16 or 20 bytes
// per cell and shot), and each value is formed by the same expressions, so the
// results are the two-pass form's bit for bit.
template <int PHYS>
__global__ __launch_bounds__(256) void k_adj_fused(const AdjArgs a) {
    constexpr int TX = 64, TZ = 16;
    constexpr int HX = TX + 8, HZ = TZ + 8;    // products: the tile with a 4-cell halo
    constexpr int PX = TX + 16, PZ = TZ + 16;  // source field: the tile with an 8-cell halo
    constexpr int NT = (PHYS == 2) ? 5 : 4;
    __shared__ float sP[PZ][PX];
    __shared__ float sT[NT][HZ][HX];
    const int nxt = (a.nxp + TX - 1) / TX, nzt = (a.nzp + TZ - 1) / TZ;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= nxt * nzt * a.nshots) return;  // block-uniform
    const int tile = b / a.nshots, s = b - tile * a.nshots;
    const int tzi = tile / nxt, txi = tile - tzi * nxt;
    const int x0 = txi * TX, z0 = tzi * TZ;
    const int tid = threadIdx.x;
    const int P = a.P, nzp = a.nzp, nxp = a.nxp;
    const float *Pp = a.P_ + (long long)s * a.psstride;
    const float *lam = a.cur + (long long)s * a.sstride;
    for (int i = tid; i < PZ * PX; i += 256) {
        const int r = i / PX, q = i - r * PX;
        const int gz = z0 - 8 + r, gx = x0 - 8 + q;
        sP[r][q] = (gz >= 0 && gz < nzp && gx >= 0 && gx < nxp) ? Pp[(long long)gz * P + gx] : 0.0f;
    }
    __syncthreads();
    for (int i = tid; i < HZ * HX; i += 256) {
        const int r = i / HX, q = i - r * HX;
        const int gz = z0 - 4 + r, gx = x0 - 4 + q;
        float o[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        if (gz >= 4 && gz < nzp - 4 && gx >= 4 && gx < nxp - 4) {
            const long long gi = (long long)gz * P + gx;
            const auto Pf = [&](int dx, int dz) { return sP[r + 4 + dz][q + 4 + dx]; };
            adj_products_cell<PHYS>(Pf, a.eps[gi], a.del[gi], PHYS == 2 ? a.cs[gi] : 1.0f,
                                    PHYS == 2 ? a.sn[gi] : 0.0f, lam[gi], a, o);
        }
#pragma unroll
        for (int k = 0; k < NT; ++k) sT[k][r][q] = o[k];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < (TX * TZ) / 256; ++u) {
        const int c = tid + 256 * u;
        const int r = c / TX, q = c - r * TX;
        const int ix = x0 + q, iz = z0 + r;
        if (ix < 4 || ix >= nxp - 4 || iz < 4 || iz >= nzp - 4) continue;
        const auto T = [&](int k, int dx, int dz) { return sT[k][r + 4 + dz][q + 4 + dx]; };
        adj_finish(a, s, ix, iz, adj_operator<PHYS>(T, a));
    }
}

// ---------------------------------------------------------------- utilities
__global__ void k_expand(const float *m, int nz, int nx, int N, float *out, int P) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    const int nxp = nx + 2 * N;
    if (ix >= P) return;
    int jz = iz - N, jx = ix - N;
    jz = jz < 0 ? 0 : (jz > nz - 1 ? nz - 1 : jz);
    jx = jx < 0 ? 0 : (jx > nx - 1 ? nx - 1 : jx);
    out[(long long)iz * P + ix] = (ix < nxp) ? m[(long long)jz * nx + jx] : 0.0f;
}
__global__ void k_copy_padded(const float *m, int nzp, int nxp, float *out, int P) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= P) return;
    out[(long long)iz * P + ix] = (ix < nxp) ? m[(long long)iz * nxp + ix] : 0.0f;
}
__global__ void k_pack_padded(const float *U, int nzp, int nxp, int P, float *out) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= nxp) return;
    out[(long long)iz * nxp + ix] = U[(long long)iz * P + ix];
}
// coefficient fields: v2 = (vp*vp)*(dt*dt); cs/sn = cosf/sinf(theta)
__global__ void k_coeffs(int phys, const float *vp, const float *th, float dt, long long n, float *v2, float *cs,
                         float *sn) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = vp[i];
    v2[i] = (v * v) * (dt * dt);
    if (phys == 2) {
        const float t = th[i];
        cs[i] = cosf(t);
        sn[i] = sinf(t);
    }
}
// CPML profiles, the FP32 formula of src/CudaKernels.py:33-90
__device__ __forceinline__ void cpml_ab(int i, int n, int N, float h, float vpc, float d0, float fp, float dt,
                                        float *a, float *b) {
    float d = 0.0f, alpha = 0.0f;
    if (i < N) {
        const float pos = ((N - i - 1.0f) * h) / (N * h);
        d = d0 / (2.0f * N * h) * (pos * pos) * vpc;
        alpha = 3.1415926535f * fp * (1.0f - pos * pos);
    } else if (i >= n - N) {
        const float pos = ((float)(i - n + N) * h) / (N * h);
        d = d0 / (2.0f * N * h) * (pos * pos) * vpc;
        alpha = 3.1415926535f * fp * (1.0f - pos * pos);
    }
    *a = expf(-(d + alpha) * dt);
    *b = (fabsf(d + alpha) > 1e-10f) ? (d / (d + alpha)) * (*a - 1.0f) : 0.0f;
}
__global__ void k_cpml_coeffs(const float *vp, int nzp, int nxp, int P, int N, float dx, float dz, float dt, float d0,
                              float fp, float *ax, float *bx, float *az, float *bz) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= P) return;
    const long long i = (long long)iz * P + ix;
    if (ix >= nxp) {
        ax[i] = 1.f;
        bx[i] = 0.f;
        az[i] = 1.f;
        bz[i] = 0.f;
        return;
    }
    float a, b;
    cpml_ab(ix, nxp, N, dx, vp[i], d0, fp, dt, &a, &b);
    ax[i] = a;
    bx[i] = b;
    cpml_ab(iz, nzp, N, dz, vp[i], d0, fp, dt, &a, &b);
    az[i] = a;
    bz[i] = b;
}
__global__ void k_inject_point(float *U, long long sstride, const int *ssz, const int *ssx, int P, const float *wav,
                               int k, int nshots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    U[(long long)s * sstride + (long long)ssz[s] * P + ssx[s]] += wav[k];
}
// max |S_0| over the interior: the field at step 0 is the injected source
// sample (word 0 of the step's kSmaxSlots; the others are cleared)
__global__ void k_smax0(unsigned int *smax, int nt, const int *ssz, const int *ssx, int N, int nzp, int nxp,
                        const float *wav, int nshots, float bound) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const bool in = ssz[s] >= N && ssz[s] < nzp - N && ssx[s] >= N && ssx[s] < nxp - N;
    smax[(long long)s * nt * kSmaxSlots] = __float_as_uint(bound > 0.0f ? bound : (in ? fabsf(wav[0]) : 0.0f));
}
// one thread per shot so duplicate receiver cells accumulate without races
__global__ void k_inject_rec(float *U, long long sstride, int P, const int *rz, const int *rx, int nrec,
                             const float *data, long long dstride, int row, float dxdz, int nshots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    for (int r = 0; r < nrec; ++r)
        U[(long long)s * sstride + (long long)rz[r] * P + rx[r]] +=
            data[(long long)s * dstride + (long long)row * nrec + r] / dxdz;
}
__global__ void k_reduce_shots(const float *parts, long long sstride, int nshots, long long n, float *out,
                               int accumulate) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = accumulate ? out[i] : 0.0f;
    for (int s = 0; s < nshots; ++s) acc += parts[(long long)s * sstride + i];
    out[i] = acc;
}
// the same from padded per-shot planes (sstride apart, pitch P, interior at
// row N, column N) into the interior [nz][nx]: the shots are added in the
// same order, so the sums are those of k_reduce_shots over interior copies
__global__ void k_reduce_shots_pad(const float *parts, long long sstride, int nshots, int nz, int nx, int P, int N,
                                   float *out, int accumulate) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nz * nx) return;
    const int iz = (int)(i / nx), ix = (int)(i - (long long)iz * nx);
    const long long pi = (long long)(iz + N) * P + ix + N;
    float acc = accumulate ? out[i] : 0.0f;
    for (int s = 0; s < nshots; ++s) acc += parts[(long long)s * sstride + pi];
    out[i] = acc;
}
__global__ void k_maxabs(const float *x, long long n, unsigned int *out) {
    __shared__ float red[256];
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}
// mode 1: image / (ilum + 1e-5 max ilum) (src/Migration.py:1086);
// mode 2: image / ilum where ilum > 1e-12, else 0 (cpp/Migration.cu:470-484)
__global__ void k_normalize(float *image, const float *ilum, long long n, int mode, const unsigned int *maxbits) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mode == 1) {
        const float mx = __uint_as_float(*maxbits);
        image[i] = image[i] / (ilum[i] + 1e-5f * mx);
    } else if (mode == 2) {
        image[i] = (ilum[i] > 1e-12f) ? image[i] / ilum[i] : 0.0f;
    }
}
// block-wide float64 sum of one value per thread (256 threads), fixed order:
// a wavefront butterfly, then the four wavefront sums in order
__device__ __forceinline__ double block_sum256(double v) {
    __shared__ double wsum[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    return ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}
// residual = dobs - dcal and, per block, the float64 sum of residual^2 / 2
// over a grid-stride share into part[blockIdx.x]; every partial is written
__global__ void k_residual(const float *dobs, const float *dcal, float *res, long long n, double *part) {
    double sum = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float r = dobs[i] - dcal[i];
        if (res) res[i] = r;
        sum += 0.5 * (double)r * (double)r;
    }
    const double b = block_sum256(sum);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}
// per-block partial sums of x^2 in float64 (the same fixed order as k_residual)
__global__ void k_sumsq(const float *x, long long n, double *part) {
    double sum = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        sum += (double)x[i] * (double)x[i];
    const double b = block_sum256(sum);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}
// *out = the sum of the nparts partials (one block, fixed order)
__global__ void k_sum_parts(const double *part, int nparts, double *out) {
    double sum = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) sum += part[i];
    const double b = block_sum256(sum);
    if (threadIdx.x == 0) *out = b;
}
// U = (U * cerx[x]) * cerz[z] over a padded field (the product and its order
// as the step kernels damp a field they save)
__global__ void k_cerjan_scale(float *U, int nzp, int nxp, int P, const float *cerx, const float *cerz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x >= nxp) return;
    float *u = U + (long long)z * P + x;
    *u = (*u * cerx[x]) * cerz[z];
}
// ilum[iz][ix] += S^2 on the interior of a padded field of pitch P
__global__ void k_ilum_acc(const float *S, int P, int N, int nz, int nx, float *ilum) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= nx) return;
    const float v = S[(long long)(iz + N) * P + ix + N];
    float *acc = ilum + (long long)iz * nx + ix;
    *acc += v * v;
}
// *flag = 1 if any value is NaN or Inf
__global__ void k_nonfinite(const float *x, long long n, int *flag) {
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        bad |= !isfinite(x[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) *flag = 1;
}
// multiparameter partials on interior arrays, float64 arithmetic
// (src/CudaKernels.py:1174-1379)
template <int PHYS>
__global__ void k_grad_multi(const float *S, int sP, long long ss, const float *lam, int lP, long long ls, int nz,
                             int nx, float dx, float dz, const float *eps, const float *del, const float *theta, int mP,
                             float *geps, float *gdel, float *gth) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    const int s = blockIdx.z;
    if (ix < 4 || ix >= nx - 4 || iz < 4 || iz >= nz - 4) return;
    const long long i = (long long)iz * nx + ix;  // interior index (models, partials)
    const long long n = sP;
    const float *c = S + (long long)s * ss + (long long)iz * sP + ix;
    const long long go = (long long)s * nz * nx;
    const long long mi = (long long)iz * mP + ix;  // model index
    const double pxx = (double)((KC0 * c[0] + KC1 * (c[1] + c[-1]) + KC2 * (c[2] + c[-2]) + KC3 * (c[3] + c[-3]) +
                                 KC4 * (c[4] + c[-4])) /
                                (dx * dx));
    const double pzz = (double)((KC0 * c[0] + KC1 * (c[n] + c[-n]) + KC2 * (c[2 * n] + c[-2 * n]) +
                                 KC3 * (c[3 * n] + c[-3 * n]) + KC4 * (c[4 * n] + c[-4 * n])) /
                                (dz * dz));
    const double px =
        (double)((KA1 * (c[1] - c[-1]) + KA2 * (c[2] - c[-2]) + KA3 * (c[3] - c[-3]) + KA4 * (c[4] - c[-4])) / dx);
    const double pz = (double)((KA1 * (c[n] - c[-n]) + KA2 * (c[2 * n] - c[-2 * n]) + KA3 * (c[3 * n] - c[-3 * n]) +
                                KA4 * (c[4 * n] - c[-4 * n])) /
                               dz);
    const double e = eps[mi], dl = del[mi];
    const double adj = lam[(long long)s * ls + (long long)iz * lP + ix];
    if (PHYS == 1) {
        const double num = -2.0 * (e - dl) * (px * px) * (pz * pz);
        const double den =
            (1.0 + 2.0 * e) * (px * px * px * px) + (pz * pz * pz * pz) + 2.0 * (1.0 + dl) * (px * px) * (pz * pz);
        const double dn_e = -2.0 * px * px * pz * pz, dn_d = 2.0 * px * px * pz * pz;
        const double dd_e = 2.0 * px * px * px * px, dd_d = 2.0 * px * px * pz * pz;
        double sde = 0, sdd = 0;
        if (fabs(den) > 1.0e-150) {
            sde = (dn_e * den - num * dd_e) / (den * den);
            sdd = (dn_d * den - num * dd_d) / (den * den);
        }
        geps[go + i] = (float)((double)geps[go + i] + adj * ((-2.0 - sde) * pxx - sde * pzz));
        gdel[go + i] = (float)((double)gdel[go + i] + adj * (-sdd * (pxx + pzz)));
    } else {
        double acc = 0.0;
        const float aa[4] = {KA1, KA2, KA3, KA4};
        float accf = 0.0f;
        for (int ii = 1; ii <= 4; ++ii)
            for (int jj = 1; jj <= 4; ++jj)
                accf += aa[ii - 1] * aa[jj - 1] *
                        (c[jj * n + ii] - c[-jj * n + ii] + c[-jj * n - ii] - c[jj * n - ii]);
        acc = (double)(accf / (dz * dx));
        const double pxz = acc;
        const double th = theta[mi];
        const double cs = cos(th), sn = sin(th);
        const double h = px * cs - pz * sn, q = px * sn + pz * cs;
        const double h2 = h * h, q2 = q * q;
        const double num = -2.0 * (e - dl) * h2 * q2;
        const double den = (1.0 + 2.0 * e) * h2 * h2 + q2 * q2 + 2.0 * (1.0 + dl) * h2 * q2;
        double sde = 0, sdd = 0, sdt = 0;
        if (fabs(den) >= 1.0e-150) {
            const double dn_e = -2.0 * h2 * q2, dn_d = 2.0 * h2 * q2;
            const double dd_e = 2.0 * h2 * h2, dd_d = 2.0 * h2 * q2;
            const double dn_t = -2.0 * (e - dl) * (-2.0 * h * q * q * q + 2.0 * h * h * h * q);
            const double dd_t = -4.0 * (1.0 + 2.0 * e) * h * h * h * q + 4.0 * h * q * q * q +
                                2.0 * (1.0 + dl) * (-2.0 * h * q * q * q + 2.0 * h * h * h * q);
            sde = (dn_e * den - num * dd_e) / (den * den);
            sdd = (dn_d * den - num * dd_d) / (den * den);
            sdt = (dn_t * den - num * dd_t) / (den * den);
        }
        const double s2t = sin(2.0 * th), c2t = cos(2.0 * th);
        const double dA = 2.0 * e * s2t - sdt, dB = -2.0 * e * s2t - sdt, dC = 4.0 * e * c2t;
        geps[go + i] = (float)((double)geps[go + i] +
                               adj * ((-(2.0 * cs * cs + sde)) * pxx - (2.0 * sn * sn + sde) * pzz + 2.0 * s2t * pxz));
        gdel[go + i] = (float)((double)gdel[go + i] + adj * (-sdd * (pxx + pzz)));
        gth[go + i] = (float)((double)gth[go + i] + adj * (dA * pxx + dB * pzz + dC * pxz));
    }
}

// ---------------------------------------------------------------- launchers
int tile_count() { return kNumTiles; }
TileShape tile_shape(int idx) {
    if (idx < 0 || idx >= kNumTiles) idx = 0;
    return kTiles[idx];
}
bool tile_fuses_psi(int idx) { return tile_shape(idx).V == -1 || tile_shape(idx).V == -3; }
bool tile_pads_acc(int idx) {
    const int v = tile_shape(idx).V;
    return v == 0 || v == -2 || v == -3 || v == -4;
}
bool tile_is_tblock(int idx) { return idx >= 0 && idx < kNumTiles && kTiles[idx].V == -4; }

int launch_step_acoustic_cerjan(int, const StepArgs &, hipStream_t, const hipEvent_t *);
int launch_step_acoustic_cpml(int, const StepArgs &, hipStream_t, const hipEvent_t *);
int launch_step_vti_cerjan(int, const StepArgs &, hipStream_t, const hipEvent_t *);
int launch_step_vti_cpml(int, const StepArgs &, hipStream_t, const hipEvent_t *);
int launch_step_tti_cerjan(int, const StepArgs &, hipStream_t, const hipEvent_t *);
int launch_step_tti_cpml(int, const StepArgs &, hipStream_t, const hipEvent_t *);

int launch_multi_acoustic_cerjan(int, const StepArgs *, const int *, int, int, int, const int *, int, hipStream_t);
int launch_multi_acoustic_cpml(int, const StepArgs *, const int *, int, int, int, const int *, int, hipStream_t);
int launch_multi_vti_cerjan(int, const StepArgs *, const int *, int, int, int, const int *, int, hipStream_t);
int launch_multi_vti_cpml(int, const StepArgs *, const int *, int, int, int, const int *, int, hipStream_t);
int launch_multi_tti_cerjan(int, const StepArgs *, const int *, int, int, int, const int *, int, hipStream_t);
int launch_multi_tti_cpml(int, const StepArgs *, const int *, int, int, int, const int *, int, hipStream_t);
int launch_step_multi(int phys, bool cpml, int tile, const StepArgs *list, const int *boff, int ngrid, int nblocks,
                      int k, const int *pboff, int npblocks, hipStream_t st) {
    if (phys == 0)
        return cpml ? launch_multi_acoustic_cpml(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st)
                    : launch_multi_acoustic_cerjan(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st);
    if (phys == 1)
        return cpml ? launch_multi_vti_cpml(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st)
                    : launch_multi_vti_cerjan(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st);
    return cpml ? launch_multi_tti_cpml(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st)
                : launch_multi_tti_cerjan(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st);
}
int march_chunks(int tile) { return tile_shape(tile).V == -1 ? tile_shape(tile).TZ : 0; }
bool tile_has_multi(int tile) { return tile == 0 || march_chunks(tile) > 0; }

int launch_step(int phys, bool cpml, int tile, const StepArgs &a, hipStream_t st, const hipEvent_t *ev) {
    if (phys == 0)
        return cpml ? launch_step_acoustic_cpml(tile, a, st, ev) : launch_step_acoustic_cerjan(tile, a, st, ev);
    if (phys == 1) return cpml ? launch_step_vti_cpml(tile, a, st, ev) : launch_step_vti_cerjan(tile, a, st, ev);
    return cpml ? launch_step_tti_cpml(tile, a, st, ev) : launch_step_tti_cerjan(tile, a, st, ev);
}

int launch_psi(const PsiArgs &a, hipStream_t st) {
    if (a.N <= 4) return 0;
    long long nx_threads, n;  // the z-columns start on a block boundary
    int nsegz;
    psi_threads(a.nzp, a.nxp, a.N, &nx_threads, &nsegz, &n);
    const int bps = (int)((n + 255) / 256);
    const long long blocks = (long long)(bps + 7) / 8 * 8 * a.nshots;
    hipLaunchKernelGGL(k_psi, dim3((unsigned)blocks), dim3(256), 0, st, a, nx_threads, nsegz, bps);
    return (int)hipGetLastError();
}
int psi_blocks_per_shot(int nzp, int nxp, int N) {
    if (N <= 4) return 0;
    long long nx_threads, n;
    int nsegz;
    psi_threads(nzp, nxp, N, &nx_threads, &nsegz, &n);
    return (int)((n + 255) / 256);
}
int launch_psi_multi(const StepArgs *list, const int *pboff, int ngrid, int npblocks, int k, hipStream_t st) {
    if (npblocks <= 0) return 0;
    hipLaunchKernelGGL(k_psi_multi, dim3((unsigned)npblocks), dim3(256), 0, st, list, pboff, ngrid, k);
    return (int)hipGetLastError();
}

int launch_adj_step(int phys, const AdjArgs &a, hipStream_t st) {
    const long long tiles = (long long)((a.nxp + 63) / 64) * ((a.nzp + 15) / 16) * a.nshots;
    const dim3 g((unsigned)((tiles + 7) / 8 * 8));  // a multiple of 8 for the XCD remap
    if (phys == 1) hipLaunchKernelGGL(k_adj_fused<1>, g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_adj_fused<2>, g, dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

int launch_expand(const float *m, int nz, int nx, int N, float *out, int P, hipStream_t st) {
    hipLaunchKernelGGL(k_expand, dim3((P + 255) / 256, nz + 2 * N), dim3(256), 0, st, m, nz, nx, N, out, P);
    return (int)hipGetLastError();
}
int launch_copy_padded(const float *m, int nzp, int nxp, float *out, int P, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_padded, dim3((P + 255) / 256, nzp), dim3(256), 0, st, m, nzp, nxp, out, P);
    return (int)hipGetLastError();
}
int launch_pack_padded(const float *U, int nzp, int nxp, int P, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_padded, dim3((nxp + 255) / 256, nzp), dim3(256), 0, st, U, nzp, nxp, P, out);
    return (int)hipGetLastError();
}
int launch_coeffs(int phys, const float *vp, const float *th, float dt, int nzp, int nxp, int P, float *v2, float *cs,
                  float *sn, hipStream_t st) {
    const long long n = (long long)nzp * P;
    (void)nxp;
    hipLaunchKernelGGL(k_coeffs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, phys, vp, th, dt, n, v2, cs, sn);
    return (int)hipGetLastError();
}
int launch_cpml_coeffs(const float *vp, int nzp, int nxp, int P, int N, float dx, float dz, float dt, float d0,
                       float fp, float *ax, float *bx, float *az, float *bz, hipStream_t st) {
    hipLaunchKernelGGL(k_cpml_coeffs, dim3((P + 255) / 256, nzp), dim3(256), 0, st, vp, nzp, nxp, P, N, dx, dz, dt,
                       d0, fp, ax, bx, az, bz);
    return (int)hipGetLastError();
}
int launch_inject_point(float *U, long long sstride, const int *ssz, const int *ssx, int P, const float *wav, int k,
                        int nshots, hipStream_t st) {
    hipLaunchKernelGGL(k_inject_point, dim3((nshots + 63) / 64), dim3(64), 0, st, U, sstride, ssz, ssx, P, wav, k,
                       nshots);
    return (int)hipGetLastError();
}
int launch_smax0(unsigned int *smax, int nt, const int *ssz, const int *ssx, int N, int nzp, int nxp,
                 const float *wav, int nshots, float bound, hipStream_t st) {
    hipLaunchKernelGGL(k_smax0, dim3((nshots + 63) / 64), dim3(64), 0, st, smax, nt, ssz, ssx, N, nzp, nxp, wav,
                       nshots, bound);
    return (int)hipGetLastError();
}
int launch_inject_rec(float *U, long long sstride, int P, const int *rz, const int *rx, int nrec, const float *data,
                      long long dstride, int row, float dxdz, int nshots, hipStream_t st) {
    hipLaunchKernelGGL(k_inject_rec, dim3((nshots + 63) / 64), dim3(64), 0, st, U, sstride, P, rz, rx, nrec, data,
                       dstride, row, dxdz, nshots);
    return (int)hipGetLastError();
}
int launch_reduce_shots(const float *parts, long long sstride, int nshots, int n, float *out, int accumulate,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_shots, dim3((n + 255) / 256), dim3(256), 0, st, parts, sstride, nshots, (long long)n,
                       out, accumulate);
    return (int)hipGetLastError();
}
int launch_reduce_shots_pad(const float *parts, long long sstride, int nshots, int nz, int nx, int P, int N,
                            float *out, int accumulate, hipStream_t st) {
    const long long n = (long long)nz * nx;
    hipLaunchKernelGGL(k_reduce_shots_pad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, parts, sstride, nshots,
                       nz, nx, P, N, out, accumulate);
    return (int)hipGetLastError();
}
int launch_maxabs(const float *x, int n, float *out, hipStream_t st) {
    (void)hipMemsetAsync(out, 0, sizeof(float), st);
    int blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_maxabs, dim3(blocks), dim3(256), 0, st, x, (long long)n, reinterpret_cast<unsigned int *>(out));
    return (int)hipGetLastError();
}
int launch_normalize(float *image, const float *ilum, int n, int mode, const float *maxv, hipStream_t st) {
    hipLaunchKernelGGL(k_normalize, dim3((n + 255) / 256), dim3(256), 0, st, image, ilum, (long long)n, mode,
                       reinterpret_cast<const unsigned int *>(maxv));
    return (int)hipGetLastError();
}
int misfit_parts(long long n) {
    long long blocks = (n + 255) / 256;
    if (blocks > kMisfitParts) blocks = kMisfitParts;
    if (blocks < 1) blocks = 1;
    return (int)blocks;
}
int launch_residual(const float *dobs, const float *dcal, float *res, long long n, double *part, double *out,
                    hipStream_t st) {
    const int blocks = misfit_parts(n);
    hipLaunchKernelGGL(k_residual, dim3((unsigned)blocks), dim3(256), 0, st, dobs, dcal, res, n, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (const double *)part, blocks, out);
    return (int)hipGetLastError();
}
int launch_sumsq(const float *x, long long n, double *part, double *out, hipStream_t st) {
    const int blocks = misfit_parts(n);
    hipLaunchKernelGGL(k_sumsq, dim3((unsigned)blocks), dim3(256), 0, st, x, n, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (const double *)part, blocks, out);
    return (int)hipGetLastError();
}
int launch_cerjan_scale(float *U, int nzp, int nxp, int P, const float *cerx, const float *cerz, hipStream_t st) {
    hipLaunchKernelGGL(k_cerjan_scale, dim3((nxp + 255) / 256, nzp), dim3(256), 0, st, U, nzp, nxp, P, cerx, cerz);
    return (int)hipGetLastError();
}
int launch_ilum_acc(const float *S, int P, int N, int nz, int nx, float *ilum, hipStream_t st) {
    hipLaunchKernelGGL(k_ilum_acc, dim3((nx + 255) / 256, nz), dim3(256), 0, st, S, P, N, nz, nx, ilum);
    return (int)hipGetLastError();
}
int launch_nonfinite(const float *x, long long n, int *flag, hipStream_t st) {
    long long blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_nonfinite, dim3((unsigned)blocks), dim3(256), 0, st, x, n, flag);
    return (int)hipGetLastError();
}
int launch_grad_multi(int phys, const float *S, int sP, long long ss, const float *lam, int lP, long long ls,
                      int nshots, int nz, int nx, float dx, float dz, const float *eps, const float *del,
                      const float *theta, int mP, float *geps, float *gdel, float *gth, hipStream_t st) {
    dim3 g((nx + 255) / 256, nz, nshots);
    if (phys == 1)
        hipLaunchKernelGGL(k_grad_multi<1>, g, dim3(256), 0, st, S, sP, ss, lam, lP, ls, nz, nx, dx, dz, eps, del,
                           theta, mP, geps, gdel, gth);
    else
        hipLaunchKernelGGL(k_grad_multi<2>, g, dim3(256), 0, st, S, sP, ss, lam, lP, ls, nz, nx, dx, dz, eps, del,
                           theta, mP, geps, gdel, gth);
    return (int)hipGetLastError();
}

}  // namespace pam
