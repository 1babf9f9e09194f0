// pam_tile.cuh -- the LDS-tiled fused time step k_tile.
//
// One 256-thread block advances one TX x TZ tile (TX = 64 columns) of one
// shot by one time step.  The block first issues every global load it needs
// at once -- the current field U_k over the tile plus a 4-cell halo (into
// LDS), and its own cells' U_{k-1} and model coefficients (into registers) --
// so that each wavefront keeps many loads in flight, then computes from LDS.
// Each thread owns 4 consecutive columns (one float4) of R = TZ/16
// consecutive rows and keeps a register window of the R+8 rows its
// z-derivatives need; x-derivatives come from the LDS row with its halo.
// TTI also stages the first x-derivative of every tile row (with the z-halo)
// in LDS, so that the mixed derivative is d/dz of it.
//
// The rest of the step is fused into the same pass: Cerjan damping (the
// damping the reference applies to U_k after the update is re-applied when
// U_k is read back as U_{k-1}), the CPML zeta update with the psi written by
// k_psi this step, the source injection of the next step (added to the value
// written), the receiver gather of this step, and for the adjoint passes the
// data injection, snapshots, illumination and imaging / gradient terms.
//
// Blocks are ordered tile-major with the shots of a tile adjacent, so the
// model coefficients of a tile are fetched from HBM once and reused from L2
// by the other shots; tiles run in bands of eight tile columns so that the
// halo rows a tile shares with its z-neighbour are reused from L2 too; the
// XCD remap gives each XCD a contiguous range.
#pragma once
#include "pam_step.cuh"

namespace pam {

// The per-step arguments of one tile pass: step, flags, the two wavefields
// and the new psi (k_tile takes them from its StepArgs; the group kernel
// trades the buffers by step parity)
struct TileStep {
    int k, flags;
    const float *cur;
    float *prv;
    const float *psix, *psiz;
};

// Tile of block b: tiles in bands of kBand tile columns, x fastest within a
// band, then z (the z-neighbour of a tile, which shares 8 halo rows with it,
// comes kBand tiles later, while those rows are still in L2, and the
// x-neighbour right after it), the shots of a tile adjacent.
template <int TZ>
__device__ __forceinline__ void tile_of(const StepArgs &a, const int b, int &x0, int &z0, int &s) {
    constexpr int kBand = 8;
    const int tile = b / a.nshots;
    s = b - tile * a.nshots;
    const int band = tile / (kBand * a.nzt);
    const int rem = tile - band * (kBand * a.nzt);
    const int wb = min(kBand, a.nxt - band * kBand);
    const int tzi = (wb == kBand) ? rem / kBand : rem / wb;
    const int txi = band * kBand + (rem - tzi * wb);
    x0 = txi * 64;
    z0 = tzi * TZ;
}

// true when the 64 x TZ tile at (x0, z0) lies max(N, 4) cells or more inside
// every edge: no rims, no absorbing layer (Cerjan factors 1, no CPML strips),
// no rows or columns beyond the grid
template <int TZ>
__device__ __forceinline__ bool tile_inner(const StepArgs &a, int x0, int z0) {
    const int m = max(a.N, 4);
    return x0 >= m && x0 + 64 <= a.nxp - m && z0 >= m && z0 + TZ <= a.nzp - m;
}

// The derivatives of one row of 4 cells: zw[0..8] the 9-row z-window of U_k
// at the cells' columns (the row itself at zw[4]), d2x / d1x the undivided x
// derivatives; c0 = U_k, pzz / pxx second and pz / px first derivatives
// (pz / px only for VTI / TTI, else 0).  Shared by the tile bodies, so that
// every instance rounds the same way.
template <int PHYS>
__device__ __forceinline__ void tile_derivs(const float (*zw)[4], const float *d2x, const float *d1x,
                                            const StepArgs &a, float *c0, float *pzz, float *pxx, float *pz,
                                            float *px) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float cc = zw[4][j];
        const float u1p = zw[5][j], u1m = zw[3][j];
        const float u2p = zw[6][j], u2m = zw[2][j];
        const float u3p = zw[7][j], u3m = zw[1][j];
        const float u4p = zw[8][j], u4m = zw[0][j];
        c0[j] = cc;
        pzz[j] = (KC0 * cc + KC1 * (u1p + u1m) + KC2 * (u2p + u2m) + KC3 * (u3p + u3m) + KC4 * (u4p + u4m)) * a.idz2;
        pxx[j] = d2x[j] * a.idx2;
        px[j] = 0.f;
        pz[j] = 0.f;
        if (PHYS >= 1) {
            pz[j] = (KA1 * (u1p - u1m) + KA2 * (u2p - u2m) + KA3 * (u3p - u3m) + KA4 * (u4p - u4m)) * a.idz;
            px[j] = d1x[j] * a.idx;
        }
    }
}

// One block's work: block b (in [0, nxt * nzt * nshots)) of the launch
// described by a and the step arguments m.  INNER: the tile satisfies
// tile_inner, so the per-cell edge tests and the absorbing-layer terms drop
// out (the values are the same bits: the factors there are exactly 1).  FM:
// the flags this instance implements (the step's flags are masked by it, so
// the code of the others drops out).
template <int PHYS, bool CPML, int TZ, int FM, bool INNER = false>
__device__ __forceinline__ void tile_body_at(const StepArgs &a, const int x0, const int z0, const int s,
                                             const TileStep m);
template <int PHYS, bool CPML, int TZ, int FM, bool INNER = false>
__device__ __forceinline__ void tile_body(const StepArgs &a, const int b, const TileStep m) {
    int x0, z0, s;
    tile_of<TZ>(a, b, x0, z0, s);
    tile_body_at<PHYS, CPML, TZ, FM, INNER>(a, x0, z0, s, m);
}
template <int PHYS, bool CPML, int TZ, int FM, bool INNER>
__device__ __forceinline__ void tile_body_at(const StepArgs &a, const int x0, const int z0, const int s,
                                             const TileStep m) {
    constexpr int TX = 64;
    constexpr int R = TZ / 16;        // rows per thread
    constexpr int ROWS = TZ + 8;      // staged rows (4-row halo each side)
    constexpr int Q = (TX + 8) / 4;   // float4 per staged row (4-cell halo each side)
    constexpr int NL = (ROWS * Q + 255) / 256;
    __shared__ float4 su[ROWS * Q];
    __shared__ float4 sd[(PHYS == 2) ? ROWS * (TX / 4) : 1];

    const int P = a.P, nzp = a.nzp, nxp = a.nxp, N = a.N;
    const long long so = (long long)s * a.sstride;
    const float *__restrict__ cur = m.cur + so;
    float *__restrict__ prv = m.prv + so;
    const float *__restrict__ psix = m.psix + so;
    const float *__restrict__ psiz = m.psiz + so;
    float *__restrict__ zetax = a.zetax + so;
    float *__restrict__ zetaz = a.zetaz + so;
    const int tid = threadIdx.x;
    const int flags = m.flags & FM;
    const int cx = tid & 15, rg = tid >> 4;
    const int col = x0 + 4 * cx;   // first of this thread's 4 columns
    const int zr0 = z0 + rg * R;   // first of this thread's R rows

    // ---- issue every load: the U_k tile with halo, U_{k-1} and the model
    float4 st[NL];
#pragma unroll
    for (int u = 0; u < NL; ++u) {
        const int i = tid + 256 * u;
        const int r = i / Q, q = i - r * Q;
        const int gz = z0 - 4 + r, gc = x0 - 4 + 4 * q;
        st[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < ROWS * Q && (INNER || (gz >= 0 && gz < nzp && gc >= 0 && gc < P)))
            st[u] = *reinterpret_cast<const float4 *>(at(cur, (unsigned)(gz * P + gc)));
    }
    float pv[R][4], v2[R][4], e[R][4], dl[R][4], cs[R][4], sn[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int z = zr0 + r;
        const bool zin = INNER || z < nzp;
        const unsigned ro = (unsigned)((zin ? z : 0) * P + col);
#pragma unroll
        for (int j = 0; j < 4; ++j) pv[r][j] = v2[r][j] = e[r][j] = dl[r][j] = cs[r][j] = sn[r][j] = 0.f;
        if (zin) {
            vld<4>(pv[r], at(prv, ro));
            vld<4>(v2[r], at(a.v2, ro));
            if (PHYS >= 1) {
                vld<4>(e[r], at(a.eps, ro));
                vld<4>(dl[r], at(a.del, ro));
            }
            if (PHYS == 2) {
                vld<4>(cs[r], at(a.cs, ro));
                vld<4>(sn[r], at(a.sn, ro));
            }
        }
    }
    // the adjoint passes' padded accumulator and snapshot cells, loaded with
    // the fields so that their latency overlaps the update (the gradient pass
    // took 71.7 instead of 61.3 us per 13-shot launch on a 451 x 1801 VTI grid
    // when they were loaded after it; the illumination pass was faster without)
    constexpr bool kEarly = FM == F_PASS_GRAD || FM == F_PASS_IMAGE;
    constexpr bool kEarlyS = kEarly && (FM & (F_IMAGE | F_GRADM)) != 0;
    float eacc[kEarly ? R : 1][4], esv[kEarlyS ? R : 1][4];
    if (kEarly && a.acc_pad) {
        const long long acc_s = (long long)s * a.acc_sstride;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int z = zr0 + r;
            if (z >= N && z < nzp - N && col + 3 >= N && col < nxp - N) {
                const unsigned ro = (unsigned)(z * P + col);
                if (FM & F_ILUM) {
                    vld<4>(eacc[r], at(a.ilum + acc_s, ro));
                } else {
                    vld<4>(eacc[r], at(a.image + acc_s, ro));
                    if (kEarlyS) snap_get4_pad(a, snap_off(a, s, m.k), s, m.k, ro, esv[kEarlyS ? r : 0]);
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NL; ++u) {
        const int i = tid + 256 * u;
        if (i < ROWS * Q) su[i] = st[u];
    }
    __syncthreads();

    if (PHYS == 2) {
        // first x-derivative (undivided) of every staged row
        for (int i = tid; i < ROWS * (TX / 4); i += 256) {
            const int r = i >> 4, c = i & 15;
            float w[12];
            const float4 l = su[r * Q + c], m = su[r * Q + c + 1], h = su[r * Q + c + 2];
            w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
            w[4] = m.x; w[5] = m.y; w[6] = m.z; w[7] = m.w;
            w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
            float d2[4], d1[4];
            xderivs<4>(w, d2, d1);
            sd[i] = make_float4(d1[0], d1[1], d1[2], d1[3]);
        }
        __syncthreads();
    }

    // ---- z-window of U_k at this thread's columns: local rows rg*R-4 .. rg*R+R+3
    float zw[R + 8][4];
#pragma unroll
    for (int i = 0; i < R + 8; ++i) {
        const float4 t = su[(rg * R + i) * Q + cx + 1];
        zw[i][0] = t.x; zw[i][1] = t.y; zw[i][2] = t.z; zw[i][3] = t.w;
    }

    bool xlive[4], xs[4];
    float fx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = col + j;
        xlive[j] = INNER || ((c >= 4) && (c < nxp - 4));
        fx[j] = (!INNER && !CPML && (flags & F_CERJAN) && c < P) ? a.cerx[c] : 1.0f;
        xs[j] = !INNER && CPML && ((c >= 4 && c < N) || (c >= nxp - N && c < nxp - 4));
    }
    const bool tile_xstrip = !INNER && CPML && (x0 < N || x0 + TX > nxp - N);
    float tmax = 0.0f;  // F_SMAX: max |new field| over this thread's interior cells

#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int z = zr0 + r;
        if (!INNER && z >= nzp) break;  // rows beyond the grid (last tile row only)
        const bool zlive = INNER || ((z >= 4) && (z < nzp - 4));
        const unsigned ro = (unsigned)(z * P + col);  // 32-bit lane offsets: SGPR-base addressing
        const float fz = (!INNER && !CPML && (flags & F_CERJAN)) ? a.cerz[z] : 1.0f;
        const int lr = rg * R + r + 4;  // staged row of z

        // x-window of row z
        float w[12];
        {
            const float4 l = su[lr * Q + cx], m = su[lr * Q + cx + 1], h = su[lr * Q + cx + 2];
            w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
            w[4] = m.x; w[5] = m.y; w[6] = m.z; w[7] = m.w;
            w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
        }
        float d2x[4], d1x[4];
        xderivs<4>(w, d2x, d1x);

        float c0[4], pzz[4], pxx[4], px[4], pz[4], pxz[4];
        tile_derivs<PHYS>(zw + r, d2x, d1x, a, c0, pzz, pxx, pz, px);
#pragma unroll
        for (int j = 0; j < 4; ++j) pxz[j] = 0.f;
        if (PHYS == 2) {
            float q[9][4];
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                if (h == 4) continue;
                const float4 t = sd[(lr - 4 + h) * (TX / 4) + cx];
                q[h][0] = t.x; q[h][1] = t.y; q[h][2] = t.z; q[h][3] = t.w;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                pxz[j] = (KA1 * (q[5][j] - q[3][j]) + KA2 * (q[6][j] - q[2][j]) + KA3 * (q[7][j] - q[1][j]) +
                          KA4 * (q[8][j] - q[0][j])) *
                         a.idxdz;
        }

        // ---- receiver gather of the input field (damped as the reference records it)
        if ((flags & F_GATHER) && a.rrow[z] >= 0) {
            const int slot = a.rrow[z];
            const int it = m.k - a.itlag;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int rr = a.rhead[(long long)slot * P + col + j];
                const float val = (c0[j] * fx[j]) * fz;
                while (rr >= 0) {
                    a.seis[(long long)s * a.dstride + (long long)it * a.nrec + rr] = val;
                    rr = a.rnext[rr];
                }
            }
        }

        if (!zlive) continue;

        // ---- CPML: zeta with the new psi, corrected derivatives
        float PXX[4], PZZ[4], PX[4], PZ[4], zx[4], zz[4], dpx[4], dpz[4];
        const bool zs = !INNER && CPML && ((z >= 4 && z < N) || (z >= nzp - N && z < nzp - 4));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            PXX[j] = pxx[j];
            PZZ[j] = pzz[j];
            PX[j] = px[j];
            PZ[j] = pz[j];
            zx[j] = zz[j] = dpx[j] = dpz[j] = 0.0f;
        }
        if (CPML && tile_xstrip) {
            float pw[12], d2p[4], d1p[4], zo[4], ax[4], bx[4];
            const float *prow = psix + (unsigned)(z * P);
            vldg<4>(pw + 4, prow, col, P);
            if (col >= 4) vld<4>(pw, prow + col - 4);
            else pw[0] = pw[1] = pw[2] = pw[3] = 0.f;
            vldg<4>(pw + 8, prow, col + 4, P);
            xderivs<4>(pw, d2p, d1p);
            vld<4>(zo, at(zetax, ro));
            vld<4>(ax, at(a.cax, ro));
            vld<4>(bx, at(a.cbx, ro));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!xs[j]) continue;
                const float dps = d1p[j] * a.idx;
                zx[j] = ax[j] * zo[j] + bx[j] * (pxx[j] + dps);
                dpx[j] = dps;
                PX[j] = px[j] + dps;
                PXX[j] = pxx[j] + dps + zx[j];
                *at(zetax, ro + j) = zx[j];
            }
        }
        if (CPML && zs) {
            float ps[9][4];
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                if (h == 4) continue;
                vld<4>(ps[h], at(psiz, (unsigned)((z - 4 + h) * P + col)));
            }
            float zo[4], az[4], bz[4];
            vld<4>(zo, at(zetaz, ro));
            vld<4>(az, at(a.caz, ro));
            vld<4>(bz, at(a.cbz, ro));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float dps = (KA1 * (ps[5][j] - ps[3][j]) + KA2 * (ps[6][j] - ps[2][j]) +
                                   KA3 * (ps[7][j] - ps[1][j]) + KA4 * (ps[8][j] - ps[0][j])) *
                                  a.idz;
                zz[j] = az[j] * zo[j] + bz[j] * (pzz[j] + dps);
                dpz[j] = dps;
                PZ[j] = pz[j] + dps;
                PZZ[j] = pzz[j] + dps + zz[j];
                if (xlive[j]) *at(zetaz, ro + j) = zz[j];
            }
        }

        // ---- the update
        float out[4], opv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool want_op = (flags & F_D2T_OP) != 0;
            float u;
            if (PHYS == 0) {
                float sum = pxx[j] + pzz[j];
                if (CPML) {
                    if (xs[j]) sum += dpx[j];
                    if (zs) sum += dpz[j];
                    if (xs[j]) sum += zx[j];
                    if (zs) sum += zz[j];
                }
                u = lf_acoustic(c0[j], pv[r][j], fx[j], fz, v2[r][j], sum, want_op, opv[j]);
            } else if (PHYS == 1) {
                u = lf_vti(c0[j], pv[r][j], fx[j], fz, v2[r][j], e[r][j], dl[r][j], PX[j], PZ[j], PXX[j], PZZ[j],
                           want_op, opv[j]);
            } else {
                u = lf_tti(c0[j], pv[r][j], fx[j], fz, v2[r][j], e[r][j], dl[r][j], cs[r][j], sn[r][j], PX[j], PZ[j],
                           PXX[j], PZZ[j], pxz[j], want_op, opv[j]);
            }
            out[j] = (u * fx[j]) * fz;
        }
        // writer-side injection of the next step's source
        if ((flags & F_SRC) && z == a.ssz[s] && m.k + 1 < a.nt) {
            const int sx = a.ssx[s];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (col + j == sx) out[j] += a.wav[m.k + 1];
        }
        // writer-side injection of the data of the next (earlier) adjoint step
        if ((flags & F_RECINJ) && a.rrow[z] >= 0 && m.k - 1 - a.itlag >= 0) {
            const int slot = a.rrow[z];
            const int tn = m.k - 1 - a.itlag;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int rr = a.rhead[(long long)slot * P + col + j];
                while (rr >= 0) {
                    out[j] += a.data[(long long)s * a.dstride + (long long)tn * a.nrec + rr] / a.dxdz;
                    rr = a.rnext[rr];
                }
            }
        }
        if (a.acc_pad && (flags & F_D2T) && (flags & F_SAVE)) {
            // padded snapshot: D_k of the interior cells (0 elsewhere), four cells in one
            // access; with F_SMAX (16-bit kinds) also max |D_k| over them
            if (z >= N && z < nzp - N && col + 3 >= N && col < nxp - N) {
                float d4[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    d4[j] = 0.0f;
                    if (col + j >= N && col + j < nxp - N) {
                        d4[j] = (flags & F_D2T_OP) ? opv[j] : out[j] - 2.0f * c0[j] + pv[r][j];
                        if (flags & F_SMAX) tmax = fmaxf(tmax, fabsf(d4[j]));
                    }
                }
                snap_put4_pad(a, snap_off(a, s, m.k), s, m.k, ro, d4);
            }
        } else if ((flags & F_SMAX) || ((flags & F_D2T) && (flags & F_SAVE)) || (flags & F_DSAVE)) {
            const bool zint = z >= N && z < nzp - N;
            const int ib = (z - N) * a.nx - N;  // interior index = ib + c (used on interior cells only)
            const long long sk = snap_off(a, s, m.k);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!(zint && col + j >= N && col + j < nxp - N)) continue;
                // interior cells are undamped: D_k = S_{k+1} - 2 S_k + S_{k-1}
                const float d = (flags & F_D2T_OP) ? opv[j] : out[j] - 2.0f * c0[j] + pv[r][j];
                const unsigned ii = (unsigned)(ib + col + j);
                if ((flags & F_D2T) && (flags & F_SAVE)) snap_put_at(a, sk, s, m.k, ii, 0u, d);
                if (flags & F_DSAVE) *at(a.dsave + (long long)s * a.dsave_sstride, ii) = d;
                tmax = fmaxf(tmax, fabsf((flags & F_D2T) ? d : out[j]));
            }
        }
        if (INNER || (xlive[0] && xlive[3])) {
            vst<4>(at(prv, ro), out);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (xlive[j]) *at(prv, ro + j) = out[j];
        }
        // ---- snapshots / illumination / imaging of the input field (after the update:
        // fewer values are live here than before it)
        if (a.acc_pad && (flags & (F_SAVE | F_ILUM | F_IMAGE | F_GRADM)) && zlive) {
            // padded accumulators and snapshots: one aligned float4 per thread
            // (the gradient reads D_k: a padded run always stores D_k)
            const bool zint = (z >= N) && (z < nzp - N);
            bool cin[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) cin[j] = zint && col + j >= N && col + j < nxp - N;
            const long long acc_s = (long long)s * a.acc_sstride;
            const long long sk = snap_off(a, s, m.k);
            if ((flags & F_SAVE) && !(flags & F_D2T)) {
                float S4[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) S4[j] = cin[j] ? c0[j] : 0.0f;
                if (zint) snap_put4_pad(a, sk, s, m.k, ro, S4);
            }
            if (cin[0] || cin[1] || cin[2] || cin[3]) {
                if (flags & F_ILUM) {
                    float v[4];
                    if (kEarly) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) v[j] = eacc[kEarly ? r : 0][j];
                    } else {
                        vld<4>(v, at(a.ilum + acc_s, ro));
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (cin[j]) v[j] += c0[j] * c0[j];
                    vst<4>(at(a.ilum + acc_s, ro), v);
                }
                if (flags & (F_IMAGE | F_GRADM)) {
                    float S4[4], v[4];
                    if (kEarlyS) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            S4[j] = esv[kEarlyS ? r : 0][j];
                            v[j] = eacc[kEarly ? r : 0][j];
                        }
                    } else {
                        snap_get4_pad(a, sk, s, m.k, ro, S4);
                        vld<4>(v, at(a.image + acc_s, ro));
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (!cin[j]) continue;
                        if (flags & F_IMAGE) v[j] += S4[j] * c0[j];
                        if (flags & F_GRADM) {
                            const float d2t = S4[j] * a.idt2;
                            v[j] += d2t * c0[j];
                        }
                    }
                    vst<4>(at(a.image + acc_s, ro), v);
                }
            }
        } else if ((flags & (F_SAVE | F_ILUM | F_IMAGE | F_GRADM)) && zlive) {
            const bool zint = (z >= N) && (z < nzp - N);
            const int ib = (z - N) * a.nx - N;  // interior index = ib + c (used on interior cells only)
            const long long acc_s = (long long)s * a.acc_sstride;  // per-shot bases: wave-uniform
            const long long sk = snap_off(a, s, m.k);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = col + j;
                const bool cin = zint && c >= N && c < nxp - N;
                const unsigned ii = (unsigned)(ib + c);
                if ((flags & F_SAVE) && !(flags & F_D2T)) {
                    if (a.save_kind == SV_F32_PAD) {
                        if (xlive[j]) snap_put_at(a, sk, s, m.k, 0u, ro + j, (c0[j] * fx[j]) * fz);
                    } else if (cin) {
                        snap_put_at(a, sk, s, m.k, ii, 0u, c0[j]);
                    }
                }
                if (!cin) continue;
                if (flags & F_ILUM) {
                    float *acc = at(a.ilum + acc_s, ii);
                    *acc += c0[j] * c0[j];
                }
                if (flags & F_IMAGE) {
                    const float S = snap_get_at(a, sk, s, m.k, ii, ro + j);
                    float *im = at(a.image + acc_s, ii);
                    *im += S * c0[j];
                }
                if (flags & F_GRADM) {
                    float d2t;
                    if (flags & F_D2T) {
                        d2t = snap_get_at(a, sk, s, m.k, ii, ro + j) * a.idt2;
                    } else {
                        const float Sp = snap_get_at(a, snap_off(a, s, m.k + 1), s, m.k + 1, ii, ro + j);
                        const float S0 = snap_get_at(a, sk, s, m.k, ii, ro + j);
                        const float Sm = snap_get_at(a, snap_off(a, s, m.k - 1), s, m.k - 1, ii, ro + j);
                        d2t = (Sp - 2.0f * S0 + Sm) * a.idt2;
                    }
                    float *im = at(a.image + acc_s, ii);
                    *im += d2t * c0[j];
                }
            }
        }
    }
    if ((flags & F_SMAX) && m.k + 1 < a.nt) block_max_to(a.smax + ((long long)s * a.nt + m.k + 1) * kSmaxSlots + blockIdx.x % kSmaxSlots, tmax);
}

// the forward, illumination and recompute instances of the 16-row tile fit 5
// waves per SIMD; the adjoint ones, which load their accumulators early, run
// at 4 (106 VGPRs for the VTI gradient: at 5 it spilled 27), as does the
// instance for every flag
template <int PHYS, bool CPML, int TZ, int WPE, int FM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(((FM == F_FWD_MASK || FM == F_PASS_ILUM || FM == F_PASS_SAVE) && TZ == 16 && PHYS < 2) ? 5 : WPE))) void k_tile(
    const StepArgs a) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= a.nxt * a.nzt * a.nshots) return;  // block-uniform
    const TileStep m{a.k, a.flags, a.cur, a.prv, a.psix, a.psiz};
    if constexpr (FM != F_ALL) {
        int x0, z0, s;
        tile_of<TZ>(a, b, x0, z0, s);
        if (tile_inner<TZ>(a, x0, z0)) {  // block-uniform
            tile_body<PHYS, CPML, TZ, FM, true>(a, b, m);
            return;
        }
    }
    tile_body<PHYS, CPML, TZ, FM>(a, b, m);
}

// ------------------------------------------------ the pipelined 16-row tile
// Interior tiles (no rim, no absorbing layer) of a launch are taken by blocks
// that advance one tile for a group of SG shots in turn, the shots of the
// group being the tile's consecutive shots.  The model coefficients of the
// tile are loaded once for the group.  While shot i is computed, the U_k tile
// of shot i+1 streams from HBM straight into the other of two LDS buffers
// (gfx950 global_load_lds_dwordx4: no registers hold it in flight) and its
// U_{k-1} into registers; shot i's result is stored after that prefetch is
// issued, so the one wait per shot (all memory operations of the wave, then
// the block barrier) only waits on operations issued before the shot's
// arithmetic.  The other tiles run the general body, one (tile, shot) per
// block, in the same launch.  Each cell is computed by the same arithmetic as
// in k_tile, so the results are those of variant 0 bit for bit.
typedef __attribute__((address_space(3))) void lds_void;

// the interior-tile rectangle [ix0, ix1) x [iz0, iz1) of the 16-row tile grid
// (tile_inner<16> holds exactly there), and the edge tiles around it
struct TileSplit {
    int ix0, ix1, iz0, iz1, nin, nedge;
};
__host__ __device__ inline TileSplit tile_split(int nxp, int nzp, int N, int nxt, int nzt) {
    const int m = N > 4 ? N : 4;
    TileSplit t;
    t.ix0 = (m + 63) / 64;
    t.ix1 = (nxp - m) / 64;
    t.iz0 = (m + 15) / 16;
    t.iz1 = (nzp - m) / 16;
    if (t.ix1 < t.ix0) t.ix1 = t.ix0;
    if (t.iz1 < t.iz0) t.iz1 = t.iz0;
    t.nin = (t.ix1 - t.ix0) * (t.iz1 - t.iz0);
    t.nedge = nxt * nzt - t.nin;
    return t;
}
// interior tile i (bands of 8 tile columns, x fastest within a band, as tile_of)
__device__ __forceinline__ void inner_tile_xz(const TileSplit &t, int i, int &x0, int &z0) {
    constexpr int kBand = 8;
    const int nw = t.ix1 - t.ix0, nh = t.iz1 - t.iz0;
    const int band = i / (kBand * nh);
    const int rem = i - band * (kBand * nh);
    const int wb = min(kBand, nw - band * kBand);
    const int tzi = rem / wb;
    x0 = (t.ix0 + band * kBand + (rem - tzi * wb)) * 64;
    z0 = (t.iz0 + tzi) * 16;
}
// edge tile i: the rows above the rectangle, the rows below it, then the
// left and right parts of the rows beside it.  Above and below, the tiles run
// in bands of 4 tile columns, x fastest within a band, so that a tile's
// z-neighbours (which share its halo rows) come a few tiles later, on the
// same XCD.
__device__ __forceinline__ void band_xz(int e, int nxt, int rows, int &txi, int &tzi) {
    constexpr int kEdgeBand = 4;
    const int band = e / (kEdgeBand * rows);
    const int rem = e - band * (kEdgeBand * rows);
    const int wb = min(kEdgeBand, nxt - band * kEdgeBand);
    tzi = rem / wb;
    txi = band * kEdgeBand + (rem - tzi * wb);
}
__device__ __forceinline__ void edge_tile_xz(const TileSplit &t, int nxt, int nzt, int i, int &x0, int &z0) {
    const int top = t.iz0 * nxt, bottom = (nzt - t.iz1) * nxt;
    int txi, tzi;
    if (i < top) {
        band_xz(i, nxt, t.iz0, txi, tzi);
    } else if (i < top + bottom) {
        band_xz(i - top, nxt, nzt - t.iz1, txi, tzi);
        tzi += t.iz1;
    } else {
        const int e = i - top - bottom;
        const int w = t.ix0 + nxt - t.ix1;
        const int r = e / w, c = e - r * w;
        tzi = t.iz0 + r;
        txi = c < t.ix0 ? c : t.ix1 + (c - t.ix0);
    }
    x0 = txi * 64;
    z0 = tzi * 16;
}

// s_barrier between two compiler barriers: no instruction, but no memory access
// is moved across (the caller waits for the accesses the barrier orders)
__device__ __forceinline__ void pipe_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// one interior tile for shots [s0, s0 + ns) (0 < ns <= SG), forward flags only.
// The U_k tile (with its halo) and the U_{k-1} cells of the next shot are
// streamed into LDS while a shot is computed (two buffers), so that no load
// in the loop writes a register the arithmetic waits for.  The buffers are
// distinct variables used in turn by an unrolled loop, so that the compiler
// can tell that the reads of one do not alias the DMA into the other (it
// would otherwise wait for that DMA before every read).
constexpr int kPipeNQ = 24 * 18;         // float4 of a staged 16-row tile with its 4-cell halo
constexpr int kPipeNB = kPipeNQ + 256;   // + U_{k-1} of the 16 x 64 cells, one float4 per thread
// The two LDS buffers of the pipelined kernels, shared by their interior and
// (fused-psi) edge blocks
static __shared__ float4 g_pipe_b0[kPipeNB], g_pipe_b1[kPipeNB];

template <int PHYS, int FM, int SG>
__device__ __forceinline__ void pipe_inner(const StepArgs &a, const TileStep m, const int x0, const int z0,
                                           const int s0, const int ns) {
    constexpr int TX = 64, ROWS = 24, Q = (TX + 8) / 4, NQ = kPipeNQ;
    float4 *const su0 = g_pipe_b0;
    float4 *const su1 = g_pipe_b1;
    __shared__ float4 sd[(PHYS == 2) ? ROWS * (TX / 4) : 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int flags = m.flags & FM;
    const int P = a.P;
    const int cx = tid & 15, rg = tid >> 4;
    const int col = x0 + 4 * cx;
    const int z = z0 + rg;
    const unsigned ro = (unsigned)(z * P + col);
    // the tile's model, once for the group
    float v2[4], e[4] = {0.f, 0.f, 0.f, 0.f}, dl[4] = {0.f, 0.f, 0.f, 0.f}, cs[4] = {0.f, 0.f, 0.f, 0.f},
          sn[4] = {0.f, 0.f, 0.f, 0.f};
    vld<4>(v2, at(a.v2, ro));
    if (PHYS >= 1) {
        vld<4>(e, at(a.eps, ro));
        vld<4>(dl, at(a.del, ro));
    }
    if (PHYS == 2) {
        vld<4>(cs, at(a.cs, ro));
        vld<4>(sn, at(a.sn, ro));
    }
    // shot s into LDS buffer buf: U_{k-1} of this thread's cells to slot NQ + tid,
    // then wave wv moves U_k tile slots [64 wv + 256 u, +64) (every address
    // inside the grid: the tile is at least max(N, 4) cells inside every edge).
    // Every wave issues two or three loads.
    auto stage = [&](int s, float4 *buf) {
        const long long so = (long long)s * a.sstride;
        __builtin_amdgcn_global_load_lds(at(m.prv + so, ro), (lds_void *)(buf + NQ + 64 * wv), 16, 0, 0);
        const float *cur = m.cur + so;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int base = 64 * wv + 256 * u;
            const int i = base + lane;
            if (base < NQ && i < NQ) {
                const int r = i / Q, q = i - r * Q;
                __builtin_amdgcn_global_load_lds(at(cur, (unsigned)((z0 - 4 + r) * P + x0 - 4 + 4 * q)),
                                                 (lds_void *)(buf + base), 16, 0, 0);
            }
        }
    };
    // what the shots need besides the fields, read before the loop (inside it, a
    // global load waited on would also wait for the DMA in flight): the first
    // receiver of each of this thread's cells and the one after it (chains of
    // more receivers on one cell are walked in the loop), the source sample,
    // and which of this thread's cells hold the source of shot s0 + i (bits
    // 4i .. 4i+3)
    int rfirst[4] = {-1, -1, -1, -1}, rsecond[4] = {-1, -1, -1, -1};
    const int rslot = (flags & F_GATHER) ? a.rrow[z] : -1;
    if (rslot >= 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            rfirst[j] = a.rhead[(long long)rslot * P + col + j];
            if (rfirst[j] >= 0) rsecond[j] = a.rnext[rfirst[j]];
        }
    }
    unsigned srcbits = 0;
    float wsrc = 0.f;
    if ((flags & F_SRC) && m.k + 1 < a.nt) {
        wsrc = a.wav[m.k + 1];
        for (int i = 0; i < ns; ++i) {
            if (z != a.ssz[s0 + i]) continue;
            const int sx = a.ssx[s0 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (col + j == sx) srcbits |= 1u << (4 * i + j);
        }
    }
    float prev_out[4];
    // shot s0 + i from buffer b; shot s0 + i + 1 is streamed into bn meanwhile
    auto shot = [&](const int i, const float4 *b, float4 *bn) {
        const int s = s0 + i;
        // this wave's DMA of shot s and its LDS reads of the previous shot
        __builtin_amdgcn_s_waitcnt(0x070);  // vmcnt(0) lgkmcnt(0)
        // every wave's: buffer b is filled, buffer bn is free.  No fence: a release
        // fence waits for every LDS write in flight, the DMA of later shots
        // included; the waits above order the accesses, and the empty asm
        // statements keep the compiler from moving memory accesses across
        // the barrier.
        pipe_barrier();
        if (i + 1 < ns) stage(s + 1, bn);
        if (i > 0) vst<4>(at(m.prv + (long long)(s - 1) * a.sstride, ro), prev_out);
        if (PHYS == 2) {
            for (int t = tid; t < ROWS * (TX / 4); t += 256) {
                const int r = t >> 4, c = t & 15;
                float w[12];
                const float4 l = b[r * Q + c], mm = b[r * Q + c + 1], h = b[r * Q + c + 2];
                w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
                w[4] = mm.x; w[5] = mm.y; w[6] = mm.z; w[7] = mm.w;
                w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
                float d2[4], d1[4];
                xderivs<4>(w, d2, d1);
                sd[t] = make_float4(d1[0], d1[1], d1[2], d1[3]);
            }
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the stores to sd
            pipe_barrier();
        }
        float zw[9][4];
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            const float4 t = b[(rg + h) * Q + cx + 1];
            zw[h][0] = t.x; zw[h][1] = t.y; zw[h][2] = t.z; zw[h][3] = t.w;
        }
        const int lr = rg + 4;
        float w[12];
        {
            const float4 l = b[lr * Q + cx], mm = b[lr * Q + cx + 1], h = b[lr * Q + cx + 2];
            w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
            w[4] = mm.x; w[5] = mm.y; w[6] = mm.z; w[7] = mm.w;
            w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
        }
        float pv[4];
        {
            const float4 t = b[NQ + tid];
            pv[0] = t.x; pv[1] = t.y; pv[2] = t.z; pv[3] = t.w;
        }
        float d2x[4], d1x[4];
        xderivs<4>(w, d2x, d1x);
        float c0[4], pzz[4], pxx[4], px[4], pz[4], pxz[4] = {0.f, 0.f, 0.f, 0.f};
        tile_derivs<PHYS>(zw, d2x, d1x, a, c0, pzz, pxx, pz, px);
        if (PHYS == 2) {
            float q[9][4];
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                if (h == 4) continue;
                const float4 t = sd[(lr - 4 + h) * (TX / 4) + cx];
                q[h][0] = t.x; q[h][1] = t.y; q[h][2] = t.z; q[h][3] = t.w;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                pxz[j] = (KA1 * (q[5][j] - q[3][j]) + KA2 * (q[6][j] - q[2][j]) + KA3 * (q[7][j] - q[1][j]) +
                          KA4 * (q[8][j] - q[0][j])) *
                         a.idxdz;
        }
        if (rslot >= 0) {  // interior cells are undamped
            float *row = a.seis + (long long)s * a.dstride + (long long)(m.k - a.itlag) * a.nrec;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (rfirst[j] < 0) continue;
                row[rfirst[j]] = c0[j];
                for (int rr = rsecond[j]; rr >= 0; rr = a.rnext[rr]) row[rr] = c0[j];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float op;
            if (PHYS == 0)
                prev_out[j] = lf_acoustic(c0[j], pv[j], 1.0f, 1.0f, v2[j], pxx[j] + pzz[j], false, op);
            else if (PHYS == 1)
                prev_out[j] = lf_vti(c0[j], pv[j], 1.0f, 1.0f, v2[j], e[j], dl[j], px[j], pz[j], pxx[j], pzz[j],
                                     false, op);
            else
                prev_out[j] = lf_tti(c0[j], pv[j], 1.0f, 1.0f, v2[j], e[j], dl[j], cs[j], sn[j], px[j], pz[j],
                                     pxx[j], pzz[j], pxz[j], false, op);
        }
        if ((srcbits >> (4 * i)) & 15u) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((srcbits >> (4 * i + j)) & 1u) prev_out[j] += wsrc;
        }
    };
    stage(s0, su0);
    // fully unrolled: every buffer reference is static
#pragma unroll
    for (int i = 0; i < SG; ++i)
        if (i < ns) shot(i, (i & 1) ? su1 : su0, (i & 1) ? su0 : su1);
    vst<4>(at(m.prv + (long long)(s0 + ns - 1) * a.sstride, ro), prev_out);
}

// One edge tile (x0, z0) of shot s with CPML, forward flags only, the psi of
// the step computed in the block instead of by a pre-pass: new psi_x on the
// tile's rows over columns x0-4 .. x0+67 and new psi_z on rows z0-4 .. z0+19
// (psi = a psi_prev + b dU, U_k staged with an 8-cell halo), written for the
// tile's own strip cells and kept in LDS for the zeta update, which then runs
// as in the 16-row tile.  The halo psi is recomputed from the previous step's
// psi (double-buffered), never read from a neighbour, and every value is
// formed by the same expressions as in k_psi and the 16-row tile, so the
// results are theirs bit for bit.
template <int PHYS>
__device__ __forceinline__ void edge_fused_body(const StepArgs &a, const TileStep m, const int x0, const int z0,
                                                const int s) {
    constexpr int SQ = 20;                  // float4 per staged row: columns x0-8 .. x0+71
    constexpr int SR = 32;                  // staged rows z0-8 .. z0+23
    constexpr int NS = SR * SQ;
    constexpr int NL = (NS + 255) / 256;
    constexpr int GX = 18;                  // psi_x groups of a row: columns x0-4 .. x0+67
    static_assert(NS <= kPipeNB && 16 * GX + 24 * 16 <= kPipeNB, "fused edge tile exceeds the pipe buffers");
    float4 *const su = g_pipe_b0;              // U_k
    float4 *const pxl = g_pipe_b1;             // new psi_x [16][GX]
    float4 *const pzl = g_pipe_b1 + 16 * GX;   // new psi_z [24][16], rows z0-4 .. z0+19
    __shared__ float4 sd[(PHYS == 2) ? 24 * 16 : 1];

    const int P = a.P, nzp = a.nzp, nxp = a.nxp, N = a.N;
    const long long so = (long long)s * a.sstride;
    const float *__restrict__ cur = m.cur + so;
    float *__restrict__ prv = m.prv + so;
    const float *__restrict__ pxi = a.psix_in + so;
    const float *__restrict__ pzi = a.psiz_in + so;
    float *__restrict__ pxo = a.psix + so;
    float *__restrict__ pzo = a.psiz + so;
    float *__restrict__ zetax = a.zetax + so;
    float *__restrict__ zetaz = a.zetaz + so;
    const int tid = threadIdx.x;
    const int flags = m.flags & F_FWD_MASK;
    const int cx = tid & 15, rg = tid >> 4;
    const int col = x0 + 4 * cx;
    const int z = z0 + rg;
    const bool xtile = x0 < N || x0 + 64 > nxp - N;  // the tile meets an x-strip (block-uniform)
    const bool ztile = z0 < N || z0 + 16 > nzp - N;  // its rows meet a z-strip (block-uniform)
    auto xstrip = [&](int c) { return (c >= 4 && c < N) || (c >= nxp - N && c < nxp - 4); };
    auto zstrip = [&](int r) { return (r >= 4 && r < N) || (r >= nzp - N && r < nzp - 4); };

    // ---- U_k rows z0-8 .. z0+23 with an 8-column halo (zero off the grid),
    // U_{k-1} and the model of this thread's cells
    float4 st[NL];
#pragma unroll
    for (int u = 0; u < NL; ++u) {
        const int i = tid + 256 * u;
        const int r = i / SQ, q = i - r * SQ;
        const int gz = z0 - 8 + r, gc = x0 - 8 + 4 * q;
        st[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < NS && gz >= 0 && gz < nzp && gc >= 0 && gc < P)
            st[u] = *reinterpret_cast<const float4 *>(at(cur, (unsigned)(gz * P + gc)));
    }
    float pv[4], v2[4], e[4], dl[4], cs[4], sn[4];
    const bool zin = z < nzp;
    const unsigned ro = (unsigned)((zin ? z : 0) * P + col);
#pragma unroll
    for (int j = 0; j < 4; ++j) pv[j] = v2[j] = e[j] = dl[j] = cs[j] = sn[j] = 0.f;
    if (zin) {
        vld<4>(pv, at(prv, ro));
        vld<4>(v2, at(a.v2, ro));
        if (PHYS >= 1) {
            vld<4>(e, at(a.eps, ro));
            vld<4>(dl, at(a.del, ro));
        }
        if (PHYS == 2) {
            vld<4>(cs, at(a.cs, ro));
            vld<4>(sn, at(a.sn, ro));
        }
    }
#pragma unroll
    for (int u = 0; u < NL; ++u) {
        const int i = tid + 256 * u;
        if (i < NS) su[i] = st[u];
    }
    __syncthreads();

    if (PHYS == 2) {
        // first x-derivative (undivided) of rows z0-4 .. z0+19
        for (int i = tid; i < 24 * 16; i += 256) {
            const int r = i >> 4, c = i & 15;
            const float4 *rw = su + (r + 4) * SQ + c + 1;
            float w[12], d2[4], d1[4];
            f4_to(w, rw[0]);
            f4_to(w + 4, rw[1]);
            f4_to(w + 8, rw[2]);
            xderivs<4>(w, d2, d1);
            sd[i] = make_float4(d1[0], d1[1], d1[2], d1[3]);
        }
    }
    if (xtile) {
        // new psi_x of rows z0 .. z0+15, columns x0-4 .. x0+67 (zero off the x-strips)
        for (int i = tid; i < 16 * GX; i += 256) {
            const int r = i / GX, g = i - r * GX;
            const int zz = z0 + r, c0 = x0 - 4 + 4 * g;
            float nv[4] = {0.f, 0.f, 0.f, 0.f};
            bool any = false;
#pragma unroll
            for (int j = 0; j < 4; ++j) any = any || xstrip(c0 + j);
            if (any && zz >= 4 && zz < nzp - 4) {
                const float4 *rw = su + (r + 8) * SQ + g;
                float w[12], d2[4], d1[4], pin[4], ax[4], bx[4];
                f4_to(w, rw[0]);
                f4_to(w + 4, rw[1]);
                f4_to(w + 8, rw[2]);
                xderivs<4>(w, d2, d1);
                const unsigned pi = (unsigned)(zz * P + c0);
                vld<4>(pin, at(pxi, pi));
                vld<4>(ax, at(a.cax, pi));
                vld<4>(bx, at(a.cbx, pi));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!xstrip(c0 + j)) continue;
                    nv[j] = ax[j] * pin[j] + bx[j] * (d1[j] * a.idx);
                    if (g >= 1 && g <= 16) *at(pxo, pi + j) = nv[j];
                }
            }
            pxl[i] = make_float4(nv[0], nv[1], nv[2], nv[3]);
        }
    }
    if (ztile) {
        // new psi_z of rows z0-4 .. z0+19, the tile's columns (zero off the z-strips)
        for (int i = tid; i < 24 * 16; i += 256) {
            const int r = i >> 4, c = i & 15;
            const int zz = z0 - 4 + r, cc = x0 + 4 * c;
            float nv[4] = {0.f, 0.f, 0.f, 0.f};
            if (zstrip(zz) && cc < nxp - 4) {
                float q[9][4], pin[4], az[4], bz[4];
#pragma unroll
                for (int h = 0; h < 9; ++h) {
                    if (h == 4) continue;
                    f4_to(q[h], su[(r + h) * SQ + c + 2]);
                }
                const unsigned pi = (unsigned)(zz * P + cc);
                vld<4>(pin, at(pzi, pi));
                vld<4>(az, at(a.caz, pi));
                vld<4>(bz, at(a.cbz, pi));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (cc + j < 4 || cc + j >= nxp - 4) continue;
                    const float d1 = KA1 * (q[5][j] - q[3][j]) + KA2 * (q[6][j] - q[2][j]) +
                                     KA3 * (q[7][j] - q[1][j]) + KA4 * (q[8][j] - q[0][j]);
                    nv[j] = az[j] * pin[j] + bz[j] * (d1 * a.idz);
                    if (r >= 4 && r < 20 && zz < nzp) *at(pzo, pi + j) = nv[j];
                }
            }
            pzl[i] = make_float4(nv[0], nv[1], nv[2], nv[3]);
        }
    }
    if (PHYS == 2 || xtile || ztile) __syncthreads();  // block-uniform
    if (!zin) return;

    // ---- the derivatives of this thread's row
    float zw[9][4];
#pragma unroll
    for (int h = 0; h < 9; ++h) f4_to(zw[h], su[(rg + 4 + h) * SQ + cx + 2]);
    float w[12];
    {
        const float4 *rw = su + (rg + 8) * SQ + cx + 1;
        f4_to(w, rw[0]);
        f4_to(w + 4, rw[1]);
        f4_to(w + 8, rw[2]);
    }
    float d2x[4], d1x[4];
    xderivs<4>(w, d2x, d1x);
    float c0[4], pzz[4], pxx[4], px[4], pz[4], pxz[4];
    tile_derivs<PHYS>(zw, d2x, d1x, a, c0, pzz, pxx, pz, px);
#pragma unroll
    for (int j = 0; j < 4; ++j) pxz[j] = 0.f;
    if (PHYS == 2) {
        float q[9][4];
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            if (h == 4) continue;
            f4_to(q[h], sd[(rg + h) * 16 + cx]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            pxz[j] = (KA1 * (q[5][j] - q[3][j]) + KA2 * (q[6][j] - q[2][j]) + KA3 * (q[7][j] - q[1][j]) +
                      KA4 * (q[8][j] - q[0][j])) *
                     a.idxdz;
    }
    // receiver gather of the input field (no Cerjan factors with CPML)
    if ((flags & F_GATHER) && a.rrow[z] >= 0) {
        const int slot = a.rrow[z];
        const int it = m.k - a.itlag;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int rr = a.rhead[(long long)slot * P + col + j];
            while (rr >= 0) {
                a.seis[(long long)s * a.dstride + (long long)it * a.nrec + rr] = c0[j];
                rr = a.rnext[rr];
            }
        }
    }
    if (z < 4 || z >= nzp - 4) return;

    // ---- CPML: zeta with the new psi, corrected derivatives
    bool xlive[4], xs[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        xlive[j] = (col + j >= 4) && (col + j < nxp - 4);
        xs[j] = xstrip(col + j);
    }
    const bool zs = zstrip(z) && col < nxp - 4;
    float PXX[4], PZZ[4], PX[4], PZ[4], dpx[4], zxv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        PXX[j] = pxx[j];
        PZZ[j] = pzz[j];
        PX[j] = px[j];
        PZ[j] = pz[j];
        dpx[j] = zxv[j] = 0.f;
    }
    if (xtile && (xs[0] || xs[1] || xs[2] || xs[3])) {
        float pw[12], d2p[4], d1p[4], zo[4], ax[4], bx[4];
        const float4 *pr = pxl + rg * GX + cx;
        f4_to(pw, pr[0]);
        f4_to(pw + 4, pr[1]);
        f4_to(pw + 8, pr[2]);
        xderivs<4>(pw, d2p, d1p);
        vld<4>(zo, at(zetax, ro));
        vld<4>(ax, at(a.cax, ro));
        vld<4>(bx, at(a.cbx, ro));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!xs[j]) continue;
            const float dps = d1p[j] * a.idx;
            zxv[j] = ax[j] * zo[j] + bx[j] * (pxx[j] + dps);
            dpx[j] = dps;
            PX[j] = px[j] + dps;
            PXX[j] = pxx[j] + dps + zxv[j];
            *at(zetax, ro + j) = zxv[j];
        }
    }
    float dpz[4] = {0.f, 0.f, 0.f, 0.f}, zzv[4] = {0.f, 0.f, 0.f, 0.f};
    if (zs) {
        float ps[9][4], zo[4], az[4], bz[4];
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            if (h == 4) continue;
            f4_to(ps[h], pzl[(rg + h) * 16 + cx]);
        }
        vld<4>(zo, at(zetaz, ro));
        vld<4>(az, at(a.caz, ro));
        vld<4>(bz, at(a.cbz, ro));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float dps = (KA1 * (ps[5][j] - ps[3][j]) + KA2 * (ps[6][j] - ps[2][j]) +
                               KA3 * (ps[7][j] - ps[1][j]) + KA4 * (ps[8][j] - ps[0][j])) *
                              a.idz;
            zzv[j] = az[j] * zo[j] + bz[j] * (pzz[j] + dps);
            dpz[j] = dps;
            PZ[j] = pz[j] + dps;
            PZZ[j] = pzz[j] + dps + zzv[j];
            if (xlive[j]) *at(zetaz, ro + j) = zzv[j];
        }
    }

    // ---- the update (no Cerjan factors), the next step's source, the store
    float out[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float op;
        if (PHYS == 0) {
            float sum = pxx[j] + pzz[j];
            if (xs[j]) sum += dpx[j];
            if (zs) sum += dpz[j];
            if (xs[j]) sum += zxv[j];
            if (zs) sum += zzv[j];
            out[j] = lf_acoustic(c0[j], pv[j], 1.0f, 1.0f, v2[j], sum, false, op);
        } else if (PHYS == 1) {
            out[j] = lf_vti(c0[j], pv[j], 1.0f, 1.0f, v2[j], e[j], dl[j], PX[j], PZ[j], PXX[j], PZZ[j], false, op);
        } else {
            out[j] = lf_tti(c0[j], pv[j], 1.0f, 1.0f, v2[j], e[j], dl[j], cs[j], sn[j], PX[j], PZ[j], PXX[j],
                            PZZ[j], pxz[j], false, op);
        }
    }
    if ((flags & F_SRC) && z == a.ssz[s] && m.k + 1 < a.nt) {
        const int sx = a.ssx[s];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (col + j == sx) out[j] += a.wav[m.k + 1];
    }
    if (xlive[0] && xlive[3]) {
        vst<4>(at(prv, ro), out);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (xlive[j]) *at(prv, ro + j) = out[j];
    }
}

// Blocks go to the XCDs round-robin (hardware block h to XCD h mod 8); XCD x
// runs interior blocks [x ni8, (x+1) ni8) first, then edge blocks
// [x ne8, (x+1) ne8): both kinds are spread over the XCDs, each XCD's share is
// contiguous (the halo and model reuse of neighbouring tiles stays in its L2),
// and the longer interior blocks start first.  Interior block b runs tile
// b / G for shot group b % G (G = ceil(nshots / SG)); edge block e runs edge
// tile e / nshots for shot e % nshots.
// (the VTI instance with fused psi runs at 4 waves per SIMD: at 5 its edge body
// spills, and the 4-wave instance measured faster)
template <int PHYS, bool CPML, int SG, bool FUSE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    PHYS < 2 ? ((FUSE && CPML && PHYS == 1) ? 4 : 5) : 1))) void k_tile_pipe(
    const StepArgs a) {
    const TileSplit t = tile_split(a.nxp, a.nzp, a.N, a.nxt, a.nzt);
    const int G = (a.nshots + SG - 1) / SG;
    const int nib = t.nin * G, neb = t.nedge * a.nshots;
    const int ni8 = (nib + 7) / 8, ne8 = (neb + 7) / 8;
    const int h = blockIdx.x, x = h & 7, q = h >> 3;
    const TileStep m{a.k, a.flags, a.cur, a.prv, a.psix, a.psiz};
    if (q < ni8) {
        const int b = x * ni8 + q;
        if (b >= nib) return;  // block-uniform
        const int ti = b / G, g = b - ti * G;
        int x0, z0;
        inner_tile_xz(t, ti, x0, z0);
        const int s0 = g * SG;
        pipe_inner<PHYS, F_FWD_MASK, SG>(a, m, x0, z0, s0, min(SG, a.nshots - s0));
        return;
    }
    const int eb = x * ne8 + (q - ni8);
    if (eb >= neb) return;  // block-uniform
    const int ei = eb / a.nshots, s = eb - ei * a.nshots;
    int x0, z0;
    edge_tile_xz(t, a.nxt, a.nzt, ei, x0, z0);
    if constexpr (CPML && FUSE)
        edge_fused_body<PHYS>(a, m, x0, z0, s);
    else
        tile_body_at<PHYS, CPML, 16, F_FWD_MASK, false>(a, x0, z0, s, m);
}

// A group of grids in one launch: block b belongs to grid g with boff[g] <= b <
// boff[g + 1], whose arguments are list[g] as of step 0 (its two wavefields and
// psi halves trade places on odd steps); k is the step of every grid.  The new
// psi comes from k_psi_multi.
template <int PHYS, bool CPML, int TZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((TZ == 16 && PHYS < 2) ? 5 : 1))) void k_tile_multi(
    const StepArgs *__restrict__ list, const int *__restrict__ boff, int ngrid, int k) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= boff[ngrid]) return;  // block-uniform
    int g = 0;
    while (g + 1 < ngrid && b >= boff[g + 1]) ++g;
    const StepArgs &a = list[g];
    TileStep m{k, a.flags, a.cur, a.prv, a.psix, a.psiz};
    if (k < a.itlag) m.flags &= ~F_GATHER;
    if (k & 1) {
        m.cur = a.prv;
        m.prv = const_cast<float *>(a.cur);
        m.psix = a.psix_in;
        m.psiz = a.psiz_in;
    }
    int x0, z0, s;
    tile_of<TZ>(a, b - boff[g], x0, z0, s);
    if (tile_inner<TZ>(a, x0, z0))  // block-uniform
        tile_body<PHYS, CPML, TZ, F_FWD_MASK, true>(a, b - boff[g], m);
    else
        tile_body<PHYS, CPML, TZ, F_FWD_MASK>(a, b - boff[g], m);
}

template <int PHYS, bool CPML, int TZ>
static int launch_tile_multi_t(const StepArgs *list, const int *boff, int ngrid, int nblocks, int k,
                               const int *pboff, int npblocks, hipStream_t st) {
    if (CPML && npblocks > 0) {
        const int r = launch_psi_multi(list, pboff, ngrid, npblocks, k, st);
        if (r) return r;
    }
    const int blocks = (nblocks + 7) / 8 * 8;
    hipLaunchKernelGGL((k_tile_multi<PHYS, CPML, TZ>), dim3((unsigned)blocks), dim3(256), 0, st, list, boff, ngrid, k);
    return (int)hipGetLastError();
}

// the pipelined variant: forward steps (flags within F_FWD_MASK) only; the
// other passes run variant 0, whose arithmetic it shares
template <int PHYS, bool CPML, int TZ, int WPE>
static int launch_tile_t(const StepArgs &a0, hipStream_t st, const hipEvent_t *ev);
template <int PHYS, bool CPML, int SG, bool FUSE>
static int launch_tile_pipe_t(const StepArgs &a0, hipStream_t st, const hipEvent_t *ev = nullptr) {
    if ((a0.flags & ~F_FWD_MASK) != 0) {
        if (CPML && FUSE) {  // the variant that updates psi itself: its pre-pass, then variant 0
            PsiArgs p;
            p.nzp = a0.nzp;
            p.nxp = a0.nxp;
            p.P = a0.P;
            p.N = a0.N;
            p.nshots = a0.nshots;
            p.sstride = a0.sstride;
            p.idx = a0.idx;
            p.idz = a0.idz;
            p.cur = a0.cur;
            p.psix_in = a0.psix_in;
            p.psiz_in = a0.psiz_in;
            p.psix = a0.psix;
            p.psiz = a0.psiz;
            p.cax = a0.cax;
            p.cbx = a0.cbx;
            p.caz = a0.caz;
            p.cbz = a0.cbz;
            const int r = launch_psi(p, st);
            if (r) return r;
        }
        return launch_tile_t<PHYS, CPML, 16, 1>(a0, st, ev);
    }
    StepArgs a = a0;
    a.nxt = (a.nxp + 63) / 64;
    a.nzt = (a.nzp + 15) / 16;
    const TileSplit t = tile_split(a.nxp, a.nzp, a.N, a.nxt, a.nzt);
    const long long G = (a.nshots + SG - 1) / SG;
    const long long ni8 = (t.nin * G + 7) / 8, ne8 = ((long long)t.nedge * a.nshots + 7) / 8;
    launch_timed(k_tile_pipe<PHYS, CPML, SG, FUSE>, dim3((unsigned)(8 * (ni8 + ne8))), st, ev, a);
    return (int)hipGetLastError();
}

template <int PHYS, bool CPML, int TZ, int WPE>
static int launch_tile_t(const StepArgs &a0, hipStream_t st, const hipEvent_t *ev) {
    StepArgs a = a0;
    a.nxt = (a.nxp + 63) / 64;
    a.nzt = (a.nzp + TZ - 1) / TZ;
    long long blocks = (long long)a.nxt * a.nzt * a.nshots;
    blocks = (blocks + 7) / 8 * 8;
    const dim3 grid((unsigned)blocks);
    // the smallest instance whose flag set holds the step's flags
    const auto within = [&](int fm) { return (a.flags & ~fm) == 0; };
    if (within(F_FWD_MASK)) {
        launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_FWD_MASK>, grid, st, ev, a);
        return (int)hipGetLastError();
    }
    if constexpr (TZ == 16) {
        if (within(F_PASS_ILUM)) {
            launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_PASS_ILUM>, grid, st, ev, a);
            return (int)hipGetLastError();
        }
        if (within(F_PASS_SAVE)) {
            launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_PASS_SAVE>, grid, st, ev, a);
            return (int)hipGetLastError();
        }
        if (within(F_PASS_GRAD)) {
            launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_PASS_GRAD>, grid, st, ev, a);
            return (int)hipGetLastError();
        }
        if (within(F_PASS_IMAGE)) {
            launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_PASS_IMAGE>, grid, st, ev, a);
            return (int)hipGetLastError();
        }
        if (within(F_PASS_SAVEALL)) {
            launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_PASS_SAVEALL>, grid, st, ev, a);
            return (int)hipGetLastError();
        }
    }
    launch_timed(k_tile<PHYS, CPML, TZ, WPE, F_ALL>, grid, st, ev, a);
    return (int)hipGetLastError();
}

}  // namespace pam
