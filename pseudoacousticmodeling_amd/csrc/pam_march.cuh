// pam_march.cuh -- the z-marching fused time step k_march.
//
// One 256-thread block owns a 64-column strip of one shot and marches down
// a range of rows in chunks of 16.  The current field U_k flows through an
// LDS ring of 64 rows (each staged row carries a 4-cell halo on each side);
// while the block computes chunk c from the ring, the global loads of chunk
// c+1 (its 16 new U_k rows, and U_{k-1} and the model coefficients of its
// cells) are already in flight in registers, so every wavefront keeps memory
// busy during the arithmetic.  Each U_k cell is read from HBM once per strip
// (plus a 4-row halo at the ends of the range); one barrier per chunk.
//
// Per cell the step is the one of k_tile (the same arithmetic, flags and
// fused epilogue): thread (cx, rg) owns columns 4cx..4cx+3 of row rg of the
// chunk.
#pragma once
#include "pam_step.cuh"

namespace pam {

constexpr int kMarchRing = 64;  // rows in the LDS ring (power of two)

template <int PHYS>
struct MarchCell {
    float pv[4], v2[4], e[4], dl[4], cs[4], sn[4];
};

template <int PHYS>
__device__ __forceinline__ void march_load_cell(MarchCell<PHYS> &m, const StepArgs &a, const float *prv, int z, int col) {
#pragma unroll
    for (int j = 0; j < 4; ++j) m.pv[j] = m.v2[j] = m.e[j] = m.dl[j] = m.cs[j] = m.sn[j] = 0.f;
    if (z < a.nzp) {
        const long long ro = (long long)z * a.P + col;
        vld<4>(m.pv, prv + ro);
        vld<4>(m.v2, a.v2 + ro);
        if (PHYS >= 1) {
            vld<4>(m.e, a.eps + ro);
            vld<4>(m.dl, a.del + ro);
        }
        if (PHYS == 2) {
            vld<4>(m.cs, a.cs + ro);
            vld<4>(m.sn, a.sn + ro);
        }
    }
}

template <int PHYS, bool CPML, int CHUNKS>
__global__ __launch_bounds__(256) void k_march(const StepArgs a) {
    constexpr int TX = 64, TC = 16;
    constexpr int Q = (TX + 8) / 4;          // float4 per staged row
    constexpr int NEW = TC * Q;              // float4 of U entering per chunk
    constexpr int NP = (NEW + 255) / 256;    // per thread
    constexpr int NPRO = ((TC + 8) * Q + 255) / 256;
    __shared__ float4 ring[kMarchRing * Q];
    __shared__ float4 sd[(PHYS == 2) ? kMarchRing * (TX / 4) : 1];

    const int nseg = a.nzt;                  // row segments
    const int nb = a.nxt * nseg * a.nshots;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= nb) return;  // block-uniform
    const int strip = b / a.nshots;
    const int s = b - strip * a.nshots;
    const int seg = strip / a.nxt;
    const int txi = strip - seg * a.nxt;
    const int x0 = txi * TX;
    const int P = a.P, nzp = a.nzp, nxp = a.nxp, N = a.N;
    const int zb = 4 + seg * CHUNKS * TC;           // first row of the range
    int ze = zb + CHUNKS * TC;                       // one past the last row
    if (ze > nzp - 4) ze = nzp - 4;
    const int nchunk = (ze - zb + TC - 1) / TC;
    const long long so = (long long)s * a.sstride;
    const float *__restrict__ cur = a.cur + so;
    float *__restrict__ prv = a.prv + so;
    const int tid = threadIdx.x;
    const int flags = a.flags;
    const int cx = tid & 15, rg = tid >> 4;
    const int col = x0 + 4 * cx;

    // ring slot of global row z (z >= -4)
#define SLOT(z) ((((z) + 8) & (kMarchRing - 1)) * Q)

    // ---- prologue: rows zb-4 .. zb+TC+3 and the cells of chunk 0
    {
        float4 st[NPRO];
#pragma unroll
        for (int u = 0; u < NPRO; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            const int gz = zb - 4 + r, gc = x0 - 4 + 4 * q;
            st[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < (TC + 8) * Q && gz >= 0 && gz < nzp && gc >= 0 && gc < P)
                st[u] = *reinterpret_cast<const float4 *>(cur + (long long)gz * P + gc);
        }
#pragma unroll
        for (int u = 0; u < NPRO; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            if (i < (TC + 8) * Q) ring[SLOT(zb - 4 + r) + q] = st[u];
        }
    }
    MarchCell<PHYS> mc, mn;
    march_load_cell<PHYS>(mc, a, prv, zb + rg, col);
    __syncthreads();

    bool xlive[4], xs[4];
    float fx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = col + j;
        xlive[j] = (c >= 4) && (c < nxp - 4);
        fx[j] = (!CPML && (flags & F_CERJAN) && c < P) ? a.cerx[c] : 1.0f;
        xs[j] = CPML && ((c >= 4 && c < N) || (c >= nxp - N && c < nxp - 4));
    }
    const bool tile_xstrip = CPML && (x0 < N || x0 + TX > nxp - N);

    for (int ch = 0; ch < nchunk; ++ch) {
        const int zc = zb + ch * TC;
        const bool more = ch + 1 < nchunk;
        // ---- prefetch chunk ch+1: U rows zc+TC+4 .. zc+2TC+3, and its cells
        float4 pf[NP];
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            const int gz = zc + TC + 4 + r, gc = x0 - 4 + 4 * q;
            pf[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (more && i < NEW && gz < nzp && gc >= 0 && gc < P)
                pf[u] = *reinterpret_cast<const float4 *>(cur + (long long)gz * P + gc);
        }
        if (more) march_load_cell<PHYS>(mn, a, prv, zc + TC + rg, col);

        if (PHYS == 2) {
            // first x-derivative of rows zc-4 .. zc+TC+3 (TC+8 rows x 16 float4)
            for (int i = tid; i < (TC + 8) * (TX / 4); i += 256) {
                const int r = i >> 4, c = i & 15;
                const int gz = zc - 4 + r;
                const float4 *rw = ring + SLOT(gz);
                float w[12];
                const float4 l = rw[c], m = rw[c + 1], h = rw[c + 2];
                w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
                w[4] = m.x; w[5] = m.y; w[6] = m.z; w[7] = m.w;
                w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
                float d2[4], d1[4];
                xderivs<4>(w, d2, d1);
                sd[((gz + 8) & (kMarchRing - 1)) * (TX / 4) + c] = make_float4(d1[0], d1[1], d1[2], d1[3]);
            }
            __syncthreads();
        }

        // ---- compute row z of chunk ch
        const int z = zc + rg;
        if (z < ze) {
            const long long ro = (long long)z * P + col;
            const float fz = (!CPML && (flags & F_CERJAN)) ? a.cerz[z] : 1.0f;
            float zw[9][4];
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                const float4 t = ring[SLOT(z - 4 + h) + cx + 1];
                zw[h][0] = t.x; zw[h][1] = t.y; zw[h][2] = t.z; zw[h][3] = t.w;
            }
            float w[12];
            {
                const float4 *rw = ring + SLOT(z);
                const float4 l = rw[cx], m = rw[cx + 1], h = rw[cx + 2];
                w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
                w[4] = m.x; w[5] = m.y; w[6] = m.z; w[7] = m.w;
                w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
            }
            float d2x[4], d1x[4];
            xderivs<4>(w, d2x, d1x);
            float c0[4], pzz[4], pxx[4], px[4], pz[4], pxz[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float cc = zw[4][j];
                c0[j] = cc;
                pzz[j] = (KC0 * cc + KC1 * (zw[5][j] + zw[3][j]) + KC2 * (zw[6][j] + zw[2][j]) +
                          KC3 * (zw[7][j] + zw[1][j]) + KC4 * (zw[8][j] + zw[0][j])) *
                         a.idz2;
                pxx[j] = d2x[j] * a.idx2;
                px[j] = pz[j] = pxz[j] = 0.f;
                if (PHYS >= 1) {
                    pz[j] = (KA1 * (zw[5][j] - zw[3][j]) + KA2 * (zw[6][j] - zw[2][j]) +
                             KA3 * (zw[7][j] - zw[1][j]) + KA4 * (zw[8][j] - zw[0][j])) *
                            a.idz;
                    px[j] = d1x[j] * a.idx;
                }
            }
            if (PHYS == 2) {
                float q[9][4];
#pragma unroll
                for (int h = 0; h < 9; ++h) {
                    if (h == 4) continue;
                    const float4 t = sd[((z - 4 + h + 8) & (kMarchRing - 1)) * (TX / 4) + cx];
                    q[h][0] = t.x; q[h][1] = t.y; q[h][2] = t.z; q[h][3] = t.w;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    pxz[j] = (KA1 * (q[5][j] - q[3][j]) + KA2 * (q[6][j] - q[2][j]) + KA3 * (q[7][j] - q[1][j]) +
                              KA4 * (q[8][j] - q[0][j])) *
                             a.idxdz;
            }

            // receiver gather of the input field (damped as the reference records it)
            if ((flags & F_GATHER) && a.rrow[z] >= 0) {
                const int slot = a.rrow[z];
                const int it = a.k - a.itlag;
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
            // snapshots / illumination / imaging of the input field
            if (flags & (F_SAVE | F_ILUM | F_IMAGE | F_GRADM)) {
                const bool zint = (z >= N) && (z < nzp - N);
                const long long ib = (long long)(z - N) * a.nx - N;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int c = col + j;
                    const bool cin = zint && c >= N && c < nxp - N;
                    const long long ii = ib + c;
                    if (flags & F_SAVE) {
                        if (a.save_kind == SV_F32_PAD) {
                            if (xlive[j]) snap_put(a, s, a.k, 0, ro + j, (c0[j] * fx[j]) * fz);
                        } else if (cin) {
                            snap_put(a, s, a.k, ii, 0, c0[j]);
                        }
                    }
                    if (!cin) continue;
                    if (flags & F_ILUM) {
                        float *acc = a.ilum + (long long)s * a.acc_sstride + ii;
                        *acc += c0[j] * c0[j];
                    }
                    if (flags & F_IMAGE) {
                        const float S = snap_get(a, s, a.k, ii, ro + j);
                        float *im = a.image + (long long)s * a.acc_sstride + ii;
                        *im += S * c0[j];
                    }
                    if (flags & F_GRADM) {
                        const float Sp = snap_get(a, s, a.k + 1, ii, ro + j);
                        const float S0 = snap_get(a, s, a.k, ii, ro + j);
                        const float Sm = snap_get(a, s, a.k - 1, ii, ro + j);
                        const float d2t = (Sp - 2.0f * S0 + Sm) * a.idt2;
                        float *im = a.image + (long long)s * a.acc_sstride + ii;
                        *im += d2t * c0[j];
                    }
                }
            }

            // CPML: zeta with the new psi, corrected derivatives
            float PXX[4], PZZ[4], PX[4], PZ[4], zx[4], zz[4], dpx[4], dpz[4];
            const bool zs = CPML && ((z >= 4 && z < N) || (z >= nzp - N && z < nzp - 4));
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
                const float *prow = a.psix + so + (long long)z * P;
                vldg<4>(pw + 4, prow, col, P);
                if (col >= 4) vld<4>(pw, prow + col - 4);
                else pw[0] = pw[1] = pw[2] = pw[3] = 0.f;
                vldg<4>(pw + 8, prow, col + 4, P);
                xderivs<4>(pw, d2p, d1p);
                vld<4>(zo, a.zetax + so + ro);
                vld<4>(ax, a.cax + ro);
                vld<4>(bx, a.cbx + ro);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!xs[j]) continue;
                    const float dps = d1p[j] * a.idx;
                    zx[j] = ax[j] * zo[j] + bx[j] * (pxx[j] + dps);
                    dpx[j] = dps;
                    PX[j] = px[j] + dps;
                    PXX[j] = pxx[j] + dps + zx[j];
                    a.zetax[so + ro + j] = zx[j];
                }
            }
            if (CPML && zs) {
                float ps[9][4];
#pragma unroll
                for (int h = 0; h < 9; ++h) {
                    if (h == 4) continue;
                    vld<4>(ps[h], a.psiz + so + (long long)(z - 4 + h) * P + col);
                }
                float zo[4], az[4], bz[4];
                vld<4>(zo, a.zetaz + so + ro);
                vld<4>(az, a.caz + ro);
                vld<4>(bz, a.cbz + ro);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float dps = (KA1 * (ps[5][j] - ps[3][j]) + KA2 * (ps[6][j] - ps[2][j]) +
                                       KA3 * (ps[7][j] - ps[1][j]) + KA4 * (ps[8][j] - ps[0][j])) *
                                      a.idz;
                    zz[j] = az[j] * zo[j] + bz[j] * (pzz[j] + dps);
                    dpz[j] = dps;
                    PZ[j] = pz[j] + dps;
                    PZZ[j] = pzz[j] + dps + zz[j];
                    if (xlive[j]) a.zetaz[so + ro + j] = zz[j];
                }
            }

            // the update
            float out[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float pe = (mc.pv[j] * fx[j]) * fz;
                float u;
                if (PHYS == 0) {
                    float sum = pxx[j] + pzz[j];
                    if (CPML) {
                        if (xs[j]) sum += dpx[j];
                        if (zs) sum += dpz[j];
                        if (xs[j]) sum += zx[j];
                        if (zs) sum += zz[j];
                    }
                    u = mc.v2[j] * sum + 2.0f * c0[j] - pe;
                } else if (PHYS == 1) {
                    const float Sd = sd_term(mc.e[j], mc.dl[j], PX[j], PZ[j]);
                    u = 2.0f * c0[j] - pe + mc.v2[j] * ((1.0f + 2.0f * mc.e[j]) + Sd) * PXX[j] +
                        mc.v2[j] * (1.0f + Sd) * PZZ[j];
                } else {
                    const float xi = PX[j] * mc.cs[j] - PZ[j] * mc.sn[j];
                    const float eta = PX[j] * mc.sn[j] + PZ[j] * mc.cs[j];
                    const float Sd = sd_term(mc.e[j], mc.dl[j], xi, eta);
                    const float Ac = (1.0f + 2.0f * mc.e[j]) * (mc.cs[j] * mc.cs[j]) + (mc.sn[j] * mc.sn[j]) + Sd;
                    const float Bc = (1.0f + 2.0f * mc.e[j]) * (mc.sn[j] * mc.sn[j]) + (mc.cs[j] * mc.cs[j]) + Sd;
                    u = 2.0f * c0[j] - pe + mc.v2[j] * Ac * PXX[j] + mc.v2[j] * Bc * PZZ[j] -
                        2.0f * mc.e[j] * mc.v2[j] * (2.0f * mc.sn[j] * mc.cs[j]) * pxz[j];
                }
                out[j] = (u * fx[j]) * fz;
            }
            if ((flags & F_SRC) && z == a.ssz[s] && a.k + 1 < a.nt) {
                const int sx = a.ssx[s];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (col + j == sx) out[j] += a.wav[a.k + 1];
            }
            if ((flags & F_RECINJ) && a.rrow[z] >= 0 && a.k - 1 - a.itlag >= 0) {
                const int slot = a.rrow[z];
                const int tn = a.k - 1 - a.itlag;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    int rr = a.rhead[(long long)slot * P + col + j];
                    while (rr >= 0) {
                        out[j] += a.data[(long long)s * a.dstride + (long long)tn * a.nrec + rr] / a.dxdz;
                        rr = a.rnext[rr];
                    }
                }
            }
            if (xlive[0] && xlive[3]) {
                vst<4>(prv + ro, out);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (xlive[j]) prv[ro + j] = out[j];
            }
        }

        // ---- the prefetched rows enter the ring
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            if (more && i < NEW) ring[SLOT(zc + TC + 4 + r) + q] = pf[u];
        }
        mc = mn;
        __syncthreads();
    }
#undef SLOT
}

template <int PHYS, bool CPML, int CHUNKS>
static int launch_march_t(const StepArgs &a0, hipStream_t st) {
    StepArgs a = a0;
    a.nxt = (a.nxp + 63) / 64;
    a.nzt = (a.nzp - 8 + CHUNKS * 16 - 1) / (CHUNKS * 16);  // row segments
    long long blocks = (long long)a.nxt * a.nzt * a.nshots;
    blocks = (blocks + 7) / 8 * 8;
    hipLaunchKernelGGL((k_march<PHYS, CPML, CHUNKS>), dim3((unsigned)blocks), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

}  // namespace pam
