// pam_march.cuh -- the z-marching fused time step k_march.
//
// One 256-thread block owns a 64-column strip of one shot and marches down
// a range of rows in chunks of 16.  The current field U_k flows through an
// LDS ring of 64 rows (each staged row carries an x-halo); while the block
// computes chunk c from the ring, the global loads of chunk c+1 (its 16 new
// U_k rows, and U_{k-1} and the model coefficients of its cells) are already
// in flight in registers, so every wavefront keeps memory busy during the
// arithmetic.  Each U_k cell is read from HBM once per strip (plus a halo of
// a few rows at the ends of the range); one barrier per chunk.
//
// CPML is fused as well: the memory variable psi of the step is computed in
// the pass that uses it.  Rows of the z-strips enter a 32-row LDS ring of new
// psi_z as the march reaches them (each row once, 4 rows ahead of its use);
// tiles of the x-strips recompute psi_x on a 4-column halo from an 8-column
// U halo.  psi is double-buffered across steps (read from psix_in/psiz_in,
// written to psix/psiz), so halo recomputation never races with a neighbour
// block's write.  zeta is read and written by the block that owns the cell.
//
// Per cell the step is the one of k_tile (the same arithmetic, flags and
// fused epilogue): thread (cx, rg) owns columns 4cx..4cx+3 of row rg of the
// chunk.
#pragma once
#include "pam_step.cuh"

namespace pam {

constexpr int kMarchRing = 64;  // rows in the LDS ring of U (power of two)
constexpr int kPsiRing = 32;    // rows in the LDS ring of new psi_z (power of two)

template <int PHYS>
struct MarchCell {
    float pv[4], v2[4], e[4], dl[4], cs[4], sn[4];
};

template <int PHYS>
__device__ __forceinline__ void march_load_cell(MarchCell<PHYS> &m, const StepArgs &a, const float *prv, int z, int col) {
#pragma unroll
    for (int j = 0; j < 4; ++j) m.pv[j] = m.v2[j] = m.e[j] = m.dl[j] = m.cs[j] = m.sn[j] = 0.f;
    if (z < a.nzp) {
        const unsigned ro = (unsigned)(z * a.P + col);
        vld<4>(m.pv, at(prv, ro));
        vld<4>(m.v2, at(a.v2, ro));
        if (PHYS >= 1) {
            vld<4>(m.e, at(a.eps, ro));
            vld<4>(m.dl, at(a.del, ro));
        }
        if (PHYS == 2) {
            vld<4>(m.cs, at(a.cs, ro));
            vld<4>(m.sn, at(a.sn, ro));
        }
    }
}


// 4 consecutive cells starting at column c: vector store when all are in the
// stencil interior [4, nxp-4), else per cell
__device__ __forceinline__ void store_live4(float *p, const float *v, int c, int nxp) {
    if (c >= 4 && c + 3 < nxp - 4) {
        vst<4>(p, v);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (c + j >= 4 && c + j < nxp - 4) p[j] = v[j];
    }
}

// The arguments of one step that change from step to step (held apart from
// StepArgs so that a group launch can read the rest in place).
struct MarchStep {
    int k, flags;
    const float *cur;
    float *prv;
    const float *pxi, *pzi;  // psi of the previous step
    float *pxo, *pzo;        // new psi
};

// One block's work: block b (in [0, nxt * nzt * nshots)) of the launch
// described by a and the step arguments m.
template <int PHYS, bool CPML, int CHUNKS, bool ADJ>
__device__ __forceinline__ void march_body(const StepArgs &a, const int b, const MarchStep m) {
    constexpr int TX = 64, TC = 16;
    constexpr int HX = CPML ? 8 : 4;       // x-halo of the staged rows
    constexpr int LA = CPML ? 8 : 4;       // rows staged above / below a chunk
    constexpr int Q = (TX + 2 * HX) / 4;   // float4 per staged row
    constexpr int CQ = HX / 4;             // float4 index of column x0 in a staged row
    constexpr int NEW = TC * Q;            // float4 of U entering per chunk
    constexpr int NP = (NEW + 255) / 256;
    constexpr int NPRO = ((TC + 2 * LA) * Q + 255) / 256;
    constexpr int G = TX / 4;              // float4 groups of the tile's own columns
    __shared__ float4 ring[kMarchRing * Q];
    __shared__ float4 sd[(PHYS == 2) ? kMarchRing * G : 1];
    __shared__ float4 pzr[CPML ? kPsiRing * G : 1];         // new psi_z rows, own columns
    __shared__ float4 pxl[CPML ? TC * (G + 2) : 1];         // new psi_x of the chunk, columns x0-4 .. x0+TX+3

    const int strip = b / a.nshots;
    const int s = b - strip * a.nshots;
    const int seg = strip / a.nxt;  // row segment
    const int txi = strip - seg * a.nxt;
    const int x0 = txi * TX;
    const int P = a.P, nzp = a.nzp, nxp = a.nxp, N = a.N;
    const int zb = 4 + seg * CHUNKS * TC;  // first row of the range
    int ze = zb + CHUNKS * TC;             // one past the last row
    if (ze > nzp - 4) ze = nzp - 4;
    const int nchunk = (ze - zb + TC - 1) / TC;
    const long long so = (long long)s * a.sstride;
    const float *__restrict__ cur = m.cur + so;
    float *__restrict__ prv = m.prv + so;
    // per-shot bases of the CPML fields; lane offsets are 32-bit (at())
    const float *__restrict__ pxi = m.pxi + so;
    const float *__restrict__ pzi = m.pzi + so;
    float *__restrict__ pxo = m.pxo + so;
    float *__restrict__ pzo = m.pzo + so;
    float *__restrict__ zetax = a.zetax + so;
    float *__restrict__ zetaz = a.zetaz + so;
    const int tid = threadIdx.x;
    const int flags = ADJ ? m.flags : (m.flags & F_FWD_MASK);  // forward-only instances drop the rest
    const int cx = tid & 15, rg = tid >> 4;
    const int col = x0 + 4 * cx;
    const bool xst = CPML && (x0 < N || x0 + TX > nxp - N);  // tile meets an x-strip (block-uniform)

#define SLOT(z) ((((z) + 8) & (kMarchRing - 1)) * Q)
#define PSLOT(z) ((((z) + 8) & (kPsiRing - 1)) * G)

    // x-halo groups of a staged row are only needed where psi_x is computed
    auto want_q = [&](int q) { return !CPML || xst || (q >= CQ - 1 && q <= CQ + G); };
    auto zstrip = [&](int z) { return CPML && ((z >= 4 && z < N) || (z >= nzp - N && z < nzp - 4)); };
    auto xstrip = [&](int c) { return CPML && ((c >= 4 && c < N) || (c >= nxp - N && c < nxp - 4)); };
    // new psi_z of the 4 cells of group c in row r: a psi_in + b dU/dz, U from the ring
    auto psiz_new = [&](float *nv, const float *pin, const float *az, const float *bz, int r, int c) {
        float uz[9][4];
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            if (h == 4) continue;
            f4_to(uz[h], ring[SLOT(r - 4 + h) + CQ + c]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float d1 = KA1 * (uz[5][j] - uz[3][j]) + KA2 * (uz[6][j] - uz[2][j]) +
                             KA3 * (uz[7][j] - uz[1][j]) + KA4 * (uz[8][j] - uz[0][j]);
            nv[j] = az[j] * pin[j] + bz[j] * (d1 * a.idz);
        }
    };
    // new psi_x of staged group g (columns x0-4+4g .. +3) of row z (chunk row r)
    // into pxl; written to psix for the tile's own columns of rows < ze
    auto psix_new_item = [&](int r, int g, int z, const float *pin, const float *ax, const float *bx) {
        const int c0 = x0 - 4 + 4 * g;
        float nv[4] = {0.f, 0.f, 0.f, 0.f};
        bool any = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) any = any || xstrip(c0 + j);
        if (any && z >= 4 && z < nzp - 4) {
            float w[12], d2[4], d1[4];
            const float4 *rw = ring + SLOT(z) + CQ - 2 + g;
            f4_to(w, rw[0]);
            f4_to(w + 4, rw[1]);
            f4_to(w + 8, rw[2]);
            xderivs<4>(w, d2, d1);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!xstrip(c0 + j)) continue;
                nv[j] = ax[j] * pin[j] + bx[j] * (d1[j] * a.idx);
                if (g >= 1 && g <= G && z < ze) *at(pxo, (unsigned)(z * P + c0 + j)) = nv[j];
            }
        }
        pxl[r * (G + 2) + g] = make_float4(nv[0], nv[1], nv[2], nv[3]);
    };

    // ---- prologue: rows zb-LA .. zb+TC+LA-1 and the cells of chunk 0
    {
        float4 st[NPRO];
#pragma unroll
        for (int u = 0; u < NPRO; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            const int gz = zb - LA + r, gc = x0 - HX + 4 * q;
            st[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < (TC + 2 * LA) * Q && gz >= 0 && gz < nzp && gc >= 0 && gc < P && want_q(q))
                st[u] = *reinterpret_cast<const float4 *>(at(cur, (unsigned)(gz * P + gc)));
        }
#pragma unroll
        for (int u = 0; u < NPRO; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            if (i < (TC + 2 * LA) * Q) ring[SLOT(zb - LA + r) + q] = st[u];
        }
    }
    MarchCell<PHYS> mc, mn;
    march_load_cell<PHYS>(mc, a, prv, zb + rg, col);
    __syncthreads();
    if (CPML) {
        // new psi_z of rows zb-4 .. zb+3 (the march computes later rows 4 ahead of use)
        for (int i = tid; i < 8 * G; i += 256) {
            const int r = zb - 4 + i / G, c = i - (i / G) * G;
            float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
            if (zstrip(r)) {
                const unsigned ro = (unsigned)(r * P + x0 + 4 * c);
                float pin[4], az[4], bz[4], nv[4];
                vld<4>(pin, at(pzi, ro));
                vld<4>(az, at(a.caz, ro));
                vld<4>(bz, at(a.cbz, ro));
                psiz_new(nv, pin, az, bz, r, c);
                o = make_float4(nv[0], nv[1], nv[2], nv[3]);
                if (r >= zb) store_live4(at(pzo, ro), nv, x0 + 4 * c, nxp);
            }
            pzr[PSLOT(r) + c] = o;
        }
    }

    bool xlive[4], xs[4];
    float fx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = col + j;
        xlive[j] = (c >= 4) && (c < nxp - 4);
        fx[j] = (!CPML && (flags & F_CERJAN) && c < P) ? a.cerx[c] : 1.0f;
        xs[j] = CPML && ((c >= 4 && c < N) || (c >= nxp - N && c < nxp - 4));
    }

    float tmax = 0.0f;  // F_SMAX: max |new field| over this thread's interior cells
    for (int ch = 0; ch < nchunk; ++ch) {
        const int zc = zb + ch * TC;
        const bool more = ch + 1 < nchunk;
        // ---- CPML loads of this chunk, issued before the prefetch so that
        // waiting for them does not wait for the prefetch (loads complete in
        // order): psi_z inputs of row zc+4+rg, psi_x inputs of row zc+rg
        // (own group, and for 32 threads a halo group), zeta of the own cells
        // rows zc-4 .. zc+TC+3 meet a z-strip: psi_z work and a barrier this chunk
        const bool zchunk = CPML && ((zc - 4 < N) || (zc + TC + 4 > nzp - N));
        const int zr = zc + 4 + rg;   // psi_z row of this thread
        const int z = zc + rg;        // row this thread updates
        const bool zs = zstrip(z) && z < ze;
        float pzin[4], paz[4], pbz[4], pxin[4], pax[4], pbx[4], hxin[4], hax[4], hbx[4];
        float zzo[4], zaz[4], zbz[4], zxo[4];
        const int hg = (tid < 16) ? 0 : G + 1;  // halo group of threads tid < 32
        const int hr = tid & 15;                // halo row of threads tid < 32
        if (CPML) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                pzin[j] = paz[j] = pbz[j] = pxin[j] = pax[j] = pbx[j] = hxin[j] = hax[j] = hbx[j] = zzo[j] = zaz[j] =
                    zbz[j] = zxo[j] = 0.f;
            if (zchunk && zstrip(zr)) {
                const unsigned ro = (unsigned)(zr * P + col);
                vld<4>(pzin, at(pzi, ro));
                vld<4>(paz, at(a.caz, ro));
                vld<4>(pbz, at(a.cbz, ro));
            }
            if (xst && z >= 4 && z < nzp - 4) {
                const unsigned ro = (unsigned)(z * P + col);
                vld<4>(pxin, at(pxi, ro));
                vld<4>(pax, at(a.cax, ro));
                vld<4>(pbx, at(a.cbx, ro));
                vld<4>(zxo, at(zetax, ro));
            }
            if (xst && tid < 32) {
                const int hz = zc + hr, hc = x0 - 4 + 4 * hg;
                if (hz >= 4 && hz < nzp - 4 && hc >= 0 && hc < P) {
                    const unsigned ro = (unsigned)(hz * P + hc);
                    vld<4>(hxin, at(pxi, ro));
                    vld<4>(hax, at(a.cax, ro));
                    vld<4>(hbx, at(a.cbx, ro));
                }
            }
            if (zs) {
                const unsigned ro = (unsigned)(z * P + col);
                vld<4>(zzo, at(zetaz, ro));
                vld<4>(zaz, at(a.caz, ro));
                vld<4>(zbz, at(a.cbz, ro));
            }
        }
        // ---- prefetch chunk ch+1: U rows zc+TC+LA .. zc+2TC+LA-1, and its cells
        float4 pf[NP];
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            const int gz = zc + TC + LA + r, gc = x0 - HX + 4 * q;
            pf[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (more && i < NEW && gz < nzp && gc >= 0 && gc < P && want_q(q))
                pf[u] = *reinterpret_cast<const float4 *>(at(cur, (unsigned)(gz * P + gc)));
        }
        if (more) march_load_cell<PHYS>(mn, a, prv, zc + TC + rg, col);

        bool sync_needed = false;
        if (PHYS == 2) {
            // first x-derivative of rows zc-4 .. zc+TC+3
            for (int i = tid; i < (TC + 8) * G; i += 256) {
                const int r = i / G, c = i - r * G;
                const int gz = zc - 4 + r;
                const float4 *rw = ring + SLOT(gz) + CQ - 1;
                float w[12], d2[4], d1[4];
                f4_to(w, rw[c]);
                f4_to(w + 4, rw[c + 1]);
                f4_to(w + 8, rw[c + 2]);
                xderivs<4>(w, d2, d1);
                sd[((gz + 8) & (kMarchRing - 1)) * G + c] = make_float4(d1[0], d1[1], d1[2], d1[3]);
            }
            sync_needed = true;
        }
        if (CPML) {
            // new psi_z of row zr (zero off the z-strips) into the ring
            float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
            if (zchunk && zstrip(zr)) {
                float nv[4];
                psiz_new(nv, pzin, paz, pbz, zr, cx);
                o = make_float4(nv[0], nv[1], nv[2], nv[3]);
                if (zr < ze) store_live4(at(pzo, (unsigned)(zr * P + col)), nv, col, nxp);
            }
            pzr[PSLOT(zr) + cx] = o;
            sync_needed = sync_needed || zchunk;
            if (xst) {
                // new psi_x of the chunk rows on columns x0-4 .. x0+TX+3, zero off the x-strips
                psix_new_item(rg, cx + 1, z, pxin, pax, pbx);
                if (tid < 32) {
                    const int hz = zc + hr;
                    psix_new_item(hr, hg, hz, hxin, hax, hbx);
                }
                sync_needed = true;
            }
        }
        if (sync_needed) __syncthreads();  // block-uniform

        // ---- compute row z of chunk ch
        if (z < ze) {
            const unsigned ro = (unsigned)(z * P + col);
            const float fz = (!CPML && (flags & F_CERJAN)) ? a.cerz[z] : 1.0f;
            float zw[9][4];
#pragma unroll
            for (int h = 0; h < 9; ++h) f4_to(zw[h], ring[SLOT(z - 4 + h) + CQ + cx]);
            float w[12];
            {
                const float4 *rw = ring + SLOT(z) + CQ - 1 + cx;
                f4_to(w, rw[0]);
                f4_to(w + 4, rw[1]);
                f4_to(w + 8, rw[2]);
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
                    f4_to(q[h], sd[((z - 4 + h + 8) & (kMarchRing - 1)) * G + cx]);
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
            // snapshots / illumination / imaging of the input field
            if (flags & (F_SAVE | F_ILUM | F_IMAGE | F_GRADM)) {
                const bool zint = (z >= N) && (z < nzp - N);
                const long long ib = (long long)(z - N) * a.nx - N;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int c = col + j;
                    const bool cin = zint && c >= N && c < nxp - N;
                    const long long ii = ib + c;
                    if ((flags & F_SAVE) && !(flags & F_D2T)) {
                        if (a.save_kind == SV_F32_PAD) {
                            if (xlive[j]) snap_put(a, s, m.k, 0, ro + j, (c0[j] * fx[j]) * fz);
                        } else if (cin) {
                            snap_put(a, s, m.k, ii, 0, c0[j]);
                        }
                    }
                    if (!cin) continue;
                    if (flags & F_ILUM) {
                        float *acc = a.ilum + (long long)s * a.acc_sstride + ii;
                        *acc += c0[j] * c0[j];
                    }
                    if (flags & F_IMAGE) {
                        const float S = snap_get(a, s, m.k, ii, ro + j);
                        float *im = a.image + (long long)s * a.acc_sstride + ii;
                        *im += S * c0[j];
                    }
                    if (flags & F_GRADM) {
                        const float Sp = (flags & F_D2T) ? 0.0f : snap_get(a, s, m.k + 1, ii, ro + j);
                        const float S0 = (flags & F_D2T) ? 0.0f : snap_get(a, s, m.k, ii, ro + j);
                        const float Sm = (flags & F_D2T) ? 0.0f : snap_get(a, s, m.k - 1, ii, ro + j);
                        const float d2t =
                            ((flags & F_D2T) ? snap_get(a, s, m.k, ii, ro + j) : (Sp - 2.0f * S0 + Sm)) * a.idt2;
                        float *im = a.image + (long long)s * a.acc_sstride + ii;
                        *im += d2t * c0[j];
                    }
                }
            }

            // CPML: zeta with the new psi, corrected derivatives
            float PXX[4], PZZ[4], PX[4], PZ[4], zx[4], zz[4], dpx[4], dpz[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                PXX[j] = pxx[j];
                PZZ[j] = pzz[j];
                PX[j] = px[j];
                PZ[j] = pz[j];
                zx[j] = zz[j] = dpx[j] = dpz[j] = 0.0f;
            }
            if (CPML && xst) {
                float pw[12], d2p[4], d1p[4];
                const float4 *pr = pxl + rg * (G + 2) + cx;
                f4_to(pw, pr[0]);
                f4_to(pw + 4, pr[1]);
                f4_to(pw + 8, pr[2]);
                xderivs<4>(pw, d2p, d1p);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!xs[j]) continue;
                    const float dps = d1p[j] * a.idx;
                    zx[j] = pax[j] * zxo[j] + pbx[j] * (pxx[j] + dps);
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
                    f4_to(ps[h], pzr[PSLOT(z - 4 + h) + cx]);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float dps = (KA1 * (ps[5][j] - ps[3][j]) + KA2 * (ps[6][j] - ps[2][j]) +
                                       KA3 * (ps[7][j] - ps[1][j]) + KA4 * (ps[8][j] - ps[0][j])) *
                                      a.idz;
                    zz[j] = zaz[j] * zzo[j] + zbz[j] * (pzz[j] + dps);
                    dpz[j] = dps;
                    PZ[j] = pz[j] + dps;
                    PZZ[j] = pzz[j] + dps + zz[j];
                    if (xlive[j]) *at(zetaz, ro + j) = zz[j];
                }
            }

            // the update
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
                    u = lf_acoustic(c0[j], mc.pv[j], fx[j], fz, mc.v2[j], sum, want_op, opv[j]);
                } else if (PHYS == 1) {
                    u = lf_vti(c0[j], mc.pv[j], fx[j], fz, mc.v2[j], mc.e[j], mc.dl[j], PX[j], PZ[j], PXX[j], PZZ[j],
                               want_op, opv[j]);
                } else {
                    u = lf_tti(c0[j], mc.pv[j], fx[j], fz, mc.v2[j], mc.e[j], mc.dl[j], mc.cs[j], mc.sn[j], PX[j],
                               PZ[j], PXX[j], PZZ[j], pxz[j], want_op, opv[j]);
                }
                out[j] = (u * fx[j]) * fz;
            }
            if ((flags & F_SRC) && z == a.ssz[s] && m.k + 1 < a.nt) {
                const int sx = a.ssx[s];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (col + j == sx) out[j] += a.wav[m.k + 1];
            }
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
            if ((flags & F_SMAX) || ((flags & F_D2T) && (flags & F_SAVE)) || (flags & F_DSAVE)) {
                const bool zint = z >= N && z < nzp - N;
                const long long ib = (long long)(z - N) * a.nx - N;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!(zint && col + j >= N && col + j < nxp - N)) continue;
                    // interior cells are undamped: D_k = S_{k+1} - 2 S_k + S_{k-1}
                    const float d = (flags & F_D2T_OP) ? opv[j] : out[j] - 2.0f * c0[j] + mc.pv[j];
                    if ((flags & F_D2T) && (flags & F_SAVE)) snap_put(a, s, m.k, ib + col + j, 0, d);
                    if (flags & F_DSAVE) a.dsave[(long long)s * a.dsave_sstride + ib + col + j] = d;
                    tmax = fmaxf(tmax, fabsf((flags & F_D2T) ? d : out[j]));
                }
            }
            if (xlive[0] && xlive[3]) {
                vst<4>(at(prv, ro), out);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (xlive[j]) *at(prv, ro + j) = out[j];
            }
        }

        // ---- the prefetched rows enter the ring
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int i = tid + 256 * u;
            const int r = i / Q, q = i - r * Q;
            if (more && i < NEW) ring[SLOT(zc + TC + LA + r) + q] = pf[u];
        }
        mc = mn;
        __syncthreads();
    }
    if ((flags & F_SMAX) && m.k + 1 < a.nt) block_max_to(a.smax + ((long long)s * a.nt + m.k + 1) * kSmaxSlots + blockIdx.x % kSmaxSlots, tmax);
#undef SLOT
#undef PSLOT
}

template <int PHYS, bool CPML, int CHUNKS, bool ADJ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((PHYS == 2) ? 1 : 3))) void k_march(
    const StepArgs a) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= a.nxt * a.nzt * a.nshots) return;  // block-uniform
    march_body<PHYS, CPML, CHUNKS, ADJ>(a, b,
                                        MarchStep{a.k, a.flags, a.cur, a.prv, a.psix_in, a.psiz_in, a.psix, a.psiz});
}

// Several grids (a group of models, each with its own shots) advanced one
// time step in one launch: block b belongs to grid g with boff[g] <= b <
// boff[g + 1], whose arguments are list[g] as of step 0 (its two wavefields
// and psi halves trade places on odd steps); k is the step of every grid.
template <int PHYS, bool CPML, int CHUNKS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((PHYS == 2) ? 1 : 3))) void k_march_multi(
    const StepArgs *__restrict__ list, const int *__restrict__ boff, int ngrid, int k) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= boff[ngrid]) return;  // block-uniform
    int g = 0;
    while (g + 1 < ngrid && b >= boff[g + 1]) ++g;
    const StepArgs &a = list[g];
    MarchStep m{k, a.flags, a.cur, a.prv, a.psix_in, a.psiz_in, a.psix, a.psiz};
    if (k < a.itlag) m.flags &= ~F_GATHER;
    if (k & 1) {
        m.cur = a.prv;
        m.prv = const_cast<float *>(a.cur);
        m.pxi = a.psix;
        m.pzi = a.psiz;
        m.pxo = const_cast<float *>(a.psix_in);
        m.pzo = const_cast<float *>(a.psiz_in);
    }
    march_body<PHYS, CPML, CHUNKS, false>(a, b - boff[g], m);
}

template <int PHYS, bool CPML, int CHUNKS>
static int launch_march_multi_t(const StepArgs *list, const int *boff, int ngrid, int nblocks, int k, hipStream_t st) {
    const int blocks = (nblocks + 7) / 8 * 8;
    hipLaunchKernelGGL((k_march_multi<PHYS, CPML, CHUNKS>), dim3((unsigned)blocks), dim3(256), 0, st, list, boff,
                       ngrid, k);
    return (int)hipGetLastError();
}

template <int PHYS, bool CPML, int CHUNKS>
static int launch_march_t(const StepArgs &a0, hipStream_t st, const hipEvent_t *ev = nullptr) {
    StepArgs a = a0;
    a.nxt = (a.nxp + 63) / 64;
    a.nzt = (a.nzp - 8 + CHUNKS * 16 - 1) / (CHUNKS * 16);  // row segments
    long long blocks = (long long)a.nxt * a.nzt * a.nshots;
    blocks = (blocks + 7) / 8 * 8;
    if ((a.flags & ~F_FWD_MASK) != 0)
        launch_timed(k_march<PHYS, CPML, CHUNKS, true>, dim3((unsigned)blocks), st, ev, a);
    else
        launch_timed(k_march<PHYS, CPML, CHUNKS, false>, dim3((unsigned)blocks), st, ev, a);
    return (int)hipGetLastError();
}

}  // namespace pam
