// pam_tblock.hip -- k_tblock: T time steps per launch for grids too small to
// fill the device (absorbing layer: Cerjan; forward flags only).
//
// A small grid (one shot of 301 x 501 cells is 150 tiles of 64 x 16) leaves
// most of the 256 CUs idle, and each one-step launch then costs about its
// launch and drain: the step is latency-bound, not bandwidth-bound.  Here a
// block stages its 64 x 16 tile with a 4T-cell halo of U_k and U_{k-1} in LDS
// and advances it T steps there, each step over a region 4 cells smaller on
// every side (overlapped tiling: the halo cells are recomputed by every
// block that needs them, so blocks never exchange data inside a launch), then
// writes the tile's last two levels to the other pair of field buffers.
// Every cell is computed by the same helpers and expressions as the 16-row
// tile (tile_derivs, xderivs, lf_*), so the results are variant 0's bit for
// bit; source injection and the receiver gather happen at every inner step.
#include "pam_variants.cuh"

namespace pam {

template <int PHYS, int T, int NTH, int TZ>
__global__ __launch_bounds__(NTH) void k_tblock(const StepArgs a, const float *__restrict__ in_cur,
                                                const float *__restrict__ in_prv, float *__restrict__ out_cur,
                                                float *__restrict__ out_prv, const int k0, const int nsub) {
    constexpr int TX = 64, HL = 4 * T;
    constexpr int W = TX + 2 * HL, H = TZ + 2 * HL, Q = W / 4;  // staged cells, float4 per staged row
    constexpr int NM = (PHYS == 0) ? 1 : (PHYS == 1) ? 3 : 5;  // model fields: v2 [, eps, delta [, cos, sin]]
    __shared__ float4 bufA[H * Q], bufB[H * Q];
    __shared__ float4 sd[(PHYS == 2) ? H * Q : 1];
    __shared__ float4 smod[NM][H * Q];             // the model over the staged region
    __shared__ float sfx[W], sfz[H];               // Cerjan factors of the staged columns / rows
    __shared__ int srh[TZ][TX], srh2[TZ][TX];      // first / second receiver of each tile cell, or -1
    const int nxt = (a.nxp + TX - 1) / TX, nzt = (a.nzp + TZ - 1) / TZ;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= nxt * nzt * a.nshots) return;  // block-uniform
    const int tile = b / a.nshots, s = b - tile * a.nshots;
    const int tzi = tile / nxt, txi = tile - tzi * nxt;
    const int gx0 = txi * TX - HL, gz0 = tzi * TZ - HL;  // grid cell of staged (0, 0)
    const int tid = threadIdx.x;
    const int P = a.P, nzp = a.nzp, nxp = a.nxp;
    const long long so = (long long)s * a.sstride;
    const int flags = a.flags & F_FWD_MASK;

    // ---- stage U_k (A), U_{k-1} (B) and the model (zero outside the grid), the
    // Cerjan factors and the tile's receivers: the inner steps then read LDS
    // only.  Every global load of a thread is issued before the first is
    // waited for (the loops have fixed trip counts and are unrolled): the
    // fields were written by other XCDs in the previous launch, so each load
    // comes from the MALL or HBM.
    constexpr int NQS = (H * Q + NTH - 1) / NTH;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 lc[NQS], lp[NQS], lm[NQS][NM];
#pragma unroll
    for (int u = 0; u < NQS; ++u) {
        const int i = tid + NTH * u;
        const int r = i / Q, q = i - r * Q;
        const int gz = gz0 + r, gx = gx0 + 4 * q;
        lc[u] = lp[u] = z4;
#pragma unroll
        for (int f = 0; f < NM; ++f) lm[u][f] = z4;
        if (i < H * Q && gz >= 0 && gz < nzp && gx >= 0 && gx < nxp) {
            const unsigned ro = (unsigned)(gz * P + gx);
            lc[u] = *reinterpret_cast<const float4 *>(at(in_cur + so, ro));
            lp[u] = *reinterpret_cast<const float4 *>(at(in_prv + so, ro));
            lm[u][0] = *reinterpret_cast<const float4 *>(at(a.v2, ro));
            if (PHYS >= 1) {
                lm[u][1 % NM] = *reinterpret_cast<const float4 *>(at(a.eps, ro));
                lm[u][2 % NM] = *reinterpret_cast<const float4 *>(at(a.del, ro));
            }
            if (PHYS == 2) {
                lm[u][3 % NM] = *reinterpret_cast<const float4 *>(at(a.cs, ro));
                lm[u][4 % NM] = *reinterpret_cast<const float4 *>(at(a.sn, ro));
            }
        }
    }
    // Cerjan factors and the first two receivers of the tile's cells
    constexpr int NF = (W + H + NTH - 1) / NTH;
    float lf[NF];
#pragma unroll
    for (int u = 0; u < NF; ++u) {
        const int i = tid + NTH * u;
        lf[u] = 1.0f;
        if (i < W) {
            const int gx = gx0 + i;
            if ((flags & F_CERJAN) && gx >= 0 && gx < P) lf[u] = a.cerx[gx];
        } else if (i < W + H) {
            const int gz = gz0 + (i - W);
            if ((flags & F_CERJAN) && gz >= 0 && gz < nzp) lf[u] = a.cerz[gz];
        }
    }
    constexpr int NR = (TZ * TX + NTH - 1) / NTH;
    int lslot[NR], lr[NR], lr2[NR];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int i = tid + NTH * u;
        const int gz = gz0 + HL + i / TX;
        lslot[u] = ((flags & F_GATHER) && i < TZ * TX && gz < nzp) ? a.rrow[gz] : -1;
    }
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int i = tid + NTH * u;
        const int gx = gx0 + HL + (i % TX);
        lr[u] = lslot[u] >= 0 ? a.rhead[(long long)lslot[u] * P + gx] : -1;
    }
#pragma unroll
    for (int u = 0; u < NR; ++u) lr2[u] = lr[u] >= 0 ? a.rnext[lr[u]] : -1;
#pragma unroll
    for (int u = 0; u < NQS; ++u) {
        const int i = tid + NTH * u;
        if (i >= H * Q) break;
        const int q = i % Q;
        const int gx = gx0 + 4 * q;
        float4 c = lc[u], p = lp[u];
        if (gx + 3 >= nxp) {  // columns past the grid (the row pitch's padding)
            if (gx + 1 >= nxp) c.y = p.y = 0.f;
            if (gx + 2 >= nxp) c.z = p.z = 0.f;
            c.w = p.w = 0.f;
        }
        bufA[i] = c;
        bufB[i] = p;
#pragma unroll
        for (int f = 0; f < NM; ++f) smod[f][i] = lm[u][f];
    }
#pragma unroll
    for (int u = 0; u < NF; ++u) {
        const int i = tid + NTH * u;
        if (i < W) sfx[i] = lf[u];
        else if (i < W + H) sfz[i - W] = lf[u];
    }
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int i = tid + NTH * u;
        if (i >= TZ * TX) break;
        srh[i / TX][i % TX] = lr[u];
        srh2[i / TX][i % TX] = lr2[u];
    }
    __syncthreads();

    // ---- one inner step: U_{k+1} from A = U_k and B = U_{k-1}, written over B,
    // on the staged region 4 st cells inside the staged edge
    auto step = [&](const int st, const float4 *A, float4 *B) {
        const int k = k0 + st - 1;  // step index of the input field
        const int r0 = 4 * st, r1 = H - 4 * st, q0 = st, q1 = Q - st;
        const int nq = q1 - q0, ncell = (r1 - r0) * nq;
        if (PHYS == 2) {
            // first x-derivative (undivided) of the rows the mixed derivative reads
            const int nrd = (r1 - r0) + 8;
            for (int i = tid; i < nrd * nq; i += NTH) {
                const int r = r0 - 4 + i / nq, q = q0 + i % nq;
                float w[12];
                const float4 l = A[r * Q + q - 1], m = A[r * Q + q], h = A[r * Q + q + 1];
                w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
                w[4] = m.x; w[5] = m.y; w[6] = m.z; w[7] = m.w;
                w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
                float d2[4], d1[4];
                xderivs<4>(w, d2, d1);
                sd[r * Q + q] = make_float4(d1[0], d1[1], d1[2], d1[3]);
            }
            __syncthreads();
        }
        for (int i = tid; i < ncell; i += NTH) {
            const int r = r0 + i / nq, q = q0 + i % nq;
            const int gz = gz0 + r, gx = gx0 + 4 * q;
            const bool in_tile = r >= HL && r < HL + TZ && q >= T && q < T + TX / 4;
            // receiver gather of the input field (the tile's own cells only)
            if (in_tile && k >= a.itlag) {
                const int tr = r - HL, tc = 4 * (q - T);
                if (srh[tr][tc] >= 0 || srh[tr][tc + 1] >= 0 || srh[tr][tc + 2] >= 0 || srh[tr][tc + 3] >= 0) {
                    const float4 cv = A[r * Q + q];
                    const float c4[4] = {cv.x, cv.y, cv.z, cv.w};
                    const float fz = sfz[r];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float fx = sfx[4 * q + j];
                        const int r1 = srh[tr][tc + j];
                        if (r1 < 0) continue;
                        const float val = (c4[j] * fx) * fz;
                        float *row = a.seis + (long long)s * a.dstride + (long long)(k - a.itlag) * a.nrec;
                        row[r1] = val;
                        // a second receiver on the cell from LDS; the chain walked
                        // in global memory only beyond it
                        for (int rr = srh2[tr][tc + j]; rr >= 0; rr = a.rnext[rr]) row[rr] = val;
                    }
                }
            }
            if (gz < 4 || gz >= nzp - 4 || gx + 3 < 4 || gx >= nxp - 4) continue;  // no live cell
            float zw[9][4];
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                const float4 t = A[(r - 4 + h) * Q + q];
                zw[h][0] = t.x; zw[h][1] = t.y; zw[h][2] = t.z; zw[h][3] = t.w;
            }
            float w[12];
            {
                const float4 l = A[r * Q + q - 1], m = A[r * Q + q], h = A[r * Q + q + 1];
                w[0] = l.x; w[1] = l.y; w[2] = l.z; w[3] = l.w;
                w[4] = m.x; w[5] = m.y; w[6] = m.z; w[7] = m.w;
                w[8] = h.x; w[9] = h.y; w[10] = h.z; w[11] = h.w;
            }
            float d2x[4], d1x[4];
            xderivs<4>(w, d2x, d1x);
            float c0[4], pzz[4], pxx[4], px[4], pz[4], pxz[4] = {0.f, 0.f, 0.f, 0.f};
            tile_derivs<PHYS>(zw, d2x, d1x, a, c0, pzz, pxx, pz, px);
            if (PHYS == 2) {
                float qd[9][4];
#pragma unroll
                for (int h = 0; h < 9; ++h) {
                    if (h == 4) continue;
                    const float4 t = sd[(r - 4 + h) * Q + q];
                    qd[h][0] = t.x; qd[h][1] = t.y; qd[h][2] = t.z; qd[h][3] = t.w;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    pxz[j] = (KA1 * (qd[5][j] - qd[3][j]) + KA2 * (qd[6][j] - qd[2][j]) +
                              KA3 * (qd[7][j] - qd[1][j]) + KA4 * (qd[8][j] - qd[0][j])) *
                             a.idxdz;
            }
            float v2[4], e[4] = {0.f, 0.f, 0.f, 0.f}, dl[4] = {0.f, 0.f, 0.f, 0.f}, cs[4] = {0.f, 0.f, 0.f, 0.f},
                  sn[4] = {0.f, 0.f, 0.f, 0.f};
            {
                const float4 t = smod[0][r * Q + q];
                v2[0] = t.x; v2[1] = t.y; v2[2] = t.z; v2[3] = t.w;
            }
            if (PHYS >= 1) {
                const float4 t1 = smod[1 % NM][r * Q + q], t2 = smod[2 % NM][r * Q + q];
                e[0] = t1.x; e[1] = t1.y; e[2] = t1.z; e[3] = t1.w;
                dl[0] = t2.x; dl[1] = t2.y; dl[2] = t2.z; dl[3] = t2.w;
            }
            if (PHYS == 2) {
                const float4 t3 = smod[3 % NM][r * Q + q], t4 = smod[4 % NM][r * Q + q];
                cs[0] = t3.x; cs[1] = t3.y; cs[2] = t3.z; cs[3] = t3.w;
                sn[0] = t4.x; sn[1] = t4.y; sn[2] = t4.z; sn[3] = t4.w;
            }
            const float4 pv4 = B[r * Q + q];
            const float pv[4] = {pv4.x, pv4.y, pv4.z, pv4.w};
            const float fz = sfz[r];
            float out[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = gx + j;
                const bool live = c >= 4 && c < nxp - 4;
                const float fx = live ? sfx[4 * q + j] : 1.0f;
                float op, u;
                if (PHYS == 0) {
                    const float sum = pxx[j] + pzz[j];
                    u = lf_acoustic(c0[j], pv[j], fx, fz, v2[j], sum, false, op);
                } else if (PHYS == 1) {
                    u = lf_vti(c0[j], pv[j], fx, fz, v2[j], e[j], dl[j], px[j], pz[j], pxx[j], pzz[j], false, op);
                } else {
                    u = lf_tti(c0[j], pv[j], fx, fz, v2[j], e[j], dl[j], cs[j], sn[j], px[j], pz[j], pxx[j], pzz[j],
                               pxz[j], false, op);
                }
                out[j] = live ? (u * fx) * fz : pv[j];
            }
            // writer-side injection of the next step's source
            if ((flags & F_SRC) && gz == a.ssz[s] && k + 1 < a.nt) {
                const int sx = a.ssx[s];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (gx + j == sx) out[j] += a.wav[k + 1];
            }
            B[r * Q + q] = make_float4(out[0], out[1], out[2], out[3]);
        }
        __syncthreads();
    };
#pragma unroll
    for (int st = 1; st <= T; ++st) {
        if (st > nsub) break;  // block-uniform
        if (st & 1)
            step(st, bufA, bufB);
        else
            step(st, bufB, bufA);
    }

    // ---- the tile's last two levels into the other buffer pair
    const float4 *Lc = (nsub & 1) ? bufB : bufA;  // U_{k0+nsub}
    const float4 *Lp = (nsub & 1) ? bufA : bufB;  // U_{k0+nsub-1}
    for (int i = tid; i < TZ * (TX / 4); i += NTH) {
        const int r = HL + i / (TX / 4), q = T + i % (TX / 4);
        const int gz = gz0 + r, gx = gx0 + 4 * q;
        if (gz >= nzp || gx >= nxp) continue;
        const unsigned ro = (unsigned)(gz * P + gx);
        const float4 c = Lc[r * Q + q], p = Lp[r * Q + q];
        if (gx + 3 < nxp) {
            *reinterpret_cast<float4 *>(at(out_cur + so, ro)) = c;
            *reinterpret_cast<float4 *>(at(out_prv + so, ro)) = p;
        } else {
            const float cc[4] = {c.x, c.y, c.z, c.w}, pp[4] = {p.x, p.y, p.z, p.w};
            for (int j = 0; j < 4 && gx + j < nxp; ++j) {
                *at(out_cur + so, ro + j) = cc[j];
                *at(out_prv + so, ro + j) = pp[j];
            }
        }
    }
}

// nsub (1 <= nsub <= tblock_steps(phys)) steps k0 .. k0 + nsub - 1 of a.nshots shots
// from (in_cur, in_prv) = (U_k0, U_k0-1) into (out_cur, out_prv) =
// (U_k0+nsub, U_k0+nsub-1); the two pairs must not overlap
int launch_tblock(int phys, const StepArgs &a0, const float *in_cur, const float *in_prv, float *out_cur,
                  float *out_prv, int k0, int nsub, hipStream_t st) {
    if (nsub < 1 || nsub > tblock_steps(phys)) return (int)hipErrorInvalidValue;
    // a block of 1024 threads (acoustic) or 512 (VTI / TTI, more registers per
    // thread): a small grid has few tiles, and more waves per tile shorten each
    // thread's serial share of the inner steps (256 threads: 3.05 us per step
    // on a 301 x 501 acoustic grid, 1024: 2.30).  Tiles of 16 rows: 8 and 32
    // gave 3.07 and 2.73 us per step there (more halo work, or fewer blocks)
    const auto grid = [&](int tz) {
        const long long tiles = (long long)((a0.nxp + 63) / 64) * ((a0.nzp + tz - 1) / tz) * a0.nshots;
        return dim3((unsigned)((tiles + 7) / 8 * 8));  // a multiple of 8 for the XCD remap
    };
    if (phys == 0)
        hipLaunchKernelGGL((k_tblock<0, 4, 1024, 16>), grid(16), dim3(1024), 0, st, a0, in_cur,
                           in_prv, out_cur, out_prv, k0, nsub);
    else if (phys == 1)
        hipLaunchKernelGGL((k_tblock<1, 4, 512, 16>), grid(16), dim3(512), 0, st, a0, in_cur, in_prv, out_cur, out_prv,
                           k0, nsub);
    else
        hipLaunchKernelGGL((k_tblock<2, 4, 512, 16>), grid(16), dim3(512), 0, st, a0, in_cur, in_prv, out_cur, out_prv,
                           k0, nsub);
    return (int)hipGetLastError();
}

}  // namespace pam
