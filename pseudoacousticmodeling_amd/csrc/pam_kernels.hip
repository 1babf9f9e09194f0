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
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "pam_internal.h"

namespace pam {

// 8th-order coefficients, src/CudaKernels.py:272-276 and :316-319
#define KC0 (-1435.0f / 504.0f)
#define KC1 (8.0f / 5.0f)
#define KC2 (-1.0f / 5.0f)
#define KC3 (8.0f / 315.0f)
#define KC4 (-1.0f / 560.0f)
#define KA1 (4.0f / 5.0f)
#define KA2 (-1.0f / 5.0f)
#define KA3 (4.0f / 105.0f)
#define KA4 (-1.0f / 280.0f)

template <int V> struct VecOf;
template <> struct VecOf<2> { typedef float2 T; };
template <> struct VecOf<4> { typedef float4 T; };

template <int V>
__device__ __forceinline__ void vld(float *d, const float *p) {
    typedef typename VecOf<V>::T T;
    const T t = *reinterpret_cast<const T *>(p);
    const float *s = reinterpret_cast<const float *>(&t);
#pragma unroll
    for (int j = 0; j < V; ++j) d[j] = s[j];
}
template <int V>
__device__ __forceinline__ void vst(float *p, const float *d) {
    typedef typename VecOf<V>::T T;
    T t;
    float *s = reinterpret_cast<float *>(&t);
#pragma unroll
    for (int j = 0; j < V; ++j) s[j] = d[j];
    *reinterpret_cast<T *>(p) = t;
}
// vector load of the V cells starting at column `col` of a row, zero when the
// vector lies at or beyond the pitch (columns >= P do not exist)
template <int V>
__device__ __forceinline__ void vldg(float *d, const float *rowp, int col, int P) {
    if (col < P) {
        vld<V>(d, rowp + col);
    } else {
#pragma unroll
        for (int j = 0; j < V; ++j) d[j] = 0.0f;
    }
}
// row segment with a 4-cell halo each side: w[0..V+7], cell j of the lane at w[4+j]
template <int V>
__device__ __forceinline__ void load_row_halo(float *w, const float *rowp, int col, int P) {
#pragma unroll
    for (int h = 0; h < 4 / V; ++h) vldg<V>(w + h * V, rowp, col - 4 + h * V, P);
    vldg<V>(w + 4, rowp, col, P);
#pragma unroll
    for (int h = 0; h < 4 / V; ++h) vldg<V>(w + 4 + V + h * V, rowp, col + V + h * V, P);
}
template <int V>
__device__ __forceinline__ void xderivs(const float *w, float *d2, float *d1) {
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const float *c = w + 4 + j;
        d2[j] = KC0 * c[0] + KC1 * (c[1] + c[-1]) + KC2 * (c[2] + c[-2]) + KC3 * (c[3] + c[-3]) +
                KC4 * (c[4] + c[-4]);
        d1[j] = KA1 * (c[1] - c[-1]) + KA2 * (c[2] - c[-2]) + KA3 * (c[3] - c[-3]) + KA4 * (c[4] - c[-4]);
    }
}

__device__ __forceinline__ float sd_term(float e, float dl, float px, float pz) {
    // Sd of src/CudaKernels.py:354-360 with the |den| > 1e-12 gate
    const float px2 = px * px, pz2 = pz * pz;
    const float num = -2.0f * (e - dl) * px2 * pz2;
    const float den = (1.0f + 2.0f * e) * (px2 * px2) + (pz2 * pz2) + 2.0f * (1.0f + dl) * px2 * pz2;
    return (fabsf(den) > 1e-12f) ? num / den : 0.0f;
}

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// works on a contiguous chunk of (tile, shot) space and its L2 sees the halo
// and model reuse.  nblocks is a multiple of 8.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    const int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename T>
__device__ __forceinline__ float sv_load(const void *p, long long i, float inv_scale);
template <>
__device__ __forceinline__ float sv_load<float>(const void *p, long long i, float) {
    return reinterpret_cast<const float *>(p)[i];
}
template <>
__device__ __forceinline__ float sv_load<__half>(const void *p, long long i, float inv_scale) {
    return __half2float(reinterpret_cast<const __half *>(p)[i]) * inv_scale;
}
template <>
__device__ __forceinline__ float sv_load<__hip_bfloat16>(const void *p, long long i, float inv_scale) {
    return __bfloat162float(reinterpret_cast<const __hip_bfloat16 *>(p)[i]) * inv_scale;
}

// snapshot value of cell (interior index ii, padded index pi) at step k for shot s
__device__ __forceinline__ float snap_get(const StepArgs &a, int s, int k, long long ii, long long pi) {
    const long long base = (long long)s * a.save_sstride + (long long)k * a.save_kstride;
    switch (a.save_kind) {
        case SV_F32_INT:
            return sv_load<float>(a.save, base + ii, 1.0f);
        case SV_F32_PAD:
            return sv_load<float>(a.save, base + pi, 1.0f);
        case SV_F16_INT:
            return sv_load<__half>(a.save, base + ii, 1.0f / a.save_scale[(long long)s * a.nt + k]);
        case SV_BF16_INT:
            return sv_load<__hip_bfloat16>(a.save, base + ii, 1.0f / a.save_scale[(long long)s * a.nt + k]);
        default:
            return 0.0f;
    }
}
__device__ __forceinline__ void snap_put(const StepArgs &a, int s, int k, long long ii, long long pi, float v) {
    const long long base = (long long)s * a.save_sstride + (long long)k * a.save_kstride;
    switch (a.save_kind) {
        case SV_F32_INT:
            reinterpret_cast<float *>(a.save)[base + ii] = v;
            break;
        case SV_F32_PAD:
            reinterpret_cast<float *>(a.save)[base + pi] = v;
            break;
        case SV_F16_INT:
            reinterpret_cast<__half *>(a.save)[base + ii] = __float2half(v * a.save_scale[(long long)s * a.nt + k]);
            break;
        case SV_BF16_INT:
            reinterpret_cast<__hip_bfloat16 *>(a.save)[base + ii] =
                __float2bfloat16(v * a.save_scale[(long long)s * a.nt + k]);
            break;
        default:
            break;
    }
}

// ---------------------------------------------------------------- the step
// PHYS: 0 acoustic, 1 VTI, 2 TTI.  CPML: memory-variable boundary (else Cerjan
// or none, by flag).  V cells per lane along x, TZ rows per wavefront.
template <int PHYS, bool CPML, int V, int TZ>
__global__ __launch_bounds__(256) void k_step(const StepArgs a) {
    constexpr int W = 64 * V;
    const int lane = threadIdx.x & 63;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int wid = b * 4 + (threadIdx.x >> 6);
    const int ntile = a.nxt * a.nzt;
    if (wid >= ntile * a.nshots) return;  // the whole wavefront leaves together
    const int tile = wid / a.nshots;
    const int s = wid - tile * a.nshots;  // shots of one tile are adjacent: the model is reused from L2
    const int tz = tile / a.nxt;
    const int tx = tile - tz * a.nxt;
    const int x0 = 4 + tx * W;
    const int z0 = 4 + tz * TZ;
    const int col = x0 + lane * V;
    const int P = a.P, nzp = a.nzp, nxp = a.nxp, N = a.N;
    const long long so = (long long)s * a.sstride;
    const float *__restrict__ cur = a.cur + so;
    float *__restrict__ prv = a.prv + so;
    const int flags = a.flags;

    // queues: U rows z-4..z+4, their first x-derivatives (VTI/TTI), second
    // x-derivatives of rows z..z+4
    float q[9][V], qd1[9][V], qd2[5][V];

#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int zr = z0 - 4 + r;
        float w[V + 8];
        if (zr < nzp) {
            load_row_halo<V>(w, cur + (long long)zr * P, col, P);
        } else {
#pragma unroll
            for (int j = 0; j < V + 8; ++j) w[j] = 0.0f;
        }
        float d2[V], d1[V];
        xderivs<V>(w, d2, d1);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            q[r][j] = w[4 + j];
            qd1[r][j] = d1[j];
            if (r >= 4) qd2[r - 4][j] = d2[j];
        }
    }

    float fx[V];
    bool live[V], xs[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const int c = col + j;
        live[j] = (c < nxp - 4);
        fx[j] = ((flags & F_CERJAN) && c < P) ? a.cerx[c] : 1.0f;
        xs[j] = CPML && ((c >= 4 && c < N) || (c >= nxp - N && c < nxp - 4));
    }
    const bool tile_live = col < nxp - 4;
    const bool tile_xstrip = CPML && (x0 < N || x0 + W > nxp - N);
    const bool inter_x = (col + V > N) && (col < nxp - N);  // some cell of the lane is interior in x

#pragma unroll
    for (int r = 0; r < TZ; ++r) {
        const int z = z0 + r;
        {   // row z+4 enters the queues
            const int ze = z + 4;
            float w[V + 8];
            if (ze < nzp) {
                load_row_halo<V>(w, cur + (long long)ze * P, col, P);
            } else {
#pragma unroll
                for (int j = 0; j < V + 8; ++j) w[j] = 0.0f;
            }
            float d2[V], d1[V];
            xderivs<V>(w, d2, d1);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                q[(r + 8) % 9][j] = w[4 + j];
                qd1[(r + 8) % 9][j] = d1[j];
                qd2[(r + 4) % 5][j] = d2[j];
            }
        }
        if (z >= nzp - 4) break;  // wavefront-uniform
        const int qc = (r + 4) % 9;
        const long long ro = (long long)z * P + col;
        const float fz = (flags & F_CERJAN) ? a.cerz[z] : 1.0f;

        float c0[V], pzz[V], pxx[V], px[V], pz[V], pxz[V];
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const float cc = q[qc][j];
            const float u1p = q[(r + 5) % 9][j], u1m = q[(r + 3) % 9][j];
            const float u2p = q[(r + 6) % 9][j], u2m = q[(r + 2) % 9][j];
            const float u3p = q[(r + 7) % 9][j], u3m = q[(r + 1) % 9][j];
            const float u4p = q[(r + 8) % 9][j], u4m = q[(r + 0) % 9][j];
            c0[j] = cc;
            pzz[j] = (KC0 * cc + KC1 * (u1p + u1m) + KC2 * (u2p + u2m) + KC3 * (u3p + u3m) + KC4 * (u4p + u4m)) *
                     a.idz2;
            pxx[j] = qd2[r % 5][j] * a.idx2;
            px[j] = 0.0f;
            pz[j] = 0.0f;
            pxz[j] = 0.0f;
            if (PHYS >= 1) {
                pz[j] = (KA1 * (u1p - u1m) + KA2 * (u2p - u2m) + KA3 * (u3p - u3m) + KA4 * (u4p - u4m)) * a.idz;
                px[j] = qd1[qc][j] * a.idx;
            }
            if (PHYS == 2) {
                pxz[j] = (KA1 * (qd1[(r + 5) % 9][j] - qd1[(r + 3) % 9][j]) +
                          KA2 * (qd1[(r + 6) % 9][j] - qd1[(r + 2) % 9][j]) +
                          KA3 * (qd1[(r + 7) % 9][j] - qd1[(r + 1) % 9][j]) +
                          KA4 * (qd1[(r + 8) % 9][j] - qd1[(r + 0) % 9][j])) *
                         a.idxdz;
            }
        }

        // ---- receiver gather of the input field (damped as the reference records it)
        if (flags & F_GATHER) {
            const int slot = a.rrow[z];
            if (slot >= 0 && tile_live) {
                const int it = a.k - a.itlag;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    if (col + j >= P) continue;
                    int rr = a.rhead[(long long)slot * P + col + j];
                    const float val = (c0[j] * fx[j]) * fz;
                    while (rr >= 0) {
                        a.seis[(long long)s * a.dstride + (long long)it * a.nrec + rr] = val;
                        rr = a.rnext[rr];
                    }
                }
            }
        }

        // ---- snapshots / illumination / imaging of the input field
        const bool zint = (z >= N) && (z < nzp - N);
        if ((flags & (F_SAVE | F_ILUM | F_IMAGE | F_GRADM)) && tile_live) {
            const long long ib = (long long)(z - N) * a.nx - N;  // interior index = ib + c
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int c = col + j;
                const bool cin = zint && c >= N && c < nxp - N;
                const long long ii = ib + c;
                if (flags & F_SAVE) {
                    if (a.save_kind == SV_F32_PAD) {
                        if (live[j]) snap_put(a, s, a.k, 0, ro + j, (c0[j] * fx[j]) * fz);
                    } else if (cin) {
                        snap_put(a, s, a.k, ii, 0, c0[j]);
                    }
                }
                if (!cin) continue;
                float *acc = nullptr;
                if (flags & F_ILUM) {
                    acc = a.ilum + (long long)s * a.acc_sstride + ii;
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
        (void)inter_x;

        if (!tile_live) continue;

        // ---- CPML: zeta with the new psi, corrected derivatives
        float PXX[V], PZZ[V], PX[V], PZ[V], zx[V], zz[V], dpx[V], dpz[V];
        const bool zs = CPML && ((z >= 4 && z < N) || (z >= nzp - N && z < nzp - 4));
#pragma unroll
        for (int j = 0; j < V; ++j) {
            PXX[j] = pxx[j];
            PZZ[j] = pzz[j];
            PX[j] = px[j];
            PZ[j] = pz[j];
            zx[j] = 0.0f;
            zz[j] = 0.0f;
            dpx[j] = 0.0f;
            dpz[j] = 0.0f;
        }
        if (CPML && tile_xstrip) {
            float w[V + 8], d2p[V], d1p[V], zo[V], ax[V], bx[V];
            load_row_halo<V>(w, a.psix + so + (long long)z * P, col, P);
            xderivs<V>(w, d2p, d1p);
            vld<V>(zo, a.zetax + so + ro);
            vld<V>(ax, a.cax + ro);
            vld<V>(bx, a.cbx + ro);
#pragma unroll
            for (int j = 0; j < V; ++j) {
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
            float ps[9][V];
#pragma unroll
            for (int h = 0; h < 9; ++h) {
                if (h == 4) continue;
                vld<V>(ps[h], a.psiz + so + (long long)(z - 4 + h) * P + col);
            }
            float zo[V], az[V], bz[V];
            vld<V>(zo, a.zetaz + so + ro);
            vld<V>(az, a.caz + ro);
            vld<V>(bz, a.cbz + ro);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const float dps = (KA1 * (ps[5][j] - ps[3][j]) + KA2 * (ps[6][j] - ps[2][j]) +
                                   KA3 * (ps[7][j] - ps[1][j]) + KA4 * (ps[8][j] - ps[0][j])) *
                                  a.idz;
                const bool c4 = (col + j >= 4) && live[j];
                zz[j] = az[j] * zo[j] + bz[j] * (pzz[j] + dps);
                dpz[j] = dps;
                PZ[j] = pz[j] + dps;
                PZZ[j] = pzz[j] + dps + zz[j];
                if (c4) a.zetaz[so + ro + j] = zz[j];
            }
        }

        // ---- the update
        float pv[V], v2[V], out[V];
        vld<V>(pv, prv + ro);
        vld<V>(v2, a.v2 + ro);
        float e[V], dl[V], cs[V], sn[V];
        if (PHYS >= 1) {
            vld<V>(e, a.eps + ro);
            vld<V>(dl, a.del + ro);
        }
        if (PHYS == 2) {
            vld<V>(cs, a.cs + ro);
            vld<V>(sn, a.sn + ro);
        }
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const float pe = (pv[j] * fx[j]) * fz;  // U_{k-1} as the reference holds it (damped twice)
            float u;
            if (PHYS == 0) {
                float sum = pxx[j] + pzz[j];
                if (CPML) {
                    if (xs[j]) sum += dpx[j];
                    if (zs) sum += dpz[j];
                    if (xs[j]) sum += zx[j];
                    if (zs) sum += zz[j];
                }
                u = v2[j] * sum + 2.0f * c0[j] - pe;
            } else if (PHYS == 1) {
                const float Sd = sd_term(e[j], dl[j], PX[j], PZ[j]);
                u = 2.0f * c0[j] - pe + v2[j] * ((1.0f + 2.0f * e[j]) + Sd) * PXX[j] + v2[j] * (1.0f + Sd) * PZZ[j];
            } else {
                const float xi = PX[j] * cs[j] - PZ[j] * sn[j];
                const float eta = PX[j] * sn[j] + PZ[j] * cs[j];
                const float Sd = sd_term(e[j], dl[j], xi, eta);
                const float Ac = (1.0f + 2.0f * e[j]) * (cs[j] * cs[j]) + (sn[j] * sn[j]) + Sd;
                const float Bc = (1.0f + 2.0f * e[j]) * (sn[j] * sn[j]) + (cs[j] * cs[j]) + Sd;
                u = 2.0f * c0[j] - pe + v2[j] * Ac * PXX[j] + v2[j] * Bc * PZZ[j] -
                    2.0f * e[j] * v2[j] * (2.0f * sn[j] * cs[j]) * pxz[j];
            }
            out[j] = (u * fx[j]) * fz;
        }
        // writer-side injection of the next step's source
        if ((flags & F_SRC) && z == a.ssz[s] && a.k + 1 < a.nt) {
            const int sx = a.ssx[s];
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (col + j == sx) out[j] += a.wav[a.k + 1];
        }
        // writer-side injection of the data of the next (earlier) adjoint step
        if (flags & F_RECINJ) {
            const int slot = a.rrow[z];
            const int tn = a.k - 1 - a.itlag;
            if (slot >= 0 && tn >= 0) {
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    if (col + j >= P) continue;
                    int rr = a.rhead[(long long)slot * P + col + j];
                    while (rr >= 0) {
                        out[j] += a.data[(long long)s * a.dstride + (long long)tn * a.nrec + rr] / a.dxdz;
                        rr = a.rnext[rr];
                    }
                }
            }
        }
        if (live[V - 1]) {
            vst<V>(prv + ro, out);
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (live[j]) prv[ro + j] = out[j];
        }
    }
}

// ---------------------------------------------------------------- CPML psi
// psi = a psi + b dU on the strips, from the current field (src/CudaKernels.py:92-163).
// Threads enumerate x-strip cells (rows [4,nzp-4) x 2(N-4) columns) followed by
// z-strip cells (2(N-4) rows x columns [4,nxp-4)).
__global__ __launch_bounds__(256) void k_psi(const PsiArgs a) {
    const int s = blockIdx.y;
    const long long so = (long long)s * a.sstride;
    const int wN = a.N - 4;
    const long long nxs = (long long)(a.nzp - 8) * 2 * wN;
    const long long nzs = (long long)2 * wN * (a.nxp - 8);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const float *U = a.cur + so;
    if (t < nxs) {
        const int row = 4 + (int)(t / (2 * wN));
        const int cc = (int)(t % (2 * wN));
        const int ix = (cc < wN) ? 4 + cc : a.nxp - a.N + (cc - wN);
        const long long i = (long long)row * a.P + ix;
        const float d1 = KA1 * (U[i + 1] - U[i - 1]) + KA2 * (U[i + 2] - U[i - 2]) + KA3 * (U[i + 3] - U[i - 3]) +
                         KA4 * (U[i + 4] - U[i - 4]);
        a.psix[so + i] = a.cax[i] * a.psix[so + i] + a.cbx[i] * (d1 * a.idx);
    } else if (t < nxs + nzs) {
        const long long u = t - nxs;
        const int rr = (int)(u / (a.nxp - 8));
        const int ix = 4 + (int)(u % (a.nxp - 8));
        const int iz = (rr < wN) ? 4 + rr : a.nzp - a.N + (rr - wN);
        const long long i = (long long)iz * a.P + ix;
        const long long P = a.P;
        const float d1 = KA1 * (U[i + P] - U[i - P]) + KA2 * (U[i + 2 * P] - U[i - 2 * P]) +
                         KA3 * (U[i + 3 * P] - U[i - 3 * P]) + KA4 * (U[i + 4 * P] - U[i - 4 * P]);
        a.psiz[so + i] = a.caz[i] * a.psiz[so + i] + a.cbz[i] * (d1 * a.idz);
    }
}

// ------------------------------------------- linearised VTI/TTI adjoint (Cerjan)
// Products from the source field P (src/CudaKernels.py:1382-1472, :1534-1633):
// tmp[0..4] = A*lam, B*lam, Q*Cx*lam, Q*Cz*lam, H*lam on the interior.
template <int PHYS>
__global__ __launch_bounds__(256) void k_adj_products(const AdjArgs a) {
    const int s = blockIdx.z;
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iz = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= a.nxp || iz >= a.nzp) return;
    const long long i = (long long)iz * a.P + ix;
    const long long pl = (long long)a.nzp * a.P;
    float *T = a.tmp + (long long)s * a.tsstride;
    float o0 = 0.f, o1 = 0.f, o2 = 0.f, o3 = 0.f, o4 = 0.f;
    if (ix >= 4 && ix < a.nxp - 4 && iz >= 4 && iz < a.nzp - 4) {
        const float *Pp = a.P_ + (long long)s * a.psstride;
        const long long P = a.P;
        const float pxx = (KC0 * Pp[i] + KC1 * (Pp[i + 1] + Pp[i - 1]) + KC2 * (Pp[i + 2] + Pp[i - 2]) +
                           KC3 * (Pp[i + 3] + Pp[i - 3]) + KC4 * (Pp[i + 4] + Pp[i - 4])) *
                          a.idx2;
        const float pzz = (KC0 * Pp[i] + KC1 * (Pp[i + P] + Pp[i - P]) + KC2 * (Pp[i + 2 * P] + Pp[i - 2 * P]) +
                           KC3 * (Pp[i + 3 * P] + Pp[i - 3 * P]) + KC4 * (Pp[i + 4 * P] + Pp[i - 4 * P])) *
                          a.idz2;
        const float px = (KA1 * (Pp[i + 1] - Pp[i - 1]) + KA2 * (Pp[i + 2] - Pp[i - 2]) +
                          KA3 * (Pp[i + 3] - Pp[i - 3]) + KA4 * (Pp[i + 4] - Pp[i - 4])) *
                         a.idx;
        const float pz = (KA1 * (Pp[i + P] - Pp[i - P]) + KA2 * (Pp[i + 2 * P] - Pp[i - 2 * P]) +
                          KA3 * (Pp[i + 3 * P] - Pp[i - 3 * P]) + KA4 * (Pp[i + 4 * P] - Pp[i - 4 * P])) *
                         a.idz;
        const float e = a.eps[i], dl = a.del[i];
        float X = px, Z = pz, c = 1.f, sn = 0.f;
        if (PHYS == 2) {
            c = a.cs[i];
            sn = a.sn[i];
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
        const float lam = a.cur[(long long)s * a.sstride + i];
        const float Q = pxx + pzz;
        if (PHYS == 1) {
            o0 = (1.0f + 2.0f * e + Sd) * lam;
            o1 = (1.0f + Sd) * lam;
        } else {
            o0 = ((1.0f + 2.0f * e) * (c * c) + (sn * sn) + Sd) * lam;
            o1 = ((1.0f + 2.0f * e) * (sn * sn) + (c * c) + Sd) * lam;
            o4 = (2.0f * e * (2.0f * sn * c)) * lam;
        }
        o2 = Q * Cx * lam;
        o3 = Q * Cz * lam;
    }
    T[i] = o0;
    T[pl + i] = o1;
    T[2 * pl + i] = o2;
    T[3 * pl + i] = o3;
    if (PHYS == 2) T[4 * pl + i] = o4;
}

// lambda_new = 2 lam - lam_prev + V (dxx(A lam) + dzz(B lam) [- dxz(H lam)] - dx(QCx lam) - dz(QCz lam)),
// Cerjan damping, data injection of the next adjoint step, imaging with S_t.
template <int PHYS>
__global__ __launch_bounds__(256) void k_adj_update(const AdjArgs a) {
    const int s = blockIdx.z;
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iz = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix < 4 || ix >= a.nxp - 4 || iz < 4 || iz >= a.nzp - 4) return;
    const long long P = a.P;
    const long long i = (long long)iz * P + ix;
    const long long pl = (long long)a.nzp * P;
    const float *T0 = a.tmp + (long long)s * a.tsstride;
    const float *T1 = T0 + pl, *T2 = T0 + 2 * pl, *T3 = T0 + 3 * pl, *T4 = T0 + 4 * pl;
    const float dxxA = (KC0 * T0[i] + KC1 * (T0[i + 1] + T0[i - 1]) + KC2 * (T0[i + 2] + T0[i - 2]) +
                        KC3 * (T0[i + 3] + T0[i - 3]) + KC4 * (T0[i + 4] + T0[i - 4])) *
                       a.idx2;
    const float dzzB = (KC0 * T1[i] + KC1 * (T1[i + P] + T1[i - P]) + KC2 * (T1[i + 2 * P] + T1[i - 2 * P]) +
                        KC3 * (T1[i + 3 * P] + T1[i - 3 * P]) + KC4 * (T1[i + 4 * P] + T1[i - 4 * P])) *
                       a.idz2;
    const float dxQ = (KA1 * (T2[i + 1] - T2[i - 1]) + KA2 * (T2[i + 2] - T2[i - 2]) + KA3 * (T2[i + 3] - T2[i - 3]) +
                       KA4 * (T2[i + 4] - T2[i - 4])) *
                      a.idx;
    const float dzQ = (KA1 * (T3[i + P] - T3[i - P]) + KA2 * (T3[i + 2 * P] - T3[i - 2 * P]) +
                       KA3 * (T3[i + 3 * P] - T3[i - 3 * P]) + KA4 * (T3[i + 4 * P] - T3[i - 4 * P])) *
                      a.idz;
    float op;
    if (PHYS == 2) {
        float acc = 0.f;
        const float aa[4] = {KA1, KA2, KA3, KA4};
#pragma unroll
        for (int ii = 1; ii <= 4; ++ii)
#pragma unroll
            for (int jj = 1; jj <= 4; ++jj)
                acc += aa[ii - 1] * aa[jj - 1] *
                       (T4[i + jj * P + ii] - T4[i - jj * P + ii] + T4[i - jj * P - ii] - T4[i + jj * P - ii]);
        op = dxxA + dzzB - acc * a.idxdz - dxQ - dzQ;
    } else {
        op = dxxA + dzzB - dxQ - dzQ;
    }
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
        const float S = a.P_[(long long)s * a.psstride + i];
        a.image[(long long)s * a.acc_sstride + ii] += S * lam;
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
// residual = dobs - dcal and a float64 sum of residual^2 / 2
__global__ void k_residual(const float *dobs, const float *dcal, float *res, long long n, double *acc) {
    __shared__ double red[256];
    double sum = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float r = dobs[i] - dcal[i];
        if (res) res[i] = r;
        sum += 0.5 * (double)r * (double)r;
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, red[0]);
}
// multiparameter partials on interior arrays, float64 arithmetic
// (src/CudaKernels.py:1174-1379)
template <int PHYS>
__global__ void k_grad_multi(const float *S, const float *lam, int nz, int nx, float dx, float dz, const float *eps,
                             const float *del, const float *theta, float *geps, float *gdel, float *gth) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix < 4 || ix >= nx - 4 || iz < 4 || iz >= nz - 4) return;
    const long long i = (long long)iz * nx + ix;
    const long long n = nx;
    const float *c = S + i;
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
    const double e = eps[i], dl = del[i];
    const double adj = lam[i];
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
        geps[i] = (float)((double)geps[i] + adj * ((-2.0 - sde) * pxx - sde * pzz));
        gdel[i] = (float)((double)gdel[i] + adj * (-sdd * (pxx + pzz)));
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
        const double th = theta[i];
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
        geps[i] = (float)((double)geps[i] +
                          adj * ((-(2.0 * cs * cs + sde)) * pxx - (2.0 * sn * sn + sde) * pzz + 2.0 * s2t * pxz));
        gdel[i] = (float)((double)gdel[i] + adj * (-sdd * (pxx + pzz)));
        gth[i] = (float)((double)gth[i] + adj * (dA * pxx + dB * pzz + dC * pxz));
    }
}

// ---------------------------------------------------------------- launchers
static const TileShape kTiles[] = {{4, 16}, {4, 32}, {2, 16}, {2, 32}, {4, 8}, {2, 8}};
int tile_count() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }
TileShape tile_shape(int idx) {
    if (idx < 0 || idx >= tile_count()) idx = 0;
    return kTiles[idx];
}

template <int PHYS, bool CPML, int V, int TZ>
static int launch_step_t(const StepArgs &a0, hipStream_t st) {
    StepArgs a = a0;
    a.nxt = (a.nxp - 8 + 64 * V - 1) / (64 * V);
    a.nzt = (a.nzp - 8 + TZ - 1) / TZ;
    const long long waves = (long long)a.nxt * a.nzt * a.nshots;
    long long blocks = (waves + 3) / 4;
    blocks = (blocks + 7) / 8 * 8;
    hipLaunchKernelGGL((k_step<PHYS, CPML, V, TZ>), dim3((unsigned)blocks), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

template <int PHYS, bool CPML>
static int launch_step_p(int tile, const StepArgs &a, hipStream_t st) {
    switch (tile) {
        case 0: return launch_step_t<PHYS, CPML, 4, 16>(a, st);
        case 1: return launch_step_t<PHYS, CPML, 4, 32>(a, st);
        case 2: return launch_step_t<PHYS, CPML, 2, 16>(a, st);
        case 3: return launch_step_t<PHYS, CPML, 2, 32>(a, st);
        case 4: return launch_step_t<PHYS, CPML, 4, 8>(a, st);
        case 5: return launch_step_t<PHYS, CPML, 2, 8>(a, st);
        default: return (int)hipErrorInvalidValue;
    }
}

int launch_step(int phys, bool cpml, int tile, const StepArgs &a, hipStream_t st) {
    if (phys == 0) return cpml ? launch_step_p<0, true>(tile, a, st) : launch_step_p<0, false>(tile, a, st);
    if (phys == 1) return cpml ? launch_step_p<1, true>(tile, a, st) : launch_step_p<1, false>(tile, a, st);
    return cpml ? launch_step_p<2, true>(tile, a, st) : launch_step_p<2, false>(tile, a, st);
}

int launch_psi(const PsiArgs &a, hipStream_t st) {
    const long long n = (long long)(a.nzp - 8) * 2 * (a.N - 4) + (long long)2 * (a.N - 4) * (a.nxp - 8);
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_psi, dim3((unsigned)((n + 255) / 256), a.nshots), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

int launch_adj_products(int phys, const AdjArgs &a, hipStream_t st) {
    dim3 g((a.nxp + 63) / 64, (a.nzp + 3) / 4, a.nshots);
    if (phys == 1) hipLaunchKernelGGL(k_adj_products<1>, g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_adj_products<2>, g, dim3(256), 0, st, a);
    return (int)hipGetLastError();
}
int launch_adj_update(int phys, const AdjArgs &a, hipStream_t st) {
    dim3 g((a.nxp + 63) / 64, (a.nzp + 3) / 4, a.nshots);
    if (phys == 1) hipLaunchKernelGGL(k_adj_update<1>, g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_adj_update<2>, g, dim3(256), 0, st, a);
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
int launch_maxabs(const float *x, int n, float *out, hipStream_t st) {
    hipMemsetAsync(out, 0, sizeof(float), st);
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
int launch_residual(const float *dobs, const float *dcal, float *res, long long n, double *acc, hipStream_t st) {
    long long blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_residual, dim3((unsigned)blocks), dim3(256), 0, st, dobs, dcal, res, n, acc);
    return (int)hipGetLastError();
}
int launch_grad_multi(int phys, const float *S, long long sk, const float *lam, long long lk, int nz, int nx, float dx,
                      float dz, const float *eps, const float *del, const float *theta, float *geps, float *gdel,
                      float *gth, hipStream_t st) {
    (void)sk;
    (void)lk;
    dim3 g((nx + 255) / 256, nz);
    if (phys == 1)
        hipLaunchKernelGGL(k_grad_multi<1>, g, dim3(256), 0, st, S, lam, nz, nx, dx, dz, eps, del, theta, geps, gdel,
                           gth);
    else
        hipLaunchKernelGGL(k_grad_multi<2>, g, dim3(256), 0, st, S, lam, nz, nx, dx, dz, eps, del, theta, geps, gdel,
                           gth);
    return (int)hipGetLastError();
}

}  // namespace pam
