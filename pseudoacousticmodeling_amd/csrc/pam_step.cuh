// pam_step.cuh -- the fused time-step kernel k_step and its device helpers,
// included by the per-physics translation units step_*.hip so that the
// 36 template instances compile in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <type_traits>

#include <hip/hip_ext.h>

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

// base + idx elements with the byte offset formed in 32 bits, so that the
// access can use a wave-uniform 64-bit base and a 32-bit lane offset (the
// offset of one shot's field stays far below 4 GiB)
template <typename T>
__device__ __forceinline__ T *at(T *base, unsigned idx) {
    typedef typename std::conditional<std::is_const<T>::value, const char, char>::type B;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(base) + idx * (unsigned)sizeof(T));
}

// Launches kernel(args...) on st; with ev (two events), the launch itself
// records the kernel's start and end in ev[0] / ev[1] (hipExtLaunchKernelGGL),
// so that the interval is the kernel's execution without the dispatch gap
// that events recorded around the launch would add.
template <typename K, typename... Args>
static inline void launch_timed(K kernel, dim3 grid, hipStream_t st, const hipEvent_t *ev, Args... args) {
    if (ev)
        hipExtLaunchKernelGGL(kernel, grid, dim3(256), 0, st, ev[0], ev[1], 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, st, args...);
}

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
__device__ __forceinline__ void f4_to(float *w, float4 t) {
    w[0] = t.x;
    w[1] = t.y;
    w[2] = t.z;
    w[3] = t.w;
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
    // v_rcp_f32 (1 ulp) instead of the IEEE division sequence: den > 1e-12 here
    return (fabsf(den) > 1e-12f) ? num * __builtin_amdgcn_rcpf(den) : 0.0f;
}

// The leapfrog update of one cell before the Cerjan factors,
// U_{k+1} = 2 U_k - (U_{k-1} fx fz) + v^2 dt^2 L(U_k)  (c0 = U_k, pv = U_{k-1}
// as stored, damped once more as the reference holds it; v2 = v^2 dt^2), and,
// when `want_op`, the operator term v^2 dt^2 L(U_k) alone into `op`.  Every
// fused multiply-add is written out, so the update rounds the same in every
// instance (forward-only or storing the operator term) whatever contraction
// the compiler would choose; the build also restricts contraction to single
// expressions.
__device__ __forceinline__ float lf_base(float c0, float pv, float fx, float fz) {
    // factors known to be 1 (CPML, interior tiles): 2 c0 - pv in one rounding
    // either way (2 c0 is exact), so both forms give the same bits
    if (__builtin_constant_p(fx) && __builtin_constant_p(fz) && fx == 1.0f && fz == 1.0f)
        return __builtin_fmaf(2.0f, c0, -pv);
    return __builtin_fmaf(-(pv * fx), fz, 2.0f * c0);
}
// acoustic: sum = U_xx + U_zz (with the CPML corrections)
__device__ __forceinline__ float lf_acoustic(float c0, float pv, float fx, float fz, float v2, float sum, bool want_op,
                                             float &op) {
    if (want_op) op = v2 * sum;
    return __builtin_fmaf(v2, sum, lf_base(c0, pv, fx, fz));
}
// VTI (src/CudaKernels.py:361): (1 + 2 eps + Sd) U_xx + (1 + Sd) U_zz
__device__ __forceinline__ float lf_vti(float c0, float pv, float fx, float fz, float v2, float e, float dl, float PX,
                                        float PZ, float PXX, float PZZ, bool want_op, float &op) {
    const float Sd = sd_term(e, dl, PX, PZ);
    const float ax = v2 * ((1.0f + 2.0f * e) + Sd), az = v2 * (1.0f + Sd);
    if (want_op) op = __builtin_fmaf(ax, PXX, az * PZZ);
    return __builtin_fmaf(az, PZZ, __builtin_fmaf(ax, PXX, lf_base(c0, pv, fx, fz)));
}
// TTI: the VTI operator in the frame rotated by theta (cs, sn = cos, sin)
__device__ __forceinline__ float lf_tti(float c0, float pv, float fx, float fz, float v2, float e, float dl, float cs,
                                        float sn, float PX, float PZ, float PXX, float PZZ, float pxz, bool want_op,
                                        float &op) {
    const float xi = PX * cs - PZ * sn;
    const float eta = PX * sn + PZ * cs;
    const float Sd = sd_term(e, dl, xi, eta);
    const float Ac = (1.0f + 2.0f * e) * (cs * cs) + (sn * sn) + Sd;
    const float Bc = (1.0f + 2.0f * e) * (sn * sn) + (cs * cs) + Sd;
    const float tx = v2 * Ac, tz = v2 * Bc, txz = (2.0f * e * v2) * (2.0f * sn * cs);
    if (want_op) op = __builtin_fmaf(-txz, pxz, __builtin_fmaf(tz, PZZ, tx * PXX));
    return __builtin_fmaf(-txz, pxz, __builtin_fmaf(tz, PZZ, __builtin_fmaf(tx, PXX, lf_base(c0, pv, fx, fz))));
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

// the scale of the compressed snapshot of step k for shot s: the max over its
// kSmaxSlots words (non-negative floats order like their bits; the loads are
// wave-uniform)
__device__ __forceinline__ float snap_scale(const StepArgs &a, int s, int k) {
    const unsigned *w = reinterpret_cast<const unsigned *>(a.save_scale) + ((long long)s * a.nt + k) * kSmaxSlots;
    unsigned m = 0u;
#pragma unroll
    for (int i = 0; i < kSmaxSlots; ++i) m = max(m, w[i]);
    return __uint_as_float(m);
}

// snapshot value of cell (interior index ii, padded index pi) at step k for shot s.
// Compressed kinds store S / max|S_k| (max over the interior, save_scale).
__device__ __forceinline__ float snap_get(const StepArgs &a, int s, int k, long long ii, long long pi) {
    const long long base = (long long)s * a.save_sstride + (long long)(k - a.save_k0) * a.save_kstride;
    switch (a.save_kind) {
        case SV_F32_INT:
            return sv_load<float>(a.save, base + ii, 1.0f);
        case SV_F32_PAD:
            return sv_load<float>(a.save, base + pi, 1.0f);
        case SV_F16_INT:
            return __half2float(reinterpret_cast<const __half *>(a.save)[base + ii]) * snap_scale(a, s, k);
        case SV_BF16_INT:
            return __bfloat162float(reinterpret_cast<const __hip_bfloat16 *>(a.save)[base + ii]) *
                   snap_scale(a, s, k);
        default:
            return 0.0f;
    }
}
// The same with the slot's element offset `off` (wave-uniform, snap_off) and
// 32-bit lane indices: the access is a uniform base plus a 32-bit lane offset.
__device__ __forceinline__ long long snap_off(const StepArgs &a, int s, int k) {
    return (long long)s * a.save_sstride + (long long)(k - a.save_k0) * a.save_kstride;
}
__device__ __forceinline__ float snap_get_at(const StepArgs &a, long long off, int s, int k, unsigned ii, unsigned pi) {
    switch (a.save_kind) {
        case SV_F32_INT:
            return *at(reinterpret_cast<const float *>(a.save) + off, ii);
        case SV_F32_PAD:
            return *at(reinterpret_cast<const float *>(a.save) + off, pi);
        case SV_F16_INT:
            return __half2float(*at(reinterpret_cast<const __half *>(a.save) + off, ii)) *
                   snap_scale(a, s, k);
        case SV_BF16_INT:
            return __bfloat162float(*at(reinterpret_cast<const __hip_bfloat16 *>(a.save) + off, ii)) *
                   snap_scale(a, s, k);
        default:
            return 0.0f;
    }
}
__device__ __forceinline__ void snap_put_at(const StepArgs &a, long long off, int s, int k, unsigned ii, unsigned pi,
                                            float v) {
    float inv = 0.0f;
    if (a.save_kind == SV_F16_INT || a.save_kind == SV_BF16_INT) {
        const float m = snap_scale(a, s, k);
        inv = scale_inv(m);
    }
    switch (a.save_kind) {
        case SV_F32_INT:
            *at(reinterpret_cast<float *>(a.save) + off, ii) = v;
            break;
        case SV_F32_PAD:
            *at(reinterpret_cast<float *>(a.save) + off, pi) = v;
            break;
        case SV_F16_INT:
            *at(reinterpret_cast<__half *>(a.save) + off, ii) = __float2half(v * inv);
            break;
        case SV_BF16_INT:
            *at(reinterpret_cast<__hip_bfloat16 *>(a.save) + off, ii) = __float2bfloat16(v * inv);
            break;
        default:
            break;
    }
}
__device__ __forceinline__ void snap_put(const StepArgs &a, int s, int k, long long ii, long long pi, float v) {
    const long long base = (long long)s * a.save_sstride + (long long)(k - a.save_k0) * a.save_kstride;
    float inv = 0.0f;
    if (a.save_kind == SV_F16_INT || a.save_kind == SV_BF16_INT) {
        const float m = snap_scale(a, s, k);
        inv = scale_inv(m);
    }
    switch (a.save_kind) {
        case SV_F32_INT:
            reinterpret_cast<float *>(a.save)[base + ii] = v;
            break;
        case SV_F32_PAD:
            reinterpret_cast<float *>(a.save)[base + pi] = v;
            break;
        case SV_F16_INT:
            reinterpret_cast<__half *>(a.save)[base + ii] = __float2half(v * inv);
            break;
        case SV_BF16_INT:
            reinterpret_cast<__hip_bfloat16 *>(a.save)[base + ii] = __float2bfloat16(v * inv);
            break;
        default:
            break;
    }
}

// Four consecutive cells (padded index pi, a multiple of 4) of a padded
// snapshot in one access: float32 as a float4, 16-bit kinds as 8 bytes, scaled
// by the step's max as the interior kinds are (the same arithmetic per cell).
__device__ __forceinline__ void snap_put4_pad(const StepArgs &a, long long off, int s, int k, unsigned pi,
                                              const float *v) {
    if (a.save_kind == SV_F32_PAD) {
        vst<4>(at(reinterpret_cast<float *>(a.save) + off, pi), v);
        return;
    }
    const float inv = scale_inv(snap_scale(a, s, k));
    unsigned short h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (a.save_kind == SV_F16_PAD)
            h[j] = __half_as_ushort(__float2half(v[j] * inv));
        else
            h[j] = __bfloat16_as_ushort(__float2bfloat16(v[j] * inv));
    }
    const uint2 w = make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16));
    *reinterpret_cast<uint2 *>(at(reinterpret_cast<unsigned short *>(a.save) + off, pi)) = w;
}
__device__ __forceinline__ void snap_get4_pad(const StepArgs &a, long long off, int s, int k, unsigned pi, float *v) {
    if (a.save_kind == SV_F32_PAD) {
        vld<4>(v, at(reinterpret_cast<const float *>(a.save) + off, pi));
        return;
    }
    const uint2 w = *reinterpret_cast<const uint2 *>(at(reinterpret_cast<const unsigned short *>(a.save) + off, pi));
    const float sc = snap_scale(a, s, k);
    const unsigned short h[4] = {(unsigned short)(w.x & 0xffffu), (unsigned short)(w.x >> 16),
                                 (unsigned short)(w.y & 0xffffu), (unsigned short)(w.y >> 16)};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (a.save_kind == SV_F16_PAD)
            v[j] = __half2float(__ushort_as_half(h[j])) * sc;
        else
            v[j] = __bfloat162float(__ushort_as_bfloat16(h[j])) * sc;
    }
}

// block-wide max of a non-negative float, then one atomicMax on its bits
// (non-negative floats order like their bit patterns); every thread of the
// block must call it
__device__ __forceinline__ void block_max_to(unsigned int *dst, float v) {
    __shared__ float red[8];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmaxf(m, red[i]);
        atomicMax(dst, __float_as_uint(m));
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
            float u, unused = 0.0f;
            if (PHYS == 0) {
                float sum = pxx[j] + pzz[j];
                if (CPML) {
                    if (xs[j]) sum += dpx[j];
                    if (zs) sum += dpz[j];
                    if (xs[j]) sum += zx[j];
                    if (zs) sum += zz[j];
                }
                u = lf_acoustic(c0[j], pv[j], fx[j], fz, v2[j], sum, false, unused);
            } else if (PHYS == 1) {
                u = lf_vti(c0[j], pv[j], fx[j], fz, v2[j], e[j], dl[j], PX[j], PZ[j], PXX[j], PZZ[j], false, unused);
            } else {
                u = lf_tti(c0[j], pv[j], fx[j], fz, v2[j], e[j], dl[j], cs[j], sn[j], PX[j], PZ[j], PXX[j], PZZ[j],
                           pxz[j], false, unused);
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

template <int PHYS, bool CPML, int V, int TZ>
static int launch_step_t(const StepArgs &a0, hipStream_t st, const hipEvent_t *ev = nullptr) {
    StepArgs a = a0;
    a.nxt = (a.nxp - 8 + 64 * V - 1) / (64 * V);
    a.nzt = (a.nzp - 8 + TZ - 1) / TZ;
    const long long waves = (long long)a.nxt * a.nzt * a.nshots;
    long long blocks = (waves + 3) / 4;
    blocks = (blocks + 7) / 8 * 8;
    launch_timed(k_step<PHYS, CPML, V, TZ>, dim3((unsigned)blocks), st, ev, a);
    return (int)hipGetLastError();
}

}  // namespace pam
