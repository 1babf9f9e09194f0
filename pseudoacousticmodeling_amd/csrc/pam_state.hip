// pam_state.hip -- float16 copies of whole wavefield states (the checkpoint
// slots of the binomial schedule): each field of each shot is stored as
// v / m in float16 with m = its max |v|, and read back as q * m.  The fields
// of a state go through one launch per pass.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstdint>

#include "pam_internal.h"

namespace pam {

// Up to kMaxFields fields of one state (the two wavefields and the CPML psi
// and zeta) go through one launch per pass, the field along blockIdx.z.
struct FieldPtrs {
    const float *x[kMaxFields];
};
struct FieldPtrsW {
    float *x[kMaxFields];
};

__global__ void k_zero_bits(unsigned int *bits, long long bfstride, int nshots, int nf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nshots * nf) bits[(long long)(i / nshots) * bfstride + i % nshots] = 0u;
}

// bits[f * bfstride + s] = max(.., max |x_f of shot s|) as float bits
// (non-negative floats order like their bit patterns); grid (blocks, nshots,
// nf).  Each block reduces its share in registers, then across its
// wavefronts, and makes one atomicMax: the blocks of a shot are kept few
// (kMaxBlocksPerShot), since every atomic on one address is serialised (1024
// blocks per shot took 304 us per field on 32 shots of 401 x 1344 cells,
// most of it waiting on the atomics).
constexpr int kMaxBlocksPerShot = 64;
__global__ __launch_bounds__(256) void k_shot_maxabs(const FieldPtrs X, long long sstride, long long n,
                                                     unsigned int *bits, long long bfstride) {
    __shared__ float red[4];
    const int s = blockIdx.y, f = blockIdx.z;
    const float *xs = X.x[f] + (long long)s * sstride;
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(xs[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(bits + (long long)f * bfstride + s, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

__global__ __launch_bounds__(256) void k_pack_half(const FieldPtrs X, long long sstride, long long n, __half *q,
                                                   long long qfstride, long long qsstride, const unsigned int *bits,
                                                   long long bfstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const float m = __uint_as_float(bits[(long long)f * bfstride + s]);
    const float inv = scale_inv(m);
    const float *xs = X.x[f] + (long long)s * sstride;
    __half *qs = q + (long long)f * qfstride + (long long)s * qsstride;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        qs[i] = __float2half(xs[i] * inv);
}

__global__ __launch_bounds__(256) void k_unpack_half(const __half *q, long long qfstride, long long qsstride,
                                                     long long n, const unsigned int *bits, long long bfstride,
                                                     const FieldPtrsW X, long long sstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const float m = __uint_as_float(bits[(long long)f * bfstride + s]);
    const __half *qs = q + (long long)f * qfstride + (long long)s * qsstride;
    float *xs = X.x[f] + (long long)s * sstride;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        xs[i] = __half2float(qs[i]) * m;
}

// The same three passes four cells per thread (a float4 of the field, four
// halfs in one 8-byte access), used when the cell count, the strides and the
// bases are multiples of four; every value goes through the same arithmetic.
__global__ __launch_bounds__(256) void k_shot_maxabs4(const FieldPtrs X, long long sstride, long long n4,
                                                      unsigned int *bits, long long bfstride) {
    __shared__ float red[4];
    const int s = blockIdx.y, f = blockIdx.z;
    const float4 *xs = reinterpret_cast<const float4 *>(X.x[f] + (long long)s * sstride);
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        const float4 v = xs[i];
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(bits + (long long)f * bfstride + s, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

__global__ __launch_bounds__(256) void k_pack_half4(const FieldPtrs X, long long sstride, long long n4, __half *q,
                                                    long long qfstride, long long qsstride, const unsigned int *bits,
                                                    long long bfstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const float m = __uint_as_float(bits[(long long)f * bfstride + s]);
    const float inv = scale_inv(m);
    const float4 *xs = reinterpret_cast<const float4 *>(X.x[f] + (long long)s * sstride);
    uint2 *qs = reinterpret_cast<uint2 *>(q + (long long)f * qfstride + (long long)s * qsstride);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        const float4 v = xs[i];
        const unsigned h0 = __half_as_ushort(__float2half(v.x * inv)), h1 = __half_as_ushort(__float2half(v.y * inv));
        const unsigned h2 = __half_as_ushort(__float2half(v.z * inv)), h3 = __half_as_ushort(__float2half(v.w * inv));
        qs[i] = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
    }
}

__global__ __launch_bounds__(256) void k_unpack_half4(const __half *q, long long qfstride, long long qsstride,
                                                      long long n4, const unsigned int *bits, long long bfstride,
                                                      const FieldPtrsW X, long long sstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const float m = __uint_as_float(bits[(long long)f * bfstride + s]);
    const uint2 *qs = reinterpret_cast<const uint2 *>(q + (long long)f * qfstride + (long long)s * qsstride);
    float4 *xs = reinterpret_cast<float4 *>(X.x[f] + (long long)s * sstride);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        const uint2 w = qs[i];
        float4 v;
        v.x = __half2float(__ushort_as_half((unsigned short)(w.x & 0xffffu))) * m;
        v.y = __half2float(__ushort_as_half((unsigned short)(w.x >> 16))) * m;
        v.z = __half2float(__ushort_as_half((unsigned short)(w.y & 0xffffu))) * m;
        v.w = __half2float(__ushort_as_half((unsigned short)(w.y >> 16))) * m;
        xs[i] = v;
    }
}

// whether every field base, the cell count and the strides allow the
// four-cell kernels (16-byte float and 8-byte half alignment)
static bool four_wide(const void *const *x, int nf, long long sstride, long long n, const void *q, long long qfstride,
                      long long qsstride) {
    if ((n | sstride | qfstride | qsstride) & 3) return false;
    if (reinterpret_cast<uintptr_t>(q) & 7) return false;
    for (int f = 0; f < nf; ++f)
        if (reinterpret_cast<uintptr_t>(x[f]) & 15) return false;
    return true;
}

// The same three passes over each field's region only (FieldRegion): block
// (j, shot, field) takes region rows j, j + gridDim.x, ..., its threads the
// float4 groups of a row; every value goes through the same arithmetic as in
// the whole-field kernels.
struct Regions {
    FieldRegion r[kMaxFields];
};
__device__ __forceinline__ int band_len(int a, int b) { return b > a ? b - a : 0; }
// the row of region row j and the float4 group of region group k
__device__ __forceinline__ int region_row(const FieldRegion &R, int j) {
    const int nr0 = band_len(R.r0[0], R.r1[0]);
    return j < nr0 ? R.r0[0] + j : R.r0[1] + (j - nr0);
}
__device__ __forceinline__ int region_group(const FieldRegion &R, int k) {
    const int ng0 = band_len(R.g0[0], R.g1[0]);
    return k < ng0 ? R.g0[0] + k : R.g0[1] + (k - ng0);
}
__device__ __forceinline__ int region_rows(const FieldRegion &R) {
    return band_len(R.r0[0], R.r1[0]) + band_len(R.r0[1], R.r1[1]);
}
__device__ __forceinline__ int region_groups(const FieldRegion &R) {
    return band_len(R.g0[0], R.g1[0]) + band_len(R.g0[1], R.g1[1]);
}

__global__ __launch_bounds__(256) void k_shot_maxabs_r(const FieldPtrs X, long long sstride, const Regions G,
                                                       unsigned int *bits, long long bfstride) {
    __shared__ float red[4];
    const int s = blockIdx.y, f = blockIdx.z;
    const FieldRegion &R = G.r[f];
    const int nr = region_rows(R), ng = region_groups(R);
    const float4 *xs = reinterpret_cast<const float4 *>(X.x[f] + (long long)s * sstride);
    float m = 0.0f;
    // four of the block's rows per pass, their loads issued together; a row
    // past the region's end reads row j again (the max is unchanged by it)
    const int G4 = gridDim.x;
    for (int j = blockIdx.x; j < nr; j += 4 * G4) {
        const float4 *xr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xr[u] = xs + (long long)region_row(R, j + u * G4 < nr ? j + u * G4 : j) * R.p4;
        for (int k = threadIdx.x; k < ng; k += blockDim.x) {
            const int g = region_group(R, k);
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = xr[u][g];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                m = fmaxf(m, fmaxf(fmaxf(fabsf(v[u].x), fabsf(v[u].y)), fmaxf(fabsf(v[u].z), fabsf(v[u].w))));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(bits + (long long)f * bfstride + s, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

__global__ __launch_bounds__(256) void k_pack_half_r(const FieldPtrs X, long long sstride, const Regions G, __half *q,
                                                     long long qfstride, long long qsstride, const unsigned int *bits,
                                                     long long bfstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const FieldRegion &R = G.r[f];
    const int nr = region_rows(R), ng = region_groups(R);
    const float inv = scale_inv(__uint_as_float(bits[(long long)f * bfstride + s]));
    const float4 *xs = reinterpret_cast<const float4 *>(X.x[f] + (long long)s * sstride);
    uint2 *qs = reinterpret_cast<uint2 *>(q + (long long)f * qfstride + (long long)s * qsstride);
    for (int j = blockIdx.x; j < nr; j += gridDim.x) {
        const long long rb = (long long)region_row(R, j) * R.p4;
        for (int k = threadIdx.x; k < ng; k += blockDim.x) {
            const long long i = rb + region_group(R, k);
            const float4 v = xs[i];
            const unsigned h0 = __half_as_ushort(__float2half(v.x * inv)), h1 = __half_as_ushort(__float2half(v.y * inv));
            const unsigned h2 = __half_as_ushort(__float2half(v.z * inv)), h3 = __half_as_ushort(__float2half(v.w * inv));
            qs[i] = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
        }
    }
}

__global__ __launch_bounds__(256) void k_unpack_half_r(const __half *q, long long qfstride, long long qsstride,
                                                       const Regions G, const unsigned int *bits, long long bfstride,
                                                       const FieldPtrsW X, long long sstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const FieldRegion &R = G.r[f];
    const int nr = region_rows(R), ng = region_groups(R);
    const float m = __uint_as_float(bits[(long long)f * bfstride + s]);
    const uint2 *qs = reinterpret_cast<const uint2 *>(q + (long long)f * qfstride + (long long)s * qsstride);
    float4 *xs = reinterpret_cast<float4 *>(X.x[f] + (long long)s * sstride);
    for (int j = blockIdx.x; j < nr; j += gridDim.x) {
        const long long rb = (long long)region_row(R, j) * R.p4;
        for (int k = threadIdx.x; k < ng; k += blockDim.x) {
            const long long i = rb + region_group(R, k);
            const uint2 w = qs[i];
            float4 v;
            v.x = __half2float(__ushort_as_half((unsigned short)(w.x & 0xffffu))) * m;
            v.y = __half2float(__ushort_as_half((unsigned short)(w.x >> 16))) * m;
            v.z = __half2float(__ushort_as_half((unsigned short)(w.y & 0xffffu))) * m;
            v.w = __half2float(__ushort_as_half((unsigned short)(w.y >> 16))) * m;
            xs[i] = v;
        }
    }
}

// y = x on the region of each field (float4 groups), shots along y, fields along z
struct CopyPtrs {
    const float *x[kMaxFields];
    float *y[kMaxFields];
};
__global__ __launch_bounds__(256) void k_copy_r(const Regions G, const CopyPtrs C, long long sstride) {
    const int s = blockIdx.y, f = blockIdx.z;
    const FieldRegion &R = G.r[f];
    const int nr = region_rows(R), ng = region_groups(R);
    const float4 *xs = reinterpret_cast<const float4 *>(C.x[f] + (long long)s * sstride);
    float4 *ys = reinterpret_cast<float4 *>(C.y[f] + (long long)s * sstride);
    for (int j = blockIdx.x; j < nr; j += gridDim.x) {
        const long long rb = (long long)region_row(R, j) * R.p4;
        for (int k = threadIdx.x; k < ng; k += blockDim.x) ys[rb + region_group(R, k)] = xs[rb + region_group(R, k)];
    }
}

static unsigned grid_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

FieldRegion region_all(int nzp, int P) {
    FieldRegion r{};
    r.r0[0] = 0;
    r.r1[0] = nzp;
    r.g0[0] = 0;
    r.g1[0] = P / 4;
    r.p4 = P / 4;
    return r;
}
FieldRegion region_xstrips(int nzp, int nxp, int P, int N) {
    FieldRegion r{};
    r.p4 = P / 4;
    r.r0[0] = 4;
    r.r1[0] = nzp - 4;
    // groups holding columns [4, N) and [nxp-N, nxp-4), one band when they meet
    const int a0 = 1, a1 = (N + 3) / 4, b0 = (nxp - N) / 4, b1 = (nxp - 4 + 3) / 4;
    if (b0 <= a1) {
        r.g0[0] = a0;
        r.g1[0] = b1;
    } else {
        r.g0[0] = a0;
        r.g1[0] = a1;
        r.g0[1] = b0;
        r.g1[1] = b1;
    }
    return r;
}
FieldRegion region_zstrips(int nzp, int P, int N) {
    FieldRegion r{};
    r.p4 = P / 4;
    r.g0[0] = 0;
    r.g1[0] = P / 4;
    if (nzp - N <= N) {
        r.r0[0] = 4;
        r.r1[0] = nzp - 4;
    } else {
        r.r0[0] = 4;
        r.r1[0] = N;
        r.r0[1] = nzp - N;
        r.r1[1] = nzp - 4;
    }
    return r;
}
static int region_max_rows(const FieldRegion *reg, int nf) {
    int m = 1;
    for (int f = 0; f < nf; ++f) {
        const FieldRegion &R = reg[f];
        const int nr = std::max(R.r1[0] - R.r0[0], 0) + std::max(R.r1[1] - R.r0[1], 0);
        m = std::max(m, nr);
    }
    return m;
}

int launch_pack_half(const float *const *x, int nf, long long sstride, long long n, int nshots, void *q,
                     long long qfstride, long long qsstride, unsigned int *bits, long long bfstride, hipStream_t st,
                     const FieldRegion *reg) {
    if (nf < 1 || nf > kMaxFields) return (int)hipErrorInvalidValue;
    FieldPtrs X{};
    for (int f = 0; f < nf; ++f) X.x[f] = x[f];
    hipLaunchKernelGGL(k_zero_bits, dim3((nshots * nf + 255) / 256), dim3(256), 0, st, bits, bfstride, nshots, nf);
    if (reg && four_wide(reinterpret_cast<const void *const *>(x), nf, sstride, n, q, qfstride, qsstride)) {
        Regions G{};
        for (int f = 0; f < nf; ++f) G.r[f] = reg[f];
        const unsigned rows = (unsigned)region_max_rows(reg, nf);
        hipLaunchKernelGGL(k_shot_maxabs_r, dim3(std::min<unsigned>(rows, (unsigned)kMaxBlocksPerShot), nshots, nf),
                           dim3(256), 0, st, X, sstride, G, bits, bfstride);
        hipLaunchKernelGGL(k_pack_half_r, dim3(rows, nshots, nf), dim3(256), 0, st, X, sstride, G,
                           reinterpret_cast<__half *>(q), qfstride, qsstride, (const unsigned int *)bits, bfstride);
        return (int)hipGetLastError();
    }
    if (four_wide(reinterpret_cast<const void *const *>(x), nf, sstride, n, q, qfstride, qsstride)) {
        const long long n4 = n / 4;
        const unsigned mb4 = std::min<unsigned>(grid_blocks(n4), (unsigned)kMaxBlocksPerShot);
        hipLaunchKernelGGL(k_shot_maxabs4, dim3(mb4, nshots, nf), dim3(256), 0, st, X, sstride, n4, bits, bfstride);
        hipLaunchKernelGGL(k_pack_half4, dim3(grid_blocks(n4), nshots, nf), dim3(256), 0, st, X, sstride, n4,
                           reinterpret_cast<__half *>(q), qfstride, qsstride, (const unsigned int *)bits, bfstride);
        return (int)hipGetLastError();
    }
    const unsigned mb = std::min<unsigned>(grid_blocks(n), (unsigned)kMaxBlocksPerShot);
    hipLaunchKernelGGL(k_shot_maxabs, dim3(mb, nshots, nf), dim3(256), 0, st, X, sstride, n, bits, bfstride);
    hipLaunchKernelGGL(k_pack_half, dim3(grid_blocks(n), nshots, nf), dim3(256), 0, st, X, sstride, n,
                       reinterpret_cast<__half *>(q), qfstride, qsstride, (const unsigned int *)bits, bfstride);
    return (int)hipGetLastError();
}

int launch_copy_region(const float *const *x, float *const *y, int nf, long long sstride, int nshots,
                       const FieldRegion *reg, hipStream_t st) {
    if (nf < 1 || nf > kMaxFields) return (int)hipErrorInvalidValue;
    bool four = sstride % 4 == 0;
    for (int f = 0; f < nf; ++f)
        four = four && reinterpret_cast<uintptr_t>(x[f]) % 16 == 0 && reinterpret_cast<uintptr_t>(y[f]) % 16 == 0;
    if (!four) {  // whole fields
        for (int f = 0; f < nf; ++f) {
            const hipError_t e = hipMemcpyAsync(y[f], x[f], (size_t)sstride * nshots * sizeof(float),
                                                hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return (int)e;
        }
        return 0;
    }
    Regions G{};
    CopyPtrs C{};
    for (int f = 0; f < nf; ++f) {
        G.r[f] = reg[f];
        C.x[f] = x[f];
        C.y[f] = y[f];
    }
    hipLaunchKernelGGL(k_copy_r, dim3((unsigned)region_max_rows(reg, nf), nshots, nf), dim3(256), 0, st, G, C,
                       sstride);
    return (int)hipGetLastError();
}

int launch_unpack_half(const void *q, long long qfstride, long long qsstride, long long n, int nshots, int nf,
                       const unsigned int *bits, long long bfstride, float *const *x, long long sstride,
                       hipStream_t st, const FieldRegion *reg) {
    if (nf < 1 || nf > kMaxFields) return (int)hipErrorInvalidValue;
    FieldPtrsW X{};
    for (int f = 0; f < nf; ++f) X.x[f] = x[f];
    if (reg && four_wide(reinterpret_cast<const void *const *>(x), nf, sstride, n, q, qfstride, qsstride)) {
        Regions G{};
        for (int f = 0; f < nf; ++f) G.r[f] = reg[f];
        hipLaunchKernelGGL(k_unpack_half_r, dim3((unsigned)region_max_rows(reg, nf), nshots, nf), dim3(256), 0, st,
                           reinterpret_cast<const __half *>(q), qfstride, qsstride, G, bits, bfstride, X, sstride);
        return (int)hipGetLastError();
    }
    if (four_wide(reinterpret_cast<const void *const *>(x), nf, sstride, n, q, qfstride, qsstride)) {
        const long long n4 = n / 4;
        hipLaunchKernelGGL(k_unpack_half4, dim3(grid_blocks(n4), nshots, nf), dim3(256), 0, st,
                           reinterpret_cast<const __half *>(q), qfstride, qsstride, n4, bits, bfstride, X, sstride);
        return (int)hipGetLastError();
    }
    hipLaunchKernelGGL(k_unpack_half, dim3(grid_blocks(n), nshots, nf), dim3(256), 0, st,
                       reinterpret_cast<const __half *>(q), qfstride, qsstride, n, bits, bfstride, X, sstride);
    return (int)hipGetLastError();
}

}  // namespace pam
