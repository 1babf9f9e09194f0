// pam_state.hip -- float16 copies of whole wavefield states (the checkpoint
// slots of the binomial schedule): each field of each shot is stored as
// v / m in float16 with m = its max |v|, and read back as q * m.  The fields
// of a state go through one launch per pass.
#include <hip/hip_fp16.h>

#include <algorithm>

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

static unsigned grid_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

int launch_pack_half(const float *const *x, int nf, long long sstride, long long n, int nshots, void *q,
                     long long qfstride, long long qsstride, unsigned int *bits, long long bfstride, hipStream_t st) {
    if (nf < 1 || nf > kMaxFields) return (int)hipErrorInvalidValue;
    FieldPtrs X{};
    for (int f = 0; f < nf; ++f) X.x[f] = x[f];
    hipLaunchKernelGGL(k_zero_bits, dim3((nshots * nf + 255) / 256), dim3(256), 0, st, bits, bfstride, nshots, nf);
    const unsigned mb = std::min<unsigned>(grid_blocks(n), (unsigned)kMaxBlocksPerShot);
    hipLaunchKernelGGL(k_shot_maxabs, dim3(mb, nshots, nf), dim3(256), 0, st, X, sstride, n, bits, bfstride);
    hipLaunchKernelGGL(k_pack_half, dim3(grid_blocks(n), nshots, nf), dim3(256), 0, st, X, sstride, n,
                       reinterpret_cast<__half *>(q), qfstride, qsstride, (const unsigned int *)bits, bfstride);
    return (int)hipGetLastError();
}

int launch_unpack_half(const void *q, long long qfstride, long long qsstride, long long n, int nshots, int nf,
                       const unsigned int *bits, long long bfstride, float *const *x, long long sstride,
                       hipStream_t st) {
    if (nf < 1 || nf > kMaxFields) return (int)hipErrorInvalidValue;
    FieldPtrsW X{};
    for (int f = 0; f < nf; ++f) X.x[f] = x[f];
    hipLaunchKernelGGL(k_unpack_half, dim3(grid_blocks(n), nshots, nf), dim3(256), 0, st,
                       reinterpret_cast<const __half *>(q), qfstride, qsstride, n, bits, bfstride, X, sstride);
    return (int)hipGetLastError();
}

}  // namespace pam
