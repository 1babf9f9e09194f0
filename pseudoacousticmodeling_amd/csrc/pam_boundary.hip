// pam_boundary.hip -- kernels of the boundary-saving source-field strategy.
//
// The forward pass keeps, for every step k, the 4-cell rims just inside the
// interior edge of the source field S_k (the cells an 8th-order stencil of
// the inner region reads but cannot recompute); src/Migration.py:58-67 and
// :288-298 keep the same four strips.  The backward pass rebuilds S_{k-1}
// from S_k and S_{k+1} with the undamped step run backwards in time,
//     S_{k-1} = v^2 dt^2 L(S_k) + 2 S_k - S_{k+1} (+ wav[k+1] at the source),
// which is exact on the inner region because the absorbing layer never acts
// there; the saved rims are then written back, and the 4-cell band just
// outside the interior is cleared so the reconstructed field is zero on the
// whole absorbing layer (the reverse step then keeps it zero there).
//
// Strip layout of one step of one shot (sb_width elements):
//   [top 4][nx] [bottom 4][nx] [left nz-8][4] [right nz-8][4]
// rows / columns in interior coordinates: top rows 0..3, bottom rows
// nz-4..nz-1, left / right columns 0..3 / nx-4..nx-1 of rows 4..nz-5.
// Element kinds: float32, or float16 / bfloat16 of v / m with m the step's
// max |v| over its strips (scale[shot][k]).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "pam_internal.h"

namespace pam {

__device__ __forceinline__ long long sb_w(int nz, int nx) { return 8LL * nx + 8LL * (nz - 8); }
long long sb_width(int nz, int nx) { return 8LL * nx + 8LL * (nz - 8); }

// interior (iz, ix) of strip element i (32-bit arithmetic: a step's strips
// hold 8 (nx + nz - 8) elements, far below 2^31 for any grid the plane guard
// admits)
__device__ __forceinline__ void sb_cell(int i, int nz, int nx, int *iz, int *ix) {
    const int tnx = 4 * nx;
    if (i < tnx) {
        *iz = i / nx;
        *ix = i - *iz * nx;
    } else if (i < 2 * tnx) {
        const int j = i - tnx;
        const int r = j / nx;
        *iz = nz - 4 + r;
        *ix = j - r * nx;
    } else {
        int j = i - 2 * tnx;
        const int side = 4 * (nz - 8);
        const bool right = j >= side;
        if (right) j -= side;
        *iz = 4 + (j >> 2);
        *ix = (right ? nx - 4 : 0) + (j & 3);
    }
}

__device__ __forceinline__ float sb_load(const void *p, int kind, long long i, float m) {
    if (kind == 1) return __half2float(reinterpret_cast<const __half *>(p)[i]) * m;
    if (kind == 2) return __bfloat162float(reinterpret_cast<const __hip_bfloat16 *>(p)[i]) * m;
    return reinterpret_cast<const float *>(p)[i];
}

// float32 strips of the padded field U (interior origin at (N, N)) into step
// k of each shot's strip buffer, one element per thread (grid: blocks, shots)
__global__ __launch_bounds__(256) void k_sb_save_f32(const float *U, long long sstride, int P, int N, int nz, int nx,
                                                     float *strips, long long strip_sstride, int k) {
    const int s = blockIdx.y;
    const int W = (int)sb_w(nz, nx);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W) return;
    const float *u = U + (long long)s * sstride + (long long)N * P + N;
    int iz, ix;
    sb_cell(i, nz, nx, &iz, &ix);
    strips[(long long)s * strip_sstride + (long long)k * W + i] = u[(long long)iz * P + ix];
}

// one block per shot: strips of the padded field U (interior origin at
// (N, N)) into step k of the shot's strip buffer; for the 16-bit kinds the
// block first finds the step's max |v| and stores it as the scale
__global__ __launch_bounds__(1024) void k_sb_save(const float *U, long long sstride, int P, int N, int nz, int nx,
                                                  int kind, void *strips, long long strip_sstride, int k,
                                                  float *scale, int nt) {
    __shared__ float red[1024];
    const int s = blockIdx.x;
    const long long W = sb_w(nz, nx);
    const float *u = U + (long long)s * sstride + (long long)N * P + N;
    const long long base = (long long)s * strip_sstride + (long long)k * W;
    float inv = 1.0f;
    if (kind != 0) {
        float m = 0.0f;
        for (long long i = threadIdx.x; i < W; i += blockDim.x) {
            int iz, ix;
            sb_cell((int)i, nz, nx, &iz, &ix);
            m = fmaxf(m, fabsf(u[(long long)iz * P + ix]));
        }
        red[threadIdx.x] = m;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
            __syncthreads();
        }
        m = red[0];
        if (threadIdx.x == 0) scale[(long long)s * nt + k] = m;
        inv = scale_inv(m);
    }
    for (long long i = threadIdx.x; i < W; i += blockDim.x) {
        int iz, ix;
        sb_cell((int)i, nz, nx, &iz, &ix);
        const float v = u[(long long)iz * P + ix];
        if (kind == 1)
            reinterpret_cast<__half *>(strips)[base + i] = __float2half(v * inv);
        else if (kind == 2)
            reinterpret_cast<__hip_bfloat16 *>(strips)[base + i] = __float2bfloat16(v * inv);
        else
            reinterpret_cast<float *>(strips)[base + i] = v;
    }
}

// Write the saved strips of step L into the padded field U, clear the 4-cell
// band just outside the interior, and add wav[wk] at a source cell that lies
// in the inner region (wk >= 0).  When D is given (D_{L+1} = S_L - 2 S_{L+1} +
// S_{L+2} on the interior, formed by the reverse step from a rebuilt S_L),
// its strip cells are re-formed from the saved strips of steps L, L+1, L+2
// and the source cell gets the same wav[wk].
__global__ __launch_bounds__(256) void k_sb_restore(float *U, long long sstride, int P, int N, int nz, int nx,
                                                    int kind, const void *strips, long long strip_sstride,
                                                    const float *scale, int nt, int L, float *D, long long dsstride,
                                                    const int *ssz, const int *ssx, const float *wav, int wk) {
    const int s = blockIdx.y;
    const long long W = sb_w(nz, nx);
    const long long wb = 4LL * (nx + 8);  // top (or bottom) band cells
    const long long nband = 2 * wb + 8LL * nz;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float *u = U + (long long)s * sstride + (long long)N * P + N;  // interior origin
    if (i < W) {
        int iz, ix;
        sb_cell((int)i, nz, nx, &iz, &ix);
        const long long base = (long long)s * strip_sstride;
        const float *sc = scale + (long long)s * nt;
        const float v0 = sb_load(strips, kind, base + (long long)L * W + i, kind ? sc[L] : 1.0f);
        u[(long long)iz * P + ix] = v0;
        if (D) {
            const float v1 = sb_load(strips, kind, base + (long long)(L + 1) * W + i, kind ? sc[L + 1] : 1.0f);
            const float v2 = sb_load(strips, kind, base + (long long)(L + 2) * W + i, kind ? sc[L + 2] : 1.0f);
            D[(long long)s * dsstride + (long long)iz * nx + ix] = v0 - 2.0f * v1 + v2;
        }
    } else if (i < W + nband) {
        int j = (int)(i - W);  // 32-bit: the band holds 8 (nx + 8) + 8 nz cells
        int iz, ix;
        if (j < 2 * (int)wb) {
            const bool bot = j >= (int)wb;
            if (bot) j -= (int)wb;
            const int r = j / (nx + 8);
            iz = (bot ? nz : -4) + r;
            ix = -4 + (j - r * (nx + 8));
        } else {
            j -= 2 * (int)wb;
            const bool right = j >= 4 * nz;
            if (right) j -= 4 * nz;
            iz = j >> 2;
            ix = (right ? nx : -4) + (j & 3);
        }
        u[(long long)iz * P + ix] = 0.0f;
    } else if (i == W + nband && wk >= 0 && wk < nt) {
        const int iz = ssz[s] - N, ix = ssx[s] - N;
        if (iz >= 4 && iz < nz - 4 && ix >= 4 && ix < nx - 4) {
            u[(long long)iz * P + ix] += wav[wk];
            if (D) D[(long long)s * dsstride + (long long)iz * nx + ix] += wav[wk];
        }
    }
}

// Random-boundary rebuild: wav[wk] added at each shot's source cell of the
// padded field U and, when D (interior [nz][nx] per shot) is given and the
// cell is interior, to D as well (wk in [0, nt))
__global__ void k_rbc_source(float *U, long long sstride, int P, float *D, long long dsstride, int N, int nz, int nx,
                             const int *ssz, const int *ssx, const float *wav, int wk, int nshots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const int z = ssz[s], x = ssx[s];
    U[(long long)s * sstride + (long long)z * P + x] += wav[wk];
    const int iz = z - N, ix = x - N;
    if (D && iz >= 0 && iz < nz && ix >= 0 && ix < nx) D[(long long)s * dsstride + (long long)iz * nx + ix] += wav[wk];
}

int launch_rbc_source(float *U, long long sstride, int P, float *D, long long dsstride, int N, int nz, int nx,
                      const int *ssz, const int *ssx, const float *wav, int wk, int nshots, hipStream_t st) {
    hipLaunchKernelGGL(k_rbc_source, dim3((nshots + 63) / 64), dim3(64), 0, st, U, sstride, P, D, dsstride, N, nz, nx,
                       ssz, ssx, wav, wk, nshots);
    return (int)hipGetLastError();
}

// dst = src on the interior, 0 elsewhere (padded [nzp][P] per shot)
__global__ void k_copy_interior(const float *src, float *dst, long long sstride, int P, int N, int nz, int nx) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    const int s = blockIdx.z;
    if (ix >= P) return;
    const long long o = (long long)s * sstride + (long long)iz * P + ix;
    const bool in = iz >= N && iz < N + nz && ix >= N && ix < N + nx;
    dst[o] = in ? src[o] : 0.0f;
}

int launch_sb_save(const float *U, long long sstride, int P, int N, int nz, int nx, int nshots, int kind,
                   void *strips, long long strip_sstride, int k, float *scale, int nt, hipStream_t st) {
    if (kind == 0) {  // no scale to find first: the strips of every shot in parallel
        const long long W = sb_width(nz, nx);
        hipLaunchKernelGGL(k_sb_save_f32, dim3((unsigned)((W + 255) / 256), nshots), dim3(256), 0, st, U, sstride, P,
                           N, nz, nx, reinterpret_cast<float *>(strips), strip_sstride, k);
        return (int)hipGetLastError();
    }
    hipLaunchKernelGGL(k_sb_save, dim3(nshots), dim3(1024), 0, st, U, sstride, P, N, nz, nx, kind, strips,
                       strip_sstride, k, scale, nt);
    return (int)hipGetLastError();
}

int launch_sb_restore(float *U, long long sstride, int P, int N, int nz, int nx, int nshots, int kind,
                      const void *strips, long long strip_sstride, const float *scale, int nt, int L, float *D,
                      long long dsstride, const int *ssz, const int *ssx, const float *wav, int wk, hipStream_t st) {
    const long long n = sb_width(nz, nx) + 8LL * (nx + 8) + 8LL * nz + 1;
    hipLaunchKernelGGL(k_sb_restore, dim3((unsigned)((n + 255) / 256), nshots), dim3(256), 0, st, U, sstride, P, N,
                       nz, nx, kind, strips, strip_sstride, scale, nt, L, D, dsstride, ssz, ssx, wav, wk);
    return (int)hipGetLastError();
}

int launch_copy_interior(const float *src, float *dst, long long sstride, int nzp, int P, int N, int nz, int nx,
                         int nshots, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_interior, dim3((P + 255) / 256, nzp, nshots), dim3(256), 0, st, src, dst, sstride, P,
                       N, nz, nx);
    return (int)hipGetLastError();
}

}  // namespace pam
