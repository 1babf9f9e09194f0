// pam_prep.hip -- preprocessing of the migration / inversion drivers on the
// GPU: model smoothing (the Python driver's 2-D Gaussian over the inner
// region with edge replication and the water column restored,
// src/utils.py:76-129; the C++ driver's separable clamped Gaussian with the
// masked cells restored, cpp/Migration.cu:171-236) and the linear top mute of
// shot gathers (src/utils.py:30-49, cpp/Migration.cu:86-115).  The
// arithmetic of each follows its reference: float64 accumulation for the
// Python rules, float32 in tap order without contraction for the C++ rules.
#include "pam_internal.h"

namespace pam {

// s = f (a parameter) or 1 / f (velocity -> slowness)
__global__ void k_slowness(const float *f, float *s, long long n, int parameter) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s[i] = parameter ? f[i] : 1.0f / f[i];
}

// first row of each column that is not water (nz when the whole column is)
__global__ void k_first_dry(const unsigned char *mask, int nz, int nx, int *first) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    int z = 0;
    if (mask)
        while (z < nz && mask[(long long)z * nx + x]) ++z;
    first[x] = mask ? z : 0;
}

// 2-D Gaussian (ks x ks weights w, float64 accumulation, divided by the
// float64 sum of the weights) on the cells a whole kernel fits
__global__ void k_smooth_conv2d(const float *s, int nz, int nx, const float *w, int ks, double total, float *conv) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int h = ks / 2;
    if (x < h || x >= nx - h || z < h || z >= nz - h) return;
    double acc = 0.0;
    for (int i = 0; i < ks; ++i) {
        const float *row = s + (long long)(z - h + i) * nx + (x - h);
        const float *wr = w + (long long)i * ks;
        for (int j = 0; j < ks; ++j) acc += (double)row[j] * (double)wr[j];
    }
    conv[(long long)z * nx + x] = (float)(acc / total);
}

// out = the smoothed value of the clamped inner cell, the original above the
// first dry cell of the column; back to velocity unless a parameter
__global__ void k_smooth_finish_py(const float *s_old, const float *conv, const int *first, int nz, int nx, int h,
                                   int parameter, float *out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x >= nx) return;
    const int zc = min(max(z, h), nz - 1 - h), xc = min(max(x, h), nx - 1 - h);
    const long long i = (long long)z * nx + x;
    const float v = (z < first[x]) ? s_old[i] : conv[(long long)zc * nx + xc];
    out[i] = parameter ? v : 1.0f / v;
}

// one separable pass: along x (dir 0) or z (dir 1), clamped indices, float32
// products and sums in tap order
__global__ void k_smooth_sep(const float *in, int nz, int nx, const float *w, int ks, int dir, float *out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x >= nx) return;
    const int h = ks / 2;
    float acc = 0.0f;
    for (int i = -h; i <= h; ++i) {
        const int xx = dir == 0 ? min(max(x + i, 0), nx - 1) : x;
        const int zz = dir == 1 ? min(max(z + i, 0), nz - 1) : z;
        acc = __fadd_rn(acc, __fmul_rn(w[i + h], in[(long long)zz * nx + xx]));
    }
    out[(long long)z * nx + x] = acc;
}

__global__ void k_smooth_finish_cpp(const float *s_old, const float *sm, const unsigned char *mask, long long n,
                                    int parameter, float *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = (mask && mask[i]) ? s_old[i] : sm[i];
    out[i] = parameter ? v : 1.0f / v;
}

// Top mute of gathers [nshots][nt][nrec] in place; t1 / t2 per (shot,
// receiver).  Python rule (float64): w = 0 before t1, (t - t1)/(t2 - t1) on
// the ramp, 1 after, applied as float64(w) * seis.  C++ rule (float32): zero
// before t1, seis * ((t - t1)/(t2 - t1)) on the ramp, unchanged after.
__global__ void k_mute_py(float *seis, int nt, int nrec, const double *t12, double dt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y, s = blockIdx.z;
    if (r >= nrec) return;
    const double t1 = t12[2 * ((long long)s * nrec + r)], t2 = t12[2 * ((long long)s * nrec + r) + 1];
    const double t = (double)it * dt;
    const double w = (t >= t2) ? 1.0 : ((t >= t1) ? (t - t1) / (t2 - t1) : 0.0);
    float *p = seis + ((long long)s * nt + it) * nrec + r;
    *p = (float)(w * (double)*p);
}

__global__ void k_mute_cpp(float *seis, int nt, int nrec, const float *t12, float dt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y, s = blockIdx.z;
    if (r >= nrec) return;
    const float t1 = t12[2 * ((long long)s * nrec + r)], t2 = t12[2 * ((long long)s * nrec + r) + 1];
    const float t = __fmul_rn((float)it, dt);
    float *p = seis + ((long long)s * nt + it) * nrec + r;
    if (t < t1)
        *p = 0.0f;
    else if (t < t2)
        *p = __fmul_rn(*p, __fdiv_rn(__fsub_rn(t, t1), __fsub_rn(t2, t1)));
}

int launch_slowness(const float *f, float *s, long long n, int parameter, hipStream_t st) {
    hipLaunchKernelGGL(k_slowness, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, s, n, parameter);
    return (int)hipGetLastError();
}
int launch_first_dry(const unsigned char *mask, int nz, int nx, int *first, hipStream_t st) {
    hipLaunchKernelGGL(k_first_dry, dim3((nx + 255) / 256), dim3(256), 0, st, mask, nz, nx, first);
    return (int)hipGetLastError();
}
int launch_smooth_conv2d(const float *s, int nz, int nx, const float *w, int ks, double total, float *conv,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_smooth_conv2d, dim3((nx + 255) / 256, nz), dim3(256), 0, st, s, nz, nx, w, ks, total, conv);
    return (int)hipGetLastError();
}
int launch_smooth_finish_py(const float *s_old, const float *conv, const int *first, int nz, int nx, int h,
                            int parameter, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_smooth_finish_py, dim3((nx + 255) / 256, nz), dim3(256), 0, st, s_old, conv, first, nz, nx,
                       h, parameter, out);
    return (int)hipGetLastError();
}
int launch_smooth_sep(const float *in, int nz, int nx, const float *w, int ks, int dir, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_smooth_sep, dim3((nx + 255) / 256, nz), dim3(256), 0, st, in, nz, nx, w, ks, dir, out);
    return (int)hipGetLastError();
}
int launch_smooth_finish_cpp(const float *s_old, const float *sm, const unsigned char *mask, long long n,
                             int parameter, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_smooth_finish_cpp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s_old, sm, mask, n,
                       parameter, out);
    return (int)hipGetLastError();
}
int launch_mute(float *seis, int nshots, int nt, int nrec, const void *t12, double dt, int cpp, hipStream_t st) {
    const dim3 g((nrec + 255) / 256, nt, nshots);
    if (cpp)
        hipLaunchKernelGGL(k_mute_cpp, g, dim3(256), 0, st, seis, nt, nrec, (const float *)t12, (float)dt);
    else
        hipLaunchKernelGGL(k_mute_py, g, dim3(256), 0, st, seis, nt, nrec, (const double *)t12, dt);
    return (int)hipGetLastError();
}

}  // namespace pam
