// pam_exact.hip -- exact discrete adjoint of the acoustic forward map, the
// backward of pam_forward (pam_forward_adjoint).
//
// With C_k the current field of step k after the injection of src[k], the
// forward map run by the step kernels is, on the live cells (4 <= i < n-4):
//
//   Cerjan  C_{k+1} = D (V Q_k + 2 C_k - D C_{k-1}) + e src[k+1],  seis_k = (D C_k)[rec]
//   CPML    psi_k  = Sx (a psi_{k-1} + b Dx C_k)            (x; z alike)
//           zeta_k = Sx (a zeta_{k-1} + b (Dxx C_k + Dx psi_k))
//           C_{k+1} = V Q_k + 2 C_k - C_{k-1} + e src[k+1],  seis_k = C_k[rec]
//
// with Q_k = Dxx C_k + Dzz C_k (+ Sx (Dx psi_x + zeta_x) + Sz (Dz psi_z + zeta_z)
// for CPML), V = vp^2 dt^2, D the Cerjan factors, a, b the CPML profiles
// (functions of vp, src/CudaKernels.py:33-90).  The reverse sweep, for
// k = nt-1 .. 0, with lambda_k = dJ/dC_k, mu = dJ/dpsi, nu = dJ/dzeta:
//
//   Cerjan  phi = D lambda_{k+1}
//           lambda_k = Dxx(V phi) + Dzz(V phi) + 2 phi - D^2 lambda_{k+2} + D R^T r_k
//   CPML    phi = lambda_{k+1},  q = V phi
//           nu_k  = Sx (q + a nu_{k+1})
//           mu_k  = Sx (a mu_{k+1} + Dx^T (Sx q + b nu_k))
//           lambda_k = Dxx(q + b nu_k) + Dzz(..) + 2 phi - lambda_{k+2}
//                      + Dx^T (b mu_k) + Dz^T (..) + R^T r_k
//
// and dJ/dV += phi Q_k, dJ/da += mu_k psi_{k-1} + nu_k zeta_{k-1},
// dJ/db += mu_k Dx C_k + nu_k (Dxx C_k + Dx psi_k).  Dxx is symmetric and Dx
// antisymmetric, so Dxx^T = Dxx and Dx^T y = -Dx y on inputs that vanish off
// the live cells.  The saved forward state of step k (C_k, psi_k, zeta_k) is
// read back from the snapshot buffers.
//
// Shots run along the last grid dimension.  The Cerjan step and the mu step
// take one thread per padded cell (blocks of 64 x 4 cells; the mu step only
// over the tiles that hold strip cells), their neighbours coming through the
// L1 / L2 caches; the CPML lambda step stages 64 x 16 tiles with their halo
// in LDS (k_exact_lambda_cpml).
#include <hip/hip_runtime.h>

#include "pam_internal.h"

namespace pam {

namespace {

constexpr float C0 = -1435.0f / 504.0f, C1 = 8.0f / 5.0f, C2 = -1.0f / 5.0f, C3 = 8.0f / 315.0f,
                C4 = -1.0f / 560.0f;
constexpr float A1 = 4.0f / 5.0f, A2 = -1.0f / 5.0f, A3 = 4.0f / 105.0f, A4 = -1.0f / 280.0f;

__device__ __forceinline__ float kc(int m) { return m == 1 ? C1 : m == 2 ? C2 : m == 3 ? C3 : C4; }
__device__ __forceinline__ float ka(int m) { return m == 1 ? A1 : m == 2 ? A2 : m == 3 ? A3 : A4; }

struct Cell {
    int iz, ix, s;
    long long c;  // index in a padded field of shot 0
    bool live;
};

__device__ __forceinline__ bool live_at(const ExactArgs &a, int iz, int ix) {
    return ix >= 4 && ix < a.nxp - 4 && iz >= 4 && iz < a.nzp - 4;
}
__device__ __forceinline__ bool xstrip(const ExactArgs &a, int iz, int ix) {
    return live_at(a, iz, ix) && (ix < a.N || ix >= a.nxp - a.N);
}
__device__ __forceinline__ bool zstrip(const ExactArgs &a, int iz, int ix) {
    return live_at(a, iz, ix) && (iz < a.N || iz >= a.nzp - a.N);
}

// d2 / d1 stencils of a saved field f (shot-offset pointer) at padded index c,
// stride st (1 along x, P along z), undivided
__device__ __forceinline__ float d2s(const float *f, long long c, long long st) {
    return C0 * f[c] + C1 * (f[c + st] + f[c - st]) + C2 * (f[c + 2 * st] + f[c - 2 * st]) +
           C3 * (f[c + 3 * st] + f[c - 3 * st]) + C4 * (f[c + 4 * st] + f[c - 4 * st]);
}
__device__ __forceinline__ float d1s(const float *f, long long c, long long st) {
    return A1 * (f[c + st] - f[c - st]) + A2 * (f[c + 2 * st] - f[c - 2 * st]) +
           A3 * (f[c + 3 * st] - f[c - 3 * st]) + A4 * (f[c + 4 * st] - f[c - 4 * st]);
}

// CPML: q + b nu_k at a neighbour (zero off the live cells), for the x or z strip
__device__ __forceinline__ float u_of(const ExactArgs &a, const float *X, const float *nu_in, const float *ca,
                                      const float *cb, int iz, int ix, long long c, bool strip) {
    const float q = a.v2[c] * X[c];  // X vanishes off the live cells
    if (!strip) return q;
    return q + cb[c] * (q + ca[c] * nu_in[c]);
}

// The 64 x 4 tiles that hold strip cells: every tile of the tile rows [0, a0)
// and [b0, nTR) (the z strips), then the tile columns [0, c0) and [d0, nTC)
// (the x strips) of the rows between; blockIdx.x runs over them in that order.
struct StripTiles {
    int nTR, nTC, a0, b0, c0, d0, rowsA, colsB;
};
__host__ StripTiles strip_tiles(const ExactArgs &a) {
    StripTiles m;
    m.nTR = (a.nzp + 3) / 4;
    m.nTC = (a.nxp + 63) / 64;
    m.a0 = (a.N + 3) / 4;          // rows < N
    m.b0 = (a.nzp - a.N) / 4;      // rows >= nzp - N
    if (m.b0 <= m.a0) m.a0 = m.b0 = m.nTR;
    m.c0 = (a.N + 63) / 64;        // columns < N
    m.d0 = (a.nxp - a.N) / 64;     // columns >= nxp - N
    if (m.d0 <= m.c0) m.c0 = m.d0 = m.nTC;
    m.rowsA = m.a0 + (m.nTR - m.b0);
    m.colsB = m.c0 + (m.nTC - m.d0);
    return m;
}

__global__ __launch_bounds__(256) void k_exact_mu(const ExactArgs a, const StripTiles t) {
    int k = blockIdx.x, by, bx;
    if (k < t.rowsA * t.nTC) {
        const int r = k / t.nTC;
        bx = k - r * t.nTC;
        by = r < t.a0 ? r : t.b0 + (r - t.a0);
    } else {
        k -= t.rowsA * t.nTC;
        const int r = k / t.colsB, cc = k - r * t.colsB;
        by = t.a0 + r;
        bx = cc < t.c0 ? cc : t.d0 + (cc - t.c0);
    }
    const int ix = bx * 64 + (threadIdx.x & 63);
    const int iz = by * 4 + (threadIdx.x >> 6), s = blockIdx.y;
    if (ix >= a.nxp || iz >= a.nzp) return;
    const bool sx = xstrip(a, iz, ix), sz = zstrip(a, iz, ix);
    if (!sx && !sz) return;
    const long long c = (long long)iz * a.P + ix;
    const long long so = (long long)s * a.sstride;
    const float *X = a.X + so;
    // u_of at a neighbour cc of a strip cell (inside the field: a strip cell is
    // at least 4 cells from every edge), zero where the neighbour is off the
    // strip; the loads are unconditional, so all neighbours' loads are issued
    // together
    auto w_at = [&](bool on, long long cc, const float *ca, const float *cb, const float *nu) {
        const float v = a.v2[cc], x = X[cc], cav = ca[cc], cbv = cb[cc], nv = nu[cc];
        const float q = v * x;
        const float w = q + cbv * (q + cav * nv);
        return on ? w : 0.0f;
    };
    if (sx) {
        const float *nu = a.nux_in + so;
        float acc = 0.0f;
#pragma unroll
        for (int m = 1; m <= 4; ++m) {
            // w(c') = Sx(c') (q + b nu_k)(c'), nu_k = Sx (q + a nu_{k+1})
            const int xm = ix - m, xp = ix + m;
            const float wm = w_at(xstrip(a, iz, xm), c - m, a.cax, a.cbx, nu);
            const float wp = w_at(xstrip(a, iz, xp), c + m, a.cax, a.cbx, nu);
            acc += ka(m) * (wm - wp);
        }
        float *mu = a.mux + so;
        mu[c] = a.cax[c] * mu[c] + acc * a.idx;
    }
    if (sz) {
        const float *nu = a.nuz_in + so;
        const long long P = a.P;
        float acc = 0.0f;
#pragma unroll
        for (int m = 1; m <= 4; ++m) {
            const int zm = iz - m, zp = iz + m;
            const float wm = w_at(zstrip(a, zm, ix), c - m * P, a.caz, a.cbz, nu);
            const float wp = w_at(zstrip(a, zp, ix), c + m * P, a.caz, a.cbz, nu);
            acc += ka(m) * (wm - wp);
        }
        float *mu = a.muz + so;
        mu[c] = a.caz[c] * mu[c] + acc * a.idz;
    }
}

__global__ __launch_bounds__(256) void k_exact_lambda(const ExactArgs a) {
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iz = blockIdx.y * 4 + (threadIdx.x >> 6), s = blockIdx.z;
    if (ix >= a.nxp || iz >= a.nzp) return;
    const long long P = a.P;
    const long long c = (long long)iz * P + ix;
    const long long so = (long long)s * a.sstride, fo = (long long)s * a.fstride;
    const float *X = a.X + so;
    float *Y = a.Y + so;
    if (!live_at(a, iz, ix)) {
        Y[c] = 0.0f;
        if (a.cpml) {
            a.nux_out[so + c] = 0.0f;
            a.nuz_out[so + c] = 0.0f;
        }
        return;
    }
    const float *C = a.C + fo;
    const float qxx = d2s(C, c, 1) * a.idx2, qzz = d2s(C, c, P) * a.idz2;
    if (!a.cpml) {
        const float D = a.cerx[ix] * a.cerz[iz];
        const float phi = D * X[c];
        // V D X at the neighbours (zero off the live cells)
        auto vdx = [&](int jz, int jx, long long cc) { return a.v2[cc] * (a.cerx[jx] * a.cerz[jz]) * X[cc]; };
        float sx = C0 * vdx(iz, ix, c), sz = C0 * vdx(iz, ix, c);
#pragma unroll
        for (int m = 1; m <= 4; ++m) {
            sx += kc(m) * (vdx(iz, ix + m, c + m) + vdx(iz, ix - m, c - m));
            sz += kc(m) * (vdx(iz + m, ix, c + m * P) + vdx(iz - m, ix, c - m * P));
        }
        Y[c] = sx * a.idx2 + sz * a.idz2 + 2.0f * phi - (D * D) * Y[c];
        a.gV[so + c] += phi * (qxx + qzz);
        return;
    }
    // ---- CPML
    const bool sxs = xstrip(a, iz, ix), szs = zstrip(a, iz, ix);
    const float *nxi = a.nux_in + so, *nzi = a.nuz_in + so;
    const float *mux = a.mux + so, *muz = a.muz + so;
    const float phi = X[c];
    const float q = a.v2[c] * phi;
    const float nx_k = sxs ? q + a.cax[c] * nxi[c] : 0.0f;
    const float nz_k = szs ? q + a.caz[c] * nzi[c] : 0.0f;
    float sx = C0 * (q + a.cbx[c] * nx_k), sz = C0 * (q + a.cbz[c] * nz_k);
    float tx = 0.0f, tz = 0.0f;
    // mu vanishes off the strips (cleared, and written on strip cells only), so
    // the D^T (b mu) terms of a cell more than 4 cells from every x (z) strip
    // are zero and are not read
    const bool nearx = ix < a.N + 4 || ix >= a.nxp - a.N - 4;
    const bool nearz = iz < a.N + 4 || iz >= a.nzp - a.N - 4;
#pragma unroll
    for (int m = 1; m <= 4; ++m) {
        sx += kc(m) * (u_of(a, X, nxi, a.cax, a.cbx, iz, ix + m, c + m, xstrip(a, iz, ix + m)) +
                       u_of(a, X, nxi, a.cax, a.cbx, iz, ix - m, c - m, xstrip(a, iz, ix - m)));
        sz += kc(m) * (u_of(a, X, nzi, a.caz, a.cbz, iz + m, ix, c + m * P, zstrip(a, iz + m, ix)) +
                       u_of(a, X, nzi, a.caz, a.cbz, iz - m, ix, c - m * P, zstrip(a, iz - m, ix)));
        // D^T (b mu) = -D (b mu)
        if (nearx) tx += ka(m) * (a.cbx[c - m] * mux[c - m] - a.cbx[c + m] * mux[c + m]);
        if (nearz) tz += ka(m) * (a.cbz[c - m * P] * muz[c - m * P] - a.cbz[c + m * P] * muz[c + m * P]);
    }
    Y[c] = sx * a.idx2 + sz * a.idz2 + 2.0f * phi - Y[c] + tx * a.idx + tz * a.idz;
    a.nux_out[so + c] = nx_k;
    a.nuz_out[so + c] = nz_k;
    float Q = qxx + qzz;
    if (sxs) {
        const float *px = a.psx + fo;
        const float dpx = d1s(px, c, 1) * a.idx;
        Q += dpx + a.zex[fo + c];
        const float pm = a.psxm ? a.psxm[fo + c] : 0.0f, zm = a.zexm ? a.zexm[fo + c] : 0.0f;
        a.gax[so + c] += mux[c] * pm + nx_k * zm;
        a.gbx[so + c] += mux[c] * (d1s(C, c, 1) * a.idx) + nx_k * (qxx + dpx);
    }
    if (szs) {
        const float *pz = a.psz + fo;
        const float dpz = d1s(pz, c, P) * a.idz;
        Q += dpz + a.zez[fo + c];
        const float pm = a.pszm ? a.pszm[fo + c] : 0.0f, zm = a.zezm ? a.zezm[fo + c] : 0.0f;
        a.gaz[so + c] += muz[c] * pm + nz_k * zm;
        a.gbz[so + c] += muz[c] * (d1s(C, c, P) * a.idz) + nz_k * (qzz + dpz);
    }
    a.gV[so + c] += phi * Q;
}

// The CPML case of k_exact_lambda on tiles of 64 x 16 cells: the values a
// cell takes from its neighbours -- u_x = V X (+ b (V X + a nu_x) on x-strip
// cells), u_z alike, the saved field C, and the b mu products near the strips --
// are staged in LDS once per block over the tile and a 4-cell halo, instead of
// being read through the caches by every neighbour; each value, and each
// cell's result, is formed by the same expressions as in k_exact_lambda.
// Thread t updates column t mod 64 of rows 4 (t / 64) .. 4 (t / 64) + 3.
constexpr int kLamTZ = 16;
__device__ __forceinline__ float d2l(const float *f, int l, int st) {
    return C0 * f[l] + C1 * (f[l + st] + f[l - st]) + C2 * (f[l + 2 * st] + f[l - 2 * st]) +
           C3 * (f[l + 3 * st] + f[l - 3 * st]) + C4 * (f[l + 4 * st] + f[l - 4 * st]);
}
__device__ __forceinline__ float d1l(const float *f, int l, int st) {
    return A1 * (f[l + st] - f[l - st]) + A2 * (f[l + 2 * st] - f[l - 2 * st]) +
           A3 * (f[l + 3 * st] - f[l - 3 * st]) + A4 * (f[l + 4 * st] - f[l - 4 * st]);
}
__global__ __launch_bounds__(256) void k_exact_lambda_cpml(const ExactArgs a) {
    // C on the tile and a 4-cell halo on every side; the x-neighbour values
    // (u_x, b mu_x) on the tile's rows with the x halo, the z-neighbour values
    // (u_z, b mu_z) on its columns with the z halo: 28.4 KB, 5 blocks per CU
    constexpr int W = 64 + 8, H = kLamTZ + 8, NS = W * H, NX = W * kLamTZ, NZ = 64 * H;
    __shared__ float sux[NX], suz[NZ], sC[NS], sbx[NX], sbz[NZ];
    const int x0 = blockIdx.x * 64, z0 = blockIdx.y * kLamTZ, s = blockIdx.z, tid = threadIdx.x;
    const long long P = a.P;
    const long long so = (long long)s * a.sstride, fo = (long long)s * a.fstride;
    const float *X = a.X + so;
    float *Y = a.Y + so;
    const float *C = a.C + fo;
    const float *nxi = a.nux_in + so, *nzi = a.nuz_in + so;
    const float *mux = a.mux + so, *muz = a.muz + so;
    // the tile has cells within 4 of an x (z) strip: the b mu terms (mu is zero
    // off the strips, so other cells' terms are zero and not read)
    const bool tnx = x0 < a.N + 4 || x0 + 63 >= a.nxp - a.N - 4;
    const bool tnz = z0 < a.N + 4 || z0 + kLamTZ - 1 >= a.nzp - a.N - 4;
    // the tile with its halo meets an x (z) strip: u_x (u_z) differs from V X there
    const bool tsx = x0 - 4 < a.N || x0 + 68 > a.nxp - a.N;
    const bool tsz = z0 - 4 < a.N || z0 + kLamTZ + 4 > a.nzp - a.N;
    // Every load of the block is issued before the first is used: the values
    // of the thread's own cells, then the tile with its halo.  Loads outside
    // the field read cell 0 and are replaced by zero (no branches, so the
    // unrolled loads stay in one basic block).
    const int lx = tid & 63, ix = x0 + lx;
    float pX[4], pV[4], pY[4], pG[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        const int iz = z0 + (tid >> 6) * 4 + rr;
        const bool lv = ix < a.nxp && iz < a.nzp && live_at(a, iz, ix);
        const long long c = lv ? (long long)iz * P + ix : 0;
        pX[rr] = X[c];
        pV[rr] = a.v2[c];
        pY[rr] = Y[c];
        pG[rr] = a.gV[so + c];
    }
    constexpr int NIT = (NS + 255) / 256;
    float rX[NIT], rV[NIT], rC[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int i = tid + it * 256, r = i / W, cc = i - r * W;
        const int iz = z0 - 4 + r, jx = x0 - 4 + cc;
        const bool in = i < NS && iz >= 0 && iz < a.nzp && jx >= 0 && jx < a.nxp;
        const long long c = in ? (long long)iz * P + jx : 0;
        const float xv = X[c], vv = a.v2[c], cv = C[c];
        rX[it] = in ? xv : 0.0f;
        rV[it] = in ? vv : 0.0f;
        rC[it] = in ? cv : 0.0f;
    }
    float ux[NIT], uz[NIT], bx[NIT], bz[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const float q = rV[it] * rX[it];  // X vanishes off the live cells
        ux[it] = q;
        uz[it] = q;
        bx[it] = 0.0f;
        bz[it] = 0.0f;
    }
    if (tsx || tnx) {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int i = tid + it * 256, r = i / W, cc = i - r * W;
            const int iz = z0 - 4 + r, jx = x0 - 4 + cc;
            const bool in = i < NS && iz >= 0 && iz < a.nzp && jx >= 0 && jx < a.nxp;
            const long long c = in ? (long long)iz * P + jx : 0;
            const float ca = a.cax[c], cb = a.cbx[c], nu = nxi[c], m = mux[c];
            const float q = ux[it];
            if (tsx && in && xstrip(a, iz, jx)) ux[it] = q + cb * (q + ca * nu);
            if (tnx && in) bx[it] = cb * m;
        }
    }
    if (tsz || tnz) {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int i = tid + it * 256, r = i / W, cc = i - r * W;
            const int iz = z0 - 4 + r, jx = x0 - 4 + cc;
            const bool in = i < NS && iz >= 0 && iz < a.nzp && jx >= 0 && jx < a.nxp;
            const long long c = in ? (long long)iz * P + jx : 0;
            const float ca = a.caz[c], cb = a.cbz[c], nu = nzi[c], m = muz[c];
            const float q = uz[it];
            if (tsz && in && zstrip(a, iz, jx)) uz[it] = q + cb * (q + ca * nu);
            if (tnz && in) bz[it] = cb * m;
        }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int i = tid + it * 256;
        if (i < NS) {
            const int r = i / W, cc = i - r * W;
            sC[i] = rC[it];
            if (r >= 4 && r < kLamTZ + 4) {
                sux[i - 4 * W] = ux[it];
                sbx[i - 4 * W] = bx[it];
            }
            if (cc >= 4 && cc < 68) {
                suz[r * 64 + cc - 4] = uz[it];
                sbz[r * 64 + cc - 4] = bz[it];
            }
        }
    }
    __syncthreads();
    if (ix >= a.nxp) return;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        const int lz = (tid >> 6) * 4 + rr, iz = z0 + lz;
        if (iz >= a.nzp) break;
        const long long c = (long long)iz * P + ix;
        if (!live_at(a, iz, ix)) {
            Y[c] = 0.0f;
            continue;
        }
        const int l = (lz + 4) * W + lx + 4, lxr = lz * W + lx + 4, lzr = (lz + 4) * 64 + lx;
        const float qxx = d2l(sC, l, 1) * a.idx2, qzz = d2l(sC, l, W) * a.idz2;
        const bool sxs = xstrip(a, iz, ix), szs = zstrip(a, iz, ix);
        // a strip cell's own inputs, each strip's in one block of loads (one
        // memory round instead of one for nu_k and one for the gradient terms)
        float xca = 0.0f, xcb = 0.0f, xnu = 0.0f, xdp = 0.0f, xze = 0.0f, xpm = 0.0f, xzm = 0.0f, xga = 0.0f,
              xgb = 0.0f, xmu = 0.0f;
        float zca = 0.0f, zcb = 0.0f, znu = 0.0f, zdp = 0.0f, zze = 0.0f, zpm = 0.0f, zzm = 0.0f, zga = 0.0f,
              zgb = 0.0f, zmu = 0.0f;
        if (sxs) {
            xca = a.cax[c];
            xcb = a.cbx[c];
            xnu = nxi[c];
            xze = a.zex[fo + c];
            xpm = a.psxm ? a.psxm[fo + c] : 0.0f;
            xzm = a.zexm ? a.zexm[fo + c] : 0.0f;
            xga = a.gax[so + c];
            xgb = a.gbx[so + c];
            xmu = mux[c];
            xdp = d1s(a.psx + fo, c, 1) * a.idx;
        }
        if (szs) {
            zca = a.caz[c];
            zcb = a.cbz[c];
            znu = nzi[c];
            zze = a.zez[fo + c];
            zpm = a.pszm ? a.pszm[fo + c] : 0.0f;
            zzm = a.zezm ? a.zezm[fo + c] : 0.0f;
            zga = a.gaz[so + c];
            zgb = a.gbz[so + c];
            zmu = muz[c];
            zdp = d1s(a.psz + fo, c, P) * a.idz;
        }
        const float phi = pX[rr];
        const float q = pV[rr] * phi;
        const float nx_k = sxs ? q + xca * xnu : 0.0f;
        const float nz_k = szs ? q + zca * znu : 0.0f;
        // off a strip nu is zero: q alone (without the zero b nu term), and the nu
        // buffers keep the zeros they were cleared to
        float sx = C0 * (sxs ? q + xcb * nx_k : q), sz = C0 * (szs ? q + zcb * nz_k : q);
        float tx = 0.0f, tz = 0.0f;
        const bool nearx = ix < a.N + 4 || ix >= a.nxp - a.N - 4;
        const bool nearz = iz < a.N + 4 || iz >= a.nzp - a.N - 4;
#pragma unroll
        for (int m = 1; m <= 4; ++m) {
            sx += kc(m) * (sux[lxr + m] + sux[lxr - m]);
            sz += kc(m) * (suz[lzr + m * 64] + suz[lzr - m * 64]);
            // D^T (b mu) = -D (b mu)
            if (nearx) tx += ka(m) * (sbx[lxr - m] - sbx[lxr + m]);
            if (nearz) tz += ka(m) * (sbz[lzr - m * 64] - sbz[lzr + m * 64]);
        }
        Y[c] = sx * a.idx2 + sz * a.idz2 + 2.0f * phi - pY[rr] + tx * a.idx + tz * a.idz;
        if (sxs) a.nux_out[so + c] = nx_k;
        if (szs) a.nuz_out[so + c] = nz_k;
        float Q = qxx + qzz;
        if (sxs) {
            Q += xdp + xze;
            a.gax[so + c] = xga + (xmu * xpm + nx_k * xzm);
            a.gbx[so + c] = xgb + (xmu * (d1l(sC, l, 1) * a.idx) + nx_k * (qxx + xdp));
        }
        if (szs) {
            Q += zdp + zze;
            a.gaz[so + c] = zga + (zmu * zpm + nz_k * zzm);
            a.gbz[so + c] = zgb + (zmu * (d1l(sC, l, W) * a.idz) + nz_k * (qzz + zdp));
        }
        a.gV[so + c] = pG[rr] + phi * Q;
    }
}

// one thread per (shot, receiver); duplicate cells accumulate through atomics
__global__ void k_exact_rec(float *Y, long long sstride, int P, const int *rz, const int *rx, int nrec,
                            const float *data, long long dstride, int row, const float *cerx, const float *cerz,
                            int nshots) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (r >= nrec || s >= nshots) return;
    const int z = rz[r], x = rx[r];
    const float w = cerx ? cerx[x] * cerz[z] : 1.0f;
    atomicAdd(Y + (long long)s * sstride + (long long)z * P + x,
              w * data[(long long)s * dstride + (long long)row * nrec + r]);
}

// d a / d vp and d b / d vp of the CPML profile along one axis (the formula of
// k_cpml_coeffs differentiated in vp)
__device__ __forceinline__ void cpml_dab(int i, int n, int N, float h, float vpc, float d0, float fp, float dt,
                                         float *da, float *db) {
    float kap = 0.0f, alpha = 0.0f;
    if (i < N) {
        const float pos = ((N - i - 1.0f) * h) / (N * h);
        kap = d0 / (2.0f * N * h) * (pos * pos);
        alpha = 3.1415926535f * fp * (1.0f - pos * pos);
    } else if (i >= n - N) {
        const float pos = ((float)(i - n + N) * h) / (N * h);
        kap = d0 / (2.0f * N * h) * (pos * pos);
        alpha = 3.1415926535f * fp * (1.0f - pos * pos);
    }
    const float d = kap * vpc, sum = d + alpha;
    const float av = expf(-sum * dt);
    *da = -kap * dt * av;
    *db = (fabsf(sum) > 1e-10f) ? kap * alpha / (sum * sum) * (av - 1.0f) + d / sum * (*da) : 0.0f;
}

__global__ void k_exact_finish(const ExactArgs a, const float *vp, float dx, float dz, float dt, float d0, float fp,
                               float *gacc) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= a.nxp || iz >= a.nzp) return;
    const long long c = (long long)iz * a.P + ix;
    float gv = 0.0f, gax = 0.0f, gbx = 0.0f, gaz = 0.0f, gbz = 0.0f;
    for (int s = 0; s < a.nshots; ++s) {
        const long long i = (long long)s * a.sstride + c;
        gv += a.gV[i];
        if (a.cpml) {
            gax += a.gax[i];
            gbx += a.gbx[i];
            gaz += a.gaz[i];
            gbz += a.gbz[i];
        }
    }
    const float v = vp[c];
    float g = gv * (2.0f * v * dt * dt);
    if (a.cpml) {
        float da, db;
        cpml_dab(ix, a.nxp, a.N, dx, v, d0, fp, dt, &da, &db);
        g += gax * da + gbx * db;
        cpml_dab(iz, a.nzp, a.N, dz, v, d0, fp, dt, &da, &db);
        g += gaz * da + gbz * db;
    }
    gacc[c] += g;
}

__global__ void k_expand_T(const float *g, int nzp, int nxp, int P, int N, float *out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int nz = nzp - 2 * N, nx = nxp - 2 * N;
    if (x >= nx || z >= nz) return;
    const int z0 = (z == 0) ? 0 : z + N, z1 = (z == nz - 1) ? nzp - 1 : z + N;
    const int x0 = (x == 0) ? 0 : x + N, x1 = (x == nx - 1) ? nxp - 1 : x + N;
    float acc = 0.0f;
    for (int zz = z0; zz <= z1; ++zz)
        for (int xx = x0; xx <= x1; ++xx) acc += g[(long long)zz * P + xx];
    out[(long long)z * nx + x] = acc;
}

}  // namespace

// blocks of 64 x 4 cells: the z-neighbours of a cell are mostly in its own
// block, so they come from L1 (rows of 256 cells fetched every z-neighbour
// through L2); only the tiles that hold strip cells are launched
int launch_exact_mu(const ExactArgs &a, hipStream_t st) {
    const StripTiles t = strip_tiles(a);
    const int nb = t.rowsA * t.nTC + (t.nTR - t.rowsA) * t.colsB;
    if (nb <= 0 || a.nshots <= 0) return 0;
    hipLaunchKernelGGL(k_exact_mu, dim3(nb, a.nshots), dim3(256), 0, st, a, t);
    return (int)hipGetLastError();
}
int launch_exact_lambda(const ExactArgs &a, hipStream_t st) {
    if (a.cpml)
        hipLaunchKernelGGL(k_exact_lambda_cpml, dim3((a.nxp + 63) / 64, (a.nzp + kLamTZ - 1) / kLamTZ, a.nshots),
                           dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_exact_lambda, dim3((a.nxp + 63) / 64, (a.nzp + 3) / 4, a.nshots), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}
int launch_exact_rec(float *Y, long long sstride, int P, const int *rz, const int *rx, int nrec, const float *data,
                     long long dstride, int row, const float *cerx, const float *cerz, int nshots, hipStream_t st) {
    if (nrec <= 0) return 0;
    hipLaunchKernelGGL(k_exact_rec, dim3((nrec + 255) / 256, nshots), dim3(256), 0, st, Y, sstride, P, rz, rx, nrec,
                       data, dstride, row, cerx, cerz, nshots);
    return (int)hipGetLastError();
}
int launch_exact_finish(const ExactArgs &a, const float *vp, float dx, float dz, float dt, float d0, float fpico,
                        float *gacc, hipStream_t st) {
    hipLaunchKernelGGL(k_exact_finish, dim3((a.nxp + 255) / 256, a.nzp), dim3(256), 0, st, a, vp, dx, dz, dt, d0,
                       fpico, gacc);
    return (int)hipGetLastError();
}
int launch_expand_T(const float *gpad, int nzp, int nxp, int P, int N, float *out, hipStream_t st) {
    const int nz = nzp - 2 * N, nx = nxp - 2 * N;
    hipLaunchKernelGGL(k_expand_T, dim3((nx + 255) / 256, nz), dim3(256), 0, st, gpad, nzp, nxp, P, N, out);
    return (int)hipGetLastError();
}

}  // namespace pam
