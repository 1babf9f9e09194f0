/*
 * pam_oracle.c -- CPU restatement of the reference 2D finite-difference engine
 * (GISIS-UFF/PseudoAcousticModeling).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.  It follows the reference's loop nests one kernel at a time
 * (race-free semantics: update pass, then damping pass), so that the HIP path
 * can be compared with it on identical inputs.
 *
 * One source, three builds (see oracle/Makefile):
 *   ORACLE_F64   store=double real=double        accuracy oracle
 *   ORACLE_F32   store=float  real=float         FP32 restatement, no FMA contraction
 *   ORACLE_MIXED store=float  real=double +OpenMP float storage with float64
 *                                                 arithmetic, rows in parallel: the
 *                                                 Numba kernels' promotion and loop
 *                                                 structure (timed CPU baseline)
 *
 * Layout: every padded field is row-major [nzp][nxp] (x fastest), nzp = nz+2N,
 * nxp = nx+2N, index iz*nxp+ix -- reference cpp/Modeling.cu:514-515 and the
 * Numba [j,i] indexing in src/utils.py:404-408.  The stencil interior is
 * 4 <= ix < nxp-4, 4 <= iz < nzp-4; the outer 4-cell rim is never updated.
 *
 * CPML memory variables are full-grid fields, zero outside their strips:
 *   psix/zetax live on rows [4,nzp-4), columns [4,N) and [nxp-N,nxp-4);
 *   psiz/zetaz live on rows [4,N) and [nzp-N,nzp-4), columns [4,nxp-4).
 * A derivative that reaches outside a strip therefore reads zero, which is what
 * the reference's compact strips read there too: never-written columns in the
 * Numba [nzp][N+4] arrays (src/utils.py:242-377), never-written cells reached
 * by the stride-N indexing of the CuPy kernels (src/CudaKernels.py:114,126),
 * and the negative-offset read of PsizFD (src/CudaKernels.py:256-258), defined
 * here as zero.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORACLE_F64)
typedef double real;
typedef double store;
#define RSQRT sqrt
#define RFABS fabs
#define REXP exp
#define RCOS cos
#define RSIN sin
#define DEN_EPS 1e-12
#elif defined(ORACLE_F32)
typedef float real;
typedef float store;
#define RFABS fabsf
#define REXP expf
#define RCOS cosf
#define RSIN sinf
#define DEN_EPS 1e-12f
#elif defined(ORACLE_MIXED)
typedef double real;
typedef float store;
#define RFABS fabs
#define REXP exp
#define RCOS cos
#define RSIN sin
#define DEN_EPS 1e-12
#else
#error "define ORACLE_F64, ORACLE_F32 or ORACLE_MIXED"
#endif

/* Row loops run on OpenMP threads on grids of at least OMP_MIN_CELLS cells;
 * smaller grids run on the calling thread, where a parallel region costs more
 * than the loop (and many small calls from several processes at once would
 * oversubscribe the cores). */
#define OMP_MIN_CELLS 65536L
#ifdef _OPENMP
#define OMP_STR(x) #x
#define OMP_ROWS(cells) _Pragma(OMP_STR(omp parallel for schedule(static) if ((cells) >= OMP_MIN_CELLS)))
#else
#define OMP_ROWS(cells)
#endif

#define R(x) ((real)(x))

/* 8th-order coefficients: src/CudaKernels.py:272-276 (second derivative) and
 * :316-319 (first derivative); the C++ decimals (cpp/Modeling.cu:500-504,
 * 560-563) are the same values. */
#define C0 (R(-1435.0) / R(504.0))
#define C1 (R(8.0) / R(5.0))
#define C2 (R(-1.0) / R(5.0))
#define C3 (R(8.0) / R(315.0))
#define C4 (R(-1.0) / R(560.0))
#define A1 (R(4.0) / R(5.0))
#define A2 (R(-1.0) / R(5.0))
#define A3 (R(4.0) / R(105.0))
#define A4 (R(-1.0) / R(280.0))

/* second derivative along a stride s, undivided */
static inline real d2(const store *U, long i, long s) {
    return C0 * U[i] + C1 * (U[i + s] + U[i - s]) + C2 * (U[i + 2 * s] + U[i - 2 * s]) +
           C3 * (U[i + 3 * s] + U[i - 3 * s]) + C4 * (U[i + 4 * s] + U[i - 4 * s]);
}
/* first derivative along a stride s, undivided */
static inline real d1(const store *U, long i, long s) {
    return A1 * (U[i + s] - U[i - s]) + A2 * (U[i + 2 * s] - U[i - 2 * s]) +
           A3 * (U[i + 3 * s] - U[i - 3 * s]) + A4 * (U[i + 4 * s] - U[i - 4 * s]);
}
/* 8th-order mixed derivative, undivided; term order of src/CudaKernels.py:405-425:
 * sum_i sum_j a_i a_j (U[+j,+i] - U[-j,+i] + U[-j,-i] - U[+j,-i]) */
static inline real dxz(const store *U, long i, long n) {
    real acc = 0;
    for (int ii = 1; ii <= 4; ++ii) {
        real ai = (ii == 1) ? A1 : (ii == 2) ? A2 : (ii == 3) ? A3 : A4;
        for (int jj = 1; jj <= 4; ++jj) {
            real aj = (jj == 1) ? A1 : (jj == 2) ? A2 : (jj == 3) ? A3 : A4;
            acc += ai * aj *
                   (U[i + jj * n + ii] - U[i - jj * n + ii] + U[i - jj * n - ii] - U[i + jj * n - ii]);
        }
    }
    return acc;
}

static inline int interior(int iz, int ix, int nzp, int nxp) {
    return ix >= 4 && ix < nxp - 4 && iz >= 4 && iz < nzp - 4;
}
static inline int in_xstrip(int ix, int nxp, int N) {
    return (ix >= 4 && ix < N) || (ix >= nxp - N && ix < nxp - 4);
}
static inline int in_zstrip(int iz, int nzp, int N) {
    return (iz >= 4 && iz < N) || (iz >= nzp - N && iz < nzp - 4);
}

/* ------------------------------------------------------------------ setup */

/* Edge-replicating pad by N on every side, corners take the corner value.
 * cpp/Modeling.cu:130-182; src/Modeling2D.py:38-54. */
void or_expand(const float *m, int nz, int nx, int N, store *out) {
    const int nxp = nx + 2 * N, nzp = nz + 2 * N;
    for (int iz = 0; iz < nzp; ++iz) {
        int jz = iz - N;
        if (jz < 0) jz = 0;
        if (jz > nz - 1) jz = nz - 1;
        for (int ix = 0; ix < nxp; ++ix) {
            int jx = ix - N;
            if (jx < 0) jx = 0;
            if (jx > nx - 1) jx = nx - 1;
            out[(long)iz * nxp + ix] = m[(long)jz * nx + jx];
        }
    }
}

/* Cerjan taper A[i] = exp(-fb^2), i < N.  With factor <= 0 the reference's own
 * form fb = (N-i)/(1.4142*3N) (cpp/Modeling.cu:111-120, src/Modeling2D.py:258-264);
 * otherwise fb = factor*(N-i). */
void or_cerjan(int N, double factor, store *A) {
    for (int i = 0; i < N; ++i) {
        real fb;
        if (factor <= 0.0)
            fb = (R(N) - R(i)) / (R(1.4142) * (R(3.0) * R(N)));
        else
            fb = R(factor) * (R(N) - R(i));
        A[i] = (store)REXP(-fb * fb);
    }
}

/* CPML damping profiles a,b per cell, from src/CudaKernels.py:33-90 and
 * src/utils.py:198-240.  Full-grid output; ax/bx meaningful on x-strip cells,
 * az/bz on z-strip cells (1/0 elsewhere). */
void or_cpml_profiles(const store *vp, int nzp, int nxp, int N, double dx, double dz, double dt,
                      double d0, double fpico, store *ax, store *bx, store *az, store *bz) {
    const real PI = R(3.1415926535);
    for (int j = 0; j < nzp; ++j)
        for (int i = 0; i < nxp; ++i) {
            long c = (long)j * nxp + i;
            real d = 0, alpha = 0, pos;
            if (i < N) {
                pos = ((R(N) - R(i) - R(1.0)) * R(dx)) / (R(N) * R(dx));
                d = R(d0) / (R(2.0) * R(N) * R(dx)) * (pos * pos) * vp[c];
                alpha = PI * R(fpico) * (R(1.0) - pos * pos);
            } else if (i >= nxp - N) {
                pos = ((R(i) - R(nxp) + R(N)) * R(dx)) / (R(N) * R(dx));
                d = R(d0) / (R(2.0) * R(N) * R(dx)) * (pos * pos) * vp[c];
                alpha = PI * R(fpico) * (R(1.0) - pos * pos);
            }
            real a = REXP(-(d + alpha) * R(dt));
            real b = 0;
            if (RFABS(d + alpha) > R(1e-10)) b = (d / (d + alpha)) * (a - R(1.0));
            ax[c] = (store)a;
            bx[c] = (store)b;
            d = 0;
            alpha = 0;
            if (j < N) {
                pos = ((R(N) - R(j) - R(1.0)) * R(dz)) / (R(N) * R(dz));
                d = R(d0) / (R(2.0) * R(N) * R(dz)) * (pos * pos) * vp[c];
                alpha = PI * R(fpico) * (R(1.0) - pos * pos);
            } else if (j >= nzp - N) {
                pos = ((R(j) - R(nzp) + R(N)) * R(dz)) / (R(N) * R(dz));
                d = R(d0) / (R(2.0) * R(N) * R(dz)) * (pos * pos) * vp[c];
                alpha = PI * R(fpico) * (R(1.0) - pos * pos);
            }
            a = REXP(-(d + alpha) * R(dt));
            b = 0;
            if (RFABS(d + alpha) > R(1e-10)) b = (d / (d + alpha)) * (a - R(1.0));
            az[c] = (store)a;
            bz[c] = (store)b;
        }
}

/* ------------------------------------------------------ source / receivers */

/* U[sz,sx] += val  -- cpp/Modeling.cu:481-486; src/Modeling2D.py:336,364 */
void or_inject(store *U, int nxp, int sz, int sx, double val) {
    U[(long)sz * nxp + sx] = (store)((real)U[(long)sz * nxp + sx] + (real)(store)val);
}

/* out[r] = U[rz[r], rx[r]]  -- cpp/Modeling.cu:488-497; src/Modeling2D.py:292-300 */
void or_gather(const store *U, int nxp, int nrec, const int *rz, const int *rx, store *out) {
    for (int r = 0; r < nrec; ++r) out[r] = U[(long)rz[r] * nxp + rx[r]];
}

/* U[rz,rx] += vals[r]/dxdz, accumulating duplicates -- src/Migration.py:238-240
 * (data / (dx*dz)); cpp/Migration.cu:493-502 multiplies by 1/(dx*dz). */
void or_inject_rec(store *U, int nxp, int nrec, const int *rz, const int *rx, const store *vals,
                   double dxdz) {
    for (int r = 0; r < nrec; ++r) {
        long c = (long)rz[r] * nxp + rx[r];
        U[c] = (store)((real)U[c] + (real)(store)((real)vals[r] / (real)dxdz));
    }
}

/* ------------------------------------------------------------ forward ops */

/* Acoustic leapfrog, src/CudaKernels.py:268-305 / src/utils.py:396-410:
 * Uf = (vp*vp)*(dt*dt)*(pxx+pzz) + 2*Uc - Uf on the interior. */
/* When set, the three update routines store only the operator term
 * V L(Uc) of the leapfrog (Uf = V L(Uc) instead of V L(Uc) + 2 Uc - Uf): the
 * second time difference S_{k+1} - 2 S_k + S_{k-1} of an undamped step, which
 * the boundary-saving rebuild needs without cancelling time levels. */
static int g_op_only = 0;
void or_set_operator_only(int on) { g_op_only = on; }

void or_update_acoustic(store *Uf, const store *Uc, const store *vp, int nzp, int nxp, double dz,
                        double dx, double dt) {
    const real idx2 = R(dx) * R(dx), idz2 = R(dz) * R(dz), dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real pxx = d2(Uc, c, 1) / idx2;
            real pzz = d2(Uc, c, nxp) / idz2;
            const real op = ((real)vp[c] * vp[c]) * dt2 * (pxx + pzz);
            Uf[c] = (store)(g_op_only ? op : op + R(2.0) * Uc[c] - Uf[c]);
        }
}

static inline real sd_term(real e, real dl, real px, real pz) {
    real px2 = px * px, pz2 = pz * pz;
    real num = R(-2.0) * (e - dl) * px2 * pz2;
#ifdef ORACLE_PLANT_SD_DEN_EPS
    /* a mistake planted on purpose, 2 (1 + eps) for 2 (1 + delta): only the
       oracle's own pin test builds this variant, to show that its criterion
       (the exact acoustic qP phase velocity to O(t^3)) catches it */
    real den = (R(1.0) + R(2.0) * e) * (px2 * px2) + (pz2 * pz2) + R(2.0) * (R(1.0) + e) * px2 * pz2;
#else
    real den = (R(1.0) + R(2.0) * e) * (px2 * px2) + (pz2 * pz2) + R(2.0) * (R(1.0) + dl) * px2 * pz2;
#endif
    return (RFABS(den) > DEN_EPS) ? num / den : R(0.0);
}

/* VTI pseudo-acoustic step, src/CudaKernels.py:307-366 / src/utils.py:412-456 /
 * cpp/Modeling.cu:554-631 (Cerjan applied separately). */
void or_update_vti(store *Uf, const store *Uc, const store *vp, const store *eps, const store *del,
                   int nzp, int nxp, double dz, double dx, double dt) {
    const real dx2 = R(dx) * R(dx), dz2 = R(dz) * R(dz), dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real pxx = d2(Uc, c, 1) / dx2, pzz = d2(Uc, c, nxp) / dz2;
            real px = d1(Uc, c, 1) / R(dx), pz = d1(Uc, c, nxp) / R(dz);
            real e = eps[c], dl = del[c];
            real Sd = sd_term(e, dl, px, pz);
            real v2 = (real)vp[c] * vp[c] * dt2;
            Uf[c] = (store)(g_op_only ? v2 * ((R(1.0) + R(2.0) * e) + Sd) * pxx + v2 * (R(1.0) + Sd) * pzz
                                      : R(2.0) * Uc[c] - Uf[c] + v2 * ((R(1.0) + R(2.0) * e) + Sd) * pxx +
                                            v2 * (R(1.0) + Sd) * pzz);
        }
}

/* TTI step, src/CudaKernels.py:368-452 / src/utils.py:458-527 / cpp/Modeling.cu:633-749.
 * theta in radians. */
void or_update_tti(store *Uf, const store *Uc, const store *vp, const store *eps, const store *del,
                   const store *theta, int nzp, int nxp, double dz, double dx, double dt) {
    const real dx2 = R(dx) * R(dx), dz2 = R(dz) * R(dz), dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real pxx = d2(Uc, c, 1) / dx2, pzz = d2(Uc, c, nxp) / dz2;
            real pxz = dxz(Uc, c, nxp) / (R(dz) * R(dx));
            real px = d1(Uc, c, 1) / R(dx), pz = d1(Uc, c, nxp) / R(dz);
            real e = eps[c], dl = del[c], th = theta[c];
            real cs = RCOS(th), sn = RSIN(th);
            real xi = px * cs - pz * sn, eta = px * sn + pz * cs;
            real Sd = sd_term(e, dl, xi, eta);
            real v2 = (real)vp[c] * vp[c] * dt2;
            real Ac = (R(1.0) + R(2.0) * e) * (cs * cs) + (sn * sn) + Sd;
            real Bc = (R(1.0) + R(2.0) * e) * (sn * sn) + (cs * cs) + Sd;
            Uf[c] = (store)(g_op_only ? v2 * Ac * pxx + v2 * Bc * pzz - R(2.0) * e * v2 * (R(2.0) * sn * cs) * pxz
                                      : R(2.0) * Uc[c] - Uf[c] + v2 * Ac * pxx + v2 * Bc * pzz -
                                            R(2.0) * e * v2 * (R(2.0) * sn * cs) * pxz);
        }
}

/* Cerjan damping of one field, src/CudaKernels.py:3-31 / src/utils.py:380-393:
 * left/right factor first, then top/bottom, so corners get the product. */
void or_cerjan_apply(store *U, int nzp, int nxp, int N, const store *A) {
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 0; j < nzp; ++j)
        for (int i = 0; i < nxp; ++i) {
            long c = (long)j * nxp + i;
            real u = U[c];
            if (i < N) u = (real)(store)(u * (real)A[i]);
            if (i >= nxp - N) u = (real)(store)(u * (real)A[nxp - 1 - i]);
            if (j < N) u = (real)(store)(u * (real)A[j]);
            if (j >= nzp - N) u = (real)(store)(u * (real)A[nzp - 1 - j]);
            U[c] = (store)u;
        }
}

/* CPML psi update, src/CudaKernels.py:92-163 / src/utils.py:242-299:
 * psi = a*psi + b*dU on the strips, from the current field. */
void or_cpml_psi(store *psix, store *psiz, const store *Uc, const store *ax, const store *bx,
                 const store *az, const store *bz, int nzp, int nxp, int N, double dx, double dz) {
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            if (in_xstrip(i, nxp, N))
                psix[c] = (store)((real)ax[c] * psix[c] + (real)bx[c] * (d1(Uc, c, 1) / R(dx)));
            if (in_zstrip(j, nzp, N))
                psiz[c] = (store)((real)az[c] * psiz[c] + (real)bz[c] * (d1(Uc, c, nxp) / R(dz)));
        }
}

/* CPML zeta update with the NEW psi, src/CudaKernels.py:165-266 / src/utils.py:301-377:
 * zeta = a*zeta + b*(d2U + dpsi). */
void or_cpml_zeta(store *zetax, store *zetaz, const store *psix, const store *psiz, const store *Uc,
                  const store *ax, const store *bx, const store *az, const store *bz, int nzp, int nxp,
                  int N, double dx, double dz) {
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            if (in_xstrip(i, nxp, N)) {
                real pxx = d2(Uc, c, 1) / (R(dx) * R(dx));
                real dps = d1(psix, c, 1) / R(dx);
                zetax[c] = (store)((real)ax[c] * zetax[c] + (real)bx[c] * (pxx + dps));
            }
            if (in_zstrip(j, nzp, N)) {
                real pzz = d2(Uc, c, nxp) / (R(dz) * R(dz));
                real dps = d1(psiz, c, nxp) / R(dz);
                zetaz[c] = (store)((real)az[c] * zetaz[c] + (real)bz[c] * (pzz + dps));
            }
        }
}

/* Acoustic CPML update, src/CudaKernels.py:454-721 / src/utils.py:768-942:
 * Uf = V*(pxx + pzz [+dpsix] [+dpsiz] [+zetax] [+zetaz]) + 2Uc - Uf. */
void or_update_acoustic_cpml(store *Uf, const store *Uc, const store *vp, const store *psix,
                             const store *psiz, const store *zetax, const store *zetaz, int nzp,
                             int nxp, int N, double dz, double dx, double dt) {
    const real dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            int xs = in_xstrip(i, nxp, N), zs = in_zstrip(j, nzp, N);
            real s = d2(Uc, c, 1) / (R(dx) * R(dx)) + d2(Uc, c, nxp) / (R(dz) * R(dz));
            if (xs) s += d1(psix, c, 1) / R(dx);
            if (zs) s += d1(psiz, c, nxp) / R(dz);
            if (xs) s += zetax[c];
            if (zs) s += zetaz[c];
            Uf[c] = (store)(((real)vp[c] * vp[c]) * dt2 * s + R(2.0) * Uc[c] - Uf[c]);
        }
}

/* VTI CPML update, src/CudaKernels.py:724-1172 / src/utils.py:944-1306:
 * px -> px + dpsix, pz -> pz + dpsiz inside Sd; pxx -> pxx + dpsix + zetax,
 * pzz -> pzz + dpsiz + zetaz.  The inverted |den| test of the CuPy lower-left
 * corner (src/CudaKernels.py:1109) is a defect and is not restated; the Numba
 * twin (src/utils.py:1256-1259) uses the same gate as every other region. */
void or_update_vti_cpml(store *Uf, const store *Uc, const store *vp, const store *eps,
                        const store *del, const store *psix, const store *psiz, const store *zetax,
                        const store *zetaz, int nzp, int nxp, int N, double dz, double dx, double dt) {
    const real dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            int xs = in_xstrip(i, nxp, N), zs = in_zstrip(j, nzp, N);
            real pxx = d2(Uc, c, 1) / (R(dx) * R(dx)), pzz = d2(Uc, c, nxp) / (R(dz) * R(dz));
            real px = d1(Uc, c, 1) / R(dx), pz = d1(Uc, c, nxp) / R(dz);
            real PXX = pxx, PZZ = pzz;
            if (xs) {
                real dps = d1(psix, c, 1) / R(dx);
                px = px + dps;
                PXX = pxx + dps + zetax[c];
            }
            if (zs) {
                real dps = d1(psiz, c, nxp) / R(dz);
                pz = pz + dps;
                PZZ = pzz + dps + zetaz[c];
            }
            real e = eps[c], dl = del[c];
            real Sd = sd_term(e, dl, px, pz);
            real v2 = (real)vp[c] * vp[c] * dt2;
            Uf[c] = (store)(R(2.0) * Uc[c] - Uf[c] + v2 * ((R(1.0) + R(2.0) * e) + Sd) * PXX +
                            v2 * (R(1.0) + Sd) * PZZ);
        }
}

/* TTI CPML update.  The reference has no TTI+CPML (src/Modeling2D.py:355-361
 * raises); this is the natural extension of the VTI-CPML form: the CPML-
 * corrected first derivatives are rotated into Sd, the corrected second
 * derivatives replace pxx/pzz, and the mixed derivative is left unstretched.
 * With theta = 0 it reduces to or_update_vti_cpml exactly. */
void or_update_tti_cpml(store *Uf, const store *Uc, const store *vp, const store *eps,
                        const store *del, const store *theta, const store *psix, const store *psiz,
                        const store *zetax, const store *zetaz, int nzp, int nxp, int N, double dz,
                        double dx, double dt) {
    const real dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            int xs = in_xstrip(i, nxp, N), zs = in_zstrip(j, nzp, N);
            real pxx = d2(Uc, c, 1) / (R(dx) * R(dx)), pzz = d2(Uc, c, nxp) / (R(dz) * R(dz));
            real pxz = dxz(Uc, c, nxp) / (R(dz) * R(dx));
            real px = d1(Uc, c, 1) / R(dx), pz = d1(Uc, c, nxp) / R(dz);
            real PXX = pxx, PZZ = pzz;
            if (xs) {
                real dps = d1(psix, c, 1) / R(dx);
                px = px + dps;
                PXX = pxx + dps + zetax[c];
            }
            if (zs) {
                real dps = d1(psiz, c, nxp) / R(dz);
                pz = pz + dps;
                PZZ = pzz + dps + zetaz[c];
            }
            real e = eps[c], dl = del[c], th = theta[c];
            real cs = RCOS(th), sn = RSIN(th);
            real xi = px * cs - pz * sn, eta = px * sn + pz * cs;
            real Sd = sd_term(e, dl, xi, eta);
            real v2 = (real)vp[c] * vp[c] * dt2;
            real Ac = (R(1.0) + R(2.0) * e) * (cs * cs) + (sn * sn) + Sd;
            real Bc = (R(1.0) + R(2.0) * e) * (sn * sn) + (cs * cs) + Sd;
            Uf[c] = (store)(R(2.0) * Uc[c] - Uf[c] + v2 * Ac * PXX + v2 * Bc * PZZ -
                            R(2.0) * e * v2 * (R(2.0) * sn * cs) * pxz);
        }
}

/* ------------------------------------------------------------ adjoint ops */

/* Linearised VTI adjoint around the source field P, src/CudaKernels.py:1382-1532 /
 * src/utils.py:530-639 / cpp/Migration.cu:573-755.  Products A*U, B*U, Q*Cx*U,
 * Q*Cz*U are formed on the interior (zero elsewhere), then
 * Uf = 2Uc - Uf + V*(dxx(AU) + dzz(BU) - dx(QCxU) - dz(QCzU)). */
void or_adjoint_vti(store *Uf, const store *Uc, const store *P, const store *vp, const store *eps,
                    const store *del, int nzp, int nxp, double dz, double dx, double dt) {
    const long n = (long)nzp * nxp;
    store *AU = (store *)calloc(n, sizeof(store)), *BU = (store *)calloc(n, sizeof(store));
    store *QX = (store *)calloc(n, sizeof(store)), *QZ = (store *)calloc(n, sizeof(store));
    const real dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real pxx = d2(P, c, 1) / (R(dx) * R(dx)), pzz = d2(P, c, nxp) / (R(dz) * R(dz));
            real px = d1(P, c, 1) / R(dx), pz = d1(P, c, nxp) / R(dz);
            real e = eps[c], dl = del[c];
            real px2 = px * px, pz2 = pz * pz, px4 = px2 * px2, pz4 = pz2 * pz2;
            real num = R(-2.0) * (e - dl) * px2 * pz2;
            real den = (R(1.0) + R(2.0) * e) * px4 + pz4 + R(2.0) * (R(1.0) + dl) * px2 * pz2;
            real Sd = 0, Cx = 0, Cz = 0;
            if (RFABS(den) > DEN_EPS) {
                Sd = num / den;
                real fac = R(4.0) * (e - dl) * ((R(1.0) + R(2.0) * e) * px4 - pz4) / (den * den);
                Cx = fac * px * pz2;
                Cz = -fac * px2 * pz;
            }
            real u = Uc[c];
            AU[c] = (store)((R(1.0) + R(2.0) * e + Sd) * u);
            BU[c] = (store)((R(1.0) + Sd) * u);
            QX[c] = (store)((pxx + pzz) * Cx * u);
            QZ[c] = (store)((pxx + pzz) * Cz * u);
        }
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real op = d2(AU, c, 1) / (R(dx) * R(dx)) + d2(BU, c, nxp) / (R(dz) * R(dz)) -
                      d1(QX, c, 1) / R(dx) - d1(QZ, c, nxp) / R(dz);
            Uf[c] = (store)(R(2.0) * Uc[c] - Uf[c] + (real)vp[c] * vp[c] * dt2 * op);
        }
    free(AU);
    free(BU);
    free(QX);
    free(QZ);
}

/* Linearised TTI adjoint, src/CudaKernels.py:1534-1715 / src/utils.py:641-765 /
 * cpp/Migration.cu:757-973: as VTI in rotated (xi, eta) plus -dxz(H*U). */
void or_adjoint_tti(store *Uf, const store *Uc, const store *P, const store *vp, const store *eps,
                    const store *del, const store *theta, int nzp, int nxp, double dz, double dx,
                    double dt) {
    const long n = (long)nzp * nxp;
    store *AU = (store *)calloc(n, sizeof(store)), *BU = (store *)calloc(n, sizeof(store));
    store *HU = (store *)calloc(n, sizeof(store));
    store *QX = (store *)calloc(n, sizeof(store)), *QZ = (store *)calloc(n, sizeof(store));
    const real dt2 = R(dt) * R(dt);
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real pxx = d2(P, c, 1) / (R(dx) * R(dx)), pzz = d2(P, c, nxp) / (R(dz) * R(dz));
            real px = d1(P, c, 1) / R(dx), pz = d1(P, c, nxp) / R(dz);
            real e = eps[c], dl = del[c], th = theta[c];
            real cs = RCOS(th), sn = RSIN(th);
            real xi = px * cs - pz * sn, eta = px * sn + pz * cs;
            real xi2 = xi * xi, eta2 = eta * eta, xi4 = xi2 * xi2, eta4 = eta2 * eta2;
            real num = R(-2.0) * (e - dl) * xi2 * eta2;
            real den = (R(1.0) + R(2.0) * e) * xi4 + eta4 + R(2.0) * (R(1.0) + dl) * xi2 * eta2;
            real Sd = 0, Cx = 0, Cz = 0;
            if (RFABS(den) > DEN_EPS) {
                Sd = num / den;
                real K = (R(1.0) + R(2.0) * e) * xi4 - eta4;
                real fac = R(4.0) * (e - dl) * xi * eta * K / (den * den);
                Cx = fac * pz;
                Cz = -fac * px;
            }
            real u = Uc[c];
            AU[c] = (store)(((R(1.0) + R(2.0) * e) * (cs * cs) + (sn * sn) + Sd) * u);
            BU[c] = (store)(((R(1.0) + R(2.0) * e) * (sn * sn) + (cs * cs) + Sd) * u);
            HU[c] = (store)((R(2.0) * e * RSIN(R(2.0) * th)) * u);
            QX[c] = (store)((pxx + pzz) * Cx * u);
            QZ[c] = (store)((pxx + pzz) * Cz * u);
        }
    OMP_ROWS((long)(nzp) * (nxp))
    for (int j = 4; j < nzp - 4; ++j)
        for (int i = 4; i < nxp - 4; ++i) {
            long c = (long)j * nxp + i;
            real op = d2(AU, c, 1) / (R(dx) * R(dx)) + d2(BU, c, nxp) / (R(dz) * R(dz)) -
                      dxz(HU, c, nxp) / (R(dz) * R(dx)) - d1(QX, c, 1) / R(dx) -
                      d1(QZ, c, nxp) / R(dz);
            Uf[c] = (store)(R(2.0) * Uc[c] - Uf[c] + (real)vp[c] * vp[c] * dt2 * op);
        }
    free(AU);
    free(BU);
    free(HU);
    free(QX);
    free(QZ);
}

/* ------------------------------------------------------ imaging / gradients */

/* Multiparameter VTI partials on interior-only [nz][nx] arrays,
 * src/CudaKernels.py:1174-1250 / src/utils.py:1309-1367.  float64 arithmetic
 * in every build, as in the reference kernel. */
void or_grad_vti(const store *P, const store *adj, store *geps, store *gdel, int nz, int nx, double dx,
                 double dz, const store *eps, const store *del) {
    OMP_ROWS((long)(nz) * (nx))
    for (int j = 4; j < nz - 4; ++j)
        for (int i = 4; i < nx - 4; ++i) {
            long c = (long)j * nx + i;
            double pxx = (double)(d2(P, c, 1) / (R(dx) * R(dx)));
            double pzz = (double)(d2(P, c, nx) / (R(dz) * R(dz)));
            double px = (double)(d1(P, c, 1) / R(dx)), pz = (double)(d1(P, c, nx) / R(dz));
            double e = eps[c], dl = del[c];
            double num = -2.0 * (e - dl) * (px * px) * (pz * pz);
            double den = (1.0 + 2.0 * e) * (px * px * px * px) + (pz * pz * pz * pz) +
                         2.0 * (1.0 + dl) * (px * px) * (pz * pz);
            double dn_e = -2.0 * px * px * pz * pz, dn_d = 2.0 * px * px * pz * pz;
            double dd_e = 2.0 * px * px * px * px, dd_d = 2.0 * px * px * pz * pz;
            double sde = 0, sdd = 0;
            if (fabs(den) > 1.0e-150) {
                sde = (dn_e * den - num * dd_e) / (den * den);
                sdd = (dn_d * den - num * dd_d) / (den * den);
            }
            double dPe = (-2.0 - sde) * pxx - sde * pzz;
            double dPd = -sdd * (pxx + pzz);
            geps[c] = (store)((double)geps[c] + (double)adj[c] * dPe);
            gdel[c] = (store)((double)gdel[c] + (double)adj[c] * dPd);
        }
}

/* Multiparameter TTI partials, src/CudaKernels.py:1252-1379 / src/utils.py:1369-1463. */
void or_grad_tti(const store *P, const store *adj, store *geps, store *gdel, store *gth, int nz,
                 int nx, double dx, double dz, const store *eps, const store *del,
                 const store *theta) {
    OMP_ROWS((long)(nz) * (nx))
    for (int j = 4; j < nz - 4; ++j)
        for (int i = 4; i < nx - 4; ++i) {
            long c = (long)j * nx + i;
            double pxx = (double)(d2(P, c, 1) / (R(dx) * R(dx)));
            double pzz = (double)(d2(P, c, nx) / (R(dz) * R(dz)));
            double pxz = (double)(dxz(P, c, nx) / (R(dz) * R(dx)));
            double px = (double)(d1(P, c, 1) / R(dx)), pz = (double)(d1(P, c, nx) / R(dz));
            double e = eps[c], dl = del[c], th = theta[c];
            double cs = cos(th), sn = sin(th);
            double h = px * cs - pz * sn, q = px * sn + pz * cs;
            double h2 = h * h, q2 = q * q;
            double num = -2.0 * (e - dl) * h2 * q2;
            double den = (1.0 + 2.0 * e) * h2 * h2 + q2 * q2 + 2.0 * (1.0 + dl) * h2 * q2;
            double sde = 0, sdd = 0, sdt = 0;
            if (fabs(den) >= 1.0e-150) {
                double dn_e = -2.0 * h2 * q2, dn_d = 2.0 * h2 * q2;
                double dd_e = 2.0 * h2 * h2, dd_d = 2.0 * h2 * q2;
                double dn_t = -2.0 * (e - dl) * (-2.0 * h * q * q * q + 2.0 * h * h * h * q);
                double dd_t = -4.0 * (1.0 + 2.0 * e) * h * h * h * q + 4.0 * h * q * q * q +
                              2.0 * (1.0 + dl) * (-2.0 * h * q * q * q + 2.0 * h * h * h * q);
                sde = (dn_e * den - num * dd_e) / (den * den);
                sdd = (dn_d * den - num * dd_d) / (den * den);
                sdt = (dn_t * den - num * dd_t) / (den * den);
            }
            double s2t = sin(2.0 * th), c2t = cos(2.0 * th);
            double dA = 2.0 * e * s2t - sdt, dB = -2.0 * e * s2t - sdt, dC = 4.0 * e * c2t;
            double dPe = (-(2.0 * cs * cs + sde)) * pxx - (2.0 * sn * sn + sde) * pzz + 2.0 * s2t * pxz;
            double dPd = -sdd * (pxx + pzz);
            double dPt = dA * pxx + dB * pzz + dC * pxz;
            geps[c] = (store)((double)geps[c] + (double)adj[c] * dPe);
            gdel[c] = (store)((double)gdel[c] + (double)adj[c] * dPd);
            gth[c] = (store)((double)gth[c] + (double)adj[c] * dPt);
        }
}

/* ------------------------------------------------------- whole steps */

/* One forward step in the reference order (src/Modeling2D.py:363-386):
 * inject into Uc, update Uf (psi, zeta, update for CPML; update, then damping
 * of Uf and Uc for Cerjan).  phys: 0 acoustic, 1 VTI, 2 TTI; abc: 0 Cerjan,
 * 1 CPML.  cp = {ax,bx,az,bz,psix,psiz,zetax,zetaz} for CPML. */
void or_forward_step(int phys, int abc, store *Uf, store *Uc, const store *vp, const store *eps,
                     const store *del, const store *theta, const store *A, store **cp, int nzp, int nxp,
                     int N, double dz, double dx, double dt, int sz, int sx, double srcval) {
    if (sz >= 0) or_inject(Uc, nxp, sz, sx, srcval);
    if (abc == 0) {
        if (phys == 0) or_update_acoustic(Uf, Uc, vp, nzp, nxp, dz, dx, dt);
        else if (phys == 1) or_update_vti(Uf, Uc, vp, eps, del, nzp, nxp, dz, dx, dt);
        else or_update_tti(Uf, Uc, vp, eps, del, theta, nzp, nxp, dz, dx, dt);
        or_cerjan_apply(Uf, nzp, nxp, N, A);
        or_cerjan_apply(Uc, nzp, nxp, N, A);
    } else {
        or_cpml_psi(cp[4], cp[5], Uc, cp[0], cp[1], cp[2], cp[3], nzp, nxp, N, dx, dz);
        or_cpml_zeta(cp[6], cp[7], cp[4], cp[5], Uc, cp[0], cp[1], cp[2], cp[3], nzp, nxp, N, dx, dz);
        if (phys == 0)
            or_update_acoustic_cpml(Uf, Uc, vp, cp[4], cp[5], cp[6], cp[7], nzp, nxp, N, dz, dx, dt);
        else if (phys == 1)
            or_update_vti_cpml(Uf, Uc, vp, eps, del, cp[4], cp[5], cp[6], cp[7], nzp, nxp, N, dz, dx,
                               dt);
        else
            or_update_tti_cpml(Uf, Uc, vp, eps, del, theta, cp[4], cp[5], cp[6], cp[7], nzp, nxp, N,
                               dz, dx, dt);
    }
}

/* Run a whole forward shot (src/Modeling2D.py:446-460 / cpp/Modeling.cu:441-468):
 * for k < nt: step, record seis[k-itlag] = Uc[rec] for k >= itlag, swap.
 * U0/U1 must be zeroed by the caller (and cp[4..7] for CPML).  If save != NULL,
 * save[k] receives the padded current field after step k (save-all). */
void or_forward_shot(int phys, int abc, store *U0, store *U1, const store *vp, const store *eps,
                     const store *del, const store *theta, const store *A, store **cp, int nzp, int nxp,
                     int N, double dz, double dx, double dt, int sz, int sx, const store *src, int nt,
                     int itlag, int nrec, const int *rz, const int *rx, store *seis, store *save) {
    store *cur = U0, *fut = U1;
    const long n = (long)nzp * nxp;
    for (int k = 0; k < nt; ++k) {
        or_forward_step(phys, abc, fut, cur, vp, eps, del, theta, A, cp, nzp, nxp, N, dz, dx, dt, sz,
                        sx, (double)src[k]);
        if (k >= itlag && seis) or_gather(cur, nxp, nrec, rz, rx, seis + (long)(k - itlag) * nrec);
        if (save) memcpy(save + (long)k * n, cur, n * sizeof(store));
        store *t = cur;
        cur = fut;
        fut = t;
    }
}
