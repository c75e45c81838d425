/*
 * CPU restatement of fteikpy 2.4.0's numeric kernels -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: the independent checker that the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product path (fteikpy_amd) never calls it and has no CPU fallback.
 *
 * Floating-point contract: every expression keeps the reference's evaluation order
 * (Python left-to-right), `x ** 2.0` is x*x, `x ** 0.5` is sqrt(x), and the file is
 * compiled with -ffp-contract=off so no FMA is formed.  This is the pure-Python
 * semantics the reference's CI runs under (NUMBA_DISABLE_JIT=1, .github/workflows/ci.yml:29)
 * up to libm pow() vs x*x/sqrt rounding (ULP level).
 *
 * Pinning (what checks each function):
 *   fteik3d / sweep3d / sweep (3D)   tests/golden vectors written by the reference itself
 *                                    (pure-Python mode), reference test sums
 *                                    (tests/test_eikonal3d.py:13-15, test_meshio.py:36-49),
 *                                    closed forms (axis / diagonal nodes, homogeneous).
 *   fteik2d / sweep2d / delta        golden vectors + test_eikonal2d.py:13-15, test_meshio.py:9-20.
 *   post-pass gradient (2D/3D)       golden vectors + test_gradient{2,3}d.py values.
 *   interp2d/3d, vinterp2d/3d        golden vectors + test_traveltime{2,3}d.py values.
 *   ray2d/3d + shrink                golden vectors + test_traveltime{2,3}d.py ray sums,
 *                                    test_meshio.py ray lengths.
 *
 * Array conventions follow the reference: slowness is cell based (nz, nx[, ny]) and C
 * ordered; traveltime is node based (nz+1, nx+1[, ny+1]).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BIG 1.0e5
#define EPS 1.0e-15
#define EPSIN 5

/* Python's builtin min/max keep the first argument unless a later one is strictly
 * smaller/larger. */
static inline double pmin2(double a, double b) { return (b < a) ? b : a; }
static inline double pmax2(double a, double b) { return (b > a) ? b : a; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------------------ */
/* 3D solver: fteikpy/_fteik/_fteik3d.py                                                */
/* ------------------------------------------------------------------------------------ */

/* _fteik3d.py:10-17 */
static double t_ana3(int i, int j, int k, double dz, double dx, double dy, double zsa,
                     double xsa, double ysa, double vzero) {
    double a = dz * ((double)i - zsa);
    double b = dx * ((double)j - xsa);
    double c = dy * ((double)k - ysa);
    return vzero * sqrt(a * a + b * b + c * c);
}

/* _fteik3d.py:20-35 */
static void t_anad3(int i, int j, int k, double dz, double dx, double dy, double zsa,
                    double xsa, double ysa, double vzero, double *t, double *tzc,
                    double *txc, double *tyc) {
    double tt = t_ana3(i, j, k, dz, dx, dy, zsa, xsa, ysa, vzero);
    if (tt > 0.0) {
        double tmp = vzero * vzero / tt;
        *tzc = ((double)i - zsa) * dz * tmp;
        *txc = ((double)j - xsa) * dx * tmp;
        *tyc = ((double)k - ysa) * dy * tmp;
    } else {
        *tzc = 0.0;
        *txc = 0.0;
        *tyc = 0.0;
    }
    *t = tt;
}

typedef struct {
    double dz, dx, dy, dz2i, dx2i, dy2i, dz2dx2, dz2dy2, dx2dy2, dsum;
} dargs3_t;

/* Candidates of the node update, _fteik3d.py:38-164: returns min(t0, t1d, t2d, t3d) and,
 * in *which, the winning candidate in the reference's priority order (:166-200) as a set of
 * axes (bit 0 z, bit 1 x, bit 2 y).  nz, nx, ny are NODE counts here as in the reference. */
static inline double cand3(const double *tt, const double *slow, const dargs3_t *d, int i,
                           int j, int k, int sgnvz, int sgnvx, int sgnvy, int sgntz, int sgntx,
                           int sgnty, int nz, int nx, int ny, double t0, int *which) {
#define TT(a, b, c) tt[((size_t)(a) * nx + (b)) * ny + (c)]
#define SL(a, b, c) slow[((size_t)(a) * (nx - 1) + (b)) * (ny - 1) + (c)]
    const double dz = d->dz, dx = d->dx, dy = d->dy;
    const double dz2i = d->dz2i, dx2i = d->dx2i, dy2i = d->dy2i;
    int i1 = i - sgnvz, j1 = j - sgnvx, k1 = k - sgnvy;

    double tv = TT(i - sgntz, j, k);
    double te = TT(i, j - sgntx, k);
    double tn = TT(i, j, k - sgnty);
    double tev = TT(i - sgntz, j - sgntx, k);
    double ten = TT(i, j - sgntx, k - sgnty);
    double tnv = TT(i - sgntz, j, k - sgnty);
    double tnve = TT(i - sgntz, j - sgntx, k - sgnty);

    int jm = imax(j - 1, 0), jp = imin(j, nx - 2);
    int km = imax(k - 1, 0), kp = imin(k, ny - 2);
    int im = imax(i - 1, 0), ip = imin(i, nz - 2);

    /* 1D operators, :78-104 */
    double vref = pmin2(pmin2(pmin2(SL(i1, jm, km), SL(i1, jm, kp)), SL(i1, jp, km)), SL(i1, jp, kp));
    double t1d1 = tv + dz * vref;
    vref = pmin2(pmin2(pmin2(SL(im, j1, km), SL(ip, j1, km)), SL(im, j1, kp)), SL(ip, j1, kp));
    double t1d2 = te + dx * vref;
    vref = pmin2(pmin2(pmin2(SL(im, jm, k1), SL(im, jp, k1)), SL(ip, jm, k1)), SL(ip, jp, k1));
    double t1d3 = tn + dy * vref;
    double t1d = pmin2(pmin2(t1d1, t1d2), t1d3);

    /* 2D operators, :107-143 */
    double t2d1 = BIG;
    vref = pmin2(SL(i1, j1, km), SL(i1, j1, kp));
    if (tv < te + dx * vref && te < tv + dz * vref) {
        double ta = tev + te - tv;
        double tb = tev - te + tv;
        double q = ta - tb;
        t2d1 = ((tb * dz2i + ta * dx2i) +
                sqrt(4.0 * (vref * vref) * (dz2i + dx2i) - dz2i * dx2i * (q * q))) /
               (dz2i + dx2i);
    }
    double t2d2 = BIG;
    vref = pmin2(SL(i1, jm, k1), SL(i1, jp, k1));
    if (tv < tn + dy * vref && tn < tv + dz * vref) {
        double ta = tv - tn + tnv;
        double tb = tn - tv + tnv;
        double q = ta - tb;
        t2d2 = ((ta * dz2i + tb * dy2i) +
                sqrt(4.0 * (vref * vref) * (dz2i + dy2i) - dz2i * dy2i * (q * q))) /
               (dz2i + dy2i);
    }
    double t2d3 = BIG;
    vref = pmin2(SL(im, j1, k1), SL(ip, j1, k1));
    if (te < tn + dy * vref && tn < te + dx * vref) {
        double ta = te - tn + ten;
        double tb = tn - te + ten;
        double q = ta - tb;
        t2d3 = ((ta * dx2i + tb * dy2i) +
                sqrt(4.0 * (vref * vref) * (dx2i + dy2i) - dx2i * dy2i * (q * q))) /
               (dx2i + dy2i);
    }
    double t2d = pmin2(pmin2(t2d1, t2d2), t2d3);

    /* 3D operator, :146-159 */
    double t3d = BIG;
    if (pmin2(t1d, t2d) > pmax2(pmax2(tv, te), tn)) {
        vref = SL(i1, j1, k1);
        double ta = te - 0.5 * tn + 0.5 * ten - 0.5 * tv + 0.5 * tev - tnv + tnve;
        double tb = tv - 0.5 * tn + 0.5 * tnv - 0.5 * te + 0.5 * tev - ten + tnve;
        double tc = tn - 0.5 * te + 0.5 * ten - 0.5 * tv + 0.5 * tnv - tev + tnve;
        double t2 = vref * vref * d->dsum * 9.0;
        double q1 = ta - tb, q2 = tb - tc, q3 = ta - tc;
        double t3 = d->dz2dx2 * (q1 * q1);
        t3 += d->dz2dy2 * (q2 * q2);
        t3 += d->dx2dy2 * (q3 * q3);
        if (t2 >= t3) {
            double t1 = tb * dz2i + ta * dx2i + tc * dy2i;
            t3d = (t1 + sqrt(t2 - t3)) / d->dsum;
        }
    }

    /* :162-200 */
    double tnew = pmin2(pmin2(pmin2(t0, t1d), t2d), t3d);
    if (tnew == t1d1) *which = 1;
    else if (tnew == t1d2) *which = 2;
    else if (tnew == t1d3) *which = 4;
    else if (tnew == t2d1) *which = 3;
    else if (tnew == t2d2) *which = 5;
    else if (tnew == t2d3) *which = 6;
    else *which = 7;
    return tnew;
#undef TT
#undef SL
}

/* Node update, _fteik3d.py:38-200. */
static inline void sweep_node3(double *tt, int32_t *sgn, const double *slow, const dargs3_t *d,
                               int i, int j, int k, int sgnvz, int sgnvx, int sgnvy,
                               int sgntz, int sgntx, int sgnty, int nz, int nx, int ny,
                               int grad) {
    size_t n = ((size_t)i * nx + j) * ny + k;
    double t0 = tt[n];
    int which;
    double tnew = cand3(tt, slow, d, i, j, k, sgnvz, sgnvx, sgnvy, sgntz, sgntx, sgnty, nz, nx,
                        ny, t0, &which);
    tt[n] = tnew;
    if (grad && tnew != t0) {
        int32_t *s = sgn + n * 3;
        s[0] = (which & 1) ? sgntz : 0;
        s[1] = (which & 2) ? sgntx : 0;
        s[2] = (which & 4) ? sgnty : 0;
    }
}

/* One sweeping set, _fteik3d.py:203-391.  `order` 0 = the reference's k,j,i nesting,
 * 1 = i,j,k nesting (same DAG, different topological order: used by the order-invariance
 * test). */
static void sweep3d(double *tt, int32_t *sgn, const double *slow, const dargs3_t *d, int nz,
                    int nx, int ny, int grad, int order) {
    /* (sgnvz, sgnvx, sgnvy) and (sgntz, sgntx, sgnty) for the 8 passes in the
     * reference's order. */
    static const int dirs[8][3] = {{1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {1, -1, -1},
                                   {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    for (int p = 0; p < 8; ++p) {
        int sz = dirs[p][0], sx = dirs[p][1], sy = dirs[p][2];
        int vz = sz > 0, vx = sx > 0, vy = sy > 0;
        /* ascending: range(1, n); descending: range(n - 2, -1, -1) */
        int i0 = sz > 0 ? 1 : nz - 2, ie = sz > 0 ? nz : -1;
        int j0 = sx > 0 ? 1 : nx - 2, je = sx > 0 ? nx : -1;
        int k0 = sy > 0 ? 1 : ny - 2, ke = sy > 0 ? ny : -1;
        if (order == 0) {
            for (int k = k0; k != ke; k += sy)
                for (int j = j0; j != je; j += sx)
                    for (int i = i0; i != ie; i += sz)
                        sweep_node3(tt, sgn, slow, d, i, j, k, vz, vx, vy, sz, sx, sy, nz, nx,
                                    ny, grad);
        } else {
            for (int i = i0; i != ie; i += sz)
                for (int j = j0; j != je; j += sx)
                    for (int k = k0; k != ke; k += sy)
                        sweep_node3(tt, sgn, slow, d, i, j, k, vz, vx, vy, sz, sx, sy, nz, nx,
                                    ny, grad);
        }
    }
}

/*
 * fteik3d, _fteik3d.py:394-490.
 * nz, nx, ny: CELL counts (shape of slow).  tt: (nz+1)(nx+1)(ny+1).  ttgrad: same x3 or NULL.
 * nsweep > 0: run exactly nsweep sets.  nsweep <= 0: run sets until one changes no node
 * (at most max_sets); *sets_done receives the number of sets executed, the final
 * no-change set included.  Returns 0, or 1 for "source out of bound".
 */
int oracle_fteik3d(const double *slow, int nz, int nx, int ny, double dz, double dx, double dy,
                   double zsrc, double xsrc, double ysrc, int nsweep, int max_sets, int grad,
                   int order, double *tt, double *ttgrad, double *vzero_out, int *sets_done) {
    int condz = 0.0 <= zsrc && zsrc <= dz * nz;
    int condx = 0.0 <= xsrc && xsrc <= dx * nx;
    int condy = 0.0 <= ysrc && ysrc <= dy * ny;
    if (!(condz && condx && condy)) return 1;

    double zsa = zsrc / dz, xsa = xsrc / dx, ysa = ysrc / dy;
    if (zsa >= nz) zsa = zsa - EPS;
    if (xsa >= nx) xsa = xsa - EPS;
    if (ysa >= ny) ysa = ysa - EPS;
    int zsi = imin((int)zsa, nz - 1), xsi = imin((int)xsa, nx - 1), ysi = imin((int)ysa, ny - 1);
    double vzero = slow[((size_t)zsi * nx + xsi) * ny + ysi];

    int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    size_t nn = (size_t)NZ * NX * NY;
    for (size_t n = 0; n < nn; ++n) tt[n] = BIG;
    int32_t *sgn = NULL;
    if (grad) {
        sgn = (int32_t *)calloc(nn * 3, sizeof(int32_t));
        memset(ttgrad, 0, nn * 3 * sizeof(double));
    }

    static const int corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                      {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (int c = 0; c < 8; ++c) {
        int i = zsi + corners[c][0], j = xsi + corners[c][1], k = ysi + corners[c][2];
        double t, tzc, txc, tyc;
        t_anad3(i, j, k, dz, dx, dy, zsa, xsa, ysa, vzero, &t, &tzc, &txc, &tyc);
        size_t n = ((size_t)i * NX + j) * NY + k;
        tt[n] = t;
        if (grad) {
            ttgrad[3 * n + 0] = tzc;
            ttgrad[3 * n + 1] = txc;
            ttgrad[3 * n + 2] = tyc;
        }
    }

    dargs3_t d;
    d.dz = dz; d.dx = dx; d.dy = dy;
    d.dz2i = 1.0 / dz / dz;
    d.dx2i = 1.0 / dx / dx;
    d.dy2i = 1.0 / dy / dy;
    d.dz2dx2 = d.dz2i * d.dx2i;
    d.dz2dy2 = d.dz2i * d.dy2i;
    d.dx2dy2 = d.dx2i * d.dy2i;
    d.dsum = d.dz2i + d.dx2i + d.dy2i;

    int done = 0;
    if (nsweep > 0) {
        for (int s = 0; s < nsweep; ++s) sweep3d(tt, sgn, slow, &d, NZ, NX, NY, grad, order);
        done = nsweep;
    } else {
        double *prev = (double *)malloc(nn * sizeof(double));
        for (int s = 0; s < max_sets; ++s) {
            memcpy(prev, tt, nn * sizeof(double));
            sweep3d(tt, sgn, slow, &d, NZ, NX, NY, grad, order);
            ++done;
            if (memcmp(prev, tt, nn * sizeof(double)) == 0) break;
        }
        free(prev);
    }
    if (sets_done) *sets_done = done;

    if (grad) {
        /* :464-488 */
        for (int i = 0; i < NZ; ++i)
            for (int j = 0; j < NX; ++j)
                for (int k = 0; k < NY; ++k) {
                    size_t n = ((size_t)i * NX + j) * NY + k;
                    int sz = sgn[3 * n], sx = sgn[3 * n + 1], sy = sgn[3 * n + 2];
                    double *g = ttgrad + 3 * n;
                    if (sz != 0) g[0] = sz * (tt[n] - tt[((size_t)(i - sz) * NX + j) * NY + k]) / dz;
                    if (sx != 0) g[1] = sx * (tt[n] - tt[((size_t)i * NX + (j - sx)) * NY + k]) / dx;
                    if (sy != 0) g[2] = sy * (tt[n] - tt[((size_t)i * NX + j) * NY + (k - sy)]) / dy;
                    double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                    if (gn > 0.0) {
                        g[0] /= gn;
                        g[1] /= gn;
                        g[2] /= gn;
                    }
                }
        free(sgn);
    }
    *vzero_out = vzero;
    return 0;
}

/* fteik3d_vectorized, _fteik3d.py:493-513: OpenMP over sources mirrors numba prange.
 * src is [nsrc,3]; outputs are stacked per source.  Returns the first nonzero status. */
int oracle_fteik3d_batch(const double *slow, int nz, int nx, int ny, double dz, double dx,
                         double dy, const double *src, int nsrc, int nsweep, int max_sets,
                         int grad, int order, int nthreads, double *tt, double *ttgrad,
                         double *vzero, int *sets_done) {
    size_t nn = (size_t)(nz + 1) * (nx + 1) * (ny + 1);
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int s = 0; s < nsrc; ++s) {
        int st = oracle_fteik3d(slow, nz, nx, ny, dz, dx, dy, src[3 * s], src[3 * s + 1],
                                src[3 * s + 2], nsweep, max_sets, grad, order, tt + nn * s,
                                grad ? ttgrad + 3 * nn * s : NULL, vzero + s,
                                sets_done ? sets_done + s : NULL);
        if (st) {
#ifdef _OPENMP
#pragma omp critical
#endif
            status = st;
        }
    }
    return status;
}

/* ------------------------------------------------------------------------------------ */
/* 2D solver: fteikpy/_fteik/_fteik2d.py                                                */
/* ------------------------------------------------------------------------------------ */

/* _fteik2d.py:11-14 */
static double t_ana2(int i, int j, double dz, double dx, double zsa, double xsa, double vzero) {
    double a = dz * ((double)i - zsa);
    double b = dx * ((double)j - xsa);
    return vzero * sqrt(a * a + b * b);
}

/* _fteik2d.py:17-30 */
static void t_anad2(int i, int j, double dz, double dx, double zsa, double xsa, double vzero,
                    double *t, double *tzc, double *txc) {
    double tt = t_ana2(i, j, dz, dx, zsa, xsa, vzero);
    if (tt > 0.0) {
        double tmp = vzero * vzero / tt;
        *tzc = ((double)i - zsa) * dz * tmp;
        *txc = ((double)j - xsa) * dx * tmp;
    } else {
        *tzc = 0.0;
        *txc = 0.0;
    }
    *t = tt;
}

/* _fteik2d.py:33-67 */
static double delta2(double t1, double tauv, double taue, double tauev, double t0c, double tzc,
                     double txc, double dzi, double dxi, double dz2i, double dx2i, double vzero,
                     double vref, int sgntz, int sgntx) {
    double ta = tauev + taue - tauv;
    double tb = tauev - taue + tauv;
    double apoly = dz2i + dx2i;
    double bpoly = 4.0 * ((double)sgntx * txc * dxi + (double)sgntz * tzc * dzi) -
                   2.0 * (ta * dx2i + tb * dz2i);
    double cpoly = ((ta * ta) * dx2i) + ((tb * tb) * dz2i) -
                   4.0 * ((double)sgntx * txc * dxi * ta + (double)sgntz * tzc * dzi * tb) +
                   4.0 * (vzero * vzero - vref * vref);
    double dpoly = bpoly * bpoly - 4.0 * apoly * cpoly;
    return dpoly >= 0.0 ? 0.5 * (sqrt(dpoly) - bpoly) / apoly + t0c : t1;
}

typedef struct {
    double dz, dx, dzi, dxi, dz2i, dx2i;
} dargs2_t;

/* Node update, _fteik2d.py:70-187 (nz, nx are node counts). */
static inline void sweep_node2(double *tt, int32_t *sgn, const double *slow, const dargs2_t *d,
                               int zsi, int xsi, double zsa, double xsa, double vzero, int i,
                               int j, int sgnvz, int sgnvx, int sgntz, int sgntx, int nz, int nx,
                               int grad) {
#define TT(a, b) tt[(size_t)(a) * nx + (b)]
#define SL(a, b) slow[(size_t)(a) * (nx - 1) + (b)]
    const double dz = d->dz, dx = d->dx;
    int i1 = i - sgnvz, j1 = j - sgnvx;
    double tv = TT(i - sgntz, j);
    double te = TT(i, j - sgntx);
    double tev = TT(i - sgntz, j - sgntx);

    double vref = pmin2(SL(i1, imax(j - 1, 0)), SL(i1, imin(j, nx - 2)));
    double t1d1 = tv + dz * vref;
    vref = pmin2(SL(imax(i - 1, 0), j1), SL(imin(i, nz - 2), j1));
    double t1d2 = te + dx * vref;
    double t1d = pmin2(t1d1, t1d2);

    double t2d = BIG;
    vref = SL(i1, j1);
    if (abs(i - zsi) > EPSIN || abs(j - xsi) > EPSIN) {
        if (tv <= te + dx * vref && te <= tv + dz * vref && te >= tev && tv >= tev) {
            double ta = tev + te - tv;
            double tb = tev - te + tv;
            double q = ta - tb;
            t2d = ((tb * d->dz2i + ta * d->dx2i) +
                   sqrt(4.0 * (vref * vref) * (d->dz2i + d->dx2i) - d->dz2i * d->dx2i * (q * q))) /
                  (d->dz2i + d->dx2i);
        } else if (te - tev <= dz * dz * vref / sqrt(dx * dx + dz * dz) && te - tev > 0.0) {
            double r = (te - tev) / dz;
            t2d = te + dx * sqrt(vref * vref - r * r);
        } else if (tv - tev <= dx * dx * vref / sqrt(dx * dx + dz * dz) && tv - tev > 0.0) {
            double r = (tv - tev) / dx;
            t2d = tv + dz * sqrt(vref * vref - r * r);
        }
    } else {
        if (tv < te + dx * vref && te < tv + dz * vref && te >= tev && tv >= tev) {
            double t0c, tzc, txc;
            t_anad2(i, j, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
            double tauv = tv - t_ana2(i - sgntz, j, dz, dx, zsa, xsa, vzero);
            double taue = te - t_ana2(i, j - sgntx, dz, dx, zsa, xsa, vzero);
            double tauev = tev - t_ana2(i - sgntz, j - sgntx, dz, dx, zsa, xsa, vzero);
            t2d = delta2(t2d, tauv, taue, tauev, t0c, tzc, txc, d->dzi, d->dxi, d->dz2i, d->dx2i,
                         vzero, vref, sgntz, sgntx);
            if (t2d < tv || t2d < te) t2d = BIG;
        }
    }

    double t0 = TT(i, j);
    double tnew = pmin2(pmin2(t0, t1d), t2d);
    TT(i, j) = tnew;
    if (grad && tnew != t0) {
        int32_t *s = sgn + ((size_t)i * nx + j) * 2;
        if (tnew == t1d1) { s[0] = sgntz; s[1] = 0; }
        else if (tnew == t1d2) { s[0] = 0; s[1] = sgntx; }
        else { s[0] = sgntz; s[1] = sgntx; }
    }
#undef TT
#undef SL
}

/* _fteik2d.py:190-287 (column-interleaved 4-direction set; nz, nx are node counts).
 * order 0 is the reference's interleave; order 1 runs the four directional passes
 * one after the other (each pass over all columns), a different update order whose
 * converged fixed point is compared against the reference's in tests. */
static void sweep2d(double *tt, int32_t *sgn, const double *slow, const dargs2_t *d, int zsi,
                    int xsi, double zsa, double xsa, double vzero, int nz, int nx, int grad,
                    int order) {
    if (order == 0) {
        for (int j = 1; j < nx; ++j) {
            for (int i = 1; i < nz; ++i)
                sweep_node2(tt, sgn, slow, d, zsi, xsi, zsa, xsa, vzero, i, j, 1, 1, 1, 1, nz, nx, grad);
            for (int i = nz - 2; i >= 0; --i)
                sweep_node2(tt, sgn, slow, d, zsi, xsi, zsa, xsa, vzero, i, j, 0, 1, -1, 1, nz, nx, grad);
        }
        for (int j = nx - 2; j >= 0; --j) {
            for (int i = 1; i < nz; ++i)
                sweep_node2(tt, sgn, slow, d, zsi, xsi, zsa, xsa, vzero, i, j, 1, 0, 1, -1, nz, nx, grad);
            for (int i = nz - 2; i >= 0; --i)
                sweep_node2(tt, sgn, slow, d, zsi, xsi, zsa, xsa, vzero, i, j, 0, 0, -1, -1, nz, nx, grad);
        }
    } else {
        static const int dirs[4][2] = {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
        for (int p = 0; p < 4; ++p) {
            int sz = dirs[p][0], sx = dirs[p][1];
            int j0 = sx > 0 ? 1 : nx - 2, je = sx > 0 ? nx : -1;
            int i0 = sz > 0 ? 1 : nz - 2, ie = sz > 0 ? nz : -1;
            for (int j = j0; j != je; j += sx)
                for (int i = i0; i != ie; i += sz)
                    sweep_node2(tt, sgn, slow, d, zsi, xsi, zsa, xsa, vzero, i, j, sz > 0,
                                sx > 0, sz, sx, nz, nx, grad);
        }
    }
}

/*
 * fteik2d, _fteik2d.py:290-641.  nz, nx: CELL counts.  tt: (nz+1)(nx+1); ttgrad x2 or NULL.
 * nsweep/max_sets/sets_done as in oracle_fteik3d.  order: see sweep2d.
 */
int oracle_fteik2d(const double *slow, int nz, int nx, double dz, double dx, double zsrc,
                   double xsrc, int nsweep, int max_sets, int grad, int order, double *tt,
                   double *ttgrad, double *vzero_out, int *sets_done) {
    int condz = 0.0 <= zsrc && zsrc <= dz * nz;
    int condx = 0.0 <= xsrc && xsrc <= dx * nx;
    if (!(condz && condx)) return 1;

    double zsa = zsrc / dz, xsa = xsrc / dx;
    if (zsa >= nz) zsa = zsa - EPS;
    if (xsa >= nx) xsa = xsa - EPS;
    int zsi = imin((int)zsa, nz - 1), xsi = imin((int)xsa, nx - 1);
    double vzero = slow[(size_t)zsi * nx + xsi];

    int NZ = nz + 1, NX = nx + 1;
    size_t nn = (size_t)NZ * NX;
#define TT(a, b) tt[(size_t)(a) * NX + (b)]
#define SL(a, b) slow[(size_t)(a) * nx + (b)]
#define SG(a, b, c) sgn[((size_t)(a) * NX + (b)) * 2 + (c)]
    for (size_t n = 0; n < nn; ++n) tt[n] = BIG;
    int32_t *sgn = NULL;
    if (grad) {
        sgn = (int32_t *)calloc(nn * 2, sizeof(int32_t));
        memset(ttgrad, 0, nn * 2 * sizeof(double));
    }

    /* :329-352 */
    double dzu = fabs(zsa - (double)zsi);
    double dzd = 1.0 - dzu;
    double dxw = fabs(xsa - (double)xsi);
    double dxe = 1.0 - dxw;
    double dzv_min = pmin2(dzu, dzd);
    double dzh_min = pmin2(dxw, dxe);
    int iflag;
    if (dzv_min < EPS && dzh_min < EPS) {
        zsa = rint(zsa);
        xsa = rint(xsa);
        iflag = 1;
    } else if (dzv_min > EPS || dzh_min > EPS) {
        zsa = dzv_min < EPS ? rint(zsa) : zsa;
        xsa = dzh_min < EPS ? rint(xsa) : xsa;
        iflag = 2;
    } else {
        zsa = rint(zsa);
        xsa = rint(xsa);
        iflag = 3;
    }

    if (iflag == 2) {
        int ntd = NZ > NX ? NZ : NX;
        double *td = (double *)malloc(ntd * sizeof(double));
        for (int n = 0; n < ntd; ++n) td[n] = BIG;
        dzu = fabs(zsa - (double)zsi);
        dzd = 1.0 - dzu;
        dxw = fabs(xsa - (double)xsi);
        dxe = 1.0 - dxw;

        static const int corners[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
        for (int c = 0; c < 4; ++c) {
            int i = zsi + corners[c][0], j = xsi + corners[c][1];
            double t, tzc, txc;
            t_anad2(i, j, dz, dx, zsa, xsa, vzero, &t, &tzc, &txc);
            TT(i, j) = t;
            if (grad) {
                ttgrad[((size_t)i * NX + j) * 2] = tzc;
                ttgrad[((size_t)i * NX + j) * 2 + 1] = txc;
            }
        }

        double dxi = 1.0 / dx;
        double dx2i = dxi / dx;
        double dzi, dz2i, t0c, tzc, txc, tauv, taue, tauev, vref;
        /* :379-435, +x along rows zsi+1 and zsi */
        td[xsi + 1] = vzero * dxe * dx;
        for (int j = xsi + 2; j < NX; ++j) {
            vref = SL(zsi, j - 1);
            td[j] = td[j - 1] + dx * vref;
            tauv = td[j] - vzero * fabs((double)j - xsa) * dx;
            tauev = td[j - 1] - vzero * fabs((double)j - xsa - 1.0) * dx;
            dzi = 1.0 / dzd;
            dz2i = dz / dzd / dzd;
            taue = TT(zsi + 1, j - 1) - t_ana2(zsi + 1, j - 1, dz, dx, zsa, xsa, vzero);
            t_anad2(zsi + 1, j, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
            TT(zsi + 1, j) = delta2(TT(zsi + 1, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, 1, 1);
            if (grad) { SG(zsi + 1, j, 0) = 1; SG(zsi + 1, j, 1) = 1; }
            if (dzu > 0.0) {
                dzi = 1.0 / dzu;
                dz2i = dz / dzu / dzu;
                taue = TT(zsi, j - 1) - t_ana2(zsi, j - 1, dz, dx, zsa, xsa, vzero);
                t_anad2(zsi, j, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
                TT(zsi, j) = delta2(TT(zsi, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, -1, 1);
                if (grad) { SG(zsi, j, 0) = -1; SG(zsi, j, 1) = 1; }
            }
        }
        /* :437-495, -x */
        td[xsi] = vzero * dxw * dx;
        for (int j = xsi - 1; j >= 0; --j) {
            vref = SL(zsi, j);
            td[j] = td[j + 1] + dx * vref;
            tauv = td[j] - vzero * fabs((double)j - xsa) * dx;
            tauev = td[j + 1] - vzero * fabs((double)j - xsa + 1.0) * dx;
            dzi = 1.0 / dzd;
            dz2i = dz / dzd / dzd;
            taue = TT(zsi + 1, j + 1) - t_ana2(zsi + 1, j + 1, dz, dx, zsa, xsa, vzero);
            t_anad2(zsi + 1, j, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
            TT(zsi + 1, j) = delta2(TT(zsi + 1, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, 1, -1);
            if (grad) { SG(zsi + 1, j, 0) = 1; SG(zsi + 1, j, 1) = -1; }
            if (dzu > 0.0) {
                dzi = 1.0 / dzu;
                dz2i = dz / dzu / dzu;
                /* the reference reads row zsi+1 here (_fteik2d.py:472-474) */
                taue = TT(zsi + 1, j + 1) - t_ana2(zsi + 1, j + 1, dz, dx, zsa, xsa, vzero);
                t_anad2(zsi, j, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
                TT(zsi, j) = delta2(TT(zsi, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, -1, -1);
                if (grad) { SG(zsi, j, 0) = -1; SG(zsi, j, 1) = -1; }
            }
        }
        /* :497-556, +z along columns xsi+1 and xsi */
        dzi = 1.0 / dz;
        dz2i = dzi / dz;
        for (int n = 0; n < ntd; ++n) td[n] = BIG;
        td[zsi + 1] = vzero * dzd * dz;
        for (int i = zsi + 2; i < NZ; ++i) {
            vref = SL(i - 1, xsi);
            td[i] = td[i - 1] + dz * vref;
            taue = td[i] - vzero * fabs((double)i - zsa) * dz;
            tauev = td[i - 1] - vzero * fabs((double)i - zsa - 1.0) * dz;
            dxi = 1.0 / dxe;
            dx2i = dx / dxe / dxe;
            tauv = TT(i - 1, xsi + 1) - t_ana2(i - 1, xsi + 1, dz, dx, zsa, xsa, vzero);
            t_anad2(i, xsi + 1, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
            TT(i, xsi + 1) = delta2(TT(i, xsi + 1), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, 1, 1);
            if (grad) { SG(i, xsi + 1, 0) = 1; SG(i, xsi + 1, 1) = 1; }
            if (dxw > 0.0) {
                dxi = 1.0 / dxw;
                dx2i = dx / dxw / dxw;
                tauv = TT(i - 1, xsi) - t_ana2(i - 1, xsi, dz, dx, zsa, xsa, vzero);
                t_anad2(i, xsi, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
                TT(i, xsi) = delta2(TT(i, xsi), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, 1, -1);
                if (grad) { SG(i, xsi, 0) = 1; SG(i, xsi, 1) = -1; }
            }
        }
        /* :558-614, -z */
        td[zsi] = vzero * dzu * dz;
        for (int i = zsi - 1; i >= 0; --i) {
            vref = SL(i, xsi);
            td[i] = td[i + 1] + dz * vref;
            taue = td[i] - vzero * fabs((double)i - zsa) * dz;
            tauev = td[i + 1] - vzero * fabs((double)i - zsa + 1.0) * dz;
            dxi = 1.0 / dxe;
            dx2i = dx / dxe / dxe;
            tauv = TT(i + 1, xsi + 1) - t_ana2(i + 1, xsi + 1, dz, dx, zsa, xsa, vzero);
            t_anad2(i, xsi + 1, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
            TT(i, xsi + 1) = delta2(TT(i, xsi + 1), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, -1, 1);
            if (grad) { SG(i, xsi + 1, 0) = -1; SG(i, xsi + 1, 1) = 1; }
            if (dxw > 0.0) {
                dxi = 1.0 / dxw;
                dx2i = dx / dxw / dxw;
                tauv = TT(i + 1, xsi) - t_ana2(i + 1, xsi, dz, dx, zsa, xsa, vzero);
                t_anad2(i, xsi, dz, dx, zsa, xsa, vzero, &t0c, &tzc, &txc);
                TT(i, xsi) = delta2(TT(i, xsi), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, -1, -1);
                if (grad) { SG(i, xsi, 0) = -1; SG(i, xsi, 1) = -1; }
            }
        }
        free(td);
    } else {
        TT((int)zsa, (int)xsa) = 0.0;
    }

    dargs2_t d;
    d.dz = dz; d.dx = dx;
    d.dzi = 1.0 / dz;
    d.dxi = 1.0 / dx;
    d.dz2i = d.dzi / dz;
    d.dx2i = d.dxi / dx;

    int done = 0;
    if (nsweep > 0) {
        for (int s = 0; s < nsweep; ++s)
            sweep2d(tt, sgn, slow, &d, zsi, xsi, zsa, xsa, vzero, NZ, NX, grad, order);
        done = nsweep;
    } else {
        double *prev = (double *)malloc(nn * sizeof(double));
        for (int s = 0; s < max_sets; ++s) {
            memcpy(prev, tt, nn * sizeof(double));
            sweep2d(tt, sgn, slow, &d, zsi, xsi, zsa, xsa, vzero, NZ, NX, grad, order);
            ++done;
            if (memcmp(prev, tt, nn * sizeof(double)) == 0) break;
        }
        free(prev);
    }
    if (sets_done) *sets_done = done;

    if (grad) {
        for (int i = 0; i < NZ; ++i)
            for (int j = 0; j < NX; ++j) {
                size_t n = (size_t)i * NX + j;
                int sz = sgn[2 * n], sx = sgn[2 * n + 1];
                double *g = ttgrad + 2 * n;
                if (sz != 0) g[0] = sz * (tt[n] - TT(i - sz, j)) / dz;
                if (sx != 0) g[1] = sx * (tt[n] - TT(i, j - sx)) / dx;
                double gn = sqrt(g[0] * g[0] + g[1] * g[1]);
                if (gn > 0.0) {
                    g[0] /= gn;
                    g[1] /= gn;
                }
            }
        free(sgn);
    }
#undef TT
#undef SL
#undef SG
    *vzero_out = vzero;
    return 0;
}

int oracle_fteik2d_batch(const double *slow, int nz, int nx, double dz, double dx,
                         const double *src, int nsrc, int nsweep, int max_sets, int grad,
                         int order, int nthreads, double *tt, double *ttgrad, double *vzero,
                         int *sets_done) {
    size_t nn = (size_t)(nz + 1) * (nx + 1);
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int s = 0; s < nsrc; ++s) {
        int st = oracle_fteik2d(slow, nz, nx, dz, dx, src[2 * s], src[2 * s + 1], nsweep,
                                max_sets, grad, order, tt + nn * s,
                                grad ? ttgrad + 2 * nn * s : NULL, vzero + s,
                                sets_done ? sets_done + s : NULL);
        if (st) {
#ifdef _OPENMP
#pragma omp critical
#endif
            status = st;
        }
    }
    return status;
}

/* ------------------------------------------------------------------------------------ */
/* Interpolation: fteikpy/_interp/                                                      */
/* ------------------------------------------------------------------------------------ */

/* np.searchsorted(a, v, side="right") - 1 */
static int ssr(const double *a, int n, double v) {
    int lo = 0, hi = n; /* first index with a[idx] > v */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

static inline double dist3(double x1, double y1, double z1, double x2, double y2, double z2) {
    double a = x1 - x2, b = y1 - y2, c = z1 - z2;
    return sqrt(a * a + b * b + c * c);
}
static inline double dist2(double x1, double y1, double x2, double y2) {
    double a = x1 - x2, b = y1 - y2;
    return sqrt(a * a + b * b);
}

/* Corner geometry shared by interp3d / vinterp3d: _interp3d.py:21-162,
 * _vinterp3d.py:29-244.  The last node has a ghost neighbour at 2*x[-1]-x[-2]
 * carrying value 1 (and distance 0 for vinterp). */
typedef struct {
    double x1, x2, y1, y2, z1, z2;
    int ok2x, ok2y, ok2z; /* neighbour i2/j2/k2 is a real node */
} cell3_t;

static cell3_t cell3(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                     int i1, int j1, int k1) {
    cell3_t c;
    c.ok2x = i1 != nx - 1;
    c.ok2y = j1 != ny - 1;
    c.ok2z = k1 != nz - 1;
    c.x1 = x[i1];
    c.x2 = c.ok2x ? x[i1 + 1] : 2.0 * c.x1 - x[nx - 2];
    c.y1 = y[j1];
    c.y2 = c.ok2y ? y[j1 + 1] : 2.0 * c.y1 - y[ny - 2];
    c.z1 = z[k1];
    c.z2 = c.ok2z ? z[k1 + 1] : 2.0 * c.z1 - z[nz - 2];
    return c;
}

/* _interp3d.py:7-174; the axes are named x, y, z there and index v[x, y, z]. */
static double interp3_one(const double *x, int nx, const double *y, int ny, const double *z,
                          int nz, const double *v, double xq, double yq, double zq,
                          double fval) {
    if (!(x[0] <= xq && xq <= x[nx - 1] && y[0] <= yq && yq <= y[ny - 1] && z[0] <= zq &&
          zq <= z[nz - 1]))
        return fval;
    int i1 = ssr(x, nx, xq), j1 = ssr(y, ny, yq), k1 = ssr(z, nz, zq);
    cell3_t c = cell3(x, nx, y, ny, z, nz, i1, j1, k1);
#define V(a, b, cc) v[((size_t)(a) * ny + (b)) * nz + (cc)]
    double v111 = V(i1, j1, k1);
    double v211 = c.ok2x ? V(i1 + 1, j1, k1) : 1.0;
    double v121 = c.ok2y ? V(i1, j1 + 1, k1) : 1.0;
    double v221 = (c.ok2x && c.ok2y) ? V(i1 + 1, j1 + 1, k1) : 1.0;
    double v112 = c.ok2z ? V(i1, j1, k1 + 1) : 1.0;
    double v212 = (c.ok2x && c.ok2z) ? V(i1 + 1, j1, k1 + 1) : 1.0;
    double v122 = (c.ok2y && c.ok2z) ? V(i1, j1 + 1, k1 + 1) : 1.0;
    double v222 = (c.ok2x && c.ok2y && c.ok2z) ? V(i1 + 1, j1 + 1, k1 + 1) : 1.0;
#undef V
    double vq = v111 * fabs((c.x2 - xq) * (c.y2 - yq) * (c.z2 - zq));
    vq += v211 * fabs((c.x1 - xq) * (c.y2 - yq) * (c.z2 - zq));
    vq += v121 * fabs((c.x2 - xq) * (c.y1 - yq) * (c.z2 - zq));
    vq += v221 * fabs((c.x1 - xq) * (c.y1 - yq) * (c.z2 - zq));
    vq += v112 * fabs((c.x2 - xq) * (c.y2 - yq) * (c.z1 - zq));
    vq += v212 * fabs((c.x1 - xq) * (c.y2 - yq) * (c.z1 - zq));
    vq += v122 * fabs((c.x2 - xq) * (c.y1 - yq) * (c.z1 - zq));
    vq += v222 * fabs((c.x1 - xq) * (c.y1 - yq) * (c.z1 - zq));
    vq /= fabs((c.x2 - c.x1) * (c.y2 - c.y1) * (c.z2 - c.z1));
    return vq;
}

void oracle_interp3d(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                     const double *v, const double *q, int nq, double fval, double *out) {
    for (int n = 0; n < nq; ++n)
        out[n] = interp3_one(x, nx, y, ny, z, nz, v, q[3 * n], q[3 * n + 1], q[3 * n + 2], fval);
}

/* _vinterp3d.py:7-264 */
static double vinterp3_one(const double *x, int nx, const double *y, int ny, const double *z,
                           int nz, const double *v, double xq, double yq, double zq,
                           double xsrc, double ysrc, double zsrc, double vzero, double fval) {
    if (!(x[0] <= xq && xq <= x[nx - 1] && y[0] <= yq && yq <= y[ny - 1] && z[0] <= zq &&
          zq <= z[nz - 1]))
        return fval;
    int xsi = ssr(x, nx, xsrc), ysi = ssr(y, ny, ysrc), zsi = ssr(z, nz, zsrc);
    int i1 = ssr(x, nx, xq), j1 = ssr(y, ny, yq), k1 = ssr(z, nz, zq);
    if (xsi == i1 && ysi == j1 && zsi == k1) return vzero * dist3(xsrc, ysrc, zsrc, xq, yq, zq);
    cell3_t c = cell3(x, nx, y, ny, z, nz, i1, j1, k1);
#define V(a, b, cc) v[((size_t)(a) * ny + (b)) * nz + (cc)]
#define D(px, py, pz) dist3(xsrc, ysrc, zsrc, px, py, pz)
    double v111 = V(i1, j1, k1), d111 = D(c.x1, c.y1, c.z1);
    double v211 = 1.0, d211 = 0.0, v121 = 1.0, d121 = 0.0, v221 = 1.0, d221 = 0.0;
    double v112 = 1.0, d112 = 0.0, v212 = 1.0, d212 = 0.0, v122 = 1.0, d122 = 0.0;
    double v222 = 1.0, d222 = 0.0;
    if (c.ok2x) { v211 = V(i1 + 1, j1, k1); d211 = D(c.x2, c.y1, c.z1); }
    if (c.ok2y) { v121 = V(i1, j1 + 1, k1); d121 = D(c.x1, c.y2, c.z1); }
    if (c.ok2x && c.ok2y) { v221 = V(i1 + 1, j1 + 1, k1); d221 = D(c.x2, c.y2, c.z1); }
    if (c.ok2z) { v112 = V(i1, j1, k1 + 1); d112 = D(c.x1, c.y1, c.z2); }
    if (c.ok2x && c.ok2z) { v212 = V(i1 + 1, j1, k1 + 1); d212 = D(c.x2, c.y1, c.z2); }
    if (c.ok2y && c.ok2z) { v122 = V(i1, j1 + 1, k1 + 1); d122 = D(c.x1, c.y2, c.z2); }
    if (c.ok2x && c.ok2y && c.ok2z) { v222 = V(i1 + 1, j1 + 1, k1 + 1); d222 = D(c.x2, c.y2, c.z2); }
#undef V
#undef D
    if (!(v111 != 0.0 && v211 != 0.0 && v121 != 0.0 && v221 != 0.0 && v112 != 0.0 &&
          v212 != 0.0 && v122 != 0.0 && v222 != 0.0))
        return vzero * dist3(xsrc, ysrc, zsrc, xq, yq, zq);
    double vq = d111 / v111 * fabs((c.x2 - xq) * (c.y2 - yq) * (c.z2 - zq));
    vq += d211 / v211 * fabs((c.x1 - xq) * (c.y2 - yq) * (c.z2 - zq));
    vq += d121 / v121 * fabs((c.x2 - xq) * (c.y1 - yq) * (c.z2 - zq));
    vq += d221 / v221 * fabs((c.x1 - xq) * (c.y1 - yq) * (c.z2 - zq));
    vq += d112 / v112 * fabs((c.x2 - xq) * (c.y2 - yq) * (c.z1 - zq));
    vq += d212 / v212 * fabs((c.x1 - xq) * (c.y2 - yq) * (c.z1 - zq));
    vq += d122 / v122 * fabs((c.x2 - xq) * (c.y1 - yq) * (c.z1 - zq));
    vq += d222 / v222 * fabs((c.x1 - xq) * (c.y1 - yq) * (c.z1 - zq));
    vq /= fabs((c.x2 - c.x1) * (c.y2 - c.y1) * (c.z2 - c.z1));
    return dist3(xsrc, ysrc, zsrc, xq, yq, zq) / vq;
}

void oracle_vinterp3d(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                      const double *v, const double *q, int nq, const double *src, double vzero,
                      double fval, double *out) {
    for (int n = 0; n < nq; ++n)
        out[n] = vinterp3_one(x, nx, y, ny, z, nz, v, q[3 * n], q[3 * n + 1], q[3 * n + 2],
                              src[0], src[1], src[2], vzero, fval);
}

/* _interp2d.py:7-74 */
static double interp2_one(const double *x, int nx, const double *y, int ny, const double *v,
                          double xq, double yq, double fval) {
    if (!(x[0] <= xq && xq <= x[nx - 1] && y[0] <= yq && yq <= y[ny - 1])) return fval;
    int i1 = ssr(x, nx, xq), j1 = ssr(y, ny, yq);
    int okx = i1 != nx - 1, oky = j1 != ny - 1;
    double x1 = x[i1], x2 = okx ? x[i1 + 1] : 2.0 * x1 - x[nx - 2];
    double y1 = y[j1], y2 = oky ? y[j1 + 1] : 2.0 * y1 - y[ny - 2];
#define V(a, b) v[(size_t)(a) * ny + (b)]
    double v11 = V(i1, j1);
    double v21 = okx ? V(i1 + 1, j1) : 1.0;
    double v12 = oky ? V(i1, j1 + 1) : 1.0;
    double v22 = (okx && oky) ? V(i1 + 1, j1 + 1) : 1.0;
#undef V
    double vq = v11 * fabs((x2 - xq) * (y2 - yq));
    vq += v21 * fabs((x1 - xq) * (y2 - yq));
    vq += v12 * fabs((x2 - xq) * (y1 - yq));
    vq += v22 * fabs((x1 - xq) * (y1 - yq));
    vq /= fabs((x2 - x1) * (y2 - y1));
    return vq;
}

void oracle_interp2d(const double *x, int nx, const double *y, int ny, const double *v,
                     const double *q, int nq, double fval, double *out) {
    for (int n = 0; n < nq; ++n) out[n] = interp2_one(x, nx, y, ny, v, q[2 * n], q[2 * n + 1], fval);
}

/* _vinterp2d.py:7-108 */
static double vinterp2_one(const double *x, int nx, const double *y, int ny, const double *v,
                           double xq, double yq, double xsrc, double ysrc, double vzero,
                           double fval) {
    if (!(x[0] <= xq && xq <= x[nx - 1] && y[0] <= yq && yq <= y[ny - 1])) return fval;
    int xsi = ssr(x, nx, xsrc), ysi = ssr(y, ny, ysrc);
    int i1 = ssr(x, nx, xq), j1 = ssr(y, ny, yq);
    if (xsi == i1 && ysi == j1) return vzero * dist2(xsrc, ysrc, xq, yq);
    int okx = i1 != nx - 1, oky = j1 != ny - 1;
    double x1 = x[i1], x2 = okx ? x[i1 + 1] : 2.0 * x1 - x[nx - 2];
    double y1 = y[j1], y2 = oky ? y[j1 + 1] : 2.0 * y1 - y[ny - 2];
#define V(a, b) v[(size_t)(a) * ny + (b)]
    double v11 = V(i1, j1), d11 = dist2(xsrc, ysrc, x1, y1);
    double v21 = 1.0, d21 = 0.0, v12 = 1.0, d12 = 0.0, v22 = 1.0, d22 = 0.0;
    if (okx) { v21 = V(i1 + 1, j1); d21 = dist2(xsrc, ysrc, x2, y1); }
    if (oky) { v12 = V(i1, j1 + 1); d12 = dist2(xsrc, ysrc, x1, y2); }
    if (okx && oky) { v22 = V(i1 + 1, j1 + 1); d22 = dist2(xsrc, ysrc, x2, y2); }
#undef V
    if (!(v11 != 0.0 && v21 != 0.0 && v12 != 0.0 && v22 != 0.0))
        return vzero * dist2(xsrc, ysrc, xq, yq);
    double vq = d11 / v11 * fabs((x2 - xq) * (y2 - yq));
    vq += d21 / v21 * fabs((x1 - xq) * (y2 - yq));
    vq += d12 / v12 * fabs((x2 - xq) * (y1 - yq));
    vq += d22 / v22 * fabs((x1 - xq) * (y1 - yq));
    vq /= fabs((x2 - x1) * (y2 - y1));
    return dist2(xsrc, ysrc, xq, yq) / vq;
}

void oracle_vinterp2d(const double *x, int nx, const double *y, int ny, const double *v,
                      const double *q, int nq, const double *src, double vzero, double fval,
                      double *out) {
    for (int n = 0; n < nq; ++n)
        out[n] = vinterp2_one(x, nx, y, ny, v, q[2 * n], q[2 * n + 1], src[0], src[1], vzero, fval);
}

/* ------------------------------------------------------------------------------------ */
/* Ray tracing: fteikpy/_fteik/_ray3d.py, _ray2d.py, _common.py                          */
/* ------------------------------------------------------------------------------------ */

/* shrink, _fteik/_common.py:4-28 (dimension-generic) */
static double shrink(const double *pcur, const double *delta, const double *lower,
                     const double *upper, int nd) {
    int condl = 0, condu = 0;
    double bl = INFINITY, bu = INFINITY;
    for (int a = 0; a < nd; ++a) {
        double tmp = pcur[a] - delta[a];
        if (tmp < lower[a]) {
            double b = (pcur[a] - lower[a]) / delta[a];
            if (!condl || b < bl) bl = b;
            condl = 1;
        }
        if (tmp > upper[a]) {
            double b = (pcur[a] - upper[a]) / delta[a];
            if (!condu || b < bu) bu = b;
            condu = 1;
        }
    }
    if (condl && condu) return pmin2(bl, bu);
    if (condl) return bl;
    if (condu) return bu;
    return 1.0;
}

/*
 * _ray3d.py:9-120.  Axes z, x, y (node axes), gradient component grids gz, gx, gy on
 * nodes.  ray: [max_step, 3] work buffer in the reference's receiver-to-source order;
 * *count receives the index of the last point (the source).  Returns 0, 2 for "end point
 * out of bound", 3 for "maximum number of steps reached".
 * One deviation: upper[a] = axis[i+1] is clamped to the last node where the reference
 * would index past the end of the axis (pure-Python IndexError).
 */
static int ray3_one(const double *z, int nz, const double *x, int nx, const double *y, int ny,
                    const double *gz, const double *gx, const double *gy, double zend,
                    double xend, double yend, double zsrc, double xsrc, double ysrc,
                    double stepsize, int max_step, int honor_grid, double *ray, int *count_out) {
    if (!(z[0] <= zend && zend <= z[nz - 1] && x[0] <= xend && xend <= x[nx - 1] &&
          y[0] <= yend && yend <= y[ny - 1]))
        return 2;
    const double *ax[3] = {z, x, y};
    const int na[3] = {nz, nx, ny};
    double lower[3], upper[3];
    int isrc = 0, jsrc = 0, ksrc = 0;
    double pcur[3] = {zend, xend, yend}, delta[3];
    if (honor_grid) {
        for (int a = 0; a < 3; ++a) {
            int i = ssr(ax[a], na[a], pcur[a]);
            lower[a] = (pcur[a] == ax[a][i]) ? ax[a][imax(i - 1, 0)] : ax[a][i];
            upper[a] = ax[a][imin(i + 1, na[a] - 1)];
        }
        isrc = ssr(z, nz, zsrc);
        jsrc = ssr(x, nx, xsrc);
        ksrc = ssr(y, ny, ysrc);
    }
    int count = 1;
    ray[0] = pcur[0]; ray[1] = pcur[1]; ray[2] = pcur[2];
    while (dist3(zsrc, xsrc, ysrc, pcur[0], pcur[1], pcur[2]) >= stepsize) {
        double g0 = interp3_one(z, nz, x, nx, y, ny, gz, pcur[0], pcur[1], pcur[2], NAN);
        double g1 = interp3_one(z, nz, x, nx, y, ny, gx, pcur[0], pcur[1], pcur[2], NAN);
        double g2 = interp3_one(z, nz, x, nx, y, ny, gy, pcur[0], pcur[1], pcur[2], NAN);
        double gn = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        double gni;
        if (gn > 0.0) gni = 1.0 / gn;
        else break;
        delta[0] = stepsize * g0 * gni;
        delta[1] = stepsize * g1 * gni;
        delta[2] = stepsize * g2 * gni;
        if (honor_grid) {
            double fac = shrink(pcur, delta, lower, upper, 3);
            for (int a = 0; a < 3; ++a) {
                pcur[a] -= fac * delta[a];
                pcur[a] = pmin2(pmax2(pcur[a], ax[a][0]), ax[a][na[a] - 1]);
            }
            if (fac < 1.0) {
                for (int a = 0; a < 3; ++a) {
                    if (fabs(pcur[a] - lower[a]) < 1.0e-8) pcur[a] = lower[a];
                    else if (fabs(pcur[a] - upper[a]) < 1.0e-8) pcur[a] = upper[a];
                }
                int idx[3];
                for (int a = 0; a < 3; ++a) {
                    int i = ssr(ax[a], na[a], pcur[a]);
                    idx[a] = i;
                    lower[a] = (pcur[a] == ax[a][i]) ? ax[a][imax(i - 1, 0)] : ax[a][i];
                    upper[a] = ax[a][imin(i + 1, na[a] - 1)];
                }
                ray[3 * count] = pcur[0]; ray[3 * count + 1] = pcur[1]; ray[3 * count + 2] = pcur[2];
                ++count;
                if (idx[0] == isrc && idx[1] == jsrc && idx[2] == ksrc) break;
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                pcur[a] -= delta[a];
                pcur[a] = pmin2(pmax2(pcur[a], ax[a][0]), ax[a][na[a] - 1]);
            }
            ray[3 * count] = pcur[0]; ray[3 * count + 1] = pcur[1]; ray[3 * count + 2] = pcur[2];
            ++count;
        }
        if (count >= max_step) return 3;
    }
    ray[3 * count] = zsrc; ray[3 * count + 1] = xsrc; ray[3 * count + 2] = ysrc;
    *count_out = count;
    return 0;
}

/* _ray3d_vectorized, _ray3d.py:123-164.  rays: [n, max_step, 3]; counts: [n];
 * status: [n] per-ray codes.  Returns the first nonzero per-ray status. */
int oracle_ray3d(const double *z, int nz, const double *x, int nx, const double *y, int ny,
                 const double *gz, const double *gx, const double *gy, const double *p, int n,
                 const double *src, double stepsize, int max_step, int honor_grid, double *rays,
                 int *counts, int *status) {
    int first = 0;
    for (int r = 0; r < n; ++r) {
        counts[r] = 0;
        int st = ray3_one(z, nz, x, nx, y, ny, gz, gx, gy, p[3 * r], p[3 * r + 1], p[3 * r + 2],
                          src[0], src[1], src[2], stepsize, max_step, honor_grid,
                          rays + (size_t)r * max_step * 3, counts + r);
        status[r] = st;
        if (st && !first) first = st;
    }
    return first;
}

/* _ray2d.py:9-91 (same deviation note as ray3_one). */
static int ray2_one(const double *z, int nz, const double *x, int nx, const double *gz,
                    const double *gx, double zend, double xend, double zsrc, double xsrc,
                    double stepsize, int max_step, int honor_grid, double *ray, int *count_out) {
    if (!(z[0] <= zend && zend <= z[nz - 1] && x[0] <= xend && xend <= x[nx - 1])) return 2;
    const double *ax[2] = {z, x};
    const int na[2] = {nz, nx};
    double lower[2], upper[2];
    int isrc = 0, jsrc = 0;
    double pcur[2] = {zend, xend}, delta[2];
    if (honor_grid) {
        for (int a = 0; a < 2; ++a) {
            int i = ssr(ax[a], na[a], pcur[a]);
            lower[a] = (pcur[a] == ax[a][i]) ? ax[a][imax(i - 1, 0)] : ax[a][i];
            upper[a] = ax[a][imin(i + 1, na[a] - 1)];
        }
        isrc = ssr(z, nz, zsrc);
        jsrc = ssr(x, nx, xsrc);
    }
    int count = 1;
    ray[0] = pcur[0]; ray[1] = pcur[1];
    while (dist2(zsrc, xsrc, pcur[0], pcur[1]) >= stepsize) {
        double g0 = interp2_one(z, nz, x, nx, gz, pcur[0], pcur[1], NAN);
        double g1 = interp2_one(z, nz, x, nx, gx, pcur[0], pcur[1], NAN);
        double gn = sqrt(g0 * g0 + g1 * g1);
        double gni;
        if (gn > 0.0) gni = 1.0 / gn;
        else break;
        delta[0] = stepsize * g0 * gni;
        delta[1] = stepsize * g1 * gni;
        if (honor_grid) {
            double fac = shrink(pcur, delta, lower, upper, 2);
            for (int a = 0; a < 2; ++a) {
                pcur[a] -= fac * delta[a];
                pcur[a] = pmin2(pmax2(pcur[a], ax[a][0]), ax[a][na[a] - 1]);
            }
            if (fac < 1.0) {
                for (int a = 0; a < 2; ++a) {
                    if (fabs(pcur[a] - lower[a]) < 1.0e-8) pcur[a] = lower[a];
                    else if (fabs(pcur[a] - upper[a]) < 1.0e-8) pcur[a] = upper[a];
                }
                int idx[2];
                for (int a = 0; a < 2; ++a) {
                    int i = ssr(ax[a], na[a], pcur[a]);
                    idx[a] = i;
                    lower[a] = (pcur[a] == ax[a][i]) ? ax[a][imax(i - 1, 0)] : ax[a][i];
                    upper[a] = ax[a][imin(i + 1, na[a] - 1)];
                }
                ray[2 * count] = pcur[0]; ray[2 * count + 1] = pcur[1];
                ++count;
                if (idx[0] == isrc && idx[1] == jsrc) break;
            }
        } else {
            for (int a = 0; a < 2; ++a) {
                pcur[a] -= delta[a];
                pcur[a] = pmin2(pmax2(pcur[a], ax[a][0]), ax[a][na[a] - 1]);
            }
            ray[2 * count] = pcur[0]; ray[2 * count + 1] = pcur[1];
            ++count;
        }
        if (count >= max_step) return 3;
    }
    ray[2 * count] = zsrc; ray[2 * count + 1] = xsrc;
    *count_out = count;
    return 0;
}

int oracle_ray2d(const double *z, int nz, const double *x, int nx, const double *gz,
                 const double *gx, const double *p, int n, const double *src, double stepsize,
                 int max_step, int honor_grid, double *rays, int *counts, int *status) {
    int first = 0;
    for (int r = 0; r < n; ++r) {
        counts[r] = 0;
        int st = ray2_one(z, nz, x, nx, gz, gx, p[2 * r], p[2 * r + 1], src[0], src[1], stepsize,
                          max_step, honor_grid, rays + (size_t)r * max_step * 2, counts + r);
        status[r] = st;
        if (st && !first) first = st;
    }
    return first;
}

/* ------------------------------------------------------------------------------------ */
/* Adjoint-state gradient of receiver traveltimes (no counterpart in the reference).     */
/* ------------------------------------------------------------------------------------ */
/*
 * Differentiates the converged fixed point of fteik3d and the receiver interpolation of
 * vinterp3d.  At a converged grid every lowered node equals one candidate of the node
 * update (cand3) for one sweep direction, evaluated on the final neighbour times; the
 * selection is recomputed here by evaluating all 8 directions and taking the first whose
 * minimum equals the node's time.  The adjoint of the triangular system is accumulated in
 * decreasing-traveltime order (every upwind neighbour of a selected stencil is earlier),
 * which is checked for every push.  Pinned by central finite differences of
 * oracle_fteik3d + vinterp3_one (tests/test_oracle_adjoint.py).
 */

/* Partial derivatives of the selected candidate: upwind node offsets by subset mask b of
 * the axes (bit 0 z, bit 1 x, bit 2 y), dT/dT_up[b], the slowness cell used and dT/ds. */
typedef struct {
    double w[8];
    size_t up[8];
    int has[8];
    size_t cell;
    double wcell;
} adj_stencil_t;

static void adj_stencil3(const double *tt, const double *slow, const dargs3_t *d, int i, int j,
                         int k, int sz, int sx, int sy, int which, int NZ, int NX, int NY,
                         adj_stencil_t *S) {
#define NODE(a, b, c) (((size_t)(a) * NX + (b)) * NY + (c))
#define CELL(a, b, c) (((size_t)(a) * (NX - 1) + (b)) * (NY - 1) + (c))
    for (int b = 0; b < 8; ++b) {
        S->w[b] = 0.0;
        S->has[b] = 0;
        S->up[b] = 0;
    }
    int dz_ = (which & 1) ? sz : 0, dx_ = (which & 2) ? sx : 0, dy_ = (which & 4) ? sy : 0;
    for (int b = 1; b < 8; ++b)
        if ((b & which) == b) {
            S->has[b] = 1;
            S->up[b] = NODE(i - ((b & 1) ? dz_ : 0), j - ((b & 2) ? dx_ : 0), k - ((b & 4) ? dy_ : 0));
        }
    int i1 = i - (sz > 0), j1 = j - (sx > 0), k1 = k - (sy > 0);
    int jm = imax(j - 1, 0), jp = imin(j, NX - 2);
    int km = imax(k - 1, 0), kp = imin(k, NY - 2);
    int im = imax(i - 1, 0), ip = imin(i, NZ - 2);
    /* the cell achieving a min(): first of equal values, like Python's min() */
    size_t c4[4];
    int nc = 0;
    double dd = 0.0;
    if (which == 1 || which == 2 || which == 4) {
        if (which == 1) {
            c4[0] = CELL(i1, jm, km); c4[1] = CELL(i1, jm, kp); c4[2] = CELL(i1, jp, km); c4[3] = CELL(i1, jp, kp);
            dd = d->dz;
        } else if (which == 2) {
            c4[0] = CELL(im, j1, km); c4[1] = CELL(ip, j1, km); c4[2] = CELL(im, j1, kp); c4[3] = CELL(ip, j1, kp);
            dd = d->dx;
        } else {
            c4[0] = CELL(im, jm, k1); c4[1] = CELL(im, jp, k1); c4[2] = CELL(ip, jm, k1); c4[3] = CELL(ip, jp, k1);
            dd = d->dy;
        }
        nc = 4;
    } else if (which != 7) {
        if (which == 3) { c4[0] = CELL(i1, j1, km); c4[1] = CELL(i1, j1, kp); }
        else if (which == 5) { c4[0] = CELL(i1, jm, k1); c4[1] = CELL(i1, jp, k1); }
        else { c4[0] = CELL(im, j1, k1); c4[1] = CELL(ip, j1, k1); }
        nc = 2;
    } else {
        c4[0] = CELL(i1, j1, k1);
        nc = 1;
    }
    size_t best = c4[0];
    for (int c = 1; c < nc; ++c)
        if (slow[c4[c]] < slow[best]) best = c4[c];
    double v = slow[best];
    S->cell = best;
    if (nc == 4) {
        /* T = T_up + d * v */
        S->w[which] = 1.0;
        S->wcell = dd;
    } else if (nc == 2) {
        /* Plane-wave operator with axes (a, b) in z, x, y order, a1 = upwind time along a,
         * b1 = along b, ab = diagonal.  All three planes read
         *   T = (a2i (ab + a1 - b1) + b2i (ab + b1 - a1) + sqrt(R)) / (a2i + b2i),
         *   R = 4 v^2 (a2i + b2i) - a2i b2i (2 (b1 - a1))^2. */
        int ba = which & 1 ? 1 : 2, bb = which & 4 ? 4 : 2;
        double a2i = ba == 1 ? d->dz2i : d->dx2i, b2i = bb == 4 ? d->dy2i : d->dx2i;
        double a1 = tt[S->up[ba]], b1 = tt[S->up[bb]];
        double q = 2.0 * (b1 - a1);
        double R = 4.0 * v * v * (a2i + b2i) - a2i * b2i * q * q;
        double sR = sqrt(R), sum = a2i + b2i;
        /* dR/da1 = 4 a2i b2i q, dR/db1 = -4 a2i b2i q */
        S->w[ba] = (a2i - b2i + 2.0 * a2i * b2i * q / sR) / sum;
        S->w[bb] = (b2i - a2i - 2.0 * a2i * b2i * q / sR) / sum;
        S->w[ba | bb] = 1.0;
        S->wcell = 4.0 * v / sR;
    } else {
        /* 3D operator: T = (tb z2i + ta x2i + tc y2i + sqrt(9 v^2 dsum - Q)) / dsum with
         * ta, tb, tc linear in the 7 upwind times (_fteik3d.py:150-152). */
        double tv = tt[S->up[1]], te = tt[S->up[2]], tn = tt[S->up[4]];
        double tev = tt[S->up[3]], tnv = tt[S->up[5]], ten = tt[S->up[6]], tnve = tt[S->up[7]];
        double ta = te - 0.5 * tn + 0.5 * ten - 0.5 * tv + 0.5 * tev - tnv + tnve;
        double tb = tv - 0.5 * tn + 0.5 * tnv - 0.5 * te + 0.5 * tev - ten + tnve;
        double tc = tn - 0.5 * te + 0.5 * ten - 0.5 * tv + 0.5 * tnv - tev + tnve;
        double Q = d->dz2dx2 * (ta - tb) * (ta - tb) + d->dz2dy2 * (tb - tc) * (tb - tc) +
                   d->dx2dy2 * (ta - tc) * (ta - tc);
        double sR = sqrt(9.0 * v * v * d->dsum - Q);
        /* dT/dta etc. */
        double ga = (d->dx2i - (d->dz2dx2 * (ta - tb) + d->dx2dy2 * (ta - tc)) / sR) / d->dsum;
        double gb = (d->dz2i - (-d->dz2dx2 * (ta - tb) + d->dz2dy2 * (tb - tc)) / sR) / d->dsum;
        double gc = (d->dy2i - (-d->dz2dy2 * (tb - tc) - d->dx2dy2 * (ta - tc)) / sR) / d->dsum;
        /* coefficients of (tv, te, tn, tev, tnv, ten, tnve) in (ta, tb, tc) */
        S->w[1] = -0.5 * ga + gb - 0.5 * gc;
        S->w[2] = ga - 0.5 * gb - 0.5 * gc;
        S->w[4] = -0.5 * ga - 0.5 * gb + gc;
        S->w[3] = 0.5 * ga + 0.5 * gb - gc;
        S->w[5] = -ga + 0.5 * gb + 0.5 * gc;
        S->w[6] = 0.5 * ga - gb + 0.5 * gc;
        S->w[7] = ga + gb + gc;
        S->wcell = 9.0 * v / sR;
    }
#undef NODE
#undef CELL
}

/* Recomputed selection: the direction (pass order of sweep3d) and candidate whose value is
 * smallest over the 8 directions, the first one on ties.  Returns the candidate's axis set,
 * 0 if the node keeps its initial value (no candidate is below BIG), and sets *stale when
 * that smallest candidate is above the node's time: the node then holds a value computed
 * from earlier neighbour times (a 2D/3D candidate whose causality condition no longer
 * holds), and its current best candidate stands in for it. */
static int adj_select3(const double *tt, const double *slow, const dargs3_t *d, int i, int j,
                       int k, int NZ, int NX, int NY, int *sz, int *sx, int *sy, int *stale) {
    static const int dirs[8][3] = {{1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {1, -1, -1},
                                   {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    double t = tt[((size_t)i * NX + j) * NY + k];
    double best = BIG;
    int bw = 0;
    for (int p = 0; p < 8; ++p) {
        int a = dirs[p][0], b = dirs[p][1], c = dirs[p][2];
        if (i - a < 0 || i - a >= NZ || j - b < 0 || j - b >= NX || k - c < 0 || k - c >= NY)
            continue;
        int which;
        double m = cand3(tt, slow, d, i, j, k, a > 0, b > 0, c > 0, a, b, c, NZ, NX, NY, BIG,
                         &which);
        if (m < best) {
            best = m;
            bw = which;
            *sz = a;
            *sx = b;
            *sy = c;
        }
    }
    if (!(best <= t)) {
        /* above the node's time: a seeded node keeps its initial value; otherwise stale */
        *stale = 1;
    } else {
        *stale = 0;
    }
    return bw;
}

static const double *g_sort_tt; /* qsort has no context argument */
static int adj_cmp_desc(const void *a, const void *b) {
    double ta = g_sort_tt[*(const size_t *)a], tb = g_sort_tt[*(const size_t *)b];
    return ta < tb ? 1 : (ta > tb ? -1 : 0);
}

/* dT/d(corner times) of vinterp3_one for one receiver, or the vzero path.  Adds
 * w * dT/dt_c into lam and w * dT/dvzero into *dvz.  Axes are absolute, v index [z][x][y]
 * (the vinterp3_one call convention with x <- z axis etc.). */
static void adj_receiver3(const double *z, int nz, const double *x, int nx, const double *y,
                          int ny, const double *tt, double zq, double xq, double yq, double zs,
                          double xs, double ys, double w, double *lam, double *dvz) {
    if (!(z[0] <= zq && zq <= z[nz - 1] && x[0] <= xq && xq <= x[nx - 1] && y[0] <= yq &&
          yq <= y[ny - 1]))
        return;
    double dist = dist3(zs, xs, ys, zq, xq, yq);
    int zsi = ssr(z, nz, zs), xsi = ssr(x, nx, xs), ysi = ssr(y, ny, ys);
    int i1 = ssr(z, nz, zq), j1 = ssr(x, nx, xq), k1 = ssr(y, ny, yq);
    if (zsi == i1 && xsi == j1 && ysi == k1) {
        *dvz += w * dist;
        return;
    }
    cell3_t c = cell3(z, nz, x, nx, y, ny, i1, j1, k1);
    double vol = fabs((c.x2 - c.x1) * (c.y2 - c.y1) * (c.z2 - c.z1));
    double a[8], t[8];
    size_t id[8];
    int ok[8];
    double V = 0.0;
    for (int q = 0; q < 8; ++q) {
        int di = q & 1, dj = (q >> 1) & 1, dk = (q >> 2) & 1;
        ok[q] = !((di && !c.ok2x) || (dj && !c.ok2y) || (dk && !c.ok2z));
        a[q] = 0.0;
        t[q] = 1.0;
        if (!ok[q]) continue;
        double pz = di ? c.x2 : c.x1, px = dj ? c.y2 : c.y1, py = dk ? c.z2 : c.z1;
        double wq = fabs(((di ? c.x1 : c.x2) - zq) * ((dj ? c.y1 : c.y2) - xq) *
                         ((dk ? c.z1 : c.z2) - yq));
        id[q] = ((size_t)(i1 + di) * nx + (j1 + dj)) * ny + (k1 + dk);
        t[q] = tt[id[q]];
        if (t[q] == 0.0) {
            *dvz += w * dist;
            return;
        }
        a[q] = dist3(zs, xs, ys, pz, px, py) * wq / vol;
        V += a[q] / t[q];
    }
    for (int q = 0; q < 8; ++q)
        if (ok[q]) lam[id[q]] += w * dist * a[q] / (V * V * t[q] * t[q]);
}

/*
 * d L / d slow for L with dL/dT_rec = dldt[s][r] (T_rec = vinterp3 of source s at receiver
 * r).  origin: grid origin; src, rec: absolute coordinates; tt: converged grids stacked per
 * source.  dlds (nz*nx*ny) is overwritten.  *orphans counts the stale nodes (see
 * adj_select3) with non-zero adjoint; *rounds_out the largest number of push rounds a
 * source needed (2 when every selected stencil reads only earlier nodes: one round and
 * the empty one that confirms it).  Returns 0, 1 for a source out of bounds, 2 if the pushes do not settle
 * within 1000 rounds.
 */
int oracle_adjoint3d(const double *slow, int nz, int nx, int ny, double dz, double dx,
                     double dy, const double *origin, const double *src, int nsrc,
                     const double *tt_all, const double *rec, int nrec, const double *dldt,
                     double *dlds, int *orphans, int *rounds_out) {
    int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    size_t nn = (size_t)NZ * NX * NY;
    dargs3_t d;
    d.dz = dz; d.dx = dx; d.dy = dy;
    d.dz2i = 1.0 / dz / dz;
    d.dx2i = 1.0 / dx / dx;
    d.dy2i = 1.0 / dy / dy;
    d.dz2dx2 = d.dz2i * d.dx2i;
    d.dz2dy2 = d.dz2i * d.dy2i;
    d.dx2dy2 = d.dx2i * d.dy2i;
    d.dsum = d.dz2i + d.dx2i + d.dy2i;
    double *za = (double *)malloc(NZ * sizeof(double));
    double *xa = (double *)malloc(NX * sizeof(double));
    double *ya = (double *)malloc(NY * sizeof(double));
    for (int i = 0; i < NZ; ++i) za[i] = origin[0] + dz * (double)i;
    for (int i = 0; i < NX; ++i) xa[i] = origin[1] + dx * (double)i;
    for (int i = 0; i < NY; ++i) ya[i] = origin[2] + dy * (double)i;
    double *lam = (double *)malloc(nn * sizeof(double));
    size_t *ord = (size_t *)malloc(nn * sizeof(size_t));
    double *pushed = (double *)malloc(nn * sizeof(double));
    unsigned char *kind = (unsigned char *)malloc(nn);
    adj_stencil_t *st = (adj_stencil_t *)malloc(nn * sizeof(adj_stencil_t));
    if (rounds_out) *rounds_out = 0;
    memset(dlds, 0, (size_t)nz * nx * ny * sizeof(double));
    int status = 0, norph = 0;
    for (int s = 0; s < nsrc && !status; ++s) {
        const double *tt = tt_all + nn * s;
        double zr = src[3 * s] - origin[0], xr = src[3 * s + 1] - origin[1],
               yr = src[3 * s + 2] - origin[2];
        if (!(0.0 <= zr && zr <= dz * nz && 0.0 <= xr && xr <= dx * nx && 0.0 <= yr &&
              yr <= dy * ny)) {
            status = 1;
            break;
        }
        double zsa = zr / dz, xsa = xr / dx, ysa = yr / dy;
        if (zsa >= nz) zsa = zsa - EPS;
        if (xsa >= nx) xsa = xsa - EPS;
        if (ysa >= ny) ysa = ysa - EPS;
        int zsi = imin((int)zsa, nz - 1), xsi = imin((int)xsa, nx - 1), ysi = imin((int)ysa, ny - 1);
        size_t scell = ((size_t)zsi * nx + xsi) * ny + ysi;
        memset(lam, 0, nn * sizeof(double));
        double dvz = 0.0;
        for (int r = 0; r < nrec; ++r)
            adj_receiver3(za, NZ, xa, NX, ya, NY, tt, rec[3 * r], rec[3 * r + 1], rec[3 * r + 2],
                          src[3 * s], src[3 * s + 1], src[3 * s + 2],
                          dldt[(size_t)s * nrec + r], lam, &dvz);
        /* selection and stencil of every node, once */
        for (size_t n = 0; n < nn; ++n) {
            int i = (int)(n / ((size_t)NX * NY)), j = (int)((n / NY) % NX), k = (int)(n % NY);
            int sz = 0, sx = 0, sy = 0, stale = 0;
            int which = adj_select3(tt, slow, &d, i, j, k, NZ, NX, NY, &sz, &sx, &sy, &stale);
            kind[n] = 0;
            if (stale || !which) {
                int di = i - zsi, dj = j - xsi, dk = k - ysi;
                if (di >= 0 && di <= 1 && dj >= 0 && dj <= 1 && dk >= 0 && dk <= 1) {
                    kind[n] = 2; /* seeded corner: T = vzero * distance */
                    continue;
                }
                if (!which) continue;
            }
            adj_stencil3(tt, slow, &d, i, j, k, sz, sx, sy, which, NZ, NX, NY, st + n);
            int causal = 1;
            for (int b = 1; b < 8; ++b)
                if (st[n].has[b] && !(tt[st[n].up[b]] < tt[n])) causal = 0;
            /* a stale node's stand-in must only read earlier nodes, else it is dropped */
            if (stale && !causal) continue;
            kind[n] = stale ? 3 : 1;
        }
        for (size_t n = 0; n < nn; ++n) ord[n] = n;
        g_sort_tt = tt;
        qsort(ord, nn, sizeof(size_t), adj_cmp_desc);
        /* lambda = g + J^T lambda: push, in decreasing-time order, the part of each node's
         * lambda not pushed yet; a candidate reading a later node (the 3D operator is not
         * strictly causal) is caught up by the next round */
        memset(pushed, 0, nn * sizeof(double));
        int round = 0;
        double scale = 0.0;
        for (;;) {
            double dmax = 0.0, lmax = 0.0;
            for (size_t o = 0; o < nn; ++o) {
                size_t n = ord[o];
                double l = lam[n] - pushed[n];
                if (l == 0.0) continue;
                pushed[n] = lam[n];
                if (fabs(l) > dmax) dmax = fabs(l);
                if (fabs(lam[n]) > lmax) lmax = fabs(lam[n]);
                if (kind[n] == 2) {
                    int i = (int)(n / ((size_t)NX * NY)), j = (int)((n / NY) % NX), k = (int)(n % NY);
                    dvz += l * t_ana3(i, j, k, dz, dx, dy, zsa, xsa, ysa, 1.0);
                    continue;
                }
                if (kind[n] == 0) {
                    if (round == 0 && tt[n] < BIG) ++norph;
                    continue;
                }
                if (kind[n] == 3 && round == 0) ++norph;
                const adj_stencil_t *S = st + n;
                dlds[S->cell] += l * S->wcell;
                for (int b = 1; b < 8; ++b)
                    if (S->has[b]) lam[S->up[b]] += l * S->w[b];
            }
            if (round == 0) scale = lmax;
            ++round;
            /* increments below the rounding of the adjoint itself change nothing */
            if (dmax <= 1e-16 * scale) break;
            if (round >= 1000) {
                status = 2;
                break;
            }
        }
        if (rounds_out && round > *rounds_out) *rounds_out = round;
        dlds[scell] += dvz;
    }
    free(za);
    free(xa);
    free(ya);
    free(lam);
    free(ord);
    free(pushed);
    free(kind);
    free(st);
    if (orphans) *orphans = norph;
    return status;
}

/*
 * Exact discrete adjoint: reverse mode through the sweep history of oracle_fteik3d in
 * converge mode.  The state before every pass is kept; passes are reversed last to first,
 * nodes in the reverse of the pass's loop order.  A node that changed in the pass has the
 * adjoint of its new value pushed through the candidate that produced it, evaluated on the
 * state after the pass (its upwind neighbours were final for the pass when it was
 * updated; the winner is recomputed as the candidate equal to the new value), and the
 * adjoint of its old value is zero; a node that did not change carries its adjoint to the
 * previous pass.  What is left on the initial state reaches vzero through the 8 seeded
 * corners.  Same arguments and outputs as oracle_adjoint3d (no stale nodes exist here);
 * *rounds_out receives the number of sets of the longest history.
 */
int oracle_adjoint3d_exact(const double *slow, int nz, int nx, int ny, double dz, double dx,
                           double dy, const double *origin, const double *src, int nsrc,
                           int max_sets, const double *rec, int nrec, const double *dldt,
                           double *dlds, int *rounds_out) {
    static const int dirs[8][3] = {{1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {1, -1, -1},
                                   {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    size_t nn = (size_t)NZ * NX * NY;
    dargs3_t d;
    d.dz = dz; d.dx = dx; d.dy = dy;
    d.dz2i = 1.0 / dz / dz;
    d.dx2i = 1.0 / dx / dx;
    d.dy2i = 1.0 / dy / dy;
    d.dz2dx2 = d.dz2i * d.dx2i;
    d.dz2dy2 = d.dz2i * d.dy2i;
    d.dx2dy2 = d.dx2i * d.dy2i;
    d.dsum = d.dz2i + d.dx2i + d.dy2i;
    double *za = (double *)malloc(NZ * sizeof(double));
    double *xa = (double *)malloc(NX * sizeof(double));
    double *ya = (double *)malloc(NY * sizeof(double));
    for (int i = 0; i < NZ; ++i) za[i] = origin[0] + dz * (double)i;
    for (int i = 0; i < NX; ++i) xa[i] = origin[1] + dx * (double)i;
    for (int i = 0; i < NY; ++i) ya[i] = origin[2] + dy * (double)i;
    double *lam = (double *)malloc(nn * sizeof(double));
    double *hist = (double *)malloc(nn * sizeof(double) * (8 * (size_t)max_sets + 1));
    memset(dlds, 0, (size_t)nz * nx * ny * sizeof(double));
    if (rounds_out) *rounds_out = 0;
    int status = 0;
    for (int s = 0; s < nsrc && !status; ++s) {
        double zr = src[3 * s] - origin[0], xr = src[3 * s + 1] - origin[1],
               yr = src[3 * s + 2] - origin[2];
        if (!(0.0 <= zr && zr <= dz * nz && 0.0 <= xr && xr <= dx * nx && 0.0 <= yr &&
              yr <= dy * ny)) {
            status = 1;
            break;
        }
        double zsa = zr / dz, xsa = xr / dx, ysa = yr / dy;
        if (zsa >= nz) zsa = zsa - EPS;
        if (xsa >= nx) xsa = xsa - EPS;
        if (ysa >= ny) ysa = ysa - EPS;
        int zsi = imin((int)zsa, nz - 1), xsi = imin((int)xsa, nx - 1), ysi = imin((int)ysa, ny - 1);
        size_t scell = ((size_t)zsi * nx + xsi) * ny + ysi;
        double vzero = slow[scell];
        /* forward, keeping the state before every pass (fteik3d, :429-462) */
        double *t0 = hist;
        for (size_t n = 0; n < nn; ++n) t0[n] = BIG;
        static const int corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                          {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        for (int c = 0; c < 8; ++c) {
            int i = zsi + corners[c][0], j = xsi + corners[c][1], k = ysi + corners[c][2];
            t0[((size_t)i * NX + j) * NY + k] = t_ana3(i, j, k, dz, dx, dy, zsa, xsa, ysa, vzero);
        }
        int npass = 0, sets = 0;
        for (int set = 0; set < max_sets; ++set) {
            int changed = 0;
            for (int p = 0; p < 8; ++p) {
                double *cur = hist + nn * (size_t)(npass + 1);
                memcpy(cur, hist + nn * (size_t)npass, nn * sizeof(double));
                int sz = dirs[p][0], sx = dirs[p][1], sy = dirs[p][2];
                int i0 = sz > 0 ? 1 : NZ - 2, ie = sz > 0 ? NZ : -1;
                int j0 = sx > 0 ? 1 : NX - 2, je = sx > 0 ? NX : -1;
                int k0 = sy > 0 ? 1 : NY - 2, ke = sy > 0 ? NY : -1;
                for (int k = k0; k != ke; k += sy)
                    for (int j = j0; j != je; j += sx)
                        for (int i = i0; i != ie; i += sz)
                            sweep_node3(cur, NULL, slow, &d, i, j, k, sz > 0, sx > 0, sy > 0, sz,
                                        sx, sy, NZ, NX, NY, 0);
                if (memcmp(cur, hist + nn * (size_t)npass, nn * sizeof(double))) changed = 1;
                ++npass;
            }
            ++sets;
            if (!changed) break;
        }
        if (rounds_out && sets > *rounds_out) *rounds_out = sets;
        const double *tt = hist + nn * (size_t)npass;
        /* receiver adjoint on the final grid */
        memset(lam, 0, nn * sizeof(double));
        double dvz = 0.0;
        for (int r = 0; r < nrec; ++r)
            adj_receiver3(za, NZ, xa, NX, ya, NY, tt, rec[3 * r], rec[3 * r + 1], rec[3 * r + 2],
                          src[3 * s], src[3 * s + 1], src[3 * s + 2],
                          dldt[(size_t)s * nrec + r], lam, &dvz);
        /* reverse the passes */
        for (int P = npass - 1; P >= 0; --P) {
            const double *before = hist + nn * (size_t)P, *after = hist + nn * (size_t)(P + 1);
            int p = P % 8;
            int sz = dirs[p][0], sx = dirs[p][1], sy = dirs[p][2];
            int i0 = sz > 0 ? 1 : NZ - 2, ie = sz > 0 ? NZ : -1;
            int j0 = sx > 0 ? 1 : NX - 2, je = sx > 0 ? NX : -1;
            int k0 = sy > 0 ? 1 : NY - 2, ke = sy > 0 ? NY : -1;
            /* reverse of `for k: for j: for i` */
            for (int k = ke - sy; k != k0 - sy; k -= sy)
                for (int j = je - sx; j != j0 - sx; j -= sx)
                    for (int i = ie - sz; i != i0 - sz; i -= sz) {
                        size_t n = ((size_t)i * NX + j) * NY + k;
                        if (after[n] == before[n]) continue;
                        double l = lam[n];
                        lam[n] = 0.0;
                        if (l == 0.0) continue;
                        int which;
                        double tn = cand3(after, slow, &d, i, j, k, sz > 0, sx > 0, sy > 0, sz, sx,
                                          sy, NZ, NX, NY, before[n], &which);
                        if (tn != after[n]) {
                            status = 2;
                            break;
                        }
                        adj_stencil_t S;
                        adj_stencil3(after, slow, &d, i, j, k, sz, sx, sy, which, NZ, NX, NY, &S);
                        dlds[S.cell] += l * S.wcell;
                        for (int b = 1; b < 8; ++b)
                            if (S.has[b]) lam[S.up[b]] += l * S.w[b];
                    }
        }
        for (int c = 0; c < 8; ++c) {
            int i = zsi + corners[c][0], j = xsi + corners[c][1], k = ysi + corners[c][2];
            dvz += lam[((size_t)i * NX + j) * NY + k] * t_ana3(i, j, k, dz, dx, dy, zsa, xsa, ysa, 1.0);
        }
        dlds[scell] += dvz;
    }
    free(za);
    free(xa);
    free(ya);
    free(lam);
    free(hist);
    return status;
}
