// FTeik 3D node operator shared by the sweep kernels and the adjoint's stencil selection.
// Args supplies dz, dx, dy, dz2i, dx2i, dy2i, dz2dx2, dz2dy2, dx2dy2 and dsum.
#pragma once

#include "fteik_device.hpp"

namespace fteik {

__device__ __forceinline__ double min4(double a, double b, double c, double d) {
    return fmin(fmin(a, b), fmin(c, d));
}

// FTeik 3D node update in oriented coordinates (a = z, b = x, c = y, all pointing
// downwind).  tv/te/tn: upwind neighbours along z/x/y; tev, ten, tnv, tnve: the
// diagonal ones; sXYZ: slowness of the cell at (a-1 | a') x (b-1 | b') x (c-1 | c'),
// with ' the downwind-side cell clamped to the grid (the reference's max(i-1,0) /
// min(i, n-2) pair).  Returns the new value; *which receives the winning candidate.
template <class Args>
__device__ __forceinline__ double node_update(const Args &A, double t0, double tv, double te,
                                              double tn, double tev, double ten, double tnv,
                                              double tnve, double s000, double s001, double s010,
                                              double s011, double s100, double s101, double s110,
                                              int *which) {
    const double dz = A.dz, dx = A.dx, dy = A.dy;
    const double dz2i = A.dz2i, dx2i = A.dx2i, dy2i = A.dy2i;

    // 1D operators (refracted times), _fteik3d.py:78-104
    const double t1d1 = tv + dz * min4(s000, s001, s010, s011);
    const double t1d2 = te + dx * min4(s000, s100, s001, s101);
    const double t1d3 = tn + dy * min4(s000, s010, s100, s110);
    const double t1d = fmin(fmin(t1d1, t1d2), t1d3);

    // 2D plane-wave operators, _fteik3d.py:107-143
    double t2d1 = kBig, t2d2 = kBig, t2d3 = kBig;
    {
        const double vref = fmin(s000, s001);  // ZX plane
        if (tv < te + dx * vref && te < tv + dz * vref) {
            const double ta = tev + te - tv, tb = tev - te + tv, q = ta - tb;
            t2d1 = ((tb * dz2i + ta * dx2i) +
                    sqrt(4.0 * (vref * vref) * (dz2i + dx2i) - dz2i * dx2i * (q * q))) /
                   (dz2i + dx2i);
        }
    }
    {
        const double vref = fmin(s000, s010);  // ZY plane
        if (tv < tn + dy * vref && tn < tv + dz * vref) {
            const double ta = tv - tn + tnv, tb = tn - tv + tnv, q = ta - tb;
            t2d2 = ((ta * dz2i + tb * dy2i) +
                    sqrt(4.0 * (vref * vref) * (dz2i + dy2i) - dz2i * dy2i * (q * q))) /
                   (dz2i + dy2i);
        }
    }
    {
        const double vref = fmin(s000, s100);  // XY plane
        if (te < tn + dy * vref && tn < te + dx * vref) {
            const double ta = te - tn + ten, tb = tn - te + ten, q = ta - tb;
            t2d3 = ((ta * dx2i + tb * dy2i) +
                    sqrt(4.0 * (vref * vref) * (dx2i + dy2i) - dx2i * dy2i * (q * q))) /
                   (dx2i + dy2i);
        }
    }
    const double t2d = fmin(fmin(t2d1, t2d2), t2d3);

    // 3D operator, _fteik3d.py:146-159
    double t3d = kBig;
    if (fmin(t1d, t2d) > fmax(fmax(tv, te), tn)) {
        const double vref = s000;
        const double ta = te - 0.5 * tn + 0.5 * ten - 0.5 * tv + 0.5 * tev - tnv + tnve;
        const double tb = tv - 0.5 * tn + 0.5 * tnv - 0.5 * te + 0.5 * tev - ten + tnve;
        const double tc = tn - 0.5 * te + 0.5 * ten - 0.5 * tv + 0.5 * tnv - tev + tnve;
        const double t2 = vref * vref * A.dsum * 9.0;
        const double q1 = ta - tb, q2 = tb - tc, q3 = ta - tc;
        double t3 = A.dz2dx2 * (q1 * q1);
        t3 += A.dz2dy2 * (q2 * q2);
        t3 += A.dx2dy2 * (q3 * q3);
        if (t2 >= t3) t3d = (tb * dz2i + ta * dx2i + tc * dy2i + sqrt(t2 - t3)) / A.dsum;
    }

    const double tnew = fmin(fmin(fmin(t0, t1d), t2d), t3d);
    // winning candidate in the reference's priority order, _fteik3d.py:166-200;
    // bit 0 = z, bit 1 = x, bit 2 = y take part
    int w;
    if (tnew == t1d1) w = 1;
    else if (tnew == t1d2) w = 2;
    else if (tnew == t1d3) w = 4;
    else if (tnew == t2d1) w = 3;
    else if (tnew == t2d2) w = 5;
    else if (tnew == t2d3) w = 6;
    else w = 7;
    *which = w;
    return tnew;
}

}  // namespace fteik
