// FTeik 3D node operator shared by the sweep kernels and the adjoint's stencil selection.
#pragma once

#include "fteik_device.hpp"

namespace fteik {

// Grid constants of the 3D operator, fixed for a solve (computed on the host).
struct Stencil3 {
    double dz, dx, dy, dz2i, dx2i, dy2i, dz2dx2, dz2dy2, dx2dy2, dsum;
    double szx, szy, sxy;        // dz2i + dx2i, dz2i + dy2i, dx2i + dy2i: 2D denominators
    double rzx, rzy, rxy, rsum;  // 1 / szx, 1 / szy, 1 / sxy, 1 / dsum, correctly rounded
};

// Grid spacings the 3D operator accepts: the division by its fixed denominators (see
// div_fixed) needs their inverse squares, sums and reciprocals to be normal doubles, which
// holds with a wide margin for spacings in [1e-100, 1e100]; calls outside are refused.
inline bool spacing_ok3(double dz, double dx, double dy) {
    auto ok = [](double h) { return h >= 1e-100 && h <= 1e100; };
    return ok(dz) && ok(dx) && ok(dy);
}

inline Stencil3 make_stencil3(double dz, double dx, double dy) {
    Stencil3 c;
    c.dz = dz; c.dx = dx; c.dy = dy;
    c.dz2i = 1.0 / dz / dz;
    c.dx2i = 1.0 / dx / dx;
    c.dy2i = 1.0 / dy / dy;
    c.dz2dx2 = c.dz2i * c.dx2i;
    c.dz2dy2 = c.dz2i * c.dy2i;
    c.dx2dy2 = c.dx2i * c.dy2i;
    c.dsum = c.dz2i + c.dx2i + c.dy2i;
    c.szx = c.dz2i + c.dx2i;
    c.szy = c.dz2i + c.dy2i;
    c.sxy = c.dx2i + c.dy2i;
    c.rzx = 1.0 / c.szx;
    c.rzy = 1.0 / c.szy;
    c.rxy = 1.0 / c.sxy;
    c.rsum = 1.0 / c.dsum;
    return c;
}

// A copy of the constants held in vector registers for the whole sweep loop: as uniform
// values they would take ~34 scalar registers (the loop's bookkeeping already fills the
// scalar file) and, with one scalar operand allowed per vector instruction, extra moves.
// The empty asm only pins each value to a VGPR pair.
__device__ __forceinline__ void pin_vgpr(double &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ Stencil3 stencil_in_vgprs(const Stencil3 &a) {
    Stencil3 c = a;
    pin_vgpr(c.dz); pin_vgpr(c.dx); pin_vgpr(c.dy);
    pin_vgpr(c.dz2i); pin_vgpr(c.dx2i); pin_vgpr(c.dy2i);
    pin_vgpr(c.dz2dx2); pin_vgpr(c.dz2dy2); pin_vgpr(c.dx2dy2); pin_vgpr(c.dsum);
    pin_vgpr(c.szx); pin_vgpr(c.szy); pin_vgpr(c.sxy);
    pin_vgpr(c.rzx); pin_vgpr(c.rzy); pin_vgpr(c.rxy); pin_vgpr(c.rsum);
    return c;
}

__device__ __forceinline__ double min4(double a, double b, double c, double d) {
    return fmin(fmin(a, b), fmin(c, d));
}

// FTeik 3D node update in oriented coordinates (a = z, b = x, c = y, all pointing
// downwind).  tv/te/tn: upwind neighbours along z/x/y; tev, ten, tnv, tnve: the
// diagonal ones; sXYZ: slowness of the cell at (a-1 | a') x (b-1 | b') x (c-1 | c'),
// with ' the downwind-side cell clamped to the grid (the reference's max(i-1,0) /
// min(i, n-2) pair).  Returns the new value; *which receives the winning candidate.
// Args derives from Stencil3.  BF: every lane computes the plane-wave candidates and their
// tests select them (no branches on the dependent chain of a wave alone on its SIMD);
// otherwise a branch skips a candidate that no lane of the wave admits.  Same results.
template <bool BF = false, class Args>
__device__ __forceinline__ double node_update(const Args &A, double t0, double tv, double te,
                                              double tn, double tev, double ten, double tnv,
                                              double tnve, double s000, double s001, double s010,
                                              double s011, double s100, double s101, double s110,
                                              int *which) {
    const double dz = A.dz, dx = A.dx, dy = A.dy;
    const double dz2i = A.dz2i, dx2i = A.dx2i, dy2i = A.dy2i;
    // the reference's (dz2i + dx2i), dz2i * dx2i and x / (dz2i + dx2i) are evaluated from
    // the precomputed sums, products and reciprocals: the same doubles.  A quotient that
    // would be infinite (infinite slowness) comes out NaN instead; both lose every fmin
    // below (t0 is finite), so the unchecked division suffices here.

    // 1D operators (refracted times), _fteik3d.py:78-104
    const double t1d1 = tv + dz * min4(s000, s001, s010, s011);
    const double t1d2 = te + dx * min4(s000, s100, s001, s101);
    const double t1d3 = tn + dy * min4(s000, s010, s100, s110);
    const double t1d = fmin(fmin(t1d1, t1d2), t1d3);

    // 2D plane-wave operators, _fteik3d.py:107-143
    double t2d1 = kBig, t2d2 = kBig, t2d3 = kBig;
    {
        const double vref = fmin(s000, s001);  // ZX plane
        const bool c = (tv < te + dx * vref) & (te < tv + dz * vref);
        if constexpr (BF) {
            const double ta = tev + te - tv, tb = tev - te + tv, q = ta - tb;
            const double v = div_fixed((tb * dz2i + ta * dx2i) + sqrt_rn(c ? 4.0 * (vref * vref) * A.szx - A.dz2dx2 * (q * q) : 1.0), A.szx, A.rzx);
            t2d1 = c ? v : kBig;
        } else if (c) {
            const double ta = tev + te - tv, tb = tev - te + tv, q = ta - tb;
            t2d1 = div_fixed((tb * dz2i + ta * dx2i) + sqrt_rn(4.0 * (vref * vref) * A.szx - A.dz2dx2 * (q * q)), A.szx, A.rzx);
        }
    }
    {
        const double vref = fmin(s000, s010);  // ZY plane
        const bool c = (tv < tn + dy * vref) & (tn < tv + dz * vref);
        if constexpr (BF) {
            const double ta = tv - tn + tnv, tb = tn - tv + tnv, q = ta - tb;
            const double v = div_fixed((ta * dz2i + tb * dy2i) + sqrt_rn(c ? 4.0 * (vref * vref) * A.szy - A.dz2dy2 * (q * q) : 1.0), A.szy, A.rzy);
            t2d2 = c ? v : kBig;
        } else if (c) {
            const double ta = tv - tn + tnv, tb = tn - tv + tnv, q = ta - tb;
            t2d2 = div_fixed((ta * dz2i + tb * dy2i) + sqrt_rn(4.0 * (vref * vref) * A.szy - A.dz2dy2 * (q * q)), A.szy, A.rzy);
        }
    }
    {
        const double vref = fmin(s000, s100);  // XY plane
        const bool c = (te < tn + dy * vref) & (tn < te + dx * vref);
        if constexpr (BF) {
            const double ta = te - tn + ten, tb = tn - te + ten, q = ta - tb;
            const double v = div_fixed((ta * dx2i + tb * dy2i) + sqrt_rn(c ? 4.0 * (vref * vref) * A.sxy - A.dx2dy2 * (q * q) : 1.0), A.sxy, A.rxy);
            t2d3 = c ? v : kBig;
        } else if (c) {
            const double ta = te - tn + ten, tb = tn - te + ten, q = ta - tb;
            t2d3 = div_fixed((ta * dx2i + tb * dy2i) + sqrt_rn(4.0 * (vref * vref) * A.sxy - A.dx2dy2 * (q * q)), A.sxy, A.rxy);
        }
    }
    const double t2d = fmin(fmin(t2d1, t2d2), t2d3);

    // 3D operator, _fteik3d.py:146-159
    double t3d = kBig;
    if (fmin(t1d, t2d) > fmax(fmax(tv, te), tn)) {
        const double vref = s000;
        // the reference's left-to-right sums te - 0.5 tn + 0.5 ten - ...: each halved term
        // is exact, so adding it in an FMA rounds the same sum once, as the separate
        // product and addition do (for times that are 0 or at least 2^-1021)
        const double ta = fma(0.5, tev, fma(-0.5, tv, fma(0.5, ten, fma(-0.5, tn, te)))) - tnv + tnve;
        const double tb = fma(0.5, tev, fma(-0.5, te, fma(0.5, tnv, fma(-0.5, tn, tv)))) - ten + tnve;
        const double tc = fma(0.5, tnv, fma(-0.5, tv, fma(0.5, ten, fma(-0.5, te, tn)))) - tev + tnve;
        const double t2 = vref * vref * A.dsum * 9.0;
        const double q1 = ta - tb, q2 = tb - tc, q3 = ta - tc;
        double t3 = A.dz2dx2 * (q1 * q1);
        t3 += A.dz2dy2 * (q2 * q2);
        t3 += A.dx2dy2 * (q3 * q3);
        // every lane of the branch takes the root (of 1.0 where the test fails), so the
        // test selects the result instead of opening another branch
        const bool ok3 = t2 >= t3;
        const double r3 = sqrt_rn(ok3 ? t2 - t3 : 1.0);
        const double v3 = div_fixed(tb * dz2i + ta * dx2i + tc * dy2i + r3, A.dsum, A.rsum);
        t3d = ok3 ? v3 : kBig;
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
