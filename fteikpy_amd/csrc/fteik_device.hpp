// Shared device/host helpers of the fteik_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fteik {

// Initial traveltime and source nudge of the reference (fteikpy/_fteik/_fteik3d.py:6-7,
// _fteik2d.py:6-8).
constexpr double kBig = 1.0e5;
constexpr double kEps = 1.0e-15;
constexpr int kEpsIn = 5;

// Python's min()/max() keep the first argument unless a later one is strictly
// smaller/larger; with NaN operands that differs from fmin/fmax, so the 2D path (whose
// source initialisation can produce inf/NaN) uses these.
__host__ __device__ __forceinline__ double pymin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double pymax(double a, double b) { return (b > a) ? b : a; }

// Upwind sign of one axis packed in 2 bits (0 -> 0, 1 -> +1, 3 -> -1); axis a uses bits
// 2a..2a+1 of the node's byte.
__host__ __device__ __forceinline__ uint8_t sgn_code(int v) {
    return (uint8_t)(v > 0 ? 1 : (v < 0 ? 3 : 0));
}
__host__ __device__ __forceinline__ int sgn_axis(uint8_t s, int axis) {
    int v = (s >> (2 * axis)) & 3;
    return v == 3 ? -1 : v;
}

// Per-source placement computed on the host exactly as the reference does
// (fteikpy/_fteik/_fteik3d.py:409-422, _fteik2d.py:302-352).
struct SrcParam {
    double zsa, xsa, ysa;  // source position in grid units (after nudges / rounding)
    int zsi, xsi, ysi;     // source cell
    int iflag;             // 2D source classification (1, 2 or 3); unused in 3D
};

// a / b, correctly rounded, for a divisor b fixed per solve with y = 1 / b correctly
// rounded.  q0 = a y is within 2 ulp of a / b; one correction with the exact remainder
// a - b q0 (an FMA) brings it within 1 ulp, and by Markstein's theorem (y within half an
// ulp of 1 / b, q within 1 ulp of a / b) a second one, q + (a - b q) y rounded once, is the
// correctly rounded quotient.  Five dependent operations instead of the generic division's
// scaled reciprocal refinement.  Valid away from overflow and the subnormal range, which
// traveltimes and inverse squared spacings stay far from.  An infinite a gives NaN where
// a / b gives +-inf (the caller must not depend on the difference).
__device__ __forceinline__ double div_fixed(double a, double b, double y) {
    const double q0 = a * y;
    const double q1 = fma(fma(-b, q0, a), y, q0);
    return fma(fma(-b, q1, a), y, q1);
}

// sqrt(x), correctly rounded, with the compiler's own refinement sequence for doubles
// without its range handling: for x in [2^-767, +inf) the library first scales only
// below 2^-767 and special-cases only 0 and +inf, so there this is the same sequence on
// the same value (rsq estimate, then the Goldschmidt/Newton steps) and returns the same
// double.  Any other argument (0, tiny, infinite, negative, NaN) takes the library sqrt; the
// branch is uniform over the wave, and the node update produces such arguments only as
// exact zeros.
__device__ __forceinline__ double sqrt_rn(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    const bool other = !(x >= 0x1p-767 && x < __builtin_inf());
    if (__builtin_expect(__any(other), 0)) {
        if (other) g = sqrt(x);
    }
    return g;
}

// sqrt_rn for code run by a single lane of its wave: the same values, with a plain branch
// (no wave vote, so loops around it stay unrollable).
__device__ __forceinline__ double sqrt_rn_lane(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    if (__builtin_expect(!(x >= 0x1p-767 && x < __builtin_inf()), 0)) g = sqrt(x);
    return g;
}

// Orders this wave's memory accesses across the call for the compiler; no hardware wait.
// A wave's LDS instructions execute in order, so after it every lane of the wave sees what
// the others wrote to LDS before it (the wave-synchronous exchange rocPRIM's wave_barrier
// provides).  Called by every lane of the wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Converts a HIP error into the library's negative return code.
inline int hip_status(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace fteik

#define FTEIK_CHECK(expr)                                   \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return ::fteik::hip_status(_e); \
    } while (0)
