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

// Converts a HIP error into the library's negative return code.
inline int hip_status(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace fteik

#define FTEIK_CHECK(expr)                                   \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return ::fteik::hip_status(_e); \
    } while (0)
