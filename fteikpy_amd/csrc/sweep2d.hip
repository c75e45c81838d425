// 2D fast-sweeping eikonal solver for MI355X (gfx950): FTeik 2D operators with the
// near-source spherical (factored) operator (fteikpy/_fteik/_fteik2d.py:33-187), the
// off-grid source line initialisation (_fteik2d.py:355-617) and the column-interleaved
// 4-direction sweep set (_fteik2d.py:190-287).
//
// The reference's 2D set is one serial dependency chain: column j's +z pass reads column
// j-1 after its -z pass, and within a pass each node reads the one before.  Exact parity
// at any nsweep needs the results of that order: each source is swept by one wave, which
// stages the two active columns in LDS and evaluates each z pass in speculative 64-node
// blocks that keep only provably exact results, finishing dense stretches in order on one
// lane (models too deep for LDS stage the same rows in global memory); sources run in
// parallel.
#include "fteik_amd.h"
#include "fteik_device.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fteik {
namespace {

struct Grid2 {
    const double *slow;  // [nz][nx] cells
    double *tt;          // [NZ][NX] nodes
    uint8_t *sgn;        // [NZ][NX] or null
    int nz, nx, NZ, NX;
    double dz, dx, dzi, dxi, dz2i, dx2i;
    double ssum, rsum;  // dz2i + dx2i (the 4-point operator's divisor), 1 / ssum rounded
    __device__ __forceinline__ double &T(int i, int j) const { return tt[(long long)i * NX + j]; }
    __device__ __forceinline__ double S(int i, int j) const { return slow[(long long)i * nx + j]; }
    __device__ __forceinline__ void setsgn(int i, int j, int sz, int sx) const {
        if (sgn) sgn[(long long)i * NX + j] = (uint8_t)(sgn_code(sz) | (sgn_code(sx) << 2));
    }
};

// _fteik2d.py:11-14
__device__ __forceinline__ double tana(int i, int j, double dz, double dx, double zsa, double xsa,
                                       double vzero) {
    const double a = dz * ((double)i - zsa), b = dx * ((double)j - xsa);
    return vzero * sqrt(a * a + b * b);
}

// _fteik2d.py:17-30: analytic time and its derivatives
__device__ __forceinline__ double tanad(int i, int j, double dz, double dx, double zsa,
                                        double xsa, double vzero, double *tzc, double *txc) {
    const double t = tana(i, j, dz, dx, zsa, xsa, vzero);
    if (t > 0.0) {
        const double tmp = vzero * vzero / t;
        *tzc = ((double)i - zsa) * dz * tmp;
        *txc = ((double)j - xsa) * dx * tmp;
    } else {
        *tzc = 0.0;
        *txc = 0.0;
    }
    return t;
}

// Factored quadratic of the spherical operator, _fteik2d.py:33-67.
__device__ __forceinline__ double delta(double t1, double tauv, double taue, double tauev,
                                        double t0c, double tzc, double txc, double dzi,
                                        double dxi, double dz2i, double dx2i, double vzero,
                                        double vref, int sgntz, int sgntx) {
    const double ta = tauev + taue - tauv;
    const double tb = tauev - taue + tauv;
    const double apoly = dz2i + dx2i;
    const double bpoly = 4.0 * ((double)sgntx * txc * dxi + (double)sgntz * tzc * dzi) -
                         2.0 * (ta * dx2i + tb * dz2i);
    const double cpoly = ((ta * ta) * dx2i) + ((tb * tb) * dz2i) -
                         4.0 * ((double)sgntx * txc * dxi * ta + (double)sgntz * tzc * dzi * tb) +
                         4.0 * (vzero * vzero - vref * vref);
    const double dpoly = bpoly * bpoly - 4.0 * apoly * cpoly;
    return dpoly >= 0.0 ? 0.5 * (sqrt(dpoly) - bpoly) / apoly + t0c : t1;
}

// The parts of a 2D node update that do not depend on the upwind z neighbour tv, from the
// reference's operands (_fteik2d.py:97-132): te / tev the upwind neighbours along x and on
// the diagonal; sza, szb the cells of the z 1D operator (slow[i1, max(j-1, 0)] and
// slow[i1, min(j, nx-1)]); sxa, sxb those of the x 1D operator (slow[max(i-1, 0), j1] and
// slow[min(i, nz-1), j1]); vref the 2D operator's cell slow[i1, j1].  Each value is the
// reference's own subexpression, evaluated in its order (stage2 computes them).
struct Node2Pre {
    double cz;    // dz * min(sza, szb): t1d1 = tv + cz
    double t1d2;  // te + dx * min(sxa, sxb)
    double A;     // te + dx * vref where te >= tev, else NaN (the 4-point test then fails)
    double B;     // dz * vref
    double K;     // 4 vref^2 (dz^-2 + dx^-2), the 4-point root's leading term
};

__device__ __forceinline__ double nan2() { return __builtin_nan(""); }

// 2D node update, _fteik2d.py:70-187: tv, te, tev the upwind neighbours, t0 the node's
// current time, q its Node2Pre (computed by stage2), vref the 2D operator's cell; far: the
// node is more than 5 nodes from the source cell along z or x (else the spherical operator
// applies).  Returns the new time; *sg receives which upwind signs to record: 0 unchanged,
// 1 z only, 2 x only, 3 both.
__device__ __forceinline__ double node2_core(const Grid2 &G, const SrcParam &P, double vzero,
                                             int i, int j, bool far, int sgntz, int sgntx,
                                             double tv, double te, double tev, double t0,
                                             const Node2Pre &q, double vref, int *sg) {
    const double dz = G.dz, dx = G.dx;
    const double t1d1 = tv + q.cz;
    const double t1d = pymin(t1d1, q.t1d2);

    double t2d = kBig;
    if (far) {
        // plane-wave: 4-point operator (tv <= te + dx vref, te <= tv + dz vref, te >= tev,
        // tv >= tev), else one of two 3-point operators.  The 4-point quotient and root
        // are the correctly rounded ones (div_fixed, sqrt_rn_lane), as the reference's /
        // and ** 0.5.
        if (tv <= q.A && te <= tv + q.B && tv >= tev) {
            const double ta = tev + te - tv, tb = tev - te + tv, d = ta - tb;
            t2d = div_fixed((tb * G.dz2i + ta * G.dx2i) +
                                sqrt_rn_lane(q.K - G.dz2i * G.dx2i * (d * d)),
                            G.ssum, G.rsum);
        } else if (te - tev <= dz * dz * vref / sqrt(dx * dx + dz * dz) && te - tev > 0.0) {
            const double r = (te - tev) / dz;
            t2d = te + dx * sqrt(vref * vref - r * r);
        } else if (tv - tev <= dx * dx * vref / sqrt(dx * dx + dz * dz) && tv - tev > 0.0) {
            const double r = (tv - tev) / dx;
            t2d = tv + dz * sqrt(vref * vref - r * r);
        }
    } else {
        // spherical operator near the source
        if (tv < te + dx * vref && te < tv + dz * vref && te >= tev && tv >= tev) {
            double tzc, txc;
            const double t0c = tanad(i, j, dz, dx, P.zsa, P.xsa, vzero, &tzc, &txc);
            const double tauv = tv - tana(i - sgntz, j, dz, dx, P.zsa, P.xsa, vzero);
            const double taue = te - tana(i, j - sgntx, dz, dx, P.zsa, P.xsa, vzero);
            const double tauev = tev - tana(i - sgntz, j - sgntx, dz, dx, P.zsa, P.xsa, vzero);
            t2d = delta(t2d, tauv, taue, tauev, t0c, tzc, txc, G.dzi, G.dxi, G.dz2i, G.dx2i,
                        vzero, vref, sgntz, sgntx);
            if (t2d < tv || t2d < te) t2d = kBig;
        }
    }
    const double tnew = pymin(pymin(t0, t1d), t2d);
    *sg = tnew == t0 ? 0 : (tnew == t1d1 ? 1 : (tnew == q.t1d2 ? 2 : 3));
    return tnew;
}

// Upwind signs of a changed node, _fteik2d.py:176-187.
// Called only where signs are recorded (G.sgn set); sg's bit 0 = z, bit 1 = x take part.
__device__ __forceinline__ void record2(const Grid2 &G, int i, int j, int sg, int sgntz,
                                        int sgntx) {
    if (sg == 0) return;
    const int sz = (sg & 1) ? sgntz : 0, sx = (sg & 2) ? sgntx : 0;
    G.sgn[(long long)i * G.NX + j] = (uint8_t)(sgn_code(sz) | (sgn_code(sx) << 2));
}

// Staged layout of the 2D sweep (in LDS, or in global memory for a model too deep for it):
// rows r = 0 .. NZ of kRow2 doubles, row r for node i = r - 1 and cell
// k = clamp(r - 1, 0, nz - 1) of the current column j:
//   0, 1  node i of the previous and of the current node column (roles swap per column)
//   2     dz * min(slow[k, max(j-1, 0)], slow[k, min(j, nx-1)])   (Node2Pre::cz)
//   3     slow[k, j1], j1 = j - sgnvx                            (vref)
//   4     dz * slow[k, j1]                                       (Node2Pre::B)
//   5     4 slow[k, j1]^2 (dz^-2 + dx^-2)                        (Node2Pre::K)
//   6     Node2Pre::t1d2 of node i (the same in both z passes)
//   7, 8  Node2Pre::A of node i in the +z and in the -z pass
// The +z pass reads node i's operands from rows i and i+1, the -z pass from rows i+1 and
// i+2, each at a fixed slot: one address register with immediate offsets.
constexpr int kRow2 = 9;

// Bytes of the staged rows for an nz-cell-deep model.
__host__ __device__ __forceinline__ long long lds2_bytes(int nz) {
    return (long long)(nz + 2) * kRow2 * (long long)sizeof(double);
}

// Operands of node i in one z pass (DZ = +1: the +z pass, rows i and i+1 of the staged
// layout; DZ = -1: the -z pass, rows i+1 and i+2), the previous column in slot PV.
struct Ops2 {
    double tev, te, t0, vref;
    Node2Pre q;
};
template <int DZ, int PV>
__device__ __forceinline__ Ops2 ops2(const double *lds, int i) {
    constexpr int R = kRow2, CU = 1 - PV;
    Ops2 o;
    if (DZ > 0) {
        const double *b = lds + R * i;
        o.tev = b[PV]; o.q.cz = b[2]; o.vref = b[3]; o.q.B = b[4]; o.q.K = b[5];
        o.te = b[R + PV]; o.t0 = b[R + CU]; o.q.t1d2 = b[R + 6]; o.q.A = b[R + 7];
    } else {
        const double *b = lds + R * (i + 1);
        o.te = b[PV]; o.t0 = b[CU]; o.q.cz = b[2]; o.vref = b[3]; o.q.B = b[4]; o.q.K = b[5];
        o.q.t1d2 = b[6]; o.q.A = b[8]; o.tev = b[R + PV];
    }
    return o;
}

// cnt consecutive nodes of one z pass from node i on (DZ the direction), in order, by one
// lane: the upwind z neighbour tv is the previous update's result, kept in a register, and
// the next node's operands are read before the current update.  tv: in, the final time of
// the node before i; out, that of the last node.  Returns nonzero if a node changed.
template <int SX, int PV, bool GRAD, int DZ, bool NEAR = true>
__device__ __forceinline__ int serial2(const Grid2 &G, const SrcParam &P, double vzero, int j,
                                       int nlo, double *lds, int i, int cnt, double &tv) {
    constexpr int CU = 1 - PV;
    int ch = 0, sg;
    Ops2 o = ops2<DZ, PV>(lds, i);
#pragma unroll 2
    for (int k = 0; k < cnt; ++k, i += DZ) {
        const Ops2 on = ops2<DZ, PV>(lds, k + 1 < cnt ? i + DZ : i);  // the last reads itself
        // NEAR false: the caller has checked that no node of the run is near the source
        const bool far = !NEAR || (unsigned)(i - nlo) > 2u * kEpsIn;
        const double tnew = node2_core(G, P, vzero, i, j, far, DZ, SX, tv, o.te, o.tev, o.t0,
                                       o.q, o.vref, &sg);
        lds[kRow2 * (i + 1) + CU] = tnew;
        if (GRAD) record2(G, i, j, sg, DZ, SX);
        ch |= sg;
        tv = tnew;
        o = on;
    }
    return ch;
}

// Speculative evaluation of n <= 64 consecutive nodes i0, i0 + DZ, ... of one z pass, node
// k on lane k; all lanes.  Within a pass node i's new time is x_i = F_i(x_{i-DZ}) with
// every other operand fixed, so the pass's results are the unique solution of that
// recurrence from the exact x before i0 (tv).  Iteration t evaluates lane k on the previous
// iteration's output of lane k-1 (lane 0 on tv; before the first iteration the outputs are
// the nodes' current times).  A lane's output is exact once its input is: if lanes 0..e-1
// are exact, lane e is exact after the next iteration, and so is every following lane up
// to the first whose output differs (bitwise) from the previous iteration's, since its
// input then already was its predecessor's exact output.  Each iteration thus finalizes at
// least one more node, at most 64 iterations; it stops once fewer than 4 nodes were
// finalized in an iteration (a dense dependency chain, cheaper on one lane).  The exact
// prefix is written back; returns its length in *done and its last time in tv, and
// nonzero on lanes whose node changed.
template <int SX, int PV, bool GRAD, int DZ, bool NEAR = true>
__device__ __forceinline__ int spec2(const Grid2 &G, const SrcParam &P, double vzero, int j,
                                     int nlo, double *lds, int i0, int n, double &tv,
                                     int *done) {
    constexpr int CU = 1 - PV, R = kRow2;
    const int lane = threadIdx.x;
    const int i = i0 + DZ * (lane < n ? lane : n - 1);
    const Ops2 o = ops2<DZ, PV>(lds, i);
    // current time of the upwind node (+z: node i-1, row i; -z: node i+1, row i+2)
    const double up = lds[R * (DZ > 0 ? i : i + 2) + CU];
    double g = lane == 0 ? tv : up, yp = o.t0, y;
    // NEAR false: the caller has checked that no node of the block is near the source
    const bool far = !NEAR || (unsigned)(i - nlo) > 2u * kEpsIn;
    int e = 0, sg;
    while (true) {
        y = node2_core(G, P, vzero, i, j, far, DZ, SX, g, o.te, o.tev, o.t0, o.q, o.vref, &sg);
        const bool changed =
            lane < n && __double_as_longlong(y) != __double_as_longlong(yp);
        const unsigned long long m = __ballot(changed) & (~0ull << e);
        const int en = m ? __ffsll((unsigned long long)m) : n;  // first changed lane + 1
        const int adv = en - e;
        e = en;
        if (e >= n || adv < 4) break;
        yp = y;
        const double s = __shfl_up(y, 1);
        if (lane > 0) g = s;
    }
    int ch = 0;
    if (lane < e) {
        lds[R * (i + 1) + CU] = y;
        if (GRAD) record2(G, i, j, sg, DZ, SX);
        ch = sg;
    }
    tv = __shfl(y, e - 1);
    *done = e;
    return ch;
}

// The +z and -z passes over node column j of one x direction (SX = sgntx), the previous
// column (final for this half set) in slot PV, column j in slot 1 - PV; _fteik2d.py:199-287.
// All lanes: each block of up to 64 nodes is first evaluated speculatively (spec2), and
// lane 0 sweeps what speculation left in order (serial2).  Returns nonzero on lanes that
// changed a node.
template <int SX, int PV, bool GRAD>
__device__ __forceinline__ int column2(const Grid2 &G, const SrcParam &P, double vzero, int j,
                                       double *lds) {
    constexpr int CU = 1 - PV;
    const int NZ = G.NZ;
    // node i is far from the source (node2_core) unless (unsigned)(i - nlo) <= 2 kEpsIn
    const int nlo = abs(j - P.xsi) > kEpsIn ? (1 << 30) : P.zsi - kEpsIn;
    int ch = 0;
    double tv = lds[kRow2 + CU];  // node 0
    for (int i0 = 1; i0 < NZ;) {  // +z: nodes 1 .. NZ-1
        const int n = min(64, NZ - i0);
        int e;
        if (i0 + n - 1 < nlo || i0 > nlo + 2 * kEpsIn)
            ch |= spec2<SX, PV, GRAD, 1, false>(G, P, vzero, j, nlo, lds, i0, n, tv, &e);
        else
            ch |= spec2<SX, PV, GRAD, 1>(G, P, vzero, j, nlo, lds, i0, n, tv, &e);
        if (e < n && threadIdx.x == 0) {
            const int a = i0 + e, b = i0 + n - 1;  // rows of the run, ascending
            if (b < nlo || a > nlo + 2 * kEpsIn)
                ch |= serial2<SX, PV, GRAD, 1, false>(G, P, vzero, j, nlo, lds, a, n - e, tv);
            else
                ch |= serial2<SX, PV, GRAD, 1>(G, P, vzero, j, nlo, lds, a, n - e, tv);
        }
        tv = __shfl(tv, 0);
        __syncthreads();
        i0 += n;
    }
    for (int i0 = NZ - 2; i0 >= 0;) {  // -z: nodes NZ-2 .. 0, tv = node NZ-1 from +z
        const int n = min(64, i0 + 1);
        int e;
        if (i0 < nlo || i0 - (n - 1) > nlo + 2 * kEpsIn)
            ch |= spec2<SX, PV, GRAD, -1, false>(G, P, vzero, j, nlo, lds, i0, n, tv, &e);
        else
            ch |= spec2<SX, PV, GRAD, -1>(G, P, vzero, j, nlo, lds, i0, n, tv, &e);
        if (e < n && threadIdx.x == 0) {
            const int a = i0 - (n - 1), b = i0 - e;  // rows of the run, ascending
            if (b < nlo || a > nlo + 2 * kEpsIn)
                ch |= serial2<SX, PV, GRAD, -1, false>(G, P, vzero, j, nlo, lds, i0 - e, n - e,
                                                       tv);
            else
                ch |= serial2<SX, PV, GRAD, -1>(G, P, vzero, j, nlo, lds, i0 - e, n - e, tv);
        }
        tv = __shfl(tv, 0);
        __syncthreads();
        i0 -= n;
    }
    return ch;
}

// Stages node column j (slot 1 - pv) with the per-row values of the layout above, the
// previous column already in slot pv; all lanes (lane l always moves rows l, l+64, ...).
template <int SX>
__device__ __forceinline__ void stage2(const Grid2 &G, int j, int pv, double *lds) {
    constexpr int R = kRow2;
    const int cu = 1 - pv, NZ = G.NZ;
    const int jl = max(j - 1, 0), jr = min(j, G.nx - 1);
#pragma unroll 4
    for (int r = threadIdx.x; r <= NZ; r += 64) {
        const int k = min(max(r - 1, 0), G.nz - 1);
        const double sl = G.S(k, jl), sr = G.S(k, jr), v = SX > 0 ? sl : sr;
        double *w = lds + R * r;
        if (r > 0) w[cu] = G.T(r - 1, j);
        w[2] = G.dz * pymin(sl, sr);
        w[3] = v;
        w[4] = G.dz * v;
        w[5] = 4.0 * (v * v) * G.ssum;
    }
    __syncthreads();
    for (int r = threadIdx.x; r <= NZ; r += 64) {
        if (r == 0) continue;
        double *w = lds + R * r;
        const double te = w[pv];
        // x 1D operator: cells max(i-1, 0) (row r-1) and min(i, nz-1) (row r)
        w[6] = te + G.dx * pymin(w[3 - R], w[3]);
        if (r >= 2) {  // +z (node i >= 1): tev = previous column at i-1, vref = cell i-1
            const double tev = w[pv - R];
            w[7] = te >= tev ? te + G.dx * w[3 - R] : nan2();
        }
        if (r < NZ) {  // -z (node i <= NZ-2): tev = previous column at i+1, vref = cell i
            const double tev = w[pv + R];
            w[8] = te >= tev ? te + G.dx * w[3] : nan2();
        }
    }
}

// Slot cu of the staged rows back to node column j; all lanes, the rows of stage2.
__device__ __forceinline__ void unstage2(const Grid2 &G, int j, int cu, const double *lds) {
    for (int r = threadIdx.x; r <= G.NZ; r += 64)
        if (r > 0) G.T(r - 1, j) = lds[kRow2 * r + cu];
}

template <int SX, bool GRAD>
__device__ __forceinline__ int column2_pv(const Grid2 &G, const SrcParam &P, double vzero, int j,
                                          int pv, double *lds) {
    return pv ? column2<SX, 1, GRAD>(G, P, vzero, j, lds) : column2<SX, 0, GRAD>(G, P, vzero, j, lds);
}

// Sweep sets of one source with the two active node columns staged at `lds` (LDS or
// global scratch, lds2_bytes(nz)), in the reference's order, _fteik2d.py:190-287, with the
// results of that order (column2).  The wave's 64 lanes move columns between the grid and the rows and
// precompute the values that do not depend on the serial chain.  Returns the number of
// sets run.
template <bool GRAD>
__device__ __forceinline__ int sets2_lds(const Grid2 &G, const SrcParam &P, double vzero,
                                         int cap, double *lds) {
    const int NX = G.NX;
    int pv = 0;  // slot of the previous column
    for (int r = threadIdx.x; r <= G.NZ; r += 64)
        if (r > 0) lds[kRow2 * r + pv] = G.T(r - 1, 0);
    int done = 0;
    for (int set = 0; set < cap; ++set) {
        int ch = 0;
        for (int j = 1; j < NX; ++j) {  // +x: previous = column j-1
            __syncthreads();
            stage2<1>(G, j, pv, lds);
            __syncthreads();
            ch |= column2_pv<1, GRAD>(G, P, vzero, j, pv, lds);
            __syncthreads();
            unstage2(G, j, 1 - pv, lds);
            pv = 1 - pv;
        }
        for (int j = NX - 2; j >= 0; --j) {  // -x: previous = column j+1
            __syncthreads();
            stage2<-1>(G, j, pv, lds);
            __syncthreads();
            ch |= column2_pv<-1, GRAD>(G, P, vzero, j, pv, lds);
            __syncthreads();
            unstage2(G, j, 1 - pv, lds);
            pv = 1 - pv;
        }
        ++done;
        // a set that changed nothing leaves every later set a no-op
        if (!__syncthreads_or(ch)) break;
    }
    return done;
}

__global__ void fill2d_kernel(double *tt, long long n, uint8_t *sgn) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        tt[i] = kBig;
        if (sgn) sgn[i] = 0;
    }
}

// Source initialisation (_fteik2d.py:355-617, by lane 0) and the sweep sets of source s
// (sets2_lds, by the whole wave) with lds2_bytes(nz) of staged rows at `rows`: the
// workgroup's LDS, or for a model too deep for it a per-source area of global memory (the
// same code; each caller passes one kind, so the accesses keep their address space).
__device__ __forceinline__ void solve_source2(const double *slow, int nz, int nx, double dz, double dx,
                              const SrcParam &P, int s, int nsweep, int max_sets,
                              double *tt_all, uint8_t *sgn_all, double *vzero_out,
                              int *sets_done, double *lds) {
    // lds: the staged rows (see above)
    Grid2 G;
    G.slow = slow;
    G.nz = nz;
    G.nx = nx;
    G.NZ = nz + 1;
    G.NX = nx + 1;
    const long long nn = (long long)G.NZ * G.NX;
    G.tt = tt_all + (long long)s * nn;
    G.sgn = sgn_all ? sgn_all + (long long)s * nn : nullptr;
    G.dz = dz;
    G.dx = dx;
    const double vzero = slow[(long long)P.zsi * nx + P.xsi];
    const int zsi = P.zsi, xsi = P.xsi, NZ = G.NZ, NX = G.NX;
    const double zsa = P.zsa, xsa = P.xsa;

    if (threadIdx.x != 0) {
        // source initialisation: lane 0 only
    } else if (P.iflag == 2) {
        const double dzu = fabs(zsa - (double)zsi), dzd = 1.0 - dzu;
        const double dxw = fabs(xsa - (double)xsi), dxe = 1.0 - dxw;
        for (int c = 0; c < 4; ++c) {  // corners (0,0) (1,0) (0,1) (1,1)
            const int i = zsi + (c & 1), j = xsi + (c >> 1);
            double tzc, txc;
            G.T(i, j) = tanad(i, j, dz, dx, zsa, xsa, vzero, &tzc, &txc);
        }
        double dxi = 1.0 / dx, dx2i = dxi / dx, dzi, dz2i, tzc, txc, t0c;
        // +x along rows zsi+1 and zsi (the 1D time td runs along row zsi's cells)
        double td_prev = vzero * dxe * dx;
        for (int j = xsi + 2; j < NX; ++j) {
            const double vref = G.S(zsi, j - 1);
            const double td = td_prev + dx * vref;
            const double tauv = td - vzero * fabs((double)j - xsa) * dx;
            const double tauev = td_prev - vzero * fabs((double)j - xsa - 1.0) * dx;
            dzi = 1.0 / dzd;
            dz2i = dz / dzd / dzd;
            double taue = G.T(zsi + 1, j - 1) - tana(zsi + 1, j - 1, dz, dx, zsa, xsa, vzero);
            t0c = tanad(zsi + 1, j, dz, dx, zsa, xsa, vzero, &tzc, &txc);
            G.T(zsi + 1, j) = delta(G.T(zsi + 1, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, 1, 1);
            G.setsgn(zsi + 1, j, 1, 1);
            if (dzu > 0.0) {
                dzi = 1.0 / dzu;
                dz2i = dz / dzu / dzu;
                taue = G.T(zsi, j - 1) - tana(zsi, j - 1, dz, dx, zsa, xsa, vzero);
                t0c = tanad(zsi, j, dz, dx, zsa, xsa, vzero, &tzc, &txc);
                G.T(zsi, j) = delta(G.T(zsi, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, -1, 1);
                G.setsgn(zsi, j, -1, 1);
            }
            td_prev = td;
        }
        // -x
        td_prev = vzero * dxw * dx;
        for (int j = xsi - 1; j >= 0; --j) {
            const double vref = G.S(zsi, j);
            const double td = td_prev + dx * vref;
            const double tauv = td - vzero * fabs((double)j - xsa) * dx;
            const double tauev = td_prev - vzero * fabs((double)j - xsa + 1.0) * dx;
            dzi = 1.0 / dzd;
            dz2i = dz / dzd / dzd;
            double taue = G.T(zsi + 1, j + 1) - tana(zsi + 1, j + 1, dz, dx, zsa, xsa, vzero);
            t0c = tanad(zsi + 1, j, dz, dx, zsa, xsa, vzero, &tzc, &txc);
            G.T(zsi + 1, j) = delta(G.T(zsi + 1, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, 1, -1);
            G.setsgn(zsi + 1, j, 1, -1);
            if (dzu > 0.0) {
                dzi = 1.0 / dzu;
                dz2i = dz / dzu / dzu;
                // row zsi+1 is read here too (_fteik2d.py:472-474)
                taue = G.T(zsi + 1, j + 1) - tana(zsi + 1, j + 1, dz, dx, zsa, xsa, vzero);
                t0c = tanad(zsi, j, dz, dx, zsa, xsa, vzero, &tzc, &txc);
                G.T(zsi, j) = delta(G.T(zsi, j), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, -1, -1);
                G.setsgn(zsi, j, -1, -1);
            }
            td_prev = td;
        }
        // +z along columns xsi+1 and xsi
        dzi = 1.0 / dz;
        dz2i = dzi / dz;
        td_prev = vzero * dzd * dz;
        for (int i = zsi + 2; i < NZ; ++i) {
            const double vref = G.S(i - 1, xsi);
            const double td = td_prev + dz * vref;
            const double taue = td - vzero * fabs((double)i - zsa) * dz;
            const double tauev = td_prev - vzero * fabs((double)i - zsa - 1.0) * dz;
            dxi = 1.0 / dxe;
            dx2i = dx / dxe / dxe;
            double tauv = G.T(i - 1, xsi + 1) - tana(i - 1, xsi + 1, dz, dx, zsa, xsa, vzero);
            t0c = tanad(i, xsi + 1, dz, dx, zsa, xsa, vzero, &tzc, &txc);
            G.T(i, xsi + 1) = delta(G.T(i, xsi + 1), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, 1, 1);
            G.setsgn(i, xsi + 1, 1, 1);
            if (dxw > 0.0) {
                dxi = 1.0 / dxw;
                dx2i = dx / dxw / dxw;
                tauv = G.T(i - 1, xsi) - tana(i - 1, xsi, dz, dx, zsa, xsa, vzero);
                t0c = tanad(i, xsi, dz, dx, zsa, xsa, vzero, &tzc, &txc);
                G.T(i, xsi) = delta(G.T(i, xsi), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, 1, -1);
                G.setsgn(i, xsi, 1, -1);
            }
            td_prev = td;
        }
        // -z
        td_prev = vzero * dzu * dz;
        for (int i = zsi - 1; i >= 0; --i) {
            const double vref = G.S(i, xsi);
            const double td = td_prev + dz * vref;
            const double taue = td - vzero * fabs((double)i - zsa) * dz;
            const double tauev = td_prev - vzero * fabs((double)i - zsa + 1.0) * dz;
            dxi = 1.0 / dxe;
            dx2i = dx / dxe / dxe;
            double tauv = G.T(i + 1, xsi + 1) - tana(i + 1, xsi + 1, dz, dx, zsa, xsa, vzero);
            t0c = tanad(i, xsi + 1, dz, dx, zsa, xsa, vzero, &tzc, &txc);
            G.T(i, xsi + 1) = delta(G.T(i, xsi + 1), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi,
                                    dz2i, dx2i, vzero, vref, -1, 1);
            G.setsgn(i, xsi + 1, -1, 1);
            if (dxw > 0.0) {
                dxi = 1.0 / dxw;
                dx2i = dx / dxw / dxw;
                tauv = G.T(i + 1, xsi) - tana(i + 1, xsi, dz, dx, zsa, xsa, vzero);
                t0c = tanad(i, xsi, dz, dx, zsa, xsa, vzero, &tzc, &txc);
                G.T(i, xsi) = delta(G.T(i, xsi), tauv, taue, tauev, t0c, tzc, txc, dzi, dxi, dz2i,
                                    dx2i, vzero, vref, -1, -1);
                G.setsgn(i, xsi, -1, -1);
            }
            td_prev = td;
        }
    } else {
        G.T((int)zsa, (int)xsa) = 0.0;
    }

    // sweep sets, _fteik2d.py:190-287 and :619-621
    G.dzi = 1.0 / dz;
    G.dxi = 1.0 / dx;
    G.dz2i = G.dzi / dz;
    G.dx2i = G.dxi / dx;
    G.ssum = G.dz2i + G.dx2i;
    G.rsum = 1.0 / G.ssum;
    const int cap = nsweep > 0 ? nsweep : max_sets;
    __syncthreads();  // lane 0's initialisation before the wave reads column 0
    const int done = G.sgn ? sets2_lds<true>(G, P, vzero, cap, lds)
                           : sets2_lds<false>(G, P, vzero, cap, lds);
    if (threadIdx.x == 0) {
        vzero_out[s] = vzero;
        sets_done[s] = done;
    }
}

// One 64-lane wave per source (block s).  scratch null: the staged rows fit the launch's
// dynamic LDS; else scratch holds lds2_bytes(nz) per source in global memory.  The bound
// lets the compiler keep the whole per-source chain in registers.
__global__ __launch_bounds__(64) void fteik2d_kernel(const double *slow, int nz, int nx, double dz,
                                                     double dx,
                               const SrcParam *prm, int nsrc, int nsweep, int max_sets,
                               double *tt_all, uint8_t *sgn_all, double *vzero_out,
                               int *sets_done, double *scratch) {
    extern __shared__ double lds2[];
    const int s = blockIdx.x;
    if (s >= nsrc) return;
    const SrcParam P = prm[s];  // in registers: the chain's stores cannot alias it
    if (!scratch)
        solve_source2(slow, nz, nx, dz, dx, P, s, nsweep, max_sets, tt_all, sgn_all, vzero_out,
                      sets_done, lds2);
    else
        solve_source2(slow, nz, nx, dz, dx, P, s, nsweep, max_sets, tt_all, sgn_all, vzero_out,
                      sets_done, scratch + (long long)s * (lds2_bytes(nz) / sizeof(double)));
}

// Batch of 2D models: block t takes source t - cum[m] of the model m with
// cum[m] <= t < cum[m+1].
struct Prob2 {
    const double *slow;
    int nz, nx;
    double dz, dx;
    const SrcParam *prm;
    double *tt;
    uint8_t *sgn;
    double *vzero;
    int *sets_done;
    double *scratch;  // staged rows of the model's sources in global memory, or null (LDS)
};
// The kernel's view of the same table: every pointer addresses device global memory
// (address space 1), which lets the inlined chain use global instead of flat instructions.
#define FTEIK_GLOBAL __attribute__((address_space(1)))
struct Prob2Dev {
    const FTEIK_GLOBAL double *slow;
    int nz, nx;
    double dz, dx;
    const FTEIK_GLOBAL SrcParam *prm;
    FTEIK_GLOBAL double *tt;
    FTEIK_GLOBAL uint8_t *sgn;
    FTEIK_GLOBAL double *vzero;
    FTEIK_GLOBAL int *sets_done;
    FTEIK_GLOBAL double *scratch;
};
static_assert(sizeof(Prob2Dev) == sizeof(Prob2), "same table layout");

// Models with a scratch area stage their rows there, the others in the launch's LDS.
__global__ __launch_bounds__(64) void fteik2d_batch_kernel(const Prob2 *probs, const int *cum,
                                                           int nmodel, int nsweep, int max_sets) {
    extern __shared__ double lds2[];
    const int t = blockIdx.x;
    if (t >= cum[nmodel]) return;
    int lo = 0, hi = nmodel - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cum[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const Prob2Dev Q = reinterpret_cast<const Prob2Dev *>(probs)[lo];
    const int s = t - cum[lo];
    const SrcParam P = ((const SrcParam *)Q.prm)[s];  // read once, before the chain
    if (!Q.scratch)
        solve_source2((const double *)Q.slow, Q.nz, Q.nx, Q.dz, Q.dx, P, s, nsweep, max_sets,
                      (double *)Q.tt, (uint8_t *)Q.sgn, (double *)Q.vzero, (int *)Q.sets_done,
                      lds2);
    else
        solve_source2((const double *)Q.slow, Q.nz, Q.nx, Q.dz, Q.dx, P, s, nsweep, max_sets,
                      (double *)Q.tt, (uint8_t *)Q.sgn, (double *)Q.vzero, (int *)Q.sets_done,
                      (double *)(Q.scratch + (long long)s * (lds2_bytes(Q.nz) / sizeof(double))));
}

// 2D gradient post-pass, _fteik2d.py:623-639 (seeded analytic components survive where
// no upwind sign was recorded).
__global__ void gradient2d_kernel(const double *tt, const uint8_t *sgn, int NZ, int NX, double dz,
                                  double dx, SrcParam P, double vzero, double *g) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (long long)NZ * NX) return;
    const int j = (int)(n % NX), i = (int)(n / NX);
    double gz = 0.0, gx = 0.0;
    if (P.iflag == 2 && i - P.zsi >= 0 && i - P.zsi <= 1 && j - P.xsi >= 0 && j - P.xsi <= 1)
        tanad(i, j, dz, dx, P.zsa, P.xsa, vzero, &gz, &gx);
    const uint8_t s = sgn[n];
    const int sz = sgn_axis(s, 0), sx = sgn_axis(s, 1);
    if (sz != 0) gz = sz * (tt[n] - tt[n - (long long)sz * NX]) / dz;
    if (sx != 0) gx = sx * (tt[n] - tt[n - sx]) / dx;
    const double gn = sqrt(gz * gz + gx * gx);
    if (gn > 0.0) {
        gz /= gn;
        gx /= gn;
    }
    g[2 * n] = gz;
    g[2 * n + 1] = gx;
}

// Host: bound check and source classification, _fteik2d.py:296-352.
bool place_source2(const double *s, int nz, int nx, double dz, double dx, SrcParam *P) {
    const double z = s[0], x = s[1];
    if (!((0.0 <= z && z <= dz * nz) && (0.0 <= x && x <= dx * nx))) return false;
    double zsa = z / dz, xsa = x / dx;
    if (zsa >= nz) zsa = zsa - kEps;
    if (xsa >= nx) xsa = xsa - kEps;
    const int zsi = std::min((int)zsa, nz - 1), xsi = std::min((int)xsa, nx - 1);
    const double dzu = std::fabs(zsa - (double)zsi), dzd = 1.0 - dzu;
    const double dxw = std::fabs(xsa - (double)xsi), dxe = 1.0 - dxw;
    const double dzv_min = pymin(dzu, dzd), dzh_min = pymin(dxw, dxe);
    int iflag;
    if (dzv_min < kEps && dzh_min < kEps) {
        zsa = std::nearbyint(zsa);
        xsa = std::nearbyint(xsa);
        iflag = 1;
    } else if (dzv_min > kEps || dzh_min > kEps) {
        zsa = dzv_min < kEps ? std::nearbyint(zsa) : zsa;
        xsa = dzh_min < kEps ? std::nearbyint(xsa) : xsa;
        iflag = 2;
    } else {
        zsa = std::nearbyint(zsa);
        xsa = std::nearbyint(xsa);
        iflag = 3;
    }
    P->zsa = zsa;
    P->xsa = xsa;
    P->ysa = 0.0;
    P->zsi = zsi;
    P->xsi = xsi;
    P->ysi = 0;
    P->iflag = iflag;
    return true;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Dynamic LDS one workgroup may use on the current device (160 KiB on MI355X), the bound
// of the staged 2D sweep: models with lds2_bytes(nz) above it sweep in global memory.
long long lds2_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
        return 65536;
    return v;
}

// Grid spacings the 2D operator accepts: its 4-point quotient divides by the fixed
// dz^-2 + dx^-2 through its reciprocal (div_fixed), which needs both to be normal doubles,
// as they are with a wide margin for spacings in [1e-100, 1e100]; calls outside are refused.
bool spacing_ok2(double dz, double dx) {
    auto ok = [](double h) { return h >= 1e-100 && h <= 1e100; };
    return ok(dz) && ok(dx);
}

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" size_t fteik2d_workspace_size(int nsrc) { return align256(sizeof(SrcParam) * nsrc) + 256; }

extern "C" int fteik2d_solve(const double *slow, int nz, int nx, double dz, double dx,
                             const double *src, int nsrc, int nsweep, int max_sets, int want_grad,
                             double *tt, uint8_t *sgn, double *vzero, int *sets_done,
                             void *workspace, size_t workspace_bytes, void *stream_) {
    if (nz < 1 || nx < 1 || nsrc < 1 || !slow || !src || !tt || !vzero || !sets_done ||
        !workspace || !spacing_ok2(dz, dx))
        return FTEIK_ERR_ARG;
    if (want_grad && !sgn) return FTEIK_ERR_ARG;
    if (nsweep <= 0 && max_sets < 1) return FTEIK_ERR_ARG;
    if (workspace_bytes < fteik2d_workspace_size(nsrc)) return FTEIK_ERR_WORKSPACE;
    hipStream_t stream = (hipStream_t)stream_;
    std::vector<SrcParam> prm(nsrc);
    for (int s = 0; s < nsrc; ++s)
        if (!place_source2(src + 2 * s, nz, nx, dz, dx, &prm[s])) return FTEIK_ERR_SOURCE_OOB;
    SrcParam *dprm = (SrcParam *)workspace;
    FTEIK_CHECK(hipMemcpyAsync(dprm, prm.data(), sizeof(SrcParam) * nsrc, hipMemcpyHostToDevice,
                               stream));
    // the host vector goes out of scope on return: finish the copy first
    FTEIK_CHECK(hipStreamSynchronize(stream));
    const long long total = (long long)(nz + 1) * (nx + 1) * nsrc;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(fill2d_kernel, dim3(blocks), dim3(256), 0, stream, tt, total,
                       want_grad ? sgn : (uint8_t *)nullptr);
    FTEIK_CHECK(hipGetLastError());
    // one wave per source (lanes of one wave running different sources would execute the
    // union of the branches their nodes take); the staged rows in LDS, or for a model too
    // deep for it in stream-ordered scratch memory
    const long long lds = lds2_bytes(nz);
    double *scratch = nullptr;
    if (lds > lds2_limit())
        FTEIK_CHECK(hipMallocAsync((void **)&scratch, (size_t)lds * nsrc, stream));
    hipLaunchKernelGGL(fteik2d_kernel, dim3(nsrc), dim3(64), scratch ? 0u : (unsigned)lds,
                       stream, slow, nz, nx, dz, dx, dprm, nsrc, nsweep, max_sets, tt,
                       want_grad ? sgn : nullptr, vzero, sets_done, scratch);
    const int st = hip_status(hipGetLastError());
    if (scratch) {
        const int sf = hip_status(hipFreeAsync(scratch, stream));
        return st ? st : sf;
    }
    return st;
}

extern "C" int fteik2d_gradient(const double *tt, const uint8_t *sgn, int nz, int nx, double dz,
                                double dx, const double *src, double vzero, double *ttgrad,
                                void *stream) {
    if (!tt || !sgn || !ttgrad || !src || nz < 1 || nx < 1) return FTEIK_ERR_ARG;
    SrcParam P;
    if (!place_source2(src, nz, nx, dz, dx, &P)) return FTEIK_ERR_SOURCE_OOB;
    const long long nn = (long long)(nz + 1) * (nx + 1);
    hipLaunchKernelGGL(gradient2d_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, tt, sgn, nz + 1, nx + 1, dz, dx, P, vzero, ttgrad);
    return hip_status(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Batch of independent 2D models (fteik_solve_batch, batch.hip)
// ---------------------------------------------------------------------------------------

#include "batch_internal.hpp"

namespace fteik {

struct Batch2 {
    std::vector<Prob2> probs;
    std::vector<SrcParam> prm;   // all sources, model after model
    std::vector<int> cum;        // [n2 + 1]
    std::vector<const fteik_problem *> src;
    char *base = nullptr;
    size_t off_probs = 0, off_cum = 0, off_prm = 0;
    int nsweep = 0, max_sets = 1;
    bool grad = false;
};

namespace {
void batch2_layout(const fteik_problem *p, int n, size_t *probs, size_t *cum, size_t *prm,
                   size_t *total) {
    int n2 = 0;
    long long ns = 0;
    for (int i = 0; i < n; ++i)
        if (p[i].ndim == 2) {
            ++n2;
            ns += p[i].nsrc;
        }
    size_t off = 0;
    *probs = off;
    off += align256(sizeof(Prob2) * n2);
    *cum = off;
    off += align256(sizeof(int) * (n2 + 1));
    *prm = off;
    off += align256(sizeof(SrcParam) * ns);
    *total = n2 ? off : 0;
}
}  // namespace

size_t batch2_bytes(const fteik_problem *probs, int nprob) {
    size_t a, b, c, t;
    batch2_layout(probs, nprob, &a, &b, &c, &t);
    return t;
}

int batch2_prepare(const fteik_problem *P, int n, int nsweep, int max_sets, int want_grad,
                   char *ws, Batch2 **out, int *failed) {
    Batch2 *b = new Batch2;
    size_t total;
    batch2_layout(P, n, &b->off_probs, &b->off_cum, &b->off_prm, &total);
    b->base = ws;
    b->nsweep = nsweep;
    b->max_sets = max_sets;
    b->grad = want_grad != 0;
    b->cum.push_back(0);
    SrcParam *prm_dev = (SrcParam *)(ws + b->off_prm);
    for (int i = 0; i < n; ++i) {
        const fteik_problem &q = P[i];
        if (q.ndim != 2) continue;
        if (q.nz < 1 || q.nx < 1 || q.nsrc < 1 || !q.slow || !q.src || !q.tt || !q.vzero ||
            !q.sets_done || !spacing_ok2(q.dz, q.dx) || (want_grad && !q.sgn) ||
            (long long)b->cum.back() + q.nsrc >= (1LL << 31)) {
            delete b;
            return FTEIK_ERR_ARG;
        }
        const size_t first = b->prm.size();
        for (int s = 0; s < q.nsrc; ++s) {
            SrcParam sp;
            if (!place_source2(q.src + 2 * s, q.nz, q.nx, q.dz, q.dx, &sp)) {
                if (failed) *failed = i;
                delete b;
                return FTEIK_ERR_SOURCE_OOB;
            }
            b->prm.push_back(sp);
        }
        Prob2 p2;
        p2.slow = q.slow;
        p2.nz = q.nz;
        p2.nx = q.nx;
        p2.dz = q.dz;
        p2.dx = q.dx;
        p2.prm = prm_dev + first;
        p2.tt = q.tt;
        p2.sgn = want_grad ? q.sgn : nullptr;
        p2.vzero = q.vzero;
        p2.sets_done = q.sets_done;
        p2.scratch = nullptr;  // set at enqueue
        b->probs.push_back(p2);
        b->cum.push_back(b->cum.back() + q.nsrc);
        b->src.push_back(&q);
    }
    *out = b;
    return 0;
}

void batch2_free(Batch2 *b) { delete b; }

int batch2_enqueue(Batch2 *b, hipStream_t stream) {
    const int n2 = (int)b->probs.size();
    if (n2 == 0) return 0;
    // Staged rows: in the launch's dynamic LDS (sized for the largest model that fits the
    // device's per-workgroup LDS) or, for models too deep for it, in one stream-ordered
    // scratch allocation carved per model
    long long lds = 0, deep = 0;
    const long long cap = lds2_limit();
    for (int m = 0; m < n2; ++m) {
        const long long l = lds2_bytes(b->probs[m].nz);
        if (l <= cap) lds = std::max(lds, l);
        else deep += l * b->src[m]->nsrc;
    }
    char *scratch = nullptr;
    if (deep) FTEIK_CHECK(hipMallocAsync((void **)&scratch, (size_t)deep, stream));
    long long off = 0;
    for (int m = 0; m < n2; ++m) {
        const long long l = lds2_bytes(b->probs[m].nz);
        b->probs[m].scratch = l <= cap ? nullptr : (double *)(scratch + off);
        if (l > cap) off += l * b->src[m]->nsrc;
    }
    int st = hip_status(hipMemcpyAsync(b->base + b->off_probs, b->probs.data(),
                                       sizeof(Prob2) * n2, hipMemcpyHostToDevice, stream));
    if (!st)
        st = hip_status(hipMemcpyAsync(b->base + b->off_cum, b->cum.data(),
                                       sizeof(int) * (n2 + 1), hipMemcpyHostToDevice, stream));
    if (!st)
        st = hip_status(hipMemcpyAsync(b->base + b->off_prm, b->prm.data(),
                                       sizeof(SrcParam) * b->prm.size(), hipMemcpyHostToDevice,
                                       stream));
    // the host arrays are freed by the caller after this returns: finish the copies
    if (!st) st = hip_status(hipStreamSynchronize(stream));
    for (int m = 0; m < n2 && !st; ++m) {
        const fteik_problem &q = *b->src[m];
        const long long total = (long long)(q.nz + 1) * (q.nx + 1) * q.nsrc;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(fill2d_kernel, dim3(blocks), dim3(256), 0, stream, q.tt, total,
                           b->grad ? q.sgn : (uint8_t *)nullptr);
        st = hip_status(hipGetLastError());
    }
    // one wave per source: sources of different models (and different branches of the node
    // update) would otherwise serialise each other inside a wave
    if (!st) {
        hipLaunchKernelGGL(fteik2d_batch_kernel, dim3(b->cum.back()), dim3(64), (unsigned)lds,
                           stream, (const Prob2 *)(b->base + b->off_probs),
                           (const int *)(b->base + b->off_cum), n2, b->nsweep, b->max_sets);
        st = hip_status(hipGetLastError());
    }
    if (scratch) {
        const int sf = hip_status(hipFreeAsync(scratch, stream));
        if (!st) st = sf;
    }
    return st;
}

}  // namespace fteik
