// Adjoint-state gradient of receiver traveltimes with respect to slowness for the 3D solver
// (no counterpart in the reference: it differentiates the converged fixed point of
// fteik3d, fteikpy/_fteik/_fteik3d.py:38-200 and :394-462, and the receiver interpolation
// vinterp3d, fteikpy/_interp/_vinterp3d.py:7-264).
//
// Selection.  At a converged grid a node lowered by the sweeps equals the smallest
// candidate of the node operator over the 8 sweep directions, evaluated on the final
// neighbour times.  adj_select_kernel recomputes that candidate (node_update, the very
// operator of the sweeps) for every node and keeps the direction and candidate (first
// direction in pass order on ties), so nothing has to be stored by the forward solve.
// Two exceptions: a seeded source corner that no candidate reaches keeps vzero * distance,
// and a "stale" node holds a value below every current candidate (a 2D/3D candidate whose
// causality condition stopped holding after the node's last update): its best current
// candidate stands in for it when that candidate reads only earlier nodes, else its
// adjoint is dropped.  Stale nodes carrying adjoint are counted and reported.
//
// Linear solve.  Differentiating T_m = F_m(T_up(m), s) gives the adjoint system
//   lambda = g + J^T lambda,   dL/ds = B^T lambda (+ the vzero terms),
// g being dL/dT at the receivers' interpolation corners.  The stencil graph has no cycles
// (a cycle is reported as an error), so it is solved exactly in reverse dependency order:
// every node counts the nodes whose stencil reads it; a node whose count is zero has its
// final lambda and pushes lambda * dT_m/dT_u to its upwind nodes and lambda * dT_m/ds to
// its slowness cell.  Rounds of such frontiers run one launch each over all sources.
// Stencil weights are recomputed from the converged times when a node is pushed.  FP64
// atomics accumulate lambda and dL/ds, so the summation order is not fixed.
//
// Exact mode (fteik3d_adjoint_history).  Stale nodes make the converged-stencil gradient
// differ from the derivative of the sweeps themselves, which is what finite differences
// of the solver measure.  A logged solve (fteik3d_solve_logged) records every node change
// (node, time before it) pass by pass; the reverse walks the passes last to first.  Within
// a pass the changes are bucketed by oriented hyperplane a + b + c and the buckets run
// downwind first: a changed node's adjoint is pushed through the candidate that produced
// its new value, recomputed on the grid as it stood after the pass (its upwind neighbours
// were final for the pass when it was updated), then the node's adjoint is cleared and
// its old time restored.  Nodes of one hyperplane never read each other, so the restore is
// safe inside a bucket.  What remains on the seeded corners reaches vzero.  This is reverse
// mode through the whole sweep history: exact wherever the sweeps are differentiable.
#include "fteik_amd.h"
#include "fteik_device.hpp"
#include "interp_device.hpp"
#include "node_update.hpp"
#include "history.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fteik {
namespace {

// Per-source constants: placement as the solver computes it, absolute source position.
struct AdjSrc {
    SrcParam p;
    double zs, xs, ys;
};

// Node classification byte: bits 0-2 candidate axes (bit 0 z, 1 x, 2 y), bits 3-5 sweep
// direction (pass order), bits 6-7 kind.
enum : unsigned { kKindNone = 0, kKindStencil = 1, kKindSeed = 2, kKindStale = 3 };
__device__ __forceinline__ uint8_t pack(unsigned which, unsigned dir, unsigned kind) {
    return (uint8_t)(which | (dir << 3) | (kind << 6));
}
// sweep direction p: z descends for p >= 4, x for odd p, y for p = 2, 3, 6, 7 (the
// reference's pass order, _fteik3d.py:217-391)
__device__ __forceinline__ int dir_sz(int p) { return (p & 4) ? -1 : 1; }
__device__ __forceinline__ int dir_sx(int p) { return (p & 1) ? -1 : 1; }
__device__ __forceinline__ int dir_sy(int p) { return (p & 2) ? -1 : 1; }

struct AdjArgs : Stencil3 {
    const double *slow;
    const double *tt;  // [nsrc][NZ][NX][NY]
    const AdjSrc *src;
    uint8_t *code;     // [nsrc][nodes]
    double *lam;       // [nsrc][nodes]
    int *cnt;          // [nsrc][nodes] dependents not pushed yet
    double *dlds;      // [ngroup][nz][nx][ny] partial sums, source s adds into s % ngroup
    long long cells;   // nz * nx * ny
    int ngroup;
    int *stale;        // stale nodes carrying adjoint
    long long nodes;
    int nz, nx, ny, NZ, NX, NY, nsrc;
};

__device__ __forceinline__ long long node_at(const AdjArgs &A, int i, int j, int k) {
    return ((long long)i * A.NX + j) * A.NY + k;
}
__device__ __forceinline__ long long cell_at(const AdjArgs &A, int i, int j, int k) {
    return ((long long)i * A.nx + j) * A.ny + k;
}
// dL/ds partial sums: sources of one hyperplane bucket or frontier hit the same cells at
// the same time, so they accumulate into separate arrays (reduced at the end).
__device__ __forceinline__ void add_dlds(const AdjArgs &A, int s, long long cell, double v) {
    atomicAdd(A.dlds + (long long)(s % A.ngroup) * A.cells + cell, v);
}
__device__ __forceinline__ void node_ijk(const AdjArgs &A, long long n, int *i, int *j, int *k) {
    *k = (int)(n % A.NY);
    const long long r = n / A.NY;
    *j = (int)(r % A.NX);
    *i = (int)(r / A.NX);
}

// Upwind node of subset mask b (bit 0 z, 1 x, 2 y) for direction signs (sz, sx, sy).
__device__ __forceinline__ long long upwind(const AdjArgs &A, int i, int j, int k, int b, int sz,
                                            int sx, int sy) {
    return node_at(A, i - ((b & 1) ? sz : 0), j - ((b & 2) ? sx : 0), k - ((b & 4) ? sy : 0));
}

// Candidates of direction p at node (i, j, k) on the converged times: the node operator of
// the sweeps with t0 = kBig, fed exactly as the sweep kernel feeds it.
__device__ inline double candidate(const AdjArgs &A, const double *tt, int i, int j, int k, int p,
                                   int *which) {
    const int sz = dir_sz(p), sx = dir_sx(p), sy = dir_sy(p);
    // cells: upwind side i1 = i - (s > 0), other side clamped to the grid
    const int i1 = i - (sz > 0), j1 = j - (sx > 0), k1 = k - (sy > 0);
    const int i2 = sz > 0 ? min(i, A.NZ - 2) : max(i - 1, 0);
    const int j2 = sx > 0 ? min(j, A.NX - 2) : max(j - 1, 0);
    const int k2 = sy > 0 ? min(k, A.NY - 2) : max(k - 1, 0);
    const double *S = A.slow;
    return node_update(A, kBig, tt[node_at(A, i - sz, j, k)], tt[node_at(A, i, j - sx, k)],
                       tt[node_at(A, i, j, k - sy)], tt[node_at(A, i - sz, j - sx, k)],
                       tt[node_at(A, i, j - sx, k - sy)], tt[node_at(A, i - sz, j, k - sy)],
                       tt[node_at(A, i - sz, j - sx, k - sy)], S[cell_at(A, i1, j1, k1)],
                       S[cell_at(A, i1, j1, k2)], S[cell_at(A, i1, j2, k1)],
                       S[cell_at(A, i1, j2, k2)], S[cell_at(A, i2, j1, k1)],
                       S[cell_at(A, i2, j1, k2)], S[cell_at(A, i2, j2, k1)], which);
}

// Partial derivatives of the selected candidate (direction p, axes `which`): dT/dT_up[b]
// for subset masks b of `which`, the slowness cell the candidate used and dT/ds there.
struct Stencil {
    double w[8];
    long long cell;
    double wcell;
};

__device__ inline void stencil(const AdjArgs &A, const double *tt, int i, int j, int k, int p,
                               int which, Stencil *S) {
    const int sz = dir_sz(p), sx = dir_sx(p), sy = dir_sy(p);
    const int i1 = i - (sz > 0), j1 = j - (sx > 0), k1 = k - (sy > 0);
    const int im = max(i - 1, 0), ip = min(i, A.NZ - 2);
    const int jm = max(j - 1, 0), jp = min(j, A.NX - 2);
    const int km = max(k - 1, 0), kp = min(k, A.NY - 2);
#pragma unroll
    for (int b = 0; b < 8; ++b) S->w[b] = 0.0;
    // the cell achieving the candidate's min(): the first of equal values, in the order
    // the reference lists them (_fteik3d.py:78-141)
    long long c[4];
    int nc;
    double step = 0.0;
    if (which == 1) {
        c[0] = cell_at(A, i1, jm, km); c[1] = cell_at(A, i1, jm, kp);
        c[2] = cell_at(A, i1, jp, km); c[3] = cell_at(A, i1, jp, kp);
        nc = 4; step = A.dz;
    } else if (which == 2) {
        c[0] = cell_at(A, im, j1, km); c[1] = cell_at(A, ip, j1, km);
        c[2] = cell_at(A, im, j1, kp); c[3] = cell_at(A, ip, j1, kp);
        nc = 4; step = A.dx;
    } else if (which == 4) {
        c[0] = cell_at(A, im, jm, k1); c[1] = cell_at(A, im, jp, k1);
        c[2] = cell_at(A, ip, jm, k1); c[3] = cell_at(A, ip, jp, k1);
        nc = 4; step = A.dy;
    } else if (which == 3) {
        c[0] = cell_at(A, i1, j1, km); c[1] = cell_at(A, i1, j1, kp); nc = 2;
    } else if (which == 5) {
        c[0] = cell_at(A, i1, jm, k1); c[1] = cell_at(A, i1, jp, k1); nc = 2;
    } else if (which == 6) {
        c[0] = cell_at(A, im, j1, k1); c[1] = cell_at(A, ip, j1, k1); nc = 2;
    } else {
        c[0] = cell_at(A, i1, j1, k1); nc = 1;
    }
    long long best = c[0];
    double v = A.slow[best];
    for (int q = 1; q < nc; ++q) {
        const double x = A.slow[c[q]];
        if (x < v) {
            v = x;
            best = c[q];
        }
    }
    S->cell = best;
    if (nc == 4) {  // T = T_up + step * v
        S->w[which] = 1.0;
        S->wcell = step;
        return;
    }
    if (nc == 2) {
        // plane-wave operator over axes a < b (z < x < y) with upwind times ta, tb and the
        // diagonal tab: T = (a2i (tab + ta - tb) + b2i (tab + tb - ta) + sqrt(R)) / (a2i + b2i)
        // with R = 4 v^2 (a2i + b2i) - a2i b2i (2 (tb - ta))^2 in all three planes
        const int ba = (which & 1) ? 1 : 2, bb = (which & 4) ? 4 : 2;
        const double a2i = ba == 1 ? A.dz2i : A.dx2i, b2i = bb == 4 ? A.dy2i : A.dx2i;
        const double ta = tt[upwind(A, i, j, k, ba, sz, sx, sy)];
        const double tb = tt[upwind(A, i, j, k, bb, sz, sx, sy)];
        const double q = 2.0 * (tb - ta);
        const double sR = sqrt(4.0 * v * v * (a2i + b2i) - a2i * b2i * q * q);
        const double sum = a2i + b2i, g = 2.0 * a2i * b2i * q / sR;
        S->w[ba] = (a2i - b2i + g) / sum;
        S->w[bb] = (b2i - a2i - g) / sum;
        S->w[ba | bb] = 1.0;
        S->wcell = 4.0 * v / sR;
        return;
    }
    // 3D operator, _fteik3d.py:146-159: T = (tb dz2i + ta dx2i + tc dy2i + sqrt(9 v^2 dsum - Q))
    // / dsum, (ta, tb, tc) linear in the 7 upwind times
    const double tv = tt[upwind(A, i, j, k, 1, sz, sx, sy)], te = tt[upwind(A, i, j, k, 2, sz, sx, sy)];
    const double tn = tt[upwind(A, i, j, k, 4, sz, sx, sy)], tev = tt[upwind(A, i, j, k, 3, sz, sx, sy)];
    const double tnv = tt[upwind(A, i, j, k, 5, sz, sx, sy)], ten = tt[upwind(A, i, j, k, 6, sz, sx, sy)];
    const double tnve = tt[upwind(A, i, j, k, 7, sz, sx, sy)];
    const double ta = te - 0.5 * tn + 0.5 * ten - 0.5 * tv + 0.5 * tev - tnv + tnve;
    const double tb = tv - 0.5 * tn + 0.5 * tnv - 0.5 * te + 0.5 * tev - ten + tnve;
    const double tc = tn - 0.5 * te + 0.5 * ten - 0.5 * tv + 0.5 * tnv - tev + tnve;
    const double Q = A.dz2dx2 * (ta - tb) * (ta - tb) + A.dz2dy2 * (tb - tc) * (tb - tc) +
                     A.dx2dy2 * (ta - tc) * (ta - tc);
    const double sR = sqrt(9.0 * v * v * A.dsum - Q);
    const double ga = (A.dx2i - (A.dz2dx2 * (ta - tb) + A.dx2dy2 * (ta - tc)) / sR) / A.dsum;
    const double gb = (A.dz2i - (A.dz2dy2 * (tb - tc) - A.dz2dx2 * (ta - tb)) / sR) / A.dsum;
    const double gc = (A.dy2i + (A.dz2dy2 * (tb - tc) + A.dx2dy2 * (ta - tc)) / sR) / A.dsum;
    // coefficients of tv, te, tn, tev, tnv, ten, tnve in (ta, tb, tc)
    S->w[1] = -0.5 * ga + gb - 0.5 * gc;
    S->w[2] = ga - 0.5 * gb - 0.5 * gc;
    S->w[4] = -0.5 * ga - 0.5 * gb + gc;
    S->w[3] = 0.5 * ga + 0.5 * gb - gc;
    S->w[5] = -ga + 0.5 * gb + 0.5 * gc;
    S->w[6] = 0.5 * ga - gb + 0.5 * gc;
    S->w[7] = ga + gb + gc;
    S->wcell = 9.0 * v / sR;
}

// Selection of every (source, node) and count of the nodes reading each node.
__global__ void adj_select_kernel(AdjArgs A) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nodes * A.nsrc) return;
    const int s = (int)(g / A.nodes);
    const long long n = g - (long long)s * A.nodes;
    const double *tt = A.tt + (long long)s * A.nodes;
    int i, j, k;
    node_ijk(A, n, &i, &j, &k);
    const double t = tt[n];
    double best = kBig;
    int bw = 0, bp = 0;
    for (int p = 0; p < 8; ++p) {
        const int sz = dir_sz(p), sx = dir_sx(p), sy = dir_sy(p);
        if (i - sz < 0 || i - sz >= A.NZ || j - sx < 0 || j - sx >= A.NX || k - sy < 0 ||
            k - sy >= A.NY)
            continue;
        int w;
        const double m = candidate(A, tt, i, j, k, p, &w);
        if (m < best) {
            best = m;
            bw = w;
            bp = p;
        }
    }
    const SrcParam &P = A.src[s].p;
    const int di = i - P.zsi, dj = j - P.xsi, dk = k - P.ysi;
    const bool corner = di >= 0 && di <= 1 && dj >= 0 && dj <= 1 && dk >= 0 && dk <= 1;
    unsigned kind = kKindNone;
    if (bw != 0 && best <= t) {
        kind = kKindStencil;
    } else if (corner) {
        kind = kKindSeed;
    } else if (bw != 0) {
        // stale: the stand-in must read only earlier nodes
        kind = kKindStale;
        const int sz = dir_sz(bp), sx = dir_sx(bp), sy = dir_sy(bp);
        for (int b = 1; b < 8; ++b)
            if ((b & bw) == b && !(tt[upwind(A, i, j, k, b, sz, sx, sy)] < t)) kind = kKindNone;
    }
    A.code[g] = pack(bw, bp, kind);
    if (kind == kKindStencil || kind == kKindStale) {
        const int sz = dir_sz(bp), sx = dir_sx(bp), sy = dir_sy(bp);
        for (int b = 1; b < 8; ++b)
            if ((b & bw) == b)
                atomicAdd(A.cnt + (long long)s * A.nodes + upwind(A, i, j, k, b, sz, sx, sy), 1);
    }
}

// Adjoint source of every receiver: dT_rec/dT at the corners of its cell (apparent
// velocity interpolation, _vinterp3d.py:26-262), or dT_rec/dvzero = distance when the
// receiver lies in the source cell or next to a zero-time corner.  Receivers outside the
// grid (fill value) contribute nothing.
__global__ void adj_receiver_kernel(AdjArgs A, Axis Z, Axis X, Axis Y, const double *rec,
                                    int nrec, const double *dldt) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)nrec * A.nsrc) return;
    const int s = (int)(g / nrec), r = (int)(g - (long long)s * nrec);
    const double wr = dldt[g];
    if (wr == 0.0) return;
    const double zq = rec[3 * r], xq = rec[3 * r + 1], yq = rec[3 * r + 2];
    if (!(Z.contains(zq) && X.contains(xq) && Y.contains(yq))) return;
    const AdjSrc S = A.src[s];
    const double dist = dist3(S.zs, S.xs, S.ys, zq, xq, yq);
    const long long scell = cell_at(A, S.p.zsi, S.p.xsi, S.p.ysi);
    const Corner cz = corner(Z, Z.cell(zq)), cx = corner(X, X.cell(xq)), cy = corner(Y, Y.cell(yq));
    if (cz.i1 == Z.cell(S.zs) && cx.i1 == X.cell(S.xs) && cy.i1 == Y.cell(S.ys)) {
        add_dlds(A, s, scell, wr * dist);
        return;
    }
    const double *tt = A.tt + (long long)s * A.nodes;
    const double vol = fabs((cz.x2 - cz.x1) * (cx.x2 - cx.x1) * (cy.x2 - cy.x1));
    double a[8], t[8];
    long long id[8];
    double V = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int di = c & 1, dj = (c >> 1) & 1, dk = (c >> 2) & 1;
        id[c] = -1;
        a[c] = 0.0;
        t[c] = 1.0;
        if ((di && !cz.has2) || (dj && !cx.has2) || (dk && !cy.has2)) continue;  // ghost corner
        id[c] = node_at(A, cz.i1 + di, cx.i1 + dj, cy.i1 + dk);
        t[c] = tt[id[c]];
        if (t[c] == 0.0) {  // the source sits on this node (_vinterp3d.py:246-250)
            add_dlds(A, s, scell, wr * dist);
            return;
        }
        const double w = fabs(((di ? cz.x1 : cz.x2) - zq) * ((dj ? cx.x1 : cx.x2) - xq) *
                              ((dk ? cy.x1 : cy.x2) - yq));
        a[c] = dist3(S.zs, S.xs, S.ys, di ? cz.x2 : cz.x1, dj ? cx.x2 : cx.x1,
                     dk ? cy.x2 : cy.x1) * w / vol;
        V += a[c] / t[c];
    }
    // T = dist / V with V = sum a_c / t_c: dT/dt_c = dist a_c / (V t_c)^2
    for (int c = 0; c < 8; ++c)
        if (id[c] >= 0)
            atomicAdd(A.lam + (long long)s * A.nodes + id[c],
                      wr * dist * a[c] / ((V * t[c]) * (V * t[c])));
}

// First frontier: the nodes no stencil reads.
__global__ void adj_frontier_init_kernel(AdjArgs A, unsigned *front, int *nfront) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nodes * A.nsrc) return;
    if (A.cnt[g] == 0) front[atomicAdd(nfront, 1)] = (unsigned)g;
}

// One round: every frontier node holds its final lambda; push it, release upwind nodes
// whose last reader this was.
__global__ void adj_round_kernel(AdjArgs A, const unsigned *front, const int *nfront_p,
                                 unsigned *next, int *nnext) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= *nfront_p) return;
    const long long g = front[f];
    const int s = (int)(g / A.nodes);
    const long long n = g - (long long)s * A.nodes;
    const uint8_t code = A.code[g];
    const unsigned kind = code >> 6, which = code & 7u;
    const double lam = A.lam[g];
    if (kind == kKindNone) {
        // a reached node that carries adjoint but has no usable stencil: a dropped stale node
        if (lam != 0.0 && A.tt[g] < kBig) atomicAdd(A.stale, 1);
        return;
    }
    int i, j, k;
    node_ijk(A, n, &i, &j, &k);
    if (kind == kKindSeed) {
        // seeded corner: T = vzero * distance (_fteik3d.py:10-17, :439-451)
        if (lam != 0.0) {
            const SrcParam &P = A.src[s].p;
            const double a = A.dz * ((double)i - P.zsa), b = A.dx * ((double)j - P.xsa),
                         c = A.dy * ((double)k - P.ysa);
            add_dlds(A, s, cell_at(A, P.zsi, P.xsi, P.ysi), lam * sqrt(a * a + b * b + c * c));
        }
        return;
    }
    if (kind == kKindStale && lam != 0.0) atomicAdd(A.stale, 1);
    const int p = (code >> 3) & 7;
    const double *tt = A.tt + (long long)s * A.nodes;
    Stencil S;
    stencil(A, tt, i, j, k, p, (int)which, &S);
    if (lam != 0.0) add_dlds(A, s, S.cell, lam * S.wcell);
    const int sz = dir_sz(p), sx = dir_sx(p), sy = dir_sy(p);
    for (int b = 1; b < 8; ++b) {
        if ((b & (int)which) != b) continue;
        const long long u = (long long)s * A.nodes + upwind(A, i, j, k, b, sz, sx, sy);
        if (lam != 0.0 && S.w[b] != 0.0) atomicAdd(A.lam + u, lam * S.w[b]);
        if (atomicSub(A.cnt + u, 1) == 1) next[atomicAdd(nnext, 1)] = (unsigned)u;
    }
}

// Exact mode: oriented hyperplane of every change of one pass, counted per plane.
__device__ __forceinline__ int plane_of(const AdjArgs &A, long long n, int p) {
    int i, j, k;
    node_ijk(A, n, &i, &j, &k);
    const int a = dir_sz(p) > 0 ? i : A.NZ - 1 - i;
    const int b = dir_sx(p) > 0 ? j : A.NX - 1 - j;
    const int c = dir_sy(p) > 0 ? k : A.NY - 1 - k;
    return a + b + c;
}

// Bucketing of one pass's changes by hyperplane: few bins (nz + nx + ny + 1), many
// entries, so each workgroup counts its kHistItems * 256 entries in LDS and touches every
// global bin once.
constexpr unsigned kHistUnused = 0xffffffffu;  // padding entry of a reserved chunk
constexpr int kHistThreads = 256;
constexpr int kHistItems = 16;
constexpr int kMaxBins = 8192;

__global__ __launch_bounds__(kHistThreads) void hist_count_kernel(
    AdjArgs A, const unsigned *node, unsigned long long first, unsigned n, int p, int nbins,
    int *bins) {
    __shared__ int local[kMaxBins];
    for (int b = threadIdx.x; b < nbins; b += kHistThreads) local[b] = 0;
    __syncthreads();
    const unsigned q0 = blockIdx.x * (kHistThreads * kHistItems) + threadIdx.x;
#pragma unroll
    for (int it = 0; it < kHistItems; ++it) {
        const unsigned q = q0 + it * kHistThreads;
        if (q < n) {
            const unsigned g = node[first + q];
            if (g != kHistUnused)
                atomicAdd(local + plane_of(A, (long long)(g % (unsigned long long)A.nodes), p), 1);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += kHistThreads)
        if (local[b]) atomicAdd(bins + b, local[b]);
}

// Same blocks as hist_count_kernel: ranks inside the block from LDS, one global range
// reservation per (block, bin).
__global__ __launch_bounds__(kHistThreads) void hist_scatter_kernel(
    AdjArgs A, const unsigned *node, unsigned long long first, unsigned n, int p, int nbins,
    int *cursor, unsigned *perm) {
    __shared__ int local[kMaxBins];
    for (int b = threadIdx.x; b < nbins; b += kHistThreads) local[b] = 0;
    __syncthreads();
    const unsigned q0 = blockIdx.x * (kHistThreads * kHistItems) + threadIdx.x;
    int h[kHistItems], r[kHistItems];
#pragma unroll
    for (int it = 0; it < kHistItems; ++it) {
        const unsigned q = q0 + it * kHistThreads;
        h[it] = -1;
        if (q < n) {
            const unsigned g = node[first + q];
            if (g != kHistUnused) {
                h[it] = plane_of(A, (long long)(g % (unsigned long long)A.nodes), p);
                r[it] = atomicAdd(local + h[it], 1);
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += kHistThreads)
        if (local[b]) local[b] = atomicAdd(cursor + b, local[b]);  // block's base in bin b
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kHistItems; ++it)
        if (h[it] >= 0) perm[local[h[it]] + r[it]] = q0 + it * kHistThreads;
}

// One hyperplane of pass p, reversed: push the adjoint of every node it changed, restore
// the node's old time.
__global__ void hist_reverse_kernel(AdjArgs A, const unsigned *node, const double *old,
                                    unsigned long long first, const unsigned *perm, int count,
                                    int p, double *tt_all, int *mismatch) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const unsigned long long e = first + perm[q];
    const long long g = node[e];
    const int s = (int)(g / A.nodes);
    const long long n = g - (long long)s * A.nodes;
    double *tt = tt_all + (long long)s * A.nodes;
    const double lam = A.lam[g];
    if (lam != 0.0) {
        int i, j, k;
        node_ijk(A, n, &i, &j, &k);
        int which;
        const double m = candidate(A, tt, i, j, k, p, &which);
        if (m != tt[n]) atomicAdd(mismatch, 1);
        Stencil S;
        stencil(A, tt, i, j, k, p, which, &S);
        add_dlds(A, s, S.cell, lam * S.wcell);
        const int sz = dir_sz(p), sx = dir_sx(p), sy = dir_sy(p);
        for (int b = 1; b < 8; ++b)
            if ((b & which) == b && S.w[b] != 0.0)
                atomicAdd(A.lam + (long long)s * A.nodes + upwind(A, i, j, k, b, sz, sx, sy),
                          lam * S.w[b]);
        A.lam[g] = 0.0;
    }
    tt[n] = old[e];
}

// Exact mode: the adjoint left on the seeded corners reaches vzero (T = vzero * distance).
__global__ void hist_seed_kernel(AdjArgs A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nsrc * 8) return;
    const int s = t >> 3, c = t & 7;
    const SrcParam &P = A.src[s].p;
    const int i = P.zsi + (c & 1), j = P.xsi + ((c >> 1) & 1), k = P.ysi + ((c >> 2) & 1);
    const double lam = A.lam[(long long)s * A.nodes + node_at(A, i, j, k)];
    if (lam == 0.0) return;
    const double a = A.dz * ((double)i - P.zsa), b = A.dx * ((double)j - P.xsa),
                 cc = A.dy * ((double)k - P.ysa);
    add_dlds(A, s, cell_at(A, P.zsi, P.xsi, P.ysi), lam * sqrt(a * a + b * b + cc * cc));
}

// dlds = sum of the partial arrays, in group order.
__global__ void dlds_reduce_kernel(const double *part, int ngroup, long long cells, double *dlds) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    double v = 0.0;
    for (int g = 0; g < ngroup; ++g) v += part[(long long)g * cells + c];
    dlds[c] = v;
}

// Clears a frontier counter on the device.
__global__ void adj_clear_kernel(int *p) { *p = 0; }

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

bool place_source3(double z, double x, double y, int nz, int nx, int ny, double dz, double dx,
                   double dy, SrcParam *P) {
    if (!((0.0 <= z && z <= dz * nz) && (0.0 <= x && x <= dx * nx) && (0.0 <= y && y <= dy * ny)))
        return false;
    double zsa = z / dz, xsa = x / dx, ysa = y / dy;
    if (zsa >= nz) zsa = zsa - kEps;
    if (xsa >= nx) xsa = xsa - kEps;
    if (ysa >= ny) ysa = ysa - kEps;
    P->zsa = zsa;
    P->xsa = xsa;
    P->ysa = ysa;
    P->zsi = std::min((int)zsa, nz - 1);
    P->xsi = std::min((int)xsa, nx - 1);
    P->ysi = std::min((int)ysa, ny - 1);
    P->iflag = 0;
    return true;
}

}  // namespace
}  // namespace fteik

using namespace fteik;

namespace {

int dlds_groups(int nsrc) { return std::min(nsrc, 32); }

void fill_args(AdjArgs &A, const double *slow, int nz, int nx, int ny, double dz, double dx,
               double dy, const double *tt, int nsrc) {
    A.slow = slow; A.tt = tt; A.nodes = (long long)(nz + 1) * (nx + 1) * (ny + 1);
    A.cells = (long long)nz * nx * ny;
    A.ngroup = dlds_groups(nsrc);
    A.nz = nz; A.nx = nx; A.ny = ny; A.NZ = nz + 1; A.NX = nx + 1; A.NY = ny + 1; A.nsrc = nsrc;
    static_cast<Stencil3 &>(A) = make_stencil3(dz, dx, dy);
}

// Placement of every source from its absolute position (relative = absolute - origin,
// exactly as the solve receives it).
int place_sources(const double *origin, const double *src, int nsrc, int nz, int nx, int ny,
                  double dz, double dx, double dy, std::vector<AdjSrc> &hs) {
    hs.resize(nsrc);
    for (int s = 0; s < nsrc; ++s) {
        const double *q = src + 3 * s;
        if (!place_source3(q[0] - origin[0], q[1] - origin[1], q[2] - origin[2], nz, nx, ny, dz,
                           dx, dy, &hs[s].p))
            return FTEIK_ERR_SOURCE_OOB;
        hs[s].zs = q[0];
        hs[s].xs = q[1];
        hs[s].ys = q[2];
    }
    return 0;
}

}  // namespace

extern "C" size_t fteik3d_adjoint_history_workspace_size(int nz, int nx, int ny, int nsrc,
                                                         long long events) {
    if (nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || events < 0) return 0;
    const size_t total = (size_t)(nz + 1) * (nx + 1) * (ny + 1) * nsrc;
    const size_t nbins = (size_t)(nz + nx + ny + 1);
    const size_t part = (size_t)dlds_groups(nsrc) * nz * nx * ny * sizeof(double);
    return align256(sizeof(AdjSrc) * nsrc) + align256(total * sizeof(double)) +
           2 * align256(nbins * sizeof(int)) + align256((size_t)events * sizeof(unsigned)) +
           align256(part) + 256;
}

extern "C" int fteik3d_adjoint_history(const double *slow, int nz, int nx, int ny, double dz,
                                       double dx, double dy, const double *origin,
                                       const double *src, int nsrc, double *tt,
                                       const void *history, const double *rec, int nrec,
                                       const double *dldt, double *dlds, void *workspace,
                                       size_t workspace_bytes, void *stream_) {
    if (!slow || !origin || !src || !tt || !history || !dlds || !workspace ||
        (nrec > 0 && (!rec || !dldt)) || nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || nrec < 0 ||
        !spacing_ok3(dz, dx, dy))
        return FTEIK_ERR_ARG;
    hipStream_t stream = (hipStream_t)stream_;
    HistHeader hdr;
    FTEIK_CHECK(hipMemcpyAsync(&hdr, history, sizeof(hdr), hipMemcpyDeviceToHost, stream));
    FTEIK_CHECK(hipStreamSynchronize(stream));
    const long long nodes = (long long)(nz + 1) * (nx + 1) * (ny + 1);
    if (hdr.nodes != nodes || hdr.nsrc != nsrc || hdr.npass < 0 || hdr.npass > hdr.max_pass)
        return FTEIK_ERR_ARG;  // not the history of this grid and source set
    if (hdr.count > (unsigned long long)hdr.cap) return FTEIK_ERR_WORKSPACE;  // incomplete log
    const long long events = (long long)hdr.count;
    if (workspace_bytes < fteik3d_adjoint_history_workspace_size(nz, nx, ny, nsrc, events))
        return FTEIK_ERR_WORKSPACE;
    std::vector<AdjSrc> hs;
    if (int st = place_sources(origin, src, nsrc, nz, nx, ny, dz, dx, dy, hs)) return st;
    const HistView hv = hist_view(const_cast<void *>(history), hdr.cap, hdr.max_pass);
    std::vector<unsigned long long> seg(hdr.npass + 1);
    FTEIK_CHECK(hipMemcpyAsync(seg.data(), hv.seg, sizeof(unsigned long long) * (hdr.npass + 1),
                               hipMemcpyDeviceToHost, stream));

    const long long total = nodes * nsrc;
    const int nbins = nz + nx + ny + 1;
    if (nbins > kMaxBins) return FTEIK_ERR_ARG;  // hyperplanes per pass beyond the LDS bins
    char *p = (char *)workspace;
    AdjSrc *dsrc = (AdjSrc *)p;
    p += align256(sizeof(AdjSrc) * nsrc);
    double *lam = (double *)p;
    p += align256(total * sizeof(double));
    int *bins = (int *)p;
    p += align256(nbins * sizeof(int));
    int *cursor = (int *)p;
    p += align256(nbins * sizeof(int));
    unsigned *perm = (unsigned *)p;
    p += align256((size_t)events * sizeof(unsigned));
    double *part = (double *)p;
    p += align256((size_t)dlds_groups(nsrc) * nz * nx * ny * sizeof(double));
    int *mismatch = (int *)p;

    AdjArgs A{};
    fill_args(A, slow, nz, nx, ny, dz, dx, dy, tt, nsrc);
    A.src = dsrc;
    A.lam = lam;
    A.dlds = part;
    const Axis Z{origin[0], dz, nz + 1}, X{origin[1], dx, nx + 1}, Y{origin[2], dy, ny + 1};

    FTEIK_CHECK(hipMemcpyAsync(dsrc, hs.data(), sizeof(AdjSrc) * nsrc, hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemsetAsync(lam, 0, total * sizeof(double), stream));
    FTEIK_CHECK(hipMemsetAsync(mismatch, 0, sizeof(int), stream));
    FTEIK_CHECK(hipMemsetAsync(part, 0, (size_t)A.ngroup * A.cells * sizeof(double), stream));
    if (nrec > 0) {
        const long long nr = (long long)nrec * nsrc;
        hipLaunchKernelGGL(adj_receiver_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0,
                           stream, A, Z, X, Y, rec, nrec, dldt);
        FTEIK_CHECK(hipGetLastError());
    }
    FTEIK_CHECK(hipStreamSynchronize(stream));  // seg, hs

    std::vector<int> counts(nbins), start(nbins);
    for (int P = hdr.npass - 1; P >= 0; --P) {
        const unsigned long long first = seg[P], last = seg[P + 1];
        if (last <= first) continue;
        const unsigned n = (unsigned)(last - first);
        const int pdir = P % 8;
        const unsigned nb = (n + kHistThreads * kHistItems - 1) / (kHistThreads * kHistItems);
        FTEIK_CHECK(hipMemsetAsync(bins, 0, nbins * sizeof(int), stream));
        hipLaunchKernelGGL(hist_count_kernel, dim3(nb), dim3(kHistThreads), 0, stream, A, hv.node,
                           first, n, pdir, nbins, bins);
        FTEIK_CHECK(hipGetLastError());
        FTEIK_CHECK(hipMemcpyAsync(counts.data(), bins, nbins * sizeof(int), hipMemcpyDeviceToHost,
                                   stream));
        FTEIK_CHECK(hipStreamSynchronize(stream));
        // buckets laid out downwind first
        int off = 0;
        for (int h = nbins - 1; h >= 0; --h) {
            start[h] = off;
            off += counts[h];
        }
        FTEIK_CHECK(hipMemcpyAsync(cursor, start.data(), nbins * sizeof(int), hipMemcpyHostToDevice,
                                   stream));
        hipLaunchKernelGGL(hist_scatter_kernel, dim3(nb), dim3(kHistThreads), 0, stream, A,
                           hv.node, first, n, pdir, nbins, cursor, perm);
        FTEIK_CHECK(hipGetLastError());
        for (int h = nbins - 1; h >= 0; --h) {
            const int c = counts[h];
            if (!c) continue;
            hipLaunchKernelGGL(hist_reverse_kernel, dim3((c + 255) / 256), dim3(256), 0, stream, A,
                               hv.node, hv.old, first, perm + start[h], c, pdir, tt, mismatch);
            FTEIK_CHECK(hipGetLastError());
        }
        // start[] is reused for the next pass only after the sync that follows its count
    }
    hipLaunchKernelGGL(hist_seed_kernel, dim3((nsrc * 8 + 255) / 256), dim3(256), 0, stream, A);
    FTEIK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(dlds_reduce_kernel, dim3((unsigned)((A.cells + 255) / 256)), dim3(256), 0,
                       stream, part, A.ngroup, A.cells, dlds);
    FTEIK_CHECK(hipGetLastError());
    int bad = 0;
    FTEIK_CHECK(hipMemcpyAsync(&bad, mismatch, sizeof(int), hipMemcpyDeviceToHost, stream));
    FTEIK_CHECK(hipStreamSynchronize(stream));
    // every logged change must be reproduced by its recomputed candidate
    if (bad) return FTEIK_ERR_UNSETTLED;
    return 0;
}

extern "C" size_t fteik3d_adjoint_workspace_size(int nz, int nx, int ny, int nsrc) {
    if (nz < 1 || nx < 1 || ny < 1 || nsrc < 1) return 0;
    const size_t total = (size_t)(nz + 1) * (nx + 1) * (ny + 1) * nsrc;
    const size_t part = (size_t)dlds_groups(nsrc) * nz * nx * ny * sizeof(double);
    return align256(sizeof(AdjSrc) * nsrc) + align256(total) + align256(total * sizeof(double)) +
           align256(total * sizeof(int)) + 2 * align256(total * sizeof(unsigned)) +
           align256(part) + 256;
}

extern "C" int fteik3d_adjoint(const double *slow, int nz, int nx, int ny, double dz, double dx,
                               double dy, const double *origin, const double *src, int nsrc,
                               const double *tt, const double *rec, int nrec, const double *dldt,
                               double *dlds, int *stale_out, int *rounds_out, void *workspace,
                               size_t workspace_bytes, void *stream_) {
    if (!slow || !origin || !src || !tt || !dlds || !workspace || (nrec > 0 && (!rec || !dldt)) ||
        nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || nrec < 0 || !spacing_ok3(dz, dx, dy))
        return FTEIK_ERR_ARG;
    const long long nodes = (long long)(nz + 1) * (nx + 1) * (ny + 1);
    const long long total = nodes * nsrc;
    if (total > 0xffffffffLL) return FTEIK_ERR_ARG;  // frontier entries are 32-bit
    if (workspace_bytes < fteik3d_adjoint_workspace_size(nz, nx, ny, nsrc)) return FTEIK_ERR_WORKSPACE;
    std::vector<AdjSrc> hs;
    if (int st = place_sources(origin, src, nsrc, nz, nx, ny, dz, dx, dy, hs)) return st;
    hipStream_t stream = (hipStream_t)stream_;
    char *p = (char *)workspace;
    AdjSrc *dsrc = (AdjSrc *)p;
    p += align256(sizeof(AdjSrc) * nsrc);
    uint8_t *code = (uint8_t *)p;
    p += align256(total);
    double *lam = (double *)p;
    p += align256(total * sizeof(double));
    int *cnt = (int *)p;
    p += align256(total * sizeof(int));
    unsigned *fa = (unsigned *)p;
    p += align256(total * sizeof(unsigned));
    unsigned *fb = (unsigned *)p;
    p += align256(total * sizeof(unsigned));
    double *part = (double *)p;
    p += align256((size_t)dlds_groups(nsrc) * nz * nx * ny * sizeof(double));
    int *ctr = (int *)p;  // [0], [1]: frontier sizes; [2]: stale count

    AdjArgs A{};
    fill_args(A, slow, nz, nx, ny, dz, dx, dy, tt, nsrc);
    A.src = dsrc; A.code = code; A.lam = lam; A.cnt = cnt; A.dlds = part; A.stale = ctr + 2;
    const Axis Z{origin[0], dz, nz + 1}, X{origin[1], dx, nx + 1}, Y{origin[2], dy, ny + 1};

    FTEIK_CHECK(hipMemcpyAsync(dsrc, hs.data(), sizeof(AdjSrc) * nsrc, hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemsetAsync(lam, 0, total * sizeof(double), stream));
    FTEIK_CHECK(hipMemsetAsync(cnt, 0, total * sizeof(int), stream));
    FTEIK_CHECK(hipMemsetAsync(ctr, 0, 4 * sizeof(int), stream));
    FTEIK_CHECK(hipMemsetAsync(part, 0, (size_t)A.ngroup * A.cells * sizeof(double), stream));
    const unsigned nb = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(adj_select_kernel, dim3(nb), dim3(256), 0, stream, A);
    FTEIK_CHECK(hipGetLastError());
    if (nrec > 0) {
        const long long nr = (long long)nrec * nsrc;
        hipLaunchKernelGGL(adj_receiver_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0,
                           stream, A, Z, X, Y, rec, nrec, dldt);
        FTEIK_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(adj_frontier_init_kernel, dim3(nb), dim3(256), 0, stream, A, fa, ctr);
    FTEIK_CHECK(hipGetLastError());
    int nfront = 0;
    FTEIK_CHECK(hipMemcpyAsync(&nfront, ctr, sizeof(int), hipMemcpyDeviceToHost, stream));
    FTEIK_CHECK(hipStreamSynchronize(stream));  // also keeps hs alive for the copy above
    long long processed = 0;
    int rounds = 0, cur = 0;
    while (nfront > 0) {
        processed += nfront;
        hipLaunchKernelGGL(adj_clear_kernel, dim3(1), dim3(1), 0, stream, ctr + (1 - cur));
        hipLaunchKernelGGL(adj_round_kernel, dim3((unsigned)((nfront + 255) / 256)), dim3(256), 0,
                           stream, A, fa, ctr + cur, fb, ctr + (1 - cur));
        FTEIK_CHECK(hipGetLastError());
        FTEIK_CHECK(hipMemcpyAsync(&nfront, ctr + (1 - cur), sizeof(int), hipMemcpyDeviceToHost, stream));
        FTEIK_CHECK(hipStreamSynchronize(stream));
        std::swap(fa, fb);
        cur = 1 - cur;
        ++rounds;
    }
    hipLaunchKernelGGL(dlds_reduce_kernel, dim3((unsigned)((A.cells + 255) / 256)), dim3(256), 0,
                       stream, part, A.ngroup, A.cells, dlds);
    FTEIK_CHECK(hipGetLastError());
    int stale = 0;
    FTEIK_CHECK(hipMemcpyAsync(&stale, ctr + 2, sizeof(int), hipMemcpyDeviceToHost, stream));
    FTEIK_CHECK(hipStreamSynchronize(stream));
    if (rounds_out) *rounds_out = rounds;
    if (stale_out) *stale_out = stale;
    // every node is released exactly once unless the stencil graph has a cycle
    if (processed != total) return FTEIK_ERR_UNSETTLED;
    return 0;
}
