// 3D fast-sweeping eikonal solver for MI355X (gfx950): FTeik first-order operators
// (fteikpy/_fteik/_fteik3d.py:38-200) swept over 8 directions per set
// (_fteik3d.py:203-391), with source seeding (_fteik3d.py:394-456) and the gradient
// post-pass (_fteik3d.py:464-488).
//
// Schedule.  Within one directional pass a node reads only itself and 7 upwind neighbours,
// all on strictly lower anti-diagonal levels of that direction.  Every topological order
// of a pass therefore yields the reference's state bit for bit (FP contraction is disabled
// at build time).  The grid is cut into bricks of tile_a (z) x 8 (x) x 8 (y) nodes in the
// pass's oriented coordinates; bricks on one brick anti-diagonal are independent and are
// processed by one launch (one 64-lane wave per brick and source).  Inside a brick lane
// (bb, cc) handles column (x, y) = (b0+bb, c0+cc) and, at step s, plane p = s - bb - cc:
// all lanes of a step lie on one node hyperplane, so every upwind value was produced at an
// earlier step.  z planes are staged through an LDS ring with coalesced global loads and
// written back coalesced once all lanes are past them.
#include "fteik_amd.h"
#include "fteik_device.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fteik {
namespace {

constexpr int kLanesB = 8;      // brick width along x (oriented b)
constexpr int kLanesC = 8;      // brick width along y (oriented c)
constexpr int kSkew = kLanesB + kLanesC - 2;  // steps between first and last lane = 14
constexpr int kPlane = (kLanesB + 1) * (kLanesC + 1);  // 81 entries: node or cell block
constexpr int kPrefetch = 1;    // planes resident in LDS ahead of lane (0,0)
constexpr int kDepth = 4;       // further planes in flight in registers (loop unrolled by it)
// a slot is rewritten at the end of step s for plane s+kPrefetch; its previous plane was
// last read (as the upwind plane of lane (7,7)) at step s-1, before that step's barrier
constexpr int kRing = kPrefetch + kSkew + 2;

struct SweepArgs {
    const double *slow;
    double *tt;
    uint8_t *sgn;
    uint8_t *chg;           // per node: bit p set if the node changed in the last pass p
    uint8_t *tsum;          // per source and tile: OR of the tile nodes' change bits
    int *changed;
    const int *list;        // sources still changing (compacted), first *nlist entries
    const int *nlist;
    const int *sets_done;   // sets completed per source (0 during the first set)
    int pass;               // 0..7, position of this pass in the set
    long long node_stride;  // nodes per source grid
    int nz, nx, ny;         // cells
    int NZ, NX, NY;         // nodes
    int sz, sx, sy;         // pass direction (sgnt of the reference)
    int tile_a;             // z extent of a tile / brick
    int nba, nbb, nbc;      // tile counts along z, x, y (memory grid, nodes 0..N-1)
    int diag;               // brick anti-diagonal of this launch
    double dz, dx, dy, dz2i, dx2i, dy2i, dz2dx2, dz2dy2, dx2dy2, dsum;
};

__device__ __forceinline__ double min4(double a, double b, double c, double d) {
    return fmin(fmin(a, b), fmin(c, d));
}

// FTeik 3D node update in oriented coordinates (a = z, b = x, c = y, all pointing
// downwind).  tv/te/tn: upwind neighbours along z/x/y; tev, ten, tnv, tnve: the
// diagonal ones; sXYZ: slowness of the cell at (a-1 | a') x (b-1 | b') x (c-1 | c'),
// with ' the downwind-side cell clamped to the grid (the reference's max(i-1,0) /
// min(i, n-2) pair).  Returns the new value; *which receives the winning candidate.
__device__ __forceinline__ double node_update(const SweepArgs &A, double t0, double tv, double te,
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

// Each workgroup runs W waves on the same brick for W different sources (from the
// compacted list of sources still changing): the slowness block of every staged plane is
// loaded once into LDS and shared, each wave owns the traveltime block of its source.
//
// Bricks are the tiles of one fixed grid in memory coordinates (tile_a x 8 x 8 nodes), so
// every pass visits the same tiles whatever its direction; a per-source byte per tile
// records in which of the last 8 passes any of its nodes changed.  A tile whose own byte
// and the bytes of its upwind neighbour tiles are all zero contains only nodes that the
// node-level rule below would skip, so the whole brick is skipped.
template <bool GRAD, int W>
__global__ __launch_bounds__(64 * W) void sweep3d_brick_kernel(const SweepArgs A) {
    constexpr int kEnt = kPlane * (W + 1);  // staged entries per plane: W node blocks + cells
    constexpr int kThreads = 64 * W;
    constexpr int NE = (kEnt + kThreads - 1) / kThreads;
    const int nact = *A.nlist;
    const int g0 = blockIdx.y * W;
    if (g0 >= nact) return;  // block-uniform: every source of this group has converged

    // brick (BA, BB, BC) on anti-diagonal A.diag, enumerated with BA outermost
    int t = blockIdx.x, BA = 0, BB = 0, BC = 0;
    {
        const int d = A.diag;
        const int amin = max(0, d - (A.nbb - 1) - (A.nbc - 1));
        const int amax = min(A.nba - 1, d);
        for (int ba = amin; ba <= amax; ++ba) {
            const int e = d - ba;
            const int bmin = max(0, e - (A.nbc - 1)), bmax = min(A.nbb - 1, e);
            const int cnt = bmax - bmin + 1;
            if (t < cnt) {
                BA = ba;
                BB = bmin + t;
                BC = e - BB;
                break;
            }
            t -= cnt;
        }
    }

    const int NX = A.NX, NY = A.NY;
    // oriented brick origin: tile BA/BB/BC of the memory grid seen from this direction
    const int a0 = A.tile_a * BA - (A.sz > 0 ? 0 : A.tile_a * A.nba - A.NZ);
    const int b0 = kLanesB * BB - (A.sx > 0 ? 0 : kLanesB * A.nbb - NX);
    const int c0 = kLanesC * BC - (A.sy > 0 ? 0 : kLanesC * A.nbc - NY);
    const int a_first = max(1, a0);
    const int na = min(A.NZ - 1, a0 + A.tile_a - 1) - a_first + 1;  // planes to update
    if (na <= 0 || b0 + kLanesB - 1 < 1 || c0 + kLanesC - 1 < 1) return;
    // tile of the memory grid and its upwind neighbours
    const int ti = A.sz > 0 ? BA : A.nba - 1 - BA;
    const int tj = A.sx > 0 ? BB : A.nbb - 1 - BB;
    const int tk = A.sy > 0 ? BC : A.nbc - 1 - BC;
    const long long tiles = (long long)A.nba * A.nbb * A.nbc;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ int wave_on[W];
    {
        const bool has = g0 + wave < nact;
        const int s0 = has ? A.list[g0 + wave] : 0;
        bool on = has;
        if (has && A.sets_done[s0] > 0) {
            // lanes 0..7 read the tile byte and its 7 upwind neighbours
            unsigned any = 0;
            if (lane < 8) {
                const int ii = ti - ((lane & 1) ? A.sz : 0);
                const int jj = tj - ((lane & 2) ? A.sx : 0);
                const int kk = tk - ((lane & 4) ? A.sy : 0);
                if (ii >= 0 && ii < A.nba && jj >= 0 && jj < A.nbb && kk >= 0 && kk < A.nbc)
                    any = A.tsum[s0 * tiles + ((long long)ii * A.nbb + jj) * A.nbc + kk];
            }
            on = __any(any != 0u);
        }
        if (lane == 0) wave_on[wave] = on ? 1 : 0;
    }
    __syncthreads();
    bool block_on = false;
#pragma unroll
    for (int w = 0; w < W; ++w) block_on |= wave_on[w] != 0;
    if (!block_on) return;  // block-uniform

    const bool has_src = wave_on[wave] != 0;
    const int src = has_src ? A.list[g0 + wave] : 0;
    const int bb = lane >> 3, cc = lane & 7;
    const int b = b0 + bb, c = c0 + cc;
    const bool inb = has_src && b >= 1 && b <= NX - 1 && c >= 1 && c <= NY - 1;

    // oriented -> memory index maps
    auto node_i = [&](int a) { return A.sz > 0 ? a : A.NZ - 1 - a; };
    auto node_j = [&](int x) { return A.sx > 0 ? x : NX - 1 - x; };
    auto node_k = [&](int y) { return A.sy > 0 ? y : NY - 1 - y; };
    auto cell_i = [&](int a) { return A.sz > 0 ? a : A.nz - 1 - a; };
    auto cell_j = [&](int x) { return A.sx > 0 ? x : A.nx - 1 - x; };
    auto cell_k = [&](int y) { return A.sy > 0 ? y : A.ny - 1 - y; };

    double *tt = A.tt + (long long)src * A.node_stride;
    uint8_t *sg = GRAD ? A.sgn + (long long)src * A.node_stride : nullptr;
    const long long plane_nodes = (long long)NX * NY;
    const long long plane_cells = (long long)A.nx * A.ny;

    // Staging entries of this thread: e = tid + 64 W n.  e < 81 W: node (eb, ec) of the
    // block of wave e / 81; then the 81 cells.  Every entry reads a valid address (nodes
    // outside the grid read the slowness base and are replaced by Big).
    const double *ent_base[NE];
    const uint8_t *ent_chg[NE];
    bool ent_node[NE], ent_big[NE], ent_use[NE], ent_c[NE];
#pragma unroll
    for (int n = 0; n < NE; ++n) {
        const int e = tid + kThreads * n;
        ent_use[n] = e < kEnt;
        ent_c[n] = e < kPlane * W;
        ent_node[n] = false;
        ent_big[n] = false;
        ent_base[n] = A.slow;
        ent_chg[n] = A.chg;
        if (e < kPlane * W) {
            const int w = e / kPlane, r = e - w * kPlane;
            const int eb = b0 - 1 + r / 9, ec = c0 - 1 + r % 9;
            if (wave_on[w] && eb >= 0 && eb <= NX - 1 && ec >= 0 && ec <= NY - 1) {
                ent_node[n] = true;
                const long long off = (long long)A.list[g0 + w] * A.node_stride +
                                      (long long)node_j(eb) * NY + node_k(ec);
                ent_base[n] = A.tt + off;
                ent_chg[n] = A.chg + off;
            } else {
                ent_big[n] = true;
            }
        } else if (e < kEnt) {
            const int r = e - kPlane * W;
            const int cb = min(max(b0 - 1 + r / 9, 0), A.nx - 1);
            const int ccl = min(max(c0 - 1 + r % 9, 0), A.ny - 1);
            ent_base[n] = A.slow + (long long)cell_j(cb) * A.ny + cell_k(ccl);
        }
    }

    __shared__ double ring[kRing][kEnt];
    __shared__ uint8_t ringc[kRing][kPlane * W];

    // plane p of the brick (a = a_first + p; p = -1 is the upwind plane) lives in slot
    // (p+1) % kRing
    auto load_plane = [&](int p, double *v, unsigned *cv) {
        const int a = a_first + p;
        const long long tb = (long long)node_i(a) * plane_nodes;
        const long long sb = (long long)cell_i(min(a, A.nz - 1)) * plane_cells;
#pragma unroll
        for (int n = 0; n < NE; ++n) {
            const double x = ent_base[n][ent_node[n] ? tb : sb];
            v[n] = ent_big[n] ? kBig : x;
            const unsigned cb = ent_chg[n][ent_node[n] ? tb : 0];
            cv[n] = ent_node[n] ? cb : 0u;
        }
    };
    auto store_plane = [&](int p, const double *v, const unsigned *cv) {
        double *slot = ring[(p + 1) % kRing];
        uint8_t *cslot = ringc[(p + 1) % kRing];
#pragma unroll
        for (int n = 0; n < NE; ++n) {
            if (ent_use[n]) slot[tid + kThreads * n] = v[n];
            if (ent_c[n]) cslot[tid + kThreads * n] = (uint8_t)cv[n];
        }
    };

    for (int p = -1; p < min(kPrefetch, na); ++p) {
        double v[NE];
        unsigned cv[NE];
        load_plane(p, v, cv);
        store_plane(p, v, cv);
    }
    // planes kPrefetch .. kPrefetch+kDepth-1 in flight in registers
    double rv[kDepth][NE];
    unsigned rc[kDepth][NE];
#pragma unroll
    for (int k = 0; k < kDepth; ++k)
        if (kPrefetch + k < na) load_plane(kPrefetch + k, rv[k], rc[k]);
    __syncthreads();
    // From the second set on, a node none of whose 7 upwind neighbours changed during the
    // last 8 passes sees exactly the candidates of its previous update in this direction,
    // while its own time can only have decreased since: it cannot change and is skipped.
    const bool may_skip = has_src && A.sets_done[src] > 0;
    const unsigned pbit = 1u << A.pass;
    uint8_t *cg = A.chg + (long long)src * A.node_stride;

    const int tblk = wave * kPlane, sblk = W * kPlane;
    const int self = (bb + 1) * 9 + (cc + 1);  // this lane's node in a block
    const int up_b = bb * 9 + (cc + 1);        // b-1
    const int up_c = (bb + 1) * 9 + cc;        // c-1
    const int up_bc = bb * 9 + cc;             // b-1, c-1
    const long long col_off = inb ? (long long)node_j(b) * NY + node_k(c) : 0;
    const uint8_t code_z = sgn_code(A.sz), code_x = (uint8_t)(sgn_code(A.sx) << 2),
                  code_y = (uint8_t)(sgn_code(A.sy) << 4);

    unsigned hist = 0;   // bit k: this lane changed its node k steps ago
    unsigned chist = 0;  // bit k: this lane changed its node's change bits k steps ago
    bool any_change = false;
    const int nsteps = na + kSkew;
    for (int s0 = 0; s0 < nsteps; s0 += kDepth) {
#pragma unroll
        for (int k = 0; k < kDepth; ++k) {
            const int s = s0 + k;
            const int p = s - bb - cc;
            bool changed_now = false, cchanged = false;
            if (inb && p >= 0 && p < na) {
                double *cur = ring[(p + 1) % kRing];
                const double *prv = ring[p % kRing];
                uint8_t *curc = ringc[(p + 1) % kRing] + tblk;
                const uint8_t *prvc = ringc[p % kRing] + tblk;
                const unsigned up = curc[up_b] | curc[up_c] | curc[up_bc] | prvc[self] |
                                    prvc[up_b] | prvc[up_c] | prvc[up_bc];
                const double t0 = cur[tblk + self];
                if (!may_skip || up != 0u) {
                    int which;
                    const double tnew = node_update(
                        A, t0, prv[tblk + self], cur[tblk + up_b], cur[tblk + up_c],
                        prv[tblk + up_b], cur[tblk + up_bc], prv[tblk + up_c], prv[tblk + up_bc],
                        prv[sblk + up_bc], prv[sblk + up_b], prv[sblk + up_c], prv[sblk + self],
                        cur[sblk + up_bc], cur[sblk + up_b], cur[sblk + up_c], &which);
                    cur[tblk + self] = tnew;
                    if (tnew != t0) {
                        changed_now = true;
                        if (GRAD) {
                            uint8_t code = 0;
                            if (which & 1) code |= code_z;
                            if (which & 2) code |= code_x;
                            if (which & 4) code |= code_y;
                            sg[(long long)node_i(a_first + p) * plane_nodes + col_off] = code;
                        }
                    }
                }
                const unsigned cs = curc[self];
                const unsigned cn = (cs & ~pbit) | (changed_now ? pbit : 0u);
                if (cn != cs) {
                    curc[self] = (uint8_t)cn;
                    cchanged = true;
                }
            }
            hist = (hist << 1) | (changed_now ? 1u : 0u);
            chist = (chist << 1) | (cchanged ? 1u : 0u);
            any_change |= changed_now;

            // write back the plane every lane has finished (only what changed)
            const int pw = s - kSkew;
            if (pw >= 0 && pw < na && inb) {
                const long long g = (long long)node_i(a_first + pw) * plane_nodes + col_off;
                if ((hist >> (kSkew - bb - cc)) & 1u) tt[g] = ring[(pw + 1) % kRing][tblk + self];
                if ((chist >> (kSkew - bb - cc)) & 1u) cg[g] = ringc[(pw + 1) % kRing][tblk + self];
            }

            // plane s+kPrefetch leaves the registers for its slot, whose previous plane was
            // last read one step ago; its register set takes the plane kDepth further on
            const int pl = s + kPrefetch;
            if (pl < na) store_plane(pl, rv[k], rc[k]);
            if (pl + kDepth < na) load_plane(pl + kDepth, rv[k], rc[k]);
            __syncthreads();
        }
    }

    if (has_src) {
        const bool wave_changed = __any(any_change);
        if (lane == 0) {
            if (wave_changed) A.changed[src] = 1;
            uint8_t *ts = A.tsum + src * tiles + ((long long)ti * A.nbb + tj) * A.nbc + tk;
            const unsigned old = *ts;
            const unsigned nw = (old & ~pbit) | (wave_changed ? pbit : 0u);
            if (nw != old) *ts = (uint8_t)nw;
        }
    }
}

// tt = Big everywhere (and signs cleared), _fteik3d.py:429-437
__global__ void fill_kernel(double *tt, long long n, uint8_t *sgn, uint8_t *chg) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        tt[i] = kBig;
        chg[i] = 0;
        if (sgn) sgn[i] = 0;
    }
}

__global__ void zero_bytes_kernel(uint8_t *p, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

__device__ __forceinline__ double t_ana3(int i, int j, int k, double dz, double dx, double dy,
                                         const SrcParam &s, double vzero) {
    const double a = dz * ((double)i - s.zsa), b = dx * ((double)j - s.xsa),
                 c = dy * ((double)k - s.ysa);
    return vzero * sqrt(a * a + b * b + c * c);
}

// Seeds the 8 corners of each source cell (_fteik3d.py:439-456), records vzero, resets
// the per-source set bookkeeping.  One thread per (source, corner).
__global__ void seed3d_kernel(const double *slow, int nz, int nx, int ny, double dz, double dx,
                              double dy, const SrcParam *prm, int nsrc, double *tt,
                              long long node_stride, double *vzero, int *sets_done,
                              int *changed, int *active, int *list, int *nlist) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nsrc * 8) return;
    const int s = g >> 3, c = g & 7;
    const SrcParam P = prm[s];
    const double vz = slow[((long long)P.zsi * nx + P.xsi) * ny + P.ysi];
    // corner order of the reference: (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (1,0,1) (0,1,1) (1,1,1)
    const int di = (c == 1 || c == 4 || c == 5 || c == 7);
    const int dj = (c == 2 || c == 4 || c == 6 || c == 7);
    const int dk = (c == 3 || c == 5 || c == 6 || c == 7);
    const int i = P.zsi + di, j = P.xsi + dj, k = P.ysi + dk;
    tt[(long long)s * node_stride + ((long long)i * (nx + 1) + j) * (ny + 1) + k] =
        t_ana3(i, j, k, dz, dx, dy, P, vz);
    if (c == 0) {
        vzero[s] = vz;
        sets_done[s] = 0;
        changed[s] = 0;
        active[s] = 1;
        list[s] = s;
        if (s == 0) *nlist = nsrc;
    }
}

// After each set: count it for every active source, retire sources whose set changed
// nothing (further sets would repeat it exactly), clear the change flags and rebuild the
// compacted list of sources the next set sweeps.  One block.
__global__ void set_bookkeeping_kernel(int nsrc, int *sets_done, int *changed, int *active,
                                       int *list, int *nlist) {
    for (int s = threadIdx.x; s < nsrc; s += blockDim.x) {
        if (active[s]) {
            sets_done[s] += 1;
            if (!changed[s]) active[s] = 0;
        }
        changed[s] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int n = 0;
        for (int s = 0; s < nsrc; ++s)
            if (active[s]) list[n++] = s;
        *nlist = n;
    }
}

// Gradient post-pass, _fteik3d.py:464-488: components with a recorded upwind sign are
// one-sided differences, the others keep the analytic seed value (0 away from the source
// cell); the result is normalised.  One thread per node.
__global__ void gradient3d_kernel(const double *tt, const uint8_t *sgn, int NZ, int NX, int NY,
                                  double dz, double dx, double dy, SrcParam P, double vzero,
                                  double *g) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nn = (long long)NZ * NX * NY;
    if (n >= nn) return;
    const int k = (int)(n % NY), j = (int)((n / NY) % NX), i = (int)(n / ((long long)NX * NY));
    double gz = 0.0, gx = 0.0, gy = 0.0;
    const int di = i - P.zsi, dj = j - P.xsi, dk = k - P.ysi;
    if (di >= 0 && di <= 1 && dj >= 0 && dj <= 1 && dk >= 0 && dk <= 1) {
        // analytic derivative of the seed, _fteik3d.py:20-35
        const double t = t_ana3(i, j, k, dz, dx, dy, P, vzero);
        if (t > 0.0) {
            const double tmp = vzero * vzero / t;
            gz = ((double)i - P.zsa) * dz * tmp;
            gx = ((double)j - P.xsa) * dx * tmp;
            gy = ((double)k - P.ysa) * dy * tmp;
        }
    }
    const uint8_t s = sgn[n];
    const double tc = tt[n];
    const int sz = sgn_axis(s, 0), sx = sgn_axis(s, 1), sy = sgn_axis(s, 2);
    if (sz != 0) gz = sz * (tc - tt[n - (long long)sz * NX * NY]) / dz;
    if (sx != 0) gx = sx * (tc - tt[n - (long long)sx * NY]) / dx;
    if (sy != 0) gy = sy * (tc - tt[n - sy]) / dy;
    const double gn = sqrt(gz * gz + gx * gx + gy * gy);
    if (gn > 0.0) {
        gz /= gn;
        gx /= gn;
        gy /= gn;
    }
    g[3 * n] = gz;
    g[3 * n + 1] = gx;
    g[3 * n + 2] = gy;
}

// Host: source placement and bound check, _fteik3d.py:402-422.
bool place_source3(const double *s, int nz, int nx, int ny, double dz, double dx, double dy,
                   SrcParam *P) {
    const double z = s[0], x = s[1], y = s[2];
    const bool ok = (0.0 <= z && z <= dz * nz) && (0.0 <= x && x <= dx * nx) &&
                    (0.0 <= y && y <= dy * ny);
    if (!ok) return false;
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

int ceil_div(int a, int b) { return (a + b - 1) / b; }

int bricks_on_diag(int d, int nba, int nbb, int nbc) {
    int total = 0;
    const int amin = std::max(0, d - (nbb - 1) - (nbc - 1)), amax = std::min(nba - 1, d);
    for (int ba = amin; ba <= amax; ++ba) {
        const int e = d - ba;
        const int bmin = std::max(0, e - (nbc - 1)), bmax = std::min(nbb - 1, e);
        if (bmax >= bmin) total += bmax - bmin + 1;
    }
    return total;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

template <bool G, int W>
void launch_brick_t(dim3 grid, hipStream_t st, const SweepArgs &a) {
    hipLaunchKernelGGL((sweep3d_brick_kernel<G, W>), grid, dim3(64 * W), 0, st, a);
}

void launch_brick(bool grad, int group, dim3 grid, hipStream_t st, const SweepArgs &a) {
    switch (group * 2 + (grad ? 1 : 0)) {
        case 2: launch_brick_t<false, 1>(grid, st, a); break;
        case 3: launch_brick_t<true, 1>(grid, st, a); break;
        case 4: launch_brick_t<false, 2>(grid, st, a); break;
        case 5: launch_brick_t<true, 2>(grid, st, a); break;
        case 8: launch_brick_t<false, 4>(grid, st, a); break;
        case 9: launch_brick_t<true, 4>(grid, st, a); break;
        case 16: launch_brick_t<false, 8>(grid, st, a); break;
        default: launch_brick_t<true, 8>(grid, st, a); break;
    }
}

struct Ws3 {
    SrcParam *prm;
    int *changed, *active, *list, *nlist;
    uint8_t *chg, *tsum;
};

Ws3 carve3(void *ws, int nsrc, long long nodes) {
    char *p = (char *)ws;
    Ws3 w;
    w.prm = (SrcParam *)p;
    p += align256(sizeof(SrcParam) * nsrc);
    w.changed = (int *)p;
    p += align256(sizeof(int) * nsrc);
    w.active = (int *)p;
    p += align256(sizeof(int) * nsrc);
    w.list = (int *)p;
    p += align256(sizeof(int) * nsrc);
    w.nlist = (int *)p;
    p += 256;
    w.chg = (uint8_t *)p;
    p += align256((size_t)nodes * nsrc);
    w.tsum = (uint8_t *)p;
    return w;
}

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" const char *fteik_version(void) { return "fteik_amd 0.1.0 (gfx950)"; }

// Tiles of the memory grid (z extent tile_a is at least 1): an upper bound on the tile
// count for any tile_a, used to size the per-source tile bytes.
static size_t max_tiles(int nz, int nx, int ny) {
    return (size_t)(nz + 1) * ((nx + 1 + kLanesB - 1) / kLanesB) * ((ny + 1 + kLanesC - 1) / kLanesC);
}

extern "C" size_t fteik3d_workspace_size(int nz, int nx, int ny, int nsrc) {
    const size_t nodes = (size_t)(nz + 1) * (nx + 1) * (ny + 1);
    return align256(sizeof(SrcParam) * nsrc) + 3 * align256(sizeof(int) * nsrc) + 256 +
           align256(nodes * nsrc) + align256(max_tiles(nz, nx, ny) * nsrc);
}

extern "C" int fteik3d_solve(const double *slow, int nz, int nx, int ny, double dz, double dx,
                             double dy, const double *src, int nsrc, int nsweep, int max_sets,
                             int want_grad, int tile_a, int group, double *tt, uint8_t *sgn,
                             double *vzero, int *sets_done, void *workspace,
                             size_t workspace_bytes, int use_graph, void *stream_) {
    if (nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || nsrc > 65535 || !slow || !src || !tt ||
        !vzero || !sets_done || !workspace)
        return FTEIK_ERR_ARG;
    if (want_grad && !sgn) return FTEIK_ERR_ARG;
    if (!(dz > 0.0) || !(dx > 0.0) || !(dy > 0.0)) return FTEIK_ERR_ARG;
    if (workspace_bytes < fteik3d_workspace_size(nz, nx, ny, nsrc)) return FTEIK_ERR_WORKSPACE;
    if (nsweep <= 0 && max_sets < 1) return FTEIK_ERR_ARG;
    if (group == 0) group = nsrc >= 64 ? 4 : (nsrc >= 8 ? 2 : 1);
    if (group != 1 && group != 2 && group != 4 && group != 8) return FTEIK_ERR_ARG;
    hipStream_t stream = (hipStream_t)stream_;

    std::vector<SrcParam> prm(nsrc);
    for (int s = 0; s < nsrc; ++s)
        if (!place_source3(src + 3 * s, nz, nx, ny, dz, dx, dy, &prm[s]))
            return FTEIK_ERR_SOURCE_OOB;

    const int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    const long long nn = (long long)NZ * NX * NY;
    Ws3 w = carve3(workspace, nsrc, nn);
    FTEIK_CHECK(hipMemcpyAsync(w.prm, prm.data(), sizeof(SrcParam) * nsrc, hipMemcpyHostToDevice,
                               stream));
    // the host vector goes out of scope on return: finish the copy first
    FTEIK_CHECK(hipStreamSynchronize(stream));

    {
        const long long total = nn * nsrc;
        const int threads = 256;
        const long long want = (total + threads - 1) / threads;
        const int blocks = (int)std::min<long long>(want, 65536);
        hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(threads), 0, stream, tt, total,
                           want_grad ? sgn : (uint8_t *)nullptr, w.chg);
        FTEIK_CHECK(hipGetLastError());
        hipLaunchKernelGGL(seed3d_kernel, dim3((nsrc * 8 + 255) / 256), dim3(256), 0, stream,
                           slow, nz, nx, ny, dz, dx, dy, w.prm, nsrc, tt, nn, vzero, sets_done,
                           w.changed, w.active, w.list, w.nlist);
        FTEIK_CHECK(hipGetLastError());
    }

    SweepArgs a;
    a.slow = slow;
    a.tt = tt;
    a.sgn = want_grad ? sgn : nullptr;
    a.changed = w.changed;
    a.list = w.list;
    a.nlist = w.nlist;
    a.chg = w.chg;
    a.tsum = w.tsum;
    a.sets_done = sets_done;
    a.node_stride = nn;
    a.nz = nz; a.nx = nx; a.ny = ny;
    a.NZ = NZ; a.NX = NX; a.NY = NY;
    if (tile_a <= 0) tile_a = nsrc >= 32 ? NZ : std::min(NZ, 16);
    a.tile_a = std::min(tile_a, NZ);
    a.nba = ceil_div(NZ, a.tile_a);
    a.nbb = ceil_div(NX, kLanesB);
    a.nbc = ceil_div(NY, kLanesC);
    {
        const long long nt = (long long)a.nba * a.nbb * a.nbc * nsrc;
        hipLaunchKernelGGL(zero_bytes_kernel, dim3((unsigned)std::min<long long>((nt + 255) / 256, 65536)),
                           dim3(256), 0, stream, w.tsum, nt);
        FTEIK_CHECK(hipGetLastError());
    }
    a.dz = dz; a.dx = dx; a.dy = dy;
    a.dz2i = 1.0 / dz / dz;
    a.dx2i = 1.0 / dx / dx;
    a.dy2i = 1.0 / dy / dy;
    a.dz2dx2 = a.dz2i * a.dx2i;
    a.dz2dy2 = a.dz2i * a.dy2i;
    a.dx2dy2 = a.dx2i * a.dy2i;
    a.dsum = a.dz2i + a.dx2i + a.dy2i;
    const int ndiag = a.nba + a.nbb + a.nbc - 2;
    std::vector<int> nb(ndiag);
    for (int d = 0; d < ndiag; ++d) nb[d] = bricks_on_diag(d, a.nba, a.nbb, a.nbc);

    // pass order of the reference, _fteik3d.py:217-391
    static const int dirs[8][3] = {{1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {1, -1, -1},
                                   {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    auto enqueue_set = [&]() -> int {
        for (int p = 0; p < 8; ++p) {
            a.sz = dirs[p][0];
            a.sx = dirs[p][1];
            a.sy = dirs[p][2];
            a.pass = p;
            for (int d = 0; d < ndiag; ++d) {
                a.diag = d;
                launch_brick(want_grad != 0, group, dim3(nb[d], ceil_div(nsrc, group)), stream, a);
            }
        }
        hipLaunchKernelGGL(set_bookkeeping_kernel, dim3(1), dim3(256), 0, stream, nsrc, sets_done,
                           w.changed, w.active, w.list, w.nlist);
        return hip_status(hipGetLastError());
    };

    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    if (use_graph && stream) {  // the legacy default stream cannot be captured
        hipStreamCaptureMode mode = hipStreamCaptureModeThreadLocal;
        FTEIK_CHECK(hipStreamBeginCapture(stream, mode));
        int st = enqueue_set();
        hipError_t e = hipStreamEndCapture(stream, &graph);
        if (st) return st;
        FTEIK_CHECK(e);
        FTEIK_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    auto run_set = [&]() -> int {
        if (exec) return hip_status(hipGraphLaunch(exec, stream));
        return enqueue_set();
    };

    int status = 0;
    if (nsweep > 0) {
        for (int s = 0; s < nsweep && !status; ++s) status = run_set();
    } else {
        int nact = nsrc;
        for (int s = 0; s < max_sets && nact > 0 && !status; ++s) {
            status = run_set();
            if (status) break;
            hipError_t e = hipMemcpyAsync(&nact, w.nlist, sizeof(int), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            status = hip_status(e);
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return status;
}

extern "C" int fteik3d_gradient(const double *tt, const uint8_t *sgn, int nz, int nx, int ny,
                                double dz, double dx, double dy, const double *src, double vzero,
                                double *ttgrad, void *stream_) {
    if (!tt || !sgn || !ttgrad || !src || nz < 1 || nx < 1 || ny < 1) return FTEIK_ERR_ARG;
    SrcParam P;
    if (!place_source3(src, nz, nx, ny, dz, dx, dy, &P)) return FTEIK_ERR_SOURCE_OOB;
    const long long nn = (long long)(nz + 1) * (nx + 1) * (ny + 1);
    hipLaunchKernelGGL(gradient3d_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream_, tt, sgn, nz + 1, nx + 1, ny + 1, dz, dx, dy, P,
                       vzero, ttgrad);
    return hip_status(hipGetLastError());
}
