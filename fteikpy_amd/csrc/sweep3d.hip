// 3D fast-sweeping eikonal solver for MI355X (gfx950): FTeik first-order operators
// (fteikpy/_fteik/_fteik3d.py:38-200) swept over 8 directions per set
// (_fteik3d.py:203-391), with source seeding (_fteik3d.py:394-456) and the gradient
// post-pass (_fteik3d.py:464-488).
//
// Exactness.  Within one directional pass a node reads only itself and 7 upwind
// neighbours, all on strictly lower anti-diagonal levels of that direction, so every
// topological order of a pass yields the reference's state bit for bit (FP contraction is
// disabled at build time).
//
// Bricks.  The grid is cut into tiles of tile_a (z) x 8 (x) x 8 (y) nodes fixed in memory
// coordinates; in each pass a tile is processed as one brick by one 64-lane wave per
// source.  Lane (bb, cc) owns column (x, y) = (b0+bb, c0+cc) of the pass's oriented
// coordinates and, at step s, works on plane p = s - bb - cc: all lanes of a step lie on one
// node hyperplane, so every upwind value was produced at an earlier step.  Traveltimes and
// change bytes of z planes are staged through an LDS ring with coalesced global loads (2
// planes in flight in registers) and written back coalesced once every lane is past them;
// each lane reads the slowness cells around its node from global memory (cached, shared by
// all sources).  W waves of a workgroup run the same brick for W sources.
//
// Skipping.  Per node, one byte records in which of the last 8 passes the node changed; a
// node whose 7 upwind neighbours all read zero is skipped (its candidates equal those of
// its previous update in this direction and its own time can only have decreased).  Per
// tile, one byte ORs its nodes' bytes; a tile whose byte and its upwind neighbours' bytes
// are zero is skipped whole.  In the first set the bytes only accumulate (the seeded
// corners start with all bits set), so a zero byte means a node still at its initial 1e5:
// a node whose 7 upwind neighbours all hold 1e5 builds only candidates >= 1e5 and keeps its
// time, and the same skips apply from the first set on (see guard3d_kernel for the bound
// on the slowness this needs).
//
// Schedules.  "overlapped": one launch list per sweep set in which each launch runs every
// brick, of any pass, whose inputs are final (schedule.hpp).  "diagonal": one launch per
// brick anti-diagonal and pass.  Both are ordered by the stream (and replayed from a
// captured graph).  "persistent": one cooperative launch per set walks the same list with
// a grid-wide barrier between its entries.
#include "fteik_amd.h"
#include "fteik_device.hpp"
#include "node_update.hpp"
#include "history.hpp"
#include "schedule.hpp"


#include <algorithm>
#include <climits>
#include <cstring>
#include <memory>
#include <mutex>
#include <cmath>
#include <vector>

namespace fteik {
namespace {

constexpr int kLanesB = sched::kLanesB;  // brick width along x (oriented b)
constexpr int kLanesC = sched::kLanesC;  // brick width along y (oriented c)
constexpr int kSkew = kLanesB + kLanesC - 2;  // steps between first and last lane = 14
constexpr int kPlane = (kLanesB + 1) * (kLanesC + 1);  // 81 entries: node or cell block
constexpr int kPrefetch = 1;    // planes resident in LDS ahead of lane (0,0)
constexpr int kDepth = 2;       // further planes in flight in registers (loop unrolled by it)
// a slot is rewritten at the end of step s for plane s+kPrefetch; its previous plane was
// last read (as the upwind plane of lane (7,7)) at step s-1, before that step's barrier
constexpr int kRing = kPrefetch + kSkew + 2;

// LDS slot pitches: 8 mod 16 (doubles, resp. dwords) spreads the lanes' skewed accesses
// over the banks.
__host__ __device__ constexpr int dpitch(int n) { return n + ((8 - n % 16) + 16) % 16; }
__host__ __device__ constexpr int bpitch(int n) {
    return 4 * ((n + 3) / 4 + ((8 - ((n + 3) / 4) % 16) + 16) % 16);
}

struct SweepArgs : Stencil3 {
    const double *slow;
    double *tt;
    uint8_t *sgn;
    uint8_t *chg;           // per node: bit p set if the node changed in the last pass p
    uint8_t *tsum;          // per source and tile: bit p set if a node of the tile changed
                            // in the latest pass p (bits only accumulate in the first set)
    uint64_t *faces;        // per source and tile: byte f = OR of the change bits of the
                            // tile's face f (0/1 first/last node plane, 2/3 first/last x
                            // column, 4/5 first/last y column); null: tile bytes only
    int *changed;
    const int *list;        // sources still changing (compacted), first *nlist entries
    const int *nlist;
    const int *sets_done;   // sets completed per source (0 during the first set)
    const int *guard;       // 0: the first set may skip (every cell's slowness is large
                            // enough for untouched neighbours to stay untouched)
    long long node_stride;  // nodes per source grid
    int nz, nx, ny;         // cells
    int NZ, NX, NY;         // nodes
    int nsrc;               // sources of the model (list, sets_done entries)
    int tile_a;             // z extent of a tile / brick
    int nba, nbb, nbc;      // tile counts along z, x, y (memory grid, nodes 0..N-1)
    int ntiles, ndiag;
    int gmax;               // source groups (workgroups per brick)
    int pass;               // pass-by-pass schedule: pass of this launch
    int diag;               // pass-by-pass schedule: brick anti-diagonal of this launch;
                            // overlapped schedule: first entry of this launch in tasks
    const unsigned long long *tasks;  // overlapped schedule: packed bricks (pass, BA, BB,
                                      // BC) by launch; null: pass-by-pass
    unsigned *log_node;     // logged solve: changed node (source * nodes + node) ...
    double *log_old;        // ... and its time before the change
    unsigned long long *log_count;
    long long log_cap;
};

using sched::decode_brick;
using sched::pass_sx;
using sched::pass_sy;
using sched::pass_sz;

// One brick for W sources (wave w handles list[g0 + w]).  Every early return is uniform
// across the workgroup.
// Logged solve: a wave reserves history entries kLogChunk at a time (one counter update
// per chunk); the unused tail of a chunk is filled with kLogUnused, which the reverse skips.
constexpr unsigned kLogChunk = 256;
constexpr unsigned kLogUnused = 0xffffffffu;
__device__ __forceinline__ void log_pad(const SweepArgs &A, unsigned long long pos,
                                        unsigned long long end, int lane) {
    for (unsigned long long e = pos + lane; e < end; e += 64)
        if (e < (unsigned long long)A.log_cap) A.log_node[e] = kLogUnused;
}

// Skip test of a brick for source s0: slot q = 0 reads the tile byte, q = 1..7 what the 7
// upwind neighbour tiles (bit 0: along z, 1: x, 2: y) hold next to it: the face toward
// this tile (its z face for the neighbours below or above, its x face for the x and xy
// neighbours, a superset of the edge or corner they share), or their whole-tile byte
// without face bytes.  Zero when the neighbour lies outside the grid.
__device__ __forceinline__ unsigned neighbour_byte(const SweepArgs &A, int s0, int q, int ti, int tj,
                                                   int tk, int sz, int sx, int sy) {
    const int ii = ti - ((q & 1) ? sz : 0);
    const int jj = tj - ((q & 2) ? sx : 0);
    const int kk = tk - ((q & 4) ? sy : 0);
    if (ii < 0 || ii >= A.nba || jj < 0 || jj >= A.nbb || kk < 0 || kk >= A.nbc) return 0u;
    const long long t = s0 * (long long)A.ntiles + ((long long)ii * A.nbb + jj) * A.nbc + kk;
    if (q == 0 || A.faces == nullptr) return A.tsum[t];
    const int f = (q & 1) ? (sz > 0 ? 1 : 0) : ((q & 2) ? (sx > 0 ? 3 : 2) : (sy > 0 ? 5 : 4));
    return (unsigned)(A.faces[t] >> (8 * f)) & 0xffu;
}

// End of a sweep step (and of the ring's first fill): the next step reads ring entries
// that other threads wrote.  In a one-wave workgroup they are all lanes of the same wave,
// whose LDS instructions execute in order, so ordering the accesses for the compiler
// suffices (no wait for the LDS writes to complete, no barrier); wider workgroups stage
// entries across waves and wait at a workgroup barrier.
template <int W>
__device__ __forceinline__ void step_sync() {
    if constexpr (W == 1) wave_sync();
    else __syncthreads();
}

// pre < 0: the brick tests each wave's source for skipping itself; else bit w of pre says
// whether wave w's source is active (the caller tested it, the same test).
template <bool GRAD, bool LOG, int W, bool BF = false>
__device__ void brick(const SweepArgs &A, int pass, int BA, int BB, int BC, int g0, int nact,
                      int pre = -1) {
    constexpr int kThreads = 64 * W;
    constexpr int kT = kPlane * W;                           // node entries (W blocks)
    constexpr int NT = (kT + kThreads - 1) / kThreads;       // = 2
    constexpr int kPitch = dpitch(kT);
    constexpr int kCPitch = bpitch(kT);
    __shared__ double ring[kRing][kPitch];
    __shared__ uint8_t ringc[kRing][kCPitch];
    __shared__ int wave_on[W];

    const int sz = pass_sz(pass), sx = pass_sx(pass), sy = pass_sy(pass);
    const int NX = A.NX, NY = A.NY;
    // oriented brick origin: tile BA/BB/BC of the memory grid seen from this direction
    const int a0 = A.tile_a * BA - (sz > 0 ? 0 : A.tile_a * A.nba - A.NZ);
    const int b0 = kLanesB * BB - (sx > 0 ? 0 : kLanesB * A.nbb - NX);
    const int c0 = kLanesC * BC - (sy > 0 ? 0 : kLanesC * A.nbc - NY);
    const int a_first = max(1, a0);
    const int na = min(A.NZ - 1, a0 + A.tile_a - 1) - a_first + 1;  // planes to update
    if (na <= 0 || b0 + kLanesB - 1 < 1 || c0 + kLanesC - 1 < 1) return;
    // tile of the memory grid and its upwind neighbours
    const int ti = sz > 0 ? BA : A.nba - 1 - BA;
    const int tj = sx > 0 ? BB : A.nbb - 1 - BB;
    const int tk = sy > 0 ? BC : A.nbc - 1 - BC;
    const long long tiles = A.ntiles;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bb = lane >> 3, cc = lane & 7;
    const int b = b0 + bb, c = c0 + cc;

    // oriented -> memory index maps
    auto node_i = [&](int a) { return sz > 0 ? a : A.NZ - 1 - a; };
    auto node_j = [&](int x) { return sx > 0 ? x : NX - 1 - x; };
    auto node_k = [&](int y) { return sy > 0 ? y : NY - 1 - y; };
    auto cell_i = [&](int a) { return sz > 0 ? a : A.nz - 1 - a; };
    auto cell_j = [&](int x) { return sx > 0 ? x : A.nx - 1 - x; };
    auto cell_k = [&](int y) { return sy > 0 ? y : A.ny - 1 - y; };
    const long long plane_nodes = (long long)NX * NY;
    const long long plane_cells = (long long)A.nx * A.ny;

    // The brick's global loads come in batches, each issued before any of it is waited
    // on.  Batch 1: the source of each of the W blocks, the first-set test and the guard.
    // Every source runs the first set, so "not the first set" is read from any source's
    // set count (that of source g0) rather than behind the list entry.
    // nact < 0: the caller left the count of sources still changing to this round (the
    // list entries are read ahead of it, within the list's nsrc entries).
    int srcw[W];  // source of block w, -1 for none
#pragma unroll
    for (int w = 0; w < W; ++w) srcw[w] = g0 + w < A.nsrc ? A.list[g0 + w] : -1;
    const bool first = A.sets_done[g0] == 0;  // g0 < nsrc for every launched block
    const bool guard_clear = *A.guard == 0;
    if (nact < 0) nact = *A.nlist;
    if (g0 >= nact) return;  // no source left in this group (uniform)
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (g0 + w >= nact) srcw[w] = -1;
    const int s0 = srcw[wave];
    const bool has = s0 >= 0;
    // skipping is allowed from the second set on, and in the first set while the guard holds
    const bool may_skip = has && (!first || guard_clear);

    // Batch 2: the skip bytes of this wave's source (lanes 0..7).  A wave with a source
    // sweeps unless it may skip and neither the tile nor an upwind neighbour holds a change
    // next to it (pre >= 0: the caller decided).
    {
        const unsigned skip_byte =
            may_skip && lane < 8 ? neighbour_byte(A, s0, lane, ti, tj, tk, sz, sx, sy) : 0u;
        bool on = has;
        if (may_skip) on = __any(skip_byte != 0u);
        if (pre >= 0) on = ((pre >> wave) & 1) != 0;
        if (lane == 0) wave_on[wave] = on ? 1 : 0;
    }
    __syncthreads();
    bool block_on = false;
#pragma unroll
    for (int w = 0; w < W; ++w) block_on |= wave_on[w] != 0;
    if (!block_on) return;

    const bool has_src = wave_on[wave] != 0;
    const int src = has_src ? s0 : 0;
    // ing: this lane's column is in the grid; inb: its nodes are updated in this pass (the
    // oriented planes b = 0 and c = 0 are not)
    const bool ing = has_src && b >= 0 && b <= NX - 1 && c >= 0 && c <= NY - 1;
    const bool inb = ing && b >= 1 && c >= 1;
    double *tt = A.tt + (long long)src * A.node_stride;
    uint8_t *sg = GRAD ? A.sgn + (long long)src * A.node_stride : nullptr;

    // Batch 3, active bricks only, every load issued before any is waited on: the first
    // planes for the ring, the planes in flight in registers, this lane's slowness cells.
    // Staging: node entries e = tid + 64 W n (< 81 W; node (eb, ec) of block e / 81),
    // traveltime and change byte, of the blocks whose source sweeps; nodes outside the
    // grid read as Big.
    long long t_off[NT];
    bool t_use[NT], t_ok[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int e = tid + kThreads * n;
        t_use[n] = e < kT;
        t_ok[n] = false;
        t_off[n] = 0;
        if (e < kT) {
            const int w = e / kPlane, r = e - w * kPlane;
            const int eb = b0 - 1 + r / 9, ec = c0 - 1 + r % 9;
            if (wave_on[w] && eb >= 0 && eb <= NX - 1 && ec >= 0 && ec <= NY - 1) {
                t_ok[n] = true;
                t_off[n] = (long long)srcw[w] * A.node_stride + (long long)node_j(eb) * NY +
                           node_k(ec);
            }
        }
    }

    // plane p of the brick (a = a_first + p; p = -1 is the upwind plane) lives in slot
    // (p+1) % kRing
    struct Stage {
        double t[NT];
        unsigned c[NT];
    };
    auto load_plane = [&](int p, Stage &v) {
        const int a = a_first + p;
        const long long tb = (long long)node_i(a) * plane_nodes;
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            v.t[n] = kBig;
            v.c[n] = 0u;
            if (t_ok[n]) {
                v.t[n] = A.tt[t_off[n] + tb];
                v.c[n] = A.chg[t_off[n] + tb];
            }
        }
    };
    auto store_plane = [&](int p, const Stage &v) {
        double *slot = ring[(p + 1) % kRing];
        uint8_t *cslot = ringc[(p + 1) % kRing];
#pragma unroll
        for (int n = 0; n < NT; ++n)
            if (t_use[n]) {
                slot[tid + kThreads * n] = v.t[n];
                cslot[tid + kThreads * n] = (uint8_t)v.c[n];
            }
    };
    // planes -1 .. kPrefetch-1 go to the ring first, planes kPrefetch .. kPrefetch+kDepth-1
    // stay in flight in registers
    Stage v0[kPrefetch + 1];
#pragma unroll
    for (int p = -1; p < kPrefetch; ++p)
        if (p < na) load_plane(p, v0[p + 1]);
    Stage rv[kDepth];
#pragma unroll
    for (int k = 0; k < kDepth; ++k)
        if (kPrefetch + k < na) load_plane(kPrefetch + k, rv[k]);

    // Slowness: each lane streams the 4 cells (b-1 | b') x (c-1 | c') around its column (' =
    // clamped to the grid's last cell) on the cell plane a' = min(a, nz - 1) of its current
    // node from global memory, one step ahead; the previous step's 4 are the upwind plane's
    // (a - 1 = the previous a').  Read-only data shared by every source, it stays in cache,
    // and keeping it out of the LDS ring lets 3 workgroups' waves share a SIMD.  Offsets
    // within a cell plane fit 32 bits (checked on the host).
    int scol[4];
    {
        const int bm = min(max(b - 1, 0), A.nx - 1), bp = min(max(b, 0), A.nx - 1);
        const int cm = min(max(c - 1, 0), A.ny - 1), cp = min(max(c, 0), A.ny - 1);
        scol[0] = cell_j(bm) * A.ny + cell_k(cm);
        scol[1] = cell_j(bm) * A.ny + cell_k(cp);
        scol[2] = cell_j(bp) * A.ny + cell_k(cm);
        scol[3] = cell_j(bp) * A.ny + cell_k(cp);
    }
    auto load_cells = [&](int p, double (&v)[4]) {
        if (inb) {
            const int pc = min(max(p, -1), na - 1);
            const long long pb = (long long)cell_i(min(a_first + pc, A.nz - 1)) * plane_cells;
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = A.slow[pb + scol[q]];
        }
    };
    double sp[4] = {0.0, 0.0, 0.0, 0.0}, sc[4] = {0.0, 0.0, 0.0, 0.0};
    load_cells(-bb - cc - 1, sp);
    load_cells(-bb - cc, sc);

#pragma unroll
    for (int p = -1; p < kPrefetch; ++p)
        if (p < na) store_plane(p, v0[p + 1]);
    step_sync<W>();
    // From the second set on, a node none of whose 7 upwind neighbours changed during the
    // last 8 passes sees exactly the candidates of its previous update in this direction,
    // while its own time can only have decreased since: it cannot change and is skipped.
    // In the first set the bytes only accumulate, and the same test finds 7 neighbours at
    // their initial time (valid when the guard holds).
    const unsigned pbit = 1u << pass;
    const unsigned keep = first ? 0xffu : ~pbit;  // bits kept besides this pass's
    uint8_t *cg = A.chg + (long long)src * A.node_stride;

    const Stencil3 C = stencil_in_vgprs(A);
    const int tblk = wave * kPlane;
    const int self = (bb + 1) * 9 + (cc + 1);  // this lane's node in a block
    const int up_b = bb * 9 + (cc + 1);        // b-1
    const int up_c = (bb + 1) * 9 + cc;        // c-1
    const int up_bc = bb * 9 + cc;             // b-1, c-1
    const long long col_off = inb ? (long long)node_j(b) * NY + node_k(c) : 0;
    const uint8_t code_z = sgn_code(sz), code_x = (uint8_t)(sgn_code(sx) << 2),
                  code_y = (uint8_t)(sgn_code(sy) << 4);

    // LOG: this wave's reserved history chunk [log_pos, log_end)
    unsigned long long log_pos = 0, log_end = 0;
    unsigned hist = 0;   // bit k: this lane changed its node k steps ago
    unsigned chist = 0;  // bit k: this lane changed its node's change bits k steps ago
    bool any_change = false;
    // face bytes of this tile: final change bytes of its nodes on the first/last node
    // plane, and (per lane, applied at the end) of the lane's column
    unsigned face_zlo = 0, face_zhi = 0, col_or = 0;
    const int tile_ilo = ti * A.tile_a, tile_ihi = min(A.NZ, tile_ilo + A.tile_a) - 1;
    if (a0 <= 0 && ing) {
        // the tile holds the oriented plane a = 0, which no pass in this direction updates:
        // its bytes (upwind plane, slot 0) still belong in the faces other tiles read
        const unsigned c0b = ringc[0][tblk + self];
        const int mi = node_i(0);
        col_or |= c0b;
        if (mi == tile_ilo) face_zlo |= c0b;
        if (mi == tile_ihi) face_zhi |= c0b;
    }
    const int nsteps = na + kSkew;
    // ring slots advance by one per step: this lane's current plane p = s - bb - cc sits in
    // slot (p+1) mod kRing, the plane written back at step s in slot (s-kSkew+1) mod kRing
    int cslot = (1 - bb - cc + 2 * kRing) % kRing;
    int wslot = (1 - kSkew + 2 * kRing) % kRing;
    for (int s0 = 0; s0 < nsteps; s0 += kDepth) {
#pragma unroll
        for (int k = 0; k < kDepth; ++k) {
            const int s = s0 + k;
            const int p = s - bb - cc;
            bool changed_now = false, cchanged = false;
            double t_prev = 0.0;  // LOG: time before the change
            const int pslot = cslot == 0 ? kRing - 1 : cslot - 1;
            if (inb && (unsigned)p < (unsigned)na) {
                double *cur = ring[cslot];
                const double *prv = ring[pslot];
                uint8_t *curc = ringc[cslot] + tblk;
                const uint8_t *prvc = ringc[pslot] + tblk;
                const unsigned up = curc[up_b] | curc[up_c] | curc[up_bc] | prvc[self] |
                                    prvc[up_b] | prvc[up_c] | prvc[up_bc];
                const double t0 = cur[tblk + self];
                if (!may_skip || up != 0u) {
                    int which;
                    const double tnew = node_update<BF>(
                        C, t0, prv[tblk + self], cur[tblk + up_b], cur[tblk + up_c],
                        prv[tblk + up_b], cur[tblk + up_bc], prv[tblk + up_c], prv[tblk + up_bc],
                        sp[0], sp[1], sp[2], sp[3], sc[0], sc[1], sc[2], &which);
                    cur[tblk + self] = tnew;
                    if (tnew != t0) {
                        changed_now = true;
                        t_prev = t0;
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
                const unsigned cn = (cs & keep) | (changed_now ? pbit : 0u);
                curc[self] = (uint8_t)cn;  // (unchanged bytes rewritten: no branch)
                cchanged = cn != cs;
            }
            hist = (hist << 1) | (changed_now ? 1u : 0u);
            chist = (chist << 1) | (cchanged ? 1u : 0u);
            any_change |= changed_now;
            if (LOG) {
                // append this wave's changes in lane order into its reserved chunk
                const unsigned long long m = __ballot(changed_now);
                if (m) {
                    const unsigned cnt = (unsigned)__popcll(m);
                    if (log_end - log_pos < cnt) {
                        log_pad(A, log_pos, log_end, lane);
                        unsigned long long base = 0;
                        if (lane == 0) base = atomicAdd(A.log_count, (unsigned long long)kLogChunk);
                        log_pos = __shfl(base, 0);
                        log_end = log_pos + kLogChunk;
                    }
                    if (changed_now) {
                        const unsigned long long e =
                            log_pos + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
                        if (e < (unsigned long long)A.log_cap) {
                            A.log_node[e] = (unsigned)((long long)src * A.node_stride +
                                                       (long long)node_i(a_first + p) * plane_nodes +
                                                       col_off);
                            A.log_old[e] = t_prev;
                        }
                    }
                    log_pos += cnt;
                }
            }

            // write back the plane every lane has finished (only what changed)
            const int pw = s - kSkew;
            if (pw >= 0 && pw < na && ing) {
                const int mi = node_i(a_first + pw);
                const unsigned cfin = ringc[wslot][tblk + self];
                if (inb) {
                    const long long g = (long long)mi * plane_nodes + col_off;
                    if ((hist >> (kSkew - bb - cc)) & 1u) tt[g] = ring[wslot][tblk + self];
                    if ((chist >> (kSkew - bb - cc)) & 1u) cg[g] = (uint8_t)cfin;
                }
                col_or |= cfin;
                if (mi == tile_ilo) face_zlo |= cfin;
                if (mi == tile_ihi) face_zhi |= cfin;
            }
            cslot = cslot == kRing - 1 ? 0 : cslot + 1;
            wslot = wslot == kRing - 1 ? 0 : wslot + 1;
            // this plane's cells become the upwind plane's; fetch the next plane's
#pragma unroll
            for (int q = 0; q < 4; ++q) sp[q] = sc[q];
            load_cells(p + 1, sc);

            // plane s+kPrefetch leaves the registers for its slot, whose previous plane was
            // last read one step ago; its register set takes the plane kDepth further on
            const int pl = s + kPrefetch;
            if (pl < na) store_plane(pl, rv[k]);
            if (pl + kDepth < na) load_plane(pl + kDepth, rv[k]);
            step_sync<W>();
        }
    }

    if (LOG) log_pad(A, log_pos, log_end, lane);
    if (A.faces != nullptr && has_src) {
        // the bytes of every node of the tile went in, updated in this pass or not
        const int mj = ing ? node_j(b) : -1, mk = ing ? node_k(c) : -1;
        unsigned f[6] = {face_zlo, face_zhi,
                         mj == tj * kLanesB ? col_or : 0u,
                         mj == min(NX, tj * kLanesB + kLanesB) - 1 ? col_or : 0u,
                         mk == tk * kLanesC ? col_or : 0u,
                         mk == min(NY, tk * kLanesC + kLanesC) - 1 ? col_or : 0u};
        unsigned long long packed = 0;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            unsigned v = f[q];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o);
            packed |= (unsigned long long)(v & 0xffu) << (8 * q);
        }
        if (lane == 0) A.faces[src * tiles + ((long long)ti * A.nbb + tj) * A.nbc + tk] = packed;
    }
    if (has_src) {
        const bool wave_changed = __any(any_change);
        if (lane == 0) {
            if (wave_changed) A.changed[src] = 1;
            uint8_t *ts = A.tsum + src * tiles + ((long long)ti * A.nbb + tj) * A.nbc + tk;
            const unsigned old = *ts;
            const unsigned nw = (old & keep) | (wave_changed ? pbit : 0u);
            if (nw != old) *ts = (uint8_t)nw;
        }
    }
}

// Diagonal schedule: one launch per (pass, brick anti-diagonal); blockIdx.x = brick on the
// diagonal, blockIdx.y = source group.
// The logged variant (history writes) needs a few VGPRs over the 168 that 3 waves per SIMD
// allow: held to 168 it spills 4 (12 bytes per lane) and runs 4% faster than at 2 waves
// (8-source workgroups are held to 2 waves by their LDS anyway).
// BF: the node update without branches around the plane-wave candidates, for solves of so
// few sources that each wave runs alone on its SIMD (one source per workgroup).
template <bool GRAD, bool LOG, int W, bool BF = false>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu((LOG && W <= 4) ? 3 : 1, (LOG && W <= 4) ? 3 : 10)))
void sweep3d_diag_kernel(const SweepArgs A) {
    int pass = A.pass, BA, BB, BC;
    if (A.tasks)
        sched::unpack_task(A.tasks[A.diag + blockIdx.x], &pass, &BA, &BB, &BC);
    else
        decode_brick(A.diag, blockIdx.x, A.nba, A.nbb, A.nbc, &BA, &BB, &BC);
    // the count of sources still changing is read with the brick's first loads
    brick<GRAD, LOG, W, BF>(A, pass, BA, BB, BC, blockIdx.y * W, -1);
}

// Batch of models: launch l of a sweep set covers the bricks of launch l of every
// model.  cum[m] = first block of model m (non-decreasing, cum[nmodel] = grid size); a
// model's blocks enumerate (brick of this launch, source group), source group fastest; the
// bricks of model m in this launch are tasks[toff[m] ...] (packed pass and oriented brick,
// each model's overlapped schedule).
template <bool GRAD, int W>
__global__ __launch_bounds__(64 * W) void sweep3d_multi_kernel(const SweepArgs *models,
                                                               const int *cum, const int *toff,
                                                               const unsigned long long *tasks,
                                                               int nmodel) {
    const int b = blockIdx.x;
    int lo;  // the last model whose first block is <= b (cum is non-decreasing, cum[0] = 0)
    if (nmodel <= 64) {
        // one load per lane and a vote instead of a search's dependent loads
        const int i = threadIdx.x & 63;
        const bool le = i < nmodel && cum[i] <= b;
        lo = (int)__popcll(__ballot(le)) - 1;
    } else {
        int l = 0, hi = nmodel - 1;
        while (l < hi) {
            const int mid = (l + hi + 1) >> 1;
            if (cum[mid] <= b) l = mid;
            else hi = mid - 1;
        }
        lo = l;
    }
    const SweepArgs A = models[lo];
    const int local = b - cum[lo];
    const int t = local / A.gmax, g = local - t * A.gmax;
    int pass, BA, BB, BC;
    sched::unpack_task(tasks[toff[lo] + t], &pass, &BA, &BB, &BC);
    brick<GRAD, false, W>(A, pass, BA, BB, BC, g * W, -1);
}

// Logged solve: closes the current pass's segment of the history.
__global__ void log_mark_kernel(HistHeader *hdr, unsigned long long *seg,
                                const unsigned long long *count) {
    const int P = hdr->npass;
    if (P < hdr->max_pass) seg[P + 1] = *count;
    hdr->npass = P + 1;
}

// sources up to which the node update runs without branches (see solve3d_impl)
constexpr int kBranchFreeSources = 4;

// default group (0): from kWideFirstSources sources the first kDenseSets sweep sets run 4
// sources per workgroup, the later ones 2 (see solve3d_impl)
constexpr int kWideFirstSources = 32;
constexpr int kDenseSets = 2;
constexpr int kWideGroup = 4;

// a grid-barrier wait longer than this gives up (the launch then ends and reports it)
constexpr unsigned long long kBarWaitSeconds = 5;

// Grid-wide barrier of the persistent schedule's cooperative launch (tree layout and
// counter choice in schedule.hpp, shared with its CPU model).  Every thread fences its
// writes before arriving and after leaving (release / acquire at agent scope).
// A wait that outlasts wait_ticks (wall clock) sets the give-up flag (word 1 of line 0),
// which every waiting workgroup also polls: the launch then ends (every workgroup leaves its
// loop), and the host reports the set's result as invalid.  The cooperative launch keeps
// every workgroup resident, so a wait ends normally within microseconds.
struct TreeBar {
    unsigned *mem;       // line 0: generation, give-up flag; then the counters, level by level
    sched::BarTree t;
    unsigned long long wait_ticks;
};

// Returns false once any wait of the launch gave up.
__device__ bool tree_barrier(const TreeBar &B, unsigned target) {
    __shared__ int s_ok;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        int idx = (int)blockIdx.x;
        bool last = true;
        for (int k = 0; k < B.t.levels && last; ++k) {
            const sched::BarCounter c = sched::bar_counter(B.t, k, idx);
            unsigned *ctr = B.mem + (size_t)c.line * sched::kBarLine;
            const unsigned old =
                __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            last = old == (unsigned)(c.members - 1);
            if (last) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            idx = c.next;
        }
        unsigned *flag = B.mem + 1;
        if (last) {
            __hip_atomic_store(B.mem, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(B.mem, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
                if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
                if (wall_clock64() - t0 > B.wait_ticks) {
                    __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        s_ok = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
    }
    __syncthreads();
    __threadfence();
    return s_ok != 0;
}

// Persistent schedule: one cooperative launch per sweep set, every workgroup resident (the
// runtime refuses the launch otherwise).  The launch walks the set's launch list -- the
// overlapped list (first[l] .. first[l+1] bricks of tasks) or, without one, one entry per
// (pass, brick anti-diagonal) -- and between entries waits at a grid-wide barrier, which
// orders every write of one entry before every read of the next as a launch boundary does.
// Within an entry, workgroups take (brick, source group) items from the entry's counter
// (zeroed before the launch), so a workgroup whose bricks are skipped takes the next.
// Every workgroup runs every entry's loop and barrier: no early exit.
// No register limit that would spill: a kernel using scratch can have fewer waves resident
// than the occupancy query reports (the runtime's scratch slots bound them), and a
// workgroup left waiting for a slot would never reach the grid barrier (the host refuses
// the schedule for a build whose kernel uses scratch).
template <bool GRAD, int W>
__global__ __launch_bounds__(64 * W)
void sweep3d_coop_kernel(const SweepArgs A, const int *first, int nentry, unsigned gy,
                         unsigned *ticket, TreeBar bar) {
    __shared__ unsigned s_item, s_base;
    if (threadIdx.x == 0)
        s_base = __hip_atomic_load(bar.mem, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const unsigned base = s_base;  // read by every workgroup before any arrives
    for (int l = 0; l < nentry; ++l) {
        int nb, f0 = 0, pass = 0, d = 0;
        if (first) {
            f0 = first[l];
            nb = first[l + 1] - f0;
        } else {
            pass = l / A.ndiag;
            d = l - pass * A.ndiag;
            nb = sched::bricks_on_diag(d, A.nba, A.nbb, A.nbc);
        }
        const unsigned nitems = (unsigned)nb * gy;
        // (an item is taken only when the previous one is done: taking the next one ahead
        // leaves it waiting behind a running brick at the end of an entry)
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(ticket + l, 1u);
            __syncthreads();
            const unsigned i = s_item;
            __syncthreads();  // s_item is rewritten only after every thread has read it
            if (i >= nitems) break;
            const int x = (int)(i % (unsigned)nb), y = (int)(i / (unsigned)nb);
            int p = pass, BA, BB, BC;
            if (first)
                sched::unpack_task(A.tasks[f0 + x], &p, &BA, &BB, &BC);
            else
                decode_brick(d, x, A.nba, A.nbb, A.nbc, &BA, &BB, &BC);
            brick<GRAD, false, W>(A, p, BA, BB, BC, y * W, -1);
        }
        if (!tree_barrier(bar, base + (unsigned)l + 1u)) break;  // uniform in the workgroup
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
// the per-source set bookkeeping.  The corners' change bytes, and the tile and face bytes
// of the tiles of the first set's geometry that hold them, start with every bit set; the
// first-set guard starts cleared.  One thread per (source, corner).
struct SeedArgs {
    const SrcParam *prm;
    double *tt;
    double *vzero;
    int *sets_done, *changed, *active, *list, *nlist;
    uint8_t *chg, *tsum;
    uint64_t *faces;      // null: no face bytes in the first set
    int *guard;
    long long node_stride;
    int nsrc, nz, nx, ny;
    int tile_a, nbb, nbc, ntiles;  // first set's tile geometry
    double dz, dx, dy;
};

__global__ void seed3d_kernel(const double *slow, SeedArgs A) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nsrc * 8) return;
    const int s = g >> 3, c = g & 7;
    const SrcParam P = A.prm[s];
    const int nx = A.nx, ny = A.ny;
    const double vz = slow[((long long)P.zsi * nx + P.xsi) * ny + P.ysi];
    // corner order of the reference: (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (1,0,1) (0,1,1) (1,1,1)
    const int di = (c == 1 || c == 4 || c == 5 || c == 7);
    const int dj = (c == 2 || c == 4 || c == 6 || c == 7);
    const int dk = (c == 3 || c == 5 || c == 6 || c == 7);
    const int i = P.zsi + di, j = P.xsi + dj, k = P.ysi + dk;
    const long long n = (long long)s * A.node_stride + ((long long)i * (nx + 1) + j) * (ny + 1) + k;
    A.tt[n] = t_ana3(i, j, k, A.dz, A.dx, A.dy, P, vz);
    A.chg[n] = 0xff;
    const long long t = (long long)s * A.ntiles +
                        ((long long)(i / A.tile_a) * A.nbb + j / kLanesB) * A.nbc + k / kLanesC;
    A.tsum[t] = 0xff;
    if (A.faces) A.faces[t] = ~0ull;
    if (c == 0) {
        A.vzero[s] = vz;
        A.sets_done[s] = 0;
        A.changed[s] = 0;
        A.active[s] = 1;
        A.list[s] = s;
        if (s == 0) {
            *A.nlist = A.nsrc;
            *A.guard = 0;
        }
    }
}

// First-set skip guard.  A node whose 7 upwind neighbours all hold the initial 1e5 builds
// candidates 1e5 + dz vref (1D), 1e5 + 2 vref dz dx / sqrt(dz^2 + dx^2) (2D, likewise for
// the other planes) and 1e5 + 3 vref / sqrt(1/dz^2 + 1/dx^2 + 1/dy^2) (3D), each added term
// >= sqrt(2) vref min(dz, dx, dy) in exact arithmetic, while rounding the 2D and 3D forms
// moves them by at most ~10 ulp(1e5) = 1.5e-10.  With every cell's slowness times the
// smallest spacing at least 1e-9 (vref is a cell slowness), every candidate stays >= 1e5 and
// the node keeps its time, so it may be skipped.  A cell below that (or not a number) sets
// the guard and the first set updates every node.
constexpr double kGuardTime = 1e-9;
__global__ void guard3d_kernel(const double *slow, long long ncell, double hmin, int *guard) {
    bool bad = false;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x)
        bad |= !(slow[c] * hmin >= kGuardTime);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(guard, 1);
}

// After each set: count it for every active source, retire sources whose set changed
// nothing (further sets would repeat it exactly), clear the change flags and rebuild the
// compacted list of sources the next set sweeps.  One block.
__device__ void bookkeeping(int nsrc, int *sets_done, int *changed, int *active, int *list,
                            int *nlist) {
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

__global__ void set_bookkeeping_kernel(int nsrc, int *sets_done, int *changed, int *active,
                                       int *list, int *nlist) {
    bookkeeping(nsrc, sets_done, changed, active, list, nlist);
}

// Per-model bookkeeping of a batch: one block per model.
struct BookArgs {
    int nsrc;
    int *sets_done, *changed, *active, *list, *nlist;
};
__global__ void set_bookkeeping_multi_kernel(const BookArgs *b) {
    const BookArgs B = b[blockIdx.x];
    bookkeeping(B.nsrc, B.sets_done, B.changed, B.active, B.list, B.nlist);
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

using sched::bricks_on_diag;

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Source groups (grid.y) of every sweep launch of an instantiated set graph, from old_gy to
// gy: the kernel nodes whose grid.y is old_gy, other than the set's bookkeeping kernel
// (one block).
int set_graph_groups(hipGraph_t graph, hipGraphExec_t exec, unsigned old_gy, unsigned gy) {
    size_t n = 0;
    FTEIK_CHECK(hipGraphGetNodes(graph, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    FTEIK_CHECK(hipGraphGetNodes(graph, nodes.data(), &n));
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType t;
        FTEIK_CHECK(hipGraphNodeGetType(nodes[i], &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams p;
        FTEIK_CHECK(hipGraphKernelNodeGetParams(nodes[i], &p));
        if (p.gridDim.y != old_gy || (p.gridDim.x == 1 && p.gridDim.y == 1)) continue;
        p.gridDim.y = gy;
        FTEIK_CHECK(hipGraphExecKernelNodeSetParams(exec, nodes[i], &p));
    }
    return 0;
}

// A captured sweep set, destroyed on every return path.
struct Graphs {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Graphs() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Set graphs of single-model solves kept across calls.  A later solve replays a captured
// set when its launches would take the same arguments: the key holds everything those
// derive from (device, caller buffers, workspace, grid, spacing, source count, brick
// height, group, gradient flag, schedule).  A solve takes its entry out of the cache while
// it uses it, so two threads never share one, and puts it back when it returns; beyond
// kGraphCacheMax entries the least recently used is destroyed once its last replay ended.
struct SetGraphKey {
    int dev;
    const double *slow;
    const double *tt;
    const uint8_t *sgn;
    const int *sets_done;
    const void *ws;
    int nz, nx, ny, nsrc, tile_a, group, grad, schedule, overlapped;
    uint64_t dz, dx, dy;  // bit patterns of the spacings
};

bool same_key(const SetGraphKey &a, const SetGraphKey &b) {
    return a.dev == b.dev && a.slow == b.slow && a.tt == b.tt && a.sgn == b.sgn &&
           a.sets_done == b.sets_done && a.ws == b.ws && a.nz == b.nz && a.nx == b.nx &&
           a.ny == b.ny && a.nsrc == b.nsrc && a.tile_a == b.tile_a && a.group == b.group &&
           a.grad == b.grad && a.schedule == b.schedule && a.overlapped == b.overlapped &&
           a.dz == b.dz && a.dx == b.dx && a.dy == b.dy;
}

struct SetGraph {
    SetGraphKey key{};
    Graphs g;
    hipEvent_t done = nullptr;           // recorded after the last replay of a solve
    std::vector<hipGraphNode_t> sweep;   // the sweep launches: every kernel node but one
    bool known = false;                  // exactly one bookkeeping node was found
    unsigned gy = 0;                     // source groups the sweep launches have now
    ~SetGraph() {
        if (done) {
            (void)hipEventSynchronize(done);
            (void)hipEventDestroy(done);
        }
    }
};

constexpr size_t kGraphCacheMax = 8;
std::mutex g_graph_mu;
// least recently used first; never destroyed at exit (the HIP runtime may be gone by then)
std::vector<std::unique_ptr<SetGraph>> &g_graphs = *new std::vector<std::unique_ptr<SetGraph>>;
long long g_graph_hits = 0, g_graph_misses = 0;

std::unique_ptr<SetGraph> graph_checkout(const SetGraphKey &k) {
    std::lock_guard<std::mutex> lock(g_graph_mu);
    for (size_t i = g_graphs.size(); i-- > 0;)
        if (same_key(g_graphs[i]->key, k)) {
            std::unique_ptr<SetGraph> g = std::move(g_graphs[i]);
            g_graphs.erase(g_graphs.begin() + (long)i);
            ++g_graph_hits;
            return g;
        }
    ++g_graph_misses;
    return nullptr;
}

void graph_checkin(std::unique_ptr<SetGraph> g) {
    std::vector<std::unique_ptr<SetGraph>> drop;  // destroyed outside the lock
    {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        for (size_t i = g_graphs.size(); i-- > 0;)
            if (same_key(g_graphs[i]->key, g->key)) {
                drop.push_back(std::move(g_graphs[i]));
                g_graphs.erase(g_graphs.begin() + (long)i);
            }
        g_graphs.push_back(std::move(g));
        while (g_graphs.size() > kGraphCacheMax) {
            drop.push_back(std::move(g_graphs.front()));
            g_graphs.erase(g_graphs.begin());
        }
    }
}

// The sweep launches of a captured set: every kernel node except the set's bookkeeping
// kernel, recognised by its function.  Without exactly one such node the graph is used
// for this solve only.
int find_sweep_nodes(SetGraph &s) {
    size_t n = 0;
    FTEIK_CHECK(hipGraphGetNodes(s.g.graph, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    FTEIK_CHECK(hipGraphGetNodes(s.g.graph, nodes.data(), &n));
    int book = 0;
    s.sweep.clear();
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType t;
        FTEIK_CHECK(hipGraphNodeGetType(nodes[i], &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams p;
        FTEIK_CHECK(hipGraphKernelNodeGetParams(nodes[i], &p));
        if (p.func == reinterpret_cast<void *>(set_bookkeeping_kernel)) ++book;
        else s.sweep.push_back(nodes[i]);
    }
    s.known = book == 1;
    if (!s.known) s.sweep.clear();
    return 0;
}

// Source groups (grid.y) of every sweep launch of a set whose sweep nodes are known.
int set_sweep_groups(SetGraph &s, unsigned gy) {
    if (s.gy == gy) return 0;
    for (hipGraphNode_t node : s.sweep) {
        hipKernelNodeParams p;
        FTEIK_CHECK(hipGraphKernelNodeGetParams(node, &p));
        p.gridDim.y = gy;
        FTEIK_CHECK(hipGraphExecKernelNodeSetParams(s.g.exec, node, &p));
    }
    s.gy = gy;
    return 0;
}

template <bool G, int W>
void launch_diag_t(dim3 grid, hipStream_t st, const SweepArgs &a, bool bf) {
    if constexpr (W == 1) {
        if (bf) {
            hipLaunchKernelGGL((sweep3d_diag_kernel<G, false, 1, true>), grid, dim3(64), 0, st, a);
            return;
        }
    }
    hipLaunchKernelGGL((sweep3d_diag_kernel<G, false, W>), grid, dim3(64 * W), 0, st, a);
}

template <bool G, int W>
void launch_diag_log_t(dim3 grid, hipStream_t st, const SweepArgs &a) {
    hipLaunchKernelGGL((sweep3d_diag_kernel<false, true, W>), grid, dim3(64 * W), 0, st, a);
}

// Persistent schedule: the set's cooperative launch (error code of the launch call).
template <bool G, int W>
void launch_coop_t(int blocks, hipStream_t st, const SweepArgs &a, const int *first, int nentry,
                   unsigned gy, unsigned *ticket, const TreeBar &bar, hipError_t *err) {
    SweepArgs args = a;
    const int *f = first;
    int n = nentry;
    unsigned g = gy;
    unsigned *t = ticket;
    TreeBar b = bar;
    void *params[] = {&args, &f, &n, &g, &t, &b};
    *err = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(sweep3d_coop_kernel<G, W>),
                                      dim3(blocks), dim3(64 * W), params, 0, st);
}

#define FTEIK_DISPATCH_W(FN, grad, group, ...)                          \
    switch ((group) * 2 + ((grad) ? 1 : 0)) {                           \
        case 2: FN<false, 1>(__VA_ARGS__); break;                       \
        case 3: FN<true, 1>(__VA_ARGS__); break;                        \
        case 4: FN<false, 2>(__VA_ARGS__); break;                       \
        case 5: FN<true, 2>(__VA_ARGS__); break;                        \
        case 8: FN<false, 4>(__VA_ARGS__); break;                       \
        case 9: FN<true, 4>(__VA_ARGS__); break;                        \
        case 16: FN<false, 8>(__VA_ARGS__); break;                      \
        default: FN<true, 8>(__VA_ARGS__); break;                       \
    }

// Workgroups of the persistent schedule's launch: as many as can be resident at once.
template <bool G, int W>
int pers_blocks_t() {
    int dev = 0, cus = 0, per = 0, coop = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
        !coop)
        return -1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, sweep3d_coop_kernel<G, W>, 64 * W,
                                                     0) != hipSuccess ||
        per < 1)
        return -1;
    // every workgroup must be resident at once: refuse a kernel that needs scratch memory
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(sweep3d_coop_kernel<G, W>)) !=
            hipSuccess ||
        fa.localSizeBytes != 0)
        return -1;
    return cus * per;
}

int pers_blocks(bool grad, int group) {
    switch (group * 2 + (grad ? 1 : 0)) {
        case 2: return pers_blocks_t<false, 1>();
        case 3: return pers_blocks_t<true, 1>();
        case 4: return pers_blocks_t<false, 2>();
        case 5: return pers_blocks_t<true, 2>();
        case 8: return pers_blocks_t<false, 4>();
        case 9: return pers_blocks_t<true, 4>();
        case 16: return pers_blocks_t<false, 8>();
        default: return pers_blocks_t<true, 8>();
    }
}

struct Ws3 {
    SrcParam *prm;
    int *changed, *active, *list, *nlist, *guard;
    int *first;         // persistent schedule: launch list boundaries (overlapped list)
    unsigned *ticket;   // persistent schedule: item counter per launch list entry
    unsigned *bar;      // persistent schedule: grid barrier lines (sched::kBarLines lines of 128 bytes)
    uint8_t *chg, *tsum;
    uint64_t *faces;
    unsigned long long *tasks;
};

using sched::max_diag;
using sched::max_tiles;

// Face bytes are kept for bricks of at least kFaceMinTile planes (tiles of any smaller
// height use the whole-tile bytes of their neighbours).
constexpr int kFaceMinTile = 8;
size_t face_tiles(int nz, int nx, int ny) {
    return (size_t)((nz + 1 + kFaceMinTile - 1) / kFaceMinTile) *
           ((nx + 1 + kLanesB - 1) / kLanesB) * ((ny + 1 + kLanesC - 1) / kLanesC);
}

// The overlapped schedule's brick list (8 bricks per tile) is kept for bricks of at least
// kOverlapMinTile planes; shorter bricks run pass by pass.
constexpr int kOverlapMinTile = kFaceMinTile;

size_t ws_bytes(int nz, int nx, int ny, int nsrc, Ws3 *w, char *base) {
    const size_t nodes = (size_t)(nz + 1) * (nx + 1) * (ny + 1);
    const size_t tiles = max_tiles(nz, nx, ny);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    char *prm = take(sizeof(SrcParam) * nsrc);
    char *changed = take(sizeof(int) * nsrc);
    char *active = take(sizeof(int) * nsrc);
    char *list = take(sizeof(int) * nsrc);
    char *misc = take(sizeof(int) * 8);  // nlist, guard
    // a set's launch list has at most 8 entries per anti-diagonal of the tile grid
    const size_t entries = 8 * max_diag(nz, nx, ny);
    char *first = take(sizeof(int) * (entries + 1));
    char *ticket = take(sizeof(unsigned) * entries);
    char *bar = take(sizeof(unsigned) * sched::kBarLine * sched::kBarLines);
    char *chg = take(nodes * nsrc);
    char *tsum = take(tiles * nsrc);
    char *faces = take(sizeof(uint64_t) * face_tiles(nz, nx, ny) * nsrc);
    char *tasks = take(sizeof(unsigned long long) * 8 * face_tiles(nz, nx, ny));
    if (w) {
        w->prm = (SrcParam *)prm;
        w->changed = (int *)changed;
        w->active = (int *)active;
        w->list = (int *)list;
        w->nlist = (int *)misc;
        w->guard = (int *)misc + 1;
        w->first = (int *)first;
        w->ticket = (unsigned *)ticket;
        w->bar = (unsigned *)bar;
        w->chg = (uint8_t *)chg;
        w->tsum = (uint8_t *)tsum;
        w->faces = (uint64_t *)faces;
        w->tasks = (unsigned long long *)tasks;
    }
    return off;
}

// Seeds every source of one model and evaluates the first-set guard (after the byte
// arrays were cleared).  `a` carries the first set's tile geometry.
int launch_seed3d(const SweepArgs &a, const Ws3 &w, const double *slow, int nsrc, double *vzero,
                  int *sets_done, hipStream_t stream) {
    SeedArgs sa;
    sa.prm = w.prm;
    sa.tt = a.tt;
    sa.vzero = vzero;
    sa.sets_done = sets_done;
    sa.changed = w.changed;
    sa.active = w.active;
    sa.list = w.list;
    sa.nlist = w.nlist;
    sa.chg = w.chg;
    sa.tsum = w.tsum;
    sa.faces = a.faces;
    sa.guard = w.guard;
    sa.node_stride = a.node_stride;
    sa.nsrc = nsrc;
    sa.nz = a.nz; sa.nx = a.nx; sa.ny = a.ny;
    sa.tile_a = a.tile_a; sa.nbb = a.nbb; sa.nbc = a.nbc; sa.ntiles = a.ntiles;
    sa.dz = a.dz; sa.dx = a.dx; sa.dy = a.dy;
    hipLaunchKernelGGL(seed3d_kernel, dim3((nsrc * 8 + 255) / 256), dim3(256), 0, stream, slow, sa);
    FTEIK_CHECK(hipGetLastError());
    const long long ncell = (long long)a.nz * a.nx * a.ny;
    const double hmin = std::min(a.dz, std::min(a.dx, a.dy));
    hipLaunchKernelGGL(guard3d_kernel, dim3((unsigned)std::min<long long>((ncell + 255) / 256, 2048)),
                       dim3(256), 0, stream, slow, ncell, hmin, w.guard);
    return hip_status(hipGetLastError());
}

// brick height default for NZ node planes.  Batches (>= 32 sources): between 20 and 28
// planes (shorter bricks are skipped more often in the sparse later sets, taller ones waste
// less on the 14-step skew), the height with the fewest steps down a column -- bricks x
// (height + 14) -- so that no short last brick pays a whole skew (the taller on ties);
// 16 planes from 16 sources, 8 for a few sources (latency bound).  Measured on MI355X.
int default_tile(int nsrc, int NZ) {
    if (nsrc < 32) return nsrc >= 16 ? 16 : 8;
    int best = 24;
    long long best_steps = -1;
    for (int t = 20; t <= 28; ++t) {
        const long long steps = (long long)ceil_div(NZ, t) * (t + 14);
        if (best_steps < 0 || steps <= best_steps) {
            best = t;
            best_steps = steps;
        }
    }
    return best;
}

// Sweep arguments of one model and its workspace carve-up (the history fields are
// cleared; the logged solve sets them).
void make_args(const double *slow, int nz, int nx, int ny, double dz, double dx, double dy,
               int nsrc, int tile_a, int group, uint8_t *sgn, double *tt, int *sets_done,
               char *workspace, SweepArgs *out, Ws3 *wout) {
    const int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    const long long nn = (long long)NZ * NX * NY;
    Ws3 w;
    ws_bytes(nz, nx, ny, nsrc, &w, workspace);
    SweepArgs a;
    a.slow = slow;
    a.tt = tt;
    a.sgn = sgn;
    a.chg = w.chg;
    a.tsum = w.tsum;
    a.faces = nullptr;
    a.changed = w.changed;
    a.list = w.list;
    a.nlist = w.nlist;
    a.sets_done = sets_done;
    a.guard = w.guard;
    a.node_stride = nn;
    a.nz = nz; a.nx = nx; a.ny = ny;
    a.NZ = NZ; a.NX = NX; a.NY = NY;
    a.nsrc = nsrc;
    if (tile_a <= 0) tile_a = default_tile(nsrc, NZ);
    a.tile_a = std::min(tile_a, NZ);
    a.nba = ceil_div(NZ, a.tile_a);
    a.nbb = ceil_div(NX, kLanesB);
    a.nbc = ceil_div(NY, kLanesC);
    a.ntiles = a.nba * a.nbb * a.nbc;
    a.ndiag = a.nba + a.nbb + a.nbc - 2;
    a.pass = 0;
    a.diag = 0;
    a.tasks = nullptr;
    a.log_node = nullptr;
    a.log_old = nullptr;
    a.log_count = nullptr;
    a.log_cap = 0;
    a.gmax = ceil_div(nsrc, group);
    if (a.tile_a >= kFaceMinTile) a.faces = w.faces;
    static_cast<Stencil3 &>(a) = make_stencil3(dz, dx, dy);
    *out = a;
    *wout = w;
}

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" const char *fteik_version(void) { return "fteik_amd 0.2.0 (gfx950)"; }

extern "C" size_t fteik3d_workspace_size(int nz, int nx, int ny, int nsrc) {
    return ws_bytes(nz, nx, ny, nsrc, nullptr, nullptr);
}

namespace {

// Overlapped-schedule brick lists by tile grid, built once per grid shape; the most recent
// shapes are kept, at most 8 of them and 256 MB of host memory together (a list costs 64
// bytes per tile).
std::mutex g_table_mu;
std::vector<std::shared_ptr<const sched::OverlapTable>> g_tables;
size_t table_bytes(const sched::OverlapTable &t) {
    return t.task.size() * sizeof(unsigned long long) + t.first.size() * sizeof(int);
}
std::shared_ptr<const sched::OverlapTable> overlap_table_for(int nba, int nbb, int nbc) {
    std::lock_guard<std::mutex> lock(g_table_mu);
    for (const auto &t : g_tables)
        if (t->nba == nba && t->nbb == nbb && t->nbc == nbc) return t;
    auto t = std::make_shared<sched::OverlapTable>();
    if (!sched::overlap_table(nba, nbb, nbc, t.get())) return nullptr;
    g_tables.push_back(t);
    size_t total = 0;
    for (const auto &u : g_tables) total += table_bytes(*u);
    while (g_tables.size() > 1 && (g_tables.size() > 8 || total > ((size_t)256 << 20))) {
        total -= table_bytes(*g_tables.front());
        g_tables.erase(g_tables.begin());  // callers holding it keep their copy alive
    }
    return t;
}

// Logged solve: initialises the history header in device memory.
__global__ void log_init_kernel(HistHeader *hdr, unsigned long long *seg, int max_pass,
                                long long cap, long long nodes, int nsrc) {
    hdr->count = 0;
    hdr->npass = 0;
    hdr->max_pass = max_pass;
    hdr->cap = cap;
    hdr->nodes = nodes;
    hdr->nsrc = nsrc;
    hdr->pad = 0;
    seg[0] = 0;
}

int solve3d_impl(const double *slow, int nz, int nx, int ny, double dz, double dx, double dy,
                 const double *src, int nsrc, int nsweep, int max_sets, int want_grad, int tile_a,
                 int group, int schedule, double *tt, uint8_t *sgn, double *vzero, int *sets_done,
                 void *workspace, size_t workspace_bytes, int use_graph, void *stream_,
                 const HistView *hv, long long hcap, long long *events) {
    if (nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || nsrc > 65535 || !slow || !src || !tt ||
        !vzero || !sets_done || !workspace)
        return FTEIK_ERR_ARG;
    if (want_grad && !sgn) return FTEIK_ERR_ARG;
    if (!spacing_ok3(dz, dx, dy)) return FTEIK_ERR_ARG;
    if ((long long)nx * ny > INT_MAX) return FTEIK_ERR_ARG;  // cell-plane offsets are 32-bit
    if (workspace_bytes < fteik3d_workspace_size(nz, nx, ny, nsrc)) return FTEIK_ERR_WORKSPACE;
    if (nsweep <= 0 && max_sets < 1) return FTEIK_ERR_ARG;
    if (schedule < 0 || schedule > 2) return FTEIK_ERR_ARG;
    // defaults measured on MI355X (160x320x320, 256 sources): 2 sources per workgroup
    // beat 4 and 8 (a wider workgroup waits at every step for its slowest source), except
    // in the first kDenseSets sets, where every source sweeps nearly every brick and 4 per
    // workgroup win (set 1 489 against 530 ms, set 2 432 against 452)
    const bool wide_first = group == 0 && nsrc >= kWideFirstSources && schedule != 0;
    if (group == 0) group = nsrc >= 8 ? 2 : 1;
    if (group != 1 && group != 2 && group != 4 && group != 8) return FTEIK_ERR_ARG;
    hipStream_t stream = (hipStream_t)stream_;

    std::vector<SrcParam> prm(nsrc);
    for (int s = 0; s < nsrc; ++s)
        if (!place_source3(src + 3 * s, nz, nx, ny, dz, dx, dy, &prm[s]))
            return FTEIK_ERR_SOURCE_OOB;

    const int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    const long long nn = (long long)NZ * NX * NY;
    // brick height: the caller's, else the library default for this many sources
    Ws3 w;
    SweepArgs a;
    make_args(slow, nz, nx, ny, dz, dx, dy, nsrc, tile_a > 0 ? tile_a : default_tile(nsrc, NZ), group,
              want_grad ? sgn : nullptr, tt, sets_done, (char *)workspace, &a, &w);
    a.log_node = hv ? hv->node : nullptr;
    a.log_old = hv ? hv->old : nullptr;
    a.log_count = hv ? &hv->hdr->count : nullptr;
    a.log_cap = hv ? hcap : 0;
    std::vector<int> nb(a.ndiag);
    for (int d = 0; d < a.ndiag; ++d) nb[d] = bricks_on_diag(d, a.nba, a.nbb, a.nbc);
    const long long gmax = a.gmax;
    // (brick, source group) items of a launch are counted in 32 bits
    if (8LL * a.ntiles * gmax >= (1LL << 31)) return FTEIK_ERR_ARG;
    // overlapped passes: bricks of at least kOverlapMinTile planes, and no history (whose
    // segments are closed pass by pass); otherwise one pass at a time
    std::shared_ptr<const sched::OverlapTable> ov;
    if (schedule != 1 && !hv && a.tile_a >= kOverlapMinTile) {
        ov = overlap_table_for(a.nba, a.nbb, a.nbc);
        if (!ov || ov->task.size() > 8 * face_tiles(nz, nx, ny)) return FTEIK_ERR_ARG;
        a.tasks = w.tasks;
    }
    // a few sources, one per workgroup: each wave runs alone on its SIMD, and the node update
    // without branches around its plane-wave candidates shortens its dependent chain (one
    // source 200^3: 27.7 against 28.3 ms); with more waves sharing a SIMD those branches
    // skip work (256 sources: 3.75 against 3.28 s branch-free)
    const bool bf = nsrc <= kBranchFreeSources && group == 1;
    // persistent schedule: one cooperative launch per set over nentry launch-list entries
    const int pblocks = schedule == 0 ? pers_blocks(want_grad != 0, group) : 0;
    if (schedule == 0 && (pblocks <= 0 || hv)) return FTEIK_ERR_UNSUPPORTED;
    const int nentry = ov ? ov->launches() : 8 * a.ndiag;
    if (schedule == 0 && (long long)nentry > 8LL * (long long)max_diag(nz, nx, ny))
        return FTEIK_ERR_ARG;
    TreeBar bar{};
    bar.mem = w.bar;
    int coop_blocks = 0;
    if (schedule == 0) {
        // no more workgroups than the largest entry has (brick, source group) items
        long long most = 0;
        if (ov) {
            for (int l = 0; l < nentry; ++l)
                most = std::max<long long>(most, ov->first[l + 1] - ov->first[l]);
        } else {
            for (int d = 0; d < a.ndiag; ++d) most = std::max<long long>(most, nb[d]);
        }
        most *= gmax;
        coop_blocks = (int)std::min<long long>(
            {(long long)pblocks, most, (long long)sched::kBarMaxBlocks});
        coop_blocks = std::max(coop_blocks, 1);
        bar.t = sched::bar_tree(coop_blocks);
        int dev = 0, khz = 0;
        FTEIK_CHECK(hipGetDevice(&dev));
        FTEIK_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        bar.wait_ticks = (unsigned long long)std::max(khz, 1) * 1000ull * kBarWaitSeconds;
    }

    FTEIK_CHECK(hipMemcpyAsync(w.prm, prm.data(), sizeof(SrcParam) * nsrc, hipMemcpyHostToDevice,
                               stream));
    if (ov)
        FTEIK_CHECK(hipMemcpyAsync(w.tasks, ov->task.data(),
                                   sizeof(unsigned long long) * ov->task.size(),
                                   hipMemcpyHostToDevice, stream));
    if (ov && schedule == 0)
        FTEIK_CHECK(hipMemcpyAsync(w.first, ov->first.data(), sizeof(int) * ov->first.size(),
                                   hipMemcpyHostToDevice, stream));
    // the host vectors go out of scope on return: finish the copies first
    FTEIK_CHECK(hipStreamSynchronize(stream));
    {
        const long long total = nn * nsrc;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, stream, tt, total,
                           want_grad ? sgn : (uint8_t *)nullptr, w.chg);
        FTEIK_CHECK(hipGetLastError());
        const long long nt = (long long)a.ntiles * nsrc;
        hipLaunchKernelGGL(zero_bytes_kernel, dim3((unsigned)std::min<long long>((nt + 255) / 256, 65536)),
                           dim3(256), 0, stream, w.tsum, nt);
        FTEIK_CHECK(hipGetLastError());
        if (a.faces) {
            const long long nf = nt * (long long)sizeof(uint64_t);
            hipLaunchKernelGGL(zero_bytes_kernel,
                               dim3((unsigned)std::min<long long>((nf + 255) / 256, 65536)),
                               dim3(256), 0, stream, (uint8_t *)a.faces, nf);
            FTEIK_CHECK(hipGetLastError());
        }
        if (int st = launch_seed3d(a, w, slow, nsrc, vzero, sets_done, stream)) return st;
        if (schedule == 0) {  // the barrier's counters start at zero (and end each barrier so)
            const long long nbar = (long long)sizeof(unsigned) * sched::kBarLine * sched::kBarLines;
            hipLaunchKernelGGL(zero_bytes_kernel, dim3((unsigned)((nbar + 255) / 256)), dim3(256), 0,
                               stream, (uint8_t *)w.bar, nbar);
            FTEIK_CHECK(hipGetLastError());
        }
        if (hv) {
            const int max_pass = 8 * (nsweep > 0 ? nsweep : max_sets);
            hipLaunchKernelGGL(log_init_kernel, dim3(1), dim3(1), 0, stream, hv->hdr, hv->seg,
                               max_pass, hcap, nn, nsrc);
            FTEIK_CHECK(hipGetLastError());
        }
    }

    // The sets run in at most two phases, each with its own source group and captured set:
    // the first kDenseSets sets at kWideGroup sources per workgroup when wide_first, every
    // set at `group` otherwise.  Source groups launched per brick (gy): all of them, until
    // a converging solve finds that at most half are still active (their sources
    // compacted to the front of the list), then those (the others' workgroups would only
    // read the count and leave).
    struct Phase {
        int group = 0;
        unsigned gy = 0;
        SetGraphKey key{};
        std::unique_ptr<SetGraph> gr;
    };
    Phase ph[2];
    ph[0].group = wide_first ? kWideGroup : group;
    ph[1].group = group;
    for (Phase &p : ph) p.gy = (unsigned)ceil_div(nsrc, p.group);
    auto phase_of = [&](int set) -> Phase & {
        return wide_first && set < kDenseSets ? ph[0] : ph[1];
    };
    auto enqueue_set = [&](const Phase &ps) -> int {
        const int grp = ps.group;
        const unsigned gy = ps.gy;
        if (schedule == 0) {
            const long long nt = (long long)sizeof(unsigned) * nentry;
            hipLaunchKernelGGL(zero_bytes_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                               stream, (uint8_t *)w.ticket, nt);
            FTEIK_CHECK(hipGetLastError());
            hipError_t err = hipSuccess;
            FTEIK_DISPATCH_W(launch_coop_t, want_grad != 0, grp, coop_blocks, stream, a,
                             ov ? w.first : nullptr, nentry, gy, w.ticket, bar, &err);
            FTEIK_CHECK(err);
        } else if (ov) {
            for (int l = 0; l < ov->launches(); ++l) {
                a.diag = ov->first[l];
                FTEIK_DISPATCH_W(launch_diag_t, want_grad != 0, grp,
                                 dim3(ov->first[l + 1] - ov->first[l], gy), stream, a, bf);
            }
        } else {
            for (int p = 0; p < 8; ++p) {
                a.pass = p;
                for (int d = 0; d < a.ndiag; ++d) {
                    a.diag = d;
                    if (hv) {
                        FTEIK_DISPATCH_W(launch_diag_log_t, false, grp, dim3(nb[d], gy), stream,
                                         a);
                    } else {
                        FTEIK_DISPATCH_W(launch_diag_t, want_grad != 0, grp, dim3(nb[d], gy),
                                         stream, a, bf);
                    }
                }
                if (hv)
                    hipLaunchKernelGGL(log_mark_kernel, dim3(1), dim3(1), 0, stream, hv->hdr,
                                       hv->seg, (const unsigned long long *)&hv->hdr->count);
            }
        }
        hipLaunchKernelGGL(set_bookkeeping_kernel, dim3(1), dim3(256), 0, stream, nsrc, sets_done,
                           w.changed, w.active, w.list, w.nlist);
        return hip_status(hipGetLastError());
    };

    // each phase's sweep set captured once, instantiated on first use, or taken from the
    // cache of earlier solves' sets (the legacy default stream cannot be captured; the
    // persistent schedule and logged solves are not kept)
    const bool graphs = use_graph && stream && schedule != 0;  // cooperative launches run direct
    const bool keep = graphs && !hv && schedule != 0;
    const int nphase = wide_first ? 2 : 1;
    if (keep) {
        int dev = 0;
        FTEIK_CHECK(hipGetDevice(&dev));
        for (int i = 2 - nphase; i < 2; ++i) {
            SetGraphKey &key = ph[i].key;
            key.dev = dev;
            key.slow = slow;
            key.tt = tt;
            key.sgn = want_grad ? sgn : nullptr;
            key.sets_done = sets_done;
            key.ws = workspace;
            key.nz = nz; key.nx = nx; key.ny = ny;
            key.nsrc = nsrc;
            key.tile_a = a.tile_a;
            key.group = ph[i].group;
            key.grad = want_grad ? 1 : 0;
            key.schedule = schedule;
            key.overlapped = ov ? 1 : 0;
            std::memcpy(&key.dz, &dz, sizeof dz);
            std::memcpy(&key.dx, &dx, sizeof dx);
            std::memcpy(&key.dy, &dy, sizeof dy);
            ph[i].gr = graph_checkout(key);
        }
    }
    auto run_set = [&](int set) -> int {
        Phase &ps = phase_of(set);
        if (graphs) {
            std::unique_ptr<SetGraph> &gr = ps.gr;
            if (!gr) {
                gr.reset(new SetGraph);
                gr->key = ps.key;
                FTEIK_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                int st = enqueue_set(ps);
                hipError_t e = hipStreamEndCapture(stream, &gr->g.graph);
                if (st) return st;
                FTEIK_CHECK(e);
                FTEIK_CHECK(hipGraphInstantiate(&gr->g.exec, gr->g.graph, nullptr, nullptr, 0));
                gr->gy = ps.gy;
                if (keep)
                    if (int st2 = find_sweep_nodes(*gr)) return st2;
            } else if (gr->gy != ps.gy) {
                // a kept set still launches the groups its last solve ended with
                if (gr->done) FTEIK_CHECK(hipEventSynchronize(gr->done));
                if (int st2 = set_sweep_groups(*gr, ps.gy)) return st2;
            }
            return hip_status(hipGraphLaunch(gr->g.exec, stream));
        }
        return enqueue_set(ps);
    };

    int status = 0;
    if (nsweep > 0) {
        for (int s = 0; s < nsweep && !status; ++s) status = run_set(s);
    } else {
        int nact = nsrc;
        for (int s = 0; s < max_sets && nact > 0 && !status; ++s) {
            status = run_set(s);
            if (status) break;
            unsigned gave_up = 0;
            hipError_t e = hipMemcpyAsync(&nact, w.nlist, sizeof(int), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess && schedule == 0)
                e = hipMemcpyAsync(&gave_up, w.bar + 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            status = hip_status(e);
            if (!status && gave_up) status = FTEIK_ERR_TIMEOUT;
            if (status) break;
            // the next set's phase launches only the groups that still hold active sources
            Phase &nx = phase_of(s + 1);
            const unsigned need = (unsigned)std::max(1, ceil_div(nact, nx.group));
            if (nact > 0 && 2 * need <= nx.gy) {
                if (nx.gr && nx.gr->known) {
                    status = set_sweep_groups(*nx.gr, need);
                } else if (nx.gr) {
                    status = set_graph_groups(nx.gr->g.graph, nx.gr->g.exec, nx.gy, need);
                    nx.gr->gy = need;
                }
                nx.gy = need;
            }
        }
    }
    if (!status && schedule == 0 && nsweep > 0) {
        // a grid-barrier wait that gave up leaves an invalid result: report it
        unsigned gave_up = 0;
        hipError_t e = hipMemcpyAsync(&gave_up, w.bar + 1, sizeof(unsigned),
                                      hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        status = hip_status(e);
        if (!status && gave_up) status = FTEIK_ERR_TIMEOUT;
    }
    for (Phase &ps : ph) {
        if (!status && keep && ps.gr && ps.gr->known) {
            // kept for later solves with the same arguments, once this one's replays are queued
            if (!ps.gr->done)
                FTEIK_CHECK(hipEventCreateWithFlags(&ps.gr->done, hipEventDisableTiming));
            FTEIK_CHECK(hipEventRecord(ps.gr->done, stream));
            graph_checkin(std::move(ps.gr));
        }
    }
    if (!status && hv) {
        unsigned long long count = 0;
        hipError_t e = hipMemcpyAsync(&count, &hv->hdr->count, sizeof(count),
                                      hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        status = hip_status(e);
        if (events) *events = (long long)count;
        if (!status && count > (unsigned long long)hcap) status = FTEIK_ERR_WORKSPACE;
    }
    return status;
}

}  // namespace

extern "C" int fteik3d_graph_cache(int clear, long long *hits, long long *misses) {
    std::vector<std::unique_ptr<SetGraph>> drop;  // destroyed outside the lock
    int held = 0;
    {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        if (clear) {
            for (auto &g : g_graphs) drop.push_back(std::move(g));
            g_graphs.clear();
        }
        if (hits) *hits = g_graph_hits;
        if (misses) *misses = g_graph_misses;
        held = (int)g_graphs.size();
    }
    return held;
}

extern "C" int fteik3d_sweep_launches(int nz, int nx, int ny, int nsrc, int tile_a,
                                      int schedule) {
    if (nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || nsrc > 65535 || tile_a < 0 || schedule < 0 ||
        schedule > 2)
        return 0;
    if (schedule == 0) return 1;
    const int NZ = nz + 1;
    const int t = std::min(tile_a > 0 ? tile_a : default_tile(nsrc, NZ), NZ);
    const int nba = ceil_div(NZ, t), nbb = ceil_div(nx + 1, kLanesB), nbc = ceil_div(ny + 1, kLanesC);
    if (schedule == 2 && t >= kOverlapMinTile) {
        const auto ov = overlap_table_for(nba, nbb, nbc);
        return ov ? ov->launches() : 0;
    }
    return 8 * (nba + nbb + nbc - 2);
}

extern "C" int fteik3d_solve(const double *slow, int nz, int nx, int ny, double dz, double dx,
                             double dy, const double *src, int nsrc, int nsweep, int max_sets,
                             int want_grad, int tile_a, int group, int schedule, double *tt,
                             uint8_t *sgn, double *vzero, int *sets_done, void *workspace,
                             size_t workspace_bytes, int use_graph, void *stream) {
    return solve3d_impl(slow, nz, nx, ny, dz, dx, dy, src, nsrc, nsweep, max_sets, want_grad,
                        tile_a, group, schedule, tt, sgn, vzero, sets_done, workspace,
                        workspace_bytes, use_graph, stream, nullptr, 0, nullptr);
}

// ---------------------------------------------------------------------------------------
// Solve without allocation: the workspace is a file-scope device array
// ---------------------------------------------------------------------------------------

namespace {

// Each device that runs the library holds one copy of this array until the process ends.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per device.  Host code uses these addresses, never the
// array's name: on the host the name is a stand-in, not device memory.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out) {
    int dev = 0;
    FTEIK_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return FTEIK_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        FTEIK_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return FTEIK_ERR_WORKSPACE;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" size_t fteik3d_scratch_capacity(void) { return kScratchBytes; }

extern "C" int fteik3d_solve_noalloc(const double *slow, int nz, int nx, int ny, double dz,
                                     double dx, double dy, const double *src, int nsrc,
                                     int nsweep, int max_sets, int want_grad, int tile_a,
                                     int group, double *tt, uint8_t *sgn, double *vzero,
                                     int *sets_done, void *stream) {
    // argument and capacity checks come before any HIP call
    if (nz < 1 || nx < 1 || ny < 1 || nsrc < 1 || nsrc > 65535 || !slow || !src || !tt ||
        !vzero || !sets_done || (want_grad && !sgn))
        return FTEIK_ERR_ARG;
    if (fteik3d_workspace_size(nz, nx, ny, nsrc) > kScratchBytes) return FTEIK_ERR_WORKSPACE;
    void *ws = nullptr;
    if (int st = scratch_address(&ws)) return st;
    // overlapped schedule launched directly (a captured graph would be instantiated, which
    // allocates); every byte of the workspace the kernels read is written by the solve's
    // own fill, clear and seed kernels or copied from the host first
    return solve3d_impl(slow, nz, nx, ny, dz, dx, dy, src, nsrc, nsweep, max_sets, want_grad,
                        tile_a, group, 2, tt, sgn, vzero, sets_done, ws, kScratchBytes, 0,
                        stream, nullptr, 0, nullptr);
}

extern "C" size_t fteik3d_history_size(long long events, int max_passes) {
    if (events < 0 || max_passes < 1) return 0;
    return hist_bytes(events, max_passes);
}

extern "C" int fteik3d_solve_logged(const double *slow, int nz, int nx, int ny, double dz,
                                    double dx, double dy, const double *src, int nsrc, int nsweep,
                                    int max_sets, int tile_a, int group, double *tt, double *vzero,
                                    int *sets_done, void *workspace, size_t workspace_bytes,
                                    void *history, size_t history_bytes, long long *events,
                                    int use_graph, void *stream) {
    if (!history || nz < 1 || nx < 1 || ny < 1 || nsrc < 1) return FTEIK_ERR_ARG;
    if (nsweep <= 0 && max_sets < 1) return FTEIK_ERR_ARG;
    // logged node indices are 32-bit (0xffffffff marks an unused entry)
    if ((long long)(nz + 1) * (nx + 1) * (ny + 1) * nsrc >= 0xffffffffLL) return FTEIK_ERR_ARG;
    const int max_pass = 8 * (nsweep > 0 ? nsweep : max_sets);
    const long long cap = hist_capacity(history_bytes, max_pass);
    if (cap < 1) return FTEIK_ERR_WORKSPACE;
    const HistView hv = hist_view(history, cap, max_pass);
    return solve3d_impl(slow, nz, nx, ny, dz, dx, dy, src, nsrc, nsweep, max_sets, 0, tile_a,
                        group, 1, tt, nullptr, vzero, sets_done, workspace, workspace_bytes,
                        use_graph, stream, &hv, cap, events);
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

// ---------------------------------------------------------------------------------------
// Batch of independent 3D models (fteik_solve_batch, batch.hip)
// ---------------------------------------------------------------------------------------

#include "batch_internal.hpp"

namespace fteik {

namespace {

template <bool G, int W>
void launch_multi_t(unsigned blocks, hipStream_t st, const SweepArgs *models, const int *cum,
                    const int *toff, const unsigned long long *tasks, int nmodel) {
    hipLaunchKernelGGL((sweep3d_multi_kernel<G, W>), dim3(blocks), dim3(64 * W), 0, st, models,
                       cum, toff, tasks, nmodel);
}

// Device layout of a batch's 3D part: model arguments, bookkeeping arguments, the active
// source count of every model (one int each, read back together), the block offsets and
// task offsets of every (launch, model), every model's overlapped brick list, then each
// model's own solve workspace.  Launches of a set: at most 8 x the longest anti-diagonal
// count (the overlapped schedule never needs more than one pass at a time).
struct Batch3Layout {
    size_t args, book, nlist, cum, toff, tasks, models, total;
    std::vector<size_t> model_off;
    int dmax;
    long long ntasks;
};

// Brick height of a batch's 3D model: its own tile_a, else the default for the batch's
// total 3D source count (every launch carries the bricks of all models, so the work per
// launch follows the total).  Measured on MI355X on the 48-model workload (32 3D models,
// 520 sources): 0.94 s against 1.04 s with each model's own count.
int batch_tile(const fteik_problem *p, int n, const fteik_problem &q) {
    if (q.tile_a > 0) return q.tile_a;
    long long nsrc3 = 0;
    for (int i = 0; i < n; ++i)
        if (p[i].ndim == 3 && p[i].nsrc > 0) nsrc3 += p[i].nsrc;
    return default_tile((int)std::min<long long>(nsrc3, INT_MAX), q.nz + 1);
}

void batch3_layout(const fteik_problem *p, int n, Batch3Layout *L) {
    int n3 = 0, dmax = 0;
    long long ntasks = 0;
    L->model_off.clear();
    size_t off = 0;
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (p[i].ndim == 3) {
            idx.push_back(i);
            ++n3;
            const int NZ = p[i].nz + 1;
            const int tile = std::max(1, std::min(batch_tile(p, n, p[i]), NZ));
            const long long nba = ceil_div(NZ, tile), nbb = ceil_div(p[i].nx + 1, kLanesB),
                            nbc = ceil_div(p[i].ny + 1, kLanesC);
            dmax = std::max(dmax, (int)(nba + nbb + nbc - 2));
            ntasks += 8 * nba * nbb * nbc;
        }
    L->dmax = dmax;
    L->ntasks = ntasks;
    const size_t rows = (size_t)8 * dmax * (n3 + 1);
    L->args = off;
    off += align256(sizeof(SweepArgs) * n3);
    L->book = off;
    off += align256(sizeof(BookArgs) * n3);
    L->nlist = off;
    off += align256(sizeof(int) * n3);
    L->cum = off;
    off += align256(sizeof(int) * rows);
    L->toff = off;
    off += align256(sizeof(int) * rows);
    L->tasks = off;
    off += align256(sizeof(unsigned long long) * (size_t)ntasks);
    L->models = off;
    for (int i : idx) {
        L->model_off.push_back(off);
        off += align256(ws_bytes(p[i].nz, p[i].nx, p[i].ny, p[i].nsrc, nullptr, nullptr));
    }
    L->total = n3 ? off : 0;
}

}  // namespace

struct Batch3 {
    std::vector<SweepArgs> args;
    std::vector<BookArgs> book;
    std::vector<std::vector<SrcParam>> prm;
    std::vector<Ws3> ws;
    std::vector<int> cum, toff;    // [launch][n3 + 1]: first block, first task of each model
    std::vector<unsigned long long> tasks;  // every model's brick list, concatenated
    int launches = 0;
    std::vector<const fteik_problem *> probs;
    Batch3Layout L;
    char *base = nullptr;
    int nsweep = 0, max_sets = 1, group = 2;
    bool grad = false;
};

size_t batch3_bytes(const fteik_problem *probs, int nprob) {
    Batch3Layout L;
    batch3_layout(probs, nprob, &L);
    return L.total;
}

int batch3_prepare(const fteik_problem *P, int n, int nsweep, int max_sets, int want_grad,
                   int group, char *ws, Batch3 **out, int *failed) {
    // one source per workgroup by default in a batch: measured on MI355X on the 48-model
    // workload, 0.91 s against 0.94 s with two (and batch_tile's brick heights)
    if (group == 0) group = 1;
    if (group != 1 && group != 2 && group != 4 && group != 8) return FTEIK_ERR_ARG;
    Batch3 *b = new Batch3;
    batch3_layout(P, n, &b->L);
    b->base = ws;
    b->nsweep = nsweep;
    b->max_sets = max_sets;
    b->group = group;
    b->grad = want_grad != 0;
    int *nlist_dev = (int *)(ws + b->L.nlist);
    int k = 0;
    for (int i = 0; i < n; ++i) {
        const fteik_problem &q = P[i];
        if (q.ndim != 3) continue;
        if (q.nz < 1 || q.nx < 1 || q.ny < 1 || q.nsrc < 1 || !q.slow || !q.src || !q.tt ||
            !q.vzero || !q.sets_done || !spacing_ok3(q.dz, q.dx, q.dy) ||
            (want_grad && !q.sgn) || (long long)q.nx * q.ny > INT_MAX) {
            delete b;
            return FTEIK_ERR_ARG;
        }
        std::vector<SrcParam> prm(q.nsrc);
        for (int s = 0; s < q.nsrc; ++s)
            if (!place_source3(q.src + 3 * s, q.nz, q.nx, q.ny, q.dz, q.dx, q.dy, &prm[s])) {
                if (failed) *failed = i;
                delete b;
                return FTEIK_ERR_SOURCE_OOB;
            }
        SweepArgs a;
        Ws3 w;
        const int tile = batch_tile(P, n, q);
        make_args(q.slow, q.nz, q.nx, q.ny, q.dz, q.dx, q.dy, q.nsrc, tile, group,
                  want_grad ? q.sgn : nullptr, q.tt, q.sets_done, ws + b->L.model_off[k], &a, &w);
        a.nlist = w.nlist = nlist_dev + k;
        if ((long long)a.ntiles * a.gmax >= (1LL << 31)) {
            delete b;
            return FTEIK_ERR_ARG;
        }
        BookArgs bk;
        bk.nsrc = q.nsrc;
        bk.sets_done = q.sets_done;
        bk.changed = w.changed;
        bk.active = w.active;
        bk.list = w.list;
        bk.nlist = w.nlist;
        b->args.push_back(a);
        b->book.push_back(bk);
        b->prm.push_back(std::move(prm));
        b->ws.push_back(w);
        b->probs.push_back(&q);
        ++k;
    }
    const int n3 = k;
    // every model sweeps on its own overlapped schedule; launch l of the batch runs launch l
    // of every model that has one
    std::vector<std::shared_ptr<const sched::OverlapTable>> tab(n3);
    std::vector<long long> base(n3);
    int nl = 0;
    for (int m = 0; m < n3; ++m) {
        const SweepArgs &a = b->args[m];
        tab[m] = overlap_table_for(a.nba, a.nbb, a.nbc);
        if (!tab[m]) {
            delete b;
            return FTEIK_ERR_ARG;
        }
        base[m] = (long long)b->tasks.size();
        b->tasks.insert(b->tasks.end(), tab[m]->task.begin(), tab[m]->task.end());
        nl = std::max(nl, tab[m]->launches());
    }
    if ((long long)b->tasks.size() != b->L.ntasks || (long long)b->tasks.size() >= (1LL << 31) ||
        nl > 8 * b->L.dmax) {
        delete b;
        return FTEIK_ERR_ARG;
    }
    b->launches = nl;
    b->cum.assign((size_t)nl * (n3 + 1), 0);
    b->toff.assign((size_t)nl * (n3 + 1), 0);
    for (int l = 0; l < nl; ++l) {
        long long c = 0;
        for (int m = 0; m < n3; ++m) {
            const size_t i = (size_t)l * (n3 + 1) + m;
            b->cum[i] = (int)c;
            const sched::OverlapTable &t = *tab[m];
            b->toff[i] = (int)(base[m] + (l < t.launches() ? t.first[l] : 0));
            if (l < t.launches()) c += (long long)(t.first[l + 1] - t.first[l]) * b->args[m].gmax;
        }
        if (c >= (1LL << 31)) {
            delete b;
            return FTEIK_ERR_ARG;
        }
        b->cum[(size_t)l * (n3 + 1) + n3] = (int)c;
    }
    *out = b;
    return 0;
}

void batch3_free(Batch3 *b) { delete b; }

int batch3_run(Batch3 *b, int use_graph, hipStream_t stream) {
    const int n3 = (int)b->args.size();
    if (n3 == 0) return 0;
    char *ws = b->base;
    SweepArgs *args_dev = (SweepArgs *)(ws + b->L.args);
    BookArgs *book_dev = (BookArgs *)(ws + b->L.book);
    int *nlist_dev = (int *)(ws + b->L.nlist);
    int *cum_dev = (int *)(ws + b->L.cum);
    int *toff_dev = (int *)(ws + b->L.toff);
    unsigned long long *tasks_dev = (unsigned long long *)(ws + b->L.tasks);
    FTEIK_CHECK(hipMemcpyAsync(args_dev, b->args.data(), sizeof(SweepArgs) * n3,
                               hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemcpyAsync(book_dev, b->book.data(), sizeof(BookArgs) * n3,
                               hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemcpyAsync(cum_dev, b->cum.data(), sizeof(int) * b->cum.size(),
                               hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemcpyAsync(toff_dev, b->toff.data(), sizeof(int) * b->toff.size(),
                               hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemcpyAsync(tasks_dev, b->tasks.data(),
                               sizeof(unsigned long long) * b->tasks.size(),
                               hipMemcpyHostToDevice, stream));
    for (int m = 0; m < n3; ++m)
        FTEIK_CHECK(hipMemcpyAsync(b->ws[m].prm, b->prm[m].data(),
                                   sizeof(SrcParam) * b->prm[m].size(), hipMemcpyHostToDevice,
                                   stream));
    // the host arrays stay alive in *b; finish the copies before the first kernel anyway so
    // that a failed launch leaves nothing pending on them
    FTEIK_CHECK(hipStreamSynchronize(stream));
    for (int m = 0; m < n3; ++m) {
        const fteik_problem &q = *b->probs[m];
        const SweepArgs &a = b->args[m];
        const long long total = a.node_stride * q.nsrc;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, stream, q.tt, total,
                           b->grad ? q.sgn : (uint8_t *)nullptr, a.chg);
        FTEIK_CHECK(hipGetLastError());
        const long long nt = (long long)a.ntiles * q.nsrc;
        hipLaunchKernelGGL(zero_bytes_kernel,
                           dim3((unsigned)std::min<long long>((nt + 255) / 256, 65536)), dim3(256),
                           0, stream, a.tsum, nt);
        FTEIK_CHECK(hipGetLastError());
        if (a.faces) {
            const long long nf = nt * (long long)sizeof(uint64_t);
            hipLaunchKernelGGL(zero_bytes_kernel,
                               dim3((unsigned)std::min<long long>((nf + 255) / 256, 65536)),
                               dim3(256), 0, stream, (uint8_t *)a.faces, nf);
            FTEIK_CHECK(hipGetLastError());
        }
        if (int st = launch_seed3d(a, b->ws[m], q.slow, q.nsrc, q.vzero, q.sets_done, stream))
            return st;
    }

    auto enqueue_set = [&]() -> int {
        for (int l = 0; l < b->launches; ++l) {
            const int blocks = b->cum[(size_t)l * (n3 + 1) + n3];
            if (blocks == 0) continue;
            FTEIK_DISPATCH_W(launch_multi_t, b->grad, b->group, (unsigned)blocks, stream,
                             args_dev, cum_dev + (size_t)l * (n3 + 1),
                             toff_dev + (size_t)l * (n3 + 1), tasks_dev, n3);
        }
        hipLaunchKernelGGL(set_bookkeeping_multi_kernel, dim3(n3), dim3(256), 0, stream,
                           book_dev);
        return hip_status(hipGetLastError());
    };

    Graphs gr;
    if (use_graph && stream) {
        FTEIK_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue_set();
        hipError_t e = hipStreamEndCapture(stream, &gr.graph);
        if (st) return st;
        FTEIK_CHECK(e);
        FTEIK_CHECK(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
    }
    auto run_set = [&]() -> int {
        if (gr.exec) return hip_status(hipGraphLaunch(gr.exec, stream));
        return enqueue_set();
    };
    int status = 0;
    if (b->nsweep > 0) {
        for (int s = 0; s < b->nsweep && !status; ++s) status = run_set();
    } else {
        std::vector<int> nact(n3);
        for (int s = 0; s < b->max_sets && !status; ++s) {
            status = run_set();
            if (status) break;
            hipError_t e = hipMemcpyAsync(nact.data(), nlist_dev, sizeof(int) * n3,
                                          hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            status = hip_status(e);
            bool any = false;
            for (int m = 0; m < n3; ++m) any |= nact[m] > 0;
            if (!any) break;
        }
    }
    return status;
}

}  // namespace fteik
