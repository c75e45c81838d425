// Brick indexing and launch lists of the 3D sweep schedules, shared by the kernels
// (sweep3d.hip) and by the CPU checks of the overlapped launch list
// (tests/schedmodel/overlap_dump.cpp), which compile this header as plain C++.
#pragma once

#include <cstddef>
#include <vector>

#if defined(__HIPCC__)
#define FTEIK_HD __host__ __device__ __forceinline__
#else
#define FTEIK_HD inline
#endif

namespace fteik {
namespace sched {

FTEIK_HD int imin(int a, int b) { return a < b ? a : b; }
FTEIK_HD int imax(int a, int b) { return a > b ? a : b; }

// pass order of the reference (_fteik3d.py:217-391): z flips at pass 4, y at 2, x at 1
FTEIK_HD int pass_sz(int p) { return (p & 4) ? -1 : 1; }
FTEIK_HD int pass_sx(int p) { return (p & 1) ? -1 : 1; }
FTEIK_HD int pass_sy(int p) { return (p & 2) ? -1 : 1; }

// Bricks (BA, BB, BC) with BA + BB + BC = d inside the nba x nbb x nbc tile grid.
FTEIK_HD int bricks_on_diag(int d, int nba, int nbb, int nbc) {
    int total = 0;
    const int amin = imax(0, d - (nbb - 1) - (nbc - 1)), amax = imin(nba - 1, d);
    for (int ba = amin; ba <= amax; ++ba) {
        const int e = d - ba;
        const int bmin = imax(0, e - (nbc - 1)), bmax = imin(nbb - 1, e);
        if (bmax >= bmin) total += bmax - bmin + 1;
    }
    return total;
}

// Brick number t of anti-diagonal d, enumerated with BA outermost (the order of
// bricks_on_diag).
FTEIK_HD void decode_brick(int d, int t, int nba, int nbb, int nbc, int *BA, int *BB, int *BC) {
    const int amin = imax(0, d - (nbb - 1) - (nbc - 1));
    const int amax = imin(nba - 1, d);
    *BA = amin;
    *BB = 0;
    *BC = 0;
    for (int ba = amin; ba <= amax; ++ba) {
        const int e = d - ba;
        const int bmin = imax(0, e - (nbc - 1)), bmax = imin(nbb - 1, e);
        const int cnt = bmax - bmin + 1;
        if (t < cnt) {
            *BA = ba;
            *BB = bmin + t;
            *BC = e - *BB;
            return;
        }
        t -= cnt;
    }
}

// Bricks are kLanesB x kLanesC node columns wide.
constexpr int kLanesB = 8, kLanesC = 8;

// Upper bounds over every tile_a >= 1 (host sizing of the workspace): tiles of the memory
// grid of an nz x nx x ny cell model, and anti-diagonals of its tile grid.
inline unsigned long long max_tiles(int nz, int nx, int ny) {
    return (unsigned long long)(nz + 1) * ((nx + 1 + kLanesB - 1) / kLanesB) *
           ((ny + 1 + kLanesC - 1) / kLanesC);
}
inline unsigned long long max_diag(int nz, int nx, int ny) {
    return (unsigned long long)(nz + 1) + (nx + 1 + kLanesB - 1) / kLanesB +
           (ny + 1 + kLanesC - 1) / kLanesC;
}

// Memory tile of oriented brick (BA, BB, BC) in pass p.
FTEIK_HD void brick_tile(int p, int BA, int BB, int BC, int nba, int nbb, int nbc, int *ti,
                         int *tj, int *tk) {
    *ti = pass_sz(p) > 0 ? BA : nba - 1 - BA;
    *tj = pass_sx(p) > 0 ? BB : nbb - 1 - BB;
    *tk = pass_sy(p) > 0 ? BC : nbc - 1 - BC;
}

// Grid-wide barrier of the persistent schedule: arrivals climb a tree of counters, kBarFan
// participants per counter (workgroups at level 0, counters of the level below above it),
// each counter on its own 128-byte line after the generation line 0, so that no counter
// takes more than kBarFan atomic adds per barrier.  The last arrival at a counter resets it
// and climbs; the last at the root publishes the barrier's generation, which the others
// poll.
constexpr int kBarFan = 16;
constexpr int kBarLine = 32;          // uints per counter line
constexpr int kBarLevels = 5;
constexpr int kBarMaxBlocks = 16384;
constexpr int kBarLines = 1100;       // generation + every level's counters for kBarMaxBlocks
struct BarTree {
    int levels;
    int off[kBarLevels];    // first line of level k
    int parts[kBarLevels];  // participants of level k
};
// Tree for `blocks` (1 .. kBarMaxBlocks) workgroups.
inline BarTree bar_tree(int blocks) {
    BarTree t{};
    int parts = blocks, line = 1, k = 0, n;
    do {
        n = (parts + kBarFan - 1) / kBarFan;
        t.off[k] = line;
        t.parts[k] = parts;
        line += n;
        parts = n;
        ++k;
    } while (n > 1);
    t.levels = k;
    return t;
}
// Counter of participant idx at level k: its line, how many arrive there, and the
// participant index it climbs with.
struct BarCounter {
    int line, members, next;
};
FTEIK_HD BarCounter bar_counter(const BarTree &t, int k, int idx) {
    const int c = idx / kBarFan;
    return {t.off[k] + c, imin(kBarFan, t.parts[k] - c * kBarFan), c};
}

// Overlapped passes: one launch list per sweep set in which every (pass, tile) brick
// runs at the earliest launch where everything it reads is final and nothing it overwrites
// is still to be read.  A brick of pass p on tile T
//  - reads the nodes and change bytes of T and of its 7 upwind tiles in pass p, so those
//    upwind bricks of pass p come in earlier launches;
//  - overwrites T, whose pass p - 1 values T's brick of pass p - 1 and the bricks of pass
//    p - 1 on T's 7 downwind tiles in pass p - 1 read, so those come in earlier launches.
// Then no two bricks of one launch touch a tile that either of them writes, and every brick
// reads exactly what it reads when each pass runs to completion before the next: the
// results are identical, node for node, with fewer launches (the next pass starts at one
// corner while the previous one is still sweeping the far side).  A sweep set ends with
// every brick of pass 7.
struct OverlapTable {
    int nba = 0, nbb = 0, nbc = 0;
    std::vector<unsigned long long> task;  // packed bricks (pack_task), grouped by launch
    std::vector<int> first;                 // launch l runs task[first[l] .. first[l+1])
    int launches() const { return (int)first.size() - 1; }
};

// Oriented brick (BA, BB, BC) of pass p in one word: 3 bits of pass, 20 bits per index.
constexpr int kTaskBits = 20;
inline unsigned long long pack_task(int p, int BA, int BB, int BC) {
    return (unsigned long long)p << (3 * kTaskBits) | (unsigned long long)BA << (2 * kTaskBits) |
           (unsigned long long)BB << kTaskBits | (unsigned long long)BC;
}
FTEIK_HD void unpack_task(unsigned long long v, int *p, int *BA, int *BB, int *BC) {
    const unsigned long long m = (1ull << kTaskBits) - 1;
    *p = (int)(v >> (3 * kTaskBits));
    *BA = (int)((v >> (2 * kTaskBits)) & m);
    *BB = (int)((v >> kTaskBits) & m);
    *BC = (int)(v & m);
}

// Launch of every (pass, tile) brick, level[p * ntiles + memory tile]; tile grids with an
// extent of 2^20 tiles or more are not representable (returns false).  Each pass visits its
// tiles in lexicographic oriented order, in which every upwind tile comes first.
inline bool overlap_levels(int nba, int nbb, int nbc, std::vector<int> *level) {
    const int lim = 1 << kTaskBits;
    if (nba < 1 || nbb < 1 || nbc < 1 || nba >= lim || nbb >= lim || nbc >= lim) return false;
    const long long ntiles = (long long)nba * nbb * nbc;
    level->assign((std::size_t)(8 * ntiles), 0);
    int *L = level->data();
    auto tile = [&](int i, int j, int k) { return ((long long)i * nbb + j) * nbc + k; };
    for (int p = 0; p < 8; ++p) {
        int *lp = L + p * ntiles;
        const int *lq = p > 0 ? L + (p - 1) * ntiles : nullptr;
        const int sz = pass_sz(p), sx = pass_sx(p), sy = pass_sy(p);
        for (int BA = 0; BA < nba; ++BA)
            for (int BB = 0; BB < nbb; ++BB)
                for (int BC = 0; BC < nbc; ++BC) {
                    int ti, tj, tk;
                    brick_tile(p, BA, BB, BC, nba, nbb, nbc, &ti, &tj, &tk);
                    int lv = 0;
                    for (int q = 1; q < 8; ++q) {
                        // upwind tile q of pass p (bit 0: z, 1: x, 2: y), already placed
                        const int i = ti - ((q & 1) ? sz : 0), j = tj - ((q & 2) ? sx : 0),
                                  k = tk - ((q & 4) ? sy : 0);
                        if (i >= 0 && i < nba && j >= 0 && j < nbb && k >= 0 && k < nbc)
                            lv = imax(lv, lp[tile(i, j, k)] + 1);
                    }
                    if (lq) {
                        const int qz = pass_sz(p - 1), qx = pass_sx(p - 1), qy = pass_sy(p - 1);
                        for (int q = 0; q < 8; ++q) {
                            // the tile itself (q = 0) and its downwind tiles of pass p - 1
                            const int i = ti + ((q & 1) ? qz : 0), j = tj + ((q & 2) ? qx : 0),
                                      k = tk + ((q & 4) ? qy : 0);
                            if (i >= 0 && i < nba && j >= 0 && j < nbb && k >= 0 && k < nbc)
                                lv = imax(lv, lq[tile(i, j, k)] + 1);
                        }
                    }
                    lp[tile(ti, tj, tk)] = lv;
                }
    }
    return true;
}

// The launch list of one sweep set: launches in level order; within a launch, bricks by
// pass, then in lexicographic oriented order.
inline bool overlap_table(int nba, int nbb, int nbc, OverlapTable *out) {
    std::vector<int> level;
    if (!overlap_levels(nba, nbb, nbc, &level)) return false;
    const long long ntiles = (long long)nba * nbb * nbc;
    int nl = 0;
    for (int v : level) nl = imax(nl, v + 1);
    out->nba = nba;
    out->nbb = nbb;
    out->nbc = nbc;
    out->first.assign((std::size_t)nl + 1, 0);
    for (int v : level) ++out->first[(std::size_t)v + 1];
    for (int l = 0; l < nl; ++l) out->first[(std::size_t)l + 1] += out->first[(std::size_t)l];
    std::vector<int> fill(out->first.begin(), out->first.end() - 1);
    out->task.assign((std::size_t)(8 * ntiles), 0);
    for (int p = 0; p < 8; ++p)
        for (int BA = 0; BA < nba; ++BA)
            for (int BB = 0; BB < nbb; ++BB)
                for (int BC = 0; BC < nbc; ++BC) {
                    int ti, tj, tk;
                    brick_tile(p, BA, BB, BC, nba, nbb, nbc, &ti, &tj, &tk);
                    const int lv = level[(std::size_t)(p * ntiles + ((long long)ti * nbb + tj) * nbc + tk)];
                    out->task[(std::size_t)fill[lv]++] = pack_task(p, BA, BB, BC);
                }
    return true;
}

}  // namespace sched
}  // namespace fteik
