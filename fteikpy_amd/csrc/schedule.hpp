// Brick and task indexing of the 3D sweep schedules, shared by the kernels (sweep3d.hip) and
// by the CPU model of the persistent schedule's dependency protocol
// (tests/schedmodel/sched_model.cpp), which compiles this header as plain C++.
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

// Persistent schedule: task t of a set is pass p = t / (ntiles G), then anti-diagonal d,
// brick i on it, source group g (G groups).  cum_nb[d] = bricks on anti-diagonals < d
// (ndiag + 1 entries).
struct Task {
    int p, d, i, g;
};
FTEIK_HD Task decode_task(long long t, int G, int ntiles, const int *cum_nb, int ndiag) {
    Task k;
    const long long per_pass = (long long)ntiles * G;
    k.p = (int)(t / per_pass);
    const long long r = t - k.p * per_pass;
    // anti-diagonal d: cum_nb[d] * G <= r < cum_nb[d+1] * G
    const int rq = (int)(r / G);
    int lo = 0, hi = ndiag - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cum_nb[mid] <= rq) lo = mid;
        else hi = mid - 1;
    }
    k.d = lo;
    const long long j = r - (long long)cum_nb[k.d] * G;
    k.i = (int)(j / G);
    k.g = (int)(j - (long long)k.i * G);
    return k;
}

// Dependency `slot` (0..29) of a task in pass p on memory tile (ti, tj, tk): slots 0..26
// are the 3x3x3 neighbourhood, which must have finished pass p - 1 (progress >= p);
// slots 27..29 the upwind face tiles along z, x, y, which must have finished pass p
// (progress >= p + 1).  Returns the tile index (ii * nbb + jj) * nbc + kk and sets *need,
// or -1 when the slot has no dependency (outside the grid, or p = 0 for the
// neighbourhood).
FTEIK_HD int task_dependency(int slot, int p, int ti, int tj, int tk, int nba, int nbb, int nbc,
                             int *need) {
    int ii, jj, kk;
    *need = 0;
    if (slot < 27) {
        if (p == 0) return -1;
        ii = ti + slot % 3 - 1;
        jj = tj + (slot / 3) % 3 - 1;
        kk = tk + slot / 9 - 1;
        *need = p;
    } else if (slot < 30) {
        const int ax = slot - 27;
        ii = ti - (ax == 0 ? pass_sz(p) : 0);
        jj = tj - (ax == 1 ? pass_sx(p) : 0);
        kk = tk - (ax == 2 ? pass_sy(p) : 0);
        *need = p + 1;
    } else {
        return -1;
    }
    if (ii < 0 || ii >= nba || jj < 0 || jj >= nbb || kk < 0 || kk >= nbc) return -1;
    return (ii * nbb + jj) * nbc + kk;
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
