/*
 * CPU model of the 3D sweep kernel's exact skipping rules -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, at the level of single nodes and tiles, the decisions sweep3d.hip's brick()
 * takes: per-node change bytes (bit p = the node changed in the latest pass p; during the
 * first set bits only accumulate and the 8 seeded corners start with every bit set), per
 * tile one byte (bit p = the tile changed in the latest pass p) and 6 face bytes (OR of the
 * node bytes on the tile's first/last z plane, x column and y column, every node of the
 * face), tiles of `tile_a` x 8 x 8 nodes fixed in memory coordinates, processed per pass
 * in an order where every upwind tile comes first.  A tile is skipped when its own byte and
 * the facing bytes of its 7 upwind neighbour tiles are zero (the z face for the neighbours
 * across z, the x face for the x and xy neighbours, the y face for the y neighbour; the
 * neighbours' whole-tile bytes when tile_a < 8, as the kernel does); a node of a processed
 * tile is skipped when its 7 upwind neighbours' bytes are zero.  Skipping starts in the
 * first set (the kernel's guard holds for the models used here).
 *
 * Every skipped node is re-evaluated with the oracle's node operator (the checker): a skip
 * is sound when that evaluation returns the node's current time.  `plant` = 1 drops the
 * z neighbour (i-sz, j, k) from the node rule, a mistake the count must catch.
 */
#include "../../oracle/fteik_oracle.c"

#define NIDX(i, j, k) (((size_t)(i) * NX + (j)) * NY + (k))

int skip_model3d(const double *slow, int nz, int nx, int ny, double dz, double dx, double dy,
                 double zsrc, double xsrc, double ysrc, int max_sets, int tile_a, int plant,
                 double *tt, int *sets_done, long long *evaluated, long long *violations) {
    double zsa = zsrc / dz, xsa = xsrc / dx, ysa = ysrc / dy;
    if (zsa >= nz) zsa -= EPS;
    if (xsa >= nx) xsa -= EPS;
    if (ysa >= ny) ysa -= EPS;
    const int zsi = imin((int)zsa, nz - 1), xsi = imin((int)xsa, nx - 1),
              ysi = imin((int)ysa, ny - 1);
    const double vzero = slow[((size_t)zsi * nx + xsi) * ny + ysi];
    const int NZ = nz + 1, NX = nx + 1, NY = ny + 1;
    const size_t nn = (size_t)NZ * NX * NY;
    const int nba = (NZ + tile_a - 1) / tile_a, nbb = (NX + 7) / 8, nbc = (NY + 7) / 8;
    const int ntiles = nba * nbb * nbc;
    const int use_faces = tile_a >= 8;
    uint8_t *chg = calloc(nn, 1), *tsum = calloc(ntiles, 1);
    uint64_t *faces = calloc(ntiles, sizeof(uint64_t));
    for (size_t n = 0; n < nn; ++n) tt[n] = BIG;
    static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                     {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (int c = 0; c < 8; ++c) {
        const int i = zsi + corner[c][0], j = xsi + corner[c][1], k = ysi + corner[c][2];
        double t, a, b, e;
        t_anad3(i, j, k, dz, dx, dy, zsa, xsa, ysa, vzero, &t, &a, &b, &e);
        tt[NIDX(i, j, k)] = t;
        chg[NIDX(i, j, k)] = 0xff;
        const int tile = ((i / tile_a) * nbb + j / 8) * nbc + k / 8;
        tsum[tile] = 0xff;
        faces[tile] = ~0ull;
    }
    dargs3_t d;
    d.dz = dz; d.dx = dx; d.dy = dy;
    d.dz2i = 1.0 / dz / dz; d.dx2i = 1.0 / dx / dx; d.dy2i = 1.0 / dy / dy;
    d.dz2dx2 = d.dz2i * d.dx2i; d.dz2dy2 = d.dz2i * d.dy2i; d.dx2dy2 = d.dx2i * d.dy2i;
    d.dsum = d.dz2i + d.dx2i + d.dy2i;
    static const int dirs[8][3] = {{1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {1, -1, -1},
                                   {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    long long evals = 0, bad = 0;
    int set = 0;
    for (; set < max_sets; ++set) {
        int set_changed = 0;
        for (int p = 0; p < 8; ++p) {
            const int sz = dirs[p][0], sx = dirs[p][1], sy = dirs[p][2];
            const unsigned pbit = 1u << p, keep = set == 0 ? 0xffu : ~pbit;
            for (int BA = 0; BA < nba; ++BA)
                for (int BB = 0; BB < nbb; ++BB)
                    for (int BC = 0; BC < nbc; ++BC) {
                        /* oriented tile origin and its memory tile, as brick() computes them */
                        const int a0 = tile_a * BA - (sz > 0 ? 0 : tile_a * nba - NZ);
                        const int b0 = 8 * BB - (sx > 0 ? 0 : 8 * nbb - NX);
                        const int c0 = 8 * BC - (sy > 0 ? 0 : 8 * nbc - NY);
                        const int ti = sz > 0 ? BA : nba - 1 - BA;
                        const int tj = sx > 0 ? BB : nbb - 1 - BB;
                        const int tk = sy > 0 ? BC : nbc - 1 - BC;
                        const int tile = (ti * nbb + tj) * nbc + tk;
                        unsigned any = 0;
                        for (int l = 0; l < 8; ++l) {
                            const int ii = ti - ((l & 1) ? sz : 0), jj = tj - ((l & 2) ? sx : 0),
                                      kk = tk - ((l & 4) ? sy : 0);
                            if (ii < 0 || ii >= nba || jj < 0 || jj >= nbb || kk < 0 || kk >= nbc)
                                continue;
                            const int t2 = (ii * nbb + jj) * nbc + kk;
                            if (l == 0 || !use_faces) {
                                any |= tsum[t2];
                            } else {
                                const int f = (l & 1) ? (sz > 0 ? 1 : 0)
                                                      : ((l & 2) ? (sx > 0 ? 3 : 2) : (sy > 0 ? 5 : 4));
                                any |= (unsigned)(faces[t2] >> (8 * f)) & 0xffu;
                            }
                        }
                        const int skip_tile = any == 0;
                        int tile_changed = 0;
                        for (int a = imax(a0, 1); a <= imin(a0 + tile_a - 1, NZ - 1); ++a)
                            for (int b = imax(b0, 1); b <= imin(b0 + 7, NX - 1); ++b)
                                for (int c = imax(c0, 1); c <= imin(c0 + 7, NY - 1); ++c) {
                                    const int i = sz > 0 ? a : NZ - 1 - a;
                                    const int j = sx > 0 ? b : NX - 1 - b;
                                    const int k = sy > 0 ? c : NY - 1 - c;
                                    const size_t n = NIDX(i, j, k);
                                    unsigned up = chg[NIDX(i, j - sx, k)] | chg[NIDX(i, j, k - sy)] |
                                                  chg[NIDX(i - sz, j - sx, k)] |
                                                  chg[NIDX(i, j - sx, k - sy)] |
                                                  chg[NIDX(i - sz, j, k - sy)] |
                                                  chg[NIDX(i - sz, j - sx, k - sy)];
                                    if (!plant) up |= chg[NIDX(i - sz, j, k)];
                                    int which;
                                    const double t0 = tt[n];
                                    const double tnew =
                                        cand3(tt, slow, &d, i, j, k, sz > 0, sx > 0, sy > 0, sz, sx,
                                              sy, NZ, NX, NY, t0, &which);
                                    int changed = 0;
                                    if (skip_tile || up == 0) {
                                        if (tnew != t0) ++bad;  /* a skip that was wrong */
                                        if (skip_tile) continue;
                                    } else {
                                        ++evals;
                                        tt[n] = tnew;
                                        changed = tnew != t0;
                                    }
                                    chg[n] = (uint8_t)((chg[n] & keep) | (changed ? pbit : 0u));
                                    tile_changed |= changed;
                                }
                        if (skip_tile) continue;
                        set_changed |= tile_changed;
                        tsum[tile] = (uint8_t)((tsum[tile] & keep) | (tile_changed ? pbit : 0u));
                        unsigned f[6] = {0, 0, 0, 0, 0, 0};
                        const int ilo = ti * tile_a, ihi = imin(NZ, ilo + tile_a) - 1;
                        const int jlo = tj * 8, jhi = imin(NX, jlo + 8) - 1;
                        const int klo = tk * 8, khi = imin(NY, klo + 8) - 1;
                        for (int i = ilo; i <= ihi; ++i)
                            for (int j = jlo; j <= jhi; ++j)
                                for (int k = klo; k <= khi; ++k) {
                                    const unsigned v = chg[NIDX(i, j, k)];
                                    if (i == ilo) f[0] |= v;
                                    if (i == ihi) f[1] |= v;
                                    if (j == jlo) f[2] |= v;
                                    if (j == jhi) f[3] |= v;
                                    if (k == klo) f[4] |= v;
                                    if (k == khi) f[5] |= v;
                                }
                        uint64_t packed = 0;
                        for (int q = 0; q < 6; ++q) packed |= (uint64_t)(f[q] & 0xffu) << (8 * q);
                        faces[tile] = packed;
                    }
        }
        if (!set_changed) {
            ++set;
            break;
        }
    }
    free(chg);
    free(tsum);
    free(faces);
    *sets_done = set;
    *evaluated = evals;
    *violations = bad;
    return 0;
}
