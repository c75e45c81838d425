// CPU model of the persistent 3D sweep schedule's dependency protocol (sweep3d.hip,
// sweep3d_persistent_kernel), with real threads standing in for resident workgroups.
//
// It runs the kernel's own task decoding and dependency slots (fteikpy_amd/csrc/
// schedule.hpp, compiled here as plain C++) with the kernel's memory orders: tasks come
// from one atomic counter in increasing order; a task waits with acquire loads until every
// dependency's per-(source group, tile) progress counter reaches its target, "runs its
// brick" -- plain reads of every node and tile byte the kernel's brick() reads (its own
// tile, the halo plane and columns of the upwind tiles, the 7 upwind tiles' bytes) and
// plain writes of its own tile -- then fences and publishes pass + 1 with a release store.
// Data are version stamps: a read must see exactly the version the sequential order gives
// (own tile: previous pass finished; upwind tiles: this pass finished, the next not
// started).  Sets are separate launches (threads joined and restarted, counters reset), as
// set_reset_kernel does on the device.
//
// Checked: every task index decodes into the grid (anti-diagonal, brick, group, tile) and
// every (pass, tile, group) is decoded exactly once per set; every progress index stays
// inside the workspace the host allocates (max_tiles x sources); every wait ends (a wait
// longer than the time limit is reported); every read sees the right version.  Built with
// -fsanitize=thread, ThreadSanitizer also reports any read or write of the stamps that the
// acquire/release chain does not order.
//
// Usage: sched_model NZ NX NY tile_a nsrc W threads sets [delay_us [plant]]
//   NZ, NX, NY node counts; W sources per workgroup (G = ceil(nsrc / W) groups); delay_us
//   an optional random per-task delay that reorders completions; plant 1 drops the upwind
//   face waits, 2 the neighbourhood waits (mistakes the checks must catch).
// Exit status 0 when every check passes.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "../../fteikpy_amd/csrc/schedule.hpp"

using namespace fteik::sched;

namespace {

struct Model {
    int NZ, NX, NY, tile_a, G, nthreads, sets, delay_us, plant;
    int nba, nbb, nbc, ntiles, ndiag;
    std::vector<int> cum;                     // cum_nb, ndiag + 1 entries
    std::vector<std::atomic<int>> progress;   // [G][ntiles]
    std::atomic<long long> queue{0};
    std::vector<int> node;                    // [G][nodes] version stamps (plain data)
    std::vector<int> tbyte;                   // [G][ntiles] tile bytes (plain data)
    std::vector<std::atomic<int>> decoded;    // [8][ntiles][G]: times decoded in this set
    std::atomic<long long> bad_reads{0}, bad_decode{0}, timeouts{0}, bad_index{0};
    unsigned long long ws_progress_entries = 0;  // what the host allocates
    int set = 0;

    long long nodes() const { return (long long)NZ * NX * NY; }
    int tile_of(int i, int j, int k) const {
        return ((i / tile_a) * nbb + j / kLanesB) * nbc + k / kLanesC;
    }
};

// brick(): oriented brick origin and the node footprint, as in sweep3d.hip
void run_brick(Model &M, int g, int p, int BA, int BB, int BC, std::mt19937 &rng) {
    const int sz = pass_sz(p), sx = pass_sx(p), sy = pass_sy(p);
    const int a0 = M.tile_a * BA - (sz > 0 ? 0 : M.tile_a * M.nba - M.NZ);
    const int b0 = kLanesB * BB - (sx > 0 ? 0 : kLanesB * M.nbb - M.NX);
    const int c0 = kLanesC * BC - (sy > 0 ? 0 : kLanesC * M.nbc - M.NY);
    int ti, tj, tk;
    brick_tile(p, BA, BB, BC, M.nba, M.nbb, M.nbc, &ti, &tj, &tk);
    const int own = (ti * M.nbb + tj) * M.nbc + tk;
    const int base = M.set * 8;
    auto mi = [&](int a) { return sz > 0 ? a : M.NZ - 1 - a; };
    auto mj = [&](int b) { return sx > 0 ? b : M.NX - 1 - b; };
    auto mk = [&](int c) { return sy > 0 ? c : M.NY - 1 - c; };
    int *nd = M.node.data() + (long long)g * M.nodes();
    int *tb = M.tbyte.data() + (long long)g * M.ntiles;
    // tile bytes: own tile (previous pass finished) and the 7 upwind tiles (this pass)
    for (int q = 0; q < 8; ++q) {
        const int ii = ti - ((q & 1) ? sz : 0), jj = tj - ((q & 2) ? sx : 0),
                  kk = tk - ((q & 4) ? sy : 0);
        if (ii < 0 || ii >= M.nba || jj < 0 || jj >= M.nbb || kk < 0 || kk >= M.nbc) continue;
        const int t = (ii * M.nbb + jj) * M.nbc + kk;
        const int want = base + p + (q == 0 ? 0 : 1);
        if (tb[t] != want) M.bad_reads++;
    }
    // nodes: planes a0-1 .. a0+tile_a-1, columns b0-1 .. b0+7, c0-1 .. c0+7, inside the grid
    for (int a = a0 - 1; a <= a0 + M.tile_a - 1; ++a) {
        if (a < 0 || a > M.NZ - 1) continue;
        for (int b = b0 - 1; b <= b0 + kLanesB - 1; ++b) {
            if (b < 0 || b > M.NX - 1) continue;
            for (int c = c0 - 1; c <= c0 + kLanesC - 1; ++c) {
                if (c < 0 || c > M.NY - 1) continue;
                const int i = mi(a), j = mj(b), k = mk(c);
                const long long n = ((long long)i * M.NX + j) * M.NY + k;
                const int want = base + p + (M.tile_of(i, j, k) == own ? 0 : 1);
                if (nd[n] != want) M.bad_reads++;
            }
        }
    }
    if (M.delay_us) std::this_thread::sleep_for(std::chrono::microseconds(rng() % (M.delay_us + 1)));
    // writes: every node of the own memory tile, and its tile byte
    const int i0 = ti * M.tile_a, j0 = tj * kLanesB, k0 = tk * kLanesC;
    for (int i = i0; i < std::min(M.NZ, i0 + M.tile_a); ++i)
        for (int j = j0; j < std::min(M.NX, j0 + kLanesB); ++j)
            for (int k = k0; k < std::min(M.NY, k0 + kLanesC); ++k)
                nd[((long long)i * M.NX + j) * M.NY + k] = base + p + 1;
    tb[own] = base + p + 1;
}

void workgroup(Model &M, unsigned seed) {
    std::mt19937 rng(seed);
    const long long per_pass = (long long)M.ntiles * M.G;
    const long long ntasks = 8 * per_pass;
    for (;;) {
        const long long t = M.queue.fetch_add(1);
        if (t >= ntasks) break;
        const Task k = decode_task(t, M.G, M.ntiles, M.cum.data(), M.ndiag);
        if (k.p < 0 || k.p > 7 || k.d < 0 || k.d >= M.ndiag || k.g < 0 || k.g >= M.G ||
            k.i < 0 || k.i >= bricks_on_diag(k.d, M.nba, M.nbb, M.nbc)) {
            M.bad_decode++;
            continue;
        }
        int BA, BB, BC, ti, tj, tk;
        decode_brick(k.d, k.i, M.nba, M.nbb, M.nbc, &BA, &BB, &BC);
        if (BA + BB + BC != k.d || BA < 0 || BA >= M.nba || BB < 0 || BB >= M.nbb || BC < 0 ||
            BC >= M.nbc) {
            M.bad_decode++;
            continue;
        }
        brick_tile(k.p, BA, BB, BC, M.nba, M.nbb, M.nbc, &ti, &tj, &tk);
        const int own = (ti * M.nbb + tj) * M.nbc + tk;
        M.decoded[((long long)k.p * M.ntiles + own) * M.G + k.g]++;
        const unsigned long long pbase = (unsigned long long)k.g * M.ntiles;
        if (pbase + M.ntiles > M.ws_progress_entries) M.bad_index++;
        // the waits of lanes 0..29
        const auto start = std::chrono::steady_clock::now();
        for (int slot = 0; slot < 30; ++slot) {
            int need = 0;
            const int idx = task_dependency(slot, k.p, ti, tj, tk, M.nba, M.nbb, M.nbc, &need);
            if (idx < 0) continue;
            if ((M.plant == 1 && slot >= 27) || (M.plant == 2 && slot < 27)) continue;
            if (idx >= M.ntiles) {
                M.bad_index++;
                continue;
            }
            while (M.progress[pbase + idx].load(std::memory_order_acquire) < need) {
                if (std::chrono::steady_clock::now() - start > std::chrono::seconds(10)) {
                    M.timeouts++;
                    break;
                }
                std::this_thread::yield();
            }
        }
        run_brick(M, k.g, k.p, BA, BB, BC, rng);
        std::atomic_thread_fence(std::memory_order_seq_cst);  // __threadfence()
        M.progress[pbase + own].store(k.p + 1, std::memory_order_release);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 9) {
        std::fprintf(stderr, "usage: %s NZ NX NY tile_a nsrc W threads sets [delay_us [plant]]\n",
                     argv[0]);
        return 2;
    }
    Model M;
    M.NZ = std::atoi(argv[1]);
    M.NX = std::atoi(argv[2]);
    M.NY = std::atoi(argv[3]);
    M.tile_a = std::min(std::atoi(argv[4]), M.NZ);
    const int nsrc = std::atoi(argv[5]), W = std::atoi(argv[6]);
    M.nthreads = std::atoi(argv[7]);
    M.sets = std::atoi(argv[8]);
    M.delay_us = argc > 9 ? std::atoi(argv[9]) : 0;
    M.plant = argc > 10 ? std::atoi(argv[10]) : 0;
    M.G = (nsrc + W - 1) / W;
    // the host's geometry (make_args) and cum_nb (solve3d_impl)
    M.nba = (M.NZ + M.tile_a - 1) / M.tile_a;
    M.nbb = (M.NX + kLanesB - 1) / kLanesB;
    M.nbc = (M.NY + kLanesC - 1) / kLanesC;
    M.ntiles = M.nba * M.nbb * M.nbc;
    M.ndiag = M.nba + M.nbb + M.nbc - 2;
    M.cum.assign(M.ndiag + 1, 0);
    for (int d = 0; d < M.ndiag; ++d)
        M.cum[d + 1] = M.cum[d] + bricks_on_diag(d, M.nba, M.nbb, M.nbc);
    if (M.cum[M.ndiag] != M.ntiles) {
        std::printf("FAIL: anti-diagonals hold %d bricks, the grid %d tiles\n", M.cum[M.ndiag],
                    M.ntiles);
        return 1;
    }
    if ((unsigned long long)M.ndiag + 1 > max_diag(M.NZ - 1, M.NX - 1, M.NY - 1) + 1) {
        std::printf("FAIL: cum_nb needs %d entries, the workspace holds %llu\n", M.ndiag + 1,
                    max_diag(M.NZ - 1, M.NX - 1, M.NY - 1) + 1);
        return 1;
    }
    // the workspace's progress array: max_tiles x sources ints (ws_bytes)
    M.ws_progress_entries = max_tiles(M.NZ - 1, M.NX - 1, M.NY - 1) * (unsigned long long)nsrc;
    M.progress = std::vector<std::atomic<int>>((size_t)M.G * M.ntiles);
    M.decoded = std::vector<std::atomic<int>>((size_t)8 * M.ntiles * M.G);
    M.node.assign((size_t)M.G * M.nodes(), 0);
    M.tbyte.assign((size_t)M.G * M.ntiles, 0);
    for (M.set = 0; M.set < M.sets; ++M.set) {
        // set_reset_kernel: queue and progress counters cleared between launches
        M.queue.store(0);
        for (auto &x : M.progress) x.store(0);
        for (auto &x : M.decoded) x.store(0);
        std::vector<std::thread> th;
        for (int w = 0; w < M.nthreads; ++w) th.emplace_back(workgroup, std::ref(M), 1234u + w);
        for (auto &t : th) t.join();
        for (size_t q = 0; q < M.decoded.size(); ++q)
            if (M.decoded[q].load() != 1) M.bad_decode++;
    }
    long long stale = 0;
    for (int v : M.node)
        if (v != 8 * M.sets) ++stale;
    std::printf("grid %dx%dx%d tile_a %d groups %d tiles %d diagonals %d threads %d sets %d: "
                "bad_reads %lld bad_decode %lld bad_index %lld timeouts %lld stale %lld\n",
                M.NZ, M.NX, M.NY, M.tile_a, M.G, M.ntiles, M.ndiag, M.nthreads, M.sets,
                M.bad_reads.load(), M.bad_decode.load(), M.bad_index.load(), M.timeouts.load(),
                stale);
    const bool ok = !M.bad_reads && !M.bad_decode && !M.bad_index && !M.timeouts && !stale;
    std::printf(ok ? "OK\n" : "FAIL\n");
    return ok ? 0 : 1;
}
