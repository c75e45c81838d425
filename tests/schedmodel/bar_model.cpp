// CPU model of the persistent schedule's grid-wide barrier (fteikpy_amd/csrc/schedule.hpp:
// bar_tree, bar_counter; the device side is tree_barrier in sweep3d.hip): one thread per
// workgroup runs `rounds` rounds of write-own-slot, barrier, check-every-slot, barrier, with
// the device's atomic operations and memory orders (the poll uses an acquire load, where
// the device polls relaxed and fences after).
// Usage: bar_model blocks rounds        -> exit 0 and "ok" when every check passed
//        bar_model layout               -> checks the line layout for 1..kBarMaxBlocks
//        bar_model giveup blocks        -> one workgroup never arrives: every other one
//                                          leaves the first barrier with the flag set,
//                                          after the (shortened) wait limit
#include "../../fteikpy_amd/csrc/schedule.hpp"

#include <atomic>
#include <chrono>
#include <string>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

using namespace fteik::sched;

static double g_wait_seconds = 5.0;

// Returns false once any wait gave up (word 1 of line 0 set), like the device barrier.
static bool barrier(const BarTree &t, std::atomic<unsigned> *mem, int blk, unsigned target) {
    std::atomic_thread_fence(std::memory_order_seq_cst);
    int idx = blk;
    bool last = true;
    for (int k = 0; k < t.levels && last; ++k) {
        const BarCounter c = bar_counter(t, k, idx);
        std::atomic<unsigned> &ctr = mem[(size_t)c.line * kBarLine];
        const unsigned old = ctr.fetch_add(1u, std::memory_order_acq_rel);
        last = old == (unsigned)(c.members - 1);
        if (last) ctr.store(0u, std::memory_order_relaxed);
        idx = c.next;
    }
    std::atomic<unsigned> &flag = mem[1];
    if (last) {
        mem[0].store(target, std::memory_order_release);
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        while (mem[0].load(std::memory_order_acquire) != target) {
            if (flag.load(std::memory_order_relaxed) != 0u) break;
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
                g_wait_seconds) {
                flag.store(1u, std::memory_order_relaxed);
                break;
            }
            std::this_thread::yield();
        }
    }
    const bool ok = flag.load(std::memory_order_relaxed) == 0u;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    return ok;
}

int main(int argc, char **argv) {
    if (argc == 2 && std::string(argv[1]) == "layout") {
        for (int b = 1; b <= kBarMaxBlocks; ++b) {
            const BarTree t = bar_tree(b);
            if (t.levels < 1 || t.levels > kBarLevels) return 3;
            int lines = 1;
            for (int k = 0; k < t.levels; ++k) {
                // every participant of level k maps to a counter of that level
                const int n = (t.parts[k] + kBarFan - 1) / kBarFan;
                int members = 0;
                for (int c = 0; c < n; ++c) {
                    const BarCounter bc = bar_counter(t, k, c * kBarFan);
                    if (bc.line != t.off[k] + c || bc.members < 1 || bc.members > kBarFan) return 4;
                    members += bc.members;
                }
                if (members != t.parts[k] || t.off[k] != lines) return 5;
                if (k + 1 < t.levels && t.parts[k + 1] != n) return 6;
                lines += n;
            }
            if (t.parts[t.levels - 1] > kBarFan || lines > kBarLines) return 7;
        }
        std::printf("ok\n");
        return 0;
    }
    if (argc == 3 && std::string(argv[1]) == "giveup") {
        const int blocks = std::atoi(argv[2]);
        if (blocks < 2 || blocks > kBarMaxBlocks) return 2;
        g_wait_seconds = 0.5;
        const BarTree t = bar_tree(blocks);
        std::vector<std::atomic<unsigned>> mem((size_t)kBarLines * kBarLine);
        for (auto &m : mem) m.store(0u);
        std::atomic<int> left{0};
        std::vector<std::thread> th;
        for (int b = 1; b < blocks; ++b)  // workgroup 0 never arrives
            th.emplace_back([&, b] {
                if (!barrier(t, mem.data(), b, 1u)) left.fetch_add(1);
            });
        for (auto &x : th) x.join();
        if (left.load() != blocks - 1 || mem[1].load() != 1u || mem[0].load() != 0u) return 11;
        std::printf("ok\n");
        return 0;
    }
    if (argc != 3) return 2;
    const int blocks = std::atoi(argv[1]), rounds = std::atoi(argv[2]);
    if (blocks < 1 || blocks > kBarMaxBlocks) return 2;
    const BarTree t = bar_tree(blocks);
    std::vector<std::atomic<unsigned>> mem((size_t)kBarLines * kBarLine);
    for (auto &m : mem) m.store(0u);
    mem[0].store(12345u);  // any starting generation
    g_wait_seconds = 60.0;
    std::vector<int> slot(blocks, -1);
    std::atomic<int> bad{0};
    const unsigned base = mem[0].load();
    std::vector<std::thread> th;
    for (int b = 0; b < blocks; ++b)
        th.emplace_back([&, b] {
            unsigned n = 0;
            for (int r = 0; r < rounds; ++r) {
                slot[b] = r;
                if (!barrier(t, mem.data(), b, base + ++n)) bad.fetch_add(1);
                for (int j = 0; j < blocks; ++j)
                    if (slot[j] != r) bad.fetch_add(1);
                if (!barrier(t, mem.data(), b, base + ++n)) bad.fetch_add(1);
            }
        });
    for (auto &x : th) x.join();
    for (size_t i = 1; i < mem.size(); ++i)
        if (mem[i].load() != 0u) return 8;  // every counter back at zero, no give-up
    if (mem[0].load() != base + 2u * (unsigned)rounds) return 9;
    if (bad.load()) return 10;
    std::printf("ok\n");
    return 0;
}
