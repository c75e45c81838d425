// Prints the overlapped sweep schedule's brick list (fteikpy_amd/csrc/schedule.hpp) for an
// nba x nbb x nbc tile grid: first line "launches tasks", then one line per launch with
// its bricks as pass,BA,BB,BC (oriented brick coordinates).
#include "../../fteikpy_amd/csrc/schedule.hpp"

#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
    if (argc != 4) return 2;
    const int nba = std::atoi(argv[1]), nbb = std::atoi(argv[2]), nbc = std::atoi(argv[3]);
    fteik::sched::OverlapTable t;
    if (!fteik::sched::overlap_table(nba, nbb, nbc, &t)) return 3;
    std::printf("%d %zu\n", t.launches(), t.task.size());
    for (int l = 0; l < t.launches(); ++l) {
        for (int e = t.first[l]; e < t.first[l + 1]; ++e) {
            int p, BA, BB, BC;
            fteik::sched::unpack_task(t.task[e], &p, &BA, &BB, &BC);
            std::printf("%s%d,%d,%d,%d", e > t.first[l] ? " " : "", p, BA, BB, BC);
        }
        std::printf("\n");
    }
    return 0;
}
