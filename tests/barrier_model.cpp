// CPU model of the grid barrier of the persistent kernels (kernels.hip,
// grid_barrier / persist_global_run): one thread per block, the barrier words
// bar[0] (arrivals) and bar[1] (generation) kept between consecutive
// launches, each block reading the generation at its start.  Checks, over
// many launches with random delays and block start orders:
//   - no block leaves barrier i before every block has arrived at it;
//   - every launch ends (no deadlock) when every block runs;
//   - bar[0] is 0 between launches.
// It also shows the one way the protocol can stall: a block that never
// starts while the others wait (a launch whose blocks are not all resident).
//
// usage: barrier_model <blocks> <launches> <max_barriers> <seed> [missing_block]
// prints "ok" or "deadlock" / "violation ..." and exits 0 / 1.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

namespace {

std::atomic<unsigned> bar[2];

// one block's barrier, as grid_barrier does it (thread 0 of the block)
bool barrier(unsigned& gen, unsigned nblocks, long spin_cap) {
    const unsigned g = gen;
    const unsigned prev = bar[0].fetch_add(1u);
    if (prev == nblocks - 1) {
        bar[0].store(0u);
        bar[1].store(g + 1);
    } else {
        long s = 0;
        while (bar[1].load() == g) {
            if (++s > spin_cap) return false;
            std::this_thread::yield();
        }
    }
    gen++;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    const unsigned nb = (unsigned)atoi(argv[1]);
    const int launches = atoi(argv[2]);
    const int maxbar = atoi(argv[3]);
    std::mt19937 rng((unsigned)atoi(argv[4]));
    const int missing = argc > 5 ? atoi(argv[5]) : -1;
    bar[0] = 0;
    bar[1] = 0;
    for (int l = 0; l < launches; l++) {
        const int nbar = 1 + (int)(rng() % (unsigned)maxbar);
        // arrived[i]: blocks that reached barrier i
        std::vector<std::atomic<unsigned>> arrived(nbar);
        for (auto& a : arrived) a = 0;
        std::atomic<int> bad{0}, stalled{0};
        std::vector<unsigned> seeds(nb);
        for (auto& s : seeds) s = rng();
        std::vector<std::thread> th;
        for (unsigned b = 0; b < nb; b++) {
            if ((int)b == missing) continue;  // this block never becomes resident
            th.emplace_back([&, b]() {
                std::mt19937 r(seeds[b]);
                std::this_thread::sleep_for(std::chrono::microseconds(r() % 200));  // late start
                unsigned gen = bar[1].load();
                for (int i = 0; i < nbar; i++) {
                    if (r() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(r() % 50));
                    arrived[i].fetch_add(1u);
                    if (!barrier(gen, nb, missing >= 0 ? 200000 : 50000000L)) {
                        stalled = 1;
                        return;
                    }
                    if (arrived[i].load() != nb) bad = 1;  // left before all arrived
                }
            });
        }
        for (auto& t : th) t.join();
        if (stalled) {
            printf("deadlock in launch %d\n", l);
            return 1;
        }
        if (bad) {
            printf("violation in launch %d\n", l);
            return 1;
        }
        if (bar[0].load() != 0) {
            printf("violation: arrivals %u after launch %d\n", bar[0].load(), l);
            return 1;
        }
    }
    printf("ok\n");
    return 0;
}
