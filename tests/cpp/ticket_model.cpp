// CPU model of the persistent kernels' ticket protocol (mppi_persistent in
// mppi_metal_amd/csrc/mppi_kernels.hip): host threads stand in for the
// workgroups of one launch and a std::atomic<uint64_t> for the ticket
// counter.  Each "workgroup" loops: t = fetch_add(counter, 1); tile = t mod P
// with P = n_tiles + G; leave when tile >= n_tiles.  Checked over many
// launches of random shapes, with the counter zeroed only when the shape
// changes (as the host does when it rebuilds a batch layout):
//   - every tile of every launch runs exactly once;
//   - after each launch the counter is a multiple of P (the next launch of
//     the same shape starts aligned);
//   - from an arbitrary (misaligned) counter value every workgroup still
//     leaves, after at most n_tiles + G + P tickets in all.
// Exit status 0 when every check holds; the first failure is printed.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

namespace {

struct Launch {
    uint64_t tickets_taken = 0;
    std::vector<uint32_t> runs;   // per tile
};

Launch run_launch(std::atomic<uint64_t>& counter, uint32_t n_tiles, uint32_t G, uint32_t seed) {
    const uint64_t P = uint64_t(n_tiles) + G;
    std::vector<std::atomic<uint32_t>> runs(n_tiles);
    for (auto& r : runs) r.store(0);
    const uint64_t start = counter.load();
    std::vector<std::thread> wgs;
    for (uint32_t g = 0; g < G; ++g)
        wgs.emplace_back([&, g] {
            std::mt19937 rng(seed * 7919u + g);
            for (;;) {
                const uint64_t t = counter.fetch_add(1, std::memory_order_relaxed);
                const uint32_t i = uint32_t(t % P);
                if (i >= n_tiles) break;
                runs[i].fetch_add(1, std::memory_order_relaxed);
                // uneven tile durations, as tiles of agents with different K and H
                if (rng() % 4 == 0) std::this_thread::yield();
            }
        });
    for (auto& w : wgs) w.join();
    Launch l;
    l.tickets_taken = counter.load() - start;
    for (auto& r : runs) l.runs.push_back(r.load());
    return l;
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    std::atomic<uint64_t> counter{0};
    int launches = 0;
    for (int shape = 0; shape < 40; ++shape) {
        const uint32_t n = 1 + rng() % 3000, G = 1 + rng() % 48;
        const uint64_t P = uint64_t(n) + G;
        counter.store(0);                        // the host zeroes on a layout change
        for (int rep = 0; rep < 6; ++rep, ++launches) {
            const Launch l = run_launch(counter, n, G, uint32_t(launches));
            for (uint32_t i = 0; i < n; ++i)
                if (l.runs[i] != 1) {
                    std::printf("FAIL launch %d (n=%u G=%u): tile %u ran %u times\n", launches, n, G, i, l.runs[i]);
                    return 1;
                }
            if (l.tickets_taken != P || counter.load() % P != 0) {
                std::printf("FAIL launch %d (n=%u G=%u): took %llu tickets, P=%llu\n", launches, n, G,
                            (unsigned long long)l.tickets_taken, (unsigned long long)P);
                return 1;
            }
        }
    }
    // Misaligned counters: no workgroup may be left looping.
    for (int trial = 0; trial < 30; ++trial) {
        const uint32_t n = 1 + rng() % 500, G = 1 + rng() % 32;
        const uint64_t P = uint64_t(n) + G;
        counter.store((uint64_t(rng()) << 20) + rng());
        const Launch l = run_launch(counter, n, G, uint32_t(1000 + trial));
        if (l.tickets_taken > uint64_t(n) + G + P) {
            std::printf("FAIL misaligned trial %d (n=%u G=%u): took %llu tickets\n", trial, n, G,
                        (unsigned long long)l.tickets_taken);
            return 1;
        }
    }
    std::printf("ticket model ok: %d aligned launches, 30 misaligned\n", launches);
    return 0;
}
