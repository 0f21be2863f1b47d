// Concurrency stress of HandleTable, built with ThreadSanitizer
// (tests/host/Makefile) and run by tests/test_host_units.py on the CPU.
// Worker threads look handlers up and use them (as Vship_ComputeSSIMU2 does)
// while other threads free ids and insert new handlers (Vship_SSIMU2Free /
// Init). Every object checks on use that it has not been destroyed; a
// use-after-free or a data race fails the run (TSan exits non-zero).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

#include "../../vship_amd/csrc/handle_table.hpp"

namespace {

constexpr unsigned kMagic = 0x5eed1234u;

struct Handler {
    std::atomic<unsigned> magic{kMagic};
    std::atomic<long> uses{0};
    ~Handler() { magic.store(0); }
    void use()
    {
        if (magic.load() != kMagic) {
            std::fprintf(stderr, "handler used after destruction\n");
            std::abort();
        }
        uses.fetch_add(1);
    }
};

}  // namespace

int main()
{
    vship::HandleTable<Handler> table;
    constexpr unsigned kIds = 8;
    for (unsigned i = 0; i < kIds; i++) table.insert(std::make_shared<Handler>());
    std::atomic<bool> stop{false};
    std::atomic<long> calls{0}, bad{0};
    std::vector<std::thread> threads;
    for (int w = 0; w < 4; w++)
        threads.emplace_back([&, w] {
            unsigned id = (unsigned)w;
            while (!stop.load()) {
                std::shared_ptr<Handler> h;
                if (table.lookup(id % kIds, &h) && h) {
                    for (int k = 0; k < 16; k++) h->use();  // the "call"
                    calls.fetch_add(1);
                } else {
                    bad.fetch_add(1);  // BadHandler: freed in between
                }
                id += 3;
            }
        });
    for (int f = 0; f < 2; f++)
        threads.emplace_back([&, f] {
            for (int it = 0; it < 20000; it++) {
                const unsigned id = (unsigned)(it * 5 + f) % kIds;
                std::shared_ptr<Handler> out;
                if (table.remove(id, &out)) {
                    out.reset();  // Free: drops the table's reference
                    table.insert(std::make_shared<Handler>());
                }
            }
        });
    threads[4].join();
    threads[5].join();
    stop.store(true);
    for (int w = 0; w < 4; w++) threads[w].join();
    std::printf("handle table stress: ok (%ld calls, %ld stale lookups)\n", calls.load(), bad.load());
    return 0;
}
