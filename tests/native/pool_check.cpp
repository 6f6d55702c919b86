// Drives the runtime's device-memory arena (sister_amd/csrc/pool.h) with host
// memory: random get/put sequences, checking alignment, disjointness, the
// budget, full coalescing of returned blocks and trim().  Prints "ok".
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <utility>
#include "../../sister_amd/csrc/pool.h"

using sister::Block;
using sister::Pool;

static int n_chunks_live = 0;
static int host_alloc(char** p, size_t n)
{
    *p = (char*)std::aligned_alloc(4096, (n + 4095) / 4096 * 4096);
    if (!*p) return -1;
    n_chunks_live++;
    return 0;
}
static void host_free(char* p)
{
    std::free(p);
    n_chunks_live--;
}

#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s line %d\n", #c, __LINE__); return 1; } } while (0)

int main()
{
    Pool pool;
    pool.dev_alloc = host_alloc;
    pool.dev_free = host_free;
    pool.chunk = 1 << 20;
    pool.budget = 8 << 20;
    std::mt19937 rng(7);
    std::vector<Block> live;
    size_t live_bytes = 0;
    for (int step = 0; step < 20000; step++) {
        const bool take = live.empty() || (rng() % 100) < 55;
        if (take) {
            const size_t n = 1 + rng() % (600 << 10);
            Block b{nullptr, 0};
            const int e = pool.get(n, b);
            if (e != 0) {
                // refused only when no free block fits and no new chunk fits the budget
                const size_t a = sister::align_up(n);
                CHECK(pool.by_size.lower_bound(a) == pool.by_size.end());
                CHECK(pool.held + a > pool.budget);
                continue;
            }
            CHECK(b.bytes >= n && b.bytes % sister::kAlign == 0);
            CHECK(((uintptr_t)b.ptr % sister::kAlign) == 0);
            for (auto& o : live) CHECK(b.ptr + b.bytes <= o.ptr || o.ptr + o.bytes <= b.ptr);
            // inside one chunk
            const int ch = pool.chunk_of(b.ptr);
            CHECK(ch >= 0 && b.ptr + b.bytes <= pool.chunks[ch].base + pool.chunks[ch].size);
            live.push_back(b);
            live_bytes += b.bytes;
        } else {
            const size_t i = rng() % live.size();
            pool.put(live[i]);
            live_bytes -= live[i].bytes;
            live[i] = live.back();
            live.pop_back();
        }
        CHECK(pool.held <= pool.budget);
        CHECK(pool.in_use == live_bytes);
        size_t free_bytes = 0;
        for (auto& kv : pool.by_addr) free_bytes += kv.second.size;
        CHECK(free_bytes + pool.in_use == pool.held);
        CHECK(pool.by_addr.size() == pool.by_size.size());
    }
    for (auto& b : live) pool.put(b);
    // every chunk is one free block again, and trim returns them all
    CHECK(pool.by_addr.size() == pool.chunks.size());
    pool.trim();
    CHECK(pool.chunks.empty() && pool.held == 0 && n_chunks_live == 0);
    // a block larger than the chunk size gets a chunk of its own; beyond the
    // budget nothing is allocated
    Block big{nullptr, 0}, huge{nullptr, 0};
    CHECK(pool.get(3 << 20, big) == 0 && pool.chunks.size() == 1 && pool.chunks[0].size == (3u << 20));
    CHECK(pool.get(6 << 20, huge) != 0);
    pool.put(big);
    pool.release_all();
    CHECK(n_chunks_live == 0);
    std::printf("ok\n");
    return 0;
}
