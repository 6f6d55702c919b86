// pool.h -- device memory arena of the runtime (host code only).
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <vector>

namespace sister {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

struct Block {
    char* ptr;
    size_t bytes;
};

// Device memory arena.  Memory is taken from the device in chunks (at least
// `chunk` bytes, within the budget) and handed out best-fit; returned blocks coalesce
// with their free neighbours of the same chunk.  Steady-state use (serving,
// repeated batches of mixed sizes) therefore never calls hipMalloc/hipFree,
// which would synchronise the device.  Chunks go back to the device only in
// trim() (entirely free chunks) and release_all().
struct Pool {
    size_t chunk = (size_t)4 << 30;   // minimum size taken from the device at once
    // device allocation hooks (hipMalloc / hipFree in the library; host
    // memory in the allocator's unit test); 0 = success
    int (*dev_alloc)(char** p, size_t bytes) = nullptr;
    void (*dev_free)(char* p) = nullptr;
    size_t budget = (size_t)64 << 30;
    size_t in_use = 0, held = 0, peak = 0;
    struct Chunk { char* base; size_t size; };
    std::vector<Chunk> chunks;
    struct Free { size_t size; int chunk; };
    std::map<char*, Free> by_addr;            // free blocks by address
    std::multimap<size_t, char*> by_size;      // and by size (best fit)

    int chunk_of(char* p) const
    {
        for (size_t i = 0; i < chunks.size(); i++)
            if (p >= chunks[i].base && p < chunks[i].base + chunks[i].size) return (int)i;
        return -1;
    }
    void erase_size(size_t n, char* p)
    {
        auto r = by_size.equal_range(n);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == p) { by_size.erase(it); return; }
    }
    void add_free(char* p, size_t n, int ch)
    {
        auto nx = by_addr.lower_bound(p);
        if (nx != by_addr.end() && nx->second.chunk == ch && p + n == nx->first) {   // merge the next block
            erase_size(nx->second.size, nx->first);
            n += nx->second.size;
            nx = by_addr.erase(nx);
        }
        if (nx != by_addr.begin()) {                                              // merge into the previous
            auto pv = std::prev(nx);
            if (pv->second.chunk == ch && pv->first + pv->second.size == p) {
                erase_size(pv->second.size, pv->first);
                pv->second.size += n;
                by_size.emplace(pv->second.size, pv->first);
                return;
            }
        }
        by_addr.emplace(p, Free{n, ch});
        by_size.emplace(n, p);
    }
    // 0 on success, -1 when the budget or the device cannot provide the block
    int get(size_t bytes, Block& out)
    {
        bytes = align_up(bytes);
        auto it = by_size.lower_bound(bytes);
        if (it == by_size.end()) {
            // requests of half a chunk or more get a chunk of their own size
            // (rounding a 3.8 GB rig workspace up to 4 GiB would leave a
            // remainder per rig that large requests cannot use)
            size_t want = bytes >= chunk / 2 ? bytes : std::max(bytes, chunk);
            if (held + want > budget) want = bytes;
            if (held + want > budget) trim();
            if (held + want > budget) return -1;
            char* p = nullptr;
            int e = dev_alloc(&p, want);
            if (e != 0 && want > bytes) {
                want = bytes;
                e = dev_alloc(&p, want);
            }
            if (e != 0) return -1;
            chunks.push_back({p, want});
            held += want;
            add_free(p, want, (int)chunks.size() - 1);
            it = by_size.lower_bound(bytes);
        }
        char* p = it->second;
        const size_t n = it->first;
        by_size.erase(it);
        auto a = by_addr.find(p);
        const int ch = a->second.chunk;
        by_addr.erase(a);
        if (n > bytes) add_free(p + bytes, n - bytes, ch);
        in_use += bytes;
        peak = std::max(peak, in_use);
        out = {p, bytes};
        return 0;
    }
    void put(const Block& b)
    {
        if (!b.ptr) return;
        in_use -= b.bytes;
        add_free(b.ptr, b.bytes, chunk_of(b.ptr));
    }
    // return entirely free chunks to the device
    void trim()
    {
        for (size_t i = chunks.size(); i-- > 0;) {
            auto a = by_addr.find(chunks[i].base);
            if (a == by_addr.end() || a->second.size != chunks[i].size) continue;
            erase_size(a->second.size, a->first);
            by_addr.erase(a);
            dev_free(chunks[i].base);
            held -= chunks[i].size;
            chunks.erase(chunks.begin() + i);
            // chunk indices above i moved down by one
            for (auto& kv : by_addr)
                if (kv.second.chunk > (int)i) kv.second.chunk--;
        }
    }
    void release_all()
    {
        for (auto& c : chunks) dev_free(c.base);
        chunks.clear();
        by_addr.clear();
        by_size.clear();
        held = in_use = 0;
    }
};


}  // namespace sister
