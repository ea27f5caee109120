// vx_hostops.cpp -- the reference C ABI's host-side utilities
// (bindings/voxel_c.h:46-65 instruction encoders, 70-76 NullBitmap,
// 97-101 DictionaryEncoding, 106-109 ThreadPool, 114 last error).  None of
// these is on the scan/sort/join/group-by data path; they exist so that Go,
// Java and C callers of the 64-symbol ABI link and behave unchanged.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_runtime.hpp"

namespace {

// Instruction::Encode, bytecode/instruction.hpp:64-71.
uint32_t enc(uint8_t op, uint8_t rd, uint8_t ra, uint8_t rb, uint16_t imm) {
    return (uint32_t)op | ((uint32_t)(rd & 0xF) << 8) | ((uint32_t)(ra & 0xF) << 12) |
           ((uint32_t)(rb & 0xF) << 16) | ((uint32_t)(imm & 0xFFF) << 20);
}

// voxel::NullBitmap (data/nulls.hpp:24-115): bit 1 = valid, all valid at
// creation, tail bits past RowCount cleared.
struct NullBits {
    std::vector<uint64_t> w;
    size_t rows = 0;
};

// DictionaryEncoding<f64>::Build (data/encoding.hpp:113-150): distinct values
// in first-seen order, then sorted ascending; indices remapped.  Values that
// compare equal share one entry (first one seen kept); every NaN is its own
// entry.
struct Dict {
    std::vector<double> dict;
    std::vector<uint32_t> idx;
};

// Fixed-size worker pool (util/thread.hpp:22): tasks are only enqueued from
// C++, so through this ABI it is created, waited on and destroyed.
struct Pool {
    std::vector<std::thread> workers;
    std::queue<std::function<void()>> tasks;
    std::mutex m;
    std::condition_variable cv, done;
    size_t active = 0;
    bool stop = false;
    explicit Pool(uint32_t n) {
        if (n == 0) n = std::thread::hardware_concurrency();
        if (n == 0) n = 1;
        for (uint32_t i = 0; i < n; ++i)
            workers.emplace_back([this] {
                for (;;) {
                    std::function<void()> t;
                    {
                        std::unique_lock<std::mutex> lk(m);
                        cv.wait(lk, [this] { return stop || !tasks.empty(); });
                        if (stop && tasks.empty()) return;
                        t = std::move(tasks.front());
                        tasks.pop();
                        ++active;
                    }
                    t();
                    {
                        std::lock_guard<std::mutex> lk(m);
                        --active;
                        if (tasks.empty() && active == 0) done.notify_all();
                    }
                }
            });
    }
    void wait_all() {
        std::unique_lock<std::mutex> lk(m);
        done.wait(lk, [this] { return tasks.empty() && active == 0; });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : workers) t.join();
    }
};

}  // namespace

extern "C" {

// ---- Instruction encoding: bindings/voxel_c.cpp:106-146 ----
uint32_t voxel_instr_vload(uint8_t vd, uint8_t ra, uint8_t seg_id, uint8_t count) {
    return enc(0x70, vd, ra, 0, (uint16_t)(((seg_id & 0xF) << 8) | count));
}
uint32_t voxel_instr_vfilter_gt(uint8_t vd, uint8_t va, uint8_t rb) { return enc(0xC5, vd, va, rb, 0); }
uint32_t voxel_instr_vfilter_ge(uint8_t vd, uint8_t va, uint8_t rb) { return enc(0xC6, vd, va, rb, 0); }
uint32_t voxel_instr_vfilter_lt(uint8_t vd, uint8_t va, uint8_t rb) { return enc(0xC3, vd, va, rb, 0); }
uint32_t voxel_instr_vfilter_le(uint8_t vd, uint8_t va, uint8_t rb) { return enc(0xC4, vd, va, rb, 0); }
uint32_t voxel_instr_vsum(uint8_t rd, uint8_t va) { return enc(0xD0, rd, va, 0, 0); }
uint32_t voxel_instr_vred_min(uint8_t rd, uint8_t va) { return enc(0xD5, rd, va, 0, 0); }
uint32_t voxel_instr_vred_max(uint8_t rd, uint8_t va) { return enc(0xD6, rd, va, 0, 0); }
uint32_t voxel_instr_vcount(uint8_t rd, uint8_t va) { return enc(0xD7, rd, va, 0, 0); }
uint32_t voxel_instr_addf(uint8_t rd, uint8_t ra, uint8_t rb) { return enc(0x2A, rd, ra, rb, 0); }
uint32_t voxel_instr_add(uint8_t rd, uint8_t ra, uint8_t rb) { return enc(0x20, rd, ra, rb, 0); }
uint32_t voxel_instr_sub(uint8_t rd, uint8_t ra, uint8_t rb) { return enc(0x21, rd, ra, rb, 0); }
uint32_t voxel_instr_mul(uint8_t rd, uint8_t ra, uint8_t rb) { return enc(0x22, rd, ra, rb, 0); }
uint32_t voxel_instr_mov(uint8_t rd, int16_t imm) { return enc(0x10, rd, 0, 0, (uint16_t)imm); }
uint32_t voxel_instr_movr(uint8_t rd, uint8_t ra) { return enc(0x11, rd, ra, 0, 0); }
uint32_t voxel_instr_cmp(uint8_t ra, uint8_t rb) { return enc(0x40, 0, ra, rb, 0); }
uint32_t voxel_instr_jnz(int16_t off) { return enc(0x52, 0, 0, 0, (uint16_t)off); }
uint32_t voxel_instr_jz(int16_t off) { return enc(0x51, 0, 0, 0, (uint16_t)off); }
uint32_t voxel_instr_jmp(int16_t off) { return enc(0x50, 0, 0, 0, (uint16_t)off); }
uint32_t voxel_instr_halt(void) { return enc(0x01, 0, 0, 0, 0); }

// ---- NullBitmap: bindings/voxel_c.cpp:151-162 ----
void* voxel_null_create(size_t count) {
    auto* b = new NullBits();
    b->rows = count;
    b->w.assign((count + 63) / 64, ~0ull);
    if (count % 64) b->w.back() &= (1ull << (count % 64)) - 1;
    return b;
}
void voxel_null_destroy(void* p) { delete static_cast<NullBits*>(p); }
int voxel_null_is_null(const void* p, size_t row) {
    auto* b = static_cast<const NullBits*>(p);
    return ((b->w[row >> 6] >> (row & 63)) & 1ull) == 0 ? 1 : 0;
}
void voxel_null_set_null(void* p, size_t row) { static_cast<NullBits*>(p)->w[row >> 6] &= ~(1ull << (row & 63)); }
void voxel_null_set_valid(void* p, size_t row) { static_cast<NullBits*>(p)->w[row >> 6] |= 1ull << (row & 63); }
size_t voxel_null_null_count(const void* p) {
    auto* b = static_cast<const NullBits*>(p);
    size_t n = 0;
    for (uint64_t x : b->w) n += (size_t)__builtin_popcountll(~x);
    if (b->rows % 64 && !b->w.empty()) {
        uint64_t tail = ~b->w.back() & ~((1ull << (b->rows % 64)) - 1);
        n -= (size_t)__builtin_popcountll(tail);
    }
    return n;
}
size_t voxel_null_valid_count(const void* p) {
    return static_cast<const NullBits*>(p)->rows - voxel_null_null_count(p);
}

// ---- DictionaryEncoding<f64>: bindings/voxel_c.cpp:206-213 ----
void* voxel_dict_create(void) { return new Dict(); }
void voxel_dict_destroy(void* p) { delete static_cast<Dict*>(p); }
void voxel_dict_build(void* p, const double* data, size_t count) {
    Dict& d = *static_cast<Dict*>(p);
    d.dict.clear();
    d.idx.clear();
    if (count == 0) return;
    // group rows holding equal values; each group's entry is its first row
    std::vector<std::pair<double, size_t>> sorted;    // (value, row)
    sorted.reserve(count);
    for (size_t i = 0; i < count; ++i) sorted.emplace_back(data[i], i);
    // rows with equal non-NaN values get the id of their first row
    std::stable_sort(sorted.begin(), sorted.end(), [](const auto& a, const auto& b) {
        bool an = a.first != a.first, bn = b.first != b.first;
        if (an || bn) return !an && bn;          // NaNs last
        return a.first < b.first;
    });
    std::vector<size_t> first_row(count);
    for (size_t k = 0; k < count;) {
        size_t j = k;
        bool nan = sorted[k].first != sorted[k].first;
        if (nan) { first_row[sorted[k].second] = sorted[k].second; ++k; continue; }
        while (j < count && sorted[j].first == sorted[k].first) ++j;
        size_t fr = sorted[k].second;                 // stable: smallest row first
        for (size_t t = k; t < j; ++t) first_row[sorted[t].second] = fr;
        k = j;
    }
    // dictionary entries: first-row values; sort ascending by value
    std::vector<size_t> reps;
    for (size_t i = 0; i < count; ++i) if (first_row[i] == i) reps.push_back(i);
    std::stable_sort(reps.begin(), reps.end(), [&](size_t a, size_t b) { return data[a] < data[b]; });
    std::vector<uint32_t> rep_index(count, 0);
    d.dict.resize(reps.size());
    for (size_t k = 0; k < reps.size(); ++k) { d.dict[k] = data[reps[k]]; rep_index[reps[k]] = (uint32_t)k; }
    d.idx.resize(count);
    for (size_t i = 0; i < count; ++i) d.idx[i] = rep_index[first_row[i]];
}
size_t voxel_dict_cardinality(const void* p) { return static_cast<const Dict*>(p)->dict.size(); }
void voxel_dict_decode_batch(void* p, double* out, size_t offset, size_t count) {
    Dict& d = *static_cast<Dict*>(p);
    for (size_t i = 0; i < count && offset + i < d.idx.size(); ++i) out[i] = d.dict[d.idx[offset + i]];
}

// ---- ThreadPool: bindings/voxel_c.cpp:218-222 ----
void* voxel_pool_create(uint32_t threads) { return new Pool(threads); }
void voxel_pool_destroy(void* p) { delete static_cast<Pool*>(p); }
void voxel_pool_wait_all(void* p) { static_cast<Pool*>(p)->wait_all(); }
size_t voxel_pool_thread_count(const void* p) { return static_cast<const Pool*>(p)->workers.size(); }

// ---- Error: bindings/voxel_c.cpp:227 ----
const char* voxel_last_error(void) { return vx::last_error(); }

}  // extern "C"
