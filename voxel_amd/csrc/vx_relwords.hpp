// vx_relwords.hpp -- the hash / sort / join words of straight-line programs
// (Engine<T> handlers, exec/engine.hpp:704-842; HashElement/HashU64
// 1051-1071) on the GPU.  The word-by-word register and slot logic is host
// code in vx_capi.cpp; these are its device operations over one engine's
// segments (DT_F64 or DT_I64 rows, 8 bytes each).
#pragma once
#include <stdint.h>

namespace vx {

struct StreamCtx;

// Grow-only device buffers of one engine's hash / sort / join words.
struct RelBufs {
    uint64_t* a1 = nullptr;
    uint64_t* a2 = nullptr;
    uint64_t* b1 = nullptr;
    uint64_t* b2 = nullptr;
    uint32_t* v1 = nullptr;
    uint32_t* v2 = nullptr;
    uint64_t cap = 0;
    unsigned long long* dev = nullptr;       // [0] count, [1] NaN flag, [2] probe result
    unsigned long long* host = nullptr;      // pinned mirror
    ~RelBufs();
    bool ensure(uint64_t n);
};

// One hash table of the engine (HashTables_[slot], engine.hpp:704-721): the
// distinct HashElement values, ascending, and the first row of each.
struct HashTab {
    uint64_t* h = nullptr;
    uint64_t* row = nullptr;
    uint64_t cap = 0;
    ~HashTab();
    bool ensure(uint64_t n);
};

constexpr uint32_t kTopkMax = 4095;          // imm12 bound of SORT_TOPK / SORT_BOTTOMK

// HashElement (engine.hpp:1051-1062): NaN -> 0, else HashU64 of the bits.
uint64_t relw_hash_element(uint64_t bits, int dtype);

// SORT_ASC / SORT_DESC: d[0, n) sorted in place by value (the total order
// of the bits: -0.0 before +0.0 ascending).  False (error set, d untouched)
// when a value is NaN or n >= 2^30.
bool relw_sort(int dtype, void* d, uint64_t n, bool desc, RelBufs& b, StreamCtx* c);

// SORT_TOPK / SORT_BOTTOMK: libstdc++'s std::partial_sort(d, d + k, d + n)
// with greater (desc) / less, step for step, so every row lands where the
// reference's puts it.  1 <= k <= min(n, kTopkMax).
bool relw_partial_sort(int dtype, void* d, uint64_t n, uint32_t k, bool desc, StreamCtx* c);

// JOIN_HASH / JOIN_NESTED (kind 0: equal pairs), JOIN_MERGE (1: sum of
// min(left run, right run) per value; NaN refused), JOIN_SEMI (2) /
// JOIN_ANTI (3): left rows with / without an equal right row.  Equality is
// C++ ==: NaN never equal, -0.0 == +0.0.
bool relw_join_count(int kind, int dtype, const void* l, uint64_t nl, const void* r, uint64_t nr, uint64_t* out,
                     RelBufs& b, StreamCtx* c);

// HASH_INIT / HASH_BUILD: t := distinct HashElement values of d[0, n) with
// their first rows; *distinct = their number.
bool relw_hash_build(int dtype, const void* d, uint64_t n, HashTab& t, uint64_t* distinct, RelBufs& b,
                     StreamCtx* c);

// HASH_PROBE / HASH_LOOKUP: first row of hash h in t (ntab entries), else
// 2^64 - 1.
bool relw_hash_probe(const HashTab& t, uint64_t ntab, uint64_t h, uint64_t* row, RelBufs& b, StreamCtx* c);

}  // namespace vx
