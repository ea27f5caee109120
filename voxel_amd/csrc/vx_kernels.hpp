// vx_kernels.hpp -- device-side entry points of libvoxel_c, shared by the
// .hip kernel files and the host runtime.  Everything here is plain C++ on the
// host side; kernels are launched only through the functions declared below.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vx {

enum DType : int { DT_F64 = 0, DT_F32 = 1, DT_I64 = 2, DT_I32 = 3, DT_U64 = 4, DT_U32 = 5 };

inline size_t dtype_size(int dt) { return (dt == DT_F64 || dt == DT_I64 || dt == DT_U64) ? 8 : 4; }

// Comparison codes of the reference ISA (bytecode/opcodes.hpp:415-420):
// EQ 0, NE 1, LT 2, LE 3, GT 4, GE 5; CMP_TRUE = no predicate.
enum Cmp : int { CMP_EQ = 0, CMP_NE = 1, CMP_LT = 2, CMP_LE = 3, CMP_GT = 4, CMP_GE = 5, CMP_TRUE = 6 };

// A predicate lowered for the device.
//  kind 0 (range): pass = (lo <= x && x <= hi) XOR inv   (NaN -> inv)
//  kind 1 (table): st = (x<t) | (x==t)<<1 | (x>t)<<2; pass = (table >> st) & 1
// The range form covers every comparison of the VM; the table form is used
// only where NaN has to fail an inequality (the JIT's ordered compares) and
// for the "lane is non-zero and not NaN" rule of VCOUNT.
struct PredDev {
    uint64_t lo, hi;   // bit patterns of T (low bytes used for 4-byte types)
    uint64_t t;        // table threshold bits
    uint32_t inv;
    uint32_t table;
};

// Per-launch aggregate, T-typed fields stored as bit patterns.
//   floats: s/c are the compensated sum pair in T; ints: s = wrapping i64 sum.
//   mn/mx: extreme passing non-NaN values; for floats mn > mx means "none".
struct ScanOut {
    uint64_t s, c;
    uint64_t cnt;
    uint64_t mn, mx;
    uint64_t pad[3];
};

// Tier-0 min/max group fold aggregate (see vx_scan.hip).
struct GroupOut {
    int64_t smin, smax;    // signed order extremes over {acc0} + group values
    uint64_t umin, umax;   // unsigned order extremes
    uint64_t last;         // last group's value (R[part] after the loop)
    uint64_t groups;
    uint64_t pad[2];
};

// Scratch owned by the runtime: per-block partials + a self-resetting ticket.
struct Scratch {
    void* partials;        // >= max_blocks * 64 bytes
    uint32_t* ticket;      // zero between launches
    size_t max_blocks;
};

struct ScanArgs {
    int dtype;             // DT_F64/F32/I64/I32
    const void* a;         // column a (device)
    const void* b;         // column b (device) or null
    uint64_t begin, end;   // row range [begin, end) of a (and b)
    PredDev pa, pb;
    int mode;              // 0: pred a on a; 1: pred a and pred b both on a; 2: pred a on a, pred b on b
    int table_b;           // pb uses the table form (mode 1/2)
    int table_a;           // pa uses the table form
    int blocks;            // 0 = pick
    int unroll;            // 16-byte loads in flight per lane and column: 2, 4, 8 (0 = default)
    int nontemporal;       // 1 = non-temporal loads, -1 = cached loads, 0 = default (non-temporal)
    // tiles / combine: 0 = default (grid-stride tiles, partials combined by a
    // second kernel), 1 = contiguous run of tiles per workgroup + last-
    // workgroup combine, 2 = grid-stride + second kernel, 3 = contiguous +
    // second kernel, 4 = grid-stride + last-workgroup combine
    int layout;
};

// Fused filter + COUNT/SUM/MIN/MAX over one range; result in *out_dev
// (device memory, ScanOut).  Returns hipSuccess or the launch error.
hipError_t launch_scan(const ScanArgs& args, const Scratch& scr, ScanOut* out_dev, hipStream_t stream);
// The same over nullable columns: rows whose validity bit (va for a, vb for
// b; NULL = all valid; words indexed from row 0, args.begin a multiple of 64)
// is 0 are skipped.  Range predicates only (no table form).
// launch_scan with the partials in the file-scope static device buffer of
// vx_scan.hip (max_parts workgroups) instead of runtime scratch, then a
// one-thread kernel writing the result as a voxel_gpu_agg (80 bytes) to
// out_dev.  Allocates nothing; split layouts (0, 2, 3) only.
constexpr size_t kNoAllocMaxParts = 4096;
hipError_t launch_scan_noalloc(const ScanArgs& args, void* out_dev, hipStream_t stream);
hipError_t launch_scan_valid(const ScanArgs& args, const uint64_t* va, const uint64_t* vb, const Scratch& scr,
                             ScanOut* out_dev, hipStream_t stream);

// Tier-0 min/max loop (VRED_MIN/MAX + CMP/JLE|JGE/MOVR fold) over groups of
// 4 eight-byte lanes starting at row off; rows >= count read as 0.
struct GroupArgs {
    int dtype;             // DT_F64 or DT_I64
    const void* a;
    uint64_t count;        // segment Count
    uint64_t off;          // first row of group 0
    uint64_t groups;       // number of groups (total - off) / 4
    PredDev p;             // filter (range or table), CMP_TRUE for none
    int use_table;
    int red;               // 2 = VRED_MIN, 3 = VRED_MAX (0 VSUM, 1 VCOUNT: launch_group_last)
    int fold;              // 2 = keep the smaller (JLE), 3 = keep the larger (JGE)
    uint64_t acc0;         // initial R[acc] bits
    int blocks;
};
hipError_t launch_group_minmax(const GroupArgs& args, const Scratch& scr, GroupOut* out_dev, hipStream_t stream);
// Sequential fold fallback when the values straddle both orders' wrap points:
// one wave computes 64 group values at a time and applies them in row order.
// result_dev receives {acc, last}.
hipError_t launch_group_fold(const GroupArgs& args, uint64_t* result_dev, hipStream_t stream);
// Value of one Tier-0 group (red 0 VSUM, 1 VCOUNT, 2 VRED_MIN, 3 VRED_MAX) at
// row `off` into *result_dev: the R[part] register after the loop.
hipError_t launch_group_last(const GroupArgs& args, uint64_t* result_dev, hipStream_t stream);

int device_cu_count();

}  // namespace vx
