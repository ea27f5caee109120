// vx_capi.cpp -- the reference C ABI's engine and JIT entry points
// (bindings/voxel_c.h:14-41), plus the voxel_gpu_* scan extensions,
// lowered to the GPU scan kernels of vx_scan.hip.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_kernels.hpp"
#include "vx_lower.hpp"
#include "vx_runtime.hpp"
#include "vx_radix.hpp"
#include "vx_relwords.hpp"
#include "vx_segagg.hpp"

namespace vx {
int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);
}

using namespace vx;

namespace {

// A segment: the reference's borrowed Segment<T>{Data, Count}
// (data/segment.hpp:25-60), mirrored in HBM on demand.
struct Seg {
    const void* host = nullptr;      // borrowed caller memory (null for device segments)
    const void* devext = nullptr;    // caller's device memory
    size_t count = 0;
    bool resident = false;           // copy once, never refresh
    DevColumn col;
};

// Engine<T>: 64 scalar registers (core/registers.hpp:20) + segments + program.
// One AggState slot (engine.hpp:347-372): the fields the aggregates share.
// The f64 sum is kept as a compensated pair; the distinct set as sorted
// order keys in one of the slot's two device buffers (`cur`).  A run works
// on a copy of the slots and writes a changed distinct set into the buffer
// that is not committed, so a program that fails part-way leaves every slot
// as it was.
struct AggSlot {
    uint64_t count = 0;
    double sh = 0.0, sl = 0.0;       // SumF64
    int64_t sumi = 0;                // SumI64
    uint64_t minv = 0, maxv = 0;     // MinVal / MaxVal bits (set from the engine type)
    uint64_t firstv = 0, lastv = 0;
    double mean = 0.0, m2 = 0.0;
    bool init = false, first_set = false, has_set = false;
    uint64_t nset = 0;               // distinct order keys, ascending, in set buffer `cur`
    int cur = 0;
};

struct SetBuf {
    uint64_t* p = nullptr;
    uint64_t cap = 0;
    ~SetBuf() { if (p) (void)hipFree(p); }
};

// One hash-table slot (HashTables_[slot], engine.hpp:704-744): the table in
// device buffer `cur` of the slot's two, `n` distinct hashes.  Like the
// AGG slots, a run works on a copy and builds into the other buffer.
struct TabSlot {
    uint64_t n = 0;
    int cur = 0;
    bool valid = false;
};

struct Engine {
    int dtype = DT_F64;
    uint64_t regs[64] = {};
    std::vector<std::unique_ptr<Seg>> segs;
    std::vector<uint32_t> code;
    Plan plan;
    AggSlot agg[16];                 // committed slot state
    SetBuf sets[16][2];
    TabSlot tab[16];                 // committed hash-table slots
    HashTab tabs[16][2];
    RelBufs rel;
    bool agg_ready = false;
    SegaggBufs segagg;
    DevColumn expr;                  // a reduced vector of a loop with vector arithmetic
    DevColumn wcol;                  // window values of a loop loading a window word
    DevColumn trap;                  // set by an INT64_MIN / -1 lane
    DevColumn stores;                // the stored vectors of a loop's VSTOREs, until the run commits them
    DevColumn fused;                 // fused expression reductions: 4 results, 4 last-group values, trap flag
};

size_t lanes_of(int dt) { return 32 / dtype_size(dt); }

// Device pointer of a segment whose rows [begin, end) are current in HBM.
const void* seg_rows(Seg& s, int dt, uint64_t begin, uint64_t end, StreamCtx* c) {
    if (s.devext) return s.devext;
    size_t es = dtype_size(dt);
    if (!s.col.ensure(s.count * es)) return nullptr;
    if (s.resident && s.col.valid) return s.col.dev;
    if (s.resident) { begin = 0; end = s.count; }
    if (end > begin) {
        char* d = static_cast<char*>(s.col.dev) + begin * es;
        const char* h = static_cast<const char*>(s.host) + begin * es;
        if (!hip_ok(hipMemcpyAsync(d, h, (end - begin) * es, hipMemcpyHostToDevice, c->stream), "copy segment to HBM"))
            return nullptr;
    }
    if (s.resident) s.col.valid = true;
    return s.col.dev;
}

bool sync_read(StreamCtx* c, void* host_dst, const void* dev_src, size_t bytes) {
    if (!hip_ok(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, c->stream), "read back result"))
        return false;
    return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// One scan launch into the stream's result slab, read back.
bool scan_once(StreamCtx* c, const ScanArgs& g, ScanOut* out) {
    if (g.end <= g.begin) {
        std::memset(out, 0, sizeof(*out));
        return true;
    }
    if (!hip_ok(launch_scan(g, c->scr, &c->dev->scan, c->stream), "scan launch")) return false;
    if (!sync_read(c, &c->host->scan, &c->dev->scan, sizeof(ScanOut))) return false;
    *out = c->host->scan;
    return true;
}

double bits_f64(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
uint64_t f64_bits(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }
float bits_f32(uint64_t b) { uint32_t u = (uint32_t)b; float f; std::memcpy(&f, &u, 4); return f; }

// ---------------------------------------------------------------------------
// Engine<T>::RunFastPath (exec/engine.hpp:109-141) on the GPU.
// ---------------------------------------------------------------------------
bool run_fast(Engine& e, StreamCtx* c) {
    const Plan& p = e.plan;
    if (p.seg >= e.segs.size()) { set_error("program reads segment %u, engine has %zu", p.seg, e.segs.size()); return false; }
    Seg& s = *e.segs[p.seg];
    const uint64_t kl = lanes_of(e.dtype);
    uint64_t total = e.regs[p.tot] < s.count ? e.regs[p.tot] : s.count;
    uint64_t off = e.regs[p.off];
    uint64_t groups = (off <= total && total - off >= kl) ? (total - off) / kl : 0;
    uint64_t end = off + groups * kl;
    ScanOut r{};
    if (groups > 0) {
        const void* d = seg_rows(s, e.dtype, off, end, c);
        if (!d) return false;
        ScanArgs g{};
        g.dtype = e.dtype; g.a = d; g.begin = off; g.end = end; g.mode = 0;
        g.pa = make_pred(e.dtype, cmp_of_filter_op(p.fop), e.regs[p.thr]);
        if (!scan_once(c, g, &r)) return false;
    }
    uint64_t accb = e.regs[p.acc];
    if (e.dtype == DT_F64) {
        double acc0 = bits_f64(accb);
        double acc = groups ? finish_sum_f64(acc0, bits_f64(r.s), bits_f64(r.c)) : acc0;
        e.regs[p.acc] = f64_bits(acc);
    } else if (e.dtype == DT_I64) {
        e.regs[p.acc] = accb + r.s;                        // wrapping i64 (engine.hpp:134-138)
    } else {                                              // i32: accumulate in T, sign-extend (registers.hpp:57-62)
        int32_t acc = (int32_t)(uint32_t)((uint32_t)accb + (uint32_t)r.s);
        e.regs[p.acc] = (uint64_t)(int64_t)acc;
    }
    e.regs[p.off] = end;
    return true;
}

// ---------------------------------------------------------------------------
// Tier-0 loop programs (dispatch loop engine.hpp:76-80 over the handlers of
// VLOAD/VFILTER/VSUM|VCOUNT|VRED_*/ADDF|ADD|CMP-select/ADD/CMP/JNZ).  A
// chain of filters zeroes a lane unless every filter passes (VFilterImpl,
// engine.hpp:1090-1110, applied to a lane already zeroed keeps it zero), so
// a reduced vector is "conjunction of its filters ? x : 0"; each reduction
// block is run as its own scan or group pass.
// ---------------------------------------------------------------------------
namespace {
bool run_tier0_block(const Plan::Block& b, const PredDev& pred, int dtype, const void* d, uint64_t count,
                     uint64_t off, uint64_t dend, uint64_t groups, uint64_t* regs, StreamCtx* c) {
    const uint64_t kl = 4;
    const bool f = dtype == DT_F64;
    const bool okpair = (b.red == 0 && b.accop == 0 && f) || (b.red == 0 && b.accop == 1 && !f) ||
                        (b.red == 1 && b.accop == 1) || (b.red >= 2 && b.accop >= 2);
    if (!okpair) { set_error("loop (reduction %u, accumulate %u) on this engine type is not lowered", b.red, b.accop); return false; }
    GroupArgs ga{};
    ga.dtype = dtype; ga.a = d; ga.count = count; ga.off = off; ga.groups = groups; ga.p = pred;
    ga.use_table = 0; ga.red = b.red; ga.fold = b.accop; ga.acc0 = regs[b.acc];

    uint64_t accb = regs[b.acc], last = 0;
    if (b.red <= 1) {
        ScanArgs g{};
        g.dtype = dtype; g.a = d; g.begin = off < dend ? off : dend; g.end = dend;
        g.pa = pred;
        if (b.red == 1) { g.mode = 1; g.pb = make_nonzero_pred(dtype); g.table_b = 1; }
        ScanOut r{};
        if (g.end > g.begin) {
            if (!hip_ok(launch_scan(g, c->scr, &c->dev->scan, c->stream), "scan launch")) return false;
        } else {
            std::memset(&c->host->scan, 0, sizeof(ScanOut));
        }
        ga.off = off + (groups - 1) * kl;
        if (!hip_ok(launch_group_last(ga, &c->dev->last, c->stream), "last-group launch")) return false;
        if (!sync_read(c, c->host, c->dev, sizeof(ResultSlab))) return false;
        if (g.end > g.begin) r = c->host->scan;
        last = c->host->last;
        if (b.red == 0 && f) accb = f64_bits(finish_sum_f64(bits_f64(accb), bits_f64(r.s), bits_f64(r.c)));
        else if (b.red == 0) accb = accb + r.s;
        else accb = accb + r.cnt;
    } else {
        if (!hip_ok(launch_group_minmax(ga, c->scr, &c->dev->group, c->stream), "group launch")) return false;
        if (!sync_read(c, &c->host->group, &c->dev->group, sizeof(GroupOut))) return false;
        GroupOut go = c->host->group;
        last = go.last;
        const uint64_t half = 1ull << 63;
        if ((uint64_t)go.smax - (uint64_t)go.smin < half) {      // unsigned: no signed overflow
            // every value lies in a signed window narrower than 2^63: the
            // wrapping CMP difference is the signed order
            accb = b.accop == 2 ? (uint64_t)go.smin : (uint64_t)go.smax;
        } else if (go.umax - go.umin < half) {
            accb = b.accop == 2 ? go.umin : go.umax;
        } else {
            if (!hip_ok(launch_group_fold(ga, c->dev->fold, c->stream), "fold launch")) return false;
            if (!sync_read(c, c->host->fold, c->dev->fold, 16)) return false;
            accb = c->host->fold[0];
            last = c->host->fold[1];
        }
    }
    regs[b.acc] = accb;
    regs[b.part] = last;
    return true;
}
}  // namespace

// Loops with vector arithmetic: the body's vector words are evaluated per row
// (lanes are independent) into the reduced register's values over every row
// the loop visits, [off, total), rows past Count loading 0; each block then
// reduces that column as an unfiltered loop would.
// The body's vector words over rows [off, off + nrows) of column col (rows at
// or past count load 0), one evaluation per reduced register, each reduced
// as an unfiltered loop would; then the loop's register effects.
// The accumulate word of a block must be one the lowering maps (ADDF for
// f64 sums, ADD for i64 sums and counts, a CMP-select fold for VRED_*).
bool block_pair_ok(const Plan::Block& b, int dtype) {
    const bool f = dtype == DT_F64;
    return (b.red == 0 && b.accop == 0 && f) || (b.red == 0 && b.accop == 1 && !f) || (b.red == 1 && b.accop == 1) ||
           (b.red >= 2 && b.accop >= 2);
}

// Loops that store nothing: the words and the reductions run in one pass
// (expr_reduce), with no evaluated column in HBM; then the loop's register
// effects.  Returns 1 when done, 0 when a VRED_MIN / VRED_MAX fold needs the
// exact sequential fold (its group values span more than 2^63 in both the
// signed and the unsigned order; nothing committed), -1 on error.
int run_expr_fused(Engine& e, uint64_t* regs, const ExprProg& prog, const void* col, uint64_t count, uint64_t off,
                   uint64_t nrows, uint64_t total, StreamCtx* c) {
    const Plan& p = e.plan;
    ExprRed rd{};
    rd.n = (int32_t)p.blocks.size();
    for (size_t q = 0; q < p.blocks.size(); ++q) {
        const Plan::Block& b = p.blocks[q];
        if (!block_pair_ok(b, e.dtype)) {
            set_error("loop (reduction %u, accumulate %u) on this engine type is not lowered", b.red, b.accop);
            return -1;
        }
        rd.vreg[q] = b.vreg;
        rd.red[q] = b.red;
    }
    // layout: ScanOut res[4] | u64 last[4] | int trap
    const size_t last_at = 4 * sizeof(ScanOut), trap_at = last_at + 32, bytes = trap_at + 8;
    if (!e.fused.ensure(bytes)) return -1;
    char* fb = static_cast<char*>(e.fused.dev);
    if (!hip_ok(launch_zero32(reinterpret_cast<uint32_t*>(fb + last_at), (32 + 8) / 4, c->stream), "clear")) return -1;
    ScanOut* res = reinterpret_cast<ScanOut*>(fb);
    if (!expr_reduce(e.dtype, col, count, off, nrows, prog, rd, res, reinterpret_cast<uint64_t*>(fb + last_at),
                     reinterpret_cast<int*>(fb + trap_at), c))
        return -1;
    struct {
        ScanOut res[4];
        uint64_t last[4];
        int trap, pad;
    } h{};
    static_assert(sizeof(h) == 4 * sizeof(ScanOut) + 40, "fused result layout");
    if (!sync_read(c, &h, fb, bytes)) return -1;
    if (h.trap) {
        set_error("a vector division or modulo of INT64_MIN by -1 (the reference's x86 build traps there)");
        return -1;
    }
    uint64_t accs[4];
    for (size_t q = 0; q < p.blocks.size(); ++q) {              // the folds first: a fallback commits nothing
        const Plan::Block& b = p.blocks[q];
        if (b.red < 2) continue;
        const ScanOut& r = h.res[q];
        const uint64_t a0 = regs[b.acc], half = 1ull << 63;
        int64_t smin = (int64_t)r.s, smax = (int64_t)r.c;
        uint64_t umin = r.cnt, umax = r.mn;
        smin = (int64_t)a0 < smin ? (int64_t)a0 : smin;          // the accumulator's own value takes part
        smax = (int64_t)a0 > smax ? (int64_t)a0 : smax;
        umin = a0 < umin ? a0 : umin;
        umax = a0 > umax ? a0 : umax;
        // every value in a window narrower than 2^63: the wrapping CMP
        // difference is that order, so the fold's result is its extreme
        if ((uint64_t)smax - (uint64_t)smin < half) accs[q] = b.accop == 2 ? (uint64_t)smin : (uint64_t)smax;
        else if (umax - umin < half) accs[q] = b.accop == 2 ? umin : umax;
        else return 0;
    }
    for (size_t q = 0; q < p.blocks.size(); ++q) {
        const Plan::Block& b = p.blocks[q];
        const ScanOut& r = h.res[q];
        uint64_t accb = regs[b.acc];
        if (b.red >= 2) {
            accb = accs[q];
        } else if (b.red == 0 && e.dtype == DT_F64) {
            accb = f64_bits(finish_sum_f64(bits_f64(accb), bits_f64(r.s), bits_f64(r.c)));
        } else if (b.red == 0) {
            accb = accb + r.s;
        } else {
            accb = accb + r.cnt;
        }
        regs[b.acc] = accb;
        regs[b.part] = h.last[q];
    }
    regs[p.off] = total;
    if (run_scalar(p.epilogue, regs) < 0) return -1;
    std::memcpy(e.regs, regs, sizeof(uint64_t) * 64);
    return 1;
}

bool run_expr_blocks(Engine& e, uint64_t* regs, const void* col, uint64_t count, uint64_t off, uint64_t nrows,
                     uint64_t total, StreamCtx* c) {
    const Plan& p = e.plan;
    const uint64_t kl = 4;
    const bool fusable = p.stores.empty() && !p.blocks.empty() && p.blocks.size() <= 4;
    if (fusable) {
        ExprProg prog{};
        prog.nops = (int32_t)p.vops.size();
        prog.vload = p.vload_dst;
        for (size_t k = 0; k < p.vops.size(); ++k) {
            const Plan::VOp& v = p.vops[k];
            prog.ops[k] = ExprOp{v.op, v.dst, v.a, v.b, (int16_t)v.cmp, v.imm, v.sreg != 0xFF ? regs[v.sreg] : 0};
        }
        for (int q = 0; q < 2; ++q) {                                 // VGATHER's segments, whole
            if (p.gsegs[q] == 0xFF || p.gsegs[q] >= e.segs.size()) continue;
            Seg& g = *e.segs[p.gsegs[q]];
            prog.gcount[q] = g.count;
            if (g.count && !(prog.gcol[q] = seg_rows(g, e.dtype, 0, g.count, c))) return false;
        }
        const int r = run_expr_fused(e, regs, prog, col, count, off, nrows, total, c);
        if (r != 0) return r > 0;
        // a VRED_* fold over values spanning both orders: the column path's exact fold
    }
    if (nrows > (1ull << 31)) { set_error("a loop with vector arithmetic over %llu rows (more than 2^31) is not lowered",
                                          (unsigned long long)nrows); return false; }
    if (!e.expr.ensure(nrows * 8)) return false;
    ExprProg prog{};
    prog.nops = (int32_t)p.vops.size();
    prog.vload = p.vload_dst;
    for (size_t k = 0; k < p.vops.size(); ++k) {
        const Plan::VOp& v = p.vops[k];
        prog.ops[k] = ExprOp{v.op, v.dst, v.a, v.b, (int16_t)v.cmp, v.imm, v.sreg != 0xFF ? regs[v.sreg] : 0};
    }
    for (int q = 0; q < 2; ++q) {                                     // VGATHER's segments, whole
        if (p.gsegs[q] == 0xFF || p.gsegs[q] >= e.segs.size()) continue;
        Seg& g = *e.segs[p.gsegs[q]];
        prog.gcount[q] = g.count;
        if (g.count && !(prog.gcol[q] = seg_rows(g, e.dtype, 0, g.count, c))) return false;
    }
    if (!e.trap.ensure(sizeof(int))) return false;
    if (!hip_ok(launch_zero32(static_cast<uint32_t*>(e.trap.dev), 1, c->stream), "clear")) return false;
    const PredDev all = make_pred(e.dtype, CMP_TRUE, 0);
    for (const Plan::Block& b : p.blocks) {
        prog.target = b.vreg;
        if (!expr_eval(e.dtype, col, count, off, nrows, prog, e.expr.dev, static_cast<int*>(e.trap.dev), c)) return false;
        if (!run_tier0_block(b, all, e.dtype, e.expr.dev, nrows, 0, nrows, nrows / kl, regs, c)) return false;
    }
    // the stored vectors, evaluated before anything is written
    const size_t es = dtype_size(e.dtype);
    if (!p.stores.empty() && !e.stores.ensure(nrows * es * p.stores.size())) return false;
    for (size_t q = 0; q < p.stores.size(); ++q) {
        prog.target = p.stores[q].vreg;
        void* dst = static_cast<char*>(e.stores.dev) + q * nrows * es;
        if (!expr_eval(e.dtype, col, count, off, nrows, prog, dst, static_cast<int*>(e.trap.dev), c)) return false;
    }
    int trapped = 0;
    if (!sync_read(c, &trapped, e.trap.dev, sizeof(int))) return false;
    if (trapped) {
        set_error("a vector division or modulo of INT64_MIN by -1 (the reference's x86 build traps there)");
        return false;
    }
    // VSTOREs in program order (a later one wins where two write a row):
    // lanes [0, count) of each group go to rows off + 4g + lane below the
    // segment's Count, in HBM and, for a segment over caller memory, back in
    // the caller's rows (the reference writes its borrowed Segment<T>::Data)
    for (size_t q = 0; q < p.stores.size(); ++q) {
        const Plan::Store& st = p.stores[q];
        if (st.seg >= e.segs.size()) continue;                 // HandleVSTORE: no such segment, nothing written
        Seg& sg = *e.segs[st.seg];
        if (off >= sg.count) continue;
        void* d = const_cast<void*>(seg_rows(sg, e.dtype, 0, sg.count, c));
        if (!d) return false;
        const void* src = static_cast<const char*>(e.stores.dev) + q * nrows * es;
        if (!expr_store(e.dtype, src, nrows, st.count, d, off, sg.count, c)) return false;
        if (!sg.devext) {
            const uint64_t end = off + nrows < sg.count ? off + nrows : sg.count;
            void* host = const_cast<void*>(sg.host);
            if (!sync_read(c, static_cast<char*>(host) + off * es, static_cast<const char*>(d) + off * es,
                           (end - off) * es))
                return false;
        }
    }
    regs[p.off] = total;
    if (run_scalar(p.epilogue, regs) < 0) return false;
    std::memcpy(e.regs, regs, sizeof(uint64_t) * 64);
    return true;
}

bool run_tier0_expr(Engine& e, uint64_t* regs, uint64_t off, uint64_t total, StreamCtx* c) {
    const Plan& p = e.plan;
    Seg& s = *e.segs[p.seg];
    const uint64_t dend = total < s.count ? total : s.count;
    const void* d = seg_rows(s, e.dtype, off < dend ? off : dend, dend, c);
    if (!d) return false;
    return run_expr_blocks(e, regs, d, s.count, off, total - off, total, c);
}

// Loops loading a window word: the word writes lane j = the window value at
// row offset + j (0 once the window passes Count) and advances the offset by
// min(4, Count - offset) (WDELTA: by the lanes it wrote), so the loop ends
// only at a total it steps onto; the window values are computed once as a
// column (the window family's kernels, WINDOW_MEAN as sum * (1 / w)) and the
// body runs on it row by row.
bool run_tier0_window(Engine& e, uint64_t* regs, StreamCtx* c) {
    const Plan& p = e.plan;
    if (e.dtype != DT_F64) { set_error("window words are lowered for f64 engines only"); return false; }
    if (p.seg >= e.segs.size()) { set_error("program reads segment %u, engine has %zu", p.seg, e.segs.size()); return false; }
    Seg& s = *e.segs[p.seg];
    const uint64_t off = regs[p.off], total = regs[p.tot], count = s.count;
    const uint64_t last = p.wkind == 5 ? (count ? count - 1 : 0) : count;   // where the partial step lands
    if (off >= last && total == off) {
        // the word writes four zero lanes and advances by 0: the body runs
        // once on them and CMP falls through
        if (!e.wcol.ensure(32)) return false;
        return run_expr_blocks(e, regs, e.wcol.dev, 0, off, 4, total, c);
    }
    uint64_t groups = 0;
    bool ok = false;
    if (off < last && total > off) {
        if (total == last) { groups = (total - off + 3) / 4; ok = true; }
        else if (total < last && (total - off) % 4 == 0) {  // full steps of 4 all the way
            ok = true;
            groups = (total - off) / 4;
        }
    }
    if (!ok) {
        set_error("the reference loop would not terminate: the window word steps the offset R%u from %llu to %llu, "
                  "never onto R%u = %llu", p.off, (unsigned long long)off, (unsigned long long)last, p.tot,
                  (unsigned long long)total);
        return false;
    }
    const void* d = seg_rows(s, e.dtype, 0, count, c);
    if (!d) return false;
    if (!e.wcol.ensure(count * 8)) return false;
    // window_launch kinds: sum, mean by 1/w, std, var, delta, order statistic
    static const int kind_of[7] = {0, 1, 6, 3, 4, 0, 7};
    const int64_t nvalid = window_launch(kind_of[p.wkind], static_cast<const double*>(d), count, count, p.window,
                                         static_cast<double*>(e.wcol.dev), c, (double)p.wq);
    if (nvalid < 0) return false;
    return run_expr_blocks(e, regs, e.wcol.dev, (uint64_t)nvalid, off, groups * 4, total, c);
}

bool run_tier0(Engine& e, StreamCtx* c) {
    const Plan& p = e.plan;
    if (e.dtype != DT_F64 && e.dtype != DT_I64) { set_error("Tier-0 loops are lowered for f64 and i64 engines only"); return false; }
    uint64_t regs[64];
    std::memcpy(regs, e.regs, sizeof(regs));
    const int pre = run_scalar(p.prefix, regs);
    if (pre < 0) return false;
    if (pre == 0) {                                       // a division by zero stopped the program
        std::memcpy(e.regs, regs, sizeof(regs));
        return true;
    }
    if (p.wkind) return run_tier0_window(e, regs, c);
    const uint64_t kl = 4;
    if (regs[p.step] != kl) { set_error("loop step register R%u holds %llu, not the lane count %llu", p.step,
                                        (unsigned long long)regs[p.step], (unsigned long long)kl); return false; }
    if (p.vcount != 0 && p.vcount < kl) { set_error("VLOAD with a partial lane count is not lowered"); return false; }
    if (p.seg >= e.segs.size()) { set_error("program reads segment %u, engine has %zu", p.seg, e.segs.size()); return false; }
    uint64_t off = regs[p.off], total = regs[p.tot];
    if (total <= off || (total - off) % kl != 0) {
        set_error("the reference loop would not terminate: (R%u - R%u) = %lld is not a positive multiple of %llu "
                  "(JNZ loops until offset == total)", p.tot, p.off, (long long)(total - off), (unsigned long long)kl);
        return false;
    }
    if (p.has_arith) return run_tier0_expr(e, regs, off, total, c);
    // each block's predicate: the conjunction of its filter chain
    std::vector<PredDev> preds(p.blocks.size());
    for (size_t k = 0; k < p.blocks.size(); ++k) {
        const Plan::Block& b = p.blocks[k];
        PredDev pr = make_pred(e.dtype, CMP_TRUE, 0);
        for (int j = 0; j < b.nf; ++j) {
            const Plan::Filt& fl = p.filts[b.f[j]];
            const PredDev q = fl.cmp == 7 ? make_never_pred(e.dtype) : make_pred(e.dtype, fl.cmp, regs[fl.thr]);
            if (j == 0) { pr = q; continue; }
            PredDev m;
            if (!merge_preds(e.dtype, pr, q, &m)) {
                set_error("a filter chain with a != filter and another filter is not lowered");
                return false;
            }
            pr = m;
        }
        pr.table = 0;
        preds[k] = pr;
    }
    Seg& s = *e.segs[p.seg];
    uint64_t dend = total < s.count ? total : s.count;    // rows past Count load as 0
    const void* d = seg_rows(s, e.dtype, off < dend ? off : dend, dend, c);
    if (!d) return false;
    const uint64_t groups = (total - off) / kl;
    for (size_t k = 0; k < p.blocks.size(); ++k)
        if (!run_tier0_block(p.blocks[k], preds[k], e.dtype, d, s.count, off, dend, groups, regs, c)) return false;
    regs[p.off] = total;
    if (run_scalar(p.epilogue, regs) < 0) return false;
    std::memcpy(e.regs, regs, sizeof(regs));
    return true;
}

// ---------------------------------------------------------------------------
// Straight-line programs: scalar words on the host, segment aggregates
// (Engine<T>::HandleAggregate, engine.hpp:499-698) on the GPU, with the
// engine's 16 persistent slots.  Sums, averages and moments are compensated
// or merged in a fixed order (within a few ulp of the reference's row-order
// f64 sums and Welford updates); counts, min/max, first/last, distinct
// counts, median and percentile are exact.
// ---------------------------------------------------------------------------
namespace {
void dd_add(double& hi, double& lo, double s, double c) {
    if (!std::isfinite(hi) || !std::isfinite(s)) { hi = hi + s; lo = 0.0; return; }
    const double t = hi + s, bp = t - hi, e = (hi - (t - bp)) + (s - bp);
    const double l = lo + e + c;
    hi = t + l;
    lo = l - (hi - t);
}

bool read_bits(StreamCtx* c, const void* dev, uint64_t row, int dtype, uint64_t* out) {
    const size_t es = dtype_size(dtype);
    uint64_t v = 0;
    if (!sync_read(c, &c->host->fold[0], static_cast<const char*>(dev) + row * es, es)) return false;
    std::memcpy(&v, &c->host->fold[0], es);
    *out = v;
    return true;
}

bool run_agg(Engine& e, uint32_t w, uint64_t* regs, AggSlot* work, StreamCtx* c) {
    const uint8_t op = op_of(w), rd = rd_of(w), seg = ra_of(w), rb = rb_of(w);
    const uint16_t imm = imm_of(w);
    if (seg >= e.segs.size()) return true;                       // missing segment: PC++ only
    Seg& s = *e.segs[seg];
    const bool F = e.dtype == DT_F64;
    const uint64_t start = regs[rb];
    const uint64_t cnt = imm == 0 ? 0xFFFFFFFFull : (uint64_t)(imm & 0xFFF);
    uint64_t end = start + cnt;
    if (end > s.count) end = s.count;
    const uint64_t n = end > start ? end - start : 0;
    const int slot = rd & 0xF;
    AggSlot& a = work[slot];
    const size_t es = dtype_size(e.dtype);
    const char* d = nullptr;
    if (n) {
        const void* base = seg_rows(s, e.dtype, start, end, c);
        if (!base) return false;
        d = static_cast<const char*>(base) + start * es;
    }
    auto scan = [&](int cmp, uint64_t tbits, ScanOut* r) {
        ScanArgs g{};
        g.dtype = e.dtype; g.a = d; g.begin = 0; g.end = n; g.mode = 0;
        g.pa = make_pred(e.dtype, cmp, tbits);
        return scan_once(c, g, r);
    };
    const double dbl_max = 1.7976931348623157e308;
    switch (op) {
    case 0xE0: {                                                 // COUNT: values != 0 (NaN counts)
        if (!a.init) { a.count = 0; a.init = true; }
        if (n) { ScanOut r{}; if (!scan(CMP_NE, 0, &r)) return false; a.count += r.cnt; }
        regs[rd] = a.count;
        return true;
    }
    case 0xE1:                                                   // SUM
    case 0xE2: {                                                 // AVG
        if (!a.init) {
            if (op == 0xE2) a.count = 0;
            a.sh = a.sl = 0.0; a.sumi = 0; a.init = true;
        }
        if (n) {
            ScanOut r{};
            if (!scan(CMP_TRUE, 0, &r)) return false;
            if (F) dd_add(a.sh, a.sl, bits_f64(r.s), bits_f64(r.c));
            else a.sumi = (int64_t)((uint64_t)a.sumi + r.s);
        }
        if (op == 0xE1) {
            regs[rd] = F ? f64_bits(a.sh + a.sl) : (uint64_t)a.sumi;
        } else {
            a.count += n;
            if (F) regs[rd] = f64_bits(a.count ? (a.sh + a.sl) / (double)a.count : 0.0);
            else regs[rd] = (uint64_t)(a.count ? a.sumi / (int64_t)a.count : 0);
        }
        return true;
    }
    case 0xE3:
    case 0xE4: {                                                 // MIN, MAX (row < current, NaN never taken)
        if (!a.init) {
            if (op == 0xE3) a.minv = F ? f64_bits(dbl_max) : (uint64_t)INT64_MAX;
            else a.maxv = F ? f64_bits(-dbl_max) : (uint64_t)INT64_MIN;
            a.init = true;
        }
        if (n) {
            ScanOut r{};
            if (!scan(CMP_TRUE, 0, &r)) return false;
            voxel_gpu_agg g{};
            fill_agg(e.dtype, r, &g);
            if (F && g.has_minmax) {
                if (op == 0xE3 && g.fmin < bits_f64(a.minv)) a.minv = f64_bits(g.fmin);
                if (op == 0xE4 && g.fmax > bits_f64(a.maxv)) a.maxv = f64_bits(g.fmax);
            } else if (!F) {
                if (op == 0xE3 && g.imin < (int64_t)a.minv) a.minv = (uint64_t)g.imin;
                if (op == 0xE4 && g.imax > (int64_t)a.maxv) a.maxv = (uint64_t)g.imax;
            }
        }
        regs[rd] = op == 0xE3 ? a.minv : a.maxv;
        return true;
    }
    case 0xE5:                                                   // FIRST
        if (!a.init) { a.first_set = false; a.init = true; }
        if (!a.first_set && n) {
            if (!read_bits(c, d, 0, e.dtype, &a.firstv)) return false;
            a.first_set = true;
        }
        regs[rd] = a.firstv;
        return true;
    case 0xE6:                                                   // LAST
        if (!a.init) a.init = true;
        if (n && !read_bits(c, d, n - 1, e.dtype, &a.lastv)) return false;
        regs[rd] = a.lastv;
        return true;
    case 0xE7:
    case 0xE8: {                                                 // STDDEV, VARIANCE: Welford, merged with Chan's formula
        if (!a.init) { a.count = 0; a.sh = a.sl = 0.0; a.sumi = 0; a.mean = 0.0; a.m2 = 0.0; a.init = true; }
        if (n) {
            Moments m;
            if (!segagg_moments(e.dtype, d, n, &m, e.segagg, c)) return false;
            const double na = (double)a.count, nb = m.n, nt = na + nb;
            const double dlt = m.mean - a.mean;
            if (a.count == 0) { a.mean = m.mean; a.m2 = m.m2; }
            else { a.mean = a.mean + dlt * (nb / nt); a.m2 = a.m2 + m.m2 + dlt * dlt * (na * (nb / nt)); }
            a.count += n;
        }
        const double var = a.count > 1 ? a.m2 / (double)a.count : 0.0;
        const double res = op == 0xE7 ? std::sqrt(var) : var;
        regs[rd] = F ? f64_bits(res) : (uint64_t)(int64_t)res;
        return true;
    }
    case 0xE9:                                                   // COUNT_DISTINCT
    case 0xEA: {                                                 // SUM_DISTINCT (the same slot set, engine.hpp:629-647)
        const char* name = op == 0xE9 ? "AGG_COUNT_DISTINCT" : "AGG_SUM_DISTINCT";
        if (!a.init) { a.has_set = true; a.nset = 0; a.init = true; }
        if (!a.has_set) { set_error("%s on a slot another aggregate started: the reference dereferences a null set", name); return false; }
        if (n) {
            if (F) {
                ScanOut r{};
                if (!scan(CMP_LE, f64_bits(INFINITY), &r)) return false;     // non-NaN rows
                if (r.cnt != n) { set_error("%s over NaN values is not lowered (each NaN is its own key in the reference)", name); return false; }
            }
            SegaggBufs& b = e.segagg;
            if (!b.ensure(n + a.nset)) return false;                           // no reallocation below
            if (!segagg_sorted_keys(e.dtype, d, n, b, c)) return false;
            uint64_t nu = 0;
            if (!segagg_unique(b.k2, n, b.k1, &nu, c)) return false;          // new distinct keys in k1
            uint64_t total = nu + a.nset;
            if (a.nset) {                                                      // merge: sort old + new, unique again
                if (!hip_ok(hipMemcpyAsync(b.k1 + nu, e.sets[slot][a.cur].p, a.nset * 8, hipMemcpyDeviceToDevice,
                                           c->stream), "copy set") ||
                    !hip_ok(launch_iota(b.v1, total, c->stream), "iota"))
                    return false;
                if (sort_pairs_u64(b.k1, b.v1, b.k2, b.v2, total, c) < 0) return false;
                if (!segagg_unique(b.k2, total, b.k1, &nu, c)) return false;
            }
            // the new set goes to the buffer that is not committed (it may
            // be this run's earlier set, already read into k1 above)
            const int out = 1 - e.agg[slot].cur;
            SetBuf& dst = e.sets[slot][out];
            if (nu > dst.cap) {
                if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return false;
                if (dst.p) (void)hipFree(dst.p);
                dst.p = nullptr; dst.cap = 0;
                if (!hip_ok(hipMalloc((void**)&dst.p, nu * 8), "hipMalloc(distinct set)")) return false;
                dst.cap = nu;
            }
            if (nu && !hip_ok(hipMemcpyAsync(dst.p, b.k1, nu * 8, hipMemcpyDeviceToDevice, c->stream), "copy set")) return false;
            a.nset = nu;
            a.cur = out;
        }
        if (op == 0xE9) { regs[rd] = a.nset; return true; }
        // the set's values summed: i64 wrapping (order-free, exact); f64
        // compensated (the reference adds them in unordered_map order)
        if (a.nset == 0) { regs[rd] = 0; return true; }
        const uint64_t* set = e.sets[slot][a.cur].p;
        ScanArgs g{};
        g.begin = 0; g.end = a.nset; g.mode = 0;
        g.pa = make_pred(e.dtype, CMP_TRUE, 0);
        ScanOut r{};
        if (F) {
            SegaggBufs& b = e.segagg;
            if (!b.ensure(a.nset)) return false;
            if (!segagg_unkey_f64(set, a.nset, b.k2, c)) return false;
            g.dtype = DT_F64; g.a = b.k2;
            if (!scan_once(c, g, &r)) return false;
            regs[rd] = f64_bits(bits_f64(r.s) + bits_f64(r.c));
        } else {
            g.dtype = DT_I64; g.a = set;                                        // key = value + 2^63 (mod 2^64)
            if (!scan_once(c, g, &r)) return false;
            regs[rd] = r.s + (a.nset << 63);
        }
        return true;
    }
    case 0xEC: {                                                 // MODE (no slot state, engine.hpp:669-681)
        if (n == 0) { regs[rd] = 0; return true; }
        uint64_t row = 0, ties = 0;
        if (!segagg_mode(e.dtype, d, n, e.segagg, &row, &ties, c)) return false;
        if (ties != 1) {
            set_error("AGG_MODE: %llu values share the largest count; the reference picks one by unordered_map "
                      "iteration order, which is not lowered", (unsigned long long)ties);
            return false;
        }
        return read_bits(c, d, row, e.dtype, &regs[rd]);
    }
    case 0xEB:
    case 0xED: {                                                 // MEDIAN, PERCENTILE (no slot state)
        if (n == 0) { regs[rd] = 0; return true; }
        if (F) {
            ScanOut r{};
            if (!scan(CMP_LE, f64_bits(INFINITY), &r)) return false;
            if (r.cnt != n) { set_error("AGG_MEDIAN / AGG_PERCENTILE over NaN values is not lowered (the reference's nth_element order is unspecified)"); return false; }
        }
        if (!segagg_sorted_keys(e.dtype, d, n, e.segagg, c)) return false;
        auto key_at = [&](uint64_t i, uint64_t* v) {
            uint64_t k;
            if (!sync_read(c, &c->host->fold[0], e.segagg.k2 + i, 8)) return false;
            std::memcpy(&k, &c->host->fold[0], 8);
            *v = order_unkey(k, e.dtype);
            return true;
        };
        if (op == 0xED) {
            const double pct = (double)(imm & 0xFF) / 100.0;
            uint64_t idx = (uint64_t)(pct * (double)n);
            if (idx >= n) idx = n - 1;
            uint64_t v;
            if (!key_at(idx, &v)) return false;
            regs[rd] = v;
        } else if (n % 2 == 0) {
            uint64_t hi, lo;
            if (!key_at(n / 2, &hi) || !key_at(n / 2 - 1, &lo)) return false;
            regs[rd] = F ? f64_bits((bits_f64(hi) + bits_f64(lo)) * 0.5) : (uint64_t)(((int64_t)hi + (int64_t)lo) / 2);
        } else {
            uint64_t v;
            if (!key_at(n / 2, &v)) return false;
            regs[rd] = v;
        }
        return true;
    }
    default:
        set_error("aggregate word 0x%02x is not lowered", op);
        return false;
    }
}
}  // namespace

// ---------------------------------------------------------------------------
// Hash / sort / join words (engine.hpp:704-842, dispatch 3286-3355) in
// straight-line programs.  The sort words reorder the segment's rows where
// the caller keeps them, as the reference's sorts write its borrowed
// Segment<T>::Data: device segments in place, host segments in HBM and then
// back to the caller's memory (and the resident mirror stays current).
// ---------------------------------------------------------------------------
namespace {
// Device rows [0, Count) of a segment that a word will reorder in place.
void* seg_rows_mut(Seg& s, int dtype, StreamCtx* c) {
    return const_cast<void*>(seg_rows(s, dtype, 0, s.count, c));
}
bool seg_write_back(Seg& s, int dtype, const void* dev, StreamCtx* c) {
    if (s.devext || s.count == 0) return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    void* host = const_cast<void*>(s.host);               // the reference sorts its borrowed rows
    return sync_read(c, host, dev, s.count * dtype_size(dtype));
}

bool run_relword(Engine& e, uint32_t w, uint64_t* regs, TabSlot* twork, StreamCtx* c) {
    const uint8_t op = op_of(w), rd = rd_of(w), ra = ra_of(w), rb = rb_of(w);
    const uint16_t imm = imm_of(w);
    const size_t nseg = e.segs.size();
    switch (op) {
    case 0xEE:
    case 0xF0: {                                            // HASH_INIT / HASH_BUILD
        if (ra >= nseg) return true;
        Seg& s = *e.segs[ra];
        const int slot = rd & 0xF;
        const int out = 1 - e.tab[slot].cur;                // never the committed table
        const void* d = s.count ? seg_rows(s, e.dtype, 0, s.count, c) : nullptr;
        if (s.count && !d) return false;
        uint64_t distinct = 0;
        if (!relw_hash_build(e.dtype, d, s.count, e.tabs[slot][out], &distinct, e.rel, c)) return false;
        twork[slot] = TabSlot{distinct, out, true};
        regs[rd] = distinct;
        return true;
    }
    case 0xEF:
    case 0xF1: {                                            // HASH_PROBE / HASH_LOOKUP
        const TabSlot& t = twork[ra & 0xF];
        uint64_t row = ~0ull;
        if (t.valid && !relw_hash_probe(e.tabs[ra & 0xF][t.cur], t.n, relw_hash_element(regs[rb], e.dtype), &row,
                                        e.rel, c))
            return false;
        regs[rd] = row;
        return true;
    }
    case 0xF2:
    case 0xF3: {                                            // SORT_ASC / SORT_DESC
        if (ra >= nseg) return true;
        Seg& s = *e.segs[ra];
        if (s.count >= 2) {
            void* d = seg_rows_mut(s, e.dtype, c);
            if (!d || !relw_sort(e.dtype, d, s.count, op == 0xF3, e.rel, c) || !seg_write_back(s, e.dtype, d, c))
                return false;
        }
        regs[rd] = s.count;
        return true;
    }
    case 0xF4:
    case 0xF5: {                                            // SORT_TOPK (greater) / SORT_BOTTOMK (less)
        const uint64_t k = regs[rb] < (uint64_t)imm ? regs[rb] : (uint64_t)imm;
        if (ra >= nseg || k == 0) return true;
        Seg& s = *e.segs[ra];
        if (k > s.count) {
            set_error("SORT_TOPK / SORT_BOTTOMK of %llu rows from a segment of %zu (the reference's partial_sort "
                      "reads past the segment there)", (unsigned long long)k, s.count);
            return false;
        }
        void* d = seg_rows_mut(s, e.dtype, c);
        if (!d || !relw_partial_sort(e.dtype, d, s.count, (uint32_t)k, op == 0xF4, c) ||
            !seg_write_back(s, e.dtype, d, c))
            return false;
        regs[rd] = k;
        return true;
    }
    case 0xFD:                                              // PARTITION_HASH
        if (ra < nseg) regs[rd] = e.segs[ra]->count;
        return true;
    default: {                                              // JOIN_HASH .. JOIN_SEMI
        if (ra >= nseg || rb >= nseg) return true;
        static const int kind_of[5] = {0, 1, 0, 3, 2};
        Seg& l = *e.segs[ra];
        Seg& r = *e.segs[rb];
        const void* dl = l.count ? seg_rows(l, e.dtype, 0, l.count, c) : nullptr;
        if (l.count && !dl) return false;
        const void* dr = r.count ? seg_rows(r, e.dtype, 0, r.count, c) : nullptr;
        if (r.count && !dr) return false;
        uint64_t m = 0;
        if (!relw_join_count(kind_of[op - 0xF6], e.dtype, dl, l.count, dr, r.count, &m, e.rel, c)) return false;
        regs[rd] = m;
        return true;
    }
    }
}
}  // namespace

bool run_straight(Engine& e, StreamCtx* c) {
    if (e.dtype != DT_F64 && e.dtype != DT_I64) {
        set_error("segment aggregate, hash, sort and join words are lowered for f64 and i64 engines only");
        return false;
    }
    if (!e.agg_ready) {                                          // AggState defaults (engine.hpp:347-360)
        const bool F = e.dtype == DT_F64;
        for (AggSlot& a : e.agg) {
            a.minv = F ? f64_bits(1.7976931348623157e308) : (uint64_t)INT64_MAX;
            a.maxv = F ? f64_bits(-1.7976931348623157e308) : (uint64_t)INT64_MIN;
        }
        e.agg_ready = true;
    }
    // registers and slots change only when the whole program succeeds
    uint64_t regs[64];
    std::memcpy(regs, e.regs, sizeof(regs));
    AggSlot work[16];
    std::copy(std::begin(e.agg), std::end(e.agg), std::begin(work));
    TabSlot twork[16];
    std::copy(std::begin(e.tab), std::end(e.tab), std::begin(twork));
    for (uint32_t w : e.plan.words) {
        if (agg_op(op_of(w))) {
            if (!run_agg(e, w, regs, work, c)) return false;
            continue;
        }
        if (rel_word_op(op_of(w))) {
            if (!run_relword(e, w, regs, twork, c)) return false;
            continue;
        }
        const int rc = run_scalar(std::vector<uint32_t>{w}, regs);
        if (rc < 0) return false;
        if (rc == 0) break;                                      // division by zero stops the program
    }
    std::copy(std::begin(work), std::end(work), std::begin(e.agg));
    std::copy(std::begin(twork), std::end(twork), std::begin(e.tab));
    std::memcpy(e.regs, regs, sizeof(regs));
    return true;
}

Engine* eng(void* h) { return static_cast<Engine*>(h); }

void engine_run(void* h) {
    set_error("%s", "");
    if (!h) { set_error("null engine"); return; }
    Engine& e = *eng(h);
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return;
    if (e.plan.kind == Plan::FAST) run_fast(e, c);
    else if (e.plan.kind == Plan::TIER0) run_tier0(e, c);
    else if (e.plan.kind == Plan::STRAIGHT) run_straight(e, c);
    else set_error("program not lowered to the GPU: %s", e.plan.why.c_str());
}

size_t engine_add(void* h, const void* data, size_t count) {
    Engine& e = *eng(h);
    auto s = std::make_unique<Seg>();
    s->host = data;
    s->count = count;
    e.segs.push_back(std::move(s));
    return e.segs.size() - 1;
}

void engine_load(void* h, const uint32_t* code, size_t len) {
    Engine& e = *eng(h);
    e.code.assign(code, code + len);
    e.plan = lower_engine_program(e.code.data(), e.code.size());
}

}  // namespace

// Result conversion and descriptor checks shared with the batch and serving
// paths (declared in vx_runtime.hpp).
namespace vx {
void fill_agg(int dtype, const ScanOut& r, voxel_gpu_agg* out) {
    std::memset(out, 0, sizeof(*out));
    out->count = r.cnt;
    if (dtype == DT_F64) {
        double s = bits_f64(r.s), cc = bits_f64(r.c);
        out->sum_hi = s; out->sum_lo = cc;
        out->sum = finish_sum_f64(0.0, s, cc);
        out->fmin = bits_f64(r.mn); out->fmax = bits_f64(r.mx);
        out->has_minmax = r.cnt > 0 && out->fmin <= out->fmax;
    } else if (dtype == DT_F32) {
        float s = bits_f32(r.s), cc = bits_f32(r.c);
        out->sum_hi = s; out->sum_lo = cc;
        out->sum = finish_sum_f32(0.0f, s, cc);
        out->fmin = bits_f32(r.mn); out->fmax = bits_f32(r.mx);
        out->has_minmax = r.cnt > 0 && out->fmin <= out->fmax;
    } else if (dtype == DT_I64) {
        out->isum = (int64_t)r.s;
        out->imin = (int64_t)r.mn; out->imax = (int64_t)r.mx;
        out->has_minmax = r.cnt > 0;
    } else {
        out->isum = (int64_t)r.s;
        out->imin = (int32_t)(uint32_t)r.mn; out->imax = (int32_t)(uint32_t)r.mx;
        out->has_minmax = r.cnt > 0;
    }
}

bool make_scan_args(const voxel_gpu_scan_desc* d, ScanArgs& g) {
    if (!d) { set_error("null scan descriptor"); return false; }
    if (d->dtype < DT_F64 || d->dtype > DT_I32) { set_error("scan dtype %d not supported", d->dtype); return false; }
    if (!d->a && d->n) { set_error("scan: column a is null"); return false; }
    if (d->op_a < 0 || d->op_a > 6 || d->op_b < 0 || d->op_b > 6) { set_error("scan: predicate code out of range"); return false; }
    g = ScanArgs{};
    g.dtype = d->dtype; g.a = d->a; g.b = d->b; g.begin = 0; g.end = d->n; g.blocks = d->blocks;
    g.unroll = d->unroll; g.nontemporal = d->nontemporal; g.layout = d->layout;
    if (g.layout < 0 || g.layout > 4) { set_error("scan: layout must be 0..4"); return false; }
    if (g.unroll != 0 && g.unroll != 2 && g.unroll != 4 && g.unroll != 8) { set_error("scan: unroll must be 2, 4 or 8"); return false; }
    g.pa = make_pred(d->dtype, d->op_a, d->t_a);
    if (d->b && d->op_b != CMP_TRUE) {
        g.mode = d->b == d->a ? 1 : 2;                    // both predicates on one column: read it once
        g.pb = make_pred(d->dtype, d->op_b, d->t_b);
    } else {
        g.mode = 0;
    }
    return true;
}


}  // namespace vx

extern "C" {

// ---- Engine (f64) : reference bindings/voxel_c.cpp:14-49 ----
void* voxel_engine_create_f64(void) { auto* e = new Engine(); e->dtype = DT_F64; return e; }
void voxel_engine_destroy(void* h) { delete eng(h); }
size_t voxel_engine_add_segment(void* h, const double* data, size_t count) { return engine_add(h, data, count); }
void voxel_engine_load_program(void* h, const uint32_t* code, size_t len) { engine_load(h, code, len); }
void voxel_engine_run(void* h) { engine_run(h); }

}  // extern "C"

namespace vx {
namespace {
// RunParallel of a Tier-0 or straight-line program (engine.hpp:143-175):
// worker t is a fresh engine whose only segment (id 0) is chunk t of
// segment seg_id (rows [t * ceil(N / threads), ...), in HBM), holding the
// program and a copy of the scalar registers; the workers run one after
// another on the GPU, and R[result] = the sum in T of their R[result] in
// worker order, starting from zero.  Workers whose chunk starts at or past N
// are skipped (the reference's worker reads out of bounds there).
void run_parallel_program(Engine& e, size_t seg_id, uint32_t threads, uint8_t rr, StreamCtx* c) {
    Seg& s = *e.segs[seg_id];
    const uint64_t n = s.count, es = dtype_size(e.dtype);
    const uint64_t chunk = (n + threads - 1) / threads;
    const char* d = nullptr;
    if (n) {
        d = static_cast<const char*>(seg_rows(s, e.dtype, 0, n, c));
        if (!d) return;
    }
    double fsum = 0.0;
    uint64_t isum = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        const uint64_t start = (uint64_t)t * chunk;
        if (start >= n) continue;
        const uint64_t cnt = start + chunk > n ? n - start : chunk;
        Engine w;
        w.dtype = e.dtype;
        auto sg = std::make_unique<Seg>();
        sg->devext = d + start * es;
        sg->count = cnt;
        w.segs.push_back(std::move(sg));
        w.code = e.code;
        w.plan = e.plan;
        std::memcpy(w.regs, e.regs, sizeof(w.regs));
        const bool ok = e.plan.kind == Plan::TIER0 ? run_tier0(w, c) : run_straight(w, c);
        if (!ok) return;
        if (e.dtype == DT_F64) fsum = fsum + bits_f64(w.regs[rr]);
        else isum += w.regs[rr];
    }
    // sort words sort each worker's chunk in place: the reference's workers
    // hold the caller's rows, so the sorted chunks go back to them
    bool sorts = false;
    for (uint32_t wd : e.code) sorts |= op_of(wd) >= 0xF2 && op_of(wd) <= 0xF5;
    if (sorts && n && !s.devext && s.host && !sync_read(c, const_cast<void*>(s.host), d, n * es)) return;
    e.regs[rr] = e.dtype == DT_F64 ? f64_bits(fsum) : isum;
}
}  // namespace
}  // namespace vx

extern "C" {

// Engine<f64>::RunParallel (engine.hpp:143-175): row chunks of ceil(N/threads),
// each runs the fast path with copied registers (dropping its own tail), and
// R[result] = sum over workers of their R[result] in T.  Workers whose chunk
// starts at or past N are skipped (the reference reads out of bounds there).
// Tier-0 and straight-line programs go through run_parallel_program.
void voxel_engine_run_parallel(void* h, size_t seg_id, uint32_t threads, uint8_t result_reg) {
    set_error("%s", "");
    if (!h) { set_error("null engine"); return; }
    Engine& e = *eng(h);
    if (e.segs.empty() || seg_id >= e.segs.size()) return;          // engine.hpp:146
    if (threads == 0) threads = std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    if (e.plan.kind == Plan::TIER0 || e.plan.kind == Plan::STRAIGHT) {
        std::lock_guard<std::mutex> lk(api_mutex());
        StreamCtx* c = ctx_for(nullptr);
        if (!c) return;
        run_parallel_program(e, seg_id, threads, result_reg & 0x3F, c);
        return;
    }
    if (e.plan.kind != Plan::FAST) { set_error("program not lowered to the GPU: %s", e.plan.why.c_str()); return; }
    if (e.plan.seg != 0) { set_error("run_parallel workers hold one segment; the program reads segment %u", e.plan.seg); return; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return;
    Seg& s = *e.segs[seg_id];
    const uint64_t n = s.count, kl = lanes_of(e.dtype);
    const uint64_t chunk = (n + threads - 1) / threads;
    const uint64_t r_off = e.regs[e.plan.off], r_tot = e.regs[e.plan.tot];
    PredDev pred = make_pred(e.dtype, cmp_of_filter_op(e.plan.fop), e.regs[e.plan.thr]);
    const void* d = nullptr;
    if (n) { d = seg_rows(s, e.dtype, 0, n, c); if (!d) return; }
    // Each worker's R[result] after its fast path; summed in T in worker
    // order.  R[acc] values (acc0 + S_t) are summed as a compensated pair.
    const uint8_t rr = result_reg & 0x3F;
    const bool res_is_off = rr == e.plan.off;
    const bool res_is_acc = rr == e.plan.acc && !res_is_off;
    double acc0 = bits_f64(e.regs[e.plan.acc]);
    double hs = 0.0, hc = 0.0, plain = 0.0;
    uint64_t isum = 0;
    auto add = [&](double v) { double t = hs + v, bp = t - hs; hc += (hs - (t - bp)) + (v - bp); hs = t; };
    for (uint32_t t = 0; t < threads; ++t) {
        uint64_t start = (uint64_t)t * chunk;
        if (start >= n) continue;
        uint64_t cnt = (start + chunk > n) ? (n - start) : chunk;
        uint64_t total = r_tot < cnt ? r_tot : cnt;
        uint64_t groups = (r_off <= total && total - r_off >= kl) ? (total - r_off) / kl : 0;
        uint64_t woff = r_off + groups * kl;
        if (!res_is_acc) {
            uint64_t v = res_is_off ? woff : e.regs[rr];
            if (e.dtype == DT_F64) plain = plain + bits_f64(v); else isum += v;
            continue;
        }
        ScanOut r{};
        if (groups) {
            ScanArgs g{};
            g.dtype = e.dtype; g.a = d; g.begin = start + r_off; g.end = start + woff; g.pa = pred;
            if (!scan_once(c, g, &r)) return;
        }
        if (e.dtype == DT_F64) {
            add(acc0);
            if (groups) { add(bits_f64(r.s)); add(bits_f64(r.c)); }
        } else {
            isum += e.regs[e.plan.acc] + r.s;
        }
    }
    if (e.dtype == DT_F64) e.regs[rr] = f64_bits(res_is_acc ? hs + hc : plain);
    else e.regs[rr] = isum;
}

void voxel_engine_set_scalar(void* h, size_t reg, uint64_t v) { eng(h)->regs[reg & 0x3F] = v; }
uint64_t voxel_engine_get_scalar(void* h, size_t reg) { return eng(h)->regs[reg & 0x3F]; }
void voxel_engine_set_scalar_f64(void* h, size_t reg, double v) { eng(h)->regs[reg & 0x3F] = f64_bits(v); }
double voxel_engine_get_scalar_f64(void* h, size_t reg) { return bits_f64(eng(h)->regs[reg & 0x3F]); }
size_t voxel_engine_k_lanes(void) { return 4; }

// ---- Engine (i64) : reference bindings/voxel_c.cpp:54-77 ----
void* voxel_engine_create_i64(void) { auto* e = new Engine(); e->dtype = DT_I64; return e; }
void voxel_engine_destroy_i64(void* h) { delete eng(h); }
size_t voxel_engine_add_segment_i64(void* h, const int64_t* data, size_t count) { return engine_add(h, data, count); }
void voxel_engine_load_program_i64(void* h, const uint32_t* code, size_t len) { engine_load(h, code, len); }
void voxel_engine_run_i64(void* h) { engine_run(h); }
void voxel_engine_set_scalar_i64(void* h, size_t reg, uint64_t v) { eng(h)->regs[reg & 0x3F] = v; }
uint64_t voxel_engine_get_scalar_i64(void* h, size_t reg) { return eng(h)->regs[reg & 0x3F]; }

// ---- JIT : reference bindings/voxel_c.cpp:80-101 ----
static double jit_run_impl(const uint32_t* code, size_t code_len, const double* data, size_t data_len,
                           uint64_t r2_in, double threshold) {
    set_error("%s", "");
    JitPlan jp = lower_jit_program(code, code_len);
    if (!jp.ok) { set_error("JIT lowering: %s", jp.why.c_str()); return 0.0; }
    uint64_t regs[64] = {};
    regs[2] = r2_in;
    regs[3] = f64_bits(threshold);
    run_prefix(jp.prefix, regs);
    uint64_t off = regs[jp.off], r2 = regs[2];
    uint64_t iters = (r2 - off + 7) >> 3;                 // x86_64.hpp:1659-1668
    if (iters == 0) { set_error("the reference JIT loop would not terminate: no rows (R2 - offset + 7 < 8)"); return 0.0; }
    uint64_t end = off + 8 * iters;
    if (end > data_len) end = data_len;                   // rows past the segment read as 0
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return 0.0;
    ScanOut r{};
    if (end > off) {
        // staging buffer kept for the life of the process (never freed at exit)
        static DevColumn* stage_p = new DevColumn();
        DevColumn& stage = *stage_p;
        if (!stage.ensure(data_len * 8)) return 0.0;
        if (!hip_ok(hipMemcpyAsync(static_cast<char*>(stage.dev) + off * 8, data + off, (end - off) * 8,
                                   hipMemcpyHostToDevice, c->stream), "copy JIT data to HBM")) return 0.0;
        int cmp = cmp_of_filter_op(jp.fop);
        ScanArgs g{};
        g.dtype = DT_F64; g.a = stage.dev; g.begin = off; g.end = end; g.mode = 0;
        g.pa = make_pred(DT_F64, cmp, regs[jp.thr], true);
        g.table_a = cmp == CMP_NE;                        // vcmppd NEQ_OQ: NaN fails
        if (!scan_once(c, g, &r)) return 0.0;
    }
    return bits_f64(r.s) + bits_f64(r.c);                 // the initial R[acc] is not added
}

double voxel_jit_run(const uint32_t* code, size_t code_len, const double* data, size_t data_len, double threshold) {
    return jit_run_impl(code, code_len, data, data_len, data_len, threshold);
}

double voxel_gpu_jit_run_r2(const uint32_t* code, size_t code_len, const double* data, size_t data_len,
                            uint64_t r2, double threshold) {
    return jit_run_impl(code, code_len, data, data_len, r2, threshold);
}

// ---- GPU extensions ----
int voxel_gpu_available(void) { return runtime_ok() ? 1 : 0; }

void* voxel_gpu_malloc(size_t bytes) {
    if (!runtime_ok()) return nullptr;
    void* p = nullptr;
    if (!hip_ok(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc")) return nullptr;
    return p;
}
void voxel_gpu_free(void* p) { if (p) (void)hipFree(p); }
int voxel_gpu_memcpy_h2d(void* dev, const void* host, size_t bytes) {
    return hip_ok(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D") ? 0 : -1;
}
int voxel_gpu_memcpy_d2h(void* host, const void* dev, size_t bytes) {
    return hip_ok(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H") ? 0 : -1;
}
int voxel_gpu_synchronize(void) { return hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize") ? 0 : -1; }

void* voxel_gpu_engine_create(int dtype) {
    if (dtype != DT_F64 && dtype != DT_I64 && dtype != DT_I32) {
        set_error("typed engine: dtype %d has no Engine<T> in the reference (f64, i64, i32 only)", dtype);
        return nullptr;
    }
    auto* e = new Engine();
    e->dtype = dtype;
    return e;
}
size_t voxel_gpu_engine_add_segment_raw(void* h, const void* data, size_t count) { return engine_add(h, data, count); }
int voxel_gpu_engine_set_resident(void* h, size_t seg, int resident) {
    Engine& e = *eng(h);
    if (seg >= e.segs.size()) { set_error("no segment %zu", seg); return -1; }
    e.segs[seg]->resident = resident != 0;
    e.segs[seg]->col.valid = false;
    return 0;
}
size_t voxel_gpu_engine_add_device_segment(void* h, const void* dev, size_t count) {
    Engine& e = *eng(h);
    auto s = std::make_unique<Seg>();
    s->devext = dev;
    s->count = count;
    e.segs.push_back(std::move(s));
    return e.segs.size() - 1;
}

int voxel_gpu_jit_lowerable(const uint32_t* code, size_t len) {
    set_error("%s", "");
    JitPlan jp = lower_jit_program(code, len);
    if (!jp.ok) set_error("JIT lowering: %s", jp.why.c_str());
    return jp.ok ? 1 : 0;
}

int voxel_gpu_engine_lowered(const void* h) {
    const Engine& e = *static_cast<const Engine*>(h);
    return e.plan.kind == Plan::FAST ? 1 : e.plan.kind == Plan::TIER0 ? 2 : e.plan.kind == Plan::STRAIGHT ? 3 : 0;
}

int voxel_gpu_scan_launch(const voxel_gpu_scan_desc* d, void* stream) {
    set_error("%s", "");
    ScanArgs g;
    if (!make_scan_args(d, g)) return -1;
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    c->last_dtype = d->dtype;
    if (g.end <= g.begin) {
        return hip_ok(hipMemsetAsync(&c->dev->scan, 0, sizeof(ScanOut), c->stream), "clear result") ? 0 : -1;
    }
    return hip_ok(launch_scan(g, c->scr, &c->dev->scan, c->stream), "scan launch") ? 0 : -1;
}

int voxel_gpu_scan_result(void* stream, voxel_gpu_agg* out) {
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (c->last_dtype < 0) { set_error("no scan launched on this stream"); return -1; }
    if (!sync_read(c, &c->host->scan, &c->dev->scan, sizeof(ScanOut))) return -1;
    fill_agg(c->last_dtype, c->host->scan, out);
    return 0;
}

int voxel_gpu_scan(const voxel_gpu_scan_desc* d, void* stream, voxel_gpu_agg* out) {
    if (voxel_gpu_scan_launch(d, stream) != 0) return -1;
    return voxel_gpu_scan_result(stream, out);
}

// No runtime_ok()/ctx_for(): those allocate the per-stream scratch on first
// use.  The device is the caller's current one.
int voxel_gpu_scan_noalloc(const voxel_gpu_scan_desc* d, voxel_gpu_agg* out_dev, void* stream) {
    set_error("%s", "");
    ScanArgs g;
    if (!make_scan_args(d, g)) return -1;
    if (!out_dev) { set_error("scan_noalloc: out_dev is null"); return -1; }
    if (d->blocks < 0 || (size_t)d->blocks > kNoAllocMaxParts) {
        set_error("scan_noalloc: %d workgroups requested, the static partial buffer holds at most %zu",
                  d->blocks, kNoAllocMaxParts);
        return -1;
    }
    if (!(g.layout == 0 || g.layout == 2 || g.layout == 3)) {
        set_error("scan_noalloc: layout %d combines in the last workgroup (needs a ticket); use layout 0, 2 or 3",
                  g.layout);
        return -1;
    }
    return hip_ok(launch_scan_noalloc(g, out_dev, static_cast<hipStream_t>(stream)), "scan_noalloc launch") ? 0 : -1;
}

}  // extern "C"
