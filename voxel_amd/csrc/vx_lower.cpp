// vx_lower.cpp -- bytecode recognition (see vx_lower.hpp).
#include "vx_lower.hpp"

namespace vx {

namespace {
bool is_prefix_op(uint8_t op) { return op == OP_MOV || op == OP_MOVR || op == OP_NOP; }

Plan none(const char* why) {
    Plan p;
    p.kind = Plan::NONE;
    p.why = why;
    return p;
}
}  // namespace

void run_prefix(const std::vector<uint32_t>& prefix, uint64_t* regs) {
    for (uint32_t w : prefix) {
        uint8_t op = op_of(w);
        if (op == OP_MOV) regs[rd_of(w)] = (uint64_t)(int64_t)simm_of(w);
        else if (op == OP_MOVR) regs[rd_of(w)] = regs[ra_of(w)];
    }
}

// Engine<T>::DetectFastPath, exec/engine.hpp:84-107, restated field by field.
static bool detect_fast_path(const uint32_t* c, size_t len, Plan& p) {
    if (len < 8) return false;
    uint32_t r0 = c[0], r1 = c[1], r2 = c[2], r3 = c[3], r4 = c[4], r5 = c[5], r6 = c[6];
    uint8_t o1 = op_of(r1);
    bool filt = o1 >= 0xC1 && o1 <= 0xC6;
    bool shape = op_of(r0) == OP_VLOAD && filt && op_of(r2) == OP_VSUM && op_of(r3) == OP_ADDF &&
                 op_of(r4) == OP_ADD && op_of(r5) == OP_CMP && op_of(r6) == OP_JNZ;
    if (!shape) return false;
    bool chain = ra_of(r1) == rd_of(r0) && ra_of(r2) == rd_of(r1) && rb_of(r3) == rd_of(r2) &&
                 ra_of(r4) == ra_of(r0) && ra_of(r5) == ra_of(r0) && simm_of(r6) == -6;
    if (!chain) return false;
    p.kind = Plan::FAST;
    p.seg = (r0 >> 28) & 0xF;
    p.thr = rb_of(r1);
    p.acc = rd_of(r3);
    p.off = ra_of(r0);
    p.tot = rb_of(r5);
    p.fop = o1;
    return true;
}

Plan lower_engine_program(const uint32_t* c, size_t len) {
    Plan p;
    if (detect_fast_path(c, len, p)) return p;

    // Tier-0 loop: prefix; VLOAD; [VFILTER*]; reduce; acc update; ADD off; CMP off,tot; JNZ; [HALT]
    size_t i = 0;
    while (i < len && is_prefix_op(op_of(c[i]))) p.prefix.push_back(c[i++]);
    const size_t L = i;
    if (i >= len || op_of(c[i]) != OP_VLOAD) return none("no VLOAD loop: program is not a lowerable scan loop");
    uint32_t vl = c[i++];
    uint8_t vA = rd_of(vl), rOff = ra_of(vl);
    p.seg = (imm_of(vl) >> 8) & 0xF;
    p.vcount = imm_of(vl) & 0xFF;
    p.off = rOff;

    uint8_t vB = vA;
    p.cmp = 6;
    p.thr = 0xFF;
    if (i < len && op_of(c[i]) >= OP_VFILTER && op_of(c[i]) <= OP_VFILTER_GE) {
        uint32_t f = c[i++];
        if (ra_of(f) != vA) return none("VFILTER does not read the loaded vector");
        if (op_of(f) == OP_VFILTER) {
            int mode = imm_of(f) & 0x7;           // HandleVFILTER, engine.hpp:2870-2894
            p.cmp = mode <= 5 ? mode : 7;
        } else {
            p.cmp = cmp_of_filter_op(op_of(f));
        }
        vB = rd_of(f);
        p.thr = rb_of(f);
    }
    if (i >= len) return none("loop body ends early");
    uint32_t rw = c[i++];
    switch (op_of(rw)) {
    case OP_VSUM: p.red = 0; break;
    case OP_VCOUNT: p.red = 1; break;
    case OP_VRED_MIN: p.red = 2; break;
    case OP_VRED_MAX: p.red = 3; break;
    default: return none("loop reduction is not VSUM/VCOUNT/VRED_MIN/VRED_MAX");
    }
    if (ra_of(rw) != vB) return none("reduction does not read the filtered vector");
    uint8_t rP = rd_of(rw);
    p.part = rP;

    if (i >= len) return none("loop body ends early");
    uint32_t aw = c[i];
    uint8_t rAcc;
    if (op_of(aw) == OP_ADDF || op_of(aw) == OP_ADD) {
        rAcc = rd_of(aw);
        bool ok = (ra_of(aw) == rAcc && rb_of(aw) == rP) || (rb_of(aw) == rAcc && ra_of(aw) == rP);
        if (!ok) return none("accumulate is not acc = acc + part");
        p.accop = op_of(aw) == OP_ADDF ? 0 : 1;
        i += 1;
    } else if (op_of(aw) == OP_CMP && i + 2 < len) {
        uint32_t jw = c[i + 1], mw = c[i + 2];
        rAcc = ra_of(aw);
        bool ok = rb_of(aw) == rP && (op_of(jw) == OP_JLE || op_of(jw) == OP_JGE) && simm_of(jw) == 2 &&
                  op_of(mw) == OP_MOVR && rd_of(mw) == rAcc && ra_of(mw) == rP;
        if (!ok) return none("compare-select accumulate is not CMP acc,part; JLE|JGE +2; MOVR acc,part");
        p.accop = op_of(jw) == OP_JLE ? 2 : 3;
        i += 3;
    } else {
        return none("loop accumulate is not ADDF/ADD/CMP-select");
    }
    p.acc = rAcc;

    if (i + 2 >= len) return none("loop control missing");
    uint32_t adw = c[i], cw = c[i + 1], jw = c[i + 2];
    if (op_of(adw) != OP_ADD || rd_of(adw) != rOff) return none("loop does not advance the offset register");
    uint8_t rStep;
    if (ra_of(adw) == rOff) rStep = rb_of(adw);
    else if (rb_of(adw) == rOff) rStep = ra_of(adw);
    else return none("loop does not advance the offset register");
    if (op_of(cw) != OP_CMP || ra_of(cw) != rOff) return none("loop does not compare the offset");
    uint8_t rTot = rb_of(cw);
    if (op_of(jw) != OP_JNZ) return none("loop does not end in JNZ");
    long target = (long)(i + 2) + simm_of(jw);
    if (target < 0 || target > (long)L) return none("loop branch does not return to VLOAD");
    // Words between the branch target and VLOAD run again each iteration;
    // they must be prefix moves that do not touch loop registers (so running
    // them again changes nothing), e.g. voxel_query's unfiltered loops that
    // branch back onto "MOV R4, lanes".
    for (long k = target; k < (long)L; ++k) {
        uint32_t w = c[k];
        uint8_t op = op_of(w);
        auto loopreg = [&](uint8_t r) { return r == rP || r == rAcc || r == rOff; };
        if (op == OP_MOV && loopreg(rd_of(w))) return none("loop re-runs a move into a loop register");
        if (op == OP_MOVR && (loopreg(rd_of(w)) || loopreg(ra_of(w)) || rd_of(w) == ra_of(w)))
            return none("loop re-runs a register move that is not idempotent");
    }
    if (target < (long)L) {
        // MOVR rd, ra with ra written earlier in [target, L) is not idempotent
        for (long k = target; k < (long)L; ++k)
            for (long q = target; q < (long)L; ++q)
                if (op_of(c[k]) == OP_MOVR && op_of(c[q]) != OP_NOP && rd_of(c[q]) == ra_of(c[k]))
                    return none("loop re-runs a register move that is not idempotent");
    }
    i += 3;
    if (i < len && op_of(c[i]) != OP_HALT) return none("instructions after the loop are not lowerable");
    p.step = rStep;
    p.tot = rTot;

    // Registers written by the loop must not be inputs of the loop.
    auto written = [&](uint8_t r) { return r == rP || r == rAcc || r == rOff; };
    if (rP == rAcc || rAcc == rOff || rP == rOff) return none("loop registers overlap");
    if (written(rStep) || written(rTot) || (p.thr != 0xFF && written(p.thr)))
        return none("loop overwrites its own bound, step or threshold");
    p.kind = Plan::TIER0;
    return p;
}

// X64Compiler fusion detection, codegen/x86_64.hpp:1600-1631, for the program
// forms the GPU lowering accepts: prefix block, fused 7-word block ending in a
// backward JNZ to its own start, then HALT or end of program.
JitPlan lower_jit_program(const uint32_t* c, size_t len) {
    JitPlan jp;
    size_t i = 0;
    while (i < len && is_prefix_op(op_of(c[i]))) jp.prefix.push_back(c[i++]);
    if (len - i < 7) { jp.why = "no fusable VLOAD/VFILTER/VSUM/ADDF/ADD/CMP/JNZ block"; return jp; }
    const uint32_t* b = c + i;
    uint8_t o1 = op_of(b[1]);
    bool filt = o1 >= 0xC1 && o1 <= 0xC6;
    bool shape = op_of(b[0]) == OP_VLOAD && filt && op_of(b[2]) == OP_VSUM && op_of(b[3]) == OP_ADDF &&
                 op_of(b[4]) == OP_ADD && op_of(b[5]) == OP_CMP;
    bool chain = ra_of(b[1]) == rd_of(b[0]) && ra_of(b[2]) == rd_of(b[1]) && rb_of(b[3]) == rd_of(b[2]) &&
                 ra_of(b[4]) == ra_of(b[0]) && ra_of(b[5]) == ra_of(b[0]);
    if (!shape || !chain) { jp.why = "no fusable VLOAD/VFILTER/VSUM/ADDF/ADD/CMP/JNZ block"; return jp; }
    if (op_of(b[6]) != OP_JNZ || simm_of(b[6]) != -6) { jp.why = "fused block does not end in JNZ -6"; return jp; }
    if (len - i > 7 && op_of(b[7]) != OP_HALT) { jp.why = "instructions after the fused loop are not lowerable"; return jp; }
    if (((b[0] >> 28) & 0xF) != 0) { jp.why = "voxel_jit_run passes one segment; VLOAD names another"; return jp; }
    jp.seg = 0;
    jp.off = ra_of(b[0]);
    jp.thr = rb_of(b[1]);
    jp.acc = rd_of(b[3]);
    jp.fop = o1;
    jp.ok = true;
    return jp;
}

}  // namespace vx
