// vx_lower.cpp -- bytecode recognition (see vx_lower.hpp).
#include "vx_lower.hpp"

#include <cmath>
#include <cstring>

#include "vx_runtime.hpp"

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

bool host_op(uint8_t op) {
    return op == OP_NOP || op == OP_MOV || op == OP_MOVR || (op >= 0x20 && op <= 0x2F) || (op >= 0x30 && op <= 0x36) ||
           (op >= 0x60 && op <= 0x6F);
}

int run_scalar(const std::vector<uint32_t>& words, uint64_t* regs) {
    auto f64 = [](uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; };
    auto bits = [](double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; };
    // The reference's x86 scalar arithmetic returns the first NaN operand,
    // quieted, when an operand is NaN.
    auto fop = [&](uint8_t op, uint64_t a, uint64_t b) -> uint64_t {
        const double fa = f64(a), fb = f64(b);
        if (std::isnan(fa)) return a | 0x0008000000000000ull;
        if (std::isnan(fb)) return b | 0x0008000000000000ull;
        return bits(op == 0x2A ? fa + fb : op == 0x2B ? fa - fb : op == 0x2C ? fa * fb : fa / fb);
    };
    for (uint32_t w : words) {
        const uint8_t op = op_of(w), rd = rd_of(w), ra = ra_of(w), rb = rb_of(w);
        const uint64_t a = regs[ra], b = regs[rb];
        const int64_t sa = (int64_t)a, sb = (int64_t)b;
        switch (op) {
        case OP_NOP: break;
        case OP_MOV: regs[rd] = (uint64_t)(int64_t)simm_of(w); break;       // HandleMOV: sign-extended imm12
        case OP_MOVR: regs[rd] = a; break;
        case 0x20: regs[rd] = a + b; break;                                  // ADD (wrapping i64)
        case 0x21: regs[rd] = a - b; break;                                  // SUB
        case 0x22: regs[rd] = a * b; break;                                  // MUL
        case 0x23:                                                           // DIV
        case 0x24:                                                           // MOD
            if (sb == 0) return 0;
            if (sa == INT64_MIN && sb == -1) { set_error("%s", "integer division INT64_MIN / -1 traps in the reference"); return -1; }
            regs[rd] = (uint64_t)(op == 0x23 ? sa / sb : sa % sb);
            break;
        case 0x25: regs[rd] = 0 - a; break;                                  // NEG
        case 0x26: regs[rd] = sa < 0 ? 0 - a : a; break;                     // ABS
        case 0x27: regs[rd] = (uint64_t)(sa < sb ? sa : sb); break;          // MIN
        case 0x28: regs[rd] = (uint64_t)(sa > sb ? sa : sb); break;          // MAX
        case 0x29:                                                           // AVG
            regs[rd] = (uint64_t)((sa >> 1) + (sb >> 1) + ((sa & 1) + (sb & 1)) / 2);
            break;
        case 0x2A: case 0x2B: case 0x2C: regs[rd] = fop(op, a, b); break;   // ADDF SUBF MULF
        case 0x2D:                                                           // DIVF
            if (f64(b) == 0.0) return 0;
            regs[rd] = fop(op, a, b);
            break;
        case 0x2E: regs[rd] = a ^ 0x8000000000000000ull; break;              // NEGF: sign flip
        case 0x2F: regs[rd] = a & 0x7FFFFFFFFFFFFFFFull; break;              // ABSF: sign clear
        case 0x30: regs[rd] = a & b; break;                                  // AND
        case 0x31: regs[rd] = a | b; break;                                  // OR
        case 0x32: regs[rd] = a ^ b; break;                                  // XOR
        case 0x33: regs[rd] = ~a; break;                                     // NOT
        case 0x34: regs[rd] = a << (b & 0x3F); break;                        // SHL
        case 0x35: regs[rd] = a >> (b & 0x3F); break;                        // SHR
        case 0x36: regs[rd] = (uint64_t)(sa >> (b & 0x3F)); break;           // SAR
        // conversions (engine.hpp:1822-1937): sign-extending / masking
        // integer widths, f32 rounding, bit copies, and trunc / round /
        // ceil / floor of the register read as f64
        case 0x60: regs[rd] = (uint64_t)(int64_t)(int8_t)(uint8_t)a; break;  // CVT_I8
        case 0x61: regs[rd] = (uint64_t)(int64_t)(int16_t)(uint16_t)a; break;// CVT_I16
        case 0x62: regs[rd] = (uint64_t)(int64_t)(int32_t)(uint32_t)a; break;// CVT_I32
        case 0x64: regs[rd] = bits((double)(float)f64(a)); break;            // CVT_F32
        case 0x66: regs[rd] = a & 0xFF; break;                               // CVT_U8
        case 0x67: regs[rd] = a & 0xFFFF; break;                             // CVT_U16
        case 0x68: regs[rd] = a & 0xFFFFFFFFull; break;                      // CVT_U32
        case 0x63: case 0x65: case 0x69: case 0x6A: case 0x6B:               // CVT_I64 CVT_F64 CVT_U64
            regs[rd] = a;                                                    // BITCAST REINTERPRET
            break;
        case 0x6C: {                                                         // TRUNC: x86 cvttsd2si
            const double t = std::trunc(f64(a));
            regs[rd] = (t != t || t >= 9223372036854775808.0 || t < -9223372036854775808.0)
                           ? 0x8000000000000000ull : (uint64_t)(int64_t)t;
            break;
        }
        case 0x6D: regs[rd] = bits(std::round(f64(a))); break;               // ROUND
        case 0x6E: regs[rd] = bits(std::ceil(f64(a))); break;                // CEIL
        case 0x6F: regs[rd] = bits(std::floor(f64(a))); break;               // FLOOR
        default: set_error("%s", "word is not a scalar instruction"); return -1;
        }
    }
    return 1;
}

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

    // Straight-line program of scalar words, segment aggregates and the
    // hash / sort / join words: the interpreter runs it word by word to HALT
    // (engine.hpp:76-80).
    {
        size_t k = 0;
        bool any_agg = false;
        for (; k < len && op_of(c[k]) != OP_HALT; ++k) {
            const uint8_t op = op_of(c[k]);
            if (agg_op(op)) {
                any_agg = true;
                continue;
            }
            if (rel_word_op(op)) {
                any_agg = true;
                continue;
            }
            if (!host_op(op)) break;
        }
        if (any_agg && (k == len || op_of(c[k]) == OP_HALT)) {
            p.kind = Plan::STRAIGHT;
            p.words.assign(c, c + k);
            return p;
        }
    }

    // Tier-0 loop: prefix; VLOAD; then filters (each reading the loaded
    // vector or a filter's output, at most two deep) and reduction blocks
    // (VSUM|VCOUNT|VRED_* of any of those vectors, then ADDF|ADD or
    // CMP/JLE|JGE/MOVR into an accumulator); ADD off; CMP off,tot; JNZ; [HALT]
    size_t i = 0;
    while (i < len && host_op(op_of(c[i]))) p.prefix.push_back(c[i++]);
    const size_t L = i;
    auto window_kind = [](uint8_t op) -> uint8_t {
        switch (op) {
        case 0xDE: return 1;                               // WINDOW_SUM
        case 0xDF: return 2;                               // WINDOW_MEAN
        case 0x97: return 3;                               // WSTD
        case 0x98: return 4;                               // WVARIANCE_W
        case 0xDD: return 5;                               // WDELTA
        case 0x99: return 6;                               // WQUANTILE
        default: return 0;
        }
    };
    if (i >= len || (op_of(c[i]) != OP_VLOAD && !window_kind(op_of(c[i]))))
        return none("no VLOAD loop: program is not a lowerable scan loop");
    uint32_t vl = c[i++];
    const uint8_t vA = rd_of(vl), rOff = ra_of(vl);
    p.seg = (imm_of(vl) >> 8) & 0xF;
    p.vcount = op_of(vl) == OP_VLOAD ? imm_of(vl) & 0xFF : 0;
    p.off = rOff;
    p.wkind = window_kind(op_of(vl));
    if (p.wkind) {
        p.window = p.wkind == 5 ? 0 : imm_of(vl) & 0xFF;
        if (p.wkind == 6) {                                // engine.hpp:2642-2664: 5-bit window (0 -> 16),
            p.window = imm_of(vl) & 0x1F;                  // 3-bit quantile q, element (q * w) / 7
            if (p.window == 0) p.window = 16;
            const int q = (imm_of(vl) >> 5) & 0x7, idx = (q * p.window) / 7;
            p.wq = (uint8_t)(idx < p.window ? idx : p.window - 1);
        }
        if (p.wkind != 5 && (p.window == 0 || (p.wkind >= 3 && p.window < 2)))
            return none("the window word is a no-op for this window (the reference loop would not end)");
        p.has_arith = true;                                // lanes evaluated row by row
    }

    // vector register -> filters applied to it (-1: not defined in the loop,
    // -2: computed by vector arithmetic)
    int chain_len[16];
    uint8_t chain[16][2];
    for (int v = 0; v < 16; ++v) chain_len[v] = -1;
    chain_len[vA] = 0;
    p.vload_dst = vA;
    bool chain_ok = true;
    bool mask_set[8] = {};                     // mask registers written earlier in the body
    auto defined = [&](uint8_t v) { return chain_len[v] != -1; };
    auto vec_arith = [](uint8_t op) {          // lowered vector words (engine.hpp:2221-2450, 2473-2590, 2671-2865)
        return op == 0x78 || (op >= 0x80 && op <= 0x8F) || (op >= 0x90 && op <= 0x96) ||
               (op >= 0xA0 && op <= 0xA7) || (op >= 0xB0 && op <= 0xB7);
    };
    // lane words (engine.hpp:1987-2006, 2124-2217, 2926-2936): they move
    // lanes inside one kL group, so the per-row evaluation reads the group's
    // other rows
    auto lane_word = [](uint8_t op) { return op == 0x72 || (op >= 0x7A && op <= 0x7F) || op == 0xC7; };
    while (i < len) {
        const uint32_t w = c[i];
        const uint8_t op = op_of(w);
        if (op >= OP_VFILTER && op <= OP_VFILTER_GE) {
            const uint8_t src = ra_of(w), dst = rd_of(w);
            if (!defined(src)) return none("VFILTER reads a vector the loop does not define");
            if (defined(dst)) return none("VFILTER overwrites a vector the loop already defines");
            if (p.vops.size() >= 16) return none("more than sixteen vector words in the loop");
            Plan::Filt f;
            if (op == OP_VFILTER) {
                int mode = imm_of(w) & 0x7;        // HandleVFILTER, engine.hpp:2870-2894
                f.cmp = mode <= 5 ? mode : 7;
            } else {
                f.cmp = cmp_of_filter_op(op);
            }
            f.thr = rb_of(w);
            if (chain_len[src] >= 0 && chain_len[src] < 2 && p.filts.size() < 4) {
                chain_len[dst] = chain_len[src] + 1;
                for (int k = 0; k < chain_len[src]; ++k) chain[dst][k] = chain[src][k];
                chain[dst][chain_len[src]] = (uint8_t)p.filts.size();
            } else {
                if (chain_len[src] >= 0) chain_ok = false;   // too deep for one range predicate
                chain_len[dst] = -2;
            }
            p.filts.push_back(f);
            p.vops.push_back(Plan::VOp{op, dst, src, 0, f.thr, f.cmp, (uint16_t)(imm_of(w) & 0x7)});
            mask_set[imm_of(w) & 0x7] = true;                  // VFilterImpl / HandleVFILTER: Mask(imm12 & 7)
            ++i;
            continue;
        }
        if (op == 0x71) {                                                   // VSTORE
            const uint8_t src = rd_of(w), ro = ra_of(w);
            const uint16_t imm = imm_of(w);
            if (!defined(src)) return none("VSTORE stores a vector the loop does not define");
            if (ro != rOff) return none("VSTORE at an offset other than the loop's offset register is not lowered");
            const uint8_t cnt = (imm & 0xFF) ? (uint8_t)(imm & 0xFF) : (uint8_t)4;
            if (cnt > 4)
                return none("VSTORE of more than kL lanes (the reference reads past the vector register into the "
                            "next ones) is not lowered");
            p.stores.push_back(Plan::Store{src, (uint8_t)((imm >> 8) & 0xF), cnt});
            p.has_arith = true;
            ++i;
            continue;
        }
        if (lane_word(op)) {
            const uint8_t dst = rd_of(w), a = ra_of(w), b = rb_of(w);
            const uint16_t imm = imm_of(w);
            const bool two = op == 0x7B || op == 0x7C || op == 0xC7;          // VPERMUTE VSHUFFLE VBLEND
            if (!defined(a) || (two && !defined(b))) return none("a lane word reads a vector the loop does not define");
            if (defined(dst)) return none("a lane word overwrites a vector the loop already defines");
            if (p.vops.size() >= 16) return none("more than sixteen vector words in the loop");
            if (op == 0xC7 && !mask_set[imm & 0x7])
                return none("VBLEND reads a mask register no earlier word of the iteration writes (the value would "
                            "come from the previous iteration)");
            int gslot = 0;
            if (op == 0x72) {
                const uint8_t cnt = imm & 0xFF, sg = (imm >> 8) & 0xF;
                if (cnt != 0 && cnt < 4)
                    return none("VGATHER of fewer than kL lanes keeps the other lanes from the previous iteration");
                if (p.gsegs[0] == 0xFF || p.gsegs[0] == sg) p.gsegs[0] = sg;
                else if (p.gsegs[1] == 0xFF || p.gsegs[1] == sg) { p.gsegs[1] = sg; gslot = 1; }
                else return none("VGATHER words of one loop read more than two segments");
            }
            const uint8_t sreg = (op == 0x7F || op == 0x7A) ? b : (uint8_t)0xFF;   // VSLIDE shift, VINSERT value
            p.vops.push_back(Plan::VOp{op, dst, a, (uint8_t)(two ? b : 0), sreg, gslot, imm});
            chain_len[dst] = -2;
            p.has_arith = true;
            ++i;
            continue;
        }
        if (vec_arith(op)) {
            const uint8_t dst = rd_of(w), a = ra_of(w), b = rb_of(w);
            const bool scal = op >= 0x90 && op <= 0x96;
            const bool unary = op == 0x85 || op == 0x86 || (op >= 0x8C && op <= 0x8E) || op == 0xA6 || op == 0xA7 ||
                               op == 0xB3;
            const bool fma = op == 0x8A || op == 0x8B;         // c = R[imm12 & 0xF]
            const bool splat = op == 0x78;                      // VSPLAT: every lane = R[ra] (engine.hpp:2106-2112)
            if ((!splat && !defined(a)) || (!scal && !unary && !splat && !defined(b)))
                return none("vector arithmetic reads a vector the loop does not define");
            if (defined(dst)) return none("vector arithmetic overwrites a vector the loop already defines");
            if (p.vops.size() >= 16) return none("more than sixteen vector words in the loop");
            const uint8_t sreg = scal ? b : fma ? (uint8_t)(imm_of(w) & 0xF) : splat ? a : (uint8_t)0xFF;
            p.vops.push_back(Plan::VOp{op, dst, (uint8_t)(splat ? 0 : a), (uint8_t)(scal || unary || splat ? 0 : b), sreg, 0,
                                       (uint16_t)(imm_of(w) & 0x7)});
            if (op >= 0xA0 && op <= 0xA5) mask_set[0] = true;                 // VCMPxx: Mask(0)
            if (op == 0xA6 || op == 0xA7) mask_set[imm_of(w) & 0x7] = true;   // VCMP(NOT)NULL: Mask(imm12 & 7)
            chain_len[dst] = -2;
            p.has_arith = true;
            ++i;
            continue;
        }
        uint8_t red;
        if (op == OP_VSUM) red = 0;
        else if (op == OP_VCOUNT) red = 1;
        else if (op == OP_VRED_MIN) red = 2;
        else if (op == OP_VRED_MAX) red = 3;
        else break;
        const uint8_t src = ra_of(w), rP = rd_of(w);
        if (!defined(src)) return none("reduction reads a vector the loop does not define");
        if (p.blocks.size() >= 4) return none("more than four reductions in the loop");
        Plan::Block b;
        b.red = red;
        b.part = rP;
        b.vreg = src;
        b.nf = chain_len[src] > 0 ? (uint8_t)chain_len[src] : 0;
        b.f[0] = chain_len[src] > 0 ? chain[src][0] : 0;
        b.f[1] = chain_len[src] > 1 ? chain[src][1] : 0;
        ++i;
        if (i >= len) return none("loop body ends early");
        const uint32_t aw = c[i];
        if (op_of(aw) == OP_ADDF || op_of(aw) == OP_ADD) {
            b.acc = rd_of(aw);
            bool ok = (ra_of(aw) == b.acc && rb_of(aw) == rP) || (rb_of(aw) == b.acc && ra_of(aw) == rP);
            if (!ok) return none("accumulate is not acc = acc + part");
            b.accop = op_of(aw) == OP_ADDF ? 0 : 1;
            i += 1;
        } else if (op_of(aw) == OP_CMP && i + 2 < len) {
            const uint32_t jw = c[i + 1], mw = c[i + 2];
            b.acc = ra_of(aw);
            bool ok = rb_of(aw) == rP && (op_of(jw) == OP_JLE || op_of(jw) == OP_JGE) && simm_of(jw) == 2 &&
                      op_of(mw) == OP_MOVR && rd_of(mw) == b.acc && ra_of(mw) == rP;
            if (!ok) return none("compare-select accumulate is not CMP acc,part; JLE|JGE +2; MOVR acc,part");
            b.accop = op_of(jw) == OP_JLE ? 2 : 3;
            i += 3;
        } else {
            return none("loop accumulate is not ADDF/ADD/CMP-select");
        }
        p.blocks.push_back(b);
    }
    if (p.blocks.empty() && p.stores.empty()) {
        if (i >= len) return none("loop body ends early");
        return none("loop reduction is not VSUM/VCOUNT/VRED_MIN/VRED_MAX");
    }
    for (const Plan::Store& st : p.stores)
        if (st.seg == p.gsegs[0] || st.seg == p.gsegs[1])
            return none("VGATHER reads a segment the loop's VSTORE writes (later iterations would read stored rows)");
    if (!p.has_arith && !chain_ok) return none("more than two chained VFILTERs");

    uint8_t rStep = 0xFF;
    if (!p.wkind) {
        if (i >= len) return none("loop control missing");
        const uint32_t adw = c[i];
        if (op_of(adw) != OP_ADD || rd_of(adw) != rOff) return none("loop does not advance the offset register");
        if (ra_of(adw) == rOff) rStep = rb_of(adw);
        else if (rb_of(adw) == rOff) rStep = ra_of(adw);
        else return none("loop does not advance the offset register");
        ++i;
    }
    if (i + 1 >= len) return none("loop control missing");
    uint32_t cw = c[i], jw = c[i + 1];
    if (op_of(cw) != OP_CMP || ra_of(cw) != rOff) return none("loop does not compare the offset");
    uint8_t rTot = rb_of(cw);
    if (op_of(jw) != OP_JNZ) return none("loop does not end in JNZ");
    long target = (long)(i + 1) + simm_of(jw);
    if (target < 0 || target > (long)L) return none("loop branch does not return to VLOAD");
    // registers the loop writes: every block's partial and accumulator, the offset
    bool wr[16] = {};
    wr[rOff] = true;
    for (const Plan::Block& b : p.blocks) {
        if (wr[b.part] || wr[b.acc] || b.part == b.acc) return none("loop registers overlap");
        wr[b.part] = wr[b.acc] = true;
    }
    // Words between the branch target and VLOAD run again each iteration;
    // they must be prefix moves that do not touch loop registers (so running
    // them again changes nothing), e.g. voxel_query's unfiltered loops that
    // branch back onto "MOV R4, lanes".
    for (long k = target; k < (long)L; ++k) {
        uint32_t w = c[k];
        uint8_t op = op_of(w);
        if (op != OP_MOV && op != OP_MOVR && op != OP_NOP) return none("loop re-runs a scalar instruction that is not a move");
        if (op == OP_MOV && wr[rd_of(w)]) return none("loop re-runs a move into a loop register");
        if (op == OP_MOVR && (wr[rd_of(w)] || wr[ra_of(w)] || rd_of(w) == ra_of(w)))
            return none("loop re-runs a register move that is not idempotent");
    }
    if (target < (long)L) {
        // MOVR rd, ra with ra written earlier in [target, L) is not idempotent
        for (long k = target; k < (long)L; ++k)
            for (long q = target; q < (long)L; ++q)
                if (op_of(c[k]) == OP_MOVR && op_of(c[q]) != OP_NOP && rd_of(c[q]) == ra_of(c[k]))
                    return none("loop re-runs a register move that is not idempotent");
    }
    i += 2;
    while (i < len && host_op(op_of(c[i]))) p.epilogue.push_back(c[i++]);
    if (i < len && op_of(c[i]) != OP_HALT) return none("instructions after the loop are not lowerable");
    p.step = rStep;
    p.tot = rTot;

    // Registers written by the loop must not be inputs of the loop.
    if ((rStep != 0xFF && wr[rStep]) || wr[rTot]) return none("loop overwrites its own bound, step or threshold");
    for (const Plan::Filt& f : p.filts)
        if (f.cmp != 7 && wr[f.thr]) return none("loop overwrites its own bound, step or threshold");
    for (const Plan::VOp& v : p.vops)
        if (v.sreg != 0xFF && v.cmp != 7 && wr[v.sreg]) return none("loop overwrites a scalar operand of its vector arithmetic");
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
