// vx_lower.hpp -- recognises the bytecode loops the reference runs natively
// and turns them into GPU plans.  Host C++ only.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace vx {

// Instruction fields, bytecode/instruction.hpp:19-56:
// [op:8][rd:4][ra:4][rb:4][imm12:12].
inline uint8_t op_of(uint32_t w) { return w & 0xFF; }
inline uint8_t rd_of(uint32_t w) { return (w >> 8) & 0xF; }
inline uint8_t ra_of(uint32_t w) { return (w >> 12) & 0xF; }
inline uint8_t rb_of(uint32_t w) { return (w >> 16) & 0xF; }
inline uint16_t imm_of(uint32_t w) { return (w >> 20) & 0xFFF; }
inline int16_t simm_of(uint32_t w) {
    uint16_t v = imm_of(w);
    return (v & 0x800) ? (int16_t)(v | 0xF000) : (int16_t)v;
}

enum Op : uint8_t {
    OP_NOP = 0x00, OP_HALT = 0x01, OP_MOV = 0x10, OP_MOVR = 0x11, OP_ADD = 0x20, OP_ADDF = 0x2A,
    OP_CMP = 0x40, OP_JNZ = 0x52, OP_JLE = 0x5A, OP_JGE = 0x5C, OP_VLOAD = 0x70,
    OP_VFILTER = 0xC0, OP_VFILTER_EQ = 0xC1, OP_VFILTER_GE = 0xC6,
    OP_VSUM = 0xD0, OP_VRED_MIN = 0xD5, OP_VRED_MAX = 0xD6, OP_VCOUNT = 0xD7,
};

// Dedicated filter opcode 0xC1..0xC6 -> kCmp code (EQ NE LT LE GT GE).
inline int cmp_of_filter_op(uint8_t op) {
    switch (op) {
    case 0xC1: return 0; case 0xC2: return 1; case 0xC3: return 2;
    case 0xC4: return 3; case 0xC5: return 4; case 0xC6: return 5;
    default: return 4;   // RunFastPath's default branch (engine.hpp:130)
    }
}

struct Plan {
    enum Kind { NONE, FAST, TIER0, STRAIGHT } kind = NONE;
    std::vector<uint32_t> words;     // STRAIGHT: scalar, AGG_* and hash / sort / join words up to HALT
    std::string why;                 // reason when NONE

    // FAST: Engine<T>::DetectFastPath (engine.hpp:84-107)
    uint8_t seg = 0, thr = 3, acc = 0, off = 1, tot = 2;
    uint8_t fop = 0xC5;

    // TIER0: prefix + loop (voxel_query.py:74-135,155-171 shapes, and their
    // generalisation: up to two chained VFILTERs, up to four reductions)
    std::vector<uint32_t> prefix;    // scalar words run before the loop (host_op)
    std::vector<uint32_t> epilogue;  // scalar words run after it, up to HALT
    uint8_t vcount = 0;              // VLOAD count field
    uint8_t step = 4;
    struct Filt {
        int cmp;                     // kCmp code, 7 never
        uint8_t thr;                 // threshold register
    };
    struct Block {                   // reduction of one vector + its accumulate
        uint8_t red;                 // 0 VSUM, 1 VCOUNT, 2 VRED_MIN, 3 VRED_MAX
        uint8_t accop;               // 0 ADDF, 1 ADD, 2 keep-smaller (JLE), 3 keep-larger (JGE)
        uint8_t part, acc;           // partial and accumulator registers
        uint8_t nf;                  // filters applied to the reduced vector (0..2) ...
        uint8_t f[2];                // ... as indices into filts, in program order
        uint8_t vreg;                // the reduced vector register
    };
    std::vector<Filt> filts;
    std::vector<Block> blocks;
    // Loops with vector arithmetic: every vector word of the body in program
    // order (arithmetic and filters), evaluated per row into the reduced
    // registers' values.
    struct VOp {
        uint8_t op, dst, a, b;       // opcode, vector registers
        uint8_t sreg;                // scalar operand / threshold register (0xFF none)
        int cmp;                     // filters: kCmp code, 7 never
        uint16_t imm = 0;            // imm12: mask register of filters / compares, VROTATE shift,
                                     // VSHUFFLE mode, VINSERT lane, VBLEND mask, VGATHER segment << 8 | count
    };
    bool has_arith = false;
    uint8_t vload_dst = 0;
    std::vector<VOp> vops;
    uint8_t gsegs[2] = {0xFF, 0xFF}; // segments VGATHER words read (a VGATHER's cmp = its index here)
    // VSTOREs of the body (engine.hpp:1971-1985), in program order: each
    // iteration writes lanes [0, count) of `vreg` to rows offset + lane of
    // segment `seg`; the lowering admits them at the loop's own offset only
    // (no iteration reads rows another one wrote).
    struct Store {
        uint8_t vreg, seg, count;
    };
    std::vector<Store> stores;
    // Loops whose load is a window word (engine.hpp:2589-2640, 3167-3226):
    // 1 WINDOW_SUM, 2 WINDOW_MEAN, 3 WSTD, 4 WVARIANCE_W, 5 WDELTA (0: VLOAD).
    // The word advances the offset itself; the loop has no ADD of it.
    uint8_t wkind = 0;
    uint8_t wq = 0;                 // WQUANTILE: order-statistic index in its window
    uint8_t window = 0;
};

// Interpreter view: fast path first (exactly DetectFastPath), then the
// Tier-0 loop shapes.
Plan lower_engine_program(const uint32_t* code, size_t len);

// JIT view (codegen/x86_64.hpp:1600-1631): the fused block may start after
// a prefix of MOV/MOVR/NOP words; only the f64 filter+sum loop fuses.
struct JitPlan {
    bool ok = false;
    std::string why;
    std::vector<uint32_t> prefix;
    uint8_t seg = 0, thr = 3, acc = 0, off = 1;
    uint8_t fop = 0xC5;
};
JitPlan lower_jit_program(const uint32_t* code, size_t len);

// Applies MOV/MOVR/NOP words to a register file (HandleMOV/HandleMOVR,
// engine.hpp:1167-1178).
void run_prefix(const std::vector<uint32_t>& prefix, uint64_t* regs);

// Scalar words a lowered Tier-0 program may run before or after its loop:
// NOP, MOV, MOVR, ADD SUB MUL DIV MOD NEG ABS MIN MAX AVG, ADDF SUBF MULF
// DIVF NEGF ABSF, AND OR XOR NOT SHL SHR SAR (engine.hpp:1167-1178,
// 1286-1505; opcodes.hpp:24-60).  Condition flags are not kept: no lowered
// program reads them outside its loop.
bool host_op(uint8_t op);

// Segment aggregate words AGG_COUNT..AGG_PERCENTILE (opcodes.hpp:228-241).
inline bool agg_op(uint8_t op) { return op >= 0xE0 && op <= 0xED; }

// Hash / sort / join words of straight-line programs: HASH_INIT, HASH_PROBE,
// HASH_BUILD, HASH_LOOKUP, SORT_ASC, SORT_DESC, SORT_TOPK, SORT_BOTTOMK,
// JOIN_HASH, JOIN_MERGE, JOIN_NESTED, JOIN_ANTI, JOIN_SEMI, PARTITION_HASH
// (opcodes.hpp:242-259).
inline bool rel_word_op(uint8_t op) { return op == 0xEE || op == 0xEF || (op >= 0xF0 && op <= 0xFA) || op == 0xFD; }

// Runs host_op words in order on a register file.  Returns 1 when all ran,
// 0 when a DIV/MOD/DIVF by zero stopped the program (the reference clears
// Running_ there and runs nothing further), -1 (error set) on a division the
// reference's host would trap on (INT64_MIN / -1).
int run_scalar(const std::vector<uint32_t>& words, uint64_t* regs);

}  // namespace vx
