// vx_expr_abi.hpp -- the program a Tier-0 loop body's vector words lower to
// (vx_lower.cpp), as the evaluation kernels read it: compiled into the library
// and also handed, as text, to the run-time compiler (vx_jit.cpp), so both
// sides see one layout.  Plain types only.
#pragma once
#include <stdint.h>

namespace vx {

// Vector words of a Tier-0 loop body evaluated per row; the lane words
// (VINSERT..VSLIDE, VBLEND, VGATHER) read the other rows of the row's kL = 4
// group, which sit in the neighbouring threads.
struct ExprOp {
    uint8_t op, dst, a, b;           // opcode, vector registers
    int16_t cmp;                     // filters: kCmp code, 7 never
    uint16_t imm;                    // mask register of filters / compares, lane words' imm12
    uint64_t s;                      // scalar operand / threshold bits
};
struct ExprProg {
    ExprOp ops[16];
    int32_t nops, vload, target, pad;
    const void* gcol[2];             // rows VGATHER reads (device; the op's cmp picks one); null: a segment
    uint64_t gcount[2];              // the engine does not have
};

// Reductions fused into one evaluation pass (expr_reduce): n of them, the
// q-th reading vector register vreg[q] with red[q] 0 = VSUM, 1 = VCOUNT,
// 2 = VRED_MIN, 3 = VRED_MAX.
struct ExprRed {
    int32_t n;
    uint8_t vreg[4], red[4];
};

// The opcodes apply_op handles by their own case; any other word in a
// program is a VFILTER_* (its kCmp in cmp).
#define VX_EXPR_CASES(X) \
    X(0x7D) X(0x7E) X(0x7F) X(0x7C) X(0x7B) X(0x7A) X(0xC7) X(0x72) \
    X(0x8F) X(0x80) X(0x81) X(0x82) X(0x83) X(0x84) X(0x85) X(0x89) \
    X(0x8A) X(0x8B) X(0x8C) X(0x8D) X(0x8E) X(0x86) X(0x87) X(0x88) \
    X(0x90) X(0x91) X(0x92) X(0x93) X(0x94) X(0x95) X(0x96) X(0x78) \
    X(0xA0) X(0xA1) X(0xA2) X(0xA3) X(0xA4) X(0xA5) X(0xA6) X(0xA7) \
    X(0xB0) X(0xB1) X(0xB2) X(0xB3) X(0xB4) X(0xB5) X(0xB6) X(0xB7)

}  // namespace vx
