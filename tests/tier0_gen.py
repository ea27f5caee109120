"""Random Tier-0 loop programs with several filters and reductions, for the
oracle pinning test (against the reference engine in oracle/_ref) and the GPU
parity test.  Words follow bytecode/instruction.hpp:19-56
([op:8][rd:4][ra:4][rb:4][imm12:12]) with the opcode values of
bytecode/opcodes.hpp.

Registers: R1 offset, R2 total, R4 step (= 4 lanes, set by the program),
R3 / R6 filter thresholds, partials R5 R7 R9 R11, accumulators R0 R8 R10 R12.
"""
from __future__ import annotations

import numpy as np

FILTER_OPS = {0: 0xC1, 1: 0xC2, 2: 0xC3, 3: 0xC4, 4: 0xC5, 5: 0xC6}   # kCmp code -> VFILTER_xx
RED_OPS = {0: 0xD0, 1: 0xD7, 2: 0xD5, 3: 0xD6}                      # VSUM VCOUNT VRED_MIN VRED_MAX
PARTS = (5, 7, 9, 11)
ACCS = (0, 8, 10, 12)
THR_REGS = (3, 6)


def w(op, rd=0, ra=0, rb=0, imm=0):
    return (op & 0xFF) | (rd & 0xF) << 8 | (ra & 0xF) << 12 | (rb & 0xF) << 16 | (imm & 0xFFF) << 20


SCALAR_OPS = [0x00, 0x10, 0x11] + list(range(0x20, 0x30)) + list(range(0x30, 0x37)) + list(range(0x60, 0x70))
SCRATCH = (13, 14, 15)


def scalar_words(rng, k, reads, writes):
    """k random scalar words (NOP MOV MOVR, integer, f64 and bitwise
    arithmetic, conversions) reading registers `reads` and writing `writes`."""
    out = []
    for _ in range(k):
        op = int(rng.choice(SCALAR_OPS))
        imm = int(rng.integers(-2048, 2048)) if op == 0x10 else 0
        out.append(w(op, rd=int(rng.choice(writes)), ra=int(rng.choice(reads)), rb=int(rng.choice(reads)), imm=imm))
    return out


def make_program(rng, is_f64=True, allow_ne=False, reds=(0, 1, 2, 3), prefix=(), epilogue=()):
    """Returns (words, spec) where spec describes the filters and blocks in
    the oracle's terms (vo_tier0_multi_*); prefix / epilogue: scalar words
    placed before the loop / between the loop and HALT."""
    code = list(prefix) + [w(0x10, rd=4, imm=4)]        # ...; MOV R4, 4
    vload = len(code)
    code.append(w(0x70, rd=0, ra=1, imm=0))              # VLOAD V0, R1, segment 0, all lanes
    nf = int(rng.integers(0, 3))
    fcmp, fsrc = [], []
    for j in range(nf):
        cmp = int(rng.choice([2, 3, 4, 5] + ([1] if allow_ne else [])))
        src = -1 if j == 0 or rng.random() < 0.3 else 0            # chain onto filter 0 or read V0
        fcmp.append(cmp)
        fsrc.append(src)
        code.append(w(FILTER_OPS[cmp], rd=1 + j, ra=0 if src < 0 else 1 + src, rb=THR_REGS[j]))
    nb = int(rng.integers(1, 4))
    bsrc, reds_used, accops = [], [], []
    for b in range(nb):
        src = int(rng.integers(-1, nf)) if nf else -1
        red = int(rng.choice(reds))
        vec = 0 if src < 0 else 1 + src
        code.append(w(RED_OPS[red], rd=PARTS[b], ra=vec))
        if red == 0:
            accop = 0 if is_f64 else 1
            code.append(w(0x2A if is_f64 else 0x20, rd=ACCS[b], ra=ACCS[b], rb=PARTS[b]))
        elif red == 1:
            accop = 1
            code.append(w(0x20, rd=ACCS[b], ra=ACCS[b], rb=PARTS[b]))
        else:
            accop = 2 if red == 2 else 3
            code.append(w(0x40, ra=ACCS[b], rb=PARTS[b]))                       # CMP acc, part
            code.append(w(0x5A if red == 2 else 0x5C, imm=2))                    # JLE|JGE +2
            code.append(w(0x11, rd=ACCS[b], ra=PARTS[b]))                        # MOVR acc, part
        bsrc.append(src)
        reds_used.append(red)
        accops.append(accop)
    code.append(w(0x20, rd=1, ra=1, rb=4))               # ADD R1, R1, R4
    code.append(w(0x40, ra=1, rb=2))                     # CMP R1, R2
    jnz = len(code)
    code.append(w(0x52, imm=(vload - jnz) & 0xFFF))      # JNZ back to VLOAD
    code += list(epilogue)
    code.append(w(0x01))                                 # HALT
    spec = {"fcmp": fcmp, "fsrc": fsrc, "bsrc": bsrc, "red": reds_used, "accop": accops}
    return np.array(code, np.uint32), spec


def oracle_run(o, x, total, spec, thr, is_f64=True, acc0=None):
    """vo_tier0_multi_* on the spec; returns (acc bits, part bits) per block
    (acc0: the accumulators' values before the loop, default 0)."""
    nf, nb = len(spec["fcmp"]), len(spec["red"])
    fcmp = np.array(spec["fcmp"] + [0], np.int32)
    fsrc = np.array(spec["fsrc"] + [0], np.int32)
    fthr = np.array(list(thr[:nf]) + [0], np.float64 if is_f64 else np.int64)
    bsrc = np.array(spec["bsrc"], np.int32)
    red = np.array(spec["red"], np.int32)
    accop = np.array(spec["accop"], np.int32)
    acc = np.array(acc0 if acc0 is not None else [0] * nb, np.uint64)
    part = np.zeros(nb, np.uint64)
    fn = o.O.vo_tier0_multi_f64 if is_f64 else o.O.vo_tier0_multi_i64
    fn(o.ptr(x), x.size, 0, total, 0, nf, o.ptr(fcmp), o.ptr(fthr), o.ptr(fsrc), nb, o.ptr(bsrc), o.ptr(red),
       o.ptr(accop), o.ptr(acc), o.ptr(part))
    return [int(a) for a in acc], [int(p) for p in part]


def run_parallel_expected(o, x, total, spec, thr, threads, rr_block, use_acc, is_f64=True):
    """RunParallel of a Tier-0 loop restated (engine.hpp:143-175): worker t
    runs the loop on chunk t of ceil(N / threads) rows as its only segment
    (R[total] copied, rows past the chunk read as zeros), and R[result] is
    the sum in T of the workers' R[result] in worker order from zero; the
    per-chunk registers come from vo_tier0_multi_* (oracle_run)."""
    n = x.size
    chunk = -(-n // threads)
    fs, isum = 0.0, 0
    for t in range(threads):
        s0 = t * chunk
        if s0 >= n:
            continue
        xc = np.ascontiguousarray(x[s0:s0 + min(chunk, n - s0)])
        acc, part = oracle_run(o, xc, total, spec, thr, is_f64)
        v = acc[rr_block] if use_acc else part[rr_block]
        if is_f64:
            fs = fs + o.bits_f64(v)
        else:
            isum = (isum + v) % 2**64
    return o.f64_bits(fs) if is_f64 else isum


def mergeable(spec):
    """True when every block's filter chain is one range (no != in a chain of two)."""
    for b, src in enumerate(spec["bsrc"]):
        chain = []
        j = src
        while j >= 0:
            chain.append(spec["fcmp"][j])
            j = spec["fsrc"][j]
        if len(chain) == 2 and 1 in chain:
            return False
    return True


AGG_OPS = list(range(0xE0, 0xEA)) + [0xEB, 0xED]          # AGG_COUNT..COUNT_DISTINCT, MEDIAN, PERCENTILE
AGG_OPS_ALL = list(range(0xE0, 0xEE))                     # ... with SUM_DISTINCT and MODE


def straight_program(rng, n, k=8, ops=AGG_OPS, scalar_reads=tuple(range(16))):
    """Straight-line program: MOVs of start rows into R1..R3, then k words
    mixing AGG words (slots R4..R8, R9 for AGG_COUNT_DISTINCT, start register
    R1..R3, row counts 0..4095 with 0 = all, a few naming a missing segment)
    and scalar arithmetic on the scratch registers."""
    code = [w(0x10, rd=r, imm=int(rng.integers(0, min(n, 2047) + 1))) for r in (1, 2, 3)]
    for _ in range(k):
        if rng.random() < 0.75:
            op = int(rng.choice(ops))
            imm = 0 if rng.random() < 0.3 else int(rng.integers(1, 4096))
            seg = 0 if rng.random() < 0.95 else 1
            # the distinct set keeps slot 9 to itself (COUNT_DISTINCT and
            # SUM_DISTINCT share it, as they share the reference's set)
            slot = 9 if op in (0xE9, 0xEA) else int(rng.integers(4, 9))
            code.append(w(op, rd=slot, ra=seg, rb=int(rng.integers(1, 4)), imm=imm))
        else:
            code += scalar_words(rng, 1, scalar_reads, SCRATCH)
    code.append(w(0x01))
    return np.array(code, np.uint32)


def last_agg_writers(words):
    """register -> opcode of the last AGG word that writes it"""
    out = {}
    for x in words:
        op = int(x) & 0xFF
        if 0xE0 <= op <= 0xED and ((int(x) >> 12) & 0xF) == 0:
            out[(int(x) >> 8) & 0xF] = op
    return out


VEC_BIN = [0x80, 0x81, 0x82, 0x83, 0x84, 0x87, 0x88, 0x89,  # VADD VSUB VMUL VDIV VMOD VMIN VMAX VAVG
           0xA0, 0xA1, 0xA2, 0xA3, 0xA4, 0xA5,              # VCMPEQ .. VCMPGE
           0xB0, 0xB1, 0xB2, 0xB4, 0xB5, 0xB6, 0xB7]        # VAND VOR VXOR VANDN VSHL VSHR VSAR
VSPLAT = 0x78
VEC_FMA = [0x8A, 0x8B]                                      # VFMA VFMS: a * b +/- R[imm12 & 0xF]
VEC_UN = [0x85, 0x86, 0x8C, 0x8D, 0x8E, 0xA6, 0xA7, 0xB3]   # VNEG VABS VSQRT VRSQRT VRCP VCMPNULL VCMPNOTNULL VNOT
VEC_SCAL = [0x90, 0x91, 0x92, 0x93, 0x94, 0x95, 0x96]      # VSADD VSSUB VSMUL VSDIV VSMOD VSMIN VSMAX
SCAL_REGS = (3, 6, 13)                                      # scalar operands / thresholds


# Lane words (engine.hpp:1987-2006, 2124-2217, 2453-2466, 2926-2936): each
# entry (opcode, kind); the ops tuple's fifth field carries imm12.
LANE_OPS = [0x7A, 0x7B, 0x7C, 0x7D, 0x7E, 0x7F, 0xC7, 0x8F, 0x72]   # VINSERT VPERMUTE VSHUFFLE VREVERSE
                                                                    # VROTATE VSLIDE VBLEND VPOW VGATHER
SLIDE_REG = 14                                                      # VSLIDE's shift register (small integers)


def make_expr_program(rng, is_f64=True, nops=4, reds=(0, 1, 2, 3), last=None, lane_words=False, gather_seg=1,
                      store_segs=(), lane_ops=None):
    """Loop with vector arithmetic and filters on V0 = VLOAD and up to three
    reductions.  Returns (words, spec) for vo_tier0_expr_*.  `last`, when
    given, is the opcode of the final vector word, whose register the first
    reduction reads.  lane_words: the body also draws the lane words of
    LANE_OPS (or lane_ops; VGATHER reading segment gather_seg), and filters
    / compares write random mask registers for VBLEND to read.  store_segs:
    one or two VSTOREs at the loop offset into these segment ids after the
    vector words (then sometimes no reduction at all)."""
    code = [w(0x10, rd=4, imm=4)]
    vload = len(code)
    code.append(w(0x70, rd=0, ra=1, imm=0))
    defined = [0]
    ops = []
    masks = set()
    for k in range(nops):
        dst = k + 1
        force_lane = last in LANE_OPS and k == nops - 1
        force_other = last is not None and last not in LANE_OPS and k == nops - 1
        if lane_words and not force_other and (force_lane or rng.random() < 0.45):
            op = last if force_lane else int(rng.choice(lane_ops or LANE_OPS))
            if op == 0xC7 and not masks:
                op = 0x7D
            a = int(rng.choice(defined))
            b = int(rng.choice(defined))
            sreg = int(rng.choice(SCAL_REGS))
            if op == 0x7A:
                lane = int(rng.integers(0, 4))
                code.append(w(op, rd=dst, ra=a, rb=sreg, imm=lane))
                ops.append((op, dst, a, 0, lane, sreg))
            elif op in (0x7B, 0x8F):
                code.append(w(op, rd=dst, ra=a, rb=b))
                ops.append((op, dst, a, b, 0, -1))
            elif op == 0x7C:
                mode = int(rng.integers(0, 4))
                code.append(w(op, rd=dst, ra=a, rb=b, imm=mode))
                ops.append((op, dst, a, b, mode, -1))
            elif op == 0x7D:
                code.append(w(op, rd=dst, ra=a))
                ops.append((op, dst, a, 0, 0, -1))
            elif op == 0x7E:
                sh = int(rng.integers(0, 8))
                code.append(w(op, rd=dst, ra=a, imm=sh))
                ops.append((op, dst, a, 0, sh, -1))
            elif op == 0x7F:
                code.append(w(op, rd=dst, ra=a, rb=SLIDE_REG))
                ops.append((op, dst, a, 0, 0, SLIDE_REG))
            elif op == 0xC7:
                m = int(rng.choice(sorted(masks)))
                code.append(w(op, rd=dst, ra=a, rb=b, imm=m))
                ops.append((op, dst, a, b, m, -1))
            else:                                                   # VGATHER: all lanes, or a missing segment
                seg = gather_seg if rng.random() < 0.85 else 7
                imm = (seg << 8) | int(rng.choice([0, 4, 9]))
                code.append(w(op, rd=dst, ra=a, imm=imm))
                ops.append((op, dst, a, 0, imm, -1))
            defined.append(dst)
            continue
        kind = rng.random()
        a = int(rng.choice(defined))
        b = int(rng.choice(defined))
        sreg = int(rng.choice(SCAL_REGS))
        if last is not None and k == nops - 1:
            kind = 0.35 if last in VEC_UN else 0.41 if last == VSPLAT else 0.0
        if kind < 0.25:
            op = int(rng.choice(VEC_BIN)) if last is None or k < nops - 1 else last
            code.append(w(op, rd=dst, ra=a, rb=b))
            ops.append((op, dst, a, b, 0, -1))
        elif kind < 0.3:
            op = int(rng.choice(VEC_FMA))
            code.append(w(op, rd=dst, ra=a, rb=b, imm=sreg))
            ops.append((op, dst, a, b, 0, sreg))
        elif kind < 0.4:
            op = int(rng.choice(VEC_UN)) if last is None or k < nops - 1 else last
            code.append(w(op, rd=dst, ra=a))
            ops.append((op, dst, a, 0, 0, -1))
        elif kind < 0.43:
            op = VSPLAT                                       # every lane = R[ra]'s bits
            code.append(w(op, rd=dst, ra=sreg))
            ops.append((op, dst, 0, 0, 0, sreg))
        elif kind < 0.75:
            op = int(rng.choice(VEC_SCAL))
            code.append(w(op, rd=dst, ra=a, rb=sreg))
            ops.append((op, dst, a, 0, 0, sreg))
        else:
            cmp = int(rng.choice([1, 2, 3, 4, 5]))
            m = int(rng.integers(0, 8)) if lane_words else 0       # the mask register it writes
            code.append(w(FILTER_OPS[cmp], rd=dst, ra=a, rb=sreg, imm=m))
            ops.append((0xC0 | (cmp + 1), dst, a, 0, cmp | (m << 4), sreg))
            masks.add(m)
        if lane_words and 0xA0 <= ops[-1][0] <= 0xA7:
            masks.add(0)                                            # VCMPxx, VCMP(NOT)NULL imm 0: mask 0
        defined.append(dst)
    for _ in range(int(rng.integers(1, 3)) if store_segs else 0):
        src = int(rng.choice(defined))
        seg = int(rng.choice(store_segs))
        imm = (seg << 8) | int(rng.choice([0, 4, 1, 2, 3]))       # (more than 4 reads past the register)
        code.append(w(0x71, rd=src, ra=1, imm=imm))
        ops.append((0x71, 0, src, 0, imm, -1))
    nb = int(rng.integers(0 if store_segs else 1, 4))
    bvreg, reds_used, accops = [], [], []
    for b in range(nb):
        vreg = int(rng.choice(defined[1:] if len(defined) > 1 else defined))
        if last is not None and b == 0:
            vreg = defined[-1]
        red = int(rng.choice(reds))
        code.append(w(RED_OPS[red], rd=PARTS[b], ra=vreg))
        if red == 0:
            accop = 0 if is_f64 else 1
            code.append(w(0x2A if is_f64 else 0x20, rd=ACCS[b], ra=ACCS[b], rb=PARTS[b]))
        elif red == 1:
            accop = 1
            code.append(w(0x20, rd=ACCS[b], ra=ACCS[b], rb=PARTS[b]))
        else:
            accop = 2 if red == 2 else 3
            code.append(w(0x40, ra=ACCS[b], rb=PARTS[b]))
            code.append(w(0x5A if red == 2 else 0x5C, imm=2))
            code.append(w(0x11, rd=ACCS[b], ra=PARTS[b]))
        bvreg.append(vreg)
        reds_used.append(red)
        accops.append(accop)
    code.append(w(0x20, rd=1, ra=1, rb=4))
    code.append(w(0x40, ra=1, rb=2))
    jnz = len(code)
    code.append(w(0x52, imm=(vload - jnz) & 0xFFF))
    code.append(w(0x01))
    return np.array(code, np.uint32), {"ops": ops, "bvreg": bvreg, "red": reds_used, "accop": accops}


def _expr_words(spec, regs, is_f64):
    ops = spec["ops"]
    code = np.zeros(5 * max(len(ops), 1), np.int32)
    dt = np.float64 if is_f64 else np.int64
    sval = np.zeros(max(len(ops), 1), dt)
    for k, (op, dst, a, b, cmp, sreg) in enumerate(ops):
        code[5 * k:5 * k + 5] = (op, dst, a, b, cmp)
        if sreg >= 0:
            sval[k] = np.uint64(regs[sreg]).view(dt)
    return code, sval


def oracle_expr_col(o, x, total, spec, regs, target, is_f64=True):
    """Lanes of vector register `target` over rows [0, total) (vo_tier0_expr_col_*)."""
    code, sval = _expr_words(spec, regs, is_f64)
    out = np.zeros(total, np.float64 if is_f64 else np.int64)
    fn = o.O.vo_tier0_expr_col_f64 if is_f64 else o.O.vo_tier0_expr_col_i64
    fn(o.ptr(x), x.size, 0, total, 0, len(spec["ops"]), o.ptr(code), o.ptr(sval), target, o.ptr(out))
    return out


def expr_traps(o, x, total, spec, regs):
    """True when an i64 VDIV/VMOD/VSDIV/VSMOD lane divides INT64_MIN by -1
    (a trap in the reference's x86 build; the product refuses such a run)."""
    lo = np.iinfo(np.int64).min
    for op, dst, a, b, cmp, sreg in spec["ops"]:
        if op not in (0x83, 0x84, 0x93, 0x94):
            continue
        xa = oracle_expr_col(o, x, total, spec, regs, a, False)
        if op in (0x83, 0x84):
            xb = oracle_expr_col(o, x, total, spec, regs, b, False)
        else:
            xb = np.full(total, np.uint64(regs[sreg]).view(np.int64))
        if np.any((xa == lo) & (xb == -1)):
            return True
    return False


def oracle_expr(o, x, total, spec, regs, is_f64=True, acc0=None):
    """vo_tier0_expr_* on the spec; scalar operands from regs (register bits)."""
    ops = spec["ops"]
    code, sval = _expr_words(spec, regs, is_f64)
    nb = len(spec["red"])
    bvreg = np.array(spec["bvreg"], np.int32)
    red = np.array(spec["red"], np.int32)
    accop = np.array(spec["accop"], np.int32)
    acc = np.array(acc0 if acc0 is not None else [0] * nb, np.uint64)
    part = np.zeros(nb, np.uint64)
    fn = o.O.vo_tier0_expr_f64 if is_f64 else o.O.vo_tier0_expr_i64
    fn(o.ptr(x), x.size, 0, total, 0, len(ops), o.ptr(code), o.ptr(sval), nb, o.ptr(bvreg), o.ptr(red),
       o.ptr(accop), o.ptr(acc), o.ptr(part))
    return [int(a) for a in acc], [int(p) for p in part]


WINDOW_OPS = {1: 0xDE, 2: 0xDF, 3: 0x97, 4: 0x98, 5: 0xDD,   # WINDOW_SUM WINDOW_MEAN WSTD WVARIANCE_W WDELTA
              6: 0x99}                                     # WQUANTILE: window = imm (5-bit w | 3-bit q << 5)


def wquantile_imm(rng):
    """A WQUANTILE imm: window 2..31 or 0 (meaning 16), quantile 0..7."""
    wq = int(rng.choice([0, 2, 3, 5, 8, 16, 31]))
    return wq | (int(rng.integers(0, 8)) << 5)


def wquantile_values(x, imm):
    """Element (q * w) / 7 of each sorted window x[s:s + w] (numpy)."""
    wn = imm & 0x1F or 16
    idx = min(((imm >> 5) & 7) * wn // 7, wn - 1)
    if x.size < wn:
        return np.zeros(0)
    return np.sort(np.lib.stride_tricks.sliding_window_view(x, wn), axis=1)[:, idx]


def make_window_program(rng, wkind, window, nops=3, reds=(0, 1, 2, 3)):
    """A loop whose load is a window word (segment 0, offset R1, the word
    itself advancing R1), then the vector words and reductions of
    make_expr_program, CMP R1, R2; JNZ; HALT (no ADD of the offset)."""
    code, spec = make_expr_program(rng, True, nops=nops, reds=reds)
    code = [int(c) for c in code]
    imm = (window & 0xFF) if wkind != 5 else 0
    code[1] = w(WINDOW_OPS[wkind], rd=0, ra=1, imm=imm)
    del code[-4]                                           # the ADD R1, R1, R4 of the VLOAD loop
    jnz = len(code) - 2
    code[jnz] = w(0x52, imm=(1 - jnz) & 0xFFF)
    return np.array(code, np.uint32), spec


def oracle_window_loop(o, x, total, spec, regs, wkind, window, acc0=None, off=0):
    """vo_tier0_window_f64 on the spec from offset `off`: (terminates,
    accumulators, partials)."""
    code, sval = _expr_words(spec, regs, True)
    nb = len(spec["red"])
    bvreg = np.array(spec["bvreg"], np.int32)
    red = np.array(spec["red"], np.int32)
    accop = np.array(spec["accop"], np.int32)
    acc = np.array(acc0 if acc0 is not None else [0] * nb, np.uint64)
    part = np.zeros(nb, np.uint64)
    ok = o.O.vo_tier0_window_f64(o.ptr(x), x.size, off, total, wkind, window, 0, len(spec["ops"]), o.ptr(code),
                                 o.ptr(sval), nb, o.ptr(bvreg), o.ptr(red), o.ptr(accop), o.ptr(acc), o.ptr(part))
    return bool(ok), [int(a) for a in acc], [int(p) for p in part]


def window_total(rng, n, wkind):
    """A loop bound the reference reaches: Count (WDELTA: Count - 1), or a
    multiple of 4 short of it."""
    end = n - 1 if wkind == 5 else n
    if rng.random() < 0.5 or end < 8:
        return end
    return 4 * int(rng.integers(1, (end - 1) // 4 + 1))


REL_OPS = [0xEE, 0xEF, 0xF0, 0xF1, 0xF2, 0xF3, 0xF4, 0xF5, 0xF6, 0xF7, 0xF8, 0xF9, 0xFA, 0xFD]
KEY_REGS = (8, 9, 10, 11, 12)          # probe keys, set by the caller
K_REGS = (13, 14)                      # row counts of the partial sorts


def rel_program(rng, nseg, k=10, ops=REL_OPS, kmax=None):
    """Straight-line program of hash / sort / join words over nseg
    segments: hash slots 0..3 (HASH_INIT/BUILD write R0..R3), probes into
    R4..R7 with keys from KEY_REGS, sorts and partial sorts (k from
    K_REGS, capped by imm), joins and PARTITION_HASH into R15..R20; now and
    then a segment id past the last one (the word does nothing).  Ends with
    HALT."""
    code = []
    built = []
    for _ in range(k):
        op = int(rng.choice(ops))
        seg = lambda: int(rng.integers(0, nseg + 1)) if rng.random() < 0.1 else int(rng.integers(0, nseg))  # noqa: E731
        if op in (0xEE, 0xF0):
            slot = int(rng.integers(0, 4))
            built.append(slot)
            code.append(w(op, rd=slot, ra=seg()))
        elif op in (0xEF, 0xF1):
            slot = int(rng.choice(built)) if built and rng.random() < 0.85 else int(rng.integers(0, 6))
            code.append(w(op, rd=int(rng.integers(4, 8)), ra=slot, rb=int(rng.choice(KEY_REGS))))
        elif op in (0xF4, 0xF5):
            imm = int(rng.integers(0, 4096)) if kmax is None else int(rng.integers(0, kmax + 1))
            code.append(w(op, rd=int(rng.integers(15, 16)), ra=seg(), rb=int(rng.choice(K_REGS)), imm=imm))
        else:
            code.append(w(op, rd=int(rng.integers(15, 16)), ra=seg(), rb=seg()))
    code.append(w(0x01))
    return np.array(code, np.uint32)
