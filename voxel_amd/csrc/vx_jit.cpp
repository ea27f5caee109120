// vx_jit.cpp -- run-time compilation of Tier-0 loop bodies (the GPU side of
// what the reference's JIT does for its fused loops, jit/compiler.hpp): a
// loop whose reductions are VSUM / VCOUNT / VRED_MIN / VRED_MAX is turned
// into HIP source in
// which every vector word is one apply_op<T, OP> call on named registers
// (straight-line code, no per-row dispatch), compiled once per program shape
// with hipRTC for the device's architecture, and kept for the process.  The
// shape is the words' opcodes and registers and the reductions; scalar
// operands, thresholds, masks and lane immediates stay kernel arguments, so
// one compiled kernel serves every value of them.  The word semantics are the
// ones the library's interpreter kernels use (vx_expr_ops.hpp, embedded here
// as text at build time).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <atomic>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_expr_abi.hpp"
#include "vx_runtime.hpp"

namespace vx {
namespace {

// The text of vx_expr_abi.hpp and vx_expr_ops.hpp (their #include and
// #pragma once lines removed), written by the build.
const char kOpsSource[] =
#include "vx_expr_src.inc"
    ;

std::atomic<int> g_mode{1};                       // 0 never, 1 loops of kJitRows rows or more, 2 always
constexpr uint64_t kJitRows = 1ull << 20;         // below this a compile costs more than it saves

struct Compiled {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};
std::map<std::string, Compiled>& cache() {        // (device, source) -> kernel; kept to process end
    static auto* m = new std::map<std::string, Compiled>();
    return *m;
}

const std::set<int>& own_cases() {
    static const std::set<int> s = {
#define VX_CASE(c) c,
        VX_EXPR_CASES(VX_CASE)
#undef VX_CASE
    };
    return s;
}

std::string hex2(int v) {
    static const char* d = "0123456789ABCDEF";
    return std::string("0x") + d[(v >> 4) & 15] + d[v & 15];
}

// The kernel for one program shape.  Rows base + j * 256 + t (j < 4) of a
// 1024-row tile are loaded together, then each row runs the words as
// straight-line code; rows past n are evaluated on 0 (so a group's four
// threads always run a word together) and not reduced.
std::string kernel_source(bool f64, const ExprProg& prog, const ExprRed& rd) {
    std::string s;
    s += "typedef __hip_internal::uint64_t uint64_t; typedef __hip_internal::int64_t int64_t;\n"
         "typedef __hip_internal::uint32_t uint32_t; typedef __hip_internal::int32_t int32_t;\n"
         "typedef __hip_internal::uint16_t uint16_t; typedef __hip_internal::int16_t int16_t;\n"
         "typedef __hip_internal::uint8_t uint8_t;\n";
    s += kOpsSource;
    s += "\ntypedef ";
    s += f64 ? "double" : "int64_t";
    s += " VxT;\n"
         "extern \"C\" __global__ __launch_bounds__(256) void vx_expr_jit(const VxT* __restrict__ col, uint64_t count,\n"
         "    uint64_t off, uint64_t n, vx::ExprProg prog, uint64_t* __restrict__ parts, uint64_t* __restrict__ last,\n"
         "    int* __restrict__ trap) {\n"
         "  using namespace vx::xo;\n"
         "  __shared__ XAcc<VxT> lds[4];\n"
         "  __shared__ GAccX glds[4];\n";
    for (int q = 0; q < rd.n; ++q) {
        const std::string Q = std::to_string(q);
        s += rd.red[q] <= 1 ? "  XAcc<VxT> a" + Q + "; xacc_init(a" + Q + ");\n"
                            : "  GAccX a" + Q + "; gacc_init(a" + Q + ");\n";
    }
    for (int k = 0; k < prog.nops; ++k)
        s += "  const VxT s" + std::to_string(k) + " = lane_of<VxT>(prog.ops[" + std::to_string(k) + "].s);\n";
    s += "  const int lane = (int)(threadIdx.x & 3);\n"
         "  bool tr = false;\n"
         "  for (uint64_t base = (uint64_t)blockIdx.x * 1024; base < n; base += (uint64_t)gridDim.x * 1024) {\n";
    for (int j = 0; j < 4; ++j) {
        const std::string J = std::to_string(j);
        s += "    const uint64_t i" + J + " = base + " + std::to_string(256 * j) + " + threadIdx.x;\n"
             "    const VxT x" + J + " = i" + J + " < n && off + i" + J + " < count ? col[off + i" + J + "] : VxT(0);\n";
    }
    for (int j = 0; j < 4; ++j) {
        const std::string J = std::to_string(j);
        bool def[16] = {};
        s += "    {\n      uint32_t m = 0; bool t = false;\n";
        s += "      const VxT v" + std::to_string(prog.vload) + " = x" + J + ";\n";
        def[prog.vload & 15] = true;
        for (int k = 0; k < prog.nops; ++k) {
            const ExprOp& o = prog.ops[k];
            const int code = own_cases().count(o.op) ? o.op : 0xC0;
            const std::string a = def[o.a & 15] ? "v" + std::to_string(o.a & 15) : "VxT(0)";
            const std::string b = def[o.b & 15] ? "v" + std::to_string(o.b & 15) : "VxT(0)";
            s += "      const VxT v" + std::to_string(o.dst & 15) + " = apply_op<VxT, " + hex2(code) + ">(prog.ops[" +
                 std::to_string(k) + "], prog, " + a + ", " + b + ", s" + std::to_string(k) + ", lane, m, t);\n";
            def[o.dst & 15] = true;
        }
        s += "      if (i" + J + " < n) {\n        tr = tr || t;\n";
        for (int q = 0; q < rd.n; ++q) {
            const std::string Q = std::to_string(q), R = "v" + std::to_string(rd.vreg[q] & 15);
            if (rd.red[q] == 0) s += "        xacc_sum(a" + Q + ", " + R + ");\n";
            else if (rd.red[q] == 1) s += "        xacc_count(a" + Q + ", " + R + ");\n";
            else
                s += "        { const uint64_t g = group_minmax(" + R + ", " + std::to_string(rd.red[q]) +
                     "); if (lane == 0) gacc_add(a" + Q + ", g); }\n";
        }
        s += "        if (i" + J + " + 4 >= n) {\n";
        for (int q = 0; q < rd.n; ++q)
            s += "          group_partial(v" + std::to_string(rd.vreg[q] & 15) + ", " + std::to_string(rd.red[q]) +
                 ", lane, &last[" + std::to_string(q) + "]);\n";
        s += "        }\n      }\n    }\n";
    }
    s += "  }\n  if (tr) *trap = 1;\n";
    for (int q = 0; q < rd.n; ++q)
        s += std::string(rd.red[q] <= 1 ? "  xacc_block_store(a" : "  gacc_block_store(a") + std::to_string(q) +
             (rd.red[q] <= 1 ? ", lds" : ", glds") + ", parts + 8 * ((uint64_t)" + std::to_string(q) +
             " * gridDim.x + blockIdx.x));\n";
    s += "}\n";
    return s;
}

// Compiles src for the current device; false with the error set.
bool compile(const std::string& src, Compiled* out) {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    hipDeviceProp_t prop;
    if (!hip_ok(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties")) return false;
    hiprtcProgram p;
    if (hiprtcCreateProgram(&p, src.c_str(), "vx_expr_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        set_error("run-time compile: hiprtcCreateProgram failed");
        return false;
    }
    const std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
    const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(p, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(p, &ls);
        std::string log(ls + 1, '\0');
        if (ls) hiprtcGetProgramLog(p, &log[0]);
        hiprtcDestroyProgram(&p);
        set_error("run-time compile of a loop body failed: %.600s", log.c_str());
        return false;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(p, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(p, code.data());
    hiprtcDestroyProgram(&p);
    if (!hip_ok(hipModuleLoadData(&out->mod, code.data()), "hipModuleLoadData")) return false;
    return hip_ok(hipModuleGetFunction(&out->fn, out->mod, "vx_expr_jit"), "hipModuleGetFunction");
}

}  // namespace

// Launches the compiled form of (prog, rd) over rows off + i, i < n, writing
// nb workgroup partials per reduction into parts (8 words each, reduction q's
// at q * nb) and the last group's values into last.  *launched = false (and
// true returned) when the mode leaves this loop to the interpreter, or, in
// mode 1, when compiling fails; false with the error set otherwise.
bool jit_expr_launch(bool f64, const void* col, uint64_t count, uint64_t off, uint64_t n, const ExprProg& prog,
                     const ExprRed& rd, uint64_t* parts, unsigned nb, uint64_t* last, int* trap, void* stream,
                     bool* launched) {
    *launched = false;
    const int mode = g_mode.load();
    if (mode == 0 || (mode == 1 && n < kJitRows)) return true;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
    const std::string src = kernel_source(f64, prog, rd);
    const std::string key = std::to_string(dev) + "\n" + src;
    auto it = cache().find(key);
    if (it == cache().end()) {
        Compiled k;
        if (!compile(src, &k)) {
            if (mode == 1) { set_error("%s", ""); return true; }     // the interpreter runs the loop instead
            return false;
        }
        it = cache().emplace(key, k).first;
    }
    ExprProg pa = prog;
    void* args[] = {(void*)&col, (void*)&count, (void*)&off, (void*)&n, (void*)&pa, (void*)&parts, (void*)&last,
                    (void*)&trap};
    if (!hip_ok(hipModuleLaunchKernel(it->second.fn, nb, 1, 1, 256, 1, 1, 0, st, args, nullptr),
                "hipModuleLaunchKernel"))
        return false;
    *launched = true;
    return true;
}

}  // namespace vx

extern "C" int voxel_gpu_set_expr_jit(int mode) {
    if (mode < 0 || mode > 2) {
        vx::set_error("expression compile mode %d (0 never, 1 large loops, 2 always)", mode);
        return -1;
    }
    return vx::g_mode.exchange(mode);
}
