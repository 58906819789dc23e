/*
 * xpu_jit.cpp - per-plan specialization of the kernels with hiprtc.
 */
#include <hip/hiprtc.h>
#include <stddef.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <memory>
#include <mutex>

#include "xpu_jit.h"

extern "C" {
extern const int xpu_jit_nsources;
extern const char *const xpu_jit_source_names[];
extern const char *const xpu_jit_source_texts[];
}

namespace {

/* hiprtc compiles without the host's C headers: the few the device code
 * includes are given here (LP64 types, as on the host) */
const char *const k_std_names[] = {
    "stdint.h", "stddef.h", "stdbool.h", "math.h", "hip/hip_runtime.h",
};
const char *const k_std_texts[] = {
    "#pragma once\n"
    "typedef signed char int8_t; typedef short int16_t; typedef int int32_t; typedef long int64_t;\n"
    "typedef unsigned char uint8_t; typedef unsigned short uint16_t; typedef unsigned int uint32_t;\n"
    "typedef unsigned long uint64_t; typedef long intptr_t; typedef unsigned long uintptr_t;\n"
    "#define INT8_MIN (-128)\n#define INT8_MAX 127\n#define UINT8_MAX 255\n"
    "#define INT16_MIN (-32768)\n#define INT16_MAX 32767\n#define UINT16_MAX 65535\n"
    "#define INT32_MIN (-2147483647-1)\n#define INT32_MAX 2147483647\n#define UINT32_MAX 4294967295U\n"
    "#define INT64_MIN (-9223372036854775807L-1)\n#define INT64_MAX 9223372036854775807L\n"
    "#define UINT64_MAX 18446744073709551615UL\n",
    "#pragma once\ntypedef unsigned long size_t; typedef long ptrdiff_t;\n"
    "#define offsetof(t, m) __builtin_offsetof(t, m)\n"
    "#ifndef NULL\n#define NULL nullptr\n#endif\n",
    "#pragma once\n",
    "#pragma once\n",          /* device math comes with hiprtc's runtime header */
    "#pragma once\n",
};

std::mutex g_lock;
std::map<std::string, std::unique_ptr<std::vector<char>>> g_code;     /* source -> code */

void
append(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

void
append(std::string &s, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    s += buf;
}

void
emit_segment(std::string &s, const xplan *P, int begin, int end)
{
    append(s, "        if (begin == %d && end == %d)\n        {\n", begin, end);
    for (int pc = begin; pc < end; pc++)
    {
        const xop &o = P->ops[pc];
        append(s, "            xop_exec<STAGED>(xop{%u, %u, %u, %u, %u, %u, (int64_t)0x%016llxULL}, "
                  "L, kds, row, kp);\n",
               o.op, o.dst, o.a, o.b, o.guard, o.aux, (unsigned long long)o.imm);
    }
    s += "            return;\n        }\n";
}

}   /* namespace */

std::string
xpu_jit_source(const xplan *P, int kind, int param)
{
    /* the plan without the session's device addresses (they come from the
     * kernel parameters), so equal plans share one compiled kernel */
    xplan copy;
    memcpy(&copy, P, sizeof(xplan));
    for (int d = 0; d <= XP_MAX_DEPTH; d++)
        copy.depth[d].inner_kds = 0;
    copy.const_pool = 0;
    size_t nbytes = offsetof(xplan, ops) + sizeof(xop) * (size_t)P->nops;
    const unsigned char *b = (const unsigned char *)&copy;

    std::string s;
    s.reserve(8 * nbytes + 4096);
    s += "#include \"xpu_device.h\"\n\n";
    append(s, "__constant__ const unsigned char xjit_plan[%zu] __attribute__((aligned(16))) = {", nbytes);
    for (size_t i = 0; i < nbytes; i++)
        append(s, "%s%u", i == 0 ? "\n    " : (i % 32 == 0 ? ",\n    " : ","), b[i]);
    s += "\n};\n\n";
    s += "struct JitProg\n{\n"
         "    static constexpr bool kUnrollRows = true;\n\n"
         "    template <bool STAGED>\n"
         "    __device__ static __forceinline__ void\n"
         "    run(cplan *P, int begin, int end, Lane &L, const kern_data_store *__restrict__ kds,\n"
         "        uint32_t row, const xpu_kparams &kp)\n"
         "    {\n";
    if (kind == XJ_SCAN_AGG)
        emit_segment(s, P, P->seg_scan.begin + P->stage_prefix, P->seg_final.end);
    else
    {
        emit_segment(s, P, P->seg_scan.begin, P->seg_scan.end);
        for (uint32_t d = 1; d <= P->nrels; d++)
        {
            emit_segment(s, P, P->depth[d].seg_hash.begin, P->depth[d].seg_hash.end);
            emit_segment(s, P, P->depth[d].seg_join.begin, P->depth[d].seg_join.end);
            emit_segment(s, P, P->depth[d].seg_oj.begin, P->depth[d].seg_oj.end);
        }
        emit_segment(s, P, P->seg_final.begin, P->seg_final.end);
        /* a projection without joins runs its quals and targets as one
         * program after the staged loads */
        if (P->nrels == 0)
            emit_segment(s, P, P->seg_scan.begin + P->stage_prefix, P->seg_final.end);
    }
    s += "    }\n};\n\n";
    append(s, "extern \"C\" __global__ void __launch_bounds__(%s)\n"
              "xpu_jit_kernel(xpu_kparams kp)\n{\n",
           kind == XJ_SCAN_AGG ? "XP_AGG_BLOCK_MAX" : "XP_BLOCK_MAX");
    int nr = (P->nregs <= XP_NREGS_SMALL) ? XP_NREGS_SMALL : XP_NREGS;
    if (kind == XJ_SCAN_AGG)
        append(s, "    scan_agg_body<%d, %d, JitProg>((cplan *)xjit_plan, kp);\n", nr, param);
    else
        append(s, "    pipeline_body<%s, %d, JitProg>((cplan *)xjit_plan, kp);\n",
               param ? "true" : "false", nr);
    s += "}\n";
    if (kind == XJ_PIPELINE && param && P->nrels > 0)
    {
        s += "\nextern \"C\" __global__ void __launch_bounds__(XQ_SCREEN_BLOCK)\n"
             "xpu_jit_screen_kernel(xpu_kparams kp)\n{\n";
        append(s, "    join_screen_body<%d, XQ_SCREEN_RPT, JitProg>((cplan *)xjit_plan, kp);\n}\n", nr);
        for (uint32_t d = 1; d <= P->nrels; d++)
        {
            append(s, "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
                      "xpu_jit_probe_kernel_%u(xpu_kparams kp)\n{\n", d);
            append(s, "    join_probe_body<%d, JitProg, %u>((cplan *)xjit_plan, kp);\n}\n", nr, d);
        }
    }
    if (kind == XJ_PIPELINE && !param && P->nrels == 0)
    {
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_scount_kernel(xpu_kparams kp)\n{\n";
        append(s, "    scan_proj_body<%d, XP_PROJ_RPT, JitProg, PM_COUNT>((cplan *)xjit_plan, kp);\n}\n", nr);
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_swrite_kernel(xpu_kparams kp)\n{\n";
        append(s, "    scan_proj_body<%d, XP_PROJ_RPT, JitProg, PM_WRITE>((cplan *)xjit_plan, kp);\n}\n", nr);
    }
    if (kind == XJ_PIPELINE && !param)
    {
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_count_kernel(xpu_kparams kp)\n{\n";
        append(s, "    pipeline_body<false, %d, JitProg, PM_COUNT>((cplan *)xjit_plan, kp);\n}\n", nr);
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_write_kernel(xpu_kparams kp)\n{\n";
        append(s, "    pipeline_body<false, %d, JitProg, PM_WRITE>((cplan *)xjit_plan, kp);\n}\n", nr);
    }
    if (kind == XJ_SCAN_AGG)
    {
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_spill_kernel(xpu_kparams kp)\n{\n";
        append(s, "    spill_agg_body<%d, JitProg>((cplan *)xjit_plan, kp);\n}\n", nr);
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_dump_kernel(xpu_kparams kp)\n{\n";
        append(s, "    dump_merge_body<%d, JitProg>((cplan *)xjit_plan, kp);\n}\n", nr);
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_AGG_BLOCK_MAX)\n"
             "xpu_jit_batch_kernel(xpu_kparams kp)\n{\n";
        append(s, "    batch_agg_body<%d, %d, JitProg>((cplan *)xjit_plan, kp);\n}\n", nr, param);
        s += "\nextern \"C\" __global__ void __launch_bounds__(XP_BLOCK_MAX)\n"
             "xpu_jit_batch_spill_kernel(xpu_kparams kp)\n{\n";
        append(s, "    batch_spill_body<%d, JitProg>((cplan *)xjit_plan, kp);\n}\n", nr);
    }
    return s;
}

bool
xpu_jit_compile(const std::string &src, std::vector<char> *code, std::string *log)
{
    std::vector<const char *> names, texts;
    for (size_t i = 0; i < sizeof(k_std_names) / sizeof(k_std_names[0]); i++)
    {
        names.push_back(k_std_names[i]);
        texts.push_back(k_std_texts[i]);
    }
    for (int i = 0; i < xpu_jit_nsources; i++)
    {
        names.push_back(xpu_jit_source_names[i]);
        texts.push_back(xpu_jit_source_texts[i]);
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "xpu_jit.hip", (int)names.size(), texts.data(),
                            names.data()) != HIPRTC_SUCCESS)
    {
        *log = "hiprtcCreateProgram failed";
        return false;
    }
    /* the same code generation flags as the ahead-of-time build */
    const char *opts[] = {
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-pass-failed",
    };
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t lsz = 0;
    hiprtcGetProgramLogSize(prog, &lsz);
    std::string lg(lsz, '\0');
    if (lsz)
        hiprtcGetProgramLog(prog, &lg[0]);
    bool ok = (rc == HIPRTC_SUCCESS);
    if (ok)
    {
        size_t csz = 0;
        hiprtcGetCodeSize(prog, &csz);
        code->resize(csz);
        hiprtcGetCode(prog, code->data());
        ok = csz > 0;
    }
    else
        lg = std::string(hiprtcGetErrorString(rc)) + ": " + lg;
    hiprtcDestroyProgram(&prog);
    *log = lg;
    return ok;
}

const std::vector<char> *
xpu_jit_code(const xplan *P, int kind, int param, std::string *log)
{
    std::string src = xpu_jit_source(P, kind, param);
    {
        std::lock_guard<std::mutex> g(g_lock);
        auto it = g_code.find(src);
        if (it != g_code.end())
            return it->second.get();
    }
    std::unique_ptr<std::vector<char>> code(new std::vector<char>());
    if (!xpu_jit_compile(src, code.get(), log))
        return nullptr;
    std::lock_guard<std::mutex> g(g_lock);
    auto &slot = g_code[src];
    if (!slot)
        slot = std::move(code);
    return slot.get();
}
