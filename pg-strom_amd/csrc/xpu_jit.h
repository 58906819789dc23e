/*
 * xpu_jit.h - per-plan specialization of the kernels (host side).
 *
 * The kernel bodies of xpu_device.h are templates over the program runner.
 * Ahead of time they run the plan through the op interpreter; here the
 * plan is emitted into a HIP source as constant bytes, each program segment
 * as straight-line calls of the op semantics (xop_exec) with constant
 * operands, and compiled with hiprtc for gfx950.  The compiler folds the op
 * dispatch, the register numbers and every metadata load of the plan, so
 * the specialized kernel keeps the register file in plain VGPRs and has no
 * per-row scalar-load chains.  Results are those of the interpreter by
 * construction: both run the same device functions.
 */
#ifndef XPU_JIT_H
#define XPU_JIT_H
#include <string>
#include <vector>

#include "xpu_program.h"

enum xpu_jit_kind
{
    XJ_SCAN_AGG = 1,        /* scan_agg_body<NR, RPT> (param = RPT) */
    XJ_PIPELINE = 2,        /* pipeline_body<PREAGG, NR> (param = PREAGG) */
};

/* HIP source of the specialized kernel "xpu_jit_kernel" */
std::string xpu_jit_source(const xplan *P, int kind, int param);

/* hiprtc compile for gfx950 into a code object; false with *log filled */
bool xpu_jit_compile(const std::string &src, std::vector<char> *code, std::string *log);

/* code object of the specialized kernel, compiled once per distinct source
 * in the process (thread-safe); nullptr with *log filled on failure */
const std::vector<char> *xpu_jit_code(const xplan *P, int kind, int param, std::string *log);

#endif
