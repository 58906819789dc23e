/*
 * xpu_compile.h - session blob -> xplan translation (host side).
 */
#ifndef XPU_COMPILE_H
#define XPU_COMPILE_H
#include <string>
#include <vector>

#include "pgstrom_xpu.h"
#include "xpu_program.h"
#include "xpu_kernels.h"

/* a kds_src column the programs load: XH_FIXED (width bytes), XH_BOOL or
 * XH_VAR (xpu_kernels.h); kind 0: not loaded */
struct xpu_src_col
{
    uint8_t     kind;
    uint8_t     width;
};

struct xpu_compiled
{
    xplan               plan;
    std::vector<char>   const_pool;         /* text constants (payload bytes) */
    std::vector<std::string> messages;      /* error texts, indexed by xop.imm */
    std::vector<char>   kds_dst_head;       /* KDS_FORMAT_ROW header of kds_dst */
    std::vector<char>   kds_final_head;     /* header of kds_final (HASH or ROW) */
    std::vector<char>   kds_fallback_head;  /* KDS_FORMAT_FALLBACK header (empty: no
                                               CPU fallback) */
    bool                has_projection = false;
    bool                has_groupby = false;
    bool                pinned_hash = false;    /* projection rows go to a device-resident
                                                   KDS_FORMAT_HASH buffer (kds_final_head) */
    uint32_t            nslots = 0;
    uint32_t            lds_entry_bytes = 0;    /* bytes per LDS group entry */
    std::vector<int64_t> params;                /* XO_PARAM values (raw register form) */
    uint64_t            param_nulls = 0;        /* bit i: parameter i is NULL */
    std::vector<xpu_src_col> src_cols;          /* by kds_src column index */
    bool                has_xact_state = false;
    std::vector<uint32_t> xids;                 /* parallelCurrentXids */
    /* GPU-Sort of the final buffer (FuncOpCode__SortKeys) */
    std::vector<xsort_key> sort_keys;
    uint32_t            sort_nwords = 0;
    /* GPU-Sort window rank (KSORT_WINDOW_FUNC__*, 0: none): the limit and
     * the key words of the partition keys and of partition + order keys */
    char        sort_wr_func = 0;
    uint32_t    sort_wr_limit = 0;
    uint32_t    sort_wr_pwords = 0;
    uint32_t    sort_wr_owords = 0;
    int64_t             sort_limit = 0;         /* gpusort_limit_count, if positive */
};

/* Validates and translates the session blob; returns 0 or an ERRCODE_*
 * with the error buffer filled. */
int xpu_compile_session(const kern_session_info *session, size_t length,
                        xpu_compiled *out, kern_errorbuf *errbuf);

std::string xpu_plan_dump(const xpu_compiled *c);

void xpu_set_error(kern_errorbuf *eb, uint32_t errcode, const char *filename,
                   int lineno, const char *funcname, const char *message);

#endif
