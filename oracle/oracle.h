/*
 * oracle.h - CPU restatement of PG-Strom's xPU scan / join / pre-aggregation
 * semantics.  TEST INFRASTRUCTURE ONLY: tests, smoke() and bench.py's
 * cpu_baseline leg load it as the checker; the product never links it.
 *
 * It consumes the same kern_session_info blob and kern_data_store inputs as
 * the HIP library (include/xpuhip.h) and produces the same output formats:
 * a KDS_FORMAT_ROW kds_dst in input-row order for scans/joins, and a
 * group table written as the reference's kds_final (HASH tuples converted to
 * ROW, gpu_service.c:2780-2818) for pre-aggregation.
 */
#ifndef PGSTROM_ORACLE_H
#define PGSTROM_ORACLE_H
#include "pgstrom_xpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_session oracle_session;

oracle_session *oracle_session_open(const kern_session_info *session, size_t length,
                                    kern_errorbuf *errbuf);
int     oracle_task_exec(oracle_session *os,
                         const kern_data_store *kds_src,
                         const kern_multirels *kmrels,
                         kern_data_store *kds_dst, size_t dst_length,
                         int nthreads,
                         kern_exec_results *results,
                         kern_errorbuf *errbuf);
int64_t oracle_task_final(oracle_session *os, kern_data_store *kds_final,
                          size_t final_length, kern_errorbuf *errbuf);
/* the CPU-fallback rows of the tasks so far (KDS_FORMAT_FALLBACK) */
int64_t oracle_session_fallback(oracle_session *os, kern_data_store *out, size_t length);
void    oracle_session_close(oracle_session *os);

uint32_t oracle_hash_any(const void *ptr, int len);

/* one expression over variable slots (the per-function checker) */
int     oracle_eval_expr(const void *kexp_blob, size_t kexp_len, int compare_nulls,
                         const char *encname, int nslots, const int32_t *slot_types, const uint8_t *slot_data,
                         const int64_t *slot_offs, int32_t res_type, uint8_t *res_image,
                         int32_t res_cap, int32_t *res_len, uint32_t *errcode, char *errmsg,
                         int errlen);

#ifdef __cplusplus
}
#endif
#endif
