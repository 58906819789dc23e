/*
 * xpuhip.h - C ABI of the MI355X xPU execution library (libxpuhip.so).
 *
 * These entry points replace the device half of PG-Strom's GPU service: what
 * gpu_service.c does between receiving an XpuCommand and writing the response
 * back (session open with kernel-pointer resolution, inner-buffer import,
 * per-chunk kernel launch with suspend/resume, final-buffer write-back).  The
 * service loop (socket, command queue, worker threads) calls them; all
 * arguments are plain pointers and sizes in the reference's own wire formats
 * (include/pgstrom_xpu.h).  Return values are the reference's ERRCODE_*:
 *   0    success
 *   'd'  ERRCODE_SUSPEND_NO_SPACE: kds_dst is full; the rows written so far are
 *        valid; call again with *resume_pos unchanged to continue the chunk
 *        (the reference resumes the kernel, gpu_service.c:3896-3908)
 *   'E'  ERRCODE_DEVICE_ERROR, with *errbuf filled
 * The library never falls back to the CPU: without a usable GPU every entry
 * point fails with 'E'.
 */
#ifndef XPUHIP_H
#define XPUHIP_H
#include "pgstrom_xpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xpuhip_session xpuhip_session;

/* Tuning of the pipeline launch; zero fields keep the library's choice for
 * the session's task kind. */
typedef struct
{
    uint32_t    block_size;         /* threads per block (64..512, multiple of 64) */
    uint32_t    rows_per_thread;    /* group-by rows per lane per tile (1, 2 or 4
                                       without joins) */
    uint32_t    lds_nents;          /* group-by LDS table entries */
    uint32_t    blocks_per_cu;      /* persistent blocks per compute unit */
    uint32_t    jit;                /* XPUHIP_JIT_*: kernel specialized for the plan */
    uint32_t    join_filter;        /* XPUHIP_JOIN_FILTER_*: inner-relation Bloom filters */
    uint32_t    join_passes;        /* XPUHIP_JOIN_PASSES_*: group-by joins in passes */
} xpuhip_tuning;

#define XPUHIP_JIT_AUTO     0       /* specialize chunks of >= 1,000,000 rows */
#define XPUHIP_JIT_ON       1       /* always specialize */
#define XPUHIP_JIT_OFF      2       /* always run the op interpreter */

/* Joins probe a Bloom filter of each inner relation's join hashes (built by
 * xpuhip_session_attach_inner) before walking its hash chain, and a group-by
 * over joins screens its rows with the depth-1 filter before the joins.
 * Results are the same either way. */
#define XPUHIP_JOIN_FILTER_AUTO 0   /* on */
#define XPUHIP_JOIN_FILTER_ON   1
#define XPUHIP_JOIN_FILTER_OFF  2

/* A group-by over joins that runs the specialized kernels (xpuhip_tuning.jit)
 * runs as a screen pass (scan quals and the depth-1 filter) and one pass per
 * depth over the rows that reach it, each its own launch over dense records,
 * instead of one kernel that carries every row through every depth.  Results
 * are the same either way. */
#define XPUHIP_JOIN_PASSES_AUTO 0   /* passes, unless the inner relations' duplicate
                                       keys need more than the record budget */
#define XPUHIP_JOIN_PASSES_ON   1
#define XPUHIP_JOIN_PASSES_OFF  2

/* Device setup; replaces gpuservSetupGpuContext (gpu_service.c:6462-6523).
 * device < 0 selects device 0. */
int     xpuhip_init(int device, kern_errorbuf *errbuf);
void    xpuhip_shutdown(void);

/* OpenSession: validates and compiles the kern_session_info blob built by
 * pgstromBuildSessionInfo (executor.c:517-801), and allocates the session's
 * kds_final when it carries a group-by.  Replaces gpuservHandleOpenSession
 * (gpu_service.c:3319-3410) and its device-pointer patching (:3156-3317). */
int     xpuhip_session_open(const kern_session_info *session, size_t length,
                            xpuhip_session **p_sess, kern_errorbuf *errbuf);

/* Validation and compilation only, without a GPU (for planners and tests):
 * returns 0 if the session would open. */
int     xpuhip_session_check(const kern_session_info *session, size_t length,
                             kern_errorbuf *errbuf);

/* Compiles (and keeps for the process) the kernel specialized for the
 * session blob with the default launch shape, so that the first task does
 * not wait for it; needs no GPU.
 * Has no counterpart in the reference, which interprets every plan. */
int     xpuhip_jit_prepare(const kern_session_info *session, size_t length,
                           kern_errorbuf *errbuf);

/* Text listing of the compiled programs of a session blob (the EXPLAIN of
 * this build's translation); returns its length, -1 on a compile error. */
int64_t xpuhip_session_explain(const kern_session_info *session, size_t length,
                               char *buf, size_t buflen, kern_errorbuf *errbuf);

/* Imports the inner relations (kern_multirels with KDS_FORMAT_HASH chunks,
 * gpu_join.c:2128-2164) into device memory for the session's hash joins;
 * replaces the shm import of gpu_service.c:2197-2354. */
int     xpuhip_session_attach_inner(xpuhip_session *sess, const kern_multirels *kmrels,
                                    size_t length, kern_errorbuf *errbuf);

/* Host-side build of one inner relation's KDS_FORMAT_HASH buffer, as the
 * backend's inner preload does it: hashes every MinimalTuple's join keys
 * (get_tuple_hashvalue, gpu_join.c:1548-1584: the device types' hash of
 * each key attribute, by its colmeta atttypid, merged by pg_hash_merge) and
 * packs the rows as kern_hashitems from the buffer's tail, chained from the
 * hash slots (__innerPreloadSetupHashBuffer, gpu_join.c:2128-2164).
 * kds: header and colmeta set up by the caller, length the buffer size,
 * hash_nslots > 0 (KDS_GET_HASHSLOT_WIDTH), nitems / usage what earlier
 * calls appended (0 at first).  mtups: nitems MinimalTuples (t_len without
 * a row id).  key_attnums: 1-based attribute numbers of the join keys.
 * nthreads <= 0 uses every core.  Needs no GPU.  Returns 0, or -1 with
 * errbuf set (buffer too small, unsupported key type, key that is still
 * compressed or external). */
int     xpuhip_inner_hash_setup(kern_data_store *kds, const void *const *mtups, uint32_t nitems,
                                const int16_t *key_attnums, uint32_t nkeys, int nthreads,
                                kern_errorbuf *errbuf);

/* Arrow IPC file chunk producer (arrowFdwLoadRecordBatch /
 * arrowFdwFillupRecordBatch, arrow_fdw.c:4018-4100): a file image is opened
 * once (its footer: schema and record batch blocks), then each record batch
 * becomes a KDS_FORMAT_ARROW chunk.  kds_head is the chunk's head with the
 * relation's colmeta (PostgreSQL types); field_index[j] names the Arrow
 * field of column j, or -1 for a column the plan does not reference.  The
 * attribute options and buffer offsets are set as
 * __arrowKdsAssignAttrOptions sets them, and only the referenced buffers
 * are copied.  xpuhip_arrow_load_batch returns the chunk size; with a null
 * kds or a too small length it only returns that size.  Uncompressed files
 * with flat fields (Int, FloatingPoint, Bool, Decimal128, Utf8 / Binary,
 * their Large forms, FixedSizeBinary); -1 with errbuf set otherwise. */
typedef struct xpuhip_arrow_file xpuhip_arrow_file;
int     xpuhip_arrow_open(const void *image, size_t length, xpuhip_arrow_file **p_file,
                          uint32_t *p_nbatches, uint32_t *p_nfields, kern_errorbuf *errbuf);
int     xpuhip_arrow_field(xpuhip_arrow_file *file, uint32_t index, char *name, size_t name_len,
                           int32_t *p_type, int32_t *p_bit_width, int32_t *p_scale);
int64_t xpuhip_arrow_load_batch(xpuhip_arrow_file *file, uint32_t rb_index,
                                const kern_data_store *kds_head, const int32_t *field_index,
                                kern_data_store *kds, size_t length, kern_errorbuf *errbuf);
void    xpuhip_arrow_close(xpuhip_arrow_file *file);

/* XpuTaskExec on a host-resident chunk (KDS_FORMAT_ARROW, or a heap-page /
 * row chunk as xpuhip_task_exec_device_tuples describes): copies it to the
 * device, runs the pipeline, and for a scan/join projection copies the rows
 * into kds_dst (KDS_FORMAT_ROW, dst_length bytes, header from the session).
 * *resume_pos must be 0 for a new chunk.  Replaces gpuservHandleGpuTaskExec
 * (gpu_service.c:4005-4290) with __gpuservLaunchGpuTaskExecKernel (:3759-4003). */
int     xpuhip_task_exec(xpuhip_session *sess,
                         const kern_data_store *kds_src, size_t src_length,
                         kern_data_store *kds_dst, size_t dst_length,
                         uint64_t *resume_pos,
                         kern_exec_results *results, kern_errorbuf *errbuf);

/* The same on a chunk already in device memory: d_kds_src is a device
 * address of a KDS_FORMAT_ARROW chunk of src_nitems rows; a projection writes
 * into the device buffer d_kds_dst of dst_length bytes. */
int     xpuhip_task_exec_device(xpuhip_session *sess,
                                const kern_data_store *d_kds_src, uint32_t src_nitems,
                                kern_data_store *d_kds_dst, size_t dst_length,
                                uint64_t *resume_pos,
                                kern_exec_results *results, kern_errorbuf *errbuf);

/* The same as xpuhip_task_exec_device for scan + projection and scan +
 * group-by sessions (no joins), allocating nothing: no hipMalloc, hipFree or
 * allocating library call.  Has no counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid it).  The task scratch is one
 * device array of 32 MiB held by each GPU for the life of the process: a
 * projection chunk needs 16 bytes per 256 rows of it (larger chunks fail
 * with 'E'); a group-by's spill and dump records take what fits, the rest
 * merge into kds_final directly.  Runs the ahead-of-time kernels on the null
 * stream and returns when they are done (xpuhip_last_kernel_ms reads 0).
 * Calls from threads of one process serialise on a lock, since they share
 * the array; calls from several processes each use their own copy. */
int     xpuhip_task_exec_device_noalloc(xpuhip_session *sess,
                                        const kern_data_store *d_kds_src, uint32_t src_nitems,
                                        kern_data_store *d_kds_dst, size_t dst_length,
                                        uint64_t *resume_pos,
                                        kern_exec_results *results, kern_errorbuf *errbuf);

/* The same on a heap-page chunk (KDS_FORMAT_BLOCK: kds_head->nitems pages of
 * BLCKSZ bytes with their line pointers, as relscan.c:413-742 builds it) or a
 * row chunk (KDS_FORMAT_ROW / KDS_FORMAT_HASH: kds_head->nitems row-index
 * entries) already in device memory at d_kds_src.  kds_head is a host copy
 * of the chunk's header and column metadata (KDS_HEAD_LENGTH bytes).  The
 * tuples the reference's scan would load - LP_NORMAL items of each page
 * (cuda_gpuscan.cu:405-534), visible rows of a row chunk (:610-700) - are
 * deformed on the device into column tiles, then run as an Arrow chunk of
 * that many rows in page / row-index order (*resume_pos counts those rows).
 * xpuhip_task_exec takes such chunks from host memory the same way. */
int     xpuhip_task_exec_device_tuples(xpuhip_session *sess,
                                       const kern_data_store *d_kds_src,
                                       const kern_data_store *kds_head,
                                       kern_data_store *d_kds_dst, size_t dst_length,
                                       uint64_t *resume_pos,
                                       kern_exec_results *results, kern_errorbuf *errbuf);

/* XpuTaskFinal of a group-by session: writes the session's kds_final into
 * kds_final (host memory) as the service writes it back, HASH converted to
 * ROW (gpu_service.c:2780-2818).  With kds_final == NULL returns the number of
 * bytes needed.  Returns the bytes written, or -1 with errbuf filled.
 * A session with FuncOpCode__SortKeys (xpucode_gpusort_keydesc) gets its
 * row index sorted by those keys on the device first, and with a positive
 * gpusort_limit_count only that many leading rows (gpuservSortingFinalBuffer,
 * gpu_service.c:5199-5345), their tuples packed as kern_buffer_simple_limit
 * packs them (the size query returns the bound before packing; the written
 * buffer's header gives its length).  The device buffer itself is not
 * reordered. */
int64_t xpuhip_task_final(xpuhip_session *sess, kern_data_store *kds_final, size_t length,
                          kern_exec_results *results, kern_errorbuf *errbuf);

/* The CPU-fallback rows of a session whose blob carries fallback_kds_head
 * (pg_strom.cpu_fallback on): the rows its tasks sent to the CPU since the
 * last call, as the KDS_FORMAT_FALLBACK chunk that gpuClientWriteBackPartial
 * returns to the backend (gpu_service.c:2853-2889; the backend re-runs them,
 * execCpuFallbackOneChunk, fallback.c:497-542).  A row falls back when its
 * depth-0 work (scan quals, the depth-1 join key hash, and without joins its
 * projection or group-by arguments) needs the bytes of a compressed or
 * external datum (HandleErrorIfCpuFallback, xpu_common.cu:1898-1955); it is
 * written with depth 0 and its depth-0 variables, and produces nothing on
 * the GPU.  The device buffer is emptied.  Returns the bytes written, 0 when
 * there are no rows, the bytes needed with kds_fallback == NULL, or -1 with
 * errbuf filled.  A fallback that does not fit the buffer (its length is the
 * session header's) fails the task with ERRCODE_DEVICE_ERROR. */
int64_t xpuhip_session_fallback(xpuhip_session *sess, kern_data_store *kds_fallback,
                                size_t length, kern_errorbuf *errbuf);

/* Releases the session and its device memory (gpuservCloseSession). */
void    xpuhip_session_close(xpuhip_session *sess);

/* One fragment of a batched launch: a scan + group-by session (no joins)
 * and its chunk already in device memory. */
typedef struct
{
    xpuhip_session *sess;
    const kern_data_store *d_kds_src;   /* KDS_FORMAT_ARROW, device address */
    uint32_t    nitems;                 /* rows of the chunk */
    uint32_t    __pad;
    kern_exec_results *results;         /* per-fragment statistics (may be NULL) */
} xpuhip_batch_task;

#define XPUHIP_BATCH_GRAPH  1   /* capture the launch sequence as a hipGraph; an
                                   identical later call replays it */

/* XpuTaskExec of many fragments in one persistent launch: the sessions must
 * share one plan shape (they may differ in parameter values, chunks and
 * kds_final).  Work items of all fragments are taken from one queue; each
 * fragment's kds_final accumulates exactly what xpuhip_task_exec_device of
 * that fragment alone would.  Launch shape from the first session's tuning.
 * Returns 0 or the first fragment's error (errbuf names the fragment).
 * Has no counterpart in the reference, which launches one kernel per chunk. */
int     xpuhip_batch_exec(const xpuhip_batch_task *tasks, uint32_t ntasks, uint32_t flags,
                          kern_errorbuf *errbuf);

/* Autotuner: runs the session's task on the caller's chunk (twice per
 * candidate: the first run compiles and warms) under each candidate launch
 * shape of its task kind - block size, rows per lane, LDS table entries,
 * blocks per CU - and keeps the fastest by kernel time.  The winner also
 * becomes the starting shape of later sessions of the same plan shape in the
 * process.  kds_final is reset afterwards.  d_dst/dst_length: the kds_dst of
 * a projection (NULL for group-by).  No counterpart in the reference. */
int     xpuhip_session_autotune(xpuhip_session *sess, const kern_data_store *d_kds_src,
                                uint32_t src_nitems, kern_data_store *d_kds_dst,
                                size_t dst_length, xpuhip_tuning *best, kern_errorbuf *errbuf);
/* The same over a whole batch (xpuhip_batch_exec without graph capture per
 * trial); every fragment's session gets the winner and is reset. */
int     xpuhip_batch_autotune(const xpuhip_batch_task *tasks, uint32_t ntasks,
                              xpuhip_tuning *best, kern_errorbuf *errbuf);

/* Clears a group-by session's kds_final (the session can run again from
 * scratch, e.g. after tuning trials). */
int     xpuhip_session_reset(xpuhip_session *sess, kern_errorbuf *errbuf);

/* ------------------------------------------------------------------
 * The GPU service as a request loop (continuous batching).  No counterpart
 * in the reference, whose service runs one task per worker thread
 * (gpu_service.c:5912-5993): requests from many backends are admitted in
 * arrival order as in-flight slots and the HBM budget allow, queued
 * group-by fragments of one plan shape are coalesced into one batched
 * launch, and admitted work overlaps on the device.  slots = 1 and
 * max_batch = 1 give the one-query-at-a-time service.
 * ------------------------------------------------------------------ */
typedef struct xpuhip_server xpuhip_server;

typedef struct
{
    uint64_t    hbm_budget;         /* bytes of requests admitted at once */
    uint32_t    slots;              /* jobs in flight (worker threads) */
    uint32_t    max_batch;          /* fragments coalesced into one launch */
} xpuhip_server_config;

#define XPUHIP_REQ_BATCHABLE    1   /* scan + group-by fragment (xpuhip_batch_exec) */

typedef struct
{
    xpuhip_session *sess;           /* the backend's session (kept open across requests) */
    const kern_data_store *d_kds_src;
    uint32_t    nitems;
    uint32_t    flags;              /* XPUHIP_REQ_* */
    kern_data_store *d_kds_dst;     /* projection output (device), or NULL */
    size_t      dst_length;
    uint64_t    hbm_bytes;          /* memory the request holds while admitted */
    uint64_t    shape;              /* requests coalesce only with equal shapes */
} xpuhip_request;

typedef struct
{
    int         rc;
    uint32_t    batched;            /* fragments of the launch it ran in (1: alone) */
    double      queue_ms;           /* submit -> admitted */
    double      exec_ms;            /* admitted -> done */
    double      total_ms;           /* submit -> done */
    kern_exec_results results;
    kern_errorbuf errbuf;
} xpuhip_request_status;

xpuhip_server *xpuhip_server_create(const xpuhip_server_config *cfg, kern_errorbuf *errbuf);
uint64_t xpuhip_server_submit(xpuhip_server *srv, const xpuhip_request *req);   /* ticket */
int     xpuhip_server_wait(xpuhip_server *srv, uint64_t ticket, xpuhip_request_status *status);
void    xpuhip_server_stats(xpuhip_server *srv, uint64_t *launches, uint64_t *coalesced);
void    xpuhip_server_destroy(xpuhip_server *srv);      /* drains, then stops */

/* Launch tuning of one session (see xpuhip_tuning). */
int     xpuhip_session_set_tuning(xpuhip_session *sess, const xpuhip_tuning *tuning);
int     xpuhip_session_get_tuning(xpuhip_session *sess, xpuhip_tuning *tuning);

/* Device memory helpers for callers that keep chunks resident in HBM. */
int     xpuhip_device_alloc(size_t length, void **p_dptr, kern_errorbuf *errbuf);
int     xpuhip_device_free(void *dptr);
int     xpuhip_memcpy_htod(void *dptr, const void *src, size_t length);
int     xpuhip_memcpy_dtoh(void *dst, const void *dptr, size_t length);
int     xpuhip_memcpy_dtod(void *dst, const void *src, size_t length);   /* copy kernel */
int     xpuhip_device_synchronize(void);

/* Kernel time of the last task executed by the calling thread, in
 * milliseconds, measured with HIP events on the stream it ran on. */
double  xpuhip_last_kernel_ms(void);

/* Launch shape and counters of the last task executed by the calling thread
 * (the reference reports the row counts in kern_exec_results; the rest has no
 * counterpart there and serves tuning and profiling). */
typedef struct
{
    double      kernel_ms;          /* as xpuhip_last_kernel_ms (all passes) */
    double      pass1_ms;           /* the pipeline / scan kernel alone */
    double      spill_ms;           /* the group-by spill pass alone */
    uint32_t    grid;               /* blocks launched */
    uint32_t    block;              /* threads per block */
    uint32_t    ntiles;             /* row tiles of the chunk */
    uint32_t    lds_nents;          /* group-by LDS table entries (0: none) */
    uint64_t    lds_bytes;          /* dynamic LDS per block */
    uint64_t    rows;               /* rows handed to the kernel */
    uint64_t    lds_miss;           /* group-by rows merged straight into kds_final */
    uint64_t    global_merges;      /* LDS entries merged into kds_final at block end */
    uint64_t    spill_rows;         /* rows aggregated by the second (spill) pass */
    uint32_t    jit;                /* 1: the kernel specialized for the plan ran */
    uint32_t    __pad;
    double      heap_ms;            /* heap-page / row chunk: its count, scan and
                                       deform passes (0 for Arrow chunks) */
    uint64_t    heap_rows;          /* tuples those passes loaded */
    double      sort_ms;            /* GPU-Sort of the last xpuhip_task_final:
                                       key, radix and permute passes */
    uint32_t    sort_items;         /* items it sorted (0: no sort) */
    uint32_t    sort_passes;        /* radix passes run (constant digits skipped) */
    uint32_t    join_slices;        /* slices of rows a multi-pass join ran in (0: none) */
    uint32_t    final_grows;        /* times the group-by kds_final grew during the task */
} xpuhip_task_stats;

int     xpuhip_last_task_stats(xpuhip_task_stats *stats);

/* Layout of this library's ABI structs as JSON (for conformance tests). */
const char *xpuhip_abi_layout_json(void);

#ifdef __cplusplus
}
#endif
#endif /* XPUHIP_H */
