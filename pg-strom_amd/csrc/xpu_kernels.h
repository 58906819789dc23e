/*
 * xpu_kernels.h - launch parameters and the per-task control block shared by
 * the HIP kernels (xpu_kernels.hip) and the host runtime (xpuhip_runtime.cpp).
 */
#ifndef XPU_KERNELS_H
#define XPU_KERNELS_H
#include <stdint.h>

#include "pgstrom_xpu.h"
#include "xpu_program.h"

/* look-back tile status words: flag(2) | rows(30) | bytes/8 (32) */
#define XLB_FLAG_SHIFT      62
#define XLB_AGG             1ULL    /* tile total published */
#define XLB_PREFIX          2ULL    /* inclusive prefix published */
#define XLB_STOP            3ULL    /* this or an earlier tile overflowed / aborted */
#define XLB_ROWS_MASK       0x3fffffffULL
#define XLB_SPIN_LIMIT      (1U << 26)

#define XLDS_EMPTY_KEY      0xffffffffffffffffULL
#define XGB_LOCKED_SLOT     0xffffffffffffffffULL
#define XSPILL_CHUNK        512     /* spill records a wave reserves at once */
#define XP_DUMP_MAXPARTS    256     /* partial-state dump partitions at most */
#define XP_BLOCK_MAX        512     /* threads per block at most */
#define XP_AGG_BLOCK_MAX    1024    /* ... of the scan + group-by kernels (one block per
                                       CU sharing the CU's LDS table) */
/* group-by joins: 64-row batches a wave screens before running the joins
 * over the rows that pass, and the static LDS of the pipeline kernel (the
 * screened-row queues and the block's counters) */
#define XQ_BATCHES          8
#define XQ_WAVE             (64 * XQ_BATCHES)
#define XP_PIPELINE_STATIC_LDS  ((XP_BLOCK_MAX / 64) * XQ_WAVE * 4 + 1024)
/* multi-pass joins: threads per block of the screen pass (wide blocks: the
 * screen holds few registers, and its blocks match the join passes' blocks
 * one to one) */
#define XQ_SCREEN_BLOCK     1024
#define XQ_SCREEN_RPT       4       /* 64-row batches per wave tile */
/* scan + projection: 64-row batches per wave tile */
#define XP_PROJ_RPT         4

/* per-task control block in device memory (cleared before each launch) */
typedef struct
{
    uint32_t    tile_ticket;        /* dynamic tile counter */
    uint32_t    resume_tile;        /* ~(first tile that did not fit); 0: none */
    uint32_t    errcode;            /* first error of the task */
    uint32_t    errmsg;             /* message id in the session's table */
    uint32_t    nitems_in;          /* rows that passed the scan quals */
    uint32_t    nitems_out;         /* rows reaching the final stage */
    uint32_t    depth_out[XP_MAX_DEPTH + 1];
    uint32_t    final_ovf;          /* a group went to the overflow table */
    uint64_t    out_nitems;         /* projection rows written */
    uint64_t    out_usage;          /* projection bytes written */
    uint64_t    lds_miss;           /* group-by rows sent straight to kds_final */
    uint64_t    global_merges;      /* group-by merges into kds_final */
    uint64_t    spill_next;         /* spill records reserved (may pass the capacity) */
    uint64_t    spill_rows;         /* rows written to the spill buffer */
} xpu_task_ctl;

/* error messages the kernels raise themselves (beyond the compiled table) */
#define XMSG_LOOKBACK_TIMEOUT   0xfff0
#define XMSG_FINAL_NO_SPACE     0xfff1
#define XMSG_TEXT_EXTERNAL      0xfff2
#define XMSG_GROUP_RETRY        0xfff3
#define XMSG_TILE_TOO_LARGE     0xfff4
#define XMSG_JOIN_RECORDS       0xfff5
#define XMSG_HEAP_ITEM          0xfff6
#define XMSG_HEAP_TUPLE         0xfff7
#define XMSG_ARROW_TYPE         0xfff8
#define XMSG_FALLBACK_FULL      0xfff9
/* numeric errors, with the reference's texts (xpu_numeric.h/.cu) */
#define XMSG_NUM_RANGE          0xffe0  /* numeric value is out of range */
#define XMSG_NUM_HEADER         0xffe1  /* corrupted numeric header */
#define XMSG_NUM_SPECIAL        0xffe2  /* unknown special numeric value */
#define XMSG_NUM_DIV0           0xffe3  /* division by zero */
#define XMSG_NUM_INT_RANGE      0xffe4  /* integer out of range */
#define XMSG_NUM_NANINF         0xffe5  /* cannot convert NaN/Inf to integer */
#define XMSG_NUM_FLOAT_RANGE    0xffe6  /* float out of range */
/* the longest varlena an Arrow column may hold (VARATT_MAX, xpu_common.h:1575) */
#define XARROW_VARATT_MAX       0x4ffffff8U

/* ------------------------------------------------------------------
 * Heap-page (KDS_FORMAT_BLOCK) and row (KDS_FORMAT_ROW / HASH) sources are
 * read into column tiles before the pipeline: a count pass over the pages
 * (or 64-row units), a scan of the unit counts, and a deform pass writing
 * the referenced attributes of every visible tuple, in page / row-index
 * order, into a KDS_FORMAT_ARROW buffer whose layout the host sets up.
 * Fixed-width attributes become Arrow values, booleans bit-packed values;
 * text/bytea attributes become (offset, length) pairs of u32 that address
 * the datum's bytes inside the source chunk (no bytes are copied): such a
 * column carries attopts.tag XARROW_TAG_HEAPREF and its extra_offset is the
 * source chunk's address minus the tile buffer's, modulo 2^64.  A length of
 * XHEAP_LEN_EXTERNAL marks a compressed or external datum, whose offset
 * addresses its varlena header.
 * ------------------------------------------------------------------ */
#define XARROW_TAG_HEAPREF      0x4845      /* not an ArrowTypeTag */
#define XHEAP_LEN_EXTERNAL      0xffffffffU
#define XHEAP_BLOCK             64          /* row chunks: one wave per 64-row unit */
#define XHEAP_DEFORM_BLOCK      256         /* heap pages: four persistent waves */
#define XHEAP_COUNT_BLOCK       256
#define XHEAP_SCAN_BLOCK        1024

enum { XH_SKIP = 0, XH_FIXED = 1, XH_BOOL = 2, XH_VAR = 3 };

/* one source attribute up to the last referenced one */
typedef struct
{
    int16_t     attlen;
    int8_t      attalign;
    uint8_t     kind;               /* XH_* */
    uint32_t    width;              /* XH_FIXED: 1, 2, 4 or 8 */
    uint64_t    values_off;         /* from the tile buffer */
    uint64_t    nullmap_off;        /* validity bits (u32 words, zeroed) */
} xheap_col;

typedef struct
{
    uint32_t    total;              /* visible tuples of the chunk */
    uint32_t    errcode;
    uint32_t    errmsg;
    uint32_t    __pad;
} xheap_status;

typedef struct
{
    const kern_data_store *src;     /* device address of the source chunk */
    char       *dst;                /* tile buffer (KDS_FORMAT_ARROW) */
    const xheap_col *cols;          /* [natts_walk] */
    uint32_t   *units;              /* [nunits + 1]: counts, then exclusive prefix */
    xheap_status *st;
    const uint32_t *xids;           /* parallelCurrentXids of the session */
    uint64_t    src_length;
    uint64_t    rowindex_off;       /* ROW / HASH: row index from the source head */
    uint32_t    format;             /* KDS_FORMAT_BLOCK, _ROW or _HASH */
    uint32_t    nitems;             /* pages (BLOCK) or tuples */
    uint32_t    nunits;             /* pages, or 64-tuple units */
    uint32_t    block_offset;
    uint32_t    src_ncols;
    uint32_t    natts_walk;
    uint32_t    nxids;
    uint32_t    capacity;           /* rows the tile buffer holds (deform pass) */
} xheap_params;

/* one fragment of a batched launch */
typedef struct
{
    const kern_data_store *kds_src;     /* KDS_FORMAT_ARROW in device memory */
    kern_data_store *kds_final;         /* the fragment's group-by result */
    xpu_task_ctl *ctl;                  /* the fragment's control block */
    const char *const_pool;
    const int64_t *params;              /* XO_PARAM values */
    uint64_t   *spill;                  /* the fragment's spill region */
    uint64_t    spill_cap;              /* records */
    uint64_t    param_nulls;
    uint32_t    nrows;
    uint32_t    __pad;
} xbatch_frag;

/* a row range of one fragment (pass 1) */
typedef struct
{
    uint32_t    frag;
    uint32_t    row_begin;
    uint32_t    row_end;
    uint32_t    __pad;
} xbatch_item;

/* a hash partition of a slice of one fragment's spill records (pass 2) */
typedef struct
{
    uint32_t    frag;
    uint16_t    part;
    uint16_t    nparts;
    uint64_t    lo;
    uint64_t    hi;
} xbatch_sitem;

#define XBATCH_SPILL_SLICE  (2ULL << 20)    /* spill records per pass-2 item */

typedef struct
{
    const xplan *plan;
    const kern_data_store *kds_src;     /* KDS_FORMAT_ARROW in device memory */
    kern_data_store *kds_out;           /* kds_dst (projection) / kds_final (group-by) */
    xpu_task_ctl *ctl;
    uint64_t   *tile_status;
    const char *const_pool;
    uint32_t    row_begin;              /* first row of this launch */
    uint32_t    row_end;                /* one past the last row */
    uint32_t    rows_per_tile;
    uint32_t    ntiles;
    uint32_t    lds_nents;              /* group-by LDS table entries (0: off) */
    uint32_t    __pad;
    uint64_t    inner_kds[XP_MAX_DEPTH + 1];    /* device address of the inner KDS of
                                                   each depth (index 1..nrels) */
    uint64_t   *spill;                  /* group-by spill records (nullptr: none) */
    uint64_t    spill_cap;              /* capacity in records */
    const int64_t *params;              /* XO_PARAM values of this launch */
    uint64_t    param_nulls;            /* bit i: parameter i is NULL */
    /* batched launches */
    const xbatch_frag *batch_frags;
    const xbatch_item *batch_items;
    xbatch_sitem *batch_sitems;
    uint32_t   *batch_ticket;           /* [0] pass-1 ticket, [1] pass-2 ticket,
                                           [2] pass-2 items */
    uint32_t    batch_nfrags;
    uint32_t    batch_nitems;
    uint32_t    batch_nsitems_cap;
    uint32_t    dump_nparts;            /* partial-state dump partitions (0: none) */
    uint64_t   *dump;                   /* partial-state records, [part][dump_cap] */
    uint64_t    dump_cap;               /* records per partition */
    uint64_t   *dump_next;              /* [dump_nparts] records reserved */
    /* join-hash Bloom filters of the inner relations (nullptr: none) */
    const uint64_t *bloom[XP_MAX_DEPTH + 1];
    uint32_t    bloom_mask[XP_MAX_DEPTH + 1];   /* words - 1 (words: a power of two) */
    /* the depth-1 filter folded to bloom_lds_words words (a power of two),
     * which the screen pass holds in LDS: a row it rejects is rejected by the
     * full filter too (nullptr: none) */
    const uint64_t *bloom_lds;
    uint32_t    bloom_lds_words;
    uint32_t    prefilter;              /* group-by joins: screen rows before the joins */
    /* multi-pass joins of a group-by: the records a pass reads and writes,
     * in per-block regions of rec_*_cap records (column-major words) */
    uint32_t   *rec_in;
    uint32_t   *rec_out;
    uint32_t   *rec_in_cnt;             /* [grid] records of each region */
    uint32_t   *rec_out_cnt;
    uint32_t    rec_in_cap;
    uint32_t    rec_out_cap;
    uint32_t    pass_depth;             /* 0: screen pass; d: join with inner relation d */
    uint32_t    screened;               /* level-1 records passed the depth-1 filter */
    /* scan + projection: rows and bytes of each count-pass block */
    uint64_t   *block_status;
    /* group-by: the session's overflow table (KDS_FORMAT_HASH, nullptr:
     * none), which takes the groups that find no room in kds_final; the
     * host grows kds_final and merges it in after the launch */
    kern_data_store *kds_ovf;
    /* the session's CPU-fallback buffer (KDS_FORMAT_FALLBACK, nullptr: none) */
    kern_data_store *kds_fallback;
} xpu_kparams;

#ifdef __cplusplus
extern "C" {
#endif
/* host-side launchers (defined in xpu_kernels.hip) */
int xpu_launch_pipeline(const xpu_kparams *kp, int nregs, int grid, int block,
                        size_t lds_bytes, bool preagg, void *stream);
/* projection in two passes: count (tile totals), tile scan, write */
int xpu_launch_projection(const xpu_kparams *kp, int nregs, int grid, int block, void *stream);
int xpu_launch_tile_scan(const xpu_kparams *kp, void *stream);
/* scan + projection: the tile scan over the count pass's wave tiles (one
 * block per count-pass block) */
int xpu_launch_tile_scan_blocks(const xpu_kparams *kp, int grid, void *stream);
/* folds the partial-state dumps of a scan + group-by task into kds_final */
int xpu_launch_dump_merge(const xpu_kparams *kp, int nregs, int grid, int block,
                          size_t lds_bytes, void *stream);
/* merges the groups of a session's overflow table into kds_final (grown
 * to hold them); the grid covers ovf_nitems groups */
int xpu_launch_final_merge_ovf(const xplan *plan, int nregs, kern_data_store *kds_final,
                               const kern_data_store *kds_ovf, uint32_t ovf_nitems,
                               xpu_task_ctl *ctl, void *stream);
/* the spill setup kernel of a batch alone (for specialized pass kernels) */
int xpu_launch_batch_setup(const xpu_kparams *kp, void *stream);
/* batched fragments: pass 1, spill setup, pass 2 (all stream-ordered) */
int xpu_launch_batch(const xpu_kparams *kp, int nregs, int rpt, int grid, int block,
                     size_t lds_bytes, void *stream);
/* second pass over the spill records of a scan + group-by launch */
int xpu_launch_spill_agg(const xpu_kparams *kp, int nregs, int grid, int block,
                         size_t lds_bytes, void *stream);
/* scan + group-by without joins; rpt = rows per lane per wave tile (1, 2, 4) */
/* sets the Bloom filter bits of the join hashes of one inner KDS_FORMAT_HASH
 * (words zeroed by the caller) and raises *maxdup to the most items that
 * share a join hash */
int xpu_launch_bloom_build(kern_data_store *inner, uint64_t nslots, uint64_t *words,
                           uint32_t mask, uint32_t *maxdup, void *stream);
/* folds a filter of nwords words into fwords words (both powers of two):
 * folded word i is the OR of words i, i + fwords, ... */
int xpu_launch_bloom_fold(const uint64_t *words, uint64_t nwords, uint64_t *folded,
                          uint64_t fwords, void *stream);
#define XQ_SCREEN_LDS_WORDS (16384)     /* 128 KB of folded depth-1 filter per screen block */
int xpu_launch_scan_agg(const xpu_kparams *kp, int nregs, int rpt, int grid, int block,
                        size_t lds_bytes, void *stream);
int xpu_launch_fill(void *ptr, uint64_t value, size_t nwords, void *stream);
/* pinned inner buffers: append a ROW chunk's rows (nrows: an upper bound
 * for the grid) to a KDS_FORMAT_HASH buffer (base[2]: reservation, see the
 * kernel), and re-chain a copied buffer into its own slots */
int xpu_launch_hash_append(kern_data_store *hk, const kern_data_store *rk, uint64_t *base,
                           uint32_t nrows, void *stream);
int xpu_launch_hash_rechain(kern_data_store *hk, uint32_t nitems, void *stream);
/* heap / row sources: count pass, unit scan (total into st->total), deform pass */
/* ---- GPU-Sort of a final buffer (xpu_sort.hip) ----
 * Each sort key becomes one to three 64-bit words whose unsigned order is
 * the key's order (NULL placement, ASC/DESC folded in); items are sorted by
 * their word tuples with a stable LSD radix sort of 8-bit digits. */
#define XSORT_MAX_KEYS      8
#define XSORT_MAX_WORDS     (3 * XSORT_MAX_KEYS)
#define XSORT_BLOCK         256
#define XSORT_ROUNDS        16                      /* items per thread of a tile */
#define XSORT_TILE          (XSORT_BLOCK * XSORT_ROUNDS)

typedef struct
{
    uint16_t    kind;           /* KSORT_KEY_KIND__* */
    uint8_t     nulls_first;
    uint8_t     order_asc;
    uint16_t    anum;           /* 1-based attribute of the kds_final tuple */
    uint8_t     word;           /* first key word */
    uint8_t     nwords;         /* 1 (<= 32-bit value), 2 (64-bit), 3 (128-bit) */
    int32_t     type_code;      /* TypeOpCode of a VREF key */
} xsort_key;

typedef struct
{
    const kern_data_store *kds;     /* the final buffer (ROW or HASH) */
    const uint64_t *rowindex;       /* its row index */
    uint32_t    n;                  /* items */
    uint32_t    nkeys;
    uint32_t    nwords;
    uint32_t    max_anum;           /* last attribute any key reads */
    xsort_key   keys[XSORT_MAX_KEYS];
    uint64_t   *words;              /* [nwords][n] */
    uint32_t   *idx;                /* [n] item numbers */
    uint64_t   *wmin;               /* [nwords] least value over the items */
    uint64_t   *wmax;               /* [nwords] greatest value over the items */
} xsort_params;

int xpu_launch_sort_keys(const xsort_params *sp, void *stream);
int xpu_launch_sort_pass(const uint64_t *words_in, const uint32_t *idx_in, uint64_t *words_out,
                         uint32_t *idx_out, uint32_t n, uint32_t nwords, uint32_t word,
                         uint64_t base, uint32_t shift, uint32_t *counts,
                         void *stream);          /* counts: 256 * (ntiles + 1) */
int xpu_launch_sort_permute(const uint64_t *rowindex, const uint32_t *idx, uint64_t *out,
                            uint32_t n, void *stream);
int xpu_launch_window_rank(const uint64_t *words, const uint64_t *ridx, uint32_t n, uint32_t pw,
                           uint32_t ow, int func, uint32_t limit, uint32_t *flag, uint32_t *cpeer,
                           uint32_t *keep, uint32_t *pos, uint32_t *tile_sum, uint32_t *nkept,
                           uint64_t *ridx_out, void *stream);
int xpu_launch_sort_limit_size(const kern_data_store *kds, const uint64_t *ridx, uint32_t nout,
                               uint64_t *offs, uint64_t *used, void *stream);
int xpu_launch_sort_limit_copy(const kern_data_store *kds, const uint64_t *ridx, uint32_t nout,
                               const uint64_t *offs, char *stage, uint64_t stage_len,
                               uint64_t *ridx_out, void *stream);

int xpu_launch_heap_count(const xheap_params *hp, void *stream);
int xpu_launch_heap_deform(const xheap_params *hp, int ncu, void *stream);
int xpu_launch_copy(void *dst, const void *src, size_t nbytes, void *stream);
int xpu_launch_finalize_dst(kern_data_store *kds_dst, const kern_data_store *head,
                            uint32_t head_len, uint64_t length, const xpu_task_ctl *ctl,
                            void *stream);
#ifdef __cplusplus
}
#endif
#endif
