/*
 * pgstrom_xpu.h
 *
 * Byte-compatible declarations of the PG-Strom xPU wire/device ABI that the
 * MI355X path accepts and produces: kern_data_store and its formats,
 * kern_tupitem / kern_hashitem, the kern_expression bytecode, the
 * kern_session_info blob, kern_exec_task / kern_exec_results, XpuCommand and
 * kern_multirels, plus the type/function opcode numbers this build executes.
 *
 * Every struct here mirrors the layout of the reference declaration cited next
 * to it (PG-Strom 6.2, src/xpu_common.h) so that a blob produced by the
 * reference's executor.c / codegen.c can be handed to this library unchanged.
 * The layouts are pinned by tests/golden/abi_layout.json, which is dumped from
 * the reference headers by oracle/ref_driver.cc.
 *
 * Plain C99: included by the C service loop, the C++ host runtime and the HIP
 * device code alike.  No PostgreSQL or CUDA headers are needed.
 */
#ifndef PGSTROM_XPU_H
#define PGSTROM_XPU_H
#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* inline helpers are usable from HIP device code as well */
#if defined(__HIPCC__)
#define XPU_INLINE      __host__ __device__ static inline
#else
#define XPU_INLINE      static inline
#endif

/* ---------------------------------------------------------------
 * Fundamental sizes and alignment (xpu_common.h:160-200)
 * --------------------------------------------------------------- */
#define XPU_BLCKSZ              8192
#define XPU_MAXIMUM_ALIGNOF     8
#define XPU_TYPEALIGN(A,LEN)    (((uint64_t)(LEN) + ((A) - 1)) & ~((uint64_t)((A) - 1)))
#define XPU_MAXALIGN(LEN)       XPU_TYPEALIGN(XPU_MAXIMUM_ALIGNOF,(LEN))
#define XPU_INTALIGN(LEN)       XPU_TYPEALIGN(4,(LEN))
#define XPU_BITMAPLEN(NATTS)    (((int)(NATTS) + 7) / 8)
#define XPU_NAMEDATALEN         64
#define XPU_L1_CACHELINE_SZ     128     /* group tuples never share a line */
#define XPU_MAXALIGNED          __attribute__((aligned(8)))

typedef uint32_t    xpu_oid_t;
typedef int64_t     gpumask_t;          /* xpu_common.h:208 */
typedef int32_t     TypeOpCode;         /* enum in xpu_common.h:383-392 */
typedef int32_t     FuncOpCode;         /* enum in xpu_common.h:397-431 */

typedef struct
{
    uint64_t    u64_lo;
    uint64_t    u64_hi;
} int128_packed_t;                      /* xpu_common.h:213-217 */

/* ---------------------------------------------------------------
 * Error / suspend codes and the error buffer (xpu_common.h:436-457)
 * --------------------------------------------------------------- */
#define ERRCODE_STROM_SUCCESS       0
#define ERRCODE_SUSPEND_FALLBACK    'f'
#define ERRCODE_SUSPEND_NO_SPACE    'd'
#define ERRCODE_DEVICE_ERROR        'E'
#define ERRCODE_DEVICE_FATAL        'F'
#define ERRCODE_IS_SUSPEND(x)       ((x) >= 'a' && (x) <= 'z')

#define KERN_ERRORBUF_FILENAME_LEN  32
#define KERN_ERRORBUF_FUNCNAME_LEN  64
#define KERN_ERRORBUF_MESSAGE_LEN   200
typedef struct
{
    uint32_t    errcode;
    int32_t     lineno;
    char        filename[KERN_ERRORBUF_FILENAME_LEN + 1];
    char        funcname[KERN_ERRORBUF_FUNCNAME_LEN + 1];
    char        message[KERN_ERRORBUF_MESSAGE_LEN + 1];
} kern_errorbuf;

/* ---------------------------------------------------------------
 * Arrow type options carried in kern_colmeta (arrow_defs.h:445-479)
 * --------------------------------------------------------------- */
#define ArrowType__Null             1
#define ArrowType__Int              2
#define ArrowType__FloatingPoint    3
#define ArrowType__Binary           4
#define ArrowType__Utf8             5
#define ArrowType__Bool             6
#define ArrowType__Decimal          7
#define ArrowType__Date             8
#define ArrowType__Decimal          7
#define ArrowType__Time             9
#define ArrowType__Timestamp        10
#define ArrowType__FixedSizeBinary  15
#define ArrowType__LargeBinary      19
#define ArrowType__LargeUtf8        20
#define ArrowPrecision__Half        0
#define ArrowPrecision__Single      1
#define ArrowPrecision__Double      2

typedef struct
{
    int32_t     tag;                    /* ArrowTypeTag */
    int16_t     unitsz;
    uint16_t    align;
    union {
        struct { uint16_t bitWidth; uint8_t is_signed; } integer;
        struct { int32_t precision; } floating_point;
        struct { uint16_t precision; uint16_t scale; uint16_t bitWidth; } decimal;
        struct { int32_t unit; } date;
        struct { int32_t unit; } time;
        struct { int32_t unit; } timestamp;
        struct { int32_t unit; } interval;
        struct { uint32_t byteWidth; } fixed_size_binary;
    } u;
} ArrowTypeOptions;

/* ---------------------------------------------------------------
 * kern_colmeta / kern_data_store (xpu_common.h:670-764)
 * --------------------------------------------------------------- */
#define TYPE_KIND__NULL         'n'
#define TYPE_KIND__BASE         'b'
#define TYPE_KIND__ARRAY        'a'
#define TYPE_KIND__COMPOSITE    'c'
#define TYPE_KIND__DOMAIN       'd'
#define TYPE_KIND__ENUM         'e'

typedef struct kern_colmeta
{
    bool        attbyval;
    int8_t      attalign;
    int16_t     attlen;
    int16_t     attnum;
    int16_t     attcacheoff;
    xpu_oid_t   atttypid;
    int32_t     atttypmod;
    int8_t      atttypkind;
    char        kds_format;
    int16_t     dtype_sizeof;
    uint32_t    kds_offset;
    uint16_t    idx_subattrs;
    uint16_t    num_subattrs;
    char        attname[XPU_NAMEDATALEN];
    ArrowTypeOptions attopts;
    int32_t     field_index;
    int64_t     virtual_offset;
    uint64_t    nullmap_offset;
    uint64_t    nullmap_length;
    uint64_t    values_offset;
    uint64_t    values_length;
    uint64_t    extra_offset;
    uint64_t    extra_length;
} kern_colmeta;

#define KDS_FORMAT_ROW          'r'
#define KDS_FORMAT_HASH         'h'
#define KDS_FORMAT_BLOCK        'b'
#define KDS_FORMAT_COLUMN       'c'
#define KDS_FORMAT_ARROW        'a'
#define KDS_FORMAT_PARQUET      'p'
#define KDS_FORMAT_FALLBACK     'f'

typedef struct kern_data_store
{
    uint64_t    length;
    uint64_t    usage;
    uint32_t    nitems;
    uint32_t    ncols;
    char        format;
    bool        has_varlena;
    bool        tdhasoid;
    xpu_oid_t   tdtypeid;
    int32_t     tdtypmod;
    xpu_oid_t   table_oid;
    uint32_t    hash_nslots;
    uint32_t    block_offset;
    uint32_t    block_nloaded;
    uint32_t    column_nrooms;
    uint32_t    arrow_virtual_usage;
    uint32_t    parquet_row_group;
    bool        parquet_parallel_load;
    bool        parquet_cache_enabled;
    uint32_t    nr_colmeta;
    kern_colmeta colmeta[1];
} kern_data_store;

/* ---------------------------------------------------------------
 * Tuple items (xpu_common.h:1199-1283)
 * --------------------------------------------------------------- */
#define HEAP_HASNULL            0x0001
#define HEAP_HASVARWIDTH        0x0002
#define HEAP_HASEXTERNAL        0x0004
#define HEAP_NATTS_MASK         0x07FF

typedef struct kern_tupitem
{
    uint32_t    t_len;          /* length up to and including the row-id */
    uint32_t    hash;
    uint8_t     has_xact_attrs;
    uint8_t     __padding__;
    uint16_t    t_infomask2;
    uint16_t    t_infomask;
    uint8_t     t_hoff;
    uint8_t     t_bits[1];
} kern_tupitem;

/* offsetof(HeapTupleHeaderData,t_infomask2) - offsetof(kern_tupitem,t_infomask2) */
#define MINIMAL_TUPLE_OFFSET    8
#define ROWID_SIZE              4
#define KERN_TUPITEM_T_BITS_OFFSET  15

typedef struct kern_hashitem
{
    uint64_t    next;           /* offset of the next item from the KDS tail */
    kern_tupitem t XPU_MAXALIGNED;
} kern_hashitem;

typedef struct kern_fallbackitem
{
    uint64_t    l_state;
    uint16_t    depth;
    uint8_t     matched;
    kern_tupitem t XPU_MAXALIGNED;
} kern_fallbackitem;

/* Transaction attributes embedded in a kern_tupitem of a KDS_FORMAT_ROW
 * source, in front of its payload (xpu_common.h:1199-1205, 1267-1283) */
typedef struct
{
    uint32_t    xmin;
    uint32_t    xmax;
    uint16_t    __padding__;
    uint16_t    ctid[3];        /* ItemPointerData: bi_hi, bi_lo, ip_posid */
} kern_tupitem_xact_attrs;

#define InvalidTransactionId    0U
#define FrozenTransactionId     2U

/* SerializedTransactionState at session_xact_state (xpu_common.h:3156-3165) */
typedef struct
{
    int32_t     xactIsoLevel;
    bool        xactDeferrable;
    uint64_t    topFullTransactionId;
    uint64_t    currentFullTransactionId;
    uint32_t    currentCommandId;
    int32_t     nParallelCurrentXids;
    uint32_t    parallelCurrentXids[1];
} SerializedTransactionState;

/* ---------------------------------------------------------------
 * Heap pages of a KDS_FORMAT_BLOCK chunk (xpu_common.h:1058-1134,
 * 1478-1483): nitems pages of BLCKSZ bytes at block_offset, their block
 * numbers (u32) at the KDS body.  A page is PageHeaderData (24 bytes) and
 * its ItemIdData line pointers; tuples are HeapTupleHeaderData.
 * --------------------------------------------------------------- */
#define XPU_PAGE_HEADER_SIZE    24      /* SizeOfPageHeaderData */
#define XPU_PAGE_PD_LOWER       12      /* offsetof(PageHeaderData, pd_lower) */
#define LP_UNUSED               0
#define LP_NORMAL               1
#define LP_REDIRECT             2
#define LP_DEAD                 3
/* ItemIdData bit fields: lp_off:15, lp_flags:2, lp_len:15 */
#define XPU_ITEMID_OFF(lp)      ((uint32_t)(lp) & 0x7fffU)
#define XPU_ITEMID_FLAGS(lp)    (((uint32_t)(lp) >> 15) & 0x3U)
#define XPU_ITEMID_LEN(lp)      ((uint32_t)(lp) >> 17)
/* HeapTupleHeaderData (xpu_common.h:906-938) */
#define XPU_HTUP_CTID           12
#define XPU_HTUP_INFOMASK2      18
#define XPU_HTUP_INFOMASK       20
#define XPU_HTUP_HOFF           22
#define XPU_HTUP_BITS           23
#define XPU_SIZEOF_HTUP_HEADER  23
/* size of an external datum (1-byte header 0x01, then its tag):
 * VARSIZE_EXTERNAL (xpu_common.h:1743-1751), 2 + sizeof(varatt_indirect)
 * or 2 + sizeof(varatt_external), 2 for other tags */
#define XPU_VARTAG_INDIRECT     1
#define XPU_VARTAG_ONDISK       18
XPU_INLINE uint32_t
XPU_VARSIZE_EXTERNAL(uint8_t tag)
{
    return tag == XPU_VARTAG_INDIRECT ? 10 : tag == XPU_VARTAG_ONDISK ? 18 : 2;
}

/* Hash-slot width chosen by the host for a KDS_FORMAT_HASH buffer
 * (xpu_common.h:1389-1396). */
XPU_INLINE uint64_t
KDS_GET_HASHSLOT_WIDTH(uint64_t nitems)
{
    if (nitems <= 5000)
        return 20000UL;
    if (nitems <= 4000000)
        return 20000UL + 2 * nitems;
    return 8020000UL + nitems;
}

XPU_INLINE uint64_t
KDS_HEAD_LENGTH(const kern_data_store *kds)
{
    return XPU_MAXALIGN(offsetof(kern_data_store, colmeta) +
                        sizeof(kern_colmeta) * kds->nr_colmeta);
}

/* ---------------------------------------------------------------
 * Opcodes.  Numbers are the enum values of the reference build with
 * WITH_POSTGIS=0 (Makefile.common:50), generated from xpu_opcodes.h by the
 * X-macros at xpu_common.h:381-431.  Only the opcodes this build executes
 * are listed; tests/test_abi.py checks each against the dump.
 * --------------------------------------------------------------- */
#define TypeOpCode__Invalid     0
#define TypeOpCode__bool        1
#define TypeOpCode__int1        2
#define TypeOpCode__int2        3
#define TypeOpCode__int4        4
#define TypeOpCode__int8        5
#define TypeOpCode__float2      6
#define TypeOpCode__float4      7
#define TypeOpCode__float8      8
#define TypeOpCode__numeric     9
#define TypeOpCode__bytea       10
#define TypeOpCode__text        11
#define TypeOpCode__bpchar      12
#define TypeOpCode__date        13
#define TypeOpCode__array       26      /* after the WITH_POSTGIS=0 types and composite */

#define FuncOpCode__ConstExpr               1
#define FuncOpCode__ParamExpr               2
#define FuncOpCode__VarExpr                 3
#define FuncOpCode__BoolExpr_And            4
#define FuncOpCode__BoolExpr_Or             5
#define FuncOpCode__BoolExpr_Not            6
#define FuncOpCode__NullTestExpr_IsNull     7
#define FuncOpCode__NullTestExpr_IsNotNull  8
#define FuncOpCode__BoolTestExpr_IsTrue     9
#define FuncOpCode__BoolTestExpr_IsNotTrue  10
#define FuncOpCode__BoolTestExpr_IsFalse    11
#define FuncOpCode__BoolTestExpr_IsNotFalse 12
#define FuncOpCode__BoolTestExpr_IsUnknown  13
#define FuncOpCode__BoolTestExpr_IsNotUnknown 14
#define FuncOpCode__DistinctFrom            15
#define FuncOpCode__CoalesceExpr            16
#define FuncOpCode__LeastExpr               17
#define FuncOpCode__GreatestExpr            18
#define FuncOpCode__CaseWhenExpr            19
#define FuncOpCode__ScalarArrayOpAny        20
#define FuncOpCode__ScalarArrayOpAll        21
#define FuncOpCode__HashedArrayOp           22
#define FuncOpCode__LoadVars                9999
#define FuncOpCode__MoveVars                10000
#define FuncOpCode__JoinQuals               10001
#define FuncOpCode__HashValue               10002
#define FuncOpCode__GiSTEval                10003
#define FuncOpCode__SaveExpr                10004
#define FuncOpCode__AggFuncs                10005
#define FuncOpCode__Projection              10006
#define FuncOpCode__SortKeys                10007
#define FuncOpCode__Packed                  10008

/* ---------------------------------------------------------------
 * kern_expression bytecode (xpu_common.h:2482-2692)
 * --------------------------------------------------------------- */
#define KERN_EXPRESSION_MAGIC       (0x4b657870)
#define KEXP_FLAG__IS_PUSHED_DOWN   0x0001U
#define SPECIAL_DEPTH__PREAGG_FINAL (-2)

typedef struct
{
    int16_t     vl_resno;
    int16_t     vl_slot_id;
} kern_varload_desc;

typedef struct
{
    int32_t     vm_offset;
    int16_t     vm_slot_id;
    bool        vm_from_xdatum;
} kern_varmove_desc;

typedef struct kern_varslot_desc
{
    TypeOpCode  vs_type_code;
    bool        vs_typbyval;
    int8_t      vs_typalign;
    int16_t     vs_typlen;
    int32_t     vs_typmod;
    int32_t     vs_offset;
    uint16_t    idx_subfield;
    uint16_t    num_subfield;
    const void *vs_ops;         /* resolved by the xPU service; ignored here */
} kern_varslot_desc;

typedef struct
{
    int16_t     fb_src_depth;
    int16_t     fb_src_resno;
    int16_t     fb_dst_resno;
    int16_t     fb_max_depth;
    uint16_t    fb_slot_id;
    int32_t     fb_kvec_offset;
} kern_fallback_desc;

typedef struct
{
    uint16_t    action;         /* KAGG_ACTION__* */
    int16_t     arg0_slot_id;
    int16_t     arg1_slot_id;
    int16_t     filter_slot_id;
    int32_t     typmod;
} kern_aggregate_desc;

/* ---------------------------------------------------------------
 * GPU-Sort key descriptors of FuncOpCode__SortKeys (xpu_common.h:2474-2520)
 * --------------------------------------------------------------- */
#define KSORT_KEY_ATTR__NULLS_FIRST         0x0400U
#define KSORT_KEY_ATTR__ORDER_ASC           0x8000U
#define KSORT_KEY_KIND__MASK                0x03ffU
#define KSORT_KEY_KIND__VREF                0
#define KSORT_KEY_KIND__PMINMAX_INT64       1
#define KSORT_KEY_KIND__PMINMAX_FP64        2
#define KSORT_KEY_KIND__PSUM_INT64          3
#define KSORT_KEY_KIND__PSUM_FP64           4
#define KSORT_KEY_KIND__PSUM_NUMERIC        5
#define KSORT_KEY_KIND__PAVG_INT64          6
#define KSORT_KEY_KIND__PAVG_FP64           7
#define KSORT_KEY_KIND__PAVG_NUMERIC        8
#define KSORT_KEY_KIND__PVARIANCE_SAMP      9
#define KSORT_KEY_KIND__NITEMS              23

#define KSORT_WINDOW_FUNC__ROW_NUMBER       'n'
#define KSORT_WINDOW_FUNC__RANK             'r'
#define KSORT_WINDOW_FUNC__DENSE_RANK       'd'

typedef struct
{
    uint16_t    kind;           /* KSORT_KEY_KIND__* */
    int8_t      nulls_first;
    int8_t      order_asc;
    uint16_t    src_anum;       /* attribute number in the kds_final tuple */
    uint16_t    buf_offset;     /* finalized-value offset past the tuple (unused here) */
    TypeOpCode  key_type_code;
    const void *key_ops;        /* resolved by the xPU service; ignored here */
} kern_sortkey_desc;

/* special values of a numeric sum state's attrs (xpu_common.h:2430-2434) */
#define __PAGG_NUMERIC_ATTRS__WEIGHT    0x00ffffU
#define __PAGG_NUMERIC_ATTRS__NAN       0x010000U
#define __PAGG_NUMERIC_ATTRS__PINF      0x020000U
#define __PAGG_NUMERIC_ATTRS__NINF      0x040000U
#define __PAGG_NUMERIC_ATTRS__MASK      0x070000U

typedef struct kern_expression
{
    uint32_t    len;
    TypeOpCode  exptype;
    const void *expr_ops;       /* resolved by the xPU service; ignored here */
    uint32_t    expflags;
    FuncOpCode  opcode;
    void       *fn_dptr;        /* resolved by the xPU service; ignored here */
    uint32_t    nr_args;
    uint32_t    args_offset;
    union {
        char    data[1] XPU_MAXALIGNED;
        struct {
            xpu_oid_t   const_type;
            bool        const_isnull;
            char        const_value[1] XPU_MAXALIGNED;
        } c;
        struct {
            uint32_t    param_id;
            char        __data[1];
        } p;
        struct {
            int32_t     var_offset;
            uint16_t    var_slot_id;
            char        __data[1];
        } v;
        struct {
            uint32_t    case_comp;
            uint32_t    case_else;
            char        data[1] XPU_MAXALIGNED;
        } casewhen;
        struct {
            uint16_t    elem_slot_id;
            char        data[1] XPU_MAXALIGNED;
        } saop;
        struct {
            const void *elem_ops;           /* device pointer (unused here) */
            int32_t     elem_type;          /* TypeOpCode of the keys */
            uint32_t    nslots;
            uint32_t    default_offset;     /* offset of the default value; 0: NULL */
            char        __padding[4];
            uint32_t    hslots[1] XPU_MAXALIGNED;
        } haop;
        struct {
            int32_t     depth;
            int32_t     nitems;
            kern_varload_desc desc[1];
        } load;
        struct {
            int32_t     depth;
            int32_t     nitems;
            kern_varmove_desc desc[1];
        } move;
        struct {
            uint16_t    sv_slot_id;
            char        data[1] XPU_MAXALIGNED;
        } save;
        struct {
            int32_t     nattrs;
            kern_aggregate_desc desc[1];
        } pagg;
        struct {
            uint32_t    hash;
            int32_t     nattrs;
            uint16_t    slot_id[1];
        } proj;
        struct {
            int32_t     nkeys;
            bool        needs_finalization;
            char        window_rank_func;       /* KSORT_WINDOW_FUNC__* or 0 */
            uint32_t    window_rank_limit;
            uint16_t    window_partby_nkeys;
            uint16_t    window_orderby_nkeys;
            kern_sortkey_desc desc[1];
        } sort;
        struct {
            uint32_t    npacked;
            uint32_t    offset[1];
        } pack;
    } u;
} kern_expression;

#define SizeOfKernExpr(PAYLOAD_SZ)  (offsetof(kern_expression, u.data) + (PAYLOAD_SZ))

/* items of a HashedArrayOp's hash table (xpu_common.h:2576-2594): 8-byte
 * aligned from the expression head; next carries flags in its low bits */
typedef struct
{
    uint32_t    hash;
    uint32_t    next;
    char        data[1] XPU_MAXALIGNED;
} kern_hashed_array_op_item;
#define KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY       0x01U
#define KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE     0x02U
#define KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK     0x07U
#define KERN_HASHED_ARRAY_OP_NITEMS_THRESHOLD   4

/* the session's database encoding (xpu_textlib.h:28-32) */
typedef struct
{
    char        encname[16];
    int32_t     enc_maxlen;
    void       *enc_mblen;          /* device pointer (unused here) */
} xpu_encode_info;
#define SizeOfKernExprVar           (offsetof(kern_expression, u.v.__data))

/* ---------------------------------------------------------------
 * Partial aggregation: actions, finals and state layouts
 * (xpu_common.h:2328-2471)
 * --------------------------------------------------------------- */
#define KAGG_ACTION__VREF           101
#define KAGG_ACTION__VREF_NOKEY     102
#define KAGG_ACTION__NROWS_ANY      201
#define KAGG_ACTION__NROWS_COND     202
#define KAGG_ACTION__PMIN_INT32     302
#define KAGG_ACTION__PMIN_INT64     303
#define KAGG_ACTION__PMIN_FP64      304
#define KAGG_ACTION__PMAX_INT32     402
#define KAGG_ACTION__PMAX_INT64     403
#define KAGG_ACTION__PMAX_FP64      404
#define KAGG_ACTION__PSUM_INT       501
#define KAGG_ACTION__PSUM_INT64     502
#define KAGG_ACTION__PSUM_FP        503
#define KAGG_ACTION__PSUM_NUMERIC   504
#define KAGG_ACTION__PAVG_INT       601
#define KAGG_ACTION__PAVG_INT64     602
#define KAGG_ACTION__PAVG_FP        603
#define KAGG_ACTION__PAVG_NUMERIC   604
#define KAGG_ACTION__STDDEV         701
#define KAGG_ACTION__COVAR          801

#define KAGG_FINAL__SIMPLE_VREF     1000
#define KAGG_FINAL__FMINMAX_INT32   1103
#define KAGG_FINAL__FMINMAX_INT64   1104
#define KAGG_FINAL__FMINMAX_FP64    1107
#define KAGG_FINAL__FSUM_INT        1200
#define KAGG_FINAL__FSUM_INT64      1201
#define KAGG_FINAL__FSUM_FP64       1203
#define KAGG_FINAL__FAVG_INT        1300
#define KAGG_FINAL__FAVG_INT64      1301
#define KAGG_FINAL__FAVG_FP64       1302

#define __PAGG_MINMAX_ATTRS__VALID  0x0001

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    int64_t     value;
} kagg_state__pminmax_int64_packed;

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    double      value;
} kagg_state__pminmax_fp64_packed;

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    int64_t     nitems;
    int64_t     sum;
} kagg_state__psum_int_packed;

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    int64_t     nitems;
    double      sum;
} kagg_state__psum_fp_packed;

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    uint64_t    nitems;
    int128_packed_t sum;
} kagg_state__psum_numeric_packed;

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    int64_t     nitems;
    double      sum_x;
    double      sum_x2;
} kagg_state__stddev_packed;

typedef struct
{
    int32_t     vl_len_;
    uint32_t    attrs;
    int64_t     nitems;
    double      sum_x;
    double      sum_xx;
    double      sum_y;
    double      sum_yy;
    double      sum_xy;
} kagg_state__covar_packed;

/* ---------------------------------------------------------------
 * Task flags (xpu_common.h:2197-2223)
 * --------------------------------------------------------------- */
#define DEVKIND__NVIDIA_GPU         0x00000001U     /* "a GPU"; the reference has one kind */
#define DEVKIND__ANY                0x00000001U
#define DEVTASK__PINNED_HASH_RESULTS 0x00001000U
#define DEVTASK__PINNED_ROW_RESULTS 0x00002000U
#define DEVTASK__MERGE_FINAL_BUFFER 0x00020000U
#define DEVTASK__SCAN               0x10000000U
#define DEVTASK__JOIN               0x20000000U
#define DEVTASK__PREAGG             0x40000000U
#define DEVTASK__MASK               0x70000000U

/* ---------------------------------------------------------------
 * Commands and the session blob (xpu_common.h:2762-2919)
 * --------------------------------------------------------------- */
#define XpuCommandTag__Success              0
#define XpuCommandTag__Error                1
#define XpuCommandTag__SuccessHalfWay       2
#define XpuCommandTag__OpenSession          100
#define XpuCommandTag__XpuTaskExec          110
#define XpuCommandTag__XpuTaskExecGpuCache  111
#define XpuCommandTag__XpuTaskFinal         119
#define XpuCommandMagicNumber               0xdeadbeafU

typedef struct kern_session_info
{
    uint64_t    query_plan_id;
    uint32_t    kcxt_kvars_nrooms;
    uint32_t    kcxt_kvars_nslots;
    uint32_t    kcxt_kvars_defs;
    uint32_t    kcxt_kvecs_bufsz;
    uint32_t    kcxt_kvecs_ndims;
    uint32_t    kcxt_extra_bufsz;
    uint32_t    cuda_stack_size;
    uint32_t    xpu_task_flags;
    gpumask_t   optimal_gpus;
    uint32_t    xpucode_load_vars_packed;
    uint32_t    xpucode_move_vars_packed;
    uint32_t    xpucode_scan_quals;
    uint32_t    xpucode_join_quals_packed;
    uint32_t    xpucode_hash_values_packed;
    uint32_t    xpucode_gist_evals_packed;
    uint32_t    xpucode_projection;
    uint32_t    xpucode_groupby_keyhash;
    uint32_t    xpucode_groupby_keyload;
    uint32_t    xpucode_groupby_keycomp;
    uint32_t    xpucode_groupby_actions;
    uint32_t    xpucode_gpusort_keydesc;
    int64_t     hostEpochTimestamp;
    uint32_t    session_xact_state;
    uint32_t    session_timezone;
    uint32_t    session_encode;
    int32_t     session_currency_frac_digits;
    uint64_t    pinned_inner_buffer_partition_size;
    uint32_t    projection_kds_dst;
    uint32_t    select_into_pathname;
    uint32_t    select_into_kds_head;
    uint32_t    select_into_projdesc;
    uint32_t    pgsql_port_number;
    uint32_t    pgsql_plan_node_id;
    uint32_t    join_inner_handle;
    uint32_t    groupby_kds_final;
    uint32_t    groupby_prepfn_bufsz;
    float       groupby_ngroups_estimation;
    uint32_t    gpusort_htup_margin;
    int32_t     gpusort_limit_count;
    uint32_t    fallback_kds_head;
    uint32_t    fallback_desc_defs;
    uint32_t    fallback_desc_nitems;
    uint32_t    nparams;
    uint32_t    poffset[1];
} kern_session_info;

typedef struct
{
    int32_t     scan_relidx;
    uint32_t    kds_src_pathname;
    uint32_t    kds_src_iovec;
    uint32_t    kds_src_offset;
    char        data[1] XPU_MAXALIGNED;
} kern_exec_task;

typedef struct
{
    uint32_t    nitems_roj;
    uint32_t    nitems_gist;
    uint32_t    nitems_out;
} kern_exec_results_stat;

typedef struct
{
    int32_t     scan_relidx;
    uint32_t    chunks_offset;
    uint32_t    chunks_nitems;
    uint32_t    ojmap_offset;
    uint32_t    ojmap_length;
    bool        right_outer_join;
    bool        final_plan_task;
    uint32_t    final_nitems;
    uint64_t    final_usage;
    uint64_t    final_total;
    uint32_t    final_sorting_msec;
    uint32_t    final_reconstruction_msec;
    uint32_t    join_reconstruction_msec;
    uint32_t    select_into_nblocks;
    uint32_t    npages_direct_read;
    uint32_t    npages_vfs_read;
    uint32_t    nchunks_parquet_read;
    uint32_t    ncaches_parquet_load;
    uint32_t    nitems_raw;
    uint32_t    nitems_in;
    uint32_t    nitems_out;
    uint32_t    num_rels;
    kern_exec_results_stat stats[1];
} kern_exec_results;

typedef struct xpu_dlist_node
{
    struct xpu_dlist_node *prev;
    struct xpu_dlist_node *next;
} xpu_dlist_node;

typedef struct
{
    uint32_t    magic;
    uint32_t    tag;
    uint64_t    length;
    void       *priv;
    xpu_dlist_node chain;
    union {
        kern_errorbuf       error;
        kern_session_info   session;
        kern_exec_task      task;
        kern_exec_results   results;
    } u;
} XpuCommand;

/* ---------------------------------------------------------------
 * Inner relations for hash joins (xpu_common.h:3486-3518)
 * --------------------------------------------------------------- */
typedef struct
{
    void       *kds_in;         /* device/host pointer slots used by the service */
    void       *kbuf_parts;
    uint64_t    kds_offset;     /* offset of the inner KDS from the kmrels head */
    uint64_t    ojmap_offset;
    uint64_t    gist_offset;
    int16_t     gist_ctid_resno;
    bool        is_nestloop;
    bool        left_outer;
    bool        right_outer;
    bool        pinned_buffer;
    uint64_t    buffer_id;
} kern_multirels_chunk;

typedef struct kern_multirels
{
    uint64_t    length;
    uint64_t    ojmap_sz;
    uint64_t    kbuf_part_offset;
    uint32_t    num_inner_rels;
    kern_multirels_chunk chunks[1];
} kern_multirels;

/* ---------------------------------------------------------------
 * PostgreSQL hash combination used for join/group keys
 * (xpu_common.h:3472-3476)
 * --------------------------------------------------------------- */
XPU_INLINE uint32_t
pg_hash_merge(uint32_t hash_prev, uint32_t hash_next)
{
    return ((hash_prev >> 3) | (hash_prev << 29)) ^ hash_next;
}

#ifdef __cplusplus
}
#endif
#endif /* PGSTROM_XPU_H */
