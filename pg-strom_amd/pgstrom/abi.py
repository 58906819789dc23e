"""ctypes mirror of include/pgstrom_xpu.h (the PG-Strom xPU ABI).

The layouts follow the reference declarations in src/xpu_common.h
(kern_colmeta :670-729, kern_data_store :739-764, kern_tupitem :1207-1237,
kern_hashitem :1273-1277, kern_expression :2596-2692, kern_session_info
:2775-2837, kern_exec_results :2856-2885, kern_multirels :3494-3518) and are
checked against tests/golden/abi_layout.json by tests/test_abi.py.
"""
import ctypes as C

MAXALIGN = 8
NAMEDATALEN = 64


def typealign(a, n):
    return (n + a - 1) & ~(a - 1)


def maxalign(n):
    return typealign(8, n)


def bitmaplen(n):
    return (n + 7) // 8


# --- error codes (xpu_common.h:436-441)
ERRCODE_STROM_SUCCESS = 0
ERRCODE_SUSPEND_FALLBACK = ord('f')
ERRCODE_SUSPEND_NO_SPACE = ord('d')
ERRCODE_DEVICE_ERROR = ord('E')
ERRCODE_DEVICE_FATAL = ord('F')

# --- KDS formats
KDS_FORMAT_ROW = b'r'
KDS_FORMAT_HASH = b'h'
KDS_FORMAT_BLOCK = b'b'
KDS_FORMAT_COLUMN = b'c'
KDS_FORMAT_ARROW = b'a'
KDS_FORMAT_FALLBACK = b'f'

TYPE_KIND__BASE = ord('b')

ArrowType__Int = 2
ArrowType__FloatingPoint = 3
ArrowType__Utf8 = 5
ArrowType__Bool = 6
ArrowType__Decimal = 7
ArrowPrecision__Single = 1
ArrowPrecision__Double = 2

HEAP_HASNULL = 0x0001
HEAP_HASVARWIDTH = 0x0002
HEAP_HASEXTERNAL = 0x0004
HEAP_NATTS_MASK = 0x07FF
MINIMAL_TUPLE_OFFSET = 8
ROWID_SIZE = 4
T_BITS_OFFSET = 15

KERN_EXPRESSION_MAGIC = 0x4b657870
KEXP_FLAG__IS_PUSHED_DOWN = 0x0001
KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY = 0x01
KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE = 0x02
KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK = 0x07
SPECIAL_DEPTH__PREAGG_FINAL = -2

DEVKIND__NVIDIA_GPU = 0x00000001
DEVTASK__PINNED_HASH_RESULTS = 0x00001000
DEVTASK__SCAN = 0x10000000
DEVTASK__JOIN = 0x20000000
DEVTASK__PREAGG = 0x40000000

XpuCommandTag__Success = 0
XpuCommandTag__Error = 1
XpuCommandTag__SuccessHalfWay = 2
XpuCommandTag__OpenSession = 100
XpuCommandTag__XpuTaskExec = 110
XpuCommandTag__XpuTaskFinal = 119
XpuCommandMagicNumber = 0xdeadbeaf

# GPU-Sort key descriptors (xpu_common.h:2474-2520)
KSORT_KEY_ATTR__NULLS_FIRST = 0x0400
KSORT_KEY_ATTR__ORDER_ASC = 0x8000
KSORT_KEY_KIND__VREF = 0
KSORT_KEY_KIND__PMINMAX_INT64 = 1
KSORT_KEY_KIND__PMINMAX_FP64 = 2
KSORT_KEY_KIND__PSUM_INT64 = 3
KSORT_KEY_KIND__PSUM_FP64 = 4
KSORT_KEY_KIND__PSUM_NUMERIC = 5
KSORT_KEY_KIND__PAVG_INT64 = 6
KSORT_KEY_KIND__PAVG_FP64 = 7
KSORT_KEY_KIND__PAVG_NUMERIC = 8
KSORT_KEY_KIND__PVARIANCE_SAMP = 9
KSORT_WINDOW_FUNC__ROW_NUMBER = ord("n")
KSORT_WINDOW_FUNC__RANK = ord("r")
KSORT_WINDOW_FUNC__DENSE_RANK = ord("d")

KAGG_ACTION__VREF = 101
KAGG_ACTION__VREF_NOKEY = 102
KAGG_ACTION__NROWS_ANY = 201
KAGG_ACTION__NROWS_COND = 202
KAGG_ACTION__PMIN_INT32 = 302
KAGG_ACTION__PMIN_INT64 = 303
KAGG_ACTION__PMIN_FP64 = 304
KAGG_ACTION__PMAX_INT32 = 402
KAGG_ACTION__PMAX_INT64 = 403
KAGG_ACTION__PMAX_FP64 = 404
KAGG_ACTION__PSUM_INT = 501
KAGG_ACTION__PSUM_INT64 = 502
KAGG_ACTION__PSUM_FP = 503
KAGG_ACTION__PSUM_NUMERIC = 504
KAGG_ACTION__PAVG_INT = 601
KAGG_ACTION__PAVG_INT64 = 602
KAGG_ACTION__PAVG_FP = 603
KAGG_ACTION__PAVG_NUMERIC = 604
KAGG_ACTION__STDDEV = 701
KAGG_ACTION__COVAR = 801

__PAGG_MINMAX_ATTRS__VALID = 0x0001
__PAGG_NUMERIC_ATTRS__WEIGHT = 0x00ffff
__PAGG_NUMERIC_ATTRS__NAN = 0x010000
__PAGG_NUMERIC_ATTRS__PINF = 0x020000
__PAGG_NUMERIC_ATTRS__NINF = 0x040000


class kern_errorbuf(C.Structure):
    _fields_ = [("errcode", C.c_uint32), ("lineno", C.c_int32),
                ("filename", C.c_char * 33), ("funcname", C.c_char * 65),
                ("message", C.c_char * 201)]

    def describe(self):
        return "errcode=%r %s:%d %s: %s" % (
            chr(self.errcode) if 32 <= self.errcode < 127 else self.errcode,
            self.filename.decode(errors="replace"), self.lineno,
            self.funcname.decode(errors="replace"),
            self.message.decode(errors="replace"))


class _ArrowInteger(C.Structure):
    _fields_ = [("bitWidth", C.c_uint16), ("is_signed", C.c_uint8)]


class _ArrowFloat(C.Structure):
    _fields_ = [("precision", C.c_int32)]


class _ArrowFixedBinary(C.Structure):
    _fields_ = [("byteWidth", C.c_uint32)]


class _ArrowDecimal(C.Structure):
    _fields_ = [("precision", C.c_uint16), ("scale", C.c_uint16), ("bitWidth", C.c_uint16)]


class _ArrowOptsUnion(C.Union):
    _fields_ = [("integer", _ArrowInteger), ("floating_point", _ArrowFloat),
                ("fixed_size_binary", _ArrowFixedBinary), ("decimal", _ArrowDecimal),
                ("_pad", C.c_uint16 * 4)]


class ArrowTypeOptions(C.Structure):
    _fields_ = [("tag", C.c_int32), ("unitsz", C.c_int16), ("align", C.c_uint16),
                ("u", _ArrowOptsUnion)]


class kern_colmeta(C.Structure):
    _fields_ = [("attbyval", C.c_bool), ("attalign", C.c_int8), ("attlen", C.c_int16),
                ("attnum", C.c_int16), ("attcacheoff", C.c_int16),
                ("atttypid", C.c_uint32), ("atttypmod", C.c_int32),
                ("atttypkind", C.c_int8), ("kds_format", C.c_char),
                ("dtype_sizeof", C.c_int16), ("kds_offset", C.c_uint32),
                ("idx_subattrs", C.c_uint16), ("num_subattrs", C.c_uint16),
                ("attname", C.c_char * NAMEDATALEN), ("attopts", ArrowTypeOptions),
                ("field_index", C.c_int32), ("virtual_offset", C.c_int64),
                ("nullmap_offset", C.c_uint64), ("nullmap_length", C.c_uint64),
                ("values_offset", C.c_uint64), ("values_length", C.c_uint64),
                ("extra_offset", C.c_uint64), ("extra_length", C.c_uint64)]


class kern_data_store_head(C.Structure):
    """kern_data_store without the colmeta[] tail."""
    _fields_ = [("length", C.c_uint64), ("usage", C.c_uint64), ("nitems", C.c_uint32),
                ("ncols", C.c_uint32), ("format", C.c_char), ("has_varlena", C.c_bool),
                ("tdhasoid", C.c_bool), ("tdtypeid", C.c_uint32), ("tdtypmod", C.c_int32),
                ("table_oid", C.c_uint32), ("hash_nslots", C.c_uint32),
                ("block_offset", C.c_uint32), ("block_nloaded", C.c_uint32),
                ("column_nrooms", C.c_uint32), ("arrow_virtual_usage", C.c_uint32),
                ("parquet_row_group", C.c_uint32), ("parquet_parallel_load", C.c_bool),
                ("parquet_cache_enabled", C.c_bool), ("nr_colmeta", C.c_uint32)]


KDS_COLMETA_OFFSET = 72   # offsetof(kern_data_store, colmeta)


def kds_head_length(nr_colmeta):
    return maxalign(KDS_COLMETA_OFFSET + C.sizeof(kern_colmeta) * nr_colmeta)


class kern_tupitem_head(C.Structure):
    _fields_ = [("t_len", C.c_uint32), ("hash", C.c_uint32), ("has_xact_attrs", C.c_uint8),
                ("__padding__", C.c_uint8), ("t_infomask2", C.c_uint16),
                ("t_infomask", C.c_uint16), ("t_hoff", C.c_uint8)]


class kern_varslot_desc(C.Structure):
    _fields_ = [("vs_type_code", C.c_int32), ("vs_typbyval", C.c_bool),
                ("vs_typalign", C.c_int8), ("vs_typlen", C.c_int16),
                ("vs_typmod", C.c_int32), ("vs_offset", C.c_int32),
                ("idx_subfield", C.c_uint16), ("num_subfield", C.c_uint16),
                ("vs_ops", C.c_void_p)]


class kern_aggregate_desc(C.Structure):
    _fields_ = [("action", C.c_uint16), ("arg0_slot_id", C.c_int16),
                ("arg1_slot_id", C.c_int16), ("filter_slot_id", C.c_int16),
                ("typmod", C.c_int32)]


class kern_session_info_head(C.Structure):
    _fields_ = [("query_plan_id", C.c_uint64), ("kcxt_kvars_nrooms", C.c_uint32),
                ("kcxt_kvars_nslots", C.c_uint32), ("kcxt_kvars_defs", C.c_uint32),
                ("kcxt_kvecs_bufsz", C.c_uint32), ("kcxt_kvecs_ndims", C.c_uint32),
                ("kcxt_extra_bufsz", C.c_uint32), ("cuda_stack_size", C.c_uint32),
                ("xpu_task_flags", C.c_uint32), ("optimal_gpus", C.c_int64),
                ("xpucode_load_vars_packed", C.c_uint32),
                ("xpucode_move_vars_packed", C.c_uint32),
                ("xpucode_scan_quals", C.c_uint32),
                ("xpucode_join_quals_packed", C.c_uint32),
                ("xpucode_hash_values_packed", C.c_uint32),
                ("xpucode_gist_evals_packed", C.c_uint32),
                ("xpucode_projection", C.c_uint32),
                ("xpucode_groupby_keyhash", C.c_uint32),
                ("xpucode_groupby_keyload", C.c_uint32),
                ("xpucode_groupby_keycomp", C.c_uint32),
                ("xpucode_groupby_actions", C.c_uint32),
                ("xpucode_gpusort_keydesc", C.c_uint32),
                ("hostEpochTimestamp", C.c_int64), ("session_xact_state", C.c_uint32),
                ("session_timezone", C.c_uint32), ("session_encode", C.c_uint32),
                ("session_currency_frac_digits", C.c_int32),
                ("pinned_inner_buffer_partition_size", C.c_uint64),
                ("projection_kds_dst", C.c_uint32), ("select_into_pathname", C.c_uint32),
                ("select_into_kds_head", C.c_uint32), ("select_into_projdesc", C.c_uint32),
                ("pgsql_port_number", C.c_uint32), ("pgsql_plan_node_id", C.c_uint32),
                ("join_inner_handle", C.c_uint32), ("groupby_kds_final", C.c_uint32),
                ("groupby_prepfn_bufsz", C.c_uint32),
                ("groupby_ngroups_estimation", C.c_float),
                ("gpusort_htup_margin", C.c_uint32), ("gpusort_limit_count", C.c_int32),
                ("fallback_kds_head", C.c_uint32), ("fallback_desc_defs", C.c_uint32),
                ("fallback_desc_nitems", C.c_uint32), ("nparams", C.c_uint32)]


SESSION_POFFSET_OFFSET = 192  # offsetof(kern_session_info, poffset)


class kern_exec_results_stat(C.Structure):
    _fields_ = [("nitems_roj", C.c_uint32), ("nitems_gist", C.c_uint32),
                ("nitems_out", C.c_uint32)]


MAX_EXEC_RESULTS_RELS = 8


class kern_exec_results(C.Structure):
    _fields_ = [("scan_relidx", C.c_int32), ("chunks_offset", C.c_uint32),
                ("chunks_nitems", C.c_uint32), ("ojmap_offset", C.c_uint32),
                ("ojmap_length", C.c_uint32), ("right_outer_join", C.c_bool),
                ("final_plan_task", C.c_bool), ("final_nitems", C.c_uint32),
                ("final_usage", C.c_uint64), ("final_total", C.c_uint64),
                ("final_sorting_msec", C.c_uint32),
                ("final_reconstruction_msec", C.c_uint32),
                ("join_reconstruction_msec", C.c_uint32),
                ("select_into_nblocks", C.c_uint32),
                ("npages_direct_read", C.c_uint32), ("npages_vfs_read", C.c_uint32),
                ("nchunks_parquet_read", C.c_uint32), ("ncaches_parquet_load", C.c_uint32),
                ("nitems_raw", C.c_uint32), ("nitems_in", C.c_uint32),
                ("nitems_out", C.c_uint32), ("num_rels", C.c_uint32),
                ("stats", kern_exec_results_stat * MAX_EXEC_RESULTS_RELS)]


class kern_multirels_chunk(C.Structure):
    _fields_ = [("kds_in", C.c_void_p), ("kbuf_parts", C.c_void_p),
                ("kds_offset", C.c_uint64), ("ojmap_offset", C.c_uint64),
                ("gist_offset", C.c_uint64), ("gist_ctid_resno", C.c_int16),
                ("is_nestloop", C.c_bool), ("left_outer", C.c_bool),
                ("right_outer", C.c_bool), ("pinned_buffer", C.c_bool),
                ("buffer_id", C.c_uint64)]


class kern_multirels_head(C.Structure):
    _fields_ = [("length", C.c_uint64), ("ojmap_sz", C.c_uint64),
                ("kbuf_part_offset", C.c_uint64), ("num_inner_rels", C.c_uint32)]


KMRELS_CHUNKS_OFFSET = 32   # offsetof(kern_multirels, chunks)

# kern_expression field offsets (xpu_common.h:2596-2692)
KEXP_U_OFFSET = 40


def kds_get_hashslot_width(nitems):
    """xpu_common.h:1389-1396"""
    if nitems <= 5000:
        return 20000
    if nitems <= 4000000:
        return 20000 + 2 * nitems
    return 8020000 + nitems
