"""Host-side mirror of the reference's kern_expression code generator and
session-blob builder, used by tests, smoke and bench to produce exactly the
kind of OpenSession payload the reference backend sends to its GPU service.

Byte layout follows:
  * codegen.c:1642-1658  magic/length trailer of every kern_expression
  * codegen.c:1660-1851  ConstExpr / ParamExpr / VarExpr encodings
  * codegen.c:1853-1900  function expressions (args appended MAXALIGN'ed)
  * codegen.c:1506-1597  SaveExpr around temporary expressions
  * codegen.c:3161-3250  LoadVars per depth, packed by depth
  * codegen.c:3437-3633  JoinQuals / HashValue, packed by depth
  * codegen.c:3881-3956  Projection
  * codegen.c:3963-4357  group-by keyhash / keyload / keycomp / AggFuncs
  * executor.c:517-801   kern_session_info blob (params, kvars defs, codes,
                         kds_dst and kds_final headers)
"""
import ctypes as C
import struct

import numpy as np

from . import abi
from . import kds as kdsmod
from .opcodes import FUNCS, SPECIAL, TYPES

TYPE_CODE = TYPES


def _pad_to(buf, align):
    n = (-len(buf)) % align
    buf += b"\0" * n


def _append_aligned(buf, data):
    """__appendBinaryStringInfo (misc.c:481-492): MAXALIGN then append."""
    _pad_to(buf, 8)
    pos = len(buf)
    buf += data
    return pos


def _finish(buf, head_pos):
    """__appendKernExpMagicAndLength (codegen.c:1642-1658)."""
    _pad_to(buf, 4)
    exptype, = struct.unpack_from("<i", buf, head_pos + 4)
    opcode, = struct.unpack_from("<i", buf, head_pos + 20)
    magic = (abi.KERN_EXPRESSION_MAGIC ^ ((exptype << 6) & 0xffffffff)
             ^ ((opcode << 14) & 0xffffffff)) & 0xffffffff
    buf += struct.pack("<I", magic)
    struct.pack_into("<I", buf, head_pos, len(buf) - head_pos)


def _kexp_head(exptype, opcode, nr_args=0, args_offset=0, flags=0):
    # len, exptype, expr_ops, expflags, opcode, fn_dptr, nr_args, args_offset
    return struct.pack("<IiQIiQII", 0, exptype, 0, flags, opcode, 0, nr_args, args_offset)


# ----------------------------------------------------------------------
# expression tree
# ----------------------------------------------------------------------
class Expr:
    typname = None
    pushed_down = False


class Var(Expr):
    def __init__(self, slot, typname):
        self.slot = slot
        self.typname = typname


class Const(Expr):
    def __init__(self, typname, value):
        self.typname = typname
        self.value = value


class Param(Expr):
    def __init__(self, typname, param_id):
        self.typname = typname
        self.param_id = param_id


class Func(Expr):
    def __init__(self, fname, *args):
        if fname not in FUNCS:
            raise KeyError("device function %s is not supported" % fname)
        self.fname = fname
        self.args = list(args)
        self.typname = FUNCS[fname][4]


class BoolOp(Expr):
    def __init__(self, kind, *args):
        assert kind in ("and", "or", "not")
        self.kind = kind
        self.args = list(args)
        self.typname = "bool"


class NullTest(Expr):
    def __init__(self, arg, isnull=True):
        self.arg = arg
        self.isnull = isnull
        self.typname = "bool"


class BoolTest(Expr):
    KINDS = ("IsTrue", "IsNotTrue", "IsFalse", "IsNotFalse", "IsUnknown", "IsNotUnknown")

    def __init__(self, kind, arg):
        assert kind in self.KINDS
        self.kind = kind
        self.arg = arg
        self.typname = "bool"


class Coalesce(Expr):
    def __init__(self, typname, *args):
        self.typname = typname
        self.args = list(args)


class DistinctFrom(Expr):
    """a IS DISTINCT FROM b over a comparator function (codegen.c:1929-1949):
    DistinctFrom with one argument, the '=' function of a and b."""

    def __init__(self, cmp):
        self.cmp = cmp
        self.typname = "bool"


class MinMax(Expr):
    """LEAST / GREATEST (codegen.c:2373-2420)."""

    def __init__(self, kind, typname, *args):
        assert kind in ("least", "greatest")
        self.kind = kind
        self.typname = typname
        self.args = list(args)


class CaseWhen(Expr):
    """CASE [key] WHEN cond THEN result ... [ELSE default] END
    (codegen.c:2701-2810).  whens: [(cond, result)].  key: (slot, expr) of
    CASE <key>: a SaveExpr of the key into the slot, which the conditions
    read as Var(slot)."""

    def __init__(self, typname, whens, default=None, key=None):
        self.typname = typname
        self.whens = list(whens)
        self.default = default
        self.key = key


ARRAY_OIDS = {"bool": 1000, "int2": 1005, "int4": 1007, "int8": 1016, "float4": 1021,
              "float8": 1022, "text": 1009, "bpchar": 1014}


def array_image(elem_typname, values):
    """A one-dimensional PostgreSQL array (ArrayType, utils/array.h) of
    elem_typname holding values (None = NULL element)."""
    oid, typlen, byval, align, _ = kdsmod.PG_TYPES[elem_typname]
    n = len(values)
    hasnull = any(v is None for v in values)
    head = 16 + 8          # vl_len_, ndim, dataoffset, elemtype; dims[1], lbound[1]
    bitmap = b""
    if hasnull:
        bits = bytearray((n + 7) // 8)
        for i, v in enumerate(values):
            if v is not None:
                bits[i >> 3] |= 1 << (i & 7)
        bitmap = bytes(bits)
    dataoffset = abi.maxalign(head + len(bitmap)) if hasnull else 0
    data = bytearray()
    for v in values:
        if v is None:
            continue
        d = encode_const_value(elem_typname, v)
        if typlen > 0 or True:
            data += b"\0" * ((-len(data)) % align)
        data += d
    start = abi.maxalign(head + len(bitmap))
    total = start + len(data)
    img = struct.pack("<IiiI", total << 2, 1 if n else 0, dataoffset, oid)
    img += struct.pack("<ii", n, 1) + bitmap
    img += b"\0" * (start - len(img))
    return img + bytes(data)


class ScalarArrayOp(Expr):
    """scalar <op> ANY / ALL (ARRAY[...]) over a constant array
    (codegen.c:2909-3060): arg 1 the array constant, arg 2 the comparator
    function of the scalar and Var(elem_slot), the array element."""

    def __init__(self, use_any, fname, scalar, elem_slot, elem_typname, values):
        self.use_any = use_any
        self.fname = fname
        self.scalar = scalar
        self.elem_slot = elem_slot
        self.elem_typname = elem_typname
        self.values = values        # list, or None for a NULL array
        self.typname = "bool"


class HashedArrayOp(Expr):
    """scalar = ANY (constant array), or CASE scalar WHEN const THEN const ...
    ELSE const END, as the reference's hash table of constant keys
    (codegen.c:2820-2906 and 2458-2660).  items: [(key or None, value or
    None)] in order; default: the value of no match (None = NULL)."""

    def __init__(self, typname, elem_typname, scalar, items, default=None, in_list=False):
        self.typname = typname
        self.elem_typname = elem_typname
        self.scalar = scalar
        self.items = list(items)
        self.default = default
        # scalar = ANY (array): the default is the zero byte of the header's
        # padding, i.e. false (codegen.c:2855)
        self.in_list = in_list


class HashValue(Expr):
    """HashValue over its arguments (pgfn_HashValue, xpu_common.cu:1363-1398)."""

    def __init__(self, *args):
        self.args = list(args)
        self.typname = "int4"


def And(*a):
    return BoolOp("and", *a)


def Or(*a):
    return BoolOp("or", *a)


def Not(a):
    return BoolOp("not", a)


def encode_const_value(typname, value):
    if typname in ("text", "bytea", "bpchar"):
        v = bytes(value)
        return struct.pack("<I", (len(v) + 4) << 2) + v
    return kdsmod.encode_datum(typname, value)


def emit_expr(buf, e, flags=0):
    """Append expression `e` (codegen_expression_walker) and return its position."""
    if e.pushed_down:
        flags |= abi.KEXP_FLAG__IS_PUSHED_DOWN
    tc = TYPE_CODE[e.typname]
    if isinstance(e, Var):
        pos = _append_aligned(buf, _kexp_head(tc, SPECIAL["VarExpr"], flags=flags) +
                              struct.pack("<iH", -1, e.slot))
        _finish(buf, pos)
        return pos
    if isinstance(e, Const):
        oid = kdsmod.PG_TYPES[e.typname][0]
        isnull = e.value is None
        head = _kexp_head(tc, SPECIAL["ConstExpr"], flags=flags) + struct.pack("<I?3x", oid, isnull)
        pos = _append_aligned(buf, head)
        if not isnull:
            buf += encode_const_value(e.typname, e.value)
        _finish(buf, pos)
        return pos
    if isinstance(e, Param):
        pos = _append_aligned(buf, _kexp_head(tc, SPECIAL["ParamExpr"], flags=flags) +
                              struct.pack("<I", e.param_id))
        _finish(buf, pos)
        return pos
    if isinstance(e, Func):
        opcode = FUNCS[e.fname][0]
        return _emit_with_args(buf, tc, opcode, e.args, flags)
    if isinstance(e, BoolOp):
        opcode = SPECIAL[{"and": "BoolExpr_And", "or": "BoolExpr_Or",
                          "not": "BoolExpr_Not"}[e.kind]]
        return _emit_with_args(buf, tc, opcode, e.args, flags)
    if isinstance(e, NullTest):
        opcode = SPECIAL["NullTestExpr_IsNull" if e.isnull else "NullTestExpr_IsNotNull"]
        return _emit_with_args(buf, tc, opcode, [e.arg], flags)
    if isinstance(e, BoolTest):
        return _emit_with_args(buf, tc, SPECIAL["BoolTestExpr_" + e.kind], [e.arg], flags)
    if isinstance(e, Coalesce):
        return _emit_with_args(buf, tc, SPECIAL["CoalesceExpr"], e.args, flags)
    if isinstance(e, HashValue):
        return _emit_with_args(buf, tc, SPECIAL["HashValue"], e.args, flags)
    if isinstance(e, DistinctFrom):
        return _emit_with_args(buf, tc, SPECIAL["DistinctFrom"], [e.cmp], flags)
    if isinstance(e, MinMax):
        return _emit_with_args(buf, tc, SPECIAL["LeastExpr" if e.kind == "least"
                                                else "GreatestExpr"], e.args, flags)
    if isinstance(e, CaseWhen):
        return _emit_casewhen(buf, tc, e, flags)
    if isinstance(e, ScalarArrayOp):
        return _emit_saop(buf, tc, e, flags)
    if isinstance(e, HashedArrayOp):
        return _emit_haop(buf, tc, e, flags)
    if isinstance(e, SaveExpr):
        pos = _append_aligned(buf, _kexp_head(tc, SPECIAL["SaveExpr"], 1, 48, flags) +
                              struct.pack("<H6x", e.slot))
        emit_expr(buf, e.arg, flags & ~abi.KEXP_FLAG__IS_PUSHED_DOWN)
        _finish(buf, pos)
        return pos
    raise TypeError(e)


def _emit_with_args(buf, tc, opcode, args, flags):
    pos = _append_aligned(buf, _kexp_head(tc, opcode, len(args), 40, flags))
    for a in args:
        emit_expr(buf, a, flags & ~abi.KEXP_FLAG__IS_PUSHED_DOWN)
    _finish(buf, pos)
    return pos


def _emit_casewhen(buf, tc, e, flags):
    sub = flags & ~abi.KEXP_FLAG__IS_PUSHED_DOWN
    pos = _append_aligned(buf, _kexp_head(tc, SPECIAL["CaseWhenExpr"], 2 * len(e.whens), 48, flags)
                          + struct.pack("<II", 0, 0))
    case_comp = case_else = 0
    args_offset = 48
    if e.key is not None:
        slot, kexpr = e.key
        _pad_to(buf, 8)
        case_comp = len(buf) - pos
        emit_expr(buf, SaveExpr(slot, kexpr), sub)
        _pad_to(buf, 8)
        args_offset = len(buf) - pos
    for cond, res in e.whens:
        emit_expr(buf, cond, sub)
        emit_expr(buf, res, sub)
    if e.default is not None:
        _pad_to(buf, 8)
        case_else = len(buf) - pos
        emit_expr(buf, e.default, sub)
    struct.pack_into("<I", buf, pos + 36, args_offset)
    struct.pack_into("<II", buf, pos + 40, case_comp, case_else)
    _finish(buf, pos)
    return pos


def _emit_array_const(buf, elem_typname, values, flags):
    """ConstExpr of an array type (exptype TypeOpCode__array)."""
    head = _kexp_head(TYPE_CODE["array"], SPECIAL["ConstExpr"], flags=flags) + \
        struct.pack("<I?3x", ARRAY_OIDS[elem_typname], values is None)
    pos = _append_aligned(buf, head)
    if values is not None:
        buf += array_image(elem_typname, values)
    _finish(buf, pos)
    return pos


def _emit_saop(buf, tc, e, flags):
    sub = flags & ~abi.KEXP_FLAG__IS_PUSHED_DOWN
    opcode = SPECIAL["ScalarArrayOpAny" if e.use_any else "ScalarArrayOpAll"]
    pos = _append_aligned(buf, _kexp_head(tc, opcode, 2, 48, flags) + struct.pack("<H6x", e.elem_slot))
    _emit_array_const(buf, e.elem_typname, e.values, sub)
    emit_expr(buf, Func(e.fname, e.scalar, Var(e.elem_slot, e.elem_typname)), sub)
    _finish(buf, pos)
    return pos


def _datum_image(typname, value):
    return encode_const_value(typname, value)


def _emit_haop(buf, tc, e, flags):
    """HashedArrayOp: header with the hash slots, then the items (hash, next
    with HAS_KEY / HAS_VALUE bits, key image, value image aligned to the
    result type), the default value, then the scalar argument."""
    from . import hashing
    sub = flags & ~abi.KEXP_FLAG__IS_PUSHED_DOWN
    nitems = len(e.items)
    nslots = nitems * 5 // 4 + 1
    head_sz = abi.maxalign(64 + 4 * nslots)
    tmp = bytearray(head_sz)
    hslots = [0] * nslots
    r_align = kdsmod.PG_TYPES[e.typname][3]
    for key, value in e.items:
        h = 0 if key is None else hashing.hash_any(_hash_image(e.elem_typname, key))
        idx = h % nslots
        nxt = hslots[idx]
        if key is not None:
            nxt |= abi.KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY
        if value is not None:
            nxt |= abi.KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE
        _pad_to(tmp, 8)
        hslots[idx] = len(tmp)
        tmp += struct.pack("<II", h, nxt)
        klen = 0
        if key is not None:
            kimg = _datum_image(e.elem_typname, key)
            tmp += kimg
            klen = len(kimg)
        if value is not None:
            gap = (-klen) % r_align
            tmp += b"\0" * gap
            tmp += _datum_image(e.typname, value)
    default_offset = 0
    if e.in_list:
        default_offset = 60     # offsetof(kern_expression, u.haop.__padding)
    elif e.default is not None:
        _pad_to(tmp, 8)
        default_offset = len(tmp)
        tmp += _datum_image(e.typname, e.default)
    _pad_to(tmp, 8)
    args_offset = len(tmp)
    tmp[0:40] = _kexp_head(tc, SPECIAL["HashedArrayOp"], 1, args_offset, flags)
    struct.pack_into("<QiII", tmp, 40, 0, TYPE_CODE[e.elem_typname], nslots, default_offset)
    for i, h in enumerate(hslots):
        struct.pack_into("<I", tmp, 64 + 4 * i, h)
    pos = _append_aligned(buf, bytes(tmp))
    emit_expr(buf, e.scalar, sub)
    _finish(buf, pos)
    return pos


def _hash_image(typname, value):
    """The bytes PostgreSQL's type hash function hashes (hashint4 etc. hash
    the value's own width; text its payload)."""
    if typname in ("text", "bytea", "bpchar"):
        return bytes(value)
    return kdsmod.encode_datum(typname, value)


class SaveExpr(Expr):
    def __init__(self, slot, arg):
        self.slot = slot
        self.arg = arg
        self.typname = arg.typname


def _packed(subs, npacked):
    """FuncOpCode__Packed wrapper (codegen.c:3203-3248): subs = {index: bytes}."""
    sz = abi.maxalign(40 + 4 + 4 * (npacked + 1))
    buf = bytearray(sz)
    offsets = [0] * (npacked + 1)
    nr_args = 0
    for idx in sorted(subs):
        offsets[idx] = _append_aligned(buf, subs[idx])
        nr_args += 1
    head = _kexp_head(TYPE_CODE["int4"], SPECIAL["Packed"], nr_args, sz)
    buf[0:40] = head
    struct.pack_into("<I", buf, 40, npacked)
    for i, o in enumerate(offsets):
        struct.pack_into("<I", buf, 44 + 4 * i, o)
    _finish(buf, 0)
    return bytes(buf)


def _loadvars(depth, items):
    """LoadVars for one depth (codegen.c:3161-3199); items: [(resno, slot)]."""
    items = sorted(items)
    n = len(items)
    args_offset = abi.maxalign(48 + 4 * n)
    buf = bytearray(_kexp_head(TYPE_CODE["int4"], SPECIAL["LoadVars"], 0, args_offset))
    buf += struct.pack("<ii", depth, n)
    for resno, slot in items:
        buf += struct.pack("<hh", resno, slot)
    buf += b"\0" * (args_offset - len(buf))
    _finish(buf, 0)
    return bytes(buf)


# ----------------------------------------------------------------------
# plan builder
# ----------------------------------------------------------------------
class Join:
    def __init__(self, depth, outer_keys, join_quals, other_quals, left_outer):
        self.depth = depth
        self.outer_keys = outer_keys
        self.join_quals = join_quals
        self.other_quals = other_quals
        self.left_outer = left_outer


class Plan:
    """Accumulates a GpuScan / GpuJoin / GpuPreAgg plan in the shape the
    reference planner hands to codegen, then serializes the session blob."""

    def __init__(self, scan_coldefs):
        self.scan_coldefs = list(scan_coldefs)      # [(name, typname)]
        self.slots = []                             # [(typname, depth, resno)]
        self.loads = {}                             # depth -> [(resno, slot)]
        self.scan_quals = None
        self.joins = []
        self.inner_coldefs = {}
        self.projection = None
        self.proj_coldefs = None
        self.groupby = None
        self.params = []
        self.query_plan_id = 0x5eed

    def _new_slot(self, typname, depth, resno):
        self.slots.append((typname, depth, resno))
        return len(self.slots) - 1

    def scan_var(self, name):
        for j, (n, t) in enumerate(self.scan_coldefs):
            if n == name:
                for s, (tt, d, r) in enumerate(self.slots):
                    if d == 0 and r == j + 1:
                        return Var(s, t)
                s = self._new_slot(t, 0, j + 1)
                self.loads.setdefault(0, []).append((j + 1, s))
                return Var(s, t)
        raise KeyError(name)

    def inner_var(self, depth, name):
        for j, (n, t) in enumerate(self.inner_coldefs[depth]):
            if n == name:
                for s, (tt, d, r) in enumerate(self.slots):
                    if d == depth and r == j + 1:
                        return Var(s, t)
                s = self._new_slot(t, depth, j + 1)
                self.loads.setdefault(depth, []).append((j + 1, s))
                return Var(s, t)
        raise KeyError(name)

    def temp_slot(self, typname, depth):
        return self._new_slot(typname, depth, 0)

    def param(self, typname, value):
        self.params.append((typname, value))
        return Param(typname, len(self.params) - 1)

    def add_inner(self, coldefs):
        depth = len(self.inner_coldefs) + 1
        self.inner_coldefs[depth] = list(coldefs)
        return depth

    def add_join(self, depth, outer_keys, join_quals, other_quals=(), left_outer=False):
        for q in other_quals:
            q.pushed_down = True
        self.joins.append(Join(depth, list(outer_keys), list(join_quals),
                               list(other_quals), left_outer))

    def set_projection(self, targets, coldefs):
        self.projection = list(targets)
        self.proj_coldefs = list(coldefs)

    def set_groupby(self, tlist):
        """tlist: [(action, expr or None, (name, typname))] in kds_final
        column order; VREF entries are grouping keys.  The reference planner
        puts partial aggregates before the keys (cuda_gpupreagg.cu:1522-1528)."""
        self.groupby = list(tlist)

    # ------------------------------------------------------------------
    def _slot_for(self, e, depth):
        """Slot of a simple Var, or of a temporary holding expression `e`;
        one temporary per distinct expression object, as
        __try_inject_temporary_expression (codegen.c:1506-1597) reuses them."""
        if isinstance(e, Var):
            return e.slot
        cache = self.__dict__.setdefault("_temp_cache", {})
        if id(e) not in cache:
            cache[id(e)] = (e, self.temp_slot(e.typname, depth))
        return cache[id(e)][1]

    def _emit_saves(self, buf, exprs, depth):
        """Wrap each expr in SaveExpr; return the slot ids."""
        slots = []
        for e in exprs:
            slot = self._slot_for(e, depth)
            emit_expr(buf, SaveExpr(slot, e))
            slots.append(slot)
        return slots

    def build(self, kds_dst_length=0, final_length=1 << 30, final_nrows=10000, xids=None,
              pinned_hash=None, pinned_nslots=137, pinned_length=512 << 20, sort_keys=None,
              sort_limit=0, window_rank=None, encoding="UTF8", fallback=False,
              fallback_length=64 << 20):
        """The session blob.  xids: parallelCurrentXids of the transaction
        (SerializedTransactionState at session_xact_state, xpu_common.h:3156-3165),
        which decide the visibility of KDS_FORMAT_ROW sources.  pinned_hash:
        join-key expressions of a projection whose rows stay on the device
        as a pinned inner buffer (DEVTASK__PINNED_HASH_RESULTS): the
        projection carries their HashValue (codegen.c:3933-3941) and
        groupby_kds_final the buffer's KDS_FORMAT_HASH header
        (executor.c:655-672).  sort_keys: GPU-Sort of the group-by results,
        [(KSORT_KEY_KIND__*, attribute number in kds_final, typname of a VREF
        key or None, asc, nulls_first)], with sort_limit rows kept
        (gpusort_limit_count); window_rank: (func, limit, partby, orderby)
        of a window rank function.  fallback: the session carries the CPU
        fallback (pg_strom.cpu_fallback on, executor.c:705-724): a
        KDS_FORMAT_FALLBACK header over the scan slot (the table's columns,
        then each depth's inner columns) of fallback_length bytes, and one
        kern_fallback_desc per loaded variable mapping it into that slot
        (__execInitTaskStateCpuFallback, executor.c:1280-1380)."""
        nrels = len(self.inner_coldefs)
        codes = {}
        # scan quals
        if self.scan_quals is not None:
            b = bytearray()
            emit_expr(b, self.scan_quals)
            codes["scan_quals"] = bytes(b)
        # joins
        if self.joins:
            jq, hv = {}, {}
            for j in self.joins:
                b = bytearray()
                quals = j.join_quals + j.other_quals
                pos = _append_aligned(b, _kexp_head(TYPE_CODE["bool"], SPECIAL["JoinQuals"],
                                                    len(quals), 40))
                for q in quals:
                    emit_expr(b, q)
                _finish(b, pos)
                jq[j.depth] = bytes(b)
                b = bytearray(_kexp_head(TYPE_CODE["int4"], SPECIAL["HashValue"],
                                         len(j.outer_keys), 40))
                for k in j.outer_keys:
                    emit_expr(b, k)
                _finish(b, 0)
                hv[j.depth] = bytes(b)
            codes["join_quals"] = _packed(jq, nrels + 1)
            codes["hash_values"] = _packed(hv, nrels + 1)
        # projection
        if self.projection is not None:
            n = len(self.projection)
            sz = abi.maxalign(48 + 2 * n)
            b = bytearray(sz)
            slots = self._emit_saves(b, self.projection, nrels + 1)
            b[0:40] = _kexp_head(TYPE_CODE["int4"], SPECIAL["Projection"],
                                 sum(1 for _ in slots), sz)
            struct.pack_into("<Ii", b, 40, 0, n)
            for i, s in enumerate(slots):
                struct.pack_into("<H", b, 48 + 2 * i, s)
            if pinned_hash:
                hv = bytearray(_kexp_head(TYPE_CODE["int4"], SPECIAL["HashValue"],
                                          len(pinned_hash), 40))
                for k in pinned_hash:
                    emit_expr(hv, k)
                _finish(hv, 0)
                hpos = _append_aligned(b, bytes(hv))
                struct.pack_into("<I", b, 40, hpos)
            _finish(b, 0)
            codes["projection"] = bytes(b)
        # group-by
        final_coldefs = None
        if self.groupby is not None:
            final_coldefs = [tuple(cd[:2]) for _, _, cd in self.groupby]
            keys = [(e, tuple(cd[:2])) for a, e, cd in self.groupby if a == abi.KAGG_ACTION__VREF]
            if keys:
                # keyhash: HashValue over SaveExpr(key)  (codegen.c:3963-4010)
                b = bytearray(40)
                key_slots = self._emit_saves(b, [e for e, _ in keys], nrels + 1)
                b[0:40] = _kexp_head(TYPE_CODE["int4"], SPECIAL["HashValue"], len(keys), 40)
                _finish(b, 0)
                codes["groupby_keyhash"] = bytes(b)
                # keyload from kds_final (codegen.c:4015-4070)
                final_slots = []
                items = []
                for (e, cd), resno in zip(keys, [i + 1 for i, (a, _, _) in enumerate(self.groupby)
                                                 if a == abi.KAGG_ACTION__VREF]):
                    s = self._new_slot(e.typname, abi.SPECIAL_DEPTH__PREAGG_FINAL, resno)
                    final_slots.append(s)
                    items.append((resno, s))
                codes["groupby_keyload"] = _loadvars(abi.SPECIAL_DEPTH__PREAGG_FINAL, items)
                # keycomp: (And) of <type>eq(input, final)  (codegen.c:4075-4160)
                cmps = []
                for (e, _), si, sf in zip(keys, key_slots, final_slots):
                    eqf = {"int1": "int1eq", "int2": "int2eq", "int4": "int4eq",
                           "int8": "int8eq", "float4": "float4eq", "float8": "float8eq",
                           "text": "text_eq", "date": None, "bool": "booleq",
                           "numeric": "numeric_eq", "bpchar": "bpchareq"}[e.typname]
                    cmps.append(Func(eqf, Var(si, e.typname), Var(sf, e.typname)))
                b = bytearray()
                emit_expr(b, cmps[0] if len(cmps) == 1 else And(*cmps))
                codes["groupby_keycomp"] = bytes(b)
            # actions (codegen.c:4200-4300)
            n = len(self.groupby)
            head_sz = abi.maxalign(44 + 12 * n)
            b = bytearray(head_sz)
            descs = []
            nr_args = 0
            emitted = set()
            for action, e, cd in self.groupby:
                a0 = -1
                if e is not None:
                    a0 = self._slot_for(e, nrels + 1)
                    if a0 not in emitted:     # codegen.c:4170-4196 reverts duplicates
                        emit_expr(b, SaveExpr(a0, e))
                        emitted.add(a0)
                        nr_args += 1
                # the argument's typmod: (name, typname, typmod) in the tlist
                descs.append((action, a0, cd[2] if len(cd) > 2 else -1))
            b[0:40] = _kexp_head(TYPE_CODE["int4"], SPECIAL["AggFuncs"], nr_args, head_sz)
            struct.pack_into("<i", b, 40, n)
            for i, (action, a0, typmod) in enumerate(descs):
                struct.pack_into("<Hhhhi", b, 44 + 12 * i, action, a0, -1, -1, typmod)
            _finish(b, 0)
            codes["groupby_actions"] = bytes(b)
        # load vars (packed by depth; codegen.c:3203-3248)
        if self.loads:
            subs = {d: _loadvars(d, items) for d, items in self.loads.items() if d >= 0}
            codes["load_vars"] = _packed(subs, nrels + 1)

        # ---- session blob (executor.c:517-801)
        nparams = len(self.params)
        session_sz = abi.SESSION_POFFSET_OFFSET + 4 * max(nparams, 1)
        buf = bytearray(abi.maxalign(session_sz))
        sess = abi.kern_session_info_head()
        poffsets = []
        for typname, value in self.params:
            if value is None:
                poffsets.append(0)
            else:
                poffsets.append(_append_aligned(buf, encode_const_value(typname, value)))
        # kvars definitions (executor.c:438-510)
        defs = bytearray()
        for typname, depth, resno in self.slots:
            oid, typlen, byval, align, tcode = kdsmod.PG_TYPES[typname]
            d = abi.kern_varslot_desc(tcode, byval, align, typlen, -1, -1, 0, 0, None)
            defs += bytes(d)
        if self.slots:
            sess.kcxt_kvars_defs = _append_aligned(buf, bytes(defs))
            sess.kcxt_kvars_nslots = len(self.slots)
            sess.kcxt_kvars_nrooms = len(self.slots)
        field = {"load_vars": "xpucode_load_vars_packed", "scan_quals": "xpucode_scan_quals",
                 "join_quals": "xpucode_join_quals_packed",
                 "hash_values": "xpucode_hash_values_packed",
                 "projection": "xpucode_projection",
                 "groupby_keyhash": "xpucode_groupby_keyhash",
                 "groupby_keyload": "xpucode_groupby_keyload",
                 "groupby_keycomp": "xpucode_groupby_keycomp",
                 "groupby_actions": "xpucode_groupby_actions"}
        for key in ("load_vars", "scan_quals", "join_quals", "hash_values", "projection",
                    "groupby_keyhash", "groupby_keyload", "groupby_keycomp", "groupby_actions"):
            if key in codes:
                setattr(sess, field[key], _append_aligned(buf, codes[key]))
        flags = abi.DEVKIND__NVIDIA_GPU
        if self.groupby is not None:
            flags |= abi.DEVTASK__PREAGG
        elif self.joins:
            flags |= abi.DEVTASK__JOIN
        else:
            flags |= abi.DEVTASK__SCAN
        if self.proj_coldefs is not None:
            head = kdsmod.kds_header(abi.KDS_FORMAT_ROW, self.proj_coldefs, kds_dst_length)
            sess.projection_kds_dst = _append_aligned(buf, bytes(head))
            if pinned_hash:
                head = kdsmod.kds_header(abi.KDS_FORMAT_HASH, self.proj_coldefs, pinned_length,
                                         pinned_nslots)
                sess.groupby_kds_final = _append_aligned(buf, bytes(head))
                flags |= abi.DEVTASK__PINNED_HASH_RESULTS
        if final_coldefs is not None:
            head = kdsmod.kds_header(abi.KDS_FORMAT_ROW, final_coldefs, 0)
            sess.projection_kds_dst = _append_aligned(buf, bytes(head))
            nslots = abi.kds_get_hashslot_width(final_nrows)
            head = kdsmod.kds_header(abi.KDS_FORMAT_HASH, final_coldefs, final_length, nslots)
            sess.groupby_kds_final = _append_aligned(buf, bytes(head))
            sess.groupby_ngroups_estimation = float(final_nrows)
            flags |= abi.DEVTASK__PINNED_HASH_RESULTS
        if sort_keys:
            sess.xpucode_gpusort_keydesc = _append_aligned(buf, _sortkeys_kexp(sort_keys,
                                                                             window_rank))
            sess.gpusort_limit_count = int(sort_limit)
            finalized = A_KSORT_PAVG_KINDS
            sess.gpusort_htup_margin = 9 * sum(1 for k in sort_keys if k[0] in finalized)
        if fallback:
            scan_slot = list(self.scan_coldefs)
            base = {0: 0}
            for d in sorted(self.inner_coldefs):
                base[d] = len(scan_slot)
                scan_slot += list(self.inner_coldefs[d])
            head = kdsmod.kds_header(abi.KDS_FORMAT_FALLBACK, scan_slot, fallback_length)
            sess.fallback_kds_head = _append_aligned(buf, bytes(head))
            descs = bytearray()
            nfb = 0
            for d in sorted(self.loads):
                if d < 0 or d not in base:
                    continue
                for resno, slot in self.loads[d]:
                    # kern_fallback_desc: src depth / resno, dst resno, max
                    # depth, slot id, kvec offset (none: no kvec buffers)
                    descs += struct.pack("<hhhhHxxi", d, resno, base[d] + resno, nrels + 1,
                                         slot, -1)
                    nfb += 1
            if nfb:
                sess.fallback_desc_defs = _append_aligned(buf, bytes(descs))
            sess.fallback_desc_nitems = nfb
        if encoding:
            # xpu_encode_info (__build_session_encode, executor.c:494-506):
            # the database encoding's name and longest character
            maxlen = {"UTF8": 4, "SQL_ASCII": 1}.get(encoding, 1)
            sess.session_encode = _append_aligned(
                buf, struct.pack("<16si4xQ", encoding.encode(), maxlen, 0))
        if xids is not None:
            xs = struct.pack("<i?3xQQIi", 1, False, 0, 0, 0, len(xids)) + \
                struct.pack("<%dI" % len(xids), *xids)
            sess.session_xact_state = _append_aligned(buf, xs)
        sess.query_plan_id = self.query_plan_id
        sess.xpu_task_flags = flags
        sess.kcxt_kvecs_ndims = nrels + 2
        sess.nparams = nparams
        _pad_to(buf, 8)
        buf[0:C.sizeof(sess)] = bytes(sess)
        for i, o in enumerate(poffsets):
            struct.pack_into("<I", buf, abi.SESSION_POFFSET_OFFSET + 4 * i, o)
        return bytes(buf)


A_KSORT_PAVG_KINDS = (abi.KSORT_KEY_KIND__PAVG_INT64, abi.KSORT_KEY_KIND__PAVG_FP64,
                      abi.KSORT_KEY_KIND__PAVG_NUMERIC)
A_KSORT_RAW_KINDS = (abi.KSORT_KEY_KIND__VREF, abi.KSORT_KEY_KIND__PMINMAX_INT64,
                     abi.KSORT_KEY_KIND__PMINMAX_FP64, abi.KSORT_KEY_KIND__PSUM_INT64,
                     abi.KSORT_KEY_KIND__PSUM_FP64, abi.KSORT_KEY_KIND__PSUM_NUMERIC)


def _sortkeys_kexp(keys, window_rank=None):
    """FuncOpCode__SortKeys (codegen.c:4360-4525): u.sort header, then one
    kern_sortkey_desc per key; finalized keys (averages) get 9 bytes of the
    tuple margin each (bool + float8)."""
    b = bytearray(_kexp_head(TYPE_CODE["int4"], SPECIAL["SortKeys"]))
    needs_fin = any(k[0] not in A_KSORT_RAW_KINDS for k in keys)
    func, wlimit, partby, orderby = window_rank or (0, 0, 0, 0)
    b += struct.pack("<i?BxxIHH", len(keys), needs_fin, func, wlimit, partby, orderby)
    margin = 0
    for kind, anum, typname, asc, nulls_first in keys:
        if kind == abi.KSORT_KEY_KIND__VREF:
            tcode = TYPE_CODE[typname]
            boff = 0
        elif kind in A_KSORT_PAVG_KINDS:
            tcode = TYPE_CODE["float8"]
            boff = margin
            margin += 9
        elif kind == abi.KSORT_KEY_KIND__PSUM_NUMERIC:
            tcode = TYPE_CODE["numeric"]
            boff = 0
        elif kind in (abi.KSORT_KEY_KIND__PMINMAX_FP64, abi.KSORT_KEY_KIND__PSUM_FP64):
            tcode = TYPE_CODE["float8"]
            boff = 0
        else:
            tcode = TYPE_CODE["int8"]
            boff = 0
        b += struct.pack("<HbbHHi4xQ", kind, 1 if nulls_first else 0, 1 if asc else 0, anum,
                         boff, tcode, 0)
    _finish(b, 0)
    return bytes(b)


def xpu_command(tag, payload):
    """Wrap a payload into an XpuCommand (xpu_common.h:2904-2919)."""
    head = struct.pack("<IIQQQQ", abi.XpuCommandMagicNumber, tag, 40 + len(payload), 0, 0, 0)
    return head + payload
