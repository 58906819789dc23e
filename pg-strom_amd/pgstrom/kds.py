"""Builders and parsers for kern_data_store buffers (host side).

Formats follow the reference (src/xpu_common.h):
  * KDS_FORMAT_ARROW  (:1536-1628): columnar, Arrow validity bitmaps and
    values/offsets buffers addressed by kern_colmeta offsets from the KDS head;
  * KDS_FORMAT_ROW    (:771-798): row-index of u64 offsets counted from the
    KDS tail, kern_tupitem payloads packed towards the tail;
  * KDS_FORMAT_HASH   (:1383-1451): hash slots of u64 kern_hashitem offsets
    (from the tail) in front of the row-index, chains through hitem->next,
    built the way gpu_join.c:2128-2164 lays out an inner relation.
"""
import ctypes as C
import struct

import numpy as np

from . import abi

# name -> (oid, typlen, typbyval, typalign, TypeOpCode)
PG_TYPES = {
    "bool": (16, 1, True, 1, 1),
    "int1": (0, 1, True, 1, 2),
    "int2": (21, 2, True, 2, 3),
    "int4": (23, 4, True, 4, 4),
    "int8": (20, 8, True, 8, 5),
    "float4": (700, 4, True, 4, 7),
    "float8": (701, 8, True, 8, 8),
    "numeric": (1700, -1, False, 4, 9),
    "bytea": (17, -1, False, 4, 10),
    "text": (25, -1, False, 4, 11),
    "bpchar": (1042, -1, False, 4, 12),
    "date": (1082, 4, True, 4, 13),
}

# Arrow type tags of the other varlena layouts (arrow_defs.h)
ARROW_FIXED_SIZE_BINARY = 15
ARROW_LARGE_BINARY = 19
ARROW_LARGE_UTF8 = 20

NP_DTYPES = {"bool": np.bool_, "int1": np.int8, "int2": np.int16, "int4": np.int32,
             "int8": np.int64, "float4": np.float32, "float8": np.float64,
             "date": np.int32}


class ArrowVarlena:
    """Text column already in Arrow form: int64 offsets[n+1] and data bytes."""

    def __init__(self, offsets, data):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.data = data

    def __len__(self):
        return len(self.offsets) - 1

    def __getitem__(self, i):
        if isinstance(i, slice):
            start, stop, step = i.indices(len(self))
            if step != 1:
                raise ValueError("ArrowVarlena slices take step 1")
            stop = max(stop, start)
            offs = self.offsets[start:stop + 1]
            return ArrowVarlena(offs - offs[0], self.data[int(offs[0]):int(offs[-1])])
        return self.data[self.offsets[i]:self.offsets[i + 1]]

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]


class DecimalValues:
    """A numeric column as an Arrow Decimal128 holds it: 128-bit integers at
    one decimal scale (value = integer * 10^-scale)."""

    def __init__(self, ints, scale, precision=38):
        self.ints = [int(v) for v in ints]
        self.scale = int(scale)
        self.precision = int(precision)

    def __len__(self):
        return len(self.ints)

    def __getitem__(self, i):
        import decimal
        return decimal.Decimal(self.ints[i]).scaleb(-self.scale)


class Column:
    """One column of a table: typed numpy values plus an optional NULL mask
    (True = NULL).  Text columns hold a list/array of bytes objects; numeric
    columns a list of varlena images / decimal values, or DecimalValues
    (an Arrow Decimal128 column in build_arrow_kds)."""

    def __init__(self, name, typname, values, nulls=None):
        self.name = name
        self.typname = typname
        if typname in ("text", "bytea", "bpchar"):
            if isinstance(values, ArrowVarlena):
                self.values = values
            else:
                self.values = list(values)
        elif typname == "numeric":
            self.values = values if isinstance(values, DecimalValues) else list(values)
        else:
            self.values = np.ascontiguousarray(values, dtype=NP_DTYPES[typname])
        n = len(self.values)
        if nulls is not None:
            nulls = np.ascontiguousarray(nulls, dtype=np.bool_)
            assert len(nulls) == n
            if not nulls.any():
                nulls = None
        self.nulls = nulls

    def __len__(self):
        return len(self.values)

    def isnull(self, i):
        return self.nulls is not None and bool(self.nulls[i])


def _fill_colmeta(cm, kds_offset, fmt, j, name, typname, attcacheoff):
    oid, typlen, byval, align, _ = PG_TYPES[typname]
    cm.attbyval = byval
    cm.attalign = align
    cm.attlen = typlen
    cm.attnum = j + 1
    cm.attcacheoff = attcacheoff
    cm.atttypid = oid
    cm.atttypmod = -1
    cm.atttypkind = abi.TYPE_KIND__BASE
    cm.kds_format = fmt
    cm.kds_offset = kds_offset
    cm.attname = name.encode()[:63]
    cm.field_index = j


def kds_header(fmt, coldefs, length=0, hash_nslots=0):
    """Header of a KDS (struct + colmeta[]) as a bytearray of
    KDS_HEAD_LENGTH bytes.  coldefs: list of (name, typname)."""
    ncols = len(coldefs)
    head_sz = abi.kds_head_length(ncols)
    buf = bytearray(head_sz)
    head = abi.kern_data_store_head.from_buffer(buf)
    head.length = length
    head.usage = 0
    head.nitems = 0
    head.ncols = ncols
    head.format = fmt
    head.tdtypeid = 0
    head.tdtypmod = -1
    head.hash_nslots = hash_nslots
    head.nr_colmeta = ncols
    attcacheoff = 0 if fmt in (abi.KDS_FORMAT_ROW, abi.KDS_FORMAT_HASH,
                               abi.KDS_FORMAT_BLOCK) else -1
    for j, (name, typname) in enumerate(coldefs):
        off = abi.KDS_COLMETA_OFFSET + j * C.sizeof(abi.kern_colmeta)
        cm = abi.kern_colmeta.from_buffer(buf, off)
        _, typlen, _, align, _ = PG_TYPES[typname]
        if typlen == -1:
            head.has_varlena = True
        # relscan.c:200-228 attcacheoff rule
        if attcacheoff < 0:
            co = -1
        elif typlen > 0:
            co = abi.typealign(align, attcacheoff)
            attcacheoff = co + typlen
        else:
            a = abi.typealign(align, attcacheoff)
            co = a if a == attcacheoff else -1
            attcacheoff = -1
        _fill_colmeta(cm, off, fmt, j, name, typname, co)
    return buf


def build_arrow_kds(columns, text_layout="utf8"):
    """Columnar KDS (KDS_FORMAT_ARROW) holding `columns` (list of Column).
    text_layout sets the Arrow layout of text/bytea columns: "utf8" (Utf8 /
    Binary, 32-bit offsets), "large" (LargeUtf8 / LargeBinary, 64-bit
    offsets) or "fixed" (FixedSizeBinary; every non-NULL value of a column
    has the same length).  Returns a bytes-like numpy uint8 array."""
    n = len(columns[0]) if columns else 0
    coldefs = [(c.name, c.typname) for c in columns]
    head = kds_header(abi.KDS_FORMAT_ARROW, coldefs)
    pieces = []
    pos = abi.typealign(64, len(head))
    layout = []
    for c in columns:
        ent = {}
        if c.nulls is not None:
            valid = np.packbits(~c.nulls, bitorder="little")
            ent["nullmap"] = (pos, valid.tobytes())
            pos = abi.typealign(64, pos + len(valid))
        if c.typname in ("text", "bytea", "bpchar"):
            if isinstance(c.values, ArrowVarlena):
                offs = c.values.offsets
                data = c.values.data
            else:
                lens = np.fromiter((0 if (c.nulls is not None and c.nulls[i]) else len(v)
                                    for i, v in enumerate(c.values)), dtype=np.int64, count=n)
                offs = np.zeros(n + 1, dtype=np.int64)
                np.cumsum(lens, out=offs[1:])
                data = b"".join(v if not (c.nulls is not None and c.nulls[i]) else b""
                                for i, v in enumerate(c.values))
            if text_layout == "fixed":
                lens = np.diff(offs)
                live = lens if c.nulls is None else lens[~c.nulls]
                w = int(live[0]) if len(live) else 1
                if len(live) and (live != w).any():
                    raise ValueError("FixedSizeBinary needs values of one length")
                raw = bytearray(w * n)
                for i in range(n):
                    if offs[i + 1] > offs[i]:
                        raw[w * i:w * (i + 1)] = data[offs[i]:offs[i + 1]]
                ent["values"] = (pos, bytes(raw))
                ent["fixed_width"] = w
                pos = abi.typealign(64, pos + len(raw))
                layout.append(ent)
                continue
            if text_layout == "large":
                ent["values"] = (pos, offs.astype(np.int64).tobytes())
                pos = abi.typealign(64, pos + 8 * (n + 1))
            else:
                assert offs[-1] < 2**31
                ent["values"] = (pos, offs.astype(np.int32).tobytes())
                pos = abi.typealign(64, pos + 4 * (n + 1))
            ent["extra"] = (pos, data)
            pos = abi.typealign(64, pos + len(data))
        elif c.typname == "bool":
            bits = np.packbits(c.values, bitorder="little")
            ent["values"] = (pos, bits.tobytes())
            pos = abi.typealign(64, pos + len(bits))
        elif c.typname == "numeric":
            if not isinstance(c.values, DecimalValues):
                raise ValueError("an Arrow numeric column is a Decimal128 (DecimalValues)")
            raw = b"".join((v & ((1 << 128) - 1)).to_bytes(16, "little") for v in c.values.ints)
            ent["values"] = (pos, raw)
            ent["decimal"] = c.values
            pos = abi.typealign(64, pos + len(raw))
        else:
            raw = c.values.tobytes()
            ent["values"] = (pos, raw)
            pos = abi.typealign(64, pos + len(raw))
        layout.append(ent)
    total = pos
    out = np.zeros(total, dtype=np.uint8)
    out[:len(head)] = np.frombuffer(bytes(head), dtype=np.uint8)
    hd = abi.kern_data_store_head.from_buffer(out)
    hd.length = total
    hd.nitems = n
    hd.usage = 0
    for j, (c, ent) in enumerate(zip(columns, layout)):
        cm = abi.kern_colmeta.from_buffer(out, abi.KDS_COLMETA_OFFSET +
                                          j * C.sizeof(abi.kern_colmeta))
        t = c.typname
        if t in ("int1", "int2", "int4", "int8", "date"):
            cm.attopts.tag = abi.ArrowType__Int
            cm.attopts.unitsz = PG_TYPES[t][1]
            cm.attopts.u.integer.bitWidth = PG_TYPES[t][1] * 8
            cm.attopts.u.integer.is_signed = 1
        elif t in ("float4", "float8"):
            cm.attopts.tag = abi.ArrowType__FloatingPoint
            cm.attopts.unitsz = PG_TYPES[t][1]
            cm.attopts.u.floating_point.precision = (
                abi.ArrowPrecision__Single if t == "float4" else abi.ArrowPrecision__Double)
        elif t in ("text", "bytea", "bpchar") and "fixed_width" in ent:
            cm.attopts.tag = ARROW_FIXED_SIZE_BINARY
            cm.attopts.unitsz = ent["fixed_width"]
            cm.attopts.u.fixed_size_binary.byteWidth = ent["fixed_width"]
        elif t in ("text", "bytea", "bpchar") and text_layout == "large":
            cm.attopts.tag = ARROW_LARGE_UTF8 if t != "bytea" else ARROW_LARGE_BINARY
            cm.attopts.unitsz = 8
        elif t in ("text", "bytea", "bpchar"):
            cm.attopts.tag = abi.ArrowType__Utf8 if t != "bytea" else 4
            cm.attopts.unitsz = 4
        elif t == "bool":
            cm.attopts.tag = abi.ArrowType__Bool
        elif t == "numeric":
            cm.attopts.tag = abi.ArrowType__Decimal
            cm.attopts.unitsz = 16
            cm.attopts.u.decimal.precision = ent["decimal"].precision
            cm.attopts.u.decimal.scale = ent["decimal"].scale
            cm.attopts.u.decimal.bitWidth = 128
        for key in ("nullmap", "values", "extra"):
            if key in ent:
                off, raw = ent[key]
                out[off:off + len(raw)] = np.frombuffer(raw, dtype=np.uint8)
                setattr(cm, key + "_offset", off)
                setattr(cm, key + "_length", len(raw))
    return out


# ----------------------------------------------------------------------
# datum encoding in heap/minimal tuples
# ----------------------------------------------------------------------
_FMT = {"bool": "<?", "int1": "<b", "int2": "<h", "int4": "<i", "int8": "<q",
        "float4": "<f", "float8": "<d", "date": "<i"}


def numeric_image(value, short_varlena=False, long_form=False):
    """PostgreSQL numeric varlena of a decimal value ('NaN', 'Infinity' and
    '-Infinity' included), as numeric_in / make_result build it
    (utils/adt/numeric.c): base-10000 digits without leading or trailing
    zero digits, the NumericShort header when the display scale and weight
    fit it (or NumericLong if long_form), under a 4-byte varlena header, or
    the 1-byte one a short datum has on disk (short_varlena)."""
    import decimal
    d = decimal.Decimal(value) if not isinstance(value, decimal.Decimal) else value
    if d.is_nan():
        body = struct.pack("<H", 0xC000)
    elif d.is_infinite():
        body = struct.pack("<H", 0xF000 if d < 0 else 0xD000)
    else:
        sign, digits, exp = d.as_tuple()
        dscale = max(0, -exp)
        # decimal digits aligned so that the point falls on a group border
        ds = "".join(map(str, digits)) or "0"
        frac = max(0, -exp)
        ds = ds + "0" * max(0, exp)
        ds = ds.rjust(frac + 1, "0")
        ip = ds[:len(ds) - frac] if frac else ds
        fp = ds[len(ds) - frac:] if frac else ""
        ip = ip.lstrip("0")
        ip = "0" * ((-len(ip)) % 4) + ip
        fp = fp + "0" * ((-len(fp)) % 4)
        groups = [int(ip[i:i + 4]) for i in range(0, len(ip), 4)] + \
                 [int(fp[i:i + 4]) for i in range(0, len(fp), 4)]
        weight = len(ip) // 4 - 1
        while groups and groups[0] == 0:
            groups.pop(0)
            weight -= 1
        while groups and groups[-1] == 0:
            groups.pop()
        if not groups:
            weight, sign = 0, 0
        if not long_form and dscale <= 63 and -64 <= weight <= 63:
            head = 0x8000 | (0x2000 if sign else 0) | (dscale << 7) | (weight & 0x7F)
            body = struct.pack("<H", head)
        else:
            body = struct.pack("<Hh", (0x4000 if sign else 0) | dscale, weight)
        body += b"".join(struct.pack("<h", g) for g in groups)
    if short_varlena and len(body) + 1 <= 127:
        return bytes([((len(body) + 1) << 1) | 1]) + body
    return struct.pack("<I", (len(body) + 4) << 2) + body


def numeric_value(image):
    """The value of a numeric varlena image as a decimal.Decimal (its display
    scale ignored)."""
    import decimal
    b0 = image[0]
    body = image[1:(b0 >> 1)] if b0 & 1 else image[4:(struct.unpack_from("<I", image)[0] >> 2)]
    head = struct.unpack_from("<H", body)[0]
    if head & 0xC000 == 0xC000:
        return decimal.Decimal({0xC000: "NaN", 0xD000: "Infinity", 0xF000: "-Infinity"}[head])
    if head & 0xC000 == 0x8000:
        weight = (head & 0x3F) - (64 if head & 0x40 else 0)
        neg = bool(head & 0x2000)
        digs = body[2:]
    else:
        weight = struct.unpack_from("<h", body, 2)[0]
        neg = head & 0xC000 == 0x4000
        digs = body[4:]
    n = len(digs) // 2
    v = decimal.Decimal(0)
    with decimal.localcontext() as ctx:
        ctx.prec = 200
        for i in range(n):
            g = struct.unpack_from("<h", digs, 2 * i)[0]
            v += decimal.Decimal(g).scaleb(4 * (weight - i))
        return -v if neg else v


def encode_datum(typname, value):
    if isinstance(value, (ExternalDatum, CompressedDatum)):
        return value.encode()
    if typname == "numeric":
        return bytes(value) if isinstance(value, (bytes, bytearray)) else numeric_image(value)
    if typname in ("text", "bytea", "bpchar"):
        v = bytes(value)
        return struct.pack("<I", (len(v) + 4) << 2) + v
    return struct.pack(_FMT[typname], value)


def form_tuple_payload(coldefs, values):
    """(t_hoff, t_infomask, body bytes) of a minimal tuple holding values
    (None = NULL), in the layout of kern_form_minimal_tuple
    (xpu_common.cu:1096-1190)."""
    natts = len(coldefs)
    hasnull = any(v is None for v in values)
    t_hoff = abi.T_BITS_OFFSET + (abi.bitmaplen(natts) if hasnull else 0)
    t_hoff = abi.maxalign(t_hoff)
    infomask = abi.HEAP_HASNULL if hasnull else 0
    body = bytearray()
    bits = bytearray(abi.bitmaplen(natts))
    pos = t_hoff
    for j, ((_, t), v) in enumerate(zip(coldefs, values)):
        if v is None:
            continue
        align = PG_TYPES[t][3]
        nxt = abi.typealign(align, pos)
        body += b"\0" * (nxt - pos)
        d = encode_datum(t, v)
        body += d
        if PG_TYPES[t][1] == -1:
            if isinstance(v, ExternalDatum):
                infomask |= abi.HEAP_HASEXTERNAL
            infomask |= abi.HEAP_HASVARWIDTH
        bits[j >> 3] |= 1 << (j & 7)
        pos = nxt + len(d)
    return t_hoff, infomask, bits if hasnull else bytearray(), bytes(body)


def form_tupitem(coldefs, values, rowid, hash_value=0):
    """A complete kern_tupitem (header, null bitmap, payload, row-id)."""
    t_hoff, infomask, bits, body = form_tuple_payload(coldefs, values)
    t_len = t_hoff + len(body) + abi.ROWID_SIZE
    buf = bytearray(t_len)
    struct.pack_into("<IIBBHHB", buf, 0, t_len, hash_value, 0, 0,
                     len(coldefs) & abi.HEAP_NATTS_MASK, infomask,
                     t_hoff + abi.MINIMAL_TUPLE_OFFSET)
    buf[abi.T_BITS_OFFSET:abi.T_BITS_OFFSET + len(bits)] = bits
    buf[t_hoff:t_hoff + len(body)] = body
    struct.pack_into("<I", buf, t_len - 4, rowid)
    return bytes(buf)


def form_minimal_tuple(coldefs, values):
    """A MinimalTuple of PostgreSQL (heap_form_minimal_tuple): the
    kern_tupitem image without its row id, t_len counting the tuple only."""
    t = bytearray(form_tupitem(coldefs, values, 0)[:-abi.ROWID_SIZE])
    struct.pack_into("<I", t, 0, len(t))
    return bytes(t)


def build_hash_kds(coldefs, rows, hashes):
    """Inner relation as KDS_FORMAT_HASH, laid out like
    __innerPreloadSetupHashBuffer (gpu_join.c:2128-2164): items are packed
    from the tail, each slot holds the offset (from the tail) of the newest
    kern_hashitem, and hitem->next chains to the previous head."""
    n = len(rows)
    nslots = abi.kds_get_hashslot_width(n)
    head = kds_header(abi.KDS_FORMAT_HASH, coldefs, hash_nslots=nslots)
    items = [form_tupitem(coldefs, r, i, h) for i, (r, h) in enumerate(zip(rows, hashes))]
    usage = sum(abi.maxalign(8 + len(t)) for t in items)
    length = len(head) + 8 * (nslots + n) + usage
    length = abi.maxalign(length)
    out = np.zeros(length, dtype=np.uint8)
    out[:len(head)] = np.frombuffer(bytes(head), dtype=np.uint8)
    hd = abi.kern_data_store_head.from_buffer(out)
    hd.length = length
    hd.nitems = n
    hd.usage = usage
    body = len(head)
    slots = np.zeros(nslots, dtype=np.uint64)
    rowidx = np.zeros(n, dtype=np.uint64)
    curr = length
    for i, (t, h) in enumerate(zip(items, hashes)):
        curr -= abi.maxalign(8 + len(t))
        self_off = length - curr
        hidx = h % nslots
        nxt = int(slots[hidx])
        slots[hidx] = self_off
        struct.pack_into("<Q", out, curr, nxt)
        out[curr + 8:curr + 8 + len(t)] = np.frombuffer(t, dtype=np.uint8)
        rowidx[i] = length - (curr + 8)
    out[body:body + 8 * nslots] = slots.view(np.uint8)
    out[body + 8 * nslots:body + 8 * (nslots + n)] = rowidx.view(np.uint8)
    return out


class PinnedInner:
    """A kern_multirels chunk naming a pinned inner buffer on the device: the
    projection results of the session whose query_plan_id is buffer_id
    (gpu_join.c:1925-1932)."""

    def __init__(self, buffer_id):
        self.buffer_id = buffer_id

    def __len__(self):
        return 0


def build_multirels(inners):
    """kern_multirels holding the inner KDS buffers (list of
    (kds_bytes, left_outer), or (PinnedInner, left_outer) for a pinned
    buffer) one after another, as gpu_join.c lays out the preload buffer
    (kds_offset counted from the kmrels head)."""
    nrels = len(inners)
    head_sz = abi.maxalign(abi.KMRELS_CHUNKS_OFFSET +
                           C.sizeof(abi.kern_multirels_chunk) * nrels)
    offs = []
    pos = head_sz
    for kds, _ in inners:
        offs.append(0 if isinstance(kds, PinnedInner) else pos)
        if not isinstance(kds, PinnedInner):
            pos = abi.typealign(64, pos + len(kds))
    out = np.zeros(pos, dtype=np.uint8)
    hd = abi.kern_multirels_head.from_buffer(out)
    hd.length = pos
    hd.num_inner_rels = nrels
    for d, ((kds, left_outer), off) in enumerate(zip(inners, offs)):
        ch = abi.kern_multirels_chunk.from_buffer(
            out, abi.KMRELS_CHUNKS_OFFSET + d * C.sizeof(abi.kern_multirels_chunk))
        ch.left_outer = bool(left_outer)
        if isinstance(kds, PinnedInner):
            ch.kds_offset = 0
            ch.pinned_buffer = True
            ch.buffer_id = kds.buffer_id
            continue
        ch.kds_offset = off
        out[off:off + len(kds)] = np.frombuffer(bytes(kds), dtype=np.uint8)
    return out


# ----------------------------------------------------------------------
# parsers
# ----------------------------------------------------------------------
def kds_colmeta(buf):
    hd = abi.kern_data_store_head.from_buffer_copy(bytes(buf[:C.sizeof(abi.kern_data_store_head)]))
    cms = []
    for j in range(hd.nr_colmeta):
        off = abi.KDS_COLMETA_OFFSET + j * C.sizeof(abi.kern_colmeta)
        cms.append(abi.kern_colmeta.from_buffer_copy(bytes(buf[off:off + C.sizeof(abi.kern_colmeta)])))
    return hd, cms


def _decode_datum(cm, raw, pos):
    """(value, new_pos) of the attribute at pos (alignment already applied)."""
    if cm.attlen == -1:
        hdr = raw[pos]
        if hdr & 0x01:      # 1-byte header
            ln = (hdr >> 1) & 0x7f
            if cm.atttypid == 1700:     # numeric: the varlena image itself
                return bytes(raw[pos:pos + ln]), pos + ln
            return bytes(raw[pos + 1:pos + ln]), pos + ln
        ln = (struct.unpack_from("<I", raw, pos)[0] >> 2) & 0x3fffffff
        if cm.atttypid == 1700:
            return bytes(raw[pos:pos + ln]), pos + ln
        return bytes(raw[pos + 4:pos + ln]), pos + ln
    fmt = {1: "<b", 2: "<h", 4: "<i", 8: "<q"}[cm.attlen]
    if cm.atttypid == 700:
        fmt = "<f"
    elif cm.atttypid == 701:
        fmt = "<d"
    elif cm.atttypid == 16:
        fmt = "<?"
    return struct.unpack_from(fmt, raw, pos)[0], pos + cm.attlen


def decode_tupitem(cms, raw, off):
    """Decode the kern_tupitem at `off` of `raw`: returns (values, rowid, hash)."""
    t_len, hashv, _, _, infomask2, infomask, t_hoff = struct.unpack_from("<IIBBHHB", raw, off)
    natts = min(infomask2 & abi.HEAP_NATTS_MASK, len(cms))
    hasnull = bool(infomask & abi.HEAP_HASNULL)
    bits = raw[off + abi.T_BITS_OFFSET: off + abi.T_BITS_OFFSET + abi.bitmaplen(natts)]
    pos = off + t_hoff - abi.MINIMAL_TUPLE_OFFSET
    vals = []
    for j in range(natts):
        cm = cms[j]
        if hasnull and not (bits[j >> 3] & (1 << (j & 7))):
            vals.append(None)
            continue
        if cm.attlen > 0 or raw[pos] == 0:
            pos = off + abi.typealign(cm.attalign, pos - off)
        v, pos = _decode_datum(cm, raw, pos)
        vals.append(v)
    rowid = struct.unpack_from("<I", raw, off + t_len - 4)[0]
    return vals, rowid, hashv


def parse_row_kds(buf):
    """List of (values, rowid) of a KDS_FORMAT_ROW (or FALLBACK-less HASH)
    buffer, in row-index order."""
    raw = memoryview(bytes(buf))
    hd, cms = kds_colmeta(raw)
    head_sz = abi.kds_head_length(hd.nr_colmeta)
    ridx = np.frombuffer(raw, dtype=np.uint64, count=hd.nitems,
                         offset=head_sz + 8 * hd.hash_nslots)
    out = []
    for i in range(hd.nitems):
        o = int(ridx[i])
        if o == 0:
            out.append(None)
            continue
        vals, rowid, _ = decode_tupitem(cms, raw, hd.length - o)
        out.append((tuple(vals), rowid))
    return out


def parse_fallback_kds(buf):
    """The items of a KDS_FORMAT_FALLBACK buffer, in row-index order:
    (depth, l_state, matched, tuple bytes without the row id, row id)
    (kern_fallbackitem, xpu_common.h:1296-1305; KDS_GET_FALLBACK_ITEM)."""
    raw = memoryview(bytes(buf))
    hd, cms = kds_colmeta(raw)
    assert hd.format == abi.KDS_FORMAT_FALLBACK
    head_sz = abi.kds_head_length(hd.nr_colmeta)
    ridx = np.frombuffer(raw, dtype=np.uint64, count=hd.nitems, offset=head_sz)
    out = []
    for i in range(hd.nitems):
        t = hd.length - int(ridx[i])
        l_state, depth, matched = struct.unpack_from("<QHB", raw, t - 16)
        t_len = struct.unpack_from("<I", raw, t)[0]
        rowid = struct.unpack_from("<I", raw, t + t_len - 4)[0]
        out.append((depth, l_state, matched, bytes(raw[t:t + t_len - 4]), rowid))
    return out


def window_headers(base, windows, base_offset, stride):
    """KDS_FORMAT_ARROW headers of row windows over one columnar chunk.

    `base` is a chunk built by build_arrow_kds (fixed-width and utf8/binary
    columns; validity bitmaps need every window to start at a multiple of 8
    rows).  Header
    i, placed `i * stride` bytes after the start of a buffer that holds `base`
    at `base_offset`, describes rows [r0, r0 + n) of `base` for windows[i] =
    (r0, n), with its column offsets pointing into `base` (as an Arrow record
    batch producer hands out slices of one file).  Returns one uint8 array of
    len(windows) * stride bytes."""
    hd = abi.kern_data_store_head.from_buffer_copy(bytes(base[:C.sizeof(abi.kern_data_store_head)]))
    head_len = abi.kds_head_length(hd.ncols)
    assert head_len <= stride and stride % 64 == 0
    out = np.zeros(len(windows) * stride, dtype=np.uint8)
    proto = np.array(base[:head_len], dtype=np.uint8)
    for i, (r0, n) in enumerate(windows):
        h0 = i * stride
        out[h0:h0 + head_len] = proto
        h = abi.kern_data_store_head.from_buffer(out, h0)
        h.nitems = n
        h.length = base_offset + hd.length - h0
        shift = base_offset - h0
        for j in range(hd.ncols):
            cm = abi.kern_colmeta.from_buffer(out, h0 + abi.KDS_COLMETA_OFFSET +
                                              j * C.sizeof(abi.kern_colmeta))
            if cm.extra_offset:
                # utf8 / binary: the window's slice of the u32 offsets; the
                # offsets stay absolute into the shared extra buffer
                cm.values_offset = shift + cm.values_offset + r0 * 4
                cm.values_length = (n + 1) * 4
                cm.extra_offset = shift + cm.extra_offset
            else:
                w = cm.attopts.unitsz
                if w <= 0:
                    raise ValueError("window_headers: column without a fixed width")
                cm.values_offset = shift + cm.values_offset + r0 * w
                cm.values_length = n * w
            if cm.nullmap_offset:
                if r0 % 8:
                    raise ValueError("window_headers: windows over NULL-bearing columns "
                                     "must start at a multiple of 8 rows")
                cm.nullmap_offset = shift + cm.nullmap_offset + r0 // 8
                cm.nullmap_length = (n + 7) // 8
    return out


# ----------------------------------------------------------------------
# heap tuples: 8 KB pages (KDS_FORMAT_BLOCK) and row chunks (KDS_FORMAT_ROW)
# ----------------------------------------------------------------------
BLCKSZ = 8192
PAGE_HEADER_SIZE = 24
HTUP_HEADER_SIZE = 23
LP_UNUSED, LP_NORMAL, LP_REDIRECT, LP_DEAD = 0, 1, 2, 3
PD_ALL_VISIBLE = 0x0004
XACT_ATTRS_SIZE = 16        # sizeof(kern_tupitem_xact_attrs)
FROZEN_XID = 2


class ExternalDatum:
    """A TOAST pointer in a text/bytea column: 1-byte header 0x01, the tag,
    then the pointer body (16 bytes for VARTAG_ONDISK)."""

    def __init__(self, tag=18):
        self.tag = tag

    def encode(self):
        body = {1: 8, 18: 16}.get(self.tag, 0)
        return bytes([0x01, self.tag]) + bytes(range(1, body + 1))


class CompressedDatum:
    """An inline compressed datum in a text/bytea column: a 4-byte header
    with the compressed bit (VARATT_IS_4B_C), va_tcinfo (raw size, method in
    the top two bits), then the compressed bytes."""

    def __init__(self, rawsize, payload=b"\x00pglz"):
        self.rawsize = rawsize
        self.payload = bytes(payload)

    def encode(self):
        total = 8 + len(self.payload)
        return struct.pack("<II", (total << 2) | 0x02, self.rawsize & 0x3fffffff) + self.payload


def _varlena(v):
    """PostgreSQL's on-page varlena form: a 1-byte header for data of up to
    126 bytes (VARATT_CAN_MAKE_SHORT), else a 4-byte header.  Returns
    (bytes, needs_alignment)."""
    if isinstance(v, ExternalDatum):
        return v.encode(), False
    if isinstance(v, CompressedDatum):
        return v.encode(), True
    v = bytes(v)
    if len(v) + 1 <= 0x7f:
        return bytes([((len(v) + 1) << 1) | 0x01]) + v, False
    return struct.pack("<I", (len(v) + 4) << 2) + v, True


def heap_payload(coldefs, values, start=0):
    """(infomask, null bitmap, data bytes) of a tuple in heap_fill_tuple's
    layout: NULLs take no space, fixed-width values are aligned to typalign,
    short varlena values are not aligned, 4-byte-header ones are.  `start`
    is the data's offset from a MAXALIGN'ed base (alignment is relative to
    it); the data bytes returned start at `start`."""
    natts = len(coldefs)
    hasnull = any(v is None for v in values)
    infomask = abi.HEAP_HASNULL if hasnull else 0
    bits = bytearray(abi.bitmaplen(natts)) if hasnull else bytearray()
    body = bytearray()
    pos = start
    for j, ((_, t), v) in enumerate(zip(coldefs, values)):
        if v is None:
            continue
        if hasnull:
            bits[j >> 3] |= 1 << (j & 7)
        _, typlen, _, align, _ = PG_TYPES[t]
        if t == "numeric" and not isinstance(v, (ExternalDatum, CompressedDatum)):
            # a numeric image goes on the page as given; a decimal value in
            # its on-page form (a 1-byte header where it fits)
            d = bytes(v) if isinstance(v, (bytes, bytearray)) else numeric_image(v, short_varlena=True)
            aligned = not (d[0] & 0x01)
            infomask |= abi.HEAP_HASVARWIDTH
            if aligned:
                nxt = abi.typealign(align, pos)
                body += b"\0" * (nxt - pos)
                pos = nxt
        elif typlen == -1:
            d, aligned = _varlena(v)
            infomask |= abi.HEAP_HASVARWIDTH
            if aligned:
                nxt = abi.typealign(align, pos)
                body += b"\0" * (nxt - pos)
                pos = nxt
        else:
            d = encode_datum(t, v)
            nxt = abi.typealign(align, pos)
            body += b"\0" * (nxt - pos)
            pos = nxt
        body += d
        pos += len(d)
    return infomask, bytes(bits), bytes(body)


def form_heap_tuple(coldefs, values, xmin=FROZEN_XID, xmax=0, ctid=(0, 0)):
    """HeapTupleHeaderData + null bitmap + data, as heap_form_tuple lays it
    out (t_hoff = MAXALIGN(23 + bitmap))."""
    natts = len(coldefs)
    hasnull = any(v is None for v in values)
    hoff = abi.maxalign(HTUP_HEADER_SIZE + (abi.bitmaplen(natts) if hasnull else 0))
    infomask, bits, body = heap_payload(coldefs, values, 0)
    blk, posid = ctid
    head = struct.pack("<IIIHHHHHB", xmin, xmax, 0, (blk >> 16) & 0xffff, blk & 0xffff, posid,
                       natts & abi.HEAP_NATTS_MASK, infomask | 0x0900, hoff)
    return head + bits + b"\0" * (hoff - HTUP_HEADER_SIZE - len(bits)) + body


def build_block_kds(columns, invisible=None, extra_items=False, max_per_page=None, seed=0,
                    empty_pages=()):
    """KDS_FORMAT_BLOCK chunk (relscan.c:554-742 layout): the rows of
    `columns` as heap tuples packed into 8 KB pages (PageAddItem order:
    line pointers up from the header, tuples MAXALIGN'ed down from the end),
    the block numbers at the KDS body, the pages at block_offset.

    invisible[i] marks rows the host's visibility check rejected: their line
    pointer is set LP_UNUSED as relscan.c:480-495 does (the tuple bytes stay
    on the page).  extra_items adds LP_DEAD items with storage and
    LP_REDIRECT items between the rows; max_per_page caps the tuples per
    page; empty_pages lists page numbers left without items.  The scan loads
    the rows that are not invisible, in row order."""
    n = len(columns[0]) if columns else 0
    coldefs = [(c.name, c.typname) for c in columns]
    rng = np.random.default_rng(seed)
    invisible = np.zeros(n, dtype=bool) if invisible is None else np.asarray(invisible, dtype=bool)

    def value(c, i):
        if c.nulls is not None and c.nulls[i]:
            return None
        v = c.values[i]
        return v.item() if hasattr(v, "item") else v

    pages = []
    page = None

    def new_page():
        p = {"buf": bytearray(BLCKSZ), "lps": [], "upper": BLCKSZ, "ntup": 0}
        pages.append(p)
        return p

    def add_item(p, raw, flags):
        alen = abi.maxalign(len(raw))
        if PAGE_HEADER_SIZE + 4 * (len(p["lps"]) + 1) > p["upper"] - alen:
            return False
        p["upper"] -= alen
        off = p["upper"]
        p["buf"][off:off + len(raw)] = raw
        p["lps"].append(off | (flags << 15) | (len(raw) << 17))
        return True

    page = new_page()
    while len(pages) - 1 in empty_pages:
        page = new_page()
    for i in range(n):
        tup = form_heap_tuple(coldefs, [value(c, i) for c in columns],
                              ctid=(len(pages) - 1, len(page["lps"]) + 1))
        if extra_items and rng.random() < 0.05:
            # a dead item keeping its storage, a redirect to the previous item
            dead = form_heap_tuple(coldefs, [value(c, i) for c in columns])
            if add_item(page, dead, LP_DEAD) and \
                    PAGE_HEADER_SIZE + 4 * (len(page["lps"]) + 1) <= page["upper"]:
                page["lps"].append(max(1, len(page["lps"]) - 1) | (LP_REDIRECT << 15))
        full = max_per_page is not None and page["ntup"] >= max_per_page
        if full or not add_item(page, tup, LP_NORMAL):
            page = new_page()
            while len(pages) - 1 in empty_pages:
                page = new_page()
            if not add_item(page, tup, LP_NORMAL):
                raise ValueError("a tuple does not fit into an empty page")
        page["ntup"] += 1
        if invisible[i]:
            page["lps"][-1] = 0         # ItemIdSetUnused
    npages = len(pages)
    head = kds_header(abi.KDS_FORMAT_BLOCK, coldefs)
    block_offset = abi.maxalign(len(head) + 4 * npages)
    length = block_offset + BLCKSZ * npages
    out = np.zeros(length, dtype=np.uint8)
    out[:len(head)] = np.frombuffer(bytes(head), dtype=np.uint8)
    hd = abi.kern_data_store_head.from_buffer(out)
    hd.length = length
    hd.nitems = npages
    hd.block_offset = block_offset
    hd.block_nloaded = npages
    blknums = np.arange(npages, dtype=np.uint32) + 1000
    out[len(head):len(head) + 4 * npages] = blknums.view(np.uint8)
    for b, p in enumerate(pages):
        buf = p["buf"]
        nlp = len(p["lps"])
        struct.pack_into("<IIHHHHHHI", buf, 0, 0, 0, 0, PD_ALL_VISIBLE,
                         PAGE_HEADER_SIZE + 4 * nlp, p["upper"], BLCKSZ, BLCKSZ | 4, 0)
        if nlp:
            buf[PAGE_HEADER_SIZE:PAGE_HEADER_SIZE + 4 * nlp] = \
                np.array(p["lps"], dtype=np.uint32).tobytes()
        o = block_offset + BLCKSZ * b
        out[o:o + BLCKSZ] = np.frombuffer(bytes(buf), dtype=np.uint8)
    return out


def form_row_tupitem(coldefs, values, rowid, xact=None):
    """kern_tupitem of a row chunk.  With xact = (xmin, xmax) the item embeds
    kern_tupitem_xact_attrs in front of its payload, laid out as the GPU
    cache writes it (gpu_cache.c:868-960): header MAXALIGN(offsetof(t_bits)
    + bitmap + 16), payload from there, INTALIGN'ed, then the row-id."""
    natts = len(coldefs)
    hasnull = any(v is None for v in values)
    bl = abi.bitmaplen(natts) if hasnull else 0
    head_sz = abi.maxalign(abi.T_BITS_OFFSET + bl + (XACT_ATTRS_SIZE if xact else 0))
    infomask, bits, body = heap_payload(coldefs, values, 0)
    t_len = abi.typealign(4, head_sz + len(body)) + abi.ROWID_SIZE
    buf = bytearray(t_len)
    struct.pack_into("<IIBBHHB", buf, 0, t_len, 0, 1 if xact else 0, 0,
                     natts & abi.HEAP_NATTS_MASK, infomask, head_sz + abi.MINIMAL_TUPLE_OFFSET)
    buf[abi.T_BITS_OFFSET:abi.T_BITS_OFFSET + len(bits)] = bits
    if xact:
        struct.pack_into("<IIHHHH", buf, head_sz - XACT_ATTRS_SIZE, xact[0], xact[1], 0, 0,
                         rowid & 0xffff, 1)
    buf[head_sz:head_sz + len(body)] = body
    struct.pack_into("<I", buf, t_len - 4, rowid)
    return bytes(buf)


def build_row_kds(columns, xact=None, holes=()):
    """KDS_FORMAT_ROW chunk of the rows of `columns` (row index of u64
    offsets from the tail, items packed towards the tail).  xact: per-row
    (xmin, xmax) to embed, or None; holes: row-index entries left 0 (no
    tuple)."""
    n = len(columns[0]) if columns else 0
    coldefs = [(c.name, c.typname) for c in columns]

    def value(c, i):
        if c.nulls is not None and c.nulls[i]:
            return None
        v = c.values[i]
        return v.item() if hasattr(v, "item") else v

    items = [form_row_tupitem(coldefs, [value(c, i) for c in columns], i,
                              None if xact is None else xact[i]) for i in range(n)]
    head = kds_header(abi.KDS_FORMAT_ROW, coldefs)
    usage = sum(abi.maxalign(len(t)) for t in items)
    length = abi.maxalign(len(head) + 8 * n + usage)
    out = np.zeros(length, dtype=np.uint8)
    out[:len(head)] = np.frombuffer(bytes(head), dtype=np.uint8)
    hd = abi.kern_data_store_head.from_buffer(out)
    hd.length = length
    hd.nitems = n
    hd.usage = usage
    rowidx = np.zeros(n, dtype=np.uint64)
    curr = length
    holes = set(holes)
    for i, t in enumerate(items):
        curr -= abi.maxalign(len(t))
        out[curr:curr + len(t)] = np.frombuffer(t, dtype=np.uint8)
        rowidx[i] = 0 if i in holes else length - curr
    out[len(head):len(head) + 8 * n] = rowidx.view(np.uint8)
    return out
