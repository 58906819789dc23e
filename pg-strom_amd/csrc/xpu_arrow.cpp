/*
 * xpu_arrow.cpp - Arrow IPC file chunk producer: one record batch of an
 * Arrow file as a KDS_FORMAT_ARROW chunk for the GPU path.
 *
 * The reference's Arrow_Fdw reads the file's footer once (schema and the
 * record batch blocks, arrow_read.c), then for every record batch sets up a
 * KDS head whose colmeta address the batch's buffers and reads the body
 * into the chunk behind the head (arrowFdwLoadRecordBatch /
 * arrowFdwFillupRecordBatch, arrow_fdw.c:4018-4100; attribute options,
 * __arrowKdsAssignAttrOptions).  This file does the same from a file in
 * memory: a small reader of the flatbuffers the IPC format uses (Footer,
 * Schema, Field, Message, RecordBatch), the type check of each referenced
 * column against its field, and the copy of the batch's buffers.  Only
 * uncompressed files with flat types are read: Int, FloatingPoint, Bool,
 * Decimal128, Utf8 / Binary, LargeUtf8 / LargeBinary and FixedSizeBinary
 * (a Date column would need the epoch shift from 1970 to 2000, which the
 * device's loads do not apply).
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdio.h>
#include <string>
#include <vector>

#include "pgstrom_xpu.h"
#include "xpuhip.h"
#include "xpu_compile.h"

namespace {

/* ------------------------------------------------------------------
 * flatbuffers: a table starts with the signed offset of its vtable
 * (table - soffset); the vtable holds its own size, the table size and one
 * 16-bit field offset per field id (0 = absent).  Offsets to tables,
 * vectors and strings are unsigned 32-bit, relative to where they are
 * stored.  Every access is checked against the buffer.
 * ------------------------------------------------------------------ */
struct FB
{
    const uint8_t *base;
    size_t      len;
    bool        bad = false;

    template <class T> T rd(size_t off)
    {
        T v{};
        if (off + sizeof(T) > len || off + sizeof(T) < off)
        {
            bad = true;
            return v;
        }
        memcpy(&v, base + off, sizeof(T));
        return v;
    }
    /* offset of field `id` within table t, or 0 if absent */
    size_t field(size_t t, int id)
    {
        int32_t so = rd<int32_t>(t);
        size_t vt = (size_t)((int64_t)t - so);
        uint16_t vlen = rd<uint16_t>(vt);
        if (bad || 4 + 2 * id + 2 > vlen)
            return 0;
        uint16_t fo = rd<uint16_t>(vt + 4 + 2 * id);
        return fo ? t + fo : 0;
    }
    template <class T> T scalar(size_t t, int id, T dflt)
    {
        size_t f = field(t, id);
        return f ? rd<T>(f) : dflt;
    }
    /* table / vector / string referenced by field `id` (0 if absent) */
    size_t ref(size_t t, int id)
    {
        size_t f = field(t, id);
        if (!f)
            return 0;
        uint32_t o = rd<uint32_t>(f);
        return o ? f + o : 0;
    }
    uint32_t vlen(size_t v) { return v ? rd<uint32_t>(v) : 0; }
    /* element i of a vector of tables */
    size_t vtable_at(size_t v, uint32_t i)
    {
        size_t e = v + 4 + 4 * (size_t)i;
        return e + rd<uint32_t>(e);
    }
    std::string str(size_t s)
    {
        if (!s)
            return std::string();
        uint32_t n = rd<uint32_t>(s);
        if (bad || s + 4 + n > len)
        {
            bad = true;
            return std::string();
        }
        return std::string((const char *)base + s + 4, n);
    }
};

/* Arrow Type union tags (Schema.fbs) */
enum { AT_Int = 2, AT_FloatingPoint = 3, AT_Binary = 4, AT_Utf8 = 5, AT_Bool = 6, AT_Decimal = 7,
       AT_Date = 8, AT_FixedSizeBinary = 15, AT_LargeBinary = 19, AT_LargeUtf8 = 20 };
enum { MH_RecordBatch = 3 };

struct ArrowField
{
    std::string name;
    int         type = 0;
    int         bit_width = 0;      /* Int / Decimal */
    bool        is_signed = true;
    int         precision = 0;      /* FloatingPoint: 0 half, 1 single, 2 double; Decimal */
    int         scale = 0;
    int         unit = 0;           /* Date */
    int         byte_width = 0;     /* FixedSizeBinary */
    bool        nested = false;
};

struct ArrowBlock
{
    int64_t offset;
    int32_t meta_len;
    int64_t body_len;
};

struct ArrowFile
{
    std::vector<uint8_t> data;
    std::vector<ArrowField> fields;
    std::vector<ArrowBlock> batches;
};

void
set_error(kern_errorbuf *eb, int line, const char *func, const std::string &msg)
{
    xpu_set_error(eb, ERRCODE_DEVICE_ERROR, "xpu_arrow.cpp", line, func, msg.c_str());
}

/* the footer of an Arrow IPC file: "ARROW1" .. footer, int32 length,
 * "ARROW1" (File.fbs Footer: version, schema, dictionaries, recordBatches) */
bool
read_footer(ArrowFile *af, std::string *err)
{
    const std::vector<uint8_t> &d = af->data;
    if (d.size() < 18 || memcmp(d.data(), "ARROW1", 6) != 0 ||
        memcmp(d.data() + d.size() - 6, "ARROW1", 6) != 0)
    {
        *err = "not an Arrow IPC file (magic)";
        return false;
    }
    int32_t flen;
    memcpy(&flen, d.data() + d.size() - 10, 4);
    if (flen <= 0 || (size_t)flen > d.size() - 18)
    {
        *err = "Arrow footer length out of the file";
        return false;
    }
    FB fb{d.data() + d.size() - 10 - flen, (size_t)flen};
    size_t footer = fb.rd<uint32_t>(0);
    size_t schema = fb.ref(footer, 1);
    size_t fields = fb.ref(schema, 1);
    uint32_t nf = fb.vlen(fields);
    for (uint32_t i = 0; i < nf && !fb.bad; i++)
    {
        size_t f = fb.vtable_at(fields, i);
        ArrowField af_;
        af_.name = fb.str(fb.ref(f, 0));
        af_.type = fb.scalar<uint8_t>(f, 2, 0);
        size_t t = fb.ref(f, 3);
        size_t children = fb.ref(f, 5);
        af_.nested = fb.vlen(children) > 0 || fb.ref(f, 4) != 0;   /* children or dictionary */
        switch (af_.type)
        {
            case AT_Int:
                af_.bit_width = fb.scalar<int32_t>(t, 0, 0);
                af_.is_signed = fb.scalar<uint8_t>(t, 1, 0) != 0;
                break;
            case AT_FloatingPoint:
                af_.precision = fb.scalar<int16_t>(t, 0, 0);
                break;
            case AT_Decimal:
                af_.precision = fb.scalar<int32_t>(t, 0, 0);
                af_.scale = fb.scalar<int32_t>(t, 1, 0);
                af_.bit_width = fb.scalar<int32_t>(t, 2, 128);
                break;
            case AT_Date:
                af_.unit = fb.scalar<int16_t>(t, 0, 1);
                break;
            case AT_FixedSizeBinary:
                af_.byte_width = fb.scalar<int32_t>(t, 0, 0);
                break;
        }
        af->fields.push_back(af_);
    }
    size_t rbs = fb.ref(footer, 3);
    uint32_t nb = fb.vlen(rbs);
    for (uint32_t i = 0; i < nb && !fb.bad; i++)
    {
        /* struct Block { long offset; int metaDataLength; (pad); long bodyLength } */
        size_t b = rbs + 4 + 24 * (size_t)i;
        ArrowBlock blk;
        blk.offset = fb.rd<int64_t>(b);
        blk.meta_len = fb.rd<int32_t>(b + 8);
        blk.body_len = fb.rd<int64_t>(b + 16);
        if (blk.offset < 8 || blk.meta_len <= 0 || blk.body_len < 0 ||
            (uint64_t)blk.offset + (uint64_t)blk.meta_len + (uint64_t)blk.body_len > d.size())
        {
            *err = "Arrow record batch block out of the file";
            return false;
        }
        af->batches.push_back(blk);
    }
    if (fb.bad)
    {
        *err = "Arrow footer is corrupted";
        return false;
    }
    return true;
}

/* the PostgreSQL type a column must have for its Arrow field, by atttypid
 * (__arrowKdsAssignAttrOptions and the type mapping of arrow_fdw.c) */
bool
type_matches(const kern_colmeta *cm, const ArrowField &f)
{
    switch (cm->atttypid)
    {
        case 16: return f.type == AT_Bool;
        case 21: return f.type == AT_Int && f.bit_width == 16;
        case 23: return f.type == AT_Int && f.bit_width == 32;
        case 20: return f.type == AT_Int && f.bit_width == 64;
        case 700: return f.type == AT_FloatingPoint && f.precision == 1;
        case 701: return f.type == AT_FloatingPoint && f.precision == 2;
        case 1700: return f.type == AT_Decimal && f.bit_width == 128;
        case 25: case 1043: case 1042:
            return f.type == AT_Utf8 || f.type == AT_LargeUtf8 || f.type == AT_FixedSizeBinary;
        case 17:
            return f.type == AT_Binary || f.type == AT_LargeBinary || f.type == AT_FixedSizeBinary;
    }
    return false;
}

/* the buffers a flat field has in a record batch: validity, then values
 * (or offsets and data) */
int
nbuffers(const ArrowField &f)
{
    return (f.type == AT_Utf8 || f.type == AT_Binary || f.type == AT_LargeUtf8 ||
            f.type == AT_LargeBinary) ? 3 : 2;
}

}   /* namespace */

struct xpuhip_arrow_file
{
    ArrowFile   af;
};

extern "C" int
xpuhip_arrow_open(const void *image, size_t length, xpuhip_arrow_file **p_file,
                  uint32_t *p_nbatches, uint32_t *p_nfields, kern_errorbuf *errbuf)
{
    if (!image || !p_file)
    {
        set_error(errbuf, __LINE__, __func__, "null argument");
        return -1;
    }
    xpuhip_arrow_file *f = new xpuhip_arrow_file;
    f->af.data.assign((const uint8_t *)image, (const uint8_t *)image + length);
    std::string err;
    if (!read_footer(&f->af, &err))
    {
        delete f;
        set_error(errbuf, __LINE__, __func__, err);
        return -1;
    }
    *p_file = f;
    if (p_nbatches)
        *p_nbatches = (uint32_t)f->af.batches.size();
    if (p_nfields)
        *p_nfields = (uint32_t)f->af.fields.size();
    return 0;
}

extern "C" int
xpuhip_arrow_field(xpuhip_arrow_file *f, uint32_t index, char *name, size_t name_len,
                   int32_t *p_type, int32_t *p_bit_width, int32_t *p_scale)
{
    if (!f || index >= f->af.fields.size())
        return -1;
    const ArrowField &af = f->af.fields[index];
    if (name && name_len)
        snprintf(name, name_len, "%s", af.name.c_str());
    if (p_type)
        *p_type = af.type;
    if (p_bit_width)
        *p_bit_width = af.type == AT_FloatingPoint ? af.precision :
                       af.type == AT_FixedSizeBinary ? af.byte_width : af.bit_width;
    if (p_scale)
        *p_scale = af.scale;
    return 0;
}

extern "C" int64_t
xpuhip_arrow_load_batch(xpuhip_arrow_file *f, uint32_t rb_index, const kern_data_store *kds_head,
                        const int32_t *field_index, kern_data_store *kds, size_t length,
                        kern_errorbuf *errbuf)
{
    if (!f || !kds_head || !field_index)
    {
        set_error(errbuf, __LINE__, __func__, "null argument");
        return -1;
    }
    const ArrowFile &af = f->af;
    if (rb_index >= af.batches.size())
    {
        set_error(errbuf, __LINE__, __func__, "record batch index out of range");
        return -1;
    }
    if (kds_head->format != KDS_FORMAT_ARROW || kds_head->nr_colmeta < kds_head->ncols)
    {
        set_error(errbuf, __LINE__, __func__, "kds head is not KDS_FORMAT_ARROW");
        return -1;
    }
    /* the batch's message: 0xFFFFFFFF, int32 length, flatbuffer Message */
    const ArrowBlock &blk = af.batches[rb_index];
    const uint8_t *msg = af.data.data() + blk.offset;
    uint32_t cont;
    int32_t mlen;
    memcpy(&cont, msg, 4);
    memcpy(&mlen, msg + 4, 4);
    size_t fb_off = 8;
    if (cont != 0xFFFFFFFFU)
    {
        /* the pre-1.0 form without the continuation marker */
        mlen = (int32_t)cont;
        fb_off = 4;
    }
    if (mlen <= 0 || fb_off + (size_t)mlen > (size_t)blk.meta_len)
    {
        set_error(errbuf, __LINE__, __func__, "record batch message out of its block");
        return -1;
    }
    FB fb{msg + fb_off, (size_t)mlen};
    size_t message = fb.rd<uint32_t>(0);
    if (fb.scalar<uint8_t>(message, 1, 0) != MH_RecordBatch)
    {
        set_error(errbuf, __LINE__, __func__, "block is not a record batch");
        return -1;
    }
    size_t rb = fb.ref(message, 2);
    const int64_t nrows = fb.scalar<int64_t>(rb, 0, 0);
    size_t nodes = fb.ref(rb, 1), buffers = fb.ref(rb, 2);
    if (fb.ref(rb, 3))
    {
        set_error(errbuf, __LINE__, __func__, "compressed record batches are not read by this build");
        return -1;
    }
    const uint32_t nnodes = fb.vlen(nodes), nbufs = fb.vlen(buffers);
    if (fb.bad || nrows < 0 || nrows > (int64_t)UINT32_MAX)
    {
        set_error(errbuf, __LINE__, __func__, "record batch metadata is corrupted");
        return -1;
    }
    /* node and first buffer of every top-level field (flat fields only) */
    std::vector<uint32_t> node_of(af.fields.size(), UINT32_MAX), buf_of(af.fields.size(), 0);
    {
        /* positions from a nested field on are not known (its children
         * take nodes and buffers of their own) */
        uint32_t n = 0, b = 0;
        for (size_t i = 0; i < af.fields.size() && !af.fields[i].nested; i++)
        {
            node_of[i] = n++;
            buf_of[i] = b;
            b += nbuffers(af.fields[i]);
        }
        if (n > nnodes || b > nbufs)
        {
            set_error(errbuf, __LINE__, __func__, "record batch has fewer nodes or buffers than the schema");
            return -1;
        }
    }
    const uint8_t *body = msg + blk.meta_len;
    /* layout: the head, then each referenced buffer at a 64-byte boundary */
    const uint64_t head_len = KDS_HEAD_LENGTH(kds_head);
    struct Piece { uint64_t dst; const uint8_t *src; uint64_t len; };
    std::vector<Piece> pieces;
    std::vector<kern_colmeta> cms(kds_head->colmeta, kds_head->colmeta + kds_head->nr_colmeta);
    uint64_t pos = (head_len + 63) & ~63ULL;
    for (uint32_t j = 0; j < kds_head->ncols; j++)
    {
        kern_colmeta *cm = &cms[j];
        cm->kds_format = KDS_FORMAT_ARROW;
        cm->nullmap_offset = cm->nullmap_length = 0;
        cm->values_offset = cm->values_length = 0;
        cm->extra_offset = cm->extra_length = 0;
        const int32_t fi = field_index[j];
        cm->field_index = fi;
        if (fi < 0)
            continue;           /* not referenced */
        if ((size_t)fi >= af.fields.size())
        {
            set_error(errbuf, __LINE__, __func__, "field index out of the schema");
            return -1;
        }
        const ArrowField &fd = af.fields[fi];
        if (fd.nested || node_of[fi] == UINT32_MAX)
        {
            set_error(errbuf, __LINE__, __func__, "nested Arrow fields are not read by this build");
            return -1;
        }
        if (!type_matches(cm, fd))
        {
            char m[160];
            snprintf(m, sizeof(m), "column %u cannot be read from Arrow field \"%s\" (type %d)",
                     j + 1, fd.name.c_str(), fd.type);
            set_error(errbuf, __LINE__, __func__, m);
            return -1;
        }
        /* __arrowKdsAssignAttrOptions */
        memset(&cm->attopts, 0, sizeof(cm->attopts));
        cm->attopts.tag = fd.type;
        switch (fd.type)
        {
            case AT_Int:
                cm->attopts.unitsz = (int16_t)(fd.bit_width / 8);
                cm->attopts.u.integer.bitWidth = (uint16_t)fd.bit_width;
                cm->attopts.u.integer.is_signed = fd.is_signed;
                break;
            case AT_FloatingPoint:
                cm->attopts.unitsz = fd.precision == 2 ? 8 : fd.precision == 1 ? 4 : 2;
                cm->attopts.u.floating_point.precision = fd.precision;
                break;
            case AT_Decimal:
                cm->attopts.unitsz = 16;
                cm->attopts.u.decimal.precision = (uint16_t)fd.precision;
                cm->attopts.u.decimal.scale = (uint16_t)fd.scale;
                cm->attopts.u.decimal.bitWidth = (uint16_t)fd.bit_width;
                break;
            case AT_FixedSizeBinary:
                cm->attopts.unitsz = (int16_t)fd.byte_width;
                cm->attopts.u.fixed_size_binary.byteWidth = (uint32_t)fd.byte_width;
                break;
            case AT_Utf8: case AT_Binary:
                cm->attopts.unitsz = 4;
                break;
            case AT_LargeUtf8: case AT_LargeBinary:
                cm->attopts.unitsz = 8;
                break;
        }
        /* struct FieldNode { long length; long null_count } */
        const size_t node = nodes + 4 + 16 * (size_t)node_of[fi];
        const int64_t nlen = fb.rd<int64_t>(node), nnull = fb.rd<int64_t>(node + 8);
        if (nlen != nrows)
        {
            set_error(errbuf, __LINE__, __func__, "field length differs from the record batch's");
            return -1;
        }
        const int nb = nbuffers(fd);
        uint64_t *offs[3] = { &cm->nullmap_offset, &cm->values_offset, &cm->extra_offset };
        uint64_t *lens[3] = { &cm->nullmap_length, &cm->values_length, &cm->extra_length };
        for (int k = 0; k < nb; k++)
        {
            /* struct Buffer { long offset; long length } */
            const size_t b = buffers + 4 + 16 * (size_t)(buf_of[fi] + k);
            const int64_t boff = fb.rd<int64_t>(b), blen = fb.rd<int64_t>(b + 8);
            if (fb.bad || boff < 0 || blen < 0 || (uint64_t)boff + (uint64_t)blen > (uint64_t)blk.body_len)
            {
                set_error(errbuf, __LINE__, __func__, "record batch buffer out of its body");
                return -1;
            }
            if (k == 0 && (nnull == 0 || blen == 0))
                continue;       /* no validity bitmap: every value is valid */
            *offs[k] = pos;
            *lens[k] = (uint64_t)blen;
            pieces.push_back(Piece{pos, body + boff, (uint64_t)blen});
            pos = (pos + (uint64_t)blen + 63) & ~63ULL;
        }
    }
    if (fb.bad)
    {
        set_error(errbuf, __LINE__, __func__, "record batch metadata is corrupted");
        return -1;
    }
    if (!kds || length < pos)
        return (int64_t)pos;    /* the size needed */
    memset(kds, 0, head_len);
    memcpy(kds, kds_head, head_len);
    memcpy(kds->colmeta, cms.data(), sizeof(kern_colmeta) * cms.size());
    kds->length = pos;
    kds->nitems = (uint32_t)nrows;
    kds->usage = 0;
    kds->format = KDS_FORMAT_ARROW;
    uint64_t prev = head_len;
    for (const Piece &p : pieces)
    {
        memset((char *)kds + prev, 0, p.dst - prev);
        memcpy((char *)kds + p.dst, p.src, p.len);
        prev = p.dst + p.len;
    }
    memset((char *)kds + prev, 0, pos - prev);
    return (int64_t)pos;
}

extern "C" void
xpuhip_arrow_close(xpuhip_arrow_file *f)
{
    delete f;
}
