/*
 * xpu_inner.cpp - host-side build of an inner relation's KDS_FORMAT_HASH
 * buffer for the GpuJoin hash probe.
 *
 * The backend preloads each inner relation into one buffer before the join
 * runs (GpuJoinInnerPreload, gpu_join.c:2180-).  Two steps of it are done
 * here, natively and over all host cores:
 *   - the hash of every row's join keys (get_tuple_hashvalue,
 *     gpu_join.c:1548-1584): ~0, merged with the device type's hash of each
 *     key (codegen.c:544-660, 0 for a NULL) by pg_hash_merge, complemented;
 *   - the layout (__innerPreloadSetupHashBuffer, gpu_join.c:2128-2164): rows
 *     are packed from the buffer's tail as kern_hashitems in input order,
 *     each slot holds the tail offset of its newest item, an item's `next`
 *     the slot's previous head, and the row index the tail offset of every
 *     item's tuple.
 *
 * The byte image equals what the serial loop of the backend writes: item
 * positions come from a prefix sum of the item sizes, and the slot chains
 * are threaded by threads that each own a range of slots and visit the rows
 * in input order.
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdio.h>
#include <thread>
#include <vector>
#include <algorithm>

#include "pgstrom_xpu.h"
#include "xpuhip.h"
#include "xpu_compile.h"

namespace {

/* PostgreSQL type OIDs of the key types this build hashes */
enum : uint32_t {
    OID_BOOL = 16, OID_BYTEA = 17, OID_INT8 = 20, OID_INT2 = 21, OID_INT4 = 23,
    OID_TEXT = 25, OID_FLOAT4 = 700, OID_FLOAT8 = 701, OID_MONEY = 790,
    OID_BPCHAR = 1042, OID_VARCHAR = 1043, OID_DATE = 1082, OID_TIME = 1083,
    OID_TIMESTAMP = 1114, OID_TIMESTAMPTZ = 1184, OID_NUMERIC = 1700, OID_UUID = 2950,
};

inline uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* Bob Jenkins' lookup3 hashlittle() with PostgreSQL's seed, as hash_any
 * computes it (xpu_common.cu:2920-3041) */
uint32_t
lookup3(const uint8_t *k, uint32_t len)
{
    uint32_t a, b, c;
    a = b = c = 0x9e3779b9U + len + 3923095U;
    auto word = [](const uint8_t *p) {
        return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    };
    while (len >= 12)
    {
        a += word(k); b += word(k + 4); c += word(k + 8);
        a -= c; a ^= rotl(c, 4);  c += b;
        b -= a; b ^= rotl(a, 6);  a += c;
        c -= b; c ^= rotl(b, 8);  b += a;
        a -= c; a ^= rotl(c, 16); c += b;
        b -= a; b ^= rotl(a, 19); a += c;
        c -= b; c ^= rotl(b, 4);  b += a;
        k += 12;
        len -= 12;
    }
    uint8_t tail[12] = {0};
    memcpy(tail, k, len);
    if (len >= 9)
        c += (word(tail + 8) & 0x00ffffffU) << 8;
    if (len >= 5)
        b += word(tail + 4);
    if (len >= 1)
        a += word(tail);
    c ^= b; c -= rotl(b, 14);
    a ^= c; a -= rotl(c, 11);
    b ^= a; b -= rotl(a, 25);
    c ^= b; c -= rotl(b, 16);
    a ^= c; a -= rotl(c, 4);
    b ^= a; b -= rotl(a, 14);
    c ^= b; c -= rotl(b, 24);
    return c;
}

/* a datum of one attribute of a minimal tuple: NULL, or its bytes */
struct Datum
{
    bool        isnull;
    const uint8_t *p;
    int         attlen;
};

/* attribute `attnum` (1-based) of a MinimalTuple, walked by the deform
 * rules of the colmeta (xpu_common.cu:60-130): fixed-width values aligned
 * by attalign, varlenas aligned unless they begin with a short header,
 * absent trailing attributes NULL */
bool
fetch_attr(const kern_data_store *kds, const uint8_t *mtup, int attnum, Datum *out,
           const char **emsg)
{
    const kern_tupitem *t = (const kern_tupitem *)mtup;
    const uint32_t natts = t->t_infomask2 & HEAP_NATTS_MASK;
    const bool hasnull = (t->t_infomask & HEAP_HASNULL) != 0;
    const uint8_t *bits = mtup + KERN_TUPITEM_T_BITS_OFFSET;
    if (t->t_hoff < MINIMAL_TUPLE_OFFSET + KERN_TUPITEM_T_BITS_OFFSET ||
        t->t_hoff - MINIMAL_TUPLE_OFFSET > t->t_len)
    {
        *emsg = "corrupted minimal tuple header";
        return false;
    }
    const uint8_t *data = mtup + t->t_hoff - MINIMAL_TUPLE_OFFSET;
    const uint8_t *end = mtup + t->t_len;
    uint64_t off = 0;
    for (int j = 0; j < attnum; j++)
    {
        const kern_colmeta *cm = &kds->colmeta[j];
        const bool isnull = (uint32_t)j >= natts || (hasnull && !(bits[j >> 3] & (1 << (j & 7))));
        if (j == attnum - 1)
        {
            out->isnull = isnull;
            out->attlen = cm->attlen;
            out->p = nullptr;
        }
        if (isnull)
            continue;
        if (cm->attlen > 0)
            off = XPU_TYPEALIGN(cm->attalign, off);
        else if (cm->attlen == -1)
        {
            if (data + off >= end)
            {
                *emsg = "minimal tuple is shorter than its attributes";
                return false;
            }
            if (data[off] == 0)   /* a pad byte: the 4-byte header is aligned */
                off = XPU_TYPEALIGN(cm->attalign, off);
        }
        else
        {
            *emsg = "cstring attributes are not supported";
            return false;
        }
        const uint8_t *p = data + off;
        uint64_t sz;
        if (cm->attlen > 0)
            sz = cm->attlen;
        else if (p[0] == 0x01)
            sz = XPU_VARSIZE_EXTERNAL(p[1]);
        else if (p[0] & 0x01)
            sz = (p[0] >> 1) & 0x7f;
        else
            sz = (((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                   ((uint32_t)p[3] << 24)) >> 2);
        if (p + sz > end)
        {
            *emsg = "minimal tuple is shorter than its attributes";
            return false;
        }
        if (j == attnum - 1)
            out->p = p;
        off += sz;
    }
    return true;
}

/* bytes of a plain varlena (after its header); false for a compressed or
 * external datum, whose bytes the backend has to detoast first */
bool
varlena_body(const uint8_t *p, const uint8_t **body, uint32_t *len)
{
    if (p[0] == 0x01)
        return false;
    if (p[0] & 0x01)
    {
        *len = ((p[0] >> 1) & 0x7f) - 1;
        *body = p + 1;
        return true;
    }
    if ((p[0] & 0x03) == 0x02)
        return false;
    *len = (((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
             ((uint32_t)p[3] << 24)) >> 2) - 4;
    *body = p + 4;
    return true;
}

/* NUMERIC_NORMALIZE (xpu_numeric.h:165-207): trailing decimal zeros of
 * the value are dropped while the scale stays non-negative */
void
numeric_normalize(int16_t *scale, __int128 *value)
{
    if (*value == 0)
    {
        *scale = 0;
        return;
    }
    while (*scale > 0 && *value % 10 == 0)
    {
        *value /= 10;
        (*scale)--;
    }
}

/* devtype_numeric_hash (codegen.c:604-628): the value as the device form
 * of __decimal_from_varlena (xpu_numeric.h:210-273), then hash of the
 * 16-bit weight xor hash of the 128-bit value, or the hash of the kind
 * byte for NaN and the infinities */
bool
numeric_hash(const uint8_t *body, uint32_t len, uint32_t *h, const char **emsg)
{
    if (len < 2)
    {
        *emsg = "corrupted numeric header";
        return false;
    }
    uint16_t n_head = (uint16_t)(body[0] | (body[1] << 8));
    if ((n_head & 0xC000) == 0xC000)
    {
        uint8_t kind;
        if (n_head == 0xC000) kind = 1;
        else if (n_head == 0xD000) kind = 2;
        else if (n_head == 0xF000) kind = 3;
        else
        {
            *emsg = "unknown special numeric value";
            return false;
        }
        *h = lookup3(&kind, 1);
        return true;
    }
    const bool is_short = (n_head & 0xC000) == 0x8000;
    int weight;
    bool neg;
    const uint8_t *digits;
    uint32_t ndigits;
    if (is_short)
    {
        weight = n_head & 0x003F;
        if (n_head & 0x0040)
            weight |= ~0x003F;
        neg = (n_head & 0x2000) != 0;
        digits = body + 2;
        ndigits = (len - 2) / 2;
    }
    else
    {
        if (len < 4)
        {
            *emsg = "corrupted numeric header";
            return false;
        }
        weight = (int16_t)(body[2] | (body[3] << 8));
        neg = (n_head & 0xC000) == 0x4000;
        digits = body + 4;
        ndigits = (len - 4) / 2;
    }
    __int128 value = 0;
    for (uint32_t i = 0; i < ndigits; i++)
    {
        if ((value >> 114) != 0)
        {
            *emsg = "numeric value is out of range";
            return false;
        }
        value = value * 10000 + (int16_t)(digits[2 * i] | (digits[2 * i + 1] << 8));
    }
    if (neg)
        value = -value;
    int16_t scale = (int16_t)(4 * ((int)ndigits - (weight + 1)));
    numeric_normalize(&scale, &value);
    uint8_t wb[2], vb[16];
    memcpy(wb, &scale, 2);
    memcpy(vb, &value, 16);
    *h = lookup3(wb, 2) ^ lookup3(vb, 16);
    return true;
}

/* the device type's hash of one non-NULL key (codegen.c:544-660) */
bool
key_hash(uint32_t typoid, const Datum &d, uint32_t *h, const char **emsg)
{
    switch (typoid)
    {
        case OID_BOOL:
        {
            uint8_t b = d.p[0] ? 1 : 0;
            *h = lookup3(&b, 1);
            return true;
        }
        case OID_INT2:
            *h = lookup3(d.p, 2);
            return true;
        case OID_INT4: case OID_FLOAT4: case OID_DATE:
            *h = lookup3(d.p, 4);
            return true;
        case OID_INT8: case OID_FLOAT8: case OID_MONEY: case OID_TIME:
        case OID_TIMESTAMP: case OID_TIMESTAMPTZ:
            *h = lookup3(d.p, 8);
            return true;
        case OID_UUID:
            *h = lookup3(d.p, 16);
            return true;
        case OID_TEXT: case OID_VARCHAR: case OID_BYTEA: case OID_BPCHAR: case OID_NUMERIC:
        {
            const uint8_t *body;
            uint32_t len;
            if (!varlena_body(d.p, &body, &len))
            {
                *emsg = "join key is a compressed or external datum (detoast it first)";
                return false;
            }
            if (typoid == OID_NUMERIC)
                return numeric_hash(body, len, h, emsg);
            if (typoid == OID_BPCHAR)
            {
                /* bpchartruelen: trailing spaces do not count */
                while (len > 0 && body[len - 1] == ' ')
                    len--;
            }
            *h = lookup3(body, len);
            return true;
        }
    }
    *emsg = "join key of a type this build does not hash";
    return false;
}

void
set_error(kern_errorbuf *eb, int line, const char *msg)
{
    xpu_set_error(eb, ERRCODE_DEVICE_ERROR, "xpu_inner.cpp", line, "xpuhip_inner_hash_setup", msg);
}

/* runs fn(lo, hi) over [0, n) on up to nthreads threads */
template <class F>
void
parallel_for(uint64_t n, int nthreads, F fn)
{
    if (nthreads <= 1 || n < 4096)
    {
        fn((uint64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++)
    {
        uint64_t lo = t * per, hi = std::min(n, lo + per);
        if (lo >= hi)
            break;
        th.emplace_back(fn, lo, hi);
    }
    for (auto &x : th)
        x.join();
}

}   /* namespace */

extern "C" int
xpuhip_inner_hash_setup(kern_data_store *kds, const void *const *mtups, uint32_t nitems,
                        const int16_t *key_attnums, uint32_t nkeys, int nthreads,
                        kern_errorbuf *errbuf)
{
    if (!kds || (nitems > 0 && !mtups) || (nkeys > 0 && !key_attnums))
    {
        set_error(errbuf, __LINE__, "null argument");
        return -1;
    }
    if (kds->format != KDS_FORMAT_HASH || kds->hash_nslots == 0)
    {
        set_error(errbuf, __LINE__, "buffer is not a KDS_FORMAT_HASH with hash slots");
        return -1;
    }
    for (uint32_t k = 0; k < nkeys; k++)
    {
        if (key_attnums[k] < 1 || key_attnums[k] > (int)kds->ncols)
        {
            set_error(errbuf, __LINE__, "join key attribute number out of range");
            return -1;
        }
    }
    if (nthreads <= 0)
        nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    const uint32_t base_nitems = kds->nitems;
    const uint64_t base_usage = kds->usage;
    const uint64_t nslots = kds->hash_nslots;

    /* hashes and item sizes */
    std::vector<uint32_t> hashes(nitems);
    std::vector<uint64_t> pos(nitems + 1, 0);
    std::vector<const char *> errs(nthreads > 0 ? (size_t)nthreads : 1, nullptr);
    const uint64_t per = (nitems + nthreads - 1) / std::max(1, nthreads);
    parallel_for(nitems, nthreads, [&](uint64_t lo, uint64_t hi) {
        const char *emsg = nullptr;
        for (uint64_t i = lo; i < hi && !emsg; i++)
        {
            const uint8_t *mt = (const uint8_t *)mtups[i];
            const kern_tupitem *t = (const kern_tupitem *)mt;
            uint32_t h = 0xffffffffU;
            for (uint32_t k = 0; k < nkeys && !emsg; k++)
            {
                Datum d;
                uint32_t x = 0;
                if (!fetch_attr(kds, mt, key_attnums[k], &d, &emsg))
                    break;
                if (!d.isnull &&
                    !key_hash(kds->colmeta[key_attnums[k] - 1].atttypid, d, &x, &emsg))
                    break;
                h = ((h >> 3) | (h << 29)) ^ x;
            }
            hashes[i] = h ^ 0xffffffffU;
            pos[i + 1] = XPU_MAXALIGN(offsetof(kern_hashitem, t) + t->t_len + ROWID_SIZE);
        }
        if (emsg)
            errs[per ? lo / per : 0] = emsg;
    });
    for (const char *e : errs)
    {
        if (e)
        {
            set_error(errbuf, __LINE__, e);
            return -1;
        }
    }
    for (uint32_t i = 0; i < nitems; i++)
        pos[i + 1] += pos[i];
    const uint64_t usage = base_usage + pos[nitems];
    const uint64_t need = KDS_HEAD_LENGTH(kds) + sizeof(uint64_t) * (nslots + base_nitems + nitems) + usage;
    if (need > kds->length)
    {
        char msg[128];
        snprintf(msg, sizeof(msg), "KDS_FORMAT_HASH buffer too small: %llu bytes needed, %llu given",
                 (unsigned long long)need, (unsigned long long)kds->length);
        set_error(errbuf, __LINE__, msg);
        return -1;
    }
    uint8_t *base = (uint8_t *)kds;
    uint8_t *tail = base + kds->length;
    uint64_t *slots = (uint64_t *)(base + KDS_HEAD_LENGTH(kds));
    uint64_t *row_index = slots + nslots;

    /* slot chains: each thread owns a range of slots and threads its rows in
     * input order, so every chain is the serial loop's */
    std::vector<uint64_t> next(nitems);
    parallel_for(nslots, nthreads, [&](uint64_t slo, uint64_t shi) {
        for (uint32_t i = 0; i < nitems; i++)
        {
            const uint64_t s = hashes[i] % nslots;
            if (s < slo || s >= shi)
                continue;
            const uint64_t self = base_usage + pos[i + 1];
            next[i] = slots[s];
            slots[s] = self;
        }
    });
    /* items and row index */
    parallel_for(nitems, nthreads, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; i++)
        {
            const uint8_t *mt = (const uint8_t *)mtups[i];
            const uint32_t t_len = ((const kern_tupitem *)mt)->t_len;
            const uint64_t self = base_usage + pos[i + 1];
            uint8_t *curr = tail - self;
            const uint64_t isz = pos[i + 1] - pos[i];
            memset(curr, 0, isz);
            kern_hashitem *hi_ = (kern_hashitem *)curr;
            hi_->next = next[i];
            memcpy(&hi_->t, mt, t_len);
            hi_->t.t_len = t_len + ROWID_SIZE;
            hi_->t.hash = hashes[i];
            const uint32_t rowid = base_nitems + (uint32_t)i;
            memcpy((uint8_t *)&hi_->t + t_len, &rowid, ROWID_SIZE);
            row_index[base_nitems + i] = (uint64_t)(tail - (uint8_t *)&hi_->t);
        }
    });
    kds->nitems = base_nitems + nitems;
    kds->usage = usage;
    return 0;
}
