/*
 * oracle.c - CPU restatement of PG-Strom's xPU scan / join / pre-aggregation
 * semantics.  TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline);
 * never linked into the product.
 *
 * It is a direct tree-walking interpreter of the kern_expression bytecode,
 * deliberately shaped unlike the GPU path (which compiles the bytecode into
 * flat register programs).  Each piece follows the reference:
 *   expression VM        xpu_common.cu:390-551 (Const/Param/Var/Bool/NullTest),
 *                        :1363-1398 (HashValue), :1400-1425 (SaveExpr),
 *                        :1442-1481 (JoinQuals statuses)
 *   datum operators      xpu_basetype.cu:325-772 (casts, NaN-aware compares,
 *                        overflow-checked arithmetic), xpu_textlib.cu:321-380
 *   hash                 xpu_common.cu:2920-3041 (lookup3, PG seed)
 *   tuple sources        cuda_gpuscan.cu:405-534 (BLOCK pages: LP_NORMAL
 *                        items), :610-700 (ROW chunks: visibility),
 *                        xpu_common.cu:194-257 (heap tuple payload)
 *   column loads         xpu_basetype.cu:242-306 (Arrow fixed width),
 *                        xpu_textlib.cu:237-274 (Arrow utf8/binary),
 *                        xpu_common.cu:60-130 (minimal-tuple deform)
 *   hash join            cuda_gpujoin.cu:304-494 (chain walk, hash compare,
 *                        join quals, left outer null-extension)
 *   projection           xpu_common.cu:1096-1190 (kern_form_minimal_tuple)
 *   group-by             cuda_gpupreagg.cu:18-223 (state init/layout),
 *                        :1196-1605 (state updates), :1607-1731 (lookup)
 *   final write-back     gpu_service.c:2780-2818 (HASH -> ROW)
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_funcs.h"

#define OMAX_SLOTS  256
/* kinds of a numeric datum (XPU_NUMERIC_KIND__*, xpu_numeric.h:36-40) */
#define ON_VALID    0
#define ON_NAN      1
#define ON_PINF     2
#define ON_NINF     3
#define ON_VARLENA  0xff
typedef unsigned __int128 ouint128_t;
static uint32_t varsize_any(const uint8_t *p);
#define OMAX_DEPTH  16
#define OMAX_ATTRS  64

typedef __int128 int128_t;

/* ------------------------------------------------------------------
 * lookup3 hash with PostgreSQL's seed (xpu_common.cu:2920-3041);
 * Bob Jenkins' published mix/final functions.
 * ------------------------------------------------------------------ */
#define ROT(x, k) (((x) << (k)) | ((x) >> (32 - (k))))
#define MIX(a, b, c) do { \
    a -= c; a ^= ROT(c, 4);  c += b; \
    b -= a; b ^= ROT(a, 6);  a += c; \
    c -= b; c ^= ROT(b, 8);  b += a; \
    a -= c; a ^= ROT(c, 16); c += b; \
    b -= a; b ^= ROT(a, 19); a += c; \
    c -= b; c ^= ROT(b, 4);  b += a; } while (0)
#define FINAL(a, b, c) do { \
    c ^= b; c -= ROT(b, 14); \
    a ^= c; a -= ROT(c, 11); \
    b ^= a; b -= ROT(a, 25); \
    c ^= b; c -= ROT(b, 16); \
    a ^= c; a -= ROT(c, 4);  \
    b ^= a; b -= ROT(a, 14); \
    c ^= b; c -= ROT(b, 24); } while (0)

uint32_t
oracle_hash_any(const void *ptr, int sz)
{
    const uint8_t *k = (const uint8_t *)ptr;
    uint32_t a, b, c, len = (uint32_t)sz;

    a = b = c = 0x9e3779b9U + len + 3923095U;
    /* little-endian byte assembly is identical for aligned and unaligned
     * inputs, so one loop covers both paths of the reference */
    while (len >= 12)
    {
        a += (uint32_t)k[0] | ((uint32_t)k[1] << 8) | ((uint32_t)k[2] << 16) | ((uint32_t)k[3] << 24);
        b += (uint32_t)k[4] | ((uint32_t)k[5] << 8) | ((uint32_t)k[6] << 16) | ((uint32_t)k[7] << 24);
        c += (uint32_t)k[8] | ((uint32_t)k[9] << 8) | ((uint32_t)k[10] << 16) | ((uint32_t)k[11] << 24);
        MIX(a, b, c);
        k += 12;
        len -= 12;
    }
    switch (len)
    {
        case 11: c += (uint32_t)k[10] << 24;    /* fall through */
        case 10: c += (uint32_t)k[9] << 16;     /* fall through */
        case 9:  c += (uint32_t)k[8] << 8;      /* fall through */
        case 8:  b += (uint32_t)k[7] << 24;     /* fall through */
        case 7:  b += (uint32_t)k[6] << 16;     /* fall through */
        case 6:  b += (uint32_t)k[5] << 8;      /* fall through */
        case 5:  b += k[4];                     /* fall through */
        case 4:  a += (uint32_t)k[3] << 24;     /* fall through */
        case 3:  a += (uint32_t)k[2] << 16;     /* fall through */
        case 2:  a += (uint32_t)k[1] << 8;      /* fall through */
        case 1:  a += k[0];
    }
    FINAL(a, b, c);
    return c;
}

/* ------------------------------------------------------------------
 * function table
 * ------------------------------------------------------------------ */
enum { K_cmp = 1, K_arith, K_unary, K_bitop, K_cast, K_textcmp, K_textlen, K_like, K_substr,
       K_numcmp, K_numarith, K_numunary, K_numcast };
enum { O_eq = 1, O_ne, O_lt, O_le, O_gt, O_ge, O_pl, O_mi, O_mul, O_div, O_mod,
       O_up, O_um, O_abs, O_not, O_and, O_or, O_xor, O_shl, O_shr, O_cast, O_len,
       O_like, O_nlike, O_ilike, O_nilike, O_substr };

typedef struct { int opcode; const char *name; int kind, op, nargs, t0, t1, t2, res; } ofunc;

#define OMAX_FB_ATTRS   256
#define T_Invalid   TypeOpCode__Invalid
#define T_bool      TypeOpCode__bool
#define T_int1      TypeOpCode__int1
#define T_int2      TypeOpCode__int2
#define T_int4      TypeOpCode__int4
#define T_int8      TypeOpCode__int8
#define T_float4    TypeOpCode__float4
#define T_float8    TypeOpCode__float8
#define T_text      TypeOpCode__text
#define T_bpchar    TypeOpCode__bpchar
#define T_numeric   TypeOpCode__numeric
#define OFUNC(OPC, NAME, KIND, OP, NARGS, A0, A1, A2, RES) \
    { OPC, #NAME, K_##KIND, O_##OP, NARGS, T_##A0, T_##A1, T_##A2, T_##RES },
static const ofunc ofunc_table[] = { OFUNC_TABLE(OFUNC) { 0, NULL, 0, 0, 0, 0, 0, 0, 0 } };

static const ofunc *
lookup_func(int opcode)
{
    for (const ofunc *f = ofunc_table; f->name; f++)
        if (f->opcode == opcode)
            return f;
    return NULL;
}

static bool is_int_type(int t) { return t == T_int1 || t == T_int2 || t == T_int4 || t == T_int8 || t == TypeOpCode__date; }
static bool is_flt_type(int t) { return t == T_float4 || t == T_float8; }
static int  type_width(int t)
{
    switch (t) {
        case T_bool: case T_int1: return 1;
        case T_int2: return 2;
        case T_int4: case T_float4: case TypeOpCode__date: return 4;
        case T_int8: case T_float8: return 8;
        default: return -1;
    }
}
static void int_range(int t, int64_t *lo, int64_t *hi)
{
    switch (t) {
        case T_int1: *lo = -128; *hi = 127; break;
        case T_int2: *lo = -32768; *hi = 32767; break;
        case T_int4: case TypeOpCode__date: *lo = INT32_MIN; *hi = INT32_MAX; break;
        default: *lo = INT64_MIN; *hi = INT64_MAX; break;
    }
}

/* ------------------------------------------------------------------
 * datum and context
 * ------------------------------------------------------------------ */
typedef struct
{
    bool        isnull;
    int64_t     i;          /* bool / ints / date */
    double      d;          /* float8 */
    float       f;          /* float4 */
    const char *p;          /* text payload (no header); numeric: its varlena */
    int32_t     len;        /* numeric varlena: its size, -1 when raw */
    /* numeric (xpu_numeric_t): kind, scale ("weight") and value; kind
     * ON_VARLENA until an operator parses the varlena at p */
    uint8_t     nk;
    int16_t     nw;
    int128_t    nv;
} od;

static int on_encode(const od *v, char *buf);
static uint32_t on_hash(const od *v);

typedef struct
{
    const kern_session_info *session;
    const kern_varslot_desc *vdesc;
    int         nslots;
    od          slots[OMAX_SLOTS];
    bool        compare_nulls;
    int         enc;        /* database encoding: ENC_NONE / ENC_SINGLE / ENC_UTF8 / ENC_OTHER */
    uint32_t    errcode;
    char        errmsg[KERN_ERRORBUF_MESSAGE_LEN + 1];
    bool        fb_stage;   /* the row's depth-0 work runs (a fallback error
                               then sends it to the CPU fallback) */
} octx;

enum { ENC_NONE = 0, ENC_SINGLE, ENC_UTF8, ENC_OTHER };

/* the session's database encoding by name (xpu_encode_catalog,
 * xpu_textlib.cu:853-900): single-byte encodings, UTF8, or another
 * multi-byte one */
static int
encoding_kind(const char *name)
{
    static const char *mb[] = { "EUC_JP", "EUC_CN", "EUC_KR", "EUC_TW", "EUC_JIS_2004",
                                "MULE_INTERNAL", "SJIS", "BIG5", "GBK", "UHC", "GB18030",
                                "JOHAB", "SHIFT_JIS_2004", NULL };
    if (!name || !*name)
        return ENC_NONE;
    if (strcmp(name, "UTF8") == 0)
        return ENC_UTF8;
    for (int i = 0; mb[i]; i++)
        if (strcmp(name, mb[i]) == 0)
            return ENC_OTHER;
    return ENC_SINGLE;
}

static int
session_encoding(const kern_session_info *s)
{
    if (!s || !s->session_encode)
        return ENC_NONE;
    char name[17];
    memcpy(name, (const char *)s + s->session_encode, 16);
    name[16] = 0;
    return encoding_kind(name);
}

/* bytes of the character at s (pg_utf8_mblen, xpu_textlib.cu:774-787) */
static int
utf8_mblen(const char *s)
{
    unsigned char c = (unsigned char)*s;
    if ((c & 0x80) == 0) return 1;
    if ((c & 0xe0) == 0xc0) return 2;
    if ((c & 0xf0) == 0xe0) return 3;
    if ((c & 0xf8) == 0xf0) return 4;
    return 1;
}

static int
mblen_of(int enc, const char *s)
{
    return enc == ENC_UTF8 ? utf8_mblen(s) : 1;
}

/* length without the trailing spaces of a bpchar (bpchar_truelen,
 * xpu_textlib.cu:59-69) */
static int
bpchar_truelen(const char *p, int len)
{
    while (len > 0 && p[len - 1] == ' ')
        len--;
    return len;
}

/* LIKE / ILIKE (GenericMatchText, xpu_textlib.cu:915-1085) without
 * recursion: the newest % is the only point a failed match resumes from,
 * since the reference's recursive search returns "abort" rather than
 * "false" once a deeper % has scanned the rest of the text.  Returns 1 /
 * 0, or -1 for an invalid escape. */
#define LIKE_UPPER(c, ic) (((ic) && (c) >= 'a' && (c) <= 'z') ? (c) + ('A' - 'a') : (c))
static int
like_match(int enc, bool ic, const char *t, int tlen, const char *p, int plen)
{
    int ti = 0, pi = 0;
    int bt_t = -1, bt_p = -1;
    char firstpat = 0;

    if (plen == 1 && p[0] == '%')
        return 1;
    for (;;)
    {
        bool mismatch = false;
        while (ti < tlen && pi < plen && !mismatch)
        {
            char c = p[pi];
            if (c == '\\')
            {
                pi++;
                if (pi >= plen)
                    return -1;
                if (LIKE_UPPER(p[pi], ic) != LIKE_UPPER(t[ti], ic))
                    mismatch = true;
                else
                {
                    ti++;
                    pi++;
                }
            }
            else if (c == '%')
            {
                pi++;
                while (pi < plen)
                {
                    if (p[pi] == '%')
                        pi++;
                    else if (p[pi] == '_')
                    {
                        if (ti >= tlen)
                            return 0;
                        ti += mblen_of(enc, t + ti);
                        pi++;
                    }
                    else
                        break;
                }
                if (pi >= plen)
                    return 1;
                if (p[pi] == '\\')
                {
                    if (plen - pi < 2)
                        return -1;
                    firstpat = LIKE_UPPER(p[pi + 1], ic);
                }
                else
                    firstpat = LIKE_UPPER(p[pi], ic);
                bt_p = pi;
                bt_t = ti;
                goto scan;
            }
            else if (c == '_')
            {
                ti += mblen_of(enc, t + ti);
                pi++;
            }
            else if (LIKE_UPPER(c, ic) != LIKE_UPPER(t[ti], ic))
                mismatch = true;
            else
            {
                ti++;
                pi++;
            }
        }
        if (!mismatch)
        {
            if (ti < tlen)
                mismatch = true;        /* end of pattern, not of text */
            else
            {
                while (pi < plen && p[pi] == '%')
                    pi++;
                return pi >= plen ? 1 : 0;
            }
        }
        /* a mismatch: resume the newest %'s scan one character on */
        if (bt_p < 0)
            return 0;
        bt_t += mblen_of(enc, t + bt_t);
    scan:
        while (bt_t < tlen && LIKE_UPPER(t[bt_t], ic) != firstpat)
            bt_t += mblen_of(enc, t + bt_t);
        if (bt_t >= tlen)
            return 0;
        ti = bt_t;
        pi = bt_p;
    }
}

static bool
oerror(octx *c, uint32_t code, const char *fmt, const char *arg)
{
    if (c->errcode == 0 || (ERRCODE_IS_SUSPEND(c->errcode) && !ERRCODE_IS_SUSPEND(code)))
    {
        c->errcode = code;
        snprintf(c->errmsg, sizeof(c->errmsg), fmt, arg);
    }
    return false;
}

#define KEXP_FIRST_ARG(k) \
    (((k)->nr_args > 0 && (k)->args_offset > 0) ? \
     (const kern_expression *)((const char *)(k) + (k)->args_offset) : NULL)
#define KEXP_NEXT_ARG(a) \
    ((const kern_expression *)((const char *)(a) + XPU_MAXALIGN((a)->len)))

static bool eval(octx *c, const kern_expression *k, od *r);

/* datum from a fixed-width / varlena image in the PostgreSQL format */
static bool
datum_from_image(octx *c, int type, const char *addr, od *r)
{
    memset(r, 0, sizeof(od));
    switch (type)
    {
        case T_bool: r->i = *(const uint8_t *)addr ? 1 : 0; break;
        case T_int1: r->i = *(const int8_t *)addr; break;
        case T_int2: { int16_t v; memcpy(&v, addr, 2); r->i = v; } break;
        case T_int4: case TypeOpCode__date: { int32_t v; memcpy(&v, addr, 4); r->i = v; } break;
        case T_int8: { int64_t v; memcpy(&v, addr, 8); r->i = v; } break;
        case T_float4: memcpy(&r->f, addr, 4); break;
        case T_float8: memcpy(&r->d, addr, 8); break;
        case TypeOpCode__numeric:
        {
            /* xpu_numeric_datum_heap_read (xpu_numeric.cu:32-44): the
             * varlena as it is, parsed by the first operator that needs
             * its value */
            uint8_t b0 = (uint8_t)addr[0];
            r->nk = ON_VARLENA;
            r->p = addr;
            r->len = (b0 == 0x01 || (b0 & 0x03) == 0x02) ? (int32_t)-1
                                                         : (int32_t)varsize_any((const uint8_t *)addr);
            break;
        }
        case T_text: case TypeOpCode__bytea: case TypeOpCode__bpchar:
        {
            uint8_t b0 = (uint8_t)addr[0];
            if (b0 == 0x01 || (b0 & 0x03) == 0x02)
            {
                /* compressed or external: kept as it is, length -1
                 * (xpu_text_datum_heap_read); its bytes are never read */
                r->len = -1;
                r->p = addr;
                break;
            }
            if (b0 & 0x01)
            {
                r->len = ((b0 >> 1) & 0x7f) - 1;
                r->p = addr + 1;
            }
            else
            {
                uint32_t h; memcpy(&h, addr, 4);
                r->len = (int32_t)((h >> 2) & 0x3fffffff) - 4;
                r->p = addr + 4;
            }
            break;
        }
        default:
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: unsupported type", "oracle");
    }
    return true;
}

/* size and bytes of a datum written into a tuple (xpu_*_datum_write); a
 * compressed or external varlena is copied as it is (xpu_textlib.cu:319-325) */
static int
datum_write(int type, const od *v, char *buf)
{
    switch (type)
    {
        case T_bool: if (buf) *(uint8_t *)buf = (uint8_t)(v->i != 0); return 1;
        case T_int1: if (buf) *(int8_t *)buf = (int8_t)v->i; return 1;
        case T_int2: if (buf) { int16_t x = (int16_t)v->i; memcpy(buf, &x, 2); } return 2;
        case T_int4: case TypeOpCode__date:
            if (buf) { int32_t x = (int32_t)v->i; memcpy(buf, &x, 4); } return 4;
        case T_int8: if (buf) memcpy(buf, &v->i, 8); return 8;
        case T_float4: if (buf) memcpy(buf, &v->f, 4); return 4;
        case T_float8: if (buf) memcpy(buf, &v->d, 8); return 8;
        case TypeOpCode__numeric:
            /* xpu_numeric_datum_write (xpu_numeric.cu:163-178): an unparsed
             * datum is copied as it is */
            if (v->nk == ON_VARLENA)
            {
                uint32_t sz = varsize_any((const uint8_t *)v->p);
                if (buf)
                    memcpy(buf, v->p, sz);
                return (int)sz;
            }
            return on_encode(v, buf);
        case T_text: case TypeOpCode__bytea: case TypeOpCode__bpchar:
            if (v->len < 0)
            {
                uint32_t sz = varsize_any((const uint8_t *)v->p);
                if (buf)
                    memcpy(buf, v->p, sz);
                return (int)sz;
            }
            if (buf)
            {
                uint32_t h = (uint32_t)(v->len + 4) << 2;
                memcpy(buf, &h, 4);
                memcpy(buf + 4, v->p, v->len);
            }
            return 4 + v->len;
    }
    return -1;
}

static uint32_t
datum_hash(int type, const od *v)
{
    switch (type)
    {
        case T_bool: { uint8_t b = v->i != 0; return oracle_hash_any(&b, 1); }
        case T_int1: { int8_t x = (int8_t)v->i; return oracle_hash_any(&x, 1); }
        case T_int2: { int16_t x = (int16_t)v->i; return oracle_hash_any(&x, 2); }
        case T_int4: case TypeOpCode__date: { int32_t x = (int32_t)v->i; return oracle_hash_any(&x, 4); }
        case T_int8: return oracle_hash_any(&v->i, 8);
        case T_float4: return oracle_hash_any(&v->f, 4);
        case T_float8: return oracle_hash_any(&v->d, 8);
        case T_text: case TypeOpCode__bytea: case TypeOpCode__bpchar:
            return oracle_hash_any(v->p, v->len);
        case TypeOpCode__numeric:
            return on_hash(v);      /* parsed by the caller */
    }
    return 0;
}

/* ------------------------------------------------------------------
 * operators
 * ------------------------------------------------------------------ */
static double as_double(int t, const od *v)
{
    if (t == T_float8) return v->d;
    if (t == T_float4) return (double)v->f;
    return (double)v->i;
}

static bool
float_cmp(int op, double x, double y)
{
    /* NaN sorts above every other value (xpu_basetype.cu:467-478) */
    switch (op)
    {
        case O_eq: return isnan(x) ? isnan(y) : (!isnan(y) && x == y);
        case O_ne: return isnan(x) ? !isnan(y) : (isnan(y) || x != y);
        case O_lt: return !isnan(x) && (isnan(y) || x < y);
        case O_le: return isnan(y) || (!isnan(x) && x <= y);
        case O_gt: return !isnan(y) && (isnan(x) || x > y);
        case O_ge: return isnan(x) || (!isnan(y) && x >= y);
    }
    return false;
}

static bool
cmp_result(int op, int comp)
{
    switch (op)
    {
        case O_eq: return comp == 0;
        case O_ne: return comp != 0;
        case O_lt: return comp < 0;
        case O_le: return comp <= 0;
        case O_gt: return comp > 0;
        case O_ge: return comp >= 0;
    }
    return false;
}

static bool is_text_type(int t)
{
    return t == T_text || t == TypeOpCode__bytea || t == T_bpchar;
}

/* A compressed or external text/bpchar/bytea needs its bytes: the CPU
 * fallback takes it (xpu_text_is_valid, xpu_textlib.h:40-70).  True, with
 * the error set, for such a datum. */
static bool
raw_text(octx *c, int type, const od *v)
{
    if (!is_text_type(type) || v->isnull || v->len >= 0)
        return false;
    oerror(c, ERRCODE_SUSPEND_FALLBACK, "%s", "text datum is compressed or external");
    return true;
}


/* ------------------------------------------------------------------
 * numeric (xpu_numeric.h, xpu_numeric.cu), restated over the datum's
 * 128-bit integer nv and decimal scale nw (the value is nv * 10^-nw, the
 * reference's "weight"), with the reference's int128 wraparound
 * ------------------------------------------------------------------ */
static int128_t on_x10(int128_t v) { return (int128_t)((ouint128_t)v * 10); }

/* NUMERIC_NORMALIZE (xpu_numeric.h:165-207): as many trailing zeros leave
 * the value as the scale has positive digits; zero gets scale 0 */
static void
on_trim(od *v)
{
    if (v->nv == 0)
    {
        v->nw = 0;
        return;
    }
    while (v->nw > 0 && v->nv % 10 == 0)
    {
        v->nv /= 10;
        v->nw--;
    }
}

/* the datum's value, parsed from its varlena where needed
 * (__decimal_from_varlena, xpu_numeric.h:210-273): NumericShort or
 * NumericLong, base-10000 digits with the leading digit at `weight`,
 * overflow when 14 high bits are in use before a digit is added */
static bool
on_parse(octx *c, od *v)
{
    if (v->nk != ON_VARLENA)
        return true;
    if (v->len < 0)
        return oerror(c, ERRCODE_SUSPEND_FALLBACK, "%s", "text datum is compressed or external");
    const uint8_t *h = (const uint8_t *)v->p;
    const uint8_t *body;
    uint32_t blen;
    if (h[0] & 1)
    {
        blen = ((h[0] >> 1) & 0x7f) - 1;
        body = h + 1;
    }
    else
    {
        uint32_t w;
        memcpy(&w, h, 4);
        blen = ((w >> 2) & 0x3fffffff) - 4;
        body = h + 4;
    }
    if (blen < 2)
        return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "corrupted numeric header");
    uint16_t head = (uint16_t)(body[0] | (body[1] << 8));
    v->nv = 0;
    v->nw = 0;
    switch (head & 0xC000)
    {
        case 0xC000:
            /* NaN, +Inf, -Inf (NUMERIC_NAN / PINF / NINF) */
            if (head == 0xC000) v->nk = ON_NAN;
            else if (head == 0xD000) v->nk = ON_PINF;
            else if (head == 0xF000) v->nk = ON_NINF;
            else return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "unknown special numeric value");
            return true;
        default:
            break;
    }
    int digit_at, weight;
    bool minus;
    if ((head & 0xC000) == 0x8000)
    {
        /* NumericShort: sign, display scale and a 7-bit weight in n_header */
        digit_at = 2;
        weight = (head & 0x0040) ? (int)(head & 0x003F) - 64 : (int)(head & 0x003F);
        minus = (head & 0x2000) != 0;
    }
    else
    {
        if (blen < 4)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "corrupted numeric header");
        digit_at = 4;
        weight = (int16_t)(body[2] | (body[3] << 8));
        minus = (head & 0xC000) == 0x4000;
    }
    int ndigits = (int)(blen - digit_at) / 2;
    int128_t acc = 0;
    for (int i = 0; i < ndigits; i++)
    {
        int16_t d = (int16_t)(body[digit_at + 2 * i] | (body[digit_at + 2 * i + 1] << 8));
        if ((acc >> 114) != 0)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "numeric value is out of range");
        acc = acc * 10000 + d;
    }
    v->nk = ON_VALID;
    v->nv = minus ? -acc : acc;
    /* digits after the decimal point: those right of digit `weight` */
    v->nw = (int16_t)(4 * (ndigits - (int16_t)(weight + 1)));
    on_trim(v);
    return true;
}

/* a and b at the larger of their scales */
static void
on_same_scale(od *a, od *b)
{
    for (; a->nw > b->nw; b->nw++)
        b->nv = on_x10(b->nv);
    for (; b->nw > a->nw; a->nw++)
        a->nv = on_x10(a->nv);
}

/* __numeric_compare (xpu_numeric.cu:478-539): NaN above +Inf above every
 * finite value above -Inf; finite values by sign first */
static int
on_cmp(od a, od b)
{
    if (a.nk == ON_NAN || b.nk == ON_NAN)
        return (a.nk == ON_NAN) - (b.nk == ON_NAN);
    if (a.nk != ON_VALID || b.nk != ON_VALID)
    {
        int ra = a.nk == ON_PINF ? 1 : a.nk == ON_NINF ? -1 : 0;
        int rb = b.nk == ON_PINF ? 1 : b.nk == ON_NINF ? -1 : 0;
        return ra > rb ? 1 : ra < rb ? -1 : 0;
    }
    int sa = a.nv > 0 ? 1 : a.nv < 0 ? -1 : 0;
    int sb = b.nv > 0 ? 1 : b.nv < 0 ? -1 : 0;
    if (sa != sb)
        return sa > sb ? 1 : -1;
    on_same_scale(&a, &b);
    return a.nv > b.nv ? 1 : a.nv < b.nv ? -1 : 0;
}

static int on_sign(const od *v) { return v->nk == ON_PINF ? 1 : v->nk == ON_NINF ? -1 : v->nk == ON_VALID ? (v->nv > 0) - (v->nv < 0) : 0; }

/* __decimal_add / _sub / _mul (xpu_numeric.h:443-605): the finite rule
 * applies as soon as ONE operand is finite (a NaN or infinity then counts
 * as zero); otherwise IEEE-like rules over NaN and the infinities */
static bool
on_addsub(bool sub, od a, od b, od *r)
{
    r->nk = ON_VALID;
    if (a.nk == ON_VALID || b.nk == ON_VALID)
    {
        on_same_scale(&a, &b);
        r->nw = a.nw;
        r->nv = sub ? (int128_t)((ouint128_t)a.nv - (ouint128_t)b.nv)
                    : (int128_t)((ouint128_t)a.nv + (ouint128_t)b.nv);
        on_trim(r);
        return true;
    }
    int ia = on_sign(&a), ib = sub ? -on_sign(&b) : on_sign(&b);
    if (a.nk == ON_NAN || b.nk == ON_NAN || ia + ib == 0)
        r->nk = ON_NAN;
    else
        r->nk = (ia + ib > 0) ? ON_PINF : ON_NINF;
    return true;
}

static bool
on_mul(od a, od b, od *r)
{
    r->nk = ON_VALID;
    if (a.nk == ON_VALID || b.nk == ON_VALID)
    {
        r->nw = (int16_t)(a.nw + b.nw);
        r->nv = (int128_t)((ouint128_t)a.nv * (ouint128_t)b.nv);
        on_trim(r);
        return true;
    }
    int prod = on_sign(&a) * on_sign(&b);
    r->nk = (a.nk == ON_NAN || b.nk == ON_NAN || prod == 0) ? ON_NAN
          : prod > 0 ? ON_PINF : ON_NINF;
    return true;
}

/* __decimal_div (xpu_numeric.h:607-700): long division of the magnitudes,
 * one decimal digit per step, until the remainder is zero or the quotient
 * reaches 2^50 */
static bool
on_div(octx *c, od a, od b, od *r)
{
    r->nk = ON_VALID;
    if (a.nk == ON_VALID && b.nk == ON_VALID)
    {
        if (b.nv == 0)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "division by zero");
        bool minus = (a.nv < 0) != (b.nv < 0);
        int128_t num = a.nv < 0 ? -a.nv : a.nv;
        int128_t den = b.nv < 0 ? -b.nv : b.nv;
        int128_t q = 0;
        int16_t w = (int16_t)(a.nw - b.nw);
        for (;;)
        {
            int128_t digit = num / den;
            q = (int128_t)((ouint128_t)q * 10 + (ouint128_t)digit);
            num -= digit * den;
            if (num == 0 || (q >> 50) != 0)
                break;
            num = on_x10(num);
            w++;
        }
        r->nv = minus ? -q : q;
        r->nw = w;
        on_trim(r);
        return true;
    }
    if (a.nk == ON_NAN || b.nk == ON_NAN)
        r->nk = ON_NAN;
    else if (a.nk == ON_VALID)
    {
        r->nv = 0;          /* finite / infinite */
        r->nw = 0;
    }
    else
    {
        int sb = on_sign(&b);
        if (b.nk == ON_VALID && sb == 0)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "division by zero");
        int q = on_sign(&a) * sb;
        r->nk = q > 0 ? ON_PINF : ON_NINF;
    }
    return true;
}

/* __xpu_numeric_to_int64 (xpu_numeric.cu:250-295): rounded half away from
 * zero at the first fractional digit, then range-checked */
static bool
on_to_int(octx *c, const od *v, int64_t lo, int64_t hi, int64_t *out)
{
    if (v->nk != ON_VALID)
        return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "cannot convert NaN/Inf to integer");
    int128_t x = v->nv;
    int w = v->nw;
    if (x != 0)
    {
        for (; w > 1; w--)
            x /= 10;
        if (w == 1)
            x = (x + (x > 0 ? 5 : -5)) / 10;
        for (; w < 0; w++)
        {
            x = on_x10(x);
            if (x < lo || x > hi)
                break;
        }
        if (x < lo || x > hi)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "integer out of range");
    }
    *out = (int64_t)x;
    return true;
}

/* __xpu_numeric_to_fp64 (xpu_numeric.cu:341-380) */
static bool
on_to_f64(octx *c, const od *v, double *out)
{
    switch (v->nk)
    {
        case ON_PINF: *out = INFINITY; return true;
        case ON_NINF: *out = -INFINITY; return true;
        case ON_NAN: *out = NAN; return true;
    }
    double f = (double)v->nv;
    if (f != 0.0)
    {
        for (int w = v->nw; w > 0; w--)
            f /= 10.0;
        for (int w = v->nw; w < 0; w++)
            f *= 10.0;
        if (isnan(f) || isinf(f))
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "float out of range");
    }
    *out = f;
    return true;
}

/* __decimal_from_float8 (xpu_numeric.h:778-846): the 53-bit significand is
 * brought to binary exponent 0 by halving / doubling, trading factors of
 * 10 into the decimal scale to keep it in range, then cut to 15 digits.
 * A float with all exponent bits set (NaN and the infinities) is NaN, and
 * zero and subnormals go through the same steps with the implicit bit. */
static void
on_from_f64(double d, od *r)
{
    uint64_t bits;
    memcpy(&bits, &d, 8);
    int e = (int)((bits >> 52) & 0x7ff);
    memset(r, 0, sizeof(od));
    if (e == 0x7ff)
    {
        r->nk = ON_NAN;
        return;
    }
    uint64_t m = (bits & 0xfffffffffffffULL) | (1ULL << 52);
    int scale = 0;
    const uint64_t limit = 0x1800000000000000ULL;
    while (e > 1075)
    {
        if (m <= limit) { m <<= 1; e--; }
        else { m /= 10; scale--; }
    }
    while (e < 1075)
    {
        if (m >= limit) { m >>= 1; e++; }
        else { m *= 10; scale++; }
    }
    while (m >= 1000000000000000ULL)
    {
        m /= 10;
        scale--;
    }
    r->nk = ON_VALID;
    r->nw = (int16_t)scale;
    r->nv = (bits >> 63) ? -(int128_t)m : (int128_t)m;
}

/* __decimal_to_varlena (xpu_numeric.h:275-372), restated on the decimal
 * digits: the magnitude's digits, padded with zeros so that the scale is a
 * multiple of 4 (which becomes the display scale), grouped in fours from
 * the right into NumericLong base-10000 digits, then a zero digit for
 * every 4 places of a negative scale.  NaN and the infinities take the
 * 6-byte PostgreSQL encodings. */
static int
on_encode(const od *v, char *buf)
{
    if (v->nk != ON_VALID)
    {
        if (buf)
        {
            uint32_t vl = 6 << 2;
            uint16_t h = v->nk == ON_PINF ? 0xD000 : v->nk == ON_NINF ? 0xF000 : 0xC000;
            memcpy(buf, &vl, 4);
            memcpy(buf + 4, &h, 2);
        }
        return 6;
    }
    od n = *v;
    bool minus = n.nv < 0;
    if (minus)
        n.nv = -n.nv;
    on_trim(&n);
    char dig[64];
    int nd = 0, w = n.nw, w4 = 0;
    if (n.nv != 0)
    {
        char rev[48];
        int k = 0;
        for (ouint128_t x = (ouint128_t)n.nv; x != 0; x /= 10)
            rev[k++] = (char)('0' + (int)(x % 10));
        /* smallest multiple of 4 not below the scale */
        w4 = w >= 0 ? (w + 3) / 4 * 4 : -((-w) / 4 * 4);
        int pad_right = w4 - w;
        int len = k + pad_right;
        int pad_left = (4 - len % 4) % 4;
        for (int i = 0; i < pad_left; i++)
            dig[nd++] = '0';
        while (k > 0)
            dig[nd++] = rev[--k];
        for (int i = 0; i < pad_right; i++)
            dig[nd++] = '0';
    }
    int ngroups = nd / 4;
    int nzero = w4 < 0 ? -w4 / 4 : 0;
    int len = 8 + 2 * (ngroups + nzero);
    if (buf)
    {
        uint32_t vl = (uint32_t)len << 2;
        uint16_t sign_dscale = (uint16_t)((w4 > 0 ? w4 : 0) | (minus ? 0x4000 : 0));
        int16_t weight = (int16_t)(ngroups - w4 / 4 - 1);
        memcpy(buf, &vl, 4);
        memcpy(buf + 4, &sign_dscale, 2);
        memcpy(buf + 6, &weight, 2);
        for (int g = 0; g < ngroups; g++)
        {
            uint16_t x = (uint16_t)((dig[4 * g] - '0') * 1000 + (dig[4 * g + 1] - '0') * 100 +
                                    (dig[4 * g + 2] - '0') * 10 + (dig[4 * g + 3] - '0'));
            memcpy(buf + 8 + 2 * g, &x, 2);
        }
        memset(buf + 8 + 2 * ngroups, 0, 2 * nzero);
    }
    return len;
}

/* xpu_numeric_datum_hash (xpu_numeric.cu:180-196) of a parsed datum */
static uint32_t
on_hash(const od *v)
{
    if (v->nk != ON_VALID)
    {
        uint8_t k = v->nk;
        return oracle_hash_any(&k, 1);
    }
    return oracle_hash_any(&v->nw, 2) ^ oracle_hash_any(&v->nv, 16);
}

/* __normalize_numeric_int128 (xpu_numeric.cu:1080-1126): the value at
 * scale `to`, digits cut (toward zero) when the scale shrinks */
static int128_t
on_rescale(int128_t v, int from, int to)
{
    for (; from < to; from++)
        v = on_x10(v);
    for (; from > to; from--)
        v /= 10;
    return v;
}

/* __numeric_typmod_weight (xpu_numeric.h:416-430) */
static int
on_typmod_scale(int32_t typmod)
{
    if (typmod < 0)
        return 15;
    int w = (((typmod - 4) & 0x7ff) ^ 1024) - 1024;
    return w < 0 ? 0 : w > 15 ? 15 : w;
}

static bool
eval_func(octx *c, const kern_expression *k, const ofunc *f, od *r)
{
    od a[3];
    const kern_expression *karg = KEXP_FIRST_ARG(k);

    memset(r, 0, sizeof(od));
    if ((int)k->nr_args != f->nargs)
        return oerror(c, ERRCODE_DEVICE_ERROR, "%s: wrong number of arguments", f->name);
    for (int i = 0; i < f->nargs; i++, karg = KEXP_NEXT_ARG(karg))
        if (!eval(c, karg, &a[i]))
            return false;
    if (a[0].isnull || (f->nargs > 1 && a[1].isnull) || (f->nargs > 2 && a[2].isnull))
    {
        /* NULL input: comparisons honour kmode_compare_nulls for eq/ne
         * (xpu_basetype.h:30-60); everything else yields NULL */
        if ((f->kind == K_cmp || f->kind == K_textcmp || f->kind == K_numcmp) && c->compare_nulls &&
            (f->op == O_eq || f->op == O_ne))
        {
            bool both = a[0].isnull && a[1].isnull;
            r->i = (f->op == O_eq) ? both : !both;
            return true;
        }
        r->isnull = true;
        return true;
    }
    if (f->kind == K_textlen && f->t0 == T_text && a[0].len < 0 && c->enc == ENC_SINGLE)
    {
        /* pgfn_textlen under a single-byte encoding counts bytes, which
         * the header of a compressed or external datum tells
         * (xpu_textlib.cu:676-703): the external pointer's va_extinfo, or
         * for a compressed datum the word that the reference's memcpy from
         * VARDATA puts into toast_compress_header's tcinfo (the one after
         * the compression info) */
        const uint8_t *h = (const uint8_t *)a[0].p;
        uint32_t w;
        memcpy(&w, h + (h[0] == 0x01 ? 6 : 8), 4);
        r->i = (int64_t)(w & 0x3fffffffU);
        return true;
    }
    {
        const int ts[3] = { f->t0, f->t1, f->t2 };
        for (int i = 0; i < f->nargs && i < 3; i++)
            if (raw_text(c, ts[i], &a[i]))
                return false;
    }
    switch (f->kind)
    {
        case K_cmp:
            if (is_flt_type(f->t0) || is_flt_type(f->t1))
            {
                if (f->t0 == T_float4 && f->t1 == T_float4)
                    r->i = float_cmp(f->op, (double)a[0].f, (double)a[1].f);
                else
                    r->i = float_cmp(f->op, as_double(f->t0, &a[0]), as_double(f->t1, &a[1]));
            }
            else
                r->i = cmp_result(f->op, a[0].i < a[1].i ? -1 : a[0].i > a[1].i ? 1 : 0);
            return true;
        case K_textcmp:
        {
            if (f->t0 == T_bpchar)
            {
                /* bpchar compares without trailing spaces (xpu_textlib.cu:59-69) */
                a[0].len = bpchar_truelen(a[0].p, a[0].len);
                a[1].len = bpchar_truelen(a[1].p, a[1].len);
            }
            int n = a[0].len < a[1].len ? a[0].len : a[1].len;
            int comp = memcmp(a[0].p, a[1].p, n);
            if (comp == 0)
                comp = a[0].len < a[1].len ? -1 : a[0].len > a[1].len ? 1 : 0;
            r->i = cmp_result(f->op, comp);
            return true;
        }
        case K_textlen:
        {
            /* pgfn_textlen / pgfn_bpcharlen (xpu_textlib.cu:583-720):
             * characters by the database encoding, up to a NUL byte */
            int len = f->t0 == T_bpchar ? bpchar_truelen(a[0].p, a[0].len) : a[0].len;
            if (c->enc == ENC_NONE)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "No encoding info was supplied");
            if (c->enc == ENC_OTHER)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "encoding not supported");
            if (c->enc == ENC_UTF8)
            {
                const char *s = a[0].p;
                int n = 0;
                while (len > 0 && *s)
                {
                    int w = utf8_mblen(s);
                    len -= w;
                    s += w;
                    n++;
                }
                len = n;
            }
            r->i = len;
            return true;
        }
        case K_like:
        {
            /* PG_TEXTLIKE_TEMPLATE (xpu_textlib.cu:1087-1155) */
            int tlen = f->t0 == T_bpchar ? bpchar_truelen(a[0].p, a[0].len) : a[0].len;
            if (c->enc == ENC_NONE || c->enc == ENC_OTHER)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", c->enc == ENC_NONE
                              ? "No encoding info was supplied" : "encoding not supported");
            int m = like_match(c->enc, f->op == O_ilike || f->op == O_nilike,
                               a[0].p, tlen, a[1].p, a[1].len);
            if (m < 0)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "invalid escape in LIKE pattern");
            r->i = (f->op == O_like || f->op == O_ilike) ? m : !m;
            return true;
        }
        case K_substr:
        {
            /* __substring_common (xpu_textlib.cu:1159-1259), 1-origin start */
            int32_t start = (int32_t)a[1].i - 1;
            int32_t count = f->nargs > 2 ? (int32_t)a[2].i : -1;
            const char *str = a[0].p;
            int32_t length = a[0].len;
            if (f->nargs > 2 && count < 0)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "negative substring length not allowed");
            if (c->enc == ENC_NONE || c->enc == ENC_OTHER)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", c->enc == ENC_NONE
                              ? "No encoding info was supplied" : "encoding not supported");
            r->p = str;
            if (c->enc == ENC_SINGLE)
            {
                if (start >= length || count == 0 || (count > 0 && start + count <= 0))
                    r->len = 0;
                else if (start < 0)
                    r->len = (count < 0 || start + count >= length) ? length
                             : (start + count < length ? start + count : length);
                else
                {
                    int64_t end = (int64_t)start + count;
                    r->len = (int32_t)((count < 0 || end > length ? length : end) - start);
                    r->p = str + start;
                }
                return true;
            }
            if (start < 0)
            {
                if (count < 0)
                {
                    r->len = length;
                    return true;
                }
                count = start + count;
                if (count < 0)
                {
                    r->len = 0;
                    return true;
                }
                start = 0;
            }
            const char *pos = str, *end = str + length, *val = NULL;
            if (pos < end)
            {
                for (int i = 0;; i++)
                {
                    int w = utf8_mblen(pos);
                    if (pos + w > end)
                        break;
                    if (i == start)
                        val = pos;
                    if (count >= 0 && i == start + count)
                        break;
                    pos += w;
                }
            }
            if (val)
            {
                r->p = val;
                r->len = (int32_t)(pos - val);
            }
            else
                r->len = 0;
            return true;
        }
        case K_arith:
            if (is_int_type(f->res))
            {
                int64_t lo, hi;
                int128_t x = a[0].i, y = a[1].i, v = 0;
                int_range(f->res, &lo, &hi);
                switch (f->op)
                {
                    case O_pl: v = x + y; break;
                    case O_mi: v = x - y; break;
                    case O_mul: v = x * y; break;
                    case O_div:
                        if (y == 0)
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: division by zero", f->name);
                        v = x / y;
                        break;
                    case O_mod:
                        /* "<type>mod : division by zero" (xpu_basetype.cu:761) */
                        if (y == 0)
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s : division by zero", f->name);
                        v = (y == -1) ? 0 : x % y;
                        break;
                }
                if (v < lo || v > hi)
                    return oerror(c, ERRCODE_DEVICE_ERROR, "%s: value out of range", f->name);
                r->i = (int64_t)v;
                return true;
            }
            if (f->res == T_float4)
            {
                float x = a[0].f, y = a[1].f, v = 0;
                switch (f->op)
                {
                    case O_pl: v = x + y; break;
                    case O_mi: v = x - y; break;
                    case O_mul: v = x * y; break;
                    case O_div:
                        if (y == 0.0f)
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: division by zero", f->name);
                        v = x / y;
                        if (v == 0.0f && x != 0.0f)
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: value out of range", f->name);
                        break;
                }
                if (isinf(v) && !isinf(x) && !isinf(y))
                    return oerror(c, ERRCODE_DEVICE_ERROR, "%s: value out of range", f->name);
                r->f = v;
                return true;
            }
            else
            {
                double x = as_double(f->t0, &a[0]), y = as_double(f->t1, &a[1]), v = 0;
                switch (f->op)
                {
                    case O_pl: v = x + y; break;
                    case O_mi: v = x - y; break;
                    case O_mul: v = x * y; break;
                    case O_div:
                        if (y == 0.0)
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: division by zero", f->name);
                        v = x / y;
                        if (v == 0.0 && x != 0.0)
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: value out of range", f->name);
                        break;
                }
                if (isinf(v) && !isinf(x) && !isinf(y))
                    return oerror(c, ERRCODE_DEVICE_ERROR, "%s: value out of range", f->name);
                r->d = v;
                return true;
            }
        case K_unary:
            if (is_int_type(f->res))
            {
                int64_t lo, hi;
                int_range(f->res, &lo, &hi);
                switch (f->op)
                {
                    case O_up: r->i = a[0].i; break;
                    case O_um:
                    case O_abs:
                        if (a[0].i == lo && (f->op == O_um || a[0].i < 0))
                            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: value out of range", f->name);
                        r->i = (f->op == O_um) ? -a[0].i : (a[0].i < 0 ? -a[0].i : a[0].i);
                        break;
                    case O_not:
                        r->i = ~a[0].i;
                        break;
                }
                return true;
            }
            if (f->res == T_float4)
                r->f = f->op == O_um ? -a[0].f : f->op == O_abs ? fabsf(a[0].f) : a[0].f;
            else
                r->d = f->op == O_um ? -a[0].d : f->op == O_abs ? fabs(a[0].d) : a[0].d;
            return true;
        case K_bitop:
        {
            int64_t x = a[0].i, y = a[1].i, v = 0;
            int bits = type_width(f->res) * 8;
            switch (f->op)
            {
                case O_and: v = x & y; break;
                case O_or:  v = x | y; break;
                case O_xor: v = x ^ y; break;
                /* C's x << y / x >> y of xpu_basetype.cu:858-866 as the host
                 * CPU runs it (PostgreSQL's int2shl etc. on x86-64): int1 /
                 * int2 / int4 are shifted as 32-bit ints by y mod 32, int8 by
                 * y mod 64, then cut to the result width */
                case O_shl:
                    v = bits == 64 ? (int64_t)((uint64_t)x << (y & 63))
                                   : (int64_t)(int32_t)((uint32_t)x << (y & 31));
                    break;
                case O_shr:
                    v = bits == 64 ? x >> (y & 63) : (int64_t)((int32_t)x >> (y & 31));
                    break;
            }
            /* wrap into the result width */
            if (bits == 8) v = (int8_t)v;
            else if (bits == 16) v = (int16_t)v;
            else if (bits == 32) v = (int32_t)v;
            r->i = v;
            return true;
        }
        case K_numcmp:
        {
            /* PG_NUMERIC_COMPARE_TEMPLATE (xpu_numeric.cu:541-568) */
            if (!on_parse(c, &a[0]) || !on_parse(c, &a[1]))
                return false;
            r->i = cmp_result(f->op, on_cmp(a[0], a[1]));
            return true;
        }
        case K_numarith:
        {
            /* pgfn_numeric_add .. _mod (xpu_numeric.cu:570-703); numeric_mod
             * runs the reference's __decimal_div, as its pgfn does */
            if (!on_parse(c, &a[0]) || !on_parse(c, &a[1]))
                return false;
            switch (f->op)
            {
                case O_pl: return on_addsub(false, a[0], a[1], r);
                case O_mi: return on_addsub(true, a[0], a[1], r);
                case O_mul: return on_mul(a[0], a[1], r);
                default: return on_div(c, a[0], a[1], r);
            }
        }
        case K_numunary:
            /* numeric_uplus returns its argument itself; uminus and abs
             * parse it (xpu_numeric.cu:705-757) */
            *r = a[0];
            if (f->op == O_up)
                return true;
            if (!on_parse(c, r))
                return false;
            if (f->op == O_um)
            {
                if (r->nk == ON_VALID)
                    r->nv = -r->nv;
                else if (r->nk == ON_PINF)
                    r->nk = ON_NINF;
                else if (r->nk == ON_NINF)
                    r->nk = ON_PINF;
            }
            else if (r->nk == ON_VALID && r->nv < 0)
                r->nv = -r->nv;
            else if (r->nk == ON_NINF)
                r->nk = ON_PINF;
            return true;
        case K_numcast:
        {
            if (f->res == T_numeric)
            {
                /* PG_INT_TO_NUMERIC / PG_FLOAT_TO_NUMERIC (xpu_numeric.cu:398-476) */
                if (f->t0 == T_float8 || f->t0 == T_float4)
                    on_from_f64(f->t0 == T_float8 ? a[0].d : (double)a[0].f, r);
                else
                {
                    r->nk = ON_VALID;
                    r->nv = a[0].i;
                    r->nw = 0;
                }
                return true;
            }
            if (!on_parse(c, &a[0]))
                return false;
            if (is_int_type(f->res))
            {
                int64_t lo, hi;
                int_range(f->res, &lo, &hi);
                return on_to_int(c, &a[0], lo, hi, &r->i);
            }
            double dv;
            if (!on_to_f64(c, &a[0], &dv))
                return false;
            if (f->res == T_float4)
                r->f = (float)dv;
            else
                r->d = dv;
            return true;
        }
        case K_cast:
        {
            int s = f->t0, d = f->res;
            /* "<src> to <dst>: out of range" (PG_SIMPLE_TYPECAST_TEMPLATE,
             * xpu_basetype.cu:343) */
            char cname[64];
            const char *sep = strstr(f->name, "_to_");
            if (sep)
                snprintf(cname, sizeof(cname), "%.*s to %s", (int)(sep - f->name), f->name, sep + 4);
            else
                snprintf(cname, sizeof(cname), "%s", f->name);
            if (d == T_bool)
            {
                r->i = a[0].i != 0;
                return true;
            }
            if (is_int_type(d))
            {
                int64_t lo, hi;
                int_range(d, &lo, &hi);
                if (is_int_type(s))
                {
                    if (a[0].i < lo || a[0].i > hi)
                        return oerror(c, ERRCODE_DEVICE_ERROR, "%s: out of range", cname);
                    r->i = a[0].i;
                    return true;
                }
                if (s == T_float4)
                {
                    float v = rintf(a[0].f);
                    if (isnan(a[0].f) || v < (float)lo || v > (float)hi ||
                        (d == T_int8 && v >= 9223372036854775808.0f))
                        return oerror(c, ERRCODE_DEVICE_ERROR, "%s: out of range", cname);
                    r->i = (int64_t)v;
                    return true;
                }
                {
                    double v = rint(a[0].d);
                    if (isnan(a[0].d) || v < (double)lo || v > (double)hi ||
                        (d == T_int8 && v >= 9223372036854775808.0))
                        return oerror(c, ERRCODE_DEVICE_ERROR, "%s: out of range", cname);
                    r->i = (int64_t)v;
                    return true;
                }
            }
            if (d == T_float4 && s == T_float8)
            {
                /* __FLOAT_FITS_IN_FLOAT4 (xpu_basetype.cu:427-435) */
                double x = a[0].d;
                if (isinf(x) || isnan(x) || x < -3.40282346638528859812e+38 ||
                    x > 3.40282346638528859812e+38)
                    return oerror(c, ERRCODE_DEVICE_ERROR, "%s: out of range", cname);
                r->f = (float)x;
            }
            else if (d == T_float4)
                r->f = (float)a[0].i;
            else
                r->d = (s == T_float4) ? (double)a[0].f : (double)a[0].i;
            return true;
        }
    }
    return oerror(c, ERRCODE_DEVICE_ERROR, "%s: unsupported function", f->name);
}

/* the type's three-way comparison (xpu_*_datum_comp): integers by value,
 * floats by < and > (NaN compares equal to everything, xpu_basetype.cu:
 * 224-238), text by bytes then length, bpchar without trailing spaces */
static int
datum_comp(int type, const od *x, const od *y)
{
    switch (type)
    {
        case T_float4:
            return x->f > y->f ? 1 : x->f < y->f ? -1 : 0;
        case T_float8:
            return x->d > y->d ? 1 : x->d < y->d ? -1 : 0;
        case TypeOpCode__numeric:
            return on_cmp(*x, *y);  /* parsed by the caller */
        case T_text: case TypeOpCode__bytea: case T_bpchar:
        {
            int xl = x->len, yl = y->len;
            if (type == T_bpchar)
            {
                xl = bpchar_truelen(x->p, xl);
                yl = bpchar_truelen(y->p, yl);
            }
            int comp = memcmp(x->p, y->p, xl < yl ? xl : yl);
            if (comp == 0)
                comp = xl < yl ? -1 : xl > yl ? 1 : 0;
            return comp;
        }
        default:
            return x->i > y->i ? 1 : x->i < y->i ? -1 : 0;
    }
}

/* value equality of two non-NULL datums as the type's '=' operator (floats
 * with NaN equal to NaN only, as float8eq) */
static bool
datum_equal(int type, const od *x, const od *y)
{
    if (type == T_float4)
        return float_cmp(O_eq, (double)x->f, (double)y->f);
    if (type == T_float8)
        return float_cmp(O_eq, x->d, y->d);
    return datum_comp(type, x, y) == 0;
}

static bool eval(octx *c, const kern_expression *k, od *r);

/* scalar op ANY / ALL (array constant), as PostgreSQL's executor defines it
 * (ExecEvalScalarArrayOp): ANY is true at the first true comparison, else
 * NULL if a comparison was NULL, else false; ALL is false at the first false
 * one, else NULL if one was NULL, else true; a NULL array is NULL.  The
 * reference's pgfn_ScalarArrayOp (xpu_common.cu:815-1011) walks the
 * elements the same way (element into the temporary slot, then the
 * comparator), but its ALL returns true where a comparison is false. */
static bool
eval_saop(octx *c, const kern_expression *k, od *r)
{
    bool use_any = (k->opcode == FuncOpCode__ScalarArrayOpAny);
    const kern_expression *karr = KEXP_FIRST_ARG(k);
    const kern_expression *kcmp = karr ? KEXP_NEXT_ARG(karr) : NULL;
    int slot = k->u.saop.elem_slot_id;
    if (!karr || !kcmp || k->nr_args != 2 || slot >= c->nslots ||
        karr->opcode != FuncOpCode__ConstExpr)
        return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "ScalarArrayOp needs a constant array");
    if (karr->u.c.const_isnull)
    {
        r->isnull = true;
        return true;
    }
    const kern_expression *kelem = KEXP_NEXT_ARG(KEXP_FIRST_ARG(kcmp));
    int etype = kelem->exptype;
    const char *img = karr->u.c.const_value;
    uint32_t hdr;
    memcpy(&hdr, img, 4);
    int32_t ndim, dataoffset, nitems = 1;
    memcpy(&ndim, img + 4, 4);
    memcpy(&dataoffset, img + 8, 4);
    for (int d = 0; d < ndim; d++)
    {
        int32_t dim;
        memcpy(&dim, img + 16 + 4 * d, 4);
        nitems *= dim;
    }
    if (ndim == 0)
        nitems = 0;
    const uint8_t *nullmap = dataoffset ? (const uint8_t *)img + 16 + 8 * ndim : NULL;
    const char *data = img + (dataoffset ? dataoffset : XPU_MAXALIGN(16 + 8 * ndim));
    int typlen = type_width(etype), align = typlen > 0 ? typlen : 4;
    if (etype == T_bool || etype == T_int1)
        align = 1;
    uint32_t off = 0;
    bool meet_null = false;
    for (int32_t i = 0; i < nitems; i++)
    {
        od *e = &c->slots[slot];
        memset(e, 0, sizeof(od));
        if (nullmap && !(nullmap[i >> 3] & (1 << (i & 7))))
            e->isnull = true;
        else
        {
            const char *addr;
            off = XPU_TYPEALIGN(align, off);
            addr = data + off;
            if (!datum_from_image(c, etype, addr, e))
                return false;
            off += typlen > 0 ? (uint32_t)typlen : (uint32_t)(e->len + (e->p - addr));
        }
        od st;
        if (!eval(c, kcmp, &st))
            return false;
        if (st.isnull)
            meet_null = true;
        else if (use_any ? st.i : !st.i)
        {
            r->i = use_any ? 1 : 0;
            return true;
        }
    }
    r->isnull = meet_null;
    r->i = meet_null ? 0 : !use_any;
    return true;
}

/* the hash table of constant keys of HashedArrayOp (codegen.c:2820-2906,
 * 2458-2660): the value of the first item, in its slot's chain order, whose
 * key equals the scalar; else the default (at default_offset, or NULL).
 * NULL scalars give NULL.  As PostgreSQL's hashed "= ANY", an IN list (its
 * default is the header's zero padding) with a NULL element gives NULL
 * rather than false when nothing matches; the reference's
 * pgfn_HashedArrayOp (xpu_common.cu:1013-1088) gives false there.  A
 * matching item without a value gives NULL. */
static bool
eval_haop(octx *c, const kern_expression *k, od *r)
{
    const kern_expression *karg = KEXP_FIRST_ARG(k);
    od sv;
    if (!karg || k->nr_args != 1)
        return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "malformed HashedArrayOp");
    if (!eval(c, karg, &sv))
        return false;
    if (sv.isnull)
    {
        r->isnull = true;
        return true;
    }
    int etype = k->u.haop.elem_type;
    uint32_t nslots = k->u.haop.nslots;
    int ewidth = type_width(etype);
    int ralign = type_width(k->exptype) > 0 ? type_width(k->exptype) : 4;
    if (k->exptype == T_bool || k->exptype == T_int1)
        ralign = 1;
    bool in_list = (k->u.haop.default_offset == offsetof(kern_expression, u.haop.__padding));
    bool has_null_key = false;
    for (uint32_t sl = 0; sl < nslots; sl++)
    {
        uint32_t next = k->u.haop.hslots[sl];
        while (next != 0)
        {
            const char *item = (const char *)k + next;
            uint32_t inext;
            memcpy(&inext, item + 4, 4);
            if (!(inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY))
                has_null_key = true;
            next = inext & ~KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK;
        }
    }
    /* the chain of the scalar's hash, in chain order */
    if (raw_text(c, etype, &sv))
        return false;
    uint32_t h = datum_hash(etype, &sv);
    uint32_t next = k->u.haop.hslots[h % nslots];
    while (next != 0)
    {
        const char *item = (const char *)k + next;
        uint32_t ihash, inext;
        memcpy(&ihash, item, 4);
        memcpy(&inext, item + 4, 4);
        if (ihash == h && (inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY))
        {
            od ev;
            if (!datum_from_image(c, etype, item + 8, &ev))
                return false;
            if (datum_equal(etype, &sv, &ev))
            {
                if (!(inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE))
                {
                    r->isnull = true;
                    return true;
                }
                uint32_t klen = ewidth > 0 ? (uint32_t)ewidth : (uint32_t)(ev.len + (ev.p - (item + 8)));
                return datum_from_image(c, k->exptype, item + 8 + XPU_TYPEALIGN(ralign, klen), r);
            }
        }
        next = inext & ~KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK;
    }
    if (in_list && has_null_key)
    {
        r->isnull = true;
        return true;
    }
    if (!k->u.haop.default_offset)
    {
        r->isnull = true;
        return true;
    }
    return datum_from_image(c, k->exptype, (const char *)k + k->u.haop.default_offset, r);
}

static const char *
session_param(const kern_session_info *s, uint32_t id)
{
    if (id >= s->nparams || s->poffset[id] == 0)
        return NULL;
    return (const char *)s + s->poffset[id];
}

static bool
eval(octx *c, const kern_expression *k, od *r)
{
    const kern_expression *karg;
    od tmp;

    memset(r, 0, sizeof(od));
    switch (k->opcode)
    {
        case FuncOpCode__ConstExpr:
            if (k->u.c.const_isnull)
            {
                r->isnull = true;
                return true;
            }
            return datum_from_image(c, k->exptype, k->u.c.const_value, r);
        case FuncOpCode__ParamExpr:
        {
            const char *p = session_param(c->session, k->u.p.param_id);
            if (!p)
            {
                r->isnull = true;
                return true;
            }
            return datum_from_image(c, k->exptype, p, r);
        }
        case FuncOpCode__VarExpr:
            if (k->u.v.var_slot_id >= c->nslots)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s: slot out of range", "VarExpr");
            *r = c->slots[k->u.v.var_slot_id];
            return true;
        case FuncOpCode__SaveExpr:
            karg = KEXP_FIRST_ARG(k);
            if (!eval(c, karg, &tmp))
                return false;
            c->slots[k->u.save.sv_slot_id] = tmp;
            *r = tmp;
            return true;
        case FuncOpCode__BoolExpr_And:
        {
            /* SQL three-valued AND, stopping at the first false
             * (xpu_common.cu:472-503) */
            bool anynull = false;
            karg = KEXP_FIRST_ARG(k);
            for (uint32_t i = 0; i < k->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                if (!eval(c, karg, &tmp))
                    return false;
                if (tmp.isnull)
                    anynull = true;
                else if (!tmp.i)
                {
                    r->i = 0;
                    return true;
                }
            }
            r->isnull = anynull;
            r->i = !anynull;
            return true;
        }
        case FuncOpCode__BoolExpr_Or:
        {
            bool anynull = false;
            karg = KEXP_FIRST_ARG(k);
            for (uint32_t i = 0; i < k->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                if (!eval(c, karg, &tmp))
                    return false;
                if (tmp.isnull)
                    anynull = true;
                else if (tmp.i)
                {
                    r->i = 1;
                    return true;
                }
            }
            r->isnull = anynull;
            r->i = 0;
            return true;
        }
        case FuncOpCode__BoolExpr_Not:
            if (!eval(c, KEXP_FIRST_ARG(k), &tmp))
                return false;
            r->isnull = tmp.isnull;
            r->i = tmp.isnull ? 0 : !tmp.i;
            return true;
        case FuncOpCode__NullTestExpr_IsNull:
        case FuncOpCode__NullTestExpr_IsNotNull:
            if (!eval(c, KEXP_FIRST_ARG(k), &tmp))
                return false;
            r->i = (k->opcode == FuncOpCode__NullTestExpr_IsNull) ? tmp.isnull : !tmp.isnull;
            return true;
        case FuncOpCode__BoolTestExpr_IsTrue:
        case FuncOpCode__BoolTestExpr_IsNotTrue:
        case FuncOpCode__BoolTestExpr_IsFalse:
        case FuncOpCode__BoolTestExpr_IsNotFalse:
        case FuncOpCode__BoolTestExpr_IsUnknown:
        case FuncOpCode__BoolTestExpr_IsNotUnknown:
            if (!eval(c, KEXP_FIRST_ARG(k), &tmp))
                return false;
            switch (k->opcode)
            {
                case FuncOpCode__BoolTestExpr_IsTrue:     r->i = !tmp.isnull && tmp.i; break;
                case FuncOpCode__BoolTestExpr_IsNotTrue:  r->i = tmp.isnull || !tmp.i; break;
                case FuncOpCode__BoolTestExpr_IsFalse:    r->i = !tmp.isnull && !tmp.i; break;
                case FuncOpCode__BoolTestExpr_IsNotFalse: r->i = tmp.isnull || tmp.i; break;
                case FuncOpCode__BoolTestExpr_IsUnknown:  r->i = tmp.isnull; break;
                default:                                  r->i = !tmp.isnull; break;
            }
            return true;
        case FuncOpCode__CoalesceExpr:
            karg = KEXP_FIRST_ARG(k);
            for (uint32_t i = 0; i < k->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                if (!eval(c, karg, &tmp))
                    return false;
                if (!tmp.isnull)
                {
                    *r = tmp;
                    return true;
                }
            }
            r->isnull = true;
            return true;
        case FuncOpCode__DistinctFrom:
        {
            /* pgfn_DistinctFrom (xpu_common.cu:613-660): the arguments of
             * the '=' function, then the function itself */
            const kern_expression *kcmp = KEXP_FIRST_ARG(k);
            od x, y;
            if (!kcmp || kcmp->nr_args != 2)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "malformed DistinctFrom");
            karg = KEXP_FIRST_ARG(kcmp);
            if (!eval(c, karg, &x) || !eval(c, KEXP_NEXT_ARG(karg), &y))
                return false;
            if (x.isnull || y.isnull)
            {
                r->i = !(x.isnull && y.isnull);
                return true;
            }
            if (!eval(c, kcmp, &tmp))
                return false;
            r->i = !tmp.i;
            return true;
        }
        case FuncOpCode__LeastExpr:
        case FuncOpCode__GreatestExpr:
        {
            /* pgfn_LeastExpr / pgfn_GreatestExpr (xpu_common.cu:683-757):
             * the least / greatest non-NULL argument by the type's
             * comparison; an equal later argument does not replace it */
            bool least = (k->opcode == FuncOpCode__LeastExpr);
            r->isnull = true;
            karg = KEXP_FIRST_ARG(k);
            for (uint32_t i = 0; i < k->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                if (!eval(c, karg, &tmp))
                    return false;
                if (tmp.isnull)
                    continue;
                if (r->isnull)
                {
                    *r = tmp;
                    continue;
                }
                if (raw_text(c, k->exptype, r) || raw_text(c, k->exptype, &tmp))
                    return false;
                /* the type's comparison parses numeric datums in place */
                if (k->exptype == T_numeric && (!on_parse(c, r) || !on_parse(c, &tmp)))
                    return false;
                int comp = datum_comp(k->exptype, r, &tmp);
                if ((least && comp > 0) || (!least && comp < 0))
                    *r = tmp;
            }
            return true;
        }
        case FuncOpCode__CaseWhenExpr:
        {
            /* pgfn_CaseWhenExpr (xpu_common.cu:759-811) */
            if (k->u.casewhen.case_comp &&
                !eval(c, (const kern_expression *)((const char *)k + k->u.casewhen.case_comp), &tmp))
                return false;
            karg = KEXP_FIRST_ARG(k);
            for (uint32_t i = 0; i + 1 < k->nr_args; i += 2, karg = KEXP_NEXT_ARG(karg))
            {
                if (!eval(c, karg, &tmp))
                    return false;
                karg = KEXP_NEXT_ARG(karg);
                if (!tmp.isnull && tmp.i)
                    return eval(c, karg, r);
            }
            if (!k->u.casewhen.case_else)
            {
                r->isnull = true;
                return true;
            }
            return eval(c, (const kern_expression *)((const char *)k + k->u.casewhen.case_else), r);
        }
        case FuncOpCode__ScalarArrayOpAny:
        case FuncOpCode__ScalarArrayOpAll:
            return eval_saop(c, k, r);
        case FuncOpCode__HashedArrayOp:
            return eval_haop(c, k, r);
        case FuncOpCode__HashValue:
        {
            uint32_t h = 0xffffffffU;
            karg = KEXP_FIRST_ARG(k);
            for (uint32_t i = 0; i < k->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                if (!eval(c, karg, &tmp))
                    return false;
                if (raw_text(c, karg->exptype, &tmp))
                    return false;
                if (!tmp.isnull && karg->exptype == T_numeric && !on_parse(c, &tmp))
                    return false;
                if (!tmp.isnull)
                    h = pg_hash_merge(h, datum_hash(karg->exptype, &tmp));
            }
            r->i = (int64_t)(h ^ 0xffffffffU);
            return true;
        }
        default:
        {
            const ofunc *f = lookup_func(k->opcode);
            if (!f)
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s: unsupported opcode", "oracle");
            return eval_func(c, k, f, r);
        }
    }
}

/* ------------------------------------------------------------------
 * loaders
 * ------------------------------------------------------------------ */
static const kern_colmeta *
kds_cmeta(const kern_data_store *kds, int resno)
{
    return &kds->colmeta[resno - 1];
}

/* Arrow column (xpu_basetype.cu:242-306, xpu_textlib.cu:237-274) */
static bool
load_arrow(octx *c, const kern_data_store *kds, int resno, uint32_t row, int type, od *r)
{
    const kern_colmeta *cm = kds_cmeta(kds, resno);
    const char *base = (const char *)kds;

    memset(r, 0, sizeof(od));
    if (cm->nullmap_offset)
    {
        uint32_t idx = row >> 3;
        if (idx >= cm->nullmap_length ||
            (((const uint8_t *)base + cm->nullmap_offset)[idx] & (1U << (row & 7))) == 0)
        {
            r->isnull = true;
            return true;
        }
    }
    if (type == T_text || type == TypeOpCode__bytea || type == T_bpchar)
    {
        /* __xpu_varlena_datum_arrow_ref (xpu_textlib.cu:19-54): by the
         * column's Arrow type; an offset out of range reads as NULL */
        uint64_t o0, o1;
        switch (cm->attopts.tag)
        {
            case ArrowType__Utf8:
            case ArrowType__Binary:
            {
                const uint32_t *off = (const uint32_t *)(base + cm->values_offset);
                if (4 * ((uint64_t)row + 2) > cm->values_length)
                {
                    r->isnull = true;
                    return true;
                }
                o0 = off[row];
                o1 = off[row + 1];
                break;
            }
            case ArrowType__LargeUtf8:
            case ArrowType__LargeBinary:
            {
                const uint64_t *off = (const uint64_t *)(base + cm->values_offset);
                if (8 * ((uint64_t)row + 2) > cm->values_length)
                {
                    r->isnull = true;
                    return true;
                }
                o0 = off[row];
                o1 = off[row + 1];
                break;
            }
            case ArrowType__FixedSizeBinary:
            {
                uint64_t w = cm->attopts.u.fixed_size_binary.byteWidth;
                if (w * ((uint64_t)row + 1) > cm->values_length)
                {
                    r->isnull = true;
                    return true;
                }
                r->p = base + cm->values_offset + w * row;
                r->len = (int32_t)w;
                return true;
            }
            default:
                return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "not a mappable Arrow data type");
        }
        /* KDS_ARROW_REF_VARLENA32/64_DATUM (xpu_common.h:1577-1626) */
        if (o0 > o1 || o1 > cm->extra_length || o1 - o0 > 0x4ffffff8U)
        {
            r->isnull = true;
            return true;
        }
        r->p = base + cm->extra_offset + o0;
        r->len = (int32_t)(o1 - o0);
        return true;
    }
    if (type == T_bool)
    {
        const uint8_t *bits = (const uint8_t *)base + cm->values_offset;
        r->i = (bits[row >> 3] >> (row & 7)) & 1;
        return true;
    }
    if (type == T_numeric)
    {
        /* xpu_numeric_datum_arrow_read (xpu_numeric.cu:46-79): a
         * Decimal128 at the column's scale */
        if (cm->attopts.tag != ArrowType__Decimal || cm->attopts.u.decimal.bitWidth != 128)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s", "not a mappable Arrow data type");
        if (16 * ((uint64_t)row + 1) > cm->values_length)
        {
            r->isnull = true;
            return true;
        }
        memcpy(&r->nv, base + cm->values_offset + 16 * (uint64_t)row, 16);
        r->nk = ON_VALID;
        r->nw = (int16_t)cm->attopts.u.decimal.scale;
        on_trim(r);
        return true;
    }
    {
        int w = type_width(type);
        if (w < 0 || (uint64_t)w * (row + 1) > cm->values_length)
        {
            r->isnull = true;
            return true;
        }
        return datum_from_image(c, type, base + cm->values_offset + (uint64_t)w * row, r);
    }
}

/* attribute `resno` of a minimal tuple (xpu_common.cu:60-130 deform) */
static bool
load_tuple_attr(octx *c, const kern_data_store *kds, const kern_tupitem *t, int resno,
                int type, od *r)
{
    int ncols = t->t_infomask2 & HEAP_NATTS_MASK;
    bool hasnull = (t->t_infomask & HEAP_HASNULL) != 0;
    const char *pos = (const char *)t + t->t_hoff - MINIMAL_TUPLE_OFFSET;

    memset(r, 0, sizeof(od));
    if (resno > ncols || resno > (int)kds->ncols)
    {
        r->isnull = true;
        return true;
    }
    for (int j = 0; j < resno; j++)
    {
        const kern_colmeta *cm = &kds->colmeta[j];
        if (hasnull && !(t->t_bits[j >> 3] & (1 << (j & 7))))
        {
            if (j == resno - 1)
            {
                r->isnull = true;
                return true;
            }
            continue;
        }
        if (cm->attlen > 0 || *(const uint8_t *)pos == 0)
            pos = (const char *)t + XPU_TYPEALIGN(cm->attalign, pos - (const char *)t);
        if (j == resno - 1)
            return datum_from_image(c, type, pos, r);
        if (cm->attlen > 0)
            pos += cm->attlen;
        else
        {
            uint8_t b0 = *(const uint8_t *)pos;
            if (b0 == 0x01)
                return oerror(c, ERRCODE_SUSPEND_FALLBACK, "%s: external datum", "oracle");
            if (b0 & 0x01)
                pos += (b0 >> 1) & 0x7f;
            else
            {
                uint32_t h; memcpy(&h, pos, 4);
                pos += (h >> 2) & 0x3fffffff;
            }
        }
    }
    r->isnull = true;
    return true;
}

static bool
exec_loadvars_arrow(octx *c, const kern_expression *kload, const kern_data_store *kds, uint32_t row)
{
    if (!kload)
        return true;
    for (int i = 0; i < kload->u.load.nitems; i++)
    {
        int resno = kload->u.load.desc[i].vl_resno;
        int slot = kload->u.load.desc[i].vl_slot_id;
        if (resno <= 0 || resno > (int)kds->ncols)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: resno out of range", "LoadVars");
        if (!load_arrow(c, kds, resno, row, c->vdesc[slot].vs_type_code, &c->slots[slot]))
            return false;
    }
    return true;
}

static bool
exec_loadvars_tuple(octx *c, const kern_expression *kload, const kern_data_store *kds,
                    const kern_tupitem *t)
{
    if (!kload)
        return true;
    for (int i = 0; i < kload->u.load.nitems; i++)
    {
        int resno = kload->u.load.desc[i].vl_resno;
        int slot = kload->u.load.desc[i].vl_slot_id;
        if (!t)
        {
            memset(&c->slots[slot], 0, sizeof(od));
            c->slots[slot].isnull = true;
            continue;
        }
        if (!load_tuple_attr(c, kds, t, resno, c->vdesc[slot].vs_type_code, &c->slots[slot]))
            return false;
    }
    return true;
}

/* ------------------------------------------------------------------
 * heap tuples of KDS_FORMAT_BLOCK pages and KDS_FORMAT_ROW/HASH row chunks
 * ------------------------------------------------------------------ */

/* one tuple of a heap / row chunk: its HeapTupleHeaderData (for a
 * kern_tupitem, the virtual header MINIMAL_TUPLE_OFFSET bytes before it) and
 * the length of its item from there */
typedef struct
{
    const uint8_t  *h;
    uint32_t        len;
} otuple;

/* VARSIZE_ANY (xpu_common.h:1812-1815, 1743-1751) */
static uint32_t
varsize_any(const uint8_t *p)
{
    if (p[0] == 0x01)
        return 2 + (p[1] == 1 ? 8 : p[1] == 18 ? 16 : 0);
    if (p[0] & 0x01)
        return p[0] >> 1;
    uint32_t w;
    memcpy(&w, p, 4);
    return (w >> 2) & 0x3fffffff;
}

/* attribute `resno` of a heap tuple, walked as __kern_extract_common_tuple
 * (xpu_common.cu:60-130) on the payload kern_extract_heap_tuple hands it
 * (:194-257): ncols = Min(natts, kds->ncols), the null bitmap when
 * HEAP_HASNULL, positions aligned relative to the payload */
static bool
load_heap_attr(octx *c, const kern_data_store *kds, const otuple *tp, int resno, int type, od *r)
{
    const uint8_t *h = tp->h;
    uint16_t im2, im;
    memcpy(&im2, h + 18, 2);
    memcpy(&im, h + 20, 2);
    int ncols = im2 & HEAP_NATTS_MASK;
    if (ncols > (int)kds->ncols)
        ncols = kds->ncols;
    const uint8_t *nullmap = (im & HEAP_HASNULL) ? h + 23 : NULL;
    uint32_t hoff = h[22];
    const uint8_t *payload = h + hoff;
    uint32_t plen = tp->len > hoff ? tp->len - hoff : 0;
    uint64_t off = 0;

    memset(r, 0, sizeof(od));
    if (hoff < 23 || hoff > tp->len)
        return oerror(c, ERRCODE_DEVICE_ERROR, "%s: heap tuple header out of its item", "oracle");
    for (int resno_j = 1; resno_j <= ncols && resno_j <= resno; resno_j++)
    {
        int j = resno_j - 1;
        const kern_colmeta *cm = &kds->colmeta[j];
        if (nullmap && !(nullmap[j >> 3] & (1 << (j & 7))))
        {
            if (resno_j == resno)
            {
                r->isnull = true;
                return true;
            }
            continue;
        }
        if (cm->attlen > 0)
            off = XPU_TYPEALIGN(cm->attalign, off);
        else if (off < plen && payload[off] == 0)
            off = XPU_TYPEALIGN(cm->attalign, off);
        if (off >= plen)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: heap tuple attribute beyond its item", "oracle");
        uint64_t sz = cm->attlen > 0 ? (uint64_t)cm->attlen : varsize_any(payload + off);
        if (off + sz > plen || (cm->attlen <= 0 && cm->attlen != -1))
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: heap tuple attribute beyond its item", "oracle");
        if (resno_j == resno)
            return datum_from_image(c, type, (const char *)payload + off, r);
        off += sz;
    }
    r->isnull = true;
    return true;
}

static bool
exec_loadvars_heap(octx *c, const kern_expression *kload, const kern_data_store *kds,
                   const otuple *tp)
{
    if (!kload)
        return true;
    for (int i = 0; i < kload->u.load.nitems; i++)
    {
        int resno = kload->u.load.desc[i].vl_resno;
        int slot = kload->u.load.desc[i].vl_slot_id;
        if (resno <= 0 || resno > (int)kds->ncols)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: resno out of range", "LoadVars");
        if (!load_heap_attr(c, kds, tp, resno, c->vdesc[slot].vs_type_code, &c->slots[slot]))
            return false;
    }
    return true;
}

/* __gpuscan_check_row_visibility (cuda_gpuscan.cu:614-642) against the
 * session's parallelCurrentXids */
static bool
row_visible(const kern_session_info *s, size_t slen, const uint8_t *t)
{
    if (!t[8])
        return true;
    uint32_t xmin, xmax, n = 0;
    const uint32_t *xids = NULL;
    const uint8_t *xa = t + t[14] - MINIMAL_TUPLE_OFFSET - sizeof(kern_tupitem_xact_attrs);
    memcpy(&xmin, xa, 4);
    memcpy(&xmax, xa + 4, 4);
    if (s->session_xact_state &&
        s->session_xact_state + offsetof(SerializedTransactionState, parallelCurrentXids) <= slen)
    {
        const SerializedTransactionState *xs =
            (const SerializedTransactionState *)((const char *)s + s->session_xact_state);
        n = xs->nParallelCurrentXids > 0 ? (uint32_t)xs->nParallelCurrentXids : 0;
        xids = xs->parallelCurrentXids;
    }
    if (xmin == InvalidTransactionId)
        return false;
    if (xmin != FrozenTransactionId)
    {
        bool found = false;
        for (uint32_t i = 0; i < n; i++)
            if (xids[i] == xmin)
                found = true;
        if (!found)
            return false;
    }
    if (xmax == InvalidTransactionId)
        return true;
    if (xmax == FrozenTransactionId)
        return false;
    for (uint32_t i = 0; i < n; i++)
        if (xids[i] == xmax)
            return false;
    return true;
}

/* The tuples a scan of the chunk loads, in order: the LP_NORMAL items of
 * each page (cuda_gpuscan.cu:479-507; PageGetMaxOffsetNumber,
 * xpu_common.h:1123-1127), or the visible entries of the row index
 * (:662-672).  Returns the count, or -1 with the error set. */
static void set_error(kern_errorbuf *eb, uint32_t code, const char *func, const char *msg);

static int64_t
collect_tuples(const kern_session_info *sess, size_t slen, const kern_data_store *kds,
               otuple **out, kern_errorbuf *eb)
{
    const uint8_t *base = (const uint8_t *)kds;
    size_t hl = KDS_HEAD_LENGTH(kds);
    size_t cap = 1024, n = 0;
    otuple *v = malloc(sizeof(otuple) * cap);

    if (kds->format == KDS_FORMAT_BLOCK)
    {
        if (kds->block_offset < hl || kds->block_offset + (uint64_t)XPU_BLCKSZ * kds->nitems > kds->length)
            goto bad;
        for (uint32_t b = 0; b < kds->nitems; b++)
        {
            const uint8_t *pg = base + kds->block_offset + (uint64_t)XPU_BLCKSZ * b;
            uint16_t pd_lower;
            memcpy(&pd_lower, pg + 12, 2);
            if (pd_lower > XPU_BLCKSZ)
                goto bad;
            uint32_t nlp = pd_lower <= 24 ? 0 : (pd_lower - 24) / 4;
            for (uint32_t k = 0; k < nlp; k++)
            {
                uint32_t lp;
                memcpy(&lp, pg + 24 + 4 * k, 4);
                if (((lp >> 15) & 3) != LP_NORMAL)
                    continue;
                uint32_t off = lp & 0x7fff, len = lp >> 17;
                if (off < 24 || len < 23 || off + len > XPU_BLCKSZ)
                    goto bad;
                if (n == cap)
                    v = realloc(v, sizeof(otuple) * (cap *= 2));
                v[n].h = pg + off;
                v[n].len = len;
                n++;
            }
        }
    }
    else
    {
        uint64_t ridx_off = hl + 8ULL * kds->hash_nslots;
        if (ridx_off + 8ULL * kds->nitems > kds->length)
            goto bad;
        for (uint32_t i = 0; i < kds->nitems; i++)
        {
            uint64_t off;
            memcpy(&off, base + ridx_off + 8ULL * i, 8);
            if (off == 0)
                continue;
            if (off < 16 || off > kds->length - (ridx_off + 8ULL * kds->nitems))
                goto bad;
            const uint8_t *t = base + kds->length - off;
            uint32_t t_len;
            memcpy(&t_len, t, 4);
            if (t_len > off || t[14] < MINIMAL_TUPLE_OFFSET + KERN_TUPITEM_T_BITS_OFFSET ||
                (uint32_t)t[14] - MINIMAL_TUPLE_OFFSET > t_len)
                goto bad;
            if (!row_visible(sess, slen, t))
                continue;
            if (n == cap)
                v = realloc(v, sizeof(otuple) * (cap *= 2));
            v[n].h = t - MINIMAL_TUPLE_OFFSET;
            v[n].len = MINIMAL_TUPLE_OFFSET + t_len;
            n++;
        }
    }
    *out = v;
    return (int64_t)n;
bad:
    free(v);
    set_error(eb, ERRCODE_DEVICE_ERROR, "collect_tuples", "page item or row-index entry out of the chunk");
    return -1;
}

/* ------------------------------------------------------------------
 * session
 * ------------------------------------------------------------------ */
typedef struct
{
    uint32_t    hash;
    uint32_t    nkeys;
    od          keys[OMAX_ATTRS];
    char       *keybuf;         /* owns copies of text key bytes */
    /* per attribute state */
    int64_t     i0[OMAX_ATTRS];     /* count / nitems / min/max int */
    int64_t     i1[OMAX_ATTRS];     /* psum_int sum */
    int128_t    s128[OMAX_ATTRS];
    double      d0[OMAX_ATTRS];     /* fp sum or fp min/max */
    double      d1[OMAX_ATTRS];
    uint32_t    attrs[OMAX_ATTRS];
} ogroup;

typedef struct
{
    ogroup    **groups;
    size_t      ngroups;
    size_t      nrooms;
    int64_t    *slots;          /* open addressing: group index + 1 */
    size_t      nslots;
} ogtable;

struct oracle_session
{
    kern_session_info *session;
    size_t      length;
    int         enc;            /* the session's database encoding (ENC_*) */
    const kern_varslot_desc *vdesc;
    int         nslots;
    int         nrels;
    const kern_expression *load[OMAX_DEPTH];
    const kern_expression *hashv[OMAX_DEPTH];
    const kern_expression *joinq[OMAX_DEPTH];
    const kern_expression *scan_quals;
    const kern_expression *proj;
    const kern_expression *kds_dst_head_dummy;
    const kern_data_store *kds_dst_head;
    const kern_data_store *kds_final_head;
    const kern_expression *gb_keyhash;
    const kern_expression *gb_keyload;
    const kern_expression *gb_keycomp;
    const kern_expression *gb_actions;
    int         key_types[OMAX_ATTRS];
    ogtable     gt;
    /* CPU fallback (executor.c:705-724): the scan slot's header and the
     * slot of each of its attributes at depth 0 (-1: NULL) */
    const kern_data_store *kds_fb_head;
    int         fb_slot[OMAX_FB_ATTRS];
    /* fallback items of the tasks so far: kern_fallbackitem images */
    char       *fb_buf;
    size_t      fb_len, fb_cap;
    uint32_t    fb_nrows;
};

static const kern_expression *
session_kexp(const kern_session_info *s, uint32_t off)
{
    return off ? (const kern_expression *)((const char *)s + off) : NULL;
}

static const kern_expression *
packed_sub(const kern_expression *kpack, int idx)
{
    if (!kpack || idx < 0 || (uint32_t)idx > kpack->u.pack.npacked)
        return NULL;
    uint32_t off = kpack->u.pack.offset[idx];
    return off ? (const kern_expression *)((const char *)kpack + off) : NULL;
}

static void
set_error(kern_errorbuf *eb, uint32_t code, const char *func, const char *msg)
{
    if (!eb)
        return;
    memset(eb, 0, sizeof(*eb));
    eb->errcode = code;
    snprintf(eb->filename, sizeof(eb->filename), "oracle.c");
    snprintf(eb->funcname, sizeof(eb->funcname), "%s", func);
    snprintf(eb->message, sizeof(eb->message), "%s", msg);
}

oracle_session *
oracle_session_open(const kern_session_info *session, size_t length, kern_errorbuf *eb)
{
    oracle_session *os = calloc(1, sizeof(oracle_session));
    kern_session_info *s = malloc(length);

    memcpy(s, session, length);
    os->session = s;
    os->enc = session_encoding(s);
    os->length = length;
    os->nslots = s->kcxt_kvars_nslots;
    if (os->nslots > OMAX_SLOTS)
    {
        set_error(eb, ERRCODE_DEVICE_ERROR, __func__, "too many slots");
        free(s); free(os);
        return NULL;
    }
    os->vdesc = s->kcxt_kvars_defs ? (const kern_varslot_desc *)((const char *)s + s->kcxt_kvars_defs) : NULL;
    {
        const kern_expression *kp = session_kexp(s, s->xpucode_load_vars_packed);
        const kern_expression *kj = session_kexp(s, s->xpucode_join_quals_packed);
        const kern_expression *kh = session_kexp(s, s->xpucode_hash_values_packed);
        os->nrels = kj ? (int)kj->u.pack.npacked - 1 : 0;
        if (os->nrels < 0) os->nrels = 0;
        for (int d = 0; d <= os->nrels && d < OMAX_DEPTH; d++)
        {
            os->load[d] = packed_sub(kp, d);
            os->joinq[d] = packed_sub(kj, d);
            os->hashv[d] = packed_sub(kh, d);
        }
    }
    os->scan_quals = session_kexp(s, s->xpucode_scan_quals);
    os->proj = session_kexp(s, s->xpucode_projection);
    os->kds_dst_head = s->projection_kds_dst ? (const kern_data_store *)((const char *)s + s->projection_kds_dst) : NULL;
    os->kds_final_head = s->groupby_kds_final ? (const kern_data_store *)((const char *)s + s->groupby_kds_final) : NULL;
    os->gb_keyhash = session_kexp(s, s->xpucode_groupby_keyhash);
    os->gb_keyload = session_kexp(s, s->xpucode_groupby_keyload);
    os->gb_keycomp = session_kexp(s, s->xpucode_groupby_keycomp);
    os->gb_actions = session_kexp(s, s->xpucode_groupby_actions);
    if (s->fallback_kds_head)
    {
        const kern_data_store *h = (const kern_data_store *)((const char *)s + s->fallback_kds_head);
        const kern_fallback_desc *fd = (const kern_fallback_desc *)((const char *)s + s->fallback_desc_defs);
        if (h->ncols > OMAX_FB_ATTRS || h->format != KDS_FORMAT_FALLBACK)
        {
            set_error(eb, ERRCODE_DEVICE_ERROR, __func__, "unsupported fallback buffer header");
            free(s); free(os);
            return NULL;
        }
        os->kds_fb_head = h;
        for (uint32_t j = 0; j < h->ncols; j++)
            os->fb_slot[j] = -1;
        for (uint32_t i = 0; i < s->fallback_desc_nitems; i++)
            if (fd[i].fb_src_depth == 0 && fd[i].fb_dst_resno >= 1 &&
                (uint32_t)fd[i].fb_dst_resno <= h->ncols && fd[i].fb_slot_id < os->nslots)
                os->fb_slot[fd[i].fb_dst_resno - 1] = fd[i].fb_slot_id;
    }
    if (os->gb_actions)
    {
        if (os->gb_actions->u.pagg.nattrs > OMAX_ATTRS)
        {
            set_error(eb, ERRCODE_DEVICE_ERROR, __func__, "too many group-by attributes");
            free(s); free(os);
            return NULL;
        }
        for (int j = 0; j < os->gb_actions->u.pagg.nattrs; j++)
        {
            const kern_aggregate_desc *d = &os->gb_actions->u.pagg.desc[j];
            os->key_types[j] = (d->arg0_slot_id >= 0) ? os->vdesc[d->arg0_slot_id].vs_type_code : 0;
        }
    }
    return os;
}

/* the CPU-fallback rows of the tasks so far as a KDS_FORMAT_FALLBACK chunk
 * (header, row index, items at the tail in input order, each tuple's row id
 * its index), as gpuClientWriteBackPartial sends them
 * (gpu_service.c:2853-2889); the store is emptied.  Returns the bytes
 * written, 0 with no rows, the bytes needed with out == NULL, or -1. */
int64_t
oracle_session_fallback(oracle_session *os, kern_data_store *out, size_t length)
{
    if (!os->kds_fb_head || os->fb_nrows == 0)
        return 0;
    const kern_data_store *h = os->kds_fb_head;
    size_t head_len = KDS_HEAD_LENGTH(h);
    uint64_t usage = 0;
    for (size_t pos = 0; pos < os->fb_len; )
    {
        uint32_t ilen;
        memcpy(&ilen, os->fb_buf + pos, 4);
        usage += ilen;
        pos += 4 + ilen;
    }
    uint64_t need = head_len + 8ULL * os->fb_nrows + usage;
    if (!out)
        return (int64_t)need;
    if (length < need)
        return -1;
    memcpy(out, h, head_len);
    out->length = need;
    out->nitems = os->fb_nrows;
    out->usage = usage;
    uint64_t *rowidx = (uint64_t *)((char *)out + head_len);
    uint64_t u = 0;
    uint32_t i = 0;
    for (size_t pos = 0; pos < os->fb_len; i++)
    {
        uint32_t ilen;
        memcpy(&ilen, os->fb_buf + pos, 4);
        u += ilen;
        char *item = (char *)out + need - u;
        memcpy(item, os->fb_buf + pos + 4, ilen);
        kern_tupitem *t = (kern_tupitem *)(item + 16);
        memcpy((char *)t + t->t_len - ROWID_SIZE, &i, ROWID_SIZE);
        rowidx[i] = u - 16;
        pos += 4 + ilen;
    }
    os->fb_len = 0;
    os->fb_nrows = 0;
    return (int64_t)need;
}

void
oracle_session_close(oracle_session *os)
{
    if (!os)
        return;
    for (size_t i = 0; i < os->gt.ngroups; i++)
    {
        free(os->gt.groups[i]->keybuf);
        free(os->gt.groups[i]);
    }
    free(os->gt.groups);
    free(os->gt.slots);
    free(os->fb_buf);
    free(os->session);
    free(os);
}

/* ------------------------------------------------------------------
 * One expression over variable slots, without a session: the checker of
 * each device function against the reference's golden vectors.  Slot i has
 * type code slot_types[i] and the PostgreSQL datum image at slot_data +
 * slot_offs[i] (SQL NULL if negative).  Returns 0 with the result's image
 * (*res_len = -1 for NULL), or 1 with the error code and message.
 * ------------------------------------------------------------------ */
int
oracle_eval_expr(const void *kexp_blob, size_t kexp_len, int compare_nulls, const char *encname,
                 int nslots, const int32_t *slot_types, const uint8_t *slot_data,
                 const int64_t *slot_offs, int32_t res_type, uint8_t *res_image, int32_t res_cap,
                 int32_t *res_len, uint32_t *errcode, char *errmsg, int errlen)
{
    static const kern_session_info empty_session;
    uint64_t *ebuf = calloc(kexp_len / 8 + 2, 8);
    octx *c = calloc(1, sizeof(octx));
    od r;
    int rc = 0;

    memcpy(ebuf, kexp_blob, kexp_len);
    c->session = &empty_session;
    c->compare_nulls = compare_nulls != 0;
    c->enc = encoding_kind(encname);
    c->nslots = nslots;
    *errcode = 0;
    for (int i = 0; i < nslots && i < OMAX_SLOTS; i++)
    {
        if (slot_offs[i] < 0)
            c->slots[i].isnull = true;
        else if (!datum_from_image(c, slot_types[i], (const char *)slot_data + slot_offs[i],
                                   &c->slots[i]))
        {
            rc = -1;
            goto out;
        }
    }
    if (!eval(c, (const kern_expression *)ebuf, &r))
    {
        *errcode = c->errcode;
        snprintf(errmsg, errlen, "%s", c->errmsg);
        rc = 1;
        goto out;
    }
    if (r.isnull)
        *res_len = -1;
    else
    {
        int n = datum_write(res_type, &r, NULL);
        if (n < 0 || n > res_cap)
        {
            rc = -1;
            goto out;
        }
        datum_write(res_type, &r, (char *)res_image);
        *res_len = n;
    }
out:
    free(c);
    free(ebuf);
    return rc;
}

static void
octx_init(octx *c, const oracle_session *os)
{
    memset(c, 0, sizeof(octx));
    c->session = os->session;
    c->vdesc = os->vdesc;
    c->nslots = os->nslots;
    c->enc = os->enc;
}

/* ------------------------------------------------------------------
 * projection: kern_form_minimal_tuple (xpu_common.cu:1096-1190)
 * ------------------------------------------------------------------ */
static int
form_tuple(const oracle_session *os, octx *c, char *buf, uint32_t rowid)
{
    const kern_expression *kp = os->proj;
    const kern_data_store *kds_dst = os->kds_dst_head;
    int nattrs = kp->u.proj.nattrs;
    bool hasnull = false;
    uint32_t t_hoff;
    uint16_t infomask = 0;
    kern_tupitem *t = (kern_tupitem *)buf;

    if ((int)kds_dst->ncols < nattrs)
        nattrs = kds_dst->ncols;
    for (int j = 0; j < nattrs; j++)
    {
        int s = kp->u.proj.slot_id[j];
        if (s >= c->nslots || c->slots[s].isnull)
        {
            hasnull = true;
            infomask |= HEAP_HASNULL;
            break;
        }
    }
    t_hoff = KERN_TUPITEM_T_BITS_OFFSET;
    if (hasnull)
        t_hoff += XPU_BITMAPLEN(nattrs);
    t_hoff = XPU_MAXALIGN(t_hoff);
    if (buf)
    {
        memset(buf, 0, t_hoff);
        t->t_infomask2 = nattrs & HEAP_NATTS_MASK;
        t->t_hoff = t_hoff + MINIMAL_TUPLE_OFFSET;
    }
    for (int j = 0; j < nattrs; j++)
    {
        const kern_colmeta *cm = &kds_dst->colmeta[j];
        int s = kp->u.proj.slot_id[j];
        if (s >= c->nslots || c->slots[s].isnull)
            continue;
        uint32_t nxt = XPU_TYPEALIGN(cm->attalign, t_hoff);
        if (buf && nxt > t_hoff)
            memset(buf + t_hoff, 0, nxt - t_hoff);
        int sz = datum_write(c->vdesc[s].vs_type_code, &c->slots[s], buf ? buf + nxt : NULL);
        if (sz < 0)
            return -1;
        if (cm->attlen == -1)
        {
            if (buf && (uint8_t)buf[nxt] == 0x01)
                infomask |= HEAP_HASEXTERNAL;       /* VARATT_IS_EXTERNAL */
            infomask |= HEAP_HASVARWIDTH;
        }
        if (buf && hasnull)
            t->t_bits[j >> 3] |= (1 << (j & 7));
        t_hoff = nxt + sz;
    }
    t_hoff += ROWID_SIZE;
    if (buf)
    {
        t->t_len = t_hoff;
        t->t_infomask = infomask;
        memcpy(buf + t_hoff - ROWID_SIZE, &rowid, ROWID_SIZE);
    }
    return t_hoff;
}

/* the CPU-fallback tuple of a row: the scan slot's attributes, those from
 * depth-0 variables with their values, the rest NULL
 * (HandleErrorIfCpuFallback at depth 0, xpu_common.cu:1898-1955, forming
 * it with __kern_form_minimal_tuple, :1096-1190) */
static int
form_fallback_tuple(const oracle_session *os, octx *c, char *buf, uint32_t rowid)
{
    const kern_data_store *h = os->kds_fb_head;
    int nattrs = (int)h->ncols;
    bool hasnull = false;
    uint32_t t_hoff;
    uint16_t infomask = 0;
    kern_tupitem *t = (kern_tupitem *)buf;

    for (int j = 0; j < nattrs; j++)
    {
        int s = os->fb_slot[j];
        if (s < 0 || s >= c->nslots || c->slots[s].isnull)
        {
            hasnull = true;
            infomask |= HEAP_HASNULL;
            break;
        }
    }
    t_hoff = KERN_TUPITEM_T_BITS_OFFSET;
    if (hasnull)
        t_hoff += XPU_BITMAPLEN(nattrs);
    t_hoff = XPU_MAXALIGN(t_hoff);
    if (buf)
    {
        memset(buf, 0, t_hoff);
        t->t_infomask2 = nattrs & HEAP_NATTS_MASK;
        t->t_hoff = t_hoff + MINIMAL_TUPLE_OFFSET;
    }
    for (int j = 0; j < nattrs; j++)
    {
        const kern_colmeta *cm = &h->colmeta[j];
        int s = os->fb_slot[j];
        if (s < 0 || s >= c->nslots || c->slots[s].isnull)
            continue;
        uint32_t nxt = XPU_TYPEALIGN(cm->attalign, t_hoff);
        if (buf && nxt > t_hoff)
            memset(buf + t_hoff, 0, nxt - t_hoff);
        int sz = datum_write(c->vdesc[s].vs_type_code, &c->slots[s], buf ? buf + nxt : NULL);
        if (sz < 0)
            return -1;
        if (cm->attlen == -1)
        {
            if (buf && (uint8_t)buf[nxt] == 0x01)
                infomask |= HEAP_HASEXTERNAL;
            infomask |= HEAP_HASVARWIDTH;
        }
        if (buf && hasnull)
            t->t_bits[j >> 3] |= (1 << (j & 7));
        t_hoff = nxt + sz;
    }
    t_hoff += ROWID_SIZE;
    if (buf)
    {
        t->t_len = t_hoff;
        t->t_infomask = infomask;
        memcpy(buf + t_hoff - ROWID_SIZE, &rowid, ROWID_SIZE);
    }
    return t_hoff;
}

/* run the projection's SaveExprs to fill the slots */
static bool
prep_projection(const oracle_session *os, octx *c)
{
    const kern_expression *kp = os->proj;
    const kern_expression *karg = KEXP_FIRST_ARG(kp);
    od tmp;
    for (uint32_t i = 0; i < kp->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
        if (!eval(c, karg, &tmp))
            return false;
    return true;
}

/* ------------------------------------------------------------------
 * group-by
 * ------------------------------------------------------------------ */
static bool
run_actions_saves(const oracle_session *os, octx *c)
{
    const kern_expression *ka = os->gb_actions;
    const kern_expression *karg = KEXP_FIRST_ARG(ka);
    od tmp;
    for (uint32_t i = 0; i < ka->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
        if (!eval(c, karg, &tmp))
            return false;
    /* a numeric sum's argument is parsed in its slot, raising its error
     * (__update_groupby__psum_numeric's xpu_numeric_validate) */
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
    {
        const kern_aggregate_desc *d = &ka->u.pagg.desc[j];
        if ((d->action == KAGG_ACTION__PSUM_NUMERIC || d->action == KAGG_ACTION__PAVG_NUMERIC) &&
            d->arg0_slot_id >= 0 && !c->slots[d->arg0_slot_id].isnull &&
            !on_parse(c, &c->slots[d->arg0_slot_id]))
            return false;
    }
    return true;
}

static bool
keys_equal(const oracle_session *os, const ogroup *g, octx *c)
{
    const kern_expression *ka = os->gb_actions;
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
    {
        const kern_aggregate_desc *d = &ka->u.pagg.desc[j];
        if (d->action != KAGG_ACTION__VREF)
            continue;
        const od *a = &c->slots[d->arg0_slot_id];
        const od *b = &g->keys[j];
        int t = os->key_types[j];
        /* kmode_compare_nulls: NULL keys group together */
        if (a->isnull || b->isnull)
        {
            if (a->isnull != b->isnull)
                return false;
            continue;
        }
        switch (t)
        {
            case T_float4: if (!float_cmp(O_eq, a->f, b->f)) return false; break;
            case T_float8: if (!float_cmp(O_eq, a->d, b->d)) return false; break;
            case T_text: case TypeOpCode__bytea: case TypeOpCode__bpchar:
                if (a->len != b->len || memcmp(a->p, b->p, a->len) != 0) return false;
                break;
            case T_numeric:
            {
                /* by value (xpu_numeric_datum_comp) */
                od x = *a, y = *b;
                if (!on_parse(c, &x) || !on_parse(c, &y) || on_cmp(x, y) != 0)
                    return false;
                break;
            }
            default: if (a->i != b->i) return false; break;
        }
    }
    return true;
}

static void
group_init(const oracle_session *os, ogroup *g)
{
    const kern_expression *ka = os->gb_actions;
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
    {
        switch (ka->u.pagg.desc[j].action)
        {
            case KAGG_ACTION__PMIN_INT32:
            case KAGG_ACTION__PMIN_INT64: g->i0[j] = INT64_MAX; break;
            case KAGG_ACTION__PMAX_INT32:
            case KAGG_ACTION__PMAX_INT64: g->i0[j] = INT64_MIN; break;
            case KAGG_ACTION__PMIN_FP64: g->d0[j] = 1.7976931348623157e308; break;
            case KAGG_ACTION__PMAX_FP64: g->d0[j] = -1.7976931348623157e308; break;
            case KAGG_ACTION__PSUM_NUMERIC:
            case KAGG_ACTION__PAVG_NUMERIC:
                /* the sum's scale, from the argument's typmod */
                g->attrs[j] = (uint32_t)on_typmod_scale(ka->u.pagg.desc[j].typmod);
                break;
            default: break;
        }
    }
}

static ogroup *
group_lookup(const oracle_session *os, ogtable *gt, octx *c, uint32_t hash)
{
    const kern_expression *ka = os->gb_actions;
    if (gt->nslots == 0 || gt->ngroups * 2 >= gt->nslots)
    {
        size_t ns = gt->nslots ? gt->nslots * 2 : 1024;
        int64_t *slots = calloc(ns, sizeof(int64_t));
        for (size_t i = 0; i < gt->ngroups; i++)
        {
            size_t h = gt->groups[i]->hash & (ns - 1);
            while (slots[h]) h = (h + 1) & (ns - 1);
            slots[h] = i + 1;
        }
        free(gt->slots);
        gt->slots = slots;
        gt->nslots = ns;
    }
    size_t h = hash & (gt->nslots - 1);
    while (gt->slots[h])
    {
        ogroup *g = gt->groups[gt->slots[h] - 1];
        if (g->hash == hash && keys_equal(os, g, c))
            return g;
        h = (h + 1) & (gt->nslots - 1);
    }
    /* insert */
    ogroup *g = calloc(1, sizeof(ogroup));
    size_t klen = 0, kpos = 0;
    g->hash = hash;
    group_init(os, g);
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
        if (ka->u.pagg.desc[j].action == KAGG_ACTION__VREF &&
            c->slots[ka->u.pagg.desc[j].arg0_slot_id].len > 0)
            klen += c->slots[ka->u.pagg.desc[j].arg0_slot_id].len;
    g->keybuf = malloc(klen + 1);
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
    {
        const kern_aggregate_desc *d = &ka->u.pagg.desc[j];
        if (d->action != KAGG_ACTION__VREF)
            continue;
        g->keys[j] = c->slots[d->arg0_slot_id];
        if (g->keys[j].p && !g->keys[j].isnull && g->keys[j].len > 0)
        {
            memcpy(g->keybuf + kpos, g->keys[j].p, g->keys[j].len);
            g->keys[j].p = g->keybuf + kpos;
            kpos += g->keys[j].len;
        }
    }
    if (gt->ngroups == gt->nrooms)
    {
        gt->nrooms = gt->nrooms ? gt->nrooms * 2 : 1024;
        gt->groups = realloc(gt->groups, sizeof(ogroup *) * gt->nrooms);
    }
    gt->groups[gt->ngroups++] = g;
    gt->slots[h] = gt->ngroups;
    return g;
}

static bool
fetch_int64(int t, const od *v, int64_t *out)
{
    if (v->isnull) return false;
    if (t == T_float4) *out = (int64_t)v->f;
    else if (t == T_float8) *out = (int64_t)v->d;
    else *out = v->i;
    return true;
}

static bool
fetch_fp64(int t, const od *v, double *out)
{
    if (v->isnull) return false;
    *out = as_double(t, v);
    return true;
}

/* __updateOneTupleGroupBy (cuda_gpupreagg.cu:1499-1605) */
static void
group_update(const oracle_session *os, ogroup *g, octx *c)
{
    const kern_expression *ka = os->gb_actions;
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
    {
        const kern_aggregate_desc *d = &ka->u.pagg.desc[j];
        int t = d->arg0_slot_id >= 0 ? c->vdesc[d->arg0_slot_id].vs_type_code : 0;
        const od *v = d->arg0_slot_id >= 0 ? &c->slots[d->arg0_slot_id] : NULL;
        int64_t iv;
        double fv;

        if (d->filter_slot_id >= 0)
        {
            const od *fl = &c->slots[d->filter_slot_id];
            if (fl->isnull || !fl->i)
                continue;
        }
        switch (d->action)
        {
            case KAGG_ACTION__NROWS_ANY:
                g->i0[j]++;
                break;
            case KAGG_ACTION__NROWS_COND:
                if (!v->isnull) g->i0[j]++;
                break;
            case KAGG_ACTION__PMIN_INT32:
            case KAGG_ACTION__PMIN_INT64:
                if (fetch_int64(t, v, &iv))
                {
                    g->attrs[j] |= __PAGG_MINMAX_ATTRS__VALID;
                    if (iv < g->i0[j]) g->i0[j] = iv;
                }
                break;
            case KAGG_ACTION__PMAX_INT32:
            case KAGG_ACTION__PMAX_INT64:
                if (fetch_int64(t, v, &iv))
                {
                    g->attrs[j] |= __PAGG_MINMAX_ATTRS__VALID;
                    if (iv > g->i0[j]) g->i0[j] = iv;
                }
                break;
            case KAGG_ACTION__PMIN_FP64:
                if (fetch_fp64(t, v, &fv))
                {
                    g->attrs[j] |= __PAGG_MINMAX_ATTRS__VALID;
                    if (fv < g->d0[j]) g->d0[j] = fv;
                }
                break;
            case KAGG_ACTION__PMAX_FP64:
                if (fetch_fp64(t, v, &fv))
                {
                    g->attrs[j] |= __PAGG_MINMAX_ATTRS__VALID;
                    if (fv > g->d0[j]) g->d0[j] = fv;
                }
                break;
            case KAGG_ACTION__PSUM_INT:
            case KAGG_ACTION__PAVG_INT:
                if (fetch_int64(t, v, &iv))
                {
                    g->i0[j]++;
                    g->i1[j] = (int64_t)((uint64_t)g->i1[j] + (uint64_t)iv);
                }
                break;
            case KAGG_ACTION__PSUM_INT64:
            case KAGG_ACTION__PAVG_INT64:
                if (fetch_int64(t, v, &iv))
                {
                    g->i0[j]++;
                    g->s128[j] += iv;
                }
                break;
            case KAGG_ACTION__PSUM_FP:
            case KAGG_ACTION__PAVG_FP:
                if (fetch_fp64(t, v, &fv))
                {
                    g->i0[j]++;
                    g->d0[j] += fv;
                }
                break;
            case KAGG_ACTION__PSUM_NUMERIC:
            case KAGG_ACTION__PAVG_NUMERIC:
                /* __update_groupby__psum_numeric (cuda_gpupreagg.cu:1408-1446):
                 * finite values at the sum's scale; NaN and the infinities
                 * flag the state (the reference ORs the kind 1 of NaN into
                 * the scale bits instead of __PAGG_NUMERIC_ATTRS__NAN) */
                if (!v->isnull && v->nk == ON_VALID)
                {
                    g->i0[j]++;
                    g->s128[j] = (int128_t)((ouint128_t)g->s128[j] +
                        (ouint128_t)on_rescale(v->nv, v->nw, (int)(g->attrs[j] & __PAGG_NUMERIC_ATTRS__WEIGHT)));
                }
                else if (!v->isnull && v->nk != ON_VARLENA)
                    g->attrs[j] |= v->nk == ON_PINF ? __PAGG_NUMERIC_ATTRS__PINF
                                 : v->nk == ON_NINF ? __PAGG_NUMERIC_ATTRS__NINF
                                 : __PAGG_NUMERIC_ATTRS__NAN;
                break;
            case KAGG_ACTION__STDDEV:
                if (fetch_fp64(t, v, &fv))
                {
                    g->i0[j]++;
                    g->d0[j] += fv;
                    g->d1[j] += fv * fv;
                }
                break;
            default:
                break;
        }
    }
}

static void
group_merge(const oracle_session *os, ogroup *dst, const ogroup *src)
{
    const kern_expression *ka = os->gb_actions;
    for (int j = 0; j < ka->u.pagg.nattrs; j++)
    {
        switch (ka->u.pagg.desc[j].action)
        {
            case KAGG_ACTION__NROWS_ANY:
            case KAGG_ACTION__NROWS_COND:
                dst->i0[j] += src->i0[j];
                break;
            case KAGG_ACTION__PMIN_INT32:
            case KAGG_ACTION__PMIN_INT64:
                dst->attrs[j] |= src->attrs[j];
                if (src->i0[j] < dst->i0[j]) dst->i0[j] = src->i0[j];
                break;
            case KAGG_ACTION__PMAX_INT32:
            case KAGG_ACTION__PMAX_INT64:
                dst->attrs[j] |= src->attrs[j];
                if (src->i0[j] > dst->i0[j]) dst->i0[j] = src->i0[j];
                break;
            case KAGG_ACTION__PMIN_FP64:
                dst->attrs[j] |= src->attrs[j];
                if (src->d0[j] < dst->d0[j]) dst->d0[j] = src->d0[j];
                break;
            case KAGG_ACTION__PMAX_FP64:
                dst->attrs[j] |= src->attrs[j];
                if (src->d0[j] > dst->d0[j]) dst->d0[j] = src->d0[j];
                break;
            case KAGG_ACTION__PSUM_INT:
            case KAGG_ACTION__PAVG_INT:
                dst->i0[j] += src->i0[j];
                dst->i1[j] = (int64_t)((uint64_t)dst->i1[j] + (uint64_t)src->i1[j]);
                break;
            case KAGG_ACTION__PSUM_INT64:
            case KAGG_ACTION__PAVG_INT64:
                dst->i0[j] += src->i0[j];
                dst->s128[j] += src->s128[j];
                break;
            case KAGG_ACTION__PSUM_NUMERIC:
            case KAGG_ACTION__PAVG_NUMERIC:
                dst->i0[j] += src->i0[j];
                dst->s128[j] = (int128_t)((ouint128_t)dst->s128[j] + (ouint128_t)src->s128[j]);
                dst->attrs[j] |= src->attrs[j] & __PAGG_NUMERIC_ATTRS__MASK;
                break;
            case KAGG_ACTION__PSUM_FP:
            case KAGG_ACTION__PAVG_FP:
                dst->i0[j] += src->i0[j];
                dst->d0[j] += src->d0[j];
                break;
            case KAGG_ACTION__STDDEV:
                dst->i0[j] += src->i0[j];
                dst->d0[j] += src->d0[j];
                dst->d1[j] += src->d1[j];
                break;
        }
    }
}

/* ------------------------------------------------------------------
 * per-row pipeline: joins (depth-first) then projection or group-by
 * ------------------------------------------------------------------ */
typedef struct
{
    char       *buf;
    size_t      len;
    size_t      cap;
    uint32_t    nrows;
} obuf;

static void
obuf_append(obuf *b, const void *p, size_t n)
{
    if (b->len + n > b->cap)
    {
        b->cap = (b->cap + n) * 2;
        b->buf = realloc(b->buf, b->cap);
    }
    memcpy(b->buf + b->len, p, n);
    b->len += n;
}

typedef struct
{
    const oracle_session *os;
    const kern_multirels *kmrels;
    octx       *c;
    obuf       *out;        /* projection: packed tuples (rowid patched later) */
    ogtable    *gt;         /* group-by */
    uint32_t    nitems_out;
    uint32_t    depth_out[OMAX_DEPTH];
} orow;

static const kern_data_store *
inner_kds(const kern_multirels *km, int depth)
{
    uint64_t off = km->chunks[depth - 1].kds_offset;
    return off ? (const kern_data_store *)((const char *)km + off) : NULL;
}

static bool emit_row(orow *w);

static bool
join_depth(orow *w, int depth)
{
    const oracle_session *os = w->os;
    octx *c = w->c;

    if (depth > os->nrels)
        return emit_row(w);

    const kern_data_store *kds = inner_kds(w->kmrels, depth);
    const kern_expression *kh = os->hashv[depth];
    const kern_expression *kj = os->joinq[depth];
    bool left_outer = w->kmrels->chunks[depth - 1].left_outer;
    bool matched = false;
    od hv;
    /* slots are overwritten by deeper depths; save ours to restore */
    od saved[OMAX_SLOTS];

    if (!kh || !eval(c, kh, &hv))
        return kh ? false : oerror(c, ERRCODE_DEVICE_ERROR, "%s: no hash value", "join");
    if (depth == 1)
        c->fb_stage = false;    /* past the row's depth-0 work */
    uint32_t hash = (uint32_t)hv.i;
    const uint64_t *hslots = (const uint64_t *)((const char *)kds + KDS_HEAD_LENGTH(kds));
    uint64_t off = hslots[hash % kds->hash_nslots];
    memcpy(saved, c->slots, sizeof(od) * c->nslots);
    while (off != 0 && off != UINT64_MAX)
    {
        const kern_hashitem *hi = (const kern_hashitem *)((const char *)kds + kds->length - off);
        off = hi->next;
        if (hi->t.hash != hash)
            continue;
        if (!exec_loadvars_tuple(c, os->load[depth], kds, &hi->t))
            return false;
        int status = 1;
        if (kj)
        {
            /* ExecGpuJoinQuals (xpu_common.cu:1442-1481) */
            const kern_expression *karg = KEXP_FIRST_ARG(kj);
            for (uint32_t i = 0; i < kj->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                od q;
                if (status < 0 && (karg->expflags & KEXP_FLAG__IS_PUSHED_DOWN))
                    continue;
                if (!eval(c, karg, &q))
                    return false;
                if (q.isnull || !q.i)
                {
                    if (!(karg->expflags & KEXP_FLAG__IS_PUSHED_DOWN))
                    {
                        status = 0;
                        break;
                    }
                    status = -1;
                }
            }
        }
        if (status != 0)
            matched = true;
        if (status > 0)
        {
            w->depth_out[depth]++;
            if (!join_depth(w, depth + 1))
                return false;
            memcpy(c->slots, saved, sizeof(od) * c->nslots);
        }
    }
    if (left_outer && !matched)
    {
        /* null-extended row, filtered by the pushed-down quals only */
        bool ok = true;
        if (!exec_loadvars_tuple(c, os->load[depth], kds, NULL))
            return false;
        if (kj)
        {
            const kern_expression *karg = KEXP_FIRST_ARG(kj);
            for (uint32_t i = 0; i < kj->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
            {
                od q;
                if (!(karg->expflags & KEXP_FLAG__IS_PUSHED_DOWN))
                    continue;
                if (!eval(c, karg, &q))
                    return false;
                if (q.isnull || !q.i)
                {
                    ok = false;
                    break;
                }
            }
        }
        if (ok)
        {
            w->depth_out[depth]++;
            if (!join_depth(w, depth + 1))
                return false;
        }
        memcpy(c->slots, saved, sizeof(od) * c->nslots);
    }
    return true;
}

static bool
emit_row(orow *w)
{
    const oracle_session *os = w->os;
    octx *c = w->c;

    if (os->gb_actions)
    {
        od hv;
        uint32_t hash = 0;
        if (os->gb_keyhash)
        {
            if (!eval(c, os->gb_keyhash, &hv))
                return false;
            hash = (uint32_t)hv.i;
        }
        if (!run_actions_saves(os, c))
            return false;
        ogroup *g = group_lookup(os, w->gt, c, hash);
        group_update(os, g, c);
        w->nitems_out++;
        return true;
    }
    if (os->proj)
    {
        if (!prep_projection(os, c))
            return false;
        int sz = form_tuple(os, c, NULL, 0);
        if (sz < 0)
            return oerror(c, ERRCODE_DEVICE_ERROR, "%s: unable to form tuple", "projection");
        char tmp[XPU_MAXALIGN(sz) + 8];
        memset(tmp, 0, sizeof(tmp));
        form_tuple(os, c, tmp, 0);
        uint32_t len32 = (uint32_t)XPU_MAXALIGN(sz);
        obuf_append(w->out, &len32, 4);
        obuf_append(w->out, tmp, len32);
        w->out->nrows++;
        w->nitems_out++;
        return true;
    }
    w->nitems_out++;
    return true;
}

typedef struct
{
    obuf        out;
    obuf        fb;         /* CPU-fallback items ([u32 length][item]) */
    ogtable     gt;
    uint32_t    nitems_in;
    uint32_t    nitems_out;
    uint32_t    depth_out[OMAX_DEPTH];
    uint32_t    errcode;
    char        errmsg[KERN_ERRORBUF_MESSAGE_LEN + 1];
} othread;

static void
run_rows(const oracle_session *os, const kern_data_store *kds_src, const otuple *tups,
         const kern_multirels *kmrels, uint32_t row_begin, uint32_t row_end, othread *th)
{
    octx *c = malloc(sizeof(octx));
    orow w;

    octx_init(c, os);
    memset(&w, 0, sizeof(w));
    w.os = os;
    w.kmrels = kmrels;
    w.c = c;
    w.out = &th->out;
    w.gt = &th->gt;
    for (uint32_t row = row_begin; row < row_end; row++)
    {
        od q;
        for (int s = 0; s < c->nslots; s++)
        {
            memset(&c->slots[s], 0, sizeof(od));
            c->slots[s].isnull = true;
        }
        c->fb_stage = true;
        bool ok = tups ? exec_loadvars_heap(c, os->load[0], kds_src, &tups[row])
                       : exec_loadvars_arrow(c, os->load[0], kds_src, row);
        if (ok && os->scan_quals)
        {
            ok = eval(c, os->scan_quals, &q);
            if (ok && (q.isnull || !q.i))
                continue;
        }
        if (ok)
        {
            th->nitems_in++;
            ok = join_depth(&w, 1);
        }
        if (!ok)
        {
            /* HandleErrorIfCpuFallback (xpu_common.cu:1898-1955): a
             * fallback error of the row's depth-0 work sends the row to
             * the CPU with depth 0; anything else fails the task */
            if (c->errcode == ERRCODE_SUSPEND_FALLBACK && c->fb_stage && os->kds_fb_head)
            {
                int sz = form_fallback_tuple(os, c, NULL, 0);
                if (sz < 0)
                {
                    oerror(c, ERRCODE_DEVICE_ERROR, "%s", "unable to form the fallback tuple");
                    break;
                }
                uint32_t ilen = (uint32_t)XPU_MAXALIGN(16 + sz);
                char item[ilen];
                memset(item, 0, ilen);      /* l_state 0, depth 0, not matched */
                form_fallback_tuple(os, c, item + 16, 0);
                obuf_append(&th->fb, &ilen, 4);
                obuf_append(&th->fb, item, ilen);
                th->fb.nrows++;
                c->errcode = 0;
                c->errmsg[0] = '\0';
                continue;
            }
            break;
        }
    }
    th->nitems_out = w.nitems_out;
    memcpy(th->depth_out, w.depth_out, sizeof(w.depth_out));
    th->errcode = c->errcode;
    memcpy(th->errmsg, c->errmsg, sizeof(th->errmsg));
    free(c);
}

int
oracle_task_exec(oracle_session *os, const kern_data_store *kds_src,
                 const kern_multirels *kmrels,
                 kern_data_store *kds_dst, size_t dst_length,
                 int nthreads, kern_exec_results *results, kern_errorbuf *eb)
{
    uint32_t nitems = kds_src->nitems;
    otuple *tups = NULL;
    othread *th;
    int rc = 0;

    if (kds_src->format == KDS_FORMAT_BLOCK || kds_src->format == KDS_FORMAT_ROW ||
        kds_src->format == KDS_FORMAT_HASH)
    {
        int64_t n = collect_tuples(os->session, os->length, kds_src, &tups, eb);
        if (n < 0)
            return ERRCODE_DEVICE_ERROR;
        nitems = (uint32_t)n;
    }
    else if (kds_src->format != KDS_FORMAT_ARROW)
    {
        set_error(eb, ERRCODE_DEVICE_ERROR, __func__, "oracle reads KDS_FORMAT_ARROW, BLOCK, ROW, HASH");
        return ERRCODE_DEVICE_ERROR;
    }
    if (os->nrels > 0 && (!kmrels || (int)kmrels->num_inner_rels < os->nrels))
    {
        set_error(eb, ERRCODE_DEVICE_ERROR, __func__, "inner relations missing");
        return ERRCODE_DEVICE_ERROR;
    }
    if (nthreads < 1)
        nthreads = 1;
    th = calloc(nthreads, sizeof(othread));
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int t = 0; t < nthreads; t++)
    {
        uint64_t b = (uint64_t)nitems * t / nthreads;
        uint64_t e = (uint64_t)nitems * (t + 1) / nthreads;
        run_rows(os, kds_src, tups, kmrels, (uint32_t)b, (uint32_t)e, &th[t]);
    }
    if (results)
    {
        memset(results, 0, sizeof(kern_exec_results));
        results->num_rels = os->nrels;
    }
    for (int t = 0; t < nthreads; t++)
    {
        if (th[t].errcode && !rc)
        {
            rc = th[t].errcode;
            set_error(eb, rc, "oracle_task_exec", th[t].errmsg);
        }
        if (results)
        {
            results->nitems_raw = nitems;
            results->nitems_in += th[t].nitems_in;
            results->nitems_out += th[t].nitems_out;
            for (int d = 1; d <= os->nrels && d <= 8; d++)
                results->stats[d - 1].nitems_out += th[t].depth_out[d];
        }
    }
    /* CPU-fallback items, in input order, to the session's store */
    if (!rc && os->kds_fb_head)
    {
        for (int t = 0; t < nthreads; t++)
        {
            if (os->fb_len + th[t].fb.len > os->fb_cap)
            {
                os->fb_cap = (os->fb_len + th[t].fb.len) * 2 + 4096;
                os->fb_buf = realloc(os->fb_buf, os->fb_cap);
            }
            memcpy(os->fb_buf + os->fb_len, th[t].fb.buf, th[t].fb.len);
            os->fb_len += th[t].fb.len;
            os->fb_nrows += th[t].fb.nrows;
        }
    }
    /* projection: write kds_dst in input order */
    if (!rc && os->proj && os->gb_actions == NULL)
    {
        const kern_data_store *head = os->kds_dst_head;
        size_t head_sz = KDS_HEAD_LENGTH(head);
        uint64_t total_rows = 0, total_bytes = 0;
        for (int t = 0; t < nthreads; t++)
        {
            total_rows += th[t].out.nrows;
            total_bytes += th[t].out.len - 4ULL * th[t].out.nrows;
        }
        if (!kds_dst || head_sz + 8 * total_rows + total_bytes > dst_length)
        {
            rc = ERRCODE_SUSPEND_NO_SPACE;
            set_error(eb, rc, __func__, "kds_dst has no space");
        }
        else
        {
            uint64_t *rowidx = (uint64_t *)((char *)kds_dst + head_sz);
            uint64_t usage = 0;
            uint32_t rowid = 0;
            memcpy(kds_dst, head, head_sz);
            kds_dst->length = dst_length;
            for (int t = 0; t < nthreads; t++)
            {
                size_t pos = 0;
                while (pos < th[t].out.len)
                {
                    uint32_t len;
                    memcpy(&len, th[t].out.buf + pos, 4);
                    pos += 4;
                    usage += len;
                    char *dst = (char *)kds_dst + dst_length - usage;
                    memcpy(dst, th[t].out.buf + pos, len);
                    kern_tupitem *ti = (kern_tupitem *)dst;
                    memcpy(dst + ti->t_len - ROWID_SIZE, &rowid, ROWID_SIZE);
                    rowidx[rowid++] = usage;
                    pos += len;
                }
            }
            kds_dst->nitems = rowid;
            kds_dst->usage = usage;
        }
    }
    /* group-by: merge thread tables into the session table */
    if (!rc && os->gb_actions)
    {
        for (int t = 0; t < nthreads; t++)
        {
            octx *c = malloc(sizeof(octx));
            octx_init(c, os);
            for (size_t i = 0; i < th[t].gt.ngroups; i++)
            {
                ogroup *src = th[t].gt.groups[i];
                const kern_expression *ka = os->gb_actions;
                for (int j = 0; j < ka->u.pagg.nattrs; j++)
                    if (ka->u.pagg.desc[j].action == KAGG_ACTION__VREF)
                        c->slots[ka->u.pagg.desc[j].arg0_slot_id] = src->keys[j];
                ogroup *dst = group_lookup(os, &os->gt, c, src->hash);
                group_merge(os, dst, src);
            }
            free(c);
        }
    }
    for (int t = 0; t < nthreads; t++)
    {
        free(th[t].out.buf);
        free(th[t].fb.buf);
        for (size_t i = 0; i < th[t].gt.ngroups; i++)
        {
            free(th[t].gt.groups[i]->keybuf);
            free(th[t].gt.groups[i]);
        }
        free(th[t].gt.groups);
        free(th[t].gt.slots);
    }
    free(th);
    free(tups);
    return rc;
}

/* ------------------------------------------------------------------
 * final buffer: group tuples in the layout of __writeOutOneTuplePreAgg
 * (cuda_gpupreagg.cu:18-223), returned as KDS_FORMAT_ROW the way the
 * service writes kds_final back (gpu_service.c:2780-2818).
 * ------------------------------------------------------------------ */
static int
write_group_tuple(const oracle_session *os, const ogroup *g, char *buf)
{
    const kern_expression *ka = os->gb_actions;
    const kern_data_store *kf = os->kds_final_head;
    int nattrs = (int)kf->ncols < ka->u.pagg.nattrs ? (int)kf->ncols : ka->u.pagg.nattrs;
    uint32_t t_hoff = XPU_MAXALIGN(KERN_TUPITEM_T_BITS_OFFSET + XPU_BITMAPLEN(nattrs));
    uint16_t infomask = HEAP_HASNULL;
    kern_tupitem *t = (kern_tupitem *)buf;

    if (buf)
    {
        memset(buf, 0, t_hoff);
        t->t_infomask2 = nattrs & HEAP_NATTS_MASK;
        t->t_hoff = t_hoff + MINIMAL_TUPLE_OFFSET;
        t->hash = g->hash;
    }
    for (int j = 0; j < nattrs; j++)
    {
        const kern_aggregate_desc *d = &ka->u.pagg.desc[j];
        const kern_colmeta *cm = &kf->colmeta[j];
        uint32_t nxt = XPU_TYPEALIGN(cm->attalign, t_hoff);
        char *p = buf ? buf + nxt : NULL;
        int nbytes = 0;

        if (buf && nxt > t_hoff)
            memset(buf + t_hoff, 0, nxt - t_hoff);
        switch (d->action)
        {
            case KAGG_ACTION__VREF:
                /* keys never set HEAP_HASVARWIDTH (cuda_gpupreagg.cu:56-70) */
                if (!g->keys[j].isnull)
                    nbytes = datum_write(os->key_types[j], &g->keys[j], p);
                break;
            case KAGG_ACTION__NROWS_ANY:
            case KAGG_ACTION__NROWS_COND:
                nbytes = 8;
                if (p) memcpy(p, &g->i0[j], 8);
                break;
            case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMIN_INT64:
            case KAGG_ACTION__PMAX_INT32: case KAGG_ACTION__PMAX_INT64:
            {
                kagg_state__pminmax_int64_packed r;
                memset(&r, 0, sizeof(r));
                r.vl_len_ = (int32_t)(sizeof(r) << 2);
                r.attrs = g->attrs[j];
                r.value = g->i0[j];
                nbytes = sizeof(r);
                if (p) memcpy(p, &r, sizeof(r));
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            case KAGG_ACTION__PMIN_FP64: case KAGG_ACTION__PMAX_FP64:
            {
                kagg_state__pminmax_fp64_packed r;
                memset(&r, 0, sizeof(r));
                r.vl_len_ = (int32_t)(sizeof(r) << 2);
                r.attrs = g->attrs[j];
                r.value = g->d0[j];
                nbytes = sizeof(r);
                if (p) memcpy(p, &r, sizeof(r));
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
            {
                kagg_state__psum_int_packed r;
                memset(&r, 0, sizeof(r));
                r.vl_len_ = (int32_t)(sizeof(r) << 2);
                r.nitems = g->i0[j];
                r.sum = g->i1[j];
                nbytes = sizeof(r);
                if (p) memcpy(p, &r, sizeof(r));
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
            case KAGG_ACTION__PSUM_NUMERIC: case KAGG_ACTION__PAVG_NUMERIC:
            {
                kagg_state__psum_numeric_packed r;
                memset(&r, 0, sizeof(r));
                r.vl_len_ = (int32_t)(sizeof(r) << 2);
                r.attrs = g->attrs[j];
                r.nitems = (uint64_t)g->i0[j];
                r.sum.u64_lo = (uint64_t)g->s128[j];
                r.sum.u64_hi = (uint64_t)(g->s128[j] >> 64);
                nbytes = sizeof(r);
                if (p) memcpy(p, &r, sizeof(r));
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
            {
                kagg_state__psum_fp_packed r;
                memset(&r, 0, sizeof(r));
                r.vl_len_ = (int32_t)(sizeof(r) << 2);
                r.nitems = g->i0[j];
                r.sum = g->d0[j];
                nbytes = sizeof(r);
                if (p) memcpy(p, &r, sizeof(r));
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            case KAGG_ACTION__STDDEV:
            {
                kagg_state__stddev_packed r;
                memset(&r, 0, sizeof(r));
                r.vl_len_ = (int32_t)(sizeof(r) << 2);
                r.nitems = g->i0[j];
                r.sum_x = g->d0[j];
                r.sum_x2 = g->d1[j];
                nbytes = sizeof(r);
                if (p) memcpy(p, &r, sizeof(r));
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            default:
                return -1;
        }
        if (buf && nbytes > 0)
            t->t_bits[j >> 3] |= (1 << (j & 7));
        t_hoff = nxt + nbytes;
    }
    t_hoff += ROWID_SIZE;
    if (buf)
    {
        t->t_len = t_hoff;
        t->t_infomask = infomask;
    }
    return t_hoff;
}

/* ------------------------------------------------------------------
 * GPU-Sort of the final buffer: the comparator of cuda_gpusort.cu:15-343
 * over the row index, then the LIMIT of gpu_service.c:5099-5197 (the rows
 * past the limit leave the row index).  qsort is not stable; neither is the
 * reference's bitonic sort, so rows whose keys compare equal may come out
 * in either order.
 * ------------------------------------------------------------------ */
/* IEEE binary16 -> float, from its fields (exact: every half is a float) */
static float
oracle_half_to_float(uint16_t h)
{
    const int sgn = h >> 15, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    float v;
    if (e == 0)
        v = ldexpf((float)m, -24);                  /* zero, subnormal */
    else if (e == 31)
        v = m ? NAN : INFINITY;
    else
        v = ldexpf((float)(1024 + m), e - 25);
    return sgn ? -v : v;
}

static const kern_data_store *sort_kds;
static const kern_expression *sort_kexp;
static int sort_ncmp;           /* keys sort_comp_rows compares (a prefix for window rank) */

/* kern_fetch_minimal_tuple_attr: attribute anum of a kds_final tuple */
static const char *
sort_fetch_attr(const kern_data_store *kds, const kern_tupitem *t, int anum)
{
    const int natts = t->t_infomask2 & HEAP_NATTS_MASK;
    const bool hasnull = (t->t_infomask & HEAP_HASNULL) != 0;
    const uint8_t *bits = (const uint8_t *)t + KERN_TUPITEM_T_BITS_OFFSET;
    const char *base = (const char *)t;
    uint32_t off = (uint32_t)t->t_hoff - MINIMAL_TUPLE_OFFSET;

    if (anum < 1 || anum > natts)
        return NULL;
    for (int j = 0; j < anum; j++)
    {
        const kern_colmeta *cm = &kds->colmeta[j];
        if (hasnull && !(bits[j >> 3] & (1 << (j & 7))))
        {
            if (j == anum - 1)
                return NULL;
            continue;
        }
        if (cm->attlen > 0 || base[off] == 0)
            off = XPU_TYPEALIGN(cm->attalign > 0 ? cm->attalign : 1, off);
        if (j == anum - 1)
            return base + off;
        off += cm->attlen > 0 ? (uint32_t)cm->attlen : varsize_any((const uint8_t *)base + off);
    }
    return NULL;
}

/* x NULL / y NULL cases shared by every key kind */
static int
sort_comp_nulls(const kern_sortkey_desc *d, bool x_ok, bool y_ok)
{
    if (x_ok && !y_ok)
        return d->nulls_first ? 1 : -1;
    if (!x_ok && y_ok)
        return d->nulls_first ? -1 : 1;
    return 0;
}

#define SORT_CMP(a, b)  ((a) < (b) ? -1 : (a) > (b) ? 1 : 0)

static int64_t
sort_i64(const char *p) { int64_t v; memcpy(&v, p, 8); return v; }
static double
sort_f64(const char *p) { double v; memcpy(&v, p, 8); return v; }

/* a numeric sum state as (kind, value): kind 0 -inf, 1 finite, 2 +inf, 3 NaN
 * (NaN above every value, as numeric's comparison) */
static int
sort_comp_numeric(const char *x, const char *y)
{
    uint32_t ax, ay;
    memcpy(&ax, x + 4, 4);
    memcpy(&ay, y + 4, 4);
    int kx = (ax & __PAGG_NUMERIC_ATTRS__NAN) ? 3 : (ax & __PAGG_NUMERIC_ATTRS__PINF) ? 2
           : (ax & __PAGG_NUMERIC_ATTRS__NINF) ? 0 : 1;
    int ky = (ay & __PAGG_NUMERIC_ATTRS__NAN) ? 3 : (ay & __PAGG_NUMERIC_ATTRS__PINF) ? 2
           : (ay & __PAGG_NUMERIC_ATTRS__NINF) ? 0 : 1;
    if (kx != ky || kx != 1)
        return SORT_CMP(kx, ky);
    __int128 vx, vy;
    memcpy(&vx, x + 16, 16);
    memcpy(&vy, y + 16, 16);
    int wx = (int16_t)(ax & __PAGG_NUMERIC_ATTRS__WEIGHT);
    int wy = (int16_t)(ay & __PAGG_NUMERIC_ATTRS__WEIGHT);
    for (; wx < wy; wx++) vx *= 10;     /* same number of decimal digits */
    for (; wy < wx; wy++) vy *= 10;
    return SORT_CMP(vx, vy);
}

/* __gpusort_comp_keys (cuda_gpusort.cu:273-342) */
static int
sort_comp_rows(const void *a, const void *b)
{
    const uint64_t ra = *(const uint64_t *)a, rb = *(const uint64_t *)b;
    const char *tail = (const char *)sort_kds + sort_kds->length;
    const kern_tupitem *tx = (const kern_tupitem *)(tail - ra);
    const kern_tupitem *ty = (const kern_tupitem *)(tail - rb);

    for (int k = 0; k < sort_ncmp; k++)
    {
        const kern_sortkey_desc *d = &sort_kexp->u.sort.desc[k];
        const char *x = sort_fetch_attr(sort_kds, tx, d->src_anum);
        const char *y = sort_fetch_attr(sort_kds, ty, d->src_anum);
        bool x_ok = x != NULL, y_ok = y != NULL;
        int comp = 0;
        switch (d->kind)
        {
            case KSORT_KEY_KIND__PMINMAX_INT64:
            case KSORT_KEY_KIND__PMINMAX_FP64:
            {
                uint32_t fx = 0, fy = 0;
                if (x) memcpy(&fx, x + 4, 4);
                if (y) memcpy(&fy, y + 4, 4);
                x_ok = x_ok && (fx & __PAGG_MINMAX_ATTRS__VALID);
                y_ok = y_ok && (fy & __PAGG_MINMAX_ATTRS__VALID);
                break;
            }
            case KSORT_KEY_KIND__PSUM_INT64:
            case KSORT_KEY_KIND__PSUM_FP64:
            case KSORT_KEY_KIND__PSUM_NUMERIC:
            case KSORT_KEY_KIND__PAVG_INT64:
            case KSORT_KEY_KIND__PAVG_FP64:
                x_ok = x_ok && sort_i64(x + 8) != 0;
                y_ok = y_ok && sort_i64(y + 8) != 0;
                break;
            default:
                break;
        }
        if (!x_ok || !y_ok)
            comp = sort_comp_nulls(d, x_ok, y_ok);
        else
        {
            switch (d->kind)
            {
                case KSORT_KEY_KIND__VREF:
                    switch (d->key_type_code)
                    {
                        case TypeOpCode__bool: comp = SORT_CMP(*(const uint8_t *)x != 0, *(const uint8_t *)y != 0); break;
                        case TypeOpCode__int1: comp = SORT_CMP(*(const int8_t *)x, *(const int8_t *)y); break;
                        case TypeOpCode__int2: comp = SORT_CMP(*(const int16_t *)x, *(const int16_t *)y); break;
                        case TypeOpCode__int4:
                        case TypeOpCode__date: comp = SORT_CMP(*(const int32_t *)x, *(const int32_t *)y); break;
                        case TypeOpCode__int8: comp = SORT_CMP(sort_i64(x), sort_i64(y)); break;
                        case TypeOpCode__float4: comp = SORT_CMP(*(const float *)x, *(const float *)y); break;
                        case TypeOpCode__float8: comp = SORT_CMP(sort_f64(x), sort_f64(y)); break;
                        case TypeOpCode__float2:
                        {
                            /* half -> float, exact */
                            uint16_t hx = *(const uint16_t *)x, hy = *(const uint16_t *)y;
                            comp = SORT_CMP(oracle_half_to_float(hx), oracle_half_to_float(hy));
                            break;
                        }
                        default: comp = 0; break;
                    }
                    break;
                case KSORT_KEY_KIND__PMINMAX_INT64:
                    comp = SORT_CMP(sort_i64(x + 8), sort_i64(y + 8));
                    break;
                case KSORT_KEY_KIND__PMINMAX_FP64:
                    comp = SORT_CMP(sort_f64(x + 8), sort_f64(y + 8));
                    break;
                case KSORT_KEY_KIND__PSUM_INT64:
                    comp = SORT_CMP(sort_i64(x + 16), sort_i64(y + 16));
                    break;
                case KSORT_KEY_KIND__PSUM_FP64:
                    comp = SORT_CMP(sort_f64(x + 16), sort_f64(y + 16));
                    break;
                case KSORT_KEY_KIND__PSUM_NUMERIC:
                    comp = sort_comp_numeric(x, y);
                    break;
                case KSORT_KEY_KIND__PAVG_INT64:
                    /* __gpusort_prep_pavg_int64 (:401-429) */
                    comp = SORT_CMP((double)sort_i64(x + 16) / (double)sort_i64(x + 8),
                                    (double)sort_i64(y + 16) / (double)sort_i64(y + 8));
                    break;
                case KSORT_KEY_KIND__PAVG_FP64:
                    comp = SORT_CMP(sort_f64(x + 16) / (double)sort_i64(x + 8),
                                    sort_f64(y + 16) / (double)sort_i64(y + 8));
                    break;
                default:
                    comp = 0;
                    break;
            }
            if (!d->order_asc)
                comp = -comp;
        }
        if (comp != 0)
            return comp;
    }
    return 0;
}

/* rows ridx[x] and ridx[y] equal on the sort's first m keys */
static bool
sort_prefix_equal(const uint64_t *ridx, uint32_t x, uint32_t y, int m)
{
    const int save = sort_ncmp;
    sort_ncmp = m;
    const bool eq = sort_comp_rows(&ridx[x], &ridx[y]) == 0;
    sort_ncmp = save;
    return eq;
}

/* Window-rank filter of the sorted row index (kern_windowrank_exec_*,
 * cuda_gpusort.cu:814-987): the first window_partby_nkeys sort keys make
 * the partition, the next window_orderby_nkeys the peers; a row is kept
 * when its 0-based row_number / rank / dense_rank within its partition is
 * below window_rank_limit - 1 (unsigned, as there), i.e. rank() <
 * window_rank_limit.  The reference finds partition and peer boundaries
 * where the 32-bit hash of those keys changes (kern_windowrank_prep_hash,
 * :1170-1311); here they are where the keys change, which is the same
 * unless two adjacent distinct keys share a hash.  The kept entries move to
 * the front in sorted order; returns their number. */
static uint32_t
window_rank_filter(uint64_t *ridx, uint32_t n)
{
    const int np = sort_kexp->u.sort.window_partby_nkeys;
    const int no = np + sort_kexp->u.sort.window_orderby_nkeys;
    const char func = sort_kexp->u.sort.window_rank_func;
    const uint32_t lim = sort_kexp->u.sort.window_rank_limit - 1U;
    uint32_t part = 0, peer = 0, dense = 0, nkept = 0;
    uint64_t *kept = malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++)
    {
        if (i == 0 || !sort_prefix_equal(ridx, i - 1, i, np))
        {
            part = peer = i;
            dense = 0;
        }
        else if (!sort_prefix_equal(ridx, i - 1, i, no))
        {
            peer = i;
            dense++;
        }
        uint32_t r;
        if (func == KSORT_WINDOW_FUNC__ROW_NUMBER)
            r = i - part;
        else if (func == KSORT_WINDOW_FUNC__RANK)
            r = peer - part;
        else
            r = dense;
        if (r < lim)
            kept[nkept++] = ridx[i];
    }
    memcpy(ridx, kept, sizeof(uint64_t) * nkept);
    free(kept);
    return nkept;
}

/* The kept rows ridx[0..nout) packed at the buffer's new tail, each with
 * its row id set to its sorted position, as kern_buffer_simple_limit
 * (cuda_gpusort.cu:1401-1454, rowid at :1446) and kern_windowrank_finalize
 * (:1314-1356) pack them; returns the new length. */
static int64_t
pack_kept_rows(kern_data_store *kds_final, uint64_t need, size_t head_sz, uint64_t *rowidx,
               uint32_t nout)
{
    uint64_t kept = 0, pos = 0;
    for (uint32_t j = 0; j < nout; j++)
        kept += XPU_MAXALIGN(((kern_tupitem *)((char *)kds_final + need - rowidx[j]))->t_len);
    char *stage = calloc(1, kept ? kept : 1);
    for (uint32_t j = 0; j < nout; j++)
    {
        const kern_tupitem *t = (const kern_tupitem *)((char *)kds_final + need - rowidx[j]);
        uint32_t sz = XPU_MAXALIGN(t->t_len);
        pos += sz;
        memcpy(stage + kept - pos, t, t->t_len);
        memcpy(stage + kept - pos + t->t_len - ROWID_SIZE, &j, ROWID_SIZE);
        rowidx[j] = pos;
    }
    uint64_t new_len = head_sz + 8ULL * nout + kept;
    memcpy((char *)kds_final + new_len - kept, stage, kept);
    free(stage);
    kds_final->nitems = nout;
    kds_final->usage = kept;
    kds_final->length = new_len;
    return (int64_t)new_len;
}

int64_t
oracle_task_final(oracle_session *os, kern_data_store *kds_final, size_t final_length,
                  kern_errorbuf *eb)
{
    const kern_data_store *head = os->kds_final_head;
    size_t head_sz;
    uint64_t usage = 0, need = 0;

    if (!head || !os->gb_actions)
    {
        set_error(eb, ERRCODE_DEVICE_ERROR, __func__, "session has no group-by");
        return -1;
    }
    head_sz = KDS_HEAD_LENGTH(head);
    for (size_t i = 0; i < os->gt.ngroups; i++)
        need += XPU_TYPEALIGN(XPU_L1_CACHELINE_SZ,
                              offsetof(kern_hashitem, t) + write_group_tuple(os, os->gt.groups[i], NULL));
    need += head_sz + 8 * os->gt.ngroups;
    if (!kds_final)
        return (int64_t)need;
    if (need > final_length)
    {
        set_error(eb, ERRCODE_SUSPEND_NO_SPACE, __func__, "kds_final has no space");
        return -1;
    }
    memcpy(kds_final, head, head_sz);
    kds_final->format = KDS_FORMAT_ROW;
    kds_final->hash_nslots = 0;
    kds_final->length = need;
    uint64_t *rowidx = (uint64_t *)((char *)kds_final + head_sz);
    for (size_t i = 0; i < os->gt.ngroups; i++)
    {
        int sz = write_group_tuple(os, os->gt.groups[i], NULL);
        uint64_t unit = XPU_TYPEALIGN(XPU_L1_CACHELINE_SZ, offsetof(kern_hashitem, t) + sz);
        usage += unit;
        kern_hashitem *hi = (kern_hashitem *)((char *)kds_final + need - usage);
        memset(hi, 0, unit);
        write_group_tuple(os, os->gt.groups[i], (char *)&hi->t);
        uint32_t rowid = (uint32_t)i;
        memcpy((char *)&hi->t + hi->t.t_len - ROWID_SIZE, &rowid, ROWID_SIZE);
        rowidx[i] = (char *)kds_final + need - (char *)&hi->t;
    }
    kds_final->nitems = (uint32_t)os->gt.ngroups;
    kds_final->usage = usage;
    if (os->session->xpucode_gpusort_keydesc)
    {
        /* gpuservSortingFinalBuffer (gpu_service.c:5199-5345) */
        sort_kds = kds_final;
        sort_kexp = (const kern_expression *)((const char *)os->session +
                                              os->session->xpucode_gpusort_keydesc);
        sort_ncmp = sort_kexp->u.sort.nkeys;
        if (kds_final->nitems > 1)
            qsort(rowidx, kds_final->nitems, sizeof(uint64_t), sort_comp_rows);
        uint32_t nout = kds_final->nitems;
        if (sort_kexp->u.sort.window_rank_func)
            nout = window_rank_filter(rowidx, kds_final->nitems);
        else if (os->session->gpusort_limit_count > 0 &&
                 (uint64_t)os->session->gpusort_limit_count < kds_final->nitems)
            nout = (uint32_t)os->session->gpusort_limit_count;
        if (nout < kds_final->nitems || (sort_kexp->u.sort.window_rank_func && nout > 0))
            return pack_kept_rows(kds_final, need, head_sz, rowidx, nout);
    }
    return (int64_t)need;
}
