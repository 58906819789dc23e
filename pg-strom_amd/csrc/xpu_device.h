/*
 * xpu_device.h - device code of the MI355X (gfx950) execution path of
 * PG-Strom's xPU scan / hash-join / pre-aggregation pipeline.  It is compiled
 * twice: ahead of time into libxpuhip.so (xpu_kernels.hip, programs run by
 * the op interpreter) and at session time per plan (xpu_jit.cpp, the same
 * code with the plan baked in as constant data and the program emitted as
 * straight-line op calls, which the compiler folds).
 *
 * One persistent kernel per task (the role of kern_gpujoin_main,
 * cuda_gpujoin.cu:1050-1279), redesigned for CDNA4:
 *
 *  - Expressions run as compiled register programs (xpu_program.h) with a
 *    wave-uniform program counter; there are no device function pointers,
 *    no alloca'd kern_context and no per-depth kvec rings in global memory
 *    (cuda_gpujoin.cu:1093-1096): a row stays in one lane from the column
 *    load to the final stage, its values in VGPRs.
 *  - Hash joins walk the inner KDS_FORMAT_HASH chains lane-by-lane in a
 *    depth-first state machine (one outer row per lane, any number of inner
 *    matches per depth, LEFT OUTER null extension), cuda_gpujoin.cu:304-494.
 *  - Projection output is stream-compacted with a wave64 shuffle scan and a
 *    decoupled look-back across tiles, so rows land in kds_dst in input order
 *    and an overflow ('d') leaves a clean resume point at a tile boundary
 *    (the reference suspends whole blocks, cuda_gpujoin.cu:836-878).
 *  - Group-by keeps a per-block partial table in LDS keyed by the packed
 *    grouping keys, updated with LDS atomics, and merges it into the
 *    reference-format kds_final (KDS_FORMAT_HASH group tuples) with 64-bit
 *    global atomics once per block; rows that miss the LDS table update
 *    kds_final directly (cuda_gpupreagg.cu:1607-1731 semantics).
 */
#ifndef XPU_DEVICE_H
#define XPU_DEVICE_H
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpu_kernels.h"


/* The plan and the headers of the source and inner KDS are read-only while
 * a kernel runs.  Reading them through the constant address space lets the
 * backend use scalar (SMEM) loads for the uniform op and column-metadata
 * fetches even after the kernel's own stores and atomics, which would
 * otherwise force them onto the vector memory path. */
#if defined(__HIP_DEVICE_COMPILE__)
#define XCONST              __attribute__((address_space(4)))
#else
#define XCONST              /* the host pass only parses device code */
#endif
typedef const XCONST xplan cplan;
typedef const XCONST kern_data_store ckds;
typedef const XCONST kern_colmeta ccolmeta;

/* ------------------------------------------------------------------
 * small helpers
 * ------------------------------------------------------------------ */
#define RNULL(L, r)         (((L).nmask >> (r)) & 1U)
#define SETNULL(L, r, n)    ((L).nmask = ((L).nmask & ~(1U << (r))) | ((uint32_t)(n) << (r)))
/* length word of a text pair holding a compressed or external datum */
#define XTEXT_RAW           (-1)
/* kinds of a numeric register triple (the reference's XPU_NUMERIC_KIND__*) */
#define XNUM_VALID          0
#define XNUM_NAN            1
#define XNUM_PINF           2
#define XNUM_NINF           3
#define XNUM_VARLENA        0xff

/* The register file is a separate local array (a homogeneous alloca that
 * the backend keeps in VGPRs, indexed with uniform register numbers); a
 * struct holding it together with the scalars could not be split. */
struct Lane
{
    int64_t    *R;              /* -> int64_t[XP_NREGS] in the kernel frame */
    uint32_t    nmask;          /* bit r: R[r] is NULL */
    uint32_t    err;            /* message id + 1 of the first error in this lane */
    uint32_t    dead;           /* the row failed the scan quals (XO_KILL): it
                                   raises no errors in the final stage */
};

__device__ __forceinline__ double   as_f64(int64_t v) { return __longlong_as_double(v); }
__device__ __forceinline__ int64_t  f64_bits(double d) { return __double_as_longlong(d); }
__device__ __forceinline__ float    as_f32(int64_t v) { return __int_as_float((int)(uint32_t)v); }
__device__ __forceinline__ int64_t  f32_bits(float f) { return (int64_t)(uint32_t)__float_as_int(f); }

__device__ __forceinline__ uint64_t
xalign(uint64_t a, uint64_t x)
{
    return (x + a - 1) & ~(a - 1);
}

__device__ __forceinline__ bool
int_in_class(int cls, int64_t v)
{
    switch (cls)
    {
        case XT_BOOL:
        case XT_INT1: return v >= -128 && v <= 127;
        case XT_INT2: return v >= -32768 && v <= 32767;
        case XT_INT4: return v >= INT32_MIN && v <= INT32_MAX;
    }
    return true;
}

__device__ __forceinline__ int64_t
int_class_min(int cls)
{
    switch (cls)
    {
        case XT_BOOL:
        case XT_INT1: return -128;
        case XT_INT2: return -32768;
        case XT_INT4: return INT32_MIN;
    }
    return INT64_MIN;
}

__device__ __forceinline__ int64_t
wrap_class(int cls, int64_t v)
{
    switch (cls)
    {
        case XT_BOOL:
        case XT_INT1: return (int8_t)v;
        case XT_INT2: return (int16_t)v;
        case XT_INT4: return (int32_t)v;
    }
    return v;
}

/* NaN-aware float ordering (xpu_basetype.cu:467-478) */
__device__ __forceinline__ bool
fcmp(int op, double x, double y)
{
    bool xn = isnan(x), yn = isnan(y);
    switch (op)
    {
        case XC_EQ: return xn ? yn : (!yn && x == y);
        case XC_NE: return xn ? !yn : (yn || x != y);
        case XC_LT: return !xn && (yn || x < y);
        case XC_LE: return yn || (!xn && x <= y);
        case XC_GT: return !yn && (xn || x > y);
        case XC_GE: return xn || (!yn && x >= y);
    }
    return false;
}

__device__ __forceinline__ bool
cmp_by(int op, int c)
{
    switch (op)
    {
        case XC_EQ: return c == 0;
        case XC_NE: return c != 0;
        case XC_LT: return c < 0;
        case XC_LE: return c <= 0;
        case XC_GT: return c > 0;
        case XC_GE: return c >= 0;
    }
    return false;
}

/* ------------------------------------------------------------------
 * lookup3 hash with PostgreSQL's seed (xpu_common.cu:2920-3041).
 * Fixed-width keys of up to 8 bytes never reach the 12-byte mixing loop,
 * so their hash is the final() mix of the two assembled words.
 * ------------------------------------------------------------------ */
#define XROT(x, k) (((x) << (k)) | ((x) >> (32 - (k))))
#define XMIX(a, b, c) do { \
    a -= c; a ^= XROT(c, 4);  c += b; \
    b -= a; b ^= XROT(a, 6);  a += c; \
    c -= b; c ^= XROT(b, 8);  b += a; \
    a -= c; a ^= XROT(c, 16); c += b; \
    b -= a; b ^= XROT(a, 19); a += c; \
    c -= b; c ^= XROT(b, 4);  b += a; } while (0)
#define XFINAL(a, b, c) do { \
    c ^= b; c -= XROT(b, 14); \
    a ^= c; a -= XROT(c, 11); \
    b ^= a; b -= XROT(a, 25); \
    c ^= b; c -= XROT(b, 16); \
    a ^= c; a -= XROT(c, 4);  \
    b ^= a; b -= XROT(a, 14); \
    c ^= b; c -= XROT(b, 24); } while (0)

__device__ __forceinline__ uint32_t
hash_fixed(uint64_t v, int len)
{
    uint32_t a, b, c;
    a = b = c = 0x9e3779b9U + (uint32_t)len + 3923095U;
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    if (len < 4)
        lo &= (1U << (8 * len)) - 1;
    if (len <= 4)
        hi = 0;
    else if (len < 8)
        hi &= (1U << (8 * (len - 4))) - 1;
    a += lo;
    b += hi;
    XFINAL(a, b, c);
    return c;
}

__device__ uint32_t
hash_bytes(const uint8_t *k, uint32_t len)
{
    uint32_t a, b, c;
    a = b = c = 0x9e3779b9U + len + 3923095U;
    while (len >= 12)
    {
        a += (uint32_t)k[0] | ((uint32_t)k[1] << 8) | ((uint32_t)k[2] << 16) | ((uint32_t)k[3] << 24);
        b += (uint32_t)k[4] | ((uint32_t)k[5] << 8) | ((uint32_t)k[6] << 16) | ((uint32_t)k[7] << 24);
        c += (uint32_t)k[8] | ((uint32_t)k[9] << 8) | ((uint32_t)k[10] << 16) | ((uint32_t)k[11] << 24);
        XMIX(a, b, c);
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
    XFINAL(a, b, c);
    return c;
}

__device__ __forceinline__ uint32_t
hash_merge(uint32_t prev, uint32_t next)
{
    return ((prev >> 3) | (prev << 29)) ^ next;
}

/* ------------------------------------------------------------------
 * minimal-tuple attribute fetch (xpu_common.cu:60-130 deform rules)
 * ------------------------------------------------------------------ */
__device__ __forceinline__ void
ldtup(Lane &L, int d, const kern_tupitem *t, const kern_data_store *kds, int resno, int type)
{
    if (!t)
    {
        SETNULL(L, d, 1);
        return;
    }
    uint16_t infomask2 = t->t_infomask2;
    uint16_t infomask = t->t_infomask;
    int ncols = infomask2 & HEAP_NATTS_MASK;
    bool hasnull = (infomask & HEAP_HASNULL) != 0;
    const char *base = (const char *)t;
    uint32_t pos = (uint32_t)t->t_hoff - MINIMAL_TUPLE_OFFSET;

    if (resno > ncols || resno > (int)kds->ncols)
    {
        SETNULL(L, d, 1);
        return;
    }
    for (int j = 0; j < resno; j++)
    {
        ccolmeta *cm = &((ckds *)(uintptr_t)kds)->colmeta[j];
        int attlen = cm->attlen;
        if (hasnull && !((t->t_bits[j >> 3] >> (j & 7)) & 1))
        {
            if (j == resno - 1)
            {
                SETNULL(L, d, 1);
                return;
            }
            continue;
        }
        if (attlen > 0 || base[pos] == 0)
            pos = (uint32_t)xalign(cm->attalign, pos);
        const char *p = base + pos;
        if (j == resno - 1)
        {
            int64_t v = 0;
            switch (type)
            {
                case TypeOpCode__bool:
                case TypeOpCode__int1: v = *(const int8_t *)p; break;
                case TypeOpCode__int2: v = *(const int16_t *)p; break;
                case TypeOpCode__int4:
                case TypeOpCode__date: v = *(const int32_t *)p; break;
                case TypeOpCode__float4: v = (int64_t)*(const uint32_t *)p; break;
                case TypeOpCode__int8:
                case TypeOpCode__float8: v = *(const int64_t *)p; break;
                case TypeOpCode__numeric:
                {
                    /* kept in its varlena form, addressed by its header */
                    const uint8_t b0 = (uint8_t)p[0];
                    L.R[d] = (int64_t)(uintptr_t)p;
                    L.R[d + 1] = (b0 == 0x01 || (b0 & 0x03) == 0x02) ? XTEXT_RAW : 0;
                    L.R[d + 2] = XNUM_VARLENA;
                    SETNULL(L, d, 0);
                    return;
                }
                default:
                {
                    uint8_t b0 = (uint8_t)p[0];
                    if (b0 == 0x01 || (b0 & 0x03) == 0x02)
                    {
                        if (!L.err)
                            L.err = XMSG_TEXT_EXTERNAL + 1;
                        SETNULL(L, d, 1);
                        return;
                    }
                    if (b0 & 0x01)
                    {
                        L.R[d] = (int64_t)(uintptr_t)(p + 1);
                        L.R[d + 1] = ((b0 >> 1) & 0x7f) - 1;
                    }
                    else
                    {
                        uint32_t h = *(const uint32_t *)p;
                        L.R[d] = (int64_t)(uintptr_t)(p + 4);
                        L.R[d + 1] = (int64_t)((h >> 2) & 0x3fffffff) - 4;
                    }
                    SETNULL(L, d, 0);
                    return;
                }
            }
            L.R[d] = v;
            SETNULL(L, d, 0);
            return;
        }
        if (attlen > 0)
            pos += attlen;
        else
        {
            uint8_t b0 = (uint8_t)p[0];
            if (b0 == 0x01)
            {
                if (!L.err)
                    L.err = XMSG_TEXT_EXTERNAL + 1;
                SETNULL(L, d, 1);
                return;
            }
            if (b0 & 0x01)
                pos += (b0 >> 1) & 0x7f;
            else
                pos += ((*(const uint32_t *)p) >> 2) & 0x3fffffff;
        }
    }
    SETNULL(L, d, 1);
}

/* ------------------------------------------------------------------
 * text helpers of the string operators
 * ------------------------------------------------------------------ */
/* bytes of the character at s: pg_utf8_mblen (xpu_textlib.cu:774-787), or 1
 * for a single-byte database encoding */
__device__ __forceinline__ int
txt_mblen(const uint8_t *s, bool utf8)
{
    if (!utf8)
        return 1;
    const uint8_t c = *s;
    if ((c & 0x80) == 0) return 1;
    if ((c & 0xe0) == 0xc0) return 2;
    if ((c & 0xf0) == 0xe0) return 3;
    if ((c & 0xf8) == 0xf0) return 4;
    return 1;
}

/* length without trailing spaces (bpchar_truelen, xpu_textlib.cu:59-69) */
__device__ __forceinline__ int64_t
txt_truelen(const uint8_t *s, int64_t len)
{
    while (len > 0 && s[len - 1] == ' ')
        len--;
    return len;
}

__device__ __forceinline__ int
txt_bytecmp(const uint8_t *s1, int64_t l1, const uint8_t *s2, int64_t l2)
{
    const int64_t n = l1 < l2 ? l1 : l2;
    int c = 0;
    for (int64_t i = 0; i < n && c == 0; i++)
        c = (int)s1[i] - (int)s2[i];
    /* the loop's exit value stays a per-lane integer: where the caller only
     * tests c != 0 (bpcharne specialised for a plan), gfx950 code
     * generation otherwise carried "bytes differ" out of the loop in a
     * lane mask written by the last iteration, in which lanes that had
     * already left the loop read 0 */
    asm volatile("" : "+v"(c));
    if (c == 0)
        c = l1 < l2 ? -1 : (l1 > l2 ? 1 : 0);
    return c;
}

__device__ __forceinline__ uint8_t
like_fold(uint8_t c, bool ic)
{
    return (ic && c >= 'a' && c <= 'z') ? (uint8_t)(c - ('a' - 'A')) : c;
}

/* states of like_match's walk */
enum { XLK_UNDECIDED = -2, XLK_STEP = 0, XLK_PERCENT, XLK_MISMATCH, XLK_SCAN };

/* LIKE / ILIKE with the result of GenericMatchText (xpu_textlib.cu:915-1085)
 * and its error cases, without recursion: the newest % is the only point a
 * mismatch resumes from, since the reference's recursive search answers
 * "abort" rather than "false" once a deeper % has scanned to the end of the
 * text.  Returns 1 / 0, or -1 for an invalid escape.  One flat loop over the
 * walk's states: each step advances the text or pattern position or the
 * resume point, or moves to the next state, and a match attempt starts at
 * most once per text position, so (tlen + 2)(plen + 2) times four bounds
 * the steps.  (A nesting of loops with early returns for the same walk gave
 * wrong answers on gfx950 for lanes whose wave mates took other paths.) */
__device__ int
like_match(const uint8_t *t, int64_t tlen, const uint8_t *p, int64_t plen, bool ic, bool utf8)
{
    if (plen == 1 && p[0] == '%')
        return 1;
    int64_t ti = 0, pi = 0, bt_t = -1, bt_p = -1;
    uint8_t firstpat = 0;
    int state = XLK_STEP, res = XLK_UNDECIDED;
    const int64_t limit = 4 * (tlen + 2) * (plen + 2);
    for (int64_t iter = 0; res == XLK_UNDECIDED; iter++)
    {
        if (iter > limit)
            res = 0;
        else if (state == XLK_STEP)
        {
            if (ti < tlen && pi < plen)
            {
                const uint8_t c = p[pi];
                if (c == '\\')
                {
                    if (pi + 1 >= plen)
                        res = -1;
                    else if (like_fold(p[pi + 1], ic) != like_fold(t[ti], ic))
                        state = XLK_MISMATCH;
                    else
                    {
                        ti++;
                        pi += 2;
                    }
                }
                else if (c == '%')
                {
                    pi++;
                    state = XLK_PERCENT;
                }
                else if (c == '_')
                {
                    ti += txt_mblen(t + ti, utf8);
                    pi++;
                }
                else if (like_fold(c, ic) != like_fold(t[ti], ic))
                    state = XLK_MISMATCH;
                else
                {
                    ti++;
                    pi++;
                }
            }
            else if (ti < tlen)
                state = XLK_MISMATCH;
            else if (pi < plen && p[pi] == '%')
                pi++;
            else
                res = pi >= plen ? 1 : 0;
        }
        else if (state == XLK_PERCENT)
        {
            /* after a %: further % and _ (each _ takes a character), then
             * the first pattern character a match may start with */
            if (pi >= plen)
                res = 1;
            else if (p[pi] == '%')
                pi++;
            else if (p[pi] == '_')
            {
                if (ti >= tlen)
                    res = 0;
                else
                {
                    ti += txt_mblen(t + ti, utf8);
                    pi++;
                }
            }
            else if (p[pi] == '\\' && plen - pi < 2)
                res = -1;
            else
            {
                firstpat = like_fold(p[pi] == '\\' ? p[pi + 1] : p[pi], ic);
                bt_p = pi;
                bt_t = ti;
                state = XLK_SCAN;
            }
        }
        else if (state == XLK_MISMATCH)
        {
            if (bt_p < 0)
                res = 0;
            else
            {
                bt_t += txt_mblen(t + bt_t, utf8);
                state = XLK_SCAN;
            }
        }
        else
        {
            /* XLK_SCAN: the next text position the newest % can resume at */
            if (bt_t >= tlen)
                res = 0;
            else if (like_fold(t[bt_t], ic) != firstpat)
                bt_t += txt_mblen(t + bt_t, utf8);
            else
            {
                ti = bt_t;
                pi = bt_p;
                state = XLK_STEP;
            }
        }
    }
    return res;
}

/* __substring_common (xpu_textlib.cu:1159-1259) on 0-based start and count
 * (-1: to the end): the byte range (*off, *len) of str */
__device__ __forceinline__ void
substr_range(const uint8_t *str, int64_t length, int32_t start, int32_t count, bool utf8,
             int64_t *off, int64_t *len)
{
    *off = 0;
    if (!utf8)
    {
        if (start >= length || count == 0 || (count > 0 && (int64_t)start + count <= 0))
            *len = 0;
        else if (start < 0)
            *len = (count < 0 || (int64_t)start + count >= length) ? length : (int64_t)start + count;
        else
        {
            const int64_t end = (count < 0 || (int64_t)start + count > length) ? length
                                                                                : (int64_t)start + count;
            *off = start;
            *len = end - start;
        }
        return;
    }
    if (start < 0)
    {
        if (count < 0)
        {
            *len = length;
            return;
        }
        count = start + count;
        if (count < 0)
        {
            *len = 0;
            return;
        }
        start = 0;
    }
    int64_t pos = 0, val = -1;
    for (int64_t i = 0; pos < length; i++)
    {
        const int w = txt_mblen(str + pos, true);
        if (pos + w > length)
            break;
        if (i == start)
            val = pos;
        if (count >= 0 && i == (int64_t)start + count)
            break;
        pos += w;
    }
    if (val >= 0)
    {
        *off = val;
        *len = pos - val;
    }
    else
        *len = 0;
}

/* a varlena image's payload (1- or 4-byte header) */
__device__ __forceinline__ const uint8_t *
varlena_payload(const uint8_t *v, int64_t *len)
{
    if (v[0] & 0x01)
    {
        *len = ((v[0] >> 1) & 0x7f) - 1;
        return v + 1;
    }
    uint32_t h;
    __builtin_memcpy(&h, v, 4);
    *len = (int64_t)((h >> 2) & 0x3fffffff) - 4;
    return v + 4;
}

/* a fixed-width image of 1/2/4/8 bytes, sign-extended (float bits kept) */
__device__ __forceinline__ int64_t
load_fixed(const uint8_t *v, int width)
{
    switch (width)
    {
        case 1: return (int8_t)v[0];
        case 2: { int16_t x; __builtin_memcpy(&x, v, 2); return x; }
        case 4: { int32_t x; __builtin_memcpy(&x, v, 4); return x; }
        default: { int64_t x; __builtin_memcpy(&x, v, 8); return x; }
    }
}

/* ------------------------------------------------------------------
 * the register-program interpreter (one inlined copy per kernel)
 * ------------------------------------------------------------------ */
#define RAISE(g, m) do { if (!RNULL(L, g) && L.R[g] != 0 && !L.err && !L.dead) \
                                L.err = (uint32_t)(m) + 1; } while (0)

/* The NaN that x86 SSE arithmetic gives, where the reference's float
 * results are computed (its pgfn_float* code on the backend's CPU): the
 * first NaN operand, quieted, else the default NaN of an invalid operation,
 * whose sign is set.  CDNA's own NaN results can differ in sign: it
 * subtracts by adding the negated operand, and its default NaN is
 * positive. */
__device__ __forceinline__ double
nan_as_x86_f64(double x, double y, double v)
{
    if (!__builtin_isnan(v))
        return v;
    uint64_t b = 0xfff8000000000000ULL;
    if (__builtin_isnan(x))
        b = (uint64_t)__double_as_longlong(x) | 0x0008000000000000ULL;
    else if (__builtin_isnan(y))
        b = (uint64_t)__double_as_longlong(y) | 0x0008000000000000ULL;
    return __longlong_as_double((long long)b);
}

__device__ __forceinline__ float
nan_as_x86_f32(float x, float y, float v)
{
    if (!__builtin_isnan(v))
        return v;
    uint32_t b = 0xffc00000U;
    if (__builtin_isnan(x))
        b = (uint32_t)__float_as_int(x) | 0x00400000U;
    else if (__builtin_isnan(y))
        b = (uint32_t)__float_as_int(y) | 0x00400000U;
    return __int_as_float((int)b);
}

/* A text register pair (R[r], R[r + 1]) with a negative length holds a
 * compressed or external varlena as it is, R[r] addressing its header (the
 * reference's xpu_text_t with length -1).  Projection copies it unchanged;
 * an op that needs its bytes raises "text datum is compressed or external",
 * which the session's CPU fallback takes (xpu_textlib.h:52-58). */
#define TXT_RAW(r)          (L.R[(r) + 1] < 0)

/* VARSIZE_ANY of a varlena header (1B_E: the TOAST pointer's size by tag) */
__device__ __forceinline__ uint32_t
varsize_any(const uint8_t *p)
{
    const uint8_t b0 = p[0];
    if (b0 == 0x01)
        return XPU_VARSIZE_EXTERNAL(p[1]);
    if (b0 & 0x01)
        return b0 >> 1;
    return (((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
             ((uint32_t)p[3] << 24)) >> 2) & 0x3fffffff;
}

/* ------------------------------------------------------------------
 * numeric (xpu_numeric.h, xpu_numeric.cu).  A numeric value occupies a
 * register triple:
 *   R[r], R[r + 1]  a finite value's 128-bit integer (low, high word), or
 *                   for a datum still in its varlena form the address of
 *                   its header in R[r] and R[r + 1] < 0 when that datum is
 *                   compressed or external;
 *   R[r + 2]        the kind (low byte, XNUM_*) and, for a finite value,
 *                   its 16-bit decimal scale in bits 16..31: the value is
 *                   R[r] * 10^-scale (the reference's "weight").
 * Datums are parsed where an operator needs their value, as the reference
 * validates its xpu_numeric_t; one that is only moved keeps its varlena
 * form and is written back byte for byte (xpu_numeric_datum_write).
 * ------------------------------------------------------------------ */
typedef __int128 xi128;
typedef unsigned __int128 xu128;

struct xnum
{
    xi128   v;
    int     kind;
    int16_t scale;
};

/* NUMERIC_NORMALIZE (xpu_numeric.h:165-207): trailing zeros leave the
 * value while the scale stays non-negative; zero has scale 0 */
__device__ __forceinline__ void
num_normalize(xnum &x)
{
    if (x.v == 0)
    {
        x.scale = 0;
        return;
    }
    while (x.scale > 0 && x.v % 10 == 0)
    {
        x.v /= 10;
        x.scale = (int16_t)(x.scale - 1);
    }
}

__device__ __forceinline__ xi128 num_mul10(xi128 v) { return (xi128)((xu128)v * 10); }

/* __decimal_from_varlena (xpu_numeric.h:210-273); returns a message id + 1
 * on error.  A compressed or external datum goes to the CPU fallback. */
static __device__ __noinline__ uint32_t
num_parse(const uint8_t *hdr, xnum *x)
{
    int64_t len;
    if (hdr[0] == 0x01 || (hdr[0] & 0x03) == 0x02)
        return XMSG_TEXT_EXTERNAL + 1;
    const uint8_t *p = varlena_payload(hdr, &len);
    if (len < 2)
        return XMSG_NUM_HEADER + 1;
    const uint16_t n_head = (uint16_t)(p[0] | (p[1] << 8));
    x->v = 0;
    x->scale = 0;
    if ((n_head & 0xC000) == 0xC000)
    {
        if (n_head == 0xC000)
            x->kind = XNUM_NAN;
        else if (n_head == 0xD000)
            x->kind = XNUM_PINF;
        else if (n_head == 0xF000)
            x->kind = XNUM_NINF;
        else
            return XMSG_NUM_SPECIAL + 1;
        return 0;
    }
    const bool is_short = (n_head & 0xC000) == 0x8000;
    const int hsz = is_short ? 2 : 4;
    if (len < hsz)
        return XMSG_NUM_HEADER + 1;
    int weight;
    bool neg;
    if (is_short)
    {
        weight = n_head & 0x003F;
        if (n_head & 0x0040)
            weight |= ~0x003F;
        neg = (n_head & 0x2000) != 0;
    }
    else
    {
        weight = (int16_t)(p[2] | (p[3] << 8));
        neg = (n_head & 0xC000) == 0x4000;
    }
    const int ndigits = (int)((len - hsz) / 2);
    xi128 v = 0;
    for (int i = 0; i < ndigits; i++)
    {
        if ((v >> 114) != 0)
            return XMSG_NUM_RANGE + 1;
        const int16_t dig = (int16_t)(p[hsz + 2 * i] | (p[hsz + 2 * i + 1] << 8));
        v = v * 10000 + dig;
    }
    if (neg)
        v = -v;
    x->kind = XNUM_VALID;
    x->v = v;
    x->scale = (int16_t)(4 * (ndigits - (int16_t)(weight + 1)));
    num_normalize(*x);
    return 0;
}

/* the value of numeric register triple r; a message id + 1 on error */
__device__ __forceinline__ uint32_t
num_read(const Lane &L, int r, xnum *x)
{
    const int kind = (int)(L.R[r + 2] & 0xff);
    if (kind == XNUM_VARLENA)
    {
        if (L.R[r + 1] < 0)
            return XMSG_TEXT_EXTERNAL + 1;
        return num_parse((const uint8_t *)(uintptr_t)L.R[r], x);
    }
    x->kind = kind;
    x->scale = (int16_t)(L.R[r + 2] >> 16);
    x->v = (xi128)(((xu128)(uint64_t)L.R[r + 1] << 64) | (uint64_t)L.R[r]);
    return 0;
}

__device__ __forceinline__ void
num_store(Lane &L, int d, const xnum &x)
{
    const xi128 v = x.kind == XNUM_VALID ? x.v : 0;
    L.R[d] = (int64_t)(uint64_t)(xu128)v;
    L.R[d + 1] = (int64_t)(uint64_t)((xu128)v >> 64);
    L.R[d + 2] = (int64_t)((uint32_t)x.kind | ((uint32_t)(uint16_t)(x.kind == XNUM_VALID ? x.scale : 0) << 16));
}

/* a numeric register triple holding the varlena at hdr (raw: compressed or
 * external) */
__device__ __forceinline__ void
num_store_varlena(Lane &L, int d, const uint8_t *hdr, bool raw)
{
    L.R[d] = (int64_t)(uintptr_t)hdr;
    L.R[d + 1] = raw ? XTEXT_RAW : 0;
    L.R[d + 2] = XNUM_VARLENA;
}

/* align two finite values to the larger scale (the loops of
 * __decimal_add / __numeric_compare) */
__device__ __forceinline__ void
num_align(xnum &a, xnum &b)
{
    while (a.scale > b.scale)
    {
        b.v = num_mul10(b.v);
        b.scale = (int16_t)(b.scale + 1);
    }
    while (a.scale < b.scale)
    {
        a.v = num_mul10(a.v);
        a.scale = (int16_t)(a.scale + 1);
    }
}

/* __numeric_compare (xpu_numeric.cu:478-539) */
static __device__ __noinline__ int
num_compare(xnum a, xnum b)
{
    if (a.kind != XNUM_VALID)
    {
        if (a.kind == XNUM_NAN)
            return b.kind == XNUM_NAN ? 0 : 1;
        if (a.kind == XNUM_PINF)
            return b.kind == XNUM_NAN ? -1 : (b.kind == XNUM_PINF ? 0 : 1);
        return b.kind == XNUM_NINF ? 0 : -1;
    }
    if (b.kind != XNUM_VALID)
        return b.kind == XNUM_NINF ? 1 : -1;
    if ((a.v > 0 && b.v <= 0) || (a.v == 0 && b.v < 0))
        return 1;
    if ((b.v > 0 && a.v <= 0) || (b.v == 0 && a.v < 0))
        return -1;
    num_align(a, b);
    return a.v > b.v ? 1 : (a.v < b.v ? -1 : 0);
}

__device__ __forceinline__ bool num_is_neg(const xnum &x) { return x.kind == XNUM_VALID && x.v < 0; }
__device__ __forceinline__ bool num_is_pos(const xnum &x) { return x.kind == XNUM_VALID && x.v > 0; }

/* __decimal_add / __decimal_sub / __decimal_mul / __decimal_div
 * (xpu_numeric.h:443-776).  As there, add, sub and mul take the finite
 * path when either operand is finite (a NaN or infinity then counts as 0),
 * and numeric_mod runs the division (xpu_numeric.cu:681-702). */
static __device__ __noinline__ uint32_t
num_arith(int op, xnum a, xnum b, xnum *r)
{
    r->scale = 0;
    r->v = 0;
    if (op == XO_NUM_ADD || op == XO_NUM_SUB)
    {
        if (a.kind == XNUM_VALID || b.kind == XNUM_VALID)
        {
            num_align(a, b);
            r->kind = XNUM_VALID;
            r->scale = a.scale;
            r->v = (xi128)(op == XO_NUM_ADD ? (xu128)a.v + (xu128)b.v : (xu128)a.v - (xu128)b.v);
            num_normalize(*r);
            return 0;
        }
        const bool sub = (op == XO_NUM_SUB);
        if (a.kind == XNUM_NAN || b.kind == XNUM_NAN)
            r->kind = XNUM_NAN;
        else if (a.kind == XNUM_PINF)
            r->kind = (b.kind == (sub ? XNUM_PINF : XNUM_NINF)) ? XNUM_NAN : XNUM_PINF;
        else if (a.kind == XNUM_NINF)
            r->kind = (b.kind == (sub ? XNUM_NINF : XNUM_PINF)) ? XNUM_NAN : XNUM_NINF;
        else if (b.kind == XNUM_PINF)
            r->kind = sub ? XNUM_NINF : XNUM_PINF;
        else
            r->kind = sub ? XNUM_PINF : XNUM_NINF;
        return 0;
    }
    if (op == XO_NUM_MUL)
    {
        if (a.kind == XNUM_VALID || b.kind == XNUM_VALID)
        {
            r->kind = XNUM_VALID;
            r->scale = (int16_t)(a.scale + b.scale);
            r->v = (xi128)((xu128)a.v * (xu128)b.v);
            num_normalize(*r);
            return 0;
        }
        const xnum &x = (a.kind == XNUM_PINF || a.kind == XNUM_NINF) && b.kind != XNUM_NAN ? b : a;
        const bool a_inf = (a.kind == XNUM_PINF || a.kind == XNUM_NINF) && b.kind != XNUM_NAN;
        const bool b_inf = !a_inf && a.kind != XNUM_NAN && (b.kind == XNUM_PINF || b.kind == XNUM_NINF);
        if (a_inf || b_inf)
        {
            const int inf = a_inf ? a.kind : b.kind;
            const xnum &o = a_inf ? b : a;
            (void)x;
            const bool oneg = o.kind == XNUM_NINF || num_is_neg(o);
            const bool opos = o.kind == XNUM_PINF || num_is_pos(o);
            if (oneg)
                r->kind = (inf == XNUM_PINF) ? XNUM_NINF : XNUM_PINF;
            else if (opos)
                r->kind = inf;
            else
                r->kind = XNUM_NAN;
        }
        else
            r->kind = XNUM_NAN;
        return 0;
    }
    /* division, and numeric_mod */
    if (a.kind == XNUM_VALID && b.kind == XNUM_VALID)
    {
        xi128 rem = a.v, div = b.v, ival = 0;
        int16_t weight = (int16_t)(a.scale - b.scale);
        bool negative = false;
        if (div == 0)
            return XMSG_NUM_DIV0 + 1;
        if (rem < 0)
        {
            rem = -rem;
            if (div < 0)
                div = -div;
            else
                negative = true;
        }
        else if (div < 0)
        {
            negative = true;
            div = -div;
        }
        for (;;)
        {
            const xi128 q = rem / div;
            ival = (xi128)((xu128)ival * 10 + (xu128)q);
            rem -= q * div;
            if (rem == 0 || (ival >> 50) != 0)
                break;
            rem = num_mul10(rem);
            weight = (int16_t)(weight + 1);
        }
        if (negative)
            ival = -ival;
        r->kind = XNUM_VALID;
        r->scale = weight;
        r->v = ival;
        num_normalize(*r);
        return 0;
    }
    if (a.kind == XNUM_NAN || b.kind == XNUM_NAN)
        r->kind = XNUM_NAN;
    else if (a.kind == XNUM_PINF || a.kind == XNUM_NINF)
    {
        const bool apos = (a.kind == XNUM_PINF);
        if (b.kind == XNUM_PINF)
            r->kind = apos ? XNUM_PINF : XNUM_NINF;
        else if (b.kind == XNUM_NINF)
            r->kind = apos ? XNUM_NINF : XNUM_PINF;
        else if (b.v > 0)
            r->kind = apos ? XNUM_PINF : XNUM_NINF;
        else if (b.v < 0)
            r->kind = apos ? XNUM_NINF : XNUM_PINF;
        else
            return XMSG_NUM_DIV0 + 1;
    }
    else
        r->kind = XNUM_VALID;   /* finite / infinite: 0 */
    return 0;
}

/* __xpu_numeric_to_int64 (xpu_numeric.cu:250-295): rounded at the first
 * fractional digit, range-checked against [lo, hi] */
static __device__ __noinline__ uint32_t
num_to_int(const xnum x, int64_t lo, int64_t hi, int64_t *out)
{
    if (x.kind != XNUM_VALID)
        return XMSG_NUM_NANINF + 1;
    xi128 ival = x.v;
    int weight = x.scale;
    if (ival != 0)
    {
        while (weight > 0)
        {
            if (weight == 1)
                ival += (ival > 0 ? 5 : -5);
            ival /= 10;
            weight--;
        }
        while (weight < 0)
        {
            ival = num_mul10(ival);
            weight++;
            if (ival < lo || ival > hi)
                break;
        }
        if (ival < lo || ival > hi)
            return XMSG_NUM_INT_RANGE + 1;
    }
    *out = (int64_t)ival;
    return 0;
}

/* __xpu_numeric_to_fp64 (xpu_numeric.cu:341-380) */
static __device__ __noinline__ uint32_t
num_to_f64(const xnum x, double *out)
{
    if (x.kind == XNUM_VALID)
    {
        double f = (double)x.v;
        int weight = x.scale;
        if (f != 0.0)
        {
            for (; weight > 0; weight--)
                f /= 10.0;
            for (; weight < 0; weight++)
                f *= 10.0;
            if (__builtin_isnan(f) || __builtin_isinf(f))
                return XMSG_NUM_FLOAT_RANGE + 1;
        }
        *out = f;
    }
    else if (x.kind == XNUM_PINF)
        *out = __builtin_inf();
    else if (x.kind == XNUM_NINF)
        *out = -__builtin_inf();
    else
        *out = __builtin_nan("");
    return 0;
}

/* __decimal_from_float8 (xpu_numeric.h:778-846): the 52-bit fraction
 * rescaled by powers of 2 and 10 to the binary exponent of 1, then cut to
 * 15 decimal digits.  As there, any float whose exponent bits are all ones
 * becomes NaN, and zero takes the path of the normal numbers. */
static __device__ __noinline__ xnum
num_from_f64(double dv)
{
    const uint64_t bits = (uint64_t)__double_as_longlong(dv);
    uint64_t frac = bits & ((1ULL << 52) - 1);
    int32_t expo = (int32_t)((bits >> 52) & 0x7ff);
    const bool sign = (bits >> 63) != 0;
    int weight = 0;
    xnum x;
    x.v = 0;
    x.scale = 0;
    if (expo == 0x7ff)
    {
        x.kind = XNUM_NAN;
        return x;
    }
    frac |= (1ULL << 52);
    while (expo > 1023 + 52)
    {
        if (frac <= 0x1800000000000000ULL)
        {
            frac *= 2;
            expo--;
        }
        else
        {
            frac /= 10;
            weight--;
        }
    }
    while (expo < 1023 + 52)
    {
        if (frac >= 0x1800000000000000ULL)
        {
            frac /= 2;
            expo++;
        }
        else
        {
            frac *= 10;
            weight++;
        }
    }
    while (frac >= 1000000000000000ULL)
    {
        frac /= 10;
        weight--;
    }
    x.kind = XNUM_VALID;
    x.scale = (int16_t)weight;
    x.v = sign ? -(xi128)frac : (xi128)frac;
    return x;
}

/* xpu_numeric_datum_hash (xpu_numeric.cu:180-196) */
static __device__ __noinline__ uint32_t
num_hash(const xnum x)
{
    if (x.kind != XNUM_VALID)
    {
        const uint8_t k = (uint8_t)x.kind;
        return hash_bytes(&k, 1);
    }
    uint8_t w[2], v[16];
    __builtin_memcpy(w, &x.scale, 2);
    __builtin_memcpy(v, &x.v, 16);
    return hash_bytes(w, 2) ^ hash_bytes(v, 16);
}

/* __decimal_to_varlena (xpu_numeric.h:275-372): the long form with a 4-byte
 * header; the last base-10000 digit is split off so the scale becomes a
 * multiple of 4, which is also the display scale written.  Returns the
 * size; writes only if buf.  NaN and the infinities are the 6-byte
 * PostgreSQL encodings (NUMERIC_NAN / PINF / NINF). */
static __device__ __noinline__ uint32_t
num_to_varlena(const xnum x, uint8_t *buf)
{
    if (x.kind != XNUM_VALID)
    {
        if (buf)
        {
            const uint16_t h = x.kind == XNUM_PINF ? 0xD000 : (x.kind == XNUM_NINF ? 0xF000 : 0xC000);
            const uint32_t vl = 6u << 2;
            __builtin_memcpy(buf, &vl, 4);
            __builtin_memcpy(buf + 4, &h, 2);
        }
        return 6;
    }
    xnum n = x;
    const bool negative = n.v < 0;
    if (negative)
        n.v = -n.v;
    num_normalize(n);
    int weight = n.scale;
    xi128 value = n.v;
    uint16_t digs[12];
    int ndigits = 0;
    if (value != 0)
    {
        int mod = -1;
        switch (weight % 4)
        {
            case -1: case 3:
                mod = (int)(value % 1000) * 10; value /= 1000; weight += 1; break;
            case -2: case 2:
                mod = (int)(value % 100) * 100; value /= 100; weight += 2; break;
            case -3: case 1:
                mod = (int)(value % 10) * 1000; value /= 10; weight += 3; break;
        }
        if (mod >= 0)
            digs[ndigits++] = (uint16_t)mod;
    }
    else
        weight = 0;
    while (value != 0 && ndigits < 12)
    {
        digs[ndigits++] = (uint16_t)(value % 10000);
        value /= 10000;
    }
    const int nzero = weight < 0 ? -weight / 4 : 0;
    const uint32_t len = 8 + 2 * (uint32_t)(ndigits + nzero);
    if (buf)
    {
        const uint32_t vl = len << 2;
        __builtin_memcpy(buf, &vl, 4);
        uint16_t sign_dscale = (uint16_t)(weight > 0 ? weight : 0);
        if (negative)
            sign_dscale |= 0x4000;
        const int16_t n_weight = (int16_t)(ndigits - (weight / 4) - 1);
        __builtin_memcpy(buf + 4, &sign_dscale, 2);
        __builtin_memcpy(buf + 6, &n_weight, 2);
        /* digits most significant first: digs[] holds them least first */
        for (int i = 0; i < ndigits; i++)
            __builtin_memcpy(buf + 8 + 2 * i, &digs[ndigits - 1 - i], 2);
        for (int i = 0; i < nzero; i++)
        {
            buf[8 + 2 * (ndigits + i)] = 0;
            buf[9 + 2 * (ndigits + i)] = 0;
        }
    }
    return len;
}

/* xpu_numeric_datum_write (xpu_numeric.cu:163-178): a datum still in its
 * varlena form is copied as it is (compressed and external ones too),
 * anything else is encoded */
__device__ __forceinline__ uint32_t
num_reg_varlena_size(const Lane &L, int r)
{
    if ((L.R[r + 2] & 0xff) == XNUM_VARLENA)
        return varsize_any((const uint8_t *)(uintptr_t)L.R[r]);
    xnum x;
    num_read(L, r, &x);
    return num_to_varlena(x, nullptr);
}

__device__ __forceinline__ uint32_t
num_reg_varlena_write(const Lane &L, int r, uint8_t *buf)
{
    if ((L.R[r + 2] & 0xff) == XNUM_VARLENA)
    {
        const uint8_t *s = (const uint8_t *)(uintptr_t)L.R[r];
        const uint32_t sz = varsize_any(s);
        for (uint32_t i = 0; i < sz; i++)
            buf[i] = s[i];
        return sz;
    }
    xnum x;
    num_read(L, r, &x);
    return num_to_varlena(x, buf);
}

/* the attrs flag of a NaN or infinite value in a numeric sum
 * (__PAGG_NUMERIC_ATTRS__*) */
__device__ __forceinline__ uint32_t
num_special_flag(int kind)
{
    return kind == XNUM_PINF ? __PAGG_NUMERIC_ATTRS__PINF
         : kind == XNUM_NINF ? __PAGG_NUMERIC_ATTRS__NINF : __PAGG_NUMERIC_ATTRS__NAN;
}

/* __normalize_numeric_int128 (xpu_numeric.cu:1080-1126): a finite value
 * rescaled to the scale of a numeric sum (truncated when digits go) */
static __device__ __noinline__ xi128
num_rescale(xi128 v, int scale_d, int scale_s)
{
    for (; scale_d > scale_s; scale_s++)
        v = num_mul10(v);
    for (; scale_d < scale_s; scale_s--)
        v /= 10;
    return v;
}

/* STAGED: the caller has loaded the columns of the plan's stage[] list into
 * their registers, so the program's XLD_STAGED loads are skipped */
template <bool STAGED>
__device__ __forceinline__ void
xop_exec(const xop o, Lane &L, const kern_data_store *__restrict__ kds, uint32_t row,
         const xpu_kparams &kp)
{
    const char *kbase = (const char *)kds;
    ckds *kmeta = (ckds *)(uintptr_t)kds;
    const char *const_pool = kp.const_pool;
    {
        const int d = o.dst, a = o.a, b = o.b, g = o.guard;
        switch (o.op)
        {
            case XO_LDCOL:
            {
                if (STAGED && (o.imm & XLD_STAGED))
                    break;
                ccolmeta *cm = &kmeta->colmeta[o.aux];
                uint64_t nm_off = cm->nullmap_offset;
                uint64_t v_off = cm->values_offset;
                uint64_t v_len = cm->values_length;
                int w = (int)(o.imm & XLD_WIDTH_MASK);
                bool isnull = false;
                if (nm_off)
                {
                    uint32_t idx = row >> 3;
                    isnull = (idx >= cm->nullmap_length) ||
                             !((((const uint8_t *)kbase)[nm_off + idx] >> (row & 7)) & 1);
                }
                /* the value is read whenever it lies inside the column, so
                 * its load does not wait for the validity bit */
                const bool inside = (uint64_t)w * (row + 1) <= v_len;
                int64_t v = 0;
                if (inside)
                {
                    const char *p = kbase + v_off + (uint64_t)w * row;
                    switch (w)
                    {
                        case 1: v = *(const int8_t *)p; break;
                        case 2: v = *(const int16_t *)p; break;
                        case 4: v = (b == XD_F32) ? (int64_t)*(const uint32_t *)p
                                                  : (int64_t)*(const int32_t *)p; break;
                        default: v = *(const int64_t *)p; break;
                    }
                }
                isnull = isnull || !inside;
                L.R[d] = isnull ? 0 : v;
                SETNULL(L, d, isnull);
                break;
            }
            case XO_LDCOL_BOOL:
            {
                ccolmeta *cm = &kmeta->colmeta[o.aux];
                bool isnull = false;
                if (cm->nullmap_offset)
                {
                    uint32_t idx = row >> 3;
                    isnull = (idx >= cm->nullmap_length) ||
                             !((((const uint8_t *)kbase)[cm->nullmap_offset + idx] >> (row & 7)) & 1);
                }
                if ((uint64_t)(row >> 3) >= cm->values_length)
                    isnull = true;
                L.R[d] = isnull ? 0 : ((((const uint8_t *)kbase)[cm->values_offset + (row >> 3)] >> (row & 7)) & 1);
                SETNULL(L, d, isnull);
                break;
            }
            case XO_LDCOL_VAR:
            {
                ccolmeta *cm = &kmeta->colmeta[o.aux];
                bool isnull = false;
                if (cm->nullmap_offset)
                {
                    uint32_t idx = row >> 3;
                    isnull = (idx >= cm->nullmap_length) ||
                             !((((const uint8_t *)kbase)[cm->nullmap_offset + idx] >> (row & 7)) & 1);
                }
                int64_t ptr = 0, len = 0;
                if (cm->attopts.tag == XARROW_TAG_HEAPREF)
                {
                    /* column tiles of a heap / row chunk: (offset, length)
                     * pairs into the source chunk */
                    if (8ULL * ((uint64_t)row + 1) > cm->values_length)
                        isnull = true;
                    if (!isnull)
                    {
                        uint2 pr = ((const uint2 *)(kbase + cm->values_offset))[row];
                        if (pr.y == XHEAP_LEN_EXTERNAL)
                        {
                            /* kept as it is: the offset addresses its header */
                            const uint8_t *vp = (const uint8_t *)kbase + cm->extra_offset + pr.x;
                            if ((uint64_t)pr.x + 2 > cm->extra_length ||
                                (uint64_t)pr.x + varsize_any(vp) > cm->extra_length)
                                isnull = true;
                            else
                            {
                                ptr = (int64_t)(uintptr_t)vp;
                                len = XTEXT_RAW;
                            }
                        }
                        else if ((uint64_t)pr.x + pr.y > cm->extra_length)
                            isnull = true;
                        else
                        {
                            ptr = (int64_t)((uintptr_t)kbase + cm->extra_offset + pr.x);
                            len = pr.y;
                        }
                    }
                }
                else if (cm->attopts.tag == ArrowType__Utf8 || cm->attopts.tag == ArrowType__Binary)
                {
                    /* KDS_ARROW_REF_VARLENA32_DATUM (xpu_common.h:1577-1601) */
                    if (4ULL * ((uint64_t)row + 2) > cm->values_length)
                        isnull = true;
                    if (!isnull)
                    {
                        const uint32_t *off = (const uint32_t *)(kbase + cm->values_offset);
                        uint32_t o0 = off[row], o1 = off[row + 1];
                        if (o0 > o1 || o1 > cm->extra_length || o1 - o0 > XARROW_VARATT_MAX)
                            isnull = true;
                        else
                        {
                            ptr = (int64_t)(uintptr_t)(kbase + cm->extra_offset + o0);
                            len = o1 - o0;
                        }
                    }
                }
                else if (cm->attopts.tag == ArrowType__LargeUtf8 ||
                         cm->attopts.tag == ArrowType__LargeBinary)
                {
                    /* KDS_ARROW_REF_VARLENA64_DATUM (xpu_common.h:1603-1626) */
                    if (8ULL * ((uint64_t)row + 2) > cm->values_length)
                        isnull = true;
                    if (!isnull)
                    {
                        const uint64_t *off = (const uint64_t *)(kbase + cm->values_offset);
                        uint64_t o0 = off[row], o1 = off[row + 1];
                        if (o0 > o1 || o1 > cm->extra_length || o1 - o0 > XARROW_VARATT_MAX)
                            isnull = true;
                        else
                        {
                            ptr = (int64_t)(uintptr_t)(kbase + cm->extra_offset + o0);
                            len = (int64_t)(o1 - o0);
                        }
                    }
                }
                else if (cm->attopts.tag == ArrowType__FixedSizeBinary)
                {
                    /* KDS_ARROW_REF_SIMPLE_DATUM with the byte width
                     * (xpu_textlib.cu:30-35, xpu_common.h:1556-1573) */
                    const uint64_t w = cm->attopts.u.fixed_size_binary.byteWidth;
                    if (w * ((uint64_t)row + 1) > cm->values_length)
                        isnull = true;
                    if (!isnull)
                    {
                        ptr = (int64_t)(uintptr_t)(kbase + cm->values_offset + w * row);
                        len = (int64_t)w;
                    }
                }
                else
                {
                    /* "not a mappable Arrow data type" (xpu_textlib.cu:49-51) */
                    if (!L.err)
                        L.err = XMSG_ARROW_TYPE + 1;
                    isnull = true;
                }
                L.R[d] = ptr;
                L.R[d + 1] = len;
                SETNULL(L, d, isnull);
                break;
            }
            case XO_LDTUP:
                ldtup(L, d, (const kern_tupitem *)(uintptr_t)L.R[a],
                      (const kern_data_store *)(uintptr_t)kp.inner_kds[o.aux],
                      (int)o.imm, b);
                break;
            case XO_CONST:
                L.R[d] = o.imm;
                SETNULL(L, d, o.aux & 1);
                break;
            case XO_CONST_TEXT:
                L.R[d] = (int64_t)(uintptr_t)(const_pool + (o.imm & 0xffffffffLL));
                L.R[d + 1] = (int64_t)((uint64_t)o.imm >> 32);
                SETNULL(L, d, 0);
                break;
            case XO_MOV:
                L.R[d] = L.R[a];
                SETNULL(L, d, RNULL(L, a));
                break;
            case XO_MOV2:
                L.R[d] = L.R[a];
                L.R[d + 1] = L.R[a + 1];
                SETNULL(L, d, RNULL(L, a));
                break;
            case XO_SETNULL:
                L.R[d] = 0;
                SETNULL(L, d, 1);
                break;
            case XO_CMP_I:
            case XO_CMP_F32:
            case XO_CMP_F64:
            case XO_CMP_TEXT:
            {
                int op = o.aux & 0xff;
                bool na = RNULL(L, a), nb = RNULL(L, b);
                if (na || nb)
                {
                    /* kmode_compare_nulls for eq/ne (xpu_basetype.h:30-60) */
                    if ((o.aux & 0x100) && (op == XC_EQ || op == XC_NE))
                    {
                        bool both = na && nb;
                        L.R[d] = (op == XC_EQ) ? both : !both;
                        SETNULL(L, d, 0);
                    }
                    else
                    {
                        L.R[d] = 0;
                        SETNULL(L, d, 1);
                    }
                    break;
                }
                bool r;
                if (o.op == XO_CMP_I)
                {
                    int64_t x = L.R[a], y = L.R[b];
                    r = cmp_by(op, x < y ? -1 : (x > y ? 1 : 0));
                }
                else if (o.op == XO_CMP_F32)
                    r = fcmp(op, (double)as_f32(L.R[a]), (double)as_f32(L.R[b]));
                else if (o.op == XO_CMP_F64)
                    r = fcmp(op, as_f64(L.R[a]), as_f64(L.R[b]));
                else
                {
                    if (TXT_RAW(a) || TXT_RAW(b))
                    {
                        RAISE(g, XMSG_TEXT_EXTERNAL);
                        L.R[d] = 0;
                        SETNULL(L, d, 1);
                        break;
                    }
                    const uint8_t *s1 = (const uint8_t *)(uintptr_t)L.R[a];
                    const uint8_t *s2 = (const uint8_t *)(uintptr_t)L.R[b];
                    int64_t l1 = L.R[a + 1], l2 = L.R[b + 1];
                    if (o.aux & XCMP_BPCHAR)
                    {
                        l1 = txt_truelen(s1, l1);
                        l2 = txt_truelen(s2, l2);
                    }
                    r = cmp_by(op, txt_bytecmp(s1, l1, s2, l2));
                }
                L.R[d] = r;
                SETNULL(L, d, 0);
                break;
            }
            case XO_ADD_I: case XO_SUB_I: case XO_MUL_I: case XO_DIV_I: case XO_MOD_I:
            {
                bool n = RNULL(L, a) | RNULL(L, b);
                int64_t x = L.R[a], y = L.R[b], v = 0;
                int cls = o.aux;
                if (!n)
                {
                    bool ovf = false;
                    switch (o.op)
                    {
                        case XO_ADD_I: ovf = __builtin_add_overflow(x, y, &v); break;
                        case XO_SUB_I: ovf = __builtin_sub_overflow(x, y, &v); break;
                        case XO_MUL_I: ovf = __builtin_mul_overflow(x, y, &v); break;
                        case XO_DIV_I:
                            if (y == 0)
                            {
                                RAISE(g, o.imm & 0xffff);
                                break;
                            }
                            if (y == -1)
                            {
                                ovf = (x == INT64_MIN);
                                v = -x;
                            }
                            else
                                v = x / y;
                            break;
                        case XO_MOD_I:
                            if (y == 0)
                            {
                                RAISE(g, o.imm & 0xffff);
                                break;
                            }
                            v = (y == -1) ? 0 : x % y;
                            break;
                    }
                    if (ovf || !int_in_class(cls, v))
                        RAISE(g, (o.imm >> 16) & 0xffff);
                }
                L.R[d] = v;
                SETNULL(L, d, n);
                break;
            }
            case XO_NEG_I: case XO_ABS_I:
            {
                bool n = RNULL(L, a);
                int64_t x = L.R[a], v = 0;
                if (!n)
                {
                    if (x == int_class_min(o.aux) && (o.op == XO_NEG_I || x < 0))
                        RAISE(g, o.imm);
                    v = (o.op == XO_NEG_I) ? -x : (x < 0 ? -x : x);
                }
                L.R[d] = v;
                SETNULL(L, d, n);
                break;
            }
            case XO_NOT_I:
                L.R[d] = wrap_class(o.aux, ~L.R[a]);
                SETNULL(L, d, RNULL(L, a));
                break;
            case XO_AND_I: case XO_OR_I: case XO_XOR_I: case XO_SHL_I: case XO_SHR_I:
            {
                bool n = RNULL(L, a) | RNULL(L, b);
                int64_t x = L.R[a], y = L.R[b], v;
                int bits = (o.aux == XT_INT2) ? 16 : (o.aux == XT_INT4) ? 32 :
                           (o.aux == XT_INT8) ? 64 : 8;
                switch (o.op)
                {
                    case XO_AND_I: v = x & y; break;
                    case XO_OR_I:  v = x | y; break;
                    case XO_XOR_I: v = x ^ y; break;
                    /* C's shifts of xpu_basetype.cu:858-866 as PostgreSQL's
                     * executor runs them: int1/int2/int4 promoted to 32-bit
                     * and shifted by y mod 32, int8 by y mod 64 */
                    case XO_SHL_I:
                        v = bits == 64 ? (int64_t)((uint64_t)x << (y & 63))
                                       : (int64_t)(int32_t)((uint32_t)x << (y & 31));
                        break;
                    default:
                        v = bits == 64 ? x >> (y & 63) : (int64_t)((int32_t)x >> (y & 31));
                        break;
                }
                L.R[d] = wrap_class(o.aux, v);
                SETNULL(L, d, n);
                break;
            }
            case XO_ADD_F32: case XO_SUB_F32: case XO_MUL_F32: case XO_DIV_F32:
            {
                bool n = RNULL(L, a) | RNULL(L, b);
                float x = as_f32(L.R[a]), y = as_f32(L.R[b]), v = 0.0f;
                if (!n)
                {
                    switch (o.op)
                    {
                        case XO_ADD_F32: v = x + y; break;
                        case XO_SUB_F32: v = x - y; break;
                        case XO_MUL_F32: v = x * y; break;
                        default:
                            if (y == 0.0f)
                                RAISE(g, (o.imm >> 16) & 0xffff);
                            v = x / y;
                            if (v == 0.0f && x != 0.0f)
                                RAISE(g, o.imm & 0xffff);
                            break;
                    }
                    if (isinf(v) && !isinf(x) && !isinf(y))
                        RAISE(g, o.imm & 0xffff);
                    v = nan_as_x86_f32(x, y, v);
                }
                L.R[d] = f32_bits(v);
                SETNULL(L, d, n);
                break;
            }
            case XO_ADD_F64: case XO_SUB_F64: case XO_MUL_F64: case XO_DIV_F64:
            {
                bool n = RNULL(L, a) | RNULL(L, b);
                double x = as_f64(L.R[a]), y = as_f64(L.R[b]), v = 0.0;
                if (!n)
                {
                    switch (o.op)
                    {
                        case XO_ADD_F64: v = x + y; break;
                        case XO_SUB_F64: v = x - y; break;
                        case XO_MUL_F64: v = x * y; break;
                        default:
                            if (y == 0.0)
                                RAISE(g, (o.imm >> 16) & 0xffff);
                            v = x / y;
                            if (v == 0.0 && x != 0.0)
                                RAISE(g, o.imm & 0xffff);
                            break;
                    }
                    if (isinf(v) && !isinf(x) && !isinf(y))
                        RAISE(g, o.imm & 0xffff);
                    v = nan_as_x86_f64(x, y, v);
                }
                L.R[d] = f64_bits(v);
                SETNULL(L, d, n);
                break;
            }
            case XO_NEG_F32: L.R[d] = f32_bits(-as_f32(L.R[a])); SETNULL(L, d, RNULL(L, a)); break;
            case XO_ABS_F32: L.R[d] = f32_bits(fabsf(as_f32(L.R[a]))); SETNULL(L, d, RNULL(L, a)); break;
            case XO_NEG_F64: L.R[d] = f64_bits(-as_f64(L.R[a])); SETNULL(L, d, RNULL(L, a)); break;
            case XO_ABS_F64: L.R[d] = f64_bits(fabs(as_f64(L.R[a]))); SETNULL(L, d, RNULL(L, a)); break;
            case XO_I2I:
            {
                bool n = RNULL(L, a);
                if (!n && !int_in_class(o.aux, L.R[a]))
                    RAISE(g, o.imm);
                L.R[d] = L.R[a];
                SETNULL(L, d, n);
                break;
            }
            case XO_F32_I:
            case XO_F64_I:
            {
                bool n = RNULL(L, a);
                int64_t v = 0;
                if (!n)
                {
                    /* rint then range check (xpu_basetype.cu:345-400) */
                    bool bad;
                    if (o.op == XO_F32_I)
                    {
                        float x = as_f32(L.R[a]);
                        float r = rintf(x);
                        bad = isnan(x) || r < (float)int_class_min(o.aux) ||
                              (o.aux == XT_INT8 ? r >= 9223372036854775808.0f
                                                : r > -(float)int_class_min(o.aux) - 1.0f);
                        v = bad ? 0 : (int64_t)r;
                    }
                    else
                    {
                        double x = as_f64(L.R[a]);
                        double r = rint(x);
                        bad = isnan(x) || r < (double)int_class_min(o.aux) ||
                              (o.aux == XT_INT8 ? r >= 9223372036854775808.0
                                                : r > -(double)int_class_min(o.aux) - 1.0);
                        v = bad ? 0 : (int64_t)r;
                    }
                    if (bad)
                        RAISE(g, o.imm);
                }
                L.R[d] = v;
                SETNULL(L, d, n);
                break;
            }
            case XO_I_F32: L.R[d] = f32_bits((float)L.R[a]); SETNULL(L, d, RNULL(L, a)); break;
            case XO_I_F64: L.R[d] = f64_bits((double)L.R[a]); SETNULL(L, d, RNULL(L, a)); break;
            case XO_F32_F64: L.R[d] = f64_bits((double)as_f32(L.R[a])); SETNULL(L, d, RNULL(L, a)); break;
            case XO_F64_F32:
            {
                bool n = RNULL(L, a);
                double x = as_f64(L.R[a]);
                if (!n && (isinf(x) || isnan(x) || x < -3.40282346638528859812e+38 ||
                           x > 3.40282346638528859812e+38))
                    RAISE(g, o.imm);
                L.R[d] = f32_bits((float)x);
                SETNULL(L, d, n);
                break;
            }
            case XO_I_BOOL: L.R[d] = (L.R[a] != 0); SETNULL(L, d, RNULL(L, a)); break;
            case XO_TEXTLEN:
            {
                /* pgfn_textlen / pgfn_bpcharlen (xpu_textlib.cu:583-720):
                 * characters by the database encoding, up to a NUL byte */
                const uint8_t *s = (const uint8_t *)(uintptr_t)L.R[a];
                int64_t len = L.R[a + 1];
                bool n = RNULL(L, a);
                if (!n && len < 0)
                {
                    if ((o.aux & (XTXT_BPCHAR | XENC_UTF8)) == 0)
                    {
                        /* a single-byte encoding counts bytes, which the
                         * header tells: the external pointer's va_extinfo,
                         * or for a compressed datum the word the reference
                         * copies from VARDATA into toast_compress_header's
                         * tcinfo (the one after the compression info) */
                        const uint8_t *w = s + ((s[0] == 0x01) ? 6 : 8);
                        len = (int64_t)((((uint32_t)w[0] | ((uint32_t)w[1] << 8) |
                                          ((uint32_t)w[2] << 16) | ((uint32_t)w[3] << 24))) &
                                        0x3fffffffU);
                        L.R[d] = len;
                        SETNULL(L, d, 0);
                        break;
                    }
                    RAISE(g, XMSG_TEXT_EXTERNAL);
                    n = true;
                }
                if (!n && (o.aux & XTXT_BPCHAR))
                    len = txt_truelen(s, len);
                if (!n && (o.aux & XENC_UTF8))
                {
                    int64_t cnt = 0, i = 0;
                    while (i < len && s[i] != 0)
                    {
                        i += txt_mblen(s + i, true);
                        cnt++;
                    }
                    len = cnt;
                }
                L.R[d] = n ? 0 : len;
                SETNULL(L, d, n);
                break;
            }
            case XO_LIKE:
            {
                bool n = RNULL(L, a) | RNULL(L, b);
                int64_t v = 0;
                if (!n && (TXT_RAW(a) || TXT_RAW(b)))
                {
                    RAISE(g, XMSG_TEXT_EXTERNAL);
                    n = true;
                }
                if (!n)
                {
                    const uint8_t *t = (const uint8_t *)(uintptr_t)L.R[a];
                    int64_t tlen = L.R[a + 1];
                    if (o.aux & XTXT_BPCHAR)
                        tlen = txt_truelen(t, tlen);
                    const int m = like_match(t, tlen, (const uint8_t *)(uintptr_t)L.R[b], L.R[b + 1],
                                             (o.aux & XLIKE_ICASE) != 0, (o.aux & XENC_UTF8) != 0);
                    if (m < 0)
                        RAISE(g, o.imm);
                    v = (m > 0) != ((o.aux & XLIKE_NEGATE) != 0);
                }
                L.R[d] = v;
                SETNULL(L, d, n);
                break;
            }
            case XO_SUBSTR:
            {
                const bool has_count = (o.imm & XSUB_HAS_COUNT) != 0;
                const int rc = (int)(o.imm & 0xff);
                bool n = RNULL(L, a) | RNULL(L, b) | (has_count && RNULL(L, rc));
                int64_t off = 0, len = 0;
                const uint8_t *str = (const uint8_t *)(uintptr_t)L.R[a];
                if (!n && TXT_RAW(a))
                {
                    RAISE(g, XMSG_TEXT_EXTERNAL);
                    n = true;
                }
                if (!n)
                {
                    const int32_t count = has_count ? (int32_t)L.R[rc] : -1;
                    if (has_count && count < 0)
                        RAISE(g, (o.imm >> 16) & 0xffff);
                    else
                        substr_range(str, L.R[a + 1], (int32_t)L.R[b] - 1, count,
                                     (o.aux & XENC_UTF8) != 0, &off, &len);
                }
                L.R[d] = (int64_t)(uintptr_t)(str + off);
                L.R[d + 1] = len;
                SETNULL(L, d, n);
                break;
            }
            case XO_AND3:
            {
                bool na = RNULL(L, a), nb = RNULL(L, b);
                bool fa = !na && L.R[a] == 0, fb = !nb && L.R[b] == 0;
                if (fa || fb)
                {
                    L.R[d] = 0;
                    SETNULL(L, d, 0);
                }
                else
                {
                    L.R[d] = !(na || nb);
                    SETNULL(L, d, na || nb);
                }
                break;
            }
            case XO_OR3:
            {
                bool na = RNULL(L, a), nb = RNULL(L, b);
                bool ta = !na && L.R[a] != 0, tb = !nb && L.R[b] != 0;
                if (ta || tb)
                {
                    L.R[d] = 1;
                    SETNULL(L, d, 0);
                }
                else
                {
                    L.R[d] = 0;
                    SETNULL(L, d, na || nb);
                }
                break;
            }
            case XO_NOT3:
                L.R[d] = RNULL(L, a) ? 0 : (L.R[a] == 0);
                SETNULL(L, d, RNULL(L, a));
                break;
            case XO_ISNULL:
                L.R[d] = RNULL(L, a);
                SETNULL(L, d, 0);
                break;
            case XO_ISNOTNULL:
                L.R[d] = !RNULL(L, a);
                SETNULL(L, d, 0);
                break;
            case XO_BOOLTEST:
            {
                bool n = RNULL(L, a), v = L.R[a] != 0, r;
                switch (o.aux)
                {
                    case 0: r = !n && v; break;
                    case 1: r = n || !v; break;
                    case 2: r = !n && !v; break;
                    case 3: r = n || v; break;
                    case 4: r = n; break;
                    default: r = !n; break;
                }
                L.R[d] = r;
                SETNULL(L, d, 0);
                break;
            }
            case XO_COALESCE:
                if (RNULL(L, d))
                {
                    L.R[d] = L.R[a];
                    if (b)
                        L.R[d + 1] = L.R[a + 1];
                    if (b > 1)
                        L.R[d + 2] = L.R[a + 2];
                    SETNULL(L, d, RNULL(L, a));
                }
                break;
            case XO_G_NOTFALSE:
                L.R[d] = (L.R[b] != 0) && (RNULL(L, a) || L.R[a] != 0);
                SETNULL(L, d, 0);
                break;
            case XO_G_NOTTRUE:
                L.R[d] = (L.R[b] != 0) && (RNULL(L, a) || L.R[a] == 0);
                SETNULL(L, d, 0);
                break;
            case XO_G_ISNULL:
                L.R[d] = (L.R[b] != 0) && RNULL(L, a);
                SETNULL(L, d, 0);
                break;
            case XO_G_POSITIVE:
                L.R[d] = (L.R[b] != 0) && !RNULL(L, a) && L.R[a] > 0;
                SETNULL(L, d, 0);
                break;
            case XO_G_NONZERO:
                L.R[d] = (L.R[b] != 0) && !RNULL(L, a) && L.R[a] != 0;
                SETNULL(L, d, 0);
                break;
            case XO_JQ_STATUS:
                if (L.R[b] != 0 && (RNULL(L, a) || L.R[a] == 0))
                    L.R[d] = o.imm;
                break;
            case XO_HASH_INIT:
                L.R[d] = 0xffffffffLL;
                SETNULL(L, d, 0);
                break;
            case XO_HASH_ADD:
                if (!RNULL(L, a))
                {
                    uint32_t h;
                    if (o.aux == 0 && TXT_RAW(a))
                    {
                        RAISE(g, XMSG_TEXT_EXTERNAL);
                        break;
                    }
                    if (o.aux == XHASH_NUMERIC)
                    {
                        xnum x;
                        const uint32_t e = num_read(L, a, &x);
                        if (e)
                        {
                            RAISE(g, e - 1);
                            break;
                        }
                        h = num_hash(x);
                    }
                    else if (o.aux == 0)
                        h = hash_bytes((const uint8_t *)(uintptr_t)L.R[a], (uint32_t)L.R[a + 1]);
                    else
                        h = hash_fixed((uint64_t)L.R[a], o.aux);
                    L.R[d] = hash_merge((uint32_t)L.R[d], h);
                }
                break;
            case XO_HASH_FIN:
                L.R[d] = (uint32_t)L.R[d] ^ 0xffffffffU;
                break;
            case XO_KILL:
                L.dead = RNULL(L, a) || L.R[a] == 0;
                break;
            case XO_PARAM:
                L.R[d] = kp.params[o.imm];
                SETNULL(L, d, (kp.param_nulls >> o.imm) & 1);
                break;
            case XO_G_TRUE:
                L.R[d] = (L.R[b] != 0) && !RNULL(L, a) && L.R[a] != 0;
                SETNULL(L, d, 0);
                break;
            case XO_SELECT:
                if (L.R[g] != 0)
                {
                    L.R[d] = L.R[a];
                    if (o.aux & 3)
                        L.R[d + 1] = L.R[a + 1];
                    if (o.aux & 2)
                        L.R[d + 2] = L.R[a + 2];
                    SETNULL(L, d, RNULL(L, a));
                }
                break;
            case XO_DISTINCT:
            {
                /* pgfn_DistinctFrom (xpu_common.cu:613-660) */
                const bool na = RNULL(L, a), nb = RNULL(L, b);
                L.R[d] = (na || nb) ? (int64_t)!(na && nb) : (int64_t)(L.R[d] == 0);
                SETNULL(L, d, 0);
                break;
            }
            case XO_MINMAX:
            {
                /* pgfn_LeastExpr / pgfn_GreatestExpr (xpu_common.cu:683-757):
                 * the type's comparison; floats by < and > only */
                if (RNULL(L, a))
                    break;
                if ((o.aux & XMM_DOMAIN) == XD_NUM)
                {
                    /* the comparison parses both datums in place, so the
                     * kept value is in its parsed form */
                    if (RNULL(L, d))
                    {
                        L.R[d] = L.R[a];
                        L.R[d + 1] = L.R[a + 1];
                        L.R[d + 2] = L.R[a + 2];
                        SETNULL(L, d, 0);
                        break;
                    }
                    xnum x, y;
                    uint32_t e = num_read(L, d, &x);
                    if (!e)
                        e = num_read(L, a, &y);
                    if (e)
                    {
                        RAISE(g, e - 1);
                        break;
                    }
                    const int c = num_compare(x, y);
                    int32_t tk = ((o.aux & XMM_GREATEST) ? (c < 0) : (c > 0)) ? 1 : 0;
                    asm volatile("" : "+v"(tk));
                    num_store(L, d, tk ? y : x);
                    break;
                }
                bool take = RNULL(L, d);
                if (!take)
                {
                    int c;
                    const int dom = o.aux & XMM_DOMAIN;
                    if (dom == XD_INT)
                        c = L.R[d] > L.R[a] ? 1 : (L.R[d] < L.R[a] ? -1 : 0);
                    else if (dom == XD_F32)
                    {
                        const float x = as_f32(L.R[d]), y = as_f32(L.R[a]);
                        c = x > y ? 1 : (x < y ? -1 : 0);
                    }
                    else if (dom == XD_F64)
                    {
                        const double x = as_f64(L.R[d]), y = as_f64(L.R[a]);
                        c = x > y ? 1 : (x < y ? -1 : 0);
                    }
                    else
                    {
                        if (TXT_RAW(d) || TXT_RAW(a))
                        {
                            RAISE(g, XMSG_TEXT_EXTERNAL);
                            break;
                        }
                        const uint8_t *s1 = (const uint8_t *)(uintptr_t)L.R[d];
                        const uint8_t *s2 = (const uint8_t *)(uintptr_t)L.R[a];
                        int64_t l1 = L.R[d + 1], l2 = L.R[a + 1];
                        if (o.aux & XMM_BPCHAR)
                        {
                            l1 = txt_truelen(s1, l1);
                            l2 = txt_truelen(s2, l2);
                        }
                        c = txt_bytecmp(s1, l1, s2, l2);
                    }
                    take = (o.aux & XMM_GREATEST) ? (c < 0) : (c > 0);
                }
                /* the choice is made opaque before both words of a text
                 * value move: with the branches of the comparison above,
                 * gfx950 code generation otherwise lost the pointer word's
                 * update for lanes that took a uniform (constant) argument */
                int32_t tk = take ? 1 : 0;
                asm volatile("" : "+v"(tk));
                const int64_t w0 = tk ? L.R[a] : L.R[d];
                L.R[d] = w0;
                if ((o.aux & XMM_DOMAIN) == XD_TEXT)
                {
                    const int64_t w1 = tk ? L.R[a + 1] : L.R[d + 1];
                    L.R[d + 1] = w1;
                }
                if (tk)
                    SETNULL(L, d, 0);
                break;
            }
            case XO_HAOP:
            {
                /* pgfn_HashedArrayOp (xpu_common.cu:1013-1088) over the
                 * expression's bytes in the constant pool; equality as the
                 * type's '=' (floats: NaN equals NaN only) */
                const uint8_t *kx = (const uint8_t *)const_pool + o.imm;
                const int kkind = o.aux & 7, kwidth = (o.aux >> 3) & 15;
                const int rkind = (o.aux >> 7) & 7, rwidth = (o.aux >> 10) & 15;
                const bool text_res = (rkind == XHA_KIND_TEXT);
                bool n = RNULL(L, a);
                const uint8_t *val = nullptr;
                if (!n && kkind == XHA_KIND_TEXT && TXT_RAW(a))
                {
                    RAISE(g, XMSG_TEXT_EXTERNAL);
                    n = true;
                }
                if (!n)
                {
                    uint32_t nslots, default_offset, nitems;
                    __builtin_memcpy(&nslots, kx + 52, 4);
                    __builtin_memcpy(&default_offset, kx + 56, 4);
                    /* the item count: the compiler stores it in the
                     * copy's elem_ops field, the bound of a chain walk */
                    __builtin_memcpy(&nitems, kx + 40, 4);
                    const uint32_t h = (kkind == XHA_KIND_TEXT)
                        ? hash_bytes((const uint8_t *)(uintptr_t)L.R[a], (uint32_t)L.R[a + 1])
                        : hash_fixed((uint64_t)L.R[a], kwidth);
                    uint32_t next;
                    __builtin_memcpy(&next, kx + 64 + 4 * (h % nslots), 4);
                    bool found = false;
                    for (uint32_t step = 0; next != 0 && step < nitems && !found; step++)
                    {
                        const uint8_t *item = kx + next;
                        uint32_t ihash, inext;
                        __builtin_memcpy(&ihash, item, 4);
                        __builtin_memcpy(&inext, item + 4, 4);
                        if (ihash == h && (inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY))
                        {
                            const uint8_t *key = item + 8;
                            int64_t klen = kwidth;
                            bool eq;
                            if (kkind == XHA_KIND_TEXT)
                            {
                                const uint8_t *kp8 = varlena_payload(key, &klen);
                                eq = txt_bytecmp((const uint8_t *)(uintptr_t)L.R[a], L.R[a + 1],
                                                 kp8, klen) == 0;
                                klen += (int64_t)(kp8 - key);
                            }
                            else if (kkind == XHA_KIND_F32)
                                eq = fcmp(XC_EQ, (double)as_f32(L.R[a]),
                                          (double)as_f32((uint32_t)load_fixed(key, 4)));
                            else if (kkind == XHA_KIND_F64)
                                eq = fcmp(XC_EQ, as_f64(L.R[a]), as_f64(load_fixed(key, 8)));
                            else
                                eq = L.R[a] == load_fixed(key, kwidth);
                            if (eq)
                            {
                                found = true;
                                if (inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE)
                                {
                                    const int ralign = text_res ? 4 : rwidth;
                                    val = key + ((klen + ralign - 1) / ralign) * ralign;
                                }
                                else
                                    n = true;
                            }
                        }
                        next = inext & ~KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK;
                    }
                    if (!found)
                    {
                        if (o.aux & XHA_NULL_MISS)
                            n = true;
                        else if (default_offset)
                            val = kx + default_offset;
                        else
                            n = true;
                    }
                }
                if (!n && val)
                {
                    if (text_res)
                    {
                        int64_t len;
                        const uint8_t *pl = varlena_payload(val, &len);
                        L.R[d] = (int64_t)(uintptr_t)pl;
                        L.R[d + 1] = len;
                    }
                    else
                        L.R[d] = (rkind == XHA_KIND_F32) ? (int64_t)(uint32_t)load_fixed(val, 4)
                                                         : load_fixed(val, rwidth);
                }
                else
                    L.R[d] = 0;
                SETNULL(L, d, n || !val);
                break;
            }
            case XO_TXTCHK:
                if (!RNULL(L, a) && TXT_RAW(a))
                    RAISE(g, XMSG_TEXT_EXTERNAL);
                break;
            case XO_NUM_VL:
            {
                /* a varlena addressed by its payload (heap tiles, tuples): a
                 * 1-byte header is odd, the top byte of a 4-byte header of
                 * a datum below 4 MB even; a raw datum is addressed by its
                 * header already */
                const uint8_t *pl = (const uint8_t *)(uintptr_t)L.R[a];
                const bool n = RNULL(L, a);
                if (n)
                    num_store_varlena(L, d, nullptr, false);
                else if (L.R[a + 1] < 0)
                    num_store_varlena(L, d, pl, true);
                else
                    num_store_varlena(L, d, (pl[-1] & 1) ? pl - 1 : pl - 4, false);
                SETNULL(L, d, n);
                break;
            }
            case XO_CONST_NUM:
                num_store_varlena(L, d, (const uint8_t *)const_pool + o.imm, false);
                SETNULL(L, d, 0);
                break;
            case XO_LDCOL_DEC:
            {
                /* xpu_numeric_datum_arrow_read (xpu_numeric.cu:46-79):
                 * Decimal128 at the column's scale, normalized; heap tiles
                 * keep the varlena */
                ccolmeta *cm = &kmeta->colmeta[o.aux];
                bool isnull = false;
                if (cm->nullmap_offset)
                {
                    const uint32_t idx = row >> 3;
                    isnull = (idx >= cm->nullmap_length) ||
                             !((((const uint8_t *)kbase)[cm->nullmap_offset + idx] >> (row & 7)) & 1);
                }
                if (cm->attopts.tag == XARROW_TAG_HEAPREF)
                {
                    if (8ULL * ((uint64_t)row + 1) > cm->values_length)
                        isnull = true;
                    const uint8_t *vp = nullptr;
                    bool raw = false;
                    if (!isnull)
                    {
                        const uint2 pr = ((const uint2 *)(kbase + cm->values_offset))[row];
                        vp = (const uint8_t *)kbase + cm->extra_offset + pr.x;
                        raw = (pr.y == XHEAP_LEN_EXTERNAL);
                        if (raw ? ((uint64_t)pr.x + 2 > cm->extra_length ||
                                   (uint64_t)pr.x + varsize_any(vp) > cm->extra_length)
                                : ((uint64_t)pr.x + pr.y > cm->extra_length || pr.x < 1))
                            isnull = true;
                        else if (!raw)
                            vp = (vp[-1] & 1) ? vp - 1 : vp - 4;
                    }
                    num_store_varlena(L, d, isnull ? nullptr : vp, raw);
                    SETNULL(L, d, isnull);
                    break;
                }
                if (cm->attopts.tag != ArrowType__Decimal || cm->attopts.u.decimal.bitWidth != 128)
                {
                    if (!L.err)
                        L.err = XMSG_ARROW_TYPE + 1;
                    isnull = true;
                }
                if (16ULL * ((uint64_t)row + 1) > cm->values_length)
                    isnull = true;
                xnum x;
                x.kind = XNUM_VALID;
                x.v = 0;
                x.scale = (int16_t)cm->attopts.u.decimal.scale;
                if (!isnull)
                {
                    const uint64_t *q = (const uint64_t *)(kbase + cm->values_offset + 16ULL * row);
                    x.v = (xi128)(((xu128)q[1] << 64) | q[0]);
                    num_normalize(x);
                }
                num_store(L, d, x);
                SETNULL(L, d, isnull);
                break;
            }
            case XO_NUM_PARSE:
            {
                /* the datum parsed in place, as xpu_numeric_validate does
                 * to a numeric sum's argument */
                if (RNULL(L, a))
                    break;
                xnum x;
                const uint32_t e = num_read(L, a, &x);
                if (e)
                {
                    RAISE(g, e - 1);
                    break;
                }
                num_store(L, d, x);
                break;
            }
            case XO_MOV3:
                L.R[d] = L.R[a];
                L.R[d + 1] = L.R[a + 1];
                L.R[d + 2] = L.R[a + 2];
                SETNULL(L, d, RNULL(L, a));
                break;
            case XO_NUM_CMP:
            {
                const int op = o.aux & 0xff;
                const bool na = RNULL(L, a), nb = RNULL(L, b);
                if (na || nb)
                {
                    if ((o.aux & 0x100) && (op == XC_EQ || op == XC_NE))
                    {
                        const bool both = na && nb;
                        L.R[d] = (op == XC_EQ) ? both : !both;
                        SETNULL(L, d, 0);
                    }
                    else
                    {
                        L.R[d] = 0;
                        SETNULL(L, d, 1);
                    }
                    break;
                }
                xnum x, y;
                uint32_t e = num_read(L, a, &x);
                if (!e)
                    e = num_read(L, b, &y);
                if (e)
                {
                    RAISE(g, e - 1);
                    L.R[d] = 0;
                    SETNULL(L, d, 1);
                    break;
                }
                const int c = num_compare(x, y);
                bool v;
                switch (op)
                {
                    case XC_EQ: v = (c == 0); break;
                    case XC_NE: v = (c != 0); break;
                    case XC_LT: v = (c < 0); break;
                    case XC_LE: v = (c <= 0); break;
                    case XC_GT: v = (c > 0); break;
                    default: v = (c >= 0); break;
                }
                L.R[d] = v;
                SETNULL(L, d, 0);
                break;
            }
            case XO_NUM_ADD: case XO_NUM_SUB: case XO_NUM_MUL: case XO_NUM_DIV:
            {
                const bool n = RNULL(L, a) || RNULL(L, b);
                xnum r;
                r.kind = XNUM_VALID;
                r.v = 0;
                r.scale = 0;
                if (!n)
                {
                    xnum x, y;
                    uint32_t e = num_read(L, a, &x);
                    if (!e)
                        e = num_read(L, b, &y);
                    if (!e)
                        e = num_arith(o.op, x, y, &r);
                    if (e)
                        RAISE(g, e - 1);
                }
                num_store(L, d, r);
                SETNULL(L, d, n);
                break;
            }
            case XO_NUM_NEG: case XO_NUM_ABS:
            {
                /* pgfn_numeric_uminus / pgfn_numeric_abs (xpu_numeric.cu:
                 * 715-757); NaN stays NaN */
                const bool n = RNULL(L, a);
                xnum x;
                x.kind = XNUM_VALID;
                x.v = 0;
                x.scale = 0;
                if (!n)
                {
                    const uint32_t e = num_read(L, a, &x);
                    if (e)
                        RAISE(g, e - 1);
                    else if (o.op == XO_NUM_NEG)
                    {
                        if (x.kind == XNUM_VALID)
                            x.v = -x.v;
                        else if (x.kind == XNUM_PINF)
                            x.kind = XNUM_NINF;
                        else if (x.kind == XNUM_NINF)
                            x.kind = XNUM_PINF;
                    }
                    else
                    {
                        if (x.kind == XNUM_VALID && x.v < 0)
                            x.v = -x.v;
                        else if (x.kind == XNUM_NINF)
                            x.kind = XNUM_PINF;
                    }
                }
                num_store(L, d, x);
                SETNULL(L, d, n);
                break;
            }
            case XO_NUM_INT:
            {
                const bool n = RNULL(L, a);
                int64_t v = 0;
                if (!n)
                {
                    xnum x;
                    uint32_t e = num_read(L, a, &x);
                    if (!e)
                        e = num_to_int(x, int_class_min(o.aux),
                                       o.aux == XT_INT8 ? INT64_MAX : -(int_class_min(o.aux) + 1), &v);
                    if (e)
                    {
                        RAISE(g, e - 1);
                        v = 0;
                    }
                }
                L.R[d] = v;
                SETNULL(L, d, n);
                break;
            }
            case XO_NUM_F64: case XO_NUM_F32:
            {
                const bool n = RNULL(L, a);
                double f = 0.0;
                if (!n)
                {
                    xnum x;
                    uint32_t e = num_read(L, a, &x);
                    if (!e)
                        e = num_to_f64(x, &f);
                    if (e)
                    {
                        RAISE(g, e - 1);
                        f = 0.0;
                    }
                }
                L.R[d] = (o.op == XO_NUM_F64) ? f64_bits(f) : f32_bits((float)f);
                SETNULL(L, d, n);
                break;
            }
            case XO_INT_NUM:
            {
                xnum x;
                x.kind = XNUM_VALID;
                x.v = (xi128)L.R[a];
                x.scale = 0;
                num_store(L, d, x);
                SETNULL(L, d, RNULL(L, a));
                break;
            }
            case XO_F64_NUM: case XO_F32_NUM:
            {
                const double f = (o.op == XO_F64_NUM) ? as_f64(L.R[a]) : (double)as_f32(L.R[a]);
                num_store(L, d, num_from_f64(f));
                SETNULL(L, d, RNULL(L, a));
                break;
            }
            default:
                break;
        }
    }
}

/* runs the ops [begin, end) of the plan, fetched one by one (wave-uniform
 * program counter, scalar loads of each op) */
struct Interp
{
    static constexpr bool kUnrollRows = false;

    template <bool STAGED>
    __device__ static __forceinline__ void
    run(cplan *P, int begin, int end, Lane &L, const kern_data_store *__restrict__ kds,
        uint32_t row, const xpu_kparams &kp)
    {
        for (int pc = begin; pc < end; pc++)
            xop_exec<STAGED>(P->ops[pc], L, kds, row, kp);
    }
};

/* ------------------------------------------------------------------
 * wave helpers (wave64)
 * ------------------------------------------------------------------ */
__device__ __forceinline__ bool
wave_any(bool p)
{
    return __ballot(p) != 0;
}

__device__ __forceinline__ uint64_t
wave_incl_scan_u64(uint64_t v)
{
    int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        uint64_t t = __shfl_up(v, off, 64);
        if (lane >= off)
            v += t;
    }
    return v;
}

__device__ __forceinline__ void
record_error(xpu_task_ctl *ctl, uint32_t msgid, uint32_t code)
{
    /* a datum the GPU cannot read is the CPU fallback's error
     * (SUSPEND_FALLBACK, xpu_textlib.h:52-58), also where the session
     * has no fallback buffer to take the row */
    if (msgid == XMSG_TEXT_EXTERNAL)
        code = ERRCODE_SUSPEND_FALLBACK;
    if (atomicCAS(&ctl->errcode, 0U, code) == 0U)
        __hip_atomic_store(&ctl->errmsg, msgid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* ------------------------------------------------------------------
 * projection: kern_form_minimal_tuple (xpu_common.cu:1096-1190)
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t
form_tuple_size(const xattr *attrs, int nattrs, const Lane &L)
{
    bool hasnull = false;
    for (int j = 0; j < nattrs; j++)
    {
        int r = attrs[j].reg;
        if (r == 0xff || RNULL(L, r))
            hasnull = true;
    }
    uint32_t t_hoff = KERN_TUPITEM_T_BITS_OFFSET + (hasnull ? XPU_BITMAPLEN(nattrs) : 0);
    t_hoff = (uint32_t)xalign(8, t_hoff);
    for (int j = 0; j < nattrs; j++)
    {
        xattr pa = attrs[j];
        if (pa.reg == 0xff || RNULL(L, pa.reg))
            continue;
        t_hoff = (uint32_t)xalign(pa.attalign, t_hoff);
        if (pa.width > 0)
            t_hoff += pa.width;
        else if (pa.domain == XD_NUM)
            t_hoff += num_reg_varlena_size(L, pa.reg);
        else if (TXT_RAW(pa.reg))
            t_hoff += varsize_any((const uint8_t *)(uintptr_t)L.R[pa.reg]);
        else
            t_hoff += 4 + (uint32_t)L.R[pa.reg + 1];
    }
    return t_hoff + ROWID_SIZE;
}

__device__ __forceinline__ void
form_tuple_write(const xattr *attrs, int nattrs, const Lane &L, char *buf, uint32_t unit,
                 uint32_t rowid, uint32_t hash)
{
    bool hasnull = false;
    for (int j = 0; j < nattrs; j++)
    {
        int r = attrs[j].reg;
        if (r == 0xff || RNULL(L, r))
            hasnull = true;
    }
    uint32_t t_hoff = KERN_TUPITEM_T_BITS_OFFSET + (hasnull ? XPU_BITMAPLEN(nattrs) : 0);
    t_hoff = (uint32_t)xalign(8, t_hoff);
    /* the tuple unit (MAXALIGN'ed) is cleared 8 bytes at a time */
    for (uint32_t i = 0; i < unit; i += 8)
        *(uint64_t *)(buf + i) = 0;
    uint16_t infomask = hasnull ? HEAP_HASNULL : 0;
    uint8_t *bits = (uint8_t *)buf + KERN_TUPITEM_T_BITS_OFFSET;
    for (int j = 0; j < nattrs; j++)
    {
        xattr pa = attrs[j];
        if (pa.reg == 0xff || RNULL(L, pa.reg))
            continue;
        t_hoff = (uint32_t)xalign(pa.attalign, t_hoff);
        char *p = buf + t_hoff;
        int64_t v = L.R[pa.reg];
        switch (pa.width)
        {
            case 1: *(int8_t *)p = (int8_t)v; t_hoff += 1; break;
            case 2: *(int16_t *)p = (int16_t)v; t_hoff += 2; break;
            case 4: *(int32_t *)p = (int32_t)v; t_hoff += 4; break;
            case 8: *(int64_t *)p = v; t_hoff += 8; break;
            default:
            {
                const uint8_t *s = (const uint8_t *)(uintptr_t)v;
                if (pa.domain == XD_NUM)
                {
                    /* only a datum still in its varlena form addresses
                     * bytes; a parsed value holds its integer there */
                    if ((L.R[pa.reg + 2] & 0xff) == XNUM_VARLENA && L.R[pa.reg + 1] < 0 &&
                        s[0] == 0x01)
                        infomask |= HEAP_HASEXTERNAL;
                    t_hoff += num_reg_varlena_write(L, pa.reg, (uint8_t *)p);
                    infomask |= HEAP_HASVARWIDTH;
                    break;
                }
                if (TXT_RAW(pa.reg))
                {
                    /* a compressed or external datum, as it is
                     * (xpu_text_datum_write, xpu_textlib.cu:319-325) */
                    const uint32_t sz = varsize_any(s);
                    for (uint32_t i = 0; i < sz; i++)
                        p[i] = s[i];
                    t_hoff += sz;
                    if (s[0] == 0x01)
                        infomask |= HEAP_HASEXTERNAL;
                    infomask |= HEAP_HASVARWIDTH;
                    break;
                }
                uint32_t len = (uint32_t)L.R[pa.reg + 1];
                *(uint32_t *)p = (len + 4) << 2;
                for (uint32_t i = 0; i < len; i++)
                    p[4 + i] = s[i];
                t_hoff += 4 + len;
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
        }
        if (hasnull)
            bits[j >> 3] |= (uint8_t)(1 << (j & 7));
    }
    t_hoff += ROWID_SIZE;
    kern_tupitem *t = (kern_tupitem *)buf;
    t->t_len = t_hoff;
    t->hash = hash;
    t->t_infomask2 = (uint16_t)(nattrs & HEAP_NATTS_MASK);
    t->t_infomask = infomask;
    t->t_hoff = (uint8_t)((uint32_t)xalign(8, KERN_TUPITEM_T_BITS_OFFSET +
                                          (hasnull ? XPU_BITMAPLEN(nattrs) : 0)) +
                          MINIMAL_TUPLE_OFFSET);
    *(uint32_t *)(buf + t_hoff - ROWID_SIZE) = rowid;
}

__device__ __forceinline__ uint32_t
proj_tuple_size(cplan *P, const Lane &L)
{
    return form_tuple_size(P->proj, (int)P->proj_nattrs, L);
}

__device__ __forceinline__ void
proj_tuple_write(cplan *P, const Lane &L, char *buf, uint32_t unit, uint32_t rowid)
{
    /* rows of a pinned inner buffer carry their join-key hash */
    form_tuple_write(P->proj, (int)P->proj_nattrs, L, buf, unit, rowid,
                     P->r_projhash ? (uint32_t)L.R[P->r_projhash] : 0U);
}

/* ------------------------------------------------------------------
 * CPU fallback: a row whose depth-0 work (scan quals, the depth-1 join key
 * hash, and without joins its projection or group-by arguments) raised
 * "compressed or external" becomes a kern_fallbackitem of the session's
 * KDS_FORMAT_FALLBACK buffer: depth 0, l_state 0, not matched, and a
 * minimal tuple of the scan slot with the row's depth-0 variables
 * (HandleErrorIfCpuFallback and __writeOutCpuFallbackTuple,
 * xpu_common.cu:1832-1965).  The backend re-runs it from its WHERE clause
 * (execCpuFallbackOneTuple, fallback.c:441-470), which the reference also
 * does for the depth it records; such a row produces nothing on the GPU.
 * ------------------------------------------------------------------ */
#define XFB_ITEM_HEAD   16      /* offsetof(kern_fallbackitem, t) */

__device__ __forceinline__ bool
fallback_write(cplan *P, const Lane &L, kern_data_store *kfb)
{
    const int nattrs = (int)P->fb_nattrs;
    const uint32_t req = (uint32_t)xalign(8, XFB_ITEM_HEAD + form_tuple_size(P->fb, nattrs, L));
    const unsigned long long usage =
        atomicAdd((unsigned long long *)&kfb->usage, (unsigned long long)req) + req;
    const uint64_t head = KDS_HEAD_LENGTH(kfb);
    uint32_t cur = __hip_atomic_load(&kfb->nitems, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t old;
    do
    {
        old = cur;
        if (head + 8ULL * ((uint64_t)old + 1) + usage > kfb->length)
            return false;
        cur = atomicCAS(&kfb->nitems, old, old + 1);
    } while (cur != old);
    char *item = (char *)kfb + kfb->length - usage;
    *(uint64_t *)item = 0;                  /* l_state */
    *(uint64_t *)(item + 8) = 0;            /* depth 0, not matched */
    form_tuple_write(P->fb, nattrs, L, item + XFB_ITEM_HEAD, req - XFB_ITEM_HEAD, old, 0U);
    ((uint64_t *)((char *)kfb + head))[old] = usage - XFB_ITEM_HEAD;
    return true;
}

/* true if the lane's error sent its row to the CPU-fallback buffer (the
 * error is then cleared); a full buffer fails the task */
__device__ __forceinline__ bool
row_fallback(cplan *P, Lane &L, const xpu_kparams &kp, bool write)
{
    if (L.err != XMSG_TEXT_EXTERNAL + 1 || kp.kds_fallback == nullptr || P->fb_nattrs == 0)
        return false;
    if (write && !fallback_write(P, L, kp.kds_fallback))
        record_error(kp.ctl, XMSG_FALLBACK_FULL, ERRCODE_DEVICE_ERROR);
    L.err = 0;
    return true;
}

/* ------------------------------------------------------------------
 * group-by: kds_final group tuples (cuda_gpupreagg.cu:18-223, 1112-1191)
 * ------------------------------------------------------------------ */
__device__ __forceinline__ int64_t
val_as_i64(int dom, int64_t raw)
{
    if (dom == XD_F32) return (int64_t)as_f32(raw);
    if (dom == XD_F64) return (int64_t)as_f64(raw);
    return raw;
}

__device__ __forceinline__ double
val_as_f64(int dom, int64_t raw)
{
    if (dom == XD_F32) return (double)as_f32(raw);
    if (dom == XD_F64) return as_f64(raw);
    return (double)raw;
}

/* payload offset of grouping key k inside a group tuple, or -1 if NULL;
 * keys follow the states and NULL keys still take their alignment
 * (cuda_gpupreagg.cu:44-70) */
__device__ __forceinline__ uint32_t
gb_tuple_size(cplan *P, const Lane &L)
{
    uint32_t t_hoff = P->gb_t_hoff + P->gb_states_end;
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action != KAGG_ACTION__VREF)
            continue;
        xattr ka = P->keys[ag.key_index];
        t_hoff = (uint32_t)xalign(ka.attalign, t_hoff);
        if (!RNULL(L, ka.reg))
            t_hoff += (ka.width > 0) ? ka.width
                    : (ka.domain == XD_NUM) ? num_reg_varlena_size(L, ka.reg)
                    : 4 + (uint32_t)L.R[ka.reg + 1];
    }
    return t_hoff + ROWID_SIZE;
}

__device__ __forceinline__ bool
gb_keys_match(cplan *P, const Lane &L, const kern_tupitem *t)
{
    const char *base = (const char *)t;
    uint32_t pos = P->gb_t_hoff + P->gb_states_end;
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action != KAGG_ACTION__VREF)
            continue;
        xattr ka = P->keys[ag.key_index];
        pos = (uint32_t)xalign(ka.attalign, pos);
        bool tnull = !((t->t_bits[j >> 3] >> (j & 7)) & 1);
        bool lnull = RNULL(L, ka.reg);
        if (tnull || lnull)
        {
            if (tnull != lnull)
                return false;   /* NULLs group together (kmode_compare_nulls) */
            continue;
        }
        const char *p = base + pos;
        int64_t v = L.R[ka.reg];
        switch (ka.width)
        {
            case 1: if (*(const int8_t *)p != (int8_t)v) return false; pos += 1; break;
            case 2: if (*(const int16_t *)p != (int16_t)v) return false; pos += 2; break;
            case 4:
                if (ka.domain == XD_F32)
                {
                    if (!fcmp(XC_EQ, (double)__int_as_float(*(const int32_t *)p), (double)as_f32(v)))
                        return false;
                }
                else if (*(const int32_t *)p != (int32_t)v)
                    return false;
                pos += 4;
                break;
            case 8:
                if (ka.domain == XD_F64)
                {
                    if (!fcmp(XC_EQ, __longlong_as_double(*(const int64_t *)p), as_f64(v)))
                        return false;
                }
                else if (*(const int64_t *)p != v)
                    return false;
                pos += 8;
                break;
            default:
            {
                if (ka.domain == XD_NUM)
                {
                    /* numeric keys group by value (xpu_numeric_datum_comp) */
                    xnum x, y;
                    if (num_parse((const uint8_t *)p, &x) || num_read(L, ka.reg, &y) ||
                        num_compare(x, y) != 0)
                        return false;
                    pos += varsize_any((const uint8_t *)p);
                    break;
                }
                uint32_t len = (uint32_t)L.R[ka.reg + 1];
                uint32_t hdr = *(const uint32_t *)p;
                if (((hdr >> 2) & 0x3fffffff) != len + 4)
                    return false;
                const uint8_t *s = (const uint8_t *)(uintptr_t)v;
                for (uint32_t i = 0; i < len; i++)
                    if ((uint8_t)p[4 + i] != s[i])
                        return false;
                pos += 4 + len;
                break;
            }
        }
    }
    return true;
}

__device__ __forceinline__ void
gb_tuple_write(cplan *P, const Lane &L, kern_hashitem *hi, uint32_t unit,
               uint32_t hash, uint32_t rowid, uint64_t next)
{
    char *ubuf = (char *)hi;
    for (uint32_t i = 0; i < unit; i += 8)
        *(uint64_t *)(ubuf + i) = 0;
    hi->next = next;
    char *buf = (char *)&hi->t;
    kern_tupitem *t = &hi->t;
    uint16_t infomask = HEAP_HASNULL;
    uint8_t *bits = t->t_bits;
    uint32_t nattrs = P->gb_nattrs;
    uint32_t t_hoff = P->gb_t_hoff;
    /* partial states, in their initial values */
    for (int j = 0; j < (int)nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF)
            continue;
        char *s = buf + t_hoff + ag.state_off;
        bits[j >> 3] |= (uint8_t)(1 << (j & 7));
        switch (ag.action)
        {
            case KAGG_ACTION__NROWS_ANY:
            case KAGG_ACTION__NROWS_COND:
                break;
            case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMIN_INT64:
                *(int32_t *)s = ag.state_size << 2;
                *(int64_t *)(s + 8) = INT64_MAX;
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PMAX_INT32: case KAGG_ACTION__PMAX_INT64:
                *(int32_t *)s = ag.state_size << 2;
                *(int64_t *)(s + 8) = INT64_MIN;
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PMIN_FP64:
                *(int32_t *)s = ag.state_size << 2;
                *(double *)(s + 8) = 1.7976931348623157e308;
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PMAX_FP64:
                *(int32_t *)s = ag.state_size << 2;
                *(double *)(s + 8) = -1.7976931348623157e308;
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PSUM_NUMERIC: case KAGG_ACTION__PAVG_NUMERIC:
                /* attrs: the sum's scale (cuda_gpupreagg.cu:173-183) */
                *(int32_t *)s = ag.state_size << 2;
                *(uint32_t *)(s + 4) = ag.num_scale;
                infomask |= HEAP_HASVARWIDTH;
                break;
            default:
                *(int32_t *)s = ag.state_size << 2;
                infomask |= HEAP_HASVARWIDTH;
                break;
        }
    }
    /* grouping keys */
    uint32_t pos = t_hoff + P->gb_states_end;
    for (int j = 0; j < (int)nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action != KAGG_ACTION__VREF)
            continue;
        xattr ka = P->keys[ag.key_index];
        pos = (uint32_t)xalign(ka.attalign, pos);
        if (RNULL(L, ka.reg))
            continue;
        bits[j >> 3] |= (uint8_t)(1 << (j & 7));
        char *p = buf + pos;
        int64_t v = L.R[ka.reg];
        switch (ka.width)
        {
            case 1: *(int8_t *)p = (int8_t)v; pos += 1; break;
            case 2: *(int16_t *)p = (int16_t)v; pos += 2; break;
            case 4: *(int32_t *)p = (int32_t)v; pos += 4; break;
            case 8: *(int64_t *)p = v; pos += 8; break;
            default:
            {
                if (ka.domain == XD_NUM)
                {
                    pos += num_reg_varlena_write(L, ka.reg, (uint8_t *)p);
                    infomask |= HEAP_HASVARWIDTH;
                    break;
                }
                uint32_t len = (uint32_t)L.R[ka.reg + 1];
                const uint8_t *s = (const uint8_t *)(uintptr_t)v;
                *(uint32_t *)p = (len + 4) << 2;
                for (uint32_t i = 0; i < len; i++)
                    p[4 + i] = s[i];
                pos += 4 + len;
                break;
            }
        }
    }
    pos += ROWID_SIZE;
    t->t_len = pos;
    t->hash = hash;
    t->t_infomask2 = (uint16_t)(nattrs & HEAP_NATTS_MASK);
    t->t_infomask = infomask;
    t->t_hoff = (uint8_t)(t_hoff + MINIMAL_TUPLE_OFFSET);
    *(uint32_t *)(buf + pos - ROWID_SIZE) = rowid;
}

/* LDS partial table: struct of arrays over n entries (layout in xplan) */
struct LdsTable
{
    unsigned long long *keys;   /* [n] packed keys, XLDS_EMPTY_KEY if free */
    unsigned long long *f64;    /* [lds_n64][n] */
    uint32_t   *rows;           /* [n] rows aggregated into the entry */
    uint32_t   *f32;            /* [lds_n32][n] */
    uint32_t    n;
};

__device__ __forceinline__ LdsTable
lds_table_at(cplan *P, unsigned long long *base, uint32_t n)
{
    LdsTable T;
    T.n = n;
    T.keys = base;
    T.f64 = base + n;
    T.rows = (uint32_t *)(T.f64 + (uint64_t)P->lds_n64 * n);
    T.f32 = T.rows + n;
    return T;
}

/* A row may use the LDS table when every aggregate argument is non-NULL and
 * every FILTER is true: then one row counter stands for count(*), count(x)
 * and the n of every sum and average.  The argument of a 128-bit sum must
 * also lie in [-2^32, 2^32), the range the table's two-field sum holds. */
__device__ __forceinline__ bool
gb_row_fast(cplan *P, const Lane &L)
{
    bool ok = true;
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF)
            continue;
        if (ag.filter_reg != 0xff)
            ok = ok && !RNULL(L, ag.filter_reg) && L.R[ag.filter_reg] != 0;
        if (ag.arg_reg != 0xff)
            ok = ok && !RNULL(L, ag.arg_reg);
        if ((ag.action == KAGG_ACTION__PSUM_INT64 || ag.action == KAGG_ACTION__PAVG_INT64) &&
            ag.arg_reg != 0xff)
        {
            const int64_t hi = val_as_i64(ag.arg_domain, L.R[ag.arg_reg]) >> 32;
            ok = ok && (hi == 0 || hi == -1);
        }
    }
    return ok;
}

/* merge one contribution (a row, or an LDS entry) into a group's states
 * (the state updates of cuda_gpupreagg.cu:1196-1494, with the int128 sum as
 * a low-word atomic plus carry into the high word, xpu_common.h:3659-3689) */
__device__ __forceinline__ void
gb_merge_global(cplan *P, const Lane &L, kern_tupitem *t, const LdsTable *T,
                int lds_slot)
{
    char *payload = (char *)t + P->gb_t_hoff;
    const bool from_lds = (lds_slot >= 0);
    const uint32_t rows = from_lds ? T->rows[lds_slot] : 0;
    if (from_lds && rows == 0)
        return;
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF)
            continue;
        char *s = payload + ag.state_off;
        const unsigned long long *F64 = from_lds && ag.lds64 >= 0
            ? T->f64 + (uint64_t)ag.lds64 * T->n + lds_slot : nullptr;
        const uint32_t *F32 = from_lds && ag.lds32 >= 0
            ? T->f32 + (uint64_t)ag.lds32 * T->n + lds_slot : nullptr;
        bool has = from_lds;
        int64_t iv = 0;
        double fv = 0.0;
        if (!from_lds)
        {
            bool pass = (ag.filter_reg == 0xff) ||
                        (!RNULL(L, ag.filter_reg) && L.R[ag.filter_reg] != 0);
            if (!pass)
                continue;
            if (ag.arg_reg != 0xff && !RNULL(L, ag.arg_reg))
            {
                has = true;
                iv = val_as_i64(ag.arg_domain, L.R[ag.arg_reg]);
                fv = val_as_f64(ag.arg_domain, L.R[ag.arg_reg]);
            }
        }
        const unsigned long long n = from_lds ? rows : (has ? 1ULL : 0ULL);
        switch (ag.action)
        {
            case KAGG_ACTION__NROWS_ANY:
            case KAGG_ACTION__NROWS_COND:
            {
                unsigned long long c = from_lds ? rows :
                    ((ag.action == KAGG_ACTION__NROWS_ANY || has) ? 1ULL : 0ULL);
                if (c)
                    atomicAdd((unsigned long long *)s, c);
                break;
            }
            case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMIN_INT64:
            case KAGG_ACTION__PMAX_INT32: case KAGG_ACTION__PMAX_INT64:
            {
                if (!has)
                    break;
                long long v = !from_lds ? (long long)iv :
                    (F32 ? (long long)(int32_t)F32[0] : (long long)F64[0]);
                atomicOr((unsigned int *)(s + 4), __PAGG_MINMAX_ATTRS__VALID);
                if (ag.action == KAGG_ACTION__PMIN_INT32 || ag.action == KAGG_ACTION__PMIN_INT64)
                    atomicMin((long long *)(s + 8), v);
                else
                    atomicMax((long long *)(s + 8), v);
                break;
            }
            case KAGG_ACTION__PMIN_FP64: case KAGG_ACTION__PMAX_FP64:
            {
                if (!has)
                    break;
                double v = from_lds ? __longlong_as_double((long long)F64[0]) : fv;
                atomicOr((unsigned int *)(s + 4), __PAGG_MINMAX_ATTRS__VALID);
                if (ag.action == KAGG_ACTION__PMIN_FP64)
                    atomicMin((double *)(s + 8), v);
                else
                    atomicMax((double *)(s + 8), v);
                break;
            }
            case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
            {
                if (!n)
                    break;
                unsigned long long sum = from_lds ? F64[0] : (unsigned long long)iv;
                atomicAdd((unsigned long long *)(s + 8), n);
                atomicAdd((unsigned long long *)(s + 16), sum);
                break;
            }
            case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
            {
                if (!n)
                    break;
                /* an LDS entry holds the sum of the values' low 32-bit
                 * halves and the count of negative values (every value lies
                 * in [-2^32, 2^32)): lo - count * 2^32 as 128 bits */
                unsigned long long lo, hi;
                if (from_lds)
                {
                    const unsigned long long sl = F64[0];
                    const unsigned long long sub = (unsigned long long)F32[0] << 32;
                    lo = sl - sub;
                    hi = sl < sub ? ~0ULL : 0ULL;
                }
                else
                {
                    lo = (unsigned long long)iv;
                    hi = iv < 0 ? ~0ULL : 0ULL;
                }
                atomicAdd((unsigned long long *)(s + 8), n);
                unsigned long long old = atomicAdd((unsigned long long *)(s + 16), lo);
                hi += (old + lo < old) ? 1ULL : 0ULL;
                if (hi)
                    atomicAdd((unsigned long long *)(s + 24), hi);
                break;
            }
            case KAGG_ACTION__PSUM_NUMERIC: case KAGG_ACTION__PAVG_NUMERIC:
            {
                /* __update_groupby__psum_numeric (cuda_gpupreagg.cu:1408-1446):
                 * a finite value rescaled to the sum's scale; NaN and the
                 * infinities set their flag (the argument was parsed by
                 * XO_NUM_PARSE, so no error is left here) */
                if (!has)
                    break;
                xnum x;
                if (num_read(L, ag.arg_reg, &x))
                    break;
                if (x.kind == XNUM_VALID)
                {
                    const xi128 v = num_rescale(x.v, ag.num_scale, x.scale);
                    const unsigned long long lo = (unsigned long long)(xu128)v;
                    unsigned long long hi = (unsigned long long)((xu128)v >> 64);
                    atomicAdd((unsigned long long *)(s + 8), 1ULL);
                    const unsigned long long old = atomicAdd((unsigned long long *)(s + 16), lo);
                    hi += (old + lo < old) ? 1ULL : 0ULL;
                    if (hi)
                        atomicAdd((unsigned long long *)(s + 24), hi);
                }
                else
                    atomicOr((unsigned int *)(s + 4), num_special_flag(x.kind));
                break;
            }
            case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
            {
                if (!n)
                    break;
                double sum = from_lds ? __longlong_as_double((long long)F64[0]) : fv;
                atomicAdd((unsigned long long *)(s + 8), n);
                atomicAdd((double *)(s + 16), sum);
                break;
            }
            case KAGG_ACTION__STDDEV:
            {
                if (!n)
                    break;
                double sx = from_lds ? __longlong_as_double((long long)F64[0]) : fv;
                double sxx = from_lds ? __longlong_as_double((long long)F64[T->n]) : fv * fv;
                atomicAdd((unsigned long long *)(s + 8), n);
                atomicAdd((double *)(s + 16), sx);
                atomicAdd((double *)(s + 24), sxx);
                break;
            }
        }
    }
}

__device__ __forceinline__ kern_hashitem *
hashitem_at(kern_data_store *kds, uint64_t off)
{
    return (kern_hashitem *)((char *)kds + kds->length - off);
}

/* ------------------------------------------------------------------
 * join-hash Bloom filter of one inner relation
 *
 * Register-blocked (Putze, Sanders, Singler, "Cache-, Hash- and
 * Space-Efficient Bloom Filters", WEA 2007): a key selects one 64-bit word
 * and sets four bits in it, so a probe is one 8-byte load.  The key is the
 * 32-bit join hash that the bucket walk compares (kern_hashitem.t.hash):
 * every hash that can find an item has its bits set, so a clear bit proves
 * that the walk would end without a candidate.  The mixing function is the
 * 64-bit finalizer of MurmurHash3 (Appleby, public domain).
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint64_t
bloom_mix(uint32_t h)
{
    uint64_t x = (uint64_t)h;
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

/* word index from the low bits (mask < 2^32), bit positions from bits 40..63 */
__device__ __forceinline__ uint64_t
bloom_bits(uint64_t x)
{
    return (1ULL << ((x >> 40) & 63)) | (1ULL << ((x >> 46) & 63)) |
           (1ULL << ((x >> 52) & 63)) | (1ULL << (x >> 58));
}

__device__ __forceinline__ bool
bloom_maybe(const uint64_t *__restrict__ words, uint32_t mask, uint32_t h)
{
    uint64_t x = bloom_mix(h);
    uint64_t m = bloom_bits(x);
    return (words[x & mask] & m) == m;
}

/* Sets the bits of every item reachable from the hash slots [s0, s1), and
 * raises *maxdup to the most items of one chain that share a join hash (a
 * probe finds at most that many candidates, since items of one hash share a
 * chain); a chain of more than 64 items counts as its length, which bounds
 * the same from above. */
__device__ __forceinline__ void
bloom_build_range(kern_data_store *ik, uint64_t *words, uint32_t mask, uint64_t s0, uint64_t s1,
                  uint32_t *maxdup)
{
    const uint64_t *slots = (const uint64_t *)((const char *)ik + KDS_HEAD_LENGTH(ik));
    for (uint64_t s = s0; s < s1; s++)
    {
        const uint64_t head = slots[s];
        uint64_t off = head;
        uint64_t steps = (uint64_t)ik->nitems + 1;     /* as many as a probe inspects */
        uint32_t len = 0;
        while (off != 0 && off != ~0ULL && off < ik->length && steps-- > 0)
        {
            const kern_hashitem *hi = hashitem_at(ik, off);
            uint64_t x = bloom_mix(hi->t.hash);
            atomicOr((unsigned long long *)&words[x & mask], (unsigned long long)bloom_bits(x));
            off = hi->next;
            len++;
        }
        uint32_t most = len;
        if (len <= 64)
        {
            most = 0;
            uint64_t a = head;
            for (uint32_t i = 0; i < len; i++)
            {
                const kern_hashitem *ha = hashitem_at(ik, a);
                uint32_t same = 0;
                uint64_t c = a;
                for (uint32_t j = i; j < len; j++)
                {
                    const kern_hashitem *hc = hashitem_at(ik, c);
                    same += (hc->t.hash == ha->t.hash);
                    c = hc->next;
                }
                most = same > most ? same : most;
                a = ha->next;
            }
        }
        if (most)
            atomicMax(maxdup, most);
    }
}

/* Reserves a group tuple of `unit` bytes (a multiple of 128) in a
 * KDS_FORMAT_HASH buffer, whose tuples grow from the tail and row index
 * from the head.  Every unit is at least 128 bytes, so tuples ending `end`
 * bytes from the tail number at most end/128, and the test keeps room for
 * their row-index entries too.  Space is taken with one atomic add: the
 * reservations that pass the test are those ending below a bound, a prefix
 * of the atomic order, so a failed one leaves no hole among the tuples; only
 * kds->usage runs past them (the host recomputes it from the row index
 * before it grows the buffer).  The row id is taken after the space, so the
 * row index has no hole either (the reference instead locks nitems around
 * both, cuda_gpupreagg.cu:1141-1168). */
__device__ __forceinline__ kern_hashitem *
gb_alloc(kern_data_store *kds, uint32_t unit, uint32_t *p_rowid)
{
    const uint64_t fixed = KDS_HEAD_LENGTH(kds) + 8ULL * kds->hash_nslots;
    const unsigned long long end =
        atomicAdd((unsigned long long *)&kds->usage, (unsigned long long)unit) + unit;
    if (fixed + end + 8ULL * (end / XPU_L1_CACHELINE_SZ) > kds->length)
        return nullptr;
    *p_rowid = atomicAdd(&kds->nitems, 1U);
    return hashitem_at(kds, end);
}

/* merges the partial states of group tuple src into group tuple dst (both
 * of this plan's layout, xpu_common.h:2400-2471): what the final merge of
 * partial results does per group (gpu_service.c:5429-5569) */
__device__ __forceinline__ void
gb_merge_tuple(cplan *P, kern_tupitem *dst, const kern_tupitem *src)
{
    char *d = (char *)dst + P->gb_t_hoff;
    const char *s = (const char *)src + P->gb_t_hoff;
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF)
            continue;
        char *ds = d + ag.state_off;
        const char *ss = s + ag.state_off;
        switch (ag.action)
        {
            case KAGG_ACTION__NROWS_ANY:
            case KAGG_ACTION__NROWS_COND:
            {
                const unsigned long long c = *(const unsigned long long *)ss;
                if (c)
                    atomicAdd((unsigned long long *)ds, c);
                break;
            }
            case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMIN_INT64:
            case KAGG_ACTION__PMAX_INT32: case KAGG_ACTION__PMAX_INT64:
            {
                if (!(*(const uint32_t *)(ss + 4) & __PAGG_MINMAX_ATTRS__VALID))
                    break;
                const long long v = *(const long long *)(ss + 8);
                atomicOr((unsigned int *)(ds + 4), __PAGG_MINMAX_ATTRS__VALID);
                if (ag.action == KAGG_ACTION__PMIN_INT32 || ag.action == KAGG_ACTION__PMIN_INT64)
                    atomicMin((long long *)(ds + 8), v);
                else
                    atomicMax((long long *)(ds + 8), v);
                break;
            }
            case KAGG_ACTION__PMIN_FP64: case KAGG_ACTION__PMAX_FP64:
            {
                if (!(*(const uint32_t *)(ss + 4) & __PAGG_MINMAX_ATTRS__VALID))
                    break;
                const double v = *(const double *)(ss + 8);
                atomicOr((unsigned int *)(ds + 4), __PAGG_MINMAX_ATTRS__VALID);
                if (ag.action == KAGG_ACTION__PMIN_FP64)
                    atomicMin((double *)(ds + 8), v);
                else
                    atomicMax((double *)(ds + 8), v);
                break;
            }
            case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
            {
                const unsigned long long n = *(const unsigned long long *)(ss + 8);
                if (!n)
                    break;
                atomicAdd((unsigned long long *)(ds + 8), n);
                atomicAdd((unsigned long long *)(ds + 16), *(const unsigned long long *)(ss + 16));
                break;
            }
            case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
            {
                const unsigned long long n = *(const unsigned long long *)(ss + 8);
                if (!n)
                    break;
                const unsigned long long lo = *(const unsigned long long *)(ss + 16);
                unsigned long long hi = *(const unsigned long long *)(ss + 24);
                atomicAdd((unsigned long long *)(ds + 8), n);
                const unsigned long long old = atomicAdd((unsigned long long *)(ds + 16), lo);
                hi += (old + lo < old) ? 1ULL : 0ULL;
                if (hi)
                    atomicAdd((unsigned long long *)(ds + 24), hi);
                break;
            }
            case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
            {
                const unsigned long long n = *(const unsigned long long *)(ss + 8);
                if (!n)
                    break;
                atomicAdd((unsigned long long *)(ds + 8), n);
                atomicAdd((double *)(ds + 16), *(const double *)(ss + 16));
                break;
            }
            case KAGG_ACTION__PSUM_NUMERIC: case KAGG_ACTION__PAVG_NUMERIC:
            {
                const uint32_t special = *(const uint32_t *)(ss + 4) & __PAGG_NUMERIC_ATTRS__MASK;
                if (special)
                    atomicOr((unsigned int *)(ds + 4), special);
                const unsigned long long n = *(const unsigned long long *)(ss + 8);
                if (!n)
                    break;
                const unsigned long long lo = *(const unsigned long long *)(ss + 16);
                unsigned long long hi = *(const unsigned long long *)(ss + 24);
                atomicAdd((unsigned long long *)(ds + 8), n);
                const unsigned long long old = atomicAdd((unsigned long long *)(ds + 16), lo);
                hi += (old + lo < old) ? 1ULL : 0ULL;
                if (hi)
                    atomicAdd((unsigned long long *)(ds + 24), hi);
                break;
            }
            case KAGG_ACTION__STDDEV:
            {
                const unsigned long long n = *(const unsigned long long *)(ss + 8);
                if (!n)
                    break;
                atomicAdd((unsigned long long *)(ds + 8), n);
                atomicAdd((double *)(ds + 16), *(const double *)(ss + 16));
                atomicAdd((double *)(ds + 24), *(const double *)(ss + 24));
                break;
            }
        }
    }
}

/* loads the grouping keys of group tuple t into their registers: the
 * inverse of the key part of gb_tuple_write (a text key's register
 * addresses its bytes inside t) */
__device__ __forceinline__ void
gb_tuple_keys(cplan *P, Lane &L, const kern_tupitem *t)
{
    const char *base = (const char *)t;
    uint32_t pos = P->gb_t_hoff + P->gb_states_end;
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action != KAGG_ACTION__VREF)
            continue;
        xattr ka = P->keys[ag.key_index];
        pos = (uint32_t)xalign(ka.attalign, pos);
        const bool tnull = !((t->t_bits[j >> 3] >> (j & 7)) & 1);
        SETNULL(L, ka.reg, tnull);
        if (tnull)
            continue;
        const char *p = base + pos;
        switch (ka.width)
        {
            case 1: L.R[ka.reg] = *(const int8_t *)p; pos += 1; break;
            case 2: L.R[ka.reg] = *(const int16_t *)p; pos += 2; break;
            case 4:
                L.R[ka.reg] = ka.domain == XD_F32 ? (int64_t)*(const uint32_t *)p
                                                  : (int64_t)*(const int32_t *)p;
                pos += 4;
                break;
            case 8: L.R[ka.reg] = *(const int64_t *)p; pos += 8; break;
            default:
            {
                if (ka.domain == XD_NUM)
                {
                    num_store_varlena(L, ka.reg, (const uint8_t *)p, false);
                    pos += varsize_any((const uint8_t *)p);
                    break;
                }
                const uint32_t len = ((*(const uint32_t *)p >> 2) & 0x3fffffff) - 4;
                L.R[ka.reg] = (int64_t)(uintptr_t)(p + 4);
                L.R[ka.reg + 1] = len;
                pos += 4 + len;
                break;
            }
        }
    }
}

/* the group tuple of the key registers on the chain from head, or null */
__device__ __forceinline__ kern_hashitem *
gb_chain_find(cplan *P, const Lane &L, kern_data_store *kds, uint64_t head, uint32_t hash)
{
    for (uint64_t off = head; off != 0; )
    {
        kern_hashitem *hi = hashitem_at(kds, off);
        const uint32_t h2 = __hip_atomic_load(&hi->t.hash, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (h2 == hash && gb_keys_match(P, L, &hi->t))
            return hi;
        off = __hip_atomic_load(&hi->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return nullptr;
}

/* Find-or-insert the group of the key registers of every pending lane and
 * merge its contribution: the row in the registers, LDS entry lds_slot of
 * T, or the partial states of group tuple src.  A wave-level retry loop in
 * which a lane either finds the group, or takes the slot lock once and
 * inserts it, so lanes of one wave never wait on each other's locks (the
 * reference does the same at block level, cuda_gpupreagg.cu:1607-1731).
 * A group that finds no room in kds_final goes to the overflow table
 * kds_ovf, when there is one: once a reservation in kds_final has failed,
 * every later one fails too, so a group lives in exactly one of the two
 * tables.  Without room in either, the contribution is dropped and the task
 * fails with ERRCODE_SUSPEND_NO_SPACE. */
__device__ __forceinline__ void
gb_global(cplan *P, Lane &L, kern_data_store *kds_final, uint32_t hash, bool pending,
          const LdsTable *T, int lds_slot, xpu_task_ctl *ctl,
          kern_data_store *kds_ovf = nullptr, const kern_tupitem *src = nullptr)
{
    if (!wave_any(pending))
        return;
    uint32_t iter = 0;

    if (P->nkeys == 0)
    {
        /* no GROUP BY: the single group tuple of a ROW-format kds_final */
        if (pending)
        {
            uint64_t *rowidx = (uint64_t *)((char *)kds_final + KDS_HEAD_LENGTH(kds_final));
            kern_tupitem *t = (kern_tupitem *)((char *)kds_final + kds_final->length - rowidx[0]);
            if (src)
                gb_merge_tuple(P, t, src);
            else
                gb_merge_global(P, L, t, T, lds_slot);
        }
        return;
    }
    kern_data_store *kds = kds_final;
    uint64_t *slot = (uint64_t *)((char *)kds + KDS_HEAD_LENGTH(kds)) + (hash % kds->hash_nslots);
    while (wave_any(pending))
    {
        if (pending)
        {
            /* Slots and the fields of published group tuples are read with
             * relaxed device-coherent loads: the inserting lane writes a tuple
             * into a fresh 128-byte unit no other XCD has cached and publishes
             * it with a release store, so no per-row acquire (an L2
             * invalidate on CDNA) is needed on the lookup path. */
            uint64_t head = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            kern_hashitem *found = nullptr;
            if (head != XGB_LOCKED_SLOT)
            {
                found = gb_chain_find(P, L, kds, head, hash);
                if (!found)
                {
                    /* A miss is confirmed after an acquire: the relaxed
                     * walk reads a group's key bytes with plain loads, which
                     * may hit a stale copy of its unit in this XCD's L2
                     * (the unit was zero-filled before the group was
                     * inserted from another XCD), and a false miss would
                     * insert the group twice.  The acquire pairs with the
                     * release that published `head`; misses are rare next
                     * to hits, so lookups stay fence-free. */
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    found = gb_chain_find(P, L, kds, head, hash);
                }
                if (!found &&
                    atomicCAS((unsigned long long *)slot, (unsigned long long)head,
                              (unsigned long long)XGB_LOCKED_SLOT) == head)
                {
                    /* slot locked: nobody inserted since `head` was read */
                    uint32_t tsz = gb_tuple_size(P, L);
                    uint32_t unit = (uint32_t)xalign(XPU_L1_CACHELINE_SZ,
                                                     offsetof(kern_hashitem, t) + tsz);
                    uint32_t rowid = 0;
                    kern_hashitem *hi = gb_alloc(kds, unit, &rowid);
                    if (!hi)
                    {
                        __hip_atomic_store(slot, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (kds_ovf && kds != kds_ovf)
                        {
                            kds = kds_ovf;
                            slot = (uint64_t *)((char *)kds + KDS_HEAD_LENGTH(kds)) +
                                   (hash % kds->hash_nslots);
                        }
                        else
                        {
                            record_error(ctl, XMSG_FINAL_NO_SPACE, ERRCODE_SUSPEND_NO_SPACE);
                            pending = false;
                        }
                    }
                    else
                    {
                        gb_tuple_write(P, L, hi, unit, hash, rowid, head);
                        uint64_t *rowidx = (uint64_t *)((char *)kds + KDS_HEAD_LENGTH(kds)) +
                                           kds->hash_nslots;
                        rowidx[rowid] = (uint64_t)((char *)kds + kds->length - (char *)&hi->t);
                        __hip_atomic_store(slot, (uint64_t)((char *)kds + kds->length - (char *)hi),
                                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                        if (kds != kds_final)
                            ctl->final_ovf = 1U;
                        found = hi;
                    }
                }
            }
            if (found)
            {
                if (src)
                    gb_merge_tuple(P, &found->t, src);
                else
                    gb_merge_global(P, L, &found->t, T, lds_slot);
                pending = false;
            }
        }
        if (++iter > (1U << 24))
        {
            if (pending)
                record_error(ctl, XMSG_GROUP_RETRY, ERRCODE_DEVICE_ERROR);
            break;
        }
    }
}

/* the groups of the overflow table kds_ovf, merged into kds_final (grown by
 * the host to hold them): one lane per overflow group */
template <int NR>
__device__ __forceinline__ void
final_merge_ovf_body(cplan *P, kern_data_store *kds_final, const kern_data_store *kds_ovf,
                     xpu_task_ctl *ctl)
{
    int64_t Rfile[NR];
    Lane L;
    L.R = Rfile;
    L.nmask = ~1U;
    L.err = 0;
    L.dead = 0;
#pragma unroll
    for (int i = 0; i < NR; i++)
        L.R[i] = 0;
    L.R[XP_REG_TRUE] = 1;
    const uint32_t n = kds_ovf->nitems;
    const uint64_t *rowidx = (const uint64_t *)((const char *)kds_ovf + KDS_HEAD_LENGTH(kds_ovf)) +
                             kds_ovf->hash_nslots;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const uint32_t i = base + threadIdx.x;
        const bool pending = i < n;
        const kern_tupitem *t = nullptr;
        uint32_t hash = 0;
        if (pending)
        {
            t = (const kern_tupitem *)((const char *)kds_ovf + kds_ovf->length - rowidx[i]);
            gb_tuple_keys(P, L, t);
            hash = t->hash;
        }
        gb_global(P, L, kds_final, hash, pending, nullptr, -1, ctl, nullptr, t);
    }
}

__device__ __forceinline__ uint64_t
gb_pack_key(cplan *P, const Lane &L)
{
    uint64_t pk = 0;
    int shift = 0;
    for (int k = 0; k < (int)P->nkeys; k++)
    {
        int bits = P->key_bits[k];
        int r = P->keys[k].reg;
        uint64_t v = RNULL(L, r) ? 0 : ((uint64_t)L.R[r] & ((bits >= 64) ? ~0ULL : ((1ULL << bits) - 1)));
        pk |= v << shift;
        shift += bits;
    }
    for (int k = 0; k < (int)P->nkeys; k++)
        pk |= (uint64_t)RNULL(L, P->keys[k].reg) << (shift + k);
    return pk;
}

__device__ __forceinline__ void
gb_unpack_key(cplan *P, Lane &L, uint64_t pk)
{
    int shift = 0;
    for (int k = 0; k < (int)P->nkeys; k++)
        shift += P->key_bits[k];
    int vshift = 0;
    for (int k = 0; k < (int)P->nkeys; k++)
    {
        int bits = P->key_bits[k];
        int r = P->keys[k].reg;
        uint64_t v = (pk >> vshift) & ((bits >= 64) ? ~0ULL : ((1ULL << bits) - 1));
        int64_t sv;
        if (P->keys[k].domain == XD_INT && bits < 64)
            sv = (int64_t)(v << (64 - bits)) >> (64 - bits);     /* sign extend */
        else
            sv = (int64_t)v;
        L.R[r] = sv;
        SETNULL(L, r, (pk >> (shift + k)) & 1);
        vshift += bits;
    }
}

/* PG hash of the grouping keys held in registers (HashValue semantics) */
__device__ __forceinline__ uint32_t
gb_hash_regs(cplan *P, const Lane &L)
{
    uint32_t h = 0xffffffffU;
    for (int k = 0; k < (int)P->nkeys; k++)
    {
        xattr ka = P->keys[k];
        if (RNULL(L, ka.reg))
            continue;
        uint32_t x;
        if (ka.domain == XD_NUM)
        {
            /* by value, as HashValue hashes a numeric: equal values of
             * different images share a slot.  A datum that does not parse
             * has sent its row to the CPU fallback at its load. */
            xnum v;
            x = num_read(L, ka.reg, &v) ? 0 : num_hash(v);
        }
        else if (ka.domain == XD_TEXT)
            x = hash_bytes((const uint8_t *)(uintptr_t)L.R[ka.reg], (uint32_t)L.R[ka.reg + 1]);
        else
            x = hash_fixed((uint64_t)L.R[ka.reg], ka.width);
        h = hash_merge(h, x);
    }
    return h ^ 0xffffffffU;
}

__device__ __forceinline__ uint32_t
lds_hash(uint64_t k)
{
    k *= 0x9E3779B97F4A7C15ULL;
    return (uint32_t)(k >> 32);
}

/* returns the LDS slot of packed key pk, inserting it; -1 if the probe fails */
__device__ __forceinline__ int
lds_probe(LdsTable &T, uint64_t pk)
{
    uint32_t h = (uint32_t)(((uint64_t)lds_hash(pk) * T.n) >> 32);
    for (int i = 0; i < 8; i++)
    {
        unsigned long long cur = T.keys[h];
        if (cur == pk)
            return (int)h;
        if (cur == XLDS_EMPTY_KEY)
        {
            unsigned long long old = atomicCAS(&T.keys[h], XLDS_EMPTY_KEY, (unsigned long long)pk);
            if (old == XLDS_EMPTY_KEY || old == pk)
                return (int)h;
        }
        if (++h == T.n)
            h = 0;
    }
    return -1;
}

/* adds a row that passed gb_row_fast into LDS entry slot */
__device__ __forceinline__ void
lds_update(cplan *P, const Lane &L, LdsTable &T, int slot)
{
    atomicAdd(&T.rows[slot], 1U);
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF || ag.arg_reg == 0xff)
            continue;
        unsigned long long *F64 = T.f64 + (uint64_t)ag.lds64 * T.n + slot;
        uint32_t *F32 = T.f32 + (uint64_t)ag.lds32 * T.n + slot;
        int64_t raw = L.R[ag.arg_reg];
        switch (ag.action)
        {
            case KAGG_ACTION__PMIN_INT32:
                atomicMin((int *)F32, (int)val_as_i64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PMAX_INT32:
                atomicMax((int *)F32, (int)val_as_i64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PMIN_INT64:
                atomicMin((long long *)F64, (long long)val_as_i64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PMAX_INT64:
                atomicMax((long long *)F64, (long long)val_as_i64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PMIN_FP64:
                atomicMin((double *)F64, val_as_f64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PMAX_FP64:
                atomicMax((double *)F64, val_as_f64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
                atomicAdd(F64, (unsigned long long)val_as_i64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
            {
                /* v in [-2^32, 2^32) (gb_row_fast): the sum of the low
                 * 32-bit halves and the count of negative values hold the
                 * 128-bit sum exactly below 2^32 rows, with no returning
                 * atomic */
                int64_t v = val_as_i64(ag.arg_domain, raw);
                atomicAdd(F64, (unsigned long long)(uint32_t)v);
                if (v < 0)
                    atomicAdd(F32, 1U);
                break;
            }
            case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
                atomicAdd((double *)F64, val_as_f64(ag.arg_domain, raw));
                break;
            case KAGG_ACTION__STDDEV:
            {
                double v = val_as_f64(ag.arg_domain, raw);
                atomicAdd((double *)F64, v);
                atomicAdd((double *)(F64 + T.n), v * v);
                break;
            }
        }
    }
}

__device__ void
lds_init(cplan *P, LdsTable &T)
{
    for (uint32_t i = threadIdx.x; i < T.n; i += blockDim.x)
    {
        T.keys[i] = XLDS_EMPTY_KEY;
        T.rows[i] = 0;
    }
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF)
            continue;
        unsigned long long init64 = 0;
        uint32_t init32 = 0;
        switch (ag.action)
        {
            case KAGG_ACTION__PMIN_INT32: init32 = (uint32_t)INT32_MAX; break;
            case KAGG_ACTION__PMAX_INT32: init32 = (uint32_t)INT32_MIN; break;
            case KAGG_ACTION__PMIN_INT64: init64 = (unsigned long long)INT64_MAX; break;
            case KAGG_ACTION__PMAX_INT64: init64 = (unsigned long long)INT64_MIN; break;
            /* as the reference's initial states (cuda_gpupreagg.cu:18-223) */
            case KAGG_ACTION__PMIN_FP64:
                init64 = (unsigned long long)__double_as_longlong(1.7976931348623157e308); break;
            case KAGG_ACTION__PMAX_FP64:
                init64 = (unsigned long long)__double_as_longlong(-1.7976931348623157e308); break;
        }
        if (ag.lds64 >= 0)
        {
            int nf = (ag.action == KAGG_ACTION__STDDEV) ? 2 : 1;
            for (int f = 0; f < nf; f++)
            {
                unsigned long long *A = T.f64 + (uint64_t)(ag.lds64 + f) * T.n;
                for (uint32_t i = threadIdx.x; i < T.n; i += blockDim.x)
                    A[i] = init64;
            }
        }
        if (ag.lds32 >= 0)
        {
            uint32_t *A = T.f32 + (uint64_t)ag.lds32 * T.n;
            for (uint32_t i = threadIdx.x; i < T.n; i += blockDim.x)
                A[i] = init32;
        }
    }
}

/* merges every entry of the block's LDS table into kds_final (after a
 * barrier that ends all updates of the table) */
template <int NR>
__device__ __forceinline__ void
lds_flush(cplan *P, LdsTable &T, kern_data_store *kds_final, xpu_task_ctl *ctl,
          kern_data_store *kds_ovf = nullptr)
{
    int64_t Rfile[NR];
    Lane L;
    L.R = Rfile;
    L.nmask = ~1U;
    L.err = 0;
    L.dead = 0;
#pragma unroll
    for (int i = 0; i < NR; i++)
        L.R[i] = 0;
    L.R[XP_REG_TRUE] = 1;
    uint32_t nflush = 0;
    for (uint32_t base = 0; base < T.n; base += blockDim.x)
    {
        uint32_t h = base + threadIdx.x;
        bool pending = false;
        uint32_t hash = 0;
        if (h < T.n)
        {
            unsigned long long key = T.keys[h];
            if (key != XLDS_EMPTY_KEY)
            {
                gb_unpack_key(P, L, key);
                hash = gb_hash_regs(P, L);
                pending = true;
                nflush++;
            }
        }
        gb_global(P, L, kds_final, hash, pending, &T, pending ? (int)h : -1, ctl, kds_ovf);
    }
    if (nflush)
        atomicAdd((unsigned long long *)&ctl->global_merges, (unsigned long long)nflush);
}

/* ------------------------------------------------------------------
 * partial-state dumps: instead of merging its LDS entries into kds_final
 * with device atomics (about nine per entry), a block writes them as
 * records, hash-partitioned, and one block per partition later folds a
 * partition's records in LDS and merges each group once.  A record is the
 * LDS entry: packed key, u64 fields, row count, u32 fields.
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t
dump_partition(uint64_t pk, uint32_t nparts)
{
    uint64_t h = pk * 0x9FB21C651E98DF25ULL;
    return (uint32_t)(((h >> 32) * (uint64_t)nparts) >> 32);
}

__device__ __forceinline__ int
dump_rec_words(cplan *P)
{
    return 1 + P->lds_n64 + (1 + P->lds_n32 + 1) / 2;
}

/* writes the block's table as partial-state records (entries beyond a
 * partition's capacity merge into kds_final directly) */
template <int NR>
__device__ void
lds_dump(cplan *P, LdsTable &T, const xpu_kparams &kp)
{
    __shared__ uint32_t s_cnt[XP_DUMP_MAXPARTS];
    __shared__ uint64_t s_base[XP_DUMP_MAXPARTS];
    const uint32_t np = kp.dump_nparts;
    const int rw = dump_rec_words(P);
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x)
        s_cnt[i] = 0;
    __syncthreads();
    for (uint32_t h = threadIdx.x; h < T.n; h += blockDim.x)
    {
        unsigned long long key = T.keys[h];
        if (key != XLDS_EMPTY_KEY && T.rows[h] != 0)
            atomicAdd(&s_cnt[dump_partition(key, np)], 1U);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x)
    {
        uint32_t c = s_cnt[i];
        s_base[i] = c ? atomicAdd((unsigned long long *)&kp.dump_next[i], (unsigned long long)c) : 0;
        s_cnt[i] = 0;
    }
    __syncthreads();
    int64_t Rfile[NR];
    Lane L;
    L.R = Rfile;
    L.nmask = ~1U;
    L.err = 0;
    L.dead = 0;
#pragma unroll
    for (int i = 0; i < NR; i++)
        L.R[i] = 0;
    L.R[XP_REG_TRUE] = 1;
    uint32_t ndirect = 0;
    for (uint32_t base = 0; base < T.n; base += blockDim.x)
    {
        const uint32_t h = base + threadIdx.x;
        bool pending = false;
        uint32_t hash = 0;
        if (h < T.n)
        {
            unsigned long long key = T.keys[h];
            if (key != XLDS_EMPTY_KEY && T.rows[h] != 0)
            {
                const uint32_t part = dump_partition(key, np);
                const uint64_t idx = s_base[part] + atomicAdd(&s_cnt[part], 1U);
                if (idx < kp.dump_cap)
                {
                    uint64_t *rec = kp.dump + ((uint64_t)part * kp.dump_cap + idx) * rw;
                    rec[0] = key;
                    for (int f = 0; f < (int)P->lds_n64; f++)
                        rec[1 + f] = T.f64[(uint64_t)f * T.n + h];
                    uint32_t *r32 = (uint32_t *)(rec + 1 + P->lds_n64);
                    r32[0] = T.rows[h];
                    for (int f = 0; f < (int)P->lds_n32; f++)
                        r32[1 + f] = T.f32[(uint64_t)f * T.n + h];
                }
                else
                {
                    gb_unpack_key(P, L, key);
                    hash = gb_hash_regs(P, L);
                    pending = true;
                    ndirect++;
                }
            }
        }
        gb_global(P, L, kp.kds_out, hash, pending, &T, pending ? (int)h : -1, kp.ctl, kp.kds_ovf);
    }
    if (ndirect)
        atomicAdd((unsigned long long *)&kp.ctl->global_merges, (unsigned long long)ndirect);
}

/* folds a partial-state record into LDS entry slot */
__device__ __forceinline__ void
lds_merge_entry(cplan *P, LdsTable &T, int slot, const uint64_t *rec)
{
    const uint32_t *r32 = (const uint32_t *)(rec + 1 + P->lds_n64);
    atomicAdd(&T.rows[slot], r32[0]);
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        xagg ag = P->aggs[j];
        if (ag.action == KAGG_ACTION__VREF)
            continue;
        unsigned long long *F64 = T.f64 + (uint64_t)ag.lds64 * T.n + slot;
        uint32_t *F32 = T.f32 + (uint64_t)ag.lds32 * T.n + slot;
        const uint64_t v64 = ag.lds64 >= 0 ? rec[1 + ag.lds64] : 0;
        const uint32_t v32 = ag.lds32 >= 0 ? r32[1 + ag.lds32] : 0;
        switch (ag.action)
        {
            case KAGG_ACTION__PMIN_INT32:
                atomicMin((int *)F32, (int)v32);
                break;
            case KAGG_ACTION__PMAX_INT32:
                atomicMax((int *)F32, (int)v32);
                break;
            case KAGG_ACTION__PMIN_INT64:
                atomicMin((long long *)F64, (long long)v64);
                break;
            case KAGG_ACTION__PMAX_INT64:
                atomicMax((long long *)F64, (long long)v64);
                break;
            case KAGG_ACTION__PMIN_FP64:
                atomicMin((double *)F64, __longlong_as_double((long long)v64));
                break;
            case KAGG_ACTION__PMAX_FP64:
                atomicMax((double *)F64, __longlong_as_double((long long)v64));
                break;
            case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
                atomicAdd(F64, (unsigned long long)v64);
                break;
            case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
                atomicAdd(F64, (unsigned long long)v64);
                atomicAdd(F32, v32);
                break;
            case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
                atomicAdd((double *)F64, __longlong_as_double((long long)v64));
                break;
            case KAGG_ACTION__STDDEV:
                atomicAdd((double *)F64, __longlong_as_double((long long)v64));
                atomicAdd((double *)(F64 + T.n), __longlong_as_double((long long)rec[2 + ag.lds64]));
                break;
        }
    }
}

/* one block per partition: fold the partition's records in LDS, merge each
 * group into kds_final (a table that fills up is merged and restarted) */
template <int NR, class PROG>
__device__ __forceinline__ void
dump_merge_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    LdsTable T = lds_table_at(P, lds_dyn, kp.lds_nents);
    if (T.n == 0 || kp.dump == nullptr)
        return;
    const int rw = dump_rec_words(P);
    for (uint32_t part = blockIdx.x; part < kp.dump_nparts; part += gridDim.x)
    {
        lds_init(P, T);
        __syncthreads();
        const uint64_t reserved = __hip_atomic_load(&kp.dump_next[part], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t n = reserved < kp.dump_cap ? reserved : kp.dump_cap;
        const uint64_t *recs = kp.dump + (uint64_t)part * kp.dump_cap * rw;
        for (uint64_t i0 = 0; i0 < n; i0 += blockDim.x)
        {
            const uint64_t i = i0 + threadIdx.x;
            const uint64_t *rec = recs + i * rw;
            bool fail = false;
            if (i < n)
            {
                int slot = lds_probe(T, rec[0]);
                if (slot >= 0)
                    lds_merge_entry(P, T, slot, rec);
                else
                    fail = true;
            }
            while (__syncthreads_or(fail))
            {
                /* the table is full: merge it, start over, then fold the
                 * records that found no room (each round places at least
                 * one of them in the emptied table) */
                lds_flush<NR>(P, T, kp.kds_out, kp.ctl, kp.kds_ovf);
                __syncthreads();
                lds_init(P, T);
                __syncthreads();
                if (fail)
                {
                    int slot = lds_probe(T, rec[0]);
                    if (slot >= 0)
                    {
                        lds_merge_entry(P, T, slot, rec);
                        fail = false;
                    }
                }
            }
        }
        __syncthreads();
        lds_flush<NR>(P, T, kp.kds_out, kp.ctl, kp.kds_ovf);
        __syncthreads();
    }
}

/* ------------------------------------------------------------------
 * staged columns: a wave's loads of the columns the plan stages
 * (XLD_STAGED LDCOLs) for 64*RPT consecutive rows, issued together before
 * any loaded value is used, so a lane keeps nstage*RPT loads in flight
 * instead of one dependent load per column per row.  fill() sets a row's
 * registers with the NULL rules of XO_LDCOL; the program then runs with
 * STAGED=true, which skips the staged LDCOLs.
 * ------------------------------------------------------------------ */
template <int RPT>
__device__ __forceinline__ uint32_t
pick_u32(const uint32_t (&a)[RPT], int i)
{
    uint32_t v = a[0];
#pragma unroll
    for (int k = 1; k < RPT; k++)
        v = (i == k) ? a[k] : v;
    return v;
}

template <int RPT>
struct StagedCols
{
    static_assert(XP_MAX_STAGE * RPT <= 32, "validity bits of a tile fit one word");
    /* geometry of the staged columns (uniform: scalar registers) */
    const char *kbase;
    int         nstage;
    const char *vbase[XP_MAX_STAGE];
    uint64_t    vlen[XP_MAX_STAGE], nmoff[XP_MAX_STAGE], nmlen[XP_MAX_STAGE];
    int         w[XP_MAX_STAGE], reg[XP_MAX_STAGE];
    bool        f32[XP_MAX_STAGE];
    /* the rows [base, base + 64*RPT): raw words, and the lane's validity
     * bits (bit st*RPT + i: row i of column st is not NULL by its bitmap);
     * 'full': every row of the tile lies inside every staged column */
    uint64_t    base;
    bool        full;
    uint32_t    lo[XP_MAX_STAGE][RPT], hi[XP_MAX_STAGE][RPT];
    uint32_t    vbits;

    __device__ __forceinline__ void
    setup(cplan *P, const kern_data_store *kds)
    {
        kbase = (const char *)kds;
        ckds *kmeta = (ckds *)(uintptr_t)kds;
        nstage = (int)P->nstage;
#pragma unroll
        for (int st = 0; st < XP_MAX_STAGE; st++)
        {
            vbase[st] = kbase;
            vlen[st] = nmoff[st] = nmlen[st] = 0;
            w[st] = 1;
            reg[st] = 0;
            f32[st] = false;
            if (st < nstage)
            {
                const xstage sd = P->stage[st];
                ccolmeta *cm = &kmeta->colmeta[sd.col];
                vbase[st] = kbase + cm->values_offset;
                vlen[st] = cm->values_length;
                nmoff[st] = cm->nullmap_offset;
                nmlen[st] = cm->nullmap_length;
                w[st] = sd.width;
                reg[st] = sd.reg;
                f32[st] = sd.f32 != 0;
            }
        }
    }

    /* A tile that lies inside every staged column (all but the last tile)
     * loads without per-lane bounds checks, so no load waits on another; the
     * tail tile checks each element. */
    __device__ __forceinline__ void
    load(uint64_t tile_base, uint64_t row_end)
    {
        const int lane = __lane_id();
        const uint64_t wrows = 64ULL * RPT;
        base = tile_base;
        full = base + wrows <= row_end;
#pragma unroll
        for (int st = 0; st < XP_MAX_STAGE; st++)
            if (st < nstage)
                full = full && (uint64_t)w[st] * (base + wrows) <= vlen[st];
#pragma unroll
        for (int st = 0; st < XP_MAX_STAGE; st++)
        {
#pragma unroll
            for (int i = 0; i < RPT; i++)
            {
                lo[st][i] = 0;
                hi[st][i] = 0;
            }
        }
        if (full)
        {
#pragma unroll
            for (int st = 0; st < XP_MAX_STAGE; st++)
            {
                if (st < nstage)
                {
                    const char *vb = vbase[st];
                    const int ww = w[st];
#pragma unroll
                    for (int i = 0; i < RPT; i++)
                    {
                        const char *p = vb + (uint64_t)ww * (base + (uint64_t)i * 64 + lane);
                        if (ww == 8)
                        {
                            uint2 u = *(const uint2 *)p;
                            lo[st][i] = u.x;
                            hi[st][i] = u.y;
                        }
                        else if (ww == 4)
                            lo[st][i] = *(const uint32_t *)p;
                        else if (ww == 2)
                            lo[st][i] = (uint32_t)(int32_t)*(const int16_t *)p;
                        else
                            lo[st][i] = (uint32_t)(int32_t)*(const int8_t *)p;
                    }
                }
            }
        }
        else
        {
#pragma unroll
            for (int st = 0; st < XP_MAX_STAGE; st++)
            {
                if (st < nstage)
                {
                    const char *vb = vbase[st];
                    const uint64_t vl = vlen[st];
                    const int ww = w[st];
#pragma unroll
                    for (int i = 0; i < RPT; i++)
                    {
                        const uint64_t row = base + (uint64_t)i * 64 + lane;
                        if (row < row_end && (uint64_t)ww * (row + 1) <= vl)
                        {
                            const char *p = vb + (uint64_t)ww * row;
                            if (ww == 8)
                            {
                                uint2 u = *(const uint2 *)p;
                                lo[st][i] = u.x;
                                hi[st][i] = u.y;
                            }
                            else if (ww == 4)
                                lo[st][i] = *(const uint32_t *)p;
                            else if (ww == 2)
                                lo[st][i] = (uint32_t)(int32_t)*(const int16_t *)p;
                            else
                                lo[st][i] = (uint32_t)(int32_t)*(const int8_t *)p;
                        }
                    }
                }
            }
        }
        /* validity bitmaps: one 64-bit word per 64 rows, read with scalar
         * loads when the tile is 64-row aligned inside the bitmap, else one
         * byte per lane; kept as one bit per (column, row batch) */
        const bool nm_words = ((base & 63) == 0);
        vbits = ~0U;
#pragma unroll
        for (int st = 0; st < XP_MAX_STAGE; st++)
        {
            const uint64_t nm_off = nmoff[st];
            if (st >= nstage || !nm_off)
                continue;
#pragma unroll
            for (int i = 0; i < RPT; i++)
            {
                const uint64_t byte = (base + (uint64_t)i * 64) >> 3;
                uint32_t bit;
                if (nm_words && (nm_off & 7) == 0 && byte + 8 <= nmlen[st])
                {
                    const uint64_t word = *(const XCONST uint64_t *)(uintptr_t)(kbase + nm_off + byte);
                    bit = (uint32_t)(word >> lane) & 1;
                }
                else
                {
                    const uint64_t row = base + (uint64_t)i * 64 + lane;
                    const uint64_t idx = row >> 3;
                    bit = idx < nmlen[st] ? (((const uint8_t *)kbase)[nm_off + idx] >> (row & 7)) & 1 : 0;
                }
                vbits &= ~((bit ^ 1U) << (st * RPT + i));
            }
        }
    }

    /* the staged registers of row base + 64*i + lane */
    __device__ __forceinline__ void
    fill(Lane &L, int i, uint64_t row, bool valid) const
    {
        const int lane = __lane_id();
#pragma unroll
        for (int st = 0; st < XP_MAX_STAGE; st++)
        {
            if (st < nstage)
            {
                const int ww = w[st];
                const uint32_t l = pick_u32<RPT>(lo[st], i);
                const uint32_t h = pick_u32<RPT>(hi[st], i);
                /* a full tile needs no per-row bounds test */
                bool isnull = full ? false : (!valid || (uint64_t)ww * (row + 1) > vlen[st]);
                isnull = isnull || !((vbits >> (st * RPT + i)) & 1);
                int64_t v = (ww == 8) ? (int64_t)(((uint64_t)h << 32) | l)
                                      : (f32[st] ? (int64_t)l : (int64_t)(int32_t)l);
                L.R[reg[st]] = isnull ? 0 : v;
                SETNULL(L, reg[st], isnull);
            }
        }
    }
};

/* ------------------------------------------------------------------
 * decoupled look-back over tiles (thread 0 of the block)
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint64_t
lb_pack(uint64_t flag, uint64_t rows, uint64_t bytes)
{
    return (flag << XLB_FLAG_SHIFT) | ((rows & XLB_ROWS_MASK) << 32) | ((bytes >> 3) & 0xffffffffULL);
}

/* ------------------------------------------------------------------
 * the pipeline kernel
 * ------------------------------------------------------------------ */
#define IT_END      1U          /* chain exhausted (offsets are 8-aligned, never 1) */

/* how a projection launch finds its tiles' output offsets */
enum
{
    PM_LOOKBACK = 0,    /* one kernel: decoupled look-back over tile status words */
    PM_COUNT = 1,       /* count pass: tile totals only, no output */
    PM_WRITE = 2,       /* write pass: offsets from the scanned tile totals */
};


/* a lane's registers before a row: all NULL but R0, and TRUE */
template <int NR>
__device__ __forceinline__ void
lane_reset(Lane &L)
{
    L.nmask = ~1U;
    L.err = 0;
    L.dead = 0;
#pragma unroll
    for (int i = 0; i < NR; i++)
        L.R[i] = 0;
    L.R[XP_REG_TRUE] = 1;
}

/* The scan quals of one row and, for a row that passes them and with
 * 'hash', the depth-1 key hash (*h); an error is reported and fails the
 * row. */
template <int NR, class PROG, bool STAGED>
__device__ __forceinline__ bool
screen_row(cplan *P, const xpu_kparams &kp, Lane &L, uint32_t row, bool hash, uint32_t *h)
{
    PROG::template run<STAGED>(P, P->seg_scan.begin, P->seg_scan.end, L, kp.kds_src, row, kp);
    if (L.err)
    {
        if (!row_fallback(P, L, kp, true))
            record_error(kp.ctl, L.err - 1, ERRCODE_DEVICE_ERROR);
        return false;
    }
    if (RNULL(L, P->r_scan_ok) || L.R[P->r_scan_ok] == 0)
        return false;
    if (hash)
    {
        PROG::template run<STAGED>(P, P->depth[1].seg_hash.begin, P->depth[1].seg_hash.end, L,
                                   kp.kds_src, row, kp);
        if (L.err)
        {
            if (!row_fallback(P, L, kp, true))
                record_error(kp.ctl, L.err - 1, ERRCODE_DEVICE_ERROR);
            return false;
        }
        *h = (uint32_t)L.R[P->r_hash];
    }
    return true;
}

/*
 * The pipeline kernel: scan quals, the joins of every depth, then either the
 * group-by (PREAGG) or the projection into kds_dst.
 *
 * A lane runs one outer row through a small state machine of phases (scan,
 * then per depth: bucket walk, join quals, outer-join fill) whose segments
 * run only for the lanes they apply to; the wave loops until every lane has
 * exhausted its bucket chains.
 *
 * Group-by joins first screen the rows: the scan quals (and, with
 * kp.prefilter, the depth-1 key hash) run for every row, and a row goes on,
 * through a per-wave LDS queue, only if it passes the quals and its hash may
 * be in the depth-1 relation (its Bloom filter, kp.bloom[1]).  The joins then
 * run over 64 queued rows at a time, so a wave's bucket walks are not spent
 * on rows that the quals or a selective first join reject.  Rows of an outer
 * join at depth 1 are not screened by the filter.
 */
template <bool PREAGG, int NR, class PROG, int PM = PM_LOOKBACK>
__device__ __forceinline__ void
pipeline_body(cplan *P, const xpu_kparams &kp)
{
    /* statistics are counted once: by the count pass of a two-pass projection */
    constexpr bool COUNTS = (PM != PM_WRITE);
    extern __shared__ unsigned long long lds_dyn[];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_stop;
    __shared__ uint64_t s_wave_cnt[XP_BLOCK_MAX / 64];
    __shared__ uint64_t s_wave_bytes[XP_BLOCK_MAX / 64];
    __shared__ uint64_t s_base_rows;
    __shared__ uint64_t s_base_bytes;
    __shared__ uint32_t s_depth_out[XP_MAX_DEPTH + 1];
    __shared__ uint32_t s_queue[PREAGG ? (XP_BLOCK_MAX / 64) * XQ_WAVE : 1];

    const kern_data_store *__restrict__ kds = kp.kds_src;
    xpu_task_ctl *ctl = kp.ctl;
    const int tid = threadIdx.x;
    const int nrels = P->nrels;
    const int lane = __lane_id();
    const int wid = tid >> 6;
    const int nwaves = blockDim.x >> 6;

    LdsTable T = lds_table_at(P, lds_dyn, PREAGG ? kp.lds_nents : 0);
    if (PREAGG && T.n > 0)
    {
        lds_init(P, T);
        __syncthreads();
    }

    uint32_t cnt_in = 0, cnt_out = 0, lds_miss = 0;
    if (tid <= XP_MAX_DEPTH)
        s_depth_out[tid] = 0;
    __syncthreads();

    /* queue of screened rows (group-by joins only) */
    const bool qmode = PREAGG && nrels > 0;
    const bool qbloom = qmode && kp.prefilter != 0 && kp.bloom[1] != nullptr &&
                        !P->depth[1].left_outer;
    uint32_t *wq = &s_queue[PREAGG ? wid * XQ_WAVE : 0];

    /* projection output of the current tile */
    uint64_t lane_cnt = 0, lane_bytes = 0, lane_wr_rows = 0, lane_wr_bytes = 0;
    bool stop_writes = false;

    /* one outer row through the joins and the final stage; 'queued' rows
     * passed the scan quals (counted) and the depth-1 filter already.
     * A sweep runs, for depth 1..nrels, the bucket walk, the join quals of
     * the current candidate and the outer-join fill, then the final stage;
     * the wave sweeps again while a lane still has candidates.  The
     * specialized kernel unrolls the depths (nrels is a constant there), so
     * every segment and chain offset is addressed statically. */
    auto run_row = [&](Lane &L, const uint32_t row, const bool valid, const int pass,
                       const bool queued) {
        lane_reset<NR>(L);
        int cd = 0;
        uint32_t it[XP_MAX_DEPTH + 2];   /* chain offsets (inner KDS < 4GB) */
        uint32_t matched = 0;
#pragma unroll
        for (int i = 0; i < XP_MAX_DEPTH + 2; i++)
            it[i] = 0;

        /* an error ends the lane's row and is reported for the task, or in
         * the row's depth-0 work (fb) sends it to the CPU fallback: written
         * by the pass that keeps the row's output, dropped by a count pass */
        auto run_seg = [&](const xseg seg, const bool act, const bool fb = false) -> bool {
            if (!act)
                return false;
            PROG::template run<false>(P, seg.begin, seg.end, L, kds, row, kp);
            if (L.err)
            {
                if (!(fb && row_fallback(P, L, kp, PM != PM_COUNT && pass == 0)))
                    record_error(ctl, L.err - 1, ERRCODE_DEVICE_ERROR);
                L.err = 0;
                return true;
            }
            return false;
        };
        /* drops the exhausted depths d, d-1, ... of the current candidate */
        auto unwind = [&](const int d) {
            for (int dd = d; dd >= 1; dd--)
            {
                if (dd == cd && it[dd] == IT_END)
                {
                    matched &= ~(1U << dd);
                    it[dd] = 0;
                    cd--;
                }
            }
        };
        /* the next item of inner relation d whose hash is h, from offset off */
        auto chain_find = [&](const int d, uint64_t off, const uint32_t h) -> uint32_t {
            kern_data_store *ik = (kern_data_store *)(uintptr_t)kp.inner_kds[d];
            uint32_t steps = ik->nitems;    /* chains are acyclic: bound the walk */
            while (off != 0 && off != ~0ULL && off < ik->length &&
                   hashitem_at(ik, off)->t.hash != h && steps-- > 0)
                off = hashitem_at(ik, off)->next;
            return (off == 0 || off == ~0ULL || off >= ik->length) ? IT_END : (uint32_t)off;
        };
        auto depth_step = [&](const int d) {
            kern_data_store *ik = (kern_data_store *)(uintptr_t)kp.inner_kds[d];
            /* bucket walk to the first candidate */
            {
                const bool act = (cd == d && it[d] == 0);
                const bool err = run_seg(P->depth[d].seg_hash, act, d == 1);
                if (act && !err)
                {
                    const uint32_t h = (uint32_t)L.R[P->r_hash];
                    const uint64_t *bf = kp.bloom[d];
                    uint64_t head = 0;
                    if (!bf || (queued && d == 1) || bloom_maybe(bf, kp.bloom_mask[d], h))
                        head = ((const uint64_t *)((const char *)ik + KDS_HEAD_LENGTH(ik)))[h % ik->hash_nslots];
                    it[d] = chain_find(d, head, h);
                }
                if (err)
                    cd = 0;
            }
            /* join quals of the candidate; the next candidate is found now so
             * that exhausted chains unwind without another sweep */
            {
                const bool act = (cd == d && it[d] > IT_END);
                if (act)
                {
                    L.R[P->r_tuple] = (int64_t)(uintptr_t)&hashitem_at(ik, it[d])->t;
                    SETNULL(L, P->r_tuple, 0);
                }
                const bool err = run_seg(P->depth[d].seg_join, act);
                if (act)
                {
                    const kern_hashitem *cur = hashitem_at(ik, it[d]);
                    const int64_t status = err ? 0 : L.R[P->r_status];
                    it[d] = chain_find(d, cur->next, cur->t.hash);
                    if (status != 0)
                        matched |= 1U << d;
                    if (status > 0)
                    {
                        if (pass == 0 && COUNTS)
                            atomicAdd(&s_depth_out[d], 1U);
                        cd = d + 1;
                        it[d + 1] = 0;
                    }
                }
                if (err)
                    cd = 0;
            }
            /* outer-join fill of a row without matches */
            {
                const bool act = (cd == d && it[d] == IT_END && P->depth[d].left_outer &&
                                  !((matched >> d) & 1));
                if (act)
                {
                    L.R[P->r_tuple] = 0;
                    SETNULL(L, P->r_tuple, 1);
                }
                const bool err = run_seg(P->depth[d].seg_oj, act);
                bool adv = false;
                if (act)
                {
                    matched |= 1U << d;
                    if (!err && L.R[P->r_status] > 0)
                    {
                        if (pass == 0 && COUNTS)
                            atomicAdd(&s_depth_out[d], 1U);
                        cd = d + 1;
                        it[d + 1] = 0;
                        adv = true;
                    }
                }
                if (!adv && cd == d && it[d] == IT_END)
                    unwind(d);
                if (err)
                    cd = 0;
            }
        };

        /* scan quals */
        {
            const bool err = run_seg(P->seg_scan, valid, true);
            const bool ok = valid && !err && !RNULL(L, P->r_scan_ok) && L.R[P->r_scan_ok] != 0;
            if (ok && pass == 0 && COUNTS && !queued)
                cnt_in++;
            cd = ok ? 1 : 0;
        }
        for (;;)
        {
            if constexpr (PROG::kUnrollRows)
            {
#pragma unroll
                for (int d = 1; d <= XP_MAX_DEPTH; d++)
                    if (d <= nrels)
                        depth_step(d);
            }
            else
            {
                for (int d = 1; d <= nrels; d++)
                    depth_step(d);
            }
            /* the final stage */
            const bool act = (cd == nrels + 1);
            const bool err = run_seg(P->seg_final, act, nrels == 0);
            if (act && !err)
            {
                if (pass == 0 && COUNTS)
                    cnt_out++;
                if (PREAGG)
                {
                    bool direct = true;
                    if (T.n > 0 && gb_row_fast(P, L))
                    {
                        int slot = lds_probe(T, gb_pack_key(P, L));
                        if (slot >= 0)
                        {
                            lds_update(P, L, T, slot);
                            direct = false;
                        }
                    }
                    if (direct)
                        lds_miss++;
                    uint32_t hash = direct ? gb_hash_regs(P, L) : 0;
                    gb_global(P, L, kp.kds_out, hash, direct, &T, -1, ctl, kp.kds_ovf);
                }
                else
                {
                    uint32_t sz = (uint32_t)xalign(8, proj_tuple_size(P, L));
                    if (pass == 0)
                    {
                        lane_cnt++;
                        lane_bytes += sz;
                    }
                    else if (!stop_writes)
                    {
                        kern_data_store *dst = kp.kds_out;
                        lane_wr_bytes += sz;
                        uint64_t usage = s_base_bytes + lane_wr_bytes;
                        uint64_t rowid = s_base_rows + lane_wr_rows;
                        char *tp = (char *)dst + dst->length - usage;
                        proj_tuple_write(P, L, tp, sz, (uint32_t)rowid);
                        uint64_t *rowidx = (uint64_t *)((char *)dst + KDS_HEAD_LENGTH(dst));
                        rowidx[rowid] = usage;
                        lane_wr_rows++;
                    }
                }
            }
            if (act)
            {
                cd = nrels;
                unwind(nrels);
            }
            if (err)
                cd = 0;
            if (!wave_any(cd > 0))
                break;
        }
    };

    /* the write pass stops at the first tile that does not fit kds_dst */
    uint32_t stop_tile = kp.ntiles;
    if (PM == PM_WRITE)
    {
        uint32_t rt = __hip_atomic_load(&ctl->resume_tile, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (rt)
            stop_tile = 0xffffffffU - rt;
    }
    uint32_t k_iter = 0;
    for (;;)
    {
        if (tid == 0)
        {
            uint32_t err = __hip_atomic_load(&ctl->errcode, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (PM == PM_LOOKBACK)
                s_tile = err ? kp.ntiles : atomicAdd(&ctl->tile_ticket, 1U);
            else
            {
                /* no inter-tile dependency: static assignment */
                uint64_t t = (uint64_t)blockIdx.x + (uint64_t)k_iter * gridDim.x;
                s_tile = (err || t >= stop_tile) ? kp.ntiles : (uint32_t)t;
            }
        }
        __syncthreads();
        const uint32_t tile = s_tile;
        k_iter++;
        if (tile >= kp.ntiles)
            break;

        /* rows of this tile: PREAGG handles rows_per_tile / blockDim rows per lane */
        const uint32_t rpt = kp.rows_per_tile / blockDim.x;
        lane_cnt = lane_bytes = lane_wr_rows = lane_wr_bytes = 0;
        stop_writes = false;

        if (qmode)
        {
            /* screen up to XQ_BATCHES row batches into the wave's queue, then
             * run the joins over the queued rows 64 at a time */
            const uint64_t tile_base = (uint64_t)kp.row_begin + (uint64_t)tile * kp.rows_per_tile;
            for (uint32_t r0 = 0; r0 < rpt; r0 += XQ_BATCHES)
            {
                const uint32_t nb = rpt - r0 < XQ_BATCHES ? rpt - r0 : XQ_BATCHES;
                uint32_t q_count = 0;      /* wave-uniform */
                auto enqueue = [&](const bool cand, const uint32_t row) {
                    const uint64_t m = __ballot(cand);
                    if (cand)
                        wq[q_count + (uint32_t)__popcll(m & ((1ULL << lane) - 1))] = row;
                    q_count += (uint32_t)__popcll(m);
                };
                for (uint32_t ri = r0; ri < r0 + nb; ri++)
                {
                    const uint64_t row64 = tile_base + (uint64_t)ri * blockDim.x + tid;
                    const bool valid = row64 < kp.row_end;
                    int64_t Rfile[NR];
                    Lane L;
                    L.R = Rfile;
                    lane_reset<NR>(L);
                    uint32_t h = 0;
                    bool ok = valid && screen_row<NR, PROG, false>(P, kp, L, (uint32_t)row64, qbloom, &h);
                    if (ok && COUNTS)
                        cnt_in++;
                    if (ok && qbloom)
                        ok = bloom_maybe(kp.bloom[1], kp.bloom_mask[1], h);
                    enqueue(ok, (uint32_t)row64);
                }
                __builtin_amdgcn_wave_barrier();
                for (uint32_t q0 = 0; q0 < q_count; q0 += 64)
                {
                    const bool has = q0 + (uint32_t)lane < q_count;
                    const uint32_t qrow = has ? wq[q0 + lane] : 0;
                    int64_t Rfile[NR];
                    Lane L;
                    L.R = Rfile;
                    run_row(L, qrow, has, 0, true);
                }
                __builtin_amdgcn_wave_barrier();
            }
            continue;
        }

        for (int pass = 0; pass < (PREAGG ? 1 : 2); pass++)
        for (uint32_t ri = 0; ri < rpt; ri++)
        {
            const uint64_t row64 = (uint64_t)kp.row_begin + (uint64_t)tile * kp.rows_per_tile +
                                   (uint64_t)ri * blockDim.x + tid;
            const bool valid = row64 < kp.row_end;
            const uint32_t row = (uint32_t)row64;

            int64_t Rfile[NR];
            Lane L;
            L.R = Rfile;

            /* a scan-only projection keeps its registers from pass 0; the
             * count pass of a two-pass projection writes nothing */
            if (!PREAGG && pass == 1 && (nrels == 0 || PM == PM_COUNT))
                break;

            run_row(L, row, valid, pass, false);

            if (!PREAGG && pass == 0)
            {
                /* ---- tile prefix: wave scan, block scan, look-back ---- */
                uint64_t ic = wave_incl_scan_u64(lane_cnt);
                uint64_t ib = wave_incl_scan_u64(lane_bytes);
                if (lane == 63)
                {
                    s_wave_cnt[wid] = ic;
                    s_wave_bytes[wid] = ib;
                }
                __syncthreads();
                if (tid == 0)
                {
                    uint64_t tc = 0, tb = 0;
                    for (int w = 0; w < nwaves; w++)
                    {
                        uint64_t c = s_wave_cnt[w], b = s_wave_bytes[w];
                        s_wave_cnt[w] = tc;
                        s_wave_bytes[w] = tb;
                        tc += c;
                        tb += b;
                    }
                    /* exclusive prefix of earlier tiles */
                    uint64_t er = 0, eb = 0;
                    bool stop = false;
                    uint64_t *st = kp.tile_status;
                    if (PM == PM_COUNT)
                    {
                        /* tile totals, scanned by xpu_tile_scan_kernel */
                        st[2ULL * tile] = tc;
                        st[2ULL * tile + 1] = tb;
                        stop = true;
                    }
                    else if (PM == PM_WRITE)
                    {
                        er = st[2ULL * tile];
                        eb = st[2ULL * tile + 1];
                        stop = (tile >= stop_tile);
                    }
                    else if (tile > 0)
                    {
                        __hip_atomic_store(&st[tile], lb_pack(XLB_AGG, tc, tb),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        int64_t j = (int64_t)tile - 1;
                        uint32_t spins = 0;
                        while (j >= 0)
                        {
                            uint64_t w = __hip_atomic_load(&st[j], __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
                            uint64_t flag = w >> XLB_FLAG_SHIFT;
                            if (flag == 0)
                            {
                                if (++spins > XLB_SPIN_LIMIT)
                                {
                                    record_error(ctl, XMSG_LOOKBACK_TIMEOUT, ERRCODE_DEVICE_ERROR);
                                    stop = true;
                                    break;
                                }
                                __builtin_amdgcn_s_sleep(1);
                                continue;
                            }
                            if (flag == XLB_STOP)
                            {
                                stop = true;
                                break;
                            }
                            er += (w >> 32) & XLB_ROWS_MASK;
                            eb += (w & 0xffffffffULL) << 3;
                            if (flag == XLB_PREFIX)
                                break;
                            j--;
                        }
                    }
                    kern_data_store *dst = kp.kds_out;
                    if (PM == PM_LOOKBACK && !stop &&
                        KDS_HEAD_LENGTH(dst) + 8 * (er + tc) + eb + tb > dst->length)
                    {
                        stop = true;
                        /* several tiles may see the overflow (a successor can read
                         * this tile's AGG word before its STOP); the resume point
                         * is the earliest: keep max(~tile) */
                        atomicMax(&ctl->resume_tile, 0xffffffffU - tile);
                        if (tile == 0)
                            record_error(ctl, XMSG_TILE_TOO_LARGE, ERRCODE_DEVICE_ERROR);
                    }
                    if (PM == PM_LOOKBACK)
                    {
                        if (stop || __hip_atomic_load(&ctl->errcode, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT) != 0)
                        {
                            stop = true;
                            __hip_atomic_store(&st[tile], lb_pack(XLB_STOP, 0, 0),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        else
                        {
                            __hip_atomic_store(&st[tile], lb_pack(XLB_PREFIX, er + tc, eb + tb),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            atomicMax((unsigned long long *)&ctl->out_nitems,
                                      (unsigned long long)(er + tc));
                            atomicMax((unsigned long long *)&ctl->out_usage,
                                      (unsigned long long)(eb + tb));
                        }
                    }
                    else if (__hip_atomic_load(&ctl->errcode, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT) != 0)
                        stop = true;
                    s_base_rows = er;
                    s_base_bytes = eb;
                    s_stop = stop;
                }
                __syncthreads();
                stop_writes = s_stop;
                /* this lane's first output row / byte offset within the tile */
                uint64_t excl_c = ic - lane_cnt + s_wave_cnt[wid];
                uint64_t excl_b = ib - lane_bytes + s_wave_bytes[wid];
                lane_wr_rows = excl_c;
                lane_wr_bytes = excl_b;
                if (!stop_writes && nrels == 0 && lane_cnt > 0)
                {
                    /* scan-only projections emit at most one row per lane and
                     * its registers are still live: write now, skip pass 1 */
                    kern_data_store *dst = kp.kds_out;
                    uint64_t usage = s_base_bytes + excl_b + lane_bytes;
                    uint64_t rowid = s_base_rows + excl_c;
                    char *tp = (char *)dst + dst->length - usage;
                    proj_tuple_write(P, L, tp, (uint32_t)lane_bytes, (uint32_t)rowid);
                    uint64_t *rowidx = (uint64_t *)((char *)dst + KDS_HEAD_LENGTH(dst));
                    rowidx[rowid] = usage;
                }
            }
        }
        if (!PREAGG)
            __syncthreads();    /* s_base_* / s_wave_* reuse by the next tile */
    }
    if (PREAGG && T.n > 0)
    {
        __syncthreads();
        lds_flush<NR>(P, T, kp.kds_out, ctl, kp.kds_ovf);
    }
    if (cnt_in)
        atomicAdd(&ctl->nitems_in, cnt_in);
    if (cnt_out)
        atomicAdd(&ctl->nitems_out, cnt_out);
    if (lds_miss)
        atomicAdd((unsigned long long *)&ctl->lds_miss, (unsigned long long)lds_miss);
    __syncthreads();
    if (tid >= 1 && tid <= nrels && s_depth_out[tid])
        atomicAdd(&ctl->depth_out[tid], s_depth_out[tid]);
}

/* ------------------------------------------------------------------
 * multi-pass joins of a group-by
 *
 * The rows of a slice are screened by one kernel (the scan quals and the
 * depth-1 Bloom filter) into level-1 records; pass d then joins the level-d
 * records with inner relation d and writes level-(d+1) records, or, at the
 * last depth, runs the final stage into the group-by.  Each pass is a
 * launch of its own, so it holds only its own registers and runs at the
 * occupancy they allow, and its bucket walks run over dense batches of 64
 * records instead of over the few lanes of a wave whose rows survived.
 * The screen reads only the columns its own program needs (the specialized
 * kernel drops the other loads); a pass reads the rest for the rows that
 * reach it.
 *
 * A level-d record holds the row and the matched item of each depth below d
 * (0: the NULL row of an outer join).  A pass restores a record's registers
 * by running the scan program for the row and the join program of each
 * depth below d with its recorded item: every register is then what the
 * single-kernel pipeline would hold at depth d.
 *
 * Records are column-major within per-block regions: block b of every pass
 * reads region b of its input and appends to region b of its output, so
 * appends reserve room with LDS atomics only.  A region of level d+1 holds
 * as many records as region b of level d times the most items that share a
 * join hash at depth d, so it cannot overflow (a full region is reported,
 * never truncated).
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t *
rec_word(uint32_t *base, uint32_t nwords, uint32_t cap, uint32_t b, uint32_t k, uint32_t r)
{
    return base + ((uint64_t)b * nwords + k) * cap + r;
}

/* The screen pass: wave tiles of 64*RPT rows (static assignment of the
 * tiles to blocks and waves), their staged loads issued together (the
 * specialized kernel keeps only the columns the screen program reads), the
 * scan quals and depth-1 filter probes of every row, then the appends of
 * the rows that pass.  Few registers: a wide block keeps many waves in
 * flight. */
template <int NR, int RPT, class PROG>
__device__ __forceinline__ void
join_screen_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    __shared__ uint32_t s_cnt;
    xpu_task_ctl *ctl = kp.ctl;
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);     /* uniform */
    const int nwaves = blockDim.x >> 6;
    const uint32_t b = blockIdx.x;
    const bool bloom = kp.prefilter && kp.bloom[1] != nullptr && !P->depth[1].left_outer;

    /* the folded depth-1 filter in LDS: a first probe that keeps most
     * rejected rows off the full filter in L2 */
    const uint32_t lwords = (bloom && kp.bloom_lds) ? kp.bloom_lds_words : 0;
    for (uint32_t i = tid; i < lwords; i += blockDim.x)
        lds_dyn[i] = kp.bloom_lds[i];
    if (tid == 0)
        s_cnt = 0;
    __syncthreads();
    uint32_t cnt_in = 0;
    bool full = false;
    StagedCols<RPT> S;
    S.setup(P, kp.kds_src);
    const uint64_t wrows = 64ULL * RPT;
    const uint64_t nwt = ((uint64_t)(kp.row_end - kp.row_begin) + wrows - 1) / wrows;
    for (uint64_t wt = (uint64_t)b * nwaves + wid; wt < nwt; wt += (uint64_t)gridDim.x * nwaves)
    {
        const uint64_t base = (uint64_t)kp.row_begin + wt * wrows;
        S.load(base, kp.row_end);
        bool ok[RPT];
        uint64_t bw[RPT], bm[RPT];
        auto screen_step = [&](const int i) __attribute__((always_inline)) {
            const uint64_t row = base + (uint64_t)i * 64 + lane;
            const bool valid = row < kp.row_end;
            int64_t Rfile[NR];
            Lane L;
            L.R = Rfile;
            lane_reset<NR>(L);
            S.fill(L, i, row, valid);
            uint32_t h = 0;
            ok[i] = valid && screen_row<NR, PROG, true>(P, kp, L, (uint32_t)row, bloom, &h);
            if (ok[i])
                cnt_in++;
            bw[i] = bm[i] = 0;
            if (ok[i] && bloom)
            {
                const uint64_t x = bloom_mix(h);
                bm[i] = bloom_bits(x);
                const bool lds_pass = lwords == 0 ||
                    (lds_dyn[x & (lwords - 1)] & bm[i]) == bm[i];
                if (lds_pass)
                    bw[i] = (lwords - 1 == kp.bloom_mask[1]) ? bm[i]     /* the whole filter */
                                                            : kp.bloom[1][x & kp.bloom_mask[1]];
            }
        };
#pragma unroll
        for (int i = 0; i < RPT; i++)
            screen_step(i);
        /* one reservation for the tile's rows that pass */
        bool cand[RPT];
        uint64_t m[RPT];
        uint32_t total = 0;
#pragma unroll
        for (int i = 0; i < RPT; i++)
        {
            cand[i] = ok[i] && (bw[i] & bm[i]) == bm[i];
            m[i] = __ballot(cand[i]);
            total += (uint32_t)__popcll(m[i]);
        }
        if (total == 0)
            continue;
        uint32_t pos = 0;
        if (lane == 0)
            pos = atomicAdd(&s_cnt, total);
        pos = __builtin_amdgcn_readfirstlane(pos);
#pragma unroll
        for (int i = 0; i < RPT; i++)
        {
            const uint32_t r = pos + (uint32_t)__popcll(m[i] & ((1ULL << lane) - 1));
            if (cand[i] && r >= kp.rec_out_cap)
                full = true;
            else if (cand[i])
                *rec_word(kp.rec_out, 1, kp.rec_out_cap, b, 0, r) = (uint32_t)(base + (uint64_t)i * 64 + lane);
            pos += (uint32_t)__popcll(m[i]);
        }
    }
    if (full)
        record_error(ctl, XMSG_JOIN_RECORDS, ERRCODE_DEVICE_ERROR);
    if (cnt_in)
        atomicAdd(&ctl->nitems_in, cnt_in);
    __syncthreads();
    if (tid == 0)
        kp.rec_out_cnt[b] = s_cnt < kp.rec_out_cap ? s_cnt : kp.rec_out_cap;
}

/* The join pass of depth D (0: kp.pass_depth) over the records of block
 * b's region: restore each record's registers, walk its bucket chain, run
 * the join quals of every candidate, then append a level-(D+1) record per
 * match (after the next depth's filter) or, at the last depth, run the
 * final stage into the group-by. */
template <int NR, class PROG, int D>
__device__ __forceinline__ void
join_probe_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    __shared__ uint32_t s_cnt;
    __shared__ uint32_t s_depth_out;
    const kern_data_store *__restrict__ kds = kp.kds_src;
    xpu_task_ctl *ctl = kp.ctl;
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);     /* uniform */
    const int nwaves = blockDim.x >> 6;
    const uint32_t b = blockIdx.x;
    const int d = D ? D : (int)kp.pass_depth;
    const int nrels = P->nrels;
    const bool last = (d == nrels);
    const uint32_t nrec = kp.rec_in_cnt[b] < kp.rec_in_cap ? kp.rec_in_cnt[b] : kp.rec_in_cap;
    const uint32_t nw_in = (uint32_t)d;     /* row, items of depths 1..d-1 */
    const uint32_t nw_out = nw_in + 1;

    /* the group-by table pays its init and merge only for a block with more
     * records than entries */
    LdsTable T = lds_table_at(P, lds_dyn, (last && nrec > kp.lds_nents) ? kp.lds_nents : 0);
    if (T.n > 0)
        lds_init(P, T);
    if (tid == 0)
    {
        s_cnt = 0;
        s_depth_out = 0;
    }
    __syncthreads();

    kern_data_store *ik = (kern_data_store *)(uintptr_t)kp.inner_kds[d];
    const bool left_outer = P->depth[d].left_outer;
    const uint64_t *bf = (d == 1 && kp.screened) ? nullptr : kp.bloom[d];
    const bool nbloom = !last && kp.prefilter && kp.bloom[d + 1] != nullptr &&
                        !P->depth[d + 1].left_outer;
    uint32_t cnt_out = 0, lds_miss = 0, depth_out = 0;
    bool full = false;

    for (uint32_t r0 = (uint32_t)wid * 64; r0 < nrec; r0 += (uint32_t)nwaves * 64)
    {
        const uint32_t r = r0 + lane;
        const bool valid = r < nrec;
        int64_t Rfile[NR];
        Lane L;
        L.R = Rfile;
        lane_reset<NR>(L);
        uint32_t row = 0;
        bool alive = valid;
        /* an error ends the record and is reported for the task; the
         * depth-1 key hash of a level-1 record (the row alone) may send the
         * row to the CPU fallback instead */
        auto run_seg = [&](const xseg seg, const bool act, const bool fb = false) -> bool {
            if (!act)
                return false;
            PROG::template run<false>(P, seg.begin, seg.end, L, kds, row, kp);
            if (L.err)
            {
                if (!(fb && row_fallback(P, L, kp, true)))
                    record_error(ctl, L.err - 1, ERRCODE_DEVICE_ERROR);
                L.err = 0;
                return true;
            }
            return false;
        };
        /* the record's registers: the scan program for its row, then the
         * join program of each depth below d with its recorded item */
        if (valid)
            row = *rec_word(kp.rec_in, nw_in, kp.rec_in_cap, b, 0, r);
        alive = alive && !run_seg(P->seg_scan, alive);
        for (int j = 1; j < XP_MAX_DEPTH && j < d; j++)
        {
            const uint32_t itj = valid ? *rec_word(kp.rec_in, nw_in, kp.rec_in_cap, b, j, r) : 0;
            if (alive)
            {
                kern_data_store *jk = (kern_data_store *)(uintptr_t)kp.inner_kds[j];
                L.R[P->r_tuple] = itj ? (int64_t)(uintptr_t)&hashitem_at(jk, itj)->t : 0;
                SETNULL(L, P->r_tuple, itj == 0);
            }
            const bool e1 = run_seg(P->depth[j].seg_join, alive && itj != 0);
            const bool e2 = run_seg(P->depth[j].seg_oj, alive && itj == 0);
            alive = alive && !e1 && !e2;
        }
        /* bucket walk to the first candidate */
        uint32_t h = 0;
        alive = alive && !run_seg(P->depth[d].seg_hash, alive, d == 1);
        if (alive)
            h = (uint32_t)L.R[P->r_hash];
        uint64_t off = 0;
        if (alive && (!bf || bloom_maybe(bf, kp.bloom_mask[d], h)))
            off = ((const uint64_t *)((const char *)ik + KDS_HEAD_LENGTH(ik)))[h % ik->hash_nslots];
        uint32_t steps = ik->nitems;    /* chains are acyclic: bound the walk */
        while (off != 0 && off != ~0ULL && off < ik->length &&
               hashitem_at(ik, off)->t.hash != h && steps-- > 0)
            off = hashitem_at(ik, off)->next;
        uint32_t cur = (off == 0 || off == ~0ULL || off >= ik->length) ? IT_END : (uint32_t)off;
        bool matched = false;

        for (;;)
        {
            const bool join = alive && cur > IT_END;
            const bool fill = alive && !join && left_outer && !matched;
            if (!wave_any(join || fill))
                break;
            /* the join quals of the candidate, or the outer-join NULL row */
            bool emit = false;
            uint32_t emit_it = 0;
            if (join)
            {
                L.R[P->r_tuple] = (int64_t)(uintptr_t)&hashitem_at(ik, cur)->t;
                SETNULL(L, P->r_tuple, 0);
            }
            else if (fill)
            {
                L.R[P->r_tuple] = 0;
                SETNULL(L, P->r_tuple, 1);
            }
            const bool ej = run_seg(P->depth[d].seg_join, join);
            const bool eo = run_seg(P->depth[d].seg_oj, fill);
            if (join)
            {
                const kern_hashitem *hi = hashitem_at(ik, cur);
                const int64_t status = ej ? 0 : L.R[P->r_status];
                if (status != 0)
                    matched = true;
                emit = status > 0;
                emit_it = cur;
                /* the next candidate with the same hash */
                uint64_t nx = hi->next;
                uint32_t ns = ik->nitems;
                while (nx != 0 && nx != ~0ULL && nx < ik->length &&
                       hashitem_at(ik, nx)->t.hash != h && ns-- > 0)
                    nx = hashitem_at(ik, nx)->next;
                cur = (nx == 0 || nx == ~0ULL || nx >= ik->length) ? IT_END : (uint32_t)nx;
            }
            if (fill)
            {
                matched = true;
                emit = !eo && L.R[P->r_status] > 0;
                emit_it = 0;
            }
            if (ej || eo)
            {
                alive = false;
                emit = false;
            }
            if (emit)
                depth_out++;
            if (last)
            {
                const bool fin = emit && !run_seg(P->seg_final, emit);
                if (fin)
                {
                    cnt_out++;
                    bool direct = true;
                    if (T.n > 0 && gb_row_fast(P, L))
                    {
                        int slot = lds_probe(T, gb_pack_key(P, L));
                        if (slot >= 0)
                        {
                            lds_update(P, L, T, slot);
                            direct = false;
                        }
                    }
                    if (direct)
                        lds_miss++;
                    uint32_t hash = direct ? gb_hash_regs(P, L) : 0;
                    gb_global(P, L, kp.kds_out, hash, direct, &T, -1, ctl, kp.kds_ovf);
                }
            }
            else
            {
                /* the next depth's filter, then the append */
                bool out = emit;
                if (out && nbloom)
                {
                    if (run_seg(P->depth[d + 1].seg_hash, true))
                    {
                        out = false;
                        alive = false;
                    }
                    else
                        out = bloom_maybe(kp.bloom[d + 1], kp.bloom_mask[d + 1], (uint32_t)L.R[P->r_hash]);
                }
                const uint64_t m = __ballot(out);
                if (m)
                {
                    uint32_t pos0 = 0;
                    if (lane == 0)
                        pos0 = atomicAdd(&s_cnt, (uint32_t)__popcll(m));
                    pos0 = __shfl(pos0, 0);
                    const uint32_t q = pos0 + (uint32_t)__popcll(m & ((1ULL << lane) - 1));
                    if (out && q >= kp.rec_out_cap)
                        full = true;
                    else if (out)
                    {
                        for (uint32_t k = 0; k < nw_in; k++)
                            *rec_word(kp.rec_out, nw_out, kp.rec_out_cap, b, k, q) =
                                *rec_word(kp.rec_in, nw_in, kp.rec_in_cap, b, k, r);
                        *rec_word(kp.rec_out, nw_out, kp.rec_out_cap, b, nw_in, q) = emit_it;
                    }
                }
            }
        }
    }
    if (full)
        record_error(ctl, XMSG_JOIN_RECORDS, ERRCODE_DEVICE_ERROR);
    if (depth_out)
        atomicAdd(&s_depth_out, depth_out);
    if (T.n > 0)
    {
        __syncthreads();
        lds_flush<NR>(P, T, kp.kds_out, ctl, kp.kds_ovf);
    }
    if (cnt_out)
        atomicAdd(&ctl->nitems_out, cnt_out);
    if (lds_miss)
        atomicAdd((unsigned long long *)&ctl->lds_miss, (unsigned long long)lds_miss);
    __syncthreads();
    if (tid == 0)
    {
        if (!last)
            kp.rec_out_cnt[b] = s_cnt < kp.rec_out_cap ? s_cnt : kp.rec_out_cap;
        if (s_depth_out)
            atomicAdd(&ctl->depth_out[d], s_depth_out);
    }
}

/* ------------------------------------------------------------------
 * scan + projection for plans without joins
 *
 * Two passes over the same wave tiles of 64*RPT rows, each block owning a
 * contiguous run of tiles (so tile order is row order): the count pass
 * stores each tile's output rows and bytes and each block's totals; the
 * tile scan turns them into every tile's exclusive prefix and finds the
 * first tile that does not fit kds_dst; the write pass runs the rows again
 * and writes each output tuple at its tile's prefix plus the outputs of
 * the rows before it in the tile, so kds_dst holds the rows in scan order.
 * The staged loads of a tile are issued together, as in the scan +
 * group-by kernel, and no pass waits on another block.
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t
wave_incl_scan_u32(uint32_t v)
{
    int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off)
            v += t;
    }
    return v;
}

template <int NR, int RPT, class PROG, int PM>
__device__ __forceinline__ void
scan_proj_body(cplan *P, const xpu_kparams &kp)
{
    __shared__ uint64_t s_tot[2 * (XP_BLOCK_MAX / 64)];
    const kern_data_store *__restrict__ kds = kp.kds_src;
    xpu_task_ctl *ctl = kp.ctl;
    kern_data_store *dst = kp.kds_out;
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);     /* uniform */
    const int nwaves = blockDim.x >> 6;
    const uint64_t wrows = 64ULL * RPT;
    const uint64_t nwt = kp.ntiles;
    const uint64_t per = (nwt + gridDim.x - 1) / gridDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * per;
    const uint64_t t1 = t0 + per < nwt ? t0 + per : nwt;
    uint64_t *st = kp.tile_status;
    uint64_t stop = nwt;
    if (PM == PM_WRITE)
    {
        uint32_t rt = __hip_atomic_load(&ctl->resume_tile, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (rt)
            stop = 0xffffffffU - rt;
        /* the rows and bytes kds_dst ends up with */
        if (blockIdx.x == 0 && tid == 0 && stop < nwt)
        {
            ctl->out_nitems = st[2 * stop];
            ctl->out_usage = st[2 * stop + 1];
        }
    }
    StagedCols<RPT> S;
    S.setup(P, kds);
    const int pc_begin = P->seg_scan.begin + P->stage_prefix;
    const int pc_end = P->seg_final.end;
    uint64_t *rowidx = (uint64_t *)((char *)dst + KDS_HEAD_LENGTH(dst));
    uint32_t cnt = 0;
    uint64_t wc = 0, wb = 0;
    for (uint64_t wt = t0 + wid; wt < t1; wt += nwaves)
    {
        if (PM == PM_WRITE && wt >= stop)
            break;
        const uint64_t base = (uint64_t)kp.row_begin + wt * wrows;
        S.load(base, kp.row_end);
        /* count pass: the tile's totals so far; write pass: the outputs
         * before the current row batch */
        uint64_t tr = 0, tb = 0;
        if (PM == PM_WRITE)
        {
            tr = st[2 * wt];
            tb = st[2 * wt + 1];
        }
        auto row_step = [&](const int i) __attribute__((always_inline))
        {
            const uint64_t row = base + (uint64_t)i * 64 + lane;
            const bool valid = row < kp.row_end;
            int64_t Rfile[NR];
            Lane L;
            L.R = Rfile;
            lane_reset<NR>(L);
            S.fill(L, i, row, valid);
            if (valid)
                PROG::template run<true>(P, pc_begin, pc_end, L, kds, (uint32_t)row, kp);
            bool ok = valid && !L.dead;
            if (valid && L.err)
            {
                /* the write pass writes a CPU-fallback row once (the
                 * count pass drops it) */
                if (!row_fallback(P, L, kp, PM == PM_WRITE) && PM == PM_COUNT)
                    record_error(ctl, L.err - 1, ERRCODE_DEVICE_ERROR);
                ok = false;
            }
            const uint32_t sz = ok ? (uint32_t)xalign(8, proj_tuple_size(P, L)) : 0;
            const uint64_t m = __ballot(ok);
            const uint32_t ib = wave_incl_scan_u32(sz);
            if (PM == PM_COUNT)
            {
                if (ok)
                    cnt++;
            }
            else if (ok)
            {
                const uint64_t rowid = tr + (uint64_t)__popcll(m & ((1ULL << lane) - 1));
                const uint64_t usage = tb + ib;
                char *tp = (char *)dst + dst->length - usage;
                proj_tuple_write(P, L, tp, sz, (uint32_t)rowid);
                rowidx[rowid] = usage;
            }
            tr += (uint64_t)__popcll(m);
            tb += (uint64_t)__shfl(ib, 63);
        };
        if constexpr (PROG::kUnrollRows)
        {
#pragma unroll
            for (int i = 0; i < RPT; i++)
                row_step(i);
        }
        else
        {
#pragma unroll 1
            for (int i = 0; i < RPT; i++)
                row_step(i);
        }
        if (PM == PM_COUNT)
        {
            if (lane == 0)
            {
                st[2 * wt] = tr;
                st[2 * wt + 1] = tb;
            }
            wc += tr;
            wb += tb;
        }
    }
    if (PM == PM_COUNT)
    {
        if (lane == 0)
        {
            s_tot[2 * wid] = wc;
            s_tot[2 * wid + 1] = wb;
        }
        __syncthreads();
        if (tid == 0)
        {
            uint64_t r = 0, bsum = 0;
            for (int w = 0; w < nwaves; w++)
            {
                r += s_tot[2 * w];
                bsum += s_tot[2 * w + 1];
            }
            kp.block_status[2 * blockIdx.x] = r;
            kp.block_status[2 * blockIdx.x + 1] = bsum;
        }
        if (cnt)
        {
            atomicAdd(&ctl->nitems_in, cnt);
            atomicAdd(&ctl->nitems_out, cnt);
        }
    }
}

/* The tile scan of a scan + projection: block b sums the totals of the
 * count-pass blocks before it, then scans the tiles block b of the count
 * pass owned into exclusive prefixes, and reports the first tile that does
 * not fit kds_dst.  The last block stores the grand totals. */
__device__ __forceinline__ void
tile_scan_blocks_body(uint64_t *st, const uint64_t *bt, uint32_t nwt, const kern_data_store *dst,
                      xpu_task_ctl *ctl)
{
    __shared__ uint64_t s_wr[XP_BLOCK_MAX / 64], s_wb[XP_BLOCK_MAX / 64];
    __shared__ uint64_t s_carry_r, s_carry_b;
    __shared__ uint32_t s_stop;
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wid = tid >> 6;
    const int nwaves = blockDim.x >> 6;
    const uint32_t b = blockIdx.x;
    const uint64_t per = ((uint64_t)nwt + gridDim.x - 1) / gridDim.x;
    const uint64_t t0 = (uint64_t)b * per;
    const uint64_t t1 = t0 + per < nwt ? t0 + per : nwt;

    /* the outputs of the blocks before b */
    uint64_t pr = 0, pb = 0;
    for (uint32_t j = tid; j < b; j += blockDim.x)
    {
        pr += bt[2 * j];
        pb += bt[2 * j + 1];
    }
    pr = wave_incl_scan_u64(pr);
    pb = wave_incl_scan_u64(pb);
    if (lane == 63)
    {
        s_wr[wid] = pr;
        s_wb[wid] = pb;
    }
    __syncthreads();
    if (tid == 0)
    {
        uint64_t r = 0, bb = 0;
        for (int w = 0; w < nwaves; w++)
        {
            r += s_wr[w];
            bb += s_wb[w];
        }
        s_carry_r = r;
        s_carry_b = bb;
        s_stop = nwt;
    }
    __syncthreads();
    const uint64_t head = KDS_HEAD_LENGTH(dst);
    const uint64_t len = dst->length;
    for (uint64_t base = t0; base < t1; base += blockDim.x)
    {
        const uint64_t t = base + tid;
        const uint64_t r = t < t1 ? st[2 * t] : 0;
        const uint64_t by = t < t1 ? st[2 * t + 1] : 0;
        const uint64_t ir = wave_incl_scan_u64(r);
        const uint64_t ib = wave_incl_scan_u64(by);
        if (lane == 63)
        {
            s_wr[wid] = ir;
            s_wb[wid] = ib;
        }
        __syncthreads();
        uint64_t wr = s_carry_r, wb = s_carry_b;
        for (int w = 0; w < wid; w++)
        {
            wr += s_wr[w];
            wb += s_wb[w];
        }
        const uint64_t er = wr + ir - r;
        const uint64_t eb = wb + ib - by;
        if (t < t1)
        {
            st[2 * t] = er;
            st[2 * t + 1] = eb;
            if (head + 8 * (er + r) + eb + by > len)
                atomicMin(&s_stop, (uint32_t)t);
        }
        __syncthreads();
        if (tid == (int)blockDim.x - 1)
        {
            s_carry_r = er + r;
            s_carry_b = eb + by;
        }
        __syncthreads();
    }
    if (tid == 0)
    {
        const uint32_t stop = s_stop;
        if (stop < nwt)
        {
            atomicMax(&ctl->resume_tile, 0xffffffffU - stop);
            if (stop == 0)
                record_error(ctl, XMSG_TILE_TOO_LARGE, ERRCODE_DEVICE_ERROR);
        }
        if (b == gridDim.x - 1)
        {
            ctl->out_nitems = s_carry_r;
            ctl->out_usage = s_carry_b;
        }
    }
}

/* ------------------------------------------------------------------
 * scan + group-by kernel for plans without joins
 *
 * Each wave owns tiles of 64*RPT consecutive rows (static assignment, so
 * no block barrier or ticket per tile).  The tile prologue issues the loads
 * of every staged column for all RPT rows of the lane before any loaded
 * value is used: a lane keeps nstage*RPT loads in flight instead of one
 * dependent load per column per row.  The rows are then run through the
 * fused scan-quals + aggregate-argument program one at a time and folded
 * into the block's LDS partial table.
 * ------------------------------------------------------------------ */
/* marks the unused records [pos, end) of a wave's spill reservation */
__device__ __forceinline__ void
spill_fill_holes(const xpu_kparams &kp, uint64_t pos, uint64_t end, int rec_words)
{
    for (uint64_t i = pos + __lane_id(); i < end; i += 64)
        kp.spill[i * rec_words] = XLDS_EMPTY_KEY;
}

/* counters of one wave over the ranges it processed */
struct ScanAggWave
{
    uint32_t    cnt_ok;         /* rows that passed the scan quals */
    uint32_t    lds_miss;       /* rows merged into kds_final directly */
    uint32_t    nspilled;       /* rows written to the spill buffer */
};

/* the wave tiles wt_first, wt_first + wt_stride, ... of the rows
 * [kp.row_begin, kp.row_end) of kp.kds_src, folded into T (whose entries
 * belong to kp.kds_out) */
template <int NR, int RPT, class PROG>
__device__ __forceinline__ void
scan_agg_range(cplan *P, const xpu_kparams &kp, LdsTable &T, uint64_t wt_first,
               uint64_t wt_stride, ScanAggWave &W)
{
    const kern_data_store *__restrict__ kds = kp.kds_src;
    xpu_task_ctl *ctl = kp.ctl;
    const int lane = __lane_id();
    StagedCols<RPT> S;
    S.setup(P, kds);
    const int pc_begin = P->seg_scan.begin + P->stage_prefix;
    const int pc_end = P->seg_final.end;

    /* rows that find no room in the LDS table are appended to the spill
     * buffer (a second kernel aggregates them by hash partition); each wave
     * reserves XSPILL_CHUNK records at a time, so appending costs one device
     * atomic per chunk */
    const bool can_spill = P->spill_ok && kp.spill != nullptr && kp.lds_nents > 0;
    const int rec_words = 1 + P->spill_nregs;
    uint64_t sp_pos = 0, sp_end = 0;
    const uint64_t wrows = 64ULL * RPT;
    const uint64_t nwt = ((uint64_t)(kp.row_end - kp.row_begin) + wrows - 1) / wrows;
    uint32_t iter = 0;
    for (uint64_t wt = wt_first; wt < nwt; wt += wt_stride, iter++)
    {
        /* another wave's error ends the task (checked now and then) */
        if ((iter & 31) == 31 &&
            __hip_atomic_load(&ctl->errcode, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            break;
        const uint64_t base = (uint64_t)kp.row_begin + wt * wrows;
        S.load(base, kp.row_end);

        /* one row of the lane's tile (a lambda: the specialized kernel
         * unrolls the rows, so the staged words stay in registers; the
         * interpreter keeps one copy of the program loop) */
        auto row_step = [&](const int i) __attribute__((always_inline))
        {
            const uint64_t row = base + (uint64_t)i * 64 + lane;
            const bool valid = row < kp.row_end;
            int64_t Rfile[NR];
            Lane L;
            L.R = Rfile;
            L.nmask = ~1U;
            L.err = 0;
            L.dead = 0;
#pragma unroll
            for (int r = 0; r < NR; r++)
                L.R[r] = 0;
            L.R[XP_REG_TRUE] = 1;
            S.fill(L, i, row, valid);
            if (valid)
                PROG::template run<true>(P, pc_begin, pc_end, L, kds, (uint32_t)row, kp);
            bool ok = valid && !L.dead;
            if (valid && L.err)
            {
                if (!row_fallback(P, L, kp, true))
                    record_error(ctl, L.err - 1, ERRCODE_DEVICE_ERROR);
                ok = false;
            }
            if (ok)
                W.cnt_ok++;
            bool direct = ok;
            bool spill = false;
            uint64_t pk = 0;
            if (ok && T.n > 0 && gb_row_fast(P, L))
            {
                pk = gb_pack_key(P, L);
                int slot = lds_probe(T, pk);
                if (slot >= 0)
                {
                    lds_update(P, L, T, slot);
                    direct = false;
                }
                else if (can_spill)
                {
                    spill = true;
                    direct = false;
                }
            }
            if (can_spill)
            {
                const uint64_t m = __ballot(spill);
                if (m)
                {
                    const uint64_t k = __popcll(m);
                    if (sp_pos + k > sp_end)
                    {
                        spill_fill_holes(kp, sp_pos, sp_end, rec_words);
                        uint64_t b0 = 0;
                        if (lane == 0)
                            b0 = atomicAdd((unsigned long long *)&ctl->spill_next,
                                           (unsigned long long)XSPILL_CHUNK);
                        b0 = __shfl(b0, 0);
                        sp_pos = b0 < kp.spill_cap ? b0 : kp.spill_cap;
                        sp_end = b0 + XSPILL_CHUNK < kp.spill_cap ? b0 + XSPILL_CHUNK : kp.spill_cap;
                    }
                    const uint64_t pos = sp_pos + __popcll(m & ((1ULL << lane) - 1));
                    if (spill)
                    {
                        if (pos < sp_end)
                        {
                            uint64_t *r = kp.spill + pos * rec_words;
                            r[0] = pk;
                            for (int j = 0; j < (int)P->spill_nregs; j++)
                                r[1 + j] = (uint64_t)L.R[P->spill_regs[j]];
                            W.nspilled++;
                        }
                        else
                            direct = true;      /* spill buffer full */
                    }
                    sp_pos = sp_pos + k < sp_end ? sp_pos + k : sp_end;
                }
            }
            if (direct)
                W.lds_miss++;
            if (wave_any(direct))
            {
                uint32_t hash = direct ? gb_hash_regs(P, L) : 0;
                gb_global(P, L, kp.kds_out, hash, direct, &T, -1, ctl, kp.kds_ovf);
            }
        };
        if constexpr (PROG::kUnrollRows)
        {
#pragma unroll
            for (int i = 0; i < RPT; i++)
                row_step(i);
        }
        else
        {
#pragma unroll 1
            for (int i = 0; i < RPT; i++)
                row_step(i);
        }
    }
    if (can_spill)
        spill_fill_holes(kp, sp_pos, sp_end, rec_words);
}

/* adds a wave's counters to a task's control block */
__device__ __forceinline__ void
scan_agg_report(xpu_task_ctl *ctl, ScanAggWave &W)
{
    if (W.cnt_ok)
    {
        atomicAdd(&ctl->nitems_in, W.cnt_ok);
        atomicAdd(&ctl->nitems_out, W.cnt_ok);
    }
    if (W.lds_miss)
        atomicAdd((unsigned long long *)&ctl->lds_miss, (unsigned long long)W.lds_miss);
    if (W.nspilled)
        atomicAdd((unsigned long long *)&ctl->spill_rows, (unsigned long long)W.nspilled);
    W.cnt_ok = W.lds_miss = W.nspilled = 0;
}

/* one task: every block takes the wave tiles blockIdx.x * nwaves + wid,
 * strided by the grid's wave count */
template <int NR, int RPT, class PROG>
__device__ __forceinline__ void
scan_agg_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    /* wave-uniform (the backend cannot prove threadIdx.x >> 6 uniform, and
     * everything derived from it, tile base included, would go divergent) */
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    LdsTable T = lds_table_at(P, lds_dyn, kp.lds_nents);
    if (T.n > 0)
        lds_init(P, T);
    __syncthreads();
    ScanAggWave W = {0, 0, 0};
    scan_agg_range<NR, RPT, PROG>(P, kp, T, (uint64_t)blockIdx.x * nwaves + wid,
                                  (uint64_t)gridDim.x * nwaves, W);
    __syncthreads();
    if (T.n > 0)
    {
        if (kp.dump)
            lds_dump<NR>(P, T, kp);
        else
            lds_flush<NR>(P, T, kp.kds_out, kp.ctl, kp.kds_ovf);
    }
    scan_agg_report(kp.ctl, W);
}


/* ------------------------------------------------------------------
 * second pass of a scan + group-by launch: the spill records
 *
 * Every block owns one hash partition of the packed keys (the number of
 * partitions is chosen so that a partition's groups fit one LDS table) and
 * a slice of the records; it aggregates the records of its partition in
 * LDS and merges the table into kds_final like the first pass.
 * ------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t
spill_partition(uint64_t pk, uint32_t nparts)
{
    uint64_t h = pk * 0xD6E8FEB86659FD93ULL;
    return (uint32_t)(((h >> 32) * (uint64_t)nparts) >> 32);
}

/* the spill records [lo, hi) of kp.spill whose key falls in partition part
 * of nparts, folded into T (merged into kp.kds_out by the caller) */
template <int NR>
__device__ __forceinline__ uint32_t
spill_agg_range(cplan *P, const xpu_kparams &kp, LdsTable &T, uint64_t lo, uint64_t hi,
                uint32_t part, uint32_t nparts)
{
    xpu_task_ctl *ctl = kp.ctl;
    const int rec_words = 1 + P->spill_nregs;
    int64_t Rfile[NR];
    Lane L;
    L.R = Rfile;
    L.nmask = ~1U;
    L.err = 0;
    L.dead = 0;
#pragma unroll
    for (int r = 0; r < NR; r++)
        L.R[r] = 0;
    L.R[XP_REG_TRUE] = 1;
    /* FILTERs passed for every spilled row */
    for (int j = 0; j < (int)P->gb_nattrs; j++)
    {
        const int fr = P->aggs[j].filter_reg;
        if (fr != 0xff)
        {
            L.R[fr] = 1;
            SETNULL(L, fr, 0);
        }
    }
    uint32_t lds_miss = 0;
    for (uint64_t base = lo; base < hi; base += blockDim.x)
    {
        const uint64_t i = base + threadIdx.x;
        bool mine = false;
        uint64_t pk = XLDS_EMPTY_KEY;
        if (i < hi)
        {
            pk = kp.spill[i * rec_words];
            mine = (pk != XLDS_EMPTY_KEY) && spill_partition(pk, nparts) == part;
        }
        bool direct = false;
        if (mine)
        {
            gb_unpack_key(P, L, pk);
            for (int j = 0; j < (int)P->spill_nregs; j++)
            {
                const int r = P->spill_regs[j];
                L.R[r] = (int64_t)kp.spill[i * rec_words + 1 + j];
                SETNULL(L, r, 0);
            }
            int slot = lds_probe(T, pk);
            if (slot >= 0)
                lds_update(P, L, T, slot);
            else
                direct = true;
        }
        if (direct)
            lds_miss++;
        if (wave_any(direct))
        {
            uint32_t hash = direct ? gb_hash_regs(P, L) : 0;
            gb_global(P, L, kp.kds_out, hash, direct, &T, -1, ctl, kp.kds_ovf);
        }
    }
    return lds_miss;
}

/* partitions of a task's spill: about half an LDS table of groups each */
__device__ __forceinline__ uint32_t
spill_nparts(uint32_t ngroups, uint32_t table_n, uint32_t maxparts)
{
    uint32_t nparts = (uint32_t)((2ULL * ngroups + table_n - 1) / table_n);
    return nparts < 1 ? 1 : (nparts > maxparts ? maxparts : nparts);
}

/* the spill records [lo, hi) of kp.spill, every key, folded into T; a full
 * table is dumped (partial-state records) and restarted */
template <int NR>
__device__ __forceinline__ void
spill_fold_range(cplan *P, const xpu_kparams &kp, LdsTable &T, uint64_t lo, uint64_t hi)
{
    const int rec_words = 1 + P->spill_nregs;
    int64_t Rfile[NR];
    Lane L;
    L.R = Rfile;
    L.nmask = ~1U;
    L.err = 0;
    L.dead = 0;
#pragma unroll
    for (int r = 0; r < NR; r++)
        L.R[r] = 0;
    L.R[XP_REG_TRUE] = 1;
    for (uint64_t base = lo; base < hi; base += blockDim.x)
    {
        const uint64_t i = base + threadIdx.x;
        uint64_t pk = XLDS_EMPTY_KEY;
        if (i < hi)
            pk = kp.spill[i * rec_words];
        bool fail = false;
        if (pk != XLDS_EMPTY_KEY)
        {
            gb_unpack_key(P, L, pk);
            for (int j = 0; j < (int)P->spill_nregs; j++)
            {
                const int r = P->spill_regs[j];
                L.R[r] = (int64_t)kp.spill[i * rec_words + 1 + j];
                SETNULL(L, r, 0);
            }
            int slot = lds_probe(T, pk);
            if (slot >= 0)
                lds_update(P, L, T, slot);
            else
                fail = true;
        }
        while (__syncthreads_or(fail))
        {
            lds_dump<NR>(P, T, kp);
            __syncthreads();
            lds_init(P, T);
            __syncthreads();
            if (fail)
            {
                int slot = lds_probe(T, pk);
                if (slot >= 0)
                {
                    lds_update(P, L, T, slot);
                    fail = false;
                }
            }
        }
    }
}

template <int NR, class PROG>
__device__ __forceinline__ void
spill_agg_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    xpu_task_ctl *ctl = kp.ctl;
    LdsTable T = lds_table_at(P, lds_dyn, kp.lds_nents);
    if (T.n > 0)
        lds_init(P, T);
    __syncthreads();
    const uint64_t reserved = __hip_atomic_load(&ctl->spill_next, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t nrec = reserved < kp.spill_cap ? reserved : kp.spill_cap;
    if (nrec == 0 || T.n == 0 || kp.spill == nullptr)
        return;
    if (kp.dump)
    {
        /* each block folds one slice of the records (read once) */
        const uint64_t lo = nrec * blockIdx.x / gridDim.x;
        const uint64_t hi = nrec * (blockIdx.x + 1) / gridDim.x;
        spill_fold_range<NR>(P, kp, T, lo, hi);
        __syncthreads();
        lds_dump<NR>(P, T, kp);
        return;
    }
    const uint32_t nparts = spill_nparts(kp.kds_out->nitems, T.n,
                                         gridDim.x < 64 ? gridDim.x : 64);
    const uint32_t part = blockIdx.x % nparts;
    const uint32_t kblk = blockIdx.x / nparts;
    const uint32_t nblk = (gridDim.x - part + nparts - 1) / nparts;
    const uint64_t lo = nrec * kblk / nblk;
    const uint64_t hi = nrec * (kblk + 1) / nblk;
    uint32_t lds_miss = spill_agg_range<NR>(P, kp, T, lo, hi, part, nparts);
    __syncthreads();
    lds_flush<NR>(P, T, kp.kds_out, ctl, kp.kds_ovf);
    if (lds_miss)
        atomicAdd((unsigned long long *)&ctl->lds_miss, (unsigned long long)lds_miss);
}

/* ------------------------------------------------------------------
 * batched fragments: many scan + group-by tasks of one plan shape (their
 * own chunk, kds_final, parameters and control block each) in one
 * persistent launch
 *
 * Pass 1: blocks take work items (a row range of one fragment; items are
 * ordered by fragment) from a ticket; a block's LDS table holds one
 * fragment at a time and is merged into that fragment's kds_final when the
 * block moves on.  Rows without room spill into the fragment's own region
 * of the spill buffer.  A one-block setup kernel then lists the spill work
 * (fragment, hash partition, slice of records) and pass 2 takes those items
 * from a second ticket.
 * ------------------------------------------------------------------ */

/* the task parameters of fragment f on top of the launch's common ones */
__device__ __forceinline__ xpu_kparams
batch_kparams(const xpu_kparams &kp, const xbatch_frag &f)
{
    xpu_kparams k = kp;
    k.kds_src = f.kds_src;
    k.kds_out = f.kds_final;
    k.ctl = f.ctl;
    k.const_pool = f.const_pool;
    k.params = f.params;
    k.param_nulls = f.param_nulls;
    k.spill = f.spill;
    k.spill_cap = f.spill_cap;
    k.row_begin = 0;
    k.row_end = f.nrows;
    return k;
}

template <int NR, int RPT, class PROG>
__device__ __forceinline__ void
batch_agg_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    __shared__ uint32_t s_item;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    LdsTable T = lds_table_at(P, lds_dyn, kp.lds_nents);
    if (T.n > 0)
        lds_init(P, T);
    ScanAggWave W = {0, 0, 0};
    int cur = -1;
    for (;;)
    {
        if (threadIdx.x == 0)
            s_item = atomicAdd(kp.batch_ticket, 1U);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= kp.batch_nitems)
            break;
        const xbatch_item item = kp.batch_items[it];
        if ((int)item.frag != cur)
        {
            if (cur >= 0)
            {
                const xbatch_frag &fc = kp.batch_frags[cur];
                if (T.n > 0)
                {
                    lds_flush<NR>(P, T, fc.kds_final, fc.ctl);
                    __syncthreads();
                    lds_init(P, T);
                }
                scan_agg_report(fc.ctl, W);
            }
            __syncthreads();
            cur = (int)item.frag;
        }
        xpu_kparams k = batch_kparams(kp, kp.batch_frags[item.frag]);
        k.row_begin = item.row_begin;
        k.row_end = item.row_end;
        scan_agg_range<NR, RPT, PROG>(P, k, T, wid, nwaves, W);
        __syncthreads();
    }
    if (cur >= 0)
    {
        const xbatch_frag &fc = kp.batch_frags[cur];
        if (T.n > 0)
            lds_flush<NR>(P, T, fc.kds_final, fc.ctl);
        scan_agg_report(fc.ctl, W);
    }
}

/* lists the spill work of a batch (one thread) */
__device__ __forceinline__ void
batch_spill_setup_body(const xpu_kparams &kp)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    uint32_t n = 0;
    for (uint32_t f = 0; f < kp.batch_nfrags; f++)
    {
        const xbatch_frag &fr = kp.batch_frags[f];
        if (!fr.spill || kp.lds_nents == 0)
            continue;
        const uint64_t reserved = __hip_atomic_load(&fr.ctl->spill_next, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t nrec = reserved < fr.spill_cap ? reserved : fr.spill_cap;
        if (nrec == 0)
            continue;
        const uint32_t nparts = spill_nparts(fr.kds_final->nitems, kp.lds_nents, 64);
        uint64_t nslices = (nrec + XBATCH_SPILL_SLICE - 1) / XBATCH_SPILL_SLICE;
        if (nslices > 64)
            nslices = 64;
        for (uint32_t p = 0; p < nparts; p++)
            for (uint64_t sl = 0; sl < nslices; sl++)
            {
                if (n >= kp.batch_nsitems_cap)
                    break;
                xbatch_sitem si;
                si.frag = f;
                si.part = (uint16_t)p;
                si.nparts = (uint16_t)nparts;
                si.lo = nrec * sl / nslices;
                si.hi = nrec * (sl + 1) / nslices;
                kp.batch_sitems[n++] = si;
            }
    }
    kp.batch_ticket[1] = 0;
    kp.batch_ticket[2] = n;
}

template <int NR, class PROG>
__device__ __forceinline__ void
batch_spill_body(cplan *P, const xpu_kparams &kp)
{
    extern __shared__ unsigned long long lds_dyn[];
    __shared__ uint32_t s_item;
    LdsTable T = lds_table_at(P, lds_dyn, kp.lds_nents);
    if (T.n == 0)
        return;
    const uint32_t nsitems = __hip_atomic_load(&kp.batch_ticket[2], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    for (;;)
    {
        if (threadIdx.x == 0)
            s_item = atomicAdd(&kp.batch_ticket[1], 1U);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= nsitems)
            break;
        const xbatch_sitem si = kp.batch_sitems[it];
        const xpu_kparams k = batch_kparams(kp, kp.batch_frags[si.frag]);
        lds_init(P, T);
        __syncthreads();
        uint32_t miss = spill_agg_range<NR>(P, k, T, si.lo, si.hi, si.part, si.nparts);
        __syncthreads();
        lds_flush<NR>(P, T, k.kds_out, k.ctl);
        if (miss)
            atomicAdd((unsigned long long *)&k.ctl->lds_miss, (unsigned long long)miss);
        __syncthreads();
    }
}

/* ------------------------------------------------------------------
 * the scan between the count and write passes of a projection: exclusive
 * prefixes of the tile totals (rows, bytes) in place, and the first tile
 * that does not fit kds_dst as the resume point (one block)
 * ------------------------------------------------------------------ */
__device__ __forceinline__ void
tile_scan_body(uint64_t *st, uint32_t ntiles, const kern_data_store *dst, xpu_task_ctl *ctl)
{
    __shared__ uint64_t s_wr[XP_BLOCK_MAX / 64], s_wb[XP_BLOCK_MAX / 64];
    __shared__ uint64_t s_carry_r, s_carry_b;
    __shared__ uint32_t s_stop;
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wid = tid >> 6;
    if (tid == 0)
    {
        s_carry_r = 0;
        s_carry_b = 0;
        s_stop = ntiles;
    }
    __syncthreads();
    const uint64_t head = KDS_HEAD_LENGTH(dst);
    const uint64_t len = dst->length;
    for (uint32_t base = 0; base < ntiles; base += blockDim.x)
    {
        const uint32_t t = base + tid;
        const uint64_t r = t < ntiles ? st[2ULL * t] : 0;
        const uint64_t b = t < ntiles ? st[2ULL * t + 1] : 0;
        const uint64_t ir = wave_incl_scan_u64(r);
        const uint64_t ib = wave_incl_scan_u64(b);
        if (lane == 63)
        {
            s_wr[wid] = ir;
            s_wb[wid] = ib;
        }
        __syncthreads();
        uint64_t wr = s_carry_r, wb = s_carry_b;
        for (int w = 0; w < wid; w++)
        {
            wr += s_wr[w];
            wb += s_wb[w];
        }
        const uint64_t er = wr + ir - r;
        const uint64_t eb = wb + ib - b;
        if (t < ntiles)
        {
            st[2ULL * t] = er;
            st[2ULL * t + 1] = eb;
            if (head + 8 * (er + r) + eb + b > len)
                atomicMin(&s_stop, t);
        }
        __syncthreads();
        if (tid == (int)blockDim.x - 1)
        {
            s_carry_r = er + r;
            s_carry_b = eb + b;
        }
        __syncthreads();
    }
    if (tid == 0)
    {
        const uint32_t stop = s_stop;
        if (stop < ntiles)
        {
            ctl->resume_tile = 0xffffffffU - stop;
            if (stop == 0)
                record_error(ctl, XMSG_TILE_TOO_LARGE, ERRCODE_DEVICE_ERROR);
            ctl->out_nitems = st[2ULL * stop];
            ctl->out_usage = st[2ULL * stop + 1];
        }
        else
        {
            ctl->out_nitems = s_carry_r;
            ctl->out_usage = s_carry_b;
        }
    }
}

#endif /* XPU_DEVICE_H */
