/*
 * xpu_compile.cpp - translation of a kern_session_info blob into the flat
 * register programs executed by the MI355X pipeline kernel.
 *
 * Input contract: the blob the reference backend builds in
 * pgstromBuildSessionInfo (executor.c:517-801) from the bytecode emitted by
 * codegen.c.  Each kern_expression is checked the way __KEXP_IS_VALID
 * (xpu_common.h:2697-2723) checks it (bounds and the per-node magic) before it
 * is translated; anything this build cannot execute is rejected here, at
 * session open, with ERRCODE_DEVICE_ERROR instead of failing inside a kernel.
 */
#include "xpu_compile.h"

#include <stdio.h>
#include <string.h>

#include "xpu_kernels.h"
#include "xpu_opcodes_table.h"

namespace {

enum { K_cmp = 1, K_arith, K_unary, K_bitop, K_cast, K_textcmp, K_textlen, K_like, K_substr,
       K_numcmp, K_numarith, K_numunary, K_numcast };
enum { O_eq = 1, O_ne, O_lt, O_le, O_gt, O_ge, O_pl, O_mi, O_mul, O_div, O_mod,
       O_up, O_um, O_abs, O_not, O_and, O_or, O_xor, O_shl, O_shr, O_cast, O_len,
       O_like, O_nlike, O_ilike, O_nilike, O_substr };

struct devfunc { int opcode; const char *name; int kind, op, nargs, t0, t1, t2, res; };

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
#define DEVFUNC(OPC, NAME, KIND, OP, NARGS, A0, A1, A2, RES) \
    { OPC, #NAME, K_##KIND, O_##OP, NARGS, T_##A0, T_##A1, T_##A2, T_##RES },
const devfunc devfunc_table[] = { XPU_FUNCTION_TABLE(DEVFUNC) };

const devfunc *
lookup_devfunc(int opcode)
{
    for (const devfunc &f : devfunc_table)
        if (f.opcode == opcode)
            return &f;
    return nullptr;
}

int
type_domain(int t)
{
    switch (t)
    {
        case TypeOpCode__bool: case TypeOpCode__int1: case TypeOpCode__int2:
        case TypeOpCode__int4: case TypeOpCode__int8: case TypeOpCode__date:
            return XD_INT;
        case TypeOpCode__float4: return XD_F32;
        case TypeOpCode__float8: return XD_F64;
        case TypeOpCode__text: case TypeOpCode__bytea: case TypeOpCode__bpchar:
            return XD_TEXT;
        case TypeOpCode__numeric:
            return XD_NUM;
    }
    return -1;
}

/* registers a value of a domain occupies */
int
dom_regs(int dom)
{
    return dom == XD_TEXT ? 2 : dom == XD_NUM ? 3 : 1;
}

/* the move op of a domain */
int
mov_op(int dom)
{
    return dom == XD_TEXT ? XO_MOV2 : dom == XD_NUM ? XO_MOV3 : XO_MOV;
}

/* __numeric_typmod_weight (xpu_numeric.h:416-430): the scale of a numeric
 * sum, from the argument's typmod (15 digits when it has none) */
int
numeric_typmod_scale(int32_t typmod)
{
    int w = 15;
    if (typmod >= 0)
    {
        w = (((typmod - 4) & 0x7ff) ^ 1024) - 1024;
        w = w < 0 ? 0 : (w > 15 ? 15 : w);
    }
    return w;
}

int
type_width(int t)
{
    switch (t)
    {
        case TypeOpCode__bool: case TypeOpCode__int1: return 1;
        case TypeOpCode__int2: return 2;
        case TypeOpCode__int4: case TypeOpCode__float4: case TypeOpCode__date: return 4;
        case TypeOpCode__int8: case TypeOpCode__float8: return 8;
    }
    return -1;
}

int
int_class(int t)
{
    switch (t)
    {
        case TypeOpCode__bool: return XT_BOOL;
        case TypeOpCode__int1: return XT_INT1;
        case TypeOpCode__int2: return XT_INT2;
        case TypeOpCode__int4: case TypeOpCode__date: return XT_INT4;
        default: return XT_INT8;
    }
}

struct Reg { int r; int dom; int type; };

/* the session's database encoding (xpu_encode_info, xpu_textlib.h:28-32)
 * by name: single-byte encodings count one character per byte; UTF8 by
 * pg_utf8_mblen; other multi-byte encodings are not executed here */
enum { ENC_NONE = 0, ENC_SINGLE, ENC_UTF8, ENC_OTHER };

int
encoding_kind(const kern_session_info *s, size_t length)
{
    static const char *mb[] = { "EUC_JP", "EUC_CN", "EUC_KR", "EUC_TW", "EUC_JIS_2004",
                                "MULE_INTERNAL", "SJIS", "BIG5", "GBK", "UHC", "GB18030",
                                "JOHAB", "SHIFT_JIS_2004", nullptr };
    if (!s->session_encode || s->session_encode + 16 > length)
        return ENC_NONE;
    char name[17];
    memcpy(name, (const char *)s + s->session_encode, 16);
    name[16] = 0;
    if (!name[0])
        return ENC_NONE;
    if (strcmp(name, "UTF8") == 0)
        return ENC_UTF8;
    for (int i = 0; mb[i]; i++)
        if (strcmp(name, mb[i]) == 0)
            return ENC_OTHER;
    return ENC_SINGLE;
}

/* kind and byte width of a HashedArrayOp key or result type (XHA_KIND_*) */
bool
haop_kind(int t, int *kind, int *width)
{
    switch (type_domain(t))
    {
        case XD_INT: *kind = XHA_KIND_INT; *width = type_width(t); return *width > 0;
        case XD_F32: *kind = XHA_KIND_F32; *width = 4; return true;
        case XD_F64: *kind = XHA_KIND_F64; *width = 8; return true;
        case XD_TEXT: *kind = XHA_KIND_TEXT; *width = 0; return true;
    }
    return false;
}

struct Compiler
{
    const kern_session_info *s;
    size_t      length;
    xpu_compiled *out;
    xplan      *P;
    std::vector<xop> ops;
    int         slot_reg[256];      /* -1 until first use */
    int         param_index[XP_MAX_PARAMS];     /* param_id -> XO_PARAM index, -1 */
    int         slot_type[256];
    int         nslots = 0;
    int         slot_top = 3;       /* R0 TRUE, R1 result, R2 tuple */
    int         temp_next = XP_NREGS;   /* temporaries grow down per segment */
    int         temp_min = XP_NREGS;
    bool        compare_nulls = false;
    int         enc = 0;            /* ENC_*: the session's database encoding */
    std::string err;

    bool fail(const std::string &m)
    {
        if (err.empty())
            err = m;
        return false;
    }

    /* temporaries: allocated downward from the top of the register file and
     * released at the end of each segment */
    int alloc_reg(int n = 1)
    {
        temp_next -= n;
        if (temp_next < slot_top)
        {
            fail("expression needs more than the per-lane register file");
            temp_next = XP_NREGS - n;
        }
        if (temp_next < temp_min)
            temp_min = temp_next;
        return temp_next;
    }

    /* slot registers: allocated upward on first use, live for the task */
    int slot_register(int slot)
    {
        if (slot_reg[slot] < 0)
        {
            int dom = type_domain(slot_type[slot]);
            slot_reg[slot] = slot_top;
            slot_top += dom_regs(dom);
            if (slot_top > temp_next)
                fail("expression needs more than the per-lane register file");
        }
        return slot_reg[slot];
    }

    int msg(const std::string &m)
    {
        for (size_t i = 0; i < out->messages.size(); i++)
            if (out->messages[i] == m)
                return (int)i;
        out->messages.push_back(m);
        return (int)out->messages.size() - 1;
    }

    void emit(int op, int dst, int a = 0, int b = 0, int guard = XP_REG_TRUE,
              int aux = 0, int64_t imm = 0)
    {
        xop o;
        memset(&o, 0, sizeof(o));
        o.op = op;
        o.dst = dst;
        o.a = a;
        o.b = b;
        o.guard = guard;
        o.aux = aux;
        o.imm = imm;
        ops.push_back(o);
    }

    xseg begin_seg()
    {
        xseg g;
        g.begin = (uint16_t)ops.size();
        g.end = g.begin;
        temp_next = XP_NREGS;
        return g;
    }

    void end_seg(xseg &g)
    {
        g.end = (uint16_t)ops.size();
    }

    /* ------------------------------------------------------------------ */
    const kern_expression *kexp_at(uint32_t off)
    {
        if (off == 0)
            return nullptr;
        if (off + sizeof(kern_expression) > length)
        {
            fail("xpucode offset out of the session blob");
            return nullptr;
        }
        return (const kern_expression *)((const char *)s + off);
    }

    bool check_node(const kern_expression *parent, const kern_expression *k)
    {
        const char *base = (const char *)s;
        if ((const char *)k < base || (const char *)k + offsetof(kern_expression, u) > base + length)
            return fail("kern_expression out of the session blob");
        if (k->len < offsetof(kern_expression, u) + 4 ||
            (const char *)k + k->len > base + length)
            return fail("kern_expression has a broken length");
        if (parent && (const char *)k + k->len > (const char *)parent + parent->len)
            return fail("kern_expression argument overruns its parent");
        uint32_t magic;
        memcpy(&magic, (const char *)k + k->len - 4, 4);
        uint32_t expect = KERN_EXPRESSION_MAGIC ^ ((uint32_t)k->exptype << 6) ^
                          ((uint32_t)k->opcode << 14);
        if (magic != expect)
            return fail("kern_expression magic mismatch (corrupted xpucode)");
        return true;
    }

    const kern_expression *first_arg(const kern_expression *k)
    {
        if (k->nr_args == 0 || k->args_offset == 0)
            return nullptr;
        return (const kern_expression *)((const char *)k + k->args_offset);
    }

    const kern_expression *next_arg(const kern_expression *a)
    {
        return (const kern_expression *)((const char *)a + XPU_MAXALIGN(a->len));
    }

    const kern_expression *packed_sub(const kern_expression *kp, int idx)
    {
        if (!kp || kp->opcode != FuncOpCode__Packed || idx < 0 ||
            (uint32_t)idx > kp->u.pack.npacked)
            return nullptr;
        uint32_t off = kp->u.pack.offset[idx];
        if (off == 0)
            return nullptr;
        const kern_expression *k = (const kern_expression *)((const char *)kp + off);
        return check_node(kp, k) ? k : nullptr;
    }

    /* ------------------------------------------------------------------ */
    Reg reg_of_slot(int slot)
    {
        int dom = type_domain(slot_type[slot]);
        if (dom < 0)
        {
            fail("slot of a type this build does not execute");
            return Reg{0, 0, 0};
        }
        return Reg{slot_register(slot), dom, slot_type[slot]};
    }

    Reg const_value(int type, bool isnull, const char *val)
    {
        int dom = type_domain(type);
        if (dom < 0)
        {
            fail("constant of unsupported type");
            return Reg{0, 0, 0};
        }
        int r = alloc_reg(dom_regs(dom));
        if (isnull)
        {
            emit(XO_CONST, r, 0, 0, XP_REG_TRUE, 1, 0);
            return Reg{r, dom, type};
        }
        if (dom == XD_NUM)
        {
            /* the varlena as it is, header included: it is moved and
             * written back unparsed (pgfn_ConstExpr reads it like a heap
             * datum) */
            uint8_t b0 = (uint8_t)val[0];
            if (b0 == 0x01 || (b0 & 0x03) == 0x02)
            {
                fail("compressed or external constant");
                return Reg{r, dom, type};
            }
            uint32_t sz;
            if (b0 & 0x01)
                sz = (b0 >> 1) & 0x7f;
            else
            {
                uint32_t h;
                memcpy(&h, val, 4);
                sz = (h >> 2) & 0x3fffffff;
            }
            if (sz < 1 || (uint64_t)(val - (const char *)s) + sz > length)
            {
                fail("numeric constant out of the session blob");
                return Reg{r, dom, type};
            }
            while (out->const_pool.size() % 8)
                out->const_pool.push_back(0);
            uint64_t off = out->const_pool.size();
            out->const_pool.insert(out->const_pool.end(), val, val + sz);
            while (out->const_pool.size() % 8)
                out->const_pool.push_back(0);
            emit(XO_CONST_NUM, r, 0, 0, XP_REG_TRUE, 0, (int64_t)off);
            return Reg{r, dom, type};
        }
        int64_t imm = 0;
        switch (type)
        {
            case TypeOpCode__bool: imm = *(const uint8_t *)val ? 1 : 0; break;
            case TypeOpCode__int1: imm = *(const int8_t *)val; break;
            case TypeOpCode__int2: { int16_t v; memcpy(&v, val, 2); imm = v; } break;
            case TypeOpCode__int4: case TypeOpCode__date: { int32_t v; memcpy(&v, val, 4); imm = v; } break;
            case TypeOpCode__int8: memcpy(&imm, val, 8); break;
            case TypeOpCode__float4: { uint32_t v; memcpy(&v, val, 4); imm = v; } break;
            case TypeOpCode__float8: memcpy(&imm, val, 8); break;
            default:
            {
                /* varlena: the payload goes to the constant pool */
                uint8_t b0 = (uint8_t)val[0];
                const char *p;
                uint32_t len;
                if (b0 == 0x01 || (b0 & 0x03) == 0x02)
                {
                    fail("compressed or external constant");
                    return Reg{r, dom, type};
                }
                if (b0 & 0x01)
                {
                    len = ((b0 >> 1) & 0x7f) - 1;
                    p = val + 1;
                }
                else
                {
                    uint32_t h;
                    memcpy(&h, val, 4);
                    len = ((h >> 2) & 0x3fffffff) - 4;
                    p = val + 4;
                }
                uint64_t off = out->const_pool.size();
                out->const_pool.insert(out->const_pool.end(), p, p + len);
                while (out->const_pool.size() % 8)
                    out->const_pool.push_back(0);
                emit(XO_CONST_TEXT, r, 0, 0, XP_REG_TRUE, 0, (int64_t)(off | ((uint64_t)len << 32)));
                return Reg{r, dom, type};
            }
        }
        emit(XO_CONST, r, 0, 0, XP_REG_TRUE, 0, imm);
        return Reg{r, dom, type};
    }

    Reg to_f64(Reg x)
    {
        if (x.dom == XD_F64)
            return x;
        int r = alloc_reg();
        emit(x.dom == XD_F32 ? XO_F32_F64 : XO_I_F64, r, x.r);
        return Reg{r, XD_F64, TypeOpCode__float8};
    }

    /* compile expression k; errors in it are raised only where R[guard] */
    Reg expr(const kern_expression *parent, const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        if (!err.empty() || !check_node(parent, k))
            return bad;
        switch (k->opcode)
        {
            case FuncOpCode__ConstExpr:
                return const_value(k->exptype, k->u.c.const_isnull, k->u.c.const_value);
            case FuncOpCode__ParamExpr:
            {
                uint32_t id = k->u.p.param_id;
                if (id >= s->nparams)
                {
                    fail("ParamExpr refers to an undefined parameter");
                    return bad;
                }
                uint32_t off = s->poffset[id];
                if (off && off >= length)
                {
                    fail("parameter offset out of the session blob");
                    return bad;
                }
                int dom = type_domain(k->exptype);
                if (dom < 0 || dom == XD_TEXT || dom == XD_NUM || id >= XP_MAX_PARAMS)
                    return const_value(k->exptype, off == 0, (const char *)s + off);
                /* fixed width: a run-time value of the launch */
                if (param_index[id] < 0)
                {
                    if (out->params.size() >= XP_MAX_PARAMS)
                        return const_value(k->exptype, off == 0, (const char *)s + off);
                    int64_t v = 0;
                    const char *val = (const char *)s + off;
                    if (off)
                    {
                        switch (k->exptype)
                        {
                            case TypeOpCode__bool: v = *(const uint8_t *)val ? 1 : 0; break;
                            case TypeOpCode__int1: v = *(const int8_t *)val; break;
                            case TypeOpCode__int2: { int16_t x; memcpy(&x, val, 2); v = x; } break;
                            case TypeOpCode__int4: case TypeOpCode__date:
                                { int32_t x; memcpy(&x, val, 4); v = x; } break;
                            case TypeOpCode__float4: { uint32_t x; memcpy(&x, val, 4); v = x; } break;
                            default: memcpy(&v, val, 8); break;
                        }
                    }
                    param_index[id] = (int)out->params.size();
                    if (!off)
                        out->param_nulls |= 1ULL << out->params.size();
                    out->params.push_back(v);
                }
                int r = alloc_reg();
                emit(XO_PARAM, r, 0, 0, XP_REG_TRUE, 0, param_index[id]);
                return Reg{r, dom, k->exptype};
            }
            case FuncOpCode__VarExpr:
            {
                int slot = k->u.v.var_slot_id;
                if (slot >= nslots)
                {
                    fail("VarExpr refers to an undefined slot");
                    return bad;
                }
                return reg_of_slot(slot);
            }
            case FuncOpCode__SaveExpr:
            {
                int slot = k->u.save.sv_slot_id;
                if (slot >= nslots || k->nr_args != 1)
                {
                    fail("malformed SaveExpr");
                    return bad;
                }
                Reg v = expr(k, first_arg(k), guard);
                Reg sr = reg_of_slot(slot);
                if (v.r != sr.r)
                {
                    if (v.dom != sr.dom)
                    {
                        fail("SaveExpr type mismatch");
                        return bad;
                    }
                    emit(mov_op(sr.dom), sr.r, v.r);
                }
                return sr;
            }
            case FuncOpCode__BoolExpr_And:
            case FuncOpCode__BoolExpr_Or:
            {
                bool is_and = (k->opcode == FuncOpCode__BoolExpr_And);
                int acc = alloc_reg();
                int gr = (k->nr_args > 1) ? alloc_reg() : XP_REG_TRUE;
                int mark = temp_next;
                int g = guard;
                emit(XO_CONST, acc, 0, 0, XP_REG_TRUE, 0, is_and ? 1 : 0);
                const kern_expression *a = first_arg(k);
                for (uint32_t i = 0; i < k->nr_args; i++, a = next_arg(a))
                {
                    Reg v = expr(k, a, g);
                    emit(is_and ? XO_AND3 : XO_OR3, acc, acc, v.r);
                    temp_next = mark;
                    if (i + 1 < k->nr_args)
                    {
                        /* later arguments run only where the result is still open */
                        emit(is_and ? XO_G_NOTFALSE : XO_G_NOTTRUE, gr, acc, g);
                        g = gr;
                    }
                }
                temp_next = acc;
                return Reg{acc, XD_INT, TypeOpCode__bool};
            }
            case FuncOpCode__BoolExpr_Not:
            {
                int r = alloc_reg();
                int mark = temp_next;
                Reg v = expr(k, first_arg(k), guard);
                emit(XO_NOT3, r, v.r);
                temp_next = mark;
                return Reg{r, XD_INT, TypeOpCode__bool};
            }
            case FuncOpCode__NullTestExpr_IsNull:
            case FuncOpCode__NullTestExpr_IsNotNull:
            {
                int r = alloc_reg();
                int mark = temp_next;
                Reg v = expr(k, first_arg(k), guard);
                emit(k->opcode == FuncOpCode__NullTestExpr_IsNull ? XO_ISNULL : XO_ISNOTNULL, r, v.r);
                temp_next = mark;
                return Reg{r, XD_INT, TypeOpCode__bool};
            }
            case FuncOpCode__BoolTestExpr_IsTrue:
            case FuncOpCode__BoolTestExpr_IsNotTrue:
            case FuncOpCode__BoolTestExpr_IsFalse:
            case FuncOpCode__BoolTestExpr_IsNotFalse:
            case FuncOpCode__BoolTestExpr_IsUnknown:
            case FuncOpCode__BoolTestExpr_IsNotUnknown:
            {
                int r = alloc_reg();
                int mark = temp_next;
                Reg v = expr(k, first_arg(k), guard);
                emit(XO_BOOLTEST, r, v.r, 0, XP_REG_TRUE,
                     k->opcode - FuncOpCode__BoolTestExpr_IsTrue);
                temp_next = mark;
                return Reg{r, XD_INT, TypeOpCode__bool};
            }
            case FuncOpCode__CoalesceExpr:
            {
                int dom = type_domain(k->exptype);
                if (dom < 0)
                {
                    fail("COALESCE of unsupported type");
                    return bad;
                }
                int r = alloc_reg(dom_regs(dom));
                int gr = (k->nr_args > 1) ? alloc_reg() : XP_REG_TRUE;
                int mark = temp_next;
                int g = guard;
                emit(XO_SETNULL, r);
                const kern_expression *a = first_arg(k);
                for (uint32_t i = 0; i < k->nr_args; i++, a = next_arg(a))
                {
                    Reg v = expr(k, a, g);
                    if (v.dom != dom)
                    {
                        fail("COALESCE argument type mismatch");
                        return bad;
                    }
                    emit(XO_COALESCE, r, v.r, dom_regs(dom) - 1);
                    temp_next = mark;
                    if (i + 1 < k->nr_args)
                    {
                        emit(XO_G_ISNULL, gr, r, g);
                        g = gr;
                    }
                }
                temp_next = r;
                return Reg{r, dom, (int)k->exptype};
            }
            case FuncOpCode__HashValue:
            {
                int h = alloc_reg();
                int mark = temp_next;
                emit(XO_HASH_INIT, h);
                const kern_expression *a = first_arg(k);
                for (uint32_t i = 0; i < k->nr_args; i++, a = next_arg(a), temp_next = mark)
                {
                    Reg v = expr(k, a, guard);
                    int w = (v.dom == XD_TEXT) ? 0 : (v.dom == XD_NUM) ? XHASH_NUMERIC
                                                                        : type_width(a->exptype);
                    if (v.dom != XD_TEXT && w < 0)
                    {
                        fail("hash key of unsupported type");
                        return bad;
                    }
                    emit(XO_HASH_ADD, h, v.r, v.dom, XP_REG_TRUE, w);
                }
                emit(XO_HASH_FIN, h);
                return Reg{h, XD_INT, TypeOpCode__int4};
            }
            case FuncOpCode__DistinctFrom:
                return distinct_from(k, guard);
            case FuncOpCode__LeastExpr:
            case FuncOpCode__GreatestExpr:
                return minmax(k, guard);
            case FuncOpCode__CaseWhenExpr:
                return casewhen(k, guard);
            case FuncOpCode__ScalarArrayOpAny:
            case FuncOpCode__ScalarArrayOpAll:
                return scalar_array_op(k, guard);
            case FuncOpCode__HashedArrayOp:
                return hashed_array_op(k, guard);
            default:
                return func(k, guard);
        }
    }

    /* text functions count characters by the session's encoding; the
     * reference raises this message when it has none (xpu_textlib.cu:600) */
    bool need_encoding(const std::string &fname)
    {
        if (enc == ENC_NONE)
            return fail("No encoding info was supplied");
        if (enc == ENC_OTHER)
            return fail(fname + ": this database encoding is not supported by this build");
        return true;
    }

    bool in_node(const kern_expression *k, uint32_t off, size_t need)
    {
        return off >= offsetof(kern_expression, u) && off + need <= k->len;
    }

    /* IS DISTINCT FROM (pgfn_DistinctFrom, xpu_common.cu:613-660): the
     * arguments of its '=' function, the function, then the NULL rules */
    Reg distinct_from(const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        const kern_expression *kcmp = first_arg(k);
        if (k->nr_args != 1 || !kcmp || !check_node(k, kcmp) || kcmp->nr_args != 2)
        {
            fail("malformed DistinctFrom");
            return bad;
        }
        const devfunc *f = lookup_devfunc(kcmp->opcode);
        if (!f || f->nargs != 2 || f->res != T_bool)
        {
            fail("DistinctFrom needs a supported comparison function");
            return bad;
        }
        int r = alloc_reg();
        int mark = temp_next;
        Reg av[3];
        const kern_expression *ka = first_arg(kcmp);
        av[0] = expr(kcmp, ka, guard);
        av[1] = expr(kcmp, next_arg(ka), guard);
        if (!err.empty())
            return bad;
        func_op(f, av, r, guard);
        emit(XO_DISTINCT, r, av[0].r, av[1].r);
        temp_next = mark;
        return Reg{r, XD_INT, TypeOpCode__bool};
    }

    /* LEAST / GREATEST (xpu_common.cu:683-757) */
    Reg minmax(const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        int dom = type_domain(k->exptype);
        if (dom < 0)
        {
            fail("LEAST/GREATEST of unsupported type");
            return bad;
        }
        int r = alloc_reg(dom_regs(dom));
        int mark = temp_next;
        int flags = dom | (k->opcode == FuncOpCode__GreatestExpr ? XMM_GREATEST : 0) |
                    (k->exptype == T_bpchar ? XMM_BPCHAR : 0);
        emit(XO_SETNULL, r);
        const kern_expression *a = first_arg(k);
        for (uint32_t i = 0; i < k->nr_args; i++, a = next_arg(a))
        {
            Reg v = expr(k, a, guard);
            if (!err.empty())
                return bad;
            if (v.dom != dom)
            {
                fail("LEAST/GREATEST argument type mismatch");
                return bad;
            }
            emit(XO_MINMAX, r, v.r, 0, guard, flags);
            temp_next = mark;
        }
        return Reg{r, dom, (int)k->exptype};
    }

    /* CASE [key] WHEN ... THEN ... [ELSE] END (pgfn_CaseWhenExpr,
     * xpu_common.cu:759-811): an arm's result is evaluated, and may raise
     * errors, only on the rows no earlier arm took */
    Reg casewhen(const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        int dom = type_domain(k->exptype);
        if (dom < 0 || (k->nr_args % 2) != 0)
        {
            fail("CASE of unsupported type or malformed CaseWhenExpr");
            return bad;
        }
        int r = alloc_reg(dom_regs(dom));
        int gopen = alloc_reg();
        int gthen = alloc_reg();
        int mark = temp_next;
        emit(XO_SETNULL, r);
        emit(XO_MOV, gopen, guard);
        if (k->u.casewhen.case_comp)
        {
            if (!in_node(k, k->u.casewhen.case_comp, offsetof(kern_expression, u)))
            {
                fail("CASE key out of the expression");
                return bad;
            }
            const kern_expression *kc =
                (const kern_expression *)((const char *)k + k->u.casewhen.case_comp);
            if (!check_node(k, kc) || kc->opcode != FuncOpCode__SaveExpr)
            {
                fail("CASE key must be a SaveExpr");
                return bad;
            }
            expr(k, kc, gopen);
            temp_next = mark;
        }
        const kern_expression *a = first_arg(k);
        for (uint32_t i = 0; i + 1 < k->nr_args && err.empty(); i += 2)
        {
            Reg c = expr(k, a, gopen);
            emit(XO_G_TRUE, gthen, c.r, gopen);
            emit(XO_G_NOTTRUE, gopen, c.r, gopen);
            temp_next = mark;
            a = next_arg(a);
            Reg v = expr(k, a, gthen);
            if (err.empty() && v.dom != dom)
            {
                fail("CASE result type mismatch");
                return bad;
            }
            emit(XO_SELECT, r, v.r, 0, gthen, dom_regs(dom) - 1);
            temp_next = mark;
            a = next_arg(a);
        }
        if (k->u.casewhen.case_else && err.empty())
        {
            if (!in_node(k, k->u.casewhen.case_else, offsetof(kern_expression, u)))
            {
                fail("CASE ELSE out of the expression");
                return bad;
            }
            const kern_expression *ke =
                (const kern_expression *)((const char *)k + k->u.casewhen.case_else);
            if (!check_node(k, ke))
                return bad;
            Reg v = expr(k, ke, gopen);
            if (err.empty() && v.dom != dom)
            {
                fail("CASE ELSE type mismatch");
                return bad;
            }
            emit(XO_SELECT, r, v.r, 0, gopen, dom_regs(dom) - 1);
        }
        temp_next = r;
        return Reg{r, dom, (int)k->exptype};
    }

    /* scalar op ANY / ALL over a constant array (pgfn_ScalarArrayOp,
     * xpu_common.cu:815-1011), unrolled: per element the element slot takes
     * the element and the comparator runs; ANY accumulates with OR and
     * ALL with AND in SQL three-valued logic, and later elements run only
     * where the result is still open.  That is PostgreSQL's ANY / ALL
     * (the reference's ALL returns true where a comparison is false). */
#define XP_MAX_SAOP_ITEMS   128
    Reg scalar_array_op(const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        const bool use_any = (k->opcode == FuncOpCode__ScalarArrayOpAny);
        const kern_expression *karr = first_arg(k);
        const kern_expression *kcmp = karr ? next_arg(karr) : nullptr;
        if (k->nr_args != 2 || !karr || !check_node(k, karr) || !check_node(k, kcmp) ||
            kcmp->nr_args != 2)
        {
            fail("malformed ScalarArrayOp");
            return bad;
        }
        if (karr->opcode != FuncOpCode__ConstExpr || karr->exptype != TypeOpCode__array)
        {
            fail("ScalarArrayOp over a non-constant array is not supported by this build");
            return bad;
        }
        int slot = k->u.saop.elem_slot_id;
        const kern_expression *kelem = first_arg(kcmp) ? next_arg(first_arg(kcmp)) : nullptr;
        if (slot >= nslots || !kelem || !check_node(kcmp, kelem) ||
            kelem->opcode != FuncOpCode__VarExpr || kelem->u.v.var_slot_id != slot)
        {
            fail("ScalarArrayOp comparator must read the element slot");
            return bad;
        }
        int etype = slot_type[slot], edom = type_domain(etype);
        if (edom < 0)
        {
            fail("ScalarArrayOp element of unsupported type");
            return bad;
        }
        int acc = alloc_reg();
        int gr = alloc_reg();
        int mark = temp_next;
        if (karr->u.c.const_isnull)
        {
            emit(XO_CONST, acc, 0, 0, XP_REG_TRUE, 1, 0);
            temp_next = mark;
            return Reg{acc, XD_INT, TypeOpCode__bool};
        }
        /* the array image (ArrayType, utils/array.h) */
        const char *img = karr->u.c.const_value;
        const char *kend = (const char *)karr + karr->len - 4;
        if (img + 16 > kend)
        {
            fail("ScalarArrayOp array constant out of its expression");
            return bad;
        }
        uint32_t hdr;
        int32_t ndim, dataoffset;
        memcpy(&hdr, img, 4);
        memcpy(&ndim, img + 4, 4);
        memcpy(&dataoffset, img + 8, 4);
        const char *aend = img + ((hdr >> 2) & 0x3fffffff);
        if ((hdr & 3) != 0 || aend > kend || ndim < 0 || ndim > 6 || img + 16 + 8 * ndim > aend)
        {
            fail("ScalarArrayOp array constant is malformed");
            return bad;
        }
        int64_t nitems = ndim > 0 ? 1 : 0;
        for (int d = 0; d < ndim; d++)
        {
            int32_t dim;
            memcpy(&dim, img + 16 + 4 * d, 4);
            nitems *= dim < 0 ? 0 : dim;
        }
        if (nitems > XP_MAX_SAOP_ITEMS)
        {
            fail("ScalarArrayOp over more than 128 elements is not supported by this build");
            return bad;
        }
        const uint8_t *nullmap = dataoffset ? (const uint8_t *)img + 16 + 8 * ndim : nullptr;
        const char *data = img + (dataoffset ? dataoffset : XPU_MAXALIGN(16 + 8 * ndim));
        if ((nullmap && (const char *)nullmap + (nitems + 7) / 8 > aend) || data > aend)
        {
            fail("ScalarArrayOp array constant is malformed");
            return bad;
        }
        const int typlen = type_width(etype);
        const int align = (etype == T_bool || etype == TypeOpCode__int1) ? 1
                        : typlen > 0 ? typlen : 4;
        int er = slot_register(slot);
        emit(XO_CONST, acc, 0, 0, XP_REG_TRUE, 0, use_any ? 0 : 1);
        int g = guard;
        size_t off = 0;
        for (int64_t i = 0; i < nitems && err.empty(); i++)
        {
            if (nullmap && !(nullmap[i >> 3] & (1 << (i & 7))))
                emit(XO_CONST, er, 0, 0, XP_REG_TRUE, 1, 0);
            else
            {
                /* the element, aligned unless a varlena starts here without a pad byte */
                if (typlen > 0 || data[off] == 0)
                    off = XPU_TYPEALIGN(align, off);
                const char *ep = data + off;
                size_t esz;
                if (typlen > 0)
                    esz = typlen;
                else
                {
                    uint8_t b0 = (uint8_t)ep[0];
                    if (b0 & 1)
                        esz = (b0 >> 1) & 0x7f;
                    else
                    {
                        uint32_t h;
                        memcpy(&h, ep, 4);
                        esz = (h >> 2) & 0x3fffffff;
                    }
                }
                if (ep + esz > aend || esz == 0)
                {
                    fail("ScalarArrayOp array element out of the array");
                    return bad;
                }
                temp_next = mark;
                Reg cv = const_value(etype, false, ep);
                emit(mov_op(edom), er, cv.r);
                off += esz;
            }
            temp_next = mark;
            Reg v = expr(k, kcmp, g);
            emit(use_any ? XO_OR3 : XO_AND3, acc, acc, v.r);
            temp_next = mark;
            if (i + 1 < nitems)
            {
                emit(use_any ? XO_G_NOTTRUE : XO_G_NOTFALSE, gr, acc, g);
                g = gr;
            }
        }
        temp_next = mark;
        return Reg{acc, XD_INT, TypeOpCode__bool};
    }

    /* HashedArrayOp (codegen.c:2820-2906 and 2458-2660): the expression's
     * hash table of constant keys, copied into the constant pool and walked
     * on the device by XO_HAOP.  The table is checked here: every chain stays
     * inside the expression and the walk of all chains takes at most as many
     * steps as there are items. */
    Reg hashed_array_op(const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        const uint32_t nslots = k->u.haop.nslots;
        const uint32_t head = (uint32_t)offsetof(kern_expression, u.haop.hslots);
        const uint32_t aoff = k->args_offset;
        int kkind, kwidth, rkind, rwidth;
        if (k->nr_args != 1 || nslots == 0 || nslots > (1U << 20) ||
            (uint64_t)head + 4ULL * nslots > aoff || aoff > k->len)
        {
            fail("malformed HashedArrayOp");
            return bad;
        }
        if (!haop_kind(k->u.haop.elem_type, &kkind, &kwidth) ||
            !haop_kind(k->exptype, &rkind, &rwidth))
        {
            fail("HashedArrayOp of unsupported type");
            return bad;
        }
        const char *base = (const char *)k;
        const uint32_t dflt = k->u.haop.default_offset;
        const bool in_list = (dflt == (uint32_t)offsetof(kern_expression, u.haop.__padding));
        if (dflt && !in_list && (dflt < head + 4 * nslots || dflt + 8 > aoff))
        {
            fail("HashedArrayOp default value out of the expression");
            return bad;
        }
        uint32_t nitems = 0;
        bool null_key = false;
        const uint32_t max_items = aoff / 8;
        for (uint32_t sl = 0; sl < nslots; sl++)
        {
            uint32_t next = k->u.haop.hslots[sl];
            while (next != 0)
            {
                if ((next & KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK) || next < head + 4 * nslots ||
                    next + 8 > aoff || ++nitems > max_items)
                {
                    fail("HashedArrayOp hash chain out of the expression");
                    return bad;
                }
                uint32_t inext;
                memcpy(&inext, base + next + 4, 4);
                uint32_t pos = next + 8;
                if (inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY)
                {
                    uint32_t klen = (uint32_t)kwidth;
                    if (kkind == XHA_KIND_TEXT)
                    {
                        uint8_t b0 = (uint8_t)base[pos];
                        uint32_t h;
                        memcpy(&h, base + pos, 4);
                        klen = (b0 & 1) ? ((b0 >> 1) & 0x7f) : ((h >> 2) & 0x3fffffff);
                        if (b0 == 0x01 || (b0 & 3) == 2)
                            klen = 0;
                    }
                    if (klen == 0 || pos + klen > aoff)
                    {
                        fail("HashedArrayOp key out of the expression");
                        return bad;
                    }
                    pos += klen;
                }
                else
                    null_key = true;
                if (inext & KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE)
                {
                    uint32_t ralign = rkind == XHA_KIND_TEXT ? 4 : (uint32_t)rwidth;
                    uint32_t koff = pos - (next + 8);
                    pos = next + 8 + (koff + ralign - 1) / ralign * ralign;
                    if (pos + (rkind == XHA_KIND_TEXT ? 4 : (uint32_t)rwidth) > aoff)
                    {
                        fail("HashedArrayOp value out of the expression");
                        return bad;
                    }
                }
                next = inext & ~KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK;
            }
        }
        /* the header and items into the constant pool, with the item count
         * in the (device-unused) elem_ops field */
        uint64_t poff = out->const_pool.size();
        out->const_pool.insert(out->const_pool.end(), base, base + aoff);
        memcpy(&out->const_pool[poff + 40], &nitems, 4);
        while (out->const_pool.size() % 8)
            out->const_pool.push_back(0);
        int r = alloc_reg(rkind == XHA_KIND_TEXT ? 2 : 1);
        int mark = temp_next;
        Reg sv = expr(k, first_arg(k), guard);
        if (!err.empty())
            return bad;
        if (sv.dom != type_domain(k->u.haop.elem_type))
        {
            fail("HashedArrayOp scalar type mismatch");
            return bad;
        }
        int aux = XHA_KEY(kkind, kwidth) | XHA_RES(rkind, rwidth) |
                  (in_list && null_key && k->exptype == T_bool ? XHA_NULL_MISS : 0);
        emit(XO_HAOP, r, sv.r, 0, XP_REG_TRUE, aux, (int64_t)poff);
        temp_next = mark;
        return Reg{r, rkind == XHA_KIND_TEXT ? XD_TEXT : type_domain(k->exptype), (int)k->exptype};
    }

    Reg func(const kern_expression *k, int guard)
    {
        Reg bad{0, 0, 0};
        const devfunc *f = lookup_devfunc(k->opcode);
        if (!f)
        {
            char buf[96];
            snprintf(buf, sizeof(buf), "device function opcode %d is not supported", (int)k->opcode);
            fail(buf);
            return bad;
        }
        if ((int)k->nr_args != f->nargs)
        {
            fail(std::string(f->name) + ": wrong number of arguments");
            return bad;
        }
        /* the result register comes first, so the arguments' temporaries
         * are released once the operator has consumed them */
        int r = alloc_reg(dom_regs(type_domain(f->res)));
        int mark = temp_next;
        Reg a[3];
        const kern_expression *ka = first_arg(k);
        for (int i = 0; i < f->nargs; i++, ka = next_arg(ka))
            a[i] = expr(k, ka, guard);
        if (!err.empty())
            return bad;
        Reg res = func_op(f, a, r, guard);
        temp_next = mark;
        return res;
    }

    Reg func_op(const devfunc *f, Reg *a, int r, int guard)
    {
        Reg bad{0, 0, 0};
        int cmpop = -1;
        switch (f->op)
        {
            case O_eq: cmpop = XC_EQ; break;
            case O_ne: cmpop = XC_NE; break;
            case O_lt: cmpop = XC_LT; break;
            case O_le: cmpop = XC_LE; break;
            case O_gt: cmpop = XC_GT; break;
            case O_ge: cmpop = XC_GE; break;
        }
        int cn = compare_nulls ? 0x100 : 0;
        std::string fname(f->name);
        switch (f->kind)
        {
            case K_cmp:
                if (a[0].dom == XD_INT && a[1].dom == XD_INT)
                    emit(XO_CMP_I, r, a[0].r, a[1].r, XP_REG_TRUE, cmpop | cn);
                else if (a[0].dom == XD_F32 && a[1].dom == XD_F32)
                    emit(XO_CMP_F32, r, a[0].r, a[1].r, XP_REG_TRUE, cmpop | cn);
                else
                {
                    Reg x = to_f64(a[0]), y = to_f64(a[1]);
                    emit(XO_CMP_F64, r, x.r, y.r, XP_REG_TRUE, cmpop | cn);
                }
                return Reg{r, XD_INT, TypeOpCode__bool};
            case K_textcmp:
                /* bpchar compares without trailing spaces (xpu_textlib.cu:59-69) */
                emit(XO_CMP_TEXT, r, a[0].r, a[1].r, XP_REG_TRUE,
                     cmpop | cn | (f->t0 == T_bpchar ? XCMP_BPCHAR : 0));
                return Reg{r, XD_INT, TypeOpCode__bool};
            case K_textlen:
            {
                /* pgfn_textlen / pgfn_bpcharlen (xpu_textlib.cu:583-720) */
                if (!need_encoding(fname))
                    return bad;
                emit(XO_TEXTLEN, r, a[0].r, 0, XP_REG_TRUE,
                     (enc == ENC_UTF8 ? XENC_UTF8 : 0) | (f->t0 == T_bpchar ? XTXT_BPCHAR : 0));
                return Reg{r, XD_INT, TypeOpCode__int4};
            }
            case K_like:
            {
                /* PG_TEXTLIKE_TEMPLATE / PG_BPCHARLIKE_TEMPLATE (xpu_textlib.cu:1087-1155) */
                if (!need_encoding(fname))
                    return bad;
                int flags = (enc == ENC_UTF8 ? XENC_UTF8 : 0) |
                            (f->t0 == T_bpchar ? XTXT_BPCHAR : 0) |
                            ((f->op == O_ilike || f->op == O_nilike) ? XLIKE_ICASE : 0) |
                            ((f->op == O_nlike || f->op == O_nilike) ? XLIKE_NEGATE : 0);
                emit(XO_LIKE, r, a[0].r, a[1].r, guard, flags, msg("invalid escape in LIKE pattern"));
                return Reg{r, XD_INT, TypeOpCode__bool};
            }
            case K_substr:
            {
                /* pgfn_substring / pgfn_substring_nolen (xpu_textlib.cu:1261-1322) */
                if (!need_encoding(fname))
                    return bad;
                int64_t imm = f->nargs > 2
                    ? ((int64_t)a[2].r | XSUB_HAS_COUNT |
                       ((int64_t)msg("negative substring length not allowed") << 16))
                    : 0;
                emit(XO_SUBSTR, r, a[0].r, a[1].r, guard, enc == ENC_UTF8 ? XENC_UTF8 : 0, imm);
                return Reg{r, XD_TEXT, TypeOpCode__text};
            }
            case K_numcmp:
                /* PG_NUMERIC_COMPARE_TEMPLATE (xpu_numeric.cu:541-568) */
                emit(XO_NUM_CMP, r, a[0].r, a[1].r, guard, cmpop | cn);
                return Reg{r, XD_INT, TypeOpCode__bool};
            case K_numarith:
            {
                /* pgfn_numeric_add .. _mod (xpu_numeric.cu:570-703); the
                 * reference's numeric_mod computes the division */
                int op = f->op == O_pl ? XO_NUM_ADD : f->op == O_mi ? XO_NUM_SUB :
                         f->op == O_mul ? XO_NUM_MUL : XO_NUM_DIV;
                emit(op, r, a[0].r, a[1].r, guard);
                return Reg{r, XD_NUM, TypeOpCode__numeric};
            }
            case K_numunary:
                /* numeric_uplus returns its argument as it is */
                if (f->op == O_up)
                    emit(XO_MOV3, r, a[0].r);
                else
                    emit(f->op == O_um ? XO_NUM_NEG : XO_NUM_ABS, r, a[0].r, 0, guard);
                return Reg{r, XD_NUM, TypeOpCode__numeric};
            case K_numcast:
            {
                int dd = type_domain(f->res);
                if (f->res == TypeOpCode__numeric)
                {
                    if (a[0].dom == XD_INT)
                        emit(XO_INT_NUM, r, a[0].r);
                    else
                        emit(a[0].dom == XD_F64 ? XO_F64_NUM : XO_F32_NUM, r, a[0].r);
                    return Reg{r, XD_NUM, TypeOpCode__numeric};
                }
                if (dd == XD_INT)
                    emit(XO_NUM_INT, r, a[0].r, 0, guard, int_class(f->res));
                else
                    emit(dd == XD_F64 ? XO_NUM_F64 : XO_NUM_F32, r, a[0].r, 0, guard);
                return Reg{r, dd, f->res};
            }
            case K_arith:
                if (type_domain(f->res) == XD_INT)
                {
                    static const int iops[] = {0, 0, 0, 0, 0, 0, 0, XO_ADD_I, XO_SUB_I, XO_MUL_I,
                                               XO_DIV_I, XO_MOD_I};
                    bool divlike = (f->op == O_div || f->op == O_mod);
                    /* message texts of xpu_basetype.cu:670, :761 */
                    int m = msg(fname + (f->op == O_mod ? " : division by zero" :
                                         divlike ? ": division by zero" : ": value out of range"));
                    int m2 = msg(fname + ": value out of range");
                    emit(iops[f->op], r, a[0].r, a[1].r, guard, int_class(f->res),
                         (int64_t)m | ((int64_t)m2 << 16));
                    return Reg{r, XD_INT, f->res};
                }
                if (f->res == TypeOpCode__float4)
                {
                    static const int fops[] = {0, 0, 0, 0, 0, 0, 0, XO_ADD_F32, XO_SUB_F32,
                                               XO_MUL_F32, XO_DIV_F32};
                    int m = msg(fname + ": value out of range");
                    int m2 = msg(fname + ": division by zero");
                    emit(fops[f->op], r, a[0].r, a[1].r, guard, 0, (int64_t)m | ((int64_t)m2 << 16));
                    return Reg{r, XD_F32, TypeOpCode__float4};
                }
                else
                {
                    static const int dops[] = {0, 0, 0, 0, 0, 0, 0, XO_ADD_F64, XO_SUB_F64,
                                               XO_MUL_F64, XO_DIV_F64};
                    Reg x = to_f64(a[0]), y = to_f64(a[1]);
                    int m = msg(fname + ": value out of range");
                    int m2 = msg(fname + ": division by zero");
                    emit(dops[f->op], r, x.r, y.r, guard, 0, (int64_t)m | ((int64_t)m2 << 16));
                    return Reg{r, XD_F64, TypeOpCode__float8};
                }
            case K_unary:
                if (a[0].dom == XD_INT)
                {
                    int op = f->op == O_up ? XO_MOV : f->op == O_um ? XO_NEG_I :
                             f->op == O_abs ? XO_ABS_I : XO_NOT_I;
                    emit(op, r, a[0].r, 0, guard, int_class(f->res), msg(fname + ": value out of range"));
                    return Reg{r, XD_INT, f->res};
                }
                if (a[0].dom == XD_F32)
                    emit(f->op == O_up ? XO_MOV : f->op == O_um ? XO_NEG_F32 : XO_ABS_F32, r, a[0].r);
                else
                    emit(f->op == O_up ? XO_MOV : f->op == O_um ? XO_NEG_F64 : XO_ABS_F64, r, a[0].r);
                return Reg{r, a[0].dom, f->res};
            case K_bitop:
            {
                int op = f->op == O_and ? XO_AND_I : f->op == O_or ? XO_OR_I :
                         f->op == O_xor ? XO_XOR_I : f->op == O_shl ? XO_SHL_I : XO_SHR_I;
                emit(op, r, a[0].r, a[1].r, XP_REG_TRUE, int_class(f->res));
                return Reg{r, XD_INT, f->res};
            }
            case K_cast:
            {
                int sd = a[0].dom, dd = type_domain(f->res);
                /* "<src> to <dst>: out of range" (xpu_basetype.cu:343) */
                std::string cname = fname;
                size_t sep = cname.find("_to_");
                if (sep != std::string::npos)
                    cname = cname.substr(0, sep) + " to " + cname.substr(sep + 4);
                int m = msg(cname + ": out of range");
                if (f->res == TypeOpCode__bool)
                    emit(XO_I_BOOL, r, a[0].r);
                else if (sd == XD_INT && dd == XD_INT)
                    emit(XO_I2I, r, a[0].r, 0, guard, int_class(f->res), m);
                else if (sd == XD_F32 && dd == XD_INT)
                    emit(XO_F32_I, r, a[0].r, 0, guard, int_class(f->res), m);
                else if (sd == XD_F64 && dd == XD_INT)
                    emit(XO_F64_I, r, a[0].r, 0, guard, int_class(f->res), m);
                else if (sd == XD_INT && dd == XD_F32)
                    emit(XO_I_F32, r, a[0].r);
                else if (sd == XD_INT && dd == XD_F64)
                    emit(XO_I_F64, r, a[0].r);
                else if (sd == XD_F32 && dd == XD_F64)
                    emit(XO_F32_F64, r, a[0].r);
                else if (sd == XD_F64 && dd == XD_F32)
                    emit(XO_F64_F32, r, a[0].r, 0, guard, 0, m);
                else
                {
                    fail(fname + ": unsupported cast");
                    return bad;
                }
                return Reg{r, dd, f->res};
            }
        }
        fail(fname + ": unsupported function kind");
        return bad;
    }

    /* LoadVars of one depth */
    bool loadvars(const kern_expression *kload, int depth, int r_tuple, bool nulls)
    {
        if (!kload)
            return true;
        if (kload->opcode != FuncOpCode__LoadVars)
            return fail("LoadVars expected");
        for (int i = 0; i < kload->u.load.nitems; i++)
        {
            int resno = kload->u.load.desc[i].vl_resno;
            int slot = kload->u.load.desc[i].vl_slot_id;
            if (slot < 0 || slot >= nslots || resno <= 0)
                return fail("LoadVars descriptor out of range");
            Reg sr = reg_of_slot(slot);
            if (sr.dom < 0)
                return fail("LoadVars of an unsupported type");
            if (nulls)
            {
                emit(XO_SETNULL, sr.r);
                continue;
            }
            if (depth == 0)
            {
                if (sr.dom == XD_TEXT)
                    emit(XO_LDCOL_VAR, sr.r, 0, 0, XP_REG_TRUE, resno - 1);
                else if (sr.dom == XD_NUM)
                    emit(XO_LDCOL_DEC, sr.r, 0, 0, XP_REG_TRUE, resno - 1);
                else if (sr.type == TypeOpCode__bool)
                    emit(XO_LDCOL_BOOL, sr.r, 0, 0, XP_REG_TRUE, resno - 1);
                else
                    emit(XO_LDCOL, sr.r, 0, sr.dom, XP_REG_TRUE, resno - 1, type_width(sr.type));
            }
            else
                emit(XO_LDTUP, sr.r, r_tuple, sr.type, XP_REG_TRUE, depth, resno);
        }
        return true;
    }

    /* ExecGpuJoinQuals (xpu_common.cu:1442-1481) as a status register */
    bool joinquals(const kern_expression *kj, int r_status, bool other_only)
    {
        emit(XO_CONST, r_status, 0, 0, XP_REG_TRUE, 0, 1);
        if (!kj)
            return true;
        if (kj->opcode != FuncOpCode__JoinQuals)
            return fail("JoinQuals expected");
        int g = XP_REG_TRUE;
        const kern_expression *a = first_arg(kj);
        for (uint32_t i = 0; i < kj->nr_args; i++, a = next_arg(a))
        {
            if (!check_node(kj, a))
                return false;
            bool pushed = (a->expflags & KEXP_FLAG__IS_PUSHED_DOWN) != 0;
            if (other_only && !pushed)
                continue;
            int gq = g;
            if (pushed && !other_only)
            {
                /* pushed-down quals are skipped once the status is -1 */
                gq = alloc_reg();
                emit(XO_G_POSITIVE, gq, r_status, g);
            }
            Reg v = expr(kj, a, gq);
            if (v.dom != XD_INT)
                return fail("join qual is not boolean");
            emit(XO_JQ_STATUS, r_status, v.r, gq, XP_REG_TRUE, 0,
                 (pushed && !other_only) ? -1 : 0);
            if (!pushed || other_only)
            {
                /* an unmatched plain join qual ends the evaluation */
                int ng = alloc_reg();
                emit(XO_G_NONZERO, ng, r_status, g);
                g = ng;
            }
        }
        return true;
    }
};

int
layout_state_size(int action)
{
    switch (action)
    {
        case KAGG_ACTION__NROWS_ANY:
        case KAGG_ACTION__NROWS_COND:   return 8;
        case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMIN_INT64:
        case KAGG_ACTION__PMAX_INT32: case KAGG_ACTION__PMAX_INT64:
            return sizeof(kagg_state__pminmax_int64_packed);
        case KAGG_ACTION__PMIN_FP64: case KAGG_ACTION__PMAX_FP64:
            return sizeof(kagg_state__pminmax_fp64_packed);
        case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
            return sizeof(kagg_state__psum_int_packed);
        case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
        case KAGG_ACTION__PSUM_NUMERIC: case KAGG_ACTION__PAVG_NUMERIC:
            return sizeof(kagg_state__psum_numeric_packed);
        case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
            return sizeof(kagg_state__psum_fp_packed);
        case KAGG_ACTION__STDDEV:
            return sizeof(kagg_state__stddev_packed);
    }
    return -1;
}

/* fields an action keeps in the LDS partial table, besides the entry's row
 * count (rows that reach the table have every argument non-NULL and every
 * FILTER true, so count(*), count(x) and the n of sums and averages are all
 * the row count); false if the action has no LDS form */
bool
lds_fields(int action, int *n64, int *n32)
{
    *n64 = *n32 = 0;
    switch (action)
    {
        case KAGG_ACTION__NROWS_ANY: case KAGG_ACTION__NROWS_COND:
            return true;
        case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMAX_INT32:
            *n32 = 1; return true;
        case KAGG_ACTION__PMIN_INT64: case KAGG_ACTION__PMAX_INT64:
        case KAGG_ACTION__PMIN_FP64: case KAGG_ACTION__PMAX_FP64:
        case KAGG_ACTION__PSUM_INT: case KAGG_ACTION__PAVG_INT:
        case KAGG_ACTION__PSUM_FP: case KAGG_ACTION__PAVG_FP:
            *n64 = 1; return true;
        case KAGG_ACTION__PSUM_INT64: case KAGG_ACTION__PAVG_INT64:
            *n64 = 1; *n32 = 1; return true;    /* low-half sum + count of negatives */
        case KAGG_ACTION__STDDEV:
            *n64 = 2; return true;
    }
    return false;
}

}   /* namespace */

void
xpu_set_error(kern_errorbuf *eb, uint32_t errcode, const char *filename, int lineno,
              const char *funcname, const char *message)
{
    if (!eb)
        return;
    memset(eb, 0, sizeof(*eb));
    eb->errcode = errcode;
    eb->lineno = lineno;
    snprintf(eb->filename, sizeof(eb->filename), "%s", filename);
    snprintf(eb->funcname, sizeof(eb->funcname), "%s", funcname);
    snprintf(eb->message, sizeof(eb->message), "%s", message);
}

int
xpu_compile_session(const kern_session_info *s, size_t length, xpu_compiled *out,
                    kern_errorbuf *eb)
{
    Compiler C;
    xplan *P = &out->plan;

    memset(P, 0, sizeof(xplan));
    for (int i = 0; i < XP_MAX_PARAMS; i++)
        C.param_index[i] = -1;
    C.s = s;
    C.length = length;
    C.out = out;
    C.P = P;
    C.enc = encoding_kind(s, length);
#define FAIL(msg) do { xpu_set_error(eb, ERRCODE_DEVICE_ERROR, "xpu_compile.cpp", __LINE__, \
                                     __func__, (msg)); return ERRCODE_DEVICE_ERROR; } while (0)
    if (length < offsetof(kern_session_info, poffset) ||
        length < offsetof(kern_session_info, poffset) + 4ULL * s->nparams)
        FAIL("session blob is too short");
    if (s->xpucode_gist_evals_packed != 0)
        FAIL("GiST joins are not supported by this build");
    if (s->select_into_kds_head != 0)
        FAIL("SELECT INTO direct is not supported by this build");
    P->task_kind = s->xpu_task_flags & DEVTASK__MASK;

    /* kvars slots -> registers */
    C.nslots = s->kcxt_kvars_nslots;
    if (C.nslots > 256)
        FAIL("too many kvars slots");
    out->nslots = C.nslots;
    if (C.nslots > 0)
    {
        if (!s->kcxt_kvars_defs ||
            s->kcxt_kvars_defs + sizeof(kern_varslot_desc) * (uint64_t)C.nslots > length)
            FAIL("kvars slot definitions out of the session blob");
        const kern_varslot_desc *vd = (const kern_varslot_desc *)((const char *)s + s->kcxt_kvars_defs);
        for (int i = 0; i < C.nslots; i++)
        {
            /* registers are given on first use; a slot of a type this build
             * does not execute is an error only if something uses it */
            C.slot_type[i] = vd[i].vs_type_code;
            C.slot_reg[i] = -1;
        }
    }

    /* structure of the pipeline */
    const kern_expression *kload = C.kexp_at(s->xpucode_load_vars_packed);
    const kern_expression *kjoin = C.kexp_at(s->xpucode_join_quals_packed);
    const kern_expression *khash = C.kexp_at(s->xpucode_hash_values_packed);
    if (kload && !C.check_node(nullptr, kload))
        FAIL(C.err.c_str());
    if (kjoin && !C.check_node(nullptr, kjoin))
        FAIL(C.err.c_str());
    if (khash && !C.check_node(nullptr, khash))
        FAIL(C.err.c_str());
    int nrels = 0;
    if (kjoin)
        nrels = (int)kjoin->u.pack.npacked - 1;
    else if (khash)
        nrels = (int)khash->u.pack.npacked - 1;
    if (nrels < 0)
        nrels = 0;
    if (nrels > XP_MAX_DEPTH)
        FAIL("too many inner relations");
    P->nrels = nrels;
    /* results are read by the kernel right after their segment, so they all
     * share R1; the inner tuple pointer is R2 (a depth's inner values are in
     * their slot registers once seg_join has run, and the probe hash is the
     * hash stored in the matched item) */
    P->r_scan_ok = XP_REG_RESULT;
    P->r_hash = XP_REG_RESULT;
    P->r_status = XP_REG_RESULT;
    P->r_keyhash = XP_REG_RESULT;
    P->r_tuple = XP_REG_TUPLE;

    /* depth 0: loads + scan quals */
    P->seg_scan = C.begin_seg();
    if (!C.loadvars(C.packed_sub(kload, 0), 0, 0, false))
        FAIL(C.err.c_str());
    {
        const kern_expression *kq = C.kexp_at(s->xpucode_scan_quals);
        if (kq)
        {
            Reg v = C.expr(nullptr, kq, XP_REG_TRUE);
            if (C.err.empty() && v.dom != XD_INT)
                C.fail("scan quals are not boolean");
            if (C.err.empty())
                C.emit(XO_MOV, P->r_scan_ok, v.r);
        }
        else
            C.emit(XO_CONST, P->r_scan_ok, 0, 0, XP_REG_TRUE, 0, 1);
        if (nrels == 0)
        {
            /* the final stage runs in the same pass over the row */
            C.emit(XO_KILL, 0, P->r_scan_ok);
            P->fused = 1;
        }
    }
    C.end_seg(P->seg_scan);
    if (!C.err.empty())
        FAIL(C.err.c_str());

    /* CPU fallback (executor.c:705-724, xpu_common.cu:1898-1955): a row
     * that falls back at depth 0 becomes a tuple of the scan slot whose
     * attributes from depth-0 variables hold their values; the rest are
     * NULL */
    if (s->fallback_kds_head)
    {
        if ((uint64_t)s->fallback_kds_head + sizeof(kern_data_store) > length)
            FAIL("fallback buffer header out of the session blob");
        const kern_data_store *h = (const kern_data_store *)((const char *)s + s->fallback_kds_head);
        const size_t hl = KDS_HEAD_LENGTH(h);
        if ((uint64_t)s->fallback_kds_head + hl > length || h->nr_colmeta < h->ncols)
            FAIL("fallback buffer header out of the session blob");
        if (h->format != KDS_FORMAT_FALLBACK)
            FAIL("fallback buffer header is not KDS_FORMAT_FALLBACK");
        if (h->ncols > XP_MAX_PROJ)
            FAIL("too many attributes in the CPU-fallback tuple");
        if (h->length < hl + 4096)
            FAIL("fallback buffer is too small");
        out->kds_fallback_head.assign((const char *)h, (const char *)h + hl);
        P->fb_nattrs = h->ncols;
        for (uint32_t j = 0; j < h->ncols; j++)
        {
            xattr *fa = &P->fb[j];
            fa->reg = 0xff;
            fa->domain = 0;
            fa->attalign = h->colmeta[j].attalign;
            fa->width = h->colmeta[j].attlen > 0 ? (int8_t)h->colmeta[j].attlen : (int8_t)-1;
        }
        const uint32_t nfb = s->fallback_desc_nitems;
        if (nfb > 0 && (!s->fallback_desc_defs ||
                        (uint64_t)s->fallback_desc_defs + (uint64_t)nfb * sizeof(kern_fallback_desc) > length))
            FAIL("fallback descriptors out of the session blob");
        const kern_fallback_desc *fd = (const kern_fallback_desc *)((const char *)s + s->fallback_desc_defs);
        for (uint32_t i = 0; i < nfb; i++)
        {
            if (fd[i].fb_dst_resno < 1 || (uint32_t)fd[i].fb_dst_resno > h->ncols)
                FAIL("fallback descriptor out of the scan slot");
            if (fd[i].fb_src_depth != 0)
                continue;
            const int slot = fd[i].fb_slot_id;
            if (slot >= (int)C.nslots)
                FAIL("fallback descriptor slot out of range");
            if (C.slot_reg[slot] < 0)
                continue;           /* not loaded at depth 0: NULL */
            Reg r = C.reg_of_slot(slot);
            xattr *fa = &P->fb[fd[i].fb_dst_resno - 1];
            const int attlen = h->colmeta[fd[i].fb_dst_resno - 1].attlen;
            const bool varlena = (r.dom == XD_TEXT || r.dom == XD_NUM);
            if (varlena != (attlen == -1) || (!varlena && attlen != type_width(r.type)))
                FAIL("CPU-fallback attribute does not match its variable's type");
            fa->reg = (uint8_t)r.r;
            fa->domain = (uint8_t)r.dom;
        }
    }

    /* joins */
    for (int d = 1; d <= nrels; d++)
    {
        const kern_expression *kh = C.packed_sub(khash, d);
        const kern_expression *kj = C.packed_sub(kjoin, d);
        if (!kh)
            FAIL("only hash joins are supported (no hash keys for a depth)");
        P->depth[d].seg_hash = C.begin_seg();
        {
            Reg h = C.expr(nullptr, kh, XP_REG_TRUE);
            if (C.err.empty())
                C.emit(XO_MOV, P->r_hash, h.r);
        }
        C.end_seg(P->depth[d].seg_hash);
        P->depth[d].seg_join = C.begin_seg();
        C.loadvars(C.packed_sub(kload, d), d, P->r_tuple, false);
        C.joinquals(kj, P->r_status, false);
        C.end_seg(P->depth[d].seg_join);
        P->depth[d].seg_oj = C.begin_seg();
        C.loadvars(C.packed_sub(kload, d), d, P->r_tuple, true);
        C.joinquals(kj, P->r_status, true);
        C.end_seg(P->depth[d].seg_oj);
        if (!C.err.empty())
            FAIL(C.err.c_str());
    }

    /* final stage */
    const kern_expression *kproj = C.kexp_at(s->xpucode_projection);
    const kern_expression *kactions = C.kexp_at(s->xpucode_groupby_actions);
    if (s->projection_kds_dst)
    {
        if (s->projection_kds_dst + sizeof(kern_data_store) > length)
            FAIL("projection_kds_dst out of the session blob");
        const kern_data_store *h = (const kern_data_store *)((const char *)s + s->projection_kds_dst);
        size_t hl = KDS_HEAD_LENGTH(h);
        if (s->projection_kds_dst + hl > length)
            FAIL("projection_kds_dst out of the session blob");
        out->kds_dst_head.assign((const char *)h, (const char *)h + hl);
    }
    if (!kactions && (s->xpu_task_flags & DEVTASK__PINNED_HASH_RESULTS))
    {
        /* results kept on the device as an inner buffer of a later join
         * (executor.c:655-672): the buffer's KDS_FORMAT_HASH header is the
         * session's groupby_kds_final */
        if (!s->groupby_kds_final || s->groupby_kds_final + sizeof(kern_data_store) > length)
            FAIL("pinned hash results without a buffer header");
        const kern_data_store *kf = (const kern_data_store *)((const char *)s + s->groupby_kds_final);
        size_t hl = KDS_HEAD_LENGTH(kf);
        if (s->groupby_kds_final + hl > length || kf->format != KDS_FORMAT_HASH ||
            kf->hash_nslots == 0)
            FAIL("pinned hash results need a KDS_FORMAT_HASH buffer header");
        out->kds_final_head.assign((const char *)kf, (const char *)kf + hl);
        out->pinned_hash = true;
    }
    if (kactions)
    {
        /* ---- group-by ---- */
        if (!C.check_node(nullptr, kactions) || kactions->opcode != FuncOpCode__AggFuncs)
            FAIL("malformed AggFuncs");
        if (!s->groupby_kds_final || s->groupby_kds_final + sizeof(kern_data_store) > length)
            FAIL("group-by without kds_final header");
        const kern_data_store *kf = (const kern_data_store *)((const char *)s + s->groupby_kds_final);
        size_t hl = KDS_HEAD_LENGTH(kf);
        if (s->groupby_kds_final + hl > length)
            FAIL("kds_final header out of the session blob");
        out->kds_final_head.assign((const char *)kf, (const char *)kf + hl);
        out->has_groupby = true;
        int nattrs = kactions->u.pagg.nattrs;
        if (nattrs > XP_MAX_PROJ || nattrs > (int)kf->ncols)
            FAIL("too many group-by attributes");
        P->gb_nattrs = nattrs;

        P->seg_final = C.begin_seg();
        /* The key hash expression (HashValue over the grouping keys) only
         * places groups in kds_final's hash slots, which the write-back turns
         * into a ROW buffer; the kernels hash the key registers themselves,
         * and only for rows that reach kds_final, so it is checked here but
         * not compiled. */
        const kern_expression *kkh = C.kexp_at(s->xpucode_groupby_keyhash);
        if (kkh && (!C.check_node(nullptr, kkh) || kkh->opcode != FuncOpCode__HashValue))
            FAIL("malformed group-by key hash");
        /* aggregate arguments (SaveExpr children of AggFuncs) */
        {
            const kern_expression *a = C.first_arg(kactions);
            for (uint32_t i = 0; i < kactions->nr_args; i++, a = C.next_arg(a))
                C.expr(kactions, a, XP_REG_TRUE);
        }
        /* a text grouping key is hashed and compared by its bytes: one held
         * compressed or external falls back to the CPU before that */
        for (int j = 0; j < nattrs; j++)
        {
            const kern_aggregate_desc *d = &kactions->u.pagg.desc[j];
            if (d->action == KAGG_ACTION__VREF && d->arg0_slot_id >= 0 &&
                d->arg0_slot_id < C.nslots)
            {
                Reg r = C.reg_of_slot(d->arg0_slot_id);
                if (r.dom == XD_TEXT)
                    C.emit(XO_TXTCHK, r.r, r.r, 0, XP_REG_TRUE);
            }
            /* a numeric sum's argument is parsed where it is loaded, and
             * raises its error there (xpu_numeric_validate) */
            if ((d->action == KAGG_ACTION__PSUM_NUMERIC || d->action == KAGG_ACTION__PAVG_NUMERIC) &&
                d->arg0_slot_id >= 0 && d->arg0_slot_id < C.nslots)
            {
                Reg r = C.reg_of_slot(d->arg0_slot_id);
                if (r.dom == XD_NUM)
                    C.emit(XO_NUM_PARSE, r.r, r.r, 0, XP_REG_TRUE);
            }
        }
        C.end_seg(P->seg_final);
        if (!C.err.empty())
            FAIL(C.err.c_str());

        /* state layout (__writeOutOneTuplePreAgg, cuda_gpupreagg.cu:18-223) */
        uint32_t off = 0;
        bool seen_key = false;
        int nkeys = 0, key_bits = 0;
        bool lds_ok = true;
        int lds_n64 = 0, lds_n32 = 0;
        for (int j = 0; j < nattrs; j++)
        {
            const kern_aggregate_desc *d = &kactions->u.pagg.desc[j];
            const kern_colmeta *cm = &kf->colmeta[j];
            xagg *ag = &P->aggs[j];
            ag->action = d->action;
            ag->arg_reg = 0xff;
            ag->filter_reg = 0xff;
            ag->key_index = 0xff;
            ag->lds64 = -1;
            ag->lds32 = -1;
            if (d->arg0_slot_id >= 0)
            {
                if (d->arg0_slot_id >= C.nslots)
                    FAIL("aggregate argument slot out of range");
                Reg r = C.reg_of_slot(d->arg0_slot_id);
                if (r.dom < 0)
                    FAIL("aggregate argument of unsupported type");
                ag->arg_reg = r.r;
                ag->arg_domain = r.dom;
            }
            if (d->filter_slot_id >= 0)
            {
                if (d->filter_slot_id >= C.nslots)
                    FAIL("aggregate filter slot out of range");
                ag->filter_reg = C.slot_reg[d->filter_slot_id];
            }
            if (d->action == KAGG_ACTION__VREF)
            {
                if (nkeys >= XP_MAX_KEYS)
                    FAIL("too many grouping keys");
                if (d->arg0_slot_id < 0)
                    FAIL("grouping key without a slot");
                int t = C.slot_type[d->arg0_slot_id];
                xattr *ka = &P->keys[nkeys];
                ka->reg = ag->arg_reg;
                ka->domain = ag->arg_domain;
                ka->attalign = cm->attalign;
                ka->width = (ka->domain == XD_TEXT || ka->domain == XD_NUM) ? -1 : type_width(t);
                if (ka->width > 0 && cm->attlen != ka->width)
                    FAIL("grouping key width differs from the kds_final column");
                if (ka->width < 0)
                    lds_ok = false;
                else
                {
                    P->key_bits[nkeys] = 8 * ka->width;
                    key_bits += 8 * ka->width + 1;
                }
                ag->key_index = nkeys++;
                seen_key = true;
                continue;
            }
            if (d->action == KAGG_ACTION__VREF_NOKEY)
                FAIL("VREF_NOKEY is not supported");
            int sz = layout_state_size(d->action);
            if (sz < 0)
                FAIL("aggregate action not supported by this build");
            if (seen_key)
                FAIL("partial aggregates must precede the grouping keys in kds_final");
            off = XPU_TYPEALIGN(cm->attalign, off);
            if (off % 8 != 0)
                FAIL("misaligned partial aggregate state");
            ag->state_off = off;
            ag->state_size = sz;
            off += sz;
            if (d->action == KAGG_ACTION__PSUM_NUMERIC || d->action == KAGG_ACTION__PAVG_NUMERIC)
            {
                if (ag->arg_reg == 0xff || ag->arg_domain != XD_NUM)
                    FAIL("numeric sum of a non-numeric argument");
                ag->num_scale = (uint16_t)numeric_typmod_scale(d->typmod);
            }
            int f64, f32;
            if (!lds_fields(d->action, &f64, &f32))
                lds_ok = false;
            if (f64 > 0)
                ag->lds64 = lds_n64;
            if (f32 > 0)
                ag->lds32 = lds_n32;
            lds_n64 += f64;
            lds_n32 += f32;
        }
        P->nkeys = nkeys;
        P->gb_states_end = off;
        P->gb_t_hoff = XPU_MAXALIGN(KERN_TUPITEM_T_BITS_OFFSET + XPU_BITMAPLEN(nattrs));
        P->lds_fast = (lds_ok && key_bits <= 63) ? 1 : 0;
        /* registers a spill record carries: the distinct aggregate arguments */
        if (P->lds_fast && nkeys > 0)
        {
            int n = 0;
            bool fits = true;
            for (int j = 0; j < nattrs; j++)
            {
                const xagg *ag = &P->aggs[j];
                if (ag->action == KAGG_ACTION__VREF || ag->arg_reg == 0xff)
                    continue;
                bool seen = false;
                for (int k = 0; k < n; k++)
                    seen = seen || P->spill_regs[k] == ag->arg_reg;
                if (seen)
                    continue;
                if (n >= XP_MAX_SPILL)
                    fits = false;
                else
                    P->spill_regs[n++] = ag->arg_reg;
            }
            P->spill_ok = fits ? 1 : 0;
            P->spill_nregs = fits ? (uint8_t)n : 0;
        }
        P->lds_n64 = lds_n64;
        P->lds_n32 = lds_n32;
        P->lds_entry_bytes = 8 * (1 + lds_n64) + 4 * (1 + lds_n32);
        out->lds_entry_bytes = P->lds_entry_bytes;
        if (kf->format == KDS_FORMAT_HASH && kf->hash_nslots == 0)
            FAIL("kds_final in HASH format without hash slots");
        if (nkeys > 0 && kf->format != KDS_FORMAT_HASH)
            FAIL("GROUP BY needs a KDS_FORMAT_HASH kds_final");
    }
    else if (kproj)
    {
        /* ---- projection ---- */
        if (!C.check_node(nullptr, kproj) || kproj->opcode != FuncOpCode__Projection)
            FAIL("malformed Projection");
        if (out->kds_dst_head.empty())
            FAIL("projection without kds_dst header");
        const kern_data_store *h = (const kern_data_store *)out->kds_dst_head.data();
        out->has_projection = true;
        P->seg_final = C.begin_seg();
        const kern_expression *a = C.first_arg(kproj);
        for (uint32_t i = 0; i < kproj->nr_args; i++, a = C.next_arg(a))
            C.expr(kproj, a, XP_REG_TRUE);
        /* pinned inner buffer: the projection's HashValue over the join keys
         * (codegen.c:3933-3941) becomes each row's kern_tupitem.hash
         * (cuda_gpujoin.cu:799-829) */
        if (s->xpu_task_flags & DEVTASK__PINNED_HASH_RESULTS)
        {
            if (kproj->u.proj.hash == 0)
                FAIL("pinned hash results without a projection hash");
            const kern_expression *kh = (const kern_expression *)((const char *)kproj +
                                                                  kproj->u.proj.hash);
            if (!C.check_node(kproj, kh) || kh->opcode != FuncOpCode__HashValue)
                FAIL("malformed projection hash");
            Reg h = C.expr(kproj, kh, XP_REG_TRUE);
            C.emit(XO_MOV, XP_REG_RESULT, h.r);
            P->r_projhash = XP_REG_RESULT;
        }
        C.end_seg(P->seg_final);
        if (!C.err.empty())
            FAIL(C.err.c_str());
        int nattrs = kproj->u.proj.nattrs;
        if (nattrs > (int)h->ncols)
            nattrs = h->ncols;
        if (nattrs > XP_MAX_PROJ)
            FAIL("too many projected attributes");
        P->proj_nattrs = nattrs;
        for (int j = 0; j < nattrs; j++)
        {
            int slot = kproj->u.proj.slot_id[j];
            const kern_colmeta *cm = &h->colmeta[j];
            xattr *pa = &P->proj[j];
            if (slot >= C.nslots)
            {
                pa->reg = 0xff;     /* considered NULL (xpu_common.cu:1117) */
                continue;
            }
            Reg r = C.reg_of_slot(slot);
            if (r.dom < 0)
                FAIL("projection of an unsupported type");
            pa->reg = r.r;
            pa->domain = r.dom;
            pa->attalign = cm->attalign;
            pa->width = (r.dom == XD_TEXT || r.dom == XD_NUM) ? -1 : type_width(r.type);
            if (pa->width > 0 && cm->attlen != pa->width)
                FAIL("projected value width differs from the kds_dst column");
        }
    }
    else
        FAIL("session has neither a projection nor a group-by");

    if (C.ops.size() > XP_MAX_OPS)
        FAIL("compiled program is too long");
    if (!C.err.empty())
        FAIL(C.err.c_str());
    /* pick the register file: temporaries were numbered down from
     * XP_NREGS; move them right above the slots' registers */
    {
        int ntemps = XP_NREGS - C.temp_min;
        int nr = (C.slot_top + ntemps <= XP_NREGS_SMALL) ? XP_NREGS_SMALL : XP_NREGS;
        if (C.slot_top + ntemps > XP_NREGS)
            FAIL("expression needs more than the per-lane register file");
        int shift = XP_NREGS - nr;
        for (xop &o : C.ops)
        {
            if (o.dst >= C.temp_min) o.dst -= shift;
            if (o.a >= C.temp_min) o.a -= shift;
            if (o.b >= C.temp_min && o.op != XO_HASH_ADD && o.op != XO_LDCOL &&
                o.op != XO_LDTUP && o.op != XO_COALESCE)
                o.b -= shift;
            if (o.guard >= C.temp_min) o.guard -= shift;
        }
        P->nregs = nr;
    }
    /* depth-0 fixed-width columns: the scan kernel loads them for a whole
     * tile of rows before running the program (all loads in flight at once)
     * and the program's own loads of them become no-ops there */
    bool prefix = true;
    for (int pc = P->seg_scan.begin; pc < P->seg_scan.end; pc++)
    {
        xop &o = C.ops[pc];
        if (o.op != XO_LDCOL || P->nstage >= XP_MAX_STAGE)
        {
            prefix = false;
            continue;
        }
        if (prefix)
            P->stage_prefix++;
        xstage *st = &P->stage[P->nstage++];
        st->col = o.aux;
        st->width = (uint8_t)(o.imm & XLD_WIDTH_MASK);
        st->f32 = (o.b == XD_F32) ? 1 : 0;
        st->reg = o.dst;
        o.imm |= XLD_STAGED;
    }
    P->nops = C.ops.size();
    memcpy(P->ops, C.ops.data(), sizeof(xop) * C.ops.size());

    /* source columns the programs read (heap / row chunks are deformed into
     * exactly these) */
    for (const xop &o : C.ops)
    {
        uint8_t kind, width = 0;
        if (o.op == XO_LDCOL)
        {
            kind = XH_FIXED;
            width = (uint8_t)(o.imm & XLD_WIDTH_MASK);
        }
        else if (o.op == XO_LDCOL_BOOL)
        {
            kind = XH_BOOL;
            width = 1;
        }
        else if (o.op == XO_LDCOL_VAR || o.op == XO_LDCOL_DEC)
            kind = XH_VAR;      /* numeric on heap pages: a varlena */
        else
            continue;
        if (o.aux >= out->src_cols.size())
            out->src_cols.resize(o.aux + 1, xpu_src_col{0, 0});
        xpu_src_col &sc = out->src_cols[o.aux];
        if (sc.kind && (sc.kind != kind || sc.width != width))
            FAIL("a source column is loaded with two types");
        sc.kind = kind;
        sc.width = width;
    }
    /* GPU-Sort of the group-by results (codegen.c:4360-4525): each key is
     * a column of the final tuple or a partial aggregate state; finalized
     * values (averages) are computed while the keys are encoded, so the
     * tuple margin (gpusort_htup_margin) is not used */
    if (s->xpucode_gpusort_keydesc)
    {
        const kern_expression *ks = C.kexp_at(s->xpucode_gpusort_keydesc);
        if (!ks || !C.check_node(nullptr, ks) || ks->opcode != FuncOpCode__SortKeys)
            FAIL("malformed SortKeys");
        if (!out->has_groupby)
            FAIL("GPU-Sort of scan or join results is not supported by this build");
        const char wf = ks->u.sort.window_rank_func;
        if (wf != 0 && wf != KSORT_WINDOW_FUNC__ROW_NUMBER && wf != KSORT_WINDOW_FUNC__RANK &&
            wf != KSORT_WINDOW_FUNC__DENSE_RANK)
            FAIL("unknown GPU-Sort window rank function");
        /* codegen.c:4505-4512: a window rank and a LIMIT never come together */
        if (wf != 0 && s->gpusort_limit_count > 0)
            FAIL("GPU-Sort window rank with a LIMIT");
        const int nk = ks->u.sort.nkeys;
        if (wf != 0 && (ks->u.sort.window_partby_nkeys < 1 || ks->u.sort.window_orderby_nkeys < 1 ||
                        ks->u.sort.window_partby_nkeys + ks->u.sort.window_orderby_nkeys > nk))
            FAIL("GPU-Sort window rank key counts out of range");
        if (nk < 1 || nk > XSORT_MAX_KEYS)
            FAIL("GPU-Sort key count out of range");
        if (offsetof(kern_expression, u.sort.desc) + (size_t)nk * sizeof(kern_sortkey_desc) > ks->len)
            FAIL("SortKeys descriptors out of the expression");
        const kern_data_store *kf = (const kern_data_store *)out->kds_final_head.data();
        uint32_t nw = 0;
        for (int k = 0; k < nk; k++)
        {
            const kern_sortkey_desc &d = ks->u.sort.desc[k];
            if (d.src_anum < 1 || d.src_anum > kf->ncols || d.src_anum > kf->nr_colmeta)
                FAIL("GPU-Sort key attribute out of range");
            const kern_colmeta &cm = kf->colmeta[d.src_anum - 1];
            xsort_key sk;
            memset(&sk, 0, sizeof(sk));
            sk.kind = d.kind;
            sk.nulls_first = d.nulls_first ? 1 : 0;
            sk.order_asc = d.order_asc ? 1 : 0;
            sk.anum = d.src_anum;
            sk.word = (uint8_t)nw;
            sk.type_code = d.key_type_code;
            switch (d.kind)
            {
                case KSORT_KEY_KIND__VREF:
                {
                    int width = 0;
                    switch (d.key_type_code)
                    {
                        case TypeOpCode__bool: case TypeOpCode__int1: width = 1; break;
                        case TypeOpCode__int2: case TypeOpCode__float2: width = 2; break;
                        case TypeOpCode__int4: case TypeOpCode__float4:
                        case TypeOpCode__date: width = 4; break;
                        case TypeOpCode__int8: case TypeOpCode__float8: width = 8; break;
                        default:
                            FAIL("GPU-Sort key type is not supported by this build");
                    }
                    if (cm.attlen != width)
                        FAIL("GPU-Sort key type does not match the final buffer column");
                    sk.nwords = width == 8 ? 2 : 1;
                    break;
                }
                case KSORT_KEY_KIND__PMINMAX_INT64:
                case KSORT_KEY_KIND__PMINMAX_FP64:
                case KSORT_KEY_KIND__PSUM_INT64:
                case KSORT_KEY_KIND__PSUM_FP64:
                case KSORT_KEY_KIND__PAVG_INT64:
                case KSORT_KEY_KIND__PAVG_FP64:
                case KSORT_KEY_KIND__PSUM_NUMERIC:
                    if (cm.attlen != -1)
                        FAIL("GPU-Sort aggregate key is not a partial state column");
                    /* the numeric key's words are the raw int128 sum, which
                     * orders states only at one scale: those of sum(int8)
                     * (PSUM_INT64, weight 0); the reference rescales by the
                     * weight first (cuda_gpusort.cu:205-222) */
                    if (d.kind == KSORT_KEY_KIND__PSUM_NUMERIC &&
                        (d.src_anum > (int)P->gb_nattrs ||
                         P->aggs[d.src_anum - 1].action != KAGG_ACTION__PSUM_INT64))
                        FAIL("GPU-Sort numeric key must be a sum(int8) state in this build");
                    sk.nwords = d.kind == KSORT_KEY_KIND__PSUM_NUMERIC ? 3 : 2;
                    break;
                default:
                    FAIL("GPU-Sort key kind is not supported by this build");
            }
            nw += sk.nwords;
            out->sort_keys.push_back(sk);
            if (wf != 0 && k + 1 == ks->u.sort.window_partby_nkeys)
                out->sort_wr_pwords = nw;
            if (wf != 0 && k + 1 == ks->u.sort.window_partby_nkeys + ks->u.sort.window_orderby_nkeys)
                out->sort_wr_owords = nw;
        }
        out->sort_nwords = nw;
        out->sort_wr_func = wf;
        out->sort_wr_limit = ks->u.sort.window_rank_limit;
        out->sort_limit = s->gpusort_limit_count > 0 ? s->gpusort_limit_count : 0;
    }
    /* parallelCurrentXids, for the visibility of KDS_FORMAT_ROW sources */
    if (s->session_xact_state)
    {
        const size_t xoff = s->session_xact_state;
        const size_t xids_at = offsetof(SerializedTransactionState, parallelCurrentXids);
        if (xoff + xids_at > length)
            FAIL("session_xact_state out of the session blob");
        const SerializedTransactionState *xs =
            (const SerializedTransactionState *)((const char *)s + xoff);
        int32_t n = xs->nParallelCurrentXids;
        if (n < 0 || xoff + xids_at + 4ULL * (uint32_t)n > length)
            FAIL("session_xact_state out of the session blob");
        out->has_xact_state = true;
        out->xids.assign(xs->parallelCurrentXids, xs->parallelCurrentXids + n);
    }
    return 0;
#undef FAIL
}

/* textual dump of a compiled plan (debugging and CPU-side tests) */
std::string
xpu_plan_dump(const xpu_compiled *c)
{
    static const char *names[XO_NUM_OPS] = {
        "NOP", "LDCOL", "LDCOL_BOOL", "LDCOL_VAR", "LDTUP", "CONST", "CONST_TEXT", "MOV", "MOV2",
        "SETNULL", "CMP_I", "CMP_F32", "CMP_F64", "CMP_TEXT", "ADD_I", "SUB_I", "MUL_I", "DIV_I",
        "MOD_I", "NEG_I", "ABS_I", "AND_I", "OR_I", "XOR_I", "NOT_I", "SHL_I", "SHR_I",
        "ADD_F32", "SUB_F32", "MUL_F32", "DIV_F32", "ADD_F64", "SUB_F64", "MUL_F64", "DIV_F64",
        "NEG_F32", "ABS_F32", "NEG_F64", "ABS_F64", "I2I", "F32_I", "F64_I", "I_F32", "I_F64",
        "F32_F64", "F64_F32", "I_BOOL", "TEXTLEN", "AND3", "OR3", "NOT3", "ISNULL", "ISNOTNULL",
        "BOOLTEST", "COALESCE", "G_NOTFALSE", "G_NOTTRUE", "G_ISNULL", "G_POSITIVE", "G_NONZERO",
        "JQ_STATUS", "HASH_INIT", "HASH_ADD", "HASH_FIN", "KILL", "PARAM", "G_TRUE", "SELECT",
        "DISTINCT", "MINMAX", "LIKE", "SUBSTR", "HAOP", "TXTCHK"};
    const xplan *P = &c->plan;
    std::string s;
    char buf[256];
    snprintf(buf, sizeof(buf), "nregs=%u nrels=%u nops=%u\n", P->nregs, P->nrels, P->nops);
    s += buf;
    auto seg = [&](const char *name, xseg g) {
        snprintf(buf, sizeof(buf), "[%s %u..%u]\n", name, g.begin, g.end);
        s += buf;
        for (int pc = g.begin; pc < g.end; pc++)
        {
            const xop &o = P->ops[pc];
            snprintf(buf, sizeof(buf), "  %3d %-10s d=%u a=%u b=%u g=%u aux=%u imm=%lld\n", pc,
                     o.op < XO_NUM_OPS ? names[o.op] : "?", o.dst, o.a, o.b, o.guard, o.aux,
                     (long long)o.imm);
            s += buf;
        }
    };
    seg("scan", P->seg_scan);
    for (uint32_t d = 1; d <= P->nrels; d++)
    {
        seg("hash", P->depth[d].seg_hash);
        seg("join", P->depth[d].seg_join);
        seg("outer", P->depth[d].seg_oj);
    }
    seg("final", P->seg_final);
    for (uint32_t i = 0; i < P->nstage; i++)
    {
        snprintf(buf, sizeof(buf), "stage[%u] col=%u width=%u f32=%u reg=%u\n", i, P->stage[i].col,
                 P->stage[i].width, P->stage[i].f32, P->stage[i].reg);
        s += buf;
    }
    if (c->has_groupby)
    {
        snprintf(buf, sizeof(buf), "groupby nkeys=%u lds_fast=%u lds_entry_bytes=%u (u64 fields %u, u32 fields %u)"
                 " spill=%u (%u regs)\n", P->nkeys, P->lds_fast, P->lds_entry_bytes, P->lds_n64,
                 P->lds_n32, P->spill_ok, P->spill_nregs);
        s += buf;
    }
    for (uint32_t j = 0; j < P->proj_nattrs; j++)
    {
        snprintf(buf, sizeof(buf), "proj[%u] reg=%u dom=%u align=%d width=%d\n", j, P->proj[j].reg,
                 P->proj[j].domain, P->proj[j].attalign, P->proj[j].width);
        s += buf;
    }
    return s;
}
