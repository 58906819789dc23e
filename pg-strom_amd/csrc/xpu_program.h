/*
 * xpu_program.h - the compiled form of a PG-Strom session for the MI355X
 * pipeline kernel (shared by the host compiler and the HIP kernels).
 *
 * At session open the kern_expression trees of the session blob (scan quals,
 * per-depth hash values and join quals, projection, group-by key hash and
 * aggregate arguments) are translated into straight-line "xop" programs over
 * a small file of 64-bit per-lane registers with a per-lane NULL mask.  The
 * kernel walks an op list with a wave-uniform program counter, so dispatch is
 * scalar and the register file stays in VGPRs (dynamic indices are uniform);
 * the reference instead calls device function pointers recursively
 * (EXEC_KERN_EXPRESSION, xpu_common.h:2694) with per-thread alloca stacks.
 *
 * Short-circuit semantics of AND/OR/CASE (an argument that is never evaluated
 * by the reference cannot raise an error) are kept with guard registers:
 * every op that may raise an error names a guard register, and only lanes
 * whose guard is true report the error.
 */
#ifndef XPU_PROGRAM_H
#define XPU_PROGRAM_H
#include <stdint.h>

#define XP_NREGS            32      /* largest per-lane register file */
#define XP_NREGS_SMALL      16      /* register file kept in VGPRs */
#define XP_REG_RESULT       1       /* scan result / hash / join status / key hash */
#define XP_REG_TUPLE        2       /* inner tuple being joined */
#define XP_REG_TRUE         0       /* R0 is the constant TRUE (guard "none") */
#define XP_MAX_DEPTH        8       /* inner relations */
#define XP_MAX_PROJ         64      /* projected / group-by attributes */
#define XP_MAX_KEYS         8
#define XP_MAX_OPS          2048
#define XP_MAX_STAGE        6       /* depth-0 columns loaded by the tile prologue */
#define XP_MAX_SPILL        8       /* registers carried by a group-by spill record */
#define XP_MAX_PARAMS       64      /* fixed-width query parameters per session */

/* XO_LDCOL imm: low byte = width; XLD_STAGED = the scan kernel's prologue has
 * already put the value into the destination register */
#define XLD_WIDTH_MASK      0xff
#define XLD_STAGED          0x100

/* value domains of a register */
enum {
    XD_INT = 0,     /* bool, int1/2/4/8, date: sign-extended int64 */
    XD_F32 = 1,     /* float4 bits in the low word */
    XD_F64 = 2,     /* float8 bits */
    XD_TEXT = 3,    /* pointer in R[r], length in R[r+1] */
    XD_NUM = 4,     /* numeric: R[r], R[r+1], R[r+2] (value or varlena, kind | scale) */
};

/* comparison operators (aux low byte of XO_CMP*) */
enum { XC_EQ = 0, XC_NE, XC_LT, XC_LE, XC_GT, XC_GE };

/* text operator flags (aux of XO_TEXTLEN, XO_CMP_TEXT, XO_LIKE, XO_SUBSTR) */
#define XENC_UTF8           0x01    /* characters by UTF-8 byte patterns (else 1 byte) */
#define XTXT_BPCHAR         0x02    /* bpchar: trailing spaces do not count */
#define XCMP_BPCHAR         0x200   /* XO_CMP_TEXT: bpchar operands */
#define XLIKE_ICASE         0x04    /* ILIKE: ASCII letters compare upper-cased */
#define XLIKE_NEGATE        0x08    /* NOT LIKE */
#define XSUB_HAS_COUNT      0x100
#define XMM_DOMAIN          0x07    /* XO_MINMAX: XD_* of the values */
#define XMM_GREATEST        0x10    /* XO_MINMAX: keep the greatest */
#define XMM_BPCHAR          0x20
/* XO_HAOP aux: key kind and width, result kind and width, flags */
#define XHA_KIND_INT        1
#define XHA_KIND_F32        2
#define XHA_KIND_F64        3
#define XHA_KIND_TEXT       4
#define XHA_KEY(kind, width)    ((kind) | ((width) << 3))
#define XHA_RES(kind, width)    (((kind) | ((width) << 3)) << 7)
#define XHA_NULL_MISS       0x4000  /* no match gives NULL (IN list with a NULL item) */

#define XHASH_NUMERIC       0x100   /* XO_HASH_ADD of a numeric triple */

/* integer result classes for range checks */
enum { XT_BOOL = 0, XT_INT1, XT_INT2, XT_INT4, XT_INT8 };

enum xop_code {
    XO_NOP = 0,
    /* loads (never raise errors) */
    XO_LDCOL,           /* dst <- Arrow fixed-width column aux; imm = byte width,
                           b = domain, null from the validity bitmap */
    XO_LDCOL_BOOL,      /* dst <- Arrow bit-packed boolean column aux */
    XO_LDCOL_VAR,       /* dst,dst+1 <- Arrow utf8/binary column aux */
    XO_LDTUP,           /* dst <- attribute imm (1-based) of the minimal tuple whose
                           pointer is in R[a], inner KDS of depth aux, b = type code;
                           text occupies dst,dst+1; R[a] == 0 loads NULL */
    XO_CONST,           /* dst <- imm; aux & 1: NULL */
    XO_CONST_TEXT,      /* dst <- const pool base + (imm & 0xffffffff),
                           dst+1 <- imm >> 32 */
    XO_MOV,             /* dst <- R[a] with null flag */
    XO_MOV2,            /* dst,dst+1 <- R[a],R[a+1] with null flag */
    XO_SETNULL,         /* dst <- NULL */
    /* comparisons: dst(bool) <- R[a] op R[b]; aux = op | (compare_nulls << 8) */
    XO_CMP_I, XO_CMP_F32, XO_CMP_F64, XO_CMP_TEXT,
    /* integer arithmetic with range check into class aux; guard */
    XO_ADD_I, XO_SUB_I, XO_MUL_I, XO_DIV_I, XO_MOD_I,
    XO_NEG_I, XO_ABS_I,
    XO_AND_I, XO_OR_I, XO_XOR_I, XO_NOT_I, XO_SHL_I, XO_SHR_I,
    /* float arithmetic with overflow/zero-divide checks; guard */
    XO_ADD_F32, XO_SUB_F32, XO_MUL_F32, XO_DIV_F32,
    XO_ADD_F64, XO_SUB_F64, XO_MUL_F64, XO_DIV_F64,
    XO_NEG_F32, XO_ABS_F32, XO_NEG_F64, XO_ABS_F64,
    /* conversions; aux = integer class of the target where relevant; guard */
    XO_I2I, XO_F32_I, XO_F64_I, XO_I_F32, XO_I_F64, XO_F32_F64, XO_F64_F32, XO_I_BOOL,
    XO_TEXTLEN,
    /* three-valued logic */
    XO_AND3,            /* dst <- R[a] AND R[b] */
    XO_OR3,
    XO_NOT3,            /* dst <- NOT R[a] */
    XO_ISNULL, XO_ISNOTNULL,
    XO_BOOLTEST,        /* aux = kind 0..5 (IsTrue .. IsNotUnknown) */
    XO_COALESCE,        /* if dst is NULL: dst <- R[a] (with null flag); text aware via b */
    /* guards: dst <- R[b] && predicate(R[a]) (never NULL) */
    XO_G_NOTFALSE,      /* R[a] is NULL or true */
    XO_G_NOTTRUE,       /* R[a] is NULL or false */
    XO_G_ISNULL,        /* R[a] is NULL */
    XO_G_POSITIVE,      /* R[a] > 0 */
    XO_G_NONZERO,       /* R[a] != 0 */
    /* join-qual status: if R[b] && (R[a] NULL or false): dst <- imm */
    XO_JQ_STATUS,
    /* hashing: dst <- 0xffffffff; dst <- merge(dst, hash(R[a])) unless NULL;
       aux = hash width in bytes (0 = text, XHASH_NUMERIC); dst <- dst ^ 0xffffffff */
    XO_HASH_INIT, XO_HASH_ADD, XO_HASH_FIN,
    /* end of the scan quals when the final stage follows in the same pass:
       a lane whose R[a] is not true raises no further errors */
    XO_KILL,
    /* dst <- query parameter imm of the launch (fixed-width parameters are
       run-time values, so plans differing only in them share a kernel) */
    XO_PARAM,
    /* guard: dst <- R[b] && R[a] is true (not NULL and not zero) */
    XO_G_TRUE,
    /* conditional move: where R[guard] is true, dst <- R[a] with its null
       flag (aux & 1: text, two registers) */
    XO_SELECT,
    /* IS DISTINCT FROM: dst holds the '=' result of R[a], R[b]; dst <- false
       if both are NULL, true if one is, else NOT dst (never NULL) */
    XO_DISTINCT,
    /* LEAST / GREATEST step: if R[a] is not NULL and dst is NULL or R[a]
       compares below (above) it, dst <- R[a]; aux = XD_* domain | XMM_* */
    XO_MINMAX,
    /* dst <- LIKE of text R[a],R[a+1] and pattern R[b],R[b+1]; aux = XLIKE_*;
       an invalid escape raises message imm where R[guard] */
    XO_LIKE,
    /* dst,dst+1 <- substring of text R[a],R[a+1] from 1-based start R[b];
       imm = count register | XSUB_HAS_COUNT | message << 16 (a negative
       count raises the message where R[guard]); aux = XENC_* */
    XO_SUBSTR,
    /* dst <- value of the HashedArrayOp table at const pool offset imm for
       the key R[a] (hash-slot walk of the reference's kern_expression
       layout); aux = XHA_* key / result description */
    XO_HAOP,
    /* R[a] (text) holding a compressed or external datum raises "text datum
       is compressed or external" where R[guard]: a grouping key needs its
       bytes */
    XO_TXTCHK,
    /* numeric (register triples, XD_NUM); ops that parse a datum raise its
       error where R[guard] */
    XO_NUM_VL,          /* dst triple <- the varlena of text-form pair R[a],R[a+1]
                           (payload address and length, or raw) */
    XO_CONST_NUM,       /* dst triple <- varlena at const pool offset imm */
    XO_LDCOL_DEC,       /* dst triple <- Arrow Decimal128 column aux at its scale */
    XO_MOV3,            /* dst,dst+1,dst+2 <- R[a],R[a+1],R[a+2] with null flag */
    XO_NUM_CMP,         /* dst(bool) <- R[a] op R[b]; aux as XO_CMP_* */
    XO_NUM_ADD, XO_NUM_SUB, XO_NUM_MUL, XO_NUM_DIV,     /* dst triple; MOD is DIV */
    XO_NUM_NEG, XO_NUM_ABS,
    XO_NUM_INT,         /* dst <- integer of R[a]; aux = integer class */
    XO_NUM_F64, XO_NUM_F32,
    XO_INT_NUM,         /* dst triple <- integer R[a] */
    XO_F64_NUM, XO_F32_NUM,
    XO_NUM_PARSE,       /* dst triple <- R[a] parsed (raises its error where R[guard]) */
    XO_NUM_OPS
};

typedef struct
{
    uint16_t    op;
    uint8_t     dst;
    uint8_t     a;
    uint8_t     b;
    uint8_t     guard;
    uint16_t    aux;
    int64_t     imm;
} xop;      /* 16 bytes */

/* a program segment [begin, end) of the op list */
typedef struct
{
    uint16_t    begin;
    uint16_t    end;
} xseg;

/* where one value lives and how it is written into a tuple */
typedef struct
{
    uint8_t     reg;        /* register (first of two for text) */
    uint8_t     domain;     /* XD_* */
    int8_t      attalign;   /* destination column alignment */
    int8_t      width;      /* fixed byte width, or -1 for varlena */
} xattr;

/* group-by aggregate actions, in kds_final column order */
typedef struct
{
    uint16_t    action;     /* KAGG_ACTION__* */
    uint8_t     arg_reg;    /* register of arg0 (0xff: none) */
    uint8_t     arg_domain;
    uint8_t     filter_reg; /* 0xff: none */
    uint8_t     key_index;  /* VREF: index among the grouping keys */
    uint16_t    state_size; /* bytes of the partial state in the tuple */
    uint32_t    state_off;  /* offset of the state from the payload start */
    int8_t      lds64;      /* 8-byte LDS field array of this action (-1: none) */
    int8_t      lds32;      /* 4-byte LDS field array of this action (-1: none) */
    uint16_t    num_scale;  /* PSUM/PAVG_NUMERIC: decimal scale of the sum */
} xagg;

/* a depth-0 fixed-width column loaded by the scan kernel's tile prologue */
typedef struct
{
    uint16_t    col;        /* kds_src column index */
    uint8_t     width;      /* 1, 2, 4 or 8 bytes */
    uint8_t     f32;        /* float4: zero-extend the raw bits */
    uint8_t     reg;        /* destination register */
    uint8_t     __pad[3];
} xstage;

/* the compiled plan, uploaded once per session */
typedef struct
{
    uint32_t    task_kind;          /* DEVTASK__* */
    uint32_t    nops;
    uint32_t    nrels;
    uint32_t    nregs;
    xseg        seg_scan;           /* depth-0 loads + scan quals */
    uint32_t    nstage;             /* columns of stage[] */
    xstage      stage[XP_MAX_STAGE];
    uint32_t    fused;              /* nrels == 0: seg_scan ends with XO_KILL and
                                       seg_final follows it directly */
    uint32_t    stage_prefix;       /* leading ops of seg_scan that are staged
                                       loads (the scan kernel starts after them) */
    uint8_t     r_scan_ok;          /* bool register, or XP_REG_TRUE */
    uint8_t     r_hash;             /* outer key hash of the depth being probed */
    uint8_t     r_status;           /* join status: 1 / 0 / -1 as ExecGpuJoinQuals */
    uint8_t     r_tuple;            /* pointer to the inner tuple being joined */
    struct {
        xseg    seg_hash;           /* outer key hash of this depth */
        xseg    seg_join;           /* inner loads + join quals */
        xseg    seg_oj;             /* NULL inner loads + pushed-down quals only,
                                       for the LEFT OUTER null-extended row */
        uint8_t left_outer;
        uint8_t __pad1[7];
        uint64_t inner_kds;         /* device address of the inner KDS_FORMAT_HASH */
    } depth[XP_MAX_DEPTH + 1];      /* index 1..nrels */
    xseg        seg_final;          /* projection or group-by argument programs */
    /* projection */
    uint32_t    proj_nattrs;
    xattr       proj[XP_MAX_PROJ];
    /* group-by */
    uint8_t     r_keyhash;
    uint8_t     nkeys;
    uint8_t     lds_fast;           /* keys pack into 63 bits: LDS partial table */
    uint8_t     r_projhash;         /* projection into a pinned hash buffer: the
                                       register of the row's join-key hash (0: none) */
    uint32_t    gb_nattrs;
    uint32_t    gb_t_hoff;          /* MAXALIGN(offsetof(t_bits) + bitmap) */
    uint32_t    gb_states_end;      /* payload offset after the last state */
    /* LDS partial table, struct of arrays over its entries: u64 packed key,
     * u64 fields [lds_n64], u32 row count, u32 fields [lds_n32] */
    uint16_t    lds_n64;
    uint16_t    lds_n32;
    uint32_t    lds_entry_bytes;
    /* spill records of rows that find no room in the LDS table: packed key,
     * then the listed argument registers (8 bytes each) */
    uint8_t     spill_ok;
    uint8_t     spill_nregs;
    uint8_t     spill_regs[XP_MAX_SPILL];
    uint8_t     __pad3[6];
    xattr       keys[XP_MAX_KEYS];
    uint8_t     key_bits[XP_MAX_KEYS];  /* packed width of each key in bits */
    xagg        aggs[XP_MAX_PROJ];
    /* CPU fallback: the attributes of the session's KDS_FORMAT_FALLBACK
     * tuple (the scan slot); reg 0xff for those a depth-0 row leaves NULL */
    uint32_t    fb_nattrs;          /* 0: the session has no CPU fallback */
    uint32_t    __pad4;
    xattr       fb[XP_MAX_PROJ];
    uint64_t    const_pool;         /* device address of text constants */
    xop         ops[XP_MAX_OPS];
} xplan;

#endif /* XPU_PROGRAM_H */
