"""Generate the device-function tables of this build from the opcode numbers
in tests/golden/abi_layout.json (dumped from the reference's xpu_opcodes.h,
WITH_POSTGIS=0).

Outputs (committed, regenerated only when the supported set changes):
  csrc/xpu_opcodes_table.h   C table {opcode, kind, op, arg types, result}
  pgstrom/opcodes.py         the same table for the Python codegen mirror

Only the numbers come from the dump; the classification of each function
(comparison / arithmetic / cast ..., argument and result types) is this
build's own, following PostgreSQL's operator semantics.
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

INT_TYPES = {"1": "int1", "2": "int2", "4": "int4", "8": "int8"}
FLT_TYPES = {"2": "float2", "4": "float4", "8": "float8"}
WIDTH = {"int1": 1, "int2": 2, "int4": 4, "int8": 8, "float2": 2, "float4": 4, "float8": 8}
CMP = ["eq", "ne", "lt", "le", "gt", "ge"]
ARITH = ["pl", "mi", "mul", "div", "mod"]
UNARY = ["up", "um", "abs", "not"]
BITOP = ["and", "or", "xor", "shl", "shr"]


def classify(name):
    """(kind, op, arg_types, result_type) or None if unsupported."""
    m = re.fullmatch(r"(int|float)([1248])([1248]?)(eq|ne|lt|le|gt|ge|pl|mi|mul|div|mod)", name)
    if m:
        fam, a, b, op = m.groups()
        tab = INT_TYPES if fam == "int" else FLT_TYPES
        if a not in tab or (b and b not in tab):
            return None
        t1 = tab[a]
        t2 = tab[b] if b else t1
        if "float2" in (t1, t2):
            return None
        if op in CMP:
            return ("cmp", op, (t1, t2), "bool")
        if op == "mod" and fam != "int":
            return None
        res = t1 if WIDTH[t1] >= WIDTH[t2] else t2
        return ("arith", op, (t1, t2), res)
    m = re.fullmatch(r"(int|float)([248])(up|um|abs|not|and|or|xor|shl|shr)", name)
    if m:
        fam, a, op = m.groups()
        t = (INT_TYPES if fam == "int" else FLT_TYPES)[a]
        if t == "float2":
            return None
        if op in ("up", "um", "abs"):
            return ("unary", op, (t,), t)
        if fam != "int":
            return None
        if op == "not":
            return ("unary", op, (t,), t)
        if op in ("shl", "shr"):
            return ("bitop", op, (t, "int4"), t)
        return ("bitop", op, (t, t), t)
    m = re.fullmatch(r"(int[1248]|float[48])_to_(int[1248]|float[48]|bool)", name)
    if m:
        s, d = m.groups()
        if d == "bool" and s != "int4":
            return None
        return ("cast", "cast", (s,), d)
    if name in ("text_eq", "text_ne", "text_lt", "text_le", "text_gt", "text_ge"):
        return ("textcmp", name[5:], ("text", "text"), "bool")
    m = re.fullmatch(r"bpchar(eq|ne|lt|le|gt|ge)", name)
    if m:
        return ("textcmp", m.group(1), ("bpchar", "bpchar"), "bool")
    if name == "booleq":
        return ("cmp", "eq", ("bool", "bool"), "bool")
    if name == "textlen":
        return ("textlen", "len", ("text",), "int4")
    if name == "bpcharlen":
        return ("textlen", "len", ("bpchar",), "int4")
    # LIKE / NOT LIKE / ILIKE / NOT ILIKE (xpu_textlib.cu:1087-1155)
    m = re.fullmatch(r"(like|textlike|notlike|textnlike|texticlike|texticnlike|"
                     r"bpcharlike|bpcharnlike|bpchariclike|bpcharicnlike)", name)
    if m:
        neg = "nlike" in name or name == "notlike"
        op = ("nilike" if neg else "ilike") if "ic" in name else ("nlike" if neg else "like")
        a0 = "bpchar" if name.startswith("bpchar") else "text"
        return ("like", op, (a0, "text"), "bool")
    # numeric (xpu_numeric.cu): compares, arithmetic, unary, casts
    m = re.fullmatch(r"numeric_(eq|ne|lt|le|gt|ge)", name)
    if m:
        return ("numcmp", m.group(1), ("numeric", "numeric"), "bool")
    m = re.fullmatch(r"numeric_(add|sub|mul|div|mod)", name)
    if m:
        op = {"add": "pl", "sub": "mi", "mul": "mul", "div": "div", "mod": "mod"}[m.group(1)]
        return ("numarith", op, ("numeric", "numeric"), "numeric")
    m = re.fullmatch(r"numeric_(uplus|uminus|abs)", name)
    if m:
        op = {"uplus": "up", "uminus": "um", "abs": "abs"}[m.group(1)]
        return ("numunary", op, ("numeric",), "numeric")
    m = re.fullmatch(r"numeric_to_(int[1248]|float[48])", name)
    if m:
        return ("numcast", "cast", ("numeric",), m.group(1))
    m = re.fullmatch(r"(int[1248]|float[48])_to_numeric", name)
    if m:
        return ("numcast", "cast", (m.group(1),), "numeric")
    if name in ("substr", "substring"):
        return ("substr", "substr", ("text", "int4", "int4"), "text")
    if name in ("substr_nolen", "substring_nolen"):
        return ("substr", "substr", ("text", "int4"), "text")
    return None


def main():
    with open(os.path.join(ROOT, "tests", "golden", "abi_layout.json")) as f:
        d = json.load(f)
    rows = []
    for k, v in d.items():
        if not k.startswith("func:"):
            continue
        name = k[5:]
        c = classify(name)
        if c:
            rows.append((v, name) + c)
    rows.sort()
    # --- C header
    with open(os.path.join(HERE, "csrc", "xpu_opcodes_table.h"), "w") as f:
        f.write("/* Generated by pg-strom_amd/gen_opcodes.py from tests/golden/abi_layout.json.\n"
                " * Device functions executed by this build: opcode number (identical to the\n"
                " * reference FuncOpCode enum, xpu_common.h:397-431 over xpu_opcodes.h),\n"
                " * operation class, argument and result types. */\n")
        f.write("#ifndef XPU_OPCODES_TABLE_H\n#define XPU_OPCODES_TABLE_H\n")
        f.write("/* XFUNC(opcode, name, kind, op, nargs, arg0, arg1, arg2, result) */\n")
        f.write("#define XPU_FUNCTION_TABLE(XFUNC) \\\n")
        for v, name, kind, op, args, res in rows:
            a0 = args[0]
            a1 = args[1] if len(args) > 1 else "Invalid"
            a2 = args[2] if len(args) > 2 else "Invalid"
            f.write("    XFUNC(%d, %s, %s, %s, %d, %s, %s, %s, %s) \\\n"
                    % (v, name, kind, op, len(args), a0, a1, a2, res))
        f.write("\n#endif\n")
    # --- Python
    with open(os.path.join(HERE, "pgstrom", "opcodes.py"), "w") as f:
        f.write('"""Generated by pg-strom_amd/gen_opcodes.py from tests/golden/abi_layout.json:\n'
                "device functions executed by this build, keyed by the reference's device\n"
                'function name (xpu_opcodes.h), with the reference FuncOpCode numbers."""\n')
        f.write("# name: (opcode, kind, op, arg types, result type)\n")
        f.write("FUNCS = {\n")
        for v, name, kind, op, args, res in rows:
            f.write("    %r: (%d, %r, %r, %r, %r),\n" % (name, v, kind, op, args, res))
        f.write("}\n")
        types = {k[5:]: v for k, v in d.items() if k.startswith("type:")}
        f.write("TYPES = %r\n" % dict(sorted(types.items(), key=lambda kv: kv[1])))
        specials = ["ConstExpr", "ParamExpr", "VarExpr", "BoolExpr_And", "BoolExpr_Or",
                    "BoolExpr_Not", "NullTestExpr_IsNull", "NullTestExpr_IsNotNull",
                    "BoolTestExpr_IsTrue", "BoolTestExpr_IsNotTrue", "BoolTestExpr_IsFalse",
                    "BoolTestExpr_IsNotFalse", "BoolTestExpr_IsUnknown",
                    "BoolTestExpr_IsNotUnknown", "DistinctFrom", "CoalesceExpr",
                    "LeastExpr", "GreatestExpr", "CaseWhenExpr", "ScalarArrayOpAny",
                    "ScalarArrayOpAll", "HashedArrayOp", "LoadVars", "MoveVars",
                    "JoinQuals", "HashValue", "SaveExpr", "AggFuncs", "Projection", "SortKeys",
                    "Packed"]
        f.write("SPECIAL = {\n")
        for s in specials:
            f.write("    %r: %d,\n" % (s, d["func:" + s]))
        f.write("}\n")
    print("%d device functions" % len(rows))


if __name__ == "__main__":
    main()
