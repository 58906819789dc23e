"""Write tests/golden/expr_vectors.json and tests/golden/func_catalog.json.

expr_vectors.json: every device function this build executes (pgstrom
opcodes.FUNCS), and the composite constructs of exprcases.templates,
evaluated by the reference's own device code - EXEC_KERN_EXPRESSION over the
pgfn_* of xpu_common.cu / xpu_basetype.cu / xpu_textlib.cu / xpu_numeric.cu,
compiled as host C++ into oracle/_ref/libpgstrom_ref.so by `make -C oracle
ref` (ref_eval_expr in oracle/ref_driver.cc) - on the edge-case inputs of
exprcases.VALUES.  Each case records the result's datum image, NULL, or the
error code and message the reference raised.

func_catalog.json: the reference's device-function list as its
xpu_opcodes.h declares it (device name, opcode, SQL name, argument types).

Run in the build container (it needs /root/reference):
    make -C oracle ref && python3 tests/golden/gen_expr_vectors.py
"""
import ctypes as C
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "pg-strom_amd"))
sys.path.insert(0, HERE)

from pgstrom import codegen as cg, opcodes  # noqa: E402
import exprcases as X  # noqa: E402

TYPE_NAME = {v: k for k, v in opcodes.TYPES.items()}


class RefLib:
    def __init__(self):
        L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libpgstrom_ref.so"))
        vp = C.c_void_p
        L.ref_eval_expr.restype = C.c_int
        L.ref_eval_expr.argtypes = [vp, C.c_size_t, C.c_int, C.c_char_p, C.c_int, vp, vp, vp, vp,
                                    vp, C.c_int32, vp, vp, C.c_char_p, C.c_int]
        L.ref_func_catalog_json.restype = C.c_char_p
        self.L = L

    def eval(self, blob, compare_nulls, slot_types, images):
        n = len(slot_types)
        types = (C.c_int32 * max(n, 1))(*[opcodes.TYPES[t] for t in slot_types])
        data = bytearray()
        offs = []
        for img in images:
            if img is None:
                offs.append(-1)
                continue
            data += b"\0" * ((-len(data)) % 8)
            offs.append(len(data))
            data += img
        data += b"\0" * 16
        offs = (C.c_int64 * max(n, 1))(*offs)
        dbuf = C.create_string_buffer(bytes(data), len(data))
        rt, rl, ec = C.c_int32(), C.c_int32(), C.c_uint32()
        img = C.create_string_buffer(8192)
        msg = C.create_string_buffer(256)
        rc = self.L.ref_eval_expr(blob, len(blob), compare_nulls, X.ENCODING.encode(), n, types,
                                  dbuf, offs,
                                  C.byref(rt), img, 8192, C.byref(rl), C.byref(ec), msg, 256)
        if rc < 0:
            raise RuntimeError("reference cannot evaluate this expression")
        if rc == 1:
            return ["E", ec.value, msg.value.decode()], None
        if rl.value < 0:
            return None, None
        return img.raw[:rl.value].hex(), TYPE_NAME.get(rt.value, rt.value)


def blob_of(expr):
    b = bytearray()
    cg.emit_expr(b, expr)
    return bytes(b)


def main():
    ref = RefLib()
    catalog = json.loads(ref.L.ref_func_catalog_json())
    with open(os.path.join(HERE, "func_catalog.json"), "w") as f:
        json.dump(catalog, f, indent=0, sort_keys=True)
        f.write("\n")
    values = {t: [v.hex() for v in vs] for t, vs in X.VALUES.items()}
    funcs = {}
    for name in sorted(opcodes.FUNCS, key=lambda n: opcodes.FUNCS[n][0]):
        opc, kind, op, args, res = opcodes.FUNCS[name]
        expr = cg.Func(name, *[cg.Var(i, t) for i, t in enumerate(args)])
        blob = blob_of(expr)
        cases, rtype = [], None
        lists = X.func_arg_lists(kind, list(args))
        for tup in X.arg_tuples(list(args), kind):
            imgs = [None if i is None else lists[j][i] for j, i in enumerate(tup)]
            if X.host_traps(name, args, imgs):
                continue
            out, t = ref.eval(blob, 0, list(args), imgs)
            rtype = rtype or t
            cases.append(list(tup) + [out])
        funcs[name] = {"opcode": opc, "args": list(args), "result": rtype or res, "cases": cases}
    templates = {}
    for name, expr, slots, cmpn in X.templates(cg):
        blob = blob_of(expr)
        cases, rtype = [], None
        combos = [[]]
        for t in slots:
            combos = [c + [i] for c in combos for i in range(len(X.TEMPLATE_VALUES[t]))]
        for combo in combos:
            imgs = [X.TEMPLATE_VALUES[t][i] for t, i in zip(slots, combo)]
            if X.template_traps(name, imgs):
                continue
            out, t = ref.eval(blob, cmpn, slots, imgs)
            rtype = rtype or t
            cases.append(combo + [out])
        templates[name] = {"slots": slots, "compare_nulls": cmpn, "result": rtype or expr.typname,
                           "cases": cases}
    with open(os.path.join(HERE, "expr_vectors.json"), "w") as f:
        f.write('{"values": %s,\n' % json.dumps(values))
        f.write('"funcs": {\n')
        f.write(",\n".join('%s: %s' % (json.dumps(k), json.dumps(v, separators=(",", ":")))
                           for k, v in funcs.items()))
        f.write('},\n"templates": {\n')
        f.write(",\n".join('%s: %s' % (json.dumps(k), json.dumps(v, separators=(",", ":")))
                           for k, v in templates.items()))
        f.write("}}\n")
    print("%d functions, %d templates, %d cases" % (
        len(funcs), len(templates),
        sum(len(v["cases"]) for v in funcs.values()) +
        sum(len(v["cases"]) for v in templates.values())))


if __name__ == "__main__":
    main()
