"""Write tests/golden/pg_hash_any.json: the reference's own pg_hash_any
(xpu_common.cu:2920-3041, built into oracle/_ref/libpgstrom_ref.so by
`make -C oracle ref`) over byte strings of every length 0..64 and the 1/2/4/8
byte images of boundary integers.  Run in the build container."""
import ctypes
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libpgstrom_ref.so"))
    lib.ref_pg_hash_any.restype = ctypes.c_uint32
    lib.ref_pg_hash_any.argtypes = [ctypes.c_char_p, ctypes.c_int]
    cases = []
    for n in range(0, 65):
        data = bytes((i * 37 + n * 11 + 5) & 0xff for i in range(n))
        cases.append(data)
    for fmt, vals in (("<b", [-128, -1, 0, 1, 127]),
                      ("<h", [-32768, -1, 0, 1, 32767]),
                      ("<i", [-2**31, -1, 0, 1, 7, 500000, 2**31 - 1]),
                      ("<q", [-2**63, -1, 0, 1, 10**12, 2**63 - 1]),
                      ("<f", [0.0, -0.0, 1.5, float("inf")]),
                      ("<d", [0.0, -0.0, 3.25, -1e300])):
        for v in vals:
            cases.append(struct.pack(fmt, v))
    out = [{"hex": c.hex(), "hash": lib.ref_pg_hash_any(c, len(c))} for c in cases]
    with open(os.path.join(HERE, "pg_hash_any.json"), "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()
