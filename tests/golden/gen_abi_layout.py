"""Write tests/golden/abi_layout.json: struct layouts and opcode numbers of
the reference ABI, dumped by oracle/_ref/libpgstrom_ref.so (built by
`make -C oracle ref` from the reference headers).  Run in the build container,
where the reference sources exist; the JSON is the committed fixture."""
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libpgstrom_ref.so"))
    lib.ref_abi_json.restype = ctypes.c_char_p
    d = json.loads(lib.ref_abi_json())
    with open(os.path.join(HERE, "abi_layout.json"), "w") as f:
        json.dump(d, f, indent=0, sort_keys=True)
        f.write("\n")
    print("wrote %d entries" % len(d), file=sys.stderr)


if __name__ == "__main__":
    main()
