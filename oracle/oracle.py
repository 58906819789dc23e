"""ctypes binding of liboracle.so (the CPU restatement; test infrastructure).
Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.oracle_session_open.restype = C.c_void_p
        L.oracle_session_open.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
        L.oracle_task_exec.restype = C.c_int
        L.oracle_task_exec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_size_t, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_task_final.restype = C.c_int64
        L.oracle_task_final.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.oracle_session_fallback.restype = C.c_int64
        L.oracle_session_fallback.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.oracle_session_close.argtypes = [C.c_void_p]
        L.oracle_hash_any.restype = C.c_uint32
        L.oracle_hash_any.argtypes = [C.c_void_p, C.c_int]
        vp = C.c_void_p
        L.oracle_eval_expr.restype = C.c_int
        L.oracle_eval_expr.argtypes = [vp, C.c_size_t, C.c_int, C.c_char_p, C.c_int, vp, vp, vp,
                                       C.c_int32, vp, C.c_int32, vp, vp, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return C.cast(C.c_char_p(bytes(a)), C.c_void_p).value


class OracleSession:
    def __init__(self, session_blob):
        from pgstrom import abi
        self.abi = abi
        self._blob = np.frombuffer(bytes(session_blob), dtype=np.uint8).copy()
        self.err = abi.kern_errorbuf()
        self.h = lib().oracle_session_open(self._blob.ctypes.data, len(self._blob),
                                           C.byref(self.err))
        if not self.h:
            raise RuntimeError("oracle_session_open: " + self.err.describe())

    def exec(self, kds_src, kmrels=None, dst_length=0, nthreads=1):
        abi = self.abi
        dst = np.zeros(dst_length, dtype=np.uint8) if dst_length else None
        res = abi.kern_exec_results()
        rc = lib().oracle_task_exec(self.h, _ptr(kds_src), _ptr(kmrels),
                                    _ptr(dst), dst_length, nthreads, C.byref(res),
                                    C.byref(self.err))
        return rc, dst, res

    def final(self):
        need = lib().oracle_task_final(self.h, None, 0, C.byref(self.err))
        if need < 0:
            raise RuntimeError(self.err.describe())
        buf = np.zeros(need, dtype=np.uint8)
        got = lib().oracle_task_final(self.h, buf.ctypes.data, need, C.byref(self.err))
        if got < 0:
            raise RuntimeError(self.err.describe())
        return buf[:got]        # a LIMIT packs the kept tuples into fewer bytes

    def fallback(self):
        """The CPU-fallback rows of the tasks so far (KDS_FORMAT_FALLBACK
        bytes), or None."""
        need = lib().oracle_session_fallback(self.h, None, 0)
        if need <= 0:
            return None
        buf = np.zeros(need, dtype=np.uint8)
        got = lib().oracle_session_fallback(self.h, buf.ctypes.data, need)
        if got != need:
            raise RuntimeError("oracle_session_fallback")
        return buf

    def close(self):
        if self.h:
            lib().oracle_session_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


def hash_any(data):
    b = bytes(data)
    return lib().oracle_hash_any(C.c_char_p(b), len(b))


def eval_expr(blob, slot_type_codes, images, res_type_code, compare_nulls=0, encoding="UTF8"):
    """One expression over variable slots (oracle_eval_expr): returns the
    result's datum image, None for NULL, or ("E", errcode, message)."""
    n = len(slot_type_codes)
    types = (C.c_int32 * max(n, 1))(*slot_type_codes)
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
    rl, ec = C.c_int32(), C.c_uint32()
    out = C.create_string_buffer(8192)
    msg = C.create_string_buffer(256)
    b = bytes(blob)
    rc = lib().oracle_eval_expr(b, len(b), compare_nulls, (encoding or "").encode(), n, types,
                                dbuf, offs, res_type_code, out,
                                8192, C.byref(rl), C.byref(ec), msg, 256)
    if rc < 0:
        raise RuntimeError("oracle cannot evaluate this expression")
    if rc == 1:
        return ("E", ec.value, msg.value.decode())
    if rl.value < 0:
        return None
    return out.raw[:rl.value]
