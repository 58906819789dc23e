"""ctypes binding of libxpuhip.so (include/xpuhip.h), the MI355X xPU library.

Loading fails with an exception when the shared library is missing: there is
no CPU fallback behind this module."""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(HERE), "libxpuhip.so")

EXPORTS = [
    "xpuhip_init", "xpuhip_shutdown", "xpuhip_session_open", "xpuhip_session_check",
    "xpuhip_session_attach_inner", "xpuhip_task_exec", "xpuhip_task_exec_device",
    "xpuhip_task_exec_device_tuples", "xpuhip_task_exec_device_noalloc",
    "xpuhip_task_final", "xpuhip_session_fallback", "xpuhip_session_close",
    "xpuhip_session_set_tuning",
    "xpuhip_session_get_tuning", "xpuhip_device_alloc", "xpuhip_device_free",
    "xpuhip_memcpy_htod", "xpuhip_memcpy_dtoh", "xpuhip_memcpy_dtod", "xpuhip_device_synchronize",
    "xpuhip_last_kernel_ms", "xpuhip_abi_layout_json", "xpuhip_session_explain",
    "xpuhip_last_task_stats", "xpuhip_jit_prepare", "xpuhip_batch_exec", "xpuhip_session_reset",
    "xpuhip_session_autotune", "xpuhip_batch_autotune",
    "xpuhip_server_create", "xpuhip_server_submit", "xpuhip_server_wait", "xpuhip_server_stats",
    "xpuhip_server_destroy", "xpuhip_inner_hash_setup",
    "xpuhip_arrow_open", "xpuhip_arrow_field", "xpuhip_arrow_load_batch", "xpuhip_arrow_close",
]
REQ_BATCHABLE = 1
BATCH_GRAPH = 1

JIT_AUTO, JIT_ON, JIT_OFF = 0, 1, 2
JOIN_FILTER_AUTO, JOIN_FILTER_ON, JOIN_FILTER_OFF = 0, 1, 2
JOIN_PASSES_AUTO, JOIN_PASSES_ON, JOIN_PASSES_OFF = 0, 1, 2


class XpuError(RuntimeError):
    def __init__(self, code, errbuf):
        self.code = code
        self.errbuf = errbuf
        super().__init__(errbuf.describe())


class xpuhip_tuning(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("rows_per_thread", C.c_uint32),
                ("lds_nents", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("jit", C.c_uint32), ("join_filter", C.c_uint32), ("join_passes", C.c_uint32)]


class xpuhip_batch_task(C.Structure):
    _fields_ = [("sess", C.c_void_p), ("d_kds_src", C.c_void_p), ("nitems", C.c_uint32),
                ("_pad", C.c_uint32), ("results", C.c_void_p)]


class xpuhip_server_config(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_uint32), ("max_batch", C.c_uint32)]


class xpuhip_request(C.Structure):
    _fields_ = [("sess", C.c_void_p), ("d_kds_src", C.c_void_p), ("nitems", C.c_uint32),
                ("flags", C.c_uint32), ("d_kds_dst", C.c_void_p), ("dst_length", C.c_size_t),
                ("hbm_bytes", C.c_uint64), ("shape", C.c_uint64)]


class xpuhip_request_status(C.Structure):
    _fields_ = [("rc", C.c_int), ("batched", C.c_uint32), ("queue_ms", C.c_double),
                ("exec_ms", C.c_double), ("total_ms", C.c_double),
                ("results", abi.kern_exec_results), ("errbuf", abi.kern_errorbuf)]


class xpuhip_task_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("pass1_ms", C.c_double), ("spill_ms", C.c_double),
                ("grid", C.c_uint32), ("block", C.c_uint32),
                ("ntiles", C.c_uint32), ("lds_nents", C.c_uint32), ("lds_bytes", C.c_uint64),
                ("rows", C.c_uint64), ("lds_miss", C.c_uint64), ("global_merges", C.c_uint64),
                ("spill_rows", C.c_uint64), ("jit", C.c_uint32), ("_pad", C.c_uint32),
                ("heap_ms", C.c_double), ("heap_rows", C.c_uint64),
                ("sort_ms", C.c_double), ("sort_items", C.c_uint32), ("sort_passes", C.c_uint32),
                ("join_slices", C.c_uint32), ("final_grows", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libxpuhip.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        vp, sz = C.c_void_p, C.c_size_t
        L.xpuhip_init.argtypes = [C.c_int, vp]
        L.xpuhip_session_open.argtypes = [vp, sz, C.POINTER(vp), vp]
        L.xpuhip_session_check.argtypes = [vp, sz, vp]
        L.xpuhip_inner_hash_setup.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, C.c_int, vp]
        L.xpuhip_arrow_open.argtypes = [vp, sz, C.POINTER(vp), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32), vp]
        L.xpuhip_arrow_field.argtypes = [vp, C.c_uint32, C.c_char_p, sz, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.xpuhip_arrow_load_batch.restype = C.c_int64
        L.xpuhip_arrow_load_batch.argtypes = [vp, C.c_uint32, vp, vp, vp, sz, vp]
        L.xpuhip_arrow_close.argtypes = [vp]
        L.xpuhip_session_attach_inner.argtypes = [vp, vp, sz, vp]
        L.xpuhip_task_exec.argtypes = [vp, vp, sz, vp, sz, C.POINTER(C.c_uint64), vp, vp]
        L.xpuhip_task_exec_device.argtypes = [vp, vp, C.c_uint32, vp, sz,
                                              C.POINTER(C.c_uint64), vp, vp]
        L.xpuhip_task_exec_device_tuples.argtypes = [vp, vp, vp, vp, sz, C.POINTER(C.c_uint64),
                                                     vp, vp]
        L.xpuhip_task_exec_device_noalloc.argtypes = [vp, vp, C.c_uint32, vp, sz,
                                                      C.POINTER(C.c_uint64), vp, vp]
        L.xpuhip_task_final.restype = C.c_int64
        L.xpuhip_task_final.argtypes = [vp, vp, sz, vp, vp]
        L.xpuhip_session_fallback.restype = C.c_int64
        L.xpuhip_session_fallback.argtypes = [vp, vp, sz, vp]
        L.xpuhip_session_close.argtypes = [vp]
        L.xpuhip_session_set_tuning.argtypes = [vp, vp]
        L.xpuhip_session_get_tuning.argtypes = [vp, vp]
        L.xpuhip_device_alloc.argtypes = [sz, C.POINTER(vp), vp]
        L.xpuhip_device_free.argtypes = [vp]
        L.xpuhip_memcpy_htod.argtypes = [vp, vp, sz]
        L.xpuhip_memcpy_dtoh.argtypes = [vp, vp, sz]
        L.xpuhip_memcpy_dtod.argtypes = [vp, vp, sz]
        L.xpuhip_last_kernel_ms.restype = C.c_double
        L.xpuhip_abi_layout_json.restype = C.c_char_p
        L.xpuhip_session_explain.restype = C.c_int64
        L.xpuhip_session_explain.argtypes = [vp, sz, vp, sz, vp]
        L.xpuhip_last_task_stats.argtypes = [vp]
        L.xpuhip_jit_prepare.argtypes = [vp, sz, vp]
        L.xpuhip_batch_exec.argtypes = [vp, C.c_uint32, C.c_uint32, vp]
        L.xpuhip_session_reset.argtypes = [vp, vp]
        L.xpuhip_session_autotune.argtypes = [vp, vp, C.c_uint32, vp, sz, vp, vp]
        L.xpuhip_batch_autotune.argtypes = [vp, C.c_uint32, vp, vp]
        L.xpuhip_server_create.restype = vp
        L.xpuhip_server_create.argtypes = [vp, vp]
        L.xpuhip_server_submit.restype = C.c_uint64
        L.xpuhip_server_submit.argtypes = [vp, vp]
        L.xpuhip_server_wait.argtypes = [vp, C.c_uint64, vp]
        L.xpuhip_server_stats.argtypes = [vp, vp, vp]
        L.xpuhip_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _buf(a):
    a = np.ascontiguousarray(np.frombuffer(bytes(a), dtype=np.uint8)
                             if not isinstance(a, np.ndarray) else a)
    return a


def explain(blob):
    """Listing of the programs the library compiles from a session blob."""
    err = abi.kern_errorbuf()
    b = _buf(blob)
    n = lib().xpuhip_session_explain(b.ctypes.data, len(b), None, 0, C.byref(err))
    if n < 0:
        raise XpuError(abi.ERRCODE_DEVICE_ERROR, err)
    out = C.create_string_buffer(n + 1)
    lib().xpuhip_session_explain(b.ctypes.data, len(b), out, n + 1, C.byref(err))
    return out.value.decode()


def init(device=0):
    err = abi.kern_errorbuf()
    rc = lib().xpuhip_init(device, C.byref(err))
    if rc:
        raise XpuError(rc, err)


def jit_prepare(blob):
    """Compiles the kernel specialized for a session blob (no GPU needed)."""
    b = _buf(blob)
    err = abi.kern_errorbuf()
    rc = lib().xpuhip_jit_prepare(b.ctypes.data, len(b), C.byref(err))
    if rc:
        raise XpuError(rc, err)


def session_check(blob):
    err = abi.kern_errorbuf()
    b = _buf(blob)
    rc = lib().xpuhip_session_check(b.ctypes.data, len(b), C.byref(err))
    return rc, err


def inner_hash_setup(coldefs, mtups, key_attnums, nthreads=0, nslots=None):
    """KDS_FORMAT_HASH buffer of an inner relation built by the library's
    xpuhip_inner_hash_setup (needs no GPU): coldefs are the relation's
    (name, type) columns, mtups its rows as MinimalTuple bytes (t_len
    without a row id), key_attnums the 1-based join key attributes.  The
    buffer is sized exactly, with KDS_GET_HASHSLOT_WIDTH slots unless nslots
    is given."""
    from . import kds as K
    n = len(mtups)
    if nslots is None:
        nslots = abi.kds_get_hashslot_width(n)
    head = K.kds_header(abi.KDS_FORMAT_HASH, coldefs, hash_nslots=nslots)
    lens = np.fromiter((len(t) for t in mtups), dtype=np.uint64, count=n)
    usage = int(((lens + 8 + abi.ROWID_SIZE + 7) & ~np.uint64(7)).sum())
    length = abi.maxalign(len(head) + 8 * (nslots + n) + usage)
    out = np.zeros(length, dtype=np.uint8)
    out[:len(head)] = np.frombuffer(bytes(head), dtype=np.uint8)
    abi.kern_data_store_head.from_buffer(out).length = length
    packed = np.frombuffer(b"".join(bytes(t) for t in mtups) or b"\0", dtype=np.uint8)
    ptrs = np.zeros(max(1, n), dtype=np.uint64)
    if n:
        ptrs[1:] = np.cumsum(lens[:-1])
        ptrs += np.uint64(packed.ctypes.data)
    keys = np.asarray(key_attnums, dtype=np.int16)
    err = abi.kern_errorbuf()
    rc = lib().xpuhip_inner_hash_setup(out.ctypes.data, ptrs.ctypes.data, n, keys.ctypes.data,
                                       len(keys), int(nthreads), C.byref(err))
    if rc != 0:
        raise XpuError(err.errcode, err)
    return out


class ArrowFile:
    """An Arrow IPC file opened by the library's chunk producer
    (xpuhip_arrow_open); batch(i, coldefs, field_names) returns record batch
    i as a KDS_FORMAT_ARROW chunk with the relation's columns coldefs
    [(name, typname)], each read from the Arrow field of the same name
    (or of field_names[j]; None = not referenced).  Needs no GPU."""

    def __init__(self, image):
        self._img = np.frombuffer(bytes(image), dtype=np.uint8)
        self.h = C.c_void_p()
        nb, nf = C.c_uint32(), C.c_uint32()
        err = abi.kern_errorbuf()
        rc = lib().xpuhip_arrow_open(self._img.ctypes.data, len(self._img), C.byref(self.h),
                                     C.byref(nb), C.byref(nf), C.byref(err))
        if rc != 0:
            raise XpuError(err.errcode, err)
        self.nbatches = nb.value
        self.fields = []
        for i in range(nf.value):
            name = C.create_string_buffer(256)
            t, w, sc = C.c_int32(), C.c_int32(), C.c_int32()
            lib().xpuhip_arrow_field(self.h, i, name, 256, C.byref(t), C.byref(w), C.byref(sc))
            self.fields.append((name.value.decode(), t.value, w.value, sc.value))

    def batch(self, rb_index, coldefs, field_names=None):
        from . import kds as K
        names = [f[0] for f in self.fields]
        field_names = field_names or [n for n, _ in coldefs]
        index = np.array([-1 if n is None else names.index(n) for n in field_names],
                         dtype=np.int32)
        head = np.frombuffer(bytes(K.kds_header(abi.KDS_FORMAT_ARROW, coldefs)), dtype=np.uint8)
        err = abi.kern_errorbuf()
        n = lib().xpuhip_arrow_load_batch(self.h, rb_index, head.ctypes.data, index.ctypes.data,
                                          None, 0, C.byref(err))
        if n < 0:
            raise XpuError(err.errcode, err)
        out = np.zeros(n, dtype=np.uint8)
        n = lib().xpuhip_arrow_load_batch(self.h, rb_index, head.ctypes.data, index.ctypes.data,
                                          out.ctypes.data, len(out), C.byref(err))
        if n < 0:
            raise XpuError(err.errcode, err)
        return out

    def close(self):
        if self.h:
            lib().xpuhip_arrow_close(self.h)
            self.h = C.c_void_p()


class DeviceBuffer:
    """A device allocation holding a copy of a host buffer (or empty)."""

    def __init__(self, nbytes, host=None):
        err = abi.kern_errorbuf()
        p = C.c_void_p()
        rc = lib().xpuhip_device_alloc(max(nbytes, 8), C.byref(p), C.byref(err))
        if rc:
            raise XpuError(rc, err)
        self.ptr = p.value
        self.nbytes = nbytes
        if host is not None:
            h = _buf(host)
            assert len(h) <= nbytes
            if lib().xpuhip_memcpy_htod(self.ptr, h.ctypes.data, len(h)):
                raise RuntimeError("xpuhip_memcpy_htod failed")

    def to_host(self, nbytes=None, offset=0):
        n = self.nbytes if nbytes is None else nbytes
        out = np.empty(n, dtype=np.uint8)
        if lib().xpuhip_memcpy_dtoh(out.ctypes.data, self.ptr + offset, n):
            raise RuntimeError("xpuhip_memcpy_dtoh failed")
        return out

    def free(self):
        if self.ptr:
            lib().xpuhip_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Session:
    def __init__(self, blob):
        self.err = abi.kern_errorbuf()
        self._blob = _buf(blob)
        h = C.c_void_p()
        rc = lib().xpuhip_session_open(self._blob.ctypes.data, len(self._blob), C.byref(h),
                                       C.byref(self.err))
        if rc:
            raise XpuError(rc, self.err)
        self.h = h.value

    def attach_inner(self, kmrels):
        self._kmrels = _buf(kmrels)
        rc = lib().xpuhip_session_attach_inner(self.h, self._kmrels.ctypes.data,
                                               len(self._kmrels), C.byref(self.err))
        if rc:
            raise XpuError(rc, self.err)

    def set_tuning(self, **kw):
        t = xpuhip_tuning(**kw)
        rc = lib().xpuhip_session_set_tuning(self.h, C.byref(t))
        if rc:
            raise ValueError("invalid tuning %r" % kw)

    def get_tuning(self):
        t = xpuhip_tuning()
        lib().xpuhip_session_get_tuning(self.h, C.byref(t))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def exec(self, kds_src, dst_length=0, resume_pos=0):
        """Host-resident chunk; returns (rc, kds_dst array or None, results, resume_pos)."""
        src = _buf(kds_src)
        dst = np.zeros(dst_length, dtype=np.uint8) if dst_length else None
        res = abi.kern_exec_results()
        pos = C.c_uint64(resume_pos)
        rc = lib().xpuhip_task_exec(self.h, src.ctypes.data, len(src),
                                    dst.ctypes.data if dst is not None else None, dst_length,
                                    C.byref(pos), C.byref(res), C.byref(self.err))
        if rc and rc != abi.ERRCODE_SUSPEND_NO_SPACE:
            raise XpuError(rc, self.err)
        return rc, dst, res, pos.value

    def exec_device(self, d_src, nitems, d_dst=None, dst_length=0, resume_pos=0):
        res = abi.kern_exec_results()
        pos = C.c_uint64(resume_pos)
        rc = lib().xpuhip_task_exec_device(self.h, d_src.ptr if d_src else None, nitems,
                                           d_dst.ptr if d_dst else None, dst_length,
                                           C.byref(pos), C.byref(res), C.byref(self.err))
        if rc and rc != abi.ERRCODE_SUSPEND_NO_SPACE:
            raise XpuError(rc, self.err)
        return rc, res, pos.value

    def exec_device_noalloc(self, d_src, nitems, d_dst=None, dst_length=0, resume_pos=0):
        """exec_device through the entry point that allocates nothing."""
        res = abi.kern_exec_results()
        pos = C.c_uint64(resume_pos)
        rc = lib().xpuhip_task_exec_device_noalloc(self.h, d_src.ptr if d_src else None, nitems,
                                                   d_dst.ptr if d_dst else None, dst_length,
                                                   C.byref(pos), C.byref(res), C.byref(self.err))
        if rc and rc != abi.ERRCODE_SUSPEND_NO_SPACE:
            raise XpuError(rc, self.err)
        return rc, res, pos.value

    def exec_device_tuples(self, d_src, head, d_dst=None, dst_length=0, resume_pos=0):
        """Heap-page / row chunk in device memory; head: host bytes of its
        header (at least KDS_HEAD_LENGTH)."""
        hb = _buf(head)
        res = abi.kern_exec_results()
        pos = C.c_uint64(resume_pos)
        rc = lib().xpuhip_task_exec_device_tuples(self.h, d_src.ptr, hb.ctypes.data,
                                                  d_dst.ptr if d_dst else None, dst_length,
                                                  C.byref(pos), C.byref(res), C.byref(self.err))
        if rc and rc != abi.ERRCODE_SUSPEND_NO_SPACE:
            raise XpuError(rc, self.err)
        return rc, res, pos.value

    def autotune(self, d_src, nitems, d_dst=None, dst_length=0):
        """Tries the candidate launch shapes on this chunk; keeps the fastest."""
        t = xpuhip_tuning()
        rc = lib().xpuhip_session_autotune(self.h, d_src.ptr, nitems, d_dst.ptr if d_dst else None,
                                           dst_length, C.byref(t), C.byref(self.err))
        if rc:
            raise XpuError(rc, self.err)
        return {k: getattr(t, k) for k, _ in t._fields_}

    def reset(self):
        """Clears the session's kds_final."""
        rc = lib().xpuhip_session_reset(self.h, C.byref(self.err))
        if rc:
            raise XpuError(rc, self.err)

    def final(self):
        need = lib().xpuhip_task_final(self.h, None, 0, None, C.byref(self.err))
        if need < 0:
            raise XpuError(abi.ERRCODE_DEVICE_ERROR, self.err)
        buf = np.zeros(need, dtype=np.uint8)
        got = lib().xpuhip_task_final(self.h, buf.ctypes.data, need, None, C.byref(self.err))
        if got < 0:
            raise XpuError(abi.ERRCODE_DEVICE_ERROR, self.err)
        return buf[:got]        # a LIMIT packs the kept tuples into fewer bytes

    def fallback(self):
        """The CPU-fallback rows the session's tasks wrote since the last
        call (KDS_FORMAT_FALLBACK bytes), or None."""
        need = lib().xpuhip_session_fallback(self.h, None, 0, C.byref(self.err))
        if need < 0:
            raise XpuError(abi.ERRCODE_DEVICE_ERROR, self.err)
        if need == 0:
            return None
        buf = np.zeros(need, dtype=np.uint8)
        got = lib().xpuhip_session_fallback(self.h, buf.ctypes.data, need, C.byref(self.err))
        if got < 0:
            raise XpuError(abi.ERRCODE_DEVICE_ERROR, self.err)
        return buf[:got]

    def close(self):
        if getattr(self, "h", None):
            lib().xpuhip_session_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


def batch_exec(tasks, graph=True):
    """One launch over many (Session, DeviceBuffer, nitems) fragments of one
    plan shape; returns the per-fragment kern_exec_results."""
    arr = (xpuhip_batch_task * len(tasks))()
    results = [abi.kern_exec_results() for _ in tasks]
    for i, (sess, dbuf, n) in enumerate(tasks):
        arr[i].sess = sess.h
        arr[i].d_kds_src = dbuf if isinstance(dbuf, int) else dbuf.ptr
        arr[i].nitems = n
        arr[i].results = C.addressof(results[i])
    err = abi.kern_errorbuf()
    rc = lib().xpuhip_batch_exec(arr, len(tasks), BATCH_GRAPH if graph else 0, C.byref(err))
    if rc:
        raise XpuError(rc, err)
    return results


class Server:
    """The request loop of the GPU service (include/xpuhip.h, xpuhip_server_*):
    slots = 1, max_batch = 1 is the one-query-at-a-time service."""

    def __init__(self, hbm_budget, slots, max_batch):
        cfg = xpuhip_server_config(hbm_budget, slots, max_batch)
        err = abi.kern_errorbuf()
        self.h = lib().xpuhip_server_create(C.byref(cfg), C.byref(err))
        if not self.h:
            raise XpuError(abi.ERRCODE_DEVICE_ERROR, err)

    def submit(self, sess, d_src, nitems, d_dst=None, dst_length=0, hbm_bytes=0, shape=0,
               batchable=False):
        r = xpuhip_request(sess.h, d_src if isinstance(d_src, int) else d_src.ptr, nitems,
                           REQ_BATCHABLE if batchable else 0,
                           (d_dst if isinstance(d_dst, int) else d_dst.ptr) if d_dst else None,
                           dst_length, hbm_bytes, shape)
        return lib().xpuhip_server_submit(self.h, C.byref(r))

    def wait(self, ticket):
        st = xpuhip_request_status()
        if lib().xpuhip_server_wait(self.h, ticket, C.byref(st)):
            raise KeyError(ticket)
        return st

    def stats(self):
        a, b = C.c_uint64(), C.c_uint64()
        lib().xpuhip_server_stats(self.h, C.byref(a), C.byref(b))
        return {"launches": a.value, "coalesced": b.value}

    def close(self):
        if getattr(self, "h", None):
            lib().xpuhip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def batch_autotune(tasks):
    arr = (xpuhip_batch_task * len(tasks))()
    for i, (sess, dbuf, n) in enumerate(tasks):
        arr[i].sess = sess.h
        arr[i].d_kds_src = dbuf if isinstance(dbuf, int) else dbuf.ptr
        arr[i].nitems = n
    t = xpuhip_tuning()
    err = abi.kern_errorbuf()
    rc = lib().xpuhip_batch_autotune(arr, len(tasks), C.byref(t), C.byref(err))
    if rc:
        raise XpuError(rc, err)
    return {k: getattr(t, k) for k, _ in t._fields_}


def last_kernel_ms():
    return lib().xpuhip_last_kernel_ms()


def last_task_stats():
    """Launch shape and counters of the calling thread's last task."""
    st = xpuhip_task_stats()
    lib().xpuhip_last_task_stats(C.byref(st))
    return {k: getattr(st, k) for k, _ in st._fields_}


def abi_layout_json():
    return lib().xpuhip_abi_layout_json().decode()
