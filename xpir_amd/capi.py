"""ctypes binding of libxpir_amd.so (the C ABI in include/xpir_amd.h).

Used by the tests and bench.py to drive the HIP path.  Loading fails loudly when the
library has not been built: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxpir_amd.so")

XPG_OK, XPG_EINVAL, XPG_ENOMEM, XPG_EDEVICE, XPG_ESIZE, XPG_EBUSY = 0, -1, -2, -3, -4, -5

# every entry point include/xpir_amd.h declares
EXPORTS = [
    "xpg_params_from_string", "xpg_polys_for_bytes", "xpg_reply_polys", "xpg_ctx_create",
    "xpg_ctx_destroy", "xpg_last_error", "xpg_ctx_stream", "xpg_ctx_set_option", "xpg_sync", "xpg_dev_alloc",
    "xpg_dev_free", "xpg_h2d", "xpg_d2h", "xpg_host_alloc", "xpg_host_free", "xpg_h2d_async",
    "xpg_db_import", "xpg_db_import_dev",
    "xpg_db_import_ntt_words", "xpg_db_export_ntt_words", "xpg_db_save", "xpg_db_load", "xpg_db_polys",
    "xpg_db_reply_polys", "xpg_db_levels",
    "xpg_db_query_words", "xpg_db_raw_bytes", "xpg_db_free", "xpg_reply", "xpg_reply_dev", "xpg_reply_dev_noalloc",
    "xpg_reply_batch_dev", "xpg_timing_enable", "xpg_timing_read", "xpg_timing_read_main", "xpg_timing_read_batched", "xpg_graph_stats", "xpg_autotune", "xpg_tune_set",
    "xpg_tune_get", "xpg_tune_clear", "xpg_keygen",
    "xpg_query_gen_dev", "xpg_random_bytes_dev", "xpg_decrypt_dev", "xpg_server_create", "xpg_server_destroy",
    "xpg_server_submit", "xpg_server_poll", "xpg_server_release", "xpg_server_now_us", "xpg_server_stats",
]


class XpgParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n", C.c_uint32), ("logn", C.c_uint32), ("m", C.c_uint32),
                ("bits", C.c_uint32), ("Berr", C.c_uint32), ("A", C.c_int32), ("n0", C.c_uint64),
                ("moduli", C.c_uint64 * 4), ("abs_bytes", C.c_uint64), ("ct_bytes", C.c_uint64)]


class XpgRequest(C.Structure):
    _fields_ = [("db", C.c_void_p), ("query_dev", C.c_void_p), ("reply_dev", C.c_void_p)]


class XpgServerConfig(C.Structure):
    _fields_ = [("reply_pool_bytes", C.c_uint64), ("workers", C.c_uint32), ("max_batch", C.c_uint32),
                ("heavy_steps", C.c_uint64)]


class XpgCompletion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("reply_dev", C.c_void_p), ("reply_words", C.c_uint64),
                ("submit_us", C.c_double), ("start_us", C.c_double), ("done_us", C.c_double),
                ("status", C.c_int)]


class XpgError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m xpir_amd.build` "
                          "(the HIP path has no fallback)")
    L = C.CDLL(LIB_PATH)
    vp, u64, u32 = C.c_void_p, C.c_uint64, C.c_uint32
    P = C.POINTER(XpgParams)
    sig = {
        "xpg_params_from_string": (C.c_int, [C.c_char_p, u64, P]),
        "xpg_polys_for_bytes": (u32, [P, u64]),
        "xpg_reply_polys": (u32, [P, u32, u32]),
        "xpg_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "xpg_ctx_destroy": (None, [vp]),
        "xpg_last_error": (C.c_char_p, [vp]),
        "xpg_ctx_stream": (vp, [vp]),
        "xpg_ctx_set_option": (C.c_int, [vp, C.c_char_p, C.c_int64]),
        "xpg_sync": (C.c_int, [vp]),
        "xpg_dev_alloc": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "xpg_dev_free": (C.c_int, [vp, vp]),
        "xpg_h2d": (C.c_int, [vp, vp, vp, u64]),
        "xpg_d2h": (C.c_int, [vp, vp, vp, u64]),
        "xpg_host_alloc": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "xpg_host_free": (C.c_int, [vp, vp]),
        "xpg_h2d_async": (C.c_int, [vp, vp, vp, u64]),
        "xpg_db_import": (C.c_int, [vp, P, vp, u64, u64, u32, u32, C.POINTER(u32), C.POINTER(vp)]),
        "xpg_db_import_dev": (C.c_int, [vp, P, vp, u64, u64, u32, u32, C.POINTER(u32), C.POINTER(vp)]),
        "xpg_db_import_ntt_words": (C.c_int, [vp, P, vp, u64, u32, u64, u32, C.POINTER(u32), C.POINTER(vp)]),
        "xpg_db_export_ntt_words": (C.c_int, [vp, vp, vp, u64]),
        "xpg_db_save": (C.c_int, [vp, vp, C.c_char_p, vp]),
        "xpg_db_load": (C.c_int, [vp, P, C.c_char_p, vp, C.POINTER(vp)]),
        "xpg_db_levels": (u32, [vp, C.POINTER(u32), u32]),
        "xpg_db_polys": (u32, [vp]),
        "xpg_db_reply_polys": (u32, [vp]),
        "xpg_db_query_words": (u64, [vp]),
        "xpg_db_raw_bytes": (u64, [vp]),
        "xpg_db_free": (None, [vp]),
        "xpg_reply": (C.c_int, [vp, vp, vp, u64, vp, u64]),
        "xpg_reply_dev": (C.c_int, [vp, vp, vp, vp]),
        "xpg_reply_dev_noalloc": (C.c_int, [vp, vp, vp, vp]),
        "xpg_reply_batch_dev": (C.c_int, [vp, C.POINTER(XpgRequest), u32]),
        "xpg_timing_enable": (C.c_int, [vp, C.c_int]),
        "xpg_timing_read": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(u64), C.c_int]),
        "xpg_timing_read_main": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(u64), C.c_int]),
        "xpg_timing_read_batched": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(u64), C.POINTER(u64), C.c_int]),
        "xpg_graph_stats": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
        "xpg_autotune": (C.c_int, [vp, vp, u32, vp, u32, C.POINTER(u32)]),
        "xpg_tune_set": (C.c_int, [vp, vp, u32]),
        "xpg_tune_get": (C.c_int, [vp, vp, u32, C.POINTER(u32)]),
        "xpg_tune_clear": (C.c_int, [vp]),
        "xpg_keygen": (C.c_int, [P, u64, vp]),
        "xpg_query_gen_dev": (C.c_int, [vp, P, vp, u32, C.POINTER(u32), u64, u64, vp]),
        "xpg_decrypt_dev": (C.c_int, [vp, P, vp, vp, u64, vp]),
        "xpg_random_bytes_dev": (C.c_int, [vp, u64, u64, vp, u64]),
        "xpg_server_create": (C.c_int, [C.c_int, C.POINTER(XpgServerConfig), C.POINTER(vp)]),
        "xpg_server_destroy": (None, [vp]),
        "xpg_server_submit": (C.c_int, [vp, vp, vp, C.POINTER(u64)]),
        "xpg_server_poll": (C.c_int, [vp, C.POINTER(XpgCompletion), u32, C.POINTER(u32), C.c_int64]),
        "xpg_server_release": (C.c_int, [vp, u64]),
        "xpg_server_now_us": (C.c_double, [vp]),
        "xpg_server_stats": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def params(desc: str, n0: int) -> XpgParams:
    prm = XpgParams()
    rc = lib().xpg_params_from_string(desc.encode(), n0, C.byref(prm))
    if rc != XPG_OK:
        raise XpgError(f"unsupported cryptosystem {desc!r} (n0={n0}): status {rc}")
    return prm


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class XpgTuneEntry(C.Structure):
    """One tuned MAC launch shape (include/xpir_amd.h, xpg_tune_entry)."""
    _fields_ = [("kind", C.c_uint32), ("n", C.c_uint32), ("m", C.c_uint32), ("P", C.c_uint32),
                ("n_l", C.c_uint64), ("groups", C.c_uint64), ("qn", C.c_uint32),
                ("kernel", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("c", C.c_int32),
                ("ms", C.c_float), ("default_ms", C.c_float)]

    def to_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        e = cls()
        for f, _ in cls._fields_:
            setattr(e, f, d.get(f, 0))
        return e


class DeviceBuffer:
    """Device allocation owned by a Context (freed explicitly or with the context)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx, self.nbytes = ctx, nbytes
        p = C.c_void_p()
        ctx._check(lib().xpg_dev_alloc(ctx.h, nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._check(lib().xpg_h2d(self.ctx.h, self.ptr, _ptr(arr), arr.nbytes))
        return self

    def download(self, dtype=np.uint64, count=None) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        count = self.nbytes // itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        self.ctx._check(lib().xpg_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().xpg_dev_free(self.ctx.h, self.ptr)
            self.ptr = None


class Database:
    def __init__(self, ctx: "Context", handle: int, prm: XpgParams, dims):
        self.ctx, self.h, self.prm, self.dims = ctx, handle, prm, list(dims)

    @property
    def polys(self) -> int:
        return lib().xpg_db_polys(self.h)

    @property
    def reply_polys(self) -> int:
        return lib().xpg_db_reply_polys(self.h)

    @property
    def query_words(self) -> int:
        return lib().xpg_db_query_words(self.h)

    @property
    def reply_words(self) -> int:
        return self.reply_polys * 2 * self.prm.m * self.prm.n

    @property
    def raw_bytes(self) -> int:
        return lib().xpg_db_raw_bytes(self.h)

    def export_words(self) -> np.ndarray:
        units = int(np.prod(self.dims, dtype=np.int64))
        out = np.empty(units * self.polys * self.prm.m * self.prm.n, dtype=np.uint64)
        self.ctx._check(lib().xpg_db_export_ntt_words(self.ctx.h, self.h, _ptr(out), out.size))
        return out

    def free(self):
        if self.h:
            lib().xpg_db_free(self.h)
            self.h = None


class Context:
    """One GPU, one HIP stream (xpg_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().xpg_ctx_create(device, C.byref(h))
        if rc != XPG_OK:
            raise XpgError(f"xpg_ctx_create({device}) failed: {rc}")
        self.h = h.value

    def _check(self, rc: int):
        if rc != XPG_OK:
            raise XpgError(f"status {rc}: {lib().xpg_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().xpg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self._check(lib().xpg_sync(self.h))

    def set_option(self, key: str, value: int):
        self._check(lib().xpg_ctx_set_option(self.h, key.encode(), value))

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def download_ptr(self, ptr: int, count: int, dtype=np.uint64) -> np.ndarray:
        """copy `count` items from a device pointer the library handed out"""
        out = np.empty(count, dtype=dtype)
        self._check(lib().xpg_d2h(self.h, _ptr(out), ptr, out.nbytes))
        return out

    @staticmethod
    def _dims(dims):
        arr = (C.c_uint32 * len(dims))(*dims)
        return arr

    def import_db(self, prm: XpgParams, elems: np.ndarray, n_elems: int, elem_bytes: int, dims,
                  alpha: int = 1) -> Database:
        elems = np.ascontiguousarray(elems, dtype=np.uint8)
        assert elems.size == n_elems * elem_bytes
        h = C.c_void_p()
        self._check(lib().xpg_db_import(self.h, C.byref(prm), _ptr(elems), n_elems, elem_bytes, alpha,
                                        len(dims), self._dims(dims), C.byref(h)))
        return Database(self, h.value, prm, dims)

    def import_db_dev(self, prm: XpgParams, elems_dev: DeviceBuffer, n_elems: int, elem_bytes: int, dims,
                      alpha: int = 1) -> Database:
        h = C.c_void_p()
        self._check(lib().xpg_db_import_dev(self.h, C.byref(prm), elems_dev.ptr, n_elems, elem_bytes, alpha,
                                            len(dims), self._dims(dims), C.byref(h)))
        return Database(self, h.value, prm, dims)

    def import_ntt_words(self, prm: XpgParams, words: np.ndarray, units: int, P: int, raw_elem_bytes: int,
                         dims) -> Database:
        words = np.ascontiguousarray(words, dtype=np.uint64)
        h = C.c_void_p()
        self._check(lib().xpg_db_import_ntt_words(self.h, C.byref(prm), _ptr(words), units, P, raw_elem_bytes,
                                                  len(dims), self._dims(dims), C.byref(h)))
        return Database(self, h.value, prm, dims)

    @staticmethod
    def _digest(digest):
        if digest is None:
            return None
        d = bytes(digest)
        if len(d) != 32:
            raise ValueError("digest is 32 bytes")
        return d  # passed as a pointer to its bytes

    def save_db(self, db: Database, path: str, digest: bytes = None):
        """write the NTT database to `path` (xpg_db_save); digest: 32 bytes naming the raw database"""
        self._check(lib().xpg_db_save(self.h, db.h, os.fsencode(path), self._digest(digest)))

    def load_db(self, prm: XpgParams, path: str, digest: bytes = None) -> Database:
        """a database from a file xpg_db_save wrote, checked against prm (and digest)"""
        h = C.c_void_p()
        self._check(lib().xpg_db_load(self.h, C.byref(prm), os.fsencode(path), self._digest(digest), C.byref(h)))
        dims = (C.c_uint32 * 16)()
        d = lib().xpg_db_levels(h.value, dims, 16)
        return Database(self, h.value, prm, list(dims[:d]))

    def reply(self, db: Database, query: np.ndarray) -> np.ndarray:
        query = np.ascontiguousarray(query, dtype=np.uint64)
        out = np.empty(db.reply_words, dtype=np.uint64)
        self._check(lib().xpg_reply(self.h, db.h, _ptr(query), query.size, _ptr(out), out.size))
        return out

    def reply_dev(self, db: Database, query_dev: DeviceBuffer, reply_dev: DeviceBuffer):
        self._check(lib().xpg_reply_dev(self.h, db.h, query_dev.ptr, reply_dev.ptr))

    def reply_dev_noalloc(self, db: Database, query_dev: DeviceBuffer, reply_dev: DeviceBuffer):
        """xpg_reply_dev_noalloc: the same reply with nothing allocated (scratch from the
        library's static device array); raises XpgError with XPG_ESIZE when it needs more."""
        self._check(lib().xpg_reply_dev_noalloc(self.h, db.h, query_dev.ptr, reply_dev.ptr))

    def reply_batch_dev(self, reqs):
        arr = (XpgRequest * len(reqs))()
        for i, (db, q, r) in enumerate(reqs):
            arr[i].db, arr[i].query_dev, arr[i].reply_dev = db.h, q.ptr, r.ptr
        self._check(lib().xpg_reply_batch_dev(self.h, arr, len(reqs)))

    def timing(self, on=True):
        """True/1: events around every MAC level and main kernel; 2: the main kernels
        alone; False/0: off"""
        self._check(lib().xpg_timing_enable(self.h, 2 if on == 2 else 1 if on else 0))

    def timing_read(self, reset: bool = True):
        ms, n = C.c_double(), C.c_uint64()
        self._check(lib().xpg_timing_read(self.h, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def timing_read_main(self, reset: bool = True):
        """(ms, launches) of the dominant MAC kernel alone, without its scan."""
        ms, n = C.c_double(), C.c_uint64()
        self._check(lib().xpg_timing_read_main(self.h, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def timing_read_batched(self, reset: bool = True):
        """(ms, launches, exact chain steps) of the register-blocked MAC kernel's launches."""
        ms, n, st = C.c_double(), C.c_uint64(), C.c_uint64()
        self._check(lib().xpg_timing_read_batched(self.h, C.byref(ms), C.byref(n), C.byref(st),
                                                  1 if reset else 0))
        return ms.value, n.value, st.value

    def autotune(self, db: "Database", max_batch: int = 0):
        """Time every MAC kernel candidate on db's level shapes (and batches of 2..max_batch
        queries on level 0), check their words agree, record the fastest per shape in the
        device's table; returns the recorded entries as dicts."""
        cap = 16 + 8
        arr = (XpgTuneEntry * cap)()
        n = C.c_uint32(0)
        self._check(lib().xpg_autotune(self.h, db.h, max_batch, arr, cap, C.byref(n)))
        return [arr[i].to_dict() for i in range(min(n.value, cap))]

    def tune_get(self):
        n = C.c_uint32(0)
        self._check(lib().xpg_tune_get(self.h, None, 0, C.byref(n)))
        arr = (XpgTuneEntry * max(1, n.value))()
        self._check(lib().xpg_tune_get(self.h, arr, n.value, C.byref(n)))
        return [arr[i].to_dict() for i in range(n.value)]

    def tune_set(self, entries):
        arr = (XpgTuneEntry * max(1, len(entries)))(*[XpgTuneEntry.from_dict(e) for e in entries])
        self._check(lib().xpg_tune_set(self.h, arr, len(entries)))

    def tune_clear(self):
        self._check(lib().xpg_tune_clear(self.h))

    def tune_save(self, path: str):
        """Write the device's tuning table as JSON (a list of entries)."""
        import json
        with open(path, "w") as f:
            json.dump(self.tune_get(), f, indent=1)

    def tune_load(self, path: str):
        """Install a table written by tune_save."""
        import json
        with open(path) as f:
            self.tune_set(json.load(f))

    def graph_stats(self):
        """(replies captured into a HIP graph, replies replayed from one)"""
        c, r = C.c_uint64(), C.c_uint64()
        self._check(lib().xpg_graph_stats(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def query_gen_dev(self, prm: XpgParams, sk_dev: DeviceBuffer, dims, index: int, noise_seed: int,
                      out: DeviceBuffer):
        self._check(lib().xpg_query_gen_dev(self.h, C.byref(prm), sk_dev.ptr, len(dims), self._dims(dims),
                                            index, noise_seed, out.ptr))

    def random_bytes_dev(self, seed: int, offset: int, dst: DeviceBuffer, nbytes: int):
        self._check(lib().xpg_random_bytes_dev(self.h, seed, offset, dst.ptr, nbytes))

    def decrypt_dev(self, prm: XpgParams, sk_dev: DeviceBuffer, ct_ptr: int, n_ct: int, out: DeviceBuffer):
        """Decrypt n_ct ciphertexts at device address ct_ptr into out (n_ct * A*n/8 bytes)."""
        assert out.nbytes >= n_ct * prm.A * prm.n // 8
        self._check(lib().xpg_decrypt_dev(self.h, C.byref(prm), sk_dev.ptr, ct_ptr, n_ct, out.ptr))

    def extract_reply_dev(self, prm: XpgParams, sk_dev: DeviceBuffer, reply_ptr: int, dims, elem_bytes: int) -> bytes:
        """The element bytes of a d-level reply at device address reply_ptr, decrypted on the
        GPU level by level: level-l plaintexts are the bytes of the level-(l-1) ciphertexts
        (PIRReplyExtraction_internal.cpp:41-106)."""
        ct_bytes = 16 * prm.m * prm.n
        abs_bytes = prm.A * prm.n // 8
        counts = [-(-elem_bytes // abs_bytes)]
        for _ in range(1, len(dims)):
            counts.append(-(-counts[-1] * ct_bytes // abs_bytes))
        cur, held = reply_ptr, None
        try:
            for lvl in range(len(dims) - 1, -1, -1):
                out = self.alloc(counts[lvl] * abs_bytes)
                self.decrypt_dev(prm, sk_dev, cur, counts[lvl], out)
                if held is not None:
                    self.sync()
                    held.free()
                held, cur = out, out.ptr
            return held.download(np.uint8)[:elem_bytes].tobytes()
        finally:
            if held is not None:
                held.free()


def keygen(prm: XpgParams, seed: int) -> np.ndarray:
    sk = np.empty(prm.m * prm.n, dtype=np.uint64)
    rc = lib().xpg_keygen(C.byref(prm), seed, _ptr(sk))
    if rc != XPG_OK:
        raise XpgError("keygen")
    return sk


class Server:
    """Continuous-batching reply server (xpg_server_*)."""

    def __init__(self, device: int = 0, reply_pool_bytes: int = 1 << 30, workers: int = 2, max_batch: int = 64,
                 heavy_steps: int = 0):
        cfg = XpgServerConfig(reply_pool_bytes, workers, max_batch, heavy_steps)
        h = C.c_void_p()
        rc = lib().xpg_server_create(device, C.byref(cfg), C.byref(h))
        if rc != XPG_OK:
            raise XpgError(f"xpg_server_create failed: {rc}")
        self.h = h.value
        self._buf = (XpgCompletion * 256)()

    def submit(self, db: "Database", query_dev: "DeviceBuffer"):
        """ticket, or None when admission is refused (reply memory full)"""
        t = C.c_uint64()
        rc = lib().xpg_server_submit(self.h, db.h, query_dev.ptr, C.byref(t))
        if rc == XPG_EBUSY:
            return None
        if rc != XPG_OK:
            raise XpgError(f"submit failed: {rc}")
        return t.value

    def poll(self, timeout_us: int = 0):
        n = C.c_uint32()
        rc = lib().xpg_server_poll(self.h, self._buf, 256, C.byref(n), timeout_us)
        if rc != XPG_OK:
            raise XpgError(f"poll failed: {rc}")
        return [XpgCompletion.from_buffer_copy(self._buf[i]) for i in range(n.value)]

    def release(self, ticket: int):
        rc = lib().xpg_server_release(self.h, ticket)
        if rc != XPG_OK:
            raise XpgError(f"release: status {rc}")

    def now_us(self) -> float:
        return lib().xpg_server_now_us(self.h)

    def stats(self):
        b, r, u = C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib().xpg_server_stats(self.h, C.byref(b), C.byref(r), C.byref(u))
        return {"batches": b.value, "batched_requests": r.value, "budget_used": u.value}

    def close(self):
        if self.h:
            lib().xpg_server_destroy(self.h)
            self.h = None
