"""ctypes view of the checker libraries (TEST INFRASTRUCTURE ONLY).

``Oracle`` wraps liboracle.so, the CPU restatement of XPIR's NFLLWE reply path
(xpir_oracle.c).  ``RefNFL`` wraps oracle/_ref/libxpir_ref.so, the reference's own
NFLlib compiled from its sources by oracle/Makefile (absent where the reference
tree is not available; tests that need it skip).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this.

Parity unpinned: query generation (``Oracle.make_query`` over xo_encrypt_bit), which
is checked only by decryption round trips; see the header of xpir_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libxpir_ref.so")

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class XoParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n", C.c_uint32), ("logn", C.c_uint32),
                ("m", C.c_uint32), ("bits", C.c_uint32), ("Berr", C.c_uint32),
                ("A", C.c_int32), ("p", C.c_uint64 * 4)]


class XoTables(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32)] + [
        (nm, C.POINTER(C.c_uint64)) for nm in ("phi", "phiq", "w", "wq", "iw", "iwq", "ipt", "iptq")]


def build():
    """Compile liboracle.so (and the reference build when the reference tree exists)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir("/root/reference/crypto"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _load(path):
    return C.CDLL(path)


class Oracle:
    """Parameters + tables for one cryptosystem string and n0, with numpy helpers."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(LIB):
                build()
            L = _load(LIB)
            L.xo_params_init.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(XoParams)]
            L.xo_tables_init.argtypes = [C.POINTER(XoParams), C.POINTER(XoTables)]
            L.xo_tables_free.argtypes = [C.POINTER(XoTables)]
            for f in ("xo_mulmod", "xo_shoup_quot"):
                getattr(L, f).restype = C.c_uint64
            L.xo_mulmod.argtypes = [C.c_uint64] * 3
            L.xo_shoup_quot.argtypes = [C.c_uint64] * 2
            L.xo_mulmod_shoup.restype = C.c_uint64
            L.xo_mulmod_shoup.argtypes = [C.c_uint64] * 4
            L.xo_mac_step.restype = C.c_uint64
            L.xo_mac_step.argtypes = [C.c_uint64] * 5
            L.xo_modulus.restype = C.c_uint64
            L.xo_modulus.argtypes = [C.c_uint32]
            L.xo_root.restype = C.c_uint64
            L.xo_root.argtypes = [C.c_uint32]
            L.xo_inv8192.restype = C.c_uint64
            L.xo_inv8192.argtypes = [C.c_uint32]
            L.xo_ntt.argtypes = [u64p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_uint32, C.c_uint64]
            L.xo_ntt_powphi.argtypes = [C.POINTER(XoTables), C.POINTER(XoParams), u64p]
            L.xo_invntt_powinvphi.argtypes = [C.POINTER(XoTables), C.POINTER(XoParams), u64p]
            L.xo_absorb_polys.restype = C.c_uint64
            L.xo_absorb_polys.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
            L.xo_absorb.argtypes = [u8p, C.c_uint64, C.POINTER(XoParams), u64p]
            L.xo_import.argtypes = [C.POINTER(XoTables), C.POINTER(XoParams), u8p, C.c_uint64, u64p]
            L.xo_reply_level.argtypes = [C.POINTER(XoParams), u64p, u64p, C.c_uint64, C.c_uint64,
                                         C.c_uint32, u64p, C.c_int]
            L.xo_reply_level_chains.argtypes = L.xo_reply_level.argtypes + [C.POINTER(C.c_double)]
            L.xo_recursion.restype = C.c_uint32
            L.xo_recursion.argtypes = [C.POINTER(XoTables), C.POINTER(XoParams), u64p, C.c_uint64,
                                       C.c_uint32, u64p]
            L.xo_next_polys.restype = C.c_uint32
            L.xo_next_polys.argtypes = [C.POINTER(XoParams), C.c_uint32]
            L.xo_rand_word.restype = C.c_uint64
            L.xo_rand_word.argtypes = [C.c_uint64, C.c_uint64]
            L.xo_rand_bytes.argtypes = [C.c_uint64, C.c_uint64, u8p, C.c_uint64]
            L.xo_keygen.argtypes = [C.POINTER(XoParams), C.c_uint64, u64p]
            L.xo_encrypt_bit.argtypes = [C.POINTER(XoTables), C.POINTER(XoParams), u64p, C.c_uint32,
                                         C.c_uint64, C.c_uint64, u64p]
            L.xo_decrypt.restype = C.c_int
            L.xo_decrypt.argtypes = [C.POINTER(XoTables), C.POINTER(XoParams), u64p, u64p, u8p]
            cls._lib = L
        return cls._lib

    def __init__(self, desc: str, n0: int):
        L = self.lib()
        self.prm = XoParams()
        rc = L.xo_params_init(desc.encode(), n0, C.byref(self.prm))
        if rc != 0:
            raise ValueError(f"bad cryptosystem {desc!r} (n0={n0}): {rc}")
        self.tab = XoTables()
        if L.xo_tables_init(C.byref(self.prm), C.byref(self.tab)) != 0:
            raise MemoryError("tables")
        self.n, self.m, self.A = self.prm.n, self.prm.m, self.prm.A
        self.p = [self.prm.p[i] for i in range(self.m)]
        self.mn = self.n * self.m

    def __del__(self):
        try:
            self.lib().xo_tables_free(C.byref(self.tab))
        except Exception:
            pass

    # --- sizes ---
    def polys(self, nbytes: int) -> int:
        return int(self.lib().xo_absorb_polys(nbytes, self.A, self.n))

    def next_polys(self, P: int) -> int:
        return int(self.lib().xo_next_polys(C.byref(self.prm), P))

    # --- database staging ---
    @staticmethod
    def aggregate_units(data: np.ndarray, n_elems: int, elem_bytes: int, alpha: int,
                        zero_pad: bool = False) -> np.ndarray:
        """Import units of alpha elements each (rows of alpha * elem_bytes bytes), as
        importDataNFL (PIRReplyGeneratorNFL_internal.cpp:86-105) fills them with
        DBHandler::readAggregatedStream (DBHandler.cpp:3-26): one buffer reused for every
        unit; a partial last unit (r = n_elems mod alpha elements) reads its r elements
        into slots 0..r-1, then zeroes alpha - r slots starting at slot r-1 (the last real
        element's), so slot alpha-1 keeps the previous unit's last element (zero when
        there is no previous unit).  zero_pad: the last unit's missing slots are zero
        and its elements kept instead."""
        data = np.ascontiguousarray(data, dtype=np.uint8)[: n_elems * elem_bytes]
        units = -(-n_elems // alpha)
        out = np.zeros((units, alpha * elem_bytes), dtype=np.uint8)
        out.reshape(-1)[: data.size] = data
        r = n_elems % alpha
        if r and not zero_pad:
            last = out[units - 1]
            last[(r - 1) * elem_bytes:] = 0
            if units > 1:
                last[(alpha - 1) * elem_bytes:] = out[units - 2][(alpha - 1) * elem_bytes:]
        return out

    # --- kernels ---
    def absorb(self, data: np.ndarray) -> np.ndarray:
        data = np.ascontiguousarray(data, dtype=np.uint8)
        out = np.zeros(self.polys(data.size) * self.mn, dtype=np.uint64)
        self.lib().xo_absorb(data, data.size, C.byref(self.prm), out)
        return out

    def import_element(self, data: np.ndarray) -> np.ndarray:
        data = np.ascontiguousarray(data, dtype=np.uint8)
        out = np.zeros(self.polys(data.size) * self.mn, dtype=np.uint64)
        self.lib().xo_import(C.byref(self.tab), C.byref(self.prm), data, data.size, out)
        return out

    def ntt_powphi(self, poly: np.ndarray) -> np.ndarray:
        x = np.array(poly, dtype=np.uint64, copy=True)
        self.lib().xo_ntt_powphi(C.byref(self.tab), C.byref(self.prm), x)
        return x

    def invntt(self, poly: np.ndarray) -> np.ndarray:
        x = np.array(poly, dtype=np.uint64, copy=True)
        self.lib().xo_invntt_powinvphi(C.byref(self.tab), C.byref(self.prm), x)
        return x

    def tables(self, name: str) -> np.ndarray:
        ptr = getattr(self.tab, name)
        return np.ctypeslib.as_array(ptr, shape=(self.mn,)).copy()

    def reply_level(self, db: np.ndarray, q: np.ndarray, n_l: int, groups: int, P: int,
                    threads: int = 0) -> np.ndarray:
        db = np.ascontiguousarray(db, dtype=np.uint64)
        q = np.ascontiguousarray(q, dtype=np.uint64)
        assert db.size == groups * n_l * P * self.mn and q.size == n_l * 2 * self.mn
        out = np.zeros(groups * P * 2 * self.mn, dtype=np.uint64)
        self.lib().xo_reply_level(C.byref(self.prm), db, q, n_l, groups, P, out, threads)
        return out

    def reply_level_chains(self, db: np.ndarray, q: np.ndarray, n_l: int, groups: int, P: int,
                           threads: int = 0) -> np.ndarray:
        """reply_level with the host threads split over slot blocks (all cores busy at P = 1)."""
        return self.reply_level_chains_timed(db, q, n_l, groups, P, threads)[0]

    def reply_level_chains_timed(self, db, q, n_l, groups, P, threads=0):
        """(reply words, seconds spent in the MAC loops)"""
        db = np.ascontiguousarray(db, dtype=np.uint64)
        q = np.ascontiguousarray(q, dtype=np.uint64)
        assert db.size == groups * n_l * P * self.mn and q.size == n_l * 2 * self.mn
        out = np.zeros(groups * P * 2 * self.mn, dtype=np.uint64)
        secs = C.c_double()
        self.lib().xo_reply_level_chains(C.byref(self.prm), db, q, n_l, groups, P, out, threads,
                                         C.byref(secs))
        return out, secs.value

    def recursion(self, reply: np.ndarray, groups: int, P: int) -> tuple[np.ndarray, int]:
        reply = np.ascontiguousarray(reply, dtype=np.uint64)
        P2 = self.next_polys(P)
        out = np.zeros(groups * P2 * self.mn, dtype=np.uint64)
        self.lib().xo_recursion(C.byref(self.tab), C.byref(self.prm), reply, groups, P, out)
        return out, P2

    # --- seeded stream + client side ---
    def rand_bytes(self, seed: int, offset: int, nbytes: int) -> np.ndarray:
        out = np.zeros(nbytes, dtype=np.uint8)
        self.lib().xo_rand_bytes(seed, offset, out, nbytes)
        return out

    def keygen(self, seed: int) -> np.ndarray:
        sk = np.zeros(self.mn, dtype=np.uint64)
        self.lib().xo_keygen(C.byref(self.prm), seed, sk)
        return sk

    def encrypt_bit(self, sk: np.ndarray, bit: int, noise_seed: int, elem: int) -> np.ndarray:
        out = np.zeros(2 * self.mn, dtype=np.uint64)
        self.lib().xo_encrypt_bit(C.byref(self.tab), C.byref(self.prm), sk, bit, noise_seed, elem, out)
        return out

    def decrypt(self, sk: np.ndarray, ct: np.ndarray) -> np.ndarray:
        out = np.zeros(self.A * self.n // 8, dtype=np.uint8)
        rc = self.lib().xo_decrypt(C.byref(self.tab), C.byref(self.prm), sk,
                                   np.ascontiguousarray(ct, dtype=np.uint64), out)
        if rc != 0:
            raise ValueError("decrypt")
        return out

    # --- whole query / reply (test helpers) ---
    def make_query(self, sk, dims, index, noise_seed):
        """Query for `index` over the grid `dims`: per level, dims[l] encryptions of 0/1
        (PIRQueryGenerator_internal.cpp:37-70).  Element numbering is global."""
        coords, x = [], index
        for nl in dims:
            coords.append(x % nl)
            x //= nl
        levels, elem = [], 0
        for lvl, nl in enumerate(dims):
            q = np.zeros((nl, 2 * self.mn), dtype=np.uint64)
            for j in range(nl):
                q[j] = self.encrypt_bit(sk, 1 if j == coords[lvl] else 0, noise_seed, elem)
                elem += 1
            levels.append(q.reshape(-1))
        return levels

    def generate_reply(self, db_words: np.ndarray, P0: int, dims, query_levels, threads: int = 0):
        """Full d-level reply (PIRReplyGeneratorNFL_internal.cpp:340-445)."""
        data, P = db_words, P0
        for lvl, nl in enumerate(dims):
            groups = int(np.prod(dims[lvl + 1:], dtype=np.int64)) if lvl + 1 < len(dims) else 1
            rep = self.reply_level(data, query_levels[lvl], nl, groups, P, threads)
            if lvl + 1 < len(dims):
                data, P = self.recursion(rep, groups, P)
            else:
                return rep, P


def extract_reply(orc: Oracle, sk, reply: np.ndarray, dims, elem_bytes: int) -> bytes:
    """Decrypt a d-level reply back to the element bytes.  Level-d ciphertexts decrypt
    to the bytes of level-(d-1) ciphertexts, and so on down to the element."""
    ct_words = 2 * orc.mn
    ct_bytes = ct_words * 8
    abs_bytes = orc.A * orc.n // 8
    d = len(dims)
    # ciphertext count per level, bottom up (computeReplySizeInChunks)
    counts = [-(-elem_bytes // abs_bytes)]
    for _ in range(1, d):
        counts.append(-(-counts[-1] * ct_bytes // abs_bytes))
    data = reply.reshape(-1, ct_words)
    for lvl in range(d - 1, -1, -1):
        assert data.shape[0] == counts[lvl], (data.shape, counts, lvl)
        plain = b"".join(orc.decrypt(sk, c).tobytes() for c in data)
        if lvl == 0:
            return plain[:elem_bytes]
        need = counts[lvl - 1] * ct_bytes
        data = np.frombuffer(plain[:need], dtype=np.uint64).reshape(-1, ct_words)
    raise AssertionError("unreachable")


class RefNFL:
    """The reference's own NFLlib (oracle/_ref/libxpir_ref.so) for one (n, bits)."""

    _lib = None

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(REF_LIB)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = _load(REF_LIB)
            L.xr_create.restype = C.c_void_p
            L.xr_create.argtypes = [C.c_uint, C.c_uint]
            L.xr_destroy.argtypes = [C.c_void_p]
            L.xr_import.restype = C.c_int64
            L.xr_import.argtypes = [C.c_void_p, C.c_uint, u8p, C.c_uint64, C.c_uint64, u64p, C.c_uint64]
            L.xr_ntt_powphi.argtypes = [C.c_void_p, u64p]
            L.xr_invntt_powinvphi.argtypes = [C.c_void_p, u64p]
            L.xr_mulandadd_shoup.restype = C.c_uint64
            L.xr_mulandadd_shoup.argtypes = [C.c_uint64] * 5
            L.xr_mulmod_shoup.restype = C.c_uint64
            L.xr_mulmod_shoup.argtypes = [C.c_uint64] * 4
            L.xr_modulus.restype = C.c_uint64
            L.xr_modulus.argtypes = [C.c_void_p, C.c_uint]
            L.xr_reply_level.argtypes = [C.c_void_p, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint,
                                         u64p, C.c_int]
            L.xr_reply_level_timed.argtypes = [C.c_void_p, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint,
                                               u64p, C.c_int, C.POINTER(C.c_double)]
            L.xr_aggregate.restype = C.c_int64
            L.xr_aggregate.argtypes = [u8p, u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, u8p,
                                       C.c_uint64]
            L.xr_import_range.restype = C.c_int
            L.xr_import_range.argtypes = [C.c_void_p, C.c_uint, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                          C.c_uint64, C.c_uint, u64p, C.c_int]
            L.xr_level0_stream.restype = C.c_int
            L.xr_level0_stream.argtypes = [C.c_void_p, C.c_uint, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                           C.c_uint64, C.c_uint64, C.c_uint, u64p, u64p, C.c_int]
            cls._lib = L
        return cls._lib

    def __init__(self, n: int, bits: int):
        self.h = self.lib().xr_create(n, bits)
        self.n, self.m = n, bits // 60
        self.mn = n * self.m

    def __del__(self):
        try:
            self.lib().xr_destroy(self.h)
        except Exception:
            pass

    def import_buffers(self, data: np.ndarray, nbufs: int, A: int) -> np.ndarray:
        data = np.ascontiguousarray(data, dtype=np.uint8)
        assert data.size % nbufs == 0
        per = data.size // nbufs
        cap = (-(-data.size * 8 // (A * self.n)) + 1) * self.mn
        out = np.zeros(cap, dtype=np.uint64)
        P = self.lib().xr_import(self.h, A, data, nbufs, per, out, cap)
        assert P >= 0
        return out[:P * self.mn]

    def ntt_powphi(self, poly):
        x = np.array(poly, dtype=np.uint64, copy=True)
        self.lib().xr_ntt_powphi(self.h, x)
        return x

    @classmethod
    def aggregate(cls, elements, alpha: int, offset: int = 0, bytes_per_file=None) -> np.ndarray:
        """The reference's alpha aggregation (DBHandler::readAggregatedStream over a
        DBVectorProcessor, in importDataNFL's loop with one reused buffer): one row of
        alpha * fileByteSize bytes per aggregate."""
        sizes = np.array([len(e) for e in elements], dtype=np.uint64)
        offs = np.zeros(len(elements), dtype=np.uint64)
        offs[1:] = np.cumsum(sizes)[:-1]
        data = np.frombuffer(b"".join(bytes(e) for e in elements) + b"\0", dtype=np.uint8).copy()
        maxb = int(sizes.max())
        fb = min(maxb if bytes_per_file is None else bytes_per_file, maxb - offset)
        runs = -(-len(elements) // alpha)
        out = np.zeros(runs * alpha * fb + 1, dtype=np.uint8)
        got = cls.lib().xr_aggregate(data, offs, sizes, len(elements), alpha, offset,
                                     maxb if bytes_per_file is None else bytes_per_file, out, out.size)
        assert got == runs, (got, runs)
        return out[:runs * alpha * fb].reshape(runs, alpha * fb)

    def import_range(self, A: int, db_seed: int, n_elems: int, elem_bytes: int, first: int, count: int,
                     P0: int, threads: int = 8) -> np.ndarray:
        """NTT words of elements [first, first + count) of the seeded database (zero polys
        past n_elems), the reference's deserializeDataNFL per element."""
        out = np.zeros(count * P0 * self.mn, dtype=np.uint64)
        rc = self.lib().xr_import_range(self.h, A, db_seed, n_elems, elem_bytes, first, count, P0, out, threads)
        assert rc == 0, "an element imported to another poly count"
        return out

    def level0_stream(self, A: int, db_seed: int, n_elems: int, elem_bytes: int, n0: int, g0: int, g1: int,
                      P0: int, q0: np.ndarray, threads: int = 8) -> np.ndarray:
        """Level-0 replies of groups [g0, g1) of a d >= 2 grid, the groups' elements made
        from the seeded byte stream and imported one group at a time (the reference's
        deserializeDataNFL + mulandaddShoup), so the whole NTT database is never held."""
        q0 = np.ascontiguousarray(q0, dtype=np.uint64)
        assert q0.size == n0 * 2 * self.mn
        out = np.zeros((g1 - g0) * P0 * 2 * self.mn, dtype=np.uint64)
        rc = self.lib().xr_level0_stream(self.h, A, db_seed, n_elems, elem_bytes, n0, g0, g1, P0, q0, out, threads)
        assert rc == 0, "an element imported to another poly count"
        return out

    def invntt(self, poly):
        x = np.array(poly, dtype=np.uint64, copy=True)
        self.lib().xr_invntt_powinvphi(self.h, x)
        return x

    def reply_level(self, db, q, n_l, groups, P, threads=1):
        return self.reply_level_timed(db, q, n_l, groups, P, threads)[0]

    def reply_level_timed(self, db, q, n_l, groups, P, threads=1):
        """(reply words, seconds spent in the MAC loops)"""
        db = np.ascontiguousarray(db, dtype=np.uint64)
        q = np.ascontiguousarray(q, dtype=np.uint64)
        out = np.zeros(groups * P * 2 * self.mn, dtype=np.uint64)
        secs = C.c_double()
        self.lib().xr_reply_level_timed(self.h, db, q, n_l, groups, P, out, threads, C.byref(secs))
        return out, secs.value
