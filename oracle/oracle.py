"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
CPU-baseline leg of bench.py, to check results.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    lib = ctypes.CDLL(LIB_PATH)
    ll, i32, dbl, cp = ctypes.c_longlong, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.orc_isqrt.argtypes, lib.orc_isqrt.restype = [ll], ll
    lib.orc_mertens.argtypes, lib.orc_mertens.restype = [ll, i32, cp, cp, cp, cp, cp], ll
    lib.orc_mertens_compressed.argtypes, lib.orc_mertens_compressed.restype = [ll, i32], ll
    lib.orc_mertens_small.argtypes, lib.orc_mertens_small.restype = [ll], ll
    lib.orc_mertens_recurrence.argtypes, lib.orc_mertens_recurrence.restype = [ll, ll], ll
    lib.orc_time_mertens.argtypes, lib.orc_time_mertens.restype = [ll, i32, P(ll)], dbl
    lib.orc_plan.argtypes = [ll, cp, cp, cp, cp, cp, P(i32), P(ll), P(ll), P(ll), P(ll), P(dbl)]
    lib.orc_plan.restype = None
    lib.orc_mobius_eratosthenes.argtypes, lib.orc_mobius_eratosthenes.restype = [ll, vp], None
    lib.orc_mobius_bytelog_block.argtypes, lib.orc_mobius_bytelog_block.restype = [ll, ll, ll, vp], None
    lib.orc_mobius_bytelog_odd.argtypes, lib.orc_mobius_bytelog_odd.restype = [ll, ll, i32, vp], None
    lib.orc_finalize_byte.argtypes, lib.orc_finalize_byte.restype = [i32, ll], i32
    lib.orc_piece_closed.argtypes, lib.orc_piece_closed.restype = [i32, ll], ll
    lib.orc_piece_lut.argtypes, lib.orc_piece_lut.restype = [i32, ll], ll
    lib.orc_s2_interval.argtypes, lib.orc_s2_interval.restype = [i32, ll, ll, ll, vp, P(ll), P(ll)], None
    lib.orc_dense_table.argtypes, lib.orc_dense_table.restype = [ll, i32, ll, ll, vp, vp], None
    lib.orc_compressed_table.argtypes = [ll, ll, i32, ll, ll, vp, vp, P(ll)]
    lib.orc_compressed_table.restype = None
    lib.orc_g_single.argtypes, lib.orc_g_single.restype = [ll, dbl, ll, vp, ll, vp, P(ll), P(ll)], None
    lib.orc_g_comb2.argtypes, lib.orc_g_comb2.restype = [ll, dbl, ll, vp, ll, vp, P(ll), P(ll)], None
    lib.orc_g_brute.argtypes, lib.orc_g_brute.restype = [ll, ll, vp, ll, P(ll), P(ll)], None
    _lib = lib
    return lib


def _b(s):
    return None if s is None else s.encode()


def _i128(lo, hi):
    return (hi.value << 64) | (lo.value & ((1 << 64) - 1))


def mertens(X, threads=0, use_chmt=None, cx=None, chmt_u_cap=None, fx=None, u_cap=None):
    return load().orc_mertens(X, threads, _b(use_chmt), _b(cx), _b(chmt_u_cap), _b(fx), _b(u_cap))


def mertens_compressed(X, threads=0):
    return load().orc_mertens_compressed(X, threads)


def mertens_small(X):
    return load().orc_mertens_small(X)


def mertens_recurrence(X, base=1 << 20):
    return load().orc_mertens_recurrence(X, base)


def time_mertens(X, threads=0):
    out = ctypes.c_longlong()
    t = load().orc_time_mertens(X, threads, ctypes.byref(out))
    return out.value, t


def plan(X, use_chmt=None, cx=None, chmt_u_cap=None, fx=None, u_cap=None):
    s, u, ud, mu, N, c = (ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(),
                          ctypes.c_longlong(), ctypes.c_double())
    load().orc_plan(X, _b(use_chmt), _b(cx), _b(chmt_u_cap), _b(fx), _b(u_cap),
                    *(ctypes.byref(v) for v in (s, u, ud, mu, N, c)))
    return dict(streaming=bool(s.value), u=u.value, u_dense=ud.value, mu_limit=mu.value, N=N.value, cx=c.value)


def isqrt(n):
    return load().orc_isqrt(n)


def mobius(limit):
    out = np.empty(limit + 1, dtype=np.int8)
    load().orc_mobius_eratosthenes(limit, out.ctypes.data)
    return out


def mobius_bytelog_block(low, length, sieve_limit):
    out = np.empty(length, dtype=np.int8)
    load().orc_mobius_bytelog_block(low, length, sieve_limit, out.ctypes.data)
    return out


def mobius_bytelog_odd(half_u, u, threads=0):
    out = np.zeros(half_u + 1, dtype=np.int8)
    load().orc_mobius_bytelog_odd(half_u, u, threads, out.ctypes.data)
    return out


def finalize_byte(acc, n):
    return load().orc_finalize_byte(acc, n)


def piece_closed(piece, q):
    return load().orc_piece_closed(piece, q)


def piece_lut(piece, q):
    return load().orc_piece_lut(piece, q)


def s2_interval(piece, first, last, y, mu):
    lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
    mu = np.ascontiguousarray(mu, dtype=np.int8)
    load().orc_s2_interval(piece, first, last, y, mu.ctypes.data, ctypes.byref(lo), ctypes.byref(hi))
    return _i128(lo, hi)


def dense_table(limit, threads=0, max_mu=0, max_M=-1):
    mu_limit = min(max_mu, limit) if max_mu > 0 else limit
    m_limit = min(max_M, limit) if max_M >= 0 else limit
    mu = np.empty(mu_limit + 1, dtype=np.int8)
    M = np.empty(m_limit + 1, dtype=np.int16)
    load().orc_dense_table(limit, threads, max_mu, max_M, mu.ctypes.data, M.ctypes.data)
    return mu, M


def compressed_table(total, dense, threads=0, max_mu=0, query_limit=None):
    q = total if query_limit is None else query_limit
    M = np.empty(q + 1, dtype=np.int32)
    mu = np.empty(q + 1, dtype=np.int8)
    lim = ctypes.c_longlong()
    load().orc_compressed_table(total, dense, threads, max_mu, q, M.ctypes.data, mu.ctypes.data, ctypes.byref(lim))
    return M, mu, lim.value


def g_single(y, cx, u, M, mu):
    lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
    M = np.ascontiguousarray(M, dtype=np.int16)
    mu = np.ascontiguousarray(mu, dtype=np.int8)
    load().orc_g_single(y, cx, u, M.ctypes.data, len(M) - 1, mu.ctypes.data, ctypes.byref(lo), ctypes.byref(hi))
    return _i128(lo, hi)


def g_comb2(y, cx, u, M, mu):
    lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
    M = np.ascontiguousarray(M, dtype=np.int16)
    mu = np.ascontiguousarray(mu, dtype=np.int8)
    load().orc_g_comb2(y, cx, u, M.ctypes.data, len(M) - 1, mu.ctypes.data, ctypes.byref(lo), ctypes.byref(hi))
    return _i128(lo, hi)


def g_brute(y, u, prefix):
    lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
    prefix = np.ascontiguousarray(prefix, dtype=np.int32)
    load().orc_g_brute(y, u, prefix.ctypes.data, len(prefix) - 1, ctypes.byref(lo), ctypes.byref(hi))
    return _i128(lo, hi)
