"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the benchmark's
CPU-baseline leg, as the checker.  The product never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Item(ctypes.Structure):
    """One worktodo item (Types.hpp:14-26)."""
    _fields_ = [("b", ctypes.c_int64), ("c", ctypes.c_int64), ("P0", ctypes.c_uint64),
                ("P1", ctypes.c_uint64), ("N0", ctypes.c_int64), ("N1", ctypes.c_int64),
                ("cls_mod", ctypes.c_uint64), ("cls_res", ctypes.c_uint64)]


class Relation(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint64), ("M", ctypes.c_uint64), ("q", ctypes.c_uint64)]


class Entry(ctypes.Structure):
    _fields_ = [("jcount", ctypes.c_uint64), ("m", ctypes.c_uint64), ("stride", ctypes.c_uint32),
                ("parity", ctypes.c_uint32)]


class EntryQ(ctypes.Structure):
    _fields_ = [("qp", ctypes.c_uint32), ("ncls", ctypes.c_uint32), ("mask", ctypes.c_uint32),
                ("s", ctypes.c_uint32), ("m", ctypes.c_uint64), ("Gc", ctypes.c_uint64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, i32, sz = ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
        P = ctypes.POINTER
        L.or_mont_mul.restype = u64
        L.or_mont_mul.argtypes = [u64, u64, u64]
        L.or_powmod.restype = u64
        L.or_powmod.argtypes = [u64, u64, u64]
        L.or_is_prime64.restype = i32
        L.or_is_prime64.argtypes = [u64]
        L.or_multiplicative_order.restype = u64
        L.or_multiplicative_order.argtypes = [u64, u64]
        L.or_jacobi.restype = i32
        L.or_jacobi.argtypes = [u64, u64]
        L.or_run_bsgs.restype = i32
        L.or_run_bsgs.argtypes = [u64, P(Item), P(Relation)]
        L.or_sieve_primes.restype = sz
        L.or_sieve_primes.argtypes = [u64, u64, ctypes.c_void_p, sz]
        L.or_sieve_survivors.restype = sz
        L.or_sieve_survivors.argtypes = [u64, u64, ctypes.c_void_p, sz]
        L.or_gpu_classes.restype = i32
        L.or_gpu_classes.argtypes = [P(Item), P(u64), ctypes.c_void_p, P(ctypes.c_uint32)]
        L.or_prep.restype = i32
        L.or_prep.argtypes = [u64, P(Item), i32, P(Entry)]
        L.or_prep_q.restype = i32
        L.or_prep_q.argtypes = [u64, P(Item), u64, i32, P(EntryQ)]
        L.or_process_range.restype = sz
        L.or_process_range.argtypes = [P(Item), u64, u64, u64, i32, ctypes.c_void_p, sz, P(u64)]
        L.or_results_json.restype = sz
        L.or_results_json.argtypes = [P(Item), ctypes.c_void_p, sz, ctypes.c_char_p, ctypes.c_char_p, sz]
        _lib = L
    return _lib


def make_item(b, c, P0, P1, N0, N1, mod=1, res=0):
    return Item(b, c, P0, P1, N0, N1, mod, res)


def mont_mul(a, b, n):
    return lib().or_mont_mul(a, b, n)


def powmod(a, e, n):
    return lib().or_powmod(a, e, n)


def is_prime64(n):
    return bool(lib().or_is_prime64(n))


def multiplicative_order(b, p):
    return lib().or_multiplicative_order(b, p)


def jacobi(a, n):
    return lib().or_jacobi(a, n)


def run_bsgs(p, item):
    r = Relation()
    if lib().or_run_bsgs(p, ctypes.byref(item), ctypes.byref(r)):
        return (r.p, r.M, r.q)
    return None


def sieve_primes(lo, hi):
    n = lib().or_sieve_primes(lo, hi, None, 0)
    out = np.empty(n, dtype=np.uint64)
    if n:
        lib().or_sieve_primes(lo, hi, out.ctypes.data, n)
    return out


def sieve_survivors(lo, hi):
    n = lib().or_sieve_survivors(lo, hi, None, 0)
    out = np.empty(n, dtype=np.uint64)
    if n:
        lib().or_sieve_survivors(lo, hi, out.ctypes.data, n)
    return out


def gpu_classes(item):
    Q = ctypes.c_uint64()
    cls = (ctypes.c_uint32 * 64)()
    n = ctypes.c_uint32()
    ok = lib().or_gpu_classes(ctypes.byref(item), ctypes.byref(Q), cls, ctypes.byref(n))
    return bool(ok), Q.value, list(cls[: n.value])


def prep(x, item, need_prime=True):
    e = Entry()
    r = lib().or_prep(x, ctypes.byref(item), int(need_prime), ctypes.byref(e))
    return r, e


def prep_q(x, item, Qmax, need_prime=True):
    e = EntryQ()
    r = lib().or_prep_q(x, ctypes.byref(item), Qmax, int(need_prime), ctypes.byref(e))
    return r, e


def process_range(item, lo, hi, chunk_ints=1 << 26, nthreads=1, cap=1 << 20):
    out = (Relation * cap)()
    seen = ctypes.c_uint64()
    n = lib().or_process_range(ctypes.byref(item), lo, hi, chunk_ints, nthreads, out, cap, ctypes.byref(seen))
    if n == ctypes.c_size_t(-1).value:
        raise RuntimeError("relation buffer too small")
    return [(out[i].p, out[i].M, out[i].q) for i in range(n)], seen.value


def results_json(item, rels, timestamp="2026-01-01 00:00:00 UTC+0"):
    arr = (Relation * max(1, len(rels)))()
    for i, (p, M, q) in enumerate(rels):
        arr[i] = Relation(p, M, q)
    ts = timestamp.encode()
    need = lib().or_results_json(ctypes.byref(item), arr, len(rels), ts, None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    lib().or_results_json(ctypes.byref(item), arr, len(rels), ts, buf, need + 1)
    return buf.value.decode()
