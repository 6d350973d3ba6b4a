"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32, i32, u64, dbl = ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        vp = ctypes.c_void_p
        L.or_nck.argtypes, L.or_nck.restype = [u32, u32], u32
        L.or_count.argtypes, L.or_count.restype = [u32, i32, u32, i32], u32
        L.or_fill_cache.argtypes, L.or_fill_cache.restype = [i32, i32, vp], i32
        L.or_rep.argtypes, L.or_rep.restype = [u32, i32], u32
        L.or_orbit_size.argtypes, L.or_orbit_size.restype = [u32, i32], i32
        L.or_orbit_list.argtypes, L.or_orbit_list.restype = [i32, vp], u64
        L.or_ctx_create.argtypes, L.or_ctx_create.restype = [i32, i32], vp
        L.or_ctx_destroy.argtypes, L.or_ctx_destroy.restype = [vp], None
        L.or_ctx_num_orbits.argtypes, L.or_ctx_num_orbits.restype = [vp], u64
        L.or_ctx_orbits.argtypes, L.or_ctx_orbits.restype = [vp, vp], None
        L.or_ctx_set_q.argtypes, L.or_ctx_set_q.restype = [vp, vp], None
        L.or_ctx_get_q.argtypes, L.or_ctx_get_q.restype = [vp, vp], None
        L.or_ctx_iterate.argtypes, L.or_ctx_iterate.restype = [vp], dbl
        L.or_ctx_log_den.argtypes, L.or_ctx_log_den.restype = [vp, vp], None
        L.or_ctx_kl.argtypes, L.or_ctx_kl.restype = [vp, vp], None
        L.or_ctx_rate.argtypes, L.or_ctx_rate.restype = [vp], dbl
        L.or_ctx_max_kl.argtypes, L.or_ctx_max_kl.restype = [vp], dbl
        L.or_ctx_sample_pass.argtypes, L.or_ctx_sample_pass.restype = [vp, u64, u64], dbl
        L.or_full_output_law.argtypes, L.or_full_output_law.restype = [i32, i32, vp, vp], None
        L.or_full_kl.argtypes, L.or_full_kl.restype = [i32, i32, vp, vp], None
        L.or_full_dual.argtypes, L.or_full_dual.restype = [i32, i32, vp, dbl, vp], dbl
        L.or_num_threads.argtypes, L.or_num_threads.restype = [], i32
        L.or_entropy_x.argtypes, L.or_entropy_x.restype = [i32, i32, u32], dbl
        L.or_kl_x.argtypes, L.or_kl_x.restype = [i32, i32, u32, vp], dbl
        L.or_sparse_output_law.argtypes, L.or_sparse_output_law.restype = [i32, i32, u64, vp, vp, vp], None
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nck(n, k):
    return lib().or_nck(n, k)


def count(x, n, y, k):
    return lib().or_count(x, n, y, k)


def cache(n, k):
    out = np.empty(1 << (n + k), dtype=np.uint32)
    assert lib().or_fill_cache(n, k, _ptr(out)) == 0
    return out


def rep(x, n):
    return lib().or_rep(x, n)


def orbit_size(x, n):
    return lib().or_orbit_size(x, n)


def orbit_list(n):
    cnt = lib().or_orbit_list(n, None)
    out = np.empty(cnt, dtype=np.uint32)
    lib().or_orbit_list(n, _ptr(out))
    return out


def full_output_law(n, k, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    q = np.empty(1 << k)
    lib().or_full_output_law(n, k, _ptr(w), _ptr(q))
    return q


def full_kl(n, k, logq):
    logq = np.ascontiguousarray(logq, dtype=np.float64)
    D = np.empty(1 << n)
    lib().or_full_kl(n, k, _ptr(logq), _ptr(D))
    return D


def full_dual(n, k, theta, beta, want_grad=True):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    g = np.empty(1 << k) if want_grad else None
    U = lib().or_full_dual(n, k, _ptr(theta), float(beta), _ptr(g) if want_grad else None)
    return (U, g) if want_grad else U


def entropy_x(n, k, x):
    """sum_y c(x,y) ln c(x,y) (nats) for one input x, any n <= 32"""
    return lib().or_entropy_x(n, k, int(x))


def kl_x(n, k, x, logq):
    """D(x||q) (nats) for one input x given ln q over all 2^k outputs"""
    logq = np.ascontiguousarray(logq, dtype=np.float64)
    return lib().or_kl_x(n, k, int(x), _ptr(logq))


def sparse_output_law(n, k, xs, w):
    """q(y) = sum_i w[i] c(xs[i], y) / C(n,k) over all 2^k outputs"""
    xs = np.ascontiguousarray(xs, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    assert xs.shape == w.shape
    q = np.empty(1 << k)
    lib().or_sparse_output_law(n, k, xs.size, _ptr(xs), _ptr(w), _ptr(q))
    return q


class BA:
    """The reference Blahut–Arimoto method over complement/reversal orbits."""

    def __init__(self, n, k):
        self.n, self.k = n, k
        self._c = lib().or_ctx_create(n, k)
        if not self._c:
            raise ValueError(f"oracle: unsupported (n,k)=({n},{k})")
        self.num_orbits = lib().or_ctx_num_orbits(self._c)

    def close(self):
        if self._c:
            lib().or_ctx_destroy(self._c)
            self._c = None

    def __del__(self):
        self.close()

    def orbits(self):
        out = np.empty(self.num_orbits, dtype=np.uint32)
        lib().or_ctx_orbits(self._c, _ptr(out))
        return out

    def set_q(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        assert Q.shape == (self.num_orbits,)
        lib().or_ctx_set_q(self._c, _ptr(Q))

    def get_q(self):
        out = np.empty(self.num_orbits)
        lib().or_ctx_get_q(self._c, _ptr(out))
        return out

    def iterate(self):
        return lib().or_ctx_iterate(self._c)

    def log_den(self):
        out = np.empty(1 << (self.k - 1))
        lib().or_ctx_log_den(self._c, _ptr(out))
        return out

    def kl(self):
        out = np.empty(self.num_orbits)
        lib().or_ctx_kl(self._c, _ptr(out))
        return out

    def rate(self):
        return lib().or_ctx_rate(self._c)

    def max_kl(self):
        return lib().or_ctx_max_kl(self._c)

    def sample_pass(self, first, count):
        return lib().or_ctx_sample_pass(self._c, first, count)
