"""ctypes binding of libbdc.so (the MI355X Blahut–Arimoto engine, include/bdc.h).

Host-side mirror of the reference's host function set (backend/BAA.cuh:16-29,
295-362) over the C ABI.  There is no CPU fallback: if the library or a GPU
is missing, every call raises.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbdc.so")
CSRC = os.path.join(HERE, "csrc")

PREC_FP64 = 0
PREC_SPLIT16 = 1

OK = 0
ERR_ARG, ERR_HIP, ERR_MEMORY, ERR_IO, ERR_STATE, ERR_NUMERIC = -1, -2, -3, -4, -5, -6

# every symbol include/bdc.h declares (checked by the CPU test suite)
EXPORTS = [
    "bdc_last_error", "bdc_version", "bdc_generate_cache", "bdc_write_caches", "bdc_create", "bdc_destroy",
    "bdc_supported", "bdc_num_orbits", "bdc_get_orbits", "bdc_set_p_uniform", "bdc_set_p_orbits",
    "bdc_get_p_orbits", "bdc_set_p_dirichlet", "bdc_iterate", "bdc_certify", "bdc_output_law", "bdc_kl",
    "bdc_entropy_term", "bdc_dual", "bdc_refine_dual", "bdc_certify_q", "bdc_log_output", "bdc_sync",
    "bdc_kernel_time", "bdc_reset_kernel_timers", "bdc_work_model", "bdc_split_config", "bdc_sample_inputs",
    "bdc_iterate_noalloc",
]
SAMPLE_ENTROPY, SAMPLE_P, SAMPLE_KL = 0, 1, 2


class BdcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bdc error {code}: {msg}")
        self.code = code


class IterStats(ctypes.Structure):
    _fields_ = [("mx_bits", ctypes.c_double), ("max_d", ctypes.c_double), ("log_z", ctypes.c_double),
                ("info", ctypes.c_double)]


class Certificate(ctypes.Structure):
    _fields_ = [("lower", ctypes.c_double), ("upper_cert", ctypes.c_double), ("sum_p", ctypes.c_double),
                ("upper_ref", ctypes.c_double)]


_lib = None


def build(jobs=8):
    subprocess.check_call(["make", "-s", "-C", CSRC, f"-j{jobs}"])


def lib():
    """Load libbdc.so (never falls back to anything else)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BdcError(ERR_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        i32, u32, u64, dbl, vp, cp = (ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_char_p)
        L.bdc_last_error.restype = cp
        L.bdc_version.restype = i32
        L.bdc_generate_cache.argtypes = [i32, i32, vp]
        L.bdc_write_caches.argtypes = [i32, cp]
        L.bdc_create.argtypes = [i32, i32, cp, ctypes.POINTER(vp)]
        L.bdc_destroy.argtypes = [vp]
        L.bdc_supported.argtypes = [i32, i32]
        L.bdc_num_orbits.argtypes = [vp, ctypes.POINTER(u64)]
        L.bdc_get_orbits.argtypes = [vp, vp, u64]
        L.bdc_set_p_uniform.argtypes = [vp]
        L.bdc_set_p_orbits.argtypes = [vp, vp, u64]
        L.bdc_get_p_orbits.argtypes = [vp, vp, u64]
        L.bdc_set_p_dirichlet.argtypes = [vp, dbl, u64]
        L.bdc_iterate.argtypes = [vp, i32, i32, dbl, vp, ctypes.POINTER(i32)]
        L.bdc_iterate_noalloc.argtypes = [vp, i32, i32, vp]
        L.bdc_certify.argtypes = [vp, ctypes.POINTER(Certificate)]
        L.bdc_output_law.argtypes = [vp, i32, vp, vp]
        L.bdc_kl.argtypes = [vp, i32, vp, vp]
        L.bdc_entropy_term.argtypes = [vp, vp]
        L.bdc_sample_inputs.argtypes = [vp, i32, vp, u64, vp]
        L.bdc_dual.argtypes = [vp, i32, vp, dbl, ctypes.POINTER(dbl), vp]
        L.bdc_refine_dual.argtypes = [vp, i32, vp, dbl, i32, dbl, vp]
        L.bdc_certify_q.argtypes = [vp, vp, ctypes.POINTER(dbl)]
        L.bdc_log_output.argtypes = [vp, vp]
        L.bdc_sync.argtypes = [vp]
        L.bdc_kernel_time.argtypes = [vp, cp, ctypes.POINTER(dbl), ctypes.POINTER(i32)]
        L.bdc_reset_kernel_timers.argtypes = [vp, i32]
        L.bdc_work_model.argtypes = [vp, i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.bdc_split_config.argtypes = [vp] + [ctypes.POINTER(i32)] * 5
        for name in EXPORTS:
            if name != "bdc_last_error":
                getattr(L, name).restype = i32
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise BdcError(rc, lib().bdc_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def generate_cache(n, k):
    """cache_{n,k} counts (uint32, 2^(n+k), index (x<<k)|y) computed on the GPU."""
    out = np.empty(1 << (n + k), dtype=np.uint32)
    _check(lib().bdc_generate_cache(n, k, _ptr(out)))
    return out


def write_caches(n_max, directory):
    _check(lib().bdc_write_caches(n_max, directory.encode()))


def supported(n, k):
    return bool(lib().bdc_supported(n, k))


class Engine:
    """One BDC_{n,k} context: caches, orbits, p and the BA step on one GPU."""

    def __init__(self, n, k, cache_dir=None):
        self.n, self.k = n, k
        h = ctypes.c_void_p()
        _check(lib().bdc_create(n, k, cache_dir.encode() if cache_dir else None, ctypes.byref(h)))
        self._h = h
        cnt = ctypes.c_uint64()
        _check(lib().bdc_num_orbits(self._h, ctypes.byref(cnt)))
        self.num_orbits = cnt.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().bdc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def orbits(self):
        out = np.empty(self.num_orbits, dtype=np.uint32)
        _check(lib().bdc_get_orbits(self._h, _ptr(out), self.num_orbits))
        return out

    def set_p_uniform(self):
        _check(lib().bdc_set_p_uniform(self._h))

    def set_p_orbits(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        _check(lib().bdc_set_p_orbits(self._h, _ptr(q), q.size))

    def get_p_orbits(self):
        out = np.empty(self.num_orbits)
        _check(lib().bdc_get_p_orbits(self._h, _ptr(out), self.num_orbits))
        return out

    def set_p_dirichlet(self, alpha, seed):
        _check(lib().bdc_set_p_dirichlet(self._h, float(alpha), int(seed)))

    def iterate(self, iters, precision=PREC_FP64, tol=0.0):
        stats = (IterStats * max(iters, 1))()
        done = ctypes.c_int()
        _check(lib().bdc_iterate(self._h, precision, iters, float(tol), ctypes.cast(stats, ctypes.c_void_p),
                                 ctypes.byref(done)))
        a = np.array([(s.mx_bits, s.max_d, s.log_z, s.info) for s in stats[:done.value]], dtype=np.float64)
        return a.reshape(-1, 4)

    def iterate_noalloc(self, iters, precision=PREC_FP64):
        """bdc_iterate_noalloc: the same steps with no device allocation"""
        stats = (IterStats * max(iters, 1))()
        _check(lib().bdc_iterate_noalloc(self._h, precision, iters, ctypes.cast(stats, ctypes.c_void_p)))
        a = np.array([(s.mx_bits, s.max_d, s.log_z, s.info) for s in stats[:iters]], dtype=np.float64)
        return a.reshape(-1, 4)

    def certify(self):
        c = Certificate()
        _check(lib().bdc_certify(self._h, ctypes.byref(c)))
        return {"lower": c.lower, "upper_cert": c.upper_cert, "sum_p": c.sum_p, "upper_ref": c.upper_ref}

    def output_law(self, w, precision=PREC_FP64):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == 1 << self.n
        q = np.empty(1 << self.k)
        _check(lib().bdc_output_law(self._h, precision, _ptr(w), _ptr(q)))
        return q

    def kl(self, log_q, precision=PREC_FP64):
        log_q = np.ascontiguousarray(log_q, dtype=np.float64)
        assert log_q.size == 1 << self.k
        d = np.empty(1 << self.n)
        _check(lib().bdc_kl(self._h, precision, _ptr(log_q), _ptr(d)))
        return d

    def entropy_term(self):
        h = np.empty(1 << self.n)
        _check(lib().bdc_entropy_term(self._h, _ptr(h)))
        return h

    def sample(self, what, xs):
        """values at selected inputs (SAMPLE_ENTROPY, SAMPLE_P or SAMPLE_KL)"""
        xs = np.ascontiguousarray(xs, dtype=np.uint32)
        out = np.empty(xs.size)
        _check(lib().bdc_sample_inputs(self._h, what, _ptr(xs), xs.size, _ptr(out)))
        return out

    def dual(self, theta, beta, precision=PREC_FP64, want_grad=True):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        u = ctypes.c_double()
        g = np.empty(1 << self.k) if want_grad else None
        _check(lib().bdc_dual(self._h, precision, _ptr(theta), float(beta), ctypes.byref(u),
                              _ptr(g) if want_grad else None))
        return (u.value, g) if want_grad else u.value

    def refine_dual(self, theta, beta, steps, lr, precision=PREC_FP64):
        th = np.array(theta, dtype=np.float64, copy=True)
        uh = np.empty(max(steps, 1))
        _check(lib().bdc_refine_dual(self._h, precision, _ptr(th), float(beta), int(steps), float(lr), _ptr(uh)))
        return th, uh[:steps]

    def certify_q(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        u = ctypes.c_double()
        _check(lib().bdc_certify_q(self._h, _ptr(theta), ctypes.byref(u)))
        return u.value

    def log_output(self):
        th = np.empty(1 << self.k)
        _check(lib().bdc_log_output(self._h, _ptr(th)))
        return th

    def sync(self):
        _check(lib().bdc_sync(self._h))

    def timers(self, enable=True):
        _check(lib().bdc_reset_kernel_timers(self._h, 1 if enable else 0))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        cnt = ctypes.c_int()
        _check(lib().bdc_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def work_model(self, precision=PREC_FP64):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().bdc_work_model(self._h, precision, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def split_config(self):
        """{'g1_limbs', 'g1_bits', 'g2_limbs', 'g2_bits', 'residuals'} of the split16 path."""
        v = [ctypes.c_int() for _ in range(5)]
        _check(lib().bdc_split_config(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("g1_limbs", "g1_bits", "g2_limbs", "g2_bits", "residuals"), (x.value for x in v)))
