"""ctypes binding of libjc_limber.so (C ABI in include/jc_limber.h).

The library must have been built (``__graft_entry__.build()`` or
``make -C jax_cosmo_amd/csrc``); there is no CPU fallback.
"""
import ctypes
import hashlib
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libjc_limber.so")

JC_DEVICE_PTRS = 1

ERRORS = {-1: ValueError, -2: MemoryError, -3: NotImplementedError, -4: RuntimeError}

EXPORTS = (
    "jc_ctx_create", "jc_ctx_destroy", "jc_ctx_set_stream", "jc_ctx_reset_stream", "jc_synchronize",
    "jc_tracers_register", "jc_tracers_register_catalogs", "jc_tracers_register_host", "jc_tracers_release",
    "jc_angular_cl",
    "jc_angular_cl_noalloc", "jc_angular_cl_jvp",
    "jc_angular_cl_vjp", "jc_matter_power", "jc_background", "jc_background_ex", "jc_set_growth", "jc_set_tiles", "jc_get_tiles", "jc_set_tables_variant",
    "jc_set_forward_options", "jc_set_models", "jc_sparse_inv", "jc_gaussian_loglike", "jc_gaussian_loglike_grad",
    "jc_dense_cholesky", "jc_dense_solve", "jc_gaussian_loglike_dense", "jc_gaussian_loglike_grad_dense",
    "jc_sigma8_norm", "jc_debug_tables", "jc_profile_enable", "jc_profile_select", "jc_profile_read", "jc_last_error", "jc_version",
)
KERNELS = ("tables", "wl", "limber", "ehtab")


class TracerDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("nz_kind", ctypes.c_int32),
        ("probe", ctypes.c_int32),
        ("bias_kind", ctypes.c_int32),
        ("nz_params", ctypes.c_double * 4),
        ("zmax", ctypes.c_double),
        ("gals_per_arcmin2", ctypes.c_double),
        ("sigma_e", ctypes.c_double),
        ("m_bias", ctypes.c_double),
        ("bias_params", ctypes.c_double * 3),
        ("nz_shifted", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


assert ctypes.sizeof(TracerDesc) == 112

_lib = None
_lib_lock = threading.Lock()

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32

# jc_pz_fn: int (*)(void* user, const double* z, double* pz, int64_t n)
PZ_FN = ctypes.CFUNCTYPE(ctypes.c_int, _P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _I64)


def _pz_callback(dist, errors):
    """A jc_pz_fn that evaluates ``dist.pz_fn`` (numpy, float64) at the
    library's redshifts; an exception is kept in ``errors`` and reported to
    the library as a failure."""
    def fn(user, z, out, n):
        try:
            zz = np.ctypeslib.as_array(z, shape=(n,)).copy()
            v = np.broadcast_to(np.asarray(dist.pz_fn(zz), dtype=np.float64), (n,))
            np.ctypeslib.as_array(out, shape=(n,))[:] = v
            return 0
        except Exception as e:  # noqa: BLE001 - reported through the C error path
            errors.append(e)
            return 1
    return PZ_FN(fn)


def load_library(path=LIB_PATH):
    """Load the HIP library and declare its signatures (raises if absent)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                "libjc_limber.so is not built (%s); run __graft_entry__.build()" % path)
        # The torch wheel ships its own HIP runtime under the same soname as
        # the system one this library links.  Loaded first, torch's copy
        # satisfies this library's dependency too, so the process runs one HIP
        # runtime; loaded the other way round, two runtimes would each try to
        # own the device and the second would see none.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(path)
        sig = {
            "jc_ctx_create": [ctypes.c_int, ctypes.POINTER(_P)],
            "jc_ctx_destroy": [_P],
            "jc_ctx_set_stream": [_P, _P],
            "jc_ctx_reset_stream": [_P],
            "jc_synchronize": [_P],
            "jc_tracers_register": [_P, _P, _I64, ctypes.POINTER(_I64)],
            "jc_tracers_register_catalogs": [_P, _P, _I64, _P, _P, ctypes.POINTER(_I64)],
            "jc_tracers_register_host": [_P, _P, _I64, _P, _P, _P, _P, ctypes.POINTER(_I64)],
            "jc_tracers_release": [_P, _I64],
            "jc_angular_cl": [_P, _P, _I64, _I64, _P, _I64, _I64, _P, _I32],
            "jc_angular_cl_noalloc": [_P, _P, _I64, _I64, _P, _I64, _I64, _P, _I32],
            "jc_angular_cl_jvp": [_P, _P, _P, _I64, _I64, _P, _I64, _I64, _P, _P, _I32],
            "jc_angular_cl_vjp": [_P, _P, _I64, _I64, _P, _I64, _I64, _P, _P, _I32],
            "jc_matter_power": [_P, _P, _P, _P, _I64, _P, _P],
            "jc_background": [_P, _P, _P, _I64, _P, _P],
            "jc_background_ex": [_P, _P, _P, _I64, _P, _P, _P, _P],
            "jc_set_growth": [_P, _I32, ctypes.c_double],
            "jc_set_tiles": [_P, _I32, _I32],
            "jc_get_tiles": [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)],
            "jc_set_tables_variant": [_P, _I32],
            "jc_set_forward_options": [_P, _I32, _I32],
            "jc_set_models": [_P, _I32, _I32],
            "jc_sparse_inv": [_P, _P, _I64, _I64, _P, _P, _I32],
            "jc_gaussian_loglike": [_P, _P, _I64, _I64, _I64, _P, _P, ctypes.c_double, _P, _P, _I32],
            "jc_gaussian_loglike_grad": [_P, _P, _I64, _I64, _P, _I64, _I64, _P, _P, ctypes.c_double, _P, _P,
                                         _I32],
            "jc_dense_cholesky": [_P, _P, _I64, _P, _P, _I32],
            "jc_dense_solve": [_P, _P, _I64, _P, _I64, _I32, _P, _I32],
            "jc_gaussian_loglike_dense": [_P, _P, _I64, _I64, _P, _P, ctypes.c_double, _P, _P, _I32],
            "jc_gaussian_loglike_grad_dense": [_P, _P, _I64, _I64, _P, _I64, _I64, _P, _P, ctypes.c_double, _P, _P,
                                               _I32],
            "jc_debug_tables": [_P, _P, _I64, _I64, _P],
            "jc_sigma8_norm": [_P, _P, _P],
            "jc_profile_enable": [_P, _I32],
            "jc_profile_select": [_P, ctypes.c_uint32],
            "jc_profile_read": [_P, _P, _P],
            "jc_last_error": [],
            "jc_version": [],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "jc_last_error" else ctypes.c_int
        _lib = lib
        return lib


def check(rc):
    if rc != 0:
        msg = load_library().jc_last_error().decode()
        raise ERRORS.get(rc, RuntimeError)(msg)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """A device context: HIP stream, resident tables and tracer sets."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = _P()
        check(self.lib.jc_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        self._sets = {}
        self._set_pairs = {}
        self._host_refs = {}  # set id -> the host n(z) functions the library calls back
        self._lock = threading.RLock()
        self._models = [0, 0, None]  # transfer, nonlinear, gamma (the library's defaults)
        self.tables_variant = 0      # the table kernel's shape (jc_set_tables_variant)

    def close(self):
        if self.handle:
            self.lib.jc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- tracers
    def tracer_set(self, entries):
        """Register (once) a list of tracer entries; returns the set id."""
        def cat_key(e):
            cat = e.get("catalog")
            if cat is None:
                return None
            return hashlib.sha1(cat[0].tobytes() + cat[1].tobytes()).hexdigest()

        hosts = [e.get("host_pz") for e in entries]
        key = tuple((e["probe"], e["kind"], e["nz_kind"], tuple(e["nz_params"]), e["nz_shifted"], e["zmax"],
                     e["gals"], e["sigma_e"], e["m_bias"], e["bias_kind"], tuple(e["bias_params"]), cat_key(e),
                     None if h is None else ("host", id(h)))
                    for e, h in zip(entries, hosts))
        with self._lock:
            if key in self._sets:
                return self._sets[key]
            arr = (TracerDesc * len(entries))()
            for i, e in enumerate(entries):
                d = arr[i]
                d.kind = e["kind"]
                d.nz_kind = e["nz_kind"]
                d.probe = e["probe"]
                d.bias_kind = e["bias_kind"]
                for q in range(4):
                    d.nz_params[q] = e["nz_params"][q]
                d.zmax = e["zmax"]
                d.gals_per_arcmin2 = e["gals"]
                d.sigma_e = e["sigma_e"]
                d.m_bias = e["m_bias"]
                for q in range(3):
                    d.bias_params[q] = e["bias_params"][q]
                d.nz_shifted = e["nz_shifted"]
                d.reserved = 0
            sid = _I64()
            cats = [e.get("catalog") for e in entries]
            if any(h is not None for h in hosts):
                # user-defined n(z): the library calls back into pz_fn; the
                # callbacks and distributions live as long as the context
                errors = []
                cbs = [None if h is None else _pz_callback(h, errors) for h in hosts]
                fns = (PZ_FN * len(entries))(*[cb if cb is not None else PZ_FN() for cb in cbs])
                flat = [None if c is None else np.ascontiguousarray(np.concatenate(c)) for c in cats]
                ptrs = (_P * len(entries))(*[None if f is None else f.ctypes.data for f in flat])
                lens = (_I64 * len(entries))(*[0 if c is None else len(c[0]) for c in cats])
                rc = self.lib.jc_tracers_register_host(self.handle, ctypes.cast(arr, _P), len(entries),
                                                       ctypes.cast(ptrs, _P), ctypes.cast(lens, _P),
                                                       ctypes.cast(fns, _P), None, ctypes.byref(sid))
                if rc != 0 and errors:
                    raise errors[0]
                check(rc)
                self._host_refs[sid.value] = (hosts, cbs, fns, errors)
            elif any(c is not None for c in cats):
                flat = [None if c is None else np.ascontiguousarray(np.concatenate(c)) for c in cats]
                ptrs = (_P * len(entries))(*[None if f is None else f.ctypes.data for f in flat])
                lens = (_I64 * len(entries))(*[0 if c is None else len(c[0]) for c in cats])
                check(self.lib.jc_tracers_register_catalogs(self.handle, ctypes.cast(arr, _P), len(entries),
                                                            ctypes.cast(ptrs, _P), ctypes.cast(lens, _P),
                                                            ctypes.byref(sid)))
            else:
                check(self.lib.jc_tracers_register(self.handle, ctypes.cast(arr, _P), len(entries),
                                                   ctypes.byref(sid)))
            self._sets[key] = sid.value
            self._set_pairs[sid.value] = len(entries) * (len(entries) + 1) // 2
            return sid.value

    def _check_pairs(self, set_id, n_pairs):
        """Host buffers are sized from ``n_pairs``: it must be the set's."""
        want = self._set_pairs.get(set_id)
        if want is not None and n_pairs != want:
            raise ValueError("tracer set %d has %d pairs, not %d" % (set_id, want, n_pairs))

    def set_stream(self, stream_handle):
        """Use a caller's HIP stream (0 = the null stream)."""
        check(self.lib.jc_ctx_set_stream(self.handle, _P(stream_handle)))

    def reset_stream(self):
        """Back to the context's own stream (host-pointer calls sync on it)."""
        check(self.lib.jc_ctx_reset_stream(self.handle))

    def synchronize(self):
        check(self.lib.jc_synchronize(self.handle))

    # -------------------------------------------------------------- host-pointer calls
    def angular_cl(self, cosmo, set_id, ell, n_z, n_pairs):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(-1, 8)
        ell = np.ascontiguousarray(ell, dtype=np.float64).reshape(-1)
        B = cosmo.shape[0]
        self._check_pairs(set_id, n_pairs)
        out = np.empty((B, n_pairs, ell.shape[0]), dtype=np.float32)
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_angular_cl(self.handle, _ptr(cosmo), B, set_id, _ptr(ell),
                                         ell.shape[0], n_z, _ptr(out), 0))
        return out

    def angular_cl_jvp(self, cosmo, tangent, set_id, ell, n_z, n_pairs):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(-1, 8)
        tangent = np.ascontiguousarray(tangent, dtype=np.float64).reshape(-1, 7)
        ell = np.ascontiguousarray(ell, dtype=np.float64).reshape(-1)
        B = cosmo.shape[0]
        if tangent.shape[0] != B:
            raise ValueError("tangent has %d rows for %d cosmologies" % (tangent.shape[0], B))
        self._check_pairs(set_id, n_pairs)
        out = np.empty((B, n_pairs, ell.shape[0]), dtype=np.float32)
        dot = np.empty_like(out)
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_angular_cl_jvp(self.handle, _ptr(cosmo), _ptr(tangent), B, set_id,
                                             _ptr(ell), ell.shape[0], n_z, _ptr(out), _ptr(dot), 0))
        return out, dot

    def angular_cl_vjp(self, cosmo, set_id, ell, n_z, cl_bar):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(-1, 8)
        ell = np.ascontiguousarray(ell, dtype=np.float64).reshape(-1)
        cl_bar = np.ascontiguousarray(cl_bar, dtype=np.float32)
        B = cosmo.shape[0]
        P = self._set_pairs.get(set_id)
        if P is not None and cl_bar.size != B * P * ell.shape[0]:
            raise ValueError("cotangent has %d values, expected [%d, %d, %d]"
                             % (cl_bar.size, B, P, ell.shape[0]))
        th = np.empty((B, 7), dtype=np.float64)
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_angular_cl_vjp(self.handle, _ptr(cosmo), B, set_id, _ptr(ell),
                                             ell.shape[0], n_z, _ptr(cl_bar), _ptr(th), 0))
        return th

    def matter_power(self, cosmo, k, a):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(8)
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1)
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        plin = np.empty_like(k)
        pnl = np.empty_like(k)
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_matter_power(self.handle, _ptr(cosmo), _ptr(k), _ptr(a), k.shape[0],
                                           _ptr(plin), _ptr(pnl)))
        return plin, pnl

    def background(self, cosmo, a):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(8)
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        chi = np.empty_like(a)
        g = np.empty_like(a)
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_background(self.handle, _ptr(cosmo), _ptr(a), a.shape[0], _ptr(chi),
                                         _ptr(g)))
        return chi, g

    def background_ex(self, cosmo, a):
        """FP64 (chi, D, dlnD/dlna, transverse comoving distance) at the scale
        factors a, under the context's growth model."""
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(8)
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        out = [np.empty_like(a) for _ in range(4)]
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_background_ex(self.handle, _ptr(cosmo), _ptr(a), a.shape[0],
                                            *[_ptr(o) for o in out]))
        return tuple(out)

    def set_growth(self, gamma=None):
        """Growth model of later calls: the growth ODE (gamma None, a Cosmology
        without gamma) or dlnD/dlna = Omega_m(a)^gamma."""
        if gamma is None:
            check(self.lib.jc_set_growth(self.handle, 0, 0.0))
        else:
            check(self.lib.jc_set_growth(self.handle, 1, float(gamma)))
        self._models[2] = None if gamma is None else float(gamma)

    def sigma8_norm(self, cosmo):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(8)
        out = ctypes.c_double()
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_sigma8_norm(self.handle, _ptr(cosmo), ctypes.byref(out)))
        return out.value

    def debug_tables(self, cosmo, set_id, n_z):
        cosmo = np.ascontiguousarray(cosmo, dtype=np.float64).reshape(8)
        out = np.zeros(1024 + 8 * (n_z + 1), dtype=np.float64)
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_debug_tables(self.handle, _ptr(cosmo), set_id, n_z, _ptr(out)))
        return out

    def profile(self, on, kernels=None):
        """Per-kernel event timing on or off; kernels: names from KERNELS to
        time (None: every kernel)."""
        mask = (1 << len(KERNELS)) - 1 if kernels is None else sum(1 << KERNELS.index(k) for k in kernels)
        check(self.lib.jc_profile_select(self.handle, mask))
        check(self.lib.jc_profile_enable(self.handle, 1 if on else 0))

    def profile_read(self):
        """{kernel: (total_ms, launches)} since the last read."""
        ms = np.zeros(len(KERNELS), dtype=np.float64)
        n = np.zeros(len(KERNELS), dtype=np.int64)
        with self._lock:
            check(self.lib.jc_profile_read(self.handle, _ptr(ms), _ptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNELS)}

    def set_tiles(self, waves, ctile):
        check(self.lib.jc_set_tiles(self.handle, waves, ctile))

    def set_tables_variant(self, v):
        check(self.lib.jc_set_tables_variant(self.handle, int(v)))
        self.tables_variant = int(v)

    def set_models(self, transfer=-1, nonlinear=-1):
        """Transfer function (0 eisenhu_osc, 1 eisenhu) and halofit
        prescription (0 takahashi2012, 1 smith2003) of later calls; -1 keeps."""
        check(self.lib.jc_set_models(self.handle, int(transfer), int(nonlinear)))
        if transfer >= 0:
            self._models[0] = int(transfer)
        if nonlinear >= 0:
            self._models[1] = int(nonlinear)

    def models(self):
        """(transfer, nonlinear, gamma) of later calls, to restore with
        set_models / set_growth."""
        return tuple(self._models)

    def set_forward_options(self, transfer_table=-1, matrix_cores=-1):
        """Forward-call options (1 on, 0 off, -1 unchanged): T(k) from a
        per-cosmology table instead of the EH98 closed form per point, and the
        pair contraction on the matrix cores instead of per-lane FMAs."""
        check(self.lib.jc_set_forward_options(self.handle, int(transfer_table), int(matrix_cores)))

    def get_tiles(self):
        w = _I32()
        c = _I32()
        check(self.lib.jc_get_tiles(self.handle, ctypes.byref(w), ctypes.byref(c)))
        return w.value, c.value

    # -------------------------------------------------------------- device-pointer calls
    def angular_cl_dev(self, cosmo_ptr, B, set_id, ell_ptr, n_ell, n_z, out_ptr, stream=None):
        with self._lock:
            if stream is not None:
                self.set_stream(stream)
            else:
                self.reset_stream()
            check(self.lib.jc_angular_cl(self.handle, _P(cosmo_ptr), B, set_id, _P(ell_ptr), n_ell,
                                         n_z, _P(out_ptr), JC_DEVICE_PTRS))

    def angular_cl_noalloc_dev(self, cosmo_ptr, B, set_id, ell_ptr, n_ell, n_z, out_ptr, stream=None):
        """jc_angular_cl_noalloc with device pointers: kernel launches only,
        so a caller may capture it in a graph."""
        with self._lock:
            if stream is not None:
                self.set_stream(stream)
            else:
                self.reset_stream()
            check(self.lib.jc_angular_cl_noalloc(self.handle, _P(cosmo_ptr), B, set_id, _P(ell_ptr), n_ell,
                                                 n_z, _P(out_ptr), JC_DEVICE_PTRS))

    def angular_cl_vjp_dev(self, cosmo_ptr, B, set_id, ell_ptr, n_ell, n_z, cbar_ptr, theta_ptr,
                           stream=None):
        with self._lock:
            if stream is not None:
                self.set_stream(stream)
            else:
                self.reset_stream()
            check(self.lib.jc_angular_cl_vjp(self.handle, _P(cosmo_ptr), B, set_id, _P(ell_ptr),
                                             n_ell, n_z, _P(cbar_ptr), _P(theta_ptr), JC_DEVICE_PTRS))

    def angular_cl_jvp_dev(self, cosmo_ptr, tangent_ptr, B, set_id, ell_ptr, n_ell, n_z, out_ptr,
                           dot_ptr, stream=None):
        with self._lock:
            if stream is not None:
                self.set_stream(stream)
            else:
                self.reset_stream()
            check(self.lib.jc_angular_cl_jvp(self.handle, _P(cosmo_ptr), _P(tangent_ptr), B, set_id,
                                             _P(ell_ptr), n_ell, n_z, _P(out_ptr), _P(dot_ptr),
                                             JC_DEVICE_PTRS))


_contexts = {}
_ctx_lock = threading.Lock()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _sparse_ok(C, P=None, n_ell=None):
    C = _f64(C)
    if C.ndim != 3 or C.shape[0] != C.shape[1]:
        raise ValueError("expected a square sparse matrix [P][P][n_ell], got shape %s" % (C.shape,))
    if P is not None and C.shape != (P, P, n_ell):
        raise ValueError("expected a sparse matrix [%d][%d][%d], got shape %s" % (P, P, n_ell, C.shape))
    return C


class _Likelihood:
    """Likelihood entry points of a Context (host arrays)."""

    def sparse_inv(self, C):
        """(inverse [P][P][n_ell], log det) of a block-diagonal covariance."""
        C = _sparse_ok(C)
        P, _, n_ell = C.shape
        out = np.empty_like(C)
        ld = ctypes.c_double()
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_sparse_inv(self.handle, _ptr(C), P, n_ell, _ptr(out), ctypes.byref(ld), 0))
        return out, ld.value

    def gaussian_loglike(self, mu, data, prec, c0=0.0, want_dmu=False):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        if mu.ndim == 2:
            mu = mu[None]
        B, P, n_ell = mu.shape
        data = _f64(data).reshape(P, n_ell)
        prec = _sparse_ok(prec, P, n_ell)
        ll = np.empty(B)
        dmu = np.empty_like(mu) if want_dmu else None
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_gaussian_loglike(self.handle, _ptr(mu), B, P, n_ell, _ptr(data), _ptr(prec),
                                               float(c0), _ptr(ll), _ptr(dmu) if want_dmu else None, 0))
        return (ll, dmu) if want_dmu else ll

    def gaussian_loglike_grad(self, cosmo, set_id, ell, n_z, data, prec, c0=0.0):
        cosmo = _f64(cosmo).reshape(-1, 8)
        ell = _f64(ell).reshape(-1)
        B = cosmo.shape[0]
        P = self._set_pairs.get(set_id, prec.shape[0] if hasattr(prec, "shape") else 0)
        prec = _sparse_ok(prec, P, ell.shape[0])
        data = _f64(data).reshape(P, ell.shape[0])
        ll = np.empty(B)
        g = np.empty((B, 7))
        with self._lock:
            self.reset_stream()
            check(self.lib.jc_gaussian_loglike_grad(self.handle, _ptr(cosmo), B, set_id, _ptr(ell), ell.shape[0], n_z,
                                                    _ptr(data), _ptr(prec), float(c0), _ptr(ll), _ptr(g), 0))
        return ll, g

    def gaussian_loglike_grad_dev(self, cosmo_ptr, B, set_id, ell_ptr, n_ell, n_z, data_ptr, prec_ptr, c0,
                                  ll_ptr, grad_ptr, stream=None):
        with self._lock:
            if stream is not None:
                self.set_stream(stream)
            else:
                self.reset_stream()
            check(self.lib.jc_gaussian_loglike_grad(self.handle, _P(cosmo_ptr), B, set_id, _P(ell_ptr), n_ell, n_z,
                                                    _P(data_ptr), _P(prec_ptr), float(c0), _P(ll_ptr),
                                                    _P(grad_ptr), JC_DEVICE_PTRS))


    # ---------------------------------------------------------- dense covariances
    def dense_cholesky(self, C):
        """Lower Cholesky factor of a dense covariance [N][N], kept on the
        device: (L as a CUDA float64 tensor, log det C)."""
        import torch

        C = _f64(C)
        if C.ndim != 2 or C.shape[0] != C.shape[1]:
            raise ValueError("expected a square dense covariance, got shape %s" % (C.shape,))
        N = C.shape[0]
        stream = torch.cuda.current_stream()
        L = torch.from_numpy(C).to("cuda")
        ld = ctypes.c_double()
        with self._lock:
            self.set_stream(stream.cuda_stream)
            try:
                check(self.lib.jc_dense_cholesky(self.handle, _P(L.data_ptr()), N, _P(L.data_ptr()),
                                                 ctypes.byref(ld), JC_DEVICE_PTRS))
            finally:
                self.reset_stream()
        return L, ld.value

    def dense_solve(self, L, R, mode=2):
        """L^-1 R (mode 0), L^-T R (1) or C^-1 R (2) for right-hand sides as
        rows of R [nrhs][N] (a CUDA float64 tensor is solved in place on the
        device and returned; a numpy array gives a new numpy array)."""
        import torch

        N = L.shape[0]
        on_dev = isinstance(R, torch.Tensor)
        X = R.to(device=L.device, dtype=torch.float64).contiguous() if on_dev else \
            torch.as_tensor(_f64(R).reshape(-1, N)).to(L.device)
        X = X.reshape(-1, N)
        with self._lock:
            self.set_stream(torch.cuda.current_stream().cuda_stream)
            try:
                check(self.lib.jc_dense_solve(self.handle, _P(L.data_ptr()), N, _P(X.data_ptr()), X.shape[0], int(mode),
                                              _P(X.data_ptr()), JC_DEVICE_PTRS))
            finally:
                self.reset_stream()
        return X if on_dev else X.cpu().numpy()

    def gaussian_loglike_dense(self, mu, data, L, c0=0.0, want_dmu=False):
        """-1/2 r^T C^-1 r + c0 for means mu [B][N] (or [N]) and C = L L^T
        (L from dense_cholesky); with want_dmu also C^-1 r [B][N]."""
        import torch

        N = L.shape[0]
        dev = L.device
        mu_t = torch.as_tensor(np.asarray(mu, dtype=np.float32).reshape(-1, N)).to(dev)
        B = mu_t.shape[0]
        data_t = torch.as_tensor(_f64(data).reshape(N)).to(dev)
        ll = torch.empty(B, dtype=torch.float64, device=dev)
        dmu = torch.empty((B, N), dtype=torch.float32, device=dev) if want_dmu else None
        with self._lock:
            self.set_stream(torch.cuda.current_stream().cuda_stream)
            try:
                check(self.lib.jc_gaussian_loglike_dense(self.handle, _P(mu_t.data_ptr()), B, N, _P(data_t.data_ptr()),
                                                         _P(L.data_ptr()), float(c0), _P(ll.data_ptr()),
                                                         _P(dmu.data_ptr()) if want_dmu else None, JC_DEVICE_PTRS))
            finally:
                self.reset_stream()
        ll = ll.cpu().numpy()
        return (ll, dmu.cpu().numpy()) if want_dmu else ll

    def gaussian_loglike_grad_dense(self, cosmo, set_id, ell, n_z, data, L, c0=0.0):
        """(log-likelihood [B], its gradient [B][7]) of the means
        angular_cl(cosmo_b) against data with C = L L^T."""
        import torch

        dev = L.device
        cosmo = _f64(cosmo).reshape(-1, 8)
        ell = _f64(ell).reshape(-1)
        B = cosmo.shape[0]
        N = L.shape[0]
        P = self._set_pairs.get(set_id)
        if P is None or P * ell.shape[0] != N:
            raise ValueError("the factor is %d x %d, the mean has pairs x ell = %s x %d" % (N, N, P, ell.shape[0]))
        # the checks the host-pointer entry points make on the cosmologies
        # (the device-pointer call takes them as given)
        ok = np.isfinite(cosmo).all(axis=1) & (cosmo[:, 0] + cosmo[:, 1] > 0) & (cosmo[:, 1] > 0) & \
            (cosmo[:, 2] > 0) & (cosmo[:, 4] > 0)
        if not ok.all():
            raise ValueError("cosmology %d: parameters must be finite and Omega_b, Omega_m, h, sigma8 positive"
                             % int(np.flatnonzero(~ok)[0]))
        cp = torch.as_tensor(cosmo).to(dev)
        el = torch.as_tensor(ell).to(dev)
        data_t = torch.as_tensor(_f64(data).reshape(N)).to(dev)
        ll = torch.empty(B, dtype=torch.float64, device=dev)
        g = torch.empty((B, 7), dtype=torch.float64, device=dev)
        with self._lock:
            self.set_stream(torch.cuda.current_stream().cuda_stream)
            try:
                check(self.lib.jc_gaussian_loglike_grad_dense(
                    self.handle, _P(cp.data_ptr()), B, set_id, _P(el.data_ptr()), ell.shape[0], n_z,
                    _P(data_t.data_ptr()), _P(L.data_ptr()), float(c0), _P(ll.data_ptr()), _P(g.data_ptr()),
                    JC_DEVICE_PTRS))
            finally:
                self.reset_stream()
        return ll.cpu().numpy(), g.cpu().numpy()


for _name in ("sparse_inv", "gaussian_loglike", "gaussian_loglike_grad", "gaussian_loglike_grad_dev",
              "dense_cholesky", "dense_solve", "gaussian_loglike_dense", "gaussian_loglike_grad_dense"):
    setattr(Context, _name, getattr(_Likelihood, _name))


def default_context(device=0):
    with _ctx_lock:
        if device not in _contexts:
            _contexts[device] = Context(device)
        return _contexts[device]
