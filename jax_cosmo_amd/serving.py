"""Served C_ell / gradient requests: continuous batching on one GPU.

``Server`` wraps the C ABI request loop (jc_server_*): requests for different
cosmologies, tracer subsets of a resident survey pool, ell lists, Simpson
sizes and (optionally) the 7-parameter adjoint are admitted into batches as
slots and device memory free up; uneven requests share one pass of persistent
kernels replayed from a captured hipGraph per slot.
"""
import ctypes
import threading

import numpy as np

from . import _native
from ._native import Context, check, default_context

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", _I32), ("max_grad", _I32), ("max_ell", _I32), ("slots", _I32),
                ("waves", _I32), ("wl_ctile", _I32), ("use_graph", _I32), ("n_nz", _I32),
                ("n_z", _I32 * 4), ("mem_budget", _I64)]


class Request(ctypes.Structure):
    _fields_ = [("cosmo", ctypes.c_double * 8), ("n_tracers", _I32), ("tracers", _I32 * 10),
                ("n_ell", _I32), ("n_z", _I32), ("want_grad", _I32), ("reserved", _I32 * 2),
                ("ell", _P), ("cl_bar", _P), ("cl_out", _P), ("grad_out", _P)]


class Completion(ctypes.Structure):
    _fields_ = [("id", _I64), ("status", _I32), ("batch_size", _I32), ("latency_us", ctypes.c_double),
                ("t_submit_us", ctypes.c_double), ("t_done_us", ctypes.c_double)]


# jc_completion as a numpy record (poll_array)
COMPLETION_DTYPE = np.dtype([("id", "<i8"), ("status", "<i4"), ("batch_size", "<i4"), ("latency_us", "<f8"),
                             ("t_submit_us", "<f8"), ("t_done_us", "<f8")])
assert COMPLETION_DTYPE.itemsize == ctypes.sizeof(Completion)


assert ctypes.sizeof(ServerConfig) == 56
assert ctypes.sizeof(Request) == 160
assert ctypes.sizeof(Completion) == 40

_SIG = {
    "jc_server_create": [_P, _I64, _P, ctypes.POINTER(_P)],
    "jc_server_submit": [_P, _P, ctypes.POINTER(_I64)],
    "jc_server_submit_many": [_P, _P, _I64, _P],
    "jc_server_poll": [_P, _P, _I64, ctypes.POINTER(_I64)],
    "jc_server_wait": [_P, _I64, _I32],
    "jc_server_stats": [_P, _P, _I64],
    "jc_server_inject_failure": [_P, _I32, _I32],
    "jc_server_destroy": [_P],
}


def _lib():
    lib = _native.load_library()
    for n, a in _SIG.items():
        f = getattr(lib, n)
        if f.argtypes is None:
            f.argtypes = a
            f.restype = ctypes.c_int
    return lib


class Server:
    """Continuous-batching server over a pool of tracers (<= 10).

    pool_probes: probe list whose bins form the resident survey pool; a
    request names a subset of pool indices."""

    def __init__(self, pool_probes, ctx=None, max_batch=256, max_grad=64, max_ell=256, slots=2,
                 waves=0, wl_ctile=0, use_graph=True, n_z=(256, 512, 1024), mem_budget=0,
                 transfer_fn=None, nonlinear_fn=None, gamma=None, tables_variant=None):
        from . import autotune, power, transfer
        from .angular_cl import tracer_entries

        # the server keeps the physics models chosen here (angular_cl.py:49-51 arguments)
        tm = transfer.transfer_model(transfer_fn or transfer.Eisenstein_Hu)
        hm = power.halofit_model(nonlinear_fn or power.halofit)

        t = autotune.lookup_server(tuple(n_z), max_batch)
        if waves == 0 and wl_ctile == 0 and t:
            waves, wl_ctile = int(t["waves"]), int(t["wl_ctile"])
        if tables_variant is None:  # the table kernel's shape (Context.set_tables_variant)
            tables_variant = int(t.get("tables_variant", 0)) if t else 0
        self.ctx = ctx or default_context()
        self.ctx.set_models(tm, hm)
        self.ctx.set_growth(gamma)  # the growth index of every request (None: the growth ODE)
        self.lib = _lib()
        self.pool = tracer_entries(pool_probes)
        self.set_id = self.ctx.tracer_set(self.pool)
        cfg = ServerConfig()
        cfg.max_batch, cfg.max_grad, cfg.max_ell, cfg.slots = max_batch, max_grad, max_ell, slots
        cfg.waves, cfg.wl_ctile, cfg.use_graph = waves, wl_ctile, 1 if use_graph else 0
        cfg.n_nz = len(n_z)
        for i, v in enumerate(n_z):
            cfg.n_z[i] = v
        cfg.mem_budget = mem_budget
        h = _P()
        # the server takes the table kernel's shape from the context at creation
        prev_variant = self.ctx.tables_variant
        self.ctx.set_tables_variant(tables_variant)
        try:
            check(self.lib.jc_server_create(self.ctx.handle, self.set_id, ctypes.byref(cfg), ctypes.byref(h)))
        finally:
            self.ctx.set_tables_variant(prev_variant)
        self.handle = h
        self.max_ell = max_ell
        # arrays the server writes into or reads from, held until their
        # request completes: per id for single submissions; per id range
        # (start -> [end, Prepared, requests still out]) for bulk ones
        self._keep = {}
        self._ranges = {}
        self._lock = threading.Lock()

    def submit(self, cosmo, tracers, ell, n_z=512, cl_bar=None):
        """Queue one request; returns (id, cl_out array, grad_out array or None).
        The arrays are filled when the completion for id is polled."""
        r, keep = self._request(cosmo, tracers, ell, n_z, cl_bar)
        rid = _I64()
        with self._lock:
            check(self.lib.jc_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid)))
            self._keep[rid.value] = keep
        return rid.value, keep[2], keep[3]

    def prepare(self, reqs):
        """Pack request dicts (keys of submit) into one descriptor array for
        submit_prepared; returns a Prepared with .outs = [(cl_out, grad_out)]."""
        arr = (Request * len(reqs))()
        keeps = []
        for i, q in enumerate(reqs):
            r, keep = self._request(q["cosmo"], q["tracers"], q["ell"], q.get("n_z", 512), q.get("cl_bar"))
            arr[i] = r
            keeps.append(keep)
        return Prepared(arr, keeps)

    def submit_prepared(self, prep, start=0, end=None):
        """Queue requests start..end-1 of a Prepared in one call (in order);
        returns their ids."""
        end = len(prep.keeps) if end is None else end
        n = end - start
        ids = np.empty(max(n, 0), dtype=np.int64)
        if n <= 0:
            return ids
        ptr = ctypes.addressof(prep.arr) + start * ctypes.sizeof(Request)
        with self._lock:
            check(self.lib.jc_server_submit_many(self.handle, _P(ptr), n, ids.ctypes.data))
            # one call's ids are consecutive (they are assigned under the
            # server's lock, in order)
            first = int(ids[0])
            assert int(ids[-1]) == first + n - 1
            self._ranges[first] = [first + n, prep, n]
        return ids

    def _request(self, cosmo, tracers, ell, n_z, cl_bar):
        tracers = list(tracers)
        ell = np.ascontiguousarray(ell, dtype=np.float64)
        T = len(tracers)
        P = T * (T + 1) // 2
        r = Request()
        for i in range(8):
            r.cosmo[i] = float(cosmo[i])
        r.n_tracers = T
        for i, t in enumerate(tracers):
            r.tracers[i] = int(t)
        r.n_ell = len(ell)
        r.n_z = int(n_z)
        out = np.empty((P, len(ell)), dtype=np.float32)
        grad = None
        cb = None
        if cl_bar is not None:
            cb = np.ascontiguousarray(cl_bar, dtype=np.float32).reshape(P, len(ell))
            grad = np.empty(7, dtype=np.float64)
            r.want_grad = 1
            r.cl_bar = cb.ctypes.data
            r.grad_out = grad.ctypes.data
        r.ell = ell.ctypes.data
        r.cl_out = out.ctypes.data
        return r, (ell, cb, out, grad)

    def poll_array(self, max_n=4096):
        """Completions since the previous poll as a numpy record array
        (COMPLETION_DTYPE: id, status, batch_size, latency_us, t_submit_us,
        t_done_us); the outputs of each listed request are in place."""
        if getattr(self, "_pollbuf", None) is None or len(self._pollbuf) < max_n:
            self._pollbuf = np.zeros(max_n, dtype=COMPLETION_DTYPE)
        buf = self._pollbuf
        n = _I64()
        check(self.lib.jc_server_poll(self.handle, buf.ctypes.data, max_n, ctypes.byref(n)))
        res = buf[:n.value].copy()
        if n.value:
            self._release(res["id"])
        return res

    def _release(self, ids):
        with self._lock:
            if self._keep:
                for rid in ids.tolist():
                    self._keep.pop(rid, None)
            if self._ranges:
                starts = np.fromiter(self._ranges.keys(), dtype=np.int64, count=len(self._ranges))
                starts.sort()
                ends = np.array([self._ranges[st][0] for st in starts.tolist()], dtype=np.int64)
                pos = np.searchsorted(starts, ids, side="right") - 1
                inr = pos >= 0
                inr[inr] &= ids[inr] < ends[pos[inr]]
                hit, cnt = np.unique(starts[pos[inr]], return_counts=True)
                for st, c in zip(hit.tolist(), cnt.tolist()):
                    rg = self._ranges[st]
                    rg[2] -= c
                    if rg[2] <= 0:
                        del self._ranges[st]

    def poll(self, max_n=4096):
        """Completions since the previous poll, one dict each (keys of
        COMPLETION_DTYPE)."""
        arr = self.poll_array(max_n)
        return [dict(id=int(r["id"]), status=int(r["status"]), batch_size=int(r["batch_size"]),
                     latency_us=float(r["latency_us"]), t_submit_us=float(r["t_submit_us"]),
                     t_done_us=float(r["t_done_us"])) for r in arr]

    def wait(self, outstanding=0, timeout_ms=60000):
        check(self.lib.jc_server_wait(self.handle, outstanding, timeout_ms))

    def inject_failure(self, batches=1, status=-4):
        """Test hook: the next `batches` batches fail at launch with status
        (-2 out of memory, -4 runtime error); the server carries on."""
        check(self.lib.jc_server_inject_failure(self.handle, int(batches), int(status)))

    def stats(self):
        """Counters since creation; wait_ms / service_ms are summed over the
        requests done (submission to batch launch, launch to completion),
        busy_ms over the batches' device spans, launch_ms over their host
        launch calls."""
        st = np.zeros(11)
        check(self.lib.jc_server_stats(self.handle, st.ctypes.data, 11))
        keys = ("batches", "done", "rejected", "max_batch", "mean_batch", "busy_ms", "graph_replays",
                "direct_launches", "wait_ms", "service_ms", "launch_ms")
        return dict(zip(keys, st.tolist()))

    def close(self):
        if self.handle:
            self.lib.jc_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Prepared:
    """Request descriptors packed for bulk submission, with the arrays they
    point to (inputs and the output arrays the server fills)."""

    def __init__(self, arr, keeps):
        self.arr = arr
        self.keeps = keeps

    @property
    def outs(self):
        return [(k[2], k[3]) for k in self.keeps]


# ---------------------------------------------------------------- request stream
BOX_LO = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
BOX_HI = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])


def make_stream(n, seed=3, n_pool=10, grad_frac=0.3, n_z_choices=(256, 512, 1024), ell_range=(20, 200)):
    """Synthetic request stream: cosmology uniform in the box, 1..10 bins of
    the pool, 20..200 log-spaced ell in [10, 5000], n_z from the choices, a
    30 % chance of the 7-parameter gradient (seeded cotangent)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        th = BOX_LO + (BOX_HI - BOX_LO) * rng.random(7)
        cosmo = np.array([th[0], th[1], th[2], th[3], th[4], 0.0, th[5], th[6]])
        T = int(rng.integers(1, n_pool + 1))
        tracers = np.sort(rng.choice(n_pool, size=T, replace=False)).tolist()
        n_ell = int(rng.integers(ell_range[0], ell_range[1] + 1))
        ell = np.geomspace(10.0, 5000.0, n_ell)
        n_z = int(rng.choice(n_z_choices))
        cb = None
        if rng.random() < grad_frac:
            P = T * (T + 1) // 2
            cb = rng.normal(size=(P, n_ell)).astype(np.float32)
        reqs.append(dict(cosmo=cosmo, tracers=tracers, ell=ell, n_z=n_z, cl_bar=cb))
    return reqs
