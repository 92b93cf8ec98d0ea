"""Tile autotuner for the batched and served Limber paths.

Tunables:
  waves     -- waves per block of the Limber kernel, each taking a share of
               the a-nodes of the same 64 (cosmology, ell) columns (1, 2, 4);
  wl_ctile  -- cosmologies per block of the weak-lensing kernel, which reuse
               every resident-table element they load (1, 4, 8).
The batch tuner times each candidate with HIP events recorded around the
launches (jc_profile_*), the server tuner by draining a sample stream.
Results are cached as JSON keyed by device and shape.
"""
import json
import os
import time

import numpy as np

from ._native import default_context

WAVES = (1, 2, 4)
CTILES = (1, 4, 8, 16)
TABLE_VARIANTS = (1, 2, 3)


def cache_path():
    """JSON cache of tuned tiles: $JAX_COSMO_AMD_TUNE_CACHE, else the file
    shipped next to the package (tuned on MI355X)."""
    return os.environ.get("JAX_COSMO_AMD_TUNE_CACHE",
                          os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json"))


_parsed = {"key": None, "db": {}}  # the cache file last parsed, by (path, mtime, size)


def _load():
    """The tuning cache (read-only: lookups run on every batch call, so the
    file is parsed again only when it changes)."""
    p = cache_path()
    try:
        st = os.stat(p)
        key = (p, st.st_mtime_ns, st.st_size)
        if _parsed["key"] != key:
            with open(p) as f:
                db = json.load(f)
            _parsed["key"], _parsed["db"] = key, db
        return _parsed["db"]
    except (OSError, ValueError):
        return {}


def _store(db):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def _device_name():
    try:
        import torch

        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:  # pragma: no cover
        return "gpu"


def batch_key(B, n_ell, n_z, T):
    b = 1 << max(0, int(np.ceil(np.log2(max(B, 1)))))  # power-of-two bucket
    return "%s|batch|B%d|L%d|N%d|T%d" % (_device_name(), b, n_ell, n_z, T)


def lookup_batch(B, n_ell, n_z, T):
    return _load().get(batch_key(B, n_ell, n_z, T))


def tune_batch(probes, B, n_ell=100, n_z=512, reps=3, seed=0, ctx=None):
    """Time every (waves, wl_ctile) on a synthetic batch of the given shape;
    cache and return the best {"waves", "wl_ctile", "ms"}."""
    from .angular_cl import _get_cl_ordering, tracer_entries

    ctx = ctx or default_context()
    sid = ctx.tracer_set(tracer_entries(probes))
    P = len(_get_cl_ordering(probes))
    T = sum(p.n_tracers for p in probes)
    rng = np.random.default_rng(seed)
    lo = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
    hi = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])
    params = np.zeros((B, 8))
    params[:, [0, 1, 2, 3, 4, 6, 7]] = lo + (hi - lo) * rng.random((B, 7))
    ell = np.geomspace(10.0, 5000.0, n_ell)
    results = []
    old = ctx.get_tiles()

    def timed():
        ctx.angular_cl(params, sid, ell, n_z, P)  # warm
        ctx.profile(True)
        ctx.profile_read()
        for _ in range(reps):
            ctx.angular_cl(params, sid, ell, n_z, P)
        prof = ctx.profile_read()
        ctx.profile(False)
        return {k: v[0] / reps for k, v in prof.items()}

    try:
        # the three kernels are independent tunables: tables variant first,
        # then (waves, wl_ctile) with the best variant
        tv = []
        for v in TABLE_VARIANTS:
            ctx.set_tables_variant(v)
            tv.append((timed()["tables"], v))
        best_v = min(tv)[1]
        ctx.set_tables_variant(best_v)
        for w in WAVES:
            for ct in CTILES:
                ctx.set_tiles(w, ct)
                t = timed()
                results.append({"waves": w, "wl_ctile": ct, "tables_variant": best_v,
                                "ms": t["limber"] + t["wl"] + t["tables"], "limber_ms": t["limber"],
                                "wl_ms": t["wl"], "tables_ms": t["tables"]})
    finally:
        ctx.set_tiles(*old)
        ctx.set_tables_variant(0)
    best = min(results, key=lambda r: r["ms"])
    best["tables_candidates"] = [{"variant": v, "ms": ms} for ms, v in tv]
    db = dict(_load())
    db[batch_key(B, n_ell, n_z, T)] = dict(best, candidates=results, when=time.time())
    _store(db)
    return best


def apply_batch(ctx, B, n_ell, n_z, T):
    """Use cached tiles for this shape, if any (0 = library default)."""
    t = lookup_batch(B, n_ell, n_z, T)
    if t:
        ctx.set_tiles(int(t["waves"]), int(t["wl_ctile"]))
        ctx.set_tables_variant(int(t.get("tables_variant", 0)))
    else:
        ctx.set_tiles(0, 0)
        ctx.set_tables_variant(0)
    return t


def server_key(n_z, max_batch):
    return "%s|server|N%s|C%d" % (_device_name(), ",".join(str(v) for v in n_z), max_batch)


def lookup_server(n_z, max_batch):
    return _load().get(server_key(n_z, max_batch))


def tune_server(pool_probes, n_requests=2048, n_z=(256, 512, 1024), max_batch=512, max_grad=160,
                max_ell=200, slots=2, seed=11):
    """Drain a sample request stream through servers built with each
    (waves, wl_ctile), then each table-kernel shape with the best tiles;
    cache and return the fastest.  Each configuration queues the whole
    stream at once (bulk submission, as a loaded server sees it) four
    times and keeps the best of the last three passes."""
    from .serving import Server, make_stream

    reqs = make_stream(n_requests, seed=seed, n_z_choices=n_z)

    def drain(**kw):
        with Server(pool_probes, max_batch=max_batch, max_grad=max_grad, max_ell=max_ell, slots=slots,
                    n_z=n_z, **kw) as srv:
            prep = srv.prepare(reqs)
            ts = []
            for rep in range(4):
                t0 = time.perf_counter()
                srv.submit_prepared(prep)
                done = 0
                while done < len(reqs):
                    done += srv.poll_array().size
                ts.append(time.perf_counter() - t0)
        return min(ts[1:])

    results = []
    for w in WAVES:
        for ct in CTILES:
            results.append({"waves": w, "wl_ctile": ct, "s": drain(waves=w, wl_ctile=ct, tables_variant=0)})
    best = min(results, key=lambda r: r["s"])
    tv = [(best["s"], 0)]
    for v in TABLE_VARIANTS:
        tv.append((drain(waves=best["waves"], wl_ctile=best["wl_ctile"], tables_variant=v), v))
    best = dict(best, s=min(tv)[0], tables_variant=min(tv)[1],
                tables_candidates=[{"variant": v, "s": dt} for dt, v in tv])
    db = dict(_load())
    db[server_key(n_z, max_batch)] = dict(best, candidates=results, when=time.time())
    _store(db)
    return best
