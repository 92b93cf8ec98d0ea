"""CPU tests of the host side: the C ABI library loads and exports every
declared symbol, the probe -> tracer-descriptor mapping, the host covariance
and likelihood code against the oracle and the reference's known answers,
and the loud failure when no GPU is present."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc
from jax_cosmo_amd import _native
from oracle import jc_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "jc_limber.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(jc_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    syms = header_symbols()
    assert len(syms) >= 16
    for s in syms:
        assert hasattr(lib, s), s
    from jax_cosmo_amd import serving

    assert set(syms) == set(_native.EXPORTS) | set(serving._SIG)
    assert lib.jc_version() == 2


def test_tracer_desc_layout():
    import ctypes

    assert ctypes.sizeof(_native.TracerDesc) == 112
    assert _native.TracerDesc.nz_params.offset == 16
    assert _native.TracerDesc.bias_params.offset == 80
    assert _native.TracerDesc.nz_shifted.offset == 104


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure mode")
def test_no_gpu_fails_loudly():
    with pytest.raises(RuntimeError, match="device"):
        _native.Context(0)


def c2_probes():
    wl = jc.probes.WeakLensing(
        [jc.redshift.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)])
    nc = jc.probes.NumberCounts(
        [jc.redshift.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0) for mu in np.linspace(0.3, 1.5, 5)],
        jc.bias.constant_linear_bias(1.5))
    return [wl, nc]


def test_tracer_entries_order_and_kinds():
    e = jc.cl.tracer_entries(c2_probes())
    assert len(e) == 10
    assert [x["kind"] for x in e] == [0] * 5 + [1] * 5
    assert [x["probe"] for x in e] == [0] * 5 + [1] * 5
    assert e[2]["nz_params"][:3] == (1.0, 2.0, 1.0)
    assert e[7]["bias_params"][0] == 1.5 and e[7]["bias_kind"] == 0 and e[7]["nz_params"][1] == 0.1
    assert e[0]["bias_kind"] == 3  # no intrinsic alignments
    assert e[0]["gals"] == 10.0 and e[0]["sigma_e"] == 0.26
    assert jc.cl._get_cl_ordering(c2_probes())[:3] == [(0, 0), (0, 1), (0, 2)]
    assert len(jc.cl._get_cl_ordering(c2_probes())) == 55


def test_unsupported_inputs_raise():
    with pytest.raises(NotImplementedError):
        jc.cl.tracer_entries([jc.probes.WeakLensing([jc.redshift.delta_nz(1.0)],
                                                    ia_bias=jc.bias.constant_linear_bias(1.0))])
    with pytest.raises(NotImplementedError):
        jc.cl.tracer_entries([jc.probes.NumberCounts([jc.redshift.delta_nz(1.0)], jc.bias.constant_linear_bias(1.0))])
    e = jc.cl.tracer_entries([jc.probes.WeakLensing([jc.redshift.kde_nz(np.array([0.5, 0.7]), 1.0, bw=0.1)])])
    assert e[0]["nz_kind"] == 4 and e[0]["nz_params"][0] == 0.1
    np.testing.assert_array_equal(e[0]["catalog"][1], [1.0, 1.0])
    with pytest.raises(NotImplementedError):
        jc.cl.tracer_entries([jc.probes.NumberCounts([jc.redshift.smail_nz(1, 2, 1)], lambda cosmo, z: z)])
    # gamma selects the growth index model (core.py:13, 59-60)
    cg = jc.Cosmology(0.3, 0.05, 0.7, 0.96, 0.8, 0.0, -1.0, 0.0, gamma=0.55)
    assert cg.gamma == 0.55 and cg._flags["gamma_growth"]
    assert jc.Planck15().gamma is None and not jc.Planck15()._flags["gamma_growth"]

    class NoShape(jc.redshift.redshift_distribution):
        pass

    with pytest.raises(NotImplementedError):
        jc.cl.tracer_entries([jc.probes.WeakLensing([NoShape()])])
    with pytest.raises(NotImplementedError):
        jc.cl.tracer_entries([jc.probes.WeakLensing([jc.redshift.smail_nz(1, 2, 1)] * 11)])


def test_user_defined_nz_entries():
    # a subclass with its own pz_fn (redshift.py:15-31) becomes a host n(z)
    # bin: the library calls that pz_fn; a systematic_shift of it keeps the
    # parent's pz_fn with the shift in the descriptor
    class Wide(jc.redshift.redshift_distribution):
        def pz_fn(self, z):
            return z * z * np.exp(-z / 0.4)

    w = Wide(zmax=6.0)
    sh = jc.redshift.systematic_shift(w, 0.05)
    e = jc.cl.tracer_entries([jc.probes.WeakLensing([w, sh]),
                              jc.probes.NumberCounts([jc.redshift.smail_nz(1, 2, 1)],
                                                     jc.bias.constant_linear_bias(1.0))])
    assert [x["nz_kind"] for x in e] == [jc.redshift.NZ_HOST, jc.redshift.NZ_HOST, jc.redshift.NZ_SMAIL]
    assert e[0]["host_pz"] is w and e[1]["host_pz"] is w and e[2]["host_pz"] is None
    assert e[1]["nz_shifted"] == 1 and e[1]["nz_params"][3] == 0.05 and e[0]["zmax"] == 6.0
    # normalised by the 256-interval Simpson over [0, zmax] (redshift.py:27-31)
    from scipy.integrate import simpson

    z = np.linspace(0.0, 6.0, 257)
    norm = simpson(z * z * np.exp(-z / 0.4), x=z)
    zq = np.array([0.3, 1.0])
    np.testing.assert_allclose(w(zq), zq * zq * np.exp(-zq / 0.4) / norm, rtol=1e-13)


def test_nz_host_normalisation_matches_oracle():
    for nz, onz in [(jc.redshift.smail_nz(1.0, 2.0, 0.5), O.smail_nz(1.0, 2.0, 0.5)),
                    (jc.redshift.gaussian_nz(0.3, 0.1), O.gaussian_nz(0.3, 0.1))]:
        z = np.linspace(0.0, 3.0, 31)
        np.testing.assert_allclose(nz(z), onz(torch.tensor(z)).numpy(), rtol=1e-13, atol=1e-300)


def test_noise_and_covariance_match_oracle():
    probes = c2_probes()
    ell = np.geomspace(10.0, 5000.0, 20)
    rng = np.random.default_rng(0)
    sig = np.abs(rng.normal(size=(55, 20))) * 1e-8
    n = jc.cl.noise_cl(ell, probes)
    owl = O.WeakLensing([O.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)])
    onc = O.NumberCounts([O.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0) for mu in np.linspace(0.3, 1.5, 5).tolist()],
                         1.5)
    on = O.noise_cl(torch.tensor(ell), [owl, onc]).numpy()
    np.testing.assert_allclose(n, on, rtol=1e-14)
    for sparse in (True, False):
        c1 = jc.cl.gaussian_cl_covariance(ell, probes, sig, n, f_sky=0.25, sparse=sparse)
        c2 = O.gaussian_cl_covariance(torch.tensor(ell), [owl, onc], torch.tensor(sig), torch.tensor(n),
                                      f_sky=0.25, sparse=sparse).numpy()
        np.testing.assert_allclose(c1, c2, rtol=1e-13)


def test_sparse_known_answers():
    # reference tests/test_sparse.py:8-84 known answers
    X = np.array([[[1, 2, 3], [4, 5, 6], [-1, -2, -3]], [[1, 2, 3], [-4, -5, -6], [7, 8, 9]]])
    ans = np.array([[1, 0, 0, 4, 0, 0, -1, 0, 0], [0, 2, 0, 0, 5, 0, 0, -2, 0], [0, 0, 3, 0, 0, 6, 0, 0, -3],
                    [1, 0, 0, -4, 0, 0, 7, 0, 0], [0, 2, 0, 0, -5, 0, 0, 8, 0], [0, 0, 3, 0, 0, -6, 0, 0, 9]])
    np.testing.assert_array_equal(jc.sparse.to_dense(X), ans)
    with pytest.raises(ValueError):
        jc.sparse.to_dense([1, 2, 3])
    Xs = np.array([[[1.0, 1.0], [1.0, 1.0]], [[1.0, 1.0], [2.0, 2.0]]])
    np.testing.assert_allclose(jc.sparse.inv(Xs), [[[2.0, 2.0], [-1.0, -1.0]], [[-1.0, -1.0], [1.0, 1.0]]])
    with pytest.raises(ValueError):
        jc.sparse.inv(np.ones((2, 3, 4)))
    X1 = [[[1.0, 2], [3, 4], [5, 6]], [[4, 5], [6, 7], [8, 9]]]
    X2 = [[[1.0, -2], [3, -4]], [[5, 4], [6, -7]], [[5, 6], [9, 8]]]
    X1d, X2d = jc.sparse.to_dense(X1), jc.sparse.to_dense(X2)
    v1, v2 = np.arange(6), np.arange(4)
    np.testing.assert_allclose(X2d @ v2, jc.sparse.dot(X2, v2))
    np.testing.assert_allclose(v2 @ X1d, jc.sparse.dot(v2, X1))
    np.testing.assert_allclose(X1d @ X2d, jc.sparse.dot(X1, X2d))
    np.testing.assert_allclose(X1d @ X2d, jc.sparse.dot(X1d, X2))
    np.testing.assert_allclose(X1d @ X2d, jc.sparse.to_dense(jc.sparse.dot(X1, X2)))
    np.testing.assert_allclose(X1d @ (X2d @ X1d) @ X2d, jc.sparse.dot(X1d, jc.sparse.dot(X2, X1), X2d))
    Xd = np.array([[[1, 2, 3], [4, 5, 6], [-1, 7, -2]], [[1, 2, 3], [-4, -5, -6], [2, -3, 9]],
                   [[7, 8, 9], [5, -4, 6], [-3, -2, -1]]], dtype=float)
    assert jc.sparse.det(Xd) == pytest.approx(np.linalg.det(jc.sparse.to_dense(Xd)), rel=1e-10)
    assert jc.sparse.det(-Xd) == pytest.approx(-jc.sparse.det(Xd), rel=1e-12)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_server_client_releases_arrays_per_completion():
    # the client's bookkeeping (no device): arrays of a bulk submission are
    # held until every request of its id range has completed; single
    # submissions by id; ids outside every range touch nothing
    import threading

    from jax_cosmo_amd import serving

    srv = object.__new__(serving.Server)
    srv._lock = threading.Lock()
    srv._keep = {200: "single"}
    srv._ranges = {10: [14, "prepA", 4], 20: [22, "prepB", 2]}
    srv._release(np.array([11, 13, 200, 21, 300], dtype=np.int64))
    assert srv._keep == {}
    assert srv._ranges == {10: [14, "prepA", 2], 20: [22, "prepB", 1]}
    srv._release(np.array([10, 12, 17, 5], dtype=np.int64))  # 17 and 5 lie in no range
    assert srv._ranges == {20: [22, "prepB", 1]}
    srv._release(np.array([20], dtype=np.int64))
    assert srv._ranges == {}
    assert serving.COMPLETION_DTYPE.itemsize == 40


def test_tuning_cache_is_reread_only_when_the_file_changes(tmp_path, monkeypatch):
    from jax_cosmo_amd import autotune

    monkeypatch.setenv("JAX_COSMO_AMD_TUNE_CACHE", str(tmp_path / "tiles.json"))
    assert autotune._load() == {}
    autotune._store({"k1": {"waves": 2}})
    first = autotune._load()
    assert first == {"k1": {"waves": 2}}
    assert autotune._load() is first  # unchanged file: the parsed copy
    autotune._store({"k1": {"waves": 2}, "k2": {"waves": 4}})
    assert autotune._load() == {"k1": {"waves": 2}, "k2": {"waves": 4}}
    (tmp_path / "tiles.json").write_text("{not json")
    assert autotune._load() == {}
