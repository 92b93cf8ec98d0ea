"""Writes the golden C_ell fixtures of this directory from the FP64 oracle.

Run from the repository root:  python tests/golden/make_golden.py
The reference itself cannot be run (it needs JAX, absent here), so these
vectors come from oracle/jc_oracle.py, which is pinned component by component
in tests/test_oracle_pins.py.  Fixtures hold inputs and FP64 outputs only.
"""
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import jc_oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
P15 = [0.2589, 0.0486, 0.6774, 0.9667, 0.8159, 0.0, -1.0, 0.0]


def c2_probes():
    wl = O.WeakLensing([O.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)],
                       sigma_e=0.26)
    nc = O.NumberCounts([O.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0)
                         for mu in np.linspace(0.3, 1.5, 5).tolist()], 1.5)
    return [wl, nc]


def main():
    torch.set_num_threads(min(8, os.cpu_count() or 1))
    # C1: Planck15, one Smail WL bin (1, 2, 1, zmax 4), 50 ell in [10, 3000], N = 256
    ell1 = np.geomspace(10.0, 3000.0, 50)
    cl1 = O.angular_cl(O.Cosmo(*P15), torch.tensor(ell1), [O.WeakLensing([O.smail_nz(1.0, 2.0, 1.0, zmax=4.0)])],
                       N=256).numpy()
    np.savez(os.path.join(HERE, "c1_planck15.npz"), params=np.array(P15), ell=ell1, n_z=256, cl=cl1)
    # C2: Planck15, 10 tracers, 100 ell in [10, 5000], N = 512
    ell2 = np.geomspace(10.0, 5000.0, 100)
    cl2 = O.angular_cl(O.Cosmo(*P15), torch.tensor(ell2), c2_probes(), N=512).numpy()
    np.savez(os.path.join(HERE, "c2_planck15.npz"), params=np.array(P15), ell=ell2, n_z=512, cl=cl2)
    # three cosmologies from the batched box (default_rng(0)), C2 tracers
    lo = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
    hi = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])
    th = lo + (hi - lo) * np.random.default_rng(0).random((3, 7))
    params = np.zeros((3, 8))
    params[:, [0, 1, 2, 3, 4, 6, 7]] = th
    cls = np.stack([O.angular_cl(O.Cosmo(*p), torch.tensor(ell2), c2_probes(), N=512).numpy() for p in params])
    np.savez(os.path.join(HERE, "c2_box3.npz"), params=params, ell=ell2, n_z=512, cl=cls)
    # gradient of sum(C_ell * w) for C1 (torch autograd through the oracle)
    w = np.random.default_rng(5).normal(size=cl1.shape)
    pv = torch.tensor(P15, dtype=torch.float64, requires_grad=True)
    s = (O.angular_cl(O.Cosmo.from_vector(pv), torch.tensor(ell1),
                      [O.WeakLensing([O.smail_nz(1.0, 2.0, 1.0, zmax=4.0)])], N=256) * torch.tensor(w)).sum()
    s.backward()
    np.savez(os.path.join(HERE, "c1_grad.npz"), params=np.array(P15), ell=ell1, n_z=256, weights=w,
             grad=pv.grad.numpy()[[0, 1, 2, 3, 4, 6, 7]])
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
