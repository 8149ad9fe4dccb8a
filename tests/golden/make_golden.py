"""Writes the committed golden fixtures: small recordings in the batch (site-grouped)
layout with the canonical outputs from the CPU oracle, which tests/test_oracle.py pins
independently (integer brute force, exact rationals, k-d tree, closed forms, link and
paging properties).  Run from the repository root:  python tests/golden/make_golden.py"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ironclust_amd import irc2, synth  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.helpers import int_dataset  # noqa: E402


def save(name, ds):
    t = {kk: (torch.as_tensor(v) if kk in ("F1", "F2", "site1", "site2") and v is not None else v)
         for kk, v in ds.items()}
    lay = irc2.site_layout(t)
    ref = O.sort_batch(ds)
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        F=lay["F"].numpy(), col_spike=lay["col_spike"].numpy(), site_ptr=lay["site_ptr"],
                        site_n_primary=lay["site_n_primary"], lim=np.asarray(ds["lim"]),
                        link=np.asarray(ds["link"]), n_spikes=ds["n_spikes"], k=ds["k"],
                        rho=ref["rho"], delta=ref["delta"], nneigh=ref["nneigh"], knn=ref["knn"])


if __name__ == "__main__":
    save("drift_f2_2k", synth.make_dataset(2000, 3, 8, 6, 3, k=10, drift_step=0.4, n_links=3, band=2,
                                           seed=101, f2=True))
    save("int_ties_1k", int_dataset(N=1000, S=2, C=4, D=3, k=8, seed=102, f2=True, lo=-2, hi=3))
    save("sparse_k_gt_cands", synth.make_dataset(300, 30, 5, 6, 2, k=12, drift_step=0.2, n_links=2, band=1,
                                                 seed=103))
    save("plain_d24_k30", synth.make_dataset(3000, 2, 24, 1, 8, k=30, seed=104))
    print("written")
