import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def load_golden(name):
    """Golden fixture -> (System, reference outputs dict)."""
    import numpy as np
    from agbnp3_amd.synth import System
    d = np.load(os.path.join(GOLDEN, f"golden_{name}.npz"))
    s = System(natoms=int(d["natoms"]), x=d["x"], y=d["y"], z=d["z"], r=d["r"], charge=d["charge"],
               eps_in=float(d["eps"][0]), eps_out=float(d["eps"][1]), igamma=d["igamma"], sgamma=d["sgamma"],
               ialpha=d["ialpha"], salpha=d["salpha"], hbtype=d["hbtype"], hbcorr=d["hbcorr"],
               ihydrogen=d["ihydrogen"], conn_nne=d["conn_nne"], conn_neighl1=d["conn_neighl1"],
               info=dict(seed=int(d["seed"])))
    names = ("egb", "evdw", "ecorr_vdw", "ecav", "ecorr_cav", "ehb", "mol_volume")
    ref = dict(zip(names, d["energies"]))
    for k in ("sp", "br", "surf_area", "dgbdr", "dvwdr", "decav", "dehb"):
        ref[k] = d[k]
    return s, ref


GOLDEN_CASES = ("globular_n2000", "n600_s11", "n300_s7", "n400_corr")
