"""Evaluate a system from the reference's text input dump (AGBNP_WRITE_DATA
format, agbnp3.c:319-336) on the GPU and print its energies.

usage: python -m agbnp3_amd.run_dump DUMP [--bonds BONDS.npz] [--eps-in 1] [--eps-out 80]

BONDS.npz (optional) holds `nne` and `neighl1` (NeighList layout, caller
order); without it the bonds are inferred from distances.
"""
import argparse
import json

import numpy as np

from . import api, sysio


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("dump")
    ap.add_argument("--bonds")
    ap.add_argument("--eps-in", type=float, default=1.0)
    ap.add_argument("--eps-out", type=float, default=80.0)
    a = ap.parse_args(argv)
    bonds = None
    if a.bonds:
        z = np.load(a.bonds, allow_pickle=False)
        bonds = (z["nne"], z["neighl1"])
    s = sysio.read_agbnp_dump(a.dump, bonds=bonds, eps_in=a.eps_in, eps_out=a.eps_out)
    inst = api.Instance(s)
    try:
        out = inst.ener()
    finally:
        inst.delete()
    print(json.dumps({k: out[k] for k in api.ENERGY_NAMES} | {"natoms": s.natoms}))


if __name__ == "__main__":
    main()
