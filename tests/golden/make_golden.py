"""Writes the golden vectors under tests/golden/ from the reference library.

Runs in the build container only (needs /root/reference).  The reference is
compiled from its own sources by oracle/Makefile into oracle/_ref/ and run
single-threaded (OMP_NUM_THREADS=1), because its E_HB reads a stale array
index that depends on the per-thread water-site order (agbnp3_cpu_simd_kernels.c:2854).

Each fixture holds the agbnp3_new() inputs (caller order), the reference
outputs of one agbnp3_ener() call and provenance (sha256 of the reference
sources, compiler flags).

usage: python tests/golden/make_golden.py
"""
import hashlib
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from agbnp3_amd import synth  # noqa: E402
import oracle  # noqa: E402

REF_DIR = os.environ.get("AGBNP3_REF_DIR", "/root/reference")
CASES = [  # name, seed, natoms, correction-parameter variant
    ("globular_n2000", 20260001, 2000, False),
    ("n600_s11", 11, 600, False),
    ("n300_s7", 7, 300, False),
    ("n400_corr", 23, 400, True),
]


def with_corrections(s, seed):
    """Non-zero sgamma/salpha and other dielectrics, so E_corr_cav,
    E_corr_vdW and the dielectric factor are pinned too."""
    import dataclasses
    rng = np.random.default_rng(seed)
    heavy = np.ones(s.natoms, bool)
    heavy[s.ihydrogen] = False
    sg = np.where(heavy, 0.02, 0.0)
    sa = np.where(heavy, rng.uniform(-0.2, 0.0, s.natoms), 0.0)
    return dataclasses.replace(s, sgamma=sg, salpha=sa, eps_in=2.0, eps_out=78.5)


def ref_digest():
    h = hashlib.sha256()
    for f in sorted(os.listdir(REF_DIR)):
        if f.endswith((".c", ".h")):
            with open(os.path.join(REF_DIR, f), "rb") as fh:
                h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()


def main():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    digest = ref_digest()
    for name, seed, n, corr in CASES:
        s = synth.generate(seed, n)
        if corr:
            s = with_corrections(s, seed)
        res, timing = oracle.run_reference(s, f"/tmp/agb_golden_{name}", threads=1)
        out = dict(
            natoms=n, seed=seed, eps=np.array([s.eps_in, s.eps_out]),
            x=s.x, y=s.y, z=s.z, r=s.r, charge=s.charge, igamma=s.igamma, sgamma=s.sgamma,
            ialpha=s.ialpha, salpha=s.salpha, hbtype=s.hbtype, hbcorr=s.hbcorr,
            ihydrogen=s.ihydrogen, conn_nne=s.conn_nne, conn_neighl1=s.conn_neighl1,
            energies=np.array([res[k] for k in oracle.ENERGY_NAMES]),
            sp=res["sp"], br=res["br"], surf_area=res["surf_area"],
            dgbdr=res["dgbdr"], dvwdr=res["dvwdr"], decav=res["decav"], dehb=res["dehb"],
            ref_sha256=np.array(digest), ref_flags=np.array("gcc -O3 -fPIC -fopenmp -DUSE_SSE; OMP_NUM_THREADS=1"),
        )
        np.savez_compressed(os.path.join(HERE, f"golden_{name}.npz"), **out)
        hb = {int(t): int((s.hbtype == t).sum()) for t in np.unique(s.hbtype)}
        print(name, "written; hbtypes", hb, "E", dict(zip(oracle.ENERGY_NAMES, np.round(out["energies"], 6))))


if __name__ == "__main__":
    main()
