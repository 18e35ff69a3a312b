"""Binary system files shared by the C drivers (layout read by oracle/ref_driver.c)."""
import numpy as np

MAGIC = 0x41474233


def write_system(path, s):
    n = s.natoms
    with open(path, "wb") as f:
        np.array([MAGIC, n, s.nhydrogen, s.conn_neighl1.size], np.int32).tofile(f)
        np.array([s.eps_in, s.eps_out], np.float64).tofile(f)
        for a in (s.x, s.y, s.z, s.r, s.charge, s.igamma, s.sgamma, s.ialpha, s.salpha, s.hbcorr):
            np.ascontiguousarray(a, np.float64).tofile(f)
        for a in (s.hbtype, s.ihydrogen, s.conn_nne, s.conn_neighl1):
            np.ascontiguousarray(a, np.int32).tofile(f)


ENERGY_NAMES = ("egb", "evdw", "ecorr_vdw", "ecav", "ecorr_cav", "ehb", "mol_volume")


def read_results(path, natoms):
    """Outputs written by the reference driver: energies and per-atom arrays."""
    a = np.fromfile(path, np.float64)
    n = natoms
    out = dict(zip(ENERGY_NAMES, a[:7]))
    o = 7
    for k in ("sp", "br", "surf_area"):
        out[k] = a[o:o + n]
        o += n
    for k in ("dgbdr", "dvwdr", "decav", "dehb"):
        out[k] = a[o:o + 3 * n].reshape(n, 3)
        o += 3 * n
    return out
