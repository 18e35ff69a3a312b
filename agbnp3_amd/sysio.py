"""Binary system files shared by the C drivers (layout read by oracle/ref_driver.c)."""
import os

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


# ------------------------------------------------------------ text dumps
# The reference's input dump (compiled with AGBNP_WRITE_DATA, agbnp3.c:319-336):
#   natoms
#   iat x y z R charge igamma sgamma ialpha salpha idelta sdelta hbtype hbcorr
#     (one line per atom in internal order: heavy atoms first, then
#      hydrogens, each in caller order; R = radius + 0.5; printf "%f")
#   nhydrogen
#   ihydrogen[k]   (one line each, caller indices)
# It holds no connection table; `read_agbnp_dump` either takes the bonds or
# infers them from distances (see `infer_bonds`).

def write_agbnp_dump(path, s):
    """Write system `s` (caller order) in the reference's dump format."""
    from . import synth  # noqa: F401  (System type)
    n = s.natoms
    ish = np.zeros(n, bool)
    ish[s.ihydrogen] = True
    int2ext = np.concatenate([np.flatnonzero(~ish), np.flatnonzero(ish)])
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for k, e in enumerate(int2ext):
            f.write("%d %f %f %f %f %f %f %f %f %f %f %f %d %f\n" % (
                k, s.x[e], s.y[e], s.z[e], s.r[e] + 0.5, s.charge[e], s.igamma[e], s.sgamma[e], s.ialpha[e],
                s.salpha[e], 0.0, 0.0, int(s.hbtype[e]), s.hbcorr[e]))
        f.write(f"{s.nhydrogen}\n")
        for h in s.ihydrogen:
            f.write(f"{int(h)}\n")


def infer_bonds(x, y, z, is_h, heavy_cut=1.95, sulfur_cut=2.15, h_cut=1.3, r=None):
    """Covalent bonds from distances: each hydrogen to its nearest heavy atom
    within `h_cut`; heavy pairs closer than `heavy_cut` (`sulfur_cut` when
    either radius is 1.75 A or more, as for sulfur).  Returns (nne, flat)
    in NeighList layout, neighbours in increasing index order."""
    n = len(x)
    p = np.stack([x, y, z], 1)
    heavy = np.flatnonzero(~is_h)
    nbrs = [[] for _ in range(n)]
    ph = p[heavy]
    for a in range(n):
        d = np.sqrt(((ph - p[a]) ** 2).sum(1))
        if is_h[a]:
            k = int(np.argmin(d)) if heavy.size else -1
            if k >= 0 and d[k] < h_cut:
                nbrs[a].append(int(heavy[k]))
                nbrs[int(heavy[k])].append(a)
        else:
            for k in np.flatnonzero(d > 0.0):
                b = int(heavy[k])
                if b <= a:
                    continue
                cut = heavy_cut
                if r is not None and max(r[a], r[b]) >= 1.75:
                    cut = sulfur_cut
                if d[k] < cut:
                    nbrs[a].append(b)
                    nbrs[b].append(a)
    nne = np.array([len(v) for v in nbrs], np.int32)
    flat = np.array([b for v in nbrs for b in sorted(v)], np.int32)
    return nne, flat


def read_agbnp_dump(path, bonds=None, eps_in=1.0, eps_out=80.0):
    """A System (caller order) from the reference's dump.  `bonds`: optional
    (nne, flat) connection table in caller order; otherwise inferred."""
    from .synth import System
    with open(path) as f:
        tok = f.read().split()
    n = int(tok[0])
    rows = np.array(tok[1:1 + 14 * n], dtype=object).reshape(n, 14)
    pos = 1 + 14 * n
    nh = int(tok[pos])
    ihyd = np.array([int(t) for t in tok[pos + 1:pos + 1 + nh]], np.int32)
    ish = np.zeros(n, bool)
    ish[ihyd] = True
    int2ext = np.concatenate([np.flatnonzero(~ish), np.flatnonzero(ish)])
    vals = rows[:, 1:12].astype(np.float64)
    out = np.zeros((n, 11))
    out[int2ext] = vals
    hbt = np.zeros(n, np.int32)
    hbt[int2ext] = rows[:, 12].astype(np.int64)
    hbc = np.zeros(n)
    hbc[int2ext] = rows[:, 13].astype(np.float64)
    x, y, z, R = out[:, 0], out[:, 1], out[:, 2], out[:, 3]
    r = R - 0.5
    nne, flat = bonds if bonds is not None else infer_bonds(x, y, z, ish, r=r)
    return System(natoms=n, x=x.copy(), y=y.copy(), z=z.copy(), r=r, charge=out[:, 4].copy(), eps_in=eps_in,
                  eps_out=eps_out, igamma=out[:, 5].copy(), sgamma=out[:, 6].copy(), ialpha=out[:, 7].copy(),
                  salpha=out[:, 8].copy(), hbtype=hbt, hbcorr=hbc, ihydrogen=ihyd,
                  conn_nne=np.asarray(nne, np.int32), conn_neighl1=np.asarray(flat, np.int32),
                  info=dict(source=os.path.abspath(path)))
