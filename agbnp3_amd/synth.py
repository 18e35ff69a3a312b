"""Seeded synthetic solutes (the globular-solute recipe) via libagbnp3_synth.so.

`generate(seed, natoms)` returns a `System`: the arguments of agbnp3_new()
as numpy arrays (caller atom order) plus the connection table in NeighList
layout (`conn_nne` counts and the flat `conn_neighl1` list).
"""
import ctypes
import dataclasses
import os

import numpy as np

from . import _paths

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_paths.LIBSYNTH):
            raise RuntimeError(f"{_paths.LIBSYNTH} is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(_paths.LIBSYNTH)
        _lib.agbsynth_generate.restype = ctypes.c_int
        _lib.agbsynth_generate.argtypes = [ctypes.c_uint64, ctypes.c_int] + [ctypes.c_void_p] * 14 + [
            ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]
        _lib.agbsynth_perturb.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double] + [ctypes.c_void_p] * 3
        _lib.agbsynth_sizes.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclasses.dataclass
class System:
    natoms: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    r: np.ndarray
    charge: np.ndarray
    eps_in: float
    eps_out: float
    igamma: np.ndarray
    sgamma: np.ndarray
    ialpha: np.ndarray
    salpha: np.ndarray
    hbtype: np.ndarray
    hbcorr: np.ndarray
    ihydrogen: np.ndarray
    conn_nne: np.ndarray
    conn_neighl1: np.ndarray
    info: dict = dataclasses.field(default_factory=dict)

    @property
    def nhydrogen(self):
        return int(self.ihydrogen.size)

    def with_coords(self, x, y, z):
        return dataclasses.replace(self, x=np.ascontiguousarray(x, np.float64),
                                   y=np.ascontiguousarray(y, np.float64),
                                   z=np.ascontiguousarray(z, np.float64))


def generate(seed, natoms, eps_in=1.0, eps_out=80.0):
    lib = _load()
    n = int(natoms)
    d = [np.zeros(n, np.float64) for _ in range(10)]
    x, y, z, r, q, ig, sg, ia, sa, hc = d
    hb = np.zeros(n, np.int32)
    ish = np.zeros(n, np.int32)
    nne = np.zeros(n, np.int32)
    cap = 8 * n + 16
    nl = np.zeros(cap, np.int32)
    nc = ctypes.c_int(0)
    info = np.zeros(5, np.float64)
    rc = lib.agbsynth_generate(ctypes.c_uint64(int(seed) & (2**64 - 1)), n, _p(x), _p(y), _p(z), _p(r), _p(q),
                               _p(ig), _p(sg), _p(ia), _p(sa), _p(hb), _p(hc), _p(ish), _p(nne), _p(nl), cap,
                               ctypes.byref(nc), _p(info))
    if rc != 0:
        raise RuntimeError(f"agbsynth_generate(seed={seed}, natoms={n}) failed with {rc}")
    return System(natoms=n, x=x, y=y, z=z, r=r, charge=q, eps_in=float(eps_in), eps_out=float(eps_out),
                  igamma=ig, sgamma=sg, ialpha=ia, salpha=sa, hbtype=hb, hbcorr=hc,
                  ihydrogen=np.nonzero(ish)[0].astype(np.int32), conn_nne=nne,
                  conn_neighl1=nl[:nc.value].copy(),
                  info=dict(seed=int(seed), sphere_radius=float(info[0]), min_nonbonded=float(info[1]),
                            backtracks=int(info[2]), nheavy=int(info[3]), segments=int(info[4])))


def perturb(seed, x, y, z, sigma):
    """Return perturbed copies: coordinates + N(0, sigma) from a seeded stream."""
    lib = _load()
    x, y, z = (np.array(a, np.float64, copy=True) for a in (x, y, z))
    lib.agbsynth_perturb(ctypes.c_uint64(int(seed)), x.size, float(sigma), _p(x), _p(y), _p(z))
    return x, y, z


def sizes(seed, ninst, nmin, nmax):
    lib = _load()
    out = np.zeros(int(ninst), np.int32)
    lib.agbsynth_sizes(ctypes.c_uint64(int(seed)), int(ninst), int(nmin), int(nmax), _p(out))
    return out
