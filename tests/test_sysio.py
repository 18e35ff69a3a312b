"""The reference's text input dump (AGBNP_WRITE_DATA, agbnp3.c:319-336) as an
input path: writer/reader round trip, bond inference, and the oracle on a
system read back from a dump."""
import numpy as np

import oracle
from agbnp3_amd import synth, sysio


def conn_sets(nne, flat):
    out, o = [], 0
    for c in nne:
        out.append(set(int(v) for v in flat[o:o + c]))
        o += c
    return out


def test_dump_round_trip(tmp_path):
    s = synth.generate(20260001, 400)
    p = tmp_path / "sys.dump"
    sysio.write_agbnp_dump(str(p), s)
    lines = p.read_text().splitlines()
    assert int(lines[0]) == s.natoms and int(lines[1 + s.natoms]) == s.nhydrogen
    # internal order: heavy atoms first, each group in caller order
    first = lines[1].split()
    heavy0 = int(np.flatnonzero(~np.isin(np.arange(s.natoms), s.ihydrogen))[0])
    assert abs(float(first[1]) - s.x[heavy0]) < 1e-6 and abs(float(first[4]) - (s.r[heavy0] + 0.5)) < 1e-6
    t = sysio.read_agbnp_dump(str(p))
    for k in ("x", "y", "z", "r", "charge", "igamma", "sgamma", "ialpha", "salpha", "hbcorr"):
        assert np.allclose(getattr(t, k), getattr(s, k), atol=1e-6, rtol=0), k
    assert np.array_equal(t.hbtype, s.hbtype)
    assert np.array_equal(t.ihydrogen, s.ihydrogen)
    # the generator's bonds are recovered from distances alone
    assert conn_sets(t.conn_nne, t.conn_neighl1) == conn_sets(s.conn_nne, s.conn_neighl1)


def test_system_from_dump_gives_the_same_energies(tmp_path):
    s = synth.generate(20260007, 300)
    p = tmp_path / "sys.dump"
    sysio.write_agbnp_dump(str(p), s)
    t = sysio.read_agbnp_dump(str(p), bonds=(s.conn_nne, s.conn_neighl1))
    a, b = oracle.Oracle(s).ener(), oracle.Oracle(t).ener()
    for k in oracle.ENERGY_NAMES:
        assert abs(a[k] - b[k]) <= 1e-4 * max(1.0, abs(a[k])), k
