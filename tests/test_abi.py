"""The C-ABI boundary: headers, exported symbols, NeighList layout, host-side
table builder (CPU only, no GPU calls)."""
import ctypes
import os
import re


import numpy as np

from agbnp3_amd import _paths, api


def declared_functions(header):
    text = open(os.path.join(_paths.INCLUDE_DIR, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|void|void \*)\s*\*?\s*(\w+)\s*\(", text, flags=re.M))


def test_library_exports_every_declared_symbol():
    api.load()
    lib = ctypes.CDLL(_paths.LIBAGBNP3)
    nb = ctypes.CDLL(_paths.LIBNBLIST)
    decl = declared_functions("agbnp3.h") | declared_functions("agbnp3_amd.h")
    assert {"agbnp3_initialize", "agbnp3_new", "agbnp3_ener", "agbnp3_delete", "agbnp3_terminate",
            "agbnp_initialize", "agbnp_terminate", "agbnp_delete"} <= decl
    for name in decl:
        assert hasattr(lib, name), name
    assert set(api.EXPORTED) == decl
    for name in declared_functions("nblist.h"):
        assert hasattr(nb, name), name
    assert set(api.NBLIST_EXPORTED) == declared_functions("nblist.h")


def test_single_precision_interface_flavour_exports_the_same_symbols():
    api.load()
    lib = ctypes.CDLL(_paths.LIBAGBNP3_SP)
    for name in declared_functions("agbnp3.h") | declared_functions("agbnp3_amd.h"):
        assert hasattr(lib, name), name


def test_library_is_built_for_gfx950():
    # the offload bundle of our own build names its target triple
    blob = open(_paths.LIBAGBNP3, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_neighlist_struct_matches_c_layout():
    # nblist.h:39-70 on LP64: 2 ints, 3 pointers, int, 2 pointers, int, 2 pointers, int, 2 pointers
    assert ctypes.sizeof(api.NeighList) == 8 + 24 + 8 + 16 + 8 + 16 + 8 + 16
    assert api.NeighList.neighl.offset == 24
    assert api.NeighList.data_index.offset == ctypes.sizeof(api.NeighList) - 8


def test_nblist_roundtrip_on_cpu():
    import numpy as np
    nne = np.array([2, 1, 1], np.int32)
    flat = np.array([1, 2, 0, 0], np.int32)
    nl = api.make_conntbl(nne, flat)
    assert nl.natoms == 3 and nl.neighl_size == 4
    assert [nl.neighl[0][k] for k in range(2)] == [1, 2]
    assert nl.neighl[2][0] == 0
    api._nb.nblist_delete_neighbor_list(ctypes.byref(nl))


def test_host_q4_tables_are_bit_identical_to_the_oracle():
    import oracle
    from conftest import GOLDEN_CASES, load_golden
    for case in GOLDEN_CASES:
        s, _ = load_golden(case)
        yo, y2o, ro = oracle.Oracle(s).tables()
        yg, y2g, dx, dxi = api.q4_tables(ro)
        assert np.array_equal(yg, yo), case
        assert np.array_equal(y2g, y2o), case
        assert dx == np.float32(20.0) / np.float32(511)


def test_distance_tests_are_compiled_without_contraction(tmp_path):
    """The neighbour-list distance test must round like the reference's
    separate FP32 multiplies and adds: the device code of the kernel may hold
    no fused multiply-add (it has no other floating-point work)."""
    import subprocess
    from agbnp3_amd import build
    for src, kernels in (("k_setup.hip", ("k_nblist",)),):
        out = tmp_path / (src + ".s")
        cmd = ([build.HIPCC] + build.HIP_FLAGS + build.HIP_FILE_FLAGS.get(src, []) +
               [f"-I{_paths.INCLUDE_DIR}", f"-I{_paths.CSRC_DIR}", "--cuda-device-only", "-S",
                os.path.join(_paths.CSRC_DIR, src), "-o", str(out)])
        subprocess.run(cmd, check=True, capture_output=True)
        asm = out.read_text()
        for k in kernels:
            bodies = [b for b in asm.split("\n_ZN3agb")[1:] if b.split("\n", 1)[0].lstrip("0123456789").startswith(k)]
            assert bodies, (src, k)
            for b in bodies:
                body = b.split("s_endpgm", 1)[0]
                assert "v_fma" not in body and "v_fmac" not in body, (src, k)
