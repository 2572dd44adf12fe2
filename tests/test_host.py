"""The product library on the CPU: it loads, exports every symbol its header declares, and its
host-side md-core (initial conditions, hash, list capacity) equals the oracle's bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

import micrographia_amd as rh
from micrographia_amd import _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()
    if not os.path.exists(_lib.LIB_PATH):
        from micrographia_amd import build

        build.build()


def header_symbols():
    text = open(os.path.join(ROOT, "include", "rahman_hip.h")).read()
    return sorted(set(re.findall(r"\b(rh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_error_reporting():
    lib = _lib.load()
    assert b"gfx950" in lib.rh_version()
    with pytest.raises(rh.RahmanError):
        _lib.check(lib.rh_engine3_step(None, 1))
    assert b"null engine" in lib.rh_last_error()


@pytest.mark.parametrize("n,density,t,seed", [(4000, 0.8, 1.0, 3), (262144, 0.8442, 0.72, 1), (1500, 0.5, 1.3, 9)])
def test_init3_matches_the_oracle(n, density, t, seed):
    box = rh.cube_for(n, density)
    assert list(box) == list(O.cube_for(n, density))
    cells = rh.fcc_cells(n, box)
    assert list(cells) == list(O.fcc_cells(n, box))
    pos = rh.fcc(n, box, cells, 0.05, seed)
    assert np.array_equal(pos, O.fcc(n, box, cells, 0.05, seed))
    sys = rh.System3.new(box, 0.005, rh.Pair.lj(2.5), [1.0])
    sp = np.zeros(n, np.uint32)
    vel = rh.maxwell(sys, sp, t, seed)
    osys = O.system3(box)
    assert np.array_equal(vel, O.maxwell3(osys, sp, t, seed))
    st = rh.State3(pos, vel, sp)
    assert st.hash() == O.hash3(pos, vel, sp)


def test_init3_mixture_with_pinned_species():
    n = 3000
    box = rh.cube_for(n, 0.6)
    sys = rh.System3.new(box, 0.004, rh.Pair.lj(2.5), [1.0, 2.5, 0.7]).with_pinned(2)
    sp = (np.arange(n) % 3).astype(np.uint32)
    vel = rh.maxwell(sys, sp, 1.1, 17)
    osys = O.system3(box, dt=0.004, masses=(1.0, 2.5, 0.7))
    osys.pinned = 1 << 2
    assert np.array_equal(vel, O.maxwell3(osys, sp, 1.1, 17))
    st = rh.State3(rh.fcc(n, box, rh.fcc_cells(n, box), 0.05, 2), vel, sp)
    rh.set_temperature(st, sys, 0.6)
    ovel = O.maxwell3(osys, sp, 1.1, 17)
    O.set_temperature3(osys, ovel, sp, 0.6)
    assert np.array_equal(st.vel, ovel)


def test_init2_and_the_golden_setup_match_the_oracle():
    # crates/md/src/checks.rs:156-164, built with the product's own md-core
    box = rh.box_for(3000, 0.6, 1.0)
    sys = rh.System.new(box, 0.005, rh.Pair.lj(2.5), [1.0, 2.0]).with_pair(0, 1, rh.Pair(0.7, 1.05, 2.5))
    st = rh.fluid(sys, 3000, 1.0, 2026)
    st.species[::4] = 1
    st.vel = rh.maxwell2(sys, st.species, 1.0, 2026)
    osys, opos, ovel, osp = O.golden_setup()
    assert list(box) == [osys.box_len[0], osys.box_len[1]]
    assert np.array_equal(st.pos, opos)
    assert np.array_equal(st.vel, ovel)
    assert np.array_equal(st.species, osp)
    assert st.hash() == O.hash2(opos, ovel, osp)


def test_verlet_capacity_rule():
    """engine3.rs:61-68 without the 128 MiB clamp: 4M atoms keep a full-size list."""
    n = 4194304
    sys = rh.System3.new(rh.cube_for(n, 0.8442), 0.005, rh.Pair.lj(2.5), [1.0])
    m = rh.verlet3_with_skin(sys, n, 0.3)
    r = 2.8
    want = int(np.ceil(4 / 3 * np.pi * r ** 3 * max(2 * 0.8442, 1.2))) + 8
    assert m.kind == 1 and m.cap in (want, want + 1) and abs(m.skin - 0.3) < 1e-15


def test_division_floor_bounds():
    import ctypes

    f = ctypes.c_float()
    _lib.check(_lib.load().rh_pair_division_floor(1.0, 1.0, 2.5, ctypes.byref(f)))
    assert 0 < f.value < 0.01  # r below 0.1 sigma falls back to IEEE division
    _lib.check(_lib.load().rh_pair_division_floor(1e-35, 1.0, 2.5, ctypes.byref(f)))
    assert f.value == float("inf")  # tiny forces: never the fast path


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every field of the ctypes mirrors sits at the header's offset (gcc's offsetof), and the
    structs have the header's size, so no call reads past or beside what Python passes."""
    import shutil
    import subprocess

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    structs = {"rh_system2": _lib.CSystem2, "rh_system3": _lib.CSystem3, "rh_mode": _lib.CMode,
               "rh_ledger2": _lib.CLedger2, "rh_ledger3": _lib.CLedger3, "rh_list_stats": _lib.CListStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rahman_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(__file__), "..", "include")
    subprocess.run([gcc, "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in out:
        if not line:
            continue
        cname, fname, value = line.split()
        py = structs[cname]
        got = ctypes.sizeof(py) if fname == "sizeof" else getattr(py, fname).offset
        assert got == int(value), (cname, fname, got, value)
        seen += 1
    assert seen == sum(len(p._fields_) + 1 for p in structs.values())
