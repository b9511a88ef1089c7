"""CPU checks of the product library: it loads and exports every function include/lmff_hip.h
declares (no compute calls: there is no GPU here), and the host-side helpers."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import REPO
from lmff import capi, structures as S

HEADER = os.path.join(REPO, "include", "lmff_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(lmff_[a-z_0-9]+)\s*\(", txt)
    return sorted(set(names))


def test_header_lists_the_binding():
    assert declared_functions() == sorted(capi.SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    lib.lmff_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.lmff_version()


def test_no_fallback_when_library_missing(tmp_path):
    old = capi._lib
    capi._lib = None
    try:
        with pytest.raises(RuntimeError):
            capi.load(str(tmp_path / "missing.so"))
    finally:
        capi._lib = old


def test_generators_shapes():
    s = S.config_c1()
    assert s.n == 4800 and s.boxhi[0] == pytest.approx(73.8) and s.boxhi[1] == pytest.approx(85.22)
    assert abs(S.temperature_of(s, s.v) - 300.0) < 1e-9
    assert np.abs((s.v * s.masses()[:, None]).sum(0)).max() < 1e-9
    t = S.twisted_bilayer(7, 3, seed=1)
    assert 2000 <= t.n <= 80000 and min(t.boxhi[:2] - t.boxlo[:2]) >= 36.5
    assert abs(S.twist_angle(2, 1) - np.radians(21.786789298261816)) < 1e-12
    # every atom of a perfect twisted cell has three intralayer neighbours at 1.42 A
    t0 = S.twisted_bilayer(3, 2, jitter=0.0, temperature=None)
    d = t0.x[:, None, :2] - t0.x[None, :, :2]
    L = (t0.boxhi - t0.boxlo)[:2]
    d -= L * np.round(d / L)
    same = t0.mol[:, None] == t0.mol[None, :]
    r = np.sqrt((d ** 2).sum(-1))
    nb = ((r < 1.6) & (r > 0.1) & same).sum(1)
    assert (nb == 3).all()


def test_write_data_round_trip(tmp_path):
    """write_lammps_data -> read_lammps_data gives back the same atoms, bit for bit."""
    from lmff import structures as S

    s = S.layered_stack(15, 9, ["G", "BN"], seed=8)
    path = str(tmp_path / "d.data")
    S.write_lammps_data(s, path, step=120)
    r = S.read_lammps_data(path)
    assert open(path).readline().strip().endswith("timestep = 120")
    assert r.n == s.n
    o = np.argsort(s.tag)
    assert np.array_equal(r.tag, s.tag[o]) and np.array_equal(r.mol, s.mol[o]) and np.array_equal(r.type, s.type[o])
    assert np.array_equal(r.x, s.x[o]) and np.array_equal(r.v, s.v[o])
    assert np.array_equal(r.boxlo, s.boxlo) and np.array_equal(r.boxhi, s.boxhi)
    assert r.meta["masses"] == {t + 1: float(f"{m:g}") for t, m in enumerate(s.mass_per_type)}
