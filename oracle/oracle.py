"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module; the product path never does.  The C source
(``lmff_oracle.c``) restates the reference's Tersoff + ILP force path in the
reference's own push form; see its header for the file:line map.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblmff_oracle.so")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_ll = C.POINTER(C.c_longlong)


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "lmff_oracle.c")):
        subprocess.run(["make", "-s", "-C", HERE, "-B"] if force else ["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_int,
                                 C.c_double, C.c_char_p, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_error.restype = C.c_char_p
        L.orc_error.argtypes = [C.c_void_p]
        L.orc_cutmax.restype = C.c_double
        L.orc_cutmax.argtypes = [C.c_void_p]
        L.orc_cutneighsq.restype = C.c_double
        L.orc_cutneighsq.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_tersoff_param.argtypes = [C.c_void_p, C.c_int, _d]
        L.orc_ilp_param.argtypes = [C.c_void_p, C.c_int, _d]
        L.orc_set_atoms.argtypes = [C.c_void_p, C.c_int, _d, _i, _ll, _ll, _d, _d, _i]
        L.orc_get_x.argtypes = [C.c_void_p, _d]
        L.orc_put_x.argtypes = [C.c_void_p, _d]
        L.orc_pbc.argtypes = [C.c_void_p]
        L.orc_build_list.argtypes = [C.c_void_p]
        L.orc_list_size.restype = C.c_longlong
        L.orc_list_size.argtypes = [C.c_void_p]
        L.orc_get_list.argtypes = [C.c_void_p, _i, _i, _i]
        L.orc_compute.argtypes = [C.c_void_p, C.c_int, C.c_int, _d, _d, _d, _d, _d]
        L.orc_normals.argtypes = [C.c_void_p, _d, _i, _d, _d]
        L.orc_run_nve.argtypes = [C.c_void_p, _d, _d, _d, C.c_double, C.c_int, C.c_int, C.c_int, _d]
        L.orc_tap.restype = C.c_double
        L.orc_tap.argtypes = [C.c_double, C.c_double]
        L.orc_dtap.restype = C.c_double
        L.orc_dtap.argtypes = [C.c_double, C.c_double]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleError(RuntimeError):
    pass


class Oracle:
    """One oracle instance = one set of pair coefficients + one system."""

    def __init__(self, ntypes: int, tersoff_file: str, tersoff_elems, ilp_file: str, ilp_elems,
                 ilp_cut: float = 16.0, tap_flag: int = 1, skin: float = 2.0, threads: int = 1):
        L = lib()
        err = C.create_string_buffer(512)
        self.h = L.orc_create(ntypes, tersoff_file.encode(), " ".join(tersoff_elems).encode(), ilp_file.encode(),
                              " ".join(ilp_elems).encode(), float(ilp_cut), int(tap_flag), float(skin), err, 512)
        if not self.h:
            raise OracleError(err.value.decode())
        self.ntypes = ntypes
        L.orc_set_threads(self.h, int(threads))
        self.n = 0

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise OracleError(lib().orc_error(self.h).decode())
        return rc

    @property
    def cutmax(self) -> float:
        return lib().orc_cutmax(self.h)

    def set_threads(self, n: int):
        lib().orc_set_threads(self.h, int(n))

    def set_system(self, s):
        self.n = s.n
        self.x = np.ascontiguousarray(s.x, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        tag = np.ascontiguousarray(s.tag, dtype=np.int64)
        mol = np.ascontiguousarray(s.mol, dtype=np.int64)
        lo = np.ascontiguousarray(s.boxlo, dtype=np.float64)
        hi = np.ascontiguousarray(s.boxhi, dtype=np.float64)
        per = np.ascontiguousarray(s.periodic, dtype=np.int32)
        self._keep = (t, tag, mol, lo, hi, per)
        self._check(lib().orc_set_atoms(self.h, self.n, _p(self.x, C.c_double), _p(t, C.c_int), _p(tag, C.c_longlong),
                                        _p(mol, C.c_longlong), _p(lo, C.c_double), _p(hi, C.c_double), _p(per, C.c_int)))

    def pbc(self):
        lib().orc_pbc(self.h)

    def get_x(self):
        x = np.zeros((self.n, 3))
        lib().orc_get_x(self.h, _p(x, C.c_double))
        return x

    def put_x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib().orc_put_x(self.h, _p(x, C.c_double))

    def build_list(self):
        self._check(lib().orc_build_list(self.h))

    def neighbor_list(self):
        ne = lib().orc_list_size(self.h)
        offs = np.zeros(self.n + 1, dtype=np.int32)
        ent = np.zeros(max(ne, 1), dtype=np.int32)
        sh = np.zeros((max(ne, 1), 3), dtype=np.int32)
        lib().orc_get_list(self.h, _p(offs, C.c_int), _p(ent, C.c_int), _p(sh, C.c_int))
        return offs, ent[:ne], sh[:ne]

    def compute(self, eflag: int = 1, vflag: int = 2):
        n = self.n
        f = np.zeros((n, 3))
        eng = np.zeros(4)
        vir = np.zeros(6)
        eatom = np.zeros(n) if eflag & 2 else None
        vatom = np.zeros((n, 6)) if vflag & 4 else None
        null = C.POINTER(C.c_double)()
        self._check(lib().orc_compute(self.h, eflag, vflag, _p(f, C.c_double), _p(eng, C.c_double), _p(vir, C.c_double),
                                      _p(eatom, C.c_double) if eatom is not None else null,
                                      _p(vatom, C.c_double) if vatom is not None else null))
        return {"f": f, "energy": eng[0], "e_vdw": eng[1], "e_rep": eng[2], "e_tersoff": eng[3], "virial": vir,
                "eatom": eatom, "vatom": vatom}

    def normals(self):
        n = self.n
        nrm = np.zeros((n, 3))
        cnt = np.zeros(n, dtype=np.int32)
        dn = np.zeros((n, 3, 3, 3))
        dri = np.zeros((n, 3, 3))
        self._check(lib().orc_normals(self.h, _p(nrm, C.c_double), _p(cnt, C.c_int), _p(dn, C.c_double),
                                      _p(dri, C.c_double)))
        return nrm, cnt, dn, dri

    def energy_forces(self, x=None):
        """Convenience: builds the list from the given (or current) positions and evaluates."""
        if x is not None:
            self.put_x(x)
        self.build_list()
        return self.compute(1, 2)

    def run_nve(self, x, v, mass_per_type, dt: float, nsteps: int, rebuild_every: int = 10, thermo_every: int = 10):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        m = np.ascontiguousarray(mass_per_type, dtype=np.float64)
        nth = nsteps // max(thermo_every, 1) + 2
        th = np.zeros((nth, 3))
        k = self._check(lib().orc_run_nve(self.h, _p(x, C.c_double), _p(v, C.c_double), _p(m, C.c_double), float(dt),
                                          int(nsteps), int(rebuild_every), int(thermo_every), _p(th, C.c_double)))
        return x, v, th[:k]


def tap(r: float, rcut: float) -> float:
    return lib().orc_tap(r, rcut)


def dtap(r: float, rcut: float) -> float:
    return lib().orc_dtap(r, rcut)
