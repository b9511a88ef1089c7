"""ctypes binding of liblmff_hip.so (the C ABI declared in include/lmff_hip.h).

The library is the only compute path: if it is missing or cannot be loaded this
module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "liblmff_hip.so")
POTENTIALS = os.path.join(PKG, "potentials")

LMFF_FP64 = 0
LMFF_MIXED = 1
LMFF_NVE = 0
LMFF_LANGEVIN = 1

ERRORS = {1: "ARG", 2: "FILE", 3: "PARAM", 4: "HIP", 5: "NORMAL", 6: "SHORT", 7: "NONFINITE", 8: "MEMORY",
          9: "BOX", 10: "STATE"}

SYMBOLS = [
    "lmff_create", "lmff_destroy", "lmff_last_error", "lmff_version", "lmff_settings", "lmff_coeff_tersoff",
    "lmff_coeff_ilp", "lmff_neighbor", "lmff_precision", "lmff_masses", "lmff_pair_compute", "lmff_system_add",
    "lmff_system_remove", "lmff_system_integrator", "lmff_system_natoms", "lmff_system_get", "lmff_system_set",
    "lmff_system_peratom", "lmff_evaluate", "lmff_system_thermo", "lmff_run", "lmff_profile",
    "lmff_profile_read", "lmff_memory", "lmff_system_bytes", "lmff_debug_ilp_list", "lmff_debug_normals",
    "lmff_pair_stats", "lmff_system_set_step", "lmff_peratom_on_last_step", "lmff_fire_defaults",
    "lmff_minimize", "lmff_minimize_cg", "lmff_system_get_image", "lmff_system_set_image",
    "lmff_pair_compute_noalloc", "lmff_noalloc_scratch_bytes", "lmff_tersoff_cutmax",
    "lmff_neighbor_stats",
]

STOP = {0: "max iterations", 1: "max force evaluations", 2: "energy tolerance", 3: "force tolerance",
        4: "search direction is not downhill", 5: "linesearch alpha is zero", 6: "forces are zero",
        7: "quadratic factors are zero", 11: "max iterations with v.f negative"}


class FireParams(C.Structure):
    _fields_ = [("dmax", C.c_double), ("delaystep", C.c_int), ("dtgrow", C.c_double), ("dtshrink", C.c_double),
                ("alpha0", C.c_double), ("alphashrink", C.c_double), ("tmax", C.c_double), ("tmin", C.c_double),
                ("halfstepback", C.c_int), ("delaystep_start", C.c_int), ("max_vdotf_negatif", C.c_int),
                ("norm", C.c_int)]


class MinStats(C.Structure):
    _fields_ = [("stop", C.c_int), ("iterations", C.c_int64), ("evaluations", C.c_int64),
                ("e_initial", C.c_double), ("e_next_to_last", C.c_double), ("e_final", C.c_double),
                ("fnorm2_initial", C.c_double), ("fnorm2_final", C.c_double),
                ("fnorminf_initial", C.c_double), ("fnorminf_final", C.c_double), ("dt_final", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["stop_text"] = STOP.get(self.stop, str(self.stop))
        return d


class LmffError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"LMFF_ERR_{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Thermo(C.Structure):
    _fields_ = [("step", C.c_int64), ("pe", C.c_double), ("ke", C.c_double), ("e_vdw", C.c_double),
                ("e_rep", C.c_double), ("e_tersoff", C.c_double), ("temp", C.c_double),
                ("virial", C.c_double * 6)]

    def as_dict(self):
        return {"step": self.step, "pe": self.pe, "ke": self.ke, "e_vdw": self.e_vdw, "e_rep": self.e_rep,
                "e_tersoff": self.e_tersoff, "temp": self.temp, "virial": list(self.virial)}


_lib = None
_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_int = C.POINTER(C.c_int)


def load(path: str = LIB_PATH):
    """Loads the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run python lammps-lmff_amd/build.py")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.lmff_create.argtypes = [C.POINTER(vp), C.c_int]
    L.lmff_destroy.argtypes = [vp]
    L.lmff_destroy.restype = None
    L.lmff_last_error.argtypes = [vp]
    L.lmff_last_error.restype = C.c_char_p
    L.lmff_version.restype = C.c_char_p
    L.lmff_settings.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.lmff_coeff_tersoff.argtypes = [vp, C.c_char_p, C.POINTER(C.c_char_p)]
    L.lmff_coeff_ilp.argtypes = [vp, C.c_char_p, C.POINTER(C.c_char_p)]
    L.lmff_neighbor.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int]
    L.lmff_precision.argtypes = [vp, C.c_int]
    L.lmff_tersoff_cutmax.argtypes = [vp, _d]
    L.lmff_neighbor_stats.argtypes = [vp, C.c_int, _i64, _i64]
    L.lmff_masses.argtypes = [vp, _d]
    L.lmff_pair_compute.argtypes = [vp, C.c_int64, _d, _i32, _i64, _i64, _d, _d, _int, C.c_int, C.c_int, C.c_int,
                                    _d, _d, _d, _d, _d, _d]
    # device pointers (the allocation-free entry point takes device arrays)
    L.lmff_pair_compute_noalloc.argtypes = [vp, C.c_int64, vp, vp, vp, vp, _d, _d, _int, C.c_int, C.c_int, vp, vp]
    L.lmff_noalloc_scratch_bytes.restype = C.c_int64
    L.lmff_noalloc_scratch_bytes.argtypes = []
    L.lmff_system_add.argtypes = [vp, C.c_int64, _d, _d, _int, _i32, _i64, _i64, _d, _d, _int]
    L.lmff_system_remove.argtypes = [vp, C.c_int]
    L.lmff_system_integrator.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.lmff_system_natoms.argtypes = [vp, C.c_int, _i64]
    L.lmff_system_get.argtypes = [vp, C.c_int, _d, _d, _d]
    L.lmff_system_set.argtypes = [vp, C.c_int, _d, _d]
    L.lmff_system_get_image.argtypes = [vp, C.c_int, _i32]
    L.lmff_system_set_image.argtypes = [vp, C.c_int, _i32]
    L.lmff_system_peratom.argtypes = [vp, C.c_int, _d, _d]
    L.lmff_evaluate.argtypes = [vp, C.c_int, C.c_int]
    L.lmff_system_thermo.argtypes = [vp, C.c_int, C.POINTER(Thermo)]
    L.lmff_run.argtypes = [vp, C.c_int64, C.c_int64, C.c_int, C.POINTER(Thermo), C.c_int64]
    L.lmff_profile.argtypes = [vp, C.c_int]
    L.lmff_profile_read.argtypes = [vp, _d, _d, _i64]
    L.lmff_memory.argtypes = [vp, _i64, _d]
    L.lmff_system_bytes.argtypes = [vp, C.c_int64, C.c_int]
    L.lmff_system_bytes.restype = C.c_int64
    L.lmff_debug_ilp_list.argtypes = [vp, C.c_int, C.c_int64, _i32, C.c_int, _int]
    L.lmff_debug_normals.argtypes = [vp, C.c_int, _d, _i32]
    L.lmff_pair_stats.argtypes = [vp, _i64]
    L.lmff_system_set_step.argtypes = [vp, C.c_int, C.c_int64]
    L.lmff_peratom_on_last_step.argtypes = [vp, C.c_int]
    L.lmff_fire_defaults.argtypes = [C.POINTER(FireParams)]
    L.lmff_fire_defaults.restype = None
    L.lmff_minimize.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.POINTER(FireParams), C.POINTER(MinStats),
                                C.c_int]
    L.lmff_minimize_cg.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_double, C.c_int,
                                   C.POINTER(MinStats)]
    _lib = L
    return L


def _p(a, t):
    if a is None:
        return C.POINTER(t)()
    return a.ctypes.data_as(C.POINTER(t))


def _names(names):
    arr = (C.c_char_p * len(names))()
    for k, s in enumerate(names):
        arr[k] = s.encode()
    return arr


class Context:
    """One lmff context (HIP stream + resident systems) on a device."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.lmff_create(C.byref(h), int(device))
        if rc:
            raise LmffError(rc, "lmff_create failed (no GPU?)")
        self.h = h
        self.ntypes = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.lmff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc:
            raise LmffError(rc, self.L.lmff_last_error(self.h).decode())

    # ---- set-up (pair_style / pair_coeff / neighbor / mass)
    def settings(self, ntypes: int, ilp_cut: float = 16.0, tap_flag: int = 1):
        self.ntypes = ntypes
        self._chk(self.L.lmff_settings(self.h, ntypes, float(ilp_cut), int(tap_flag)))

    def coeff_tersoff(self, file: str, elements):
        self._chk(self.L.lmff_coeff_tersoff(self.h, file.encode(), _names(elements)))

    def coeff_ilp(self, file: str, elements):
        self._chk(self.L.lmff_coeff_ilp(self.h, file.encode(), _names(elements)))

    def neighbor(self, skin: float = 2.0, every: int = 10, delay: int = 0, check: int = 0):
        self._chk(self.L.lmff_neighbor(self.h, float(skin), int(every), int(delay), int(check)))

    def precision(self, mode):
        self._chk(self.L.lmff_precision(self.h, LMFF_MIXED if mode in (1, "mixed") else LMFF_FP64))

    def masses(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        self._chk(self.L.lmff_masses(self.h, _p(m, C.c_double)))

    # ---- Pair::compute on host arrays
    def pair_compute(self, s, eflag: int = 1, vflag: int = 2, new_list: int = 1, x=None):
        n = s.n
        x = np.ascontiguousarray(s.x if x is None else x, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        tag = np.ascontiguousarray(s.tag, dtype=np.int64)
        mol = np.ascontiguousarray(s.mol, dtype=np.int64)
        lo = np.ascontiguousarray(s.boxlo, dtype=np.float64)
        hi = np.ascontiguousarray(s.boxhi, dtype=np.float64)
        per = np.ascontiguousarray(s.periodic, dtype=np.int32)
        f = np.zeros((n, 3))
        eng = np.zeros(1)
        vir = np.zeros(6)
        pv = np.zeros(2)
        eatom = np.zeros(n) if eflag & 2 else None
        vatom = np.zeros((n, 6)) if vflag & 4 else None
        self._chk(self.L.lmff_pair_compute(self.h, n, _p(x, C.c_double), _p(t, C.c_int32), _p(tag, C.c_int64),
                                           _p(mol, C.c_int64), _p(lo, C.c_double), _p(hi, C.c_double),
                                           _p(per, C.c_int), int(new_list), int(eflag), int(vflag),
                                           _p(f, C.c_double), _p(eng, C.c_double), _p(vir, C.c_double),
                                           _p(eatom, C.c_double), _p(vatom, C.c_double), _p(pv, C.c_double)))
        return {"f": f, "energy": float(eng[0]), "e_vdw": float(pv[0]), "e_rep": float(pv[1]),
                "e_tersoff": float(eng[0] - pv[0] - pv[1]), "virial": vir, "eatom": eatom, "vatom": vatom}

    def pair_compute_noalloc(self, n: int, x_dev: int, type_dev: int, tag_dev: int, mol_dev: int, boxlo, boxhi,
                             periodic, eflag: int, vflag: int, f_dev: int, globals_dev: int):
        """The allocation-free entry point: every array argument is a device address (int),
        f_dev accumulates forces [n][3], globals_dev receives {eng_vdwl, virial[6], pvector[2]};
        0 for an output leaves it out."""
        lo = np.ascontiguousarray(boxlo, dtype=np.float64)
        hi = np.ascontiguousarray(boxhi, dtype=np.float64)
        per = np.ascontiguousarray(periodic, dtype=np.int32)
        vp = C.c_void_p
        self._chk(self.L.lmff_pair_compute_noalloc(self.h, int(n), vp(x_dev), vp(type_dev), vp(tag_dev), vp(mol_dev),
                                                   _p(lo, C.c_double), _p(hi, C.c_double), _p(per, C.c_int),
                                                   int(eflag), int(vflag), vp(f_dev or None), vp(globals_dev or None)))

    # ---- resident systems
    def add_system(self, s, v=None) -> int:
        x = np.ascontiguousarray(s.x, dtype=np.float64)
        vv = s.v if v is None else v
        vv = None if vv is None else np.ascontiguousarray(vv, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        tag = np.ascontiguousarray(s.tag, dtype=np.int64)
        mol = np.ascontiguousarray(s.mol, dtype=np.int64)
        lo = np.ascontiguousarray(s.boxlo, dtype=np.float64)
        hi = np.ascontiguousarray(s.boxhi, dtype=np.float64)
        per = np.ascontiguousarray(s.periodic, dtype=np.int32)
        sid = C.c_int(-1)
        self._chk(self.L.lmff_system_add(self.h, s.n, _p(lo, C.c_double), _p(hi, C.c_double), _p(per, C.c_int),
                                         _p(t, C.c_int32), _p(tag, C.c_int64), _p(mol, C.c_int64),
                                         _p(x, C.c_double), _p(vv, C.c_double), C.byref(sid)))
        return sid.value

    def remove_system(self, sid: int):
        self._chk(self.L.lmff_system_remove(self.h, int(sid)))

    def integrator(self, sid: int, style=LMFF_NVE, dt: float = 0.001, t_target: float = 300.0, damp: float = 1.0,
                   seed: int = 1):
        self._chk(self.L.lmff_system_integrator(self.h, int(sid), int(style), float(dt), float(t_target),
                                                float(damp), int(seed)))

    def natoms(self, sid: int) -> int:
        n = C.c_int64()
        self._chk(self.L.lmff_system_natoms(self.h, int(sid), C.byref(n)))
        return n.value

    def get(self, sid: int):
        n = self.natoms(sid)
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        self._chk(self.L.lmff_system_get(self.h, int(sid), _p(x, C.c_double), _p(v, C.c_double), _p(f, C.c_double)))
        return x, v, f

    def set(self, sid: int, x=None, v=None):
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        self._chk(self.L.lmff_system_set(self.h, int(sid), _p(x, C.c_double), _p(v, C.c_double)))

    def get_image(self, sid: int):
        """Image flags [n][3] (box crossings counted when positions are wrapped)."""
        im = np.zeros((self.natoms(sid), 3), dtype=np.int32)
        self._chk(self.L.lmff_system_get_image(self.h, int(sid), _p(im, C.c_int32)))
        return im

    def set_image(self, sid: int, image):
        im = np.ascontiguousarray(image, dtype=np.int32)
        self._chk(self.L.lmff_system_set_image(self.h, int(sid), _p(im, C.c_int32)))

    def peratom(self, sid: int, energy=True, virial=True):
        n = self.natoms(sid)
        e = np.zeros(n) if energy else None
        w = np.zeros((n, 6)) if virial else None
        self._chk(self.L.lmff_system_peratom(self.h, int(sid), _p(e, C.c_double), _p(w, C.c_double)))
        return e, w

    def set_step(self, sid: int, step: int):
        """reset_timestep for one system (the Langevin noise is keyed by the step)."""
        self._chk(self.L.lmff_system_set_step(self.h, int(sid), int(step)))

    def peratom_on_last_step(self, enable: bool = True):
        """The last step of every run also tallies per-atom energy and virial."""
        self._chk(self.L.lmff_peratom_on_last_step(self.h, 1 if enable else 0))

    def minimize(self, etol: float, ftol: float, maxiter: int, nsys: int = 1, **fire):
        """min_style fire + minimize etol ftol maxiter (every resident system); keyword
        arguments override LAMMPS' FIRE defaults (dmax, delaystep, ..., norm 0/1/2)."""
        p = FireParams()
        self.L.lmff_fire_defaults(C.byref(p))
        for k, v in fire.items():
            if not hasattr(p, k):
                raise ValueError(f"unknown FIRE parameter {k}")
            setattr(p, k, v)
        out = (MinStats * max(nsys, 1))()
        self._chk(self.L.lmff_minimize(self.h, float(etol), float(ftol), int(maxiter), C.byref(p), out, int(nsys)))
        return [out[k].as_dict() for k in range(nsys)]

    def minimize_cg(self, etol: float, ftol: float, maxiter: int, maxeval: int, dmax: float = 0.1, norm: int = 0):
        """min_style cg (quadratic line search) + minimize etol ftol maxiter maxeval; one system."""
        out = MinStats()
        self._chk(self.L.lmff_minimize_cg(self.h, float(etol), float(ftol), int(maxiter), int(maxeval), float(dmax),
                                          int(norm), C.byref(out)))
        return out.as_dict()

    def evaluate(self, eflag: int = 1, vflag: int = 2):
        self._chk(self.L.lmff_evaluate(self.h, int(eflag), int(vflag)))

    def thermo(self, sid: int) -> dict:
        t = Thermo()
        self._chk(self.L.lmff_system_thermo(self.h, int(sid), C.byref(t)))
        return t.as_dict()

    def run(self, nsteps: int, thermo_every: int = 0, use_graph: bool = True, nsys: int = 1):
        nrec = 0
        if nsteps > 0:
            nrec = (nsteps // thermo_every if thermo_every > 0 else 0)
            if thermo_every <= 0 or nsteps % thermo_every:
                nrec += 1
        buf = (Thermo * max(nrec * nsys, 1))()
        self._chk(self.L.lmff_run(self.h, int(nsteps), int(thermo_every), int(bool(use_graph)), buf, nrec * nsys))
        return [[buf[k * nrec + r].as_dict() for r in range(nrec)] for k in range(nsys)]

    def neighbor_stats(self, reset: bool = False):
        """(list builds, dangerous builds) since the last reset."""
        b = np.zeros(1, dtype=np.int64)
        d = np.zeros(1, dtype=np.int64)
        self._chk(self.L.lmff_neighbor_stats(self.h, int(reset), _p(b, C.c_int64), _p(d, C.c_int64)))
        return int(b[0]), int(d[0])

    def profile(self, enable: bool = True):
        self._chk(self.L.lmff_profile(self.h, int(enable)))

    def profile_read(self):
        tot = C.c_double()
        ms = np.zeros(7)
        cnt = np.zeros(7, dtype=np.int64)
        self._chk(self.L.lmff_profile_read(self.h, C.byref(tot), _p(ms, C.c_double), _p(cnt, C.c_int64)))
        return tot.value, ms, cnt

    def memory(self):
        b = C.c_int64()
        pa = C.c_double()
        self._chk(self.L.lmff_memory(self.h, C.byref(b), C.byref(pa)))
        return b.value, pa.value

    def system_bytes(self, natoms: int, nlayers: int = 2) -> int:
        return int(self.L.lmff_system_bytes(self.h, int(natoms), int(nlayers)))

    def ilp_list(self, sid: int, atom: int, cap: int = 8192):
        rows = np.zeros((cap, 4), dtype=np.int32)
        n = C.c_int()
        self._chk(self.L.lmff_debug_ilp_list(self.h, int(sid), int(atom), _p(rows, C.c_int32), cap, C.byref(n)))
        return rows[: min(n.value, cap)]

    def pair_stats(self):
        out = np.zeros(4, dtype=np.int64)
        self._chk(self.L.lmff_pair_stats(self.h, _p(out, C.c_int64)))
        return {"ilp_entries": int(out[0]), "ilp_pairs": int(out[1]), "tersoff_nbrs": int(out[2]),
                "short_entries": int(out[3])}

    def normals(self, sid: int):
        n = self.natoms(sid)
        nv = np.zeros((n, 3))
        cnt = np.zeros(n, dtype=np.int32)
        self._chk(self.L.lmff_debug_normals(self.h, int(sid), _p(nv, C.c_double), _p(cnt, C.c_int32)))
        return nv, cnt


def potential(name: str) -> str:
    return os.path.join(POTENTIALS, name)


def setup_graphene(ctx: Context, ntypes: int = 1, elements_tersoff=("C(O)",), elements_ilp=("C",),
                   tersoff_file: str = "C.opt.tersoff", masses=(12.0107,), ilp_cut: float = 16.0, tap: int = 1,
                   skin: float = 2.0, every: int = 10, delay: int = 0, check: int = 0, precision=LMFF_FP64):
    """pair_style hybrid/overlay tersoff ilp/graphene/hbn <cut> <tap> + pair_coeff + neighbor + mass."""
    ctx.settings(ntypes, ilp_cut, tap)
    ctx.coeff_tersoff(potential(tersoff_file), list(elements_tersoff))
    ctx.coeff_ilp(potential("BNCH.ILP"), list(elements_ilp))
    ctx.neighbor(skin, every, delay, check)
    ctx.masses(list(masses))
    ctx.precision(precision)
    return ctx


def setup_for(ctx: Context, s, precision=LMFF_FP64, **kw):
    """Coefficients matching a generated System: C-only -> C.opt.tersoff C(O); C/B/N -> BNC.tersoff."""
    if s.elements == ["C"]:
        return setup_graphene(ctx, 1, ("C(O)",), ("C",), "C.opt.tersoff", (12.0107,), precision=precision, **kw)
    return setup_graphene(ctx, len(s.elements), tuple(s.elements), tuple(s.elements), "BNC.tersoff",
                          tuple(s.mass_per_type), precision=precision, **kw)
