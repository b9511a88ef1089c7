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
        L.orc_compute_lmff.argtypes = [C.c_void_p, C.c_int, _d, _d]
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

    def compute_lmff(self, eflag: int = 1):
        """The reference's fused LMFF CPU path restated (LMFF/pair_lmff.cpp:1121-1613): forces and
        energies only (the CPU baseline's path); the same results as compute() up to rounding."""
        f = np.zeros((self.n, 3))
        eng = np.zeros(4)
        self._check(lib().orc_compute_lmff(self.h, int(eflag), _p(f, C.c_double), _p(eng, C.c_double)))
        return {"f": f, "energy": eng[0], "e_vdw": eng[1], "e_rep": eng[2], "e_tersoff": eng[3]}

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


FIRE_DEFAULTS = dict(dmax=0.1, delaystep=20, dtgrow=1.1, dtshrink=0.5, alpha0=0.25, alphashrink=0.99, tmax=10.0,
                     tmin=0.02, halfstepback=1, delaystep_start=1, max_vdotf_negatif=2000, norm=0)


def fire_minimize(o: Oracle, x0, mass_per_atom, dt: float, etol: float, ftol: float, maxiter: int,
                  skin: float = 2.0, **params):
    """FIRE on the oracle's forces: MinFire::iterate (LMFF/min_fire.cpp:104-460) with the
    default semi-implicit Euler integrator and half step back, velocities zeroed at the start
    (MinFire::setup_style, :83-88), lists rebuilt when an atom has moved more than skin/2
    (Min::init sets every 1, delay 0, check yes: LMFF/min.cpp:190-196; Neighbor::check_distance).
    Energy and force tolerances as in :420-455 (EPS_ENERGY 1e-8).  Returns the stop code
    (0 max iterations, 2 energy tolerance, 3 force tolerance, 11 v.f negative), counts,
    energies and the final positions."""
    p = dict(FIRE_DEFAULTS)
    p.update(params)
    ftm2v = 1.0 / 1.0364269e-4
    x = np.array(x0, dtype=np.float64)
    m = np.asarray(mass_per_atom, dtype=np.float64)[:, None]
    v = np.zeros_like(x)
    o.put_x(x)
    o.pbc()
    o.build_list()
    x = o.get_x()
    xhold = x.copy()
    r = o.compute(1, 0)
    f, ecur = r["f"], r["energy"]
    e_init = ecur
    dtcur, dtmax, dtmin, alpha = dt, p["tmax"] * dt, p["tmin"] * dt, p["alpha0"]
    ts = last_negative = start = 0
    negatif = 0
    dtv = 0.0
    stop, niter, neval = 0, 0, 0
    eprev = ecur

    def fnorm(ff):
        if p["norm"] == 1:
            return (ff * ff).sum(1).max()
        if p["norm"] == 2:
            return np.abs(ff).max()
        return (ff * ff).sum()

    for _ in range(maxiter):
        ts += 1
        niter += 1
        vdotf = float((v * f).sum())
        mix = False
        if vdotf > 0.0:
            negatif = 0
            vdotv, fdotf = float((v * v).sum()), float((f * f).sum())
            scale1 = 1.0 - alpha
            scale2 = 0.0 if fdotf <= 1e-20 else alpha * np.sqrt(vdotv / fdotf)
            mix = True
            if ts - last_negative > p["delaystep"]:
                dtcur = min(dtcur * p["dtgrow"], dtmax)
                alpha *= p["alphashrink"]
        else:
            last_negative = ts
            if not (ts - start < p["delaystep"] and p["delaystep_start"]):
                alpha = p["alpha0"]
                if dtcur * p["dtshrink"] >= dtmin:
                    dtcur *= p["dtshrink"]
            negatif += 1
            if p["max_vdotf_negatif"] > 0 and negatif > p["max_vdotf_negatif"]:
                stop = 11
                break
            if p["halfstepback"]:
                x -= 0.5 * dtv * v
            v[:] = 0.0
        vmax = float(np.abs(v).max()) if v.size else 0.0
        dtv = dtcur
        if dtv * vmax > p["dmax"]:
            dtv = p["dmax"] / vmax
        v += (dtv * ftm2v / m) * f
        if mix:
            v = scale1 * v + scale2 * f
        x += dtv * v
        if ((x - xhold) ** 2).sum(1).max() > (0.5 * skin) ** 2:
            o.put_x(x)
            o.pbc()
            o.build_list()
            x = o.get_x()
            xhold = x.copy()
        else:
            o.put_x(x)
        r = o.compute(1, 0)
        f = r["f"]
        neval += 1
        eprev, ecur = ecur, r["energy"]
        if etol > 0.0 and ts - last_negative > p["delaystep"] and \
                abs(ecur - eprev) < etol * 0.5 * (abs(ecur) + abs(eprev) + 1.0e-8):
            stop = 2
            break
        if ftol > 0.0 and fnorm(f) < ftol * ftol:
            stop = 3
            break
    return {"stop": stop, "iterations": niter, "evaluations": neval, "e_initial": e_init, "e_final": ecur,
            "e_next_to_last": eprev, "fnorm2_final": float(np.sqrt((f * f).sum())), "x": x, "f": f}


def cg_minimize(o: Oracle, x0, etol: float, ftol: float, maxiter: int, maxeval: int, skin: float = 2.0,
                dmax: float = 0.1, norm: int = 0):
    """Conjugate gradients on the oracle's forces: MinCG::iterate (LMFF/min_cg.cpp:40-190,
    Polak-Ribiere beta >= 0, restart every 3N iterations and when g.h <= 0) with
    MinLineSearch::linemin_quadratic and alpha_step (LMFF/min_linesearch.cpp:324-503, 837-868:
    alphamax = min(1, dmax/max|h|), backtracking by 1/2 with slope 0.4, quadratic projection
    when the secant model fits within 0.1).  Lists are rebuilt when an atom has moved more than
    skin/2 (Min::init); the start of the line search follows a wrap by its periodic image
    (FixMinimize::reset_coords).  Stop codes as Min::stopstrings."""
    EMACH, EPS_QUAD = 1.0e-8, 1.0e-28
    x = np.array(x0, dtype=np.float64)
    o.put_x(x)
    o.pbc()
    o.build_list()
    x = o.get_x()
    xhold = x.copy()
    L = np.asarray(o._keep[4]) - np.asarray(o._keep[3])
    per = np.asarray(o._keep[5]).astype(bool)
    state = {"x": x, "xhold": xhold, "neval": 0}

    def evaluate(xn):
        if ((xn - state["xhold"]) ** 2).sum(1).max() > (0.5 * skin) ** 2:
            o.put_x(xn)
            o.pbc()
            o.build_list()
            xw = o.get_x()
            state["xhold"] = xw.copy()
            state["wrap"] = xw - xn
            xn = xw
        else:
            o.put_x(xn)
        state["x"] = xn
        r = o.compute(1, 0)
        return r["energy"], r["f"]

    def fnorm(ff):
        if norm == 1:
            return (ff * ff).sum(1).max()
        if norm == 2:
            return np.abs(ff).max()
        return (ff * ff).sum()

    ecur, f = evaluate(x)
    e_init = ecur
    g = f.copy()
    h = f.copy()
    gg = float((f * f).sum())
    nlimit = min(2147483647, 3 * len(x))

    def alpha_step(xs, alpha):
        state["neval"] += 1
        xn = xs + alpha * h if alpha > 0.0 else xs.copy()
        return evaluate(xn)

    def image_follow(xs):
        d = state["x"] - xs
        for k in range(3):
            if per[k]:
                xs[:, k] += L[k] * np.rint(d[:, k] / L[k])
        return xs

    def linemin(eoriginal):
        nonlocal f
        fdothall = float((f * h).sum())
        if fdothall <= 0.0:
            return 4, eoriginal
        hmax = float(np.abs(h).max())
        if hmax == 0.0:
            return 6, eoriginal
        alphamax = min(1.0, dmax / hmax)
        xs = state["x"].copy()
        alpha, fhprev, engprev, alphaprev = alphamax, fdothall, eoriginal, 0.0
        while True:
            e, f = alpha_step(xs, alpha)
            xs = image_follow(xs)
            fh = float((f * h).sum())
            delfh = fh - fhprev
            if abs(fh) < EPS_QUAD or abs(delfh) < EPS_QUAD:
                e, f = alpha_step(xs, 0.0)
                return 7, e
            relerr = abs(1.0 - (0.5 * (alpha - alphaprev) * (fh + fhprev) + e) / engprev)
            a0 = alpha - (alpha - alphaprev) * fh / delfh
            if relerr <= 0.1 and 0.0 < a0 < alphamax:
                e, f = alpha_step(xs, a0)
                xs = image_follow(xs)
                if e - eoriginal < EMACH:
                    return 0, e
            de_ideal = -0.4 * alpha * fdothall
            if e - eoriginal <= de_ideal:
                return 0, e
            fhprev, engprev, alphaprev = fh, e, alpha
            alpha *= 0.5
            if alpha <= 0.0 or de_ideal >= -EMACH:
                e, f = alpha_step(xs, 0.0)
                return 5, e

    stop, niter, eprev = 0, 0, ecur
    for _ in range(maxiter):
        niter += 1
        eprev = ecur
        rc, ecur = linemin(ecur)
        if rc:
            stop = rc
            break
        if state["neval"] >= maxeval:
            stop = 1
            break
        if abs(ecur - eprev) < etol * 0.5 * (abs(ecur) + abs(eprev) + 1.0e-8):
            stop = 2
            break
        dot0, dot1 = float((f * f).sum()), float((f * g).sum())
        if ftol > 0.0 and fnorm(f) < ftol * ftol:
            stop = 3
            break
        beta = max(0.0, (dot0 - dot1) / gg)
        if (niter + 1) % nlimit == 0:
            beta = 0.0
        gg = dot0
        g = f.copy()
        h = g + beta * h
        if float((g * h).sum()) <= 0.0:
            h = g.copy()
    return {"stop": stop, "iterations": niter, "evaluations": state["neval"], "e_initial": e_init, "e_final": ecur,
            "e_next_to_last": eprev, "fnorm2_final": float(np.sqrt((f * f).sum())), "x": state["x"], "f": f}
