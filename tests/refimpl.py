"""Independent all-pairs numpy evaluation of the LMFF energy (test infrastructure).

Written from the published functional forms, not from the oracle's code:
  Tersoff (J. Tersoff, PRB 37, 6991 (1988); LAMMPS pair_style tersoff docs):
    E = 1/2 sum_i sum_{j!=i} fC(r_ij) [fR(r_ij) + b_ij fA(r_ij)],
    fR = A exp(-lam1 r), fA = -B exp(-lam2 r),
    b_ij = (1 + (beta zeta_ij)^n)^(-1/2n),
    zeta_ij = sum_{k!=i,j} fC(r_ik) g(cos theta_ijk) exp((lam3 (r_ij - r_ik))^m),
    g = gamma (1 + c^2/d^2 - c^2/(d^2 + (cos theta - h)^2)).
  ILP (Leven, Maaravi, Azuri, Kronik, Hod, JCTC 12, 2896 (2016); Ouyang et al.,
  Nano Lett. 18, 6009 (2018)), for atoms in different layers within rc:
    E_rep = sum over ordered pairs Tap(r) exp(-lambda (r - z0)) (eps/2 + C exp(-rho_ij^2/delta^2)),
            rho_ij^2 = r^2 - (n_i . r_ij)^2,
    E_vdw = -sum over unordered pairs Tap(r) C6 / (r^6 (1 + exp(-d (r/(sR reff) - 1)))),
    n_i   = unit normal of the plane through i's (up to three) nearest same-layer
            neighbours within rcut: (x2-x1) x (x3-x1) for three, (x1-xi) x (x2-xi) for
            two, z for fewer (the sign is irrelevant: n enters squared).
  Tap(x) = 20x^7 - 70x^6 + 84x^5 - 35x^4 + 1, x = r/rc.
Parameters are read by a minimal parser of the same LAMMPS potential files.

Two conventions are taken from the reference rather than from the published forms, because
they decide what a tally or an asymmetric parameter table gives:
  * which atom of an unordered pair computes it (and so whose parameter row is used) when a
    full list is halved: the tag-parity rule of pair_tersoff.cpp:150-158 and
    pair_ilp_graphene_hbn.cpp:488-496 (owner_is_first below);
  * per-atom energies: Pair::ev_tally / ev_tally_xyz give half of every tallied energy to each
    of the two atoms (LMFF/pair.cpp:939-941, 1087-1089).  Tersoff tallies the repulsive pair
    energy (pair_tersoff.cpp:164-174) and 1/2 b_ij f_A of each ordered pair (214-224), ILP the
    van der Waals half pair (ilp.cpp:536) and the repulsion of each ordered pair (671).
"""
from __future__ import annotations

import numpy as np


def read_table(path: str, nkey: int, ncol: int):
    rows = {}
    with open(path) as fh:
        toks = []
        for line in fh:
            line = line.split("#", 1)[0]
            toks += line.split()
            while len(toks) >= ncol:
                row, toks = toks[:ncol], toks[ncol:]
                rows[tuple(row[:nkey])] = [float(v) for v in row[nkey:]]
    return rows


def tersoff_params(path: str, elems):
    raw = read_table(path, 3, 17)
    out = {}
    for key, v in raw.items():
        if all(k in elems for k in key):
            m, gamma, lam3, c, d, h, n, beta, lam2, B, R, D, lam1, A = v
            out[key] = dict(m=m, gamma=gamma, lam3=lam3, c=c, d=d, h=h, n=n, beta=beta, lam2=lam2, B=B, R=R, D=D,
                            lam1=lam1, A=A)
    return out


def ilp_params(path: str, elems):
    raw = read_table(path, 2, 13)
    out = {}
    for key, v in raw.items():
        if all(k in elems for k in key):
            z0, alpha, delta, eps, C, d, sR, reff, C6, S, rcut = v
            meV = 1e-3 * S
            out[key] = dict(z0=z0, lam=alpha / z0, delta=delta, eps=eps * meV, C=C * meV, d=d, seff=sR * reff,
                            C6=C6 * meV, rcut=rcut)
    return out


def tap(r, rc):
    x = r / rc
    t = 20 * x**7 - 70 * x**6 + 84 * x**5 - 35 * x**4 + 1
    return np.where(x < 1.0, t, 0.0)


def min_image(s):
    """all-pairs displacement x_i - x_j under the minimum-image convention, [n,n,3]"""
    x = s.x
    d = x[:, None, :] - x[None, :, :]
    L = s.boxhi - s.boxlo
    for k in range(3):
        if s.periodic[k]:
            d[:, :, k] -= L[k] * np.round(d[:, :, k] / L[k])
    return d


def fc(r, R, D):
    return np.where(r < R - D, 1.0, np.where(r > R + D, 0.0, 0.5 * (1.0 - np.sin(0.5 * np.pi * (r - R) / D))))


def owner_is_first(ti, tj):
    """True where the pair (i, j) of a full list is computed by i: tags of opposite parity are
    kept by the larger tag, tags of equal parity by the smaller one (pair_tersoff.cpp:150-158,
    pair_ilp_graphene_hbn.cpp:488-496).  Equal tags (an atom and its own image) never occur
    here: every periodic side exceeds twice the cutoff."""
    ti = np.asarray(ti, dtype=np.int64)
    tj = np.asarray(tj, dtype=np.int64)
    assert not np.any(ti == tj)
    odd = (ti + tj) % 2 == 1
    return np.where(ti > tj, odd, ~odd)


def _share(eatom, i, j, e, split):
    """Per-atom tally of energy e of the pair (i, j): split "half" as Pair::ev_tally, "first"
    (everything to i: a counter-example for the tests)."""
    if split == "half":
        eatom[i] += 0.5 * e
        eatom[j] += 0.5 * e
    elif split == "first":
        eatom[i] += e
    else:
        raise ValueError(split)


def tersoff_energy(s, tfile, elem_of_type, peratom=False, split="half", parity=True):
    """Total Tersoff energy; with peratom, (E, eatom).  parity=False averages the repulsive
    energy of a pair over both parameter rows (i,j,j) and (j,i,i) instead of taking the owner's
    (a counter-example for asymmetric tables)."""
    elems = [e for e in elem_of_type]
    P = tersoff_params(tfile, set(elems))
    el = [elems[t - 1] for t in s.type]
    d = min_image(s)
    r = np.sqrt((d * d).sum(-1))
    n = s.n
    cutmax = max(p["R"] + p["D"] for p in P.values())
    E = 0.0
    eatom = np.zeros(n)
    nb = [np.nonzero((r[i] < cutmax) & (np.arange(n) != i))[0] for i in range(n)]
    for i in range(n):
        for j in nb[i]:
            # repulsive: once per unordered pair, with the parameters of the atom that computes it
            if j > i:
                rij = r[i, j]
                if parity:
                    o, q = (i, j) if owner_is_first(s.tag[i], s.tag[j]) else (j, i)
                    rows = [P[(el[o], el[q], el[q])]]
                else:
                    rows = [P[(el[i], el[j], el[j])], P[(el[j], el[i], el[i])]]
                er = 0.0
                for pr in rows:
                    if rij < pr["R"] + pr["D"]:
                        er += fc(rij, pr["R"], pr["D"]) * pr["A"] * np.exp(-pr["lam1"] * rij) / len(rows)
                if er != 0.0:
                    E += er
                    _share(eatom, i, j, er, split)
            pij = P[(el[i], el[j], el[j])]
            rij = r[i, j]
            if rij >= pij["R"] + pij["D"]:
                continue
            zeta = 0.0
            for k in nb[i]:
                if k == j:
                    continue
                p = P[(el[i], el[j], el[k])]
                rik = r[i, k]
                if rik >= p["R"] + p["D"]:
                    continue
                cos = np.dot(-d[i, j], -d[i, k]) / (rij * rik)
                g = p["gamma"] * (1 + p["c"] ** 2 / p["d"] ** 2 - p["c"] ** 2 / (p["d"] ** 2 + (cos - p["h"]) ** 2))
                zeta += fc(rik, p["R"], p["D"]) * g * np.exp((p["lam3"] * (rij - rik)) ** int(p["m"]))
            b = (1.0 + (pij["beta"] * zeta) ** pij["n"]) ** (-0.5 / pij["n"])
            f = fc(rij, pij["R"], pij["D"])
            ea = -0.5 * f * b * pij["B"] * np.exp(-pij["lam2"] * rij)
            E += ea
            _share(eatom, i, j, ea, split)
    return (float(E), eatom) if peratom else float(E)


def normals(s, ifile, elem_of_type):
    P = ilp_params(ifile, set(elem_of_type))
    el = [elem_of_type[t - 1] for t in s.type]
    d = min_image(s)
    r = np.sqrt((d * d).sum(-1))
    n = s.n
    nv = np.zeros((n, 3))
    cnt = np.zeros(n, dtype=int)
    for i in range(n):
        rc = np.array([P[(el[i], el[j])]["rcut"] for j in range(n)])
        sel = np.nonzero((r[i] > 0) & (r[i] < rc) & (s.mol == s.mol[i]))[0]
        cnt[i] = len(sel)
        pos = -d[i, sel]  # neighbour positions relative to i
        if len(sel) <= 1:
            v = np.array([0.0, 0.0, 1.0])
        elif len(sel) == 2:
            v = np.cross(pos[0], pos[1])
        elif len(sel) == 3:
            v = np.cross(pos[1] - pos[0], pos[2] - pos[0])
        else:
            raise ValueError("more than 3 normal neighbours")
        nv[i] = v / np.linalg.norm(v)
    return nv, cnt


def ilp_energy(s, ifile, elem_of_type, rc=16.0, tap_flag=1, peratom=False, split="half", parity=True):
    """returns (E_vdw, E_rep), with peratom (E_vdw, E_rep, eatom).  The van der Waals energy of
    an unordered pair uses the parameter row of the atom that computes it (parity=False: the
    average of both rows, a counter-example for asymmetric tables); the repulsion of the
    ordered pair (i, j) uses row (i, j) and i's normal."""
    P = ilp_params(ifile, set(elem_of_type))
    el = [elem_of_type[t - 1] for t in s.type]
    nv, _ = normals(s, ifile, elem_of_type)
    d = min_image(s)
    r = np.sqrt((d * d).sum(-1))
    other = s.mol[:, None] != s.mol[None, :]
    mask = other & (r < rc)
    ii, jj = np.nonzero(mask)
    keys = list(zip([el[i] for i in ii], [el[j] for j in jj]))
    get = lambda name: np.array([P[k][name] for k in keys])
    rr = r[ii, jj]
    T = tap(rr, rc) if tap_flag else np.ones_like(rr)
    dd = d[ii, jj]
    pn = (nv[ii] * dd).sum(-1)
    rho2 = rr * rr - pn * pn
    erep = T * np.exp(-get("lam") * (rr - get("z0"))) * (0.5 * get("eps") + get("C") * np.exp(-rho2 / get("delta") ** 2))
    evdw = -T * get("C6") / (rr**6 * (1.0 + np.exp(-get("d") * (rr / get("seff") - 1.0))))
    # every ordered pair evaluated with its own row (i, j); an unordered pair's vdW energy is
    # the owner's row, i.e. the ordered pair (i, j) whose i owns it
    wv = owner_is_first(s.tag[ii], s.tag[jj]).astype(np.float64) if parity else np.full(len(ii), 0.5)
    ev = float((wv * evdw).sum())
    er = float(erep.sum())
    if not peratom:
        return ev, er
    eatom = np.zeros(s.n)
    if split == "half":
        # each ordered pair carries w*evdw (its share of the unordered vdW pair) + erep; both
        # halves of each go to i and j
        e = wv * evdw + erep
        np.add.at(eatom, ii, 0.5 * e)
        np.add.at(eatom, jj, 0.5 * e)
    elif split == "first":
        np.add.at(eatom, ii, wv * evdw + erep)
    else:
        raise ValueError(split)
    return ev, er, eatom
