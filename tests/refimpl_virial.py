"""Independent per-atom virial of the LMFF force field for one element (test infrastructure).

The forces come from automatic differentiation of the published energy forms (the same forms
as tests/refimpl.py), not from the analytic derivatives the oracle and the HIP kernels use:
every term's positions are separate leaf copies, so the gradient on a copy is that term's
force on that atom alone.  Only the assignment of each term's virial to atoms is taken from
the reference, i.e. its tally calls:

  Tersoff (Baseline/MANYBODY/pair_tersoff.cpp): repulsive pair -> ev_tally(i, j) halves
    (pair_tersoff.cpp:150-170); force_zeta -> ev_tally(i, j, ..., -fpair, -delr1) halves
    (pair_tersoff.cpp:218-232); each attractive() triplet -> v_tally3(i, j, k, fj, fk, delr1,
    delr2) thirds (pair_tersoff.cpp:236-259).
  ILP (Baseline/USER-MISC/pair_ilp_graphene_hbn.cpp): van der Waals half pairs -> ev_tally
    (i, j) halves (536); repulsion of the ordered pair (i, j) -> ev_tally_xyz(i, j, fkc, del)
    halves (671), and for each normal neighbour k of i -> ev_tally_xyz(k, j, fk, x_k - x_j)
    halves (667).  The d n_i / d x_i force of two-neighbour normals is applied but not
    tallied (ilp.cpp:640-642).

With LAMMPS' layout per atom: xx, yy, zz, xy, xz, yz.  Open boundaries only.
"""
from __future__ import annotations

import numpy as np
import torch

from refimpl import ilp_params, tersoff_params

f64 = torch.float64


def _v(d, f):
    return torch.stack([d[:, 0] * f[:, 0], d[:, 1] * f[:, 1], d[:, 2] * f[:, 2],
                        d[:, 0] * f[:, 1], d[:, 0] * f[:, 2], d[:, 1] * f[:, 2]], 1)


def _leaf(t):
    return t.detach().clone().requires_grad_(True)


def peratom_virial(s, tfile: str, telem: str, ifile: str, ielem: str, rc: float = 16.0, tap_flag: int = 1,
                   k_partner: str = "j"):
    """k_partner: the atom that shares the virial of a normal-derivative force with k: "j" as
    in ilp.cpp:667, or "i" (the LMFF style's choice, kept here only as a counter-example)."""
    assert not any(s.periodic), "open boundaries only"
    X = torch.tensor(np.asarray(s.x), dtype=f64)
    mol = torch.tensor(np.asarray(s.mol))
    n = X.shape[0]
    W = torch.zeros(n, 6, dtype=f64)

    def tally(idx, v, w):
        W.index_add_(0, idx, w * v.detach())

    # ------------------------------------------------------------------ Tersoff
    tp = tersoff_params(tfile, {telem})[(telem, telem, telem)]
    R, D = tp["R"], tp["D"]

    def fcut(r):
        mid = 0.5 * (1.0 - torch.sin(0.5 * np.pi * (r - R) / D))
        return torch.where(r < R - D, torch.ones_like(r), torch.where(r > R + D, torch.zeros_like(r), mid))

    def fa(r):
        return -tp["B"] * torch.exp(-tp["lam2"] * r) * fcut(r)

    def bij(z):
        return (1.0 + (tp["beta"] * z) ** tp["n"]) ** (-0.5 / tp["n"])

    def gijk(cos):
        c2, d2 = tp["c"] ** 2, tp["d"] ** 2
        return tp["gamma"] * (1.0 + c2 / d2 - c2 / (d2 + (tp["h"] - cos) ** 2))

    dist = (X[:, None, :] - X[None, :, :]).norm(dim=-1)  # exact, not through a matrix product
    near = (dist < R + D) & ~torch.eye(n, dtype=torch.bool)
    I, J = torch.nonzero(near, as_tuple=True)

    # repulsive, each unordered pair once: E = fC A exp(-lam1 r)
    m = I < J
    i, j = I[m], J[m]
    xi, xj = _leaf(X[i]), _leaf(X[j])
    r = (xi - xj).norm(dim=1)
    (fcut(r) * tp["A"] * torch.exp(-tp["lam1"] * r)).sum().backward()
    v = _v(X[i] - X[j], -xi.grad)
    tally(i, v, 0.5)
    tally(j, v, 0.5)

    # bond order of each ordered pair (i, j): E = 1/2 fA(r_ij) b(zeta_ij)
    P = I.shape[0]
    trip_p, trip_k = [], []
    for p in range(P):
        ks = torch.nonzero(near[I[p]], as_tuple=True)[0]
        ks = ks[ks != J[p]]
        trip_p.append(torch.full((ks.shape[0],), p, dtype=torch.long))
        trip_k.append(ks)
    tp_idx = torch.cat(trip_p) if trip_p else torch.zeros(0, dtype=torch.long)
    tk_idx = torch.cat(trip_k) if trip_k else torch.zeros(0, dtype=torch.long)
    ti_idx, tj_idx = I[tp_idx], J[tp_idx]

    def zeta_terms(xi_, xj_, xk_):
        dij, dik = xj_ - xi_, xk_ - xi_
        rij, rik = dij.norm(dim=1), dik.norm(dim=1)
        cos = (dij * dik).sum(1) / (rij * rik)
        arg = tp["lam3"] * (rij - rik)
        ex = torch.exp(arg ** int(tp["m"]) if int(tp["m"]) == 3 else arg)
        return fcut(rik) * gijk(cos) * ex

    with torch.no_grad():
        zeta = torch.zeros(P, dtype=f64).index_add_(0, tp_idx, zeta_terms(X[ti_idx], X[tj_idx], X[tk_idx]))
    # force_zeta: the r_ij dependence at fixed zeta
    xi, xj = _leaf(X[I]), _leaf(X[J])
    r = (xj - xi).norm(dim=1)
    (0.5 * fa(r) * bij(zeta)).sum().backward()
    v = _v(X[I] - X[J], -xi.grad)
    tally(I, v, 0.5)
    tally(J, v, 0.5)
    # the zeta dependence, per triplet
    xi, xj, xk = _leaf(X[ti_idx]), _leaf(X[tj_idx]), _leaf(X[tk_idx])
    z = torch.zeros(P, dtype=f64).index_add(0, tp_idx, zeta_terms(xi, xj, xk))
    r_ij = (X[J] - X[I]).norm(dim=1)
    (0.5 * fa(r_ij) * bij(z)).sum().backward()
    v = _v(X[tj_idx] - X[ti_idx], -xj.grad) + _v(X[tk_idx] - X[ti_idx], -xk.grad)
    for idx in (ti_idx, tj_idx, tk_idx):
        tally(idx, v, 1.0 / 3.0)

    # ------------------------------------------------------------------ ILP
    ip = ilp_params(ifile, {ielem})[(ielem, ielem)]

    def tap(r):
        if not tap_flag:
            return torch.ones_like(r)
        x = r / rc
        return 20 * x**7 - 70 * x**6 + 84 * x**5 - 35 * x**4 + 1

    same = mol[:, None] == mol[None, :]
    nbr = [torch.nonzero((dist[a] > 0) & (dist[a] < ip["rcut"]) & same[a], as_tuple=True)[0] for a in range(n)]
    assert all(len(q) <= 3 for q in nbr)
    pair = (dist < rc) & ~same

    # van der Waals, each unordered pair once
    I, J = torch.nonzero(pair & (torch.arange(n)[:, None] < torch.arange(n)[None, :]), as_tuple=True)
    xi, xj = _leaf(X[I]), _leaf(X[J])
    r = (xi - xj).norm(dim=1)
    (-tap(r) * ip["C6"] / (r**6 * (1.0 + torch.exp(-ip["d"] * (r / ip["seff"] - 1.0))))).sum().backward()
    v = _v(X[I] - X[J], -xi.grad)
    tally(I, v, 0.5)
    tally(J, v, 0.5)

    # repulsion of each ordered pair (i, j) with the normal of i, grouped by i's neighbour count
    I, J = torch.nonzero(pair, as_tuple=True)
    cnt = torch.tensor([len(nbr[a]) for a in I.tolist()])
    for c in (0, 1, 2, 3):
        sel = torch.nonzero(cnt == c, as_tuple=True)[0]
        if sel.numel() == 0:
            continue
        i, j = I[sel], J[sel]
        xi, xj = _leaf(X[i]), _leaf(X[j])
        if c >= 2:
            K = torch.stack([X[nbr[a]] for a in i.tolist()])  # [m, c, 3]
            xk = _leaf(K)
            if c == 3:
                nv = torch.cross(xk[:, 1] - xk[:, 0], xk[:, 2] - xk[:, 0], dim=1)
            else:
                xib = _leaf(X[i])  # the normal's own dependence on x_i: applied, not tallied
                nv = torch.cross(xk[:, 0] - xib, xk[:, 1] - xib, dim=1)
            nv = nv / nv.norm(dim=1, keepdim=True)
        else:
            nv = torch.tensor([0.0, 0.0, 1.0], dtype=f64).expand(i.shape[0], 3)
        d = xi - xj
        r = d.norm(dim=1)
        pn = (nv * d).sum(1)
        rho2 = r * r - pn * pn
        e = tap(r) * torch.exp(-ip["lam"] * (r - ip["z0"])) * (0.5 * ip["eps"] + ip["C"] * torch.exp(-rho2 / ip["delta"] ** 2))
        e.sum().backward()
        fkc = xj.grad  # the reference's fkc: f[j] -= fkc
        v = _v(X[i] - X[j], fkc)
        tally(i, v, 0.5)
        tally(j, v, 0.5)
        if c >= 2:
            for q in range(c):
                k = torch.tensor([int(nbr[a][q]) for a in i.tolist()])
                other = j if k_partner == "j" else i
                v = _v(X[k] - X[other], -xk.grad[:, q])
                tally(k, v, 0.5)
                tally(other, v, 0.5)
    return W.numpy()
