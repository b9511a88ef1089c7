"""Seeded generators for the layered systems the LMFF force path is run on.

Every generator returns a ``System`` holding LAMMPS-style per-atom arrays
(``x`` [n,3] float64, ``type`` 1-based int32, ``tag`` 1-based int64,
``mol`` = layer id int64) plus the orthogonal box.  Units are LAMMPS
``metal`` (Angstrom, ps, eV, g/mol).

Geometry conventions
--------------------
* Rectangular 4-atom graphene cell a = 2.46 A (x, zigzag) by b = 4.261 A (y),
  basis (0,0), (1/2,1/6), (1/2,1/2), (0,2/3) in fractions of (a, b); the
  C-C bond is 1.4203 A.
* hBN uses the same basis strained to the graphene lattice, B on (0,0) and
  (1/2,1/2), N on (1/2,1/6) and (0,2/3).
* AB stacking shifts every other layer by b/3 along y.
* x and y are periodic; z is open (the reference decks put >= 20 A of
  vacuum along z, which is beyond the 18 A neighbor cutoff, so no image in z
  ever interacts).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# LAMMPS metal units (reference: LMFF/update.cpp:167-172)
BOLTZ = 8.617343e-5
MVV2E = 1.0364269e-4
FTM2V = 1.0 / 1.0364269e-4

MASS = {"C": 12.0107, "B": 10.811, "N": 14.0067}

A_CELL = 2.46
B_CELL = 4.261
BASIS = np.array([[0.0, 0.0], [0.5, 1.0 / 6.0], [0.5, 0.5], [0.0, 2.0 / 3.0]])
HBN_SPECIES = ("B", "N", "B", "N")


@dataclass
class System:
    x: np.ndarray
    type: np.ndarray
    tag: np.ndarray
    mol: np.ndarray
    boxlo: np.ndarray
    boxhi: np.ndarray
    periodic: tuple
    elements: list          # element name of each atom type, index 0 = type 1
    v: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def mass_per_type(self) -> np.ndarray:
        return np.array([MASS[e] for e in self.elements], dtype=np.float64)

    def masses(self) -> np.ndarray:
        return self.mass_per_type[self.type - 1]

    def copy(self) -> "System":
        return System(self.x.copy(), self.type.copy(), self.tag.copy(), self.mol.copy(),
                      self.boxlo.copy(), self.boxhi.copy(), tuple(self.periodic),
                      list(self.elements), None if self.v is None else self.v.copy(),
                      dict(self.meta))


def _layer_species(kind: str):
    if kind == "G":
        return ("C", "C", "C", "C")
    if kind == "BN":
        return HBN_SPECIES
    raise ValueError(f"unknown layer kind {kind!r}")


def layered_stack(nx: int, ny: int, layers, spacing: float = 3.35, jitter: float = 0.02,
                  seed: int = 12345, temperature: float | None = 300.0,
                  elements=None, z0: float = 0.0, a: float = A_CELL, b: float = B_CELL) -> System:
    """Stack of rectangular-cell layers, ``layers`` a sequence of "G"/"BN".

    ``elements`` lists the element of each atom type (default: C only for an
    all-graphene stack, else C, B, N).  Positions are jittered by a seeded
    Gaussian of width ``jitter`` and velocities drawn from Maxwell-Boltzmann at
    ``temperature`` with zero net momentum, rescaled to the exact temperature.
    """
    layers = list(layers)
    if elements is None:
        elements = ["C"] if all(l == "G" for l in layers) else ["C", "B", "N"]
    tmap = {e: k + 1 for k, e in enumerate(elements)}
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel()], axis=1).astype(np.float64)
    xs, ts, ms = [], [], []
    for li, kind in enumerate(layers):
        spec = _layer_species(kind)
        shift_y = (b / 3.0) if (li % 2 == 1) else 0.0
        for bi in range(4):
            fx = (cells[:, 0] + BASIS[bi, 0]) * a
            fy = (cells[:, 1] + BASIS[bi, 1]) * b + shift_y
            pos = np.stack([fx, fy, np.full_like(fx, z0 + li * spacing)], axis=1)
            xs.append(pos)
            ts.append(np.full(len(fx), tmap[spec[bi]], dtype=np.int32))
            ms.append(np.full(len(fx), li + 1, dtype=np.int64))
    # atoms are ordered layer by layer, then by basis site, then by cell
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    m = np.concatenate(ms)
    Lx, Ly = nx * a, ny * b
    x[:, 1] = np.mod(x[:, 1], Ly)
    rng = np.random.Generator(np.random.PCG64(seed))
    if jitter > 0:
        x = x + rng.normal(0.0, jitter, size=x.shape)
    x[:, 0] = np.mod(x[:, 0], Lx)
    x[:, 1] = np.mod(x[:, 1], Ly)
    ztop = z0 + (len(layers) - 1) * spacing
    sysm = System(x=x, type=t, tag=np.arange(1, len(x) + 1, dtype=np.int64), mol=m,
                  boxlo=np.array([0.0, 0.0, z0 - 10.0]), boxhi=np.array([Lx, Ly, ztop + 10.0]),
                  periodic=(1, 1, 0), elements=list(elements),
                  meta={"nx": nx, "ny": ny, "layers": layers, "seed": seed})
    if temperature is not None:
        sysm.v = maxwell_boltzmann(sysm, temperature, rng)
    return sysm


def maxwell_boltzmann(s: System, temperature: float, rng) -> np.ndarray:
    m = s.masses()
    sigma = np.sqrt(BOLTZ * temperature / (m * MVV2E))
    v = rng.normal(0.0, 1.0, size=(s.n, 3)) * sigma[:, None]
    p = (v * m[:, None]).sum(axis=0)
    v -= p / m.sum()
    t_now = temperature_of(s, v)
    if t_now > 0:
        v *= math.sqrt(temperature / t_now)
    return v


def kinetic_energy(s: System, v: np.ndarray) -> float:
    return 0.5 * MVV2E * float((s.masses()[:, None] * v * v).sum())


def temperature_of(s: System, v: np.ndarray) -> float:
    dof = 3 * s.n - 3
    return 2.0 * kinetic_energy(s, v) / (dof * BOLTZ)


# ---------------------------------------------------------------------------
# Commensurate twisted bilayer graphene
# ---------------------------------------------------------------------------

def twist_angle(m: int, n: int) -> float:
    """Commensurate twist angle (radians) of the (m, n) moire cell."""
    c = (m * m + 4 * m * n + n * n) / (2.0 * (m * m + m * n + n * n))
    return math.acos(max(-1.0, min(1.0, c)))


def _rot(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, -s], [s, c]])


def twisted_bilayer(m: int, n: int, min_atoms: int = 2000, max_atoms: int = 80000,
                    min_side: float = 36.5, spacing: float = 3.35, a: float = 2.46,
                    jitter: float = 0.02, seed: int = 0, temperature: float | None = 300.0,
                    target_atoms: int | None = None) -> System:
    """Twisted bilayer graphene on the rectangular (m, n) commensurate cell.

    Layer 1 is the hexagonal lattice a1 = a(1,0), a2 = a(1/2, sqrt3/2) with a
    2-atom basis; layer 2 is layer 1 rotated so that its lattice vector
    n*a1 + m*a2 lands on m*a1 + n*a2.  The hexagonal moire cell (T1, R60 T1) is
    turned into the rectangular cell |T1| x sqrt3|T1| (4(m^2+mn+n^2) atoms per
    layer), rotated so T1 lies on x, and replicated so every side is at least
    ``min_side`` and the atom count lies in [min_atoms, max_atoms] (as close to
    ``target_atoms`` as the replication allows, when given).
    """
    if not (1 <= n < m):
        raise ValueError("need 1 <= n < m")
    a1 = np.array([a, 0.0])
    a2 = np.array([0.5 * a, 0.5 * math.sqrt(3.0) * a])
    basis = [np.zeros(2), (a1 + a2) / 3.0]
    T1 = m * a1 + n * a2
    v2 = n * a1 + m * a2
    theta = math.atan2(T1[1], T1[0]) - math.atan2(v2[1], v2[0])
    phi = -math.atan2(T1[1], T1[0])          # puts T1 on +x
    L = float(np.linalg.norm(T1))
    Lx, Ly = L, math.sqrt(3.0) * L
    per_layer = 4 * (m * m + m * n + n * n)
    layers = []
    for rot in (0.0, theta):
        R = _rot(phi) @ _rot(rot)
        A = np.stack([R @ a1, R @ a2], axis=1)
        kmax = int(math.ceil(2.5 * (Ly / a))) + 2
        ii, jj = np.meshgrid(np.arange(-kmax, kmax + 1), np.arange(-kmax, kmax + 1), indexing="ij")
        lat = ii.ravel()[:, None] * A[:, 0][None, :] + jj.ravel()[:, None] * A[:, 1][None, :]
        pts = []
        for b in basis:
            p = lat + (R @ b)[None, :]
            fx = p[:, 0] / Lx
            fy = p[:, 1] / Ly
            keep = (fx >= -1e-9) & (fx < 1 - 1e-9) & (fy >= -1e-9) & (fy < 1 - 1e-9)
            pts.append(p[keep])
        p = np.concatenate(pts)
        p[:, 0] = np.mod(p[:, 0], Lx)
        p[:, 1] = np.mod(p[:, 1], Ly)
        key = np.round(p / 1e-6).astype(np.int64)
        _, uniq = np.unique(key, axis=0, return_index=True)
        p = p[np.sort(uniq)]
        if len(p) != per_layer:
            raise RuntimeError(f"twist ({m},{n}) built {len(p)} atoms per layer, expected {per_layer}")
        layers.append(p)
    rx = max(1, math.ceil(min_side / Lx))
    ry = max(1, math.ceil(min_side / Ly))
    cell_atoms = 2 * per_layer
    want = target_atoms if target_atoms is not None else min_atoms
    while rx * ry * cell_atoms < want:
        if rx * Lx <= ry * Ly:
            rx += 1
        else:
            ry += 1
    if rx * ry * cell_atoms > max_atoms:
        raise ValueError(f"twist ({m},{n}) cannot fit {max_atoms} atoms with sides >= {min_side}")
    xs, ms = [], []
    for li, p in enumerate(layers):
        for ix in range(rx):
            for iy in range(ry):
                q = p + np.array([ix * Lx, iy * Ly])
                xs.append(np.column_stack([q, np.full(len(q), li * spacing)]))
                ms.append(np.full(len(q), li + 1, dtype=np.int64))
    x = np.concatenate(xs)
    mol = np.concatenate(ms)
    rng = np.random.Generator(np.random.PCG64(seed))
    if jitter > 0:
        x = x + rng.normal(0.0, jitter, size=x.shape)
    BX, BY = rx * Lx, ry * Ly
    x[:, 0] = np.mod(x[:, 0], BX)
    x[:, 1] = np.mod(x[:, 1], BY)
    s = System(x=x, type=np.ones(len(x), dtype=np.int32), tag=np.arange(1, len(x) + 1, dtype=np.int64),
               mol=mol, boxlo=np.array([0.0, 0.0, -10.0]), boxhi=np.array([BX, BY, spacing + 10.0]),
               periodic=(1, 1, 0), elements=["C"],
               meta={"m": m, "n": n, "theta_deg": math.degrees(theta), "rep": (rx, ry), "seed": seed})
    if temperature is not None:
        s.v = maxwell_boltzmann(s, temperature, rng)
    return s


def config_c1(seed: int = 12345) -> System:
    """AB bilayer graphene, 2 x (30x20 cells) = 4,800 C atoms."""
    return layered_stack(30, 20, ["G", "G"], seed=seed)


def config_c2(seed: int = 12346) -> System:
    """Graphene/hBN heterobilayer, 220x130 cells per layer = 228,800 atoms."""
    return layered_stack(220, 130, ["G", "BN"], seed=seed)


def config_c3(seed: int = 12347) -> System:
    """10-layer alternating graphene/hBN stack, 260x260 cells = 2,704,000 atoms."""
    return layered_stack(260, 260, ["G", "BN"] * 5, seed=seed)


def config_c5(seed: int = 12349) -> System:
    """4-layer alternating graphene/hBN, 1118x1118 cells = 19,998,784 atoms."""
    return layered_stack(1118, 1118, ["G", "BN"] * 2, seed=seed)


def read_lammps_data(path: str) -> System:
    """Reads an ``atom_style full`` LAMMPS data file (Atoms + optional Velocities).

    Follows the section layout LAMMPS' ``read_data`` accepts (LMFF/read_data.cpp:116):
    header counts and box bounds, ``Masses``, ``Atoms # full`` rows
    ``id mol type q x y z [ix iy iz]`` and ``Velocities`` rows ``id vx vy vz``.  The image flags
    (0 where a row has none) are returned as ``meta["image"]``, [n][3] in id order.
    """
    with open(path) as fh:
        lines = fh.read().splitlines()
    natoms = ntypes = 0
    lo = np.zeros(3)
    hi = np.zeros(3)
    masses = {}
    atoms = []
    vel = {}
    section = None
    for raw in lines[1:]:
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        words = line.split()
        if len(words) >= 2 and words[1] == "atoms" and section is None:
            natoms = int(words[0])
            continue
        if len(words) >= 3 and words[1] == "atom" and words[2] == "types":
            ntypes = int(words[0])
            continue
        if len(words) == 4 and words[2].endswith("lo") and words[3].endswith("hi"):
            d = "xyz".index(words[2][0])
            lo[d], hi[d] = float(words[0]), float(words[1])
            continue
        if words[0] in ("Masses", "Atoms", "Velocities", "Bonds", "Angles", "Pair"):
            section = words[0]
            continue
        if section == "Masses":
            masses[int(words[0])] = float(words[1])
        elif section == "Atoms":
            atoms.append(words)
        elif section == "Velocities":
            vel[int(words[0])] = [float(w) for w in words[1:4]]
    if len(atoms) != natoms:
        raise ValueError(f"{path}: expected {natoms} atoms, read {len(atoms)}")
    ids = np.array([int(a[0]) for a in atoms], dtype=np.int64)
    mol = np.array([int(a[1]) for a in atoms], dtype=np.int64)
    typ = np.array([int(a[2]) for a in atoms], dtype=np.int32)
    x = np.array([[float(a[4]), float(a[5]), float(a[6])] for a in atoms])
    img = np.array([[int(w) for w in a[7:10]] if len(a) >= 10 else [0, 0, 0] for a in atoms],
                   dtype=np.int32).reshape(-1, 3)
    order = np.argsort(ids, kind="stable")
    s = System(x=x[order], type=typ[order], tag=ids[order], mol=mol[order], boxlo=lo, boxhi=hi,
               periodic=(1, 1, 1), elements=["C"] * max(ntypes, 1),
               meta={"masses": masses, "image": img[order]})
    if vel:
        s.v = np.array([vel[int(t)] for t in s.tag])
    return s


def write_lammps_data(s: System, path: str, step: int = 0) -> None:
    """Writes ``s`` as an ``atom_style full`` data file in the layout of LAMMPS' ``write_data``
    (LMFF/write_data.cpp:236-420, AtomVecFull::write_data): header with the timestep, counts,
    box, ``Masses``, ``Atoms # full`` rows ``id mol type q x y z ix iy iz`` (q 0; image flags from
    ``meta["image"]`` when present, else 0) and ``Velocities``, numbers in ``%-1.16e`` so that a
    read gives the same doubles back."""
    order = np.argsort(s.tag, kind="stable")
    img = np.asarray(s.meta.get("image", ()), dtype=np.int64).reshape(-1, 3)
    if img.shape[0] != s.n:  # none, or left over from a system of another size
        img = np.zeros((s.n, 3), dtype=np.int64)
    with open(path, "w") as fh:
        fh.write(f"LAMMPS data file via write_data, lmff, timestep = {int(step)}\n\n")
        fh.write(f"{s.n} atoms\n{len(s.elements)} atom types\n\n")
        for d, ax in enumerate("xyz"):
            fh.write(f"{s.boxlo[d]:<1.16e} {s.boxhi[d]:<1.16e} {ax}lo {ax}hi\n")
        fh.write("\nMasses\n\n")
        for t, m in enumerate(s.mass_per_type, start=1):
            fh.write(f"{t} {m:g}\n")
        fh.write("\nAtoms # full\n\n")
        for i in order:
            x = s.x[i]
            fh.write(f"{int(s.tag[i])} {int(s.mol[i])} {int(s.type[i])} {0.0:<1.16e} "
                     f"{x[0]:<1.16e} {x[1]:<1.16e} {x[2]:<1.16e} {img[i, 0]} {img[i, 1]} {img[i, 2]}\n")
        if s.v is not None:
            fh.write("\nVelocities\n\n")
            for i in order:
                v = s.v[i]
                fh.write(f"{int(s.tag[i])} {v[0]:<1.16e} {v[1]:<1.16e} {v[2]:<1.16e}\n")


def read_dump_custom(path: str):
    """Reads a text ``dump custom`` file (DumpCustom::header_item, LMFF/dump_custom.cpp:422-440):
    returns a list of (timestep, column names, float array [natoms][ncolumns])."""
    out = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = 0
    while k < len(lines):
        if lines[k].startswith("ITEM: TIMESTEP"):
            step = int(lines[k + 1])
            n = int(lines[k + 3])
            cols = lines[k + 8].split()[2:]
            rows = np.array([[float(w) for w in lines[k + 9 + r].split()] for r in range(n)]).reshape(n, len(cols))
            out.append((step, cols, rows))
            k += 9 + n
        else:
            k += 1
    return out
