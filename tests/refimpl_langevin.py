"""Restatement of the device thermostat in numpy (test infrastructure).

FixLangevin::post_force with the default uniform noise (LMFF/fix_langevin.cpp:297-303 for the
prefactors, 854-859 for the noise, 876-880 for the sum):
    f += gfactor1 * v + gfactor2 * sqrt(T) * (U - 0.5),
    gfactor1 = -m / damp / ftm2v,
    gfactor2 = sqrt(m) * sqrt(24 kB / damp / dt / mvv2e) / ftm2v,
with U uniform on (0, 1).  The reference draws U from a per-rank Marsaglia stream (RanMars,
fix_langevin.cpp:85), which depends on the rank layout; the build keys it instead by
(seed, step, atom tag) with Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11), written here from the paper: ten rounds of
    (L0, R0, L1, R1) -> (hi(M1 R1) ^ L0' ^ k0, lo(M1 R1), hi(M0 R0) ^ L1' ^ k1, lo(M0 R0)),
M0 = 0xD2511F53, M1 = 0xCD9E8D57, the key bumped by (0x9E3779B9, 0xBB67AE85) between rounds,
with the counter {step low word, step high word, tag, w} (w = 0x4c4d4646 for the step draws,
0x53455455 for the draw on the set-up force) and U_k = (out_k + 1/2) 2^-32.
"""
from __future__ import annotations

import numpy as np

FTM2V = 1.0 / 1.0364269e-4
MVV2E = 1.0364269e-4
BOLTZ = 8.617343e-5
M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
STEP_WORD, SETUP_WORD = 0x4C4D4646, 0x53455455
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: uint array [..., 4], key: uint array [..., 2] (broadcast) -> uint32 [..., 4]."""
    c = [np.asarray(ctr, dtype=np.uint64)[..., q] & MASK for q in range(4)]
    k0 = np.asarray(key, dtype=np.uint64)[..., 0] & MASK
    k1 = np.asarray(key, dtype=np.uint64)[..., 1] & MASK
    for r in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [(hi1 ^ c[1] ^ k0) & MASK, lo1, (hi0 ^ c[3] ^ k1) & MASK, lo0]
        k0 = (k0 + np.uint64(W0)) & MASK
        k1 = (k1 + np.uint64(W1)) & MASK
    return np.stack(c, -1).astype(np.uint32)


def uniforms(step, tags, seed, word=STEP_WORD):
    """U - 0.5 for the three components of every atom, [n, 3]."""
    tags = np.asarray(tags, dtype=np.uint64)
    n = tags.shape[0]
    ctr = np.zeros((n, 4), dtype=np.uint64)
    ctr[:, 0] = np.uint64(step) & MASK
    ctr[:, 1] = np.uint64(step) >> np.uint64(32)
    ctr[:, 2] = tags & MASK
    ctr[:, 3] = word
    key = np.array([int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF], dtype=np.uint64)
    out = philox4x32_10(ctr, key).astype(np.float64)
    return (out[:, :3] + 0.5) * 2.0 ** -32 - 0.5


def langevin_force(v, mass, tags, step, seed, t_target, damp, dt, word=STEP_WORD, kb_factor=24.0):
    """Drag plus noise of FixLangevin::post_force for velocities v [n, 3] and masses [n].
    kb_factor 24 is the uniform-noise default; 2 belongs to the gjf variant (a counter-example
    here)."""
    m = np.asarray(mass, dtype=np.float64)[:, None]
    gfactor1 = -m / damp / FTM2V
    gfactor2 = np.sqrt(m) * np.sqrt(kb_factor * BOLTZ / damp / dt / MVV2E) / FTM2V
    return gfactor1 * v + gfactor2 * np.sqrt(t_target) * uniforms(step, tags, seed, word)


def verlet_langevin(x, v, mass, tags, forces, nsteps, seed, t_target, damp, dt, step0=0, kb_factor=24.0):
    """Velocity Verlet with fix langevin (Verlet::setup + Verlet::run order: set-up force with
    post_force, then per step initial_integrate, force, post_force, final_integrate).  forces(x)
    gives the conservative force; the noise of the step that starts at counter s is keyed by s."""
    x = np.array(x, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    dtf = 0.5 * dt * FTM2V
    m = np.asarray(mass, dtype=np.float64)[:, None]
    f = forces(x) + langevin_force(v, mass, tags, step0, seed, t_target, damp, dt, SETUP_WORD, kb_factor)
    for k in range(nsteps):
        v = v + dtf / m * f
        x = x + dt * v
        f = forces(x) + langevin_force(v, mass, tags, step0 + k, seed, t_target, damp, dt, STEP_WORD, kb_factor)
        v = v + dtf / m * f
    return x, v
