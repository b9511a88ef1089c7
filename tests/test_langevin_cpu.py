"""Pins of the thermostat restatement (tests/refimpl_langevin.py), CPU only: Philox4x32-10
against the published known-answer vectors of its authors' Random123 library, and the
fluctuation-dissipation balance of the uniform-noise Langevin force (free particles relax to
the target temperature; the gjf prefactor, 2 kB instead of 24 kB, does not)."""
import numpy as np
import pytest

import refimpl_langevin as L


@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, out):
    got = L.philox4x32_10(np.array([ctr], dtype=np.uint64), np.array(key, dtype=np.uint64))[0]
    assert [int(g) for g in got] == out


def test_uniforms_are_uniform_and_distinct_per_key():
    tags = np.arange(1, 200001)
    u = L.uniforms(7, tags, seed=12345)
    assert u.min() > -0.5 and u.max() < 0.5
    assert abs(u.mean()) < 3e-3 and abs(u.var() - 1.0 / 12.0) < 1e-3
    # a different step, word or seed gives an unrelated stream
    for other in (L.uniforms(8, tags, 12345), L.uniforms(7, tags, 12346), L.uniforms(7, tags, 12345, L.SETUP_WORD)):
        assert abs(np.corrcoef(u.ravel(), other.ravel())[0, 1]) < 1e-2


def _free_particle_temperature(kb_factor, t_target=300.0, damp=0.1, dt=0.001, n=20000, steps=3000):
    mass = np.full(n, 12.0107)
    tags = np.arange(1, n + 1)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    temps = []
    burn = 1000
    dtf = 0.5 * dt * L.FTM2V
    f = L.langevin_force(v, mass, tags, 0, 99, t_target, damp, dt, L.SETUP_WORD, kb_factor)
    for k in range(steps):
        v = v + dtf / mass[:, None] * f
        x = x + dt * v
        f = L.langevin_force(v, mass, tags, k, 99, t_target, damp, dt, L.STEP_WORD, kb_factor)
        v = v + dtf / mass[:, None] * f
        if k >= burn and k % 20 == 0:
            ke = 0.5 * L.MVV2E * (mass[:, None] * v * v).sum()
            temps.append(2.0 * ke / (3 * n * L.BOLTZ))
    return float(np.mean(temps))


def test_uniform_noise_holds_the_target_temperature():
    t = _free_particle_temperature(24.0)
    assert abs(t / 300.0 - 1.0) < 0.02
    # the prefactor is what is pinned: the gjf variant's 2 kB is twelve times too little noise
    assert _free_particle_temperature(2.0) < 0.2 * 300.0
