"""Pins the oracle's FP64 Ewald against the reference's known answer
(test/ewald_test.cpp:126-156, EwaldSummations.jl) and checks the
potential-difference trial energy against two full energies."""
import json
import pathlib

import numpy as np
import pytest

import oracle_lib as O

FX = json.loads((pathlib.Path(__file__).parent / "golden" / "ewald_100.json").read_text())


@pytest.mark.parametrize("alpha", FX["alphas"])
def test_ewald_known_answer(alpha):
    r = np.array(FX["r"])
    q = np.array(FX["q"])
    E = O.ewald_energy(r, q, FX["L"], FX["s"], alpha, FX["eps"])
    assert abs(E / (4 * np.pi) - FX["E_over_4pi"]) < FX["tol"]


def test_ewald_potential_alpha_independent():
    r = np.array(FX["r"])
    q = np.array(FX["q"])
    pots = [O.ewald_potential(r, q, FX["L"], FX["s"], a, FX["potential_point"]) for a in FX["alphas"]]
    assert max(pots) - min(pots) < 1e-10


def test_shift_energy_by_potential_difference():
    L = 20.0
    r, q = O.reference_test_system(300, L)
    idx, d = O.reference_test_moves(300, L, 3)
    dE = O.ewald_shift_chain(r, q, L, 5.0, 1.0, idx, d, np.zeros(3))
    for k in range(3):
        r2 = r.copy()
        r2[idx[k]] = np.mod(r[idx[k]] + d[k], L)
        two = O.ewald_energy(r2, q, L, 5.0, 1.0) - O.ewald_energy(r, q, L, 5.0, 1.0)
        assert abs(two - dE[k]) < 1e-10
