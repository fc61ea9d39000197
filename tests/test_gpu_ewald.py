"""FP64 Ewald on the GPU against the reference's known answer
(test/ewald_test.cpp:126-156) and the oracle's Ewald.

Bars: E/(4 pi) of the 100-charge fixture within the fixture's 1e-10 for every
alpha; the potential at the fixture point independent of alpha to 1e-10;
energy and a chained trial-move dE sequence equal to the oracle's to 1e-10
relative (energy) and 1e-10 absolute (dE).
"""
import json
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O
import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu
FX = json.loads((pathlib.Path(__file__).parent / "golden" / "ewald_100.json").read_text())


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("alpha", FX["alphas"])
def test_gpu_ewald_known_answer(alpha):
    log("gpu ewald known answer alpha", alpha)
    e = P.Ewald(FX["L"], FX["s"], alpha, FX["eps"], np.array(FX["q"]), np.array(FX["r"]))
    assert abs(e.compute_energy() / (4 * np.pi) - FX["E_over_4pi"]) < FX["tol"]
    e.close()


def test_gpu_ewald_potential_alpha_independent_and_matches_oracle():
    q, r = np.array(FX["q"]), np.array(FX["r"])
    pt = np.array(FX["potential_point"])
    pots = []
    for a in FX["alphas"]:
        e = P.Ewald(FX["L"], FX["s"], a, FX["eps"], q, r)
        pots.append(e.compute_potential(pt)[0])
        e.close()
    assert max(pots) - min(pots) < 1e-10
    assert abs(pots[-1] - O.ewald_potential(r, q, FX["L"], FX["s"], FX["alphas"][-1], pt)) < 1e-10


def test_gpu_ewald_energy_and_shift_chain_match_oracle():
    log("gpu ewald vs oracle")
    L = 20.0
    r, q = O.reference_test_system(2000, L)
    e = P.Ewald(L, 5.0, 1.0, 1.0, q, r)
    E = e.compute_energy()
    Eo = O.ewald_energy(r, q, L, 5.0, 1.0)
    assert abs(E - Eo) < 1e-10 * abs(Eo)
    idx, d = O.reference_test_moves(2000, L, 20)
    apply = np.arange(20) % 3 != 0
    dE = e.shift_chain(idx, d, apply)
    dEo = O.ewald_shift_chain(r, q, L, 5.0, 1.0, idx, d, apply.astype(np.float64))
    assert np.abs(dE - dEo).max() < 1e-10, (dE, dEo)
    # after the applied moves the total energy is the energy of the moved system
    r2 = r.copy()
    for k in range(20):
        if apply[k]:
            r2[idx[k]] = np.mod(r2[idx[k]] + d[k], L)
    assert abs(e.compute_energy() - O.ewald_energy(r2, q, L, 5.0, 1.0)) < 1e-10 * abs(Eo)
    e.close()
