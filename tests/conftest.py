"""Shared test setup: import paths, the `gpu` marker, common model helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FEMN = os.path.join(GOLDEN, "femn")
AU_TO_EV = 27.211385

# Wannier centres of the FeMn fixtures: 4 FCC sites x 2 spins
# (scripts/FeMn_WX_setup.py:133-137, 406)
FEMN_ATOMS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
FEMN_CENTERS = np.asfortranarray(np.concatenate([FEMN_ATOMS, FEMN_ATOMS]).T)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def femn_model(delta="0.900"):
    from mepinterp_amd.formats import read_hr_dat, read_latt
    tag = {"0.900": "0.9", "1.000": "1.0", "1.100": "1.1"}[delta]
    d = os.path.join(FEMN, f"delta{delta}")
    R, H, deg = read_hr_dat(os.path.join(d, "FeMn_hr.dat"))
    latt = read_latt(os.path.join(d, f"FeMn_d{tag}.latt"))
    return dict(R_frac=R, H=H, r=None, wf_centers=FEMN_CENTERS, a_latt=latt, deg=deg,
                name=f"FeMn_d{tag}")


def femn_golden(delta="0.900"):
    s = {"0.900": "09", "1.000": "10", "1.100": "11"}[delta]
    d = os.path.join(FEMN, f"delta{delta}")
    return (np.load(os.path.join(d, "ahc_tens.npy")),
            np.load(os.path.join(d, f"ahcVELO_sol{s}.npy")))


def oracle_model(m, **kw):
    import oracle as O
    return O.Model(m["R_frac"], m["H"], r=m.get("r"), wf_centers=m.get("wf_centers"),
                   a_latt=m["a_latt"], **kw)


def rand_kpts(nk, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(-0.5, 0.5, size=(3, nk)))


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle as O
    O.build()
    return O
