"""The hipvbscf driver end to end on the device with the H2 deck (6-31G**, STR=FULL, P-LBFGS from the
core-Hamiltonian block guess): the energy and iteration count must equal the oracle's P-LBFGS
restatement on the same integrals and start (1e-10 Eh; the guess orbitals may differ in sign, which
the optimiser's iterates follow), the Coulson-Chirgwin weights must sum to 1, and a restart from
the written .orb file through $GUS (GUESS=READ) must stop at once at the same energy."""
import os
import re
import shutil

import numpy as np
import pytest

from oracle.plbfgs import plbfgs
import xmvb_amd as vb
from cli_util import H2_DECK, deck_variant, diagnostics, run_cli

pytestmark = pytest.mark.gpu


def _weights(stdout):
    rows, on = [], False
    for line in stdout.splitlines():
        if "Coulson-Chirgwin weight" in line:
            on = True
            continue
        if on:
            m = re.fullmatch(r"\s*(\d+)\s+(\S+)\s+(\S+)\s*", line)
            if not m:
                break
            rows.append((float(m.group(2)), float(m.group(3))))
    return rows


@pytest.fixture(scope="module")
def oracle_run():
    d = vb.h2_deck()
    ref = plbfgs(d["n_bf"], d["n_aux"], d["topo"], d["S"], d["h"], d["B"], d["C0"], support=d["support"])
    return d, ref


@pytest.mark.parametrize("solver", ["DENSE", "DAVIDSON"])
def test_h2_deck_end_to_end(tmp_path, oracle_run, solver):
    d, ref = oracle_run
    deck = deck_variant(H2_DECK, str(tmp_path / "h2.xmi"), {"EIGENSOLVER=DENSE": f"EIGENSOLVER={solver}"})
    print(f"case cli h2 {solver}", flush=True)
    rc, out, err = run_cli(deck, timeout=300)
    assert rc == 0, out[-2000:] + err
    dg = diagnostics(out)
    e_ref = ref["energy"] + d["e_nuc"]
    print(f"cli E={dg['total_energy_hartree']} ({dg['scf_iterations']} it, {dg['termination']}); "
          f"oracle E={e_ref:.12f} ({ref['iterations']} it)", flush=True)
    assert dg["converged"] == "1" and dg["termination"] == ref["termination"]
    assert int(dg["scf_iterations"]) == ref["iterations"]
    assert abs(float(dg["total_energy_hartree"]) - e_ref) <= 1e-10
    assert os.path.exists(tmp_path / "h2.orb") and dg["orbital_output"].endswith("h2.orb")
    w = _weights(out)
    assert len(w) == 3
    assert abs(sum(x[1] for x in w) - 1.0) <= 1e-9


def test_restart_from_orbital_file(tmp_path, oracle_run):
    d, ref = oracle_run
    first = deck_variant(H2_DECK, str(tmp_path / "h2.xmi"))
    rc, out, err = run_cli(first, timeout=300)
    assert rc == 0, err
    e1 = float(diagnostics(out)["total_energy_hartree"])
    orb = open(tmp_path / "h2.orb").read()
    second = deck_variant(H2_DECK, str(tmp_path / "h2_restart.xmi"), {"ISCF=8": "ISCF=8 GUESS=READ"},
                          append="$GUS\n" + orb + "$END\n")
    print("case cli h2 restart", flush=True)
    rc, out, err = run_cli(second, timeout=300)
    assert rc == 0, err
    dg = diagnostics(out)
    assert "guess: READ" in out
    assert int(dg["scf_iterations"]) <= 1
    assert abs(float(dg["total_energy_hartree"]) - e1) <= 1e-9


def test_h2_state_average_matches_oracle(tmp_path):
    """NSTATE=2 (equal weights) on the H2 deck with the dense eigensolver: the state-averaged
    P-LBFGS run equals the oracle's on the same integrals and start, and both state energies are
    reported."""
    d = vb.h2_deck()
    ref = plbfgs(d["n_bf"], d["n_aux"], d["topo"], d["S"], d["h"], d["B"], d["C0"], support=d["support"],
                 states=[0, 1], weights=[1.0, 1.0])
    deck = deck_variant(H2_DECK, str(tmp_path / "h2sa.xmi"), {"NAE=2": "NAE=2 NSTATE=2"})
    print("case cli h2 NSTATE=2", flush=True)
    rc, out, err = run_cli(deck, timeout=300)
    assert rc == 0, out[-2000:] + err
    dg = diagnostics(out)
    e_ref = ref["energy"] + d["e_nuc"]
    print(f"cli E={dg['total_energy_hartree']} ({dg['scf_iterations']} it); oracle E={e_ref:.12f} "
          f"({ref['iterations']} it)", flush=True)
    assert dg["states"] == "2" and dg["converged"] == "1"
    assert int(dg["scf_iterations"]) == ref["iterations"]
    assert abs(float(dg["total_energy_hartree"]) - e_ref) <= 1e-10
    es = [float(m) for m in re.findall(r"State \d+ total energy\s+(\S+) Eh", out)]
    assert len(es) == 2 and es[0] < es[1]
    assert abs(0.5 * (es[0] + es[1]) - e_ref) <= 1e-9
