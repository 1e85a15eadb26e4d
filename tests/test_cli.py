"""The hipvbscf command-line driver on the CPU (--dry-run: deck, integrals, supports and guess, no
device work) and its rejections of options HIPVBSCF does not implement, in the manner of
run_cuvbscf (src/cli/cuvbscf.cpp:38-50) and main (src/cli/main.cpp:8-20: "XMVB-CPP error: ...",
status 2). The H2 deck's integrals, supports and core-Hamiltonian block guess must equal the ones
the Python deck builder makes from the same library (orbital signs are arbitrary)."""
import os

import numpy as np
import pytest

import xmvb_amd as vb
from cli_util import H2_DECK, deck_variant, diagnostics, guess_rows, run_cli


def test_h2_dry_run_matches_the_deck_builder():
    rc, out, err = run_cli("--dry-run", H2_DECK)
    assert rc == 0, err
    assert "basis functions: 10" in out and "structures (STR=FULL): 3" in out and "determinants: 4" in out
    d = vb.h2_deck()
    assert f"factors of rank {d['n_aux']}" in out
    assert "Orbital 1 support: 1 2 3 4 5" in out and "Orbital 2 support: 6 7 8 9 10" in out
    rows = guess_rows(out)
    assert len(rows) == 2
    for k, row in enumerate(rows):
        c = np.zeros(d["n_bf"])
        for idx, v in row.items():
            c[idx - 1] = v
        ref = d["C0"][k]
        sign = np.sign(c @ ref)
        assert np.abs(sign * c - ref).max() <= 2e-10  # printed with 10 decimals


def test_read_guess_from_gus_block(tmp_path):
    gus = """$GUS
    5     5
# ORBITAL 1 NAO = 5
  0.5000000000      1   0.5000000000      2   0.0000000000      3   0.0000000000      4
  0.1000000000      5
# ORBITAL 2 NAO = 5
  0.5000000000      6   0.5000000000      7   0.0000000000      8   0.0000000000      9
 -0.1000000000     10
$END
"""
    deck = deck_variant(H2_DECK, str(tmp_path / "h2_read.xmi"), append=gus)
    rc, out, err = run_cli("--dry-run", deck)
    assert rc == 0, err
    assert "guess: READ" in out
    d = vb.h2_deck()
    S = d["S"]
    rows = guess_rows(out)
    for k, row in enumerate(rows):
        c = np.zeros(10)
        for idx, v in row.items():
            c[idx - 1] = v
        raw = np.zeros(10)
        raw[5 * k:5 * k + 5] = [0.5, 0.5, 0.0, 0.0, 0.1 if k == 0 else -0.1]
        ref = raw / np.sqrt(raw @ S @ raw)  # normalised on its support, as the reference's guess is
        assert np.abs(c - ref).max() <= 2e-10


def test_oeo_support_is_every_function(tmp_path):
    deck = deck_variant(H2_DECK, str(tmp_path / "h2_oeo.xmi"), {"ORBTYP=HAO": "ORBTYP=OEO"})
    rc, out, err = run_cli("--dry-run", deck)
    assert rc == 0, err
    assert "Orbital 1 support: 1 2 3 4 5 6 7 8 9 10" in out


@pytest.mark.parametrize("replace,message", [
    ({"HIPVBSCF": "CUVBSCF"}, "requires a HIPVBSCF input"),
    ({"ISCF=8": "ISCF=9"}, "P-LBFGS (ISCF=8) only"),
    ({"ISCF=8": "ISCF=5"}, "unsupported ISCF value 5"),
    ({"INT=LIBCINT": "INT=RI"}, "auxiliary basis"),
    ({"BASIS=6-31G**": "BASIS=CC-PVDZ"}, "not available"),
    ({"EIGENSOLVER=DENSE": "EIGENSOLVER=LANCZOS"}, "unsupported EIGENSOLVER value"),
    ({"NAE=2": "NAE=2 NSTATE=9"}, "1 to 8 averaged states"),
    ({"EIGENSOLVER=DENSE": "EIGENSOLVER=DAVIDSON NSTATE=2"}, "needs EIGENSOLVER=DENSE"),
    ({"NAE=2": "NAE=2 WSTATE(1)=0.5,0.25"}, "only equal positive WSTATE weights"),
    ({"NAE=2": "NAE=2 WSTATE(2)=0.5,0.5"}, "supports only WSTATE(1)"),
    ({"FRGTYP=ATOM": "FRGTYP=SAO"}, "FRGTYP=ATOM only"),
    ({"HIPVBSCF": "HIPVBSCF CASSCF"}, "only one electronic method"),
    ({"HIPVBSCF": "HIPVBSCF TBVBSCF"}, "no longer supported"),
    ({"H 0.0 0.0 0.74": "Q 0.0 0.0 0.74"}, "unknown element"),
])
def test_unsupported_options_are_rejected(tmp_path, replace, message):
    deck = deck_variant(H2_DECK, str(tmp_path / "bad.xmi"), replace)
    rc, out, err = run_cli("--dry-run", deck)
    assert rc == 2
    assert err.startswith("XMVB-CPP error:") and message in err, err


def test_explicit_structures_are_rejected(tmp_path):
    deck = deck_variant(H2_DECK, str(tmp_path / "str.xmi"), append="$STR\n1 2\n$END\n")
    rc, out, err = run_cli("--dry-run", deck)
    assert rc == 2 and "explicit $STR" in err


def test_missing_deck_is_an_error():
    rc, out, err = run_cli("--dry-run", "/nonexistent/deck.xmi")
    assert rc == 2 and "failed to open input file" in err
