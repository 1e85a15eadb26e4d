"""C ABI: the engine library loads on a CPU-only host, exports every symbol include/xmvb_amd.h
declares, and its host-side entry points ($CTRL flag, option checks, synthetic problems) behave
like the reference's (no GPU calls here)."""
import hashlib
import json
import os
import re

import numpy as np
import pytest

import xmvb_amd as vb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xmvb_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vbe_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = vb.library()
    declared = header_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(vb.EXPORTED_SYMBOLS) == declared


@pytest.mark.parametrize("line,method", [("VBSCF", "VBSCF"), ("hipvbscf EIGENSOLVER=DENSE", "HIPVBSCF"),
                                         ("CUVBSCF MOLDEN ITMAX=10", "CUVBSCF"), ("", "VBSCF"),
                                         ("casscf", "CASSCF"), ("DFT", "DFT"), ("HIPVBSCF HIPVBSCF", "HIPVBSCF")])
def test_ctrl_method_flag(line, method):
    assert vb.parse_ctrl_method(line) == method


@pytest.mark.parametrize("line,status", [("VBSCF HIPVBSCF", 1), ("HIPVBSCF CUVBSCF", 1), ("TBVBSCF", 5),
                                         ("VBPDFT", 5)])
def test_ctrl_method_flag_rejections(line, status):
    # conflicting methods throw std::invalid_argument in apply_ctrl_flag (src/input/deck/model.cpp:360-381)
    with pytest.raises(vb.VbeError) as e:
        vb.parse_ctrl_method(line)
    assert e.value.status == status


def test_hipvbscf_option_rejections():
    vb.check_hipvbscf_options(eigensolver=0, ri=True, n_states=1, n_inactive_electrons=0)
    vb.check_hipvbscf_options(eigensolver=1, ri=True, n_states=1, n_inactive_electrons=0)
    vb.check_hipvbscf_options(eigensolver=0, ri=True, n_states=3, n_inactive_electrons=0)  # NSTATE=3, dense
    for kwargs in (dict(eigensolver=2), dict(ri=False), dict(n_states=2, eigensolver=1), dict(n_states=0),
                   dict(n_states=9), dict(n_inactive_electrons=2)):
        with pytest.raises(vb.VbeError) as e:
            vb.check_hipvbscf_options(**kwargs)
        assert e.value.status == 5


def test_synthetic_generator_is_deterministic_and_well_formed():
    a = vb.synthetic_problem(123, 20, 4, 30)
    b = vb.synthetic_problem(123, 20, 4, 30)
    c = vb.synthetic_problem(124, 20, 4, 30)
    for f in ("S", "h", "B", "C"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    assert not np.array_equal(a.C, c.C)
    assert np.allclose(a.S, a.S.T) and np.linalg.eigvalsh(a.S).min() >= 1.0 - 1e-12
    assert np.allclose(a.h, a.h.T)
    assert np.diag(a.h).mean() < -1.0


def test_synthetic_scan_scaling():
    near = vb.synthetic_problem(5, 10, 2, 4, scan_r=0.0)
    far = vb.synthetic_problem(5, 10, 2, 4, scan_r=2.0)
    f = np.exp(-0.5 * 2.0)
    assert np.allclose(far.h[:5, 5:], f * near.h[:5, 5:]) and np.allclose(far.h[:5, :5], near.h[:5, :5])
    assert np.array_equal(far.S, near.S)


def test_synthetic_golden_checksums():
    # tests/golden/synthetic.json pins the medium-config generator output (seed 0)
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "synthetic.json")))
    for key, rec in gold.items():
        sp = vb.synthetic_problem(rec["seed"], rec["n_bf"], rec["n_act"], rec["n_aux"])
        for f in ("S", "h", "B", "C"):
            arr = np.ascontiguousarray(getattr(sp, f).ravel(order="F") if f in "Sh" else getattr(sp, f).ravel())
            assert hashlib.sha256(arr.tobytes()).hexdigest() == rec[f], (key, f)


def test_topology_view_roundtrip():
    t = vb.Topology.build("COV", 6, 6, 1)
    v = t.view()
    assert v.n_structures == 5 and v.n_determinants == 20 and v.n_alpha_strings == 20
    assert v.alpha_masks[0] == t.alpha_masks[0] and v.term_offsets[20] == len(t.term_structure)
