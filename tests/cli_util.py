"""Helpers for running the hipvbscf command-line driver in tests."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xmvb-cpp_amd", "bin", "hipvbscf")
H2_DECK = os.path.join(ROOT, "tests", "decks", "h2_631gss.xmi")


def run_cli(*args, timeout=120):
    if not os.path.exists(CLI):
        raise FileNotFoundError(f"{CLI} is not built (run __graft_entry__.build())")
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout, p.stderr


def diagnostics(stdout: str) -> dict:
    """key=value lines of the run diagnostics block."""
    out = {}
    for line in stdout.splitlines():
        m = re.fullmatch(r"([a-z_]+)=(.*)", line.strip())
        if m:
            out[m.group(1)] = m.group(2)
    return out


def guess_rows(stdout: str) -> list:
    """'(coefficient, AO index)' rows printed by --dry-run, as {index: coefficient} per orbital."""
    rows, on = [], False
    for line in stdout.splitlines():
        if line.strip().startswith("Guess orbitals"):
            on = True
            continue
        if on and re.match(r"\s*\d+:", line):
            vals = line.split(":", 1)[1].split()
            rows.append({int(vals[k + 1]): float(vals[k]) for k in range(0, len(vals), 2)})
    return rows


def deck_variant(src: str, dst: str, replace: dict = None, append: str = ""):
    text = open(src).read()
    for a, b in (replace or {}).items():
        assert a in text, a
        text = text.replace(a, b)
    with open(dst, "w") as f:
        f.write(text + append)
    return dst
