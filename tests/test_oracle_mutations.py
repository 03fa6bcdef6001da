"""The oracle's known-answer pins catch a one-constant typo in any model (CPU).

Each case copies oracle/fw_oracle.c to a temporary directory, changes one
numeric literal inside one model's ionic sweep (a float constant scaled by
1.01, zero replaced by 0.0625, or an integer constant incremented), builds
that copy and runs the model's known-answer comparison
(tests/test_oracle_known_answers.py) against it.  Every literal of every
sweep is planted in turn (comments and array indices excluded), and every
planted mistake must push the comparison past its tolerance, except the
literals listed in UNDETECTABLE, which no comparison of outputs can see:
  * luo_rudy91_2d.py's calc_ina initialises beta_h and beta_J to 0 and then
    assigns both on either branch of u >= -40 (dead stores);
  * the second term of LR91's beta_K1, exp(0.06175 * (u - E_K1 - 594.31)),
    is below 1e-12 of the sum it is added to at every potential from -110 to
    +45 mV, so a change to its constants changes no FP64 result.
Parameter and state slots (p[k], st[k]) are planted too, each moved to a
neighbouring slot.  A parameter swap is invisible, and allowed, when the two
slots hold the same default value (Barkley b = eap = 0.02, Bueno-Orovio
theta_v_m = theta_o = 0.006) or when the reference kernel never reads the
parameter (UNUSED_PARAMS).
The mistakes live only in the temporary copies; the oracle is untouched.
"""
from __future__ import annotations

import ctypes
import random
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

from tests import test_oracle_known_answers as KT

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "fw_oracle.c"
SWEEPS = {"AlievPanfilov2D": "ionic_ap", "Barkley2D": "ionic_barkley",
          "MitchellSchaeffer2D": "ionic_ms", "FentonKarma2D": "ionic_fk",
          "BuenoOrovio2D": "ionic_bo", "LuoRudy912D": "ionic_lr91", "TP062D": "ionic_tp06",
          "Courtemanche2D": "ionic_crn"}
# parameters in the reference kernel's signature that its body never uses:
# LR91's cao (luo_rudy91_2d.py:510 signature, unused); TP06's Bufc and Kbufc
# feed only calc_cai, whose result the kernel discards (tp06_2d.py:1098);
# Courtemanche's calc_ikr overrides gkr (courtemanche_2d.py:401) and Vc enters
# only through Vj, Vup and Vrel
UNUSED_PARAMS = {"LuoRudy912D": {"cao"}, "TP062D": {"Bufc", "Kbufc"},
                 "Courtemanche2D": {"gkr", "Vc"}}

# (sweep, code line as written, literal) -> planted mistakes no output shows
UNDETECTABLE = {
    ("LuoRudy912D", "double alpha_h = 0, beta_h = 0, beta_J = 0, alpha_J = 0;", "0"): 2,
    ("LuoRudy912D", "exp(0.06175 * (uc - E_K1 - 594.31))) /", "0.06175"): 1,
    ("LuoRudy912D", "exp(0.06175 * (uc - E_K1 - 594.31))) /", "594.31"): 1,
}
NUM = re.compile(r"(?<![\w.])(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+|\d+)"
                 r"(?![\w.])")


def _code_spans(src: str, start: int, end: int):
    """[a, b) ranges of src[start:end] outside comments."""
    spans, i = [], start
    for m in re.finditer(r"/\*.*?\*/|//[^\n]*", src[start:end], re.S):
        spans.append((i, start + m.start()))
        i = start + m.end()
    spans.append((i, end))
    return spans


def _span(src: str, fn: str):
    start = src.index(f"static void {fn}(")
    nxt = re.search(r"\n(static |int |void )", src[start + 1:])
    return start, (start + 1 + nxt.start() if nxt else len(src))


def _literals(src: str, fn: str):
    """(position, text) of numeric literals in the body of static function
    `fn` up to the next top-level function, outside comments and brackets."""
    start, end = _span(src, fn)
    out = []
    for a, b in _code_spans(src, start, end):
        for m in NUM.finditer(src, a, b):
            before = src[:m.start()].rstrip()
            if before.endswith("["):  # p[3], st[0]: slots, handled by _slots
                continue
            out.append((m.start(), m.group(0)))
    return out


def _slots(src: str, fn: str, n_par: int, n_state: int):
    """(position, text, replacement) of every parameter and state slot
    p[k] / st[k] in the sweep, each moved to a neighbouring slot that
    exists (k + 1, or k - 1 for the last)."""
    start, end = _span(src, fn)
    out = []
    for m in re.finditer(r"\b(p|st)\[(\d+)\]", src[start:end]):
        k = int(m.group(2))
        n = n_par if m.group(1) == "p" else n_state
        if n < 2:  # a single slot has no neighbour to confuse it with
            continue
        j = k + 1 if k + 1 < n else k - 1
        out.append((start + m.start(2), m.group(2), str(j), m.group(1), k))
    return out


def _mutate(text: str) -> str:
    if re.fullmatch(r"\d+", text):
        return str(int(text) + 1)
    x = float(text)
    return repr(x * 1.01) if x != 0.0 else "0.0625"


def _cases():
    """(sweep, position, text, replacement): every literal and every slot."""
    src = SRC.read_text()
    out = []
    for cls, fn in SWEEPS.items():
        t = KT.TABLES[cls]
        out += [(cls, pos, text, _mutate(text)) for pos, text in _literals(src, fn)]
        for pos, text, rep, kind, k in _slots(src, fn, len(t["scalars"]), len(t["state_args"])):
            if kind == "p":
                j = int(rep)
                if (t["scalar_args"][k] in UNUSED_PARAMS.get(cls, ()) or
                        t["scalars"][k] == t["scalars"][j]):
                    continue  # invisible by construction (module docstring)
            out.append((cls, pos, text, rep))
    return out


CASES = _cases()


def _build(tmp: Path, k: int, pos: int, text: str, rep: str) -> Path:
    src = SRC.read_text()
    mutated = src[:pos] + rep + src[pos + len(text):]
    d = tmp / f"m{k}"
    d.mkdir()
    (d / "fw_oracle.c").write_text(mutated)
    shutil.copy(SRC.with_suffix(".h"), d / "fw_oracle.h")
    out = d / f"libmut{k}.so"
    r = subprocess.run(["gcc", "-O0", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                        "-std=c11", "-shared", "-o", str(out), str(d / "fw_oracle.c"), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _load(path: Path):
    L = ctypes.CDLL(str(path))
    D = ctypes.POINTER(ctypes.c_double)
    L.orc_ionic.argtypes = [ctypes.c_int, D, D, ctypes.POINTER(D),
                            ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_double, D]
    L.orc_ionic.restype = ctypes.c_int
    return L


def test_every_sweep_has_constants_and_slots_to_plant_mistakes_in():
    counts = {cls: sum(1 for c, *_ in CASES if c == cls) for cls in SWEEPS}
    # every parameter and state slot of each model is read by its sweep
    for cls, fn in SWEEPS.items():
        t = KT.TABLES[cls]
        used = {(m.group(1), int(m.group(2))) for m in
                re.finditer(r"\b(p|st)\[(\d+)\]", SRC.read_text()[slice(*_span(SRC.read_text(), fn))])}
        assert {k for w, k in used if w == "p"} == set(range(len(t["scalars"]))), cls
        assert {k for w, k in used if w == "st"} == set(range(len(t["state_args"]))), cls
    assert all(n >= 3 for n in counts.values()), counts
    assert len(CASES) > 700


def test_planted_constant_typos_fail_the_known_answer_pins(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is needed to build the planted copies")
    src = SRC.read_text()
    lines = src.splitlines()
    with ThreadPoolExecutor(max_workers=8) as ex:
        libs = list(ex.map(lambda a: _build(tmp_path, *a),
                           [(k, pos, text, rep) for k, (_, pos, text, rep) in enumerate(CASES)]))
    missed = {}
    for (cls, pos, text, rep), path in zip(CASES, libs):
        L = _load(path)
        if cls in KT.PHYS:
            dev = KT.phys_deviation(cls, L)
        else:
            dev = KT.light_deviation(cls, L) / KT.LIGHT_TOL
        if not dev > 1.0:
            line = src[:pos].count("\n")
            key = (cls, lines[line].strip(), text)
            missed.setdefault(key, []).append(f"fw_oracle.c:{line + 1} {text} -> {rep}")
    unexpected = {k: v for k, v in missed.items() if len(v) > UNDETECTABLE.get(k, 0)}
    assert not unexpected, unexpected
