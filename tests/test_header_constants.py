"""The ctypes binding's constants are the C header's: every `PAM_*` enumerator
that include/pam.h gives an explicit value has the same value in
pseudoacousticmodeling_amd/_lib.py (options, storage kinds, physics, boundaries,
layouts, preprocessing modes).  CPU only: the header is read as text."""
import os
import re

from pseudoacousticmodeling_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))


def _header_enumerators():
    text = open(os.path.join(HERE, "..", "include", "pam.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    out = {}
    for body in re.findall(r"enum\s*\w*\s*\{(.*?)\}", text, flags=re.S):
        for name, value in re.findall(r"\b(PAM_[A-Z0-9_]+)\s*=\s*(-?\d+)", body):
            out[name] = int(value)
    return out


def test_enumerators_match_the_binding():
    enums = _header_enumerators()
    assert {"PAM_OPT_LANES", "PAM_OPT_SNAPSHOTS", "PAM_STORE_RANDOM_BOUNDARY"} <= set(enums)
    missing = [n for n in enums if n.startswith(("PAM_OPT_", "PAM_STORE_")) and not hasattr(_lib, n)]
    assert not missing, missing
    wrong = {n: (v, getattr(_lib, n)) for n, v in enums.items() if hasattr(_lib, n) and getattr(_lib, n) != v}
    assert not wrong, wrong
