"""Every entry point that include/g12kba.h declares is described in the
integration guide (a new export without documentation fails here)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_exported_symbol_is_documented():
    header = open(os.path.join(ROOT, "include", "g12kba.h")).read()
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    names = sorted(set(re.findall(r"\b(g12_[a-z0-9_]+)\(", header)))
    assert len(names) > 60
    missing = [n for n in names if n not in guide]
    assert not missing, missing
