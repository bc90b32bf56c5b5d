"""The oracle reproduces the committed C1 digests (tests/golden/c1_digests.json,
written by tests/golden/make_c1_digests.py): any change to the CPU
restatement's C1 tree or traversals shows up here."""
import importlib.util
import json
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


def load_maker():
    spec = importlib.util.spec_from_file_location("make_c1_digests", GOLDEN / "make_c1_digests.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_oracle_matches_committed_c1_digests():
    want = json.loads((GOLDEN / "c1_digests.json").read_text())
    got = load_maker().oracle_outputs()
    assert got == want
