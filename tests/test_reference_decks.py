"""The regression cases stored as data reproduce hippoLBM's own regression decks.

When the reference tree is present (it is not on GPU hosts), every
``regression-test/*.msp`` is read unmodified and must configure exactly the run
its rendered counterpart configures: the same domain, parameters, operator
lists, obstacles, initial distribution and checker pins.  Both front ends are
checked: the Python deck reader and the C++ operator layer's reader
(``hlbm_run --dump-deck``, compared after the deck reader normalises both).
Skipped when the reference tree is absent.
"""
import dataclasses
import glob
import json
import os
import subprocess

import pytest

import render_decks
from conftest import DECKS, ROOT
from hippolbm_amd import deck as dk

REF_DIR = "/root/reference/regression-test"
REF_DECKS = sorted(glob.glob(os.path.join(REF_DIR, "*.msp")))
RUN = os.path.join(ROOT, "hippolbm_amd", "hlbm_run")

pytestmark = pytest.mark.skipif(not REF_DECKS, reason="reference regression decks not present")


def test_every_reference_case_is_stored():
    names = {os.path.splitext(os.path.basename(p))[0] for p in REF_DECKS}
    assert names == set(render_decks.cases())


@pytest.mark.parametrize("path", REF_DECKS, ids=[os.path.basename(p) for p in REF_DECKS])
def test_rendered_case_configures_the_reference_run(path):
    name = os.path.basename(path)
    ref = dataclasses.asdict(dk.load_msp(path))
    ours = dataclasses.asdict(dk.load_msp(os.path.join(DECKS, name)))
    assert ours == ref
    # the checker pins are the reference's numbers
    assert ours["checker"] == ref["checker"] and ours["end_iteration"] == ref["end_iteration"]


@pytest.mark.parametrize("path", REF_DECKS, ids=[os.path.basename(p) for p in REF_DECKS])
def test_cpp_reader_sees_the_same_run(path):
    if not os.path.exists(RUN):
        pytest.skip("hlbm_run not built")

    def tree(p):
        out = subprocess.run([RUN, p, "--dump-deck"], capture_output=True, text=True, check=True).stdout
        return json.loads(out)

    a, b = tree(path), tree(os.path.join(DECKS, os.path.basename(path)))
    # the C++ reader keeps scalars as text: bring both through the deck model
    da = dataclasses.asdict(dk.from_dict(_typed(a)))
    db = dataclasses.asdict(dk.from_dict(_typed(b)))
    da["name"] = db["name"] = ""
    assert da == db


def _typed(v):
    if isinstance(v, dict):
        return {k: _typed(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_typed(x) for x in v]
    if isinstance(v, str):
        if v in ("true", "false"):
            return v == "true"
        return v
    return v
