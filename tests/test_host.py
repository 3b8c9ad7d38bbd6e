"""The C++ operator layer's deck reader (no GPU): hlbm_run --dump-deck renders
every regression deck exactly as a YAML loader reads it, and a deck naming an
operator the layer does not provide, or missing a required slot, is refused
before any device work."""
import glob
import json
import os
import subprocess

import pytest
import yaml

from conftest import DECKS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "hippolbm_amd", "hlbm_run")
DECK_FILES = sorted(glob.glob(os.path.join(DECKS, "*.msp")))


def norm(v):
    if isinstance(v, dict):
        return {str(k): norm(x) for k, x in v.items()}
    if isinstance(v, list):
        return [norm(x) for x in v]
    if v is None:
        return None
    if isinstance(v, bool):
        return str(v).lower()
    return v


def same(a, b):
    """a: the reader's JSON (scalars as text); b: PyYAML's typed values."""
    if isinstance(b, dict):
        return isinstance(a, dict) and list(a) == list(b) and all(same(a[k], b[k]) for k in b)
    if isinstance(b, list):
        return isinstance(a, list) and len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    if b is None:
        return a is None
    if isinstance(b, (int, float)):
        return float(a) == float(b)
    return a.strip() == str(b).strip()


@pytest.mark.parametrize("path", DECK_FILES, ids=[os.path.basename(p) for p in DECK_FILES])
def test_reader_matches_yaml(path):
    out = subprocess.run([RUN, path, "--dump-deck"], capture_output=True, text=True, check=True).stdout
    assert same(json.loads(out), norm(yaml.safe_load(open(path))))


def test_unknown_operator_is_refused(tmp_path):
    deck = tmp_path / "bad.msp"
    deck.write_text(open(DECK_FILES[0]).read() + "\nboundary_conditions:\n  - write_vtk_slices\n")
    r = subprocess.run([RUN, str(deck)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "unsupported operator 'write_vtk_slices'" in r.stdout


def test_required_analysis_slot_is_checked_before_device_work(tmp_path):
    # plane_velocity_profile's dimension slot is REQUIRED (plane_velocity_profile.cu:160-161)
    deck = tmp_path / "bad.msp"
    deck.write_text(open(DECK_FILES[0]).read() + "\nanalysis:\n  - plane_velocity_profile:\n     dump_file: p.csv\n")
    r = subprocess.run([RUN, str(deck)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "the dimension slot is required" in r.stdout
    assert not (tmp_path / "hippoLBMOutputDir").exists()
