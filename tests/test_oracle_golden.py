"""The oracle against the committed golden vectors of the reference
(tests/golden/ref_r5l26_moist_gaussian.json, written by
tests/golden/make_golden.py from the reference's own compiled code).

Runs without the reference sources: the grid and the initial state are
rebuilt and checked against the digests of what the reference was given,
then every operator output and the three-step composite are compared by
sha256 of their float64 bytes (bit for bit).
"""
import json
import os

import numpy as np
import pytest

import parity_cases as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_r5l26_moist_gaussian.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def case():
    grid = P.case_grid()
    return grid, P.case_state(grid)


def check(name, arr, want):
    got = P.digest(arr)
    if got["sha256"] != want["sha256"]:
        a = np.asarray(arr, dtype=np.float64).ravel()
        diffs = [(i, a[i].hex(), h) for i, h in want["sample"] if a[i].hex() != h]
        pytest.fail(f"{name}: digest differs ({len(diffs)} of {len(want['sample'])} samples differ: {diffs[:3]})")


def test_inputs_match_what_the_reference_was_given(golden, case):
    grid, state = case
    mine = P.input_digests(grid, state)
    assert set(mine) == set(golden["inputs"])
    for k, want in golden["inputs"].items():
        assert mine[k]["sha256"] == want["sha256"], k


def test_operators_match_golden(golden, case):
    grid, state = case
    out = P.oracle_operators(grid, state)
    assert set(out) == set(golden["operators"])
    for k, want in golden["operators"].items():
        check(k, out[k], want)


def test_composite_matches_golden(golden, case):
    grid, state = case
    out = P.oracle_composite(grid, state)
    assert set(out) == set(golden["composite"])
    for k, want in golden["composite"].items():
        check(k, out[k], want)
