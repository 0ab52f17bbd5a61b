"""Writes the golden vectors of the reference's own compiled code.

Runs the cases of tests/parity_cases.py through the reference (oracle/_ref,
built from the sources under /root/reference by `make -C oracle ref`) and
stores digests -- sha256 of the float64 bytes plus a few exact samples -- of
every output, and of the grid and initial state both sides start from, in
tests/golden/ref_r5l26_moist_gaussian.json.  The reference itself never
travels: test_oracle_golden.py checks the oracle against these digests.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import conftest  # noqa: E402,F401  (puts the repo and oracle/ on sys.path)
import parity_cases as P  # noqa: E402
import refharness as R  # noqa: E402

OUT = os.path.join(HERE, "ref_r5l26_moist_gaussian.json")


def main():
    if not R.available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference sources are present")
    grid = P.case_grid()
    state = P.case_state(grid)
    sz = R.sizes()
    assert grid.dims.s == sz["s"] and grid.dims.v == sz["v"]
    doc = {
        "what": "outputs of the reference's compiled functions (oracle/_ref) on tests/parity_cases.py",
        "case": {"res_id": 5, "n_layers": 26, "moist": True, "orography": "gaussian", "n_lloyd": 2000,
                 "dt": P.DT, "n_steps": P.N_STEPS, "operator_seed": P.OPERATOR_SEED},
        "inputs": P.input_digests(grid, state),
        "operators": {k: P.digest(v) for k, v in R.reference_operators(grid, state).items()},
        "composite": {k: P.digest(v) for k, v in R.reference_composite(grid, state).items()},
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
