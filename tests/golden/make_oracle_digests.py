"""Writes tests/golden/oracle_digests.json: the oracle's output digests of
every parity case and of a short MCMC run, which
tests/test_oracle_regression.py compares against.

Run from the repository root after building the oracle (make -C oracle):
    python -m tests.golden.make_oracle_digests
"""
import json
import os

from tests.cases import CASES
from tests.oracle_lib import load_oracle
from tests.test_oracle_regression import GOLDEN, case_digests, kept_output_digests, mcmc_digests


def main():
    lib = load_oracle()
    out = {"cases": {name: case_digests(lib, name) for name in sorted(CASES)},
           "kept_outputs": kept_output_digests(lib), "mcmc": mcmc_digests(lib)}
    with open(GOLDEN, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {os.path.relpath(GOLDEN)}: {len(out['cases'])} cases")


if __name__ == "__main__":
    main()
