"""Writes tests/golden/underflow_rate_case.json: one particle and one alignment column
on which the lowest gamma rate category's partials underflow.

The state is particle 770 of the 200-taxon x 10,000-site CAT benchmark population
(smccat.simulate.prior_particles with seed 1 over 4,096 particles, K ~ U{5..120})
and column 0 of that benchmark's simulated alignment (simulate_config("c4")), with
the class that column is assigned to.  Its gamma shape is 0.0576, so the lowest of
the four rates is 8e-15: P is the identity up to round-off there and the
variable column's likelihood under that rate is below the range of a double.
The file is data written by this package's own simulator; nothing of the
reference is run.

Usage:  python tests/golden/make_underflow_case.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "smc-cat_amd"))

from smccat import simulate as sim  # noqa: E402

PARTICLE, SITE = 770, 0


def main():
    codes, cfg = sim.simulate_config("c4")
    T, S = codes.shape
    batch = sim.prior_particles(np.random.default_rng(1), cfg["particles"], T, S, num_rates=cfg["rates"],
                                invar=cfg["invar"], k_range=cfg["k_range"])
    i, s = PARTICLE, SITE
    k = int(batch.assignment[i, s])
    out = {
        "taxa": int(T),
        "num_rates": int(cfg["rates"]),
        "column": codes[:, s].astype(int).tolist(),
        "children": batch.children[i].astype(int).tolist(),
        "branch_len": batch.branch_len[i].tolist(),
        "base": batch.base[i].tolist(),
        "logits": batch.logits[i, k].tolist(),
        "pinv": float(batch.pinv[i]),
        "shape": float(batch.shape[i]),
    }
    with open(os.path.join(HERE, "underflow_rate_case.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
