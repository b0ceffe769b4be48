"""Writes configs/cfg4_replicates.json: 64 independent replicate configs.

Master seed 2024; replicate r draws N = round(exp(U(ln 2e4, ln 5e5))), grid
g = round(sqrt(N / 0.25)) (density 0.25), regular rate U[0.01, 0.05], high rate
U[0.1, 0.4], death rate U[0.05, 0.3], seed 1000 + r, 2,400 hours; everything else as the
10,000-agent reference-style config (configs/cfg1_10k.json).
"""
import copy
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    base = json.load(open(os.path.join(HERE, "configs", "cfg1_10k.json")))
    rng = np.random.default_rng(2024)
    out = []
    for r in range(64):
        c = copy.deepcopy(base)
        n = int(round(math.exp(rng.uniform(math.log(2e4), math.log(5e5)))))
        c["population"]["Auto"]["number_of_agents"] = n
        c["geography_parameters"]["grid_size"] = int(round(math.sqrt(n / 0.25)))
        c["disease"]["regular_transmission_rate"] = float(rng.uniform(0.01, 0.05))
        c["disease"]["high_transmission_rate"] = float(rng.uniform(0.1, 0.4))
        c["disease"]["death_rate"] = float(rng.uniform(0.05, 0.3))
        c["hours"] = 2400
        c["seed"] = 1000 + r
        out.append(c)
    with open(os.path.join(HERE, "configs", "cfg4_replicates.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("total agents", sum(c["population"]["Auto"]["number_of_agents"] for c in out))


if __name__ == "__main__":
    main()
