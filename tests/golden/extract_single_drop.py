"""Writes tests/golden/v1_single_drop.json from the reference's recorded single-drop results
(results/v1_single_drop.json: a sphere driven into a stiff ground contact, stepped by SolverMuJoCo
with the euler integrator; gold run at dt 2e-5 and five fixed steps).  Only the numbers are kept.

    python tests/golden/extract_single_drop.py /root/reference
"""
import json
import os
import sys


def main(ref_root: str) -> None:
    d = json.load(open(os.path.join(ref_root, "results", "v1_single_drop.json")))
    out = {"source": "results/v1_single_drop.json of the reference (values copied as data; produced by "
                     "scripts/rl/adaptive_expts/v1_single_drop.py with SolverMuJoCo, integrator euler)",
           # the scene of scripts/rl/adaptive_expts/v1_single_drop.py:28-50
           "scene": {"radius": 0.05, "z0": d["meta"]["Z0"], "vz": d["meta"]["VZ"], "ke": d["meta"]["ke"], "kd": 0.0,
                     "mu": 0.0, "margin": 0.005, "dt_outer": 0.01, "duration": 0.4},
           "gold_dt": d["meta"]["gold_dt"], "gold_zmin": d["meta"]["gold_zmin"], "fixed": d["fixed"]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "v1_single_drop.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
