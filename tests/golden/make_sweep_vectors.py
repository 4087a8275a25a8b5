"""Write tests/golden/sweep_vectors.json from the reference's own code, built from
source into oracle/_ref/libref_ssp.so (oracle/Makefile, target ref):

* logarithm10_time_bins(t_max, dt) (Percolation_FD/stochasticSPDP.cpp:336-377) for
  a set of (t_max, dt);
* var_mean_incremental_surv_runs (Percolation_FD/stochasticSPDP.cpp:173-207)
  folding seeded replicate series one after another into a running table.

    python tests/golden/make_sweep_vectors.py

The vectors are data: inputs and the reference's outputs.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

TIME_BIN_CASES = [(1e4, 0.1), (1e5, 0.005), (2000.0, 0.2), (1.5, 0.1), (50.0, 1.0), (1e3, 0.05),
                  (91201.0, 0.1), (0.5, 0.1), (3e4, 0.02)]


def replicate_series(seed, n_rep, n_meas):
    """n_rep replicates x n_meas rows of (#active, mean): alive up to a random
    measurement, zero after it, as a replicate whose time loop ended there."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n_rep, n_meas, 2))
    for k in range(n_rep):
        end = int(rng.integers(0, n_meas + 4))
        for m in range(min(end, n_meas)):
            x[k, m, 0] = float(rng.integers(1, 4096))
            x[k, m, 1] = float(rng.random() * 0.5)
    return x


def first_table(x0):
    """The first replicate's table (rietkerk_bjork_FKE.cpp:4581-4596 /
    stochasticSPDP.cpp:1068-1076): mean, 0, #surviving in {0, 1}, #active."""
    tab = np.zeros((x0.shape[0], 5))
    tab[:, 1] = x0[:, 1]
    tab[:, 3] = (x0[:, 0] > 0).astype(float)
    tab[:, 4] = x0[:, 0]
    return tab


def main():
    from oracle import oracle as o
    if o.build_ref() is None:
        raise SystemExit("the reference sources are not present: cannot build oracle/_ref")
    bins = [{"t_max": t, "dt": dt, "t": o.ref_log10_time_bins(t, dt).tolist()} for t, dt in TIME_BIN_CASES]
    runs = []
    for seed, n_rep, n_meas in [(11, 9, 24), (12, 16, 40), (13, 3, 7)]:
        x = replicate_series(seed, n_rep, n_meas)
        tab = first_table(x[0])
        for k in range(1, n_rep):
            o.ref_var_mean_surv(tab, x[k])
        runs.append({"seed": seed, "series": x.tolist(), "table": tab[:, 1:].tolist()})
    out = {"source": "Percolation_FD/stochasticSPDP.cpp built from source (oracle/Makefile ref)",
           "time_bins": bins, "replicate_stats": runs}
    with open(os.path.join(HERE, "sweep_vectors.json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")
    print("wrote", os.path.join(HERE, "sweep_vectors.json"))


if __name__ == "__main__":
    main()
