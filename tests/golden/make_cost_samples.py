"""Regenerate the committed cost model and its sample fixture (needs a GPU).

    python -m kclistpp_amd.train_cost_model --out kclistpp_amd/cost_model.kclcm \
        --samples /tmp/samples.npz
    python tests/golden/make_cost_samples.py /tmp/samples.npz

The fixture keeps 6000 rows (seeded choice) of the k >= 5 training samples:
features and labels (log2(1 + build cycles), log2(1 + counting cycles))
measured on an MI355X by the enumeration kernel's own per-task timers.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(path: str) -> None:
    d = np.load(path)
    x, y = d["x"], d["y"]
    rng = np.random.default_rng(0)
    idx = np.nonzero(x[:, 7] >= 5)[0]
    sel = np.sort(rng.choice(idx, 6000, replace=False))
    np.savez_compressed(os.path.join(HERE, "cost_samples_k5to7.npz"), x=x[sel], y=y[sel])


if __name__ == "__main__":
    main(sys.argv[1])
