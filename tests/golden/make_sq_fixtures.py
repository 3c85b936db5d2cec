"""Writes the full-size FP64 oracle fixtures of one SQ density step
(scf_iter 1, the smooth model Veff of sqdft_amd.configs) for the 32-atom Al
cell (32^3 nodes, nloc 17, npl 48), the 216-atom LiH cell (48^3 nodes,
nloc 15, npl 64) and the 256-atom Al cell (64^3 nodes, nloc 17, npl 48):
rho, lambda_f, Ebs, Eent, Lanczos counts -- for al256 rho and the counts of
every 8th node only (the global values still cover every node).

Run from the repository root: python tests/golden/make_sq_fixtures.py [al32|lih216|al256 ...]
(al32 takes ~7 min, lih216 ~32 min, al256 ~55 min on 8 cores)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]
import helpers as H  # noqa: E402
from sqdft_amd import configs  # noqa: E402


def main(names):
    for name in names:
        cfg = configs.CONFIGS[name]
        S = configs.build(name, seed=0)
        veff = configs.smooth_veff(S)
        osys = H.oracle_for(S, cfg["Rcut"], cfg["T"])
        t0 = time.time()
        out = osys.density(veff, scf_iter=1, nthreads=0)
        assert out["status"] >= 0, out["status"]
        path = os.path.join(ROOT, "tests", "golden", f"sq_{name}.npz")
        stride = 8 if name == "al256" else 1
        np.savez_compressed(path, rho=out["rho"][::stride], lambda_f=out["lambda_f"], ebs=out["ebs"],
                            eent=out["eent"], iters=out["iters"][::stride].astype(np.int16),
                            fermi_iters=out["fermi_iters"], stride=stride)
        print(name, f"{time.time() - t0:.0f} s", "lambda_f", out["lambda_f"], "->", path, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["al32", "lih216"])
