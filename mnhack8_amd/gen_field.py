#!/usr/bin/env python3
"""gen_field.py -- per-element material fields for the heat solver.

Same command line and outputs as the reference's scripts/gen_field.py
(scripts/gen_field.py:63-118): --n --model --out [--k --k2 --period --rhoc
--mu --sigma --rhoc-sigma --L --upscale --seed], writing <out>_k.bin,
<out>_rhoc.bin (float64, e = ix*n*n + iy*n + iz) and <out>_meta.json.
Two models are added for the synthetic workloads: iid_lognormal (--clip) and
block_lognormal (--block).
"""
import argparse
import os
import sys

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from mnhack8_amd import fields  # noqa: E402

MODELS = ["homogeneous", "quadrants", "lognormal", "layered", "iid_lognormal", "block_lognormal"]


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--n", type=int, required=True)
    p.add_argument("--model", choices=MODELS, required=True)
    p.add_argument("--out", required=True)
    p.add_argument("--k", type=float, default=1.0)
    p.add_argument("--k2", type=float, default=100.0, help="layered: second conductivity")
    p.add_argument("--period", type=int, default=4, help="layered: slab thickness in elements")
    p.add_argument("--rhoc", type=float, default=1.0)
    p.add_argument("--mu", type=float, default=0.0, help="lognormal: mean of ln k")
    p.add_argument("--sigma", type=float, default=1.0, help="lognormal: std of ln k")
    p.add_argument("--rhoc-sigma", type=float, default=0.0, help="lognormal: std of ln rhoc (same field)")
    p.add_argument("--L", type=float, default=0.125, help="lognormal: correlation length (domain = 1)")
    p.add_argument("--upscale", type=int, default=1, help="lognormal: generate at (r*n)^3, upscale r^3 blocks")
    p.add_argument("--clip", type=float, default=3.0, help="iid_lognormal: clip the normal deviate to +-clip")
    p.add_argument("--block", type=int, default=4, help="block_lognormal: block edge in elements")
    p.add_argument("--seed", type=int, default=1)
    a = p.parse_args(argv)
    n = a.n
    meta = vars(a).copy()
    if a.model == "homogeneous":
        k, rhoc = fields.homogeneous(n, a.k, a.rhoc)
    elif a.model == "quadrants":
        if n % 2:
            print(f"WARNING: n={n} is odd; the quadrants model cannot match the prototype "
                  f"(interfaces at 0.5 fall inside elements). Fine as a negative control.", file=sys.stderr)
        k, rhoc = fields.quadrants(n)
    elif a.model == "layered":
        k, rhoc = fields.layered(n, a.k, a.k2, a.period, a.rhoc)
    elif a.model == "lognormal":
        k, rhoc = fields.lognormal(n, a.mu, a.sigma, a.rhoc, a.rhoc_sigma, a.L, a.upscale, a.seed)
        if a.upscale > 1:
            meta["upscale_rule"] = {"k": "harmonic", "rhoc": "arithmetic"}
        meta["L_over_h"] = a.L * n
    elif a.model == "iid_lognormal":
        k, rhoc = fields.iid_lognormal(n, a.mu, a.sigma, a.clip, a.rhoc, a.seed)
    else:
        k, rhoc = fields.block_lognormal(n, a.mu, a.sigma, a.block, a.rhoc, a.seed)
    m = fields.write_field(a.out, k, rhoc, meta)
    print(f"{a.model}: n={n}  k in [{m['kmin']:.4g}, {m['kmax']:.4g}] (contrast {m['contrast']:.3g})  "
          f"rhoc in [{m['rhocmin']:.4g}, {m['rhocmax']:.4g}]  -> {a.out}_{{k,rhoc}}.bin")


if __name__ == "__main__":
    main()
