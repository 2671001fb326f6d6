"""Two SCF iterations of the 128^3 x 400 ABA triblock configuration (box 6.0, segments
100/200/100, chiN 25, kappa 25, Anderson m = 20, lamellar seed), for counter profiles of
the line passes at that size:  python -m adscf_amd.profile_large 32|64"""
import sys

import adscf_amd
from adscf_amd import fields


def main(prec_arg: str) -> None:
    n, L, ns = 128, 6.0, 400
    segs = fields.block_segments([0.25, 0.5, 0.25], ns)
    w = fields.lamellar_seed(n, L, 25.0, sigma=0.1, seed=2007)
    prec = adscf_amd.FP32 if prec_arg == "32" else adscf_amd.FP64
    with adscf_amd.Context(n, n, n, L, L, L, 1.0 / ns, precision=prec) as ctx:
        ctx.set_interactions(25.0, [[0.0, 25.0], [25.0, 0.0]], [1.0, 1.0])
        ctx.scf_setup([1, 2, 1], segs, ensemble=1, rho=1.0, mixing_dim=20, lam_mix=0.5)
        ctx.scf_set_field(w)
        st = ctx.scf_iterate(2, tol=-1.0)
        print("done", st, flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "32")
