"""SCF iterations of the 64^3 x 200 diblock workload of bench.py (seeded cylinder fields,
Anderson m = 20) without graph replays or sweep timing, for kernel traces and counter
profiles:  ADSCF_GRAPH=0 python -m adscf_amd.profile_metric 32|64 [iterations]
The first 21 iterations mix without the Anderson solve; later ones run every kernel of the
steady-state iteration."""
import sys

import adscf_amd
from adscf_amd import fields

N, L, NS, FA, CHIN, KAPPA, M, LAM = 64, 4.5, 200, 0.3, 20.0, 25.0, 20, 0.5


def main(prec_arg: str, iters: int) -> None:
    w0 = fields.cylinder_seed(N, L, CHIN, sigma=0.2, seed=2007)
    segs = fields.block_segments([FA, 1 - FA], NS)
    prec = adscf_amd.FP32 if prec_arg == "32" else adscf_amd.FP64
    with adscf_amd.Context(N, N, N, L, L, L, 1.0 / NS, precision=prec) as ctx:
        ctx.set_interactions(KAPPA, [[0.0, CHIN], [CHIN, 0.0]], [1.0, 1.0])
        ctx.scf_setup([1, 2], segs, ensemble=1, rho=1.0, mixing_dim=M, lam_mix=LAM)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(iters, tol=-1.0)
        print("done", st, flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "32", int(sys.argv[2]) if len(sys.argv) > 2 else 30)
