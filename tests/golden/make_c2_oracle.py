"""Write tests/golden/c2_oracle_4000.npz: the NumPy FP64 oracle over the whole
collisional nonlinear EPW run (nx = 64 on [0, 2 pi/0.3), nv = 1024 on
[-6.4, 6.4], f0 = (1 + 0.05 cos kx) exp(-v^2/2)/sqrt(2 pi), e0 = 0, Dougherty
nu = 1e-3 every step, dt = 0.05, 4000 leapfrog steps).

Stored (all float64): mean_e2 [4000] (mean over x of E^2 after each step),
e_snap [40][64] (E after steps 100, 200, ..., 4000), f_rows [4][1024] (f at
x rows 0, 16, 32, 48 after the last step), n_final [64] (trapz density after
the last step) and n_initial [64].  The oracle needs about a minute here;
the GPU tests compare against these values instead of re-running it.

    python tests/golden/make_c2_oracle.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import vpfp_oracle as O  # noqa: E402

K, NX, NV, VMAX, AMP, NU, DT, STEPS = 0.3, 64, 1024, 6.4, 0.05, 1e-3, 0.05, 4000
ROWS = (0, 16, 32, 48)


def main():
    g = O.Grid.make(NX, NV, 2 * np.pi / K, VMAX)
    f0 = O.perturbed_maxwellian(g.x, g.v, K, AMP)
    _, f, eh = O.run_steps(g, np.zeros(NX), f0, STEPS, DT, nu=NU, fp_type="dg")
    out = os.path.join(HERE, "c2_oracle_4000.npz")
    np.savez_compressed(out, mean_e2=np.mean(eh ** 2, axis=1), e_snap=eh[99::100].copy(),
                        f_rows=f[list(ROWS)].copy(), n_final=O.compute_charges(f, g.dv),
                        n_initial=O.compute_charges(f0, g.dv))
    print("wrote", out)


if __name__ == "__main__":
    main()
