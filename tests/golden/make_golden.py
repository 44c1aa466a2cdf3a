"""Writes tests/golden/small_cases.npz from the FP64 oracle (run on the CPU).

The case is a 6x5x4 box, FCD interaction with +-3 periodic images, the
wrapper's Si/Ci tables, density seed 21 (the same inputs as
tests/test_gpu_parity.py::test_against_committed_fixture).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import dda_oracle as O  # noqa: E402


def main():
    tables = O.sici_tables()
    shape = (6, 5, 4)
    rng = np.random.default_rng(21)
    n = int(np.prod(shape))
    diel = np.repeat(rng.uniform(0.0, 1.0, n), 3)
    a = O.wrapper_args(list(shape), 15.0, diel, [1.0 + 0j, 5.96 + 3.89j], [0, 0, 1], [1, 0, 0], 542.0,
                       symmetry_axes=[2.5, 2.0], symmetry_method="None", apc_max_m=3, apc_max_n=3,
                       apc_method_name="FCD", tables=tables)
    m = O.OracleDDAModel(*a)
    x = np.random.default_rng(99).standard_normal(3 * n) + 1j * np.random.default_rng(98).standard_normal(3 * n)
    y = m.apply(x)
    m.solveElectricField(np.zeros(3 * n, complex), 100000, 1e-12)
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), x=x, y_fcd=y, e_fcd=m.getElectricField(),
                        obj=m.calculateObjective())


if __name__ == "__main__":
    main()
