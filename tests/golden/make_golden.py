"""Writes the golden fixtures of this directory from the oracle (the CPU restatement of the reference's
assembly, FEKernels.h:317-508). Run from the repository root after building:

    python tests/golden/make_golden.py

Each fixture holds the generator parameters, the generated mesh and displacements, the BSR pattern and the
FP64 residual and BSR values, all in full. The 64 x 64 order-1 case is the reference's CPU-runnable configuration
(unit square, u,v ~ U(+-1e-3) seed 1, E = 70e9, nu = 0.3, t = 1; 148 996 BSR values).

    python tests/golden/make_golden.py [name ...]     # only the named fixtures
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("gpu-fem-project_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

E, NU, T = 70e9, 0.3, 1.0

# name: (kind, order, size parameters, jitter, mesh seed, state seed)
CASES = {
    "square_p1_4x3": ("square", 1, (4, 3), 0.2, 11, 12),
    "square_p2_3x3": ("square", 2, (3, 3), 0.2, 21, 22),
    "annulus_p3_2x8": ("annulus", 3, (2, 8), 0.1, 31, 32),
    "square_p4_2x2": ("square", 4, (2, 2), 0.2, 41, 42),
    "annulus_p4_2x6": ("annulus", 4, (2, 6), 0.1, 51, 52),
    "c1_square_p1_64x64": ("square", 1, (64, 64), 0.0, 0, 1),
}


def build_mesh(F, kind, order, size, jitter, seed):
    if kind == "square":
        return F.square_mesh(size[0], order, jitter=jitter, seed=seed, ny=size[1])
    return F.annulus_mesh(size[0], size[1], order, jitter=jitter, seed=seed)


def main():
    import fem_amd as F
    import oracle as O

    F.load()
    O.load()
    names = sys.argv[1:] or list(CASES)
    for name in names:
        kind, order, size, jitter, mseed, useed = CASES[name]
        m = build_mesh(F, kind, order, size, jitter, mseed)
        u = F.fill_uniform(2 * m.num_nodes, -1e-3, 1e-3, useed)
        rowp, cols = O.bsr_pattern(m)
        bmap, misses = O.element_bcsr_map(m, rowp, cols)
        assert misses == 0
        res, vals = O.jacobian(m, u, E, NU, T, bmap, int(rowp[-1]))
        out = dict(kind=np.array(kind), order=np.array(order), size=np.array(size), jitter=np.array(jitter),
                   mesh_seed=np.array(mseed), state_seed=np.array(useed), material=np.array([E, NU, T]),
                   num_nodes=np.array(m.num_nodes), num_elements=np.array(m.num_elements),
                   residual=res, rowp=rowp, coords=m.coords, conn=m.conn, states=u, cols=cols, values=vals)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {m.num_elements} elements, {m.num_nodes} nodes, {int(rowp[-1])} blocks -> "
              f"{os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
