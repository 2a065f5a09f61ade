"""Reader of the full-size golden fixtures (written by make_fullsize.py) and
the comparison of a nodal solution with one of their cases."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def path(workload):
    return os.path.join(HERE, f"fullsize_{workload}.npz")


def load(workload):
    """The fixture of a workload, or None when it has not been generated."""
    p = path(workload)
    if not os.path.exists(p):
        return None
    with np.load(p, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def compare(fx, u_nodal, case=0):
    """Differences of the nodal solution u (n_nodes x ndof) from fixture case:
      sample_max    max over the sampled nodes of |u - u_ref|, over max |u_ref|
      sample_l2     ||u - u_ref|| over the sampled nodes, over ||u_ref|| there
      norm2         | ||u|| - ||u_ref|| | / ||u_ref|| over all DOFs
      checksum      | w.u - w.u_ref | / (||w|| ||u_ref||) over all DOFs (seeded w)
    """
    u = np.asarray(u_nodal, dtype=np.float64)
    us = u[fx["sample_nodes"]]
    ref = fx["u_sample"][case]
    w = np.random.default_rng(int(fx["check_seed"])).uniform(-1.0, 1.0, size=u.size)
    n2 = float(fx["norm2"][case])
    return dict(sample_max=float(np.abs(us - ref).max() / fx["maxabs"][case]),
                sample_l2=float(np.linalg.norm(us - ref) / np.linalg.norm(ref)),
                norm2=float(abs(np.linalg.norm(u) - n2) / n2),
                checksum=float(abs(np.dot(w, u.ravel()) - fx["checksum"][case]) / (np.linalg.norm(w) * n2)))
