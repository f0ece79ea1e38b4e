"""Writes c0_eigenvalues.json: the 34 lowest eigenvalues of the 2-atom bcc
system (32^3 grid, FD8 + V_loc + KB nonlocal, Gamma point) from ARPACK
(scipy.sparse.linalg.eigsh, implicitly restarted Lanczos) over the oracle's
H.psi.  ARPACK is independent of both LOBPCG implementations, so these values
pin their converged eigenvalues.

    python tests/golden/make_c0_eigenvalues.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import conftest  # noqa: E402
import oracle  # noqa: E402
from scipy.sparse.linalg import LinearOperator, eigsh  # noqa: E402

qm = conftest._load_pkg()
s = qm.system.config(0)
n = s.ng


def mm(V):
    V = np.asarray(V)
    if V.ndim == 1:
        return oracle.hpsi(s, V[None, :])[0]
    return oracle.hpsi(s, np.ascontiguousarray(V.T)).T


A = LinearOperator((n, n), matvec=mm, matmat=mm, dtype=np.float64)
w = np.sort(eigsh(A, k=34, which="SA", tol=1e-14, ncv=100)[0])
with open(os.path.join(HERE, "c0_eigenvalues.json"), "w") as f:
    json.dump({"system": s.name, "method": "scipy eigsh (ARPACK), tol 1e-14, ncv 100",
               "eigenvalues_Ha": [float(x) for x in w]}, f, indent=1)
print(w)
