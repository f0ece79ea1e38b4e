"""Writes tests/golden/opea_small12_kpoint.bin: the inputs of one small complex
k-point H.psi (12^3 grid, 2 atoms with l <= 2 projectors, 2 SoA complex bands)
and the oracle's FP64 result, for the C++ consumer test
(tests/consumer/opea_consumer.cpp), which binds the C ABI the way a QUMASUN
host binds its LOBPCG operator slot.

Layout (little endian): b"QMOPEA01"; int32 nx, ny, nz, nb, natoms, ik;
float64 hx, hy, hz; float64 k[3]; int32 nproj[natoms]; int64 vox_off[natoms+1];
int32 vox_xyz[3 V]; int64 tab_off[natoms+1]; float64 table[T]; float64 energy[P];
int32 proj_off[natoms+1]; float64 vloc[Ng]; float64 psi[2 nb Ng]; float64 hpsi[2 nb Ng].

    python tests/golden/make_opea_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import conftest  # noqa: E402,F401  (registers qumasun_amd)
import oracle  # noqa: E402

import qumasun_amd as qm  # noqa: E402

OUT = os.path.join(HERE, "opea_small12_kpoint.bin")


def main():
    s = qm.system.small_system(n=12, L=4.0, natoms=2, rc=1.4, nbands=2, ncomp=2, k_red=(0.25, -0.5, 0.125), seed=11)
    psi = s.initial_bands()
    k = s.kpoints[0]
    want = oracle.hpsi(s, psi, k=k)
    with open(OUT, "wb") as f:
        f.write(b"QMOPEA01")
        np.array([s.nx, s.ny, s.nz, s.nbands, s.natoms, 0], np.int32).tofile(f)
        np.array(list(s.h) + list(k), np.float64).tofile(f)
        for arr, dt in ((s.nproj, np.int32), (s.vox_off, np.int64), (s.vox_xyz, np.int32), (s.tab_off, np.int64),
                        (s.table, np.float64), (s.energy, np.float64), (s.proj_off, np.int32),
                        (s.vloc, np.float64), (psi, np.float64), (want, np.float64)):
            np.ascontiguousarray(arr, dtype=dt).tofile(f)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
