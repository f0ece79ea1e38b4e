"""The C++ consumer (tests/consumer/opea_consumer.cpp) on the GPU: a program
that includes include/qumasun_hip.h and binds the LOBPCG operator slot
OpeA(SoAComplex* Ax, const SoAComplex* x, int num) to qm_hpsi_apply_f64 as a
QUMASUN host would, checked inside the program against the oracle's FP64
H.psi stored in tests/golden/opea_small12_kpoint.bin (<= 1e-12 relative), on
the whole block and on the sub-block &x[1], plus the status-code error path."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "consumer", "opea_consumer")
FIXTURE = os.path.join(HERE, "golden", "opea_small12_kpoint.bin")


def test_consumer_fixture_present():
    """The fixture holds the header the consumer reads (CPU check)."""
    with open(FIXTURE, "rb") as f:
        assert f.read(8) == b"QMOPEA01"


@pytest.mark.gpu
def test_consumer_opea_binding():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run qumasun-workshop-template_amd/build.py")
    print("case C++ consumer OpeA binding, 12^3 complex k-point, 2 bands", flush=True)
    r = subprocess.run([BIN, FIXTURE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    assert r.stdout.count("max rel err") == 2
