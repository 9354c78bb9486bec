"""CPU-side checks of the C ABI: the library loads, exports what include/femgpu.h declares, and
its host mesh services agree with the oracle.  No GPU compute is called here."""
import subprocess

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A


def test_library_exports_every_declared_symbol():
    A.load()
    out = subprocess.check_output(["nm", "-D", "--defined-only", A.LIB_PATH]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    declared = A.declared_symbols()
    assert len(declared) >= 30
    assert [d for d in declared if d not in exported] == []
    assert sorted(e for e in exported if e.startswith("femgpu") and e not in declared) == []


def test_status_strings():
    lib = A.load()
    assert lib.femgpu_status_string(A.NOT_SPD) == b"matrix not SPD"
    assert lib.femgpu_status_string(99) == b"unknown status"


@pytest.mark.parametrize("kind,nb", [(0, (4, 3, 5)), (1, (4, 3, 2)), (2, (5, 4, 3)), (1, (6, 5, 4))])
def test_make_rect_and_skin_match_oracle(oracle, kind, nb):
    c, e = L.make_rect(kind, nb)
    c2, e2 = oracle.make_rect(kind, nb)
    assert np.array_equal(c, c2) and np.array_equal(e, e2)
    assert np.array_equal(L.skin_faces(kind, len(c), e), oracle.skin_faces(kind, len(c), e))


def test_make_rect_rejects_bad_sizes():
    with pytest.raises(L.FemGpuError):
        L.make_rect(L.HEXA, (1, 4, 4))
