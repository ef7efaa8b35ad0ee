"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/ifdda.h; the Fortran drop-in exports the reference's
subroutine names.  No compute calls (no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "if-dda_amd", "libifdda_hip.so")
FLIB = os.path.join(ROOT, "if-dda_amd", "libifdda_dropin.so")


def declared():
    src = open(os.path.join(ROOT, "include", "ifdda.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ifdda_[a-z0-9_]+)\s*\(", src)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_header_declares_the_api():
    names = declared()
    for n in ("ifdda_create", "ifdda_green_build", "ifdda_matvec", "ifdda_matvec_device",
              "ifdda_solve", "ifdda_solve_device", "ifdda_produitfftmatvect3",
              "ifdda_produitfftmatvectopt", "ifdda_greencalculfft", "ifdda_greencalculquadfft",
              "ifdda_inverserig", "ifdda_inverserigopt"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build() first"
    C.CDLL(LIB)  # loads without a GPU
    missing = [n for n in declared() if n not in exported(LIB)]
    assert not missing, missing


def test_fortran_dropin_exports_reference_names():
    assert os.path.exists(FLIB)
    C.CDLL(FLIB)
    syms = exported(FLIB)
    for n in ("produitfftmatvect3_", "produitfftmatvectopt_", "greencalculfft_",
              "greencalculquadfft_", "inverserig_", "inverserigopt_", "diffractefft2d_",
              "computegcfft2d_", "calculforcefft_", "calculforcefftopt_", "derivativefield2_"):
        assert n in syms


def test_python_binding_loads():
    import ifdda
    L = ifdda.lib()
    assert L.ifdda_last_error() is not None


def test_no_oracle_in_product():
    """The product never links or imports the oracle."""
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in out and "cdmref" not in out
    pkg = os.path.join(ROOT, "if-dda_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".f90", ".cpp")):
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "import oracle" not in txt and "from oracle" not in txt, f
                assert "liboracle" not in txt and "libcdmref" not in txt, f


@pytest.mark.parametrize("n,L", [(1, 2), (4, 8), (7, 15), (10, 20), (13, 27), (20, 40),
                                 (37, 75), (48, 96), (64, 128), (100, 200)])
def test_padded_lengths_rule(n, L):
    """Smallest 2,3,5-smooth length >= 2n (restated from the plan rule)."""
    def smooth(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1
    m = 2 * n
    while not smooth(m):
        m += 1
    assert m == L
