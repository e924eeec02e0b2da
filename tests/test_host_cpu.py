"""Host-side checks that need no GPU: the C ABI exports, the positional
rhic-conf readers against the reference's own configuration files, and the
bench's refusal of more than one GPU."""
import ctypes as C
import os
import re

import pytest

import gpuvh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONF = os.path.join(ROOT, "tests", "golden", "rhic-conf")


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(gvh_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 15
    lib = C.CDLL(gpuvh.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_read_reference_lattice_properties():
    L = gpuvh.read_lattice_parameters(CONF)
    assert (L.numLatticePointsX, L.numLatticePointsY, L.numLatticePointsRapidity) == (101, 101, 1)
    assert (L.numComputationalLatticePointsX, L.numComputationalLatticePointsY,
            L.numComputationalLatticePointsRapidity) == (105, 105, 5)
    assert L.numProperTimePoints == 10000
    assert (L.latticeSpacingX, L.latticeSpacingY, L.latticeSpacingRapidity,
            L.latticeSpacingProperTime) == (0.15, 0.15, 0.1, 0.05)


def test_read_reference_hydro_properties():
    """Line 5 is read as the regulation temperature although the shipped file
    names it initializePiNavierStokes (HydroParameters.cpp:83-87)."""
    H = gpuvh.read_hydro_parameters(CONF)
    assert (H.initialProperTimePoint, H.shearViscosityToEntropyDensity,
            H.freezeoutTemperatureGeV, H.initializePimunuNavierStokes,
            H.regulationTemperatureGeV) == (0.5, 0.2, 0.155, 0, 0.0)


def test_read_reference_ic_properties_past_end():
    """The shipped ic.properties has nine lines and no final newline; the
    tenth read fails without clearing the string, so sourceType repeats the
    last value: atoi("5.9") = 5."""
    I = gpuvh.read_ic_parameters(CONF)
    assert I.initialConditionType == 0
    assert I.numberOfSourceFiles == 80
    assert I.numberOfNucleonsPerNuclei == 20
    assert I.initialEnergyDensity == 0.5
    assert I.scatteringCrossSectionNN == 62.0
    assert I.impactParameter == 0.0
    assert I.fractionOfBinaryCollisions == 1.0
    assert I.rapidityVariance == 0.16
    assert I.rapidityMean == 5.9
    assert I.sourceType == 5


def test_positional_reader_rules(tmp_path):
    """Keys are ignored, whitespace anywhere is stripped, a line without '='
    is taken whole; with a final newline a missing line reads as empty (0)."""
    (tmp_path / "lattice.properties").write_text(
        "foo = 7\nbar=9\n 3 \nx=12\n  a = 0 . 2 5\nb=0.5\nc=0.75\n")
    L = gpuvh.read_lattice_parameters(str(tmp_path))
    assert (L.numLatticePointsX, L.numLatticePointsY, L.numLatticePointsRapidity) == (7, 9, 3)
    assert L.numProperTimePoints == 12
    assert (L.latticeSpacingX, L.latticeSpacingY, L.latticeSpacingRapidity) == (0.25, 0.5, 0.75)
    assert L.latticeSpacingProperTime == 0.0


def test_missing_config_is_an_error(tmp_path):
    with pytest.raises(gpuvh.GvhError, match="no configuration file"):
        gpuvh.read_hydro_parameters(str(tmp_path))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_more_than_one_gpu(n):
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says
    so, before touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and ("--gpus %d" % n) in r.stderr
    assert r.stdout == ""
