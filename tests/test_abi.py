"""The C-ABI libraries load and export every symbol include/*.h declares (CPU only; no
compute calls).  Also pins the reference's symbol set: every extern "C" name of the
reference's 3D and 2D shims and spacetime.cc (listed here from the reference sources,
src/prop3DAco*DenQ_DEO2_FDTD.cc, src/prop2DAco*DenQ_DEO2_FDTD.cc and src/spacetime.cc:228-330)
is exported."""
import ctypes
import os
import re
import subprocess

import pytest

from wavefd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

HEADER_OF = {
    "iso": "prop3DAcoIsoDenQ_DEO2_FDTD.h",
    "vti": "prop3DAcoVTIDenQ_DEO2_FDTD.h",
    "tti": "prop3DAcoTTIDenQ_DEO2_FDTD.h",
    "spacetime": "spacetime.h",
    "codec": "wavefd_codec.h",
    "ginsu": "wavefd_ginsu.h",
    "lcodec": "wavefd_lcodec.h",
    "iso2d": "prop2DAcoIsoDenQ_DEO2_FDTD.h",
    "vti2d": "prop2DAcoVTIDenQ_DEO2_FDTD.h",
    "tti2d": "prop2DAcoTTIDenQ_DEO2_FDTD.h",
}

REFERENCE_SYMBOLS = {
    "iso": ["alloc", "free", "SetupDtOmegaInvQ", "TimeStep", "ScaleSpatialDerivatives",
            "ForwardBornInjection_V", "ForwardBornInjection_VB", "ForwardBornInjection_B",
            "AdjointBornAccumulation_V", "AdjointBornAccumulation_VB", "AdjointBornAccumulation_B",
            "AdjointBornAccumulation_wavefieldsep", "getNx", "getNy", "getNz", "getV", "getB",
            "getDtOmegaInvQ", "getPSpace", "getPCur", "getPOld"],
    "vti": ["alloc", "free", "SetupDtOmegaInvQ", "TimeStepLinear", "TimeStep", "ScaleSpatialDerivatives",
            "ForwardBornInjection_V", "ForwardBornInjection_VEA", "AdjointBornAccumulation_V",
            "AdjointBornAccumulation_wavefieldsep_V", "AdjointBornAccumulation_VEA", "getNx", "getNy", "getNz",
            "getV", "getEps", "getEta", "getB", "getF", "getPSpace", "getMSpace", "getPCur", "getPOld",
            "getMCur", "getMOld"],
    "tti": ["alloc", "free", "SetupDtOmegaInvQ", "TimeStep", "ScaleSpatialDerivatives",
            "ForwardBornInjection_V", "ForwardBornInjection_VEA", "AdjointBornAccumulation_V",
            "AdjointBornAccumulation_wavefieldsep_V", "AdjointBornAccumulation_VEA", "getNx", "getNy", "getNz",
            "getV", "getEps", "getEta", "getSinTheta", "getCosTheta", "getSinPhi", "getCosPhi", "getB", "getF",
            "getDtOmegaInvQ", "getPSpace", "getMSpace", "getPCur", "getPOld", "getMCur", "getMOld"],
    "iso2d": ["alloc", "free", "SetupDtOmegaInvQ", "TimeStep", "ScaleSpatialDerivatives",
              "ForwardBornInjection_V", "ForwardBornInjection_VB", "ForwardBornInjection_B",
              "AdjointBornAccumulation_V", "AdjointBornAccumulation_VB", "AdjointBornAccumulation_B",
              "AdjointBornAccumulation_wavefieldsep", "getNx", "getNz", "getV", "getB", "getDtOmegaInvQ",
              "getPSpace", "getPCur", "getPOld"],
    "vti2d": ["alloc", "free", "SetupDtOmegaInvQ", "TimeStepLinear", "TimeStep", "ScaleSpatialDerivatives",
              "ForwardBornInjection_V", "ForwardBornInjection_VEA", "AdjointBornAccumulation_V",
              "AdjointBornAccumulation_wavefieldsep_V", "AdjointBornAccumulation_VEA", "getNx", "getNz", "getV",
              "getEps", "getEta", "getB", "getF", "getDtOmegaInvQ", "getPSpace", "getMSpace", "getPCur",
              "getPOld", "getMCur", "getMOld", "applyFirstDerivatives2D_PlusHalf"],
    "tti2d": ["alloc", "free", "SetupDtOmegaInvQ", "TimeStep", "ScaleSpatialDerivatives",
              "ForwardBornInjection_V", "ForwardBornInjection_VEA", "AdjointBornAccumulation_V",
              "AdjointBornAccumulation_wavefieldsep_V", "AdjointBornAccumulation_VEA", "getNx", "getNz", "getV",
              "getEps", "getEta", "getSinTheta", "getCosTheta", "getB", "getF", "getDtOmegaInvQ", "getPSpace",
              "getMSpace", "getPCur", "getPOld", "getMCur", "getMOld"],
    "spacetime": ["injectdata_float", "injectdata_double", "extractdata_float", "extractdata_double",
                  "interpadjoint_1d_float", "interpadjoint_1d_double", "interpforward_1d_float",
                  "interpforward_1d_double", "interpadjoint_nd_float", "interpadjoint_nd_double",
                  "interpforward_nd_float", "interpforward_nd_double"],
}


def declared(kind):
    text = open(os.path.join(INC, HEADER_OF[kind])).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text))
    names -= {"extern", "if", "defined"}
    if kind in _lib.PREFIXES and not kind.endswith("2d"):
        add = open(os.path.join(INC, "wavefd_amd.h")).read()
        add = re.sub(r"/\*.*?\*/", "", add, flags=re.S)
        for nm in re.findall(r"P##_(\w+)\s*\(", add):
            names.add(_lib.PREFIXES[kind] + "_" + nm)
        for nm in re.findall(r"\b(" + _lib.PREFIXES[kind] + r"_\w+)\s*\(", add):
            names.add(nm)
    return names


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


ALL = ["iso", "vti", "tti", "spacetime", "codec", "ginsu", "lcodec", "iso2d", "vti2d", "tti2d"]


@pytest.mark.parametrize("kind", ALL)
def test_library_exports_every_declared_symbol(kind):
    path = _lib.lib_path(kind)
    assert os.path.exists(path), "build first"
    names = declared(kind)
    assert len(names) >= 4
    missing = names - exported(path)
    assert not missing, missing


@pytest.mark.parametrize("kind", ["iso", "vti", "tti", "spacetime", "iso2d", "vti2d", "tti2d"])
def test_reference_symbol_set_is_covered(kind):
    pre = "" if kind == "spacetime" else _lib.PREFIXES[kind] + "_"
    names = {pre + s for s in REFERENCE_SYMBOLS[kind]}
    assert names <= declared(kind)
    assert names <= exported(_lib.lib_path(kind))


@pytest.mark.parametrize("kind", ALL)
def test_ctypes_table_matches_headers(kind):
    table = set(_lib.signatures(kind))
    assert table == declared(kind), (table ^ declared(kind))


@pytest.mark.parametrize("kind", ALL)
def test_library_loads_without_gpu(kind):
    lib = _lib.load(kind)
    assert isinstance(lib.dll, ctypes.CDLL)


def test_product_does_not_reference_oracle():
    """Nothing under wavefd_amd/ or include/ may import, link or load the oracle."""
    for top in ("wavefd_amd", "include"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, top)):
            for f in files:
                if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp")):
                    text = open(os.path.join(dirpath, f), errors="replace").read()
                    assert "oracle" not in text.lower(), os.path.join(dirpath, f)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti", "iso2d", "vti2d", "tti2d"])
def test_setup_q_rejects_small_freq_without_exiting(kind):
    """freqQ < FLT_EPSILON: the reference prints its message and calls exit(EXIT_FAILURE)
    (src/absorb.h:34-39, :98-103).  The drop-in prints the same message, records it for
    LastError and returns, leaving the handle alone (it is never dereferenced here: a null
    handle proves that, and no GPU is needed)."""
    lib = _lib.load(kind)
    lib.ClearError()
    lib.SetupDtOmegaInvQ(None, 0.0, 0.1, 100.0)
    msg = _lib.last_error(kind)
    assert "freqQ [0.000000] is too small" in msg
    lib.ClearError()
    assert _lib.last_error(kind) == ""
