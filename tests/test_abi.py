"""CPU checks of the drop-in boundary: libm2.so loads, exports every entry
point declared in include/mersenne2.h, its host-only calls work, and the
`mersenne2` command line keeps the reference's argv grammar and exit codes
(mersenne2_gpu.cpp:3618-3716)."""
import os
import re
import subprocess

import pytest

import mersenne2gpu_amd as m2
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mersenne2.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(m2_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(m2.SIGNATURES)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", m2.LIB_PATH], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (m2_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    m2.lib()  # loads with the ctypes signatures


def test_host_only_entry_points():
    assert m2.transform_log2(19937) == 9
    assert m2.transform_log2(27899687) == 20
    assert m2.transform_log2(449846671) == 24
    with pytest.raises(m2.M2Error):
        m2.transform_log2(2)
    f20 = m2.footprint(27899687)
    assert 24 * 2 ** 19 < f20 < 200 * 2 ** 20


def _cli(*args):
    return subprocess.run([m2.CLI_PATH] + list(args), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=60)


def test_cli_usage_and_errors():
    r = _cli()
    assert r.returncode == 1 and r.stderr.startswith("Usage: ")
    r = _cli("19937", "prp", "100", "extra")
    assert r.returncode == 1 and "Unknown argument: extra" in r.stderr
    for flag, msg in [("-d", "Missing device id after -d"), ("-N", "Missing max iterations after -N"),
                      ("-B", "Missing carry_pairs after -B"), ("-W", "Missing work-group size after -W")]:
        r = _cli("19937", flag)
        assert r.returncode == 1 and msg in r.stderr


@pytest.mark.skipif(m2.device_count() > 0, reason="needs a machine without a GPU")
def test_cli_without_gpu_fails_like_the_reference():
    r = _cli("19937", "prp")
    assert r.returncode == 1
    assert r.stdout.startswith("p=19937, ln=9, transform=256 (n=512)")
    assert "No GPU device found" in r.stderr
