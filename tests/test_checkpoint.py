"""Checkpoint files: host-side format checks (CPU) and save / resume through
the engine and the command line (GPU)."""
import os
import struct
import subprocess

import pytest

import mersenne2gpu_amd as m2
from conftest import load_fixture


def _fnv1a(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h ^= b
        h = (h * 1099511628211) & ((1 << 64) - 1)
    return h


def _write(path, p, mode, it, residue: bytes, res64=None, magic=b"M2CKPT01", corrupt=False):
    r64 = int.from_bytes(residue[:8].ljust(8, b"\0"), "little") if res64 is None else res64
    body = magic + struct.pack("<IIQQQ", p, mode, it, r64, len(residue)) + residue
    tail = struct.pack("<Q", _fnv1a(body))
    data = bytearray(body + tail)
    if corrupt:
        data[len(body) // 2] ^= 1
    path.write_bytes(bytes(data))


def test_checkpoint_info_reads_header(tmp_path):
    f = tmp_path / "a.m2ck"
    r = (123456789).to_bytes((127 + 7) // 8, "little")
    _write(f, 127, 1, 77, r)
    info = m2.checkpoint_info(str(f))
    assert info == {"p": 127, "mode": "prp", "iteration": 77, "res64": "00000000075BCD15"}


@pytest.mark.parametrize("kind", ["corrupt", "magic", "res64", "length", "missing"])
def test_checkpoint_info_refuses_bad_files(tmp_path, kind):
    f = tmp_path / "b.m2ck"
    r = (5).to_bytes(16, "little")
    if kind == "corrupt":
        _write(f, 127, 1, 3, r, corrupt=True)
    elif kind == "magic":
        _write(f, 127, 1, 3, r, magic=b"M2CKPT99")
    elif kind == "res64":
        _write(f, 127, 1, 3, r, res64=6)
    elif kind == "length":
        _write(f, 127, 1, 3, r + b"\0")
    with pytest.raises(m2.M2Error):
        m2.checkpoint_info(str(f))


@pytest.mark.gpu
def test_checkpoint_save_load_continues_exactly(gpu_ctx, tmp_path):
    c = [c for c in load_fixture("c1.json")["cases"] if c["p"] == 19937 and c["mode"] == "prp"][0]
    f = str(tmp_path / "t.m2ck")
    with m2.Test(gpu_ctx, 19937, "prp") as a:
        a.run(700)
        a.save_checkpoint(f)
    assert m2.checkpoint_info(f)["iteration"] == 700
    with m2.Test(gpu_ctx, 19937, "prp") as b:
        b.load_checkpoint(f)
        assert b.iteration == 700
        b.run(2000 - 700)
        assert b.res64() == c["checkpoints"]["2000"]["res64"]
    with m2.Test(gpu_ctx, 19937, "ll") as wrong:
        with pytest.raises(m2.M2Error):
            wrong.load_checkpoint(f)


@pytest.mark.gpu
def test_checkpoint_multipass_plan(gpu_ctx, tmp_path):
    c = load_fixture("c2.json")["cases"][0]
    f = str(tmp_path / "c.m2ck")
    with m2.Test(gpu_ctx, c["p"], "prp") as a:
        a.run(500)
        a.save_checkpoint(f)
    with m2.Test(gpu_ctx, c["p"], "prp") as b:
        b.load_checkpoint(f)
        b.run(500)
        assert b.res64() == c["checkpoints"]["1000"]["res64"]


@pytest.mark.gpu
def test_cli_checkpoint_and_resume(tmp_path):
    f = tmp_path / "cli.m2ck"
    r = subprocess.run([m2.CLI_PATH, "19937", "prp", "500", "-N", "1500", "--checkpoint", str(f)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert m2.checkpoint_info(str(f))["iteration"] == 1500
    r = subprocess.run([m2.CLI_PATH, "19937", "prp", "500", "--resume", str(f), "--res64-at", "2000"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "resumed from %s at iteration 1500" % f in r.stdout
    assert "Res64@2000: B5DBC55987E2FD16" in r.stdout
    assert r.stdout.splitlines()[-1] == "Res64: 0000000000000009"
