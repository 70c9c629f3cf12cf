"""ORACLE — test infrastructure only.

Runs the reference program (oracle/_ref/mersenne2_ref, built unmodified from
its own source by `make -C oracle ref`) with the read-back interposer
(oracle/_ref/libm2_refdump.so) preloaded, and turns the digit state it reads
back after its last iteration into the canonical residue, Res64 and SHA-256,
the same quantities this project's fixtures hold. The reference prints none of
them (mersenne2_gpu.cpp:4855-4882).

    python oracle/run_reference.py 19937 prp            # a whole test
    python oracle/run_reference.py 27899687 prp -N 2000  # 2000 iterations

Prints one JSON line: p, mode, iterations, res64, sha256, the reference's
own stdout lines (its verdict and it/s) and its wall time. M2_REF_CPU=1 is
passed through to the interposer (GPU enumeration mapped to CPU devices).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "_ref", "mersenne2_ref")
DUMP_LIB = os.path.join(HERE, "_ref", "libm2_refdump.so")


def transformsize(p: int) -> int:
    """mersenne2_gpu.cpp:286-293."""
    ln = 3
    while ln + 2 * ((p >> ln) + 1) >= 92:
        ln += 1
    return ln


def bit_positions(p: int, n: int):
    return [-(-p * j // n) for j in range(n + 1)]


def residue_from_digits(p: int, digits) -> int:
    """R = (sum d_j 2^ceil(p j / n)) mod 2^p - 1, by halving the digit range
    (O(M(p) log n) instead of n big additions)."""
    n = len(digits)
    B = bit_positions(p, n)

    def comb(lo, hi):
        if hi - lo <= 16:
            # digits may exceed their width: add each at its position
            v = 0
            for j in range(lo, hi):
                v += digits[j] << (B[j] - B[lo])
            return v
        mid = (lo + hi) // 2
        return comb(lo, mid) + (comb(mid, hi) << (B[mid] - B[lo]))

    m = (1 << p) - 1
    r = comb(0, n)
    while r >> p:
        r = (r & m) + (r >> p)
    return 0 if r == m else r


def digits_from_dump(raw: bytes, h: int):
    import numpy as np
    if len(raw) == 24 * h:
        a = np.frombuffer(raw, dtype=np.uint64).reshape(h, 3)
        return [int(x) for x in a[:, :2].reshape(-1)]
    if len(raw) == 16 * h:
        a = np.frombuffer(raw, dtype=np.uint64)
        return [int(x) for x in a]
    raise ValueError("read-back of %d bytes does not match h = %d elements" % (len(raw), h))


def run(p: int, mode: str, iters: int = 0, extra=(), timeout: float = 3600.0):
    if not (os.path.exists(REF) and os.path.exists(DUMP_LIB)):
        raise FileNotFoundError("build the reference and the interposer first: make -C oracle ref")
    ln = transformsize(p)
    h = 1 << (ln - 1)
    with tempfile.TemporaryDirectory() as td:
        dump = os.path.join(td, "z.bin")
        env = dict(os.environ, LD_PRELOAD=DUMP_LIB, M2_REF_DUMP=dump)
        cmd = [REF, str(p), mode] + (["-N", str(iters)] if iters else []) + list(extra)
        t0 = time.perf_counter()
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env, timeout=timeout)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("reference exited %d: %s %s" % (r.returncode, r.stdout[-500:], r.stderr[-500:]))
        with open(dump, "rb") as f:
            raw = f.read()
    R = residue_from_digits(p, digits_from_dump(raw, h))
    rb = R.to_bytes((p + 7) // 8, "little")
    lines = r.stdout.splitlines()
    return {"p": p, "mode": mode, "iterations": iters or (p if mode == "prp" else p - 2), "ln": ln,
            "res64": "%016X" % (R & ((1 << 64) - 1)), "sha256": hashlib.sha256(rb).hexdigest(),
            "verdict_line": next((l for l in lines if l.startswith("2^")), None),
            "device_line": next((l for l in lines if l.startswith("device=")), None),
            "wall_s": round(wall, 3)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("p", type=int)
    ap.add_argument("mode", choices=["ll", "prp"])
    ap.add_argument("-N", type=int, default=0, help="iterations (0: the whole test)")
    ap.add_argument("extra", nargs="*", help="further reference flags")
    a = ap.parse_args(argv)
    print(json.dumps(run(a.p, a.mode, a.N, a.extra)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
