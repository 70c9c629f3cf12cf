"""Exact residues of the Lucas-Lehmer and base-3 PRP iterations mod 2^p - 1.

Independent of the transform: the state after k iterations is
  PRP: 3^(2^k) mod M_p (from 3, x -> x^2)
  LL : s_k with s_0 = 4, s_{k+1} = s_k^2 - 2 mod M_p
which is what the reference's exact debug path computes (mersenne_step_ref,
mersenne2_gpu.cpp:3424-3432, starting values :3908). Large exponents use GMP
through ctypes (libgmp.so.10); small ones use Python integers.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import hashlib
from typing import Dict, Iterable, List

_GMP = None


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


class _Funcs:
    """Attribute access to libgmp's mpz_* entry points without the __gmpz_ prefix."""

    def __init__(self, lib):
        self._lib = lib

    def __getattr__(self, name):
        return getattr(self._lib, "__gmpz_" + name)


def _gmp():
    global _GMP
    if _GMP is None:
        name = ctypes.util.find_library("gmp") or "libgmp.so.10"
        g = ctypes.CDLL(name)
        P = ctypes.POINTER(_Mpz)
        for fn, args in {
            "__gmpz_init": [P], "__gmpz_clear": [P], "__gmpz_set_ui": [P, ctypes.c_ulong],
            "__gmpz_mul": [P, P, P], "__gmpz_add": [P, P, P], "__gmpz_sub": [P, P, P],
            "__gmpz_sub_ui": [P, P, ctypes.c_ulong], "__gmpz_add_ui": [P, P, ctypes.c_ulong],
            "__gmpz_fdiv_q_2exp": [P, P, ctypes.c_ulong], "__gmpz_fdiv_r_2exp": [P, P, ctypes.c_ulong],
            "__gmpz_cmp": [P, P], "__gmpz_cmp_ui": [P, ctypes.c_ulong], "__gmpz_setbit": [P, ctypes.c_ulong],
            "__gmpz_sizeinbase": [P, ctypes.c_int],
            "__gmpz_export": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_size_t,
                              ctypes.c_int, ctypes.c_size_t, P],
        }.items():
            getattr(g, fn).argtypes = args
        getattr(g, "__gmpz_cmp").restype = ctypes.c_int
        getattr(g, "__gmpz_cmp_ui").restype = ctypes.c_int
        getattr(g, "__gmpz_sizeinbase").restype = ctypes.c_size_t
        _GMP = _Funcs(g)
    return _GMP


class MersenneGMP:
    """x mod 2^p - 1 held in a GMP integer, squared in place."""

    def __init__(self, p: int, x0: int):
        g = _gmp()
        self.p = p
        self.x, self.t, self.m = _Mpz(), _Mpz(), _Mpz()
        for v in (self.x, self.t, self.m):
            g.init(ctypes.byref(v))
        g.set_ui(ctypes.byref(self.m), 0)
        g.setbit(ctypes.byref(self.m), p)
        g.sub_ui(ctypes.byref(self.m), ctypes.byref(self.m), 1)
        g.set_ui(ctypes.byref(self.x), x0)

    def __del__(self):
        g = _GMP
        if g is not None:
            for v in (self.x, self.t, self.m):
                g.clear(ctypes.byref(v))

    def step(self, sub: int) -> None:
        g, x, t, m, p = _gmp(), ctypes.byref(self.x), ctypes.byref(self.t), ctypes.byref(self.m), self.p
        g.mul(x, x, x)
        g.fdiv_q_2exp(t, x, p)
        g.fdiv_r_2exp(x, x, p)
        g.add(x, x, t)
        if g.cmp(x, m) >= 0:
            g.sub(x, x, m)
        if sub:
            if g.cmp_ui(x, sub) < 0:
                g.add(x, x, m)
            g.sub_ui(x, x, sub)

    def residue_bytes(self) -> bytes:
        g = _gmp()
        nbytes = (self.p + 7) // 8
        buf = (ctypes.c_ubyte * (nbytes + 8))()
        cnt = ctypes.c_size_t(0)
        g.export(buf, ctypes.byref(cnt), -1, 1, 0, 0, ctypes.byref(self.x))
        return bytes(buf[:nbytes])


class MersennePy:
    def __init__(self, p: int, x0: int):
        self.p, self.m, self.x = p, (1 << p) - 1, x0

    def step(self, sub: int) -> None:
        self.x = (self.x * self.x - sub) % self.m

    def residue_bytes(self) -> bytes:
        return self.x.to_bytes((self.p + 7) // 8, "little")


def run(p: int, mode: str, checkpoints: Iterable[int], keep_bytes: Iterable[int] = ()) -> Dict[int, dict]:
    """Res64 (and SHA-256 of the canonical residue) after each checkpoint k.

    mode "prp": x0 = 3, x -> x^2;  mode "ll": x0 = 4, x -> x^2 - 2.
    """
    ks: List[int] = sorted(set(int(k) for k in checkpoints))
    keep = set(int(k) for k in keep_bytes)
    x0, sub = (3, 0) if mode == "prp" else (4, 2)
    eng = MersenneGMP(p, x0) if p > 4000 else MersennePy(p, x0)
    out: Dict[int, dict] = {}
    done = 0
    for k in ks:
        for _ in range(k - done):
            eng.step(sub)
        done = k
        r = eng.residue_bytes()
        rec = {"res64": "%016X" % int.from_bytes(r[:8].ljust(8, b"\0"), "little"),
               "sha256": hashlib.sha256(r).hexdigest()}
        if k in keep:
            rec["residue_hex"] = r[::-1].hex().upper()
        out[k] = rec
    return out
