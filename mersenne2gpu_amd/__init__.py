"""mersenne2 on MI355X: Python view of the libm2 C ABI (include/mersenne2.h).

The engine itself is C++/HIP (libm2.so, built in-tree by
mersenne2gpu_amd.build); this module only binds it with ctypes so that tests,
the benchmark and the serving loop can drive it. There is no fallback: if
libm2.so is missing, importing `lib()` raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, List, Optional, Sequence, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libm2.so")
CLI_PATH = os.path.join(HERE, "mersenne2")

M2_OK = 0
MODE_LL = 0
MODE_PRP = 1
ERRORS = {-1: "M2_E_INVAL", -2: "M2_E_DEVICE", -3: "M2_E_NOMEM", -4: "M2_E_HIP", -5: "M2_E_RANGE",
          -6: "M2_E_STATE"}

# Every entry point of include/mersenne2.h with its ctypes signature.
_c = ctypes
_P = _c.c_void_p
SIGNATURES: Dict[str, Tuple[object, List[object]]] = {
    "m2_last_error": (_c.c_char_p, []),
    "m2_device_count": (_c.c_int, [_c.POINTER(_c.c_int)]),
    "m2_open": (_c.c_int, [_c.c_int, _c.POINTER(_P)]),
    "m2_close": (_c.c_int, [_P]),
    "m2_device_name": (_c.c_int, [_P, _c.c_char_p, _c.c_size_t]),
    "m2_device_props": (_c.c_int, [_P, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint32),
                                   _c.POINTER(_c.c_uint32)]),
    "m2_transform_log2": (_c.c_int, [_c.c_uint64, _c.POINTER(_c.c_int)]),
    "m2_test_footprint": (_c.c_int, [_c.c_uint64, _c.POINTER(_c.c_uint64)]),
    "m2_test_create": (_c.c_int, [_P, _c.c_uint64, _c.c_int, _P, _c.POINTER(_P)]),
    "m2_test_destroy": (_c.c_int, [_P]),
    "m2_test_run": (_c.c_int, [_P, _c.c_uint64]),
    "m2_test_sync": (_c.c_int, [_P]),
    "m2_test_iteration": (_c.c_int, [_P, _c.POINTER(_c.c_uint64)]),
    "m2_test_total_iters": (_c.c_int, [_P, _c.POINTER(_c.c_uint64)]),
    "m2_test_res64": (_c.c_int, [_P, _c.POINTER(_c.c_uint64)]),
    "m2_test_residue": (_c.c_int, [_P, _c.c_char_p, _c.c_size_t]),
    "m2_test_set_residue": (_c.c_int, [_P, _c.c_char_p, _c.c_size_t, _c.c_uint64]),
    "m2_test_is_prime": (_c.c_int, [_P, _c.POINTER(_c.c_int)]),
    "m2_test_time": (_c.c_int, [_P, _c.c_uint64, _c.POINTER(_c.c_double)]),
    "m2_test_kernel_times": (_c.c_int, [_P, _c.c_uint64, _c.POINTER(_c.c_double), _c.c_int, _c.POINTER(_c.c_int),
                                        _c.c_char_p, _c.c_size_t]),
    "m2_test_kernel_profile": (_c.c_int, [_P, _c.c_uint64, _c.POINTER(_c.c_double), _c.POINTER(_c.c_uint64),
                                          _c.c_int, _c.POINTER(_c.c_int), _c.c_char_p, _c.c_size_t]),
    "m2_test_plan": (_c.c_int, [_P, _c.c_char_p, _c.c_size_t]),
    "m2_debug_fused_spin_limit": (_c.c_int, [_c.c_uint32, _c.POINTER(_c.c_uint32)]),
    "m2_debug_product_rate": (_c.c_int, [_P, _c.POINTER(_c.c_double)]),
    "m2_test_run_checked": (_c.c_int, [_P, _c.c_uint64, _c.c_uint32, _c.c_uint32, _c.POINTER(_c.c_uint64)]),
    "m2_debug_gerbicz_fault": (_c.c_int, [_c.c_uint64]),
    "m2_checkpoint_save": (_c.c_int, [_P, _c.c_char_p]),
    "m2_checkpoint_load": (_c.c_int, [_P, _c.c_char_p]),
    "m2_checkpoint_info": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_int),
                                      _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64)]),
    "m2_autotune": (_c.c_int, [_P, _c.c_uint64, _c.c_uint32, _P, _c.c_char_p, _c.c_size_t]),
    "m2_batch_create": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_int),
                                   _c.POINTER(_c.c_uint64), _P, _c.POINTER(_P)]),
    "m2_batch_destroy": (_c.c_int, [_P]),
    "m2_batch_run": (_c.c_int, [_P, _c.c_uint64]),
    "m2_batch_sync": (_c.c_int, [_P]),
    "m2_batch_iteration": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_c.c_uint64)]),
    "m2_batch_res64": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_c.c_uint64)]),
    "m2_batch_residue": (_c.c_int, [_P, _c.c_size_t, _c.c_char_p, _c.c_size_t]),
    "m2_batch_is_prime": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_c.c_int)]),
    "m2_server_create": (_c.c_int, [_P, _c.c_uint64, _c.c_uint32, _P, _c.POINTER(_P)]),
    "m2_server_destroy": (_c.c_int, [_P]),
    "m2_server_submit": (_c.c_int, [_P, _c.c_uint64, _c.c_int, _c.c_uint64, _c.POINTER(_c.c_uint64)]),
    "m2_server_step": (_c.c_int, [_P, _c.POINTER(_c.c_int)]),
    "m2_server_poll": (_c.c_int, [_P, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_int),
                                  _c.c_size_t, _c.POINTER(_c.c_size_t)]),
    "m2_server_stats": (_c.c_int, [_P, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64),
                                   _c.POINTER(_c.c_uint64)]),
    "m2_square_noalloc": (_c.c_int, [_P, _c.c_uint64, _c.c_int, _c.c_char_p, _c.c_size_t, _c.c_uint64,
                                     _c.c_char_p, _c.c_size_t]),
}

_LIB = None


class M2Error(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = ""
        if _LIB is not None:
            raw = _LIB.m2_last_error()
            msg = raw.decode() if raw else ""
        super().__init__("%s: %s (%s)" % (where, ERRORS.get(code, code), msg))
        self.code = code


def lib():
    """The loaded libm2.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libm2.so not built: run `python -m mersenne2gpu_amd.build` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _chk(rc: int, where: str) -> None:
    if rc != M2_OK:
        raise M2Error(rc, where)


def transform_log2(p: int) -> int:
    v = ctypes.c_int(0)
    _chk(lib().m2_transform_log2(p, ctypes.byref(v)), "m2_transform_log2")
    return v.value


def footprint(p: int) -> int:
    v = ctypes.c_uint64(0)
    _chk(lib().m2_test_footprint(p, ctypes.byref(v)), "m2_test_footprint")
    return v.value


def device_count() -> int:
    v = ctypes.c_int(0)
    _chk(lib().m2_device_count(ctypes.byref(v)), "m2_device_count")
    return v.value


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _chk(lib().m2_open(device, ctypes.byref(h)), "m2_open")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().m2_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _chk(lib().m2_device_name(self.h, buf, 256), "m2_device_name")
        return buf.value.decode()

    def props(self) -> Dict[str, int]:
        cu, mem, wg, lds = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        _chk(lib().m2_device_props(self.h, ctypes.byref(cu), ctypes.byref(mem), ctypes.byref(wg), ctypes.byref(lds)),
             "m2_device_props")
        return {"compute_units": cu.value, "global_mem": mem.value, "max_wg": wg.value, "lds_bytes": lds.value}

    def test(self, p: int, mode: str = "prp", graph_iters: int = 0) -> "Test":
        return Test(self, p, mode, graph_iters)

    def product_rate(self) -> float:
        """Measured v_mad_u64_u32 issue rate of the device, lane operations per
        second (m2_debug_product_rate)."""
        v = ctypes.c_double()
        _chk(lib().m2_debug_product_rate(self.h, ctypes.byref(v)), "m2_debug_product_rate")
        return v.value


def debug_gerbicz_fault(block: int) -> None:
    """Test hook: flip one residue bit after block `block` (1-based) of the
    next run_checked call, once (m2_debug_gerbicz_fault)."""
    _chk(lib().m2_debug_gerbicz_fault(int(block)), "m2_debug_gerbicz_fault")


def checkpoint_info(path: str) -> Dict[str, object]:
    """Header of a checkpoint file (m2_checkpoint_info; no device needed)."""
    p, mode, it, r = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint64()
    _chk(lib().m2_checkpoint_info(os.fsencode(path), ctypes.byref(p), ctypes.byref(mode), ctypes.byref(it),
                                  ctypes.byref(r)), "m2_checkpoint_info")
    return {"p": p.value, "mode": "prp" if mode.value == MODE_PRP else "ll", "iteration": it.value,
            "res64": "%016X" % r.value}


def autotune(ctx: "Context", p: int, ntests: int = 1) -> Tuple[Dict[str, int], str]:
    """Plan options the autotuner picks for p's transform length and a batch of
    ntests tests, and its per-candidate report (m2_autotune)."""
    o = _Opts()
    buf = ctypes.create_string_buffer(8192)
    _chk(lib().m2_autotune(ctx.h, int(p), int(ntests), ctypes.byref(o), buf, 8192), "m2_autotune")
    return o.as_dict(), buf.value.decode()


class _Opts(ctypes.Structure):
    _fields_ = [("wg", ctypes.c_uint32), ("carry_wg", ctypes.c_uint32), ("carry_pairs", ctypes.c_uint32),
                ("local_stage_cap", ctypes.c_uint32), ("graph_iters", ctypes.c_uint32),
                ("max_radix", ctypes.c_uint32), ("autotune", ctypes.c_uint32), ("no_fuse", ctypes.c_uint32)]

    def as_dict(self) -> Dict[str, int]:
        return {k: getattr(self, k) for k, _ in self._fields_}


def make_opts(**kw) -> "_Opts":
    o = _Opts()
    for k, v in kw.items():
        setattr(o, k, int(v or 0))
    return o


class Test:
    """One LL or PRP test resident on the device (mirror of m2_test_*)."""

    def __init__(self, ctx: Context, p: int, mode: str = "prp", graph_iters: int = 0, **opts):
        self.p = int(p)
        self.mode = mode
        m = _mode(mode)
        opts = make_opts(graph_iters=graph_iters, **opts)
        h = ctypes.c_void_p()
        _chk(lib().m2_test_create(ctx.h, self.p, m, ctypes.byref(opts), ctypes.byref(h)), "m2_test_create")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().m2_test_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, iters: int) -> None:
        _chk(lib().m2_test_run(self.h, int(iters)), "m2_test_run")

    def sync(self) -> None:
        _chk(lib().m2_test_sync(self.h), "m2_test_sync")

    @property
    def iteration(self) -> int:
        v = ctypes.c_uint64()
        _chk(lib().m2_test_iteration(self.h, ctypes.byref(v)), "m2_test_iteration")
        return v.value

    @property
    def total_iters(self) -> int:
        v = ctypes.c_uint64()
        _chk(lib().m2_test_total_iters(self.h, ctypes.byref(v)), "m2_test_total_iters")
        return v.value

    def res64(self) -> str:
        v = ctypes.c_uint64()
        _chk(lib().m2_test_res64(self.h, ctypes.byref(v)), "m2_test_res64")
        return "%016X" % v.value

    def residue(self) -> bytes:
        n = (self.p + 7) // 8
        buf = ctypes.create_string_buffer(n)
        _chk(lib().m2_test_residue(self.h, buf, n), "m2_test_residue")
        return buf.raw[:n]

    def set_residue(self, r: bytes, iteration: int) -> None:
        _chk(lib().m2_test_set_residue(self.h, r, len(r), int(iteration)), "m2_test_set_residue")

    def is_prime(self) -> bool:
        v = ctypes.c_int()
        _chk(lib().m2_test_is_prime(self.h, ctypes.byref(v)), "m2_test_is_prime")
        return bool(v.value)

    def save_checkpoint(self, path: str) -> None:
        _chk(lib().m2_checkpoint_save(self.h, os.fsencode(path)), "m2_checkpoint_save")

    def load_checkpoint(self, path: str) -> None:
        _chk(lib().m2_checkpoint_load(self.h, os.fsencode(path)), "m2_checkpoint_load")

    def run_checked(self, iters: int, block: int, blocks_per_check: int = 10) -> int:
        """Advance a PRP test by `iters` iterations under a Gerbicz check
        (m2_test_run_checked); returns the number of failed checks."""
        e = ctypes.c_uint64()
        _chk(lib().m2_test_run_checked(self.h, int(iters), int(block), int(blocks_per_check), ctypes.byref(e)),
             "m2_test_run_checked")
        return e.value

    def time(self, iters: int) -> float:
        """Device milliseconds for `iters` iterations (hipEvents)."""
        v = ctypes.c_double()
        _chk(lib().m2_test_time(self.h, int(iters), ctypes.byref(v)), "m2_test_time")
        return v.value

    def kernel_times(self, iters: int) -> Dict[str, float]:
        """Mean device milliseconds per launch of each kernel over `iters` iterations."""
        ms = (ctypes.c_double * 16)()
        cnt = ctypes.c_int()
        names = ctypes.create_string_buffer(1024)
        _chk(lib().m2_test_kernel_times(self.h, int(iters), ms, 16, ctypes.byref(cnt), names, 1024),
             "m2_test_kernel_times")
        keys = names.value.decode().split("\n")[: cnt.value]
        return {k: v for k, v in zip(keys, ms[: cnt.value])}

    def kernel_profile(self, iters: int) -> Dict[str, tuple]:
        """(total device milliseconds, launches) of each kernel over `iters` iterations."""
        ms = (ctypes.c_double * 16)()
        nl = (ctypes.c_uint64 * 16)()
        cnt = ctypes.c_int()
        names = ctypes.create_string_buffer(1024)
        _chk(lib().m2_test_kernel_profile(self.h, int(iters), ms, nl, 16, ctypes.byref(cnt), names, 1024),
             "m2_test_kernel_profile")
        keys = names.value.decode().split("\n")[: cnt.value]
        return {k: (ms[i], nl[i]) for i, k in enumerate(keys)}

    def plan(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        _chk(lib().m2_test_plan(self.h, buf, 4096), "m2_test_plan")
        return buf.value.decode()


def _mode(m) -> int:
    if m in ("prp", MODE_PRP):
        return MODE_PRP
    if m in ("ll", MODE_LL):
        return MODE_LL
    raise ValueError("mode must be 'll' or 'prp': %r" % (m,))


class Batch:
    """Many resident tests advanced together (mirror of m2_batch_*).

    iters[i] (optional) is the iteration test i stops at; 0 or None means its
    full test (p PRP iterations, p - 2 LL steps)."""

    def __init__(self, ctx: Context, ps: Sequence[int], modes, iters: Optional[Sequence[int]] = None,
                 local_stage_cap: int = 0, **opts):
        n = len(ps)
        if isinstance(modes, str):
            modes = [modes] * n
        self.ps = [int(p) for p in ps]
        self.modes = list(modes)
        P = (ctypes.c_uint64 * max(n, 1))(*self.ps)
        M = (ctypes.c_int * max(n, 1))(*[_mode(m) for m in self.modes])
        It = None
        if iters is not None:
            It = (ctypes.c_uint64 * max(n, 1))(*[int(k or 0) for k in iters])
        opts = make_opts(local_stage_cap=local_stage_cap, **opts)
        h = ctypes.c_void_p()
        _chk(lib().m2_batch_create(ctx.h, n, P, M, It, ctypes.byref(opts), ctypes.byref(h)), "m2_batch_create")
        self.h = h

    def __len__(self) -> int:
        return len(self.ps)

    def close(self) -> None:
        if self.h:
            lib().m2_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_iters: int = 0) -> None:
        _chk(lib().m2_batch_run(self.h, int(max_iters)), "m2_batch_run")

    def sync(self) -> None:
        _chk(lib().m2_batch_sync(self.h), "m2_batch_sync")

    def iteration(self, i: int) -> int:
        v = ctypes.c_uint64()
        _chk(lib().m2_batch_iteration(self.h, i, ctypes.byref(v)), "m2_batch_iteration")
        return v.value

    def res64(self, i: int) -> str:
        v = ctypes.c_uint64()
        _chk(lib().m2_batch_res64(self.h, i, ctypes.byref(v)), "m2_batch_res64")
        return "%016X" % v.value

    def residue(self, i: int) -> bytes:
        n = (self.ps[i] + 7) // 8
        buf = ctypes.create_string_buffer(n)
        _chk(lib().m2_batch_residue(self.h, i, buf, n), "m2_batch_residue")
        return buf.raw[:n]

    def is_prime(self, i: int) -> bool:
        v = ctypes.c_int()
        _chk(lib().m2_batch_is_prime(self.h, i, ctypes.byref(v)), "m2_batch_is_prime")
        return bool(v.value)


class Server:
    """Request queue over resident tests (mirror of m2_server_*)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int = 0, slice_iters: int = 0, local_stage_cap: int = 0,
                 **opts):
        opts = make_opts(local_stage_cap=local_stage_cap, **opts)
        h = ctypes.c_void_p()
        _chk(lib().m2_server_create(ctx.h, int(hbm_budget_bytes), int(slice_iters), ctypes.byref(opts),
                                    ctypes.byref(h)), "m2_server_create")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().m2_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, p: int, mode="prp", iters: int = 0) -> int:
        v = ctypes.c_uint64()
        _chk(lib().m2_server_submit(self.h, int(p), _mode(mode), int(iters), ctypes.byref(v)), "m2_server_submit")
        return v.value

    def step(self) -> bool:
        """One admit / slice / retire round; True while work remains."""
        v = ctypes.c_int()
        _chk(lib().m2_server_step(self.h, ctypes.byref(v)), "m2_server_step")
        return bool(v.value)

    def poll(self, max_items: int = 4096) -> List[Tuple[int, str, bool]]:
        ids = (ctypes.c_uint64 * max_items)()
        r64 = (ctypes.c_uint64 * max_items)()
        ver = (ctypes.c_int * max_items)()
        cnt = ctypes.c_size_t()
        _chk(lib().m2_server_poll(self.h, ids, r64, ver, max_items, ctypes.byref(cnt)), "m2_server_poll")
        return [(ids[i], "%016X" % r64[i], bool(ver[i])) for i in range(cnt.value)]

    def stats(self) -> Dict[str, int]:
        q, a, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _chk(lib().m2_server_stats(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(b)), "m2_server_stats")
        return {"queued": q.value, "active": a.value, "bytes_in_use": b.value}


def square_noalloc(ctx: Context, p: int, mode: str, residue: bytes, iters: int) -> bytes:
    """The residue after `iters` iterations from `residue`, through the entry
    point that allocates nothing (m2_square_noalloc; n <= 2^12)."""
    n = (int(p) + 7) // 8
    out = ctypes.create_string_buffer(n)
    _chk(lib().m2_square_noalloc(ctx.h, int(p), _mode(mode), bytes(residue), len(residue), int(iters), out, n),
         "m2_square_noalloc")
    return out.raw[:n]


def run_checkpoints(ctx: Context, p: int, mode: str, checkpoints: Sequence[int]) -> Dict[int, str]:
    """Res64 after each checkpoint iteration."""
    out: Dict[int, str] = {}
    with Test(ctx, p, mode) as t:
        done = 0
        for k in sorted(set(int(k) for k in checkpoints)):
            t.run(k - done)
            done = k
            out[k] = t.res64()
    return out
