"""ctypes binding of the oracle (oracle/liboracle.so); used only by tests,
__graft_entry__.smoke() and bench.py's CPU baseline."""
import ctypes


class GF(ctypes.Structure):
    _fields_ = [("s0", ctypes.c_uint64), ("s1", ctypes.c_uint64), ("t0", ctypes.c_uint32), ("t1", ctypes.c_uint32)]


class Oracle:
    def __init__(self, path):
        L = ctypes.CDLL(path)
        u64, u32, i = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        for name, res, args in [
            ("m2o_add61", u64, [u64, u64]), ("m2o_sub61", u64, [u64, u64]), ("m2o_mul61", u64, [u64, u64]),
            ("m2o_shl61", u64, [u64, ctypes.c_uint]), ("m2o_add31", u32, [u32, u32]), ("m2o_sub31", u32, [u32, u32]),
            ("m2o_mul31", u32, [u32, u32]), ("m2o_shl31", u32, [u32, ctypes.c_uint]),
            ("m2o_gf_mul", None, [ctypes.POINTER(GF)] * 3),
            ("m2o_transformsize", i, [u32]),
            ("m2o_digit_tables", i, [u32, i, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
            ("m2o_twiddles", None, [i, ctypes.POINTER(GF)]), ("m2o_root", None, [i, ctypes.POINTER(GF)]),
            ("m2o_forward", None, [i, ctypes.POINTER(GF), ctypes.POINTER(GF)]),
            ("m2o_square_half", None, [i, ctypes.POINTER(GF), ctypes.POINTER(GF)]),
            ("m2o_inverse", None, [i, ctypes.POINTER(GF), ctypes.POINTER(GF)]),
            ("m2o_garner", None, [u64, u32, ctypes.POINTER(u64)]),
            ("m2o_test_new", ctypes.c_void_p, [u32, i, u32]), ("m2o_test_free", None, [ctypes.c_void_p]),
            ("m2o_test_run", None, [ctypes.c_void_p, u64]), ("m2o_test_res64", u64, [ctypes.c_void_p]),
            ("m2o_test_residue", i, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
            ("m2o_test_verdict", i, [ctypes.c_void_p]), ("m2o_test_ln", i, [ctypes.c_void_p]),
            ("m2o_test_digits", None, [ctypes.c_void_p, ctypes.POINTER(u64)]),
            ("m2o_digits_residue", i, [u32, i, ctypes.POINTER(u64), ctypes.c_char_p, ctypes.c_size_t]),
            ("m2o_set_threads", i, [i]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    def __getattr__(self, name):
        return getattr(self.L, "m2o_" + name)

    # helpers -------------------------------------------------------------
    def tables(self, p, ln):
        n = 1 << ln
        wd, w61, w31 = (ctypes.create_string_buffer(n) for _ in range(3))
        rc = self.L.m2o_digit_tables(p, ln, wd, w61, w31)
        return rc, list(wd.raw), list(w61.raw), list(w31.raw)

    def twiddles(self, ln):
        w = (GF * (1 << ln))()
        self.L.m2o_twiddles(ln, w)
        return w

    def run(self, p, mode, checkpoints, carry_pairs=64):
        t = self.L.m2o_test_new(p, 1 if mode == "prp" else 0, carry_pairs)
        out, done = {}, 0
        try:
            for k in sorted(set(checkpoints)):
                self.L.m2o_test_run(t, k - done)
                done = k
                out[k] = "%016X" % self.L.m2o_test_res64(t)
            verdict = self.L.m2o_test_verdict(t)
            n = (p + 7) // 8
            buf = ctypes.create_string_buffer(max(n, 8))
            self.L.m2o_test_residue(t, buf, max(n, 8))
            residue = buf.raw[:n]
        finally:
            self.L.m2o_test_free(t)
        return out, verdict, residue
