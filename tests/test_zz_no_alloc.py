"""m2_square_noalloc (the entry point that allocates nothing) against the
test path (m2_test_set_residue + m2_test_run + m2_test_residue), which runs
the same single-workgroup kernel on buffers the library allocates. Bit-exact
residues, no tolerance; the scratch array keeps each call's contents, so the
calls run back to back with shrinking inputs.
"""
import random

import pytest

import mersenne2gpu_amd as m2

pytestmark = pytest.mark.gpu


def _alone(ctx, p, mode, r, iters):
    with m2.Test(ctx, p, mode) as t:
        t.set_residue(r, 0)
        t.run(iters)
        return t.residue()


def _rand_residue(p, seed):
    return random.Random(seed).getrandbits(p).to_bytes((p + 7) // 8, "little")


def test_noalloc_matches_the_test_path_on_successive_calls(gpu_ctx):
    # largest length first (n = 2^12), then smaller inputs on the same array
    cases = [(159739, "prp", 300), (81919, "ll", 200), (19937, "prp", 5000), (4423, "ll", 4096), (127, "prp", 9),
             (7, "ll", 5)]
    for i, (p, mode, iters) in enumerate(cases):
        r = _rand_residue(p, i)
        print("noalloc p=%d %s iters=%d" % (p, mode, iters), flush=True)
        got = m2.square_noalloc(gpu_ctx, p, mode, r, iters)
        assert got == _alone(gpu_ctx, p, mode, r, iters), (p, mode, iters)


def test_noalloc_known_answers(gpu_ctx):
    # a whole PRP test of the Mersenne prime 2^19937 - 1 from 3 ends at 9;
    # 3 squarings of 3 mod 2^127 - 1 give 3^8
    p = 19937
    got = m2.square_noalloc(gpu_ctx, p, "prp", (3).to_bytes((p + 7) // 8, "little"), p)
    assert int.from_bytes(got, "little") == 9
    got = m2.square_noalloc(gpu_ctx, 127, "prp", (3).to_bytes(16, "little"), 3)
    assert int.from_bytes(got, "little") == 3 ** 8
    # zero iterations: the canonical form of the input (2^p - 1 is 0)
    got = m2.square_noalloc(gpu_ctx, 127, "prp", ((1 << 127) - 1).to_bytes(16, "little"), 0)
    assert int.from_bytes(got, "little") == 0


def test_noalloc_refuses_what_does_not_fit(gpu_ctx):
    p = 159763  # n = 2^13: beyond the scratch array
    with pytest.raises(m2.M2Error) as ei:
        m2.square_noalloc(gpu_ctx, p, "prp", (3).to_bytes((p + 7) // 8, "little"), 1)
    assert ei.value.code == -5 and "scratch" in str(ei.value)
    with pytest.raises(m2.M2Error) as ei:
        m2.square_noalloc(gpu_ctx, 19937, "prp", b"\x03", 1)  # residue shorter than ceil(p/8)
    assert ei.value.code == -1
    # and the array still serves a valid call afterwards
    r = _rand_residue(19937, 7)
    assert m2.square_noalloc(gpu_ctx, 19937, "prp", r, 10) == _alone(gpu_ctx, 19937, "prp", r, 10)
