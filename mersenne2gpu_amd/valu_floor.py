"""VALU instruction floor of the iteration's kernels against the issued count.

    python -m mersenne2gpu_amd.valu_floor

Compiles csrc/probes/valu_floor.hip for gfx950 (hipcc, no GPU needed) and
counts the VALU instructions of each arithmetic primitive in the assembly: a
radix-r butterfly with and without twiddle products in either field, the
squaring of a spectral pair, and the carry of one 8-digit segment (unweight,
CRT and normalisation, in the 64-bit form that digit widths 16..32 take, as
norm64 the one of widths 32..44 and as norm128 the 128-bit form of the
others; ripple; weight), each minus a skeleton kernel with the
same loads and stores. A kernel's floor per wave is the sum of the
primitives each of its threads runs (half the waves of every kernel work on
the M61 field and half on the M31 field; every thread carries one segment),
i.e. the arithmetic alone, without addressing, LDS traffic, staging or loop
control. The issued counts are SQ_INSTS_VALU / SQ_WAVES of rocprofv3 runs,
from profiles/c2_prp_n2p20_sq_counters.md and profiles/ll_n2p24_sq_counters.md
(override with --issued kernel=value).
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import tempfile

from . import build as B

PROBE = os.path.join(B.CSRC, "probes", "valu_floor.hip")

# Threads of one field role per kernel: the primitives each runs, by count.
# Names: fwd/inv, radix, t (twiddle products) / u (level 0, none); sq: one
# squaring pair; carry parts are run by every thread (both roles).
PLANS = {
    "n=2^20 PRP (p=27899687)": {
        "k_centre<1024, 10>": ({"fwd4t": 5, "inv4t": 5, "sq": 2}, {}),
        "k_outer_fused<9, 2, 10>": ({"inv8t": 2, "inv8u": 1, "fwd8u": 1, "fwd8t": 2},
                                    {"norm": 1, "ripple": 1, "weight": 1}),
        "k_outer_fwd<9, 2, 10>": ({"fwd8u": 1, "fwd8t": 2}, {"ripple": 1, "weight": 1}),
        "k_outer_inv<9, 2, 10>": ({"inv8u": 1, "inv8t": 2}, {"norm": 1}),
    },
    "n=2^24 LL (p=449846671)": {
        "k_centre<512, 10>": ({"fwd2t": 4, "fwd8t": 3, "inv8t": 3, "inv2t": 4, "sq": 4}, {}),
        "k_outer_inv<7, 4, 16>": ({"inv8t": 2, "inv2u": 4}, {"norm": 1, "ripple": 0.75}),
        "k_outer_fwd<7, 4, 16>": ({"fwd2u": 4, "fwd8t": 2}, {"ripple": 0.25, "weight": 1}),
        "k_mid<true, 6, 5, 10>": ({"fwd8t": 2}, {}),
        "k_mid<false, 6, 5, 10>": ({"inv8t": 2}, {}),
    },
}

# SQ_INSTS_VALU per wave, rocprofv3 (profiles named above)
ISSUED = {
    "k_centre<1024, 10>": 2402, "k_outer_fused<9, 2, 10>": 3204, "k_outer_fwd<9, 2, 10>": 1538,
    "k_outer_inv<9, 2, 10>": 1904,
    "k_centre<512, 10>": 4501, "k_outer_inv<7, 4, 16>": 1893, "k_outer_fwd<7, 4, 16>": 1466,
    "k_mid<true, 6, 5, 10>": 1083, "k_mid<false, 6, 5, 10>": 1091,
}


def compile_probe(workdir: str) -> str:
    cmd = [B.hipcc(), "-x", "hip", "--offload-arch=" + B.ARCH, "-O3", "-std=c++17", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "--save-temps", "-I", B.CSRC, "-I", B.INCLUDE, "-c", PROBE, "-o",
           os.path.join(workdir, "probe.o")]
    r = subprocess.run(cmd, cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError("probe build failed:\n" + r.stdout)
    for f in os.listdir(workdir):
        if f.endswith("-hip-amdgcn-amd-amdhsa-%s.s" % B.ARCH):
            with open(os.path.join(workdir, f)) as fh:
                return fh.read()
    raise RuntimeError("no device assembly written")


def valu_count(asm: str, name: str) -> int:
    m = re.search(r"^%s:[^\n]*\n(.*?)\bs_endpgm" % re.escape(name), asm, re.S | re.M)
    if not m:
        raise KeyError(name)
    return sum(1 for line in m.group(1).splitlines() if re.match(r"\s+v_", line))


def primitives(asm: str) -> dict:
    c = lambda n: valu_count(asm, n)
    out = {}
    for f in ("61", "31"):
        for r in ("8", "4", "2"):
            sk = c("skel%s_%s" % (r, f))
            for kind in ("fwd%st", "inv%st", "fwd%su", "inv%su"):
                name = (kind % r) + "_" + f
                try:
                    out[name] = c(name) - sk
                except KeyError:
                    pass
        out["sq_" + f] = c("sq_" + f) - c("skel2_" + f)
    sk = c("skel_carry")
    out["norm"] = c("carry_norm") - sk
    out["norm64"] = c("carry_norm64") - sk
    out["norm128"] = c("carry_norm128") - sk
    out["weight"] = c("carry_w") - sk
    out["ripple"] = c("carry_rw") - c("carry_w")
    out["carry_all"] = c("carry_all") - sk
    return out


def floors(prim: dict) -> dict:
    res = {}
    for plan, kernels in PLANS.items():
        for k, (field, carry) in kernels.items():
            per = {f: sum(n * prim["%s_%s" % (p, f)] for p, n in field.items()) for f in ("61", "31")}
            cw = sum(n * prim[p] for p, n in carry.items())
            res[k] = (plan, per["61"], per["31"], cw, (per["61"] + per["31"]) / 2.0 + cw)
    return res


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--issued", action="append", default=[], help="kernel=VALU instructions per wave")
    args = ap.parse_args(argv)
    issued = dict(ISSUED)
    for kv in args.issued:
        k, v = kv.rsplit("=", 1)
        issued[k] = float(v)
    with tempfile.TemporaryDirectory() as d:
        prim = primitives(compile_probe(d))
    print("primitive VALU instructions (gfx950, -O3):")
    for k in sorted(prim):
        print("  %-10s %5d" % (k, prim[k]))
    print("\n| workload | kernel | M61 thread | M31 thread | carry | floor / wave | issued / wave | issued / floor |")
    print("|---|---|---:|---:|---:|---:|---:|---:|")
    for k, (plan, p61, p31, cw, fl) in floors(prim).items():
        iss = issued.get(k)
        print("| %s | %s | %d | %d | %d | %.0f | %s | %s |" % (
            plan, k, p61, p31, cw, fl, "%.0f" % iss if iss else "-", "%.2f" % (iss / fl) if iss else "-"))


if __name__ == "__main__":
    main()
