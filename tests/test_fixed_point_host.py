"""The fixed-point accumulator of ETRAJ_PMD_FIXED (fix_from_double /
fix_to_double in etraj.jl_amd/csrc/etraj_engine.hpp), compiled for the host
and checked against exact rational arithmetic:

  * fix_from_double(v) == round(v * 2^100) with ties to even, as a 128-bit
    two's-complement integer, for values across the whole range, at the
    rounding boundary and on both sides of the range limit 2^26;
  * fix_to_double(q) == the double nearest to q / 2^100 (ties to even);
  * the sum of the integers of a set of values is the same in any order and
    equals the exact sum of the rounded values.
"""
import math
import os
import random
import shutil
import subprocess
from fractions import Fraction

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r"""
#include "etraj_engine.hpp"
#include <cstdio>
#include <cstdlib>
int main() {
    char op[8];
    while (std::scanf("%7s", op) == 1) {
        if (op[0] == 'f') {                  // f <double as hex float>
            double v; std::scanf("%la", &v);
            bool ok; etraj::Fix128 q = etraj::fix_from_double(v, ok);
            std::printf("%d %llu %llu\n", ok ? 1 : 0, (unsigned long long)q.lo, (unsigned long long)q.hi);
        } else {                             // t <lo> <hi>
            unsigned long long lo, hi; std::scanf("%llu %llu", &lo, &hi);
            std::printf("%a\n", etraj::fix_to_double(lo, hi));
        }
    }
    return 0;
}
"""

M128 = (1 << 128) - 1


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("fix")
    (d / "fix.cpp").write_text(SRC)
    exe = d / "fix"
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    subprocess.run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I",
                    os.path.join(rocm, "include"), "-I", os.path.join(ROOT, "etraj.jl_amd", "csrc"),
                    str(d / "fix.cpp"), "-o", str(exe)], check=True, capture_output=True)

    def run(lines):
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                             check=True).stdout.split("\n")
        return [x for x in out if x]
    return run


def exact_q(v):
    """round(v 2^100), ties to even (Python's round of a Fraction)."""
    return round(Fraction(v) * 2 ** 100)


def values(seed=5, n=4000):
    rnd = random.Random(seed)
    vs = [0.0, 2.0 ** -100, 3 * 2.0 ** -102, 2.0 ** -101, 3 * 2.0 ** -101, 5 * 2.0 ** -102, 2.0 ** -102,
          math.nextafter(2.0 ** 26, 0), 1.0, 0.1, 1e-16, 1e-30, 7.9e-31]
    vs += [-v for v in vs]
    for _ in range(n):
        vs.append(rnd.choice([-1, 1]) * rnd.random() * 2.0 ** rnd.randint(-105, 25))
    return vs


def test_from_double_is_round_to_nearest_even(prog):
    vs = values()
    out = prog([f"f {v.hex()}" for v in vs])
    for v, line in zip(vs, out):
        ok, lo, hi = (int(x) for x in line.split())
        assert ok == 1
        q = exact_q(v)
        assert (hi << 64 | lo) == q & M128, v


def test_range_limit(prog):
    out = prog([f"f {v.hex()}" for v in (2.0 ** 26, -2.0 ** 26, math.inf, math.nan, 3e9)])
    assert [line.split()[0] for line in out] == ["0"] * 5


def test_to_double_is_correctly_rounded(prog):
    rnd = random.Random(11)
    qs = [1, -1, 3, (1 << 100), -(1 << 100), (1 << 126) - 1, -((1 << 126) - 1), (1 << 64) - 1, 1 << 64,
          (1 << 64) + 1, (1 << 117) + (1 << 64) + 1]
    qs += [rnd.choice([-1, 1]) * rnd.getrandbits(rnd.randint(1, 126)) for _ in range(3000)]
    out = prog([f"t {q & ((1 << 64) - 1)} {(q & M128) >> 64}" for q in qs])
    for q, line in zip(qs, out):
        assert float.fromhex(line) == float(Fraction(q, 2 ** 100)), q


def test_sums_are_order_independent_and_exact(prog):
    vs = values(seed=9, n=3000)
    out = prog([f"f {v.hex()}" for v in vs])
    qs = [(int(hi) << 64 | int(lo)) for _, lo, hi in (line.split() for line in out)]
    total = sum(qs) & M128
    rnd = random.Random(3)
    for _ in range(3):
        order = list(range(len(qs)))
        rnd.shuffle(order)
        acc = 0
        for i in order:
            acc = (acc + qs[i]) & M128            # 128-bit wrap-around, as the device adds
        assert acc == total
    signed = total - (1 << 128) if total >> 127 else total
    assert signed == sum(exact_q(v) for v in vs)
    # the exact sum of the rounded values is within n half-units of the exact sum
    assert abs(Fraction(signed, 2 ** 100) - sum(Fraction(v) for v in vs)) <= Fraction(len(vs), 2 ** 101)
