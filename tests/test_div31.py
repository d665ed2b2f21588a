"""Host check of the multiply-shift divisor the path kernels use for their claim and pixel index arithmetic
(makeDiv31 in cloudrendering_amd/csrc/vpt_types.h): n / d == (n * m) >> s for every numerator below 2^31, checked
on divisors of every bit length, powers of two and their neighbours, and the numerators where rounding errors would
show first (multiples of d and their predecessors near 2^31)."""
import os
import shutil
import subprocess

import pytest

SRC = r'''
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>
#include "vpt_types.h"
using namespace crvpt;
static bool check(uint32_t d, uint32_t n) {
    const Div31 q = makeDiv31(d);
    return uint32_t((uint64_t(n) * q.m) >> q.s) == n / d;
}
int main() {
    std::mt19937_64 rng(7);
    std::vector<uint32_t> ds;
    for (uint32_t b = 0; b < 32; b++) {
        const uint32_t p = 1u << b;
        for (int64_t o = -2; o <= 2; o++) {
            const int64_t d = int64_t(p) + o;
            if (d >= 1 && d <= 0xFFFFFFFFll) ds.push_back(uint32_t(d));
        }
        for (int k = 0; k < 200; k++) ds.push_back(uint32_t(p | (rng() & (p - 1))) | 1u);
    }
    for (uint32_t d : {3u, 5u, 7u, 256u * 255u, 1048576u, 2073600u, 1920u * 1080u, 65u, 64u * 1023u}) ds.push_back(d);
    long bad = 0;
    for (uint32_t d : ds) {
        const uint32_t top = 0x7FFFFFFFu;
        for (uint32_t n : {0u, 1u, d - 1u, d, d + 1u, top, top - 1u, top - top % d, top - top % d - 1u})
            if (n <= top && !check(d, n)) bad++;
        for (int k = 0; k < 2000; k++) {
            const uint32_t n = uint32_t(rng() & top);
            if (!check(d, n)) bad++;
            const uint64_t mult = (uint64_t(n) / d) * d;  // a multiple of d and its predecessor
            if (mult <= top && !check(d, uint32_t(mult))) bad++;
            if (mult >= 1 && mult - 1 <= top && !check(d, uint32_t(mult - 1))) bad++;
        }
    }
    std::printf("%zu divisors, %ld mismatches\n", ds.size(), bad);
    return bad != 0;
}
'''


def test_div31_exact_below_2_31(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    csrc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cloudrendering_amd", "csrc")
    src = tmp_path / "div31.cpp"
    src.write_text(SRC)
    exe = tmp_path / "div31"
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", csrc, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
