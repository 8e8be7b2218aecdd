"""The branch-free stencil forms the Jacobian kernel uses round exactly like
the per-kind stencil (dt_stencil.hpp): dstencil_sel for every kind and
dstencil_c4c2 for the two central kinds, checked bit for bit on the CPU on
random and special samples, built with contraction off as the kernels are
(reference stencils: src/headers/jacobian_macros.h:4-195)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "disptools_amd", "csrc")

PROGRAM = r"""
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include "dt_stencil.hpp"
using namespace dt;

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static float sample() {
    static const float special[] = {0.0f, -0.0f, 1.0f, -1.0f, 1e-30f, -1e-30f, 3e38f, -3e38f, 1e-45f, 0.5f};
    const uint64_t r = next();
    if ((r & 15) == 0) return special[(r >> 4) % 10];
    const int e = (int)((r >> 8) % 40) - 20;
    return std::ldexp((float)((r >> 20) % 2000001) / 1e6f - 1.0f, e);
}
static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0 || (std::isnan(a) && std::isnan(b)); }

int main() {
    long bad = 0;
    for (long i = 0; i < 2000000; ++i) {
        const float m2 = sample(), m1 = sample(), c0 = sample(), p1 = sample(), p2 = sample();
        const float inv = std::fabs(sample()) + 1e-3f;
        for (int k = K_C4; k <= K_BWD; ++k) {
            const double ref = dstencil<double>(k, m2, m1, c0, p1, p2, inv);
            if (!same(dstencil_sel(stencil_sel(k), m2, m1, c0, p1, p2, inv), ref)) ++bad;
            if (k <= K_C2 && !same(dstencil_c4c2(k == K_C4, m2, m1, p1, p2, inv), ref)) ++bad;
        }
    }
    std::printf("%s %ld\n", bad ? "mismatch" : "ok", bad);
    return bad ? 1 : 0;
}
"""


def test_branch_free_stencils_round_like_the_per_kind_stencil(tmp_path):
    cxx = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if cxx is None:
        pytest.skip("hipcc not found")
    src = tmp_path / "forms.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "forms"
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
