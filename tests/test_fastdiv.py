"""The precomputed divisors that split a two-kernel tile index
(csrc/fdtd_fastdiv.hpp): the quotient must equal integer division for every
numerator the kernels pass (0 <= n < 2^31).  The header needs no HIP, so the
check is a small C++ program built with g++ on the host: every divisor up to
4096 against numerators around each multiple and at the range ends, plus
powers of two, their neighbours and random divisors against random and edge
numerators."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "3d_fdtd_version008_amd", "csrc")

PROGRAM = r"""
#include "fdtd_fastdiv.hpp"
#include <cstdio>
#include <cstdint>
#include <vector>

using fdtdk::fdiv;
using fdtdk::make_fastdiv;

static_assert(fdiv(1000, make_fastdiv(7)) == 142, "constexpr evaluation");

static uint64_t st = 0x9e3779b97f4a7c15ull;
static unsigned rnd31() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (unsigned)(st >> 33);  // 31 bits
}

static int check(unsigned d, unsigned n) {
    const unsigned q = fdiv(n, make_fastdiv(d));
    if (q != n / d) {
        std::printf("FAIL d=%u n=%u got %u want %u\n", d, n, q, n / d);
        return 1;
    }
    return 0;
}

int main() {
    const unsigned top = 0x7fffffffu;
    int bad = 0;
    std::vector<unsigned> ds;
    for (unsigned d = 1; d <= 4096; ++d) ds.push_back(d);
    for (int b = 12; b <= 31; ++b) {
        const unsigned p = 1u << b;
        ds.push_back(p - 1); ds.push_back(p); if (b < 31) ds.push_back(p + 1);
    }
    for (int r = 0; r < 2000; ++r) ds.push_back(1 + rnd31() % top);
    for (unsigned d : ds) {
        bad += check(d, 0); bad += check(d, top); bad += check(d, top - 1);
        // around multiples of d
        for (int r = 0; r < 24; ++r) {
            const unsigned q = rnd31() / d;
            const uint64_t base = (uint64_t)q * d;
            for (int o = -1; o <= 1; ++o) {
                const int64_t n = (int64_t)base + o;
                if (n >= 0 && n <= (int64_t)top) bad += check(d, (unsigned)n);
            }
        }
        for (int r = 0; r < 64; ++r) bad += check(d, rnd31());
        if (bad) break;
    }
    std::printf("%s %zu divisors\n", bad ? "FAILED" : "ok", ds.size());
    return bad ? 1 : 0;
}
"""


def test_fastdiv_matches_integer_division():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "fd.cpp")
        exe = os.path.join(tmp, "fd")
        with open(src, "w") as f:
            f.write(PROGRAM)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.startswith("ok"), r.stdout
