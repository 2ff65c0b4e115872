"""The Python restatement of std::mt19937 + std::uniform_real_distribution used to
predict PdfBase::fillMCDataSimple's events, checked against the C++ standard's known
answer and against libstdc++ itself (a small program built with g++)."""
import os
import shutil
import subprocess

import pytest

from _stdrandom import MT19937, uniform


def test_mt19937_standard_known_answer():
    """[rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995."""
    g = MT19937()
    for _ in range(9999):
        g()
    assert g() == 4123659995


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_uniform_draws_match_libstdcxx(tmp_path):
    """The draw pattern of fillMCDataSimple (a [0, 1) uniform, then a [-1/2, 1/2) one) for
    two seeds, bit for bit against libstdc++."""
    src = tmp_path / "draws.cpp"
    src.write_text(
        "#include <cstdio>\n#include <cstdlib>\n#include <random>\n"
        "int main(int argc, char **argv) {\n"
        "  std::mt19937 g((unsigned)std::strtoul(argv[1], nullptr, 10));\n"
        "  std::uniform_real_distribution<> a(0.0, 1.0), b(-.5, .5);\n"
        "  for(int i = 0; i < 2000; i++) { double u = a(g); double v = b(g); std::printf(\"%a %a\\n\", u, v); }\n"
        "}\n")
    exe = tmp_path / "draws"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    for seed in (1, 20231, 4294967295):
        out = subprocess.run([str(exe), str(seed)], check=True, capture_output=True, text=True).stdout.split()
        g = MT19937(seed)
        for k in range(2000):
            u, v = uniform(g, 0.0, 1.0), uniform(g, -0.5, 0.5)
            assert float.fromhex(out[2 * k]) == u and float.fromhex(out[2 * k + 1]) == v, (seed, k)
