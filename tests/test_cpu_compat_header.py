"""The C++ compatibility header compiles against the C ABI and links the library
(host-only calls: fermi and dipole; no GPU work)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "g12kba_compat.hpp"
#include <cstdio>
int main() {
    std::vector<std::vector<std::complex<double>>> G(3, std::vector<std::complex<double>>(4));
    G[1][0] = {0.0, 1.0};
    std::vector<double> p(3, 0.0);
    dipole(p, G);
    std::printf("%.6f %.6f\n", fermi(10.0, 10.0, 0.5), p[1]);
    return 0;
}
'''


def test_compat_header_compiles_and_links():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "t.cpp")
    open(src, "w").write(SRC)
    lib = os.path.join(ROOT, "g12kba_amd", "lib")
    exe = os.path.join(d, "t")
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-L", lib, "-lg12kba",
                           f"-Wl,-rpath,{lib}", "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    f, p1 = map(float, out.stdout.split())
    assert f == 0.5
    assert abs(p1 - 0.5) < 1e-12  # N = 2: 2*(1 - 0 - 0.5)*(1 - 0)/2
