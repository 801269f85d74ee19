"""The C++ compatibility header compiles against the C ABI and links the library
(host-only calls: fermi, dipole, make_G0 and init_G_test; no GPU work)."""
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
    std::vector<double> eps(2, 0.0);
    std::vector<std::complex<double>> g0, gt(36);
    make_G0(g0, 3, 2, 2, 1.0, 1.0, 0.5, 0.0, eps);   // 2 of 6 levels filled
    init_G_test(gt, 3, 2);                           // site 0 of both orbitals
    double tr0 = 0, trt = 0;
    for (int i = 0; i < 6; ++i) { tr0 += g0[i * 7].imag(); trt += gt[i * 7].imag(); }
    std::printf("%.6f %.6f %.1f %.1f\n", tr0, trt, gt[0].imag(), gt[3 * 7].imag());
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
    lines = out.stdout.split("\n")
    f, p1 = map(float, lines[0].split())
    assert f == 0.5
    assert abs(p1 - 0.5) < 1e-12  # N = 2: 2*(1 - 0 - 0.5)*(1 - 0)/2
    tr0, trt, g00, g33 = map(float, lines[1].split())
    assert abs(tr0 - 2) < 1e-6 and trt == 2 and g00 == 1 and g33 == 1
