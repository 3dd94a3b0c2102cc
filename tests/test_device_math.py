"""Host builds of the product's __host__ __device__ math (ugks_device.hpp)
against library and oracle values: the modified Bessel function I0 of the
Cercignani-Lampis wall against scipy.special.i0 over both of its branches,
and the scattering kernel against the oracle's independent restatement.
Needs hipcc (skipped without it); no GPU is used."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import special

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)

PROGRAM = r"""
#include "ugks_device.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
int main(int argc, char **argv) {
    if (!strcmp(argv[1], "i0")) {
        for (int i = 2; i < argc; i++) printf("%.17g\n", ugks::bessel_i0(atof(argv[i])));
    } else {   // cl uin.x uin.y uout.x uout.y an at Tw
        ugks::d2 a{atof(argv[2]), atof(argv[3])}, b{atof(argv[4]), atof(argv[5])};
        printf("%.17g\n", ugks::cl_kernel(a, b, atof(argv[6]), atof(argv[7]), atof(argv[8])));
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("devmath")
    src, exe = d / "t.cpp", d / "t"
    src.write_text(PROGRAM)
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-I", os.path.join(ROOT, "pyugks_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    return str(exe)


def test_bessel_i0_both_branches(prog):
    xs = np.concatenate([np.linspace(0.0, 25.0, 51), [24.999, 25.001], np.linspace(25.5, 700.0, 40), [-7.25]])
    out = subprocess.run([prog, "i0", *[repr(float(x)) for x in xs]], capture_output=True, text=True, check=True)
    got = np.array([float(t) for t in out.stdout.split()])
    assert np.max(np.abs(got / special.i0(xs) - 1)) < 3e-15


def test_cl_kernel_matches_oracle(prog):
    H = O.helpers()
    rng = np.random.default_rng(3)
    for _ in range(30):
        uin = (-rng.uniform(0.05, 3.0), rng.uniform(-3, 3))
        uout = (rng.uniform(0.05, 3.0), rng.uniform(-3, 3))
        an, at, Tw = rng.uniform(0.1, 1.0), rng.uniform(0.1, 1.0), rng.uniform(0.5, 2.0)
        args = [repr(float(x)) for x in (*uin, *uout, an, at, Tw)]
        got = float(subprocess.run([prog, "cl", *args], capture_output=True, text=True, check=True).stdout)
        ref = H.or_cl_kernel(O.V2(*uin), O.V2(*uout), an, at, Tw)
        assert abs(got - ref) <= 1e-13 * abs(ref) + 1e-300
