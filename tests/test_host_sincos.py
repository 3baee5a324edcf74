"""The FP64 joint-angle sin/cos of the compiled kernels (sincos_q in
gp_device.h: three-part pi/2 reduction and Taylor polynomials), built for the
host through the HIP stand-in, against the x87 extended-precision sin and cos
(numpy longdouble, 64-bit mantissa) on 400,000 angles: random ones in
[-1e4, 1e4] and near every multiple of pi/4 up to 1e3 (where the reduction
and the quadrant choice matter), plus large angles up to 1e14 (which first
lose a multiple of 2 pi held in two parts), and infinities and NaN (NaN out).
Bound: 4e-16 absolute (two units in the last place of a value of size one)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fp64_joint_sincos(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    if np.finfo(np.longdouble).nmant < 63:
        pytest.skip("no extended-precision long double here")
    exe = str(tmp_path / "sincos")
    p = subprocess.run([gxx, "-std=c++17", "-O2", "-I", os.path.join(ROOT, "tests", "host"), "-x", "c++",
                        os.path.join(ROOT, "tests", "host", "sincos_host_main.cpp"), "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-3000:]
    rng = np.random.default_rng(0)
    quarter = np.arange(-1273, 1274) * (np.pi / 4)
    near = (quarter[:, None] + rng.uniform(-1e-6, 1e-6, (quarter.size, 40))).ravel()
    big = rng.uniform(8e5, 1e7, 20000) * rng.choice([-1.0, 1.0], 20000)
    huge = 10.0 ** rng.uniform(7, 14, 20000)
    x = np.concatenate([rng.uniform(-1e4, 1e4, 300000), near[:60000], big, huge,
                        [0.0, -0.0, 1e-300, 3e5, 823549.0, 823551.0, 1e6, -2e7, 1e12]])
    special = np.array([np.inf, -np.inf, np.nan])
    x = np.concatenate([x, special])
    x.tofile(tmp_path / "x.bin")
    p = subprocess.run([exe, str(tmp_path / "x.bin"), str(tmp_path / "y.bin")], capture_output=True, timeout=300)
    assert p.returncode == 0
    y = np.fromfile(tmp_path / "y.bin", dtype=np.float64)
    s, c = y[:x.size], y[x.size:]
    assert np.isnan(s[-3:]).all() and np.isnan(c[-3:]).all()
    x, s, c = x[:-3], s[:-3], c[:-3]
    xl = x.astype(np.longdouble)
    es = np.abs(s.astype(np.longdouble) - np.sin(xl)).max()
    ec = np.abs(c.astype(np.longdouble) - np.cos(xl)).max()
    print(f"max |sin error| {float(es):.3g}, max |cos error| {float(ec):.3g}")
    assert es <= 4e-16 and ec <= 4e-16
