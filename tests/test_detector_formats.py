"""Detector files: the Python writers (3d_fdtd_version008_amd/detectors.py)
must produce the bytes the reference's C++ stream output produces.  The
reference writes its TLS CSV rows with iostream manipulators
(std::setprecision(18) in PopulationTimeSeriesDetector / PolarizationDiagnostics,
src/two_level_detectors.hpp:91-103, 250-257; fixed 6 / scientific 12 / fixed 8
in TLSGlobalHistory, :323-331).  A small C++ program applies those stream
settings to the same records with g++'s libstdc++, and its output is compared
byte for byte with the writer's, on values chosen to exercise the
conversions (zero, negative, tiny and huge magnitudes, exact integers,
fractions near rounding boundaries)."""
import importlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

PROGRAM = r"""
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>
int main(int argc, char** argv) {
    const double dt = std::strtod(argv[1], nullptr);
    const long every = std::atol(argv[2]);
    std::vector<double> v;
    double x;
    while (std::scanf("%lf", &x) == 1) v.push_back(x);
    std::ofstream fi(std::string(argv[3]) + "/inv.csv"), fp(std::string(argv[3]) + "/pol.csv"),
        fh(std::string(argv[3]) + "/hist.csv");
    fi << "t, total_Nu, total_Ng, inversion, inv_fraction\n";
    fp << "t, Pz_center, dPz_dt_center, total_P_energy\n";
    fh << "step,t_fs,Nu_total,Ng_total,inversion_total,Nu_percent,Ng_percent,inv_norm\n";
    for (size_t r = 0; r * 5 + 4 < v.size(); ++r) {
        const double nu = v[r * 5], ng = v[r * 5 + 1], u = v[r * 5 + 2], pz = v[r * 5 + 3], dpz = v[r * 5 + 4];
        const size_t n = r * every;
        const double t = n * dt;
        const double inversion = nu - ng, denom = nu + ng;
        const double frac = (denom > 1e-30) ? (inversion / denom) : 0.0;
        fi << std::setprecision(18) << t << "," << nu << "," << ng << "," << inversion << "," << frac << "\n";
        fp << std::setprecision(18) << t << "," << pz << "," << dpz << "," << u << "\n";
        const double ntot = nu + ng;
        double a = 0.0, b = 0.0, c = 0.5;
        if (ntot > 1e-30) { a = nu / ntot; b = ng / ntot; c = (inversion / ntot + 1.0) / 2.0; }
        const double t_fs = n * dt * 1e15;
        fh << n << "," << std::setprecision(6) << std::fixed << t_fs << "," << std::setprecision(12)
           << std::scientific << nu << "," << ng << "," << inversion << "," << std::setprecision(8) << std::fixed
           << a << "," << b << "," << c << "\n";
    }
    return 0;
}
"""


def test_tls_csv_bytes_match_iostream():
    det = importlib.import_module("3d_fdtd_version008_amd.detectors")
    rng = np.random.default_rng(11)
    rec = [
        [2.5e24 * 1e-21, 7.5e23 * 1e-21, -1.25e-30, 3.0e-12, 4.0e3],
        [0.0, 0.0, 0.0, 0.0, 0.0],
        [1.0, 2.0, 3.0, 4.0, 5.0],
        [1e-300, 3e-310, -1e300, -0.0, 1.5],
        [123456789.0, 0.1, 1.0 / 3.0, 2.0 / 3.0, 0.5 + 2 ** -40],
    ]
    rec += (rng.standard_normal((20, 5)) * 10.0 ** rng.integers(-20, 25, size=(20, 5))).tolist()
    rec = np.abs(np.array(rec))  # totals are non-negative in the detector; keep the signs in U_int, Pz
    rec[:, 2:] *= np.where(rng.random((len(rec), 3)) < 0.5, -1.0, 1.0)
    dt, every = 1.9066e-17, 10
    with tempfile.TemporaryDirectory() as t:
        src, exe = os.path.join(t, "w.cpp"), os.path.join(t, "w")
        open(src, "w").write(PROGRAM)
        subprocess.run(["g++", "-O1", "-std=c++17", src, "-o", exe], check=True)
        inp = " ".join("%.17g" % x for x in rec.ravel())
        subprocess.run([exe, "%.17g" % dt, str(every), t], input=inp, text=True, check=True)
        paths = det.write_tls_csv(t, rec, dt, every)
        for mine, ref in zip(paths, ("inv.csv", "pol.csv", "hist.csv")):
            assert open(mine, "rb").read() == open(os.path.join(t, ref), "rb").read(), ref
