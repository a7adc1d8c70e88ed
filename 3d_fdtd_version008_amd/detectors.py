"""Detector output files in the reference's formats, written from the
device's records (so the reference's Py_Visualization scripts read them
unchanged).

TLS diagnostics (records from Scene.tls_records, one every save_every steps
from step 0):

- <root>/populations/inversion_time.csv, PopulationTimeSeriesDetector
  (src/two_level_detectors.hpp:17-109): header
  "t, total_Nu, total_Ng, inversion, inv_fraction", rows written with
  std::setprecision(18) (the %.18g conversion);
- <root>/polarization/polarization_time.csv, PolarizationDiagnostics
  (src/two_level_detectors.hpp:205-263): header
  "t, Pz_center, dPz_dt_center, total_P_energy", %.18g;
- <root>/tls_global_history.csv, TLSGlobalHistory
  (src/two_level_detectors.hpp:265-335): header
  "step,t_fs,Nu_total,Ng_total,inversion_total,Nu_percent,Ng_percent,inv_norm",
  the step, t in fs with 6 decimals, totals in scientific notation with 12
  digits, fractions with 8 decimals.

Each value is derived from the record exactly as the reference derives it
from its own sums (inversion = Nu - Ng, fraction = inversion / (Nu + Ng)
when the sum exceeds 1e-30, and so on).
"""
from __future__ import annotations

import os
from typing import Sequence

import numpy as np


def _g18(x: float) -> str:
    return "%.18g" % x


def write_tls_csv(root: str, records: np.ndarray, dt: float, save_every: int) -> Sequence[str]:
    """Writes the three TLS CSV files under `root`; returns their paths."""
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 5)
    pop_dir = os.path.join(root, "populations")
    pol_dir = os.path.join(root, "polarization")
    os.makedirs(pop_dir, exist_ok=True)
    os.makedirs(pol_dir, exist_ok=True)
    p_inv = os.path.join(pop_dir, "inversion_time.csv")
    p_pol = os.path.join(pol_dir, "polarization_time.csv")
    p_hist = os.path.join(root, "tls_global_history.csv")
    with open(p_inv, "w") as fi, open(p_pol, "w") as fp, open(p_hist, "w") as fh:
        fi.write("t, total_Nu, total_Ng, inversion, inv_fraction\n")
        fp.write("t, Pz_center, dPz_dt_center, total_P_energy\n")
        fh.write("step,t_fs,Nu_total,Ng_total,inversion_total,Nu_percent,Ng_percent,inv_norm\n")
        for r, (nu, ng, u_int, pz, dpz) in enumerate(rec):
            n = r * save_every
            t = n * dt
            inversion = nu - ng
            denom = nu + ng
            frac = inversion / denom if denom > 1e-30 else 0.0
            fi.write(f"{_g18(t)},{_g18(nu)},{_g18(ng)},{_g18(inversion)},{_g18(frac)}\n")
            fp.write(f"{_g18(t)},{_g18(pz)},{_g18(dpz)},{_g18(u_int)}\n")
            ntot = nu + ng
            nu_pc, ng_pc, inv_norm = 0.0, 0.0, 0.5
            if ntot > 1e-30:
                nu_pc = nu / ntot
                ng_pc = ng / ntot
                inv_norm = (inversion / ntot + 1.0) / 2.0
            t_fs = n * dt * 1e15
            fh.write(f"{n},{t_fs:.6f},{nu:.12e},{ng:.12e},{inversion:.12e},{nu_pc:.8f},{ng_pc:.8f},{inv_norm:.8f}\n")
    return [p_inv, p_pol, p_hist]
