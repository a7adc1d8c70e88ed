"""Checks of the oracle against what the mathematics fixes, independent of the
reference: the Courant time step, closed-form CPML profiles and source waveforms, the explicit-Euler
free decay of the upper population, discrete div H = 0 of the Yee update,
and the DFT monitor (which has no reference counterpart) against a direct
sum over a probe trace and against the closed-form DFT of a sampled
sinusoid.  Each is chosen so a dropped term, a wrong sign or index fails."""
import math

import numpy as np
import pytest

C0 = 299792458.0
EPS0 = 8.854187817e-12
MU0 = 4.0 * math.pi * 1e-7


def test_cpml_profiles_closed_form(pkg, oracle_mod):
    sp = pkg.scene.uniform_sphere_case(core=10, npml=6, steps=1)
    sp = sp.copy(cpml_m=3.0, cpml_kappa_max=7.0, cpml_alpha0=0.1, cpml_rerr=1e-9)
    o = oracle_mod.CpuSim(sp)
    N, npml, d, dt = o.shape[0], sp.npml, sp.dx_uniform, o.dt
    smax = (3.0 + 1.0) * math.log(1.0 / 1e-9) * EPS0 * C0 / (2.0 * npml * d)
    kx, bE, cE, bH, cH = (o.array(n) for n in ("kx", "cpml_bEx", "cpml_cEx", "cpml_bHx", "cpml_cHx"))
    for n in range(N):
        dd = min(n, N - 1 - n)
        if dd >= npml:
            assert kx[n] == 1.0 and bE[n] == 1.0 and cE[n] == 0.0
            continue
        r = (npml - (dd + 0.5)) / npml
        kap = 1 + 6.0 * r ** 3
        sig = smax * kap * r ** 3
        alf = 0.1 * (C0 / d) * (1 - r)
        for norm, b, c in ((EPS0, bE, cE), (MU0, bH, cH)):
            sn = sig / norm
            be = math.exp(-(sn / kap + alf) * dt)
            ce = sn * (be - 1) / (kap * (sn + kap * alf))
            assert b[n] == pytest.approx(be, rel=1e-13)
            assert c[n] == pytest.approx(ce, rel=1e-12)
        assert kx[n] == pytest.approx(kap, rel=1e-14)


def test_waveforms_closed_form(pkg, oracle_mod):
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import cases
    for sp in (cases.pec_cavity(), cases.plane_waves()):
        o = oracle_mod.CpuSim(sp)
        dt = o.dt
        for q in range(o.info()[7]):
            v = o.source(q, 0)
            f0, tau, t0, I0 = v[15], v[16], v[17], v[18]
            kind = int(v[0])
            wf = (sp.dipoles + sp.planes)[q].waveform
            for n in (0, 7, 55, 130):
                t = (n + 0.5) * dt
                if wf == 0:
                    a = math.pi * f0 * (t - t0)
                    w = I0 * (1 - 2 * a * a) * math.exp(-a * a)
                elif wf == 1:
                    w = I0 * math.sin(2 * math.pi * f0 * (t - t0)) * math.exp(-((t - t0) ** 2) / tau ** 2)
                elif wf == 2:
                    x = (t - t0) / tau
                    w = I0 * (1 - 2 * x * x) * math.exp(-x * x)
                else:
                    env = 0.5 * (1 - math.cos(math.pi * t / tau)) if t < tau else 1.0
                    w = I0 * env * math.sin(2 * math.pi * f0 * t)
                val = o.source(q, n)[14]
                if kind == 0:
                    expect = w / v[6]
                else:
                    pw = sp.planes[q - len(sp.dipoles)]
                    dirn = pw.direction
                    dspace = {0: sp.dx_uniform, 1: sp.dx_uniform, 2: sp.dy_uniform, 3: sp.dy_uniform}.get(dirn, sp.dz_uniform)
                    sign = 1.0 if dirn in (0, 2, 4) else -1.0
                    expect = 2 * w / (376.730313668 * dspace) * sign
                assert val == pytest.approx(expect, rel=1e-12, abs=1e-300)


def test_population_free_decay(pkg, oracle_mod):
    """No source: Ez stays 0, P stays 0, and Nu_n = Nu_0 (1 - dt/tau)^n."""
    S = pkg.scene
    sp = S.SceneSpec(nx_core=8, ny_core=8, nz_core=8, npml=4, dx_uniform=10e-9, dy_uniform=10e-9, dz_uniform=10e-9,
                     structures=[S.Structure(S.SHAPE_BOX, 1.5, (10e-9, 70e-9, 10e-9, 70e-9, 10e-9, 70e-9), 0)],
                     tls_materials=[S.TLSMaterial(1.0e-6, 1e13, 3e-15, 1e24, 0.7)], n_steps=200)
    o = oracle_mod.CpuSim(sp)
    nu0 = o.array("Nu").max()
    o.step(200)
    nu = o.array("Nu")
    ng = o.array("Ng")
    mask = o.array("Ndip") > 0
    assert nu0 == 0.7 * 1e24
    expect = nu0 * (1 - o.dt / 3e-15) ** 200
    assert np.allclose(nu[mask], expect, rtol=1e-12, atol=0)
    assert np.allclose(nu[mask] + ng[mask], 1e24, rtol=1e-13)
    assert np.abs(o.array("Pz")).max() == 0.0


def test_divergence_free_H(pkg, oracle_mod):
    """Away from the CPML the Yee H update keeps the discrete div H at roundoff."""
    sp = pkg.scene.uniform_sphere_case(core=20, npml=6, steps=120)
    o = oracle_mod.CpuSim(sp)
    o.step(120)
    Hx, Hy, Hz = o.array("Hx"), o.array("Hy"), o.array("Hz")
    # H lives on cell faces (Hx at i, j+1/2, k+1/2): div H uses forward differences
    s, u = slice(8, -8), slice(9, -7)
    div = ((Hx[u, s, s] - Hx[s, s, s]) + (Hy[s, u, s] - Hy[s, s, s]) + (Hz[s, s, u] - Hz[s, s, s]))
    scale = max(np.abs(Hx).max(), np.abs(Hy).max(), np.abs(Hz).max())
    assert scale > 0
    assert np.abs(div).max() <= 1e-12 * scale


def test_dft_matches_direct_sum_of_probe_trace(pkg, oracle_mod):
    """A 1-cell DFT monitor equals sum_n F(t_n) e^{-i w t_n} dt over the Ez
    probe trace of the same cell (t_n = (n+1) dt)."""
    sp = pkg.scene.uniform_sphere_case(core=12, npml=5, steps=300)
    o0 = oracle_mod.CpuSim(sp)
    ijk, _ = o0.probe_index(1)
    freqs = (C0 / 0.9e-6, C0 / 1.0e-6, C0 / 1.3e-6)
    sp.dfts = [pkg.scene.DFTMonitor(2, (ijk[0], ijk[0] + 1, ijk[1], ijk[1] + 1, ijk[2], ijk[2] + 1), freqs),
               pkg.scene.DFTMonitor(4, (ijk[0], ijk[0] + 1, ijk[1], ijk[1] + 1, ijk[2], ijk[2] + 1), freqs)]
    sp.probes[1] = pkg.scene.Probe(sp.probes[1].x, sp.probes[1].y, sp.probes[1].z, (2, 4))
    o = oracle_mod.CpuSim(sp)
    o.step(sp.n_steps)
    dt = o.dt
    n = np.arange(sp.n_steps)
    for m, (comp, toff) in enumerate(((0, 1.0), (1, 0.5))):
        trace = o.probe(1, comp)
        acc = o.dft(m).reshape(len(freqs), 1, 2)
        for f_i, f in enumerate(freqs):
            w = 2 * math.pi * f
            t = (n + toff) * dt
            X = np.sum(trace * np.exp(-1j * w * t) * dt)
            got = acc[f_i, 0, 0] + 1j * acc[f_i, 0, 1]
            assert abs(got - X) <= 1e-12 * max(abs(X), np.abs(trace).sum() * dt)


def test_dft_of_sampled_sinusoid_closed_form():
    """The accumulation rule sum cos(w t)dt, -sin(w t)dt applied to a sampled
    sinusoid reproduces the closed-form geometric sum (pins the sign of the
    imaginary part and the time base used by the monitors)."""
    dt, f0, f, N = 1e-17, 2.9e14, 3.1e14, 4000
    n = np.arange(N)
    t = (n + 1.0) * dt
    F = np.cos(2 * math.pi * f0 * t)
    w = 2 * math.pi * f
    acc = complex(np.sum(F * np.cos(w * t) * dt), np.sum(F * -np.sin(w * t) * dt))

    def geo(om):
        z = np.exp(1j * om * dt)
        return dt * z * (1 - z ** N) / (1 - z)

    closed = 0.5 * (geo(2 * math.pi * (f0 - f)) + geo(-2 * math.pi * (f0 + f)))
    assert abs(acc - closed) <= 1e-9 * abs(closed)


def test_dt_closed_form(pkg, oracle_mod):
    """make_context's time step (src/params.hpp:264-265) is the 3-D Courant
    limit scaled by S over the smallest spacing of each axis: on a uniform
    cube dt = S d / (c0 sqrt 3); on an anisotropic uniform grid and on the
    graded z axis it is S / (c0 sqrt(sum 1/d_min^2)), with d_min taken from
    the scene's parameters (not from the spacing arrays the setup builds).
    The product's setup, the oracle and the compiled reference harness must
    all give it; a wrong axis, a max in place of the min or a dropped square
    moves dt by far more than the 4-ulp bar."""
    import dataclasses

    d = 20e-9
    cube = pkg.scene.uniform_sphere_case(core=16, npml=4, d=d, steps=2)
    aniso = dataclasses.replace(cube, dx_uniform=10e-9, dy_uniform=20e-9, dz_uniform=35e-9)
    graded = pkg.scene.gain_slab_case(n_total=64, npml=8, steps=2, lateral=20e-9, fine=5e-9, coarse=20e-9,
                                      fine_len=100e-9)
    cases = [
        (cube, cube.S * d / (C0 * math.sqrt(3.0))),
        (aniso, aniso.S / (C0 * math.sqrt(1 / 10e-9 ** 2 + 1 / 20e-9 ** 2 + 1 / 35e-9 ** 2))),
        (graded, graded.S / (C0 * math.sqrt(2 / 20e-9 ** 2 + 1 / 5e-9 ** 2))),
    ]
    impls = ["oracle"] + (["ref"] if oracle_mod.available("ref") else [])
    for sp, want in cases:
        got = {"setup": pkg.Setup(sp).dt}
        for impl in impls:
            got[impl] = oracle_mod.CpuSim(sp, impl).dt
        for who, v in got.items():
            assert abs(v - want) <= 4 * math.ulp(want), (sp.name, who, v, want)
        # and below the 1-D limit of the finest axis
        assert max(got.values()) < min(sp.dx_uniform, sp.dy_uniform, sp.dz_uniform, 5e-9 if sp is graded else 1.0) / C0
