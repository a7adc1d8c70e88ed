"""Batches: many domains stepped by one persistent launch per phase.

A domain's fields and probe traces are identical (bitwise, every storage
format) whether it runs alone or in a batch, under every tiling, in graph
replay and when other members join or leave between calls; DFT accumulators
(whose phases the batch computes on the device) agree to 1e-12 relative L2.
FP64 batch members match the CPU oracle to 1e-12."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_specs(pkg, n, steps, seed=7):
    """Batch-workload domains at test sizes (axes 32..48 cells)."""
    return [pkg.scene.batch_case(i, seed=seed, steps=steps, lo=32, hi=48) for i in range(n)]


def _state(sc, setup):
    st = {f: sc.field(f) for f in FIELDS}
    if setup.tls_on:
        st["Nu"] = sc.field("Nu")
        st["Pz"] = sc.field("Pz")
    for c in range(setup.n_channels):
        st[f"probe{c}"] = sc.probe(c)
    return st


def _dfts(sc, setup):
    return [sc.dft(m) for m in range(setup.n_dfts)]


def _alone(pkg, setups, precision, steps):
    res = []
    for st in setups:
        sc = pkg.Scene(st, precision)
        sc.set_kernel_mode(0)
        sc.step(steps)
        res.append((_state(sc, st), _dfts(sc, st)))
        sc.close()
    return res


def _check_same(ref, got, tag):
    for (sa, da), (sb, db) in zip(ref, got):
        diff = [k for k in sa if not np.array_equal(sa[k], sb[k])]
        assert not diff, f"{tag}: differs in {diff}"
        for x, y in zip(da, db):
            assert rel_l2(y, x) <= 1e-12, f"{tag}: DFT {rel_l2(y, x)}"


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_batch_offset_widths_agree(pkg, precision):
    """The batch passes address cells by 32-bit byte offsets when every
    member fits, by 64-bit element offsets when a member asks for them
    (mode 0, + 131072): the same bits either way."""
    specs = _small_specs(pkg, 4, steps=60, seed=11)
    setups = [pkg.Setup(sp) for sp in specs]
    runs = []
    for force64 in (False, True):
        _say(f"[batch offsets] {precision} 64-bit={force64}")
        b = pkg.Batch(precision)
        scs = [pkg.Scene(st, precision) for st in setups]
        if force64:
            scs[0].set_kernel_mode(0, 0x20000 + 32)
        for sc in scs:
            b.add(sc)
        b.step(60)
        runs.append([(_state(sc, st), _dfts(sc, st)) for sc, st in zip(scs, setups)])
        b.close()
        for sc in scs:
            sc.close()
    _check_same(runs[0], runs[1], f"{precision} offsets")


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp16"])
def test_batch_equals_alone(pkg, precision):
    specs = _small_specs(pkg, 6, steps=80)
    setups = [pkg.Setup(sp) for sp in specs]
    _say(f"[batch] {precision} domains={[st.shape for st in setups]}")
    ref = _alone(pkg, setups, precision, 80)
    # (rows, work-groups per CU, tiles per counter increment, planes); chunk 0:
    # one work-group per tile instead of persistent work-groups
    for tuning, graph in [((4, 4, 4, 1), False), ((1, 2, 1, 3), False), ((2, 8, 16, 8), True), ((4, 4, 0, 1), False),
                          ((1, 4, 0, 3), True)]:
        _say(f"  tuning={tuning} graph={graph}")
        b = pkg.Batch(precision)
        scs = [pkg.Scene(st, precision) for st in setups]
        for sc in scs:
            b.add(sc)
        b.set_tuning(*tuning)
        b.set_graph(graph)
        for _ in range(4):
            b.step(20)  # graph mode: captured once, replayed three times
        got = [(_state(sc, st), _dfts(sc, st)) for sc, st in zip(scs, setups)]
        _check_same(ref, got, f"{precision} {tuning} graph={graph}")
        b.close()
        for sc in scs:
            sc.close()


def test_batch_membership_changes(pkg):
    """Members join and leave between calls (continuous batching): every
    domain still ends bitwise equal to its solo run."""
    specs = _small_specs(pkg, 4, steps=60, seed=11)
    setups = [pkg.Setup(sp) for sp in specs]
    ref = _alone(pkg, setups, "fp32", 60)
    b = pkg.Batch("fp32")
    scs = [pkg.Scene(st, "fp32") for st in setups]
    b.add(scs[0])
    b.step(20)
    b.add(scs[1])
    b.add(scs[2])
    b.step(20)
    b.remove(scs[0])
    b.add(scs[3])
    b.step(20)   # scs[0] at 40 steps, 1 and 2 at 40, 3 at 20
    b.add(scs[0])
    b.step(20)   # 0, 1, 2 done (60); 3 at 40
    b.remove(scs[0]); b.remove(scs[1]); b.remove(scs[2])
    b.step(20)   # 3 done
    got = [(_state(sc, st), _dfts(sc, st)) for sc, st in zip(scs, setups)]
    _check_same(ref, got, "membership")
    b.close()


def test_batch_fp64_matches_oracle(pkg, oracle_mod):
    specs = _small_specs(pkg, 3, steps=120, seed=3)
    setups = [pkg.Setup(sp) for sp in specs]
    b = pkg.Batch("fp64")
    scs = [pkg.Scene(st, "fp64") for st in setups]
    for sc in scs:
        b.add(sc)
    b.step(120)
    for sp, st, sc in zip(specs, setups, scs):
        ref = oracle_mod.CpuSim(sp, "oracle")
        ref.step(120)
        worst = max(rel_l2(sc.field(f), ref.array(f)) for f in FIELDS)
        _say(f"  {sp.name}: worst field rel L2 {worst:.2e}")
        assert worst <= 1e-12
        ref.close()


def test_batch_autotune_keeps_results(pkg):
    specs = _small_specs(pkg, 5, steps=0, seed=5)
    setups = [pkg.Setup(sp) for sp in specs]
    b = pkg.Batch("fp32")
    scs = [pkg.Scene(st, "fp32") for st in setups]
    for sc in scs:
        b.add(sc)
    tun, taken = b.autotune(trial_steps=2)
    _say(f"  autotune -> rows, wg/cu, chunk, planes = {tun} after {taken} steps")
    # a warm-up call, 24 persistent tilings and 6 with one work-group per tile
    assert taken == 2 * (1 + 24 + 6) and tun == b.tuning()
    b.step(4)
    total = taken + 4
    ref = _alone(pkg, setups, "fp32", total)
    got = [(_state(sc, st), _dfts(sc, st)) for sc, st in zip(scs, setups)]
    _check_same(ref, got, "autotune")
    # a second batch with the same key reuses the winner without trials
    b2 = pkg.Batch("fp32")
    scs2 = [pkg.Scene(st, "fp32") for st in setups]
    for sc in scs2:
        b2.add(sc)
    tun2, taken2 = b2.autotune(trial_steps=2)
    assert taken2 == 0 and tun2 == tun
