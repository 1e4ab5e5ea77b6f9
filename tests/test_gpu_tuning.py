"""GPU: tuning knobs of the bunched walk (leaf buckets, bunch and super-group
sizes, fused or two-kernel) and the autotuner.

A leaf bucket replaces a small cell by its bodies, which is exact; so with
theta = 0 (every cell opened) each setting must give the direct sum, and at
the usual theta the error against the FP64 direct sum can only shrink.
"""
import numpy as np
import pytest

import oracle as O
from hierarquicos_amd import OctreeType, direct_forces, ics

pytestmark = pytest.mark.gpu


def _tree(n, seed, method=12):
    m, q, _ = ics.plummer(n, seed=seed)
    x, y, z = q.T.copy()
    return m, x, y, z, OctreeType().pre_init(m, method).init(x, y, z)


@pytest.mark.parametrize("leaf", [0, 3, 16, 64])
@pytest.mark.parametrize("fused", [True, False])
def test_theta_zero_is_direct_for_every_bucket(leaf, fused):
    m, x, y, z, t = _tree(6000, 21)
    t.tune(leaf=leaf, fused=fused)
    Fd = O.direct_targets(m, x, y, z, np.arange(6000), 1.0, 1e-6)
    F64 = t.forces_bunched(0.0, 1.0, 1e-6, "fp64").T
    assert float(np.max(np.linalg.norm(F64 - Fd, axis=1) / np.linalg.norm(Fd, axis=1))) < 1e-12
    Fm = t.forces_bunched(0.0, 1.0, 1e-6, "mixed").T
    Fmd = direct_forces(m, x, y, z, 1.0, 1e-6, precision="mixed")
    rms, mx = O.rel_errors(Fm, Fmd)
    assert rms < 2e-6 and mx < 2e-5, (rms, mx)


def test_buckets_do_not_increase_error():
    m, x, y, z, t = _tree(40000, 22)
    sample = np.random.default_rng(1).choice(40000, 2048, replace=False)
    Fd = O.direct_targets(m, x, y, z, sample, 1.0, 1e-6)
    errs, cells, parts = {}, {}, {}
    for leaf in (0, 4, 16):
        t.tune(leaf=leaf)
        F = t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T
        errs[leaf] = O.rel_errors(F[sample], Fd)[0]
        inf = t.info()
        cells[leaf], parts[leaf] = inf["exec_cells"], inf["exec_parts"]
    # buckets trade cell interactions for body interactions
    assert cells[16] < cells[4] < cells[0] and parts[16] > parts[4] > parts[0], (cells, parts)
    assert errs[4] <= errs[0] * 1.001 and errs[16] <= errs[4] * 1.001, errs
    assert errs[0] < 1e-4


def test_fused_and_two_kernel_agree_with_buckets():
    # the two-kernel walk has no far field: compare with it switched off
    m, x, y, z, t = _tree(30000, 23)
    t.tune(leaf=6, far=0)
    a = t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T
    t.tune(fused=False)
    b = t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T
    rel = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    assert float(rel.max()) < 1e-5


def test_autotune_picks_a_valid_setting(tmp_path):
    """The autotuner's setting lies on its axes, keeps the error within 2% of
    the starting setting's on its sample (and within budget here), goes into
    the tuning table, and a second handle of the same workload takes it from
    the table at its first bunched walk without tuning; a knob its caller set
    stays as set.  The table survives a save / clear / load round trip."""
    from hierarquicos_amd import tune_table_clear, tune_table_load, tune_table_save
    tune_table_clear()
    m, x, y, z, t = _tree(100000, 24)
    sample = np.arange(0, 100000, 97)
    Fd = O.direct_targets(m, x, y, z, sample, 1.0, 1e-6)
    e_start = O.rel_errors(t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T[sample], Fd)[0]
    r = t.autotune(0.25, 1.0, 1e-6, "mixed", reps=1)
    assert r["bunch"] in (32, 48, 64) and r["super_size"] in (128, 256, 512)
    assert r["leaf"] in (0, 2, 4, 8, 16, 32, 64) and r["split"] in (0, 1, 2) and r["ms"] > 0
    for k in ("bunch", "super_size", "leaf", "split"):
        assert t.get_tune(k) == r[k]
    F = t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T
    e_tuned = O.rel_errors(F[sample], Fd)[0]
    assert e_tuned < 1e-4 and e_tuned < 1.1 * e_start, (e_tuned, e_start)
    # a second handle of the same workload: the table's setting, no tuning
    t2 = OctreeType().pre_init(m, 12).init(x, y, z)
    t2.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert [t2.get_tune(k) for k in ("bunch", "super_size", "leaf", "split")] == \
        [r["bunch"], r["super_size"], r["leaf"], r["split"]]
    # a knob the caller set is kept
    t3 = OctreeType().pre_init(m, 12).tune(leaf=5).init(x, y, z)
    t3.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert t3.get_tune("leaf") == 5 and t3.get_tune("super_size") == r["super_size"]
    # another angle is another class: defaults
    t4 = OctreeType().pre_init(m, 12).init(x, y, z)
    t4.forces_bunched(0.16, 1.0, 1e-6, "mixed")
    assert t4.get_tune("leaf") == 0 and t4.get_tune("super_size") == 256
    # save, clear, load
    path = tmp_path / "tune.txt"
    tune_table_save(path)
    tune_table_clear()
    t5 = OctreeType().pre_init(m, 12).init(x, y, z)
    t5.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert t5.get_tune("leaf") == 0 and t5.get_tune("super_size") == 256
    tune_table_load(path)
    t6 = OctreeType().pre_init(m, 12).init(x, y, z)
    t6.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert [t6.get_tune(k) for k in ("bunch", "super_size", "leaf", "split")] == \
        [r["bunch"], r["super_size"], r["leaf"], r["split"]]
    tune_table_clear()


def test_tune_table_rejects_bad_rows(tmp_path):
    from hierarquicos_amd import HQError, tune_table_load
    path = tmp_path / "bad.txt"
    path.write_text("17 0 8 50 1 18 64 128 4\n")
    with pytest.raises(HQError):
        tune_table_load(path)


@pytest.mark.parametrize("method,leaf", [(10, 0), (11, 4), (12, 2)])
def test_split_walk_bitidentical_to_fused(method, leaf):
    """The split evaluation (walk records its batches; one wavefront per batch;
    partials summed in order) gives the fused kernel's forces bit for bit."""
    m, q, _ = ics.plummer(20000, seed=5)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    ta = OctreeType().pre_init(m, method, precision="mixed").tune(leaf=leaf, split=0)
    tb = OctreeType().pre_init(m, method, precision="mixed").tune(leaf=leaf, split=2)
    ta.init(x, y, z)
    tb.init(x, y, z)
    Fa = ta.forces_bunched(0.25, 1.0, 1e-4, "mixed")
    Fb1 = tb.forces_bunched(0.25, 1.0, 1e-4, "mixed")   # measures the lists (fused)
    Fb2 = tb.forces_bunched(0.25, 1.0, 1e-4, "mixed")   # split
    assert ta.info()["split_walks"] == 0 and tb.info()["split_walks"] == 1
    assert np.array_equal(Fa, Fb1) and np.array_equal(Fa, Fb2)
    assert tb.info()["exec_cells"] == ta.info()["exec_cells"]


def _far_rows(method, theta2, tiers, seed=25, n=200000):
    """(far, inner, RMS error, RMS difference from the walk without a far
    field, RMS error without it, far cells) per tier setting, the errors
    against the FP64 direct sum on a sample of targets."""
    m, x, y, z, _ = _tree(n, seed)
    sample = np.arange(0, n, 101)
    Fd = O.direct_targets(m, x, y, z, sample, 1.0, 1e-6)
    t = OctreeType().pre_init(m, method, precision="mixed").init(x, y, z)
    t.tune(far=0)
    Fn = t.forces_bunched(theta2, 1.0, 1e-6, "mixed").T
    assert t.info()["far_cells"] == 0
    e_near, _ = O.rel_errors(Fn[sample], Fd)
    rows = []
    for far, inner in tiers:
        t.tune(far=far, far_inner=inner)
        Ff = t.forces_bunched(theta2, 1.0, 1e-6, "mixed").T
        e_far, _ = O.rel_errors(Ff[sample], Fd)
        d, _ = O.rel_errors(Ff[sample], Fn[sample])
        rows.append((far, inner, e_far, d, e_near, t.info()["far_cells"]))
    print(method, theta2, rows)
    return rows


@pytest.mark.parametrize("method", [10, 11, 12])
def test_far_field_error_budget(method):
    """Far cells through each bunch's local expansion against the same walk
    with every cell evaluated per body, at theta = 0.5: the outer tier alone
    (order 5 beyond 5 bunch radii), both tiers (the default: order 6 from 3
    radii) and the inner tier alone (order 6 from 3 radii out).  The outer
    tier adds a small part of the multipole error; the inner one, whose
    truncation at 3 radii is about 4e-5 RMS whatever the multipole order,
    keeps the octupole walk within the 1e-4 budget against the FP64 direct
    sum."""
    rows = _far_rows(method, 0.25, [(5.0, 0.0), (5.0, 3.0), (100.0, 3.0)])
    (_, _, e5, d5, e0, c5), (_, _, e53, d53, _, c53), (_, _, e3, d3, _, c3) = rows
    assert d5 < 0.35 * e0 and e5 < 1.05 * e0 + 1e-6, rows[0]
    for e, d in ((e53, d53), (e3, d3)):
        assert d < 0.8 * e0 and d < 5e-5 and e < 1.3 * e0, rows
    if method == 12:
        assert e53 < 1e-4, rows
    # the inner tier takes the cells between 3 and 5 radii too; the tiers
    # only change the orders the same cells are expanded to
    assert c53 > c5 > 0 and c3 == c53, rows


def test_far_field_tiers_at_a_wider_angle():
    """theta^2 = 0.5: the multipole error grows faster than the expansions'
    truncation, and both tiers add a small part of it."""
    rows = _far_rows(12, 0.5, [(5.0, 0.0), (5.0, 3.0)])
    for _, _, e, d, e0, c in rows:
        assert c > 0 and d < 0.35 * e0 and e < 1.05 * e0, rows


def test_fused_walk_stack_retries():
    """The fused walk first runs with a short LDS stack; when a bunch
    overflows it, the walk runs again with the long stack (the same bits as
    the split walk, which always has the long one) and the handle starts
    there next time.  When the long stack overflows too, the deep stack
    (4096 entries) takes the walk, far field included, and the handle starts
    there until a walk needs well under the long one."""
    m, q, _ = ics.plummer(20000, seed=5)
    x, y, z = (np.ascontiguousarray(q[:, c]) for c in range(3))
    ta = OctreeType().pre_init(m, 12, precision="mixed").tune(split=0).init(x, y, z)
    tb = OctreeType().pre_init(m, 12, precision="mixed").tune(split=2).init(x, y, z)
    Fa1 = ta.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert ta.info()["max_stack"] > 1152 and ta.info()["far_cells"] > 0   # needed the long stack
    Fa2 = ta.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    tb.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    Fb = tb.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert tb.info()["split_walks"] == 1
    assert np.array_equal(Fa1, Fb) and np.array_equal(Fa2, Fb)
    assert ta.info()["walk_stack"] == 1536 and tb.info()["walk_stack"] == 1536
    # deeper than either: the deep stack, within the budget
    m, x, y, z, t = _tree(200000, 25)
    F = t.forces_bunched(0.1, 1.0, 1e-6, "mixed").T
    inf = t.info()
    assert inf["far_cells"] > 0 and inf["walk_stack"] == 4096 and inf["max_stack"] > 1536
    sample = np.arange(0, 200000, 101)
    Fd = O.direct_targets(m, x, y, z, sample, 1.0, 1e-6)
    assert O.rel_errors(F[sample], Fd)[0] < 2e-5
    # the handle starts with the deep stack next time (the same forces), and
    # steps back down at a wider angle
    assert np.array_equal(t.forces_bunched(0.1, 1.0, 1e-6, "mixed").T, F)
    assert t.info()["walk_stack"] == 4096
    t.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    t.forces_bunched(0.25, 1.0, 1e-6, "mixed")
    assert t.info()["walk_stack"] < 4096
    # the two-kernel walk (no far field) stays within the budget too
    t.tune(fused=False)
    F2 = t.forces_bunched(0.1, 1.0, 1e-6, "mixed").T
    assert t.info()["far_cells"] == 0 and t.info()["walk_stack"] == 0
    assert O.rel_errors(F2[sample], Fd)[0] < 2e-5


def test_headline_settings_within_budget():
    """The settings of the 1M headline (super-groups of 128, outer tier from 5
    and inner tier from 3 bunch radii, leaf bucket 4) keep the octupole walk
    within the 1e-4 RMS budget against the FP64 direct sum, here at 2e5."""
    m, x, y, z, _ = _tree(200000, 25)
    t = OctreeType().pre_init(m, 12, precision="mixed")
    t.tune(super_size=128, far=5.0, far_inner=3.0, leaf=4)
    t.init(x, y, z)
    F = t.forces_bunched(0.25, 1.0, 1e-6, "mixed").T
    assert t.info()["far_cells"] > 0
    sample = np.arange(0, 200000, 101)
    Fd = O.direct_targets(m, x, y, z, sample, 1.0, 1e-6)
    rms, _ = O.rel_errors(F[sample], Fd)
    assert rms < 1e-4, rms
