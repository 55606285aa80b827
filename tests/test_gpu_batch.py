"""Batched parameter sweeps: every instance of a batch follows the same
trajectory as the same system run alone."""
import numpy as np
import pytest

from polyswift_plusplus_amd import BatchSCFT, DiblockSCFT, Mixing, fields

pytestmark = pytest.mark.gpu

SPECS = [dict(shape=(16, 16, 16), cell=(0.3,) * 3, fA=0.30, ds=0.05, chiN=14.0),
         dict(shape=(16, 16, 16), cell=(0.3,) * 3, fA=0.45, ds=0.05, chiN=20.0),
         dict(shape=(32, 16, 8), cell=(0.25, 0.3, 0.35), fA=0.35, ds=0.05, chiN=17.0),
         dict(shape=(12, 12, 12), cell=(0.4,) * 3, fA=0.40, ds=0.05, chiN=12.0),
         dict(shape=(16, 16, 16), cell=(0.28,) * 3, fA=0.25, ds=0.05, chiN=25.0)]
MIX = Mixing("anderson", lambda0=0.2, depth=5, beta=1.0, anderson_start=0.05)


def _fields(i, sp):
    return fields.seeded_fields(sp["shape"], sp["cell"], sigma=0.3, seed=100 + i, seed_kind="lamellar")


def _single(sp, i, prec, n, fields_out=False):
    e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"], prec=prec)
    e.setFields(*_fields(i, sp))
    h = e.iterate(n, MIX)
    if fields_out:
        h["fields"] = e.getFields()
    e.close()
    return h


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_matches_single_instances(prec, use_graph, monkeypatch):
    """Like for like: the batch runs its quadrature after the sweeps, so the
    single-instance references do too (PSX_OVERLAP_QUAD=0, the default).  The
    batch's four streams change no arithmetic: F and the fields agree bit for
    bit, across two iterate calls, with and without the graph."""
    print(f"batch {prec} graph={use_graph}", flush=True)
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "0")
    b = BatchSCFT(SPECS, prec=prec)
    assert b.nclasses == 3
    for i, sp in enumerate(SPECS):
        b.instances[i].setFields(*_fields(i, sp))
    h1 = b.iterate(3, MIX, use_graph=use_graph)
    h2 = b.iterate(5, MIX, use_graph=use_graph)
    for i, sp in enumerate(SPECS):
        ref = _single(sp, i, prec, 8, fields_out=True)
        F = np.r_[h1["F"][i], h2["F"][i]]
        assert np.max(np.abs(F - ref["F"])) <= 1e-14 * np.max(np.abs(ref["F"])), i
        for got, want in zip(b.instances[i].getFields(), ref["fields"]):
            assert np.array_equal(got, want), i
    b.close()


MANY = [dict(shape=s, cell=(0.3,) * 3, fA=0.25 + 0.05 * (i % 4), ds=0.05, chiN=12.0 + i)
        for i, s in enumerate([(8, 8, 8), (12, 12, 12), (16, 16, 16), (10, 10, 10), (16, 8, 8),
                               (24, 12, 6), (18, 18, 18), (6, 6, 6), (20, 20, 20)])]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_of_nine_sizes_in_groups(use_graph, monkeypatch):
    """Nine grid sizes are nine size classes, run as three groups of at most
    four (one graph each): every instance still follows its single-engine
    trajectory bit for bit."""
    print(f"batch of 9 sizes graph={use_graph}", flush=True)
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "0")
    b = BatchSCFT(MANY, prec="fp64")
    assert b.nclasses == 9
    for i, sp in enumerate(MANY):
        b.instances[i].setFields(*_fields(i, sp))
    h1 = b.iterate(2, MIX, use_graph=use_graph)
    h2 = b.iterate(3, MIX, use_graph=use_graph)
    for i, sp in enumerate(MANY):
        ref = _single(sp, i, "fp64", 5, fields_out=True)
        F = np.r_[h1["F"][i], h2["F"][i]]
        assert np.max(np.abs(F - ref["F"])) <= 1e-14 * np.max(np.abs(ref["F"])), i
        for got, want in zip(b.instances[i].getFields(), ref["fields"]):
            assert np.array_equal(got, want), i
    b.close()


def test_batch_recaptures_after_engine_buffers_change(monkeypatch):
    """A deeper Anderson history on the second call reallocates every
    instance's ring: the batch graph is captured again, not replayed with the
    old pointers."""
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "0")
    deep = Mixing("anderson", lambda0=0.2, depth=9, beta=1.0, anderson_start=0.05)
    b = BatchSCFT(SPECS, prec="fp64")
    for i, sp in enumerate(SPECS):
        b.instances[i].setFields(*_fields(i, sp))
    h1 = b.iterate(3, MIX)
    h2 = b.iterate(4, deep)
    for i, sp in enumerate(SPECS):
        e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"])
        e.setFields(*_fields(i, sp))
        r1 = e.iterate(3, MIX)
        r2 = e.iterate(4, deep)
        F, ref = np.r_[h1["F"][i], h2["F"][i]], np.r_[r1["F"], r2["F"]]
        assert np.max(np.abs(F - ref)) <= 1e-14 * np.max(np.abs(ref)), i
        e.close()
    b.close()


def test_batch_autotune_keeps_results():
    b = BatchSCFT(SPECS, prec="fp64")
    for i, sp in enumerate(SPECS):
        b.instances[i].setFields(*_fields(i, sp))
    choice = b.autotune(3)
    print("autotune choices per size class:", choice, flush=True)
    h = b.iterate(6, MIX)
    for i, sp in enumerate(SPECS):
        ref = _single(sp, i, "fp64", 6)
        assert np.max(np.abs(h["F"][i] - ref["F"])) <= 1e-11 * np.max(np.abs(ref["F"])), i
    b.close()


def test_batch_rejects_mixed_step_counts():
    from polyswift_plusplus_amd import PsxError
    specs = [dict(SPECS[0]), dict(SPECS[1], ds=0.025)]
    with pytest.raises(PsxError):
        BatchSCFT(specs)


def test_batch_instance_precision_switch_refused():
    """An instance cannot change precision on its own: the class launches
    write the batch precision's slices, which the instance's quadrature must
    read.  The batch switches all of them together."""
    print("batch instance precision switch", flush=True)
    from polyswift_plusplus_amd import PsxError
    b = BatchSCFT(SPECS[:2], prec="fp32")
    try:
        with pytest.raises(PsxError) as ex:
            b.instances[0].setPrecision("fp64")
        assert ex.value.status == 7
        b.instances[0].setPrecision("fp32")      # no change: accepted
        b.setPrecision("fp64")
        for i, sp in enumerate(SPECS[:2]):
            b.instances[i].setFields(*_fields(i, sp))
        h = b.iterate(2, MIX)
        for i, sp in enumerate(SPECS[:2]):
            ref = _single(sp, i, "fp64", 2)
            assert np.max(np.abs(h["F"][i] - ref["F"])) <= 1e-14 * np.max(np.abs(ref["F"]))
    finally:
        b.close()
