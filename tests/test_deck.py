"""Input-deck subset reader (txpp .pre files) on CPU: expressions,
conditionals, macros, disabled blocks, the deck -> engine mapping, and the
reference's own melt decks when the reference tree is present."""
import os

import pytest

from polyswift_plusplus_amd import deck
from polyswift_plusplus_amd.deck import DeckError, evaluate

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
MACROS = os.path.join(DECKS, "macros")
REF = "/root/reference"


def test_expressions():
    env = {"chiNAB": 24.0, "NLEN": 100, "NZ": 1}
    assert evaluate("chiNAB/float(NLEN)", env) == 0.24
    assert evaluate("NZ == 1 and NLEN > 10", env) is True
    assert evaluate("-(2 + 3) * 2 ** 3 % 7", env) == -(2 + 3) * 2 ** 3 % 7
    assert evaluate("sqrt(16) + abs(-1) + min(3, 4)", env) == 8.0


@pytest.mark.parametrize("bad", ["__import__('os')", "NLEN.real", "open('x')", "[x for x in y]",
                                 "lambda: 1", "undefined_name + 1"])
def test_expressions_execute_nothing_else(bad):
    with pytest.raises(DeckError):
        evaluate(bad, {"NLEN": 100})


def test_diblock_fixture():
    cfg = deck.configure(deck.load(os.path.join(DECKS, "melt_diblock.pre")))
    assert cfg.shape == (16, 12, 10) and cfg.cell == (0.3, 0.25, 0.35)
    assert [(b.species, b.head, b.tail) for b in cfg.blocks] == [("A", (), (1,)), ("B", (0,), ())]
    assert cfg.blocks[0].lengthfrac == 0.30 and cfg.blocks[1].lengthfrac == pytest.approx(0.7)
    assert cfg.ds == 0.05
    assert cfg.chiN == pytest.approx(15.0, rel=1e-15)
    assert cfg.lambdas == (0.1, 0.05)           # the $ else branch (NZ != 1)
    assert cfg.nsteps == 12 and cfg.seed == 4242 and cfg.history_period == 2
    assert cfg.species_fields == ("totA", "totB")


def test_star_fixture_with_macro_file():
    root = deck.load(os.path.join(DECKS, "melt_star.pre"), search_path=[MACROS])
    cfg = deck.configure(root)
    assert cfg.shape == (16, 12, 10) and cfg.cell == (0.3, 0.3, 0.3)
    assert cfg.block_names == ["armA", "armB", "centreA"]
    assert [(b.species, b.head, b.tail, b.bSegRatio) for b in cfg.blocks] == [
        ("A", (), (1, 2), 1.0), ("B", (), (0, 2), 1.2), ("A", (0, 1), (), 1.0)]
    assert cfg.chiN == pytest.approx(14.0) and cfg.noise == 0.005


def test_disabled_block_is_dropped():
    root = deck.load(os.path.join(DECKS, "melt_diblock.pre"))
    inter = root.one("EffHamil").one("Interaction")
    assert inter.children == []


def test_missing_macro_file():
    with pytest.raises(DeckError, match="not found"):
        deck.parse("$ import nowhere\n")


@pytest.mark.parametrize("text,msg", [
    ("<Grid g>\n numCellsGlobal = [4 4 4]\n", "not closed"),
    ("$ if 1 == 1\n", "endif"),
    ("</Grid>\n", "unbalanced"),
    ("what is this\n", "cannot parse"),
])
def test_malformed(text, msg):
    with pytest.raises(DeckError, match=msg):
        deck.parse(text)


def test_unsupported_physics_raises():
    base = open(os.path.join(DECKS, "melt_diblock.pre")).read()
    with pytest.raises(DeckError, match="Solvent"):
        deck.configure(deck.parse(base + "\n<Solvent s>\n kind = simpleSolvent\n</Solvent>\n"))
    extra = base[base.index("<Polymer"):base.index("</Polymer>") + len("</Polymer>")]
    with pytest.raises(DeckError, match="1 to 8 Polymers"):
        deck.configure(deck.parse(base + "".join("\n" + extra.replace("diblock1", f"copy{i}")
                                                 for i in range(8))))
    with pytest.raises(DeckError, match="blockCopolymer only"):
        deck.configure(deck.parse(base + "\n" + extra.replace("diblock1", "d2").replace(
            "kind = blockCopolymer", "kind = polydisperse")))


def test_blend_deck():
    """A diblock with an A homopolymer of half its length (the shape of the
    reference's pstests/tests/diblock/abHomopMix.pre, with a constant chi)."""
    cfg = deck.configure(deck.load(os.path.join(DECKS, "ab_homop_mix.pre")))
    assert cfg.is_blend and len(cfg.polymers) == 2
    d, h = cfg.polymers
    assert (d.volfrac, d.length, h.volfrac, h.length) == (0.9, 100.0, 0.1, 50.0)
    assert [(b.species, b.lengthfrac, b.head, b.tail) for b in d.blocks] == [
        ("A", 0.5, (), (1,)), ("B", 0.5, (0,), ())]
    assert [(b.species, b.lengthfrac, b.head, b.tail) for b in h.blocks] == [("A", 1.0, (), ())]
    assert h.block_names == ["blockHomoA"]
    assert cfg.chiN == pytest.approx(0.12 * 100) and cfg.ds == 0.05
    single = deck.configure(deck.load(os.path.join(DECKS, "melt_diblock.pre")))
    assert not single.is_blend and len(single.polymers) == 1


# ---- the reference's own decks (read from the reference tree when present)
REF_DECKS = os.path.join(REF, "pstests", "tests", "diblock")
needs_ref = pytest.mark.skipif(not os.path.isdir(REF_DECKS), reason="reference tree not present")


@needs_ref
def test_reference_diblock_deck():
    cfg = deck.configure(deck.load(os.path.join(REF_DECKS, "diblock.pre")))
    assert cfg.shape == (64, 64, 1) and cfg.cell == (0.1, 0.1, 0.1)
    assert [(b.species, b.lengthfrac, b.head, b.tail) for b in cfg.blocks] == [
        ("A", 0.5, (), (1,)), ("B", 0.5, (0,), ())]
    assert cfg.ds == 0.05 and cfg.chiN == pytest.approx(12.0)
    assert cfg.lambdas == (0.4, 0.2) and cfg.noise == 0.002
    assert cfg.nsteps == 200 and cfg.seed == 38383


@needs_ref
def test_reference_triblock_and_star_decks():
    tri = deck.configure(deck.load(os.path.join(REF_DECKS, "triblock.pre")))
    assert [(b.species, b.lengthfrac, b.head, b.tail) for b in tri.blocks] == [
        ("A", 0.25, (), (1,)), ("B", 0.5, (0,), (2,)), ("A", 0.25, (1,), ())]
    assert tri.chiN == pytest.approx(20.0) and tri.nsteps == 400
    star = deck.configure(deck.load(os.path.join(REF_DECKS, "star3ab.pre")))
    assert [(b.species, b.head, b.tail) for b in star.blocks] == [
        ("A", (), (1, 2)), ("B", (), (0, 2)), ("A", (0, 1), ())]
    asym = deck.configure(deck.load(os.path.join(REF_DECKS, "asymmetric.pre")))
    assert asym.shape == (128, 128, 1) and asym.blocks[0].bSegRatio == 3.16227
    assert asym.chiN == pytest.approx(12.5)


@needs_ref
def test_reference_example_with_expressions():
    path = os.path.join(REF, "psexamples", "PSimBase", "linearTriblock", "linearTriblock.pre")
    cfg = deck.configure(deck.load(path))
    assert cfg.shape == (100, 100, 1)
    assert [b.lengthfrac for b in cfg.blocks] == [0.2, 0.6, 0.2]
    assert cfg.chiN == pytest.approx(24.0, rel=1e-14)


@needs_ref
@pytest.mark.parametrize("name", ["chargedAB.pre", "abSolventMix.pre", "abHomopMix.pre",
                                  "polydBulk.pre", "pspi.pre", "tri3abc.pre"])
def test_reference_decks_outside_the_subset_raise(name):
    """Decks with solvents, charges, polydispersity, chi functions (the blend
    abHomopMix.pre ramps its chi with an STFunc) or three species parse, then
    stop with a DeckError naming the feature."""
    path = os.path.join(REF_DECKS, name)
    try:
        root = deck.load(path)
    except DeckError:
        return
    with pytest.raises(DeckError):
        deck.configure(root)


@needs_ref
def test_reference_blend_deck_polymers():
    """abHomopMix.pre's two Polymers configure as a blend once its chi
    ramp (an STFunc) is replaced by a constant chi
    (pstests/tests/diblock/abHomopMix.pre:117-160)."""
    root = deck.load(os.path.join(REF_DECKS, "abHomopMix.pre"))
    inter = root.all("EffHamil")[0].all("Interaction")[0]
    with pytest.raises(DeckError, match="chi functions"):
        deck.configure(root)
    inter.children = []
    with pytest.raises(DeckError, match="chi not set"):
        deck.configure(root)
    inter.attrs["chi"] = 0.12
    cfg = deck.configure(root)
    assert cfg.chiN == pytest.approx(12.0)
    assert cfg.is_blend
    assert [(p.name, p.volfrac, p.length, len(p.blocks)) for p in cfg.polymers] == [
        ("diblock1", 0.9, 100.0, 2), ("homopolymerA", 0.1, 50.0, 1)]
    assert cfg.polymers[1].blocks[0].species == "A" and cfg.polymers[1].blocks[0].lengthfrac == 1.0


@needs_ref
def test_reference_spectral_filter_decks():
    """multispecf.pre: steepest descent, then a multiSpecFilter on both
    fields every 100 steps from step 200 (pstests/tests/diblock/multispecf.pre:
    73-103)."""
    cfg = deck.configure(deck.load(os.path.join(REF_DECKS, "multispecf.pre")))
    assert cfg.shape == (128, 128, 1) and cfg.chiN == pytest.approx(18.0)
    assert cfg.lambdas == (0.2, 0.1) and cfg.noise == 0.02
    (f,) = cfg.filters
    assert (f.fields, f.cutoffFactor, f.filterStrength, f.specCellSizes) == ("AB", 0.6, 0.0, (4, 4, 1))
    s = f.schedule.resolve(cfg.nsteps)
    assert [t for t in range(1, cfg.nsteps + 1) if s.due(t)] == [200, 300, 400]
    path = os.path.join(REF, "psexamples", "PSimBase", "spectralFilter", "spectralFilter.pre")
    ex = deck.configure(deck.load(path))
    assert len(ex.filters) == 1 and ex.filters[0].specCellSizes is not None


def test_filter_before_the_field_update():
    """A filter listed before the steepest-descent Updater in updaterSequence
    runs between the step's densities and its update (flagged ``before``);
    after it, on the updated fields."""
    macros = [os.path.join(DECKS, "macros")]
    after = deck.configure(deck.load(os.path.join(DECKS, "melt_filter.pre"), search_path=macros))
    before = deck.configure(deck.load(os.path.join(DECKS, "melt_filter_before.pre"), search_path=macros))
    (fa,), (fb,) = after.filters, before.filters
    assert not fa.before and fb.before
    s = fb.schedule.resolve(before.nsteps)
    assert [t for t in range(1, before.nsteps + 1) if s.due(t)] == [3, 7, 11]


def test_update_sequence_without_the_field_update():
    text = open(os.path.join(DECKS, "melt_filter.pre")).read().replace(
        "updaterSequence = [wAwB filt]", "updaterSequence = [filt]")
    root = deck.parse(text, search_path=[os.path.join(DECKS, "macros")], here=DECKS)
    with pytest.raises(DeckError, match="does not name the field update"):
        deck.configure(root)
