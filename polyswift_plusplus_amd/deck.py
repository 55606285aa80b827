"""Reader for the subset of PolySwift++ input decks (``.pre``, txpp syntax)
that describes a bulk block-copolymer melt, and a runner that maps such a
deck onto the engine.

The txpp subset (the forms the reference's decks use, e.g.
pstests/tests/diblock/diblock.pre:1-146, psexamples/PSimBase/linearTriblock):

* ``# ...`` comments;
* ``$ NAME = expr`` variables, with Python-style arithmetic (``chiNAB/float(NLEN)``);
* ``$ if cond`` / ``$ elif cond`` / ``$ else`` / ``$ endif``;
* ``$ import NAME``: macro definitions ``<macro name(a, b)> ... </macro>``
  from ``NAME.mac`` on the search path; a call ``name(x, y)`` on its own line
  expands the body with the arguments substituted;
* nested blocks ``<Kind name> ... </Kind>``; ``<Off Kind name>`` disables one;
* ``key = value`` with numbers, names, strings, variable references and
  ``[a b c]`` lists.

The one macro the melt decks need, ``setupPS(nx, ny, nz, dx, dy, dz, debug)``
of psutils/macros/pseudoSpecSetup.mac:32-99, is built in (it declares the
uniform grid ``mainGrid``, numCellsGlobal = [nx ny nz], cellSizes =
[dx dy dz]); a ``pseudoSpecSetup.mac`` found on the search path is used
instead.

Mapping to the engine (``configure``):

* Grid ``numCellsGlobal`` / ``cellSizes`` (psgrid/PsUniCartGrid.cpp:30-43);
* the single ``blockCopolymer`` Polymer: its ``flexPseudoSpec`` Blocks with
  ``lengthfrac``, ``ds``, ``bSegRatio`` (default 1), ``headjoined`` /
  ``tailjoined`` (pspolymer/PsBlockCopolymer.cpp, psbase/PsBlockBase.cpp:
  110-131); the species of a block is its ``scfield``: the first field of the
  ``flory`` Interaction's ``scfields`` is A, the second B;
* chiN = chi * length of the first Polymer (its scale length,
  pseffhamil/PsFloryInteraction.cpp:160-161);
* the ``steepestDescent`` Updater's ``relaxlambdas`` and ``noise``
  (pseffhamil/PsSteepDUpdater.cpp:56-63); ``nsteps``, ``randomSeed``;
* ``simpleSpecFilter`` / ``multiSpecFilter`` Updaters in the EffHamil's
  ``updaterSequence`` (pseffhamil/PsEffHamil.cpp:185-203): their
  ``updatefields``, ``cutoffFactor``, ``filterStrength``, ``specCellSizes``
  and the ``applyStart`` / ``applyFrequency`` / ``applyEnd`` gate
  (psbase/PsUpdater.cpp:60-118), applied to the engine's conjugate fields
  at each step they are due at.  A filter after the steepest-descent
  Updater acts on the updated fields; one before it acts between the
  step's densities (computed first, pstrol/PsDomain.cpp:354-365) and the
  field update, which then starts from the filtered fields
  (pseffhamil/PsSteepDUpdater.cpp:170-184).

Anything else that changes the physics (solvents, several polymers,
charged or polydisperse blocks, walls, chi functions) raises ``DeckError``.
"""
from __future__ import annotations

import ast
import dataclasses
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from .updaters import SpecFilterUpdater, UpdaterSchedule


class DeckError(ValueError):
    """The deck is malformed or uses a feature outside the supported subset."""


# ---------------------------------------------------------------------------
# expressions
# ---------------------------------------------------------------------------
_FUNCS = {"float": float, "int": int, "abs": abs, "min": min, "max": max, "round": round,
          "sqrt": math.sqrt, "exp": math.exp, "log": math.log, "sin": math.sin,
          "cos": math.cos, "tan": math.tan, "str": str}
_CONSTS = {"pi": math.pi, "True": True, "False": False}
_BIN = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b, ast.Mult: lambda a, b: a * b,
        ast.Div: lambda a, b: a / b, ast.FloorDiv: lambda a, b: a // b, ast.Mod: lambda a, b: a % b,
        ast.Pow: lambda a, b: a ** b}
_CMP = {ast.Eq: lambda a, b: a == b, ast.NotEq: lambda a, b: a != b, ast.Lt: lambda a, b: a < b,
        ast.LtE: lambda a, b: a <= b, ast.Gt: lambda a, b: a > b, ast.GtE: lambda a, b: a >= b}


def evaluate(expr: str, env: dict):
    """Evaluate a txpp expression (arithmetic, comparisons, and/or/not,
    float/int/abs/min/max/round/sqrt/exp/log/sin/cos/tan, pi) with the
    variables of env.  Nothing else is executed."""
    try:
        tree = ast.parse(expr.strip(), mode="eval")
    except SyntaxError as e:
        raise DeckError(f"bad expression {expr!r}: {e.msg}") from None

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float, str, bool)):
            return n.value
        if isinstance(n, ast.Name):
            if n.id in env:
                return env[n.id]
            if n.id in _CONSTS:
                return _CONSTS[n.id]
            raise DeckError(f"undefined variable {n.id!r} in {expr!r}")
        if isinstance(n, ast.BinOp) and type(n.op) in _BIN:
            return _BIN[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp):
            v = ev(n.operand)
            if isinstance(n.op, ast.USub):
                return -v
            if isinstance(n.op, ast.UAdd):
                return +v
            if isinstance(n.op, ast.Not):
                return not v
        if isinstance(n, ast.Compare):
            left = ev(n.left)
            for op, right in zip(n.ops, n.comparators):
                r = ev(right)
                if type(op) not in _CMP or not _CMP[type(op)](left, r):
                    return False
                left = r
            return True
        if isinstance(n, ast.BoolOp):
            vals = [ev(v) for v in n.values]
            return all(vals) if isinstance(n.op, ast.And) else any(vals)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS \
                and not n.keywords:
            return _FUNCS[n.func.id](*[ev(a) for a in n.args])
        raise DeckError(f"unsupported construct in expression {expr!r}")

    return ev(tree)


# ---------------------------------------------------------------------------
# deck tree
# ---------------------------------------------------------------------------
@dataclass
class Node:
    """One ``<Kind name>`` block: attributes in file order and child blocks."""
    kind: str
    name: str
    attrs: dict = field(default_factory=dict)
    children: list = field(default_factory=list)

    def all(self, kind: str) -> list:
        return [c for c in self.children if c.kind == kind]

    def one(self, kind: str, name: str | None = None) -> "Node":
        cands = [c for c in self.all(kind) if name is None or c.name == name]
        if len(cands) != 1:
            raise DeckError(f"expected one <{kind}{' ' + name if name else ''}> in <{self.kind} "
                            f"{self.name}>, found {len(cands)}")
        return cands[0]

    def get(self, key, default=None):
        return self.attrs.get(key, default)


def _strip_comment(line: str) -> str:
    out, q = [], None
    for ch in line:
        if q:
            if ch == q:
                q = None
        elif ch in "\"'":
            q = ch
        elif ch == "#":
            break
        out.append(ch)
    return "".join(out).rstrip()


_WORD = re.compile(r"^[A-Za-z_][A-Za-z0-9_]*$")
_NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")


def _scalar(tok: str, env: dict):
    tok = tok.strip()
    if not tok:
        return ""
    if len(tok) >= 2 and tok[0] == tok[-1] and tok[0] in "\"'":
        inner = tok[1:-1]
        # txpp strings such as "'off'" carry their own quotes
        if len(inner) >= 2 and inner[0] == inner[-1] and inner[0] in "\"'":
            inner = inner[1:-1]
        return inner
    if _NUM.match(tok):
        v = float(tok)
        return int(v) if re.match(r"^[-+]?\d+$", tok) else v
    if _WORD.match(tok):
        if tok in env:
            return env[tok]
        return tok
    try:
        return evaluate(tok, env)
    except DeckError:
        return tok


def _value(text: str, env: dict):
    text = text.strip()
    if text.startswith("[") and text.endswith("]"):
        toks = [t for t in re.split(r"[\s,]+", text[1:-1].strip()) if t]
        return [_scalar(t, env) for t in toks]
    return _scalar(text, env)


_BUILTIN_MACROS = {
    # setupPS(nx, ny, nz, dx, dy, dz, debug): the uniform grid of the
    # pseudo-spectral melt decks (psutils/macros/pseudoSpecSetup.mac:32-99
    # declares it together with decomposition, MPI and FFT objects, which a
    # single-GPU engine does not need).
    "setupPS": (["nx", "ny", "nz", "dx", "dy", "dz", "debug"],
                ["defaultGrid = mainGrid",
                 "<Grid mainGrid>",
                 "  kind = uniCartGrid",
                 "  numCellsGlobal = [nx ny nz]",
                 "  cellSizes = [dx dy dz]",
                 "</Grid>"]),
}


class _Parser:
    def __init__(self, search_path):
        self.search_path = list(search_path)
        self.macros = {}
        self.env = {}

    def _import(self, name: str, here: str):
        for d in [here] + self.search_path:
            path = os.path.join(d, name + ".mac")
            if os.path.exists(path):
                self._read_macros(open(path).read().splitlines(), os.path.dirname(path))
                return
        if name == "pseudoSpecSetup":
            for k, v in _BUILTIN_MACROS.items():
                self.macros.setdefault(k, v)
            return
        raise DeckError(f"$ import {name}: {name}.mac not found on the search path")

    def _read_macros(self, lines, here):
        i = 0
        while i < len(lines):
            s = _strip_comment(lines[i]).strip()
            m = re.match(r"^<macro\s+([A-Za-z_]\w*)\s*\(([^)]*)\)\s*>$", s)
            if m:
                params = [p.strip() for p in m.group(2).split(",") if p.strip()]
                body = []
                i += 1
                while i < len(lines) and not _strip_comment(lines[i]).strip().startswith("</macro>"):
                    body.append(lines[i])
                    i += 1
                self.macros[m.group(1)] = (params, body)
            elif s.startswith("$"):
                self._dollar(s[1:].strip(), here, [True])
            i += 1

    def _dollar(self, s: str, here: str, active: list):
        """Variable and import lines (conditionals are handled by the caller)."""
        if not active[-1]:
            return
        m = re.match(r"^import\s+(\S+)$", s)
        if m:
            self._import(m.group(1), here)
            return
        m = re.match(r"^([A-Za-z_]\w*)\s*=\s*(.+)$", s)
        if m:
            self.env[m.group(1)] = _value(m.group(2), self.env)
            return
        raise DeckError(f"unsupported txpp line: $ {s}")

    def parse(self, lines, here, root: Node):
        stack = [root]
        # conditional stack: (branch active, some branch already taken, parent active)
        cond = [(True, True, True)]
        for lineno, raw in enumerate(lines, 1):
            s = _strip_comment(raw).strip()
            if not s:
                continue
            active = cond[-1][0]
            if s.startswith("$"):
                body = s[1:].strip()
                m = re.match(r"^(if|elif)\s+(.+)$", body)
                if m and m.group(1) == "if":
                    parent = active
                    ok = parent and bool(evaluate(m.group(2), self.env))
                    cond.append((ok, ok, parent))
                    continue
                if m and m.group(1) == "elif":
                    _, taken, parent = cond[-1]
                    ok = parent and not taken and bool(evaluate(m.group(2), self.env))
                    cond[-1] = (ok, taken or ok, parent)
                    continue
                if body == "else":
                    _, taken, parent = cond[-1]
                    cond[-1] = (parent and not taken, True, parent)
                    continue
                if body == "endif":
                    if len(cond) == 1:
                        raise DeckError(f"line {lineno}: $ endif without $ if")
                    cond.pop()
                    continue
                self._dollar(body, here, [active])
                continue
            if not active:
                continue
            m = re.match(r"^<(Off\s+)?([A-Za-z_]\w*)\s+([^>\s]+)\s*>$", s)
            if m:
                node = Node(m.group(2), m.group(3))
                node.off = bool(m.group(1))
                stack[-1].children.append(node)
                stack.append(node)
                continue
            m = re.match(r"^</([A-Za-z_]\w*)(\s+[^>\s]+)?\s*>$", s)
            if m:
                if len(stack) == 1 or stack[-1].kind != m.group(1):
                    raise DeckError(f"line {lineno}: unbalanced </{m.group(1)}>")
                stack.pop()
                continue
            m = re.match(r"^([A-Za-z_]\w*)\s*\((.*)\)$", s)
            if m and m.group(1) in self.macros:
                params, body = self.macros[m.group(1)]
                args = [a.strip() for a in m.group(2).split(",")] if m.group(2).strip() else []
                if len(args) != len(params):
                    raise DeckError(f"line {lineno}: {m.group(1)} takes {len(params)} arguments")
                vals = {p: _scalar(a, self.env) for p, a in zip(params, args)}
                expanded = []
                for bl in body:
                    t = bl
                    for p, v in vals.items():
                        rep = f"'{v}'" if isinstance(v, str) and not _WORD.match(v) else str(v)
                        t = re.sub(rf"\b{re.escape(p)}\b", rep, t)
                    expanded.append(t)
                sub = _Parser(self.search_path)
                sub.macros, sub.env = self.macros, self.env
                tmp = Node("macro", m.group(1))
                sub.parse(expanded, here, tmp)
                stack[-1].children.extend(tmp.children)
                stack[-1].attrs.update(tmp.attrs)
                continue
            m = re.match(r"^([A-Za-z_]\w*)\s*=\s*(.*)$", s)
            if m:
                stack[-1].attrs[m.group(1)] = _value(m.group(2), self.env)
                continue
            raise DeckError(f"line {lineno}: cannot parse {s!r}")
        if len(stack) != 1:
            raise DeckError(f"<{stack[-1].kind} {stack[-1].name}> is not closed")
        if len(cond) != 1:
            raise DeckError("$ if without $ endif")
        _drop_off(root)
        return root


def _drop_off(node: Node):
    node.children = [c for c in node.children if not getattr(c, "off", False)]
    for c in node.children:
        _drop_off(c)


def parse(text: str, search_path=(), here: str = ".", variables: dict | None = None) -> Node:
    """Parse deck text into a tree rooted at Node('deck', '')."""
    p = _Parser(search_path)
    if variables:
        p.env.update(variables)
    root = p.parse(text.splitlines(), here, Node("deck", ""))
    root.variables = dict(p.env)
    return root


def load(path: str, search_path=(), variables: dict | None = None) -> Node:
    """Parse a .pre file; macros are looked up next to it, then on search_path."""
    return parse(open(path).read(), search_path, os.path.dirname(os.path.abspath(path)), variables)


# ---------------------------------------------------------------------------
# deck -> engine configuration
# ---------------------------------------------------------------------------
@dataclass
class MeltConfig:
    """Everything the engine needs from a bulk block-copolymer melt deck."""
    shape: tuple
    cell: tuple
    blocks: list            # polyswift_plusplus_amd.scft.Block
    block_names: list
    ds: float
    chiN: float
    lambdas: tuple          # relaxlambdas (lambda0, lambda1)
    noise: float
    nsteps: int
    seed: int | None
    species_fields: tuple   # (A field name, B field name)
    history_period: int | None
    constraint_field: str = "defaultPressure"
    dump_period: int | None = None
    filters: list = field(default_factory=list)   # SpecFilterUpdater, in sequence order
    # monomerDens PhysFields in deck order: each draws N uniforms for its
    # random initial conjugate field (psphysf/PsMonomerDens.cpp:116-126)
    init_fields: tuple = ()
    # every Polymer (BlendPolymer) in deck order; blocks / block_names above
    # are the first one's
    polymers: list = field(default_factory=list)

    @property
    def is_blend(self) -> bool:
        return len(self.polymers) > 1 or any(p.volfrac != 1.0 for p in self.polymers)


@dataclass
class BlendPolymer:
    """One <Polymer> of the deck: volfrac, length and its blocks (head/tail
    indices within the polymer)."""
    name: str
    volfrac: float
    length: float
    blocks: list
    block_names: list


def _polymer_blocks(bnodes, fields, ds_vals):
    """The flexPseudoSpec Blocks of one Polymer as engine Blocks (joins by
    index within the polymer); adds each block's ds to ds_vals."""
    from .scft import Block
    names = [b.name for b in bnodes]
    idx = {nm: i for i, nm in enumerate(names)}
    blocks = []
    for b in bnodes:
        if b.get("kind") != "flexPseudoSpec":
            raise DeckError(f"Block kind {b.get('kind')!r} unsupported (flexPseudoSpec only)")
        sc = b.get("scfield")
        if sc not in fields:
            raise DeckError(f"block {b.name}: scfield {sc!r} is not one of the flory scfields")

        def joins(key):
            v = b.get(key)
            if v is None:
                raise DeckError(f"block {b.name}: {key} not set")
            v = v if isinstance(v, list) else [v]
            out = []
            for nm in v:
                if nm == "freeEnd":
                    continue
                if nm not in idx:
                    raise DeckError(f"block {b.name}: {key} names unknown block {nm!r}")
                out.append(idx[nm])
            return tuple(out)

        ds_vals.add(float(b.get("ds")))
        # lengthRatio (length / scale length) enters the engine per polymer
        blocks.append(Block("A" if sc == fields[0] else "B", float(b.get("lengthfrac")),
                            float(b.get("bSegRatio", 1.0)), joins("headjoined"), joins("tailjoined")))
    return blocks, names


_SUPPORTED_TOP = {"Grid", "Decomp", "Comm", "FFT", "PhysField", "EffHamil", "Polymer", "History",
                  "XSim", "XVar"}


def configure(root: Node) -> MeltConfig:
    """Map a parsed deck onto the engine (see the module docstring)."""
    for c in root.children:
        if c.kind not in _SUPPORTED_TOP:
            raise DeckError(f"<{c.kind} {c.name}> is outside the bulk block-copolymer melt subset")
    grids = root.all("Grid")
    if not grids:
        raise DeckError("no <Grid> block (call setupPS or declare one)")
    gname = root.get("defaultGrid")
    grid = next((g for g in grids if g.name == gname), grids[0])
    n = grid.get("numCellsGlobal")
    h = grid.get("cellSizes")
    if not (isinstance(n, list) and isinstance(h, list) and len(n) in (1, 2, 3) and len(n) == len(h)):
        raise DeckError(f"<Grid {grid.name}> needs numCellsGlobal and cellSizes of equal length")
    shape = tuple(int(x) for x in n) + (1,) * (3 - len(n))
    cell = tuple(float(x) for x in h) + (1.0,) * (3 - len(h))

    polys = root.all("Polymer")
    if not 1 <= len(polys) <= 8:
        raise DeckError(f"1 to 8 Polymers supported (a melt or a blend of block copolymers), "
                        f"found {len(polys)}")
    for poly in polys:
        if poly.get("kind") != "blockCopolymer":
            raise DeckError(f"Polymer kind {poly.get('kind')!r} unsupported (blockCopolymer only)")
    # the first polymer's length scales every length (pspolymer/PsPolymerHldr.cpp:69-83)
    length = float(polys[0].get("length", 1.0))

    hams = root.all("EffHamil")
    if len(hams) != 1:
        raise DeckError("one <EffHamil> expected")
    ham = hams[0]
    for c in ham.children:
        if c.kind not in ("Updater", "Interaction"):
            raise DeckError(f"<{c.kind} {c.name}> in EffHamil unsupported")
    inters = ham.all("Interaction")
    if len(inters) != 1 or inters[0].get("kind") != "flory":
        raise DeckError("one flory Interaction expected")
    inter = inters[0]
    if inter.children:
        raise DeckError("chi functions (STFunc) unsupported")
    fields = inter.get("scfields")
    if not (isinstance(fields, list) and len(fields) == 2):
        raise DeckError("flory scfields must name two fields")
    if inter.get("chi") is None:
        raise DeckError(f"flory Interaction {inter.name}: chi not set")
    chiN = float(inter.get("chi")) * length
    ups = ham.all("Updater")
    sds = [u for u in ups if u.get("kind") == "steepestDescent"]
    if len(sds) != 1:
        raise DeckError("one steepestDescent Updater expected")
    sd = sds[0]
    lam = sd.get("relaxlambdas", [0.1, 0.0])
    if not isinstance(lam, list):
        lam = [lam]
    lam = (float(lam[0]), float(lam[1]) if len(lam) > 1 else 0.0)
    noise = float(sd.get("noise", 0.0))
    seq = ham.get("updaterSequence", [u.name for u in ups])
    seq = seq if isinstance(seq, list) else [seq]
    if sd.name not in seq:
        raise DeckError(f"updaterSequence does not name the field update {sd.name!r}")
    by_name = {u.name: u for u in ups}
    filters = []
    for pos, nm in enumerate(seq):
        u = by_name.get(nm)
        if u is None:
            raise DeckError(f"updaterSequence names an undefined Updater {nm!r}")
        if u is sd:
            continue
        kind = u.get("kind")
        if kind == "poisson" or kind == "poissonUpdater":
            raise DeckError("the Poisson updater needs charged species, outside the melt subset")
        if kind not in ("simpleSpecFilter", "multiSpecFilter"):
            raise DeckError(f"Updater kind {kind!r} unsupported")
        if u.get("fftKind") is None:
            raise DeckError(f"Updater {nm!r}: fftKind not set")
        ufields = u.get("updatefields")
        ufields = ufields if isinstance(ufields, list) else [ufields]
        which = ""
        for fname in ufields:
            if fname not in fields:
                raise DeckError(f"Updater {nm!r}: updatefield {fname!r} is not a flory scfield")
            which += "A" if fname == fields[0] else "B"
        cells = None
        if kind == "multiSpecFilter":
            cells = u.get("specCellSizes")
            if not isinstance(cells, list):
                raise DeckError(f"Updater {nm!r}: specCellSizes not set")
            cells = tuple(int(c) for c in cells) + (1,) * (3 - len(cells))
        sched = UpdaterSchedule(int(u.get("applyStart", 0)), int(u.get("applyFrequency", 1)),
                                None if u.get("applyEnd") is None else int(u.get("applyEnd")))
        filters.append(SpecFilterUpdater(which, float(u.get("cutoffFactor", 0.0)),
                                         float(u.get("filterStrength", 0.0)), cells, sched, nm,
                                         before=pos < seq.index(sd.name)))

    ds_vals = set()
    polymers = []
    for poly in polys:
        bnodes = poly.all("Block")
        if not bnodes:
            raise DeckError(f"Polymer {poly.name} has no blocks")
        pblocks, pnames = _polymer_blocks(bnodes, fields, ds_vals)
        polymers.append(BlendPolymer(poly.name, float(poly.get("volfrac", 1.0)),
                                     float(poly.get("length", 1.0)), pblocks, pnames))
    blocks, names = polymers[0].blocks, polymers[0].block_names
    if len(ds_vals) != 1:
        raise DeckError(f"blocks must share one ds, found {sorted(ds_vals)}")

    if len(ds_vals) != 1:
        raise DeckError(f"blocks must share one ds, found {sorted(ds_vals)}")
    hist = [hh for hh in root.all("History") if hh.get("kind") == "freeEnergy"]
    period = int(hist[0].get("updatePeriodicity", 1)) if hist else None
    seed = root.get("randomSeed")
    cons = [pf.name for pf in root.all("PhysField") if pf.get("kind") == "constraint"]
    init_fields = []
    for pf in root.all("PhysField"):
        if pf.get("kind") != "monomerDens":
            continue
        if pf.get("initOption", "random") != "random":
            raise DeckError(f"PhysField {pf.name}: initOption {pf.get('initOption')!r} unsupported "
                            "(random only)")
        init_fields.append(pf.name)
    dump = root.get("dumpPeriodicity")
    return MeltConfig(shape=shape, cell=cell, blocks=blocks, block_names=names, ds=ds_vals.pop(),
                      polymers=polymers,
                      chiN=chiN, lambdas=lam, noise=noise, nsteps=int(root.get("nsteps", 1)),
                      seed=None if seed is None else int(seed), species_fields=tuple(fields),
                      history_period=period, constraint_field=cons[0] if cons else "defaultPressure",
                      dump_period=None if dump is None else int(dump), filters=filters,
                      init_fields=tuple(init_fields))


def build(cfg: MeltConfig, prec="fp64", stream=None):
    """The engine for a configured deck: one block copolymer, or a blend."""
    from .scft import BlendSCFT, BlockCopolymerSCFT, Polymer
    if cfg.is_blend:
        return BlendSCFT(cfg.shape, cfg.cell,
                         [Polymer(p.volfrac, p.length, tuple(p.blocks)) for p in cfg.polymers],
                         cfg.ds, cfg.chiN, prec=prec, stream=stream)
    return BlockCopolymerSCFT(cfg.shape, cfg.cell, cfg.blocks, cfg.ds, cfg.chiN, prec=prec,
                              stream=stream)


def reference_initial_fields(cfg: MeltConfig, rng) -> tuple:
    """(w_A, w_B) as the reference initialises them: every monomerDens
    PhysField in deck order takes N uniforms u of theRng and sets its
    conjugate field to 1.0 * (u + (-0.5)) (PsGridField::resetRandom,
    psbase/PsGridField.cpp:389-402; psphysf/PsMonomerDens.cpp:116-126)."""
    n = int(np.prod(cfg.shape))
    out = {}
    for name in cfg.init_fields:
        out[name] = (1.0 * (rng.uniform(n) + (-0.5))).reshape(cfg.shape)
    a, b = cfg.species_fields
    if a not in out or b not in out:
        raise DeckError("the flory scfields must be monomerDens PhysFields")
    return out[a], out[b]


def run(deck, iters: int | None = None, prec="fp64", sigma: float = 0.1, fields_seed=None,
        use_graph: bool = True, tol: float = 0.0, prefix: str | None = None,
        restore_step: int | None = None, search_path=(), init: str = "reference",
        noise: bool = True):
    """Run a deck: its steepest-descent relaxation (with the deck's noise and
    spectral filters) for ``nsteps`` (or ``iters``) iterations.

    The reference's random stream (MT19937 seeded with ``randomSeed``,
    0 when the deck sets none) feeds the initial fields and the noise, in the
    reference's order.  Initial fields: ``init="reference"`` draws them as the
    reference does (U(-0.5, 0.5) per monomerDens PhysField); ``init="gaussian"``
    uses the project's seeded N(0, sigma) fields (Philox keyed by
    ``fields_seed`` or randomSeed); ``restore_step`` reads the PhysField files
    ``<prefix>_<field>_<1000000+step>.h5`` of that step instead (the
    reference's ``-r`` restart, which draws no initial fields).  With
    ``prefix`` the PhysFields are dumped every ``dumpPeriodicity`` steps,
    after the update, as the reference's main loop does
    (polyswift/polyswift.cxx:544-588).  ``noise=False`` leaves the noise term
    out (noise-free runs)."""
    from . import fields as F
    from . import physfield
    from .scft import Mixing, Rng
    root = load(deck, search_path) if isinstance(deck, str) else deck
    cfg = configure(root)
    e = build(cfg, prec)
    names = (*cfg.species_fields, cfg.constraint_field)
    rng = Rng(cfg.seed if cfg.seed is not None else 0)
    step0 = 0
    if restore_step is not None:
        if prefix is None:
            raise DeckError("restore_step needs the dump prefix")
        physfield.restore(e, prefix, restore_step, names)
        step0 = int(restore_step)
    elif init == "reference":
        e.setFields(*reference_initial_fields(cfg, rng))
    elif init == "gaussian":
        seed = fields_seed if fields_seed is not None else (cfg.seed or 0)
        wA, wB = F.seeded_fields(cfg.shape, cfg.cell, sigma=sigma, seed=seed)
        e.setFields(wA, wB)
    else:
        raise DeckError(f"init must be 'reference' or 'gaussian', not {init!r}")
    if noise and cfg.noise != 0.0:
        e.setNoise(cfg.noise, rng)
    n = int(iters or cfg.nsteps)
    mix = Mixing("steepest", *cfg.lambdas)
    period = cfg.dump_period if prefix else None
    filters = [dataclasses.replace(f, schedule=f.schedule.resolve(cfg.nsteps, step0))
               for f in cfg.filters]
    pre = [f for f in filters if f.before]
    post = [f for f in filters if not f.before]
    hist = {"F": [], "Q": [], "err": []}
    t, done = step0, 0
    while done < n:
        if any(f.schedule.due(t + 1) for f in pre):
            # A step with filters before its field update runs split: the
            # densities of the unfiltered fields, the filters, then the
            # steepest-descent update from the filtered fields.  Its residual
            # is not read back (NaN in the history).
            e.update()
            Q = e.calcBigQ()
            for f in pre:
                f.update(e, t + 1)
            F = e.steepestDescent(*cfg.lambdas)
            hist["F"].append(F)
            hist["Q"].append(Q)
            hist["err"].append(float("nan"))
            t += 1
            done += 1
            chunk = h_iters = 1
        else:
            chunk = n - done
            if period:
                chunk = min(chunk, period - t % period)
            for f in post:
                nd = f.schedule.next_due(t)
                if nd is not None:
                    chunk = min(chunk, nd - t)
            for f in pre:              # stop before the step it is due at
                nd = f.schedule.next_due(t)
                if nd is not None:
                    chunk = min(chunk, nd - 1 - t)
            h = e.iterate(chunk, mix, tol=tol, use_graph=use_graph)
            for k in hist:
                hist[k].extend(h[k])
            h_iters = h["iters"]
            t += h_iters
            done += h_iters
        if h_iters == chunk:
            for f in post:             # after the field update, in updaterSequence order
                f.update(e, t)
        if period and t % period == 0:
            physfield.dump(e, prefix, t, names)
        if h_iters < chunk:        # converged
            break
    out = {k: np.asarray(v) for k, v in hist.items()}
    out.update(iters=done, step=t, config=cfg, engine=e, rng=rng,
               noise=cfg.noise if noise else 0.0)
    return out
