"""The launch lists of a sweep, on the CPU: a model of sweep_launches and
range_fio (polyswift_plusplus_amd/csrc/psx_scft.hip) in both sweep modes,
executed symbolically.

Contour step s applies the axis round trips D_0 (z), D_1 (y), D_2 (x) in the
order 0, 1, 2 when s is even and 2, 1, 0 when s is odd; line mode fuses the
last pass of a step with the first of the next (same axis), plane mode groups
every run of z and y passes into one plane launch.  Executing the rows
(in, pre, round trips, post, mid, pre2, round trips, post2, out) must give
every stored slice as W D_z D_y D_x W of the one before, with no row needing
a pointer it does not have.  The GPU parity tests check the C++ itself
(tests/test_gpu_plane.py, tests/test_gpu_parity.py)."""
import pytest

PLANE = 3


def sweep_launches(na, Ns, plane):
    ops = []
    for st in range(Ns):
        for p in range(na):
            a = p if st % 2 == 0 else na - 1 - p
            ops.append(dict(axis=a, step=st, first=p == 0, last=p == na - 1))
    out, i = [], 0
    while i < len(ops):
        if plane and na == 3 and ops[i]["axis"] <= 1:
            j = i
            while j < len(ops) and ops[j]["axis"] <= 1:
                j += 1
            o = dict(ops[i], axis=PLANE, last=(j - i == 2) and ops[j - 1]["last"])
            out.append((o, j - i, ops[i:j]))
            i = j
        elif ops[i]["last"] and i + 1 < len(ops) and ops[i + 1]["axis"] == ops[i]["axis"]:
            out.append((ops[i], 2, ops[i:i + 2]))
            i += 2
        else:
            out.append((ops[i], 1, ops[i:i + 1]))
            i += 1
    return out


def range_fio(na, Ns, plane):
    rows = []
    for o, cnt, sub in sweep_launches(na, Ns, plane):
        f = dict(in_=("S", o["step"]) if o["first"] else "tmp", pre="W" if o["first"] else None)
        if o["axis"] == PLANE:
            assert cnt in (2, 4)
            if cnt == 2:
                f.update(emul=("P", o["step"]), post="W" if o["last"] else None,
                         out=("S", o["step"] + 1) if o["last"] else "tmp")
            else:   # (y, z last | z first, y)
                assert not o["first"] and sub[1]["last"] and sub[2]["first"] and not sub[3]["last"]
                f.update(emul=("P", o["step"]), post="W", mid=("S", o["step"] + 1), pre2="W",
                         emul2=("P", o["step"] + 1), out="tmp")
        else:
            f["emul"] = (o["axis"], o["step"])
            if cnt == 1:
                f.update(post="W" if o["last"] else None,
                         out=("S", o["step"] + 1) if o["last"] else "tmp")
            else:
                s2, last2 = o["step"] + 1, na == 1
                f.update(post="W", mid=("S", s2), pre2="W", emul2=(o["axis"], s2),
                         post2="W" if last2 else None, out=("S", s2 + 1) if last2 else "tmp")
        rows.append((f, o["axis"]))
    return rows


def execute(na, Ns, plane):
    store = {("S", 0): ()}
    rows = range_fio(na, Ns, plane)
    for f, ax in rows:
        for k in ("in_", "out", "emul"):
            assert f.get(k) is not None, (k, f)
        if f.get("emul2") is not None:
            assert f.get("mid") is not None

        def rt(v, emul):
            if ax == PLANE:
                return v + (("D", 0, emul[1]), ("D", 1, emul[1]))
            return v + (("D", emul[0], emul[1]),)

        v = store[f["in_"]]
        if f.get("pre"):
            v += ("W",)
        v = rt(v, f["emul"])
        if f.get("post"):
            v += ("W",)
        if f.get("emul2") is not None:
            store[f["mid"]] = v
            if f.get("pre2"):
                v += ("W",)
            v = rt(v, f["emul2"])
            if f.get("post2"):
                v += ("W",)
        store[f["out"]] = v
    for s in range(1, Ns + 1):
        prev, cur = store[("S", s - 1)], store[("S", s)]
        assert cur[:len(prev)] == prev, s
        step = cur[len(prev):]
        assert step[0] == "W" and step[-1] == "W", step
        assert sorted(step[1:-1]) == sorted(("D", a, s - 1) for a in range(na)), (s, step)
    return rows


@pytest.mark.parametrize("plane", [False, True])
@pytest.mark.parametrize("na", [1, 2, 3])
def test_every_slice_is_one_contour_step_of_the_last(na, plane):
    for Ns in range(1, 40):
        execute(na, Ns, plane)


def test_launch_counts():
    # line mode: 2 Ns + 1 launches per 3D sweep; plane mode: Ns + 1
    # (51 plane launches and 50 fused x launches at Ns = 100)
    assert len(range_fio(3, 100, False)) == 201
    rows = range_fio(3, 100, True)
    assert len(rows) == 101
    assert sum(1 for _, ax in rows if ax == PLANE) == 51
    for Ns in range(2, 40):
        assert len(range_fio(3, Ns, True)) == Ns + 1


def test_plane_launches_alternate_with_fused_x():
    axes = [ax for _, ax in range_fio(3, 10, True)]
    assert axes == [PLANE, 2] * 5 + [PLANE]
