"""CPU checks of the K-step pass planner (vl_plan_passes, the host code behind the passes of
vl_sim_create and vl_batch_build) on channel-like flag maps with bodies.  The GPU tests show
bit equality with single steps on particular domains; these pin the plan's invariants on
many: strips that cover only interior nodes and whose dependency footprints are interior, every
interior node covered once or listed once, ring-edge bits that cover every uncovered neighbour
of a covered node (the ring stores write the populations pointing across the marked edges), and
far tasks first with no uncovered node within K of their outputs."""
import numpy as np
import pytest

import vortexlab_amd as vl
from vortexlab_amd import api


def tb_halo(k, v):
    return (k + v - 1) // v * v


def tb_out(k, v):
    return (64 * v - tb_halo(k, v) - k) // v * v


def channel_flags(nx, ny, bodies, rects=()):
    """0 interior, 1 boundary (the domain edge, or next to a solid node), 2 solid."""
    yy, xx = np.mgrid[0:ny, 0:nx]
    solid = np.zeros((ny, nx), bool)
    for cx, cy, r in bodies:
        solid |= (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r
    for x0, x1, y0, y1 in rects:
        solid[y0:y1, x0:x1] = True
    near = np.zeros_like(solid)
    p = np.pad(solid, 1)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            near |= p[1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    edge = np.zeros_like(solid)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    fl = np.where(solid, 2, np.where(near | edge, 1, 0)).astype(np.uint8)
    return fl


def random_case(seed):
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(90, 320)), int(rng.integers(30, 130))
    bodies = [(rng.uniform(0.2, 0.8) * nx, rng.uniform(0.25, 0.75) * ny, rng.uniform(2.0, 0.2 * ny))
              for _ in range(int(rng.integers(0, 3)))]
    rects = []
    if rng.random() < 0.5:
        x0 = int(rng.integers(5, nx - 20))
        y0 = int(rng.integers(3, ny - 10))
        rects.append((x0, x0 + int(rng.integers(2, 15)), y0, y0 + int(rng.integers(2, 8))))
    return channel_flags(nx, ny, bodies, rects)


def joined_groups(tasks, nfar, waves):
    """The joined groups of a plan: (first task, end) for each, checked to be whole workgroups at
    the start of the far range and of the near range."""
    joined = (tasks[:, 3] >> 13) & 1
    groups = []
    for b, e in ((0, nfar), (nfar, len(tasks))):
        n = b
        while n < e and joined[n]:
            n += 1
        assert (n - b) % waves == 0, "joined tasks fill whole workgroups"
        assert not joined[n:e].any(), "joined groups come first in their range"
        groups += [(g, g + waves) for g in range(b, n, waves)]
    return groups


def check_plan(fl, k, v, rows, x0, waves=0):
    """The plan's invariants; waves > 0 plans joined strips for workgroups of that many waves."""
    ny, nx = fl.shape
    tasks, nfar, listed, ncov = api.plan_passes(vl.backend(), fl, k, v, rows, x0, join=waves > 0,
                                                waves=waves if waves else 0)
    interior = fl == 0
    fluid = fl != 2
    covered = np.zeros((ny, nx), np.int32)
    owner = -np.ones((ny, nx), np.int64)
    groups = joined_groups(tasks, nfar, waves) if waves else []
    group_of = -np.ones(len(tasks), np.int64)
    for gi, (g0, g1) in enumerate(groups):
        group_of[g0:g1] = gi
        grp = tasks[g0:g1]
        assert (grp[:, 1] == grp[0, 1]).all() and (grp[:, 2] == grp[0, 2]).all()
        assert ((grp[:, 3] >> 12) & 3 == 3).all(), "joined tasks are lockstep"
        loaded = grp[:, 0] - ((grp[:, 3] >> 16) & 0xFF)  # each wave's first loaded column
        assert (np.diff(loaded) == 64 * v).all(), "the waves load adjacent strips"
        assert loaded[0] % v == 0, "aligned vector loads"
        ends = grp[:, 0] + (grp[:, 3] & 0xFF)
        assert (grp[1:, 0] == ends[:-1]).all(), "the waves' outputs are contiguous"
        # outputs where level K is valid: K in from each end of the loaded columns
        assert grp[0, 0] - loaded[0] >= k and loaded[-1] + 64 * v - ends[-1] >= k
    for i, (tx, ty, tz, tw) in enumerate(tasks):
        w, edges = int(tw) & 0xFF, (int(tw) >> 8) & 15
        if group_of[i] < 0:
            # the lockstep bit only as the caller's grouping sets it (waves > 0), never joined
            assert (int(tw) >> (13 if waves else 12)) == 0
            assert 1 <= w <= tb_out(k, v) and w % v == 0 and (tx - tb_halo(k, v)) % v == 0
        else:
            assert 1 <= w <= 64 * v and w % v == 0
        assert 1 <= tz <= rows
        covered[ty:ty + tz, tx:tx + w] += 1
        owner[ty:ty + tz, tx:tx + w] = i
        # the dependency footprint (K - 1 around the outputs) lies inside the domain's interior
        fy0, fy1, fx0, fx1 = ty - (k - 1), ty + tz + k - 1, tx - (k - 1), tx + w + k - 1
        assert fy0 >= 1 and fy1 <= ny - 1 and fx0 >= 1 and fx1 <= nx - 1
        assert interior[fy0:fy1, fx0:fx1].all(), f"task {i}: footprint not interior"
    assert covered.max(initial=0) <= 1, "tasks overlap"
    assert int(covered.sum()) == ncov
    # every interior node is covered or listed, once
    want = np.flatnonzero((interior & (covered == 0)).ravel())
    got_y, got_x = (listed >> 16).astype(np.int64), (listed & 0xFFFF).astype(np.int64)
    assert np.array_equal(np.sort(got_y * nx + got_x), want)
    assert np.all(np.diff(listed.astype(np.int64)) > 0), "listed nodes in (y, x) order"
    # ring edges: every uncovered fluid neighbour u of a covered node r lies outside r's task
    # across sides whose bits are set
    uncovered = fluid & (covered == 0)
    ys, xs = np.nonzero(covered)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx == 0 and dy == 0:
                continue
            uy, ux = ys + dy, xs + dx
            hit = uncovered[uy, ux]
            for ry, rx, qy, qx in zip(ys[hit], xs[hit], uy[hit], ux[hit]):
                tx, ty, tz, tw = tasks[owner[ry, rx]]
                w, edges = int(tw) & 0xFF, (int(tw) >> 8) & 15
                sides = [(qy < ty, 1), (qy > ty + tz - 1, 2), (qx < tx, 4), (qx > tx + w - 1, 8)]
                assert any(out for out, _ in sides)
                for out, bit in sides:
                    assert not out or edges & bit, f"ring node ({rx},{ry}): edge bit {bit} missing"
    # far tasks first: no uncovered fluid node within K of their outputs, and no ring edges; a
    # joined group is near when one of its waves is
    def near_box(t):
        tx, ty, tz, tw = (int(a) for a in t)
        w = tw & 0xFF
        return uncovered[max(ty - k, 0):ty + tz + k, max(tx - k, 0):tx + w + k].any()

    for i, t in enumerate(tasks):
        edges = (int(t[3]) >> 8) & 15
        if i < nfar:
            assert not near_box(t) and edges == 0, f"far task {i} near an uncovered node"
        elif group_of[i] < 0:
            assert near_box(t), f"near task {i} has no uncovered node within K"
        else:
            g0, g1 = groups[group_of[i]]
            assert any(near_box(u) for u in tasks[g0:g1]), f"near group at {g0} has no uncovered node within K"
    for b, e in ((0, nfar), (nfar, len(tasks))):
        single = [(int(t[1]), int(t[0])) for i, t in enumerate(tasks[b:e], b) if group_of[i] < 0]
        heads = [(int(tasks[g0][1]), int(tasks[g0][0])) for g0, _ in groups if b <= g0 < e]
        assert single == sorted(single) and heads == sorted(heads)
    return len(tasks), nfar, len(listed), ncov


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("k,v,rows", [(2, 1, 7), (3, 1, 20), (4, 1, 3), (5, 1, 20), (3, 2, 24), (4, 2, 1)])
def test_plan_invariants(seed, k, v, rows):
    fl = random_case(seed)
    x0 = 32 if seed % 2 else tb_halo(k, v) + (seed % 3) * v
    ntask, nfar, nlisted, ncov = check_plan(fl, k, v, rows, x0)
    assert ncov > 0 or fl.shape[0] < 2 * k + 3


def test_plan_of_the_channel_with_a_cylinder():
    """A 400 x 100 channel with a D = 30 cylinder: most of the interior is covered, the listed
    nodes are the margins along the walls, inlet and outlet and around the body."""
    fl = channel_flags(400, 100, [(100.0, 50.0, 15.0)])
    ntask, nfar, nlisted, ncov = check_plan(fl, 5, 1, 20, 32)
    assert ncov > 0.8 * int((fl == 0).sum())
    assert 0 < nfar < ntask


def test_plan_rejects_bad_arguments():
    fl = channel_flags(50, 20, [])
    be = vl.backend()
    for k, v, x0 in ((1, 1, 32), (6, 1, 32), (3, 3, 32), (4, 1, 2), (3, 2, 33), (4, 2, 5)):
        with pytest.raises(api.VortexError):
            api.plan_passes(be, fl, k, v, 20, x0)


def launch_groups(ranges, waves):
    """The workgroups of launches over `ranges`: W consecutive tasks from each begin."""
    for b, e in ranges:
        for g in range(b, e, waves):
            yield g, min(g + waves, e)


def check_lockstep(tasks, ranges, waves, owner=None):
    """Every workgroup that holds a lockstep task is full, all its tasks are marked and share
    first row, row count and owner (so its waves meet the same barriers); every full group of
    that kind is marked; no marked task is left out of every launch."""
    lock = (tasks[:, 3] >> 12) & 1
    launched = np.zeros(len(tasks), bool)
    nmarked = 0
    for g0, g1 in launch_groups(ranges, waves):
        launched[g0:g1] = True
        grp = tasks[g0:g1]
        same = (g1 - g0 == waves and (grp[:, 1] == grp[0, 1]).all() and (grp[:, 2] == grp[0, 2]).all()
                and (owner is None or (owner[g0:g1] == owner[g0]).all()))
        if lock[g0:g1].any():
            assert lock[g0:g1].all() and same, f"workgroup at {g0}: waves would meet different barriers"
            nmarked += 1
        else:
            assert not same, f"workgroup at {g0}: could sweep in step but is not marked"
    assert not (lock.astype(bool) & ~launched).any()
    return nmarked


def wide_case(seed):
    """A channel wide enough for joined strips, with bodies and a wall block."""
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(600, 1500)), int(rng.integers(40, 220))
    bodies = [(rng.uniform(0.1, 0.9) * nx, rng.uniform(0.25, 0.75) * ny, rng.uniform(2.0, 0.2 * ny))
              for _ in range(int(rng.integers(1, 4)))]
    x0 = int(rng.integers(5, nx - 40))
    y0 = int(rng.integers(3, ny - 12))
    return channel_flags(nx, ny, bodies, [(x0, x0 + int(rng.integers(2, 30)), y0, y0 + int(rng.integers(2, 9)))])


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("k,v,rows,waves", [(5, 1, 20, 8), (4, 1, 16, 4), (3, 2, 24, 8), (2, 1, 5, 2)])
def test_joined_plan_invariants(seed, k, v, rows, waves):
    """Joined strips (a workgroup's waves sweep one strip, handing each other their edge values)
    keep every invariant of the plan, and the rows a body takes from a joined strip's wide
    footprint go to single-wave strips over its columns, not to the single-step list."""
    fl = wide_case(seed)
    ntask, nfar, nlisted, ncov = check_plan(fl, k, v, rows, 32, waves=waves)
    _, _, nlisted1, ncov1 = check_plan(fl, k, v, rows, 32)
    tasks, _, _, _ = api.plan_passes(vl.backend(), fl, k, v, rows, 32, waves=waves, join=True)
    gout = (64 * v * waves - tb_halo(k, v) - k) // v * v
    if fl.shape[1] >= 32 + gout + 2 * k:
        assert ((tasks[:, 3] >> 13) & 1).any(), "some joined strips on a channel this wide"
    # the fallback strips lie on the single-wave grid, cut at the joined strips' ends, so they
    # leave no more single-step nodes than the single-wave plan
    assert nlisted <= nlisted1 and ncov >= ncov1


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("k,v,rows,waves", [(5, 1, 20, 8), (4, 1, 16, 4), (3, 2, 24, 8), (2, 1, 3, 8)])
@pytest.mark.parametrize("join", [False, True])
def test_lockstep_groups_of_the_launched_ranges(seed, k, v, rows, waves, join):
    """vl_sim_create's lockstep bits hold over the ranges it launches, far and near tasks apart
    (also when the far count is no multiple of the workgroup width), with and without joined
    strips."""
    fl = wide_case(100 + seed) if join else random_case(100 + seed)
    tasks, nfar, _, _ = api.plan_passes(vl.backend(), fl, k, v, rows, 32, waves=waves, join=join)
    ranges = [(0, nfar), (nfar, len(tasks))]
    check_lockstep(tasks, ranges, waves)
    # the same bits re-derived by the host-only entry agree, and pass its own check
    again = api.pass_lockstep(vl.backend(), tasks, ranges, waves)
    assert np.array_equal(again, tasks)
    api.pass_lockstep(vl.backend(), tasks, ranges, waves, check_only=True)


@pytest.mark.parametrize("nx,ny,r,k,rows", [(1400, 300, 20, 5, 20), (800, 200, 12, 4, 16), (1200, 400, 50, 3, 12)])
def test_lockstep_check_rejects_groups_across_launch_ranges(nx, ny, r, k, rows):
    """Bits grouped over the far / near split but launched as one list (the mistake the check
    guards against) leave a workgroup whose waves meet different barriers: on these channels the
    far count is no multiple of 8."""
    be = vl.backend()
    fl = channel_flags(nx, ny, [(0.3 * nx, 0.5 * ny, r)])
    tasks, nfar, _, _ = api.plan_passes(be, fl, k, 1, rows, 32, waves=8)
    assert nfar % 8 != 0
    with pytest.raises(AssertionError):
        check_lockstep(tasks, [(0, len(tasks))], 8)
    with pytest.raises(api.VortexError):
        api.pass_lockstep(be, tasks, [(0, len(tasks))], 8, check_only=True)
    api.pass_lockstep(be, tasks, [(0, nfar), (nfar, len(tasks))], 8, check_only=True)


@pytest.mark.parametrize("seed", range(4))
def test_lockstep_groups_of_a_batch(seed):
    """The batch's task list: each case's tasks in (row, column) order, cases one after another;
    a group's pass launches a prefix ending at a case boundary.  Marked workgroups never span two
    cases, nor a prefix end."""
    be = vl.backend()
    rng = np.random.default_rng(seed)
    parts, owner = [], []
    for c in range(int(rng.integers(3, 9))):
        fl = random_case(300 + 10 * seed + c)
        t, _, _, _ = api.plan_passes(be, fl, 5, 1, 20, 32)
        t = t[np.lexsort((t[:, 0], t[:, 1]))]
        parts.append(t)
        owner += [c] * len(t)
    tasks = np.concatenate(parts)
    owner = np.array(owner, dtype=np.int32)
    ends = np.cumsum([len(p) for p in parts])
    ranges = [(0, int(e)) for e in ends]
    marked = api.pass_lockstep(be, tasks, ranges, 8, owner=owner)
    for r in ranges:
        check_lockstep(marked, [r], 8, owner)
    # without the owner, groups of equal rows may span two cases: the check with owners refuses
    # such bits whenever they occur
    loose = api.pass_lockstep(be, tasks, ranges, 8)
    spans = any(len(set(owner[g:g + 8])) > 1 and (loose[g:g + 8, 3] >> 12 & 1).all()
                for g in range(0, len(tasks) - 7, 8))
    if spans:
        with pytest.raises(api.VortexError):
            api.pass_lockstep(be, loose, ranges, 8, owner=owner, check_only=True)
