"""The host voxel rule (svdag_gen_voxelize, the rule of the triangle-surface
workload and of the device voxeliser) against its geometric definition.

A voxel (i, j, k) of a 2^d grid is emitted for a triangle with normal N when
(a) along the dominant axis a of N (largest |N_a|; ties to x, then y), the
    segment through the voxel centre from face to face meets the triangle's
    plane (closed: a plane through a face emits both voxels);
(b) in each axis plane (xy, yz, zx) where the triangle is not edge-on, every
    projected edge has some vertex of the voxel centre's half-voxel diamond
    (the centre moved by 1/2 along either in-plane axis) on the side of the
    edge where the triangle lies, or on the edge;
(c) along the two other axes the voxel lies within one voxel of the
    triangle's bounding box.

Pinned two ways, integer output with tolerance 0:
- exhaustively over every voxel of small grids, in exact integer arithmetic
  (coordinates on a 1/8 lattice scaled by 16), for seeded soups of small,
  large, sliver, axis-aligned, degenerate, out-of-grid and tied-normal
  triangles (the dominant axis then follows the tie order); the
  lattice makes plane-through-face and centre-on-edge ties common, so the
  closed boundaries are exercised;
- closed forms: a right triangle (legs L) parallel to an axis plane covers the
  voxels with i + j <= L - 1/2 in one layer, two layers when it lies on a face;
  a triangle far larger than the grid on a tilted plane covers exactly the
  voxel k = floor(plane(centre)) of every column.
"""
import numpy as np
import pytest

S = 16  # scale: coordinates on a 1/8 lattice and voxel centres become integers


def decode(keys, depth):
    keys = np.asarray(keys, dtype=np.uint64)
    xyz = np.zeros((len(keys), 3), dtype=np.int64)
    for b in range(depth):
        for a in range(3):
            xyz[:, a] |= ((keys >> np.uint64(3 * b + 2 - a)) & np.uint64(1)).astype(np.int64) << b
    return {tuple(v) for v in xyz}


def rule(tris, depth):
    import dag_compression_fast_amd as host
    return decode(host.gen_voxelize(np.asarray(tris, dtype=np.float64), depth, threads=2), depth)


def cross2(ax, ay, bx, by):
    return ax * by - ay * bx


def brute_force(tri, side):
    """Voxels of one triangle by the definition in the module docstring,
    evaluated at every voxel of the grid in exact integer arithmetic."""
    P = np.rint(np.asarray(tri, dtype=np.float64) * S).astype(np.int64)
    assert np.array_equal(P, np.asarray(tri) * S), "vertices must lie on the 1/8 lattice"
    N = np.cross(P[1] - P[0], P[2] - P[0])
    if not N.any():
        return set()
    aN = np.abs(N)
    a = 0 if aN[0] >= aN[1] and aN[0] >= aN[2] else (1 if aN[1] >= aN[2] else 2)
    g = np.indices((side, side, side)).reshape(3, -1).T.astype(np.int64)  # voxel indices
    C = S * g + S // 2  # centres
    keep = np.ones(len(g), dtype=bool)
    # (a) the plane meets the centre segment along a: opposite signs (or a zero)
    # at the two face centres.
    lo, hi = C.copy(), C.copy()
    lo[:, a] -= S // 2
    hi[:, a] += S // 2
    keep &= np.sign((lo - P[0]) @ N) * np.sign((hi - P[0]) @ N) <= 0
    # (b) per axis plane (i0, i1), unless edge-on: each edge has a diamond
    # vertex on the triangle's side of it (the side of the third vertex).
    for i0, i1 in ((0, 1), (1, 2), (2, 0)):
        for e in range(3):
            A0, A1, V = P[e], P[(e + 1) % 3], P[(e + 2) % 3]
            ex, ey = A1[i0] - A0[i0], A1[i1] - A0[i1]
            inner = np.sign(cross2(ex, ey, V[i0] - A0[i0], V[i1] - A0[i1]))
            if inner == 0:
                break  # the projection is a segment or a point: edge-on, no test
            ok = np.zeros(len(g), dtype=bool)
            for dx, dy in ((S // 2, 0), (-S // 2, 0), (0, S // 2), (0, -S // 2)):
                ok |= inner * cross2(ex, ey, C[:, i0] + dx - A0[i0], C[:, i1] + dy - A0[i1]) >= 0
            keep &= ok
    # (c) within one voxel of the bounding box along the other two axes.
    for b in ((a + 1) % 3, (a + 2) % 3):
        lo_b = int(np.floor(P[:, b].min() / S)) - 1
        hi_b = int(np.floor(P[:, b].max() / S)) + 1
        keep &= (g[:, b] >= lo_b) & (g[:, b] <= hi_b)
    return {tuple(v) for v in g[keep]}


def lattice(x):
    return np.round(np.asarray(x, dtype=np.float64) * 8) / 8


def soup(rng, side, count):
    out = []
    for i in range(count):
        kind = i % 7
        if kind == 0:  # small, inside
            v0 = rng.uniform(0, side, 3)
            tri = [v0, v0 + rng.uniform(-3, 3, 3), v0 + rng.uniform(-3, 3, 3)]
        elif kind == 1:  # large, straddling the grid
            tri = [rng.uniform(-4, side + 4, 3) for _ in range(3)]
        elif kind == 2:  # sliver: a long edge and a nearby third vertex
            v0, v1 = rng.uniform(0, side, 3), rng.uniform(0, side, 3)
            tri = [v0, v1, v0 + (v1 - v0) * rng.uniform(0.2, 0.8) + rng.uniform(-0.4, 0.4, 3)]
        elif kind == 3:  # parallel to an axis plane
            tri = [rng.uniform(0, side, 3) for _ in range(3)]
            ax = rng.integers(0, 3)
            for t in tri:
                t[ax] = tri[0][ax]
        elif kind == 4:  # an edge parallel to an axis
            v0 = rng.uniform(0, side, 3)
            v1 = v0.copy()
            v1[rng.integers(0, 3)] = rng.uniform(0, side)
            tri = [v0, v1, rng.uniform(0, side, 3)]
        elif kind == 5:  # degenerate on the lattice: collinear, or a point
            v0, d = rng.uniform(0, side, 3), lattice(rng.uniform(-2, 2, 3))
            tri = [v0, v0 + d, v0 + 2 * d] if i % 14 == 5 else [v0, v0, v0]
        else:  # normal with tied largest components: |Nx| = |Ny|, or all equal
            v0 = rng.uniform(0, side, 3)
            d, e = lattice(rng.uniform(0.5, 4, 2)) * rng.choice([-1, 1], 2)
            p = rng.permutation(3)
            u1 = np.zeros(3)
            u1[p[[0, 1]]] = (d, -d)
            u2 = np.zeros(3)
            if i % 14 == 6:
                u2[p[2]] = e
            else:
                u2[p[[0, 2]]] = (e, -e)
            tri = [v0, v0 + u1, v0 + u2]
        out.append(lattice(tri))
    return np.array(out, dtype=np.float64)


@pytest.mark.parametrize("depth,count,seed", [(3, 300, 1), (4, 240, 2)])
def test_rule_equals_exhaustive_exact_definition(depth, count, seed):
    side = 1 << depth
    tris = soup(np.random.default_rng(seed), side, count)
    ties = 0
    for i, tri in enumerate(tris):
        exp = brute_force(tri, side)
        got = rule(tri[None], depth)
        assert got == exp, f"triangle {i} {tri.tolist()}: rule - definition {sorted(got - exp)[:5]}, " \
                           f"definition - rule {sorted(exp - got)[:5]}"
        ties += len(exp) > 0
    assert ties > count // 2, "most triangles emit voxels"
    # The soup as one call: the union.
    union = set().union(*(brute_force(t, side) for t in tris))
    assert rule(tris, depth) == union


def right_triangle(L, layer, axes, ccw=True):
    """Right triangle with legs L from the origin in the plane of axes (i0, i1)
    at coordinate `layer` along the third axis."""
    pts = [(0.0, 0.0), (L, 0.0), (0.0, L)]
    if not ccw:
        pts = pts[::-1]
    out = []
    for p, q in pts:
        v = [layer] * 3
        v[axes[0]], v[axes[1]] = p, q
        out.append(v)
    return np.array(out, dtype=np.float64)


@pytest.mark.parametrize("axes", [(0, 1), (1, 2), (2, 0), (1, 0)])
@pytest.mark.parametrize("ccw", [True, False])
def test_right_triangle_closed_form(axes, ccw):
    depth, side = 4, 16
    third = 3 - axes[0] - axes[1]
    for L, layer, layers in ((10.0, 2.5, (2,)), (10.25, 2.5, (2,)), (10.5, 5.75, (5,)), (10.0, 3.0, (2, 3)),
                             (3.0, 0.0, (0,)), (40.0, 7.5, (7,))):
        exp = set()
        for i in range(side):
            for j in range(side):
                if i + j <= L - 0.5:
                    for k in layers:
                        v = [0, 0, 0]
                        v[axes[0]], v[axes[1]], v[third] = i, j, k
                        exp.add(tuple(v))
        got = rule(right_triangle(L, layer, axes, ccw)[None], depth)
        assert got == exp, f"L={L} layer={layer}: {len(got)} voxels, closed form {len(exp)}"
    # Triangular numbers: legs 10 -> 55 voxels, legs 10.5 -> 66 (x + y = 10 on the closed boundary).
    assert len(rule(right_triangle(10.0, 2.5, axes, ccw)[None], depth)) == 55
    assert len(rule(right_triangle(10.5, 2.5, axes, ccw)[None], depth)) == 66


@pytest.mark.parametrize("seed", range(6))
def test_big_tilted_triangle_one_voxel_per_column(seed):
    rng = np.random.default_rng(100 + seed)
    depth, side = 4, 16
    alpha, beta = rng.uniform(0.3, 0.9, 2) * rng.choice([-1, 1], 2)
    gamma = side / 2 + rng.uniform(-0.5, 0.5)
    perm = rng.permutation(3)  # the plane's dominant axis

    def plane(x, y):
        return alpha * x + beta * y + gamma

    tri = np.zeros((3, 3))
    for r, (x, y) in enumerate(((-400.0, -400.0), (1200.0, -400.0), (-400.0, 1200.0))):
        tri[r, perm] = (x, y, plane(x, y))
    exp = set()
    for i in range(side):
        for j in range(side):
            k = int(np.floor(plane(i + 0.5, j + 0.5)))
            if 0 <= k < side:
                v = np.zeros(3, dtype=np.int64)
                v[perm] = (i, j, k)
                exp.add(tuple(v))
    got = rule(tri[None], depth)
    assert got == exp
