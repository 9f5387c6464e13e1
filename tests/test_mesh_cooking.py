"""The library's host-side mesh cooking (pxg_cook_mesh: TriangleMesh::from_triangles,
geometry/triangle_mesh.h:75-170) against the oracle's restatement, bit for bit: welded vertices,
triangle indices, face normals, adjacency, bounds and the non-manifold edge count, for the
reference tests' meshes, jittered soups and random soups with duplicate, collapsing and
non-manifold triangles. Host code only: no device is touched."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi


def soups():
    rng = np.random.default_rng(17)
    out = [scenes.tiled_floor(16), scenes.ledge(), scenes.bumpy_terrain(), scenes.two_triangle_floor()]
    g = scenes.tiled_floor(9)
    soup = [g.positions[i] + rng.uniform(-2e-5, 2e-5, 3).astype(np.float32) for i in g.indices]
    out.append(_abi.TriangleMesh(soup, np.arange(len(soup))))
    for _ in range(4):
        p = rng.normal(size=(60, 3)).astype(np.float32)
        p = np.concatenate([p, p[:20] + rng.uniform(-1e-5, 1e-5, (20, 3)).astype(np.float32)])
        ind = rng.integers(0, len(p), 3 * 150)
        ind[:30] = np.repeat(ind[:10], 3)  # collapsed triangles
        ind[30:39] = [0, 1, 2, 1, 0, 3, 0, 1, 4]  # an edge owned three times
        out.append(_abi.TriangleMesh(p, ind))
    return out


@pytest.mark.parametrize("k", range(9))
def test_cooking_matches_the_oracle_bit_for_bit(oracle, k):
    from physxcoopa_amd.world import cook_mesh
    mesh = soups()[k]
    want = oracle.cook_mesh(mesh)
    got = cook_mesh(mesh)
    assert got["vertices"].tobytes() == want.vertices.tobytes()
    assert got["indices"].tobytes() == want.indices.tobytes()
    assert got["normals"].tobytes() == want.normals.tobytes()
    assert got["adjacency"].tobytes() == want.adjacency.tobytes()
    assert got["bounds"].tobytes() == want.bounds.tobytes()
    assert got["non_manifold_edges"] == want.non_manifold_edges
    assert got["n_nodes"] == len(want.node_bounds)


def test_cooking_refuses_bad_input():
    import ctypes
    from physxcoopa_amd._lib import lib
    from physxcoopa_amd.world import cook_mesh
    with pytest.raises(RuntimeError):
        cook_mesh(_abi.TriangleMesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [0, 1, 3]))
    d = _abi.MeshDesc()
    d.n_positions = 3  # positions missing
    assert lib().pxg_cook_mesh(ctypes.byref(d), None, None, None, None, None) == _abi.PXG_ERR_INVALID
