"""topologyIsSane (A4) through the C-ABI host check, and the oracle's
restatements of it and of MimeticIPEvaluator::gravityTerm (A24), pinned by
closed forms and constructed counter-examples."""
import numpy as np
import pytest



def _dup_neighbour(g, c, rng):
    """Copy of g.nb in which one internal face of cell c points at another of
    c's neighbours (two faces between the same pair of cells)."""
    nb = np.array(g.nb, dtype=np.int32)
    hs = [h for h in range(g.facepos[c], g.facepos[c + 1]) if nb[h] >= 0]
    a, b = rng.choice(hs, 2, replace=False)
    nb[a] = nb[b]
    return nb


@pytest.mark.parametrize("kind", ["cartesian", "corner_point"])
def test_topology_is_sane_matches_oracle(pkg, orc, kind):
    g = pkg.grids.cartesian(5, 4, 3) if kind == "cartesian" else \
        pkg.grids.corner_point(6, 5, 4, 10.0, 10.0, 2.0, seed=3)
    assert orc.topology_is_sane(g.num_cells, g.facepos, g.nb)
    assert pkg.mhs.topology_is_sane(g)
    rng = np.random.default_rng(5)
    for c in rng.choice(g.num_cells, 6, replace=False):
        nb = _dup_neighbour(g, int(c), rng)

        class G2:
            num_cells, facepos = g.num_cells, g.facepos
        G2.nb = nb
        assert not orc.topology_is_sane(g.num_cells, g.facepos, nb)
        assert not pkg.mhs.topology_is_sane(G2)


def test_topology_two_cells_sharing_two_faces(pkg, orc):
    # two cells of three faces each: two faces shared, one boundary face each
    class G:
        num_cells = 2
        facepos = np.array([0, 3, 6])
        nb = np.array([1, 1, -1, 0, 0, -1])
    assert not orc.topology_is_sane(2, G.facepos, G.nb)
    assert not pkg.mhs.topology_is_sane(G)
    G.nb = np.array([1, -1, -1, 0, -1, -1])
    assert orc.topology_is_sane(2, G.facepos, G.nb)
    assert pkg.mhs.topology_is_sane(G)


def test_oracle_gravity_term_closed_form(pkg, orc):
    """Box cells: the face centroid sits half a cell width from the cell
    centroid along the face's axis, so gterm = omega * (+-d/2) * g_axis."""
    sp = (2.0, 3.0, 0.5)
    g = pkg.grids.cartesian(3, 2, 2, *sp)
    rng = np.random.default_rng(1)
    om = rng.uniform(0.5, 2.0, g.num_cells)
    grav = np.array([0.3, -1.1, 9.81])
    got = orc.gravity_term(g, grav, om)
    cell = g.hf_cell()
    side = np.arange(g.num_half_faces) - np.asarray(g.facepos)[cell]   # local face: x-, x+, y-, y+, z-, z+
    axis, sign = side // 2, np.where(side % 2 == 0, -1.0, 1.0)
    want = om[cell] * sign * np.asarray(sp)[axis] / 2.0 * grav[axis]
    assert np.allclose(got, want, rtol=1e-14, atol=1e-14)
    # the gravity terms of opposite faces of a box cancel
    for c in range(g.num_cells):
        assert abs(got[g.facepos[c]:g.facepos[c + 1]].sum()) < 1e-12
