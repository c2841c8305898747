"""Host side of the adaptive-mesh setup in libmeltpooldg_amd (no GPU needed): the band
refinement with 2:1 balance, the node numbering and the hanging-node constraint lines, each
against the oracle's independent construction (Python mesh generator, coordinate-keyed nodes,
point-location constraints)."""
import numpy as np
import pytest

import orc

CASES = [
    (1, 2, 2, (0.1, 0.1, 0.1), 0.2, 0.03),
    (2, 3, 2, (0.2, 0.3, 0.25), 0.3, 0.04),
    (3, 2, 2, (0.1, 0.1, 0.1), 0.2, 0.03),
    (3, 4, 1, (0.5, 0.5, 0.5), 0.3, 0.05),
]


def _lib():
    import meltpooldg_amd as mf
    try:
        mf.lib()
    except ImportError as e:  # pragma: no cover
        pytest.skip(str(e))
    return mf


@pytest.mark.parametrize("case", CASES)
def test_band_refinement_matches_reference_generator(case):
    mf = _lib()
    _, nc, lev, ce, r, hw = case
    got = mf.refine_band_cells(nc, lev, ce, r, hw)
    ref = orc.refine_band_cells(nc, lev, ce, r, hw)
    assert len(got) == len(ref)
    assert set(map(tuple, got)) == set(map(tuple, ref))


def _key(x):
    return tuple(np.round(np.asarray(x) * 2 ** 24).astype(np.int64))


@pytest.mark.parametrize("case", CASES)
def test_nodes_and_constraints_match_oracle(case):
    mf = _lib()
    degree, nc, lev, ce, r, hw = case
    cells = mf.refine_band_cells(nc, lev, ce, r, hw)
    s = mf.AdaptiveScratchData(degree, nc, cells)
    am = orc.AdaptiveMesh(degree, nc, cells)
    assert s.n_dofs == am.n_nodes and s.n_cells == len(cells)
    xs = s.support_points()
    pos = {_key(x): i for i, x in enumerate(am.xyz)}
    perm = np.array([pos[_key(x)] for x in xs])          # device node -> oracle node
    assert len(set(perm.tolist())) == s.n_dofs
    node_line, ptr, col, val = s.constraint_lines()
    hanging = node_line >= 0
    assert np.array_equal(hanging.astype(np.uint8), am.hanging[perm])
    assert s.constraint_info()["n_hanging"] == int(am.hanging.sum())
    # every line: same masters (mapped) and weights
    for i in np.nonzero(hanging)[0]:
        l = node_line[i]
        mine = {int(perm[c]): v for c, v in zip(col[ptr[l]:ptr[l + 1]], val[ptr[l]:ptr[l + 1]])}
        o = perm[i]
        ref = {int(c): v for c, v in zip(am.line_col[am.line_ptr[o]:am.line_ptr[o + 1]],
                                         am.line_val[am.line_ptr[o]:am.line_ptr[o + 1]])}
        ref = {c: v for c, v in ref.items() if abs(v) > 1e-15}
        mine = {c: v for c, v in mine.items() if abs(v) > 1e-15}
        assert mine.keys() == ref.keys()
        assert max(abs(mine[c] - ref[c]) for c in ref) < 1e-13
    assert np.array_equal(s.default_dirichlet_mask(), am.dirichlet[perm])
    # cell node tables agree through the map
    assert 1 <= s.constraint_info()["n_colours"] <= 64


def test_invalid_adaptive_meshes_are_rejected():
    mf = _lib()
    with pytest.raises(mf.MFError):
        mf.AdaptiveScratchData(2, 2, np.array([[0, 0, 0, 0], [0, 0, 0, 0]]))   # duplicate
    with pytest.raises(mf.MFError):
        mf.AdaptiveScratchData(2, 2, np.array([[0, 5, 0, 0]]))                  # outside
    # a level-0 cell touching level-2 cells: hanging nodes would need a chain through a leaf
    # that does not exist, and the class colouring would not be conflict-free
    fine = [[2, 4 + i, j, k] for i in range(4) for j in range(4) for k in range(4)]
    with pytest.raises(mf.MFError, match="2:1"):
        mf.AdaptiveScratchData(2, (2, 1, 1), np.array([[0, 0, 0, 0]] + fine), hi=(2.0, 1.0, 1.0))
    # a level-0 cell next to level-1 cells (one level apart) is accepted
    mid = [[1, 2 + i, j, k] for i in range(4) for j in range(2) for k in range(2)]
    s = mf.AdaptiveScratchData(2, (3, 1, 1), np.array([[0, 0, 0, 0]] + mid), hi=(3.0, 1.0, 1.0))
    assert s.constraint_info()["n_hanging"] > 0
