"""Device adaptive setup (mf_adaptive_dev.hip: band refinement with 2:1 balance, node
identification, hanging-node constraint lines, their transposition, the hanging-cell tables)
against the host path: the same cells, DoF numbering, support points, boundary flags,
hanging-node constraints, Dirichlet rows and operator (bit for bit)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = {
    "q1": (1, 2, 2, (0.1, 0.1, 0.1), 0.2, 0.03, (1.0, 1.0, 1.0)),
    "q2": (2, 3, 2, (0.2, 0.3, 0.25), 0.3, 0.04, (1.0, 1.3, 0.8)),
    "q3": (3, 2, 2, (0.1, 0.1, 0.1), 0.2, 0.03, (1.0, 1.0, 1.0)),
    "q3b": (3, 4, 2, (0.5, 0.5, 0.5), 0.3, 0.05, (1.0, 1.0, 1.0)),
}


REFINE = [
    ((32, 32, 32), 2, (0.5, 0.5, 0.5), 0.3, 0.05, (1.0, 1.0, 1.0)),
    ((5, 7, 6), 3, (0.2, 0.3, 0.25), 0.3, 0.04, (1.0, 1.3, 0.8)),
    ((4, 4, 4), 4, (0.1, 0.1, 0.1), 0.2, 0.03, (1.0, 1.0, 1.0)),
    ((3, 3, 3), 0, (0.5, 0.5, 0.5), 0.3, 0.05, (1.0, 1.0, 1.0)),
    ((6, 6, 6), 2, (5.0, 5.0, 5.0), 0.1, 0.01, (1.0, 1.0, 1.0)),   # band misses the box
]


@pytest.mark.parametrize("case", REFINE)
def test_device_refinement_matches_host(case):
    import meltpooldg_amd as mf
    nc, lev, ce, r, hw, hi = case
    try:
        mf.set_adaptive_setup(0)
        ch = mf.refine_band_cells(nc, lev, ce, r, hw, hi=hi)
    finally:
        mf.set_adaptive_setup(1)
    cd = mf.refine_band_cells(nc, lev, ce, r, hw, hi=hi)
    assert cd.shape == ch.shape
    assert np.array_equal(cd, ch)


@pytest.mark.parametrize("name", list(CASES))
def test_device_setup_matches_host(name):
    import meltpooldg_amd as mf
    degree, nc, lev, ce, r, hw, hi = CASES[name]
    st = torch.cuda.current_stream()
    try:
        mf.set_adaptive_setup(0)
        cells = mf.refine_band_cells(nc, lev, ce, r, hw, hi=hi)
    finally:
        mf.set_adaptive_setup(1)
    try:
        mf.set_adaptive_setup(0)
        sh = mf.AdaptiveScratchData(degree, nc, cells, hi=hi, stream=st)
    finally:
        mf.set_adaptive_setup(1)
    sd = mf.AdaptiveScratchData(degree, nc, cells, hi=hi, stream=st)
    assert sd.n_dofs == sh.n_dofs
    assert np.array_equal(sd.support_points(), sh.support_points())
    assert np.array_equal(sd.default_dirichlet_mask(), sh.default_dirichlet_mask())
    assert sd.constraint_info() == sh.constraint_info()
    for a, b in zip(sd.constraint_lines(), sh.constraint_lines()):
        assert np.array_equal(a, b)
    x = torch.tensor(np.random.default_rng(3).uniform(-1, 1, sd.n_dofs), dtype=torch.float64,
                     device="cuda")
    ys, ds = [], []
    for s in (sd, sh):
        op = mf.HeatDiffuseOperator(s, rho_cp=1.3, k=0.7, dt=1e-2)
        y = torch.zeros_like(x)
        op.vmult(y, x)
        ys.append(y)
        d = torch.zeros_like(x)
        op.compute_inverse_diagonal_from_matrixfree(d)
        ds.append(d)
    assert torch.equal(ys[0], ys[1])
    # the condensed diagonal adds the hanging-cell slots with atomics: equal up to the order
    torch.testing.assert_close(ds[0], ds[1], rtol=1e-13, atol=0)


def test_device_setup_rejects_unbalanced_mesh():
    import meltpooldg_amd as mf
    fine = [[2, 4 + i, j, k] for i in range(4) for j in range(4) for k in range(4)]
    with pytest.raises(mf.MFError, match="2:1"):
        mf.AdaptiveScratchData(2, (2, 1, 1), np.array([[0, 0, 0, 0]] + fine), hi=(2.0, 1.0, 1.0),
                               stream=torch.cuda.current_stream())
