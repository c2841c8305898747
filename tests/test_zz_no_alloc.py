"""mf_operator_compute_inverse_diagonal_noalloc through the C ABI against the allocating
mf_operator_compute_inverse_diagonal: bit-identical inverse diagonals on two calls in a row, the
second on a smaller problem, for operators that use the relative threshold (the reduction whose
partials sit in the static scratch array) and one that does not."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _advection(mf, degree, ncell, dt, seed):
    s = mf.ScratchData(degree, ncell, stream=torch.cuda.current_stream())
    s.set_dirichlet_mask(s.boundary_mask())
    op = mf.AdvectionDiffusionOperator(s, theta=0.5, diffusivity=0.01, mass_factor=1.0, dt=dt)
    vel = torch.tensor(np.random.default_rng(seed).uniform(-1, 1, 3 * s.n_dofs),
                       dtype=torch.float64, device="cuda")
    op.set_fields(velocity=vel)
    return s, op, vel


def _both(mf, op):
    lib = mf.lib()
    ref = torch.full((op.n_rows,), float("nan"), dtype=torch.float64, device="cuda")
    got = torch.full_like(ref, float("nan"))
    print(f"noalloc: {type(op).__name__} rows={op.n_rows}", flush=True)
    assert lib.mf_operator_compute_inverse_diagonal(op._h, ref.data_ptr()) == 0
    rc = lib.mf_operator_compute_inverse_diagonal_noalloc(op._h, got.data_ptr())
    assert rc == 0, mf.lib().mf_last_error()
    torch.cuda.synchronize()
    assert torch.isfinite(got).all()
    assert torch.equal(got, ref)
    return got


def test_noalloc_inverse_diagonal_matches_allocating_entry_point():
    import meltpooldg_amd as mf
    # first call: a large diagonal (small dt), the second a smaller problem with other entries
    s1, op1, v1 = _advection(mf, 2, 10, 1e-5, 1)
    d1 = _both(mf, op1)
    s2, op2, v2 = _advection(mf, 1, 4, 0.3, 2)
    d2 = _both(mf, op2)
    assert op2.n_rows < op1.n_rows
    # the normal vector operator (three blocks, relative threshold) and a heat operator
    # (absolute threshold: no scratch) after them
    nv = mf.NormalVectorOperator(s2, filter_parameter=2.0)
    _both(mf, nv)
    heat = mf.HeatDiffuseOperator(s2, rho_cp=1.0, k=1.0, dt=0.1)
    _both(mf, heat)
    # repeated on the first operator: the shared array keeps no state that matters
    assert torch.equal(_both(mf, op1), d1)


def test_noalloc_inverse_diagonal_rejects_null():
    import meltpooldg_amd as mf
    assert mf.lib().mf_operator_compute_inverse_diagonal_noalloc(None, None) == mf.MF_ERR_ARG
