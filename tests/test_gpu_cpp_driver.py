"""A compiled C++ program that uses the library through include/meltpooldg_amd.hpp (the C++
layer a MeltPoolDG caller would include) runs one vmult and one Jacobi-GMRES solve of the tanh
heat Jacobian on the GPU; its results are checked against the oracle: vmult within 1e-12,
the solution within 1e-10 with the oracle's iteration count +-1, and max_iterations = 1 throws
mf::NoConvergence."""
import os
import subprocess

import numpy as np
import pytest

import orc
from mfcases import rel_err

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hpp_driver")


def test_cpp_caller_vmult_and_solve(tmp_path):
    assert os.path.exists(DRIVER), "tests/cpp/hpp_driver is built by __graft_entry__.build()"
    deg, nc = 2, (6, 5, 4)
    m = orc.make_mesh(3, deg, nc)
    n = orc.n_dofs(m)
    mask = orc.boundary_mask(m)
    T = 1700.0 + 150.0 * orc.random_field(m, 21)
    src = orc.random_field(m, 22)
    b = orc.random_field(m, 23)
    b[mask == 1] = 0.0
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([deg, *nc], dtype=np.int32).tofile(f)
        mask.astype(np.uint8).tofile(f)
        for a in (T, src, b):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    print(f"cpp driver: Q{deg} {nc}", flush=True)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    assert int(np.frombuffer(raw[:8], dtype=np.int64)[0]) == n
    dst = np.frombuffer(raw[8:8 + 8 * n], dtype=np.float64)
    x = np.frombuffer(raw[8 + 8 * n:8 + 16 * n], dtype=np.float64)
    its, threw = np.frombuffer(raw[8 + 16 * n:], dtype=np.int32)
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                        tanh_Tc=1700.0, tanh_width=50.0, dt=1e-3)
    ref = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, mask=mask, T=T, T_old=T)
    assert rel_err(dst, ref.vmult(src)) < 1e-12
    x_ref, it_ref, _, st = ref.solve(b, solver=orc.GMRES, rel_tol=1e-10,
                                     inv_diag=ref.inverse_diagonal())
    assert st == 0 and abs(int(its) - it_ref) <= 1, (its, it_ref)
    assert rel_err(x, x_ref) < 1e-10
    assert threw == 1


CUT_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "cut_driver")
GOLDEN_ONE_PHASE_CUT = [4.04890e-03, 4.01890e-03, 3.99116e-03, 3.96522e-03, 3.94086e-03]


def test_cpp_caller_cut_heat_steps(tmp_path):
    """tests/cpp/cut_driver runs the one_phase_cut golden case (extruded to 3D) through
    mf::HeatCutOperator: its norms match the reference's output to the printed digits (+-1 in
    the sixth, see tests/test_oracle_cut.py) and the oracle's loop to 1e-9."""
    assert os.path.exists(CUT_DRIVER), "tests/cpp/cut_driver is built by __graft_entry__.build()"
    out = tmp_path / "cut.bin"
    print("cpp cut driver: one_phase_cut golden, 17 x 17 x 1", flush=True)
    r = subprocess.run([CUT_DRIVER, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    ns = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    norms = np.frombuffer(raw[4:4 + 8 * ns], dtype=np.float64)
    np.testing.assert_allclose(norms, GOLDEN_ONE_PHASE_CUT, rtol=3e-6)
    m = orc.make_mesh(3, 1, (17, 17, 1), (0.0, 0.0, 0.0), (0.1, 0.1, 1.0))
    X = orc.node_coords(m)
    p = orc.cut_params(2.0, rho_cp=7850.0 * 460.0, k=55.563, theta=0.5, gamma_A=1.5,
                       gamma_M=0.75, h=0.1 / 17)
    _, ref = orc.cut_time_steps(m, 0.05 - X[:, 1], p, X[:, 0].copy(), ns)
    np.testing.assert_allclose(norms, ref, rtol=1e-9)
