"""CPU-side checks of the device library: it loads, exports every entry point the C header
declares, and its host logic (mesh setup, 1D tables, argument validation, synthetic inputs)
agrees with the oracle.  No GPU compute is called here."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import orc

mf = pytest.importorskip("meltpooldg_amd")


def test_library_exports_every_header_function():
    L = mf.lib()
    names = mf.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_tables_match_oracle(degree):
    s = mf.ScratchData(degree, (2, 3, 4), hi=(1.0, 2.0, 0.5))
    S, DS, Dq, w, gl = s.shape_info()
    S2, DS2, Dq2, w2 = orc.shape_matrices(degree)
    assert np.allclose(S, S2, atol=1e-14) and np.allclose(DS, DS2, atol=1e-13)
    assert np.allclose(Dq, Dq2, atol=1e-13) and np.allclose(w, w2, atol=1e-15)
    assert np.allclose(gl, orc.gauss_lobatto(degree + 1), atol=1e-15)
    # collocation identity used by the kernels: Dq S = DS
    assert np.allclose(Dq @ S, DS, atol=1e-12)
    m = orc.make_mesh(3, degree, (2, 3, 4), (0, 0, 0), (1.0, 2.0, 0.5))
    assert s.n_dofs == orc.n_dofs(m) and s.n_cells == orc.n_cells(m)
    assert np.array_equal(s.support_points(), orc.node_coords(m))
    assert np.array_equal(s.boundary_mask(), orc.boundary_mask(m))
    assert np.array_equal(s.random_field(9), orc.random_field(m, 9))


def test_argument_errors_are_reported():
    with pytest.raises(mf.MFError) as e:
        mf.ScratchData(4, 3)
    assert e.value.code == mf.MF_ERR_ARG and "degree" in str(e.value)
    with pytest.raises(mf.MFError):
        mf.ScratchData(2, (0, 1, 1))
    s = mf.ScratchData(2, 2)
    with pytest.raises(mf.MFError) as e:
        mf.AdvectionDiffusionOperator(s, diffusivity=-1.0)
    assert "diffusivity" in str(e.value)
    op = mf.HeatDiffuseOperator(s)
    with pytest.raises(mf.MFError) as e:
        op.reset_time_increment(0.0)
    assert e.value.code == mf.MF_ERR_ARG
    with pytest.raises(mf.MFError):
        op.reset_dof_index(2)
    # vmult before reset_time_increment is a state error (AssertThrowZeroTimeIncrement)
    L = mf.lib()
    code = L.mf_operator_vmult(op._h, C.c_void_p(8), C.c_void_p(8))
    assert code == mf.MF_ERR_STATE
    assert b"time increment" in L.mf_last_error()
    # tanh law without T is a state error
    op2 = mf.HeatDiffuseOperator(s, law=mf.LAW_TANH, tanh_width=50.0, dt=1e-3)
    assert L.mf_operator_vmult(op2._h, C.c_void_p(8), C.c_void_p(8)) == mf.MF_ERR_STATE


def test_too_large_mesh_rejected():
    with pytest.raises(mf.MFError):
        mf.ScratchData(3, (1000, 1000, 1000))


def test_boundary_argument_errors():
    """mf_operator_set_boundary validation (host side only)."""
    s = mf.ScratchData(2, (2, 2, 2))
    op = mf.HeatDiffuseOperator(s, dt=1e-3)
    with pytest.raises(mf.MFError) as e:
        op.set_boundary(convection_sides=[1], neumann_sides={1: 2.0})
    assert "Neumann" in str(e.value)
    with pytest.raises(mf.MFError):
        op.set_boundary(convection_sides=[6])
    with pytest.raises(mf.MFError):
        op.set_boundary(convection_sides=[0], convection_coefficient=-1.0)
    op.set_boundary(convection_sides=[0, 5], convection_coefficient=1.0)
    op.clear_boundary()
    adv = mf.AdvectionDiffusionOperator(s, dt=1e-3)
    L = mf.lib()
    b = mf.HeatBoundary()
    b.convection_sides = 1
    assert L.mf_operator_set_boundary(adv._h, C.byref(b)) == mf.MF_ERR_ARG


def test_levelset_projection_host_logic():
    """Level-set projections: argument validation, block sizes and the normalisation tolerance
    against the oracle's compute_numerical_zero_of_norm (host side only)."""
    s = mf.ScratchData(2, (3, 2, 2), lo=(0.0, 0.0, 0.0), hi=(1e-4, 2e-4, 5e-5))
    nop = mf.NormalVectorOperator(s)
    cop = mf.CurvatureOperator(s)
    n = s.n_dofs
    assert nop.n_rows == 3 * n and cop.n_rows == n
    vol = 1e-4 * 2e-4 * 5e-5
    assert cop.normal_tolerance == pytest.approx(orc.lib().orc_numerical_zero_of_norm(3, vol),
                                                 rel=1e-15)
    assert cop.normal_tolerance == pytest.approx(1e-7)
    assert mf.CurvatureOperator(mf.ScratchData(1, (2, 2, 2))).normal_tolerance == pytest.approx(1e-3)
    with pytest.raises(mf.MFError) as e:
        mf.NormalVectorOperator(s, narrow_band=True, narrow_band_threshold=0.0)
    assert "narrow band threshold" in str(e.value)
    with pytest.raises(mf.MFError):
        mf.CurvatureOperator(s, filter_parameter=-1.0)
    # a narrow band needs its level set before any application (no device work is started)
    band = mf.NormalVectorOperator(s, narrow_band=True)
    L = mf.lib()
    fake = C.c_void_p(256)
    assert L.mf_operator_vmult(band._h, fake, fake) == mf.MF_ERR_STATE
    assert "nullptr" in L.mf_last_error().decode()
    # the projections take their field from set_level_set only, and no heat source term
    assert L.mf_operator_set_fields(nop._h, None, None, None) == mf.MF_ERR_ARG
    assert L.mf_operator_create_rhs(nop._h, fake, fake, fake) == mf.MF_ERR_ARG
    heat = mf.HeatDiffuseOperator(s, dt=1.0)
    assert L.mf_operator_set_level_set(heat._h, None) == mf.MF_ERR_ARG


def test_reinit_operator_host_logic():
    """Olsson reinitialization operator: argument validation and the pseudo time step scale
    (eps of the smallest cell), host side only."""
    s = mf.ScratchData(2, (4, 2, 2), hi=(1.0, 0.5, 0.5))
    op = mf.ReinitializationOperator(s)
    h = np.sqrt((0.25 ** 2 * 3) / 3)   # diameter / sqrt(3) of a 0.25^3 cell
    assert op.min_epsilon == pytest.approx(0.5 * h)
    assert mf.ReinitializationOperator(s, thickness_type=mf.THICKNESS_ABSOLUTE,
                                       thickness_value=0.01).min_epsilon == pytest.approx(0.01)
    assert mf.ReinitializationOperator(s, thickness_type=mf.THICKNESS_CELLS,
                                       thickness_value=3.0).min_epsilon == pytest.approx(0.5 * h)
    with pytest.raises(mf.MFError) as e:
        mf.ReinitializationOperator(s, thickness_value=0.0)
    assert "larger than zero" in str(e.value)
    with pytest.raises(mf.MFError):
        mf.ReinitializationOperator(s, thickness_type=7)
    L = mf.lib()
    fake = C.c_void_p(256)
    # no time increment / no normal vector: errors before any device work
    assert L.mf_operator_vmult(op._h, fake, fake) == mf.MF_ERR_STATE
    op.reset_time_increment(0.1)
    assert L.mf_operator_vmult(op._h, fake, fake) == mf.MF_ERR_STATE
    assert "normal vector" in L.mf_last_error().decode()
    heat = mf.HeatDiffuseOperator(s, dt=1.0)
    assert L.mf_operator_set_normal_vector(heat._h, None) == mf.MF_ERR_ARG


def test_vtu_and_restart_without_fields(tmp_path):
    """Mesh-only VTU output and a restart file without vectors or mask need no device: the
    written mesh is parsed back, the restart rebuilds the same scratch."""
    import re
    s = mf.ScratchData(2, (3, 2, 1), lo=(0.0, 0.5, 0.0), hi=(1.5, 1.5, 0.25))
    path = tmp_path / "mesh.vtu"
    L = mf.lib()
    assert L.mf_write_vtu(s._h, str(path).encode(), 0, None, None, None) == mf.MF_OK
    raw = open(path, "rb").read()
    head = raw[:raw.index(b"<AppendedData")].decode()
    assert int(re.search(r'NumberOfPoints="(\d+)"', head).group(1)) == s.n_dofs
    assert int(re.search(r'NumberOfCells="(\d+)"', head).group(1)) == s.n_cells * 8
    off = int(re.search(r'Name="Points" NumberOfComponents="3" format="appended" '
                        r'offset="(\d+)"', head).group(1))
    us = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    nb = int(np.frombuffer(raw[us + off:us + off + 8], dtype=np.uint64)[0])
    pts = np.frombuffer(raw[us + off + 8:us + off + 8 + nb], dtype=np.float64).reshape(-1, 3)
    assert np.array_equal(pts, s.support_points())
    rp = tmp_path / "mesh.mfr"
    assert L.mf_restart_save(s._h, str(rp).encode(), 0, None, None) == mf.MF_OK
    s2 = mf.restart_load(rp)
    assert (s2.n_dofs, s2.n_cells, s2.degree, s2.ncell) == (s.n_dofs, s.n_cells, 2, (3, 2, 1))
    assert s2.lo == (0.0, 0.5, 0.0) and s2.hi == (1.5, 1.5, 0.25)
    assert np.array_equal(s2.support_points(), s.support_points())
    bad = tmp_path / "bad.mfr"
    bad.write_bytes(b"not a restart file")
    h = C.c_void_p()
    assert L.mf_restart_load_scratch(str(bad).encode(), C.byref(h)) == mf.MF_ERR_ARG


def test_cpp_header_compiles(tmp_path):
    """include/meltpooldg_amd.hpp (the C++ layer over the C ABI) parses and type-checks:
    its classes are inline, so including it checks them; the vector template is instantiated
    explicitly."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "use_hpp.cpp"
    src.write_text('#include "meltpooldg_amd.hpp"\n'
                   "template class mf::DeviceVector<double>;\n"
                   "template class mf::DeviceVector<float>;\n")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    r = subprocess.run([hipcc, "-fsyntax-only", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
                        "-I/opt/rocm/include", "-I" + inc, str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]


class _FakeDeviceTensor:
    """Stands in for a device tensor in the argument checks (nothing is dereferenced)."""

    def __init__(self, n, dtype="float64", index=0, contiguous=True):
        import torch
        self._n, self.dtype = n, getattr(torch, dtype)
        self.device = type("D", (), {"index": index})()
        self._contiguous = contiguous
        self.is_cuda = True

    def numel(self):
        return self._n

    def is_contiguous(self):
        return self._contiguous

    def data_ptr(self):
        return 0x1000


def test_vector_arguments_are_checked_before_any_call():
    """Length, dtype, layout and device of every vector are checked on the host: a float32 or
    short tensor handed to an FP64 entry point raises ValueError instead of reaching a kernel."""
    import torch
    ok = _FakeDeviceTensor(10)
    assert mf._ptr(ok, 10, "float64", 0).value == 0x1000
    with pytest.raises(ValueError, match="dtype"):
        mf._ptr(_FakeDeviceTensor(10, "float32"), 10)
    with pytest.raises(ValueError, match="entries"):
        mf._ptr(_FakeDeviceTensor(9), 10)
    with pytest.raises(ValueError, match="entries"):
        mf._ptr(_FakeDeviceTensor(30), 10)
    with pytest.raises(ValueError, match="device 1"):
        mf._ptr(_FakeDeviceTensor(10, index=1), 10, "float64", 0)
    with pytest.raises(ValueError, match="contiguous"):
        mf._ptr(_FakeDeviceTensor(10, contiguous=False), 10)
    with pytest.raises(ValueError, match="device tensors"):
        mf._ptr(torch.zeros(10, dtype=torch.float64), 10)
    assert mf._ptr(_FakeDeviceTensor(4, "float32"), 4, "float32").value == 0x1000
    # through the operator wrappers: the check fires before the C call
    s = mf.ScratchData(1, 2)
    op = mf.HeatDiffuseOperator(s, dt=1.0)
    assert op.n_rows == s.n_dofs == 27
    with pytest.raises(ValueError, match="entries"):
        op.vmult(_FakeDeviceTensor(27), _FakeDeviceTensor(26))
    with pytest.raises(ValueError, match="dtype"):
        op.vmult_f32(_FakeDeviceTensor(27), _FakeDeviceTensor(27))
    with pytest.raises(ValueError, match="entries"):
        op.set_fields(velocity=_FakeDeviceTensor(27))
    with pytest.raises(ValueError, match="entries"):
        op.set_fields(k_cell=_FakeDeviceTensor(27))
    with pytest.raises(ValueError, match="entries"):
        mf.LinearSolver.solve(op, _FakeDeviceTensor(27), _FakeDeviceTensor(28))


def test_cpp_driver_is_built_against_the_library():
    """The compiled C++ caller (tests/cpp/hpp_driver, built by build()) resolves
    libmeltpooldg_amd.so from the tree."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    drv = os.path.join(root, "tests", "cpp", "hpp_driver")
    assert os.path.exists(drv), "run __graft_entry__.build() first"
    r = subprocess.run(["ldd", drv], capture_output=True, text=True, timeout=60)
    line = [ln for ln in r.stdout.splitlines() if "libmeltpooldg_amd.so" in ln]
    assert line and "not found" not in line[0]
    assert os.path.realpath(line[0].split("=>")[1].split("(")[0].strip()) == \
        os.path.realpath(os.path.join(root, "meltpooldg_amd", "libmeltpooldg_amd.so"))
