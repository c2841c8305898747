"""Formats around the device path: VTU output of device fields (GenericDataOut / DataOut
write_vtu, source/post_processing/generic_data_out.cpp) and restart files (Restart::
serialize_internal / deserialize_internal, source/utilities/restart.cpp:115-186).

The VTU file is parsed back here (XML header + appended raw blocks): one point per DoF at its
support point, the fields bit for bit, degree^3 hexahedra per cell that tile the domain (their
volumes sum to its volume and each is a positive box).  A restart round trip rebuilds the same
DoF numbering and Dirichlet mask (operator applications agree bit for bit) and returns the
vectors bit for bit; mismatched requests fail with an error."""
import re

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def read_vtu(path):
    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData encoding=\"raw\">")
    us = raw.index(b"_", start) + 1
    head = raw[:start].decode()
    pts = int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))
    cells = int(re.search(r'NumberOfCells="(\d+)"', head).group(1))
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="([^"]+)" NumberOfComponents="(\d)" '
                         r'format="appended" offset="(\d+)"/>', head):
        typ, name, nc, off = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
        pos = us + off
        nbytes = int(np.frombuffer(raw[pos:pos + 8], dtype=np.uint64)[0])
        dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[typ]
        a = np.frombuffer(raw[pos + 8:pos + 8 + nbytes], dtype=dt)
        arrays[name] = a.reshape(-1, nc) if nc > 1 else a
    return pts, cells, arrays


def _check_tiling(arrays, n_sub, volume):
    conn = arrays["connectivity"].reshape(n_sub, 8)
    X = arrays["Points"]
    assert np.array_equal(arrays["offsets"], 8 * np.arange(1, n_sub + 1))
    assert np.all(arrays["types"] == 12)
    lo, hi = X[conn[:, 0]], X[conn[:, 6]]
    assert np.all(hi > lo)
    # corners of an axis-aligned box in VTK order
    for v, (a, b, c) in enumerate([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                                   (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]):
        expect = np.stack([hi[:, 0] if a else lo[:, 0], hi[:, 1] if b else lo[:, 1],
                           hi[:, 2] if c else lo[:, 2]], axis=1)
        assert np.allclose(X[conn[:, v]], expect, rtol=0, atol=1e-14)
    assert abs(np.prod(hi - lo, axis=1).sum() - volume) < 1e-12 * volume


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_vtu_structured(tmp_path, degree):
    import meltpooldg_amd as mf
    s = mf.ScratchData(degree, (3, 2, 2), hi=(1.0, 0.8, 0.5), stream=torch.cuda.current_stream())
    n = s.n_dofs
    X = s.support_points()
    u = torch.tensor(np.sin(X[:, 0]) + X[:, 2], dtype=torch.float64, device="cuda")
    w = torch.tensor(X.T.ravel().copy(), dtype=torch.float64, device="cuda")
    path = tmp_path / "out.vtu"
    mf.write_vtu(s, path, {"temperature": u, "normal": w})
    pts, cells, arr = read_vtu(path)
    assert pts == n and cells == s.n_cells * degree ** 3
    assert np.array_equal(arr["Points"], X)
    assert np.array_equal(arr["temperature"], u.cpu().numpy())
    assert np.array_equal(arr["normal"], X)
    _check_tiling(arr, cells, 1.0 * 0.8 * 0.5)


def test_vtu_adaptive(tmp_path):
    import meltpooldg_amd as mf
    cells = mf.refine_band_cells(2, 2, (0.3, 0.4, 0.35), 0.25, 0.04)
    s = mf.AdaptiveScratchData(2, 2, cells, stream=torch.cuda.current_stream())
    X = s.support_points()
    u = torch.tensor(X[:, 1], dtype=torch.float64, device="cuda")
    path = tmp_path / "a.vtu"
    mf.write_vtu(s, path, {"y": u})
    pts, ncells, arr = read_vtu(path)
    assert pts == s.n_dofs and ncells == len(cells) * 8
    assert np.array_equal(arr["y"], X[:, 1])
    _check_tiling(arr, ncells, 1.0)


def _apply(mf, s, x):
    op = mf.HeatDiffuseOperator(s, rho_cp=1.3, k=0.7, dt=1e-2)
    y = torch.zeros_like(x)
    op.vmult(y, x)
    return y


def test_restart_structured(tmp_path):
    import meltpooldg_amd as mf
    st = torch.cuda.current_stream()
    s = mf.ScratchData(2, (4, 3, 2), lo=(0.1, 0.0, -0.2), hi=(1.0, 0.9, 0.4), stream=st)
    rng = np.random.default_rng(3)
    mask = (rng.uniform(size=s.n_dofs) < 0.2).astype(np.uint8)
    s.set_dirichlet_mask(mask)
    a = torch.tensor(rng.uniform(-1, 1, s.n_dofs), dtype=torch.float64, device="cuda")
    b = torch.tensor(rng.uniform(-1, 1, 3 * s.n_dofs), dtype=torch.float64, device="cuda")
    path = tmp_path / "state.mfr"
    mf.restart_save(s, path, [a, b])
    a2, b2 = torch.zeros_like(a), torch.zeros_like(b)
    s2 = mf.restart_load(path, [a2, b2], stream=st)
    assert (s2.n_dofs, s2.n_cells, s2.degree, s2.ncell) == (s.n_dofs, s.n_cells, 2, (4, 3, 2))
    assert np.array_equal(s2.support_points(), s.support_points())
    assert torch.equal(a2, a) and torch.equal(b2, b)
    assert torch.equal(_apply(mf, s2, a2), _apply(mf, s, a))  # same numbering and mask
    with pytest.raises(mf.MFError):
        mf.restart_load(path, [a2], stream=st)                  # wrong vector count
    with pytest.raises(mf.MFError):
        mf.restart_load(path, [a2, a2], stream=st)              # wrong length


def test_restart_adaptive(tmp_path):
    import meltpooldg_amd as mf
    st = torch.cuda.current_stream()
    cells = mf.refine_band_cells(2, 2, (0.3, 0.4, 0.35), 0.25, 0.04)
    s = mf.AdaptiveScratchData(2, 2, cells, stream=st)
    x = torch.tensor(np.random.default_rng(4).uniform(-1, 1, s.n_dofs), dtype=torch.float64,
                     device="cuda")
    path = tmp_path / "a.mfr"
    mf.restart_save(s, path, [x])
    x2 = torch.zeros_like(x)
    s2 = mf.restart_load(path, [x2], stream=st)
    assert isinstance(s2, mf.AdaptiveScratchData)
    assert s2.n_dofs == s.n_dofs and np.array_equal(s2.cells, s.cells)
    assert np.array_equal(s2.support_points(), s.support_points())
    assert torch.equal(x2, x)
    assert torch.equal(_apply(mf, s2, x2), _apply(mf, s, x))


def _load_vectors_into(mf, s, path, vectors):
    arr_v = (mf.C.c_void_p * len(vectors))(*[s._v(v, v.numel()) for v in vectors])
    arr_l = (mf.C.c_int64 * len(vectors))(*[v.numel() for v in vectors])
    return mf.lib().mf_restart_load_vectors(s._h, str(path).encode(), len(vectors), arr_v, arr_l)


def test_restart_rejects_other_meshes_and_short_files(tmp_path):
    """Vectors load only into the mesh they were saved on: a shifted box with the same counts,
    or a band mesh with the same cell and DoF counts but other cells, is refused; a truncated
    file is refused before anything is sized from its counts."""
    import meltpooldg_amd as mf
    st = torch.cuda.current_stream()
    s = mf.ScratchData(2, (4, 3, 2), lo=(0.1, 0.0, -0.2), hi=(1.0, 0.9, 0.4), stream=st)
    a = torch.tensor(np.random.default_rng(5).uniform(-1, 1, s.n_dofs), dtype=torch.float64,
                     device="cuda")
    path = tmp_path / "s.mfr"
    mf.restart_save(s, path, [a])
    same = mf.ScratchData(2, (4, 3, 2), lo=(0.1, 0.0, -0.2), hi=(1.0, 0.9, 0.4), stream=st)
    a2 = torch.zeros_like(a)
    assert _load_vectors_into(mf, same, path, [a2]) == mf.MF_OK and torch.equal(a2, a)
    shifted = mf.ScratchData(2, (4, 3, 2), lo=(0.2, 0.0, -0.2), hi=(1.1, 0.9, 0.4), stream=st)
    assert _load_vectors_into(mf, shifted, path, [a2]) == mf.MF_ERR_ARG
    # band meshes mirrored in x: equal counts, other cells
    c1 = mf.refine_band_cells(2, 2, (0.3, 0.4, 0.35), 0.25, 0.04)
    c2 = mf.refine_band_cells(2, 2, (0.7, 0.4, 0.35), 0.25, 0.04)
    b1 = mf.AdaptiveScratchData(2, 2, c1, stream=st)
    b2 = mf.AdaptiveScratchData(2, 2, c2, stream=st)
    assert (b1.n_cells, b1.n_dofs) == (b2.n_cells, b2.n_dofs)
    assert not np.array_equal(np.asarray(c1), np.asarray(c2))
    x = torch.tensor(np.random.default_rng(6).uniform(-1, 1, b1.n_dofs), dtype=torch.float64,
                     device="cuda")
    bpath = tmp_path / "b.mfr"
    mf.restart_save(b1, bpath, [x])
    x2 = torch.zeros_like(x)
    assert _load_vectors_into(mf, b2, bpath, [x2]) == mf.MF_ERR_ARG
    assert _load_vectors_into(mf, b1, bpath, [x2]) == mf.MF_OK and torch.equal(x2, x)
    # truncated files: inside the cell list, and inside the vector
    raw = bpath.read_bytes()
    for cut in (200, len(raw) - 16):
        short = tmp_path / f"short{cut}.mfr"
        short.write_bytes(raw[:cut])
        with pytest.raises(mf.MFError):
            mf.restart_load(short, [x2], stream=st)
