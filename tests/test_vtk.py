"""Legacy VTK frames and the VTK restart IC (CPU only).

The writer follows IOManager::WriteVtk (src/io_manager.cpp:218-315) and the
reader + VtkIC follow src/ics/vtk_ic.cpp:42-239.  The reference writes through
the VTK library, which is not in this image, so no reference-written file
exists to compare against: the checks below parse our files with an
independent legacy-format reader written here, restate the frame layout and
the bilinear restart mapping in numpy, and use closed forms (a bilinear field
is reproduced exactly by bilinear interpolation).
"""
import os

import numpy as np
import pytest

import cases

RNG = np.random.default_rng(7)


def _parse_legacy(path):
    """Minimal independent reader for our legacy structured-grid files."""
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    assert line().startswith("# vtk DataFile Version")
    line()
    fmt = line().strip()
    assert line().strip() == "DATASET STRUCTURED_GRID"
    ni, nj, nk = map(int, line().split()[1:])
    assert nk == 1
    n = ni * nj

    def values(count):
        nonlocal pos
        if fmt == "BINARY":
            v = np.frombuffer(data, dtype=">f8", count=count, offset=pos).astype(np.float64)
            pos += 8 * count
            assert data[pos:pos + 1] == b"\n"
            pos += 1
            return v
        out = []
        while len(out) < count:
            out.extend(float(t) for t in line().split())
        return np.array(out)

    kw, np_, typ = line().split()
    assert kw == "POINTS" and int(np_) == n and typ == "double"
    points = values(3 * n).reshape(nj, ni, 3)
    kw, cnt = line().split()
    assert kw == "POINT_DATA" and int(cnt) == n
    kw, _, narr = line().split()
    assert kw == "FIELD"
    arrays = {}
    for _ in range(int(narr)):
        name, nc, nt, typ = line().split()
        assert int(nt) == n and typ == "double"
        arrays[name] = values(n * int(nc)).reshape(nj, ni, int(nc))
    return fmt, ni, nj, points, arrays


def _straight_r(L, M):
    """r(l, m) of the straight channel r in [0.2, 0.8] (src/grid.cpp:31-37 op order)."""
    dy = 1.0 / M
    m = np.arange(M + 3)
    frac = (m - 1) * dy
    return (1.0 - frac) * 0.2 + frac * 0.8


def _random_phys(rows, cols):
    ph = [RNG.uniform(0.5, 2.0, (rows, cols)) for _ in range(10)]
    for p in ph:
        p[0, :] = p[-1, :] = 0.0
        p[:, 0] = p[:, -1] = 0.0
    return ph


@pytest.mark.parametrize("binary", [True, False])
def test_frame_layout(mpd, tmp_path, binary):
    L, M = 12, 6
    text = cases.channel_config(L, M)
    ph = _random_phys(L + 2, M + 3)
    path = tmp_path / "f.vtk"
    mpd.write_vtk_fields(text, ph, path, binary=binary)
    fmt, ni, nj, pts, arr = _parse_legacy(path)
    assert fmt == ("BINARY" if binary else "ASCII")
    assert (ni, nj) == (L + 1, M + 1)
    assert list(arr) == ["Rho", "Energy", "MPI_Rank", "Hphi_times_r", "Velocity", "MagneticField"]
    i = np.arange(ni)
    li = np.minimum(i, L - 1)  # the last point column repeats row L-1
    r = _straight_r(L, M)[1:M + 2]
    assert np.array_equal(pts[:, :, 0], np.broadcast_to(i * (1.0 / L), (nj, ni)))
    assert np.array_equal(pts[:, :, 1], np.broadcast_to(r[:, None], (nj, ni)))
    assert not pts[:, :, 2].any()

    def g(k):  # global [l_g][m_g] -> frame [j][i]
        return ph[k][li + 1][:, 1:M + 2].T

    assert np.array_equal(arr["Rho"][:, :, 0], g(0))
    assert np.array_equal(arr["Energy"][:, :, 0], g(4))
    assert not arr["MPI_Rank"].any()
    assert np.array_equal(arr["Hphi_times_r"][:, :, 0], g(9) * r[:, None])
    assert np.array_equal(arr["Velocity"], np.stack([g(1), g(2), g(3)], axis=-1))
    assert np.array_equal(arr["MagneticField"], np.stack([g(7), g(8), g(9)], axis=-1))


def _restart_config(L, M, path, exact=None):
    base = cases.channel_config(L, M)
    i = base.index("initial_conditions:")
    j = base.index("boundary_conditions:")
    extra = "" if exact is None else f"  vtk_exact_z: {str(exact).lower()}\n"
    return base[:i] + f"initial_conditions:\n  vtk_file: \"{path}\"\n{extra}" + base[j:]


def _bilinear(data, i_frac, j_frac):
    """VtkIC::Impl::Interp restated (src/ics/vtk_ic.cpp:137-160), same op order."""
    nj, ni = data.shape
    i_frac = min(max(i_frac, 0.0), float(ni - 1))
    j_frac = min(max(j_frac, 0.0), float(nj - 1))
    i0 = min(int(i_frac), ni - 2)
    j0 = min(int(j_frac), nj - 2)
    ti, tj = i_frac - i0, j_frac - j0
    v00, v10 = data[j0, i0], data[j0, i0 + 1]
    v01, v11 = data[j0 + 1, i0], data[j0 + 1, i0 + 1]
    return (1.0 - ti) * (1.0 - tj) * v00 + ti * (1.0 - tj) * v10 + (1.0 - ti) * tj * v01 + ti * tj * v11


def _expected_ic(frame_arrays, L, M, ni, nj, exact, gamma):
    names = {0: ("Rho", 0), 1: ("Velocity", 0), 2: ("Velocity", 1), 3: ("Velocity", 2), 4: ("Energy", 0),
             7: ("MagneticField", 0), 8: ("MagneticField", 1), 9: ("MagneticField", 2)}
    i_scale = ((ni - 2) if exact else (ni - 1)) / (L - 1.0)
    j_scale = (nj - 1) / float(M)
    out = [np.zeros((L + 2, M + 3)) for _ in range(10)]
    for l in range(1, L + 1):
        for m in range(1, M + 2):
            i_f, j_f = (l - 1) * i_scale, (m - 1) * j_scale
            for k, (nm, c) in names.items():
                out[k][l, m] = _bilinear(frame_arrays[nm][:, :, c], i_f, j_f)
            out[5][l, m] = (gamma - 1.0) * out[0][l, m] * out[4][l, m]
            out[6][l, m] = out[5][l, m] + 0.5 * (out[7][l, m] * out[7][l, m] + out[8][l, m] * out[8][l, m] +
                                                 out[9][l, m] * out[9][l, m])
    return out


@pytest.mark.parametrize("L0,M0,L1,M1,exact", [
    (12, 6, 12, 6, False),   # same size, the reference's axial mapping
    (12, 6, 12, 6, True),    # same size, exact axial mapping
    (10, 5, 23, 11, False),  # remap to a finer grid
    (20, 8, 9, 4, True),     # remap to a coarser grid
])
def test_restart_matches_restated_mapping(mpd, tmp_path, L0, M0, L1, M1, exact):
    ph = _random_phys(L0 + 2, M0 + 3)
    path = tmp_path / "ic.vtk"
    mpd.write_vtk_fields(cases.channel_config(L0, M0), ph, path)
    _, ni, nj, _, arr = _parse_legacy(path)
    _, got = mpd.initial_state(_restart_config(L1, M1, path, exact))
    want = _expected_ic(arr, L1, M1, ni, nj, exact, float(cases.GAMMA_53))
    for k in range(10):
        assert np.array_equal(got[k], want[k]), k


def test_same_size_exact_restart_is_a_copy(mpd, tmp_path):
    L, M = 12, 6
    ph = _random_phys(L + 2, M + 3)
    path = tmp_path / "ic.vtk"
    mpd.write_vtk_fields(cases.channel_config(L, M), ph, path)
    _, got = mpd.initial_state(_restart_config(L, M, path, exact=True))
    for k in (0, 1, 2, 3, 4, 7, 8, 9):
        assert np.array_equal(got[k][1:L + 1, 1:M + 2], ph[k][1:L + 1, 1:M + 2]), k


def test_reference_mapping_on_a_bilinear_field(mpd, tmp_path):
    """Closed form: a field bilinear in the point indices is reproduced at the
    fractional index the reference maps each cell to."""
    L0, M0, L1, M1 = 10, 5, 17, 9
    ni, nj = L0 + 1, M0 + 1
    phys = _random_phys(L0 + 2, M0 + 3)
    # Rho(i, j) = 1 + 0.25 i + 0.5 j + 0.125 i j on every point, i.e. on global
    # cells l_g = i (the repeated last point column is consistent because the
    # field is defined on l_g = min(i, L0 - 1) -- use only j-dependence there).
    for l in range(1, L0 + 1):
        for m in range(1, M0 + 2):
            i, j = l - 1, m - 1
            phys[0][l, m] = 1.0 + 0.25 * i + 0.5 * j + 0.125 * i * j
    path = tmp_path / "lin.vtk"
    mpd.write_vtk_fields(cases.channel_config(L0, M0), phys, path)
    _, got = mpd.initial_state(_restart_config(L1, M1, path))
    i_scale, j_scale = (ni - 1) / (L1 - 1.0), (nj - 1) / M1
    for l in range(1, L1 + 1):
        for m in range(1, M1 + 2):
            i_f, j_f = (l - 1) * i_scale, (m - 1) * j_scale
            i_f = min(i_f, ni - 1.0)
            if i_f > L0 - 1:  # between the last row and its repeat: constant in i
                i_f = L0 - 1.0
            want = 1.0 + 0.25 * i_f + 0.5 * j_f + 0.125 * i_f * j_f
            assert abs(got[0][l, m] - want) <= 1e-13 * abs(want), (l, m)


def test_restart_errors(mpd, tmp_path):
    bad = tmp_path / "bad.vtk"
    bad.write_text("not a vtk file\n")
    with pytest.raises(mpd.MpdError, match="not a legacy VTK file"):
        mpd.initial_state(_restart_config(12, 6, bad))
    with pytest.raises(mpd.MpdError, match="cannot open"):
        mpd.initial_state(_restart_config(12, 6, tmp_path / "missing.vtk"))
    # A frame without MagneticField.
    path = tmp_path / "f.vtk"
    ph = _random_phys(14, 9)
    mpd.write_vtk_fields(cases.channel_config(12, 6), ph, path, binary=False)
    text = open(path).read()
    cut = text.index("MagneticField")
    text = text[:cut].replace("FIELD FieldData 6", "FIELD FieldData 5")
    path.write_text(text)
    with pytest.raises(mpd.MpdError, match="'MagneticField' not found"):
        mpd.initial_state(_restart_config(12, 6, path))
