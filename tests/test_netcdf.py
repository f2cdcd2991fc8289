"""NetCDF trajectory input (host only, no GPU): LAMMPS-style files in the
classic formats are written here -- CDF-1 and CDF-2 by scipy.io.netcdf_file,
CDF-5 by the small writer below, which follows the published format -- and
read back through trj_nc_* (raw frames, then the read_nc_cfg semantics:
species by first appearance among the selected types, origin shift, zero cell
length -> non-periodic axis of length span + 8)."""
import os
import struct

import numpy as np
import pytest

from trj_analysis_amd import trjhip as T

scipy_io = pytest.importorskip("scipy.io")


def _frames(nframes=3, natoms=50, seed=5):
    rng = np.random.default_rng(seed)
    L = np.array([10.0, 12.0, 0.0], dtype=np.float32)  # z non-periodic (LAMMPS writes 0)
    org = np.array([-5.0, 1.0, 0.0], dtype=np.float32)
    types = rng.integers(1, 4, size=natoms).astype(np.int32)
    types[:3] = [2, 3, 1]  # first appearance: 2, 3, 1
    coords = np.empty((nframes, natoms, 3), dtype=np.float32)
    for f in range(nframes):
        coords[f, :, 0] = org[0] + rng.random(natoms) * L[0]
        coords[f, :, 1] = org[1] + rng.random(natoms) * L[1]
        coords[f, :, 2] = -7.0 + rng.random(natoms) * 20.0
    return coords, types, org, L


def _write_scipy(path, version):
    coords, types, org, L = _frames()
    nf, na, _ = coords.shape
    with scipy_io.netcdf_file(path, "w", version=version) as nc:
        nc.createDimension("frame", None)
        nc.createDimension("atom", na)
        nc.createDimension("spatial", 3)
        nc.createDimension("cell_spatial", 3)
        t = nc.createVariable("time", "f", ("frame",))
        t.units = "lj"
        co = nc.createVariable("cell_origin", "f", ("frame", "cell_spatial"))
        cl = nc.createVariable("cell_lengths", "f", ("frame", "cell_spatial"))
        ty = nc.createVariable("type", "i", ("frame", "atom"))
        xyz = nc.createVariable("coordinates", "f", ("frame", "atom", "spatial"))
        for f in range(nf):
            t[f] = float(f)
            co[f] = org
            cl[f] = L
            ty[f] = types
            xyz[f] = coords[f]
    return coords, types, org, L


def _name5(s):
    b = s.encode()
    return struct.pack(">q", len(b)) + b + b"\0" * ((4 - len(b) % 4) % 4)


def _write_cdf5(path):
    """CDF-5 (64-bit data): dims, no attributes, four record variables."""
    coords, types, org, L = _frames()
    nf, na, _ = coords.shape
    dims = [("frame", 0), ("atom", na), ("spatial", 3), ("cell_spatial", 3)]
    # (name, dimids, nc_type, per-record bytes, data per frame)
    vars_ = [("cell_origin", [0, 3], 5, 12, lambda f: org.astype(">f4").tobytes()),
             ("cell_lengths", [0, 3], 5, 12, lambda f: L.astype(">f4").tobytes()),
             ("type", [0, 1], 4, 4 * na, lambda f: types.astype(">i4").tobytes()),
             ("coordinates", [0, 1, 2], 5, 12 * na, lambda f: coords[f].astype(">f4").tobytes())]

    def header(begins):
        h = b"CDF\x05" + struct.pack(">q", nf)
        h += struct.pack(">i", 0x0A) + struct.pack(">q", len(dims))
        for nm, ln in dims:
            h += _name5(nm) + struct.pack(">q", ln)
        h += struct.pack(">i", 0) + struct.pack(">q", 0)  # no global attributes
        h += struct.pack(">i", 0x0B) + struct.pack(">q", len(vars_))
        for (nm, dd, ty, vs, _), b in zip(vars_, begins):
            h += _name5(nm) + struct.pack(">q", len(dd)) + b"".join(struct.pack(">q", d) for d in dd)
            h += struct.pack(">i", 0) + struct.pack(">q", 0)  # no attributes
            h += struct.pack(">i", ty) + struct.pack(">q", (vs + 3) // 4 * 4) + struct.pack(">q", b)
        return h

    hlen = len(header([0] * len(vars_)))
    begins, off = [], hlen
    for v in vars_:
        begins.append(off)
        off += (v[3] + 3) // 4 * 4
    body = b""
    for f in range(nf):
        for v in vars_:
            d = v[4](f)
            body += d + b"\0" * ((4 - len(d) % 4) % 4)
    with open(path, "wb") as fh:
        fh.write(header(begins) + body)
    return coords, types, org, L


@pytest.mark.parametrize("fmt", ["cdf1", "cdf2", "cdf5"])
def test_raw_frames(tmp_path, fmt):
    path = str(tmp_path / f"t_{fmt}.nc")
    if fmt == "cdf5":
        coords, types, org, L = _write_cdf5(path)
    else:
        coords, types, org, L = _write_scipy(path, 1 if fmt == "cdf1" else 2)
    with T.NcTrajectory(path) as tr:
        assert tr.version == {"cdf1": 1, "cdf2": 2, "cdf5": 5}[fmt]
        assert (tr.nframes, tr.natoms) == coords.shape[:2]
        assert tr.coord_dtype == T.TRJ_F32 and tr.has_type and tr.has_cell_lengths and tr.has_cell_origin
        for f in range(tr.nframes):
            pos, ty, o, c = tr.read(f)
            np.testing.assert_array_equal(pos, coords[f].astype(np.float64))
            np.testing.assert_array_equal(ty, types)
            np.testing.assert_array_equal(o, org.astype(np.float64))
            np.testing.assert_array_equal(c, L.astype(np.float64))
        with pytest.raises(T.TrjError):
            tr.read(tr.nframes)


def test_frame_ref_semantics(tmp_path):
    path = str(tmp_path / "t.nc")
    coords, types, org, L = _write_scipy(path, 2)
    with T.NcTrajectory(path) as tr:
        # select LAMMPS types 1 and 3: species by first appearance -> 3 is 1, 1 is 2
        pos, it, box, per = tr.frame_ref(1, [1, 3])
        keep = (types == 1) | (types == 3)
        want_it = np.where(types[keep] == 3, 1, 2)
        np.testing.assert_array_equal(it, want_it)
        np.testing.assert_array_equal(per, [1, 1, 0])
        c = coords[1][keep].astype(np.float64)
        np.testing.assert_array_equal(pos[0], c[:, 0] - np.float64(org[0]))
        np.testing.assert_array_equal(pos[1], c[:, 1] - np.float64(org[1]))
        span = coords[1][:, 2].astype(np.float64)
        assert box[0] == np.float64(L[0]) and box[1] == np.float64(L[1])
        assert box[2] == pytest.approx(span.max() - span.min() + 8.0)
        np.testing.assert_allclose(pos[2], c[:, 2] - span.min() + 4.0)
        assert pos[2].min() >= 4.0 and pos[2].max() <= box[2] - 4.0 + 1e-12
        # the map is fixed by the first frame read
        _, it0, _, _ = tr.frame_ref(0, [1, 3])
        np.testing.assert_array_equal(it0, want_it)
        with pytest.raises(T.TrjError):
            tr.frame_ref(0, [1, 2])


def test_missing_species_and_bad_files(tmp_path):
    path = str(tmp_path / "t.nc")
    _write_scipy(path, 1)
    with T.NcTrajectory(path) as tr:
        with pytest.raises(T.TrjError):
            tr.frame_ref(0, [1, 7])  # type 7 never appears
    h5 = str(tmp_path / "h5.nc")
    with open(h5, "wb") as fh:
        fh.write(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(T.TrjError, match="HDF5"):
        T.NcTrajectory(h5)
    junk = str(tmp_path / "junk.nc")
    with open(junk, "wb") as fh:
        fh.write(b"not netcdf")
    with pytest.raises(T.TrjError):
        T.NcTrajectory(junk)
    with pytest.raises(T.TrjError):
        T.NcTrajectory(str(tmp_path / "absent.nc"))
    assert os.path.exists(path)


def test_time_and_periodic_flags(tmp_path):
    path = str(tmp_path / "t.nc")
    coords, types, org, L = _write_scipy(path, 2)
    with T.NcTrajectory(path) as tr:
        assert tr.has_time
        assert [tr.time(f) for f in range(tr.nframes)] == [0.0, 1.0, 2.0]
        with pytest.raises(T.TrjError):
            tr.time(tr.nframes)
        sel = [1, 2, 3]
        # x given as non-periodic by the caller: no origin shift, length kept,
        # coordinates centred in [0, L)
        pos, it, box, per = tr.frame_ref(2, sel, periodic=(False, True, True))
        np.testing.assert_array_equal(per, [0, 1, 0])
        assert box[0] == np.float64(L[0])
        x = coords[2][:, 0].astype(np.float64)
        lo, hi = pos[0].min(), pos[0].max()
        assert lo == pytest.approx(0.5 * (L[0] - (x.max() - x.min())))
        assert hi == pytest.approx(L[0] - lo)
        np.testing.assert_allclose(np.diff(pos[0]), np.diff(x), atol=1e-12)
        np.testing.assert_array_equal(pos[1], coords[2][:, 1].astype(np.float64) - np.float64(org[1]))


@pytest.mark.parametrize("qtype", ["f", "d"])
def test_charges_follow_selected_atoms(tmp_path, qtype):
    """q(frame, atom) comes back with the selected atoms, in their order
    (netcdf.f90:404-408, 657-658); a file without q reports has_charge 0 and
    refuses a charge read."""
    path = str(tmp_path / "q.nc")
    coords, types, org, L = _frames(nframes=2, natoms=40, seed=9)
    rng = np.random.default_rng(2)
    q = rng.uniform(-1.0, 1.0, size=(2, 40))
    with scipy_io.netcdf_file(path, "w", version=2) as nc:
        nc.createDimension("frame", None)
        nc.createDimension("atom", 40)
        nc.createDimension("spatial", 3)
        nc.createDimension("cell_spatial", 3)
        co = nc.createVariable("cell_origin", "f", ("frame", "cell_spatial"))
        cl = nc.createVariable("cell_lengths", "f", ("frame", "cell_spatial"))
        ty = nc.createVariable("type", "i", ("frame", "atom"))
        xyz = nc.createVariable("coordinates", "f", ("frame", "atom", "spatial"))
        qv = nc.createVariable("q", qtype, ("frame", "atom"))
        for f in range(2):
            co[f] = org
            cl[f] = L
            ty[f] = types
            xyz[f] = coords[f]
            qv[f] = q[f]
    stored = q.astype(np.float32 if qtype == "f" else np.float64).astype(np.float64)
    with T.NcTrajectory(path) as tr:
        assert tr.has_charge
        for f in range(2):
            pos, it, box, per, qq = tr.frame_ref(f, [3, 1], charges=True)
            sel = np.isin(types, [3, 1])
            np.testing.assert_array_equal(qq, stored[f][sel])
            pos2, it2, _, _ = tr.frame_ref(f, [3, 1])
            np.testing.assert_array_equal(pos, pos2)
            np.testing.assert_array_equal(it, it2)
    p2 = str(tmp_path / "noq.nc")
    _write_scipy(p2, 1)
    with T.NcTrajectory(p2) as tr:
        assert not tr.has_charge
        with pytest.raises(T.TrjError):
            tr.frame_ref(0, [1, 2, 3], charges=True)


@pytest.mark.parametrize("bad", [np.nan, 1e12, -3e9])
def test_corrupt_type_values_refused(tmp_path, bad):
    """A type variable stored as float with a NaN or a huge value (a corrupt
    file) is refused with TRJ_EINVAL before it can size the species map."""
    path = str(tmp_path / "bad.nc")
    coords, types, org, L = _frames(nframes=1, natoms=20)
    with scipy_io.netcdf_file(path, "w", version=1) as nc:
        nc.createDimension("frame", None)
        nc.createDimension("atom", 20)
        nc.createDimension("spatial", 3)
        nc.createDimension("cell_spatial", 3)
        cl = nc.createVariable("cell_lengths", "f", ("frame", "cell_spatial"))
        ty = nc.createVariable("type", "d", ("frame", "atom"))
        xyz = nc.createVariable("coordinates", "f", ("frame", "atom", "spatial"))
        t = types.astype(np.float64)
        t[7] = bad
        cl[0] = np.array([10.0, 12.0, 9.0], dtype=np.float32)
        ty[0] = t
        xyz[0] = coords[0]
    with T.NcTrajectory(path) as tr:
        with pytest.raises(T.TrjError) as e:
            tr.frame_ref(0, [1, 2])
        assert e.value.code == T.TRJ_EINVAL
        assert "type value of atom 7" in str(e.value)
        with pytest.raises(T.TrjError) as e:
            tr.read(0)
        assert e.value.code == T.TRJ_EINVAL


def test_selected_type_out_of_range_refused(tmp_path):
    path = str(tmp_path / "t.nc")
    _write_scipy(path, 1)
    with T.NcTrajectory(path) as tr:
        with pytest.raises(T.TrjError) as e:
            tr.frame_ref(0, [1, 2_000_000_000])
        assert e.value.code == T.TRJ_EINVAL
