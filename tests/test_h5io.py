"""The reference's HDF5 formats (libgtt_h5), checked against an independent
path: the tests read and write the files themselves through the HDF5
high-level C API (ctypes), using the dataset names, ranks, types and
ordering of the reference's loader (xray/CBCT_acquisition.h:177-333,
xray/CBCT_binning.h:106-157), the reference's binning writer
(python/l1shroud.py:139-150) and image writer (proton/hdf5_utils.h:17-45)."""
import ctypes
import os

import numpy as np
import pytest

from gtt import h5io

pytestmark = pytest.mark.skipif(not h5io.available(), reason="HDF5 libraries not present")

hid_t = ctypes.c_int64
hsize_t = ctypes.c_ulonglong


class H5:
    """Minimal independent HDF5 access (HDF5 1.10 C API)."""

    def __init__(self):
        h5io.load()
        self.L = ctypes.CDLL(os.path.join(h5io.H5LIB, "libhdf5.so"), mode=ctypes.RTLD_GLOBAL)
        self.HL = ctypes.CDLL(os.path.join(h5io.H5LIB, "libhdf5_hl.so"), mode=ctypes.RTLD_GLOBAL)
        self.L.H5open()
        self.L.H5Fcreate.restype = hid_t
        self.L.H5Fopen.restype = hid_t
        self.L.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]
        self.L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
        self.L.H5Fclose.argtypes = [hid_t]
        self.UINT = hid_t.in_dll(self.L, "H5T_NATIVE_UINT_g").value
        self.FLOAT = hid_t.in_dll(self.L, "H5T_NATIVE_FLOAT_g").value
        H = self.HL
        H.H5LTmake_dataset.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(hsize_t),
                                       hid_t, ctypes.c_void_p]
        H.H5LTread_dataset.argtypes = [hid_t, ctypes.c_char_p, hid_t, ctypes.c_void_p]
        H.H5LTget_dataset_ndims.argtypes = [hid_t, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        H.H5LTget_dataset_info.argtypes = [hid_t, ctypes.c_char_p, ctypes.POINTER(hsize_t),
                                           ctypes.c_void_p, ctypes.c_void_p]
        H.H5LTget_attribute_float.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p]
        H.H5LTget_attribute_uint.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p]
        H.H5LTget_attribute_string.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]

    def create(self, path):
        return self.L.H5Fcreate(os.fsencode(path), 0x0002, 0, 0)

    def open(self, path):
        return self.L.H5Fopen(os.fsencode(path), 0x0000, 0)

    def close(self, f):
        self.L.H5Fclose(f)

    def write(self, f, name, arr, uint=False):
        a = np.ascontiguousarray(arr, np.uint32 if uint else np.float32)
        dims = (hsize_t * max(a.ndim, 1))(*(a.shape or (1,)))
        rc = self.HL.H5LTmake_dataset(f, name.encode(), max(a.ndim, 1), dims,
                                      self.UINT if uint else self.FLOAT, a.ctypes.data)
        assert rc >= 0, name

    def read(self, f, name, uint=False):
        rank = ctypes.c_int()
        assert self.HL.H5LTget_dataset_ndims(f, name.encode(), ctypes.byref(rank)) >= 0, name
        dims = (hsize_t * rank.value)()
        assert self.HL.H5LTget_dataset_info(f, name.encode(), dims, None, None) >= 0
        out = np.empty(tuple(dims), np.uint32 if uint else np.float32)
        assert self.HL.H5LTread_dataset(f, name.encode(), self.UINT if uint else self.FLOAT,
                                        out.ctypes.data) >= 0
        return out

    def attr_float(self, f, obj, name, n):
        out = np.empty(n, np.float32)
        assert self.HL.H5LTget_attribute_float(f, obj.encode(), name.encode(), out.ctypes.data) >= 0
        return out


@pytest.fixture(scope="module")
def h5():
    return H5()


def scan(seed=0, nu=12, nv=7, n=5):
    r = np.random.default_rng(seed)
    return (r.random((n, nv, nu)).astype(np.float32), np.linspace(0, 300, n).astype(np.float32),
            r.uniform(-2, 2, (n, 2)).astype(np.float32))


def test_acquisition_written_in_reference_layout(h5, tmp_path):
    proj, ang, off = scan()
    path = tmp_path / "acq.hdf5"
    h5io.save_acquisition(path, proj, ang, 1000.0, 1500.0, (199.68, 149.76), off)
    f = h5.open(path)
    try:
        assert h5.read(f, "/geometry_dataformat_version", uint=True).tolist() == [2]
        assert h5.read(f, "/projection_dataformat_version", uint=True).tolist() == [1]
        p = h5.read(f, "/projections")
        assert p.shape == (5, 7, 12) and np.array_equal(p, proj)   # [views][v][u]
        assert np.array_equal(h5.read(f, "/angles"), ang)
        assert np.array_equal(h5.read(f, "/offsetx"), off[:, 0])
        assert np.array_equal(h5.read(f, "/offsety"), off[:, 1])
        assert h5.read(f, "/SAD").tolist() == [1000.0] and h5.read(f, "/SDD").tolist() == [1500.0]
        assert np.allclose(h5.read(f, "/FOV"), [199.68, 149.76])
    finally:
        h5.close(f)
    a = h5io.load_acquisition(path)
    assert np.array_equal(a["projections"], proj) and np.array_equal(a["angles"], ang)
    assert np.array_equal(a["offsets"], off) and a["airscan"] is None


def test_acquisition_spacing_fallback_and_airscan(h5, tmp_path):
    proj, ang, off = scan(1, nu=10, nv=6, n=4)
    air = np.random.default_rng(3).random((2, 6, 10)).astype(np.float32)
    path = tmp_path / "old.hdf5"
    f = h5.create(path)
    h5.write(f, "/geometry_dataformat_version", [2], uint=True)
    h5.write(f, "/projection_dataformat_version", [1], uint=True)
    h5.write(f, "/angles", ang)
    h5.write(f, "/offsetx", off[:, 0])
    h5.write(f, "/offsety", off[:, 1])
    h5.write(f, "/projections", proj)
    h5.write(f, "/airscan", air)
    h5.write(f, "/SAD", [800.0])
    h5.write(f, "/SDD", [1200.0])
    h5.write(f, "/spacing", [0.5, 0.25])       # no /FOV: FOV = spacing x (nu, nv)
    h5.close(f)
    a = h5io.load_acquisition(path)
    assert np.allclose(a["fov"], [0.5 * 10, 0.25 * 6])
    assert a["sad"] == 800.0 and a["sdd"] == 1200.0
    assert np.array_equal(a["airscan"], air) and np.array_equal(a["projections"], proj)


def test_acquisition_errors(h5, tmp_path):
    with pytest.raises(RuntimeError, match="Projection file does not exist"):
        h5io.load_acquisition(tmp_path / "missing.hdf5")
    path = tmp_path / "v3.hdf5"
    f = h5.create(path)
    h5.write(f, "/geometry_dataformat_version", [3], uint=True)
    h5.close(f)
    with pytest.raises(RuntimeError, match="wrong geometry data version inside hdf5 file, found: 3, needed: 2"):
        h5io.load_acquisition(path)
    path = tmp_path / "nooffsety.hdf5"
    f = h5.create(path)
    h5.write(f, "/geometry_dataformat_version", [2], uint=True)
    h5.write(f, "/angles", [0.0, 1.0])
    h5.write(f, "/offsetx", [0.0, 0.0])
    h5.close(f)
    with pytest.raises(RuntimeError, match="Error getting /offsety dataset info from file."):
        h5io.load_acquisition(path)


def test_binning_python_writer_layout(h5, tmp_path):
    # written the way python/l1shroud.py:save_binning writes it
    binning = np.array([0, 1, 2, 0, 1, 2, 2, 5])    # view -> bin; 5 = excluded
    nbins = 3
    path = tmp_path / "bins.hdf5"
    f = h5.create(path)
    index = np.arange(len(binning), dtype=np.uint32)
    for i in range(nbins):
        h5.write(f, "/bin_%d" % (i + 1), index[binning == i], uint=True)
    h5.write(f, "/numBins", [nbins], uint=True)
    h5.write(f, "/binning_dataformat_version", [1], uint=True)
    h5.close(f)
    bins = h5io.load_binning(path)
    assert [b.tolist() for b in bins] == [[0, 3], [1, 4], [2, 5, 6]]
    out = tmp_path / "bins2.hdf5"
    h5io.save_binning(out, bins)
    f = h5.open(out)
    try:
        assert h5.read(f, "/numBins", uint=True).tolist() == [3]
        assert h5.read(f, "/binning_dataformat_version", uint=True).tolist() == [1]
        assert h5.read(f, "/bin_3", uint=True).tolist() == [2, 5, 6]
    finally:
        h5.close(f)


def test_image_writer_layout_and_attributes(h5, tmp_path):
    img = np.random.default_rng(9).random((4, 5, 6)).astype(np.float32)   # (nz, ny, nx)
    path = tmp_path / "image.hdf5"
    h5io.save_image(path, img, dims_mm=(12.0, 10.0, 20.0), origin=(1.0, 2.0, 3.0),
                    arguments="--size 6,5,4", iterations=7)
    f = h5.open(path)
    try:
        assert np.array_equal(h5.read(f, "/image"), img)               # dims reversed: [z][y][x]
        assert np.allclose(h5.attr_float(f, "/image", "dimensions", 3), [12.0, 10.0, 20.0])
        # element_size_um[D-1-i] = dimensions[i] / size(i): (x 2, y 2, z 5) stored reversed
        assert np.allclose(h5.attr_float(f, "/image", "element_size_um", 3), [5.0, 2.0, 2.0])
        assert np.allclose(h5.attr_float(f, "/image", "origin", 3), [1.0, 2.0, 3.0])
        it = np.zeros(1, np.uint32)
        assert h5.HL.H5LTget_attribute_uint(f, b"/image", b"iterations", it.ctypes.data) >= 0
        assert it[0] == 7
        s = ctypes.create_string_buffer(64)
        assert h5.HL.H5LTget_attribute_string(f, b"/image", b"input", s) >= 0
        assert s.value == b"--size 6,5,4"
    finally:
        h5.close(f)
