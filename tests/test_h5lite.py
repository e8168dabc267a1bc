"""The HDF5 checkpoint without h5py (cupymag_amd/utils/h5lite.py): round
trips, the reference's layout (dataset ``data``, root attributes ``shape``
and ``dtype``), and interoperability with the HDF5 C library when the image
has one (it reads our files; we read its chunked, deflated and shuffled
files)."""
import ctypes as C
import glob

import numpy as np
import pytest

from cupymag_amd.utils import h5lite
from cupymag_amd.utils.magnetization_io import read_array, write_array

_LIBS = sorted(glob.glob("/opt/conda/lib/libhdf5.so.*") + glob.glob("/usr/lib/x86_64-linux-gnu/libhdf5*.so.*"))
_LIBS = [p for p in _LIBS if "_hl" not in p and "fortran" not in p and "cpp" not in p]


def _libhdf5():
    if not _LIBS:
        pytest.skip("no HDF5 C library in this image")
    L = C.CDLL(_LIBS[0])
    hid = C.c_int64
    L.H5open()
    for f in ("H5Fopen", "H5Fcreate", "H5Dopen2", "H5Dcreate2", "H5Dget_space", "H5Aopen", "H5Aget_type",
              "H5Screate_simple", "H5Pcreate"):
        getattr(L, f).restype = hid
    L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid, hid]
    L.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    L.H5Dcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid, hid]
    L.H5Dget_space.argtypes = [hid]
    L.H5Sget_simple_extent_dims.argtypes = [hid, C.c_void_p, C.c_void_p]
    L.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    L.H5Aopen.argtypes = [hid, C.c_char_p, hid]
    L.H5Aread.argtypes = [hid, hid, C.c_void_p]
    L.H5Aget_type.argtypes = [hid]
    L.H5Tget_size.argtypes = [hid]
    L.H5Tget_size.restype = C.c_size_t
    L.H5Screate_simple.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    L.H5Pcreate.argtypes = [hid]
    L.H5Pset_chunk.argtypes = [hid, C.c_int, C.c_void_p]
    L.H5Pset_deflate.argtypes = [hid, C.c_uint]
    L.H5Pset_shuffle.argtypes = [hid]
    for f in ("H5Dclose", "H5Sclose", "H5Fclose", "H5Aclose", "H5Tclose", "H5Pclose"):
        getattr(L, f).argtypes = [hid]
    L.NATIVE_DOUBLE = hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    L.NATIVE_LLONG = hid.in_dll(L, "H5T_NATIVE_LLONG_g").value
    L.DCREATE = hid.in_dll(L, "H5P_CLS_DATASET_CREATE_ID_g").value
    return L


@pytest.mark.parametrize("shape", [(5, 3), (300001, 3), (7,), (64, 6)])
def test_round_trip(tmp_path, shape):
    a = np.random.default_rng(0).standard_normal(shape)
    fn = tmp_path / "m.h5"
    h5lite.write(fn, a, attrs={"shape": a.shape, "dtype": str(a.dtype)})
    assert fn.read_bytes()[:8] == b"\x89HDF\r\n\x1a\n"
    np.testing.assert_array_equal(h5lite.read(fn), a)
    assert h5lite.read_attrs(fn) == {"shape": shape, "dtype": "float64"}
    with pytest.raises(KeyError):
        h5lite.read(fn, "other")


def test_checkpoint_api(tmp_path):
    m = np.random.default_rng(1).standard_normal((100, 3))
    write_array(m, tmp_path / "last_m")          # suffix added as in the reference
    assert (tmp_path / "last_m.h5").exists()
    np.testing.assert_array_equal(read_array(tmp_path / "last_m.h5", 100), m)
    assert read_array(tmp_path / "last_m.h5", 99) is None   # shape mismatch -> uniform start
    assert read_array(None, 100) is None


def test_hdf5_library_reads_our_file(tmp_path):
    L = _libhdf5()
    for shape in [(5, 3), (300001, 3)]:
        a = np.random.default_rng(2).standard_normal(shape)
        fn = str(tmp_path / "ours.h5")
        h5lite.write(fn, a, attrs={"shape": a.shape, "dtype": "float64"})
        f = L.H5Fopen(fn.encode(), 0, 0)
        assert f >= 0
        ds = L.H5Dopen2(f, b"data", 0)
        sp = L.H5Dget_space(ds)
        dims = (C.c_uint64 * 2)()
        assert L.H5Sget_simple_extent_dims(sp, dims, None) == 2 and tuple(dims) == shape
        out = np.empty(shape)
        assert L.H5Dread(ds, L.NATIVE_DOUBLE, 0, 0, 0, out.ctypes.data) == 0
        np.testing.assert_array_equal(out, a)
        at = L.H5Aopen(f, b"shape", 0)
        sh = (C.c_int64 * 2)()
        assert L.H5Aread(at, L.NATIVE_LLONG, sh) == 0 and tuple(sh) == shape
        at2 = L.H5Aopen(f, b"dtype", 0)
        ty = L.H5Aget_type(at2)
        buf = C.create_string_buffer(L.H5Tget_size(ty) + 1)
        assert L.H5Aread(at2, ty, buf) == 0 and buf.value == b"float64"
        for h, fn_ in ((ty, "H5Tclose"), (at2, "H5Aclose"), (at, "H5Aclose"), (sp, "H5Sclose"), (ds, "H5Dclose"),
                       (f, "H5Fclose")):
            assert getattr(L, fn_)(h) == 0


@pytest.mark.parametrize("shape,chunk,shuffle", [((1000, 3), (128, 3), False), ((777, 3), (100, 3), True),
                                                 ((5000, 3), (5000, 3), False)])
def test_we_read_hdf5_library_files(tmp_path, shape, chunk, shuffle):
    L = _libhdf5()
    a = np.random.default_rng(3).standard_normal(shape)
    fn = str(tmp_path / "theirs.h5")
    f = L.H5Fcreate(fn.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    sp = L.H5Screate_simple(2, (C.c_uint64 * 2)(*shape), None)
    pl = L.H5Pcreate(L.DCREATE)
    L.H5Pset_chunk(pl, 2, (C.c_uint64 * 2)(*chunk))
    if shuffle:
        L.H5Pset_shuffle(pl)
    L.H5Pset_deflate(pl, 6)
    ds = L.H5Dcreate2(f, b"data", L.NATIVE_DOUBLE, sp, 0, pl, 0)
    assert L.H5Dwrite(ds, L.NATIVE_DOUBLE, 0, 0, 0, a.ctypes.data) == 0
    for h, fn_ in ((ds, "H5Dclose"), (pl, "H5Pclose"), (sp, "H5Sclose"), (f, "H5Fclose")):
        getattr(L, fn_)(h)
    np.testing.assert_array_equal(h5lite.read(fn), a)
