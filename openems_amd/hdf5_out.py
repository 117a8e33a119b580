"""Field dumps as openEMS HDF5 files (dump FileType 1), through the HDF5 C
library (libhdf5, loaded with ctypes; the image carries it under /opt/conda).

Layout of HDF5_File_Writer and ProcessFields (tools/hdf5_file_writer.cpp:111-346,
Common/processfields.cpp:142-160, processfields_td.cpp:67-75,
processfields_fd.cpp:212-230), which the reference's readers expect
(matlab/ReadHDF5FieldData.m, ReadHDF5Mesh.m):

* root attribute ``openEMS_HDF5_version`` = 0.2 (double);
* ``/Mesh/x``, ``/Mesh/y``, ``/Mesh/z``: float32 mesh lines of the dump box
  (dual lines where the dump uses the dual mesh) times the drawing unit;
* time domain: ``/FieldData/TD/<step, 8 digits>`` float32 [3][nz][ny][nx], attribute
  ``time`` (float32, seconds);
* frequency domain: ``/FieldData/FD`` attribute ``frequency`` (double array),
  ``/FieldData/FD/f<n>_real`` and ``f<n>_imag`` float32 [3][nz][ny][nx], each
  with attribute ``frequency`` (float32).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

_hid = C.c_int64
_hsize = C.c_uint64
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5F_ACC_TRUNC = 0x0002
_H5F_ACC_RDONLY = 0x0000
_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    cands = [os.environ.get("OPENEMS_HDF5_LIB", ""), ctypes.util.find_library("hdf5") or "",
             "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            lib = C.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise RuntimeError(f"the HDF5 C library is not available ({err}); set OPENEMS_HDF5_LIB")
    sig = {
        "H5open": (C.c_int, []),
        "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
        "H5Fclose": (C.c_int, [_hid]),
        "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
        "H5Gopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Gclose": (C.c_int, [_hid]),
        "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
        "H5Sclose": (C.c_int, [_hid]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
        "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Dwrite": (C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dread": (C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dclose": (C.c_int, [_hid]),
        "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
        "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Awrite": (C.c_int, [_hid, _hid, C.c_void_p]),
        "H5Aread": (C.c_int, [_hid, _hid, C.c_void_p]),
        "H5Aget_space": (_hid, [_hid]),
        "H5Aclose": (C.c_int, [_hid]),
        "H5Lexists": (C.c_int, [_hid, C.c_char_p, _hid]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.H5open() < 0:
        raise RuntimeError("H5open failed")
    # errors come back as negative returns; keep the library's stack printer quiet
    lib.H5Eset_auto2.restype = C.c_int
    lib.H5Eset_auto2.argtypes = [_hid, C.c_void_p, C.c_void_p]
    lib.H5Eset_auto2(0, None, None)
    lib.T_FLOAT = _hid.in_dll(lib, "H5T_NATIVE_FLOAT_g").value
    lib.T_DOUBLE = _hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    _lib = lib
    return lib


def _ok(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def _space(shape):
    L = _load()
    dims = (_hsize * len(shape))(*[int(v) for v in shape])
    return _ok(L.H5Screate_simple(len(shape), dims, None), "H5Screate_simple")


def _write_attr(obj, name, arr):
    L = _load()
    arr = np.ascontiguousarray(arr)
    t = L.T_FLOAT if arr.dtype == np.float32 else L.T_DOUBLE
    sp = _space([arr.size])
    a = _ok(L.H5Acreate2(obj, name.encode(), t, sp, _H5P_DEFAULT, _H5P_DEFAULT), "H5Acreate2 " + name)
    _ok(L.H5Awrite(a, t, arr.ctypes.data), "H5Awrite " + name)
    L.H5Aclose(a)
    L.H5Sclose(sp)


def _write_data(loc, name, arr, attrs=()):
    L = _load()
    arr = np.ascontiguousarray(arr, np.float32)
    sp = _space(arr.shape)
    d = _ok(L.H5Dcreate2(loc, name.encode(), L.T_FLOAT, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
            "H5Dcreate2 " + name)
    _ok(L.H5Dwrite(d, L.T_FLOAT, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, arr.ctypes.data), "H5Dwrite " + name)
    for an, av in attrs:
        _write_attr(d, an, av)
    L.H5Dclose(d)
    L.H5Sclose(sp)


def _group(f, path):
    """Open or create a group path such as /FieldData/TD."""
    L = _load()
    cur = ""
    g = None
    for part in [p for p in path.split("/") if p]:
        cur += "/" + part
        if g is not None:
            L.H5Gclose(g)
        if L.H5Lexists(f, cur.encode(), _H5P_DEFAULT) > 0:
            g = _ok(L.H5Gopen2(f, cur.encode(), _H5P_DEFAULT), "H5Gopen2 " + cur)
        else:
            g = _ok(L.H5Gcreate2(f, cur.encode(), _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT), "H5Gcreate2 " + cur)
    return g


def _mesh_lines(op, dump):
    """Mesh lines of the dump box as HDF5_File_Writer::WriteRectMesh stores
    them: float32 line times float32 drawing unit."""
    gd = np.float32(float(op.scene.get("delta_unit", 1.0)))
    out = []
    for n in range(3):
        pos = range(int(dump.start[n]), int(dump.stop[n]) + 1)
        out.append(np.array([np.float32(op.disc_line(n, i, dump.dual_mesh)) * gd for i in pos], np.float32))
    return out


def _new_file(path, op, dump):
    L = _load()
    f = _ok(L.H5Fcreate(path.encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), "H5Fcreate " + path)
    _write_attr(f, "openEMS_HDF5_version", np.array([0.2], np.float64))
    g = _group(f, "/Mesh")
    for name, lines in zip(("x", "y", "z"), _mesh_lines(op, dump)):
        _write_data(g, name, lines)
    L.H5Gclose(g)
    return f


def _zyx(field):
    """[3][nx][ny][nz] -> the writer's [3][nz][ny][nx]."""
    return np.ascontiguousarray(np.transpose(np.asarray(field, np.float32), (0, 3, 2, 1)))


def write_td_dump(path, op, dump, ts, times, fields):
    """A time-domain dump (processing.FieldDumpTD) and its collected samples
    as <path> (conventionally <name>.h5)."""
    L = _load()
    f = _new_file(path, op, dump)
    g = _group(f, "/FieldData/TD")
    for k in range(len(ts)):
        _write_data(g, "%08d" % int(ts[k]), _zyx(fields[k]), [("time", np.array([times[k]], np.float32))])
    L.H5Gclose(g)
    L.H5Fclose(f)
    return path


def write_fd_dump(path, op, dump, fd, freqs):
    """A frequency-domain dump (processing.FieldDumpFD), fd[nfreq][3][nx][ny][nz]
    complex64, as <path>."""
    L = _load()
    f = _new_file(path, op, dump)
    g = _group(f, "/FieldData/FD")
    _write_attr(g, "frequency", np.asarray(freqs, np.float64))
    for n, fr in enumerate(freqs):
        a = [("frequency", np.array([fr], np.float32))]
        _write_data(g, f"f{n}_real", _zyx(np.real(fd[n])), a)
        _write_data(g, f"f{n}_imag", _zyx(np.imag(fd[n])), a)
    L.H5Gclose(g)
    L.H5Fclose(f)
    return path


def read_dataset(path, name):
    """(float32 array, {attribute: value}) of one dataset; attributes read
    are those named in the layout above."""
    L = _load()
    f = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen " + path)
    d = _ok(L.H5Dopen2(f, name.encode(), _H5P_DEFAULT), "H5Dopen2 " + name)
    sp = L.H5Dget_space(d)
    nd = L.H5Sget_simple_extent_ndims(sp)
    dims = (_hsize * nd)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    out = np.zeros(tuple(int(v) for v in dims), np.float32)
    _ok(L.H5Dread(d, L.T_FLOAT, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data), "H5Dread")
    L.H5Sclose(sp)
    attrs = {}
    for an in ("time", "frequency"):
        if _attr_exists(d, an):
            attrs[an] = read_attribute_of(d, an, np.float32)
    L.H5Dclose(d)
    L.H5Fclose(f)
    return out, attrs


def _attr_exists(obj, name):
    L = _load()
    fn = getattr(L, "H5Aexists")
    fn.restype = C.c_int
    fn.argtypes = [_hid, C.c_char_p]
    return fn(obj, name.encode()) > 0


def read_attribute_of(obj, name, dtype):
    L = _load()
    a = _ok(L.H5Aopen(obj, name.encode(), _H5P_DEFAULT), "H5Aopen " + name)
    sp = L.H5Aget_space(a)
    dims = (_hsize * 1)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    out = np.zeros(int(dims[0]), dtype)
    _ok(L.H5Aread(a, L.T_FLOAT if dtype == np.float32 else L.T_DOUBLE, out.ctypes.data), "H5Aread")
    L.H5Sclose(sp)
    L.H5Aclose(a)
    return out


def read_attribute(path, obj_path, name, dtype=np.float64, group=False):
    """An attribute of the root ("/"), a group (group=True) or a dataset."""
    L = _load()
    f = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen " + path)
    if obj_path == "/":
        obj, close = f, None
    elif group:
        obj, close = _ok(L.H5Gopen2(f, obj_path.encode(), _H5P_DEFAULT), "H5Gopen2 " + obj_path), L.H5Gclose
    else:
        obj, close = _ok(L.H5Dopen2(f, obj_path.encode(), _H5P_DEFAULT), "H5Dopen2 " + obj_path), L.H5Dclose
    out = read_attribute_of(obj, name, dtype)
    if close:
        close(obj)
    L.H5Fclose(f)
    return out


class _GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", _hsize), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


def list_group(path, group):
    """Names of the links in a group, in name order."""
    L = _load()
    for name, res, args in (("H5Gget_info", C.c_int, [_hid, C.POINTER(_GInfo)]),
                            ("H5Lget_name_by_idx", C.c_ssize_t,
                             [_hid, C.c_char_p, C.c_int, C.c_int, _hsize, C.c_char_p, C.c_size_t, _hid])):
        getattr(L, name).restype = res
        getattr(L, name).argtypes = args
    f = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen " + path)
    g = _ok(L.H5Gopen2(f, group.encode(), _H5P_DEFAULT), "H5Gopen2 " + group)
    info = _GInfo()
    _ok(L.H5Gget_info(g, C.byref(info)), "H5Gget_info")
    out = []
    buf = C.create_string_buffer(256)
    for k in range(info.nlinks):
        n = _ok(L.H5Lget_name_by_idx(g, b".", 0, 0, k, buf, 256, _H5P_DEFAULT), "H5Lget_name_by_idx")
        out.append(buf.value[:n].decode())
    L.H5Gclose(g)
    L.H5Fclose(f)
    return out


def exists(path, name):
    L = _load()
    f = _ok(L.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen " + path)
    r = L.H5Lexists(f, name.encode(), _H5P_DEFAULT) > 0
    L.H5Fclose(f)
    return r
