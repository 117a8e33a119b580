"""Field dumps as VTK files (dump FileType 0), the layout ProcessFieldsTD /
ProcessFieldsFD write through VTK_File_Writer (tools/vtk_file_writer.cpp,
Common/processfields_td.cpp:58-64, processfields_fd.cpp:112-209):

* an XML RectilinearGrid file (.vtr) per sample, binary data (base64, no
  compression), coordinates the dump box's mesh lines times the drawing unit,
  one point-data vector array named "E-Field" or "H-Field", x fastest;
* time domain: ``<name>_<step, 10 digits>.vtr``;
* frequency domain, per frequency f (written as an integer when it is one,
  else with 6 decimals): 21 phase snapshots ``<name>_f=<f>_p=<deg, 3 digits>.vtr``
  holding real(F exp(i 2 pi p / 21)) in complex<float>, then
  ``<name>_f=<f>_abs.vtr`` (|F|) and ``<name>_f=<f>_arg.vtr`` (arg F).
"""
from __future__ import annotations

import base64
import math
import xml.etree.ElementTree as ET

import numpy as np

FIELD_NAMES = {0: "E-Field", 1: "H-Field"}


def _b64(arr):
    raw = np.ascontiguousarray(arr).tobytes()
    return base64.b64encode(np.uint32(len(raw)).tobytes() + raw).decode("ascii")


def write_vtr(path, coords, name, vec):
    """One RectilinearGrid file: coords (x, y, z) 1-D, vec [3][nx][ny][nz]."""
    nx, ny, nz = (len(c) for c in coords)
    ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    pts = np.ascontiguousarray(np.transpose(np.asarray(vec, np.float32), (3, 2, 1, 0)))   # [z][y][x][3]
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">',
             f'  <RectilinearGrid WholeExtent="{ext}">',
             f'    <Piece Extent="{ext}">',
             f'      <PointData Vectors="{name}">',
             f'        <DataArray type="Float32" Name="{name}" NumberOfComponents="3" format="binary">',
             "          " + _b64(pts),
             "        </DataArray>",
             "      </PointData>",
             "      <Coordinates>"]
    for c in coords:
        lines += ['        <DataArray type="Float64" format="binary">',
                  "          " + _b64(np.asarray(c, np.float64)),
                  "        </DataArray>"]
    lines += ["      </Coordinates>", "    </Piece>", "  </RectilinearGrid>", "</VTKFile>", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))
    return path


def read_vtr(path):
    """(coords (x, y, z), name, vec [3][nx][ny][nz]) of a file write_vtr wrote."""
    root = ET.parse(path).getroot()
    piece = root.find("RectilinearGrid/Piece")
    arr = piece.find("PointData/DataArray")

    def dec(el, dtype):
        raw = base64.b64decode(el.text.strip())
        n = int(np.frombuffer(raw[:4], np.uint32)[0])
        return np.frombuffer(raw[4:4 + n], dtype)
    coords = [dec(c, np.float64) for c in piece.findall("Coordinates/DataArray")]
    nx, ny, nz = (len(c) for c in coords)
    pts = dec(arr, np.float32).reshape(nz, ny, nx, 3)
    return coords, arr.get("Name"), np.transpose(pts, (3, 2, 1, 0))


def _coords(op, dump):
    gd = float(op.scene.get("delta_unit", 1.0))
    return [np.array([op.disc_line(n, i, dump.dual_mesh) * gd for i in range(int(dump.start[n]), int(dump.stop[n]) + 1)])
            for n in range(3)]


def _freq_str(f):
    return str(int(f)) if f == int(f) else "%f" % f


def write_td_dump(base, op, dump, ts, fields):
    """<base>_<step>.vtr for every sample; returns the paths."""
    coords = _coords(op, dump)
    name = FIELD_NAMES[dump.field]
    return [write_vtr("%s_%010d.vtr" % (base, int(t)), coords, name, fields[k]) for k, t in enumerate(ts)]


def write_fd_dump(base, op, dump, fd, freqs, phases=21):
    """Phase snapshots, magnitude and phase files per frequency; returns the paths."""
    coords = _coords(op, dump)
    name = FIELD_NAMES[dump.field]
    out = []
    for n, f in enumerate(freqs):
        F = np.asarray(fd[n], np.complex64)
        fs = _freq_str(f)
        for p in range(phases):
            angle = 2.0 * math.pi * p / phases
            e = np.exp(np.complex64(1j * angle))                           # exp(i angle) in complex<float>
            re = (F.real * e.real - F.imag * e.imag).astype(np.float32)
            deg = int(angle * 180 / math.pi)
            out.append(write_vtr(f"{base}_f={fs}_p={deg:03d}.vtr", coords, name, re))
        out.append(write_vtr(f"{base}_f={fs}_abs.vtr", coords, name, np.abs(F).astype(np.float32)))
        out.append(write_vtr(f"{base}_f={fs}_arg.vtr", coords, name, np.angle(F).astype(np.float32)))
    return out
