"""Output and restart files of a time-stepping run.

``write_vtr`` writes a VTK XML rectilinear-grid file with the fields of
``Print_Result_VTK_2D`` (PrintResult.f90:165-190, written through
VTR_mod.f90:79-180,212-262,330-345,395-410): the cell centres as the mesh
(extent 1..ni x 1..nj x 1..1), point data Velocity (u, v, 0) with its
X_Velocity / Y_Velocity components, Pressure, Levelset, SolidLvs, Vof and
SolidVof, every value as Float32, i running fastest.  ``format="ascii"``
gives the reference's text form; ``"binary"`` the inline base64 form of the
same arrays (a UInt32 byte count before each one), which ParaView reads as
well and which is ~4x smaller and much faster to write at 2048^2.  As in the
reference, u and v are the face values U(1:ni, 1:nj) and V(1:ni, 1:nj) put
at the cell centres.

``save_checkpoint`` / ``load_checkpoint`` hold a context's device state
(sg_state_save / sg_state_load: FP64 as held, so a resumed run repeats the
uninterrupted one bit for bit) with the host's step counter and time in a
JSON header, in place of the Tecplot-ASCII restart of RunAgain
(Solver_Mod.f90:164-195).
"""
import base64
import json
import os
import struct

import numpy as np

VTR_FIELDS = ("Velocity", "X_Velocity", "Y_Velocity", "Pressure", "Levelset", "SolidLvs", "Vof", "SolidVof")


def vtr_name(directory, itt, prefix="FlowField"):
    """The reference's file name: <dir>/Paraview/<prefix>_<itt:08d>.vtr."""
    return os.path.join(directory, "Paraview", f"{prefix}_{int(itt):08d}.vtr")


def _array(f, name, a, ncomp, fmt):
    a = np.ascontiguousarray(a, dtype=np.float32)
    f.write(f'        <DataArray type="Float32" Name="{name}" NumberOfComponents="{ncomp}" format="{fmt}">\n')
    if fmt == "ascii":
        flat = a.ravel()
        for k in range(0, flat.size, 8):
            f.write(" ".join(f"{float(x):.9g}" for x in flat[k:k + 8]) + "\n")
    else:
        raw = a.astype("<f4").tobytes()
        f.write(base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii") + "\n")
    f.write("        </DataArray>\n")


def write_vtr(path, grid, u, v, p, phi, phiS, vof, vofS, fmt="ascii"):
    """u, v, p ghosted ((nj+2, ni+2), the device layout of U, V, P); phi,
    phiS, vof, vofS cell arrays (nj, ni)."""
    if fmt not in ("ascii", "binary"):
        raise ValueError("fmt must be 'ascii' or 'binary'")
    ni, nj = grid.ni, grid.nj
    uc = np.asarray(u)[1:nj + 1, 1:ni + 1]
    vc = np.asarray(v)[1:nj + 1, 1:ni + 1]
    pc = np.asarray(p)[1:nj + 1, 1:ni + 1]
    for name, a in (("phi", phi), ("phiS", phiS), ("vof", vof), ("vofS", vofS)):
        if np.shape(a) != (nj, ni):
            raise ValueError(f"{name} must be a cell array of shape {(nj, ni)}")
    vel = np.stack([uc, vc, np.zeros_like(uc)], axis=-1)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        attrs = 'byte_order="LittleEndian" ' if fmt == "binary" else ""
        f.write(f'<VTKFile type="RectilinearGrid" version="0.1" {attrs}format="{fmt}">\n')
        f.write(f' <RectilinearGrid WholeExtent="1 {ni} 1 {nj} 1 1">\n')
        f.write(f'  <Piece Extent="1 {ni} 1 {nj} 1 1">\n')
        f.write("   <Coordinates>\n")
        _array(f, "X_COORDINATES", grid.xc, 1, fmt)
        _array(f, "Y_COORDINATES", grid.yc, 1, fmt)
        _array(f, "Z_COORDINATES", np.zeros(1), 1, fmt)
        f.write("   </Coordinates>\n")
        f.write("   <PointData>\n")
        _array(f, "Velocity", vel, 3, fmt)
        _array(f, "X_Velocity", uc, 1, fmt)
        _array(f, "Y_Velocity", vc, 1, fmt)
        _array(f, "Pressure", pc, 1, fmt)
        _array(f, "Levelset", phi, 1, fmt)
        _array(f, "SolidLvs", phiS, 1, fmt)
        _array(f, "Vof", vof, 1, fmt)
        _array(f, "SolidVof", vofS, 1, fmt)
        f.write("   </PointData>\n")
        f.write("  </Piece>\n")
        f.write(" </RectilinearGrid>\n")
        f.write("</VTKFile>\n")


def read_vtr(path):
    """The arrays of a file written by write_vtr, as float32 numpy arrays
    (flattened, i fastest; Velocity as (n, 3)), and the extent."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("RectilinearGrid/Piece")
    ext = [int(t) for t in piece.get("Extent").split()]
    out = {}
    for da in piece.iter("DataArray"):
        text = da.text.strip()
        if da.get("format") == "ascii":
            a = np.array(text.split(), dtype=np.float32)
        else:
            raw = base64.b64decode(text)
            n = struct.unpack("<I", raw[:4])[0]
            a = np.frombuffer(raw[4:4 + n], dtype="<f4").copy()
        k = int(da.get("NumberOfComponents"))
        out[da.get("Name")] = a.reshape(-1, k) if k > 1 else a
    return out, ext


_CKPT_MAGIC = b"SGCKPT01"


def save_checkpoint(path, pj, itt, time, extra=None):
    """Device state of Projection pj (sg_state_save) + itt, time and any
    JSON-able extra, in one file."""
    blob = pj.state_save()
    meta = json.dumps(dict(itt=int(itt), time=float(time), ni=pj.ni, nj=pj.nj, extra=extra or {})).encode()
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(_CKPT_MAGIC)
        f.write(struct.pack("<QQ", len(meta), len(blob)))
        f.write(meta)
        f.write(blob)
    os.replace(tmp, path)


def load_checkpoint(path, pj):
    """Restores the device state of a file written by save_checkpoint into
    Projection pj (same grid, geometry already set); returns the header."""
    with open(path, "rb") as f:
        if f.read(8) != _CKPT_MAGIC:
            raise ValueError(f"{path}: not a checkpoint of this library")
        nmeta, nblob = struct.unpack("<QQ", f.read(16))
        meta = json.loads(f.read(nmeta).decode())
        blob = f.read(nblob)
    if len(blob) != nblob:
        raise ValueError(f"{path}: truncated")
    if (meta["ni"], meta["nj"]) != (pj.ni, pj.nj):
        raise ValueError(f"{path}: saved on a {meta['ni']}x{meta['nj']} grid, not {pj.ni}x{pj.nj}")
    pj.state_load(blob)
    return meta
