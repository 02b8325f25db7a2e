"""Output of the engine state in the reference's formats, without VTK.

* VTU point data with the reference's array names and layout
  (VtkParticleWriter::appendNodes, rw/vtkParticleWriter.cpp:39-272): points
  are the current positions, every array is Float64; vector arrays
  Displacement, Velocity, Force_Density and Force (= f V) have 3 components,
  Fixity, Particle_ID, Zone_ID, Nodal_Volume, Damage_Z and Theta one.  The
  file is the XML UnstructuredGrid format (ASCII, 17 significant digits), so
  ParaView and VTK read it as they read the reference's output.
* Strain and stress at the mesh quadrature points (the Strain_Stress tag,
  VtkParticleWriter::appendStrainStress, rw/vtkParticleWriter.cpp:444-495).
* The compressive-test curve as CSV (t, wall_penetration, tot_reaction_force):
  the values DEMModel::ppCompressiveTest computes (demModel.cpp:2202-2250),
  written as numbers (the reference's format string writes literal text).
"""
from __future__ import annotations

import numpy as np

from . import capi

VECTOR_TAGS = ("Displacement", "Velocity", "Force_Density", "Force")
SCALAR_TAGS = ("Fixity", "Particle_ID", "Zone_ID", "Nodal_Volume", "Damage_Z", "Theta")
DEFAULT_TAGS = VECTOR_TAGS + SCALAR_TAGS


def _array(name, data, ncomp):
    flat = np.asarray(data, dtype=np.float64).reshape(-1)
    body = " ".join(f"{v:.17g}" for v in flat)
    head = f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">'
    return head + body + "</DataArray>\n"


def state_arrays(eng, spec, tags=DEFAULT_TAGS):
    """Point coordinates and the requested point arrays of an engine's state."""
    vol = spec.vol
    f = eng.field(capi.FIELD_F)
    out = {}
    for t in tags:
        if t == "Displacement":
            out[t] = eng.field(capi.FIELD_U)
        elif t == "Velocity":
            out[t] = eng.field(capi.FIELD_V)
        elif t == "Force_Density":
            out[t] = f
        elif t == "Force":
            out[t] = f * vol[:, None]
        elif t == "Fixity":
            out[t] = spec.fix.astype(np.float64)
        elif t == "Particle_ID":
            out[t] = spec.pt_id.astype(np.float64)
        elif t == "Zone_ID":
            out[t] = np.array([p["zone"] for p in spec.particles], np.float64)[spec.pt_id]
        elif t == "Nodal_Volume":
            out[t] = vol
        elif t == "Damage_Z":
            out[t] = eng.field(capi.FIELD_Z).astype(np.float64)
        elif t == "Theta":
            out[t] = eng.field(capi.FIELD_THETA)
        else:
            raise ValueError(f"unknown output tag {t}")
    return eng.field(capi.FIELD_X), out


def write_vtu(path, eng, spec, tags=DEFAULT_TAGS):
    x, arrays = state_arrays(eng, spec, tags)
    n = len(x)
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" '
                 'byte_order="LittleEndian">\n<UnstructuredGrid>\n'
                 f'<Piece NumberOfPoints="{n}" NumberOfCells="0">\n<PointData>\n')
        for t in tags:
            fh.write(_array(t, arrays[t], 3 if t in VECTOR_TAGS else 1))
        fh.write("</PointData>\n<Points>\n")
        fh.write(_array("Points", x, 3))
        fh.write('</Points>\n<Cells>\n<DataArray type="Int64" Name="connectivity" format="ascii"></DataArray>\n'
                 '<DataArray type="Int64" Name="offsets" format="ascii"></DataArray>\n'
                 '<DataArray type="UInt8" Name="types" format="ascii"></DataArray>\n</Cells>\n'
                 "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def read_vtu(path):
    """Points and point arrays of a VTU written by write_vtu (or any ASCII
    Float64 VTU): (points [n][3], {name: array})."""
    import xml.etree.ElementTree as ET

    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    n = int(piece.get("NumberOfPoints"))

    def arr(el):
        nc = int(el.get("NumberOfComponents", "1"))
        v = np.array((el.text or "").split(), dtype=np.float64)
        return v.reshape(n, nc) if nc > 1 else v

    pts = arr(piece.find("Points/DataArray"))
    data = {el.get("Name"): arr(el) for el in piece.find("PointData")}
    return pts, data


def write_strain_vtu(path, xq, strain, stress):
    """The Strain_Stress output (VtkParticleWriter::appendStrainStress,
    rw/vtkParticleWriter.cpp:444-495): quadrature points as points, Strain
    and Stress with 6 components (xx yy zz yz xz xy)."""
    xq = np.asarray(xq)
    n = len(xq)
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" '
                 'byte_order="LittleEndian">\n<UnstructuredGrid>\n'
                 f'<Piece NumberOfPoints="{n}" NumberOfCells="0">\n<PointData>\n')
        names = ' '.join(f'ComponentName{i}="{c}"' for i, c in enumerate(("xx", "yy", "zz", "yz", "xz", "xy")))
        for tag, a in (("Strain", strain), ("Stress", stress)):
            body = " ".join(f"{float(v):.9g}" for v in np.asarray(a).reshape(-1))
            fh.write(f'<DataArray type="Float64" Name="{tag}" NumberOfComponents="6" {names} format="ascii">'
                     + body + "</DataArray>\n")
        fh.write("</PointData>\n<Points>\n")
        fh.write(_array("Points", xq, 3))
        fh.write('</Points>\n<Cells>\n<DataArray type="Int64" Name="connectivity" format="ascii"></DataArray>\n'
                 '<DataArray type="Int64" Name="offsets" format="ascii"></DataArray>\n'
                 '<DataArray type="UInt8" Name="types" format="ascii"></DataArray>\n</Cells>\n'
                 "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def restart_from_vtu(eng, spec, path, step):
    """DEMModel::restart (demModel.cpp:84-103) from a VTU written by
    write_vtu or by the reference: current positions from the points,
    Displacement (or x - xRef when absent) and Velocity (required), as
    VtkParticleReader::readNodes (vtkParticleReader.cpp:41-115) reads them;
    then the step counter, time and the controller's restart rule.  Call
    before integrate_begin."""
    pts, data = read_vtu(path)
    if len(pts) != spec.n_nodes:
        raise ValueError(f"{path}: {len(pts)} points, the model has {spec.n_nodes} nodes")
    if "Velocity" not in data:
        raise ValueError(f"{path}: restart file does not have Velocity data")
    x = np.asarray(pts, np.float64).reshape(-1, 3)
    u = np.asarray(data["Displacement"], np.float64).reshape(-1, 3) if "Displacement" in data else x - spec.x_ref
    eng.set_field(capi.FIELD_X, x)
    eng.set_field(capi.FIELD_U, u)
    eng.set_field(capi.FIELD_V, np.asarray(data["Velocity"], np.float64).reshape(-1, 3))
    eng.restart(step)


def write_compressive_csv(path, rows):
    """rows: iterable of (t, wall_penetration, tot_reaction_force)."""
    with open(path, "w") as fh:
        fh.write("t, wall_penetration, tot_reaction_force\n")
        for t, pen, reac in rows:
            fh.write(f"{t:.17g}, {pen:.17g}, {reac:.17g}\n")
