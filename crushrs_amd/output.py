"""Output formats of a solve: the Parquet history tables of
src/output/history.rs:1-103 and the VTK series of src/output/vtk.rs:1-115,
written from the histories and frames the GPU recorded
(Results.history_*, Results.frame_*).

Parquet goes through pyarrow; VTK is the XML unstructured-grid format
(binary appended, UInt64 headers) with a `.pvd` collection, as ParaView,
VisIt, PyVista and meshio read it.
"""
from __future__ import annotations

from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from . import Model, Results


def accelerometer_names(model: Model, n: int) -> list[str]:
    """`<part>_rear` for each group's accelerometer node, <part> the first
    part whose elements hold the node (the names add_vehicle_accelerometer
    gives, src/vehicle.rs:792-822)."""
    names = model.part_names()
    hexes = model.hexes()
    hex_part = np.ctypeslib.as_array(model.view.hex_part, shape=(model.n_hexes,))
    out = []
    for g in range(n):
        node = model.view.groups[g].accel_node
        if node < 0:
            out.append(f"group{g}")
            continue
        parts = hex_part[np.any(hexes == node, axis=1)]
        out.append(f"{names[int(parts.min())]}_rear" if len(parts) else f"group{g}")
    return out


def _xyz(prefix: str, a: np.ndarray) -> dict:
    return {f"{prefix}x": a[:, 0], f"{prefix}y": a[:, 1], f"{prefix}z": a[:, 2]}


def history_tables(model: Model, results: Results, accelerometers: Optional[Sequence[str]] = None) -> dict:
    """The three tables of write_history (src/output/history.rs:36-103) as
    column dicts: `nodes` (one row per sample and accelerometer: global
    position, displacement, velocity, acceleration and the same in the
    accelerometer's body-fixed frame, 25 columns), `frames` (frame origin
    and axes) and `parts` (mass, mean displacement, velocity, acceleration
    and kinetic energy per part)."""
    if results.history_time is None:
        raise ValueError("the solve recorded no history (settings.history_steps = 0)")
    k, na = results.accel.shape[0], results.accel.shape[1]
    names = list(accelerometers) if accelerometers is not None else accelerometer_names(model, na)
    acc = results.accel
    t = np.repeat(results.history_time, na)
    step = np.repeat(results.history_step, na).astype(np.int64)
    who = np.tile(np.array(names, dtype=object), k)
    node = np.tile(np.array([model.view.groups[g].accel_node for g in range(na)], dtype=np.int64), k)
    flat = acc.reshape(k * na, -1)
    axes = flat[:, 15:24].reshape(-1, 3, 3)

    def local(v):
        return np.einsum("kij,kj->ki", axes, v)

    nodes = {"time": t, "step": step, "accelerometer": who, "node": node}
    nodes.update(_xyz("", flat[:, 0:3]))
    nodes.update(_xyz("u", flat[:, 3:6]))
    nodes.update(_xyz("v", flat[:, 6:9]))
    nodes.update(_xyz("a", flat[:, 9:12]))
    nodes.update(_xyz("lu", local(flat[:, 3:6])))
    nodes.update(_xyz("lv", local(flat[:, 6:9])))
    nodes.update(_xyz("la", local(flat[:, 9:12])))
    framed = np.tile(np.array([model.view.groups[g].frame_origin >= 0 for g in range(na)]), k)
    frames = {"time": t[framed], "step": step[framed], "accelerometer": who[framed]}
    frames.update(_xyz("o", flat[framed, 12:15]))
    for i, ax in enumerate(("ex", "ey", "ez")):
        frames.update(_xyz(f"{ax}_", axes[framed, i, :]))
    parts_names = model.part_names()
    P = results.parts
    n_parts = P.shape[1]
    pf = P.reshape(k * n_parts, -1)
    parts = {"time": np.repeat(results.history_time, n_parts),
             "part": np.tile(np.array(parts_names[:n_parts], dtype=object), k),
             "mass": pf[:, 0]}
    parts.update(_xyz("u", pf[:, 1:4]))
    parts.update(_xyz("v", pf[:, 4:7]))
    parts.update(_xyz("a", pf[:, 7:10]))
    parts["kinetic_energy"] = pf[:, 10]
    return {"nodes": nodes, "frames": frames, "parts": parts}


def write_history(model: Model, results: Results, base, accelerometers: Optional[Sequence[str]] = None) -> list[Path]:
    """`<base>.nodes.parquet`, `<base>.frames.parquet`, `<base>.parts.parquet`
    (write_history, src/output/history.rs:36-103); returns the paths."""
    import pyarrow as pa
    import pyarrow.parquet as pq
    tables = history_tables(model, results, accelerometers)
    out = []
    for name in ("nodes", "frames", "parts"):
        cols = {}
        for c, v in tables[name].items():
            cols[c] = pa.array(list(v), type=pa.string()) if v.dtype == object else pa.array(v)
        path = Path(f"{base}.{name}.parquet")
        path.parent.mkdir(parents=True, exist_ok=True)
        pq.write_table(pa.table(cols), path)
        out.append(path)
    return out


def _block(a: np.ndarray) -> bytes:
    b = np.ascontiguousarray(a).tobytes()
    return np.uint64(len(b)).tobytes() + b


def write_vtu(model: Model, results: Results, k: int, path) -> None:
    """Frame k as a `.vtu` unstructured grid of hexahedra (write_vtu,
    src/output/vtk.rs:34-99): deformed points, point data `displacement`,
    cell data `plastic_strain`, `stress` (6 components), `part`, `eroded`."""
    if results.frame_time is None:
        raise ValueError("the solve recorded no frames (settings.frame_steps = 0)")
    X = model.nodes().astype(np.float32)
    u = results.frame_displacement[k].astype(np.float32)
    hexes = model.hexes().astype(np.int64)
    n_pts, n_cells = X.shape[0], hexes.shape[0]
    hex_part = np.ctypeslib.as_array(model.view.hex_part, shape=(n_cells,)).astype(np.int32)
    arrays = [
        ("Points", _block((X + u).reshape(-1))),
        ("displacement", _block(u.reshape(-1))),
        ("connectivity", _block(hexes.reshape(-1))),
        ("offsets", _block(8 * np.arange(1, n_cells + 1, dtype=np.int64))),
        ("types", _block(np.full(n_cells, 12, np.uint8))),
        ("plastic_strain", _block(results.frame_plastic_strain[k].astype(np.float32))),
        ("stress", _block(results.frame_stress[k].astype(np.float32).reshape(-1))),
        ("part", _block(hex_part)),
        ("eroded", _block(results.frame_eroded[k].astype(np.int32))),
    ]
    off, pos = {}, 0
    for name, b in arrays:
        off[name] = pos
        pos += len(b)
    xml = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        "  <UnstructuredGrid>\n"
        f'    <Piece NumberOfPoints="{n_pts}" NumberOfCells="{n_cells}">\n'
        "      <Points>\n"
        f'        <DataArray type="Float32" NumberOfComponents="3" format="appended" offset="{off["Points"]}"/>\n'
        "      </Points>\n"
        '      <PointData Vectors="displacement">\n'
        f'        <DataArray type="Float32" Name="displacement" NumberOfComponents="3" format="appended" '
        f'offset="{off["displacement"]}"/>\n'
        "      </PointData>\n"
        "      <Cells>\n"
        f'        <DataArray type="Int64" Name="connectivity" format="appended" offset="{off["connectivity"]}"/>\n'
        f'        <DataArray type="Int64" Name="offsets" format="appended" offset="{off["offsets"]}"/>\n'
        f'        <DataArray type="UInt8" Name="types" format="appended" offset="{off["types"]}"/>\n'
        "      </Cells>\n"
        '      <CellData Scalars="plastic_strain">\n'
        f'        <DataArray type="Float32" Name="plastic_strain" format="appended" offset="{off["plastic_strain"]}"/>\n'
        f'        <DataArray type="Float32" Name="stress" NumberOfComponents="6" format="appended" '
        f'offset="{off["stress"]}"/>\n'
        f'        <DataArray type="Int32" Name="part" format="appended" offset="{off["part"]}"/>\n'
        f'        <DataArray type="Int32" Name="eroded" format="appended" offset="{off["eroded"]}"/>\n'
        "      </CellData>\n"
        "    </Piece>\n"
        "  </UnstructuredGrid>\n"
        '  <AppendedData encoding="raw">\n   _'
    )
    path = Path(path)
    with open(path, "wb") as f:
        f.write(xml.encode())
        for _, b in arrays:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def write_series(model: Model, results: Results, base) -> Path:
    """Every frame as `<base>_NNNN.vtu` plus `<base>.pvd` (write_series,
    src/output/vtk.rs:101-115); returns the .pvd path."""
    base = Path(base)
    base.parent.mkdir(parents=True, exist_ok=True)
    stem = base.name
    lines = ['<?xml version="1.0"?>\n<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">\n'
             "  <Collection>\n"]
    for k in range(len(results.frame_time)):
        name = f"{stem}_{k:04d}.vtu"
        write_vtu(model, results, k, base.parent / name)
        lines.append(f'    <DataSet timestep="{results.frame_time[k]!r}" file="{name}"/>\n')
    lines.append("  </Collection>\n</VTKFile>\n")
    pvd = base.parent / f"{stem}.pvd"
    pvd.write_text("".join(lines))
    return pvd


def read_vtu(path) -> dict:
    """The arrays of a `.vtu` written by write_vtu, by name (for checks)."""
    import re
    data = Path(path).read_bytes()
    head, _, rest = data.partition(b'<AppendedData encoding="raw">\n   _')
    types = {"Float32": np.float32, "Int64": np.int64, "UInt8": np.uint8, "Int32": np.int32}
    out = {}
    for m in re.finditer(rb'<DataArray type="(\w+)"(?: Name="(\w+)")?[^>]*offset="(\d+)"', head):
        ty, name, off = m.group(1).decode(), (m.group(2) or b"Points").decode(), int(m.group(3))
        n = int(np.frombuffer(rest[off:off + 8], np.uint64)[0])
        out[name] = np.frombuffer(rest[off + 8:off + 8 + n], types[ty])
    return out
