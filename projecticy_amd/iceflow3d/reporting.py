"""Output formats of the 3D dam-break harness (iceflow3d/reporting.py:1-256).

- ``write_vti``: one VTK XML ImageData file per sample, cell data in SI cell
  coordinates, arrays appended raw (little-endian, UInt64 block headers),
  x fastest with vector components interleaved, so ParaView reads the
  files the reference writes and these alike;
- ``write_pvd``: the ParaView time-series collection over those files;
- ``render_frame``: a PNG of the sample (mid-depth slices of phase and ice,
  temperature and speed, and a top view of the ice) with matplotlib;
- ``FrameGif``: an animated GIF of the PNG frames, built when closed and
  reading one frame at a time.
"""

from pathlib import Path
import struct
import xml.etree.ElementTree as ET

import numpy as np


def _vtk_payload(values):
    """(VTK type name, component count, bytes) of one cell array: integers as
    Int32, everything else as Float32, x fastest."""
    values = np.asarray(values)
    integer = np.issubdtype(values.dtype, np.integer) or values.dtype == np.bool_
    dtype = np.dtype("<i4") if integer else np.dtype("<f4")
    components = values.shape[3] if values.ndim == 4 else 1
    axes = (2, 1, 0, 3) if values.ndim == 4 else (2, 1, 0)
    payload = np.ascontiguousarray(values.astype(dtype, copy=False).transpose(axes))
    return ("Int32" if integer else "Float32"), components, payload


def write_vti(path, fields, dx):
    """Write ``fields`` (name -> [nx, ny, nz] or [nx, ny, nz, c] arrays) as VTK
    ImageData cell data with spacing ``dx`` metres (reporting.py:48-92)."""
    shape = np.asarray(fields["water_phase"]).shape
    extent = " ".join(f"0 {n}" for n in shape)
    root = ET.Element("VTKFile", type="ImageData", version="1.0", byte_order="LittleEndian", header_type="UInt64")
    image = ET.SubElement(root, "ImageData", WholeExtent=extent, Origin="0 0 0", Spacing=f"{dx} {dx} {dx}")
    piece = ET.SubElement(image, "Piece", Extent=extent)
    cell_data = ET.SubElement(piece, "CellData", Scalars="water_phase", Vectors="velocity_m_s")
    blocks, offset = [], 0
    for name, values in fields.items():
        kind, components, payload = _vtk_payload(values)
        ET.SubElement(cell_data, "DataArray", type=kind, Name=name, NumberOfComponents=str(components),
                      format="appended", offset=str(offset))
        blocks.append(payload)
        offset += 8 + payload.nbytes
    header = ET.tostring(root, encoding="utf-8")
    closing = b"</VTKFile>"
    if not header.endswith(closing):
        raise RuntimeError("unexpected XML serialisation of the VTK header")
    with Path(path).open("wb") as stream:
        stream.write(b'<?xml version="1.0"?>\n')
        stream.write(header[: -len(closing)])
        stream.write(b'<AppendedData encoding="raw">_')
        for block in blocks:
            stream.write(struct.pack("<Q", block.nbytes))
            stream.write(block.tobytes())
        stream.write(b"</AppendedData></VTKFile>\n")


def read_vti(path):
    """The cell arrays of a file written by ``write_vti``, as [nx, ny, nz(, c)]
    arrays (for checks and post-processing)."""
    raw = Path(path).read_bytes()
    start = raw.index(b'<AppendedData encoding="raw">_') + len(b'<AppendedData encoding="raw">_')
    header = raw[: raw.index(b"<AppendedData")] + b"</VTKFile>"
    root = ET.fromstring(header.split(b"\n", 1)[1])
    extent = [int(v) for v in root.find("ImageData").get("WholeExtent").split()]
    shape = tuple(extent[1::2])
    out = {}
    for array in root.iter("DataArray"):
        at = start + int(array.get("offset"))
        size = struct.unpack("<Q", raw[at : at + 8])[0]
        dtype = "<i4" if array.get("type") == "Int32" else "<f4"
        flat = np.frombuffer(raw[at + 8 : at + 8 + size], dtype=dtype)
        components = int(array.get("NumberOfComponents"))
        if components == 1:
            out[array.get("Name")] = flat.reshape(shape[::-1]).transpose(2, 1, 0)
        else:
            out[array.get("Name")] = flat.reshape(*shape[::-1], components).transpose(2, 1, 0, 3)
    return out


def write_pvd(path, samples):
    """ParaView collection of (time_s, file name) samples (reporting.py:95-111)."""
    root = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
    collection = ET.SubElement(root, "Collection")
    for time_s, name in samples:
        ET.SubElement(collection, "DataSet", timestep=f"{time_s:.16g}", group="", part="0", file=str(name))
    ET.ElementTree(root).write(path, encoding="utf-8", xml_declaration=True)


def render_frame(path, config, scales, fields, time_s):
    """PNG of one sample: mid-depth phase/ice, temperature and speed slices
    and the ice seen from above (the reference draws the first panel as a 3D
    voxel surface; the slices carry the same information at a fraction of
    the cost)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    phi, ids = np.asarray(fields["water_phase"]), np.asarray(fields["solid_body_id"])
    mid = phi.shape[2] // 2
    mm = scales.dx_m * 1000.0
    lengths = np.array(phi.shape) * mm
    side = (0, lengths[0], 0, lengths[1])
    solid = ids[:, :, mid] != -1
    air = (phi[:, :, mid] < 0.5) & ~solid
    temperature = np.where(air, config.thermal.initial_air_temperature_c, fields["temperature_c"][:, :, mid])
    speed = np.linalg.norm(np.asarray(fields["velocity_m_s"])[:, :, mid], axis=-1)
    fig, axes = plt.subplots(2, 2, figsize=(13, 7), constrained_layout=True)
    panels = (
        (axes[0, 0], np.where(ids[:, :, mid] >= 0, 2.0, phi[:, :, mid]), "Blues", 0.0, 2.0, "Water phase, ice = 2"),
        (axes[0, 1], temperature, "inferno", 0.0, config.thermal.initial_water_temperature_c, "Temperature [degC]"),
        (axes[1, 0], speed, "viridis", 0.0, 0.8, "Speed [m/s]"),
    )
    for axis, values, cmap, lo, hi, label in panels:
        masked = np.ma.masked_where(solid & (label != "Water phase, ice = 2"), values)
        image = axis.imshow(masked.T, origin="lower", extent=side, cmap=cmap, vmin=lo, vmax=hi, aspect="equal")
        axis.set(xlabel="x [mm]", ylabel="y [mm]", title=f"{label}, z = {(mid + 0.5) * mm:.2f} mm")
        fig.colorbar(image, ax=axis, shrink=0.8)
    top = (ids >= 0).any(axis=1)
    axes[1, 1].imshow(top.T, origin="lower", extent=(0, lengths[0], 0, lengths[2]), cmap="Oranges", vmin=0,
                      vmax=1.5, aspect="equal")
    axes[1, 1].set(xlabel="x [mm]", ylabel="z [mm]", title="Ice seen from above")
    fig.suptitle(f"3D water and melting ice | t = {time_s:.6f} s")
    fig.savefig(path, dpi=100)
    plt.close(fig)


class FrameGif:
    """An animated GIF of PNG frames, written when closed from the frames on
    disk, one decoded frame at a time."""

    def __init__(self, path, duration_ms=120):
        self.path = Path(path)
        self.frames = []
        self.duration_ms = duration_ms

    def append(self, png):
        self.frames.append(Path(png))

    def close(self):
        if not self.frames:
            self.path.unlink(missing_ok=True)
            return
        from PIL import Image

        def converted(paths):
            for p in paths:
                with Image.open(p) as image:
                    yield image.convert("RGB").convert("P", palette=Image.Palette.ADAPTIVE, colors=256)

        frames = converted(self.frames)
        first = next(frames)
        first.save(self.path, save_all=True, append_images=frames, duration=self.duration_ms, loop=0, disposal=2)
        self.frames = []
