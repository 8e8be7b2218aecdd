"""RVF vector-field files (the reference's ``src/rvf_io.c:11-131``).

Format, as the reference C writes and reads it:

* a text header of two lines, ``"%zu %zu %zu\\n%f %f %f\\n"`` -- the sizes
  ``nx ny nz`` and the spacing ``dx dy dz`` (C ``%f``: six decimals);
* then ``3 * nx * ny * nz`` little-endian doubles, the three components of
  each voxel together (x, y, z), voxels with x fastest, then y, then z.

The reference's Python reader (``disptools/io.py:167-191``) reads the data
with the library's float type and reshapes it as ``(nx, ny, nz, 3)``; that
matches its C writer only for FLOAT_SIZE=64 and cubic volumes.  These
functions follow the C reader and writer.

Fields are numpy arrays ``[3, nz, ny, nx]`` (the library's layout) or
``[nz, ny, nx, 3]`` (the layout of ``sitk.GetArrayFromImage`` of a vector
image, ``layout="zyx3"``).
"""

import numpy as np

__all__ = ["write_rvf", "read_rvf", "encode_rvf", "decode_rvf"]


def _header(shape_xyz, spacing):
    nx, ny, nz = (int(v) for v in shape_xyz)
    # the spacing is the library's float, printed with %f
    dx, dy, dz = (float(np.float32(v)) for v in spacing)
    return ("%d %d %d\n%f %f %f\n" % (nx, ny, nz, dx, dy, dz)).encode("ascii")


def encode_rvf(field, spacing=(1.0, 1.0, 1.0), layout="3zyx"):
    """Bytes of an RVF file holding `field`."""
    f = np.asarray(field)
    if layout == "3zyx":
        if f.ndim != 4 or f.shape[0] != 3:
            raise ValueError("field must be [3, nz, ny, nx]")
        inter = np.moveaxis(f, 0, -1)
    elif layout == "zyx3":
        if f.ndim != 4 or f.shape[-1] != 3:
            raise ValueError("field must be [nz, ny, nx, 3]")
        inter = f
    else:
        raise ValueError("layout must be '3zyx' or 'zyx3'")
    nz, ny, nx = inter.shape[:3]
    data = np.ascontiguousarray(inter, dtype="<f8").tobytes()
    return _header((nx, ny, nz), spacing) + data


def decode_rvf(buf, dtype=np.float32, layout="3zyx"):
    """(field, spacing) from the bytes of an RVF file."""
    mv = memoryview(buf)
    end1 = bytes(mv[:256]).find(b"\n")
    if end1 < 0:
        raise ValueError("rvf: expected size (3 integers)")
    sizes = bytes(mv[:end1]).split()
    end2 = bytes(mv[end1 + 1:end1 + 257]).find(b"\n")
    if len(sizes) != 3 or end2 < 0:
        raise ValueError("rvf: expected size (3 integers)")
    spacing_f = bytes(mv[end1 + 1:end1 + 1 + end2]).split()
    if len(spacing_f) != 3:
        raise ValueError("rvf: expected spacing (3 floats)")
    nx, ny, nz = (int(v) for v in sizes)
    # the C reader parses the spacing with %f into a float
    spacing = tuple(float(np.float32(float(v))) for v in spacing_f)
    off = end1 + 1 + end2 + 1
    count = 3 * nx * ny * nz
    if len(mv) - off < 8 * count:
        raise ValueError("rvf: truncated data (%d of %d bytes)" % (len(mv) - off, 8 * count))
    data = np.frombuffer(mv[off:off + 8 * count], dtype="<f8").reshape(nz, ny, nx, 3)
    if layout == "zyx3":
        return np.array(data, dtype=dtype), spacing
    if layout != "3zyx":
        raise ValueError("layout must be '3zyx' or 'zyx3'")
    return np.ascontiguousarray(np.moveaxis(data, -1, 0), dtype=dtype), spacing


def write_rvf(filename, field, spacing=(1.0, 1.0, 1.0), layout="3zyx"):
    """Write `field` to `filename` in RVF format (``write_rvf``, rvf_io.c:11-52)."""
    with open(filename, "wb") as fh:
        fh.write(encode_rvf(field, spacing, layout))


def read_rvf(filename, dtype=np.float32, layout="3zyx"):
    """Read an RVF file (``read_rvf``, rvf_io.c:57-131): (field, spacing)."""
    with open(filename, "rb") as fh:
        return decode_rvf(fh.read(), dtype, layout)
