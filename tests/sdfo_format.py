"""Independent numpy restatement of the "SDFO" 0.1.0 layout (SdfIO.cpp:60-275).

Test infrastructure: writes the exact bytes the reference's saveSdf /
saveSparseSdf produce for given fields, and parses headers, so the library's
writer and reader can be checked byte for byte.
"""
import struct

import numpy as np

VERSION = 100
FLAG_HESSIAN, FLAG_SPARSE, FLAG_CONTACT = 1, 2, 4


def header(shape, bmin, bmax, voxel, flags, reserved=b""):
    h = b"SDFO" + struct.pack("<I", VERSION) + struct.pack("<3I", *shape)
    h += struct.pack("<3d", *bmin) + struct.pack("<3d", *bmax) + struct.pack("<3d", *voxel)
    h += struct.pack("<I", flags)
    h += reserved.ljust(32, b"\0")
    assert len(h) == 128
    return h


def dense_bytes(shape, bmin, bmax, voxel, phi, n, H, witness, cand=None):
    """saveSdf (SdfIO.cpp:93-183)."""
    if cand is None:
        h = header(shape, bmin, bmax, voxel, FLAG_HESSIAN, struct.pack("<I", 1))
    else:
        h = header(shape, bmin, bmax, voxel, FLAG_HESSIAN | FLAG_CONTACT,
                   struct.pack("<II", 3, cand["stride"]))
    body = [np.asarray(phi, "<f8").tobytes(), np.asarray(n, "<f8").tobytes(),
            np.asarray(H, "<f8").tobytes(), np.asarray(witness, "<f8").tobytes()]
    if cand is not None:
        body += [np.asarray(cand["count"], "<u4").tobytes(), np.asarray(cand["phi"], "<f8").tobytes(),
                 np.asarray(cand["n"], "<f8").tobytes(), np.asarray(cand["branch"], "<i4").tobytes()]
    return h + b"".join(body)


def sparse_bytes(shape, bmin, bmax, voxel, block_dim, halo, band, bijk, phi, n, H=None, cand=None):
    """saveSparseSdf (SdfIO.cpp:185-275)."""
    contact = cand is not None and len(bijk) > 0
    flags = FLAG_SPARSE | (FLAG_HESSIAN if H is not None else 0) | (FLAG_CONTACT if contact else 0)
    res = struct.pack("<IIII", 4 if contact else 2, block_dim, halo, len(bijk)) + struct.pack("<d", band)
    if contact:
        res += struct.pack("<I", cand["stride"])
    out = [header(shape, bmin, bmax, voxel, flags, res)]
    for b in range(len(bijk)):
        out.append(struct.pack("<3iI", *[int(x) for x in bijk[b]], 0))
        out.append(np.asarray(phi[b], "<f4").tobytes())
        out.append(np.asarray(n[b], "<f4").tobytes())
        if H is not None:
            out.append(np.asarray(H[b], "<f4").tobytes())
        if contact:
            out.append(np.asarray(cand["count"][b], "u1").tobytes())
            out.append(np.asarray(cand["phi"][b], "<f4").tobytes())
            out.append(np.asarray(cand["n"][b], "<f4").tobytes())
            out.append(np.asarray(cand["branch"][b], "<i4").tobytes())
    return b"".join(out)


def parse_header(b):
    magic = b[:4]
    version, nx, ny, nz = struct.unpack("<4I", b[4:20])
    bmin = struct.unpack("<3d", b[20:44])
    bmax = struct.unpack("<3d", b[44:68])
    voxel = struct.unpack("<3d", b[68:92])
    (flags,) = struct.unpack("<I", b[92:96])
    return dict(magic=magic, version=version, shape=(nx, ny, nz), bmin=bmin, bmax=bmax,
                voxel=voxel, flags=flags, reserved=b[96:128])


def python_reader_phi(b):
    """The prototype's TrilinearSDFGrid._load (sdf_grid.py:32-64): phi block
    reshaped (nz, ny, nx) and transposed to [i, j, k]."""
    hd = parse_header(b)
    nx, ny, nz = hd["shape"]
    phi = np.frombuffer(b[128:128 + 8 * nx * ny * nz], dtype=np.float64).copy()
    return phi.reshape((nz, ny, nx)).transpose((2, 1, 0))
