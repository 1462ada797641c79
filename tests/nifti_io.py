"""Independent NIfTI-1 single-file writer/reader for the host-layer tests.

Written from the NIfTI-1 header specification (348-byte header, "n+1" magic,
data at vox_offset); used to make inputs for the C++ reader and the CLIs and to
read back what the CLIs write.  Volumes are numpy arrays indexed [x, y, z].
"""
import gzip
import struct

import numpy as np

_DT = {np.dtype(np.uint8): (2, 8), np.dtype(np.int16): (4, 16), np.dtype(np.int32): (8, 32),
       np.dtype(np.float32): (16, 32), np.dtype(np.float64): (64, 64)}
_NP = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64}


def quaternion_matrix(b, c, d, pixdim, qfac=1.0):
    """Rotation*scale block of qto_xyz (NIfTI-1 method 2) in float64."""
    a = np.sqrt(max(0.0, 1.0 - (b * b + c * c + d * d)))
    R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                  [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                  [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b]])
    return R * np.array([pixdim[0], pixdim[1], pixdim[2] * qfac])[None, :]


def write_nifti(path, vol, pixdim=(1.0, 1.0, 1.0), qoffset=(0.0, 0.0, 0.0), quatern=(0.0, 0.0, 0.0),
                qform_code=1, dtype=np.float32, big_endian=False, qfac=1.0):
    vol = np.asarray(vol)
    code, bitpix = _DT[np.dtype(dtype)]
    e = ">" if big_endian else "<"
    h = bytearray(348)
    struct.pack_into(e + "i", h, 0, 348)
    dim = [3, vol.shape[0], vol.shape[1], vol.shape[2], 1, 1, 1, 1]
    struct.pack_into(e + "8h", h, 40, *dim)
    struct.pack_into(e + "hh", h, 70, code, bitpix)
    struct.pack_into(e + "8f", h, 76, qfac, pixdim[0], pixdim[1], pixdim[2], 1.0, 1.0, 1.0, 1.0)
    struct.pack_into(e + "f", h, 108, 352.0)
    struct.pack_into(e + "hh", h, 252, qform_code, 0)
    struct.pack_into(e + "6f", h, 256, *quatern, *qoffset)
    h[344:348] = b"n+1\x00"
    data = np.asarray(vol, dtype=np.dtype(dtype).newbyteorder(e)).ravel(order="F").tobytes()
    blob = bytes(h) + b"\x00\x00\x00\x00" + data
    if str(path).endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(blob)
    else:
        with open(path, "wb") as f:
            f.write(blob)


def read_nifti(path):
    """Returns (vol[x, y, z] float64, qto_xyz 4x4 float64)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rb") as f:
        blob = f.read()
    e = "<" if struct.unpack_from("<i", blob, 0)[0] == 348 else ">"
    assert struct.unpack_from(e + "i", blob, 0)[0] == 348
    assert blob[344:347] == b"n+1"
    dim = struct.unpack_from(e + "8h", blob, 40)
    code, _ = struct.unpack_from(e + "hh", blob, 70)
    pix = struct.unpack_from(e + "8f", blob, 76)
    off = int(struct.unpack_from(e + "f", blob, 108)[0])
    b, c, d, qx, qy, qz = struct.unpack_from(e + "6f", blob, 256)
    n = dim[1] * dim[2] * dim[3]
    vol = np.frombuffer(blob, dtype=np.dtype(_NP[code]).newbyteorder(e), count=n, offset=off)
    vol = vol.astype(np.float64).reshape(dim[1:4], order="F")
    q = np.eye(4)
    q[:3, :3] = quaternion_matrix(b, c, d, pix[1:4], -1.0 if pix[0] < 0 else 1.0)
    q[:3, 3] = (qx, qy, qz)
    return vol, q
