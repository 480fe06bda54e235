"""Reader/writer for the SWFX named-array fixture container.

Layout (little-endian): b"SWFX", u32 version=1, then records of
u32 name_length, name, u8 dtype (0 f64, 1 i64, 2 u64, 3 u8), u64 count, data.
The C++ writer lives in oracle/fixture_io.hpp.
"""
from __future__ import annotations

import struct

import numpy as np

_DTYPES = {0: np.float64, 1: np.int64, 2: np.uint64, 3: np.uint8}
_CODES = {np.dtype(np.float64): 0, np.dtype(np.int64): 1, np.dtype(np.uint64): 2, np.dtype(np.uint8): 3}


def read(path: str) -> dict[str, np.ndarray]:
    with open(path, "rb") as handle:
        data = handle.read()
    if data[:4] != b"SWFX":
        raise ValueError(f"{path}: not an SWFX container")
    (version,) = struct.unpack_from("<I", data, 4)
    if version != 1:
        raise ValueError(f"{path}: unsupported SWFX version {version}")
    offset = 8
    out: dict[str, np.ndarray] = {}
    while offset < len(data):
        (length,) = struct.unpack_from("<I", data, offset)
        offset += 4
        name = data[offset:offset + length].decode()
        offset += length
        dtype = _DTYPES[data[offset]]
        offset += 1
        (count,) = struct.unpack_from("<Q", data, offset)
        offset += 8
        nbytes = count * np.dtype(dtype).itemsize
        out[name] = np.frombuffer(data, dtype=dtype, count=count, offset=offset).copy()
        offset += nbytes
    return out


def write(path: str, arrays: dict[str, np.ndarray]) -> None:
    with open(path, "wb") as handle:
        handle.write(b"SWFX")
        handle.write(struct.pack("<I", 1))
        for name, value in arrays.items():
            value = np.ascontiguousarray(value)
            encoded = name.encode()
            handle.write(struct.pack("<I", len(encoded)))
            handle.write(encoded)
            handle.write(struct.pack("<B", _CODES[value.dtype]))
            handle.write(struct.pack("<Q", value.size))
            handle.write(value.tobytes())
