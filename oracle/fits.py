"""Minimal FITS primary-HDU reader (test infrastructure only).

astropy is not installed; the reference fixtures under
runs/checkpoints/s2_gate_314_* are simple 2-D primary HDUs (BITPIX -32), so a
header-card parser plus a big-endian data read is enough.  Used only by
tests/golden/make_golden.py to turn those fixtures into .npz vectors.
"""

from __future__ import annotations

import numpy as np

_BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def read_fits(path):
    """Return (data as float32 or float64 ndarray, header dict)."""
    with open(path, "rb") as handle:
        raw = handle.read()
    header = {}
    offset = 0
    done = False
    while not done:
        block = raw[offset : offset + _BLOCK]
        if len(block) < _BLOCK:
            raise ValueError(f"truncated FITS header: {path}")
        offset += _BLOCK
        for i in range(0, _BLOCK, 80):
            card = block[i : i + 80].decode("ascii", errors="replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] != "= ":
                continue
            value = card[10:].split("/")[0].strip()
            if value.startswith("'"):
                value = value.strip("'").strip()
            elif value in ("T", "F"):
                value = value == "T"
            else:
                try:
                    value = int(value)
                except ValueError:
                    try:
                        value = float(value)
                    except ValueError:
                        pass
            header[key] = value
    bitpix = int(header["BITPIX"])
    naxis = int(header["NAXIS"])
    shape = tuple(int(header[f"NAXIS{i}"]) for i in range(naxis, 0, -1))
    count = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=_DTYPES[bitpix], count=count, offset=offset).reshape(shape)
    bzero = float(header.get("BZERO", 0.0))
    bscale = float(header.get("BSCALE", 1.0))
    if bitpix > 0:
        out = data.astype(np.float64) * bscale + bzero
        return out.astype(np.float32), header
    return data.astype(np.float32 if bitpix == -32 else np.float64), header
