"""Simple-FITS I/O for the resident path: header probe, raw u16 reads into
caller buffers (pinned host memory for the upload ring) and the direct output
writer.

Follows the reference's bounded fast path (src/glass/io/fits_fast.py:44-165,
:206-233, :379-437) and its direct primary-HDU writer
(src/glass/io/fits_io.py:9-16, :95-151; count maps rounded with np.rint,
resident_cuda.py:4798-4800):
* the primary header is read in 2880-byte blocks of 80-character cards up to
  END; the first occurrence of a keyword wins; values parse as quoted strings
  (with doubled quotes), T/F, integers or floats (D exponents accepted);
* a simple 2-D image (SIMPLE = T, NAXIS = 2, PCOUNT = 0, GCOUNT = 1) with a
  known BITPIX, data right after the header;
* the GPU decode path takes BITPIX = 16, BSCALE = 1, BZERO = 32768, no BLANK:
  its data bytes are handed over undecoded;
* written files hold SIMPLE/BITPIX/NAXIS/NAXIS1/NAXIS2, the caller's cards,
  END, big-endian data and zero padding to a 2880-byte multiple, written to a
  temporary name and renamed.
"""

from __future__ import annotations

import os
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

BLOCK = 2880
CARD = 80

_BITPIX_DTYPES = {8: "u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_WRITE_DTYPES = {
    np.dtype(np.uint8): (8, ">u1"),
    np.dtype(np.int16): (16, ">i2"),
    np.dtype(np.int32): (32, ">i4"),
    np.dtype(np.float32): (-32, ">f4"),
    np.dtype(np.float64): (-64, ">f8"),
}
_WRITE_CHUNK_BYTES = 16 * 1024 * 1024


class FitsUnsupported(ValueError):
    """The file is not a simple image this reader handles."""


@dataclass(frozen=True)
class SimpleFitsImageSpec:
    path: Path
    bitpix: int
    width: int
    height: int
    data_offset: int
    dtype: np.dtype
    bscale: float = 1.0
    bzero: float = 0.0
    blank: Any = None

    @property
    def shape(self) -> tuple[int, int]:
        return self.height, self.width


def _value_field(field: str) -> str:
    """The value part of a card (before an unquoted '/')."""
    out = []
    quoted = False
    i = 0
    while i < len(field):
        ch = field[i]
        if ch == "'":
            out.append(ch)
            if quoted and i + 1 < len(field) and field[i + 1] == "'":
                out.append("'")
                i += 2
                continue
            quoted = not quoted
        elif ch == "/" and not quoted:
            break
        else:
            out.append(ch)
        i += 1
    return "".join(out).strip()


def parse_card_value(card: str) -> Any:
    if len(card) < 10 or card[8] != "=":
        return None
    raw = _value_field(card[10:])
    if not raw:
        return None
    if raw[0] == "'":
        chars = []
        i = 1
        while i < len(raw):
            if raw[i] == "'":
                if i + 1 < len(raw) and raw[i + 1] == "'":
                    chars.append("'")
                    i += 2
                    continue
                break
            chars.append(raw[i])
            i += 1
        return "".join(chars).rstrip()
    if raw == "T":
        return True
    if raw == "F":
        return False
    text = raw.replace("D", "E").replace("d", "e")
    try:
        if any(t in text for t in (".", "E", "e")):
            return float(text)
        return int(text)
    except ValueError:
        return raw


def read_primary_header(path) -> tuple[dict[str, Any], int]:
    """(first value of each keyword, bytes of header blocks)."""
    header: dict[str, Any] = {}
    consumed = 0
    with open(path, "rb") as handle:
        while True:
            block = handle.read(BLOCK)
            if not block:
                break
            if len(block) != BLOCK:
                raise FitsUnsupported("truncated FITS header block")
            consumed += BLOCK
            for off in range(0, BLOCK, CARD):
                try:
                    card = block[off:off + CARD].decode("ascii", errors="strict")
                except UnicodeDecodeError as exc:
                    raise FitsUnsupported("FITS primary header is not ASCII") from exc
                key = card[:8].strip()
                if key == "END":
                    return header, consumed
                if key and key not in header:
                    value = parse_card_value(card)
                    if value is not None:
                        header[key] = value
    raise FitsUnsupported("FITS primary header END card not found")


def simple_fits_image_spec(path) -> SimpleFitsImageSpec:
    target = Path(path)
    header, data_offset = read_primary_header(target)
    if header.get("SIMPLE") is not True:
        raise FitsUnsupported("not a simple primary FITS image")
    if int(header.get("NAXIS", -1)) != 2:
        raise FitsUnsupported("only 2D primary images are supported")
    bitpix = int(header.get("BITPIX", 0))
    if bitpix not in _BITPIX_DTYPES:
        raise FitsUnsupported(f"unsupported BITPIX={bitpix}")
    width, height = int(header.get("NAXIS1", 0)), int(header.get("NAXIS2", 0))
    if width <= 0 or height <= 0:
        raise FitsUnsupported("invalid NAXIS1/NAXIS2 dimensions")
    if int(header.get("PCOUNT", 0) or 0) != 0 or int(header.get("GCOUNT", 1) or 1) != 1:
        raise FitsUnsupported("random-groups FITS images are not supported")
    dtype = np.dtype(_BITPIX_DTYPES[bitpix])
    if target.stat().st_size < data_offset + width * height * dtype.itemsize:
        raise FitsUnsupported("FITS image data is truncated")
    return SimpleFitsImageSpec(target, bitpix, width, height, data_offset, dtype,
                               float(header.get("BSCALE", 1.0) or 1.0), float(header.get("BZERO", 0.0) or 0.0),
                               header.get("BLANK"))


def u16_gpu_eligibility(spec: SimpleFitsImageSpec, expected_shape=None) -> dict[str, Any]:
    """Whether the raw bytes can go to the GPU undecoded (BITPIX 16, BSCALE 1,
    BZERO 32768, no BLANK, expected shape)."""
    reason = ""
    if expected_shape is not None and spec.shape != tuple(expected_shape):
        reason = "shape_mismatch"
    elif spec.bitpix != 16:
        reason = f"bitpix_not_16:{spec.bitpix}"
    elif spec.bscale != 1.0:
        reason = f"bscale_not_1:{spec.bscale:g}"
    elif spec.bzero != 32768.0:
        reason = f"bzero_not_32768:{spec.bzero:g}"
    elif spec.blank is not None:
        reason = "blank_present"
    return {"path": str(spec.path), "eligible": not reason, "reason": reason, "bitpix": spec.bitpix,
            "bscale": spec.bscale, "bzero": spec.bzero, "blank_present": spec.blank is not None,
            "shape": {"height": spec.height, "width": spec.width}, "data_offset": spec.data_offset}


def read_u16be_raw_into(spec: SimpleFitsImageSpec, out) -> int:
    """Read the undecoded big-endian data bytes of an eligible image into the
    writable contiguous byte buffer `out` (numpy array or memoryview, e.g. a
    pinned host tensor's .numpy()).  Returns the bytes read."""
    elig = u16_gpu_eligibility(spec)
    if not elig["eligible"]:
        raise FitsUnsupported(f"not eligible for the raw u16 path: {elig['reason']}")
    nbytes = spec.width * spec.height * 2
    view = memoryview(out).cast("B")
    if view.nbytes < nbytes:
        raise ValueError("output buffer is smaller than the image data")
    fd = os.open(spec.path, os.O_RDONLY)
    try:
        done = 0
        while done < nbytes:
            got = os.preadv(fd, [view[done:nbytes]], spec.data_offset + done)
            if got <= 0:
                raise FitsUnsupported("FITS image data is truncated")
            done += got
    finally:
        os.close(fd)
    return nbytes


def read_image(path, dtype=np.float32) -> np.ndarray:
    """Decoded image (BSCALE/BZERO applied in float32, BLANK -> NaN)."""
    spec = simple_fits_image_spec(path)
    raw = np.memmap(spec.path, dtype=spec.dtype, mode="r", offset=spec.data_offset, shape=spec.shape)
    data = np.asarray(raw, dtype=np.float32)
    if spec.blank is not None:
        data[np.asarray(raw == spec.blank)] = np.nan
    if spec.bscale != 1.0 or spec.bzero != 0.0:
        data *= np.float32(spec.bscale)
        data += np.float32(spec.bzero)
    return np.asarray(data, dtype=dtype)


def _format_card(key: str, value: Any) -> str:
    key = str(key).upper()[:8]
    if isinstance(value, bool):
        text = ("T" if value else "F").rjust(20)
    elif isinstance(value, (int, np.integer)):
        text = str(int(value)).rjust(20)
    elif isinstance(value, (float, np.floating)):
        v = float(value)
        text = repr(v).upper() if np.isfinite(v) else repr(v)
        if "." not in text and "E" not in text and np.isfinite(v):
            text += ".0"
        text = text.rjust(20)
    else:
        text = "'" + str(value).replace("'", "''").ljust(8) + "'"
    card = f"{key:<8}= {text}"
    if len(card) > CARD:
        raise ValueError(f"FITS card for {key} is longer than 80 characters")
    return card.ljust(CARD)


def header_bytes(bitpix: int, width: int, height: int, header: dict | None = None) -> bytes:
    cards = [_format_card("SIMPLE", True), _format_card("BITPIX", bitpix), _format_card("NAXIS", 2),
             _format_card("NAXIS1", width), _format_card("NAXIS2", height)]
    reserved = {"SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "END"}
    for key, value in (header or {}).items():
        if value is not None and len(str(key)) <= 8 and str(key).upper() not in reserved:
            cards.append(_format_card(str(key), value))
    cards.append("END".ljust(CARD))
    text = "".join(cards).encode("ascii")
    return text + b" " * ((BLOCK - len(text) % BLOCK) % BLOCK)


def write_fits_data(path, data, header: dict | None = None, dtype=np.float32) -> None:
    """Direct simple-primary writer for 2-D images of the listed dtypes."""
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    arr = np.asarray(data)
    if arr.ndim != 2:
        raise ValueError("the FITS writer supports only 2D primary images")
    dt = np.dtype(dtype)
    if dt not in _WRITE_DTYPES:
        raise ValueError(f"unsupported FITS output dtype {dt}")
    bitpix, storage = _WRITE_DTYPES[dt]
    rows = max(1, _WRITE_CHUNK_BYTES // max(1, arr.shape[1] * dt.itemsize))
    data_bytes = arr.shape[0] * arr.shape[1] * dt.itemsize
    tmp = target.with_name(target.name + ".tmp")
    with open(tmp, "wb") as handle:
        handle.write(header_bytes(bitpix, arr.shape[1], arr.shape[0], header))
        for y0 in range(0, arr.shape[0], rows):
            slab = np.ascontiguousarray(np.asarray(arr[y0:y0 + rows], dtype=dt).astype(storage, copy=False))
            handle.write(memoryview(slab).cast("B"))
        pad = (BLOCK - data_bytes % BLOCK) % BLOCK
        if pad:
            handle.write(b"\0" * pad)
    os.replace(tmp, target)


def count_map_for_write(data, dtype=np.int16) -> np.ndarray:
    """Count maps are written rounded to nearest (np.rint)."""
    return np.rint(np.asarray(data)).astype(dtype, copy=False)


def write_u16be_bzero(path, adu, header: dict | None = None) -> None:
    """A BITPIX 16 / BZERO 32768 image of unsigned 16-bit ADU (the camera raw
    layout the GPU decode path takes)."""
    a = np.asarray(adu, dtype=np.uint16)
    if a.ndim != 2:
        raise ValueError("the FITS writer supports only 2D primary images")
    stored = (a.astype(np.int32) - 32768).astype(">i2")
    hdr = {"BZERO": 32768, "BSCALE": 1}
    hdr.update(header or {})
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    with open(target, "wb") as handle:
        handle.write(header_bytes(16, a.shape[1], a.shape[0], hdr))
        handle.write(stored.tobytes())
        pad = (BLOCK - stored.nbytes % BLOCK) % BLOCK
        if pad:
            handle.write(b"\0" * pad)
