"""Simple-FITS I/O for the resident path: header probe, raw u16 reads into
caller buffers (pinned host memory for the upload ring) and the direct output
writer.

The reader is written from the FITS standard (version 4.0, sections 3.1-3.3,
4.1-4.2 and 5.2) for the files the reference's bounded fast path accepts
(src/glass/io/fits_fast.py:44-165, :206-233): a primary HDU holding one 2-D
image.  The header is a sequence of 2880-byte blocks of 80-byte keyword
records, ended by the END record; the data begin at the next block boundary.
Only the first record of a keyword counts.  Values follow the fixed/free
format grammar of section 4.2: quoted strings ('' is a literal quote,
trailing blanks are not significant), logicals T/F, integers and reals (E or D
exponent); a '/' outside a string starts the comment.

The writer (src/glass/io/fits_io.py:9-16, :95-151; count maps rounded with
np.rint, resident_cuda.py:4798-4800) emits SIMPLE/BITPIX/NAXIS/NAXIS1/NAXIS2,
the caller's records, END, big-endian data and zero padding to a 2880-byte
multiple, through a temporary name and a rename.
"""

from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

BLOCK = 2880   # bytes per header / data block (standard section 3.1)
CARD = 80      # bytes per keyword record
_CARDS_PER_BLOCK = BLOCK // CARD

# BITPIX -> big-endian storage type (standard table 8).
_STORAGE = {8: np.dtype("u1"), 16: np.dtype(">i2"), 32: np.dtype(">i4"), 64: np.dtype(">i8"),
            -32: np.dtype(">f4"), -64: np.dtype(">f8")}
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


# Value field of a keyword record (bytes 11-80), standard section 4.2.
_VALUE = re.compile(
    r"""\s*(?:
        '(?P<string>(?:[^']|'')*)'
      | (?P<logical>[TF])(?=\s|/|$)
      | (?P<int>[+-]?\d+)(?=\s|/|$)
      | (?P<real>[+-]?(?:\d+\.?\d*|\.\d+)(?:[EeDd][+-]?\d+)?)(?=\s|/|$)
    )?\s*(?P<rest>.*)$""",
    re.X,
)


def parse_card_value(card: str) -> Any:
    """Value of one keyword record, or None for commentary records and
    undefined values.  A value outside the grammar is returned as its text."""
    if len(card) < 10 or card[8] != "=":
        return None
    m = _VALUE.match(card[10:])
    if m.group("string") is not None:
        return m.group("string").replace("''", "'").rstrip()
    if m.group("logical") is not None:
        return m.group("logical") == "T"
    if m.group("int") is not None:
        return int(m.group("int"))
    if m.group("real") is not None:
        return float(m.group("real").translate(str.maketrans("Dd", "Ee")))
    rest = m.group("rest")
    text = rest.split("/", 1)[0].strip() if not rest.startswith("/") else ""
    return text or None


@dataclass
class FitsHeader:
    """Keyword -> value of a primary header (first record of each keyword
    wins) and the byte offset of the data that follow it."""

    values: dict[str, Any] = field(default_factory=dict)
    data_offset: int = 0

    def number(self, key: str, default: float) -> float:
        v = self.values.get(key, default)
        return default if v is None or isinstance(v, (bool, str)) else v


def _keyword_records(handle):
    """Yield the 80-byte records of consecutive header blocks."""
    while True:
        block = handle.read(BLOCK)
        if len(block) != BLOCK:
            raise FitsUnsupported("FITS header ends inside a 2880-byte block (truncated file)"
                                  if block else "FITS header has no END record")
        for k in range(_CARDS_PER_BLOCK):
            yield block[k * CARD:(k + 1) * CARD]


def read_header(path) -> FitsHeader:
    header = FitsHeader()
    with open(path, "rb") as handle:
        for count, record in enumerate(_keyword_records(handle), start=1):
            if not record.isascii():
                raise FitsUnsupported("FITS header record holds non-ASCII bytes")
            card = record.decode("ascii")
            key = card[:8].rstrip()
            if key == "END":
                blocks = -(-count // _CARDS_PER_BLOCK)
                header.data_offset = blocks * BLOCK
                return header
            if key and key not in header.values:
                value = parse_card_value(card)
                if value is not None:
                    header.values[key] = value
    raise AssertionError("unreachable")


def read_primary_header(path) -> tuple[dict[str, Any], int]:
    """(keyword -> value, header bytes) of the primary HDU."""
    h = read_header(path)
    return h.values, h.data_offset


@dataclass(frozen=True)
class ImageLayout:
    """Where and how a primary 2-D image is stored."""

    path: Path
    shape: tuple[int, int]     # (NAXIS2, NAXIS1) = (height, width)
    bitpix: int
    data_offset: int
    bscale: float = 1.0
    bzero: float = 0.0
    blank: Any = None

    @property
    def height(self) -> int:
        return self.shape[0]

    @property
    def width(self) -> int:
        return self.shape[1]

    @property
    def dtype(self) -> np.dtype:
        return _STORAGE[self.bitpix]

    @property
    def data_bytes(self) -> int:
        return self.shape[0] * self.shape[1] * self.dtype.itemsize


def simple_fits_image_spec(path) -> ImageLayout:
    """Probe a primary-HDU 2-D image; FitsUnsupported for anything else."""
    target = Path(path)
    h = read_header(target)
    v = h.values
    naxis = v.get("NAXIS")
    if v.get("SIMPLE") is not True or naxis is None:
        raise FitsUnsupported("file does not start a conforming primary HDU (SIMPLE = T, NAXIS)")
    if naxis != 2:
        raise FitsUnsupported(f"primary HDU has NAXIS = {naxis}; only 2-D images are read")
    bitpix = v.get("BITPIX")
    if bitpix not in _STORAGE:
        raise FitsUnsupported(f"unsupported BITPIX={bitpix}")
    width, height = v.get("NAXIS1"), v.get("NAXIS2")
    if not (isinstance(width, int) and isinstance(height, int) and width > 0 and height > 0):
        raise FitsUnsupported(f"image axes NAXIS1={width}, NAXIS2={height} are not positive integers")
    if v.get("GROUPS") is True or h.number("PCOUNT", 0) != 0 or h.number("GCOUNT", 1) != 1:
        raise FitsUnsupported("random-groups records (GROUPS/PCOUNT/GCOUNT) are not images")
    layout = ImageLayout(target, (height, width), bitpix, h.data_offset,
                         float(h.number("BSCALE", 1.0) or 1.0), float(h.number("BZERO", 0.0) or 0.0),
                         v.get("BLANK"))
    if target.stat().st_size < layout.data_offset + layout.data_bytes:
        raise FitsUnsupported("file is shorter than its header says (truncated image data)")
    return layout


def u16_gpu_eligibility(spec: ImageLayout, expected_shape=None) -> dict[str, Any]:
    """Whether the raw bytes can go to the GPU undecoded (BITPIX 16, BSCALE 1,
    BZERO 32768, no BLANK, expected shape)."""
    checks = (
        (expected_shape is not None and spec.shape != tuple(expected_shape), "shape_mismatch"),
        (spec.bitpix != 16, f"bitpix_not_16:{spec.bitpix}"),
        (spec.bscale != 1.0, f"bscale_not_1:{spec.bscale:g}"),
        (spec.bzero != 32768.0, f"bzero_not_32768:{spec.bzero:g}"),
        (spec.blank is not None, "blank_present"),
    )
    reason = next((why for failed, why in checks if failed), "")
    return {"path": str(spec.path), "eligible": not reason, "reason": reason, "bitpix": spec.bitpix,
            "bscale": spec.bscale, "bzero": spec.bzero, "blank_present": spec.blank is not None,
            "shape": {"height": spec.height, "width": spec.width}, "data_offset": spec.data_offset}


def read_u16be_raw_into(spec: ImageLayout, out) -> int:
    """Read the undecoded big-endian data bytes of an eligible image into the
    writable contiguous byte buffer `out` (numpy array or memoryview, e.g. a
    pinned host tensor's .numpy()).  Returns the bytes read."""
    elig = u16_gpu_eligibility(spec)
    if not elig["eligible"]:
        raise FitsUnsupported(f"not eligible for the raw u16 path: {elig['reason']}")
    nbytes = spec.data_bytes
    view = memoryview(out).cast("B")
    if view.nbytes < nbytes:
        raise ValueError("output buffer is smaller than the image data")
    fd = os.open(spec.path, os.O_RDONLY)
    try:
        done = 0
        while done < nbytes:
            got = os.preadv(fd, [view[done:nbytes]], spec.data_offset + done)
            if got <= 0:
                raise FitsUnsupported("file is shorter than its header says (truncated image data)")
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
