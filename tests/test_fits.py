"""Simple-FITS reader / writer of the resident path (glass_amd/fits.py), on
the CPU: the reference's own FITS files (tests/golden, copied byte for byte
by make_golden.py), the card grammar, the raw u16 layout and write/read round
trips."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from glass_amd import fits as F
from glass_amd.synthetic import encode_fits_u16be_bzero

GOLDEN = Path(__file__).resolve().parent / "golden"
LIGHT = GOLDEN / "s2_gate_314_light_000.fits"
CALIBRATED = GOLDEN / "s2_gate_314_calibrated_F000010.fits"


def test_reads_the_reference_files(golden_314):
    spec = F.simple_fits_image_spec(LIGHT)
    assert (spec.bitpix, spec.shape, spec.data_offset) == (-32, (24, 24), 2880)
    header, consumed = F.read_primary_header(LIGHT)
    assert consumed == 2880
    assert header["IMAGETYP"] == "light" and header["INSTRUME"] == "GLASS-SYNTH"
    assert header["EXPTIME"] == 60.0 and header["XBINNING"] == 1 and header["SIMPLE"] is True
    assert np.array_equal(F.read_image(LIGHT), golden_314["raw_light"][0])
    assert np.array_equal(F.read_image(CALIBRATED), golden_314["calibrated_light"][0])
    assert not F.u16_gpu_eligibility(spec)["eligible"]
    assert F.u16_gpu_eligibility(spec)["reason"] == "bitpix_not_16:-32"


def test_rewrite_reproduces_the_reference_data_bytes(tmp_path):
    header, _ = F.read_primary_header(LIGHT)
    extra = {k: v for k, v in header.items() if k not in ("SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2")}
    out = tmp_path / "light.fits"
    F.write_fits_data(out, F.read_image(LIGHT), extra, dtype=np.float32)
    ours, ref = out.read_bytes(), LIGHT.read_bytes()
    assert len(ours) == len(ref) == 5760
    assert ours[2880:] == ref[2880:]  # big-endian data and zero padding
    # the header cards carry the same keywords and values (comments aside)
    assert F.read_primary_header(out)[0] == header
    # the value columns of the fixed-format cards line up with the reference's
    for i in range(0, 2880, 80):
        card = ref[i:i + 80].decode()
        if card.startswith("END"):
            break
        key = card[:8].strip()
        mine = next(ours[j:j + 80].decode() for j in range(0, 2880, 80) if ours[j:j + 8].decode().strip() == key)
        assert mine[:30] == card[:30], key


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16, np.int32, np.uint8])
def test_write_read_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(3)
    data = (rng.normal(0, 1000, (7, 13))).astype(dtype) if np.dtype(dtype).kind == "f" else \
        rng.integers(0, 120, (7, 13)).astype(dtype)
    path = tmp_path / "x.fits"
    F.write_fits_data(path, data, {"OBJECT": "M31 'core'", "RATIO": 1.5, "FLAG": False, "N": 7}, dtype=dtype)
    assert path.stat().st_size % 2880 == 0
    header, _ = F.read_primary_header(path)
    assert header["OBJECT"] == "M31 'core'" and header["RATIO"] == 1.5 and header["FLAG"] is False
    # images materialise in float32, as the reference's fast reader does
    assert np.array_equal(F.read_image(path), data.astype(np.float32))
    raw = np.fromfile(path, dtype=np.dtype(data.dtype).newbyteorder(">"), count=data.size, offset=2880)
    assert np.array_equal(raw.reshape(data.shape), data)


def test_card_grammar():
    assert F.parse_card_value("KEY     = 'it''s' / comment with ' quote") == "it's"
    assert F.parse_card_value("KEY     =              1.5D+02 / d exponent") == 150.0
    assert F.parse_card_value("KEY     =                   -3") == -3
    assert F.parse_card_value("KEY     = 'a/b'") == "a/b"
    assert F.parse_card_value("COMMENT this is no value card") is None
    assert F.parse_card_value("KEY     =   / empty") is None
    assert F.parse_card_value("KEY     =                    T/no blank before the comment") is True
    assert F.parse_card_value("KEY     =                 .5e1") == 5.0
    assert F.parse_card_value("KEY     = 'trailing   '") == "trailing"
    assert F.parse_card_value("KEY     = '  leading'") == "  leading"
    assert F.parse_card_value("KEY     = (1.0, 2.0) / complex stays text") == "(1.0, 2.0)"
    assert F.parse_card_value("KEY     =                  +12") == 12


def test_non_image_headers_are_refused(tmp_path):
    def write(name, cards):
        text = "".join(F._format_card(k, v) for k, v in cards) + "END".ljust(80)
        path = tmp_path / name
        path.write_bytes(text.encode().ljust(2880) + bytes(2880 * 2))
        return path

    cube = write("cube.fits", [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 3), ("NAXIS1", 2), ("NAXIS2", 2),
                               ("NAXIS3", 2)])
    with pytest.raises(F.FitsUnsupported, match="2-D"):
        F.simple_fits_image_spec(cube)
    groups = write("groups.fits", [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 2),
                                   ("NAXIS2", 2), ("GROUPS", True), ("PCOUNT", 1), ("GCOUNT", 1)])
    with pytest.raises(F.FitsUnsupported, match="random-groups"):
        F.simple_fits_image_spec(groups)
    not_simple = write("ns.fits", [("SIMPLE", False), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 2), ("NAXIS2", 2)])
    with pytest.raises(F.FitsUnsupported, match="SIMPLE"):
        F.simple_fits_image_spec(not_simple)
    bad_bitpix = write("bp.fits", [("SIMPLE", True), ("BITPIX", 12), ("NAXIS", 2), ("NAXIS1", 2), ("NAXIS2", 2)])
    with pytest.raises(F.FitsUnsupported, match="BITPIX"):
        F.simple_fits_image_spec(bad_bitpix)
    # 36 records fill the first block exactly: END opens the second, the data the third
    many = [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 2), ("NAXIS2", 2)] + \
        [(f"K{i}", i) for i in range(31)]
    spec = F.simple_fits_image_spec(write("many.fits", many))
    assert spec.data_offset == 2 * 2880 and spec.shape == (2, 2)


def test_header_errors(tmp_path):
    no_end = tmp_path / "noend.fits"
    no_end.write_bytes(F.header_bytes(16, 4, 4)[:2880].replace(b"END" + b" " * 77, b" " * 80))
    with pytest.raises(F.FitsUnsupported, match="END"):
        F.read_primary_header(no_end)
    short = tmp_path / "short.fits"
    short.write_bytes(b"SIMPLE  =                    T")
    with pytest.raises(F.FitsUnsupported, match="truncated"):
        F.read_primary_header(short)
    truncated = tmp_path / "trunc.fits"
    truncated.write_bytes(F.header_bytes(-32, 100, 100) + b"\0" * 100)
    with pytest.raises(F.FitsUnsupported, match="truncated"):
        F.simple_fits_image_spec(truncated)
    # the first occurrence of a keyword wins
    dup = tmp_path / "dup.fits"
    cards = F._format_card("SIMPLE", True) + F._format_card("BITPIX", 8) + F._format_card("NAXIS", 2) + \
        F._format_card("NAXIS1", 2) + F._format_card("NAXIS2", 2) + F._format_card("OBJECT", "first") + \
        F._format_card("OBJECT", "second") + "END".ljust(80)
    dup.write_bytes(cards.encode().ljust(2880) + bytes(2880))
    assert F.read_primary_header(dup)[0]["OBJECT"] == "first"


def test_raw_u16_layout_and_eligibility(tmp_path):
    rng = np.random.default_rng(9)
    adu = rng.integers(0, 65536, (5, 11), dtype=np.uint16)
    adu[0, :4] = [0, 1, 32767, 65535]
    path = tmp_path / "raw.fits"
    F.write_u16be_bzero(path, adu, {"EXPTIME": 30.0})
    spec = F.simple_fits_image_spec(path)
    assert F.u16_gpu_eligibility(spec, (5, 11))["eligible"]
    assert F.u16_gpu_eligibility(spec, (11, 5))["reason"] == "shape_mismatch"
    buf = np.zeros(adu.size * 2 + 6, np.uint8)
    assert F.read_u16be_raw_into(spec, buf) == adu.size * 2
    raw = buf[:adu.size * 2]
    # the GPU decode: ((b0 << 8) | b1) ^ 0x8000 is the unsigned ADU
    decoded = ((raw[0::2].astype(np.uint16) << 8) | raw[1::2]) ^ np.uint16(0x8000)
    assert np.array_equal(decoded.reshape(adu.shape), adu)
    assert raw.tobytes() == bytes(encode_fits_u16be_bzero(adu))
    # the decoded image (BZERO applied) is the ADU as float32
    assert np.array_equal(F.read_image(path), adu.astype(np.float32))
    with pytest.raises(ValueError):
        F.read_u16be_raw_into(spec, np.zeros(10, np.uint8))
    f32 = tmp_path / "f32.fits"
    F.write_fits_data(f32, adu.astype(np.float32))
    with pytest.raises(F.FitsUnsupported, match="bitpix_not_16"):
        F.read_u16be_raw_into(F.simple_fits_image_spec(f32), buf)


def test_count_maps_round_to_nearest_even():
    got = F.count_map_for_write(np.array([0.5, 1.5, 2.5, 2.6, -0.5], np.float32))
    assert got.dtype == np.int16 and got.tolist() == [0, 2, 2, 3, 0]


def test_write_resident_outputs(tmp_path):
    from glass_amd.pipeline import StackResult, write_resident_outputs

    rng = np.random.default_rng(1)
    shape = (9, 12)
    res = StackResult(master=rng.normal(100, 5, shape).astype(np.float32),
                      weight=rng.uniform(0, 8, shape).astype(np.float32),
                      coverage=np.full(shape, 7.0, np.float32), low=np.full(shape, 0.5, np.float32),
                      high=np.full(shape, 1.5, np.float32), dq=np.full(shape, 0x341, np.uint32),
                      geometric=np.full(shape, 8.0, np.float32), weights=np.ones(8, np.float32))
    prof = write_resident_outputs(res, tmp_path, "H")
    assert sorted(prof["paths"]) == ["coverage", "dq", "high_rejection", "low_rejection", "master", "weight"]
    master = tmp_path / "resident_master_H.fits"
    assert np.array_equal(F.read_image(master), res.master)
    assert F.read_primary_header(master)[0]["IMAGETYP"] == "master"
    cov = F.simple_fits_image_spec(tmp_path / "resident_coverage_map_H.fits")
    assert cov.bitpix == 16 and F.read_primary_header(cov.path)[0]["DTYPE"] == "int16"
    assert np.all(F.read_image(tmp_path / "resident_low_rejection_map_H.fits") == 0.0)   # rint(0.5) = 0
    assert np.all(F.read_image(tmp_path / "resident_high_rejection_map_H.fits") == 2.0)  # rint(1.5) = 2
    assert np.all(F.read_image(tmp_path / "resident_dq_map_H.fits") == 0x341)
    assert F.read_primary_header(tmp_path / "resident_dq_map_H.fits")[0]["DQFLAGS"] == \
        "NO_DATA,WARP_EDGE,LOW_REJECTED,HIGH_REJECTED"
