"""Problem files (.gpp): round trips through the library's writer and
readers, an independent Python parser of the documented layout and checksum
(so the format is pinned by something other than the code that writes it),
corruption handling, and the reference-produced fixture
tests/golden/ref_near_pole_16x3.gpp checked against the oracle.  CPU only:
gpp_file_write / gpp_file_info / gpp_file_read touch no device."""
import os
import struct

import numpy as np
import pytest

import gpp_inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_near_pole_16x3.gpp")
INPUTS = ("inv_igp_index", "wx_array", "vcoul", "aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array")
OUTS = ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")
MASK64 = (1 << 64) - 1


def parse_independently(path):
    """Reads a .gpp file following the layout stated in include/gpp.h."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"GPPSIG01"
    version, flags = struct.unpack_from("<II", raw, 8)
    nb, nv, nc, ng, nf, zero = struct.unpack_from("<6i", raw, 16)
    occ, = struct.unpack_from("<d", raw, 40)
    payload, checksum = struct.unpack_from("<QQ", raw, 48)
    assert version == 1 and zero == 0 and raw[64:128] == bytes(64)
    body = raw[128:]
    assert len(body) == payload and payload % 64 == 0
    words = np.frombuffer(body, "<u8").tolist()
    lanes = []
    for k in range(4):
        h = 0xcbf29ce484222325
        for w in words[k::4]:
            h = ((h ^ w) * 0x100000001b3) & MASK64
        lanes.append(h)
    h = 0xcbf29ce484222325
    for x in lanes:
        h = ((h ^ x) * 0x100000001b3) & MASK64
    assert h == checksum
    sections = [("inv_igp_index", "<i4", ng), ("wx_array", "<f8", nf * (nb if flags & 1 else 1)),
                ("vcoul", "<f8", nc), ("aqsmtemp", "<c16", nb * nc), ("aqsntemp", "<c16", nb * nc),
                ("I_eps_array", "<c16", ng * nc), ("wtilde_array", "<c16", ng * nc)]
    if flags & 2:
        sections += [("achtemp", "<c16", nf), ("asxtemp", "<c16", nf), ("acht_n1_loc", "<c16", nb),
                     ("achstemp", "<c16", 1)]
    out, off = {}, 0
    for name, dt, n in sections:
        nbytes = np.dtype(dt).itemsize * n
        out[name] = np.frombuffer(body, dt, n, off)
        pad_end = (off + nbytes + 63) // 64 * 64
        assert body[off + nbytes:pad_end] == bytes(pad_end - off - nbytes)
        off = pad_end
    assert off == payload
    return (nb, nv, nc, ng, nf, occ, flags), out


def _problem(pygpp, a, nv):
    nb, nc = a["aqsmtemp"].shape
    return pygpp.Problem(nb, nv, nc, a["I_eps_array"].shape[0], np.shape(a["wx_array"])[-1], 0.75)


def _same_inputs(a, b):
    for k in INPUTS:
        np.testing.assert_array_equal(np.asarray(a[k]).ravel(), np.asarray(b[k]).ravel(), err_msg=k)


def test_round_trip_shared_wx(gpplib, tmp_path):
    import pygpp
    a = gpp_inputs.seeded_inputs(7, 130, 9, nfreq=4, seed=61)
    p = _problem(pygpp, a, 2)
    f = str(tmp_path / "a.gpp")
    pygpp.write_problem(f, p, a)
    info = pygpp.file_info(f)
    assert info["problem"] == p and info["flags"] == 0
    assert os.path.getsize(f) == 128 + info["payload_bytes"]
    q, b, exp = pygpp.read_problem(f)
    assert q == p and exp is None
    _same_inputs(a, b)
    sizes, c = parse_independently(f)
    assert sizes == (7, 2, 130, 9, 4, 0.75, 0)
    _same_inputs(a, c)


def test_round_trip_band_wx_and_expected(gpplib, oracle, tmp_path):
    import pygpp
    a = gpp_inputs.seeded_inputs(6, 77, 5, nfreq=3, seed=62)
    a["wx_array"] = gpp_inputs.per_band_wx(6, 3, seed=62, levels=2)
    ref = oracle.sigma(a, nvband=1)
    p = _problem(pygpp, a, 1)
    f = str(tmp_path / "b.gpp")
    pygpp.write_problem(f, p, a, expected=dict(ref, acht_n1_loc=None))
    info = pygpp.file_info(f)
    assert info["flags"] == pygpp.GPP_FILE_WX_PER_BAND | pygpp.GPP_FILE_HAS_EXPECTED
    q, b, exp = pygpp.read_problem(f)
    assert b["wx_array"].shape == (6, 3)
    _same_inputs(a, b)
    np.testing.assert_array_equal(exp["achtemp"], ref["achtemp"])
    np.testing.assert_array_equal(exp["asxtemp"], ref["asxtemp"])
    assert exp["achstemp"] == ref["achstemp"]
    assert np.all(np.isnan(exp["acht_n1_loc"]))  # not given to the writer
    _, c = parse_independently(f)
    _same_inputs(a, c)
    np.testing.assert_array_equal(c["achtemp"], ref["achtemp"])


def _status_of(fn):
    import pygpp
    with pytest.raises(pygpp.GppError) as e:
        fn()
    return e.value.status


def test_corruption_detected(gpplib, tmp_path):
    import pygpp
    a = gpp_inputs.seeded_inputs(4, 64, 3, seed=63)
    p = _problem(pygpp, a, 1)
    good = str(tmp_path / "good.gpp")
    pygpp.write_problem(good, p, a)
    raw = bytearray(open(good, "rb").read())

    def variant(name, data):
        path = str(tmp_path / name)
        open(path, "wb").write(bytes(data))
        return path

    flipped = bytearray(raw)
    flipped[128 + 700] ^= 0x10
    assert _status_of(lambda: pygpp.read_problem(variant("flip.gpp", flipped))) == pygpp.GPP_ERR_IO
    assert _status_of(lambda: pygpp.read_problem(variant("trunc.gpp", raw[:-64]))) == pygpp.GPP_ERR_IO
    assert _status_of(lambda: pygpp.file_info(variant("short.gpp", raw[:100]))) == pygpp.GPP_ERR_IO
    magic = bytearray(raw)
    magic[0:8] = b"NOTAGPP!"
    assert _status_of(lambda: pygpp.file_info(variant("magic.gpp", magic))) == pygpp.GPP_ERR_IO
    sizes = bytearray(raw)
    struct.pack_into("<i", sizes, 24, -5)  # ncouls
    assert _status_of(lambda: pygpp.file_info(variant("sizes.gpp", sizes))) == pygpp.GPP_ERR_IO
    ver = bytearray(raw)
    struct.pack_into("<I", ver, 8, 2)
    assert _status_of(lambda: pygpp.file_info(variant("ver.gpp", ver))) == pygpp.GPP_ERR_IO
    assert _status_of(lambda: pygpp.file_info(str(tmp_path / "missing.gpp"))) == pygpp.GPP_ERR_IO
    # the intact file still reads
    pygpp.read_problem(good)


def test_writer_rejects_bad_inputs(gpplib, tmp_path):
    import pygpp
    a = gpp_inputs.seeded_inputs(3, 40, 4, seed=64)
    a["inv_igp_index"] = np.array([1, -2, 3, 4], np.int32)
    p = _problem(pygpp, a, 1)
    assert _status_of(lambda: pygpp.write_problem(str(tmp_path / "x.gpp"), p, a)) == pygpp.GPP_ERR_BAD_INDEX
    a = gpp_inputs.seeded_inputs(3, 40, 4, seed=64)
    bad = pygpp.Problem(3, 4, 40, 4, 3)  # nvband > number_bands
    assert _status_of(lambda: pygpp.write_problem(str(tmp_path / "y.gpp"), bad, a)) == pygpp.GPP_ERR_INVALID_ARGUMENT
    assert _status_of(lambda: pygpp.write_problem(str(tmp_path / "no_dir" / "z.gpp"), p, a)) == pygpp.GPP_ERR_IO


def test_reference_fixture_file(gpplib, oracle):
    """The fixture's expected outputs came from the compiled reference; the
    oracle reproduces them bit for bit from the file's inputs, and the
    independent parser agrees with the library's reader."""
    import pygpp
    p, a, exp = pygpp.read_problem(GOLDEN)
    assert (p.number_bands, p.nvband, p.ncouls, p.ngpown, p.nfreq) == (16, 3, 96, 12, 3)
    o = oracle.sigma(a, nvband=p.nvband)
    for k in OUTS:
        np.testing.assert_array_equal(np.atleast_1d(o[k]), np.atleast_1d(exp[k]), err_msg=k)
    _, c = parse_independently(GOLDEN)
    _same_inputs(a, c)
