"""DaVis containers on the CPU: the oracle's decoders, corrections and
geometries against LaVision's lvpyio frames (the reference's own fixtures,
tests/golden/davis/), the container parsers of pivtools_amd.readers, and the
library's host decompressors (token stream, LZ4) against the oracle."""

import os
import struct
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ingest_oracle as I  # noqa: E402
from pivtools_amd import readers as R  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden", "davis")


def _unpack(npz, case, dest):
    """Write one recording of an archive (keys '{case}__file__{path}', '/'
    written as '~') under dest; return the .set path."""
    prefix = f"{case}__file__"
    set_path = None
    for k in npz.files:
        if not k.startswith(prefix):
            continue
        rel = k[len(prefix):].replace("~", "/")
        p = os.path.join(dest, rel)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        npz[k].tofile(p)
        if p.endswith(".set"):
            set_path = p
    return set_path


def _transformer_cases():
    z = np.load(os.path.join(GOLD, "transformers_lvpyio.npz"))
    return sorted({k.split("__")[0] for k in z.files})


def oracle_read(rec, stream, image):
    """The oracle's frame: decode, corrections in LaVision's order, geometry."""
    s = rec.streams[stream]
    h, w = s.stored_shape
    img = I.decode_davis(rec.payload(stream, image), s.decoder, w, h)
    corr = {t: R._correction_image(s.correction_paths[t], s.stored_shape) for t in s.correction_paths}
    img = I.correct_davis(img, corr.get("intensity-correction-1"), corr.get("dark-image-subtraction"),
                          corr.get("background-subtraction"))
    geo = [t for t in s.transformers if t in R.GEOMETRIES]
    return I.geometry_davis(img, geo[0] if geo else None)


@pytest.mark.parametrize("case", _transformer_cases())
def test_set_transformers_against_lvpyio(case, tmp_path):
    """Every geometry, correction and order of the reference's lvpyio oracle:
    the parser finds the stream, encoding, shape and transformers, and the
    oracle's frames equal lvpyio's bit for bit."""
    z = np.load(os.path.join(GOLD, "transformers_lvpyio.npz"))
    rec = R.SetRecording(_unpack(z, case, str(tmp_path)))
    ref = z[f"{case}__lvpyio"]
    assert len(rec.streams) == 1 and rec.n_images == ref.shape[0]
    assert rec.streams[0].shape == ref.shape[1:]
    for b in range(rec.n_images):
        np.testing.assert_array_equal(oracle_read(rec, 0, b), ref[b])


def test_geometry_index_maps_by_hand():
    """The 3 x 5 frame of the reference test, each geometry written by hand
    (where the sensor's top-left pixel lands), not from numpy's rot90."""
    x = np.arange(15, dtype=np.uint16).reshape(3, 5)
    assert I.geometry_davis(x, "rotate-cw-90")[0].tolist() == [10, 5, 0]
    assert I.geometry_davis(x, "rotate-ccw-90")[0].tolist() == [4, 9, 14]
    assert I.geometry_davis(x, "rotate-180")[0].tolist() == [14, 13, 12, 11, 10]
    assert I.geometry_davis(x, "vertical-mirror")[0].tolist() == [10, 11, 12, 13, 14]
    assert I.geometry_davis(x, "horinzontal-mirror")[0].tolist() == [4, 3, 2, 1, 0]
    assert I.geometry_davis(x, "rotate-cw-90-vertical-mirror")[0].tolist() == [14, 9, 4]
    assert I.geometry_davis(x, "rotate-ccw-90-vertical-mirror")[0].tolist() == [0, 5, 10]


@pytest.mark.parametrize("enc", ["mono-10pmsb", "mono-12pmsb", "mono-12packed"])
def test_packed_encodings_against_lvpyio(enc):
    z = np.load(os.path.join(GOLD, "packed_lvpyio.npz"))
    key = enc.replace("-", "_")
    np.testing.assert_array_equal(I.decode_davis(z[key + "__payload"], enc, 64, 48), z[key + "__lvpyio"])


def test_packed_encodings_hand_computed_bytes():
    """The reference tests' worked groups."""
    assert I.decode_davis(bytes([0xA5, 0x3C, 0x96, 0x0F, 0xE1]), "mono-10pmsb", 4, 1).tolist() == [[660, 969, 387, 993]]
    two = bytes([0x34, 0x12, 0xAB, 0xFF, 0x0F, 0x80])
    # p0 = b0 << 4 | b1 >> 4, p1 = (b1 & 0x0F) << 8 | b2
    assert I.decode_davis(two, "mono-12pmsb", 2, 2).tolist() == [[0x341, 0x2AB], [0xFF0, 0xF80]]
    # p0 = b0 << 4 | (b1 & 0x0F), p1 = b2 << 4 | (b1 >> 4)
    assert I.decode_davis(two, "mono-12packed", 2, 2).tolist() == [[0x342, 0xAB1], [0xFFF, 0x800]]
    assert I.decode_davis(bytes([0x34, 0x12, 0xAB]), "mono-12p", 2, 1).tolist() == [[564, 2737]]


@pytest.mark.parametrize("enc,bits,msb", [("mono-10p", 10, False), ("mono-10pmsb", 10, True),
                                          ("mono-12p", 12, False), ("mono-12pmsb", 12, True)])
def test_bit_stream_round_trip(enc, bits, msb):
    """The PFNC packings against an independent bit-by-bit packer."""
    v = np.random.default_rng(bits + msb).integers(0, 1 << bits, 8 * 12).astype(np.uint16)
    np.testing.assert_array_equal(I.decode_davis(I.pack_bits(v, bits, msb), enc, 12, 8).reshape(-1), v)


def test_planted_mistakes_fail():
    """A swapped nibble order or a geometry sign slip is caught by the lvpyio
    frames (the pins discriminate)."""
    z = np.load(os.path.join(GOLD, "packed_lvpyio.npz"))
    good = I.decode_davis(z["mono_12packed__payload"], "mono-12packed", 64, 48)
    wrong = I.decode_davis(z["mono_12packed__payload"], "mono-12pmsb", 64, 48)
    assert np.array_equal(good, z["mono_12packed__lvpyio"]) and not np.array_equal(wrong, good)
    x = np.arange(20, dtype=np.uint16).reshape(4, 5)
    assert not np.array_equal(I.geometry_davis(x, "rotate-cw-90"), I.geometry_davis(x, "rotate-ccw-90"))


def _core_or_skip():
    from pivtools_amd import _native
    try:
        return _native.core()
    except OSError as exc:       # the HIP runtime cannot be loaded here
        pytest.skip(f"libpivtools_amd.so not loadable: {exc}")


def test_im7_fixtures_against_lvpyio():
    """.im7 pack 1 (token stream, running value carried across frames; a high
    nibble 0x8 is a delta) and version-1 pack 20 (one LZ4 block per frame),
    read by the library's host decoders, equal lvpyio's frames."""
    _core_or_skip()
    z = np.load(os.path.join(GOLD, "im7_lvpyio.npz"))
    for name in ("carry_3frames", "high_nibble_8", "three_frames"):
        import tempfile
        with tempfile.NamedTemporaryFile(suffix=".im7", delete=False) as f:
            f.write(z[f"{name}__im7"].tobytes())
        try:
            frames = R.read_im7_frames(f.name)
        finally:
            os.unlink(f.name)
        np.testing.assert_array_equal(frames, z[f"{name}__lvpyio"])


def test_set_imx_against_lvpyio(tmp_path):
    """An IMX-compressed mono-12p stream (107 x 5312): the library's token
    decoder gives the payload words, the oracle decodes them, and the frame
    equals lvpyio's."""
    _core_or_skip()
    z = np.load(os.path.join(GOLD, "set_imx_lvpyio.npz"))
    for k in z.files:
        if k == "lvpyio":
            continue
        p = os.path.join(str(tmp_path), k.replace("~", "/"))
        os.makedirs(os.path.dirname(p), exist_ok=True)
        z[k].tofile(p)
    rec = R.SetRecording(os.path.join(str(tmp_path), "fixture.set"))
    s = rec.streams[0]
    assert s.codec == "IMX" and s.decoder == "mono-12p"
    img = oracle_read(rec, 0, 0)
    np.testing.assert_array_equal(img, z["lvpyio"][: img.shape[0], : img.shape[1]])


def test_host_token_decoder_against_oracle():
    """The library's token decoder against the oracle's on random streams of
    every token kind, and its failure codes."""
    _core_or_skip()
    rng = np.random.default_rng(5)
    for trial in range(20):
        toks = bytearray()
        val = 0                      # the decoder's running value starts at 0
        n = 0
        while n < 300:
            kind = rng.integers(0, 4)
            if kind == 0:
                d = int(rng.integers(-100, 100))
                if d in (0x7F, -128, -127, 127):
                    continue
                if not 0 <= val + d <= 65535:
                    continue
                toks.append(d & 0xFF)
                val += d
                n += 1
            elif kind == 1:
                val = int(rng.integers(0, 65536))
                toks += bytes([0x80, val & 0xFF, val >> 8])
                n += 1
            elif kind == 2:
                cnt = int(rng.integers(0, 6))
                toks += bytes([0x7F, cnt])
                for _ in range(cnt):
                    val = int(rng.integers(0, 65536))
                    toks += bytes([val & 0xFF, val >> 8])
                n += cnt
            else:
                toks.append(0x81)
                k = int(rng.integers(1, 5))
                for q in range(k):
                    hi = int(rng.integers(-8, 8))
                    if not 0 <= val + hi <= 65535:
                        hi = 0
                    if q == k - 1:            # a low nibble 0x8 ends the run
                        toks.append(((hi & 0xF) << 4) | 8)
                        val += hi
                        n += 1
                    else:
                        lo = int(rng.integers(-7, 8))
                        if not 0 <= val + hi + lo <= 65535:
                            lo = 0
                        toks.append(((hi & 0xF) << 4) | (lo & 0xF))
                        val += hi + lo
                        n += 2
        want, off, prev = I.davis_tokens(toks, n, 0, 0)
        got, goff, gprev = R.davis_tokens(np.frombuffer(bytes(toks), np.uint8), n, 0, 0)
        np.testing.assert_array_equal(got, want)
        assert (goff, gprev) == (off, prev)
    with pytest.raises(IOError, match="ended inside a token"):
        R.davis_tokens(bytes([0x80, 0x01]), 1)
    with pytest.raises(IOError, match="runs past"):
        R.davis_tokens(bytes([0x7F, 3, 0, 0, 0, 0, 0, 0]), 2)
    with pytest.raises(IOError, match="range"):
        R.davis_tokens(bytes([0xFF]), 1)


def test_host_lz4_against_oracle_and_hand_blocks():
    """LZ4 blocks: a hand-built block (literals, an overlapping match, 255-run
    length extensions) and random repetitive data compressed by a greedy
    encoder written here, through the library and the oracle."""
    _core_or_skip()
    # "abcabcabcabc" + 300 x "z": literals 'abc', match offset 3 length 9,
    # then literal 'z' and a match offset 1 length 299 (4 + 15 + 255 + 25)
    blk = bytes([0x35]) + b"abc" + struct.pack("<H", 3) + bytes([0x1F]) + b"z" + struct.pack("<H", 1) + \
        bytes([255, 25]) + bytes([0x10]) + b"!"
    want = b"abcabcabcabc" + b"z" * 300 + b"!"
    assert I.lz4_block(blk, len(want)) == want
    assert R.lz4_block(blk, len(want)) == want

    def greedy(data: bytes) -> bytes:
        out = bytearray()
        i, lit0 = 0, 0
        last = {}
        while i + 4 <= len(data) - 5:
            key = data[i:i + 4]
            j = last.get(key)
            last[key] = i
            if j is None or i - j > 65535:
                i += 1
                continue
            ml = 4
            while i + ml < len(data) - 5 and data[j + ml] == data[i + ml]:
                ml += 1
            lit = data[lit0:i]

            def ext(n):
                return bytes([255] * (n // 255) + [n % 255])
            out.append((min(len(lit), 15) << 4) | min(ml - 4, 15))
            if len(lit) >= 15:
                out += ext(len(lit) - 15)
            out += lit + struct.pack("<H", i - j)
            if ml - 4 >= 15:
                out += ext(ml - 19)
            i += ml
            lit0 = i
        lit = data[lit0:]
        out.append(min(len(lit), 15) << 4)
        if len(lit) >= 15:
            out += bytes([255] * ((len(lit) - 15) // 255) + [(len(lit) - 15) % 255])
        return bytes(out + lit)

    rng = np.random.default_rng(3)
    for _ in range(5):
        data = bytes(np.repeat(rng.integers(0, 6, 900).astype(np.uint8), rng.integers(1, 9, 900)))
        blk = greedy(data)
        assert I.lz4_block(blk, len(data)) == data
        assert R.lz4_block(blk, len(data)) == data
    with pytest.raises(IOError):
        R.lz4_block(bytes([0x04]) + struct.pack("<H", 9), 8)     # match before the output start


def test_set_refusals(tmp_path):
    """The parser refuses what the reference refuses, by name: an index word
    it does not read, an unknown codec, an unknown transformer."""
    z = np.load(os.path.join(GOLD, "transformers_lvpyio.npz"))
    set_path = _unpack(z, "g_raw_16_bit_none", str(tmp_path))
    idx = os.path.join(os.path.dirname(set_path), "rec", "Frame0-0.ims")
    data = bytearray(open(idx, "rb").read())
    good = bytes(data)
    data[5 * 4:5 * 4 + 4] = struct.pack("<i", 8)
    open(idx, "wb").write(bytes(data))
    with pytest.raises(ValueError, match="bits per decoded pixel"):
        R.SetRecording(set_path)
    data = bytearray(good)
    data[24:28] = b"zlib"
    open(idx, "wb").write(bytes(data))
    with pytest.raises(ValueError, match="compression"):
        R.SetRecording(set_path)
    open(idx, "wb").write(good)
    man = os.path.join(os.path.dirname(set_path), "rec", "StreamSet.xml")
    txt = open(man).read().replace("</StreamCfg>", '<Transformer ID="mystery" FilePrefix="T9"><ContentPurpose '
                                   'StartFrame="0" EndFrame="0"/></Transformer></StreamCfg>')
    open(man, "w").write(txt)
    with pytest.raises(ValueError, match="unknown transformer"):
        R.SetRecording(set_path)


def test_zlib_im7_pack2(tmp_path):
    """.im7 pack 2, one zlib block per row behind an int32 size, against the
    pixels written."""
    rng = np.random.default_rng(9)
    px = rng.integers(0, 4096, (2, 6, 10)).astype("<u2")
    hdr = struct.pack("<hhhhiiiihhh", 0, 2, -4, 0, 10, 6, 1, 2, 0, 0, 0).ljust(256, b"\0")
    body = b"".join(struct.pack("<i", len(c)) + c for c in (zlib.compress(r.tobytes()) for r in px.reshape(-1, 10)))
    p = tmp_path / "z.im7"
    p.write_bytes(hdr + body)
    np.testing.assert_array_equal(R.read_im7_frames(p), px)
