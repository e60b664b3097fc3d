"""The image loader behind load_image_texture (src/utils.cc:236-240; the
reference decodes with the vendored stb_image): baseline JPEG and PPM.

Pinned three ways: against stb_image itself (the reference compiled from its
sources, oracle/_ref) on the reference's own earth map; against an
independent decoder (PIL / libjpeg) on synthetic JPEGs covering every code
path (4:4:4, 4:2:2, 4:2:0, grayscale, restart intervals, multi-scan); and on
closed cases (a flat image decodes to its colour; PPM round trips).  The
inverse DCT and colour conversion are this build's own (double precision,
rounded), so values may differ from a given decoder by a unit or two."""
import ctypes as C
import io
import os

import numpy as np
import pytest

EARTH = "/root/reference/resources/earthmap.jpg"


def load(tpt, path):
    L = tpt.lib()
    L.tpt_image_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.tpt_image_free.argtypes = [C.c_void_p]
    p, w, h = C.c_void_p(), C.c_int(), C.c_int()
    rc = L.tpt_image_load(str(path).encode(), C.byref(p), C.byref(w), C.byref(h))
    if rc != 0:
        raise tpt.TptError(L.tpt_last_error().decode())
    img = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(h.value, w.value, 3)).copy()
    L.tpt_image_free(p)
    return img


@pytest.mark.skipif(not os.path.exists(EARTH), reason="the reference tree is not here")
def test_earthmap_matches_stb_image(tpt, reference):
    mine = load(tpt, EARTH)
    L = reference.L
    L.ref_load_image.argtypes = [C.c_char_p, C.c_void_p, C.c_long, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]
    buf = np.zeros(1024 * 512 * 4, np.uint8)
    w, h, c = C.c_int(), C.c_int(), C.c_int()
    assert L.ref_load_image(EARTH.encode(), buf.ctypes.data, buf.size, C.byref(w), C.byref(h), C.byref(c)) == 0
    assert (w.value, h.value, c.value) == (1024, 512, 3) and mine.shape == (512, 1024, 3)
    ref = buf[:w.value * h.value * 3].reshape(h.value, w.value, 3)
    d = np.abs(mine.astype(int) - ref.astype(int))
    print("earth: max", d.max(), "mean", d.mean(), "exact", (d == 0).mean())
    assert d.max() <= 4 and d.mean() < 0.05 and (d == 0).mean() > 0.95


def smooth_image(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w] / max(h, w)
    img = np.zeros((h, w, 3))
    for ch in range(3):
        a, b, c = rng.uniform(1, 6, 3)
        img[..., ch] = 127.5 + 100 * np.sin(a * x + b * y + c) * np.cos(b * x - a * y)
    img += rng.normal(0, 6, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("mode,sub,extra", [("RGB", 0, {}), ("RGB", 1, {}), ("RGB", 2, {}), ("L", 0, {}),
                                            ("RGB", 0, {"restart_marker_blocks": 3}),
                                            ("RGB", 2, {"restart_marker_rows": 1}),
                                            ("RGB", 0, {"optimize": True})])
def test_synthetic_jpegs_match_libjpeg(tpt, tmp_path, mode, sub, extra):
    Image = pytest.importorskip("PIL.Image")
    src = smooth_image(83, 117, 5 + sub)  # sizes that are not multiples of the MCU
    im = Image.fromarray(src if mode == "RGB" else src[..., 0], mode)
    path = tmp_path / "t.jpg"
    try:
        im.save(path, "JPEG", quality=92, subsampling=sub, **extra)
    except (TypeError, ValueError, OSError) as e:
        pytest.skip(f"encoder option not available: {e}")
    mine = load(tpt, path)
    ref = np.asarray(Image.open(path).convert("RGB"))
    d = np.abs(mine.astype(int) - ref.astype(int))
    print(mode, sub, extra, "max", d.max(), "mean", d.mean())
    if sub == 0:
        assert d.max() <= 3 and d.mean() < 0.3
    else:
        # chroma replicated here, interpolated ("fancy upsampling") by libjpeg
        assert d.mean() < 3.0 and np.percentile(d, 99) <= 16
    # close to the source image as well
    assert np.abs(mine.astype(int) - (np.repeat(src[..., :1], 3, 2) if mode == "L" else src).astype(int)).mean() < 8


def test_flat_image_decodes_exactly(tpt, tmp_path):
    Image = pytest.importorskip("PIL.Image")
    for colour in [(0, 0, 0), (255, 255, 255), (128, 128, 128), (200, 30, 90)]:
        path = tmp_path / "flat.jpg"
        Image.new("RGB", (40, 24), colour).save(path, "JPEG", quality=100, subsampling=0)
        mine = load(tpt, path)
        assert mine.shape == (24, 40, 3)
        assert np.abs(mine.astype(int) - np.array(colour)).max() <= 1, colour


def test_ppm_round_trip(tpt, tmp_path):
    img = smooth_image(7, 5, 1)
    p6 = tmp_path / "a.ppm"
    p6.write_bytes(b"P6\n# comment\n5 7\n255\n" + img.tobytes())
    assert np.array_equal(load(tpt, p6), img)
    p3 = tmp_path / "b.ppm"
    p3.write_text("P3\n5 7 255\n" + " ".join(str(v) for v in img.ravel()) + "\n")
    assert np.array_equal(load(tpt, p3), img)


def test_bad_files_are_refused(tpt, tmp_path):
    bad = tmp_path / "x.jpg"
    bad.write_bytes(b"\xff\xd8\xff\xc2\x00\x0b\x08\x00\x10\x00\x10\x01\x01\x11\x00")  # progressive SOF2
    with pytest.raises(tpt.TptError, match="sequential"):
        load(tpt, bad)
    with pytest.raises(tpt.TptError, match="cannot open"):
        load(tpt, tmp_path / "missing.jpg")
    junk = tmp_path / "j.bin"
    junk.write_bytes(b"hello")
    with pytest.raises(tpt.TptError, match="not a JPEG"):
        load(tpt, junk)


@pytest.mark.parametrize("header", [b"P6 6148914691236517206 1 255\n", b"P6 1 6148914691236517206 255\n",
                                    b"P6 70000 1 255\n", b"P6 20000 20000 255\n",
                                    b"P3 99999999999999999999999 2 255\n"])
def test_oversized_ppm_headers_are_refused(tpt, tmp_path, header):
    """Sizes whose pixel count would wrap, or exceed the loader's bound
    (65535 per side, 2^28 pixels), are rejected while parsing the header."""
    p = tmp_path / "big.ppm"
    p.write_bytes(header + b"\x00" * 64)
    with pytest.raises(tpt.TptError, match="PPM"):
        load(tpt, p)


def jpeg_segments(data):
    """(marker, payload offset, payload length) of every header segment before
    the first scan's entropy-coded data."""
    out, pos = [], 2
    while pos + 4 <= len(data):
        assert data[pos] == 0xFF
        m = data[pos + 1]
        n = (data[pos + 2] << 8) | data[pos + 3]
        out.append((m, pos + 4, n - 2))
        if m == 0xDA:
            break
        pos += 2 + n
    return out


def crafted_jpeg(tmp_path, dc_value=None, ac_value=None):
    """A valid baseline JPEG whose Huffman tables then map every code of the
    DC (or AC) tables to an out-of-range magnitude category."""
    Image = pytest.importorskip("PIL.Image")
    path = tmp_path / "c.jpg"
    Image.fromarray(smooth_image(16, 16, 3)).save(path, "JPEG", quality=90, subsampling=0)
    data = bytearray(path.read_bytes())
    for m, off, n in jpeg_segments(data):
        if m != 0xC4:
            continue
        p = off
        while p < off + n:
            tc = data[p] >> 4
            counts = data[p + 1:p + 17]
            total = sum(counts)
            vals = range(p + 17, p + 17 + total)
            for v in vals:
                if tc == 0 and dc_value is not None:
                    data[v] = dc_value
                if tc == 1 and ac_value is not None and data[v] != 0x00:
                    data[v] = ac_value
            p += 17 + total
    out = tmp_path / "crafted.jpg"
    out.write_bytes(bytes(data))
    return out


@pytest.mark.parametrize("dc", [12, 31, 40, 255])
def test_crafted_dc_category_is_refused(tpt, tmp_path, dc):
    with pytest.raises(tpt.TptError, match="DC magnitude"):
        load(tpt, crafted_jpeg(tmp_path, dc_value=dc))


def test_crafted_ac_category_is_refused(tpt, tmp_path):
    with pytest.raises(tpt.TptError, match="AC magnitude"):
        load(tpt, crafted_jpeg(tmp_path, ac_value=0x0F))


def test_oversized_jpeg_frame_is_refused(tpt, tmp_path):
    Image = pytest.importorskip("PIL.Image")
    path = tmp_path / "f.jpg"
    Image.fromarray(smooth_image(8, 8, 4)).save(path, "JPEG", quality=90, subsampling=0)
    data = bytearray(path.read_bytes())
    for m, off, n in jpeg_segments(data):
        if m == 0xC0:
            data[off + 1:off + 5] = b"\xff\xff\xff\xff"  # 65535 x 65535
    bad = tmp_path / "huge.jpg"
    bad.write_bytes(bytes(data))
    with pytest.raises(tpt.TptError, match="too large"):
        load(tpt, bad)


def test_mutated_files_fail_cleanly(tpt, tmp_path):
    """Seeded byte mutations of a small JPEG and PPM: each either decodes to an
    image of the size it reports or is refused with an error; none crashes."""
    Image = pytest.importorskip("PIL.Image")
    src = tmp_path / "s.jpg"
    Image.fromarray(smooth_image(24, 40, 9)).save(src, "JPEG", quality=85, subsampling=2, restart_marker_blocks=2)
    base = [src.read_bytes(), b"P6\n5 7\n255\n" + smooth_image(7, 5, 2).tobytes()]
    rng = np.random.default_rng(77)
    decoded = refused = 0
    for it in range(600):
        data = bytearray(base[it % 2])
        for _ in range(int(rng.integers(1, 6))):
            k = int(rng.integers(0, len(data)))
            op = int(rng.integers(0, 3))
            if op == 0:
                data[k] = int(rng.integers(0, 256))
            elif op == 1:
                del data[k:]
                if len(data) < 2:
                    data = bytearray(base[it % 2][:2])
            else:
                data.insert(k, int(rng.integers(0, 256)))
        p = tmp_path / f"m{it % 2}.bin"
        p.write_bytes(bytes(data))
        try:
            img = load(tpt, p)
            assert img.ndim == 3 and img.shape[2] == 3
            decoded += 1
        except tpt.TptError:
            refused += 1
    assert decoded + refused == 600 and refused > 0
