"""Post-processing on the CPU checker (ToneMap.cpp:5-73, PostProcessing.cpp:16-45, Texture.cpp:19-81)
against closed forms, and the library's PNG writer decoded independently (zlib, crc32)."""
import struct
import zlib

import numpy as np
import pytest

import oracle
from lambda_amd import abi, gpu


def _img(vals):
    a = np.asarray(vals, dtype=np.float32)
    out = np.ones(a.shape[:-1] + (4,), dtype=np.float32)
    out[..., :3] = a
    return out


def test_aces_keeps_grey_neutral_when_corrected_and_tints_it_as_written():
    """The fitted ACES matrices' rows sum to 1, so a grey stays grey when they multiply by rows;
    applied transposed as written (ToneMap.cpp:19-24) the column sums differ and greys tint. Large
    greys approach the fit's limit 1/0.983729 > 1 and clamp to 1; black maps below 0 and clamps to 0."""
    greys = np.array([0.0, 0.02, 0.18, 0.5, 1.0, 4.0, 1e4], dtype=np.float32)
    img = _img(np.repeat(greys[None, :, None], 3, axis=2))
    dc, _ = abi.post_desc(clamp=False, gamma=False, estimator=abi.ESTIMATOR_CORRECTED)
    df, _ = abi.post_desc(clamp=False, gamma=False, estimator=abi.ESTIMATOR_FAITHFUL)
    c, _ = oracle.post(img, dc)
    f, _ = oracle.post(img, df)
    c, f = c[0, :, :3], f[0, :, :3]
    np.testing.assert_allclose(c[:, 0], c[:, 1], rtol=2e-5, atol=1e-6)
    np.testing.assert_allclose(c[:, 1], c[:, 2], rtol=2e-5, atol=1e-6)
    assert np.all(np.diff(c[:, 0]) > 0)
    assert c[-1, 0] == pytest.approx(1 / 0.983729, rel=1e-3)
    assert c[0, 0] < 0
    assert np.abs(f[3] - f[3].mean()).max() > 1e-3  # tinted
    # closed form of the corrected curve on greys: RRT(g) (both matrices preserve greys)
    g = greys[3]
    rrt = (g * (g + 0.0245786) - 0.000090537) / (g * (0.983729 * g + 0.4329510) + 0.238081)
    assert c[3, 0] == pytest.approx(rrt, rel=1e-5)
    dcl, _ = abi.post_desc(clamp=True, gamma=False)
    cl, _ = oracle.post(img, dcl)
    assert cl[0, 0, 0] == 0.0 and cl[0, -1, 0] == 1.0


def test_gamma_and_packing():
    """PackChannels: (int)(c * 256) clamped to [0, 255]; gamma c^(1/2.2) before it."""
    vals = np.array([0.0, 0.25, 0.5, 0.999, 1.0, -1.0, 3.0], dtype=np.float32)
    img = _img(np.repeat(vals[None, :, None], 3, axis=2))
    d, _ = abi.post_desc(tonemap=abi.TONEMAP_NONE, clamp=False, gamma=False)
    _, b = oracle.post(img, d)
    assert b[0, :, 0].tolist() == [0, 64, 128, 255, 255, 0, 255]
    assert (b[..., 3] == 255).all()
    d, _ = abi.post_desc(tonemap=abi.TONEMAP_NONE, clamp=True, gamma=True)
    _, b = oracle.post(img, d)
    want = [min(255, int(np.float32(v) ** np.float32(1 / 2.2) * 256)) if v > 0 else 0 for v in np.clip(vals, 0, 1)]
    assert b[0, :, 0].tolist() == want


def test_convolution_filter_edges():
    """An identity kernel leaves the image; a 3x1 box on a single bright pixel in the first column
    spreads it to its neighbour; as written the left neighbour of column 0 wraps (unsigned) to the
    last column, corrected clamps it to column 0."""
    img = _img(np.zeros((1, 5, 3)))
    img[0, 0, :3] = 3.0
    ident = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 0]], dtype=np.float32)
    d, k = abi.post_desc(tonemap=abi.TONEMAP_NONE, clamp=False, gamma=False, kernel=ident)
    a, _ = oracle.post(img, d)
    np.testing.assert_array_equal(a[..., :3], img[..., :3])
    box = np.ones((1, 3), dtype=np.float32)
    for est, col0, col4 in ((abi.ESTIMATOR_CORRECTED, 6.0, 0.0), (abi.ESTIMATOR_FAITHFUL, 3.0, 0.0)):
        d, k = abi.post_desc(tonemap=abi.TONEMAP_NONE, clamp=False, gamma=False, kernel=box, intensity=1.0,
                             estimator=est)
        a, _ = oracle.post(img, d)
        # corrected: column 0 sees (clamped 0, 0, 1) = 3 + 3 + 0; as written: (wrapped 4, 0, 1) = 0 + 3 + 0
        assert a[0, 0, 0] == pytest.approx(col0) and a[0, 1, 0] == pytest.approx(3.0)
        assert a[0, 4, 0] == pytest.approx(col4 if est == abi.ESTIMATOR_CORRECTED else 0.0)


def test_png_writer_is_a_valid_png(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, size=(37, 70001 // 37 // 4 + 5, 4), dtype=np.uint8)  # > one stored block
    p = tmp_path / "a.png"
    gpu.write_png(p, img)
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body)
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[0][1][:10])
    assert (w, h, depth, ctype) == (img.shape[1], img.shape[0], 8, 6)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    np.testing.assert_array_equal(rows[:, 1:].reshape(h, w, 4), img)
