"""Oracle pinning for the film output (camera/Film.cu:120-362,
camera/BMP_Format.cu).

The reference's film code needs the CUDA headers (Film.cu includes
cuda_libraries/math.cu, which includes thrust and curand), so it cannot be
built here; the oracle's operators are pinned instead by:
  * the published closed forms, evaluated here with numpy in float64:
    Reinhard et al. 2002 eq. 4 (extended operator, white = max luminance),
    Hable's filmic curve (A..F = .15 .5 .1 .2 .02 .3, exposure 2, W = 11.2),
    Hill's fitted ACES RRT+ODT with its input/output matrices, and the
    Reinhard-Jodie composition as the reference writes it (lerp with equal
    end points, i.e. pixel / (1 + luminance));
  * values the mathematics fixes (the brightest channel under equalize and
    gamma is exactly 1, the brightest luminance under Reinhard maps to 1,
    black stays black);
  * the byte conversion against x86's truncating conversion rules;
  * the BMP and PPM pixel arrays decoded independently (BITMAPINFOHEADER
    rules: bottom-up rows, B G R, 4-byte row padding; netpbm P6)."""
import struct

import numpy as np
import pytest

import oracle

W_LUM = np.array([0.212671, 0.715160, 0.072169], np.float32).astype(np.float64)


def film(w, h, seed=3, hi=6.0, zeros=True):
    rng = np.random.default_rng(seed)
    s = rng.random((h * w, 3)) * hi * rng.random((h * w, 1))
    if zeros:
        s[rng.random(h * w) < 0.1] = 0.0
    return s


def lum(p):
    p = np.asarray(p, np.float64)
    return p[..., 0] * W_LUM[0] + p[..., 1] * W_LUM[1] + p[..., 2] * W_LUM[2]


def reinhard_2002(p, l_white):
    L = lum(p)[..., None]
    Ld = L * (1 + L / (l_white * l_white)) / (1 + L)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = np.where(L > 0, p * (Ld / L), p)
    return np.where(L > 0, np.clip(out, 0, 1), out)


def hable(x):
    A, B, C, D, E, F = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F


def aces_hill(p):
    m_in = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
    m_out = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])
    v = p @ m_in.T
    v = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    return np.clip(v @ m_out.T, 0, 1)


@pytest.mark.parametrize("samples", [1, 7, 64])
def test_maxima(samples):
    s = film(13, 9)
    m, ml = oracle.film_maxima(s, samples)
    assert m == s.max() / samples
    assert ml == pytest.approx(lum(s).max() / samples, rel=1e-15)


def test_maxima_with_valid_counts():
    """nvalid: pixel sum = sums * samples / nvalid (0 where nvalid is 0)."""
    s = film(8, 8)
    nv = np.random.default_rng(5).integers(0, 17, 64).astype(np.uint32)
    est = np.where(nv[:, None] > 0, s * (16 / np.maximum(nv, 1)[:, None].astype(np.float64)), 0.0)
    assert oracle.film_maxima(s, 16, nv) == oracle.film_maxima(est, 16)
    a, ma = oracle.film_image(s, 8, 8, 16, "reinhard", nvalid=nv)
    b, mb = oracle.film_image(est, 8, 8, 16, "reinhard")
    assert np.array_equal(a, b) and np.array_equal(ma, mb)


@pytest.mark.parametrize("mode", oracle.TONE_MAPPINGS)
def test_operator_closed_forms(mode):
    w, h, n = 11, 7, 3
    s = film(w, h, hi=9.0)
    _, mapped = oracle.film_image(s, w, h, n, mode)
    mapped = mapped.reshape(-1, 3)
    p = s / n
    m, ml = s.max() / n, lum(s).max() / n
    if mode == "equalize":
        want = p / m
    elif mode == "reinhard":
        want = reinhard_2002(p, ml)
    elif mode == "reinhardJodie":
        want = np.clip(p / (1 + lum(p))[:, None], 0, 1)
    elif mode == "ACESfilmic":
        want = aces_hill(p)
    elif mode == "uncharted2":
        want = hable(2 * p) / hable(11.2)
    else:
        want = (p / m) ** (1 / 2.2)
    # float-narrowed constants (0.983729f, 0.15f ...) move values by ~1e-8
    tol = 1e-6 if mode in ("ACESfilmic", "uncharted2") else 1e-12
    assert np.allclose(mapped, want, rtol=tol, atol=tol)


def test_fixed_points():
    w, h = 9, 4
    s = film(w, h)
    for mode in ("equalize", "gammaClamp"):
        px, mapped = oracle.film_image(s, w, h, 5, mode, "ppm")
        i = np.argmax(s)
        assert mapped[i] == 1.0 and px[i] == 255
    # Reinhard: the pixel of largest luminance maps to luminance 1
    _, mapped = oracle.film_image(s, w, h, 5, "reinhard")
    k = np.argmax(lum(s))
    assert lum(mapped.reshape(-1, 3)[k]) == pytest.approx(1.0, rel=1e-12) or mapped.reshape(-1, 3)[k].max() == 1.0
    # black pixels give black bytes under every operator
    z = np.zeros((w * h, 3))
    z[0] = 1.0
    for mode in oracle.TONE_MAPPINGS:
        px, _ = oracle.film_image(z, w, h, 1, mode, "ppm")
        assert not px[3:].any(), mode


def test_uncharted2_black_is_not_exactly_zero_but_gives_byte_zero():
    """The float constants make f(0) = D*E/(D*F) - E/F a rounding residue,
    not 0; it stays far below one byte step."""
    v = oracle.tonemap_pixel([0, 0, 0], 1, 1, 1, "uncharted2")
    assert np.all(np.abs(v) < 1e-7)
    assert [oracle.to_byte(x) for x in v] == [0, 0, 0]


def test_byte_conversion_follows_x86():
    """(unsigned char)(v * 255): truncation to int32, low byte kept; out of
    range and NaN give INT_MIN, whose low byte is 0."""
    cases = {0.0: 0, 0.5: 127, 1.0: 255, 0.999: 254, -0.001: 0, -0.01: 254, 1.2: 306 & 255,
             2.0: 510 & 255, float("nan"): 0, 1e10: 0, -1e10: 0, float("inf"): 0}
    for v, b in cases.items():
        assert oracle.to_byte(v) == b, v


def test_uncharted2_wraps_bright_pixels():
    """No clamp: values above 1 wrap modulo 256 as the reference's bytes do."""
    v = oracle.tonemap_pixel([40, 40, 40], 1, 1, 1, "uncharted2")
    assert v[0] > 1.0
    assert oracle.to_byte(v[0]) == int(v[0] * 255) % 256


def decode_bmp(data):
    assert data[:2] == b"BM"
    off, = struct.unpack_from("<I", data, 10)
    hsize, w, h, planes, bpp, comp = struct.unpack_from("<IiiHHI", data, 14)
    assert (hsize, planes, bpp, comp) == (40, 1, 24, 0) and h > 0
    stride = (w * 24 + 31) // 32 * 4
    img = np.zeros((h, w, 3), np.uint8)
    for r in range(h):
        row = np.frombuffer(data, np.uint8, w * 3, off + r * stride).reshape(w, 3)
        img[h - 1 - r] = row[:, ::-1]  # bottom-up, B G R
        pad = data[off + r * stride + 3 * w: off + (r + 1) * stride]
        assert pad == b"\0" * len(pad)
    assert len(data) == off + h * stride
    return img


def decode_ppm(data):
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255"
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)


@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (3, 2), (4, 4), (5, 3), (7, 5), (16, 2)])
def test_bmp_and_ppm_decode_to_the_same_image(w, h):
    s = film(w, h, seed=w * 31 + h)
    for mode in oracle.TONE_MAPPINGS:
        bmp, mapped = oracle.film_image(s, w, h, 2, mode, "bmp")
        ppm, _ = oracle.film_image(s, w, h, 2, mode, "ppm")
        a = decode_bmp(oracle.image_header("bmp", w, h) + bmp.tobytes())
        b = decode_ppm(oracle.image_header("ppm", w, h) + ppm.tobytes())
        assert np.array_equal(a, b)
        want = np.array([oracle.to_byte(v) for v in mapped], np.uint8).reshape(h, w, 3)
        assert np.array_equal(b, want)


def test_bmp_header_fields():
    """BitmapImage's header, including its Image_Size of (w + 4 - w % 4) * h
    (not the padded pixel array's size) and File_Size = 54 + Image_Size."""
    for w, h in [(5, 3), (8, 2), (1920, 1080)]:
        hd = oracle.image_header("bmp", w, h)
        assert len(hd) == 54
        fsize, r1, r2, off = struct.unpack_from("<IHHI", hd, 2)
        f = struct.unpack_from("<IiiHHIIiiII", hd, 14)
        isz = (w + 4 - w % 4) * h
        assert (fsize, r1, r2, off) == (54 + isz, 0, 0, 54)
        assert f == (40, w, h, 1, 24, 0, isz, 3780, 3780, 0, 0)
    assert oracle.image_header("ppm", 1920, 1080) == b"P6\n1920 1080\n255\n"
