"""Image writers of libnlpf (util::write_image, crates/akari_render/src/util/mod.rs:57-121) through
the host entry point, which needs no GPU: the PNG is decoded by PIL and compared with the film
oracle's quantiser; the EXR is parsed here from the OpenEXR 2.0 file layout."""
import os
import struct

import numpy as np
import pytest
from PIL import Image

import film_oracle as F
import nlpf as N

f32 = np.float32


def make_image(W, H, seed):
    rng = np.random.default_rng(seed)
    img = np.ones((H, W, 4), f32)
    img[..., :3] = rng.uniform(-0.2, 1.3, (H, W, 3)).astype(f32)
    flat = img.reshape(-1, 4)
    flat[0, :3] = [np.nan, 1.0, 0.0]
    flat[-1, :3] = [np.inf, -np.inf, 0.0031308]
    img[..., 3] = rng.random((H, W)).astype(f32)
    return img


def read_exr(path):
    """Minimal reader of single-part, uncompressed scanline OpenEXR files (OpenEXR file layout:
    magic, version, attribute list, offset table, chunks of (y, size, channel planes))."""
    b = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", b, 0)
    assert magic == 20000630 and version & 0xFF == 2 and version >> 8 == 0
    at, attrs = 8, {}
    while b[at] != 0:
        name_end = b.index(b"\0", at)
        name = b[at:name_end].decode()
        type_end = b.index(b"\0", name_end + 1)
        typ = b[name_end + 1:type_end].decode()
        (size,) = struct.unpack_from("<i", b, type_end + 1)
        attrs[name] = (typ, b[type_end + 5:type_end + 5 + size])
        at = type_end + 5 + size
    at += 1
    typ, v = attrs["channels"]
    assert typ == "chlist"
    chans, k = [], 0
    while v[k] != 0:
        e = v.index(b"\0", k)
        ptype, _plin, xs, ys = struct.unpack_from("<iBxxxii", v, e + 1)
        chans.append((v[k:e].decode(), ptype, xs, ys))
        k = e + 17
    assert attrs["compression"] == ("compression", b"\x00")
    xmin, ymin, xmax, ymax = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = xmax - xmin + 1, ymax - ymin + 1
    assert attrs["displayWindow"][1] == attrs["dataWindow"][1]
    for req in ("lineOrder", "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"):
        assert req in attrs
    offsets = struct.unpack_from(f"<{H}Q", b, at)
    out = {}
    for name, ptype, xs, ys in chans:
        assert ptype == 2 and xs == 1 and ys == 1  # FLOAT, full resolution
        out[name] = np.zeros((H, W), f32)
    for y, off in enumerate(offsets):
        ly, size = struct.unpack_from("<ii", b, off)
        assert ly == y and size == 4 * W * len(chans)
        data = np.frombuffer(b, f32, W * len(chans), off + 8).reshape(len(chans), W)
        for c, (name, *_rest) in enumerate(chans):
            out[name][y] = data[c]
    assert offsets[-1] + 8 + 4 * W * len(chans) == len(b)
    return [c[0] for c in chans], out


def test_exr_holds_the_linear_rgb_bit_for_bit(tmp_path):
    img = make_image(37, 11, 1)
    path = tmp_path / "sub" / "dir" / "out.exr"
    N.write_image_host(img, str(path))
    names, ch = read_exr(path)
    assert names == ["B", "G", "R"]
    for c, name in enumerate("RGB"):
        assert np.array_equal(ch[name].view(np.uint32), img[..., c].view(np.uint32)), name


@pytest.mark.parametrize("W,H", [(1, 1), (37, 11), (200, 150)])  # 200x150 spans two stored blocks
def test_png_matches_quantizer(tmp_path, W, H):
    img = make_image(W, H, W)
    path = tmp_path / "out.png"
    N.write_image_host(img, str(path))
    with Image.open(path) as im:
        assert im.mode == "RGB" and im.size == (W, H)
        got = np.asarray(im)
    assert np.array_equal(got, F.quantize_ldr(img))


def test_png_known_codes(tmp_path):
    vals = [0.0, 1.0, 0.5, 0.2, -1.0, np.nan, np.inf]
    img = np.array([[[v, v, v, 1.0] for v in vals]], f32)
    path = tmp_path / "k.png"
    N.write_image_host(img, str(path))
    with Image.open(path) as im:
        assert np.asarray(im)[0, :, 0].tolist() == [0, 254, 187, 123, 0, 0, 255]


def test_write_errors(tmp_path):
    img = np.zeros((2, 2, 4), f32)
    with pytest.raises(N.NlpfError) as e:
        N.write_image_host(img, "")
    assert e.value.code == 2
    d = tmp_path / "adir"
    os.makedirs(d / "x.png")  # a directory where the file should go
    with pytest.raises(N.NlpfError) as e:
        N.write_image_host(img, str(d / "x.png"))
    assert e.value.code == N.NLPF_E_IO and "cannot open" in str(e.value)
    with pytest.raises(ValueError):
        N.write_image_host(np.zeros((2, 2, 3), f32), str(tmp_path / "y.png"))
