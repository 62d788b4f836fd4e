"""Writes the image files of hip_textures.pbrt (data of this repository):
8-bit RGB PNGs (stored, zlib-compressed scanlines with filter 0) and a flat
(non run-length) Radiance RGBE environment map. Deterministic; run from any
directory: python tests/golden/pbrt/make_textures.py"""
import os
import struct
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "textures")


def write_png(path, rgb):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w, _ = rgb.shape
    raw = b"".join(b"\0" + rgb[y].tobytes() for y in range(h))

    def chunk(tag, data):
        body = tag + data
        return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
    data += chunk(b"IDAT", zlib.compress(raw, 9)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def write_hdr(path, rgb):
    """rgb: h x w x 3 floats (top row first), written flat as RGBE."""
    h, w, _ = rgb.shape
    m = rgb.max(axis=2)
    e = np.zeros((h, w), np.int32)
    mant = np.zeros((h, w, 3), np.float64)
    nz = m > 1e-32
    fr, ex = np.frexp(m[nz])
    e[nz] = ex
    mant[nz] = rgb[nz] / m[nz][:, None] * fr[:, None] * 256.0
    px = np.zeros((h, w, 4), np.uint8)
    px[..., :3] = np.clip(mant, 0, 255).astype(np.uint8)
    px[..., 3] = np.where(nz, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + px.tobytes())


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(11)
    # 8 x 8 colour checker with a random tint per cell
    cells = rng.integers(40, 250, (8, 8, 3))
    write_png(os.path.join(OUT, "checker8.png"), cells)
    # 5 x 3 odd-sized tile (mip chain with wrapped odd edges)
    write_png(os.path.join(OUT, "tile5x3.png"), rng.integers(20, 255, (3, 5, 3)))
    # 16 x 16 stripes (minification under trilinear / EWA)
    y, x = np.mgrid[0:16, 0:16]
    stripes = np.stack([(x % 4 < 2) * 200 + 30, (y % 4 < 2) * 180 + 40, ((x + y) % 3) * 90 + 20], axis=2)
    write_png(os.path.join(OUT, "stripes16.png"), stripes)
    # 8 x 8 tangent-space normal map, tilted up to ~35 degrees (encoded (n + 1) / 2)
    xy = rng.uniform(-0.5, 0.5, (8, 8, 2))
    n = np.concatenate([xy, np.ones((8, 8, 1))], axis=2)
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    write_png(os.path.join(OUT, "bumps8.png"), np.round((n + 1.0) * 0.5 * 255.0))
    # 16 x 8 sky: brighter towards the zenith, a warm patch near the horizon
    v = (np.arange(8) + 0.5) / 8.0
    u = (np.arange(16) + 0.5) / 16.0
    sky = np.zeros((8, 16, 3))
    sky[..., 0] = 0.3 + 0.5 * (1.0 - v)[:, None]
    sky[..., 1] = 0.4 + 0.5 * (1.0 - v)[:, None]
    sky[..., 2] = 0.6 + 0.6 * (1.0 - v)[:, None]
    sky[3:5, 4:7] = (3.0, 2.4, 1.5)
    sky *= (0.9 + 0.2 * u)[None, :, None]
    write_hdr(os.path.join(OUT, "sky16x8.hdr"), sky)


if __name__ == "__main__":
    main()
