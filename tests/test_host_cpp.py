"""Runs the C++ host-surface checks (tests/cpp/test_host.cpp, built into pilot_amd/lib)."""

import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from pilot_amd import PACKAGE_DIR

EXE = os.path.join(PACKAGE_DIR, "lib", "test_host")


def _run(mode, env=None):
    assert os.path.exists(EXE), "build the host surface first (make -C pilot_amd/csrc host)"
    p = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=300,
                       env=None if env is None else {**os.environ, **env})
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr


def _frame_bytes(w, h):
    """8-bit RGBA of the synthetic frame test_host.cpp writes: round(clamp(v, 0, 1) * 255)."""
    x, y, c = np.meshgrid(np.arange(w), np.arange(h), np.arange(4), indexing="xy")
    v = ((x * 37 + y * 101 + c * 53) % 256).astype(np.float32) / np.float32(255)  # shape (h, w, 4)
    v[0, 0] = [-0.5, 1.7, np.nan, 0.5]
    v = np.nan_to_num(np.clip(v, 0, 1), nan=0.0) * np.float32(255)
    return np.floor(v + np.float32(0.5)).astype(np.uint8)  # values are >= 0: half away from zero


def _png_chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body), kind
        chunks.append((kind, body))
        pos += 12 + n
    return chunks


def test_host_surface_cpu(tmp_path):
    _run("cpu", {"BEAMS_TEST_IMAGE_DIR": str(tmp_path)})
    w, h = 5, 3
    want = _frame_bytes(w, h)
    assert want[0, 0].tolist() == [0, 255, 0, 128] and want[2, 4, 1] == (4 * 37 + 2 * 101 + 53) % 256

    chunks = _png_chunks((tmp_path / "frame.png").read_bytes())
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    assert struct.unpack(">IIBBBBB", chunks[0][1]) == (w, h, 8, 6, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[1][1]), np.uint8).reshape(h, 1 + 4 * w)
    assert (raw[:, 0] == 0).all()  # filter type 0 on every scanline
    np.testing.assert_array_equal(raw[:, 1:].reshape(h, w, 4), want)

    ppm = (tmp_path / "frame.ppm").read_bytes()
    head = b"P6\n5 3\n255\n"
    assert ppm[:len(head)] == head
    np.testing.assert_array_equal(np.frombuffer(ppm[len(head):], np.uint8).reshape(h, w, 3), want[..., :3])

    # 160 x 120 RGBA: 76920 scanline bytes, two stored deflate blocks
    big = zlib.decompress(dict(_png_chunks((tmp_path / "frame_big.png").read_bytes()))[b"IDAT"])
    bw, bh = 160, 120
    rows = np.frombuffer(big, np.uint8).reshape(bh, 1 + 4 * bw)
    assert (rows[:, 0] == 0).all()
    np.testing.assert_array_equal(rows[:, 1:].reshape(bh, bw, 4), _frame_bytes(bw, bh))



@pytest.mark.gpu
def test_host_surface_gpu():
    _run("gpu")
