"""Result files and error metrics in the reference application's formats.

save_result writes what RTLibExtOPX7TestApplication::SaveResultImage writes
(RTLibExtOPX7TestApplication.cpp:2557-2624) for one accumulated image:

  result_<tracer>_<spp>.bin   float32 RGB of sum / spp, W * H pixels, in the
                              reverse of the frame buffer's pixel order (the
                              reference writes pixel i of the buffer at
                              N - 1 - i)
  result_<tracer>_<spp>.png   8-bit RGBA of min(gamma(sum / spp), 1) * 255.99
                              in the same reversed order, gamma
                              v <= 0.00174 ? 32 v : v^(1/2.2)
  config_<tracer>_<spp>.json  ImageConfigData (Core/Scene.h:88-128):
                              PngFilePath, ExrFilePath, BinFilePath, Width,
                              Height, Samples, Time (ms), EnableVis and, for
                              guided tracers, Custom {NumIterations,
                              TimesForItrerations[i] (ms)}

The OpenEXR file the reference also writes is not produced (no EXR writer
here); its path is still recorded.  compare() computes rMAE and SMAPE as
RTLibExtOPX7Comp.cpp:127-143 does over the whole image.
"""
from __future__ import annotations

import json
import struct
import zlib
from pathlib import Path

import numpy as np


def linear_to_gamma(v):
    v = np.asarray(v, np.float32)
    return np.where(v <= np.float32(0.00174), np.float32(32.0) * v,
                    np.power(np.maximum(v, 0), np.float32(1.0 / 2.2), dtype=np.float32)).astype(np.float32)


def _png(path: Path, rgba: np.ndarray):
    """Minimal PNG writer: 8-bit RGBA, no filtering, zlib level 6.  rgba[0] is the top row."""
    h, w, _ = rgba.shape
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def _buffer_order(img):
    """(H, W, 3) image with row 0 at the bottom -> (N, 3) in frame-buffer order."""
    return np.ascontiguousarray(img, np.float32).reshape(-1, 3)


def save_result(directory, tracer, accum_sum, spp, time_ms, times_for_iterations=None, enable_vis=False):
    """Write the reference's .bin, .png and config .json for an accumulated
    image (accum_sum: H x W x 3 sums, row 0 at the bottom, as
    Context.accum() returns).  Returns the dict written to the .json."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    h, w, _ = accum_sum.shape
    n = h * w
    buf = _buffer_order(accum_sum)
    hdr = (buf[::-1] / np.float32(spp)).astype(np.float32)
    stem = f"result_{tracer}_{spp}"
    bin_path, png_path, exr_path = d / f"{stem}.bin", d / f"{stem}.png", d / f"{stem}.exr"
    hdr.tofile(bin_path)
    g = np.minimum(linear_to_gamma(hdr), np.float32(1.0))
    rgba = np.empty((n, 4), np.uint8)
    rgba[:, :3] = (np.float32(255.99) * g).astype(np.uint8)
    rgba[:, 3] = 255
    _png(png_path, rgba.reshape(h, w, 4))
    cfg = {"PngFilePath": str(png_path), "ExrFilePath": str(exr_path), "BinFilePath": str(bin_path),
           "Width": int(w), "Height": int(h), "Samples": int(spp), "Time": int(time_ms), "EnableVis": bool(enable_vis)}
    if times_for_iterations is not None:
        custom = {"NumIterations": len(times_for_iterations)}
        for i, t in enumerate(times_for_iterations):
            custom[f"TimesForItrerations[{i}]"] = float(t)
        cfg["Custom"] = custom
    (d / f"config_{tracer}_{spp}.json").write_text(json.dumps(cfg, indent=4))
    return cfg


def load_bin(path, width, height):
    """A .bin written by save_result (or the reference): (N, 3) float32 in file order."""
    a = np.fromfile(path, np.float32)
    assert a.size == 3 * width * height, (a.size, width, height)
    return a.reshape(-1, 3)


def compare(base, comp):
    """(rMAE, SMAPE) of comp against base, (N, 3) float32 each, as the
    reference's comparison tool computes them over the whole image: pixels
    with a NaN or a zero-norm base value are skipped; SMAPE is divided by
    the pixel count, rMAE by the summed mean norm."""
    b = np.asarray(base, np.float32).reshape(-1, 3)
    c = np.asarray(comp, np.float32).reshape(-1, 3)
    ok = ~(np.isnan(b).any(1) | np.isnan(c).any(1))
    nb = np.sqrt(np.sum(b * b, axis=1, dtype=np.float32)).astype(np.float32)
    nc = np.sqrt(np.sum(c * c, axis=1, dtype=np.float32)).astype(np.float32)
    ok &= nb > 0
    diff = np.sqrt(np.sum((b - c) ** 2, axis=1, dtype=np.float32)).astype(np.float32)
    smape = np.float32(np.sum(diff[ok] / nb[ok], dtype=np.float32)) / np.float32(b.shape[0])
    mean = np.sum((nb[ok] + nc[ok]) * np.float32(0.5), dtype=np.float32)
    rmae = np.float32(np.sum(diff[ok], dtype=np.float32)) / mean
    return float(rmae), float(smape)
