"""Film checkpoints on the host (rp_checkpoint_write / rp_checkpoint_read,
counterpart of src/checkpoint.cpp): SHA-256 against hashlib, the file layout
against an independent writer written from its specification in
include/rp_io.h, and the rejection cases of the reference's checkpoint tests
(corrupt, truncated, wrong version, inconsistent provenance, nonfinite sums,
failed publication keeps the destination)."""
import hashlib
import os
import struct

import numpy as np
import pytest

import rypath_amd as rp

MAGIC = b"RPAFILM\n"


def encode(width, height, samples, seed, randomization, identity, segments, planes, fmt=1):
    head = MAGIC + struct.pack("<IIII", fmt, width, height, randomization) + struct.pack("<QQ", seed, samples)
    head += identity + struct.pack("<I", len(segments))
    for b, e, dev in segments:
        d = dev.encode()
        head += struct.pack("<QQI", b, e, len(d)) + d
    body = head + np.ascontiguousarray(planes, "<f8").tobytes()
    return body + hashlib.sha256(body).digest()


def sample_film(w=5, h=3, samples=7, seed=9):
    rng = np.random.default_rng(1)
    planes = np.zeros((6, w * h))
    planes[0::2] = rng.uniform(0, 10, (3, w * h)) * samples  # sums
    planes[1::2] = rng.normal(0, 1e-14, (3, w * h))          # corrections
    ident = hashlib.sha256(b"scene").digest()
    segs = [(0, 3, "MI355X gfx950"), (3, samples, "other device")]
    return w, h, samples, seed, 2, ident, segs, planes


@pytest.mark.parametrize("n", [0, 1, 55, 56, 64, 65, 1000, 4097])
def test_sha256_matches_hashlib(n):
    data = bytes((i * 31 + 7) & 255 for i in range(n))
    assert rp.sha256(data) == hashlib.sha256(data).digest()


def test_layout_round_trip_is_byte_identical(tmp_path):
    w, h, samples, seed, rnd, ident, segs, planes = sample_film()
    blob = encode(w, h, samples, seed, rnd, ident, segs, planes)
    src = tmp_path / "a.ckpt"
    src.write_bytes(blob)
    film = rp.Film.read(src)
    info = film.info()
    assert (info["width"], info["height"], info["samples"], info["seed"]) == (w, h, samples, seed)
    assert info["randomization"] == rnd and info["segments"] == 2 and info["scene_identity"] == ident
    mean = ((planes[0::2] + planes[1::2]) / samples).astype(np.float32)  # (3, npx)
    np.testing.assert_array_equal(film.resolve(), mean.T.reshape(h, w, 3))
    dst = tmp_path / "b.ckpt"
    film.write(dst)
    assert dst.read_bytes() == blob
    assert not [p for p in os.listdir(tmp_path) if p.startswith(".rypath-checkpoint")]


def test_empty_film(tmp_path):
    f = rp.Film(4, 2)
    assert f.info()["samples"] == 0
    with pytest.raises(rp.RypathError):
        f.resolve()
    f.write(tmp_path / "e.ckpt")
    g = rp.Film.read(tmp_path / "e.ckpt")
    assert g.info()["samples"] == 0 and g.info()["width"] == 4


def corrupt_cases():
    w, h, samples, seed, rnd, ident, segs, planes = sample_film()
    good = encode(w, h, samples, seed, rnd, ident, segs, planes)
    flipped = bytearray(good)
    flipped[len(good) // 2] ^= 1
    nan_planes = planes.copy()
    nan_planes[0, 0] = np.nan
    return [
        ("checksum", bytes(flipped)),
        ("truncated", good[:-40]),
        ("magic", b"X" + good[1:]),
        ("format", encode(w, h, samples, seed, rnd, ident, segs, planes, fmt=2)),
        ("gap", encode(w, h, samples, seed, rnd, ident, [(0, 3, "a"), (4, samples, "b")], planes)),
        ("coverage", encode(w, h, samples, seed, rnd, ident, [(0, 3, "a")], planes)),
        ("nonfinite", encode(w, h, samples, seed, rnd, ident, segs, nan_planes)),
        ("size", encode(w, h + 1, samples, seed, rnd, ident, segs, planes)),
        ("randomization", encode(w, h, samples, seed, 9, ident, segs, planes)),
        ("empty-nonzero", encode(w, h, 0, seed, rnd, ident, [], planes)),
    ]


@pytest.mark.parametrize("name,blob", corrupt_cases(), ids=[c[0] for c in corrupt_cases()])
def test_corrupt_checkpoints_are_rejected(tmp_path, name, blob):
    p = tmp_path / "c.ckpt"
    p.write_bytes(blob)
    with pytest.raises(rp.RypathError) as e:
        rp.Film.read(p)
    assert e.value.code == rp.FAILURE


def test_failed_publication_keeps_destination(tmp_path):
    w, h, samples, seed, rnd, ident, segs, planes = sample_film()
    src = tmp_path / "a.ckpt"
    src.write_bytes(encode(w, h, samples, seed, rnd, ident, segs, planes))
    film = rp.Film.read(src)
    with pytest.raises(rp.RypathError):
        film.write(tmp_path / "missing" / "x.ckpt")
    target = tmp_path / "dir.ckpt"
    target.mkdir()
    with pytest.raises(rp.RypathError):
        film.write(target)
    assert target.is_dir()
    with pytest.raises(rp.RypathError):
        rp.Film.read(tmp_path)  # not a regular file
    assert not [p for p in os.listdir(tmp_path) if p.startswith(".rypath-checkpoint")]
