"""Film checkpoints on the host (rp_checkpoint_write / rp_checkpoint_read,
counterpart of src/checkpoint.cpp): SHA-256 against hashlib, the RYPCHK
format-3 layout against an independent encoder and decoder written here from
the reference's specification (src/checkpoint.cpp:17-122, 319-455,
src/checkpoint_contract.h:14), the rejection cases of the reference's
checkpoint tests (corrupt, truncated, wrong format or contract word,
inconsistent provenance or counts, nonfinite sums), the carry-over
refusals (filters, splats, compatibility output, samplers this library does
not render), the earlier format 1 still read, and failed publication
keeping the destination."""
import hashlib
import os
import struct

import numpy as np
import pytest

import rypath_amd as rp

MAGIC3 = b"RYPCHK\r\n"
CONTRACT = (3, 12, 29, 11, 2)  # format, identity, renderer, sampling, accumulation
MAGIC1 = b"RPAFILM\n"
PIXEL = 216


def encode(width, height, samples, seed, randomization, identity, segments, planes, fmt=1):
    """This library's earlier format 1 (read only)."""
    head = MAGIC1 + struct.pack("<IIII", fmt, width, height, randomization) + struct.pack("<QQ", seed, samples)
    head += identity + struct.pack("<I", len(segments))
    for b, e, dev in segments:
        d = dev.encode()
        head += struct.pack("<QQI", b, e, len(d)) + d
    body = head + np.ascontiguousarray(planes, "<f8").tobytes()
    return body + hashlib.sha256(body).digest()


def _str(s):
    b = s.encode()
    return struct.pack("<I", len(b)) + b


def encode3(width, height, samples, seed, randomization, identity, segments, planes, *, job=None, batch=None,
            sampler=1, material_seed=0, model=1, exposure=1.0, packet=1, compat=0, max_component=float("inf"),
            filt=(0, 0.5, 0.5, 0.5), counts=None, extra=None, weights=None, contract=CONTRACT, trailing=b""):
    """Format 3 written field by field from src/checkpoint.cpp:85-131."""
    npx = width * height
    meta = struct.pack("<IIQQQIIiI", width, height, samples if job is None else job,
                       (samples or 1) if batch is None else batch, seed, sampler, randomization, material_seed, model)
    meta += struct.pack("<dBBd", exposure, packet, compat, max_component)
    meta += struct.pack("<Ifff", *filt) + identity + struct.pack("<I", len(segments))
    for b, e, build, dev, arch, osname in segments:
        meta += struct.pack("<QQ", b, e) + _str(build) + _str(dev) + _str(arch) + _str(osname)
    meta += trailing
    counts = np.full(npx, samples, np.uint64) if counts is None else counts
    extra = np.zeros((npx, 18)) if extra is None else extra
    if weights is None:
        weights = np.zeros((npx, 2))
        if packet:
            weights[:, 0] = counts.astype(np.float64)
    rec = np.zeros(npx, np.dtype([("sum", "<f8", 3), ("corr", "<f8", 3), ("count", "<u8"), ("extra", "<f8", 18),
                                  ("w", "<f8", 2)]))
    planes = np.asarray(planes, np.float64).reshape(6, npx)
    rec["sum"] = planes[0::2].T
    rec["corr"] = planes[1::2].T
    rec["count"], rec["extra"], rec["w"] = counts, extra, weights
    assert rec.dtype.itemsize == PIXEL
    payload = rec.tobytes()
    head = MAGIC3 + struct.pack("<5I", *contract) + struct.pack("<QQ", len(meta), len(payload))
    assert len(head) == 44
    body = head + meta + payload
    return body + hashlib.sha256(body).digest()


def decode3(blob):
    """Parses a format-3 file into a dict (the reader of src/checkpoint.cpp:319-455)."""
    assert blob[:8] == MAGIC3
    words = struct.unpack_from("<5I", blob, 8)
    meta_size, payload = struct.unpack_from("<QQ", blob, 28)
    assert len(blob) == 44 + meta_size + payload + 32
    assert hashlib.sha256(blob[:-32]).digest() == blob[-32:]
    m = blob[44:44 + meta_size]
    out = dict(words=words)
    (out["width"], out["height"], out["job"], out["batch"], out["seed"], out["sampler"], out["randomization"],
     out["material_seed"], out["model"]) = struct.unpack_from("<IIQQQIIiI", m, 0)
    o = struct.calcsize("<IIQQQIIiI")
    out["exposure"], out["packet"], out["compat"], out["max_component"] = struct.unpack_from("<dBBd", m, o)
    o += struct.calcsize("<dBBd")
    out["filter"] = struct.unpack_from("<Ifff", m, o)
    o += 16
    out["identity"] = m[o:o + 32]
    o += 32
    (n,) = struct.unpack_from("<I", m, o)
    o += 4
    segs = []
    for _ in range(n):
        b, e = struct.unpack_from("<QQ", m, o)
        o += 16
        strs = []
        for _ in range(4):
            (k,) = struct.unpack_from("<I", m, o)
            strs.append(m[o + 4:o + 4 + k].decode())
            o += 4 + k
        segs.append((b, e, *strs))
    assert o == meta_size == 118 + sum(32 + sum(len(x.encode()) for x in s[2:]) for s in segs)
    out["segments"] = segs
    rec = np.frombuffer(blob[44 + meta_size:44 + meta_size + payload],
                        np.dtype([("sum", "<f8", 3), ("corr", "<f8", 3), ("count", "<u8"), ("extra", "<f8", 18),
                                  ("w", "<f8", 2)]))
    out["pixels"] = rec
    return out


def sample_film(w=5, h=3, samples=7, seed=9):
    rng = np.random.default_rng(1)
    planes = np.zeros((6, w * h))
    planes[0::2] = rng.uniform(0, 10, (3, w * h)) * samples  # sums
    planes[1::2] = rng.normal(0, 1e-14, (3, w * h))          # corrections
    ident = hashlib.sha256(b"scene").digest()
    segs = [(0, 3, "rypath_amd C ABI 1", "AMD Instinct MI355X", "gfx950:sramecc+:xnack-", "Linux"),
            (3, samples, "other build", "other device", "gfx942", "Linux")]
    return w, h, samples, seed, 2, ident, segs, planes


@pytest.mark.parametrize("n", [0, 1, 55, 56, 64, 65, 1000, 4097])
def test_sha256_matches_hashlib(n):
    data = bytes((i * 31 + 7) & 255 for i in range(n))
    assert rp.sha256(data) == hashlib.sha256(data).digest()


@pytest.mark.parametrize("packet,sampler,model", [(1, 1, 1), (0, 0, 0), (0, 1, 1)])
def test_format3_round_trip_is_byte_identical(tmp_path, packet, sampler, model):
    w, h, samples, seed, rnd, ident, segs, planes = sample_film()
    blob = encode3(w, h, samples, seed, rnd, ident, segs, planes, job=64, batch=8, sampler=sampler,
                   material_seed=-3, model=model, packet=packet, max_component=float("inf"))
    src = tmp_path / "a.ckpt"
    src.write_bytes(blob)
    film = rp.Film.read(src)
    info = film.info()
    assert (info["width"], info["height"], info["samples"], info["seed"]) == (w, h, samples, seed)
    assert info["randomization"] == rnd and info["segments"] == 2 and info["scene_identity"] == ident
    assert (info["sampler_kind"], info["model"], info["packet"], info["format"]) == (sampler, model, packet, 3)
    assert (info["job_samples"], info["batch_size"]) == (64, 8)
    mean = ((planes[0::2] + planes[1::2]) / samples).astype(np.float32)  # (3, npx)
    np.testing.assert_array_equal(film.resolve(), mean.T.reshape(h, w, 3))
    dst = tmp_path / "b.ckpt"
    film.write(dst)
    assert dst.read_bytes() == blob
    assert not [p for p in os.listdir(tmp_path) if p.startswith(".rypath-checkpoint")]


def test_written_file_decodes_with_the_reference_layout(tmp_path):
    """A film written here parses with the format-3 reader restated above:
    contract words, unit box filter, one weight per sample for a packet film,
    zero splat and compatibility sums, and the job settings."""
    w, h, samples, seed, rnd, ident, segs, planes = sample_film()
    src = tmp_path / "old.ckpt"
    src.write_bytes(encode(w, h, samples, seed, rnd, ident, [(0, samples, "MI355X gfx950")], planes))
    film = rp.Film.read(src)
    assert film.info()["format"] == 1
    film.set_job(100, 10)
    film.write(tmp_path / "new.ckpt")
    d = decode3((tmp_path / "new.ckpt").read_bytes())
    assert d["words"] == CONTRACT
    assert (d["width"], d["height"], d["job"], d["batch"], d["seed"], d["randomization"]) == (w, h, 100, 10, seed, rnd)
    assert (d["sampler"], d["model"], d["packet"], d["compat"], d["exposure"]) == (1, 1, 1, 0, 1.0)
    assert d["filter"] == (0, 0.5, 0.5, 0.5) and d["max_component"] == float("inf") and d["identity"] == ident
    assert d["segments"] == [(0, samples, "", "MI355X gfx950", "", "")]
    p = d["pixels"]
    assert (p["count"] == samples).all() and (p["extra"] == 0).all()
    assert (p["w"][:, 0] == samples).all() and (p["w"][:, 1] == 0).all()
    np.testing.assert_array_equal(p["sum"], planes[0::2].T)
    np.testing.assert_array_equal(p["corr"], planes[1::2].T)
    with pytest.raises(rp.RypathError):
        film.set_job(samples - 1, 1)  # fewer than the film holds


def test_empty_film(tmp_path):
    f = rp.Film(4, 2)
    assert f.info()["samples"] == 0
    with pytest.raises(rp.RypathError):
        f.resolve()
    f.write(tmp_path / "e.ckpt")
    d = decode3((tmp_path / "e.ckpt").read_bytes())
    assert d["job"] == 1 and (d["pixels"]["count"] == 0).all() and (d["pixels"]["w"] == 0).all()
    g = rp.Film.read(tmp_path / "e.ckpt")
    assert g.info()["samples"] == 0 and g.info()["width"] == 4


def corrupt_cases():
    w, h, samples, seed, rnd, ident, segs, planes = sample_film()
    args = (w, h, samples, seed, rnd, ident)
    good = encode3(*args, segs, planes)
    flipped = bytearray(good)
    flipped[len(good) // 2] ^= 1
    nan_planes = planes.copy()
    nan_planes[0, 0] = np.nan
    counts = np.full(w * h, samples, np.uint64)
    counts[4] -= 1
    extra = np.zeros((w * h, 18))
    extra[2, 5] = 0.25
    weights = np.zeros((w * h, 2))
    weights[:, 0] = samples * 1.5
    seg1 = [segs[0]]
    gap = [segs[0], (4, samples) + segs[1][2:]]
    return [
        # the reference's own rejections
        ("checksum", bytes(flipped)),
        ("truncated", good[:-40]),
        ("magic", b"X" + good[1:]),
        ("format", encode3(*args, segs, planes, contract=(2,) + CONTRACT[1:])),
        ("contract", encode3(*args, segs, planes, contract=CONTRACT[:2] + (28,) + CONTRACT[3:])),
        ("gap", encode3(*args, gap, planes)),
        ("coverage", encode3(*args, seg1, planes)),
        ("nonfinite", encode3(*args, segs, nan_planes)),
        ("randomization", encode3(w, h, samples, seed, 9, ident, segs, planes)),
        ("empty-nonzero", encode3(w, h, 0, seed, rnd, ident, [], planes, job=4)),
        ("count-above-job", encode3(*args, segs, planes, job=samples - 1)),
        ("uneven-counts", encode3(*args, segs, planes, counts=counts)),
        ("packet-flag", encode3(*args, segs, planes, packet=2)),
        ("demo-packet", encode3(*args, segs, planes, model=0, packet=1)),
        ("demo-exposure", encode3(*args, segs, planes, model=0, packet=0, exposure=2.0)),
        ("trailing-metadata", encode3(*args, segs, planes, trailing=b"\0")),
        ("zero-batch", encode3(*args, segs, planes, batch=0)),
        # what this library's film cannot hold
        ("gaussian-filter", encode3(*args, segs, planes, filt=(2, 1.5, 1.5, 0.5))),
        ("wide-box", encode3(*args, segs, planes, filt=(0, 1.0, 1.0, 0.5))),
        ("splats", encode3(*args, segs, planes, extra=extra)),
        ("weights", encode3(*args, segs, planes, weights=weights)),
        ("halton", encode3(*args, segs, planes, sampler=2)),
        ("compatibility", encode3(*args, segs, planes, compat=1, max_component=10.0)),
        # format 1
        ("format1-checksum", encode(w, h, samples, seed, rnd, ident, [(0, samples, "a")], planes)[:-1] + b"\0"),
        ("format1-version", encode(w, h, samples, seed, rnd, ident, [(0, samples, "a")], planes, fmt=2)),
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
    src.write_bytes(encode3(w, h, samples, seed, rnd, ident, segs, planes))
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
