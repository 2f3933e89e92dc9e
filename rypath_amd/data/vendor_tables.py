"""Re-derive the vendored third-party data tables from the reference checkout.

The outputs are committed; this script only documents (and can repeat) how
they were produced.  Inputs (paths under the reference checkout):

* third_party/pbrt/sobol_tables.inc -- PBRT v4 b0bdf461 Sobol direction words
  (1024 x 52 u32) and the interval-index matrices (25 x 52 and 26 x 52 u64).
  Apache-2.0 (PBRT) and MIT (Gruenschloss); data of Joe & Kuo (2008).
* third_party/pbrt/srgb.bin -- PBRT v4 sRGB-to-sigmoid coefficient table
  (64 scale nodes + 3 x 64^3 x 3 binary32, little endian).  Apache-2.0.
* third_party/cie/CIE_xyz_1931_2deg.csv -- CIE 1931 2-degree observer,
  1 nm, 360..830 nm.  CC BY-SA 4.0 (CIE 2019, DOI 10.25039/CIE.DS.xvudnb9b).
* third_party/pbrt/d65.inc -- PBRT v4 stdillum-D65 (CIE standard illuminant
  D65, normalised), 471 binary32 values at 360..830 nm.  Apache-2.0.

Outputs (this directory): sobol_tables.bin (234208 B: directions u32,
intervals u64, inverse u64, little endian), srgb_sigmoid.bin (verbatim copy),
cie_xyz_1931_2deg.csv (verbatim copy), d65.bin (471 binary32, little endian).
"""
import hashlib
import re
import shutil
import struct
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def parse_sobol(text):
    body = "\n".join(line for line in text.splitlines() if not line.lstrip().startswith("//"))
    blocks = re.findall(r"\{([^{}]*)\}", body)
    if len(blocks) != 3:
        raise SystemExit(f"expected 3 arrays, found {len(blocks)}")
    vals = [[int(tok.rstrip("ULul"), 0) for tok in re.findall(r"\b(?:0x[0-9a-fA-F]+|\d+)[ULul]*", b)] for b in blocks]
    return vals


def parse_hex_floats(text):
    body = "\n".join(line for line in text.splitlines() if not line.lstrip().startswith("//"))
    return [float.fromhex(tok) for tok in re.findall(r"(-?0x[0-9a-fA-F.]+p[-+]?\d+)f?", body)]


def main(ref=Path("/root/reference")):
    inc = (ref / "third_party/pbrt/sobol_tables.inc").read_bytes()
    sha = hashlib.sha256(inc).hexdigest()
    assert sha == "349a336cd9dff8bdc06cd30490b1cf887c3f041c5fd00ff3bba5ea9e7f38aeab", sha
    directions, intervals, inverse = parse_sobol(inc.decode())
    assert len(directions) == 1024 * 52, len(directions)
    assert len(intervals) == 25 * 52, len(intervals)
    assert len(inverse) == 26 * 52, len(inverse)
    blob = struct.pack("<%dI" % len(directions), *directions)
    blob += struct.pack("<%dQ" % len(intervals), *intervals)
    blob += struct.pack("<%dQ" % len(inverse), *inverse)
    assert len(blob) == 234208
    (HERE / "sobol_tables.bin").write_bytes(blob)
    shutil.copyfile(ref / "third_party/pbrt/srgb.bin", HERE / "srgb_sigmoid.bin")
    shutil.copyfile(ref / "third_party/cie/CIE_xyz_1931_2deg.csv", HERE / "cie_xyz_1931_2deg.csv")
    d65 = parse_hex_floats((ref / "third_party/pbrt/d65.inc").read_text())
    assert len(d65) == 471, len(d65)
    (HERE / "d65.bin").write_bytes(struct.pack("<471f", *d65))
    for name in ("sobol_tables.bin", "srgb_sigmoid.bin", "cie_xyz_1931_2deg.csv", "d65.bin"):
        print(name, hashlib.sha256((HERE / name).read_bytes()).hexdigest())


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
