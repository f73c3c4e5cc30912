#!/usr/bin/env python3
"""Extract the reference's own golden data for the noise path into
tests/golden/ as small data fixtures.

Reads (as data, nothing is executed):
  src/liboslnoise/oslnoise_test.cpp   golden vectors of test_perlin/test_cell/
                                      test_hash (:86-365, eps 1e-3 at :28)
  testsuite/hash/ref/out.txt          exact hash() ints (shader test.osl,
                                      testshade -g 2 2 -center)
  testsuite/transformc/ref/out.txt    color-space transforms, 3 decimals
  testsuite/<noise tests>/ref/*.tif   uint8 reference images (512x512 RGB)
  testsuite/blackbody/ref/out.exr     half RGB image 1000x64 (ZIP-compressed
                                      OpenEXR, decoded here)

Run in a container that has the reference mounted at /root/reference:
    python tests/golden/make_fixtures.py
"""
import json
import os
import re

import numpy as np

REF = os.environ.get("OSL_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_oslnoise_test():
    src = open(os.path.join(REF, "src/liboslnoise/oslnoise_test.cpp")).read()
    out = {}
    # static float results_Nd[...] = { ... };  static Vec3 vresults_Nd[...] = {...};
    for fn in ("test_perlin", "test_cell", "test_hash"):
        body = src[src.index("\n" + fn + "()"):]
        body = body[:body.index("\n}\n")]
        tab = {}
        for m in re.finditer(r"static (float|Vec3) (v?results_\dd)\[[^\]]*\]\s*=\s*\{(.*?)\};",
                             body, re.S):
            kind, name, vals = m.groups()
            vals = vals.replace("Vec3", "")
            nums = [float(x) for x in re.findall(r"-?\d+(?:\.\d*)?(?:e-?\d+)?", vals)]
            if kind == "Vec3":
                nums = [nums[i:i + 3] for i in range(0, len(nums), 3)]
            tab[name] = nums
        out[fn] = tab
    return out


def parse_hash_out():
    rows = []
    for line in open(os.path.join(REF, "testsuite/hash/ref/out.txt")):
        m = re.match(r"hash\((.*)\) = (-?\d+)", line.strip())
        if not m:
            continue
        args, val = m.groups()
        rows.append({"args": args, "value": int(val)})
    return rows


def parse_transformc_out():
    """testsuite/transformc/ref/out.txt (shader test.osl, testshade 1x1:
    u = v = 0.5, Dx(u) = Dy(v) = 1), printed with %0.3f."""
    num = r"(-?\d+\.\d+)"
    tri = r"\s+".join([num] * 3)
    rows = {"pairs": [], "special": []}
    for line in open(os.path.join(REF, "testsuite/transformc/ref/out.txt")):
        line = line.strip()
        m = re.match(r"(\w+) " + tri + r" = (\w+) " + tri + r"$", line)
        if m:
            g = m.groups()
            rows["pairs"].append({"from": g[0], "in": [float(x) for x in g[1:4]],
                                  "to": g[4], "out": [float(x) for x in g[5:8]]})
            continue
        m = re.match(r'transformc \("(\w+)","(\w+)",\((.*)\)\) = (.*)$', line)
        if m:
            frm, to, a, b = m.groups()
            ins = [[float(x) for x in part.split()] for part in a.split("/")]
            outs = [[float(x) for x in part.split()] for part in b.split("/")]
            rows["special"].append({"from": frm, "to": to, "in": ins, "out": outs})
    return rows


def read_exr_half(rel):
    """A scanline OpenEXR file of HALF channels with ZIP (16 lines per
    block) or ZIPS (1 line) compression, as float16 [H, W, C] in the file's
    channel order (alphabetical): header attributes, the line offset table,
    then per block zlib data whose bytes are delta-coded and split into
    even / odd halves before the per-line, per-channel pixel rows."""
    import struct
    import zlib
    d = open(os.path.join(REF, rel), "rb").read()
    assert struct.unpack("<ii", d[:8]) == (20000630, 2)
    pos, attrs = 8, {}
    while True:
        e = d.index(b"\0", pos)
        name = d[pos:e].decode()
        pos = e + 1
        if not name:
            break
        e = d.index(b"\0", pos)
        typ = d[pos:e].decode()
        size = struct.unpack("<i", d[e + 1:e + 5])[0]
        attrs[name] = (typ, d[e + 5:e + 5 + size])
        pos = e + 5 + size
    chl, chans = attrs["channels"][1], []
    q = 0
    while chl[q] != 0:
        e = chl.index(b"\0", q)
        chans.append((chl[q:e].decode(), struct.unpack("<i", chl[e + 1:e + 5])[0]))
        q = e + 17
    assert all(t == 1 for _, t in chans), "HALF channels only"
    comp = attrs["compression"][1][0]
    lines_per_block = {2: 1, 3: 16}[comp]
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    nblocks = (h + lines_per_block - 1) // lines_per_block
    offsets = struct.unpack("<%dQ" % nblocks, d[pos:pos + 8 * nblocks])
    img = np.zeros((h, w, len(chans)), np.float16)
    for off in offsets:
        y, size = struct.unpack("<ii", d[off:off + 8])
        raw = np.frombuffer(zlib.decompress(d[off + 8:off + 8 + size]), np.uint8)
        t = (np.cumsum(raw.astype(np.int64)) - 128 * np.arange(len(raw))) % 256
        t[0] = raw[0]
        t = t.astype(np.uint8)
        half = (len(t) + 1) // 2
        out = np.empty_like(t)
        out[0::2] = t[:half]
        out[1::2] = t[half:]
        nl = min(lines_per_block, h - (y - y0))
        px = out.view("<f2").reshape(nl, len(chans), w)
        img[y - y0:y - y0 + nl] = px.transpose(0, 2, 1)
    return img, [c for c, _ in chans]


def read_tif(rel):
    from PIL import Image
    im = Image.open(os.path.join(REF, rel))
    return np.asarray(im.convert("RGB"), dtype=np.uint8)


def main():
    vec = {"oslnoise_test": parse_oslnoise_test(), "hash_out": parse_hash_out(),
           "source": {"oslnoise_test": "src/liboslnoise/oslnoise_test.cpp:86-365",
                      "hash_out": "testsuite/hash/ref/out.txt"}}
    with open(os.path.join(HERE, "noise_vectors.json"), "w") as f:
        json.dump(vec, f, indent=1)
    tc = parse_transformc_out()
    tc["source"] = "testsuite/transformc/ref/out.txt"
    with open(os.path.join(HERE, "transformc.json"), "w") as f:
        json.dump(tc, f, indent=1)
    imgs = {}
    for name in ("noise", "cellnoise", "hashnoise", "noise-simplex", "noise-gabor",
                 "noise-gabor2d-filter", "pnoise-gabor", "pnoise", "pnoise-perlin", "pnoise-cell",
                 "pnoise-generic"):
        imgs[name.replace("-", "_")] = read_tif("testsuite/%s/ref/out.tif" % name)
    bb, chans = read_exr_half("testsuite/blackbody/ref/out.exr")
    assert chans == ["B", "G", "R"]
    imgs["blackbody_rgb_half"] = bb[..., ::-1]  # R, G, B
    np.savez_compressed(os.path.join(HERE, "noise_images.npz"), **imgs)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
