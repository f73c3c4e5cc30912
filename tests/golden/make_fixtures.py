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
                 "noise-gabor2d-filter", "pnoise-gabor"):
        imgs[name.replace("-", "_")] = read_tif("testsuite/%s/ref/out.tif" % name)
    np.savez_compressed(os.path.join(HERE, "noise_images.npz"), **imgs)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
