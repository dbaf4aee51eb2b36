"""Writes the golden frames of tests/scenes.py GOLDEN_SCENES, and the
digests of its FULL_SIZE_SCENES (full_sizes.json).

Renders each scene through a CPU build of the reference (oracle/_ref,
built by `make -C oracle ref` from the reference sources) with one OpenMP
thread -- the reference's row-parallel loop races on shared 8-pixel stores
(triangles.c:402, 434) -- and stores colour (packed RGBA8) and depth in one
compressed .npz per scene together with the scene parameters.

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py [NAME ...]

With names, only those scenes are written.  A scene listed in
scenes.GOLDEN_REF_OVERRIDES is rendered by the reference with those extra
parameters (the file keeps the scene's own parameters).  The bilinear
scenes (scenes.BILINEAR_GOLDEN_NAMES) are rendered by oracle/_ref's build of
the reference with the one-token colour-lerp fix (`make -C oracle
ref-bilinear`): the unmodified build's bilinear output is undefined.
"""
import json
import os
import sys

os.environ["OMP_NUM_THREADS"] = "1"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import pixelforge_amd as pf  # noqa: E402
import scenes  # noqa: E402

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref.so")
REF_BILINEAR_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref_bilinear.so")


def full_sizes(only):
    """Digests of the benchmark configurations at full size
    (scenes.FULL_SIZE_SCENES) into full_sizes.json."""
    path = os.path.join(HERE, "full_sizes.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    for name, kw in scenes.FULL_SIZE_SCENES:
        if only and name not in only:
            continue
        so = REF_BILINEAR_SO if name in scenes.FULL_SIZE_BILINEAR else REF_SO
        color, depth, err = pf.render(scenes.params(kw), so)
        if err:
            sys.exit(f"{name}: reference reported error {err}")
        c, d = scenes.digest(color, depth)
        out[name] = {"params": kw, "color_sha256": c, "depth_sha256": d,
                     "reference": "bilinear colour-lerp fix" if name in scenes.FULL_SIZE_BILINEAR else "plain"}
        print(name, color.shape)
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    for so in (REF_SO, REF_BILINEAR_SO):
        if not os.path.exists(so):
            sys.exit(f"{so} missing: run make -C oracle ref")
    only = set(sys.argv[1:])
    full_sizes(only)
    for name, kw in scenes.GOLDEN_SCENES:
        if only and name not in only:
            continue
        p = scenes.params(dict(kw, **scenes.GOLDEN_REF_OVERRIDES.get(name, {})))
        so = REF_BILINEAR_SO if name in scenes.BILINEAR_GOLDEN_NAMES else REF_SO
        color, depth, err = pf.render(p, so)
        if err:
            sys.exit(f"{name}: reference reported error {err}")
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), color=color, depth=depth,
                            params=json.dumps(kw))
        print(name, color.shape)


if __name__ == "__main__":
    main()
