"""Writes tests/golden/hbr_catmark.npz from the reference's own regression
data: the Catmark shapes of regression/hbr_regression (control meshes and
tags from regression/shapes/<name>.h) and their Hbr baselines
(regression/hbr_regression/baseline/<name>_level<l>.obj: every vertex of
uniform subdivision level l + 1, in Hbr's creation order).

Only numbers are kept: control points, faces, crease / corner tags, the
boundary and creasing options the tags select, and the baseline positions
of levels 1-3 (float32).  Shapes with hierarchical edits (an Hbr-only
feature) are left out.

Run in a container that has the reference tree:

    python tests/golden/make_hbr_golden.py /root/reference
"""
import os
import re
import sys

import numpy as np

SHAPES = ["catmark_chaikin0", "catmark_chaikin1", "catmark_cube", "catmark_cube_corner0",
          "catmark_cube_corner1", "catmark_cube_corner2", "catmark_cube_corner3",
          "catmark_cube_corner4", "catmark_cube_creases0", "catmark_cube_creases1",
          "catmark_dart_edgecorner", "catmark_dart_edgeonly", "catmark_edgecorner",
          "catmark_edgeonly", "catmark_flap", "catmark_flap2", "catmark_pyramid",
          "catmark_pyramid_creases0", "catmark_pyramid_creases1", "catmark_tent",
          "catmark_tent_creases0", "catmark_tent_creases1"]
LEVELS = (1, 2, 3)

# regression/common/far_utils.h:84-94: "interpolateboundary" 0 none, 1 edge
# and corner, 2 edge only (default edge only); mapped to the table
# compiler's numbering 0 none, 1 edge only, 2 edge and corner
BOUNDARY = {0: 0, 1: 2, 2: 1}


def obj_text(header_path):
    """The OBJ text of a regression/shapes header (a C string literal)."""
    out = []
    for line in open(header_path):
        m = re.match(r'\s*"(.*)"\s*$', line.rstrip("\n"))
        if m:
            out.append(m.group(1).replace("\\n", "\n"))
    return "".join(out)


def parse_obj(text):
    verts, faces, tags = [], [], []
    for line in text.splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            verts.append([float(x) for x in tok[1:4]])
        elif tok[0] == "f":
            faces.append([int(x.split("/")[0]) - 1 for x in tok[1:]])
        elif tok[0] == "t":
            ni, nf, ns = (int(x) for x in tok[2].split("/"))
            args = tok[3:]
            tags.append((tok[1], [int(x) for x in args[:ni]],
                         [float(x) for x in args[ni:ni + nf]], args[ni + nf:ni + nf + ns]))
    return np.array(verts, np.float64), faces, tags


def main(ref_root):
    out = {}
    for name in SHAPES:
        verts, faces, tags = parse_obj(obj_text(os.path.join(ref_root, "regression", "shapes",
                                                             name + ".h")))
        pairs, pw, corners, cw = [], [], [], []
        boundary, creasing = 1, 0
        for tname, ints, floats, strings in tags:
            if tname == "crease":
                for j in range(0, len(ints) - 1, 2):
                    pairs.append((ints[j], ints[j + 1]))
                    pw.append(max(0.0, floats[j] if len(floats) > 1 else floats[0]))
            elif tname == "corner":
                for j, v in enumerate(ints):
                    corners.append(v)
                    cw.append(max(0.0, floats[j] if len(floats) > 1 else floats[0]))
            elif tname == "interpolateboundary":
                boundary = BOUNDARY[ints[0]]
            elif tname == "creasemethod":
                creasing = 1 if strings and strings[0] == "chaikin" else 0
            else:
                raise ValueError(f"{name}: unhandled tag {tname}")
        out[f"{name}/positions"] = verts.astype(np.float32)
        out[f"{name}/face_sizes"] = np.array([len(f) for f in faces], np.int32)
        out[f"{name}/face_verts"] = np.array([v for f in faces for v in f], np.int32)
        out[f"{name}/crease_pairs"] = np.array(pairs, np.int32).reshape(-1, 2)
        out[f"{name}/crease_weights"] = np.array(pw, np.float32)
        out[f"{name}/corner_verts"] = np.array(corners, np.int32)
        out[f"{name}/corner_weights"] = np.array(cw, np.float32)
        out[f"{name}/options"] = np.array([boundary, creasing], np.int32)
        for level in LEVELS:
            base = os.path.join(ref_root, "regression", "hbr_regression", "baseline",
                                f"{name}_level{level - 1}.obj")
            bverts, _, _ = parse_obj(open(base).read())
            out[f"{name}/level{level}"] = bverts.astype(np.float32)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hbr_catmark.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes,", len(SHAPES), "shapes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
