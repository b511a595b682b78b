"""Writes tests/golden/bfr_tess_quad.npz: Bfr::Tessellation::GetCoords<float>
of a Catmark quad at uniform rates, computed by the reference's own Bfr
(built into oracle/_ref from the reference sources).  Data only: the rate
and the (u, v) pairs.

    python tests/golden/make_tess_golden.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from oracle import ref  # noqa: E402

RATES = (1, 2, 3, 4, 5, 7, 8, 16, 33)


def main():
    out = {f"rate{r}": ref.tess_uniform_coords(4, r) for r in RATES}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bfr_tess_quad.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
