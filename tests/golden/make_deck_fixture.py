"""Convert the reference's shipped input deck into a compressed fixture.

Reads epi5_MT/input_2layer_nfix16.dat (17 whitespace-separated columns per
cell, parsed by CellManager::_load_from_file, cellmanager.cpp:243-244) from
the reference tree and writes the numeric columns to
tests/golden/deck_2layer_nfix16.npz.  The fixture is data only: the deck's
rows, unchanged, as arrays.

    python tests/golden/make_deck_fixture.py [/path/to/reference]
"""
import os
import sys

import numpy as np

COLS = ["index", "state", "radius", "ageb", "agek", "ca2p", "x", "y", "z",
        "ca2p_avg", "rest_div_times", "ex_fat", "in_fat", "touch", "spr_nat_len",
        "pair", "fix_origin"]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "epi5_MT", "input_2layer_nfix16.dat")
    rows = np.loadtxt(src, dtype=np.float64)
    assert rows.shape[1] == len(COLS)
    out = {}
    for i, name in enumerate(COLS):
        col = rows[:, i]
        if name in ("index", "state", "rest_div_times", "touch", "pair", "fix_origin"):
            out[name] = col.astype(np.int32)
        else:
            out[name] = col
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "deck_2layer_nfix16.npz")
    np.savez_compressed(dst, **out)
    print(dst, rows.shape)


if __name__ == "__main__":
    main()
