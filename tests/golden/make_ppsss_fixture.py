"""Convert the reference's example parameter file into a fixture.

Reads epi_intg/ppsss (one key=value line per CalcParams key, in the order
Params::generate_paramfile writes them, epi_intg/Params.cpp:98-110) from the
reference tree and writes its (key, value text) pairs, in file order, to
tests/golden/ppsss.json.  The fixture is data only: the file's pairs,
unchanged.

    python tests/golden/make_ppsss_fixture.py [/path/to/reference]
"""
import json
import os
import sys


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "epi_intg", "ppsss")
    pairs = []
    with open(src) as f:
        for line in f:
            line = line.rstrip("\n")
            if "=" in line:
                k, v = line.split("=", 1)
                pairs.append([k, v])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ppsss.json")
    with open(out, "w") as f:
        f.write('{"source": "epi_intg/ppsss", "pairs": [\n')
        f.write(",\n".join(json.dumps(p) for p in pairs))
        f.write("\n]}\n")
    print(f"{len(pairs)} pairs -> {out}")


if __name__ == "__main__":
    main()
