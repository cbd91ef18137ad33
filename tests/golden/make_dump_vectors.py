"""Writes tests/golden/dump_vectors.json: what the reference's own Python
dump reader (scripts/readPhantomDump.py of the reference tree, read_dump)
extracts from full dumps written by phantom_amd.dumpfile.write_fulldump
for two small seeded states (a hydro box and an MHD + dust box).

Run in a container where the reference tree is present:
    python tests/golden/make_dump_vectors.py [path/to/reference]
The fixture holds the reference reader's header sections and per-array
values; tests/test_dumpfile.py checks that our writer still produces
files the reference reader reads this way, and that our reader agrees."""
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests.golden import dump_states  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sys.path.insert(0, os.path.join(ref, "scripts"))
    import readPhantomDump  # the reference reader

    out = {}
    for name, make in dump_states.STATES.items():
        s = make()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "dump_00000")
            dump_states.write(path, s)
            r = readPhantomDump.read_dump(path)
        entry = {"desc": r["desc"]}
        for sec in ("part_numbers", "part_numbers_long", "quantities", "units"):
            entry[sec] = {k: (np.asarray(v).tolist() if np.ndim(v) else float(v) if isinstance(v, float) else int(v)
                              if np.issubdtype(type(v), np.integer) else float(v))
                          for k, v in r[sec].items()}
        entry["blocks"] = []
        for b in r["blocks"]:
            entry["blocks"].append({"dim": int(b["dim"]), "nintdb": int(b["nintdb"]), "ndouble": int(b["ndouble"]),
                                    "nsingle": int(b["nsingle"]),
                                    "data": {k: np.asarray(v).tolist() for k, v in b["data"].items()}})
        out[name] = entry
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dump_vectors.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", dst)


if __name__ == "__main__":
    main()
