"""Phantom full dump files (phantom_amd.dumpfile, the format of
write_fulldump / read_dump, readwrite_dumps.f90:45-716).

The fixture tests/golden/dump_vectors.json holds what the reference's own
Python dump reader (scripts/readPhantomDump.py, read_dump) extracted from
files our writer produced for two seeded states (tests/golden/
make_dump_vectors.py): the header sections by tag and every array by block
and tag.  Our writer must still produce those files (same header entries,
same arrays, same values) and our reader must read them the same way.
Round trips check the state arrays the SPH path uses, and corrupted files
are refused with the reference's error texts."""
import json
import os
import struct

import numpy as np
import pytest

from phantom_amd import dumpfile
from tests.golden import dump_states

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def vectors():
    with open(os.path.join(HERE, "golden", "dump_vectors.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", list(dump_states.STATES))
def test_reference_reader_view(name, vectors, tmp_path):
    ref = vectors[name]
    s = dump_states.STATES[name]()
    path = str(tmp_path / "dump_00000")
    dump_states.write(path, s)
    d = dumpfile.read_dump(path)
    assert d["fileid"].strip() == ref["desc"]
    h = d["header"]
    # the reference reader's sections: default int, int*8, default real, real*8
    for sec, k in (("part_numbers", dumpfile.I_INT), ("part_numbers_long", dumpfile.I_INT8),
                   ("quantities", dumpfile.I_REAL), ("units", dumpfile.I_REAL8)):
        ours = {}
        for t, v in h.entries[k]:
            ours.setdefault(t, []).append(v)
        assert set(ours) == set(ref[sec]), sec
        for t, want in ref[sec].items():
            got = ours[t]
            np.testing.assert_array_equal(np.array(got if len(got) > 1 else got[0]), np.array(want), err_msg=t)
    assert len(d["blocks"]) == len(ref["blocks"])
    for b, rb in zip(d["blocks"], ref["blocks"]):
        assert set(b) == set(rb["data"])
        for t, want in rb["data"].items():
            np.testing.assert_array_equal(b[t], np.array(want, dtype=b[t].dtype), err_msg=t)
        assert all(a.size == rb["dim"] for a in b.values())


@pytest.mark.parametrize("name", list(dump_states.STATES))
def test_state_round_trip(name, tmp_path):
    s = dump_states.STATES[name]()
    p = s["params"]
    path = str(tmp_path / "dump_00001")
    dump_states.write(path, s)
    st = dumpfile.state_from_dump(dumpfile.read_dump(path), maxvxyzu=p.maxvxyzu)
    np.testing.assert_array_equal(st["xyzh"][:, :3], s["xyzh"][:, :3])
    # h is stored as real*4 (readwrite_dumps.f90:218)
    np.testing.assert_array_equal(st["xyzh"][:, 3], s["xyzh"][:, 3].astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(st["vxyzu"], s["vxyzu"])
    np.testing.assert_array_equal(st["alphaind"][:, 0], s["alphaind"][:, 0])
    assert st["time"] == s["time"]
    assert st["hfact"] == p.hfact and st["gamma"] == p.gamma and st["massoftype_gas"] == p.massoftype_gas
    assert st["polyk"] == pytest.approx(p.polyk, rel=1e-15)
    if p.mhd:
        # B = Bevol*rho(h) is written; dividing by rho of the real*4 h
        # recovers Bevol up to the rounding of h
        rel = np.abs(st["Bevol"][:, :3] - s["Bevol"][:, :3]) / np.maximum(np.abs(s["Bevol"][:, :3]), 1e-300)
        assert rel.max() < 1e-6
        np.testing.assert_array_equal(st["Bevol"][:, 3], s["Bevol"][:, 3])
    if p.use_dustfrac:
        np.testing.assert_array_equal(st["dustfrac"], s["dustfrac"])


def test_corrupt_files_refused(tmp_path):
    s = dump_states.hydro_box()
    path = str(tmp_path / "dump_00002")
    dump_states.write(path, s)
    raw = open(path, "rb").read()
    # truncated: end of file in the data
    bad = str(tmp_path / "trunc")
    open(bad, "wb").write(raw[:len(raw) - 100])
    with pytest.raises(dumpfile.DumpError, match="end of file|record lengths"):
        dumpfile.read_dump(bad)
    # the magic number of another byte order
    swapped = bytearray(raw)
    swapped[4:8] = struct.pack(">i", dumpfile.INT1)
    bad2 = str(tmp_path / "endian")
    open(bad2, "wb").write(bytes(swapped))
    with pytest.raises(dumpfile.DumpError, match="wrong endian"):
        dumpfile.read_dump(bad2)
    # single-precision default real: the first record is 20 bytes
    bad3 = str(tmp_path / "real4")
    rec = struct.pack("<ifiii", dumpfile.INT1, float(dumpfile.INT2), dumpfile.INT2, dumpfile.IVERSION, dumpfile.INT1O)
    open(bad3, "wb").write(struct.pack("<i", len(rec)) + rec + struct.pack("<i", len(rec)) + raw[32:])
    with pytest.raises(dumpfile.DumpError, match="real size"):
        dumpfile.read_dump(bad3)
