"""Two small seeded particle states written to full dumps for the dump
format vectors (make_dump_vectors.py) and tests/test_dumpfile.py."""
import numpy as np

from phantom_amd import abi, dumpfile, setups


def hydro_box():
    s = setups.uniform_box(nx=4, seed=21, isothermal=False)
    n = s["xyzh"].shape[0]
    rng = np.random.default_rng(5)
    s["divcurlv"] = rng.normal(0.0, 1.0, (n, 1)).astype(np.float32)
    s["time"] = 0.125
    return s


def mhd_dust_box():
    s = setups.mhd_dust_box(nx=4, seed=22, jitter=0.05)
    n = s["xyzh"].shape[0]
    rng = np.random.default_rng(6)
    s["dustfrac"] = s["dustevol"] ** 2 / (1.0 + s["dustevol"] ** 2)
    s["tstop"] = rng.uniform(0.01, 1.0, n)
    s["deltav"] = rng.normal(0.0, 1e-3, (n, 3))
    s["divcurlB"] = rng.normal(0.0, 1.0, (n, 4)).astype(np.float32)
    s["Bevol"][:, 3] = rng.normal(0.0, 0.1, n)
    s["time"] = 2.5
    return s


STATES = {"hydro_box": hydro_box, "mhd_dust_box": mhd_dust_box}


def write(path, s):
    p = s["params"]
    dumpfile.write_fulldump(path, p, s["xyzh"], s["vxyzu"], time=s["time"], alphaind=s["alphaind"],
                            divcurlv=s.get("divcurlv"), Bevol=s.get("Bevol"), divcurlB=s.get("divcurlB"),
                            dustfrac=s.get("dustfrac"), tstop=s.get("tstop"), deltav=s.get("deltav"),
                            ident=dumpfile.fileident(mhd=bool(p.mhd), onefluid_dust=bool(p.use_dustfrac)))
