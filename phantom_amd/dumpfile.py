"""Phantom full dump files for the particle state of the SPH path.

The format is the tagged binary format that write_fulldump and read_dump
use (readwrite_dumps.f90:45-334, 477-716), built from the dump_utils
routines (utils_dumpfiles.f90): Fortran unformatted sequential records,
each framed by its length in bytes as a 4-byte integer before and after,
little-endian, default integer 4 bytes and default real 8 bytes (the
reference's double-precision build).  In order:

1. int1, real(int2), int2, iversion, int1o (open_dumpfile_w,
   utils_dumpfiles.f90:1111-1139), then the 100-character file
   identification (fileident, readwrite_dumps_common.f90:32-69);
2. the header (write_header, utils_dumpfiles.f90:1502-1605): for each of
   the eight data types (default integer, int*1, int*2, int*4, int*8,
   default real, real*4, real*8) a record with the number of entries and,
   when non-zero, a record of 16-character tags and a record of values;
   the entries are those of fill_header (readwrite_dumps_common.f90:261-390);
3. the number of array blocks, narraylengths * nblocks (2, or 4 with MHD);
4. per array block: the array length (int*8) and the number of arrays of
   each data type (write_block_header, utils_dumpfiles.f90:1980-1991);
5. per array block and data type, per array: a tag record and a value
   record (write_array, utils_dumpfiles.f90:1612-1973).

Block 1 holds the hydro arrays in write_fulldump's order
(readwrite_dumps.f90:172-262); block 2 the sink particles (none here);
block 4, with MHD, B = Bevol*rho, psi and divcurlB.  read_dump returns
every array it finds, by block and tag, so files written by the reference
with other arrays are read as well.
"""
import struct

import numpy as np

INT1, INT2 = 60769, 60878        # utils_dumpfiles.f90:24
INT1O, INT2O = 690706, 780806    # utils_dumpfiles.f90:25
IVERSION = 1                     # utils_dumpfiles.f90:21
LENTAG = 16                      # utils_dumpfiles.f90:14
LENID = 100                      # utils_dumpfiles.f90:15
# data types in file order (utils_dumpfiles.f90:28-36)
I_INT, I_INT1, I_INT2, I_INT4, I_INT8, I_REAL, I_REAL4, I_REAL8 = range(8)
DTYPES = ["<i4", "<i1", "<i2", "<i4", "<i8", "<f8", "<f4", "<f8"]

ERR_ENDIAN, ERR_VERSION, ERR_REALSIZE, ERR_INTSIZE = 2, 3, 4, 5  # utils_dumpfiles.f90:40-48


class DumpError(IOError):
    pass


# ---------------------------------------------------------------- records
class _Writer:
    def __init__(self, f):
        self.f = f

    def record(self, payload):
        n = len(payload)
        self.f.write(struct.pack("<i", n))
        self.f.write(payload)
        self.f.write(struct.pack("<i", n))

    def array(self, a, dtype):
        self.record(np.ascontiguousarray(a, dtype=dtype).tobytes())


class _Reader:
    def __init__(self, f):
        self.f = f

    def record(self):
        head = self.f.read(4)
        if len(head) < 4:
            raise DumpError("end of file reached in data read")
        (n,) = struct.unpack("<i", head)
        payload = self.f.read(n)
        tail = self.f.read(4)
        if len(payload) != n or len(tail) < 4:
            raise DumpError("end of file reached in data read")
        (n2,) = struct.unpack("<i", tail)
        if n2 != n:
            raise DumpError("inconsistent Fortran record lengths")
        return payload

    def array(self, dtype):
        return np.frombuffer(self.record(), dtype=dtype).copy()


def tag(label):
    """A 16-character tag, blank padded (tag, utils_dumpfiles.f90:264-270)."""
    return label[:LENTAG].ljust(LENTAG).encode("ascii")


def fileident(mhd=False, gravity=False, onefluid_dust=False, version="2025.0.0", gitsha="00000000",
              date="01/01/2025", time="00:00:00.000"):
    """FT:Phantom:<version>:<gitsha> (hydro|mhd+clean<options>): <date> <time>
    (readwrite_dumps_common.f90:32-69)."""
    opts = ""
    if gravity:
        opts += "+grav"
    if onefluid_dust:
        opts += "+1dust"
    s = "FT:Phantom:" + version + ":" + gitsha
    if mhd:
        s += " (mhd+clean" + opts + ")  : " + date + " " + time
    else:
        s += " (hydro" + opts + "): " + date + " " + time
    return s[:LENID].ljust(LENID)


# ---------------------------------------------------------------- header
class Header:
    """Tagged header entries by data type: entries[k] is a list of
    (tag, value); a tag repeated over consecutive entries is an array
    (add_to_header of an array, utils_dumpfiles.f90:866-909)."""

    def __init__(self):
        self.entries = [[] for _ in range(8)]

    def add(self, k, label, value):
        vals = np.atleast_1d(np.asarray(value))
        for v in vals:
            self.entries[k].append((label, v))

    def get(self, label, default=None):
        """A value (scalar) or the array of values carrying a tag, from the
        first data type that has it."""
        for ent in self.entries:
            vals = [v for t, v in ent if t == label]
            if vals:
                return vals[0] if len(vals) == 1 else np.array(vals)
        return default

    def write(self, w):
        for k in range(8):
            ent = self.entries[k]
            w.array([len(ent)], "<i4")
            if ent:
                w.record(b"".join(tag(t) for t, _ in ent))
                w.array([v for _, v in ent], DTYPES[k])

    @classmethod
    def read(cls, r):
        h = cls()
        for k in range(8):
            (n,) = r.array("<i4")
            if n > 0:
                raw = r.record()
                tags = [raw[i * LENTAG:(i + 1) * LENTAG].decode("ascii").strip() for i in range(n)]
                vals = r.array(DTYPES[k])
                h.entries[k] = list(zip(tags, vals))
        return h


def _fill_header(p, n, time, maxtypes, nblocks=1, extra=None):
    """The entries fill_header writes for a gas-only run
    (readwrite_dumps_common.f90:261-390)."""
    h = Header()
    npartoftype = np.zeros(maxtypes, np.int64)
    npartoftype[0] = n
    ndust = 1 if (p.use_dustfrac and p.ndustsmall > 0) else 0
    h.add(I_INT, "nparttot", n)
    h.add(I_INT, "ntypes", maxtypes)
    h.add(I_INT, "npartoftype", npartoftype)
    h.add(I_INT, "nblocks", nblocks)
    h.add(I_INT, "nptmass", 0)
    h.add(I_INT, "ndustlarge", 0)
    h.add(I_INT, "ndustsmall", ndust)
    h.add(I_INT, "idust", 2)
    h.add(I_INT, "idtmax_n", 1)
    h.add(I_INT, "idtmax_frac", 0)
    h.add(I_INT, "idumpfile", 0)
    h.add(I_INT, "majorv", 2025)
    h.add(I_INT, "minorv", 0)
    h.add(I_INT, "microv", 0)
    h.add(I_INT8, "nparttot", n)
    h.add(I_INT8, "ntypes", maxtypes)
    h.add(I_INT8, "npartoftype", npartoftype)
    h.add(I_INT4, "iexternalforce", 0)
    h.add(I_INT4, "ieos", p.ieos)
    # write_headeropts_eos (eos.f90:1690-1711)
    h.add(I_INT, "isink", 0)
    h.add(I_REAL, "gamma", p.gamma)
    h.add(I_REAL, "RK2", 1.5 * p.polyk)
    h.add(I_REAL, "polyk2", 0.0)
    h.add(I_REAL, "qfacdisc", 0.75)
    h.add(I_REAL, "qfacdisc2", 0.75)
    h.add(I_REAL, "cs_min", 0.0)
    h.add(I_REAL, "time", time)
    h.add(I_REAL, "dtmax", p.dtmax)
    h.add(I_REAL, "rhozero", 0.0)
    h.add(I_REAL, "hfact", p.hfact)
    h.add(I_REAL, "tolh", p.tolh)
    h.add(I_REAL, "C_cour", p.C_cour)
    h.add(I_REAL, "C_force", p.C_force)
    h.add(I_REAL, "alpha", p.alpha)
    h.add(I_REAL, "alphau", p.alphau)
    h.add(I_REAL, "alphaB", p.alphaB)
    massoftype = np.zeros(maxtypes)
    massoftype[0] = p.massoftype_gas
    h.add(I_REAL, "massoftype", massoftype)
    h.add(I_REAL, "Bextx", 0.0)
    h.add(I_REAL, "Bexty", 0.0)
    h.add(I_REAL, "Bextz", 0.0)
    h.add(I_REAL, "dum", 0.0)
    if p.periodic:
        for t, v in zip(("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"),
                        (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax)):
            h.add(I_REAL, t, v)
    h.add(I_REAL, "get_conserv", -1.0)
    h.add(I_REAL, "mtot_in", 0.0)
    h.add(I_REAL, "etot_in", 0.0)
    h.add(I_REAL, "angtot_in", 0.0)
    h.add(I_REAL, "totmom_in", 0.0)
    h.add(I_REAL, "mdust_in", np.zeros(max(ndust, 1)))
    if ndust:
        h.add(I_REAL, "grainsize", p.grainsize)
        h.add(I_REAL, "graindens", p.graindens)
    units = dict(udist=1.0, umass=1.0, utime=1.0, umagfd=1.0)
    if extra:
        units.update({k: v for k, v in extra.items() if k in units})
    for t in ("udist", "umass", "utime", "umagfd"):
        h.add(I_REAL8, t, units[t])
    return h


# ---------------------------------------------------------------- write
def write_fulldump(path, params, xyzh, vxyzu, time=0.0, alphaind=None, divcurlv=None, Bevol=None, divcurlB=None,
                   divBsymm=None, dustfrac=None, tstop=None, deltav=None, poten=None, ibin=None, iorig=None,
                   maxtypes=8, ident=None, units=None):
    """Write a full dump of the particle state.  Arrays in the reference's
    storage layout as numpy arrays of shape (n, ncomp); Bevol (n, 4) is
    written as B = Bevol*rho (rho from h) and psi (readwrite_dumps.f90:
    305-313); ibin as dt = dtmax/2**ibin (real*4, :222-226)."""
    p = params
    n = xyzh.shape[0]
    mhd = bool(p.mhd) and Bevol is not None
    dust = bool(p.use_dustfrac and p.ndustsmall > 0) and dustfrac is not None
    if iorig is None:
        iorig = np.arange(1, n + 1, dtype=np.int64)
    # block 1 arrays by data type, in write_fulldump's call order
    b1 = [[] for _ in range(8)]
    for c, t in enumerate("xyz"):
        b1[I_REAL].append((t, xyzh[:, c]))
    if dust:
        b1[I_REAL].append(("dustfrac", np.asarray(dustfrac).reshape(n)))
        if tstop is not None:
            b1[I_REAL].append(("tstop", np.asarray(tstop).reshape(n)))
        if deltav is not None:
            for c, t in enumerate(("deltavx", "deltavy", "deltavz")):
                b1[I_REAL].append((t, deltav[:, c]))
    for c, t in enumerate(("vx", "vy", "vz", "u")[:vxyzu.shape[1]]):
        b1[I_REAL].append((t, vxyzu[:, c]))
    b1[I_REAL4].append(("h", xyzh[:, 3]))
    if alphaind is not None:
        b1[I_REAL4].append(("alpha", alphaind[:, 0]))
    if divcurlv is not None:
        dcv = divcurlv.reshape(n, -1)
        for c, t in enumerate(("divv", "curlvx", "curlvy", "curlvz")[:dcv.shape[1]]):
            b1[I_REAL4].append((t, dcv[:, c]))
    if p.gravity and poten is not None:
        b1[I_REAL4].append(("poten", poten))
    if p.ind_timesteps and ibin is not None:
        b1[I_REAL4].append(("dt", p.dtmax / 2.0 ** np.asarray(ibin, np.float64)))
    b1[I_INT8].append(("iorig", iorig))
    blocks = [(n, b1), (0, [[] for _ in range(8)])]
    if mhd:
        b4 = [[] for _ in range(8)]
        rho = p.massoftype_gas * (p.hfact / np.abs(xyzh[:, 3])) ** 3
        for c, t in enumerate(("Bx", "By", "Bz")):
            b4[I_REAL].append((t, Bevol[:, c] * rho))
        b4[I_REAL].append(("psi", Bevol[:, 3]))
        if divcurlB is not None:
            for c, t in enumerate(("divB", "curlBx", "curlBy", "curlBz")):
                b4[I_REAL4].append((t, divcurlB[:, c]))
        elif divBsymm is not None:
            b4[I_REAL4].append(("divBsymm", divBsymm))
        blocks += [(0, [[] for _ in range(8)]), (n, b4)]
    hdr = _fill_header(p, n, time, maxtypes, extra=units)
    ident = ident or fileident(mhd=mhd, gravity=bool(p.gravity), onefluid_dust=dust)
    with open(path, "wb") as f:
        w = _Writer(f)
        w.record(struct.pack("<idiii", INT1, float(INT2), INT2, IVERSION, INT1O))
        w.record(ident[:LENID].ljust(LENID).encode("ascii"))
        hdr.write(w)
        w.array([len(blocks)], "<i4")  # narraylengths * nblocks, nblocks = 1
        for length, arrs in blocks:
            w.record(struct.pack("<q", length) + struct.pack("<8i", *[len(a) for a in arrs]))
        for length, arrs in blocks:
            for k in range(8):
                for t, a in arrs[k]:
                    w.record(tag(t))
                    w.array(np.asarray(a).reshape(length), DTYPES[k])


# ---------------------------------------------------------------- read
def read_dump(path):
    """Read a full dump: returns dict(fileid, header (a Header), blocks (a
    list of dicts tag -> numpy array, one per array block), nparttot)."""
    with open(path, "rb") as f:
        r = _Reader(f)
        first = r.record()
        if len(first) != 24:
            raise DumpError("default real size wrong")
        int1i, r1, int2i, iver, int3i = struct.unpack("<idiii", first)
        if int1i != INT1 and int1i != INT1O:
            raise DumpError("wrong endian?")
        if iver == INT1O:
            iver = 0
        if iver > IVERSION:
            raise DumpError("file format version newer than current code can read")
        fileid = r.record().decode("ascii")
        if int2i != INT2 and int2i != INT2O:
            raise DumpError("default real size wrong")
        if int3i != INT1O:
            raise DumpError("default int size wrong")
        if fileid[1] not in "Tt":
            raise DumpError("routine requires tagged format but not detected")
        hdr = Header.read(r)
        (number,) = r.array("<i4")
        nblocks = int(hdr.get("nblocks", 1))
        narraylengths = number // nblocks
        if narraylengths < 1 or narraylengths > 4:
            raise DumpError(f"error 7 in readdump, narraylengths={narraylengths}")
        out = [dict() for _ in range(narraylengths)]
        for _ in range(nblocks):
            heads = []
            for _j in range(narraylengths):
                raw = r.record()
                (length,) = struct.unpack("<q", raw[:8])
                nums = struct.unpack("<8i", raw[8:40])
                heads.append((length, nums))
            for j, (length, nums) in enumerate(heads):
                for k in range(8):
                    for _i in range(nums[k]):
                        t = r.record().decode("ascii").strip()
                        a = r.array(DTYPES[k])
                        if a.size != length:
                            raise DumpError(f"array {t}: {a.size} values for a block of length {length}")
                        out[j][t] = np.concatenate([out[j][t], a]) if t in out[j] else a
    return dict(fileid=fileid, header=hdr, blocks=out, nparttot=int(hdr.get("nparttot", 0)))


def state_from_dump(d, maxvxyzu=4):
    """The hot path's arrays from a read dump: xyzh (n,4) with h from the
    real*4 h, vxyzu (n,maxvxyzu), alphaind (n,3) float32 (alpha only) and,
    where present, Bevol (n,4) = (B/rho, psi), dustfrac and time."""
    b1 = d["blocks"][0]
    hdr = d["header"]
    n = b1["x"].shape[0]
    xyzh = np.stack([b1["x"], b1["y"], b1["z"], b1["h"].astype(np.float64)], axis=1)
    vxyzu = np.stack([b1[t] for t in ("vx", "vy", "vz", "u")[:maxvxyzu]], axis=1)
    alphaind = np.zeros((n, 3), np.float32)
    if "alpha" in b1:
        alphaind[:, 0] = b1["alpha"]
    out = dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, time=float(hdr.get("time", 0.0)),
               hfact=float(hdr.get("hfact")), gamma=float(hdr.get("gamma")),
               polyk=2.0 / 3.0 * float(hdr.get("RK2")),  # read_headeropts_eos (eos.f90:1727-1729)
               massoftype_gas=float(np.atleast_1d(hdr.get("massoftype"))[0]))
    if len(d["blocks"]) >= 4 and "Bx" in d["blocks"][3]:
        b4 = d["blocks"][3]
        hfact = float(hdr.get("hfact"))
        pmass = float(np.atleast_1d(hdr.get("massoftype"))[0])
        rho = pmass * (hfact / np.abs(xyzh[:, 3])) ** 3
        out["Bevol"] = np.stack([b4["Bx"] / rho, b4["By"] / rho, b4["Bz"] / rho, b4["psi"]], axis=1)
    if "dustfrac" in b1:
        out["dustfrac"] = b1["dustfrac"].copy()
    return out
