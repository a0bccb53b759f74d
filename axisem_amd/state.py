"""Reader/writer of axh state files (the uploads of include/axihip.h).

Format (little endian)::

    "AXHSTATE" | u32 version = 1 | u32 nrecords
    per record: char name[48] | u32 dtype (1 f32, 2 f64, 3 i32) | u32 ndim |
                u64 dims[4] | u64 nbytes | data

``dims_i32`` holds the 22 int32 fields of ``axh_dims`` and ``dims_f64`` its two
doubles (deltat, magnitude).  Arrays keep the reference's Fortran layout; they
are returned as flat numpy arrays plus their Fortran dims.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

_DT = {1: np.float32, 2: np.float64, 3: np.int32}
_DT_INV = {np.dtype(np.float32): 1, np.dtype(np.float64): 2, np.dtype(np.int32): 3}

DIMS_FIELDS = ("abi_version", "npol", "nel_solid", "nel_fluid", "nglob_solid",
               "nglob_fluid", "nel_bdry", "src_type", "scheme", "anel",
               "coarse_grained", "n_sls", "niter", "seis_it", "num_rec", "nelsrc",
               "fluid_src")

SRC_NAMES = {0: "monopole", 1: "dipole", 2: "quadpole"}
SCHEME_NAMES = {0: "newmark2", 1: "symplec4", 2: "ML_SO4m5", 3: "ML_SO6m7"}
NSTAGES = {0: 1, 1: 4, 2: 5, 3: 7}


@dataclass
class State:
    dims: dict
    arrays: dict = field(default_factory=dict)   # name -> (flat ndarray, fortran dims)

    def __getitem__(self, name):
        return self.arrays[name][0]

    def shape(self, name):
        return self.arrays[name][1]

    def get(self, name, default=None):
        a = self.arrays.get(name)
        return default if a is None else a[0]

    @property
    def nstages(self) -> int:
        return NSTAGES[self.dims["scheme"]]

    def gll_points(self) -> int:
        """Element-local GLL points advanced per step (25 (nel_solid+nel_fluid))."""
        return 25 * (self.dims["nel_solid"] + self.dims["nel_fluid"])


def read_state(path: str) -> State:
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != b"AXHSTATE":
        raise ValueError("not an axh state file: " + path)
    ver, nrec = struct.unpack_from("<II", buf, 8)
    if ver != 1:
        raise ValueError("unsupported state file version %d" % ver)
    off = 16
    arrays = {}
    dims = {}
    for _ in range(nrec):
        name = buf[off:off + 48].split(b"\0", 1)[0].decode()
        dt, nd = struct.unpack_from("<II", buf, off + 48)
        d4 = struct.unpack_from("<4Q", buf, off + 56)
        (nb,) = struct.unpack_from("<Q", buf, off + 88)
        off += 96
        a = np.frombuffer(buf, dtype=_DT[dt], count=nb // np.dtype(_DT[dt]).itemsize,
                          offset=off).copy()
        off += nb
        if name == "dims_i32":
            for k, v in zip(DIMS_FIELDS, a):
                dims[k] = int(v)
        elif name == "dims_f64":
            dims["deltat"], dims["magnitude"] = float(a[0]), float(a[1])
        else:
            arrays[name] = (a, tuple(int(x) for x in d4[:nd]))
    return State(dims, arrays)


def read_dims(path: str) -> dict:
    """Only the dims of a state file (skips the array payloads)."""
    dims = {}
    with open(path, "rb") as f:
        head = f.read(16)
        if head[:8] != b"AXHSTATE":
            raise ValueError("not an axh state file: " + path)
        _, nrec = struct.unpack_from("<II", head, 8)
        for _ in range(nrec):
            h = f.read(96)
            name = h[:48].split(b"\0", 1)[0].decode()
            dt, _nd = struct.unpack_from("<II", h, 48)
            (nb,) = struct.unpack_from("<Q", h, 88)
            if name == "dims_i32":
                a = np.frombuffer(f.read(nb), np.int32)
                for k, v in zip(DIMS_FIELDS, a):
                    dims[k] = int(v)
            elif name == "dims_f64":
                a = np.frombuffer(f.read(nb), np.float64)
                dims["deltat"], dims["magnitude"] = float(a[0]), float(a[1])
            else:
                f.seek(nb, 1)
            if len(dims) == len(DIMS_FIELDS) + 2:
                break
    return dims


def write_state(st: State, path: str) -> None:
    recs = []
    di = np.zeros(22, np.int32)
    for k, name in enumerate(DIMS_FIELDS):
        di[k] = st.dims[name]
    recs.append(("dims_i32", di, (22,)))
    recs.append(("dims_f64", np.array([st.dims["deltat"], st.dims["magnitude"]]), (2,)))
    for name, (a, fd) in st.arrays.items():
        recs.append((name, a, fd))
    with open(path, "wb") as f:
        f.write(b"AXHSTATE" + struct.pack("<II", 1, len(recs)))
        for name, a, fd in recs:
            a = np.ascontiguousarray(a)
            d4 = list(fd) + [0] * (4 - len(fd))
            f.write(name.encode().ljust(48, b"\0"))
            f.write(struct.pack("<II4QQ", _DT_INV[a.dtype], len(fd), *d4, a.nbytes))
            f.write(a.tobytes())
