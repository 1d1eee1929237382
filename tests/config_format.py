"""Reader/writer of the reference's binary configuration file
(LGT_Code/modules/io/config_IO.c:40-117 writeConfig, :537-621 readConfig) for
the tests: little-endian int32 DIM, int32 SUN, int32 extents[DIM] (t first),
double plaquette, then per site (z fastest) and direction c0..c3 as doubles."""
import numpy as np

HEADER = np.dtype([("dim", "<i4"), ("sun", "<i4"), ("L", "<i4", (4,)), ("plaq", "<f8")])


def read(path):
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:HEADER.itemsize], dtype=HEADER)[0]
    L = tuple(int(x) for x in h["L"])
    vol = int(np.prod(L))
    U = np.frombuffer(raw[HEADER.itemsize:], dtype="<f8")
    assert U.size == 16 * vol, (U.size, vol)
    return int(h["dim"]), int(h["sun"]), L, float(h["plaq"]), U.reshape(vol, 4, 4)


def write(path, L, plaq, U, dim=4, sun=2):
    h = np.zeros(1, dtype=HEADER)
    h["dim"], h["sun"], h["L"], h["plaq"] = dim, sun, L, plaq
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(np.ascontiguousarray(U, dtype="<f8").tobytes())
