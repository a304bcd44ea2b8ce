"""Minimal FITS primary-HDU reader/writer in numpy, independent of the
library's C implementation, used to cross-check it."""
import numpy as np

_DT = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def read(path):
    raw = open(path, "rb").read()
    cards = {}
    pos = 0
    end = False
    while not end:
        block = raw[pos:pos + 2880].decode("ascii")
        pos += 2880
        for i in range(36):
            c = block[80 * i:80 * i + 80]
            key = c[:8].strip()
            if key == "END":
                end = True
                break
            if c[8:10] == "= ":
                val = c[10:].split("/")[0].strip()
                cards[key] = val
    bitpix = int(cards["BITPIX"])
    naxis = int(cards["NAXIS"])
    shape = [int(cards[f"NAXIS{i + 1}"]) for i in range(naxis)]
    n = int(np.prod(shape)) if shape else 0
    a = np.frombuffer(raw, dtype=_DT[bitpix], count=n, offset=pos).astype(np.float64)
    if "BSCALE" in cards or "BZERO" in cards:
        a = a * float(cards.get("BSCALE", 1.0)) + float(cards.get("BZERO", 0.0))
    return a.reshape(shape[::-1]), cards


def write(path, arr, bitpix):
    arr = np.asarray(arr)
    cards = ["SIMPLE  = %20s" % "T", "BITPIX  = %20d" % bitpix, "NAXIS   = %20d" % arr.ndim]
    for i, s in enumerate(arr.shape[::-1]):
        cards.append("NAXIS%-3d= %20d" % (i + 1, s))
    cards.append("END")
    hdr = "".join(c.ljust(80) for c in cards)
    hdr = hdr.ljust((len(hdr) + 2879) // 2880 * 2880)
    data = arr.astype(_DT[bitpix]).tobytes()
    data = data + b"\0" * ((-len(data)) % 2880)
    with open(path, "wb") as f:
        f.write(hdr.encode("ascii") + data)
