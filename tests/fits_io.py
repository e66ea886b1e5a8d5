"""Small reader for the FITS binary tables the library writes (tests only).

Reads the first BINTABLE extension: header keywords and every column as a
flat numpy array in element order (row-major over the table, so a full-sky
HEALPix column comes back in pixel order, as healpy's read_map(nest=True)
returns it).  Stands in for healpy, which is not installed here.
"""
import numpy as np

_DT = {"E": ">f4", "D": ">f8", "J": ">i4", "K": ">i8"}


def _cards(block):
    for i in range(0, len(block), 80):
        yield block[i:i + 80].decode("ascii")


def _parse_value(v):
    v = v.strip()
    if v.startswith("'"):
        s, i = "", 1
        while i < len(v):
            if v[i] == "'":
                if i + 1 < len(v) and v[i + 1] == "'":
                    s += "'"
                    i += 2
                    continue
                break
            s += v[i]
            i += 1
        return s.rstrip()
    v = v.split("/")[0].strip()
    if v == "T":
        return True
    if v == "F":
        return False
    try:
        return int(v)
    except ValueError:
        return float(v)


def _read_header(f):
    hdr, comments = {}, []
    while True:
        block = f.read(2880)
        if len(block) < 2880:
            raise ValueError("truncated FITS header")
        done = False
        for c in _cards(block):
            key = c[:8].strip()
            if key == "END":
                done = True
                break
            if key == "COMMENT":
                comments.append(c[8:].rstrip())
            elif c.startswith("HIERARCH "):
                k, v = c[9:].split("=", 1)
                hdr[k.strip()] = _parse_value(v)
            elif c[8:10] == "= ":
                hdr[key] = _parse_value(c[10:])
        if done:
            return hdr, comments


def read_bintable(path):
    """Returns (header dict, comment lines, {column name: flat array})."""
    with open(path, "rb") as f:
        primary, _ = _read_header(f)
        assert primary.get("SIMPLE") is True
        hdr, comments = _read_header(f)
        assert hdr["XTENSION"] == "BINTABLE"
        ncol, nrow, rowb = hdr["TFIELDS"], hdr["NAXIS2"], hdr["NAXIS1"]
        fields = []
        for i in range(1, ncol + 1):
            form = hdr["TFORM%d" % i]
            rep, t = int(form[:-1] or 1), form[-1]
            fields.append((hdr["TTYPE%d" % i], rep, _DT[t]))
        data = np.frombuffer(f.read(nrow * rowb), dtype=np.uint8)
        assert data.size == nrow * rowb
        rows = data.reshape(nrow, rowb)
        cols, off = {}, 0
        for name, rep, dt in fields:
            size = np.dtype(dt).itemsize * rep
            cols[name] = rows[:, off:off + size].copy().view(dt).reshape(-1).astype(dt[1:])
            off += size
        return hdr, comments, cols


def read_map(path, field=0):
    """Like healpy.read_map(path, nest=True, field=field) for these files."""
    hdr, _, cols = read_bintable(path)
    names = list(cols)
    v = cols[names[field]].astype(np.float64)
    v[v == np.float32(-1.6375e30)] = -1.6375e30
    return v
