"""Reader of legacy BINARY structured-grid VTK files, written independently
of the library's writer (tests only)."""
import numpy as np


def read_vtk(path):
    """Independent reader of a legacy BINARY structured-grid file."""
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        while data[pos:pos + 1] in (b"\n", b" "):
            pos += 1
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    def block(n, ty):
        nonlocal pos
        dt = np.dtype(">f4" if ty == "float" else ">f8")
        a = np.frombuffer(data, dtype=dt, count=n, offset=pos)
        pos += n * dt.itemsize
        return a.astype(np.float64)

    assert line() == "# vtk DataFile Version 2.0"
    line()
    assert line() == "BINARY"
    assert line() == "DATASET STRUCTURED_GRID"
    dims = tuple(int(v) for v in line().split()[1:])
    _, npts, ty = line().split()
    pts = block(3 * int(npts), ty).reshape(-1, 3)
    fields = {}
    if pos < len(data) and line().startswith("POINT_DATA"):
        while True:
            while pos < len(data) and data[pos:pos + 1] in (b"\n", b" "):
                pos += 1
            if pos >= len(data):
                break
            _, name, fty, _ = line().split()
            assert line() == "LOOKUP_TABLE default"
            fields[name] = block(int(npts), fty)
    return dims, pts, fields
