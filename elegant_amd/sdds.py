"""SDDS-lite: reading and writing SDDS version 1 files (ASCII and binary,
row- or column-major, little-endian) with parameters and columns, and the
particle files elegant exchanges on the tracked path:

* phase-space output (dump_phase_space, sdds_support.c:1585-1589): columns
  x, xp, y, yp, t = s/(c beta), p = Po(1+delta), particleID; parameters Step,
  pCentral, Charge, Particles (sdds_support.c:150-170);
* lost particles (dump_lost_particles, sdds_support.c:1706-1712): x, xp, y,
  yp, s, p, particleID, Pass (sdds_support.c:199-213);
* wake Green functions (columns t and W, as doc/examples/LCLS/wakes/*.sdds)
  for WAKE elements.

Arrays, associates and compressed data are outside this subset.
"""
import math
import re
import struct

import numpy as np

C_MKS = 2.99792458e8

_TYPES = {
    "double": ("d", np.float64), "float": ("f", np.float32), "long": ("i", np.int32),
    "ulong": ("I", np.uint32), "long64": ("q", np.int64), "ulong64": ("Q", np.uint64),
    "short": ("h", np.int16), "ushort": ("H", np.uint16), "character": ("c", "S1"), "string": (None, object),
}


class Page:
    def __init__(self, parameters=None, columns=None):
        self.parameters = dict(parameters or {})   # name -> value
        self.columns = dict(columns or {})         # name -> numpy array

    def rows(self):
        return len(next(iter(self.columns.values()))) if self.columns else 0


class SDDS:
    """Layout (ordered definitions) plus pages of data."""

    def __init__(self):
        self.description = None
        self.parameters = []   # [(name, type, dict of other fields)]
        self.columns = []
        self.pages = []

    def define_parameter(self, name, type_="double", **fields):
        self.parameters.append((name, type_, fields))

    def define_column(self, name, type_="double", **fields):
        self.columns.append((name, type_, fields))


def _namelist(text):
    """fields of one '&kind a=b, c="d e", &end' definition"""
    out = {}
    for m in re.finditer(r'(\w+)\s*=\s*("(?:[^"\\]|\\.)*"|[^,\s&]+)', text):
        v = m.group(2)
        if v.startswith('"'):
            v = v[1:-1].replace('\\"', '"')
        out[m.group(1).lower()] = v
    return out


def _split_ascii(line):
    toks = []
    for m in re.finditer(r'"((?:[^"\\]|\\.)*)"|(\S+)', line):
        toks.append(m.group(1).replace('\\"', '"') if m.group(1) is not None else m.group(2))
    return toks


def read(path):
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    header = []
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].decode("latin-1").rstrip("\r")
        pos = end + 1
        header.append(line)
        if line.lstrip().lower().startswith("&data"):
            # a definition may continue to its &end
            while "&end" not in line.lower():
                end = data.index(b"\n", pos)
                line += " " + data[pos:end].decode("latin-1").rstrip("\r")
                pos = end + 1
            break
    if not header[0].startswith("SDDS1"):
        raise ValueError("%s: not an SDDS1 file" % path)
    s = SDDS()
    text = "\n".join(header[1:])
    datamode = {}
    for m in re.finditer(r"&(\w+)(.*?)&end", text, re.S | re.I):
        kind, body = m.group(1).lower(), _namelist(m.group(2))
        name = body.pop("name", None)
        if kind == "parameter":
            s.parameters.append((name, body.pop("type", "double"), body))
        elif kind == "column":
            s.columns.append((name, body.pop("type", "double"), body))
        elif kind == "description":
            s.description = body
        elif kind == "data":
            datamode = body
        elif kind in ("array", "associate"):
            raise ValueError("%s: SDDS %s definitions are not supported" % (path, kind))
    mode = datamode.get("mode", "binary").lower()
    if mode == "ascii":
        _read_ascii(s, data[pos:].decode("latin-1"), int(datamode.get("no_row_counts", "0")))
    else:
        _read_binary(s, data, pos, int(datamode.get("column_major_order", "0")))
    return s


def _convert(type_, tok):
    if type_ == "string":
        return tok
    if type_ == "character":
        return tok[:1]
    if type_ in ("double", "float"):
        return float(tok)
    return int(tok)


def _read_ascii(s, text, no_row_counts):
    lines = [l for l in text.replace("\r", "").split("\n") if l.strip() and not l.lstrip().startswith("!")]
    i = 0
    while i < len(lines):
        page = Page()
        for name, type_, fields in s.parameters:
            if "fixed_value" in fields:
                page.parameters[name] = _convert(type_, fields["fixed_value"])
                continue
            tok = lines[i].strip()
            i += 1
            if type_ == "string" and tok.startswith('"'):
                tok = _split_ascii(tok)[0]
            page.parameters[name] = _convert(type_, tok)
        if s.columns:
            if no_row_counts:
                raise ValueError("no_row_counts ASCII data is not supported")
            n = int(lines[i].split()[0])
            i += 1
            cols = [[] for _ in s.columns]
            got = 0
            toks = []
            while got < n:
                toks += _split_ascii(lines[i])
                i += 1
                while len(toks) >= len(s.columns) and got < n:
                    for k, (name, type_, _) in enumerate(s.columns):
                        cols[k].append(_convert(type_, toks[k]))
                    toks = toks[len(s.columns):]
                    got += 1
            for k, (name, type_, _) in enumerate(s.columns):
                page.columns[name] = np.array(cols[k], dtype=_TYPES[type_][1])
        s.pages.append(page)


def _read_binary(s, data, pos, column_major):
    def take(fmt):
        nonlocal pos
        v = struct.unpack_from("<" + fmt, data, pos)
        pos += struct.calcsize("<" + fmt)
        return v[0]

    def take_string():
        nonlocal pos
        n = take("i")
        v = data[pos:pos + n].decode("latin-1")
        pos += n
        return v

    while pos < len(data):
        page = Page()
        n = take("i")
        for name, type_, fields in s.parameters:
            if "fixed_value" in fields:
                page.parameters[name] = _convert(type_, fields["fixed_value"])
            elif type_ == "string":
                page.parameters[name] = take_string()
            else:
                page.parameters[name] = take(_TYPES[type_][0])
        if column_major:
            for name, type_, _ in s.columns:
                if type_ == "string":
                    page.columns[name] = np.array([take_string() for _ in range(n)], dtype=object)
                else:
                    fmt, dt = _TYPES[type_]
                    page.columns[name] = np.frombuffer(data, dtype=np.dtype(dt).newbyteorder("<"), count=n,
                                                       offset=pos).copy()
                    pos += n * np.dtype(dt).itemsize
        else:
            cols = {name: [] for name, _, _ in s.columns}
            for _ in range(n):
                for name, type_, _ in s.columns:
                    cols[name].append(take_string() if type_ == "string" else take(_TYPES[type_][0]))
            for name, type_, _ in s.columns:
                page.columns[name] = np.array(cols[name], dtype=_TYPES[type_][1])
        s.pages.append(page)


def _fmt_def(kind, name, type_, fields):
    extra = "".join(', %s="%s"' % (k, v) if " " in str(v) or "," in str(v) else ", %s=%s" % (k, v)
                    for k, v in fields.items())
    return "&%s name=%s, type=%s%s, &end" % (kind, name, type_, extra)


def write(path, s, mode="ascii", column_major=False):
    out = ["SDDS1"]
    if s.description:
        out.append("&description " + ", ".join('%s="%s"' % kv for kv in s.description.items()) + ", &end")
    out += [_fmt_def("parameter", n, t, f) for n, t, f in s.parameters]
    out += [_fmt_def("column", n, t, f) for n, t, f in s.columns]
    out.append("&data mode=%s%s, &end" % (mode, ", column_major_order=1" if column_major and mode == "binary" else ""))
    head = ("\n".join(out) + "\n").encode("latin-1")
    with open(path, "wb") as f:
        f.write(head)
        for k, page in enumerate(s.pages):
            n = page.rows()
            if mode == "ascii":
                lines = ["! page number %d" % (k + 1)]
                for name, type_, fields in s.parameters:
                    if "fixed_value" in fields:
                        continue
                    v = page.parameters[name]
                    lines.append('"%s"' % v if type_ == "string" else
                                 (repr(float(v)) if type_ in ("double", "float") else str(int(v))))
                if s.columns:
                    lines.append("%d" % n)
                    cols = [page.columns[name] for name, _, _ in s.columns]
                    for i in range(n):
                        lines.append(" ".join('"%s"' % c[i] if t == "string" else
                                              (repr(float(c[i])) if t in ("double", "float") else str(int(c[i])))
                                              for c, (_, t, _) in zip(cols, s.columns)))
                f.write(("\n".join(lines) + "\n").encode("latin-1"))
            else:
                f.write(struct.pack("<i", n))
                for name, type_, fields in s.parameters:
                    if "fixed_value" in fields:
                        continue
                    v = page.parameters[name]
                    if type_ == "string":
                        b = str(v).encode("latin-1")
                        f.write(struct.pack("<i", len(b)) + b)
                    else:
                        f.write(struct.pack("<" + _TYPES[type_][0], v))
                if column_major:
                    for name, type_, _ in s.columns:
                        f.write(np.ascontiguousarray(page.columns[name], dtype=np.dtype(_TYPES[type_][1])
                                                     .newbyteorder("<")).tobytes())
                else:
                    rec = np.zeros(n, dtype=[(name, np.dtype(_TYPES[t][1]).newbyteorder("<"))
                                             for name, t, _ in s.columns])
                    for name, _, _ in s.columns:
                        rec[name] = page.columns[name]
                    f.write(rec.tobytes())


# ---------------------------------------------------------------- particles

def phase_space(rows, Po, step=1, charge=0.0):
    """an SDDS holding rows (n, >=7) as elegant's phase-space output"""
    s = SDDS()
    for n, t in (("Step", "long"), ("pCentral", "double"), ("Charge", "double"), ("Particles", "long")):
        s.define_parameter(n, t)
    for n, t in (("x", "double"), ("xp", "double"), ("y", "double"), ("yp", "double"), ("t", "double"),
                 ("p", "double"), ("particleID", "ulong64")):
        s.define_column(n, t)
    p = Po * (1 + rows[:, 5])
    cols = {"x": rows[:, 0], "xp": rows[:, 1], "y": rows[:, 2], "yp": rows[:, 3],
            "t": rows[:, 4] / (C_MKS * p / np.sqrt(p * p + 1)), "p": p,
            "particleID": rows[:, 6].astype(np.uint64)}
    s.pages.append(Page({"Step": step, "pCentral": Po, "Charge": charge, "Particles": rows.shape[0]}, cols))
    return s


def rows_from_phase_space(page, Po):
    """rows (n, 8) from a phase-space page (delta = p/Po - 1, s = t c beta)"""
    p = np.asarray(page.columns["p"], dtype=np.float64)
    n = p.size
    rows = np.zeros((n, 8))
    rows[:, 0] = page.columns["x"]
    rows[:, 1] = page.columns["xp"]
    rows[:, 2] = page.columns["y"]
    rows[:, 3] = page.columns["yp"]
    rows[:, 4] = np.asarray(page.columns["t"]) * (C_MKS * p / np.sqrt(p * p + 1))
    rows[:, 5] = p / Po - 1
    rows[:, 6] = page.columns["particleID"]
    rows[:, 7] = -1
    return rows


def lost_particles(rows_lost, step=1):
    """an SDDS holding lost rows (s = loss position, delta = Po(1+delta),
    column 7 = pass) as elegant's loss output"""
    s = SDDS()
    s.define_parameter("Step", "long")
    for n, t in (("x", "double"), ("xp", "double"), ("y", "double"), ("yp", "double"), ("s", "double"),
                 ("p", "double"), ("particleID", "ulong64"), ("Pass", "long")):
        s.define_column(n, t)
    r = rows_lost
    s.pages.append(Page({"Step": step}, {"x": r[:, 0], "xp": r[:, 1], "y": r[:, 2], "yp": r[:, 3], "s": r[:, 4],
                                         "p": r[:, 5], "particleID": r[:, 6].astype(np.uint64),
                                         "Pass": r[:, 7].astype(np.int32)}))
    return s


def wake_table(path, tcol="t", wcol="W", page=0):
    """(t, W) arrays of a longitudinal wake Green function file"""
    s = read(path)
    pg = s.pages[page]
    return np.asarray(pg.columns[tcol], dtype=np.float64), np.asarray(pg.columns[wcol], dtype=np.float64)


def impedance_table(real=None, imag=None, page=0):
    """(f, Zreal, Zimag) for a ZLONGIT with BROAD_BAND=0 from elegant's
    ZREAL/ZIMAG file specifications, each "file=fcol+zcol" (the table
    reader's form, zlongit.c:640-643); a missing one gives None.  Both
    files must share their frequencies (count and spacing are checked by
    the reference, zlongit.c:665-674; here the values must agree)."""
    out, f = [], None
    for spec in (real, imag):
        if spec is None:
            out.append(None)
            continue
        path, _, cols = spec.partition("=")
        fcol, _, zcol = cols.partition("+")
        if not fcol or not zcol:
            raise ValueError("impedance file specification must be file=freqColumn+valueColumn: %r" % spec)
        pg = read(path).pages[page]
        fi = np.asarray(pg.columns[fcol], dtype=np.float64)
        if f is not None and not np.array_equal(f, fi):
            raise ValueError("ZREAL and ZIMAG tables have different frequencies")
        f = fi
        out.append(np.asarray(pg.columns[zcol], dtype=np.float64))
    if f is None:
        raise ValueError("give ZREAL and/or ZIMAG")
    return f, out[0], out[1]
