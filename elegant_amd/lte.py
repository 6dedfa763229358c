"""Parser for the subset of elegant's lattice (.lte) language that the
tracked path needs (mad_parse.c, get_beamline.c): element definitions
"LABEL: TYPE, PARAM=value, ...", beamlines "LABEL: LINE=(...)" with
repetition "n*item", reversal "-item" and nesting, "&" continuations, "!"
comments, quoted strings and quoted RPN expressions, and USE.

Type names are matched as elegant's tell_type does (mad_parse.c:669-745):
the first min(len(token), len(name)) characters must agree with exactly one
entity name (so DRIFT, QUADRUPOLE and MARKER name DRIF, QUAD and MARK); an
earlier element's label may also serve as the type (a copy with
overrides).  A reversed line reverses its order and swaps the entrance and
exit edge parameters of bends and the edge flags of KQUAD
(mad_parse.c:564-571, 1397-1430).  Everything is upper-cased outside quotes.

Only the element types the GPU path tracks become elements:
DRIF, EDRIFT, KQUAD, KSEXT, KOCT, CSBEND, RFCA, WAKE, ZLONGIT, MAXAMP, MARK,
CHARGE.  Others raise, unless listed in `ignore` (dropped) or mapped by
`transmute` (as elegant's transmute_elements: e.g. {"QUAD": "KQUAD"}).
"""
import math
import os
import re

from . import _abi as A
from . import elements as E

# elegant's entity names, in its table order (track_data.c:67-202): the
# abbreviation rule depends on the whole list
ENTITY_NAMES = (
    "LINE", "QUAD", "SBEN", "RBEN", "DRIF", "SEXT", "OCTU", "MULT", "SOLE", "HKICK", "VKICK", "RFCA", "ELSE",
    "HMON", "VMON", "MONI", "RCOL", "ECOL", "MARK", "MATR", "ALPH", "RFDF", "RFTMEZ0", "RMDF", "TMCF", "CEPL",
    "WATCH", "TWPL", "MALIGN", "TWLA", "PEPPOT", "ENERGY", "MAXAMP", "ROTATE", "TRCOUNT", "RECIRC", "QUFRINGE",
    "SCRAPER", "CENTER", "BUMPER", "KSEXT", "KSBEND", "KQUAD", "MAGNIFY", "SAMPLE", "KICKER", "SCATTER",
    "NIBEND", "KPOLY", "NISEPT", "RAMPRF", "RAMPP", "STRAY", "CSBEND", "TWMTA", "MATTER", "RFMODE", "TRFMODE",
    "ZLONGIT", "SREFFECTS", "MODRF", "BMAPXY", "ZTRANSVERSE", "IBSCATTER", "FMULT", "WAKE", "TRWAKE", "TUBEND",
    "CHARGE", "PFILTER", "HISTOGRAM", "CSRCSBEND", "CSRDRIFT", "RFCW", "REMCOR", "MAPSOLENOID", "REFLECT",
    "CLEAN", "TWISS", "WIGGLER", "SCRIPT", "FLOOR", "LTHINLENS", "LMIRROR", "EMATRIX", "FRFMODE", "FTRFMODE",
    "TFBPICKUP", "TFBDRIVER", "LSCDRIFT", "DSCATTER", "LSRMDLTR", "POLYSERIES", "RFTM110", "CWIGGLER",
    "EDRIFT", "SCMULT", "ILMATRIX", "TSCATTER", "KQUSE", "UKICKMAP", "MBUMPER", "EMITTANCE", "MHISTOGRAM",
    "FTABLE", "KOCT", "RIMULT", "GFWIGGLER", "MRFDF", "CORGPIPE", "LRWAKE", "EHKICK", "EVKICK", "EKICKER",
    "BMXYZ", "BRAT", "BGGEXP", "BRANCH", "IONEFFECTS", "SLICE", "SPEEDBUMP", "CCBEND", "HKPOLY", "BOFFAXE",
    "APCONTOUR", "TAPERAPC", "TAPERAPE", "TAPERAPR", "SHRFDF", "GKICKMAP", "BEAMBEAM", "CPICKUP", "CKICKER",
    "LGBEND", "CORGPLATES",
)

SUPPORTED = {"DRIF": A.EH_T_DRIF, "EDRIFT": A.EH_T_EDRIFT, "KQUAD": A.EH_T_KQUAD, "KSEXT": A.EH_T_KSEXT,
             "KOCT": A.EH_T_KOCT, "CSBEND": A.EH_T_CSBEND, "RFCA": A.EH_T_RFCA, "WAKE": A.EH_T_WAKE,
             "ZLONGIT": A.EH_T_ZLONGIT, "MAXAMP": A.EH_T_MAXAMP, "MARK": A.EH_T_MARK, "CHARGE": A.EH_T_CHARGE,
             "MALIGN": A.EH_T_MALIGN, "MONI": A.EH_T_MONI, "HMON": A.EH_T_HMON, "VMON": A.EH_T_VMON,
             "HKICK": A.EH_T_HKICK, "VKICK": A.EH_T_VKICK, "RECIRC": A.EH_T_RECIRC,
             "RCOL": A.EH_T_RCOL}

_RPN_UNARY = {"sqrt": math.sqrt, "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin,
              "acos": math.acos, "atan": math.atan, "exp": math.exp, "ln": math.log, "abs": abs,
              "chs": lambda a: -a, "sqr": lambda a: a * a}
_RPN_BINARY = {"+": lambda a, b: a + b, "-": lambda a, b: a - b, "*": lambda a, b: a * b,
               "/": lambda a, b: a / b, "pow": math.pow}


class LatticeError(ValueError):
    pass


def rpn(expr):
    """value of a quoted RPN expression ("0.5 2 /", "pi 4 / sin")"""
    st = []
    for tok in expr.split():
        t = tok.lower()
        if t in _RPN_BINARY:
            b, a = st.pop(), st.pop()
            st.append(_RPN_BINARY[t](a, b))
        elif t in _RPN_UNARY:
            st.append(_RPN_UNARY[t](st.pop()))
        elif t == "pi":
            st.append(math.pi)
        else:
            st.append(float(tok))
    if len(st) != 1:
        raise LatticeError("bad RPN expression %r" % expr)
    return st[0]


def entity_type(token):
    """tell_type's matching of a type token against the entity names"""
    found = None
    for name in ENTITY_NAMES:
        m = min(len(token), len(name))
        if token[:m] == name[:m]:
            if found is not None:
                raise LatticeError("element type %s is ambiguous" % token)
            if len(token) == len(name):
                return name
            found = name
    return found


def _statements(text):
    """logical statements: comments removed, '&' continuations joined,
    upper-cased outside quotes"""
    out, cur = [], ""
    for raw in text.splitlines():
        line, q = [], False
        for ch in raw:
            if ch == '"':
                q = not q
            elif ch == "!" and not q:
                break
            line.append(ch if q else ch.upper())
        line = "".join(line).rstrip()
        if line.endswith("&"):
            cur += line[:-1] + " "
            continue
        cur += line
        if cur.strip():
            out.append(cur.strip())
        cur = ""
    if cur.strip():
        out.append(cur.strip())
    return out


def _split_top(s, sep=","):
    parts, depth, q, cur = [], 0, False, ""
    for ch in s:
        if ch == '"':
            q = not q
        if not q:
            if ch == "(":
                depth += 1
            elif ch == ")":
                depth -= 1
            elif ch == sep and depth == 0:
                parts.append(cur)
                cur = ""
                continue
        cur += ch
    parts.append(cur)
    return [p.strip() for p in parts if p.strip()]


def _value(tok):
    if tok.startswith('"') and tok.endswith('"'):
        inner = tok[1:-1]
        try:
            return rpn(inner)
        except (LatticeError, ValueError, IndexError):
            return inner
    try:
        return float(tok) if re.search(r"[.eE]", tok) else int(tok)
    except ValueError:
        return tok


class Lattice:
    def __init__(self):
        self.elements = {}   # label -> (entity type, {PARAM: value})
        self.lines = {}      # label -> list of items
        self.order = []      # labels in definition order
        self.use = None
        self.directory = "."

    def parse(self, text):
        for st in _statements(text):
            # a statement with a "label:" prefix defines an element or line;
            # only one without it can be the USE or RETURN command, as
            # tell_type decides from the token before the colon
            # (mad_parse.c:669-713)
            m = re.match(r'^"?([^:"]+)"?\s*:\s*(.*)$', st)
            if not m:
                if re.match(r"^USE\s*(,|$)", st):
                    self.use = _split_top(st, ",")[1] if "," in st else None
                    continue
                if re.match(r"^RETURN\s*$", st):
                    break
                raise LatticeError("cannot parse %r" % st)
            label, body = m.group(1).strip(), m.group(2).strip()
            lm = re.match(r"^LINE\s*=\s*\((.*)\)\s*$", body)
            if lm:
                self.lines[label] = self._line_items(lm.group(1))
                self.order.append(label)
                continue
            parts = _split_top(body, ",")
            tname = parts[0]
            params = {}
            for p in parts[1:]:
                if "=" not in p:
                    raise LatticeError("%s: parameter %r has no value" % (label, p))
                k, v = p.split("=", 1)
                params[k.strip()] = _value(v.strip())
            if tname in self.elements:          # copy of an earlier element
                etype, base = self.elements[tname]
                params = dict(base, **params)
            else:
                etype = entity_type(tname)
                if etype is None:
                    raise LatticeError("%s: unknown element type %s" % (label, tname))
            self.elements[label] = (etype, params)
            self.order.append(label)
        return self

    def _line_items(self, s):
        items = []
        for tok in _split_top(s, ","):
            rev = False
            mult = 1
            m = re.match(r"^(-?)\s*(\d+)\s*\*\s*(.*)$", tok)
            if m:
                rev = m.group(1) == "-"
                mult = int(m.group(2))
                tok = m.group(3).strip()
            if tok.startswith("-"):
                rev = not rev
                tok = tok[1:].strip()
            if tok.startswith("(") and tok.endswith(")"):
                sub = self._line_items(tok[1:-1])
                items.append(("group", sub, rev, mult))
            else:
                if len(tok) >= 2 and tok[0] == '"' and tok[-1] == '"':
                    tok = tok[1:-1]   # quoted names, as in their definitions
                items.append(("name", tok, rev, mult))
        return items

    def expand(self, name=None, _depth=0):
        """flattened list of (label, reversed) for a line (default: USE, else
        the last line defined)"""
        if name is None:
            name = self.use or ([l for l in self.order if l in self.lines] or [None])[-1]
            if name is None:
                raise LatticeError("no beamline defined")
        if _depth > 64:
            raise LatticeError("beamline nesting too deep (recursive definition?)")
        if name in self.lines:
            return self._expand_items(self.lines[name], _depth)
        if name in self.elements:
            return [(name, False)]
        raise LatticeError("undefined item %s" % name)

    def _expand_items(self, items, depth):
        out = []
        for kind, what, rev, mult in items:
            seq = self._expand_items(what, depth + 1) if kind == "group" else self.expand(what, depth + 1)
            if rev:
                seq = [(n, not r) for n, r in reversed(seq)]
            out += seq * mult
        return out

    def beamline(self, name=None, transmute=None, ignore=()):
        """an elements.Beamline of the expanded line"""
        transmute = {k.upper(): v.upper() for k, v in (transmute or {}).items()}
        ignore = {t.upper() for t in ignore}
        els, keep = [], []
        for label, rev in self.expand(name):
            etype, params = self.elements[label]
            etype = transmute.get(etype, etype)
            if etype in ignore:
                continue
            if etype not in SUPPORTED:
                raise LatticeError("%s: element type %s is not on the tracked path (ignore or transmute it)"
                                   % (label, etype))
            e = self._make(label, SUPPORTED[etype], params, keep)
            if rev:
                _reverse_edges(e)
            els.append(e)
        return E.Beamline(els, keep=keep)

    def _make(self, label, tcode, params, keep):
        kw = {}
        for k, v in params.items():
            if tcode == A.EH_T_DRIF and k == "L":
                kw["L"] = v
            elif tcode == A.EH_T_WAKE and k in ("INPUTFILE", "TCOLUMN", "WCOLUMN"):
                continue
            elif tcode in (A.EH_T_MONI, A.EH_T_HMON, A.EH_T_VMON) and k in (
                    "READOUT", "XREADOUT", "YREADOUT", "CO_FITPOINT", "STORE_TURN_BY_TURN"):
                continue  # readouts only: nothing on the tracked path
            elif tcode == A.EH_T_RFCA and k == "FIDUCIAL":
                kw[k] = str(v).split(",")[0].lower().replace("+", "")
            else:
                kw[k] = v
        if tcode == A.EH_T_WAKE:
            from . import sdds
            path = os.path.join(self.directory, str(params.get("INPUTFILE", "")))
            t, W = sdds.wake_table(path, str(params.get("TCOLUMN", "t")), str(params.get("WCOLUMN", "W")))
            table = E.WakeTable(t, W)
            keep.append(table)
            return E.WAKE(table, label, **kw)
        try:
            return E.make(tcode, label, **kw)
        except (KeyError, ValueError) as err:
            raise LatticeError("%s: %s" % (label, err))


def _reverse_edges(e):
    """entrance/exit swap of a reversed element (mad_parse.c:564-571)"""
    if e.type == A.EH_T_CSBEND:
        b = e.p.csbend
        for a, c in (("e1", "e2"), ("h1", "h2"), ("edge1_effects", "edge2_effects"), ("fint1", "fint2"),
                     ("edge1_kick_limit", "edge2_kick_limit")):
            x, y = getattr(b, a), getattr(b, c)
            setattr(b, a, y)
            setattr(b, c, x)
    elif e.type == A.EH_T_KQUAD:
        q = e.p.kquad
        q.edge1_effects, q.edge2_effects = q.edge2_effects, q.edge1_effects


def load(path):
    """parse a .lte file"""
    with open(path) as f:
        lat = Lattice().parse(f.read())
    lat.directory = os.path.dirname(os.path.abspath(path))
    return lat


def dumps(beamline, line_name="RING"):
    """.lte text for a Beamline of supported elements (every parameter that
    differs from the type's default written out)"""
    inv = {v: k for k, v in SUPPORTED.items()}
    lines, names = [], []
    seen = {}
    for i, e in enumerate(beamline.elements):
        label = "E%d_%s" % (i, e.name.decode() or inv[e.type])
        names.append(label)
        parts = [inv[e.type]]
        if e.type in A.PARAM_STRUCT:
            field, _ = A.PARAM_STRUCT[e.type]
            sub = getattr(e.p, field)
            ref = getattr(E.make(e.type).p, field)
            for key, attr in E._NAMES[e.type].items():
                if key in ("N_KICKS",) and e.type == A.EH_T_CSBEND:
                    continue   # alias of N_SLICES for CSBEND
                v = getattr(sub, attr)
                if key == "OPEN_SIDE" and e.type in (A.EH_T_MAXAMP, A.EH_T_RCOL):
                    if v:
                        parts.append('OPEN_SIDE="%s"' % {c: t for t, c in E.OPEN_SIDE_CODES.items()}[v])
                    continue
                if isinstance(v, (int, float)) and v != getattr(ref, attr):
                    parts.append("%s=%r" % (key, v))
        if e.type == A.EH_T_WAKE:
            raise LatticeError("WAKE elements need their table file; write it with sdds.write first")
        lines.append("%s: %s" % (label, ", ".join(parts)))
    body = []
    for k in range(0, len(names), 8):
        body.append(", ".join(names[k:k + 8]))
    lines.append("%s: LINE=(%s)" % (line_name, ", &\n    ".join(body)))
    return "\n".join(lines) + "\n"
