"""Reader for gimic.inp keyword files (getkw syntax, src/gimic.in:59-124).

    key = value            value: number (1.d-8 style allowed), on/off,
                           true/false, "string" or bare word, [a, b, ...]
    Section(arg) { ... }   nested sections, arg optional
    # comment

Returns a dict: top-level keywords plus one dict per section; a section's
argument is stored under '_arg'.  Only keywords present in the file are
returned; `with_defaults` adds the schema defaults of src/gimic.in.

`parse_stream` reads the keyword stream that gimic.bin takes on its
standard input (the reference's frontend parses gimic.inp against its
template and pipes the result; getkw.f90:1526-1646 reads it) into the same
dict; `parse_any` takes either form.
"""
import re

TOKEN = re.compile(r'"[^"]*"|\[|\]|\{|\}|=|,|\(|\)|[^\s\[\]{}=,()"]+')

DEFAULTS = {
    "title": "", "debug": 0, "basis": "mol", "xdens": "XDENS", "magnet_axis": "",
    "openshell": False, "dryrun": False,
    "Advanced": {"screening": False, "screening_thrs": 1.0e-8, "spherical": True, "GIAO": True,
                 "diamag": True, "paramag": True, "lip_order": 3},
    "Essential": {"acid": False, "jmod": False, "prop": False},
    "Grid": {"_arg": "std", "type": "even", "gauss_order": 7},
}


def _value(tok):
    if tok.startswith('"'):
        return tok[1:-1]
    low = tok.lower()
    if low in ("on", "true", "yes"):
        return True
    if low in ("off", "false", "no"):
        return False
    try:
        if re.fullmatch(r"[-+]?\d+", tok):
            return int(tok)
        return float(tok.replace("d", "e").replace("D", "E"))
    except ValueError:
        return tok


def parse(text):
    lines = [l.split("#", 1)[0] if '"' not in l.split("#", 1)[0] or l.count('"') % 2 == 0
             else l for l in text.splitlines()]
    toks = TOKEN.findall("\n".join(lines))
    root = {}
    stack = [root]
    i = 0
    while i < len(toks):
        t = toks[i]
        if t == "}":
            stack.pop()
            i += 1
            continue
        if i + 1 < len(toks) and toks[i + 1] == "=":
            key = t
            i += 2
            if toks[i] == "[":
                vals = []
                i += 1
                while toks[i] != "]":
                    if toks[i] != ",":
                        vals.append(_value(toks[i]))
                    i += 1
                i += 1
                stack[-1][key] = vals
            else:
                stack[-1][key] = _value(toks[i])
                i += 1
            continue
        # section header: Name [ ( arg ) ] {
        name = t
        sect = {}
        i += 1
        if toks[i] == "(":
            sect["_arg"] = _value(toks[i + 1])
            i += 3
        if toks[i] != "{":
            raise ValueError(f"gimic.inp: expected '{{' after section {name}")
        i += 1
        stack[-1][name] = sect
        stack.append(sect)
    return root


def with_defaults(kw):
    out = {k: (dict(v) if isinstance(v, dict) else v) for k, v in DEFAULTS.items()}
    for k, v in kw.items():
        if isinstance(v, dict):
            out.setdefault(k, {}).update(v)
        else:
            out[k] = v
    return out


_STREAM_SCALAR = {"INT": int, "DBL": lambda t: float(t.replace("d", "e").replace("D", "E")),
                  "BOOL": lambda t: t.strip().lower() in ("t", "true", ".true.", "1", "yes", "on"),
                  "STR": str}


def parse_stream(text):
    """Keyword stream (getkw.f90:1526-1646; written by the frontend's
    getkw.py, Section/Keyword __str__) -> the dict of parse().

    Section:  SECT <name> <nsect> <set>
              ARG T|F KW <nkw>
              [the section's argument, as a keyword]
              <nkw keywords> <nsect sections>
    Keyword:  <TYPE> <name> <n> <set>, then n value lines; n = 0: no value
              (left out of the dict), STR with n = -1: the empty string.
    """
    lines = text.splitlines()
    pos = 0

    def line():
        nonlocal pos
        if pos >= len(lines):
            raise ValueError("keyword stream: unexpected end")
        pos += 1
        return lines[pos - 1]

    def keyword():
        f = line().split()
        if len(f) < 4:
            raise ValueError(f"keyword stream: bad keyword header {' '.join(f)!r}")
        typ, name, n = f[0], f[1], int(f[2])
        base = typ[:-6] if typ.endswith("_ARRAY") else typ
        if base not in _STREAM_SCALAR:
            raise ValueError(f"keyword stream: invalid type {typ} for {name}")
        if typ == "STR" and n == -1:
            return name, ""
        if n <= 0:
            return name, None
        if base == "STR":  # one record per value, read whole
            vals = [line() for _ in range(n)]
        else:  # list-directed: n values, possibly over several records
            vals = []
            while len(vals) < n:
                vals += line().replace(",", " ").split()
            vals = [_STREAM_SCALAR[base](v) for v in vals[:n]]
        return name, (vals if typ.endswith("_ARRAY") or n > 1 else vals[0])

    def section():
        f = line().split()
        if len(f) < 3 or f[0] != "SECT":
            raise ValueError(f"keyword stream: expected SECT, got {' '.join(f)!r}")
        name, nsect = f[1], int(f[2])
        g = line().split()
        if len(g) < 4 or g[0] != "ARG" or g[2] != "KW":
            raise ValueError(f"keyword stream: expected ARG ... KW, got {' '.join(g)!r}")
        d = {}
        if g[1].upper().startswith("T"):
            _, v = keyword()
            if v is not None:
                d["_arg"] = v
        for _ in range(int(g[3])):
            k, v = keyword()
            if v is not None:
                d[k] = v
        for _ in range(nsect):
            k, v = section()
            d[k] = v
        return name, d

    _, top = section()
    return top


def parse_any(text):
    """gimic.inp text or the keyword stream"""
    return parse_stream(text) if text.lstrip().startswith("SECT ") else parse(text)
