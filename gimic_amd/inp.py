"""Reader for gimic.inp keyword files (getkw syntax, src/gimic.in:59-124).

    key = value            value: number (1.d-8 style allowed), on/off,
                           true/false, "string" or bare word, [a, b, ...]
    Section(arg) { ... }   nested sections, arg optional
    # comment

Returns a dict: top-level keywords plus one dict per section; a section's
argument is stored under '_arg'.  Only keywords present in the file are
returned; `with_defaults` adds the schema defaults of src/gimic.in.
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
