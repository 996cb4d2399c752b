"""The keyword stream gimic.bin reads on standard input (getkw.f90:1526-1646),
as the reference's frontend writes it (getkw.py, Section/Keyword __str__):
the reader gives the dict that parse() gives for the same gimic.inp."""
import pytest

from conftest import golden
from gimic_amd import inp

# a stream in the frontend's layout: set flags True/False, a section with an
# argument, an unset keyword (n = 0), an empty string (n = -1), arrays with
# one value per line, a Fortran D exponent
STREAM = """SECT top 2 True
ARG F KW 6
STR calc 1 True
integral
STR title -1 False
INT debug 1 True
1
BOOL openshell 1 True
False
STR magnet_axis 0 False
DBL_ARRAY magnet 3 True
0.0
0.0
1.0
SECT Grid 0 True
ARG T KW 3
STR Grid 1 True
bond
INT_ARRAY bond 2 True
2
1
DBL_ARRAY width 2 True
-1.25614
6.0
INT gauss_order 1 True
9
SECT Advanced 0 True
ARG F KW 2
DBL screening_thrs 1 True
1.d-8
BOOL spherical 1 True
False
"""


def test_stream_reads_the_frontend_layout():
    kw = inp.parse_stream(STREAM)
    assert kw == {
        "calc": "integral", "title": "", "debug": 1, "openshell": False,
        "magnet": [0.0, 0.0, 1.0],
        "Grid": {"_arg": "bond", "bond": [2, 1], "width": [-1.25614, 6.0], "gauss_order": 9},
        "Advanced": {"screening_thrs": 1e-8, "spherical": False},
    }
    assert "magnet_axis" not in kw  # unset: left to the defaults


def _stream(name, d, set_flag=True):
    """a parse() dict written as the frontend writes its sections"""
    kws = {k: v for k, v in d.items() if not isinstance(v, dict) and k != "_arg"}
    sects = {k: v for k, v in d.items() if isinstance(v, dict)}

    def typ(v):
        x = v[0] if isinstance(v, list) else v
        t = "BOOL" if isinstance(x, bool) else "INT" if isinstance(x, int) else \
            "DBL" if isinstance(x, float) else "STR"
        return t + "_ARRAY" if isinstance(v, list) else t

    def kwtext(k, v):
        if v == "":
            return f"STR {k} -1 True\n"
        vals = v if isinstance(v, list) else [v]
        return f"{typ(v)} {k} {len(vals)} True\n" + "".join(f"{x!r}\n" if isinstance(x, float)
                                                            else f"{x}\n" for x in vals)

    out = f"SECT {name} {len(sects)} {set_flag}\n"
    out += f"ARG {'T' if '_arg' in d else 'F'} KW {len(kws)}\n"
    if "_arg" in d:
        out += kwtext(name, d["_arg"])
    out += "".join(kwtext(k, v) for k, v in kws.items())
    out += "".join(_stream(k, v) for k, v in sects.items())
    return out


@pytest.mark.parametrize("fixture,key", [("c4h4.npz", "integration_inp"),
                                         ("openshell.npz", "integration_inp"),
                                         ("openshell.npz", "threed_inp")])
def test_stream_of_fixture_inputs_equals_their_gimic_inp(fixture, key):
    text = str(golden(fixture)[key])
    want = inp.parse(text)
    stream = _stream("top", want)
    assert inp.parse_stream(stream) == want
    assert inp.parse_any(stream) == want and inp.parse_any(text) == want


def test_stream_errors_are_reported():
    with pytest.raises(ValueError, match="expected SECT"):
        inp.parse_stream("calc = integral\n")
    with pytest.raises(ValueError, match="unexpected end"):
        inp.parse_stream("SECT top 0 True\nARG F KW 1\nINT debug 1 True\n")
    with pytest.raises(ValueError, match="invalid type"):
        inp.parse_stream("SECT top 0 True\nARG F KW 1\nCPLX z 1 True\n1\n")
