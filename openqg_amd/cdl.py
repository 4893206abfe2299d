"""Read the CDL (netCDF text) configuration files of an OpenQG experiment.

OpenQG's experiment harness turns every `.cdl` file into a netCDF file with
`ncgen` before the model reads it (run_model.py:126-140).  The configuration
files hold only dimensions, scalar or 1-D variables and their values in the
`data:` section, so this module reads the text directly: `read_cdl(path)`
returns a dict of each variable's value (a float, an int, a string, or a list
for a 1-D variable) plus `"__dims__"`, the dimensions.  Variables declared
`integer`/`int` come back as ints, the others as floats; character variables
as stripped strings.
"""
from __future__ import annotations

import re

_TYPES = {"double": float, "float": float, "real": float, "int": int, "integer": int, "short": int, "long": int,
          "byte": int, "char": str}


def _strip_comments(text: str) -> str:
    out = []
    for line in text.splitlines():
        q = False
        for k, ch in enumerate(line):
            if ch == '"':
                q = not q
            if not q and line.startswith("//", k):
                line = line[:k]
                break
        out.append(line)
    return "\n".join(out)


def _statements(block: str):
    """Split on ';' outside quoted strings."""
    cur, q = [], False
    for ch in block:
        if ch == '"':
            q = not q
        if ch == ";" and not q:
            yield "".join(cur).strip()
            cur = []
        else:
            cur.append(ch)
    rest = "".join(cur).strip()
    if rest:
        yield rest


def parse_cdl(text: str) -> dict:
    text = _strip_comments(text)
    m = re.search(r"\{(.*)\}", text, re.S)
    if not m:
        raise ValueError("not a CDL file: no { ... } body")
    body = m.group(1)
    sections = re.split(r"^\s*(dimensions|variables|data)\s*:", body, flags=re.M)
    parts = dict(zip(sections[1::2], sections[2::2]))
    dims, types, out = {}, {}, {}
    for st in _statements(parts.get("dimensions", "")):
        for d in st.split(","):
            if "=" in d:
                k, v = d.split("=", 1)
                v = v.strip()
                dims[k.strip()] = 0 if v.upper() == "UNLIMITED" else int(v)
    for st in _statements(parts.get("variables", "")):
        mm = re.match(r"(\w+)\s+(.*)", st)
        if not mm or mm.group(1) not in _TYPES:
            continue   # attributes (name:attr = ...) are not needed
        for decl in re.split(r",(?![^(]*\))", mm.group(2)):
            name = decl.split("(")[0].strip()
            if name:
                types[name] = (_TYPES[mm.group(1)], "(" in decl)
    for st in _statements(parts.get("data", "")):
        if "=" not in st:
            continue
        k, v = st.split("=", 1)
        name = k.strip()
        conv, is_array = types.get(name, (float, False))
        v = v.strip()
        if conv is str:
            strs = re.findall(r'"((?:[^"\\]|\\.)*)"', v)
            out[name] = "".join(strs).strip()
            continue
        vals = [x.strip() for x in v.split(",") if x.strip()]
        conv_vals = [int(float(x)) if conv is int else float(x.rstrip("dDfF") if x[-1:] in "fF" else x.replace("d", "e").replace("D", "e"))
                     for x in vals]
        out[name] = conv_vals if (is_array or len(conv_vals) > 1) else conv_vals[0]
    out["__dims__"] = dims
    return out


def read_cdl(path: str) -> dict:
    with open(path) as f:
        return parse_cdl(f.read())
