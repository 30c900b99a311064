"""Reader/writer for the Fortran namelist groups of MADWAVE3's input.dat.

The reference reads its configuration with Fortran namelist I/O from input.dat
(group list: SRC/mod_gridYpara_01y2.f:92-109, mod_pot_01y2.f:62-65,
mod_absorcion_01y2.f:42-43, mod_colini_01y2.f:36, mod_flux_01y2.f:72-73,
main_madwave3.f:110, mod_PsiE_01y2.f:30).  This module parses the subset of
namelist syntax those files use: ``&group key=value, ... /`` with
case-insensitive keys, ``!`` comments, Fortran ``d`` exponents, logical and
quoted-string values, and arrays given as ``key(i)=v`` or ``key=v1,v2``.
Each reader of the reference supplies defaults before the read; callers here
do the same with :func:`group`.
"""
import re

__all__ = ["parse", "group", "dump_group", "NamelistError"]


class NamelistError(ValueError):
    pass


_num = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?$")
_int = re.compile(r"^[+-]?\d+$")


def _strip_comment(line):
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


def _value(tok):
    t = tok.strip()
    if not t:
        return None
    if t[0] in "'\"":
        return t[1:-1] if len(t) >= 2 and t[-1] == t[0] else t[1:]
    low = t.lower()
    if low in (".true.", ".t.", "t", "true"):
        return True
    if low in (".false.", ".f.", "f", "false"):
        return False
    if "*" in t:
        n, v = t.split("*", 1)
        if _int.match(n.strip()):
            return [_value(v)] * int(n)
    if _int.match(t):
        return int(t)
    if _num.match(t):
        return float(t.replace("d", "e").replace("D", "e"))
    raise NamelistError("cannot read namelist value %r" % tok)


def _split_values(text):
    vals, cur, q = [], [], None
    for ch in text:
        if q:
            cur.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            cur.append(ch)
        elif ch in ", \t\n":
            if cur:
                vals.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
    if cur:
        vals.append("".join(cur))
    out = []
    for v in vals:
        x = _value(v)
        if isinstance(x, list):
            out.extend(x)
        elif x is not None:
            out.append(x)
    return out


def parse(text):
    """Return {group_name_lower: {key_lower: value}} for every group in text.

    A scalar key holds a scalar; ``key=v1,v2`` holds a list; ``key(i)=v``
    stores into a dict {i: v} under the key (1-based, as in Fortran).
    """
    groups = {}
    body = "\n".join(_strip_comment(l) for l in text.splitlines())
    for m in re.finditer(r"&(\w+)(.*?)(?:^\s*/|/\s*$|\s/\s|&end)", body,
                         re.S | re.M | re.I):
        name = m.group(1).lower()
        content = m.group(2)
        entries = {}
        parts = re.split(r"([A-Za-z_]\w*(?:\s*\(\s*\d+\s*\))?)\s*=", content)
        for i in range(1, len(parts) - 1, 2):
            key = parts[i].strip().lower().replace(" ", "")
            vals = _split_values(parts[i + 1])
            km = re.match(r"(\w+)\((\d+)\)", key)
            if km:
                d = entries.setdefault(km.group(1), {})
                if not isinstance(d, dict):
                    d = {j + 1: v for j, v in enumerate(d if isinstance(d, list) else [d])}
                    entries[km.group(1)] = d
                for j, v in enumerate(vals):
                    d[int(km.group(2)) + j] = v
            else:
                if len(vals) == 0:
                    continue
                entries[key] = vals[0] if len(vals) == 1 else vals
        groups[name] = entries
    return groups


def group(groups, name, defaults=None, required=True):
    """Values of one group over its defaults (the reference sets defaults
    before each namelist read, e.g. mod_gridYpara_01y2.f:116-119)."""
    g = groups.get(name.lower())
    if g is None:
        if required:
            raise NamelistError("input.dat has no namelist &%s" % name)
        g = {}
    out = {k.lower(): v for k, v in (defaults or {}).items()}
    out.update(g)
    return out


def _fmt(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return repr(v).replace("e", "d") if ("e" in repr(v)) else repr(v) + "d0"
    if isinstance(v, str):
        return "'%s'" % v
    if isinstance(v, (list, tuple)):
        return ", ".join(_fmt(x) for x in v)
    raise NamelistError("cannot write %r" % (v,))


def dump_group(name, values):
    """Text of one namelist group, one key per line."""
    lines = ["&%s" % name]
    for k, v in values.items():
        lines.append("  %s=%s," % (k, _fmt(v)))
    lines.append("/")
    return "\n".join(lines) + "\n"
