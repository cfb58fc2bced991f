"""Lattice files: AT JSON and AT repr, besides MATLAB .mat (lattice.load_mat).

* JSON (pyat/at/load/json.py): {"atjson": 1, "at_version", "elements":
  [element dicts with a "Class" key], "properties": {lattice attributes,
  "particle": {...}}}; an element dict holds FamName, Length and PassMethod
  first, then the other attributes sorted by name
  (lattice/elements/element_object.py:235-276).
* repr (pyat/at/load/reprfile.py): first line the repr of the lattice
  attributes, then one ``Class(args, keyword=value, ...)`` per element, the
  positional arguments being the class's build attributes
  (element_object.py:221-233). The reference reads it with ``eval``; here
  each line is parsed with ``ast`` and only literals, ``array(...)``,
  ``Particle(...)`` and the element classes are accepted.
"""

from __future__ import annotations

import ast
import json
import math
import os
import sys

import numpy as np

from .lattice import (AtError, Aperture, Corrector, Dipole, Drift, Element, Lattice, Marker,
                      Monitor, Multipole, Octupole, Particle, Quadrupole, RFCavity, Sextupole,
                      ThinMultipole, load_mat)

__all__ = ["load_json", "save_json", "load_repr", "save_repr", "load_lattice", "save_lattice",
           "element_from_dict", "element_to_dict"]

AT_VERSION = "0.6.0"

# Positional (build) attributes of each class, in order.
_BUILD = {
    Element: ("FamName",),
    Marker: ("FamName",),
    Monitor: ("FamName",),
    Aperture: ("FamName", "Limits"),
    Drift: ("FamName", "Length"),
    ThinMultipole: ("FamName", "PolynomA", "PolynomB"),
    Multipole: ("FamName", "Length", "PolynomA", "PolynomB"),
    Octupole: ("FamName", "Length", "PolynomA", "PolynomB"),
    Dipole: ("FamName", "Length", "BendingAngle", "K"),
    Quadrupole: ("FamName", "Length", "K"),
    Sextupole: ("FamName", "Length", "H"),
    RFCavity: ("FamName", "Length", "Voltage", "Frequency", "HarmNumber", "Energy"),
    Corrector: ("FamName", "Length", "KickAngle"),
}
_FILE_NAME = {Dipole: "Bend"}
_CLASSES = {cls.__name__: cls for cls in _BUILD}
_CLASSES.update({"Bend": Dipole, "Dipole": Dipole})


def _build(cls) -> tuple:
    for c in cls.__mro__:
        if c in _BUILD:
            return _BUILD[c]
    return ("FamName",)


def _items(elem: Element):
    v = dict(vars(elem))
    for k in ("FamName", "Length", "PassMethod"):
        if k in v:
            yield k, v.pop(k)
    for k in sorted(v):
        if not k.startswith("_"):
            yield k, v[k]


def _plain(v):
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.integer):
        return int(v)
    if isinstance(v, np.floating):
        return float(v)
    return v


def element_to_dict(elem: Element) -> dict:
    """Attributes of an element plus its "Class" (element_object.py:262-277)."""
    d = {k: _plain(v) for k, v in _items(elem)}
    d["Class"] = _FILE_NAME.get(type(elem), type(elem).__name__)
    return d


def element_from_dict(d: dict, index: int | None = None) -> Element:
    """Element of the class named by "Class" (or deduced from PassMethod)
    built from its build attributes, the others as keywords
    (load/utils.py:278-310)."""
    d = dict(d)
    cname = d.pop("Class", None)
    cls = _CLASSES.get(cname) if cname else None
    if cls is None:
        pm = d.get("PassMethod", "")
        cls = {"DriftPass": Drift, "RFCavityPass": RFCavity, "CorrectorPass": Corrector,
               "AperturePass": Aperture, "ThinMPolePass": ThinMultipole}.get(pm)
        if cls is None:
            if pm.startswith("BndMPole"):
                cls = Dipole
            elif pm.startswith("StrMPole"):
                cls = Multipole
            elif "Length" in d and d["Length"] != 0:
                cls = Drift if pm in ("", "IdentityPass") else Element
            else:
                cls = Marker
    args = [d.pop(a, None) for a in _build(cls)]
    while args and args[-1] is None:
        args.pop()
    if any(a is None for a in args):
        where = "" if index is None else f" #{index}"
        raise AtError(f"Element{where} of class {cls.__name__}: missing build attribute")
    for k, v in list(d.items()):
        if isinstance(v, list):
            d[k] = np.array(v, dtype=np.float64) if _numeric(v) else v
    if cls is Element:
        return Element(*args, **d)
    return cls(*args, **d)


def _numeric(v) -> bool:
    try:
        np.asarray(v, dtype=np.float64)
        return True
    except (TypeError, ValueError):
        return False


def _properties(ring: Lattice) -> dict:
    props = {"name": getattr(ring, "name", ""), "energy": float(ring.energy),
             "periodicity": int(getattr(ring, "periodicity", 1))}
    part = ring.particle
    props["particle"] = {"name": part.name, "rest_energy": float(part.rest_energy),
                         "charge": float(part.charge)}
    for k, v in vars(ring).items():
        if k.startswith("_") or k in props or k in ("particle", "in_file", "mat_key", "use",
                                                      "mat_file", "m_file", "repr_file"):
            continue
        if isinstance(v, (int, float, str, bool, np.integer, np.floating)):
            props[k] = _plain(v)
    return props


def _lattice(elements, props: dict, **kwargs) -> Lattice:
    props = dict(props)
    pdict = props.pop("particle", None)
    if isinstance(pdict, Particle):
        particle = pdict
    elif isinstance(pdict, dict):
        pd = dict(pdict)
        particle = Particle(pd.pop("name", None), **pd)
    else:
        particle = Particle("relativistic")
    props.update(kwargs)
    energy = props.pop("energy", 0.0)
    periodicity = props.pop("periodicity", 1)
    name = props.pop("name", "")
    particle = props.pop("particle", particle)
    # derived quantities are properties of the lattice, not stored attributes
    for k in [k for k in props if isinstance(getattr(Lattice, k, None), property)]:
        props.pop(k)
    return Lattice(elements, energy=energy, particle=particle, periodicity=periodicity,
                   name=name, **props)


def save_json(ring: Lattice, filename=None, compact: bool = False) -> None:
    """json.py:38-72"""
    data = {"atjson": 1, "at_version": AT_VERSION,
            "elements": [element_to_dict(e) for e in ring],
            "properties": _properties(ring)}
    indent = None if compact else 2
    if filename is None:
        print(json.dumps(data, indent=indent))
    else:
        with open(filename, "w") as f:
            json.dump(data, f, indent=indent)


def load_json(filename, **kwargs) -> Lattice:
    """AT JSON lattice (json.py:75-157; Xsuite files are not read)."""
    with open(filename) as f:
        data = json.load(f)
    if "atjson" not in data and "elements" not in data:
        raise AtError("Cannot guess the file origin: not an AT JSON file")
    elems = [element_from_dict(d, index=i) for i, d in enumerate(data["elements"])]
    return _lattice(elems, data.get("properties", {}), **kwargs)


def _repr_value(v) -> str:
    if isinstance(v, np.ndarray):
        with np.printoptions(threshold=sys.maxsize):
            body = np.array2string(v, separator=", ", max_line_width=sys.maxsize,
                                   formatter={"float_kind": lambda x: repr(float(x)),
                                              "int_kind": lambda x: repr(int(x))})
        return "array(" + " ".join(body.split()) + ")"
    if isinstance(v, float):
        return repr(v) if math.isfinite(v) else ("NaN" if math.isnan(v) else repr(v))
    return repr(v)


def element_repr(elem: Element) -> str:
    """``Class(args, keyword=value, ...)`` (element_object.py:118-123,
    221-233): the build attributes positionally, then the attributes that
    differ from an element built from those arguments alone."""
    cls = type(elem)
    attrs = dict(_items(elem))
    args = []
    for a in _build(cls):
        args.append(attrs.pop(a, getattr(elem, a, None)))
    ref = cls(*args) if cls is not Element else Element(*args)
    # MaxOrder is always written: the constructors derive it from the
    # polynomial lengths, which the keywords may change
    kws = {k: v for k, v in attrs.items()
           if k == "MaxOrder" or not (hasattr(ref, k) and _same(v, getattr(ref, k)))}
    parts = [_repr_value(a) for a in args] + [f"{k}={_repr_value(v)}" for k, v in kws.items()]
    return f"{cls.__name__}({', '.join(parts)})"


def _same(a, b) -> bool:
    try:
        return bool(np.array_equal(a, b))
    except Exception:
        return False


def save_repr(ring: Lattice, filename=None) -> None:
    """reprfile.py:77-100"""
    props = _properties(ring)
    pd = props.pop("particle")
    head = "{" + ", ".join([f"{k!r}: {_repr_value(v)}" for k, v in props.items()] +
                           [f"'particle': Particle({pd['name']!r}, rest_energy="
                            f"{pd['rest_energy']!r}, charge={pd['charge']!r})"]) + "}"
    lines = [head] + [element_repr(e) for e in ring]
    text = "\n".join(lines) + "\n"
    if filename is None:
        sys.stdout.write(text)
    else:
        with open(filename, "w") as f:
            f.write(text)


_NAMES = {"NaN": math.nan, "nan": math.nan, "inf": math.inf, "Inf": math.inf, "True": True,
          "False": False, "None": None}


def _literal(node):
    """Value of a repr-file expression node: literals, array(...) and
    Particle(...) only."""
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, (ast.List, ast.Tuple)):
        vals = [_literal(e) for e in node.elts]
        return vals if isinstance(node, ast.List) else tuple(vals)
    if isinstance(node, ast.Dict):
        return {_literal(k): _literal(v) for k, v in zip(node.keys, node.values)}
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _literal(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.Name) and node.id in _NAMES:
        return _NAMES[node.id]
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
        fname = node.func.id
        args = [_literal(a) for a in node.args]
        kws = {k.arg: _literal(k.value) for k in node.keywords}
        if fname == "array":
            dt = kws.get("dtype")
            return np.array(args[0], dtype=np.float64 if dt is None else dt)
        if fname == "Particle":
            return Particle(*args, **kws)
    if isinstance(node, ast.Name) and node.id in ("uint8", "float64", "int32"):
        return node.id
    raise AtError(f"Unsupported expression in repr file: {ast.dump(node)[:80]}")


def _element_from_repr(line: str) -> Element:
    tree = ast.parse(line.strip(), mode="eval").body
    if not (isinstance(tree, ast.Call) and isinstance(tree.func, ast.Name)
            and tree.func.id in _CLASSES):
        raise AtError(f"Not an element definition: {line[:60]!r}")
    cls = _CLASSES[tree.func.id]
    args = [_literal(a) for a in tree.args]
    kws = {k.arg: _literal(k.value) for k in tree.keywords}
    return cls(*args, **kws) if cls is not Element else Element(*args, **kws)


def load_repr(filename, **kwargs) -> Lattice:
    """AT repr file (reprfile.py:40-74), parsed without eval."""
    with open(filename) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    if not lines:
        raise AtError("empty repr file")
    props = _literal(ast.parse(lines[0], mode="eval").body)
    elems = [_element_from_repr(ln) for ln in lines[1:]]
    return _lattice(elems, props, **kwargs)


def load_lattice(filename, **kwargs) -> Lattice:
    """Dispatch on the file extension (load/allfiles.py)."""
    ext = os.path.splitext(str(filename))[1].lower()
    if ext == ".mat":
        return load_mat(filename, **kwargs)
    if ext == ".json":
        return load_json(filename, **kwargs)
    if ext == ".repr":
        return load_repr(filename, **kwargs)
    raise AtError(f"File extension {ext!r} is not supported")


def save_lattice(ring: Lattice, filename, **kwargs) -> None:
    ext = os.path.splitext(str(filename))[1].lower()
    if ext == ".json":
        return save_json(ring, filename, **kwargs)
    if ext == ".repr":
        return save_repr(ring, filename, **kwargs)
    raise AtError(f"File extension {ext!r} is not supported for saving")
