"""`raytatouille-result` JSON codec for field references and solver outputs.

Format (as read and written by the reference, src/bwpm/results.py:1-18,58-134):

- envelope ``{"format": "raytatouille-result", "schema_version": "0.1.x", "type", "data"}``;
- arrays as ``{"dtype", "shape", "values"}`` in C order, complex values as ``[re, im]``;
- non-finite floats inside arrays as ``"NaN"``, ``"Infinity"``, ``"-Infinity"``; as scalars
  ``{"float": "NaN"}``; complex scalars ``{"complex": [re, im]}``;
- floats written with Python's shortest round-trip repr, ``allow_nan=False``.

The encoder and decoder here are this package's own; the key order of a FieldReference follows
src/bwpm/results.py:137-167 so that files written here diff cleanly against the reference's.
"""

from __future__ import annotations

import json
import math
from typing import Any

import numpy as np

FORMAT = "raytatouille-result"
SCHEMA_VERSION = "0.1.13"
FIELD_REFERENCE = "FieldReference"

_ARRAY_DTYPES = frozenset({
    "bool", "int8", "int16", "int32", "int64", "uint8", "uint16", "uint32", "uint64",
    "float32", "float64", "complex64", "complex128",
})
_NONFINITE = {"NaN": math.nan, "Infinity": math.inf, "-Infinity": -math.inf}


def _item(x: float) -> float | str:
    if math.isfinite(x):
        return x
    if math.isnan(x):
        return "NaN"
    return "Infinity" if x > 0 else "-Infinity"


def _number(x: Any) -> float:
    return _NONFINITE[x] if isinstance(x, str) else float(x)


def encode(value: Any) -> Any:
    """JSON-compatible form of dicts, lists, NumPy arrays and scalars."""
    if isinstance(value, dict):
        return {str(k): encode(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [encode(v) for v in value]
    if isinstance(value, np.ndarray):
        name = str(value.dtype)
        if name not in _ARRAY_DTYPES:
            raise ValueError(f"results: arrays of dtype {name} cannot be written")
        flat = value.ravel(order="C").tolist()
        if value.dtype.kind == "f":
            values: list[Any] = [_item(x) for x in flat]
        elif value.dtype.kind == "c":
            values = [[_item(z.real), _item(z.imag)] for z in flat]
        else:
            values = flat
        return {"dtype": name, "shape": list(value.shape), "values": values}
    if isinstance(value, np.generic):
        return encode(value.item())
    if isinstance(value, float):
        return value if math.isfinite(value) else {"float": _item(value)}
    if isinstance(value, complex):
        return {"complex": [_item(value.real), _item(value.imag)]}
    return value


def decode(value: Any) -> Any:
    """Inverse of `encode`."""
    if isinstance(value, list):
        return [decode(v) for v in value]
    if not isinstance(value, dict):
        return value
    keys = set(value)
    if keys == {"dtype", "shape", "values"}:
        dtype = np.dtype(value["dtype"])
        raw = value["values"]
        if dtype.kind == "f":
            arr = np.array([_number(x) for x in raw], dtype=dtype)
        elif dtype.kind == "c":
            arr = np.array([complex(_number(a), _number(b)) for a, b in raw], dtype=dtype)
        else:
            arr = np.array(raw, dtype=dtype)
        return arr.reshape(value["shape"])
    if keys == {"float"}:
        return _number(value["float"])
    if keys == {"complex"}:
        a, b = value["complex"]
        return complex(_number(a), _number(b))
    return {k: decode(v) for k, v in value.items()}


def to_json(type_name: str, data: dict[str, Any], indent: int | None = None) -> str:
    """Envelope text of `data` under `type_name`."""
    envelope = {"format": FORMAT, "schema_version": SCHEMA_VERSION, "type": type_name,
                "data": encode(data)}
    return json.dumps(envelope, allow_nan=False, indent=indent, ensure_ascii=False)


def load_json(text: str) -> tuple[str, dict[str, Any]]:
    """(type, decoded data) of an envelope text; accepts schema versions 0.1.x only."""
    envelope = json.loads(text)
    if envelope.get("format") != FORMAT:
        raise ValueError(f"results: format must be '{FORMAT}'")
    if not str(envelope.get("schema_version", "")).startswith("0.1."):
        raise ValueError("results: only schema versions 0.1.x are read")
    data = decode(envelope["data"])
    if not isinstance(data, dict):
        raise ValueError("results: data must be an object")
    return str(envelope["type"]), data


def field_reference(case: str, description: str, source: str, parameters: dict[str, Any],
                    inputs: dict[str, Any], expected: dict[str, Any], tolerance: dict[str, float],
                    reference: dict[str, Any] | None = None,
                    reference_tolerance: dict[str, float] | None = None) -> dict[str, Any]:
    """One FieldReference record with the key order of the format."""
    return {
        "case": case,
        "description": description,
        "source": source,
        "parameters": parameters,
        "inputs": inputs,
        "expected": expected,
        "tolerance": tolerance,
        "reference": reference or {},
        "reference_tolerance": reference_tolerance or {},
    }
