"""Reading and writing `raytatouille-result` files (the reference's fixture and result format).

The format is fixed by the reference (src/bwpm/results.py:1-18 describes it, :58-134 reads and
writes it, :137-167 builds a FieldReference); files written here must load there and vice versa.

- A file is one JSON object: ``format`` ("raytatouille-result"), ``schema_version`` (0.1.x),
  ``type`` and ``data``.
- An array is ``{"dtype", "shape", "values"}`` with the values flattened in C order; a complex
  value is the pair ``[re, im]``. Inside arrays a non-finite float is one of the tokens ``"NaN"``,
  ``"Infinity"``, ``"-Infinity"``.
- A lone float that is not finite is ``{"float": token}``; a lone complex is
  ``{"complex": [re, im]}`` (tokens allowed in both parts).
- Finite floats are written with Python's shortest round-trip representation and the text never
  holds JSON's non-standard NaN/Infinity literals.

Encoding dispatches on the value's type (`functools.singledispatch`); arrays convert their
non-finite entries by index after one vectorised scan, and complex arrays go through an interleaved
float view. Decoding recognises the tagged objects by their exact key sets and rebuilds arrays from
flat float buffers.
"""

from __future__ import annotations

import json
import math
from functools import singledispatch
from typing import Any

import numpy as np

FORMAT = "raytatouille-result"
SCHEMA_VERSION = "0.1.13"
FIELD_REFERENCE = "FieldReference"

# the dtypes the format allows for arrays
ARRAY_DTYPES = frozenset(
    np.dtype(t).name
    for t in (np.bool_, np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64,
              np.float32, np.float64, np.complex64, np.complex128)
)

# non-finite float ↔ token
_TOKENS = {"NaN": math.nan, "Infinity": math.inf, "-Infinity": -math.inf}

# keys of a FieldReference record, in the order the format writes them
FIELD_REFERENCE_KEYS = ("case", "description", "source", "parameters", "inputs", "expected", "tolerance",
                        "reference", "reference_tolerance")

_ARRAY_KEYS = frozenset(("dtype", "shape", "values"))


def _token_of(x: float) -> str:
    if x != x:
        return "NaN"
    return "Infinity" if x > 0 else "-Infinity"


def _float_list(flat: np.ndarray) -> list[Any]:
    """A float64 vector as a JSON list, its non-finite entries replaced by tokens."""
    out: list[Any] = flat.tolist()
    for i in np.flatnonzero(~np.isfinite(flat)).tolist():
        out[i] = _token_of(out[i])
    return out


def _read_number(v: Any) -> float:
    if isinstance(v, str):
        if v not in _TOKENS:
            raise ValueError(f"results: {v!r} is not a number token")
        return _TOKENS[v]
    return float(v)


# ---- encoding ---------------------------------------------------------------------------------

@singledispatch
def encode(value: Any) -> Any:
    """JSON-compatible form of `value` (dicts, lists, tuples, NumPy arrays and scalars, floats,
    complex numbers); strings, ints, bools and None pass through."""
    return value


@encode.register(dict)
def _encode_dict(value: dict) -> dict:
    return {str(k): encode(v) for k, v in value.items()}


@encode.register(list)
@encode.register(tuple)
def _encode_seq(value: list | tuple) -> list:
    return [encode(v) for v in value]


@encode.register(np.generic)
def _encode_np_scalar(value: np.generic) -> Any:
    return encode(value.item())


@encode.register(float)
def _encode_float(value: float) -> Any:
    return value if math.isfinite(value) else {"float": _token_of(value)}


@encode.register(complex)
def _encode_complex(value: complex) -> dict:
    return {"complex": _float_list(np.array([value.real, value.imag], dtype=np.float64))}


@encode.register(np.ndarray)
def _encode_array(value: np.ndarray) -> dict:
    name = value.dtype.name
    if name not in ARRAY_DTYPES:
        raise ValueError(f"results: arrays of dtype {name} cannot be written")
    flat = np.ascontiguousarray(value).reshape(-1)
    if value.dtype.kind == "c":
        # re, im interleaved in one float64 vector, then cut into pairs
        inter = flat.astype(np.complex128).view(np.float64)
        vals = _float_list(inter)
        values: list[Any] = [vals[i:i + 2] for i in range(0, len(vals), 2)]
    elif value.dtype.kind == "f":
        values = _float_list(flat.astype(np.float64))
    else:
        values = flat.tolist()
    return {"dtype": name, "shape": list(value.shape), "values": values}


# ---- decoding ---------------------------------------------------------------------------------

def _decode_array(obj: dict) -> np.ndarray:
    dtype = np.dtype(obj["dtype"])
    if dtype.name not in ARRAY_DTYPES:
        raise ValueError(f"results: arrays of dtype {dtype.name} cannot be read")
    shape = tuple(int(s) for s in obj["shape"])
    raw = obj["values"]
    if dtype.kind == "c":
        parts = [_read_number(x) for pair in raw for x in pair]
        if len(parts) != 2 * len(raw):
            raise ValueError("results: complex values must be [re, im] pairs")
        flat = np.array(parts, dtype=np.float64).view(np.complex128).astype(dtype)
    elif dtype.kind == "f":
        flat = np.fromiter((_read_number(x) for x in raw), dtype=np.float64, count=len(raw)).astype(dtype)
    else:
        flat = np.array(raw, dtype=dtype)
    if flat.size != math.prod(shape):
        raise ValueError(f"results: {flat.size} values do not fill shape {list(shape)}")
    return flat.reshape(shape)


def decode(value: Any) -> Any:
    """Python/NumPy form of encoded data (the inverse of `encode`)."""
    if isinstance(value, list):
        return [decode(v) for v in value]
    if not isinstance(value, dict):
        return value
    keys = value.keys()
    if keys == _ARRAY_KEYS:
        return _decode_array(value)
    if keys == {"float"}:
        return _read_number(value["float"])
    if keys == {"complex"}:
        re, im = (_read_number(x) for x in value["complex"])
        return complex(re, im)
    return {k: decode(v) for k, v in value.items()}


# ---- files ------------------------------------------------------------------------------------

def _schema_readable(version: Any) -> bool:
    parts = str(version).split(".")
    return len(parts) == 3 and parts[0] == "0" and parts[1] == "1" and parts[2].isdigit()


def to_json(type_name: str, data: dict[str, Any], indent: int | None = None) -> str:
    """Text of a result file holding `data` as `type_name`."""
    doc = dict(zip(("format", "schema_version", "type", "data"), (FORMAT, SCHEMA_VERSION, type_name, encode(data))))
    return json.dumps(doc, allow_nan=False, indent=indent)


def load_json(text: str) -> tuple[str, dict[str, Any]]:
    """(type, data) of a result file's text; schema versions 0.1.x are read."""
    doc = json.loads(text)
    if not isinstance(doc, dict) or doc.get("format") != FORMAT:
        raise ValueError(f"results: format must be '{FORMAT}'")
    if not _schema_readable(doc.get("schema_version")):
        raise ValueError("results: only schema versions 0.1.x are read")
    data = decode(doc.get("data"))
    if not isinstance(data, dict):
        raise ValueError("results: data must be an object")
    return str(doc.get("type")), data


def field_reference(case: str, description: str, source: str, parameters: dict[str, Any],
                    inputs: dict[str, Any], expected: dict[str, Any], tolerance: dict[str, float],
                    reference: dict[str, Any] | None = None,
                    reference_tolerance: dict[str, float] | None = None) -> dict[str, Any]:
    """A FieldReference record (one test case: parameters, inputs, expected values and their
    tolerance, optionally an independent reference value and its tolerance), keys in format order."""
    values = (case, description, source, parameters, inputs, expected, tolerance, reference or {},
              reference_tolerance or {})
    return dict(zip(FIELD_REFERENCE_KEYS, values))
