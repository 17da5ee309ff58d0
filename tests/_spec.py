"""Builds mirror objects from the JSON specs of tests/golden/reference_goldens.json."""
import numpy as np

import stencils_hip as sh

DTYPES = {"f64": np.float64, "f32": np.float32, "i64": np.int64, "i32": np.int32,
          "u8": np.uint8, "bool": np.bool_}


def make_stencil(spec):
    kind = spec[0]
    if kind == "Positional":
        return sh.Positional(*[tuple(o) for o in spec[1]])
    if kind == "Rectangle":
        return sh.Rectangle(*[tuple(o) for o in spec[1]])
    if kind == "Annulus":
        return sh.Annulus(spec[1], spec[2], spec[3])
    return getattr(sh, kind)(*spec[1:])


def make_boundary(spec, dtype=None):
    kind = spec[0]
    if kind == "Remove":
        v = spec[1] if len(spec) > 1 else None
        if v is not None and dtype is not None:
            v = np.dtype(dtype).type(v)
        return sh.Remove(v)
    return getattr(sh, kind)()


def make_padding(name):
    return {"Conditional": sh.Conditional(), "Halo_out": sh.Halo("out"),
            "Halo_in": sh.Halo("in")}[name]


def make_input(case):
    if "input3" in case:
        sl = [np.asarray(s) for s in case["input3"]]
        return np.asfortranarray(np.stack(sl, axis=2))
    return np.asfortranarray(np.asarray(case["input"], dtype=DTYPES.get(case.get("dtype", "i64"))))


F = {"mean": sh.mean, "sum": sh.sum, "minimum": sh.minimum, "maximum": sh.maximum,
     "kernelproduct": sh.kernelproduct}
