"""Converts the oracle's mode representation to the ndconv API objects, and
draws random cases."""
import numpy as np

import ndconv as nc


def conv_mode(m):
    if m == "full":
        return nc.ConvMode.Full
    if m == "same":
        return nc.ConvMode.Same
    if m == "valid":
        return nc.ConvMode.Valid
    if m[0] == "custom":
        return nc.ConvMode.Custom(m[1], m[2])
    return nc.ConvMode.Explicit(m[1], m[2])


def border(b):
    if b == "zeros":
        return nc.BorderType.Zeros
    if b == "reflect":
        return nc.BorderType.Reflect
    if b == "replicate":
        return nc.BorderType.Replicate
    if b == "circular":
        return nc.BorderType.Circular
    return nc.BorderType.Const(b[1])


def padding_mode(m):
    if m == "zeros":
        return nc.PaddingMode.Zeros
    if m == "reflect":
        return nc.PaddingMode.Reflect
    if m == "replicate":
        return nc.PaddingMode.Replicate
    if m == "circular":
        return nc.PaddingMode.Circular
    if m[0] == "const":
        return nc.PaddingMode.Const(m[1])
    if m[0] == "custom":
        return nc.PaddingMode.Custom([border(b) for b in m[1]])
    return nc.PaddingMode.Explicit([[border(a), border(b)] for a, b in m[1]])


BORDER_CHOICES = ["zeros", "reflect", "replicate", "circular", "const"]


def random_border(rng, cval):
    b = BORDER_CHOICES[int(rng.integers(0, len(BORDER_CHOICES)))]
    return ("const", cval) if b == "const" else b


def random_case(rng, n, cval, max_x=9, max_k=4, allow_stride=True, signed=True):
    """(xshape, kshape, dilation, reverse, conv mode, padding mode) in the
    oracle's representation; pads may exceed the axis length.  Per-side
    constants are cval and -cval (cval twice for unsigned element types)."""
    neg = -cval if signed else cval
    xs = [int(v) for v in rng.integers(1, max_x + 1, n)]
    ks = [int(v) for v in rng.integers(1, max_k + 1, n)]
    dil = [int(v) for v in rng.integers(1, 3, n)]
    rev = bool(rng.integers(0, 2))
    r = int(rng.integers(0, 5))
    if r == 0:
        mode = "full"
    elif r == 1:
        mode = "same"
    elif r == 2:
        mode = "valid"
    elif r == 3:
        s = [int(v) for v in rng.integers(1, 3 if allow_stride else 2, n)]
        mode = ("custom", [int(v) for v in rng.integers(0, 6, n)], s)
    else:
        s = [int(v) for v in rng.integers(1, 3 if allow_stride else 2, n)]
        mode = ("explicit", [[int(a), int(b)] for a, b in rng.integers(0, 7, (n, 2))], s)
    r = int(rng.integers(0, 7))
    if r == 0:
        pm = "zeros"
    elif r == 1:
        pm = ("const", cval)
    elif r == 2:
        pm = "reflect"
    elif r == 3:
        pm = "replicate"
    elif r == 4:
        pm = "circular"
    elif r == 5:
        pm = ("custom", [random_border(rng, cval) for _ in range(n)])
    else:
        pm = ("explicit", [[random_border(rng, cval), random_border(rng, neg)] for _ in range(n)])
    return xs, ks, dil, rev, mode, pm


def rand_array(rng, shape, dtype):
    dt = np.dtype(dtype)
    if dt.kind == "c":
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)
    if dt.kind == "f":
        return rng.standard_normal(shape).astype(dt)
    info = np.iinfo(dt)
    lo, hi = max(info.min, -100), min(info.max, 100)
    a = rng.integers(lo, hi + 1, size=shape).astype(dt)
    # sprinkle zeros (zero taps are skipped by the reference)
    return np.where(rng.random(shape) < 0.1, dt.type(0), a)
