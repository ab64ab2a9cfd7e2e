"""The C-ABI boundary on the CPU: the library loads, exports every symbol the
header declares, its struct layouts match the binding, and the host-side
plan (shape arithmetic, error precedence, border index maps) agrees with the
oracle.  No kernel is launched here."""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import ndconv_oracle as O

import ndconv as nc
from ndconv import _lib as C

HEADER = os.path.join(ROOT, "include", "ndconv.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ndc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    names = _declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/ndconv.h but not exported"
    assert lib.ndc_abi_version() == 1


def test_struct_layouts_match_binding():
    lib = C.lib()
    for which, st in enumerate([C.ConvDesc, C.PlanInfo, C.ConvModeC, C.PaddingModeC, C.Border, C.Value,
                                C.BatchDesc]):
        assert lib.ndc_struct_size(which) == ctypes.sizeof(st), st.__name__
    assert lib.ndc_struct_size(99) == -1


def test_create_without_device_reports_no_device():
    # this container has no GPU: creating a handle must fail cleanly
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present")
    h = ctypes.c_void_p()
    assert C.lib().ndc_create(ctypes.byref(h)) == C.ERR_NO_DEVICE


# --- shape arithmetic ------------------------------------------------------------

def _modes(n, rng):
    yield nc.ConvMode.Full, "full"
    yield nc.ConvMode.Same, "same"
    yield nc.ConvMode.Valid, "valid"
    p = [int(v) for v in rng.integers(0, 6, n)]
    s = [int(v) for v in rng.integers(1, 4, n)]
    yield nc.ConvMode.Custom(p, s), ("custom", p, s)
    pe = [[int(a), int(b)] for a, b in rng.integers(0, 6, (n, 2))]
    yield nc.ConvMode.Explicit(pe, s), ("explicit", pe, s)


def test_plan_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(60):
        n = int(rng.integers(1, 4))
        xs = [int(v) for v in rng.integers(1, 40, n)]
        ks = [int(v) for v in rng.integers(1, 7, n)]
        dil = [int(v) for v in rng.integers(1, 4, n)]
        for mode, omode in _modes(n, rng):
            opl = O.conv_fft_plan(xs, ks, omode, dil)
            mismatch = any(opl["kdil"][i] > opl["pds"][i] for i in range(n))
            for fft in (False, True):
                if mismatch:
                    with pytest.raises(nc.MismatchShape):
                        nc.plan(xs, ks, mode, nc.PaddingMode.Zeros, dil, fft=fft)
                    continue
                p = nc.plan(xs, ks, mode, nc.PaddingMode.Zeros, dil, fft=fft)
                assert p["padding"] == opl["padding"]
                assert p["strides"] == opl["strides"]
                assert p["kernel_dilated"] == opl["kdil"]
                assert p["padded"] == opl["pds"]
                assert p["out_shape"] == opl["out_shape"]
                if fft:
                    assert p["fft_size"] == opl["fft_size"]


def test_plan_config_sizes():
    p = nc.plan([200, 5000], [11, 31], nc.ConvMode.Same, nc.PaddingMode.Custom(
        [nc.BorderType.Reflect, nc.BorderType.Circular]), 2, fft=True)
    assert p["fft_size"] == [256, 5120] and p["out_shape"] == [200, 5000]
    assert p["padding"] == [[10, 10], [30, 30]]
    p = nc.plan([10, 100, 200], [5, 11, 31], nc.ConvMode.Full, fft=True)
    assert p["fft_size"] == [18, 128, 288] and p["out_shape"] == [14, 110, 230]
    p = nc.plan([4096, 4096], [31, 31], nc.ConvMode.Custom([0, 0], [2, 2]), nc.PaddingMode.Replicate, 2)
    assert p["out_shape"] == [2018, 2018]


def test_good_size_through_the_library():
    for n in [1, 2, 3, 5, 7, 385, 425, 426, 1537, 5030, 98305, 218450, 262143, 294912]:
        p = nc.plan([n], [1], nc.ConvMode.Valid, fft=True)
        assert p["fft_size"] == [O.good_size(n)], n


# --- border maps -----------------------------------------------------------------

BORDERS = [("zeros", nc.BorderType.Zeros), (("const", -7), nc.BorderType.Const(-7)),
           ("reflect", nc.BorderType.Reflect), ("replicate", nc.BorderType.Replicate),
           ("circular", nc.BorderType.Circular)]


def _expected_map_values(n, p0, p1, fb, bb, fft_len=None):
    """Oracle padding of x = [1..n] with distinct per-side constants gives the
    value each padded cell receives; the map must reproduce it."""
    fo = ("const", -1000) if fb[0] in ("zeros",) or (isinstance(fb[0], tuple)) else fb[0]
    bo = ("const", -2000) if bb[0] in ("zeros",) or (isinstance(bb[0], tuple)) else bb[0]
    x = np.arange(1, n + 1, dtype=np.int64)
    v = O.padding(x, ("explicit", [[fo, bo]]), [[p0, p1]])
    if fft_len is not None:
        v = np.concatenate([v, np.full(fft_len - len(v), -3000)])
    return v


def _map_values(codes, fb, bb):
    out = []
    for c in codes:
        if c >= 0:
            out.append(int(c) + 1)
        elif c == -1:
            out.append(0)
        elif c == -2:
            out.append(-1000)
        elif c == -3:
            out.append(-2000)
        else:
            out.append(-3000)
    return np.array(out)


@pytest.mark.parametrize("fb,bb", list(itertools.product(BORDERS, BORDERS)),
                         ids=lambda b: str(b[0]))
def test_border_map_matches_oracle_1d(fb, bb):
    for n in [1, 2, 3, 5, 8]:
        for p0 in range(0, 12):
            for p1 in range(0, 12):
                mode = nc.PaddingMode.Explicit([[fb[1], bb[1]]])
                cm = nc.ConvMode.Explicit([[p0, p1]], [1])
                fo = ("const", -1000) if fb[0] == "zeros" or isinstance(fb[0], tuple) else fb[0]
                bo = ("const", -2000) if bb[0] == "zeros" or isinstance(bb[0], tuple) else bb[0]
                try:
                    exp = O.padding(np.arange(1, n + 1, dtype=np.int64), ("explicit", [[fo, bo]]), [[p0, p1]])
                except O.OracleError as e:
                    assert e.kind == O.INVALID_ARGUMENT
                    with pytest.raises(nc.InvalidArgument):
                        nc.border_map([n], [1], cm, mode, 0)
                    continue
                got = _map_values(nc.border_map([n], [1], cm, mode, 0), fb, bb)
                np.testing.assert_array_equal(got, exp, err_msg=f"n={n} p=({p0},{p1})")
                # FFT maps: same cells, then the zero tail up to the FFT size
                if n + p0 + p1 >= 1:
                    gf = _map_values(nc.border_map([n], [1], cm, mode, 0, fft=True), fb, bb)
                    fl = O.good_size(n + p0 + p1)
                    np.testing.assert_array_equal(gf[:n + p0 + p1], exp)
                    assert (gf[n + p0 + p1:] == -3000).all() and len(gf) == fl


# --- error precedence ---------------------------------------------------------------

def test_error_precedence():
    Z = nc.PaddingMode.Zeros
    with pytest.raises(nc.DataShape):
        nc.plan([0, 3], [2, 2], nc.ConvMode.Same, Z)
    with pytest.raises(nc.KernelShape):
        nc.plan([3, 3], [0, 2], nc.ConvMode.Same, Z)
    with pytest.raises(nc.DataShape):   # conv_fft reports a zero-size kernel as DataShape
        nc.plan([3, 3], [0, 2], nc.ConvMode.Same, Z, fft=True)
    with pytest.raises(nc.MismatchShape):
        nc.plan([3, 3], [4, 2], nc.ConvMode.Valid, Z)
    with pytest.raises(nc.InvalidArgument):
        nc.plan([3], [2], nc.ConvMode.Custom([1], [0]), Z)
    with pytest.raises(nc.InvalidArgument):
        nc.plan([3], [2], nc.ConvMode.Custom([1], [0]), Z, fft=True)
    # conv pads before checking shapes: a reflect panic wins over MismatchShape
    with pytest.raises(nc.InvalidArgument):
        nc.plan([3], [9], nc.ConvMode.Explicit([[5, 0]], [1]), nc.PaddingMode.Reflect)
    # conv_fft checks shapes first
    with pytest.raises(nc.MismatchShape):
        nc.plan([3], [9], nc.ConvMode.Explicit([[5, 0]], [1]), nc.PaddingMode.Reflect, fft=True)
    # dilation 0: fine for conv (kdil = 1), a panic for conv_fft
    assert nc.plan([3], [2], nc.ConvMode.Same, Z, dilation=0)["out_shape"] == [3]
    with pytest.raises(nc.InvalidArgument):
        nc.plan([3], [2], nc.ConvMode.Same, Z, dilation=0, fft=True)
    # integer dtypes are not FFT element types
    with pytest.raises(nc.Unsupported):
        nc.plan([3], [2], nc.ConvMode.Same, Z, fft=True, dtype=np.int32)
