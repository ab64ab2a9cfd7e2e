"""Writes the golden fixtures used by the oracle and parity tests.

reference_goldens.json
    The exact expected arrays that the reference crate's own unit tests
    assert, transcribed as data (inputs and expected outputs), each with the
    reference file:line it comes from.

torch_cases.json
    The reference's libtorch-checked cases (conv/tests.rs:36-547 and
    padding/mod.rs:763-861).  The reference computes the expected values with
    libtorch at test time; here they are regenerated once with this
    container's PyTorch on the CPU in float64 and stored as data.  Padding is
    applied explicitly with the reference's own unfold rule, so torch only
    supplies the correlation sums (and, for the padding cases, F.pad).

Run:  python tests/golden/make_fixtures.py
"""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# Golden arrays from the reference's tests
# ---------------------------------------------------------------------------

PADDING_GOLDENS = [
    # padding/mod.rs zeros_padding
    dict(src="padding/mod.rs:468-473", x=[1, 2, 3], ep=[[1, 1]], mode="zeros", expect=[0, 1, 2, 3, 0]),
    dict(src="padding/mod.rs:476-484", x=[[1, 2], [3, 4]], ep=[[1, 1], [1, 1]], mode="zeros",
         expect=[[0, 0, 0, 0], [0, 1, 2, 0], [0, 3, 4, 0], [0, 0, 0, 0]]),
    dict(src="padding/mod.rs:487-503", x=[[[1, 2]], [[3, 4]]], ep=[[1, 0], [0, 1], [1, 0]], mode="zeros",
         expect=[[[0, 0, 0], [0, 0, 0]], [[0, 1, 2], [0, 0, 0]], [[0, 3, 4], [0, 0, 0]]]),
    dict(src="padding/mod.rs:506-511", x=[1, 2, 3], ep=[[2, 1]], mode="zeros", expect=[0, 0, 1, 2, 3, 0]),
    # const_padding
    dict(src="padding/mod.rs:518-523", x=[1, 2, 3], ep=[[1, 1]], mode=["const", 7], expect=[7, 1, 2, 3, 7]),
    dict(src="padding/mod.rs:526-534", x=[[1, 2], [3, 4]], ep=[[1, 1], [1, 1]], mode=["const", 9],
         expect=[[9, 9, 9, 9], [9, 1, 2, 9], [9, 3, 4, 9], [9, 9, 9, 9]]),
    # replicate_padding
    dict(src="padding/mod.rs:541-546", x=[1, 2, 3], ep=[[1, 2]], mode="replicate", expect=[1, 1, 2, 3, 3, 3]),
    dict(src="padding/mod.rs:549-557", x=[[1, 2], [3, 4]], ep=[[1, 1], [1, 1]], mode="replicate",
         expect=[[1, 1, 2, 2], [1, 1, 2, 2], [3, 3, 4, 4], [3, 3, 4, 4]]),
    dict(src="padding/mod.rs:560-565", x=[1, 2], ep=[[3, 3]], mode="replicate", expect=[1, 1, 1, 1, 2, 2, 2, 2]),
    # reflect_padding
    dict(src="padding/mod.rs:572-577", x=[1, 2, 3, 4], ep=[[2, 2]], mode="reflect", expect=[3, 2, 1, 2, 3, 4, 3, 2]),
    dict(src="padding/mod.rs:580-593", x=[[1, 2, 3], [4, 5, 6]], ep=[[1, 1], [1, 1]], mode="reflect",
         expect=[[5, 4, 5, 6, 5], [2, 1, 2, 3, 2], [5, 4, 5, 6, 5], [2, 1, 2, 3, 2]]),
    # circular_padding
    dict(src="padding/mod.rs:600-605", x=[1, 2, 3, 4], ep=[[2, 2]], mode="circular", expect=[3, 4, 1, 2, 3, 4, 1, 2]),
    dict(src="padding/mod.rs:608-616", x=[[1, 2], [3, 4]], ep=[[1, 1], [1, 1]], mode="circular",
         expect=[[4, 3, 4, 3], [2, 1, 2, 1], [4, 3, 4, 3], [2, 1, 2, 1]]),
    dict(src="padding/mod.rs:619-625", x=[1, 2, 3], dtype="uint8", ep=[[1, 1]], mode="circular", expect=[3, 1, 2, 3, 1]),
    # custom_padding (pads come from ConvMode::Full with a 3x3 kernel: [[2,2],[2,2]])
    dict(src="padding/mod.rs:632-655", x=[[1, 2], [3, 4]], ep=[[2, 2], [2, 2]],
         mode=["custom", ["replicate", "circular"]],
         expect=[[1, 2, 1, 2, 1, 2], [1, 2, 1, 2, 1, 2], [1, 2, 1, 2, 1, 2],
                 [3, 4, 3, 4, 3, 4], [3, 4, 3, 4, 3, 4], [3, 4, 3, 4, 3, 4]]),
    dict(src="padding/mod.rs:658-681", x=[[1, 2], [3, 4]], ep=[[2, 2], [2, 2]],
         mode=["custom", ["reflect", ["const", 7]]],
         expect=[[7, 7, 0, 0, 7, 7], [7, 7, 3, 4, 7, 7], [7, 7, 1, 2, 7, 7],
                 [7, 7, 3, 4, 7, 7], [7, 7, 1, 2, 7, 7], [7, 7, 3, 4, 7, 7]]),
    # explicit_padding
    dict(src="padding/mod.rs:688-698", x=[1, 2, 3], ep=[[1, 2]],
         mode=["explicit", [[["const", 7], ["const", 9]]]], expect=[7, 1, 2, 3, 9, 9]),
    # edge_cases
    dict(src="padding/mod.rs:707-712", x=[1, 2, 3], ep=[[0, 0]], mode="zeros", expect=[1, 2, 3]),
    dict(src="padding/mod.rs:715-720", x=[42], ep=[[2, 2]], mode="replicate", expect=[42, 42, 42, 42, 42]),
]

# conv_fft/padding.rs: padding::data into an 8x8 FFT buffer, padding::kernel with dilation
FFT_BUFFER_GOLDENS = [
    dict(src="conv_fft/padding.rs:124-153", kind="data", x=[[1, 2], [3, 4]], ep=[[2, 2], [2, 2]],
         mode=["custom", [["const", 7], ["const", 8]]], fft_size=[8, 8],
         expect=[[8, 8, 7, 7, 8, 8, 0, 0], [8, 8, 7, 7, 8, 8, 0, 0], [8, 8, 1, 2, 8, 8, 0, 0],
                 [8, 8, 3, 4, 8, 8, 0, 0], [8, 8, 7, 7, 8, 8, 0, 0], [8, 8, 7, 7, 8, 8, 0, 0],
                 [0, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0, 0]]),
    dict(src="conv_fft/padding.rs:156-181", kind="kernel", k=[[1, 2, 3], [4, 5, 6], [7, 8, 9]],
         dilation=[2, 3], reverse=True, fft_size=[8, 8],
         expect=[[1, 0, 0, 2, 0, 0, 3, 0], [0] * 8, [4, 0, 0, 5, 0, 0, 6, 0], [0] * 8,
                 [7, 0, 0, 8, 0, 0, 9, 0], [0] * 8, [0] * 8, [0] * 8]),
]

OFFSET_GOLDENS = [
    dict(src="dilation/mod.rs:278-293", k=[1.0, 2.0, 3.0], dilation=[1], reverse=True, strides=[1],
         expect=[[0, 3.0], [1, 2.0], [2, 1.0]]),
    dict(src="dilation/mod.rs:296-311", k=[1.0, 2.0, 3.0], dilation=[2], reverse=True, strides=[1],
         expect=[[0, 3.0], [2, 2.0], [4, 1.0]]),
    dict(src="dilation/mod.rs:314-326", k=[1.0, 2.0, 3.0], dilation=[2], reverse=False, strides=[1],
         expect=[[0, 1.0], [2, 2.0], [4, 3.0]]),
    dict(src="dilation/mod.rs:329-345", k=[[1.0, 2.0], [3.0, 4.0]], dilation=[1, 1], reverse=True, strides=[10, 1],
         expect=[[0, 4.0], [1, 3.0], [10, 2.0], [11, 1.0]]),
    dict(src="dilation/mod.rs:348-365", k=[[1.0, 2.0], [3.0, 4.0]], dilation=[2, 3], reverse=True, strides=[10, 1],
         expect=[[0, 4.0], [3, 3.0], [20, 2.0], [23, 1.0]]),
    dict(src="dilation/mod.rs:368-376", k=[1.0, 0.0, 2.0, 0.0, 3.0], dilation=[1], reverse=True, strides=[1],
         expect_len=3),
    dict(src="dilation/mod.rs:385-394", k=[42.0], dilation=[5], reverse=True, strides=[1], expect=[[0, 42.0]]),
    dict(src="dilation/mod.rs:397-404", k=[0.0, 0.0, 0.0], dilation=[2], reverse=True, strides=[1], expect=[]),
]

CONV_GOLDENS = [
    dict(src="conv/tests.rs:556-563", x=[42], k=[2], mode="same", padmode="zeros", expect=[84]),
    dict(src="conv/tests.rs:566-573", x=[1, 2, 3, 4], k=[3], mode="same", padmode="zeros", expect=[3, 6, 9, 12]),
    dict(src="conv/tests.rs:576-583", x=[[1, 2], [3, 4]], k=[[1]], mode="same", padmode="zeros",
         expect=[[1, 2], [3, 4]]),
]

# conv_fft/tests.rs vs_conv: the FFT result rounded must equal direct conv.
# (case data only; the expected values are the direct convolution itself)
FFT_VS_CONV_CASES = [
    dict(src="conv_fft/tests.rs:152-170", x=[1, 2, 3, 4, 5, 6], k=[1, 1, 1, 1], dilation=2, mode="same",
         padmode="zeros", dtype="float32"),
    dict(src="conv_fft/tests.rs:240-258", x=[1, 2, 3, 4, 5], k=[1, 2, 1], mode="full", padmode="zeros",
         dtype="float64"),
    dict(src="conv_fft/tests.rs:261-279", x=[1, 2, 3, 4, 5, 6], k=[1, 1, 1], mode="valid", padmode="zeros",
         dtype="float32"),
    dict(src="conv_fft/tests.rs:288-306", x=[[1, 2], [3, 4], [5, 6], [7, 8], [9, 10], [11, 12]], k=[[1, 0], [3, 1]],
         mode="same", padmode="replicate", dtype="float64"),
    dict(src="conv_fft/tests.rs:309-337", x=[[1, 2], [3, 4]], k=[[1, 0], [3, 1]], dilation=2, reverse=False,
         mode=["custom", [3, 3], [2, 2]], padmode="replicate", dtype="float64"),
    dict(src="conv_fft/tests.rs:377-395", x=[[1, 2], [3, 4]], k=[[1, 1], [1, 1]], mode="full", padmode="zeros",
         dtype="float32"),
    dict(src="conv_fft/tests.rs:398-416", x=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], k=[[1, 1], [1, 1]], mode="valid",
         padmode="zeros", dtype="float64"),
    dict(src="conv_fft/tests.rs:425-446", x=[[[1, 2], [3, 4]], [[5, 6], [7, 8]]],
         k=[[[1, 1, 1], [1, 1, 1], [1, 1, 1]], [[1, 1, 1], [1, 1, 1], [1, 1, 1]]], mode="same", padmode="zeros",
         dtype="float32"),
    dict(src="conv_fft/tests.rs:510-528", x=[[[1, 2]], [[3, 4]]], k=[[[1, 1]], [[1, 1]]], mode="full",
         padmode="zeros", dtype="float64"),
    dict(src="conv_fft/tests.rs:531-549", x=[[[1, 2, 3], [4, 5, 6]], [[7, 8, 9], [10, 11, 12]]],
         k=[[[1, 1]], [[1, 1]]], mode="valid", padmode="zeros", dtype="float32"),
    dict(src="conv_fft/tests.rs:558-576", x=[[1, 2, 3], [4, 5, 6]], k=[[1, 1], [1, 1]], mode="same",
         padmode="replicate", dtype="float32"),
    dict(src="conv_fft/tests.rs:579-597", x=[[1, 2, 3], [4, 5, 6]], k=[[1, 1], [1, 1]], mode="same",
         padmode="zeros", dtype="float64"),
    dict(src="conv_fft/tests.rs:600-618", x=[[1, 2], [3, 4]], k=[[1, 1], [1, 1]], mode="full",
         padmode=["const", 7], dtype="float32"),
    dict(src="conv_fft/tests.rs:803-821", x=[42], k=[2], mode="same", padmode="zeros", dtype="float32"),
    dict(src="conv_fft/tests.rs:210-237", x=[0.0, 0.1, 0.3, 0.4] * 4, k=[0.1, 0.3, 0.6, 0.3, 0.1], mode="same",
         padmode="circular", dtype="float32", tol=1e-6, no_round=True),
]

# conv_fft/tests.rs complex cases (conv_fft against itself in the reference;
# here they are checked against the direct complex sum)
COMPLEX_CASES = [
    dict(src="conv_fft/tests.rs:173-207",
         x=[[1.0, 0.5], [2.0, -0.3], [3.0, 0.8], [4.0, -0.2], [5.0, 0.6], [6.0, -0.4]],
         k=[[1.0, 0.1], [1.0, -0.1], [1.0, 0.2], [1.0, -0.2]], xshape=[6], kshape=[4], dilation=2,
         mode="same", padmode="zeros"),
    dict(src="conv_fft/tests.rs:340-374",
         x=[[1.0, 0.2], [2.0, -0.3], [3.0, 0.5], [4.0, -0.1]],
         k=[[1.0, 0.1], [0.0, 0.2], [3.0, -0.2], [1.0, 0.15]], xshape=[2, 2], kshape=[2, 2], dilation=2,
         reverse=False, mode=["custom", [3, 3], [2, 2]], padmode="replicate"),
    dict(src="conv_fft/tests.rs:449-507",
         x=[[1.0, 0.3], [2.0, -0.2], [3.0, 0.5], [4.0, -0.4], [5.0, 0.1], [6.0, -0.3], [7.0, 0.6], [8.0, -0.1]],
         k=[[1.0, 0.05], [1.0, -0.05], [1.0, 0.1], [1.0, -0.1], [1.0, 0.15], [1.0, -0.15], [1.0, 0.2], [1.0, -0.2],
            [1.0, 0.05], [1.0, -0.05], [1.0, 0.1], [1.0, -0.1], [1.0, 0.15], [1.0, -0.15], [1.0, 0.2],
            [1.0, -0.2], [1.0, 0.05], [1.0, -0.05]],
         xshape=[2, 2, 2], kshape=[2, 3, 3], mode="same", padmode="zeros"),
]

# The doc examples (lib.rs:23-50, conv/mod.rs:102-109, dilation/mod.rs:90-102,140-154)
DOC_CASES = [
    dict(src="lib.rs:27-38", x=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], k=[[1, 1], [1, 1]], mode="same", padmode="zeros"),
    dict(src="conv/mod.rs:106-108", x=[[1, 2, 3], [4, 5, 6]], k=[[1, 1], [1, 1]], mode="same", padmode="zeros"),
    dict(src="dilation/mod.rs:94-101", x=[1, 2, 3, 4, 5], k=[1, 1, 1], dilation=2, mode="same", padmode="zeros"),
    dict(src="dilation/mod.rs:144-153", x=[1, 2, 3, 4, 5], k=[1, 2, 3], reverse=False, mode="same",
         padmode="zeros"),
]


# ---------------------------------------------------------------------------
# libtorch-checked cases regenerated with torch (CPU, float64)
# ---------------------------------------------------------------------------

TORCH_CONV_CASES = [
    # conv/tests.rs vs_torch: (x, k, dilation, reverse, mode)
    dict(src="conv/tests.rs:45-66", x=[1, 2, 3, 4, 5], k=[1, 2, 1], mode="full"),
    dict(src="conv/tests.rs:69-90", x=[[1, 2], [3, 4]], k=[[1, 1], [1, 1]], mode="full"),
    dict(src="conv/tests.rs:93-114", x=[[[1, 2]], [[3, 4]]], k=[[[1, 1]], [[1, 1]]], mode="full"),
    dict(src="conv/tests.rs:123-143", x=[1, 2, 3, 4, 5], k=[1, 2, 1], mode="same"),
    dict(src="conv/tests.rs:146-176", x=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], k=[[1, 0, -1], [2, 0, -2], [1, 0, -1]],
         mode="same"),
    dict(src="conv/tests.rs:179-202", x=[[[1, 2, 3]], [[4, 5, 6]]], k=[[[1, 1, 1]]], mode="same"),
    dict(src="conv/tests.rs:211-231", x=[1, 2, 3, 4, 5], k=[1, 2, 1], mode="valid"),
    dict(src="conv/tests.rs:234-254", x=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], k=[[1, 1], [1, 1]], mode="valid"),
    dict(src="conv/tests.rs:257-277", x=[[[1, 2, 3], [4, 5, 6]], [[7, 8, 9], [10, 11, 12]]], k=[[[1, 1]], [[1, 1]]],
         mode="valid"),
    dict(src="conv/tests.rs:286-313", x=[1, 2, 3, 4, 5, 6], k=[1, 1, 1], mode=["custom", [1], [2]]),
    dict(src="conv/tests.rs:316-343", x=[[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12]], k=[[1, 1], [1, 1]],
         mode=["custom", [1, 1], [2, 2]]),
    dict(src="conv/tests.rs:346-373", x=[1, 2, 3, 4, 5, 6, 7, 8, 9], k=[1, 2, 1], mode=["custom", [2], [3]]),
    dict(src="conv/tests.rs:382-409", x=[1, 2, 3, 4, 5, 6], k=[1, 1, 2], dilation=2, reverse=False,
         mode=["custom", [4], [2]]),
    dict(src="conv/tests.rs:412-436", x=[[1, 1, 1], [1, 1, 1], [1, 1, 2]], k=[[2, 1, 1], [1, 1, 1]], dilation=2,
         reverse=False, mode="same"),
    dict(src="conv/tests.rs:439-469", x=[[[1, 2], [3, 4]], [[5, 6], [7, 8]]],
         k=[[[1, 1, 1], [1, 1, 1], [1, 1, 1]], [[1, 1, 1], [1, 1, 1], [1, 1, 1]]], dilation=2, reverse=False,
         mode=["custom", [2, 2, 2], [1, 2, 1]]),
    dict(src="conv/tests.rs:478-515", x=[1, 2, 3, 4, 5, 6], k=[1, 1, 2], dilation=2, reverse=True,
         mode=["custom", [4], [2]]),
    dict(src="conv/tests.rs:518-545", x=[1, 2, 3, 4, 5, 6], k=[1, 1, 2], dilation=2, reverse=False,
         mode=["custom", [4], [2]]),
]

TORCH_PAD_CASES = [
    # padding/mod.rs:763-801 aligned_with_libtorch: (array, kernel for unfold, ConvMode)
    dict(src="padding/mod.rs:766-778", x=[[[1, 2, 3], [3, 4, 5]], [[5, 6, 7], [7, 8, 9]]], kshape=[3, 3, 3],
         mode="same"),
    dict(src="padding/mod.rs:781-789", x=[[1, 2], [3, 4]], kshape=[2, 2], mode="full"),
    dict(src="padding/mod.rs:792-800", x=[1, 2, 3], kshape=[4], mode="same"),
]


def _unfold_pads(mode, kd):
    if mode == "full":
        return [[k - 1, k - 1] for k in kd], [1] * len(kd)
    if mode == "same":
        return [[(k - 1) // 2 + 1, (k - 1) // 2] if k % 2 == 0 else [(k - 1) // 2] * 2 for k in kd], [1] * len(kd)
    if mode == "valid":
        return [[0, 0] for _ in kd], [1] * len(kd)
    if mode[0] == "custom":
        return [[p, p] for p in mode[1]], list(mode[2])
    return [list(p) for p in mode[1]], list(mode[2])


def torch_conv_expected(case):
    import torch
    import torch.nn.functional as F
    x = torch.tensor(np.array(case["x"], dtype=np.float64))
    k = np.array(case["k"], dtype=np.float64)
    n = x.dim()
    d = case.get("dilation", 1)
    dil = [d] * n if isinstance(d, int) else list(d)
    reverse = case.get("reverse", True)
    if reverse:
        k = k[tuple(slice(None, None, -1) for _ in range(n))].copy()
    kd = [k.shape[i] * dil[i] - dil[i] + 1 for i in range(n)]
    pads, strides = _unfold_pads(case["mode"], kd)
    flat = []
    for i in reversed(range(n)):
        flat += [pads[i][0], pads[i][1]]
    xp = F.pad(x.reshape([1, 1] + list(x.shape)), flat, mode="constant", value=0.0)
    w = torch.tensor(k).reshape([1, 1] + list(k.shape))
    fn = {1: F.conv1d, 2: F.conv2d, 3: F.conv3d}[n]
    y = fn(xp, w, stride=strides, dilation=dil)
    return y.reshape(y.shape[2:]).numpy().round().astype(np.int64).tolist()


def torch_pad_expected(case, mode):
    import torch
    import torch.nn.functional as F
    x = torch.tensor(np.array(case["x"], dtype=np.float64))
    n = x.dim()
    kd = case["kshape"]
    pads, _ = _unfold_pads(case["mode"], kd)
    flat = []
    for i in reversed(range(n)):
        flat += [pads[i][0], pads[i][1]]
    t = x.reshape([1, 1] + list(x.shape))
    if mode == "zeros":
        y = F.pad(t, flat, mode="constant", value=0.0)
    elif mode == "const7":
        y = F.pad(t, flat, mode="constant", value=7.0)
    else:
        y = F.pad(t, flat, mode=mode)
    return pads, y.reshape(y.shape[2:]).numpy().round().astype(np.int64).tolist()


def main():
    goldens = dict(padding=PADDING_GOLDENS, fft_buffer=FFT_BUFFER_GOLDENS, offsets=OFFSET_GOLDENS,
                   conv=CONV_GOLDENS, fft_vs_conv=FFT_VS_CONV_CASES, complex_cases=COMPLEX_CASES,
                   doc_cases=DOC_CASES)
    with open(os.path.join(HERE, "reference_goldens.json"), "w") as f:
        json.dump(goldens, f, indent=1)

    conv_cases = []
    for c in TORCH_CONV_CASES:
        c = dict(c)
        c["expect"] = torch_conv_expected(c)
        conv_cases.append(c)
    pad_cases = []
    for c in TORCH_PAD_CASES:
        for m in ["zeros", "const7", "replicate", "reflect", "circular"]:
            pads, exp = torch_pad_expected(c, m)
            pad_cases.append(dict(src=c["src"], x=c["x"], ep=pads,
                                  mode={"zeros": "zeros", "const7": ["const", 7]}.get(m, m), expect=exp))
    with open(os.path.join(HERE, "torch_cases.json"), "w") as f:
        json.dump(dict(conv=conv_cases, padding=pad_cases), f, indent=1)
    print("wrote", len(conv_cases), "torch conv cases and", len(pad_cases), "torch padding cases")


if __name__ == "__main__":
    main()
