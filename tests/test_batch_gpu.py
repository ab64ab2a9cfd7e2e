"""The batched rank-1 path on the GPU: every pair must match the FP64 oracle's
conv_fft of that pair within 1e-5 normwise relative error (the reference's
TOLERANCE_F32, conv_fft/tests.rs:15), whichever form computes it: FP32 direct,
the BF16x3 / BF16x6 matrix-core direct forms, FP32 overlap-save FFT, the
autotuned mixes, or the unbatched per-pair path."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from modes import conv_mode, padding_mode, random_border
from oracle import ndconv_oracle as O

import ndconv as nc

pytestmark = pytest.mark.gpu

TOL = 1e-5   # every form, the BF16x3 and BF16x6 matrix-core forms included
ALGOS = ["auto", "auto_bf16x3", "fft", "direct", "bf16x3", "bf16x6", "plain"]


def oracle_pairs(xs, ks, mode, pm, dil, rev):
    return [O.conv_fft(x.astype(np.float64), k.astype(np.float64), mode, pm, dil, rev) for x, k in zip(xs, ks)]


def check(xs, ks, mode, pm, dil=1, rev=True, algos=ALGOS, tol=TOL):
    refs = oracle_pairs(xs, ks, mode, pm, dil, rev)   # the oracle runs first: it raises on invalid input
    worst = {}
    for algo in algos:
        got = nc.conv_fft_batch(xs, ks, conv_mode(mode), padding_mode(pm), dil, rev, algo)
        assert len(got) == len(refs)
        w = 0.0
        for i, (g, r) in enumerate(zip(got, refs)):
            assert g.shape == r.shape, (algo, i)
            e = O.relerr(g, r)[0]
            assert e <= tol, (algo, i, len(xs[i]), len(ks[i]), e)
            w = max(w, e)
        worst[algo] = w
    return worst


@pytest.mark.parametrize("pm", ["circular", "zeros", "replicate", "reflect", ("const", 0.75)])
def test_batch_same_mode_borders(gpu, pm):
    rng = np.random.default_rng(31)
    n = rng.integers(1, 3000, 40)
    k = rng.integers(1, 500, 40)
    if pm == "reflect":
        k = np.minimum(k, 2 * n - 1)   # reflect needs pads shorter than the signal
    xs = [rng.uniform(-1, 1, int(a)).astype(np.float32) for a in n]
    ks = [rng.uniform(-1, 1, int(b)).astype(np.float32) for b in k]
    print(pm, check(xs, ks, "same", pm))


def test_batch_front_pad_longer_than_signal(gpu):
    # Same + Circular with kernels much longer than the signal: the front pad
    # wraps once and the cells before p0 - n stay 0
    rng = np.random.default_rng(32)
    n = rng.integers(1, 40, 30)
    k = rng.integers(60, 1023, 30)
    xs = [rng.uniform(-1, 1, int(a)).astype(np.float32) for a in n]
    ks = [rng.uniform(-1, 1, int(b)).astype(np.float32) for b in k]
    print(check(xs, ks, "same", "circular"))


@pytest.mark.parametrize("seed", range(4))
def test_batch_random_modes(gpu, seed):
    rng = np.random.default_rng(100 + seed)
    for case in range(4):
        count = int(rng.integers(1, 25))
        dil = int(rng.integers(1, 4))
        rev = bool(rng.integers(0, 2))
        s = int(rng.integers(1, 4))
        kmax = int(rng.integers(2, 300))
        n = rng.integers(1, 4000, count)
        k = rng.integers(1, kmax, count)
        choice = int(rng.integers(0, 5))
        if choice == 0:
            mode = "full"
        elif choice == 1:
            mode = "valid"
            n = np.maximum(n, (k - 1) * dil + 1)
        elif choice == 2:
            mode = "same"
        elif choice == 3:
            mode = ("custom", [int(rng.integers(0, 200))], [s])
        else:
            mode = ("explicit", [[int(rng.integers(0, 300)), int(rng.integers(0, 300))]], [s])
        pm = ("explicit", [[random_border(rng, 0.5), random_border(rng, -0.25)]])
        xs = [rng.uniform(-1, 1, int(a)).astype(np.float32) for a in n]
        ks = [rng.uniform(-1, 1, int(b)).astype(np.float32) for b in k]
        try:
            oracle_pairs(xs, ks, mode, pm, dil, rev)
        except O.OracleError as e:
            with pytest.raises(getattr(nc, e.kind)):
                nc.conv_fft_batch(xs, ks, conv_mode(mode), padding_mode(pm), dil, rev)
            continue
        print(seed, case, mode, check(xs, ks, mode, pm, dil, rev))


def test_batch_long_signals_and_kernels(gpu):
    # the largest FFT block (8192 samples), kernels up to 1023 taps and
    # signals needing hundreds of blocks
    rng = np.random.default_rng(33)
    n = np.array([262144, 100000, 65537, 8191, 300, 20000])
    k = np.array([1023, 1000, 512, 700, 1023, 3])
    xs = [rng.uniform(-1, 1, int(a)).astype(np.float32) for a in n]
    ks = [rng.uniform(-1, 1, int(b)).astype(np.float32) for b in k]
    print(check(xs, ks, "same", "circular", algos=["auto", "fft", "direct", "bf16x3", "bf16x6"]))


@pytest.mark.parametrize("dil", [2, 3, 5])
def test_batch_dilated_matrix_core(gpu, dil):
    # dilated kernels up to the longest dilated length of a matrix-core item
    # (1024), so their taps are staged in up to three rounds, both orientations
    rng = np.random.default_rng(40 + dil)
    kmax = (1024 - 1) // dil + 1
    k = np.concatenate([[kmax, kmax - 1, 2, 1], rng.integers(1, kmax + 1, 8)])
    n = rng.integers(4200, 20000, len(k))
    xs = [rng.uniform(-1, 1, int(a)).astype(np.float32) for a in n]
    ks = [rng.uniform(-1, 1, int(b)).astype(np.float32) for b in k]
    for rev in (True, False):
        for mode, pm in (("same", "circular"), ("full", "zeros")):
            print(dil, rev, mode, check(xs, ks, mode, pm, dil, rev, algos=["bf16x3", "bf16x6", "auto_bf16x3"]))


def test_batch_errors_follow_the_reference(gpu):
    x = [np.ones(10, np.float32), np.ones(0, np.float32)]
    k = [np.ones(3, np.float32), np.ones(3, np.float32)]
    with pytest.raises(nc.DataShape):
        nc.conv_fft_batch(x, k, nc.ConvMode.Same, nc.PaddingMode.Zeros)
    x = [np.ones(10, np.float32), np.ones(5, np.float32)]
    k = [np.ones(3, np.float32), np.ones(0, np.float32)]
    with pytest.raises(nc.DataShape):          # conv_fft reports an empty kernel as DataShape
        nc.conv_fft_batch(x, k, nc.ConvMode.Same, nc.PaddingMode.Zeros)
    k = [np.ones(3, np.float32), np.ones(20, np.float32)]
    with pytest.raises(nc.MismatchShape):
        nc.conv_fft_batch(x, k, nc.ConvMode.Valid, nc.PaddingMode.Zeros)
    # reflect: a front pad of at least n + back pad panics in the reference
    x = [np.ones(10, np.float32), np.ones(1, np.float32)]
    with pytest.raises(O.OracleError):
        O.conv_fft(x[1], k[1], "same", "reflect")
    with pytest.raises(nc.InvalidArgument):
        nc.conv_fft_batch(x, k, nc.ConvMode.Same, nc.PaddingMode.Reflect)
    assert nc.conv_fft_batch([], [], nc.ConvMode.Same, nc.PaddingMode.Zeros) == []


def test_uneven_batch_full_size(gpu):
    """The 4096-pair uneven batch (seed 5): every pair of every form against
    the FP64 oracle at 1e-5, and graph replays against eager runs."""
    import torch
    from bench import uneven_batch
    n, k, x, kv = uneven_batch(5, 4096)
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev)
    kt = torch.from_numpy(kv).to(dev)
    outs = {}
    for algo in ["auto_bf16x3", "fft", "direct", "bf16x3", "bf16x6", "auto"]:
        bp = nc.BatchPlan(n, k, nc.ConvMode.Same, nc.PaddingMode.Circular, algo=algo)
        y = bp.run(xt, kt)
        torch.cuda.synchronize()
        outs[algo] = y.cpu().numpy()
        if algo == "auto_bf16x3":
            rep = bp.report()
            print({kk: rep[kk] for kk in ("items", "lines", "grid", "costs")})
            for r in range(3):
                yg = bp.run(xt, kt, graph=True)
                torch.cuda.synchronize()
                np.testing.assert_array_equal(yg.cpu().numpy(), outs["auto_bf16x3"], err_msg=f"replay {r}")
                bp.run(xt, kt)   # eager work between replays
            torch.cuda.synchronize()
            ylens, yoff = bp.y_lens, np.concatenate([[0], np.cumsum(bp.y_lens)])
        bp.close()
    xo = np.concatenate([[0], np.cumsum(n)])
    ko = np.concatenate([[0], np.cumsum(k)])
    algos = list(outs)

    def errors(i):
        ref = O.conv_fft(x[xo[i]:xo[i + 1]].astype(np.float64), kv[ko[i]:ko[i + 1]].astype(np.float64), "same",
                         "circular")
        assert len(ref) == ylens[i]
        return [O.relerr(outs[a][yoff[i]:yoff[i + 1]], ref)[0] for a in algos]

    with ThreadPoolExecutor(8) as ex:
        errs = np.array(list(ex.map(errors, range(len(n)))))
    for j, algo in enumerate(algos):
        i = int(np.argmax(errs[:, j]))
        print(f"{algo}: worst normwise rel err {errs[i, j]:.3e} (pair {i}: n {n[i]}, k {k[i]}), "
              f"median {np.median(errs[:, j]):.3e}")
        assert errs[i, j] <= TOL, (algo, i, int(n[i]), int(k[i]), errs[i, j])


@pytest.mark.parametrize("algo", ["auto_bf16x3", "fft", "direct", "bf16x3", "bf16x6"])
def test_batch_nonfinite_samples(gpu, algo):
    """A pair with an infinite or NaN sample: the reference's conv_fft of it
    (conv_fft/mod.rs:185-292) transforms every sample into every bin, so its
    whole output is NaN.  The batch computes each output from a window of the
    padded signal, so at least every output whose window reaches the sample
    is non-finite; and the other pairs of the batch stay exact."""
    rng = np.random.default_rng(41)
    n = [3000, 700, 5000, 257, 9000, 1500]
    k = [31, 200, 3, 64, 501, 17]
    xs = [rng.uniform(-1, 1, a).astype(np.float32) for a in n]
    ks = [rng.uniform(-1, 1, b).astype(np.float32) for b in k]
    bad = {0: (1200, np.inf), 2: (4000, np.nan), 4: (10, -np.inf)}
    for i, (pos, v) in bad.items():
        xs[i][pos] = v
    got = nc.conv_fft_batch(xs, ks, nc.ConvMode.Same, nc.PaddingMode.Circular, algo=algo)
    for i in range(len(n)):
        if i in bad:
            pos = bad[i][0]
            kl = k[i]
            p0 = (kl - 1) // 2 + ((kl - 1) % 2)       # Same: front pad
            # output o reads padded positions o .. o + k - 1, i.e. samples o - p0 .. o - p0 + k - 1
            lo, hi = max(0, pos + p0 - kl + 1), min(n[i] - 1, pos + p0)
            assert not np.isfinite(got[i][lo:hi + 1]).any(), (algo, i)
        else:
            ref = O.conv_fft(xs[i].astype(np.float64), ks[i].astype(np.float64), "same", "circular")
            assert O.relerr(got[i], ref)[0] <= TOL, (algo, i)
