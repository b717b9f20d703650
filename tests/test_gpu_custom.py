"""User-defined device integrands (include/viltrum_amd/device.hpp) in a program written
the way the reference's users write integrands (tests/cpp/custom.hip): a functor equal
to a built-in polynomial gives that family's bins and training pairs bit for bit (FP64
reference order, FP32, and 4 dims through the region-major kernel); lambdas in the
adapter's call forms f(x, y) and f(std::array) give the closed-form bins; a batch of one
functor type gives its single calls; as the inner integrand of the Fubini form a functor
gives the equal built-in polynomial's bins bit for bit; over an infinite range a functor
of a sequence gives the built-in Dyson series' bins bit for bit, and the reference's own
documented series its closed form; forms and region dims a hook was not built for, and
bins of a type not instantiated, are refused."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "custom")


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    assert os.path.exists(EXE), "build tests/cpp/custom (python tests/cpp/build.py)"
    d = str(tmp_path_factory.mktemp("custom"))
    out = subprocess.run([EXE, d], capture_output=True, text=True, timeout=300, check=True).stdout
    lines = dict(line.split(" ", 1) for line in out.strip().split("\n"))
    return d, lines


def _bins(d, name, dtype, shape):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dtype).reshape(shape)


def _edges(n):
    return np.arange(n + 1, dtype=np.float64) / n


@pytest.mark.parametrize("case", ["poly2_f64", "poly2_f32", "poly4_f32", "batch3_f32", "fubini3_f32", "dyson_seq2_f32",
                                  "dyson_seq4_f32", "dyson_seq1_opt_f32", "seq_refused", "unsupported"])
def test_flags(run, case):
    assert run[1][case].split()[0] == "1", case


def test_training_pairs_match_builtin(run):
    ok, n = run[1]["pairs_f32"].split()
    assert ok == "1" and int(n) > 255


def test_quadratic_lambda_is_exact(run):
    """(1 + x - x^2/2)(2 - y + 3y^2): the control variate interpolates it exactly, so every
    bin is the closed form up to rounding whatever the residual samples."""
    b = _bins(run[0], "quad2_f64", np.float64, (16, 16))
    e = _edges(16)
    gx = lambda x: x + x ** 2 / 2 - x ** 3 / 6  # noqa: E731
    gy = lambda y: 2 * y - y ** 2 / 2 + y ** 3  # noqa: E731
    truth = np.outer(np.diff(gx(e)), np.diff(gy(e))) * 256
    assert np.max(np.abs(b - truth)) < 1e-12 * np.max(np.abs(truth))


def test_transcendental_lambda_matches_closed_form(run):
    """sin(3x) e^y, FP32, 1024 regions and 16 samples per bin: the estimator's per-bin
    error (FP32 rounding and the residual's noise) measured 1.0e-7 of the scale, mean
    -3e-9; the bounds leave a factor of 20."""
    b = _bins(run[0], "trans2_f32", np.float32, (32, 32)).astype(np.float64)
    e = _edges(32)
    truth = np.outer(np.diff(-np.cos(3 * e) / 3), np.diff(np.exp(e))) * 1024
    assert np.all(np.isfinite(b))
    scale = np.max(np.abs(truth))
    rmse, bias = np.sqrt(np.mean((b - truth) ** 2)) / scale, np.mean(b - truth) / scale
    print("trans2_f32 relative rmse %.3g, mean error %.3g" % (rmse, bias))
    assert rmse < 2e-6 and abs(bias) < 1e-7, (rmse, bias)


def test_array_form_three_dims(run):
    """f(std::array) = x y + 0.75 t over three dims with two bin dims: quadratic, so the
    control variate is exact up to FP32 rounding."""
    b = _bins(run[0], "arr3_f32", np.float32, (12, 10)).astype(np.float64)
    ex, ey = _edges(12), _edges(10)
    truth = (np.outer(np.diff(ex ** 2 / 2), np.diff(ey ** 2 / 2)) + 0.375 * np.outer(np.diff(ex), np.diff(ey))) * 120
    assert np.max(np.abs(b - truth)) < 2e-6 * np.max(np.abs(truth))


def test_reference_series_of_a_sequence_matches_closed_form(run):
    """The reference's documented integrand of an infinite range (a series cut by Russian
    roulette, decaying factor d = 0.5) over the primary infinite range, regions over entry
    0, 16 bins: E[f | x0] = [x0 < d] / (1 - d), so bins 0..7 hold 2 and bins 8..15 hold 0
    (every region node there is 0, so the control variate and the residual vanish).  The
    series' sd given x0 < d is 2.78 (simulated, 2e6 draws); 4096 residual samples per bin
    give 0.043 per bin: the bounds are about 6 sd per bin and 5 sd for the mean of eight."""
    b = _bins(run[0], "series_seq1", np.float32, (16,)).astype(np.float64)
    print("series_seq1", np.round(b, 3))
    assert np.all(np.isfinite(b))
    assert np.max(np.abs(b[:8] - 2.0)) < 0.25 and abs(np.mean(b[:8]) - 2.0) < 0.08
    assert np.max(np.abs(b[8:])) < 0.05
