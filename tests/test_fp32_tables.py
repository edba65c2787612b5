"""The FP32 device path's fitted forms: the committed coefficient tables are
what their generators produce, and the generators' FP32 evaluations stay
within their stated error of the FP64 references (the oracle's NFW shape,
scipy's erfcx)."""
import os
import re
import sys

import numpy as np
import scipy.special as sp

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "y3_cluster_cpp_amd", "csrc")
sys.path.insert(0, CSRC)

import fit_erfcx  # noqa: E402
import fit_nfw_table  # noqa: E402

FLOAT = r"(-?\d+\.?\d*(?:e[-+]?\d+)?)f"


def _nfw_inc():
    with open(os.path.join(CSRC, "y3g_nfw_table.inc")) as f:
        text = f.read()
    knots = re.findall(r"\{" + FLOAT + r", " + FLOAT + r"\}", text)
    return np.array([[float(a), float(b)] for a, b in knots], dtype=np.float32)


def test_nfw_table_is_generator_output():
    tab = _nfw_inc()
    v, d = fit_nfw_table.knots()
    assert tab.shape == (len(v), 2)
    assert np.array_equal(tab[:, 0], v.astype(np.float32))
    assert np.array_equal(tab[:, 1], d.astype(np.float32))


def test_nfw_table_against_oracle():
    tab = _nfw_inc()
    xs = np.concatenate([np.geomspace(1e-5, 3e4, 1500), np.linspace(0.9, 1.1, 201)])
    got = np.array([fit_nfw_table.eval_f32(x * x, tab[:, 0], tab[:, 1]) for x in xs])
    ref = np.array([O.lib.y3o_nfw_g(x) for x in xs])
    assert np.max(np.abs(got / ref - 1)) < 1e-6


def test_nfw_generator_closed_form_is_oracle():
    xs = np.geomspace(1e-4, 1e4, 200)
    assert np.allclose(fit_nfw_table.g64(xs), [O.lib.y3o_nfw_g(x) for x in xs], rtol=1e-13, atol=0)


def test_erfcx_coefficients_in_device_header():
    with open(os.path.join(CSRC, "y3g_device.hpp")) as f:
        text = f.read()
    body = text[text.index("float erfcx_pos_f(float x)\n{"):]
    body = body[:body.index("return")]
    lits = [float(a) for a in re.findall(FLOAT, body)]
    # the literals are 2, 2 (t = (x - 2) / (x + 2)) then c10 ... c0 in Horner order
    coef = np.float32(fit_erfcx.coefficients())
    assert np.array_equal(np.float32(lits[2:]), np.abs(coef[::-1]))


def test_erfcx_fp32_accuracy():
    pc = fit_erfcx.coefficients()
    xs = np.concatenate([np.linspace(0, 10, 20001), np.geomspace(10, 1e7, 2000)])
    assert np.max(np.abs(fit_erfcx.erfcx_f32(xs, pc) / sp.erfcx(xs) - 1)) < 3e-7
