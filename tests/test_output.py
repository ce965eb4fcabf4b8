"""Candidate properties and the value(error) text of the ACCEL files
(host code, no GPU needed).

calc_props / calc_rzwerrs of the library against the oracle's restatement
(src/characteristics.c:195-396; same formulas, so equal to the last bit)
and closed forms; nice_output against values worked by hand from the rules
of src/output.c:6-216 (journal notation: value rounded at the error's
leading digit(s), error digits in parentheses, scientific form when the
error's exponent requires it)."""
import numpy as np
import pytest

from oracle import oracle as O

A = pytest.importorskip("presto_amd.accel")


@pytest.mark.parametrize("val,err,numerr,expect", [
    (12.3456, 0.0123, 1, " 12.35(1)"),
    (12.3456, 0.0123, 2, " 12.346(12)"),
    (1234.5678, 0.5, 1, " 1234.6(5)"),
    (123456.0, 2345.0, 1, " 1.23(2)x10^5"),
    (-3.3, 0.25, 2, "-3.30(25)"),
    (0.00012345, 0.0000031, 2, " 0.0001235(31)"),
    (5e-7, 3e-9, 1, " 5.00(3)x10^-7"),
    (0.5, 0.0999, 1, " 0.5(1)"),      # 0.999 rounds up to the next decade
    (7.0, 0.0, 1, " 7(0)"),
])
def test_nice_output_known_answers(val, err, numerr, expect):
    assert A.nice_output(val, err, numerr) == expect


def test_nice_output_right_justifies():
    assert A.nice_output(12.3456, 0.0123, 1, 14) == " " * 5 + " 12.35(1)"


def _rd(rng):
    return O.ORDerivs(pow=float(rng.uniform(10, 1e4)), phs=float(rng.uniform(0, 6.28)),
                      dpow=float(rng.normal()), dphs=float(rng.normal(-3, 1)),
                      d2pow=float(-rng.uniform(1e2, 1e5)), d2phs=float(rng.normal()),
                      locpow=float(rng.uniform(0.5, 3.0)))


def test_calc_props_and_rzwerrs_equal_oracle():
    rng = np.random.default_rng(3)
    for _ in range(200):
        o = _rd(rng)
        d = A.RDerivs(*[getattr(o, f) for f, _ in O.ORDerivs._fields_])
        r, z, w = rng.uniform(100, 1e6), rng.uniform(-200, 200), rng.choice([0.0, 37.5])
        po, pa = O.calc_props(o, r, z, w), A.calc_props(d, r, z, w)
        for f, _ in O.OFourierProps._fields_:
            assert getattr(po, f) == getattr(pa, f), f
        T = rng.uniform(10, 1e4)
        eo, ea = O.calc_rzwerrs(po, T), A.calc_rzwerrs(pa, T)
        for f, _ in O.ORzwErrs._fields_:
            assert getattr(eo, f) == getattr(ea, f), f


def test_calc_props_rejects_zero_power():
    with pytest.raises(A.AccelError):
        A.calc_props(A.RDerivs(0.0, 0, 0, 0, 0, 0, 1.0), 10.0, 0.0, 0.0)
    with pytest.raises(A.AccelError):
        A.calc_props(A.RDerivs(5.0, 0, 0, 0, 0, 0, 0.0), 10.0, 0.0, 0.0)
