"""Rate expressions (Rate::Expr) on the CPU: the parser against the
reference's own parser tests, the postfix lowering against the reference's
conversion tree, the oracle's interpreter against the reference's evaluation
values and exact probability laws, exp/pow against mpmath and the C99
special cases, and the C ABI's validation (no device needed)."""

from __future__ import annotations

import ctypes
import json
import math
import warnings
from pathlib import Path

import numpy as np
import pytest

from rebop_amd import _native
from rebop_amd import expr as E
from rebop_amd.gillespie import Gillespie

GOLDEN = Path(__file__).resolve().parent / "golden"
TABLES = json.loads((GOLDEN / "reference_tables.json").read_text())
C, SP = _native.REBOP_EXPR_CONST, _native.REBOP_EXPR_SPECIES


# ---- parser (src/expr.rs:140-425) ---------------------------------------


@pytest.mark.parametrize("text,shown", TABLES["expr_parse"]["cases"])
def test_parse_and_display_match_reference_tests(text, shown):
    assert str(E.parse(text)) == shown


@pytest.mark.parametrize("text", TABLES["expr_parse"]["fail"] + ["2^3^4", "-infect", " 3", "3 ",
                                                                 "max (1,2)", "exp 3", "2e", "A +",
                                                                 "(A", "A)", "", "+A", "3 4"])
def test_parse_failures(text):
    """The reference's failures (expr.rs:403-408), and inputs its grammar
    rejects: chained powers, '-inf' eating a name's prefix, unconsumed
    whitespace, a space before '(' of max/exp, a dangling exponent."""
    with pytest.raises(E.RateParseError):
        E.parse(text)


@pytest.mark.parametrize("name", TABLES["expr_parse"]["variables"])
def test_float_like_names_are_variables(name):
    """expr.rs:410-424: `variable` is tried before `constant`."""
    node = E.parse(name)
    assert node.kind == "Variable" and node.name == name


def test_signed_literals_and_infinities():
    assert E.parse("-2").kind == "Constant" and E.parse("-2").value == -2.0
    assert E.parse("- 2").kind == "Neg"
    assert E.parse("-inf").value == -math.inf and str(E.parse("-inf")) == "-inf"
    assert E.parse("+infinity").value == math.inf
    assert str(E.parse("3 - -2")) == "(3 - -2)"
    assert str(E.parse("2*-3")) == "(2 * -3)"
    assert str(E.parse("1e-7")) == "0.0000001"
    assert str(E.parse("1.")) == "1"
    assert str(E.parse("-A^2")) == "(-(A ^ 2))"
    assert str(E.parse("(2^3)^4")) == "((2 ^ 3) ^ 4)"


# ---- PExpr::to_expr (src/expr.rs:57-105) --------------------------------


def _postfix_of_tree(tree, consts):
    kind = tree[0]
    if kind == "Constant":
        consts.append(tree[1])
        return [C | ((len(consts) - 1) << 8)]
    if kind == "Concentration":
        return [SP | (tree[1] << 8)]
    out = []
    for sub in tree[1:]:
        out += _postfix_of_tree(sub, consts)
    return out + [getattr(_native, "REBOP_EXPR_" + kind.upper())]


def test_conversion_matches_reference_tree():
    t = TABLES["expr_eval"]["conversion"]
    code, consts = E.to_postfix(E.parse(t["expr"]), t["species"], {})
    want_consts: list = []
    assert code == _postfix_of_tree(t["tree"], want_consts)
    assert consts == want_consts


def test_conversion_parameters_and_missing_parameter():
    code, consts = E.to_postfix(E.parse("k * A"), {"A": 0}, {"k": 0.4})
    assert code == [C | 0, SP | 0, _native.REBOP_EXPR_MUL] and consts == [0.4]
    with pytest.raises(ValueError, match="Parameter k should have a value"):
        E.to_postfix(E.parse("k * A"), {"A": 0}, {})


# ---- evaluation in the oracle (Expr::eval, src/expr.rs:24-38) -----------


def _oracle_expr_net(oracle, text, species, S):
    code, consts = E.to_postfix(E.parse(text), species, {})
    n = oracle.OracleNet(S)
    n.add_expr(code, consts, [0] * S)
    return n


def test_eval_reference_value(oracle):
    """expr.rs:481-499: the value Rust computes, to the last bit."""
    t = TABLES["expr_eval"]["conversion"]
    n = _oracle_expr_net(oracle, t["expr"], t["species"], 6)
    assert n.rate(0, t["counts"]) == t["value"]


@pytest.mark.parametrize("text,a,value", TABLES["expr_eval"]["max"])
def test_eval_max_reference_values(oracle, text, a, value):
    n = _oracle_expr_net(oracle, text, {"A": 0}, 1)
    assert n.rate(0, [a]) == value


def test_eval_operators_against_python(oracle):
    """Each operation on random counts equals Python's IEEE arithmetic."""
    rng = np.random.default_rng(5)
    sp = {"A": 0, "B": 1}
    for _ in range(200):
        a, b = (int(v) for v in rng.integers(-50, 200, 2))
        for text, f in [("A + B", lambda: a + b), ("A - B", lambda: float(a) - b),
                        ("A * B", lambda: float(a) * b), ("A / (B + 0.5)", lambda: a / (b + 0.5)),
                        ("-A", lambda: -float(a)), ("max(A, B)", lambda: float(max(a, b))),
                        ("min(A, B)", lambda: float(min(a, b))), ("A * 0.1 - B / 3", lambda: a * 0.1 - b / 3)]:
            n = _oracle_expr_net(oracle, text, sp, 2)
            assert n.rate(0, [a, b]) == f(), (text, a, b)


def test_malformed_programs_rejected_by_oracle(oracle):
    n = oracle.OracleNet(2)
    ADD = _native.REBOP_EXPR_ADD
    for code, consts in [([], []), ([ADD], []), ([C | 0], []), ([SP | (2 << 8)], []),
                         ([C | 0, C | 0], [1.0]), ([99], []), ([C | 0] * 33 + [ADD] * 32, [1.0])]:
        with pytest.raises(oracle.OracleError):
            n.add_expr(code, consts, [0, 0])


def test_expression_equal_to_lma_gives_identical_trajectories(oracle):
    """'0.5 * A' is the same f64 product as LMA(0.5, A): bit-identical paths."""
    x0 = np.array([40, 0], dtype=np.int64)
    a = oracle.OracleNet(2)
    a.add_lma(0.5, [1, 0], [-1, 1])
    a.add_lma(2.0, [0, 0], [1, 0])
    b = oracle.OracleNet(2)
    code, consts = E.to_postfix(E.parse("0.5 * A"), {"A": 0, "B": 1}, {})
    b.add_expr(code, consts, [-1, 1])
    code, consts = E.to_postfix(E.parse("2"), {"A": 0, "B": 1}, {})
    b.add_expr(code, consts, [1, 0])
    ra = a.run_grid(3, x0, 10.0, 20, n_traj=64)
    rb = b.run_grid(3, x0, 10.0, 20, n_traj=64)
    assert np.array_equal(ra[1], rb[1]) and ra[3] == rb[3]


def test_expression_birth_process_is_poisson(oracle):
    """tests/test_rebop.py:99-106 as a law: rate 'B' with B = 1 fixed makes
    A(t) ~ Poisson(t); rate '0.25 * B' with B = 2 gives Poisson(t / 2)."""
    from scipy import stats

    for text, b, lam in [("B", 1, 1.0), ("0.25 * B", 2, 0.5)]:
        n = oracle.OracleNet(2)
        code, consts = E.to_postfix(E.parse(text), {"A": 0, "B": 1}, {})
        n.add_expr(code, consts, [1, 0])
        _, out, _, _, _ = n.run_grid(17, np.array([0, b]), 10.0, 10, n_traj=20000, nthreads=8)
        for i in (2, 5, 10):
            a = out[i, 0]
            k = np.arange(a.max() + 1)
            exp_p = stats.poisson.pmf(k, lam * i)
            obs = np.bincount(a, minlength=len(k))
            keep = exp_p * len(a) >= 5
            chi = stats.chisquare(np.append(obs[keep], obs[~keep].sum()),
                                  np.append(exp_p[keep], 1 - exp_p[keep].sum()) * len(a))
            assert chi.pvalue > 1e-4, (text, i, chi)
        assert np.all(out[:, 1] == b)


# ---- exp / pow / max / min (rebop_expr.h) -------------------------------


def _ulps(got, ref):
    import mpmath as mp

    r = float(ref)
    if got == r:
        return 0.0
    return float(abs(mp.mpf(got) - ref) / mp.mpf(math.ulp(r)))


def test_exp64_accuracy_against_mpmath(oracle):
    import mpmath as mp

    mp.mp.prec = 120
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-708, 709.7, 3000), rng.uniform(-1, 1, 3000),
                         [0.5 * math.log(2), -0.5 * math.log(2), 1e-300, -1e-17, 700.0, -700.0]])
    worst = max(_ulps(oracle.exp64(x), mp.exp(mp.mpf(x))) for x in xs)
    assert worst <= 1.0


def test_exp64_special_values(oracle):
    assert oracle.exp64(0.0) == 1.0 and oracle.exp64(-0.0) == 1.0
    assert oracle.exp64(1.0) == math.e
    assert oracle.exp64(math.inf) == math.inf and oracle.exp64(-math.inf) == 0.0
    assert math.isnan(oracle.exp64(math.nan))
    assert oracle.exp64(709.79) == math.inf
    assert oracle.exp64(-745.2) == 0.0
    assert oracle.exp64(-745.13) == 5e-324  # the least subnormal
    assert oracle.exp64(7.0) == pytest.approx(1096.6331584284585, rel=2**-52)


def test_pow64_integer_powers_are_exact_when_representable(oracle):
    rng = np.random.default_rng(2)
    for _ in range(2000):
        x = int(rng.integers(-60, 61))
        n = int(rng.integers(0, 9))
        if x == 0 and n == 0:
            continue
        assert oracle.pow64(float(x), float(n)) == float(x ** n), (x, n)
    assert oracle.pow64(7.0, 11.0) == 1977326743.0
    assert oracle.pow64(2.0, -3.0) == 0.125


def test_pow64_accuracy_against_mpmath(oracle):
    """Non-integral exponents go through exp(y ln x): the error bound grows
    with |y ln x| (the rounding of the product)."""
    import mpmath as mp

    mp.mp.prec = 120
    rng = np.random.default_rng(3)
    for _ in range(3000):
        x = 10 ** rng.uniform(-5, 5)
        y = rng.uniform(-10, 10)
        ref = mp.power(mp.mpf(x), mp.mpf(y))
        if not (1e-300 < ref < 1e300):
            continue
        assert _ulps(oracle.pow64(x, y), ref) <= 2.0 * max(1.0, abs(y * math.log(x))), (x, y)


inf, nan = math.inf, math.nan
# C99 Annex F.9.4.4 (pow), written out case by case: (x, y, result)
POW_SPECIAL = [
    (nan, 0.0, 1.0), (nan, -0.0, 1.0), (1.0, nan, 1.0), (1.0, inf, 1.0), (-1.0, inf, 1.0),
    (-1.0, -inf, 1.0), (0.0, -3.0, inf), (-0.0, -3.0, -inf), (0.0, -2.0, inf), (-0.0, -2.0, inf),
    (0.0, -0.5, inf), (-0.0, -inf, inf), (0.0, 3.0, 0.0), (-0.0, 3.0, -0.0), (-0.0, 2.0, 0.0),
    (-0.0, 0.5, 0.0), (0.5, -inf, inf), (2.0, -inf, 0.0), (0.5, inf, 0.0), (2.0, inf, inf),
    (-inf, -3.0, -0.0), (-inf, -2.0, 0.0), (-inf, -0.5, 0.0), (-inf, 3.0, -inf), (-inf, 2.0, inf),
    (-inf, 0.5, inf), (inf, -1.0, 0.0), (inf, 0.5, inf), (-2.0, 0.5, nan), (-2.0, 3.0, -8.0),
    (-2.0, -2.0, 0.25), (2.0, nan, nan), (nan, 1.0, nan),
]


@pytest.mark.parametrize("x,y,want", POW_SPECIAL)
def test_pow64_special_cases(oracle, x, y, want):
    got = oracle.pow64(x, y)
    if math.isnan(want):
        assert math.isnan(got)
    else:
        assert got == want and math.copysign(1.0, got) == math.copysign(1.0, want)


def test_fmax_fmin_nan_operands(oracle):
    """f64::max / f64::min: a NaN operand yields the other operand."""
    assert oracle.fmax64(nan, 2.0) == 2.0 and oracle.fmax64(2.0, nan) == 2.0
    assert oracle.fmin64(nan, -1.0) == -1.0 and oracle.fmin64(-1.0, nan) == -1.0
    assert math.isnan(oracle.fmax64(nan, nan))
    assert oracle.fmax64(3.0, 2.5) == 3.0 and oracle.fmin64(3.0, 2.5) == 2.5


# ---- C ABI validation (host side, no device needed) ----------------------


def _abi_net(S=2):
    from rebop_amd.engine import GillespieNet

    return GillespieNet(np.zeros(S, dtype=np.int64), seed=1)


@pytest.mark.parametrize("code,consts,status", [
    ([], [], _native.REBOP_ERR_VALUE),
    ([_native.REBOP_EXPR_ADD], [], _native.REBOP_ERR_VALUE),
    ([C | (1 << 8)], [1.0], _native.REBOP_ERR_INDEX),
    ([SP | (2 << 8)], [], _native.REBOP_ERR_INDEX),
    ([C, C], [1.0], _native.REBOP_ERR_VALUE),
    ([77], [], _native.REBOP_ERR_VALUE),
    ([C] * 33 + [_native.REBOP_EXPR_ADD] * 32, [1.0], _native.REBOP_ERR_VALUE),
])
def test_abi_rejects_malformed_expressions(code, consts, status):
    g = _abi_net()
    lib = _native.load()
    c = np.ascontiguousarray(code, dtype=np.uint32)
    k = np.ascontiguousarray(consts, dtype=np.float64)
    d = np.zeros(2, dtype=np.int64)
    rc = lib.rebop_gillespie_add_reaction_expr(
        g.handle, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(c),
        k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(k),
        d.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(d))
    assert rc == status
    assert g.nb_reactions() == 0


def test_abi_accepts_expression_and_counts_it():
    from rebop_amd.engine import RateExpr

    g = _abi_net()
    code, consts = E.to_postfix(E.parse("0.2 * A * B / (5 + B)"), {"A": 0, "B": 1}, {})
    g.add_reaction(RateExpr.from_postfix(code, consts), [-1, 1])
    assert g.nb_reactions() == 1
    with pytest.raises(_native.RebopError):
        g.add_reaction(RateExpr.from_postfix(code, consts), [-1, 1, 0])


# ---- Python surface (tests/test_rebop.py:92-166, pyo3_gillespie.rs) -------


@pytest.mark.parametrize("bad", TABLES["python_rate_errors"]["not_understood"])
def test_python_rate_not_understood(bad):
    s = Gillespie()
    with pytest.raises(ValueError, match=TABLES["python_rate_errors"]["not_understood_message"]):
        s.add_reaction(bad, [], ["A"])


def test_python_missing_parameter():
    s = Gillespie()
    s.add_reaction("k", [], ["A"])
    with pytest.raises(ValueError, match=TABLES["python_rate_errors"]["missing_parameter_message"]):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s.run({}, 10, 10)


def test_python_str_shows_expressions():
    s = Gillespie()
    s.add_reaction(14, [], ["A"])
    s.add_reaction("0.2 * B * C / (5 + C)", ["B"], ["D"])
    assert str(s) == ("3 species and 2 reactions\n"
                      " --> A @ LMA(14)\n"
                      "B --> D @ (((0.2 * B) * C) / (5 + C))\n")
