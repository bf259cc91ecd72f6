"""The define_system analogue on the CPU: the oracle's restatement of the
macro's arithmetic (src/gillespie_macro.rs) pinned by closed forms, brute
force, the macro's own unit tests (:173-260) and a two-sample test against the
run-time semantics; the text front end; the C-ABI argument checks."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.pyoracle import OracleNet, choose_cumrate_sum, choose_first_above
from rebop_amd import GillespieNet, Rate, _native, define_system
from rebop_amd.system import parse_system
from tests.networks import build_oracle, dimers

# ---- the oracle's define_system arithmetic ---------------------------------


def _rate(k, reactants, x):
    n = OracleNet.system(len(x))
    n.add_system(k, reactants, [])
    return n.rate(0, x)


def test_integer_mass_action_products():
    """_rate_lma!: x (x-1) ... (x-m+1) in isize, converted once, then
    `rate * lma * lma` in written order."""
    assert _rate(2.0, [(0, 2)], [5, 0]) == 40.0           # 2 * 5 * 4
    assert _rate(2.0, [(0, 1), (0, 1)], [5, 0]) == 50.0   # A + A is A^2, not A (A-1)
    assert _rate(2.0, [(0, 0)], [5, 0]) == 10.0           # the literal 0: the loop 1..0 is empty, factor x
    assert _rate(2.0, [(0, 3)], [2, 0]) == 0.0            # 2 * 1 * 0
    assert _rate(2.0, [(0, 3)], [5, 0]) == 120.0
    assert _rate(1.5, [], [5, 0]) == 1.5                  # no reactant: the rate itself
    # written order: (0.1 * 3) * 20 != (0.1 * 20) * 3 in f64
    assert _rate(0.1, [(1, 1), (0, 2)], [5, 3]) == (0.1 * 3.0) * 20.0 != (0.1 * 20.0) * 3.0
    assert _rate(0.1, [(0, 2), (1, 1)], [5, 3]) == (0.1 * 20.0) * 3.0


def test_integer_product_is_converted_once():
    """k * f64(x (x-1)) differs from the run-time chain (k * x) * (x-1) for
    some inputs; the system form must give the former."""
    found = 0
    for x in range(2, 400):
        for k in (0.1, 0.3, 0.7, 1.1, 2.9):
            sys_v = k * float(x * (x - 1))
            chain = (k * float(x)) * float(x - 1)
            assert _rate(k, [(0, 2)], [x]) == sys_v
            found += sys_v != chain
    assert found > 10  # the two forms really differ on these inputs


def test_integer_product_wraps_like_a_release_build():
    """isize arithmetic wraps: 2^40 (2^40 - 1) mod 2^64 is -2^40 as a signed value."""
    x = 1 << 40
    p = (x * (x - 1)) % (1 << 64)
    signed = p - (1 << 64) if p >= 1 << 63 else p
    assert signed == -(1 << 40)
    assert _rate(1.0, [(0, 2)], [x]) == float(signed)
    # negative counts: (-3)(-4) = 12
    assert _rate(1.0, [(0, 2)], [-3]) == 12.0


def test_first_match_choice_brute_force():
    """_choice!: the first r with chosen < cum[r] (strict), n if none; the
    run-time choose_cumrate_sum counts cum[r] < chosen instead."""
    cum = [0.0, 0.0, 1.0, 1.0, 3.0]
    assert choose_first_above(0.0, cum) == 2    # zero-propensity leading reactions skipped
    assert choose_cumrate_sum(0.0, cum) == 0    # the run-time quirk picks reaction 0
    assert choose_first_above(1.0, cum) == 4    # ties go past the equal sums
    assert choose_cumrate_sum(1.0, cum) == 2
    assert choose_first_above(3.0, cum) == 5    # none
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = int(rng.integers(1, 9))
        c = np.cumsum(rng.integers(0, 3, size=n).astype(float))
        chosen = float(rng.integers(0, int(c[-1]) + 2)) + rng.choice([0.0, 0.5])
        first = next((r for r in range(n) if chosen < c[r]), n)
        assert choose_first_above(chosen, c) == first
        if chosen not in set(c.tolist()):  # off the ties, both rules agree
            assert choose_cumrate_sum(chosen, c) == min(first, n)


def test_system_net_jump_and_refusals():
    n = OracleNet.system(3)
    n.add_system(1.0, [(0, 1), (1, 2)], [(2, 1), (0, 1)])  # A + 2B => C + A
    x = n.affect(0, [5, 5, 5])
    assert x.tolist() == [5, 3, 6]
    with pytest.raises(Exception):
        n.add_lma(1.0, [1, 0, 0], [-1, 0, 0])
    with pytest.raises(Exception):
        n.add_system(1.0, [(3, 1)], [])
    m = OracleNet(2)
    with pytest.raises(Exception):
        m.add_system(1.0, [(0, 1)], [])


# ---- the macro's own unit tests (gillespie_macro.rs:173-260) on the oracle --


def _system_oracle(species, reactions):
    idx = {s: i for i, s in enumerate(species)}
    n = OracleNet.system(len(species))
    for k, re, pr in reactions:
        n.add_system(k, [(idx[s], m) for s, m in re], [(idx[s], m) for s, m in pr])
    return n


SIR_SYS = (["S", "I", "R"], [(0.1 / 10000.0, [("S", 1), ("I", 1)], [("I", 1), ("I", 1)]),
                             (0.01, [("I", 1)], [("R", 1)])])
DIMERS_SYS = (["gene", "mRNA", "protein", "dimer"], [
    (25.0, [("gene", 1)], [("gene", 1), ("mRNA", 1)]),
    (1000.0, [("mRNA", 1)], [("mRNA", 1), ("protein", 1)]),
    (0.001, [("protein", 2)], [("dimer", 1)]),
    (0.1, [("mRNA", 1)], []),
    (1.0, [("protein", 1)], []),
])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_macro_sir(seed):
    n = _system_oracle(*SIR_SYS)
    tr = n.state([9999, 1, 0], seed)
    tr.advance_until(1000.0)
    assert tr.x.sum() == 10000 and tr.t == 1000.0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_macro_dimers(seed):
    n = _system_oracle(*DIMERS_SYS)
    tr = n.state([1, 0, 0, 0], seed)
    tr.advance_until(1.0)
    assert tr.x[0] == 1 and 1000 < tr.x[3] < 10000


def test_macro_birth_death_and_forgotten_parameter():
    n = _system_oracle(["A"], [(10.0, [], [("A", 1)]), (0.1, [("A", 1)], [])])
    tr = n.state([0], 5)
    tr.advance_until(100.0)
    assert 50 < tr.x[0] < 200
    # r_death never set: NaN, so no reaction happens and t reaches tmax
    n = _system_oracle(["A"], [(10.0, [], [("A", 1)]), (math.nan, [("A", 1)], [])])
    tr = n.state([0], 5)
    tr.advance_until(100.0)
    assert abs(tr.t - 100.0) < np.finfo(float).eps and tr.x[0] == 0 and tr.draws == 0


def test_macro_no_reactions():
    n = OracleNet.system(3)
    tr = n.state([42, 1337, 0], 1)
    tr.advance_until(1e20)
    assert abs(tr.t - 1e20) < np.finfo(float).eps and tr.x.tolist() == [42, 1337, 0]


def test_macro_semantics_same_law_as_run_time(nthreads):
    """The macro's arithmetic changes roundings and tie-breaks only: the dimer
    count at t = 1 has the same law as with the run-time semantics
    (two-sample Kolmogorov-Smirnov, independent seeds)."""
    from scipy.stats import ks_2samp

    n = 4000
    sysnet = _system_oracle(*DIMERS_SYS)
    _, a, _, _, _ = sysnet.run_grid(11, [1, 0, 0, 0], 1.0, 1, n_traj=n, nthreads=nthreads)
    rt = build_oracle(dimers())
    _, b, _, _, _ = rt.run_grid(12, [1, 0, 0, 0], 1.0, 1, n_traj=n, nthreads=nthreads)
    assert ks_2samp(a[-1, 3], b[-1, 3]).pvalue > 1e-3
    assert ks_2samp(a[-1, 2], b[-1, 2]).pvalue > 1e-3


# ---- the text front end ------------------------------------------------------

LIB_EXAMPLE = """
    r_tx r_tl r_dim r_decay_mRNA r_decay_prot;
    Dimers { gene, mRNA, protein, dimer }
    transcription   : gene      => gene + mRNA      @ r_tx
    translation     : mRNA      => mRNA + protein   @ r_tl
    dimerization    : 2 protein => dimer            @ r_dim
    decay_mRNA      : mRNA      =>                  @ r_decay_mRNA
    decay_protein   : protein   =>                  @ r_decay_prot
"""


def test_parse_reference_examples():
    spec = parse_system(LIB_EXAMPLE)  # src/lib.rs:23-32
    assert spec.name == "Dimers"
    assert spec.params == ("r_tx", "r_tl", "r_dim", "r_decay_mRNA", "r_decay_prot")
    assert spec.species == ("gene", "mRNA", "protein", "dimer")
    assert [r.name for r in spec.reactions] == ["transcription", "translation", "dimerization",
                                                "decay_mRNA", "decay_protein"]
    assert spec.reactions[2].reactants == (("protein", 2),)
    assert spec.reactions[2].products == (("dimer", 1),)
    assert spec.reactions[3].products == ()
    sir = parse_system("r_inf r_heal; SIR { S, I, R } infection : S + I => 2 I @ r_inf "
                       "healing : I => R @ r_heal")  # src/lib.rs:66-71, one line
    assert sir.reactions[0].reactants == (("S", 1), ("I", 1))
    assert sir.reactions[0].products == (("I", 2),)
    empty = parse_system(";\n FooBarBuz { Foo, Bar, Buz }")  # gillespie_macro.rs:245-249
    assert empty.params == () and empty.reactions == ()
    bd = parse_system("r_birth r_death; BirthDeath { A } birth: => A @ r_birth // a comment\n"
                      " death: A => @ r_death")
    assert bd.reactions[0].reactants == () and bd.reactions[1].products == ()


def test_rate_expressions_evaluate_in_f64():
    S = define_system("r1 r2; SIR { S, I, R } inf : S + I => I + I @ 0.1 / 10000. "
                      "heal : I => R @ -(r1 - 2.) * r2 + 1e-3")
    s = S.with_parameters(0.5, 4.0)
    assert s.rates() == [0.1 / 10000.0, -(0.5 - 2.0) * 4.0 + 1e-3]
    z = define_system("k; Z { A } r : A => @ k / 0.")
    assert z.with_parameters(1.0).rates() == [math.inf]
    assert math.isnan(z.new().rates()[0])


@pytest.mark.parametrize("text,match", [
    ("k; X { A } r : A => @ A * k", "reads species"),
    ("k; X { A } r : B => @ k", "unknown species"),
    ("k; X { A } r : A => @ q", "unknown parameter"),
    ("k; X { t } r : t => @ k", "reserved"),
    ("k; X { A, k } r : A => @ k", "already"),
    ("k; X { A } r : A => @ k r : A => @ k", "already"),
    ("k; X { A } r : 65 A => @ k", "beyond"),
    ("k; X { A } r : 2.5 A => @ k", "not an integer"),
    ("k; X { A } r : A => @ k $", "unexpected character"),
    ("k X { A }", "expected ';'"),
])
def test_front_end_errors(text, match):
    with pytest.raises(ValueError, match=match):
        define_system(text)


def test_generated_class_fields():
    D = define_system(LIB_EXAMPLE)
    d = D.new()
    assert d.gene == 0 and d.t == 0.0 and math.isnan(d.r_tx)
    e = D.with_parameters(25.0, 1000.0, 0.001, 0.1, 1.0)
    assert (e.r_tx, e.r_decay_prot) == (25.0, 1.0)
    with pytest.raises(TypeError):
        D.with_parameters(1.0)
    e.seed(7)
    assert e._draws == 0 and e._seed == 7
    assert repr(e).startswith("Dimers { gene: 0, mRNA: 0, protein: 0, dimer: 0, r_tx: 25.0")


# ---- C-ABI argument checks (host-only calls) ---------------------------------


def test_system_handle_argument_checks():
    g = GillespieNet.system([0, 0], seed=1)
    with pytest.raises(_native.RebopError) as e:
        g.add_reaction(Rate.lma(1.0, [1, 0]), [-1, 1])
    assert e.value.code == _native.REBOP_ERR_UNSUPPORTED
    with pytest.raises(_native.RebopValueError) as e:
        g.add_system_reaction(1.0, [(2, 1)], [])
    assert e.value.code == _native.REBOP_ERR_INDEX
    with pytest.raises(_native.RebopValueError) as e:
        g.add_system_reaction(1.0, [(0, _native.REBOP_SYSTEM_MAX_COUNT + 1)], [])
    assert e.value.code == _native.REBOP_ERR_VALUE
    with pytest.raises(_native.RebopValueError):
        g.add_system_reaction(math.inf, [(0, 1)], [])
    g.add_system_reaction(math.nan, [(0, 1)], [(1, 1)])  # a parameter never set
    g.add_system_reaction(2.0, [(0, 0), (0, 1)], [(1, 3)])
    assert g.nb_reactions() == 2
    h = GillespieNet([0, 0], seed=1)
    with pytest.raises(_native.RebopError) as e:
        h.add_system_reaction(1.0, [(0, 1)], [])
    assert e.value.code == _native.REBOP_ERR_UNSUPPORTED
