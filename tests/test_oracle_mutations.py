"""The oracle's pins catch plausible mistakes: each mutation below is planted in
an in-memory copy of oracle/rectv_oracle.py (on the CPU) and must make at least
one pin of tests/test_oracle_solve_pins.py fail, while the unmodified oracle
passes them all.  The mutations are the ones a restatement of the reference can
plausibly get wrong without any operator-level check noticing:

  * a time-reversed basis: phi and conj(phi) swapped in mulphi
    (kernels_operators.cuh:135-149);
  * the rho rule inverted (rectv.cu:231-232);
  * the sign of the dual update flipped (admm.cu:107);
  * the left halo slab taken from the iterate being written instead of the
    previous one (rectv.cu:170-171);
  * frame times rounded instead of C integer division (kernels_operators.cuh:113).
"""
import os
import types

import pytest

import test_oracle_solve_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "rectv_oracle.py")

MUTATIONS = {
    "time_reversed_basis": ("    ph = phij if c == 1 else np.conj(phij)\n",
                            "    ph = np.conj(phij) if c == 1 else phij\n"),
    "rho_rule_inverted": ("        if r > 10 * s:\n            rho *= 2\n        elif s > 10 * r:\n            rho *= 0.5\n",
                          "        if r > 10 * s:\n            rho *= 0.5\n        elif s > 10 * r:\n            rho *= 2\n"),
    "dual_update_sign": ("    mu += rho * h2 - rho * psi                                  # (:107)\n",
                         "    mu -= rho * h2 - rho * psi                                  # (:107)\n"),
    "left_halo_from_new_iterate": ("                fext[0] = f[lo - 1]\n",
                                   "                fext[0] = fn[lo - 1]\n"),
    "frame_times_rounded": ("    return (np.arange(m) * ntheta) // m\n",
                            "    return np.round(np.arange(m) * ntheta / m).astype(np.int64)\n"),
}


def mutated(name):
    old, new = MUTATIONS[name]
    src = open(SRC).read()
    assert src.count(old) == 1, "mutation site of %s not found in the oracle" % name
    mod = types.ModuleType("rectv_oracle_" + name)
    mod.__file__ = SRC
    exec(compile(src.replace(old, new), "rectv_oracle[%s]" % name, "exec"), mod.__dict__)
    return mod


def failing_pins(O):
    out = []
    for pin in P.PINS:
        try:
            pin(O)
        except (AssertionError, FloatingPointError, ValueError):
            out.append(pin.__name__)
    return out


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_mutation_is_caught(name):
    caught = failing_pins(mutated(name))
    assert caught, "mutation %s passes every pin" % name


def test_unmutated_copy_passes_every_pin():
    src = open(SRC).read()
    mod = types.ModuleType("rectv_oracle_copy")
    mod.__file__ = SRC
    exec(compile(src, "rectv_oracle[copy]", "exec"), mod.__dict__)
    assert failing_pins(mod) == []
