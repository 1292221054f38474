"""The oracle's pins must fail on a wrong oracle.

Each case below plants one plausible mistake (a swapped wavenumber or corner,
a flipped sign, a dropped source term, a wrong factor or convention) in a copy
of oracle/qss_oracle.cpp, builds the copy in a temporary directory and runs
the pins that cover that code (tests/test_oracle_fields.py, the gather pin of
tests/test_oracle_push.py, tests/test_oracle_beam.py) against it: every
mutant must be caught by at least one of them.  Only the copy is changed; the
oracle itself is not touched."""
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from tests.test_oracle_beam import check_beam, check_bins
from tests.test_oracle_fields import check_guess, check_iterate
from tests.test_oracle_push import check_external_b, check_gather

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")
SRC = os.path.join(ORACLE_DIR, "qss_oracle.cpp")

GUESS = [check_guess]
ITER = [check_iterate]
MUTANTS = {
    # predictor (Guess)
    "guess Ey with akz": ("double ey = -eyp + aky * (rb + rbp + rp + rho) / (dens + ak2);",
                          "double ey = -eyp + akz * (rb + rbp + rp + rho) / (dens + ak2);", GUESS),
    "guess Ez without the previous beam": (
        "double ez = -ezp + akz * (rb + rbp + rp + rho) / (dens + ak2);",
        "double ez = -ezp + akz * (rb + rp + rho) / (dens + ak2);", GUESS),
    "guess Jy from Ez": ("double jy = jyp + hx * dens * (ey + eyp) / 2.;",
                         "double jy = jyp + hx * dens * (ez + ezp) / 2.;", GUESS),
    "guess Jx without dens": ("double jx = jxp + hx * dens * (ex + exp_) / 2.;",
                              "double jx = jxp + hx * (ex + exp_) / 2.;", GUESS),
    "guess Bx sign": ("S[SP_BX][n1] = -aky / ak2 * jz + akz / ak2 * jy;",
                      "S[SP_BX][n1] = aky / ak2 * jz + akz / ak2 * jy;", GUESS),
    "guess By without the beam current": (
        "S[SP_BY][n1] = (-akz * (jx + S[SP_JXB][n1]) + dens * ez) / ak2;",
        "S[SP_BY][n1] = (-akz * jx + dens * ez) / ak2;", GUESS),
    "guess Bz sign of the dens term": (
        "S[SP_BZ][n1] = (aky * (jx + S[SP_JXB][n1]) - dens * ey) / ak2;",
        "S[SP_BZ][n1] = (aky * (jx + S[SP_JXB][n1]) + dens * ey) / ak2;", GUESS),
    "guess rho without the previous beam": (
        "- (rb + rbp) * dens * hx * hx / denominator", "- (rb + rb) * dens * hx * hx / denominator",
        GUESS),
    "guess Ex without the previous beam": (
        "2. * (2. * rp + rb + rbp) * dens * hx", "2. * (2. * rp + rb + rb) * dens * hx", GUESS),
    # corrector (Iterate)
    "iterate Ey from the new rho": (
        "double newEy = -eyp + aky * (rb + rbp + rp + S[SP_RHO][n1]) / (dens + ak2);",
        "double newEy = -eyp + aky * (rb + rbp + rp + rho) / (dens + ak2);", ITER),
    "iterate Ez with aky": (
        "double newEz = -ezp + akz * (rb + rbp + rp + S[SP_RHO][n1]) / (dens + ak2);",
        "double newEz = -ezp + aky * (rb + rbp + rp + S[SP_RHO][n1]) / (dens + ak2);", ITER),
    "iterate Bx with Jy and Jz swapped": (
        "S[SP_BX][n1] = -aky / ak2 * S[SP_JZ][n1] + akz / ak2 * S[SP_JY][n1];",
        "S[SP_BX][n1] = -aky / ak2 * S[SP_JY][n1] + akz / ak2 * S[SP_JZ][n1];", ITER),
    "iterate By from Ey": (
        "S[SP_BY][n1] = (-akz * (S[SP_JX][n1] + S[SP_JXB][n1]) + dens * newEz) / ak2;",
        "S[SP_BY][n1] = (-akz * (S[SP_JX][n1] + S[SP_JXB][n1]) + dens * newEy) / ak2;", ITER),
    "iterate continuity with aky and akz swapped": (
        "hx * (aky * (jy + S[SP_JYP][n1]) + akz * (jz + S[SP_JZP][n1])) / 2.;",
        "hx * (akz * (jy + S[SP_JYP][n1]) + aky * (jz + S[SP_JZP][n1])) / 2.;", ITER),
    "iterate residual not normalised": ("err = std::max(err, dif / mx);",
                                        "err = std::max(err, dif);", ITER),
    "psi with rho + Jx": ("ps[n1] = (S[SP_RHO][n1] - S[SP_JX][n1]) / ak2;",
                          "ps[n1] = (S[SP_RHO][n1] + S[SP_JX][n1]) / ak2;", ITER),
    # gather
    "CIC corners (j+1,k) and (j,k+1) swapped": (
        "f[c] = acc * F[c][n00] + apc * F[c][n10] + acp * F[c][n01] + app * F[c][n11];",
        "f[c] = acc * F[c][n00] + acp * F[c][n10] + apc * F[c][n01] + app * F[c][n11];",
        [lambda mk: check_gather(mk, 1)]),
    "push field from the C layer only": (
        "for (long n = 0; n < ncell; ++n) F[c][n] = 0.5 * (o->P[c][n] + o->C[c][n]);",
        "for (long n = 0; n < ncell; ++n) F[c][n] = o->C[c][n];",
        [lambda mk: check_gather(mk, 0), lambda mk: check_gather(mk, 1)]),
    "NGP gather from the next node": ("for (int c = 0; c < 6; ++c) f[c] = F[c][n00];",
                                      "for (int c = 0; c < 6; ++c) f[c] = F[c][n00 + 1];",
                                      [lambda mk: check_gather(mk, 0)]),
    "external B dropped": ("  by += p.b_ext[1];\n", "\n", [check_external_b]),
    # frozen driver
    "beam R = sigma": ("double Ry = std::sqrt(2.0) * B.sigma_r;", "double Ry = B.sigma_r;",
                       [check_beam, check_bins]),
    "beam one layer late": ("    double xi = (double)(L - ci) * p.hx;\n    for (int b = 0;",
                            "    double xi = (double)(L - ci - 1) * p.hx;\n    for (int b = 0;",
                            [check_beam, check_bins]),
    "beam density at the cell centre": ("          double yc = o->hy * cj - p.ly / 2.0;",
                                        "          double yc = o->hy * (cj + 0.5) - p.ly / 2.0;",
                                        [check_beam]),
    "binned profile with the Gaussian's exponent halved": (
        "return o->bins[ib] * std::exp(-arg);", "return o->bins[ib] * std::exp(-0.5 * arg);",
        [check_bins]),
}


def _build(name, tmp):
    old, new, _ = MUTANTS[name]
    text = open(SRC).read()
    assert text.count(old) == 1, "mutation site not unique: %s" % name
    d = os.path.join(tmp, "m%d" % sorted(MUTANTS).index(name))
    os.makedirs(d)
    src = os.path.join(d, "qss_oracle.cpp")
    with open(src, "w") as f:
        f.write(text.replace(old, new))
    so = os.path.join(d, "liboracle.so")
    cmd = [os.environ.get("CXX", "g++"), "-O1", "-std=c++17", "-fPIC", "-fopenmp",
           "-ffp-contract=off", "-shared", "-I", ORACLE_DIR, src, "-o", so]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return so


@pytest.fixture(scope="module")
def mutant_libs():
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")
    try:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            libs = dict(zip(MUTANTS, ex.map(lambda n: _build(n, tmp), MUTANTS)))
        yield libs
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(oracle_mod, mutant_libs, name):
    L = oracle_mod.load_library(mutant_libs[name])
    make = lambda p: oracle_mod.Oracle(p, history=(), library=L)
    caught = []
    for check in MUTANTS[name][2]:
        try:
            check(make)
        except AssertionError as e:
            caught.append(str(e).splitlines()[0][:120])
    assert caught, "no pin fails on the mutant '%s'" % name
