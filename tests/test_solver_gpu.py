"""The reference's own GPU solvers (Glucose GPU and rel-newtech GPU), built
from their sources against this library (oracle/solver.mk -> oracle/_ref/),
solve instances with
the MI355X engine behind GpuClauseSharer: satisfiable answers come with a
model that satisfies every clause, unsatisfiable ones are pigeonhole
formulas, and the run reports clauses through the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOLVERS = {name: os.path.join(ROOT, "oracle", "_ref", name)
           for name in ("glucose-gpu", "rel-newtech-gpu")}


def needs(name):
    return pytest.mark.skipif(
        not os.path.exists(SOLVERS[name]),
        reason=f"oracle/_ref/{name} is built from the reference sources (oracle/solver.mk), "
               "which are not present here")


def write_cnf(path, nvars, clauses):
    with open(path, "w") as f:
        f.write(f"p cnf {nvars} {len(clauses)}\n")
        for c in clauses:
            f.write(" ".join(str(x) for x in c) + " 0\n")


def planted_3sat(rng, n, ratio):
    sol = rng.integers(0, 2, size=n + 1).astype(bool)
    out = []
    while len(out) < int(ratio * n):
        vs = rng.choice(np.arange(1, n + 1), size=3, replace=False)
        lits = [int(v) if rng.random() < 0.5 else -int(v) for v in vs]
        if any((l > 0) == sol[abs(l)] for l in lits):
            out.append(lits)
    return out


def pigeonhole(p, h):
    var = lambda i, j: i * h + j + 1
    cl = [[var(i, j) for j in range(h)] for i in range(p)]
    for j in range(h):
        for a in range(p):
            for b in range(a + 1, p):
                cl.append([-var(a, j), -var(b, j)])
    return p * h, cl


def solve(path, solver="glucose-gpu", threads=4, env=None):
    r = subprocess.run([SOLVERS[solver], "-model", "-verb=1", f"-thread-count={threads}",
                        "-write-stats-period-sec=-1", path],
                       capture_output=True, text=True, timeout=240,
                       env=None if env is None else {**os.environ, **env})
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    return r


@pytest.mark.parametrize("solver", [pytest.param(n, marks=needs(n)) for n in SOLVERS])
def test_reference_solver_sat_instances_with_model(tmp_path, solver):
    rng = np.random.default_rng(4)
    for n, ratio in ((150, 4.0), (250, 4.2), (600, 4.2)):
        cls = planted_3sat(rng, n, ratio)
        path = str(tmp_path / f"planted_{n}.cnf")
        write_cnf(path, n, cls)
        r = solve(path, solver)
        assert r.returncode == 10, r.returncode
        assert "s SATISFIABLE" in r.stdout
        model = set()
        for line in r.stdout.splitlines():
            if line.startswith("v "):
                model |= {int(x) for x in line[2:].split() if x != "0"}
        for c in cls:
            assert any(l in model for l in c), c


@pytest.mark.parametrize("solver,screen",
                         [pytest.param(n, "auto", marks=needs(n)) for n in SOLVERS]
                         + [pytest.param("glucose-gpu", s, marks=needs("glucose-gpu"))
                            for s in ("0", "1")])
def test_reference_solver_unsat_pigeonhole(tmp_path, solver, screen):
    # gpuRun screened against the table's aggregate: "auto" (the default)
    # picks the form by the runs' device times, "1" always, "0" never
    n, cls = pigeonhole(8, 7)
    path = str(tmp_path / "php_8_7.cnf")
    write_cnf(path, n, cls)
    r = solve(path, solver, env={"GPUSHARESAT_SCREEN": screen})
    assert r.returncode == 20, r.returncode
    assert "s UNSATISFIABLE" in r.stdout
    if solver != "glucose-gpu":
        return  # only Glucose GPU prints its sharing statistics at exit
    # the solver threads shared clauses through the GPU and got reports back
    stat = lambda name: sum(int(x) for x in re.findall(rf'"{name}":(\d+)', r.stdout))
    assert stat("gpuRuns") > 0 and stat("gpuClausesAdded") > 0, r.stdout[-2000:]
    assert stat("reportedClauses") > 0
    assert stat("clauseTestsOnAssigs") > 0
